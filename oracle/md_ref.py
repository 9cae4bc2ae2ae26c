"""TEST INFRASTRUCTURE ONLY: numpy restatement of the reference's rigid-body
NVE step (nve_1, source/evolve.f90:621-1062, with nosquish, :8107-8204, sites,
molecular_sites.f90:14-111, and the kinetic energies of utility.f90:164-194),
used by tests/ to check the device driver.  Forces come from a callable,
normally the CPU oracle of the three-body term.  Arrays: com, vel, jbody
[nm,3]; quat [nm,4]; site offsets and site forces [nm,ns,3]."""
from __future__ import annotations

import numpy as np


def quat_to_rot(q):
    """rott of quat_to_rot (molecular_sites.f90:95-111) as [nm,3,3] row-major."""
    q0, q1, q2, q3 = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    r = np.empty((len(q), 3, 3))
    r[:, 0, 0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3
    r[:, 0, 1] = 2.0 * (q1 * q2 - q0 * q3)
    r[:, 0, 2] = 2.0 * (q1 * q3 + q0 * q2)
    r[:, 1, 0] = 2.0 * (q1 * q2 + q0 * q3)
    r[:, 1, 1] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3
    r[:, 1, 2] = 2.0 * (q2 * q3 - q0 * q1)
    r[:, 2, 0] = 2.0 * (q1 * q3 - q0 * q2)
    r[:, 2, 1] = 2.0 * (q2 * q3 + q0 * q1)
    r[:, 2, 2] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3
    return r


def sites(q, pcoord):
    """xs(js,j) = R(q_j) pcoord(:,js) (transform iway=1, molecular_sites.f90:14-82)."""
    return np.einsum("mab,sb->msa", quat_to_rot(q), pcoord)


def nosquish(dt, q, p, pmi, mrot=10):
    """nosquish (evolve.f90:8107-8204), vectorised over molecules."""
    q, p = q.copy(), p.copy()
    tr = dt / mrot

    def rot(q, p, a, b, c, d, sa, sb, sc, sd, z):
        cs, sn = np.cos(z), np.sin(z)
        qn = np.empty_like(q)
        pn = np.empty_like(p)
        # component k of the new vector = cs * v[k] + sign[k] * sn * v[perm[k]]
        perm = (a, b, c, d)
        sg = (sa, sb, sc, sd)
        for k in range(4):
            qn[:, k] = cs * q[:, k] + sg[k] * sn * q[:, perm[k]]
            pn[:, k] = cs * p[:, k] + sg[k] * sn * p[:, perm[k]]
        return qn, pn

    for _ in range(mrot):
        z = 0.125 / pmi[2] * tr * (-p[:, 0] * q[:, 3] + p[:, 1] * q[:, 2] - p[:, 2] * q[:, 1] + p[:, 3] * q[:, 0])
        q, p = rot(q, p, 3, 2, 1, 0, -1, 1, -1, 1, z)
        z = 0.125 / pmi[1] * tr * (-p[:, 0] * q[:, 2] - p[:, 1] * q[:, 3] + p[:, 2] * q[:, 0] + p[:, 3] * q[:, 1])
        q, p = rot(q, p, 2, 3, 0, 1, -1, -1, 1, 1, z)
        z = 0.25 / pmi[0] * tr * (-p[:, 0] * q[:, 1] + p[:, 1] * q[:, 0] + p[:, 2] * q[:, 3] - p[:, 3] * q[:, 2])
        q, p = rot(q, p, 1, 0, 3, 2, -1, 1, 1, -1, z)
        z = 0.125 / pmi[1] * tr * (-p[:, 0] * q[:, 2] - p[:, 1] * q[:, 3] + p[:, 2] * q[:, 0] + p[:, 3] * q[:, 1])
        q, p = rot(q, p, 2, 3, 0, 1, -1, -1, 1, 1, z)
        z = 0.125 / pmi[2] * tr * (-p[:, 0] * q[:, 3] + p[:, 1] * q[:, 2] - p[:, 2] * q[:, 1] + p[:, 3] * q[:, 0])
        q, p = rot(q, p, 3, 2, 1, 0, -1, 1, -1, 1, z)
    return q, p


def _force_torque(site, fsite, fcom):
    F = fsite.sum(1) + fcom
    T = np.cross(site, fsite).sum(1)
    return F, T


def _q_of_j(q, j):
    """Quaternion momentum 2 S(q) j (evolve.f90:738-741)."""
    q0, q1, q2, q3 = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    x, y, z = j[:, 0], j[:, 1], j[:, 2]
    return np.stack([2 * (-q1 * x - q2 * y - q3 * z), 2 * (q0 * x - q3 * y + q2 * z),
                     2 * (q3 * x + q0 * y - q1 * z), 2 * (-q2 * x + q1 * y + q0 * z)], 1)


def _j_of_p(q, p):
    """Body angular momentum 0.5 S(q)^T p (evolve.f90:795-804)."""
    q0, q1, q2, q3 = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    p0, p1, p2, p3 = p[:, 0], p[:, 1], p[:, 2], p[:, 3]
    return np.stack([0.5 * (-q1 * p0 + q0 * p1 + q3 * p2 - q2 * p3),
                     0.5 * (-q2 * p0 - q3 * p1 + q0 * p2 + q1 * p3),
                     0.5 * (-q3 * p0 + q2 * p1 - q1 * p2 + q0 * p3)], 1)


def _kick(st, site, fsite, fcom, dt, totm):
    F, T = _force_torque(site, fsite, fcom)
    tq = np.einsum("mba,mb->ma", quat_to_rot(st["quat"]), T)  # R^T t, body-frame torque
    p = _q_of_j(st["quat"], st["jbody"]) + 0.5 * dt * _q_of_j(st["quat"], tq)
    v = st["vel"] + 0.5 * F * dt / totm
    return p, v


def nve_step(st, forces, force_fn, dt, pcoord, totm, pmi, box):
    """One step; st: dict(com, quat, vel, jbody); forces: (fsite, fcom) at st.
    Returns (new state, new forces, (E3, Kt, Kr) in kcal/mol)."""
    site = sites(st["quat"], pcoord)
    p, v = _kick(st, site, forces[0], forces[1], dt, totm)
    com = st["com"] + dt * v
    q, p = nosquish(dt, st["quat"], p, pmi)
    new = {"com": com, "quat": q, "vel": v, "jbody": _j_of_p(q, p)}
    site = sites(q, pcoord)
    fsite, fcom, e3 = force_fn(new["com"], site)
    p, v = _kick(new, site, fsite, fcom, dt, totm)
    new["vel"] = v
    new["jbody"] = _j_of_p(new["quat"], p)
    if box > 0:
        d = new["com"] / box
        new["com"] = new["com"] - box * (np.sign(d) * np.floor(np.abs(d) + 0.5))  # NINT: half away from zero
    kt = 0.5 * totm * (v ** 2).sum() / 418.4
    kr = 0.5 * (new["jbody"] ** 2 / np.asarray(pmi)[None, :]).sum() / 418.4
    return new, (fsite, fcom), (e3, kt, kr)
