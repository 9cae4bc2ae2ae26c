"""Rigid-water geometry: quaternions -> lab-frame site offsets, and the seeded
synthetic systems the benchmarks and tests use.

``quat_to_rot`` and ``site_offsets`` restate molecular_sites.f90:14-111
(``sites`` / ``transform(iway=1)`` / ``quat_to_rot``).  ``principal_coords``
restates shift_to_com + principle_moment_inertia (initialization.f90:728-845)
for the shipped FIELD, whose inertia tensor is already diagonal, so the Jacobi
rotation is the identity and the principal-frame coordinates are the FIELD
coordinates minus the centre of mass.
"""
from __future__ import annotations

import math

import numpy as np

from .params import field_path, read_field_sites

# A**3 per molecule at 0.997 g/cm^3; L = (N * V1)^(1/3) gives the box lengths
# 12.42 / 19.71 / 24.84 A quoted for 64 / 256 / 512 waters.
VOLUME_PER_WATER = 29.915


def box_length(n: int) -> float:
    return (n * VOLUME_PER_WATER) ** (1.0 / 3.0)


def quat_to_rot(q: np.ndarray) -> np.ndarray:
    """Rotation matrices [..., 3, 3] from quaternions [..., 4] = (q0,q1,q2,q3)
    (molecular_sites.f90:95-111; row-major rott(1..9))."""
    q0, q1, q2, q3 = (q[..., k] for k in range(4))
    r = np.empty(q.shape[:-1] + (3, 3))
    r[..., 0, 0] = q0**2 + q1**2 - q2**2 - q3**2
    r[..., 0, 1] = 2.0 * (q1 * q2 - q0 * q3)
    r[..., 0, 2] = 2.0 * (q1 * q3 + q0 * q2)
    r[..., 1, 0] = 2.0 * (q1 * q2 + q0 * q3)
    r[..., 1, 1] = q0**2 - q1**2 + q2**2 - q3**2
    r[..., 1, 2] = 2.0 * (q2 * q3 - q0 * q1)
    r[..., 2, 0] = 2.0 * (q1 * q3 - q0 * q2)
    r[..., 2, 1] = 2.0 * (q2 * q3 + q0 * q1)
    r[..., 2, 2] = q0**2 - q1**2 - q2**2 + q3**2
    return r


def principal_coords(path=None) -> tuple[list[str], np.ndarray, np.ndarray]:
    """Site names, principal-frame site coordinates [ns,3] (A) and masses."""
    names, xyz, mass, _ = read_field_sites(path or field_path())
    com = (mass[:, None] * xyz).sum(0) / mass.sum()
    pc = xyz - com
    inertia = np.zeros((3, 3))
    for k in range(len(mass)):
        inertia += mass[k] * (np.dot(pc[k], pc[k]) * np.eye(3) - np.outer(pc[k], pc[k]))
    off = inertia - np.diag(np.diag(inertia))
    if np.abs(off).max() > 1e-12 * np.abs(inertia).max():
        raise ValueError("FIELD inertia tensor is not diagonal; principal rotation not implemented")
    return names, pc, mass


def site_offsets(quats: np.ndarray, pcoord: np.ndarray) -> np.ndarray:
    """Lab-frame site offsets [nm, ns, 3] = R(q) . pcoord (transform iway=1)."""
    rot = quat_to_rot(quats)
    return np.einsum("mab,sb->msa", rot, pcoord)


def random_quaternions(n: int, rng: np.random.Generator) -> np.ndarray:
    """Uniform random unit quaternions (Shoemake's method)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def jittered_lattice(n: int, L: float, rng: np.random.Generator, jitter: float = 0.3) -> np.ndarray:
    """COM positions [n,3] in [-L/2, L/2): a simple cubic lattice of
    nc = ceil(n^(1/3)) points per side (spacing L/nc); when nc^3 > n a seeded
    random subset of n lattice points is kept (the same rule for 256 waters on
    a 7^3 lattice); each point gets a uniform +-jitter A displacement per axis."""
    nc = int(round(n ** (1.0 / 3.0)))
    if nc**3 < n:
        nc += 1
    g = (np.arange(nc) + 0.5) * (L / nc) - 0.5 * L
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    if len(pts) > n:
        pts = pts[np.sort(rng.permutation(len(pts))[:n])]
    pts = pts + rng.uniform(-jitter, jitter, size=pts.shape)
    return pts - L * np.round(pts / L)


class WaterBox:
    """One periodic box of rigid waters in the reference's array layout.

    ``x,y,z`` [nm] COM in A; ``xs,ys,zs`` Fortran-order site offsets with
    leading dimension ``ld`` (``xs[is + ld*i]``), A; ``box`` in A (<=0: open).
    """

    def __init__(self, com: np.ndarray, quats: np.ndarray, box: float, pcoord=None, ld: int = 28):
        if pcoord is None:
            pcoord = principal_coords()[1]
        self.com = np.ascontiguousarray(com, dtype=np.float64)
        self.quats = np.ascontiguousarray(quats, dtype=np.float64)
        self.box = float(box)
        self.pcoord = np.asarray(pcoord, dtype=np.float64)
        self.ld = int(ld)
        self.refresh_sites()

    @property
    def nm(self) -> int:
        return len(self.com)

    def refresh_sites(self):
        off = site_offsets(self.quats, self.pcoord)  # [nm, ns, 3]
        ns = off.shape[1]
        if ns > self.ld:
            raise ValueError("leading dimension smaller than site count")
        buf = np.zeros((3, self.nm, self.ld))
        buf[:, :, :ns] = off.transpose(2, 0, 1)
        self.site = off
        self.xs, self.ys, self.zs = (np.ascontiguousarray(buf[a].reshape(-1)) for a in range(3))

    @property
    def x(self):
        return np.ascontiguousarray(self.com[:, 0])

    @property
    def y(self):
        return np.ascontiguousarray(self.com[:, 1])

    @property
    def z(self):
        return np.ascontiguousarray(self.com[:, 2])


def lattice_box(n: int, seed: int, jitter: float = 0.3, box: float | None = None) -> WaterBox:
    """Seeded jittered-lattice box at 0.997 g/cm^3 with random orientations."""
    rng = np.random.default_rng(seed)
    L = box_length(n) if box is None else box
    com = jittered_lattice(n, L, rng, jitter)
    q = random_quaternions(n, rng)
    return WaterBox(com, q, L)


def trimer_set(n: int, seed: int, ravg_lo: float = 2.5, ravg_hi: float = 9.0,
               dmin: float = 2.3) -> list[WaterBox]:
    """Isolated water trimers (open boundary) with R_avg ~ U(lo, hi) A and every
    COM-COM separation >= dmin A.  The triangle shape is drawn by rejection:
    two side ratios ~ U(0.7, 1.3) around R_avg with the triangle inequality."""
    rng = np.random.default_rng(seed)
    pcoord = principal_coords()[1]
    out = []
    while len(out) < n:
        ravg = rng.uniform(ravg_lo, ravg_hi)
        s = rng.uniform(0.7, 1.3, size=3)
        s = s / s.mean() * ravg
        a, b, c = s  # |AB|, |AC|, |BC|
        if a + b <= c * 1.02 or a + c <= b * 1.02 or b + c <= a * 1.02 or s.min() < dmin:
            continue
        # place A at origin, B on x, C in the xy-plane, then rotate randomly
        cx = (a * a + b * b - c * c) / (2 * a)
        cy2 = b * b - cx * cx
        if cy2 <= 0:
            continue
        pts = np.array([[0.0, 0.0, 0.0], [a, 0.0, 0.0], [cx, math.sqrt(cy2), 0.0]])
        pts -= pts.mean(0)
        rot = quat_to_rot(random_quaternions(1, rng))[0]
        pts = pts @ rot.T
        q = random_quaternions(3, rng)
        out.append(WaterBox(pts, q, 0.0, pcoord=pcoord))
    return out


def uneven_boxes(nbox: int, seed: int, nmin: int = 64, nmax: int = 1024) -> list[WaterBox]:
    """``nbox`` independent periodic boxes, N log-uniform in [nmin, nmax]."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=nbox)).astype(int)
    return [lattice_box(int(n), seed=int(rng.integers(1 << 31))) for n in ns]
