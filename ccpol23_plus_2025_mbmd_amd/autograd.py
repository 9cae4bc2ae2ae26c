"""Differentiable three-body energy for PyTorch.

``ThreeBodyEnergy`` is a torch.autograd.Function over the library's batched
entry points.  Forward evaluates E_b = qfs2 + qfs3 (kcal/mol) of every box b;
the adjoint of E with respect to the site positions is the analytic force of
the evaluation kernels (site and COM forces, threebody_potential.f90:462-537,
1011-1166), and the adjoint with respect to the linear SR/LR coefficients is
the per-box coefficient gradient dE/dc, which the FS3 kernel's gradient variant
recomputes triplet by triplet (nothing per triplet is kept between calls).
Gradients with respect to orientations follow from the site offsets through
``sites_from_quaternions``, which is an ordinary differentiable torch function.

Units: positions in A, energies in kcal/mol; the library's forces are in
10 J mol^-1 A^-1 (418.4 of them per kcal mol^-1 A^-1).
"""
from __future__ import annotations

import numpy as np
import torch
from torch.autograd.function import once_differentiable

from .threebody import NLIN, ThreeBody

KCAL_PER_FORCE_UNIT = 1.0 / 418.4


def sites_from_quaternions(quat: torch.Tensor, pcoord: torch.Tensor) -> torch.Tensor:
    """Lab-frame site offsets [nm, ns, 3] = R(q) pcoord (molecular_sites.f90:66-111);
    quat [nm, 4] = (q0, q1, q2, q3), pcoord [ns, 3] principal-frame sites."""
    q0, q1, q2, q3 = quat.unbind(-1)
    r = torch.stack([
        torch.stack([q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)], -1),
        torch.stack([2 * (q1 * q2 + q0 * q3), q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * (q2 * q3 - q0 * q1)], -1),
        torch.stack([2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3], -1),
    ], -2)  # [nm, 3, 3]
    return torch.einsum("mab,sb->msa", r, pcoord)


class ThreeBodyEnergy(torch.autograd.Function):
    """E [nbox] = ThreeBodyEnergy.apply(com, site, coef, tb, mol_start, box_len)

    com [nm, 3] and site [nm, ns>=17, 3] (A, float64, cuda); coef [728] the SR
    then LR linear coefficients (cs3_s, cs3_l) used for this call; tb a
    ThreeBody context (its precision decides FP64 or mixed); mol_start [nbox+1]
    and box_len [nbox] (A, <= 0 open) describe the packed boxes.

    The coefficients are loaded into tb, which keeps them for its later calls;
    they are copied to the host (a stream synchronisation) only when coef is
    not the tensor of tb's previous load or has been modified since.  The
    gradients are first derivatives only (the adjoint is the kernels' analytic
    force and dE/dc, not a differentiable graph): differentiating them again
    raises."""

    @staticmethod
    def forward(ctx, com, site, coef, tb: ThreeBody, mol_start, box_len):
        nm = com.shape[0]
        if not (coef is tb._coef_ref and coef._version == tb._coef_version):
            coef_np = coef.detach().to("cpu", torch.float64).numpy()
            tb.set_coefficients(coef_np[:NLIN], coef_np[NLIN:])
            tb._coef_ref, tb._coef_version = coef, coef._version
        com_t = com.detach().t().contiguous()
        ns = site.shape[1]
        site_t = site.detach().permute(2, 0, 1).contiguous()
        want_phi = ctx.needs_input_grad[2]
        if want_phi:
            fsite, f2, f3, en, phi = tb.eval_batch_grad_device(mol_start, box_len, com_t, site_t)
        else:
            fsite, f2, f3, en = tb.eval_batch_device(mol_start, box_len, com_t, site_t)
            phi = None
        ms = np.asarray(mol_start, dtype=np.int64)
        counts = torch.as_tensor(np.diff(ms), device=com.device)
        mol_box = torch.repeat_interleave(torch.arange(len(ms) - 1, device=com.device), counts)
        # dE/dsite = -F_site, dE/dcom = -(sum of site forces + COM forces), kcal/mol/A
        dsite = -(fsite[:, :, :ns].permute(1, 2, 0)) * KCAL_PER_FORCE_UNIT
        dcom = -(fsite.sum(2) + f2 + f3).t() * KCAL_PER_FORCE_UNIT
        ctx.save_for_backward(dcom, dsite, mol_box, phi if phi is not None else torch.empty(0))
        ctx.want_phi = want_phi
        return en[:, 0] + en[:, 1]

    @staticmethod
    @once_differentiable
    def backward(ctx, gE):
        dcom, dsite, mol_box, phi = ctx.saved_tensors
        g = gE[mol_box]
        gcom = dcom * g[:, None] if ctx.needs_input_grad[0] else None
        gsite = dsite * g[:, None, None] if ctx.needs_input_grad[1] else None
        gcoef = gE @ phi if ctx.want_phi else None
        return gcom, gsite, gcoef, None, None, None


def threebody_energy(tb: ThreeBody, com, quat, coef, pcoord, mol_start, box_len):
    """E [nbox] as a differentiable function of COM positions [nm, 3], quaternions
    [nm, 4] and the 728 coefficients; pcoord [ns, 3] principal-frame sites."""
    site = sites_from_quaternions(quat, pcoord)
    return ThreeBodyEnergy.apply(com, site, coef, tb, mol_start, box_len)
