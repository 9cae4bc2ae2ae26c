"""Refit of the 728 linear FS3 coefficients to reference energies.

The reference fitted ``cs3_s`` / ``cs3_l`` offline (its design-matrix code is
commented out, threebody_potential.f90:1071,1304,1547,1801) and ships the
result as ``data3b_full*`` tables.  Here the pieces run on the GPU:

* the design matrix Phi[b, n] = dE_b/dc_n of every configuration b (one box
  each), from the FS3 gradient kernels (``ThreeBody.eval_batch_grad_device``);
  E is linear in c, so E_b = Phi[b] . c exactly;
* the normal equations Phi^T W Phi c = Phi^T W y, scaled to unit diagonal and
  lightly regularised, with the Gram matrix on the FP64 matrix cores
  (``ThreeBody.gram``) and a Cholesky solve;
* the fitted coefficients written back as ``data3b_full`` tables
  (``formats.write_data3b``) that the reference and this library both read.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .formats import write_data3b
from .params import NLIN, Data3B, ThreeBodyParams


def design_matrix(tb, mol_start, box_len, com, site):
    """Phi [nbox, 728] (kcal/mol per unit coefficient, SR then LR) and the FS3
    energies [nbox] at the context's current coefficients."""
    _, _, _, en, phi = tb.eval_batch_grad_device(mol_start, box_len, com, site)
    return phi, en[:, 1]


def solve_coefficients(tb, phi, y, weights=None, ridge: float = 1e-13, refine: int = 2):
    """Weighted least squares for c: minimises sum_b w_b (Phi[b] c - y_b)^2.
    The columns of Phi differ in scale by many orders of magnitude, so the
    normal equations are scaled to unit diagonal first (D G D, D = diag(G)^-1/2)
    and regularised there by ``ridge`` (relative to 1); columns no configuration
    touches (coefficients of absent site-type triples) stay 0.  ``refine``
    steps of iterative refinement reuse the Cholesky factor.  Returns
    (c [728] float64 cuda tensor, residual RMS in kcal/mol)."""
    import torch
    dev = phi.device
    y = torch.as_tensor(y, dtype=torch.float64, device=dev)
    sw = None if weights is None else torch.as_tensor(weights, dtype=torch.float64, device=dev).sqrt()
    a = phi if sw is None else phi * sw[:, None]
    yw = y if sw is None else y * sw
    g = tb.gram(a.contiguous())
    dg = torch.diagonal(g)
    d = torch.where(dg > 0, 1.0 / torch.sqrt(dg.clamp_min(1e-300)), torch.zeros_like(dg))
    gs = d[:, None] * g * d[None, :] + ridge * torch.eye(g.shape[0], dtype=g.dtype, device=dev)
    chol = torch.linalg.cholesky(gs)
    c = d * torch.cholesky_solve((d * (a.T @ yw))[:, None], chol)[:, 0]
    # iterative refinement on the weighted residual (the same factor)
    for _ in range(refine):
        r = yw - a @ c
        c = c + d * torch.cholesky_solve((d * (a.T @ r))[:, None], chol)[:, 0]
    res = float(torch.sqrt(torch.mean((phi @ c - y) ** 2)))
    return c, res


def fitted_params(p: ThreeBodyParams, c) -> ThreeBodyParams:
    """The parameter set p with its SR and LR linear coefficients replaced by c [728]."""
    c = np.asarray(c.detach().cpu() if hasattr(c, "detach") else c, dtype=np.float64)
    if c.shape != (2 * NLIN,):
        raise ValueError(f"expected {2 * NLIN} coefficients, got {c.shape}")
    sr = Data3B(beta=p.sr.beta, gamma=p.sr.gamma, r0=p.sr.r0, cs=c[:NLIN].copy())
    lr = Data3B(beta=p.lr.beta, gamma=p.lr.gamma, r0=p.lr.r0, cs=c[NLIN:].copy())
    return p.replace(sr=sr, lr=lr)


def write_fit(directory, p: ThreeBodyParams, c) -> tuple[Path, Path]:
    """Write the fitted tables as ``data3b_full_SR_5p5`` and ``data3b_full_LR_4p5``
    in ``directory`` (the names the reference README stages)."""
    q = fitted_params(p, c)
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    sr, lr = d / "data3b_full_SR_5p5", d / "data3b_full_LR_4p5"
    write_data3b(sr, q.sr)
    write_data3b(lr, q.lr)
    return sr, lr
