"""torch.autograd wrapper of the discrete adjoint (lbm_adjoint).

``MeanDrag.apply(s, uin, setup)`` returns J = scale * mean over steps [n0, n1] of the drag
F_x for a solid-fraction field ``s`` (nz, ny, nx) and an inlet profile ``uin`` (nz, ny), both
float64 tensors.  The forward pass runs the HIP forward + checkpointed reverse sweep once and
keeps dJ/ds and dJ/du_in for backward, so ``J.backward()`` costs nothing more.  Inlet
parameterisations (e.g. Legendre coefficients) compose with ordinary torch operations.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import lbm as L


@dataclass
class AdjointSetup:
    cfg: L.LbmConfig          # plain mode, FP64 or FP32 storage; the inlet is replaced by uin
    n0: int
    n1: int
    scale: float = 1.0        # e.g. 2/(u0^2 * area) for a drag coefficient
    checkpoint_every: int = 0
    mask: np.ndarray | None = None
    last: dict | None = None  # timings and J of the last call


class MeanDrag(torch.autograd.Function):
    @staticmethod
    def forward(ctx, s: torch.Tensor, uin: torch.Tensor, setup: AdjointSetup):
        cfg = L.LbmConfig.from_buffer_copy(setup.cfg)
        cfg.inlet = L.INLET_PROFILE
        if setup.mask is not None:
            cfg.body = L.BODY_MASK
        with L.Solver(cfg) as sol:
            if setup.mask is not None:
                sol.set_mask(setup.mask)
            sol.set_solid_fraction(s.detach().cpu().double().numpy())
            sol.set_inlet_profile(uin.detach().cpu().double().numpy())
            sol.initialize()
            J, ds, du, tf, ta = sol.adjoint(setup.n0, setup.n1, setup.scale, setup.checkpoint_every)
        setup.last = {"J": J, "ms_forward": tf, "ms_adjoint": ta}
        ctx.save_for_backward(torch.from_numpy(ds).to(s.device, s.dtype), torch.from_numpy(du).to(uin.device, uin.dtype))
        return torch.tensor(J, dtype=s.dtype, device=s.device)

    @staticmethod
    def backward(ctx, gJ):
        ds, du = ctx.saved_tensors
        return gJ * ds, gJ * du, None


def legendre_basis(ny: int, nz: int, ncoef: int = 8, dim: int = 3) -> torch.Tensor:
    """phi_k(y, z) of the parabolic x Legendre inlet (walls half a node outside): (ncoef, nz, ny)."""
    LY = [0, 1, 0, 2, 1, 0, 3, 0]
    LZ = [0, 0, 1, 0, 1, 2, 0, 3]
    ey = (np.arange(ny) + 0.5) / ny
    ez = (np.arange(nz) + 0.5) / nz
    by = 4 * ey * (1 - ey)
    bz = 4 * ez * (1 - ez) if dim == 3 else np.ones(nz)
    P = np.polynomial.legendre.legval
    out = np.zeros((ncoef, nz, ny))
    for k in range(ncoef):
        if dim == 3:
            cy = np.zeros(LY[k] + 1)
            cy[LY[k]] = 1
            cz = np.zeros(LZ[k] + 1)
            cz[LZ[k]] = 1
            out[k] = (bz * P(2 * ez - 1, cz))[:, None] * (by * P(2 * ey - 1, cy))[None, :]
        else:
            c = np.zeros(k + 1)
            c[k] = 1
            out[k] = (by * P(2 * ey - 1, c))[None, :]
    return torch.from_numpy(out)
