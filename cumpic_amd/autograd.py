"""torch.autograd binding of the reverse-mode step (cumc_tape_run / cumc_tape_backward).

``Rollout.apply(v0, solver, problem, nsteps, every, dts)`` maps the interior initial velocity
field ``v0`` (shape (nz, ny, nx, 3), float64) of ``problem`` to the interior final primitive
state (nz, ny, nx, 8) and cell-centred field (nz, ny, nx, 3) after initial_update and ``nsteps``
RUNGEKUTTA steps on the GPU, with the time steps frozen to ``dts`` (the CFL choice is a max over
cells and is not differentiated).  The backward pass runs the hand-written adjoint kernels with
checkpoint-every-``every`` recomputation.  Tensors cross the C ABI as host arrays.
"""
from __future__ import annotations

import numpy as np
import torch

from .ics import ix


def _interior(cfg):
    return (slice(ix(1), ix(cfg.nz) + 1), slice(ix(1), ix(cfg.ny) + 1), slice(ix(1), ix(cfg.nx) + 1))


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, solver, problem, nsteps, every, dts):
        cfg = problem.cfg
        inner = _interior(cfg)
        prim = problem.prim.copy()
        prim[inner + (slice(1, 4),)] = v0.detach().to("cpu", torch.float64).numpy()
        solver.upload(prim, problem.eps)
        used = solver.tape_run(int(nsteps), int(every), None if dts is None else np.asarray(dts, dtype=np.float64))
        st = solver.download()
        ctx.solver, ctx.cfg, ctx.device = solver, cfg, v0.device
        ctx.dts = used
        p = torch.from_numpy(np.ascontiguousarray(st["prim"][inner])).to(v0.device)
        b = torch.from_numpy(np.ascontiguousarray(st["bcell"][inner])).to(v0.device)
        return p, b

    @staticmethod
    def backward(ctx, g_prim, g_bcell):
        cfg = ctx.cfg
        inner = _interior(cfg)
        sp = np.zeros(cfg.shape + (8,))
        sb = np.zeros(cfg.shape + (3,))
        if g_prim is not None:
            sp[inner] = g_prim.detach().to("cpu", torch.float64).numpy()
        if g_bcell is not None:
            sb[inner] = g_bcell.detach().to("cpu", torch.float64).numpy()
        gp, _ = ctx.solver.tape_backward(seed_prim=sp, seed_bcell=sb)
        gv = torch.from_numpy(np.ascontiguousarray(gp[inner + (slice(1, 4),)])).to(ctx.device)
        return gv, None, None, None, None, None


def cell_volume(problem) -> float:
    """dV of the uniform Cartesian grid of problem."""
    return float((problem.xf[4] - problem.xf[3]) * (problem.yf[4] - problem.yf[3]) * (problem.zf[4] - problem.zf[3]))


def magnetic_energy(bcell_interior: torch.Tensor, problem) -> torch.Tensor:
    """J = sum over interior cells of |B|^2 / 2 * dV."""
    return 0.5 * (bcell_interior * bcell_interior).sum() * cell_volume(problem)


def initial_velocity(problem) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(problem.prim[_interior(problem.cfg) + (slice(1, 4),)]))
