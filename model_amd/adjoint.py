"""Gradients through the hand-written discrete adjoint of the dry step (dyn_adjoint_record /
dyn_adjoint_backward), as plain functions and as a torch.autograd.Function.

The differentiated map takes the initial u, v, w, theta' (each set into all three buffers of its
prognostic, as the model's initialisation does) to the current-buffer fields after nsteps steps.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import ids
from .core import DynCore, _check

FIELDS = (ids.U, ids.V, ids.W, ids.TH)
_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def set_initial(core: DynCore, fields) -> None:
    """Set u, v, w, theta' (arrays [NX, NY, NZ]) into all three buffers of each prognostic."""
    for f, a in zip(FIELDS, fields):
        a = np.ascontiguousarray(a, dtype=np.float64)
        for role in (ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P):
            core.set_field(f, role, a)


def final_fields(core: DynCore):
    return tuple(core.get_field(f, ids.ROLE_CUR) for f in FIELDS)


def record(core: DynCore, nsteps: int, every: int) -> int:
    """Run nsteps from the context's state, keeping snapshots every `every` steps (0: the longest
    interval that fits in half of the free device memory).  Returns the interval in use."""
    _check(core.L.dyn_adjoint_record(core.h, nsteps, every))
    n = core.L.dyn_adjoint_interval(core.h)
    _check(min(n, 0))
    return n


def backward(core: DynCore, seeds):
    """dJ/d(initial u, v, w, theta') from dJ/d(final current u, v, w, theta') (None = zero)."""
    shape = (core.case.nx, core.case.ny, core.case.nz)
    seeds = [None if s is None else np.ascontiguousarray(s, dtype=np.float64).reshape(shape) for s in seeds]
    grads = [np.zeros(shape) for _ in FIELDS]
    _check(core.L.dyn_adjoint_backward(core.h, *[_ptr(s) for s in seeds], *[_ptr(g) for g in grads]))
    return grads


def vorticity_weights(case, radius=400e3, depth=1500.0):
    """Weights (Wu, Wv) with J = sum(Wu * u) + sum(Wv * v) = the area-weighted mean relative vorticity
    (dv/dx - du/dy at the cell corners) over the levels below `depth` within `radius` of the domain
    centre, on the interior columns.  J is linear in (u, v), so these are dJ/du and dJ/dv."""
    NX, NY, NZ, dx = case.nx, case.ny, case.nz, case.dx
    zsu = case.columns["zsu"]
    xc = 0.5 * (NX - 6 - 1) * dx
    yc = 0.5 * (NY - 1) * dx
    Wu = np.zeros((NX, NY, NZ))
    Wv = np.zeros((NX, NY, NZ))
    ii, jj = np.meshgrid(np.arange(4, NX - 3), np.arange(1, NY - 1), indexing="ij")
    xcorner = (ii - 3 - 0.5) * dx
    ycorner = (jj - 0.5) * dx
    inside = np.hypot(xcorner - xc, ycorner - yc) <= radius
    levels = [k for k in range(1, NZ - 1) if zsu[k] < depth]
    count = inside.sum() * len(levels)
    if count == 0:
        raise ValueError("empty objective region")
    w = 1.0 / (count * dx)
    for k in levels:
        for i, j in zip(ii[inside], jj[inside]):
            # zeta(i-1/2, j-1/2) = (v(i,j) - v(i-1,j)) / dx - (u(i,j) - u(i,j-1)) / dy
            Wv[i, j, k] += w
            Wv[i - 1, j, k] -= w
            Wu[i, j, k] -= w
            Wu[i, j - 1, k] += w
    return Wu, Wv


try:
    import torch

    class Trajectory(torch.autograd.Function):
        """(u0, v0, w0, th0) -> final (u, v, w, th) of `nsteps` dry steps on `core` (float64 CPU tensors).
        backward() runs the hand-written adjoint over the trajectory recorded by forward()."""

        @staticmethod
        def forward(ctx, core, nsteps, every, u0, v0, w0, th0):
            set_initial(core, [t.detach().cpu().numpy() for t in (u0, v0, w0, th0)])
            record(core, nsteps, every)
            ctx.core = core
            return tuple(torch.from_numpy(a) for a in final_fields(core))

        @staticmethod
        def backward(ctx, gu, gv, gw, gth):
            seeds = [None if g is None else g.detach().cpu().numpy() for g in (gu, gv, gw, gth)]
            grads = backward(ctx.core, seeds)
            return (None, None, None) + tuple(torch.from_numpy(g) for g in grads)

except ImportError:  # torch is only needed for the autograd wrapper
    Trajectory = None
