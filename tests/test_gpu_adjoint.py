"""The hand-written discrete adjoint of the dry step (needs an MI355X).

Checks: directional derivatives <dJ/dx0, d> against central finite differences of the forward model
(eps = 1e-6, FP64) on a reduced vortex-in-shear grid, for seeded random directions over every point of
u, v, w, theta'; the FP32 adjoint against the FP64 one; the torch.autograd.Function wrapper against the
plain calls.  J is the area-weighted mean relative vorticity in the lowest 1.5 km within 400 km of the
centre (linear in the final u, v).
"""
import numpy as np
import pytest

from model_amd import configs
from model_amd.adjoint import FIELDS, backward, final_fields, record, set_initial, vorticity_weights
from model_amd.core import DYN_FP32, DynCore

pytestmark = pytest.mark.gpu


def _case():
    return configs.vortex_in_shear(nx=48, ny=48, nz=20, steps=8)


def _initial(c):
    return [c.fields["u"], c.fields["v"], c.fields["w"], c.fields["th"]]


def _objective(core, Wu, Wv):
    u, v, _, _ = final_fields(core)
    return float(np.sum(Wu * u) + np.sum(Wv * v))


@pytest.mark.parametrize("nsteps,every", [(1, 1), (6, 4)], ids=["1-step", "6-steps-every-4"])
def test_adjoint_matches_central_differences(nsteps, every):
    c = _case()
    core = DynCore(c)
    x0 = _initial(c)
    Wu, Wv = vorticity_weights(c, radius=400e3, depth=1500.0)
    set_initial(core, x0)
    record(core, nsteps, every)
    grads = backward(core, [Wu, Wv, None, None])
    rng = np.random.Generator(np.random.PCG64(20261206))
    eps = 1e-6
    for _ in range(3):
        dirs = [rng.standard_normal(x.shape) for x in x0]
        adj = float(sum(np.sum(g * d) for g, d in zip(grads, dirs)))

        def J(sign):
            set_initial(core, [x + sign * eps * d for x, d in zip(x0, dirs)])
            core.step(nsteps)
            core.sync()
            return _objective(core, Wu, Wv)

        fd = (J(+1) - J(-1)) / (2 * eps)
        print(f"nsteps={nsteps}: adjoint {adj:.12e} central difference {fd:.12e} rel {abs(adj - fd) / abs(fd):.2e}")
        assert abs(adj - fd) <= 1e-6 * abs(fd)
    core.close()


def test_fp32_adjoint_against_fp64_adjoint():
    c = _case()
    Wu, Wv = vorticity_weights(c)
    out = {}
    for prec in (0, DYN_FP32):
        core = DynCore(c, precision=prec)
        record(core, 4, 2)
        out[prec] = backward(core, [Wu, Wv, None, None])
        core.close()
    for f, a, b in zip(FIELDS, out[0], out[DYN_FP32]):
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-300)
        print("fp32 vs fp64 adjoint", f, rel)
        assert rel < 1e-3


def test_torch_autograd_wrapper():
    torch = pytest.importorskip("torch")
    from model_amd.adjoint import Trajectory
    c = _case()
    core = DynCore(c)
    Wu, Wv = vorticity_weights(c)
    x0 = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in _initial(c)]
    u, v, w, th = Trajectory.apply(core, 3, 2, *x0)
    J = (torch.from_numpy(Wu) * u).sum() + (torch.from_numpy(Wv) * v).sum()
    J.backward()
    core2 = DynCore(c)
    record(core2, 3, 2)
    ref = backward(core2, [Wu, Wv, None, None])
    for t, r in zip(x0, ref):
        assert np.array_equal(t.grad.numpy(), r)


def test_checkpoint_interval_does_not_change_the_gradient():
    """Every checkpoint interval (recomputed segments, the record pass's own last segment, the whole
    trajectory kept, the automatic interval) gives the same gradient bit for bit, also on a second
    backward sweep over the same record."""
    c = _case()
    Wu, Wv = vorticity_weights(c)
    core = DynCore(c)
    x0 = _initial(c)
    ref = None
    for every in (1, 4, 6, 0):
        set_initial(core, x0)
        used = record(core, 6, every)
        assert used == (6 if every == 0 else every)
        for sweep in range(2):
            grads = backward(core, [Wu, Wv, None, None])
            if ref is None:
                ref = grads
            for f, a, b in zip(FIELDS, ref, grads):
                assert np.array_equal(a, b), (every, sweep, f)
    core.close()


def test_adjoint_with_half_slab_j_pass():
    """The transposed pressure solve through the half-slab j pass gives the gradient of the one-slab j
    pass to rounding (1e-12 relative), FP64. The half slabs need a power-of-two half length: ny = 64
    (32 = 8 x 4, which takes the radix-8 pass)."""
    from model_amd import core
    c = configs.vortex_in_shear(nx=48, ny=64, nz=20, steps=8)
    Wu, Wv = vorticity_weights(c)
    out = {}
    for jpass in (1, 3):
        g = DynCore(c)
        g.tune_set(core.DYN_TUNE_PRESSURE_JPASS, jpass)
        record(g, 3, 2)
        out[jpass] = backward(g, [Wu, Wv, None, None])
        g.close()
    for f, a, b in zip(FIELDS, out[1], out[3]):
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-300)
        assert rel < 1e-12, (f, rel)
