"""The hand-written discrete adjoint of the dry step (needs an MI355X).

Checks:
- against the oracle's discrete adjoint (oracle/adjoint.cpp: the oracle's arithmetic recorded on a tape
  and swept backwards on the CPU, the FP64 reference discrete adjoint), entry by entry over the four
  gradients, to 1e-10 of the largest;
- against the oracle's tangent-linear model (oracle/tangent.cpp: the oracle's own arithmetic with dual
  numbers, every upwind sign frozen at the base trajectory as the adjoint freezes it), the FP64 reference
  derivative: <dJ/dx0, d> for seeded random directions, and single gradient entries against tangents of
  basis directions, to 1e-10;
- against central finite differences of the forward model (eps = 1e-6) from a spun-up state, where no
  upwind sign sits at its switch point (in the initial state w' = 0 exactly, and the sign of zero is
  taken as -1, so differences across +-eps straddle that switch at every vertical face);
- the full 192x192x40 case (360 steps, a checkpoint every 20): 16 seeded directions against central
  differences, and the Taylor remainder J(x + eps d) - J(x) - eps <g, d>, which must fall as eps^2;
- the FP32 adjoint against the FP64 one, the torch.autograd.Function wrapper against the plain calls,
  and the checkpoint interval and j-pass form leaving the gradient unchanged.
J is the area-weighted mean relative vorticity in the lowest 1.5 km within 400 km of the centre (linear
in the final u, v).
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


def _spun_up(core, steps=2):
    """The state after `steps` forward steps from the case (w' nonzero away from the lid)."""
    core.step(steps)
    core.sync()
    return [a.copy() for a in final_fields(core)]


def _directions(c, rng, n):
    """Seeded random directions over every point of u, v, w, theta', with w left alone at the lid
    levels 0, 1 and NZ-1 (the boundary update overwrites them, and they are exactly zero)."""
    out = []
    for _ in range(n):
        d = [rng.standard_normal((c.nx, c.ny, c.nz)) for _ in FIELDS]
        d[2][:, :, [0, 1, c.nz - 1]] = 0.0
        out.append(d)
    return out


@pytest.mark.parametrize("spin_up", [0, 2], ids=["initial-state", "spun-up"])
def test_adjoint_matches_the_oracle_tangent(spin_up):
    """The reference discrete derivative: the oracle's tangent-linear model in FP64."""
    from oracle.pyoracle import Tangent
    c = _case()
    core = DynCore(c)
    x0 = _spun_up(core, spin_up) if spin_up else _initial(c)
    Wu, Wv = vorticity_weights(c, radius=400e3, depth=1500.0)
    set_initial(core, x0)
    record(core, 6, 4)
    grads = backward(core, [Wu, Wv, None, None])
    core.close()
    gmax = max(np.abs(g).max() for g in grads)

    def tangent_dJ(dirs):
        t = Tangent(c)
        for f, a, d in zip(FIELDS, x0, dirs):
            for role in (0, 1, 2):
                t.set_field(f, role, a, d)
        t.step(6)
        return float(np.sum(Wu * t.get_field(FIELDS[0], 1)[1]) + np.sum(Wv * t.get_field(FIELDS[1], 1)[1]))

    rng = np.random.Generator(np.random.PCG64(20261207))
    for dirs in _directions(c, rng, 4):
        adj = float(sum(np.sum(g * d) for g, d in zip(grads, dirs)))
        tl = tangent_dJ(dirs)
        print(f"spin-up {spin_up}: adjoint {adj:.15e} tangent {tl:.15e} rel {abs(adj - tl) / abs(tl):.2e}")
        assert abs(adj - tl) <= 1e-10 * abs(tl)
    # single entries of the gradient: basis directions at the largest entries of each field
    for fi in range(4):
        flat = np.argsort(np.abs(grads[fi]).ravel())[-2:]
        for n in flat:
            e = [np.zeros_like(x) for x in x0]
            e[fi].ravel()[n] = 1.0
            tl = tangent_dJ(e)
            got = grads[fi].ravel()[n]
            print(f"spin-up {spin_up}: field {fi} entry {n}: adjoint {got:.15e} tangent {tl:.15e}")
            assert abs(got - tl) <= 1e-10 * gmax


@pytest.mark.parametrize("spin_up", [0, 2], ids=["initial-state", "spun-up"])
def test_adjoint_matches_the_oracle_discrete_adjoint(spin_up):
    """The FP64 reference discrete adjoint (oracle/adjoint.cpp: the oracle's own arithmetic on a tape,
    swept backwards on the CPU), entry by entry over all four gradients: 32x32x12, 3 steps with a
    checkpoint every 2."""
    from oracle.pyoracle import Adjoint
    c = configs.vortex_in_shear(nx=32, ny=32, nz=12, steps=8)
    core = DynCore(c)
    x0 = _spun_up(core, spin_up) if spin_up else _initial(c)
    Wu, Wv = vorticity_weights(c, radius=400e3, depth=1500.0)
    set_initial(core, x0)
    record(core, 3, 2)
    grads = backward(core, [Wu, Wv, None, None])
    core.close()
    ref = Adjoint(c, x0)
    ref.step(3)
    want = ref.gradient([Wu, Wv, None, None])
    ref.close()
    gmax = max(np.abs(g).max() for g in want)
    for f, got, exp in zip(FIELDS, grads, want):
        err = np.abs(got - exp).max() / gmax
        print(f"spin-up {spin_up}: field {f}: max |HIP - oracle adjoint| / max|g| = {err:.2e}")
        assert err <= 1e-10, (f, err)


@pytest.mark.parametrize("nz", [10, 19])
def test_adjoint_level_counts_against_the_oracle_discrete_adjoint(nz):
    """The transposed vertical solves walk the levels in chunks of 8, two chunks a loop trip, then what is
    left: nz - 2 = 8 takes only the partial chunk after no loop trip, 17 only whole loop trips (12, in the
    test above, a full and a partial chunk).  From the initial state, 2 steps with a checkpoint every 2,
    against the FP64 reference discrete adjoint on the CPU."""
    from oracle.pyoracle import Adjoint
    c = configs.vortex_in_shear(nx=32, ny=32, nz=nz, steps=8)
    core = DynCore(c)
    x0 = _initial(c)
    Wu, Wv = vorticity_weights(c, radius=400e3, depth=1500.0)
    set_initial(core, x0)
    record(core, 2, 2)
    grads = backward(core, [Wu, Wv, None, None])
    core.close()
    ref = Adjoint(c, x0)
    ref.step(2)
    want = ref.gradient([Wu, Wv, None, None])
    ref.close()
    gmax = max(np.abs(g).max() for g in want)
    for f, got, exp in zip(FIELDS, grads, want):
        err = np.abs(got - exp).max() / gmax
        print(f"nz {nz}: field {f}: max |HIP - oracle adjoint| / max|g| = {err:.2e}")
        assert err <= 1e-10, (nz, f, err)


@pytest.mark.parametrize("nsteps,every", [(1, 1), (6, 4)], ids=["1-step", "6-steps-every-4"])
def test_adjoint_matches_central_differences(nsteps, every):
    c = _case()
    core = DynCore(c)
    x0 = _spun_up(core)
    Wu, Wv = vorticity_weights(c, radius=400e3, depth=1500.0)
    set_initial(core, x0)
    record(core, nsteps, every)
    grads = backward(core, [Wu, Wv, None, None])
    rng = np.random.Generator(np.random.PCG64(20261206))
    eps = 1e-6
    for dirs in _directions(c, rng, 3):
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


def test_c4_full_size_gradient():
    """The adjoint configuration at full size: 192x192x40, dx = 20 km, dt = 60 s, 360 steps (6 h), a
    checkpoint every 20 steps, FP64; dJ/d(theta', u', v') at t = 0 along 16 seeded directions against
    central differences (eps = 1e-6), and the Taylor remainder of one direction over eps = 1e-2 .. 1e-4.
    Over 360 steps a few upwind signs flip inside +-eps somewhere in the trajectory, and each flip adds a
    first-order kink to the difference quotient, so the per-direction agreement is reported and bounded
    in distribution; the remainder must still fall as eps^2 (flips inside +-eps are themselves
    proportional to eps)."""
    c = configs.vortex_in_shear()
    assert (c.nx, c.ny, c.nz, c.dt, c.steps) == (198, 192, 40, 60.0, 360)
    core = DynCore(c)
    x0 = _initial(c)
    Wu, Wv = vorticity_weights(c, radius=400e3, depth=1500.0)
    set_initial(core, x0)
    assert record(core, c.steps, 20) == 20
    grads = backward(core, [Wu, Wv, None, None])

    def J(dirs, eps):
        set_initial(core, [x + eps * d for x, d in zip(x0, dirs)])
        core.step(c.steps)
        core.sync()
        return _objective(core, Wu, Wv)

    rng = np.random.Generator(np.random.PCG64(20261202 + 4))
    eps = 1e-6
    rels = []
    first = None
    for n in range(16):
        dirs = [rng.standard_normal(x.shape) for x in x0]
        dirs[2][:] = 0.0  # theta', u', v' only
        first = first or dirs
        adj = float(sum(np.sum(g * d) for g, d in zip(grads, dirs)))
        fd = (J(dirs, eps) - J(dirs, -eps)) / (2 * eps)
        rels.append(abs(adj - fd) / abs(fd))
        print(f"C4 direction {n}: adjoint {adj:.12e} central difference {fd:.12e} rel {rels[-1]:.2e}", flush=True)
    rels = np.array(rels)
    print(f"C4: median rel {np.median(rels):.2e}, worst {rels.max():.2e}, within 1e-6: {(rels <= 1e-6).sum()}/16")
    J0 = J(first, 0.0)
    adj = float(sum(np.sum(g * d) for g, d in zip(grads, first)))
    epss = [1e-2, 1e-3, 1e-4]
    rem = [abs(J(first, e) - J0 - e * adj) for e in epss]
    slopes = [np.log10(rem[i] / rem[i + 1]) for i in range(len(rem) - 1)]
    print("C4 Taylor remainders", ["%.3e" % r for r in rem], "slopes", ["%.2f" % s for s in slopes])
    core.close()
    assert np.median(rels) <= 1e-5 and rels.max() <= 1e-3
    assert all(1.8 <= s <= 2.4 for s in slopes)
