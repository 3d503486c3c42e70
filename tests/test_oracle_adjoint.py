"""The oracle's discrete adjoint (oracle/adjoint.cpp), pinned on the CPU.

The adjoint build compiles the oracle's own source with tape-backed values and sweeps the tape
backwards, so it is the transpose of the same discrete map the tangent-linear model differentiates
forwards (oracle/tangent.cpp, itself pinned against central differences in test_oracle_tangent.py).
Here: its values are the oracle bit for bit, <dJ/dx0, d> equals the tangent of J along d for seeded
random directions, and single gradient entries equal the tangents of their basis directions.  The HIP
adjoint is then held to this gradient entry by entry on the GPU (tests/test_gpu_adjoint.py).  J is the
vorticity objective of the adjoint configuration, linear in the final u, v.
"""
import numpy as np
import pytest

from model_amd import configs, ids
from model_amd.adjoint import vorticity_weights
from oracle.pyoracle import Adjoint, Oracle, Tangent

F = (ids.U, ids.V, ids.W, ids.TH)
STEPS = 3


@pytest.fixture(scope="module")
def run():
    c = configs.vortex_in_shear(nx=24, ny=24, nz=12, steps=8)
    x0 = [c.fields["u"], c.fields["v"], c.fields["w"], c.fields["th"]]
    Wu, Wv = vorticity_weights(c, radius=400e3, depth=1500.0)
    a = Adjoint(c, x0)
    a.step(STEPS)
    final = {f: a.get_field(f) for f in F}
    g = a.gradient([Wu, Wv, None, None])
    a.close()
    return c, x0, Wu, Wv, final, g


def _tangent_dJ(c, x0, dirs, Wu, Wv):
    t = Tangent(c)
    for f, a, d in zip(F, x0, dirs):
        for r in (ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P):
            t.set_field(f, r, a, d)
    t.step(STEPS)
    return float(np.sum(Wu * t.get_field(ids.U, ids.ROLE_CUR)[1]) + np.sum(Wv * t.get_field(ids.V, ids.ROLE_CUR)[1]))


def test_adjoint_values_are_the_oracle_bit_for_bit(run):
    c, _, _, _, final, _ = run
    o = Oracle(c)
    o.step(STEPS)
    for f in F:
        assert np.array_equal(final[f], o.get_field(f, ids.ROLE_CUR)), ids.FIELD_NAMES[f]


def test_adjoint_is_the_transpose_of_the_tangent(run):
    c, x0, Wu, Wv, _, g = run
    rng = np.random.Generator(np.random.PCG64(20261207))
    for _ in range(3):
        dirs = [rng.standard_normal(x.shape) for x in x0]
        lhs = sum(float(np.sum(gg * d)) for gg, d in zip(g, dirs))
        rhs = _tangent_dJ(c, x0, dirs, Wu, Wv)
        assert abs(lhs - rhs) <= 1e-11 * abs(rhs), (lhs, rhs)


@pytest.mark.parametrize("field", [0, 1, 3], ids=["u", "v", "theta"])
def test_gradient_entries_against_basis_tangents(run, field):
    c, x0, Wu, Wv, _, g = run
    gmax = max(np.abs(gg).max() for gg in g)
    flat = np.argsort(np.abs(g[field]).ravel())[-2:]
    for idx in flat:
        dirs = [np.zeros_like(x) for x in x0]
        dirs[field].ravel()[idx] = 1.0
        assert abs(g[field].ravel()[idx] - _tangent_dJ(c, x0, dirs, Wu, Wv)) <= 1e-12 * gmax
