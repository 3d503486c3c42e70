"""Batched ensembles (needs an MI355X): members of different horizontal sizes advanced by one launch per
pass must give, member by member, the same bits as each member stepped alone, and at the ensemble
configuration's own sizes each FP64 member must match the oracle."""
import numpy as np
import pytest

from model_amd import configs, ids
from model_amd.core import DYN_FP32, DynCore, Ensemble

pytestmark = pytest.mark.gpu

FIELDS = (ids.U, ids.V, ids.W, ids.TH)


def _cases():
    return [configs.vortex_in_shear(nx=n, ny=m, nz=20, seed_offset=10 + k) for k, (n, m) in enumerate([(48, 64), (64, 48), (80, 80)])]


@pytest.mark.parametrize("groups", [1, 2, 3], ids=["1-group", "2-groups", "3-groups"])
@pytest.mark.parametrize("precision", [DYN_FP32, 0], ids=["fp32", "fp64"])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_ensemble_matches_members_alone(precision, graph, groups):
    cases = _cases()
    alone = [DynCore(c, precision=precision) for c in cases]
    for g in alone:
        g.step(5)
        g.sync()
    members = [DynCore(c, precision=precision) for c in cases]
    ens = Ensemble(members)
    assert ens.groups == min(4, len(members))
    ens.set_groups(groups)
    assert ens.groups == groups
    ens.step(5, graph=graph)
    ens.sync()
    for a, b in zip(alone, members):
        for f in FIELDS:
            for role in (ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P):
                x, y = a.get_field(f, role), b.get_field(f, role)
                assert np.array_equal(x, y), (f, role, np.abs(x - y).max())
    ens.close()


@pytest.mark.parametrize("precision", [DYN_FP32, 0], ids=["fp32", "fp64"])
def test_ensemble_with_a_prime_length_member(precision):
    """A member whose i transform length has a prime factor above 5 (ni = 46 = 2 x 23, nj = 38 = 2 x 19)
    cannot join the batched pressure solve, so the ensemble solves each member's pressure on its own and
    batches the rest of the step: every member still gives the bits of its solo run."""
    sizes = [(48, 64), (46, 38), (64, 48)]
    cases = [configs.vortex_in_shear(nx=n, ny=m, nz=20, seed_offset=30 + k) for k, (n, m) in enumerate(sizes)]
    alone = [DynCore(c, precision=precision) for c in cases]
    for g in alone:
        g.step(4)
        g.sync()
    members = [DynCore(c, precision=precision) for c in cases]
    ens = Ensemble(members)
    ens.step(3)
    ens.step(1, graph=False)
    ens.sync()
    for a, b in zip(alone, members):
        for f in FIELDS:
            for role in (ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P):
                x, y = a.get_field(f, role), b.get_field(f, role)
                assert np.array_equal(x, y), (f, role, np.abs(x - y).max())
    ens.close()


def test_ensemble_autotune_keeps_state_and_values():
    cases = _cases()
    alone = [DynCore(c, precision=DYN_FP32) for c in cases]
    for g in alone:
        g.step(4)
        g.sync()
    members = [DynCore(c, precision=DYN_FP32) for c in cases]
    ens = Ensemble(members)
    shape = ens.autotune()
    assert shape[0] in (2, 4, 8) and shape[2] in (2, 4, 8)
    ens.step(4)
    ens.sync()
    for a, b in zip(alone, members):
        for f in FIELDS:
            assert np.array_equal(a.get_field(f, ids.ROLE_CUR), b.get_field(f, ids.ROLE_CUR))
    ens.close()


def test_c3_members_against_the_oracle():
    """Four members of the ensemble configuration (horizontal sizes from {96,128,160,192}, nz = 40,
    dx = 20 km, dry vortex in shear, seeded noise per member), batched in FP64 for 5 steps: every member
    matches the oracle to 1e-11 in every prognostic; the same batch in FP32 reports its per-field error
    against the FP64 batch."""
    from oracle.pyoracle import Oracle
    sizes = [(96, 128), (128, 96), (160, 192), (192, 160)]
    cases = [configs.vortex_in_shear(nx=n, ny=m, nz=40, seed_offset=100 + k) for k, (n, m) in enumerate(sizes)]
    members = [DynCore(c) for c in cases]
    ens = Ensemble(members)
    ens.step(5)
    ens.sync()
    for c, g in zip(cases, members):
        o = Oracle(c, variant="omp")
        o.step(5)
        for f in FIELDS:
            ref = o.get_field(f, ids.ROLE_CUR)
            e = np.abs(g.get_field(f, ids.ROLE_CUR) - ref).max() / np.abs(ref).max()
            print(f"member {c.nx - 6}x{c.ny}: {ids.FIELD_NAMES[f]} rel {e:.2e}")
            assert e <= 1e-11
    members32 = [DynCore(c, precision=DYN_FP32) for c in cases]
    ens32 = Ensemble(members32)
    ens32.step(5)
    ens32.sync()
    for c, a, b in zip(cases, members, members32):
        row = {}
        for f in FIELDS:
            ref = a.get_field(f, ids.ROLE_CUR)[3:-3, :, 1:-1]
            got = b.get_field(f, ids.ROLE_CUR)[3:-3, :, 1:-1]
            row[ids.FIELD_NAMES[f]] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        print(f"FP32 member {c.nx - 6}x{c.ny}: " + " ".join(f"{k} {v:.2e}" for k, v in row.items()))
        assert row["th"] < 1e-5 and row["w"] < 1e-3
    ens.close()
    ens32.close()
