"""bench.py's host logic on the CPU: the one-GPU refusal, the counter-summary lookup behind
roofline.traffic (used only for the kernel sources it was measured on), kernel-name parsing and the
roofline arithmetic."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_kernel_base_name():
    assert bench.kernel_base_name("void dyn::(anonymous namespace)::k_adv_scalars<float, 8, true, 1>(dyn::Geo, float)") == \
        "k_adv_scalars"
    assert bench.kernel_base_name("dyn::(anonymous namespace)::k_copy_fields(dyn::FieldCopies, long long)") == "k_copy_fields"


def _summary(tmp_path, digest, kernels):
    path = tmp_path / "traffic.json"
    path.write_text(json.dumps({"csrc_sha256": digest, "kernels": kernels}))
    return str(path)


@pytest.mark.parametrize("record", sorted(bench.KERNEL_SYMBOL))
def test_traffic_lookup_for_every_timed_kernel(tmp_path, monkeypatch, record):
    sym = f"void dyn::(anonymous namespace)::{bench.KERNEL_SYMBOL[record]}<float, 8, true, 1>(dyn::Geo)"
    monkeypatch.setattr(bench, "TRAFFIC_PROFILE", _summary(tmp_path, bench.csrc_digest(), {sym: {"bytes": 123456}}))
    assert bench.measured_traffic(record) == (123456, "measured on these sources")


def test_stale_or_ambiguous_summary_gives_no_traffic(tmp_path, monkeypatch):
    sym = "void dyn::(anonymous namespace)::k_adv_scalars<float, 8, true, 1>(dyn::Geo)"
    monkeypatch.setattr(bench, "TRAFFIC_PROFILE", _summary(tmp_path, "0" * 64, {sym: {"bytes": 1}}))
    b, status = bench.measured_traffic("adv_scalars")
    assert b is None and status.startswith("stale")
    two = {sym: {"bytes": 1}, sym.replace("8, true", "4, true"): {"bytes": 2}}
    monkeypatch.setattr(bench, "TRAFFIC_PROFILE", _summary(tmp_path, bench.csrc_digest(), two))
    assert bench.measured_traffic("adv_scalars")[0] is None


def test_roofline_arithmetic(monkeypatch):
    monkeypatch.setattr(bench, "measured_traffic", lambda name: (None, "none"))
    # 3 launches of 1e6 points each in 0.3 ms: 21 passes x 4 B x 1e6 / 0.1 ms = 840 GB/s
    roof = bench.roofline([("adv_scalars", 3, 0.3, 3e6), ("post_uvw", 3, 0.1, 3e6)], 4, 1)
    assert roof["kernel"] == "adv_scalars" and roof["field_passes_per_point"] == 21
    assert roof["achieved"] == pytest.approx(840.0) and roof["frac"] == pytest.approx(0.105)
