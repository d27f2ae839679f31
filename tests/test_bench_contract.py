"""bench.py's derived figures and its JSON contract pieces that need no GPU:
the roofline arithmetic, the sort/build HBM fractions, the traffic figures it
reads from the committed rocprofv3 summaries, and the one-GPU guard."""
import subprocess
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO))
import bench  # noqa: E402


def _leg(**kw):
    leg = {"far_field": "monopole", "interactions": 700 * 2**20,
           "interactions_per_step_timed": 7.5e8, "tree_nodes": 3_000_000,
           "internal_nodes": 750_000,
           "phase_ms_per_step_eager": {"sort": 0.25, "build": 0.4, "force": 0.6}}
    leg.update(kw)
    return leg


def test_roofline_uses_the_timed_launches():
    leg = _leg()
    r = bench.roofline_walk(leg, "group")
    # 7.5e8 interactions x 14 flops in 0.6 ms
    assert r["achieved"] == pytest.approx(7.5e8 * 14 / 0.6e-3 / 1e12)
    assert r["frac"] == pytest.approx(r["achieved"] / 157.3)
    assert r["peak"] == 157.3 and r["unit"] == "TFLOP/s"
    assert r["interactions_per_launch"] == 7.5e8 and r["avg_launch_ms"] == 0.6


def test_roofline_reports_the_per_body_test_equivalent():
    """With the EXACT traversal's interaction count on the same state, the
    roofline also gives the rate of the interactions the per-body test keeps."""
    leg = _leg(interactions_per_body=720.0, interactions_per_body_exact=380.0)
    r = bench.roofline_walk(leg, "group")
    assert r["exact_over_group_interactions"] == pytest.approx(380 / 720)
    assert r["achieved_exact_equivalent"] == pytest.approx(
        7.5e8 * 380 / 720 * 14 / 0.6e-3 / 1e12)
    assert "exact_over_group_interactions" not in bench.roofline_walk(_leg(), "group")


def test_learned_roofline_counts_near_cell_quadrupole_terms():
    """The learned walk's flops: 14 per interaction plus 22 for each term that
    also adds a near cell's quadrupole term at the body."""
    r = bench.roofline_walk(_leg(far_field="learned", quad_terms_per_step_timed=2e8), "group")
    assert r["achieved"] == pytest.approx((7.5e8 * 14 + 2e8 * 22) / 0.6e-3 / 1e12)
    assert r["quad_terms_per_launch"] == 2e8


def test_roofline_traffic_comes_from_the_profiles_per_leg():
    """Each leg's traffic is the committed counter summary's line for its own
    walk kernel: the plain walk has one; the learned walk and the far-field
    apply have none for the current code (their counter passes were not
    taken), so their traffic is null rather than another kernel's."""
    plain = bench.roofline_walk(_leg(), "group")["traffic"]
    learned = bench.roofline_walk(_leg(far_field="learned"), "group")["traffic"]
    assert plain is not None and 1e8 < plain < 1e9
    assert learned is None
    assert bench.traffic_from_profile("k_mlp_apply") is None
    assert bench.traffic_from_profile("no_such_kernel") is None


def test_sort_build_hbm_fractions():
    n = 1 << 20
    h = bench.hbm_sort_build(_leg(), n)
    assert h["sort"]["bytes"] == (8 * 24 + 92) * n
    # one 64-byte record per non-empty slot: at most internal nodes + bodies
    assert h["build"]["bytes"] == 32 * n + 64 * (750_000 + n)
    assert h["sort"]["achieved_tb_s"] == pytest.approx(h["sort"]["bytes"] / 0.25e-3 / 1e12)
    assert h["build"]["frac"] == pytest.approx(h["build"]["achieved_tb_s"] / 8.0)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_equal_error_interpolation():
    pts = [{"theta": 0.5, "rms": 1.4e-2, "ms_per_step": 1.4},
           {"theta": 0.4, "rms": 9.0e-3, "ms_per_step": 1.7},
           {"theta": 0.3, "rms": 5.0e-3, "ms_per_step": 2.1}]
    learned = {"force_error": {"rms": 9.0e-3}, "ms_per_step": 1.7}
    e = bench.equal_error(pts, learned)
    assert e["plain_theta"] == pytest.approx(0.4) and e["plain_ms_per_step"] == pytest.approx(1.7)
    assert e["learned_over_plain"] == pytest.approx(1.0)
    # halfway in log error between 0.3 and 0.4
    learned = {"force_error": {"rms": (5.0e-3 * 9.0e-3) ** 0.5}, "ms_per_step": 1.6}
    e = bench.equal_error(pts, learned)
    assert e["plain_theta"] == pytest.approx(0.35)
    assert e["plain_ms_per_step"] == pytest.approx(1.9)
    assert e["learned_over_plain"] == pytest.approx(1.9 / 1.6)
    # outside the measured range: no claim
    e = bench.equal_error(pts, {"force_error": {"rms": 1e-3}, "ms_per_step": 1.0})
    assert e["learned_over_plain"] is None


def test_cpu_baseline_helpers():
    """The CPU figures name the host and carry their units; the all-pairs rate
    is that of the reference's single-threaded step on a bounded sample."""
    h = bench.host_cpu()
    assert h["host_logical_cpus"] >= 1 and h["allowed_cpus"] >= 1
    assert 1 <= bench.cpu_threads() <= 256
    sys.path.insert(0, str(REPO / "n-body-2d_amd"))
    from nbody2d import ics

    m, p, v = ics.uniform_disk(512, seed=2)
    rate, k, secs = bench.cpu_allpairs_rate(m, p, v, n_sample=256)
    assert k == 256 and secs > 0 and rate == pytest.approx(256 * 255 / secs)
