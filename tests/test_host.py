"""C++ host adapter: universe CSV, config.json, epsilon, the Simulation state
machine and the factory (driven through the headless driver nbody2d-run).

CPU tests pin the loader/writer on the reference's own universe files
(tests/golden/*.csv, copied from universe-db/) and epsilon's exact branch on
the oracle; GPU tests run whole simulations from a config file."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import nbody2d
import oracle_lib as O
from nbody2d import ics

REPO = Path(__file__).resolve().parents[1]
RUN = REPO / "n-body-2d_amd" / "nbody2d" / "nbody2d-run"
GOLDEN = REPO / "tests" / "golden"


def run(*args, check=True):
    r = subprocess.run([str(RUN), *map(str, args)], capture_output=True, text=True, timeout=600)
    if check and r.returncode != 0:
        raise AssertionError(f"nbody2d-run failed: {r.stderr}")
    return r


def read_csv(path):
    lines = Path(path).read_text().splitlines()
    assert lines[0] == "id,mass,x,y,vel_x,vel_y"
    rows = [ln.split(",") for ln in lines[1:]]
    ids = [r[0] for r in rows]
    vals = np.array([[float(x) for x in r[1:]] for r in rows]) if rows else np.zeros((0, 5))
    return ids, vals


def write_csv(path, mass, pos, vel, ids=None):
    with open(path, "w") as f:
        f.write("id,mass,x,y,vel_x,vel_y\n")
        for i in range(len(mass)):
            name = ids[i] if ids else f"b{i}"
            vals = (mass[i], pos[i, 0], pos[i, 1], vel[i, 0], vel[i, 1])
            f.write(name + "," + ",".join(repr(float(x)) for x in vals) + "\n")


def config(tmp, infile, outfile, **sim):
    s = {"timestep": 1e-3, "iterations": 10, "algorithm": "barnes-hut gpu", "theta": 0.5,
         "softening_factor": 0.0, "threads": 4}
    s.update(sim)
    cfg = {"IO": {"universe_infile": str(infile), "universe_outfile": str(outfile),
                  "echo_config": False, "echo_bodies": False},
           "Simulation": s,
           "Graphics": {"enabled": False, "resolution": [1280, 720], "fps": 60, "vsync": False,
                        "pixel_scale": 1.0, "show_grid": False, "show_commands_panel": False,
                        "show_config_panel": False, "show_stats_panel": False,
                        "panel_update_hz": 1}}
    p = Path(tmp) / "config.json"
    p.write_text(json.dumps(cfg))
    return p


def test_csv_reference_unit_file_stod_prefix_rule(tmp_path):
    out = tmp_path / "u.csv"
    run("--echo-csv", GOLDEN / "universe-unit-test.csv", out)
    ids, v = read_csv(out)
    assert ids == ["BodyAlpha", "BodyBeta", "BodyGamma"]
    assert v[0].tolist() == [100.0, 15.0, -5.0, 30.0, 2000.0]
    assert v[1].tolist() == [6.5e-23, 2.96e-14, -5e-15, 6.2e-5, 2.97]  # "2.97-e4" -> 2.97
    assert v[2].tolist() == [6.5e23, 2.96e14, -5e15, 6.2e5, 2.97e4]


def test_csv_round_trip_is_lossless(tmp_path):
    rng = np.random.default_rng(0)
    m = rng.lognormal(0, 3, 200)
    p = rng.normal(0, 1e5, (200, 2))
    v = rng.normal(0, 1e-3, (200, 2))
    src = tmp_path / "a.csv"
    write_csv(src, m, p, v)
    out = tmp_path / "b.csv"
    run("--echo-csv", src, out)
    ids, vals = read_csv(out)
    assert np.array_equal(vals[:, 0], m)
    assert np.array_equal(vals[:, 1:3], p) and np.array_equal(vals[:, 3:5], v)
    run("--echo-csv", out, tmp_path / "c.csv")
    assert (tmp_path / "c.csv").read_text() == out.read_text()


def test_csv_duplicate_ids_and_bad_numbers_are_errors(tmp_path):
    bad = tmp_path / "dup.csv"
    bad.write_text("id,mass,x,y,vel_x,vel_y\na,1,0,0,0,0\na,1,1,1,0,0\n")
    r = run("--echo-csv", bad, tmp_path / "o.csv", check=False)
    assert r.returncode == 1 and "duplicate ID" in r.stderr
    bad.write_text("id,mass,x,y,vel_x,vel_y\na,heavy,0,0,0,0\n")
    r = run("--echo-csv", bad, tmp_path / "o.csv", check=False)
    assert r.returncode == 1 and "cannot parse number" in r.stderr


def test_softening_exact_branch_matches_oracle():
    ids, vals = read_csv(GOLDEN / "solar-system-plus.csv")
    eps = json.loads(run("--softening", GOLDEN / "solar-system-plus.csv", 0.01).stdout)["epsilon"]
    pos = np.ascontiguousarray(vals[:, 1:3])
    ref = O.load().or_plummer_softening_exact(len(pos), O.ptr(pos), 0.01, 1e6)
    assert eps == pytest.approx(ref, rel=1e-14)


def test_config_check_and_validation(tmp_path):
    src = tmp_path / "u.csv"
    m, p, v = ics.exp_disk(100, seed=1)
    write_csv(src, m, p, v)
    cfg = config(tmp_path, src, tmp_path / "out.csv", iterations=1e7, algorithm="Barnes-Hut GPU")
    info = json.loads(run(cfg, "--check").stdout)
    assert info == {"bodies": 100, "iterations": 10000000, "algorithm": "Barnes-Hut GPU"}
    for bad, msg in [({"timestep": 0.0}, "timestep"), ({"algorithm": "naive"}, "algorithm"),
                     ({"theta": 101}, "theta"), ({"softening_factor": 0.3}, "softening_factor"),
                     ({"threads": 0}, "threads"), ({"traversal": "fast"}, "traversal")]:
        cfg = config(tmp_path, src, tmp_path / "out.csv", **bad)
        r = run(cfg, "--check", check=False)
        assert r.returncode == 1 and msg in r.stderr, (bad, r.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("algorithm,traversal", [("barnes-hut gpu", "exact"),
                                                 ("Barnes-Hut", "group"),
                                                 ("all pairs", "exact")])
def test_driver_matches_library(tmp_path, algorithm, traversal):
    m, p, v = ics.exp_disk(3000, seed=2)
    src, out = tmp_path / "in.csv", tmp_path / "out.csv"
    write_csv(src, m, p, v)
    cfg = config(tmp_path, src, out, iterations=7, algorithm=algorithm, G=1.0, softening=0.01,
                 theta=0.6, traversal=traversal, timestep=2e-3)
    print(f"[gpu] driver {algorithm} {traversal}", flush=True)
    stats = json.loads(run(cfg).stdout)
    assert stats["iterations"] == 7 and stats["epsilon"] == 0.01
    ids, vals = read_csv(out)
    ap = algorithm == "all pairs"
    A = nbody2d.abi
    kw = dict(G=1.0, dt=2e-3, theta=0.6, softening=0.01,
              traversal=A.TRAV_GROUP if traversal == "group" else A.TRAV_EXACT,
              algorithm=A.ALGO_ALL_PAIRS if ap else A.ALGO_BARNES_HUT,
              integrator=A.INTEG_DK if ap else A.INTEG_KD)
    if algorithm.lower() != "barnes-hut gpu":  # the CPU paths' physics
        kw["force_kernel"] = A.FORCE_REF_CPU
    if algorithm.lower() == "barnes-hut":
        kw.update(mac=A.MAC_DIAG, cells=A.CELLS_REF_CPU)
    with nbody2d.Context(m, p, v, **kw) as c:
        c.step(7)
        gp, gv = c.state()
    assert np.array_equal(vals[:, 1:3], gp) and np.array_equal(vals[:, 3:5], gv)


@pytest.mark.gpu
@pytest.mark.parametrize("algorithm", ["Barnes-Hut", "barnes-hut gpu"])
def test_driver_softening_factor_exact_branch(tmp_path, algorithm):
    """epsilon from softening_factor (Simulation.cpp:82-126): with n^2 <= 1e6
    the driver takes the exact branch, factor x (sum of pair distances) / n,
    equal to the oracle's; the run then equals the library run with that
    absolute epsilon, bit for bit."""
    m, p, v = ics.exp_disk(800, seed=5)
    src, out = tmp_path / "in.csv", tmp_path / "out.csv"
    write_csv(src, m, p, v)
    cfg = config(tmp_path, src, out, iterations=5, algorithm=algorithm, G=1.0,
                 softening_factor=0.01, theta=0.5, timestep=1e-3)
    print(f"[gpu] driver softening_factor {algorithm}", flush=True)
    stats = json.loads(run(cfg).stdout)
    want = O.load().or_plummer_softening_exact(len(p), O.ptr(np.ascontiguousarray(p)), 0.01, 1e6)
    assert want > 0 and stats["epsilon"] == pytest.approx(want, rel=1e-14)
    A = nbody2d.abi
    kw = dict(G=1.0, dt=1e-3, theta=0.5, softening=stats["epsilon"])
    if algorithm == "Barnes-Hut":
        kw.update(force_kernel=A.FORCE_REF_CPU, mac=A.MAC_DIAG, cells=A.CELLS_REF_CPU)
    with nbody2d.Context(m, p, v, **kw) as c:
        c.step(5)
        gp, gv = c.state()
    ids, vals = read_csv(out)
    assert np.array_equal(vals[:, 1:3], gp) and np.array_equal(vals[:, 3:5], gv)


@pytest.mark.gpu
def test_run_pause_set_timestep_from_second_thread(tmp_path):
    m, p, v = ics.exp_disk(2000, seed=3)
    src, out = tmp_path / "in.csv", tmp_path / "out.csv"
    write_csv(src, m, p, v)
    cfg = config(tmp_path, src, out, iterations=40, G=1.0, softening=0.01, timestep=1e-3)
    print("[gpu] driver run/pause", flush=True)
    r = json.loads(run(cfg, "--run-pause").stdout)
    assert r["finished"] and r["iteration"] == 40
    k = r["paused_at"]
    assert 1 <= k <= 40
    # The step is latched after each iteration (post_iteration,
    # Simulation.cpp:68-72): the first step after the resume still runs with
    # the old dt, the rest with the new one.
    old = min(k + 1, 40)
    assert r["simulated_elapsed_s"] == pytest.approx(old * 1e-3 + (40 - old) * 2e-3, rel=1e-12)
