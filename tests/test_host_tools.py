"""Host-side pieces of the measurement and batching tools (no GPU): the
benchmark's byte models, its traffic file, its catalogue comparison and its
refusal of several GPUs, and the committed autotuner table."""

import importlib.util
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_algorithmic_bytes_closed_forms():
    """CIC: positions read once (12 B per particle) and each sub-grid written
    once (4 B per cell); binning: the half-complex modes of every sub-grid
    read once (8 B per mode)."""
    b = _bench()
    nd = 512
    ng, nh = nd**3, nd * nd * (nd // 2 + 1)
    assert b.algorithmic_bytes("cic_deposit", nd, 0, 1, 4) == 12 * ng + 4 * ng * 2
    assert b.algorithmic_bytes("cic_deposit", nd, 0, 0, 8) == 24 * ng + 8 * ng
    assert b.algorithmic_bytes("power_bins", nd, 0, 1, 4) == 8 * nh * 2
    assert b.algorithmic_bytes("no_such_stage", nd, 0, 1, 4) is None


def test_traffic_file_matches_its_counters():
    """profiles/traffic_nd512.json: traffic = FETCH_SIZE x (copy bytes /
    copy FETCH_SIZE) + WRITE_SIZE x (copy bytes / copy WRITE_SIZE), in
    bytes, for the stage the benchmark names."""
    b = _bench()
    with open(os.path.join(ROOT, "profiles", "traffic_nd512.json")) as f:
        t = json.load(f)
    copy_kib = t["copy_mib"] * 1024
    expect = (t["fetch_kib"] * copy_kib / t["copy_fetch_kib"] +
              t["write_kib"] * copy_kib / t["copy_write_kib"]) * 1024
    assert abs(t["traffic_bytes"] / expect - 1) < 1e-4
    assert t["algorithmic_bytes"] == b.algorithmic_bytes(t["stage"], t["nd"], 0, 1, 4)
    assert b.measured_traffic(t["stage"], t["nd"]) == t["traffic_bytes"]
    assert b.measured_traffic(t["stage"], t["nd"] * 2) is None


def test_catalogue_agreement_counts():
    b = _bench()
    rng = np.random.default_rng(3)
    x = np.unique(rng.integers(0, 512, (400, 3)).astype(np.int32), axis=0)
    rng.shuffle(x)
    n = rng.integers(1, 40, len(x)).astype(np.int32)
    m = (rng.random(len(x)) + 1.0).astype(np.float32)
    # the second catalogue: the same halos, one count changed, two halos
    # dropped, two extra, and the first ten in another place
    x2, n2, m2 = x.copy(), n.copy(), m.astype(np.float64) * 1.001
    n2[50] += 1
    x2 = np.concatenate([x2[:-2], np.array([[600, 1, 1], [601, 1, 1]], np.int32)])
    n2 = np.concatenate([n2[:-2], np.array([1, 1], np.int32)])
    m2 = np.concatenate([m2[:-2], np.ones(2)])
    x2[:10] = x2[:10][::-1]
    n2[:10] = n2[:10][::-1]
    m2[:10] = m2[:10][::-1]
    r = b.catalogue_agreement((x, n, m), (x2, n2, m2))
    assert r["nh_fp32"] == len(x) and r["nh_fp64"] == len(x2)
    assert r["same_centre"] == len(x) - 2
    assert r["same_centre_and_count"] == len(x) - 3
    assert r["same_place"] == len(x) - 3 - 10
    assert abs(r["mass_rel_diff_max"] - 0.001) < 1e-5


def test_bench_refuses_several_gpus():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr)


def test_committed_tuning_table():
    from exshalos_amd.batch import HW_QUEUES, load_table

    t = load_table()
    assert t["device"].startswith("MI355X")
    assert 1 <= t["workers"] <= 16 and str(t["workers"]) in t["mcells_per_s_by_workers"]
    assert t["hw_queues"] == HW_QUEUES
    assert set(t["sizes"]) == {"64", "96", "128", "192"}
    for v in t["sizes"].values():
        assert v["deposit_tile"] in (16, 3216, 32)
        assert str(v["deposit_tile"]) in v["ms_per_box"]
