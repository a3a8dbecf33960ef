"""bench.py's measurement contract, on the CPU: the one-GPU refusal, the
traffic figure it reads from the committed PMC summary, its flop counts
against a brute-force count, and the CPU baseline leg's fields."""
import itertools
import json
import os
import subprocess
import sys

import numpy as np

import bench
from mshqc_amd import synthetic as syn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    """--gpus 2 exits with status 2 and a JSON line naming the reason, before
    anything touches a device."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU only" in line["error"]


def test_traffic_reads_the_committed_summary():
    """The roofline's traffic is the committed summary's per-launch figure,
    and that figure is the sum of its scaled FETCH_SIZE and WRITE_SIZE rows."""
    t = bench.measured_traffic("pair", "mp2_o100", "bf16x4")
    assert t is not None
    with open(os.path.join(ROOT, "profiles", "pmc_pair_mp2_o100_bf16x4.md")) as f:
        row = next(ln for ln in f if ln.startswith("| `pair_bf16_kernel") and "scaled bytes" in ln)
    fetch, write = (float(x) for x in row.strip().strip("|").split("|")[1:3])
    assert abs(t - (fetch + write)) <= 1e-3 * t
    assert bench.measured_traffic("pair", "no_such_workload", "bf16x4") is None


def test_work_counts_match_brute_force():
    """rmp2_work counts the i <= j pair blocks' products (2 flop per Q term);
    ump2_work the i < j same-spin blocks of each spin plus every
    opposite-spin pair."""
    Q, o, v = 5, 4, 3
    n = sum(1 for i, j in itertools.product(range(o), repeat=2) if i <= j for _ in range(v * v))
    assert bench.rmp2_work(Q, o, v) == n * 2 * Q
    oa, va, ob, vb = 4, 3, 3, 2
    same = sum(1 for i, j in itertools.product(range(oa), repeat=2) if i < j) * va * va
    same += sum(1 for i, j in itertools.product(range(ob), repeat=2) if i < j) * vb * vb
    opp = oa * ob * va * vb
    assert bench.ump2_work(Q, oa, va, ob, vb) == (same + opp) * 2 * Q


def test_cpu_baseline_fields_and_extrapolation():
    """The CPU leg on a tiny input: the fields the bench line carries, a
    sample no larger than one pair per first index, and positive
    extrapolated times."""
    B, eo, ev = syn.rmp2_inputs(16, 6, 10, 3)
    cb = bench.cpu_baseline(B, eo, ev, npairs=100, nref=2)
    assert set(cb) >= {"value", "unit", "cores", "kind", "sample", "reference_energy_loop"}
    assert cb["kind"] == "port" and cb["unit"] == "s" and cb["cores"] >= 1
    assert cb["value"] > 0 and cb["reference_energy_loop"]["value"] > 0
    assert cb["sample"].startswith("6 of 36 ordered")
    assert cb["reference_energy_loop"]["sample"].startswith("2 of 36 ordered")
    assert np.isfinite(cb["value"])
