"""Benchmark of the CD-MP2 energy + gradient on one MI355X.

Workload (default): closed-shell CD-RMP2 energy plus gradient
(dE/dB^Q_ia and dE/deps_p) at Q=4000, o=100, v=1100 on seeded synthetic
B^Q_ia and orbital energies, inputs resident in HBM before timing starts.
One step = one full E+grad evaluation.  Both the FP64 form and the
split-BF16 form (when requested/available) are timed; the JSON line reports
the form named by --precision and carries the other as a comparison.

Prints ONE JSON line on rank 0 (this build runs on one GPU only).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse_precision(s: str) -> int:
    if s in ("fp64", "f64", "0"):
        return 0
    if s.startswith("bf16x"):
        return int(s[5:])
    raise ValueError(f"unknown precision {s}")


def prec_name(p: int) -> str:
    return "f64" if p == 0 else f"bf16x{p}"


def rmp2_work(Q, o, v):
    """Algorithmic flops of the energy: [o(o+1)/2] v^2 2Q (i <= j pair blocks)."""
    return o * (o + 1) / 2 * v * v * 2.0 * Q


def time_steps(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def cpu_baseline(B, eo, ev, npairs, seed=7):
    """The oracle's E+grad per-pair work on a bounded sample of ordered pairs,
    extrapolated linearly to all o^2 ordered pairs (per-pair work is uniform)."""
    import oracle

    Q, o, v = B.shape
    rng = np.random.default_rng(seed)
    rows = rng.permutation(o)[:npairs]
    cols = rng.integers(0, o, size=npairs)
    pairs = np.stack([rows, cols], axis=1).astype(np.int32)
    t0 = time.perf_counter()
    oracle.rmp2_grad_pairs(B, eo, ev, pairs)
    dt = time.perf_counter() - t0
    total = dt * (o * o) / npairs
    return {
        "value": total,
        "unit": "s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"{npairs} of {o * o} ordered (i,j) pairs (G, amplitudes, energy and "
                  f"dE/dB row update per pair), {dt:.1f} s measured, extrapolated x{o * o / npairs:.0f}",
    }


def measured_traffic(kernel: str, workload: str, precision: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE scaled by the calibration kernel), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}_{precision}.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16x5", help="fp64 | bf16xK (K slices)")
    ap.add_argument("--compare", default="fp64", help="second precision timed beside it ('' for none)")
    ap.add_argument("--workload", default="mp2_o100")
    ap.add_argument("--cpu-pairs", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB moved by one plain device copy before timing (a known-size "
                         "kernel for scaling PMC byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are out of "
                                   "scope (see DESIGN.md for how the path would shard)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from mshqc_amd import hip, synthetic as syn

    w = syn.WORKLOADS[args.workload]
    assert w["kind"] == "rmp2"
    Q, o, v = w["Q"], w["o"], w["v"]
    log(f"[bench] generating {args.workload}: Q={Q} o={o} v={v}")
    B, eo, ev = syn.workload(args.workload)
    dB = torch.from_numpy(B).cuda()
    deo = torch.from_numpy(eo).cuda()
    dev = torch.from_numpy(ev).cuda()
    ctx = hip.context(0)
    W = rmp2_work(Q, o, v)
    if args.copy_calibration:
        # A plain streaming kernel of known size with a name of its own
        # (sqrt): it reads and writes nel * 8 bytes once each.
        nel = args.copy_calibration * (1 << 20) // 8
        src = torch.full((nel,), 2.0, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.sqrt(src, out=dst)
        torch.cuda.synchronize()
        log(f"[bench] calibration sqrt: {nel * 8} bytes read, {nel * 8} bytes written")
        del src, dst

    results = {}
    precs = [parse_precision(args.precision)]
    if args.compare:
        precs.append(parse_precision(args.compare))
    for p in precs:
        log(f"[bench] E+grad {prec_name(p)}: warmup {args.warmup}, steps {args.steps}")

        def step():
            return ctx.rmp2_grad(dB, deo, dev, precision=p)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        out = None
        for _ in range(args.steps):
            out = step()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / args.steps
        prof = {k: ctx.profile_get(k) for k in (0, 1, 2)}
        ctx.profile(False)
        E = out[0]
        log(f"[bench] {prec_name(p)}: {sec:.4f} s/step, E = {E!r}, pair kernel {prof[0]}, gemm {prof[1]}")
        results[p] = dict(sec=sec, E=E, prof=prof, dEdB=out[1], deo=out[2], dev=out[3])

    main_p = precs[0]
    r = results[main_p]
    sec = r["sec"]
    npair_launch, pair_ms = r["prof"][0]
    ngemm, gemm_ms = r["prof"][1]
    steps = args.steps
    pair_ms_step = pair_ms / steps
    gemm_ms_step = gemm_ms / steps
    # Dominant kernel: the pair-block (ia|jb) kernel. Algorithmic flops per step: W.
    peak = FP64_PEAK_TFLOPS if main_p == 0 else BF16_PEAK_TFLOPS
    pair_tflops = W / (pair_ms_step * 1e-3) / 1e12 if pair_ms_step > 0 else 0.0
    hw_factor = 1 if main_p == 0 else main_p * (main_p + 1) // 2
    line = {
        "metric": "CD-MP2 E+grad seconds & TFLOP/s, split-BF16 vs FP64 on MI355X at <=1 uEh error",
        "value": sec,
        "unit": "s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec_name(main_p),
        "data": "synthetic (seeded generators: B=0.05 N(0,1) exp(-2a/v), eps_i~U[-1.5,-0.3], "
                "eps_a=0.1+3(a/v)^1.5)",
        "config": {"workload": f"{args.workload}: CD-RMP2 energy + gradient (dE/dB, dE/deps)",
                   "Q": Q, "o": o, "v": v, "precision": prec_name(main_p),
                   "parallelism": "single GPU"},
        "tflops": 3 * W / sec / 1e12,
        "energy": r["E"],
        "kernels": {
            "pair_kernel": {"launches_per_step": npair_launch / steps, "ms_per_step": pair_ms_step,
                            "algorithmic_flop_per_step": W},
            "gemm": {"launches_per_step": ngemm / steps, "ms_per_step": gemm_ms_step,
                     "algorithmic_flop_per_step": 2 * W},
        },
        "roofline": {
            "kernel": "pair-block (ia|jb) kernel with fused denominator epilogue",
            "bound": "mfma",
            "achieved": pair_tflops,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": (pair_tflops * hw_factor) / peak,
            "traffic": measured_traffic("pair", args.workload, prec_name(main_p)),
            "traffic_unit": "bytes per launch (HBM read + write)",
        },
    }
    if len(precs) > 1:
        other = results[precs[1]]
        dE = r["E"] - other["E"]
        line["compare"] = {
            "precision": prec_name(precs[1]),
            "seconds": other["sec"],
            "energy": other["E"],
            "abs_dE": abs(dE),
            "rel_dE": abs(dE) / abs(other["E"]),
            "speedup": other["sec"] / sec,
            "grad_rel_err": float(torch.linalg.norm(r["dEdB"] - other["dEdB"]).item()
                                  / torch.linalg.norm(other["dEdB"]).item()),
        }
    if not args.no_cpu:
        log(f"[bench] cpu baseline on {args.cpu_pairs} sampled pairs")
        line["cpu_baseline"] = cpu_baseline(B, eo, ev, args.cpu_pairs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
