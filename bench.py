"""Benchmark of the CD-MP2 energy + gradient on one MI355X.

Workload (default, `mp2_o100`): closed-shell CD-RMP2 energy plus gradient
(dE/dB^Q_ia and dE/deps_p) at Q=4000, o=100, v=1100 on seeded synthetic
B^Q_ia and orbital energies, inputs resident in HBM before timing starts.
One step = one full E+grad evaluation.  The split-BF16 form (4 slices) is
timed and reported; the FP64 form is timed beside it as the comparison.
The RMP2 workloads also time the energy alone in both forms (`energy_only`:
the pair kernel with its fused denominator epilogue, no amplitudes or
gradient); at o=60, v=540, Q=2400 the FP64 energy is the plain form's
output.

Other workloads (`--workload`):
  ump2_o120  CD-UMP2 energy + gradient, o_a=120, v_a=1380, o_b=118,
             v_b=1382, Q=5600 (same-spin and opposite-spin blocks)
  jk600      Fock J/K build from L^Q (n=600, Q=2400): general-P J + K, the
             occupied-factor K (C_occ 600x60) and the AO->MO transform to
             B^Q_ia; one step = those three calls
  mp2_o60, mp2_o40  smaller CD-RMP2 E+grad workloads
  transform_o100  AO->MO transform B = C_occ^T L C_vir at n=1200, Q=4000,
             o=100, v=1100 (L of 46 GB generated on the device), split-BF16
             (4 slices) beside FP64, with the |dE| it adds to the energy

Prints ONE JSON line (this build runs on one GPU only).
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
METRIC = "CD-MP2 E+grad seconds & TFLOP/s, split-BF16 vs FP64 on MI355X at <=1 uEh error"
DATA = ("synthetic (seeded generators: B=0.05 N(0,1) exp(-2a/v), eps_i~U[-1.5,-0.3], "
        "eps_a=0.1+3(a/v)^1.5; L=sym(0.1 N(0,1) exp(-|mu-nu|/50)), C_occ=QR(N(0,1)))")


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


def ump2_work(Q, oa, va, ob, vb):
    """[oa(oa-1)/2 va^2 + ob(ob-1)/2 vb^2 + oa ob va vb] 2Q."""
    return (oa * (oa - 1) / 2 * va * va + ob * (ob - 1) / 2 * vb * vb + oa * ob * va * vb) * 2.0 * Q


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


def host_cpu():
    """The host's logical CPU count and model name (from /proc/cpuinfo)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"logical_cpus": os.cpu_count(), "model": model}


def cpu_baseline(B, eo, ev, npairs, nref, seed=7):
    """CPU legs on a bounded sample of ordered pairs (i, j), one pair per first
    index i (a fixed random j each), extrapolated linearly to all o^2 ordered
    pairs (the per-pair work is uniform):
      * value: the oracle's E+grad per pair (the pair block G, amplitudes,
        energy and the dE/dB row update) -- the same work as one GPU step;
      * reference_energy_loop: the reference's energy-only loop as written
        (cholesky_rmp2.cc:118-155, rank-1 updates, t2 store left out) on the
        first `nref` of the sampled pairs."""
    import oracle

    Q, o, v = B.shape
    npairs = min(npairs, o)
    rng = np.random.default_rng(seed)
    rows = rng.permutation(o)[:npairs]
    cols = rng.integers(0, o, size=npairs)
    pairs = np.stack([rows, cols], axis=1).astype(np.int32)
    t0 = time.perf_counter()
    oracle.rmp2_grad_pairs(B, eo, ev, pairs)
    dt = time.perf_counter() - t0
    nref = min(nref, npairs)
    t0 = time.perf_counter()
    oracle.rmp2_energy_pairs_rank1(B, eo, ev, pairs[:nref])
    dt_ref = time.perf_counter() - t0
    return {
        "value": dt * (o * o) / npairs,
        "unit": "s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "host": host_cpu(),
        "sample": f"{npairs} of {o * o} ordered (i,j) pairs ({100.0 * npairs / (o * o):.1f}%; G, amplitudes, "
                  f"energy and dE/dB row update per pair), {dt:.1f} s measured, extrapolated x{o * o / npairs:.0f}",
        "reference_energy_loop": {
            "value": dt_ref * (o * o) / nref,
            "unit": "s",
            "what": "energy only, the reference's rank-1 pair loop (cholesky_rmp2.cc:118-155) restated, "
                    "no t2 store",
            "sample": f"{nref} of {o * o} ordered pairs, {dt_ref:.1f} s measured, extrapolated x{o * o / nref:.0f}",
        },
    }


def timed(torch, ctx, fn, steps, warmup):
    """Seconds per step of fn() (wall clock over all steps, synchronized on
    both sides) and the per-kernel-class profile of the timed steps; the
    profile's "step_ms" holds each step's device time from HIP events on the
    stream the library runs on (it follows torch's current stream)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ctx.profile(True)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    out = None
    for a, b in evs:
        a.record()
        out = fn()
        b.record()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    prof = {k: ctx.profile_get(k) for k in (0, 1, 2)}
    prof["step_ms"] = [a.elapsed_time(b) for a, b in evs]
    ctx.profile(False)
    return sec, out, prof


def step_stats(step_ms):
    """Median, min and max of the per-step device times, in seconds."""
    x = sorted(step_ms)
    n = len(x)
    med = x[n // 2] if n % 2 else 0.5 * (x[n // 2 - 1] + x[n // 2])
    return {"median": med * 1e-3, "min": x[0] * 1e-3, "max": x[-1] * 1e-3, "n": n,
            "timer": "HIP events around each step on the library's stream"}


def calibrate(torch, mib):
    """A plain streaming kernel of known size with a name of its own (sqrt):
    it reads and writes mib MiB once each; PMC byte counters are scaled by it."""
    nel = mib * (1 << 20) // 8
    src = torch.full((nel,), 2.0, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.sqrt(src, out=dst)
    torch.cuda.synchronize()
    log(f"[bench] calibration sqrt: {nel * 8} bytes read, {nel * 8} bytes written")


def base_line(args, sec, dtype, config, flop, step_ms=None):
    line = {
        "metric": METRIC,
        "value": sec,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": DATA,
        "config": config,
        "tflops": flop / sec / 1e12 if flop else None,
    }
    if step_ms:
        line["step_seconds"] = step_stats(step_ms)
    return line


def mp2_line(args, results, precs, W3, config, energy_of, grads_of):
    """JSON line of an MP2 E+grad workload (RMP2 or UMP2)."""
    import torch

    main_p = precs[0]
    r = results[main_p]
    steps = args.steps
    npair_launch, pair_ms = r["prof"][0]
    ngemm, gemm_ms = r["prof"][1]
    pair_ms_step, gemm_ms_step = pair_ms / steps, gemm_ms / steps
    W = W3 / 3
    peak = FP64_PEAK_TFLOPS if main_p == 0 else BF16_PEAK_TFLOPS
    pair_tflops = W / (pair_ms_step * 1e-3) / 1e12 if pair_ms_step > 0 else 0.0
    hw_factor = 1 if main_p == 0 else main_p * (main_p + 1) // 2
    gemm_hw = 1 if main_p == 0 else 6  # the gradient GEMMs keep min(k, 3) slices: 6 products
    line = base_line(args, r["sec"], prec_name(main_p), config, W3, r["prof"]["step_ms"])
    line["energy"] = energy_of(r["out"])
    line["kernels"] = {
        "pair_kernel": {"launches_per_step": npair_launch / steps, "ms_per_step": pair_ms_step,
                        "algorithmic_flop_per_step": W, "hardware_products_per_flop": hw_factor},
        "gemm": {"timed_calls_per_step": ngemm / steps, "ms_per_step": gemm_ms_step,
                 "algorithmic_flop_per_step": 2 * W, "hardware_products_per_flop": gemm_hw,
                 "note": "gradient contractions; a timed call is one GEMM in FP64 and one occupied "
                         "block pair (its one or two GEMMs and the slicing of the amplitude block) "
                         "in the split form"},
    }
    line["roofline"] = {
        "kernel": "pair-block (ia|jb) kernel with fused denominator epilogue",
        "bound": "mfma",
        "achieved": pair_tflops,
        "peak": peak,
        "unit": "TFLOP/s",
        # achieved counts algorithmic flops only; in the split form each one
        # costs hw_factor BF16 MFMA products
        "frac": pair_tflops / peak,
        "frac_hardware_products": pair_tflops * hw_factor / peak,
        "achieved_vs_fp64_peak": pair_tflops / FP64_PEAK_TFLOPS,
        "traffic": measured_traffic("pair", args.workload, prec_name(main_p)),
        "traffic_unit": "bytes per launch beyond L2: FETCH_SIZE (misses of the L2, served by the MALL or "
                        "HBM) + WRITE_SIZE, each scaled by a known-size copy",
    }
    if len(precs) > 1:
        other = results[precs[1]]
        Ea, Eb = energy_of(r["out"]), energy_of(other["out"])
        ga, gb = grads_of(r["out"]), grads_of(other["out"])
        num = sum(float(torch.linalg.norm(x - y).item()) ** 2 for x, y in zip(ga, gb)) ** 0.5
        den = sum(float(torch.linalg.norm(y).item()) ** 2 for y in gb) ** 0.5
        line["compare"] = {
            "precision": prec_name(precs[1]),
            "seconds": other["sec"],
            "energy": Eb,
            "abs_dE": abs(Ea - Eb),
            "rel_dE": abs(Ea - Eb) / abs(Eb),
            "speedup": other["sec"] / r["sec"],
            "steps": len(other["prof"]["step_ms"]),
            "grad_rel_err": num / den,
        }
    return line


def energy_only_leg(args, torch, ctx, precs, dB, deo, dev, W):
    """The CD-RMP2 energy alone (no amplitudes, no gradient: the pair kernel
    with its fused denominator epilogue) in each timed precision, at most 5
    steps after one warm-up, with the energy difference against FP64 when
    both forms run."""
    out = {}
    for p in precs:
        sec, E, prof = timed(torch, ctx, lambda: ctx.rmp2_energy(dB, deo, dev, precision=p),
                             min(args.steps, 5), min(args.warmup, 1))
        log(f"[bench] energy only {prec_name(p)}: {sec:.4f} s/step, E = {E!r}")
        out[prec_name(p)] = {"seconds": sec, "energy": E, "tflops": W / sec / 1e12,
                             "step_seconds": step_stats(prof["step_ms"])}
    if 0 in precs and len(precs) > 1:
        other = prec_name(precs[0] if precs[0] != 0 else precs[1])
        out["abs_dE"] = abs(out[other]["energy"] - out["f64"]["energy"])
        out["speedup"] = out["f64"]["seconds"] / out[other]["seconds"]
    return out


def run_mp2(args, torch, hip, syn, w):
    ctx = hip.context(0)
    if args.zblock:
        ctx.set_zblock(int(args.zblock))
    precs = [parse_precision(args.precision)]
    if args.compare:
        precs.append(parse_precision(args.compare))
    if w["kind"] == "rmp2":
        Q, o, v = w["Q"], w["o"], w["v"]
        log(f"[bench] generating {args.workload}: Q={Q} o={o} v={v}")
        B, eo, ev = syn.workload(args.workload)
        dB, deo, dev = (torch.from_numpy(x).cuda() for x in (B, eo, ev))
        W3 = 3 * rmp2_work(Q, o, v)
        config = {"workload": f"{args.workload}: CD-RMP2 energy + gradient (dE/dB, dE/deps)",
                  "Q": Q, "o": o, "v": v, "parallelism": "single GPU"}

        def step_of(p):
            return lambda: ctx.rmp2_grad(dB, deo, dev, precision=p)

        def energy_of(out):
            return out[0]

        def grads_of(out):
            return [out[1]]
    else:
        Q, oa, va, ob, vb = w["Q"], w["oa"], w["va"], w["ob"], w["vb"]
        log(f"[bench] generating {args.workload}: Q={Q} oa={oa} va={va} ob={ob} vb={vb}")
        (Ba, eoa, eva), (Bb, eob, evb) = syn.workload(args.workload)
        dev_in = [torch.from_numpy(x).cuda() for x in (Ba, Bb, eoa, eva, eob, evb)]
        del Ba, Bb
        W3 = 3 * ump2_work(Q, oa, va, ob, vb)
        config = {"workload": f"{args.workload}: CD-UMP2 energy + gradient (same and opposite spin)",
                  "Q": Q, "o_alpha": oa, "v_alpha": va, "o_beta": ob, "v_beta": vb,
                  "parallelism": "single GPU"}

        def step_of(p):
            return lambda: ctx.ump2_grad(*dev_in, precision=p)

        def energy_of(out):
            return float(sum(out[0]))

        def grads_of(out):
            return [out[1][0], out[1][1]]
    if args.copy_calibration:
        calibrate(torch, args.copy_calibration)
    results = {}
    for k, p in enumerate(precs):
        # the comparison leg: at most 5 timed steps (its per-step times vary by
        # well under 1%), so the default run stays within a few minutes
        steps, warmup = (args.steps, args.warmup) if k == 0 else (min(args.steps, 5), min(args.warmup, 1))
        log(f"[bench] E+grad {prec_name(p)}: warmup {warmup}, steps {steps}")
        sec, out, prof = timed(torch, ctx, step_of(p), steps, warmup)
        log(f"[bench] {prec_name(p)}: {sec:.4f} s/step, E = {energy_of(out)!r}, pair kernel {prof[0]}, "
            f"gemm {prof[1]}")
        results[p] = dict(sec=sec, out=out, prof=prof)
    config["precision"] = prec_name(precs[0])
    line = mp2_line(args, results, precs, W3, config, energy_of, grads_of)
    if w["kind"] == "rmp2" and not args.no_energy_only:
        line["energy_only"] = energy_only_leg(args, torch, ctx, precs, dB, deo, dev, W3 / 3)
    if w["kind"] == "rmp2" and not args.no_cpu:
        log(f"[bench] cpu baseline on {args.cpu_pairs} sampled pairs ({args.cpu_ref_pairs} for the reference loop)")
        line["cpu_baseline"] = cpu_baseline(B, eo, ev, args.cpu_pairs, args.cpu_ref_pairs)
    return line


def run_jk(args, torch, hip, syn, w):
    """General-P J + K, occupied-factor J + K, restricted open-shell J + Ka + Kb
    and the AO->MO transform on L^Q."""
    ctx = hip.context(0)
    Q, n, o = w["Q"], w["n"], w["o"]
    log(f"[bench] generating {args.workload}: Q={Q} n={n} o={o}")
    L, C, P = syn.workload(args.workload)
    dL = torch.from_numpy(L).cuda()
    del L
    dC, dP = torch.from_numpy(C).cuda(), torch.from_numpy(P).cuda()
    Cvir = torch.from_numpy(np.ascontiguousarray(np.linalg.qr(
        np.random.default_rng(w["seed"] + 1).standard_normal((n, n - o)))[0])).cuda()
    parts = {
        "jk_general_P": (lambda: ctx.jk(dL, dP, screen=True), 4.0 * Q * n ** 3),
        "jk_occupied_factor": (lambda: ctx.jk_factor(dL, dC), 4.0 * Q * n * n * o),
        # restricted open shell, o alpha and o - 2 beta electrons (a triplet)
        "jk_rohf": (lambda: ctx.jk_rohf(dL, dC, o, o - 2), 2.0 * Q * n * n * o + 2.0 * Q * n * n * (2 * o - 2)),
        "transform": (lambda: ctx.transform(dL, dC, Cvir), 2.0 * Q * n * n * o + 2.0 * Q * n * o * (n - o)),
    }
    if args.copy_calibration:
        calibrate(torch, args.copy_calibration)
    out = {}
    total = 0.0
    for name, (fn, flops) in parts.items():
        sec, _, prof = timed(torch, ctx, fn, args.steps, args.warmup)
        j_ms = prof[2][1] / args.steps
        out[name] = {"seconds": sec, "flop": flops, "tflops": flops / sec / 1e12,
                     "j_kernel_ms": j_ms if j_ms > 0 else None, "step_seconds": step_stats(prof["step_ms"])}
        total += sec
        log(f"[bench] {name}: {sec * 1e3:.2f} ms/step, {flops / sec / 1e12:.1f} TFLOP/s, J kernels {prof[2]}")
    # one pass over the mu <= nu half of L, plus the X_q = L^q C_occ the
    # densities gamma_q = <L^q C_occ, C_occ> are read from
    jb = 4.0 * Q * n * (n + 1) + 8.0 * Q * n * o
    j_ms = out["jk_occupied_factor"]["j_kernel_ms"]
    config = {"workload": f"{args.workload}: CD-RHF/ROHF Fock pieces from L", "Q": Q, "n": n, "o": o,
              "parallelism": "single GPU"}
    line = base_line(args, total, "f64", config, sum(f for _, f in parts.values()))
    line["step_seconds"] = {name: v["step_seconds"] for name, v in out.items()}
    line["kernels"] = out
    if j_ms:
        achieved = jb / (j_ms * 1e-3) / 1e9
        line["roofline"] = {"kernel": "J (occupied factor): densities from the exchange's L^q C_occ, then one "
                                      "accumulation pass over the upper half of each symmetric L^Q",
                            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    return line


def run_transform(args, torch, hip, syn, w):
    """The AO->MO transform at the precision workload's shape: one step = one
    mshqc_hip_transform call; L = sym(0.1 N(0,1) exp(-|mu-nu|/50)) generated
    on the device from a seeded torch generator (46 GB, too large to build on
    the host in a bench run), C = Q factor of a seeded Gaussian n x n."""
    ctx = hip.context(0)
    Q, n, o = w["Q"], w["n"], w["o"]
    v = n - o
    log(f"[bench] generating {args.workload}: Q={Q} n={n} o={o} v={v} on the device")
    g = torch.Generator(device="cuda").manual_seed(w["seed"])
    idx = torch.arange(n, device="cuda", dtype=torch.float64)
    env = torch.exp(-(idx[:, None] - idx[None, :]).abs() / 50.0)
    L = torch.empty((Q, n, n), dtype=torch.float64, device="cuda")
    for q0 in range(0, Q, 100):
        A = 0.1 * torch.randn((min(100, Q - q0), n, n), generator=g, dtype=torch.float64, device="cuda") * env
        L[q0:q0 + A.shape[0]] = 0.5 * (A + A.transpose(1, 2))
    del A
    C = np.linalg.qr(np.random.default_rng(w["seed"]).standard_normal((n, n)))[0]
    Co = torch.from_numpy(np.ascontiguousarray(C[:, :o])).cuda()
    Cv = torch.from_numpy(np.ascontiguousarray(C[:, o:])).cuda()
    eo, ev = syn.gen_eps(o, v, np.random.default_rng(w["seed"] + 1))
    deo, dev = torch.from_numpy(eo).cuda(), torch.from_numpy(ev).cuda()
    p = parse_precision(args.precision)
    p = p if p else 4
    res = {}
    for prec in (p, 0):
        sec, B, prof = timed(torch, ctx, lambda: ctx.transform(L, Co, Cv, precision=prec), args.steps, args.warmup)
        res[prec] = (sec, B, prof)
        log(f"[bench] transform {prec_name(prec)}: {sec * 1e3:.2f} ms/step, GEMMs {prof[1]}")
    E = {prec: ctx.rmp2_energy(res[prec][1], deo, dev) for prec in res}
    flop = 2.0 * Q * n * n * o + 2.0 * Q * n * o * v
    bytes_alg = 8.0 * Q * n * n + 8.0 * Q * o * v
    sec = res[p][0]
    config = {"workload": f"{args.workload}: AO->MO transform B = C_occ^T L C_vir", "Q": Q, "n": n, "o": o, "v": v,
              "parallelism": "single GPU", "precision": prec_name(p)}
    line = base_line(args, sec, prec_name(p), config, flop, res[p][2]["step_ms"])
    # At k slices the GEMMs issue k(k+1)/2 BF16 products per flop: 10 at
    # k = 4 is 444 hardware flop per algorithmic byte, past the ridge of
    # 2.5 PF/s / 8 TB/s, so the dominant GEMM is priced against the MFMA peak;
    # the HBM view of the whole call is kept beside it.
    n_gemm, gemm_ms = res[p][2][1]
    gemm_s = gemm_ms * 1e-3 / args.steps
    achieved = flop / gemm_s / 1e12
    hbm = bytes_alg / sec / 1e9
    line["roofline"] = {"kernel": f"gemm_bf16_nt_kernel<{p}> (both transform GEMMs of every Q chunk)",
                        "bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / BF16_PEAK_TFLOPS,
                        "frac_hardware_products": achieved * p * (p + 1) // 2 / BF16_PEAK_TFLOPS,
                        "gemm_seconds_per_step": gemm_s, "traffic": None,
                        "hbm_view": {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm / HBM_PEAK_GBS,
                                     "algorithmic_bytes": bytes_alg,
                                     "note": "whole call; L read once (8 Q n^2) + B written once (8 Q o v)"}}
    line["compare"] = {"precision": "f64", "seconds": res[0][0], "speedup": res[0][0] / sec,
                       "tflops": flop / res[0][0] / 1e12,
                       "B_rel_err": (torch.abs(res[p][1] - res[0][1]).max() / torch.abs(res[0][1]).max()).item(),
                       "energy_f64_B": E[0], "energy_split_B": E[p], "abs_dE": abs(E[p] - E[0])}
    return line


def omp2_flops(Q, oa, va, ob, vb, swapped):
    """FP64 flops the OMP2 amplitude/OPDM call executes: the (ia|kb) GEMMs
    over all ordered occupied pairs of each spin block (twice for the
    opposite-spin block when its G_oo^b needs the spin-swapped pass), and the
    OPDM contractions G_oo = sum T T^T, G_vv = sum T T^T per spin block."""
    f = 2.0 * Q * ((oa * va) ** 2 + (ob * vb) ** 2 + oa * va * ob * vb * (2 if swapped else 1))
    for oX, vX in ((oa, va), (ob, vb)):
        f += 2.0 * oX ** 3 * vX ** 2 + 2.0 * oX ** 2 * vX ** 3
    f += 2.0 * oa * ob * va * vb * (oa + ob + va + vb)
    return f


def run_omp2(args, torch, hip, syn, w):
    """One CD-OMP2 iteration's amplitude work: t2 per block, its energy and the
    four OPDM blocks (the part of cholesky_omp2.cc:94-326 that scales)."""
    ctx = hip.context(0)
    Q, oa, va, ob, vb = w["Q"], w["oa"], w["va"], w["ob"], w["vb"]
    log(f"[bench] generating {args.workload}: Q={Q} oa={oa} va={va} ob={ob} vb={vb}")
    (Ba, eoa, eva), (Bb, eob, evb) = syn.workload(args.workload)
    dev = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (Ba, Bb, eoa, eva, eob, evb)]
    sec, out, prof = timed(torch, ctx, lambda: ctx.omp2_opdm(*dev), args.steps, args.warmup)
    # the opposite-spin pass is repeated with the spins swapped when its
    # amplitude block (oa va vb per beta index) cannot hold every beta index
    swapped = hip.ZBLOCK_DEFAULT // (oa * va * vb) < ob
    flop = omp2_flops(Q, oa, va, ob, vb, swapped)
    E = out[0]
    log(f"[bench] omp2_opdm: {sec * 1e3:.1f} ms/step, {flop / sec / 1e12:.1f} TFLOP/s FP64, "
        f"E = {E}, pair GEMMs {prof[0]}, OPDM GEMMs {prof[1]}")
    config = {"workload": f"{args.workload}: CD-OMP2 amplitudes, energy and OPDM per iteration", "Q": Q,
              "oa": oa, "va": va, "ob": ob, "vb": vb, "parallelism": "single GPU"}
    line = base_line(args, sec, "f64", config, flop, prof["step_ms"])
    achieved = flop / sec / 1e12
    line["roofline"] = {"kernel": "FP64 MFMA GEMMs (pair integrals + OPDM contractions)", "bound": "mfma",
                        "achieved": achieved, "peak": 78.6, "unit": "TFLOP/s", "frac": achieved / 78.6,
                        "traffic": None}
    line["energy_components"] = list(E)
    return line


def run_cd(args, torch, hip, syn, w):
    """Pivoted Cholesky of an ERI-shaped PSD matrix M = L^T L (n^2 x n^2) on
    the device (cholesky_eri.cc:40-203)."""
    ctx = hip.context(0)
    Q0, n, thr = w["Q"], w["n"], w["thr"]
    np2 = n * n
    log(f"[bench] generating {args.workload}: M = L^T L, L [{Q0}, {np2}]")
    Ls = torch.from_numpy(syn.workload(args.workload)).cuda()
    M = (Ls.t() @ Ls).contiguous()
    del Ls
    sec, out, prof = timed(torch, ctx, lambda: ctx.cholesky(M, thr), args.steps, args.warmup)
    L, conv, res = out
    q = L.shape[0]
    # HBM bytes the algorithm moves: step k reads k earlier vectors and the
    # pivot column, writes one vector, scans and updates the diagonal
    traffic = sum(8.0 * np2 * (k + 1 + 1 + 3) for k in range(q))
    achieved = traffic / sec / 1e9
    log(f"[bench] cholesky: {q} vectors (converged={conv}, max residual {res:.2e}) in {sec * 1e3:.1f} ms, "
        f"{achieved:.0f} GB/s algorithmic")
    config = {"workload": f"{args.workload}: pivoted Cholesky of an n^2 x n^2 PSD matrix", "n": n,
              "rank": Q0, "threshold": thr, "vectors": q, "parallelism": "single GPU"}
    line = base_line(args, sec, "f64", config, 0.0, prof["step_ms"])
    line["roofline"] = {"kernel": "pivot scan + new vector (reads the earlier vectors)", "bound": "hbm",
                        "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    return line


def rmp3_flops(Q, o, v):
    """FP64 flops of the CD-RMP3 call: (ac|bd) from Lvv (2 v^4 Q), the
    particle-particle contraction (2 o^2 v^4), (ki|lj) and the hole-hole term
    (2 o^4 Q + 2 o^4 v^2), (ia|jb) (2 (ov)^2 Q) and the ring product 2 (ov)^3."""
    return 2.0 * v ** 4 * Q + 2.0 * o * o * v ** 4 + 2.0 * o ** 4 * (Q + v * v) + 2.0 * (o * v) ** 2 * Q \
        + 2.0 * (o * v) ** 3


def run_rmp3(args, torch, hip, syn, w):
    """CD-RMP3 second-order amplitudes and energy (cholesky_rmp3.cc:79-200)
    from device MO blocks; the transforms and t1 are made before timing."""
    ctx = hip.context(0)
    Q, n, o = w["Q"], w["n"], w["o"]
    v = n - o
    log(f"[bench] generating {args.workload}: Q={Q} n={n} o={o} v={v}")
    L, C, eo, ev = syn.workload(args.workload)
    dL = torch.from_numpy(L).cuda()
    del L
    Co = torch.from_numpy(np.ascontiguousarray(C[:, :o])).cuda()
    Cv = torch.from_numpy(np.ascontiguousarray(C[:, o:])).cuda()
    deo, dev = torch.from_numpy(eo).cuda(), torch.from_numpy(ev).cuda()
    B, Loo, Lvv = ctx.transform(dL, Co, Cv), ctx.transform(dL, Co, Co), ctx.transform(dL, Cv, Cv)
    t1 = ctx.rmp2_t2(B, deo, dev)
    sec, out, prof = timed(torch, ctx, lambda: ctx.rmp3(B, Loo, Lvv, t1, deo, dev), args.steps, args.warmup)
    flop = rmp3_flops(Q, o, v)
    log(f"[bench] rmp3: {sec * 1e3:.1f} ms/step, {flop / sec / 1e12:.1f} TFLOP/s FP64, E3 = {out[0]!r}, "
        f"GEMMs {prof[1]}")
    config = {"workload": f"{args.workload}: CD-RMP3 second-order amplitudes and energy", "Q": Q, "n": n,
              "o": o, "v": v, "parallelism": "single GPU"}
    line = base_line(args, sec, "f64", config, flop, prof["step_ms"])
    achieved = flop / sec / 1e12
    line["roofline"] = {"kernel": "FP64 MFMA GEMMs ((ac|bd) build and ladder contractions)", "bound": "mfma",
                        "achieved": achieved, "peak": 78.6, "unit": "TFLOP/s", "frac": achieved / 78.6,
                        "traffic": None}
    line["e_mp3"] = out[0]
    return line


def ump3_flops(Q, oa, va, ob, vb):
    """FP64 flops of the CD-UMP3 call.  Per spin: (ia|jb) unscrambled and
    scrambled (4 (ov)^2 Q), (kl|ij) and the hole-hole term (2 o^4 (Q + v^2)),
    the (ab|cd) ladder through the factors and (kj|bc) (6 o^2 v^2 Q), the ring
    product (2 (ov)^3).  Opposite spin: (ia|jb) (2 oa va ob vb Q), (ki|lj) and
    its term (2 (oa ob)^2 (Q + va vb)), (ac|bd) and its ladder
    (2 (va vb)^2 (Q + oa ob)), the two rings (2 oa va ob vb (oa va + ob vb))."""
    f = 0.0
    for o, v in ((oa, va), (ob, vb)):
        f += 4.0 * (o * v) ** 2 * Q + 2.0 * o ** 4 * (Q + v * v) + 6.0 * (o * v) ** 2 * Q + 2.0 * (o * v) ** 3
    ova, ovb = oa * va, ob * vb
    f += 2.0 * ova * ovb * Q + 2.0 * (oa * ob) ** 2 * (Q + va * vb) + 2.0 * (va * vb) ** 2 * (Q + oa * ob) \
        + 2.0 * ova * ovb * (ova + ovb)
    return f


def run_ump3(args, torch, hip, syn, w):
    """CD-UMP3 energy components (cholesky_ump3.cc:160-637) from device MO
    blocks of both spins; the transforms are made before timing."""
    ctx = hip.context(0)
    Q, n, oa, ob = w["Q"], w["n"], w["oa"], w["ob"]
    va, vb = n - oa, n - ob
    log(f"[bench] generating {args.workload}: Q={Q} n={n} oa={oa} va={va} ob={ob} vb={vb}")
    L, spa, spb = syn.workload(args.workload)
    dL = torch.from_numpy(L).cuda()
    del L
    blocks = []
    for (C, eo, ev), o in ((spa, oa), (spb, ob)):
        Co = torch.from_numpy(np.ascontiguousarray(C[:, :o])).cuda()
        Cv = torch.from_numpy(np.ascontiguousarray(C[:, o:])).cuda()
        blocks.append((ctx.transform(dL, Co, Cv), ctx.transform(dL, Co, Co), ctx.transform(dL, Cv, Cv),
                       torch.from_numpy(eo).cuda(), torch.from_numpy(ev).cuda()))
    (Ba, Looa, Lvva, eoa, eva), (Bb, Loob, Lvvb, eob, evb) = blocks
    sec, out, prof = timed(torch, ctx, lambda: ctx.ump3(Ba, Bb, Looa, Loob, Lvva, Lvvb, eoa, eva, eob, evb),
                           args.steps, args.warmup)
    flop = ump3_flops(Q, oa, va, ob, vb)
    log(f"[bench] ump3: {sec * 1e3:.1f} ms/step, {flop / sec / 1e12:.1f} TFLOP/s FP64, E3 = {list(out)!r}, "
        f"GEMMs {prof[1]}")
    config = {"workload": f"{args.workload}: CD-UMP3 energy components", "Q": Q, "n": n, "oa": oa, "va": va,
              "ob": ob, "vb": vb, "parallelism": "single GPU"}
    line = base_line(args, sec, "f64", config, flop, prof["step_ms"])
    achieved = flop / sec / 1e12
    line["roofline"] = {"kernel": "FP64 MFMA GEMMs ((ac|bd) build, ladders, rings)", "bound": "mfma",
                        "achieved": achieved, "peak": 78.6, "unit": "TFLOP/s", "frac": achieved / 78.6,
                        "traffic": None}
    line["energy_components"] = list(out)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16x4", help="fp64 | bf16xK (K slices)")
    ap.add_argument("--compare", default="fp64", help="second precision timed beside it ('' for none)")
    ap.add_argument("--workload", default="mp2_o100")
    ap.add_argument("--cpu-pairs", type=int, default=100,
                    help="ordered pairs (at most one per first index) timed by the CPU E+grad leg")
    ap.add_argument("--cpu-ref-pairs", type=int, default=100,
                    help="of those, pairs timed by the reference's energy-only loop")
    ap.add_argument("--zblock", type=float, default=0,
                    help="cap (doubles) of the gradient amplitude block per occupied block pair "
                         "(0: the library default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-energy-only", action="store_true",
                    help="skip the energy-only leg of the RMP2 workloads")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB moved by one plain device kernel before timing (a known-size "
                         "kernel for scaling PMC byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the measured value is one GPU's throughput, and "
                                   "runs on more than one GPU are out of scope",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from mshqc_amd import hip, synthetic as syn

    w = syn.WORKLOADS[args.workload]
    if w["kind"] in ("rmp2", "ump2"):
        line = run_mp2(args, torch, hip, syn, w)
    elif w["kind"] == "omp2":
        line = run_omp2(args, torch, hip, syn, w)
    elif w["kind"] == "cd":
        line = run_cd(args, torch, hip, syn, w)
    elif w["kind"] == "rmp3":
        line = run_rmp3(args, torch, hip, syn, w)
    elif w["kind"] == "ump3":
        line = run_ump3(args, torch, hip, syn, w)
    elif w["kind"] == "transform":
        line = run_transform(args, torch, hip, syn, w)
    else:
        line = run_jk(args, torch, hip, syn, w)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
