#!/usr/bin/env python3
"""ScatCov forward + VJP evaluations per second on one MI355X.

Workload: one HEALPix NESTED map at nside 1024, J = 9 scales, L = 4
orientations, kernel 5, the complete plan (11354 coefficients).  A step is one
forward sweep plus one VJP of the synthesis loss sum |S(x) - S(t)|^2 (its
cotangent 2 (S(x) - S(t)) computed in the step), with the maps resident in HBM.
The map x and the target t follow the lognormal recipe: g ~ N(0, 1) at nside
32, copied to the NESTED children, plus N(0, 0.1^2) per pixel, x = exp(g / 2)
- mean (seeds 1 and 2).

The headline value is the mixed-precision engine (FP32 bulk fields and
operators, FP64 sky-wide sums); the plain FP64 engine runs the same steps for
the ratio, and its statistics are the reference for the mixed path's error.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "stl-dev-mpi_amd"))

METRIC = "fwd+VJP evals/s, nside 1024 L=4: mixed vs FP64 GPU at ≤1e-6 rel err; HBM GB/s"
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published
FAMILIES = ["mean", "var", "S1", "S2", "S3", "S4"]
ERROR_BAR = 1e-6          # the metric's relative-error bar, per statistic family and on the gradient
SPEED_RATIO_TARGET = 3.0  # mixed / FP64 evaluations per second asked for
HOST_THREADS = 16         # host threads per GPU on a one-GPU MI355X host (its CPU share)
HEADLINE = ROOT / "tests" / "golden" / "headline"


def lognormal_map(nside, seed, base=32, sigma=0.1):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(12 * base * base)
    g = np.repeat(g, (nside // base) ** 2) + sigma * rng.standard_normal(12 * nside * nside)
    x = np.exp(0.5 * g)
    return x - x.mean()


def family_errors(values, ref, plan):
    fam = np.array([FAMILIES.index(c["family"]) for c in plan])
    out = {}
    for f, name in enumerate(FAMILIES):
        sel = fam == f
        scale = np.abs(ref[..., sel]).max()
        if name == "mean":
            scale = np.sqrt(np.abs(ref[..., fam == 1]).max())
        out[name] = float(np.abs(values[..., sel] - ref[..., sel]).max() / max(scale, 1e-300))
    return out


def run_engine(precision, nside, J, x, t, steps, warmup, timing, fields="auto", extra_cot=None):
    """Timed forward + VJP steps of one engine; with extra_cot, one more VJP of that cotangent
    on the last forward (for the parity report, after the timed region)."""
    import torch
    from stl_port import NativeEngine

    eng = NativeEngine(nside, J, batch=1, backend="gpu", precision=precision, timing=timing, fields=fields)
    xd = torch.from_numpy(x).to("cuda:0").reshape(1, -1)
    td = torch.from_numpy(t).to("cuda:0").reshape(1, -1)
    target, _ = eng.forward_device(td, generation=0, objective="target")
    grad = torch.empty_like(xd)
    values = None

    def step(gen):
        nonlocal values
        values, _ = eng.forward_device(xd, generation=gen)
        cot = 2.0 * (values - target)
        eng.vjp_device(cot, grad, generation=gen)

    for i in range(warmup):
        step(i + 1)
    torch.cuda.synchronize()
    info0 = eng.info() if timing else None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step(warmup + i + 1)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    extra_grad = None
    if extra_cot is not None:
        g2 = torch.empty_like(xd)
        eng.vjp_device(extra_cot.reshape(values.shape), g2, generation=warmup + steps)
        extra_grad = g2.cpu().numpy().ravel()
    info1 = eng.info()
    timers = {}
    if timing:
        for name, v in info1["timers"].items():
            v0 = info0["timers"].get(name, {"ms": 0.0, "launches": 0, "bytes": 0.0})
            timers[name] = {k: v[k] - v0[k] for k in ("ms", "launches", "bytes")}
    out = dict(elapsed=elapsed, values=values.copy(), grad=grad.cpu().numpy().ravel(), plan=eng.plan,
               setup_s=info1["setup_seconds"], device_bytes=info1["device_memory"]["total_bytes"],
               peak_bytes=info1["device_memory"]["peak_bytes"], fields_bytes=info1["device_memory"]["fields_bytes"],
               fields=info1["fields"], timers=timers, conv_variants=info1["conv_variants"], extra_grad=extra_grad)
    eng.close()
    del xd, td, grad
    torch.cuda.empty_cache()
    return out


ENSEMBLE = [(128, 8), (256, 4), (512, 2)]  # (nside, maps), J = log2(nside) - 1


def run_ensemble(steps, warmup):
    """Forward + VJP of 8 maps at nside 128, 4 at 256 and 2 at 512 (lognormal
    recipe, per-map seeds), mixed precision: as one ensemble (each group's HIP
    graph on its own stream, the groups concurrent) and, for comparison, as
    three standalone batched engines run one after the other."""
    import torch
    from stl_port import NativeEngine, NativeEnsemble

    xs = [np.stack([lognormal_map(n, 100 + 10 * g + b) for b in range(m)]) for g, (n, m) in enumerate(ENSEMBLE)]
    ts = [np.stack([lognormal_map(n, 200 + 10 * g + b) for b in range(m)]) for g, (n, m) in enumerate(ENSEMBLE)]
    xd = [torch.from_numpy(x).cuda() for x in xs]
    td = [torch.from_numpy(t).cuda() for t in ts]
    out = [torch.empty_like(x) for x in xd]
    res = {}
    with NativeEnsemble(ENSEMBLE, precision="mixed") as ens:
        target = ens.forward_device(td)

        def step():
            vals = ens.forward_device(xd)
            ens.vjp_device([2.0 * (v - t) for v, t in zip(vals, target)], out)

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        res["ensemble_ms"] = 1e3 * (time.perf_counter() - t0) / steps
    engines = [NativeEngine(n, n.bit_length() - 2, batch=m, precision="mixed") for n, m in ENSEMBLE]
    targets = [e.forward_device(t, objective="target")[0] for e, t in zip(engines, td)]

    def step_separate(gen):
        for e, x, t, o in zip(engines, xd, targets, out):
            v, _ = e.forward_device(x, generation=gen)
            e.vjp_device(2.0 * (v - t), o, generation=gen)

    for i in range(warmup):
        step_separate(i + 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step_separate(warmup + i + 1)
    torch.cuda.synchronize()
    res["separate_engines_ms"] = 1e3 * (time.perf_counter() - t0) / steps
    for e in engines:
        e.close()
    maps = sum(m for _, m in ENSEMBLE)
    return {"groups": [{"nside": n, "maps": m, "J": n.bit_length() - 2} for n, m in ENSEMBLE],
            "evals_per_s": round(1e3 / res["ensemble_ms"], 2), "maps_per_s": round(maps * 1e3 / res["ensemble_ms"], 1),
            "ms_per_eval": round(res["ensemble_ms"], 3), "separate_engines_ms_per_eval": round(res["separate_engines_ms"], 3),
            "precision": "mixed", "steps": steps}


def run_synthesis(nside, J, iters):
    """L-BFGS synthesis on the device (stlx_synthesize): target the lognormal
    map (seed 1), start N(0, 1) (seed 2), ergodic weights (the loss is
    sum |S(x) - S(t)|^2 / var up to the reference's normalisation), history 10,
    mixed-precision forward + VJP, FP64 L-BFGS."""
    import torch
    from stl_port import NativeEngine, ergodic_weights

    t = lognormal_map(nside, 1)
    x0 = np.random.default_rng(2).standard_normal(12 * nside * nside)
    with NativeEngine(nside, J, precision="mixed") as eng:
        target, var = eng.forward_device(torch.from_numpy(t).cuda().reshape(1, -1), moments=True)
        w = ergodic_weights(eng.native_pixels(), var[0])
        x = torch.from_numpy(x0).cuda().reshape(1, -1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        losses = eng.synthesize(x, target, weights=w, max_iter=iters, history_size=10)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
    return {"nside": nside, "J": J, "iterations": iters, "evaluations": int(len(losses)), "seconds": round(sec, 3),
            "ms_per_evaluation": round(1e3 * sec / len(losses), 3), "first_loss": float(losses[0]),
            "last_loss": float(losses[-1]), "history": 10, "weights": "ergodic"}


def roofline(timers, steps, workload=None):
    """Dominant kernel (family at one level) by total time: algorithmic bytes
    per launch over its mean launch duration (HIP events on the engine stream).
    `traffic` is that kernel's HBM bytes per launch from profiles/traffic.json
    (PMC counters of the same workload), else null."""
    name, t = max(timers.items(), key=lambda kv: kv[1]["ms"])
    per_launch_ms = t["ms"] / t["launches"]
    bytes_per_launch = t["bytes"] / t["launches"]
    achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    traffic = None
    prof = ROOT / "profiles" / "traffic.json"
    if prof.is_file():
        measured = json.loads(prof.read_text())
        if workload is not None and measured.get("workload") == workload:
            traffic = measured.get(name, {}).get("bytes_per_launch")
    total_bytes = sum(v["bytes"] for v in timers.values()) / steps
    total_ms = sum(v["ms"] for v in timers.values()) / steps
    by_level = {}
    for k, v in timers.items():
        lev = k.rsplit("@L", 1)[-1] if "@L" in k else "-"
        by_level[lev] = by_level.get(lev, 0.0) + v["ms"] / steps
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
            "bytes_per_launch": bytes_per_launch, "ms_per_launch": per_launch_ms,
            "eval_algorithmic_GB": round(total_bytes / 1e9, 3), "eval_kernel_ms": round(total_ms, 3),
            "eval_GBps": round(total_bytes / (total_ms * 1e-3) / 1e9, 1),
            "kernel_ms_by_level": {k: round(v, 4) for k, v in sorted(by_level.items())},
            "kernels": {k: {"ms_per_launch": round(v["ms"] / v["launches"], 4),
                            "GBps": round(v["bytes"] / v["launches"] / (v["ms"] / v["launches"] * 1e-3) / 1e9, 1)}
                        for k, v in sorted(timers.items()) if v["launches"]}}


def compulsory_bytes(nside, J, elem_bytes, L=4):
    """Compulsory HBM bytes of one forward + VJP, independent of how the operators are stored:
    fields 33.5 N0 elements, the convolution coefficients (44.1 used offsets x 2L values per
    row of every level, read forward and reverse) and the decimation weights (30 per coarse row,
    7.5 per fine row transposed, with a 4-byte index), all at elem_bytes per value.  It gives
    51.05 GB (FP32) and 101.1 GB (FP64) at nside 1024, J = 9."""
    N = [12 * (nside >> n) ** 2 for n in range(J)]
    fields = elem_bytes * 33.5 * N[0]
    conv = 2 * elem_bytes * sum(N) * 44.1 * 2 * L
    dec = sum((30 * N[n + 1] + 7.5 * N[n]) * (4 + elem_bytes) for n in range(J - 1))
    return fields + conv + dec


def eval_roofline(nside, J, elem_bytes, evals_per_s, copy_gbps):
    """Whole-evaluation rate against the compulsory bytes: achieved GB/s and its fractions of
    the 8 TB/s spec and of the measured device copy."""
    gb = compulsory_bytes(nside, J, elem_bytes) / 1e9
    achieved = gb * evals_per_s
    out = {"compulsory_GB": round(gb, 2), "achieved_GBps": round(achieved, 1),
           "frac_of_peak": round(achieved / HBM_PEAK_GBPS, 4)}
    if copy_gbps:
        out["frac_of_measured_copy"] = round(achieved / copy_gbps, 4)
    return out


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def level_work(nside, J):
    """Work of one forward + VJP in units of pixel x branch-row pairs: sum_n N_n (n + 1)^2 (the
    S3/S4 products grow with the square of a level's branch rows, the convolutions with their
    number; the pairs term dominates from J = 5 on)."""
    return float(sum(12 * (nside >> n) ** 2 * (n + 1) ** 2 for n in range(J)))


def cpu_baseline():
    """The FP64 oracle (the C++ restatement of the reference's CPU engine; its geometry rebuilt
    at every application, like the reference's per-block rebuild, crates/stl-core/src/
    engine.rs:837-862) on the host: forward + VJP of the lognormal maps timed at nside 256
    (J 7) and 512 (J 8) on HOST_THREADS threads (about a minute); a power law t = c W^p in the
    work W of level_work through the two points gives the nside 1024, J 9 rate (extrapolated,
    labelled so).  The reference's own per-core anchor is timed too: nside 64, L = 2, J = 5 on
    one thread (3.26 s per evaluation on one EPYC 9655 core, docs/validation.md:162)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    threads = min(HOST_THREADS, os.cpu_count() or 1)

    def timed(nside, J, L=4):
        x, t = lognormal_map(nside, 1), lognormal_map(nside, 2)
        e = oracle.Engine(nside, J, L=L, cache_rows=0)
        e.load(t)
        target = e.forward()
        t0 = time.perf_counter()
        e.load(x)
        v = e.forward()
        e.gradient(2.0 * (v - target))
        return time.perf_counter() - t0

    oracle.set_threads(1)
    anchor = timed(64, 5, L=2)
    oracle.set_threads(threads)
    points = [(256, 7, timed(256, 7)), (512, 8, timed(512, 8))]
    (n1, j1, t1), (n2, j2, t2) = points
    p = np.log(t2 / t1) / np.log(level_work(n2, j2) / level_work(n1, j1))
    s_eval = float(t2 * (level_work(1024, 9) / level_work(n2, j2)) ** p)
    return {"value": round(1.0 / s_eval, 6), "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"forward+VJP of the FP64 oracle (geometry rebuilt per application) on {threads} threads of "
                      f"{os.cpu_count()} ({cpu_model()}): measured "
                      + ", ".join(f"nside {n} J {j}: {t:.2f} s" for n, j, t in points)
                      + f"; nside 1024 J 9 extrapolated by the power law W^{p:.3f} through them: {s_eval:.0f} s/eval",
            "measured": [{"nside": n, "J": j, "s_per_eval": round(t, 3)} for n, j, t in points],
            "extrapolated": True, "extrapolated_nside1024_s_per_eval": round(s_eval, 1),
            "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count(),
            "anchor": {"nside": 64, "L": 2, "J": 5, "threads": 1, "s_per_eval": round(anchor, 3),
                       "reference_s_per_eval": 3.26, "reference_source": "docs/validation.md:162 (1 EPYC 9655 core)",
                       "ratio": round(anchor / 3.26, 3)}}


def headline_parity(values, grad, plan, nside, J):
    """Mixed engine against the FP64 oracle's vectors of the same map (tests/golden/headline,
    written by its make_vectors.py): per-family errors of the statistics and the errors of the
    gradient of the vectors' cotangent, projected on 8 directions and at a pixel sample."""
    case = HEADLINE / f"bench_n{nside}"
    if not (case / "meta.json").is_file():
        return None
    meta = json.loads((case / "meta.json").read_text())
    if meta["J"] != J:
        return None
    sys.path.insert(0, str(HEADLINE))
    import recipe
    ref = np.load(case / "values.npy")
    out = {"statistics": family_errors(values.reshape(-1), ref, plan)}
    proj = np.array([float(grad @ recipe.direction(nside, i)) for i in range(recipe.NDIR)])
    rproj = np.load(case / "grad_proj.npy")
    out["gradient_projections"] = float(np.abs(proj - rproj).max() / np.abs(rproj).max())
    ids = np.load(case / "grad_ids.npy")
    out["gradient_sample"] = float(np.abs(grad[ids] - np.load(case / "grad_sample.npy")).max() / meta["grad_max_abs"])
    return out


def copy_bandwidth(gib=2, reps=5):
    """Measured HBM bandwidth of torch's device-to-device copy (bytes read + written / time)."""
    import torch
    n = (gib << 30) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda:0").uniform_()
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return round(2.0 * n * 4 / (ms * 1e-3) / 1e9, 1)


def run_2048(fields):
    """The largest single-map configuration: nside 2048, J = 10, mixed precision, 5 timed steps."""
    x2, t2 = lognormal_map(2048, 3), lognormal_map(2048, 4)
    r2 = run_engine("mixed", 2048, 10, x2, t2, 5, 1, timing=False, fields=fields)
    return {"nside": 2048, "J": 10, "precision": "mixed", "fields": r2["fields"],
            "evals_per_s": round(5 / r2["elapsed"], 3), "ms_per_eval": round(1e3 * r2["elapsed"] / 5, 3),
            "setup_s": round(r2["setup_s"], 3), "device_bytes": r2["device_bytes"],
            "peak_device_bytes": r2["peak_bytes"], "fields_bytes": r2["fields_bytes"],
            "coefficients": len(r2["plan"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nside", type=int, default=1024)
    ap.add_argument("--J", type=int, default=9)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-2048", action="store_true", help="skip the nside 2048 (J=10) single-map rates")
    ap.add_argument("--synthesis-iters", type=int, default=20,
                    help="L-BFGS iterations of the synthesis measurement at the bench nside (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for {args.gpus}); the whole sky fits one "
              f"MI355X, and the multi-GPU sky split is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    import torch
    from stl_port.native import stream_copy_gbps
    assert torch.cuda.is_available(), "bench.py needs a GPU"

    x, t = lognormal_map(args.nside, 1), lognormal_map(args.nside, 2)
    # The oracle vectors' cotangent of this map (for the parity report, outside the timed steps).
    sys.path.insert(0, str(HEADLINE))
    import recipe
    # Throughput with the forward and the VJP replayed as HIP graphs; then the
    # same steps launched eagerly with HIP events around every kernel, for the
    # per-kernel durations of the roofline.
    mixed = run_engine("mixed", args.nside, args.J, x, t, args.steps, args.warmup, timing=False)
    extra = recipe.cotangent(len(mixed["plan"]), 1)
    print(f"[bench] mixed: {args.steps / mixed['elapsed']:.2f} evals/s", file=sys.stderr, flush=True)
    tsteps = max(3, min(args.steps, 10))
    mixed_t = run_engine("mixed", args.nside, args.J, x, t, tsteps, 1, timing=True, extra_cot=extra)
    print(f"[bench] mixed, per-kernel events: {tsteps / mixed_t['elapsed']:.2f} evals/s", file=sys.stderr,
          flush=True)
    fp64 = fp64_t = None
    if not args.no_fp64:
        fp64 = run_engine("fp64", args.nside, args.J, x, t, args.steps, args.warmup, timing=False)
        print(f"[bench] fp64: {args.steps / fp64['elapsed']:.2f} evals/s", file=sys.stderr, flush=True)
        fp64_t = run_engine("fp64", args.nside, args.J, x, t, tsteps, 1, timing=True, extra_cot=extra)
    ensemble = None if args.no_ensemble else run_ensemble(max(3, args.steps // 2), 2)
    if ensemble:
        print(f"[bench] ensemble: {ensemble['evals_per_s']} evals/s", file=sys.stderr, flush=True)
    synthesis = run_synthesis(args.nside, args.J, args.synthesis_iters) if args.synthesis_iters > 0 else None
    if synthesis:
        print(f"[bench] synthesis: {synthesis['ms_per_evaluation']} ms/evaluation", file=sys.stderr, flush=True)
    big = None
    if not args.no_2048:
        # cfg: nside 2048 with the pyramid levels recomputed in the reverse; the keep store beside it.
        big = run_2048("recompute")
        print(f"[bench] nside 2048 (recompute): {big['evals_per_s']} evals/s", file=sys.stderr, flush=True)
        big["keep"] = run_2048("keep")
        print(f"[bench] nside 2048 (keep): {big['keep']['evals_per_s']} evals/s", file=sys.stderr, flush=True)
    copy_gbps = copy_bandwidth()
    stream_gbps = round(stream_copy_gbps(), 1)
    print(f"[bench] copy bandwidth: torch {copy_gbps} GB/s, stream kernel {stream_gbps} GB/s", file=sys.stderr,
          flush=True)
    cpu = None if args.no_cpu else cpu_baseline()

    value = args.steps / mixed["elapsed"]
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mixed["elapsed"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk, f64 sums",
        "data": "synthetic: lognormal HEALPix NESTED maps (x seed 1, target seed 2)",
        "config": {"workload": f"ScatCov forward+VJP of sum|S(x)-S(t)|^2, HEALPix NESTED nside {args.nside}, "
                               f"J={args.J}, L=4, kernel 5, 1 map, {len(mixed['plan'])} coefficients",
                   "nside": args.nside, "J": args.J, "L": 4, "batch": 1, "precision": "mixed",
                   "fields": mixed["fields"]},
        "roofline": roofline(mixed_t["timers"], tsteps, {"nside": args.nside, "J": args.J, "precision": "mixed"}),
        "cpu_baseline": cpu,
        "setup_s": round(mixed["setup_s"], 3),
        "device_bytes": mixed["device_bytes"],
        "peak_device_bytes": mixed["peak_bytes"],
        "ms_per_step_with_kernel_events": round(1e3 * mixed_t["elapsed"] / tsteps, 3),
        # per level: forward z, forward branch, adjoint z, adjoint branch kernel variant
        "conv_variants": mixed["conv_variants"],
        "ensemble": ensemble,
        "synthesis": synthesis,
        "nside2048": big,
    }
    r = line["roofline"]
    r["copy_GBps_measured"] = copy_gbps
    r["stream_copy_GBps_measured"] = stream_gbps
    r["frac_of_measured_copy"] = round(r["achieved"] / max(copy_gbps, stream_gbps), 4)
    attainable = max(copy_gbps, stream_gbps)
    line["eval_roofline"] = eval_roofline(args.nside, args.J, 4, value, attainable)
    if big is not None:
        big["eval_roofline"] = eval_roofline(2048, 10, 4, big["evals_per_s"], attainable)
        big["keep"]["eval_roofline"] = eval_roofline(2048, 10, 4, big["keep"]["evals_per_s"], attainable)
    # Mixed engine against the FP64 oracle's vectors of this map, and the bar of the metric.
    vs_oracle = headline_parity(mixed_t["values"], mixed_t["extra_grad"], mixed["plan"], args.nside, args.J)
    if vs_oracle is not None:
        line["mixed_error_vs_oracle"] = vs_oracle
        worst_grad = max(vs_oracle["gradient_projections"], vs_oracle["gradient_sample"])
        line["error_bar"] = {"bar": ERROR_BAR, "against": "FP64 oracle vectors of the bench map",
                             "statistics_within": max(vs_oracle["statistics"].values()) <= ERROR_BAR,
                             "gradient_within": worst_grad <= ERROR_BAR}
    if fp64 is not None:
        v64 = args.steps / fp64["elapsed"]
        line["fp64"] = {"value": round(v64, 3), "ms_per_step": round(1e3 * fp64["elapsed"] / args.steps, 3),
                        "roofline": roofline(fp64_t["timers"], tsteps), "setup_s": round(fp64["setup_s"], 3),
                        "device_bytes": fp64["device_bytes"],
                        "eval_roofline": eval_roofline(args.nside, args.J, 8, v64, attainable)}
        p64 = headline_parity(fp64_t["values"], fp64_t["extra_grad"], fp64["plan"], args.nside, args.J)
        if p64 is not None:
            line["fp64"]["error_vs_oracle"] = p64
        line["mixed_over_fp64"] = round(value / v64, 3)
        line["mixed_over_fp64_target"] = SPEED_RATIO_TARGET
        line["mixed_over_fp64_meets_target"] = value / v64 >= SPEED_RATIO_TARGET
        line["mixed_error_vs_fp64"] = family_errors(mixed["values"], fp64["values"], mixed["plan"])
        g64 = fp64["grad"]
        line["mixed_error_vs_fp64"]["gradient"] = float(np.abs(mixed["grad"] - g64).max() / np.abs(g64).max())
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
