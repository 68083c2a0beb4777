"""stl-synth: synthesis of maps whose ScatCov statistics match a target map's.

Restates the single-group path of the reference's stl-synth (crates/stl-cli/src/lib.rs:39-160
usage, :562-1313 run) over the MI355X engine: the target (NESTED `.npy` / `.fits`) standardised
by its mean and std, initial maps of white noise or a coarser `--init` (coarse to fine), the
target statistics and optional ergodic and family weights, L-BFGS on the device
(NativeEngine.synthesize: one map, or an ensemble whose mean statistics match the target), and
the reference's outputs: PREFIX.json, PREFIX_target.fits, PREFIX_synthesis.fits (one map),
PREFIX[_mMMM]_synthesis_nested.npy and PREFIX[_mMMM]_synthesis_standardized.npy.

The extra loss terms of the reference's stl-synth (scatter, level power, kappa, kurtosis,
Minkowski, marginal), its preconditioners, --freeze-nside, --groups > 1 and the conditional
--init noise are refused with a message.

    python -m stl_port.synth --target map_nside256.npy --out run --niter 50 --ensemble 4
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

from .maps import SlicedMap, initial_maps, moments, nside_of, read_map
from .native import NativeEngine, ergodic_weights
from .stats import default_j

FAMILIES = ("mean", "var", "S1", "S2", "S3", "S4")

USAGE = """\
stl-synth --target MAP [options]

  --target MAP          NESTED full-sky map (.npy or HEALPix .fits), standardised
  --init MAP            initial map (NESTED, nside <= target); default white noise; with
                        --ensemble, '{m}' in MAP gives one init per member (3-digit index)
  --init-units U        'target' (default: standardised like the target) or 'standardized'
  --init-noise X        noise std added to --init: 'auto' (default, RMS of the target's
                        missing small scales when the init nside is lower), a number, or 0
  --seed N              white-noise seed (default 11)
  --J N                 scales (default log2(nside) - 1)
  --niter N             L-BFGS iterations (default 50)
  --history-size N      L-BFGS history (default 10)
  --weights W           'none' (default, sum |S - S_target|^2) or 'ergodic' (reduced chi^2)
  --family-weights F    extra factors on whole families, e.g. 'S1=10,S2=10'
  --ensemble B          B maps whose MEAN statistics match the target; member m starts from
                        seed + m (default 1)
  --out PREFIX          output prefix (default synthesis_rust_nside{nside})
  --precision P         'mixed' (default) or 'fp64'
  --device N            GPU (default 0)
  --quiet
  accepted without effect: --gpu, --ranks N, --block-pixels N, --conv-chunk N, --L 4,
                        --kernel-size 5, --groups 1
"""

_ZERO_TERMS = ("--scatter-weight", "--level-power-weight", "--kappa-weight", "--kappa-s3-weight",
               "--kurtosis-weight", "--minkowski-weight", "--marginal-weight", "--marginal-tail-weight")


def parse_family_weights(v):
    """'S1=10,S2=10' -> [("S1", 10.0), ("S2", 10.0)] (lib.rs:162-175)."""
    out = []
    for item in v.split(","):
        if "=" not in item:
            raise ValueError(f"--family-weights: '{item}' is not FAMILY=FACTOR")
        name, f = item.split("=", 1)
        if name.strip() not in FAMILIES:
            raise ValueError(f"--family-weights: unknown family '{name}'")
        try:
            factor = float(f.strip())
        except ValueError:
            raise ValueError(f"--family-weights: '{f}' is not a number") from None
        if not (np.isfinite(factor) and factor >= 0.0):
            raise ValueError(f"--family-weights: factor {factor} must be finite and >= 0")
        out.append((name.strip(), factor))
    return out


def parse_args(argv):
    a = dict(target=None, init=None, init_units="target", init_noise="auto", seed=11, J=None, niter=50,
             history_size=10, weights="none", family_weights=[], ensemble=1, out=None, precision="mixed",
             device=0, quiet=False)
    it = iter(argv)
    for k in it:
        def value():
            v = next(it, None)
            if v is None:
                raise ValueError(f"{k} needs a value")
            return v
        if k == "--target":
            a["target"] = value()
        elif k == "--init":
            a["init"] = value()
        elif k == "--init-units":
            a["init_units"] = value()
            if a["init_units"] not in ("target", "standardized"):
                raise ValueError("--init-units: 'target' or 'standardized'")
        elif k == "--init-noise":
            v = value()
            if v == "conditional":
                raise ValueError("--init-noise conditional is not supported by this library")
            a["init_noise"] = v if v == "auto" else float(v)
        elif k == "--seed":
            a["seed"] = int(value())
        elif k == "--J":
            a["J"] = int(value())
        elif k == "--niter":
            a["niter"] = int(value())
        elif k == "--history-size":
            a["history_size"] = int(value())
        elif k == "--weights":
            a["weights"] = value()
            if a["weights"] not in ("none", "ergodic"):
                raise ValueError("--weights: 'none' or 'ergodic'")
        elif k == "--family-weights":
            a["family_weights"] = parse_family_weights(value())
        elif k == "--ensemble":
            a["ensemble"] = int(value())
            if a["ensemble"] < 1:
                raise ValueError("--ensemble must be >= 1")
        elif k == "--out":
            a["out"] = value()
        elif k == "--precision":
            a["precision"] = value()
        elif k == "--device":
            a["device"] = int(value())
        elif k == "--quiet":
            a["quiet"] = True
        elif k == "--gpu":
            pass
        elif k in ("--ranks", "--block-pixels", "--conv-chunk"):
            value()
        elif k in ("--L", "--kernel-size", "--groups"):
            v = int(value())
            if v != {"--L": 4, "--kernel-size": 5, "--groups": 1}[k]:
                raise ValueError(f"{k} {v} is not supported by this library")
        elif k in _ZERO_TERMS:
            if float(value()) != 0.0:
                raise ValueError(f"{k}: this loss term is not supported by this library")
        else:
            raise ValueError(f"{k}: not supported by this library")
    if a["target"] is None:
        raise ValueError("--target is required")
    return a


def synthesize(args, library=None):
    """Runs one synthesis of the parsed options `args`; writes the outputs and returns the report."""
    import torch

    started = time.perf_counter()
    target_path = os.fspath(args["target"])
    raw = read_map(target_path, library)
    npix = raw.size
    nside = nside_of(npix)
    j = default_j(nside) if args["J"] is None else int(args["J"])
    members = int(args["ensemble"])
    mean, std = moments(raw)
    target = (raw - mean) / std
    del raw
    initial, init_info = initial_maps(target, members, args["seed"], args["init"], args["init_units"],
                                      args["init_noise"], mean, std, library)
    prec, dev = args["precision"], int(args["device"])
    probe = NativeEngine(nside, j, batch=1, precision=prec, device=dev, library=library)
    try:
        target_stats, target_var = probe.forward(target, moments=True)
        native = probe.native_pixels()
        plan = probe.plan
        gpu_info = {k: v for k, v in probe.info().items() if k in ("version", "precision", "device_memory")}
    finally:
        probe.close()
    ncoef = len(plan)
    weights = np.ones(ncoef)
    if args["weights"] == "ergodic":
        weights = ergodic_weights(native, target_var[0]) * members
    for fam, f in args["family_weights"]:
        weights[[c["family"] == fam for c in plan]] *= f
    engine = NativeEngine(nside, j, batch=members, precision=prec, device=dev, library=library)
    try:
        setup_seconds = time.perf_counter() - started
        x = torch.tensor(initial, dtype=torch.float64, device=f"cuda:{dev}")
        t0 = time.perf_counter()
        history = engine.synthesize(x, target_stats[0] if members > 1 else target_stats, weights,
                                    max_iter=args["niter"], history_size=args["history_size"],
                                    ensemble=members > 1)
        seconds = time.perf_counter() - t0
        xs = x.cpu().numpy()
        values, _ = engine.forward(xs)
    finally:
        engine.close()
    resid = values.mean(axis=0) - target_stats[0] if members > 1 else values - target_stats
    final_loss = float((weights * np.abs(resid) ** 2).sum())

    prefix = args["out"] or f"synthesis_rust_nside{nside}"

    def member_file(m, suffix):
        return f"{prefix}{suffix}" if members == 1 else f"{prefix}_m{m:03d}{suffix}"

    outputs = [(f"{prefix}_target.fits", "TARGET", target * std + mean)]
    for m in range(members):
        if members == 1:
            outputs.append((f"{prefix}_synthesis.fits", "SYNTHESIS", xs[m] * std + mean))
        outputs.append((member_file(m, "_synthesis_nested.npy"), "", xs[m] * std + mean))
        outputs.append((member_file(m, "_synthesis_standardized.npy"), "", xs[m]))
    for path, column, values_m in outputs:
        out = SlicedMap(path, False, library)
        out.create(npix, column)
        out.write(0, values_m)
    history = [float(v) for v in history]
    report = {
        "mode": "mi355x",
        "nside": nside,
        "target": {"file": target_path, "mean": mean, "std": std},
        "J": j, "L": 4, "kernel_size": 5, "fewer": False,
        "coefficients": ncoef,
        "niter": args["niter"],
        "evaluations": len(history),
        "variables": "map",
        "ranks": 1,
        "ensemble": {
            "members": members, "groups": 1, "ranks_per_group": 1,
            "loss": "sum_i w_i |mean over members of S_i - S_target,i|^2",
            "scatter_weight": 0.0,
            "member_files": [member_file(m, "_synthesis_nested.npy") for m in range(members)],
        },
        "backend": "gpu",
        "gpu": gpu_info,
        "setup_seconds": round(setup_seconds, 2),
        "seconds": round(seconds, 2),
        "seconds_per_evaluation": round(seconds / max(len(history), 1), 3),
        "loss_first": history[0] if history else None,
        "loss_last": history[-1] if history else None,
        "loss_final_map": final_loss,
        "weights": args["weights"],
        "family_weights": {f: w for f, w in args["family_weights"]},
        "loss_history": history,
        "fits": [f"{prefix}_target.fits", f"{prefix}_synthesis.fits"],
        "init": init_info,
    }
    with open(f"{prefix}.json", "w") as f:
        json.dump(report, f, indent=2)
    if not args["quiet"]:
        print(f"[stl-synth] {len(history)} evaluations, loss {report['loss_first']} -> {report['loss_last']}, "
              f"{seconds:.2f} s", file=sys.stderr)
    return report


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or "-h" in argv or "--help" in argv:
        print(USAGE, file=sys.stderr)
        return 0 if argv else 2
    try:
        args = parse_args(argv)
    except ValueError as e:
        print(f"{e}\n\n{USAGE}", file=sys.stderr)
        return 2
    try:
        synthesize(args)
    except (ValueError, RuntimeError, OSError) as e:
        print(f"stl-synth: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
