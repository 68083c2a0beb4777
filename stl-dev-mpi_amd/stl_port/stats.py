"""stl-stats: the ScatCov statistics of NESTED maps as the reference's JSON report.

Restates crates/stl-cli/src/bin/stl-stats.rs and `statistics_of_maps`
(crates/stl-cli/src/lib.rs:1323-1407) over the MI355X engine: every map (one nside) is
standardised by its own mean and std, or by those of `--norm`, and evaluated with the plan of
stl-synth; the report holds the plan and, per map, its moments and the statistics (re, im).

    python -m stl_port.stats --out stats.json --norm target.npy -- m000.npy m001.npy
"""
from __future__ import annotations

import json
import os
import sys
import time

from .maps import map_len, moments, nside_of, read_map
from .native import NativeEngine

USAGE = """\
stl-stats --out STATS.json [--norm MAP] [engine options] -- MAP [MAP ...]

  --out FILE     JSON: the plan (family, scales, orientations, native level) and, per map, its
                 moments and statistics (re, im)
  --norm MAP     standardise every map with the mean and std of MAP (e.g. the target), so that
                 the statistics of all the maps are in the same units; default: its own moments
  engine options: --J N (default log2(nside) - 1), --L 4, --kernel-size 5, --precision mixed|fp64,
                 --device N, --quiet; accepted and without effect here (one GPU holds the sky):
                 --gpu, --ranks N, --block-pixels N, --conv-chunk N
"""

_PLAN_KEYS = ("family", "native_dg", "j1", "j2", "j3", "l1", "l2", "l3")


def default_j(nside):
    """Scales of a map of nside: log2(nside) - 1, at least 1 (lib.rs:1341)."""
    return max(int(nside).bit_length() - 2, 1)


def statistics_of_maps(maps, norm=None, J=None, L=4, kernel_size=5, fewer=False, precision="mixed", device=0,
                       quiet=True, library=None):
    """The stl-stats report (a dict, the reference's JSON layout) of the map files `maps`."""
    maps = [os.fspath(m) for m in maps]
    if not maps:
        raise ValueError("at least one map is required")
    if L != 4 or kernel_size != 5 or fewer:
        raise ValueError("this engine computes the plan of L = 4, kernel size 5 and the full S3/S4 "
                         "definitions only")
    npix = map_len(maps[0], library)
    nside = nside_of(npix)
    j = default_j(nside) if J is None else int(J)
    started = time.perf_counter()
    engine = NativeEngine(nside, j, batch=1, precision=precision, device=device, library=library)
    try:
        info = engine.info()
        gpu = {"device": int(device), "precision": precision, "library_sha256": info.get("library_sha256"),
               "device_memory": info.get("device_memory"), "version": info.get("version")}
        setup_seconds = time.perf_counter() - started

        def read(path):
            if map_len(path, library) != npix:
                raise ValueError(f"{path}: not the nside of {maps[0]}")
            raw = read_map(path, library)
            mean, std = moments(raw)
            return raw, mean, std

        fixed = None
        if norm is not None:
            _, m0, s0 = read(os.fspath(norm))
            fixed = (m0, s0)
        out = []
        for path in maps:
            raw, mean, std = read(path)
            m0, s0 = fixed if fixed is not None else (mean, std)
            x = (raw - m0) / s0
            t = time.perf_counter()
            values, _ = engine.forward(x)
            seconds = time.perf_counter() - t
            if not quiet:
                print(f"[stl-stats] {path}: {seconds:.1f} s", file=sys.stderr)
            st = values[0]
            out.append({"file": path, "mean": mean, "std": std, "seconds": seconds,
                        "re": [float(v) for v in st.real], "im": [float(v) for v in st.imag]})
        plan = [{k: c[k] for k in _PLAN_KEYS} for c in engine.plan]
    finally:
        engine.close()
    normalisation = ({"file": os.fspath(norm), "mean": fixed[0], "std": fixed[1]} if fixed is not None
                     else "own moments")
    return {"nside": nside, "J": j, "L": L, "kernel_size": kernel_size, "fewer": fewer,
            "normalisation": normalisation, "ranks": 1, "gpu": gpu, "setup_seconds": setup_seconds,
            "seconds": time.perf_counter() - started, "plan": plan, "maps": out}


def parse_engine_options(args):
    """Engine options shared by stl-stats and stl-synth: a dict of statistics_of_maps keywords."""
    opts = {}
    it = iter(args)
    for k in it:
        def value():
            v = next(it, None)
            if v is None:
                raise ValueError(f"{k} needs a value")
            return v
        if k == "--J":
            opts["J"] = int(value())
        elif k == "--L":
            opts["L"] = int(value())
        elif k == "--kernel-size":
            opts["kernel_size"] = int(value())
        elif k == "--fewer":
            opts["fewer"] = True
        elif k == "--precision":
            opts["precision"] = value()
        elif k == "--device":
            opts["device"] = int(value())
        elif k == "--quiet":
            opts["quiet"] = True
        elif k == "--gpu":
            pass
        elif k in ("--ranks", "--block-pixels", "--conv-chunk", "--gpu-block-rows", "--gpu-cache",
                   "--geometry-cache", "--gpu-conv", "--gpu-fields"):
            value()
        else:
            raise ValueError(f"unknown option {k}")
    return opts


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if "-h" in argv or "--help" in argv:
        print(USAGE, file=sys.stderr)
        return 0
    if "--" not in argv:
        print(USAGE, file=sys.stderr)
        return 2
    split = argv.index("--")
    maps = argv[split + 1:]
    out = norm = None
    engine = []
    it = iter(argv[:split])
    for k in it:
        if k == "--out":
            out = next(it, None)
        elif k == "--norm":
            norm = next(it, None)
        else:
            engine.append(k)
    if out is None or not maps:
        print(f"--out and at least one map are required\n\n{USAGE}", file=sys.stderr)
        return 2
    try:
        opts = parse_engine_options(engine)
    except ValueError as e:
        print(e, file=sys.stderr)
        return 2
    opts.setdefault("quiet", False)
    try:
        report = statistics_of_maps(maps, norm=norm, **opts)
    except (ValueError, RuntimeError, OSError) as e:
        print(f"stl-stats: {e}", file=sys.stderr)
        return 1
    try:
        with open(out, "w") as f:
            json.dump(report, f)
    except OSError as e:
        print(f"stl-stats: {out}: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
