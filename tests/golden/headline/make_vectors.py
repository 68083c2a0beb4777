"""FP64 oracle vectors at the benchmark's sizes (nside 128 ... 1024), for the GPU parity tests.

The oracle (oracle/stl_oracle.cpp, the FP64 restatement of the reference's CPU engine,
crates/stl-core/src/engine.rs) takes minutes per evaluation at nside 512 and 1024, too long
for a GPU test, so this script runs it once in the build container and commits what the
tests compare against.  Per map it writes, under tests/golden/headline/<case>/:

- values.npy     complex statistics [ncoef] (the complete plan, L = 4, kernel 5, cosmo);
- variance.npy   the spatial variances of the terms (forward with moments), for the cases
                 with moments = true;
- grad_proj.npy  the map gradient of Re sum(conj(cot) values) projected on 8 directions
                 (dot products with recipe.direction(nside, i), i = 0..7);
- grad_ids.npy / grad_sample.npy  the gradient at a fixed sample of pixels;
- meta.json      nside, J, seeds, ncoef, the SHA-256 of the map's bytes (the tests check it
                 before comparing) and the oracle's wall time.

The maps, cotangents, directions and pixel samples come from recipe.py, which the tests
import, so both sides draw the same numbers.  Cases:

- bench_n1024: the benchmark's map (lognormal recipe, seed 1), nside 1024, J = 9, moments;
- ens_n128_b{0..7}, ens_n256_b{0..3}, ens_n512_b{0,1}: the uneven ensemble's 14 maps
  (8 x nside 128, 4 x 256, 2 x 512; J = log2(nside) - 1; seeds 100 + 10 g + b).

Usage: python tests/golden/headline/make_vectors.py [case-prefix ...]  (threads:
STL_ORACLE_THREADS, default all).  Re-running overwrites the cases named.
"""
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(HERE))

import oracle  # noqa: E402
import recipe  # noqa: E402


def run_case(case):
    nside, J, seed = case["nside"], case["J"], case["seed"]
    x = recipe.lognormal(nside, seed)
    t0 = time.perf_counter()
    eng = oracle.Engine(nside, J)
    eng.load(x)
    if case["moments"]:
        values, variance = eng.forward(moments=True)
    else:
        values, variance = eng.forward(), None
    values = values.reshape(-1)
    t1 = time.perf_counter()
    cot = recipe.cotangent(values.size, seed)
    grad = eng.gradient(cot).reshape(-1)
    t2 = time.perf_counter()
    del eng
    out = HERE / case["name"]
    out.mkdir(exist_ok=True)
    np.save(out / "values.npy", values)
    if variance is not None:
        np.save(out / "variance.npy", variance.reshape(-1))
    np.save(out / "grad_proj.npy", np.array([float(grad @ recipe.direction(nside, i)) for i in range(recipe.NDIR)]))
    ids = recipe.sample_ids(nside)
    np.save(out / "grad_ids.npy", ids)
    np.save(out / "grad_sample.npy", grad[ids])
    meta = dict(case, ncoef=int(values.size), map_sha256=hashlib.sha256(x.tobytes()).hexdigest(),
                grad_max_abs=float(np.abs(grad).max()), oracle_forward_s=round(t1 - t0, 1),
                oracle_vjp_s=round(t2 - t1, 1), oracle_threads=oracle.lib().or_max_threads())
    (out / "meta.json").write_text(json.dumps(meta, indent=1) + "\n")
    print(f"{case['name']}: ncoef {values.size}, forward {t1 - t0:.1f} s, vjp {t2 - t1:.1f} s", flush=True)


def main(prefixes):
    if os.environ.get("STL_ORACLE_THREADS") is None:
        oracle.set_threads(os.cpu_count() or 1)
    for case in recipe.CASES:
        if not prefixes or any(case["name"].startswith(p) for p in prefixes):
            run_case(case)


if __name__ == "__main__":
    main(sys.argv[1:])
