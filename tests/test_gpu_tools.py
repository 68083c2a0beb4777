"""stl-stats and stl-synth command lines on the GPU engine, against the FP64 oracle.

The Takahashi nside-64 map of the reference's examples (tests/golden/takahashi) is the target:
stl-stats must report, for each map, the statistics the oracle computes for it standardised by
the --norm map (1e-10 relative per family, the FP64 parity bar), with the reference's report
layout (crates/stl-cli/src/lib.rs:1338-1407); stl-synth must start from the initial maps it
documents (coarse-to-fine --init with auto noise), report as first loss the ensemble loss the
oracle gives for them, decrease it, and write the reference's output files.
"""
import json

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN
from stl_port import maps as M
from stl_port import stats as S
from stl_port import synth as Y

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TARGET = GOLDEN / "takahashi" / "takahashi_r000_zs1_nside64_nested.npy"
FAMILY_NAMES = ["mean", "var", "S1", "S2", "S3", "S4"]


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def family_errors(values, ref, plan):
    fam = np.array([FAMILY_NAMES.index(c["family"]) for c in plan])
    out = {}
    for f, name in enumerate(FAMILY_NAMES):
        sel = fam == f
        scale = np.abs(ref[sel]).max()
        if name == "mean":
            scale = np.sqrt(np.abs(ref[fam == 1]).max())
        out[name] = float(np.abs(values[sel] - ref[sel]).max() / scale)
    return out


def test_stats_report_matches_oracle(tmp_path):
    target = np.load(TARGET)
    noise = M.pixel_normal(3, 0, target.size) * 2e-4 + 1e-5
    M.write_map(tmp_path / "noise.npy", noise)
    M.write_map(tmp_path / "target.fits", target, "TARGET", double=True)
    out = tmp_path / "stats.json"
    print("[gpu] stl-stats nside 64 fp64", flush=True)
    rc = S.main(["--out", str(out), "--norm", str(TARGET), "--precision", "fp64", "--quiet", "--",
                 str(TARGET), str(tmp_path / "noise.npy"), str(tmp_path / "target.fits")])
    assert rc == 0
    r = json.loads(out.read_text())
    assert (r["nside"], r["J"], r["L"], r["kernel_size"], r["fewer"], r["ranks"]) == (64, 5, 4, 5, False, 1)
    assert len(r["plan"]) == 2522 and set(r["plan"][0]) == {"family", "native_dg", "j1", "j2", "j3", "l1", "l2",
                                                             "l3"}
    mean, std = M.moments(target)
    assert r["normalisation"] == {"file": str(TARGET), "mean": pytest.approx(mean, rel=1e-15, abs=1e-30),
                                  "std": pytest.approx(std, rel=1e-15)}
    ref = O.Engine(64, 5)
    for entry, raw in zip(r["maps"], (target, noise, target)):
        assert entry["mean"] == pytest.approx(raw.mean(), rel=1e-12, abs=1e-18)
        assert entry["std"] == pytest.approx(raw.std(), rel=1e-12)
        ref.load((raw - mean) / std)
        expect = ref.forward()[0]
        got = np.array(entry["re"]) + 1j * np.array(entry["im"])
        errs = family_errors(got, expect, r["plan"])
        assert max(errs.values()) < 1e-10, (entry["file"], errs)


def test_synth_ensemble_from_coarse_init(tmp_path):
    target = np.load(TARGET)
    mean, std = M.moments(target)
    t = (target - mean) / std
    coarse = target.reshape(-1, 16).mean(axis=1)  # nside 16 by NESTED groups of 16
    M.write_map(tmp_path / "coarse.npy", coarse)
    prefix = str(tmp_path / "run")
    args = Y.parse_args(["--target", str(TARGET), "--init", str(tmp_path / "coarse.npy"), "--ensemble", "2",
                         "--niter", "8", "--seed", "4", "--family-weights", "S1=10,S3=0.5", "--precision", "fp64",
                         "--J", "5", "--out", prefix, "--quiet"])
    print("[gpu] stl-synth nside 64 ensemble 2 fp64", flush=True)
    report = Y.synthesize(args)
    hist = report["loss_history"]
    assert report["evaluations"] == len(hist) >= 2
    assert report["loss_last"] < 0.5 * report["loss_first"]
    # First loss: sum_i w_i |mean_m S_m,i - S_i(target)|^2 of the documented initial maps.
    x0, info = M.initial_maps(t, 2, 4, str(tmp_path / "coarse.npy"), "target", "auto", mean, std)
    assert info["nside"] == 16 and info["noise_std"] > 0
    ref = O.Engine(64, 5)
    ref.load(t)
    tstats = ref.forward()[0]
    plan = O.plan(5, 4)
    ref2 = O.Engine(64, 5, B=2)
    ref2.load(x0)
    s0 = ref2.forward()
    w = np.ones(len(tstats))
    fams = [FAMILY_NAMES[int(c[0])] for c in plan]  # column 0 of the oracle's plan: the family
    w[[f == "S1" for f in fams]] *= 10.0
    w[[f == "S3" for f in fams]] *= 0.5
    loss0 = float((w * np.abs(s0.mean(axis=0) - tstats) ** 2).sum())
    assert hist[0] == pytest.approx(loss0, rel=1e-9)
    # Outputs: the reference's files, in physical and standardised units.
    with open(f"{prefix}.json") as f:
        assert json.load(f)["ensemble"]["members"] == 2
    assert np.allclose(M.read_map(f"{prefix}_target.fits"), target, rtol=1e-6, atol=0)
    for m in range(2):
        nested = np.load(f"{prefix}_m{m:03d}_synthesis_nested.npy")
        stdz = np.load(f"{prefix}_m{m:03d}_synthesis_standardized.npy")
        assert nested.shape == (target.size,) and np.allclose(nested, stdz * std + mean, rtol=0, atol=1e-15)
    # The final loss is that of the written maps.
    x1 = np.stack([np.load(f"{prefix}_m{m:03d}_synthesis_standardized.npy") for m in range(2)])
    ref2.load(x1)
    loss1 = float((w * np.abs(ref2.forward().mean(axis=0) - tstats) ** 2).sum())
    assert report["loss_final_map"] == pytest.approx(loss1, rel=1e-8)
