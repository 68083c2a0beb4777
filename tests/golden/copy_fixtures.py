"""Copy the subset of the reference's golden fixtures used by the tests.

The reference dumps its fixtures from STL-Dev with `tools/dump_reference.py`
(`fixtures/<suite>/<case>/*.npy` + `meta.json`, one `index.json` per suite).
They are data (inputs and expected outputs); this script copies a subset of
them here unchanged, with the `index.json` filtered to the copied cases.

Usage: python tests/golden/copy_fixtures.py [/path/to/reference/fixtures]
"""
import json
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent

# Cases kept per suite: every stats and synthesis case; geometry and block cases that pin
# each primitive at small and large nside while keeping the copy small.
KEEP = {
    "stats": lambda name: True,
    "geometry": lambda name: (
        name in {"centres_n8", "centres_n32", "centres_n1024", "ring2nest_o1", "ring2nest_o3", "ring2nest_o10",
                 "interp_n1", "interp_n2", "interp_n8", "interp_n32", "interp_n1024",
                 "down_n4_s0", "down_n8_s0", "down_n8_s256", "down_n8_s448", "down_n8_s704",
                 "down_n32_s0", "down_n32_s768", "down_n32_s12032", "down_n128_s0", "down_n128_s196352",
                 "down_n1024_s0", "down_n1024_s5241856"}
        or name.startswith(("rotation_", "kernel_grid_", "neighbours_"))
        or (name.startswith("conv_") and ("_k5_" in name or name in {
            "conv_n8_b64_s256_k3_G2_cosmo", "conv_n32_b256_s4864_k3_G2_cosmo",
            "conv_n128_b256_s65536_k3_G2_cosmo", "conv_n1024_b256_s4194304_k3_G2_cosmo"}))
    ),
    "block": lambda name: name.startswith(("wavelet_", "decimate_n16", "decimate_n32", "level_n16", "level_n32",
                                           "wave_n16", "wave_n32")),
    "synthesis": lambda name: True,
}

# Single data files copied unchanged: the nside-64 NESTED convergence map of the reference's
# statistics and synthesis examples (Takahashi et al. 2017, r000, zs1).
FILES = ["takahashi/takahashi_r000_zs1_nside64_nested.npy"]


def main(src):
    src = Path(src)
    for suite, keep in KEEP.items():
        index = json.loads((src / suite / "index.json").read_text())
        kept = [case for case in index if keep(case["name"])]
        dst = HERE / suite
        dst.mkdir(exist_ok=True)
        for case in kept:
            if (dst / case["name"]).exists():
                shutil.rmtree(dst / case["name"])
            shutil.copytree(src / suite / case["name"], dst / case["name"])
            for f in (dst / case["name"]).iterdir():
                f.chmod(0o644)
            (dst / case["name"]).chmod(0o755)
        (dst / "index.json").write_text(json.dumps(kept, indent=1) + "\n")
        print(f"{suite}: {len(kept)} of {len(index)} cases")
    for name in FILES:
        dst = HERE / name
        dst.parent.mkdir(exist_ok=True)
        shutil.copyfile(src / name, dst)
        dst.chmod(0o644)
        print(f"{name}: copied")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/fixtures")
