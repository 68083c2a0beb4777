"""Write reference_symbols.json: the C symbols the reference's Python binding binds eagerly.

The reference binding (python/stl_ensemble/stl_port/native.py:55-116) sets argtypes on every
name of its signature table when it loads the library, so a replacement library must export
all of them.  This script reads that table as text and keeps only the names (data for the
library test); it imports nothing from the reference.

Usage: python tests/golden/binding/extract_symbols.py [/root/reference]
"""
import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main(root):
    text = (Path(root) / "python/stl_ensemble/stl_port/native.py").read_text().splitlines()
    table = "\n".join(text[54:116])  # lines 55-116: the signature table of _library()
    names = sorted(set(re.findall(r'"(stlp_[a-z0-9_]+)"\s*:', table)))
    out = {"source": "python/stl_ensemble/stl_port/native.py:55-116", "symbols": names}
    (HERE / "reference_symbols.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"{len(names)} symbols")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
