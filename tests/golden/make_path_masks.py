"""Extract the PATH_MASKS table (voxelis/src/spatial/voxtree.rs:41-105) from the
reference source text into tests/golden/path_masks.json.

Run in the build container where the reference is mounted; the JSON it writes is
the committed fixture (data only: the 7x6 table of integers).
"""
import json
import os
import re
import sys

SRC = "/root/reference/voxelis/src/spatial/voxtree.rs"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "path_masks.json")


def main():
    text = open(SRC).read()
    start = text.index("= [", text.index("const PATH_MASKS"))
    end = text.index("\n];", start)
    body = text[start:end]
    lits = re.findall(r"0b([01_]+)", body)
    vals = [int(l.replace("_", ""), 2) for l in lits]
    if len(vals) != 7 * 6:
        sys.exit(f"expected 42 entries, found {len(vals)}")
    table = [vals[i * 6:(i + 1) * 6] for i in range(7)]
    json.dump({"source": "voxelis/src/spatial/voxtree.rs:41-105", "path_masks": table},
              open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
