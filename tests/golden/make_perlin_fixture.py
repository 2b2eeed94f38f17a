"""Writes tests/golden/perlin_perm.json: the Perlin lattice permutation of
YMRT's noise textures (perlin_noise_permute_table, src/TextureEval.cuh:19-53),
read from the reference source as data.  Run in a container that has the
reference tree:  python tests/golden/make_perlin_fixture.py [reference root]

The table is stored there as 512 entries (the 256-entry permutation twice);
the fixture keeps all 512 so the test can check the repetition too.
"""
import json
import os
import re
import sys

ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(ROOT, "src", "TextureEval.cuh")


def main():
    text = open(SRC).read()
    m = re.search(r"perlin_noise_permute_table\s*\[[^\]]*\]\s*=\s*\{([^}]*)\}", text)
    if not m:
        raise SystemExit("permutation table not found in " + SRC)
    values = [int(v) for v in re.findall(r"\d+", m.group(1))]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "perlin_perm.json")
    with open(out, "w") as f:
        json.dump({"source": "src/TextureEval.cuh:19-53", "count": len(values), "table": values}, f)
        f.write("\n")
    print("wrote %d entries to %s" % (len(values), out))


if __name__ == "__main__":
    main()
