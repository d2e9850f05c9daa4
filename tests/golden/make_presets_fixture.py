"""Writes ref_presets_2d.fae: the 2D flame presets of the reference's preset
library that this engine renders (assets/presets.fae entries 0-4 and 7:
Grand Julian, Julian Disc, Plastic, Spherical3, Square Tile, Flower),
copied unchanged as a JSON array. Run in a container that holds the
reference checkout; the tests read only the written file."""
import json
import os
import sys

SRC = "/root/reference/assets/presets.fae"
KEEP = [0, 1, 2, 3, 4, 7]


def main():
    presets = json.load(open(SRC))
    out = [presets[i] for i in KEEP]
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_presets_2d.fae")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst, [p["flame"]["name"] for p in out], file=sys.stderr)


if __name__ == "__main__":
    main()
