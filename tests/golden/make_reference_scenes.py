"""Writes tests/golden/scenes/: the reference's five scene files
(data/scenes/*.json), kept byte for byte as loader fixtures (scene data:
configuration, materials and blocks of a WC dam break, a DP column
collapse, a hydrostatic tank and two DP scenes with static rigid
intruders).  Run once in a checkout that has the reference beside it."""
import json
import os
import shutil
import sys

SRC = "/root/reference/data/scenes"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scenes")


def main(src=SRC):
    os.makedirs(DST, exist_ok=True)
    for name in sorted(os.listdir(src)):
        if not name.endswith(".json"):
            continue
        with open(os.path.join(src, name)) as f:
            json.load(f)  # a scene file parses as JSON
        shutil.copyfile(os.path.join(src, name), os.path.join(DST, name))
        print(name)


if __name__ == "__main__":
    main(*sys.argv[1:])
