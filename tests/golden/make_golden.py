"""Writes tests/golden/oracle_hashes.json: SHA-256 of every output of the
oracle for the hand-made scenes and a 32-path version of the 1024x1024
workload. tests/test_golden.py checks the oracle still produces them.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from helpers import oracle_frame  # noqa: E402
from pathfinder_amd import synth  # noqa: E402

KINDS = {"microlines": oracle.MICROLINES, "fills": oracle.FILLS, "tiles": oracle.TILES,
         "zbuffer": oracle.ZBUFFER, "first_tile_map": oracle.FIRST_TILE_MAP,
         "alpha_tiles": oracle.ALPHA_TILES, "backdrops": oracle.BACKDROPS}


def scenes():
    out = dict(synth.tiny_scenes())
    out["small32"] = synth.small_scene(paths=32, size=256)
    return out


def digest(scene, size, bg):
    cmds, o = oracle_frame(scene, size, bg, threads=2)
    rec = {}
    for _, b in cmds["batches"]:
        for name, what in KINDS.items():
            a = o.read(b["batch_id"], what)
            rec["%s[%d]" % (name, b["batch_id"])] = [int(a.size),
                                                     hashlib.sha256(a.tobytes()).hexdigest()]
    m = o.read_mask()
    fb = o.read_framebuffer()
    rec["mask"] = [int(m.size), hashlib.sha256(m.tobytes()).hexdigest()]
    rec["framebuffer"] = [int(fb.size), hashlib.sha256(fb.tobytes()).hexdigest()]
    return rec


if __name__ == "__main__":
    out = {name: digest(*v) for name, v in sorted(scenes().items())}
    with open(os.path.join(HERE, "oracle_hashes.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", len(out), "scenes")
