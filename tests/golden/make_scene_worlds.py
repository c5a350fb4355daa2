"""Writes tests/golden/scenes/<scene>.npz: every shipped reference scene folder (assets/scenes) packed into the
world arrays the library and the oracle consume (header, statF, statU, initial poses, velocities and
articulation state).  Run in a container that holds the reference tree:

    python tests/golden/make_scene_worlds.py /root/reference/assets/scenes

The files are inputs (scene data in the packed layout), so GPU parity can step the real scenes on a machine
without the reference tree; the CPU test test_scene_fixtures_match_a_fresh_pack guards them against
packer changes.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from steel_webgpu_amd import pack  # noqa: E402
from steel_webgpu_amd.scene import loadScene  # noqa: E402

FIELDS = ("header", "statF", "statU", "poses0", "vel0", "art0")


def main(root):
    out = os.path.join(HERE, "scenes")
    os.makedirs(out, exist_ok=True)
    for name in sorted(os.listdir(root)):
        if name == "shared" or not os.path.isfile(os.path.join(root, name, "scene.json")):
            continue
        pw = pack.pack_scene(loadScene(os.path.join(root, name)))
        np.savez_compressed(os.path.join(out, name + ".npz"), **{k: getattr(pw, k) for k in FIELDS})
        print(name, pw.NB, pw.NS)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/assets/scenes")
