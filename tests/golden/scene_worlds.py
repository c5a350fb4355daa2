"""Loads the packed reference scene fixtures written by make_scene_worlds.py."""
import glob
import os

import numpy as np

from steel_webgpu_amd import pack

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("header", "statF", "statU", "poses0", "vel0", "art0")


def names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "scenes", "*.npz")))


def load(name):
    """A PackedWorld holding the fixture's arrays (enough for Batch and the oracle)."""
    z = np.load(os.path.join(HERE, "scenes", name + ".npz"))  # allow_pickle stays False
    pw = pack.PackedWorld()
    for k in FIELDS:
        setattr(pw, k, z[k])
    pw.NB, pw.NS, pw.NA, pw.NL, pw.ND = (pw.h(f) for f in ("W_NB", "W_NS", "W_NA", "W_NL", "W_ND"))
    return pw
