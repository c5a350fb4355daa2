"""Writes the oracle checkpoints the GPU parity tests start from: the FP32 oracle's full state of a world after a
number of frames from rest, so a test can step the GPU and the oracle from the same settled state without
re-running the oracle's approach to it.

    python tests/golden/make_checkpoints.py

Each checkpoint is tests/golden/checkpoints/<name>_f<frame>.npz holding the packed world (header, statF, statU and the
initial state: the packing computes hull planes with library float code whose last bits may differ between hosts, so
the world travels with its state) and the state arrays (poses, vel, art, anchor keys and values in the kernels'
layout, the frame clock and stats words) with the frame number."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle_bind import OracleWorld  # noqa: E402
from steel_webgpu_amd import pack, workloads  # noqa: E402

# name -> (world, frames from rest)
CHECKPOINTS = {
    "arm_hulls": (lambda: pack.pack_scene(workloads.arm_hulls(), capacity_policy="scaled"), 400),
}


def checkpoint_path(name):
    return os.path.join(HERE, "checkpoints", f"{name}_f{CHECKPOINTS[name][1]}.npz")


def save(name):
    make, frames = CHECKPOINTS[name]
    pw = make()
    o = OracleWorld(pw)
    o.frames(frames)
    assert o.warm.size <= 1 or pw.h("W_WH") == 0, "warm-start tables are not checkpointed"
    np.savez_compressed(checkpoint_path(name), frame=np.int64(frames), poses=o.poses, vel=o.vel, art=o.art, akey=o.akey, aval=o.aval,
                        clock=o.count, stats=o.stats, **{"world_" + k: getattr(pw, k) for k in WORLD_FIELDS})
    print(f"{name}: frame {frames}, contacts max {int(o.stats[1])}, dense rows max {int(o.stats[7])}")


WORLD_FIELDS = ("header", "statF", "statU", "poses0", "vel0", "art0")


def world(name):
    """The checkpoint's packed world (enough for Batch and the oracle)."""
    z = np.load(checkpoint_path(name), allow_pickle=False)
    pw = pack.PackedWorld()
    for k in WORLD_FIELDS:
        setattr(pw, k, z["world_" + k])
    pw.NB, pw.NS, pw.NA, pw.NL, pw.ND = (pw.h(f) for f in ("W_NB", "W_NS", "W_NA", "W_NL", "W_ND"))
    return pw


def load_into(o, name):
    """Put a checkpoint into a float OracleWorld of the same world."""
    z = np.load(checkpoint_path(name), allow_pickle=False)
    o.poses[:] = z["poses"]; o.vel[:] = z["vel"]; o.art[:] = z["art"]
    o.akey[:] = z["akey"]; o.aval[:] = z["aval"]; o.count[:] = z["clock"]; o.stats[:] = z["stats"]
    return int(z["frame"])


if __name__ == "__main__":
    os.makedirs(os.path.join(HERE, "checkpoints"), exist_ok=True)
    for n in CHECKPOINTS:
        save(n)
