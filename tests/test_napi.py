"""The Node-API binding (steel_webgpu_amd/js): loads under the image's Node, maps a missing GPU to the
reference's UnsupportedError (src/device.js:1-4), and its worldHeader() helper writes the same 160 header words
as the Python packer from the reference-named layout and constants."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from steel_webgpu_amd import pack, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "steel_webgpu_amd", "steel_napi.node")
JS = os.path.join(ROOT, "steel_webgpu_amd", "js", "steel_amd.js")

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON), reason="node or the addon is absent")


def _node(script, stdin=""):
    r = subprocess.run(["node", "-e", script], input=stdin, capture_output=True, text=True, timeout=60, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_missing_gpu_raises_unsupported_error():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    out = _node(f"""const s = require({json.dumps(JS)});
try {{ s.initDevice(0); console.log('opened'); }}
catch (e) {{ console.log(JSON.stringify({{u: e instanceof s.UnsupportedError, e: e instanceof Error, name: e.name, msg: e.message}})); }}""")
    r = json.loads(out)
    assert r == {"u": True, "e": True, "name": "UnsupportedError", "msg": r["msg"]} and "device" in r["msg"]


def _reference_c(pw):
    """The Sim constructor's C object (sim.js:456-488) for a packed world: its section offsets and the constants,
    named as the reference names them."""
    F = pack.FIELDS
    hv = lambda k: int(pw.header[F["W_" + k]])
    hf = lambda k: float(np.uint32(pw.header[F["W_" + k]]).view(np.float32))
    C = {}
    for k in ("NB", "NS", "NA", "NL", "ND", "NDA", "JM", "MAXPATH", "MAXLP", "NCOLW", "ROWS", "PO", "PATCH", "EXTRA", "REST", "DHDR",
              "DPS", "NMIMIC", "NEST", "MO", "TORS", "TO", "VELITERS", "COMPL", "CO", "DRIVE_ROWS", "FUSE_VLIM", "VEL_DRIVE_ACTIVE",
              "NLOOP", "ANG", "CTRL_KIND", "CTRL_A", "CTRL_B", "FMODE", "FMO", "ANG_ART", "EO", "VLIM_ON", "VLIM_FREE", "JFL", "TGT_HELD",
              "SLEEP", "FRICTION_COMBINE", "RESTITUTION_COMBINE"):
        C[k] = hv(k)
    for k in ("F_CFM", "F_OMEGA", "F_SPEC", "F_GAP_GATE", "F_FGAP", "F_FSCALE", "F_ANCHOR_BETA", "F_BP_CELL", "F_REST_SPEED", "F_JL_GAP",
              "F_JL_SPEED", "F_VLIM_FRAC", "F_SLEEP_LIN", "F_SLEEP_ANG", "F_SLEEP_TIME", "F_CREEP"):
        C[k] = hf(k)
    C.update(pw.SF)
    C.update(pw.SU)
    C.update(pw.P)
    C.update({k: v for k, v in pw.A.items()})
    extra = {"NH": hv("NH"), "NQ": hv("NQ"), "nstatf": hv("NSTATF"), "nstatu": hv("NSTATU"), "nart": hv("NART"), "fps": hf("FPS"),
             "substeps": hv("SUBSTEPS"), "iterations": hv("ITERS"), "gravity": [hf("GX"), hf("GY"), hf("GZ")], "regularization": hf("REG"),
             "warmScale": hf("WARM_SCALE"), "maxDepen": hf("MAX_DEPEN"), "baumgarte": hf("BETA"), "warmSlots": hv("WH"),
             "capacities": {k: hv(k) for k in ("MAXP", "MAXC", "MAXU", "UH", "AH", "MAXDENSE")},
             "path": {"driven": hv("DRIVEN"), "frames": hv("PATH_FRAMES"), "integral": hv("INTEGRAL_COUNT")}}
    return C, extra


@pytest.mark.parametrize("make", [lambda: workloads.boxes64(n=8), lambda: workloads.quadruped_world(2), lambda: workloads.arm_hulls(nlinks=6, nhulls=10)])
def test_world_header_helper_matches_the_python_packer(make):
    pw = pack.pack_scene(make())
    C, extra = _reference_c(pw)
    out = _node(f"""const s = require({json.dumps(JS)});
let d = ''; process.stdin.on('data', (c) => {{ d += c; }}); process.stdin.on('end', () => {{
  const a = JSON.parse(d); console.log(JSON.stringify(Array.from(s.worldHeader(a.C, a.extra)))); }});""", json.dumps({"C": C, "extra": extra}))
    got = np.array(json.loads(out), np.uint32)
    assert np.array_equal(got, pw.header.astype(np.uint32)), np.nonzero(got != pw.header)[0]


@pytest.mark.gpu
def test_node_steps_two_worlds_bitexact_with_the_python_host(tmp_path):
    """Two packed worlds created, stepped (12 frames) and read back through the addon under Node match the same
    batch stepped through the ctypes host bit for bit (poses, velocities, articulation state)."""
    from steel_webgpu_amd import sim
    pws = [pack.pack_scene(workloads.quadruped_world(2)), pack.pack_scene(workloads.arm_hulls(nlinks=6, nhulls=10, fps=200, substeps=4))]
    for i, pw in enumerate(pws):
        for n, dt in (("header", np.uint32), ("statF", np.float32), ("statU", np.uint32), ("poses0", np.float32), ("vel0", np.float32),
                      ("art0", np.float32)):
            a = np.ascontiguousarray(getattr(pw, n), dt)
            if a.size == 0:
                a = np.zeros(8, dt)
            a.tofile(tmp_path / f"w{i}_{n}.bin")
    out = _node(f"""const fs = require('fs'); const s = require({json.dumps(JS)});
const dir = {json.dumps(str(tmp_path))};
const ld = (i, n, T) => {{ const b = fs.readFileSync(`${{dir}}/w${{i}}_${{n}}.bin`); return new T(b.buffer.slice(b.byteOffset, b.byteOffset + b.length)); }};
const worlds = [0, 1].map((i) => ({{ header: ld(i, 'header', Uint32Array), statF: ld(i, 'statF', Float32Array), statU: ld(i, 'statU', Uint32Array),
  poses0: ld(i, 'poses0', Float32Array), vel0: ld(i, 'vel0', Float32Array), art0: ld(i, 'art0', Float32Array) }}));
const dev = s.initDevice(0); const b = new s.Batch(dev, worlds);
b.step(12);
for (const i of [0, 1]) for (const n of ['poses', 'vel', 'art']) {{
  const dst = new Uint8Array(b.bufferSize(i, n)); b.read(i, n, dst); fs.writeFileSync(`${{dir}}/out${{i}}_${{n}}.bin`, dst); }}
b.destroy(); dev.close(); console.log('ok');""")
    assert out.strip() == "ok"
    b = sim.Batch(pws, block=64)
    b.step(12)
    for i in range(2):
        for n in ("poses", "vel", "art"):
            want = b.read(i, n, dtype=np.uint8)
            got = np.fromfile(tmp_path / f"out{i}_{n}.bin", np.uint8)
            assert np.array_equal(got, want), (i, n)
