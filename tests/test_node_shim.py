"""The JavaScript surface (splat.js_amd/index.mjs -> N-API addon -> C ABI).

CPU: the addon loads under this image's Node and every binding resolves.
GPU: a Node process trains the reference's gradcheck rig through
createTrainer()/GSTrainer and its results equal the ctypes path bit for bit.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import gsabi
from helpers import rig_setup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")


@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_addon_loads_in_node():
    out = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "test_addon_load.mjs")],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert out.returncode == 0, out.stdout
    assert json.loads(out.stdout.strip().splitlines()[-1])["ok"] is True


@pytest.mark.gpu
@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_node_trainer_matches_ctypes(tmp_path):
    params, _, cams, images, W, H, radius, opts = rig_setup(0)
    steps = 4
    (tmp_path / "params.f32").write_bytes(params.astype(np.float32).tobytes())
    for i, im in enumerate(images):
        (tmp_path / f"img{i}.f32").write_bytes(np.asarray(im["rgb"], np.float32).tobytes())
    cfg = {"opts": dict(opts, shDeg=0), "cams": cams, "images": [{"tw": im["tw"], "th": im["th"]} for im in images],
           "maxW": W, "maxH": H, "radius": radius, "steps": steps}
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    out = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "run_trainer.mjs"), str(tmp_path)],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    res = json.loads(out.stdout.strip().splitlines()[-1])
    t = gsabi.GSTrainer(shDeg=0, **opts)
    t.setup({"data": params, "n": params.shape[0]}, cams, images, W, H, radius)
    for _ in range(steps):
        t.stepOnce(0)
    loss = t.readLoss()
    img = t.renderView(t.camMeta[0])
    data = t.readGaussians()["data"]
    psnr = t.evalCamPsnr(0)
    t.destroy()
    js_params = np.frombuffer((tmp_path / "out_params.f32").read_bytes(), np.float32).reshape(-1, 16)
    js_img = np.frombuffer((tmp_path / "out_img.f32").read_bytes(), np.float32).reshape(img.shape)
    assert res["iter"] == steps and res["n"] == params.shape[0]
    np.testing.assert_array_equal(js_params.view(np.uint32), data.view(np.uint32))
    np.testing.assert_array_equal(js_img.view(np.uint32), img.view(np.uint32))
    assert res["loss"] == loss and res["psnr"] == psnr
