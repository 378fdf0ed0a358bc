"""The JavaScript surface (splat.js_amd/index.mjs -> N-API addon -> C ABI).

CPU: the addon loads under this image's Node and every binding resolves.
GPU: a Node process trains the reference's gradcheck rig through
createTrainer()/GSTrainer (after an autotuning sweep, at 8-pixel tiles with
128-key batches) and its results equal the ctypes path bit for bit.
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
           "maxW": W, "maxH": H, "radius": radius, "steps": steps, "tile": 8, "batch": 128}
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    out = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "run_trainer.mjs"), str(tmp_path)],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    res = json.loads(out.stdout.strip().splitlines()[-1])
    t = gsabi.GSTrainer(shDeg=0, **opts)
    t.setup({"data": params, "n": params.shape[0]}, cams, images, W, H, radius)
    t.setRasterShape(8, 128)
    for _ in range(steps):
        t.stepOnce(0)
    loss = t.readLoss()
    img = t.renderView(t.camMeta[0])
    data = t.readGaussians()["data"]
    psnr = t.evalCamPsnr(0)
    t.destroy()
    js_params = np.frombuffer((tmp_path / "out_params.f32").read_bytes(), np.float32).reshape(-1, 16)
    js_img = np.frombuffer((tmp_path / "out_img.f32").read_bytes(), np.float32).reshape(img.shape)
    assert res["iter"] == steps and res["n"] == params.shape[0] and res["tune"] == 9
    np.testing.assert_array_equal(js_params.view(np.uint32), data.view(np.uint32))
    np.testing.assert_array_equal(js_img.view(np.uint32), img.view(np.uint32))
    assert res["loss"] == loss and res["psnr"] == psnr


@pytest.mark.gpu
@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_node_async_readback_and_host_fields(tmp_path):
    """Config 1 (20k Gaussians, 320x240) from Node: 100 stepOnce calls are
    only enqueued; 100 reads of iter/n/cap do not wait for the GPU;
    readLoss() returns a pending Promise while those steps run, and the event
    loop keeps turning until it settles; the results equal the ctypes
    path's; destroy() frees at once and a new trainer works right after."""
    from helpers import config1_setup
    p, sh, cams, images, W, H, radius = config1_setup()
    steps = 100
    (tmp_path / "params.f32").write_bytes(p.astype(np.float32).tobytes())
    (tmp_path / "sh.f32").write_bytes(sh.astype(np.float32).tobytes())
    for i, im in enumerate(images):
        (tmp_path / f"img{i}.f32").write_bytes(np.asarray(im["rgb"], np.float32).tobytes())
    cfg = {"opts": {"shDeg": 2}, "cams": cams, "images": [{"tw": im["tw"], "th": im["th"]} for im in images],
           "maxW": W, "maxH": H, "radius": radius, "steps": steps, "sh": True}
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    out = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "async_readback.mjs"), str(tmp_path)],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    res = json.loads(out.stdout.strip().splitlines()[-1])
    print(res)
    assert res["iter"] == steps + 1
    assert res["pendingAfterCall"] is True
    # the enqueue returned long before the GPU finished the queued steps
    assert res["waitMs"] > res["readLossCallMs"]
    assert res["iterReadsMs"] < 20.0
    assert res["lost"] == "destroyed" and "destroyed" in res["afterDestroy"]
    t = gsabi.GSTrainer(shDeg=2)
    t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, W, H, radius)
    t.stepOnce(0)
    first = t.readLoss()
    for _ in range(steps):
        t.stepOnce(0)
    loss = t.readLoss()
    data = t.readGaussians()["data"]
    psnr = t.evalCamPsnr(0)
    t.destroy()
    js = np.frombuffer((tmp_path / "out_params.f32").read_bytes(), np.float32).reshape(-1, 16)
    np.testing.assert_array_equal(js.view(np.uint32), data.view(np.uint32))
    # the loss statistics are double atomics (summation order varies)
    for a, b in ((res["first"], first), (res["loss"], loss), (res["loss2"], first)):
        assert abs(a - b) <= 1e-12 * abs(b)
    assert abs(res["psnr"] - psnr) <= 1e-9 * abs(psnr)
