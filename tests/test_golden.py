"""Committed golden fixtures (tests/golden/, written by make_golden.py).

CPU: the f32 restatement still reproduces them bit for bit.
GPU: one trainer step through the C ABI reproduces them bit for bit.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import gsabi
from helpers import OracleTrainer, config1_setup, rig_setup

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _rig(deg):
    return np.load(os.path.join(HERE, f"rig_deg{deg}.npz"))


@pytest.mark.parametrize("deg", [0, 3])
def test_oracle_reproduces_rig_fixture(oracle_lib, deg):
    g = _rig(deg)
    params, sh, cams, images, W, H, radius, opts = rig_setup(deg)
    np.testing.assert_array_equal(_bits(params), _bits(g["params_in"]))
    ref = OracleTrainer(params, sh, cams, images, radius, opts=opts, shDeg=deg)
    ref.iter = int(g["start_iter"])
    r = ref.step(0)
    for k in ("ids", "keys", "end", "accum"):
        np.testing.assert_array_equal(r[k], g[k], err_msg=k)
    np.testing.assert_array_equal(r["ts"], g["tile_start"])
    for k in ("img", "T", "gradF"):
        np.testing.assert_array_equal(_bits(r[k].astype(np.float32)), _bits(g[k]), err_msg=k)
    np.testing.assert_array_equal(_bits(ref.P), _bits(g["params_out"]))
    if deg:
        np.testing.assert_array_equal(_bits(r["shGrad"]), _bits(g["shGrad"]))
        np.testing.assert_array_equal(_bits(ref.SH), _bits(g["sh_out"]))


def test_oracle_reproduces_config1_hashes(oracle_lib):
    g = json.load(open(os.path.join(HERE, "config1.json")))
    p, sh, cams, images, W, H, radius = config1_setup()
    ref = OracleTrainer(p, sh, cams, images, radius, shDeg=2)
    ref.iter = g["start_iter"]
    r = ref.step(0)
    assert len(r["ids"]) == g["num_keys"] and int((r["cnt"] > 0).sum()) == g["visible"]
    got = {"ids": _sha(r["ids"]), "keys": _sha(r["keys"]), "img": _sha(r["img"].astype(np.float32)),
           "T": _sha(r["T"]), "end": _sha(r["end"]), "accum": _sha(r["accum"]),
           "gradF14": _sha(np.ascontiguousarray(r["gradF"][:, :14].astype(np.float32))),
           "shGrad": _sha(r["shGrad"]), "params_out": _sha(ref.P), "sh_out": _sha(ref.SH)}
    assert got == g["sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("deg", [0, 3])
def test_gpu_reproduces_rig_fixture(deg):
    g = _rig(deg)
    params, sh, cams, images, W, H, radius, opts = rig_setup(deg)
    t = gsabi.GSTrainer(shDeg=deg, **opts)
    gs = {"data": params, "n": params.shape[0]}
    if deg:
        gs["sh"] = sh
    t.setup(gs, cams, images, W, H, radius)
    t.iter = int(g["start_iter"])
    d = t.step_debug(0)
    out = t.readGaussians()
    t.destroy()
    np.testing.assert_array_equal(d["ids"], g["ids"])
    np.testing.assert_array_equal(d["keys"], g["keys"])
    np.testing.assert_array_equal(_bits(d["img"]), _bits(g["img"]))
    np.testing.assert_array_equal(_bits(d["T"]), _bits(g["T"]))
    np.testing.assert_array_equal(d["end"], g["end"])
    np.testing.assert_array_equal(d["accum"], g["accum"])
    np.testing.assert_array_equal(_bits(d["gradF"][:, :14]), _bits(g["gradF"][:, :14]))
    np.testing.assert_array_equal(_bits(out["data"]), _bits(g["params_out"]))
    if deg:
        np.testing.assert_array_equal(_bits(d["shGrad"]), _bits(g["shGrad"]))
        np.testing.assert_array_equal(_bits(out["sh"]), _bits(g["sh_out"]))


@pytest.mark.gpu
def test_gpu_reproduces_config1_hashes():
    g = json.load(open(os.path.join(HERE, "config1.json")))
    p, sh, cams, images, W, H, radius = config1_setup()
    t = gsabi.GSTrainer(shDeg=2)
    t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, W, H, radius)
    t.iter = g["start_iter"]
    d = t.step_debug(0)
    out = t.readGaussians()
    t.destroy()
    assert d["numKeys"] == g["num_keys"]
    got = {"ids": _sha(d["ids"]), "keys": _sha(d["keys"]), "img": _sha(d["img"]), "T": _sha(d["T"]),
           "end": _sha(d["end"]), "accum": _sha(d["accum"]),
           "gradF14": _sha(np.ascontiguousarray(d["gradF"][:, :14])), "shGrad": _sha(d["shGrad"]),
           "params_out": _sha(out["data"]), "sh_out": _sha(out["sh"])}
    assert got == g["sha256"]
