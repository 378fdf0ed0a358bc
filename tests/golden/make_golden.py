"""Writes the golden fixtures under tests/golden/ from the f32 CPU
restatement (oracle/). The reference ships no trainer vectors (its only
trainer test is a PSNR floor in a browser), so the fixtures pin our own
restatement; test_oracle.py pins the restatement itself against closed
forms, textbook constants and finite differences.

    python tests/golden/make_golden.py

Fixtures:
  rig_deg0.npz   the reference's gradcheck rig (gradcheck.js:139-185), one
                 step at iteration 0: inputs, sort keys/ids, image, T, end,
                 fixed-point gradient sums, parameter gradients, and the
                 parameters after Adam (all f32 bits)
  rig_deg3.npz   gradCheckSH's rig (gradcheck.js:192-201), SH degree 3 with
                 all bands active (iteration 5000)
  config1.json   config 1 (20k Gaussians, 320x240, SH 2, iteration 2500):
                 SHA-256 of each output array, plus key/visible counts
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("oracle", os.path.join("splat.js_amd", "py"), "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

from helpers import OracleTrainer, config1_setup, rig_setup  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def rig_fixture(deg, start_iter):
    params, sh, cams, images, W, H, radius, opts = rig_setup(deg)
    ref = OracleTrainer(params, sh, cams, images, radius, opts=opts, shDeg=deg)
    ref.iter = start_iter
    r = ref.step(0)
    out = {"params_in": params, "target_rgb": np.asarray(images[0]["rgb"], np.float32),
           "ids": r["ids"], "keys": r["keys"], "tile_start": r["ts"], "img": r["img"].astype(np.float32),
           "T": r["T"], "end": r["end"], "accum": r["accum"], "gradF": r["gradF"].astype(np.float32),
           "params_out": ref.P, "start_iter": np.int64(start_iter)}
    if deg > 0:
        out["sh_in"] = sh
        out["shGrad"] = r["shGrad"]
        out["sh_out"] = ref.SH
    return out


def config1_fixture():
    p, sh, cams, images, W, H, radius = config1_setup()
    ref = OracleTrainer(p, sh, cams, images, radius, shDeg=2)
    ref.iter = 2500
    r = ref.step(0)
    return {"start_iter": 2500, "num_keys": int(len(r["ids"])), "visible": int((r["cnt"] > 0).sum()),
            "sha256": {"ids": sha(r["ids"]), "keys": sha(r["keys"]), "img": sha(r["img"].astype(np.float32)),
                       "T": sha(r["T"]), "end": sha(r["end"]), "accum": sha(r["accum"]),
                       "gradF14": sha(np.ascontiguousarray(r["gradF"][:, :14].astype(np.float32))),
                       "shGrad": sha(r["shGrad"]), "params_out": sha(ref.P), "sh_out": sha(ref.SH)}}


def main():
    np.savez_compressed(os.path.join(HERE, "rig_deg0.npz"), **rig_fixture(0, 0))
    np.savez_compressed(os.path.join(HERE, "rig_deg3.npz"), **rig_fixture(3, 5000))
    with open(os.path.join(HERE, "config1.json"), "w") as f:
        json.dump(config1_fixture(), f, indent=1)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
