"""The drop-in boundary's behaviour through the C ABI (ctypes binding):
the reference's default key budget, input validation of the SH array, and
the asynchronous readbacks (enqueue, then wait) against the synchronous
ones on twin trainers."""
import numpy as np
import pytest

import gsabi
from helpers import config1_setup

pytestmark = pytest.mark.gpu


def _setup(t, n=2000):
    p, sh, cams, images, W, H, radius = config1_setup(n=n)
    t.setup({"data": p, "n": n, "sh": sh}, cams, images, W, H, radius)
    return p, sh, cams, images, W, H, radius


def test_default_entry_budget_follows_the_reference():
    """trainer.js:48-49: 12M keys unless opts.maxSplats is set, then
    max(12M, 24 * maxSplats)."""
    for kw, want in (({}, 12_000_000), ({"maxSplats": 100_000}, 12_000_000),
                     ({"maxSplats": 1_000_000}, 24_000_000), ({"entriesCap": 777}, 777)):
        t = gsabi.GSTrainer(shDeg=2, **kw)
        _setup(t)
        assert t.info().entriesCap == want, kw
        t.destroy()


def test_setup_rejects_sh_of_another_length():
    p, sh, cams, images, W, H, radius = config1_setup(n=500)
    t = gsabi.GSTrainer(shDeg=2)
    for bad in (sh[:, :9], sh[:400], np.zeros((500, 45), np.float32)):
        with pytest.raises(gsabi.GsError, match="gaussians.sh"):
            t.setup({"data": p, "n": 500, "sh": bad}, cams, images, W, H, radius)
    t.setup({"data": p, "n": 500, "sh": sh}, cams, images, W, H, radius)
    t.destroy()


def test_async_readbacks_equal_the_synchronous_ones():
    a, b = gsabi.GSTrainer(shDeg=2), gsabi.GSTrainer(shDeg=2)
    for t in (a, b):
        _setup(t)
        t.iter = 2500
    for _ in range(3):
        a.stepOnce(0)
        b.stepOnce(0)
    pend = b.read_loss_async()
    # more work may be enqueued while the readback is pending
    a_loss = a.readLoss()
    for t in (a, b):
        t.stepOnce(0)
    b_loss, b_over = pend.wait()
    assert abs(a_loss - b_loss) <= 1e-12 * abs(a_loss) and b_over == 0
    assert b.read_loss_async().wait()[0] is not None
    assert a.readLoss() is not None
    pg = b.read_gaussians_async()
    pp = b.eval_cam_psnr_async(0)
    ga = a.readGaussians()
    pa = a.evalCamPsnr(0)
    gb = pg.wait()
    pb = pp.wait()
    np.testing.assert_array_equal(ga["data"].view(np.uint32), gb["data"].view(np.uint32))
    np.testing.assert_array_equal(ga["sh"].view(np.uint32), gb["sh"].view(np.uint32))
    assert abs(pa - pb) <= 1e-9 * abs(pa)
    # evalCamPsnr reset the loss window: nothing seen since (the reference's null)
    assert a.readLoss() is None and b.read_loss_async().wait()[0] is None
    # a pending readback outlives its trainer
    b.stepOnce(0)
    late = b.read_loss_async()
    b.destroy()
    assert late.wait()[0] is not None
    a.destroy()
