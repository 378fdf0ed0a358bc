// The trainer's asynchronous surface from Node: stepOnce only enqueues, the
// host fields read without waiting for the GPU, readLoss() returns a pending
// Promise while steps are queued and the event loop keeps running until it
// settles, destroy() frees the device state at once (device.lost settles
// with 'destroyed') and a second trainer can be created right after.
// Inputs are raw files in the directory argv[2] (tests/test_node_shim.py).
import fs from 'fs';
import path from 'path';
import { createTrainer } from '../../splat.js_amd/index.mjs';

const dir = process.argv[2];
const cfg = JSON.parse(fs.readFileSync(path.join(dir, 'cfg.json'), 'utf8'));
const f32 = (name) => {
  const b = fs.readFileSync(path.join(dir, name));
  return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4);
};
const ms = (a, b) => Number(b - a) / 1e6;
const now = () => process.hrtime.bigint();

const main = async () => {
  const data = f32('params.f32');
  const sh = cfg.sh ? f32('sh.f32') : null;
  const n = data.length / 16;
  const images = cfg.images.map((im, i) => ({ tw: im.tw, th: im.th, rgb: f32(`img${i}.f32`) }));
  const trainer = await createTrainer(cfg.opts);
  trainer.setup({ data, n, sh }, cfg.cams, images, cfg.maxW, cfg.maxH, cfg.radius);
  trainer.stepOnce(0);
  const first = await trainer.readLoss();

  const t0 = now();
  for (let k = 0; k < cfg.steps; k++) trainer.stepOnce(0);
  const t1 = now();
  let iter = 0;
  for (let k = 0; k < 100; k++) iter = trainer.iter + trainer.n * 0 + trainer.cap * 0;
  const t2 = now();
  let settled = false;
  const p = trainer.readLoss().then((v) => { settled = true; return v; });
  const t3 = now();
  const pendingAfterCall = !settled;
  let ticks = 0;
  const iv = setInterval(() => { ticks++; }, 0);
  const loss = await p;
  clearInterval(iv);
  const t4 = now();
  const g = await trainer.readGaussians();
  const psnr = await trainer.evalCamPsnr(0);
  fs.writeFileSync(path.join(dir, 'out_params.f32'), Buffer.from(g.data.buffer));
  trainer.destroy();
  const lost = await trainer.device.lost;
  let afterDestroy = '';
  try { trainer.stepOnce(0); } catch (e) { afterDestroy = e.message; }

  const second = await createTrainer(cfg.opts);
  second.setup({ data, n, sh }, cfg.cams, images, cfg.maxW, cfg.maxH, cfg.radius);
  second.stepOnce(0);
  const loss2 = await second.readLoss();
  second.destroy();

  console.log(JSON.stringify({
    first, loss, loss2, psnr, iter, pendingAfterCall, ticks, lost: lost.reason, afterDestroy,
    enqueueMs: ms(t0, t1), iterReadsMs: ms(t1, t2), readLossCallMs: ms(t2, t3), waitMs: ms(t3, t4),
  }));
};
main().catch((e) => { console.error(e.stack || String(e)); process.exit(1); });
