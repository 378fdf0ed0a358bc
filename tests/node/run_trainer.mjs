// Drives the JS GSTrainer surface (splat.js_amd/index.mjs) through the N-API
// addon: createTrainer -> setup -> autotune -> setRasterShape -> stepOnce x N ->
// readLoss -> renderView -> readGaussians -> evalCamPsnr. Inputs and outputs are raw little-endian
// files in the directory given as argv[2] (written/read by
// tests/test_node_shim.py).
import fs from 'fs';
import path from 'path';
import { createTrainer } from '../../splat.js_amd/index.mjs';

const dir = process.argv[2];
const cfg = JSON.parse(fs.readFileSync(path.join(dir, 'cfg.json'), 'utf8'));
const f32 = (name) => {
  const b = fs.readFileSync(path.join(dir, name));
  return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4);
};

const main = async () => {
  const trainer = await createTrainer(cfg.opts);
  const data = f32('params.f32');
  const n = data.length / 16;
  const images = cfg.images.map((im, i) => ({ tw: im.tw, th: im.th, rgb: f32(`img${i}.f32`) }));
  trainer.setup({ data, n }, cfg.cams, images, cfg.maxW, cfg.maxH, cfg.radius);
  // the autotuner leaves the state as it was; then the shape under test
  const tune = trainer.autotune([0], 1);
  trainer.setRasterShape(cfg.tile, cfg.batch);
  for (let k = 0; k < cfg.steps; k++) trainer.stepOnce(0);
  await trainer.device.queue.onSubmittedWorkDone();
  const loss = await trainer.readLoss();
  const img = trainer.renderView(trainer.camMeta[0]);
  const g = await trainer.readGaussians();
  const psnr = await trainer.evalCamPsnr(0);
  fs.writeFileSync(path.join(dir, 'out_params.f32'), Buffer.from(g.data.buffer));
  fs.writeFileSync(path.join(dir, 'out_img.f32'), Buffer.from(img.buffer));
  console.log(JSON.stringify({ loss, psnr, iter: trainer.iter, n: g.n, tune: tune.length }));
};
main().catch((e) => { console.error(e.stack || String(e)); process.exit(1); });
