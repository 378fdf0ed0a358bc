// Loads the JS shim and the N-API addon without touching the GPU: every
// binding resolves against libsplatgs.so, and createTrainer is exported.
import assert from 'assert';
import { createRequire } from 'module';
import * as api from '../../splat.js_amd/index.mjs';

const require = createRequire(import.meta.url);
const addon = require('../../splat.js_amd/addon/gs_napi.node');
assert.strictEqual(addon.loaded(), true, addon.loadError());
for (const f of ['create', 'setup', 'step', 'readLoss', 'render', 'readGaussians', 'evalCamPsnr',
  'setHoldout', 'setExcluded', 'setIter', 'info', 'sync', 'refine', 'exportPly', 'evalCamPsnrRefined',
  'getCamera', 'readCamGrads', 'destroy', 'stepStats', 'readLossAsync', 'evalCamPsnrAsync', 'readGaussiansAsync',
  'stepViews', 'setRasterShape', 'autotune']) {
  assert.strictEqual(typeof addon[f], 'function', f);
}
assert.strictEqual(typeof api.createTrainer, 'function');
assert.strictEqual(typeof api.GSTrainer.create, 'function');
for (const m of ['setup', 'stepOnce', 'readLoss', 'renderView', 'renderTrainCam', 'readGaussians', 'evalCamPsnr',
  'evalCamPsnrRefined', 'refine', 'exportPly', 'readCamGrads', 'stepStats', 'destroy', 'stepViews',
  'setRasterShape', 'autotune']) {
  assert.strictEqual(typeof api.GSTrainer.prototype[m], 'function', m);
}
console.log(JSON.stringify({ ok: true }));
