// splat.js_amd — MI355X-native trainer step for splat.js.
// Layer-1 entry point with the reference's name and shape (src/index.js:56-58).
import { GSTrainer } from './js/trainer.mjs';

export { GSTrainer, STRIDE, TILE, LOSS, ACCUM, SORT } from './js/trainer.mjs';

/** a trainer on the MI355X (the reference binds one to a WebGPU device) */
export function createTrainer(opts = {}) {
  return GSTrainer.create(opts);
}
