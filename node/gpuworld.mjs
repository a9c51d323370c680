// ES-module face of gpuworld.js (the reference imports gpuworld as an ES module).
import gw from './gpuworld.js';

export const {
  GPU_STRIDES, C_PIN, C_ANGLE, C_GRIP,
  isAvailable, create, upload, uploadControls, packControls, step, download, mirrorInto,
  groupCollPoints, destroy, rollout, readState,
} = gw;
export default gw;
