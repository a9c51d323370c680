// ES-module face of brain.js (the reference imports net.js and es.js as ES modules).
import brain from './brain.js';

export const {
  HIDDEN, OUT_DIM, inputDim, weightCount, createPopulation, setGenomes, getGenomes, forwardInto,
  createES, beginGeneration, update, esState, trainerStreamState, destroy,
} = brain;
export default brain;
