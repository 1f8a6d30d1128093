'use strict';
/*
 * JavaScript host of the MI355X Gaussian ray tracer: the roles of the
 * reference's TypeScript host (src/main.ts frame loop, src/camera.ts
 * WASDCamera, src/GS_ply.ts loader) over the Node-API addon lib/gtrace.node
 * and the C-ABI of include/gt.h.
 *
 *   const gt = require('.../node');
 *   const { raw, n } = gt.loadPly('scene.ply');          // GS_ply.ts loadPly
 *   const camera = new gt.WASDCamera({ position: [0, 0, 3], target: [0, 0, 0] });
 *   const tracer = new gt.GaussianTracer(raw, n, 0.2);   // gaussianSceneToBvh(alpha_min)
 *   const rgba = tracer.render(camera, { alpha_min: 0.2, hit_array_size: 6, T_min: 0.03,
 *                                        debugMode: 0, width: 800, height: 600 });
 *
 * rgba is a Float32Array of width*height*4 floats, pixel (x, y) at
 * 4*(y*width + x), alpha 1 (src/GS_shader.ts:300-301).
 */
const path = require('path');

const addon = require(path.join(__dirname, '..', 'lib', 'gtrace.node'));

const DEFAULTS = Object.freeze({
  // src/main.ts:448-465
  hFov: 60.0, t_min: 1e-3, t_max: 1e5, T_min: 0.03, sh_degree_max: 3, earlyStop: 0, debugMode: 0,
  alpha_min: 0.2, hit_array_size: 6,
});

/* WASDCamera (src/camera.ts:126-204) placed at `position` looking at
 * `target`; `matrix` is the column-major camera-to-world matrix the
 * reference writes to the uniform block after the first update(). */
class WASDCamera {
  constructor({ position = [0, 0, 3], target = [0, 0, 0] } = {}) {
    this.position = Array.from(position);
    this.target = Array.from(target);
    this.matrix = addon.cameraLookAt(this.position, this.target);
  }

  lookAt(position, target) {
    this.position = Array.from(position);
    this.target = Array.from(target);
    this.matrix = addon.cameraLookAt(this.position, this.target);
    return this;
  }
}

/* loadPly (src/GS_ply.ts:58-132, 171-196): PLY-raw records, 59 floats each */
function loadPly(file) {
  return addon.plyRead(String(file));
}

/* The project's seeded synthetic scene generator (same streams as the
 * Python and C++ callers). */
function generateScene(params) {
  return addon.generateScene(params);
}

class GaussianTracer {
  /* raw: Float32Array of n*59 PLY-raw floats; alphaMin: the proxy filter /
   * size threshold of gaussianSceneToBvh (src/main.ts:30-111);
   * opts.gpuBuild: build the proxies and the BVH on the GPU (same results,
   * milliseconds instead of seconds for a million Gaussians) */
  constructor(raw, n, alphaMin = DEFAULTS.alpha_min, opts = {}) {
    this.n = n;
    this.alphaMin = alphaMin;
    this.handle = addon.sceneCreate(raw, n, alphaMin, !!opts.gpuBuild);
  }

  /* render(camera, {alpha_min, hit_array_size, T_min, debugMode, ...}):
   * one traceRays of the whole image (src/main.ts:520-534, 563-593).
   * A different alpha_min rebuilds the proxies and the BVH, as the
   * reference must call gaussianSceneToBvh again for it. */
  render(camera, opts = {}) {
    const o = Object.assign({}, DEFAULTS, opts);
    if (!(o.width >= 1 && o.height >= 1)) throw new RangeError('render: opts.width and opts.height are required');
    if (o.alpha_min !== this.alphaMin) {
      addon.sceneSetAlphaMin(this.handle, o.alpha_min);
      this.alphaMin = o.alpha_min;
    }
    const toWorld = camera instanceof Float32Array ? camera : camera.matrix;
    return addon.render(this.handle, toWorld, o);
  }

  info() {
    return addon.sceneInfo(this.handle);
  }

  destroy() {
    if (this.handle) addon.sceneDestroy(this.handle);
    this.handle = null;
  }
}

/* The screen pass (src/toneMapper.ts:11-26): 8-bit RGBA of a rendered frame
 * as the canvas shows it.  mode 0 = the shader's output (hdrColor), 1 = its
 * exposure curve 1 - exp(-1.66 hdr), 2 = that with gamma 1/2.2. */
function toneMap(rgba, width, height, mode = 0) {
  return addon.toneMap(rgba, width, height, mode);
}

/* The Save button (src/main.ts:606-611): a frame as a PNG file. */
function savePng(path, rgba, width, height, mode = 0) {
  addon.writePng(path, addon.toneMap(rgba, width, height, mode), width, height);
}

module.exports = {
  WASDCamera, GaussianTracer, loadPly, generateScene, toneMap, savePng, DEFAULTS, RAW_FLOATS: 59,
  version: addon.version, addon,
};
