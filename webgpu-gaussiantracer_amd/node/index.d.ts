// Types of the JavaScript host (index.js) over the Node-API addon.

export declare const RAW_FLOATS: 59;

export interface RenderOptions {
  width: number;
  height: number;
  alpha_min?: number;       // default 0.2  (src/main.ts:464)
  hit_array_size?: number;  // default 6    (src/main.ts:465), 1..32
  T_min?: number;           // default 0.03
  debugMode?: 0 | 1 | 2;    // 0 render, 1 closest proxy, 2 any-hit count
  hFov?: number;            // degrees, default 60
  t_min?: number;           // default 1e-3
  t_max?: number;           // default 1e5
  sh_degree_max?: number;   // 0..3, default 3
  earlyStop?: number;       // rounds, <= 0: never
  antialiasing?: boolean;   // the reference's compiled-out AA path: jittered rays (default false)
  accumulatedFrames?: number; // frames already averaged (the uniform of main.ts:455, 597)
  prev?: Float32Array;      // the previous average, required when accumulatedFrames > 0
  out?: Float32Array;       // width*height*4 floats to render into and return (reused by a render loop)
}

export declare const DEFAULTS: Readonly<Required<Omit<RenderOptions, 'width' | 'height'>>>;

export interface PlyScene {
  raw: Float32Array;        // n * 59 PLY-raw floats
  n: number;
  shDegree: number;
}

export interface GenerateParams {
  n: number;
  nClusters?: number;
  boxHalf?: number;
  clusterSigma?: number;
  logScaleMean?: number;
  logScaleStd?: number;
  logitMean?: number;
  logitStd?: number;
  dcStd?: number;
  restStd?: number;
  seed?: number;
}

export interface SceneInfo {
  nGaussians: number;
  nProxies: number;
  nNodes: number;
  maxStack: number;
  bvhDepth: number;
  deviceBytes: number;
  buildMs: number;
}

export declare class WASDCamera {
  constructor(opts?: { position?: ArrayLike<number>; target?: ArrayLike<number> });
  position: number[];
  target: number[];
  matrix: Float32Array;     // 16 floats, column-major camera-to-world
  lookAt(position: ArrayLike<number>, target: ArrayLike<number>): this;
}

export declare class GaussianTracer {
  constructor(raw: Float32Array, n: number, alphaMin?: number, opts?: { gpuBuild?: boolean });
  render(camera: WASDCamera | Float32Array, opts: RenderOptions): Float32Array;  // width*height*4
  info(): SceneInfo;
  destroy(): void;
}

export declare function loadPly(file: string): PlyScene;
export declare function generateScene(params: GenerateParams): Float32Array;
export declare function version(): string;
/** 8-bit RGBA of a rendered frame as the reference's canvas shows it (mode 0: the shader's hdrColor; 1: exposure; 2: exposure + gamma). */
export declare function toneMap(rgba: Float32Array, width: number, height: number, mode?: 0 | 1 | 2): Uint8Array;
/** The reference's Save button: a frame as a PNG file. */
export declare function savePng(path: string, rgba: Float32Array, width: number, height: number, mode?: 0 | 1 | 2): void;
