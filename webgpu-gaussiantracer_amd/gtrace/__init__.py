"""Python binding of libgt.so (the MI355X Gaussian ray tracer C-ABI).

Mirrors the reference's per-frame surface
    render(camera, {alpha_min, hit_array_size, T_min, debugMode}) -> RGBA
(src/main.ts:448-534, 542-600) with
    Scene            the acceleration container (gaussianSceneToBvh, main.ts:30-111)
    Uniforms         the 112-B uniform block (GS_shader.ts:96-108, main.ts:520-534)
    Scene.render()   traceRays(W, H) + readback (pass_encoder.ts:35-65)
The library must be built (webgpu-gaussiantracer_amd/build_lib.py); there is
no fallback path: a missing libgt.so raises at import.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GTRACE_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libgt.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libgt.so not built at {LIB_PATH}; run __graft_entry__.build()")
# One HIP runtime per process: PyTorch ships its own libamdhip64.so.7 and
# loads it under a different file name, so libgt (which needs the soname
# libamdhip64.so.7) must be loaded after torch to bind to torch's copy;
# torch streams and device pointers passed to libgt then belong to it.
try:
    import torch  # noqa: F401
except ImportError:
    pass
_lib = C.CDLL(LIB_PATH)

GT_OK = 0
GT_RAW_FLOATS = 59
GT_MAX_K = 32
GT_NSTATS = 8
SENTINEL = 0xFFFFFFFF
STAT_NAMES = ["rays", "rounds", "h_eval", "h_comp", "proxies", "hits", "nodes", "overflow"]


class GTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gt error {code}: {msg}")
        self.code = code


class Uniforms(C.Structure):
    _fields_ = [
        ("to_world", C.c_float * 16), ("camera_aspect", C.c_float * 2),
        ("t_min", C.c_float), ("t_max", C.c_float), ("T_min", C.c_float), ("alpha_min", C.c_float),
        ("k", C.c_uint32), ("sh_degree_max", C.c_uint32), ("accumulated_frames", C.c_uint32),
        ("early_stop", C.c_uint32), ("debug_mode", C.c_uint32), ("_pad", C.c_uint32),
    ]


assert C.sizeof(Uniforms) == 112


class Frame(C.Structure):
    _fields_ = [
        ("rgba", C.c_void_p), ("hit_ids", C.c_void_p), ("accept", C.c_void_p),
        ("n_rounds", C.c_void_p), ("T", C.c_void_p), ("stats", C.c_void_p),
        ("max_rounds", C.c_uint32), ("comp_cap", C.c_uint32), ("comp_ids", C.c_void_p), ("n_comp", C.c_void_p),
        ("antialiasing", C.c_uint32), ("_pad", C.c_uint32),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("n_gaussians", C.c_uint32), ("n_proxies", C.c_uint32), ("n_nodes", C.c_uint32),
        ("max_stack", C.c_uint32), ("bvh_depth", C.c_uint32), ("_pad", C.c_uint32),
        ("device_bytes", C.c_uint64), ("build_ms", C.c_double),
    ]


class Tuning(C.Structure):
    _fields_ = [("path", C.c_uint32), ("refill_min", C.c_uint32), ("comp_min", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("tile_w", C.c_uint32), ("stack_lds", C.c_uint32),
                ("rects", C.c_uint32), ("colors", C.c_uint32)]


class AutotuneResult(C.Structure):
    _fields_ = [("best", Tuning), ("best_ms", C.c_double), ("first_ms", C.c_double), ("n_tried", C.c_uint32),
                ("_pad", C.c_uint32)]


class Adam(C.Structure):
    _fields_ = [("lr", C.c_float * GT_RAW_FLOATS), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("step", C.c_uint32)]


class GenParams(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("n_clusters", C.c_uint32), ("box_half", C.c_float),
        ("cluster_sigma", C.c_float), ("log_scale_mean", C.c_float), ("log_scale_std", C.c_float),
        ("logit_mean", C.c_float), ("logit_std", C.c_float), ("dc_std", C.c_float),
        ("rest_std", C.c_float), ("seed", C.c_uint64),
    ]


_P = C.c_void_p
_lib.gt_last_error.restype = C.c_char_p
_lib.gt_version.restype = C.c_char_p
_lib.gt_uniforms_init.argtypes = [C.POINTER(Uniforms), C.POINTER(C.c_float), C.c_double, C.c_uint32, C.c_uint32,
                                  C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint32,
                                  C.c_int32, C.c_uint32]
_lib.gt_camera_look_at.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
_lib.gt_scene_create_raw.argtypes = [_P, C.c_uint32, C.c_double, C.POINTER(_P)]
_lib.gt_scene_create_packed.argtypes = [_P, C.c_uint32, C.c_double, C.POINTER(_P)]
_lib.gt_scene_set_alpha_min.argtypes = [_P, C.c_double]
_lib.gt_scene_create_device.argtypes = [_P, C.c_uint32, C.c_double, _P, C.POINTER(_P)]
_lib.gt_scene_update_device.argtypes = [_P, _P, _P]
_lib.gt_scene_refit_device.argtypes = [_P, _P, _P, C.POINTER(C.c_uint32)]
_lib.gt_scene_create_raw_gpu.argtypes = [_P, C.c_uint32, C.c_double, C.POINTER(_P)]
_lib.gt_scene_destroy.argtypes = [_P]
_lib.gt_scene_destroy.restype = None
_lib.gt_scene_get_info.argtypes = [_P, C.POINTER(SceneInfo)]
_lib.gt_scene_get_host_data.argtypes = [_P, _P, _P, _P, _P]
_lib.gt_host_prepare.argtypes = [_P, C.c_uint32, C.c_double, _P, _P, _P, _P, C.POINTER(C.c_uint32)]
_lib.gt_debug_face_test.argtypes = [_P, _P, _P, C.c_uint32, C.c_float, C.c_float, C.c_float, _P, _P]
_lib.gt_scene_set_tuning.argtypes = [_P, C.POINTER(Tuning)]
_lib.gt_scene_get_tuning.argtypes = [_P, C.POINTER(Tuning)]
_lib.gt_batch_create.argtypes = [_P, C.c_uint32, _P, _P, _P, _P, C.POINTER(_P)]
_lib.gt_batch_render.argtypes = [_P, C.c_int, _P]
_lib.gt_batch_rays.argtypes = [_P, C.POINTER(C.c_uint64)]
_lib.gt_batch_render_noalloc.argtypes = [_P, C.c_uint32, _P, _P, _P, _P, _P]
_lib.gt_batch_destroy.argtypes = [_P]
_lib.gt_batch_destroy.restype = None
_lib.gt_autotune_frame.argtypes = [_P, C.POINTER(Uniforms), C.c_uint32, C.c_uint32, _P, C.c_uint32, _P,
                                   C.POINTER(AutotuneResult)]
_lib.gt_autotune_batch.argtypes = [_P, _P, C.c_int, C.c_uint32, _P, C.POINTER(AutotuneResult)]
_lib.gt_backward.argtypes = [_P, C.POINTER(Uniforms), C.c_uint32, C.c_uint32, _P, _P, _P, C.c_uint32, _P, _P,
                             C.c_uint64, _P]
_lib.gt_render.argtypes = [_P, C.POINTER(Uniforms), C.c_uint32, C.c_uint32, C.POINTER(Frame), C.c_int, _P]
_lib.gt_adam_step.argtypes = [_P, _P, _P, _P, C.c_uint32, C.POINTER(Adam), _P]
_lib.gt_generate_scene.argtypes = [C.POINTER(GenParams), _P]
_lib.gt_ply_read.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32),
                             C.POINTER(C.c_uint32)]
_lib.gt_free.argtypes = [_P]
_lib.gt_free.restype = None
_lib.gt_tonemap.argtypes = [_P, C.c_uint32, C.c_uint32, C.c_uint32, _P, C.c_int, _P]
_lib.gt_write_png.argtypes = [C.c_char_p, _P, C.c_uint32, C.c_uint32]

TONEMAP_IDENTITY, TONEMAP_EXPOSURE, TONEMAP_EXPOSURE_GAMMA = 0, 1, 2

EXPORTED = [
    "gt_uniforms_init", "gt_camera_look_at", "gt_scene_create_raw", "gt_scene_create_packed",
    "gt_scene_set_alpha_min", "gt_scene_create_device", "gt_scene_create_raw_gpu", "gt_scene_update_device", "gt_scene_refit_device", "gt_scene_destroy", "gt_scene_get_info", "gt_scene_get_host_data",
    "gt_render", "gt_backward", "gt_autotune_frame", "gt_autotune_batch", "gt_adam_step", "gt_batch_create", "gt_batch_render", "gt_batch_rays", "gt_batch_destroy", "gt_batch_render_noalloc",
    "gt_scene_set_tuning", "gt_scene_get_tuning", "gt_host_prepare", "gt_debug_face_test", "gt_generate_scene", "gt_ply_read", "gt_free", "gt_tonemap", "gt_write_png", "gt_last_error", "gt_version",
]


def lib():
    return _lib


def _check(rc):
    if rc != GT_OK:
        raise GTError(rc, _lib.gt_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def last_error():
    return _lib.gt_last_error().decode(errors="replace")


def camera_look_at(pos, target):
    """WASDCamera matrix (camera.ts) looking from pos at target, column-major f32[16]."""
    p = (C.c_float * 3)(*pos)
    t = (C.c_float * 3)(*target)
    m = (C.c_float * 16)()
    _lib.gt_camera_look_at(p, t, m)
    return np.array(m[:], dtype=np.float32)


def make_uniforms(to_world, width, height, hfov_deg=60.0, t_min=1e-3, t_max=1e5, T_min=0.03,
                  alpha_min=0.2, k=6, sh_degree_max=3, early_stop=0, debug_mode=0, accumulated_frames=0):
    """upadateUniformData (src/main.ts:520-534) with the reference's defaults (448-464)."""
    u = Uniforms()
    tw = (C.c_float * 16)(*[float(x) for x in np.asarray(to_world, dtype=np.float32).ravel()])
    _lib.gt_uniforms_init(C.byref(u), tw, float(hfov_deg), width, height, t_min, t_max, T_min, alpha_min,
                          k, sh_degree_max, early_stop, debug_mode)
    u.accumulated_frames = accumulated_frames
    return u


def uniforms_as_dict(u):
    return {
        "to_world": np.array(u.to_world[:], dtype=np.float32), "aspect": (u.camera_aspect[0], u.camera_aspect[1]),
        "t_min": u.t_min, "t_max": u.t_max, "T_min": u.T_min, "alpha_min": u.alpha_min, "k": u.k,
        "sh_degree_max": u.sh_degree_max, "early_stop": u.early_stop, "debug_mode": u.debug_mode,
    }


# Seeded synthetic scenes of the benchmark configurations
CONFIGS = {
    # 20k Gaussians uniform in a 4^3 box, log-scale N(-3, 0.5), logit N(0, 1.5)
    "c1": dict(n=20000, n_clusters=0, box_half=2.0, cluster_sigma=0.0, log_scale_mean=-3.0,
               log_scale_std=0.5, logit_mean=0.0, logit_std=1.5),
    # 1M Gaussians, 64 clusters in a 20^3 box, offsets N(0, 0.8), log-scale N(-4, 0.6), logit N(0.5, 1.5)
    "c2": dict(n=1000000, n_clusters=64, box_half=10.0, cluster_sigma=0.8, log_scale_mean=-4.0,
               log_scale_std=0.6, logit_mean=0.5, logit_std=1.5),
    # 200k Gaussians, one cluster N(0, 0.8) at the origin, log-scale N(-3.5, 0.5)
    "c4": dict(n=200000, n_clusters=1, box_half=0.0, cluster_sigma=0.8, log_scale_mean=-3.5,
               log_scale_std=0.5, logit_mean=0.5, logit_std=1.5),
    # 3M Gaussians, 256 clusters in a 40^3 box, log-scale N(-4.5, 0.7)
    "c5": dict(n=3000000, n_clusters=256, box_half=20.0, cluster_sigma=0.8, log_scale_mean=-4.5,
               log_scale_std=0.7, logit_mean=0.5, logit_std=1.5),
}


def generate_raw(n=None, seed=0, config=None, **kw):
    """PLY-raw parameters (n x 59 float32) from the library's seeded generator."""
    d = dict(n_clusters=0, box_half=2.0, cluster_sigma=0.8, log_scale_mean=-3.0, log_scale_std=0.5,
             logit_mean=0.0, logit_std=1.5, dc_std=0.5, rest_std=0.05)
    if config is not None:
        d.update(CONFIGS[config])
    d.update(kw)
    if n is not None:
        d["n"] = n
    g = GenParams(seed=seed, **d)
    raw = np.zeros((g.n, GT_RAW_FLOATS), dtype=np.float32)
    _check(_lib.gt_generate_scene(C.byref(g), _ptr(raw)))
    return raw


def read_ply(path):
    """PLY-raw records of a 3DGS .ply (GS_ply.ts semantics) and its SH degree."""
    p = C.POINTER(C.c_float)()
    n = C.c_uint32()
    deg = C.c_uint32()
    _check(_lib.gt_ply_read(path.encode(), C.byref(p), C.byref(n), C.byref(deg)))
    try:
        raw = np.ctypeslib.as_array(p, shape=(n.value * GT_RAW_FLOATS,)).copy().reshape(n.value, GT_RAW_FLOATS)
    finally:
        _lib.gt_free(C.cast(p, C.c_void_p))
    return raw, deg.value


def tonemap(rgba, mode=TONEMAP_IDENTITY):
    """8-bit RGBA (H, W, 4) of a float frame as the reference's canvas shows it
    (src/toneMapper.ts:11-26): the curve of mode on RGB, alpha 255."""
    rgba = np.ascontiguousarray(rgba, np.float32)
    H, W = rgba.shape[0], rgba.shape[1]
    assert rgba.shape == (H, W, 4)
    out = np.empty((H, W, 4), np.uint8)
    _check(_lib.gt_tonemap(_ptr(rgba), W, H, mode, _ptr(out), 0, None))
    return out


def tonemap_device(rgba_ptr, width, height, out_ptr, mode=TONEMAP_IDENTITY, stream_ptr=None):
    """The same on device buffers (float RGBA in, 8-bit RGBA out), asynchronous on the stream."""
    _check(_lib.gt_tonemap(rgba_ptr, width, height, mode, out_ptr, 1, stream_ptr))


def write_png(path, rgba8):
    """PNG file of an (H, W, 4) uint8 image (the reference's Save button, src/main.ts:606-611)."""
    rgba8 = np.ascontiguousarray(rgba8, np.uint8)
    H, W = rgba8.shape[0], rgba8.shape[1]
    assert rgba8.shape == (H, W, 4)
    _check(_lib.gt_write_png(os.fsencode(path), _ptr(rgba8), W, H))


def debug_face_test(w2o, o, d, tmin=1e-3, tmax=1e5, margin_scale=1.0):
    """Candidate and exact face masks of (proxy, ray) pairs (kernel pre-test self-check)."""
    w2o = np.ascontiguousarray(w2o, np.float32).reshape(-1, 12)
    o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
    n = w2o.shape[0]
    cand = np.zeros(n, np.uint32)
    exact = np.zeros(n, np.uint32)
    _check(_lib.gt_debug_face_test(_ptr(w2o), _ptr(o), _ptr(d), n, tmin, tmax, margin_scale, _ptr(cand),
                                   _ptr(exact)))
    return cand, exact


def host_prepare(raw, alpha_min):
    """Host stage of scene creation (no GPU needed): packed records and proxies."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    n = raw.shape[0]
    packed = np.zeros((n, 80), np.float32)
    ids = np.zeros(n, np.uint32)
    o2w = np.zeros((n, 12), np.float32)
    w2o = np.zeros((n, 12), np.float32)
    npx = C.c_uint32()
    _check(_lib.gt_host_prepare(_ptr(raw), n, float(alpha_min), _ptr(packed), _ptr(ids), _ptr(o2w), _ptr(w2o),
                                C.byref(npx)))
    m = npx.value
    return dict(packed=packed, proxy_ids=ids[:m].copy(), o2w=o2w[:m].copy(), w2o=w2o[:m].copy())


class Scene:
    """Acceleration container over the Gaussians (gaussianSceneToBvh, src/main.ts:30-111).

    raw / packed: host arrays, built on the host (binned SAH), or for raw
    with gpu_build=True uploaded and built on the GPU.  raw_dev: a
    device pointer to n x 59 PLY-raw floats (e.g. a torch tensor's
    data_ptr()), built on the GPU (gt_scene_create_device); update_device()
    rebuilds it after the parameters change."""

    def __init__(self, raw=None, alpha_min=0.2, packed=None, raw_dev=None, n=None, stream_ptr=None,
                 gpu_build=False):
        self._h = C.c_void_p()
        if raw_dev is not None:
            if n is None:
                raise ValueError("n is required with raw_dev")
            _check(_lib.gt_scene_create_device(raw_dev, int(n), float(alpha_min), stream_ptr, C.byref(self._h)))
        elif raw is not None:
            raw = np.ascontiguousarray(raw, dtype=np.float32)
            assert raw.ndim == 2 and raw.shape[1] == GT_RAW_FLOATS
            create = _lib.gt_scene_create_raw_gpu if gpu_build else _lib.gt_scene_create_raw
            _check(create(_ptr(raw), raw.shape[0], float(alpha_min), C.byref(self._h)))
        elif packed is not None:
            packed = np.ascontiguousarray(packed, dtype=np.float32).reshape(-1, 80)
            _check(_lib.gt_scene_create_packed(_ptr(packed), packed.shape[0], float(alpha_min), C.byref(self._h)))
        else:
            raise ValueError("raw or packed is required")

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _lib.gt_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_alpha_min(self, alpha_min):
        _check(_lib.gt_scene_set_alpha_min(self._h, float(alpha_min)))

    def update_device(self, raw_dev, stream_ptr=None):
        """Rebuilds a device-built scene from updated device parameters (same n)."""
        _check(_lib.gt_scene_update_device(self._h, raw_dev, stream_ptr))

    def refit_device(self, raw_dev, stream_ptr=None):
        """Updates the scene from device parameters (same n) keeping its tree
        (boxes refit); renders equal a fresh build's.  Returns True when the
        set of proxies changed and the scene was rebuilt on the GPU instead."""
        rebuilt = C.c_uint32(0)
        _check(_lib.gt_scene_refit_device(self._h, raw_dev, stream_ptr, C.byref(rebuilt)))
        return bool(rebuilt.value)

    def set_tuning(self, path=0, refill_min=0, comp_min=0, blocks_per_cu=0, tile_w=0, stack_lds=0, rects=0,
                   colors=0):
        """Launch tuning (results do not depend on it); 0 = library default.
        rects: screen-rectangle culling of one-camera frames, 0 when the frame
        has at least a quarter as many rays as the scene has proxies, 1 always,
        2 never.  colors: one-camera renders composite per-Gaussian colours
        computed for the frame's origin before the traversal, 0 when the frame
        has at least half as many rays as the scene has Gaussians, 1 always,
        2 never (each hit evaluates its SH record)."""
        _check(_lib.gt_scene_set_tuning(self._h, C.byref(Tuning(path, refill_min, comp_min, blocks_per_cu, tile_w,
                                                                stack_lds, rects, colors))))

    def get_tuning(self):
        t = Tuning()
        _check(_lib.gt_scene_get_tuning(self._h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def info(self):
        i = SceneInfo()
        _check(_lib.gt_scene_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in SceneInfo._fields_ if not f.startswith("_")}

    def host_data(self):
        inf = self.info()
        n, npx = inf["n_gaussians"], inf["n_proxies"]
        packed = np.zeros((n, 80), np.float32)
        ids = np.zeros(npx, np.uint32)
        o2w = np.zeros((npx, 12), np.float32)
        w2o = np.zeros((npx, 12), np.float32)
        _check(_lib.gt_scene_get_host_data(self._h, _ptr(packed), _ptr(ids), _ptr(o2w), _ptr(w2o)))
        return dict(packed=packed, proxy_ids=ids, o2w=o2w, w2o=w2o)

    def render(self, u, width, height, max_rounds=0, want_T=False, want_stats=False, checkpoint_cap=0,
               antialiasing=False, rgba_prev=None, out=None):
        """traceRays(W, H) with host buffers: returns rgba (H, W, 4) and, on request,
        hit lists (H*W, max_rounds, k), accept masks, round counts, T, stats and
        the backward checkpoint (composited ids per ray, capacity checkpoint_cap;
        checkpoint_cap=-1 returns only the per-ray counts).  antialiasing: the
        reference's jittered-ray path; with u.accumulated_frames > 0 the frame
        is averaged into rgba_prev (the previous average).  out: a float32
        (height, width, 4) array to render into (a render loop reuses it)."""
        R = width * height
        if out is not None:
            assert out.dtype == np.float32 and out.shape == (height, width, 4) and out.flags.c_contiguous
            rgba = out
        else:
            rgba = np.zeros((height, width, 4), np.float32)
        if rgba_prev is not None:
            rgba[...] = np.asarray(rgba_prev, np.float32).reshape(height, width, 4)
        out = {"rgba": rgba}
        f = Frame()
        f.rgba = _ptr(rgba)
        f.antialiasing = 1 if antialiasing else 0
        if max_rounds:
            hits = np.zeros((R, max_rounds, u.k), np.uint32)
            acc = np.zeros((R, max_rounds), np.uint32)
            rounds = np.zeros(R, np.uint32)
            f.hit_ids, f.accept, f.n_rounds = _ptr(hits), _ptr(acc), _ptr(rounds)
            f.max_rounds = max_rounds
            out.update(hit_ids=hits, accept=acc, n_rounds=rounds)
        if want_T or max_rounds:
            T = np.zeros(R, np.float32)
            f.T = _ptr(T)
            out["T"] = T
        if want_stats:
            st = np.zeros(GT_NSTATS, np.uint64)
            f.stats = _ptr(st)
            out["stats"] = st
        if checkpoint_cap:
            ncomp = np.zeros(R, np.uint32)
            f.n_comp = _ptr(ncomp)
            out.update(n_comp=ncomp)
            if checkpoint_cap > 0:
                cids = np.zeros((R, checkpoint_cap), np.uint32)
                f.comp_ids, f.comp_cap = _ptr(cids), checkpoint_cap
                out.update(comp_ids=cids)
        _check(_lib.gt_render(self._h, C.byref(u), width, height, C.byref(f), 0, None))
        if want_stats:
            out["stats"] = dict(zip(STAT_NAMES, [int(x) for x in out["stats"]]))
        return out

    def render_device(self, u, width, height, rgba_ptr, stream_ptr=None, stats_ptr=None, comp_ids_ptr=None,
                      n_comp_ptr=None, comp_cap=0, antialiasing=False, hit_ids_ptr=None, accept_ptr=None,
                      n_rounds_ptr=None, T_ptr=None, max_rounds=0):
        """Asynchronous traceRays into caller-owned device memory (GT_BUF_DEVICE),
        optionally writing the backward checkpoint (or, with antialiasing,
        averaging the jittered frame into rgba), and the per-ray outputs of
        gt_frame: hit lists (R x max_rounds x k ids), accept masks
        (R x max_rounds), round counts and final T (R each)."""
        f = Frame()
        f.rgba = rgba_ptr
        f.antialiasing = 1 if antialiasing else 0
        f.stats = stats_ptr
        f.comp_ids, f.n_comp, f.comp_cap = comp_ids_ptr, n_comp_ptr, comp_cap
        f.hit_ids, f.accept, f.n_rounds, f.T = hit_ids_ptr, accept_ptr, n_rounds_ptr, T_ptr
        f.max_rounds = max_rounds
        _check(_lib.gt_render(self._h, C.byref(u), width, height, C.byref(f), 1, stream_ptr))


class Batch:
    """Many cameras rendered by one launch (optionally a replayed hipGraph).
    frames: list of dicts with device pointers rgba (required) and optional
    hit_ids, accept, n_rounds, T, max_rounds; stats_ptr accumulates counters."""

    def __init__(self, scene, uniforms, sizes, frames, stats_ptr=None):
        n = len(uniforms)
        U = (Uniforms * n)(*uniforms)
        Wa = (C.c_uint32 * n)(*[w for w, _ in sizes])
        Ha = (C.c_uint32 * n)(*[h for _, h in sizes])
        F = (Frame * n)()
        for i, fr in enumerate(frames):
            F[i].rgba = fr["rgba"]
            F[i].hit_ids = fr.get("hit_ids")
            F[i].accept = fr.get("accept")
            F[i].n_rounds = fr.get("n_rounds")
            F[i].T = fr.get("T")
            F[i].max_rounds = fr.get("max_rounds", 0)
            F[i].antialiasing = 1 if fr.get("antialiasing") else 0
        F[0].stats = stats_ptr
        self._scene = scene  # keeps the scene alive
        self._h = C.c_void_p()
        _check(_lib.gt_batch_create(scene.handle, n, U, Wa, Ha, F, C.byref(self._h)))

    def render(self, stream_ptr=None, use_graph=False):
        _check(_lib.gt_batch_render(self._h, 1 if use_graph else 0, stream_ptr))

    def rays(self):
        r = C.c_uint64()
        _check(_lib.gt_batch_rays(self._h, C.byref(r)))
        return r.value

    def close(self):
        if self._h:
            _lib.gt_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NOALLOC_MAX_CAMERAS = 256


def batch_render_noalloc(scene, uniforms, sizes, frames, stats_ptr=None, stream_ptr=None):
    """The cameras of a Batch rendered by one call that allocates no device
    memory (gt_batch_render_noalloc); frames as for Batch.  Asynchronous."""
    n = len(uniforms)
    U = (Uniforms * n)(*uniforms)
    Wa = (C.c_uint32 * n)(*[w for w, _ in sizes])
    Ha = (C.c_uint32 * n)(*[h for _, h in sizes])
    F = (Frame * n)()
    for i, fr in enumerate(frames):
        F[i].rgba = fr["rgba"]
        F[i].hit_ids = fr.get("hit_ids")
        F[i].accept = fr.get("accept")
        F[i].n_rounds = fr.get("n_rounds")
        F[i].T = fr.get("T")
        F[i].max_rounds = fr.get("max_rounds", 0)
    F[0].stats = stats_ptr
    _check(_lib.gt_batch_render_noalloc(scene.handle, n, U, Wa, Ha, F, stream_ptr))


def _autotune_dict(r):
    return dict(best={f: getattr(r.best, f) for f, _ in Tuning._fields_}, best_ms=r.best_ms, first_ms=r.first_ms,
                n_tried=r.n_tried)


def autotune_frame(scene, u, width, height, rgba_ptr, iters=3, stream_ptr=None):
    """Picks the fastest launch tuning for this frame and stores it in the scene."""
    r = AutotuneResult()
    _check(_lib.gt_autotune_frame(scene.handle, C.byref(u), width, height, rgba_ptr, iters, stream_ptr, C.byref(r)))
    return _autotune_dict(r)


def autotune_batch(scene, batch, use_graph=True, iters=3, stream_ptr=None):
    """Picks the fastest launch tuning for this batch and stores it in the scene."""
    r = AutotuneResult()
    _check(_lib.gt_autotune_batch(scene.handle, batch._h, 1 if use_graph else 0, iters, stream_ptr, C.byref(r)))
    return _autotune_dict(r)


def autotune_batch_k(scene, uniforms, sizes, ks=(4, 8, 16, 32), tile_ws=(4, 8, 16, 32), iters=2,
                     stream_ptr=None):
    """Opt-in choice of hit_array_size for a batch, with the ray-tile shape.

    k changes the image (round boundaries move the T_min checks and the 1e-4
    relaunch window), so it is the caller's to accept: for each k the
    cameras are rendered one gt_render at a time (the plain kernel at that
    k) and as one batch under a hipGraph for each tile shape of the
    persistent packet kernel; the batch must equal the plain renders bit
    for bit, and the fastest tile shape is kept.  Returns one dict per k
    (ms per batch, the plain launches' ms, tile_w, equal) and the fastest k.
    Uses torch for the device buffers; leaves the scene's tuning as found."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    stream = torch.cuda.current_stream()
    saved = scene.get_tuning()
    out = []

    def timed(fn, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    try:
        for k in ks:
            us = []
            for u in uniforms:
                v = Uniforms.from_buffer_copy(u)
                v.k = k
                us.append(v)
            plain = [torch.empty((h, w, 4), dtype=torch.float32, device=dev) for w, h in sizes]
            bufs = [torch.empty((h, w, 4), dtype=torch.float32, device=dev) for w, h in sizes]
            scene.set_tuning(**saved)

            def render_plain():
                for v, (w, h), b in zip(us, sizes, plain):
                    scene.render_device(v, w, h, b.data_ptr(), stream_ptr)
            render_plain()
            plain_ms = timed(render_plain, iters)
            best = None
            for tw in tile_ws:
                scene.set_tuning(path=4, tile_w=tw)
                b = Batch(scene, us, sizes, [dict(rgba=x.data_ptr()) for x in bufs])
                b.render(stream_ptr, use_graph=True)  # capture
                ms = timed(lambda: b.render(stream_ptr, use_graph=True), iters)
                equal = all(torch.equal(x.view(torch.int32), y.view(torch.int32)) for x, y in zip(bufs, plain))
                b.close()
                if not equal:
                    raise GTError(GT_OK - 1, f"batch at k={k} tile_w={tw} differs from the plain renders")
                if best is None or ms < best[0]:
                    best = (ms, tw)
            out.append(dict(k=k, ms=best[0], tile_w=best[1], plain_ms=plain_ms, equal_to_plain=True))
            del plain, bufs
    finally:
        scene.set_tuning(**saved)
    fastest = min(out, key=lambda r: r["ms"])["k"]
    return out, fastest


def backward(scene, u, width, height, dl_drgba_ptr, grads_ptr, comp_ids_ptr=None, n_comp_ptr=None, comp_cap=0,
             overflow_ptr=None, budget_bytes=0, stream_ptr=None):
    """Accumulates dLoss/draw (n x 59, device) for one frame; device pointers."""
    _check(_lib.gt_backward(scene.handle, C.byref(u), width, height, dl_drgba_ptr, comp_ids_ptr, n_comp_ptr,
                            comp_cap, grads_ptr, overflow_ptr, budget_bytes, stream_ptr))


def adam_step(params_ptr, grads_ptr, m_ptr, v_ptr, n, lr, step, beta1=0.9, beta2=0.999, eps=1e-15, stream_ptr=None):
    """One in-place Adam update of n x 59 device parameters; lr: 59 per-field rates."""
    a = Adam()
    lr = np.broadcast_to(np.asarray(lr, np.float32), (GT_RAW_FLOATS,))
    for i in range(GT_RAW_FLOATS):
        a.lr[i] = float(lr[i])
    a.beta1, a.beta2, a.eps, a.step = beta1, beta2, eps, int(step)
    _check(_lib.gt_adam_step(params_ptr, grads_ptr, m_ptr, v_ptr, int(n), C.byref(a), stream_ptr))
