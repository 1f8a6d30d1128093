"""ctypes binding of liboracle.so, the CPU restatement of the reference path.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  See gt_oracle.h for the
reference files each function restates.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"liboracle.so not built at {LIB_PATH}; run __graft_entry__.build()")
_lib = C.CDLL(LIB_PATH)

RAW = 59
PACKED = 80
NSTATS = 8
SENTINEL = 0xFFFFFFFF
STAT_NAMES = ["rays", "rounds", "h_eval", "h_comp", "proxies", "hits", "ties", "overflow"]


class Scene(C.Structure):
    _fields_ = [("n", C.c_uint32), ("packed", C.c_void_p), ("n_proxies", C.c_uint32),
                ("proxy_ids", C.c_void_p), ("o2w", C.c_void_p), ("w2o", C.c_void_p), ("bvh", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [("to_world", C.c_float * 16), ("aspect", C.c_float * 2),
                ("t_min", C.c_float), ("t_max", C.c_float), ("T_min", C.c_float), ("alpha_min", C.c_float),
                ("k", C.c_uint32), ("sh_degree_max", C.c_uint32), ("accumulated_frames", C.c_uint32),
                ("early_stop", C.c_uint32), ("debug_mode", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("antialias", C.c_uint32)]


_P = C.c_void_p
_lib.gto_activate.argtypes = [_P, C.c_uint32, _P, _P]
_lib.gto_build_proxies.argtypes = [_P, _P, C.c_uint32, C.c_double, _P, _P, _P]
_lib.gto_build_proxies.restype = C.c_uint32
_lib.gto_activate_packed.argtypes = [_P, C.c_uint32, _P]
_lib.gto_activate_packed.restype = None
_lib.gto_build_proxies_packed.argtypes = [_P, _P, C.c_uint32, C.c_double, _P, _P, _P]
_lib.gto_build_proxies_packed.restype = C.c_uint32
_lib.gto_render.argtypes = [C.POINTER(Scene), C.POINTER(Params), C.c_int, C.c_int, _P, C.c_uint32,
                            _P, _P, _P, C.c_uint32, _P, _P, _P]
_lib.gto_reeval_f64.argtypes = [C.POINTER(Scene), C.POINTER(Params), _P, C.c_uint32, _P, C.c_uint32, _P, _P,
                                _P, _P]
_lib.gto_reeval_f64.restype = None
_lib.gto_bvh_build.argtypes = [C.POINTER(Scene)]
_lib.gto_bvh_build.restype = C.c_void_p
_lib.gto_bvh_free.argtypes = [C.c_void_p]
_lib.gto_bvh_free.restype = None
_lib.gto_set_mutation.argtypes = [C.c_int]
_lib.gto_set_mutation.restype = None
_lib.gto_ray_dir.argtypes = [C.POINTER(Params), C.c_uint32, C.c_uint32, _P]
_lib.gto_ray_dir.restype = None
_lib.gto_exp64.argtypes = [C.c_double]
_lib.gto_exp64.restype = C.c_double
_lib.gto_log64.argtypes = [C.c_double]
_lib.gto_log64.restype = C.c_double
_lib.gto_response.argtypes = [_P, _P, _P]
_lib.gto_response.restype = C.c_double
_lib.gto_radiance.argtypes = [_P, C.c_uint32, _P, _P]
_lib.gto_radiance.restype = None
_lib.gto_tri_hit.argtypes = [_P, _P, C.c_float, C.c_float, C.c_int, _P]
_lib.gto_tri_hit.restype = C.c_int


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def set_mutation(m):
    _lib.gto_set_mutation(int(m))


def activate(raw):
    raw = np.ascontiguousarray(raw, np.float32)
    n = raw.shape[0]
    packed = np.zeros((n, PACKED), np.float32)
    act = np.zeros((n, 8), np.float64)
    _lib.gto_activate(_p(raw), n, _p(packed), _p(act))
    return packed, act


def build_proxies(raw, act, alpha_min):
    raw = np.ascontiguousarray(raw, np.float32)
    n = raw.shape[0]
    ids = np.zeros(n, np.uint32)
    o2w = np.zeros((n, 12), np.float32)
    w2o = np.zeros((n, 12), np.float32)
    npx = _lib.gto_build_proxies(_p(raw), _p(act), n, float(alpha_min), _p(ids), _p(o2w), _p(w2o))
    return ids[:npx].copy(), o2w[:npx].copy(), w2o[:npx].copy()


def build_proxies_packed(packed, alpha_min):
    """Proxies of packed 320-B records (n x 80 floats) from their f32 fields."""
    packed = np.ascontiguousarray(packed, np.float32).reshape(-1, PACKED)
    n = packed.shape[0]
    act = np.zeros((n, 8), np.float64)
    _lib.gto_activate_packed(_p(packed), n, _p(act))
    ids = np.zeros(n, np.uint32)
    o2w = np.zeros((n, 12), np.float32)
    w2o = np.zeros((n, 12), np.float32)
    npx = _lib.gto_build_proxies_packed(_p(packed), _p(act), n, float(alpha_min), _p(ids), _p(o2w), _p(w2o))
    return act, ids[:npx].copy(), o2w[:npx].copy(), w2o[:npx].copy()


class OracleScene:
    """Activated records + proxies of a PLY-raw scene (or, with packed=, of
    packed 320-B records whose proxies come from the f32 fields), built by
    the oracle."""

    def __init__(self, raw=None, alpha_min=0.2, packed=None):
        if packed is not None:
            self.raw = None
            self.packed = np.ascontiguousarray(packed, np.float32).reshape(-1, PACKED)
            self.act, self.proxy_ids, self.o2w, self.w2o = build_proxies_packed(self.packed, alpha_min)
        else:
            self.raw = np.ascontiguousarray(raw, np.float32)
            self.packed, self.act = activate(self.raw)
            self.proxy_ids, self.o2w, self.w2o = build_proxies(self.raw, self.act, alpha_min)
        self._s = Scene(self.packed.shape[0], _p(self.packed), len(self.proxy_ids), _p(self.proxy_ids),
                        _p(self.o2w), _p(self.w2o), None)

    def prebuild_bvh(self):
        """Builds the oracle BVH once so later renders skip the build."""
        if not self._s.bvh:
            self._s.bvh = _lib.gto_bvh_build(C.byref(self._s))

    def __del__(self):
        try:
            if self._s.bvh:
                _lib.gto_bvh_free(self._s.bvh)
                self._s.bvh = None
        except Exception:
            pass

    @property
    def c(self):
        return self._s


def params_from_uniforms(u, width, height, antialias=False):
    p = Params()
    p.antialias = 1 if antialias else 0
    for i in range(16):
        p.to_world[i] = u.to_world[i]
    p.aspect[0], p.aspect[1] = u.camera_aspect[0], u.camera_aspect[1]
    p.t_min, p.t_max, p.T_min, p.alpha_min = u.t_min, u.t_max, u.T_min, u.alpha_min
    p.k, p.sh_degree_max, p.accumulated_frames = u.k, u.sh_degree_max, u.accumulated_frames
    p.early_stop, p.debug_mode = u.early_stop, u.debug_mode
    p.width, p.height = width, height
    return p


def render(scene, params, max_rounds=0, use_bvh=True, threads=None, ray_list=None, rgba_prev=None):
    """Returns dict(rgba (R,4), T, n_rounds, hit_ids (R,MR,k), accept (R,MR), stats).
    rgba_prev: the antialiasing running average the new frame is averaged into."""
    if threads is None:
        threads = os.cpu_count() or 1
    R = params.width * params.height if ray_list is None else len(ray_list)
    rl = None if ray_list is None else np.ascontiguousarray(ray_list, np.uint32)
    rgba = np.zeros((R, 4), np.float32) if rgba_prev is None else np.array(rgba_prev, np.float32).reshape(R, 4)
    T = np.zeros(R, np.float32)
    rounds = np.zeros(R, np.uint32)
    hits = acc = None
    if max_rounds:
        hits = np.zeros((R, max_rounds, params.k), np.uint32)
        acc = np.zeros((R, max_rounds), np.uint32)
    st = np.zeros(NSTATS, np.uint64)
    rc = _lib.gto_render(C.byref(scene.c), C.byref(params), int(use_bvh), int(threads), _p(rl),
                         0 if rl is None else len(rl), _p(rgba), _p(T), _p(rounds), max_rounds, _p(hits), _p(acc),
                         _p(st))
    if rc != 0:
        raise RuntimeError(f"oracle render failed: {rc}")
    return dict(rgba=rgba, T=T, n_rounds=rounds, hit_ids=hits, accept=acc,
                stats=dict(zip(STAT_NAMES, [int(x) for x in st])))


def reeval_f64(scene, params, n_rounds, hit_ids, accept, ray_list=None):
    """FP64 pixels (R,3) and T of given hit lists and accept masks."""
    R = len(n_rounds)
    MR = hit_ids.shape[1]
    rl = None if ray_list is None else np.ascontiguousarray(ray_list, np.uint32)
    rgb = np.zeros((R, 3), np.float64)
    T = np.zeros(R, np.float64)
    _lib.gto_reeval_f64(C.byref(scene.c), C.byref(params), _p(rl), 0 if rl is None else len(rl),
                        _p(np.ascontiguousarray(n_rounds, np.uint32)), MR,
                        _p(np.ascontiguousarray(hit_ids, np.uint32)), _p(np.ascontiguousarray(accept, np.uint32)),
                        _p(rgb), _p(T))
    return rgb, T


def ray_dir(params, px, py):
    d = np.zeros(3, np.float32)
    _lib.gto_ray_dir(C.byref(params), px, py, _p(d))
    return d


def exp64(x):
    return _lib.gto_exp64(float(x))


def log64(x):
    return _lib.gto_log64(float(x))


def response(rec, o, d):
    rec = np.ascontiguousarray(rec, np.float32)
    return _lib.gto_response(_p(rec), _p(np.asarray(o, np.float32)), _p(np.asarray(d, np.float32)))


def radiance(rec, deg, d):
    rec = np.ascontiguousarray(rec, np.float32)
    out = np.zeros(3, np.float32)
    _lib.gto_radiance(_p(rec), deg, _p(np.ascontiguousarray(d, np.float32)), _p(out))
    return out


def tri_hit(o, d, tmin, tmax, face):
    t = np.zeros(1, np.float32)
    h = _lib.gto_tri_hit(_p(np.ascontiguousarray(o, np.float32)), _p(np.ascontiguousarray(d, np.float32)),
                         float(tmin), float(tmax), int(face), _p(t))
    return bool(h), float(t[0])
