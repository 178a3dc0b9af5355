"""CPU ORACLE - test infrastructure only.

ctypes bindings of liboracle.so (oracle_geom.cpp, oracle_pt.cpp), the CPU
restatement of the reference's path-tracing shaders.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker
or the CPU baseline; the product library never does.
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


def _load():
    if not LIB_PATH.exists():
        from .build import build
        build(verbose=False)
    return ctypes.CDLL(str(LIB_PATH))


_lib = _load()
_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p


def _sig(name, args, res=None):
    fn = getattr(_lib, name)
    fn.argtypes = args
    fn.restype = res
    return fn


_sig("orc_rand_u32", [ctypes.c_uint32], ctypes.c_uint32)
_sig("orc_rand_f32", [ctypes.c_uint32], ctypes.c_float)
_sig("orc_rand_u32_many", [_u32p, ctypes.c_uint64, _u32p])
_sig("orc_rand_f32_many", [_u32p, ctypes.c_uint64, _f32p])
_sig("orc_create_ray", [_f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                        _f32p, _f32p])
_sig("orc_create_rays", [_f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                         _f32p])
_sig("orc_intersect_brute", [_f32p, _u32p, ctypes.c_uint64, _f32p, ctypes.c_uint64, _vp])
_sig("orc_bvh_create", [_f32p, ctypes.c_uint64, _u32p, ctypes.c_uint64], _vp)
_sig("orc_bvh_destroy", [_vp])
_sig("orc_bvh_intersect", [_vp, _f32p, ctypes.c_uint64, _vp])
_sig("orc_num_threads", [], ctypes.c_int)
_sig("orc_scene_create", [_f32p, ctypes.c_uint64, _u32p, ctypes.c_uint64, _u32p, _u32p,
                          ctypes.c_uint64, _u32p, ctypes.c_uint64, _u32p, ctypes.c_uint32,
                          ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                          ctypes.c_uint32], _vp)
_sig("orc_scene_destroy", [_vp])
_sig("orc_scene_set_alpha_test", [_vp, ctypes.c_int], None)
_sig("orc_scene_set_object_space", [_vp, _f32p, _vp], None)
_sig("orc_trace_debug", [_vp, _f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                         ctypes.c_float, ctypes.c_float, _f32p], None)
_sig("orc_trace_plain", [_vp, _f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                         ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                         ctypes.c_int, _f32p, _f32p, _f32p], ctypes.c_uint64)
_sig("orc_oren_nayar", [_f32p, _f32p, ctypes.c_float, _f32p, _f32p])
_sig("orc_reflectance", [ctypes.c_float, ctypes.c_float], ctypes.c_float)
_sig("orc_sample_vndf_ggx", [ctypes.c_uint32, _f32p, ctypes.c_float, ctypes.c_float, _f32p])
_sig("orc_cosine_hemisphere", [ctypes.c_uint32, _f32p, _f32p, _f32p])
_sig("orc_high_colour", [ctypes.c_uint32, _f32p])
_sig("orc_background", [_vp, _f32p, _f32p])

HIT_DTYPE = np.dtype([("prim", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
MISS = 0xFFFFFFFF


def _f(a):
    return a.ctypes.data_as(_f32p)


def _u(a):
    return a.ctypes.data_as(_u32p)


def num_threads() -> int:
    return _lib.orc_num_threads()


def rand_u32(seeds) -> np.ndarray:
    s = np.ascontiguousarray(seeds, np.uint32).reshape(-1)
    out = np.empty_like(s)
    _lib.orc_rand_u32_many(_u(s), len(s), _u(out))
    return out


def rand_f32(seeds) -> np.ndarray:
    s = np.ascontiguousarray(seeds, np.uint32).reshape(-1)
    out = np.empty(len(s), np.float32)
    _lib.orc_rand_f32_many(_u(s), len(s), _f(out))
    return out


def camera_floats(camera_bytes: bytes) -> np.ndarray:
    a = np.frombuffer(camera_bytes, np.float32)
    assert a.size == 32
    return np.ascontiguousarray(a)


def create_rays(camera_bytes: bytes, width: int, height: int, tmin=1e-4, tmax=1000.0):
    cam = camera_floats(camera_bytes)
    rays = np.empty((width * height, 8), np.float32)
    _lib.orc_create_rays(_f(cam), width, height, tmin, tmax, _f(rays))
    return rays


def _vidx(vertices, indices):
    v = np.ascontiguousarray(vertices, np.float32)
    if v.shape[1] == 3:
        v = np.ascontiguousarray(np.concatenate([v, np.zeros((len(v), 1), np.float32)], 1))
    idx = None if indices is None else np.ascontiguousarray(indices, np.uint32).reshape(-1)
    ntri = len(idx) // 3 if idx is not None else len(v) // 3
    return v, idx, ntri


def intersect_brute(vertices, indices, rays) -> np.ndarray:
    v, idx, ntri = _vidx(vertices, indices)
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    hits = np.empty(len(r), HIT_DTYPE)
    _lib.orc_intersect_brute(_f(v), None if idx is None else _u(idx), ntri, _f(r), len(r),
                             hits.ctypes.data_as(_vp))
    return hits


class Bvh:
    def __init__(self, vertices, indices):
        v, idx, ntri = _vidx(vertices, indices)
        self._keep = (v, idx)
        self._h = _lib.orc_bvh_create(_f(v), len(v), None if idx is None else _u(idx), ntri)

    def intersect(self, rays) -> np.ndarray:
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.empty(len(r), HIT_DTYPE)
        _lib.orc_bvh_intersect(self._h, _f(r), len(r), hits.ctypes.data_as(_vp))
        return hits

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.orc_bvh_destroy(self._h)
            self._h = None


def _solid(rgba):
    return np.asarray(rgba, np.uint8).reshape(1, 1, 4)


class Scene:
    """Oracle copy of a scene: same vertices, materials and texel values.

    Texture arrays follow the order diffuse[], emission[], attributes[],
    recolour (default 1x1 white), 6 background faces (default black)."""

    def __init__(self, spec, object_space=None):
        """object_space: (per-primitive object-space triangles (n, 9), per-
        primitive mirrored flag) of an instanced scene, or None."""
        v, idx, ntri = _vidx(spec.vertices, spec.indices)
        images = list(spec.diffuse) + list(spec.emission) + list(spec.attributes)
        diffuse_base = 0
        emission_base = len(spec.diffuse)
        attr_base = emission_base + len(spec.emission)
        recolour = len(images)
        images.append(spec.recolour if spec.recolour is not None else _solid((255, 255, 255, 255)))
        bg_base = len(images)
        if spec.background is not None:
            images.extend(list(spec.background))
        else:
            images.extend([_solid((0, 0, 0, 255))] * 6)
        descs, pools, off = [], [], 0
        for im in images:
            a = np.ascontiguousarray(im, np.uint8)
            h, w = a.shape[:2]
            descs += [off, w, h]
            pools.append(a.reshape(-1, 4).view(np.uint32).reshape(-1))
            off += w * h
        texels = np.ascontiguousarray(np.concatenate(pools))
        d = np.asarray(descs, np.uint32)
        mi = np.ascontiguousarray(spec.material_idx, np.uint32)
        mats = np.ascontiguousarray(spec.materials, np.uint32).reshape(-1, 8)
        self._keep = (v, idx, mi, mats, texels, d)
        self._h = _lib.orc_scene_create(_f(v), len(v), None if idx is None else _u(idx), ntri,
                                        _u(mi), _u(mats), len(mats), _u(texels), len(texels),
                                        _u(d), len(images), diffuse_base, emission_base,
                                        attr_base, recolour, bg_base)
        _lib.orc_scene_set_alpha_test(self._h, 0 if getattr(spec, "opaque", True) else 1)
        if object_space is not None:
            obj, mirrored = object_space
            obj = np.ascontiguousarray(obj, np.float32).reshape(ntri, 9)
            mir = np.ascontiguousarray(mirrored, np.uint8).reshape(ntri)
            self._keep = self._keep + (obj, mir)
            _lib.orc_scene_set_object_space(self._h, _f(obj), mir.ctypes.data_as(_vp))
        self.spec = spec

    def trace_plain(self, camera_bytes: bytes, width: int, height: int, quality: int = 1,
                    samples: int = 1, tmin: float = 1e-4, tmax: float = 1000.0, bounces: int = 5,
                    rr_start: Optional[int] = None, seed_offset: int = 0, fp64: bool = False,
                    row0: int = 0, rows: Optional[int] = None):
        """Returns (colour, normal, albedo, path_vertices) for rows
        [row0, row0 + rows) of a width x height image."""
        rows = height - row0 if rows is None else rows
        cam = camera_floats(camera_bytes)
        c = np.empty((rows, width, 4), np.float32)
        n = np.empty_like(c)
        a = np.empty_like(c)
        verts = _lib.orc_trace_plain(self._h, _f(cam), width, height, row0, rows, quality, samples,
                                     tmin, tmax, bounces,
                                     0xFFFFFFFF if rr_start is None else rr_start, seed_offset,
                                     1 if fp64 else 0, _f(c), _f(n), _f(a))
        return c, n, a, verts

    def trace_debug(self, camera_bytes: bytes, width: int, height: int, kind: int,
                    tmin: float = 1e-4, tmax: float = 1000.0) -> np.ndarray:
        """The debugging shaders (src/debug/*.wesl): kind 0 front face,
        1 reflectance, 2 tangent; (height, width, 4) colour."""
        cam = camera_floats(camera_bytes)
        c = np.empty((height, width, 4), np.float32)
        _lib.orc_trace_debug(self._h, _f(cam), width, height, kind, tmin, tmax, _f(c))
        return c

    def background(self, direction) -> np.ndarray:
        d = np.ascontiguousarray(direction, np.float32)
        out = np.empty(3, np.float32)
        _lib.orc_background(self._h, _f(d), _f(out))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.orc_scene_destroy(self._h)
            self._h = None


def oren_nayar(incoming, reflected, roughness, color) -> np.ndarray:
    out = np.empty(3, np.float32)
    _lib.orc_oren_nayar(_f(np.asarray(incoming, np.float32)), _f(np.asarray(reflected, np.float32)),
                        roughness, _f(np.asarray(color, np.float32)), _f(out))
    return out


def reflectance(cosine: float, ref_idx: float) -> float:
    return _lib.orc_reflectance(cosine, ref_idx)


def sample_vndf_ggx(seed: int, view, rx: float, ry: float) -> np.ndarray:
    out = np.empty(3, np.float32)
    _lib.orc_sample_vndf_ggx(seed, _f(np.asarray(view, np.float32)), rx, ry, _f(out))
    return out


def cosine_hemisphere(seed: int, normal, tangent) -> np.ndarray:
    out = np.empty(3, np.float32)
    _lib.orc_cosine_hemisphere(seed, _f(np.asarray(normal, np.float32)),
                               _f(np.asarray(tangent, np.float32)), _f(out))
    return out


def high_colour(seed: int):
    out = np.empty(4, np.float32)
    _lib.orc_high_colour(seed, _f(out))
    return out[:3], float(out[3])


# --------------------------------------------------------- ReSTIR / EOF
_sig("orc_normalised_to_u32", [_f32p], ctypes.c_uint32)
_sig("orc_u32_to_normalised", [ctypes.c_uint32, _f32p])
_sig("orc_merge", [_vp, _vp, ctypes.c_float, ctypes.c_uint32], ctypes.c_int)
_sig("orc_update", [_vp, _vp, ctypes.c_float, ctypes.c_uint32])
_sig("orc_temporal", [_vp, _vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32])
_sig("orc_spatial", [_vp, _f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                     ctypes.c_uint32, ctypes.c_uint32, _f32p, ctypes.POINTER(ctypes.c_int32)])
_sig("orc_eof_gi", [_vp, ctypes.c_uint64, ctypes.c_uint32])
_sig("orc_eof_world", [_u32p, ctypes.c_uint64])

# bindings.wesl:158-169 Reservoir (64 bytes)
RESERVOIR_DTYPE = np.dtype([("visible_point", "<f4", 3), ("visible_normal", "<u4"),
                            ("sample_point", "<f4", 3), ("sample_normal", "<u4"),
                            ("out_radiance", "<f4", 3), ("ty", "<u4"), ("roughness", "<f4"),
                            ("w", "<f4"), ("packed_confidence_valid", "<u4"), ("W", "<f4")])
assert RESERVOIR_DTYPE.itemsize == 64
# bindings.wesl:33-40 WorldMarcovChainState (56 bytes)
WORLD_DTYPE = np.dtype([("secondary_hash", "<u4"), ("num_samples", "<u4"), ("luminance", "<f4", 3),
                        ("lock", "<u4"), ("chain_light_source", "<f4", 3),
                        ("chain_mean_cosine", "<f4"), ("chain_weight_sum", "<f4"),
                        ("chain_num_samples", "<u4"), ("chain_score", "<f4"), ("timeout", "<u4")])
assert WORLD_DTYPE.itemsize == 56


def normalised_to_u32(n) -> int:
    return int(_lib.orc_normalised_to_u32(_f(np.ascontiguousarray(n, np.float32))))


def u32_to_normalised(u: int) -> np.ndarray:
    out = np.empty(3, np.float32)
    _lib.orc_u32_to_normalised(int(u), _f(out))
    return out


def merge(res: np.ndarray, r: np.ndarray, p: float, seed: int) -> bool:
    """In place on a 1-element RESERVOIR_DTYPE array `res`."""
    return bool(_lib.orc_merge(res.ctypes.data_as(_vp), r.ctypes.data_as(_vp), p, seed))


def update(res: np.ndarray, s: np.ndarray, w: float, seed: int) -> None:
    _lib.orc_update(res.ctypes.data_as(_vp), s.ctypes.data_as(_vp), w, seed)


def temporal(gi: np.ndarray, gi_old: np.ndarray, width: int, height: int, seed_offset: int):
    out = np.ascontiguousarray(gi.copy())
    old = np.ascontiguousarray(gi_old)
    _lib.orc_temporal(out.ctypes.data_as(_vp), old.ctypes.data_as(_vp), width, height, seed_offset)
    return out


def spatial(gi: np.ndarray, info: np.ndarray, width: int, height: int, is_samples: int,
            is_space: int, seed_offset: int):
    """Returns (rgba image, written mask)."""
    g = np.ascontiguousarray(gi)
    inf = np.ascontiguousarray(info, np.float32).reshape(-1)
    out = np.zeros((height, width, 4), np.float32)
    written = np.zeros((height, width), np.int32)
    _lib.orc_spatial(g.ctypes.data_as(_vp), _f(inf), width, height, is_samples, is_space,
                     seed_offset, _f(out), written.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out, written.astype(bool)


def eof_gi(res: np.ndarray, cap: int) -> np.ndarray:
    out = np.ascontiguousarray(res.copy())
    _lib.orc_eof_gi(out.ctypes.data_as(_vp), len(out), cap)
    return out


def eof_world(w: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(w.copy())
    _lib.orc_eof_world(out.view(np.uint32).ctypes.data_as(_u32p), len(out))
    return out


_sig("orc_trace_guided", [_vp, _f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                          ctypes.c_uint32, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                          ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                          _vp, _vp, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint64, _vp,
                          ctypes.c_uint64, _f32p, _f32p, _f32p, _f32p], ctypes.c_uint64)


class GuidedState:
    """Persistent state of the guided mode for the oracle (one frame's
    inputs/outputs, WGSL layouts)."""

    def __init__(self, width, height, world_size, screen=True, spatial=True):
        n = width * height
        self.world = np.zeros(max(world_size, 1), WORLD_DTYPE)
        self.world_old = np.zeros(max(world_size, 1), WORLD_DTYPE)
        self.mc_old = np.zeros((n if screen else 1, 8), np.uint32)
        self.mc = np.zeros_like(self.mc_old)
        self.gi = np.zeros(n if spatial else 1, RESERVOIR_DTYPE)
        self.info = np.zeros((n, 9), np.float32)


def trace_guided(scene: "Scene", camera_bytes: bytes, width: int, height: int, state: GuidedState,
                 quality=1, samples=4, tmin=1e-4, tmax=1000.0, bounces=5, rr_start=None,
                 seed_offset=0, restir=2, subgroup=64):
    cam = camera_floats(camera_bytes)
    c = np.empty((height, width, 4), np.float32)
    nrm = np.empty_like(c)
    alb = np.empty_like(c)
    v = _lib.orc_trace_guided(scene._h, _f(cam), width, height, quality, samples, tmin, tmax,
                              bounces, 0xFFFFFFFF if rr_start is None else rr_start, seed_offset,
                              restir, subgroup, state.world.ctypes.data_as(_vp),
                              state.world_old.ctypes.data_as(_vp), len(state.world),
                              _u(state.mc_old), _u(state.mc), len(state.mc_old),
                              state.gi.ctypes.data_as(_vp), len(state.gi), _f(state.info), _f(c),
                              _f(nrm), _f(alb))
    return c, nrm, alb, v


# ---------------------------------------------------------------- neural cache
_f64p = ctypes.POINTER(ctypes.c_double)
_sig("orc_nrc_num_weights", [], ctypes.c_int)
_sig("orc_nrc_encode", [_f32p, ctypes.c_int, _f64p])
_sig("orc_nrc_forward", [_f64p, _f32p, ctypes.c_uint64, ctypes.c_int, _f64p])
_sig("orc_nrc_grad", [_f64p, _f32p, _f32p, ctypes.c_uint64, ctypes.c_int, _f64p, _f64p])
_sig("orc_nrc_adam", [_f32p, _f32p, _f32p, _f32p, ctypes.c_uint64, ctypes.c_float,
                      ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                      ctypes.c_float])

NRC_NUM_WEIGHTS = _lib.orc_nrc_num_weights()
NRC_EXACT, NRC_BF16 = 0, 1


def _d(a):
    return a.ctypes.data_as(_f64p)


def nrc_encode(x, mode=NRC_EXACT) -> np.ndarray:
    """320 encoded features of one 13-input row."""
    xv = np.ascontiguousarray(x, np.float32)
    out = np.empty(320, np.float64)
    _lib.orc_nrc_encode(_f(xv), mode, _d(out))
    return out


def nrc_forward(weights, inputs, mode=NRC_EXACT) -> np.ndarray:
    w = np.ascontiguousarray(weights, np.float64)
    x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 13)
    out = np.empty((len(x), 3), np.float64)
    _lib.orc_nrc_forward(_d(w), _f(x), len(x), mode, _d(out))
    return out


def nrc_grad(weights, inputs, targets, mode=NRC_EXACT):
    """(gradient of the mean relative L2 loss, loss)."""
    w = np.ascontiguousarray(weights, np.float64)
    x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 13)
    t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
    g = np.empty(NRC_NUM_WEIGHTS, np.float64)
    loss = np.zeros(1, np.float64)
    _lib.orc_nrc_grad(_d(w), _f(x), _f(t), len(x), mode, _d(g), _d(loss))
    return g, float(loss[0])


def nrc_adam(w, m, v, grad, lr, beta1, beta2, eps, bias1, bias2):
    """In-place Adam step on float32 arrays (the GPU's float arithmetic)."""
    g = np.ascontiguousarray(grad, np.float32)
    _lib.orc_nrc_adam(_f(w), _f(m), _f(v), _f(g), len(w), lr, beta1, beta2, eps, bias1, bias2)


_sig("orc_spectral_pick", [_f32p, ctypes.c_float, ctypes.c_float, ctypes.c_uint32, _f32p],
     ctypes.c_float)


def spectral_pick(picked: float, lo: float, hi: float, seed: int):
    """Spectral quality at one vertex: (ior, colour weight, new picked)."""
    p = np.array([picked], np.float32)
    c = np.empty(3, np.float32)
    ri = _lib.orc_spectral_pick(_f(p), lo, hi, seed, _f(c))
    return float(ri), c, float(p[0])


# ---------------------------------------------------------------- guided-mode hooks
_sig("orc_strip_nan", [ctypes.c_float], ctypes.c_float)
_sig("orc_grid_key", [_f32p, _f32p, _f32p, _f32p, ctypes.c_uint32, _u32p])
_sig("orc_get_kappa", [_u32p], ctypes.c_float)
_sig("orc_expm1", [ctypes.c_float], ctypes.c_float)
_sig("orc_x_div_expm1", [ctypes.c_float], ctypes.c_float)
_sig("orc_vmf_pdf_many", [_f32p, ctypes.c_float, _f32p, ctypes.c_uint64, _f32p])
_sig("orc_vmf_sample_many", [_f32p, ctypes.c_float, _u32p, ctypes.c_uint64, _f32p])
_sig("orc_update_state", [_f32p, _f32p, ctypes.c_float, ctypes.c_uint32, ctypes.c_int, _u32p,
                          ctypes.c_float, ctypes.c_float, _u32p])
_sig("orc_resample_mcmm", [_u32p, _u32p, _f32p], ctypes.c_int)
_sig("orc_all_vmf_pdfs", [_u32p, ctypes.c_float, _f32p, _f32p], ctypes.c_float)
_sig("orc_grid_op", [ctypes.c_int, _vp, _vp, ctypes.c_uint32, _f32p, _f32p, _f32p, _f32p,
                     ctypes.c_int, _u32p, _f32p, _u32p, _f32p])

NUM_RESAMPLES = 21


def _v3(a):
    return np.ascontiguousarray(a, np.float32).reshape(3)


def chain_words(light_source=(0.0, 0.0, 0.0), mean_cosine=0.0, weight_sum=0.0, num_samples=0,
                score=0.0) -> np.ndarray:
    """One MarcovChainState (bindings.wesl:3-9) as 8 u32 words (last = pad)."""
    w = np.zeros(8, np.uint32)
    w[:5] = np.array([*light_source, mean_cosine, weight_sum], np.float32).view(np.uint32)
    w[5] = num_samples
    w[6] = np.float32(score).view(np.uint32)
    return w


def chain_fields(w) -> dict:
    w = np.asarray(w, np.uint32)
    f = w[:7].view(np.float32)
    return {"light_source": f[:3].copy(), "mean_cosine": float(f[3]), "weight_sum": float(f[4]),
            "num_samples": int(w[5]), "score": float(f[6])}


def strip_nan(x: float) -> float:
    return float(_lib.orc_strip_nan(x))


def grid_key(pos, normal, incoming, cam, table: int) -> dict:
    out = np.zeros(6, np.uint32)
    _lib.orc_grid_key(_f(_v3(pos)), _f(_v3(normal)), _f(_v3(incoming)), _f(_v3(cam)), table,
                      _u(out))
    return {"idx": int(out[0]), "secondary": int(out[1]), "primary": int(out[2]),
            "loc": out[3:6].view(np.int32).copy()}


def get_kappa(chain) -> float:
    return float(_lib.orc_get_kappa(_u(np.ascontiguousarray(chain, np.uint32))))


def expm1(x: float) -> float:
    return float(_lib.orc_expm1(x))


def x_div_expm1(x: float) -> float:
    return float(_lib.orc_x_div_expm1(x))


def vmf_pdf(mean, kappa: float, dirs) -> np.ndarray:
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    out = np.empty(len(d), np.float32)
    _lib.orc_vmf_pdf_many(_f(_v3(mean)), kappa, _f(d), len(d), _f(out))
    return out


def vmf_sample(mean, kappa: float, seeds) -> np.ndarray:
    """Unnormalised Xs of general.wesl:797-808, one per seed."""
    s = np.ascontiguousarray(seeds, np.uint32).reshape(-1)
    out = np.empty((len(s), 3), np.float32)
    _lib.orc_vmf_sample_many(_f(_v3(mean)), kappa, _u(s), len(s), _f(out))
    return out


def update_state(x1, x2, p_hat: float, seed: int, not_vmf: bool, chain, confidence: float,
                 sum_mc: float) -> np.ndarray:
    out = np.zeros(8, np.uint32)
    _lib.orc_update_state(_f(_v3(x1)), _f(_v3(x2)), p_hat, seed, int(not_vmf),
                          _u(np.ascontiguousarray(chain, np.uint32)), confidence, sum_mc, _u(out))
    return out


def resample_mcmm(chains, seed: int):
    """(selected index or -1, seed after, sum_mc) over 21 chains."""
    c = np.ascontiguousarray(chains, np.uint32).reshape(NUM_RESAMPLES, 8)
    s = np.array([seed], np.uint32)
    sm = np.zeros(1, np.float32)
    k = _lib.orc_resample_mcmm(_u(c), _u(s), _f(sm))
    return int(k), int(s[0]), float(sm[0])


def all_vmf_pdfs(chains, sum_mc: float, hit, d) -> float:
    c = np.ascontiguousarray(chains, np.uint32).reshape(NUM_RESAMPLES, 8)
    return float(_lib.orc_all_vmf_pdfs(_u(c), sum_mc, _f(_v3(hit)), _f(_v3(d))))


GRID_LOAD, GRID_STORE, GRID_RADIANCE, GRID_ADD = 0, 1, 2, 3


def grid_op(op: int, world: np.ndarray, world_old: np.ndarray, cam, pos, normal, incoming,
            force_old=False, chain=None, radiance=None):
    """One sequential world-grid call on WORLD_DTYPE arrays (world in place)."""
    assert world.dtype == WORLD_DTYPE and world_old.dtype == WORLD_DTYPE
    assert len(world) == len(world_old)
    cin = np.ascontiguousarray(chain if chain is not None else np.zeros(8), np.uint32)
    rin = _v3(radiance if radiance is not None else (0, 0, 0))
    cout = np.zeros(8, np.uint32)
    rout = np.zeros(3, np.float32)
    _lib.orc_grid_op(op, world.ctypes.data_as(_vp), world_old.ctypes.data_as(_vp), len(world),
                     _f(_v3(cam)), _f(_v3(pos)), _f(_v3(normal)), _f(_v3(incoming)), int(force_old),
                     _u(cin), _f(rin), _u(cout), _f(rout))
    if op == GRID_LOAD:
        return cout
    if op == GRID_RADIANCE:
        return rout
    return None


_sig("orc_screen_draw", [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                         ctypes.c_uint32, _u32p])


def screen_draw(seed: int, x: int, y: int, width: int, subgroup: int = 64) -> np.ndarray:
    """fill_mc_screen_space_states' draws: [lane pixel, lane valid, 12 indices, seed after]."""
    out = np.zeros(15, np.uint32)
    _lib.orc_screen_draw(seed, x, y, width, subgroup, _u(out))
    return out
