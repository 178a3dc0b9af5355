"""Python mirror of the reference's host API over the C ABI (include/lambent.h).

Names and argument meaning follow the reference crate:
``Material.new`` (lib.rs:132-162), ``EmissionScale`` (lib.rs:96-121),
``MaterialType`` (lib.rs:194-203), ``refractive_indices`` (lib.rs:30-56),
``Camera.from_proj_view`` (camera.rs:13-42), ``dispatch_size`` (lib.rs:207-219),
``TextureLoader`` (textures.rs:12-91), ``RayTracer``/``Low``/``Medium``/``High``
(lib.rs:260-329, path_tracing/mod.rs:66-147), ``PathTracerOpts`` /
``RestirTemporalMode`` (path_tracing/mod.rs:7-34), ``DataBuffers`` /
``BufferType`` / ``AdvanceOptions`` (data_buffer/mod.rs:18-472).

Every call goes to liblambent_hip.so; there is no Python or CPU fallback.  A
missing library raises at import of this module.
"""
from __future__ import annotations

import ctypes
import enum
import io
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

_PKG = Path(__file__).resolve().parent
_LIB_PATH = Path(os.environ.get("LMB_LIBRARY", _PKG / "liblambent_hip.so"))

if not _LIB_PATH.exists():
    raise ImportError(
        f"{_LIB_PATH} is missing: build it with `python -m lambent_amd.build` "
        "(the HIP library is required; there is no fallback)")

_lib = ctypes.CDLL(str(_LIB_PATH))

# ----------------------------------------------------------------- errors
LMB_OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_SINGULAR_PROJECTION = -2
ERR_SINGULAR_VIEW = -3
ERR_RESERVED_INDEX = -4
ERR_OUT_OF_MEMORY = -5
ERR_DEVICE = -6
ERR_BVH_TOO_DEEP = -7
ERR_TEXTURE = -8


class LambentError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class CameraError(LambentError):
    """camera.rs:63-66 CameraError (ProjectionUninvertable / ViewUninvertable)."""


class TextureError(LambentError):
    """textures.rs:93-97 TextureError."""


_lib.lmb_last_error.restype = ctypes.c_char_p
_lib.lmb_version.restype = ctypes.c_char_p


def _check(code: int) -> None:
    if code == LMB_OK:
        return
    msg = _lib.lmb_last_error().decode(errors="replace")
    if code in (ERR_SINGULAR_PROJECTION, ERR_SINGULAR_VIEW):
        raise CameraError(code, msg)
    if code == ERR_TEXTURE:
        raise TextureError(code, msg)
    raise LambentError(code, msg)


def version() -> str:
    return _lib.lmb_version().decode()


def library_path() -> str:
    return str(_LIB_PATH)


# ---------------------------------------------------------------- ctypes
_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p


class _MaterialC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "tex_pos_1", "tex_pos_2", "tex_pos_3", "tex_pos_recolour",
        "tex_idx_diffuse_emission", "tex_idx_attributes_ty", "emission_scale",
        "refractive_index")]


class _CameraC(ctypes.Structure):
    _fields_ = [("projection_inverse", ctypes.c_float * 16),
                ("view_inverse", ctypes.c_float * 16)]


class _SceneDescC(ctypes.Structure):
    _fields_ = [("vertices", _f32p), ("num_vertices", ctypes.c_uint64),
                ("indices", _u32p), ("num_indices", ctypes.c_uint64),
                ("material_idx", _u32p), ("materials", ctypes.POINTER(_MaterialC)),
                ("num_materials", ctypes.c_uint64), ("textures", _vp), ("flags", ctypes.c_uint32)]


class _GeometryDescC(ctypes.Structure):
    _fields_ = [("vertices", _f32p), ("num_vertices", ctypes.c_uint64),
                ("indices", _u32p), ("num_indices", ctypes.c_uint64),
                ("aabbs", _f32p), ("num_aabbs", ctypes.c_uint64)]


@dataclass
class AabbGeometry:
    """Procedural geometry (wgpu BlasGeometries::AabbGeometries): boxes as
    (n, 6) floats, min xyz then max xyz.  With the default intersection
    handler (default_intersection_handler.wesl) no box is ever hit."""
    boxes: np.ndarray


class _InstanceC(ctypes.Structure):
    _fields_ = [("geometry", ctypes.c_uint32), ("transform", ctypes.c_float * 12),
                ("material_idx", _u32p), ("mask", ctypes.c_uint32)]


class _InstancedDescC(ctypes.Structure):
    _fields_ = [("geometries", ctypes.POINTER(_GeometryDescC)), ("num_geometries", ctypes.c_uint32),
                ("instances", ctypes.POINTER(_InstanceC)), ("num_instances", ctypes.c_uint32),
                ("materials", ctypes.POINTER(_MaterialC)), ("num_materials", ctypes.c_uint64),
                ("textures", _vp), ("flags", ctypes.c_uint32)]


class _BvhStatsC(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_uint64), ("num_nodes", ctypes.c_uint64),
                ("max_depth", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("sah_cost", ctypes.c_double), ("build_ms", ctypes.c_double),
                ("device_bytes", ctypes.c_uint64)]


class _KernelStatsC(ctypes.Structure):
    _fields_ = [("extend_ms", ctypes.c_double), ("shade_ms", ctypes.c_double),
                ("extend_launches", ctypes.c_uint64), ("shade_launches", ctypes.c_uint64),
                ("path_vertices", ctypes.c_uint64), ("cache_infer_ms", ctypes.c_double),
                ("cache_train_ms", ctypes.c_double), ("cache_infer_launches", ctypes.c_uint64),
                ("cache_train_steps", ctypes.c_uint64)]


class _BatchConfigC(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_uint32), ("refill", ctypes.c_uint32),
                ("tile_w", ctypes.c_uint32), ("tile_h", ctypes.c_uint32),
                ("pass_pixels", ctypes.c_uint32), ("use_graph", ctypes.c_uint32)]


class _NrcOptsC(ctypes.Structure):
    _fields_ = [("learning_rate", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("epsilon", ctypes.c_float),
                ("train_paths", ctypes.c_uint32), ("train_steps", ctypes.c_uint32),
                ("seed", ctypes.c_uint32), ("ema_decay", ctypes.c_float),
                ("defer_specular", ctypes.c_uint32), ("overlap_training", ctypes.c_uint32),
                ("stratify", ctypes.c_uint32)]


class _NrcStatsC(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("records", ctypes.c_uint64),
                ("queries", ctypes.c_uint64), ("loss", ctypes.c_float),
                ("reserved", ctypes.c_uint32)]


class _TracerOptsC(ctypes.Structure):
    _fields_ = [("quality", ctypes.c_uint32), ("samples", ctypes.c_uint32),
                ("t_min", ctypes.c_float), ("t_max", ctypes.c_float),
                ("restir", ctypes.c_uint32), ("bounces", ctypes.c_uint32),
                ("rr_start", ctypes.c_uint32), ("guiding", ctypes.c_uint32)]


def _sig(name, argtypes, restype=ctypes.c_int32):
    fn = getattr(_lib, name)
    fn.argtypes = argtypes
    fn.restype = restype
    return fn


_sig("lmb_emission_scale_pack", [ctypes.c_float, ctypes.c_float, _u32p])
_sig("lmb_material_new", [_f32p, _f32p, _f32p, _u16p, ctypes.c_uint16, ctypes.c_int32,
                          ctypes.c_int32, ctypes.c_uint32, _f32p, ctypes.c_uint32,
                          ctypes.POINTER(_MaterialC)])
_sig("lmb_camera_from_proj_view", [_f32p, _f32p, ctypes.POINTER(_CameraC)])
_sig("lmb_camera_update_from_proj_view", [ctypes.POINTER(_CameraC), _f32p, _f32p])
_sig("lmb_dispatch_size", [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p], None)
_sig("lmb_textures_create", [ctypes.POINTER(_vp)])
_sig("lmb_textures_add", [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u8p, _u32p])
_sig("lmb_textures_set_recolour", [_vp, ctypes.c_uint32, ctypes.c_uint32, _u8p])
_sig("lmb_textures_set_background", [_vp, ctypes.c_uint32, _u8p])
_sig("lmb_textures_destroy", [_vp])
_sig("lmb_scene_create", [ctypes.POINTER(_SceneDescC), ctypes.POINTER(_vp)])
_sig("lmb_scene_destroy", [_vp])
_sig("lmb_scene_create_instanced", [ctypes.POINTER(_InstancedDescC), ctypes.POINTER(_vp)])
_sig("lmb_scene_update_instances", [_vp, ctypes.POINTER(_InstanceC), ctypes.c_uint32])
_sig("lmb_scene_bvh_stats", [_vp, ctypes.POINTER(_BvhStatsC)])
_sig("lmb_scene_intersect", [_vp, _f32p, ctypes.c_uint64, _vp])
_sig("lmb_scene_intersect_device", [_vp, _vp, ctypes.c_uint64, _vp, _vp])
_sig("lmb_scene_intersect_device_noalloc", [_vp, _vp, ctypes.c_uint64, _vp, _vp])
_sig("lmb_primary_hits", [_vp, ctypes.POINTER(_CameraC), ctypes.c_uint32, ctypes.c_uint32,
                          ctypes.c_float, ctypes.c_float, _u32p, _f32p])
_sig("lmb_sync", [_vp])

# Entry points of the tracer and per-frame state (declared in lambent.h);
# bound lazily so that a partial library still imports for the CPU tests.
_OPTIONAL = {
    "lmb_tracer_opts_default": ([ctypes.POINTER(_TracerOptsC)], None),
    "lmb_tracer_create": ([ctypes.POINTER(_TracerOptsC), ctypes.POINTER(_vp)], ctypes.c_int32),
    "lmb_tracer_destroy": ([_vp], ctypes.c_int32),
    "lmb_data_buffers_create": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.POINTER(_vp)], ctypes.c_int32),
    "lmb_data_buffers_resize": ([_vp, ctypes.c_uint32, ctypes.c_uint32], ctypes.c_int32),
    "lmb_data_buffers_destroy": ([_vp], ctypes.c_int32),
    "lmb_data_buffers_output": ([_vp, ctypes.c_uint32, ctypes.POINTER(_vp),
                                 ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int32),
    "lmb_data_buffers_read_output": ([_vp, ctypes.c_uint32, _f32p], ctypes.c_int32),
    "lmb_data_buffers_pack_rgb": ([_vp, ctypes.c_uint32, _vp, _vp], ctypes.c_int32),
    "lmb_frame_ring_create": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.POINTER(_vp)], ctypes.c_int32),
    "lmb_frame_ring_destroy": ([_vp], ctypes.c_int32),
    "lmb_frame_ring_push": ([_vp, _vp, _vp], ctypes.c_int32),
    "lmb_frame_ring_count": ([_vp, _u32p], ctypes.c_int32),
    "lmb_frame_ring_average": ([_vp, _vp], ctypes.c_int32),
    "lmb_frame_ring_output": ([_vp, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_uint64)],
                              ctypes.c_int32),
    "lmb_frame_ring_read_average": ([_vp, _f32p], ctypes.c_int32),
    "lmb_data_buffers_state_size": ([_vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)],
                                    ctypes.c_int32),
    "lmb_data_buffers_read_state": ([_vp, ctypes.c_uint32, _vp, ctypes.c_uint64],
                                    ctypes.c_int32),
    "lmb_data_buffers_write_state": ([_vp, ctypes.c_uint32, _vp, ctypes.c_uint64],
                                     ctypes.c_int32),
    "lmb_data_buffers_path_vertices": ([_vp, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int32),
    "lmb_data_buffers_set_profiling": ([_vp, ctypes.c_int32], ctypes.c_int32),
    "lmb_data_buffers_kernel_stats": ([_vp, ctypes.POINTER(_KernelStatsC)], ctypes.c_int32),
    "lmb_trace": ([_vp, _vp, ctypes.POINTER(_CameraC), _vp, ctypes.c_uint32, _vp],
                  ctypes.c_int32),
    "lmb_temporal_resample": ([_vp, ctypes.c_uint32, _vp], ctypes.c_int32),
    "lmb_spatial_resample": ([_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _vp],
                             ctypes.c_int32),
    "lmb_advance_frame": ([_vp, ctypes.c_uint32, ctypes.c_uint32, _vp], ctypes.c_int32),
    "lmb_nrc_opts_default": ([ctypes.POINTER(_NrcOptsC)], None),
    "lmb_nrc_create": ([ctypes.POINTER(_NrcOptsC), ctypes.POINTER(_vp)], ctypes.c_int32),
    "lmb_nrc_destroy": ([_vp], ctypes.c_int32),
    "lmb_nrc_get_weights": ([_vp, _f32p], ctypes.c_int32),
    "lmb_nrc_set_weights": ([_vp, _f32p], ctypes.c_int32),
    "lmb_nrc_infer": ([_vp, _f32p, ctypes.c_uint64, _f32p], ctypes.c_int32),
    "lmb_nrc_gradient": ([_vp, _f32p, _f32p, ctypes.c_uint64, _f32p, _f32p], ctypes.c_int32),
    "lmb_nrc_train": ([_vp, _f32p, _f32p, ctypes.c_uint64, _f32p], ctypes.c_int32),
    "lmb_nrc_apply_gradient": ([_vp, _f32p], ctypes.c_int32),
    "lmb_nrc_set_opts": ([_vp, ctypes.POINTER(_NrcOptsC)], ctypes.c_int32),
    "lmb_nrc_stats_get": ([_vp, ctypes.POINTER(_NrcStatsC)], ctypes.c_int32),
    "lmb_trace_nrc": ([_vp, _vp, ctypes.POINTER(_CameraC), _vp, ctypes.c_uint32, _vp, _vp],
                      ctypes.c_int32),
    "lmb_trace_nrc_train": ([_vp, _vp, ctypes.POINTER(_CameraC), _vp, ctypes.c_uint32, _vp,
                             _vp], ctypes.c_int32),
    "lmb_batch_config_default": ([ctypes.POINTER(_BatchConfigC)], None),
    "lmb_batch_default_sweep": ([ctypes.POINTER(_BatchConfigC), ctypes.c_uint32], ctypes.c_uint32),
    "lmb_batch_create": ([_vp, ctypes.c_uint32, ctypes.POINTER(_TracerOptsC),
                          ctypes.POINTER(_BatchConfigC), ctypes.POINTER(_vp)], ctypes.c_int32),
    "lmb_batch_destroy": ([_vp], ctypes.c_int32),
    "lmb_batch_set_config": ([_vp, ctypes.POINTER(_BatchConfigC)], ctypes.c_int32),
    "lmb_batch_get_config": ([_vp, ctypes.POINTER(_BatchConfigC)], ctypes.c_int32),
    "lmb_batch_set_camera": ([_vp, ctypes.c_uint32, ctypes.POINTER(_CameraC)], ctypes.c_int32),
    "lmb_batch_trace": ([_vp, ctypes.c_uint32, _vp], ctypes.c_int32),
    "lmb_batch_read_output": ([_vp, ctypes.c_uint32, _f32p], ctypes.c_int32),
    "lmb_batch_path_vertices": ([_vp, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int32),
    "lmb_batch_autotune": ([_vp, ctypes.POINTER(_BatchConfigC), ctypes.c_uint32, ctypes.c_uint32,
                            ctypes.POINTER(_BatchConfigC), ctypes.POINTER(ctypes.c_double)],
                           ctypes.c_int32),
}
for _name, (_args, _res) in _OPTIONAL.items():
    if hasattr(_lib, _name):
        _sig(_name, _args, _res)


def exported_symbols() -> list[str]:
    """Names declared in lambent.h that the loaded library exports."""
    return [n for n in HEADER_FUNCTIONS if hasattr(_lib, n)]


HEADER_FUNCTIONS = [
    "lmb_last_error", "lmb_version", "lmb_emission_scale_pack", "lmb_material_new",
    "lmb_camera_from_proj_view", "lmb_camera_update_from_proj_view", "lmb_dispatch_size",
    "lmb_textures_create", "lmb_textures_add", "lmb_textures_set_recolour",
    "lmb_textures_set_background", "lmb_textures_destroy", "lmb_scene_create",
    "lmb_scene_destroy", "lmb_scene_bvh_stats", "lmb_scene_intersect",
    "lmb_scene_intersect_device", "lmb_scene_intersect_device_noalloc", "lmb_sync",
    "lmb_primary_hits", "lmb_scene_create_instanced", "lmb_scene_update_instances",
    *list(_OPTIONAL),
]


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_f32p)


def _uptr(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


# -------------------------------------------------------------- materials
class MaterialType(enum.IntEnum):
    """lib.rs:194-203."""
    Diffuse = 0
    Metallic = 1
    Transparent = 2


class refractive_indices:  # noqa: N801 - mirrors the Rust module name
    """lib.rs:30-56 (ranges are (n at 760 nm, n at 340 nm))."""
    VACUUM = (1.0, 1.0)
    AIR = (1.00027531, 1.000287)
    ICE = (1.3057, 1.3267)
    WATER = (1.33, 1.3442)
    GLASS = (1.4540, 1.4787)
    DIAMOND = (2.4063, 2.4922)

    @staticmethod
    def refractive_index_between(in_front, behind):
        f32 = np.float32
        return (float(f32(behind[0]) / f32(in_front[0])), float(f32(behind[1]) / f32(in_front[1])))


@dataclass(frozen=True)
class EmissionScale:
    """lib.rs:96-121: two bf16 scales, packed emissive | diffuse << 16."""
    emissive_texture_scale: float = 0.0
    diffuse_texture_scale: float = 0.0

    @staticmethod
    def from_emission_scale(scale: float) -> "EmissionScale":
        return EmissionScale(float(scale), 0.0)

    @staticmethod
    def from_f32(emissive_texture_scale: float, diffuse_texture_scale: float) -> "EmissionScale":
        return EmissionScale(float(emissive_texture_scale), float(diffuse_texture_scale))

    def pack(self) -> int:
        out = ctypes.c_uint32()
        _check(_lib.lmb_emission_scale_pack(self.emissive_texture_scale,
                                            self.diffuse_texture_scale, ctypes.byref(out)))
        return out.value


class Material:
    """lib.rs:83-163: the 32-byte material record."""
    __slots__ = ("_c",)
    DTYPE = np.dtype([(n, "<u4") for n, _ in _MaterialC._fields_])

    def __init__(self, c: _MaterialC):
        self._c = c

    @staticmethod
    def new(tex_pos_1: Sequence[float], tex_pos_2: Sequence[float], tex_pos_3: Sequence[float],
            tex_pos_recolour: Sequence[int], tex_idx_diffuse: int,
            tex_idx_emission: Optional[int], tex_idx_attributes: Optional[int],
            emission_scale: EmissionScale, refractive_index: Optional[Sequence[float]],
            ty: MaterialType) -> "Material":
        p1 = np.asarray(tex_pos_1, np.float32)
        p2 = np.asarray(tex_pos_2, np.float32)
        p3 = np.asarray(tex_pos_3, np.float32)
        rc = np.asarray(tex_pos_recolour, np.uint16)
        ior = None if refractive_index is None else np.asarray(refractive_index, np.float32)
        c = _MaterialC()
        _check(_lib.lmb_material_new(
            _fptr(p1), _fptr(p2), _fptr(p3), rc.ctypes.data_as(_u16p), int(tex_idx_diffuse),
            -1 if tex_idx_emission is None else int(tex_idx_emission),
            -1 if tex_idx_attributes is None else int(tex_idx_attributes),
            emission_scale.pack(), None if ior is None else _fptr(ior), int(ty),
            ctypes.byref(c)))
        return Material(c)

    def words(self) -> tuple:
        return tuple(getattr(self._c, n) for n, _ in _MaterialC._fields_)

    def to_bytes(self) -> bytes:
        return bytes(self._c)

    def __eq__(self, other):
        return isinstance(other, Material) and self.words() == other.words()

    def __repr__(self):
        return "Material(" + ", ".join(f"{w:#010x}" for w in self.words()) + ")"


def materials_buffer(materials: Sequence[Material]) -> bytes:
    """lib.rs:184-192 `<[Material]>::buffer_descriptor` contents."""
    return b"".join(m.to_bytes() for m in materials)


def descriptor(materials: Sequence[Material]) -> bytes:
    """README.md:21-23 name of `<[Material]>::buffer_descriptor` (lib.rs:184-192)."""
    return materials_buffer(materials)


def materials_array(materials: Sequence[Material]) -> np.ndarray:
    return np.frombuffer(materials_buffer(materials), dtype=np.uint32).reshape(-1, 8).copy()


# ----------------------------------------------------------------- camera
class Camera:
    """camera.rs:6-70: inverse projection and view matrices (column-major)."""

    def __init__(self, c: _CameraC):
        self._c = c

    @staticmethod
    def _cols(m) -> np.ndarray:
        a = np.asarray(m, dtype=np.float32)
        if a.shape != (4, 4):
            raise ValueError("matrices are [[f32; 4]; 4] (four columns)")
        return np.ascontiguousarray(a.reshape(16))

    @staticmethod
    def from_proj_view(proj, view) -> "Camera":
        """`proj`/`view` are 4 columns of 4 floats, like cgmath's `.into()`."""
        p, v = Camera._cols(proj), Camera._cols(view)
        c = _CameraC()
        _check(_lib.lmb_camera_from_proj_view(_fptr(p), _fptr(v), ctypes.byref(c)))
        return Camera(c)

    @staticmethod
    def new(proj, view) -> "Camera":
        """README.md:48-50 name of from_proj_view."""
        return Camera.from_proj_view(proj, view)

    def buffer_descriptor(self) -> bytes:
        """camera.rs:59-70: the 128-byte uniform buffer contents."""
        return bytes(self._c)

    descriptor = buffer_descriptor  # README.md:48-50 name

    def update_from_proj_view(self, proj=None, view=None) -> None:
        p = None if proj is None else Camera._cols(proj)
        v = None if view is None else Camera._cols(view)
        _check(_lib.lmb_camera_update_from_proj_view(
            ctypes.byref(self._c), None if p is None else _fptr(p),
            None if v is None else _fptr(v)))

    @property
    def projection_inverse(self) -> np.ndarray:
        return np.array(self._c.projection_inverse, np.float32).reshape(4, 4)

    @property
    def view_inverse(self) -> np.ndarray:
        return np.array(self._c.view_inverse, np.float32).reshape(4, 4)

    def to_bytes(self) -> bytes:
        return bytes(self._c)


def perspective(fovy_deg: float, aspect: float, near: float, far: float) -> np.ndarray:
    """cgmath::perspective (OpenGL clip space), returned as 4 columns, f32."""
    f32 = np.float32
    fovy = f32(np.deg2rad(f32(fovy_deg)))
    f = f32(1.0) / f32(np.tan(fovy / f32(2.0)))
    n, fa, a = f32(near), f32(far), f32(aspect)
    m = np.zeros((4, 4), np.float32)
    m[0, 0] = f / a
    m[1, 1] = f
    m[2, 2] = (fa + n) / (n - fa)
    m[2, 3] = f32(-1.0)
    m[3, 2] = (f32(2.0) * fa * n) / (n - fa)
    return m


def look_to_rh(eye, direction, up) -> np.ndarray:
    """cgmath::Matrix4::look_to_rh, returned as 4 columns, f32."""
    f32 = np.float32
    e = np.asarray(eye, f32)
    f = np.asarray(direction, f32)
    f = f / f32(np.sqrt(f32(np.dot(f, f))))
    s = np.cross(f, np.asarray(up, f32)).astype(f32)
    s = s / f32(np.sqrt(f32(np.dot(s, s))))
    u = np.cross(s, f).astype(f32)
    m = np.zeros((4, 4), np.float32)
    m[0] = (s[0], u[0], -f[0], 0)
    m[1] = (s[1], u[1], -f[1], 0)
    m[2] = (s[2], u[2], -f[2], 0)
    m[3] = (-np.dot(e, s), -np.dot(e, u), np.dot(e, f), 1)
    return m


def look_at_rh(eye, centre, up) -> np.ndarray:
    """cgmath::Matrix4::look_at_rh: look_to_rh towards centre - eye."""
    e = np.asarray(eye, np.float32)
    return look_to_rh(e, np.asarray(centre, np.float32) - e, up)


@dataclass(frozen=True)
class DispatchSize:
    width: int
    height: int


def dispatch_size(width: int, height: int) -> DispatchSize:
    """lib.rs:207-219."""
    gx, gy = ctypes.c_uint32(), ctypes.c_uint32()
    _lib.lmb_dispatch_size(width, height, ctypes.byref(gx), ctypes.byref(gy))
    return DispatchSize(gx.value, gy.value)


# --------------------------------------------------------------- textures
class TextureKind(enum.IntEnum):
    DIFFUSE = 0
    EMISSION = 1
    ATTRIBUTES = 2


class TextureLoader:
    """textures.rs:12-91: decodes images to RGBA8 (the PNG decode uses Pillow,
    where the reference uses the `image` crate)."""

    def __init__(self):
        self.images: list[np.ndarray] = []

    def load(self, path: str) -> None:
        with open(path, "rb") as fh:
            self.load_from_bytes(fh.read())

    def load_from_bytes(self, data: bytes) -> None:
        try:
            from PIL import Image
        except ImportError as exc:  # pragma: no cover
            raise TextureError(ERR_TEXTURE, f"no image decoder available: {exc}") from exc
        try:
            img = Image.open(io.BytesIO(data)).convert("RGBA")
        except Exception as exc:
            raise TextureError(ERR_TEXTURE, f"could not decode image: {exc}") from exc
        self.images.append(np.asarray(img, dtype=np.uint8).copy())

    def load_rgba8(self, rgba: np.ndarray) -> None:
        a = np.ascontiguousarray(rgba, dtype=np.uint8)
        if a.ndim != 3 or a.shape[2] != 4:
            raise TextureError(ERR_TEXTURE, "expected an (h, w, 4) uint8 array")
        self.images.append(a)

    def raw(self) -> list:
        """README.md:36-41: the decoded RGBA8 images, in load order."""
        return list(self.images)

    def create_textures(self, emission: Sequence[np.ndarray] = (),
                        attributes: Sequence[np.ndarray] = ()) -> "TextureSet":
        """textures.rs:41-90: the loaded images as the diffuse textures (with
        optional emission and attribute arrays) of a texture set."""
        return TextureSet(self.images, emission, attributes)

    create_texture = create_textures  # README.md:41-43 name


class TextureSet:
    """bind_group_from_textures (textures.rs:151-243): diffuse, emission and
    attribute arrays, optional recolour texture and the background cube."""

    def __init__(self, diffuse: Sequence[np.ndarray] = (), emission: Sequence[np.ndarray] = (),
                 attributes: Sequence[np.ndarray] = (), recolour: Optional[np.ndarray] = None,
                 background: Optional[np.ndarray] = None):
        h = _vp()
        _check(_lib.lmb_textures_create(ctypes.byref(h)))
        self._h = h
        for kind, imgs in ((TextureKind.DIFFUSE, diffuse), (TextureKind.EMISSION, emission),
                           (TextureKind.ATTRIBUTES, attributes)):
            for img in imgs:
                self.add(kind, img)
        if recolour is not None:
            a = np.ascontiguousarray(recolour, np.uint8)
            _check(_lib.lmb_textures_set_recolour(self._h, a.shape[1], a.shape[0],
                                                  a.ctypes.data_as(_u8p)))
        if background is not None:
            a = np.ascontiguousarray(background, np.uint8)
            if a.ndim != 4 or a.shape[0] != 6 or a.shape[1] != a.shape[2] or a.shape[3] != 4:
                raise TextureError(ERR_TEXTURE, "background must be (6, n, n, 4) uint8")
            _check(_lib.lmb_textures_set_background(self._h, a.shape[1], a.ctypes.data_as(_u8p)))

    def add(self, kind: TextureKind, img: np.ndarray) -> int:
        a = np.ascontiguousarray(img, np.uint8)
        if a.ndim != 3 or a.shape[2] != 4:
            raise TextureError(ERR_TEXTURE, "expected an (h, w, 4) uint8 array")
        idx = ctypes.c_uint32()
        _check(_lib.lmb_textures_add(self._h, int(kind), a.shape[1], a.shape[0],
                                     a.ctypes.data_as(_u8p), ctypes.byref(idx)))
        return idx.value

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _lib.lmb_textures_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solid_rgba8(r: int, g: int, b: int, a: int = 255, size: int = 1) -> np.ndarray:
    img = np.empty((size, size, 4), np.uint8)
    img[...] = (r, g, b, a)
    return img


def uniform_background(value: float, size: int = 1) -> np.ndarray:
    c = int(round(min(max(value, 0.0), 1.0) * 255))
    return np.broadcast_to(np.array([c, c, c, 255], np.uint8), (6, size, size, 4)).copy()


# ------------------------------------------------------------------ scene
@dataclass
class BvhStats:
    num_triangles: int
    num_nodes: int
    max_depth: int
    sah_cost: float
    build_ms: float
    device_bytes: int


SCENE_ALPHA_TEST = 1  # LMB_SCENE_ALPHA_TEST


class Scene:
    """Geometry + materials + textures on the GPU with a compressed 8-wide
    BVH (replaces the BLAS/TLAS build and bind groups 0 and 2).  ``opaque``:
    the OPAQUE geometry flag of the reference's BLAS (main.rs:469); False
    makes every candidate hit pass the any-hit alpha test of
    general.wesl:1057-1066."""

    def __init__(self, vertices: np.ndarray, indices: Optional[np.ndarray],
                 material_idx: np.ndarray, materials: Sequence[Material] | np.ndarray,
                 textures: Optional[TextureSet] = None, opaque: bool = True):
        v = np.ascontiguousarray(vertices, np.float32)
        if v.ndim != 2 or v.shape[1] not in (3, 4):
            raise ValueError("vertices must be (n, 3) or (n, 4) float32")
        if v.shape[1] == 3:
            v = np.ascontiguousarray(np.concatenate([v, np.zeros((len(v), 1), np.float32)], 1))
        idx = None if indices is None else np.ascontiguousarray(indices, np.uint32).reshape(-1)
        mi = np.ascontiguousarray(material_idx, np.uint32).reshape(-1)
        if isinstance(materials, np.ndarray):
            mats = np.ascontiguousarray(materials, np.uint32).reshape(-1, 8)
        else:
            mats = materials_array(materials)
        self._keep = (v, idx, mi, mats, textures)
        d = _SceneDescC()
        d.vertices = _fptr(v)
        d.num_vertices = len(v)
        d.indices = None if idx is None else _uptr(idx)
        d.num_indices = 0 if idx is None else len(idx)
        d.material_idx = _uptr(mi)
        d.materials = mats.ctypes.data_as(ctypes.POINTER(_MaterialC))
        d.num_materials = len(mats)
        d.textures = None if textures is None else textures.handle
        d.flags = 0 if opaque else SCENE_ALPHA_TEST
        h = _vp()
        _check(_lib.lmb_scene_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.num_triangles = (len(idx) // 3) if idx is not None else len(v) // 3

    @classmethod
    def from_instances(cls, geometries, instances, materials: Sequence[Material] | np.ndarray,
                       textures: Optional[TextureSet] = None, opaque: bool = True) -> "Scene":
        """TLAS instances (main.rs:560-563): geometries = [(vertices, indices
        or None) or AabbGeometry], instances = [(geometry, 3x4 row-major
        transform, material_idx per primitive, mask)], baked into one BVH."""
        keep = []
        geo = (_GeometryDescC * max(len(geometries), 1))()
        for k, gd in enumerate(geometries):
            if isinstance(gd, AabbGeometry):
                b = np.ascontiguousarray(gd.boxes, np.float32).reshape(-1, 6)
                keep.append(b)
                geo[k].aabbs, geo[k].num_aabbs = _fptr(b), len(b)
                continue
            v, idx = gd
            v = np.ascontiguousarray(v, np.float32).reshape(-1, 4)
            idx = None if idx is None else np.ascontiguousarray(idx, np.uint32).reshape(-1)
            keep += [v, idx]
            geo[k].vertices, geo[k].num_vertices = _fptr(v), len(v)
            geo[k].indices = None if idx is None else _uptr(idx)
            geo[k].num_indices = 0 if idx is None else len(idx)
        ins = (_InstanceC * max(len(instances), 1))()
        for k, (g, tf, mi, mask) in enumerate(instances):
            mi = np.ascontiguousarray(mi, np.uint32).reshape(-1)
            keep.append(mi)
            ins[k].geometry = int(g)
            ins[k].transform[:] = [float(x) for x in np.asarray(tf, np.float32).reshape(12)]
            ins[k].material_idx = _uptr(mi)
            ins[k].mask = int(mask)
        mats = (np.ascontiguousarray(materials, np.uint32).reshape(-1, 8)
                if isinstance(materials, np.ndarray) else materials_array(materials))
        d = _InstancedDescC()
        d.geometries, d.num_geometries = geo, len(geometries)
        d.instances, d.num_instances = ins, len(instances)
        d.materials = mats.ctypes.data_as(ctypes.POINTER(_MaterialC))
        d.num_materials = len(mats)
        d.textures = None if textures is None else textures.handle
        d.flags = 0 if opaque else SCENE_ALPHA_TEST
        h = _vp()
        _check(_lib.lmb_scene_create_instanced(ctypes.byref(d), ctypes.byref(h)))
        self = cls.__new__(cls)
        self._keep = (keep, mats, textures)
        self._h = h
        self.num_triangles = sum(
            (len(np.asarray(geometries[g][1]).reshape(-1)) // 3 if geometries[g][1] is not None
             else len(np.asarray(geometries[g][0]).reshape(-1, 4)) // 3)
            for g, _, _, mask in instances
            if int(mask) & 0xFF and not isinstance(geometries[g], AabbGeometry))
        return self

    def update_instances(self, instances) -> None:
        """New transforms and masks for every instance of a scene made by
        from_instances (the TLAS rebuild of main.rs:560-563): instances =
        [(geometry, 3x4 transform, material_idx (ignored), mask)] in creation
        order.  Returns when the scene is ready."""
        ins = (_InstanceC * max(len(instances), 1))()
        for k, (g, tf, _mi, mask) in enumerate(instances):
            ins[k].geometry = int(g)
            ins[k].transform[:] = [float(x) for x in np.asarray(tf, np.float32).reshape(12)]
            ins[k].material_idx = None
            ins[k].mask = int(mask)
        _check(_lib.lmb_scene_update_instances(self._h, ins, len(instances)))

    @property
    def handle(self):
        return self._h

    def bvh_stats(self) -> BvhStats:
        s = _BvhStatsC()
        _check(_lib.lmb_scene_bvh_stats(self._h, ctypes.byref(s)))
        return BvhStats(s.num_triangles, s.num_nodes, s.max_depth, s.sah_cost, s.build_ms,
                        s.device_bytes)

    def intersect(self, rays: np.ndarray) -> np.ndarray:
        """rays: (n, 8) float32 (origin, tmin, dir, tmax) -> (n,) structured hits."""
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.empty(len(r), dtype=HIT_DTYPE)
        _check(_lib.lmb_scene_intersect(self._h, _fptr(r), len(r), hits.ctypes.data_as(_vp)))
        return hits

    def intersect_device(self, rays_dev: int, n: int, hits_dev: int, stream=None,
                         allocate: bool = True) -> None:
        """Closest hits of n device rays (32 B each) into device hits (16 B
        each), enqueued on `stream`.  allocate=False takes the allocation-free
        entry point (static scratch; calls on one GPU must not overlap)."""
        fn = _lib.lmb_scene_intersect_device if allocate else _lib.lmb_scene_intersect_device_noalloc
        _check(fn(self._h, rays_dev, n, hits_dev, stream))

    def primary_hits(self, camera: Camera, width: int, height: int, t_min: float = 1e-4,
                     t_max: float = 1000.0):
        prim = np.empty(width * height, np.uint32)
        t = np.empty(width * height, np.float32)
        _check(_lib.lmb_primary_hits(self._h, ctypes.byref(camera._c), width, height, t_min,
                                     t_max, _uptr(prim), _fptr(t)))
        return prim.reshape(height, width), t.reshape(height, width)

    def close(self):
        if self._h:
            _lib.lmb_scene_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


HIT_DTYPE = np.dtype([("prim", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
MISS = 0xFFFFFFFF


def sync(stream=None) -> None:
    _check(_lib.lmb_sync(stream))


# ----------------------------------------------------------------- tracer
class Quality(enum.IntEnum):
    """path_tracing/mod.rs:66-147: High (spectral), Medium, Low; SPECTRAL (no
    reference counterpart): four wavelengths per path with Cauchy dispersion."""
    LOW = 0
    MEDIUM = 1
    HIGH = 2
    SPECTRAL = 3
    DEBUG_FRONT_FACE = 16
    DEBUG_REFLECTANCE = 17
    DEBUG_TANGENT = 18


Low, Medium, High, Spectral = Quality.LOW, Quality.MEDIUM, Quality.HIGH, Quality.SPECTRAL


class debug:  # noqa: N801 - mirrors the Rust module name
    """src/debug/mod.rs:1-58: shaders that show the camera ray's closest hit
    (front face / back face shaded by 1 / (t + 1), the material's IOR pair,
    the triangle's tangent)."""
    FrontFace = Quality.DEBUG_FRONT_FACE
    Reflectance = Quality.DEBUG_REFLECTANCE
    Tangent = Quality.DEBUG_TANGENT


class RestirTemporalMode(enum.IntEnum):
    """path_tracing/mod.rs:25-34."""
    Disabled = 0
    PerSample = 1
    PerPixel = 2


class Guiding(enum.IntEnum):
    OFF = 0
    MARKOV_CHAIN = 1


@dataclass
class PathTracerOpts:
    """path_tracing/mod.rs:7-23 (+ the bounce count, Russian roulette start
    and guiding switch that the reference fixes at compile time)."""
    samples: int = 4
    tmin: float = 0.0001
    tmax: float = 1000.0
    restir_temporal: RestirTemporalMode = RestirTemporalMode.PerPixel
    bounces: int = 5
    rr_start: Optional[int] = None
    guiding: Guiding = Guiding.MARKOV_CHAIN


def _tracer_opts_c(quality: Quality, opts: PathTracerOpts) -> _TracerOptsC:
    c = _TracerOptsC()
    c.quality = int(quality)
    c.samples = int(opts.samples)
    c.t_min = float(opts.tmin)
    c.t_max = float(opts.tmax)
    c.restir = int(opts.restir_temporal)
    c.bounces = int(opts.bounces)
    c.rr_start = 0xFFFFFFFF if opts.rr_start is None else int(opts.rr_start)
    c.guiding = int(opts.guiding)
    return c


class Pipeline:
    """The ray-tracing pipeline (lib.rs:331-376 create_pipeline result)."""

    def __init__(self, quality: Quality, opts: PathTracerOpts):
        c = _tracer_opts_c(quality, opts)
        h = _vp()
        _check(_lib.lmb_tracer_create(ctypes.byref(c), ctypes.byref(h)))
        self._h = h
        self.quality = quality
        self.opts = opts

    def dispatch(self, scene: Scene, camera: Camera, buffers: "DataBuffers", seed_offset: int = 0,
                 stream=None) -> None:
        """general.wesl:506 rt_main over dispatch_size(w, h) with the
        seed_offset immediate (examples/cornell_box/src/main.rs:723-731)."""
        _check(_lib.lmb_trace(self._h, scene.handle, ctypes.byref(camera._c), buffers.handle,
                              int(seed_offset) & 0xFFFFFFFF, stream))

    def dispatch_cached(self, scene: Scene, camera: Camera, buffers: "DataBuffers",
                        cache: "NeuralCache", seed_offset: int = 0, stream=None) -> None:
        """One frame of the tracer with the neural radiance cache (paths end
        at their second vertex in the cache; about cache.opts.train_paths of
        them train it).  No counterpart in the reference."""
        _check(_lib.lmb_trace_nrc(self._h, scene.handle, ctypes.byref(camera._c), buffers.handle,
                                  int(seed_offset) & 0xFFFFFFFF, cache.handle, stream))

    def train_cached(self, scene: Scene, camera: Camera, buffers: "DataBuffers",
                     cache: "NeuralCache", seed_offset: int = 0, stream=None) -> None:
        """A training frame: about cache.opts.train_paths paths at random
        pixels train the cache; no image.  No counterpart in the reference."""
        _check(_lib.lmb_trace_nrc_train(self._h, scene.handle, ctypes.byref(camera._c),
                                        buffers.handle, int(seed_offset) & 0xFFFFFFFF,
                                        cache.handle, stream))

    def close(self):
        if self._h:
            _lib.lmb_tracer_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RayTracer:
    """lib.rs:260-329 RayTracer<S>; `RayTracer(Medium).create_pipeline(opts)`."""

    def __init__(self, quality: Quality = Quality.MEDIUM):
        self.quality = Quality(quality)

    def create_pipeline(self, opts: Optional[PathTracerOpts] = None) -> Pipeline:
        return Pipeline(self.quality, opts or PathTracerOpts())


class BufferType(enum.IntFlag):
    """data_buffer/mod.rs:18-27."""
    MARKOV_CHAIN_SCREEN_SPACE = 1 << 0
    SPATIAL_RESAMPLING = 1 << 1
    MARKOV_CHAIN_WORLD_SPACE = 1 << 2

    @classmethod
    def all(cls) -> "BufferType":
        return cls(7)


@dataclass(frozen=True)
class AdvanceOptions:
    """data_buffer/mod.rs:442-472."""
    confidence_cap: int


AdvanceOptions.RESPONSIVE = AdvanceOptions(2)
AdvanceOptions.SEMI_INTERACTIVE = AdvanceOptions(30)
AdvanceOptions.SLOW = AdvanceOptions(100)


class StateBuffer(enum.IntEnum):
    GI_RESERVOIRS = 0
    OLD_GI_RESERVOIRS = 1
    INFO = 2
    MARKOV_CHAINS = 3
    OLD_MARKOV_CHAINS = 4
    WORLD_MARKOV_CHAINS = 5
    OLD_WORLD_MARKOV_CHAINS = 6


class Output(enum.IntEnum):
    COLOR = 0
    NORMAL = 1
    ALBEDO = 2


class FrameRing:
    """The example application's frame ring and average (main.rs:377-413,
    699-748; average_to_buf.wgsl): push() keeps a trace's colour output in the
    next of `frames` slots, average() returns the mean of the filled slots as
    packed RGB (h, w, 3) float32.  No counterpart in the library API."""

    def __init__(self, width: int, height: int, frames: int):
        h = _vp()
        _check(_lib.lmb_frame_ring_create(width, height, frames, ctypes.byref(h)))
        self._h, self.width, self.height = h, width, height

    def push(self, buffers: "DataBuffers", stream=None) -> None:
        _check(_lib.lmb_frame_ring_push(self._h, buffers.handle, stream))

    def __len__(self) -> int:
        n = ctypes.c_uint32()
        _check(_lib.lmb_frame_ring_count(self._h, ctypes.byref(n)))
        return n.value

    def average(self, stream=None) -> np.ndarray:
        _check(_lib.lmb_frame_ring_average(self._h, stream))
        out = np.empty((self.height, self.width, 3), np.float32)
        _check(_lib.lmb_frame_ring_read_average(self._h, _fptr(out)))
        return out

    def output_device_ptr(self) -> tuple[int, int]:
        p, n = _vp(), ctypes.c_uint64()
        _check(_lib.lmb_frame_ring_output(self._h, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def average_on_device(self, stream=None) -> None:
        """Enqueues the average into the device buffer of output_device_ptr()
        without reading it back."""
        _check(_lib.lmb_frame_ring_average(self._h, stream))

    def close(self):
        if self._h:
            _lib.lmb_frame_ring_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DataBuffers:
    """data_buffer/mod.rs:103-440 DataBuffers (plus the output images)."""

    def __init__(self, width: int, height: int, world_space_buffer_size: int = 1,
                 buffers: BufferType = BufferType.all()):
        h = _vp()
        _check(_lib.lmb_data_buffers_create(width, height, world_space_buffer_size, int(buffers),
                                            ctypes.byref(h)))
        self._h = h
        self.width, self.height = max(width, 1), max(height, 1)

    @property
    def handle(self):
        return self._h

    def resize(self, width: int, height: int) -> None:
        _check(_lib.lmb_data_buffers_resize(self._h, width, height))
        self.width, self.height = max(width, 1), max(height, 1)

    def advance_frame(self, buffers_to_temporally_advance: BufferType,
                      options: AdvanceOptions = AdvanceOptions.SEMI_INTERACTIVE,
                      stream=None) -> None:
        _check(_lib.lmb_advance_frame(self._h, int(buffers_to_temporally_advance),
                                      int(options.confidence_cap), stream))

    def temporal_resample(self, seed_offset: int = 0, stream=None) -> None:
        _check(_lib.lmb_temporal_resample(self._h, int(seed_offset) & 0xFFFFFFFF, stream))

    def spatial_resample(self, is_samples: int = 8, is_space: int = 61, seed_offset: int = 0,
                         stream=None) -> None:
        _check(_lib.lmb_spatial_resample(self._h, is_samples, is_space,
                                         int(seed_offset) & 0xFFFFFFFF, stream))

    def output(self, which: Output = Output.COLOR) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), np.float32)
        _check(_lib.lmb_data_buffers_read_output(self._h, int(which), _fptr(out)))
        return out

    def output_device_ptr(self, which: Output = Output.COLOR) -> tuple[int, int]:
        p, n = _vp(), ctypes.c_uint64()
        _check(_lib.lmb_data_buffers_output(self._h, int(which), ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def state_bytes(self, which: StateBuffer) -> bytes:
        n = ctypes.c_uint64()
        _check(_lib.lmb_data_buffers_state_size(self._h, int(which), ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(_lib.lmb_data_buffers_read_state(self._h, int(which), buf, n.value))
        return buf.raw

    def write_state(self, which: StateBuffer, data: bytes) -> None:
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        _check(_lib.lmb_data_buffers_write_state(self._h, int(which), buf, len(data)))

    def set_profiling(self, enable: bool = True) -> None:
        _check(_lib.lmb_data_buffers_set_profiling(self._h, 1 if enable else 0))

    def kernel_stats(self) -> dict:
        """Summed extend / shade launch durations (HIP events) since the last
        call, and the path vertices issued meanwhile."""
        st = _KernelStatsC()
        _check(_lib.lmb_data_buffers_kernel_stats(self._h, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in _KernelStatsC._fields_}

    def path_vertices(self) -> int:
        n = ctypes.c_uint64()
        _check(_lib.lmb_data_buffers_path_vertices(self._h, ctypes.byref(n)))
        return n.value

    def close(self):
        if self._h:
            _lib.lmb_data_buffers_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- neural cache
NRC_NUM_WEIGHTS = 37056
NRC_INPUTS = 13


@dataclass
class NeuralCacheOpts:
    """lmb_nrc_opts: Adam parameters, training paths and optimiser steps per
    cached frame (steps 1..256), weight-initialisation seed, EMA decay of the
    inference weights, the spread rule's bound (defer_specular) and whether
    the training runs on the cache's own stream beside the frame's samples
    (overlap_training), and whether the first scatter of a sample draws from
    the pixel's scrambled Sobol sequence (stratify)."""
    learning_rate: float = 1e-3
    beta1: float = 0.9
    beta2: float = 0.99
    epsilon: float = 1e-8
    train_paths: int = 65536
    train_steps: int = 4
    seed: int = 1
    ema_decay: float = 0.0
    defer_specular: int = 2
    overlap_training: bool = True
    stratify: bool = True


def _nrc_opts_c(o: NeuralCacheOpts) -> _NrcOptsC:
    c = _NrcOptsC()
    c.learning_rate, c.beta1, c.beta2, c.epsilon = (o.learning_rate, o.beta1, o.beta2, o.epsilon)
    c.train_paths, c.train_steps, c.seed = o.train_paths, o.train_steps, o.seed
    c.ema_decay = o.ema_decay
    c.defer_specular = int(o.defer_specular)
    c.overlap_training = 1 if o.overlap_training else 0
    c.stratify = 1 if o.stratify else 0
    return c


class NeuralCache:
    """The neural radiance cache (MLP on the matrix cores, trained online)."""

    def __init__(self, opts: Optional[NeuralCacheOpts] = None):
        o = opts or NeuralCacheOpts()
        c = _nrc_opts_c(o)
        h = _vp()
        _check(_lib.lmb_nrc_create(ctypes.byref(c), ctypes.byref(h)))
        self._h = h
        self.opts = o

    def set_opts(self, opts: NeuralCacheOpts) -> None:
        """Learning-rate schedules, freezing (train_paths=0): weights and the
        optimiser state are kept."""
        c = _nrc_opts_c(opts)
        _check(_lib.lmb_nrc_set_opts(self._h, ctypes.byref(c)))
        self.opts = opts

    @property
    def handle(self):
        return self._h

    def weights(self) -> np.ndarray:
        w = np.empty(NRC_NUM_WEIGHTS, np.float32)
        _check(_lib.lmb_nrc_get_weights(self._h, _fptr(w)))
        return w

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        if w.size != NRC_NUM_WEIGHTS:
            raise ValueError(f"expected {NRC_NUM_WEIGHTS} weights")
        _check(_lib.lmb_nrc_set_weights(self._h, _fptr(w)))

    def infer(self, inputs: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, NRC_INPUTS)
        out = np.empty((len(x), 3), np.float32)
        _check(_lib.lmb_nrc_infer(self._h, _fptr(x), len(x), _fptr(out)))
        return out

    def gradient(self, inputs: np.ndarray, targets: np.ndarray):
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, NRC_INPUTS)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        g = np.empty(NRC_NUM_WEIGHTS, np.float32)
        loss = ctypes.c_float()
        _check(_lib.lmb_nrc_gradient(self._h, _fptr(x), _fptr(t), len(x), _fptr(g),
                                     ctypes.byref(loss)))
        return g, loss.value

    def train(self, inputs: np.ndarray, targets: np.ndarray) -> float:
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, NRC_INPUTS)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        loss = ctypes.c_float()
        _check(_lib.lmb_nrc_train(self._h, _fptr(x), _fptr(t), len(x), ctypes.byref(loss)))
        return loss.value

    def apply_gradient(self, grad: np.ndarray) -> None:
        g = np.ascontiguousarray(grad, np.float32)
        _check(_lib.lmb_nrc_apply_gradient(self._h, _fptr(g)))

    def stats(self) -> dict:
        st = _NrcStatsC()
        _check(_lib.lmb_nrc_stats_get(self._h, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in _NrcStatsC._fields_ if n != "reserved"}

    def close(self):
        if self._h:
            _lib.lmb_nrc_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- batched tracer
class _BatchItemC(ctypes.Structure):
    _fields_ = [("scene", _vp), ("camera", _CameraC), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32)]


@dataclass
class BatchConfig:
    """lmb_batch_config: launch shape of the batched wavefronts."""
    blocks_per_cu: int = 8
    refill: int = 16
    tile_w: int = 8
    tile_h: int = 8
    pass_pixels: int = 0
    use_graph: bool = True

    def _c(self) -> _BatchConfigC:
        c = _BatchConfigC()
        c.blocks_per_cu, c.refill = self.blocks_per_cu, self.refill
        c.tile_w, c.tile_h, c.pass_pixels = self.tile_w, self.tile_h, self.pass_pixels
        c.use_graph = 1 if self.use_graph else 0
        return c

    @staticmethod
    def _from(c: _BatchConfigC) -> "BatchConfig":
        return BatchConfig(c.blocks_per_cu, c.refill, c.tile_w, c.tile_h, c.pass_pixels,
                           bool(c.use_graph))

    @staticmethod
    def default_sweep() -> list:
        n = _lib.lmb_batch_default_sweep(None, 0)
        arr = (_BatchConfigC * n)()
        _lib.lmb_batch_default_sweep(arr, n)
        return [BatchConfig._from(c) for c in arr]


class Batch:
    """Many (scene, camera, resolution) images rendered by one set of
    persistent wavefront launches (no reference counterpart)."""

    def __init__(self, items, opts: PathTracerOpts, quality: Quality = Quality.MEDIUM,
                 config: Optional[BatchConfig] = None):
        self._items = list(items)  # keeps the scenes alive
        arr = (_BatchItemC * len(self._items))()
        for k, (scene, cam, w, h) in enumerate(self._items):
            arr[k].scene = scene.handle
            arr[k].camera = cam._c
            arr[k].width, arr[k].height = int(w), int(h)
        oc = _tracer_opts_c(quality, opts)
        cc = (config or BatchConfig())._c()
        h = _vp()
        _check(_lib.lmb_batch_create(arr, len(self._items), ctypes.byref(oc), ctypes.byref(cc),
                                     ctypes.byref(h)))
        self._h = h

    def __len__(self):
        return len(self._items)

    @property
    def config(self) -> BatchConfig:
        c = _BatchConfigC()
        _check(_lib.lmb_batch_get_config(self._h, ctypes.byref(c)))
        return BatchConfig._from(c)

    def set_config(self, config: BatchConfig) -> None:
        c = config._c()
        _check(_lib.lmb_batch_set_config(self._h, ctypes.byref(c)))

    def set_camera(self, item: int, camera: Camera) -> None:
        _check(_lib.lmb_batch_set_camera(self._h, item, ctypes.byref(camera._c)))

    def trace(self, seed_offset: int = 0, stream=None) -> None:
        _check(_lib.lmb_batch_trace(self._h, int(seed_offset) & 0xFFFFFFFF, stream))

    def output(self, item: int) -> np.ndarray:
        _, _, w, h = self._items[item]
        out = np.empty((h, w, 4), np.float32)
        _check(_lib.lmb_batch_read_output(self._h, item, _fptr(out)))
        return out

    def path_vertices(self) -> int:
        n = ctypes.c_uint64()
        _check(_lib.lmb_batch_path_vertices(self._h, ctypes.byref(n)))
        return n.value

    def autotune(self, candidates=None, frames: int = 2):
        """(best config, ms per frame of every candidate)."""
        cands = list(candidates) if candidates is not None else BatchConfig.default_sweep()
        arr = (_BatchConfigC * len(cands))(*[c._c() for c in cands])
        ms = (ctypes.c_double * len(cands))()
        best = _BatchConfigC()
        _check(_lib.lmb_batch_autotune(self._h, arr, len(cands), frames, ctypes.byref(best), ms))
        return BatchConfig._from(best), list(ms)

    def close(self):
        if self._h:
            _lib.lmb_batch_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
