"""ctypes binding of librefrakt_amd.so and a Python mirror of the reference's
flame API (src/flame.hpp: flame_xform, flame, load_flame, warmup,
draw_to_bins; src/main.cpp: the DE and tonemap passes).

The library is built in-tree by ``refrakt_amd.build`` (or
``__graft_entry__.build()``); importing this module never builds it and fails
loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librefrakt_amd.so")

RF_OK = 0
RF_ENGINE_COMPAT = 0
RF_ENGINE_FAST = 1
RF_HIST_AUTO, RF_HIST_QUAD, RF_HIST_WAVEAGG, RF_HIST_NAIVE, RF_HIST_LDS, RF_HIST_SORT = 0, 1, 2, 3, 4, 5
MAX_XFORMS = 32
MAX_VARS = 16
MAX_PARAMS = 8


class RefraktError(RuntimeError):
    pass


class VarTerm(C.Structure):
    _fields_ = [("id", C.c_int32), ("weight", C.c_float), ("params", C.c_float * MAX_PARAMS)]


class XformDesc(C.Structure):
    _fields_ = [("affine", C.c_float * 6), ("post", C.c_float * 6), ("has_post", C.c_int32),
                ("n_vars", C.c_int32), ("vars", VarTerm * MAX_VARS), ("weight", C.c_float),
                ("color", C.c_float), ("color_speed", C.c_float), ("rotation_frequency", C.c_float),
                ("opacity", C.c_float)]


class FlameDesc(C.Structure):
    _fields_ = [("n_xforms", C.c_int32), ("xforms", XformDesc * MAX_XFORMS), ("has_final", C.c_int32),
                ("final_xform", XformDesc), ("palette", (C.c_float * 4) * 256), ("size", C.c_uint32 * 2),
                ("center", C.c_float * 2), ("scale", C.c_float), ("rotate", C.c_float),
                ("estimator_min", C.c_int32), ("estimator_radius", C.c_int32),
                ("estimator_curve", C.c_float), ("gamma", C.c_float), ("vibrancy", C.c_float),
                ("brightness", C.c_float)]


class SimParams(C.Structure):
    _fields_ = [("total_particles", C.c_uint64), ("temporal_samples", C.c_uint32),
                ("shuffle_count", C.c_uint32), ("seed", C.c_uint64), ("device", C.c_int32),
                ("engine", C.c_int32), ("hist_mode", C.c_int32), ("block_waves", C.c_int32)]


class DEParams(C.Structure):
    _fields_ = [("estimator_radius", C.c_int32), ("estimator_min", C.c_int32), ("estimator_curve", C.c_float)]


class TonemapParams(C.Structure):
    _fields_ = [("gamma", C.c_float), ("brightness", C.c_float), ("vibrancy", C.c_float),
                ("scale_constant", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("warmup_passes", C.c_uint32),
                ("draw_passes", C.c_uint32), ("draw_calls", C.c_uint32), ("tss_width", C.c_float),
                ("density_estimation", C.c_int32), ("de", DEParams), ("tm", TonemapParams)]


class Stats(C.Structure):
    _fields_ = [("drawn", C.c_uint64), ("iterations", C.c_uint64), ("ms_warmup", C.c_double),
                ("ms_draw", C.c_double), ("ms_density", C.c_double), ("ms_tonemap", C.c_double)]


class FrameDescC(C.Structure):
    _fields_ = [("flame", C.POINTER(FlameDesc)), ("width", C.c_uint32), ("height", C.c_uint32),
                ("samples", C.c_uint64)]


class BatchConfig(C.Structure):
    _fields_ = [("block_waves", C.c_int32), ("hist_mode", C.c_int32), ("iters", C.c_uint32), ("fuse", C.c_uint32)]


_STRUCTS = [VarTerm, XformDesc, FlameDesc, SimParams, DEParams, TonemapParams, RenderParams, Stats, FrameDescC,
            BatchConfig]

_P = C.c_void_p
_SIGS = {
    "rf_create": (C.c_int, [C.POINTER(SimParams), C.POINTER(_P)]),
    "rf_destroy": (None, [_P]),
    "rf_set_sim_parameters": (C.c_int, [_P, C.c_uint64, C.c_uint32, C.c_uint32]),
    "rf_set_engine": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "rf_last_error": (C.c_char_p, [_P]),
    "rf_last_draw_ms": (C.c_double, [_P]),
    "rf_stream": (_P, [_P]),
    "rf_flame_load": (C.c_int, [_P, C.c_char_p, C.POINTER(_P)]),
    "rf_flame_parse": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(_P)]),
    "rf_flame_parse_desc": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(FlameDesc), C.c_char_p, C.c_size_t]),
    "rf_flame_create": (C.c_int, [_P, C.POINTER(FlameDesc), C.POINTER(_P)]),
    "rf_flame_update": (C.c_int, [_P, C.POINTER(FlameDesc)]),
    "rf_flame_get_desc": (C.c_int, [_P, C.POINTER(FlameDesc)]),
    "rf_flame_destroy": (None, [_P]),
    "rf_flame_param_block": (C.c_int, [C.POINTER(FlameDesc), C.POINTER(C.c_float), C.c_int]),
    "rf_flame_validate": (C.c_int, [C.POINTER(FlameDesc), C.c_char_p, C.c_size_t]),
    "rf_fast_perms_download": (C.c_int, [_P, C.POINTER(C.c_uint16)]),
    "rf_hist_create": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    "rf_hist_wrap": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    "rf_hist_destroy": (None, [_P]),
    "rf_hist_zero": (C.c_int, [_P]),
    "rf_hist_download": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "rf_hist_upload": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "rf_hist_device_ptr": (_P, [_P]),
    "rf_warmup": (C.c_int, [_P, _P, C.c_uint32, C.c_float]),
    "rf_draw_to_bins": (C.c_int, [_P, _P, _P, C.c_uint32, C.POINTER(C.c_uint64)]),
    "rf_warmup_rows": (C.c_int, [_P, _P, C.c_uint32, C.c_float, C.POINTER(C.c_uint32)]),
    "rf_draw_to_bins_rows": (C.c_int, [_P, _P, _P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "rf_draw_to_bins_noalloc": (C.c_int, [_P, _P, _P, C.c_uint32, C.POINTER(C.c_uint64)]),
    "rf_density_estimate": (C.c_int, [_P, _P, C.POINTER(DEParams), _P]),
    "rf_tonemap": (C.c_int, [_P, _P, _P, C.c_uint32, C.c_uint32, C.POINTER(TonemapParams)]),
    "rf_density_tonemap": (C.c_int, [_P, _P, C.POINTER(DEParams), C.POINTER(TonemapParams), _P]),
    "rf_render": (C.c_int, [_P, _P, C.POINTER(RenderParams), C.POINTER(C.c_float), C.POINTER(C.c_float),
                            C.POINTER(Stats)]),
    "rf_particles_download": (C.c_int, [_P, _P, C.POINTER(C.c_float)]),
    "rf_rng_download": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "rf_shuffle_download": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "rf_sample_points_download": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "rf_flame_hist_mode": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "rf_flame_sort_overflow": (C.c_double, [_P]),
    "rf_batch_create": (C.c_int, [_P, C.POINTER(FrameDescC), C.c_uint32, C.POINTER(_P)]),
    "rf_batch_destroy": (None, [_P]),
    "rf_batch_set_config": (C.c_int, [_P, C.POINTER(BatchConfig)]),
    "rf_batch_get_config": (C.c_int, [_P, C.POINTER(BatchConfig)]),
    "rf_batch_tune": (C.c_int, [_P, C.c_double, C.POINTER(BatchConfig), C.POINTER(C.c_double)]),
    "rf_batch_run": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "rf_batch_frame_stats": (C.c_int, [_P, _P, _P, _P]),
    "rf_batch_hist_ptr": (_P, [_P, C.c_uint32]),
    "rf_batch_image_ptr": (_P, [_P, C.c_uint32]),
    "rf_batch_download": (C.c_int, [_P, C.c_uint32, C.c_int32, _P]),
    "rf_denoiser_create": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "rf_denoiser_destroy": (None, [_P]),
    "rf_denoiser_num_params": (C.c_size_t, [_P]),
    "rf_denoiser_get_params": (C.c_int, [_P, _P, C.c_size_t]),
    "rf_denoiser_set_params": (C.c_int, [_P, _P, C.c_size_t]),
    "rf_denoiser_train_step": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                         C.POINTER(C.c_float)]),
    "rf_denoiser_infer": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "rf_denoiser_grads": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _P,
                                    C.POINTER(C.c_float)]),
    "rf_log_density_input": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_uint32, C.c_double]),
    "rf_nn_conv3x3": (C.c_int, [_P, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                C.c_int32]),
    "rf_nn_conv3x3_wgrad": (C.c_int, [_P, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_int32,
                                      C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_int32]),
    "rf_nn_flip_weights": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]),
    "rf_nn_adam": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_float,
                             C.c_float, C.c_float, C.c_int32]),
    "rf_rgba8_from_f32": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P]),
    "rf_rgba8_device": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "rf_write_png": (C.c_int, [C.c_char_p, _P, C.c_uint32, C.c_uint32]),
    "rf_write_exr": (C.c_int, [C.c_char_p, _P, C.c_uint32, C.c_uint32, C.c_int32]),
    "rf_half_from_float": (C.c_uint16, [C.c_float]),
    "rf_screen_affine": (C.c_int, [C.POINTER(FlameDesc), C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]),
    "rf_set_flame_programs": (C.c_int, [_P, C.c_int32]),
    "rf_batch_program_info": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "rf_flame_program_info": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "rf_flame_program_source": (C.c_int, [C.POINTER(FlameDesc), C.c_int32, C.c_int32, C.c_char_p, C.c_size_t]),
    "rf_flame_program_build": (C.c_int, [C.POINTER(FlameDesc), C.c_int32, C.c_int32, C.c_char_p, C.c_void_p,
                                         C.c_size_t, C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t,
                                         C.POINTER(C.c_double)]),
    "rf_variation_count": (C.c_int, []),
    "rf_variation_name": (C.c_char_p, [C.c_int]),
    "rf_variation_id": (C.c_int, [C.c_char_p]),
    "rf_variation_num_params": (C.c_int, [C.c_int]),
    "rf_variation_param_name": (C.c_char_p, [C.c_int, C.c_int]),
    "rf_variation_flags": (C.c_int, [C.c_int]),
    "rf_eval_xform": (C.c_int, [_P, C.POINTER(XformDesc), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                C.POINTER(C.c_float), C.c_uint32]),
    "rf_atomic_probe": (C.c_int, [_P, C.c_int32, C.c_uint64, C.c_uint64, C.POINTER(C.c_double)]),
    "rf_struct_size": (C.c_size_t, [C.c_int]),
}

_lib = None


def lib():
    """Loads librefrakt_amd.so (built in-tree); raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RefraktError(f"{LIB_PATH} is missing: run refrakt_amd.build (or __graft_entry__.build())")
        # torch ships its own HIP runtime under the file name libamdhip64.so
        # (soname libamdhip64.so.7).  Loading torch first makes this library's
        # libamdhip64.so.7 dependency resolve to that same runtime; the other
        # order puts two HIP runtimes in one process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        for i, st in enumerate(_STRUCTS):
            if L.rf_struct_size(i) != C.sizeof(st):
                raise RefraktError(f"ABI mismatch for {st.__name__}: C {L.rf_struct_size(i)} vs ctypes {C.sizeof(st)}")
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


# ------------------------------------------------------------ registry
def variation_names() -> List[str]:
    L = lib()
    return [L.rf_variation_name(i).decode() for i in range(L.rf_variation_count())]


def variation_id(name: str) -> int:
    return lib().rf_variation_id(name.encode())


def variation_params(name: str) -> List[str]:
    L = lib()
    i = variation_id(name)
    if i < 0:
        raise KeyError(name)
    return [L.rf_variation_param_name(i, k).decode() for k in range(L.rf_variation_num_params(i))]


# ---------------------------------------------------------- flame model
@dataclass
class Xform:
    """flame_xform (src/flame.hpp:21-37)."""
    affine: List[float] = field(default_factory=lambda: [1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    post: Optional[List[float]] = None
    variations: Dict[str, float] = field(default_factory=dict)
    var_param: Dict[str, float] = field(default_factory=dict)
    weight: float = 0.0
    color: float = 0.0
    color_speed: float = 0.0
    rotation_frequency: float = 0.0
    opacity: float = 0.0

    def to_c(self) -> XformDesc:
        x = XformDesc()
        x.affine[:] = [float(v) for v in self.affine]
        if self.post is not None:
            x.post[:] = [float(v) for v in self.post]
            x.has_post = 1
        if len(self.variations) > MAX_VARS:
            raise RefraktError("too many variations in one xform")
        x.n_vars = len(self.variations)
        for t, (name, w) in enumerate(self.variations.items()):
            vid = variation_id(name)
            if vid < 0:
                raise RefraktError(f"unknown variation {name}")
            x.vars[t].id = vid
            x.vars[t].weight = float(w)
            for k, pn in enumerate(variation_params(name)):
                x.vars[t].params[k] = float(self.var_param.get(pn, 0.0))
        x.weight, x.color, x.color_speed = float(self.weight), float(self.color), float(self.color_speed)
        x.rotation_frequency, x.opacity = float(self.rotation_frequency), float(self.opacity)
        return x

    @staticmethod
    def from_c(x: XformDesc) -> "Xform":
        names = variation_names()
        out = Xform(affine=list(x.affine), post=list(x.post) if x.has_post else None,
                    weight=x.weight, color=x.color, color_speed=x.color_speed,
                    rotation_frequency=x.rotation_frequency, opacity=x.opacity)
        for t in range(x.n_vars):
            name = names[x.vars[t].id]
            out.variations[name] = x.vars[t].weight
            for k, pn in enumerate(variation_params(name)):
                out.var_param[pn] = x.vars[t].params[k]
        return out


@dataclass
class Flame:
    """flame (src/flame.hpp:41-60): xforms, palette, view and DE/tonemap settings."""
    xforms: List[Xform] = field(default_factory=list)
    palette: np.ndarray = field(default_factory=lambda: np.zeros((256, 4), np.float32))
    size: List[int] = field(default_factory=lambda: [1024, 1024])
    center: List[float] = field(default_factory=lambda: [0.0, 0.0])
    scale: float = 256.0
    rotate: float = 0.0
    final_xform: Optional[Xform] = None
    estimator_min: int = 0
    estimator_radius: int = 9
    estimator_curve: float = 0.4
    gamma: float = 2.2
    vibrancy: float = 1.0
    brightness: float = 4.0

    def to_c(self) -> FlameDesc:
        d = FlameDesc()
        if not 1 <= len(self.xforms) <= MAX_XFORMS:
            raise RefraktError("a flame needs 1..32 xforms")
        d.n_xforms = len(self.xforms)
        for i, x in enumerate(self.xforms):
            d.xforms[i] = x.to_c()
        if self.final_xform is not None:
            d.has_final = 1
            d.final_xform = self.final_xform.to_c()
        pal = np.ascontiguousarray(self.palette, dtype=np.float32).reshape(256, 4)
        C.memmove(d.palette, pal.ctypes.data, pal.nbytes)
        d.size[:] = [int(v) for v in self.size]
        d.center[:] = [float(v) for v in self.center]
        d.scale, d.rotate = float(self.scale), float(self.rotate)
        d.estimator_min, d.estimator_radius = int(self.estimator_min), int(self.estimator_radius)
        d.estimator_curve, d.gamma = float(self.estimator_curve), float(self.gamma)
        d.vibrancy, d.brightness = float(self.vibrancy), float(self.brightness)
        return d

    @staticmethod
    def from_c(d: FlameDesc) -> "Flame":
        pal = np.ctypeslib.as_array(d.palette).reshape(256, 4).astype(np.float32).copy()
        return Flame(xforms=[Xform.from_c(d.xforms[i]) for i in range(d.n_xforms)], palette=pal,
                     size=list(d.size), center=list(d.center), scale=d.scale, rotate=d.rotate,
                     final_xform=Xform.from_c(d.final_xform) if d.has_final else None,
                     estimator_min=d.estimator_min, estimator_radius=d.estimator_radius,
                     estimator_curve=d.estimator_curve, gamma=d.gamma, vibrancy=d.vibrancy,
                     brightness=d.brightness)


def to_rgba8(img: np.ndarray) -> np.ndarray:
    """(H, W, 4) float image -> (H, W, 4) uint8, the screenshot conversion."""
    img = np.ascontiguousarray(img, np.float32)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError("expected an (H, W, 4) image")
    out = np.empty(img.shape, np.uint8)
    rc = lib().rf_rgba8_from_f32(img.ctypes.data, img.shape[1], img.shape[0], out.ctypes.data)
    if rc != RF_OK:
        raise RefraktError(f"rf_rgba8_from_f32 failed ({rc})")
    return out


def write_png(path: str, rgba8: np.ndarray) -> None:
    """Writes an (H, W, 4) uint8 image as a PNG (first row at the top)."""
    rgba8 = np.ascontiguousarray(rgba8, np.uint8)
    if rgba8.ndim != 3 or rgba8.shape[2] != 4:
        raise ValueError("expected an (H, W, 4) uint8 image")
    rc = lib().rf_write_png(os.fsencode(path), rgba8.ctypes.data, rgba8.shape[1], rgba8.shape[0])
    if rc != RF_OK:
        raise RefraktError(f"rf_write_png failed ({rc}) for {path}")


def write_exr(path: str, rgba: np.ndarray, half: bool = True) -> None:
    """Writes an (H, W, 4) float RGBA image as an uncompressed OpenEXR file
    (first row at the top; channels HALF, or FLOAT when half is False)."""
    rgba = np.ascontiguousarray(rgba, np.float32)
    if rgba.ndim != 3 or rgba.shape[2] != 4:
        raise ValueError("expected an (H, W, 4) float image")
    rc = lib().rf_write_exr(os.fsencode(path), rgba.ctypes.data, rgba.shape[1], rgba.shape[0], 1 if half else 0)
    if rc != RF_OK:
        raise RefraktError(f"rf_write_exr failed ({rc}) for {path}")


def param_block(flame: Flame) -> np.ndarray:
    """The reference's flat parameter block (src/flame.cpp:33-103)."""
    d = flame.to_c()
    L = lib()
    n = L.rf_flame_param_block(C.byref(d), None, 0)
    out = np.zeros(n, np.float32)
    L.rf_flame_param_block(C.byref(d), _fptr(out), n)
    return out


def screen_affine(flame: Flame, width: int, height: int) -> np.ndarray:
    """draw_to_bins' screen-space affine (src/flame.cpp:289-296)."""
    out = np.zeros(6, np.float32)
    d = flame.to_c()
    if lib().rf_screen_affine(C.byref(d), width, height, _fptr(out)) != RF_OK:
        raise RefraktError("rf_screen_affine failed")
    return out


def program_source(flame: "Flame", block_waves: int = 8, hist_mode: int = 5) -> str:
    """HIP source of the flame program of `flame` (one entry point)."""
    L = lib()
    d = flame.to_c()
    n = L.rf_flame_program_source(C.byref(d), block_waves, hist_mode, None, 0)
    if n < 0:
        raise RefraktError(f"rf_flame_program_source failed ({n})")
    buf = C.create_string_buffer(n)
    L.rf_flame_program_source(C.byref(d), block_waves, hist_mode, buf, n)
    return buf.value.decode()


def program_build(flame: "Flame", block_waves: int = 8, hist_mode: int = 5, arch: str = "gfx950"):
    """Compiles the flame program for `arch` without a device; returns (code
    object bytes, compiler log, seconds)."""
    L = lib()
    d = flame.to_c()
    size = C.c_size_t(0)
    log = C.create_string_buffer(1 << 16)
    sec = C.c_double(0.0)
    cap = 16 << 20
    code = C.create_string_buffer(cap)
    rc = L.rf_flame_program_build(C.byref(d), block_waves, hist_mode, arch.encode(), code, cap, C.byref(size), log,
                                  len(log), C.byref(sec))
    if rc != RF_OK:
        raise RefraktError(f"rf_flame_program_build failed ({rc}): {log.value.decode(errors='replace')}")
    return code.raw[:size.value], log.value.decode(errors="replace"), sec.value


def validate(flame: Flame) -> str:
    """Empty string if rf_flame_create would accept the flame, else the reason."""
    d = flame.to_c()
    err = C.create_string_buffer(256)
    rc = lib().rf_flame_validate(C.byref(d), err, 256)
    return "" if rc == RF_OK else err.value.decode()


def parse_flam3(xml: str) -> Flame:
    """flame::load_flame's reader (src/flame.cpp:160-226) on XML text."""
    d = FlameDesc()
    raw = xml.encode()
    err = C.create_string_buffer(512)
    rc = lib().rf_flame_parse_desc(raw, len(raw), C.byref(d), err, 512)
    if rc != RF_OK:
        raise RefraktError(f"flam3 parse failed ({rc}): {err.value.decode()}")
    return Flame.from_c(d)


def load_flame(path: str) -> Flame:
    with open(path, "r", encoding="utf-8", errors="replace") as f:
        return parse_flam3(f.read())


# -------------------------------------------------------------- runtime
class Context:
    """Simulation state of flame::set_sim_parameters plus one HIP stream."""

    def __init__(self, total_particles: int, temporal_samples: int = 512, shuffle_count: int = 1024,
                 seed: int = 1, device: int = 0, engine: int = RF_ENGINE_FAST, hist_mode: int = RF_HIST_AUTO,
                 block_waves: int = 0, flame_programs: Optional[bool] = None):
        self.L = lib()
        sp = SimParams(total_particles, temporal_samples, shuffle_count, seed, device, engine, hist_mode, block_waves)
        h = C.c_void_p()
        rc = self.L.rf_create(C.byref(sp), C.byref(h))
        if rc != RF_OK:
            raise RefraktError(f"rf_create failed ({rc}): {self.L.rf_last_error(None).decode()}")
        self.h = h
        if flame_programs is not None:
            self.set_flame_programs(flame_programs)
        self.total_particles = total_particles
        self.temporal_samples = temporal_samples
        self.shuffle_count = shuffle_count

    def check(self, rc):
        if rc != RF_OK:
            raise RefraktError(f"refrakt_amd error {rc}: {self.L.rf_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.rf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sim_parameters(self, total_particles, temporal_samples, shuffle_count):
        self.check(self.L.rf_set_sim_parameters(self.h, total_particles, temporal_samples, shuffle_count))
        self.total_particles, self.temporal_samples, self.shuffle_count = total_particles, temporal_samples, shuffle_count

    def last_draw_ms(self) -> float:
        """Device time of the last draw_to_bins' chaos-game launches (HIP events)."""
        return self.L.rf_last_draw_ms(self.h)

    @property
    def stream_ptr(self) -> int:
        return self.L.rf_stream(self.h) or 0

    def set_flame_programs(self, enable: bool):
        """Fast engine: run per-structure compiled flame programs (True, the
        default) or the precompiled interpreter (False)."""
        self.check(self.L.rf_set_flame_programs(self.h, 1 if enable else 0))

    def set_engine(self, engine, hist_mode=RF_HIST_QUAD):
        self.check(self.L.rf_set_engine(self.h, engine, hist_mode))

    def flame(self, flame: Flame) -> "FlameHandle":
        return FlameHandle(self, flame)

    def histogram(self, width: int, height: int, device_ptr: Optional[int] = None) -> "Histogram":
        """A W x H RGBA float histogram; with device_ptr, wraps caller-owned device memory."""
        return Histogram(self, width, height, device_ptr)

    def shuffle_rows(self) -> np.ndarray:
        ppt = self.total_particles // self.temporal_samples
        out = np.zeros((self.shuffle_count, ppt), np.uint32)
        self.check(self.L.rf_shuffle_download(self.h, _uptr(out)))
        return out

    def rng_states(self) -> np.ndarray:
        out = np.zeros((self.total_particles, 4), np.uint32)
        self.check(self.L.rf_rng_download(self.h, _uptr(out)))
        return out

    def fast_perms(self, block_waves: int = 4) -> np.ndarray:
        out = np.zeros((8, 64 * block_waves), np.uint16)
        self.check(self.L.rf_fast_perms_download(self.h, out.ctypes.data_as(C.POINTER(C.c_uint16))))
        return out

    def sample_points(self) -> np.ndarray:
        ppt = self.total_particles // self.temporal_samples
        out = np.zeros((ppt, 4), np.float32)
        self.check(self.L.rf_sample_points_download(self.h, _fptr(out)))
        return out

    def eval_xform(self, xf: Xform, pts: np.ndarray, rands: np.ndarray) -> np.ndarray:
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        n = pts.shape[0]
        rands = np.ascontiguousarray(rands, np.float32).reshape(n, 16)
        out = np.zeros((n, 4), np.float32)
        xd = xf.to_c()
        self.check(self.L.rf_eval_xform(self.h, C.byref(xd), _fptr(pts), _fptr(rands), _fptr(out), n))
        return out

    def atomic_probe(self, hist_mode: int, footprint_bins: int, samples: int) -> float:
        ms = C.c_double()
        self.check(self.L.rf_atomic_probe(self.h, hist_mode, footprint_bins, samples, C.byref(ms)))
        return ms.value

    def density_tonemap(self, hist: "Histogram", de: Optional[DEParams], tm: TonemapParams, out_ptr: int):
        self.check(self.L.rf_density_tonemap(self.h, hist.h, C.byref(de) if de is not None else None,
                                             C.byref(tm), C.c_void_p(out_ptr)))

    def density_estimate(self, hist: "Histogram", de: DEParams, out_ptr: int):
        self.check(self.L.rf_density_estimate(self.h, hist.h, C.byref(de), C.c_void_p(out_ptr)))

    def tonemap(self, in_ptr: int, out_ptr: int, width: int, height: int, tm: TonemapParams):
        self.check(self.L.rf_tonemap(self.h, C.c_void_p(in_ptr), C.c_void_p(out_ptr), width, height, C.byref(tm)))

    def rgba8(self, in_ptr: int, out_ptr: int, width: int, height: int):
        """RGBA float image -> RGBA8 on the GPU (device pointers)."""
        self.check(self.L.rf_rgba8_device(self.h, C.c_void_p(in_ptr), C.c_void_p(out_ptr), width, height))

    def denoiser(self, seed: int = 1) -> "Denoiser":
        return Denoiser(self, seed)

    def log_density_input(self, hist_ptr: int, out_ptr: int, npix: int, scale: float):
        """bf16 16-channel denoiser input from a device RGBA histogram."""
        self.check(self.L.rf_log_density_input(self.h, C.c_void_p(hist_ptr), C.c_void_p(out_ptr), npix, scale))

    def batch(self, frames) -> "Batch":
        """Batched rendering of (flame, width, height, samples) frames."""
        return Batch(self, frames)

    def render(self, fh: "FlameHandle", params: RenderParams, want_hist=True, want_image=True):
        W, H = params.width, params.height
        hist = np.zeros((H, W, 4), np.float32) if want_hist else None
        img = np.zeros((H, W, 4), np.float32) if want_image else None
        st = Stats()
        self.check(self.L.rf_render(self.h, fh.h, C.byref(params), _fptr(hist) if want_hist else None,
                                    _fptr(img) if want_image else None, C.byref(st)))
        return hist, img, st


class FlameHandle:
    """A flame bound to a context: warmup / draw_to_bins (src/flame.cpp:228-330)."""

    def __init__(self, ctx: Context, flame: Flame):
        self.ctx = ctx
        self.flame = flame
        h = C.c_void_p()
        self._desc = flame.to_c()
        ctx.check(ctx.L.rf_flame_create(ctx.h, C.byref(self._desc), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.rf_flame_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, flame: Flame):
        self.flame = flame
        self._desc = flame.to_c()
        self.ctx.check(self.ctx.L.rf_flame_update(self.h, C.byref(self._desc)))

    def warmup(self, num_passes: int, tss_width: float = 1.2 / 60.0, rows: Optional[np.ndarray] = None):
        if rows is None:
            self.ctx.check(self.ctx.L.rf_warmup(self.ctx.h, self.h, num_passes, tss_width))
        else:
            r = np.ascontiguousarray(rows, np.uint32).reshape(-1)
            assert r.size == 2 * (num_passes + 1)
            self.ctx.check(self.ctx.L.rf_warmup_rows(self.ctx.h, self.h, num_passes, tss_width, _uptr(r)))

    def draw_to_bins(self, hist: "Histogram", num_iter: int, rows: Optional[np.ndarray] = None) -> int:
        drawn = C.c_uint64()
        if rows is None:
            self.ctx.check(self.ctx.L.rf_draw_to_bins(self.ctx.h, self.h, hist.h, num_iter, C.byref(drawn)))
        else:
            r = np.ascontiguousarray(rows, np.uint32).reshape(-1)
            assert r.size == 2 * num_iter
            self.ctx.check(self.ctx.L.rf_draw_to_bins_rows(self.ctx.h, self.h, hist.h, num_iter, _uptr(r),
                                                           C.byref(drawn)))
        return drawn.value

    def draw_to_bins_noalloc(self, hist: "Histogram", num_iter: int) -> int:
        """rf_draw_to_bins_noalloc: the draw without any device allocation
        (fast engine, direct histogram modes)."""
        drawn = C.c_uint64()
        self.ctx.check(self.ctx.L.rf_draw_to_bins_noalloc(self.ctx.h, self.h, hist.h, num_iter, C.byref(drawn)))
        return drawn.value

    def particles(self) -> np.ndarray:
        out = np.zeros((self.ctx.total_particles, 4), np.float32)
        self.ctx.check(self.ctx.L.rf_particles_download(self.ctx.h, self.h, _fptr(out)))
        return out

    def hist_mode(self):
        """(strategy of the next draw, LDS hit fraction of the last LDS draw or -1)."""
        frac = C.c_double(-1.0)
        m = self.ctx.L.rf_flame_hist_mode(self.h, C.byref(frac))
        if m < 0:
            self.ctx.check(m)
        return m, frac.value

    def sort_overflow(self) -> float:
        """Overflow fraction of the last sort-mode draw (-1 if none)."""
        return float(self.ctx.L.rf_flame_sort_overflow(self.h))

    def program_info(self):
        """(1 if the last fast-engine launch ran a flame program, 0 if it was
        interpreted, -1 if none ran; seconds spent compiling its program)."""
        sec = C.c_double(0.0)
        used = self.ctx.L.rf_flame_program_info(self.h, C.byref(sec))
        return used, sec.value

    def param_block(self) -> np.ndarray:
        return param_block(self.flame)

    def screen_affine(self, width: int, height: int) -> np.ndarray:
        return screen_affine(self.flame, width, height)


class Histogram:
    """A W x H RGBA float32 bin array in device memory (flame::bin_t)."""

    def __init__(self, ctx: Context, width: int, height: int, device_ptr: Optional[int] = None):
        self.ctx = ctx
        self.width, self.height = width, height
        h = C.c_void_p()
        if device_ptr is None:
            ctx.check(ctx.L.rf_hist_create(ctx.h, width, height, C.byref(h)))
        else:
            ctx.check(ctx.L.rf_hist_wrap(ctx.h, C.c_void_p(device_ptr), width, height, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.rf_hist_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def zero(self):
        self.ctx.check(self.ctx.L.rf_hist_zero(self.h))

    def download(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.float32)
        self.ctx.check(self.ctx.L.rf_hist_download(self.h, _fptr(out)))
        return out

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float32).reshape(self.height, self.width, 4)
        self.ctx.check(self.ctx.L.rf_hist_upload(self.h, _fptr(a)))

    @property
    def device_ptr(self) -> int:
        return self.ctx.L.rf_hist_device_ptr(self.h)


def render(ctx: Context, flame: Flame, width: int, height: int, warmup_passes: int = 19, draw_passes: int = 64,
           draw_calls: int = 1, density_estimation: bool = True, scale_constant: float = 1e-4,
           tss_width: float = 1.2 / 60.0):
    """render(flame, params) -> (histogram, tonemapped image, stats)."""
    fh = ctx.flame(flame)
    rp = RenderParams(width, height, warmup_passes, draw_passes, draw_calls, tss_width, int(density_estimation),
                      DEParams(flame.estimator_radius, flame.estimator_min, flame.estimator_curve),
                      TonemapParams(flame.gamma, flame.brightness, flame.vibrancy, scale_constant))
    try:
        return ctx.render(fh, rp)
    finally:
        fh.close()


class Batch:
    """Many frames in one persistent chaos-game launch plus per-frame DE and
    tonemap, captured in a HIP graph (rf_batch_*)."""

    def __init__(self, ctx: Context, frames):
        self.ctx = ctx
        self.sizes = [(int(w), int(h)) for _, w, h, _ in frames]
        descs = [f.to_c() for f, _, _, _ in frames]
        arr = (FrameDescC * len(frames))()
        for i, (d, (_, w, h, n)) in enumerate(zip(descs, frames)):
            arr[i].flame = C.pointer(d)
            arr[i].width, arr[i].height, arr[i].samples = int(w), int(h), int(n)
        h = _P()
        ctx.check(ctx.L.rf_batch_create(ctx.h, arr, len(frames), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.rf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, block_waves=8, hist_mode=RF_HIST_LDS, iters=256, fuse=20):
        cfg = BatchConfig(block_waves, hist_mode, iters, fuse)
        self.ctx.check(self.ctx.L.rf_batch_set_config(self.h, C.byref(cfg)))

    def config(self) -> dict:
        cfg = BatchConfig()
        self.ctx.check(self.ctx.L.rf_batch_get_config(self.h, C.byref(cfg)))
        return {"block_waves": cfg.block_waves, "hist_mode": cfg.hist_mode, "iters": cfg.iters, "fuse": cfg.fuse}

    def tune(self, sample_fraction: float = 0.125):
        cfg, ms = BatchConfig(), C.c_double()
        self.ctx.check(self.ctx.L.rf_batch_tune(self.h, sample_fraction, C.byref(cfg), C.byref(ms)))
        return self.config(), ms.value

    def run(self) -> float:
        ms = C.c_double()
        self.ctx.check(self.ctx.L.rf_batch_run(self.h, C.byref(ms)))
        return ms.value

    def program_info(self):
        """(1 if the chaos-game launch runs a flame program, else 0; compile seconds)."""
        sec = C.c_double(0.0)
        used = self.ctx.L.rf_batch_program_info(self.h, C.byref(sec))
        return used, sec.value

    def frame_stats(self) -> dict:
        n = len(self.sizes)
        drawn = np.zeros(n, np.uint64)
        chaos = np.zeros(n, np.float64)
        image = np.zeros(n, np.float64)
        self.ctx.check(self.ctx.L.rf_batch_frame_stats(self.h, drawn.ctypes.data, chaos.ctypes.data,
                                                       image.ctypes.data))
        return {"drawn": drawn, "chaos_done_ms": chaos, "image_done_ms": image}

    def _download(self, i: int, which: int) -> np.ndarray:
        w, h = self.sizes[i]
        out = np.zeros((h, w, 4), np.float32)
        self.ctx.check(self.ctx.L.rf_batch_download(self.h, i, which, out.ctypes.data))
        return out

    def histogram(self, i: int) -> np.ndarray:
        return self._download(i, 0)

    def image(self, i: int) -> np.ndarray:
        return self._download(i, 1)

    def hist_ptr(self, i: int) -> int:
        return self.ctx.L.rf_batch_hist_ptr(self.h, i)

    def image_ptr(self, i: int) -> int:
        return self.ctx.L.rf_batch_image_ptr(self.h, i)


class Denoiser:
    """The U-Net denoiser (rf_denoiser_*).  Tensors are passed as device pointers."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        h = _P()
        ctx.check(ctx.L.rf_denoiser_create(ctx.h, seed, C.byref(h)))
        self.h = h
        self.n_params = int(ctx.L.rf_denoiser_num_params(h))

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.rf_denoiser_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        self.ctx.check(self.ctx.L.rf_denoiser_get_params(self.h, out.ctypes.data, self.n_params))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        self.ctx.check(self.ctx.L.rf_denoiser_set_params(self.h, p.ctypes.data, p.size))

    def train_step(self, x_ptr: int, t_ptr: int, N: int, H: int, W: int, lr: float, want_loss=True):
        loss = C.c_float(0.0)
        self.ctx.check(self.ctx.L.rf_denoiser_train_step(self.h, C.c_void_p(x_ptr), C.c_void_p(t_ptr), N, H, W,
                                                         lr, C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def infer(self, x_ptr: int, y_ptr: int, N: int, H: int, W: int):
        self.ctx.check(self.ctx.L.rf_denoiser_infer(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), N, H, W))

    def grads(self, x_ptr: int, t_ptr: int, N: int, H: int, W: int):
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float(0.0)
        self.ctx.check(self.ctx.L.rf_denoiser_grads(self.h, C.c_void_p(x_ptr), C.c_void_p(t_ptr), N, H, W,
                                                    g.ctypes.data, C.byref(loss)))
        return g, loss.value
