"""ctypes binding of the CPU oracle (oracle/liboracle.so) — test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAX_XF, MAX_TERMS, MAX_PRM = 32, 16, 8


class Term(C.Structure):
    _fields_ = [("var", C.c_int32), ("weight", C.c_float), ("prm", C.c_float * MAX_PRM)]


class OXform(C.Structure):
    _fields_ = [("aff", C.c_float * 6), ("post", C.c_float * 6), ("has_post", C.c_int32), ("n_terms", C.c_int32),
                ("terms", Term * MAX_TERMS), ("weight", C.c_float), ("color", C.c_float),
                ("color_speed", C.c_float), ("opacity", C.c_float), ("rotation_frequency", C.c_float)]


class OFlame(C.Structure):
    _fields_ = [("n_xforms", C.c_int32), ("xf", OXform * MAX_XF), ("has_final", C.c_int32), ("fin", OXform),
                ("pal", (C.c_float * 4) * 256)]


class Jsf(C.Structure):
    _fields_ = [("a", C.c_uint32), ("b", C.c_uint32), ("c", C.c_uint32), ("d", C.c_uint32)]


class Rng(C.Structure):
    _fields_ = [("kind", C.c_int), ("jsf", C.POINTER(Jsf)), ("table", C.POINTER(C.c_float)), ("pos", C.c_int)]


_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_U = C.POINTER(C.c_uint32)
_SIGS = {
    "orc_jsf_seed": (None, [C.POINTER(Jsf), C.c_uint32]),
    "orc_jsf_seed_all": (None, [_U, C.c_uint64]),
    "orc_jsf_next_a": (C.c_uint32, [C.POINTER(Jsf)]),
    "orc_jsf_next_d": (C.c_uint32, [C.POINTER(Jsf)]),
    "orc_jsf_randf": (C.c_float, [C.POINTER(Jsf)]),
    "orc_philox4x32_10": (None, [_U, _U, _U]),
    "orc_jsf_from_philox": (None, [C.POINTER(Jsf), C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
    "orc_var_count": (C.c_int, []),
    "orc_var_name": (C.c_char_p, [C.c_int]),
    "orc_var_id": (C.c_int, [C.c_char_p]),
    "orc_var_nparams": (C.c_int, [C.c_int]),
    "orc_var_param": (C.c_char_p, [C.c_int, C.c_int]),
    "orc_var_eval": (None, [C.c_int, C.c_double, C.c_double, C.c_double, _F, _F, C.POINTER(Rng), _D, _D]),
    "orc_apply_f32": (None, [C.POINTER(OXform), C.c_float, C.c_float, C.c_float, C.c_int, C.POINTER(Rng), _F]),
    "orc_apply_f64": (None, [C.POINTER(OXform), C.c_double, C.c_double, C.c_double, C.c_int, C.POINTER(Rng), _D]),
    "orc_select": (C.c_int, [C.POINTER(OFlame), C.c_float]),
    "orc_cdf": (None, [C.POINTER(OFlame), _F]),
    "orc_animate": (None, [C.POINTER(OFlame), C.c_uint32, C.c_uint32, C.c_float, C.POINTER(OFlame)]),
    "orc_screen_affine": (None, [C.c_uint32, C.c_uint32, _U, _F, C.c_float, C.c_float, _F]),
    "orc_hammersley": (None, [C.c_uint32, _F]),
    "orc_pal_index": (C.c_int, [C.c_float]),
    "orc_compat_pass": (None, [C.POINTER(OFlame), C.c_int, C.c_uint32, C.c_uint32, _F, _F, _U, _U, C.c_uint32,
                               C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32, _F, _D, _U,
                               C.POINTER(C.c_uint64)]),
    "orc_fast_run": (None, [C.POINTER(OFlame), C.c_int, C.c_uint32, C.c_uint64, C.c_int,
                            C.POINTER(C.c_uint16), C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                            C.c_uint32, C.c_uint32, _F, _F, _D, _U, C.POINTER(C.c_uint64)]),
    "orc_ref_passes_mt": (C.c_uint64, [C.POINTER(OFlame), C.c_uint32, C.c_uint32, _F, _F, _U, _U, _U, C.c_uint32,
                                       C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32, _F, _F,
                                       C.c_int]),
    "orc_chaos_f64": (C.c_uint64, [C.POINTER(OFlame), C.c_uint32, C.c_uint32, _D, C.c_uint64, C.c_uint32,
                                   C.c_uint32, C.c_uint64, C.c_int, _D]),
    "orc_density": (None, [_F, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_float, C.c_int, _D]),
    "orc_tonemap": (None, [_D, _D, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for n, (res, args) in _SIGS.items():
            fn = getattr(L, n)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def fp(a):
    return a.ctypes.data_as(_F)


def dp(a):
    return a.ctypes.data_as(_D)


def up(a):
    return a.ctypes.data_as(_U)


def var_id(name: str) -> int:
    return lib().orc_var_id(name.encode())


def var_params(name: str):
    L = lib()
    i = var_id(name)
    return [L.orc_var_param(i, k).decode() for k in range(L.orc_var_nparams(i))]


def xform_to_c(x) -> OXform:
    """refrakt_amd.Xform (or any object with the same fields) -> oracle xform."""
    o = OXform()
    o.aff[:] = [float(v) for v in x.affine]
    if x.post is not None:
        o.post[:] = [float(v) for v in x.post]
        o.has_post = 1
    o.n_terms = len(x.variations)
    for t, (name, w) in enumerate(x.variations.items()):
        vid = var_id(name)
        if vid < 0:
            raise KeyError(name)
        o.terms[t].var = vid
        o.terms[t].weight = float(w)
        for k, pn in enumerate(var_params(name)):
            o.terms[t].prm[k] = float(x.var_param.get(pn, 0.0))
    o.weight, o.color, o.color_speed = float(x.weight), float(x.color), float(x.color_speed)
    o.opacity, o.rotation_frequency = float(x.opacity), float(x.rotation_frequency)
    return o


def flame_to_c(f) -> OFlame:
    o = OFlame()
    o.n_xforms = len(f.xforms)
    for i, x in enumerate(f.xforms):
        o.xf[i] = xform_to_c(x)
    if f.final_xform is not None:
        o.has_final = 1
        o.fin = xform_to_c(f.final_xform)
    pal = np.ascontiguousarray(f.palette, np.float32).reshape(256, 4)
    C.memmove(o.pal, pal.ctypes.data, pal.nbytes)
    return o


def screen_affine(f, W, H):
    out = np.zeros(6, np.float32)
    size = np.array(f.size, np.uint32)
    center = np.array(f.center, np.float32)
    lib().orc_screen_affine(W, H, up(size), fp(center), float(f.scale), float(f.rotate), fp(out))
    return out


def hammersley(count):
    out = np.zeros((count, 4), np.float32)
    lib().orc_hammersley(count, fp(out))
    return out


def jsf_seed_states(n):
    out = np.zeros((n, 4), np.uint32)
    lib().orc_jsf_seed_all(up(out), n)
    return out


def apply_f64(x, pts, rands=None):
    """Applies one xform in double precision to pts (n x 3); rands n x 16 table."""
    L = lib()
    ox = xform_to_c(x)
    pts = np.asarray(pts, np.float64).reshape(-1, 3)
    n = pts.shape[0]
    rt = np.zeros((n, 16), np.float32) if rands is None else np.ascontiguousarray(rands, np.float32).reshape(n, 16)
    out = np.zeros((n, 4), np.float64)
    o = np.zeros(4, np.float64)
    for i in range(n):
        row = np.ascontiguousarray(rt[i])
        rng = Rng(1, None, fp(row), 0)
        L.orc_apply_f64(C.byref(ox), pts[i, 0], pts[i, 1], pts[i, 2], 0, C.byref(rng), dp(o))
        out[i] = o
    return out


def var_eval(name, x, y, w=1.0, params=None, affine=(1, 0, 0, 1, 0, 0), rands=None):
    L = lib()
    prm = np.zeros(MAX_PRM, np.float32)
    if params:
        for k, pn in enumerate(var_params(name)):
            prm[k] = params.get(pn, 0.0)
    aff = np.array(affine, np.float32)
    rt = np.zeros(16, np.float32) if rands is None else np.asarray(rands, np.float32)
    rng = Rng(1, None, fp(rt), 0)
    ox, oy = C.c_double(), C.c_double()
    L.orc_var_eval(var_id(name), float(x), float(y), float(w), fp(prm), fp(aff), C.byref(rng), C.byref(ox),
                   C.byref(oy))
    return ox.value, oy.value


def animate(flame_c, i, T, tss):
    """Temporal sample i of T (animate.tpl.glsl:18-96): a copy of flame_c with
    rotated affines."""
    out = OFlame()
    lib().orc_animate(C.byref(flame_c), i, T, tss, C.byref(out))
    return out


def compat_pass(flame_c, T, ppt, pos_in, pos_out, shuf, rs, in_row, out_row, random_read, random_write, first_run,
                do_draw, W=1, H=1, ss=None, bins=None, hits=None, flames_per_sample=None):
    L = lib()
    drawn = C.c_uint64()
    ss = np.zeros(6, np.float32) if ss is None else np.ascontiguousarray(ss, np.float32)
    if flames_per_sample is not None:
        arr = (OFlame * T)(*flames_per_sample)
        fl, stride = arr, 1
    else:
        fl, stride = C.byref(flame_c), 0
    L.orc_compat_pass(fl, stride, T, ppt, fp(pos_in), fp(pos_out), up(shuf), up(rs), in_row, out_row,
                      int(random_read), int(random_write), int(first_run), int(do_draw), W, H, fp(ss),
                      dp(bins) if bins is not None else None, up(hits) if hits is not None else None,
                      C.byref(drawn))
    return drawn.value


def ref_passes_mt(flame_c, T, ppt, pos_a, pos_b, shuf, rs, rows, n_passes, random_read, random_write, first_run,
                  do_draw, W=1, H=1, ss=None, bins=None, threads=1):
    """n_passes reference passes on `threads` threads (orc_ref_passes_mt);
    pass k reads pos_a if k is even, else pos_b.  Returns the binned count."""
    L = lib()
    ss = np.zeros(6, np.float32) if ss is None else np.ascontiguousarray(ss, np.float32)
    rows = np.ascontiguousarray(rows, np.uint32)
    return L.orc_ref_passes_mt(C.byref(flame_c), T, ppt, fp(pos_a), fp(pos_b), up(shuf), up(rs), up(rows), n_passes,
                               int(random_read), int(random_write), int(first_run), int(do_draw), W, H, fp(ss),
                               fp(bins) if bins is not None else None, int(threads))


def fast_run(flame_c, ppt, P, waves, perms, seed, launch, n_iter, first_run, do_draw, W, H, ss, pos, bins=None,
             hits=None):
    L = lib()
    drawn = C.c_uint64()
    perms = np.ascontiguousarray(perms, np.uint16)
    L.orc_fast_run(C.byref(flame_c), 0, ppt, P, waves, perms.ctypes.data_as(C.POINTER(C.c_uint16)), seed, launch,
                   n_iter, int(first_run), int(do_draw), W, H, fp(np.ascontiguousarray(ss, np.float32)), fp(pos),
                   dp(bins) if bins is not None else None, up(hits) if hits is not None else None, C.byref(drawn))
    return drawn.value


def chaos_f64(flame, W, H, particles, fuse, iters, seed=1, threads=None, ss=None):
    """FP64 oracle histogram (H x W x 4 doubles, bin layout) and binned count."""
    L = lib()
    fc = flame_to_c(flame)
    if ss is None:
        ss = screen_affine(flame, W, H)
    ssd = np.ascontiguousarray(ss, np.float64)
    bins = np.zeros((H, W, 4), np.float64)
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    n = L.orc_chaos_f64(C.byref(fc), W, H, dp(ssd), particles, fuse, iters, seed, threads, dp(bins))
    return bins, n


def density(hist, R, rmin, curve, fp32=True):
    L = lib()
    hist = np.ascontiguousarray(hist, np.float32)
    H, W = hist.shape[:2]
    out = np.zeros((H, W, 4), np.float64)
    L.orc_density(fp(hist), W, H, int(R), int(rmin), float(curve), int(fp32), dp(out))
    return out


def tonemap(img, gamma, brightness, vibrancy, k=1e-4):
    L = lib()
    img = np.ascontiguousarray(img, np.float64)
    out = np.zeros_like(img)
    L.orc_tonemap(dp(img), dp(out), img.size // 4, gamma, brightness, vibrancy, k)
    return out
