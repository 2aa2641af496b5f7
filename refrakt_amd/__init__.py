"""refrakt_amd: a fractal-flame renderer for AMD MI355X (gfx950).

The hot path (chaos game -> RGBA FP32 histogram -> density estimation ->
tonemap) runs in hand-written HIP kernels behind the C ABI declared in
include/refrakt_amd.h; this package is the Python host mirror of the
reference's flame API used by the tests and the benchmark.
"""
from .api import (  # noqa: F401
    Context, DEParams, Flame, FlameDesc, FlameHandle, Histogram, RefraktError, RenderParams, Stats,
    TonemapParams, Xform, XformDesc, lib, load_flame, param_block, parse_flam3, program_build, program_source, render,
    screen_affine, to_rgba8, validate, variation_id, variation_names, write_exr, write_png,
    variation_params, RF_ENGINE_COMPAT, RF_ENGINE_FAST, RF_HIST_AUTO, RF_HIST_LDS, RF_HIST_NAIVE, RF_HIST_SORT, RF_HIST_QUAD,
    RF_HIST_WAVEAGG,
)
