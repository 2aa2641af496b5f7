"""Builds librefrakt_amd.so (HIP kernels for gfx950 + host library) in-tree.

Run as ``python -m refrakt_amd.build`` or through ``__graft_entry__.build()``.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librefrakt_amd.so")
SOURCES = ["rf_k_compat.hip", "rf_k_fast.hip", "rf_k_misc.hip", "rf_host.cpp", "rf_flam3.cpp", "rf_output.cpp",
           "rf_k_batch.hip", "rf_batch.cpp", "rf_k_nn.hip", "rf_denoiser.cpp", "rf_k_sort.hip"]
ARCH = os.environ.get("RF_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # determinism: no implicit contraction; fmaf is written where the reference uses fma
    "-ffp-contract=off",
    # division and sqrt within the ~2.5 ulp GLSL allows; the few results that
    # must match the CPU bit for bit call the correctly rounded forms explicitly
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    # packed FP32 ops are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def generate():
    try:
        import yaml  # noqa: F401
    except ImportError:
        return  # generated files are committed
    sys.path.insert(0, HERE)
    try:
        import gen_variations
        gen_variations.main([])
    finally:
        sys.path.pop(0)


def build(verbose=False):
    generate()
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    jobs = []
    for src in SOURCES:
        obj = os.path.join(build_dir, src.rsplit(".", 1)[0] + ".o")
        jobs.append(([hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], obj))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 4)) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
