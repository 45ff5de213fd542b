"""Build the HIP library ``tbplas_amd/_build/libtbplas_amd.so`` for gfx950.

Run ``python -m tbplas_amd.build`` (or ``__graft_entry__.build()``).  The
sources are compiled in parallel with hipcc and linked into one shared
library that exports the C ABI of ``include/tbplas_amd.h``.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libtbplas_amd.so")
# the longest compiles first: the one-term kernels, then the tile-pass instances
SOURCES = ["tbg_pass_f32_t5_real.hip", "tbg_pass_f32_t5_cplx.hip", "tbg_pass_f64_t5_real.hip",
           "tbg_pass_f64_t5_cplx.hip", "tbg_prop.hip", "tbg_pass_f32_t4_real.hip",
           "tbg_pass_f32_t4_cplx.hip", "tbg_pass_f64_t4_real.hip", "tbg_pass_f64_t4_cplx.hip",
           "tbg_pass_t2.hip",
           "tbg_pass_f32_t3_real.hip", "tbg_pass_f32_t3_cplx.hip", "tbg_pass_f64_t3_real.hip",
           "tbg_pass_f64_t3_cplx.hip", "tbg_super.hip", "tbg_api.hip", "tbg_cond.hip",
           "tbg_block.hip", "tbg_mlp.hip", "tbg_csr.hip", "tbg_hall.hip", "tbg_static.hip"]
ARCH = os.environ.get("TBG_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for all device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n"
                           f"{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    """Compile every HIP source and link the shared library."""
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
