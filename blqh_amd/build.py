"""Build libblqh.so (HIP kernels + C ABI) and the blqh CLI for gfx950, in-tree.

    python -m blqh_amd.build            # compile everything
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")
LIB = os.path.join(LIBDIR, "libblqh.so")
CLI = os.path.join(BINDIR, "blqh")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are disabled for all device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
SOURCES = ["basis.hip", "hamil.hip", "panel.hip", "tiles.hip", "lanczos.hip", "batch.hip", "mlp.hip", "obs.hip", "dense.hip", "records.cpp", "api.cpp"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp] + objs +
                          ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "cli.cpp")
    if os.path.exists(cli_src):
        for exe, defs in ((CLI, []), (CLI + "_spin", ["-DBLQH_SPINFUL"])):
            subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include")] + defs +
                                  [cli_src, "-o", exe, "-L", LIBDIR, "-lblqh", "-Wl,-rpath,$ORIGIN/../lib"])
    hc_src = os.path.join(ROOT, "tests", "hostcheck", "hostcheck.cpp")
    if os.path.exists(hc_src):
        # CPU mirror of the device pipeline, used by the CPU test suite only
        hc_dir = os.path.join(ROOT, "tests", "hostcheck", "_build")
        os.makedirs(hc_dir, exist_ok=True)
        subprocess.check_call([HIPCC] + COMMON + ["-x", "hip", "-shared", hc_src, "-o",
                                                  os.path.join(hc_dir, "libhostcheck.so"), "-L", LIBDIR, "-lblqh",
                                                  f"-Wl,-rpath,{LIBDIR}"])
    if verbose:
        print(f"built {LIB}" + (f" and {CLI}" if os.path.exists(cli_src) else ""))
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
