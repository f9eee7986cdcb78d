"""Builds the native libraries in-tree.

libpvlbm.so  — the gfx950 HIP kernels and the C ABI of include/pvlbm.h (hipcc)
libpvnet.so  — the warm-start network's gfx950 kernels and the C ABI of include/pvnet.h (hipcc)
libpvgen.so  — synthetic vessel mesh generators of include/pvgen.h (g++)
lbm_run      — command-line driver over the C++ facade include/pvlbm.hpp (g++)

Every HIP translation unit is compiled with -ffp-contract=off so that the FP64 path keeps the
reference's operation order without fused multiply-adds, and with packed FP32 instructions
disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

LIBS = {
    "libpvlbm.so": ["capi.cpp", "k_step.hip", "k_setup.hip", "noalloc.hip"],
    "libpvnet.so": ["k_net.hip"],
}

# per-library flags: the network's MFMA accumulators stay in VGPRs (the AGPR form made the compiler
# copy every accumulator between the two register files around each MFMA of the k-loops)
LIB_FLAGS = {
    "libpvnet.so": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]))
    return r


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    for lib, srcs in LIBS.items():
        out = os.path.join(HERE, lib)
        tmp = out + ".tmp"
        _run([hipcc] + HIP_FLAGS + LIB_FLAGS.get(lib, []) + [os.path.join(CSRC, s) for s in srcs] + ["-o", tmp],
             verbose)
        os.replace(tmp, out)
    gen = os.path.join(HERE, "libpvgen.so")
    _run(["g++", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-Wall",
          os.path.join(CSRC, "geomgen.cpp"), "-o", gen + ".tmp"], verbose)
    os.replace(gen + ".tmp", gen)
    # the command-line driver over the C++ facade (include/pvlbm.hpp)
    exe = os.path.join(HERE, "lbm_run")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "lbm_run.cpp"), "-L", HERE, "-lpvlbm", "-Wl,-rpath,$ORIGIN",
          "-o", exe + ".tmp"], verbose)
    os.replace(exe + ".tmp", exe)


def build_test_shims(verbose=False):
    """Compile the test shim that drives the C++ facade's loop in-process (test infrastructure)."""
    src = os.path.join(ROOT, "tests", "facade", "facade_iterate.cpp")
    out = os.path.join(ROOT, "tests", "facade", "libfacade_iterate.so")
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"), src,
          "-L", HERE, "-lpvlbm", "-Wl,-rpath,$ORIGIN/../../pv_dec_original_amd", "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)


def build_oracle(verbose=False):
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)


if __name__ == "__main__":
    build(verbose=True)
    build_oracle(verbose=True)
    build_test_shims(verbose=True)
