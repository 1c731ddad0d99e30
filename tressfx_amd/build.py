"""Builds the gfx950 shared library libtressfx_amd.so in-tree with hipcc.

Usage: python -m tressfx_amd.build   (or tressfx_amd.build.build_library())
"""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libtressfx_amd.so")

SOURCES = [
    "tfx_capi.hip",
    "tfx_sim_kernels.hip",
    "tfx_fused_kernel.hip",
    "tfx_sdf_kernels.hip",
    "tfx_batch.hip",
    "tfx_mesh.hip",
    "tfx_server.hip",
    "tfx_asset.hip",
]

# Kernel sources built a second time for TFX_ARITH_FAST: their kernels go to
# namespace tfx::fast and their C entry points get a _fast suffix
# (tfx_device.h).  That build divides by the hardware reciprocal (a / b =
# a * v_rcp(b)), takes square roots and reciprocal square roots with the
# hardware instructions (-fapprox-func: no range scaling around them), and
# fuses multiply-adds (-ffp-contract=fast, overriding the exact build's off).
FAST_SOURCES = [
    "tfx_sim_kernels.hip",
    "tfx_fused_kernel.hip",
    "tfx_sdf_kernels.hip",
]
FAST_FLAGS = ["-DTFX_FAST_BUILD", "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fapprox-func",
              "-ffp-contract=fast"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every multiply-add is evaluated as written (the parity
# contract with the FP32 transcription).  Packed FP32 instructions are
# disabled for the device pass.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(REPO_DIR, "include"),
]


CPP_TEST_SRC = os.path.join(REPO_DIR, "tests", "cpp", "shim_test.cpp")
CPP_TEST_BIN = os.path.join(REPO_DIR, "tests", "cpp", "shim_test")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def cpp_test_command(out=CPP_TEST_BIN):
    """g++ line of the C++ shim test program: the C++ classes of
    include/tressfx_amd.hpp over libtressfx_amd.so (found next to the package
    through a relative rpath)."""
    rpath = os.path.relpath(PKG_DIR, os.path.dirname(os.path.abspath(out)))
    return ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
            "-I", os.path.join(REPO_DIR, "include"), "-I", os.path.join(ROCM, "include"), CPP_TEST_SRC,
            "-L", PKG_DIR, "-ltressfx_amd", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
            "-Wl,-rpath,$ORIGIN/" + rpath, "-o", out]


def build_cpp_test(verbose=False, out=CPP_TEST_BIN):
    cmd = cpp_test_command(out)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError("g++ failed building the C++ shim test")
    return out


def build_library(verbose=False):
    """Every source to an object (the FAST_SOURCES twice), in parallel, then
    one link."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor
    compile_flags = [f for f in FLAGS if f != "-shared"]
    tmp = LIB_PATH + ".tmp"
    with tempfile.TemporaryDirectory() as work:
        jobs = [(src, os.path.join(work, os.path.splitext(src)[0] + ".o"), []) for src in SOURCES]
        jobs += [(src, os.path.join(work, os.path.splitext(src)[0] + "_fast.o"), FAST_FLAGS) for src in FAST_SOURCES]

        def compile_one(job):
            src, obj, extra = job
            cmd = [HIPCC] + compile_flags + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            return src, subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)

        workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 4)))
        with ThreadPoolExecutor(max_workers=workers) as pool:
            for src, res in pool.map(compile_one, jobs):
                if res.returncode != 0:
                    sys.stderr.write(res.stdout)
                    raise RuntimeError(f"hipcc failed compiling {src}")
        cmd = [HIPCC] + FLAGS + [obj for _, obj, _ in jobs] + ["-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout)
            raise RuntimeError("hipcc failed linking libtressfx_amd.so")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_cpp_test(verbose=True))
