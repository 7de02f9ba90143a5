"""Build recipe of the HIP library (``libbmg.so``) for gfx950.

Compiled in-tree with hipcc so that the shared object travels with the
repository.  ``-ffp-contract=off`` keeps every product and sum rounded in the
order the source writes it (the FP64 path is bit-compatible with the CPU
restatement of the reference); packed FP32 instructions are disabled.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["csrc/bmg_kernels.hip", "csrc/bmg_capi.hip", "csrc/bmg_batch.hip",
           "csrc/bmg_fft.hip", "csrc/unet_kernels.hip", "csrc/unet.hip"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
LIBS = ["-L" + os.path.join(ROCM, "lib"), "-lrocfft",
        "-Wl,-rpath," + os.path.join(ROCM, "lib")]
LIB = os.path.join(HERE, "libbmg.so")
ARCH = os.environ.get("BMG_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


CPP_TESTS = [os.path.join(ROOT, "tests", "cpp", t)
             for t in ("test_hipmg", "test_poisson", "test_split")]


def build(verbose=False):
    cmd = (["hipcc"] + FLAGS + ["-I" + os.path.join(ROOT, "include"),
                                "-o", LIB + ".tmp"]
           + [os.path.join(HERE, s) for s in SOURCES] + LIBS)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr)
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(" ".join(cmd))
    # C++ programs driving the adapters of host/ (GPU tests run them)
    for exe in CPP_TESTS:
        cmd = ["g++", "-O2", "-std=c++17",
               "-I" + os.path.join(ROOT, "include"),
               "-I" + os.path.join(HERE, "host"), exe + ".cpp", "-o", exe,
               "-L" + HERE, "-lbmg", "-Wl,-rpath,$ORIGIN/../../briscola_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("g++ failed:\n" + r.stderr)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
