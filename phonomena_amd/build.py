"""Build libphonofdtd.so for gfx950 in-tree (phonomena_amd/lib/).

    python -m phonomena_amd.build

hipcc only (no CMake): one shared library from the HIP translation units in
phonomena_amd/csrc.  Flags:
  -ffp-contract=off   the FP64 mode must keep the reference's expression
                      order; FP32 code uses explicit fused multiply-adds
  -packed-fp32-ops    off: scalar FP32 code is not turned into packed ops
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib", "libphonofdtd.so")
SOURCES = ["pf_api.hip", "pf_batch.hip", "pf_diag.hip", "pf_spectrum.hip"]
ARCH = os.environ.get("PF_OFFLOAD_ARCH", "gfx950")

FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function"]


def build(verbose=False):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["hipcc"] + FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    msgs = [ln for ln in (r.stdout + r.stderr).splitlines() if noise not in ln and ln.strip()]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
