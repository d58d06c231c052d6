"""Compile the HIP library ``libflexmm_amd.so`` in-tree for gfx950.

Used by ``__graft_entry__.build()`` and ``python -m flexmm_amd.build``.  The
build needs only ``hipcc`` (ROCm); no GPU is required to compile.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libflexmm_amd.so")
SOURCES = ["prims.hip", "tree.hip", "fmm.hip"]
ARCH = os.environ.get("FLEXMM_AMD_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 VALU forms are kept out of the generated code on purpose
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the tree's midpoint recurrence and all kernels spell out their FMAs
    "-ffp-contract=off",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the ROCm toolchain is required to build flexmm_amd")
    return path


def build(verbose=False):
    """Compile all sources into ``flexmm_amd/libflexmm_amd.so`` (always rebuilds)."""
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=CSRC, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed (%d):\n%s" % (res.returncode, out))
    os.replace(tmp, LIB)
    if verbose and out.strip():
        print(out)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
