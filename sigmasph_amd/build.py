"""Build libsigmasph.so (HIP, gfx950) in-tree with hipcc.

Run as ``python -m sigmasph_amd.build``; ``__graft_entry__.build()`` calls `build()`.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "sigmasph.hip")
DEPS = [os.path.join(HERE, "csrc", "sph_device.h"), os.path.join(HERE, "csrc", "sph_solid.h"),
        os.path.join(HERE, "csrc", "sph_mlp.h"),
        os.path.join(ROOT, "include", "sigmasph.h")]
LIB = os.path.join(HERE, "libsigmasph.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 VALU ops stay off on this target (they are generated from ordinary
    # float code and have produced silent wrong results on these machines)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build(verbose=False):
    """Compile the library unconditionally (the build check compiles every time)."""
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building %s" % SRC)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
