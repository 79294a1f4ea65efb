"""Build the gfx950 library climage_amd/lib/libclimage_hip.so with hipcc.

Every translation unit is compiled with -ffp-contract=off: the kernels must
round every FP32 multiply and add separately (see device_common.h), and the
host code that prepares per-frame constants must do the same.
Packed FP32 instructions are disabled (-packed-fp32-ops).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libclimage_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", SRC,
]


def sources():
    out = []
    for f in sorted(os.listdir(SRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(SRC, f))
    return out


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=True, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, len(srcs))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, err in results:
        if verbose and err.strip():
            # hipcc's host pass warns that it does not know the packed-fp32 feature
            lines = [l for l in err.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if lines:
                print("\n".join(lines), file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_examples(verbose)
    return LIB


def build_examples(verbose=True):
    """C++ drivers of the typed API (include/*.hpp), linked against the library."""
    ex_dir = os.path.join(HERE, "examples")
    for f in sorted(os.listdir(ex_dir)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(LIB_DIR, f[:-4])
        cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(ex_dir, f),
               "-o", exe, "-L", LIB_DIR, "-lclimage_hip", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {f}:\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"built {exe}")


if __name__ == "__main__":
    build()
