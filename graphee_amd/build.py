"""Builds libgmp.so (the engine's HIP kernels + C ABI) in-tree for gfx950.

Every translation unit under csrc/ is compiled by hipcc as HIP with
  -ffp-contract=off                              (no implicit FMA: the GraphMP-compatible
                                                  apps reproduce the reference's separate
                                                  multiply and add roundings)
  -fhip-fp32-correctly-rounded-divide-sqrt       (IEEE float division, as on the CPU)
  -Xclang -target-feature -Xclang -packed-fp32-ops  (no packed FP32 instructions)
and linked into graphee_amd/libgmp.so.  Objects go to graphee_amd/_build/.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgmp.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-x", "hip", *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_apps(verbose)
    return LIB


APPS = os.path.join(HERE, "apps")
BIN = os.path.join(HERE, "bin")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_apps(verbose=True):
    """GraphMP-style apps (host C++ over the C ABI) -> graphee_amd/bin/gmp_<app>."""
    os.makedirs(BIN, exist_ok=True)
    for src in sorted(f for f in os.listdir(APPS) if f.endswith(".cpp")):
        out = os.path.join(BIN, "gmp_" + src[:-4])
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", os.path.join(APPS, src), f"-L{HERE}",
               "-lgmp", "-Wl,-rpath,$ORIGIN/..", "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src}:\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"built {out}")


if __name__ == "__main__":
    build()
    sys.exit(0)
