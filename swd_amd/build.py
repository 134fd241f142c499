"""Build recipe for libswd.so (HIP kernels for gfx950 + the C ABI).

``python -m swd_amd.build`` (or ``__graft_entry__.build()``) compiles every
source in ``swd_amd/csrc`` with hipcc and links ``swd_amd/libswd.so`` in tree,
so the library travels with the repository snapshot.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libswd.so")
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}"]
# Device code: gfx950 only; packed-FP32 instruction selection disabled.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src + ".o")
    path = os.path.join(CSRC, src)
    cmd = [HIPCC, "-c", path, "-o", obj] + COMMON
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", "-c", path, "-o", obj] + COMMON + DEVICE
    else:
        cmd += ["-D__HIP_PLATFORM_AMD__"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        lines = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


SAN_LIB = os.path.join(HERE, "libswd_san.so")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]


def build_sanitized(verbose: bool = True) -> str:
    """libswd_san.so: the host C++ sources (engine, server and arenas, MEX
    argument marshalling, model preparation) compiled by g++ with
    AddressSanitizer and UndefinedBehaviorSanitizer, linked with the HIP
    objects of the normal build (device code is not instrumented).  Load it
    with SWD_LIBRARY=<path> and LD_PRELOAD of gcc's libasan."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    hip = [os.path.join(OBJ, f + ".o") for f in srcs if f.endswith(".hip")]
    if not all(os.path.exists(o) for o in hip):
        build(verbose=False)
    gxx = shutil.which("g++") or "g++"

    def comp(f):
        obj = os.path.join(OBJ, f + ".san.o")
        cmd = [gxx, "-std=c++17", "-O1", "-g", "-fPIC", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
               "-I/opt/rocm/include", f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}", "-c",
               os.path.join(CSRC, f), "-o", obj] + SAN_FLAGS
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {f}:\n{' '.join(cmd)}\n{r.stderr}")
        return obj

    cpp = [f for f in srcs if f.endswith(".cpp")]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cpp))) as ex:
        objs = list(ex.map(comp, cpp))
    tmp = SAN_LIB + ".tmp"
    cmd = [gxx, "-shared", "-o", tmp] + objs + hip + SAN_FLAGS + ["-L/opt/rocm/lib", "-lamdhip64",
                                                                  "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, SAN_LIB)
    if verbose:
        print(f"built {SAN_LIB} (host code with ASan + UBSan)")
    return SAN_LIB


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        build_sanitized()
    else:
        build()
