"""Build libginf_bp.so (HIP kernels for gfx950 + C ABI + C++ drop-in API).

Usage: python -m gridinfcuda_amd.build [-j N]

Objects go to gridinfcuda_amd/_build/, the library to
gridinfcuda_amd/libginf_bp.so (in-tree, so it travels with the repository
snapshot).  Cross-compiles without a GPU.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libginf_bp.so")
ARCH = os.environ.get("GINF_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -ffp-contract=off: no FMA contraction, so float sums round like the
# reference's x86-64 SSE code.  Packed FP32 instructions are disabled for
# the device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{INCLUDE}", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources() -> list[str]:
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src: str, build_dir: str = BUILD, extra=()) -> str:
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, *extra, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {os.path.basename(src)}")
    return obj


def _headers_mtime() -> float:
    names = [os.path.join(CSRC, n) for n in os.listdir(CSRC) if n.endswith((".h", ".inc"))]
    names += [os.path.join(INCLUDE, n) for n in os.listdir(INCLUDE) if n.endswith(".h")]
    return max((os.path.getmtime(n) for n in names), default=0.0)


def build(jobs: int | None = None, out: str = LIB, build_dir: str = BUILD, extra=(), incremental: bool = False) -> str:
    """Compile every source and link `out`; `extra` adds compiler flags
    (e.g. -D options for an experimental variant built to another path).
    incremental: keep objects newer than their source and every header."""
    os.makedirs(build_dir, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    hdr = _headers_mtime()

    def one(src):
        obj = os.path.join(build_dir, os.path.basename(src) + ".o")
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(src), hdr):
            return obj
        return _compile(src, build_dir, tuple(extra))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, srcs))
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, out)
    if out == LIB:
        build_dropin_test()
    return out


def build_dropin_test() -> list[str]:
    """Link the C++ callers of the drop-in API under tests/cpp/ against the
    library with the host compiler, one executable each in
    gridinfcuda_amd/bin/."""
    cdir = os.path.join(ROOT, "tests", "cpp")
    os.makedirs(os.path.join(PKG, "bin"), exist_ok=True)
    exes = []
    for name in sorted(os.listdir(cdir)):
        if not name.endswith(".cc"):
            continue
        exe = os.path.join(PKG, "bin", name[:-3])
        cmd = ["g++", "-std=c++17", "-O2", f"-I{INCLUDE}", os.path.join(cdir, name), f"-L{PKG}", "-lginf_bp",
               "-Wl,-rpath,$ORIGIN/..", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"drop-in test link failed: {name}")
        exes.append(exe)
    return exes


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-o", default=LIB, help="output library path")
    ap.add_argument("-D", action="append", default=[], help="extra preprocessor definition")
    ap.add_argument("--incremental", action="store_true", help="recompile only sources newer than their objects")
    args = ap.parse_args()
    bdir = BUILD if args.o == LIB else os.path.join(os.path.dirname(os.path.abspath(args.o)), "_build_variant")
    print(build(args.j, args.o, bdir, [f"-D{d}" for d in args.D], incremental=args.incremental))
