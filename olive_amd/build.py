"""Build libolive_hip.so in-tree with hipcc for gfx950.

Usage: python -m olive_amd.build [--jobs N] [--verbose]

Every .hip file under csrc/ is compiled to an object (in parallel), gen.cpp
(pure host code) likewise, and everything is linked into
olive_amd/libolive_hip.so next to this file.  Objects live in
olive_amd/_build/ (git-ignored).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libolive_hip.so")
ARCH = os.environ.get("OLIVE_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def common_flags():
    return [
        "-O3", "-std=c++17", "-fPIC",
        f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
        "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
    ]


def device_flags():
    # packed fp32 ops disabled on purpose for all device code
    return [f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def compile_one(src: str, verbose: bool, build_dir: str = BUILD, defines=()) -> str:
    base = os.path.splitext(os.path.basename(src))[0]
    obj = os.path.join(build_dir, base + ".o")
    dflags = [f"-D{d}" for d in defines]
    if src.endswith(".hip"):
        cmd = [hipcc(), "-c", "-x", "hip", src, "-o", obj] + common_flags() + device_flags() + dflags
    else:
        cmd = [hipcc(), "-c", src, "-o", obj] + common_flags() + dflags
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip() and verbose:
        print(r.stderr, file=sys.stderr)
    return obj


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def build(jobs: int = 8, verbose: bool = False, defines=(), out: str = LIB) -> str:
    """Compile and link.  `defines` (NAME=VALUE) override the tuning macros
    of the sources; `out` names the library (default: the in-tree one)."""
    build_dir = BUILD if out == LIB else out + ".objs"
    os.makedirs(build_dir, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose, build_dir, defines), srcs))
    tmp = out + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-lpthread", "-Wl,--no-undefined"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--define", action="append", default=[], help="NAME=VALUE tuning macro")
    ap.add_argument("--out", default=LIB)
    a = ap.parse_args(argv)
    print(build(a.jobs, a.verbose, a.define, os.path.abspath(a.out)))


if __name__ == "__main__":
    main()
