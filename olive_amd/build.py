"""Build libolive_hip.so in-tree with hipcc for gfx950.

Usage: python -m olive_amd.build [--jobs N] [--verbose]

Every .hip file under csrc/ is compiled to an object (in parallel), gen.cpp
(pure host code) likewise, and everything is linked into
olive_amd/libolive_hip.so next to this file.  Objects live in
olive_amd/_build/ (git-ignored).

build_apps() compiles the programs written against the Olive C++ API
(olive_amd/include) into olive_amd/_apps/: the examples under
olive_amd/examples/, and -- where the reference checkout is present
(OLIVE_REFERENCE, default /root/reference) -- its unchanged callers BFS.cu,
SSSP.cu and PageRank.cu as ref_BFS, ref_SSSP, ref_PageRank.  A caller is
compiled through a symbolic link in a scratch folder, so that its
`#include "oliver.h"` resolves to olive_amd/include/oliver.h and not to the
header beside it; the source itself is read where it lies.
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
INCLUDE = os.path.join(HERE, "include")
EXAMPLES = os.path.join(HERE, "examples")
APPS = os.path.join(HERE, "_apps")
REFERENCE = os.environ.get("OLIVE_REFERENCE", "/root/reference")
REFERENCE_CALLERS = ("BFS", "SSSP", "PageRank")
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


def _compile_app(src: str, out: str, verbose: bool) -> str:
    cmd = ([hipcc(), "-x", "hip", src, "-o", out, "-O3", "-std=c++17",
            f"-I{INCLUDE}", f"-I{os.path.join(ROOT, 'include')}"] + device_flags()
           + [f"-L{HERE}", "-lolive_hip", "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return out


def reference_callers():
    """Paths of the reference's caller programs, or [] if absent."""
    paths = [os.path.join(REFERENCE, n + ".cu") for n in REFERENCE_CALLERS]
    return paths if all(os.path.isfile(p) for p in paths) else []


def build_apps(jobs: int = 8, verbose: bool = False) -> list:
    """Compile the examples and, if present, the reference's callers against
    olive_amd/include and libolive_hip.so (which must be built)."""
    if not os.path.exists(LIB):
        raise RuntimeError("build libolive_hip.so first")
    os.makedirs(APPS, exist_ok=True)
    jobs_list = []
    for f in sorted(os.listdir(EXAMPLES)):
        if f.endswith(".hip"):
            jobs_list.append((os.path.join(EXAMPLES, f), os.path.join(APPS, f[:-4])))
    links = []
    callers = reference_callers()
    if callers:
        src_dir = os.path.join(APPS, "_src")
        os.makedirs(src_dir, exist_ok=True)
        for p in callers:
            name = os.path.splitext(os.path.basename(p))[0]
            link = os.path.join(src_dir, name + ".cu")
            if os.path.lexists(link):
                os.remove(link)
            os.symlink(p, link)
            links.append(link)
            jobs_list.append((link, os.path.join(APPS, "ref_" + name)))
    try:
        with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
            outs = list(ex.map(lambda j: _compile_app(j[0], j[1], verbose), jobs_list))
    finally:
        for link in links:
            os.remove(link)
        if links:
            os.rmdir(os.path.dirname(links[0]))
    return outs


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--define", action="append", default=[], help="NAME=VALUE tuning macro")
    ap.add_argument("--out", default=LIB)
    ap.add_argument("--apps", action="store_true", help="also build olive_amd/_apps/")
    a = ap.parse_args(argv)
    print(build(a.jobs, a.verbose, a.define, os.path.abspath(a.out)))
    if a.apps:
        for p in build_apps(a.jobs, a.verbose):
            print(p)


if __name__ == "__main__":
    main()
