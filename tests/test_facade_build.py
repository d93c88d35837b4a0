"""The Olive C++ API (olive_amd/include) compiles for gfx950: the examples
and -- where the reference checkout is present -- its unchanged callers
BFS.cu, SSSP.cu and PageRank.cu (the programs of the reference that
instantiate Oliver<...>: BFS.cu:103, SSSP.cu:125, PageRank.cu:110).
CPU only: compiles to object files, runs nothing."""
import os
import subprocess

import pytest

from olive_amd import build as olive_build

EXAMPLES = sorted(f for f in os.listdir(olive_build.EXAMPLES) if f.endswith(".hip"))


def _compile(src, tmp_path):
    obj = os.path.join(str(tmp_path), os.path.basename(src) + ".o")
    cmd = ([olive_build.hipcc(), "-x", "hip", "-c", src, "-o", obj, "-O1", "-std=c++17", "-Werror=return-type",
            f"-I{olive_build.INCLUDE}", f"-I{os.path.join(olive_build.ROOT, 'include')}"]
           + olive_build.device_flags())
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert os.path.getsize(obj) > 0


@pytest.mark.parametrize("name", EXAMPLES)
def test_example_compiles(name, tmp_path):
    _compile(os.path.join(olive_build.EXAMPLES, name), tmp_path)


@pytest.mark.parametrize("name", olive_build.REFERENCE_CALLERS)
def test_reference_caller_compiles_unchanged(name, tmp_path):
    callers = olive_build.reference_callers()
    if not callers:
        pytest.skip("reference checkout not present")
    src = [p for p in callers if os.path.basename(p) == name + ".cu"][0]
    # A link, not a copy: `#include "oliver.h"` then resolves through -I to
    # olive_amd/include instead of to the header beside the original file.
    link = os.path.join(str(tmp_path), name + ".cu")
    os.symlink(src, link)
    _compile(link, tmp_path)


def test_facade_has_no_cuda_includes():
    for f in os.listdir(olive_build.INCLUDE):
        with open(os.path.join(olive_build.INCLUDE, f)) as fh:
            text = fh.read()
        assert "cuda_runtime" not in text and "cuda.h" not in text, f
