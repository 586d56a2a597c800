"""Bounds-checked CPU emulations of the grid model's kernels (tests/emul/):
the K_mm tile layout and convolution blocks, the dense-stream wavefront
mapping, the along-strike K_fm / K_mf tables with their F and G blocks and
the role table, at the shapes of the composite and 65536-unknown models and
at odd ones.  Each restates the kernels' index arithmetic over buffers that
fail on an out-of-range index or an unfilled (NaN) value and compares the
result with a dense product of a synthetic translation-invariant pair
function (relative error < 1e-13, tables and expansions exact).  They run
before a launch shape is first sent to a GPU."""
import os
import shutil
import subprocess

import pytest

EMUL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emul")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("name,flags", [("grid_emul", ["-fopenmp"]), ("strike_emul", [])])
def test_kernel_index_emulation(tmp_path, name, flags):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall"] + flags +
                   ["-o", exe, os.path.join(EMUL, name + ".cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert r.stdout.strip().splitlines()[-1] == "ok"
