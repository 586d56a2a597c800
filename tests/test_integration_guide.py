"""The integration guide's examples against the code they describe: its C blocks
compile against include/oetqf_amd.h (gcc -fsyntax-only, with the example's
variables declared around them), its Julia ccall matches the header's
signature, and every oq.<name> and Ensemble / Group method it uses exists in
the Python package."""
import os
import re
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GUIDE = os.path.join(ROOT, "INTEGRATION.md")

_DECLS = """#include <stdio.h>
#include <stdint.h>
#include "oetqf_amd.h"
void example(void) {
    oq_fault_mesh mf; oq_hex8_mesh ma; oq_quadrature quad; oq_rs_props fault; oq_visc_props visc;
    oq_solver_opts opts;
    double lam = 0, mu = 0, t = 0, tend = 1, *v = 0, *theta = 0, *eps = 0, *sigma = 0, *delta = 0, *dv = 0,
           *dtheta = 0, *deps = 0, *dsigma = 0, *ddelta = 0, *u0 = 0, *step_t = 0, *step_vmax = 0,
           *snap_t = 0, *snap_u = 0;
    int64_t stride = 1, nstep = 1, nst = 0, nsn = 0;
"""


def _guide():
    return open(GUIDE).read()


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_c_examples_compile_against_the_header(tmp_path):
    blocks = re.findall(r"```c\n(.*?)```", _guide(), re.S)
    assert blocks
    body = "\n".join(b.replace('#include "oetqf_amd.h"\n', "") for b in blocks)
    src = tmp_path / "guide.c"
    src.write_text(_DECLS + body + "\n}\n")
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-Wno-unused-variable", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_julia_ccall_example_matches_the_header():
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_julia_binding_signatures import _header, _problems
    bad, seen, n = _problems(_guide(), _header())
    assert n >= 1 and not bad, bad


def test_python_names_exist(oq):
    text = _guide()
    names = set(re.findall(r"\boq\.([A-Za-z_][A-Za-z0-9_]*)", text))
    assert names
    assert [n for n in names if not hasattr(oq, n)] == []
    for cls, var in ((oq.Ensemble, "ens"), (oq.Group, "grp")):
        methods = set(re.findall(r"\b%s\.([A-Za-z_][A-Za-z0-9_]*)\(" % var, text))
        assert methods
        assert [m for m in methods if not hasattr(cls, m)] == [], cls.__name__
