"""CPU tests of the C ABI boundary: the library loads, exports exactly what
include/voluma_amd.h declares, and the host-side mirror keeps the reference's
defaults (src/fvm/linalg/LinearSolver.h:15-31). No GPU calls."""
import ctypes
import os
import re
import subprocess

import voluma_amd as va
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "voluma_amd.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|void|const char\*)\s+(vol_[A-Za-z_0-9]+)\(", txt, re.M)))


def test_header_declares_and_library_exports_every_symbol():
    names = declared()
    assert len(names) >= 20
    lib = va.lib()
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", va.LIB_PATH], capture_output=True, text=True).stdout
    exported = sorted(set(re.findall(r" T (vol_[A-Za-z_0-9]+)$", out, re.M)))
    assert exported == names
    assert sorted(va.EXPORTED) == names


def test_controls_defaults_match_reference():
    c = va._Controls()
    va.lib().vol_controls_default(ctypes.byref(c))
    assert (c.solver, c.preconditioner) == (b"PCG", b"diagonal")
    assert (c.tolerance, c.relTol, c.maxIter, c.minIter, c.nSweeps) == (1e-8, 0.0, 1000, 0, 1)
    py = va.SolverControls()
    assert (py.solver, py.preconditioner, py.tolerance, py.relTol, py.maxIter, py.minIter, py.nSweeps) == \
        ("PCG", "diagonal", 1e-8, 0.0, 1000, 0, 1)
    assert (c.gamgPreSweeps, c.gamgPostSweeps, c.gamgOmega) == (py.gamgPreSweeps, py.gamgPostSweeps, py.gamgOmega)


def test_struct_layouts(tmp_path):
    """ctypes mirrors agree with the C header: size and every field offset, from gcc."""
    fields = {"vol_solver_controls": va._Controls, "vol_solver_performance": va._Perf}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "voluma_amd.h"', "int main(void) {"]
    for cname, cls in fields.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n") if l)
    for cname, cls in fields.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)
