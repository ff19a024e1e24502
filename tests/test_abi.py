"""C ABI checks that need no GPU: the library loads, exports every entry
point include/smili_hip.h declares, and rejects invalid descriptors before
its first HIP call (error statuses instead of the reference's silent NaNs or
Python -1 returns, lbfgs.py:186-201)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from smili_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "smili_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(smili_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, n
    assert lib.smili_abi_version() == 3


def test_struct_layouts_match_header_sizes():
    # pointer-size fields in declared order: 10 doubles + 5 pointers (reg)
    assert C.sizeof(_lib.SmiliReg) == 18 * 8
    assert C.sizeof(_lib.SmiliTune) == 12 * 4


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof / offsetof of every ABI struct as gcc lays out the header."""
    structs = {"smili_reg": _lib.SmiliReg, "smili_data": _lib.SmiliData, "smili_desc": _lib.SmiliDesc,
               "smili_movie_desc": _lib.SmiliMovieDesc, "smili_tune": _lib.SmiliTune}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "smili_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("  return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        cname, what, val = line.split()
        py = structs[cname]
        got = C.sizeof(py) if what == "size" else getattr(py, what).offset
        assert got == int(val), (cname, what)


def _desc(nx=4, ny=4, nuv=3, bad=None):
    d = _lib.SmiliDesc()
    keep = []

    def arr(a, t):
        a = np.ascontiguousarray(a, dtype=t)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int32 if t == np.int32 else C.c_double))

    d.nx, d.ny = nx, ny
    xi = np.tile(np.arange(1, nx + 1), ny)
    yi = np.repeat(np.arange(1, ny + 1), nx)
    if bad == "xidx":
        xi[3] = nx + 1
    d.npix = nx * ny
    d.xidx, d.yidx = arr(xi, np.int32), arr(yi, np.int32)
    d.nuv = nuv
    d.u, d.v = arr(np.zeros(nuv), np.float64), arr(np.zeros(nuv), np.float64)
    idx = np.array([1, -2, 3], np.int32)
    if bad == "uvidx":
        idx[1] = 7
    var = np.ones(3)
    if bad == "var":
        var[2] = 0.0
    d.data.namp = 3
    d.data.uvidxamp = arr(idx, np.int32)
    d.data.vamp = arr(np.ones(3), np.float64)
    d.data.varamp = arr(var, np.float64)
    d.data.wamp = 1.0
    if bad == "nodata":
        d.data.namp = 0
    if bad == "l1":
        d.reg.l1_l = 1.0
    return d, keep


@pytest.mark.parametrize("bad,msg", [("xidx", "xidx/yidx out of range"), ("uvidx", "uv index out of range"),
                                     ("var", "variance must be positive"), ("nodata", "no data"),
                                     ("l1", "l1_wgt missing")])
def test_invalid_descriptor_rejected(bad, msg):
    lib = _lib.load()
    d, keep = _desc(bad=bad)
    h = C.c_void_p()
    rc = lib.smili_problem_create(C.byref(d), None, C.byref(h))
    assert rc == -1
    assert msg in lib.smili_last_error().decode()
    assert not h.value


def test_null_arguments():
    lib = _lib.load()
    assert lib.smili_problem_create(None, None, None) == -1
    assert lib.smili_cost_grad(None, None, None, None, None) == -1
    assert lib.smili_synchronize(None) == -1
    lib.smili_problem_destroy(None)        # no-op


def test_host_wrapper_fails_loudly_without_library(tmp_path):
    with pytest.raises(OSError):
        _lib.load(str(tmp_path / "missing.so"))


def _movie(bad=None):
    d, keep = _desc(nuv=3)
    md = _lib.SmiliMovieDesc()
    md.frame = d
    md.nz = 2
    nuvs = np.array([2, 1], np.int32)
    if bad == "nuvs":
        nuvs[1] = 2
    keep.append(nuvs)
    md.nuvs = nuvs.ctypes.data_as(C.POINTER(C.c_int32))
    if bad == "nz":
        md.nz = 0
    if bad == "rt":
        md.rt_l = 1.0
    if bad == "lc":
        lc = np.array([1.0, 0.0])
        keep.append(lc)
        md.lc_tgtfd = lc.ctypes.data_as(C.POINTER(C.c_double))
    return md, keep


@pytest.mark.parametrize("bad,msg", [("nuvs", "sum of nuvs"), ("nz", "nz must be"), ("rt", "rt_wgt missing"),
                                     ("lc", "lc_tgtfd must be finite and nonzero")])
def test_invalid_movie_rejected(bad, msg):
    lib = _lib.load()
    md, keep = _movie(bad)
    h = C.c_void_p()
    assert lib.smili_movie_create(C.byref(md), None, C.byref(h)) == -1
    assert msg in lib.smili_last_error().decode()
    assert not h.value


def test_invalid_batch_member_named():
    lib = _lib.load()
    good, k1 = _desc()
    bad, k2 = _desc(bad="var")
    arr = (_lib.SmiliDesc * 2)(good, bad)
    h = C.c_void_p()
    assert lib.smili_batch_create(arr, 2, None, C.byref(h)) == -1
    assert "problem 1: " in lib.smili_last_error().decode()
    assert lib.smili_batch_create(arr, 0, None, C.byref(h)) == -1
    assert not h.value
