"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its struct layout matches the Python binding (no GPU compute calls)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from exagoop_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "exagoop_amd.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z_0-9]+)\s*\(", txt)) - {"mpm_wall_udf_fn"})


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_option_numbers_match_header():
    txt = open(HEADER).read()
    defs = dict((k, int(v)) for k, v in re.findall(r"#define MPM_OPT_([A-Z_]+) (\d+)", txt))
    assert len(defs) >= 7
    for name, value in defs.items():
        assert getattr(capi, "OPT_" + name) == value, name


def test_specs_layout_matches_c():
    fields = [f[0] for f in capi.Specs._fields_]
    prog = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(){",
            'printf("%zu\\n", sizeof(mpm_specs));']
    for f in fields:
        prog.append(f'printf("%zu\\n", offsetof(mpm_specs, {f}));')
    prog.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(src, "w").write("\n".join(prog))
        subprocess.check_call(["gcc", src, "-o", exe])
        out = [int(x) for x in subprocess.check_output([exe]).split()]
    assert out[0] == ctypes.sizeof(capi.Specs)
    for f, off in zip(fields, out[1:]):
        assert getattr(capi.Specs, f).offset == off, f


def test_specs_default_matches_reference_defaults():
    lib = capi.load()
    s = capi.Specs()
    lib.mpm_specs_default(ctypes.byref(s))
    p = capi.default_specs()
    for f, _ in capi.Specs._fields_:
        a, b = getattr(s, f), getattr(p, f)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), f
        else:
            assert a == b, f
    assert s.alpha_pic_flip == 0.95 and s.mass_tolerance == 1e-8 and s.CFL == 0.1
    assert list(s.gravity) == [0.0, -9.81, 0.0] and s.num_redist == 10


@pytest.mark.parametrize("order,n,per,exp", [
    (2, (64, 64, 4), (0, 0, 1), [2, 2, 2]),
    (2, (64, 64, 4), (0, 0, 0), [2, 2, 1]),   # non-periodic n < 5 falls back to linear
    (3, (128, 64, 64), (0, 0, 0), [3, 3, 3]),
    (3, (2, 8, 8), (1, 0, 0), [1, 3, 3]),     # periodic n < 3 falls back
    (1, (8, 8, 8), (0, 0, 0), [1, 1, 1]),
])
def test_order_directional(order, n, per, exp):
    lib = capi.load()
    s = capi.default_specs()
    s.order_scheme = order
    s.ncells[:] = list(n)
    s.periodic[:] = list(per)
    out = (ctypes.c_int * 3)()
    assert lib.mpm_order_directional(ctypes.byref(s), out) == 0
    assert list(out) == exp


def test_bad_order_is_an_error_not_an_abort():
    lib = capi.load()
    s = capi.default_specs()
    s.order_scheme = 4
    out = (ctypes.c_int * 3)()
    assert lib.mpm_order_directional(ctypes.byref(s), out) == -3
