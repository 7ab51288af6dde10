"""CPU checks of the C-ABI library: it loads, exports every declared symbol, and the
ctypes struct mirrors match the C layout.  No compute calls (no GPU needed)."""
import ctypes as C
import os
import re

from cudasirecon_amd import engine
from cudasirecon_amd._abi import SimrFit, SimrParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(inc, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{}]*\)\s*;", src):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof"):
                continue
            names.add(name)
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = engine.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match():
    L = engine.lib()
    assert L.simr_params_size() == C.sizeof(SimrParams)
    assert L.simr_fit_size() == C.sizeof(SimrFit)


def test_defaults_match_python_mirror():
    from cudasirecon_amd._abi import default_params as py_default

    a = engine.default_params().as_dict()
    b = py_default().as_dict()
    for k in a:
        if isinstance(a[k], list):
            assert all(abs(x - y) < 1e-6 for x, y in zip(a[k], b[k])), k
        else:
            assert abs(a[k] - b[k]) < 1e-6, (k, a[k], b[k])


def test_footprint_and_errors_without_gpu():
    from cudasirecon_amd._abi import set_params
    from cudasirecon_amd.synth import make_otf_table

    p = engine.default_params()
    _, og = make_otf_table(3, 1.42, 1.515, 515, 0.08)
    set_params(p, nx=512, ny=512, nz=48, dxy=0.08, dz=0.125, wave=515, na=1.42, nimm=1.515, bRadAvgOTF=1, **og)
    fb = engine.footprint(p)
    # bands 758 MB + assembly stages ~2.1 GB + output 201 MB + raw 755 MB
    assert 3.0e9 < fb < 6.0e9
    q = p.copy()
    q.nphases = 4
    assert engine.footprint(q) == 0
    assert b"nphases" in engine.lib().simr_last_error()
    q = p.copy()
    q.nx = 500 + 7  # not 2^a 3^b 5^c
    assert engine.footprint(q) == 0
