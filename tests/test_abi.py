"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(dpf_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("dpf_topology_create", "dpf_steady_forward", "dpf_steady_pullback",
              "dpf_residuals", "dpf_hm_forward", "dpf_hm_pullback"):
        assert n in names


def test_library_exports_every_declared_symbol(pkg):
    from dpfehm_jl_amd import _lib
    L = _lib.lib()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    assert not (declared_symbols() - set(_lib.SIGNATURES))


def test_opts_defaults(pkg):
    from dpfehm_jl_amd import _lib
    o = _lib.default_opts()
    assert o.reltol == 1e-10 and o.maxiter == 2000 and o.precision == _lib.DPF_FP64
    assert o.inner_reltol == 1e-5 and o.cheb_degree == 0 and o.memory == _lib.DPF_HOST
    assert ctypes.sizeof(_lib.Opts) == 56


def test_unknown_kwarg_rejected_before_gpu(pkg):
    from dpfehm_jl_amd import api
    import pytest
    with pytest.raises(TypeError):
        api._opts({"reltol": 1e-8, "bogus": 1})
    o = api._opts({"reltol": 1e-8, "linear_solver": None, "precision": "mixed"})
    assert o.reltol == 1e-8 and o.precision == 1
