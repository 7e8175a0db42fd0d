"""The C ABI library loads and exports every declared symbol; host-side
logic that needs no GPU (threshold of the sqrt predicate, parameters,
error paths)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from gpu_self_join_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "gdsjoin.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "defined", "sizeof"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    declared = _declared_functions()
    assert "GDSJoinPy" in declared and "copyResultIntoPythonArray" in declared
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == declared


def test_params_default_follow_params_h():
    p = _lib.default_params()
    # params.h:40-60 as shipped
    assert (p.ilp, p.shortcircuit, p.reorder, p.stamp, p.queryreorder) == (8, 1, 1, 0, 1)
    assert p.k == 0  # resolved to min(dims, 6) at join time


@pytest.mark.parametrize("eps", [1.0, 0.5, 0.1, 3.7, 1e-3, 123.456, 2.0 ** -20])
def test_eps_threshold_f32(eps):
    L = _lib.load()
    e = np.float32(eps)
    t = np.float32(L.gdsj_eps_threshold_f32(C.c_float(e)))
    assert np.sqrt(t) <= e
    assert np.sqrt(np.nextafter(t, np.float32(np.inf))) > e
    # every float around eps^2: (x <= t) == (sqrt(x) <= eps)
    base = np.array([t], np.float32).view(np.int32)[0]
    xs = (base + np.arange(-4096, 4097, dtype=np.int32)).view(np.float32)
    assert np.array_equal(xs <= t, np.sqrt(xs) <= e)


@pytest.mark.parametrize("eps", [1.0, 0.5, 0.1, 3.7, 123.456, 7.25e-5])
def test_eps_threshold_f64(eps):
    L = _lib.load()
    t = L.gdsj_eps_threshold_f64(C.c_double(eps))
    assert np.sqrt(t) <= eps
    assert np.sqrt(np.nextafter(t, np.inf)) > eps
    base = np.array([t], np.float64).view(np.int64)[0]
    xs = (base + np.arange(-2048, 2049, dtype=np.int64)).view(np.float64)
    assert np.array_equal(xs <= t, np.sqrt(xs) <= eps)


def test_unknown_param_field_rejected():
    with pytest.raises(TypeError):
        _lib.default_params(no_such_field=1)


def test_gdsjoin_argument_errors():
    from gpu_self_join_amd import gdsjoin
    with pytest.raises(ValueError):
        gdsjoin([1.0, 2.0, 3.0], 1.0, 2, "float")
    with pytest.raises(ValueError):
        gdsjoin([1.0, 2.0], 1.0, 2, "half")
