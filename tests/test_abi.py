"""The C-ABI library loads, exports exactly what include/dyn_nb.h declares,
and reports errors as status codes (no GPU compute calls here)."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "dyn_nb.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dyn_[a-z0-9_]+)\s*\(", txt)))


def test_library_loads_and_exports_every_declared_symbol():
    from dynamics_amd import api
    syms = header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(api.lib, s), s
    assert sorted(api.EXPORTED) == syms


def test_version_and_device_count_never_fail():
    from dynamics_amd import api
    assert api.lib.dyn_version() >= 100
    assert api.device_count() >= 0


def test_grid_size_rule():
    from dynamics_amd import api
    assert api.pme_grid_n((39.4, 39.4, 39.4), 1.0) == (40, 40, 40)
    assert api.pme_grid_n((64.0, 64.0, 64.0), 1.0) == (64, 64, 64)
    assert api.pme_grid_n((61.0, 7.0, 13.1), 1.0) == (64, 8, 15)
    with pytest.raises(api.DynError):
        api.pme_grid_n((10.0, 10.0, 10.0), 0.0)


def test_errors_are_status_codes_not_aborts():
    from dynamics_amd import api
    h = C.c_void_p()
    # null output pointer, bad device: error codes with a message
    assert api.lib.dyn_nb_create(None, 0, 1, 1, 10.0, 0.3, None) == api.DYN_ERR_ARG
    assert b"null" in api.lib.dyn_global_last_error()
    assert api.lib.dyn_nb_create(C.byref(h), 99, 1, 1, 10.0, 0.3, None) != api.DYN_OK
    assert api.lib.dyn_nb_create(C.byref(h), 0, 1, 1, -1.0, 0.3, None) != api.DYN_OK
    # null handles
    assert api.lib.dyn_nb_compute(None, None, None, None, None, None,
                                  np.ones(3, np.float32), 0, 0, 0.0, None, None, None, None,
                                  None) == api.DYN_ERR_ARG
    assert api.lib.dyn_engine_compute(None, 7) == api.DYN_ERR_ARG
    api.lib.dyn_nb_destroy(None)
    api.lib.dyn_engine_destroy(None)
    api.lib.dyn_pme_destroy(None)
