"""The C ABI: every function include/sphx.h declares is exported by the HIP
library (loaded without a GPU) and by the oracle; the product loader binds only
the HIP build."""
import os
import re

import pytest

import sphinxsys_amd
from sphinxsys_amd import _capi
from tests._util import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sphx.h")).read()
    decl = r"^\s*(?:const\s+char\s*\*|sphx_status)\s*(sphx_[a-z0-9_]+)\s*\("
    return sorted(set(re.findall(decl, text, flags=re.M)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_capi.SYMBOLS)


def test_hip_library_exports_every_declared_symbol():
    assert os.path.exists(sphinxsys_amd.LIBRARY_PATH), "run __graft_entry__.build() first"
    lib = sphinxsys_amd.load()
    assert lib.backend == "hip-gfx950"
    for name in declared_symbols():
        assert hasattr(lib.lib, name), name


def test_oracle_exports_every_declared_symbol(oracle):
    for name in declared_symbols():
        assert hasattr(oracle.lib, name), name
    assert oracle.backend == "oracle-cpu"


def test_loader_refuses_missing_library(monkeypatch, tmp_path):
    monkeypatch.setattr(sphinxsys_amd, "LIBRARY_PATH", str(tmp_path / "libsphx.so"))
    monkeypatch.setattr(sphinxsys_amd, "_lib", None)
    with pytest.raises(sphinxsys_amd.SphxError):
        sphinxsys_amd.load()


def test_hip_context_without_gpu_reports_error():
    """No GPU in this container: creating a context must fail with a status,
    never fall back to the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = sphinxsys_amd.load()
    import ctypes as C
    h = C.c_void_p()
    status = lib.lib.sphx_context_create(0, C.byref(h))
    assert status != 0
