"""The C-ABI library loads (no GPU needed) and exports every entry point the
public header declares, with the Python binding covering all of them."""

import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "probit_amd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pgp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "pgp_laplace_fit" in names and "pgp_predict" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from probit_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.fail("libprobit_amd.so is not built (run python -m probit_amd.build)")
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_binding_covers_header():
    from probit_amd import _lib
    assert sorted(_lib.EXPORTED) == declared_functions()


def test_size_queries_without_gpu():
    from probit_amd import _lib
    lib = _lib.load()
    assert lib.pgp_workspace_size(100, 2, 0) > 128 * 128 * 16
    assert lib.pgp_workspace_size(100, 2, 1) < lib.pgp_workspace_size(100, 2, 0)
    assert lib.pgp_workspace_size(0, 2, 0) == -1
    assert lib.pgp_predict_workspace_size(100, 50, 0) > 0
    assert lib.pgp_batcher_table_size(16) > 0


def test_abi_rejects_bad_arguments_without_touching_the_gpu():
    import ctypes as C

    from probit_amd import _lib
    lib = _lib.load()
    assert lib.pgp_potrf(0, None, 64, 64, None, None, None, None) < 0
    assert lib.pgp_potrf(7, None, 64, 64, None, None, None, None) < 0
    assert lib.pgp_laplace_fit(None, None, None, None, None, None, 0, None) < 0
    assert b"NULL" in lib.pgp_last_error() or len(lib.pgp_last_error()) > 0
    h = C.c_void_p()
    assert lib.pgp_batcher_create(0, None, None, 0, None, C.byref(h)) < 0


def test_product_path_refuses_without_gpu(monkeypatch):
    import torch

    from probit_amd import _lib
    monkeypatch.setattr(torch.cuda, "is_available", lambda: False)
    with pytest.raises(RuntimeError, match="GPU"):
        _lib.lib()


def test_product_package_never_imports_the_oracle():
    pkg = ROOT / "probit_amd"
    for p in pkg.rglob("*.py"):
        txt = p.read_text()
        assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+(\w+)", txt, flags=re.M), p
