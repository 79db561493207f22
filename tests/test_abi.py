"""The C ABI library (libaimdhf.so) loads and exports every entry point that include/aimdhf.h declares."""

import ctypes
import os
import re

import pytest

from tests.conftest import PKG, ROOT


def _declared():
    hdr = open(os.path.join(ROOT, "include", "aimdhf.h")).read()
    return sorted(set(re.findall(r"\bint\s+(aimdhf_\w+)\s*\(", hdr)))


def test_header_declares_the_boundary():
    names = _declared()
    for n in ("aimdhf_plan_create", "aimdhf_set_geometry", "aimdhf_jk", "aimdhf_one_electron",
              "aimdhf_grad_one", "aimdhf_grad_two", "aimdhf_plan_destroy", "aimdhf_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from aimd_amd import _lib
    L = _lib.lib()
    for n in _declared():
        assert hasattr(L, n), n
    assert set(_declared()) == set(_lib.SIGNATURES)
    assert L.aimdhf_lmax() == 4


def test_last_error_is_callable_without_a_device():
    from aimd_amd import _lib
    buf = ctypes.create_string_buffer(64)
    assert _lib.lib().aimdhf_last_error(buf, 64) == 0


def test_null_arguments_are_rejected_before_any_device_call():
    from aimd_amd import _lib
    L = _lib.lib()
    assert L.aimdhf_plan_create(0, 0, None, None, None, None, None, None, None, None) == _lib.AIMDHF_E_ARG
    out = ctypes.c_void_p()
    assert L.aimdhf_plan_create(0, 1, None, None, None, None, None, None, None, ctypes.byref(out)) == _lib.AIMDHF_E_ARG
    assert L.aimdhf_jk(None, 1, None, None, None, None) == _lib.AIMDHF_E_ARG
    assert L.aimdhf_set_geometry(None, None, 1e-12, None) == _lib.AIMDHF_E_ARG
    with pytest.raises(ValueError):
        _lib.check(_lib.AIMDHF_E_ARG)
    with pytest.raises(MemoryError):
        _lib.check(_lib.AIMDHF_E_NOMEM)


def test_product_fails_loudly_without_a_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from aimd_amd.qc.basis import build_basis
    from aimd_amd.qc.engine import Engine
    b = build_basis(["H", "H"], [[0, 0, 0], [0, 0, 1.4]], "sto-3g")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        Engine(b)


def test_profile_kind_names_cover_every_kind():
    from aimd_amd.qc.engine import Engine
    assert len(Engine.KIND_NAMES) == Engine.NKIND
    assert len(set(Engine.KIND_NAMES)) == Engine.NKIND
    assert not set(Engine.FOCK_KINDS) & set(Engine.GRAD_KINDS)
