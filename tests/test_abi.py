"""The C-ABI library loads and exports every symbol include/eup.h declares (CPU-only checks)."""
import ctypes
import os
import re

import pytest

from euplectella_hpc_amd import lib as eup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    with open(os.path.join(ROOT, "include", "eup.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(eup_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(eup.EXPORTS)


def test_library_exports_every_symbol():
    L = eup.load()
    for name in header_functions():
        assert hasattr(L, name), name
    assert L.eup_abi_version() == 1


def test_struct_layout_matches_header():
    # eup_config: 13 x int32 (52 B) + pad to 8 + 3 doubles
    assert ctypes.sizeof(eup.EupConfig) == 80
    assert eup.EupConfig.omega.offset == 56


def test_create_rejects_bad_config_without_touching_the_gpu():
    L = eup.load()
    cfg = eup.EupConfig()
    cfg.l, cfg.m, cfg.n = 1, 4, 4
    cfg.omega = 1.0
    h = ctypes.c_void_p()
    rc = L.eup_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == -1
    assert b"must be >= 2" in L.eup_last_error(None)
    cfg.l = 8
    cfg.mode, cfg.storage = eup.MODE_PARITY, eup.STORE_F32
    cfg.front, cfg.rear = eup.FACE_OUTFLOW, eup.FACE_INFLOW
    assert L.eup_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"parity mode requires EUP_STORE_F64" in L.eup_last_error(None)
    cfg.storage = eup.STORE_F64
    cfg.up = 1
    assert L.eup_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"periodic" in L.eup_last_error(None)


def test_null_handle_calls_fail_cleanly():
    L = eup.load()
    assert L.eup_step(None, 1) == -1
    assert L.eup_sync(None) == -1
    L.eup_destroy(None)


def test_lattice_requires_a_device_on_cpu_hosts():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(eup.EupError) as ei:
        eup.Lattice(8, 8, 8, svisc=0.1, u_inflow=0.05)
    assert "EUP_EHIP" in str(ei.value)


def test_fortran_module_declares_every_entry_point():
    """euplectella_hpc_amd/fortran/eup_mod.F90 binds every function of include/eup.h."""
    with open(os.path.join(ROOT, "euplectella_hpc_amd", "fortran", "eup_mod.F90")) as fh:
        text = fh.read()
    bound = sorted(set(re.findall(r"bind\(c,\s*name='(eup_[a-z_]+)'\)", text)))
    assert bound == header_functions()
