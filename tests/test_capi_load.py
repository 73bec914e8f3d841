"""The C-ABI library loads without a GPU and exports every entry point that
include/wetcloth_amd.h declares (no compute calls are made here)."""
import ctypes
import os
import subprocess

import pytest


def test_header_declares_entry_points(wc):
    syms = wc.declared_symbols()
    for must in ("wc_create", "wc_destroy", "wc_step", "wc_pcg_solve_csr",
                 "wc_mic0_factor_csr", "wc_mic0_apply_csr", "wc_spmv_csr",
                 "wc_set_particles", "wc_get_particles", "wc_project"):
        assert must in syms
    assert len(syms) >= 20


def test_library_exports_every_declared_symbol(wc):
    lib = wc.lib()
    missing = [s for s in wc.declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", wc.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(wc.declared_symbols()) <= exported


def test_library_has_gfx950_code_object(wc):
    data = open(wc.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_errors_without_device_are_reported(wc):
    if os.environ.get("HIP_VISIBLE_DEVICES") is None:
        try:
            import torch
            if torch.cuda.is_available():
                pytest.skip("a GPU is present")
        except Exception:
            pass
    with pytest.raises(wc.WetclothError):
        wc.Scene((8, 8, 8), 1.0 / 8)
