"""C-ABI boundary checks that need no GPU: the library loads, exports every entry point declared
in include/vib.h, and the host-only helpers (whitening, preintegration) match the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "vib.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vib_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(abi):
    lib = abi.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (vib_[a-z_0-9]+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    assert lib.vib_abi_version() == 1


def test_create_without_device_fails_cleanly(abi):
    import ctypes as C
    lib = abi.load_library()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("device present")
    except Exception:
        pass
    h = C.c_void_p()
    assert lib.vib_create(C.byref(h), 0) == 13  # VIB_ERR_NO_DEVICE


def test_footprint_is_positive(abi):
    from vibpy import synth
    w = synth.generate(K=5, L=100, seed=3)
    s = w.cstruct()
    import ctypes as C
    n = abi.load_library().vib_window_footprint(C.byref(s), 0)
    assert n > 8 * (w.F * 10 + w.L * 7)
