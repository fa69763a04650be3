"""CPU checks of the C ABI library: it builds, loads, and exports every entry
point include/*.h declares; without a GPU mafm_create reports NO_DEVICE
(no compute call is made here)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")


def _declared():
    names = set()
    for fn in os.listdir(INCLUDE):
        if fn.endswith(".h"):
            txt = open(os.path.join(INCLUDE, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(mafm_\w+)\s*\(", txt))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    from mechafm_amd import abi
    if not os.path.exists(abi.LIB_PATH):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "mechafm_amd", "csrc"), "-j8"],
                              stdout=subprocess.DEVNULL)
    return abi.lib()


def test_headers_declare_entry_points():
    names = _declared()
    assert "mafm_scan" in names and "mafm_sim_open" in names and len(names) >= 20


@pytest.mark.parametrize("name", _declared())
def test_symbol_exported(lib, name):
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_binding_covers_every_symbol():
    from mechafm_amd import abi
    assert set(_declared()) == set(abi.SIGNATURES)


def test_create_without_device(lib):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert lib.mafm_create(0, ctypes.byref(h)) == 6      # MAFM_ERR_NO_DEVICE
    assert not h.value


def test_null_arguments_rejected(lib):
    assert lib.mafm_scan(None, None, None) == 1           # MAFM_ERR_INVALID
    assert lib.mafm_sim_open(None, None, None, None) == 1
