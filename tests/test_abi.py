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
    h = ctypes.c_void_p()
    rc = lib.mafm_create(0, ctypes.byref(h))
    if rc == 0:                                          # a GPU is present
        lib.mafm_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == 6                                       # MAFM_ERR_NO_DEVICE
    assert not h.value


def test_null_arguments_rejected(lib):
    assert lib.mafm_scan(None, None, None) == 1           # MAFM_ERR_INVALID
    assert lib.mafm_sim_open(None, None, None, None) == 1
    assert lib.mafm_scan_host(None, None, None, None) == 1
    assert lib.mafm_scan_host_query(None, None) == 1
    assert lib.mafm_scan_host_wait(None, None) == 1
    assert lib.mafm_valu_calibration(None, 0, 1, None, None) == 1


STRUCTS = [("mafm_site", "Site"), ("mafm_tip_dummy", "TipDummy"), ("mafm_system_desc", "SystemDesc"),
           ("mafm_grid_desc", "GridDesc"), ("mafm_scan_desc", "ScanDesc"), ("mafm_scan_stats", "ScanStats"),
           ("mafm_device_props", "DeviceProps"), ("mafm_potential_desc", "PotentialDesc"),
           ("mafm_adjoint_desc", "AdjointDesc"), ("mafm_adjoint_out", "AdjointOut"), ("mafm_sim_info", "SimInfo"),
           ("mafm_run_opts", "RunOpts"), ("mafm_deltaf_desc", "DeltaFDesc"),
           ("mafm_deltaf_weights_out", "DeltaFWeights")]


def test_struct_layouts_match_headers(tmp_path):
    """Every ctypes mirror has the C struct's size and field offsets (a C
    program compiled against include/ prints them); the record dtype too."""
    from mechafm_amd import abi
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "mafm.h"', '#include "mafm_sim.h"',
             "int main(void) {"]
    expect = []
    for cname, pname in STRUCTS:
        cls = getattr(abi, pname)
        lines.append(f'    printf("%zu\\n", sizeof({cname}));')
        expect.append(ctypes.sizeof(cls))
        for field, _ in cls._fields_:
            lines.append(f'    printf("%zu\\n", offsetof({cname}, {field}));')
            expect.append(getattr(cls, field).offset)
    lines.append('    printf("%zu\\n", sizeof(mafm_record));')
    expect.append(abi.RECORD_DTYPE.itemsize)
    for field in abi.RECORD_DTYPE.names:
        lines.append(f'    printf("%zu\\n", offsetof(mafm_record, {field}));')
        expect.append(abi.RECORD_DTYPE.fields[field][1])
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = str(tmp_path / "layout")
    subprocess.check_call(["gcc", "-std=c99", "-I", INCLUDE, "-o", exe, str(src)])
    got = [int(v) for v in subprocess.check_output([exe]).split()]
    assert got == expect
