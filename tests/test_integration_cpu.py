"""The maintainer patch (oracle/integration.patch: a `lower()` / `lowerFlex()`
on each reference Interaction subclass and Simulation::run handing its pixel
loop -- rigid or flexible -- to libmafm) applied to the reference sources,
compiled and linked against libmafm.so by the oracle/Makefile recipe.  Built only, never
run: this checks that the patch applies to the reference as it is and that
the C ABI it calls links."""
import os
import subprocess

import pytest

from conftest import ROOT

REF = "/root/reference/src"
LIB = os.path.join(ROOT, "mechafm_amd", "lib", "libmafm.so")


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference sources are not in this container")
def test_patched_reference_compiles_and_links():
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "mechafm_amd", "csrc"), "-j8"],
                              stdout=subprocess.DEVNULL)
    subprocess.check_call(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "integration"],
                          stdout=subprocess.DEVNULL)
    exe = os.path.join(ROOT, "oracle", "_ref", "mechafm-integrated")
    assert os.path.exists(exe)
    syms = subprocess.check_output(["nm", "-C", exe], text=True)
    assert "Simulation::runOnDevice()" in syms
    assert "Simulation::runFlexibleOnDevice()" in syms
    for name in ("mafm_create", "mafm_set_systems", "mafm_scan", "mafm_records", "mafm_destroy",
                 "mafm_scan_flexible", "mafm_flex_state"):
        assert f"U {name}" in syms, name
    deps = subprocess.check_output(["ldd", exe], text=True)
    assert "libmafm.so" in deps and "not found" not in deps


def test_patch_touches_only_declared_files():
    """The patch adds one header and edits interactions.hpp, simulation.hpp
    and simulation.cpp; nothing else of the reference changes."""
    text = open(os.path.join(ROOT, "oracle", "integration.patch")).read()
    files = sorted({l.split()[1].split("/", 1)[1] for l in text.splitlines() if l.startswith("+++ ")})
    assert files == ["src/interactions.hpp", "src/mafm_lowering.hpp", "src/simulation.cpp", "src/simulation.hpp"]
