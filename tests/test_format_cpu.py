"""The output writer's number formatting (host/format.hpp) produces exactly
the characters of glibc's printf "%6.3f" / "%8.4f" / "%d" that the
reference writes (simulation.cpp:183-194): a compiled checker compares the
two on special values, decimal ties and random doubles."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fixed_point_formatting_equals_printf(tmp_path):
    exe = str(tmp_path / "format_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "format_check.cpp")])
    r = subprocess.run([exe, "300000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
