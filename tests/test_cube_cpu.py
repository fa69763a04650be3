"""The cube reader's parallel number parser (host/cube_values.hpp) against a
sequential strtod parse (tests/cube_values_check.cpp)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_cube_values_match_strtod(tmp_path):
    exe = str(tmp_path / "cube_values_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-pthread", "-o", exe,
                           os.path.join(HERE, "cube_values_check.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
