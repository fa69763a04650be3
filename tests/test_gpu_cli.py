"""The command-line drop-in for bin/mechafm-omp (mechafm.cpp:209-233):
`mechafm_amd/lib/mechafm INPUT [OUTDIR]` run as a subprocess on the
reference's own example job.  Every per-height file equals the one the
reference program writes (lines sorted: the reference writes them in thread
order); statistics.txt has the reference's lines with the exact step total;
the console carries the reference's warnings on stderr, its settings echo
and closing lines on stdout; a bad input ends with the reference's error
line and exit code 1."""
import gzip
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "mechafm_amd", "lib", "mechafm")
SRC = os.path.join(GOLDEN, "example_graphene")


def _job(tmp_path):
    job = tmp_path / "job"
    job.mkdir()
    for fn in ("input.scan", "graphene.xyz", "parameters.dat"):
        shutil.copy(os.path.join(SRC, fn), job / fn)
    return job


def _expected():
    files, name = {}, None
    with gzip.open(os.path.join(SRC, "outputs.txt.gz"), "rt") as f:
        for line in f.read().splitlines():
            if line.startswith("## "):
                name = line[3:]
                files[name] = []
            else:
                files[name].append(line)
    return files


@pytest.mark.parametrize("precision", ["plain64", "fast64"])
def test_cli_writes_the_reference_files(tmp_path, precision):
    job = _job(tmp_path)
    out = tmp_path / "out"
    out.mkdir()
    run = subprocess.run([CLI, "input.scan", str(out) + "/", "--precision", precision], cwd=job,
                         capture_output=True, text=True, timeout=240)
    assert run.returncode == 0, run.stderr
    expected = _expected()
    got = sorted(f for f in os.listdir(out) if f.startswith("scan-"))
    assert got == sorted(expected)
    for fn, lines in expected.items():
        assert sorted(open(out / fn).read().splitlines()) == lines, fn
    total = sum(int(l.split()[-1]) for ls in expected.values() for l in ls)
    stats = open(out / "statistics.txt").read().splitlines()
    assert stats[:5] == ["Simulation run finished", "Statistics:", "    Computed 52111 tip positions",
                         f"    Needed {total} minimization steps in total",
                         f"    Which means approximately {total / 52111:.2f} minimization steps per tip position"]
    assert stats[5].startswith("    The simulation wall time is ") and stats[5].endswith(" seconds")
    assert stats[6].startswith("    The entire simulation took ") and len(stats) == 7
    # console: the reference's warning block on stderr and settings echo on stdout
    # (its credits banner between them is not reproduced), then its run lines
    ref = open(os.path.join(SRC, "console.txt")).read()
    warn = ref.split("+ - - -")[0]
    echo = "+- \n+- Input settings" + ref.split("+- \n+- Input settings")[1]
    assert run.stderr == warn
    assert echo in run.stdout
    for line in ("+- 3D data grid is: 41 x 31 x 41 (52111 in total)", "+- Starting simulation",
                 "+- Simulation run finished", "+- Statistics:", "+-     Computed 52111 tip positions",
                 f"+-     Needed {total} minimization steps in total",
                 "+- How many x,y points did each process handle:", "+-     Process  0:   1271 x,y points"):
        assert line + "\n" in run.stdout, line
    assert run.stdout.index(echo) < run.stdout.index("+- Starting simulation") < run.stdout.index("+- Statistics:")
    # progress while the scan runs, at every tenth as the reference prints it
    # (simulation.cpp:131-137, "+- Finished %4.1f %% of the simulation")
    marks = [run.stdout.index(f"+- Finished {10.0 * q:4.1f} % of the simulation\n") for q in range(1, 11)]
    assert marks == sorted(marks) and run.stdout.index("+- Starting simulation") < marks[0]
    assert marks[-1] < run.stdout.index("+- Statistics:") and run.stdout.count("+- Finished") == 10


@pytest.mark.parametrize("mutate,message", [
    (None, "The file {job}/missing.scan does not exist!"),
    (("minimiser    FIRE", "minimiser    BFGS"), "Unrecognised minimiser type!"),
    (("graphene.xyz", "nothere.xyz"), "No such file: nothere.xyz!"),
])
def test_cli_bad_input_exits_with_the_reference_error(tmp_path, mutate, message):
    job = _job(tmp_path)
    name = "input.scan"
    if mutate is None:
        name = str(job / "missing.scan")
    else:
        text = open(job / "input.scan").read()
        assert mutate[0] in text
        open(job / "input.scan", "w").write(text.replace(mutate[0], mutate[1]))
    run = subprocess.run([CLI, name, str(tmp_path / "out") + "/"], cwd=job, capture_output=True, text=True,
                         timeout=120)
    assert run.returncode == 1
    assert run.stderr.endswith("+- ERROR (on process 0): " + message.format(job=job) + "\n"), run.stderr
    assert not os.path.exists(tmp_path / "out") or not os.listdir(tmp_path / "out")


def test_cli_command_line_errors():
    run = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert run.returncode == 1 and run.stderr == "+- ERROR (on process 0): Specify an input file to be read!\n"
    run = subprocess.run([CLI, "a", "b", "c"], capture_output=True, text=True, timeout=60)
    assert run.returncode == 1 and run.stderr == "+- ERROR (on process 0): Too many command line arguments!\n"
