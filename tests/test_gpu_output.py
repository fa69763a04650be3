"""Output written while the batch scans (mafm_batch_run_output).

The 64 molecules of the batched workload (2 x 2 pixels each, all 40
heights) are scanned in groups while a writer thread writes the finished
groups' per-height files.  Every file, and every record the jobs then serve
from the host, equals what a plain batch run followed by mafm_sim_write_output
gives; statistics.txt counts each job's own iterations."""
import gzip
import os
import re

import numpy as np
import pytest

from conftest import materialize_scale
from mechafm_amd import abi

pytestmark = pytest.mark.gpu


def _files(folder):
    out = {}
    for fn in sorted(os.listdir(folder)):
        if fn.startswith("scan-"):
            raw = open(os.path.join(folder, fn), "rb").read()
            out[fn.replace(".gz", "")] = gzip.decompress(raw) if fn.endswith(".gz") else raw
    return out


def _jobs(ctx, inputs, root, gzip_on):
    sims = []
    for m, inp in enumerate(inputs):
        text = open(inp).read()
        text = re.sub(r"^gzip\s+\w+", "gzip on" if gzip_on else "gzip off", text, flags=re.M)
        open(inp, "w").write(text)
        sims.append(abi.Simulation(inp, ctx, outdir=os.path.join(root, f"mol{m:02d}")))
    return sims


@pytest.mark.parametrize("gzip_on", [False, True])
def test_streamed_output_equals_write_after_scan(gpu_ctx, tmp_path, gzip_on):
    inputs = materialize_scale("c3_batch", str(tmp_path / "in"))
    ref_sims = _jobs(gpu_ctx, inputs, str(tmp_path / "ref"), gzip_on)
    abi.batch_run(ref_sims, "fast64", options=abi.SCAN_REPRODUCIBLE)
    ref_recs = [s.records() for s in ref_sims]
    for s in ref_sims:
        s.write_output(threads=4)
    sims = _jobs(gpu_ctx, inputs, str(tmp_path / "out"), gzip_on)
    st = abi.batch_run_output(sims, "fast64", groups=5, threads=4, options=abi.SCAN_REPRODUCIBLE)
    assert st["positions"] == sum(len(r) for r in ref_recs)
    assert st["total_steps"] == sum(int(r["n_steps"].sum()) for r in ref_recs)
    for m, (s, rr) in enumerate(zip(sims, ref_recs)):
        np.testing.assert_array_equal(s.records(), rr)
        a, b = _files(str(tmp_path / "out" / f"mol{m:02d}")), _files(str(tmp_path / "ref" / f"mol{m:02d}"))
        assert len(a) == 40 and a == b
        stats = open(tmp_path / "out" / f"mol{m:02d}" / "statistics.txt").read().splitlines()
        assert stats[3] == f"    Needed {int(rr['n_steps'].sum())} minimization steps in total"
    # the files can be written again from the host records
    sims[1].write_output(threads=2)
    assert _files(str(tmp_path / "out" / "mol01")) == _files(str(tmp_path / "ref" / "mol01"))
    # the records now live on the host: device-side post-processing asks for another run
    with pytest.raises(abi.MafmError, match="on the host only"):
        sims[0].write_deltaf()


def test_streamed_output_reports_a_bad_folder(gpu_ctx, tmp_path):
    inputs = materialize_scale("c3_batch", str(tmp_path / "in"))[:3]
    sims = _jobs(gpu_ctx, inputs, str(tmp_path / "out"), False)
    os.chmod(tmp_path / "out" / "mol01", 0o500)
    try:
        if os.access(tmp_path / "out" / "mol01", os.W_OK):
            pytest.skip("running with permissions that ignore the read-only folder")
        with pytest.raises(abi.MafmError, match="cannot open output"):
            abi.batch_run_output(sims, "fast64", groups=3)
    finally:
        os.chmod(tmp_path / "out" / "mol01", 0o700)
