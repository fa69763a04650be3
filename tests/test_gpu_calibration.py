"""mafm_valu_calibration through the C ABI: each kind of calibration chain
runs and reports its known VALU wave-instruction count (32 768 wave64s, 8
chains per lane, one instruction per step for the FMA chains and two for
v_rsq + add), and a kind outside 0..2 is refused."""
import pytest

from mechafm_amd import abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,per_step", [(0, 1), (1, 1), (2, 2)])
def test_calibration_chains(gpu_ctx, kind, per_step):
    r = gpu_ctx.valu_calibration(kind, 64)
    assert r["wave_instructions"] == 32768 * 8 * 64 * per_step
    assert r["ms"] > 0


def test_calibration_rejects_bad_arguments(gpu_ctx):
    with pytest.raises(abi.MafmError, match="calibration kind"):
        gpu_ctx.valu_calibration(3)
    with pytest.raises(abi.MafmError, match="iterations"):
        gpu_ctx.valu_calibration(0, 0)
