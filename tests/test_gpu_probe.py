"""The VALU issue probe behind the kernels' issue-rate figures."""
import pytest

from xitils_amd import api

pytestmark = pytest.mark.gpu


def test_valu_rate_is_a_plausible_issue_rate():
    """Wave-level VALU instructions per second on 256 CUs at 2.4 GHz: more than
    one per CU per cycle (a wave64 fp32 instruction takes two cycles of a
    SIMD: the published 157.3 TFLOP/s is 128 FMA lanes per CU per cycle), no
    more than two, and repeatable."""
    dev = api.Device(0)
    r1, r2 = dev.valu_rate(), dev.valu_rate()
    print(f"[gpu] VALU issue: {r1 / 1e9:.1f} / {r2 / 1e9:.1f} G wave-instructions/s", flush=True)
    assert 256 * 2.4e9 < r1 <= 2 * 256 * 2.4e9
    assert abs(r1 - r2) < 0.1 * r1
