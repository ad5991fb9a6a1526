"""The HIP wake path against the reference itself: the recorded per-step
state of the reference's own procedures (tests/golden/refprobe_rotor_wake.txt,
see tests/test_oracle_refprobe.py) replayed through the C ABI on the GPU."""
import pytest

from tests.test_oracle_refprobe import load_fixture, replay

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_wake_matches_reference_step_by_step(precision):
    """FP64: unique-segment sums in a different order than the reference's
    ring-form loops, so agreement is to round-off (1e-12 of scale).  Mixed
    precision: this wake has fewer than 1024 nodes, so the FP64 small-target
    kernel runs and the same bound holds."""
    from volcanor_amd import vlm
    ctx = vlm.Context(0)
    ctx.set_precision(precision)
    h, steps = load_fixture()
    rel = replay(ctx, h, steps)
    for k, v in rel.items():
        assert v < 1e-12, (k, v)
