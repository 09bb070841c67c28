"""The bench's stream plan (CPU): the warm-up spans two fills of the slots by default, the timed steps and the tail
keep every slot busy through the last timed completion, and the 32x16-patch HMC workload defaults to 256 slots."""
import sys

import pytest

sys.path.insert(0, __import__("os").path.dirname(__import__("os").path.dirname(__file__)))
import bench  # noqa: E402


def plan(argv):
    old = sys.argv
    sys.argv = ["bench.py"] + argv
    try:
        a = bench.parse()
    finally:
        sys.argv = old
    return a, bench.stream_plan(a)


def test_default_plan_spans_two_fills():
    a, (W, K, E, Ew, tail, total) = plan([])
    assert (W, K, E) == (5, 20, 96)
    assert W * Ew >= 2 * a.slots and Ew == 410
    assert tail == a.slots
    assert total == W * Ew + K * E + tail == 4994


@pytest.mark.parametrize("fills", [1.0, 1.5, 2.0, 3.0])
def test_warmup_fills(fills):
    a, (W, K, E, Ew, tail, total) = plan(["--warmup-fills", str(fills), "--warmup", "3", "--slots", "100"])
    assert W * Ew >= fills * 100 and W * (Ew - 1) < fills * 100 or Ew == E


def test_c3_defaults_to_256_slots():
    a, (W, K, E, Ew, tail, total) = plan(["--workload", "c3", "--warmup", "1", "--steps", "1", "--evals", "64",
                                          "--tail", "0", "--warmup-fills", "1"])
    assert a.slots == 256 and total == 256 + 64
