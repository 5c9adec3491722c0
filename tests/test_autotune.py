"""Host logic of the leaf-size autotuner (haccsr.autotune.tune_leaf_size)
on synthetic step times: the search extends past an edge until the fastest
leaf size is bracketed, times the midpoints next to it, and picks the
smallest leaf size within the tie band of the fastest."""
import pytest

from haccsr import autotune


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


@pytest.fixture
def clock(monkeypatch):
    c = FakeClock()
    monkeypatch.setattr(autotune.time, "perf_counter", c)
    return c


def stepper(clock, cost):
    seen = []

    def run_step(nd):
        seen.append(nd)
        clock.t += cost(nd)
    return run_step, seen


def test_minimum_past_the_last_candidate_is_bracketed(clock):
    # a V with its minimum at 1800: beyond the default candidates
    run, seen = stepper(clock, lambda nd: 1.0 + abs(nd - 1800) * 1e-3)
    best, times = autotune.tune_leaf_size(run, log=None)
    order = sorted(times)
    k = order.index(best)
    assert 0 < k < len(order) - 1, "the pick must have a slower neighbour on each side"
    assert max(times) > 1024
    assert abs(best - 1800) <= 300


def test_minimum_before_the_first_candidate_is_bracketed(clock):
    run, seen = stepper(clock, lambda nd: 1.0 + abs(nd - 60) * 1e-3)
    best, times = autotune.tune_leaf_size(run, log=None)
    assert min(times) < 128 and best < 128


def test_flat_minimum_picks_the_smallest_tied_size(clock):
    # 640, 768 and 896 within 0.2% of each other; 896 the fastest (640 and
    # 896 are the refinement midpoints around the fastest candidate, 768)
    cost = {128: 1.30, 192: 1.20, 256: 1.10, 384: 1.05, 512: 1.02, 640: 1.0015,
            768: 1.001, 896: 1.000, 1024: 1.01}
    run, seen = stepper(clock, lambda nd: cost.get(nd, 1.02))
    best, times = autotune.tune_leaf_size(run, log=None)
    assert 640 in times and 896 in times  # the midpoints next to 768
    assert best == 640


def test_midpoints_next_to_the_fastest_are_timed(clock):
    run, seen = stepper(clock, lambda nd: 1.0 + abs(nd - 450) * 1e-3)
    best, times = autotune.tune_leaf_size(run, log=None)
    assert (384 + 512) // 2 in times
    assert best in (448, 512, 384)
