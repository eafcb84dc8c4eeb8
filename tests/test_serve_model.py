"""The serving-loop queueing model against cases whose latencies follow by hand."""
import pytest

from morphogenesis_amd import serve_model as sm

C = 0.002  # seconds of GPU time per key


def test_a_request_alone_takes_its_keys_times_the_key_cost():
    for steps in (1, 4, 8):
        lat = sm.simulate([(1.0, 48)], cap=96, key_s=C, steps=steps)
        assert lat.tolist() == pytest.approx([48 * C])


def test_requests_that_fit_share_every_step():
    lat = sm.simulate([(0.0, 30), (0.0, 60)], cap=96, key_s=C)
    assert lat.tolist() == pytest.approx([90 * C, 90 * C])


def test_the_cap_queues_the_next_request_fifo():
    # 60 + 60 > 96: the second starts when the first finishes, and the small
    # request behind it waits too (head of line), then shares its steps.
    lat = sm.simulate([(0.0, 60), (0.0, 60), (0.0, 3)], cap=96, key_s=C)
    assert lat.tolist() == pytest.approx([60 * C, 60 * C + 63 * C, 60 * C + 63 * C])


def test_a_later_arrival_joins_at_the_next_step():
    # 40 keys from t = 0, steps of 10 C alone; a 3-key request at t = 5 C joins
    # at the second step's start (10 C). Three steps of 43 C / 4 are shared,
    # then the 3-key request's last step runs alone (3 C / 4).
    lat = sm.simulate([(0.0, 40), (5 * C, 3)], cap=96, key_s=C)
    t1 = 10 * C + 3 * 43 * C / 4
    t2 = t1 + 3 * C / 4 - 5 * C
    assert lat.tolist() == pytest.approx([t1, t2])


def test_oversized_requests_are_refused():
    with pytest.raises(ValueError):
        sm.simulate([(0.0, 97)], cap=96, key_s=C)


def test_latency_grows_with_key_cost_and_load():
    reqs = sm.arrivals(5.0, 300.0, seed=1)
    assert all(3 <= k <= 96 and k % 3 == 0 for _, k in reqs)
    p = [sm.percentiles(sm.simulate(reqs, 96, c))["p99_ms"] for c in (1e-3, 2e-3, 3e-3)]
    assert p[0] < p[1] < p[2]


def test_policies_differ_only_when_the_head_is_blocked():
    # 80 keys hold the cap; a 30-key request waits; a 3-key one behind it fits.
    reqs = [(0.0, 80), (0.0, 30), (0.0, 3)]
    fifo = sm.simulate(reqs, 96, C, policy="fifo")
    skip = sm.simulate(reqs, 96, C, policy="skip")
    small = sm.simulate(reqs, 96, C, policy="smallest")
    # FIFO: 80 alone, then 30 + 3 together.
    assert fifo.tolist() == pytest.approx([80 * C, 80 * C + 33 * C, 80 * C + 33 * C])
    # Skip runs the 3-key request beside the 80 keys, then the 30.
    assert skip.tolist() == pytest.approx([83 * C, 83 * C, 83 * C + 30 * C])
    # Smallest first takes 3 + 30, then the 80 keys alone.
    assert small.tolist() == pytest.approx([33 * C, 33 * C, 33 * C + 80 * C])
    # Without a blocked head all three agree.
    free = [(0.0, 10), (0.001, 20), (0.5, 5)]
    ref = sm.simulate(free, 96, C)
    for policy in sm.POLICIES:
        assert sm.simulate(free, 96, C, policy=policy).tolist() == pytest.approx(ref.tolist())
    with pytest.raises(ValueError):
        sm.simulate(free, 96, C, policy="lifo")
