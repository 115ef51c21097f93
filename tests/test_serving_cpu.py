"""Serving stream and queue model (CPU)."""

import numpy as np

from ttnumpy_amd.serving import ServingConfig, fifo_latencies, request_stream


def test_request_stream_is_seeded_and_in_range():
    cfg = ServingConfig(n_requests=2000, rate=50.0, seed=3)
    a, b = request_stream(cfg), request_stream(cfg)
    assert [(x.level, x.theta, x.offset) for x in a] == [(x.level, x.theta, x.offset) for x in b]
    offs = np.array([x.offset for x in a])
    assert offs[0] == 0.0 and np.all(np.diff(offs) >= 0.0)
    # exponential gaps at rate 50: mean gap 0.02 s within a few standard errors
    gaps = np.diff(offs)
    assert abs(gaps.mean() - 0.02) < 4 * 0.02 / np.sqrt(len(gaps))
    assert {x.level for x in a} == set(cfg.levels)
    th = np.array([x.theta for x in a])
    assert th.min() >= 0.0 and th.max() < cfg.theta_max


def test_fifo_latencies_closed_form():
    # arrivals 0, 1, 1.5, 10; service 2, 2, 2, 1: starts 0, 2, 4, 10
    finish, lat = fifo_latencies([0.0, 1.0, 1.5, 10.0], [2.0, 2.0, 2.0, 1.0])
    assert finish == [2.0, 4.0, 6.0, 11.0]
    assert lat == [2.0, 3.0, 4.5, 1.0]
    # an idle server: latency equals service time
    f2, l2 = fifo_latencies([0.0, 5.0, 9.0], [1.0, 1.0, 1.0])
    assert l2 == [1.0, 1.0, 1.0] and f2 == [1.0, 6.0, 10.0]
