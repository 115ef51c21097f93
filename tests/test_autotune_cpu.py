"""The GEMM-tiling autotuner's host logic: shape recording and the choice of
the limit from timings (no kernel runs here)."""

from ttnumpy_amd import autotune as AT
from ttnumpy_amd import device as D


def test_recording_counts_shapes_and_restores():
    with AT.recording() as rec:
        D.record_gemm(1, 100, 70, 8, False, True)
        D.record_gemm(1, 100, 70, 8, False, True, count=4)
        D.record_gemm(3, 5, 6, 7, True, True)
        with AT.recording() as inner:
            D.record_gemm(2, 2, 2, 2, True, False)
        assert inner == {(2, 2, 2, 2, True, False): 1}
    assert rec == {(1, 100, 70, 8, False, True): 5, (3, 5, 6, 7, True, True): 1}
    assert D._GEMM_RECORD[0] is None
    D.record_gemm(1, 1, 1, 1, True, True)     # no recorder: nothing kept


def test_tiles64():
    assert AT.tiles64((1, 64, 64, 9, True, True)) == 1
    assert AT.tiles64((2, 65, 64, 9, True, True)) == 4
    assert AT.tiles64((3, 1, 200, 9, True, True)) == 12


def test_choose_limit_splits_at_the_crossover():
    # 32 x 32 wins below 10 tiles, 64 x 64 from 10 on
    small = (1, 64 * 3, 64 * 3, 64, False, True)       # 9 tiles
    mid = (1, 64 * 5, 64 * 2, 64, False, True)         # 10 tiles
    big = (1, 64 * 20, 64 * 20, 64, False, True)       # 400 tiles
    times = {small: (1.0, 2.0), mid: (3.0, 2.5), big: (9.0, 5.0)}
    calls = {small: 10, mid: 1, big: 2}
    limit, ms = AT.choose_limit(times, calls)
    assert limit == 10
    assert ms == 10 * 1.0 + 2.5 + 2 * 5.0
    assert AT.predicted_ms(times, calls, AT.ALL_64) == 10 * 2.0 + 2.5 + 10.0
    assert AT.predicted_ms(times, calls, AT.ALL_32) == 10 * 1.0 + 3.0 + 18.0
    # all 64 x 64 best: limit 0
    assert AT.choose_limit({small: (2.0, 1.0)}, {small: 1})[0] == 0
