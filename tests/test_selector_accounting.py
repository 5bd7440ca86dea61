"""The selector trainer's switching-aware accounting (the reference's
loss_ratio_real, model/1-test_dtree.py:266-287 and 312-313) and the replay
of the device's switching-aware choice, on hand-sized sequences (CPU only)."""
import numpy as np

import train_selector as ts


def test_real_loss_ratio_hand_example():
    # one sequence of 3 calls; slot 0 is a column slot, 5 row SpMSpV, 7 SpMV
    times = np.full((3, 8), 100.0)
    single = np.full((3, 8), 100.0)
    times[:, 0], single[:, 0] = [10, 30, 10], [9, 28, 9]
    times[:, 5], single[:, 5] = [20, 12, 20], [15, 6, 15]
    seq = np.zeros(3, np.int64)
    # oracle per call: 0, 5, 0 (plain 10 + 12 + 10 = 32); charged: c_single
    # first, then switches: 9 + 12 + 10 = 31
    # choice 0,0,0: 9 + 28 + 9 = 46 -> (46 - 31) / 32
    r = ts.real_loss_ratio(times, single, seq, np.array([0, 0, 0]))
    assert abs(r - (46 - 31) / 32) < 1e-12
    # choice 0,5,0 (the oracle's): ratio 0
    assert abs(ts.real_loss_ratio(times, single, seq, np.array([0, 5, 0]))) < 1e-12
    # two sequences average their ratios
    t2 = np.vstack([times, times])
    s2 = np.vstack([single, single])
    sq = np.array([0, 0, 0, 1, 1, 1])
    r2 = ts.real_loss_ratio(t2, s2, sq, np.array([0, 0, 0, 0, 5, 0]))
    assert abs(r2 - ((46 - 31) / 32 + 0.0) / 2) < 1e-12


def test_switching_choices_keep_family_within_penalty():
    z = np.zeros((3, 8), np.float32)
    z[0, 5] = 3.0          # row slot first
    z[1, 1] = 2.0          # column slot on top, row slot 4 within 0.5
    z[1, 4] = 1.6
    z[2, 7] = 5.0          # SpMV far ahead
    z[2, 5] = 1.0
    seq = np.zeros(3, np.int64)
    np.testing.assert_array_equal(ts.switching_choices(z, seq, 0.0), [5, 1, 7])
    np.testing.assert_array_equal(ts.switching_choices(z, seq, 0.5), [5, 4, 7])
    # a new sequence forgets the previous call
    np.testing.assert_array_equal(ts.switching_choices(z, np.array([0, 1, 1]), 0.5), [5, 1, 7])
