"""fitPeaks' job construction (PeakFit.c:321-351) on the host: sorted
peaks, ROIs merged while the next one starts at or before the current end,
clipped to the profile; jobs above the kernel's peak cap split."""
from midas_amd import _lib
from midas_amd.peakfit import fit_jobs


def test_disjoint_rois_stay_apart():
    assert fit_jobs(200, [50, 150], 20) == [(30, 70, [0]), (130, 170, [1])]


def test_overlap_merges_and_end_moves():
    # sorted order 10, 12, 14: all start before the running end
    jobs = fit_jobs(100, [10, 14, 60, 30, 12], 5)
    assert jobs == [(5, 19, [0, 4, 1]), (25, 35, [3]), (55, 65, [2])]


def test_touching_start_equal_to_end_merges():
    # second ROI starts exactly at the first ROI's end (start <= end)
    assert fit_jobs(100, [20, 30], 5) == [(15, 35, [0, 1])]
    assert fit_jobs(100, [20, 31], 5) == [(15, 25, [0]), (26, 36, [1])]


def test_clipped_to_the_profile():
    assert fit_jobs(50, [2, 48], 10) == [(0, 12, [0]), (38, 49, [1])]


def test_long_chains_split_at_the_cap():
    cap = _lib.MC_FIT_MAX_PEAKS
    bins = list(range(10, 10 + 3 * (2 * cap + 1), 3))
    jobs = fit_jobs(500, bins, 20)
    assert [len(m) for _, _, m in jobs] == [cap, cap, 1]
    assert sum((m for _, _, m in jobs), []) == list(range(len(bins)))
    for lo, hi, m in jobs:
        assert lo == max(0, bins[m[0]] - 20) and hi == bins[m[-1]] + 20
