"""Live peak updates and peak snapping of the frame stream's fitter
(IntegratorFitPeaksGPUStream.cu:1712-1776, 2633-2675).  CPU only: no frame
is fitted here (tests/test_gpu_fit.py fits frames)."""
import numpy as np

from midas_amd.stream_fits import StreamPeakFitter, auto_roi, snap_peaks


def test_snap_within_two_bins():
    r = 10.0 + 0.5 * np.arange(20)
    assert snap_peaks(r, [12.1, 10.0, 25.0, 19.9, 9.1], 0.5) == \
        [(4, 12.0), (0, 10.0), (19, 19.5), (0, 10.0)]  # 25.0 is 5.5 away; 19.9 -> 19.5


def test_auto_roi_from_widest_peak():
    y = np.ones(200)
    y[90:111] = 11.0        # FWHM estimate 22 samples (crossings at 89 and 111)
    y[150:153] = 11.0
    assert auto_roi(y, [100, 151]) == 33
    assert auto_roi(np.ones(200), [100]) == 15


def test_peak_update_file(tmp_path):
    f = StreamPeakFitter(np.arange(10.0), 1.0, out_dir=tmp_path)
    try:
        assert not f.enabled and f.fit_bin is None
        assert not f.poll_update()
        (tmp_path / "peak_update.txt").write_text("replace\n3.25\n7.5\n")
        assert f.poll_update()
        assert not (tmp_path / "peak_update.txt").exists()
        assert f.enabled and f.locations == [3.25, 7.5]
        assert (tmp_path / "fit.bin").exists() and (tmp_path / "fit_curves.bin").exists()
        assert (tmp_path / "active_peaks.txt").read_text() == "3.2500\n7.5000\n"
        (tmp_path / "peak_update.txt").write_text("append\n1.0\nnot a number\n2.0\n")
        assert f.poll_update() and f.locations == [3.25, 7.5, 1.0, 2.0]
        (tmp_path / "peak_update.txt").write_text("anything\n8.0\n")
        assert f.poll_update() and f.locations == [8.0]
        (tmp_path / "peak_update.txt").write_text("append\n")
        assert not f.poll_update() and f.locations == [8.0]
    finally:
        f.close()


def test_fit_settings_from_a_parameter_file(tmp_path):
    from midas_amd.params import parse_integrator_params
    pf_ = tmp_path / "ps.txt"
    pf_.write_text("RBinSize 0.5\nPeakLocation 12.5\nPeakLocation 30.25\nFitROIAuto 1\n"
                   "FitROIPadding 12\nSNIPIterations 30\nDoSmoothing 1\nMultiplePeaks 1\n")
    P = parse_integrator_params(pf_)
    assert (P.FitROIAuto, P.DoSmoothing, P.MultiplePeaks, P.DoPeakFit) == (1, 1, 1, 1)
    f = StreamPeakFitter.from_params(P, np.arange(100.0), out_dir=tmp_path)
    try:
        assert f.locations == [12.5, 30.25] and f.roi_auto and f.padding == 12
        assert f.snip == 30 and f.r_bin_size == 0.5 and f.enabled
        assert (tmp_path / "active_peaks.txt").read_text() == "12.5000\n30.2500\n"
    finally:
        f.close()
