"""Per-frame lineout fits of the frame stream (IntegratorFitPeaksGPUStream.cu
:1712-1776, 2599-2700).

The reference GPU stream fits each frame's area-weighted lineout on the CPU
right after the frame, on the processing thread: at given peak locations
(each snapped to the nearest R bin when within two bins; ROI half width
FitROIPadding, or with FitROIAuto max(15, 1.5 x the widest estimated FWHM)),
or, with AutoDetectPeaks, by fitPeaksAutoDetect.  It appends 7 doubles per
fitted peak to ``fit.bin`` and polls ``peak_update.txt`` for live changes of
the peak list: the first line says ``replace`` (default) or ``append``, the
remaining lines are R values; the file is removed after reading, fitting is
switched on, and ``active_peaks.txt`` lists the active peaks ("%.4f" lines).

``StreamPeakFitter`` keeps these semantics but takes the fits off the frame
path: frames are queued (``submit``) and a worker thread fits every queued
frame's windows in one GPU launch per window length (``fit_windows``).
Use it from ``FrameServer(on_result=...)`` with row 1 (area-weighted) of the
frame's lineouts, the profile the stream fits.  Rows for ``_caked_peaks.h5``
go to an optional ``ResultFiles`` (``peaks``).
"""
from __future__ import annotations

import os
import queue
import threading
from pathlib import Path

import numpy as np

from . import peakfit as pf

MAX_PEAK_LOCATIONS = 100


def snap_peaks(r, locations, r_bin_size):
    """Bins of the locations within 2 bins of a bin centre (the first
    nearest bin wins), .cu:2633-2650: [(bin, radius)]."""
    r = np.asarray(r, dtype=np.float64)
    out = []
    for p in locations:
        d = np.abs(r - p)
        b = int(np.argmin(d))
        if d[b] < r_bin_size * 2.0:
            out.append((b, float(r[b])))
    return out


def auto_roi(profile, bins):
    """FitROIAuto, .cu:2659-2675: max(15, int(1.5 x the widest FWHM
    estimated on +-50 bins around each peak))."""
    y = np.asarray(profile, dtype=np.float64)
    n = y.size
    widest = 0.0
    for b in bins:
        lo, hi = max(0, b - 50), min(n - 1, b + 50)
        if lo >= hi:
            continue
        fw, _, _ = pf.estimate_initial_params(y[lo:hi + 1], b - lo)
        widest = max(widest, fw)
    return max(15, int(widest * 1.5))


class StreamPeakFitter:
    def __init__(self, r_centers, r_bin_size: float, peak_locations=(), *,
                 auto_detect: int = 0, snip_iterations: int = 50, fit_roi_padding: int = 20,
                 fit_roi_auto: bool = False, out_dir=".", results=None, max_batch: int = 64,
                 ctx=None):
        self.r = np.asarray(r_centers, dtype=np.float64)
        self.r_bin_size = float(r_bin_size)
        self.locations = [float(v) for v in peak_locations][:MAX_PEAK_LOCATIONS]
        self.auto_detect = int(auto_detect)
        self.snip = int(snip_iterations)
        self.padding = int(fit_roi_padding)
        self.roi_auto = bool(fit_roi_auto)
        self.dir = Path(out_dir)
        self.results = results
        self.max_batch = max_batch
        self.ctx = ctx
        self.enabled = bool(self.locations) or self.auto_detect > 0
        self.fit_bin = open(self.dir / "fit.bin", "wb") if self.enabled else None
        self.fit_curves = open(self.dir / "fit_curves.bin", "wb") if self.enabled else None
        if self.enabled and self.locations:
            self._write_active()
        self.fitted: dict[int, np.ndarray] = {}  # frame -> [n, 7]
        self._q: queue.Queue = queue.Queue()
        self._err: BaseException | None = None
        self._th = threading.Thread(target=self._loop, daemon=True)
        self._th.start()

    @classmethod
    def from_params(cls, P, r_centers, **kw):
        """The stream's fit settings of a parameter file (.cu:1930-1966):
        PeakLocation, AutoDetectPeaks, SNIPIterations, FitROIPadding,
        FitROIAuto; bins of RBinSize."""
        return cls(r_centers, P.RBinSize, P.PeakLocations, auto_detect=P.AutoDetectPeaks,
                   snip_iterations=P.SNIPIterations, fit_roi_padding=P.FitROIPadding,
                   fit_roi_auto=bool(P.FitROIAuto), **kw)

    # ---- live updates (check_peak_update, .cu:1712-1776) ----
    def _write_active(self):
        with open(self.dir / "active_peaks.txt", "w") as f:
            for v in self.locations:
                f.write(f"{v:.4f}\n")

    def poll_update(self) -> bool:
        path = self.dir / "peak_update.txt"
        try:
            with open(path) as f:
                lines = f.read().splitlines()
        except OSError:
            return False
        replace = not (lines and "append" in lines[0])
        new = []
        for ln in lines[1:]:
            if len(new) >= MAX_PEAK_LOCATIONS:
                break
            try:
                new.append(float(ln.split()[0]))
            except (ValueError, IndexError):
                continue
        try:
            os.remove(path)
        except OSError:
            pass
        if not new:
            return False
        if replace:
            self.locations = []
        self.locations = (self.locations + new)[:MAX_PEAK_LOCATIONS]
        self.enabled = True
        if self.fit_bin is None:
            self.fit_bin = open(self.dir / "fit.bin", "wb")
        if self.fit_curves is None:
            self.fit_curves = open(self.dir / "fit_curves.bin", "wb")
        self._write_active()
        return True

    # ---- fitting ----
    def submit(self, frame_idx: int, profile) -> None:
        """Queues one frame's area-weighted lineout [n_r]."""
        if self._err is not None:
            raise RuntimeError("peak fitting failed") from self._err
        self._q.put((int(frame_idx), np.asarray(profile, dtype=np.float64).copy()))

    def _fit_batch(self, batch):
        """The frames' fits, in frame order: [(frame, rows [k, 7])]."""
        if self.auto_detect > 0:
            return [(f, self._auto_rows(y)) for f, y in batch]
        pk = snap_peaks(self.r, self.locations, self.r_bin_size)
        if not pk:
            return [(f, np.zeros((0, 7))) for f, _ in batch]
        bins = [b for b, _ in pk]
        wins, owners = [], []
        for k, (f, y) in enumerate(batch):
            roi = auto_roi(y, bins) if self.roi_auto else self.padding
            for j, b in enumerate(bins):
                lo, hi = max(0, b - roi), min(self.r.size - 1, b + roi)
                wins.append((lo, hi, self.r[b]))
                owners.append((k, j))
        rows = np.zeros((len(batch), len(bins), 7))
        by_len: dict[int, list] = {}
        for w, o in zip(wins, owners):
            by_len.setdefault(w[1] - w[0] + 1, []).append((w, o))
        for L, items in by_len.items():
            X = np.stack([self.r[lo:hi + 1] for (lo, hi, _), _ in items])
            Y = np.stack([batch[k][1][lo:hi + 1] for (lo, hi, _), (k, _) in items])
            G = np.array([c for (_, _, c), _ in items])
            res = pf.fit_windows(X, Y, G, self.r_bin_size, model="tch", max_iter=200,
                                 ctx=self.ctx)
            for row, (_, (k, j)) in zip(res, items):
                if row[7] > 0:
                    rows[k, j] = row[:7]
        return [(f, rows[k]) for k, (f, _) in enumerate(batch)]

    def _auto_rows(self, y):
        out, n = pf.fit_peaks_auto(self.r, y, self.auto_detect, self.r_bin_size, self.padding,
                                   self.snip, ctx=self.ctx)
        return out[:n]

    def _emit(self, frame, rows):
        self.fitted[frame] = rows
        if rows.shape[0] and self.fit_bin is not None:
            self.fit_bin.write(np.ascontiguousarray(rows, dtype="<f8").tobytes())
            self.fit_bin.flush()
        if self.results is not None and rows.shape[0]:
            self.results.peaks(frame, 0, np.arange(rows.shape[0]), rows)

    def _loop(self):
        while True:
            item = self._q.get()
            if item is None:
                return
            batch = [item]
            while len(batch) < self.max_batch:
                try:
                    nxt = self._q.get_nowait()
                except queue.Empty:
                    break
                if nxt is None:
                    self._q.put(None)
                    break
                batch.append(nxt)
            try:
                self.poll_update()
                if self.enabled:
                    for f, rows in self._fit_batch(batch):
                        self._emit(f, rows)
            except BaseException as e:  # reported by submit() and close()
                self._err = e
                return

    def close(self):
        self._q.put(None)
        self._th.join()
        for f in (self.fit_bin, self.fit_curves):
            if f is not None:
                f.close()
        self.fit_bin = self.fit_curves = None
        if self._err is not None:
            raise RuntimeError("peak fitting failed") from self._err

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
