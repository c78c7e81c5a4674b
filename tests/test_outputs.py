"""Result files (mc_out_*): the HDF5 reader used here is first checked on a
file written by libhdf5, then every file the output set writes is read back
and compared with the reference's contents as oracle/outputs.py restates
them.  CPU only: the output set makes no GPU call."""
import math
import struct
from pathlib import Path

import numpy as np
import pytest

from oracle import outputs as ref_out
from tests import h5lite

GOLDEN = Path(__file__).resolve().parent / "golden"


# ---- the reader, on a file written by libhdf5 --------------------------------

def test_reader_on_libhdf5_file():
    """The reference's FF_HEDM/Example/consolidated_Output.h5 (a fixture of
    its own regression test, tests/test_ff_hedm.py:464): the parameters
    agree with FF_HEDM/Example/Parameters.txt (px 200 at :53, Lsd 1e6 at
    :56, Wavelength 0.22291 at :85), and the per-grain scalar datasets agree
    with the rows of the grains table stored elsewhere in the file."""
    f = h5lite.File(GOLDEN / "ref_consolidated_Output.h5")
    assert f.dataset("/parameters/px").read() == 200.0
    assert f.dataset("/parameters/Distance").read() == 1e6
    assert f.dataset("/parameters/Wavelength").read() == 0.22291
    assert f.dataset("/parameters/IDsFileName").read() == "SpotsToIndex.csv"
    names = list(f.attrs("/grains")["column_names"])
    table = f.dataset("/grains/summary").read()
    assert table.shape == (3, 47) and len(names) == 47
    for gid in (51, 60, 62):
        row = table[table[:, 0] == gid][0]
        g = f"/grains/grain_{gid:04d}"
        assert f.dataset(g + "/grain_id").read() == gid
        assert f.dataset(g + "/confidence").read() == row[names.index("Confidence")]
        assert f.dataset(g + "/radius").read() == row[names.index("Radius")]
        assert f.dataset(g + "/phase_nr").read() == int(row[names.index("PhaseNr")])
    assert list(f.dataset("/spots_to_index/data").read()) == list(range(50, 68))
    assert "Each row is a grain" in f.attrs("/grain_ids_key")["description"]


# ---- the output set -----------------------------------------------------------

@pytest.fixture(scope="module")
def lib():
    from midas_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libmidas_caking.so is not built")
    return _lib.lib()


def _params(**kw):
    from midas_amd.params import IntegrationParams
    P = IntegrationParams(NrPixelsY=64, NrPixelsZ=64, pxY=200.0, pxZ=200.0, Lsd=1e6,
                          RMin=10.0, RMax=30.0, RBinSize=2.0, EtaMin=-180.0, EtaMax=180.0,
                          EtaBinSize=45.0)
    for k, v in kw.items():
        setattr(P, k, v)
    return P


def _cakes(n, nr, ne, seed):
    rng = np.random.default_rng(seed)
    c = rng.gamma(2.0, 50.0, (n, nr, ne)).astype(np.float32)
    c[:, 0, 1] = 0.0           # empty bins are skipped by the CPU lineout
    c[:, 2, 3] = np.nan        # mask-contaminated bins
    if n > 1:
        c[1, 4, :] = 0.0       # a row without values
    return c


def test_caked_hdf_and_lineouts(lib, tmp_path):
    from midas_amd.outputs import ResultFiles
    from midas_amd.params import bin_edges
    P = _params(OmegaStart=-10.0, OmegaStep=0.25, Lam=0.17, Polariz=0.95, SHpL=0.003,
                U=1.1, V=-0.1, W=0.06, X=0.5, Y=0.25, Z=0.125)
    RLo, RHi, ELo, EHi = bin_edges(P)
    nr, ne = RLo.size, ELo.size
    F, N = 7, 3
    cakes = _cakes(F, nr, ne, 1)
    rng = np.random.default_rng(2)
    lo = rng.random((F, 3, nr)).astype(np.float32)
    metas = rng.random((F, 4)) * 100
    reta = rng.random((5, nr, ne))
    frame_nrs = [0, 1, 2, 3, 4, 5, 6]
    files = {"caked_hdf", "per_frame", "sum_frames", "omegas", "lineout", "stream", "int2d"}
    with ResultFiles(tmp_path, "scan_00017.zip", P, files=files, omega_sum_frames=N,
                     reta=reta) as out:
        for i in range(F):
            out.frame(frame_nrs[i], cakes[i], lo[i], meta=metas[i])
    omegas = [P.OmegaStart + i * P.OmegaStep for i in frame_nrs]
    f = h5lite.File(tmp_path / "scan_00017.zip.caked.hdf")
    assert sorted(f.keys("/")) == ["InstrumentParameters", "IntegrationResult", "OmegaSumFrame",
                                   "Omegas", "REtaMap", "SumFrames"]
    for i in range(F):
        d = f.dataset(f"/IntegrationResult/FrameNr_{i}")
        np.testing.assert_array_equal(d.read(), cakes[i].astype(np.float64))
        assert d.attrs["omega"][0] == omegas[i]
        assert d.attrs["Header"] == "Radius,Eta" and d.attrs["Units"] == "Pixels,Degrees"
    want = ref_out.omega_sum_chunks(frame_nrs, omegas, cakes, N, metas)
    assert sorted(f.keys("/OmegaSumFrame")) == sorted(w["name"] for w in want)
    for w in want:
        d = f.dataset("/OmegaSumFrame/" + w["name"])
        np.testing.assert_allclose(d.read(), w["data"], rtol=1e-12, equal_nan=True)
        for k in ("LastFrameNumber", "Number Of Frames Summed", "FirstOme", "LastOme",
                  "Temperature", "Pressure", "I", "I0"):
            np.testing.assert_allclose(d.attrs[k][0], w[k], rtol=1e-12, err_msg=k)
    s = f.dataset("/SumFrames")
    np.testing.assert_allclose(s.read(), cakes.astype(np.float64).sum(0), rtol=1e-12,
                               equal_nan=True)
    assert s.attrs["nFrames"][0] == F
    np.testing.assert_array_equal(f.dataset("/Omegas").read(), omegas)
    assert f.dataset("/Omegas").attrs["Units"] == "Degrees"
    r = f.dataset("/REtaMap")
    np.testing.assert_array_equal(r.read(), reta)
    assert r.attrs["nEtaBins"][0] == ne and r.attrs["nRBins"][0] == nr
    assert r.attrs["Header"] == "Radius,2Theta,Eta,BinArea,Q"
    assert r.attrs["Units"] == "Pixels,Degrees,Degrees,Pixels,InvAngstrom"
    inst = {k: f.dataset("/InstrumentParameters/" + k).read()[0] for k in
            ("Polariz", "Lam", "SH_L", "U", "V", "W", "X", "Y", "Z", "Distance")}
    assert inst == {"Polariz": 0.95, "Lam": 0.17, "SH_L": 0.003, "U": 1.1, "V": -0.1,
                    "W": 0.06, "X": 0.5, "Y": 0.25, "Z": 0.125, "Distance": 1e6}
    # lineouts: <prefix> = name without its extension
    rc = (RLo + RHi) / 2
    lb = np.fromfile(tmp_path / "scan_00017_lineout.bin").reshape(F, nr)
    xy = (tmp_path / "scan_00017_lineout.xy").read_text()
    want_xy = ""
    for i in range(F):
        l1 = ref_out.lineout_cpu_mean(cakes[i])
        np.testing.assert_allclose(lb[i], l1, rtol=1e-14)
        want_xy += ref_out.lineout_xy(l1, rc, P.pxY, P.Lsd)
    assert xy == want_xy
    st = np.fromfile(tmp_path / "lineout.bin").reshape(F, nr, 2)
    sm = np.fromfile(tmp_path / "lineout_simple_mean.bin").reshape(F, nr, 2)
    for i in range(F):
        np.testing.assert_array_equal(st[i], ref_out.stream_pairs(rc, lo[i, 1]))
        np.testing.assert_array_equal(sm[i], ref_out.stream_pairs(rc, lo[i, 2]))
    i2 = np.fromfile(tmp_path / "Int2D.bin").reshape(F, nr, ne)
    np.testing.assert_array_equal(i2, cakes.astype(np.float64))


def test_oracle_lineout_matches_c_oracle(oracle_lib):
    """The FP64 lineout restatement agrees with the C oracle's CPU lineout."""
    cakes = _cakes(3, 10, 8, 5).astype(np.float64)
    for c in cakes:
        lc, _, _ = oracle_lib.lineouts(c, np.ones_like(c))
        np.testing.assert_allclose(ref_out.lineout_cpu_mean(c), lc, rtol=1e-15)


def test_q_mode_lineout_xy(lib, tmp_path):
    from midas_amd.outputs import ResultFiles
    P = _params(QMin=1.0, QMax=4.0, QBinSize=0.5, Wavelength=0.2)
    assert P.q_mode
    from midas_amd.params import bin_edges
    RLo, RHi, ELo, EHi = bin_edges(P)
    cakes = _cakes(2, RLo.size, ELo.size, 3)
    with ResultFiles(tmp_path, "q.h5", P, files={"lineout"}) as out:
        for i in range(2):
            out.frame(i, cakes[i])
    want = "".join(ref_out.lineout_xy(ref_out.lineout_cpu_mean(c), (RLo + RHi) / 2, P.pxY,
                                      P.Lsd, True, P.QMin, P.QBinSize) for c in cakes)
    assert (tmp_path / "q_lineout.xy").read_text() == want


def test_peak_files(lib, tmp_path):
    from midas_amd.outputs import ResultFiles
    from midas_amd.params import bin_edges
    P = _params(Lam=0.172978)
    RLo, RHi, ELo, EHi = bin_edges(P)
    nr, ne = RLo.size, ELo.size
    rng = np.random.default_rng(7)
    n = 40
    fr = rng.integers(0, 3, n)
    eb = rng.integers(0, ne, n)
    pk = rng.integers(0, 4, n)
    p = rng.random((n, 7)) * [100, 30, 5, 3, 0.2, 1, 50]
    p[::7, 1] = 0.0  # not fitted: skipped
    lfits = rng.random((2, 7))
    with ResultFiles(tmp_path, "/some/dir/run_0001.tif", P, files={"peaks"},
                     source="run_0001.zip") as out:
        for i in range(3):
            out.frame(i, _cakes(1, nr, ne, i)[0])
        out.lineout_fits(0, lfits)
        out.peaks(fr, eb, pk, p)
    csv = (tmp_path / "run_0001_fit_per_eta.csv").read_text()
    want = ref_out.FIT_PER_ETA_HEADER
    rows = []
    for i in range(n):
        ec = (ELo[eb[i]] + EHi[eb[i]]) / 2
        line = ref_out.fit_per_eta_line(fr[i], eb[i], ec, pk[i], p[i], P.pxY, P.Lsd)
        if line:
            want += line
            rows.append(ref_out.pfio_row(fr[i], eb[i], pk[i], ec, p[i], P.pxY, P.Lsd, P.Lam))
    assert csv == want
    np.testing.assert_array_equal(np.fromfile(tmp_path / "run_0001_fit.bin").reshape(2, 7),
                                  lfits)
    f = h5lite.File(tmp_path / "run_0001_caked_peaks.h5")
    assert f.attrs("/metadata")["zarr_source"] == "run_0001.zip"
    tth = np.arctan((RLo + RHi) / 2 * P.pxY / P.Lsd) * 180 / math.pi
    np.testing.assert_allclose(f.dataset("/metadata/tth_axis").read(), tth, rtol=1e-15)
    np.testing.assert_array_equal(f.dataset("/metadata/eta_axis").read(), (ELo + EHi) / 2)
    assert list(f.dataset("/metadata/frame_keys").read()) == ["0", "1", "2"]
    for k in ("frame_idx", "eta_idx", "peak_nr"):
        d = f.dataset("/peaks/" + k)
        assert d.dtype == ("i", 4)
        assert list(d.read()) == [r[k] for r in rows]
    for k in ("eta_deg", "center_2theta", "area", "sig", "gam", "FWHM_deg", "eta_mix",
              "d_spacing_A", "chi_sq"):
        np.testing.assert_allclose(f.dataset("/peaks/" + k).read(), [r[k] for r in rows],
                                   rtol=1e-14, err_msg=k)
    assert list(f.dataset("/peaks/frame_key").read()) == [str(r["frame_idx"]) for r in rows]


def test_peaks_from_ring_fits(lib, tmp_path):
    """pV singlets keep centre and SNR only, when SNR > 1
    (IntegratorZarrOMP.c:2013-2026)."""
    from midas_amd.outputs import ResultFiles
    P = _params()
    fits = np.zeros((2, 8, 3, 8))
    fits[..., 1] = 15.0
    fits[..., 6] = 3.0
    fits[..., 0] = 99.0
    fits[0, 1, 2, 6] = 0.5      # SNR below 1: dropped
    fits[1, 4, 0, 7] = -1.0     # failed fit: dropped
    with ResultFiles(tmp_path, "r.bin", P, files={"peaks"}) as out:
        out.peaks_from_ring_fits(fits, frame0=10)
    lines = (tmp_path / "r_fit_per_eta.csv").read_text().splitlines()[1:]
    assert len(lines) == 2 * 8 * 3 - 2
    first = lines[0].split(",")
    assert first[0] == "10" and float(first[7]) == 0.0 and float(first[12]) == 3.0


def test_many_members_and_strings(lib, tmp_path):
    """Groups larger than one symbol-table node and string datasets larger
    than one global heap collection."""
    from midas_amd.outputs import ResultFiles
    P = _params(RMax=14.0, EtaBinSize=180.0)
    F = 70000
    c = np.ones((2, 2), np.float32)
    with ResultFiles(tmp_path, "big.dat", P, files={"peaks"}) as out:
        for i in range(F):
            out.frame(i, c)
    f = h5lite.File(tmp_path / "big_caked_peaks.h5")
    keys = f.dataset("/metadata/frame_keys").read()
    assert len(keys) == F and keys[0] == "0" and keys[-1] == str(F - 1)
    assert not (tmp_path / "big.dat.caked.hdf").exists()  # no caked file asked for
    # More members than one B-tree node holds (2 x 256 symbol-table nodes of
    # 128): a two-level group tree.
    n = 66000
    with ResultFiles(tmp_path, "many.dat", P, files={"per_frame"}, omega_sum_frames=0) as out:
        for i in range(n):
            out.frame(i, c * i)
    f = h5lite.File(tmp_path / "many.dat.caked.hdf")
    ks = f.keys("/IntegrationResult")
    assert len(ks) == n and ks == sorted(ks, key=lambda k: k.encode())
    for i in (0, 65535, 65536, n - 1):
        assert f.dataset(f"/IntegrationResult/FrameNr_{i}").read()[0, 0] == float(i)


def test_errors(lib, tmp_path):
    from midas_amd._lib import MidasError
    from midas_amd.outputs import ResultFiles
    P = _params()
    (tmp_path / "x_lineout.bin").mkdir()  # the file name is taken by a directory
    with pytest.raises(MidasError, match="x_lineout.bin"):
        ResultFiles(tmp_path, "x.zip", P, files={"lineout"})
    with pytest.raises(ValueError):
        ResultFiles(tmp_path, "x", P, files={"nonsense"})
    out = ResultFiles(tmp_path, "s.bin", P, files={"stream"})
    from midas_amd.params import bin_edges
    RLo, _, ELo, _ = bin_edges(P)
    with pytest.raises(MidasError):
        out.frame(0, np.zeros((RLo.size, ELo.size), np.float32))  # stream needs lineouts
    out.close()
