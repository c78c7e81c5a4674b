"""The C oracle against the reference's mapper/integrator compiled in place
(oracle/_ref).  Skipped where the reference sources are absent."""
import numpy as np
import pytest

from helpers import panel_shifts_for, small_cases, to_oracle


@pytest.mark.parametrize("name", list(small_cases()))
def test_map_byte_identical(name, ref_lib):
    P = small_cases()[name]
    rng = np.random.default_rng(21)
    mask = (rng.random((P.NrPixelsZ, P.NrPixelsY)) < 0.015).astype(np.float64)
    sh = panel_shifts_for(name)
    a = ref_lib.build_map(to_oracle(P), mask=mask, panel_shifts=sh)
    b = ref_lib.ref_build_map(to_oracle(P), mask=mask, panel_shifts=sh)
    np.testing.assert_array_equal(a.pairs, b.pairs)
    np.testing.assert_array_equal(a.flags, b.flags)
    assert a.entries.tobytes() == b.entries.tobytes()


def test_map_with_distortion_and_residual_maps(ref_lib):
    P = small_cases()["tilt"]
    rng = np.random.default_rng(4)
    dy = rng.normal(0, 0.05, (P.NrPixelsZ, P.NrPixelsY))
    dz = rng.normal(0, 0.05, (P.NrPixelsZ, P.NrPixelsY))
    res = rng.normal(0, 0.02, (P.NrPixelsZ, P.NrPixelsY))
    a = ref_lib.build_map(to_oracle(P), dist_y=dy, dist_z=dz, residual=res)
    b = ref_lib.ref_build_map(to_oracle(P), dist_y=dy, dist_z=dz, residual=res)
    assert a.entries.tobytes() == b.entries.tobytes()


@pytest.mark.parametrize("name", ["flat", "subpix", "tilt"])
def test_integration_matches_integration_apply_map(name, ref_lib):
    """integration_apply_map (IntegrationCore.c) precomputes the bilinear
    weights, IntegratorZarrOMP.c multiplies them inline; the oracle follows
    the latter, so fractional coordinates agree to rounding and integer
    coordinates bit-exactly."""
    P = small_cases()[name]
    rng = np.random.default_rng(5)
    m = ref_lib.build_map(to_oracle(P))
    img = rng.poisson(80.0, (P.NrPixelsZ, P.NrPixelsY)) - 7.5
    for grad in (False, True):
        c1, a1 = ref_lib.integrate(m, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                                   gradient=grad)
        c2, a2 = ref_lib.ref_integrate(m, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                                       gradient=grad)
        np.testing.assert_array_equal(a1, a2)
        if P.SubPixelLevel == 1 and not grad:
            assert c1.tobytes() == c2.tobytes()
        else:
            np.testing.assert_allclose(c1, c2, rtol=1e-12, atol=1e-10, equal_nan=True)


def test_geometry_vectors(ref_lib):
    P = to_oracle(small_cases()["tilt"])
    rng = np.random.default_rng(8)
    Y = rng.uniform(-5, 85, 200)
    Z = rng.uniform(-5, 70, 200)
    R1, E1 = ref_lib.pixel_to_reta(P, Y, Z)
    R2, E2 = ref_lib.ref_pixel_to_reta(P, Y, Z)
    assert R1.tobytes() == R2.tobytes() and E1.tobytes() == E2.tobytes()
