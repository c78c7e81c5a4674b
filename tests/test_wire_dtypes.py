"""The TCP front end converts a wire frame to the geometry's pixel type on
the host; only conversions that keep every value are accepted (the
reference converts every wire type to float, IntegratorFitPeaksGPUStream.cu
:1468-1494, 1497-1630).  CPU only."""
import numpy as np
import pytest

from midas_amd.server import wire_refusal


@pytest.mark.parametrize("code,dt", [(1, np.uint16), (0, np.uint16), (1, np.int32),
                                     (1, np.float32), (7, np.int64), (4, np.float64),
                                     (6, np.int32), (6, np.int64), (6, np.float32),
                                     (6, np.float64), (2, np.int64)])
def test_value_preserving_conversions_are_accepted(code, dt):
    assert wire_refusal(code, np.dtype(dt)) is None


@pytest.mark.parametrize("code,dt", [(6, np.uint16), (6, np.uint32), (4, np.uint16),
                                     (7, np.uint16), (7, np.float32), (3, np.int32),
                                     (5, np.float32), (2, np.int32), (9, np.uint16)])
def test_lossy_conversions_are_refused(code, dt):
    why = wire_refusal(code, np.dtype(dt))
    assert why and ("dtype" in why)
