import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libmidas_caking.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as orc
    orc.lib()
    return orc


@pytest.fixture(scope="session")
def ref_lib():
    from oracle import oracle as orc
    if orc.ref_lib() is None:
        pytest.skip("reference sources are not present here (they are read only in the "
                    "build container)")
    return orc
