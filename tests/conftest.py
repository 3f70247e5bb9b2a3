import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs an MI355X (gfx950) and the built HIP library')
    config.addinivalue_line('markers', 'slow: long-running test')


@pytest.fixture(scope='session')
def ref():
    """The reference C kernels (oracle/_ref); skipped where not built."""
    from oracle import refbuild
    if not refbuild.available():
        try:
            refbuild.build()
        except Exception:
            pass
    if not refbuild.available():
        pytest.skip('reference build oracle/_ref not available')
    refbuild.lib()
    return refbuild


@pytest.fixture
def rng():
    import numpy as np
    return np.random.default_rng(42)
