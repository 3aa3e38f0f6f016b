import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle_mod():
    """The oracle front end, with the restatement built (and the
    reference-compiled library where the reference checkout exists)."""
    from oracle import oracle as O
    if not os.path.exists(O.ORACLE_SO) or (os.path.isdir(os.path.join(O.REF_ROOT, "src"))
                                           and not os.path.exists(O.REF_SO)):
        O.build()
    return O
