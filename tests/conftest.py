import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbwhip.so)")
    config.addinivalue_line("markers", "slow: large CPU case")


def _ensure_built():
    """Build the libraries only if they are missing (a fresh checkout on a
    build machine); GPU test runs are expected to find them prebuilt."""
    from bigwham_dev_amd import build as b
    if not (b.LIB / "libbwhost.so").exists():
        b.build_host()
    if not (b.LIB / "libbwhip.so").exists():
        b.build_hip()
    import build_oracle
    if not (ROOT / "oracle" / "lib" / "liboracle.so").exists():
        build_oracle.build_oracle()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
