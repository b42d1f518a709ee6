import gzip
import os
import shutil
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "maximal-kplex-enumeration_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs (KP_SLOW=1)")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("KP_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow oracle run; set KP_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def datasets(tmp_path_factory):
    """Decompresses the two graphs shipped with the reference (dataset/*.bin)."""
    d = tmp_path_factory.mktemp("datasets")
    out = {}
    for name in ("soc-epinions", "email-euall"):
        dst = d / f"{name}.bin"
        with gzip.open(os.path.join(GOLDEN, f"{name}.bin.gz"), "rb") as fi, open(dst, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        out[name] = str(dst)
    return out
