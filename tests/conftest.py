import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    import oracle_api
    oracle_api.lib()
    return oracle_api


@pytest.fixture(scope="session")
def fixture_file(tmp_path_factory):
    """Decompressed copy of a reference fixture file (the library reads plain files)."""
    import gzip
    import shutil

    import mmp_io
    base = tmp_path_factory.mktemp("fixtures")
    cache = {}

    def get(system, name):
        key = (system, name)
        if key not in cache:
            dst = base / f"{system}_{name}"
            with gzip.open(mmp_io.fixture_path(system, name), "rb") as fi, open(dst, "wb") as fo:
                shutil.copyfileobj(fi, fo)
            cache[key] = str(dst)
        return cache[key]
    return get
