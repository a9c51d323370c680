import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    import json
    cache = {}

    def load(name):
        if name not in cache:
            with open(os.path.join(ROOT, "tests", "golden", name)) as f:
                cache[name] = json.load(f)
        return cache[name]
    return load
