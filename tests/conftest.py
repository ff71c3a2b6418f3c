import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libabc.so)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def golden_recovered():
    return dict(np.load(os.path.join(GOLDEN, "recovered.npz")))


@pytest.fixture(scope="session")
def golden_summary():
    return dict(np.load(os.path.join(GOLDEN, "summary_stats.npz")))


@pytest.fixture(scope="session")
def golden_capture():
    return np.load(os.path.join(GOLDEN, "capture.npz"))["betas"]


@pytest.fixture(scope="session")
def engine_factory():
    from abc_inference_transcription_amd.engine import Engine
    made = []

    def make(design):
        e = Engine(design)
        made.append(e)
        return e
    yield make
    for e in made:
        e.close()
