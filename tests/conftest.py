import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "broadphase-gpu_amd"))
sys.path.insert(0, ROOT)

from stq_amd import scenes  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def boxes_from_json(blist):
    b = scenes.empty_boxes(len(blist))
    for i, e in enumerate(blist):
        b["min"][i] = e["min"]
        b["max"][i] = e["max"]
        b["vids"][i] = e["vids"]
    return b


@pytest.fixture(scope="session")
def known_answers():
    with open(os.path.join(GOLDEN, "known_answers.json")) as fh:
        data = json.load(fh)
    return {k: (boxes_from_json(v["boxes"]), np.array(v["pairs"], dtype=np.int32).reshape(-1, 2))
            for k, v in data.items()}


@pytest.fixture(scope="session")
def rounding_vectors():
    with open(os.path.join(GOLDEN, "rounding.json")) as fh:
        return json.load(fh)
