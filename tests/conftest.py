import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def kats():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as f:
        return json.load(f)


def piece(pattern: str) -> np.ndarray:
    """ASCII tetromino -> (rows, cols, 1) int32 grid ('#' = 1)."""
    lines = pattern.strip().split("\n")
    g = np.zeros((len(lines), max(len(l) for l in lines), 1), dtype=np.int32)
    for i, line in enumerate(lines):
        for j, ch in enumerate(line):
            if ch == "#":
                g[i, j, 0] = 1
    return g


def grid_from_rows(rows, shape):
    g = np.zeros(shape, dtype=np.int32)
    for i, r in enumerate(rows):
        for j, tok in enumerate(r.split()):
            if tok != ".":
                g[i, j, 0] = int(tok)
    return g
