import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402

load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def fmm():
    import fmm_gen_amd
    return fmm_gen_amd


@pytest.fixture(scope="session")
def rt():
    from fmm_gen_amd import runtime
    return runtime


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    return O


def int_mats(rng, m, n, k, lo=-4, hi=4):
    A = np.asfortranarray(rng.integers(lo, hi + 1, (m, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(lo, hi + 1, (k, n)).astype(np.float64))
    C = np.asfortranarray(rng.integers(lo, hi + 1, (m, n)).astype(np.float64))
    return A, B, C
