"""Shared test set-up: paths, the `gpu` marker, and loaders.

`-m "not gpu"` tests run on the CPU only (oracle vs. the reference's
known-answer vectors, host logic, the C-ABI library's exported symbols).
`-m gpu` tests call the HIP library through its C ABI on a real GPU and
compare it with the oracle.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "gempic.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")


def golden(name):
    import json
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)
