"""Shared fixtures: golden reference cases unpacked into temporary directories,
the FP64 oracle (the checker) and the GPU library (the product)."""
import gzip
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU checks")


_CACHE = {}


def unpack_case(name, dest_root):
    """Copy tests/golden/<name> to dest_root/<name> with train.in decompressed."""
    src = os.path.join(GOLDEN, name)
    dst = os.path.join(dest_root, name)
    if os.path.isdir(dst):
        return dst
    os.makedirs(dst)
    for f in os.listdir(src):
        p = os.path.join(src, f)
        if f.endswith(".gz"):
            with gzip.open(p, "rb") as fi, open(os.path.join(dst, f[:-3]), "wb") as fo:
                fo.write(fi.read())
        else:
            shutil.copyfile(p, os.path.join(dst, f))
    return dst


@pytest.fixture(scope="session")
def golden_root(tmp_path_factory):
    return str(tmp_path_factory.mktemp("golden"))


@pytest.fixture(scope="session")
def golden(golden_root):
    def get(name):
        return unpack_case(name, golden_root)
    return get


@pytest.fixture(scope="session")
def oracle_case(golden):
    """Cached oracle Case per golden case name."""
    from oracle import oracle as O

    def get(name):
        if name not in _CACHE:
            _CACHE[name] = O.Case(golden(name))
        return _CACHE[name]
    return get


def read_ga_out(path):
    rows = []
    for line in open(path):
        v = line.split()
        if v:
            rows.append((int(v[0]), float(v[1]), np.array([float(x) for x in v[2:]], np.float32)))
    return rows


def print_case(*args):
    """Name the case on stdout before a GPU step (flushed)."""
    print("[case]", *args, flush=True)
