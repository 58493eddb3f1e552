import gzip
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dynex-neuromorphic-chip_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CNF_NAMES = {
    n: f"transformed_barthel_n_{n}_r_8.000_p0_0.080_instance_001.cnf" for n in (100, 200, 500, 1000, 10000)
}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdnx on cuda:0)")


@pytest.fixture(scope="session")
def cnf_dir(tmp_path_factory):
    """The reference's five DIMACS fixtures (cnf/*.cnf), unpacked."""
    d = tmp_path_factory.mktemp("cnf")
    for n, name in CNF_NAMES.items():
        with gzip.open(os.path.join(GOLDEN, "cnf", name + ".gz"), "rb") as f:
            (d / name).write_bytes(f.read())
    return d


def cnf_path(cnf_dir, n):
    return str(cnf_dir / CNF_NAMES[n])


def planted_cp(m, k=3):
    return np.arange(0, k * m + 1, k, dtype=np.int32)


def random_state(rng, n, m, xl_hi=5.0):
    """A state with V in [-1,1] (with exact +-1 and 0 sprinkled in), Xs in
    [0,1] and Xl in [1, xl_hi]."""
    v = rng.uniform(-1, 1, n)
    idx = rng.choice(n, size=max(1, n // 10), replace=False)
    v[idx] = rng.choice([-1.0, 0.0, 1.0], size=idx.size)
    xs = rng.uniform(0, 1, m)
    xl = rng.uniform(1, xl_hi, m)
    return np.concatenate([v, xs, xl]).astype(np.float64)

