"""Shared fixtures.

``-m "not gpu"`` tests exercise the CPU oracle against the reference's own
known-answer vectors and closed forms, the host logic and the C ABI's symbol
table.  ``-m gpu`` tests call the HIP library through the C ABI and compare it
with the oracle.
"""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np
import pytest

# PyTorch ships its own HIP runtime with the same soname as /opt/rocm's.  The
# first one loaded serves every later library in the process, and torch cannot
# start its GPU side on top of the other one, so torch is imported before the
# library is loaded whenever it is installed (tests hand torch device buffers
# to the library).
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "flexpart-gpu_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    """Import flexpart-gpu_amd/ as the package ``flexpart_gpu_amd``."""
    name = "flexpart_gpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def fp():
    return load_package()


@pytest.fixture(scope="session")
def workloads(fp):
    import flexpart_gpu_amd.workloads as w  # noqa: F401  (package loaded by fp)
    return w


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib("fpm")
    return oracle


def met_dict(met, bracket=0):
    u0, v0, w0 = met.levels[bracket]
    u1, v1, w1 = met.levels[bracket + 1]
    return dict(shape=met.shape, u0=u0, v0=v0, w0=w0, u1=u1, v1=v1, w1=w1,
                level_heights=met.level_heights)


def pbl_dict(pbl):
    return dict(shape=pbl.shape, ustar=pbl.ustar, wstar=pbl.wstar, hmix=pbl.hmix,
                oli=pbl.oli, precip=pbl.precip)


def assert_particles_equal(a: np.ndarray, b: np.ndarray, what: str = ""):
    """Bitwise equality of two particle arrays, with a readable first mismatch."""
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    av = a.view(np.uint32).reshape(len(a), 24)
    bv = b.view(np.uint32).reshape(len(b), 24)
    bad = np.nonzero((av != bv).any(axis=1))[0]
    if len(bad):
        i = bad[0]
        names = [n for n in a.dtype.names for _ in range(4 if n == "mass" else 1)]
        cols = np.nonzero(av[i] != bv[i])[0]
        detail = ", ".join(f"{names[c]}: {a[i:i+1].view(np.uint32).reshape(24)[c]:#x} vs "
                           f"{b[i:i+1].view(np.uint32).reshape(24)[c]:#x}" for c in cols[:6])
        raise AssertionError(f"{what}: {len(bad)} of {len(a)} slots differ; first slot {i}: {detail}\n"
                             f"A={a[i]}\nB={b[i]}")
