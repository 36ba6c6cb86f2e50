"""pytest configuration: the `gpu` marker and shared fixtures.

Tests marked `gpu` need an MI355X and call the library through its C ABI;
everything else runs on the CPU (oracle pinning, host logic, ABI exports).
"""
import os
import sys

import pytest

# The oracle's MKL tile kernels must run on MKL's sequential threading layer,
# which MKL fixes at its first routine call in the process.
os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X GPU; calls the HIP library")


@pytest.fixture(scope="session")
def plasma():
    """The loaded C library, initialised once for the GPU tests."""
    import plasma_amd
    lib = plasma_amd.load()
    rc = lib.plasma_init()
    assert rc == plasma_amd.PlasmaSuccess, "plasma_init failed: %d" % rc
    yield lib
    lib.plasma_finalize()


@pytest.fixture()
def nb_guard(plasma):
    """Restore the tile size / tuning flags a test changes."""
    import plasma_amd as pa
    params = (pa.PlasmaNb, pa.PlasmaTuning, pa.PlasmaLookahead,
              pa.PlasmaUpdateDepth, pa.PlasmaGraphCapture,
              pa.PlasmaSplitUpdate, pa.PlasmaHalfInput)
    saved = [pa.get_param(p)[1] for p in params]
    # a test that sets nb means that nb: the autotuned table (on by default,
    # as in the reference) is switched off unless the test enables it
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    yield
    for p, v in zip(params, saved):
        pa.set_param(p, v)
