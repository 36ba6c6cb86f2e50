"""The oracle's MKL backend and MKL's threading layer (CPU only).

MKL fixes its threading layer at its first routine call.  In a process that
imported torch first (torch loads its libgomp with RTLD_GLOBAL), MKL's
Intel-OpenMP layer forks its teams through libiomp5 while its
omp_get_thread_num / omp_get_num_threads calls bind to libgomp, so every
thread of a team computes the same share and the results are wrong: the
oracle's dsposv then returned iter = -31 (8+ threads) or info = 129 (one
thread) on a kappa = 1e2 system whose correct answer is two iterations.

These tests run that process mix in a fresh interpreter:
  * an MKL routine called first without the sequential layer: the oracle
    must refuse its MKL backend (never compute on the broken layer);
  * the test helpers' MKL handle called first: it selects the sequential
    layer, and the oracle's dsposv gives the same iteration count and
    solution as its built-in tile kernels.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MKL_RT = "/opt/conda/lib/libmkl_rt.so.1"

pytestmark = pytest.mark.skipif(not os.path.exists(MKL_RT),
                                reason="MKL not present")

_PRELUDE = """
import ctypes, sys
sys.path.insert(0, %r)
import torch                      # libgomp loaded globally first
import numpy as np
""" % ROOT


def _run(body, env_layer):
    env = dict(os.environ)
    env.pop("MKL_THREADING_LAYER", None)
    if env_layer:
        env["MKL_THREADING_LAYER"] = env_layer
    r = subprocess.run([sys.executable, "-c", _PRELUDE + body], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=600)
    return r.returncode, r.stdout


def test_oracle_refuses_threaded_mkl():
    body = """
m = ctypes.CDLL(%r)
a = np.zeros(16)
iseed = (ctypes.c_int * 4)(0, 0, 0, 1)
m.LAPACKE_dlarnv(1, iseed, 16, a.ctypes.data_as(ctypes.c_void_p))
import oracle as ora
try:
    ora.set_blas(ora.MKL)
except RuntimeError as e:
    print("REFUSED", e)
    sys.exit(0)
print("ACCEPTED")
sys.exit(1)
""" % MKL_RT
    # the oracle module sets the variable itself only when it is unset, and
    # it is imported after MKL already started: the layer stays threaded
    rc, out = _run(body, env_layer=None)
    assert rc == 0 and "REFUSED" in out, out


def test_helper_mkl_first_keeps_oracle_exact():
    body = """
from tests import _util as U
from tests.test_oracle import _cond_spd
U.tester_spd(16)                  # an MKL routine before the oracle
import oracle as ora
A = _cond_spd(512, 1e2, seed=7)
B = U.rhs(512, 5, seed=8)
assert ora.set_blas(ora.MKL) == ora.MKL
res = []
for t in (1, 8):
    ora.set_threads(t)
    res.append(ora.dsposv(ora.LOWER, A, B, 128))
ora.set_blas(ora.BUILTIN)
Xb, itb, infob = ora.dsposv(ora.LOWER, A, B, 128)
for X, it, info in res:
    assert info == 0 and infob == 0, (info, infob)
    assert it == itb == 2, (it, itb)
    assert np.abs(X - Xb).max() <= 1e-13 * np.abs(Xb).max()
print("OK")
"""
    rc, out = _run(body, env_layer=None)
    assert rc == 0 and "OK" in out, out
