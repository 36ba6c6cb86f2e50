"""Autotuned schedules: a table loaded with plasma_tuning_load sets nb,
update depth and lookahead of tuned factorizations (reference
plasma_tune_potrf, control/tuning.c:367-374), and the results stay those of
the oracle."""
import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa
from tests import _util as U

pytestmark = pytest.mark.gpu

L = pa.PlasmaLower


def test_table_applies(plasma, nb_guard, tmp_path):
    lib = pa.load()
    table = tmp_path / "t.txt"
    table.write_text("# routine precision n_max nb update_depth lookahead\n"
                     "potrf d 600 64 3 0\n"
                     "potrf d 5000 96 1 1\n"
                     "potrf s 5000 32 2 1\n")
    assert lib.plasma_tuning_load(str(table).encode()) == 0
    try:
        pa.set_param(pa.PlasmaTuning, pa.PlasmaEnabled)
        ora.set_blas(ora.MKL)
        for n, nb, depth, la in ((500, 64, 3, 0), (1000, 96, 1, 1),
                                 (7000, 96, 1, 1)):
            A = U.tester_spd(n, n)
            a = A.copy(order="F")
            assert pa.potrf(L, n, a, n) == 0
            assert pa.get_param(pa.PlasmaNb)[1] == nb
            assert pa.get_param(pa.PlasmaUpdateDepth)[1] == depth
            assert pa.get_param(pa.PlasmaLookahead)[1] == la
            if n <= 1000:
                ref, _ = ora.potrf(L, A, nb)
                assert U.factor_error(a, ref, True) <= 1e-12
        assert lib.plasma_tuning_load(str(tmp_path / "missing").encode()) \
            == pa.PlasmaErrorIllegalValue
    finally:
        lib.plasma_tuning_load(None)
        pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)


def test_disabled_keeps_context(plasma, nb_guard, tmp_path):
    lib = pa.load()
    table = tmp_path / "t.txt"
    table.write_text("potrf d 100000 64 1 0\n")
    assert lib.plasma_tuning_load(str(table).encode()) == 0
    try:
        pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
        pa.set_param(pa.PlasmaNb, 128)
        A = U.tester_spd(300, 300)
        a = A.copy(order="F")
        assert pa.potrf(L, 300, a, 300) == 0
        assert pa.get_param(pa.PlasmaNb)[1] == 128
    finally:
        lib.plasma_tuning_load(None)


def test_table_rows_per_launch_mode(plasma, nb_guard, tmp_path):
    """Rows tagged eager / graph apply in their launch mode
    (PlasmaGraphCapture); untagged rows serve both; a precision with rows of
    the other mode only falls back to them; an unknown tag drops the row."""
    lib = pa.load()
    table = tmp_path / "t.txt"
    table.write_text("potrf d 5000 64 3 0 eager\n"
                     "potrf d 5000 96 2 1 graph\n"
                     "potrf d 9000 32 1 1 sometimes\n"
                     "potrf s 5000 128 4 2 graph\n")
    assert lib.plasma_tuning_load(str(table).encode()) == 0
    try:
        pa.set_param(pa.PlasmaTuning, pa.PlasmaEnabled)
        ora.set_blas(ora.MKL)
        n = 700
        A = U.tester_spd(n, n)
        for graph, prec, want in ((0, "d", (64, 3, 0)), (1, "d", (96, 2, 1)),
                                  (0, "s", (128, 4, 2)),
                                  (1, "s", (128, 4, 2))):
            pa.set_param(pa.PlasmaGraphCapture,
                         pa.PlasmaEnabled if graph else pa.PlasmaDisabled)
            a = (A if prec == "d" else A.astype(np.float32)).copy(order="F")
            assert pa.potrf(L, n, a, n, prec=prec) == 0
            got = (pa.get_param(pa.PlasmaNb)[1],
                   pa.get_param(pa.PlasmaUpdateDepth)[1],
                   pa.get_param(pa.PlasmaLookahead)[1])
            assert got == want, (graph, prec, got)
            if prec == "d":
                ref, _ = ora.potrf(L, A, want[0])
                assert U.factor_error(a, ref, True) <= 1e-12
    finally:
        lib.plasma_tuning_load(None)
        pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
        pa.set_param(pa.PlasmaGraphCapture, pa.PlasmaDisabled)
