"""CPU tests of the C-ABI boundary: the shared library loads and exports
every function the public headers declare, the headers compile as C and
C++, the ctypes mirror matches the C struct layout, the host-side descriptor
logic follows the reference (control/descriptor.c), and the library refuses
to initialise without a GPU instead of falling back to the CPU."""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as ora
import plasma_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(INC, "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
        for m in re.finditer(r"\b(plasma_\w+)\s*\(", text):
            name = m.group(1)
            before = text[:m.start()].rstrip()
            # a declaration starts a statement with a return type
            if re.search(r"(int|void|size_t|plasma_desc_t|plasma_enum_t|"
                         r"plasma_context_t\s*\*)\s*$", before):
                names.add(name)
    return names


def test_library_exports_every_declared_function():
    lib_syms = subprocess.run(["nm", "-D", "--defined-only", pa.LIB_PATH],
                              capture_output=True, text=True,
                              check=True).stdout
    exported = {l.split()[-1] for l in lib_syms.splitlines() if " T " in l}
    declared = _declared_functions()
    assert len(declared) >= 45
    missing = sorted(declared - exported)
    assert not missing, missing


def test_python_mirror_covers_headers():
    declared = _declared_functions()
    missing = sorted(n for n in declared if n not in pa.PROTOTYPES
                     and n != "plasma_context_self")
    assert not missing, missing
    lib = pa.load()
    for name in pa.PROTOTYPES:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("compiler", ["gcc -x c -std=c99",
                                      "g++ -x c++ -std=c++11"])
def test_headers_compile(tmp_path, compiler):
    src = tmp_path / "t.c"
    src.write_text('#include "plasma.h"\nint main(void) { return 0; }\n')
    cmd = compiler.split() + ["-Wall", "-Werror", "-I", INC, "-fsyntax-only",
                              str(src)]
    subprocess.run(cmd, check=True)


def test_desc_struct_layout(tmp_path):
    fields = [f[0] for f in pa.plasma_desc_t._fields_]
    src = tmp_path / "off.c"
    body = "".join('printf("%%zu ", offsetof(plasma_desc_t, %s));' % f
                   for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n'
                   '#include "plasma.h"\nint main(void){%s'
                   'printf("%%zu\\n", sizeof(plasma_desc_t));return 0;}\n'
                   % body)
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-I", INC, str(src), "-o", str(exe)], check=True)
    out = list(map(int, subprocess.run([str(exe)], capture_output=True,
                                       text=True).stdout.split()))
    want = [getattr(pa.plasma_desc_t, f).offset for f in fields]
    assert out[:-1] == want
    assert out[-1] == ctypes.sizeof(pa.plasma_desc_t)


def _offsets_from_header(tmp_path, kind, m, n, nb):
    """Tile offsets computed by include/plasma_descriptor.h on the host."""
    src = tmp_path / "tiles.c"
    src.write_text(r'''
#include <stdio.h>
#include <stdlib.h>
#include "plasma.h"
int main(int argc, char **argv) {
    int kind = atoi(argv[1]), m = atoi(argv[2]), n = atoi(argv[3]);
    int nb = atoi(argv[4]);
    plasma_desc_t A;
    if (kind == PlasmaGeneral)
        plasma_desc_general_init(PlasmaRealDouble, NULL, nb, nb, m, n, 0, 0,
                                 m, n, &A);
    else
        plasma_desc_triangular_init(PlasmaRealDouble, kind, NULL, nb, nb, m,
                                    n, 0, 0, m, n, &A);
    for (int j = 0; j < A.nt; j++)
        for (int i = 0; i < A.mt; i++) {
            if ((kind == PlasmaLower && j > i) || (kind == PlasmaUpper && i > j))
                continue;
            printf("%d %d %zu %d %d %d\n", i, j, plasma_tile_offset(A, i, j),
                   plasma_tile_mmain(A, i), plasma_tile_mview(A, i),
                   plasma_tile_nview(A, j));
        }
    return 0;
}
''')
    exe = tmp_path / "tiles"
    lib = os.path.dirname(pa.LIB_PATH)
    subprocess.run(["gcc", "-I", INC, str(src), "-o", str(exe), "-L", lib,
                    "-lplasma_amd", "-Wl,-rpath," + lib], check=True)
    out = subprocess.run([str(exe), str(kind), str(m), str(n), str(nb)],
                         capture_output=True, text=True, check=True).stdout
    return [tuple(map(int, l.split())) for l in out.splitlines()]


@pytest.mark.parametrize("kind", [pa.PlasmaGeneral, pa.PlasmaLower,
                                  pa.PlasmaUpper])
@pytest.mark.parametrize("m,n,nb", [(100, 100, 16), (257, 257, 64),
                                    (64, 64, 16), (70, 45, 16)])
def test_tile_addressing_matches_oracle_layout(tmp_path, kind, m, n, nb):
    """The product's tile offsets (host copy of the device formula) place
    every element where the oracle's tile image has it."""
    if kind != pa.PlasmaGeneral and m != n:
        pytest.skip("triangular descriptors are square")
    A = np.asfortranarray(np.arange(m * n, dtype=np.float64).reshape(n, m).T)
    oregion = {pa.PlasmaGeneral: 0, pa.PlasmaLower: ora.LOWER,
               pa.PlasmaUpper: ora.UPPER}[kind]
    img = ora.lapack2tile(oregion, A, nb)
    for (i, j, off, ld, mv, nv) in _offsets_from_header(tmp_path, kind, m, n, nb):
        tile = A[i * nb:i * nb + mv, j * nb:j * nb + nv]
        got = img[off:off + ld * nv].reshape(nv, ld).T[:mv, :]
        assert np.array_equal(got, tile), (i, j)


def test_host_descriptor_logic():
    lib = pa.load()
    d = pa.plasma_desc_t()
    assert lib.plasma_desc_general_init(pa.PlasmaRealDouble, None, 16, 16,
                                        100, 70, 0, 0, 100, 70,
                                        ctypes.byref(d)) == 0
    assert (d.mt, d.nt, d.gmt, d.gnt) == (7, 5, 7, 5)
    assert (d.A21, d.A12, d.A22) == (96 * 64, 96 * 64 + 4 * 64,
                                     96 * 64 + 4 * 64 + 96 * 6)
    assert lib.plasma_desc_check(d) == 0
    v = lib.plasma_desc_view(d, 16, 32, 50, 20)
    assert (v.i, v.j, v.m, v.n, v.mt, v.nt) == (16, 32, 50, 20, 4, 2)
    bad = pa.plasma_desc_t.from_buffer_copy(d)
    bad.i = 3  # not tile aligned (reference descriptor.c:309-312)
    assert lib.plasma_desc_check(bad) == pa.PlasmaErrorIllegalValue
    bad = pa.plasma_desc_t.from_buffer_copy(d)
    bad.m = 200
    assert lib.plasma_desc_check(bad) == pa.PlasmaErrorIllegalValue
    s, r = pa.plasma_sequence_t(), pa.plasma_request_t()
    assert lib.plasma_sequence_init(ctypes.byref(s)) == 0
    assert lib.plasma_request_init(ctypes.byref(r)) == 0
    assert lib.plasma_request_fail(ctypes.byref(s), ctypes.byref(r), -103) == -103
    assert s.status == r.status == -103
    assert lib.plasma_uplo_const(b"l") == pa.PlasmaLower
    assert lib.plasma_uplo_const(b"U") == pa.PlasmaUpper


def test_not_initialised_and_no_gpu_fallback():
    """Without plasma_init the setters report NotInitialized; on a host
    without a GPU plasma_init fails with PlasmaErrorEnvironment (there is
    no CPU path)."""
    code = r'''
import ctypes, plasma_amd as pa
lib = pa.load()
v = ctypes.c_int()
assert lib.plasma_set(pa.PlasmaNb, 128) == pa.PlasmaErrorNotInitialized
assert lib.plasma_get(pa.PlasmaNb, ctypes.byref(v)) == pa.PlasmaErrorNotInitialized
a = (ctypes.c_double * 4)()
assert lib.plasma_dposv(122, 2, 1, a, 2, a, 2) == pa.PlasmaErrorNotInitialized
rc = lib.plasma_init()
import torch
if not torch.cuda.is_available():
    assert rc == pa.PlasmaErrorEnvironment, rc
print("ok")
'''
    env = dict(os.environ, PYTHONPATH=ROOT)
    out = subprocess.run(["python", "-c", code], capture_output=True,
                         text=True, env=env, timeout=300)
    assert out.returncode == 0 and "ok" in out.stdout, out.stderr[-2000:]


def test_potrf_uninitialised_exits_like_reference():
    """plasma_dpotrf before plasma_init terminates with EXIT_FAILURE
    (reference compute/zpotrf.c:81 via plasma_fatal_error)."""
    code = ("import ctypes, plasma_amd as pa\n"
            "a = (ctypes.c_double * 4)()\n"
            "pa.load().plasma_dpotrf(122, 2, a, 2)\n"
            "print('returned')\n")
    env = dict(os.environ, PYTHONPATH=ROOT)
    out = subprocess.run(["python", "-c", code], capture_output=True,
                         text=True, env=env, timeout=300)
    assert out.returncode == 1 and "returned" not in out.stdout
    assert "PLASMA FATAL ERROR" in out.stderr
