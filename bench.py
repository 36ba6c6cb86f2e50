#!/usr/bin/env python3
"""Benchmark of the Cholesky solve path on one MI355X.

Headline metric: time to an FP64 backward error for A X = B at n = 32768,
nrhs = 64, mixed-precision plasma_dsposv (FP32 tile Cholesky on MFMA, FP64
iterative refinement) against plain FP64 plasma_dposv, both reported as
FP64-equivalent TFLOP/s, flops_dpotrf(n) + flops_dpotrs(n, nrhs) per solve
(the reference tester's convention, test/flops.h:366-417 and
test/test_zcposv.c:314).

Matrices (synthetic, generated on the GPU, resident in HBM):
  * A = Q diag(s) Q^T, Q the product of 64 Householder reflectors with
    N(0,1) vectors, s log-spaced on [1/kappa, 1] (so ||A||_2 = 1 and
    cond_2(A) = kappa), kappa in {1e2, 1e4, 1e6}; the headline value is the
    mixed solve at the largest kappa;
  * A = G G^T / n + I (G iid N(0,1)) for the plain FP64 drop-in line.
B is iid N(0,1).  Every solve is checked afterwards with the tester's
scaled residual ||B - A X||_inf / (n ||A||_inf ||X||_inf eps64) < 10.

Timed steps run with PlasmaGraphCapture enabled (the factorization's
launch sequence replays from a captured hipGraph) and no per-launch
profiling; a second, profiled pass of the same solves with the same
schedules (capture stays enabled, so the library reads the same rows of its
tuned table; profiling bypasses the graphs), launched one by one with HIP
events around every launch, gives the per-kernel-class times and the
roofline.  Both times are printed.

One mixed step is one plasma_dsposv call on device-resident LAPACK arrays
(translation to tiles, FP32 factor, refinement, translation back, final
synchronisation).  One plain step is the same work as plasma_dposv on
device-resident arrays (tr2desc + ge2desc, tile Cholesky, two sweeps,
desc2tr + desc2ge); the input arrays are only read, so every step solves
the same system.

Also printed: per kappa, plasma_dhposv on the same system (bf16-input
factor of the equilibrated matrix, GMRES-IR; inner GMRES iterations, -1 for
the FP64 fallback); the dominant kernel's roofline (the FP32 trailing
update of the mixed solve; HIP events on the stream it runs on),
per-kernel-class times, and the CPU baseline: the oracle's dsposv (PLASMA's tile algorithm
restated, OpenMP tasks, sequential MKL tile kernels) timed on this host on
a smaller n.  Beside the headline: plain FP64 lines at n = 8192 / 16384,
the batched solver, the Cholesky adjoint, host-pointer rates, and the
complex precisions (zposv, cposv, zcposv at n = 16384 with the complex
update's roofline and the oracle's zposv on this host).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK = 78.6e12      # MI355X dense FP64 matrix peak (FLOP/s)
FP32_MFMA_PEAK = 157.3e12     # MI355X dense FP32 matrix peak (FLOP/s)
HALF_MFMA_PEAK = 2.5e15       # MI355X dense bf16 / fp16 matrix peak (FLOP/s)
# 16-bit MFMA products per FP32 product of the split updates
SPLIT_PRODUCTS = {"bf16": 6, "fp16": 3}
HBM_PEAK = 8.0e12             # bytes/s
EPS64 = 2.0 ** -53

PROF_CLASSES = ["trailing_update", "potrf_diag", "trsm_panel", "sweeps",
                "layout", "mixed", "adjoint"]


def flops_posv(n, nrhs):
    n, nrhs = float(n), float(nrhs)
    potrf = n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0
    potrs = 2.0 * n * n * nrhs
    return potrf + potrs


def householder_spd(n, kappa, nrefl=64, seed=0, device="cuda"):
    """A = Q diag(s) Q^T with Q = H_1 ... H_nrefl, H_i = I - tau_i v_i v_i^T,
    v_i iid N(0,1) (torch generator `seed`), s_i = kappa^(-i/(n-1)).

    Q is applied in compact WY form, Q = I - V T V^T (T upper triangular,
    the forward recurrence T(1:i-1, i) = -tau_i T(1:i-1, 1:i-1) V^T v_i), so
    A = D + W M W^T with W = [V, D V] and M = [[C, -T], [-T^T, 0]],
    C = T (V^T D V) T^T: one rank-2*nrefl update of the diagonal.  The
    result is symmetrised exactly.  Returns (A, s) as float64 tensors."""
    import torch
    f64 = torch.float64
    g = torch.Generator(device=device).manual_seed(seed)
    V = torch.randn(nrefl, n, dtype=f64, device=device, generator=g).T
    i = torch.arange(n, dtype=f64, device=device)
    s = torch.pow(torch.tensor(float(kappa), dtype=f64, device=device),
                  -i / max(n - 1, 1))
    gram = (V.T @ V).cpu()
    tau = 2.0 / gram.diagonal()
    T = torch.zeros(nrefl, nrefl, dtype=f64)
    for j in range(nrefl):
        T[j, j] = tau[j]
        if j:
            T[:j, j] = -tau[j] * (T[:j, :j] @ gram[:j, j])
    U = V * s[:, None]
    C = T @ (V.T @ U).cpu() @ T.T
    M = torch.zeros(2 * nrefl, 2 * nrefl, dtype=f64)
    M[:nrefl, :nrefl] = C
    M[:nrefl, nrefl:] = -T
    M[nrefl:, :nrefl] = -T.T
    W = torch.cat([V, U], dim=1)
    A = (W @ M.to(device)) @ W.T
    del W, U, V
    A.diagonal().add_(s)
    A = (A + A.T).mul_(0.5)
    return A, s


def ggt_spd(n, seed=0, device="cuda"):
    """A = G G^T / n + I, G iid N(0,1) (torch generator `seed`)."""
    import torch
    g = torch.Generator(device=device).manual_seed(seed)
    G = torch.randn(n, n, dtype=torch.float64, device=device, generator=g)
    A = (G @ G.T).div_(n)
    A.diagonal().add_(1.0)
    return A


def batched_systems(count, seed=0, device="cuda"):
    """The batched workload: `count` systems of order n_i ~ U{32..512},
    A_i = exp(-|x - y|^2 / (2 * 0.2^2)) on n_i points ~ U[0,1]^3 plus
    1e-3 I, one right-hand side ~ N(0,1); packed column major (ld n_i).
    Returns (n, A, aoff, B, boff) with A, B float64 tensors on the device
    and n, aoff, boff host numpy arrays."""
    import numpy as np
    import torch
    g = torch.Generator(device=device).manual_seed(seed)
    rng = np.random.default_rng(seed)
    n = rng.integers(32, 513, count).astype(np.int32)
    aoff = np.zeros(count, dtype=np.uint64)
    boff = np.zeros(count, dtype=np.uint64)
    aoff[1:] = np.cumsum(n.astype(np.uint64) ** 2)[:-1]
    boff[1:] = np.cumsum(n.astype(np.uint64))[:-1]
    A = torch.empty(int((n.astype(np.uint64) ** 2).sum()), dtype=torch.float64,
                    device=device)
    B = torch.randn(int(n.sum()), dtype=torch.float64, device=device,
                    generator=g)
    for i in range(count):
        k = int(n[i])
        x = torch.rand(k, 3, dtype=torch.float64, device=device, generator=g)
        K = torch.exp(-torch.cdist(x, x).pow_(2) / (2 * 0.2 ** 2))
        K.diagonal().add_(1e-3)
        A[int(aoff[i]):int(aoff[i]) + k * k] = K.reshape(-1)
    return n, A, aoff, B, boff


def scaled_residual(A, X, B):
    """||B - A X||_inf / (n ||A||_inf ||X||_inf eps64); X, B column-major
    n x nrhs stored as (nrhs, n) tensors, A symmetric."""
    n = A.shape[0]
    R = B.T - A @ X.T
    an = A.abs().sum(1).max().item()
    xn = X.T.abs().sum(1).max().item()
    rn = R.abs().sum(1).max().item()
    return rn / (n * an * xn) / EPS64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--nb", type=int, default=0,
                    help="pin the tile size (with --lookahead, update depth "
                         "2); 0 keeps the library's autotuned schedules")
    ap.add_argument("--lookahead", type=int, default=1)
    ap.add_argument("--kappa", type=str, default="1e2,1e4,1e6",
                    help="condition numbers of the mixed-precision sweep; "
                         "the headline value is the last one")
    ap.add_argument("--no-ggt", action="store_true",
                    help="skip the plain FP64 line on A = G G^T/n + I")
    ap.add_argument("--cpu-n", type=int, default=16384,
                    help="order of the CPU-baseline sample (0 skips it)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--batch", type=int, default=4096,
                    help="systems of the batched line (0 skips it)")
    ap.add_argument("--adjoint-n", type=str, default="4096,16384",
                    help="orders of the Cholesky-adjoint lines (empty "
                         "skips them)")
    ap.add_argument("--split-update", choices=("bf16", "fp16", "off"),
                    default=None,
                    help="FP32 trailing updates on 16-bit MFMA from exact "
                         "operand splits: three bf16 terms (six products) "
                         "or row-scaled two-term fp16 (three products); "
                         "default: the library's setting")
    ap.add_argument("--config2", type=str, default="8192,16384",
                    help="orders of the plain FP64 dpotrf+dpotrs lines on "
                         "G G^T/n + I (tuned nb; empty skips them)")
    ap.add_argument("--complex-n", type=int, default=16384,
                    help="order of the complex-precision lines (zposv, "
                         "cposv, zcposv; 0 skips them)")
    ap.add_argument("--complex-cpu-n", type=int, default=6144,
                    help="order of the oracle zposv CPU sample of the "
                         "complex lines (0 skips it)")
    ap.add_argument("--no-half", action="store_true",
                    help="skip the half-precision-input GMRES-IR solves "
                         "(plasma_dhposv) of the kappa sweep")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-pointer (PCIe-inclusive) timings")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the solves, copy MIB MiB device to device "
                         "once (a kernel of known traffic for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def host_cpu():
    """Model, sockets and cores of this host (lscpu), the CPU share this
    process may use, and the BLAS the oracle's tile kernels call."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket",
                     "Thread(s) per core", "CPU(s)"):
                info[k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    info["affinity_cpus"] = len(os.sched_getaffinity(0))
    try:
        lib = ctypes.CDLL("/opt/conda/lib/libmkl_rt.so.1")
        buf = ctypes.create_string_buffer(256)
        lib.mkl_get_version_string(buf, 256)
        info["blas"] = buf.value.decode().strip()
    except (OSError, AttributeError):
        info["blas"] = "built-in loops"
    return info


def cpu_baseline(n, nrhs, nb, kappa, threads):
    """The oracle (PLASMA's tile algorithms restated: OpenMP tasks,
    sequential MKL tile kernels) on this host, on the same kind of matrices
    at order n: dsposv on the Householder matrix of condition kappa and
    plain dposv on G G^T/n + I; one warm-up and one timed run each (a
    bounded sample, about 10-20 s)."""
    import numpy as np
    import oracle as ora
    backend = ora.set_blas(ora.MKL)
    if threads <= 0:
        # the GPU box's CPU share for one GPU is 16 cores, whatever the
        # affinity mask shows
        threads = min(16, len(os.sched_getaffinity(0)))
    ora.set_threads(threads)
    import torch
    A, _ = householder_spd(n, kappa, seed=1)
    g = torch.Generator(device="cuda").manual_seed(2)
    B = torch.randn(nrhs, n, dtype=torch.float64, device="cuda", generator=g)
    A = np.asfortranarray(A.cpu().numpy())
    B = np.asfortranarray(B.cpu().numpy().T)
    G = np.asfortranarray(ggt_spd(n, seed=3).cpu().numpy())
    res = {}
    for name in ("dsposv", "dposv"):
        ts = []
        for _ in range(2):
            t0 = time.perf_counter()
            if name == "dsposv":
                _, it, info = ora.dsposv(ora.LOWER, A, B, nb)
            else:
                _, _, info = ora.posv(ora.LOWER, G, B, nb)
            ts.append(time.perf_counter() - t0)
            assert info == 0
        res[name] = (ts[-1], it if name == "dsposv" else None)
    t, iters = res["dsposv"]
    tp, _ = res["dposv"]
    cpu = host_cpu()
    return {
        "value": flops_posv(n, nrhs) / t / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": "oracle dsposv n=%d nrhs=%d nb=%d lower, Householder "
                  "A cond %.0e, OpenMP tile tasks on %d threads, %s "
                  "sequential tile kernels, one run after a warm-up "
                  "(%.2f s per solve, %d refinement iterations)"
                  % (n, nrhs, nb, kappa, threads,
                     "MKL" if backend == ora.MKL else "built-in", t, iters),
        "plain_dposv": {"value": flops_posv(n, nrhs) / tp / 1e12,
                        "unit": "TFLOP/s", "seconds": round(tp, 3),
                        "workload": "oracle dposv n=%d nrhs=%d nb=%d on "
                                    "G G^T/n + I" % (n, nrhs, nb)},
        "host": cpu,
    }


class Bench:
    def __init__(self, args):
        import torch
        import plasma_amd as pa
        self.torch, self.pa, self.args = torch, pa, args
        self.lib = lib = pa.load()
        rc = lib.plasma_init()
        if rc != 0:
            raise SystemExit("plasma_init failed: %d" % rc)
        if args.split_update is not None:
            pa.set_param(pa.PlasmaSplitUpdate,
                         {"bf16": pa.PlasmaSplitBf16,
                          "fp16": pa.PlasmaSplitFp16,
                          "off": pa.PlasmaDisabled}[args.split_update])
        if args.nb > 0:
            pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
            pa.set_param(pa.PlasmaNb, args.nb)
            pa.set_param(pa.PlasmaLookahead, args.lookahead)
            pa.set_param(pa.PlasmaUpdateDepth, 2)

    def split_mode(self):
        """The FP32 update the library runs: "bf16", "fp16" or None (FP32
        MFMA); plasma_get reports the split in force."""
        pa = self.pa
        v = pa.get_param(pa.PlasmaSplitUpdate)[1]
        return {pa.PlasmaSplitFp16: "fp16",
                pa.PlasmaSplitBf16: "bf16"}.get(v)

    def schedule(self, prec, n):
        """(nb, update depth, lookahead) of a factorization of order n: the
        pinned one, or the autotuned table's row the library applies in
        graph mode, the mode of every solve here (among the rows of the
        precision for graph replay or for any mode, else the eager rows: the
        first with n <= n_max, else the last)."""
        if self.args.nb > 0:
            return self.args.nb, 2, self.args.lookahead
        path = os.environ.get("PLASMA_TUNING_FILE") or os.path.join(
            ROOT, "plasma_amd", "tuning", "mi355x.txt")
        rows, eager = [], []
        for line in open(path):
            f = line.split()
            if 6 <= len(f) <= 8 and f[0] == "potrf" and f[1] == prec:
                r = tuple(int(x) for x in f[2:6])
                (eager if f[6:7] == ["eager"] else rows).append(r)
        rows = sorted(rows or eager)
        if not rows:
            raise SystemExit("no %s rows in %s" % (prec, path))
        for n_max, nb, depth, la in rows:
            if n <= n_max:
                return nb, depth, la
        return rows[-1][1:]

    def prof_reset(self):
        self.lib.plasma_profile_enable(0)
        self.lib.plasma_profile_enable(1)

    def prof_read(self, steps):
        lib = self.lib
        out = {}
        for cls, name in enumerate(PROF_CLASSES):
            a, f, b, c = (ctypes.c_double(), ctypes.c_double(),
                          ctypes.c_double(), ctypes.c_long())
            lib.plasma_profile_query(cls, ctypes.byref(a), ctypes.byref(f),
                                     ctypes.byref(b), ctypes.byref(c))
            if c.value:
                out[name] = {"ms": a.value, "flops": f.value,
                             "bytes": b.value, "launches": c.value}
        lib.plasma_profile_enable(0)
        return out

    def timed(self, step, label, profile=True):
        torch, args = self.torch, self.args
        for _ in range(args.warmup):
            step()
        print("bench: %s warmup done" % label, file=sys.stderr, flush=True)
        if profile:
            self.prof_reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            step()
            print("bench: %s step %d done" % (label, i), file=sys.stderr,
                  flush=True)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        return ms, (self.prof_read(args.steps) if profile else {})

    def plain(self, A0, B0, label, profile=True):
        """Plain FP64 dposv on device-resident LAPACK arrays."""
        torch, pa, lib = self.torch, self.pa, self.lib
        n, nrhs = A0.shape[0], B0.shape[0]
        # the tile interface takes nb from the descriptor: apply the
        # schedule the LAPACK-layout driver would pick
        nb, depth, la = self.schedule("d", n)
        pa.set_param(pa.PlasmaUpdateDepth, depth)
        pa.set_param(pa.PlasmaLookahead, la)
        Aout = torch.empty_like(A0)
        X = torch.empty_like(B0)
        A, B = pa.plasma_desc_t(), pa.plasma_desc_t()

        def step():
            assert lib.plasma_desc_triangular_create(
                pa.PlasmaRealDouble, pa.PlasmaLower, nb, nb, n, n, 0, 0, n,
                n, ctypes.byref(A)) == 0
            assert lib.plasma_desc_general_create(
                pa.PlasmaRealDouble, nb, nb, n, nrhs, 0, 0, n, nrhs,
                ctypes.byref(B)) == 0
            seq, req = pa.plasma_sequence_t(), pa.plasma_request_t()
            lib.plasma_sequence_init(ctypes.byref(seq))
            lib.plasma_request_init(ctypes.byref(req))
            s, r = ctypes.byref(seq), ctypes.byref(req)
            lib.plasma_omp_dtr2desc(A0.data_ptr(), n, A, s, r)
            lib.plasma_omp_dge2desc(B0.data_ptr(), n, B, s, r)
            lib.plasma_omp_dposv(pa.PlasmaLower, A, B, s, r)
            lib.plasma_omp_ddesc2tr(A, Aout.data_ptr(), n, s, r)
            lib.plasma_omp_ddesc2ge(B, X.data_ptr(), n, s, r)
            status = lib.plasma_sequence_sync(s)
            lib.plasma_desc_destroy(ctypes.byref(A))
            lib.plasma_desc_destroy(ctypes.byref(B))
            if status != 0:
                raise SystemExit("plasma dposv failed: %d" % status)

        ms, prof = self.timed(step, label, profile)
        del Aout
        res = scaled_residual(A0, X, B0)
        return ms, prof, res

    def mixed(self, A0, B0, label, profile=True, routine="plasma_dsposv"):
        """plasma_dsposv (or plasma_dhposv, same arguments) on
        device-resident LAPACK arrays."""
        torch, pa, lib = self.torch, self.pa, self.lib
        fn = getattr(lib, routine)
        n, nrhs = A0.shape[0], B0.shape[0]
        X = torch.empty_like(B0)
        it = ctypes.c_int(0)
        iters = []

        def step():
            rc = fn(pa.PlasmaLower, n, nrhs, A0.data_ptr(), n,
                    B0.data_ptr(), n, X.data_ptr(), n, ctypes.byref(it))
            if rc != 0:
                raise SystemExit("%s failed: %d" % (routine, rc))
            iters.append(it.value)

        ms, prof = self.timed(step, label, profile)
        res = scaled_residual(A0, X, B0)
        return ms, prof, res, iters[-1]


def host_line(bench, A0, B0, kappa):
    """The drop-in rate with host arrays (PCIe-inclusive): plasma_dposv and
    plasma_dsposv on page-locked host copies of the headline system, the
    library staging A and B through HBM.  One warm-up, then the mean of
    `steps` calls (the input copied back between calls, untimed)."""
    import torch
    pa, lib, steps = bench.pa, bench.lib, bench.args.steps
    n, nrhs = A0.shape[0], B0.shape[0]
    hA0 = torch.empty(A0.shape, dtype=A0.dtype, pin_memory=True)
    hB0 = torch.empty(B0.shape, dtype=B0.dtype, pin_memory=True)
    hA0.copy_(A0)
    hB0.copy_(B0)
    hA = torch.empty_like(hA0, pin_memory=True)
    hB = torch.empty_like(hB0, pin_memory=True)
    hX = torch.empty_like(hB0, pin_memory=True)
    it = ctypes.c_int(0)
    out = {"workload": "headline system (cond %.0e) with page-locked host "
                       "arrays: the library copies A and B to HBM and the "
                       "results back inside the timed call" % kappa}
    for name in ("plain", "mixed"):
        ts = []
        for rep in range(steps + 1):
            hA.copy_(hA0)
            hB.copy_(hB0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if name == "plain":
                rc = lib.plasma_dposv(pa.PlasmaLower, n, nrhs, hA.data_ptr(),
                                      n, hB.data_ptr(), n)
            else:
                rc = lib.plasma_dsposv(pa.PlasmaLower, n, nrhs,
                                       hA.data_ptr(), n, hB.data_ptr(), n,
                                       hX.data_ptr(), n, ctypes.byref(it))
            ts.append(time.perf_counter() - t0)
            if rc != 0:
                raise SystemExit("host-pointer %s solve failed: %d"
                                 % (name, rc))
        t = sum(ts[1:]) / steps
        out[name + "_ms"] = round(t * 1e3, 2)
        out[name + "_tflops"] = round(flops_posv(n, nrhs) / t / 1e12, 3)
        print("bench: host-pointer %s done" % name, file=sys.stderr,
              flush=True)
    del hA0, hB0, hA, hB, hX
    return out


def config2_line(bench, orders, nrhs):
    """Plain FP64 plasma_dposv (dpotrf + dpotrs) on A = G G^T/n + I at the
    given orders, autotuned schedule, graph mode: rate against the FP64
    matrix peak."""
    import torch
    out = {}
    g = torch.Generator(device="cuda").manual_seed(7)
    for n in orders:
        A0 = ggt_spd(n, seed=n)
        B0 = torch.randn(nrhs, n, dtype=torch.float64, device="cuda",
                         generator=g)
        torch.cuda.synchronize()
        bench.pa.set_param(bench.pa.PlasmaGraphCapture,
                           bench.pa.PlasmaEnabled)
        ms, _, res = bench.plain(A0, B0, "config2 n=%d" % n, profile=False)
        tf = flops_posv(n, nrhs) / (ms * 1e-3)
        out["n=%d" % n] = {
            "ms": round(ms, 3), "tflops": round(tf / 1e12, 3),
            "frac_fp64_peak": round(tf / FP64_MFMA_PEAK, 4),
            "residual_eps": round(res, 4),
            "schedule": dict(zip(("nb", "update_depth", "lookahead"),
                                 bench.schedule("d", n)))}
        del A0, B0
        torch.cuda.empty_cache()
    return out


def complex_line(bench, n, nrhs, steps, cpu_n, threads):
    """The complex precisions on a Hermitian positive definite A = G G^H/n
    + I (G complex N(0,1)), nrhs complex N(0,1) right-hand sides, lower,
    device-resident: plasma_zposv and plasma_cposv (rates in the
    reference's complex flop count, zpotrf 4 n^3 / 3 + zpotrs 8 n^2 nrhs,
    against the FP64 / FP32 matrix peak that their four-real-product MFMA
    updates run on), plasma_zcposv (same flop count, its speedup over
    zposv), the update kernel's roofline from a profiled pass, and the
    oracle's zposv timed on this host at order cpu_n."""
    import torch
    pa = bench.pa
    lib = bench.lib
    fl = 4.0 * n ** 3 / 3.0 + 8.0 * n * n * nrhs
    out = {"workload": "Hermitian A = G G^H/n + I (G complex N(0,1)), "
                       "n=%d, nrhs=%d, lower; B complex N(0,1)" % (n, nrhs)}
    g = torch.Generator(device="cuda").manual_seed(11)

    def herm(dt):
        G = torch.randn(n, n, dtype=dt, device="cuda", generator=g)
        A = (G @ G.conj().T) / n
        A.diagonal().add_(1.0)
        # column major on the device: row j of the tensor is column j
        return A.T.contiguous()

    def resid(Ac, X, B, eps):
        """the tester's scaled residual in units of the solve's eps"""
        A = Ac.T.to(torch.complex128)
        Xt, Bt = X.T.to(torch.complex128), B.T.to(torch.complex128)
        R = Bt - A @ Xt
        return (R.abs().sum(1).max() / (n * A.abs().sum(1).max()
                * Xt.abs().sum(1).max() * eps)).item()

    zms = None
    for prec, dt, peak in (("z", torch.complex128, FP64_MFMA_PEAK),
                           ("c", torch.complex64, FP32_MFMA_PEAK)):
        A0 = herm(dt)
        B0 = torch.randn(nrhs, n, dtype=dt, device="cuda", generator=g)
        W, X = torch.empty_like(A0), torch.empty_like(B0)
        posv = getattr(lib, "plasma_%sposv" % prec)
        times = []
        for it in range(steps + 2):
            W.copy_(A0)
            X.copy_(B0)
            torch.cuda.synchronize()
            if it == steps + 1:
                bench.prof_reset()
            t0 = time.perf_counter()
            rc = posv(pa.PlasmaLower, n, nrhs, W.data_ptr(), n, X.data_ptr(),
                      n)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            if rc != 0:
                raise SystemExit("%sposv failed: %d" % (prec, rc))
            if it == steps + 1:
                prof = bench.prof_read(1)
            elif it > 0:
                times.append(t1 - t0)
            print("bench: %sposv n=%d step %d done" % (prec, n, it),
                  file=sys.stderr, flush=True)
        ms = sum(times) / len(times) * 1e3
        if prec == "z":
            zms = ms
        out[prec + "posv"] = {
            "ms": round(ms, 3), "tflops": round(fl / (ms * 1e-3) / 1e12, 3),
            "frac_peak": round(fl / (ms * 1e-3) / peak, 4),
            "residual_eps": round(resid(A0, X, B0, 2.0 ** -53 if prec == "z"
                                        else 2.0 ** -24), 4),
            "roofline": roofline(prof, "cjobs_kernel<UpdateDec> (complex "
                                 "herk + gemm, four real MFMA products per "
                                 "complex product)", peak),
            "kernels": classes(prof, 1)}
        del A0, B0, W, X
        torch.cuda.empty_cache()
    A0 = herm(torch.complex128)
    B0 = torch.randn(nrhs, n, dtype=torch.complex128, device="cuda",
                     generator=g)
    X = torch.empty_like(B0)
    it_c = ctypes.c_int(0)
    times = []
    for it in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc = lib.plasma_zcposv(pa.PlasmaLower, n, nrhs, A0.data_ptr(), n,
                               B0.data_ptr(), n, X.data_ptr(), n,
                               ctypes.byref(it_c))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if rc != 0:
            raise SystemExit("zcposv failed: %d" % rc)
        if it > 0:
            times.append(t1 - t0)
        print("bench: zcposv n=%d step %d done" % (n, it), file=sys.stderr,
              flush=True)
    ms = sum(times) / len(times) * 1e3
    out["zcposv"] = {"ms": round(ms, 3),
                     "tflops": round(fl / (ms * 1e-3) / 1e12, 3),
                     "iterations": it_c.value,
                     "residual_eps": round(resid(A0, X, B0, 2.0 ** -53), 4),
                     "speedup_vs_zposv": round(zms / ms, 3)}
    del A0, B0, X
    torch.cuda.empty_cache()
    if cpu_n > 0:
        import numpy as np
        import oracle as ora
        if threads <= 0:
            threads = min(16, len(os.sched_getaffinity(0)))
        ora.set_threads(threads)
        rng = np.random.default_rng(0)
        G = rng.standard_normal((cpu_n, cpu_n)) + \
            1j * rng.standard_normal((cpu_n, cpu_n))
        A = G @ G.conj().T / cpu_n + np.eye(cpu_n)
        B = rng.standard_normal((cpu_n, nrhs)) + \
            1j * rng.standard_normal((cpu_n, nrhs))
        del G
        t0 = time.perf_counter()
        _, _, info = ora.zposv(ora.LOWER, A, B, 128)
        t1 = time.perf_counter()
        cfl = 4.0 * cpu_n ** 3 / 3.0 + 8.0 * cpu_n * cpu_n * nrhs
        out["cpu_baseline"] = {
            "value": round(cfl / (t1 - t0) / 1e12, 4), "unit": "TFLOP/s",
            "cores": threads, "kind": "port",
            "sample": "oracle zposv (PLASMA's complex tile algorithm "
                      "restated, OpenMP tasks, built-in complex tile loops) "
                      "n=%d nrhs=%d nb=128, one run, %.1f s, info %d"
                      % (cpu_n, nrhs, t1 - t0, info)}
    return out


def classes(prof, steps):
    out = {}
    for name, p in prof.items():
        a, f, b = p["ms"], p["flops"], p["bytes"]
        out[name] = {"ms_per_step": round(a / steps, 3),
                     "launches_per_step": p["launches"] // steps,
                     "tflops": round(f / (a * 1e-3) / 1e12, 2)
                     if a and f else None,
                     "gbps": round(b / (a * 1e-3) / 1e9, 1)
                     if a and b else None}
    return out


def batched_line(bench, count, steps):
    """plasma_{d,s}posv_vbatched on the batched workload: mean time of one
    launch over `steps` calls (inputs restored between calls, untimed),
    FP64-equivalent rate sum(n^3/3 + n^2/2 + n/6 + 2 n^2), and the largest
    scaled residual of 64 sampled systems."""
    import numpy as np
    import torch
    pa, lib = bench.pa, bench.lib
    n, A0, aoff, B0, boff = batched_systems(count, seed=3)
    nf = n.astype(np.float64)
    flops = float((nf ** 3 / 3 + nf ** 2 / 2 + nf / 6 + 2 * nf ** 2).sum())
    info = np.zeros(count, dtype=np.int32)
    out = {"workload": "%d SPD systems, n ~ U{32..512}, RBF(0.2) on U[0,1]^3 "
                       "+ 1e-3 I, one right-hand side each" % count}
    for prec, dt in (("d", torch.float64), ("s", torch.float32)):
        # separate working copies (to() returns the input itself when the
        # type already matches, and the solve overwrites A and B)
        A = torch.empty(A0.shape, dtype=dt, device=A0.device)
        B = torch.empty(B0.shape, dtype=dt, device=B0.device)
        fn = getattr(lib, "plasma_%sposv_vbatched" % prec)
        ts = []
        for it in range(steps + 1):
            A.copy_(A0)
            B.copy_(B0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rc = fn(pa.PlasmaLower, count, n.ctypes.data, A.data_ptr(),
                    aoff.ctypes.data, 1, B.data_ptr(), boff.ctypes.data,
                    info.ctypes.data)
            ts.append(time.perf_counter() - t0)
            if rc != 0 or info.any():
                raise SystemExit("batched %sposv failed: %d" % (prec, rc))
        t = sum(ts[1:]) / steps
        worst = 0.0
        eps = 2.0 ** -53 if prec == "d" else 2.0 ** -24
        for i in np.linspace(0, count - 1, 64).astype(int):
            k, ao, bo = int(n[i]), int(aoff[i]), int(boff[i])
            M = A0[ao:ao + k * k].reshape(k, k)
            x = B[bo:bo + k].to(torch.float64)
            r = B0[bo:bo + k] - M @ x
            res = (r.abs().max() / (k * M.abs().sum(1).max() *
                                    x.abs().max())).item() / eps
            worst = max(worst, res)
        # algorithmic traffic: each system's stored triangle read and its
        # factor written (n^2 elements), its right-hand side read and
        # written
        es = 8 if prec == "d" else 4
        nbytes = float(((nf ** 2) + 2 * nf).sum()) * es
        peak = FP64_MFMA_PEAK if prec == "d" else FP32_MFMA_PEAK
        out[{"d": "fp64", "s": "fp32"}[prec]] = {
            "ms": round(t * 1e3, 3),
            "tflops": round(flops / t / 1e12, 3),
            "frac_mfma_peak": round(flops / t / peak, 4),
            "gbps": round(nbytes / t / 1e9, 1),
            "frac_hbm": round(nbytes / t / HBM_PEAK, 4),
            "max_residual_eps": round(worst, 4)}
        del A, B
    return out


def gp_kernel(n, seed=0, dtype=None):
    """K = RBF(X, lengthscale 1) + 0.1 I with X (n x 8) iid N(0,1), and
    y ~ N(0,1): the Gaussian-process log-likelihood workload."""
    import torch
    dtype = dtype or torch.float64
    g = torch.Generator(device="cuda").manual_seed(seed)
    X = torch.randn(n, 8, dtype=torch.float64, device="cuda", generator=g)
    sq = (X * X).sum(1)
    D2 = (sq[:, None] + sq[None, :] - 2.0 * (X @ X.T)).clamp_min(0.0)
    K = torch.exp(-0.5 * D2)
    K.diagonal().add_(0.1)
    y = torch.randn(1, n, dtype=torch.float64, device="cuda", generator=g)
    return K.to(dtype).contiguous(), y.to(dtype).contiguous()


def adjoint_line(bench, n, steps):
    """Gaussian-process log-likelihood and its gradient dL/dK on the
    device: forward = potrf + log det + potrs, backward =
    plasma_?potrf_adjoint (the log-det term, Lbar = -diag(1/L_ii)) +
    plasma_?potrs_adjoint (the quadratic term, Xbar = -y/2).  Mean times
    over `steps` calls; the check is max |G K + I/2 - a y^T / 2| (the
    closed form 1/2 (a a^T - K^{-1}) multiplied by K)."""
    import torch
    pa, lib = bench.pa, bench.lib
    out = {"workload": "GP log-likelihood gradient, K = RBF(X (n x 8) "
                       "N(0,1), lengthscale 1) + 0.1 I, n = %d, y ~ N(0,1)"
                       % n}
    for prec, dt in (("d", torch.float64), ("s", torch.float32)):
        K, y = gp_kernel(n, seed=5, dtype=dt)
        F = torch.empty_like(K)
        a = torch.empty_like(y)
        G = torch.empty_like(K)
        xb = torch.empty_like(y)
        potrf = getattr(lib, "plasma_%spotrf" % prec)
        potrs = getattr(lib, "plasma_%spotrs" % prec)
        fadj = getattr(lib, "plasma_%spotrf_adjoint" % prec)
        sadj = getattr(lib, "plasma_%spotrs_adjoint" % prec)
        tf, tb = [], []
        for it in range(steps + 1):
            F.copy_(K)
            a.copy_(y)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rc = potrf(pa.PlasmaLower, n, F.data_ptr(), n)
            rc = rc or potrs(pa.PlasmaLower, n, 1, F.data_ptr(), n,
                             a.data_ptr(), n)
            d = F.diagonal()
            loglik = (-0.5 * (y * a).sum() - d.log().sum()).item()
            t1 = time.perf_counter()
            if rc != 0:
                raise SystemExit("GP forward failed: %d" % rc)
            G.zero_()
            G.diagonal().copy_(-1.0 / d)
            xb.copy_(y).mul_(-0.5)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            rc = fadj(pa.PlasmaLower, n, F.data_ptr(), n, G.data_ptr(), n)
            rc = rc or sadj(pa.PlasmaLower, n, 1, F.data_ptr(), n,
                            a.data_ptr(), n, xb.data_ptr(), n, G.data_ptr(),
                            n)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            if rc != 0:
                raise SystemExit("GP backward failed: %d" % rc)
            tf.append(t1 - t0)
            tb.append(t3 - t2)
            print("bench: adjoint %s step %d done" % (prec, it),
                  file=sys.stderr, flush=True)
        f_ms = sum(tf[1:]) / steps * 1e3
        b_ms = sum(tb[1:]) / steps * 1e3
        Gd, Kd = G.double(), K.double()
        R = Gd @ Kd - 0.5 * (a.double().T @ y.double())
        R.diagonal().add_(0.5)
        out[{"d": "fp64", "s": "fp32"}[prec]] = {
            "forward_ms": round(f_ms, 2), "backward_ms": round(b_ms, 2),
            "backward_over_forward": round(b_ms / f_ms, 2),
            "loglik": loglik,
            "check_max_abs": float(R.abs().max().item())}
        del K, F, G, Gd, Kd, R
        torch.cuda.empty_cache()
    return out


def pmc_traffic(name):
    """HBM bytes per launch of a kernel, from the summary of the two
    rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE, each scaled by the same
    counter's reading for a copy of known size) kept in profiles/<name>.md
    as a line 'traffic_bytes_per_launch: <bytes>'; None when absent."""
    import re
    path = os.path.join(ROOT, "profiles", name + ".md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", text, re.M)
    return float(m.group(1)) if m else None


def roofline(prof, kernel, peak, traffic=None):
    p = prof.get("trailing_update")
    if not p:
        return None
    achieved = p["flops"] / (p["ms"] * 1e-3)
    return {"kernel": kernel, "bound": "mfma",
            "achieved": round(achieved / 1e12, 3),
            "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "avg_launch_ms": round(p["ms"] / p["launches"], 4),
            "traffic": traffic}


def split_roofline(r, split):
    """The FP32 update's roofline entry: FP32-algorithmic flops per launch
    over the launch time.  On FP32 MFMA the peak is the FP32 matrix peak.
    A split update runs each FP32 product as p 16-bit products (six bf16 or
    three fp16), so its ceiling is the 16-bit dense peak / p in
    FP32-algorithmic flops, and frac is also the fraction of the 16-bit
    peak its issued MFMA work reaches (reported beside it)."""
    if r is None:
        return None
    if not split:
        r["kernel"] += ": tg::kernel<float,...,UpdateDecode> on FP32 MFMA"
        return r
    if split == "bf16":
        r["kernel"] += (": split_update_kernel, operands split into three "
                        "bf16 terms, v_mfma_f32_16x16x32_bf16")
    else:
        r["kernel"] += (": split16_update_kernel, row-scaled operands split "
                        "into two fp16 terms, global_load_lds_dwordx4 "
                        "fragments, v_mfma_f32_16x16x32_f16")
    products = SPLIT_PRODUCTS[split]
    work = r["achieved"] * 1e12 * products
    r["peak"] = round(HALF_MFMA_PEAK / products / 1e12, 1)
    r["peak_basis"] = ("%s dense MFMA peak %.0f TFLOP/s / %d products per "
                       "FP32 product" % (split, HALF_MFMA_PEAK / 1e12,
                                         products))
    r["frac"] = round(r["achieved"] * 1e12 * products / HALF_MFMA_PEAK, 4)
    r["fp32_mfma_peak_frac"] = round(r["achieved"] * 1e12 / FP32_MFMA_PEAK, 4)
    r["mfma_work"] = {"unit": "TFLOP/s", "precision": split,
                      "products_per_fp32_product": products,
                      "achieved": round(work / 1e12, 1),
                      "peak": HALF_MFMA_PEAK / 1e12,
                      "frac": round(work / HALF_MFMA_PEAK, 4)}
    return r


def orders(text):
    return [int(x) for x in text.split(",") if x.strip()]


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this benchmark runs on one GPU; --gpus %d is not "
              "supported (multi-GPU sharding is design-only)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    bench = Bench(args)
    n, nrhs = args.n, args.nrhs
    fl = flops_posv(n, nrhs)
    kappas = [float(k) for k in args.kappa.split(",") if k]
    g = torch.Generator(device="cuda").manual_seed(1)
    B0 = torch.randn(nrhs, n, dtype=torch.float64, device="cuda", generator=g)
    torch.cuda.synchronize()

    sweep = {}
    head = None
    host = None
    for kappa in kappas:
        A0, _ = householder_spd(n, kappa, seed=0)
        torch.cuda.synchronize()
        tag = "kappa=%.0e" % kappa
        # timed: the library replaying its captured factorization graphs,
        # no per-launch profiling (the library's fastest mode)
        bench.pa.set_param(bench.pa.PlasmaGraphCapture,
                           bench.pa.PlasmaEnabled)
        ms_m, _, res_m, iters = bench.mixed(A0, B0, "mixed " + tag,
                                            profile=False)
        ms_p, _, res_p = bench.plain(A0, B0, "plain " + tag, profile=False)
        # the half-precision-input factors with GMRES-IR on the same system
        half = {}
        for fmt, val in () if args.no_half else (
                ("bf16", bench.pa.PlasmaSplitBf16),
                ("fp16", bench.pa.PlasmaSplitFp16)):
            bench.pa.set_param(bench.pa.PlasmaHalfInput, val)
            half[fmt] = bench.mixed(A0, B0, "%s gmres-ir %s" % (fmt, tag),
                                    profile=False, routine="plasma_dhposv")
        # profiled: the same solves and schedules (graph capture stays
        # enabled, so the library picks the same rows of its table) launched
        # one by one, since profiling bypasses the graphs; every launch
        # bracketed by HIP events (kernel classes and the roofline)
        pms_m, prof_m, _, _ = bench.mixed(A0, B0, "mixed profiled " + tag)
        pms_p, prof_p, _ = bench.plain(A0, B0, "plain profiled " + tag)
        if kappa == kappas[-1] and not args.no_host:
            host = host_line(bench, A0, B0, kappa)
        del A0
        sweep[tag] = {
            "mixed_tflops": round(fl / (ms_m * 1e-3) / 1e12, 3),
            "mixed_ms": round(ms_m, 3),
            "iterations": iters,
            "mixed_residual_eps": round(res_m, 4),
            "plain_tflops": round(fl / (ms_p * 1e-3) / 1e12, 3),
            "plain_ms": round(ms_p, 3),
            "plain_residual_eps": round(res_p, 4),
            "speedup": round(ms_p / ms_m, 3),
            "profiled_mixed_ms": round(pms_m, 3),
            "profiled_plain_ms": round(pms_p, 3),
        }
        for fmt, (ms_h, _, res_h, it_h) in half.items():
            sweep[tag].update({
                fmt + "_gmres_ms": round(ms_h, 3),
                fmt + "_gmres_tflops": round(fl / (ms_h * 1e-3) / 1e12, 3),
                fmt + "_gmres_inner_iterations": it_h,
                fmt + "_gmres_residual_eps": round(res_h, 4)})
        head = (kappa, ms_m, prof_m, res_m, iters, ms_p, prof_p, res_p,
                pms_m, pms_p)
        print("bench: %s %s" % (tag, json.dumps(sweep[tag])),
              file=sys.stderr, flush=True)

    ggt = None
    if not args.no_ggt:
        A0 = ggt_spd(n, seed=0)
        torch.cuda.synchronize()   # the library's streams read A0 next
        bench.pa.set_param(bench.pa.PlasmaGraphCapture,
                           bench.pa.PlasmaEnabled)
        ms_g, _, res_g = bench.plain(A0, B0, "plain G G^T/n + I",
                                     profile=False)
        _, prof_g, _ = bench.plain(A0, B0, "plain profiled G G^T/n + I")
        del A0
        ggt = {"workload": "plain FP64 dposv (drop-in) n=%d nrhs=%d lower, "
                           "A = G G^T/n + I" % (n, nrhs),
               "tflops": round(fl / (ms_g * 1e-3) / 1e12, 3),
               "ms": round(ms_g, 3), "residual_eps": round(res_g, 4),
               "roofline": roofline(
                   prof_g, "upd64_kernel (FP64 MFMA "
                   "syrk+gemm)", FP64_MFMA_PEAK),
               "kernels": classes(prof_g, args.steps)}

    (kappa, ms_m, prof_m, res_m, iters, ms_p, prof_p, res_p, pms_m,
     pms_p) = head
    line = {
        "metric": "dposv n=32768 time-to-FP64 backward error: mixed-prec vs "
                  "plain FP64; TFLOP/s",
        "value": round(fl / (ms_m * 1e-3) / 1e12, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_m, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "plasma_dsposv n=%d nrhs=%d lower, A = Q "
                               "diag(s) Q^T (64 Householder reflectors), "
                               "cond %.0e; FP32 factor, FP64 refinement"
                               % (n, nrhs, kappa),
                   "n": n, "nrhs": nrhs, "uplo": "lower", "kappa": kappa,
                   "schedule_fp32_factor": dict(zip(
                       ("nb", "update_depth", "lookahead"),
                       bench.schedule("s", n))),
                   "schedule_fp64_plain": dict(zip(
                       ("nb", "update_depth", "lookahead"),
                       bench.schedule("d", n))),
                   "tuning": "pinned" if args.nb > 0 else "autotuned"},
        "iterations": iters,
        "residual_eps": round(res_m, 4),
        "backward_error_ok": res_m < 10.0,
        "timing": "graph mode (captured factorization graphs, no "
                  "per-launch profiling); roofline and kernels from a "
                  "profiled eager pass of the same solves right after",
        "profiled_ms_per_step": round(pms_m, 3),
        "plain_fp64": {"tflops": round(fl / (ms_p * 1e-3) / 1e12, 3),
                       "ms": round(ms_p, 3), "residual_eps": round(res_p, 4),
                       "profiled_ms": round(pms_p, 3),
                       "roofline": roofline(
                           prof_p, "upd64_kernel (FP64 "
                           "MFMA syrk+gemm)", FP64_MFMA_PEAK),
                       "kernels": classes(prof_p, args.steps)},
        "speedup_vs_plain": round(ms_p / ms_m, 3),
        "kappa_sweep": sweep,
        "roofline": split_roofline(
            roofline(prof_m, "trailing update of the FP32 factor",
                     FP32_MFMA_PEAK,
                     pmc_traffic("trailing_update_fp32_dsposv_"
                                 "n32768_pmc_traffic")),
            bench.split_mode()),
        "kernels": classes(prof_m, args.steps),
        "plain_ggt": ggt,
        "batched": batched_line(bench, args.batch, args.steps)
        if args.batch > 0 else None,
        "adjoint": {"n=%d" % k: adjoint_line(bench, k, args.steps)
                    for k in orders(args.adjoint_n)} or None,
        "config2": config2_line(bench, orders(args.config2), nrhs) or None,
        "complex": complex_line(bench, args.complex_n, nrhs, args.steps,
                                args.complex_cpu_n, args.cpu_threads)
        if args.complex_n > 0 else None,
        "host_pointers": host,
    }
    if args.calibrate_copy > 0:
        nel = args.calibrate_copy * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        line["calibration_copy_bytes"] = 2 * nel * 8
        del src, dst
    if args.cpu_n > 0:
        print("bench: cpu baseline n=%d" % args.cpu_n, file=sys.stderr,
              flush=True)
        line["cpu_baseline"] = cpu_baseline(args.cpu_n, nrhs, 256, kappa,
                                            args.cpu_threads)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    bench.lib.plasma_finalize()


if __name__ == "__main__":
    main()
