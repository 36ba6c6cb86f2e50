#!/usr/bin/env python3
"""Benchmark of the Cholesky solve path on one MI355X.

Workload (the headline metric's configuration): plasma_dposv, n = 32768,
nrhs = 64, lower, A = G G^T / n + I with G iid N(0,1) (torch generator seed
0), B iid N(0,1).  Inputs are generated on the GPU and stay resident in HBM.
One step is exactly the work of plasma_dposv on device-resident LAPACK
arrays: tile descriptors from the pool, tr2desc(A) + ge2desc(B), tile
Cholesky + two triangular sweeps, desc2tr / desc2ge into output arrays,
then the sequence is synchronised.  The input arrays are only read, so every
step factors the same matrix.

Printed JSON: TFLOP/s with the reference tester's flop count
flops_dpotrf(n) + flops_dpotrs(n, nrhs) (test/flops.h:366-417), the scaled
residual of the last solution, the trailing-update kernel's roofline
(HIP events on its own stream) and the CPU baseline: the oracle (PLASMA's
tile algorithm restated, OpenMP tasks, sequential MKL tile kernels) timed
on this host on a smaller n.
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
HBM_PEAK = 8.0e12             # bytes/s


def flops_posv(n, nrhs):
    n, nrhs = float(n), float(nrhs)
    potrf = n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0
    potrs = 2.0 * n * n * nrhs
    return potrf + potrs


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--lookahead", type=int, default=1)
    ap.add_argument("--cpu-n", type=int, default=16384,
                    help="order of the CPU-baseline sample (0 skips it)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def cpu_baseline(n, nrhs, nb, threads):
    """Oracle dposv on the host: median of 2 runs after one warm-up."""
    import numpy as np
    import torch
    import oracle as ora
    backend = ora.set_blas(ora.MKL)
    if threads <= 0:
        threads = min(16, len(os.sched_getaffinity(0)))
    ora.set_threads(threads)
    g = torch.Generator(device="cuda").manual_seed(1)
    G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    A = (G @ G.T).div_(n)
    A.diagonal().add_(1.0)
    B = torch.randn(nrhs, n, dtype=torch.float64, device="cuda", generator=g)
    A = np.asfortranarray(A.cpu().numpy())
    B = np.asfortranarray(B.cpu().numpy().T)
    del G
    times = []
    for i in range(3):
        t0 = time.perf_counter()
        _, X, info = ora.posv(ora.LOWER, A, B, nb)
        times.append(time.perf_counter() - t0)
        assert info == 0
    t = sorted(times[1:])[len(times[1:]) // 2]
    return {
        "value": flops_posv(n, nrhs) / t / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": "oracle dposv n=%d nrhs=%d nb=%d lower, OpenMP tile tasks, "
                  "%s sequential tile kernels, median of 2 runs after 1 "
                  "warm-up (%.2f s per solve)"
                  % (n, nrhs, nb, "MKL" if backend == ora.MKL else "built-in",
                     t),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not "
              "supported (multi-GPU sharding is design-only)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    import plasma_amd as pa

    lib = pa.load()
    rc = lib.plasma_init()
    if rc != 0:
        raise SystemExit("plasma_init failed: %d" % rc)
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    pa.set_param(pa.PlasmaNb, args.nb)
    pa.set_param(pa.PlasmaLookahead, args.lookahead)
    lib.plasma_profile_enable.argtypes = [ctypes.c_int]
    lib.plasma_profile_query.argtypes = [
        ctypes.c_int, ctypes.POINTER(ctypes.c_double),
        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
        ctypes.POINTER(ctypes.c_long)]

    n, nrhs, nb = args.n, args.nrhs, args.nb
    gen = torch.Generator(device="cuda").manual_seed(0)
    G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen)
    A0 = (G @ G.T).div_(n)
    A0.diagonal().add_(1.0)
    del G
    # column-major n x nrhs: row j of the tensor is column j
    B0 = torch.randn(nrhs, n, dtype=torch.float64, device="cuda", generator=gen)
    Aout = torch.empty_like(A0)
    X = torch.empty_like(B0)
    torch.cuda.synchronize()

    A = pa.plasma_desc_t()
    B = pa.plasma_desc_t()

    def step():
        assert lib.plasma_desc_triangular_create(
            pa.PlasmaRealDouble, pa.PlasmaLower, nb, nb, n, n, 0, 0, n, n,
            ctypes.byref(A)) == 0
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

    for _ in range(args.warmup):
        step()
    print("bench: warmup done (n=%d nb=%d)" % (n, nb), file=sys.stderr,
          flush=True)
    lib.plasma_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        print("bench: step %d done" % i, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps

    def prof(cls):
        a, f, b, c = (ctypes.c_double(), ctypes.c_double(), ctypes.c_double(),
                      ctypes.c_long())
        lib.plasma_profile_query(cls, ctypes.byref(a), ctypes.byref(f),
                                 ctypes.byref(b), ctypes.byref(c))
        return a.value, f.value, b.value, c.value

    upd_ms, upd_flops, _, upd_launches = prof(0)
    classes = {}
    for cls, name in enumerate(["trailing_update", "potrf_diag", "trsm_panel",
                                "sweeps", "layout", "mixed"]):
        a, f, b, c = prof(cls)
        if c:
            classes[name] = {"ms_per_step": round(a / args.steps, 3),
                             "launches_per_step": c // args.steps,
                             "tflops": round(f / (a * 1e-3) / 1e12, 2)
                             if a and f else None,
                             "gbps": round(b / (a * 1e-3) / 1e9, 1)
                             if a and b else None}
    lib.plasma_profile_enable(0)

    # verification of the last solution: scaled residual (FP64, on the GPU)
    Xc = X.T
    Bc = B0.T
    R = Bc - A0 @ Xc
    an = A0.abs().sum(1).max().item()
    xn = Xc.abs().sum(1).max().item()
    rn = R.abs().sum(1).max().item()
    eps = 2.0 ** -53
    residual = rn / (n * an * xn) / eps

    value = flops_posv(n, nrhs) / (ms * 1e-3) / 1e12
    achieved = upd_flops / (upd_ms * 1e-3) if upd_ms else 0.0
    line = {
        "metric": "dposv n=32768 time-to-FP64 backward error: mixed-prec vs "
                  "plain FP64; TFLOP/s",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain FP64 dposv (drop-in) n=%d nrhs=%d "
                               "lower, A = G G^T/n + I" % (n, nrhs),
                   "n": n, "nrhs": nrhs, "nb": nb, "uplo": "lower",
                   "lookahead": args.lookahead, "precision": "fp64"},
        "residual_eps": round(residual, 4),
        "backward_error_ok": residual < 10.0,
        "roofline": {
            "kernel": "trailing_update_kernel<double> (FP64 MFMA syrk+gemm)",
            "bound": "mfma",
            "achieved": round(achieved / 1e12, 3),
            "peak": FP64_MFMA_PEAK / 1e12,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_MFMA_PEAK, 4),
            "avg_launch_ms": round(upd_ms / upd_launches, 4)
            if upd_launches else None,
            "traffic": None,
        },
        "kernels": classes,
    }
    if args.cpu_n > 0:
        print("bench: cpu baseline n=%d" % args.cpu_n, file=sys.stderr,
              flush=True)
        line["cpu_baseline"] = cpu_baseline(args.cpu_n, nrhs, 256,
                                            args.cpu_threads)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    lib.plasma_finalize()


if __name__ == "__main__":
    main()
