"""Autotuner of the factorization schedule on the GPU it runs on.

    python -m plasma_amd.tune [--out FILE] [--sizes 2048,4096,...]
                              [--modes eager,graph]

For each launch mode (eager launches, or captured graphs replayed:
PlasmaGraphCapture), precision (d, s) and order n it times
plasma_{d,s}potrf on device-resident A = G G^T / n + I (G iid N(0,1)) for
every tile size in {128, 256, 512}, update depth in {1, 2, 3, 4}, lookahead
depth in {0, 1, 2, 3} (how many column groups the panel stream may run
ahead of the bulk updates) and, in single precision, the MFMA form of the
trailing updates (the operand splits bf16 x3 / six products and row-scaled
fp16 x2 / three products); each schedule is timed as the faster of two
runs after one warm-up, tuning disabled while measuring.  It writes the
fastest schedule per (mode, precision, n) as the table the library reads
at plasma_init (control/tuning.c):

    potrf <d|s> <n_max> <nb> <update_depth> <lookahead> <eager|graph> [split]

Each row covers the orders up to its n; orders above the largest measured
n use the last row.  The two modes differ: a replayed graph has no launch
gaps, which favours deeper lookahead.
"""
import argparse
import ctypes
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuning", "mi355x.txt")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--sizes", default="2048,4096,8192,16384,32768")
    ap.add_argument("--nbs", default="128,256,512")
    ap.add_argument("--depths", default="1,2,3,4")
    ap.add_argument("--lookaheads", default="0,1,2,3")
    ap.add_argument("--splits", default="bf16,fp16",
                    help="FP32 update operand forms to tune (bf16, fp16, "
                         "fp32)")
    ap.add_argument("--precs", default="d,s",
                    help="precisions to tune (d, s)")
    ap.add_argument("--modes", default="eager,graph",
                    help="launch modes to tune: eager (PlasmaGraphCapture "
                         "disabled) and/or graph (enabled, captured graphs "
                         "replayed)")
    args = ap.parse_args()
    import torch
    sys.path.insert(0, os.path.dirname(HERE))
    import plasma_amd as pa
    lib = pa.load()
    if lib.plasma_init() != 0:
        raise SystemExit("plasma_init failed")
    pa.set_param(pa.PlasmaTuning, pa.PlasmaDisabled)
    sizes = [int(x) for x in args.sizes.split(",")]
    nbs = [int(x) for x in args.nbs.split(",")]
    depths = [int(x) for x in args.depths.split(",")]
    las = [int(x) for x in args.lookaheads.split(",")]
    modes = args.modes.split(",")
    split_val = {"bf16": pa.PlasmaSplitBf16, "fp16": pa.PlasmaSplitFp16,
                 "fp32": pa.PlasmaDisabled}
    rows = []
    precs = args.precs.split(",")
    for mode, prec, dt in [(m, p, t) for m in modes
                           for p, t in (("d", torch.float64),
                                        ("s", torch.float32)) if p in precs]:
        pa.set_param(pa.PlasmaGraphCapture,
                     pa.PlasmaEnabled if mode == "graph"
                     else pa.PlasmaDisabled)
        potrf = getattr(lib, "plasma_%spotrf" % prec)
        for n in sizes:
            g = torch.Generator(device="cuda").manual_seed(n)
            G = torch.randn(n, n, dtype=torch.float64, device="cuda",
                            generator=g)
            A0 = ((G @ G.T) / n + torch.eye(n, dtype=torch.float64,
                                            device="cuda")).to(dt)
            del G
            A = torch.empty_like(A0)
            best = None
            splits = args.splits.split(",") if prec == "s" else [None]
            for nb, depth, la, sp in [(a, b, c, d) for a in nbs
                                      for b in depths for c in las
                                      for d in splits]:
                pa.set_param(pa.PlasmaNb, nb)
                pa.set_param(pa.PlasmaUpdateDepth, depth)
                pa.set_param(pa.PlasmaLookahead, la)
                if sp is not None:
                    pa.set_param(pa.PlasmaSplitUpdate, split_val[sp])
                times = []
                for rep in range(3):
                    A.copy_(A0)
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    rc = potrf(pa.PlasmaLower, n, A.data_ptr(), n)
                    times.append(time.perf_counter() - t0)
                    if rc != 0:
                        raise SystemExit("potrf failed: %d" % rc)
                t = sorted(times[1:])[0]
                flops = n ** 3 / 3.0
                print("tune %s %s n=%d nb=%d depth=%d lookahead=%d "
                      "split=%s %.3f ms %.2f TFLOP/s" % (
                          mode, prec, n, nb, depth, la, sp, t * 1e3,
                          flops / t / 1e12), flush=True)
                if best is None or t < best[0]:
                    best = (t, nb, depth, la, sp)
            t, nb, depth, la, sp = best
            rows.append("potrf %s %d %d %d %d %s%s" % (
                prec, n, nb, depth, la, mode, "" if sp is None else " " + sp))
            del A, A0
    lib.plasma_finalize()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        f.write("# routine precision n_max nb update_depth lookahead "
                "mode [split]\n")
        f.write("# written by plasma_amd/tune.py (the faster of two timed "
                "runs per schedule)\n")
        for r in rows:
            f.write(r + "\n")
    print("wrote %s" % args.out)


if __name__ == "__main__":
    main()
