"""Benchmark: pairs/s of the connected kernel sum at 24^3x48 on one MI355X.

Workload (one step = one window): the sliding-hypercube chain K=8, stride n=3,
h=(1,1,1,1); window 1 = 65 pairs x 2 flavours over 30 distinct propagator blocks
(3.06 GB each, FP64, resident in HBM), random fields from seed 2 (RandomProvider,
generated on the device). Two window-batched forms run as one hipGraph per
window: FP32 products with the plain QED kernel (form 1), and the same with the
MLP surrogate's channels (form 2, network trained first over the lattice's
reach, seed 4). The plain FP64 per-pair path runs on the same fields as the
reference; every pair sum of both forms is compared to it. The headline value is
the faster form whose largest per-pair relative error is within 1e-3.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--epochs E]
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DIMS = (48, 24, 24, 24)
XUNIT = 105.658 / 197.3269804 * 0.1  # m_mu a, a = 0.1 fm
MI355X_HBM = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by a 1 GiB copy's reading), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_24x48_window.txt")
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes" and f[1] == kernel:
                return float(f[2])
    except OSError:
        pass
    return None


def cpu_baseline(seconds_target=12.0):
    """The oracle (CPU restatement, FP64, OpenMP) on a bounded sample: whole pairs
    on an 8^4 lattice, repeated for about `seconds_target`, scaled per site to 24^3x48."""
    import numpy as np
    o = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    o.oracle_kernel_sum.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o.oracle_set_threads(threads)
    dims = np.array((8, 8, 8, 8), np.int32)
    V = int(np.prod(dims))
    S = np.empty(2 * 12 * 24 * V)
    Y = np.empty_like(S)
    o.oracle_fill_block(P(S), P(dims), None, ctypes.c_uint64(11))
    o.oracle_fill_block(P(Y), P(dims), None, ctypes.c_uint64(12))
    out = np.zeros(2)
    n, t0 = 0, time.time()
    while True:
        o.oracle_kernel_sum(P(S), P(Y), P(dims), P(np.zeros(4, np.int32)), P(np.ones(4, np.int32)),
                            ctypes.c_double(XUNIT), 0, P(out))
        n += 1
        if time.time() - t0 > seconds_target:
            break
    dt = (time.time() - t0) / n
    per_site = dt / V
    v_big = 48 * 24 ** 3
    return {"value": 1.0 / (per_site * v_big), "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"{n} whole pairs (both flavours) on an 8^4 lattice, {dt:.2f} s each, "
                      f"scaled by sites to 24^3x48"}


def time_form(con, descs, form, warmup, steps):
    """seconds per window (host clock around synchronised graph launches) and the
    last result"""
    import torch
    res = None
    for _ in range(warmup):
        res = con.window_kernel_sum(descs, XUNIT, form=form)
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(steps):
        res = con.window_kernel_sum(descs, XUNIT, form=form)
    torch.cuda.synchronize()
    return (time.time() - t0) / steps, res


def bench_2p2(H, torch, lat, con, args):
    """The 2+2 pieces of one source (Contractor::pieces_2p2, src/contract.cpp:116-150):
    P1 alone (pi_2p2 + p1_2p2), then P1 and four targets (p234_sites + p234_reduce per
    target). Both calls are synchronous, so the host clock brackets them. pi_2p2's
    algorithmic bytes per site: two 12 x 24 double spinor sets read (4608 B) and 16
    doubles of Pi written (128 B)."""
    V = 1
    for d in DIMS:
        V *= d
    blk = torch.empty(H.block_length(lat), dtype=torch.float64, device="cuda")
    H.fill_random_block_device(lat, 1, blk.data_ptr())
    w = (3, 5, 7, 11)
    targets = [(0, 0, 0, 0), (24, 12, 12, 12), (5, 1, 2, 3), (47, 23, 0, 9)]
    lmax = max(DIMS)
    d1 = torch.empty(64 * lmax, dtype=torch.float64, device="cuda")
    d234 = torch.empty(len(targets) * 320, dtype=torch.float64, device="cuda")

    def run(ts, k):
        for _ in range(args.warmup):
            con.pieces_2p2(blk.data_ptr(), w, ts, XUNIT, d1.data_ptr(), d234.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            con.pieces_2p2(blk.data_ptr(), w, ts, XUNIT, d1.data_ptr(), d234.data_ptr())
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k

    t_p1 = run([], max(args.steps, 10))
    log(f"P1 alone {t_p1 * 1e3:.3f} ms per source")
    t_all = run(targets, args.steps)
    per_target = (t_all - t_p1) / len(targets)
    log(f"P1 + {len(targets)} targets {t_all * 1e3:.3f} ms; {per_target * 1e3:.3f} ms per target, "
        f"{V / per_target:.3e} sites/s (four kernel evaluations each)")
    line = {"metric": "2+2 sources per second (P1 + 4 targets)", "value": 1.0 / t_all, "unit": "sources/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_all * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (device RandomProvider block, seed 1)",
            "config": {"workload": "pieces_2p2 of one source, %dx%dx%dx%d" % DIMS, "targets": len(targets)},
            "p1_ms": t_p1 * 1e3, "ms_per_target": per_target * 1e3,
            "pi_2p2_bytes_per_launch": V * (4608 + 128)}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 1 GiB device buffer three times (reads and writes 1 GiB "
                         "each): the reference point for the HBM byte counters, then exit")
    ap.add_argument("--part", choices=["kernel_sum", "2p2"], default="kernel_sum",
                    help="2p2: time the 2+2 pieces of one source (P1 and four targets) on the same "
                         "lattice and print their own line instead of the kernel-sum line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "the multi-GPU path is a design (sites sharded, one allreduce per pass).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import hlbl_amd as H

    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8, "copies": 3}), flush=True)
        return

    lat = H.Lattice(DIMS)
    qed = H.QedKernel(0)
    con = H.Contractor(lat, qed)
    if args.part == "2p2":
        bench_2p2(H, torch, lat, con, args)
        return

    # surrogate: trained on the GPU by the library's own kernels (not timed), over
    # the distances this lattice reaches (outside: the table)
    qed.mlp_set_domain(lat.reach(XUNIT))
    t0 = time.time()
    loss = qed.mlp_train(n_samples=1 << 20, epochs=args.epochs, batch=1024, lr=2e-3, seed=4)
    t_train = time.time() - t0
    log(f"surrogate trained in {t_train:.1f} s, loss {loss[0]:.3e} -> {loss[-1]:.3e}")

    # window 1 of the chain: its 65 pairs and 30 blocks
    sch = H.Schedule((1, 1, 1, 1), 3, 8)
    pairs_w = [(q, p) for w, p, parts in sch.sources() if w == 1 for q in parts]
    assert len(pairs_w) == 65
    positions = sorted({q for q, p in pairs_w} | {p for q, p in pairs_w})
    wrap = lambda v: tuple(a % L for a, L in zip(v, DIMS))
    nblk = H.block_length(lat)
    blocks = {}
    t0 = time.time()
    for pos in positions:
        d = torch.empty(nblk, dtype=torch.float64, device="cuda")
        H.RandomProvider(lat, 2).compute(wrap(pos), d.data_ptr())
        blocks[pos] = d
    torch.cuda.synchronize()
    log(f"{len(positions)} blocks ({len(positions) * nblk * 8 / 1e9:.1f} GB) generated in {time.time() - t0:.1f} s")
    descs = [(blocks[q].data_ptr(), blocks[p].data_ptr(), wrap(q), lat.site_map_diff(p, q))
             for q, p in pairs_w]

    # the two window forms, then their stage times (HIP events on the window stream)
    dt1, res1 = time_form(con, descs, 1, args.warmup, args.steps)
    dt2, res2 = time_form(con, descs, 2, args.warmup, args.steps)
    judge, nsamp = con.judge()
    _, ms1 = con.window_profile(descs, XUNIT, form=1)
    _, ms2 = con.window_profile(descs, XUNIT, form=2)
    log(f"form 1: {dt1 * 1e3:.1f} ms per window ({65 / dt1:.2f} pairs/s), stages {ms1} ms")
    log(f"form 2: {dt2 * 1e3:.1f} ms per window ({65 / dt2:.2f} pairs/s), stages {ms2} ms, "
        f"judge {judge:.2e} ({nsamp} sites)")

    # plain FP64, one pair per call, over the whole window (the reference sums)
    ref = []
    torch.cuda.synchronize()
    t0 = time.time()
    for d in descs:
        ref.append(con.kernel_sum(*d, XUNIT))
    dt64 = (time.time() - t0) / len(descs)
    ref = np.array(ref)
    err1 = np.abs(np.array(res1) - ref) / np.abs(ref)
    err2 = np.abs(np.array(res2) - ref) / np.abs(ref)
    scale = np.sqrt(np.mean(ref ** 2))
    log(f"plain FP64: {dt64 * 1e3:.1f} ms per pair ({1 / dt64:.2f} pairs/s); max rel err form1 "
        f"{err1.max():.2e} form2 {err2.max():.2e} (median {np.median(err2):.2e}, "
        f"max |diff| / rms pair sum {np.max(np.abs(np.array(res2) - ref)) / scale:.2e})")

    # headline: the faster form within the 1e-3 error bound
    forms = {1: (dt1, err1.max(), ms1), 2: (dt2, err2.max(), ms2)}
    ok = [f for f in forms if forms[f][1] <= 1e-3]
    head = min(ok, key=lambda f: forms[f][0]) if ok else 1
    dt, err, ms = forms[head]
    value = 65.0 / dt
    # roofline of the dominant kernel, pass 2, against HBM (its highest floor):
    # algorithmic bytes per launch = per pair, both FP64 blocks once (2 x 2 x 12 x 24
    # doubles per site) and the pair's coefficient field once (640 floats per site)
    V = lat.volume()
    bytes_pass2 = 65 * (2 * 2 * 12 * 24 * 8 + 640 * 4) * V
    achieved = bytes_pass2 / (ms[3] * 1e-3)
    traffic = measured_traffic("pass2_window")
    cpu = None if args.no_cpu else cpu_baseline()
    names = {1: "FP32 window-batched (pass-2 products as scaled FP16 hi/lo x3), plain QED kernel",
             2: "FP32 window-batched (pass-2 products as scaled FP16 hi/lo x3), MLP surrogate channels"}
    line = {
        "metric": "pairs/sec at 24^3x48 on 1 GPU, surrogate+FP32 vs plain FP64, rel err <=1e-3",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "24^3x48 window 1 of chain K=8 n=3 h=(1,1,1,1): 65 pairs x 2 flavours, "
                               f"30 resident FP64 blocks, seed 2; {names[head]}",
                   "lattice": list(DIMS), "pairs_per_step": 65, "xunit": XUNIT, "form": head},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": MI355X_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / MI355X_HBM, "traffic": traffic,
                     "kernel": "pass2_window", "ms": ms[3]},
        "max_pair_rel_err": float(err),
        "plain_fp64_pairs_per_s": 1.0 / dt64,
        "speedup_vs_plain_fp64": value * dt64,
        "form1_pairs_per_s": 65.0 / dt1, "form1_max_pair_rel_err": float(err1.max()),
        "form2_pairs_per_s": 65.0 / dt2, "form2_max_pair_rel_err": float(err2.max()),
        "form2_median_pair_rel_err": float(np.median(err2)),
        "surrogate_judge_max_rel": judge, "surrogate_train_s": t_train,
        "stage_ms_form1": ms1, "stage_ms_form2": ms2,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
