"""Launch-shape tuner for a contractor's kernels on this GPU.

For one lattice it builds window 1 of the chain K=8, stride n, h=(1,1,1,1) on random
blocks (RandomProvider, generated on the device): the workload the window form runs;
or, with --workload source, the call of window 1's source with the most partners (the
per-source calls a lattice uses whose windows do not fit in memory, as 32^3x64).
Then, one shape at a time (coordinate descent from the defaults), it times every
candidate and keeps the fastest:
  - pass1_batch, pass2_wg_per_cu: the window's pass-1 and pass-2 stage times (HIP
    events of hlbl_window_profile, fields kept);
  - field_waves x field_wg_per_cu, axes_wg_per_cu, near_wg_per_cu: the coefficient-field
    stage with every field built (the kept ones dropped before each run);
  - fp64_field_wg_per_cu, fp64_pass2_wg_per_cu: one pair of the plain FP64 path (host
    clock around the synchronous call).
Each time is the median of `reps` runs after one untimed run; a candidate replaces the
current shape only when it is faster by more than `margin` (1%), so that run-to-run
noise does not pick shapes. The pair sums under the
picks are compared with those under the defaults (they differ only in the order of the
FP64 sums across workgroups). Picks and every timing go to
hlbl_amd/data/tuning_gfx950.json under the lattice's key; Contractor applies them.

python -m hlbl_amd.tune --dims 48 24 24 24 --stride 3 [--reps 5] [--margin 0.01] [--out PATH]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import hlbl_amd as H  # noqa: E402

# candidates per knob, in the order they are tuned; a tuple value sets two fields
CANDIDATES = [
    ("pass1_batch", [8, 16, 24, 32, 48, 64, 1 << 20], "pass1"),
    ("pass2_wg_per_cu", [1, 2, 3, 4], "pass2"),
    (("field_waves", "field_wg_per_cu"), [(8, 1), (8, 2), (4, 1), (4, 2), (4, 3)], "fields"),
    ("axes_wg_per_cu", [1, 2, 4, 8, 16], "fields"),
    ("near_wg_per_cu", [1, 2, 4, 8], "fields"),
    ("fp64_field_wg_per_cu", [1, 2, 4, 8], "plain"),
    ("fp64_pass2_wg_per_cu", [1, 2, 4, 8], "plain"),
]
STAGE = {"pass1": 0, "fields": 2, "pass2": 3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def window1(lat, stride, seed, torch, workload="window"):
    """window 1 of the chain K=8, stride n, h=(1,1,1,1), or its source with the most
    partners: descriptors and the blocks"""
    dims = lat.global_
    sch = H.Schedule((1, 1, 1, 1), stride, 8)
    if workload == "window":
        pairs = [(q, p) for w, p, parts in sch.sources() if w == 1 for q in parts]
    else:
        _, p, parts = max((s for s in sch.sources() if s[0] == 1), key=lambda s: len(s[2]))
        pairs = [(q, p) for q in parts]
    wrap = lambda v: tuple(a % L for a, L in zip(v, dims))
    blocks = {}
    for pos in sorted({q for q, p in pairs} | {p for q, p in pairs}):
        d = torch.empty(H.block_length(lat), dtype=torch.float64, device="cuda")
        H.RandomProvider(lat, seed).compute(wrap(pos), d.data_ptr())
        blocks[pos] = d
    torch.cuda.synchronize()
    descs = [(blocks[q].data_ptr(), blocks[p].data_ptr(), wrap(q), lat.site_map_diff(p, q)) for q, p in pairs]
    return descs, blocks


def measure(con, descs, xunit, kind, reps):
    """median time (ms) of the stage(s) a knob affects"""
    ts = []
    for r in range(reps + 1):
        if kind == "plain":
            t0 = time.perf_counter()
            con.kernel_sum(*descs[0], xunit)
            t = (time.perf_counter() - t0) * 1e3
        else:
            if kind == "fields":
                con.drop_fields()
            _, ms = con.window_profile(descs, xunit, form=1)
            t = ms[STAGE[kind]]
        if r > 0:
            ts.append(t)
    return float(np.median(ts))


def tune(con, descs, xunit, reps=5, margin=0.01):
    picks = con.tuning()
    timings = {}
    for knob, cands, kind in CANDIDATES:
        names = knob if isinstance(knob, tuple) else (knob,)
        key = "+".join(names)
        timings[key] = {}
        current = tuple(picks[n] for n in names)
        times = {}
        for cand in cands:
            vals = cand if isinstance(cand, tuple) else (cand,)
            trial = dict(picks)
            trial.update(zip(names, vals))
            con.set_tuning(**trial)
            t = measure(con, descs, xunit, kind, reps)
            timings[key][str(cand)] = t
            times[vals] = t
            log(f"{key} = {cand}: {t:.3f} ms")
        best = min(times, key=times.get)
        if current in times and times[best] > (1.0 - margin) * times[current]:
            best = current
        picks.update(zip(names, best))
        con.set_tuning(**picks)
        log(f"  -> {key} = {best}")
    return picks, timings


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dims", type=int, nargs=4, default=[48, 24, 24, 24])
    ap.add_argument("--stride", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--workload", choices=["window", "source"], default="window")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--margin", type=float, default=0.01)
    ap.add_argument("--out", default=H.TUNING_PATH)
    args = ap.parse_args()
    import torch

    xunit = 105.658 / 197.3269804 * 0.1
    lat = H.Lattice(tuple(args.dims))
    qed = H.QedKernel(0)
    con = H.Contractor(lat, qed, tuning="default")
    descs, blocks = window1(lat, args.stride, args.seed, torch, args.workload)
    log(f"{args.workload} of window 1: {len(descs)} pairs over {len(blocks)} blocks of {tuple(args.dims)}")

    defaults = con.tuning()
    con.drop_fields()
    base, ms_default_cold = con.window_profile(descs, xunit, form=1)
    _, ms_default = con.window_profile(descs, xunit, form=1)
    plain_default = measure(con, descs, xunit, "plain", args.reps)
    picks, timings = tune(con, descs, xunit, args.reps, args.margin)
    con.drop_fields()
    got, ms_tuned_cold = con.window_profile(descs, xunit, form=1)
    _, ms_tuned = con.window_profile(descs, xunit, form=1)
    plain_tuned = measure(con, descs, xunit, "plain", args.reps)
    diff = float(np.max(np.abs(np.array(got) - np.array(base)) / np.abs(np.array(base))))
    ref64 = con.kernel_sum(*descs[0], xunit)
    con.set_tuning(**defaults)
    ref64_default = con.kernel_sum(*descs[0], xunit)
    diff64 = max(abs(a - b) / abs(b) for a, b in zip(ref64, ref64_default))
    log(f"defaults {defaults}: stages {ms_default} ms (fields cold {ms_default_cold[2]:.2f} ms), "
        f"plain pair {plain_default:.2f} ms")
    log(f"picks {picks}: stages {ms_tuned} ms (fields cold {ms_tuned_cold[2]:.2f} ms), plain pair "
        f"{plain_tuned:.2f} ms; window pair sums vs defaults {diff:.1e}, plain {diff64:.1e}")

    try:
        with open(args.out) as f:
            table = json.load(f)
    except (OSError, ValueError):
        table = {}
    table.setdefault("device", "gfx950 (MI355X)")
    table.setdefault("lattices", {})
    table["lattices"]["x".join(map(str, args.dims))] = {
        "workload": (f"window 1 of chain K=8 n={args.stride} h=(1,1,1,1)" if args.workload == "window" else
                     f"the source of window 1 (chain K=8 n={args.stride} h=(1,1,1,1)) with the most partners") +
                    f", {len(descs)} pairs, seed {args.seed}, form 1; plain path: its first pair",
        "command": "python -m hlbl_amd.tune --dims " + " ".join(map(str, args.dims)) +
                   f" --stride {args.stride} --seed {args.seed} --workload {args.workload} --reps {args.reps}"
                   f" --margin {args.margin}",
        "picks": picks, "defaults": defaults, "timings_ms": timings,
        "stage_ms_default": ms_default, "stage_ms_tuned": ms_tuned,
        "fields_cold_ms_default": ms_default_cold[2], "fields_cold_ms_tuned": ms_tuned_cold[2],
        "plain_pair_ms_default": plain_default, "plain_pair_ms_tuned": plain_tuned,
        "window_pair_rel_diff_vs_default": diff, "plain_pair_rel_diff_vs_default": diff64,
    }
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps({"lattice": args.dims, "picks": picks}), flush=True)
    con.close()
    qed.close()


if __name__ == "__main__":
    main()
