"""The sliding-hypercube schedule (src/schedule.cpp) against the reference's own
compiled schedule (oracle/_ref) and the closed forms of its tests
(sanity_check/test_schedule.cpp). CPU only."""
import ctypes
import itertools
import json
import os

import numpy as np
import pytest

import hlbl_amd as H
from conftest import P

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))
SIGNS = [tuple(v) for v in itertools.product((-1, 1), repeat=4)]


def ref_rows(ref, h, stride, n, origin):
    ref.ref_build_schedule.restype = ctypes.c_int
    cnt = ref.ref_build_schedule(P(np.array(h, np.int32)), stride, n, P(np.array(origin, np.int32)))
    rows = np.zeros((cnt, 10), np.int32)
    ref.ref_schedule_rows(P(rows))
    return rows


def our_rows(h, stride, n, origin):
    s = H.Schedule(h, stride, n, origin)
    rows = []
    for w, p, parts in s.sources():
        if not parts:
            rows.append([w, *p, *p, 0])
        for q in parts:
            rows.append([w, *p, *q, len(parts)])
    return np.array(rows, np.int32)


CASES = [((1, 1, 1, 1), 1, 3, (0, 0, 0, 0)), ((1, -1, 1, -1), 2, 4, (0, 0, 0, 0)),
         ((1, 1, 1, 1), 3, 8, (0, 0, 0, 0)), ((1, -1, 1, 1), 1, 6, (5, 0, 0, 0)),
         ((-1, -1, -1, -1), 3, 6, (5, 0, 0, 0)), ((1, 1, -1, -1), 2, 2, (2, 0, 5, 1)),
         ((-1, 1, -1, 1), 4, 4, (0, 0, 0, 0))]


@pytest.mark.parametrize("h,stride,n,origin", CASES)
def test_schedule_matches_reference_bit_exact(ref, h, stride, n, origin):
    assert np.array_equal(our_rows(h, stride, n, origin), ref_rows(ref, h, stride, n, origin))


@pytest.mark.parametrize("h", SIGNS)
def test_simulate_matches_reference(ref, h):
    for mode in (0, 1):
        for k in (1, 2, 5):
            want = (ctypes.c_longlong * 5)()
            assert ref.ref_simulate(P(np.array(h, np.int32)), 4, k, P(np.zeros(4, np.int32)), mode, want) == 0
            got = H.Schedule(h, 4, k).simulate(interleaved=bool(mode))
            assert list(got.values()) == list(want)


def test_closed_forms():
    g = GOLD["schedule"]
    s = H.Schedule((1, 1, 1, 1), 1, 3)
    s.check_complete()
    src = s.sources()
    counts = {}
    for w, _p, parts in src:
        c = counts.setdefault(w, [0, 0])
        c[0] += 1
        c[1] += len(parts)
    assert counts[0] == g["stride1_K3"]["window0"]
    assert all(counts[k] == g["stride1_K3"]["later"] for k in (1, 2, 3))
    big = H.Schedule((1, 1, 1, 1), 4, 30)
    big.check_complete()
    k30 = g["K30_stride4"]
    ser, inter = big.simulate(False), big.simulate(True)
    assert ser["n_propagators"] == k30["n_propagators"] and ser["n_pairs"] == k30["n_pairs"]
    assert ser["n_kernel_sums"] == k30["n_kernel_sums"] and ser["n_solves"] == 466 * 24
    assert ser["peak_cache"] == k30["peak_serial"] and inter["peak_cache"] == k30["peak_interleaved"]
    for h in SIGNS:
        for k in (1, 2, 5):
            st = H.Schedule(h, 4, k).simulate(True)
            assert st["n_propagators"] == 16 + 15 * k and st["n_pairs"] == 15 + 65 * k
            assert st["peak_cache"] == 16


def test_benchmark_chain_sizes():
    """The chains of the benchmark configurations: K=1: 31/80, K=4: 76/275, K=8: 136/535."""
    for K, props, pairs in ((1, 31, 80), (4, 76, 275), (8, 136, 535)):
        st = H.Schedule((1, 1, 1, 1), 3, K).simulate(True)
        assert (st["n_propagators"], st["n_pairs"]) == (props, pairs)


def test_resume_overlap():
    """A chain from window j's centre with K-j steps repeats only 15 pairs."""
    for h in ((1, 1, 1, 1), (1, -1, 1, -1)):
        n, K, origin = 3, 8, (2, 0, 5, 1)
        full = H.Schedule(h, n, K, origin)
        fs = full.sources()
        for j in (0, 1, 5):
            o2 = tuple(o + n * j * hh for o, hh in zip(origin, h))
            res = H.Schedule(h, n, K - j, o2)
            rs = res.sources()
            by_w_full = [s for s in fs if s[0] > j]
            by_w_res = [(w + j, p, q) for w, p, q in rs if w >= 1]
            assert by_w_full == by_w_res
            done = {frozenset((q, p)) for w, p, parts in fs if w <= j for q in parts}
            again = {frozenset(pq) for pq in res.pairs()}
            assert len(done & again) == 15
            assert len(done | again) == len(full.pairs())


def test_check_complete_detects_a_missing_pair():
    s = H.Schedule((1, 1, 1, 1), 1, 1)
    s.check_complete()
    # checked at the wrong stride, every scheduled pair is unexpected
    with pytest.raises(H.HlblError):
        H.check(H.lib().hlbl_schedule_check_complete(s.handle, 2))


def test_max_steps_without_wrap(ref):
    for h, L in (((1, 1, 1, 1), (16, 16, 16, 16)), ((1, -1, 1, -1), (12, 8, 8, 8)), ((1, 1, 1, 1), (48, 24, 24, 24))):
        want = ref.ref_max_steps_without_wrap(P(np.array(h, np.int32)), 3, P(np.array(L, np.int32)),
                                              P(np.zeros(4, np.int32)), 10000)
        assert H.max_steps_without_wrap(h, 3, L) == want
    assert 0 <= H.max_steps_without_wrap((1, 1, 1, 1), 4, (16, 16, 16, 16)) < 100


def test_bad_inputs():
    with pytest.raises(H.HlblError):
        H.Schedule((1, 2, 1, 1), 1, 1)
    with pytest.raises(H.HlblError):
        H.Schedule((1, 1, 1, 1), 0, 1)


def test_plan_2p2_known_answers():
    """plan_2p2 on the cases of sanity_check/test_schedule.cpp:134-187: two chains from
    (0,0,0,0) and (1,1,1,1), h = (1,1,1,1), one step each, on 6^4: (1,1,1,1) at chain 0
    window 1, (5,5,5,5) at chain 0 as (-1,-1,-1,-1), (2,2,2,2) at chain 1, (3,0,0,0)
    unsolved; every other solved site listed once; the report's 14 lines. A chain that
    wraps onto its own sites gets one 2+2 per site."""
    L = (6, 6, 6, 6)
    chains = [H.Schedule((1, 1, 1, 1), 1, 1, (0, 0, 0, 0)), H.Schedule((1, 1, 1, 1), 1, 1, (1, 1, 1, 1))]
    sources = [(1, 1, 1, 1), (2, 2, 2, 2), (5, 5, 5, 5), (3, 0, 0, 0)]
    plan = H.plan_2p2(chains, sources, L)
    assert plan["at"][0] == (0, 1) and plan["at"][2][0] == 0 and plan["at"][1][0] == 1
    assert sorted(plan["at"]) == [0, 1, 2]
    assert plan["unsolved"] == [3]
    # (-1,-1,-1,-1) is solved by chain 0 (unwrapped), so (5,5,5,5) is found there
    assert any(tuple(p) == (-1, -1, -1, -1) and w == plan["at"][2][1]
               for w, p, _ in chains[0].sources())
    sites = {tuple(a % 6 for a in p) for s in chains for _w, p, _ in s.sources()}
    listed = [p for p, _c, _w in plan["not_sources"]]
    assert len(listed) == len(set(listed)) == len(sites) - 3
    assert not {(1, 1, 1, 1), (2, 2, 2, 2), (5, 5, 5, 5)} & set(listed)
    lines = H.report_2p2(plan, sources)
    n = len(plan["not_sources"])
    assert len(lines) == 14
    assert lines[0] == (f"WARNING: {n} solved source positions are not 2+2 sources "
                        "(no 2+2 computed for them)")
    assert lines[11] == f"  ... and {n - 10} more"
    assert lines[12] == "WARNING: 1 2+2 sources are not solved by any chain (no 2+2 computed for them)"
    assert lines[13] == "  (3,0,0,0)"
    assert H.report_2p2(plan, []) == []
    wrapped = H.plan_2p2([H.Schedule((1, 1, 1, 1), 1, 6)], [(0, 0, 0, 0)], (4, 4, 4, 4))
    assert wrapped["at"] == {0: (0, 0)} and wrapped["unsolved"] == []


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_plan_2p2_first_solve(seed):
    """Every 2+2 source is placed at the earliest solve of its wrapped site in run order
    (brute force over the flattened solve list), duplicates keep the first index, and
    not_sources holds exactly the first solves of the other sites, in run order."""
    rng = np.random.default_rng(seed)
    L = (4, 6, 4, 6)
    chains = []
    for _ in range(3):
        h = tuple(int(v) for v in rng.choice([-1, 1], 4))
        origin = tuple(int(v) for v in rng.integers(-3, 4, 4))
        chains.append(H.Schedule(h, int(rng.integers(1, 3)), int(rng.integers(1, 4)), origin))
    solves = [(tuple(a % l for a, l in zip(p, L)), c, w) for c, s in enumerate(chains) for w, p, _ in s.sources()]
    sources = [tuple(int(v) for v in rng.integers(-6, 7, 4)) for _ in range(12)]
    sources.append(sources[0])
    plan = H.plan_2p2(chains, sources, L)
    first = {}
    for k, (p, c, w) in enumerate(solves):
        first.setdefault(p, (k, c, w))
    for i, src in enumerate(sources):
        p = tuple(a % l for a, l in zip(src, L))
        dup = any(tuple(a % l for a, l in zip(sources[j], L)) == p for j in range(i))
        if p not in first:
            assert i in plan["unsolved"] and i not in plan["at"]
        elif dup:
            assert i not in plan["at"]
        else:
            assert plan["at"][i] == first[p][1:]
    src_sites = {tuple(a % l for a, l in zip(s, L)) for s in sources}
    want = sorted((k, p, c, w) for p, (k, c, w) in first.items() if p not in src_sites)
    assert plan["not_sources"] == [(p, c, w) for _k, p, c, w in want]
