"""Parity of the HIP path against the oracle, through the C ABI (needs an MI355X).

FP64 forms: 1e-10 relative to the oracle (the reference's own tolerance,
sanity_check/test_contract_random.cpp:55-57). Integer/byte work (fields,
schedule rows): bit-exact. Mixed-precision forms: stated per test."""
import ctypes
import os

import numpy as np
import pytest

import hlbl_amd as H
from conftest import P, fill_block, oracle_pair

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def qed():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    q = H.QedKernel(0)
    yield q
    q.close()


def dev_block(lat, seed):
    d = torch.empty(H.block_length(lat), dtype=torch.float64, device="cuda")
    H.fill_random_block_device(lat, seed, d.data_ptr())
    return d


def log(*a):
    print(*a, flush=True)


def test_device_fill_bit_exact(qed, oracle):
    for dims, seed in (((4, 2, 2, 2), 9), ((8, 4, 4, 4), 1), ((6, 6, 6, 6), 3)):
        lat = H.Lattice(dims)
        d = dev_block(lat, seed)
        torch.cuda.synchronize()
        assert np.array_equal(d.cpu().numpy(), fill_block(oracle, dims, seed)), dims


def test_qed_kernel_points(qed, oracle):
    rng = np.random.default_rng(0)
    n = 200
    x = rng.uniform(-3, 3, (n, 4))
    y = rng.uniform(-3, 3, (n, 4))
    x[0] = 0
    y[1] = 0
    x[2] = [7, 0, 0, 0]  # outside the radius
    dx, dy = torch.tensor(x, device="cuda"), torch.tensor(y, device="cuda")
    dk = torch.zeros((n, 384), dtype=torch.float64, device="cuda")
    qed.kernel(dx.data_ptr(), dy.data_ptr(), n, dk.data_ptr())
    torch.cuda.synchronize()
    want = np.zeros((n, 384))
    oracle.oracle_qed_kernel(P(x), P(y), n, P(want))
    got = dk.cpu().numpy()
    scale = np.max(np.abs(want))
    # outside the radius: exactly zero; at x = 0 or y = 0 the subtraction cancels
    # L0 against itself, zero up to the rounding of the two evaluations
    assert np.all(got[2] == 0)
    assert np.max(np.abs(got[:2])) <= 1e-14 * scale
    assert np.max(np.abs(got - want)) <= 1e-12 * scale


def test_qed_channels(qed, oracle):
    """Interpolated table channels at (|x|, beta, |y|), including the edges."""
    rng = np.random.default_rng(1)
    R = 12.0 * 105.6583745 / 197.3269804
    n = 100
    acb = np.stack([rng.uniform(0, R, n), rng.uniform(0, np.pi, n), rng.uniform(0, R, n)], 1)
    acb[0] = [0, 0, 0]
    acb[1] = [R, np.pi, R]
    acb[2] = [1e-3, 1e-4, 1e-5]
    d_in = torch.tensor(acb, device="cuda")
    d_out = torch.zeros((n, 42), dtype=torch.float64, device="cuda")
    qed.channels(d_in.data_ptr(), n, d_out.data_ptr())
    torch.cuda.synchronize()
    got = d_out.cpu().numpy()
    for i in range(n):
        ch = np.zeros(42)
        oracle.oracle_channels_interp(*acb[i], P(ch))
        assert np.allclose(got[i], ch, rtol=1e-13, atol=1e-13 * np.abs(ch).max() + 1e-300), i


GOLD = __import__("json").load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))
# the oracle's tolerance against the reference's KQED-based golden sums (tests/test_oracle.py)
GOLDEN_TOL = 2e-2


def test_golden_kernel_sum_gpu(qed):
    """The FP64 GPU kernel sum on the reference's golden geometry
    (sanity_check/test_contract_random.cpp:37-58) against its golden values."""
    g = GOLD["test_contract_random"]
    lat = H.Lattice(g["lattice"])
    con = H.Contractor(lat, qed)
    S, Y = dev_block(lat, g["seed_x"]), dev_block(lat, g["seed_y"])
    got = np.array(con.kernel_sum(S.data_ptr(), Y.data_ptr(), g["gsx"], g["yv"], g["xunit"]))
    con.close()
    want = np.array(g["kernel_sum"])
    log("golden kernel sum", got, want, (got - want) / want)
    assert np.all(np.abs(got - want) <= GOLDEN_TOL * np.abs(want)), (got, want)


PAIR_CASES = [
    # the reference's golden-test geometry
    ((8, 4, 4, 4), (1, 2), (0, 0, 0, 0), (1, 1, 1, 1), 1.0),
    ((8, 4, 4, 4), (1, 2), (3, 1, 2, 0), (-1, 2, 1, -2), 0.3),
    ((8, 4, 4, 4), (5, 6), (7, 3, 0, 1), (4, -2, 2, 1), 0.05),
    ((4, 4, 4, 4), (7, 8), (2, 1, 3, 0), (2, 2, -2, 2), 0.9),
    # the plain small benchmark geometry: 8^4, xunit = m_mu a at a = 0.1 fm
    ((8, 8, 8, 8), (0, 1), (1, 1, 1, 1), (1, 1, 1, 1), 0.0535446292),
]


@pytest.mark.parametrize("dims,seeds,gsx,yv,xunit", PAIR_CASES)
def test_kernel_sum_fp64_vs_oracle(qed, oracle, dims, seeds, gsx, yv, xunit):
    log("pair", dims, seeds, gsx, yv, xunit)
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    S, Y = dev_block(lat, seeds[0]), dev_block(lat, seeds[1])
    got = np.array(con.kernel_sum(S.data_ptr(), Y.data_ptr(), gsx, yv, xunit))
    want = oracle_pair(oracle, S.cpu().numpy(), Y.cpu().numpy(), dims, gsx, yv, xunit)
    con.close()
    assert np.all(np.abs(got - want) <= 1e-10 * np.abs(want)), (got, want)


def test_kernel_sum_deterministic(qed):
    lat = H.Lattice((8, 8, 8, 8))
    con = H.Contractor(lat, qed)
    S, Y = dev_block(lat, 3), dev_block(lat, 4)
    a = con.kernel_sum(S.data_ptr(), Y.data_ptr(), (0, 0, 0, 0), (1, 1, 1, 1), 0.05)
    b = con.kernel_sum(S.data_ptr(), Y.data_ptr(), (0, 0, 0, 0), (1, 1, 1, 1), 0.05)
    assert a == b
    con.close()


def test_seam2_launchers_vs_oracle(qed, oracle):
    """cu_dzu_dzsu / gamma insertion / cu_4pt_contraction, per flavour, accumulating."""
    dims, gsx, yv, xunit = (8, 4, 4, 4), (1, 0, 2, 3), (1, -1, 1, 1), 0.4
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    S, Y = dev_block(lat, 21), dev_block(lat, 22)
    hs, hy = S.cpu().numpy(), Y.cpu().numpy()
    g = np.array(dims, np.int32)
    z4 = np.zeros(4, np.int32)
    for f in (0, 1):
        dzu = torch.zeros(6 * 144 * 2, dtype=torch.float64, device="cuda")
        dzsu = torch.zeros(4 * 144 * 2, dtype=torch.float64, device="cuda")
        con.dzu_dzsu(dzu.data_ptr(), dzsu.data_ptr(), S.data_ptr(), Y.data_ptr(), f, gsx)
        wu, ws = np.zeros(6 * 288), np.zeros(4 * 288)
        oracle.oracle_dzu_dzsu(P(hs), P(hy), f, P(g), P(g), P(z4), P(np.array(gsx, np.int32)), P(wu), P(ws))
        assert np.allclose(dzu.cpu().numpy(), wu, rtol=1e-11, atol=1e-11 * np.abs(wu).max())
        assert np.allclose(dzsu.cpu().numpy(), ws, rtol=1e-11, atol=1e-11 * np.abs(ws).max())
        gdzu = torch.zeros(6 * 4 * 288, dtype=torch.float64, device="cuda")
        gdzsu = torch.zeros(4 * 4 * 288, dtype=torch.float64, device="cuda")
        H.insert_g5_gamma(dzu.data_ptr(), dzsu.data_ptr(), gdzu.data_ptr(), gdzsu.data_ptr())
        gu, gs = np.zeros(6 * 4 * 288), np.zeros(4 * 4 * 288)
        oracle.oracle_insert_gamma(P(wu), P(ws), P(gu), P(gs))
        assert np.array_equal(gdzu.cpu().numpy(), gu.astype(np.float64)) or np.allclose(gdzu.cpu().numpy(), gu, rtol=0, atol=1e-14 * np.abs(gu).max())
        ks = torch.full((1,), 0.25, dtype=torch.float64, device="cuda")  # accumulates onto 0.25
        con.four_pt(ks.data_ptr(), gdzu.data_ptr(), gdzsu.data_ptr(), S.data_ptr(), Y.data_ptr(), f, gsx, xunit, yv)
        want = oracle.oracle_4pt(P(gu), P(gs), P(hs), P(hy), f, P(g), P(g), P(z4), P(np.array(gsx, np.int32)),
                                 P(np.array(yv, np.int32)), xunit, 0)
        assert abs(ks.item() - 0.25 - want) <= 1e-10 * abs(want)
    con.close()


def test_seam2_two_t_slabs_add_up(qed):
    """The sharded sum on one GPU: the 8x4x4x4 lattice as two t-slabs of 4x4x4x4
    (procs (2,1,1,1), proc_coords (0,..) and (1,..)). Each slab's device fill is the
    matching slice of the whole block; the slabs' partial dzu/dzsu (cu_dzu_dzsu) add up
    to the whole lattice's, and their 4-point sums (cu_4pt_contraction with the
    allreduced g_dzu/g_dzsu) add up to the whole lattice's kernel sum, per flavour."""
    dims, gsx, yv, xunit = (8, 4, 4, 4), (5, 1, 2, 3), (2, -1, 1, 2), 0.6
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    S, Y = dev_block(lat, 61), dev_block(lat, 62)
    slabs = [H.Lattice((4, 4, 4, 4), procs=(2, 1, 1, 1), proc_coords=(p, 0, 0, 0)) for p in (0, 1)]
    cons = [H.Contractor(sl, qed) for sl in slabs]
    Vs = 4 ** 4
    part = []
    for p, sl in enumerate(slabs):
        ls, ly = dev_block(sl, 61), dev_block(sl, 62)
        want_s = S.view(24, 2 * Vs, 24)[:, p * Vs:(p + 1) * Vs, :]
        assert torch.equal(ls.view(24, Vs, 24), want_s) and torch.equal(ly.view(24, Vs, 24), Y.view(24, 2 * Vs, 24)[:, p * Vs:(p + 1) * Vs, :])
        part.append((ls, ly))
    for f in (0, 1):
        dzu = torch.zeros(6 * 288, dtype=torch.float64, device="cuda")
        dzsu = torch.zeros(4 * 288, dtype=torch.float64, device="cuda")
        con.dzu_dzsu(dzu.data_ptr(), dzsu.data_ptr(), S.data_ptr(), Y.data_ptr(), f, gsx)
        su = torch.zeros_like(dzu)
        ss = torch.zeros_like(dzsu)
        for c, (ls, ly) in zip(cons, part):
            pu, ps = torch.zeros_like(dzu), torch.zeros_like(dzsu)
            c.dzu_dzsu(pu.data_ptr(), ps.data_ptr(), ls.data_ptr(), ly.data_ptr(), f, gsx)
            su += pu
            ss += ps
        assert torch.allclose(su, dzu, rtol=1e-12, atol=1e-12 * dzu.abs().max().item())
        assert torch.allclose(ss, dzsu, rtol=1e-12, atol=1e-12 * dzsu.abs().max().item())
        gdzu = torch.zeros(6 * 4 * 288, dtype=torch.float64, device="cuda")
        gdzsu = torch.zeros(4 * 4 * 288, dtype=torch.float64, device="cuda")
        H.insert_g5_gamma(dzu.data_ptr(), dzsu.data_ptr(), gdzu.data_ptr(), gdzsu.data_ptr())
        whole = torch.zeros(1, dtype=torch.float64, device="cuda")
        con.four_pt(whole.data_ptr(), gdzu.data_ptr(), gdzsu.data_ptr(), S.data_ptr(), Y.data_ptr(), f, gsx, xunit, yv)
        total = 0.0
        for c, (ls, ly) in zip(cons, part):
            k = torch.zeros(1, dtype=torch.float64, device="cuda")
            c.four_pt(k.data_ptr(), gdzu.data_ptr(), gdzsu.data_ptr(), ls.data_ptr(), ly.data_ptr(), f, gsx, xunit, yv)
            total += k.item()
        log("slabs flavour", f, "whole", whole.item(), "sum of slabs", total)
        assert abs(total - whole.item()) <= 1e-10 * abs(whole.item())
    for c in cons:
        c.close()
    con.close()


@pytest.mark.parametrize("dims,seed,gsw,f", [((6, 6, 6, 6), 3, (0, 0, 0, 0), 0),
                                             ((8, 4, 6, 6), 17, (3, 1, 5, 2), 1)])
def test_pieces_2p2_p1_vs_oracle(qed, oracle, dims, seed, gsw, f):
    """The 2+2 piece P1 (kernels.cu:241-340) against the oracle to 1e-10 of its largest
    entry: on the reference's golden geometry (6^4, seed 3, flavour 0), whose total is
    the reference's golden sum, and on a non-cubic lattice with the source off the
    origin (bins beyond L_rho zero). Two calls agree bitwise (fixed summation order)."""
    import json
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    blk = dev_block(lat, seed)
    lmax = max(dims)
    d = torch.empty(64 * lmax, dtype=torch.float64, device="cuda")
    con.pieces_2p2_p1(blk.data_ptr(), gsw, d.data_ptr(), iflavor=f)
    got = d.cpu().numpy().reshape(4, 4, 4, lmax)
    con.pieces_2p2_p1(blk.data_ptr(), gsw, d.data_ptr(), iflavor=f)
    assert np.array_equal(d.cpu().numpy().reshape(4, 4, 4, lmax), got)
    want = np.zeros((4, 4, 4, lmax))
    oracle.oracle_p1.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_int, ctypes.c_void_p]
    oracle.oracle_p1(P(blk.cpu().numpy()), P(np.array(dims, np.int32)), f, P(np.array(gsw, np.int32)),
                     lmax, P(want))
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    log("P1", dims, "max rel err", err)
    assert err <= 1e-10
    for rho, L in enumerate(dims):
        assert np.all(got[rho, :, :, L:] == 0)
    g = gold["sum_p1_2p2"]
    if (dims, seed, f) == (tuple(g["lattice"]), g["seed"], g["flavour"]):
        assert abs(got.sum() - g["sum_p1"]) <= 1e-10 * abs(g["sum_p1"])
    con.close()


def test_pieces_2p2_vs_oracle(qed, oracle):
    """All 2+2 pieces (kernels.cu:241-760) against the oracle to 1e-10 of the largest entry
    per target, on 6^4 (seed 3) with the source off the origin: P1 as the P1-only entry;
    P2_0, P2_1, P3, P4_0, P4_1 for three targets, the source itself among them (there the
    Lambda-subtracted kernel vanishes in P2 and P3)."""
    dims, seed, w, xunit = (6, 6, 6, 6), 3, (1, 2, 0, 3), 0.7
    targets = [(3, 3, 3, 3), (0, 5, 1, 2), w]
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    blk = dev_block(lat, seed)
    lmax = max(dims)
    d1 = torch.empty(64 * lmax, dtype=torch.float64, device="cuda")
    d1b = torch.empty(64 * lmax, dtype=torch.float64, device="cuda")
    d234 = torch.empty(len(targets) * 320, dtype=torch.float64, device="cuda")
    con.pieces_2p2(blk.data_ptr(), w, targets, xunit, d1.data_ptr(), d234.data_ptr())
    con.pieces_2p2_p1(blk.data_ptr(), w, d1b.data_ptr())
    assert torch.equal(d1, d1b)
    got = d234.cpu().numpy().reshape(len(targets), 320)
    oracle.oracle_p234.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int] + [ctypes.c_void_p] * 2 + \
        [ctypes.c_double, ctypes.c_void_p]
    hb = blk.cpu().numpy()
    for t, tgt in enumerate(targets):
        want = np.zeros(320)
        oracle.oracle_p234(P(hb), P(np.array(dims, np.int32)), 0, P(np.array(w, np.int32)),
                           P(np.array(tgt, np.int32)), xunit, P(want))
        err = np.max(np.abs(got[t] - want)) / np.max(np.abs(want))
        log("P234 target", tgt, "max rel err", err)
        assert err <= 1e-10
    scale = np.max(np.abs(got[2]))
    assert np.max(np.abs(got[2, :192])) <= 1e-12 * scale  # P2_0, P2_1, P3 at the source
    con.close()


def test_pieces_2p2_many_sites_per_wave(qed, oracle):
    """P2..P4 on 12x8x8x8 (6144 sites, more than the launch's waves, so a wave sums
    several sites): against the oracle to 1e-10 of the largest entry per target, and
    bitwise equal on a second call (fixed summation order)."""
    dims, seed, w, xunit = (12, 8, 8, 8), 5, (1, 2, 3, 4), 0.3
    targets = [(0, 0, 0, 0), (6, 4, 4, 4)]
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    blk = dev_block(lat, seed)
    d1 = torch.empty(64 * max(dims), dtype=torch.float64, device="cuda")
    d234 = torch.empty(len(targets) * 320, dtype=torch.float64, device="cuda")
    con.pieces_2p2(blk.data_ptr(), w, targets, xunit, d1.data_ptr(), d234.data_ptr())
    got = d234.cpu().numpy().reshape(len(targets), 320)
    con.pieces_2p2(blk.data_ptr(), w, targets, xunit, d1.data_ptr(), d234.data_ptr())
    assert np.array_equal(d234.cpu().numpy().reshape(len(targets), 320), got)
    oracle.oracle_p234.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int] + [ctypes.c_void_p] * 2 + \
        [ctypes.c_double, ctypes.c_void_p]
    hb = blk.cpu().numpy()
    for t, tgt in enumerate(targets):
        want = np.zeros(320)
        oracle.oracle_p234(P(hb), P(np.array(dims, np.int32)), 0, P(np.array(w, np.int32)),
                           P(np.array(tgt, np.int32)), xunit, P(want))
        err = np.max(np.abs(got[t] - want)) / np.max(np.abs(want))
        log("P234 12x8^3 target", tgt, "max rel err", err)
        assert err <= 1e-10
    con.close()


def test_golden_2p2_sums_gpu(qed):
    """The GPU 2+2 pieces on the reference's golden geometry
    (sanity_check/test_contract_random.cpp:72-127) against its golden sums."""
    g = GOLD["sum_p1_2p2"]
    lat = H.Lattice(g["lattice"])
    con = H.Contractor(lat, qed)
    blk = dev_block(lat, g["seed"])
    lmax = max(g["lattice"])
    d1 = torch.empty(64 * lmax, dtype=torch.float64, device="cuda")
    d234 = torch.empty(2 * 320, dtype=torch.float64, device="cuda")
    con.pieces_2p2(blk.data_ptr(), g["w"], [tuple(t) for t in g["targets"]], g["xunit"], d1.data_ptr(),
                   d234.data_ptr())
    got = d234.cpu().numpy().reshape(2, 5, 64)
    con.close()
    s23, s4 = got[:, :3].sum(), got[:, 3:].sum()
    log("golden 2+2 sums", s23, g["sum_p23"], s4, g["sum_p4"])
    assert abs(d1.cpu().numpy().sum() - g["sum_p1"]) <= 1e-10 * abs(g["sum_p1"])
    assert abs(s23 - g["sum_p23"]) <= GOLDEN_TOL * abs(g["sum_p23"])
    assert abs(s4 - g["sum_p4"]) <= GOLDEN_TOL * abs(g["sum_p4"])


def test_window_forms(qed):
    """Form 0 (FP64 per pair) equals kernel_sum; form 1 (FP32 products, FP64 Kahan
    sums) agrees with FP64 to 1e-4 relative per pair sum."""
    dims = (8, 8, 8, 8)
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    blocks = [dev_block(lat, s) for s in (31, 32, 33, 34)]
    pairs = [(blocks[0].data_ptr(), blocks[1].data_ptr(), (0, 0, 0, 0), (1, 1, 1, 1)),
             (blocks[2].data_ptr(), blocks[1].data_ptr(), (2, 2, 2, 2), (-1, -1, -1, -1)),
             (blocks[2].data_ptr(), blocks[3].data_ptr(), (2, 2, 2, 2), (1, 1, 1, 1)),
             (blocks[0].data_ptr(), blocks[3].data_ptr(), (0, 0, 0, 0), (3, 3, -3, 3))]
    xunit = 0.0535446292
    ref = [con.kernel_sum(a, b, g, y, xunit) for a, b, g, y in pairs]
    f0 = con.window_kernel_sum(pairs, xunit, form=0)
    assert np.array_equal(np.array(f0), np.array(ref))
    f1 = con.window_kernel_sum(pairs, xunit, form=1)
    f1b = con.window_kernel_sum(pairs, xunit, form=1)  # graph replay
    assert f1 == f1b
    err = np.abs(np.array(f1) - np.array(ref)) / np.abs(np.array(ref))
    log("form 1 rel err", err)
    assert np.all(err < 1e-4)
    con.close()


def test_window_field_store(qed):
    """The kept coefficient fields: windows 1 and 2 of a chain see the same 15 distinct
    yv, so window 2 after window 1 builds none and its sums equal those of window 2
    with every field rebuilt, bit for bit; a call with one new yv builds one; a changed
    xunit builds all again."""
    dims = (8, 8, 8, 8)
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    xunit, seed = 0.3, 9
    sch = H.Schedule((1, 1, 1, 1), 1, 2)
    wrap = lambda v: tuple(a % L for a, L in zip(v, dims))
    blocks = {}

    def window(k):
        pairs = [(q, p) for w, p, parts in sch.sources() if w == k for q in parts]
        for q, p in pairs:
            for pos in (q, p):
                if wrap(pos) not in blocks:
                    blocks[wrap(pos)] = dev_block(lat, H.source_seed(seed, wrap(pos)))
        return [(blocks[wrap(q)].data_ptr(), blocks[wrap(p)].data_ptr(), wrap(q), lat.site_map_diff(p, q))
                for q, p in pairs]

    w1, w2 = window(1), window(2)
    assert len(w1) == len(w2) == 65
    yv1 = {lat.site_map_diff(d[3], (0, 0, 0, 0)) for d in w1}
    assert yv1 == {lat.site_map_diff(d[3], (0, 0, 0, 0)) for d in w2}
    con.window_kernel_sum(w1, xunit, form=1)
    assert con.fields_built() == len(yv1)
    kept = con.window_kernel_sum(w2, xunit, form=1)
    assert con.fields_built() == 0
    con.drop_fields()
    fresh = con.window_kernel_sum(w2, xunit, form=1)
    assert con.fields_built() == len(yv1)
    assert kept == fresh
    # one pair with a yv outside the set: one build
    extra = w2[:3] + [(w2[0][0], w2[0][1], w2[0][2], (3, 0, 0, 0))]
    assert (3, 0, 0, 0) not in yv1
    part = con.window_kernel_sum(extra, xunit, form=1)
    assert con.fields_built() == 1
    assert part[:3] == fresh[:3]
    con.window_kernel_sum(w2, xunit * 1.5, form=1)
    assert con.fields_built() == len(yv1)
    # a larger store drops the kept fields: the next call builds all of its own
    slots = con.reserve_fields()
    assert con.reserve_fields(slots + 2) == slots + 2
    assert con.window_kernel_sum(w2, xunit, form=1) == fresh and con.fields_built() == len(yv1)
    # the plain FP64 form of the same pairs is unaffected
    ref = [con.kernel_sum(*d, xunit) for d in w2[:2]]
    err = np.abs(np.array(fresh[:2]) - np.array(ref)) / np.abs(np.array(ref))
    assert np.all(err < 1e-4)
    con.close()


def test_tuning_shapes(qed):
    """Launch shapes (hlbl_tuning) change only the order of the FP64 sums across
    workgroups: window and plain pair sums under other shapes agree with the defaults'
    to 1e-12; out-of-range shapes are refused."""
    dims = (8, 8, 8, 8)
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed, tuning="default")
    blocks = [dev_block(lat, s) for s in (41, 42, 43)]
    pairs = [(blocks[0].data_ptr(), blocks[1].data_ptr(), (0, 0, 0, 0), (1, 1, 1, 1)),
             (blocks[2].data_ptr(), blocks[1].data_ptr(), (2, 2, 2, 2), (-1, -1, -1, -1)),
             (blocks[2].data_ptr(), blocks[0].data_ptr(), (2, 2, 2, 2), (2, -2, 2, -2))]
    xunit = 0.3
    base = np.array(con.window_kernel_sum(pairs, xunit, form=1))
    base64 = np.array(con.kernel_sum(*pairs[0], xunit))
    for t in ({"pass1_batch": 1, "pass2_wg_per_cu": 3}, {"pass1_batch": 1 << 20, "field_waves": 4, "field_wg_per_cu": 3},
              {"axes_wg_per_cu": 1, "near_wg_per_cu": 8, "fp64_field_wg_per_cu": 7, "fp64_pass2_wg_per_cu": 1}):
        con.set_tuning(**t)
        got = np.array(con.window_kernel_sum(pairs, xunit, form=1))
        assert con.fields_built() == 0
        con.drop_fields()
        cold = np.array(con.window_kernel_sum(pairs, xunit, form=1))
        got64 = np.array(con.kernel_sum(*pairs[0], xunit))
        assert np.all(np.abs(got - base) <= 1e-12 * np.abs(base)), t
        assert np.all(np.abs(cold - base) <= 1e-12 * np.abs(base)), t
        assert np.all(np.abs(got64 - base64) <= 1e-12 * np.abs(base64)), t
    # a non-positive or non-finite xunit is refused by the kernel-sum and window entries
    for bad in (0.0, -0.3, float("nan"), float("inf")):
        with pytest.raises(H.HlblError):
            con.window_kernel_sum(pairs, bad, form=1)
        with pytest.raises(H.HlblError):
            con.kernel_sum(*pairs[0], bad)
    for bad in ({"field_waves": 6}, {"field_waves": 8, "field_wg_per_cu": 3}, {"pass1_batch": 0}):
        with pytest.raises(H.HlblError):
            con.set_tuning(**bad)
    con.close()


@pytest.mark.parametrize("dims", [(6, 6, 6, 6), (6, 6, 6, 12)])
def test_window_form1_short_lines(qed, dims):
    """Form 1 where a (t, x, y) line is not a whole number of the plane sums' 8-site
    chunks (z extent 6: one short chunk; 12: a full and a short one) against FP64."""
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    blocks = [dev_block(lat, s) for s in (41, 42, 43)]
    pairs = [(blocks[0].data_ptr(), blocks[1].data_ptr(), (0, 0, 0, 0), (1, 1, 1, 1)),
             (blocks[2].data_ptr(), blocks[1].data_ptr(), (1, 2, 3, 5), (-1, 2, -1, 1)),
             (blocks[2].data_ptr(), blocks[0].data_ptr(), (1, 2, 3, 5), (2, 0, 1, -2))]
    xunit = 0.0535446292
    ref = np.array([con.kernel_sum(a, b, g, y, xunit) for a, b, g, y in pairs])
    f1 = np.array(con.window_kernel_sum(pairs, xunit, form=1))
    err = np.abs(f1 - ref) / np.abs(ref)
    log(dims, "form 1 rel err", err)
    assert np.all(err < 1e-4)
    con.close()


def test_run_chain_tiers_and_direct(qed):
    """The schedule loop (src/driver.cpp:23-165): rows in order, device tier = host tier,
    pairs recomputed directly, cache peak = simulate's, window batches = pair by pair."""
    dims, xunit, seed = (6, 6, 6, 6), 1.0, 7
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    sch = H.Schedule((1, 1, 1, 1), 1, 2)
    expect = sch.simulate(True)
    dev_rows, dst = H.run_chain(con, sch, seed, xunit, residency=0, form=0, batch=0)
    host_rows, hst = H.run_chain(con, sch, seed, xunit, residency=1, form=0, batch=0)
    assert len(dev_rows) == expect["n_pairs"] == 15 + 2 * 65
    assert dst["cache_peak"] == expect["peak_cache"] and hst["cache_peak"] == expect["peak_cache"]
    assert dev_rows == host_rows
    order = [(tuple(np.mod(p, 6)), tuple(np.mod(q, 6))) for q, p in sch.pairs()]
    assert [(r["y"], r["x"]) for r in dev_rows] == order
    for i in (0, len(dev_rows) // 2, len(dev_rows) - 1):
        r = dev_rows[i]
        X = dev_block(lat, H.source_seed(seed, r["x"]))
        Y = dev_block(lat, H.source_seed(seed, r["y"]))
        direct = con.kernel_sum(X.data_ptr(), Y.data_ptr(), r["x"], r["yv"], xunit)
        assert direct == r["kernel_sum"]
        assert r["yv"] == lat.site_map_diff(r["y"], r["x"])
    win_rows, _ = H.run_chain(con, sch, seed, xunit, residency=0, form=0, batch=1)
    assert win_rows == dev_rows
    # one call per new source: same rows, and the pair loop's cache peak
    src_rows, sst = H.run_chain(con, sch, seed, xunit, residency=0, form=0, batch=2)
    assert src_rows == dev_rows and sst["cache_peak"] == expect["peak_cache"]
    con.close()


@pytest.mark.parametrize("dims,stride,form", [((8, 8, 8, 8), 1, 0), ((8, 8, 8, 8), 1, 1), ((32, 16, 16, 16), 2, 1)])
@pytest.mark.parametrize("batch", [1, 2])
def test_host_tier_batches(qed, dims, stride, form, batch):
    """The host tier (pinned blocks; src/propagator_cache.cpp:52-68) feeding batched
    calls: one call per window (batch 1) or per new source (batch 2), whose blocks are
    staged on the device from pinned memory on a second stream. Rows, sums and the
    cache peak equal the device tier's bit for bit."""
    xunit, seed = 105.658 / 197.3269804 * 0.1, 5
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    sch = H.Schedule((1, -1, 1, -1), stride, 2)
    print(f"host tier batches: dims {dims} form {form} batch {batch}", flush=True)
    dev_rows, dst = H.run_chain(con, sch, seed, xunit, residency=0, form=form, batch=batch)
    host_rows, hst = H.run_chain(con, sch, seed, xunit, residency=1, form=form, batch=batch)
    assert len(dev_rows) == 15 + 2 * 65
    assert host_rows == dev_rows
    assert hst["cache_peak"] == dst["cache_peak"]
    if batch == 2:
        assert hst["cache_peak"] == sch.simulate(True)["peak_cache"]
    log(f"host tier {dims} form {form} batch {batch}: contract {hst['seconds_contract']:.2f} s "
        f"(device tier {dst['seconds_contract']:.2f} s)")
    con.close()


@pytest.mark.skipif(not os.path.exists(H.H5_LIB_PATH), reason="libhlbl_h5.so absent (no HDF5 at build time)")
@pytest.mark.parametrize("batch", [0, 1])
def test_run_chain_writes_result_file(qed, tmp_path, batch):
    """The driver appends each window's rows to the result file as the window finishes
    (src/driver.cpp:134-137): read back with h5dump, the file holds exactly the rows
    run_chain returns (doubles bitwise), with the chain index given."""
    from test_result_file import dataset
    dims, xunit, seed = (6, 6, 6, 6), 1.0, 7
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    sch = H.Schedule((1, 1, 1, 1), 1, 2)
    path = str(tmp_path / "chain.h5")
    with H.ResultFile(path, dims, [((0, 0, 0, 0), (1, 1, 1, 1), 1, 2)], xunit) as f:
        rows, _ = H.run_chain(con, sch, seed, xunit, residency=0, form=0, batch=batch, writer=f, chain=3)
        assert f.rows() == len(rows) == 15 + 2 * 65
    for name, key in (("x", "x"), ("y", "y"), ("yv", "yv"), ("window", "window"),
                      ("kernel_sum", "kernel_sum"), ("chain", "chain")):
        _, _, vals, _ = dataset(path, "/kernel_sum/" + name, tmp_path)
        assert np.array_equal(vals, np.array([r[key] for r in rows])), name
    assert {r["chain"] for r in rows} == {3}
    con.close()


@pytest.mark.skipif(not os.path.exists(H.H5_LIB_PATH), reason="libhlbl_h5.so absent (no HDF5 at build time)")
def test_run_chains_two_chains(qed, tmp_path):
    """run_chains (src/driver.cpp:169-208) against independent runs, as the reference's
    test_driver_random.cpp:237-310: rows = each chain's own rows with its index, cache
    peak = the larger one, one file with both chains' rows and the chain-list layout."""
    from test_result_file import dataset, has_attribute
    dims, xunit, seed = (6, 6, 6, 6), 1.0, 7
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    chains = [((1, 1, 1, 1), 1, 2), ((-1, 1, -1, 1), 1, 1)]
    schedules = [H.Schedule(h, n, k) for h, n, k in chains]
    path = str(tmp_path / "two.h5")
    info = [((0, 0, 0, 0), h, n, k) for h, n, k in chains]
    with H.ResultFile(path, dims, info, xunit) as f:
        rows, st = H.run_chains(con, schedules, seed, xunit, residency=0, form=0, batch=1, writer=f)
    alone = [H.run_chain(con, s, seed, xunit, residency=0, form=0, batch=1) for s in schedules]
    assert rows == [dict(r, chain=c) for c, (rs, _) in enumerate(alone) for r in rs]
    assert st["pairs"] == len(rows) == sum(len(s.pairs()) for s in schedules)
    assert st["cache_peak"] == max(a[1]["cache_peak"] for a in alone)
    assert st["positions_solved"] == sum(a[1]["positions_solved"] for a in alone)
    _, _, ks, _ = dataset(path, "/kernel_sum/kernel_sum", tmp_path)
    _, _, ch, _ = dataset(path, "/kernel_sum/chain", tmp_path)
    assert np.array_equal(ks, np.array([r["kernel_sum"] for r in rows]))
    assert ch.tolist() == [r["chain"] for r in rows]
    assert not has_attribute(path, "direction")
    con.close()


def test_config_16x32_fp64_and_window(qed, oracle):
    """The configured plain-path workload: 16^3x32, chain K=4, n=2, h=(+1,-1,+1,-1),
    RandomProvider seed 1, a = 0.1 fm. The FP64 path against the oracle to 1e-10 on
    the first, middle and last pairs of the chain; the FP32 window form (all 65 pairs
    of window 1 in one call) against the FP64 path per pair."""
    dims = (32, 16, 16, 16)
    xunit = 105.658 / 197.3269804 * 0.1
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    sch = H.Schedule((1, -1, 1, -1), 2, 4)
    pairs = sch.pairs()
    assert len(pairs) == 275
    prov = H.RandomProvider(lat, 1)
    wrap = lambda v: tuple(a % L for a, L in zip(v, dims))
    cache = {}

    def block(pos):
        if pos not in cache:
            d = torch.empty(H.block_length(lat), dtype=torch.float64, device="cuda")
            prov.compute(wrap(pos), d.data_ptr())
            cache[pos] = d
        return cache[pos]

    for i in (0, len(pairs) // 2, len(pairs) - 1):
        q, p = pairs[i]
        S, Y = block(q), block(p)
        yv = lat.site_map_diff(p, q)
        got = np.array(con.kernel_sum(S.data_ptr(), Y.data_ptr(), wrap(q), yv, xunit))
        want = oracle_pair(oracle, S.cpu().numpy(), Y.cpu().numpy(), dims, wrap(q), yv, xunit)
        log("pair", i, q, p, "gpu", got, "oracle", want)
        assert np.all(np.abs(got - want) <= 1e-10 * np.abs(want)), (i, got, want)
    # window 1 in FP32 against FP64 per pair
    src = sch.sources()
    pw = [(q, p) for w, p, parts in src if w == 1 for q in parts]
    assert len(pw) == 65
    descs = [(block(q).data_ptr(), block(p).data_ptr(), wrap(q), lat.site_map_diff(p, q)) for q, p in pw]
    f1 = np.array(con.window_kernel_sum(descs, xunit, form=1))
    ref = np.array([con.kernel_sum(*d, xunit) for d in descs])
    err = np.abs(f1 - ref) / np.abs(ref)
    log("16^3x32 window 1, form 1: max rel err", err.max(), "median", np.median(err))
    assert err.max() < 1e-3
    con.close()


def test_config1_chain_8x8x8x8(qed, oracle):
    """Configured small plain workload (src/driver.cpp:100-132): 8^4, chain h=(1,1,1,1),
    n=1, K=1, RandomProvider seed 0, m_mu = 105.658 MeV, a = 0.1 fm, device cache. All 80
    rows: indices and order bit-exact against the schedule, every kernel sum against the
    oracle to 1e-10."""
    dims, seed = (8, 8, 8, 8), 0
    xunit = H.xunit(105.658, 0.1)
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    sch = H.Schedule((1, 1, 1, 1), 1, 1)
    rows, st = H.run_chain(con, sch, seed, xunit, residency=0, form=0, batch=0)
    con.close()
    pairs = sch.pairs()
    assert len(rows) == len(pairs) == 80 and sch.simulate(True)["n_propagators"] == 31
    assert [(r["y"], r["x"]) for r in rows] == [(tuple(np.mod(p, dims)), tuple(np.mod(q, dims))) for q, p in pairs]
    blocks = {}

    def host_block(pos):
        if pos not in blocks:
            blocks[pos] = fill_block(oracle, dims, H.source_seed(seed, pos))
        return blocks[pos]

    worst = 0.0
    for r in rows:
        assert r["yv"] == lat.site_map_diff(r["y"], r["x"])
        want = oracle_pair(oracle, host_block(r["x"]), host_block(r["y"]), dims, r["x"], r["yv"], xunit)
        got = np.array(r["kernel_sum"])
        worst = max(worst, np.max(np.abs(got - want) / np.abs(want)))
    log("config 1: 80 pairs, max rel diff to the oracle", worst)
    assert worst <= 1e-10


@pytest.mark.timeout(600)
def test_config3_24x48_window1(qed, oracle):
    """Configured window-batched workload: 24^3x48, chain K=8, n=3, h=(1,1,1,1),
    RandomProvider seed 2. Window 1's 65 pairs in the FP32 form (one call) against the
    FP64 path per pair and flavour, below 1e-3; the window's first and last pairs in FP64
    against the oracle to 1e-10 (about a minute of the oracle each)."""
    dims = (48, 24, 24, 24)
    xunit = H.xunit(105.658, 0.1)
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    sch = H.Schedule((1, 1, 1, 1), 3, 8)
    assert len(sch.pairs()) == 535 and sch.simulate(True)["n_propagators"] == 136
    pw = [(q, p) for w, p, parts in sch.sources() if w == 1 for q in parts]
    assert len(pw) == 65
    prov = H.RandomProvider(lat, 2)
    wrap = lambda v: tuple(a % L for a, L in zip(v, dims))
    cache = {}
    for q, p in pw:
        for pos in (q, p):
            if pos not in cache:
                d = torch.empty(H.block_length(lat), dtype=torch.float64, device="cuda")
                prov.compute(wrap(pos), d.data_ptr())
                cache[pos] = d
    descs = [(cache[q].data_ptr(), cache[p].data_ptr(), wrap(q), lat.site_map_diff(p, q)) for q, p in pw]
    f1 = np.array(con.window_kernel_sum(descs, xunit, form=1))
    ref = np.array([con.kernel_sum(*d, xunit) for d in descs])
    err = np.abs(f1 - ref) / np.abs(ref)
    log("24^3x48 window 1, form 1 against FP64: max rel err", err.max(), "median", np.median(err))
    assert err.max() < 1e-3
    host = {pos: cache[pos].cpu().numpy() for i in (0, len(pw) - 1) for pos in pw[i]}
    for pos in list(cache):
        del cache[pos]
    for i in (0, len(pw) - 1):
        q, p = pw[i]
        want = oracle_pair(oracle, host[q], host[p], dims, wrap(q), lat.site_map_diff(p, q), xunit)
        log(f"24^3x48 pair {i}: gpu", ref[i], "oracle", want)
        assert np.all(np.abs(ref[i] - want) <= 1e-10 * np.abs(want)), i
    con.close()


def test_config_32x64_two_chains_source_batches(qed):
    """The at-scale workload: 32^3x64 (9.66 GB blocks), two chains of K=4 with
    h=(+1,+1,+1,+1) and (-1,+1,-1,+1), n=4, RandomProvider seed 3, at most 16 blocks
    alive. A window touches 31 blocks (300 GB, more than HBM), so each chain runs the
    FP32 window form with one call per new source and all its partners (batch 2): the
    solve order and cache peak of the pair loop (the schedule's interleaved peak, 16).
    The first, middle and last pair of every window against the FP64 path: below 1e-3
    relative per pair and flavour."""
    torch.cuda.empty_cache()  # the chain's blocks are allocated outside torch's cache
    dims = (64, 32, 32, 32)
    xunit, seed = 105.658 / 197.3269804 * 0.1, 3
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    errs = []
    for h in ((1, 1, 1, 1), (-1, 1, -1, 1)):
        torch.cuda.empty_cache()  # the previous chain's check blocks back to the device
        sch = H.Schedule(h, 4, 4)
        assert sch.simulate(True)["peak_cache"] == 16
        rows, st = H.run_chain(con, sch, seed, xunit, residency=0, form=1, batch=2)
        assert len(rows) == 275 and st["cache_peak"] == 16
        slots = con.reserve_fields()
        con.drop_fields()  # the kept fields' memory back before the FP64 checks
        assert con.reserve_fields() == 0
        assert [(r["y"], r["x"]) for r in rows] == \
            [(tuple(np.mod(p, dims)), tuple(np.mod(q, dims))) for q, p in sch.pairs()]
        by_window = {}
        for i, r in enumerate(rows):
            by_window.setdefault(r["window"], []).append(i)
        for idx in by_window.values():
            for i in sorted({idx[0], idx[len(idx) // 2], idx[-1]}):
                r = rows[i]
                X = dev_block(lat, H.source_seed(seed, r["x"]))
                Y = dev_block(lat, H.source_seed(seed, r["y"]))
                ref = con.kernel_sum(X.data_ptr(), Y.data_ptr(), r["x"], r["yv"], xunit)
                errs.append(max(abs(a - b) / abs(b) for a, b in zip(r["kernel_sum"], ref)))
                del X, Y
        log("32^3x64 chain", h, "contract s", st["seconds_contract"], "solve s", st["seconds_solve"])
    log("32^3x64 form 1, sampled pairs: max rel err", max(errs), "median", np.median(errs),
        "field slots during the chains", slots)
    assert len(errs) == 30 and max(errs) < 1e-3
    con.close()


@pytest.mark.skipif(not os.path.exists(H.H5_LIB_PATH), reason="libhlbl_h5.so absent (no HDF5 at build time)")
def test_run_2p2_against_direct_pieces(qed, tmp_path):
    """The 2+2 step of a run (src/driver.cpp:58-95 with plan_2p2) on 6^4, two chains:
    one row per planned source in run order, each bitwise equal to pieces_2p2 on that
    source's RandomProvider block; the unsolved source gets none; the 2p2 group of the
    file holds the same rows."""
    from test_result_file import dataset
    dims, xunit, seed = (6, 6, 6, 6), 0.8, 5
    lat = H.Lattice(dims)
    con = H.Contractor(lat, qed)
    schedules = [H.Schedule((1, 1, 1, 1), 1, 1), H.Schedule((1, 1, 1, 1), 1, 1, (1, 1, 1, 1))]
    sources = [((5, 5, 5, 5), [(0, 0, 0, 0), (2, 3, 4, 5)]), ((2, 2, 2, 2), []),
               ((1, 1, 1, 1), [(1, 1, 1, 1)]), ((3, 0, 0, 0), [(1, 2, 3, 4)])]
    path = str(tmp_path / "r2p2.h5")
    info = [((0, 0, 0, 0), (1, 1, 1, 1), 1, 1), ((1, 1, 1, 1), (1, 1, 1, 1), 1, 1)]
    with H.ResultFile(path, dims, info, xunit, with_2p2=True) as f:
        rows, plan = H.run_2p2(con, lat, seed, schedules, sources, xunit, writer=f)
    assert plan["unsolved"] == [3]
    assert [r["w"] for r in rows] == [(5, 5, 5, 5), (1, 1, 1, 1), (2, 2, 2, 2)]  # run order
    assert [(r["chain"], r["window"]) for r in rows] == [plan["at"][i] for i in (0, 2, 1)]
    lmax = max(dims)
    for r, (site, targets) in zip(rows, [sources[0], sources[2], sources[1]]):
        blk = torch.empty(H.block_length(lat), dtype=torch.float64, device="cuda")
        H.fill_random_block_device(lat, H.source_seed(seed, site), blk.data_ptr())
        d1 = torch.empty(64 * lmax, dtype=torch.float64, device="cuda")
        d234 = torch.empty(max(1, len(targets)) * 320, dtype=torch.float64, device="cuda")
        con.pieces_2p2(blk.data_ptr(), site, targets, xunit, d1.data_ptr(), d234.data_ptr())
        assert np.array_equal(r["p1"], d1.cpu().numpy())
        assert np.array_equal(r["p234"], d234.cpu().numpy()[:320 * len(targets)])
    _, _, w, _ = dataset(path, "/2p2/w", tmp_path)
    assert w.tolist() == [list(r["w"]) for r in rows]
    _, _, p1, _ = dataset(path, "/2p2/P1", tmp_path)
    assert np.array_equal(p1.reshape(len(rows), -1), np.stack([r["p1"] for r in rows]))
    _, _, p234, _ = dataset(path, "/2p2/P234", tmp_path)
    assert np.array_equal(p234.reshape(3, -1), np.concatenate([r["p234"] for r in rows]).reshape(3, -1))
    con.close()
