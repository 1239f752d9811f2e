"""Host-side parts of the library on the CPU: symbol exports, the random-field
provider, units, lattice maps and argument checks (no GPU call)."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import hlbl_amd as H
from conftest import P, fill_block, ROOT

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))


def header_symbols():
    hdr = open(os.path.join(ROOT, "include", "hlbl_amd.h")).read()
    return list(dict.fromkeys(re.findall(r"\b(hlbl_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(H.LIB_PATH)
    names = header_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    H.lib()  # binds every signature


def test_host_fill_matches_oracle(oracle):
    for dims, seed, threads in (((4, 2, 2, 2), 1, 1), ((6, 6, 6, 6), 3, 4), ((8, 4, 4, 4), 123456789, 3)):
        lat = H.Lattice(dims)
        n = H.block_length(lat)
        out = np.empty(n)
        H.fill_random_block_host(lat, seed, out.ctypes.data, threads)
        assert np.array_equal(out, fill_block(oracle, dims, seed))


def test_source_seed_matches_oracle(oracle):
    for seed, gsx in ((0, (0, 0, 0, 0)), (2, (3, 6, 9, 12)), (7, (47, 23, 0, 5))):
        assert H.source_seed(seed, gsx) == oracle.oracle_source_seed(ctypes.c_uint64(seed),
                                                                     P(np.array(gsx, np.int32)))


def test_fill_independent_of_process_grid(oracle):
    """A block on a 2x1x1x1 process grid equals the matching half of the global one."""
    glob = fill_block(oracle, (8, 4, 4, 4), 5).reshape(2, 12, 8, 4 * 4 * 4 * 24)
    lat = H.Lattice((4, 4, 4, 4), procs=(2, 1, 1, 1), proc_coords=(1, 0, 0, 0))
    out = np.empty(H.block_length(lat))
    H.fill_random_block_host(lat, 5, out.ctypes.data, 2)
    assert np.array_equal(out.reshape(2, 12, 4, -1), glob[:, :, 4:8])


def test_xunit():
    for m, a, want in GOLD["xunit"]["cases"]:
        assert abs(H.xunit(m, a) - want) < 5e-11


def test_site_map_diff():
    for a, b, L, want in GOLD["lattice"]["site_map_diff"]:
        assert H.Lattice(L).site_map_diff(a, b) == tuple(want)
    lat = H.Lattice((8, 4, 4, 4))
    assert lat.local_index((0, 0, 0, 1)) == 1 and lat.local_index((1, 0, 0, 0)) == 64


def test_argument_errors_without_gpu():
    lib = H.lib()
    with pytest.raises(H.HlblError):
        H.check(lib.hlbl_fill_random_block_host(None, H.Geom(2, 2, 2, 2), H.Coord(0, 0, 0, 0), 0, 1))
    with pytest.raises(H.HlblError):
        H.check(lib.hlbl_fill_random_block_host(ctypes.c_void_p(1), H.Geom(0, 2, 2, 2), H.Coord(0, 0, 0, 0), 0, 1))
    with pytest.raises(H.HlblError):
        H.check(lib.hlbl_kernel_sum(None, None, None, H.Coord(), H.Coord(), 1.0, (ctypes.c_double * 2)(), None))
    assert H.lib().hlbl_mlp_param_count(None) == -1
    with pytest.raises(H.HlblError, match="null"):
        H.check(lib.hlbl_qed_kernel(None, 0, None, None, 0, None, None))


def test_block_length():
    assert H.block_length(H.Lattice((48, 24, 24, 24))) * 8 == 4608 * 663552
