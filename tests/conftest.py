import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libhlbl_ref.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture(scope="session")
def oracle():
    """The CPU restatement (test infrastructure)."""
    if not os.path.exists(ORACLE):
        pytest.fail(f"{ORACLE} missing: run python -c 'import __graft_entry__ as g; g.build()'")
    o = ctypes.CDLL(ORACLE)
    o.oracle_hash64.restype = ctypes.c_uint64
    o.oracle_hash64.argtypes = [ctypes.c_uint64]
    o.oracle_source_seed.restype = ctypes.c_uint64
    o.oracle_source_seed.argtypes = [ctypes.c_uint64, ctypes.c_void_p]
    o.oracle_fill_block.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
    o.oracle_sum_p1.restype = ctypes.c_double
    o.oracle_4pt.restype = ctypes.c_double
    o.oracle_4pt.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int] + [ctypes.c_void_p] * 5 + \
        [ctypes.c_double, ctypes.c_int]
    o.oracle_kernel_sum.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    o.oracle_channels_interp.argtypes = [ctypes.c_double] * 3 + [ctypes.c_void_p]
    o.oracle_load_table.argtypes = [ctypes.c_char_p]
    o.oracle_table_size.restype = ctypes.c_size_t
    return o


@pytest.fixture(scope="session")
def ref():
    """The reference's own schedule/gamma sources, compiled (only where the reference is present)."""
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    r = ctypes.CDLL(REF)
    return r


def fill_block(oracle, dims, seed):
    dims = np.array(dims, np.int32)
    v = int(np.prod(dims))
    out = np.empty(2 * 12 * 24 * v)
    oracle.oracle_fill_block(P(out), P(dims), None, ctypes.c_uint64(seed))
    return out


def oracle_pair(oracle, S, Y, dims, gsx, yv, xunit, mode=0):
    out = np.zeros(2)
    oracle.oracle_kernel_sum(P(S), P(Y), P(np.array(dims, np.int32)), P(np.array(gsx, np.int32)),
                             P(np.array(yv, np.int32)), ctypes.c_double(xunit), mode, P(out))
    return out
