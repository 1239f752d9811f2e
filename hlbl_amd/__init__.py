"""hlbl_amd: the per-pair connected HLbL kernel sum on MI355X (gfx950).

Python mirror of the reference's host interface for this path (ctypes over the
C ABI of include/hlbl_amd.h):

  Lattice / xunit                 include/hlbl/lattice.hpp, include/hlbl/physics.hpp
  fill_random_block, RandomProvider  src/provider.cpp:10-69
  Contractor.kernel_sum           src/contract.cpp:104-114
  cu_dzu_dzsu / cu_4pt_contraction   kernels/cuda_lattice.h:49-57
  build_schedule, check_complete, simulate, max_steps_without_wrap
                                  src/schedule.cpp:34-214

Device buffers are passed as integer addresses (e.g. torch.Tensor.data_ptr()).
The library must be built (hlbl_amd.build); there is no CPU fallback.
"""
import ctypes as C
import math
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhlbl_amd.so")


class HlblError(RuntimeError):
    pass


class Geom(C.Structure):
    _fields_ = [("T", C.c_int), ("LX", C.c_int), ("LY", C.c_int), ("LZ", C.c_int)]


class Coord(C.Structure):
    _fields_ = [("t", C.c_int), ("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class Pair(C.Structure):
    _fields_ = [("a", C.c_double), ("b", C.c_double)]


class Tuning(C.Structure):
    """Launch shapes of a contractor's kernels (hlbl_tuning of include/hlbl_amd.h)."""
    _fields_ = [(n, C.c_int) for n in ("pass1_batch", "field_waves", "field_wg_per_cu", "axes_wg_per_cu",
                                        "near_wg_per_cu", "pass2_wg_per_cu", "fp64_field_wg_per_cu",
                                        "fp64_pass2_wg_per_cu")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning_gfx950.json")


def tuned_shapes(global_dims, path=TUNING_PATH):
    """The recorded launch shapes for a lattice (from hlbl_amd.tune), or None."""
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    entry = table.get("lattices", {}).get("x".join(str(int(d)) for d in global_dims))
    return None if entry is None else entry.get("picks")


class PairDesc(C.Structure):
    _fields_ = [("d_x", C.c_void_p), ("d_y", C.c_void_p), ("gsx", Coord), ("yv", Coord)]


def _geom(v):
    return Geom(*[int(a) for a in v])


def _coord(v):
    return Coord(*[int(a) for a in v])


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "hlbl_last_error": (C.c_char_p, []),
    "hlbl_xunit": (C.c_double, [C.c_double, C.c_double]),
    "hlbl_source_seed": (C.c_uint64, [C.c_uint64, Coord]),
    "hlbl_block_length": (C.c_size_t, [Geom]),
    "hlbl_fill_random_block_host": (C.c_int, [C.c_void_p, Geom, Coord, C.c_uint64, C.c_int]),
    "hlbl_fill_random_block_device": (C.c_int, [C.c_void_p, Geom, Coord, C.c_uint64, C.c_void_p]),
    "hlbl_qed_init": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "hlbl_window_set_judge_bound": (C.c_int, [C.c_void_p, C.c_double]),
    "hlbl_window_fallback": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "hlbl_window_drop_fields": (C.c_int, [C.c_void_p]),
    "hlbl_tuning_defaults": (C.c_int, [C.c_void_p]),
    "hlbl_contractor_set_tuning": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hlbl_contractor_get_tuning": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hlbl_window_fields_built": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "hlbl_window_reserve_fields": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "hlbl_qed_init_file": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_char_p]),
    "hlbl_qed_exact": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p]),
    "hlbl_qed_free": (C.c_int, [C.c_void_p]),
    "hlbl_qed_table_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_double)]),
    "hlbl_qed_table_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "hlbl_qed_kernel": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_void_p]),
    "hlbl_qed_channels": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                    C.c_void_p]),
    "hlbl_qed_kernel_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                          C.c_void_p, C.c_void_p]),
    "hlbl_mlp_param_count": (C.c_int, [C.c_void_p]),
    "hlbl_mlp_init": (C.c_int, [C.c_void_p, C.c_uint64]),
    "hlbl_mlp_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64,
                                 C.c_void_p, C.c_void_p]),
    "hlbl_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "hlbl_mlp_set_domain": (C.c_int, [C.c_void_p, C.c_double]),
    "hlbl_mlp_set_inner_radius": (C.c_int, [C.c_void_p, C.c_double]),
    "hlbl_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "hlbl_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                   C.c_void_p]),
    "hlbl_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                     C.POINTER(C.c_float), C.c_void_p]),
    "hlbl_contractor_create": (C.c_int, [C.POINTER(C.c_void_p), Geom, Geom, Coord, C.c_void_p]),
    "hlbl_contractor_free": (C.c_int, [C.c_void_p]),
    "hlbl_kernel_sum": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, Coord, Coord, C.c_double,
                                  C.POINTER(C.c_double), C.c_void_p]),
    "hlbl_dzu_dzsu": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int, Coord, Coord, Geom, Geom, C.c_void_p]),
    "hlbl_insert_g5_gamma": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p]),
    "hlbl_pieces_2p2_p1": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, Coord, C.c_void_p,
                                     C.c_void_p]),
    "hlbl_pieces_2p2": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, Coord, C.POINTER(Coord), C.c_int,
                                  C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hlbl_4pt_contraction": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_int, Coord, Coord, Pair,
                                       Coord, Geom, Geom, C.c_void_p]),
    "hlbl_window_kernel_sum": (C.c_int, [C.c_void_p, C.POINTER(PairDesc), C.c_int, C.c_double,
                                         C.c_int, C.POINTER(C.c_double), C.c_void_p]),
    "hlbl_window_judge": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "hlbl_window_profile": (C.c_int, [C.c_void_p, C.POINTER(PairDesc), C.c_int, C.c_double,
                                      C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hlbl_schedule_build": (C.c_void_p, [Coord, C.c_int, C.c_int, Coord]),
    "hlbl_schedule_free": (None, [C.c_void_p]),
    "hlbl_schedule_n_windows": (C.c_int, [C.c_void_p]),
    "hlbl_schedule_n_sources": (C.c_int, [C.c_void_p]),
    "hlbl_schedule_source": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(Coord), C.POINTER(C.c_int)]),
    "hlbl_schedule_partner": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(Coord)]),
    "hlbl_schedule_check_complete": (C.c_int, [C.c_void_p, C.c_int]),
    "hlbl_schedule_simulate": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_longlong)]),
    "hlbl_max_steps_without_wrap": (C.c_int, [Coord, C.c_int, Coord, Coord, C.c_int]),
    "hlbl_run_chain": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, C.c_int,
                                 C.c_int, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "hlbl_run_chain_2p2": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                     C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "hlbl_run_chain_each_window": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, C.c_int,
                                             C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
}


def lib():
    """The loaded C library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HlblError(f"{LIB_PATH} is not built: run python -m hlbl_amd.build")
        L = C.CDLL(LIB_PATH)
        missing = [n for n in _SIGS if not hasattr(L, n)]
        if missing:
            raise HlblError(f"{LIB_PATH} lacks {missing}: rebuild with python -m hlbl_amd.build")
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise HlblError(lib().hlbl_last_error().decode())


def xunit(muon_mass_mev, lattice_spacing_fm):
    """m_mu a / (hbar c) (include/hlbl/physics.hpp:9-14)."""
    return lib().hlbl_xunit(muon_mass_mev, lattice_spacing_fm)


class Lattice:
    """Local/global extents and process coordinates; one process here."""

    def __init__(self, local, procs=(1, 1, 1, 1), proc_coords=(0, 0, 0, 0)):
        self.local = tuple(int(v) for v in local)
        self.procs = tuple(int(v) for v in procs)
        self.proc_coords = tuple(int(v) for v in proc_coords)
        self.global_ = tuple(l * p for l, p in zip(self.local, self.procs))

    def volume(self):
        v = 1
        for a in self.local:
            v *= a
        return v

    def local_index(self, c):
        t, x, y, z = c
        return ((t * self.local[1] + x) * self.local[2] + y) * self.local[3] + z

    def reach(self, xunit):
        """Largest distance (muon units) between two sites with minimum-image
        coordinates: xunit * |(T/2, LX/2, LY/2, LZ/2)|. Every argument of the
        QED kernel on this lattice lies within it."""
        return xunit * math.sqrt(sum((L / 2) ** 2 for L in self.global_))

    def site_map_diff(self, a, b):
        """Minimum-image a - b in [-L/2, L/2) (include/hlbl/vec4.hpp:40-44)."""
        out = []
        for ai, bi, L in zip(a, b, self.global_):
            r = (ai - bi) % L
            out.append(r - L if r >= L // 2 else r)
        return tuple(out)


def block_length(lattice):
    return int(lib().hlbl_block_length(_geom(lattice.local)))


def source_seed(seed, gsx):
    return int(lib().hlbl_source_seed(C.c_uint64(seed), _coord(gsx)))


def fill_random_block_host(lattice, seed, out_ptr, n_threads=0):
    """Host fill of one block at out_ptr (block_length doubles)."""
    check(lib().hlbl_fill_random_block_host(C.c_void_p(out_ptr), _geom(lattice.local),
                                            _coord(lattice.proc_coords), C.c_uint64(seed),
                                            n_threads))


def fill_random_block_device(lattice, seed, d_out, stream=None):
    check(lib().hlbl_fill_random_block_device(C.c_void_p(d_out), _geom(lattice.local),
                                              _coord(lattice.proc_coords), C.c_uint64(seed),
                                              stream))


class RandomProvider:
    """A different seeded random field per source position (src/provider.cpp:59-69)."""

    def __init__(self, lattice, seed=0):
        self.lattice = lattice
        self.seed = seed

    def compute(self, gsx, d_out, stream=None):
        fill_random_block_device(self.lattice, source_seed(self.seed, gsx), d_out, stream)
        return 0


TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "qed_kernel_table.bin")


def qed_exact(a, beta, b, quad=None):
    """The 42 frame components of T^I, T^II, T^III at |x| = a, angle beta, |y| = b from
    the kernel's defining integrals (host; quad = (nt, nr, ntheta, npsi) or None for the
    table's own quadrature)."""
    out = (C.c_double * 42)()
    q = None if quad is None else (C.c_int * 4)(*quad)
    check(lib().hlbl_qed_exact(float(a), float(beta), float(b), q, out))
    return list(out)


class QedKernel:
    """L^(Lambda), Lambda = 0.4 (KQED's initialise / QED_Mkernel_L2 replacement). The
    plain kernel interpolates the tabulated T tensors of table_path (default: the
    library's data/qed_kernel_table.bin)."""

    def __init__(self, device=0, table_path=None):
        h = C.c_void_p()
        if table_path is None:
            check(lib().hlbl_qed_init(C.byref(h), device))
        else:
            check(lib().hlbl_qed_init_file(C.byref(h), device, os.fsencode(table_path)))
        self.handle = h

    def close(self):
        if self.handle:
            check(lib().hlbl_qed_free(self.handle))
            self.handle = None

    def table_info(self):
        vals = [C.c_int() for _ in range(4)]
        rad = C.c_double()
        check(lib().hlbl_qed_table_info(self.handle, *[C.byref(v) for v in vals], C.byref(rad)))
        return tuple(v.value for v in vals) + (rad.value,)

    def kernel(self, d_x, d_y, n, d_kerv, mode=0, stream=None):
        check(lib().hlbl_qed_kernel(self.handle, mode, C.c_void_p(d_x), C.c_void_p(d_y), n,
                                    C.c_void_p(d_kerv), stream))

    def kernel_noalloc(self, d_x, d_y, n, d_kerv, stream=None):
        """kernel() in its plain form with the library's fixed scratch: at most 2048 pairs,
        no allocation; calls that overlap on one GPU race."""
        check(lib().hlbl_qed_kernel_noalloc(self.handle, C.c_void_p(d_x), C.c_void_p(d_y), n,
                                            C.c_void_p(d_kerv), stream))

    def channels(self, d_acb, n, d_out, mode=0, stream=None):
        check(lib().hlbl_qed_channels(self.handle, mode, C.c_void_p(d_acb), n,
                                      C.c_void_p(d_out), stream))

    # the MLP surrogate
    def mlp_param_count(self):
        return lib().hlbl_mlp_param_count(self.handle)

    def mlp_init(self, seed=0):
        check(lib().hlbl_mlp_init(self.handle, C.c_uint64(seed)))

    def mlp_train(self, n_samples=1 << 20, epochs=20, batch=1024, lr=2e-3, seed=4, stream=None):
        loss = (C.c_float * epochs)()
        check(lib().hlbl_mlp_train(self.handle, n_samples, epochs, batch, lr, C.c_uint64(seed),
                                   loss, stream))
        return list(loss)

    def mlp_set_domain(self, radius):
        """Train (and use) the surrogate on |x|, |y| <= radius (muon units); outside,
        the table's channels are used. Discards the trained state."""
        check(lib().hlbl_mlp_set_domain(self.handle, float(radius)))

    def mlp_set_inner_radius(self, inner):
        """Train (and use) the surrogate on |x|, |y| >= inner (default 0.125 muon units);
        below, the table's channels are used. Discards the trained state."""
        check(lib().hlbl_mlp_set_inner_radius(self.handle, float(inner)))

    def mlp_get_params(self):
        n = self.mlp_param_count()
        buf = (C.c_float * n)()
        check(lib().hlbl_mlp_get_params(self.handle, buf, n))
        return buf

    def mlp_set_params(self, params):
        n = self.mlp_param_count()
        buf = (C.c_float * n)(*params)
        check(lib().hlbl_mlp_set_params(self.handle, buf, n))

    def mlp_forward(self, d_in, n, d_out, precision=1, stream=None):
        check(lib().hlbl_mlp_forward(self.handle, C.c_void_p(d_in), n, C.c_void_p(d_out), precision,
                                     stream))

    def mlp_loss_grad(self, d_in, d_target, n, d_grad, stream=None):
        loss = C.c_float()
        check(lib().hlbl_mlp_loss_grad(self.handle, C.c_void_p(d_in), C.c_void_p(d_target), n,
                                       C.c_void_p(d_grad), C.byref(loss), stream))
        return loss.value


class Contractor:
    """The connected kernel sum of a pair (include/hlbl/contract.hpp:42-94)."""

    def __init__(self, lattice, qed, tuning="recorded"):
        """tuning: "recorded" (the shapes hlbl_amd.tune recorded for this lattice, else the
        defaults), "default", or a dict of hlbl_tuning fields (the rest default)."""
        h = C.c_void_p()
        check(lib().hlbl_contractor_create(C.byref(h), _geom(lattice.global_),
                                           _geom(lattice.local), _coord(lattice.proc_coords),
                                           qed.handle))
        self.handle = h
        self.lattice = lattice
        self.qed = qed
        if tuning == "recorded":
            tuning = tuned_shapes(lattice.global_)
        if isinstance(tuning, dict):
            self.set_tuning(**tuning)

    def tuning(self):
        t = Tuning()
        check(lib().hlbl_contractor_get_tuning(self.handle, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **fields):
        """Set launch shapes (hlbl_tuning fields); unnamed fields keep the defaults."""
        t = Tuning()
        check(lib().hlbl_tuning_defaults(C.byref(t)))
        for k, v in fields.items():
            if k not in dict(Tuning._fields_):
                raise HlblError(f"unknown tuning field {k}")
            setattr(t, k, int(v))
        check(lib().hlbl_contractor_set_tuning(self.handle, C.byref(t)))

    def close(self):
        if self.handle:
            check(lib().hlbl_contractor_free(self.handle))
            self.handle = None

    def kernel_sum(self, d_fwd_src, d_fwd_y, gsx, yv, xunit_, stream=None):
        """[flavour 0, flavour 1] connected kernel sum of x = gsx, y = gsx + yv (FP64)."""
        out = (C.c_double * 2)()
        check(lib().hlbl_kernel_sum(self.handle, C.c_void_p(d_fwd_src), C.c_void_p(d_fwd_y),
                                    _coord(gsx), _coord(yv), xunit_, out, stream))
        return [out[0], out[1]]

    def window_kernel_sum(self, pairs, xunit_, form=1, stream=None):
        """pairs: list of (d_x, d_y, gsx, yv). Returns [[f0, f1], ...]."""
        n = len(pairs)
        arr = (PairDesc * n)()
        for i, (dx, dy, gsx, yv) in enumerate(pairs):
            arr[i] = PairDesc(C.c_void_p(dx), C.c_void_p(dy), _coord(gsx), _coord(yv))
        out = (C.c_double * (2 * n))()
        check(lib().hlbl_window_kernel_sum(self.handle, arr, n, xunit_, form, out, stream))
        return [[out[2 * i], out[2 * i + 1]] for i in range(n)]

    def window_profile(self, pairs, xunit_, form=1):
        """One eager window with HIP events per stage: (results, ms[4]) with ms =
        {pass 1, surrogate channels, coefficient fields, pass 2}."""
        n = len(pairs)
        arr = (PairDesc * n)()
        for i, (dx, dy, gsx, yv) in enumerate(pairs):
            arr[i] = PairDesc(C.c_void_p(dx), C.c_void_p(dy), _coord(gsx), _coord(yv))
        out = (C.c_double * (2 * n))()
        ms = (C.c_double * 4)()
        check(lib().hlbl_window_profile(self.handle, arr, n, xunit_, form, out, ms))
        return [[out[2 * i], out[2 * i + 1]] for i in range(n)], list(ms)

    def judge(self):
        e = C.c_double()
        n = C.c_int()
        check(lib().hlbl_window_judge(self.handle, C.byref(e), C.byref(n)))
        return e.value, n.value

    def set_judge_bound(self, bound):
        """Form 2 returns the plain kernel's results (form 1) when the judge exceeds bound."""
        check(lib().hlbl_window_set_judge_bound(self.handle, float(bound)))

    def drop_fields(self):
        """Forget the kept coefficient fields and free their memory (the next window
        call builds its own)."""
        check(lib().hlbl_window_drop_fields(self.handle))

    def fields_built(self):
        """Coefficient fields the last window call built (not served by kept ones)."""
        n = C.c_int()
        check(lib().hlbl_window_fields_built(self.handle, C.byref(n)))
        return n.value

    def reserve_fields(self, n=0):
        """At least n slots for kept coefficient fields; returns the slots the store has."""
        s = C.c_int()
        check(lib().hlbl_window_reserve_fields(self.handle, int(n), C.byref(s)))
        return s.value

    def fell_back(self):
        """Whether the last form-2 window returned the plain kernel's results."""
        f = C.c_int()
        check(lib().hlbl_window_fallback(self.handle, C.byref(f)))
        return bool(f.value)

    # seam 2: the launchers of kernels/cuda_lattice.h
    def dzu_dzsu(self, d_dzu, d_dzsu, fwd_src, fwd_y, iflavor, gsx, stream=None):
        L = self.lattice
        check(lib().hlbl_dzu_dzsu(self.handle, C.c_void_p(d_dzu), C.c_void_p(d_dzsu),
                                  C.c_void_p(fwd_src), C.c_void_p(fwd_y), iflavor,
                                  _coord(L.proc_coords), _coord(gsx), _geom(L.global_),
                                  _geom(L.local), stream))

    def four_pt(self, d_kernel_sum, d_g_dzu, d_g_dzsu, fwd_src, fwd_y, iflavor, gsx, xunit_, yv,
                stream=None):
        L = self.lattice
        check(lib().hlbl_4pt_contraction(self.handle, C.c_void_p(d_kernel_sum),
                                         C.c_void_p(d_g_dzu), C.c_void_p(d_g_dzsu),
                                         C.c_void_p(fwd_src), C.c_void_p(fwd_y), iflavor,
                                         _coord(L.proc_coords), _coord(gsx), Pair(xunit_, 0.0),
                                         _coord(yv), _geom(L.global_), _geom(L.local), stream))


    def pieces_2p2_p1(self, d_fwd, gsw, d_p1, iflavor=0, stream=None):
        """The P1 piece of the 2+2 contribution of the block at d_fwd for source gsw
        (Contractor::pieces_2p2, src/contract.cpp:116-150) into the device array d_p1 of
        [4][4][4][Lmax] doubles, Lmax the largest global extent (synchronous)."""
        check(lib().hlbl_pieces_2p2_p1(self.handle, C.c_void_p(d_fwd), iflavor, _coord(gsw),
                                       C.c_void_p(d_p1), stream))


    def pieces_2p2(self, d_fwd, gsw, targets, xunit_, d_p1, d_p234, iflavor=0, stream=None):
        """All 2+2 pieces of the block at d_fwd for source gsw (Contractor::pieces_2p2,
        src/contract.cpp:116-150): P1 into d_p1 as pieces_2p2_p1, and for each target
        site P2_0, P2_1, P3, P4_0, P4_1 into the device array d_p234 [n][5][4][4][4]."""
        arr = (Coord * max(1, len(targets)))(*[_coord(t) for t in targets])
        check(lib().hlbl_pieces_2p2(self.handle, C.c_void_p(d_fwd), iflavor, _coord(gsw), arr,
                                    len(targets), xunit_, C.c_void_p(d_p1), C.c_void_p(d_p234), stream))


def insert_g5_gamma(d_dzu, d_dzsu, d_g_dzu, d_g_dzsu, stream=None):
    check(lib().hlbl_insert_g5_gamma(C.c_void_p(d_dzu), C.c_void_p(d_dzsu), C.c_void_p(d_g_dzu),
                                     C.c_void_p(d_g_dzsu), stream))


class Schedule:
    """A built sliding-hypercube schedule (src/schedule.cpp:61-91)."""

    def __init__(self, h, stride, n_steps, origin=(0, 0, 0, 0)):
        p = lib().hlbl_schedule_build(_coord(h), stride, n_steps, _coord(origin))
        if not p:
            raise HlblError(lib().hlbl_last_error().decode())
        self.handle = C.c_void_p(p)
        self.stride = stride

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.hlbl_schedule_free(self.handle)
            self.handle = None

    def n_windows(self):
        return lib().hlbl_schedule_n_windows(self.handle)

    def sources(self):
        """[(window, position, [partners...]), ...] in run order."""
        out = []
        L = lib()
        for i in range(L.hlbl_schedule_n_sources(self.handle)):
            w = C.c_int()
            pos = Coord()
            npart = C.c_int()
            check(L.hlbl_schedule_source(self.handle, i, C.byref(w), C.byref(pos),
                                         C.byref(npart)))
            parts = []
            for j in range(npart.value):
                q = Coord()
                check(L.hlbl_schedule_partner(self.handle, i, j, C.byref(q)))
                parts.append((q.t, q.x, q.y, q.z))
            out.append((w.value, (pos.t, pos.x, pos.y, pos.z), parts))
        return out

    def pairs(self):
        """(older, newer) in run order (all_pairs, schedule.cpp:101-109)."""
        return [(q, p) for _w, p, parts in self.sources() for q in parts]

    def check_complete(self):
        check(lib().hlbl_schedule_check_complete(self.handle, self.stride))

    def simulate(self, interleaved=True):
        st = (C.c_longlong * 5)()
        check(lib().hlbl_schedule_simulate(self.handle, 1 if interleaved else 0, st))
        keys = ("n_propagators", "n_solves", "n_pairs", "n_kernel_sums", "peak_cache")
        return dict(zip(keys, [int(v) for v in st]))


def max_steps_without_wrap(h, stride, lattice, origin=(0, 0, 0, 0), cap=10000):
    return lib().hlbl_max_steps_without_wrap(_coord(h), stride, _coord(lattice), _coord(origin),
                                             cap)


class Plan2p2(C.Structure):
    _fields_ = [("w", Coord), ("window", C.c_int), ("targets", C.POINTER(Coord)), ("n_targets", C.c_int)]


def run_chain(contractor, schedule, seed, xunit_, residency=0, form=0, batch=0, writer=None,
              chain=0, plan_2p2=None, iflavor_2p2=0):
    """The schedule loop on random fields (src/driver.cpp:23-165).

    batch 0: one pair per call; 1: one call per window (its sources solved first);
    2: one call per new source with all its partners (the cache peak of batch 0).
    With a ResultFile as writer, each window's rows are appended to it (and flushed)
    when the window is done, as the reference's driver does (src/driver.cpp:134-137);
    chain is the chain index stored in the rows.
    plan_2p2: [(site in [0, L), window, [targets]), ...], the chain's planned 2+2
    sources: computed inside the solve loop from the solved block (src/driver.cpp:58-95),
    appended to the writer (opened with_2p2) after each window's kernel-sum rows.
    Returns (rows, stats), rows dicts {x, y, yv, window, chain, kernel_sum}; with
    plan_2p2 (rows, stats, rows_2p2), rows_2p2 dicts {w, chain, window, p1, targets, p234}
    in run order."""
    n_pairs = len(schedule.pairs())
    buf = (C.c_double * (16 * max(1, n_pairs)))()
    st = (C.c_double * 5)()
    fn = user = None
    if writer is not None:
        fn = C.cast(lib_h5().hlbl_result_append_rows, C.c_void_p)
        user = writer.handle
    if plan_2p2 is None:
        check(lib().hlbl_run_chain_each_window(contractor.handle, schedule.handle, C.c_uint64(seed),
                                               xunit_, residency, form, batch, chain, fn, user, buf,
                                               n_pairs, st))
    else:
        keep = []
        plan = (Plan2p2 * max(1, len(plan_2p2)))()
        for i, (w, window, targets) in enumerate(plan_2p2):
            tg = (Coord * max(1, len(targets)))(*[_coord(t) for t in targets])
            keep.append(tg)
            plan[i] = Plan2p2(_coord(w), int(window), tg, len(targets))
        rows_2p2 = []

        def on_2p2(_user, rows, n):
            try:
                for i in range(n):
                    r = rows[i]
                    nt = r.n_targets
                    rows_2p2.append({"w": (r.w.t, r.w.x, r.w.y, r.w.z), "chain": r.chain, "window": r.window,
                                     "p1": np.ctypeslib.as_array(r.p1, (64 * max(contractor.lattice.global_),)).copy(),
                                     "targets": [(r.targets[k].t, r.targets[k].x, r.targets[k].y, r.targets[k].z)
                                                 for k in range(nt)],
                                     "p234": np.ctypeslib.as_array(r.p234, (320 * nt,)).copy() if nt else
                                     np.zeros(0)})
                if writer is not None and lib_h5().hlbl_result_append_2p2(writer.handle, rows, n) != 0:
                    return 1
                return 0
            except Exception:  # noqa: BLE001 (reported as a failed callback)
                return 1

        cb = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(Source2p2), C.c_int)(on_2p2)
        check(lib().hlbl_run_chain_2p2(contractor.handle, schedule.handle, C.c_uint64(seed), xunit_,
                                       residency, form, batch, chain, fn, user, plan, len(plan_2p2),
                                       iflavor_2p2, C.cast(cb, C.c_void_p), None, buf, n_pairs, st))
    rows = []
    for i in range(n_pairs):
        r = buf[16 * i:16 * i + 16]
        rows.append({"x": tuple(int(v) for v in r[0:4]), "y": tuple(int(v) for v in r[4:8]),
                     "yv": tuple(int(v) for v in r[8:12]), "window": int(r[12]),
                     "chain": int(r[13]), "kernel_sum": [r[14], r[15]]})
    keys = ("pairs", "positions_solved", "cache_peak", "seconds_contract", "seconds_solve")
    if plan_2p2 is None:
        return rows, dict(zip(keys, list(st)))
    return rows, dict(zip(keys, list(st))), rows_2p2


def run_chains(contractor, schedules, seed, xunit_, residency=0, form=0, batch=0, writer=None):
    """run_chains (src/driver.cpp:169-208): the chains one after the other, chain index
    c stored in their rows; rows concatenated in chain order; stats summed, the cache
    peak the largest of the chains'. With a writer, every window of every chain is
    appended as it finishes."""
    if not schedules:
        raise HlblError("run_chains: no chains")
    rows, total = [], None
    for c, sch in enumerate(schedules):
        r, st = run_chain(contractor, sch, seed, xunit_, residency, form, batch, writer, chain=c)
        rows += r
        if total is None:
            total = dict(st)
        else:
            for k, v in st.items():
                total[k] = max(total[k], v) if k == "cache_peak" else total[k] + v
    return rows, total


def plan_2p2(schedules, sources, lattice_dims):
    """plan_2p2 (src/schedule.cpp:284-317): which solves are followed by the 2+2 pieces,
    the first solve in run order (chains in input order, windows, solve order) of each
    2+2 source, sites compared wrapped onto the lattice.

    Returns {"at": {source index: (chain, window)}, "not_sources": [(wrapped site, chain,
    window), ...] (first solves of sites that are no 2+2 source), "unsolved": [source
    indices no chain solves]}."""
    wrap = lambda v: tuple(int(a) % int(L) for a, L in zip(v, lattice_dims))
    index = {}
    for i, src in enumerate(sources):
        index.setdefault(wrap(src), i)
    at, not_sources, solved = {}, [], set()
    for c, sch in enumerate(schedules):
        for window, pos, _parts in sch.sources():
            p = wrap(pos)
            if p in solved:
                continue
            solved.add(p)
            if p in index:
                at[index[p]] = (c, window)
            else:
                not_sources.append((p, c, window))
    unsolved = [i for i, src in enumerate(sources) if wrap(src) not in solved]
    return {"at": at, "not_sources": not_sources, "unsolved": unsolved}


def report_2p2(plan, sources, max_sites=10):
    """report_2p2 (src/schedule.cpp:315-347): the warning lines of a 2+2 plan, solved
    sites that are no 2+2 source, then 2+2 sources no chain solves, at most max_sites
    sites listed each and the rest counted; no lines without 2+2 sources."""
    lines = []
    if not sources:
        return lines
    ns = plan["not_sources"]
    if ns:
        lines.append(f"WARNING: {len(ns)} solved source positions are not 2+2 sources "
                     "(no 2+2 computed for them)")
        lines += [f"  ({p[0]},{p[1]},{p[2]},{p[3]}) in window {w} of chain {c}" for p, c, w in ns[:max_sites]]
        if len(ns) > max_sites:
            lines.append(f"  ... and {len(ns) - max_sites} more")
    un = plan["unsolved"]
    if un:
        lines.append(f"WARNING: {len(un)} 2+2 sources are not solved by any chain "
                     "(no 2+2 computed for them)")
        lines += ["  ({},{},{},{})".format(*sources[i]) for i in un[:max_sites]]
        if len(un) > max_sites:
            lines.append(f"  ... and {len(un) - max_sites} more")
    return lines


def run_2p2(contractor, lattice, seed, schedules, sources_2p2, xunit_, writer=None, iflavor=0,
            residency=0, form=0, batch=0):
    """A run with its 2+2 step (src/driver.cpp:58-95 with plan_2p2): the chains in turn
    through the library's solve loop, which computes P1 and the P234 rows of its targets
    at the first solve in run order of each 2+2 source (site, targets), from that solved
    block (the RandomProvider field of the run's seed). Each row carries the chain and
    window of its solve; with a ResultFile opened with_2p2 each window's 2+2 rows are
    appended after its kernel-sum rows. Returns (rows_2p2 in run order, plan)."""
    dims = tuple(lattice.global_)
    wrap = lambda v: tuple(int(a) % int(L) for a, L in zip(v, dims))
    plan = plan_2p2(schedules, [s for s, _t in sources_2p2], dims)
    rows = []
    for c, sch in enumerate(schedules):
        entries = [(wrap(sources_2p2[i][0]), w, [wrap(t) for t in sources_2p2[i][1]])
                   for i, (ci, w) in plan["at"].items() if ci == c]
        _r, _st, r2 = run_chain(contractor, sch, seed, xunit_, residency, form, batch, writer, chain=c,
                                plan_2p2=entries, iflavor_2p2=iflavor)
        rows += r2
    return rows, plan


# ---- result file (include/hlbl_h5.h; the reference's ResultWriter, src/output.cpp) ----

H5_LIB_PATH = os.path.join(_HERE, "libhlbl_h5.so")
_h5 = None


class ChainInfo(C.Structure):
    _fields_ = [("start", Coord), ("direction", Coord), ("stride", C.c_int), ("steps", C.c_int)]


class RunInfo(C.Structure):
    _fields_ = [("lattice", Coord), ("procs", Coord), ("nconf", C.c_int),
                ("chains", C.POINTER(ChainInfo)), ("n_chains", C.c_int), ("kernel", C.c_char_p),
                ("muon_mass_mev", C.c_double), ("lattice_spacing_fm", C.c_double),
                ("xunit", C.c_double), ("with_2p2", C.c_int)]


class Source2p2(C.Structure):
    _fields_ = [("w", Coord), ("chain", C.c_int), ("window", C.c_int),
                ("p1", C.POINTER(C.c_double)), ("targets", C.POINTER(Coord)), ("n_targets", C.c_int),
                ("p234", C.POINTER(C.c_double))]


def lib_h5():
    """The loaded result-file library; raises if it has not been built (it needs HDF5)."""
    global _h5
    if _h5 is None:
        if not os.path.exists(H5_LIB_PATH):
            raise HlblError(f"{H5_LIB_PATH} is not built: HDF5 headers were not found at build time "
                            "(HLBL_HDF5_ROOT)")
        L = C.CDLL(H5_LIB_PATH)
        L.hlbl_result_open.restype = C.c_void_p
        L.hlbl_result_open.argtypes = [C.c_char_p, C.POINTER(RunInfo)]
        L.hlbl_result_append.restype = C.c_int
        L.hlbl_result_append.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
        L.hlbl_result_append_2p2.restype = C.c_int
        L.hlbl_result_append_2p2.argtypes = [C.c_void_p, C.POINTER(Source2p2), C.c_int]
        L.hlbl_result_rows.restype = C.c_longlong
        L.hlbl_result_rows.argtypes = [C.c_void_p]
        L.hlbl_result_close.restype = C.c_int
        L.hlbl_result_close.argtypes = [C.c_void_p]
        L.hlbl_h5_last_error.restype = C.c_char_p
        _h5 = L
    return _h5


class ResultFile:
    """The reference's ResultWriter (src/output.cpp:131-247): one HDF5 file per run with
    the kernel_sum group (x, y, yv, chain, window, kernel_sum), the chains table and the
    run attributes; every append is flushed.

    chains: [(start, direction, stride, steps)], steps resolved. The kernel attribute
    names the QED kernel used; the default is the reference's, L^(Lambda) at
    Lambda = 0.4."""

    def __init__(self, path, lattice_dims, chains, xunit_, muon_mass_mev=0.0, lattice_spacing_fm=0.0,
                 procs=(1, 1, 1, 1), nconf=-1, kernel="LLambda0.4",
                 with_2p2=False):
        L = lib_h5()
        arr = (ChainInfo * len(chains))(*[ChainInfo(_coord(s), _coord(d), int(st), int(n))
                                          for s, d, st, n in chains])
        info = RunInfo(_coord(lattice_dims), _coord(procs), int(nconf), arr, len(chains),
                       kernel.encode(), float(muon_mass_mev), float(lattice_spacing_fm), float(xunit_),
                       1 if with_2p2 else 0)
        h = L.hlbl_result_open(os.fsencode(path), C.byref(info))
        if not h:
            raise HlblError(L.hlbl_h5_last_error().decode())
        self.handle = C.c_void_p(h)
        self.path = path

    def append(self, rows):
        """rows: dicts as run_chain returns them"""
        buf = (C.c_double * (16 * max(1, len(rows))))()
        for i, r in enumerate(rows):
            buf[16 * i:16 * i + 16] = [*r["x"], *r["y"], *r["yv"], r["window"], r.get("chain", 0),
                                       *r["kernel_sum"]]
        if lib_h5().hlbl_result_append(self.handle, buf, len(rows)) != 0:
            raise HlblError(lib_h5().hlbl_h5_last_error().decode())

    def append_2p2(self, sources):
        """ResultWriter::append_2p2 (src/output.cpp:249-297). sources: dicts {w, chain,
        window, p1 (64 * Lmax doubles), targets [(t, x, y, z), ...], p234 (n_targets * 320
        doubles)} as Contractor.pieces_2p2 gives them (copied to the host)."""
        keep, rows = [], (Source2p2 * max(1, len(sources)))()
        for i, s in enumerate(sources):
            p1 = (C.c_double * len(s["p1"]))(*[float(v) for v in s["p1"]])
            tg = (Coord * max(1, len(s["targets"])))(*[_coord(t) for t in s["targets"]])
            p234 = (C.c_double * max(1, len(s["p234"])))(*[float(v) for v in s["p234"]])
            keep += [p1, tg, p234]
            rows[i] = Source2p2(_coord(s["w"]), int(s.get("chain", 0)), int(s.get("window", 0)), p1, tg,
                                len(s["targets"]), p234)
        if lib_h5().hlbl_result_append_2p2(self.handle, rows, len(sources)) != 0:
            raise HlblError(lib_h5().hlbl_h5_last_error().decode())

    def rows(self):
        return int(lib_h5().hlbl_result_rows(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            rc = lib_h5().hlbl_result_close(self.handle)
            self.handle = None
            if rc != 0:
                raise HlblError(lib_h5().hlbl_h5_last_error().decode())

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, "handle", None) and _h5 is not None:
            _h5.hlbl_result_close(self.handle)
