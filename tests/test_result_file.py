"""The result file (include/hlbl_h5.h) against the reference's format
(src/output.cpp:131-247, README.md:266-290), read back with HDF5's own h5dump tool,
not with this library: layout, values, attributes, the chains table, the one-chain
attributes, the chunk policy, flush-per-append (a process that dies keeps what it
appended) and argument errors. CPU only."""
import json
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

import hlbl_amd as H
from conftest import ROOT

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))["result_file"]
H5DUMP = shutil.which("h5dump") or os.path.join(os.environ.get("HLBL_HDF5_ROOT", "/opt/conda"), "bin", "h5dump")

pytestmark = pytest.mark.skipif(not (os.path.exists(H.H5_LIB_PATH) and os.path.exists(H5DUMP)),
                                reason="libhlbl_h5.so or h5dump absent (no HDF5 at build time)")

CHAIN = ((1, 2, 0, 3), (1, -1, 1, -1), 2, 3)  # start, direction, stride, steps


def rows_of(n, window, chain=0, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        x = tuple(int(v) for v in rng.integers(0, 8, 4))
        y = tuple(int(v) for v in rng.integers(0, 8, 4))
        yv = tuple(int(v) for v in rng.integers(-4, 4, 4))
        out.append({"x": x, "y": y, "yv": yv, "window": window, "chain": chain,
                    "kernel_sum": [float(rng.standard_normal() * 1e3), float(rng.standard_normal() * 1e-7)]})
    return out


def header(path, obj=None):
    cmd = [H5DUMP, "-H", "-p"] + (["-d", obj] if obj else []) + [path]
    return subprocess.run(cmd, capture_output=True, text=True, check=True).stdout


def dataset(path, name, tmp_path):
    """(shape, max shape, values) of a dataset, the values from h5dump's binary output"""
    h = header(path, name)
    m = re.search(r"DATASPACE\s+SIMPLE \{ \( ([^)]*) \) / \( ([^)]*) \) \}", h)
    shape = tuple(int(v) for v in m.group(1).split(","))
    maxshape = tuple(v.strip() for v in m.group(2).split(","))
    dtype = {"H5T_STD_I32LE": "<i4", "H5T_IEEE_F64LE": "<f8"}[re.search(r"DATATYPE\s+(\S+)", h).group(1)]
    out = tmp_path / (name.strip("/").replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", str(out), path], capture_output=True, check=True)
    return shape, maxshape, np.fromfile(out, dtype=dtype).reshape(shape), h


def attribute(path, name):
    """an attribute's values: a string, or a list of numbers"""
    txt = subprocess.run([H5DUMP, "-w", "0", "-m", "%.17g", "-a", "/" + name, path],
                         capture_output=True, text=True, check=True).stdout
    data = txt[txt.index("DATA {") + 6:txt.rindex("}", 0, txt.rindex("}"))]
    s = re.search(r'"(.*)"', data, re.S)
    if s:
        return s.group(1)
    return [float(v) for v in re.sub(r"\(\d+\):", "", data).replace("}", "").split(",")]


def has_attribute(path, name):
    return f'ATTRIBUTE "{name}"' in header(path)


def write(path, chains, windows, xunit=0.0535446292, kernel=None):
    kw = {} if kernel is None else {"kernel": kernel}
    f = H.ResultFile(str(path), (8, 4, 4, 4), chains, xunit, 105.658, 0.1, procs=(1, 1, 1, 1),
                     nconf=-1, **kw)
    for w in windows:
        f.append(w)
    f.close()


def test_library_exports_every_declared_symbol():
    import ctypes
    hdr = open(os.path.join(ROOT, "include", "hlbl_h5.h")).read()
    names = list(dict.fromkeys(re.findall(r"\b(hlbl_[a-z0-9_]+)\s*\(", hdr)))
    lib = ctypes.CDLL(H.H5_LIB_PATH)
    assert len(names) >= 6
    assert not [n for n in names if not hasattr(lib, n)]


def test_layout_and_values(tmp_path):
    path = str(tmp_path / "run.h5")
    windows = [rows_of(3, 0, seed=1), rows_of(5, 1, seed=2), rows_of(0, 2), rows_of(2, 3, seed=3)]
    write(path, [CHAIN], windows)
    rows = [r for w in windows for r in w]
    n = len(rows)
    expect = {"/kernel_sum/x": np.array([r["x"] for r in rows]),
              "/kernel_sum/y": np.array([r["y"] for r in rows]),
              "/kernel_sum/yv": np.array([r["yv"] for r in rows]),
              "/kernel_sum/chain": np.array([r["chain"] for r in rows]),
              "/kernel_sum/window": np.array([r["window"] for r in rows]),
              "/kernel_sum/kernel_sum": np.array([r["kernel_sum"] for r in rows])}
    for name, (dtype, row) in GOLD["datasets"].items():
        shape, maxshape, vals, h = dataset(path, name, tmp_path)
        assert shape == (n, *row), name
        assert maxshape == ("H5S_UNLIMITED", *[str(d) for d in row]), name
        assert vals.dtype == np.dtype(dtype)
        assert np.array_equal(vals, expect[name]), name  # bitwise for the doubles
        # chunks of min(256, 1 MiB of rows) rows
        row_bytes = np.dtype(dtype).itemsize * int(np.prod(row, dtype=int))
        chunk = re.search(r"CHUNKED \( ([^)]*) \)", h).group(1)
        assert tuple(int(v) for v in chunk.split(",")) == (min(256, max(1, 2 ** 20 // row_bytes)), *row)
    shape, _, table, _ = dataset(path, GOLD["chains_table"], tmp_path)
    assert shape == (1, GOLD["chains_columns"])
    assert table[0].tolist() == [*CHAIN[0], *CHAIN[1], CHAIN[2], CHAIN[3]]


def test_attributes(tmp_path):
    path = str(tmp_path / "run.h5")
    write(path, [CHAIN], [rows_of(2, 0)])
    for name, text in GOLD["attribute_texts"].items():
        assert attribute(path, name) == text, name
    assert attribute(path, "lattice") == [8, 4, 4, 4]
    assert attribute(path, "procs") == [1, 1, 1, 1]
    assert attribute(path, "nconf") == [-1]
    assert attribute(path, "muon_mass_MeV") == [105.658]
    assert attribute(path, "lattice_spacing_fm") == [0.1]
    assert attribute(path, "xunit") == [0.0535446292]
    # one chain: also direction, origin, stride (files from before chain lists)
    assert attribute(path, "direction") == list(CHAIN[1])
    assert attribute(path, "origin") == list(CHAIN[0])
    assert attribute(path, "stride") == [CHAIN[2]]
    assert attribute(path, "kernel") == GOLD["kernel_default"]
    write(path, [CHAIN], [], kernel="LSurrogate")
    assert attribute(path, "kernel") == "LSurrogate"
    h = header(path)
    for a in GOLD["attributes_int4"] + GOLD["attributes_int1"]:
        assert re.search(rf'ATTRIBUTE "{a}" \{{\s*DATATYPE\s+H5T_STD_I32LE', h), a
    for a in GOLD["attributes_double"]:
        assert re.search(rf'ATTRIBUTE "{a}" \{{\s*DATATYPE\s+H5T_IEEE_F64LE\s+DATASPACE\s+SCALAR', h), a


def test_two_chains_have_no_single_chain_attributes(tmp_path):
    """As the reference's test_driver_random.cpp:284-310: with a chain list the
    direction/origin/stride attributes are absent and chains has one row per chain."""
    path = str(tmp_path / "run.h5")
    other = ((0, 0, 0, 0), (-1, 1, -1, 1), 1, 2)
    write(path, [CHAIN, other], [rows_of(2, 0, chain=0), rows_of(3, 0, chain=1, seed=5)])
    for a in GOLD["one_chain_only"]:
        assert not has_attribute(path, a)
    shape, _, table, _ = dataset(path, GOLD["chains_table"], tmp_path)
    assert shape == (2, 10)
    assert table[1].tolist() == [*other[0], *other[1], other[2], other[3]]
    _, _, chain, _ = dataset(path, "/kernel_sum/chain", tmp_path)
    assert chain.tolist() == [0, 0, 1, 1, 1]


def test_appends_survive_a_killed_writer(tmp_path):
    """Every append is flushed: a writer process that exits without closing the file
    leaves the windows it appended readable (the reference flushes per window so that a
    killed job keeps its finished windows)."""
    path = str(tmp_path / "killed.h5")
    w0, w1 = rows_of(4, 0, seed=11), rows_of(3, 1, seed=12)
    code = (f"import os, sys; sys.path.insert(0, {ROOT!r}); import hlbl_amd as H\n"
            f"f = H.ResultFile({path!r}, (8, 4, 4, 4), [{CHAIN!r}], 0.5)\n"
            f"f.append({w0!r})\nf.append({w1!r})\nos._exit(0)\n")
    subprocess.run([sys.executable, "-c", code], check=True, timeout=120)
    _, _, ks, _ = dataset(path, "/kernel_sum/kernel_sum", tmp_path)
    assert np.array_equal(ks, np.array([r["kernel_sum"] for r in w0 + w1]))
    _, _, win, _ = dataset(path, "/kernel_sum/window", tmp_path)
    assert win.tolist() == [0] * 4 + [1] * 3


def test_errors(tmp_path):
    with pytest.raises(H.HlblError, match="no chains"):
        H.ResultFile(str(tmp_path / "a.h5"), (8, 4, 4, 4), [], 0.5)
    with pytest.raises(H.HlblError, match="create file"):
        H.ResultFile(str(tmp_path / "missing_dir" / "a.h5"), (8, 4, 4, 4), [CHAIN], 0.5)
    f = H.ResultFile(str(tmp_path / "b.h5"), (8, 4, 4, 4), [CHAIN], 0.5)
    bad = rows_of(1, 0)
    bad[0]["window"] = 0.5
    with pytest.raises(H.HlblError, match="window is not an integer"):
        f.append(bad)
    assert f.rows() == 0
    f.append(rows_of(2, 0))
    assert f.rows() == 2
    f.close()


def test_loads_before_the_main_library():
    """The result-file library carries its own C++ runtime: loaded first in a process,
    it must not bring in an older libstdc++ from HDF5's directory that the main
    library cannot use."""
    code = (f"import sys; sys.path.insert(0, {ROOT!r}); import hlbl_amd as H\n"
            "H.lib_h5(); H.lib(); print('ok')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-2000:]


def test_2p2_group(tmp_path):
    """The 2p2 group (src/output.cpp:146-172, 249-297) read back with h5dump: shapes,
    values bitwise, pair rows per (source, target), pair_yv the minimum image of y - w
    in [-L/2, L/2), the attribute texts; without with_2p2 there is no group and
    appending 2+2 rows fails."""
    g2 = GOLD["group_2p2"]
    dims, lmax = (8, 4, 4, 4), 8
    rng = np.random.default_rng(3)
    srcs = [{"w": (1, 2, 3, 0), "chain": 0, "window": 1, "p1": rng.standard_normal(64 * lmax),
             "targets": [(7, 0, 1, 3), (5, 2, 3, 1)], "p234": rng.standard_normal(2 * 320)},
            {"w": (0, 0, 0, 0), "chain": 1, "window": 0, "p1": rng.standard_normal(64 * lmax),
             "targets": [], "p234": []},
            {"w": (2, 1, 0, 3), "chain": 1, "window": 2, "p1": rng.standard_normal(64 * lmax),
             "targets": [(2, 3, 2, 1)], "p234": rng.standard_normal(320)}]
    path = str(tmp_path / "r.h5")
    f = H.ResultFile(path, dims, [CHAIN, ((0, 0, 0, 0), (1, 1, 1, 1), 1, 2)], 0.5, with_2p2=True)
    f.append(rows_of(2, 0))
    f.append_2p2(srcs[:1])
    f.append_2p2(srcs[1:])
    f.close()
    shape, _, w, _ = dataset(path, "/2p2/w", tmp_path)
    assert w.tolist() == [list(s["w"]) for s in srcs]
    _, _, ch, _ = dataset(path, "/2p2/chain", tmp_path)
    _, _, win, _ = dataset(path, "/2p2/window", tmp_path)
    assert ch.tolist() == [0, 1, 1] and win.tolist() == [1, 0, 2]
    shape, maxshape, p1, _ = dataset(path, "/2p2/P1", tmp_path)
    assert shape == (3, 4, 4, 4, lmax) and maxshape[0] == "H5S_UNLIMITED"
    assert np.array_equal(p1, np.stack([s["p1"] for s in srcs]).reshape(3, 4, 4, 4, lmax))
    pairs = [(s["w"], t, s["p234"][320 * i:320 * (i + 1)]) for s in srcs for i, t in enumerate(s["targets"])]
    shape, _, p234, _ = dataset(path, "/2p2/P234", tmp_path)
    assert shape == (3, 5, 4, 4, 4)
    assert np.array_equal(p234, np.stack([p for _, _, p in pairs]).reshape(3, 5, 4, 4, 4))
    _, _, pw, _ = dataset(path, "/2p2/pair_w", tmp_path)
    _, _, py, _ = dataset(path, "/2p2/pair_y", tmp_path)
    _, _, pyv, _ = dataset(path, "/2p2/pair_yv", tmp_path)
    assert pw.tolist() == [list(a) for a, _, _ in pairs] and py.tolist() == [list(b) for _, b, _ in pairs]
    L = np.array(dims)
    want_yv = (np.array([b for _, b, _ in pairs]) - np.array([a for a, _, _ in pairs]) + L // 2) % L - L // 2
    assert np.array_equal(pyv, want_yv)
    for name, text in g2["attribute_texts"].items():
        assert attribute(path, "2p2/" + name) == text, name
    for name in g2["datasets"]:
        assert f'DATASET "{name.split("/")[-1]}"' in header(path)
    plain = str(tmp_path / "plain.h5")
    f = H.ResultFile(plain, dims, [CHAIN], 0.5)
    with pytest.raises(H.HlblError, match="with_2p2"):
        f.append_2p2(srcs[:1])
    f.close()
    assert 'GROUP "2p2"' not in header(plain)
