"""Build libhlbl_amd.so in-tree with hipcc for gfx950.

Usage: python -m hlbl_amd.build  (or hlbl_amd.build.build()).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhlbl_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # packed FP32 instructions off (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = (r.stdout + r.stderr).strip()
    lines = [ln for ln in msg.splitlines() if "packed-fp32-ops" not in ln]
    return src, obj, r.returncode, "\n".join(lines)


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    failed = False
    for src, _obj, rc, msg in results:
        if msg and verbose:
            print(f"[{os.path.basename(src)}]\n{msg}", file=sys.stderr)
        failed |= rc != 0
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
           *[o for _s, o, _r, _m in results], "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_h5(verbose)
    build_qedgen()
    return LIB


# The table generator (host C++, OpenMP): writes data/qed_kernel_table.bin from the
# kernel's defining integrals (about 50 minutes on 8 cores); not run by the build.
QEDGEN_SRC = os.path.join(HERE, "qedgen", "gen_qed_table.cpp")
QEDGEN = os.path.join(HERE, "qedgen", "gen_qed_table")


def build_qedgen():
    cmd = [os.environ.get("CXX", "g++"), "-O3", "-march=x86-64-v2", "-std=c++17", "-fopenmp",
           "-Wall", QEDGEN_SRC, "-o", QEDGEN + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gen_qed_table build failed:\n" + r.stdout + r.stderr)
    os.replace(QEDGEN + ".tmp", QEDGEN)
    return QEDGEN


# The result-file library (libhlbl_h5.so): host C++ linked against HDF5, kept apart so
# that libhlbl_amd.so does not depend on HDF5. Built where HDF5's headers are found
# (HLBL_HDF5_ROOT, else /opt/conda); otherwise skipped with a message.
H5_SRC = os.path.join(HERE, "h5", "result_file.cpp")
H5_LIB = os.path.join(HERE, "libhlbl_h5.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def hdf5_root():
    root = os.environ.get("HLBL_HDF5_ROOT", "/opt/conda")
    return root if os.path.exists(os.path.join(root, "include", "hdf5.h")) else None


def build_h5(verbose=True):
    root = hdf5_root()
    if root is None:
        if verbose:
            print("HDF5 headers not found (HLBL_HDF5_ROOT): libhlbl_h5.so not built", file=sys.stderr)
        return None
    tmp = H5_LIB + ".tmp"
    # The C++ runtime is linked in statically: the HDF5 directory's rpath would otherwise
    # also resolve libstdc++ there, and an older one loaded first breaks libhlbl_amd.so.
    # libhdf5 is linked by its full path, so no -L puts that directory's libstdc++ first.
    lib_dir = os.path.join(root, "lib")
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           f"-I{INCLUDE}", f"-I{os.path.join(root, 'include')}", H5_SRC, "-o", tmp,
           os.path.join(lib_dir, "libhdf5.so"), f"-Wl,-rpath,{lib_dir}",
           "-static-libstdc++", "-static-libgcc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("libhlbl_h5.so build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, H5_LIB)
    return H5_LIB


if __name__ == "__main__":
    print(build())
