"""Build the gfx950 PIC/DSMC library in-tree.

    python engineering-degree-in-plasma-simulations_amd/build.py

compiles every csrc/*.hip with hipcc for gfx950 into
engineering-degree-in-plasma-simulations_amd/libpic_mi355.so (the C ABI of
include/pic_abi.h), the host-only output writers into libpic_io.so
(include/pic_io.h), and the driver over the host classes (pic_driver).
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpic_mi355.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-munsafe-fp-atomics",
    # no packed-FP32 instructions (v_pk_*_f32) in any device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    # output writers (host only, no GPU): libpic_io.so
    io_lib = os.path.join(HERE, "libpic_io.so")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "host", "pic_io.cpp"), "-o", io_lib + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("pic_io build failed:\n" + r.stderr[-4000:])
    os.replace(io_lib + ".tmp", io_lib)
    # host classes (the reference's interfaces) + driver, linked to the libraries
    driver = os.path.join(HERE, "pic_driver")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "host", "pic_host.cpp"), os.path.join(HERE, "host", "pic_main.cpp"),
           "-o", driver + ".tmp", "-L", HERE, "-lpic_mi355", "-lpic_io", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n" + r.stderr[-4000:])
    os.replace(driver + ".tmp", driver)
    # the host classes' test executable (run by the GPU tests)
    ht = os.path.join(HERE, "pic_host_test")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "host", "pic_host.cpp"), os.path.join(HERE, "host", "pic_host_test.cpp"),
           "-o", ht + ".tmp", "-L", HERE, "-lpic_mi355", "-lpic_io", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host test build failed:\n" + r.stderr[-4000:])
    os.replace(ht + ".tmp", ht)
    if verbose:
        print("built", LIB, "and", driver)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
