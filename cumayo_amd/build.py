"""Builds the HIP engine: one shared library per MAYO parameter set.

``python -m cumayo_amd.build`` compiles ``csrc/mayo_api.hip`` four times
(``-DMAYO_VERSION=1,2,3,5``) for gfx950 into ``cumayo_amd/lib/``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "mayo_api.hip")
LIBDIR = os.path.join(HERE, "lib")
VERSIONS = (1, 2, 3, 5)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CUMAYO_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed fp32 instructions disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def lib_path(version: int) -> str:
    return os.path.join(LIBDIR, f"libcumayo_mayo{version}.so")


def build_one(version: int, extra=()) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = lib_path(version)
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, f"-DMAYO_VERSION={version}", *extra, "-o", tmp, SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for MAYO_{version}:\n{res.stderr[-4000:]}")
    os.replace(tmp, out)
    return out


FLOW_SRC = os.path.join(HERE, "csrc", "bench_flow.cpp")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def flow_path(version: int) -> str:
    return os.path.join(LIBDIR, f"bench_flow_mayo{version}")


def build_flow(version: int) -> str:
    """The reference-flow benchmark program, linked against the library."""
    out = flow_path(version)
    cmd = [HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", FLOW_SRC, f"-L{LIBDIR}",
           f"-lcumayo_mayo{version}", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for bench_flow MAYO_{version}:\n{res.stderr[-4000:]}")
    os.replace(out + ".tmp", out)
    return out


PEAK_SRC = os.path.join(HERE, "csrc", "valu_peak.hip")


def peak_path() -> str:
    return os.path.join(LIBDIR, "valu_peak")


def build_peak() -> str:
    """The VALU issue-rate program behind the roofline's measured peak."""
    os.makedirs(LIBDIR, exist_ok=True)
    out = peak_path()
    flags = [f for f in FLAGS if f not in ("-fPIC", "-shared")]
    res = subprocess.run([HIPCC, *flags, PEAK_SRC, "-o", out + ".tmp"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for valu_peak:\n{res.stderr[-4000:]}")
    os.replace(out + ".tmp", out)
    return out


CONTRACTION_SRC = os.path.join(HERE, "csrc", "contraction_bench.hip")


def contraction_path() -> str:
    return os.path.join(LIBDIR, "contraction_bench")


def build_contraction() -> str:
    """The P1-contraction measurement program: the bitsliced VALU column loop
    against the int8-MFMA GF(2)-plane form of the same product (MAYO_1)."""
    os.makedirs(LIBDIR, exist_ok=True)
    out = contraction_path()
    flags = [f for f in FLAGS if f not in ("-fPIC", "-shared")]
    res = subprocess.run([HIPCC, *flags, "-DMAYO_VERSION=1", CONTRACTION_SRC, "-o", out + ".tmp"],
                         capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for contraction_bench:\n{res.stderr[-4000:]}")
    os.replace(out + ".tmp", out)
    return out


def build_all(extra=()) -> list:
    def one(v):
        lib = build_one(v, extra)
        return [lib, build_flow(v)]
    with ThreadPoolExecutor(max_workers=4) as ex:
        fut = ex.submit(build_peak)
        fut2 = ex.submit(build_contraction)
        built = [p for pair in ex.map(one, VERSIONS) for p in pair]
        return built + [fut.result(), fut2.result()]


if __name__ == "__main__":
    for p in build_all(sys.argv[1:]):
        print(p)
