"""Build step for the HIP library (gfx950), run in-tree on the CPU host."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)                 # mpi-msfec_amd/
REPO = os.path.dirname(PKG_ROOT)
CSRC = os.path.join(PKG_ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(HERE, "libmsfec_hip.so")

SOURCES = ["msfec_q.hip", "msfec_ned.hip", "msfec_rt.hip", "msfec_uneven.hip", "msfec_nrt.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


CPP = os.path.join(PKG_ROOT, "cpp")
DRIVER = os.path.join(CPP, "q_multiscale_driver")


NED_DRIVER = os.path.join(CPP, "q_ned_driver")
RT_DRIVER = os.path.join(CPP, "rt_dq_driver")
NRT_DRIVER = os.path.join(CPP, "ned_rt_driver")
RT3_DRIVER = os.path.join(CPP, "rt_dq3_driver")


def build_driver(verbose: bool = False) -> str:
    """The QBasis, QNedBasis, RTDQBasis (2-D and 3-D) and NedRTBasis C++ shims and their drivers:
    host C++ only, linked against the HIP library through the C-ABI."""
    for shim, main, exe in (("msfec_q_basis.cpp", "q_multiscale_driver.cpp", DRIVER),
                            ("msfec_q_ned_basis.cpp", "q_ned_driver.cpp", NED_DRIVER),
                            ("msfec_rt_dq_basis.cpp", "rt_dq_driver.cpp", RT_DRIVER),
                            ("msfec_ned_rt_basis.cpp", "ned_rt_driver.cpp", NRT_DRIVER),
                            ("msfec_rt_dq3_basis.cpp", "rt_dq3_driver.cpp", RT3_DRIVER)):
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{CPP}",
               os.path.join(CPP, shim), os.path.join(CPP, "msfec_vtu.cpp"), os.path.join(CPP, main),
               f"-L{HERE}", "-l:libmsfec_hip.so", "-Wl,-rpath,$ORIGIN/../msfec_amd",
               "-o", exe + ".tmp"]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        os.replace(exe + ".tmp", exe)
    return DRIVER


def build(verbose: bool = False) -> str:
    """Every source compiled to an object in parallel, then linked."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(PKG_ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc_flags = [f for f in HIPCC_FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *cc_flags, f"-I{INCLUDE}", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        return obj

    with ThreadPoolExecutor(len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    build_driver(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
