"""Build libhyperflow_hip.so (HIP for gfx950) in-tree.

Steps: compile the mechanism topology generator (host C++), generate one
topology header per bundled mechanism, compile the kernels once per topology
plus the host C-ABI, and link the shared library next to this file.

    python -m hyperflow_amd.build [--jobs N] [--verbose]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
GEN = BUILD / "gen"
OBJ = BUILD / "obj"
LIB = PKG / "libhyperflow_hip.so"
INCLUDE = PKG.parent / "include"

# (mechanism file, topology struct name)
MECHANISMS = [
    ("H2_Li.yaml", "h2_li"),
    ("dme_sk39.yaml", "dme_sk39"),
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HF_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(map(str, cmd)), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(map(str, cmd))}")
    return r


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    GEN.mkdir(parents=True, exist_ok=True)
    OBJ.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    mechgen = BUILD / "mechgen"
    _run(["g++", "-O2", "-std=c++17", "-o", mechgen, CSRC / "mechgen.cpp", CSRC / "mechanism.cpp"], verbose)
    for fname, name in MECHANISMS:
        _run([mechgen, PKG / "mechanisms" / fname, name, GEN / f"{name}.inc"], verbose)
    (GEN / "registry.inc").write_text("".join(f"HF_TOPOLOGY({n})\n" for _, n in MECHANISMS))

    inc = ["-I", INCLUDE, "-I", CSRC, "-I", GEN]
    units = []
    for _, name in MECHANISMS:
        units.append((CSRC / "topo_instance.hip", OBJ / f"topo_{name}.o",
                      ["-x", "hip", f"-DHF_TOPO={name}", f"-DHF_TOPO_INC={name}.inc"]))
    for src in ("capi.cpp", "registry.cpp"):
        units.append((CSRC / src, OBJ / (Path(src).stem + ".o"), ["-x", "hip"]))
    for src in ("mechanism.cpp", "host_thermo.cpp"):
        units.append((CSRC / src, OBJ / (Path(src).stem + ".o"), []))

    def compile_unit(u):
        src, obj, extra = u
        if extra and extra[:2] == ["-x", "hip"]:
            cmd = [HIPCC, *HIP_FLAGS, *inc, *extra, "-c", src, "-o", obj]
        else:
            cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", *inc, "-c", src, "-o", obj]
        _run(cmd, verbose)
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_unit, units))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    build_consumer(verbose)
    return LIB


CONSUMER_SRC = PKG.parent / "tests" / "abi_consumer.cpp"
CONSUMER = BUILD / "abi_consumer"


def build_consumer(verbose=False):
    """The C++ consumer of the ABI (tests/abi_consumer.cpp), built with the
    host compiler against the header and the library, as a C++ driver links
    it. Skipped when the tests are not present."""
    if not CONSUMER_SRC.exists():
        return None
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", INCLUDE, "-I", "/opt/rocm/include",
          CONSUMER_SRC, "-o", CONSUMER, "-L", PKG, "-lhyperflow_hip", "-L", "/opt/rocm/lib", "-lamdhip64",
          "-Wl,-rpath,$ORIGIN/..:/opt/rocm/lib"], verbose)
    return CONSUMER


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    if LIB.exists():
        LIB.unlink()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args()
    if a.clean:
        clean()
    print(build(a.jobs, a.verbose))
