"""Build the native parts in-tree: libwfe.so (gfx950 HIP kernels + C ABI +
host case builder), the `wfe` driver executable, and the CPU oracle used by
the tests (oracle/liboracle.so).

Run as ``python -m wfe_amd.build_ext`` or through ``__graft_entry__.build()``.
Everything is compiled on every call (hipcc cross-compiles for gfx950 without
a GPU).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "wfe_amd"
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
OBJDIR = ROOT / "_build"
ORACLE = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC",
    # Packed FP32 instructions are kept out of the device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-Wall", "-Wno-sign-compare"]

HIP_SOURCES = ["hip/context.hip", "hip/diag_api.hip"]
HOST_SOURCES = [
    "core/config.cpp", "core/metric.cpp", "core/base_state.cpp", "core/initial_state.cpp",
    "core/boundary_host.cpp", "core/case.cpp",
    "io/input.cpp", "io/writer.cpp", "capi/case_api.cpp", "dynamics/integrator.cpp",
]
APP_SOURCES = ["apps/wfe_main.cpp", "apps/integrator_check.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if err:
        sys.stderr.write(err + "\n")


def _obj(src: str) -> Path:
    return OBJDIR / (src.replace("/", "_") + ".o")


def build(verbose: bool = False) -> dict:
    OBJDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    inc = [f"-I{CSRC}", f"-I{ROOT / 'include'}"]
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([HIPCC, *HIP_FLAGS, *inc, "-c", str(CSRC / s), "-o", str(_obj(s))])
    for s in HOST_SOURCES:
        jobs.append([CXX, *CXX_FLAGS, *inc, "-c", str(CSRC / s), "-o", str(_obj(s))])
    for s in APP_SOURCES:
        jobs.append([CXX, *CXX_FLAGS, *inc, "-c", str(CSRC / s), "-o", str(_obj(s))])
    jobs.append([CXX, "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-Wall",
                 str(ORACLE / "oracle.cpp"), str(ORACLE / "case_builder.cpp"),
                 "-o", str(ORACLE / "liboracle.so.tmp")])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    (ORACLE / "liboracle.so.tmp").replace(ORACLE / "liboracle.so")

    lib = LIBDIR / "libwfe.so"
    objs = [str(_obj(s)) for s in HIP_SOURCES + HOST_SOURCES]
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
          "-o", str(lib) + ".tmp"])
    Path(str(lib) + ".tmp").replace(lib)
    exe = LIBDIR / "wfe"
    for src, out in ((APP_SOURCES[0], exe), (APP_SOURCES[1], LIBDIR / "wfe_integrator_check")):
        _run([CXX, str(_obj(src)), "-o", str(out) + ".tmp", f"-L{LIBDIR}", "-lwfe",
              "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
        Path(str(out) + ".tmp").replace(out)
    if verbose:
        print(f"built {lib}, {exe}, {ORACLE / 'liboracle.so'}")
    return {"lib": str(lib), "exe": str(exe), "oracle": str(ORACLE / "liboracle.so")}


def clean() -> None:
    shutil.rmtree(OBJDIR, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
