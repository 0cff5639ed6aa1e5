"""Build the steps_amd HIP library (gfx950) in-tree.

    python -m steps_amd.build          # builds steps_amd/lib/libsteps_amd.so

Every HIP source is compiled with hipcc for gfx950 only, with packed FP32 VALU
ops disabled, and linked into one shared library that exports the C ABI of
include/steps_force.h.  The C++ drop-in shim (csrc/shim) is compiled against
that library into libsteps_shim_test.so for its tests.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "steps_amd"
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
OBJDIR = ROOT / "_build" / "steps_amd"
LIBNAME = "libsteps_amd.so"
# the drop-in shim built once per reference build variant, linked with the
# test harness that plays the role of StePS's globals (tests/shim)
SHIM_VARIANTS = {"r3": [], "t3": ["-DPERIODIC"], "s1r2": ["-DPERIODIC_Z", "-DEWALD_INTERPOLATION_ORDER=4"],
                 "s1r2_nolookup": ["-DPERIODIC_Z", "-DPERIODIC_Z_NOLOOKUP"],
                 "r3_bh": ["-DUSE_BH=0.5"], "t3_bh": ["-DPERIODIC", "-DUSE_BH=0.5"],
                 "r3_bh_rand": ["-DUSE_BH=0.5", "-DRANDOMIZE_BH=7"],
                 "s1r2_bh_rand": ["-DPERIODIC_Z", "-DEWALD_INTERPOLATION_ORDER=2", "-DUSE_BH=0.5", "-DRANDOMIZE_BH=7"]}

HIP_SOURCES = ["pair_kernels.hip", "engine.hip", "kdk_kernels.hip", "table_kernels.hip", "order_kernels.hip",
               "server.hip", "noalloc.hip", "bh_kernels.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 (v_pk_*_f32) stays off: scalar FP32/FP64 VALU only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build steps_amd)")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(line for line in (r.stdout + r.stderr).splitlines()
                        if "packed-fp32-ops" not in line)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{msg}")


def build(verbose: bool = False) -> Path:
    """Compile every HIP source and link libsteps_amd.so; returns its path."""
    cc = hipcc()
    OBJDIR.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    inc = [f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = OBJDIR / (Path(src).stem + ".o")
        objs.append(obj)
        jobs.append([cc, *HIPCC_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=min(len(jobs), 8)) as ex:
        for f in [ex.submit(_run, j, verbose) for j in jobs]:
            f.result()
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)], verbose)
    os.replace(tmp, out)
    build_shims(verbose)
    build_host(verbose)
    return out


def build_host(verbose: bool = False) -> None:
    """The host side of a StePS run (include/steps_host.h): the parameter-file
    and snapshot library libsteps_host.so, and the steps_run driver linked
    with it and with libsteps_amd.so."""
    host = CSRC / "host"
    common = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", f"-I{host}"]
    lib = LIBDIR / "libsteps_host.so"
    tmp = LIBDIR / "libsteps_host.so.tmp"
    _run([*common, "-fPIC", "-shared", str(host / "params.cc"), str(host / "snapio.cc"), str(host / "cosmo.cc"),
          "-o", str(tmp)], verbose)
    os.replace(tmp, lib)
    exe = LIBDIR / "steps_run"
    tmp = LIBDIR / "steps_run.tmp"
    _run([*common, str(host / "run.cc"), f"-L{LIBDIR}", "-lsteps_host", "-lsteps_amd", "-Wl,-rpath,$ORIGIN",
          "-o", str(tmp)], verbose)
    os.replace(tmp, exe)


def build_shims(verbose: bool = False) -> None:
    """The C++ drop-in shim (host code, g++) for each reference build variant,
    linked with the globals harness of tests/shim into libsteps_shim_<v>.so."""
    shim = CSRC / "shim" / "steps_forces_shim.cc"
    harness = ROOT / "tests" / "shim" / "reference_globals_harness.cc"
    jobs = []
    for name, flags in SHIM_VARIANTS.items():
        out = LIBDIR / f"libsteps_shim_{name}.so"
        tmp = LIBDIR / f"libsteps_shim_{name}.so.tmp"
        jobs.append((["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", *flags,
                      f"-I{ROOT / 'include'}", str(shim), str(harness), f"-L{LIBDIR}", "-lsteps_amd",
                      "-Wl,-rpath,$ORIGIN", "-o", str(tmp)], tmp, out))
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        for f in [ex.submit(_run, j[0], verbose) for j in jobs]:
            f.result()
    for _, tmp, out in jobs:
        os.replace(tmp, out)


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
