"""Build the native parts of fluidsim_amd in-tree.

    python -m fluidsim_amd.build          # or fluidsim_amd.build.build()

Produces
  fluidsim_amd/lib/libfluidsim_amd.so   HIP kernels + the C ABI of include/fluidsim_amd.h (hipcc, gfx950)
  fluidsim_amd/lib/libfluidsim_host.so  C++ host layer: Config, Parameter, Sim, .vtk output (g++)
  fluidsim_amd/bin/fluidsim             the solver executable (main over fsh_run)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
DEVICE_SRC = PKG / "csrc" / "device"
HOST_SRC = PKG / "csrc" / "host"
BUILD = PKG / "_build"
LIB = PKG / "lib"
BIN = PKG / "bin"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 (v_pk_*_f32) is disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds are fused only inside one source expression: every inlined
    # copy of a formula then rounds the same way (the batched and single-scene
    # kernels, and the pair loop at its two call sites, give identical bits)
    "-ffp-contract=on",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
    f"-I{DEVICE_SRC}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", f"-I{INCLUDE}", f"-I{HOST_SRC}"]

DEVICE_FILES = ["fsa_cells.hip", "fsa_fluid.hip", "fsa_virt.hip", "fsa_solid.hip", "fsa_observe.hip", "fsa_capi.hip",
                "fsa_batch.hip", "fsa_pairs.hip"]
HOST_FILES = ["config.cpp", "parameter.cpp", "sim.cpp", "host_capi.cpp"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    proc = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(map(str, cmd))}\n"
                           f"{proc.stdout}\n{proc.stderr}")
    if verbose and proc.stderr.strip():
        # hipcc's host pass warns that it does not know the packed-fp32 feature; keep the rest
        lines = [l for l in proc.stderr.splitlines() if "packed-fp32-ops" not in l]
        if any(l.strip() for l in lines):
            print("\n".join(lines), file=sys.stderr)


def build(verbose: bool = False, jobs: int = 8) -> dict[str, Path]:
    """Compile every native artefact; returns their paths."""
    if not Path(HIPCC).exists() and shutil.which(HIPCC) is None:
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    for d in (BUILD, LIB, BIN):
        d.mkdir(parents=True, exist_ok=True)

    dev_objs = [BUILD / (f + ".o") for f in DEVICE_FILES]
    host_objs = [BUILD / (f + ".o") for f in HOST_FILES]
    cmds = [[HIPCC, *HIP_FLAGS, "-c", DEVICE_SRC / f, "-o", o] for f, o in zip(DEVICE_FILES, dev_objs)]
    cmds += [[CXX, *HOST_FLAGS, "-c", HOST_SRC / f, "-o", o] for f, o in zip(HOST_FILES, host_objs)]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))

    dev_lib = LIB / "libfluidsim_amd.so"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *dev_objs, "-o", dev_lib], verbose)

    host_lib = LIB / "libfluidsim_host.so"
    _run([CXX, "-shared", "-fPIC", *host_objs, "-o", host_lib, f"-L{LIB}", "-lfluidsim_amd",
          "-Wl,-rpath,$ORIGIN"], verbose)

    exe = BIN / "fluidsim"
    _run([CXX, *HOST_FLAGS, HOST_SRC / "main.cpp", "-o", exe, f"-L{LIB}", "-lfluidsim_host", "-lfluidsim_amd",
          "-Wl,-rpath,$ORIGIN/../lib"], verbose)
    return {"device": dev_lib, "host": host_lib, "exe": exe}


if __name__ == "__main__":
    out = build(verbose=True)
    for k, v in out.items():
        print(f"{k}: {v}")
