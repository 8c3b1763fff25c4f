"""Build libfkadj.so from csrc/ with hipcc for gfx950 (in-tree, next to this file)."""
from __future__ import annotations

import glob
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(PKG), "csrc")
BUILD = os.path.join(os.path.dirname(PKG), "_build")
LIB = os.path.join(PKG, "libfkadj.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 VALU forms are disabled on purpose: ordinary float code must not be turned into
# v_pk_* instructions by the compiler
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = [os.path.join(BUILD, os.path.basename(s)[:-4] + ".o") for s in srcs]

    def cc(pair):
        src, obj = pair
        cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
        if verbose:
            print(" ".join(cmd))

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(cc, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_c_client() -> str:
    """tests/c_abi_client.c, a plain C99 caller of the C ABI (gcc; rpath to this directory)."""
    root = os.path.dirname(os.path.dirname(PKG))
    src = os.path.join(root, "tests", "c_abi_client.c")
    out_dir = os.path.join(root, "tests", "_bin")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "c_abi_client")
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(root, "include"), src,
           "-L", PKG, "-lfkadj", "-Wl,-rpath,$ORIGIN/../../adjoint-optimization_amd/fkadj", "-lm", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed for {src}:\n{r.stderr}")
    return out


def build_map_check() -> str:
    """tests/tile_map_check.cpp: the fused kernels' index maps (csrc/fk_tile_map.hpp) checked on the host."""
    root = os.path.dirname(os.path.dirname(PKG))
    src = os.path.join(root, "tests", "tile_map_check.cpp")
    out_dir = os.path.join(root, "tests", "_bin")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "tile_map_check")
    cmd = [HIPCC, *FLAGS, "-I", CSRC, src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
