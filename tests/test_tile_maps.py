"""The fused kernels' index maps, checked on the host with the kernels' own functions
(tests/tile_map_check.cpp over csrc/fk_tile_map.hpp, built by __graft_entry__.build()): thread
ownership of tile and halo cells for every tile configuration and time-blocking depth, stencil
bounds and closure, slot predicates, the LDS region load map, the XCD-aware launch map and the
split of a large batch call into contiguous runs of members."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "_bin", "tile_map_check")


def test_tile_maps():
    if not os.path.exists(CHECK):   # a host program: hipcc builds it without a GPU
        from fkadj import build as fb
        fb.build_map_check()
    r = subprocess.run([CHECK], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "tile maps ok" in r.stdout, r.stderr
