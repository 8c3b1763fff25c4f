"""The C ABI from plain C: include/fkadj.h compiles as C99 without warnings (CPU), and the
prebuilt tests/_bin/c_abi_client (tests/c_abi_client.c, built by __graft_entry__.build()) passes its
checks on the GPU: FP64 gradient vs central differences of J, FP32 vs FP64, batch / plan / server
equal to the instances alone, kept trajectory segments changing no result."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLIENT = os.path.join(ROOT, "tests", "_bin", "c_abi_client")


def test_header_is_plain_c99():
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "c_abi_client.c")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_c_client_on_gpu():
    print("[case] c_abi_client 48x80 nt=40 n_amp=4", flush=True)
    assert os.path.exists(CLIENT), "run __graft_entry__.build() first"
    r = subprocess.run([CLIENT], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "c abi ok" in r.stdout
