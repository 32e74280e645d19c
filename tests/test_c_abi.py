"""The boundary as a compiled C program sees it: tests/c/abi_check.c includes
the installed contract.h and contract_ext.h, links -lcontract from the
install prefix (build/install, the reference's install layout,
CMakeLists.txt:44-51) and checks the drop-in and extension entry points on
device buffers against tests/golden/c_abi_fixture.bin (FP64 oracle, 1e-12)."""

import filecmp
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "c", "_build", "abi_check")
FIXTURE = os.path.join(ROOT, "tests", "golden", "c_abi_fixture.bin")
PREFIX = os.path.join(ROOT, "build", "install")


def test_install_layout_matches_sources():
    for h in ("contract.h", "contract_enum.h", "contract_ext.h"):
        assert filecmp.cmp(os.path.join(ROOT, "include", h), os.path.join(PREFIX, "include", h), shallow=False), h
    assert filecmp.cmp(os.path.join(ROOT, "contract_amd", "lib", "libcontract.so"),
                       os.path.join(PREFIX, "lib", "libcontract.so"), shallow=False)


def test_c_program_links_the_installed_library():
    out = subprocess.run(["ldd", BIN], capture_output=True, text=True, check=True).stdout
    line = next(ln for ln in out.splitlines() if "libcontract.so" in ln)
    assert os.path.realpath(line.split("=>")[1].split()[0]) == os.path.realpath(
        os.path.join(PREFIX, "lib", "libcontract.so"))


def test_fixture_is_the_oracle_on_the_c_inputs():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_c_abi_fixture
    want = make_c_abi_fixture.results()
    got = np.fromfile(FIXTURE, dtype=np.complex128)
    assert got.shape == want.shape and np.array_equal(got, want)


@pytest.mark.gpu
def test_c_program_on_the_gpu(cuda):
    r = subprocess.run([BIN, FIXTURE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ALL PASS")
    assert r.stdout.count("PASS") == 10
