"""The C++ Builder facade (include/proxpde_amd_builder.hpp): the test program
tests/cpp/facade_test (built by __graft_entry__.build()) drives
Builder<RowMajor>::buildLhs(std::tuple{AssemblyStiffness...}, bcs) / buildRhs /
closeMatrix as the reference's tests do (test/poisson3dhex_q2.cpp and others)
and checks the closed matrix and b against the oracle's Builder, the
reference's known error norm of poisson3dhex_q2 (n = 4), and bitwise equality
on Q1 hex."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "facade_test")


def test_facade_program_is_built_and_linked():
    assert os.path.exists(EXE), "tests/cpp/facade_test missing: run __graft_entry__.build()"
    out = subprocess.run(["ldd", EXE], capture_output=True, text=True, check=True).stdout
    assert "libproxpde_amd.so" in out and "not found" not in out, out


@pytest.mark.gpu
def test_facade_against_oracle(gpu):
    print("case facade_test (C++ Builder facade, 4 cases)", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "FAIL" not in r.stdout
