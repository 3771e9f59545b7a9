"""The C example program (tests/c_abi/abi_example.c, built by
__graft_entry__.build) on the GPU: a plain C caller of include/blobs_gpu.h
steps two overlapping spheres and checks separation, gravity and the
contact counter."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "c_abi" / "abi_example"


def test_c_program_steps_on_the_gpu():
    assert EXE.exists(), f"{EXE} is missing: build it with __graft_entry__.build() (make -C tests/c_abi)"
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "c-abi example ok" in r.stdout
