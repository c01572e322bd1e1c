"""The Eigen drop-in programs linked against host stand-ins of the device
runtime and of the C-ABI routines (tests/eigen_glue/host_shim.cpp: host
memory for device buffers, each routine written from its BLAS definition
with Eigen on the host), run without a GPU.  They check the header's mapping
of Eigen expressions onto routine arguments -- side, stored triangle, op,
leading dimensions, scalar factors, storage-order flips -- for every case of
the device programs, against Eigen's own evaluation of the same expressions,
bit for bit on small-integer operands.  (The device kernels behind the
routines are checked by the GPU tests.)  Skipped when the programs were not
built (they need the Eigen headers at build time)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "eigen_glue", "_build")


@pytest.mark.parametrize("prog", ["test_device_glue_host", "test_device_map_host", "test_device_level3_host"])
def test_eigen_program_on_host_stand_ins(prog):
    path = os.path.join(BUILD, prog)
    if not os.path.exists(path):
        pytest.skip(f"{prog} not built (needs the Eigen headers at build time)")
    r = subprocess.run([path], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "ALL PASSED" in r.stdout and "FAIL " not in r.stdout, r.stdout[-3000:]
