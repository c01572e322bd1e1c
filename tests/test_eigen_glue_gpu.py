"""Eigen drop-in programs (compiled against the Eigen headers by build()):

* test_device_glue    -- C.device(dev).noalias() {=,+=,-=} expr on device Maps
  through Eigen's AssignmentWithDevice seam (include/eigen_amd_eigen.h);
* test_device_map     -- C = A*B, C.noalias() {=,+=,-=} expr on
  eigen_amd::DeviceMap with no .device(): matrix, vector, row-vector, tiny
  (2x3x1, 1x1) and row-major shapes, real and complex scalars (adjoint
  operands, complex alpha), and complex products through the .device() seam;
* test_device_level3  -- the level-3 views on DeviceMap: triangular solves
  and in-place products, products into a triangle, self-adjoint products
  and rank updates (transposed / adjoint / scaled views, unit diagonals,
  row-major storage, real and complex scalars);
* test_blas_redirect  -- unchanged C.noalias() += alpha*A*B with
  EIGEN_USE_BLAS routed to libeigen_amd_blas (host and device Maps);
* test_blas_level3_redirect -- unchanged matrix-vector, triangular solve,
  triangular and self-adjoint products, rank update and complex products
  with EIGEN_USE_BLAS, each BLAS symbol counted on its way to the library.

Each program compares with Eigen's own CPU evaluation (or an exact loop) on
small-integer operands, bit for bit, and prints PASS/FAIL per case."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "eigen_glue", "_build")


@pytest.mark.parametrize("prog", ["test_device_glue", "test_device_map", "test_device_level3", "test_blas_redirect",
                                  "test_blas_level3_redirect"])
def test_eigen_program(gpu, prog):
    path = os.path.join(BUILD, prog)
    if not os.path.exists(path):
        pytest.skip(f"{prog} not built (needs the Eigen headers at build time)")
    r = subprocess.run([path], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
    assert "FAIL " not in r.stdout
