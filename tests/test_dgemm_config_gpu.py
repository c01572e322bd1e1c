"""FP64 dgemm drop-in at the benchmark's sizes (runs on an MI355X).

C = 1.5*A*B + 0.5*C with A, B, C entries N(0,1) from the device generator,
M = N = K = 4096 and the tall-skinny M = 16384, N = 256, K = 16384, through
the BLAS entry dgemm_ on device pointers.  Errors are measured against the
long-double oracle on sampled rows: the rows of A and C needed are gathered on
the device and copied to the host."""
import ctypes as C

import numpy as np
import pytest

import eigen_amd as E
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,k", [(4096, 4096, 4096), (16384, 256, 16384)])
def test_dgemm_sampled_rows(gpu, m, n, k):
    import torch
    dA = torch.empty(m * k, dtype=torch.float64, device="cuda")
    dB = torch.empty(k * n, dtype=torch.float64, device="cuda")
    dC = torch.empty(m * n, dtype=torch.float64, device="cuda")
    for i, t in enumerate((dA, dB, dC)):
        E.fill(E.F64, t, t.numel(), 11 + i, E.DIST_NORMAL, 0, 1)
    rows = np.arange(0, m, m // 16, dtype=np.int64)
    idx = torch.from_numpy(rows).cuda()
    A_rows = dA.view(k, m)[:, idx].T.contiguous().cpu().numpy()   # (16, k)
    C0_rows = dC.view(n, m)[:, idx].T.contiguous().cpu().numpy()  # (16, n)
    B = dB.view(n, k).T.cpu().numpy()                               # (k, n) column-major values
    torch.cuda.synchronize()
    L = E.blas()
    i = lambda v: C.byref(C.c_int(v))
    al, be = C.c_double(1.5), C.c_double(0.5)
    L.dgemm_(b"N", b"N", i(m), i(n), i(k), C.byref(al), C.c_void_p(dA.data_ptr()), i(m), C.c_void_p(dB.data_ptr()),
             i(k), C.byref(be), C.c_void_p(dC.data_ptr()), i(m))
    got = dC.view(n, m)[:, idx].T.contiguous().cpu().numpy()
    Af = np.asfortranarray(A_rows)
    Bf = np.asfortranarray(B)
    C0f = np.asfortranarray(C0_rows)
    ref, absref = O.gemm_ref(False, False, 1.5, Af, Bf, 0.5, C0f, with_abs=True)
    u = np.finfo(np.float64).eps / 2
    err = np.linalg.norm(got - ref)
    assert err <= O.higham_mary_bound(absref, k, u)
    # normwise metric of the benchmark: ||dC|| / (K u ||A|| ||B||), rows scaled to the whole matrix
    nA = np.linalg.norm(A_rows) * np.sqrt(m / len(rows))
    assert err * np.sqrt(m / len(rows)) / (k * u * nA * np.linalg.norm(B)) <= 1.0
