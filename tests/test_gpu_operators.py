"""FP64 operator precompute on the GPU (kernel sampling, Gram matrix and the
316 blocks E_k = U^T K0_k U on the device, the eigenproblem on the host)
against the host precompute and the oracle's kernel matrix, for n = 4 and
n = 8 (bbfmm.c:1209-1421)."""
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_device_operators_n4_match_host(cuda):
    import bbfmmongpu_amd as B
    t = time.time()
    U, K, sig = B.operators(4, 32, prec="d", device=True)
    td = time.time() - t
    Uh, Kh, sigh = B.operators(4, 32, prec="d")
    print(f"n=4 device precompute {td:.2f}s")
    assert np.allclose(sig[:32], sigh[:32], rtol=1e-9)
    Um, Uhm = U.reshape(32, 64).T, Uh.reshape(32, 64).T
    assert np.abs(Um.T @ Um - np.eye(32)).max() < 1e-12
    P, Ph = Um[:, :30] @ Um[:, :30].T, Uhm[:, :30] @ Uhm[:, :30].T  # complete singular clusters
    assert np.abs(P - Ph).max() < 1e-7
    K0 = O.kernel_matrix(4).reshape(316, 64, 64)  # [k][src][fld]
    E = K.reshape(316, 32, 32)
    for k in (0, 57, 200, 315):
        assert np.allclose(E[k].T, Um.T @ K0[k].T @ Um, atol=1e-12)


def test_device_operators_n8_invariants(cuda):
    import bbfmmongpu_amd as B
    t = time.time()
    U, K, sig = B.operators(8, 256, prec="d", device=True)
    print(f"n=8 device precompute {time.time() - t:.2f}s")
    Um = U.reshape(256, 512).T
    assert np.abs(Um.T @ Um - np.eye(256)).max() < 1e-11
    assert np.all(np.diff(sig) <= 1e-12 * sig[0])
    E = K.reshape(316, 256, 256)  # E[k] = (E_k col-major)^T
    for k in (0, 100, 315):
        assert np.allclose(E[k].T, E[315 - k], atol=1e-11 * np.abs(E[k]).max())
    # two blocks against the oracle's kernel matrix
    K0 = O.kernel_matrix(8).reshape(316, 512, 512)
    for k in (0, 315):
        ref = Um.T @ K0[k].T @ Um
        assert np.abs(E[k].T - ref).max() < 1e-11 * np.abs(ref).max()
