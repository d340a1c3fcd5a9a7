"""Chebyshev order n = 8 (r = 256), the reference's second GPU configuration
(m2l_host_basic.cu:82-97): the dense-grid M2L in plain FMAs and on FP32
MFMA, and full solves with one and sixteen charge vectors, against the
oracle on identical inputs and operators.

The operators here are random (orthonormal U, Gaussian K): parity of the
implementation does not depend on them being the kernel's SVD, and the
FP64 precompute of the real ones takes tens of seconds on the host.
Tolerances (relative L2): 1e-5 fp32, 1e-12 fp64."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

N8, R8 = 512, 256


@pytest.fixture(scope="module")
def ops8():
    rng = np.random.default_rng(8)
    Q, _ = np.linalg.qr(rng.standard_normal((N8, R8)))
    U = np.asfortranarray(Q).ravel(order="F")          # n^3 x r column-major
    K = rng.standard_normal(316 * R8 * R8) / np.sqrt(R8)
    return U, K


def _dt(prec):
    return np.float32 if prec == "s" else np.float64


# (prec, form, tolerance at the L level vs the fp64 oracle on identical inputs)
CASES_M2L = [("s", 0, 1e-5), ("d", 0, 1e-12), ("s", 2, 1e-5)]


@pytest.mark.parametrize("prec,form,tol", CASES_M2L)
@pytest.mark.parametrize("nrhs", [1, 3])
def test_grid_m2l_r256_against_oracle(prec, form, tol, nrhs, ops8, cuda):
    import torch
    import bbfmmongpu_amd as B
    U, K = ops8
    N = 32768
    pos, _q = O.uniform_inputs(N, seed=3)
    plan = B.Plan(pos, 8, U, K, prec=prec, nrhs=nrhs).set(scheme=B.SCHEME_IJ, form=form)
    tree = O.Tree(pos, 0, "d")
    assert plan.maxlev == tree.maxlev == 3
    rng = np.random.default_rng(12)
    M = rng.standard_normal((nrhs, tree.ncell * R8)).astype(_dt(prec))
    Md = torch.from_numpy(M).to(cuda)
    Ld = torch.zeros_like(Md)
    plan.m2l(Md, Ld)
    torch.cuda.synchronize()
    Lg = Ld.cpu().numpy()
    s = O.levsta(3)[2] * R8
    Kd = K.astype(_dt(prec)).astype(np.float64)
    for j in range(nrhs):
        ref = O.m2l(tree, Kd, M[j].astype(np.float64), R8, "sibling")
        err = O.rel_l2(ref[s:], Lg[j, s:])
        print(f"r256 grid m2l prec={prec} form={form} rhs={j}: rel_l2={err:.3e}", flush=True)
        assert err < tol


@pytest.mark.parametrize("prec,form,nrhs,check", [
    ("s", 0, 1, (0,)), ("d", 0, 1, (0,)), ("s", 2, 16, (0, 9, 15)),
])
def test_solve_n8_against_oracle(prec, form, nrhs, check, ops8, cuda):
    """Full solve at n = 8 (P2M/M2M/L2L/L2P with 512 nodes, M2L at r = 256,
    P2P), one vector in the build's precision or sixteen vectors with the
    FP32 MFMA M2L and the MFMA P2P."""
    import torch
    import bbfmmongpu_amd as B
    U, K = ops8
    N = 32768
    pos, Q = O.uniform_inputs(N, seed=5, nrhs=nrhs)
    Q = np.atleast_2d(Q).astype(_dt(prec))
    plan = B.Plan(pos, 8, U, K, nrhs=nrhs, prec=prec, lmod=0).set(scheme=B.SCHEME_IJ, form=form)
    assert plan.maxlev == 3
    phi = plan.solve(torch.from_numpy(Q).to(cuda)).cpu().numpy().reshape(nrhs, N)
    tree = O.Tree(pos, 0, prec)
    dt = _dt(prec)
    tol = 1e-5 if prec == "s" else 1e-12
    for j in check:
        ref = O.fmm(tree, Q[j].astype(dt), 8, U.astype(dt), K.astype(dt), scheme="sibling")
        err = O.rel_l2(ref["phi"], phi[j])
        print(f"n=8 solve prec={prec} form={form} rhs={j}: rel_l2={err:.3e}", flush=True)
        assert err < tol


@pytest.fixture(scope="module")
def real_ops8(cuda):
    """The kernel's own n = 8, r = 256 operators (FP64 precompute with the
    sampling, Gram matrix and E blocks on the GPU)."""
    import bbfmmongpu_amd as B
    U, K, sigma = B.operators(8, R8, prec="d", device=True)
    return U, K, sigma


@pytest.mark.parametrize("prec,form", [("d", 0), ("s", 0), ("s", 2)])
def test_solve_n8_real_operators(prec, form, real_ops8, cuda):
    """n = 8 with the real FP64 operators: the GPU solve against the oracle's
    solve on the same operators (1e-12 fp64, 1e-5 fp32), and the oracle
    against the FP64 direct sum (the accuracy n = 8 buys: below 2e-5 on
    these points, against about 6e-4 at n = 4)."""
    import torch
    import bbfmmongpu_amd as B
    U, K, sigma = real_ops8
    assert sigma[0] > sigma[R8 - 1] > 0
    N = 32768
    pos, q = O.uniform_inputs(N, seed=9)
    dt = _dt(prec)
    q = q.astype(dt)
    plan = B.Plan(pos, 8, U, K, nrhs=1, prec=prec).set(scheme=B.SCHEME_IJ, form=form)
    assert plan.maxlev == 3
    phi = plan.solve(torch.from_numpy(q).to(cuda)).cpu().numpy()
    plan.close()
    tree = O.Tree(pos, 0, prec)
    ref = O.fmm(tree, q, 8, U.astype(dt), K.astype(dt), scheme="sibling")
    err = O.rel_l2(ref["phi"], phi)
    tol = 1e-5 if prec == "s" else 1e-12
    print(f"n=8 real operators prec={prec} form={form}: vs oracle {err:.3e}", flush=True)
    assert err < tol
    if prec == "d":
        tg = np.arange(0, N, 37, dtype=np.int32)
        d = O.direct(pos, q, tg) * q[tg]
        acc = O.rel_l2(d, ref["phi"][tg])
        print(f"n=8 oracle vs FP64 direct sum: {acc:.3e}", flush=True)
        assert acc < 2e-5
