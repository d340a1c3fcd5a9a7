"""Chebyshev orders n = 3..8 and any SVD rank, as the reference takes them at
run time (main.c:58; r = n^3/2, or n^3/2 - 1 under
DISABLE_INCREASE_CUTOFF_BY_ONE, bbfmm.c:1340-1343).  Ranks other than 32 and
256 run the any-rank M2L kernels (K in the reference's column-major layout);
every order runs its own instantiation of the pass kernels.

* The plan (plain form, all four M2L schemes, fp32 and fp64) against the
  oracle's solve on the same inputs and operators, whose L2P evaluates the
  Chebyshev interpolant (the plan's default).
* The drop-in anotherFMMCompute at n = 3 and 7 against the oracle restating
  the reference's L2P (its second-to-last z weight used twice,
  anotherDownwardPassX.c:304), and m2l_drv at odd ranks.
Tolerances: relative L2 1e-5 (fp32), 1e-12 (fp64)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from conftest import operators_upto

pytestmark = pytest.mark.gpu
TOL = {"s": 1e-5, "d": 1e-12}
CASES = [(3, 13), (3, 12), (4, 31), (5, 62), (5, 61), (6, 108), (6, 107), (7, 171), (7, 170), (8, 255)]
N_PTS = {3: 12000, 4: 12000, 5: 8000, 6: 6000, 7: 4096, 8: 4096}

_ref_cache = {}


def _inputs(n, r, prec):
    dt = np.float32 if prec == "s" else np.float64
    U, K = operators_upto(n, r)
    pos, q = O.uniform_inputs(N_PTS[n], seed=10 + n)
    return pos.astype(dt), q.astype(dt), U.astype(dt), K.astype(dt)


def _oracle(n, r, prec, l2p_exact):
    key = (n, r, prec, l2p_exact)
    if key not in _ref_cache:
        pos, q, U, K = _inputs(n, r, prec)
        T = O.Tree(pos, 1, prec)  # lmod = 1: maxlev 3 or 4, so the dense grids have two levels or more
        _ref_cache[key] = (T, O.fmm(T, q, n, U, K, scheme="sibling", l2p_exact=l2p_exact))
    return _ref_cache[key]


@pytest.mark.parametrize("prec", ["s", "d"])
@pytest.mark.parametrize("n,r", CASES)
def test_plan_orders_all_schemes(n, r, prec, cuda):
    import torch
    import bbfmmongpu_amd as B
    pos, q, U, K = _inputs(n, r, prec)
    T, ref = _oracle(n, r, prec, True)
    plan = B.Plan(pos, n, U, K, nrhs=1, lmod=1, prec=prec)
    assert plan.maxlev == T.maxlev >= 3 and plan.r == r
    qd = torch.from_numpy(q).to(cuda)
    for scheme in (B.SCHEME_BA, B.SCHEME_SI, B.SCHEME_CL, B.SCHEME_IJ):
        plan.set(scheme=scheme, form=B.FORM_PLAIN)
        phi = plan.solve(qd).cpu().numpy()
        err = O.rel_l2(ref["phi"], phi)
        print(f"n={n} r={r} prec={prec} scheme={scheme}: {err:.2e}", flush=True)
        assert err < TOL[prec]
    plan.close()


@pytest.mark.parametrize("n,r", [(5, 62), (7, 170), (8, 255)])
def test_plan_orders_vectors(n, r, cuda):
    """Groups of 4 and 16 vectors share the particle geometry in the pass
    kernels (fp32, IJ scheme), each vector equal to its own oracle solve."""
    import torch
    import bbfmmongpu_amd as B
    pos, _, U, K = _inputs(n, r, "s")
    rng = np.random.default_rng(n)
    T = O.Tree(pos, 1, "s")
    for nrhs in (4, 16):
        Q = rng.uniform(-1, 1, (nrhs, pos.shape[0])).astype(np.float32)
        plan = B.Plan(pos, n, U, K, nrhs=nrhs, lmod=1, prec="s").set(scheme=B.SCHEME_IJ, form=B.FORM_PLAIN)
        phi = plan.solve(torch.from_numpy(Q).to(cuda)).cpu().numpy()
        plan.close()
        for j in (0, nrhs - 1):
            ref = O.fmm(T, Q[j], n, U, K, scheme="sibling", l2p_exact=True)
            err = O.rel_l2(ref["phi"], phi[j])
            print(f"n={n} r={r} nrhs={nrhs} vector {j}: {err:.2e}", flush=True)
            assert err < TOL["s"]


@pytest.mark.parametrize("prec", ["s", "d"])
@pytest.mark.parametrize("n,r", [(3, 13), (7, 171)])
def test_plan_reference_l2p(n, r, prec, cuda):
    """BBFMM_OPT_L2P_REFERENCE: the reference's L2P weights at n % 4 == 3."""
    import torch
    import bbfmmongpu_amd as B
    pos, q, U, K = _inputs(n, r, prec)
    _, ref = _oracle(n, r, prec, False)
    _, exact = _oracle(n, r, prec, True)
    assert O.rel_l2(ref["phi"], exact["phi"]) > 1e-3  # the two L2P forms differ
    plan = B.Plan(pos, n, U, K, nrhs=1, lmod=1, prec=prec).set(scheme=B.SCHEME_IJ, l2p_reference=True)
    phi = plan.solve(torch.from_numpy(q).to(cuda)).cpu().numpy()
    plan.close()
    err = O.rel_l2(ref["phi"], phi)
    print(f"n={n} r={r} prec={prec} reference L2P: {err:.2e}")
    assert err < TOL[prec]


def _tree(lib, pos, n, r, maxlev):
    h = C.c_void_p()
    N = pos.shape[0]
    lib.anotherBuildFMMHierachy(1.0, n, 1, maxlev, r, C.byref(h), N, N, O._p(pos), O._p(pos))
    assert lib.bbfmm_amd_last_error() == 0
    return h


@pytest.mark.parametrize("prec", ["s", "d"])
@pytest.mark.parametrize("n,r", [(3, 13), (5, 62), (7, 171)])
def test_dropin_orders(n, r, prec, cuda):
    """anotherFMMCompute at run-time n and cutoff, schemes BA and IJ: the
    reference's field output, its L2P included."""
    import bbfmmongpu_amd as B
    lib = B.load(prec)
    pos, q, U, K = _inputs(n, r, prec)
    T, ref = _oracle(n, r, prec, False)
    h = _tree(lib, pos, n, r, T.maxlev)
    Tkz, Kt, Kw, Cw = B.weights(n, prec)
    for scheme in (0, 3):
        phi = np.zeros_like(q)
        assert lib.bbfmm_amd_set_default_scheme(scheme) == 0
        lib.anotherFMMCompute(C.byref(h), O._p(pos), O._p(pos), O._p(q), O._p(K), O._p(U), O._p(Tkz), O._p(Kt),
                              O._p(Kw), O._p(Cw), 1.0, r, n, 1, O._p(phi))
        lib.bbfmm_amd_set_default_scheme(3)
        assert lib.bbfmm_amd_last_error() == 0, lib.bbfmm_amd_last_error_string()
        err = O.rel_l2(ref["phi"], phi)
        print(f"anotherFMMCompute n={n} r={r} prec={prec} scheme={scheme}: {err:.2e}")
        assert err < TOL[prec]
    lib.anotherFMMCleanup(C.byref(h))


@pytest.mark.parametrize("prec", ["s", "d"])
@pytest.mark.parametrize("r", [13, 62, 107])
def test_m2l_drv_any_rank(r, prec, cuda):
    """m2l_drv at ranks outside {32, 256}, every scheme, against the oracle's
    list-driven M2L on random multipoles."""
    import bbfmmongpu_amd as B
    dt = np.float32 if prec == "s" else np.float64
    n = {13: 3, 62: 5, 107: 6}[r]
    _, K = operators_upto(n, r)
    K = K.astype(dt)
    pos, _ = O.uniform_inputs(6000, seed=r)
    T = O.Tree(pos.astype(dt), 1, prec)
    rng = np.random.default_rng(r)
    M = rng.standard_normal(T.ncell * r).astype(dt)
    ref = O.m2l(T, K, M, r, "basic")
    s = O.levsta(T.maxlev)[2] * r
    for scheme in (0, 1, 2, 3):
        L = np.zeros(T.ncell * r, dt)
        rc = B.m2l_drv(r, L, K, M, 2, T.maxlev, 1.0, T.center, T.ineigh, T.neighbors, T.iinter,
                       T.interaction, scheme, prec=prec)
        assert rc == 0, B.last_error(prec)
        err = O.rel_l2(ref[s:], L[s:])
        print(f"m2l_drv r={r} prec={prec} scheme={scheme}: {err:.2e}")
        assert err < TOL[prec]
