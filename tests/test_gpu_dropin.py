"""The drop-in entry points through the C ABI, as the reference's host code
calls them (anotherFMMCompute, anotherFMMInteraction, anotherDownwardPass,
m2l_drv), against the oracle; multi-vector solves; edge cases of the tree
(empty leaves, particles on the faces of the root cell, tiny N); the FP64
direct sum.  Tolerances: relative L2 1e-5 (fp32), 1e-12 (fp64)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
TOL = {"s": 1e-5, "d": 1e-12}
libc = C.CDLL(None)
libc.malloc.restype = C.c_void_p
libc.malloc.argtypes = [C.c_size_t]
libc.free.argtypes = [C.c_void_p]


def _dt(prec):
    return np.float32 if prec == "s" else np.float64


def _tree(lib, pos, maxlev):
    h = C.c_void_p()
    N = pos.shape[0]
    lib.anotherBuildFMMHierachy(1.0, 4, 1, maxlev, 32, C.byref(h), N, N, O._p(pos), O._p(pos))
    assert lib.bbfmm_amd_last_error() == 0
    return h


def _tree_fields(h, prec):
    from test_abi import _tree_struct
    real = C.c_float if prec == "s" else C.c_double
    return C.cast(h, C.POINTER(_tree_struct(real))).contents


@pytest.mark.parametrize("prec", ["s", "d"])
@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_anotherFMMCompute(prec, scheme, ops4, cuda):
    import bbfmmongpu_amd as B
    lib = B.load(prec)
    dt = _dt(prec)
    U, K, _ = ops4
    Uc, Kc = U.astype(dt), K.astype(dt)
    N = 20000
    pos, q = O.uniform_inputs(N, seed=2)
    pos, q = pos.astype(dt), q.astype(dt)
    T = O.Tree(pos, 0, prec)
    h = _tree(lib, pos, T.maxlev)
    Tkz, Kt, Kw, Cw = B.weights(4, prec)
    phi = np.zeros(N, dt)
    assert lib.bbfmm_amd_set_default_scheme(scheme) == 0
    lib.anotherFMMCompute(C.byref(h), O._p(pos), O._p(pos), O._p(q), O._p(Kc), O._p(Uc), O._p(Tkz), O._p(Kt),
                          O._p(Kw), O._p(Cw), 1.0, 32, 4, 1, O._p(phi))
    assert lib.bbfmm_amd_last_error() == 0, lib.bbfmm_amd_last_error_string()
    lib.bbfmm_amd_set_default_scheme(3)
    ref = O.fmm(T, q, 4, Uc, Kc, scheme="sibling")
    err = O.rel_l2(ref["phi"], phi)
    print(f"anotherFMMCompute prec={prec} scheme={scheme}: {err:.2e}")
    assert err < TOL[prec]
    lib.anotherFMMCleanup(C.byref(h))


@pytest.mark.parametrize("prec", ["s", "d"])
def test_interaction_then_downward(prec, ops4, cuda):
    """anotherFMMInteraction (M -> fieldval) then anotherDownwardPass
    (fieldval -> phi), with M from the oracle's upward pass."""
    import bbfmmongpu_amd as B
    lib = B.load(prec)
    dt = _dt(prec)
    U, K, _ = ops4
    Uc, Kc = U.astype(dt), K.astype(dt)
    N = 20000
    pos, q = O.uniform_inputs(N, seed=4)
    pos, q = pos.astype(dt), q.astype(dt)
    T = O.Tree(pos, 0, prec)
    ref = O.fmm(T, q, 4, Uc, Kc, scheme="sibling")
    h = _tree(lib, pos, T.maxlev)
    t = _tree_fields(h, prec)
    nbytes = ref["M"].nbytes
    p = libc.malloc(nbytes)
    C.memmove(p, ref["M"].ctypes.data, nbytes)
    t.proxysval = p
    Tkz, Kt, Kw, Cw = B.weights(4, prec)
    lib.anotherFMMInteraction(C.byref(h), O._p(Kc), O._p(Kt), O._p(Uc), O._p(Kw), 4, 1, 32, 1.0)
    assert lib.bbfmm_amd_last_error() == 0, lib.bbfmm_amd_last_error_string()
    t = _tree_fields(h, prec)
    assert not t.proxysval
    F = np.ctypeslib.as_array(C.cast(t.fieldval, C.POINTER(C.c_float if prec == "s" else C.c_double)),
                              shape=(T.ncell * 64,)).copy()
    assert O.rel_l2(ref["F"], F) < TOL[prec]
    assert C.c_double.in_dll(lib, "kernel_exec_time").value > 0
    phi = np.zeros(N, dt)
    lib.anotherDownwardPass(C.byref(h), O._p(pos), O._p(pos), O._p(Cw), O._p(Tkz), O._p(q), O._p(Uc), 32, 4, 1,
                            1.0, O._p(phi))
    assert lib.bbfmm_amd_last_error() == 0, lib.bbfmm_amd_last_error_string()
    assert not _tree_fields(h, prec).fieldval
    err = O.rel_l2(ref["phi"], phi)
    print(f"interaction+downward prec={prec}: {err:.2e}")
    assert err < TOL[prec]
    lib.anotherFMMCleanup(C.byref(h))


@pytest.mark.parametrize("prec", ["s", "d"])
@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_m2l_drv(prec, scheme, ops4, cuda):
    import bbfmmongpu_amd as B
    dt = _dt(prec)
    U, K, _ = ops4
    Kc = K.astype(dt)
    pos, _ = O.uniform_inputs(30000, seed=5)
    T = O.Tree(pos.astype(dt), 0, prec)
    rng = np.random.default_rng(3)
    M = rng.standard_normal(T.ncell * 32).astype(dt)
    Lv = np.full(T.ncell * 32, 7.0, dt)  # untouched below minlev
    rc = B.m2l_drv(32, Lv, Kc, M, 2, T.maxlev, 1.0, T.center, T.ineigh, T.neighbors, T.iinter,
                   T.interaction, scheme, prec=prec)
    assert rc == 0, B.last_error(prec)
    ref = O.m2l(O.Tree(pos, 0, "d"), Kc.astype(np.float64), M.astype(np.float64), 32, "basic")
    s = O.levsta(T.maxlev)[2] * 32
    assert np.all(Lv[:s] == 7.0)
    err = O.rel_l2(ref[s:], Lv[s:])
    print(f"m2l_drv prec={prec} scheme={scheme}: {err:.2e}")
    assert err < TOL[prec]


@pytest.mark.parametrize("prec", ["s", "d"])
def test_multi_vector_solve_matches_single(prec, ops4, cuda):
    import torch
    import bbfmmongpu_amd as B
    dt = _dt(prec)
    U, K, _ = ops4
    N = 30000
    pos, Q = O.uniform_inputs(N, seed=7, nrhs=5)
    Q = Q.astype(dt)
    multi = B.Plan(pos, 4, U, K, nrhs=5, prec=prec).set(scheme=3)
    single = B.Plan(pos, 4, U, K, nrhs=1, prec=prec).set(scheme=3)
    qd = torch.from_numpy(Q).cuda()
    pm = multi.solve(qd).cpu().numpy()
    for j in range(5):
        ps = single.solve(qd[j].contiguous()).cpu().numpy()
        assert O.rel_l2(ps, pm[j]) < (1e-6 if prec == "s" else 1e-14)


@pytest.mark.parametrize("case", ["clustered", "faces", "tiny", "deep"])
def test_edge_cases(case, ops4, cuda):
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    rng = np.random.default_rng(12)
    lmod = 0
    if case == "clustered":  # most leaves empty
        N = 20000
        pos = np.clip(0.03 * rng.standard_normal((N, 3)) + 0.2, -0.5, 0.5)
    elif case == "faces":  # particles exactly on the faces and corners of the root cell
        N = 8000
        pos = rng.random((N, 3)) - 0.5
        pos[: N // 4, 0] = 0.5
        pos[N // 4: N // 2, 1] = -0.5
        pos[N // 2: N // 2 + 10] = 0.5
    elif case == "tiny":
        N = 7
        pos = rng.random((N, 3)) - 0.5
    else:  # deeper tree than the level rule (-d 2): few particles per leaf
        N = 20000
        pos = rng.random((N, 3)) - 0.5
        lmod = 2
    q = rng.uniform(-1, 1, N)
    for prec in ("s", "d"):
        dt = _dt(prec)
        T = O.Tree(pos.astype(dt), lmod, prec)
        ref = O.fmm(T, q.astype(dt), 4, U.astype(dt), K.astype(dt), scheme="sibling")
        for scheme, form in ((0, 0), (3, 0)) + (((3, 1),) if prec == "s" else ()):
            plan = B.Plan(pos, 4, U, K, prec=prec, lmod=lmod).set(scheme=scheme, form=form)
            assert plan.maxlev == T.maxlev
            phi = plan.solve_host(q)
            err = O.rel_l2(ref["phi"], phi)
            tol = 1e-4 if form == 1 else TOL[prec]
            print(f"edge {case} prec={prec} scheme={scheme} form={form}: {err:.2e}")
            assert err < tol


def test_direct_sum_against_oracle(cuda):
    import bbfmmongpu_amd as B
    pos, q = O.uniform_inputs(20000, seed=9)
    pos[5] = pos[11]  # coincident pair
    tg = np.random.default_rng(0).choice(20000, 500, replace=False).astype(np.int32)
    got = B.direct(pos, q, tg)
    ref = O.direct(pos, q, tg) * q[tg]
    assert O.rel_l2(ref, got) < 1e-13
    full = B.direct(pos[:3000], q[:3000])
    assert O.rel_l2(O.direct(pos[:3000], q[:3000]) * q[:3000], full) < 1e-13
