"""Pointer-form loads (BBFMM_OPT_ADDRESSING = 1).  The split-BF16 M2L
kernels (source-stationary dense-grid, list) and the fp32 16-vector P2P read
through scalar-offset buffer loads while their arrays are below 2 GB and
through pointers beyond; the option forces the pointer form on small
problems so that both forms are run.  Each must give the same bits as the
buffer form (only the addressing differs) and stay within the form's bound
of the oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [1, 4, 5, 6, 7, 8])
def test_source_stationary_m2l_pointer_form(variant, ops4, cuda):
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    nrhs = 16
    pos, _ = O.uniform_inputs(65536, seed=0)
    tree = O.Tree(pos, 0, "d")
    rng = np.random.default_rng(21)
    M = rng.standard_normal((nrhs, tree.ncell * 32)).astype(np.float32)
    Md = torch.from_numpy(M).to(cuda)
    out = {}
    for flat in (False, True):
        plan = B.Plan(pos, 4, U, K, prec="s", nrhs=nrhs).set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2,
                                                           variant=variant, flat_loads=flat)
        Ld = torch.zeros_like(Md)
        plan.m2l(Md, Ld)
        torch.cuda.synchronize()
        out[flat] = Ld.cpu().numpy()
        plan.close()
    assert np.array_equal(out[False], out[True])
    s = O.levsta(tree.maxlev)[2] * 32
    Kd = K.astype(np.float32).astype(np.float64)
    for j in (0, nrhs - 1):
        ref = O.m2l(tree, Kd, M[j].astype(np.float64), 32, "basic")
        err = O.rel_l2(ref[s:], out[True][j, s:])
        print(f"pointer-form source-stationary M2L variant={variant} rhs={j}: {err:.2e}", flush=True)
        assert err < 3e-5


def test_p2p_seg16_pointer_form(ops4, cuda):
    """The whole 16-vector solve (fp32 P2P) with both forms: the same sums
    (the two instantiations may contract the 1/r geometry into FMAs
    differently, so the bound is fp32 round-off, not equality), and each
    within the split-BF16 M2L's bound of the oracle."""
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, Q = O.uniform_inputs(65536, seed=5, nrhs=16)
    Q = Q.astype(np.float32)
    qd = torch.from_numpy(Q).to(cuda)
    out = {}
    for flat in (False, True):
        plan = B.Plan(pos, 4, U, K, prec="s", nrhs=16).set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2,
                                                         p2p_form=B.P2P_FP32, flat_loads=flat)
        out[flat] = plan.solve(qd).cpu().numpy()
        plan.close()
    diff = np.abs(out[False].astype(np.float64) - out[True])
    worst = int(np.argmax(diff))
    print(f"pointer vs buffer form: {np.count_nonzero(diff)} of {diff.size} values differ, max |d| {diff.max():.3e} "
          f"at {np.unravel_index(worst, diff.shape)}, rel L2 {O.rel_l2(out[False], out[True]):.3e}", flush=True)
    assert O.rel_l2(out[False], out[True]) < 1e-6
    tree = O.Tree(pos, 0, "s")
    Uf, Kf = U.astype(np.float32), K.astype(np.float32)
    for j in (0, 15):
        ref = O.fmm(tree, Q[j], 4, Uf, Kf, scheme="sibling")
        for flat in (False, True):
            err = O.rel_l2(ref["phi"], out[flat][j])
            print(f"16-vector solve flat={flat} rhs={j}: {err:.2e}", flush=True)
            assert err < 3e-5


def test_list_m2l_pointer_form(ops4, cuda):
    """Sparse forests: the split-BF16 list M2L with both forms."""
    import bbfmmongpu_amd as B
    from bbfmmongpu_amd import workloads as W
    U, K, _ = ops4
    rng = np.random.default_rng(4)
    inst = [(W.uniform(20000, rng), 3), (W.plummer(8000, rng), 4), (W.sphere_surface(12000, rng), 4)]
    pos_list = [p for p, _ in inst]
    depths = [d for _, d in inst]
    Ntot = sum(p.shape[0] for p in pos_list)
    Q = np.random.default_rng(6).uniform(-1.0, 1.0, size=(1, Ntot)).astype(np.float32)
    out = {}
    for flat in (False, True):
        plan = B.Plan.batch(pos_list, 4, U, K, nrhs=1, depths=depths, prec="s").set(
            form=B.FORM_BF16X2, scheme=B.SCHEME_IJ, flat_loads=flat)
        out[flat] = plan.solve_host(Q).reshape(Ntot)
        plan.close()
    assert np.array_equal(out[False], out[True])
    o = 0
    Uf, Kf = U.astype(np.float32), K.astype(np.float32)
    for (p, d) in inst:
        n = p.shape[0]
        tree = O.Tree(p, 0, "s", maxlev=d)
        ref = O.fmm(tree, Q[0, o:o + n], 4, Uf, Kf, scheme="basic")
        err = O.rel_l2(ref["phi"], out[True][o:o + n])
        print(f"pointer-form list M2L instance N={n} depth={d}: {err:.2e}", flush=True)
        assert err < 3e-5
        o += n
