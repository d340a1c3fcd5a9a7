"""M2L configurations and the autotuner (bbfmm_amd_plan_tune): every
scheme / tile variant / level switch the tuner can pick computes the same
L as the oracle (relative L2 1e-5, split-BF16 3e-5), and a tuned plan
solves as before."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# (form, scheme, variant, level switch, tolerance)
CONFIGS = [
    (0, 0, 0, 2, 1e-5), (0, 0, 1, 2, 1e-5), (0, 0, 2, 2, 1e-5), (0, 1, 0, 2, 1e-5),
    (0, 3, 0, 2, 1e-5), (0, 3, 0, 3, 1e-5), (0, 2, 0, 3, 1e-5),
    (1, 3, 1, 2, 3e-5), (1, 3, 2, 2, 3e-5), (1, 3, 3, 2, 3e-5), (1, 3, 1, 3, 3e-5), (1, 3, 3, 4, 3e-5),
    (1, 3, 4, 2, 3e-5), (1, 3, 5, 2, 3e-5), (1, 3, 4, 3, 3e-5), (1, 3, 5, 4, 3e-5),
    (1, 3, 6, 2, 3e-5), (1, 3, 7, 2, 3e-5), (1, 3, 6, 3, 3e-5), (1, 3, 7, 4, 3e-5),
    (2, 3, 2, 2, 1e-5), (2, 3, 3, 3, 1e-5),
    (3, 3, 2, 2, 1e-6), (3, 3, 3, 3, 1e-6), (3, 3, 0, 2, 1e-6),
]


# source-stationary class groups with a vector count that leaves partial column tiles
RAGGED = [(1, 3, 1, 2, 3e-5, 3), (1, 3, 4, 2, 3e-5, 3), (1, 3, 5, 2, 3e-5, 5), (1, 3, 6, 2, 3e-5, 3),
          (1, 3, 7, 2, 3e-5, 5)]


@pytest.mark.parametrize("form,scheme,variant,switch,tol,nrhs",
                         [c + (16 if c[0] in (1, 3) else 2,) for c in CONFIGS] + RAGGED)
def test_m2l_configuration_against_oracle(form, scheme, variant, switch, tol, nrhs, ops4, cuda):
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, _q = O.uniform_inputs(65536, seed=0)
    plan = B.Plan(pos, 4, U, K, prec="s", nrhs=nrhs).set(scheme=scheme, form=form, variant=variant,
                                                         level_switch=switch)
    tree = O.Tree(pos, 0, "d")
    rng = np.random.default_rng(13)
    M = rng.standard_normal((nrhs, tree.ncell * 32)).astype(np.float32)
    Md = torch.from_numpy(M).to(cuda)
    Ld = torch.zeros_like(Md)
    plan.m2l(Md, Ld)
    torch.cuda.synchronize()
    Lg = Ld.cpu().numpy()
    s = O.levsta(4)[2] * 32
    Kd = K.astype(np.float32).astype(np.float64)
    for j in (0, nrhs - 1):
        ref = O.m2l(tree, Kd, M[j].astype(np.float64), 32, "basic")
        err = O.rel_l2(ref[s:], Lg[j, s:])
        print(f"form={form} scheme={scheme} variant={variant} switch={switch} nrhs={nrhs} rhs={j}: {err:.2e}",
              flush=True)
        assert err < tol
    plan.close()


@pytest.mark.parametrize("variant,upper,split,nrhs", [(7, 4, 4, 16), (4, 4, 2, 16), (4, 5, 3, 3), (1, 4, 16, 16),
                                                       (6, 7, 5, 16), (8, 4, 4, 16), (8, 5, 2, 3), (8, 1, 7, 16)])
def test_m2l_upper_levels_in_step_parts(variant, upper, split, nrhs, ops4, cuda):
    """BBFMM_OPT_M2L_UPPER_SPLIT: the launch of the levels above the deepest
    takes each class group's steps in `split` parts whose partial sums are
    added afterwards.  The deepest level keeps the bits of the single launch
    of its variant; every level matches the oracle's M2L; the upper levels
    match the unsplit upper launch to fp32 rounding (the sum order differs)."""
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, _q = O.uniform_inputs(65536, seed=0)
    tree = O.Tree(pos, 0, "d")
    rng = np.random.default_rng(23)
    M = rng.standard_normal((nrhs, tree.ncell * 32)).astype(np.float32)
    Md = torch.from_numpy(M).to(cuda)
    out = {}
    for up, sp in ((0, 1), (upper, 1), (upper, split)):
        plan = B.Plan(pos, 4, U, K, prec="s", nrhs=nrhs).set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2,
                                                             variant=variant, variant_upper=up, upper_split=sp)
        Ld = torch.zeros_like(Md)
        plan.m2l(Md, Ld)
        torch.cuda.synchronize()
        out[(up, sp)] = Ld.cpu().numpy()
        plan.close()
    ls = O.levsta(tree.maxlev)
    deep, s = ls[tree.maxlev] * 32, ls[2] * 32
    got = out[(upper, split)]
    assert np.array_equal(out[(0, 1)][:, deep:], got[:, deep:])
    d = O.rel_l2(out[(upper, 1)][:, s:deep].astype(np.float64), got[:, s:deep])
    print(f"variant={variant} upper={upper} split={split} nrhs={nrhs}: upper levels vs unsplit {d:.2e}", flush=True)
    assert d < 1e-6
    Kd = K.astype(np.float32).astype(np.float64)
    for j in (0, nrhs - 1):
        ref = O.m2l(tree, Kd, M[j].astype(np.float64), 32, "basic")
        err = O.rel_l2(ref[s:], got[j, s:])
        print(f"variant={variant} upper={upper} split={split} nrhs={nrhs} rhs={j}: {err:.2e}", flush=True)
        assert err < 3e-5


@pytest.mark.parametrize("variant,split,nrhs", [(4, 3, 16), (0, 1, 16), (5, 2, 3), (7, 4, 16)])
def test_m2l_whole_launch_in_step_parts(variant, split, nrhs, ops4, cuda):
    """BBFMM_OPT_M2L_UPPER_SPLIT without an upper-level variant: the single
    launch takes each class group's steps in `split` parts, their partial
    sums added afterwards; variant 0 does so by itself on deepest levels of
    8 K to 64 K columns (here 8192: 5 parts).  L matches the unsplit launch
    of the same tile shape to fp32 rounding and the oracle's M2L."""
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, _q = O.uniform_inputs(65536, seed=0)
    tree = O.Tree(pos, 0, "d")
    assert tree.maxlev == 4
    rng = np.random.default_rng(29)
    M = rng.standard_normal((nrhs, tree.ncell * 32)).astype(np.float32)
    Md = torch.from_numpy(M).to(cuda)
    out = {}
    one = variant if variant else 4
    for v, sp in ((one, 1), (variant, split)):
        plan = B.Plan(pos, 4, U, K, prec="s", nrhs=nrhs).set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2, variant=v,
                                                             upper_split=sp)
        Ld = torch.zeros_like(Md)
        plan.m2l(Md, Ld)
        torch.cuda.synchronize()
        out[(v, sp)] = Ld.cpu().numpy()
        plan.close()
    s = O.levsta(tree.maxlev)[2] * 32
    got = out[(variant, split)]
    d = O.rel_l2(out[(one, 1)][:, s:].astype(np.float64), got[:, s:])
    print(f"variant={variant} split={split} nrhs={nrhs}: vs unsplit {d:.2e}", flush=True)
    assert 0 < d < 1e-6 if (variant, split) != (one, 1) else d == 0
    Kd = K.astype(np.float32).astype(np.float64)
    for j in (0, nrhs - 1):
        ref = O.m2l(tree, Kd, M[j].astype(np.float64), 32, "basic")
        err = O.rel_l2(ref[s:], got[j, s:])
        assert err < 3e-5


def test_m2l_step_parts_option_range(ops4, cuda):
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, _q = O.uniform_inputs(4096, seed=0)
    plan = B.Plan(pos, 4, U, K, prec="s", nrhs=16).set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2, variant=4)
    for bad in (0, 17):
        with pytest.raises(B.BBFMMError):
            plan.set(upper_split=bad)
    plan.close()
    plan = B.Plan(pos, 4, U, K, prec="s", nrhs=16).set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2, variant=2,
                                                       upper_split=2)
    with pytest.raises(B.BBFMMError):  # step parts need the source-stationary kernel
        plan.solve_host(np.ones((16, 4096), np.float32))
    plan.close()


@pytest.mark.parametrize("variant,upper,nrhs", [(1, 5, 16), (1, 4, 16), (4, 5, 16), (5, 1, 16), (1, 5, 3), (6, 7, 16),
                                                 (1, 7, 16)])
def test_m2l_upper_levels_own_variant(variant, upper, nrhs, ops4, cuda):
    """Split-BF16 source-stationary M2L with the levels above the deepest in a
    launch of their own (BBFMM_OPT_M2L_VARIANT_UPPER): the same L as the
    oracle, and the same bits as the single launch of the deepest level's
    variant on that level."""
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, _q = O.uniform_inputs(65536, seed=0)
    tree = O.Tree(pos, 0, "d")
    rng = np.random.default_rng(17)
    M = rng.standard_normal((nrhs, tree.ncell * 32)).astype(np.float32)
    Md = torch.from_numpy(M).to(cuda)
    out = {}
    for up in (0, upper):
        plan = B.Plan(pos, 4, U, K, prec="s", nrhs=nrhs).set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2,
                                                             variant=variant, variant_upper=up)
        Ld = torch.zeros_like(Md)
        plan.m2l(Md, Ld)
        torch.cuda.synchronize()
        out[up] = Ld.cpu().numpy()
        plan.close()
    ls = O.levsta(tree.maxlev)
    deep = ls[tree.maxlev] * 32  # the deepest level's cells: same kernel, same launch shape
    assert np.array_equal(out[0][:, deep:], out[upper][:, deep:])
    s = ls[2] * 32
    Kd = K.astype(np.float32).astype(np.float64)
    for j in (0, nrhs - 1):
        ref = O.m2l(tree, Kd, M[j].astype(np.float64), 32, "basic")
        err = O.rel_l2(ref[s:], out[upper][j, s:])
        print(f"variant={variant} upper={upper} nrhs={nrhs} rhs={j}: {err:.2e}", flush=True)
        assert err < 3e-5


def test_fp16x3_scales_each_level(ops4, cuda):
    """The split-FP16 form scales K and every level's M by powers of two:
    multipoles whose size differs by 1e10 between levels (beyond fp16's
    range) keep fp32-level accuracy on every level."""
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, _q = O.uniform_inputs(65536, seed=0)
    plan = B.Plan(pos, 4, U, K, prec="s", nrhs=16).set(scheme=B.SCHEME_IJ, form=B.FORM_FP16X3)
    tree = O.Tree(pos, 0, "d")
    ls = O.levsta(5)  # first cell of levels 0..5
    rng = np.random.default_rng(17)
    M = rng.standard_normal((16, tree.ncell * 32))
    for lev, f in ((2, 1e5), (3, 1.0), (4, 1e-5)):
        M[:, ls[lev] * 32:ls[lev + 1] * 32] *= f
    M = M.astype(np.float32)
    Ld = torch.zeros((16, tree.ncell * 32), device=cuda)
    plan.m2l(torch.from_numpy(M).to(cuda), Ld)
    Lg = Ld.cpu().numpy()
    Kd = K.astype(np.float32).astype(np.float64)
    ref = O.m2l(tree, Kd, M[5].astype(np.float64), 32, "basic")
    for lev in (2, 3, 4):
        sl = slice(ls[lev] * 32, ls[lev + 1] * 32)
        err = O.rel_l2(ref[sl], Lg[5, sl])
        print(f"level {lev}: rel_l2 {err:.2e}", flush=True)
        assert err < 1e-6
    plan.close()


@pytest.mark.parametrize("form,nrhs", [(0, 1), (1, 16), (3, 16)])
def test_tuned_plan_solves_like_oracle(form, nrhs, ops4, cuda):
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, Q = O.uniform_inputs(65536, seed=2, nrhs=nrhs)
    Q = np.atleast_2d(Q).astype(np.float32)
    plan = B.Plan(pos, 4, U, K, prec="s", nrhs=nrhs).set(scheme=B.SCHEME_IJ, form=form)
    plan.solve(torch.from_numpy(Q).to(cuda))
    table = plan.tune(reps=2)
    print("tuning table:", [(int(a), int(b), int(u), int(c), round(d, 4)) for a, b, u, c, d in table], flush=True)
    assert len(table) >= 2
    best = min(table, key=lambda t: t[4])
    assert all(t[4] > 0 for t in table)
    phi = plan.solve(torch.from_numpy(Q).to(cuda)).cpu().numpy().reshape(nrhs, -1)
    tree = O.Tree(pos, 0, "d")
    ref = O.fmm(tree, Q[0].astype(np.float64), 4, U.astype(np.float32).astype(np.float64),
                K.astype(np.float32).astype(np.float64), scheme="sibling")
    err = O.rel_l2(ref["phi"], phi[0])
    print(f"tuned (best {best}) solve rel_l2 {err:.2e}", flush=True)
    assert err < (3e-5 if form == 1 else 1e-5)  # form 3 (FP16 3-term): fp32 level
    plan.close()


def test_tuned_batch_plan(ops4, cuda):
    import bbfmmongpu_amd as B
    from bbfmmongpu_amd import workloads as W
    U, K, _ = ops4
    pos_list, q_list = W.uneven_instances(count=6, nmin=4096, nmax=16384, seed=7)
    plan = B.Plan.batch(pos_list, 4, U, K, nrhs=1, prec="s")
    q = np.concatenate(q_list).astype(np.float32)
    phi0 = plan.solve_host(q)
    table = plan.tune(reps=2)
    assert sorted(int(t[1]) for t in table) == [0, 1, 2] and all(int(t[0]) == 0 for t in table)
    phi1 = plan.solve_host(q)
    assert O.rel_l2(phi0, phi1) < 1e-6
    plan.close()


@pytest.mark.parametrize("two,one,nrhs", [(6, 5, 16), (7, 4, 16), (6, 5, 3), (7, 4, 5), (8, 4, 16), (8, 7, 5)])
def test_m2l_two_tiles_per_wave_same_bits(two, one, nrhs, ops4, cuda):
    """Variants 6 and 7 give every wave of the source-stationary kernel two
    32-column tiles (2 and 4 field classes); each tile's products and their
    order are those of variants 5 and 4 (one tile per wave), so L is the same
    to the bit, also with vector counts that leave partial tiles.  Variant 8
    (both class groups in one block, sharing the source tiles through LDS)
    walks the union of the groups' step lists in the same order, the extra
    steps with zero operators, so it too gives the bits of variant 4."""
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, _q = O.uniform_inputs(65536, seed=3)
    rng = np.random.default_rng(17)
    out = {}
    for v in (two, one):
        plan = B.Plan(pos, 4, U, K, prec="s", nrhs=nrhs).set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2, variant=v)
        if v == two:
            M = rng.standard_normal((nrhs, plan.ncell * 32)).astype(np.float32)
            Md = torch.from_numpy(M).to(cuda)
        Ld = torch.zeros_like(Md)
        plan.m2l(Md, Ld)
        torch.cuda.synchronize()
        out[v] = Ld.cpu().numpy()
        plan.close()
    assert np.abs(out[one]).max() > 0
    assert np.array_equal(out[two], out[one])
