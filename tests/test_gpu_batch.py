"""Many independent problems in one plan (bbfmm_amd_batch_create): every
instance on its own sparse octree of uniform depth, all stages launched
once over all instances.  Each instance's field is checked against the
oracle's complete-octree solve at the same depth (relative L2: 1e-5 fp32,
1e-12 fp64), and instances with automatic depths (leaves of at most 32
points) against the FP64 direct sum."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _instances():
    from bbfmmongpu_amd import workloads as W
    rng = np.random.default_rng(4)
    return [(W.uniform(20000, rng), 3), (W.plummer(8000, rng), 4), (W.sphere_surface(12000, rng), 4),
            (W.uniform(3000, rng), 2), (W.plummer(30000, rng), 5)]


@pytest.mark.parametrize("prec,nrhs", [("s", 1), ("d", 1), ("s", 3)])
def test_batch_against_oracle_per_instance(prec, nrhs, ops4, cuda):
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    inst = _instances()
    pos_list = [p for p, _ in inst]
    depths = [d for _, d in inst]
    Ntot = sum(p.shape[0] for p in pos_list)
    rng = np.random.default_rng(6)
    Q = rng.uniform(-1.0, 1.0, size=(nrhs, Ntot))
    dt = np.float32 if prec == "s" else np.float64
    plan = B.Plan.batch(pos_list, 4, U, K, nrhs=nrhs, depths=depths, prec=prec)
    assert plan.depths == depths
    phi = plan.solve_host(Q.astype(dt)).reshape(nrhs, Ntot)
    tol = 1e-5 if prec == "s" else 1e-12
    o = 0
    for (p, d) in inst:
        n = p.shape[0]
        tree = O.Tree(p, 0, prec, maxlev=d)
        for j in range(nrhs):
            ref = O.fmm(tree, Q[j, o:o + n].astype(dt), 4, U.astype(dt), K.astype(dt), scheme="basic")
            err = O.rel_l2(ref["phi"], phi[j, o:o + n])
            print(f"instance N={n} depth={d} rhs={j}: {err:.2e}", flush=True)
            assert err < tol
        o += n
    plan.close()


def _sparse_instances():
    from bbfmmongpu_amd import workloads as W
    rng = np.random.default_rng(7)
    return [(W.sphere_surface(6000, rng), 5), (W.plummer(4000, rng), 4), (W.sphere_surface(2000, rng), 4),
            (W.uniform(1500, rng), 4)]


def _leaves(inst):
    total = 0
    for p, d in inst:
        c = np.clip(np.floor((p + 0.5) * (1 << d)).astype(np.int64), 0, (1 << d) - 1)
        total += np.unique((c[:, 0] << (2 * d)) + (c[:, 1] << d) + c[:, 2]).size
    return total


@pytest.mark.parametrize("prec,nrhs", [("s", 1), ("d", 1), ("s", 5)])
def test_batch_sparse_leaves_against_oracle(prec, nrhs, ops4, cuda):
    """About 3 points per leaf: P2P runs a lane per target particle (one and
    four vectors at a time); per instance against the oracle."""
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    inst = _sparse_instances()
    pos_list = [p for p, _ in inst]
    depths = [d for _, d in inst]
    Ntot = sum(p.shape[0] for p in pos_list)
    assert Ntot < 12 * _leaves(inst)  # the lane-per-particle P2P is chosen below 12 points per leaf
    Q = np.random.default_rng(8).uniform(-1.0, 1.0, size=(nrhs, Ntot))
    dt = np.float32 if prec == "s" else np.float64
    plan = B.Plan.batch(pos_list, 4, U, K, nrhs=nrhs, depths=depths, prec=prec)
    phi = plan.solve_host(Q.astype(dt)).reshape(nrhs, Ntot)
    tol = 1e-5 if prec == "s" else 1e-12
    o = 0
    for (p, d) in inst:
        n = p.shape[0]
        tree = O.Tree(p, 0, prec, maxlev=d)
        for j in sorted({0, nrhs - 1}):
            ref = O.fmm(tree, Q[j, o:o + n].astype(dt), 4, U.astype(dt), K.astype(dt), scheme="basic")
            err = O.rel_l2(ref["phi"], phi[j, o:o + n])
            print(f"sparse instance N={n} depth={d} rhs={j}: {err:.2e}", flush=True)
            assert err < tol
        o += n
    plan.close()


def test_batch_sparse_fetch_matches_fp64(ops4, cuda):
    """Sparse forests in fp32 run P2M, M2M and post-M2L on MFMA over 32 cells
    per wave; the multipoles and the post-M2L F they fetch match the fp64
    plan's (per-cell kernels) on the same instances."""
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    inst = _sparse_instances()
    pos_list = [p for p, _ in inst]
    depths = [d for _, d in inst]
    Ntot = sum(p.shape[0] for p in pos_list)
    q = np.random.default_rng(9).uniform(-1.0, 1.0, size=Ntot)
    out = {}
    for prec, dt in (("s", np.float32), ("d", np.float64)):
        plan = B.Plan.batch(pos_list, 4, U, K, nrhs=1, depths=depths, prec=prec)
        plan.solve_host(q.astype(dt))
        out[prec] = {w: plan.fetch(w).astype(np.float64) for w in ("M", "L", "F")}
        plan.close()
    for w in ("M", "L", "F"):
        err = O.rel_l2(out["d"][w], out["s"][w])
        print(f"sparse fetch {w}: {err:.2e}", flush=True)
        assert err < 1e-5


def test_batch_auto_depth_against_direct(ops4, cuda):
    import bbfmmongpu_amd as B
    from bbfmmongpu_amd import workloads as W
    U, K, _ = ops4
    pos_list, q_list = W.uneven_instances(count=8, nmin=4096, nmax=32768, seed=3)
    plan = B.Plan.batch(pos_list, 4, U, K, nrhs=1, prec="s")
    q = np.concatenate(q_list).astype(np.float32)
    phi = plan.solve_host(q)
    o = 0
    for i, (p, qi) in enumerate(zip(pos_list, q_list)):
        n = p.shape[0]
        tg = np.sort(np.random.default_rng(i).choice(n, min(n, 1024), replace=False)).astype(np.int32)
        ref = B.direct(p, qi, tg)
        err = O.rel_l2(ref, phi[o + tg])
        print(f"instance {i} N={n} depth={plan.depths[i]}: {err:.2e}", flush=True)
        assert err < 2e-3
        o += n
    assert max(plan.depths) >= 5  # the Plummer cores need deep leaves
    plan.close()


@pytest.mark.parametrize("form,nrhs,tol", [(2, 1, 1e-5), (2, 3, 1e-5), (1, 1, 3e-5), (1, 3, 3e-5), (1, 16, 3e-5)])
def test_batch_list_mfma_against_oracle(form, nrhs, tol, ops4, cuda):
    """The IJ blocking on sparse trees (class-grouped gathered columns on
    MFMA): FP32 MFMA and split BF16, per instance against the oracle."""
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    inst = _instances()
    pos_list = [p for p, _ in inst]
    depths = [d for _, d in inst]
    Ntot = sum(p.shape[0] for p in pos_list)
    Q = np.random.default_rng(16).uniform(-1.0, 1.0, size=(nrhs, Ntot)).astype(np.float32)
    plan = B.Plan.batch(pos_list, 4, U, K, nrhs=nrhs, depths=depths, prec="s").set(form=form, scheme=B.SCHEME_IJ)
    phi = plan.solve_host(Q).reshape(nrhs, Ntot)
    Ud, Kd = U.astype(np.float32).astype(np.float64), K.astype(np.float32).astype(np.float64)
    o = 0
    for (p, d) in inst:
        n = p.shape[0]
        # identical inputs: the plan computes with fp32-rounded positions
        tree = O.Tree(p.astype(np.float32).astype(np.float64), 0, "d", maxlev=d)
        for j in sorted({0, nrhs - 1}):
            ref = O.fmm(tree, Q[j, o:o + n].astype(np.float64), 4, Ud, Kd, scheme="basic")
            err = O.rel_l2(ref["phi"], phi[j, o:o + n])
            print(f"list-MFMA form={form} instance N={n} depth={d} rhs={j}: {err:.2e}", flush=True)
            assert err < tol
        o += n
    plan.close()


@pytest.mark.parametrize("form,switch,tol", [(1, 3, 3e-5), (1, 4, 3e-5), (2, 3, 1e-5)])
def test_batch_list_m2l_level_switch(form, switch, tol, ops4, cuda):
    """Sparse forests with a level switch: the basic scheme (sources by class
    slot, fp32 FMAs) on levels 2 .. switch - 1 and the list MFMA kernel from
    the switch on, per instance against the oracle."""
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    inst = _instances()
    pos_list = [p for p, _ in inst]
    depths = [d for _, d in inst]
    Ntot = sum(p.shape[0] for p in pos_list)
    Q = np.random.default_rng(18).uniform(-1.0, 1.0, size=(1, Ntot)).astype(np.float32)
    plan = B.Plan.batch(pos_list, 4, U, K, nrhs=1, depths=depths, prec="s").set(
        form=form, scheme=B.SCHEME_IJ, level_switch=switch)
    phi = plan.solve_host(Q).reshape(Ntot)
    plan.close()
    Ud, Kd = U.astype(np.float32).astype(np.float64), K.astype(np.float32).astype(np.float64)
    o = 0
    for (p, d) in inst:
        n = p.shape[0]
        tree = O.Tree(p.astype(np.float32).astype(np.float64), 0, "d", maxlev=d)
        ref = O.fmm(tree, Q[0, o:o + n].astype(np.float64), 4, Ud, Kd, scheme="basic")
        err = O.rel_l2(ref["phi"], phi[o:o + n])
        print(f"level switch {switch} form={form} instance N={n} depth={d}: {err:.2e}", flush=True)
        assert err < tol
        o += n
