"""The batch and high-order workloads at their full sizes.

* 64 uneven instances (3.9 M points, N log-uniform in [4096, 262144], half
  Plummer, half sphere surface, leaves of at most 32 points, n = 4, r = 32,
  split-BF16 list M2L, one hipGraph): every instance against the FP64
  direct sum at 256 sampled targets (relative L2 2e-4; measured 8.5e-5 at
  most), the four smallest instances against the oracle's FP64 solve on
  the same inputs and operators (3e-5), and graph replays equal to the
  eager solve.
* N = 4,194,304 uniform, depth 6, n = 8, r = 256, 16 vectors, M2L on FP32
  MFMA (the real FP64 operators, cast to fp32): vectors 0 and 15 against
  the FP64 direct sum at 2048 targets (2e-5; measured 3.5e-6) and vector 0
  against the plan's own FP64 plain solve on the FP64 operators (3e-5;
  measured 1.0e-5).

The direct sums and the FP64 plain solve are pinned to the oracle at
smaller sizes in the other test files; here they carry the check to the
sizes the oracle cannot run in seconds."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_c4_full_64_uneven_instances(ops4, cuda):
    import torch
    import bbfmmongpu_amd as B
    from bbfmmongpu_amd import workloads as W
    U, K, _ = ops4
    Uf, Kf = U.astype(np.float32), K.astype(np.float32)
    pos_list, q_list = W.uneven_instances(count=64, nmin=4096, nmax=262144, seed=0)
    Ns = [p.shape[0] for p in pos_list]
    assert len(Ns) == 64 and sum(Ns) > 3_500_000
    plan = B.Plan.batch(pos_list, 4, U, K, nrhs=1, prec="s").set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2)
    depths = plan.depths
    print(f"C4 full: N total {sum(Ns)}, depths {min(depths)}..{max(depths)}, cells {plan.ncell}", flush=True)
    q = torch.from_numpy(np.concatenate(q_list).astype(np.float32)).cuda()[None, :].contiguous()
    phi = torch.empty_like(q)
    plan.solve(q, phi)
    eager = phi.cpu().numpy()[0].copy()
    plan.set(graph=True)
    for rep in range(3):
        plan.solve(q, phi)
        torch.cuda.synchronize()
        assert np.array_equal(phi.cpu().numpy()[0], eager), f"graph replay {rep}"
    plan.close()
    offs = np.concatenate([[0], np.cumsum(Ns)])
    worst = 0.0
    for i, (p, qi) in enumerate(zip(pos_list, q_list)):
        n = p.shape[0]
        tg = np.sort(np.random.default_rng(100 + i).choice(n, min(n, 256), replace=False)).astype(np.int32)
        err = O.rel_l2(B.direct(p, qi, tg), eager[offs[i] + tg])
        worst = max(worst, err)
        assert err < 2e-4, f"instance {i} (N={n}, depth {depths[i]}): {err:.3e}"
    print(f"C4 full: worst instance vs FP64 direct {worst:.3e}", flush=True)
    for i in np.argsort(Ns)[:4]:
        p, qi = pos_list[i], q_list[i].astype(np.float32)
        tree = O.Tree(p, 0, "s", maxlev=depths[i])
        ref = O.fmm(tree, qi, 4, Uf, Kf, scheme="basic")
        err = O.rel_l2(ref["phi"], eager[offs[i]:offs[i + 1]])
        print(f"C4 full: instance {i} (N={Ns[i]}, depth {depths[i]}) vs oracle {err:.2e}", flush=True)
        assert err < 3e-5


def test_c3_full_n8_16_vectors(cuda):
    import torch
    import bbfmmongpu_amd as B
    N, nrhs = 1 << 22, 16
    rng = np.random.default_rng(0)
    pos = rng.random((N, 3)) - 0.5
    Q = rng.uniform(-1.0, 1.0, size=(nrhs, N))
    U64, K64, _ = B.operators(8, 256, prec="d", device=True)
    plan = B.Plan(pos, 8, U64.astype(np.float32), K64.astype(np.float32), nrhs=nrhs, prec="s").set(
        scheme=B.SCHEME_IJ, form=B.FORM_FP32MFMA)
    assert plan.maxlev == 6
    phi = plan.solve(torch.from_numpy(Q.astype(np.float32)).cuda()).cpu().numpy()
    plan.close()
    tg = np.sort(np.random.default_rng(1).choice(N, size=2048, replace=False)).astype(np.int32)
    for j in (0, nrhs - 1):
        err = O.rel_l2(B.direct(pos, Q[j], tg), phi[j, tg])
        print(f"C3 full: vector {j} vs FP64 direct {err:.3e}", flush=True)
        assert err < 2e-5
    p64 = B.Plan(pos, 8, U64, K64, nrhs=1, prec="d").set(scheme=B.SCHEME_IJ, form=B.FORM_PLAIN)
    phi64 = p64.solve(torch.from_numpy(Q[0]).cuda()).cpu().numpy()
    p64.close()
    err = O.rel_l2(phi64, phi[0])
    print(f"C3 full: vector 0 vs the FP64 plain solve {err:.3e}", flush=True)
    assert err < 3e-5
