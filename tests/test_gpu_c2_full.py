"""The precision workload at its full size: N = 1,048,576 uniform points,
depth 5 (about 32 per leaf), n = 4, r = 32, 16 charge vectors per solve,
split-BF16 M2L (the source-stationary MFMA kernel) with each P2P form, the
particle-row form and its fused first and last levels.  Vectors 0, 7 and 15
are compared with the oracle's FP64 solve (sibling M2L, CPU8A near field)
on the same operators: the split-BF16 products and fp32 accumulation stay
within 3e-5 relative L2 of it."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

N_FULL = 1 << 20
TOL = 3e-5


@pytest.fixture(scope="module")
def c2_reference(ops4):
    U, K, _ = ops4
    pos, Q = O.uniform_inputs(N_FULL, seed=0, nrhs=16)
    Q = Q.astype(np.float32)
    tree = O.Tree(pos, 0, "d")
    assert tree.maxlev == 5
    # the oracle on the GPU path's fp32 inputs and operators, in FP64
    Kd, Ud = K.astype(np.float32).astype(np.float64), U.astype(np.float32).astype(np.float64)
    ref = {}
    for j in (0, 7, 15):
        ref[j] = O.fmm(tree, Q[j].astype(np.float64), 4, Ud, Kd, scheme="sibling")["phi"]
        print(f"oracle fp64 solve of vector {j} done", flush=True)
    return pos, Q, ref


@pytest.mark.parametrize("p2p", [0, 1])
def test_c2_full_size_16_vectors(p2p, c2_reference, ops4, cuda):
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, Q, ref = c2_reference
    plan = B.Plan(pos, 4, U, K, nrhs=16, prec="s").set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2, p2p_form=p2p)
    assert plan.maxlev == 5
    q = torch.from_numpy(Q).to(cuda)
    phi = plan.solve(q).cpu().numpy()
    again = plan.solve(q).cpu().numpy()
    plan.close()
    assert np.array_equal(phi, again)  # deterministic
    for j, r in ref.items():
        err = O.rel_l2(r, phi[j])
        print(f"C2 full size p2p_form={p2p} vector {j}: rel L2 vs oracle fp64 = {err:.3e}", flush=True)
        assert err < TOL
