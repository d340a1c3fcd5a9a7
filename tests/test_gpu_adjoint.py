"""Adjoint of the solve (bbfmm_amd_gradient and the torch.autograd.Function):
dLoss/dcharges and dLoss/dpositions for Loss = 1/2 ||phi - phi_t||^2.

Workload: N = 262,144 uniform in [-1/2, 1/2)^3, depth 4, n = 4, r = 32,
8 charge vectors uniform [-1, 1]; phi_t is the FP64 direct-sum potential of
a second charge set (seed 2).  Checks:
  * the fp32 gradients against the analytic gradient of the FP64 direct-sum
    loss (dq = A g, dx_k = sum_v g_vk grad phi_qv(x_k) + q_vk grad phi_gv(x_k)
    with exact fields), to the FMM's own accuracy;
  * the fp64 gradients against central finite differences of the fp64 FMM
    loss (h = 1e-3) on 64 random charges (the loss is quadratic in q: exact
    up to rounding) and on 64 random coordinates kept away from leaf faces,
    so that the tree does not change, of particles with no neighbour closer
    than 10 h; a Richardson pair (h, h/2) gives the O(h^4) check.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N5, M5, DEPTH5 = 262144, 8, 4


@pytest.fixture(scope="module")
def c5(cuda):
    import bbfmmongpu_amd as B
    rng = np.random.default_rng(0)
    pos = rng.random((N5, 3)) - 0.5
    Q = rng.uniform(-1.0, 1.0, size=(M5, N5))
    Q2 = np.random.default_rng(2).uniform(-1.0, 1.0, size=(M5, N5))
    t, _ = B.direct_field(pos, Q2)
    return pos, Q, t


def _rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_autograd_against_direct_sum_gradient(c5, ops4, cuda):
    import torch
    import bbfmmongpu_amd as B
    pos, Q, t = c5
    U, K, _ = ops4
    assert B.maxlev(N5, 0) == DEPTH5
    pos_t = torch.tensor(pos, dtype=torch.float32, device=cuda, requires_grad=True)
    q_t = torch.tensor(Q, dtype=torch.float32, device=cuda, requires_grad=True)
    t_t = torch.tensor(t, dtype=torch.float32, device=cuda)
    phi = B.potential(pos_t, q_t, 4, U, K, prec="s")
    loss = 0.5 * ((phi - t_t) ** 2).sum()
    loss.backward()
    dq = q_t.grad.cpu().numpy()
    dx = pos_t.grad.cpu().numpy()
    # exact loss gradient from the FP64 direct sum
    phi_ex, _ = B.direct_field(pos, Q)
    g_ex = phi_ex - t
    print(f"potential rel err vs direct {_rel(phi.detach().cpu().numpy(), phi_ex):.2e}")
    sample = np.sort(np.random.default_rng(5).choice(N5, 256, replace=False)).astype(np.int32)
    dq_ex, grad_g = B.direct_field(pos, g_ex, sample)          # A g, grad Phi_g
    _, grad_q = B.direct_field(pos, Q, sample)                 # grad Phi_q
    dx_ex = np.einsum("vk,vkc->kc", g_ex[:, sample], grad_q) + np.einsum("vk,vkc->kc", Q[:, sample], grad_g)
    e_dq = _rel(dq[:, sample], dq_ex)
    e_dx = _rel(dx[sample], dx_ex)
    print(f"autograd vs direct-sum gradient: dq {e_dq:.2e} dx {e_dx:.2e}")
    assert e_dq < 3e-3   # measured 3.5e-5
    assert e_dx < 3e-3   # measured 1.6e-3 (the FMM's own accuracy at n = 4)


def test_fp64_gradient_against_finite_differences(c5, ops4, cuda):
    import torch
    import bbfmmongpu_amd as B
    pos, Q, t = c5
    U, K, _ = ops4
    fwd = B.Plan(pos, 4, U, K, nrhs=M5, prec="d").set(scheme=B.SCHEME_IJ, output="potential")
    adj = B.Plan(pos, 4, U, K, nrhs=2 * M5, prec="d").set(scheme=B.SCHEME_IJ)

    def loss(qv):
        phi = fwd.solve_host(qv).reshape(M5, N5)
        return 0.5 * float(((phi - t) ** 2).sum())

    phi = fwd.solve_host(Q).reshape(M5, N5)
    g = phi - t
    dq, dx = adj.gradient(torch.from_numpy(Q).cuda(), torch.from_numpy(np.ascontiguousarray(g)).cuda())
    dq, dx = dq.cpu().numpy(), dx.cpu().numpy()

    rng = np.random.default_rng(9)
    # charges: the loss is quadratic in q, so central differences are exact
    h = 1e-3
    fd_q, an_q = [], []
    for _ in range(64):
        v, k = int(rng.integers(M5)), int(rng.integers(N5))
        Qp, Qm = Q.copy(), Q.copy()
        Qp[v, k] += h
        Qm[v, k] -= h
        fd_q.append((loss(Qp) - loss(Qm)) / (2 * h))
        an_q.append(dq[v, k])
    e_q = _rel(an_q, fd_q)
    print(f"dq vs central differences (64 charges): {e_q:.2e}")
    assert e_q < 1e-7

    # coordinates away from leaf faces (the tree stays the same) of particles
    # whose nearest neighbour is at least 10 h away: the central difference
    # of a pair term at distance r carries an O((h / r)^2) relative error
    leaf = 1.0 / (1 << DEPTH5)
    picks = []
    while len(picks) < 64:
        k, c = int(rng.integers(N5)), int(rng.integers(3))
        f = ((pos[k, c] + 0.5) / leaf) % 1.0
        if min(f, 1 - f) * leaf <= 4 * h:
            continue
        d2 = ((pos - pos[k]) ** 2).sum(axis=1)
        d2[k] = np.inf
        if d2.min() >= (10 * h) ** 2:
            picks.append((k, c))

    def loss_at(k, c, d):
        p = pos.copy()
        p[k, c] += d
        fwd.update_positions(p)
        return loss(Q)

    fd1, fd2, an = [], [], []
    for k, c in picks:
        fd1.append((loss_at(k, c, h) - loss_at(k, c, -h)) / (2 * h))
        fd2.append((loss_at(k, c, h / 2) - loss_at(k, c, -h / 2)) / h)
        an.append(dx[k, c])
    fwd.update_positions(pos)
    fd1, fd2 = np.array(fd1), np.array(fd2)
    rich = (4 * fd2 - fd1) / 3
    e1, er = _rel(an, fd1), _rel(an, rich)
    print(f"dx vs central differences (64 coordinates): h=1e-3 {e1:.2e}, Richardson {er:.2e}")
    assert e1 < 2e-3   # measured 2.0e-4
    assert er < 1e-5   # measured 3.8e-7
    fwd.close()
    adj.close()


def test_gradient_argument_checks(ops4, cuda):
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos = np.random.default_rng(1).random((5000, 3)) - 0.5
    odd = B.Plan(pos, 4, U, K, nrhs=3, prec="s")
    q = torch.zeros((1, 5000), device=cuda)
    with pytest.raises(B.BBFMMError):
        odd.gradient(q, q)
    # a particle moved to another leaf is rejected and the plan keeps working
    p2 = pos.copy()
    p2[0] = -p2[0]
    with pytest.raises(B.BBFMMError, match="another leaf"):
        odd.update_positions(p2)
    odd.close()


def test_autograd_plan_cache(ops4, cuda):
    """potential() keeps its plans: a second call on the same positions
    builds nothing, positions moved inside their leaves are uploaded into the
    cached plans, a particle moved to another leaf makes a new plan; in every
    case the potentials and gradients equal those of fresh plans."""
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    U, K = U.astype(np.float32), K.astype(np.float32)
    N, m = 20000, 8
    rng = np.random.default_rng(11)
    pos = rng.random((N, 3)) - 0.5
    Q = torch.from_numpy(rng.uniform(-1, 1, (m, N)).astype(np.float32)).to(cuda)
    T = torch.from_numpy(rng.uniform(-1, 1, (m, N)).astype(np.float32)).to(cuda)

    def grads(p):
        p_t = torch.tensor(p, dtype=torch.float32, device=cuda, requires_grad=True)
        q_t = Q.clone().requires_grad_(True)
        ph = B.potential(p_t, q_t, 4, U, K, prec="s")
        (0.5 * ((ph - T) ** 2).sum()).backward()
        return ph.detach().cpu().numpy(), q_t.grad.cpu().numpy(), p_t.grad.cpu().numpy()

    def fresh(p):
        B.plan_cache.clear()
        return grads(p)

    B.plan_cache.clear()
    b0 = B.plan_cache.builds
    first = grads(pos)
    assert B.plan_cache.builds == b0 + 2  # forward (m vectors) and backward (2m)
    again = grads(pos)
    assert B.plan_cache.builds == b0 + 2
    for a, b in zip(first, again):
        assert np.array_equal(a, b)
    # inside the leaves (depth 3 here: leaves of 1/8; a shift of 1e-6 keeps
    # every particle in its leaf unless it lies within 1e-6 of a face)
    leaf = 1.0 / 8
    frac = (pos + 0.5) / leaf
    keep = np.all((frac - np.floor(frac) > 1e-4) & (np.ceil(frac) - frac > 1e-4), axis=1)
    moved = pos + np.where(keep[:, None], 1e-6, 0.0)
    b1 = B.plan_cache.builds
    cached = grads(moved)
    assert B.plan_cache.builds == b1
    ref = fresh(moved)
    for a, b in zip(cached, ref):
        assert np.array_equal(a, b)
    # one particle to another leaf: new plans
    across = moved.copy()
    across[0] = -across[0]
    if np.array_equal(np.floor((across[0] + 0.5) / leaf), np.floor((moved[0] + 0.5) / leaf)):
        across[0] = np.where(across[0] > 0, across[0] - leaf, across[0] + leaf)
    grads(moved)
    b2 = B.plan_cache.builds
    crossed = grads(across)
    assert B.plan_cache.builds == b2 + 2
    ref = fresh(across)
    for a, b in zip(crossed, ref):
        assert np.array_equal(a, b)
    B.plan_cache.clear()


def test_autograd_plan_cache_in_place_moves(ops4, cuda):
    """The same position tensor moved in place (an optimiser step): the cache
    sees the new version and moves the positions into its plans; the
    results equal those of fresh plans."""
    import torch
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    U, K = U.astype(np.float32), K.astype(np.float32)
    N, m = 20000, 8
    rng = np.random.default_rng(12)
    pos = rng.random((N, 3)) - 0.5
    Q = torch.from_numpy(rng.uniform(-1, 1, (m, N)).astype(np.float32)).to(cuda)
    T = torch.from_numpy(rng.uniform(-1, 1, (m, N)).astype(np.float32)).to(cuda)
    p_t = torch.tensor(pos, dtype=torch.float32, device=cuda, requires_grad=True)

    def step():
        p_t.grad = None
        q_t = Q.clone().requires_grad_(True)
        ph = B.potential(p_t, q_t, 4, U, K, prec="s")
        (0.5 * ((ph - T) ** 2).sum()).backward()
        return ph.detach().cpu().numpy(), q_t.grad.cpu().numpy(), p_t.grad.cpu().numpy()

    B.plan_cache.clear()
    step()
    builds = B.plan_cache.builds
    leaf = 1.0 / 8
    frac = (pos + 0.5) / leaf
    keep = np.all((frac - np.floor(frac) > 1e-4) & (np.ceil(frac) - frac > 1e-4), axis=1)
    shift = torch.from_numpy(np.where(keep[:, None], 1e-6, 0.0).astype(np.float32)).to(cuda)
    with torch.no_grad():
        p_t.add_(shift)
    moved = step()
    assert B.plan_cache.builds == builds  # uploaded into the cached plans, not rebuilt
    B.plan_cache.clear()
    ref = step()
    for a, b in zip(moved, ref):
        assert np.array_equal(a, b)
    B.plan_cache.clear()
