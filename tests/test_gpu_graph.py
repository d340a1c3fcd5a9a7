"""Whole-solve hipGraph (BBFMM_OPT_GRAPH): every replay, with eager work
and synchronisations in between and the charges changed in place between
replays, matches the eager solve of the same charges (bit for bit: the
graph runs the same kernels on the same data)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _replays_match(plan, Q, cuda):
    import torch
    q = torch.from_numpy(Q).to(cuda).contiguous()
    eager = []
    for scale in (1.0, 2.0, -0.5):
        eager.append(plan.solve((q * scale).contiguous()).cpu().numpy())
    plan.set(graph=True)
    qg = q.clone()
    phi = torch.empty_like(qg)
    for rep in range(2):
        for k, scale in enumerate((1.0, 2.0, -0.5)):
            qg.copy_(q * scale)              # eager work on the graph's input
            plan.solve(qg, phi)               # graph replay
            torch.cuda.synchronize()
            got = phi.cpu().numpy()
            err = O.rel_l2(eager[k], got)
            print(f"replay {rep}/{k}: max abs diff {np.abs(got - eager[k]).max():.3e} rel {err:.2e}", flush=True)
            assert np.array_equal(got, eager[k])
    plan.set(graph=False)


@pytest.mark.parametrize("near_stream,p2p", [(False, 0), (True, 0), (False, 1), (True, 1)])
def test_graph_single_plan_chosen_form(near_stream, p2p, ops4, cuda):
    """near_stream: P2P on a side stream, forked and joined inside the graph;
    p2p 1: the split-BF16 P2P (its charge split kernel inside the graph)."""
    import bbfmmongpu_amd as B
    U, K, _ = ops4
    pos, Q = O.uniform_inputs(65536, seed=8, nrhs=16)
    plan = B.Plan(pos, 4, U, K, nrhs=16, prec="s", lmod=1).set(scheme=B.SCHEME_IJ, form=B.FORM_BF16X2,
                                                               near_stream=near_stream, p2p_form=p2p)
    _replays_match(plan, Q.astype(np.float32), cuda)
    plan.close()


@pytest.mark.parametrize("form", [None, 1])
def test_graph_batch_plan(form, ops4, cuda):
    """A batch of sparse instances (a few points per leaf): the lane P2P and
    the MFMA pass kernels of sparse forests, and with form 1 the split-BF16
    list M2L, replayed from the graph."""
    import bbfmmongpu_amd as B
    from bbfmmongpu_amd import workloads as W
    U, K, _ = ops4
    pos_list, q_list = W.uneven_instances(count=6, nmin=4096, nmax=16384, seed=5)
    plan = B.Plan.batch(pos_list, 4, U, K, nrhs=1, prec="s")
    if form is not None:
        plan.set(scheme=B.SCHEME_IJ, form=form)
    Q = np.concatenate(q_list).astype(np.float32)[None, :]
    _replays_match(plan, Q, cuda)
    plan.close()
