"""bbfmm_amd_direct_noalloc (the FP64 direct sum without device allocation,
partial sums in the library's file-scope device array) through the C ABI,
against bbfmm_amd_direct on the same inputs: two calls in a row on one GPU,
the second with fewer sources and targets, each bit-identical to the
allocating entry point (same kernels, same summation order)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _noalloc(lib, pos_d, q_d, N, nt, tg_d, out_d):
    lib.bbfmm_amd_direct_noalloc.restype = C.c_int
    lib.bbfmm_amd_direct_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                             C.c_void_p]
    return lib.bbfmm_amd_direct_noalloc(N, pos_d, q_d, nt, tg_d, out_d, None)


def test_direct_noalloc_matches_direct_twice(cuda):
    import torch
    import bbfmmongpu_amd as B
    lib = B.load("d")
    for N, nt, seed in ((30000, 4096, 1), (20000, 1000, 2)):
        pos, q = O.uniform_inputs(N, seed=seed)
        tg = np.random.default_rng(seed).choice(N, nt, replace=False).astype(np.int32)
        want = B.direct(pos, q, targets=tg)
        pos_d = torch.from_numpy(np.ascontiguousarray(pos)).to(cuda)
        q_d = torch.from_numpy(np.ascontiguousarray(q)).to(cuda)
        tg_d = torch.from_numpy(tg).to(cuda)
        out_d = torch.full((nt,), np.nan, dtype=torch.float64, device=cuda)
        torch.cuda.synchronize()
        print(f"direct_noalloc N={N} nt={nt}", flush=True)
        rc = _noalloc(lib, pos_d.data_ptr(), q_d.data_ptr(), N, nt, tg_d.data_ptr(), out_d.data_ptr())
        assert rc == 0, lib.bbfmm_amd_last_error_string()
        torch.cuda.synchronize()
        got = out_d.cpu().numpy()
        assert np.array_equal(got, want), np.max(np.abs(got - want))
        # and against the oracle's FP64 direct sum (different summation order)
        ref = O.direct(pos, q, tg) * q[tg]
        assert O.rel_l2(ref, got) < 1e-13


def test_direct_noalloc_identity_targets(cuda):
    """targets NULL: the first nt particles are the targets."""
    import torch
    import bbfmmongpu_amd as B
    lib = B.load("d")
    N, nt = 5000, 3000
    pos, q = O.uniform_inputs(N, seed=3)
    want = B.direct(pos, q, targets=np.arange(nt, dtype=np.int32))
    pos_d = torch.from_numpy(pos).to(cuda)
    q_d = torch.from_numpy(q).to(cuda)
    out_d = torch.zeros(nt, dtype=torch.float64, device=cuda)
    rc = _noalloc(lib, pos_d.data_ptr(), q_d.data_ptr(), N, nt, None, out_d.data_ptr())
    assert rc == 0, lib.bbfmm_amd_last_error_string()
    torch.cuda.synchronize()
    assert np.array_equal(out_d.cpu().numpy(), want)
