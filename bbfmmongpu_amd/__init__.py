"""bbfmmongpu_amd -- MI355X (gfx950) black-box FMM downward pass.

Python mirror of the library's C ABI (include/bbfmm_amd.h).  The compute
path is the HIP library libbbfmm_amd_{s,d}.so built in-tree; importing this
package does not build it, and every entry point raises if it is missing.
Device buffers are torch tensors on ``cuda:0`` (plumbing only); host buffers
are numpy arrays in the reference's layouts.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._build import lib_path

__all__ = [
    "load", "Plan", "operators", "write_operators", "read_operators", "direct", "direct_field",
    "FMMPotential", "potential", "weights", "maxlev", "m2l_drv", "BBFMMError",
    "SCHEME_BA", "SCHEME_SI", "SCHEME_CL", "SCHEME_IJ", "FORM_PLAIN", "FORM_BF16X2", "FORM_FP32MFMA", "FORM_FP16X3",
]

SCHEME_BA, SCHEME_SI, SCHEME_CL, SCHEME_IJ = 0, 1, 2, 3
FORM_PLAIN, FORM_BF16X2, FORM_FP32MFMA, FORM_FP16X3 = 0, 1, 2, 3
OPT_SCHEME, OPT_FORM, OPT_TIMING, OPT_OUTPUT, OPT_GRAPH, OPT_M2L_VARIANT, OPT_LEVEL_SWITCH = 1, 2, 3, 4, 5, 6, 7
OPT_NEAR_STREAM = 8
OPT_P2P_FORM = 9
OPT_L2P_REFERENCE = 10
OPT_ADDRESSING = 11
OPT_M2L_VARIANT_UPPER = 12
OPT_M2L_UPPER_SPLIT = 13
P2P_FP32, P2P_BF16X2 = 0, 1
# "p2p_beside": P2P's own time when it runs on the side stream beside the far
# field ("p2p" is then the wait for it after the far field)
STAGES = ["gather", "p2m", "m2m", "m2l_prep", "m2l", "post_m2l", "l2l", "l2p", "p2p", "scatter", "total",
          "p2p_beside"]

_libs: dict = {}


class BBFMMError(RuntimeError):
    pass


def dtype(prec: str):
    return np.float32 if prec == "s" else np.float64


def load(prec: str = "s"):
    """The ctypes handle of libbbfmm_amd_{prec}.so (raises if not built)."""
    if prec in _libs:
        return _libs[prec]
    path = lib_path(prec)
    if not os.path.exists(path):
        raise BBFMMError(f"{path} is not built; run __graft_entry__.build() first")
    L = C.CDLL(path)
    R = C.c_float if prec == "s" else C.c_double
    vp, ip, i = C.c_void_p, C.c_int, C.c_int
    L.bbfmm_amd_plan_create.argtypes = [C.POINTER(vp), i, vp, R, i, i, i, vp, vp, i, vp]
    L.bbfmm_amd_plan_destroy.argtypes = [vp]
    L.bbfmm_amd_plan_set.argtypes = [vp, i, i]
    L.bbfmm_amd_plan_info.argtypes = [vp, vp, vp]
    L.bbfmm_amd_solve.argtypes = [vp, vp, vp]
    L.bbfmm_amd_solve_host.argtypes = [vp, vp, vp]
    L.bbfmm_amd_m2l.argtypes = [vp, vp, vp]
    L.bbfmm_amd_fetch.argtypes = [vp, i, vp]
    L.bbfmm_amd_stage_times.argtypes = [vp, vp, i]
    L.bbfmm_amd_maxlev.argtypes = [i, i]
    L.bbfmm_amd_weights.argtypes = [i, vp, vp, vp, vp]
    L.bbfmm_amd_weights.restype = None
    L.bbfmm_amd_operators.argtypes = [i, i, vp, vp, vp]
    L.bbfmm_amd_operators_device.argtypes = [i, i, vp, vp, vp]
    cp = C.c_char_p
    L.bbfmm_amd_write_operators.argtypes = [cp, cp, i, i, vp, vp]
    L.bbfmm_amd_read_operators.argtypes = [cp, cp, i, vp, vp, vp]
    L.bbfmm_amd_last_error_string.restype = C.c_char_p
    L.bbfmm_amd_direct.argtypes = [i, vp, vp, i, vp, vp]
    L.bbfmm_amd_direct_field.argtypes = [i, vp, i, vp, i, vp, vp, vp]
    L.bbfmm_amd_gradient.argtypes = [vp, vp, vp, vp, vp]
    L.bbfmm_amd_plan_update_positions.argtypes = [vp, vp]
    L.bbfmm_amd_plan_tune.argtypes = [vp, i, vp, i, vp]
    L.bbfmm_amd_batch_create.argtypes = [C.POINTER(vp), i, vp, vp, vp, i, R, i, i, vp, vp, i, vp, vp]
    L.bbfmm_amd_set_default_scheme.argtypes = [i]
    L.m2l_drv.argtypes = [i, vp, vp, vp, i, i, R, vp, vp, i, vp, vp, i, vp, i]
    for name in ("anotherFMMCompute",):
        getattr(L, name).argtypes = [C.POINTER(vp), vp, vp, vp, vp, vp, vp, vp, vp, vp, R, i, i, i, vp]
        getattr(L, name).restype = None
    L.anotherFMMInteraction.argtypes = [C.POINTER(vp), vp, vp, vp, vp, i, i, i, R]
    L.anotherFMMInteraction.restype = None
    L.anotherDownwardPass.argtypes = [C.POINTER(vp), vp, vp, vp, vp, vp, vp, i, i, i, R, vp]
    L.anotherDownwardPass.restype = None
    L.anotherBuildFMMHierachy.argtypes = [R, i, i, i, i, C.POINTER(vp), i, i, vp, vp]
    L.anotherBuildFMMHierachy.restype = None
    L.anotherFMMCleanup.argtypes = [C.POINTER(vp)]
    L.anotherFMMCleanup.restype = None
    _libs[prec] = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(L, rc):
    if rc != 0:
        raise BBFMMError(f"bbfmm_amd error {rc}: {L.bbfmm_amd_last_error_string().decode()}")
    return rc


def last_error(prec: str = "s"):
    L = load(prec)
    return L.bbfmm_amd_last_error(), L.bbfmm_amd_last_error_string().decode()


def maxlev(N: int, lmod: int = 0) -> int:
    """Level rule of the reference (bbfmm.c:151-160)."""
    return load("d").bbfmm_amd_maxlev(N, lmod)


def weights(n: int, prec: str = "s"):
    """(Tkz, Ktable, Kweights, Cweights) of ComputeWeights (bbfmm.c:922-1028)."""
    dt = dtype(prec)
    Tkz, Kt = np.zeros(n * n, dt), np.zeros(343, np.int32)
    Kw, Cw = np.zeros(n ** 3, dt), np.zeros(2 * n * n, dt)
    load(prec).bbfmm_amd_weights(n, _p(Tkz), _p(Kt), _p(Kw), _p(Cw))
    return Tkz, Kt, Kw, Cw


def operators(n: int, r: int | None = None, prec: str = "s", device: bool = False):
    """FP64 operator precompute, cast to the build's real: (U, K, sigma)
    with U n^3 x r and K 316 blocks of r x r, column-major (reference layout).
    device=True samples the kernel and forms the Gram matrix and the 316
    blocks on the GPU (the eigenproblem stays on the host)."""
    n3 = n ** 3
    r = r or n3 // 2
    dt = dtype(prec)
    U = np.zeros(n3 * r, dt)
    K = np.zeros(316 * r * r, dt)
    sig = np.zeros(n3)
    L = load(prec)
    fn = L.bbfmm_amd_operators_device if device else L.bbfmm_amd_operators
    _check(L, fn(n, r, _p(U), _p(K), _p(sig)))
    return U, K, sig


def write_operators(Kpath: str, Upath: str, n: int, U, K, prec: str = "s"):
    """Operator files in the reference's binary layout (bbfmm.c:1355-1403)."""
    dt = dtype(prec)
    U = np.ascontiguousarray(U, dt)
    K = np.ascontiguousarray(K, dt)
    r = U.size // n ** 3
    L = load(prec)
    _check(L, L.bbfmm_amd_write_operators(os.fsencode(Kpath), os.fsencode(Upath), n, r, _p(U), _p(K)))


def read_operators(Kpath: str, Upath: str, n: int, prec: str = "s"):
    """Operator files as FMMReadMatrices reads them (bbfmm.c:216-254): (U, K, r)."""
    L = load(prec)
    r = C.c_int(0)
    _check(L, L.bbfmm_amd_read_operators(os.fsencode(Kpath), os.fsencode(Upath), n, C.byref(r), None, None))
    dt = dtype(prec)
    U = np.zeros(r.value * n ** 3, dt)
    K = np.zeros(316 * r.value ** 2, dt)
    _check(L, L.bbfmm_amd_read_operators(os.fsencode(Kpath), os.fsencode(Upath), n, C.byref(r), _p(U), _p(K)))
    return U, K, r.value


def direct(pos, q, targets=None) -> np.ndarray:
    """FP64 direct sum on the GPU: q_t * potential_t for the target indices."""
    L = load("d")
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.float64)
    N = pos.shape[0]
    if targets is None:
        nt, tp, keep = N, None, None
    else:
        keep = np.ascontiguousarray(targets, dtype=np.int32)
        nt, tp = keep.size, _p(keep)
    out = np.zeros(nt)
    _check(L, L.bbfmm_amd_direct(N, _p(pos), _p(q), nt, tp, _p(out)))
    return out


def direct_field(pos, q, targets=None):
    """FP64 direct sum on the GPU for nvec charge vectors q (nvec x N):
    (potentials nvec x nt, gradients nvec x nt x 3) at the target indices."""
    L = load("d")
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(np.atleast_2d(q), dtype=np.float64)
    N = pos.shape[0]
    nvec = q.shape[0]
    if targets is None:
        nt, tp = N, None
    else:
        keep = np.ascontiguousarray(targets, dtype=np.int32)
        nt, tp = keep.size, _p(keep)
    pot = np.zeros((nvec, nt))
    grad = np.zeros((nvec, nt, 3))
    _check(L, L.bbfmm_amd_direct_field(N, _p(pos), nvec, _p(q), nt, tp, _p(pot), _p(grad)))
    return pot, grad


def _torch():
    import torch
    return torch


class Plan:
    """Persistent device plan (bbfmm_amd_plan_create), or with
    Plan.batch(...) one plan over many independent problems
    (bbfmm_amd_batch_create)."""

    def __init__(self, pos, n: int, U, K, nrhs: int = 1, lmod: int = 0, L: float = 1.0,
                 prec: str = "s", stream=None, _batch=None):
        self.prec = prec
        self.lib = load(prec)
        dt = dtype(prec)
        self.pos = np.ascontiguousarray(pos, dtype=dt).reshape(-1, 3)
        self.N = self.pos.shape[0]
        self.n = n
        self.nrhs = nrhs
        self._U = np.ascontiguousarray(U, dtype=dt)
        self._K = np.ascontiguousarray(K, dtype=dt)
        self.r = self._U.size // n ** 3
        h = C.c_void_p()
        if _batch is None:
            _check(self.lib, self.lib.bbfmm_amd_plan_create(
                C.byref(h), self.N, _p(self.pos), L, n, lmod, self.r, _p(self._K), _p(self._U), nrhs,
                C.c_void_p(stream)))
            self.sizes, self.depths = [self.N], None
        else:
            sizes, depths, leafmax = _batch
            self.sizes = [int(x) for x in sizes]
            if sum(self.sizes) != self.N:
                raise BBFMMError("instance sizes do not add up to the points given")
            ni = len(self.sizes)
            Ns = np.asarray(self.sizes, np.int32)
            dreq = None if depths is None else np.ascontiguousarray(depths, np.int32)
            dout = np.zeros(ni, np.int32)
            _check(self.lib, self.lib.bbfmm_amd_batch_create(
                C.byref(h), ni, _p(Ns), _p(self.pos), None if dreq is None else _p(dreq), leafmax, L, n,
                self.r, _p(self._K), _p(self._U), nrhs, C.c_void_p(stream), _p(dout)))
            self.depths = dout.tolist()
        self.h = h
        info = np.zeros(6, np.int32)
        dinfo = np.zeros(2)
        self.lib.bbfmm_amd_plan_info(self.h, _p(info), _p(dinfo))
        self.maxlev, self.ncell = int(info[0]), int(info[1])
        self.m2l_interactions, self.p2p_pairs = float(dinfo[0]), float(dinfo[1])

    @classmethod
    def batch(cls, pos_list, n: int, U, K, nrhs: int = 1, depths=None, leafmax: int = 32, L: float = 1.0,
              prec: str = "s", stream=None):
        """One plan over independent problems pos_list[i] (N_i x 3 each),
        each on a sparse octree of uniform depth (depths[i], or the smallest
        depth whose leaves hold at most leafmax points).  Charges and
        results are nrhs x sum(N_i), the instances concatenated."""
        sizes = [np.asarray(p).reshape(-1, 3).shape[0] for p in pos_list]
        pos = np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, 3) for p in pos_list])
        return cls(pos, n, U, K, nrhs=nrhs, L=L, prec=prec, stream=stream, _batch=(sizes, depths, leafmax))

    def set(self, scheme: int | None = None, form: int | None = None, timing: int | None = None,
            output: str | None = None, graph: bool | None = None, variant: int | None = None,
            level_switch: int | None = None, near_stream: bool | None = None, p2p_form: int | None = None,
            l2p_reference: bool | None = None, flat_loads: bool | None = None,
            variant_upper: int | None = None, upper_split: int | None = None):
        """output: "q_phi" (q .* potential, the reference's output) or "potential";
        l2p_reference: orders n = 3, 7 evaluate L2P with the reference's z
        weights (anotherDownwardPassX.c:304) instead of the interpolant;
        variant_upper: IJ split-BF16, the source-stationary variant of the
        levels above the deepest (a launch of their own; 0: as the deepest);
        upper_split: that launch takes each class group's steps in this many
        parts (1..16), their partial sums added to L afterwards;
        flat_loads: kernels with buffer-load and pointer forms take the
        pointer form (their form for arrays of 2 GB and more) at any size;
        graph: capture the whole solve into one hipGraph and replay it;
        near_stream: 16-vector MFMA solves run P2P beside the far field;
        p2p_form: 16-vector MFMA solves, P2P_FP32 (default) or P2P_BF16X2
        (the sum over sources as split-BF16 products, fp32 accumulation)."""
        if scheme is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_SCHEME, scheme))
        if form is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_FORM, form))
        if timing is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_TIMING, int(timing)))
        if output is not None:
            if output not in ("q_phi", "potential"):
                raise BBFMMError(f"unknown output {output!r}")
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_OUTPUT, int(output == "potential")))
        if variant is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_M2L_VARIANT, int(variant)))
        if level_switch is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_LEVEL_SWITCH, int(level_switch)))
        if near_stream is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_NEAR_STREAM, int(bool(near_stream))))
        if p2p_form is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_P2P_FORM, int(p2p_form)))
        if variant_upper is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_M2L_VARIANT_UPPER, int(variant_upper)))
        if upper_split is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_M2L_UPPER_SPLIT, int(upper_split)))
        if flat_loads is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_ADDRESSING, int(bool(flat_loads))))
        if l2p_reference is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_L2P_REFERENCE, int(bool(l2p_reference))))
        if graph is not None:
            _check(self.lib, self.lib.bbfmm_amd_plan_set(self.h, OPT_GRAPH, int(bool(graph))))
        return self

    def tune(self, reps: int = 3):
        """M2L autotuner: times every valid (scheme, variant, upper-level
        variant, level switch) on this plan and keeps the fastest; returns the
        table of (scheme, variant, variant_upper, level_switch, ms)."""
        table = np.zeros((64, 5))
        rows = C.c_int(0)
        _check(self.lib, self.lib.bbfmm_amd_plan_tune(self.h, reps, _p(table), 64, C.byref(rows)))
        return [tuple(r) for r in table[:min(rows.value, 64)].tolist()]

    def update_positions(self, pos):
        """Move the particles within their leaves (same tree and lists)."""
        dt = dtype(self.prec)
        p = np.ascontiguousarray(pos, dtype=dt).reshape(-1, 3)
        if p.shape[0] != self.N:
            raise BBFMMError("pos has the wrong size")
        _check(self.lib, self.lib.bbfmm_amd_plan_update_positions(self.h, _p(p)))
        self.pos = p
        return self

    def gradient(self, q, g, dq=None, dx=None):
        """Adjoint of the potentials for a plan of nrhs = 2m vectors: q, g
        (m x N torch tensors on the GPU; g = dLoss/dpotential) -> (dLoss/dq
        (m x N), dLoss/dpositions (N x 3))."""
        torch = _torch()
        want = torch.float32 if self.prec == "s" else torch.float64
        m = self.nrhs // 2
        for name, t in (("q", q), ("g", g)):
            if t.dtype != want or not t.is_contiguous() or t.numel() != m * self.N:
                raise BBFMMError(f"{name} must be a contiguous {want} tensor of {m} x {self.N}")
        if dq is None:
            dq = torch.empty_like(q)
        if dx is None:
            dx = torch.empty((self.N, 3), dtype=want, device=q.device)
        _check(self.lib, self.lib.bbfmm_amd_gradient(self.h, C.c_void_p(q.data_ptr()), C.c_void_p(g.data_ptr()),
                                                     C.c_void_p(dq.data_ptr()), C.c_void_p(dx.data_ptr())))
        return dq, dx

    def solve(self, q, phi=None):
        """q: torch tensor on the GPU (nrhs x N or N), returns phi = q .* potential."""
        torch = _torch()
        want = torch.float32 if self.prec == "s" else torch.float64
        if q.dtype != want or not q.is_contiguous():
            raise BBFMMError(f"q must be a contiguous {want} tensor")
        if q.numel() != self.N * self.nrhs:
            raise BBFMMError("q has the wrong size")
        if phi is None:
            phi = torch.empty_like(q)
        _check(self.lib, self.lib.bbfmm_amd_solve(self.h, C.c_void_p(q.data_ptr()),
                                                  C.c_void_p(phi.data_ptr())))
        return phi

    def solve_host(self, q):
        dt = dtype(self.prec)
        q = np.ascontiguousarray(q, dtype=dt)
        if q.size != self.N * self.nrhs:
            raise BBFMMError("q has the wrong size")
        phi = np.zeros_like(q)
        _check(self.lib, self.lib.bbfmm_amd_solve_host(self.h, _p(q), _p(phi)))
        return phi

    def m2l(self, M, Lout):
        _check(self.lib, self.lib.bbfmm_amd_m2l(self.h, C.c_void_p(M.data_ptr()),
                                                C.c_void_p(Lout.data_ptr())))
        return Lout

    def fetch(self, which: str):
        dt = dtype(self.prec)
        idx = {"M": 0, "L": 1, "F": 2, "phi_far": 3}[which]
        size = {0: self.ncell * self.r, 1: self.ncell * self.r, 2: self.ncell * self.n ** 3,
                3: self.N}[idx]
        out = np.zeros(size, dt)
        _check(self.lib, self.lib.bbfmm_amd_fetch(self.h, idx, _p(out)))
        return out

    def stage_times(self) -> dict:
        out = np.zeros(len(STAGES))
        k = self.lib.bbfmm_amd_stage_times(self.h, _p(out), len(STAGES))
        return dict(zip(STAGES[:k], out[:k].tolist()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.bbfmm_amd_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def m2l_drv(r, Lv, K, M, minlev, maxlev, L0, center, ineigh, neighbors, iinter, interaction,
            scheme, prec="s", pitch_neighbors=64, pitch_interaction=192) -> int:
    """The m2l_drv entry point on host arrays (Lv is written in place)."""
    L = load(prec)
    return L.m2l_drv(r, _p(Lv), _p(K), _p(M), minlev, maxlev, L0, _p(center), _p(ineigh),
                     pitch_neighbors, _p(neighbors), _p(iinter), pitch_interaction,
                     _p(interaction), scheme)


# ---------------------------------------------------------------- autograd
class _PlanCache:
    """Plans of the autograd path kept across calls.  An entry is reused while
    the order, precision, level modifier, vector count, particle count and
    stream are the same and the operators are equal: the same array objects
    as the entry's last call (taken as unmodified), or equal by value to the
    copies the entry keeps.  Positions need no transfer when they are the
    same GPU tensor object at the same version (torch counts in-place
    changes) or equal on the device; new positions are moved into the plan by
    bbfmm_amd_plan_update_positions while every particle stays in its leaf,
    and the plan is rebuilt when one leaves.  At most `size` plans are kept
    (least recently used first out)."""

    def __init__(self, size: int = 4):
        self.size = size
        # [key, U copy, K copy, (U, K) of the last call, device positions, tensor stamp, plan]
        self.entries = []
        self.builds = 0  # plans built (a test counts them)

    @staticmethod
    def _stamp(t):
        import weakref
        return (weakref.ref(t), t._version)

    @staticmethod
    def _same(stamp, t):
        return stamp is not None and stamp[0]() is t and stamp[1] == t._version

    def get(self, pos, n, U, K, prec, lmod, nrhs, stream):
        torch = _torch() if not isinstance(pos, np.ndarray) else None
        dev_pos = pos.detach() if torch is not None and pos.is_cuda else None
        N = pos.shape[0]
        key = (n, prec, lmod, nrhs, N, stream)
        hit = None
        for e in self.entries:
            if e[0] != key:
                continue
            if e[3][0] is U and e[3][1] is K:
                hit = e
                break
            Ua, Ka = np.asarray(U), np.asarray(K)
            if e[1].shape == Ua.shape and e[2].shape == Ka.shape and np.array_equal(e[1], Ua) \
                    and np.array_equal(e[2], Ka):
                hit = e
                break
        if hit is not None:
            self.entries.remove(hit)
            self.entries.append(hit)
            hit[3] = (U, K)
            plan = hit[6]
            if dev_pos is not None and hit[4] is not None:
                if self._same(hit[5], pos) or (hit[4].dtype == dev_pos.dtype and torch.equal(hit[4], dev_pos)):
                    hit[5] = self._stamp(pos)
                    return plan
            p = np.ascontiguousarray(pos.detach().cpu().numpy() if torch is not None else pos,
                                     dtype=dtype(prec)).reshape(-1, 3)
            if not np.array_equal(plan.pos, p):
                try:
                    plan.update_positions(p)
                except BBFMMError:
                    plan.close()
                    self.entries.remove(hit)
                    hit = None
            if hit is not None:
                hit[4] = dev_pos.clone() if dev_pos is not None else None
                hit[5] = self._stamp(pos) if dev_pos is not None else None
                return plan
        p = pos.detach().cpu().numpy() if torch is not None else pos
        plan = Plan(p, n, U, K, nrhs=nrhs, prec=prec, lmod=lmod, stream=stream)
        self.builds += 1
        self.entries.append([key, np.array(U, copy=True), np.array(K, copy=True), (U, K),
                             dev_pos.clone() if dev_pos is not None else None,
                             self._stamp(pos) if dev_pos is not None else None, plan])
        while len(self.entries) > self.size:
            self.entries.pop(0)[6].close()
        return plan

    def clear(self):
        for e in self.entries:
            e[6].close()
        self.entries.clear()


plan_cache = _PlanCache()


def _make_autograd():
    torch = _torch()

    class FMMPotential(torch.autograd.Function):
        """phi = potential(pos, q): the FMM potentials of m charge vectors
        (q: m x N, pos: N x 3, both on the GPU) with dLoss/dq and
        dLoss/dpos from the adjoint solve (bbfmm_amd_gradient).  The
        backward recomputes the expansions (a solve of the 2m vectors
        [q; dLoss/dphi]); nothing but the inputs is kept from the forward.
        Plans come from plan_cache: the tree and lists are built once per
        geometry, and positions that stay in their leaves are only uploaded.
        cfg = (n, U, K, prec, lmod, scheme, form)."""

        @staticmethod
        def forward(ctx, pos, q, cfg):
            n, U, K, prec, lmod, scheme, form = cfg
            m = q.shape[0]
            plan = plan_cache.get(pos, n, U, K, prec, lmod, m, torch.cuda.current_stream().cuda_stream)
            plan.set(scheme=scheme, form=form, output="potential")
            phi = plan.solve(q.detach().contiguous())
            ctx.save_for_backward(pos, q)
            ctx.cfg = cfg
            return phi

        @staticmethod
        def backward(ctx, gphi):
            pos, q = ctx.saved_tensors
            n, U, K, prec, lmod, scheme, form = ctx.cfg
            m = q.shape[0]
            plan = plan_cache.get(pos, n, U, K, prec, lmod, 2 * m, torch.cuda.current_stream().cuda_stream)
            plan.set(scheme=scheme, form=form)
            dq, dx = plan.gradient(q.detach().contiguous(), gphi.detach().contiguous())
            return dx.to(pos.dtype), dq, None

    return FMMPotential


class _Lazy:
    cls = None


def __getattr__(name):
    if name == "FMMPotential":
        if _Lazy.cls is None:
            _Lazy.cls = _make_autograd()
        return _Lazy.cls
    raise AttributeError(name)


def potential(pos, q, n: int, U, K, prec: str = "s", lmod: int = 0, scheme: int = 3,
              form: int | None = None):
    """Differentiable FMM potentials (torch): pos N x 3, q m x N on the GPU,
    returns m x N potentials; gradients flow to pos and q."""
    if form is None:  # the n = 4 precision form once the M2L has several columns per cell
        form = FORM_BF16X2 if (prec == "s" and n == 4 and q.shape[0] >= 8) else FORM_PLAIN
    return __getattr__("FMMPotential").apply(pos, q, (n, U, K, prec, lmod, scheme, form))
