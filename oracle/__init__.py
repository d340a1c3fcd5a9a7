"""Python loader for the bbFMM oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a CPU restatement of the reference's algorithm in plain C
(oracle/bbfmm_oracle.c) plus numpy for dense algebra.  Only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py may import this
module, and only as the checker; the product never does.

Layouts follow the reference (SURVEY-free summary):
  positions  N x 3 (real3, 12/24 bytes), charges N
  K          316 blocks of r x r, column-major, Ktable order (bbfmm.c:1361-1406)
  U          n^3 x r column-major (U[i + n3*j])
  M, L       [cell][r]; F [cell][n^3]; cells numbered as cell.h:27-28
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
PITCH_NEIGHBORS = 64
PITCH_INTERACTION = 192
PITCH_SOURCECLUSTERS = 32

_libs = {}


def build():
    """Compile the oracle shared libraries (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(prec: str = "s"):
    if prec not in _libs:
        path = os.path.join(HERE, f"liboracle_{prec}.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_p2p_pairs.restype = C.c_double
        _libs[prec] = L
    return _libs[prec]


def dtype(prec: str):
    return np.float32 if prec == "s" else np.float64


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _r(prec, x):
    return C.c_float(x) if prec == "s" else C.c_double(x)


def ncell(maxlev: int) -> int:
    return (8 ** (maxlev + 1) - 1) // 7


def levsta(maxlev: int):
    return [(8 ** l - 1) // 7 for l in range(maxlev + 1)]


def maxlev_rule(N: int, lmod: int = 0) -> int:
    return lib("d").orc_maxlev(C.c_int(N), C.c_int(lmod))


class Tree:
    """Complete octree with lists, as anotherBuildFMMHierachy builds it."""

    def __init__(self, pos: np.ndarray, lmod: int = 0, prec: str = "s", L: float = 1.0,
                 maxlev: int | None = None):
        self.prec = prec
        dt = dtype(prec)
        self.pos = np.ascontiguousarray(pos, dtype=dt)
        self.N = self.pos.shape[0]
        self.maxlev = maxlev if maxlev is not None else maxlev_rule(self.N, lmod)
        self.L = L
        lb = lib(prec)
        nc = ncell(self.maxlev)
        self.ncell = nc
        self.center = np.zeros((nc, 3), dt)
        lb.orc_centers(C.c_int(self.maxlev), _r(prec, L), _p(self.center))
        self.list = np.zeros(max(self.N, 1), np.int32)
        self.sta = np.zeros(nc, np.int32)
        self.end = np.zeros(nc, np.int32)
        rc = lb.orc_distribute(C.c_int(self.maxlev), _r(prec, L), _p(self.center), C.c_int(self.N),
                               _p(self.pos), _p(self.list), _p(self.sta), _p(self.end))
        if rc != 0:
            raise ValueError("particle outside the root cell")
        self.ineigh = np.zeros(nc, np.int32)
        self.iinter = np.zeros(nc, np.int32)
        self.neighbors = np.full(nc * PITCH_NEIGHBORS, -1, np.int32)
        self.interaction = np.full(nc * PITCH_INTERACTION, -1, np.int32)
        lb.orc_lists(C.c_int(self.maxlev), _r(prec, L), _p(self.center), _p(self.ineigh),
                     _p(self.neighbors), _p(self.iinter), _p(self.interaction))
        self.Ktable = np.zeros(343, np.int32)
        lb.orc_Ktable(_p(self.Ktable))
        self.Kindex = np.full(nc * PITCH_INTERACTION, -1, np.int32)
        lb.orc_Kindex(C.c_int(self.maxlev), _r(prec, L), _p(self.center), _p(self.iinter),
                      _p(self.interaction), _p(self.Ktable), _p(self.Kindex))
        nclu = (8 ** self.maxlev - 8) // 7
        self.sourceclusters = np.full(max(nclu, 1) * PITCH_SOURCECLUSTERS, -1, np.int32)
        lb.orc_sourceclusters(C.c_int(2), C.c_int(self.maxlev), _r(prec, L), _p(self.center),
                              _p(self.ineigh), _p(self.neighbors), _p(self.sourceclusters))

    def p2p_pairs(self) -> float:
        return lib(self.prec).orc_p2p_pairs(C.c_int(self.maxlev), _p(self.sta), _p(self.end),
                                            _p(self.ineigh), _p(self.neighbors))

    def m2l_interactions(self, minlev: int = 2) -> int:
        ls = levsta(self.maxlev)
        return int(self.iinter[ls[minlev]:].sum())


def sibling_tables():
    k0 = np.zeros(27, np.int32)
    k1 = np.zeros(27, np.int32)
    lib("d").orc_sibling_tables(_p(k0), _p(k1))
    return k0, k1


def weights(n: int, prec: str = "s"):
    dt = dtype(prec)
    Tkz = np.zeros(n * n, dt)
    Kt = np.zeros(343, np.int32)
    Kw = np.zeros(n ** 3, dt)
    Cw = np.zeros(2 * n * n, dt)
    lib(prec).orc_weights(C.c_int(n), _p(Tkz), _p(Kt), _p(Kw), _p(Cw))
    return Tkz, Kt, Kw, Cw


def kernel_matrix(n: int) -> np.ndarray:
    """FP64 K0, n^3 x 316 n^3 column-major (flat)."""
    n3 = n ** 3
    K0 = np.zeros(316 * n3 * n3, np.float64)
    lib("d").orc_kernel_matrix(C.c_int(n), _p(K0))
    return K0


def sym_eig(A: np.ndarray):
    m = A.shape[0]
    a = np.array(A, dtype=np.float64, order="F").ravel(order="F").copy()
    w = np.zeros(m)
    V = np.zeros(m * m)
    sweeps = lib("d").orc_sym_eig(C.c_int(m), _p(a), _p(w), _p(V))
    if sweeps < 0:
        raise RuntimeError("Jacobi eigensolver did not converge")
    return w, V.reshape(m, m).T.copy()  # columns are eigenvectors


def operators(n: int, r: int | None = None):
    """Operator precompute of ComputeKernelSVD (bbfmm.c:1209-1421) in FP64.

    Left singular vectors of K0 are the eigenvectors of K0 K0^T (Jacobi);
    returns (U flat n3*r column-major, K flat 316*r*r column-major, sigma).
    """
    n3 = n ** 3
    if r is None:
        r = n3 // 2  # bbfmm.c:1340-1343: n^3/2 - 1, then +1
    K0 = kernel_matrix(n)
    A = K0.reshape(316 * n3, n3)  # row = column index of K0; A[c, i] = K0[i, c]
    G = A.T @ A  # K0 K0^T
    w, V = sym_eig(G)
    U = np.ascontiguousarray(V[:, :r])  # n3 x r
    Uf = U.ravel(order="F").copy()
    E = np.zeros(316 * r * r)
    lib("d").orc_transfer_operators(C.c_int(n), C.c_int(r), _p(K0), _p(Uf), _p(E))
    return Uf, E, np.sqrt(np.maximum(w, 0))


def fmm(tree: Tree, q: np.ndarray, n: int, Uf: np.ndarray, K: np.ndarray,
        scheme: str = "basic", l2p_exact: bool = False):
    """Full oracle solve (anotherFMMCompute.c): returns a dict with M, L, F,
    phi_far and phi (= q*phi, the reference's output convention).
    l2p_exact: at n % 4 == 3 evaluate the L2P interpolant itself instead of
    restating the reference's z weights (anotherDownwardPassX.c:304)."""
    prec = tree.prec
    dt = dtype(prec)
    lb = lib(prec)
    n3 = n ** 3
    r = Uf.size // n3
    q = np.ascontiguousarray(q, dtype=dt)
    Uf = np.ascontiguousarray(Uf, dtype=dt)
    K = np.ascontiguousarray(K, dtype=dt)
    Tkz, Kt, Kw, Cw = weights(n, prec)
    nc = tree.ncell
    M = np.zeros(nc * r, dt)
    lb.orc_upward(C.c_int(tree.maxlev), _r(prec, tree.L), C.c_int(n), C.c_int(r), _p(tree.center),
                  _p(tree.list), _p(tree.sta), _p(tree.end), _p(tree.pos), _p(q), _p(Uf),
                  _p(Tkz), _p(Kw), _p(Cw), _p(M))
    Lv = np.zeros(nc * r, dt)
    if scheme == "basic":
        lb.orc_m2l_basic(C.c_int(r), C.c_int(2), C.c_int(tree.maxlev), _p(K), _p(M),
                         _p(tree.iinter), _p(tree.interaction), _p(tree.Kindex), _p(Lv))
    else:
        lb.orc_m2l_sibling(C.c_int(r), C.c_int(2), C.c_int(tree.maxlev), _p(K), _p(M),
                           _p(tree.sourceclusters), _p(tree.Ktable), _p(Lv))
    F = np.zeros(nc * n3, dt)
    lb.orc_postm2l(C.c_int(tree.maxlev), _r(prec, tree.L), C.c_int(n), C.c_int(r), _r(prec, 1.0),
                   _p(Uf), _p(Kw), _p(Lv), _p(F))
    F0 = F.copy()
    phi = np.zeros(tree.N, dt)
    lb.orc_set_l2p_exact(C.c_int(int(bool(l2p_exact))))
    try:
        lb.orc_downward(C.c_int(tree.maxlev), _r(prec, tree.L), C.c_int(n), _p(tree.center),
                        _p(tree.list), _p(tree.sta), _p(tree.end), _p(tree.pos), _p(q),
                        _p(Tkz), _p(Cw), _p(F), _p(phi))
    finally:
        lb.orc_set_l2p_exact(C.c_int(0))
    phi_far = phi.copy()
    lb.orc_nearfield(C.c_int(tree.maxlev), C.c_int(tree.N), _p(tree.list), _p(tree.sta),
                     _p(tree.end), _p(tree.ineigh), _p(tree.neighbors), _p(tree.pos), _p(q),
                     _p(phi))
    return {"M": M, "L": Lv, "F": F0, "F_l2l": F, "phi_far": phi_far, "phi": phi}


def m2l(tree: Tree, K: np.ndarray, M: np.ndarray, r: int, scheme: str = "basic"):
    dt = dtype(tree.prec)
    K = np.ascontiguousarray(K, dtype=dt)
    M = np.ascontiguousarray(M, dtype=dt)
    Lv = np.zeros(tree.ncell * r, dt)
    lb = lib(tree.prec)
    if scheme == "basic":
        lb.orc_m2l_basic(C.c_int(r), C.c_int(2), C.c_int(tree.maxlev), _p(K), _p(M),
                         _p(tree.iinter), _p(tree.interaction), _p(tree.Kindex), _p(Lv))
    else:
        lb.orc_m2l_sibling(C.c_int(r), C.c_int(2), C.c_int(tree.maxlev), _p(K), _p(M),
                           _p(tree.sourceclusters), _p(tree.Ktable), _p(Lv))
    return Lv


def direct(pos: np.ndarray, q: np.ndarray, targets: np.ndarray | None = None) -> np.ndarray:
    """FP64 direct sum phi_t (not multiplied by q_t)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    N = pos.shape[0]
    if targets is None:
        nt, tp = N, None
    else:
        targets = np.ascontiguousarray(targets, dtype=np.int32)
        nt, tp = targets.size, _p(targets)
    out = np.zeros(nt)
    lib("d").orc_direct(C.c_int(N), _p(pos), _p(q), C.c_int(nt), tp, _p(out))
    return out


def rel_l2(ref, x) -> float:
    ref = np.asarray(ref, np.float64)
    x = np.asarray(x, np.float64)
    return float(np.linalg.norm(ref - x) / np.linalg.norm(ref))


def uniform_inputs(N: int, seed: int = 0, nrhs: int = 1, qlo: float = -1.0, qhi: float = 1.0):
    """Synthetic inputs: positions i.i.d. U[0,1)^3 shifted into the root cell
    [-0.5,0.5)^3, then charges U[qlo,qhi) (one row per right-hand side)."""
    rng = np.random.default_rng(seed)
    pos = rng.random((N, 3)) - 0.5
    q = rng.uniform(qlo, qhi, size=(nrhs, N))
    return pos, (q[0] if nrhs == 1 else q)


# ---------------------------------------------------------------- reference
def ref_available(prec: str = "s", fast: bool = False) -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"ref_fmm_{'fast_' if fast else ''}{prec}"))


def build_ref():
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile.ref")], check=True)


def write_operator_files(dirname: str, prec: str, Uf: np.ndarray, K: np.ndarray, n: int):
    """Reference K/U file formats (bbfmm.c:1361-1406, read at :216-254)."""
    dt = dtype(prec)
    r = Uf.size // n ** 3
    with open(os.path.join(dirname, "U.bin"), "wb") as f:
        f.write(np.int32(r).tobytes())
        f.write(np.ascontiguousarray(Uf, dt).tobytes())
    with open(os.path.join(dirname, "K.bin"), "wb") as f:
        f.write(np.ascontiguousarray(K, dt).tobytes())


def ref_run(pos, q, n, Uf, K, prec="s", lmod=0, dump=True, fast=False, threads=None, reps=1,
            bind=False):
    """Run the compiled reference CPU path (oracle/_ref) on these inputs.

    reps > 1: the solve runs reps times on one tree, the first a warm-up;
    info["t_solve"] is the median of the others.  bind: OMP_PROC_BIND=close,
    OMP_PLACES=cores (the threads pinned one per core)."""
    exe = os.path.join(REF_DIR, f"ref_fmm_{'fast_' if fast else ''}{prec}")
    dt = dtype(prec)
    N = pos.shape[0]
    n3 = n ** 3
    r = Uf.size // n3
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    if bind:
        env["OMP_PROC_BIND"] = "close"
        env["OMP_PLACES"] = "cores"
    with tempfile.TemporaryDirectory() as d:
        np.ascontiguousarray(pos, dt).tofile(os.path.join(d, "pos.bin"))
        np.ascontiguousarray(q, dt).tofile(os.path.join(d, "q.bin"))
        write_operator_files(d, prec, Uf, K, n)
        res = subprocess.run([exe, d, str(N), str(n), str(lmod), "1" if dump else "0", str(int(reps))],
                             capture_output=True, text=True, env=env)
        if res.returncode != 0:
            raise RuntimeError(res.stderr[-2000:])
        info = json.loads(res.stdout.strip().splitlines()[-1])
        info["stderr"] = res.stderr
        out = {"info": info, "phi": np.fromfile(os.path.join(d, "phi.bin"), dt)}
        if dump:
            ml = info["maxlev"]
            nc = ncell(ml)
            out["center"] = np.fromfile(os.path.join(d, "ref_center.bin"), dt).reshape(nc, 3)
            out["list"] = np.fromfile(os.path.join(d, "ref_list.bin"), np.int32)
            out["sta"] = np.fromfile(os.path.join(d, "ref_sta.bin"), np.int32)
            out["end"] = np.fromfile(os.path.join(d, "ref_end.bin"), np.int32)
            for k in ("ineigh", "iinter", "neighbors", "interaction"):
                out[k] = np.fromfile(os.path.join(d, f"ref_{k}.bin"), np.int32)
            out["M"] = np.fromfile(os.path.join(d, "ref_M.bin"), dt)
            out["F"] = np.fromfile(os.path.join(d, "ref_F.bin"), dt)
            out["phi_far"] = np.fromfile(os.path.join(d, "ref_phifar.bin"), dt)
        return out
