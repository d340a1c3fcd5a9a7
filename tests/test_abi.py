"""The C-ABI libraries on the CPU: they load, export every entry point that
include/bbfmm_amd.h declares, their host-side setup (level rule, tree, lists,
Chebyshev weights, FP64 operators) agrees with the oracle, and the error
paths return codes instead of aborting.  No GPU compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import bbfmmongpu_amd as B
import oracle as O

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "bbfmm_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)  # preprocessor lines
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text))
    names.discard("if")
    return sorted(names)


@pytest.mark.parametrize("prec", ["s", "d"])
def test_library_exports_declared_symbols(prec):
    lib = B.load(prec)
    names = declared_functions()
    assert "anotherFMMCompute" in names and "m2l_drv" in names and "bbfmm_amd_solve" in names
    for name in names:
        assert hasattr(lib, name), name
    C.c_double.in_dll(lib, "kernel_exec_time")


def test_level_rule_matches_oracle():
    for N in (100, 4096, 65536, 262144, 1 << 20, 1 << 22):
        for lmod in (-1, 0, 1, 2):
            assert B.maxlev(N, lmod) == O.maxlev_rule(N, lmod)


@pytest.mark.parametrize("prec", ["s", "d"])
@pytest.mark.parametrize("n", [4, 8])
def test_weights_bit_exact_with_oracle(prec, n):
    a = B.weights(n, prec)
    b = O.weights(n, prec)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


class _Cell(C.Structure):
    _fields_ = [("neighbors", C.POINTER(C.c_int)), ("interaction", C.POINTER(C.c_int)),
                ("center", C.c_void_p), ("fieldsta", C.POINTER(C.c_int)), ("fieldend", C.POINTER(C.c_int)),
                ("sourcesta", C.POINTER(C.c_int)), ("sourceend", C.POINTER(C.c_int)),
                ("ineigh", C.POINTER(C.c_int)), ("iinter", C.POINTER(C.c_int)),
                ("pitch_neighbors", C.c_int), ("pitch_interaction", C.c_int)]


def _tree_struct(real):
    class _Tree(C.Structure):
        _fields_ = [("ncell", C.c_int), ("minlev", C.c_int), ("maxlev", C.c_int),
                    ("levsta", C.POINTER(C.c_int)), ("levend", C.POINTER(C.c_int)),
                    ("celeng", C.POINTER(real)), ("c", C.POINTER(_Cell)), ("fieldval", C.c_void_p),
                    ("sourceval", C.c_void_p), ("proxysval", C.c_void_p), ("ncell0", C.c_int),
                    ("Nf", C.c_int), ("Ns", C.c_int), ("fieldlist", C.POINTER(C.c_int)),
                    ("sourcelist", C.POINTER(C.c_int))]
    return _Tree


@pytest.mark.parametrize("prec", ["s", "d"])
@pytest.mark.parametrize("N,lmod", [(5000, 0), (30000, 0), (3000, 1)])
def test_tree_builder_bit_exact_with_oracle(prec, N, lmod):
    lib = B.load(prec)
    real = C.c_float if prec == "s" else C.c_double
    pos, _ = O.uniform_inputs(N, seed=6)
    pos = pos.astype(O.dtype(prec))
    T = O.Tree(pos, lmod, prec)
    h = C.c_void_p()
    lib.anotherBuildFMMHierachy(1.0, 4, 1, T.maxlev, 32, C.byref(h), N, N, O._p(pos), O._p(pos))
    assert lib.bbfmm_amd_last_error() == 0
    t = C.cast(h, C.POINTER(_tree_struct(real))).contents
    c = t.c.contents
    nc = t.ncell
    assert nc == T.ncell and t.maxlev == T.maxlev and t.minlev == 2
    arr = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy()
    assert np.array_equal(arr(t.fieldlist, N), T.list)
    ls = O.levsta(T.maxlev)[-1]
    assert np.array_equal(arr(c.fieldsta, nc)[ls:], T.sta[ls:])
    assert np.array_equal(arr(c.fieldend, nc)[ls:], T.end[ls:])
    ctr = np.ctypeslib.as_array(C.cast(c.center, C.POINTER(real)), shape=(nc * 3,)).reshape(nc, 3)
    assert np.array_equal(ctr, T.center)
    assert np.array_equal(arr(c.ineigh, nc), T.ineigh) and np.array_equal(arr(c.iinter, nc), T.iinter)
    nb = arr(c.neighbors, nc * 64)
    it = arr(c.interaction, nc * 192)
    for A in range(nc):
        assert np.array_equal(nb[A * 64:A * 64 + T.ineigh[A]], T.neighbors[A * 64:A * 64 + T.ineigh[A]])
        assert np.array_equal(it[A * 192:A * 192 + T.iinter[A]], T.interaction[A * 192:A * 192 + T.iinter[A]])
    lib.anotherFMMCleanup(C.byref(h))
    assert h.value is None


def test_operators_match_oracle_subspace():
    U, K, sig = B.operators(4, 32, prec="d")
    Uo, Ko, sigo = O.operators(4)
    assert np.allclose(sig[:32], sigo[:32], rtol=1e-9)
    Um, Uom = U.reshape(32, 64).T, Uo.reshape(32, 64).T
    assert np.abs(Um.T @ Um - np.eye(32)).max() < 1e-12
    P, Po = Um[:, :30] @ Um[:, :30].T, Uom[:, :30] @ Uom[:, :30].T  # complete singular clusters
    assert np.abs(P - Po).max() < 1e-7  # two Gram-based solves: ~eps (s_0 / gap)^2
    # E = U^T K0 U for the library's own U; and E_{-v} = E_v^T
    K0 = O.kernel_matrix(4).reshape(316, 64, 64)  # [k][src][fld]
    E = K.reshape(316, 32, 32)
    for k in (0, 99, 315):
        assert np.allclose(E[k].T, Um.T @ K0[k].T @ Um, atol=1e-12)
        assert np.allclose(E[k].T, E[315 - k], atol=1e-12 * np.abs(E[k]).max())


def test_error_paths_without_gpu():
    lib = B.load("s")
    U, K, _ = B.operators(4, 32, prec="s")
    pos = np.zeros((10, 3), np.float32)
    h = C.c_void_p()
    # unsupported orders (outside 3..8): rejected before any device call
    for n in (2, 9):
        rc = lib.bbfmm_amd_plan_create(C.byref(h), 10, O._p(pos), 1.0, n, 0, 32, O._p(K), O._p(U), 1, None)
        assert rc == 1 and b"not supported" in lib.bbfmm_amd_last_error_string()
    # particle outside the root cell
    bad = np.full((10, 3), 0.9, np.float32)
    rc = lib.bbfmm_amd_plan_create(C.byref(h), 10, O._p(bad), 1.0, 4, 0, 32, O._p(K), O._p(U), 1, None)
    assert rc == 5
    # m2l_drv: unsupported rank and scheme return M2L_EXIT_FAIL (the reference aborts / ignores)
    nc = O.ncell(2)
    z = np.zeros(nc * 192, np.int32)
    Lv = np.zeros(nc * 32, np.float32)
    ctr = np.zeros((nc, 3), np.float32)
    for r in (0, 513):  # any rank 1..512 runs (the any-rank kernels)
        assert B.m2l_drv(r, Lv, K, Lv, 2, 2, 1.0, ctr, z, z, z, z, 0, prec="s") == 1
        assert f"r={r}".encode() in lib.bbfmm_amd_last_error_string()
    assert B.m2l_drv(32, Lv, K, Lv, 2, 2, 1.0, ctr, z, z, z, z, 7, prec="s") == 1
    assert b"scheme" in lib.bbfmm_amd_last_error_string()
    assert lib.bbfmm_amd_set_default_scheme(9) == 1


@pytest.mark.parametrize("prec", ["s", "d"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_particle_just_below_root_cell_is_rejected(prec, axis):
    """A coordinate in (-L/2 - leaf length, -L/2) truncates to leaf index 0
    (the reference's bucketing, anotherFMMDistribute.c:267-269), so only an
    explicit lower-bound check rejects it; it must not be accepted and
    extrapolated by P2M/L2P."""
    lib = B.load(prec)
    dt = O.dtype(prec)
    U, K, _ = B.operators(4, 32, prec=prec)
    pos, _ = O.uniform_inputs(100, seed=3)
    pos = pos.astype(dt)
    pos[17, axis] = -0.5 - 1e-3  # inside the first leaf's width below the root cell
    h = C.c_void_p()
    rc = lib.bbfmm_amd_plan_create(C.byref(h), 100, O._p(pos), 1.0, 4, 0, 32, O._p(K), O._p(U), 1, None)
    assert rc == 5, lib.bbfmm_amd_last_error_string()
    t = C.c_void_p()
    one = 1.0
    lib.anotherBuildFMMHierachy(one, 4, 1, 2, 32, C.byref(t), 100, 100, O._p(pos), O._p(pos))
    assert lib.bbfmm_amd_last_error() == 5
    lib.anotherFMMCleanup(C.byref(t))
    Ns = np.array([50, 50], np.int32)
    rc = lib.bbfmm_amd_batch_create(C.byref(h), 2, O._p(Ns), O._p(pos), None, 32, one, 4, 32, O._p(K), O._p(U), 1,
                                    None, None)
    assert rc == 5, lib.bbfmm_amd_last_error_string()
    # exactly -L/2 is inside (leaf 0), as in the reference
    pos[17, axis] = -0.5
    lib.anotherBuildFMMHierachy(one, 4, 1, 2, 32, C.byref(t), 100, 100, O._p(pos), O._p(pos))
    assert lib.bbfmm_amd_last_error() == 0
    lib.anotherFMMCleanup(C.byref(t))


def test_no_gpu_is_reported_not_faked():
    """Without a visible GPU the library reports BBFMM_E_NODEVICE; there is no
    CPU fallback."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    lib = B.load("s")
    U, K, _ = B.operators(4, 32, prec="s")
    pos, _ = O.uniform_inputs(1000, seed=0)
    pos = pos.astype(np.float32)
    h = C.c_void_p()
    rc = lib.bbfmm_amd_plan_create(C.byref(h), 1000, O._p(pos), 1.0, 4, 0, 32, O._p(K), O._p(U), 1, None)
    assert rc == 4, lib.bbfmm_amd_last_error_string()


@pytest.mark.parametrize("prec", ["s", "d"])
def test_operator_files_match_reference_layout(prec, tmp_path):
    """The library writes the reference's K/U file layout byte for byte (the
    oracle's writer is the one the compiled reference reads) and reads it back."""
    U, K, _ = B.operators(4, 32, prec=prec)
    B.write_operators(str(tmp_path / "K.bin"), str(tmp_path / "U.bin"), 4, U, K, prec=prec)
    ref_dir = tmp_path / "oracle"
    ref_dir.mkdir()
    O.write_operator_files(str(ref_dir), prec, U, K, 4)
    for name in ("K.bin", "U.bin"):
        assert (tmp_path / name).read_bytes() == (ref_dir / name).read_bytes(), name
    U2, K2, r = B.read_operators(str(tmp_path / "K.bin"), str(tmp_path / "U.bin"), 4, prec=prec)
    assert r == 32 and np.array_equal(U2, U) and np.array_equal(K2, K)
    # truncated file: an error code, not a crash
    (tmp_path / "K.bin").write_bytes((tmp_path / "K.bin").read_bytes()[:100])
    with pytest.raises(B.BBFMMError):
        B.read_operators(str(tmp_path / "K.bin"), str(tmp_path / "U.bin"), 4, prec=prec)


def test_device_operators_need_a_gpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    with pytest.raises(B.BBFMMError, match="error 4"):
        B.operators(4, 32, prec="d", device=True)


def test_direct_noalloc_rejects_too_many_targets():
    """bbfmm_amd_direct_noalloc: the partial sums live in a fixed device
    array, so more targets than it holds are an argument error, raised
    before any device call."""
    lib = B.load("d")
    lib.bbfmm_amd_direct_noalloc.restype = C.c_int
    lib.bbfmm_amd_direct_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                             C.c_void_p]
    mx = 16384  # BBFMM_DIRECT_NOALLOC_MAX_TARGETS
    dummy = C.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib.bbfmm_amd_direct_noalloc(100000, dummy, dummy, mx + 1, None, dummy, None)
    assert rc == 1 and b"maximum" in lib.bbfmm_amd_last_error_string()
    rc = lib.bbfmm_amd_direct_noalloc(10, dummy, dummy, 20, None, dummy, None)  # identity targets beyond N
    assert rc == 1
