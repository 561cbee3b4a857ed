"""The oracle's BCRS storage with 4x4 blocks (lsd_mkb/lad_bcrs), pinned
against the reference's own las_bcrs_intf.c compiled into oracle/_ref:
pattern, element assembly (with and without assembly tables) and residual
bit-exact; the 4x4 LU (the reference calls LAPACK dgetrf/dgetrs, absent
here) by factor reconstruction and by solving; block Jacobi / Gauss-Seidel by
convergence to the dense solution and by the reference's sweep-count rule.
Skipped parts when oracle/_ref was not built."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import P, bcrs_assemble, bcrs_vector_problem, oracle, ref, ref_available  # noqa: E402

need_bcrs = pytest.mark.skipif(not ref_available("libref_bcrs.so"), reason="oracle/_ref not built")


class BcrsStruct(ctypes.Structure):
    """itt_bcrs_matrices (lsd_mkb/lad_bcrs/lah_bcrs.h)"""
    _fields_ = [(n, ctypes.c_int) for n in ("Max_SM_size", "Nrblocks", "Nrdofgl", "Nr_sm_blocks",
                                            "Half_bandwidth", "Max_nrneig", "Precon", "ILU_k",
                                            "ILU_nr_sm_blocks", "ILU_half_bandwidth")] + \
              [("rhs", ctypes.POINTER(ctypes.c_double)),
               ("block_val", ctypes.POINTER(ctypes.c_double)),
               ("block_col_ind", ctypes.POINTER(ctypes.c_int)),
               ("block_row_ptr", ctypes.POINTER(ctypes.c_int)),
               ("block_diag_ptr", ctypes.POINTER(ctypes.c_int)),
               ("block_diag_precon", ctypes.POINTER(ctypes.c_double)),
               ("block_diag_ips", ctypes.POINTER(ctypes.c_int)),
               ("ilu_diag_ptr", ctypes.POINTER(ctypes.c_int)),
               ("ilu_val", ctypes.POINTER(ctypes.c_double)),
               ("ilu_col_ind", ctypes.POINTER(ctypes.c_int)),
               ("ilu_row_ptr", ctypes.POINTER(ctypes.c_int)),
               ("pivots", ctypes.POINTER(ctypes.c_double))]


def vector_problem(nx, ny, nz, seed):
    mesh = mf.Mesh.box_prism(nx, ny, nz, seed=seed, jitter=0.2)
    return (mesh,) + bcrs_vector_problem(mesh.arrays(), seed)


def ref_bcrs():
    L = ref("libref_bcrs.so")
    L.lar_allocate_SM_and_LV_bcrs.restype = ctypes.c_int
    L.lar_allocate_SM_and_LV_bcrs.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2
    L.lar_assemble_SM_and_LV_bcrs.restype = ctypes.c_int
    L.lar_assemble_SM_and_LV_bcrs.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 5
    L.lar_fill_assembly_table_int_ent_bcrs.restype = ctypes.c_int
    L.lar_fill_assembly_table_int_ent_bcrs.argtypes = [ctypes.c_int] * 2 + [ctypes.c_void_p] * 3
    L.lar_assemble_SM_and_LV_with_table_bcrs.restype = ctypes.c_int
    L.lar_assemble_SM_and_LV_with_table_bcrs.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 4
    L.lar_compute_residual_bcrs.restype = None
    L.lar_compute_residual_bcrs.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 3
    L.lar_free_SM_and_LV_bcrs.restype = ctypes.c_int
    L.lar_free_SM_and_LV_bcrs.argtypes = [ctypes.c_int]
    return L


def ref_allocate(L, nn, nb_lists):
    arrs = [np.ascontiguousarray(x, np.int32) for x in nb_lists]
    ptrs = (ctypes.c_void_p * nn)(*[P(x) for x in arrs])
    nro = np.array([len(x) for x in nb_lists], np.int32)
    mid = L.lar_allocate_SM_and_LV_bcrs(4 * nn, 24, 4, P(nro), ctypes.cast(ptrs, ctypes.c_void_p))
    return mid, arrs


def ref_arrays(L, mid, nn):
    st = (BcrsStruct * 20).in_dll(L, "itv_bcrs_matrices")[mid]
    nsm = st.Nr_sm_blocks
    rp = np.ctypeslib.as_array(st.block_row_ptr, (nn + 1,)).copy()
    ci = np.ctypeslib.as_array(st.block_col_ind, (nsm,)).copy()
    val = np.ctypeslib.as_array(st.block_val, (16 * nsm,)).copy()
    rhs = np.ctypeslib.as_array(st.rhs, (4 * nn,)).copy()
    return rp, ci, val, rhs


@need_bcrs
def test_pattern_assembly_and_residual_bitexact_with_reference():
    L = ref_bcrs()
    mesh, en, nn, nb_lists, mats = vector_problem(3, 3, 2, 5)
    mid, keep = ref_allocate(L, nn, nb_lists)
    assert mid >= 0
    for e, (K, F) in zip(en, mats):
        l_bl = np.ascontiguousarray(e, np.int32)
        L.lar_assemble_SM_and_LV_bcrs(mid, 0, 6, P(l_bl), None, P(K), P(np.ascontiguousarray(F)),
                                      None)
    rp, ci, val, rhs = ref_arrays(L, mid, nn)
    orp, oci, oval, orhs = bcrs_assemble(nn, nb_lists, en, mats)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    assert np.array_equal(val, oval) and np.array_equal(rhs, orhs)
    # residual V = b - A x
    x = np.random.default_rng(1).uniform(-1, 1, 4 * nn)
    v = np.zeros(4 * nn)
    L.lar_compute_residual_bcrs(mid, 1, 0, 4 * nn, P(x), None, P(v))
    vo = np.zeros(4 * nn)
    oracle().orc_bcrs_residual(nn, P(orp), P(oci), P(oval), 1, P(x), P(orhs), P(vo))
    assert np.array_equal(v, vo)
    L.lar_free_SM_and_LV_bcrs(mid)


@need_bcrs
def test_assembly_with_tables_equals_table_less():
    L = ref_bcrs()
    mesh, en, nn, nb_lists, mats = vector_problem(2, 3, 2, 7)
    mid, keep = ref_allocate(L, nn, nb_lists)
    tab = np.zeros(6 * 6 + 6, np.int32)
    for e, (K, F) in zip(en, mats):
        l_bl = np.ascontiguousarray(e, np.int32)
        L.lar_fill_assembly_table_int_ent_bcrs(mid, 6, P(l_bl), None, P(tab))
        L.lar_assemble_SM_and_LV_with_table_bcrs(mid, 0, 6, P(tab), P(K), P(np.ascontiguousarray(F)),
                                                 None)
    rp, ci, val, rhs = ref_arrays(L, mid, nn)
    orp, oci, oval, orhs = bcrs_assemble(nn, nb_lists, en, mats)
    assert np.array_equal(val, oval) and np.array_equal(rhs, orhs)
    L.lar_free_SM_and_LV_bcrs(mid)


def test_lu4_factors_and_solves():
    o = oracle()
    rng = np.random.default_rng(3)
    for t in range(200):
        A = rng.uniform(-1, 1, (4, 4))
        if t % 3 == 0:
            A[0, 0] = 0.0  # forces a row interchange
        a = np.ascontiguousarray(A.T.ravel())  # column-major
        piv = np.zeros(4, np.int32)
        assert o.orc_lu4(P(a), P(piv)) == 0
        LU = a.reshape(4, 4).T
        Lm = np.tril(LU, -1) + np.eye(4)
        U = np.triu(LU)
        PA = A.copy()
        for j in range(4):  # the interchanges, in order
            assert piv[j] >= j
            PA[[j, piv[j]]] = PA[[piv[j], j]]
        # partial pivoting: every multiplier is at most 1 in magnitude
        assert np.abs(np.tril(LU, -1)).max() <= 1.0
        assert np.abs(Lm @ U - PA).max() <= 1e-14 * np.abs(A).max() * 4
        b = rng.uniform(-1, 1, 4)
        xb = b.copy()
        o.orc_lu4_solve(P(a), P(piv), P(xb))
        assert np.abs(xb - np.linalg.solve(A, b)).max() <= 1e-11 * max(1.0, np.abs(np.linalg.solve(A, b)).max())


def bj_gs_setup(seed):
    mesh, en, nn, nb_lists, mats = vector_problem(3, 2, 2, seed)
    rp, ci, val, rhs = bcrs_assemble(nn, nb_lists, en, mats)
    # make the diagonal blocks dominant so that both iterations converge
    diag = np.array([next(s for s in range(rp[r], rp[r + 1]) if ci[s] == r) for r in range(nn)], np.int32)
    for r in range(nn):
        val[16 * diag[r]:16 * diag[r] + 16] += (np.eye(4) * 400.0).ravel()
    lu = np.zeros(16 * nn)
    piv = np.zeros(4 * nn, np.int32)
    for r in range(nn):
        lu[16 * r:16 * r + 16] = val[16 * diag[r]:16 * diag[r] + 16]
        blk = np.ascontiguousarray(lu[16 * r:16 * r + 16])
        pv = np.zeros(4, np.int32)
        oracle().orc_lu4(P(blk), P(pv))
        lu[16 * r:16 * r + 16] = blk
        piv[4 * r:4 * r + 4] = pv
    A = np.zeros((4 * nn, 4 * nn))
    for r in range(nn):
        for s in range(rp[r], rp[r + 1]):
            A[4 * r:4 * r + 4, 4 * ci[s]:4 * ci[s] + 4] = val[16 * s:16 * s + 16].reshape(4, 4).T
    return nn, rp, ci, diag, val, lu, piv, rhs, A


@pytest.mark.parametrize("gs", [0, 1])
def test_block_jacobi_and_gauss_seidel_converge_to_the_solution(gs):
    nn, rp, ci, diag, val, lu, piv, rhs, A = bj_gs_setup(11)
    x = np.zeros(4 * nn)
    oracle().orc_bcrs_bj_gs(nn, P(rp), P(ci), P(diag), P(val), P(lu), P(piv), gs, 1, 200, P(rhs), P(x))
    xs = np.linalg.solve(A, rhs)
    assert np.abs(x - xs).max() <= 1e-12 * np.abs(xs).max()
    # one Jacobi sweep from zero is D^-1 b; Gauss-Seidel's first row too
    x1 = np.zeros(4 * nn)
    oracle().orc_bcrs_bj_gs(nn, P(rp), P(ci), P(diag), P(val), P(lu), P(piv), gs, 1, 1, P(rhs), P(x1))
    assert np.abs(x1[:4] - np.linalg.solve(A[:4, :4], rhs[:4])).max() <= 1e-14


def test_sweep_count_rule():
    """Nr_prec odd gives Nr_prec/2 + 1 forward sweeps (las_bcrs_intf.c:2960)"""
    nn, rp, ci, diag, val, lu, piv, rhs, A = bj_gs_setup(13)
    o = oracle()
    x3 = np.zeros(4 * nn)
    o.orc_bcrs_bj_gs(nn, P(rp), P(ci), P(diag), P(val), P(lu), P(piv), 1, 1, 3, P(rhs), P(x3))
    x11 = np.zeros(4 * nn)
    for _ in range(2):
        o.orc_bcrs_bj_gs(nn, P(rp), P(ci), P(diag), P(val), P(lu), P(piv), 1, 1, 1, P(rhs), P(x11))
    assert np.array_equal(x3, x11)


def test_inv4_inverts():
    """orc_inv4 after orc_lu4 (LAPACK dgetrf + dgetri's algorithm): A inv(A)
    = I to rounding; exact for permutation and power-of-two diagonal
    matrices"""
    o = oracle()
    rng = np.random.default_rng(8)
    for t in range(200):
        A = rng.uniform(-1, 1, (4, 4))
        if t % 4 == 0:
            A[0, 0] = 0.0
        a = np.ascontiguousarray(A.T.ravel())
        piv = np.zeros(4, np.int32)
        assert o.orc_lu4(P(a), P(piv)) == 0
        o.orc_inv4(P(a), P(piv))
        Ai = a.reshape(4, 4).T
        cond = np.linalg.cond(A)
        assert np.abs(A @ Ai - np.eye(4)).max() <= 1e-15 * 40 * cond
    for perm in ([1, 0, 3, 2], [3, 2, 1, 0], [2, 0, 3, 1]):
        A = np.eye(4)[perm] * 2.0
        a = np.ascontiguousarray(A.T.ravel())
        piv = np.zeros(4, np.int32)
        o.orc_lu4(P(a), P(piv))
        o.orc_inv4(P(a), P(piv))
        assert np.array_equal(a.reshape(4, 4).T, A.T / 4.0)
    A = np.diag([2.0, -0.5, 8.0, 0.25])
    a = np.ascontiguousarray(A.T.ravel())
    piv = np.zeros(4, np.int32)
    o.orc_lu4(P(a), P(piv))
    o.orc_inv4(P(a), P(piv))
    assert np.array_equal(a.reshape(4, 4).T, np.diag(1.0 / np.diag(A)))


def ilu_dense(nn, rp, ci, diag, ilu):
    """dense unit-lower L and upper U from the factors (U's diagonal blocks
    are the inverses of the stored ones)"""
    Lm, U = np.eye(4 * nn), np.zeros((4 * nn, 4 * nn))
    for r in range(nn):
        for s in range(rp[r], rp[r + 1]):
            c = ci[s]
            blk = ilu[16 * s:16 * s + 16].reshape(4, 4).T
            if c < r:
                Lm[4 * r:4 * r + 4, 4 * c:4 * c + 4] = blk
            elif c > r:
                U[4 * r:4 * r + 4, 4 * c:4 * c + 4] = blk
            else:
                U[4 * r:4 * r + 4, 4 * c:4 * c + 4] = np.linalg.inv(blk)
    return Lm, U


def test_ilu0_reproduces_the_matrix_on_its_pattern():
    """block ILU(0): (L U)_kj = A_kj for every stored block (k, j), within
    rounding; L U differs from A off the pattern (fill is dropped)"""
    nn, rp, ci, diag, val, lu, piv, rhs, A = bj_gs_setup(17)
    ilu = val.copy()
    assert oracle().orc_bcrs_ilu0(nn, P(rp), P(ci), P(diag), P(ilu)) == 0
    Lm, U = ilu_dense(nn, rp, ci, diag, ilu)
    LU = Lm @ U
    mask = np.zeros((4 * nn, 4 * nn), bool)
    for r in range(nn):
        for s in range(rp[r], rp[r + 1]):
            mask[4 * r:4 * r + 4, 4 * ci[s]:4 * ci[s] + 4] = True
    assert np.abs((LU - A)[mask]).max() <= 1e-12 * np.abs(A).max()
    assert np.abs((LU - A)[~mask]).max() > 1e-6  # fill-in exists and is dropped
    # the solves apply (L U)^-1
    b = np.random.default_rng(5).uniform(-1, 1, 4 * nn)
    v = np.zeros(4 * nn)
    oracle().orc_bcrs_rhsub(nn, P(rp), P(ci), P(diag), P(ilu), P(b), P(v))
    ref = np.linalg.solve(LU, b)
    assert np.abs(v - ref).max() <= 1e-12 * np.abs(ref).max()


def test_ilu0_of_a_block_tridiagonal_matrix_is_exact():
    """a chain of blocks has no fill: ILU(0) is the exact block LU and one
    forward/backward solve gives A^-1 b"""
    nn = 40
    rng = np.random.default_rng(6)
    nb_lists = [[j for j in (i - 1, i + 1) if 0 <= j < nn] for i in range(nn)]
    ptr = np.zeros(nn + 1, np.int32)
    ptr[1:] = np.cumsum([len(x) for x in nb_lists])
    ind = np.array([v for x in nb_lists for v in x], np.int32)
    o = oracle()
    rp = np.zeros(nn + 1, np.int32)
    nsm = o.orc_bcrs_pattern(nn, P(ptr), P(ind), P(rp), None)
    ci = np.zeros(nsm, np.int32)
    o.orc_bcrs_pattern(nn, P(ptr), P(ind), P(rp), P(ci))
    val = rng.uniform(-1, 1, 16 * nsm)
    diag = np.array([next(s for s in range(rp[r], rp[r + 1]) if ci[s] == r) for r in range(nn)], np.int32)
    for r in range(nn):
        val[16 * diag[r]:16 * diag[r] + 16] += (np.eye(4) * 5.0).ravel()
    A = np.zeros((4 * nn, 4 * nn))
    for r in range(nn):
        for s in range(rp[r], rp[r + 1]):
            A[4 * r:4 * r + 4, 4 * ci[s]:4 * ci[s] + 4] = val[16 * s:16 * s + 16].reshape(4, 4).T
    ilu = val.copy()
    assert o.orc_bcrs_ilu0(nn, P(rp), P(ci), P(diag), P(ilu)) == 0
    b = rng.uniform(-1, 1, 4 * nn)
    v = np.zeros(4 * nn)
    o.orc_bcrs_rhsub(nn, P(rp), P(ci), P(diag), P(ilu), P(b), P(v))
    x = np.linalg.solve(A, b)
    assert np.abs(v - x).max() <= 1e-13 * np.abs(x).max()


# ------------------------------------------------------------ block ILU(k)
def graph_within(nn, rp, ci, hops):
    """rows/columns within the given number of hops in the block graph"""
    G = np.zeros((nn, nn), bool)
    for r in range(nn):
        G[r, ci[rp[r]:rp[r + 1]]] = True
    R = np.eye(nn, dtype=bool)
    for _ in range(hops):
        R = (R.astype(np.int64) @ G.astype(np.int64)) > 0
    return R


@need_bcrs
@pytest.mark.parametrize("k", [1, 2])
def test_iluk_pattern_bitexact_with_reference(k):
    """the grown ILU(k) pattern equals the one the reference's own
    lar_allocate_preconditioner_bcrs builds (compiled from its sources; the
    allocation calls no LAPACK), and it is the set of blocks within k+1 hops
    of the block graph"""
    from oracle_lib import bcrs_iluk
    L = ref_bcrs()
    L.lar_allocate_preconditioner_bcrs.restype = ctypes.c_int
    L.lar_allocate_preconditioner_bcrs.argtypes = [ctypes.c_int] * 3
    mesh, en, nn, nb_lists, mats = vector_problem(3, 3, 2, 5)
    mid, keep = ref_allocate(L, nn, nb_lists)
    assert L.lar_allocate_preconditioner_bcrs(mid, 5, k) >= 0
    st = (BcrsStruct * 20).in_dll(L, "itv_bcrs_matrices")[mid]
    rrp = np.ctypeslib.as_array(st.ilu_row_ptr, (nn + 1,)).copy()
    rci = np.ctypeslib.as_array(st.ilu_col_ind, (st.ILU_nr_sm_blocks,)).copy()
    rp, ci, val, rhs = bcrs_assemble(nn, nb_lists, en, mats)
    orp, oci, diag, ilu = bcrs_iluk(nn, rp, ci, val, k)
    assert np.array_equal(orp, rrp) and np.array_equal(oci, rci)
    R = graph_within(nn, rp, ci, k + 1)
    for r in range(nn):
        assert np.array_equal(oci[orp[r]:orp[r + 1]], np.flatnonzero(R[r]))
    assert len(oci) > len(ci)
    mesh.free()


@pytest.mark.parametrize("k", [1, 2])
def test_iluk_reproduces_the_matrix_on_its_pattern(k):
    """block ILU(k): (L U)_ij = A_ij on every block of the grown pattern
    (A_ij = 0 at fill); more fill brings L U closer to A off the pattern"""
    from oracle_lib import bcrs_iluk
    nn, rp, ci, diag, val, lu, piv, rhs, A = bj_gs_setup(17)
    orp, oci, odiag, ilu = bcrs_iluk(nn, rp, ci, val, k)
    Lm, U = ilu_dense(nn, orp, oci, odiag, ilu)
    LU = Lm @ U
    mask = np.zeros((4 * nn, 4 * nn), bool)
    for r in range(nn):
        for s in range(orp[r], orp[r + 1]):
            mask[4 * r:4 * r + 4, 4 * oci[s]:4 * oci[s] + 4] = True
    assert np.abs((LU - A)[mask]).max() <= 1e-12 * np.abs(A).max()
    ilu0 = val.copy()
    assert oracle().orc_bcrs_ilu0(nn, P(rp), P(ci), P(diag), P(ilu0)) == 0
    L0, U0 = ilu_dense(nn, rp, ci, diag, ilu0)
    assert np.abs(LU - A).max() < np.abs(L0 @ U0 - A).max()


def test_iluk_with_a_full_pattern_is_exact():
    """when k+1 hops reach every block the ILU(k) factors are the exact block
    LU, and one solve gives A^-1 b"""
    from oracle_lib import bcrs_iluk
    nn, rp, ci, diag, val, lu, piv, rhs, A = bj_gs_setup(19)
    k = 12
    orp, oci, odiag, ilu = bcrs_iluk(nn, rp, ci, val, k)
    assert orp[-1] == nn * nn
    b = np.random.default_rng(8).uniform(-1, 1, 4 * nn)
    v = np.zeros(4 * nn)
    oracle().orc_bcrs_rhsub(nn, P(orp), P(oci), P(odiag), P(ilu), P(b), P(v))
    x = np.linalg.solve(A, b)
    assert np.abs(v - x).max() <= 1e-12 * np.abs(x).max()


def test_iluk_list_capacity_is_enforced():
    """a row needing more columns than R's list holds is an error (R exits)"""
    nn, rp, ci, *_ = bj_gs_setup(17)
    rp2 = np.zeros(nn + 1, np.int32)
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    assert oracle().orc_bcrs_iluk_pattern(nn, P(rp), P(ci), 1, 8, P(rp2), None) == -1
