"""The storage seam with the reference's names (include/modfem_lar.h) on the
device, against the oracle (itself bit-exact with the reference's
las_crs_intf.c, tests/test_oracle_ref.py): pattern, tables, assembled values,
Jacobi, residual and Jacobi sweep are bit-exact."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import OracleProblem, P, oracle  # noqa: E402

pytestmark = pytest.mark.gpu


def test_lar_seam_bitexact():
    mf.init()
    L = mf.lib()
    m = mf.Mesh.box_prism(6, 5, 4, seed=17, jitter=0.2)
    op = OracleProblem(m.arrays())
    lists = []
    for r in range(op.n):
        cols = op.col_ind[op.row_ptr[r]: op.row_ptr[r + 1]]
        lists.append(np.ascontiguousarray(cols[cols != r], np.int32))
    nro = np.array([len(x) for x in lists], np.int32)
    ptrs = (ctypes.POINTER(ctypes.c_int) * op.n)(*[x.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
                                                  for x in lists])
    ones = np.ones(op.n, np.int32)
    pos = np.arange(op.n, dtype=np.int32)
    mid = L.lar_allocate_SM_and_LV(2, 1, op.n, 36, op.n, 1, P(ones), P(pos), P(nro),
                                   ctypes.cast(ptrs, ctypes.c_void_p))
    assert mid >= 0, L.mf_last_error()
    rp, ci = np.zeros(op.n + 1, np.int32), np.zeros(op.nnz, np.int32)
    assert L.mf_lar_get_system(mid, P(rp), P(ci), None, None) == 0
    assert np.array_equal(rp, op.row_ptr) and np.array_equal(ci, op.col_ind)
    rng = np.random.default_rng(3)
    kappa = np.exp(rng.normal(size=op.nel))
    coef = np.zeros((op.nel, 21))
    coef[:, 0] = coef[:, 4] = coef[:, 8] = kappa
    coef[:, 19] = 1.0
    val0, rhs0 = op.assemble(mf.IND_DIFF, coef)
    assert L.lar_initialize_SM_and_LV(mid, 0) == 0
    for ie in range(op.nel + op.nbf):
        e = ie if ie < op.nel else op.bfe[ie - op.nel]
        K, F = op.element(ie, mf.IND_DIFF, coef[ie]) if ie < op.nel else op.face(ie - op.nel)
        L_bl = np.ascontiguousarray(op.node_dof[op.en[e, :6]], np.int32)
        tab = np.zeros(42, np.int32)
        assert L.lar_fill_assembly_table_int_ent(mid, 6, P(L_bl), P(np.ones(6, np.int32)), P(tab)) >= 0
        assert np.array_equal(tab, op.table[op.table_ptr[ie]: op.table_ptr[ie + 1]])
        sm = np.ascontiguousarray(K.reshape(-1))
        L.lar_assemble_SM_and_LV_with_table(mid, 0, 6, P(tab), P(sm), P(np.ascontiguousarray(F)), None)
    val, rhs = np.zeros(op.nnz), np.zeros(op.n)
    assert L.mf_lar_get_system(mid, None, None, P(val), P(rhs)) == 0
    assert np.array_equal(val, val0) and np.array_equal(rhs, rhs0)
    assert L.lar_allocate_preconditioner(mid, 1, 0) > 0
    assert L.lar_fill_preconditioner(mid) == 0
    o = oracle()
    dptr, dinv = np.zeros(op.n, np.int32), np.zeros(op.n)
    o.orc_fill_jacobi(op.n, P(op.row_ptr), P(op.col_ind), P(val0), P(dptr), P(dinv))
    x = rng.uniform(-1, 1, op.n)
    v, v0 = np.zeros(op.n), np.zeros(op.n)
    L.lar_compute_residual(mid, 1, 0, op.n, P(x), None, P(v))
    o.orc_compute_residual(op.n, P(op.row_ptr), P(op.col_ind), P(val0), 1, P(rhs0), P(x), P(v0))
    assert np.array_equal(v, v0)
    w, w0 = x.copy(), x.copy()
    L.lar_perform_BJ_or_GS_iterations(mid, 1, 0, 1, op.n, P(w), None)
    o.orc_bj_iteration(op.n, P(op.row_ptr), P(op.col_ind), P(val0), P(dptr), P(dinv), 1, P(rhs0),
                       P(w0))
    assert np.array_equal(w, w0)
    assert L.lar_allocate_preconditioner(mid, 5, 0) < 0  # ILU needs block storage
    assert L.lar_free_preconditioner(mid) == 0
    assert L.lar_free_SM_and_LV(mid) == 0


def lar_crs_matrix(L, op, coef):
    """CRS matrix through the seam: R's allocation from neighbour lists,
    per-entity tables and assembly of the oracle's entity matrices"""
    lists = []
    for r in range(op.n):
        cols = op.col_ind[op.row_ptr[r]: op.row_ptr[r + 1]]
        lists.append(np.ascontiguousarray(cols[cols != r], np.int32))
    nro = np.array([len(x) for x in lists], np.int32)
    ptrs = (ctypes.POINTER(ctypes.c_int) * op.n)(*[x.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
                                                  for x in lists])
    ones = np.ones(op.n, np.int32)
    pos = np.arange(op.n, dtype=np.int32)
    mid = L.lar_allocate_SM_and_LV(2, 1, op.n, 36, op.n, 1, P(ones), P(pos), P(nro),
                                   ctypes.cast(ptrs, ctypes.c_void_p))
    assert mid >= 0, L.mf_last_error()
    assert L.lar_initialize_SM_and_LV(mid, 0) == 0
    for ie in range(op.nel + op.nbf):
        e = ie if ie < op.nel else op.bfe[ie - op.nel]
        K, F = op.element(ie, mf.IND_DIFF, coef[ie]) if ie < op.nel else op.face(ie - op.nel)
        L_bl = np.ascontiguousarray(op.node_dof[op.en[e, :6]], np.int32)
        tab = np.zeros(42, np.int32)
        assert L.lar_fill_assembly_table_int_ent(mid, 6, P(L_bl), P(np.ones(6, np.int32)), P(tab)) >= 0
        sm = np.ascontiguousarray(K.reshape(-1))
        L.lar_assemble_SM_and_LV_with_table(mid, 0, 6, P(tab), P(sm), P(np.ascontiguousarray(F)), None)
    return mid


def diffusion_coef(op, seed):
    rng = np.random.default_rng(seed)
    kappa = np.exp(rng.normal(size=op.nel))
    coef = np.zeros((op.nel, 21))
    coef[:, 0] = coef[:, 4] = coef[:, 8] = kappa
    coef[:, 1] = coef[:, 3] = 0.1 * kappa
    coef[:, 19] = 1.0
    return coef


@pytest.mark.parametrize("dims", [(7, 6, 5), (66, 66, 1)])
def test_lar_crs_gauss_seidel_chunks_bitexact(dims):
    """CRS BLOCK_GS sweeps with the row chunks of T OpenMP threads equal the
    oracle's chunked restatement (itself bitwise R's compiled sweep on 1-4
    threads) bit for bit; the 66x66x1 box has a half bandwidth of ~4500 rows,
    beyond the kernel's LDS ring of 4096 new values, so its early rows are
    read back from device memory"""
    print("chunked GS", dims, flush=True)
    mf.init()
    L = mf.lib()
    m = mf.Mesh.box_prism(*dims, seed=23, jitter=0.2)
    op = OracleProblem(m.arrays())
    coef = diffusion_coef(op, 5)
    val0, rhs0 = op.assemble(mf.IND_DIFF, coef)
    mid = lar_crs_matrix(L, op, coef)
    assert L.lar_allocate_preconditioner(mid, 2, 0) > 0, L.mf_last_error()
    assert L.lar_fill_preconditioner(mid) == 0
    o = oracle()
    dptr, dinv = np.zeros(op.n, np.int32), np.zeros(op.n)
    o.orc_fill_jacobi(op.n, P(op.row_ptr), P(op.col_ind), P(val0), P(dptr), P(dinv))
    if dims[2] == 1:
        band = max(abs(int(c) - r) for r in range(op.n)
                   for c in op.col_ind[op.row_ptr[r]: op.row_ptr[r + 1]])
        assert band > 4096
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, op.n)
    b = rng.uniform(-1, 1, op.n)
    for T in (1, 2, 3, 7, 64, op.n):
        assert L.mf_lar_set_gs_threads(mid, T) == 0
        for use_rhs, B, sweeps in ((1, None, 2), (0, None, 1), (1, b, 1)):
            w, w0 = x.copy(), x.copy()
            L.lar_perform_BJ_or_GS_iterations(mid, use_rhs, 0, sweeps, op.n, P(w),
                                              None if B is None else P(B))
            rb = rhs0 if B is None else B
            o.orc_crs_gs_chunked(op.n, P(op.row_ptr), P(op.col_ind), P(val0), P(dptr), P(dinv),
                                 use_rhs, P(rb), T, sweeps, P(w0))
            assert np.array_equal(w, w0), (T, use_rhs, sweeps)
    assert L.mf_lar_set_gs_threads(mid, 0) < 0
    assert L.lar_free_preconditioner(mid) == 0
    assert L.lar_free_SM_and_LV(mid) == 0


def test_lar_crs_gauss_seidel_gmres_reaches_the_solution():
    """GMRES(30) over the CRS matrix with two chunked GS sweeps per
    application (as the reference's compreres smooths) reaches the direct
    solution for 1 and 4 threads; CRS GMRES with block Jacobi stays with
    mf_gmres (refused here)"""
    mf.init()
    L = mf.lib()
    m = mf.Mesh.box_prism(8, 7, 6, seed=3, jitter=0.2)
    op = OracleProblem(m.arrays())
    coef = diffusion_coef(op, 8)
    val0, rhs0 = op.assemble(mf.IND_DIFF, coef)
    A = np.zeros((op.n, op.n))
    for r in range(op.n):
        A[r, op.col_ind[op.row_ptr[r]: op.row_ptr[r + 1]]] = val0[op.row_ptr[r]: op.row_ptr[r + 1]]
    xs = np.linalg.solve(A, rhs0)
    mid = lar_crs_matrix(L, op, coef)
    its = {}
    for T in (1, 4):
        assert L.lar_allocate_preconditioner(mid, 2, 0) > 0
        assert L.lar_fill_preconditioner(mid) == 0
        assert L.mf_lar_set_gs_threads(mid, T) == 0
        x = np.zeros(op.n)
        it, rr, rs = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        rc = L.mf_lar_gmres(mid, 30, 2000, 1e-12, 2, 1, P(x), ctypes.byref(it), ctypes.byref(rr),
                            ctypes.byref(rs))
        assert rc == 1, (T, rc, L.mf_last_error())
        assert np.abs(x - xs).max() <= 1e-8 * np.abs(xs).max(), T
        its[T] = it.value
        assert L.lar_free_preconditioner(mid) == 0
    assert L.lar_allocate_preconditioner(mid, 1, 0) > 0
    assert L.lar_fill_preconditioner(mid) == 0
    assert L.mf_lar_gmres(mid, 30, 2000, 1e-12, 1, 1, P(np.zeros(op.n)), None, None, None) < 0
    assert L.lar_free_preconditioner(mid) == 0
    assert L.lar_free_SM_and_LV(mid) == 0
    print("GMRES(30) + GS(2) iterations by thread count", its, flush=True)
