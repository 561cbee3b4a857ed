"""BCRS storage with 4x4 blocks through the storage seam (include/modfem_lar.h)
on the device, against the oracle (itself bit-exact with the reference's
lsd_mkb/lad_bcrs/las_bcrs_intf.c for pattern, assembly and residual, and
pinned by the ILU(0) pattern identity for the factors,
tests/test_oracle_bcrs.py): pattern, assembly with and without tables,
residual, diagonal-block LU factors, block Jacobi and block Gauss-Seidel
sweeps, block ILU(0) and ILU(k) factors and solves are all bit-exact (the
ILU(k) pattern also against the reference's own pattern growth, through the
oracle)."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import P, bcrs_assemble, bcrs_diag_lu, bcrs_vector_problem, oracle  # noqa: E402

pytestmark = pytest.mark.gpu

LAC_STORAGE_BCRS = 3
BLOCK_JACOBI, BLOCK_GS, MULTI_ILU, BLOCK_ILU = 1, 2, 4, 5


def allocate(L, nn, nb_lists, storage=LAC_STORAGE_BCRS, block_size=4):
    arrs = [np.ascontiguousarray(x, np.int32) for x in nb_lists]
    ptrs = (ctypes.c_void_p * nn)(*[P(x) for x in arrs])
    nro = np.array([len(x) for x in nb_lists], np.int32)
    mid = L.lar_allocate_SM_and_LV(storage, 1, 4 * nn, 24, nn, block_size, None, None, P(nro),
                                   ctypes.cast(ptrs, ctypes.c_void_p))
    return mid


def build(nx, ny, nz, seed, tables, diag_shift=6.0):
    """allocate and assemble a BCRS matrix through the seam; returns the
    matrix id and the oracle's arrays"""
    mf.init()
    L = mf.lib()
    a = mf.Mesh.box_prism(nx, ny, nz, seed=seed, jitter=0.2).arrays()
    en, nn, nb_lists, mats = bcrs_vector_problem(a, seed, diag_shift)
    mid = allocate(L, nn, nb_lists)
    assert mid >= 0, L.mf_last_error()
    assert L.lar_initialize_SM_and_LV(mid, 0) == 0
    tab = np.zeros(6 * 6 + 6, np.int32)
    for e, (K, F) in zip(en, mats):
        l_bl = np.ascontiguousarray(e, np.int32)
        Fc = np.ascontiguousarray(F)
        if tables:
            assert L.lar_fill_assembly_table_int_ent(mid, 6, P(l_bl), None, P(tab)) == 1
            assert L.lar_assemble_SM_and_LV_with_table(mid, 0, 6, P(tab), P(K), P(Fc), None) == 1
        else:
            assert L.lar_assemble_SM_and_LV(mid, 0, 6, P(l_bl), None, P(K), P(Fc), None) == 1
    rp, ci, val, rhs = bcrs_assemble(nn, nb_lists, en, mats)
    return L, mid, nn, rp, ci, val, rhs


def free(L, mid):
    assert L.lar_free_preconditioner(mid) == 0
    assert L.lar_free_SM_and_LV(mid) == 0


@pytest.mark.parametrize("tables", [True, False])
def test_pattern_assembly_and_residual_bitexact(tables):
    L, mid, nn, rp, ci, val, rhs = build(4, 3, 3, 5, tables)
    st, bs, nb, ns = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
    assert L.mf_lar_get_sizes(mid, ctypes.byref(st), ctypes.byref(bs), ctypes.byref(nb),
                              ctypes.byref(ns)) == 0
    assert (st.value, bs.value, nb.value, ns.value) == (LAC_STORAGE_BCRS, 4, nn, len(ci))
    grp, gci = np.zeros(nn + 1, np.int32), np.zeros(len(ci), np.int32)
    gval, grhs = np.zeros(16 * len(ci)), np.zeros(4 * nn)
    assert L.mf_lar_get_system(mid, P(grp), P(gci), P(gval), P(grhs)) == 0
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert np.array_equal(gval, val) and np.array_equal(grhs, rhs)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, 4 * nn)
    b = rng.uniform(-1, 1, 4 * nn)
    o = oracle()
    for use_rhs, bb, ini_zero in [(1, None, 0), (1, b, 0), (0, None, 0), (1, None, 1)]:
        v, v0 = np.zeros(4 * nn), np.zeros(4 * nn)
        L.lar_compute_residual(mid, use_rhs, ini_zero, 4 * nn, P(x), P(bb) if bb is not None else None,
                               P(v))
        o.orc_bcrs_residual(nn, P(rp), P(ci), P(val), use_rhs, None if ini_zero else P(x),
                            P(bb if bb is not None else rhs), P(v0))
        assert np.array_equal(v, v0), (use_rhs, ini_zero)
    assert L.lar_free_SM_and_LV(mid) == 0


@pytest.mark.parametrize("precon", [BLOCK_JACOBI, BLOCK_GS])
def test_block_lu_and_sweeps_bitexact(precon):
    # no diagonal shift: the diagonal blocks need row interchanges
    L, mid, nn, rp, ci, val, rhs = build(5, 4, 3, 9, True, diag_shift=0.0)
    assert L.lar_allocate_preconditioner(mid, precon, 0) == nn, L.mf_last_error()
    assert L.lar_fill_preconditioner(mid) == 0, L.mf_last_error()
    diag, lu0, piv0 = bcrs_diag_lu(nn, rp, ci, val)
    lu, piv = np.zeros(16 * nn), np.zeros(4 * nn, np.int32)
    assert L.mf_lar_get_block_precon(mid, P(lu), P(piv)) == 0
    assert np.array_equal(piv, piv0) and np.array_equal(lu, lu0)
    assert (piv != np.tile(np.arange(4), nn)).any()  # row interchanges occur
    o = oracle()
    rng = np.random.default_rng(4)
    v_in = rng.uniform(-1, 1, 4 * nn)
    b = rng.uniform(-1, 1, 4 * nn)
    gs = int(precon == BLOCK_GS)
    for nr_prec, use_rhs, bb in [(1, 1, None), (2, 1, None), (3, 1, b), (4, 0, None), (5, 1, b)]:
        v, v0 = v_in.copy(), v_in.copy()
        L.lar_perform_BJ_or_GS_iterations(mid, use_rhs, 0, nr_prec, 4 * nn, P(v),
                                          P(bb) if bb is not None else None)
        o.orc_bcrs_bj_gs(nn, P(rp), P(ci), P(diag), P(val), P(lu0), P(piv0), gs, use_rhs, nr_prec,
                         P(bb if bb is not None else rhs), P(v0))
        assert np.array_equal(v, v0), (nr_prec, use_rhs)
    free(L, mid)


def test_gauss_seidel_many_levels_bitexact_and_converges():
    """a larger mesh (several hundred levels per sweep) with dominant
    diagonal blocks: bitwise the oracle, and converging to the solution"""
    L, mid, nn, rp, ci, val, rhs = build(10, 9, 6, 21, True, diag_shift=40.0)
    assert L.lar_allocate_preconditioner(mid, BLOCK_GS, 0) == nn
    assert L.lar_fill_preconditioner(mid) == 0
    diag, lu0, piv0 = bcrs_diag_lu(nn, rp, ci, val)
    v, v0 = np.zeros(4 * nn), np.zeros(4 * nn)
    L.lar_perform_BJ_or_GS_iterations(mid, 1, 1, 40, 4 * nn, P(v), None)
    oracle().orc_bcrs_bj_gs(nn, P(rp), P(ci), P(diag), P(val), P(lu0), P(piv0), 1, 1, 40, P(rhs),
                            P(v0))
    assert np.array_equal(v, v0)
    r = np.zeros(4 * nn)
    L.lar_compute_residual(mid, 1, 0, 4 * nn, P(v), None, P(r))
    assert np.abs(r).max() <= 1e-10 * np.abs(rhs).max()
    free(L, mid)


def test_errors():
    mf.init()
    L = mf.lib()
    a = mf.Mesh.box_prism(2, 2, 1, seed=1, jitter=0.0).arrays()
    en, nn, nb_lists, mats = bcrs_vector_problem(a, 1)
    # BCRS needs block size 4
    assert allocate(L, nn, nb_lists, block_size=3) < 0
    # block size 4 with undefined storage picks BCRS
    mid = allocate(L, nn, nb_lists, storage=-1)
    assert mid >= 0
    st = ctypes.c_int()
    assert L.mf_lar_get_sizes(mid, ctypes.byref(st), None, None, None) == 0 and st.value == 3
    # ILU(k) with k < 0 and additive Schwarz are not offered
    assert L.lar_allocate_preconditioner(mid, 5, -1) < 0
    assert L.lar_allocate_preconditioner(mid, 3, 0) < 0
    # all-zero matrix: singular diagonal blocks are reported
    assert L.lar_allocate_preconditioner(mid, BLOCK_JACOBI, 0) == nn
    assert L.lar_fill_preconditioner(mid) < 0
    assert b"singular" in L.mf_last_error()
    # a block pair outside the pattern
    tab = np.zeros(6, np.int32)
    far = np.array([0, nn - 1], np.int32)
    assert nn - 1 not in nb_lists[0]
    assert L.lar_fill_assembly_table_int_ent(mid, 2, P(far), None, P(tab)) < 0
    free(L, mid)


@pytest.mark.parametrize("shape,precon", [((5, 4, 3), MULTI_ILU), ((10, 9, 6), BLOCK_ILU)])
def test_block_ilu0_factors_and_solves_bitexact(shape, precon):
    L, mid, nn, rp, ci, val, rhs = build(*shape, 13, True)
    assert L.lar_allocate_preconditioner(mid, precon, 0) == nn, L.mf_last_error()
    assert L.lar_fill_preconditioner(mid) == 0, L.mf_last_error()
    diag = np.array([next(s for s in range(rp[r], rp[r + 1]) if ci[s] == r) for r in range(nn)],
                    np.int32)
    ilu0 = val.copy()
    o = oracle()
    assert o.orc_bcrs_ilu0(nn, P(rp), P(ci), P(diag), P(ilu0)) == 0
    ilu = np.zeros(16 * len(ci))
    assert L.mf_lar_get_block_precon(mid, P(ilu), None) == 0
    assert np.array_equal(ilu, ilu0)
    b = np.random.default_rng(3).uniform(-1, 1, 4 * nn)
    for bb in (b, None):
        v, v0 = np.full(4 * nn, 7.0), np.zeros(4 * nn)
        L.lar_perform_rhsub(mid, 4 * nn, P(v), P(bb) if bb is not None else None)
        o.orc_bcrs_rhsub(nn, P(rp), P(ci), P(diag), P(ilu0), P(bb if bb is not None else rhs), P(v0))
        assert np.array_equal(v, v0)
    # V and B may be the same array
    w = b.copy()
    L.lar_perform_rhsub(mid, 4 * nn, P(w), P(w))
    o.orc_bcrs_rhsub(nn, P(rp), P(ci), P(diag), P(ilu0), P(b), P(v0))
    assert np.array_equal(w, v0)
    free(L, mid)


@pytest.mark.parametrize("precon,nr_prec", [(0, 1), (BLOCK_JACOBI, 1), (BLOCK_JACOBI, 2),
                                            (BLOCK_GS, 1), (BLOCK_GS, 2), (MULTI_ILU, 1),
                                            (BLOCK_ILU, 1)])
def test_gmres_on_bcrs_solves_the_system(precon, nr_prec):
    """GMRES(30) over the storage operations reaches the dense solution to
    the solver tolerance; preconditioning cuts the iteration count"""
    L, mid, nn, rp, ci, val, rhs = build(6, 5, 4, 3, True, diag_shift=2.0)
    A = np.zeros((4 * nn, 4 * nn))
    for r in range(nn):
        for s in range(rp[r], rp[r + 1]):
            A[4 * r:4 * r + 4, 4 * ci[s]:4 * ci[s] + 4] = val[16 * s:16 * s + 16].reshape(4, 4).T
    xs = np.linalg.solve(A, rhs)
    if precon:
        assert L.lar_allocate_preconditioner(mid, precon, 0) == nn
        assert L.lar_fill_preconditioner(mid) == 0
    x = np.zeros(4 * nn)
    it, rr, res = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    assert L.mf_lar_gmres(mid, 30, 2000, 1e-10, nr_prec, 1, P(x), ctypes.byref(it), ctypes.byref(rr),
                          ctypes.byref(res)) == 1, L.mf_last_error()
    assert rr.value <= 1e-10
    assert np.abs(x - xs).max() <= 1e-10 * 120 * np.abs(xs).max()  # tol x cond(A)
    if precon in (BLOCK_GS, MULTI_ILU, BLOCK_ILU):
        assert it.value < 60, it.value
    # a given initial guess is used (ini_zero = 0)
    x = 0.5 * xs
    assert L.mf_lar_gmres(mid, 30, 2000, 1e-10, nr_prec, 0, P(x), ctypes.byref(it), ctypes.byref(rr),
                          ctypes.byref(res)) == 1
    assert np.abs(x - xs).max() <= 1e-10 * 120 * np.abs(xs).max()
    if precon:
        free(L, mid)
    else:
        assert L.lar_free_SM_and_LV(mid) == 0


def greedy_colors(nn, rp, ci):
    color = -np.ones(nn, int)
    for r in range(nn):
        used = {color[c] for c in ci[rp[r]:rp[r + 1]] if color[c] >= 0}
        c = 0
        while c in used:
            c += 1
        color[r] = c
    return color


def test_colour_ordered_gauss_seidel_is_the_permuted_sweep():
    """colour-ordered block GS equals the oracle's sequential block GS on the
    matrix renumbered colour by colour (summation order aside)"""
    L, mid, nn, rp, ci, val, rhs = build(5, 4, 3, 29, True, diag_shift=6.0)
    assert L.lar_allocate_preconditioner(mid, BLOCK_GS, 0) == nn
    assert L.lar_fill_preconditioner(mid) == 0
    color = greedy_colors(nn, rp, ci)
    assert L.mf_lar_set_gs_order(mid, 1) == color.max() + 1
    order = np.argsort(color, kind="stable")  # new position -> old row
    newpos = np.empty(nn, int)
    newpos[order] = np.arange(nn)
    # renumbered pattern and blocks
    prp = np.zeros(nn + 1, np.int32)
    pci, pval = [], []
    for rn in range(nn):
        r = order[rn]
        ent = sorted((newpos[ci[s]], s) for s in range(rp[r], rp[r + 1]))
        for c, s in ent:
            pci.append(c)
            pval.append(val[16 * s:16 * s + 16])
        prp[rn + 1] = len(pci)
    pci = np.array(pci, np.int32)
    pval = np.concatenate(pval)
    pdiag, plu, ppiv = bcrs_diag_lu(nn, prp, pci, pval)
    rng = np.random.default_rng(8)
    v_in = rng.uniform(-1, 1, 4 * nn)
    perm4 = (4 * order[:, None] + np.arange(4)).ravel()
    for nr_prec in (1, 2, 3):
        v = v_in.copy()
        L.lar_perform_BJ_or_GS_iterations(mid, 1, 0, nr_prec, 4 * nn, P(v), None)
        vp = np.ascontiguousarray(v_in[perm4])
        oracle().orc_bcrs_bj_gs(nn, P(prp), P(pci), P(pdiag), P(pval), P(plu), P(ppiv), 1, 1, nr_prec,
                                P(np.ascontiguousarray(rhs[perm4])), P(vp))
        assert np.abs(v[perm4] - vp).max() <= 1e-13 * np.abs(vp).max(), nr_prec
    # back to the reference order: bitwise the oracle again
    assert L.mf_lar_set_gs_order(mid, 0) == 1
    diag, lu0, piv0 = bcrs_diag_lu(nn, rp, ci, val)
    v, v0 = v_in.copy(), v_in.copy()
    L.lar_perform_BJ_or_GS_iterations(mid, 1, 0, 2, 4 * nn, P(v), None)
    oracle().orc_bcrs_bj_gs(nn, P(rp), P(ci), P(diag), P(val), P(lu0), P(piv0), 1, 1, 2, P(rhs), P(v0))
    assert np.array_equal(v, v0)
    free(L, mid)


def test_gmres_with_colour_ordered_gauss_seidel():
    L, mid, nn, rp, ci, val, rhs = build(6, 5, 4, 3, True, diag_shift=2.0)
    A = np.zeros((4 * nn, 4 * nn))
    for r in range(nn):
        for s in range(rp[r], rp[r + 1]):
            A[4 * r:4 * r + 4, 4 * ci[s]:4 * ci[s] + 4] = val[16 * s:16 * s + 16].reshape(4, 4).T
    xs = np.linalg.solve(A, rhs)
    assert L.lar_allocate_preconditioner(mid, BLOCK_GS, 0) == nn
    assert L.lar_fill_preconditioner(mid) == 0
    assert L.mf_lar_set_gs_order(mid, 1) > 1
    x = np.zeros(4 * nn)
    it, rr, res = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    assert L.mf_lar_gmres(mid, 30, 2000, 1e-10, 2, 1, P(x), ctypes.byref(it), ctypes.byref(rr),
                          ctypes.byref(res)) == 1
    assert np.abs(x - xs).max() <= 1e-10 * 120 * np.abs(xs).max()
    assert it.value < 60
    free(L, mid)


@pytest.mark.parametrize("k,precon", [(1, BLOCK_ILU), (2, MULTI_ILU)])
def test_block_iluk_pattern_factors_and_solves_bitexact(k, precon):
    """block ILU(k), k > 0: the grown pattern equals the oracle's (itself
    bit-exact with the reference's lar_allocate_preconditioner_bcrs), the
    factors on it and the solves are bit-exact"""
    from oracle_lib import bcrs_iluk
    L, mid, nn, rp, ci, val, rhs = build(5, 4, 3, 21, True)
    assert L.lar_allocate_preconditioner(mid, precon, k) == nn, L.mf_last_error()
    assert L.lar_fill_preconditioner(mid) == 0, L.mf_last_error()
    orp, oci, odiag, oilu = bcrs_iluk(nn, rp, ci, val, k)
    prp = np.zeros(nn + 1, np.int32)
    ns = ctypes.c_longlong()
    assert L.mf_lar_get_ilu_pattern(mid, P(prp), None, ctypes.byref(ns)) == k
    pci = np.zeros(ns.value, np.int32)
    assert L.mf_lar_get_ilu_pattern(mid, None, P(pci), None) == k
    assert np.array_equal(prp, orp) and np.array_equal(pci, oci)
    assert len(pci) > len(ci)
    ilu = np.zeros(16 * len(pci))
    assert L.mf_lar_get_block_precon(mid, P(ilu), None) == 0
    assert np.array_equal(ilu, oilu)
    o = oracle()
    b = np.random.default_rng(4).uniform(-1, 1, 4 * nn)
    v, v0 = np.full(4 * nn, 3.0), np.zeros(4 * nn)
    L.lar_perform_rhsub(mid, 4 * nn, P(v), P(b))
    o.orc_bcrs_rhsub(nn, P(orp), P(oci), P(odiag), P(oilu), P(b), P(v0))
    assert np.array_equal(v, v0)
    # a second fill gives the same factors; k = 0 goes back to the matrix pattern
    assert L.lar_fill_preconditioner(mid) == 0
    ilu2 = np.zeros_like(ilu)
    assert L.mf_lar_get_block_precon(mid, P(ilu2), None) == 0
    assert np.array_equal(ilu2, oilu)
    assert L.lar_allocate_preconditioner(mid, precon, 0) == nn
    assert L.mf_lar_get_ilu_pattern(mid, P(prp), None, ctypes.byref(ns)) == 0
    assert ns.value == len(ci)
    free(L, mid)


def test_gmres_with_block_iluk():
    """GMRES(30) with block ILU(1) and ILU(2) reaches the dense solution in
    fewer iterations than with ILU(0); a too-small reference list is an error"""
    L, mid, nn, rp, ci, val, rhs = build(6, 5, 4, 3, True, diag_shift=2.0)
    A = np.zeros((4 * nn, 4 * nn))
    for r in range(nn):
        for s in range(rp[r], rp[r + 1]):
            A[4 * r:4 * r + 4, 4 * ci[s]:4 * ci[s] + 4] = val[16 * s:16 * s + 16].reshape(4, 4).T
    xs = np.linalg.solve(A, rhs)
    its = []
    for k in (0, 1, 2):
        assert L.lar_allocate_preconditioner(mid, BLOCK_ILU, k) == nn, L.mf_last_error()
        assert L.lar_fill_preconditioner(mid) == 0
        x = np.zeros(4 * nn)
        it, rr, res = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        assert L.mf_lar_gmres(mid, 30, 2000, 1e-10, 1, 1, P(x), ctypes.byref(it), ctypes.byref(rr),
                              ctypes.byref(res)) == 1, L.mf_last_error()
        assert np.abs(x - xs).max() <= 1e-10 * 120 * np.abs(xs).max()
        its.append(it.value)
    assert its[2] <= its[1] < its[0], its
    free(L, mid)
