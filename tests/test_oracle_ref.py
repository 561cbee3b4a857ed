"""Pin the oracle against the reference's own code, compiled from its sources
into oracle/_ref (mmd_prism, apd_std_lin, lad_crs):
  - quadrature points/weights, shape functions, Jacobians, face area
    elements: bit-exact with apr_set_quadr_3D/2D and apr_elem_calc_3D;
  - jk mesh import: element node order and boundary-face list bit-exact with
    mmr_import_mesh_grad + mmr_el_node_coor + mmr_fa_neig;
  - CRS pattern, assembly tables, assembled values, Jacobi fill, residual and
    Jacobi sweep: bit-exact with las_crs_intf.c.
Skipped when oracle/_ref was not built (no reference tree).
"""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import PRISM, TETRA, OracleProblem, P, oracle, ref, ref_available  # noqa: E402

D = ctypes.c_double
I = ctypes.c_int
APC_BASE_PRISM_STD, APC_BASE_TETRA_STD = 3, 4
MMC_TRIA, MMC_QUAD = 3, 4
MMC_GRADMESH_DATA = 10

need_apstd = pytest.mark.skipif(not ref_available("libref_apstd.so"), reason="oracle/_ref not built")
need_mesh = pytest.mark.skipif(not ref_available("libref_mesh.so"), reason="oracle/_ref not built")
need_crs = pytest.mark.skipif(not ref_available("libref_crs.so"), reason="oracle/_ref not built")


def apstd():
    L = ref("libref_apstd.so")
    L.apr_gauss_init.restype = None
    L.apr_elem_calc_3D.restype = D
    L.apr_elem_calc_3D.argtypes = [I, I, ctypes.c_void_p, I] + [ctypes.c_void_p] * 13
    L.apr_gauss_init()
    return L


def random_prism(rng, distort=0.25):
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float)
    return base + rng.uniform(-distort, distort, base.shape)


def random_tet(rng, distort=0.25):
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    return base + rng.uniform(-distort, distort, base.shape)


@need_apstd
@pytest.mark.parametrize("el_type,base", [(PRISM, APC_BASE_PRISM_STD), (TETRA, APC_BASE_TETRA_STD)])
def test_quadrature_3d_bitexact(el_type, base):
    L = apstd()
    xg, wg = np.zeros(3000), np.zeros(1000)
    ng = I()
    pd = np.array([1], np.int32)
    L.apr_set_quadr_3D(base, P(pd), ctypes.byref(ng), P(xg), P(wg))
    x0, w0 = np.zeros(18), np.zeros(6)
    n0 = oracle().orc_quadr_3d(el_type, P(x0), P(w0))
    assert ng.value == n0
    assert np.array_equal(xg[: 3 * n0], x0[: 3 * n0]) and np.array_equal(wg[:n0], w0[:n0])


@need_apstd
@pytest.mark.parametrize("quad,fa_type", [(0, MMC_TRIA), (1, MMC_QUAD)])
def test_quadrature_2d_bitexact(quad, fa_type):
    L = apstd()
    xg, wg = np.zeros(450), np.zeros(150)
    ng = I()
    pd = np.array([1], np.int32)
    L.apr_set_quadr_2D(fa_type, APC_BASE_PRISM_STD, P(pd), ctypes.byref(ng), P(xg), P(wg))
    x0, w0 = np.zeros(8), np.zeros(4)
    n0 = oracle().orc_quadr_2d(quad, P(x0), P(w0))
    assert ng.value == n0
    assert np.array_equal(xg[: 2 * n0], x0[: 2 * n0]) and np.array_equal(wg[:n0], w0[:n0])


@need_apstd
@pytest.mark.parametrize("el_type,base", [(PRISM, APC_BASE_PRISM_STD), (TETRA, APC_BASE_TETRA_STD)])
def test_elem_calc_bitexact(el_type, base):
    """shape functions, physical gradients, |det| and global coordinates at
    random points of random distorted elements"""
    L = apstd()
    o = oracle()
    rng = np.random.default_rng(7)
    ns = 6 if el_type == PRISM else 4
    pd = np.array([1], np.int32)
    controls = [2] + ([3, 4, 5, 6, 7] if el_type == PRISM else [])
    for trial in range(200):
        nc = np.ascontiguousarray((random_prism(rng) if el_type == PRISM else random_tet(rng)).reshape(-1))
        eta = rng.uniform(0, 0.33, 3)  # inside the reference element (asserted)
        if el_type == PRISM:
            eta[2] = rng.uniform(-1, 1)
        for control in controls:
            out_r = [np.zeros(ns) for _ in range(4)] + [np.zeros(3), np.zeros(3)]
            out_o = [np.zeros(ns) for _ in range(4)] + [np.zeros(3), np.zeros(3)]
            vn_r = P(out_r[5]) if control > 2 else None
            det_r = L.apr_elem_calc_3D(control, 1, P(pd), base, P(eta), P(nc), None,
                                       P(out_r[0]), P(out_r[1]), P(out_r[2]), P(out_r[3]),
                                       P(out_r[4]), None, None, None, None, vn_r)
            det_o = o.orc_elem_calc_3d(el_type, control, P(eta), P(nc), P(out_o[0]), P(out_o[1]),
                                       P(out_o[2]), P(out_o[3]), P(out_o[4]),
                                       P(out_o[5]) if control > 2 else None)
            assert det_r == det_o, (trial, control)
            for a, b in zip(out_r, out_o):
                assert np.array_equal(a, b), (trial, control)


@need_apstd
def test_tet_face_area_and_normal_bitexact_with_hybrid_vertex_order():
    """tetrahedron faces (controls 3-6): the oracle uses the hybrid mesh
    module's face table {0,2,1},{0,1,3},{0,3,2},{1,2,3} (mmd_t4_prism/
    mms_prism_t4_intf.cpp:218); the compiled reference is linked with the
    prism module's {0,1,2},{0,1,3},{0,2,3},{1,2,3}.  Handing the reference
    the element with vertices 1,2 (face 0) or 2,3 (face 2) exchanged makes
    it read the hybrid table's vertices: area element and unit normal are
    then bit-exact"""
    L = apstd()
    o = oracle()
    rng = np.random.default_rng(9)
    pd = np.array([1], np.int32)
    perm = {3: [0, 2, 1, 3], 4: [0, 1, 2, 3], 5: [0, 1, 3, 2], 6: [0, 1, 2, 3]}
    for trial in range(100):
        X = random_tet(rng)
        eta = rng.uniform(0, 0.33, 3)
        for control in (3, 4, 5, 6):
            nc_o = np.ascontiguousarray(X.reshape(-1))
            nc_r = np.ascontiguousarray(X[perm[control]].reshape(-1))
            vr, vo = np.zeros(3), np.zeros(3)
            w = [np.zeros(4) for _ in range(4)] + [np.zeros(3)]
            det_r = L.apr_elem_calc_3D(control, 1, P(pd), APC_BASE_TETRA_STD, P(eta), P(nc_r),
                                       None, P(w[0]), P(w[1]), P(w[2]), P(w[3]), P(w[4]),
                                       None, None, None, None, P(vr))
            det_o = o.orc_elem_calc_3d(TETRA, control, P(eta), P(nc_o), P(w[0]), P(w[1]),
                                       P(w[2]), P(w[3]), P(w[4]), P(vo))
            assert det_r == det_o and np.array_equal(vr, vo), (trial, control)
            # outward: the normal points away from the opposite vertex
            opp = {3: 3, 4: 2, 5: 1, 6: 0}[control]
            assert vo @ (X[opp] - X[[i for i in range(4) if i != opp]].mean(0)) < 0


def write_box_jk(path, nx, ny, nz, seed):
    m = mf.Mesh.box_prism(nx, ny, nz, lx=2.0, ly=1.0, lz=0.5, seed=seed)
    m.write_jk(path)
    return m


@need_mesh
def test_jk_import_matches_reference_mesh_module(tmp_path):
    """element IDs, element-local node order and the boundary int-ent list of
    our jk import equal what the reference mesh module produces"""
    path = str(tmp_path / "box.jk")
    m = write_box_jk(path, 5, 4, 3, seed=13)
    a = m.arrays()
    L = ref("libref_mesh.so")
    mid = L.mmr_init_mesh(MMC_GRADMESH_DATA, path.encode(), None)
    assert mid > 0
    nodes = (I * 10)()
    coor = (D * 30)()
    el, e = 0, 0
    while True:
        el = L.mmr_get_next_elem_all(mid, el)
        if el == 0:
            break
        L.mmr_el_node_coor(mid, el, nodes, coor)
        assert el == e + 1
        got = [nodes[i + 1] - 1 for i in range(nodes[0])]
        assert got == list(a["el_nodes"][e, :6]), (el, got, a["el_nodes"][e])
        assert np.array_equal(np.array(coor[:18]).reshape(6, 3), a["coor"][a["el_nodes"][e, :6]])
        e += 1
    assert e == m.nel
    fn, fs = (I * 2)(), (I * 2)()
    shift = I()
    ac, bc = (D * 4)(), (D * 2)()
    fa, f = 0, 0
    while True:
        fa = L.mmr_get_next_face_all(mid, fa)
        if fa == 0:
            break
        b = L.mmr_fa_bc(mid, fa)
        if b > 0:
            L.mmr_fa_neig(mid, fa, fn, fs, ctypes.byref(shift), None, ac, bc)
            assert (fn[0] - 1, fs[0], b) == (a["bf_elem"][f], a["bf_side"][f], a["bf_bc"][f]), fa
            f += 1
    assert f == m.nbf


@need_mesh
def test_reference_golden_mesh_imports_identically():
    """the reference's own ctest mesh (Laplace_std_in_box_opencl) through both
    importers"""
    path = os.path.join(os.path.dirname(__file__), "golden", "laplace_box_mesh_jk.dat")
    m = mf.Mesh.read_jk(path)
    a = m.arrays()
    L = ref("libref_mesh.so")
    mid = L.mmr_init_mesh(MMC_GRADMESH_DATA, path.encode(), None)
    nodes = (I * 10)()
    el, e = 0, 0
    while True:
        el = L.mmr_get_next_elem_all(mid, el)
        if el == 0:
            break
        L.mmr_el_node_coor(mid, el, nodes, None)
        assert [nodes[i + 1] - 1 for i in range(6)] == list(a["el_nodes"][e, :6])
        e += 1
    assert e == m.nel == 1528


def mesh_dump_paths(tmp_path):
    import zipfile
    gold = os.path.join(os.path.dirname(__file__), "golden")
    zipfile.ZipFile(os.path.join(gold, "prism_cube", "laplace_std_mesh_prism.dat.zip")).extractall(tmp_path)
    return [os.path.join(gold, "prism_cube", "mesh_prism.dat"),
            os.path.join(gold, "dg_cube", "mesh_prism.dmp"),
            os.path.join(str(tmp_path), "mesh_prism.dat")]


@need_mesh
def test_mesh_dump_reader_matches_reference_mesh_module(tmp_path):
    """the reference's prism mesh dumps (ctest Laplace_std_in_cube
    mesh_prism.dat, Laplace_dg mesh_prism.dmp, Laplace_std mesh_prism.dat
    with 1702 nodes) through our reader and through the reference's
    mmr_read_mesh + mmr_el_node_coor + mmr_fa_neig / mmr_fa_bc: element
    order, element-local node order, coordinates and the boundary int-ent
    list (owner, side, bc) identical"""
    L = ref("libref_mesh.so")
    sizes = []
    for path in mesh_dump_paths(tmp_path):
        m = mf.Mesh.read_dat(path)
        a = m.arrays()
        mid = L.mmr_init_mesh(1, path.encode(), None)  # MMC_MOD_FEM_PRISM_DATA
        assert mid > 0
        nodes, coor = (I * 10)(), (D * 30)()
        el, e = 0, 0
        while True:
            el = L.mmr_get_next_elem_all(mid, el)
            if el == 0:
                break
            L.mmr_el_node_coor(mid, el, nodes, coor)
            assert el == e + 1
            assert [nodes[i + 1] - 1 for i in range(6)] == list(a["el_nodes"][e, :6]), (path, el)
            assert np.array_equal(np.array(coor[:18]).reshape(6, 3), a["coor"][a["el_nodes"][e, :6]])
            e += 1
        assert e == m.nel
        fn, fs, shift = (I * 2)(), (I * 2)(), I()
        ac, bc = (D * 4)(), (D * 2)()
        fa, f = 0, 0
        while True:
            fa = L.mmr_get_next_face_all(mid, fa)
            if fa == 0:
                break
            b = L.mmr_fa_bc(mid, fa)
            if b > 0:
                L.mmr_fa_neig(mid, fa, fn, fs, ctypes.byref(shift), None, ac, bc)
                assert (fn[0] - 1, fs[0], b) == (a["bf_elem"][f], a["bf_side"][f], a["bf_bc"][f])
                f += 1
        assert f == m.nbf
        sizes.append((m.nno, m.nel, m.nbf))
        m.free()
    assert sizes[0] == sizes[1] == (8, 2, 8) and sizes[2][0] == 1702


def ref_crs():
    L = ref("libref_crs.so")
    L.lar_get_SM_and_LV_crs_from_crs.argtypes = [I, I, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p]
    return L


def neighbour_lists(op):
    """per block: sorted off-diagonal neighbour blocks (sir_create's l_neig_bl)"""
    lists = []
    for r in range(op.n):
        cols = op.col_ind[op.row_ptr[r]: op.row_ptr[r + 1]]
        lists.append(np.ascontiguousarray(cols[cols != r], np.int32))
    return lists


def ref_crs_arrays(L, mid, n):
    rp, ci, va, rh = (ctypes.POINTER(I)(), ctypes.POINTER(I)(), ctypes.POINTER(D)(),
                      ctypes.POINTER(D)())
    L.lar_get_SM_and_LV_crs_from_crs(mid, 0, ctypes.byref(rp), ctypes.byref(ci), ctypes.byref(va),
                                     ctypes.byref(rh))
    nnz = rp[n]
    return (np.ctypeslib.as_array(rp, (n + 1,)).copy(), np.ctypeslib.as_array(ci, (nnz,)).copy(),
            np.ctypeslib.as_array(va, (nnz,)).copy(), np.ctypeslib.as_array(rh, (n,)).copy())


@need_crs
def test_crs_storage_bitexact_with_reference():
    m = mf.Mesh.box_prism(6, 5, 4, seed=17, jitter=0.2)
    op = OracleProblem(m.arrays())
    L = ref_crs()
    lists = neighbour_lists(op)
    nro = np.array([len(x) for x in lists], np.int32)
    ptrs = (ctypes.POINTER(I) * op.n)(*[x.ctypes.data_as(ctypes.POINTER(I)) for x in lists])
    mid = L.lar_allocate_SM_and_LV_crs(op.n, 36, P(nro), ptrs)
    assert mid >= 0
    rp, ci, _, _ = ref_crs_arrays(L, mid, op.n)
    assert np.array_equal(rp, op.row_ptr) and np.array_equal(ci, op.col_ind)
    # tables of every int-ent
    for ie in range(op.nel + op.nbf):
        e = ie if ie < op.nel else op.bfe[ie - op.nel]
        L_bl = np.ascontiguousarray(op.node_dof[op.en[e, :6]], np.int32)
        tab = np.zeros(42, np.int32)
        nrdof = np.ones(6, np.int32)
        L.lar_fill_assembly_table_int_ent_crs(mid, 6, P(L_bl), P(nrdof), P(tab))
        assert np.array_equal(tab, op.table[op.table_ptr[ie]: op.table_ptr[ie + 1]]), ie
    # assembly through the reference with the oracle's entity matrices
    rng = np.random.default_rng(3)
    kappa = np.exp(rng.normal(size=op.nel))
    coef = np.zeros((op.nel, 21))
    coef[:, 0] = coef[:, 4] = coef[:, 8] = kappa
    coef[:, 19] = 1.0
    val0, rhs0 = op.assemble(mf.IND_DIFF, coef)
    L.lar_initialize_SM_and_LV_crs(mid, 0)
    for ie in range(op.nel + op.nbf):
        if ie < op.nel:
            K, F = op.element(ie, mf.IND_DIFF, coef[ie])
        else:
            K, F = op.face(ie - op.nel)
        sm = np.ascontiguousarray(K.reshape(-1))
        rv = np.ascontiguousarray(F)
        tab = np.ascontiguousarray(op.table[op.table_ptr[ie]: op.table_ptr[ie + 1]])
        L.lar_assemble_SM_and_LV_with_table_crs(mid, 0, 6, P(tab), P(sm), P(rv), None)
    _, _, va, rh = ref_crs_arrays(L, mid, op.n)
    assert np.array_equal(va, val0) and np.array_equal(rh, rhs0)
    # Jacobi fill, residual, one BJ sweep
    L.lar_allocate_preconditioner_crs(mid, 1, 0)
    L.lar_fill_preconditioner_crs(mid)
    o = oracle()
    dptr, dinv = np.zeros(op.n, np.int32), np.zeros(op.n)
    o.orc_fill_jacobi(op.n, P(op.row_ptr), P(op.col_ind), P(val0), P(dptr), P(dinv))
    x = rng.uniform(-1, 1, op.n)
    v_r, v_o = np.zeros(op.n), np.zeros(op.n)
    L.lar_compute_residual_crs(mid, 1, 0, op.n, P(x), None, P(v_r))
    o.orc_compute_residual(op.n, P(op.row_ptr), P(op.col_ind), P(val0), 1, P(rhs0), P(x), P(v_o))
    assert np.array_equal(v_r, v_o)
    w_r, w_o = x.copy(), x.copy()
    L.lar_perform_BJ_or_GS_iterations_crs(mid, 1, 0, 1, op.n, P(w_r), None)
    o.orc_bj_iteration(op.n, P(op.row_ptr), P(op.col_ind), P(val0), P(dptr), P(dinv), 1, P(rhs0),
                       P(w_o))
    assert np.array_equal(w_r, w_o)
    L.lar_free_preconditioner_crs(mid)
    L.lar_free_SM_and_LV_crs(mid)


def ref_crs_matrix(L, op, val, rhs):
    """a CRS matrix in R's lad_crs (R's own allocation from neighbour lists)
    holding the oracle's values and rhs, written through R's accessor"""
    lists = neighbour_lists(op)
    nro = np.array([len(x) for x in lists], np.int32)
    ptrs = (ctypes.POINTER(I) * op.n)(*[x.ctypes.data_as(ctypes.POINTER(I)) for x in lists])
    mid = L.lar_allocate_SM_and_LV_crs(op.n, 36, P(nro), ptrs)
    assert mid >= 0
    L.lar_initialize_SM_and_LV_crs(mid, 0)
    rp, ci, va, rh = (ctypes.POINTER(I)(), ctypes.POINTER(I)(), ctypes.POINTER(D)(),
                      ctypes.POINTER(D)())
    L.lar_get_SM_and_LV_crs_from_crs(mid, 0, ctypes.byref(rp), ctypes.byref(ci), ctypes.byref(va),
                                     ctypes.byref(rh))
    np.ctypeslib.as_array(va, (len(val),))[:] = val
    np.ctypeslib.as_array(rh, (op.n,))[:] = rhs
    return mid


@need_crs
@pytest.mark.parametrize("nthreads", [1, 2, 3, 4])
def test_chunked_gs_bitexact_with_reference_threads(nthreads):
    """R's BLOCK_GS sweep depends on its OpenMP thread count through the row
    chunks; the oracle's chunked restatement reproduces it bit for bit for
    1..4 threads, two sweeps with the stored rhs and one without"""
    m = mf.Mesh.box_prism(7, 6, 5, seed=23, jitter=0.2)
    op = OracleProblem(m.arrays())
    rng = np.random.default_rng(5 + nthreads)
    kappa = np.exp(rng.normal(size=op.nel))
    coef = np.zeros((op.nel, 21))
    coef[:, 0] = coef[:, 4] = coef[:, 8] = kappa
    coef[:, 1] = coef[:, 3] = 0.1 * kappa
    coef[:, 19] = 1.0
    val, rhs = op.assemble(mf.IND_DIFF, coef)
    L = ref_crs()
    gomp = ctypes.CDLL("libgomp.so.1")
    mid = ref_crs_matrix(L, op, val, rhs)
    assert L.lar_allocate_preconditioner_crs(mid, 2, 0) > 0
    L.lar_fill_preconditioner_crs(mid)
    o = oracle()
    dptr, dinv = np.zeros(op.n, np.int32), np.zeros(op.n)
    o.orc_fill_jacobi(op.n, P(op.row_ptr), P(op.col_ind), P(val), P(dptr), P(dinv))
    x = rng.uniform(-1, 1, op.n)
    gomp.omp_set_num_threads(nthreads)
    try:
        for use_rhs, sweeps in ((1, 2), (0, 1)):
            w_r, w_o = x.copy(), x.copy()
            L.lar_perform_BJ_or_GS_iterations_crs(mid, use_rhs, 0, sweeps, op.n, P(w_r), None)
            o.orc_crs_gs_chunked(op.n, P(op.row_ptr), P(op.col_ind), P(val), P(dptr), P(dinv),
                                 use_rhs, P(rhs), nthreads, sweeps, P(w_o))
            assert np.array_equal(w_r, w_o), (nthreads, use_rhs)
    finally:
        gomp.omp_set_num_threads(os.cpu_count() or 1)
        L.lar_free_preconditioner_crs(mid)
        L.lar_free_SM_and_LV_crs(mid)


@pytest.mark.parametrize("nthreads", [1, 5])
def test_chunked_gs_is_gauss_seidel_per_chunk(nthreads):
    """independent of R: one chunked sweep solves (D + L_c) v = b - (A - D - L_c) v0,
    L_c the strictly lower entries whose column lies in the row's own chunk"""
    m = mf.Mesh.box_prism(5, 4, 4, seed=2, jitter=0.2)
    op = OracleProblem(m.arrays())
    rng = np.random.default_rng(11)
    coef = np.zeros((op.nel, 21))
    coef[:, 0] = coef[:, 4] = coef[:, 8] = np.exp(rng.normal(size=op.nel))
    coef[:, 19] = 1.0
    val, rhs = op.assemble(mf.IND_DIFF, coef)
    o = oracle()
    dptr, dinv = np.zeros(op.n, np.int32), np.zeros(op.n)
    o.orc_fill_jacobi(op.n, P(op.row_ptr), P(op.col_ind), P(val), P(dptr), P(dinv))
    x = rng.uniform(-1, 1, op.n)
    w = x.copy()
    o.orc_crs_gs_chunked(op.n, P(op.row_ptr), P(op.col_ind), P(val), P(dptr), P(dinv), 1, P(rhs),
                         nthreads, 1, P(w))
    A = np.zeros((op.n, op.n))
    for r in range(op.n):
        A[r, op.col_ind[op.row_ptr[r]:op.row_ptr[r + 1]]] = val[op.row_ptr[r]:op.row_ptr[r + 1]]
    chunk = np.arange(op.n) // (op.n // nthreads + 1)
    Lc = np.tril(A, -1) * (chunk[:, None] == chunk[None, :])
    D = np.diag(np.diag(A))
    lhs = (D + Lc) @ w + (A - D - Lc) @ x
    scale = np.abs(A) @ (np.abs(w) + np.abs(x)) + np.abs(rhs)
    assert (np.abs(lhs - rhs) <= 1e-13 * scale).all()
    if nthreads > 1:  # the chunks matter: plain GS differs
        w1 = x.copy()
        o.orc_crs_gs_chunked(op.n, P(op.row_ptr), P(op.col_ind), P(val), P(dptr), P(dinv), 1,
                             P(rhs), 1, 1, P(w1))
        assert np.abs(w1 - w).max() > 1e-6
