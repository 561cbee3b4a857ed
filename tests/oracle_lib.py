"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the
reference subsets compiled from the reference's own sources (oracle/_ref).
Test infrastructure: used only as the checker."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

PRISM, TETRA = 5, 7
V = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


def P(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


_orc = None


def oracle():
    global _orc
    if _orc is None:
        L = ctypes.CDLL(ORACLE)
        sig = {
            "orc_quadr_3d": (I, [I, V, V]),
            "orc_quadr_2d": (I, [I, V, V]),
            "orc_elem_calc_3d": (D, [I, I, V, V, V, V, V, V, V, V]),
            "orc_num_int_el": (I, [I, V, V, V, I, V, V]),
            "orc_face_penalty": (I, [I, I, V, D, I, V, V]),
            "orc_face_neumann": (I, [I, I, V, D, V]),
            "orc_face_neumann_exact": (I, [I, I, V, I, V]),
            "orc_dof_list": (I, [I, V, V, I, V, V]),
            "orc_crs_pattern": (ctypes.c_long, [I, V, V, V, I, V, V]),
            "orc_assembly_tables": (ctypes.c_long, [I, V, V, I, V, V, V, V, V, V]),
            "orc_integrate_assemble": (I, [I, V, V, V, V, V, I, I, V, V, V, V, V, V, V]),
            "orc_fill_jacobi": (None, [I, V, V, V, V, V]),
            "orc_compute_residual": (None, [I, V, V, V, I, V, V, V]),
            "orc_bj_iteration": (None, [I, V, V, V, V, V, I, V, V]),
            "orc_crs_gs_chunked": (None, [I, V, V, V, V, V, I, V, I, I, V]),
            "orc_gmres": (I, [I, V, V, V, V, I, I, I, D, I, V, V, V, V]),
            "orc_gmres_tol": (I, [I, V, V, V, V, I, I, I, D, D, I, V, V, V, V]),
            "orc_ilu0": (I, [I, V, V, V, V]),
            "orc_gs_sweeps": (None, [I, V, V, V, V, V, V, I, I, V, V]),
            "orc_set_gs": (None, [V, I]),
            "orc_ilu_solve": (None, [I, V, V, V, V, V]),
            "orc_drotg": (None, [V, V, V, V]),
            "orc_dg_shape": (I, [I, V, V, V, V, V]),
            "orc_dg_quadr": (I, [I, V, V]),
            "orc_dg_ns": (I, [I]),
            "orc_dg_penalty_degree": (I, [I, I]),
            "orc_dg_num_int_el": (I, [I, V, V, V, I, V, V]),
            "orc_dg_face_quadr": (I, [I, I, V, V]),
            "orc_dg_face_eta": (None, [I, V, V]),
            "orc_dg_face_ds": (D, [I, V, V, V, V]),
            "orc_dg_map_point": (I, [I, V, I, V, V, V]),
            "orc_dg_face_area": (D, [V, I]),
            "orc_dg_face_int": (I, [I, I, V, V, I, V, V, V, I, V, D, D, I, D, V, V]),
            "orc_dg_face_int2": (I, [I, I, I, V, V, I, V, V, V, I, V, D, D, I, D, V, V]),
            "orc_dg_eval": (D, [I, V, V, V, V]),
            "orc_dgt_ns": (I, [I]),
            "orc_dgt_shape": (I, [I, V, V, V, V, V]),
            "orc_dgt_exponents": (I, [I, V]),
            "orc_dgt_quadr": (I, [I, V, V]),
            "orc_dgt_num_int_el": (I, [I, V, V, V, I, V, V]),
            "orc_dgt_face_eta": (None, [I, V, V]),
            "orc_dgt_face_ds": (D, [I, V, V, V, V]),
            "orc_dgt_face_area": (D, [V, I]),
            "orc_dgx_map_point": (I, [I, I, V, V, I, I, V, V]),
            "orc_dgx_face_int": (I, [I, I, I, I, I, V, V, I, V, V, V, I, V, D, D, I, D, V, V]),
            "orc_dgt_eval": (D, [I, V, V, V, V]),
            "orc_dg_errors": (I, [I, V, V, I, V, V]),
            "orc_dg_assemble_sample": (D, [I, I, V, V, I, V, V, V, D, I]),
            "orc_bcrs_pattern": (I, [I, V, V, V, V]),
            "orc_bcrs_assemble": (None, [V, V, I, V, V, V, V, V]),
            "orc_bcrs_residual": (None, [I, V, V, V, I, V, V, V]),
            "orc_lu4": (I, [V, V]),
            "orc_lu4_solve": (None, [V, V, V]),
            "orc_bcrs_bj_gs": (None, [I, V, V, V, V, V, V, I, I, I, V, V]),
            "orc_inv4": (None, [V, V]),
            "orc_bcrs_ilu0": (I, [I, V, V, V, V]),
            "orc_bcrs_rhsub": (None, [I, V, V, V, V, V, V]),
            "orc_set_threads": (None, [I]),
            "orc_lun": (I, [I, V, V]),
            "orc_bcrs_iluk_pattern": (ctypes.c_long, [I, V, V, I, I, V, V]),
            "orc_invn": (None, [I, V, V]),
            "orc_block_ilu0": (I, [I, V, V, V, V, V, V]),
            "orc_block_ilu_solve": (None, [I, V, V, V, V, V, V, V, V, V]),
            "orc_std_errors": (I, [I, V, V, V, V, I, V]),
            "orc_set_tet_side3_scale": (None, [D]),
            "orc_dg_set_plant_downwind": (None, [I]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _orc = L
    return _orc


def ref_available(name):
    return os.path.exists(os.path.join(REF_DIR, name))


_refs = {}


def ref(name):
    """a reference subset library, loaded lazily-bound: symbols of modules
    that are not compiled in stay unresolved and are never called"""
    if name not in _refs:
        mode = os.RTLD_LAZY | os.RTLD_GLOBAL
        if name == "libref_dg.so":
            # the DG module defines the same apr_* names as the std-linear
            # one: bind its own calls to its own definitions
            mode = os.RTLD_LAZY | os.RTLD_LOCAL | os.RTLD_DEEPBIND
        _refs[name] = ctypes.CDLL(os.path.join(REF_DIR, name), mode=mode)
    return _refs[name]


class OracleProblem:
    """the oracle's view of a mesh: lists, pattern, tables, assembly"""

    def __init__(self, arrays):
        self.a = arrays
        o = oracle()
        self.nno = arrays["coor"].shape[0]
        self.nel = arrays["el_type"].shape[0]
        self.nbf = arrays["bf_elem"].shape[0]
        self.et = np.ascontiguousarray(arrays["el_type"], np.int32)
        self.en = np.ascontiguousarray(arrays["el_nodes"], np.int32)
        self.coor = np.ascontiguousarray(arrays["coor"], np.float64)
        self.bfe = np.ascontiguousarray(arrays["bf_elem"], np.int32)
        self.bfs = np.ascontiguousarray(arrays["bf_side"], np.int32)
        self.dof_node = np.zeros(self.nno, np.int32)
        self.node_dof = np.zeros(self.nno, np.int32)
        self.n = o.orc_dof_list(self.nel, P(self.et), P(self.en), self.nno,
                                P(self.dof_node), P(self.node_dof))
        self.dof_node = self.dof_node[: self.n].copy()
        self.nnz = o.orc_crs_pattern(self.nel, P(self.et), P(self.en), P(self.node_dof),
                                     self.n, None, None)
        self.row_ptr = np.zeros(self.n + 1, np.int32)
        self.col_ind = np.zeros(self.nnz, np.int32)
        o.orc_crs_pattern(self.nel, P(self.et), P(self.en), P(self.node_dof), self.n,
                          P(self.row_ptr), P(self.col_ind))
        nint = self.nel + self.nbf
        self.table_ptr = np.zeros(nint + 1, np.int64)
        tot = o.orc_assembly_tables(self.nel, P(self.et), P(self.en), self.nbf, P(self.bfe),
                                    P(self.node_dof), P(self.row_ptr), P(self.col_ind),
                                    P(self.table_ptr), None)
        self.table = np.zeros(tot, np.int32)
        o.orc_assembly_tables(self.nel, P(self.et), P(self.en), self.nbf, P(self.bfe),
                              P(self.node_dof), P(self.row_ptr), P(self.col_ind),
                              P(self.table_ptr), P(self.table))

    def assemble(self, ind, coef_el, src_model=0, bf_g=None):
        o = oracle()
        ind = np.ascontiguousarray(ind, np.int32)
        coef_el = np.ascontiguousarray(coef_el, np.float64)
        self.val = np.zeros(self.nnz)
        self.rhs = np.zeros(self.n)
        g = None if bf_g is None else np.ascontiguousarray(bf_g, np.float64)
        o.orc_integrate_assemble(self.nel, P(self.et), P(self.en), P(self.coor), P(ind),
                                 P(coef_el), src_model, self.nbf, P(self.bfe), P(self.bfs),
                                 P(g), P(self.table_ptr), P(self.table), P(self.val),
                                 P(self.rhs))
        return self.val, self.rhs

    def element(self, e, ind, coef, src_model=0):
        o = oracle()
        k = 4 if self.et[e] == TETRA else 6
        nc = np.ascontiguousarray(self.coor[self.en[e, :k]].reshape(-1))
        sm = np.zeros(36)
        rhs = np.zeros(6)
        o.orc_num_int_el(int(self.et[e]), P(nc), P(np.ascontiguousarray(ind, np.int32)),
                         P(np.ascontiguousarray(coef, np.float64)), src_model, P(sm), P(rhs))
        return sm[: k * k].reshape(k, k), rhs[:k]  # [u][w] (column-major stiff[u*k+w])

    def face(self, f, g=0.0, src_model=0):
        o = oracle()
        e = self.bfe[f]
        k = 4 if self.et[e] == TETRA else 6
        nc = np.ascontiguousarray(self.coor[self.en[e, :k]].reshape(-1))
        sm = np.zeros(36)
        rhs = np.zeros(6)
        o.orc_face_penalty(int(self.et[e]), int(self.bfs[f]), P(nc), g, src_model, P(sm), P(rhs))
        return sm[: k * k].reshape(k, k), rhs[:k]

    def face_neumann(self, f, g, src_model=0):
        """Neumann face rhs: constant flux g, or for an exact-solution source
        model the exact normal flux at every face point"""
        e = self.bfe[f]
        k = 4 if self.et[e] == TETRA else 6
        nc = np.ascontiguousarray(self.coor[self.en[e, :k]].reshape(-1))
        rhs = np.zeros(6)
        if src_model:
            oracle().orc_face_neumann_exact(int(self.et[e]), int(self.bfs[f]), P(nc), src_model, P(rhs))
        else:
            oracle().orc_face_neumann(int(self.et[e]), int(self.bfs[f]), P(nc), g, P(rhs))
        return rhs[:k]

    def gmres(self, val, rhs, precon=1, krylbas=30, max_iter=1000, tol=1e-8, tol_abs=0.0):
        o = oracle()
        x = np.zeros(self.n)
        it, rr, rs = I(), D(), D()
        conv = o.orc_gmres_tol(self.n, P(self.row_ptr), P(self.col_ind), P(val), P(rhs), precon,
                               krylbas, max_iter, tol, tol_abs, 1, P(x), ctypes.byref(it),
                               ctypes.byref(rr), ctypes.byref(rs))
        return x, dict(converged=bool(conv), iters=it.value, res_rel=rr.value, resid=rs.value)


def system_with_face_types(op, ind, coef, src_model, neumann):
    """the oracle's assembled system with the boundary faces flagged in
    `neumann` taking the exact (or zero, for a constant-source model) normal
    flux instead of the Dirichlet penalty: the assembly with penalty on every
    face, then each Neumann face's penalty block and load replaced"""
    v0, r0 = op.assemble(ind, coef, src_model=src_model)
    v0, r0 = v0.copy(), r0.copy()
    for f in np.nonzero(neumann)[0]:
        K, F = op.face(int(f), g=0.0, src_model=src_model)
        k = K.shape[0]
        t = op.table[op.table_ptr[op.nel + f]:op.table_ptr[op.nel + f + 1]]
        flux = op.face_neumann(int(f), 0.0, src_model=src_model)
        for w in range(k):
            for u in range(k):
                v0[t[w + u * k]] -= K[u][w]
            r0[t[k * k + w]] += flux[w] - F[w]
    return v0, r0


def std_errors(arrays, u_node, src_model):
    """(L2 norm, H1 seminorm) of u_h - u_exact as pdr_error_test_exact
    computes them; u_node holds one value per mesh node"""
    et = np.ascontiguousarray(arrays["el_type"], np.int32)
    en = np.ascontiguousarray(arrays["el_nodes"], np.int32)
    coor = np.ascontiguousarray(arrays["coor"], np.float64)
    u = np.ascontiguousarray(u_node, np.float64)
    out = np.zeros(2)
    oracle().orc_std_errors(len(et), P(et), P(en), P(coor), P(u), int(src_model), P(out))
    return out[0], out[1]


def oracle_gmres_crs(row_ptr, col_ind, val, rhs, precon=1, krylbas=30, max_iter=1000, tol=1e-8):
    o = oracle()
    n = row_ptr.shape[0] - 1
    x = np.zeros(n)
    it, rr, rs = I(), D(), D()
    conv = o.orc_gmres(n, P(row_ptr), P(col_ind), P(val), P(rhs), precon, krylbas, max_iter,
                       tol, 1, P(x), ctypes.byref(it), ctypes.byref(rr), ctypes.byref(rs))
    return x, dict(converged=bool(conv), iters=it.value, res_rel=rr.value, resid=rs.value)


def oracle_ilu0(row_ptr, col_ind, val):
    """ILU(0) factors on the pattern (strict L, strict U, 1/u_ii on the diagonal)"""
    lu = np.zeros_like(val)
    rc = oracle().orc_ilu0(row_ptr.shape[0] - 1, P(row_ptr), P(col_ind), P(np.ascontiguousarray(val)), P(lu))
    assert rc == 0
    return lu


def ilu_product(row_ptr, col_ind, lu, pos=None):
    """L U (dense) from factors in CRS positions; pos = elimination position
    of every row (identity when None); rows/columns in elimination order"""
    n = row_ptr.shape[0] - 1
    pos = np.arange(n) if pos is None else pos
    L, U = np.eye(n), np.zeros((n, n))
    for r in range(n):
        for p in range(row_ptr[r], row_ptr[r + 1]):
            i, j = pos[r], pos[col_ind[p]]
            if j < i:
                L[i, j] = lu[p]
            elif j > i:
                U[i, j] = lu[p]
            else:
                U[i, i] = 1.0 / lu[p]
    return L @ U



def dg_shape(pdeg, eta):
    """oracle DG prism shape functions at one point: (phi, dx, dy, dz)"""
    out = [np.zeros(128) for _ in range(4)]
    e = np.ascontiguousarray(eta, np.float64)
    ns = oracle().orc_dg_shape(pdeg, P(e), *[P(a) for a in out])
    assert ns > 0
    return [a[:ns] for a in out]


def dg_quadr(pdeg):
    xg, wg = np.zeros(3 * 768), np.zeros(768)
    ng = oracle().orc_dg_quadr(pdeg, P(xg), P(wg))
    assert ng > 0
    return xg[:3 * ng].reshape(ng, 3), wg[:ng]


def dg_num_int_el(pdeg, nc, ind, coef, src_model=0, tet=False):
    """oracle DG element integral: (stiff[ns, ns] column-major as written,
    i.e. stiff[j, i] = entry for trial j, test i; rhs[ns]); a tet takes the
    complete base of degree pdeg % 100 and its first four nodes"""
    st, rhs = np.zeros(128 * 128), np.zeros(128)
    o = oracle()
    if tet:
        ns = o.orc_dgt_num_int_el(int(pdeg) % 100, P(np.ascontiguousarray(np.asarray(nc).reshape(-1)[:12], np.float64)),
                                  P(np.ascontiguousarray(ind, np.int32)),
                                  P(np.ascontiguousarray(coef, np.float64)), src_model, P(st), P(rhs))
    else:
        ns = o.orc_dg_num_int_el(pdeg, P(np.ascontiguousarray(nc, np.float64)),
                                 P(np.ascontiguousarray(ind, np.int32)),
                                 P(np.ascontiguousarray(coef, np.float64)), src_model,
                                 P(st), P(rhs))
    assert ns > 0
    return st[:ns * ns].reshape(ns, ns), rhs[:ns]


# prism face vertices (mmd_prism/mms_prism_intf.c:98), quads cyclic
DG_FACE_V = [[0, 1, 2], [3, 4, 5], [0, 1, 4, 3], [1, 2, 5, 4], [2, 0, 3, 5]]
# tetrahedron sides (the mesh module's face-node table)
DGT_FACE_V = [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]
TETRA = 7
# complete prism base of degree p: pdeg DG_COMPLETE + p (the oracle's encoding)
DG_COMPLETE = 1000


def dg_faces(el_nodes, el_type=None):
    """faces of a prism / tet mesh, each once: (e1, s1, e2, s2) with e1 < e2,
    or (e, s, -1, -1) for a boundary face"""
    seen = {}
    faces = []
    for e, nodes in enumerate(el_nodes):
        tet = el_type is not None and el_type[e] == TETRA
        for s, fv in enumerate(DGT_FACE_V if tet else DG_FACE_V):
            key = tuple(sorted(int(nodes[v]) for v in fv))
            if key in seen:
                e1, s1 = seen.pop(key)
                faces.append((e1, s1, e, s))
            else:
                seen[key] = (e, s)
    for (e, s) in seen.values():
        faces.append((e, s, -1, -1))
    return faces


def dg_degrees(pdeg, nel, el_type=None):
    """per-element degrees, block sizes and block offsets (pdeg: one degree
    or one per element; tets take the complete base of degree pdeg % 100)"""
    pd = np.full(nel, int(pdeg)) if np.isscalar(pdeg) else np.asarray(pdeg, int)
    comp = pd > DG_COMPLETE
    px = np.where(comp, pd - DG_COMPLETE, pd % 100)
    pz = pd // 100
    ns = np.where(comp, (px + 1) * (px + 2) * (px + 3) // 6, (px + 1) * (px + 2) // 2 * (pz + 1))
    if el_type is not None:
        tet = np.asarray(el_type) == TETRA
        ns = np.where(tet, (px + 1) * (px + 2) * (px + 3) // 6, ns)
    return pd, ns, np.r_[0, np.cumsum(ns)]


def _dg_parts(arrays, pdeg, ind, coef_of, penalty_const, src_model, bval, swap=False):
    """(index vector, matrix, load) of every element and face integral of the
    oracle's DG system, element blocks at the offsets of dg_degrees; faces
    each once with the lower element as first neighbour (swap: the higher)"""
    o = oracle()
    coor, en = arrays["coor"], arrays["el_nodes"]
    et = arrays.get("el_type")
    nel = en.shape[0]
    nodes = coor[en]
    pd, ns, off = dg_degrees(pdeg, nel, et)
    tet = np.zeros(nel, bool) if et is None else np.asarray(et) == TETRA
    comp = pd > DG_COMPLETE
    pmin = np.where(tet, pd % 100, np.where(comp, pd - DG_COMPLETE, np.minimum(pd % 100, pd // 100)))
    # tets pass their complete degree p, prisms their pdeg
    pk = np.where(tet, pd % 100, pd)
    for e in range(nel):
        K, F = dg_num_int_el(int(pd[e]), nodes[e], ind, coef_of(e), src_model, tet=bool(tet[e]))
        yield np.arange(off[e], off[e + 1]), K.T, F
    bc_of = _dg_bc_types(arrays)
    st = np.zeros(80 * 80)
    rhs = np.zeros(80)
    for (e1, s1, e2, s2) in dg_faces(en, et):
        if swap and e2 >= 0:
            e1, s1, e2, s2 = e2, s2, e1, s1
        st[:] = 0.0
        rhs[:] = 0.0
        nc1 = np.ascontiguousarray(nodes[e1].ravel())
        g1 = np.ascontiguousarray(en[e1], np.int32)
        hf = np.sqrt(o.orc_dgt_face_area(P(nc1), s1) if tet[e1] else o.orc_dg_face_area(P(nc1), s1))
        c1 = np.ascontiguousarray(coef_of(e1), np.float64)
        if e2 < 0:
            pen = pmin[e1] * penalty_const
            nf = o.orc_dgx_face_int(int(tet[e1]), int(pk[e1]), int(tet[e1]), int(pk[e1]),
                                    bc_of(e1, s1), P(nc1), P(g1), s1, P(c1), None, None, 0, None,
                                    pen, hf, src_model, bval, P(st), P(rhs))
            assert nf == ns[e1]
            idx = np.arange(off[e1], off[e1 + 1])
        else:
            nc2 = np.ascontiguousarray(nodes[e2].ravel())
            g2 = np.ascontiguousarray(en[e2], np.int32)
            c2 = np.ascontiguousarray(coef_of(e2), np.float64)
            pen = min(pmin[e1], pmin[e2]) * penalty_const
            nf = o.orc_dgx_face_int(int(tet[e1]), int(pk[e1]), int(tet[e2]), int(pk[e2]), 0,
                                    P(nc1), P(g1), s1, P(c1), P(nc2), P(g2), s2, P(c2), pen, hf,
                                    src_model, bval, P(st), P(rhs))
            assert nf == ns[e1] + ns[e2]
            idx = np.r_[off[e1]:off[e1 + 1], off[e2]:off[e2 + 1]]
        yield idx, st[:nf * nf].reshape(nf, nf).T.copy(), rhs[:nf].copy()


def dg_system_dense(arrays, pdeg, ind, coef_of, penalty_const=1.0, src_model=0, bval=0.0,
                    sparse=False, swap=False):
    """the oracle's DG system: element integrals plus every face's matrices
    (each neighbour pair once, first neighbour = lower element), boundary
    faces by their flags; dense, or scipy CSR with sparse=True.  pdeg: one
    degree or one per element (blocks at the offsets of dg_degrees)"""
    nel = arrays["el_nodes"].shape[0]
    n = dg_degrees(pdeg, nel, arrays.get("el_type"))[2][-1]
    parts = _dg_parts(arrays, pdeg, ind, coef_of, penalty_const, src_model, bval, swap)
    b = np.zeros(n)
    if sparse:
        import scipy.sparse as sp
        rows, cols, vals = [], [], []
        for idx, M, F in parts:
            k = len(idx)
            rows.append(np.repeat(idx, k)); cols.append(np.tile(idx, k)); vals.append(M.ravel())
            b[idx] += F
        A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(n, n)).tocsr()
        return A, b
    A = np.zeros((n, n))
    for idx, M, F in parts:
        A[np.ix_(idx, idx)] += M
        b[idx] += F
    return A, b


def _dg_bc_types(arrays):
    """boundary face (element, side) -> 1 Dirichlet / 2 Neumann from the mesh
    bc flags (pdr_get_bc_type, pds_conv_diff_weakform.c:1240-1259: flags
    1..9 Dirichlet, 10..19 Neumann); faces without a flag are Dirichlet"""
    t = {}
    for e, s, bc in zip(arrays["bf_elem"], arrays["bf_side"], arrays["bf_bc"]):
        t[(int(e), int(s))] = 1 if bc < 10 else 2 if bc < 20 else 3
    return lambda e, s: t.get((e, s), 1)


def dg_errors(arrays, pdeg, x, src_model=1):
    """(L2 error, H1-seminorm error) of a DG solution (element blocks in
    element order) against the exact solution"""
    o = oracle()
    nodes = arrays["coor"][arrays["el_nodes"]]
    nel = nodes.shape[0]
    ns = len(x) // nel
    l2 = h1 = 0.0
    a, c = ctypes.c_double(), ctypes.c_double()
    for e in range(nel):
        o.orc_dg_errors(pdeg, P(np.ascontiguousarray(nodes[e].ravel())),
                        P(np.ascontiguousarray(x[e * ns:(e + 1) * ns])), src_model,
                        ctypes.byref(a), ctypes.byref(c))
        l2 += a.value
        h1 += c.value
    return np.sqrt(l2), np.sqrt(h1)


# --- BCRS storage with 4x4 blocks -------------------------------------------

def bcrs_vector_problem(arrays, seed, diag_shift=6.0):
    """a 4-component problem on a prism mesh: node blocks, neighbour lists
    (nodes sharing an element), 24x24 element matrices (column-major,
    uniform(-1, 1) + diag_shift I) and 24-vectors"""
    en = arrays["el_nodes"]
    nn = arrays["coor"].shape[0]
    nbs = [set() for _ in range(nn)]
    for e in en:
        for u in e:
            nbs[u].update(int(v) for v in e if v != u)
    nb_lists = [sorted(s) for s in nbs]
    rng = np.random.default_rng(seed)
    mats = []
    for _ in range(en.shape[0]):
        K = rng.uniform(-1, 1, (24, 24))
        K += np.eye(24) * diag_shift
        mats.append((np.ascontiguousarray(K.T.ravel()), rng.uniform(-1, 1, 24)))
    return en, nn, nb_lists, mats


def bcrs_flat(nb_lists):
    ptr = np.zeros(len(nb_lists) + 1, np.int32)
    ptr[1:] = np.cumsum([len(x) for x in nb_lists])
    ind = np.array([v for x in nb_lists for v in x], np.int32)
    return ptr, ind


def bcrs_assemble(nn, nb_lists, en, mats):
    """the oracle's BCRS pattern and element-by-element assembly"""
    o = oracle()
    ptr, ind = bcrs_flat(nb_lists)
    rp = np.zeros(nn + 1, np.int32)
    nsm = o.orc_bcrs_pattern(nn, P(ptr), P(ind), P(rp), None)
    ci = np.zeros(nsm, np.int32)
    o.orc_bcrs_pattern(nn, P(ptr), P(ind), P(rp), P(ci))
    val = np.zeros(16 * nsm)
    rhs = np.zeros(4 * nn)
    for e, (K, F) in zip(en, mats):
        l_bl = np.ascontiguousarray(e, np.int32)
        o.orc_bcrs_assemble(P(rp), P(ci), 6, P(l_bl), P(K), P(np.ascontiguousarray(F)), P(val), P(rhs))
    return rp, ci, val, rhs


def bcrs_diag_lu(nn, rp, ci, val):
    """diagonal positions and the oracle's LU factors of the diagonal blocks"""
    o = oracle()
    diag = np.array([next(s for s in range(rp[r], rp[r + 1]) if ci[s] == r) for r in range(nn)],
                    np.int32)
    lu = np.zeros(16 * nn)
    piv = np.zeros(4 * nn, np.int32)
    for r in range(nn):
        blk = np.ascontiguousarray(val[16 * diag[r]:16 * diag[r] + 16])
        pv = np.zeros(4, np.int32)
        assert o.orc_lu4(P(blk), P(pv)) == 0
        lu[16 * r:16 * r + 16] = blk
        piv[4 * r:4 * r + 4] = pv
    return diag, lu, piv


class BlockSystem:
    """a matrix of variable-size blocks packed as the oracle's block ILU
    reads it: row i has its diagonal block (ns[i] x ns[i]) and blocks a_ij
    for the neighbours j in nbrs[i] (listed in the given order), every block
    column-major"""

    def __init__(self, ns, nbrs, diag, off):
        """diag[i]: array [ns_i, ns_i] (row, column); off[i][q]: array
        [ns_i, ns_j] for j = nbrs[i][q]"""
        self.nb = len(ns)
        self.ns = np.asarray(ns, np.int32)
        self.xoff = np.r_[0, np.cumsum(self.ns)].astype(np.int64)
        self.nptr = np.r_[0, np.cumsum([len(x) for x in nbrs])].astype(np.int32)
        self.nbr = np.array([j for x in nbrs for j in x], np.int32)
        parts, pos = [], 0
        self.doff = np.zeros(self.nb, np.int64)
        self.ooff = np.zeros(max(1, len(self.nbr)), np.int64)
        for i in range(self.nb):
            self.doff[i] = pos
            parts.append(np.asarray(diag[i], float).T.ravel())
            pos += self.ns[i] ** 2
            for q, j in enumerate(nbrs[i]):
                self.ooff[self.nptr[i] + q] = pos
                parts.append(np.asarray(off[i][q], float).T.ravel())
                pos += self.ns[i] * self.ns[j]
        self.val = np.ascontiguousarray(np.concatenate(parts))

    def block(self, val, i, j):
        """block (i, j) of a packed value array as [rows, cols]"""
        ni = self.ns[i]
        if i == j:
            o = self.doff[i]
            return val[o:o + ni * ni].reshape(ni, ni).T
        p = next(p for p in range(self.nptr[i], self.nptr[i + 1]) if self.nbr[p] == j)
        nj = self.ns[j]
        o = self.ooff[p]
        return val[o:o + ni * nj].reshape(nj, ni).T

    def dense(self, val=None):
        val = self.val if val is None else val
        n = int(self.xoff[-1])
        A = np.zeros((n, n))
        for i in range(self.nb):
            ri = slice(self.xoff[i], self.xoff[i + 1])
            A[ri, ri] = self.block(val, i, i)
            for p in range(self.nptr[i], self.nptr[i + 1]):
                j = self.nbr[p]
                A[ri, self.xoff[j]:self.xoff[j + 1]] = self.block(val, i, j)
        return A

    def ilu0(self):
        """the oracle's block ILU(0) factors (packed like val) and its return code"""
        f = self.val.copy()
        rc = oracle().orc_block_ilu0(self.nb, P(self.ns), P(self.nptr), P(self.nbr), P(self.doff),
                                     P(self.ooff), P(f))
        return f, rc

    def ilu_solve(self, f, b):
        v = np.zeros(int(self.xoff[-1]))
        oracle().orc_block_ilu_solve(self.nb, P(self.ns), P(self.xoff), P(self.nptr), P(self.nbr),
                                     P(self.doff), P(self.ooff), P(f),
                                     P(np.ascontiguousarray(b, float)), P(v))
        return v

    def factors_dense(self, f):
        """dense L (unit lower) and U (upper, diagonal blocks inv(D_i)^-1) of
        packed factors"""
        n = int(self.xoff[-1])
        L, U = np.eye(n), np.zeros((n, n))
        for i in range(self.nb):
            ri = slice(self.xoff[i], self.xoff[i + 1])
            U[ri, ri] = np.linalg.inv(self.block(f, i, i))
            for p in range(self.nptr[i], self.nptr[i + 1]):
                j = self.nbr[p]
                rj = slice(self.xoff[j], self.xoff[j + 1])
                (L if j < i else U)[ri, rj] = self.block(f, i, j)
        return L, U


def bcrs_iluk(nn, rp, ci, val, k):
    """the oracle's block ILU(k) on BCRS storage: the grown pattern (R's list
    capacity from the most off-diagonal blocks of a row), values copied onto
    it (zero for fill), the diagonal positions and the factors"""
    o = oracle()
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    max_off = int(max(rp[r + 1] - rp[r] - 1 for r in range(nn)))
    cap = int(0.5 * (k + 2) ** 3 * max_off)
    rp2 = np.zeros(nn + 1, np.int32)
    n2 = o.orc_bcrs_iluk_pattern(nn, P(rp), P(ci), k, cap, P(rp2), None)
    assert n2 > 0
    ci2 = np.zeros(n2, np.int32)
    o.orc_bcrs_iluk_pattern(nn, P(rp), P(ci), k, cap, P(rp2), P(ci2))
    ilu = np.zeros(16 * n2)
    for r in range(nn):
        src = {int(ci[s]): s for s in range(rp[r], rp[r + 1])}
        for s2 in range(rp2[r], rp2[r + 1]):
            s = src.get(int(ci2[s2]))
            if s is not None:
                ilu[16 * s2:16 * s2 + 16] = val[16 * s:16 * s + 16]
    diag = np.array([next(s for s in range(rp2[r], rp2[r + 1]) if ci2[s] == r) for r in range(nn)],
                    np.int32)
    assert o.orc_bcrs_ilu0(nn, P(rp2), P(ci2), P(diag), P(ilu)) == 0
    return rp2, ci2, diag, ilu
