"""The oracle's DG prism element integral (apd_dg_prism), pinned three ways:
- shape functions and the prism quadrature against the reference's own
  apr_shape_fun_3D / apr_set_quadr_3D compiled into oracle/_ref
  (aps_dg_util.c:408-1496, aps_gauss_util.c:27-150), bit-exact for the
  monomials;
- the element matrices on affine prisms against closed-form integrals of
  monomials over the reference prism (exact: the rule's degree covers every
  integrand), with exponents read off the shape functions themselves;
- the Jacobian, inverse and derivative transformation on distorted prisms
  against the reference's apr_elem_calc_3D (aps_dg_util.c:112-368).
"""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from oracle_lib import P, dg_num_int_el, dg_quadr, dg_shape, ref, ref_available  # noqa: E402

PDEGS = [101, 102, 103, 201, 202, 203, 301, 302, 303]
# R's pz = 0 tensor degrees (pdeg 1, 2) and the complete base of degree
# 1..3 (the oracle's code DG_COMPLETE + p)
EXTRA = [1, 2, 1001, 1002, 1003]
REF_NODES = np.array([[0, 0, -1], [1, 0, -1], [0, 1, -1], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float)
need_dg = pytest.mark.skipif(not ref_available("libref_dg.so"), reason="oracle/_ref not built")
TENSOR = 1  # APC_BASE_TENSOR_DG, include/aph_intf.h:122
COMPLETE = 2  # APC_BASE_COMPLETE_DG, include/aph_intf.h:123


def rbase(pdeg):
    """(base type, pdeg) of a degree code as the reference takes them"""
    return (COMPLETE, pdeg - 1000) if pdeg > 1000 else (TENSOR, pdeg)


def rule_degrees(pdeg):
    """(px, pz) of the tensor rule a degree code integrates with"""
    return (pdeg - 1000, pdeg - 1000) if pdeg > 1000 else (pdeg % 100, pdeg // 100)

# slots of the coefficient-presence vector and of the coefficient vector
IND_M, IND_AXX, IND_AZZ, IND_BX, IND_BZ, IND_C, IND_QX, IND_QZ, IND_S = 1, 2, 10, 11, 13, 17, 19, 21, 22
C_A, C_B, C_C, C_Q, C_S, C_M = 0, 9, 15, 16, 19, 20


def ref_dg():
    L = ref("libref_dg.so")
    L.apr_shape_fun_3D.restype = ctypes.c_int
    L.apr_shape_fun_3D.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5
    L.apr_set_quadr_3D.restype = ctypes.c_int
    L.apr_set_quadr_3D.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 4
    L.apr_elem_calc_3D.restype = ctypes.c_double
    L.apr_elem_calc_3D.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int] + \
        [ctypes.c_void_p] * 13
    L.apr_gauss_init.restype = None
    L.apr_gauss_init()
    return L


def ref_shape(L, pdeg, eta):
    out = [np.zeros(200) for _ in range(4)]
    e = np.ascontiguousarray(eta, float)
    ns = L.apr_shape_fun_3D(*rbase(pdeg), P(e), *[P(a) for a in out])
    return [a[:ns] for a in out]


def exponents(pdeg):
    """(a, b, c) of each shape function e0^a e1^b z^c, read off its value at
    (2, 3, 5), where every monomial takes a distinct value"""
    phi = dg_shape(pdeg, [2.0, 3.0, 5.0])[0]
    ex = []
    for v in phi:
        v = int(round(v))
        abc = []
        for pr in (2, 3, 5):
            k = 0
            while v % pr == 0:
                v //= pr
                k += 1
            abc.append(k)
        assert v == 1
        ex.append(tuple(abc))
    return ex


def mono_int(a, b, c):
    """integral of e0^a e1^b z^c over {e0, e1 >= 0, e0 + e1 <= 1} x [-1, 1]"""
    if c % 2:
        return 0.0
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2) * 2.0 / (c + 1)


@need_dg
@pytest.mark.parametrize("pdeg", PDEGS + EXTRA)
def test_shape_functions_match_reference(pdeg):
    L = ref_dg()
    rng = np.random.default_rng(pdeg)
    for _ in range(8):
        eta = np.array([rng.uniform(0, 0.6), rng.uniform(0, 0.4), rng.uniform(-1, 1)])
        mine, theirs = dg_shape(pdeg, eta), ref_shape(L, pdeg, eta)
        assert len(mine[0]) == len(theirs[0])
        for m, t in zip(mine, theirs):
            assert np.array_equal(m, t)
    from oracle_lib import oracle
    assert len(mine[0]) == oracle().orc_dg_ns(pdeg)
    if pdeg > 1000:
        p = pdeg - 1000
        assert len(mine[0]) == (p + 1) * (p + 2) * (p + 3) // 6
    else:
        px, pz = pdeg % 100, pdeg // 100
        assert len(mine[0]) == (px + 1) * (px + 2) // 2 * (pz + 1)


@pytest.mark.parametrize("pdeg", EXTRA)
def test_extra_bases_are_the_expected_monomials(pdeg):
    """pz = 0: the 2D monomials of degree <= px, constant in z; complete:
    every monomial of total degree <= p, each once"""
    ex = exponents(pdeg)
    assert len(set(ex)) == len(ex)
    if pdeg > 1000:
        p = pdeg - 1000
        want = {(a, b, c) for a in range(p + 1) for b in range(p + 1) for c in range(p + 1)
                if a + b + c <= p}
    else:
        want = {(a, b, 0) for a in range(pdeg + 1) for b in range(pdeg + 1) if a + b <= pdeg}
    assert set(ex) == want


@need_dg
@pytest.mark.parametrize("pdeg", PDEGS + EXTRA)
def test_quadrature_matches_reference(pdeg):
    L = ref_dg()
    xg, wg = np.zeros(3000), np.zeros(1000)
    ng = ctypes.c_int(0)
    base, rp = rbase(pdeg)
    pv = np.array([rp], np.int32)
    L.apr_set_quadr_3D(base, P(pv), ctypes.byref(ng), P(xg), P(wg))
    x, w = dg_quadr(pdeg)
    assert len(w) == ng.value
    # 15-digit literals in the reference, Newton-computed Gauss-Legendre here
    assert np.abs(x - xg[:3 * ng.value].reshape(-1, 3)).max() <= 2e-15
    assert np.abs(w - wg[:ng.value]).max() <= 5e-15


@pytest.mark.parametrize("pdeg", PDEGS + EXTRA)
def test_quadrature_is_exact_to_its_degree(pdeg):
    x, w = dg_quadr(pdeg)
    px, pz = rule_degrees(pdeg)
    for a in range(2 * px + 1):
        for b in range(2 * px + 1 - a):
            for c in range(max(2 * pz, 1) + 1):
                q = np.sum(w * x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c)
                assert abs(q - mono_int(a, b, c)) <= 1e-13, (a, b, c)


def all_terms(rng, full=True):
    ind = np.zeros(23, np.int32)
    ind[[IND_M, IND_AXX, IND_BX, IND_C, IND_QX, IND_S]] = 1
    if full:
        ind[[IND_AZZ, IND_BZ, IND_QZ]] = 1
    coef = np.zeros(21)
    coef[C_A:C_A + 9] = rng.uniform(-0.3, 0.3, 9) + np.eye(3).ravel() * 2.0
    coef[C_B:C_B + 3] = rng.uniform(-1, 1, 3)
    coef[C_C], coef[C_M], coef[C_S] = rng.uniform(0.5, 1.5, 3)
    coef[C_Q:C_Q + 3] = rng.uniform(-1, 1, 3)
    return ind, coef


@pytest.mark.parametrize("pdeg", PDEGS + EXTRA)
@pytest.mark.parametrize("full", [True, False])
def test_affine_prism_matches_closed_form(pdeg, full):
    rng = np.random.default_rng(pdeg + 7 * full)
    A = np.eye(3) * rng.uniform(0.5, 2.0, 3) + rng.uniform(-0.3, 0.3, (3, 3))
    x0 = rng.uniform(-1, 1, 3)
    nc = x0 + REF_NODES @ A.T
    ind, coef = all_terms(rng, full)
    K, F = dg_num_int_el(pdeg, nc, ind, coef)
    ex = exponents(pdeg)
    ns = len(ex)
    det = abs(np.linalg.det(A))
    Ai = np.linalg.inv(A)  # Ai[k, c] = d eta_k / d x_c
    a = coef[C_A:C_A + 9].reshape(3, 3)  # a[r, c]: d_c on trial, d_r on test
    b, q = coef[C_B:C_B + 3], coef[C_Q:C_Q + 3]
    if not full:  # the 2D branches: a_xx, a_xy, a_yx, a_yy and b_x, b_y, q_x, q_y only
        a = a.copy()
        a[2, :] = a[:, 2] = 0.0
        b, q = b.copy(), q.copy()
        b[2] = q[2] = 0.0
    G = Ai @ a.T @ Ai.T  # G[k, l] multiplies d_k phi_j d_l phi_i

    def deriv(e, k):  # d/d eta_k of a monomial: (coefficient, exponents)
        e = list(e)
        c = e[k]
        e[k] = max(c - 1, 0)
        return c, tuple(e)

    def integ(c1, e1, c2, e2):
        if c1 == 0 or c2 == 0:
            return 0.0
        return c1 * c2 * mono_int(e1[0] + e2[0], e1[1] + e2[1], e1[2] + e2[2])

    Kx = np.zeros((ns, ns))
    Fx = np.zeros(ns)
    for j in range(ns):
        for i in range(ns):
            mij = mono_int(*[ex[i][t] + ex[j][t] for t in range(3)])
            v = (coef[C_M] + coef[C_C]) * mij
            for k in range(3):
                ck, ek = deriv(ex[j], k)
                for l in range(3):
                    cl, el = deriv(ex[i], l)
                    v += G[k, l] * integ(ck, ek, cl, el)
            bt = Ai @ b  # sum_k b_k d_k = sum_l (Ai b)_l d eta_l
            for l in range(3):
                cl, el = deriv(ex[i], l)
                v += bt[l] * integ(1, ex[j], cl, el)
            Kx[j, i] = v * det
    qt = Ai @ q
    for i in range(ns):
        f = coef[C_S] * mono_int(*ex[i])
        for l in range(3):
            cl, el = deriv(ex[i], l)
            f += qt[l] * (cl * mono_int(*el) if cl else 0.0)
        Fx[i] = f * det
    assert np.abs(K - Kx).max() <= 1e-12 * np.abs(Kx).max()
    assert np.abs(F - Fx).max() <= 1e-12 * np.abs(Fx).max()


@need_dg
@pytest.mark.parametrize("pdeg", [101, 202, 303, 102, 2, 1003])
def test_distorted_prism_geometry_matches_reference(pdeg):
    """on a prism whose quads are not planar the Jacobian varies; the
    oracle's matrix equals the one built from the reference's own
    apr_elem_calc_3D values at the reference's quadrature points"""
    L = ref_dg()
    rng = np.random.default_rng(pdeg + 1)
    nc = REF_NODES + rng.uniform(-0.15, 0.15, (6, 3))
    ind, coef = all_terms(rng, True)
    K, F = dg_num_int_el(pdeg, nc, ind, coef)
    xg, wg = np.zeros(3000), np.zeros(1000)
    ng = ctypes.c_int(0)
    base, rp = rbase(pdeg)
    pv = np.array([rp], np.int32)
    L.apr_set_quadr_3D(base, P(pv), ctypes.byref(ng), P(xg), P(wg))
    ncc = np.ascontiguousarray(nc.ravel())
    ns = K.shape[0]
    Kr, Fr = np.zeros((ns, ns)), np.zeros(ns)
    a = coef[C_A:C_A + 9].reshape(3, 3)
    for g in range(ng.value):
        phi, dx, dy, dz, xc = [np.zeros(200) for _ in range(4)] + [np.zeros(3)]
        eta = np.ascontiguousarray(xg[3 * g:3 * g + 3])
        det = L.apr_elem_calc_3D(2, 1, P(pv), base, P(eta), P(ncc), None, P(phi), P(dx), P(dy),
                                 P(dz), P(xc), None, None, None, None, None)
        vol = det * wg[g]
        grad = np.stack([dx[:ns], dy[:ns], dz[:ns]])  # grad[c, j]
        ph = phi[:ns]
        Kr += (np.einsum("rc,cj,ri->ji", a, grad, grad) + (coef[C_M] + coef[C_C]) * np.outer(ph, ph)
               + np.outer(ph, coef[C_B:C_B + 3] @ grad)) * vol
        Fr += (coef[C_Q:C_Q + 3] @ grad + coef[C_S] * ph) * vol
    assert np.abs(K - Kr).max() <= 1e-12 * np.abs(Kr).max()
    assert np.abs(F - Fr).max() <= 1e-12 * np.abs(Fr).max()


# ---------------------------------------------------------------- faces
GOLDEN_DG = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dg")


def read_error_table():
    """the reference's NIPG convergence table for DIFF_IN_CUBE
    (ctest/conv_diff/Laplace_dg_in_cube/comparison_data/diff_nipg_h_conv.txt):
    {(n, pdeg): (ndof, L2 error, H1-seminorm error)} with h = 1/n"""
    out = {}
    with open(os.path.join(GOLDEN_DG, "diff_nipg_h_conv.txt")) as f:
        for line in f:
            t = line.replace(",", ".").split()
            try:
                vals = [float(v) for v in t[:5]]
            except ValueError:
                continue
            if len(vals) < 5:
                continue
            n = int(round(1.0 / float(t[0])))
            out[(n, int(t[1]))] = (int(t[2]), float(t[3]), float(t[4]))
    return out


def cube_arrays(n, jitter=0.0, seed=0):
    """unit cube, n^3 hexes split into prisms along the (0,0)-(1,1) diagonal
    (the reference's two-prism cube mesh refined log2 n times), bottom and
    top faces Neumann (bc flag 11), sides Dirichlet (flag 1), as in the
    reference's mesh file for this test"""
    import modfem_amd as mf
    a = mf.Mesh.box_prism(n, n, n, seed=seed, jitter=jitter).arrays()
    bc = a["bf_bc"].copy()
    bc[a["bf_side"] < 2] = 11
    a["bf_bc"] = bc
    return a


def test_face_points_coincide_across_neighbours():
    """a face point mapped into the second neighbour is the same physical
    point (the check the reference itself makes, weakform_dg.c:605-612)"""
    from oracle_lib import dg_faces, oracle
    a = cube_arrays(3, jitter=0.2, seed=4)
    nodes = a["coor"][a["el_nodes"]]
    o = oracle()
    nfaces = 0
    for (e1, s1, e2, s2) in dg_faces(a["el_nodes"]):
        if e2 < 0:
            continue
        nfaces += 1
        for q in ([0.2, 0.3], [-0.5, 0.7], [0.1, 0.1]):
            eta, eta2, n, x1, x2 = [np.zeros(3) for _ in range(5)]
            o.orc_dg_face_eta(s1, P(np.array(q)), P(eta))
            rc = o.orc_dg_map_point(s1, P(np.ascontiguousarray(a["el_nodes"][e1], np.int32)), s2,
                                    P(np.ascontiguousarray(a["el_nodes"][e2], np.int32)),
                                    P(eta), P(eta2))
            assert rc == 0
            o.orc_dg_face_ds(s1, P(np.ascontiguousarray(nodes[e1].ravel())), P(eta), P(n), P(x1))
            from oracle_lib import dg_shape  # noqa: F401
            n2 = np.zeros(3)
            o.orc_dg_face_ds(s2, P(np.ascontiguousarray(nodes[e2].ravel())), P(eta2), P(n2), P(x2))
            assert np.abs(x1 - x2).max() <= 1e-14
            if s1 < 2:  # planar faces: opposite outward normals
                assert np.abs(n + n2).max() <= 1e-13
    assert nfaces > 50


@need_dg
@pytest.mark.parametrize("pdeg", PDEGS + EXTRA)
def test_face_quadrature_matches_reference(pdeg):
    L = ref_dg()
    L.apr_set_quadr_2D.restype = ctypes.c_int
    L.apr_set_quadr_2D.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4
    from oracle_lib import oracle
    for fa_type, quad in ((3, 0), (4, 1)):  # MMC_TRIA, MMC_QUAD (include/mmh_intf.h:224)
        xg, wg = np.zeros(450), np.zeros(150)
        ng = ctypes.c_int(0)
        base, rp = rbase(pdeg)
        L.apr_set_quadr_2D(fa_type, base, P(np.array([rp], np.int32)), ctypes.byref(ng),
                           P(xg), P(wg))
        mx, mw = np.zeros(128), np.zeros(64)
        n = oracle().orc_dg_face_quadr(pdeg, quad, P(mx), P(mw))
        assert n == ng.value
        assert np.abs(mx[:2 * n] - xg[:2 * n]).max() <= 2e-15
        assert np.abs(mw[:n] - wg[:n]).max() <= 5e-15


@need_dg
def test_face_area_element_and_normal_match_reference():
    """apr_elem_calc_3D with Control = 3 + side returns |ds| and the unit
    normal (aps_dg_util.c:334-367)"""
    from oracle_lib import oracle
    L = ref_dg()
    rng = np.random.default_rng(5)
    nc = np.ascontiguousarray((REF_NODES + rng.uniform(-0.15, 0.15, (6, 3))).ravel())
    pv = np.array([202], np.int32)
    for side in range(5):
        for q in ([0.2, 0.3], [-0.4, 0.6]):
            eta = np.zeros(3)
            oracle().orc_dg_face_eta(side, P(np.array(q)), P(eta))
            n, x = np.zeros(3), np.zeros(3)
            d = oracle().orc_dg_face_ds(side, P(nc), P(eta), P(n), P(x))
            phi, dx, dy, dz = [np.zeros(200) for _ in range(4)]
            nr, xr = np.zeros(3), np.zeros(3)
            dr = L.apr_elem_calc_3D(3 + side, 1, P(pv), TENSOR, P(eta), P(nc), None, P(phi),
                                    P(dx), P(dy), P(dz), P(xr), None, None, None, None, P(nr))
            assert abs(d - dr) <= 1e-15 * abs(dr) and np.array_equal(n, nr)
            assert np.abs(x - xr).max() <= 1e-15


@pytest.mark.parametrize("n,pdeg", [(1, 101), (1, 202), (1, 303), (2, 101), (2, 202), (2, 303),
                                    (4, 101), (4, 202), (4, 303), (8, 101)])
def test_reference_error_table_reproduced(n, pdeg):
    """element integrals + NIPG faces (Dirichlet sides, Neumann top and
    bottom, penalty constant 1) solved directly reproduce every printed digit
    of the reference's DIFF_IN_CUBE convergence table"""
    import scipy.sparse.linalg as spl
    import modfem_amd as mf
    from oracle_lib import dg_errors, dg_system_dense
    table = read_error_table()
    ndof, l2_ref, h1_ref = table[(n, pdeg)]
    a = cube_arrays(n)
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    A, b = dg_system_dense(a, pdeg, mf.IND_DIFF, lambda e: cc, 1.0, 1, sparse=True)
    assert len(b) == ndof
    x = spl.spsolve(A.tocsc(), b)
    l2, h1 = dg_errors(a, pdeg, x, 1)
    assert abs(l2 - l2_ref) <= 6e-10 and abs(h1 - h1_ref) <= 6e-10, (l2, h1, l2_ref, h1_ref)


def mixed_degrees(nel, seed):
    return np.random.default_rng(seed).choice([101, 202, 303], size=nel)


@pytest.mark.parametrize("seed", [1, 2])
def test_mixed_degree_system_reproduces_a_linear_solution(seed):
    """elements of degree 1, 2 and 3 side by side: a linear exact solution
    (in every element's space) with its Dirichlet data is reproduced to
    rounding on an affine prism mesh.  Faces between degrees need the rule
    of the higher one: with the lower one's rule the degree-3 test functions
    are integrated inexactly and the solution is off by far more"""
    import modfem_amd as mf
    from oracle_lib import dg_degrees, dg_system_dense, oracle
    a = mf.Mesh.box_prism(3, 2, 2, seed=seed).arrays()  # all sides Dirichlet
    nel = a["el_nodes"].shape[0]
    pd = mixed_degrees(nel, seed)
    assert len(set(pd)) == 3
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    A, b = dg_system_dense(a, pd, mf.IND_DIFF, lambda e: cc, 1.0, 9)
    x = np.linalg.solve(A, b)
    _, ns, off = dg_degrees(pd, nel)
    nodes = a["coor"][a["el_nodes"]]
    o = oracle()
    rng = np.random.default_rng(seed)
    worst = 0.0
    for e in range(nel):
        for _ in range(4):
            u, v = rng.uniform(0, 0.5, 2)
            eta = np.array([u, v, rng.uniform(-1, 1)])
            xp = np.zeros(3)
            uh = o.orc_dg_eval(int(pd[e]), P(np.ascontiguousarray(nodes[e].ravel())),
                               P(np.ascontiguousarray(x[off[e]:off[e + 1]])), P(eta), P(xp))
            worst = max(worst, abs(uh - (0.3 + xp[0] - 2 * xp[1] + 0.5 * xp[2])))
    assert worst <= 1e-11, worst


def test_mixed_degree_system_does_not_depend_on_the_first_neighbour():
    """the face rule (more points wins) and the penalty (smaller degree) are
    symmetric in the two neighbours: assembling every interior face with
    the higher element first gives the same system on planar faces (on
    warped quad faces the face size sqrt|F| depends on the first
    neighbour's vertex cycle, as in the reference)"""
    import modfem_amd as mf
    from oracle_lib import dg_system_dense
    a = mf.Mesh.box_prism(3, 2, 2, seed=3).arrays()
    nel = a["el_nodes"].shape[0]
    pd = mixed_degrees(nel, 3)
    rng = np.random.default_rng(3)
    coefs = [np.r_[np.eye(3).ravel() * rng.uniform(0.5, 2), rng.uniform(-1, 1, 3), np.zeros(9)]
             for _ in range(nel)]
    ind = mf.coeff_indicator("axx", "ayy", "azz", "bx", "by", "bz", "s")
    A1, b1 = dg_system_dense(a, pd, ind, lambda e: coefs[e], 1.5, 1)
    A2, b2 = dg_system_dense(a, pd, ind, lambda e: coefs[e], 1.5, 1, swap=True)
    assert np.abs(A1 - A2).max() <= 1e-12 * np.abs(A1).max()
    assert np.abs(b1 - b2).max() <= 1e-12 * np.abs(b1).max()


def _face_geometry(coor, nodes, side):
    """area and unit normal (pointing out of the element) of a planar face"""
    from oracle_lib import DG_FACE_V
    X = coor[nodes[DG_FACE_V[side]]]
    if len(X) == 3:
        nv = 0.5 * np.cross(X[1] - X[0], X[2] - X[0])
    else:
        nv = 0.5 * np.cross(X[2] - X[0], X[3] - X[1])
    area = np.linalg.norm(nv)
    n = nv / area
    if n @ (X.mean(0) - coor[nodes].mean(0)) < 0:
        n = -n
    return area, n


@pytest.mark.parametrize("pdeg", [101, 202, 303])
def test_upwind_convection_energy_identity(pdeg):
    """the convection part of the DG form (volume term plus the upwind face
    terms, pds_conv_diff_weakform_dg.c:819-881, and the Dirichlet inflow /
    outflow faces, :1114-1160) with constant b obeys the upwind energy
    identity B(w, w) = 1/2 sum_interior |b.n| int [w]^2 + 1/2 int_boundary
    |b.n| w^2.  For w constant per element it is a closed form in the face
    areas and normals of an affine mesh.  The downwind side, planted in the
    oracle, flips the sign of the interior part and fails"""
    import modfem_amd as mf
    from oracle_lib import dg_faces, dg_shape, dg_system_dense, oracle
    m = mf.Mesh.box_prism(3, 3, 2, lx=1.0, ly=1.3, lz=0.7, seed=4, jitter=0.0)
    a = m.arrays()
    coor, en = a["coor"], a["el_nodes"]
    nel = en.shape[0]
    rng = np.random.default_rng(pdeg)
    b = rng.uniform(-1, 1, 3)
    coef = np.r_[np.zeros(9), b, np.zeros(9)]
    ind = mf.coeff_indicator("bx", "by", "bz")
    # the constant function in the element base
    pts = np.random.default_rng(0).uniform(0, 0.45, (60, 3))  # inside the reference prism
    pts[:, 2] = 4 * pts[:, 2] - 0.9
    Phi = np.array([dg_shape(pdeg, p)[0] for p in pts])
    assert np.linalg.matrix_rank(Phi) == Phi.shape[1]
    one = np.linalg.lstsq(Phi, np.ones(len(pts)), rcond=None)[0]
    assert np.abs(Phi @ one - 1).max() <= 1e-13
    c = rng.uniform(-1, 1, nel)
    x = np.concatenate([ci * one for ci in c])
    expect = 0.0
    for (e1, s1, e2, s2) in dg_faces(en):
        area, n = _face_geometry(coor, en[e1], s1)
        jump2 = (c[e1] - c[e2]) ** 2 if e2 >= 0 else c[e1] ** 2
        expect += 0.5 * abs(b @ n) * area * jump2
    A, _ = dg_system_dense(a, pdeg, ind, lambda e: coef, penalty_const=0.0)
    got = x @ A @ x
    assert abs(got - expect) <= 1e-12 * abs(expect), (got, expect)
    try:
        oracle().orc_dg_set_plant_downwind(1)
        Ad, _ = dg_system_dense(a, pdeg, ind, lambda e: coef, penalty_const=0.0)
    finally:
        oracle().orc_dg_set_plant_downwind(0)
    assert abs(x @ Ad @ x - expect) > 0.1 * abs(expect)
    m.free()


# ---------------------------------------------- complete base, pz = 0 degrees
def index_map(sub, full):
    """positions in base `full` of the functions of base `sub` (by exponents)"""
    ef = exponents(full)
    return np.array([ef.index(t) for t in exponents(sub)])


@pytest.mark.parametrize("p", [1, 2, 3])
def test_complete_system_is_the_tensor_system_on_its_functions(p):
    """the complete base of degree p is a subset of the tensor base (p, p)
    and takes the same volume and face rules and penalty degree, so its
    assembled system (every element and face term, distorted mesh, Dirichlet
    and Neumann faces) equals the tensor system's rows and columns of those
    functions; a wrong rule, penalty or function would show here"""
    import modfem_amd as mf
    from oracle_lib import dg_system_dense, oracle
    a = cube_arrays(2, jitter=0.15, seed=p)
    nel = a["el_nodes"].shape[0]
    rng = np.random.default_rng(p)
    ind, _ = all_terms(rng, True)
    ce = [all_terms(rng, True)[1] for _ in range(nel)]
    Ac, bc = dg_system_dense(a, 1000 + p, ind, lambda e: ce[e], 1.5, 1)
    At, bt = dg_system_dense(a, 101 * p, ind, lambda e: ce[e], 1.5, 1)
    o = oracle()
    nc, nt = o.orc_dg_ns(1000 + p), o.orc_dg_ns(101 * p)
    m = index_map(1000 + p, 101 * p)
    idx = (np.arange(nel)[:, None] * nt + m[None, :]).ravel()
    assert len(bc) == nel * nc
    sub = At[np.ix_(idx, idx)]
    assert np.abs(Ac - sub).max() <= 1e-13 * np.abs(sub).max()
    assert np.abs(bc - bt[idx]).max() <= 1e-13 * np.abs(bt).max()
    assert o.orc_dg_penalty_degree(1000 + p, -1) == p
    del mf


@pytest.mark.parametrize("px", [1, 2])
def test_flat_degree_system_is_the_tensor_system_on_its_functions(px):
    """pz = 0 (pdeg 1, 2): on a mesh extruded straight up (every Jacobian
    constant in z) the one-point rule in z integrates the z-constant products
    exactly, so the system equals the tensor (px, 1) system's rows and columns
    of the z-constant functions once both use the penalty degree R gives the
    pz = 0 degrees, min(px, pz) = 0 (constant source and boundary data)"""
    import modfem_amd as mf
    from oracle_lib import dg_system_dense, oracle
    a = mf.Mesh.box_prism(2, 2, 2, seed=px).arrays()
    nel = a["el_nodes"].shape[0]
    rng = np.random.default_rng(10 + px)
    ind, _ = all_terms(rng, True)
    ce = [all_terms(rng, True)[1] for _ in range(nel)]
    o = oracle()
    assert o.orc_dg_penalty_degree(px, -1) == 0 and o.orc_dg_penalty_degree(px, 101) == 0
    # constant source and boundary data: the load integrands are z-constant too
    Af, bf = dg_system_dense(a, px, ind, lambda e: ce[e], 1.5, 0, 0.7)
    At, bt = dg_system_dense(a, 100 + px, ind, lambda e: ce[e], 0.0, 0, 0.7)
    nf, nt = o.orc_dg_ns(px), o.orc_dg_ns(100 + px)
    m = index_map(px, 100 + px)
    idx = (np.arange(nel)[:, None] * nt + m[None, :]).ravel()
    sub = At[np.ix_(idx, idx)]
    assert len(bf) == nel * nf
    assert np.abs(Af - sub).max() <= 1e-12 * np.abs(sub).max()
    assert np.abs(bf - bt[idx]).max() <= 1e-12 * np.abs(bt).max()


@pytest.mark.parametrize("pd", [1001, 1002, 1003, "mixed"])
def test_complete_base_reproduces_a_linear_solution(pd):
    """every complete degree holds the linear functions: the linear exact
    solution is reproduced to rounding, uniform and with degrees mixed"""
    import modfem_amd as mf
    from oracle_lib import dg_degrees, dg_system_dense, oracle
    a = mf.Mesh.box_prism(3, 2, 2, seed=4).arrays()
    nel = a["el_nodes"].shape[0]
    pdv = (np.random.default_rng(5).choice([1001, 1002, 1003], size=nel) if pd == "mixed"
           else np.full(nel, pd))
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    A, b = dg_system_dense(a, pdv, mf.IND_DIFF, lambda e: cc, 1.0, 9)
    x = np.linalg.solve(A, b)
    _, ns, off = dg_degrees(pdv, nel)
    nodes = a["coor"][a["el_nodes"]]
    o = oracle()
    rng = np.random.default_rng(6)
    worst = 0.0
    for e in range(nel):
        for _ in range(4):
            u, v = rng.uniform(0, 0.5, 2)
            eta = np.array([u, v, rng.uniform(-1, 1)])
            xp = np.zeros(3)
            uh = o.orc_dg_eval(int(pdv[e]), P(np.ascontiguousarray(nodes[e].ravel())),
                               P(np.ascontiguousarray(x[off[e]:off[e + 1]])), P(eta), P(xp))
            worst = max(worst, abs(uh - (0.3 + xp[0] - 2 * xp[1] + 0.5 * xp[2])))
    assert worst <= 1e-11, worst
