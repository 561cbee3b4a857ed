"""The oracle's DG on tetrahedra (no reference counterpart: the reference's
DG module is prism-only, apd_dg_prism/aph_dg_prism.h:1-3), pinned by the
mathematics:
- the volume rule integrates every monomial of degree <= 2p exactly
  (closed form a! b! c! / (a + b + c + 3)! on the reference tet);
- element matrices on an affine tet equal closed-form monomial integrals
  for every term (mass, full anisotropic diffusion, convection with the
  derivative on the test function, reaction, source);
- face points seen from both neighbours (tet-tet, tet-prism) are the same
  physical points;
- on affine hybrid meshes (prism and tet layers, degrees 1-3, uniform and
  mixed) a linear exact solution with its Dirichlet data is reproduced to
  rounding (NIPG consistency), which a wrong normal, map, face rule or
  penalty breaks."""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from oracle_lib import P, oracle  # noqa: E402


def ref_integral(a, b, c):
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


def exps(p):
    ex = np.zeros(64, np.int32)
    n = oracle().orc_dgt_exponents(p, P(ex))
    return [(int(v) & 3, (int(v) >> 2) & 3, (int(v) >> 4) & 3) for v in ex[:n]]


@pytest.mark.parametrize("p", [1, 2, 3])
def test_base_is_the_complete_monomial_base(p):
    o = oracle()
    ex = exps(p)
    assert len(ex) == o.orc_dgt_ns(p) == (p + 1) * (p + 2) * (p + 3) // 6
    assert len(set(ex)) == len(ex) and all(sum(t) <= p for t in ex)
    eta = np.array([0.21, 0.33, 0.17])
    phi, dx, dy, dz = (np.zeros(32) for _ in range(4))
    n = o.orc_dgt_shape(p, P(eta), P(phi), P(dx), P(dy), P(dz))
    for i, (a, b, c) in enumerate(ex[:n]):
        assert phi[i] == pytest.approx(eta[0] ** a * eta[1] ** b * eta[2] ** c, rel=1e-15, abs=0)
        d = a * eta[0] ** max(a - 1, 0) * eta[1] ** b * eta[2] ** c
        assert dx[i] == pytest.approx(d, rel=1e-15, abs=1e-300)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_volume_rule_is_exact_to_degree_2p(p):
    xg, wg = np.zeros(3 * 128), np.zeros(128)
    ng = oracle().orc_dgt_quadr(p, P(xg), P(wg))
    x = xg[:3 * ng].reshape(ng, 3)
    w = wg[:ng]
    assert ng == (p + 2) * (p + 1) ** 2
    assert np.all(w > 0) and np.all(x >= 0) and np.all(x.sum(1) <= 1)
    for deg in range(2 * p + 1):
        for a in range(deg + 1):
            for b in range(deg + 1 - a):
                c = deg - a - b
                q = np.sum(w * x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c)
                assert abs(q - ref_integral(a, b, c)) <= 1e-15, (a, b, c)


def closed_form_element(p, X, coef):
    """element matrices of an affine tet from exact monomial integrals"""
    ex = exps(p)
    J = np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]]).T
    det = abs(np.linalg.det(J))
    Ji = np.linalg.inv(J)  # reference gradient = Ji^T physical gradient
    A = coef[0:9].reshape(3, 3)
    bb = coef[9:12]
    cr = coef[15]   # reaction c
    s = coef[19]    # source s
    mm = coef[20]   # mass m

    def mono_int(t):
        return ref_integral(*t) if min(t) >= 0 else 0.0

    def grad_terms(e):
        a, b, c = e
        return [(a, (a - 1, b, c)), (b, (a, b - 1, c)), (c, (a, b, c - 1))]

    ns = len(ex)
    K = np.zeros((ns, ns))
    F = np.zeros(ns)
    for i, ei in enumerate(ex):      # test
        for j, ej in enumerate(ex):  # trial
            v = (mm + cr) * mono_int(tuple(np.add(ei, ej)))
            # sum_rc a_rc d_c phi_j d_r phi_i in physical derivatives
            for r in range(3):
                for cc in range(3):
                    for kr, (fi, ti) in enumerate(grad_terms(ei)):
                        for kc, (fj, tj) in enumerate(grad_terms(ej)):
                            if fi == 0 or fj == 0:
                                continue
                            w = A[r, cc] * Ji[kr, r] * Ji[kc, cc]
                            v += w * fi * fj * mono_int(tuple(np.add(ti, tj)))
            for r in range(3):
                for kr, (fi, ti) in enumerate(grad_terms(ei)):
                    if fi:
                        v += bb[r] * Ji[kr, r] * fi * mono_int(tuple(np.add(ti, ej)))
            K[j, i] = v * det
        F[i] = s * mono_int(ei) * det
    return K, F


@pytest.mark.parametrize("p", [1, 2, 3])
def test_element_matrices_match_closed_form_integrals(p):
    from oracle_lib import dg_num_int_el
    import modfem_amd as mf
    rng = np.random.default_rng(p)
    X = np.array([[0.1, 0.0, 0.2], [1.3, 0.2, 0.1], [0.2, 1.1, 0.3], [0.3, 0.2, 0.9]])
    X += rng.uniform(-0.1, 0.1, X.shape)
    coef = np.zeros(21)
    coef[0:9] = rng.uniform(-0.3, 0.3, 9) + np.eye(3).ravel() * 2
    coef[9:12] = rng.uniform(-1, 1, 3)
    coef[15] = 0.7
    coef[19] = 1.3
    coef[20] = 0.4
    ind = mf.coeff_indicator("m", "axx", "axy", "axz", "ayx", "ayy", "ayz", "azx", "azy", "azz",
                             "bx", "by", "bz", "c", "s")
    K, F = dg_num_int_el(100 * p + p, np.r_[X.ravel(), np.zeros(6)], ind, coef, 0, tet=True)
    Kc, Fc = closed_form_element(p, X, coef)
    assert np.abs(K - Kc).max() <= 1e-13 * np.abs(Kc).max()
    assert np.abs(F - Fc).max() <= 1e-13 * np.abs(Fc).max()


def hybrid_arrays(seed, jitter=0.0, n=(2, 2, 4)):
    import modfem_amd as mf
    m = mf.Mesh.box_hybrid(*n, seed=seed, jitter=jitter, tet_frac=0.5)
    a = m.arrays()
    m.free()
    assert (a["el_type"] == 7).any() and (a["el_type"] == 5).any()
    return a


def test_face_points_coincide_from_both_sides():
    from oracle_lib import DG_FACE_V, DGT_FACE_V, dg_faces
    o = oracle()
    a = hybrid_arrays(4, jitter=0.15)
    et, en, X = a["el_type"], a["el_nodes"], a["coor"]
    worst, kinds = 0.0, set()
    xg, wg = np.zeros(128), np.zeros(64)
    ng = o.orc_dg_face_quadr(303, 0, P(xg), P(wg))
    for (e1, s1, e2, s2) in dg_faces(en, et):
        if e2 < 0 or (et[e1] == 5 and et[e2] == 5):
            continue
        t1, t2 = int(et[e1] == 7), int(et[e2] == 7)
        kinds.add((t1, t2))
        nc1 = np.ascontiguousarray(X[en[e1]].ravel())
        nc2 = np.ascontiguousarray(X[en[e2]].ravel())
        g1 = np.ascontiguousarray(en[e1], np.int32)
        g2 = np.ascontiguousarray(en[e2], np.int32)
        for k in range(ng):
            eta, eta2, n1, x1, n2, x2 = (np.zeros(3) for _ in range(6))
            xf = np.ascontiguousarray(xg[2 * k:2 * k + 2])
            if t1:
                o.orc_dgt_face_eta(s1, P(xf), P(eta))
                o.orc_dgt_face_ds(s1, P(nc1), P(eta), P(n1), P(x1))
            else:
                o.orc_dg_face_eta(s1, P(xf), P(eta))
                o.orc_dg_face_ds(s1, P(nc1), P(eta), P(n1), P(x1))
            assert o.orc_dgx_map_point(t1, s1, P(g1), P(eta), t2, s2, P(g2), P(eta2)) == 0
            if t2:
                o.orc_dgt_face_ds(s2, P(nc2), P(eta2), P(n2), P(x2))
            else:
                o.orc_dg_face_ds(s2, P(nc2), P(eta2), P(n2), P(x2))
            worst = max(worst, np.abs(x1 - x2).max())
            assert np.dot(n1, n2) == pytest.approx(-1.0, abs=1e-12)
    assert kinds >= {(1, 1), (0, 1), (1, 0)} or kinds >= {(1, 1), (0, 1)} or kinds >= {(1, 1), (1, 0)}
    assert worst <= 1e-14


@pytest.mark.parametrize("degrees", ["101", "202", "303", "mixed"])
def test_hybrid_system_reproduces_a_linear_solution(degrees):
    """u = 0.3 + x - 2y + 0.5z with its Dirichlet data on every boundary
    face: the discrete solution is u to rounding on an affine hybrid mesh"""
    import modfem_amd as mf
    from oracle_lib import dg_degrees, dg_system_dense
    a = hybrid_arrays(7)
    nel = a["el_nodes"].shape[0]
    et = a["el_type"]
    pd = (np.random.default_rng(3).choice([101, 202, 303], size=nel) if degrees == "mixed"
          else np.full(nel, int(degrees)))
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    A, b = dg_system_dense(a, pd, mf.IND_DIFF, lambda e: cc, 1.0, 9)
    x = np.linalg.solve(A, b)
    _, ns, off = dg_degrees(pd, nel, et)
    nodes = a["coor"][a["el_nodes"]]
    o = oracle()
    rng = np.random.default_rng(1)
    worst = 0.0
    for e in range(nel):
        for _ in range(3):
            xp = np.zeros(3)
            dofs = np.ascontiguousarray(x[off[e]:off[e + 1]])
            if et[e] == 7:
                lam = rng.dirichlet(np.ones(4))
                eta = np.ascontiguousarray(lam[1:])
                uh = o.orc_dgt_eval(int(pd[e]) % 100, P(np.ascontiguousarray(nodes[e][:4].ravel())),
                                    P(dofs), P(eta), P(xp))
            else:
                u, v = rng.uniform(0, 0.5, 2)
                eta = np.array([u, v, rng.uniform(-1, 1)])
                uh = o.orc_dg_eval(int(pd[e]), P(np.ascontiguousarray(nodes[e].ravel())), P(dofs),
                                   P(eta), P(xp))
            worst = max(worst, abs(uh - (0.3 + xp[0] - 2 * xp[1] + 0.5 * xp[2])))
    assert worst <= 1e-10, worst
