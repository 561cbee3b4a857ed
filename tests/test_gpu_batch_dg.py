"""DG problems as batch members (no reference counterpart for batching: the
reference solves one system at a time).  A DG member is GMRES on the
block-Jacobi-preconditioned system Dinv A x = Dinv b held as one SELL
matrix, which is the reference's left-preconditioned block-Jacobi GMRES of
DG systems (lss_mkb_core.c:1861-2097, 2317-2668).

- An FP64 DG member gives the result of the DG problem's own block-Jacobi
  GMRES: same convergence, iteration count within rounding drift, solution
  within the solver tolerance.
- A mixed-precision DG member meets ||Dinv (b - A x)|| <= tol ||Dinv b|| in
  FP64, checked on the host with the oracle's assembled DG system and the
  inverses of its diagonal blocks.
- DG and std members, FP64 and mixed, give bitwise the same iterates alone
  and in a mix whose membership changes (fixed tiles)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402

pytestmark = pytest.mark.gpu

CC = np.r_[np.eye(3).ravel(), np.zeros(12)]


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cube_mesh(n, jitter=0.1, seed=0):
    """n^3 hexes split into prisms, bottom/top faces Neumann (flag 11), sides
    Dirichlet: the reference's DIFF_IN_CUBE DG mesh refined"""
    a = mf.Mesh.box_prism(n, n, n, seed=seed, jitter=jitter).arrays()
    bc = a["bf_bc"].copy()
    bc[a["bf_side"] < 2] = 11
    return mf.Mesh.from_arrays(a["coor"], a["el_type"], a["el_nodes"], a["bf_elem"],
                               a["bf_side"], bc)


def make_dg(n, pdeg, seed):
    mesh = cube_mesh(n, seed=seed)
    dg = mf.DgProblem(mesh, pdeg)
    dg.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=CC, src_model=mf.SRC_DIFF_IN_CUBE)
    dg.assemble()
    return mesh, dg


def make_std(size, seed):
    mesh = mf.Mesh.box_prism(*size, seed=seed, jitter=0.2)
    p = mf.Problem.from_mesh(mesh)
    kappa = np.exp(np.random.default_rng(seed).normal(0.0, 0.5, mesh.nel))
    cc = np.zeros(mf.NCOEF)
    cc[mf.C_S] = 1.0
    p.set_coefficients(mf.IND_DIFF, mf.COEF_KAPPA, coef_const=cc, coef_el=kappa)
    p.integrate_assemble(64)
    return mesh, p


CASES = [(3, 101, 1), (4, 101, 2), (3, 202, 3), (2, 202, 4)]


@pytest.mark.parametrize("use_graph", [False, True])
def test_dg_members_match_the_dg_block_jacobi_solve(use_graph):
    made = [make_dg(*c) for c in CASES]
    ref = [dg.gmres(precon=1, krylbas=30, max_iter=4000, tol=1e-8) for _, dg in made]
    b = mf.Batch(krylbas=30)
    for _, dg in made:
        b.add(dg, max_iter=4000, tol=1e-8)
    assert b.run(use_graph=use_graph) == 0
    for (x0, i0), (_, dg) in zip(ref, made):
        st = b.status(dg)
        assert st["done"] and st["converged"], st
        assert abs(st["iters"] - i0["iters"]) <= max(2, i0["iters"] // 50), (st, i0)
        assert rel_err(dg.solution(), x0) <= 1e-6
    b.free()
    for mesh, dg in made:
        dg.free()
        mesh.free()


def test_mixed_dg_member_meets_the_block_jacobi_measure():
    import scipy.sparse as sp
    from oracle_lib import dg_system_dense
    tol = 1e-8
    made = [make_dg(*c) for c in CASES[:3]]
    b = mf.Batch(krylbas=30)
    for _, dg in made:
        b.add(dg, max_iter=6000, tol=tol, inner_precision=32)
    while b.run(max_cycles=1, use_graph=False) > 0:
        pass
    for (mesh, dg), (n, pdeg, seed) in zip(made, CASES[:3]):
        st = b.status(dg)
        assert st["done"] and st["converged"] and st["refinements"] >= 1, st
        A, rhs = dg_system_dense(mesh.arrays(), pdeg, mf.IND_DIFF, lambda e: CC, 1.0,
                                 mf.SRC_DIFF_IN_CUBE, sparse=True)
        A = sp.csr_matrix(A)
        ns = dg.ns
        blocks = [np.linalg.inv(A[e * ns:(e + 1) * ns, e * ns:(e + 1) * ns].toarray())
                  for e in range(dg.nel)]
        Dinv = sp.block_diag(blocks, format="csr")
        x = dg.solution()
        rel = np.linalg.norm(Dinv @ (rhs - A @ x)) / np.linalg.norm(Dinv @ rhs)
        print(f"DG {pdeg} n={n}: {st['iters']} inner its, {st['refinements']} refinements, "
              f"measure {rel:.2e} (reported {st['res_rel']:.2e})")
        assert rel <= tol * 1.01, (rel, st)
    b.free()
    for mesh, dg in made:
        dg.free()
        mesh.free()


def run_alone(obj, kind, use_graph):
    b = mf.Batch(krylbas=20)
    b.set_tiles(4)
    b.add(obj, precon=1, max_iter=4000, tol=1e-9, inner_precision=kind)
    while b.run(max_cycles=1, use_graph=use_graph) > 0:
        pass
    st = b.status(obj)
    x = obj.solution()
    b.remove(obj)
    b.free()
    return st, x


@pytest.mark.parametrize("use_graph", [False, True])
def test_dg_and_std_members_same_alone_and_in_the_mix(use_graph):
    members = [make_dg(3, 101, 5) + (64,), make_std((6, 5, 4), 6) + (32,),
               make_dg(2, 202, 7) + (32,), make_std((5, 5, 5), 8) + (64,),
               make_dg(3, 202, 9) + (64,)]
    ref = [run_alone(obj, kind, use_graph) for _, obj, kind in members]
    b = mf.Batch(krylbas=20)
    b.set_tiles(4)
    order = [2, 0, 4, 1, 3]
    joined, finished = [], set()
    while len(finished) < len(members):
        if len(joined) < len(order):
            k = order[len(joined)]
            b.add(members[k][1], precon=1, max_iter=4000, tol=1e-9, inner_precision=members[k][2])
            joined.append(k)
        b.run(max_cycles=1, use_graph=use_graph)
        for k in joined:
            if k in finished:
                continue
            obj = members[k][1]
            st = b.status(obj)
            if st["done"]:
                finished.add(k)
                s0, x0 = ref[k]
                assert st["converged"] and s0["converged"], (k, st, s0)
                assert st["iters"] == s0["iters"] and st["refinements"] == s0["refinements"]
                assert np.array_equal(obj.solution(), x0), k
                b.remove(obj)
    b.free()
    for mesh, obj, _ in members:
        obj.free()
        mesh.free()


def test_dg_member_refuses_mixed_degrees_and_free():
    mesh = cube_mesh(2)
    dg = mf.DgProblem(mesh, np.array([101, 202] * (mesh.nel // 2)))
    dg.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=CC, src_model=mf.SRC_DIFF_IN_CUBE)
    dg.assemble()
    b = mf.Batch(krylbas=10)
    with pytest.raises(RuntimeError):
        b.add(dg)
    dg.free()
    mesh.free()
    mesh, dg = make_dg(2, 101, 3)
    b.add(dg)
    with pytest.raises(RuntimeError):
        mf.check(mf.lib().mf_dg_free(dg.h))
    b.run()
    b.remove(dg)
    b.free()
    dg.free()
    mesh.free()
