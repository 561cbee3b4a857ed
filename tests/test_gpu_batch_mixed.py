"""Batched GMRES with multicolour ILU(0) members and mixed-precision members
(no reference counterpart: the reference solves one system at a time,
lss_mkb_core.c:2317; its ILU(0) is lss_mkb_core.c:1947-1997 with 1x1 blocks).

- An FP64 ILU(0) member gives the result of the single ILU(0) solve in the
  same (multicolour) elimination order: same convergence, iteration count
  within rounding drift of the reduction order, solution within the solver
  tolerance.
- A mixed-precision member (FP32 inner solves, FP64 refinement) meets the
  FP64 measure ||D^-1 (b - A x)|| <= tol ||D^-1 b||, checked on the host with
  the matrix read back from the device.
- Every member, of any kind, gives bitwise the same iterates whether it runs
  alone or among other members joining and leaving (fixed tiles: a member's
  blocks then depend only on its own size)."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make_problems(sizes, seed=0, ilu_order=1):
    probs = []
    rng = np.random.default_rng(seed)
    for k, (nx, ny, nz) in enumerate(sizes):
        mesh = mf.Mesh.box_prism(nx, ny, nz, seed=200 + k, jitter=0.2,
                                 lx=1.0, ly=rng.uniform(1, 3), lz=rng.uniform(1, 2))
        p = mf.Problem.from_mesh(mesh)
        kappa = np.exp(rng.normal(0.0, 0.5, mesh.nel))
        cc = np.zeros(mf.NCOEF)
        cc[mf.C_S] = 1.0
        p.set_coefficients(mf.IND_DIFF, mf.COEF_KAPPA, coef_const=cc, coef_el=kappa)
        p.integrate_assemble(64)
        p.set_ilu_order(ilu_order)
        probs.append((mesh, p))
    return probs


def free_all(probs):
    for mesh, p in probs:
        p.free()
        mesh.free()


def fp64_measure(p, x, precon):
    """||D^-1 (b - A x)|| / ||D^-1 b|| in FP64 on the host (D = I for precon 0)"""
    val, rhs = p.system()
    rp, ci = p.pattern()
    A = sp.csr_matrix((val, ci, rp), shape=(p.n, p.n))
    d = 1.0 / A.diagonal() if precon > 0 else np.ones(p.n)
    return float(np.linalg.norm(d * (rhs - A @ x)) / np.linalg.norm(d * rhs))


SIZES = [(6, 5, 4), (10, 9, 8), (3, 3, 3), (14, 12, 10), (8, 8, 8)]


@pytest.mark.parametrize("use_graph", [False, True])
def test_ilu_members_match_single_ilu_solves(use_graph):
    probs = make_problems(SIZES, seed=1)
    ref = [p.gmres(precon=2, krylbas=30, max_iter=4000, tol=1e-8) for _, p in probs]
    b = mf.Batch(krylbas=30)
    for _, p in probs:
        b.add(p, precon=2, max_iter=4000, tol=1e-8)
    assert b.run(use_graph=use_graph) == 0
    for (x0, i0), (_, p) in zip(ref, probs):
        st = b.status(p)
        assert st["done"] and st["converged"], st
        assert abs(st["iters"] - i0["iters"]) <= max(2, i0["iters"] // 50), (st, i0)
        assert st["res_rel"] < 1e-8
        assert rel_err(p.solution(), x0) <= 1e-6
    b.free()
    free_all(probs)


def test_ilu_member_needs_multicolour_order():
    (mesh, p), = make_problems([(4, 4, 4)], ilu_order=0)
    b = mf.Batch(krylbas=10)
    with pytest.raises(RuntimeError):
        b.add(p, precon=2)
    b.free()
    free_all([(mesh, p)])


@pytest.mark.parametrize("precon", [1, 2])
def test_mixed_members_meet_the_fp64_measure(precon):
    probs = make_problems(SIZES[:4], seed=2)
    tol = 1e-8
    b = mf.Batch(krylbas=30)
    for _, p in probs:
        b.add(p, precon=precon, max_iter=6000, tol=tol, inner_precision=32)
    while b.run(max_cycles=1, use_graph=False) > 0:
        pass
    for _, p in probs:
        st = b.status(p)
        assert st["done"] and st["converged"] and st["refinements"] >= 1, st
        x = p.solution()
        rel = fp64_measure(p, x, precon)
        print(f"n={p.n} precon={precon}: {st['iters']} inner its, {st['refinements']} "
              f"refinements, measure {rel:.2e} (reported {st['res_rel']:.2e})")
        assert rel <= tol * 1.0001, (rel, st)
        assert abs(rel - st["res_rel"]) <= 1e-3 * tol
        if precon == 1:
            xp, ip = p.gmres(precon=1, krylbas=30, max_iter=6000, tol=tol)
            assert ip["converged"] and rel_err(x, xp) <= 1e-5
    b.free()
    free_all(probs)


def test_mixed_single_solve_uses_fp32_basis_and_meets_the_measure():
    """mf_gmres_mixed with Jacobi: FP32 inner basis, FP64 measure"""
    (mesh, p), = make_problems([(12, 11, 10)], seed=4)
    x, info = p.gmres_mixed(precon=1, krylbas=30, max_iter=5000, tol=1e-8)
    assert info["converged"] and info["refinements"] >= 1, info
    assert fp64_measure(p, x, 1) <= 1e-8 * 1.0001
    free_all([(mesh, p)])


KINDS = [(1, 64), (2, 64), (1, 32), (2, 32), (0, 32)]


def run_alone(p, precon, prec, use_graph):
    b = mf.Batch(krylbas=20)
    b.set_tiles(4)
    b.add(p, precon=precon, max_iter=4000 if precon else 300, tol=1e-9, inner_precision=prec)
    while b.run(max_cycles=1, use_graph=use_graph) > 0:
        pass
    st = b.status(p)
    x = p.solution()
    b.remove(p)
    b.free()
    return st, x


@pytest.mark.parametrize("use_graph", [False, True])
def test_every_kind_same_alone_and_in_the_mix(use_graph):
    probs = make_problems([(7, 6, 5), (9, 8, 6), (5, 5, 5), (11, 9, 7), (6, 6, 9)], seed=5)
    kinds = {k: KINDS[k % len(KINDS)] for k in range(len(probs))}
    ref = {k: run_alone(probs[k][1], *kinds[k], use_graph) for k in range(len(probs))}
    b = mf.Batch(krylbas=20)
    b.set_tiles(4)
    order = [3, 0, 4, 1, 2]  # joining order: one new member per cycle
    joined, finished = [], set()
    while len(finished) < len(probs):
        if len(joined) < len(order):
            k = order[len(joined)]
            b.add(probs[k][1], precon=kinds[k][0], max_iter=4000 if kinds[k][0] else 300,
                  tol=1e-9, inner_precision=kinds[k][1])
            joined.append(k)
        b.run(max_cycles=1, use_graph=use_graph)
        for k in joined:
            if k in finished:
                continue
            st = b.status(probs[k][1])
            if st["done"]:
                finished.add(k)
                s0, x0 = ref[k]
                # unpreconditioned members may end on their iteration limit
                assert st["converged"] == s0["converged"], (kinds[k], st, s0)
                assert st["converged"] or kinds[k][0] == 0, (kinds[k], st)
                assert st["iters"] == s0["iters"] and st["refinements"] == s0["refinements"]
                assert np.array_equal(probs[k][1].solution(), x0), kinds[k]
                b.remove(probs[k][1])
    b.free()
    free_all(probs)
