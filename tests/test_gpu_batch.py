"""Batched GMRES over resident problems (no reference counterpart: the
reference solves one system at a time, lss_mkb_core.c:2317).  Each member of
a batch must give the result it gives alone: same convergence, iteration
count within rounding drift of the reduction order, and solution within the
solver tolerance, whether members join at once or while others are running
(continuous batching), eagerly launched or replayed as one hipGraph."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make_problems(sizes, seed=0):
    probs = []
    rng = np.random.default_rng(seed)
    for k, (nx, ny, nz) in enumerate(sizes):
        mesh = mf.Mesh.box_prism(nx, ny, nz, seed=100 + k, jitter=0.2,
                                 lx=1.0, ly=rng.uniform(1, 3), lz=rng.uniform(1, 2))
        p = mf.Problem.from_mesh(mesh)
        kappa = np.exp(rng.normal(0.0, 0.5, mesh.nel))
        cc = np.zeros(mf.NCOEF)
        cc[mf.C_S] = 1.0
        p.set_coefficients(mf.IND_DIFF, mf.COEF_KAPPA, coef_const=cc, coef_el=kappa)
        p.integrate_assemble(64)
        probs.append((mesh, p))
    return probs


SIZES = [(6, 5, 4), (10, 9, 8), (3, 3, 3), (14, 12, 10), (8, 8, 8)]


def alone(p, m=30, tol=1e-8):
    x, info = p.gmres(precon=1, krylbas=m, max_iter=4000, tol=tol)
    return x, info


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_single_solves(use_graph):
    probs = make_problems(SIZES)
    ref = [alone(p) for _, p in probs]
    b = mf.Batch(krylbas=30)
    for _, p in probs:
        b.add(p, precon=1, max_iter=4000, tol=1e-8)
    assert b.run(use_graph=use_graph) == 0
    for (x0, i0), (_, p) in zip(ref, probs):
        st = b.status(p)
        assert st["done"] and st["converged"], st
        assert abs(st["iters"] - i0["iters"]) <= max(2, i0["iters"] // 50), (st, i0)
        assert st["res_rel"] < 1e-8
        assert rel_err(p.solution(), x0) <= 1e-6
    b.free()
    for _, p in probs:
        p.free()


def test_graph_replays_match_eager_launches():
    """the same cycle replayed from a graph and launched eagerly gives
    bitwise the same iterates (same kernels, same order)"""
    probs = make_problems(SIZES[:3], seed=3)
    out = []
    for use_graph in (False, True):
        b = mf.Batch(krylbas=20)
        b.set_tiles(8)  # same blocks in both modes (tuned tiles may differ by mode)
        for _, p in probs:
            b.add(p, precon=1, max_iter=2000, tol=1e-9)
        while b.run(max_cycles=1, use_graph=use_graph) > 0:
            pass
        out.append([(b.status(p)["iters"], p.solution()) for _, p in probs])
        b.free()
    for (i0, x0), (i1, x1) in zip(out[0], out[1]):
        assert i0 == i1 and np.array_equal(x0, x1)
    for _, p in probs:
        p.free()


def test_continuous_batching_same_alone_and_in_the_mix():
    probs = make_problems(SIZES, seed=7)
    ref = [alone(p) for _, p in probs]
    b = mf.Batch(krylbas=30)
    for _, p in probs[:2]:
        b.add(p, precon=1, max_iter=4000, tol=1e-8)
    b.run(max_cycles=1)
    b.add(probs[2][1], precon=1, max_iter=4000, tol=1e-8)  # joins while others run
    b.run(max_cycles=2)
    for _, p in probs[3:]:
        b.add(p, precon=1, max_iter=4000, tol=1e-8)
    finished = set()
    while True:
        active = b.run(max_cycles=1)
        for k, (_, p) in enumerate(probs[:3] + probs[3:]):
            if k not in finished:
                try:
                    st = b.status(p)
                except RuntimeError:
                    continue
                if st["done"]:
                    finished.add(k)
                    x = p.solution()
                    assert st["converged"]
                    assert abs(st["iters"] - ref[k][1]["iters"]) <= max(2, ref[k][1]["iters"] // 50)
                    assert rel_err(x, ref[k][0]) <= 1e-6
                    b.remove(p)  # slot freed while others still run
        if active == 0:
            break
    assert len(finished) == len(probs)
    with pytest.raises(RuntimeError):
        b.status(probs[0][1])
    b.free()
    for _, p in probs:
        p.free()


def test_member_cannot_be_freed():
    (_, p), = make_problems([(4, 4, 4)])
    b = mf.Batch(krylbas=10)
    b.add(p)
    with pytest.raises(RuntimeError):
        mf.check(mf.lib().mf_problem_free(p.h))
    b.run()
    b.remove(p)
    b.free()
    p.free()


@pytest.mark.parametrize("spb", [1, 3, 8, 32, 64])
def test_every_fixed_tile_matches_single_solves(spb):
    probs = make_problems(SIZES[:4], seed=11)
    ref = [alone(p) for _, p in probs]
    b = mf.Batch(krylbas=30)
    b.set_tiles(spb)
    for _, p in probs:
        b.add(p, precon=1, max_iter=4000, tol=1e-8)
    assert b.run(use_graph=True) == 0
    t = b.tiles()
    assert t["spb"] == spb and not t["tuned"]
    for (x0, i0), (_, p) in zip(ref, probs):
        st = b.status(p)
        assert st["done"] and st["converged"], st
        assert abs(st["iters"] - i0["iters"]) <= max(2, i0["iters"] // 50), (st, i0)
        assert rel_err(p.solution(), x0) <= 1e-6
    b.free()
    for _, p in probs:
        p.free()


def test_tile_bounds_are_checked():
    b = mf.Batch(krylbas=10)
    for bad in (-1, 65):
        with pytest.raises(RuntimeError):
            b.set_tiles(bad)
    b.free()


@pytest.mark.parametrize("use_graph", [False, True])
def test_tuned_tile_changes_no_iterate(use_graph):
    """trial cycles of the tuner are undone: a tuned batch gives bitwise the
    iterates of a batch run with the chosen tile fixed from the start; a
    second batch of the same shape takes the cached choice without trials"""
    probs = make_problems(SIZES, seed=5)
    mf.Batch.clear_tile_cache()
    b = mf.Batch(krylbas=20)
    for _, p in probs:
        b.add(p, precon=1, max_iter=2000, tol=1e-9)
    b.run(max_cycles=1, use_graph=use_graph)
    t = b.tiles()
    assert t["tuned"]
    trials = t["trials"]
    assert all(ms > 0 for ms in trials.values()), trials
    assert t["spb"] == min(trials, key=trials.get)
    for _, p in probs:
        assert b.status(p)["iters"] <= 20  # one real cycle, none from the trials
    b.run(use_graph=use_graph)
    tuned = [(b.status(p)["iters"], p.solution()) for _, p in probs]
    b.free()
    f = mf.Batch(krylbas=20)
    f.set_tiles(t["spb"])
    for _, p in probs:
        f.add(p, precon=1, max_iter=2000, tol=1e-9)
    f.run(use_graph=use_graph)
    for (i0, x0), (_, p) in zip(tuned, probs):
        assert f.status(p)["iters"] == i0 and np.array_equal(p.solution(), x0)
    f.free()
    c = mf.Batch(krylbas=20)
    for _, p in probs:
        c.add(p, precon=1, max_iter=2000, tol=1e-9)
    c.run(max_cycles=1, use_graph=use_graph)
    assert c.tiles() == t  # cached: same tile and the same recorded trials
    c.free()
    for _, p in probs:
        p.free()


def member_case(k, rng):
    """problem k of the mixed-assembly batch: prisms and tetrahedra, three
    term sets, row-wise, two-pass and atomic assembly modes"""
    kind = k % 6
    shape = (int(rng.integers(3, 9)), int(rng.integers(3, 9)), int(rng.integers(2, 7)))
    mesh = (mf.Mesh.box_tetra(*shape, seed=300 + k, jitter=0.2) if kind in (1, 4)
            else mf.Mesh.box_prism(*shape, seed=300 + k, jitter=0.2))
    p = mf.Problem.from_mesh(mesh)
    nel = mesh.nel
    kappa = np.exp(rng.normal(0.0, 0.5, nel))
    cc = np.zeros(mf.NCOEF)
    cc[mf.C_S] = 1.0
    if kind in (0, 1):
        p.set_coefficients(mf.IND_DIFF, mf.COEF_KAPPA, coef_const=cc, coef_el=kappa)
    elif kind in (2, 4):
        b = rng.uniform(-1, 1, (nel, 3))
        p.set_coefficients(mf.IND_CONV_DIFF, mf.COEF_KAPPA_B, coef_const=cc,
                           coef_el=np.concatenate([kappa, b[:, 0], b[:, 1], b[:, 2]]))
    else:
        full = rng.uniform(-0.3, 0.3, (21, nel))
        full[0] += 2.0
        full[4] += 2.0
        full[8] += 2.0
        p.set_coefficients(mf.IND_ALL, mf.COEF_FULL, coef_el=full.ravel())
    p.set_assembly_mode([0, 2, 2, 0, 2, 1][kind] if k < 6 else [2, 0, 0, 2, 0, 1][kind])
    return mesh, p


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_assembly_matches_each_problems_own(precision):
    """one launch per kernel for every row-wise and two-pass member (block
    table over the members), the atomic member on its own: every member's
    assembled system is bitwise the one its own assembly gives"""
    rng = np.random.default_rng(11)
    made = [member_case(k, rng) for k in range(12)]
    ref = []
    for _, p in made:
        p.integrate_assemble(precision)
        ref.append(p.system())
    b = mf.Batch(krylbas=10)
    for _, p in made:
        p.set_system(np.zeros(p.nnz), np.zeros(p.n))
        b.add(p, precon=1, max_iter=100, tol=1e-8)
    b.integrate_assemble(precision)
    for k, ((_, p), (v0, r0)) in enumerate(zip(made, ref)):
        v, r = p.system()
        if k % 6 == 5:  # element-parallel atomics: reproducible to rounding only
            assert np.abs(v - v0).max() <= 1e-12 * np.abs(v0).max(), k
            assert np.abs(r - r0).max() <= 1e-12 * np.abs(r0).max(), k
        else:
            assert np.array_equal(v, v0) and np.array_equal(r, r0), k
    b.free()
    for mesh, p in made:
        p.free()
        mesh.free()
