"""p-multigrid V-cycle as the GMRES preconditioner of the DG block system
(the V-cycle of lsr_mkb_core_vcycle, lsd_mkb_core/lss_mkb_core.c:1288-1470,
over levels of decreasing isotropic degree):
- the coarse operators are the Galerkin products P^T A P with P the injection
  of the nested tensor bases: every coarse block equals the finer level's
  block restricted to the coarse functions, bit for bit;
- GMRES(30) with one V-cycle per application reaches the oracle system's
  direct solution, in fewer iterations than with block Gauss-Seidel or
  block Jacobi."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from test_gpu_dg_tet import with_neumann_z  # noqa: E402
from test_oracle_dg import exponents  # noqa: E402

pytestmark = pytest.mark.gpu


def problem(pdeg, n=3, seed=5):
    mf.init()
    mesh = with_neumann_z(mf.Mesh.box_prism(n, n, n, seed=seed, jitter=0.15))
    cc = np.r_[np.eye(3).ravel(), np.zeros(12)]
    dg = mf.DgProblem(mesh, pdeg)
    dg.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc, src_model=mf.SRC_DIFF_IN_CUBE)
    dg.assemble()
    return mesh, dg, cc


def test_coarse_blocks_are_galerkin_sub_blocks():
    print("multigrid Galerkin blocks", flush=True)
    mesh, dg, _ = problem(303)
    assert dg.set_multigrid(2, 2, 2) == 3
    rows = {2: [dg.block_row(e)[0] for e in range(dg.nel)]}
    for lev in (1, 0):
        rows[lev] = [dg.multigrid_level_row(lev, e)[0] for e in range(dg.nel)]
    nb = dg.faces()[0]
    for lev in (1, 0):
        q = lev + 1
        ef = exponents(101 * (q + 1))
        m = np.array([ef.index(t) for t in exponents(101 * q)])
        for e in range(dg.nel):
            for s in range(6):
                if s > 0 and nb[e, s - 1] < 0:
                    continue
                want = rows[lev + 1][e][s][np.ix_(m, m)]
                assert np.array_equal(rows[lev][e][s], want), (lev, e, s)
    dg.free()
    mesh.free()


@pytest.mark.parametrize("pdeg", [202, 303])
def test_multigrid_gmres_reaches_the_direct_solution(pdeg):
    import scipy.sparse.linalg as spl
    from oracle_lib import dg_system_dense
    print("multigrid GMRES", pdeg, flush=True)
    mesh, dg, cc = problem(pdeg, n=4, seed=6)
    a = mesh.arrays()
    its = {}
    for precon in (1, 2, 4):
        if precon == 4:
            dg.set_multigrid(2, 2, 2)
        x, info = dg.gmres(precon=precon, krylbas=30, max_iter=4000, tol=1e-10)
        assert info["converged"], (precon, info)
        its[precon] = info["iters"]
        if precon == 4:
            xm = x
    A, b = dg_system_dense(a, pdeg, mf.IND_DIFF, lambda e: cc, 1.0, mf.SRC_DIFF_IN_CUBE, sparse=True)
    xs = spl.spsolve(A.tocsc(), b)
    print("  iterations by precon", its, flush=True)
    assert np.abs(xm - xs).max() <= 1e-7 * np.abs(xs).max()
    assert its[4] < its[2] and its[4] < its[1]
    dg.free()
    mesh.free()


def test_multigrid_rejects_what_it_does_not_cover():
    mf.init()
    mesh = mf.Mesh.box_prism(2, 2, 2, seed=1)
    for pd in (101, 102, 1003):
        dg = mf.DgProblem(mesh, pd)
        with pytest.raises(RuntimeError):
            dg.set_multigrid()
        dg.free()
    dg = mf.DgProblem(mesh, 202)
    with pytest.raises(RuntimeError):
        dg.gmres(precon=4)
    dg.free()
    mesh.free()
