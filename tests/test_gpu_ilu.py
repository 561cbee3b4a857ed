"""ILU(0) on the device (precon 2): factors in the reference's DOF order
against the oracle's restatement of the reference's block ILU(0) for 1x1
blocks (lad_block/las_block_intf.c:1409-2143), the defining property
(L U)_ij = a_ij on the pattern for the multicolour order, and GMRES with
the ILU preconditioner (compreres' BLOCK_ILU branch, lss_mkb_core.c:1947)
against the oracle's."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import OracleProblem, ilu_product, oracle_ilu0  # noqa: E402

pytestmark = pytest.mark.gpu


def conv_diff_problem(nx, ny, nz, seed):
    mesh = mf.Mesh.box_prism(nx, ny, nz, seed=seed, jitter=0.2)
    rng = np.random.default_rng(seed + 1)
    kappa = np.exp(rng.normal(0, 1, mesh.nel))
    b = rng.uniform(-1, 1, (mesh.nel, 3))
    cc = np.zeros(mf.NCOEF)
    cc[mf.C_S] = 1.0
    p = mf.Problem.from_mesh(mesh)
    p.set_coefficients(mf.IND_CONV_DIFF, mf.COEF_KAPPA_B, coef_const=cc,
                       coef_el=np.concatenate([kappa, b[:, 0], b[:, 1], b[:, 2]]))
    p.integrate_assemble(64)
    tab = np.zeros((mesh.nel, 21))
    tab[:, 0] = tab[:, 4] = tab[:, 8] = kappa
    tab[:, 9:12] = b
    tab[:, 19] = 1.0
    return mesh, p, tab


def test_reference_order_factors_match_oracle():
    mesh, p, tab = conv_diff_problem(7, 6, 5, 3)
    val, rhs = p.system()
    op = OracleProblem(mesh.arrays())
    assert np.array_equal(op.row_ptr, p.pattern()[0])
    lu0 = oracle_ilu0(op.row_ptr, op.col_ind, val)
    lu, pos, levels = p.ilu()
    assert np.array_equal(pos, np.arange(p.n))
    err = np.abs(lu - lu0).max() / np.abs(lu0).max()
    assert err <= 1e-12, err
    assert levels[0] > 10 and levels[1] > 10
    p.free()


def test_multicolour_factors_reproduce_matrix_on_pattern():
    mesh, p, _ = conv_diff_problem(6, 5, 4, 5)
    p.set_ilu_order(1)
    val, _ = p.system()
    rp, ci = p.pattern()
    lu, pos, (nl, nu, ncol) = p.ilu()
    assert sorted(pos) == list(range(p.n))
    assert nl <= ncol and nu <= ncol and ncol <= 40
    LU = ilu_product(rp, ci, lu, pos)
    rows = np.repeat(np.arange(p.n), np.diff(rp))
    err = np.abs(LU[pos[rows], pos[ci]] - val).max()
    assert err <= 1e-12 * np.abs(val).max(), err
    p.free()


def test_gmres_ilu_matches_oracle_and_multicolour_converges():
    mesh, p, tab = conv_diff_problem(9, 8, 7, 11)
    op = OracleProblem(mesh.arrays())
    v0, r0 = op.assemble(mf.IND_CONV_DIFF, tab)
    x0, info0 = op.gmres(v0, r0, precon=2, krylbas=30, max_iter=2000, tol=1e-8)
    x, info = p.gmres(precon=2, krylbas=30, max_iter=2000, tol=1e-8)
    assert info0["converged"] and info["converged"]
    assert abs(info["iters"] - info0["iters"]) <= max(2, info0["iters"] // 20), (info, info0)
    assert np.abs(x - x0).max() <= 1e-6 * np.abs(x0).max()
    _, infoj = p.gmres(precon=1, krylbas=30, max_iter=4000, tol=1e-8)
    assert info["iters"] < infoj["iters"]  # ILU(0) beats Jacobi on this system
    p.set_ilu_order(1)
    xm, infom = p.gmres(precon=2, krylbas=30, max_iter=4000, tol=1e-8)
    assert infom["converged"]
    assert np.abs(xm - x0).max() <= 1e-5 * np.abs(x0).max()
    # re-assembly invalidates the factors; the next solve refactors
    p.integrate_assemble(64)
    xr, infor = p.gmres(precon=2, krylbas=30, max_iter=4000, tol=1e-8)
    assert infor["iters"] == infom["iters"] and np.array_equal(xr, xm)
    p.free()


def test_control_file_block_ilu():
    G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mkb")
    prm = mf.mkb_read(os.path.join(G, "laplace_std_in_cube.dat"))
    prm = dict(prm, precon=5, ilu_k=0)  # BLOCK_ILU, ILU(0) as preconditioner
    mesh, p, tab = conv_diff_problem(6, 6, 6, 13)
    x, info = p.solve_mkb(prm)
    x2, info2 = p.gmres(precon=2, krylbas=prm["krylbas"], max_iter=prm["max_iter"], tol=prm["conv_meas"])
    assert info["converged"] and info["iters"] == info2["iters"] and np.array_equal(x, x2)
    p.free()


@pytest.mark.parametrize("sweeps", [1, 2])
def test_gmres_gauss_seidel_matches_oracle_in_the_same_order(sweeps):
    """precon 3: GS smoothing (BLOCK_GS) in the device's multicolour order;
    the oracle runs sequential GS in that same order"""
    from oracle_lib import oracle
    mesh, p, tab = conv_diff_problem(8, 7, 6, 17)
    p.set_gs(sweeps)
    pos, ncol = p.gs_order()
    assert sorted(pos) == list(range(p.n)) and 2 <= ncol <= 40
    x, info = p.gmres(precon=3, krylbas=30, max_iter=3000, tol=1e-8)
    op = OracleProblem(mesh.arrays())
    v0, r0 = op.assemble(mf.IND_CONV_DIFF, tab)
    order = np.empty(p.n, np.int32)
    order[pos] = np.arange(p.n, dtype=np.int32)
    oracle().orc_set_gs(order.ctypes.data_as(__import__("ctypes").c_void_p), sweeps)
    try:
        x0, info0 = op.gmres(v0, r0, precon=3, krylbas=30, max_iter=3000, tol=1e-8)
    finally:
        oracle().orc_set_gs(None, 1)
    assert info0["converged"] and info["converged"]
    assert abs(info["iters"] - info0["iters"]) <= max(2, info0["iters"] // 20), (info, info0)
    assert np.abs(x - x0).max() <= 1e-6 * np.abs(x0).max()
    p.free()


def test_control_file_block_gs():
    """the reference's own GS control file (ctest Laplace_std/mkb.dat: block
    GS, 2 sweeps) drives the solve"""
    G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mkb")
    prm = mf.mkb_read(os.path.join(G, "laplace_std.dat"))
    assert (prm["precon"], prm["nr_prec"]) == (2, 2)
    mesh, p, _ = conv_diff_problem(6, 6, 6, 19)
    x, info = p.solve_mkb(prm)
    p.set_gs(2)
    x2, info2 = p.gmres(precon=3, krylbas=prm["krylbas"], max_iter=prm["max_iter"], tol=prm["conv_meas"])
    assert info["converged"] and info["iters"] == info2["iters"] and np.array_equal(x, x2)
    p.free()
