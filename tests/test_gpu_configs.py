"""The workload configurations at (or near) their stated sizes, on the device
through the C ABI, against the oracle:

  unit-cube Poisson      20^3 hexes x 2 prisms (16,000 prisms, 9,261 DOFs,
                         nnz 178,181, 3,200 boundary faces), -Lap u = 1,
                         penalty u = 0, GMRES(30) + Jacobi to 1e-8: sizes
                         exact, assembled system <= 1e-12, iteration count
                         within 5% of the oracle's, solution <= 1e-6
  128^3 convection-      the same generator at 32^3 (65,536 prisms):
  diffusion form         lognormal kappa (sigma 1), b ~ U(-1,1)^3 per
                         element, 0.2h jitter; FP32 integration within 2e-6
                         of FP64; mixed-precision GMRES(50) + ILU(0) meets
                         the FP64 1e-8 preconditioned test (checked on the
                         host with the oracle's system and ILU) and agrees
                         with the oracle's FP64 GMRES + ILU(0) solution
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import OracleProblem, P, oracle, oracle_ilu0  # noqa: E402

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_unit_cube_poisson_20cubed():
    mesh = mf.Mesh.box_prism(20, 20, 20, seed=2017, jitter=0.0, bc=1)
    op = OracleProblem(mesh.arrays())
    assert (mesh.nel, op.n, op.nnz, mesh.nbf) == (16000, 9261, 178181, 3200)
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    cc[19] = 1.0
    p = mf.Problem.from_mesh(mesh)
    p.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc)
    p.integrate_assemble(64)
    assert np.array_equal(p.dofs(), op.dof_node)
    val, rhs = p.system()
    v0, r0 = op.assemble(mf.IND_DIFF, np.tile(cc, (op.nel, 1)))
    assert rel_err(val, v0) <= 1e-12 and rel_err(rhs, r0) <= 1e-12
    x, info = p.gmres(precon=1, krylbas=30, max_iter=500, tol=1e-8)
    x0, info0 = op.gmres(v0, r0, precon=1, krylbas=30, max_iter=500, tol=1e-8)
    print(f"unit cube: iterations gpu {info['iters']} oracle {info0['iters']}")
    assert info["converged"] and info0["converged"]
    assert abs(info["iters"] - info0["iters"]) <= max(3, info0["iters"] // 20)
    assert rel_err(x, x0) <= 1e-6
    # the solution is the discrete Poisson solution: symmetric in x, y, z
    # up to the diagonal split, positive inside
    assert x.min() > -1e-6 and 0.05 < x.max() < 0.06
    p.free()
    mesh.free()


def convdiff_32cubed(seed=2017):
    mesh = mf.Mesh.box_prism(32, 32, 32, seed=seed, jitter=0.2, bc=1)
    rng = np.random.default_rng(seed)
    kappa = np.exp(rng.normal(0.0, 1.0, mesh.nel))
    b = rng.uniform(-1, 1, (mesh.nel, 3))
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    cc[19] = 1.0
    tab = np.zeros((mesh.nel, 21))
    tab[:, 0] = tab[:, 4] = tab[:, 8] = kappa
    tab[:, 9:12] = b
    tab[:, 19] = 1.0
    return mesh, cc, np.concatenate([kappa, b[:, 0], b[:, 1], b[:, 2]]), tab


def test_convection_diffusion_32cubed_fp32_integration_and_mixed_gmres_ilu0():
    import scipy.sparse as sp
    mesh, cc, ce, tab = convdiff_32cubed()
    p = mf.Problem.from_mesh(mesh)
    p.set_coefficients(mf.IND_CONV_DIFF, mf.COEF_KAPPA_B, coef_const=cc, coef_el=ce)
    p.integrate_assemble(32)
    v32, r32 = p.system()
    p.integrate_assemble(64)
    val, rhs = p.system()
    assert rel_err(v32, val) <= 2e-6 and rel_err(r32, rhs) <= 2e-6
    op = OracleProblem(mesh.arrays())
    v0, r0 = op.assemble(mf.IND_CONV_DIFF, tab)
    assert rel_err(val, v0) <= 1e-12 and rel_err(rhs, r0) <= 1e-12
    tol = 1e-8
    x, info = p.gmres_mixed(precon=2, krylbas=50, max_iter=5000, tol=tol)
    assert info["converged"], info
    A = sp.csr_matrix((v0, op.col_ind, op.row_ptr), shape=(op.n, op.n))
    lu = oracle_ilu0(op.row_ptr, op.col_ind, v0)

    def prec(v):
        out = np.zeros(op.n)
        oracle().orc_ilu_solve(op.n, P(op.row_ptr), P(op.col_ind), P(lu),
                               P(np.ascontiguousarray(v)), P(out))
        return out
    rel = np.linalg.norm(prec(r0 - A @ x)) / np.linalg.norm(prec(r0))
    x0, info0 = op.gmres(v0, r0, precon=2, krylbas=50, max_iter=5000, tol=tol)
    print(f"32^3 conv-diff: mixed {info['iters']} its, {info['refinements']} refinements, "
          f"FP64 ILU-preconditioned residual {rel:.2e}; oracle FP64 {info0['iters']} its")
    assert rel <= tol * 1.0001
    assert info0["converged"] and rel_err(x, x0) <= 1e-5
    p.free()
    mesh.free()
