"""DG prisms with R's other bases on the device, against the oracle (whose
shape functions and rules for these bases are bit-exact with R's compiled
apr_shape_fun_3D / apr_set_quadr_3D / apr_set_quadr_2D, and whose systems are
pinned against the tensor systems they restrict, tests/test_oracle_dg.py):
- the complete base APC_BASE_COMPLETE_DG of degree 1..3 (degree code
  MF_DG_COMPLETE + p), uniform and mixed per element;
- the tensor degrees with pz = 0 (pdeg 1, 2), uniform and mixed, whose
  penalty degree min(px, pz) is zero in R (pds_conv_diff_weakform_dg.c:
  516-527).
Assembled block rows and load vector <= 1e-12 of the oracle's assembly
(anisotropic diffusion, convection with upwinding, Dirichlet and Neumann
data, penalty 1.5) from the vector-ALU and the MFMA kernels; block-Jacobi and
block-ILU(0) GMRES reach the oracle system's direct solution."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from test_gpu_dg import mixed_block_of, random_full  # noqa: E402
from test_gpu_dg_tet import IND_ALL_DG, with_neumann_z  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = {"flat1": [1], "flat2": [2], "flat_mixed": [1, 2], "complete1": [1001],
         "complete2": [1002], "complete3": [1003], "complete_mixed": [1001, 1002, 1003]}


@pytest.mark.parametrize("valu", [6, 0])
@pytest.mark.parametrize("case", list(CASES))
def test_block_system_matches_oracle(case, valu):
    from oracle_lib import dg_system_dense
    print("DG bases", case, "valu", valu, flush=True)
    mf.init()
    mesh = with_neumann_z(mf.Mesh.box_prism(3, 2, 3, seed=7, jitter=0.15))
    a = mesh.arrays()
    nel = mesh.nel
    rng = np.random.default_rng(12)
    pd = rng.choice(CASES[case], size=nel)
    ce = random_full(rng, nel)
    uniform = len(CASES[case]) == 1
    dg = mf.DgProblem(mesh, int(pd[0]) if uniform else pd)
    dg.set_coefficients(IND_ALL_DG, mf.COEF_FULL, coef_el=ce.ravel(), src_model=mf.SRC_DIFF_IN_CUBE)
    dg.set_penalty(1.5)
    dg.set_valu_blocks(valu)
    dg.assemble()
    A, b = dg_system_dense(a, pd, IND_ALL_DG, lambda e: ce[:, e], 1.5, mf.SRC_DIFF_IN_CUBE,
                           sparse=True)
    assert dg.ndof == len(b)
    err = mixed_block_of(dg, A)
    rhs = np.concatenate([dg.block_row(e)[1] for e in range(dg.nel)])
    print(f"  blocks {err:.2e}", flush=True)
    assert err <= 1e-12
    assert np.abs(rhs - b).max() <= 1e-12 * np.abs(b).max()
    dg.free()
    mesh.free()


@pytest.mark.parametrize("case,precon", [("complete_mixed", 1), ("complete_mixed", 3),
                                         ("complete3", 3), ("flat2", 1)])
def test_gmres_reaches_the_direct_solution(case, precon):
    import scipy.sparse.linalg as spl
    from oracle_lib import dg_system_dense
    print("DG bases GMRES", case, "precon", precon, flush=True)
    mf.init()
    mesh = with_neumann_z(mf.Mesh.box_prism(3, 3, 3, seed=9, jitter=0.1))
    a = mesh.arrays()
    pd = np.random.default_rng(4).choice(CASES[case], size=mesh.nel)
    cc = np.r_[np.eye(3).ravel(), np.zeros(12)]
    uniform = len(CASES[case]) == 1
    dg = mf.DgProblem(mesh, int(pd[0]) if uniform else pd)
    dg.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc, src_model=mf.SRC_DIFF_IN_CUBE)
    dg.assemble()
    if precon == 3:
        dg.fill_ilu()
    x, info = dg.gmres(precon=precon, krylbas=30, max_iter=4000, tol=1e-10)
    assert info["converged"], info
    A, b = dg_system_dense(a, pd, mf.IND_DIFF, lambda e: cc, 1.0, mf.SRC_DIFF_IN_CUBE, sparse=True)
    # the oracle system's own residual of the device solution
    assert np.linalg.norm(A @ x - b) <= 1e-8 * np.linalg.norm(b)
    if case.startswith("complete"):
        xs = spl.spsolve(A.tocsc(), b)
        assert np.abs(x - xs).max() <= 1e-7 * np.abs(xs).max()
    # (pz = 0 systems carry R's zero penalty and are ill-conditioned: the
    # residual is the measure there, under block Jacobi; their block ILU(0)
    # factors amplify it: GMRES meets 1e-10 on the preconditioned residual
    # with a true residual near 1e-6.  With pdeg 1 elements present, block-
    # Jacobi GMRES stagnates near 2e-5: without a penalty the in-plane linear
    # elements lack the stabilisation the interior-penalty form relies on at
    # low degree; their block rows are checked above)
    dg.free()
    mesh.free()


def test_bases_reject_what_they_do_not_cover():
    mf.init()
    mesh = mf.Mesh.box_prism(2, 2, 2, seed=1)
    for bad in (3, 1004, 1000, 4):
        with pytest.raises(RuntimeError):
            mf.DgProblem(mesh, bad)
    pd = np.full(mesh.nel, 1002)
    pd[0] = 202  # one family per mesh
    with pytest.raises(RuntimeError):
        mf.DgProblem(mesh, pd)
    pd = np.full(mesh.nel, 1)
    pd[0] = 101
    with pytest.raises(RuntimeError):
        mf.DgProblem(mesh, pd)
    mesh.free()
