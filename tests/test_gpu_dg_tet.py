"""DG on tetrahedra and on hybrid prism/tet meshes (no reference
counterpart: the reference's DG module is prism-only,
apd_dg_prism/aph_dg_prism.h:1-3), against the oracle's restatement, which
is pinned by closed-form integrals, quadrature exactness, face-point
coincidence and linear-solution reproduction (tests/test_oracle_dg_tet.py):
- assembled block rows and load vector <= 1e-12 of the oracle's assembly on
  tet and hybrid meshes (uniform degrees 1-3 and degrees mixed per element;
  anisotropic diffusion, convection with upwinding, Dirichlet and Neumann
  data, penalty 1.5), from the vector-ALU and the MFMA kernels;
- block-Jacobi and block-ILU(0) GMRES reach the oracle system's direct
  solution (the oracle system itself reproduces linear solutions to
  rounding on hybrid meshes)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from test_gpu_dg import mixed_block_of, random_full  # noqa: E402

pytestmark = pytest.mark.gpu

IND_ALL_DG = mf.coeff_indicator("m", "axx", "axy", "axz", "ayx", "ayy", "ayz", "azx", "azy",
                                "azz", "bx", "by", "bz", "c", "s")


def with_neumann_z(mesh):
    """the box's z = 0 and z = max faces Neumann (flag 11), the rest Dirichlet"""
    a = mesh.arrays()
    X, en, et = a["coor"], a["el_nodes"], a["el_type"]
    zmax = X[:, 2].max()
    bc = a["bf_bc"].copy()
    for k, (e, s) in enumerate(zip(a["bf_elem"], a["bf_side"])):
        fv = ([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]][s] if et[e] == 7 else
              [[0, 1, 2], [3, 4, 5], [0, 1, 4, 3], [1, 2, 5, 4], [2, 0, 3, 5]][s])
        z = X[en[e][fv], 2]
        if np.all(np.abs(z) < 1e-12) or np.all(np.abs(z - zmax) < 1e-12):
            bc[k] = 11
    m = mf.Mesh.from_arrays(a["coor"], a["el_type"], a["el_nodes"], a["bf_elem"], a["bf_side"], bc)
    mesh.free()
    return m


def meshes():
    yield "tets", with_neumann_z(mf.Mesh.box_tetra(2, 2, 2, seed=3, jitter=0.15))
    yield "hybrid", with_neumann_z(mf.Mesh.box_hybrid(2, 2, 4, seed=5, jitter=0.15, tet_frac=0.5))


@pytest.mark.parametrize("valu", [6, 0])
@pytest.mark.parametrize("degrees", ["101", "202", "303", "mixed"])
def test_block_system_matches_oracle(degrees, valu):
    from oracle_lib import dg_system_dense
    for name, mesh in meshes():
        a = mesh.arrays()
        nel = mesh.nel
        rng = np.random.default_rng(11)
        pd = (rng.choice([101, 202, 303], size=nel) if degrees == "mixed"
              else np.full(nel, int(degrees)))
        ce = random_full(rng, nel)
        dg = mf.DgProblem(mesh, pd if degrees == "mixed" else int(degrees))
        dg.set_coefficients(IND_ALL_DG, mf.COEF_FULL, coef_el=ce.ravel(), src_model=mf.SRC_DIFF_IN_CUBE)
        dg.set_penalty(1.5)
        dg.set_valu_blocks(valu)
        dg.assemble()
        A, b = dg_system_dense(a, pd, IND_ALL_DG, lambda e: ce[:, e], 1.5, mf.SRC_DIFF_IN_CUBE,
                               sparse=True)
        assert dg.ndof == len(b), name
        err = mixed_block_of(dg, A)
        rhs = np.concatenate([dg.block_row(e)[1] for e in range(dg.nel)])
        print(f"{name} {degrees} valu {valu}: blocks {err:.2e}")
        assert err <= 1e-12, name
        assert np.abs(rhs - b).max() <= 1e-12 * np.abs(b).max(), name
        dg.free()
        mesh.free()


@pytest.mark.parametrize("precon", [1, 3])
def test_gmres_reaches_the_direct_solution(precon):
    """block Jacobi and block ILU(0) GMRES on the hybrid DIFF_IN_CUBE system
    with degrees mixed per element"""
    import scipy.sparse.linalg as spl
    from oracle_lib import dg_system_dense
    mesh = with_neumann_z(mf.Mesh.box_hybrid(3, 3, 4, seed=8, jitter=0.1, tet_frac=0.5))
    a = mesh.arrays()
    pd = np.random.default_rng(4).choice([101, 202, 303], size=mesh.nel)
    cc = np.r_[np.eye(3).ravel(), np.zeros(12)]
    dg = mf.DgProblem(mesh, pd)
    dg.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc, src_model=mf.SRC_DIFF_IN_CUBE)
    dg.assemble()
    if precon == 3:
        dg.fill_ilu()
    x, info = dg.gmres(precon=precon, krylbas=30, max_iter=4000, tol=1e-10)
    assert info["converged"], info
    A, b = dg_system_dense(a, pd, mf.IND_DIFF, lambda e: cc, 1.0, mf.SRC_DIFF_IN_CUBE, sparse=True)
    xs = spl.spsolve(A.tocsc(), b)
    assert np.abs(x - xs).max() <= 1e-7 * np.abs(xs).max()
    dg.free()
    mesh.free()
