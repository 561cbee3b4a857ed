"""The reference's own known answers on the device path (libmodfem_hip.so),
from its ctest cases copied as data under tests/golden/:

  prism cube (Laplace_std_in_cube)   mesh dump reader, DIFF_IN_CUBE, bc 1
                                     Dirichlet penalty, bc 11 Neumann exact
                                     flux: the 8-value dump to 1e-9 and the
                                     L2 / H1 gates of conf.txt
  tetrahedral cube (..._tetra)       the L2 / H1 gates of conf.txt; its
                                     comparison dump is the prism cube's
                                     (not a tetrahedral answer): skipped
  LDC tetrahedral mesh               800 nodes, 2166 tetrahedra: assembled
  (Laplace_std_tetra)                system against the oracle (<= 1e-12),
                                     solution, and the L2 / H1 gates

Bars: dump values 1e-9 absolute; error norms below the gates; assembled
matrix and load max|A-A0| <= 1e-12 max|A0|.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import OracleProblem, std_errors, system_with_face_types  # noqa: E402
from test_golden import GOLD, load_dump, read_ldc_mesh  # noqa: E402

pytestmark = pytest.mark.gpu


def device_cube_solve(mesh, tol=1e-12):
    """bc >= 10 -> Neumann with the exact flux (pdr_get_bc_type), else
    Dirichlet penalty; GMRES(30) + Jacobi on the device"""
    a = mesh.arrays()
    p = mf.Problem.from_mesh(mesh)
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    p.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc, src_model=mf.SRC_DIFF_IN_CUBE)
    types = np.where(a["bf_bc"] >= 10, 2, 1).astype(np.int32)
    p.set_face_bc(types, np.zeros(mesh.nbf))
    p.integrate_assemble(64)
    x, info = p.gmres(precon=1, krylbas=30, max_iter=5000, tol=tol)
    assert info["converged"], info
    u = np.zeros(mesh.nno)
    u[p.dofs()] = x
    return p, u


def test_gpu_reproduces_reference_prism_cube_dump():
    m = mf.Mesh.read_dat(os.path.join(GOLD, "prism_cube", "mesh_prism.dat"))
    p, u = device_cube_solve(m)
    ids, ref = load_dump(os.path.join(GOLD, "prism_cube", "field.dmp"))
    print("prism cube: max |u - dump|", np.abs(u[ids - 1] - ref).max())
    assert np.abs(u[ids - 1] - ref).max() <= 1e-9
    l2, h1 = std_errors(m.arrays(), u, mf.SRC_DIFF_IN_CUBE)
    assert l2 < 0.024103 and h1 < 0.200045, (l2, h1)
    p.free()
    m.free()


def test_gpu_tetra_cube_meets_reference_error_gates():
    m = mf.Mesh.read_nas(os.path.join(GOLD, "tetra_cube", "mesh.nas"))
    p, u = device_cube_solve(m)
    l2, h1 = std_errors(m.arrays(), u, mf.SRC_DIFF_IN_CUBE)
    print(f"tetra cube: L2 {l2:.8f} (gate 0.03885) H1 {h1:.8f} (gate 0.37829)")
    assert 0.03885 - 1e-4 < l2 < 0.03885 and 0.37829 - 1e-4 < h1 < 0.37829
    p.free()
    m.free()


@pytest.mark.skip(reason="the tetrahedral case's comparison/field.dmp is byte for byte the "
                         "2-prism cube's dump (tests/test_golden.py::test_tetra_cube_comparison_"
                         "dump_is_the_prism_cube_dump); that dump is reproduced on the prism mesh "
                         "and the tetrahedral case is checked against its L2/H1 gates instead")
def test_gpu_reproduces_reference_tetra_cube_dump():
    m = mf.Mesh.read_nas(os.path.join(GOLD, "tetra_cube", "mesh.nas"))
    p, u = device_cube_solve(m)
    ids, ref = load_dump(os.path.join(GOLD, "tetra_cube", "field.dmp"))
    assert np.abs(u[ids - 1] - ref).max() <= 1e-6


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_gpu_ldc_tetra_mesh_against_oracle_and_gates(mode, tmp_path):
    m = read_ldc_mesh(tmp_path)
    a = m.arrays()
    op = OracleProblem(a)
    coef = np.zeros((op.nel, 21))
    coef[:, 0] = coef[:, 4] = coef[:, 8] = 1.0
    v0, r0 = system_with_face_types(op, mf.IND_DIFF, coef, mf.SRC_DIFF_IN_CUBE, a["bf_bc"] >= 10)
    p = mf.Problem.from_mesh(m)
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    p.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc, src_model=mf.SRC_DIFF_IN_CUBE)
    p.set_face_bc(np.where(a["bf_bc"] >= 10, 2, 1).astype(np.int32), np.zeros(m.nbf))
    p.set_assembly_mode(mode)
    p.integrate_assemble(64)
    assert np.array_equal(p.dofs(), op.dof_node)
    val, rhs = p.system()
    ev = np.abs(val - v0).max() / np.abs(v0).max()
    er = np.abs(rhs - r0).max() / np.abs(r0).max()
    print(f"LDC mode {mode}: matrix {ev:.2e} load {er:.2e}")
    assert ev <= 1e-12 and er <= 1e-12
    x, info = p.gmres(precon=1, krylbas=30, max_iter=20000, tol=1e-12)
    assert info["converged"]
    u = np.zeros(m.nno)
    u[p.dofs()] = x
    l2, h1 = std_errors(a, u, mf.SRC_DIFF_IN_CUBE)
    assert l2 < 0.000062 and h1 < 0.0066, (l2, h1)
    p.free()
    m.free()


def test_gpu_reproduces_reference_dg_cube_dump():
    """ctest Laplace_dg on the device: the 2-prism cube mesh dump, DG degree
    101, penalty constant 1, DIFF_IN_CUBE; GMRES with block Jacobi to 1e-13:
    the 12 element DOFs of comparison/field.dmp to 1e-10"""
    from test_golden import load_dg_dump
    m = mf.Mesh.read_dat(os.path.join(GOLD, "dg_cube", "mesh_prism.dmp"))
    _, ref = load_dg_dump(os.path.join(GOLD, "dg_cube", "field.dmp"))
    dg = mf.DgProblem(m, 101)
    dg.set_penalty(1.0)
    dg.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=np.r_[np.eye(3).ravel(), np.zeros(12)],
                        src_model=mf.SRC_DIFF_IN_CUBE)
    dg.assemble()
    x, info = dg.gmres(precon=1, krylbas=30, max_iter=1000, tol=1e-13)
    assert info["converged"], info
    print("DG cube: max |x - dump|", np.abs(x.reshape(ref.shape) - ref).max())
    assert np.abs(x.reshape(ref.shape) - ref).max() <= 1e-10
    dg.free()
    m.free()
