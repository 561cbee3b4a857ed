"""The reference's own seams driven by compiled C (tests/c/seam_driver.c,
built by build()): a ModFEM-style driver that includes only modfem_lsr.h,
modfem_lar.h, modfem_tmr.h and the mesh/element calls of modfem_hip.h.

  sir_create + sir_solve sequence     lsr_mkb_* over the device GMRES on the
  (solver assembling)                 reference's jk box mesh, DIFF_IN_BOX,
                                      penalty faces, the reference's control
                                      file: the reference's own dump
                                      field_box.dmp to 1e-9, and the
                                      oracle's GMRES iteration count within 5%
  GPU_ASSEMBLING create-assemble      tmr_* with the reference's two PDE
  sequence                            packings, three colours: the registered
                                      CRS values and rhs against the oracle's
                                      element assembly, <= 1e-12 relative
"""
import os
import subprocess
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import OracleProblem  # noqa: E402
from test_golden import GOLD, load_dump  # noqa: E402

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "seam_driver")


def run_driver(*args):
    if not os.path.exists(DRIVER):
        raise RuntimeError("tests/c/seam_driver is not built (run __graft_entry__.build())")
    r = subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("control", ["reference_block_storage", "crs_storage"])
def test_sir_sequence_through_lsr_seam_reproduces_reference_dump(control, tmp_path):
    """the reference's control file (GMRES(50), one block-Jacobi sweep, block
    storage: no assembly tables, entities assembled from their DOF lists),
    and the same solver with CRS storage forced (block type -2: assembly
    tables, as sir_create builds them for 1000+ DOFs)"""
    out = tmp_path / "x.bin"
    mesh_file = os.path.join(GOLD, "laplace_box_mesh_jk.dat")
    if control == "reference_block_storage":
        mkb = os.path.join(GOLD, "mkb", "laplace_std_in_cube.dat")
    else:
        mkb = tmp_path / "mkb.dat"
        mkb.write_text("FINE_LEVEL\n1 1 50\n1 1 -2\n500 0 1e-9\n0 1 0\n")
    txt = run_driver("sir", mesh_file, mkb, out)
    assert ("tables 0" in txt) == (control == "reference_block_storage")
    d = np.fromfile(out, dtype=np.float64)
    n, iters, meas = int(d[0]), int(d[1]), d[2]
    x = d[3:3 + n]
    m = mf.Mesh.read_jk(mesh_file)
    op = OracleProblem(m.arrays())
    u = np.zeros(m.nno)
    u[op.dof_node] = x
    ids, ref = load_dump(os.path.join(GOLD, "laplace_box_field.dmp"))
    err = np.abs(u[ids - 1] - ref).max()
    print(f"lsr seam: {iters} iterations, measure {meas:.2e}, max |u - dump| {err:.2e}")
    assert meas < 1e-12 and err <= 1e-9
    coef = np.zeros((op.nel, 21))
    coef[:, 0] = coef[:, 4] = coef[:, 8] = 1.0
    v0, r0 = op.assemble(mf.IND_DIFF, coef, src_model=mf.SRC_DIFF_IN_BOX)
    x0, info0 = op.gmres(v0, r0, precon=1, krylbas=50, max_iter=2000, tol=1e-12)
    assert info0["converged"] and abs(iters - info0["iters"]) <= max(3, info0["iters"] // 20)
    m.free()


@pytest.mark.parametrize("packing", [1, 2])
def test_accel_sequence_through_tmr_seam_matches_oracle(packing, tmp_path):
    out = tmp_path / "crs.bin"
    nx, ny, nz = 9, 8, 7
    run_driver("accel", nx, ny, nz, packing, out)
    d = np.fromfile(out, dtype=np.float64)
    n, nnz = int(d[0]), int(d[1])
    val, rhs = d[2:2 + nnz], d[2 + nnz:2 + nnz + n]
    pde = d[2 + nnz + n:]
    m = mf.Mesh.box_prism(nx, ny, nz, lx=1.0, ly=1.3, lz=0.7, seed=11, jitter=0.2, bc=0)
    op = OracleProblem(m.arrays())
    assert (op.n, op.nnz, op.nbf) == (n, nnz, 0)
    if packing == 1:
        coef = np.zeros((op.nel, 21))
        coef[:, 0] = coef[:, 4] = coef[:, 8] = 1.0
        v0, r0 = op.assemble(mf.IND_DIFF, coef, src_model=mf.SRC_DIFF_IN_BOX)
    else:
        p = pde.reshape(op.nel, 20)
        coef = np.zeros((op.nel, 21))
        coef[:, 0:20] = p  # axx..azz, bx..bz, tx..tz, c, qx..qz, s: the slot order
        v0, r0 = op.assemble(mf.IND_ALL, coef)
    ev = np.abs(val - v0).max() / np.abs(v0).max()
    er = np.abs(rhs - r0).max() / np.abs(r0).max()
    print(f"tmr seam packing {packing}: values {ev:.2e}, rhs {er:.2e}")
    assert ev <= 1e-12 and er <= 1e-12
    m.free()
