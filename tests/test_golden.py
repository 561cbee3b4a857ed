"""End-to-end known answer from the reference's own test tree:
ctest/conv_diff/Laplace_std_in_box_opencl -- the jk mesh mesh_jk.dat (1528
prisms, 1702 nodes, bottom bc 1, top bc 2, side bc 3 -> Dirichlet penalty),
problem DIFF_IN_BOX (u = exp(-(x^2+y^2+z^2)/100)), and its solution dump
comparison/field_box.dmp (node id, value; about 12 significant digits).
Copied as data into tests/golden/.  The oracle's integrate -> assemble ->
GMRES reproduces the dump to 1e-9 absolute."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import OracleProblem  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_dump(path):
    lines = open(path).read().split("\n")
    cnt = int(lines[1])
    ids = np.array([int(l.split()[0]) for l in lines[2: 2 + cnt]])
    vals = np.array([float(l.split()[1]) for l in lines[2: 2 + cnt]])
    return ids, vals


def test_oracle_reproduces_reference_field_dump():
    m = mf.Mesh.read_jk(os.path.join(GOLD, "laplace_box_mesh_jk.dat"))
    op = OracleProblem(m.arrays())
    coef = np.zeros((op.nel, 21))
    coef[:, 0] = coef[:, 4] = coef[:, 8] = 1.0
    val, rhs = op.assemble(mf.IND_DIFF, coef, src_model=mf.SRC_DIFF_IN_BOX)
    x, info = op.gmres(val, rhs, precon=1, krylbas=20, max_iter=1000, tol=1e-12)
    assert info["converged"]
    u = np.zeros(m.nno)
    u[op.dof_node] = x
    ids, ref = load_dump(os.path.join(GOLD, "laplace_box_field.dmp"))
    assert len(ids) == m.nno == 1702
    assert np.abs(u[ids - 1] - ref).max() <= 1e-9


def test_field_dump_read_write_reemits_the_reference_file(tmp_path):
    """the reference's own dump read into DOF order and written back is
    byte-identical (apr_write_field's format: header, entity count, node id
    order, %.12lg values with a trailing space); values below accuracy are
    written as 0; a dump missing a DOF is rejected"""
    m = mf.Mesh.read_jk(os.path.join(GOLD, "laplace_box_mesh_jk.dat"))
    src = os.path.join(GOLD, "laplace_box_field.dmp")
    x = m.read_field(src)
    ids, ref = load_dump(src)
    dof_node = np.zeros(len(x), np.int32)
    mf.check(mf.lib().mf_mesh_setup_get(m.h, dof_node.ctypes.data, None, None))
    u = np.zeros(m.nno)
    u[dof_node] = x
    assert np.array_equal(u[ids - 1], ref)
    out = tmp_path / "field.dmp"
    m.write_field(x, out)
    assert open(out).read() == open(src).read()
    m.write_field(x, out, accuracy=1.0)
    lines = open(out).read().split("\n")
    small = [l for l in lines[2:] if l and float(l.split()[1]) == 0.0]
    assert len(small) == int(np.sum(np.abs(ref) < 1.0)) and all(l.endswith(" 0 ") for l in small)
    trunc = tmp_path / "short.dmp"
    trunc.write_text("\n".join(open(src).read().split("\n")[:-3]) + "\n")
    with pytest.raises(RuntimeError):
        m.read_field(trunc)
    m.free()
