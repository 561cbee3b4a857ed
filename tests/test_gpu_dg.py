"""DG prism element integrals on the device (FP64 MFMA kernel) against the
oracle's restatement of apr_num_int_el (apd_dg_prism/aps_dg_prism_intf.c:
1239-1728): every tensor degree 101..303, every term of the DG weak form,
per-element coefficients, the 2D coefficient branches and the exact-solution
source; bar 1e-12 relative to the largest entry of each element block."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import dg_num_int_el  # noqa: E402

pytestmark = pytest.mark.gpu
PDEGS = [101, 201, 301, 102, 202, 302, 103, 203, 303]
IND_DG_ALL = mf.coeff_indicator("m", "axx", "axy", "axz", "ayx", "ayy", "ayz", "azx", "azy",
                                "azz", "bx", "by", "bz", "c", "qx", "qy", "qz", "s")


def element_nodes(arrays):
    return arrays["coor"][arrays["el_nodes"]]  # [nel][6][3]


def random_full(rng, nel):
    ce = np.zeros((mf.NCOEF, nel))
    ce[0:9] = rng.uniform(-0.2, 0.2, (9, nel)) + (np.eye(3).ravel() * 1.5)[:, None]
    ce[9:12] = rng.uniform(-1, 1, (3, nel))
    ce[15] = rng.uniform(0.1, 1.0, nel)
    ce[16:19] = rng.uniform(-1, 1, (3, nel))
    ce[19] = rng.uniform(-1, 1, nel)
    ce[20] = rng.uniform(0.5, 1.5, nel)
    return ce


def compare(dg, nodes, ind, coefs_of, src=0):
    st, rhs = dg.element_matrices()
    worst = 0.0
    for e in range(dg.nel):
        K0, F0 = dg_num_int_el(dg.pdeg, nodes[e], ind, coefs_of(e), src)
        assert K0.shape == st[e].shape
        worst = max(worst, np.abs(st[e] - K0).max() / np.abs(K0).max(),
                    np.abs(rhs[e] - F0).max() / max(np.abs(F0).max(), 1e-300))
    assert worst <= 1e-12, worst
    return st, rhs


@pytest.mark.parametrize("pdeg", PDEGS)
def test_element_matrices_every_term(pdeg):
    mesh = mf.Mesh.box_prism(3, 3, 2, seed=pdeg, jitter=0.25)
    nodes = element_nodes(mesh.arrays())
    rng = np.random.default_rng(pdeg)
    ce = random_full(rng, mesh.nel)
    dg = mf.DgProblem(mesh, pdeg)
    assert dg.ns == (pdeg % 100 + 1) * (pdeg % 100 + 2) // 2 * (pdeg // 100 + 1)
    dg.set_coefficients(IND_DG_ALL, mf.COEF_FULL, coef_el=ce.ravel())
    st, rhs = compare(dg, nodes, IND_DG_ALL, lambda e: ce[:, e])
    # device-resident integration gives the same blocks
    dg.integrate()
    st2, rhs2 = dg.get_element_matrices()
    assert np.array_equal(st, st2) and np.array_equal(rhs, rhs2)
    dg.free()


@pytest.mark.parametrize("pdeg", [202, 303])
def test_2d_branches_and_kappa_models(pdeg):
    """axx..ayy without azz, bx, by without bz, qx, qy without qz; then
    per-element kappa and b (CM_KAPPA_B) with the DIFF_IN_CUBE source"""
    mesh = mf.Mesh.box_prism(2, 3, 2, seed=7, jitter=0.2)
    nodes = element_nodes(mesh.arrays())
    cc = np.zeros(mf.NCOEF)
    cc[0:9] = [2.0, 0.3, 0.1, -0.2, 1.5, 0.2, 0.1, 0.4, 1.2]
    cc[9:12] = [0.5, -0.7, 0.9]
    cc[15], cc[19], cc[20] = 0.3, 1.0, 0.8
    cc[16:19] = [0.2, -0.1, 0.4]
    ind2 = mf.coeff_indicator("axx", "axy", "ayx", "ayy", "bx", "by", "qx", "qy", "s", "m")
    dg = mf.DgProblem(mesh, pdeg)
    dg.set_coefficients(ind2, mf.COEF_UNIFORM, coef_const=cc)
    compare(dg, nodes, ind2, lambda e: cc)
    rng = np.random.default_rng(3)
    kap = np.exp(rng.normal(0, 1, mesh.nel))
    b = rng.uniform(-1, 1, (3, mesh.nel))
    ind3 = mf.coeff_indicator("axx", "axy", "axz", "ayx", "ayy", "ayz", "azx", "azy", "azz",
                              "bx", "by", "bz", "s")

    def full(e):
        c = cc.copy()
        c[0:9] = np.eye(3).ravel() * kap[e]
        c[9:12] = b[:, e]
        return c
    dg.set_coefficients(ind3, mf.COEF_KAPPA_B, coef_const=cc,
                        coef_el=np.concatenate([kap, b.ravel()]), src_model=mf.SRC_DIFF_IN_CUBE)
    compare(dg, nodes, ind3, full, src=mf.SRC_DIFF_IN_CUBE)
    dg.free()


def test_dg_rejects_bad_input():
    mesh = mf.Mesh.box_prism(2, 2, 1, seed=1)
    with pytest.raises(mf.MFError):
        mf.DgProblem(mesh, 104)
    with pytest.raises(mf.MFError):
        mf.DgProblem(mesh, 3)
    tet = mf.Mesh.box_tetra(2, 2, 1, seed=1)
    with pytest.raises(mf.MFError):
        mf.DgProblem(tet, 101)
