"""mf_gmres_noalloc (GMRES whose work arrays live in a fixed device array)
against mf_gmres through the C ABI: two calls in a row, the second on a
smaller problem (the array keeps the first call's contents), bitwise the
same iterates and solution; a start vector; the size limit reported as an
error."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402

pytestmark = pytest.mark.gpu


def problem(nx, ny, nz, seed, bc=1):
    mesh = mf.Mesh.box_prism(nx, ny, nz, seed=seed, jitter=0.2, bc=bc)
    rng = np.random.default_rng(seed)
    kappa = np.exp(rng.normal(0, 1, mesh.nel))
    b = rng.uniform(-1, 1, (mesh.nel, 3))
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    cc[15] = 0.0 if bc else 10.0
    cc[19] = 1.0
    p = mf.Problem.from_mesh(mesh)
    ind = mf.IND_CONV_DIFF if bc else mf.coeff_indicator("axx", "ayy", "azz", "bx", "by", "bz", "c", "s")
    p.set_coefficients(ind, mf.COEF_KAPPA_B, coef_const=cc,
                       coef_el=np.concatenate([kappa, b[:, 0], b[:, 1], b[:, 2]]))
    p.integrate_assemble(64)
    return mesh, p


def noalloc(p, precon, krylbas, max_iter, tol, x0=None):
    x = np.zeros(p.n) if x0 is None else np.array(x0, np.float64)
    it, rr, rs = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    conv = mf.check(mf.lib().mf_gmres_noalloc(p.h, precon, krylbas, max_iter, tol,
                                              int(x0 is None), x.ctypes.data, ctypes.byref(it),
                                              ctypes.byref(rr), ctypes.byref(rs)))
    return x, dict(converged=bool(conv), iters=it.value, res_rel=rr.value, resid=rs.value)


@pytest.mark.parametrize("precon,bc", [(1, 1), (0, 0)])
def test_noalloc_gmres_equals_gmres_on_two_calls(precon, bc):
    for (nx, ny, nz), seed in (((14, 12, 10), 3), ((7, 6, 5), 4)):
        mesh, p = problem(nx, ny, nz, seed, bc)
        x, info = p.gmres(precon=precon, krylbas=30, max_iter=3000, tol=1e-9)
        xn, infon = noalloc(p, precon, 30, 3000, 1e-9)
        print(f"noalloc {nx}x{ny}x{nz} precon {precon}: {infon['iters']} its, "
              f"equal {np.array_equal(x, xn)}")
        assert info["converged"] and infon["converged"]
        assert infon["iters"] == info["iters"] and infon["res_rel"] == info["res_rel"]
        assert np.array_equal(xn, x)
        # the problem's own solve still works after the scratch solve
        x2, info2 = p.gmres(precon=precon, krylbas=30, max_iter=3000, tol=1e-9)
        assert np.array_equal(x2, x)
        p.free()
        mesh.free()


def test_noalloc_gmres_start_vector_and_limit():
    mesh, p = problem(9, 8, 7, 5)
    x0 = np.random.default_rng(1).uniform(-1, 1, p.n)
    x, info = p.gmres(precon=1, krylbas=20, max_iter=3000, tol=1e-9, x0=x0)
    xn, infon = noalloc(p, 1, 20, 3000, 1e-9, x0=x0)
    assert infon["iters"] == info["iters"] and np.array_equal(xn, x)
    with pytest.raises(RuntimeError):
        noalloc(p, 2, 20, 3000, 1e-9)  # ILU is not on the no-allocation path
    p.free()
    mesh.free()
    big, pb = problem(64, 64, 64, 6)
    with pytest.raises(RuntimeError, match="128 MiB"):
        noalloc(pb, 1, 63, 10, 1e-9)
    xb, infob = noalloc(pb, 1, 30, 40, 1e-9)  # fits: same iterates as mf_gmres
    xr, infor = pb.gmres(precon=1, krylbas=30, max_iter=40, tol=1e-9)
    assert infob["iters"] == infor["iters"] == 40 and np.array_equal(xb, xr)
    pb.free()
    big.free()
