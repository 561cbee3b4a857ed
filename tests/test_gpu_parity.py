"""GPU parity: the HIP path (libmodfem_hip.so through its C ABI) against the
CPU oracle on identical seeded inputs.

Bars (stated per test):
  integer maps (DOF list, CRS pattern, tables)  bit-exact
  element / face matrices (FP64)                max|K-K0| <= 1e-12 max|K0| per entity
  assembled matrix and load (FP64)              max|A-A0| <= 1e-12 max|A0|
  lar_ residual / Jacobi sweep                  bit-exact with the reference order
  GMRES solution                                both converge to the 1e-8 preconditioned
                                                residual; max|x-x0| <= 1e-6 max|x0|
  FP32 integration                              max|K32-K64| <= 2e-6 max|K64|
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modfem_amd as mf  # noqa: E402
from oracle_lib import OracleProblem, P  # noqa: E402

pytestmark = pytest.mark.gpu


def coef_table(nel, kappa=None, b=None, full=None, s=1.0):
    c = np.zeros((nel, 21))
    if full is not None:
        return full
    k = np.ones(nel) if kappa is None else kappa
    c[:, 0] = c[:, 4] = c[:, 8] = k
    if b is not None:
        c[:, 9:12] = b
    c[:, 19] = s
    return c


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def per_entity_err(G, O):
    return max(np.abs(g - o).max() / max(np.abs(o).max(), 1e-300) for g, o in zip(G, O))


CASES = {
    "poisson_uniform": dict(ind=mf.IND_DIFF, model=mf.COEF_UNIFORM),
    "poisson_kappa": dict(ind=mf.IND_DIFF, model=mf.COEF_KAPPA),
    "convdiff_kappa_b": dict(ind=mf.IND_CONV_DIFF, model=mf.COEF_KAPPA_B),
    "all_terms_full": dict(ind=mf.IND_ALL, model=mf.COEF_FULL),
    "poisson_kappa_source": dict(ind=mf.IND_DIFF, model=mf.COEF_KAPPA_S),
}


def make_case(mesh, name, seed=5):
    rng = np.random.default_rng(seed)
    nel = mesh.nel
    kappa = np.exp(rng.normal(0, 1, nel))
    b = rng.uniform(-1, 1, (nel, 3))
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    cc[19] = 1.0
    spec = CASES[name]
    if spec["model"] == mf.COEF_UNIFORM:
        return spec["ind"], spec["model"], cc, None, coef_table(nel)
    if spec["model"] == mf.COEF_KAPPA:
        return spec["ind"], spec["model"], cc, kappa, coef_table(nel, kappa)
    if spec["model"] == mf.COEF_KAPPA_B:
        ce = np.concatenate([kappa, b[:, 0], b[:, 1], b[:, 2]])
        return spec["ind"], spec["model"], cc, ce, coef_table(nel, kappa, b)
    if spec["model"] == mf.COEF_KAPPA_S:
        src = rng.uniform(-1, 1, nel)
        return spec["ind"], spec["model"], cc, np.concatenate([kappa, src]), coef_table(nel, kappa, s=src)
    full = rng.uniform(-0.3, 0.3, (nel, 21))
    full[:, 0] += 2.0
    full[:, 4] += 2.0
    full[:, 8] += 2.0
    return spec["ind"], spec["model"], cc, full.T.copy().reshape(-1), full


@pytest.fixture(scope="module")
def prism_mesh():
    return mf.Mesh.box_prism(7, 6, 5, lx=1.0, ly=2.0, lz=0.5, seed=11, jitter=0.2)


@pytest.fixture(scope="module")
def tetra_mesh():
    return mf.Mesh.box_tetra(5, 4, 4, seed=3, jitter=0.15)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("kind", ["prism", "tetra"])
def test_element_matrices(case, kind, prism_mesh, tetra_mesh):
    mesh = prism_mesh if kind == "prism" else tetra_mesh
    p = mf.Problem.from_mesh(mesh)
    op = OracleProblem(mesh.arrays())
    ind, model, cc, ce, tab = make_case(mesh, case)
    p.set_coefficients(ind, model, coef_const=cc, coef_el=ce)
    E = p.element_matrices()
    k = 6 if kind == "prism" else 4
    errs = []
    for e in range(p.nel):
        K0, F0 = op.element(e, ind, tab[e])
        K = E[e, : k * k].reshape(k, k)
        F = E[e, k * k: k * k + k]
        errs.append(max(np.abs(K - K0).max() / np.abs(K0).max(),
                        np.abs(F - F0).max() / max(np.abs(F0).max(), 1e-300)))
    assert max(errs) <= 1e-12, (case, kind, max(errs))
    p.free()


@pytest.mark.parametrize("kind", ["prism", "tetra"])
def test_face_matrices(kind, prism_mesh, tetra_mesh):
    mesh = prism_mesh if kind == "prism" else tetra_mesh
    p = mf.Problem.from_mesh(mesh)
    op = OracleProblem(mesh.arrays())
    g = np.random.default_rng(2).uniform(-1, 1, p.nbf)
    p.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=np.r_[1, 0, 0, 0, 1, 0, 0, 0, 1,
                                                                         np.zeros(10), 1, 0],
                       bf_g=g)
    Fm = p.face_matrices()
    k = 6 if kind == "prism" else 4
    for f in range(p.nbf):
        K0, F0 = op.face(f, g[f])
        K = Fm[f, : k * k].reshape(k, k)
        F = Fm[f, k * k: k * k + k]
        assert np.abs(K - K0).max() <= 1e-12 * np.abs(K0).max()
        assert np.abs(F - F0).max() <= 1e-12 * max(np.abs(F0).max(), 1e-300) + 1e-300
    p.free()


@pytest.mark.parametrize("case", list(CASES))
def test_assembled_system(case, prism_mesh):
    p = mf.Problem.from_mesh(prism_mesh)
    op = OracleProblem(prism_mesh.arrays())
    rp, ci = p.pattern()
    assert np.array_equal(rp, op.row_ptr) and np.array_equal(ci, op.col_ind)
    assert np.array_equal(p.dofs(), op.dof_node)
    for ie in range(0, p.nel + p.nbf, 7):
        t0 = op.table[op.table_ptr[ie]: op.table_ptr[ie + 1]]
        assert np.array_equal(p.table(ie), t0)
    ind, model, cc, ce, tab = make_case(prism_mesh, case)
    p.set_coefficients(ind, model, coef_const=cc, coef_el=ce)
    p.integrate_assemble(64)
    val, rhs = p.system()
    v0, r0 = op.assemble(ind, tab)
    assert rel_err(val, v0) <= 1e-12
    assert rel_err(rhs, r0) <= 1e-12
    p.free()


@pytest.mark.parametrize("src", [mf.SRC_DIFF_IN_CUBE, mf.SRC_DIFF_IN_BOX])
def test_exact_solution_source(src, prism_mesh):
    p = mf.Problem.from_mesh(prism_mesh)
    op = OracleProblem(prism_mesh.arrays())
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    p.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc, src_model=src)
    p.integrate_assemble(64)
    val, rhs = p.system()
    v0, r0 = op.assemble(mf.IND_DIFF, coef_table(p.nel, s=0.0), src_model=src)
    assert rel_err(val, v0) <= 1e-12
    assert rel_err(rhs, r0) <= 1e-12
    p.free()


@pytest.mark.parametrize("precon", [0, 1])
def test_gmres_matches_oracle(precon):
    if precon == 1:
        # Dirichlet penalty faces (1e7): needs the Jacobi preconditioner
        mesh = mf.Mesh.box_prism(10, 9, 8, seed=21, jitter=0.2)
        ind, model, cc, ce, tab = make_case(mesh, "convdiff_kappa_b", seed=9)
    else:
        # no penalty faces; diffusion + reaction keeps the unpreconditioned
        # system well conditioned
        mesh = mf.Mesh.box_prism(10, 9, 8, seed=21, jitter=0.2, bc=0)
        ind = mf.coeff_indicator("axx", "ayy", "azz", "c", "s")
        model = mf.COEF_UNIFORM
        cc = np.zeros(21)
        cc[0] = cc[4] = cc[8] = 1.0
        cc[15] = 10.0
        cc[19] = 1.0
        ce = None
        tab = np.tile(cc, (mesh.nel, 1))
    p = mf.Problem.from_mesh(mesh)
    op = OracleProblem(mesh.arrays())
    p.set_coefficients(ind, model, coef_const=cc, coef_el=ce)
    p.integrate_assemble(64)
    x, info = p.gmres(precon=precon, krylbas=30, max_iter=3000, tol=1e-8)
    v0, r0 = op.assemble(ind, tab)
    x0, info0 = op.gmres(v0, r0, precon=precon, krylbas=30, max_iter=3000, tol=1e-8)
    assert info0["converged"], info0
    assert info["converged"], info
    assert info["res_rel"] < 1e-8
    # iteration counts agree up to rounding-driven drift
    assert abs(info["iters"] - info0["iters"]) <= max(3, info0["iters"] // 20), (info, info0)
    assert rel_err(x, x0) <= 1e-6
    # device-resident solve: same iterates, solution fetched afterwards
    xn, info_n = p.gmres(precon=precon, krylbas=30, max_iter=3000, tol=1e-8, fetch=False)
    assert xn is None and info_n["iters"] == info["iters"]
    assert np.array_equal(p.solution(), x)
    p.free()


def test_gmres_dense_reference():
    """small nonsymmetric system: GMRES solution equals the dense LU solve"""
    rng = np.random.default_rng(4)
    n = 60
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - 3), min(n, i + 4)):
            A[i, j] = rng.uniform(-1, 1)
        A[i, i] = 8.0 + rng.uniform(0, 1)
    rp = [0]
    ci, v = [], []
    for i in range(n):
        nz = np.nonzero(A[i])[0]
        ci += list(nz)
        v += list(A[i, nz])
        rp.append(len(ci))
    b = rng.uniform(-1, 1, n)
    p = mf.Problem.from_crs(np.array(rp), np.array(ci), np.array(v), b)
    x, info = p.gmres(precon=1, krylbas=n, max_iter=200, tol=1e-12)
    xd = np.linalg.solve(A, b)
    assert info["converged"]
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()
    # Krylov invariance: with m >= n GMRES finishes within n steps
    assert info["iters"] <= n
    p.free()


def test_fp32_integration_error(prism_mesh):
    p = mf.Problem.from_mesh(prism_mesh)
    ind, model, cc, ce, tab = make_case(prism_mesh, "convdiff_kappa_b")
    p.set_coefficients(ind, model, coef_const=cc, coef_el=ce)
    E64 = p.element_matrices(precision=64)
    E32 = p.element_matrices(precision=32)
    err = per_entity_err(E32, E64)
    assert err <= 2e-6, err
    p.integrate_assemble(32)
    v32, r32 = p.system()
    p.integrate_assemble(64)
    v64, r64 = p.system()
    assert rel_err(v32, v64) <= 2e-6
    p.free()


def test_large_mesh_invariants():
    """size-independent properties on a 40^3 jittered box: zero row sums of the
    pure-diffusion element part, total load equals the volume, symmetric
    pattern"""
    mesh = mf.Mesh.box_prism(40, 40, 40, seed=5, jitter=0.2, bc=0)
    p = mf.Problem.from_mesh(mesh)
    kappa = np.exp(np.random.default_rng(1).normal(0, 1, p.nel))
    cc = np.zeros(21)
    cc[19] = 1.0
    p.set_coefficients(mf.IND_DIFF, mf.COEF_KAPPA, coef_const=cc, coef_el=kappa)
    p.integrate_assemble(64)
    val, rhs = p.system()
    rp, ci = p.pattern()
    rows = np.repeat(np.arange(p.n), np.diff(rp))
    rowsum = np.bincount(rows, weights=val, minlength=p.n)
    assert np.abs(rowsum).max() <= 1e-11 * np.abs(val).max()
    assert abs(rhs.sum() - 1.0) <= 1e-12
    import scipy.sparse as sp
    A = sp.csr_matrix((val, ci, rp), shape=(p.n, p.n))
    assert abs(A - A.T).max() <= 1e-13 * np.abs(val).max()
    p.free()


def test_gpu_reproduces_reference_field_dump():
    """the reference's own known answer (ctest Laplace_std_in_box_opencl) on
    the device path: jk mesh, DIFF_IN_BOX, penalty faces, GMRES"""
    from test_golden import GOLD, load_dump
    m = mf.Mesh.read_jk(os.path.join(GOLD, "laplace_box_mesh_jk.dat"))
    p = mf.Problem.from_mesh(m)
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    p.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc, src_model=mf.SRC_DIFF_IN_BOX)
    p.integrate_assemble(64)
    x, info = p.gmres(precon=1, krylbas=20, max_iter=1000, tol=1e-12)
    assert info["converged"]
    u = np.zeros(m.nno)
    u[p.dofs()] = x
    ids, ref = load_dump(os.path.join(GOLD, "laplace_box_field.dmp"))
    assert np.abs(u[ids - 1] - ref).max() <= 1e-9
    p.free()


@pytest.mark.parametrize("kind", ["prism", "tetra"])
def test_assembly_modes_agree_and_rowwise_is_reproducible(kind, prism_mesh, tetra_mesh):
    """row-wise (deterministic) and element-atomic assembly agree to 1e-12;
    row-wise is bitwise identical across runs"""
    mesh = prism_mesh if kind == "prism" else tetra_mesh
    p = mf.Problem.from_mesh(mesh)
    ind, model, cc, ce, tab = make_case(mesh, "all_terms_full")
    p.set_coefficients(ind, model, coef_const=cc, coef_el=ce)
    p.integrate_assemble(64)
    v0, r0 = p.system()
    p.integrate_assemble(64)
    v1, r1 = p.system()
    assert np.array_equal(v0, v1) and np.array_equal(r0, r1)
    p.set_assembly_mode(1)
    p.integrate_assemble(64)
    v2, r2 = p.system()
    assert rel_err(v2, v0) <= 1e-12 and rel_err(r2, r0) <= 1e-12
    p.set_assembly_mode(2)
    p.integrate_assemble(64)
    v3, r3 = p.system()
    p.integrate_assemble(64)
    v4, r4 = p.system()
    assert np.array_equal(v3, v4) and np.array_equal(r3, r4)
    assert rel_err(v3, v0) <= 1e-12 and rel_err(r3, r0) <= 1e-12
    # mixed-precision form: FP32 element rows, FP64 sums
    p.integrate_assemble(32)
    v5, r5 = p.system()
    assert rel_err(v5, v0) <= 2e-6 and rel_err(r5, r0) <= 2e-6
    p.free()


@pytest.mark.parametrize("conv_type", [0, 1])
def test_solve_with_reference_control_file(conv_type):
    """GMRES driven by the reference's own Jacobi control file
    (ctest Laplace_std_in_cube/mkb.dat): REL_RES_INI as written, and ABS_RES
    through the override argument, against the oracle with the same
    tolerances (lss_mkb_core.c:2220-2225, 2377-2378)"""
    G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mkb")
    tol = 1e-9 if conv_type == 0 else 1e-7
    prm = mf.mkb_read(os.path.join(G, "laplace_std_in_cube.dat"), -1, conv_type, tol, -1)
    assert prm["conv_type"] == conv_type and prm["precon"] == 1
    mesh = mf.Mesh.box_prism(8, 7, 6, seed=31, jitter=0.2)
    ind, model, cc, ce, tab = make_case(mesh, "convdiff_kappa_b", seed=5)
    p = mf.Problem.from_mesh(mesh)
    p.set_coefficients(ind, model, coef_const=cc, coef_el=ce)
    p.integrate_assemble(64)
    x, info = p.solve_mkb(prm)
    op = OracleProblem(mesh.arrays())
    v0, r0 = op.assemble(ind, tab)
    x0, info0 = op.gmres(v0, r0, precon=1, krylbas=prm["krylbas"], max_iter=prm["max_iter"],
                         tol=tol if conv_type == 0 else 0.0, tol_abs=tol if conv_type == 1 else 0.0)
    assert info0["converged"] and info["converged"]
    assert abs(info["iters"] - info0["iters"]) <= max(3, info0["iters"] // 20), (info, info0)
    if conv_type == 1:
        assert info["resid"] < tol
    assert rel_err(x, x0) <= 1e-5
    with pytest.raises(RuntimeError):  # additive Schwarz is not available
        p.solve_mkb(dict(prm, precon=3))
    p.free()


@pytest.mark.parametrize("precon", [1, 2])
def test_mixed_precision_gmres(precon):
    """FP64 refinement around the FP32-valued operator reaches the plain
    solve's stopping measure, checked independently in FP64 on the host:
    ||D^-1 (b - A x)|| <= tol ||D^-1 b|| with the oracle-assembled system"""
    import scipy.sparse as sp
    mesh = mf.Mesh.box_prism(12, 11, 10, seed=41, jitter=0.2)
    ind, model, cc, ce, tab = make_case(mesh, "convdiff_kappa_b", seed=13)
    p = mf.Problem.from_mesh(mesh)
    p.set_coefficients(ind, model, coef_const=cc, coef_el=ce)
    p.integrate_assemble(64)
    tol = 1e-8
    x, info = p.gmres_mixed(precon=precon, krylbas=30, max_iter=5000, tol=tol)
    assert info["converged"] and info["refinements"] >= 1, info
    op = OracleProblem(mesh.arrays())
    v0, r0 = op.assemble(ind, tab)
    A = sp.csr_matrix((v0, op.col_ind, op.row_ptr), shape=(op.n, op.n))
    if precon == 1:
        d = 1.0 / A.diagonal()
        prec = lambda v: d * v  # noqa: E731
    else:
        from oracle_lib import oracle_ilu0, oracle
        lu = oracle_ilu0(op.row_ptr, op.col_ind, v0)

        def prec(v):
            out = np.zeros(op.n)
            oracle().orc_ilu_solve(op.n, P(op.row_ptr), P(op.col_ind), P(lu),
                                   P(np.ascontiguousarray(v)), P(out))
            return out
    rel = np.linalg.norm(prec(r0 - A @ x)) / np.linalg.norm(prec(r0))
    assert rel <= tol * 1.0001, (rel, info)
    assert abs(rel - info["res_rel"]) <= 1e-3 * tol
    xp, infop = p.gmres(precon=precon, krylbas=30, max_iter=5000, tol=tol)
    # both solutions satisfy the same 1e-8 preconditioned-residual test
    assert rel_err(x, xp) <= 1e-5
    xn, info_n = p.gmres_mixed(precon=precon, krylbas=30, max_iter=5000, tol=tol, fetch=False)
    assert xn is None and np.array_equal(p.solution(), x)
    p.free()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_dirichlet_and_neumann_faces(mode):
    """per-face types: Dirichlet penalty faces with data g and Neumann faces
    with flux g (pds_conv_diff_weakform_std.c:480-522), in every assembly
    mode; the oracle system is built with penalty on every face and the
    Neumann faces' penalty replaced by their flux through the tables"""
    mesh = mf.Mesh.box_prism(5, 4, 3, seed=7, jitter=0.2)
    op = OracleProblem(mesh.arrays())
    ind, model, cc, ce, tab = make_case(mesh, "poisson_kappa", seed=3)
    rng = np.random.default_rng(8)
    types = rng.choice([1, 2], op.nbf).astype(np.int32)
    g = rng.uniform(-1, 1, op.nbf)
    p = mf.Problem.from_mesh(mesh)
    p.set_coefficients(ind, model, coef_const=cc, coef_el=ce)
    p.set_face_bc(types, g)
    p.set_assembly_mode(mode)
    p.integrate_assemble(64)
    val, rhs = p.system()
    v0, r0 = op.assemble(ind, tab, bf_g=g)
    v0, r0 = v0.copy(), r0.copy()
    k = 6
    for f in np.nonzero(types == 2)[0]:
        K, F = op.face(int(f), g=float(g[f]))
        t = op.table[op.table_ptr[op.nel + f]:op.table_ptr[op.nel + f + 1]]
        for w in range(k):
            for u in range(k):
                v0[t[w + u * k]] -= K[u][w]
            r0[t[k * k + w]] += op.face_neumann(int(f), float(g[f]))[w] - F[w]
    assert rel_err(val, v0) <= 1e-12 and rel_err(rhs, r0) <= 1e-12
    # face matrices: Neumann faces have no stiffness
    Fm = p.face_matrices()
    for f in range(op.nbf):
        if types[f] == 2:
            assert np.abs(Fm[f, :36]).max() == 0
            assert np.abs(Fm[f, 36:42] - op.face_neumann(f, float(g[f]))).max() <= 1e-14
    p.free()


@pytest.mark.parametrize("kind,mode", [("prism", 0), ("prism", 1), ("tetra", 0), ("tetra", 2)])
def test_neumann_faces_of_an_exact_solution_problem(kind, mode):
    """DIFF_IN_CUBE with Neumann faces: the face data is the exact normal
    flux grad u . n at every face point (pdr_bc_neum_coeff), not a constant;
    face right-hand sides and the assembled system against the oracle"""
    mesh = (mf.Mesh.box_prism(4, 3, 3, seed=11, jitter=0.2) if kind == "prism"
            else mf.Mesh.box_tetra(3, 3, 2, seed=11, jitter=0.2))
    op = OracleProblem(mesh.arrays())
    rng = np.random.default_rng(12)
    types = rng.choice([1, 2], op.nbf).astype(np.int32)
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    p = mf.Problem.from_mesh(mesh)
    p.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc, src_model=mf.SRC_DIFF_IN_CUBE)
    p.set_face_bc(types, np.zeros(op.nbf))
    p.set_assembly_mode(mode)
    p.integrate_assemble(64)
    Fm = p.face_matrices()
    scale = 0.0
    for f in np.nonzero(types == 2)[0]:
        ref = op.face_neumann(int(f), 0.0, src_model=mf.SRC_DIFF_IN_CUBE)
        k = len(ref)
        scale = max(scale, np.abs(ref).max())
        assert np.abs(Fm[f, k * k:k * k + k] - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())
    assert scale > 1e-3  # the flux is not zero on these faces
    # the assembled system: penalty everywhere in the oracle, then the
    # Neumann faces' penalty swapped for their exact flux
    coef = np.tile(cc, (op.nel, 1))
    v0, r0 = op.assemble(mf.IND_DIFF, coef, src_model=mf.SRC_DIFF_IN_CUBE)
    v0, r0 = v0.copy(), r0.copy()
    val, rhs = p.system()
    for f in np.nonzero(types == 2)[0]:
        K, F = op.face(int(f), g=0.0, src_model=mf.SRC_DIFF_IN_CUBE)
        k = K.shape[0]
        t = op.table[op.table_ptr[op.nel + f]:op.table_ptr[op.nel + f + 1]]
        flux = op.face_neumann(int(f), 0.0, src_model=mf.SRC_DIFF_IN_CUBE)
        for w in range(k):
            for u in range(k):
                v0[t[w + u * k]] -= K[u][w]
            r0[t[k * k + w]] += flux[w] - F[w]
    assert rel_err(val, v0) <= 1e-12 and rel_err(rhs, r0) <= 1e-12
    p.free()



def star_mesh(k=40):
    """tetrahedra around one vertex of valence k + 2: a ring of k points,
    an apex above and below; the centre's matrix row has k + 3 entries"""
    ang = 2 * np.pi * np.arange(k) / k
    ring = np.c_[np.cos(ang), np.sin(ang), np.zeros(k)]
    coor = np.vstack([[0.0, 0.0, 0.0], ring, [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    top, bot = k + 1, k + 2
    el, bf_elem, bf_side = [], [], []
    for i in range(k):
        a, b = 1 + i, 1 + (i + 1) % k
        el.append([0, a, b, top, -1, -1])   # positive volume
        bf_elem.append(2 * i); bf_side.append(3)       # (a, b, top)
        el.append([0, b, a, bot, -1, -1])
        bf_elem.append(2 * i + 1); bf_side.append(3)   # (b, a, bot)
    el = np.array(el, np.int32)
    return mf.Mesh.from_arrays(coor, np.full(len(el), mf.TETRA, np.int32), el,
                               np.array(bf_elem, np.int32), np.array(bf_side, np.int32),
                               np.ones(len(bf_elem), np.int32))


@pytest.mark.parametrize("mode", [0, 2])
def test_rows_longer_than_32_entries_assemble_through_the_atomic_path(mode):
    """a vertex of valence 42 (row of 43 entries): the row-wise and
    two-pass modes hand the mesh to the element-atomic path, whose system
    matches the oracle to 1e-12"""
    m = star_mesh(40)
    op = OracleProblem(m.arrays())
    assert np.diff(op.row_ptr).max() == 43
    p = mf.Problem.from_mesh(m)
    cc = np.zeros(21)
    cc[0] = cc[4] = cc[8] = 1.0
    cc[19] = 1.0
    p.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc)
    p.set_assembly_mode(mode)
    p.integrate_assemble(64)
    val, rhs = p.system()
    v0, r0 = op.assemble(mf.IND_DIFF, np.tile(cc, (op.nel, 1)))
    assert rel_err(val, v0) <= 1e-12 and rel_err(rhs, r0) <= 1e-12
    p.free()
    m.free()
