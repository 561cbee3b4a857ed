"""modfem_amd -- MI355X-native ModFEM hot path (integrate -> assemble -> GMRES).

Python mirror of the C ABI in ``include/modfem_hip.h`` (``libmodfem_hip.so``).
It follows the reference's call sequence for one solve
(sid_mkb/sis_mkb_intf.c: sir_create -> sir_solve): build the mesh input,
create the problem (DOF lists, CRS pattern, assembly tables), bind the PDE
coefficients, integrate + assemble on the device, then run GMRES.

The library is loaded from this directory; there is no CPU fallback: if the
HIP library is missing or fails to initialise, every call raises.
"""
import ctypes
import os

import numpy as np

__all__ = ["lib", "Mesh", "Problem", "MFError", "PRISM", "TETRA",
           "COEF_UNIFORM", "COEF_KAPPA", "COEF_KAPPA_B", "COEF_FULL", "COEF_KAPPA_S", "DgProblem",
           "SRC_CONST", "SRC_DIFF_IN_CUBE", "SRC_DIFF_IN_BOX", "IND",
           "coeff_indicator", "init"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmodfem_hip.so")

PRISM, TETRA = 5, 7
COEF_UNIFORM, COEF_KAPPA, COEF_KAPPA_B, COEF_FULL, COEF_KAPPA_S = 0, 1, 2, 3, 4
SRC_CONST, SRC_DIFF_IN_CUBE, SRC_DIFF_IN_BOX = 0, 1, 2
# coefficient slots: a[9] b[3] t[3] c q[3] s m
NCOEF = 21
C_A, C_B, C_T, C_C, C_Q, C_S, C_M = 0, 9, 12, 15, 16, 19, 20
# reference coefficient-indicator slots (pds_conv_diff_weakform.c:87-276)
IND = {"m": 1, "axx": 2, "axy": 3, "axz": 4, "ayx": 5, "ayy": 6, "ayz": 7,
       "azx": 8, "azy": 9, "azz": 10, "bx": 11, "by": 12, "bz": 13, "tx": 14,
       "ty": 15, "tz": 16, "c": 17, "l": 18, "qx": 19, "qy": 20, "qz": 21,
       "s": 22}


def coeff_indicator(*names):
    """coeff_vect_ind[23] with slot 0 set and the named slots selected."""
    ind = np.zeros(23, dtype=np.int32)
    ind[0] = 1
    for n in names:
        ind[IND[n]] = 1
    return ind


# DIFF_IN_CUBE / DIFF_IN_BOX selection (pds_conv_diff_weakform.c:103-113)
IND_DIFF = coeff_indicator("axx", "ayy", "azz", "s")
# DG degree code of the complete prism base of degree p: DG_COMPLETE + p
DG_COMPLETE = 1000
# generic conv-diff selection used for convection-diffusion workloads
IND_CONV_DIFF = coeff_indicator("axx", "axy", "axz", "ayx", "ayy", "ayz",
                                "azx", "azy", "azz", "bx", "by", "bz", "s")
# TEST_NUM_INT: every term (pds_conv_diff_weakform.c:114-160)
IND_ALL = coeff_indicator("axx", "axy", "axz", "ayx", "ayy", "ayz", "azx",
                          "azy", "azz", "bx", "by", "bz", "tx", "ty", "tz",
                          "c", "qx", "qy", "qz", "s")


class MFError(RuntimeError):
    pass


_lib = None
_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)


def _declare(L):
    I, D, LL, S = ctypes.c_int, ctypes.c_double, ctypes.c_longlong, ctypes.c_char_p
    V = ctypes.c_void_p
    sig = {
        "mf_init": (I, [I]),
        "mf_last_error": (S, []),
        "mf_version": (I, []),
        "mf_mesh_create": (I, [I, V, I, V, V, I, V, V, V]),
        "mf_mesh_read_jk": (I, [S]),
        "mf_mesh_read_nas": (I, [ctypes.c_char_p]),
        "mf_mesh_read_dat": (I, [ctypes.c_char_p]),
        "mf_mesh_write_jk": (I, [I, S]),
        "mf_mesh_box_prism": (I, [I, I, I, D, D, D, ctypes.c_ulonglong, D, I]),
        "mf_mesh_box_tetra": (I, [I, I, I, D, D, D, ctypes.c_ulonglong, D, I]),
        "mf_mesh_box_hybrid": (I, [I, I, I, D, D, D, ctypes.c_ulonglong, D, I, D]),
        "mf_mesh_sizes": (I, [I, V, V, V]),
        "mf_mesh_get": (I, [I, V, V, V, V, V, V]),
        "mf_mesh_free": (I, [I]),
        "mf_mesh_setup": (I, [I, V, V]),
        "mf_mesh_setup_get": (I, [I, V, V, V]),
        "mf_field_write": (I, [I, V, D, ctypes.c_char_p]),
        "mf_field_read": (I, [I, ctypes.c_char_p, V]),
        "mf_mesh_setup_table": (I, [I, I, V]),
        "mf_mesh_setup_order": (I, [I, V, V]),
        "mf_problem_create": (I, [I]),
        "mf_problem_create_device": (I, [I]),
        "mf_problem_create_crs": (I, [I, V, V, V, V]),
        "mf_problem_sizes": (I, [I, V, V, V, V, V]),
        "mf_problem_get_dofs": (I, [I, V]),
        "mf_problem_get_pattern": (I, [I, V, V]),
        "mf_problem_get_table": (I, [I, I, V]),
        "mf_problem_set_coefficients": (I, [I, V, I, V, V, I, V]),
        "mf_integrate_assemble": (I, [I, I]),
        "mf_integrate_assemble_async": (I, [I, I]),
        "mf_element_matrices": (I, [I, I, I, I, V]),
        "mf_face_matrices": (I, [I, I, I, V]),
        "mf_get_system": (I, [I, V, V]),
        "mf_set_system": (I, [I, V, V]),
        "mf_spmv": (I, [I, V, V]),
        "mf_fill_precon": (I, [I, I]),
        "mf_gmres": (I, [I, I, I, I, D, I, V, V, V, V]),
        "mf_gmres_noalloc": (I, [I, I, I, I, D, I, V, V, V, V]),
        "mf_true_residual": (I, [I, V, V]),
        "mf_get_solution": (I, [I, V]),
        "mf_problem_set_face_bc": (I, [I, V, V]),
        "mf_set_ilu_order": (I, [I, I]),
        "mf_set_gs": (I, [I, I]),
        "mf_gs_get": (I, [I, V, V]),
        "mf_ilu_get": (I, [I, V, V, V]),
        "mf_batch_create": (I, [I]),
        "mf_batch_add": (I, [I, I, I, I, ctypes.c_double]),
        "mf_batch_add_ex": (I, [I, I, I, I, ctypes.c_double, I]),
        "mf_batch_set_stream": (I, [I]),
        "mf_batch_add_dg": (I, [I, I, I, ctypes.c_double, I]),
        "mf_batch_remove_dg": (I, [I, I]),
        "mf_batch_status_dg": (I, [I, I, V, V, V, V, V]),
        "mf_batch_run_async": (I, [I, I]),
        "mf_batch_wait": (I, [I, V]),
        "mf_batch_status_ex": (I, [I, I, V, V, V, V, V]),
        "mf_batch_remove": (I, [I, I]),
        "mf_batch_integrate_assemble": (I, [I, I]),
        "mf_batch_run": (I, [I, I, I, V]),
        "mf_batch_status": (I, [I, I, V, V, V, V]),
        "mf_batch_free": (I, [I]),
        "mf_batch_set_tiles": (I, [I, I]),
        "mf_batch_tiles": (I, [I, V, V, V]),
        "mf_batch_tile_candidates": (I, [V]),
        "mf_batch_tile_cache_clear": (I, []),
        "mf_gmres_mixed": (I, [I, I, I, I, ctypes.c_double, V, V, V, V, V]),
        "mf_mkb_read": (I, [ctypes.c_char_p, I, I, ctypes.c_double, I, V, V]),
        "mf_solve_mkb": (I, [I, V, ctypes.c_double, I, V, V, V, V]),
        "mf_bench_integrate": (I, [I, I, I, V, V]),
        "mf_bench_spmv": (I, [I, I, V]),
        "mf_bench_spmv_ex": (I, [I, I, I, I, V]),
        "mf_last_solve_ms": (I, [I, V]),
        "mf_bench_copy": (I, [ctypes.c_longlong, I, V]),
        "mf_gmres_kernel_stats": (I, [I, V]),
        "mf_set_assembly_mode": (I, [I, I]),
        "mf_problem_free": (I, [I]),
        "mf_dg_create": (I, [I, I]),
        "mf_dg_sizes": (I, [I, V, V, V]),
        "mf_dg_set_coefficients": (I, [I, V, I, V, V, I]),
        "mf_dg_element_matrices": (I, [I, I, I, V, V]),
        "mf_dg_integrate": (I, [I]),
        "mf_dg_get_element_matrices": (I, [I, I, I, V, V]),
        "mf_dg_free": (I, [I]),
        "mf_dg_set_penalty": (I, [I, D]),
        "mf_dg_set_valu_blocks": (I, [I, I]),
        "mf_dg_set_boundary_values": (I, [I, V]),
        "mf_dg_get_faces": (I, [I, V, V, V]),
        "mf_dg_assemble": (I, [I]),
        "mf_dg_get_block_row": (I, [I, I, V, V]),
        "mf_dg_fill_precon": (I, [I]),
        "mf_dg_get_precon_block": (I, [I, I, V]),
        "mf_dg_set_gs": (I, [I, I, V]),
        "mf_dg_set_multigrid": (I, [I, I, I, I]),
        "mf_dg_multigrid_level": (I, [I, I]),
        "mf_dg_gmres": (I, [I, I, I, I, D, V, V, V, V]),
        "mf_dg_get_solution": (I, [I, V]),
        "mf_dg_bench": (I, [I, I, V]),
        "mf_dg_create_mixed": (I, [I, V]),
        "mf_dg_fill_ilu": (I, [I, V]),
        "mf_dg_get_ilu_row": (I, [I, I, V]),
        "mf_dg_ilu_apply": (I, [I, V, V]),
        "mf_dg_layout": (I, [I, V, V, V]),
        # storage seam with the reference's names (lsd_mkb/lah_intf.h)
        "lar_allocate_SM_and_LV": (I, [I, I, I, I, I, I, V, V, V, V]),
        "lar_initialize_SM_and_LV": (I, [I, I]),
        "lar_get_storage": (D, [I]),
        "lar_fill_assembly_table_int_ent": (I, [I, I, V, V, V]),
        "lar_assemble_SM_and_LV_with_table": (I, [I, I, I, V, V, V, V]),
        "lar_assemble_SM_and_LV": (I, [I, I, I, V, V, V, V, V]),
        "lar_free_SM_and_LV": (I, [I]),
        "lar_allocate_preconditioner": (I, [I, I, I]),
        "lar_fill_preconditioner": (I, [I]),
        "lar_free_preconditioner": (I, [I]),
        "lar_compute_residual": (None, [I, I, I, I, V, V, V]),
        "lar_perform_BJ_or_GS_iterations": (None, [I, I, I, I, I, V, V]),
        "lar_perform_rhsub": (None, [I, I, V, V]),
        "mf_lar_get_system": (I, [I, V, V, V, V]),
        "mf_lar_get_sizes": (I, [I, V, V, V, V]),
        "mf_lar_get_block_precon": (I, [I, V, V]),
        "mf_lar_get_ilu_pattern": (I, [I, V, V, V]),
        "mf_lar_gmres": (I, [I, I, I, D, I, I, V, V, V, V]),
        "mf_lar_bench": (I, [I, I, I, V]),
        "mf_lar_set_gs_order": (I, [I, I]),
        "mf_lar_set_gs_threads": (I, [I, I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def lib():
    """The loaded HIP library (raises MFError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MFError("libmodfem_hip.so is not built (run __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def check(rc):
    if rc < 0:
        raise MFError(lib().mf_last_error().decode())
    return rc


_inited = False


def init(device=0):
    global _inited
    if not _inited:
        check(lib().mf_init(device))
        _inited = True


def gmres_kernel_stats(enable):
    """returns accumulated {op:(launches, ms, bytes), update:(...)} and sets the
    timing switch for later solves"""
    out = np.zeros(6)
    check(lib().mf_gmres_kernel_stats(int(enable), _p(out)))
    return {"op_dots": tuple(out[0:3]), "update": tuple(out[3:6])}


MKB_KEYS = ("solver", "nr_level", "krylbas", "precon", "nr_prec", "block_type", "storage_type",
            "ilu_k", "max_iter", "conv_type", "monitor", "nr_pre_smooth", "nr_post_smooth")


def mkb_read(filename="", max_iter=-1, error_type=-1, error_tol=-1.0, monitoring=-1):
    """solver control file -> dict (keys MKB_KEYS + conv_meas)"""
    ip = np.zeros(13, np.int32)
    cm = ctypes.c_double()
    check(lib().mf_mkb_read(filename.encode(), max_iter, error_type, error_tol, monitoring,
                            _p(ip), ctypes.byref(cm)))
    d = dict(zip(MKB_KEYS, (int(v) for v in ip)))
    d["conv_meas"] = cm.value
    return d


def bench_copy(nbytes, reps=5):
    """ms per device-to-device copy of nbytes"""
    init()
    ms = ctypes.c_double()
    check(lib().mf_bench_copy(int(nbytes), reps, ctypes.byref(ms)))
    return ms.value


class Batch:
    """Batched GMRES over resident problems (continuous batching)."""

    def __init__(self, krylbas=30, own_stream=False):
        """own_stream: cycles run on a stream of the batch's own, so that
        several batches can run at once (run_async / wait)"""
        init()
        self.h = check(lib().mf_batch_create(krylbas))
        if own_stream:
            check(lib().mf_batch_set_stream(self.h))

    def add(self, problem, precon=1, max_iter=1000, tol=1e-8, inner_precision=64):
        """start a member; inner_precision=32 makes it a mixed-precision
        member (FP64 refinement around FP32 inner solves); precon 2 (ILU(0))
        needs problem.set_ilu_order(1).  A DgProblem joins preconditioned by
        block Jacobi (precon is ignored)."""
        if isinstance(problem, DgProblem):
            return check(lib().mf_batch_add_dg(self.h, problem.h, max_iter, tol, inner_precision))
        return check(lib().mf_batch_add_ex(self.h, problem.h, precon, max_iter, tol,
                                           inner_precision))

    def remove(self, problem):
        if isinstance(problem, DgProblem):
            check(lib().mf_batch_remove_dg(self.h, problem.h))
        else:
            check(lib().mf_batch_remove(self.h, problem.h))

    def integrate_assemble(self, precision=64):
        check(lib().mf_batch_integrate_assemble(self.h, precision))

    def run(self, max_cycles=1000000, use_graph=True):
        """run restart cycles; returns the number of members still iterating"""
        a = ctypes.c_int()
        check(lib().mf_batch_run(self.h, max_cycles, int(use_graph), ctypes.byref(a)))
        return a.value

    def run_async(self, use_graph=False):
        """enqueue one restart cycle; wait() completes it"""
        check(lib().mf_batch_run_async(self.h, int(use_graph)))

    def wait(self):
        """complete the cycle in flight; returns the members still iterating"""
        a = ctypes.c_int()
        check(lib().mf_batch_wait(self.h, ctypes.byref(a)))
        return a.value

    def status(self, problem):
        d, c, it, rr, nr = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double(),
                            ctypes.c_int())
        fn = lib().mf_batch_status_dg if isinstance(problem, DgProblem) else lib().mf_batch_status_ex
        check(fn(self.h, problem.h, ctypes.byref(d), ctypes.byref(c), ctypes.byref(it),
                 ctypes.byref(rr), ctypes.byref(nr)))
        return dict(done=bool(d.value), converged=bool(c.value), iters=it.value, res_rel=rr.value,
                    refinements=nr.value)

    def set_tiles(self, spb):
        """slices of 64 rows per block of every member; 0 lets the library
        tune it on the next run"""
        check(lib().mf_batch_set_tiles(self.h, spb))

    def tiles(self):
        """tile in use, whether it was tuned, and the trial cycle times (ms)
        of the candidates"""
        spb, tuned = ctypes.c_int(), ctypes.c_int()
        cand = np.zeros(16, dtype=np.int32)
        n = check(lib().mf_batch_tile_candidates(cand.ctypes.data))
        ms = np.zeros(n)
        check(lib().mf_batch_tiles(self.h, ctypes.byref(spb), ctypes.byref(tuned), ms.ctypes.data))
        return dict(spb=spb.value, tuned=bool(tuned.value),
                    trials=dict(zip(cand[:n].tolist(), ms.tolist())))

    @staticmethod
    def clear_tile_cache():
        check(lib().mf_batch_tile_cache_clear())

    def free(self):
        if self.h:
            lib().mf_batch_free(self.h)
            self.h = 0


class Mesh:
    """Mesh input (element/face lists as the approximation module sees them)."""

    def __init__(self, handle):
        self.h = handle
        nno, nel, nbf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().mf_mesh_sizes(handle, ctypes.byref(nno), ctypes.byref(nel), ctypes.byref(nbf)))
        self.nno, self.nel, self.nbf = nno.value, nel.value, nbf.value

    @classmethod
    def box_prism(cls, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, seed=0, jitter=0.0, bc=1):
        return cls(check(lib().mf_mesh_box_prism(nx, ny, nz, lx, ly, lz, seed, jitter, bc)))

    @classmethod
    def box_tetra(cls, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, seed=0, jitter=0.0, bc=1):
        return cls(check(lib().mf_mesh_box_tetra(nx, ny, nz, lx, ly, lz, seed, jitter, bc)))

    @classmethod
    def box_hybrid(cls, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, seed=0, jitter=0.0, bc=1, tet_frac=0.4):
        """z-layers of prisms and of tetrahedra (a layer is tetrahedral with
        probability tet_frac), conforming: for DG problems"""
        return cls(check(lib().mf_mesh_box_hybrid(nx, ny, nz, lx, ly, lz, seed, jitter, bc,
                                                  float(tet_frac))))

    @classmethod
    def read_jk(cls, path):
        return cls(check(lib().mf_mesh_read_jk(path.encode())))

    @classmethod
    def read_nas(cls, path):
        return cls(check(lib().mf_mesh_read_nas(str(path).encode())))

    @classmethod
    def read_dat(cls, path):
        """the reference's prism mesh dump (.dat / .dmp)"""
        return cls(check(lib().mf_mesh_read_dat(str(path).encode())))

    def ndof(self):
        n, nnz = ctypes.c_int(), ctypes.c_longlong()
        check(lib().mf_mesh_setup(self.h, ctypes.byref(n), ctypes.byref(nnz)))
        return n.value

    def write_field(self, x, path, accuracy=0.0):
        """dump x (DOF order) in the reference's field format"""
        x = np.ascontiguousarray(x, np.float64)
        check(lib().mf_field_write(self.h, _p(x), float(accuracy), str(path).encode()))

    def read_field(self, path):
        """the first solution vector of a field dump, in DOF order"""
        x = np.zeros(self.ndof())
        check(lib().mf_field_read(self.h, str(path).encode(), _p(x)))
        return x

    @classmethod
    def from_arrays(cls, coor, el_type, el_nodes, bf_elem, bf_side, bf_bc):
        coor = np.ascontiguousarray(coor, dtype=np.float64)
        el_type = np.ascontiguousarray(el_type, dtype=np.int32)
        el_nodes = np.ascontiguousarray(el_nodes, dtype=np.int32)
        bf = [np.ascontiguousarray(a, dtype=np.int32) for a in (bf_elem, bf_side, bf_bc)]
        return cls(check(lib().mf_mesh_create(coor.shape[0], _p(coor), el_type.shape[0],
                                              _p(el_type), _p(el_nodes), bf[0].shape[0],
                                              _p(bf[0]), _p(bf[1]), _p(bf[2]))))

    def write_jk(self, path):
        check(lib().mf_mesh_write_jk(self.h, path.encode()))

    def arrays(self):
        coor = np.zeros((self.nno, 3))
        et = np.zeros(self.nel, np.int32)
        en = np.zeros((self.nel, 6), np.int32)
        be = np.zeros(self.nbf, np.int32)
        bs = np.zeros(self.nbf, np.int32)
        bb = np.zeros(self.nbf, np.int32)
        check(lib().mf_mesh_get(self.h, _p(coor), _p(et), _p(en), _p(be), _p(bs), _p(bb)))
        return dict(coor=coor, el_type=et, el_nodes=en, bf_elem=be, bf_side=bs, bf_bc=bb)

    def setup(self):
        """host integer set-up: (dof_node, row_ptr, col_ind)"""
        n, nnz = ctypes.c_int(), ctypes.c_longlong()
        check(lib().mf_mesh_setup(self.h, ctypes.byref(n), ctypes.byref(nnz)))
        d = np.zeros(n.value, np.int32)
        rp = np.zeros(n.value + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        check(lib().mf_mesh_setup_get(self.h, _p(d), _p(rp), _p(ci)))
        return d, rp, ci

    def setup_order(self):
        """device-internal numbering after setup(): (row_perm, el_perm)"""
        n, nnz = ctypes.c_int(), ctypes.c_longlong()
        check(lib().mf_mesh_setup(self.h, ctypes.byref(n), ctypes.byref(nnz)))
        rp = np.zeros(n.value, np.int32)
        ep = np.zeros(self.nel, np.int32)
        check(lib().mf_mesh_setup_order(self.h, _p(rp), _p(ep)))
        return rp, ep

    def setup_table(self, int_ent):
        t = np.zeros(42, np.int32)
        k = check(lib().mf_mesh_setup_table(self.h, int_ent, _p(t)))
        return t[:k]

    def free(self):
        if self.h:
            lib().mf_mesh_free(self.h)
            self.h = 0


class DgProblem:
    """A DG prism field of degree pdeg = 100*pz + px on a prism mesh
    (apd_dg_prism), one degree for every element or one per element
    (101, 202 or 303 each); element DOF blocks in mesh element order at the
    offsets `xoff`."""

    def __init__(self, mesh, pdeg):
        init()
        if np.isscalar(pdeg):
            self.h = check(lib().mf_dg_create(mesh.h, int(pdeg)))
        else:
            pd = np.ascontiguousarray(pdeg, np.int32)
            self.h = check(lib().mf_dg_create_mixed(mesh.h, _p(pd)))
        self.pdeg = pdeg
        nel, ns, ng = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().mf_dg_sizes(self.h, ctypes.byref(nel), ctypes.byref(ns), ctypes.byref(ng)))
        self.nel, self.ns, self.ng = nel.value, ns.value, ng.value
        nd = ctypes.c_longlong()
        self.xoff = np.zeros(self.nel + 1, np.int64)
        self.soff = np.zeros(6 * self.nel + 1, np.int64)
        check(lib().mf_dg_layout(self.h, ctypes.byref(nd), _p(self.xoff), _p(self.soff)))
        self.ndof = nd.value

    def set_coefficients(self, ind, coef_model=COEF_UNIFORM, coef_const=None, coef_el=None,
                         src_model=SRC_CONST):
        ind = np.ascontiguousarray(ind, dtype=np.int32)
        cc = np.zeros(NCOEF) if coef_const is None else np.ascontiguousarray(coef_const, np.float64)
        ce = None if coef_el is None else np.ascontiguousarray(coef_el, np.float64)
        check(lib().mf_dg_set_coefficients(self.h, _p(ind), coef_model, _p(cc), _p(ce), src_model))

    def element_matrices(self, first=0, count=None):
        """(stiff[count, ns, ns], rhs[count, ns]); stiff[e, j, i]: trial j, test i"""
        count = self.nel - first if count is None else count
        st = np.zeros((count, self.ns, self.ns))
        rhs = np.zeros((count, self.ns))
        check(lib().mf_dg_element_matrices(self.h, first, count, _p(st), _p(rhs)))
        return st, rhs

    def integrate(self):
        check(lib().mf_dg_integrate(self.h))

    def get_element_matrices(self, first=0, count=None):
        count = self.nel - first if count is None else count
        st = np.zeros((count, self.ns, self.ns))
        rhs = np.zeros((count, self.ns))
        check(lib().mf_dg_get_element_matrices(self.h, first, count, _p(st), _p(rhs)))
        return st, rhs

    def set_penalty(self, penalty_const):
        check(lib().mf_dg_set_penalty(self.h, float(penalty_const)))

    def set_valu_blocks(self, max_ns):
        """blocks of at most max_ns shape functions assembled on the vector
        ALUs, larger ones on MFMA (0: all on MFMA)"""
        check(lib().mf_dg_set_valu_blocks(self.h, int(max_ns)))

    def set_boundary_values(self, val=None):
        """per (element, side) boundary data used when no exact solution is selected"""
        v = None if val is None else np.ascontiguousarray(val, np.float64).reshape(-1)
        if v is not None and v.shape[0] != 5 * self.nel:
            raise ValueError("need 5 values per element")
        check(lib().mf_dg_set_boundary_values(self.h, _p(v)))

    def faces(self):
        """(neighbour or -1 Dirichlet / -2 Neumann, orientation code, face size) per (element, side)"""
        nb = np.zeros((self.nel, 5), np.int32)
        code = np.zeros((self.nel, 5), np.int32)
        hf = np.zeros((self.nel, 5))
        check(lib().mf_dg_get_faces(self.h, _p(nb), _p(code), _p(hf)))
        return nb, code, hf

    def assemble(self):
        check(lib().mf_dg_assemble(self.h))

    def block_row(self, e):
        """blocks[6, ns, ns] with blocks[slot, j, i] = row i, column j of the
        block (slot 0 own, 1 + s the neighbour across side s), rhs[ns]; for a
        mixed problem a list of six column-major blocks (ns_e x ns_e, then
        ns_e x ns_nb across an interior side, ns_e x ns_e otherwise) as
        arrays [cols, rows]"""
        e = int(e)
        ns = int(self.xoff[e + 1] - self.xoff[e])
        if self.ns:
            blk = np.zeros((6, ns, ns))
            rhs = np.zeros(ns)
            check(lib().mf_dg_get_block_row(self.h, e, _p(blk), _p(rhs)))
            return blk, rhs
        if not hasattr(self, "_fnb"):
            self._fnb = self.faces()[0]
        cols = [ns] + [int(self.xoff[n + 1] - self.xoff[n]) if n >= 0 else ns for n in self._fnb[e]]
        flat = np.zeros(ns * sum(cols))
        rhs = np.zeros(ns)
        check(lib().mf_dg_get_block_row(self.h, e, _p(flat), _p(rhs)))
        cuts = np.r_[0, np.cumsum(cols)] * ns
        return [flat[cuts[k]:cuts[k + 1]].reshape(cols[k], ns) for k in range(6)], rhs

    def fill_precon(self):
        check(lib().mf_dg_fill_precon(self.h))

    def precon_block(self, e):
        """inverse of element e's own block (inv[j, i]: row i, column j)"""
        ns = int(self.xoff[int(e) + 1] - self.xoff[int(e)])
        inv = np.zeros((ns, ns))
        check(lib().mf_dg_get_precon_block(self.h, int(e), _p(inv)))
        return inv

    def set_multigrid(self, nu1=2, nu2=2, coarse=2):
        """p-multigrid V-cycle levels for gmres(precon=4); returns the level count"""
        return check(lib().mf_dg_set_multigrid(self.h, int(nu1), int(nu2), int(coarse)))

    def multigrid_level_row(self, level, e):
        """(blocks[6, ns, ns], rhs) of element e on p-multigrid level `level`
        (0: degree 1), its Galerkin blocks current"""
        h = check(lib().mf_dg_multigrid_level(self.h, int(level)))
        ns = (level + 2) * (level + 3) // 2 * (level + 2)
        blk, rhs = np.zeros((6, ns, ns)), np.zeros(ns)
        check(lib().mf_dg_get_block_row(h, int(e), _p(blk), _p(rhs)))
        return blk, rhs

    def fill_ilu(self):
        """block ILU(0) factors on the device; returns the forward and
        backward level counts of its schedule"""
        lev = np.zeros(2, np.int32)
        check(lib().mf_dg_fill_ilu(self.h, _p(lev)))
        return int(lev[0]), int(lev[1])

    def ilu_row(self, e):
        """factored block row of element e in the layout of block_row (slot
        0: inverse of U's diagonal block)"""
        e = int(e)
        ns = int(self.xoff[e + 1] - self.xoff[e])
        if self.ns:
            blk = np.zeros((6, ns, ns))
            check(lib().mf_dg_get_ilu_row(self.h, e, _p(blk)))
            return blk
        if not hasattr(self, "_fnb"):
            self._fnb = self.faces()[0]
        cols = [ns] + [int(self.xoff[n + 1] - self.xoff[n]) if n >= 0 else ns for n in self._fnb[e]]
        flat = np.zeros(ns * sum(cols))
        check(lib().mf_dg_get_ilu_row(self.h, e, _p(flat)))
        cuts = np.r_[0, np.cumsum(cols)] * ns
        return [flat[cuts[k]:cuts[k + 1]].reshape(cols[k], ns) for k in range(6)]

    def ilu_apply(self, b):
        """(LU)^-1 b with the device factors"""
        b = np.ascontiguousarray(b, dtype=np.float64)
        v = np.zeros(self.ndof)
        check(lib().mf_dg_ilu_apply(self.h, _p(b), _p(v)))
        return v

    def set_gs(self, nsweeps):
        """block Gauss-Seidel sweeps for precon 2; returns the number of colours"""
        nc = ctypes.c_int()
        check(lib().mf_dg_set_gs(self.h, int(nsweeps), ctypes.byref(nc)))
        return nc.value

    def gmres(self, precon=1, krylbas=30, max_iter=1000, tol=1e-8, fetch=True):
        x = np.zeros(self.ndof) if fetch else None
        it, rr, rs = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        conv = check(lib().mf_dg_gmres(self.h, precon, krylbas, max_iter, tol, _p(x),
                                       ctypes.byref(it), ctypes.byref(rr), ctypes.byref(rs)))
        return x, {"converged": bool(conv), "iters": it.value, "res_rel": rr.value,
                   "resid": rs.value}

    def solution(self):
        x = np.zeros(self.ndof)
        check(lib().mf_dg_get_solution(self.h, _p(x)))
        return x

    def bench(self, reps=3):
        """(ms per assembly, ms per preconditioned operator application)"""
        out = np.zeros(2)
        check(lib().mf_dg_bench(self.h, reps, _p(out)))
        return float(out[0]), float(out[1])

    def free(self):
        if self.h:
            lib().mf_dg_free(self.h)
            self.h = 0


class Problem:
    """One std-linear problem with its device-resident CRS system."""

    def __init__(self, handle):
        self.h = handle
        n, nnz, nel, nbf, mr = (ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int(),
                                ctypes.c_int(), ctypes.c_int())
        check(lib().mf_problem_sizes(handle, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(nel),
                                     ctypes.byref(nbf), ctypes.byref(mr)))
        self.n, self.nnz, self.nel, self.nbf, self.max_row = (n.value, nnz.value, nel.value,
                                                               nbf.value, mr.value)

    @classmethod
    def from_mesh(cls, mesh, device_setup=False):
        """device_setup: build the DOF numbering, pattern and tables on the
        GPU (bitwise the host set-up's)"""
        init()
        if device_setup:
            return cls(check(lib().mf_problem_create_device(mesh.h)))
        return cls(check(lib().mf_problem_create(mesh.h)))

    @classmethod
    def from_crs(cls, row_ptr, col_ind, val, rhs=None):
        init()
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col_ind, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=np.float64)
        b = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64)
        return cls(check(lib().mf_problem_create_crs(rp.shape[0] - 1, _p(rp), _p(ci), _p(v), _p(b))))

    def dofs(self):
        d = np.zeros(self.n, np.int32)
        check(lib().mf_problem_get_dofs(self.h, _p(d)))
        return d

    def pattern(self):
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        check(lib().mf_problem_get_pattern(self.h, _p(rp), _p(ci)))
        return rp, ci

    def table(self, int_ent):
        t = np.zeros(42, np.int32)
        k = check(lib().mf_problem_get_table(self.h, int_ent, _p(t)))
        return t[:k]

    def set_coefficients(self, ind, coef_model=COEF_UNIFORM, coef_const=None, coef_el=None,
                         src_model=SRC_CONST, bf_g=None):
        ind = np.ascontiguousarray(ind, dtype=np.int32)
        cc = np.zeros(NCOEF) if coef_const is None else np.ascontiguousarray(coef_const, np.float64)
        ce = None if coef_el is None else np.ascontiguousarray(coef_el, np.float64)
        g = None if bf_g is None else np.ascontiguousarray(bf_g, np.float64)
        check(lib().mf_problem_set_coefficients(self.h, _p(ind), coef_model, _p(cc), _p(ce),
                                                src_model, _p(g)))

    def set_assembly_mode(self, mode):
        check(lib().mf_set_assembly_mode(self.h, mode))

    def integrate_assemble(self, precision=64, wait=True):
        """wait=False only enqueues the work on the library stream (later
        library calls see the new system)"""
        if wait:
            check(lib().mf_integrate_assemble(self.h, precision))
        else:
            check(lib().mf_integrate_assemble_async(self.h, precision))

    def element_matrices(self, first=0, count=None, precision=64):
        count = self.nel - first if count is None else count
        out = np.zeros((count, 42))
        check(lib().mf_element_matrices(self.h, first, count, precision, _p(out)))
        return out

    def face_matrices(self, first=0, count=None):
        count = self.nbf - first if count is None else count
        out = np.zeros((count, 42))
        check(lib().mf_face_matrices(self.h, first, count, _p(out)))
        return out

    def system(self):
        val = np.zeros(self.nnz)
        rhs = np.zeros(self.n)
        check(lib().mf_get_system(self.h, _p(val), _p(rhs)))
        return val, rhs

    def set_system(self, val=None, rhs=None):
        v = None if val is None else np.ascontiguousarray(val, np.float64)
        b = None if rhs is None else np.ascontiguousarray(rhs, np.float64)
        check(lib().mf_set_system(self.h, _p(v), _p(b)))

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.n)
        check(lib().mf_spmv(self.h, _p(x), _p(y)))
        return y

    def gmres(self, precon=1, krylbas=30, max_iter=1000, tol=1e-8, x0=None, fetch=True):
        """fetch=False keeps the solution on the device (x is None; see
        solution()) and starts from zero"""
        if fetch:
            x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
            xp = _p(x)
        else:
            x, xp = None, None
        it, rr, rs = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        conv = check(lib().mf_gmres(self.h, precon, krylbas, max_iter, tol, int(x0 is None),
                                    xp, ctypes.byref(it), ctypes.byref(rr), ctypes.byref(rs)))
        return x, dict(converged=bool(conv), iters=it.value, res_rel=rr.value, resid=rs.value)

    def gmres_mixed(self, precon=1, krylbas=30, max_iter=1000, tol=1e-8, fetch=True):
        """FP64 refinement around GMRES with an FP32-valued operator"""
        x = np.zeros(self.n) if fetch else None
        it, rr, rs, nref = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        conv = check(lib().mf_gmres_mixed(self.h, precon, krylbas, max_iter, tol,
                                          _p(x) if fetch else None, ctypes.byref(it),
                                          ctypes.byref(rr), ctypes.byref(rs), ctypes.byref(nref)))
        return x, dict(converged=bool(conv), iters=it.value, res_rel=rr.value, resid=rs.value,
                       refinements=nref.value)

    def solve_mkb(self, params, fetch=True):
        """GMRES as the solver control parameters (mkb_read) say"""
        ip = np.array([params[k] for k in MKB_KEYS], np.int32)
        x = np.zeros(self.n) if fetch else None
        it, rr, rs = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        conv = check(lib().mf_solve_mkb(self.h, _p(ip), params["conv_meas"], 1,
                                        _p(x) if fetch else None, ctypes.byref(it),
                                        ctypes.byref(rr), ctypes.byref(rs)))
        return x, dict(converged=bool(conv), iters=it.value, res_rel=rr.value, resid=rs.value)

    def set_face_bc(self, bc_type, g):
        """per-face type (1 Dirichlet penalty, 2 Neumann flux) and data"""
        t = np.ascontiguousarray(bc_type, np.int32)
        gv = np.ascontiguousarray(g, np.float64)
        check(lib().mf_problem_set_face_bc(self.h, _p(t), _p(gv)))

    def set_gs(self, nsweeps):
        check(lib().mf_set_gs(self.h, nsweeps))

    def gs_order(self):
        """(position of every reference row in the GS order, colours)"""
        pos = np.zeros(self.n, np.int32)
        nc = ctypes.c_int()
        check(lib().mf_gs_get(self.h, _p(pos), ctypes.byref(nc)))
        return pos, nc.value

    def set_ilu_order(self, order):
        check(lib().mf_set_ilu_order(self.h, order))

    def ilu(self):
        """(factors in reference CRS positions, elimination position per row,
        (forward levels, backward levels, colours))"""
        lu = np.zeros(self.nnz)
        pos = np.zeros(self.n, np.int32)
        lev = np.zeros(3, np.int32)
        check(lib().mf_ilu_get(self.h, _p(lu), _p(pos), _p(lev)))
        return lu, pos, tuple(int(v) for v in lev)

    def solution(self):
        """host copy of the last GMRES solution"""
        x = np.zeros(self.n)
        check(lib().mf_get_solution(self.h, _p(x)))
        return x

    def true_residual(self, x):
        x = np.ascontiguousarray(x, np.float64)
        r = ctypes.c_double()
        check(lib().mf_true_residual(self.h, _p(x), ctypes.byref(r)))
        return r.value

    def bench_integrate(self, precision=64, reps=5):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().mf_bench_integrate(self.h, precision, reps, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def bench_spmv(self, reps=20, prec=None, ch=12):
        """ms per operator application; prec 0/1/2 selects FP64, FP32 values,
        FP32 values and vectors (None: the solver's FP64 operator)"""
        a = ctypes.c_double()
        if prec is None:
            check(lib().mf_bench_spmv(self.h, reps, ctypes.byref(a)))
        else:
            check(lib().mf_bench_spmv_ex(self.h, prec, ch, reps, ctypes.byref(a)))
        return a.value

    def last_solve_ms(self):
        a = ctypes.c_double()
        check(lib().mf_last_solve_ms(self.h, ctypes.byref(a)))
        return a.value

    def free(self):
        if self.h:
            lib().mf_problem_free(self.h)
            self.h = 0
