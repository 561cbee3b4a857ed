"""Benchmark of the integrate -> assemble -> GMRES hot path on one MI355X.

Workload (default): std-linear prism convection-diffusion on a 128^3-hex box
split into 2 prisms per hex (4,194,304 prisms, 2,146,689 DOFs), interior
nodes jittered by +-0.2h, per-element lognormal diffusion (mu=0, sigma=1) and
uniform random velocity in [-1,1]^3, unit source, Dirichlet penalty on all
boundary faces.  One step = FP64 integrate + assemble into CRS, Jacobi fill,
and GMRES(50) with Jacobi to a 1e-8 relative preconditioned residual (the
reference's stopping test).  Inputs are resident in device memory before the
timed region.

Prints one JSON line (see the keys below).  --gpus N with N > 1 exits: the
build uses one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s
FP64_PEAK_TFS = 78.6       # MI355X dense FP64, TFLOP/s
DG_VALU_NS = 6             # the library's default for mf_dg_set_valu_blocks
ASSEMBLY_MODE = 0          # row-wise deterministic assembly (fastest measured)
COPY_BYTES = 1 << 30        # stream-copy yardstick (read + write of 1 GiB)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_op_dots_sell_box128_pmc.txt")
METRIC = "elements integrated+assembled/s; solves/s & p99 latency at 1e-8 rel residual"


def make_workload(n, seed):
    import modfem_amd as mf
    mesh = mf.Mesh.box_prism(n, n, n, seed=seed, jitter=0.2, bc=1)
    rng = np.random.default_rng(seed + 1)
    kappa = np.exp(rng.normal(0.0, 1.0, mesh.nel))
    b = rng.uniform(-1.0, 1.0, (mesh.nel, 3))
    cc = np.zeros(mf.NCOEF)
    cc[mf.C_S] = 1.0
    coef_el = np.concatenate([kappa, b[:, 0], b[:, 1], b[:, 2]])
    return mesh, kappa, b, cc, coef_el


def _median_time(fn, reps, warmup=1):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def _ref_crs_residual_GBs(op, val, rhs, reps):
    """the reference's own lar_compute_residual_crs and one block-Jacobi sweep
    (lar_perform_BJ_or_GS_iterations_crs), compiled from las_crs_intf.c into
    oracle/_ref, on the oracle's assembled system; algorithmic bytes of the
    residual over its median time (None when oracle/_ref is absent)"""
    import ctypes
    from oracle_lib import ref, ref_available
    if not ref_available("libref_crs.so"):
        return None
    L = ref("libref_crs.so")
    I = ctypes.c_int
    P = lambda v: v.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    lists = []
    for r in range(op.n):
        cols = op.col_ind[op.row_ptr[r]: op.row_ptr[r + 1]]
        lists.append(np.ascontiguousarray(cols[cols != r], np.int32))
    nro = np.array([len(x) for x in lists], np.int32)
    ptrs = (ctypes.POINTER(I) * op.n)(*[x.ctypes.data_as(ctypes.POINTER(I)) for x in lists])
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)  # the reference prints progress on stdout
    try:
        mid = L.lar_allocate_SM_and_LV_crs(op.n, 36, P(nro), ptrs)
        rp, ci, va, rh = (ctypes.POINTER(I)(), ctypes.POINTER(I)(), ctypes.POINTER(ctypes.c_double)(),
                          ctypes.POINTER(ctypes.c_double)())
        L.lar_get_SM_and_LV_crs_from_crs.argtypes = [I, I] + [ctypes.c_void_p] * 4
        L.lar_get_SM_and_LV_crs_from_crs(mid, 0, ctypes.byref(rp), ctypes.byref(ci),
                                         ctypes.byref(va), ctypes.byref(rh))
        np.ctypeslib.as_array(va, (op.nnz,))[:] = val
        np.ctypeslib.as_array(rh, (op.n,))[:] = rhs
        x = np.random.default_rng(1).uniform(-1, 1, op.n)
        v = np.zeros(op.n)
        L.lar_compute_residual_crs.restype = None
        L.lar_compute_residual_crs.argtypes = [I, I, I, I] + [ctypes.c_void_p] * 3
        t_res = _median_time(lambda: L.lar_compute_residual_crs(mid, 1, 0, op.n, P(x), None, P(v)),
                             reps)
        L.lar_allocate_preconditioner_crs(mid, 1, 0)
        L.lar_fill_preconditioner_crs(mid)
        L.lar_perform_BJ_or_GS_iterations_crs.restype = None
        L.lar_perform_BJ_or_GS_iterations_crs.argtypes = [I, I, I, I, I, ctypes.c_void_p,
                                                          ctypes.c_void_p]
        t_bj = _median_time(lambda: L.lar_perform_BJ_or_GS_iterations_crs(mid, 1, 0, 1, op.n, P(v),
                                                                          None), reps)
        L.lar_free_SM_and_LV_crs(mid)
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
    b_res = 12.0 * op.nnz + 4.0 * (op.n + 1) + 16.0 * op.n
    return {"residual_GBs": b_res / t_res / 1e9, "residual_ms": t_res * 1e3,
            "jacobi_sweep_ms": t_bj * 1e3}


def cpu_baseline(mesh, kappa, b, gpu_iters, krylbas, sample_iters, seed):
    """the host-side baseline of the same workload: the oracle (the C
    restatement of the reference path, OpenMP) at the job's host threads and
    at one core, phases timed separately -- set-up (DOF list, pattern, tables),
    integrate+assemble, Jacobi fill, GMRES per iteration -- with one warm-up
    and the median of 5 (3 at one core); a step is integrate+assemble, fill
    and the GPU's iteration count of GMRES iterations.  The one-core
    integrate+assemble and the reference's own compiled CRS residual and
    Jacobi sweep are timed on a 32^3 box of the same generator"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import modfem_amd as mf
    from oracle_lib import OracleProblem, oracle, P
    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc

    def coef_of(kap, bb):
        c = np.zeros((len(kap), 21))
        c[:, 0] = c[:, 4] = c[:, 8] = kap
        c[:, 9:12] = bb
        c[:, 19] = 1.0
        return c

    arr = mesh.arrays()
    t0 = time.perf_counter()
    op = OracleProblem(arr)
    t_setup = time.perf_counter() - t0
    coef = coef_of(kappa, b)
    res = {"threads": threads, "nproc": nproc, "setup_s": t_setup}
    o = oracle()
    for nt, reps in ((threads, 5), (1, 3)):
        o.orc_set_threads(nt)
        key = "threads_%d" % nt
        rec = {}
        if nt == threads:
            rec["integrate_assemble_s"] = _median_time(lambda: op.assemble(mf.IND_CONV_DIFF, coef), reps)
            val, rhs = op.val.copy(), op.rhs.copy()
        else:
            # one core: a 32^3 box of the same generator, per element
            sm, kap_s, b_s, _, _ = make_workload(32, seed)
            ops = OracleProblem(sm.arrays())
            cs = coef_of(kap_s, b_s)
            t = _median_time(lambda: ops.assemble(mf.IND_CONV_DIFF, cs), reps)
            rec["integrate_assemble_s"] = t * mesh.nel / sm.nel
            rec["integrate_assemble_sample"] = "32^3 box (%d prisms), scaled per element" % sm.nel
        dinv, dptr = np.zeros(op.n), np.zeros(op.n, np.int32)
        rec["jacobi_fill_ms"] = _median_time(
            lambda: o.orc_fill_jacobi(op.n, P(op.row_ptr), P(op.col_ind), P(val), P(dptr), P(dinv)),
            reps) * 1e3
        its = sample_iters if nt == threads else max(3, sample_iters // 10)
        t0 = time.perf_counter()
        _, info = op.gmres(val, rhs, precon=1, krylbas=krylbas, max_iter=its, tol=1e-8)
        rec["gmres_ms_per_iter"] = (time.perf_counter() - t0) / max(info["iters"], 1) * 1e3
        rec["gmres_sample_iters"] = info["iters"]
        rec["step_s"] = (rec["integrate_assemble_s"] + rec["jacobi_fill_ms"] * 1e-3 +
                         gpu_iters * rec["gmres_ms_per_iter"] * 1e-3)
        rec["elements_per_s"] = mesh.nel / rec["step_s"]
        res[key] = rec
    o.orc_set_threads(threads)
    # a 64^3 box (275k rows, 5.7 M non-zeros: larger than the host caches)
    sm, kap_s, b_s, _, _ = make_workload(64, seed)
    ops = OracleProblem(sm.arrays())
    v_s, r_s = ops.assemble(mf.IND_CONV_DIFF, coef_of(kap_s, b_s))
    res["reference_crs_64cubed"] = _ref_crs_residual_GBs(ops, v_s.copy(), r_s.copy(), 5)
    full = res["threads_%d" % threads]
    return {
        "value": full["elements_per_s"], "unit": "elements/s", "cores": threads, "kind": "port",
        "sample": ("oracle (C restatement, OpenMP, %d threads of %d host cores) on the same mesh: "
                   "set-up %.2f s; integrate+assemble %.2f s, Jacobi fill %.1f ms, GMRES(%d) "
                   "%.1f ms/it over %d iterations (medians of 5 after a warm-up), extrapolated "
                   "to the GPU's %d iterations; one-core figures and the reference's compiled "
                   "CRS residual beside it" % (threads, nproc, t_setup, full["integrate_assemble_s"],
                                               full["jacobi_fill_ms"], krylbas,
                                               full["gmres_ms_per_iter"], full["gmres_sample_iters"],
                                               gpu_iters)),
        "phases": res,
    }


def measured_traffic(n, krylbas):
    """HBM bytes per launch of the operator+dots kernel from the committed
    FETCH_SIZE / WRITE_SIZE counter summary of this workload (None if absent)"""
    if n != 128 or krylbas != 50 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


# ---------------------------------------------------------------------------
# many-problem workloads: a batch of independent problems (one pipeline versus
# a loop of plain solves) and a served request stream (continuous batching)
# ---------------------------------------------------------------------------
def make_box_problem(mf, n_el, rng, seed, sigma=0.5):
    """std-linear prism Poisson on a box of about n_el prisms with random
    aspect ratios in [1,4], kappa lognormal(0, sigma) and source U(-1,1) per
    element, Dirichlet penalty on the boundary; two-pass deterministic
    assembly (element rows, then row gathers in element order: the fastest
    measured below a few million elements)"""
    asp = rng.uniform(1.0, 4.0, 3)
    sc = (n_el / 2.0 / asp.prod()) ** (1.0 / 3.0)
    nx, ny, nz = (max(1, int(round(a * sc))) for a in asp)
    mesh = mf.Mesh.box_prism(nx, ny, nz, lx=asp[0], ly=asp[1], lz=asp[2], seed=seed, jitter=0.2)
    p = mf.Problem.from_mesh(mesh)
    ce = np.concatenate([np.exp(rng.normal(0.0, sigma, mesh.nel)), rng.uniform(-1.0, 1.0, mesh.nel)])
    p.set_coefficients(mf.IND_DIFF, mf.COEF_KAPPA_S, coef_el=ce)
    p.set_assembly_mode(2)
    mesh.free()
    return p


def fp64_measure(p, precon=1):
    """||D^-1 (b - A x)|| / ||D^-1 b|| of a problem's solution, in FP64 on the
    host with the matrix read back from the device"""
    import scipy.sparse as sp
    val, rhs = p.system()
    rp, ci = p.pattern()
    A = sp.csr_matrix((val, ci, rp), shape=(p.n, p.n))
    d = 1.0 / A.diagonal() if precon > 0 else np.ones(p.n)
    x = p.solution()
    return float(np.linalg.norm(d * (rhs - A @ x)) / np.linalg.norm(d * rhs))


def bench_batch(args):
    """independent std-linear Poisson problems, element counts log-uniform in
    [el_min, el_max]: integrate+assemble+GMRES(30)+Jacobi to 1e-8 for all of
    them as one batch (device restart logic, hipGraph cycles) versus a loop of
    plain solves of the same problems; the batch in FP64 and with
    mixed-precision members (FP32 inner solves, FP64 refinement to the same
    1e-8 Jacobi-scaled FP64 measure, checked on the host)"""
    import modfem_amd as mf
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    probs = [make_box_problem(mf, float(np.exp(rng.uniform(np.log(args.el_min), np.log(args.el_max)))),
                              rng, args.seed * 1000 + k) for k in range(args.count)]
    t_setup = time.perf_counter() - t0
    m, tol, mi = 30, 1e-8, 20000
    nel = sum(p.nel for p in probs)
    ndof = sum(p.n for p in probs)
    # warm-up of both paths on the first problems
    for p in probs[:4]:
        p.integrate_assemble(64)
        p.gmres(precon=1, krylbas=m, max_iter=mi, tol=tol, fetch=False)
    wb = mf.Batch(m)
    for p in probs[:4]:
        wb.add(p, 1, mi, tol)
    wb.integrate_assemble(64)
    wb.run(use_graph=True)
    wb.free()
    # loop of plain solves
    t0 = time.perf_counter()
    loop_iters = []
    for p in probs:
        p.integrate_assemble(64)
        _, info = p.gmres(precon=1, krylbas=m, max_iter=mi, tol=tol, fetch=False)
        loop_iters.append(info["iters"])
    t_loop = time.perf_counter() - t0
    x_loop = [p.solution() for p in probs]
    # the tile tuner's trial cycles for this batch shape (cached for the
    # process; trial cycles are undone, so the solves below start afresh)
    b = mf.Batch(m)
    for p in probs:
        b.add(p, 1, mi, tol)
    t0 = time.perf_counter()
    b.run(max_cycles=1, use_graph=True)
    t_tune = time.perf_counter() - t0
    b.free()
    # one batch
    b = mf.Batch(m)
    for p in probs:
        b.add(p, 1, mi, tol)
    t0 = time.perf_counter()
    b.integrate_assemble(64)
    b.run(use_graph=True)
    t_batch = time.perf_counter() - t0
    st = [b.status(p) for p in probs]
    tiles = b.tiles()
    diff = max(float(np.linalg.norm(p.solution() - x) / np.linalg.norm(x)) for p, x in zip(probs, x_loop))
    b.free()
    # the same batch with the former fixed tile (8 slices per block)
    b = mf.Batch(m)
    b.set_tiles(8)
    for p in probs:
        b.add(p, 1, mi, tol)
    t0 = time.perf_counter()
    b.integrate_assemble(64)
    b.run(use_graph=True)
    t_fixed8 = time.perf_counter() - t0
    b.free()
    # mixed-precision members (the tile cache already holds this shape)
    b = mf.Batch(m)
    for p in probs:
        b.add(p, 1, mi, tol, inner_precision=32)
    t0 = time.perf_counter()
    b.integrate_assemble(64)
    b.run(use_graph=True)
    t_mixed = time.perf_counter() - t0
    stm = [b.status(p) for p in probs]
    b.free()
    meas = max(fp64_measure(p) for p in probs)
    out = {
        "metric": "solves/s of a batch of independent problems at 1e-8 rel residual",
        "value": len(probs) / t_mixed, "unit": "solves/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32 inner solves, f64 refinement", "data": "synthetic",
        "config": {"workload": "%d std-linear prism Poisson problems, %d-%d prisms log-uniform, "
                               "aspect 1-4, kappa lognormal(0,0.5), source U(-1,1), "
                               "GMRES(30)+Jacobi 1e-8" % (args.count, args.el_min, args.el_max),
                   "elements": nel, "dofs": ndof},
        "mixed_batch_s": t_mixed, "speedup_vs_loop": t_loop / t_mixed,
        "mixed_all_converged": all(x["converged"] for x in stm),
        "mixed_max_fp64_measure": meas,
        "mixed_iters_max": max(x["iters"] for x in stm),
        "mixed_refinements_max": max(x["refinements"] for x in stm),
        "fp64_batch_s": t_batch, "fp64_solves_per_s": len(probs) / t_batch,
        "fp64_speedup_vs_loop": t_loop / t_batch,
        "loop_s": t_loop, "loop_solves_per_s": len(probs) / t_loop,
        "elements_per_s": nel / t_mixed,
        "all_converged": all(x["converged"] for x in st),
        "iters": {"min": min(x["iters"] for x in st), "median": float(np.median([x["iters"] for x in st])),
                  "max": max(x["iters"] for x in st)},
        "max_iter_diff_vs_loop": max(abs(x["iters"] - i) for x, i in zip(st, loop_iters)),
        "max_rel_diff_vs_loop_solution": diff, "setup_s": t_setup,
        "tiles": {"slices_per_block": tiles["spb"], "tuned": tiles["tuned"],
                  "trial_cycle_ms": tiles["trials"], "tuning_s": t_tune,
                  "batch_s_with_fixed_tile_8": t_fixed8},
    }
    for p in probs:
        p.free()
    return out


def make_dg_request(mf, n_el, pdeg, rng, seed):
    """DG prism Poisson of degree pdeg on a box of about n_el prisms: random
    aspect ratios in [1,4], kappa lognormal(0, 1) and source U(-1,1) per
    element, interior-penalty flux, Dirichlet boundary"""
    asp = rng.uniform(1.0, 4.0, 3)
    sc = (n_el / 2.0 / asp.prod()) ** (1.0 / 3.0)
    nx, ny, nz = (max(1, int(round(a * sc))) for a in asp)
    mesh = mf.Mesh.box_prism(nx, ny, nz, lx=asp[0], ly=asp[1], lz=asp[2], seed=seed, jitter=0.2)
    dg = mf.DgProblem(mesh, pdeg)
    ce = np.concatenate([np.exp(rng.normal(0.0, 1.0, mesh.nel)), rng.uniform(-1.0, 1.0, mesh.nel)])
    dg.set_coefficients(mf.IND_DIFF, mf.COEF_KAPPA_S, coef_el=ce)
    mesh.free()
    return dg


def is_dg(p):
    return type(p).__name__ == "DgProblem"


def footprint(p, m, mixed=False):
    """device bytes a resident problem holds: SELL values+columns, element
    tables, vectors and the Krylov basis (and the FP32 values, inverses and
    refinement vectors of a mixed-precision member)"""
    if is_dg(p):  # blocks, their inverses, the SELL form (+ FP32 copy), vectors
        nd, nn = p.nel * p.ns, p.ns * p.ns
        sell = 6.0 * nn * p.nel
        return (8.0 * 7 * nn * p.nel + (14.0 if not mixed else 20.0) * sell
                + 8.0 * (m + 10) * nd)
    extra = 4.0 * p.nnz * 1.04 + 20.0 * p.n if mixed else 0.0
    return 12.0 * p.nnz * 1.04 + 8.0 * (m + 8) * p.n + 120.0 * p.nel + extra


def bench_serve(args):
    """Poisson-arrival stream of std-linear and DG p=1/2 Poisson requests
    (prisms, log-uniform sizes): admission by device-memory footprint and
    slot count, continuous batching between restart cycles, latency p50/p99
    and solves/s; versus serving the same stream one plain FP64 solve at a
    time (FIFO, with each request's measured solve time: std GMRES(30) with
    Jacobi, DG GMRES(30) with block Jacobi).  The server runs FP64 members
    and mixed-precision members (FP32 inner solves, FP64 refinement to the
    same FP64 measure; the std solutions checked on the host after the run)"""
    import modfem_amd as mf
    rng = np.random.default_rng(args.seed)
    drng = np.random.default_rng(args.seed + 7)
    m, tol, mi = 30, 1e-8, 20000
    probs = []
    for k in range(args.count):
        if drng.random() < args.dg_frac:
            n_el = float(np.exp(drng.uniform(np.log(args.dg_el_min), np.log(args.dg_el_max))))
            probs.append(make_dg_request(mf, n_el, 101 if k % 2 else 202, drng, args.seed * 1000 + k))
        else:
            probs.append(make_box_problem(mf, float(np.exp(rng.uniform(np.log(args.el_min),
                                                                       np.log(args.el_max)))),
                                          rng, args.seed * 1000 + k, sigma=1.0))

    def assemble(p, wait=True):
        if is_dg(p):
            p.assemble()
        else:
            p.integrate_assemble(64, wait=wait)

    def plain_solve(p):
        assemble(p)
        p.gmres(precon=1, krylbas=m, max_iter=mi, tol=tol, fetch=False)
    # sequential plain solves: service time of every request alone
    for p in probs[:3] + [q for q in probs if is_dg(q)][:2]:
        plain_solve(p)
    service = []
    for p in probs:
        t0 = time.perf_counter()
        plain_solve(p)
        service.append(time.perf_counter() - t0)
    results = []
    budget = args.hbm_gb * 1e9
    if max(footprint(p, m, True) for p in probs) > budget:
        raise SystemExit("a request is larger than the admission budget")
    def serve(arr, mixed):
        """continuous batching server over arrival times arr: latencies and span"""
        prec = 32 if mixed else 64
        b = mf.Batch(m)
        nxt, live, used, lat = 0, {}, 0.0, []
        t0 = time.perf_counter()
        while len(lat) < len(probs):
            now = time.perf_counter() - t0
            while (nxt < len(probs) and arr[nxt] <= now and len(live) < args.slots and
                   used + footprint(probs[nxt], m, mixed) <= budget):
                p = probs[nxt]
                assemble(p, wait=False)  # ordered before the next cycle
                b.add(p, 1, mi, tol, inner_precision=prec)
                live[nxt] = p
                used += footprint(p, m, mixed)
                nxt += 1
            if not live:
                time.sleep(max(0.0, min(arr[nxt] - now, 0.01)))
                continue
            b.run(max_cycles=1, use_graph=False)
            now = time.perf_counter() - t0
            for k in list(live):
                st = b.status(live[k])
                if st["done"]:
                    if not st["converged"]:
                        raise RuntimeError("request %d did not converge" % k)
                    lat.append(now - arr[k])
                    used -= footprint(live[k], m, mixed)
                    b.remove(live[k])
                    del live[k]
        span = time.perf_counter() - t0 - arr[0]
        b.free()
        return lat, span

    # a running server has tuned the batch tiles of the shapes its traffic
    # makes: untimed passes over two other arrival sequences at the highest
    # rate fill the tile cache
    t0 = time.perf_counter()
    for w in (1, 2):
        warm = np.cumsum(np.random.default_rng(args.seed + w).exponential(1.0 / max(args.rates),
                                                                           len(probs)))
        serve(warm, True)
        serve(warm, False)
    t_warm = time.perf_counter() - t0

    def served(lat, span, seq_span):
        return {"solves_per_s": len(probs) / span,
                "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                               "p99": float(np.percentile(lat, 99) * 1e3)},
                "speedup_solves_per_s": seq_span / span}
    meas = 0.0
    for lam in args.rates:
        arr = np.cumsum(rng.exponential(1.0 / lam, len(probs)))
        # sequential FIFO server
        fin, seq_lat = 0.0, []
        for a, sv in zip(arr, service):
            fin = max(fin, a) + sv
            seq_lat.append(fin - a)
        seq_span = fin - arr[0]
        lat64, span64 = serve(arr, False)
        lat, span = serve(arr, True)
        meas = max([meas] + [fp64_measure(p) for p in probs if not is_dg(p)])
        r = {"rate_per_s": lam}
        r.update(served(lat, span, seq_span))
        r.update({"fp64_members": served(lat64, span64, seq_span),
                  "sequential_solves_per_s": len(probs) / seq_span,
                  "sequential_latency_ms": {"p50": float(np.percentile(seq_lat, 50) * 1e3),
                                            "p99": float(np.percentile(seq_lat, 99) * 1e3)}})
        results.append(r)
    best = max(results, key=lambda r: r["solves_per_s"])
    out = {
        "metric": "served solves/s and p50/p99 latency at 1e-8 rel residual",
        "value": best["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32 inner solves, f64 refinement", "data": "synthetic",
        "config": {"workload": "Poisson arrivals of %d prism Poisson requests: std-linear (%d-%d "
                               "prisms) and, with probability %.2f, DG p=1 or p=2 (%d-%d prisms), "
                               "log-uniform sizes, lognormal kappa, GMRES(30)+(block) Jacobi 1e-8 "
                               "(mixed-precision members: FP32 inner solves, FP64 refinement to "
                               "the same FP64 measure), <= %d slots, %.0f GB admission budget; "
                               "sequential baseline: plain FP64 solves"
                               % (args.count, args.el_min, args.el_max, args.dg_frac,
                                  args.dg_el_min, args.dg_el_max, args.slots, args.hbm_gb),
                   "dg_requests": int(sum(is_dg(p) for p in probs))},
        "rates": results,
        "max_fp64_measure_of_served_solutions": meas,
        "sequential_mean_service_ms": float(np.mean(service) * 1e3),
        "tile_warmup_pass_s": t_warm,
    }
    for p in probs:
        p.free()
    return out


# ---------------------------------------------------------------------------
# DG workload: NIPG Poisson on a jittered prism box, p = 1, 2, 3
# ---------------------------------------------------------------------------
def dg_mesh(mf, n, seed):
    """jittered unit-cube prism box, bottom/top faces Neumann (flag 11), sides
    Dirichlet (flag 1), as the reference's DG cube test"""
    a = mf.Mesh.box_prism(n, n, n, seed=seed, jitter=0.2).arrays()
    bc = a["bf_bc"].copy()
    bc[a["bf_side"] < 2] = 11
    return mf.Mesh.from_arrays(a["coor"], a["el_type"], a["el_nodes"], a["bf_elem"], a["bf_side"],
                               bc), a


def dg_counts(pdeg):
    px, pz = pdeg % 100, pdeg // 100
    ns = (px + 1) * (px + 2) // 2 * (pz + 1)
    ngt = {1: 3, 2: 6, 3: 12}[px]
    ngq = (px + 1) * (pz + 1)
    ng = ngt * (pz + 1)
    return ns, ng, ngt, ngq


def dg_cpu_sample(a, pdeg, sample):
    """oracle (C, one core): element and face integrals of the first `sample`
    elements and of their faces"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes
    from oracle_lib import dg_faces, oracle
    en = a["el_nodes"][:sample]
    nodes = np.ascontiguousarray(a["coor"][en].reshape(sample, 18))
    gids = np.ascontiguousarray(en, np.int32)
    faces = []
    for (e1, s1, e2, s2) in dg_faces(en):
        faces.append((e1, s1, e2, s2) if e2 >= 0 else (e1, s1, -1, -1))
    faces = np.ascontiguousarray(np.array(faces, np.int32))
    ind = np.zeros(23, np.int32)
    ind[[2, 6, 10, 22]] = 1
    coef = np.zeros(21)
    coef[0] = coef[4] = coef[8] = 1.0
    pen = min(pdeg % 100, pdeg // 100) * 1.0
    P = lambda v: v.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    t0 = time.perf_counter()
    oracle().orc_dg_assemble_sample(pdeg, sample, P(nodes), P(gids), len(faces), P(faces), P(ind),
                                    P(coef), pen, 1)
    return sample / (time.perf_counter() - t0), len(faces)


def dg_work(dg, pd):
    """algorithmic counts of one assembly and one operator application of a
    DG problem with degrees pd per element: (flops, operator bytes).  Per
    quadrature point, (8 ns_test ns_trial + 32 ns_test) flops (the 4x4 point
    matrix applied to both function sets); own blocks take the volume points
    and every face's points, coupling blocks their face's points; a face
    between two degrees has the rule of the higher one"""
    nb, _, _ = dg.faces()
    p = pd % 100
    ns = (p + 1) * (p + 2) // 2 * (p + 1)
    tri = np.array([0, 3, 6, 12])
    ngv = tri[p] * (p + 1)
    inner = nb >= 0
    nbs = np.where(inner, nb, 0)
    pn = np.where(inner, p[nbs], p[:, None])
    q = np.maximum(p[:, None], pn)
    pts = np.where(np.arange(5)[None, :] < 2, tri[q], (q + 1) ** 2)
    nsn = np.where(inner, ns[nbs], 0)
    flops = ((ngv + pts.sum(1)) * (8.0 * ns * ns + 32.0 * ns)).sum()
    flops += (pts * inner * (8.0 * ns[:, None] * nsn + 32.0 * ns[:, None])).sum()
    op_bytes = 8.0 * (2 * ns * ns + (ns[:, None] * nsn).sum(1) + 2 * ns + nsn.sum(1)).sum()
    return float(flops), float(op_bytes)


def dg_work_hybrid(dg, pd, et):
    """dg_work for prism / tet meshes: tets have (p+1)(p+2)(p+3)/6 functions,
    (p+2)(p+1)^2 volume points and four triangle sides"""
    nb, _, _ = dg.faces()
    p = pd % 100
    tet = et == 7
    ns = np.where(tet, (p + 1) * (p + 2) * (p + 3) // 6, (p + 1) * (p + 2) // 2 * (p + 1))
    tri = np.array([0, 3, 6, 12])
    ngv = np.where(tet, (p + 2) * (p + 1) ** 2, tri[p] * (p + 1))
    inner = nb >= 0
    nbs = np.where(inner, nb, 0)
    pn = np.where(inner, p[nbs], p[:, None])
    q = np.maximum(p[:, None], pn)
    side = np.arange(5)[None, :]
    exists = ~(tet[:, None] & (side == 4))
    is_tri = tet[:, None] | (side < 2)
    pts = np.where(is_tri, tri[q], (q + 1) ** 2) * exists
    nsn = np.where(inner, ns[nbs], 0)
    flops = ((ngv + pts.sum(1)) * (8.0 * ns * ns + 32.0 * ns)).sum()
    flops += (pts * inner * (8.0 * ns[:, None] * nsn + 32.0 * ns[:, None])).sum()
    op_bytes = 8.0 * (2 * ns * ns + (ns[:, None] * nsn).sum(1) + 2 * ns + nsn.sum(1)).sum()
    return float(flops), float(op_bytes)


def bench_dg_hybrid(mf, args, cc):
    """config 3's element mix: z-layers of the n^3 hex box split into prisms
    or into tetrahedra (a layer is tetrahedral with probability tet_frac:
    0.18 gives about 40% tets by count), degrees drawn from {1, 2, 3} per
    element; bottom/top Neumann, sides Dirichlet; assembly, block-Jacobi
    fill and operator, and (with --dg-solve hybrid) GMRES(30) + block ILU(0)
    to 1e-8"""
    m0 = mf.Mesh.box_hybrid(args.dg_n, args.dg_n, args.dg_n, seed=args.seed, jitter=0.2,
                            tet_frac=args.dg_tet_frac)
    a = m0.arrays()
    m0.free()
    X, en, et = a["coor"], a["el_nodes"], a["el_type"]
    bc = a["bf_bc"].copy()
    zmax = X[:, 2].max()
    for k, (e, sd) in enumerate(zip(a["bf_elem"], a["bf_side"])):
        fv = ([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]][sd] if et[e] == 7 else
              [[0, 1, 2], [3, 4, 5], [0, 1, 4, 3], [1, 2, 5, 4], [2, 0, 3, 5]][sd])
        z = X[en[e][fv], 2]
        if np.all(np.abs(z) < 1e-12) or np.all(np.abs(z - zmax) < 1e-12):
            bc[k] = 11
    mesh = mf.Mesh.from_arrays(X, et, en, a["bf_elem"], a["bf_side"], bc)
    nel = mesh.nel
    pd = np.random.default_rng(args.seed + 1).choice([101, 202, 303], size=nel)
    dg = mf.DgProblem(mesh, pd)
    dg.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc, src_model=mf.SRC_DIFF_IN_CUBE)
    t0 = time.perf_counter()
    dg.assemble()
    t_first = time.perf_counter() - t0
    for _ in range(args.warmup):
        dg.assemble()
    t0 = time.perf_counter()
    dg.fill_precon()
    t_prec = time.perf_counter() - t0
    ms_asm, ms_op = dg.bench(max(args.steps, 1))
    flops, op_bytes = dg_work_hybrid(dg, pd, et)
    rec = {"elements": int(nel), "tets": int((et == 7).sum()), "prisms": int((et == 5).sum()),
           "dofs": int(dg.ndof), "assemble_ms": ms_asm, "elements_per_s": nel / (ms_asm * 1e-3),
           "assemble_tflops": flops / (ms_asm * 1e-3) / 1e12, "first_assemble_s": t_first,
           "block_jacobi_ms": t_prec * 1e3, "op_ms": ms_op,
           "op_GBs": op_bytes / (ms_op * 1e-3) / 1e9}
    if "hybrid" in args.dg_solve:
        t0 = time.perf_counter()
        rec["ilu_levels"] = dg.fill_ilu()
        rec["ilu_factor_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        x, info = dg.gmres(precon=3, krylbas=30, max_iter=args.dg_max_iter, tol=1e-8, fetch=False)
        solve_s = time.perf_counter() - t0
        rec["solve"] = {"preconditioner": "block ILU(0), element order", "gmres_iters": info["iters"],
                        "gmres_converged": info["converged"], "gmres_s": solve_s,
                        "gmres_res_rel": info["res_rel"]}
    dg.free()
    mesh.free()
    print(json.dumps({"dg_hybrid": rec}), file=sys.stderr, flush=True)
    return rec


def bench_dg(args):
    """NIPG DG Poisson (DIFF_IN_CUBE: unit diffusion, manufactured source,
    Dirichlet sides, Neumann top/bottom, penalty constant 1) on a jittered
    n^3 prism box with the degree of every element drawn from {1, 2, 3}:
    device assembly of the block system (own + rectangular coupling blocks),
    block-Jacobi fill, the block-Jacobi operator, and (on request) GMRES(30)
    to 1e-8; the uniform degrees beside it"""
    import modfem_amd as mf
    mf.init()
    mesh, a = dg_mesh(mf, args.dg_n, args.seed)
    nel = mesh.nel
    cc = np.r_[np.eye(3).ravel(), np.zeros(12)]
    pd_mixed = np.random.default_rng(args.seed).choice([101, 202, 303], size=nel)
    cases = [("mixed", pd_mixed)] + [(str(p), np.full(nel, p)) for p in args.dg_pdegs]
    res = {}
    for name, pd in cases:
        dg = mf.DgProblem(mesh, pd if name == "mixed" else int(pd[0]))
        dg.set_coefficients(mf.IND_DIFF, mf.COEF_UNIFORM, coef_const=cc, src_model=mf.SRC_DIFF_IN_CUBE)
        if args.dg_valu_ns is not None:
            dg.set_valu_blocks(args.dg_valu_ns)
        t0 = time.perf_counter()
        dg.assemble()
        t_first = time.perf_counter() - t0
        for _ in range(args.warmup):
            dg.assemble()
        t0 = time.perf_counter()
        dg.fill_precon()
        t_prec = time.perf_counter() - t0
        ms_asm, ms_op = dg.bench(max(args.steps, 1))
        # the same assembly with every block on the MFMA kernels
        dg.set_valu_blocks(0)
        ms_mfma, _ = dg.bench(max(args.steps, 1))
        dg.set_valu_blocks(DG_VALU_NS if args.dg_valu_ns is None else args.dg_valu_ns)
        dg.assemble()
        flops, op_bytes = dg_work(dg, pd)
        rec = {"pdeg": name, "dofs": int(dg.ndof), "assemble_ms": ms_asm,
               "elements_per_s": nel / (ms_asm * 1e-3), "assemble_tflops": flops / (ms_asm * 1e-3) / 1e12,
               "first_assemble_s": t_first, "block_jacobi_ms": t_prec * 1e3,
               "assemble_ms_all_mfma": ms_mfma,
               "op_ms": ms_op, "op_GBs": op_bytes / (ms_op * 1e-3) / 1e9}
        if name in args.dg_solve:
            # block Gauss-Seidel smoothing diverges on mixed-degree systems
            # (spectral radius 1.5-1.8 of the sweep): block Jacobi there;
            # block ILU(0) in element order for every case
            solves = {}
            precons = args.dg_precons or ((1, 3) if name == "mixed" else (2, 3))
            for precon in precons:
                label = {1: "block Jacobi", 2: "block Gauss-Seidel, 2 multicolour sweeps",
                         3: "block ILU(0), element order",
                         4: "p-multigrid V-cycle (2+2 block GS sweeps, 2 ILU(0) on p=1)"}[precon]
                if precon == 4:
                    if name == "mixed" or int(name) not in (202, 303):
                        continue
                    t0 = time.perf_counter()
                    rec["mg_levels"] = dg.set_multigrid(2, 2, 2)
                    rec["mg_setup_s"] = time.perf_counter() - t0
                if precon == 2:
                    rec["gs_colours"] = dg.set_gs(2)
                if precon == 3:
                    t0 = time.perf_counter()
                    rec["ilu_levels"] = dg.fill_ilu()
                    rec["ilu_factor_s"] = time.perf_counter() - t0
                t0 = time.perf_counter()
                x, info = dg.gmres(precon=precon, krylbas=30, max_iter=args.dg_max_iter, tol=1e-8,
                                   fetch=False)
                solve_s = time.perf_counter() - t0
                solves[label] = {"gmres_iters": info["iters"], "gmres_converged": info["converged"],
                                 "gmres_s": solve_s, "gmres_res_rel": info["res_rel"],
                                 "gmres_ms_per_iter": solve_s * 1e3 / max(info["iters"], 1)}
            rec["solves"] = solves
        if not args.no_cpu and name != "mixed":
            sample = min(nel, {101: 20000, 202: 4000}.get(int(name), 800))
            rate, _ = dg_cpu_sample(a, int(name), sample)
            rec["cpu_elements_per_s"] = rate
        res[name] = rec
        dg.free()
        print(json.dumps({"dg_level": rec}), file=sys.stderr, flush=True)
    if args.dg_tet_frac > 0:
        res["hybrid"] = bench_dg_hybrid(mf, args, cc)
    r = res["mixed"]
    out = {
        "metric": METRIC, "value": r["elements_per_s"], "unit": "elements integrated+assembled/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["assemble_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded jittered box, seeded degrees, manufactured exact solution)",
        "config": {"workload": "DG NIPG Poisson, %d^3 hex box x 2 prisms (%d prisms, jitter 0.2h), "
                               "degree of every element drawn uniformly from {1, 2, 3} (%d DOFs); "
                               "uniform degrees %s beside it" % (args.dg_n, nel, r["dofs"],
                                                                 list(args.dg_pdegs)),
                   "levels": res},
        "roofline": {"bound": "mfma", "achieved": r["assemble_tflops"], "peak": FP64_PEAK_TFS,
                     "unit": "TFLOP/s", "frac": r["assemble_tflops"] / FP64_PEAK_TFS,
                     "traffic": None, "kernel": "k_dg_diag + k_dg_couple, mixed degrees"},
        "roofline_op": {"bound": "hbm", "achieved": r["op_GBs"], "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": r["op_GBs"] / HBM_PEAK_GBS,
                        "kernel": "k_dg_op, mixed degrees"},
    }
    if not args.no_cpu:
        lv = [res[str(p)] for p in (101, 202, 303) if str(p) in res]
        if len(lv) == 3:
            cr = 3.0 / sum(1.0 / v["cpu_elements_per_s"] for v in lv)
            out["cpu_baseline"] = {"value": cr, "unit": "elements integrated+assembled/s", "cores": 1,
                                   "kind": "port",
                                   "sample": "oracle (C, one core) element + face integrals of the first "
                                             "20000/4000/800 elements at p=1/2/3 and their faces, "
                                             "combined for equal shares of the three degrees"}
    return out


# ---------------------------------------------------------------------------
# BCRS workload: a 4-component reaction-diffusion system in 4x4 blocks
# ---------------------------------------------------------------------------
BCRS_COUPLING = np.array([[1.0, 0.2, 0.0, -0.1], [-0.2, 1.0, 0.1, 0.0],
                          [0.0, -0.1, 1.0, 0.2], [0.1, 0.0, -0.2, 1.0]])


def bcrs_system(mf, n, seed):
    """element data of the BCRS workload: a jittered n^3 prism box, scalar
    diffusion + reaction (c = 10) + unit source element matrices from the
    device integration, coupled over 4 components as kron(K_e, C) with a
    fixed non-symmetric C; node blocks, neighbour lists flattened"""
    mesh = mf.Mesh.box_prism(n, n, n, seed=seed, jitter=0.2)
    a = mesh.arrays()
    en = np.ascontiguousarray(a["el_nodes"], np.int32)
    nn = a["coor"].shape[0]
    prob = mf.Problem.from_mesh(mesh)
    cc = np.zeros(mf.NCOEF)
    cc[[0, 4, 8]] = 1.0
    cc[mf.C_C] = 10.0
    cc[mf.C_S] = 1.0
    prob.set_coefficients(mf.coeff_indicator("axx", "ayy", "azz", "c", "s"), mf.COEF_UNIFORM,
                          coef_const=cc)
    em = prob.element_matrices()
    prob.free()
    iu, iv = np.nonzero(~np.eye(6, dtype=bool))
    u = en[:, iu].ravel().astype(np.int64)
    v = en[:, iv].ravel().astype(np.int64)
    key = np.unique(u * nn + v)
    nb_ind = (key % nn).astype(np.int32)
    nb_cnt = np.bincount((key // nn).astype(np.int64), minlength=nn).astype(np.int32)
    return mesh, en, nn, em, nb_ind, nb_cnt


def bcrs_allocate(L, nn, nb_ind, nb_cnt, alloc):
    """allocate through the lar_ seam (or the reference's own BCRS module)"""
    import ctypes
    ptr = np.zeros(nn + 1, np.int64)
    ptr[1:] = np.cumsum(nb_cnt)
    base = nb_ind.ctypes.data
    ptrs = (ctypes.c_void_p * nn)(*(base + 4 * ptr[:-1]).tolist())
    return alloc(nn, ctypes.cast(ptrs, ctypes.c_void_p))


def bcrs_assemble(en, em, assemble_el, chunk=16384):
    """kron(K_e, C) element by element, column-major 24x24, rhs kron(F_e, w)"""
    C = BCRS_COUPLING
    w = np.array([1.0, 0.5, -0.25, 1.0])
    nel = en.shape[0]
    for e0 in range(0, nel, chunk):
        e1 = min(nel, e0 + chunk)
        Ks = em[e0:e1, :36].reshape(-1, 6, 6).transpose(0, 2, 1)  # [row][col]
        K = np.einsum("eij,ab->eiajb", Ks, C).reshape(-1, 24, 24)
        Kc = np.ascontiguousarray(K.transpose(0, 2, 1)).reshape(-1, 576)  # column-major
        F = np.ascontiguousarray(np.einsum("ei,a->eia", em[e0:e1, 36:], w).reshape(-1, 24))
        for q in range(e1 - e0):
            assemble_el(en[e0 + q], Kc[q], F[q])


def bench_bcrs(args):
    """BCRS 4x4-block storage through the lar_ seam: device residual, block
    Jacobi / Gauss-Seidel sweeps, block ILU(0) factorisation and solves, and
    GMRES(30) with each preconditioner to 1e-8; the reference's own compiled
    residual (lar_compute_residual_bcrs) on the host beside it"""
    import ctypes
    import modfem_amd as mf
    mf.init()
    L = mf.lib()
    P = lambda v: None if v is None else v.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    t0 = time.perf_counter()
    mesh, en, nn, em, nb_ind, nb_cnt = bcrs_system(mf, args.bcrs_n, args.seed)
    mid = bcrs_allocate(L, nn, nb_ind, nb_cnt,
                        lambda nb, lists: L.lar_allocate_SM_and_LV(3, 1, 4 * nb, 24, nb, 4, None, None,
                                                                   P(nb_cnt), lists))
    if mid < 0:
        raise RuntimeError(L.mf_last_error().decode())
    tab = np.zeros(42, np.int32)

    def asm(l_bl, K, F):
        L.lar_fill_assembly_table_int_ent(mid, 6, P(l_bl), None, P(tab))
        L.lar_assemble_SM_and_LV_with_table(mid, 0, 6, P(tab), P(K), P(F), None)
    bcrs_assemble(en, em, asm)
    t_setup = time.perf_counter() - t0
    nblk = ctypes.c_longlong()
    L.mf_lar_get_sizes(mid, None, None, None, ctypes.byref(nblk))
    nblk = nblk.value
    reps = max(args.steps, 1) * 10
    ms = ctypes.c_double()

    def bench(what):
        rc = L.mf_lar_bench(mid, what, reps, ctypes.byref(ms))
        if rc < 0:
            raise RuntimeError(L.mf_last_error().decode())
        return ms.value
    res_ms = bench(0)
    res_bytes = 132.0 * nblk + 4.0 * (nn + 1) + 96.0 * nn
    if args.probe:
        # counter probe: 11 residual launches, then the copy yardstick
        cms = mf.bench_copy(COPY_BYTES, reps=3)
        print(json.dumps({"probe": True, "residual_launches": reps + 1, "residual_bytes": res_bytes,
                          "copy_bytes": COPY_BYTES, "copy_launches": 4,
                          "copy_GBs": 2 * COPY_BYTES / (cms * 1e-3) / 1e9}))
        L.lar_free_SM_and_LV(mid)
        return
    rec = {"blocks": nn, "stored_blocks": nblk, "dofs": 4 * nn, "setup_s": t_setup,
           "residual_ms": res_ms, "residual_GBs": res_bytes / (res_ms * 1e-3) / 1e9}
    x = np.zeros(4 * nn)
    it, rr, rs = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    cases = [("bj", 1, 1, 0), ("gs", 2, 2, 0), ("gs_colour", 2, 2, 0), ("ilu", 5, 1, 0)]
    if args.bcrs_iluk > 0:
        cases.append(("ilu%d" % args.bcrs_iluk, 5, 1, args.bcrs_iluk))
    for name, precon, nr_prec, iluk in cases:
        t0 = time.perf_counter()
        assert L.lar_allocate_preconditioner(mid, precon, iluk) == nn, L.mf_last_error()
        if iluk:
            ns = ctypes.c_longlong()
            L.mf_lar_get_ilu_pattern(mid, None, None, ctypes.byref(ns))
            rec[name + "_stored_blocks"] = ns.value
            rec[name + "_allocate_s"] = time.perf_counter() - t0
            print(json.dumps({"bcrs_" + name + "_pattern": {"stored_blocks": ns.value,
                                                           "s": rec[name + "_allocate_s"]}}),
                  file=sys.stderr, flush=True)
        if precon == 2:
            rec[name + "_colours"] = L.mf_lar_set_gs_order(mid, 1 if name == "gs_colour" else 0)
        t0 = time.perf_counter()
        if L.lar_fill_preconditioner(mid) != 0:
            raise RuntimeError(L.mf_last_error().decode())
        rec[name + "_fill_ms"] = (time.perf_counter() - t0) * 1e3
        print(json.dumps({"bcrs_" + name + "_fill_ms": rec[name + "_fill_ms"]}), file=sys.stderr,
              flush=True)
        if precon == 5:
            rec[name + "_factor_ms"] = bench(2)
            rec[name + "_solve_ms"] = bench(3)
        else:
            rec[name + "_sweep_ms"] = bench(1)
        t0 = time.perf_counter()
        conv = L.mf_lar_gmres(mid, 30, args.bcrs_max_iter, 1e-8, nr_prec, 1, P(x), ctypes.byref(it),
                              ctypes.byref(rr), ctypes.byref(rs))
        dt = time.perf_counter() - t0
        rec[name + "_gmres"] = {"converged": conv == 1, "iters": it.value, "res_rel": rr.value,
                                "s": dt, "ms_per_iter": dt * 1e3 / max(it.value, 1)}
        print(json.dumps({"bcrs_" + name: rec[name + "_gmres"]}), file=sys.stderr, flush=True)
        assert L.lar_free_preconditioner(mid) == 0
    # the residual vector of the solution, as a check of the last solve
    r = np.zeros(4 * nn)
    L.lar_compute_residual(mid, 1, 0, 4 * nn, P(x), None, P(r))
    rhs = np.zeros(4 * nn)
    L.mf_lar_get_system(mid, None, None, None, P(rhs))
    rec["true_res_rel"] = float(np.linalg.norm(r) / np.linalg.norm(rhs))
    L.lar_free_SM_and_LV(mid)
    out = {
        "metric": "BCRS 4x4 residual v = b - A x (lar_compute_residual)", "value": rec["residual_GBs"],
        "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded jittered box, device element matrices, fixed coupling)",
        "config": {"workload": "4-component reaction-diffusion in BCRS 4x4 blocks, %d^3 hex box x 2 "
                               "prisms (%d node blocks, %d stored blocks)" % (args.bcrs_n, nn, nblk),
                   "levels": rec},
        "roofline": {"bound": "hbm", "achieved": rec["residual_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": rec["residual_GBs"] / HBM_PEAK_GBS,
                     "traffic": bcrs_traffic(args.bcrs_n), "kernel": "k_bcrs_residual"},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = bcrs_cpu_baseline(en, em, nn, nb_ind, nb_cnt, res_bytes)
    print(json.dumps(out), flush=True)


def bcrs_traffic(n):
    """HBM bytes per launch of k_bcrs_residual from the committed FETCH_SIZE /
    WRITE_SIZE counter summary of this workload (None if absent)"""
    path = os.path.join(ROOT, "profiles", "k_bcrs_residual_box48_pmc.txt")
    if n != 48 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def bcrs_cpu_baseline(en, em, nn, nb_ind, nb_cnt, res_bytes):
    """the reference's own lar_compute_residual_bcrs (compiled from its
    sources into oracle/_ref) on one host thread, on the same system"""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import ref, ref_available
    if not ref_available("libref_bcrs.so"):
        return None
    R = ref("libref_bcrs.so")
    gomp = ctypes.CDLL("libgomp.so.1")
    threads = gomp.omp_get_max_threads()
    gomp.omp_set_num_threads(1)
    R.lar_allocate_SM_and_LV_bcrs.restype = ctypes.c_int
    R.lar_allocate_SM_and_LV_bcrs.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2
    R.lar_assemble_SM_and_LV_bcrs.restype = ctypes.c_int
    R.lar_assemble_SM_and_LV_bcrs.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 5
    R.lar_compute_residual_bcrs.restype = None
    R.lar_compute_residual_bcrs.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 3
    R.lar_free_SM_and_LV_bcrs.argtypes = [ctypes.c_int]
    P = lambda v: v.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    # the reference prints progress on stdout: send it to stderr, so that
    # stdout keeps the one JSON line
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        return bcrs_cpu_time(R, P, en, em, nn, nb_ind, nb_cnt, res_bytes, gomp, threads)
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)


def bcrs_cpu_time(R, P, en, em, nn, nb_ind, nb_cnt, res_bytes, gomp, threads):
    rid = bcrs_allocate(R, nn, nb_ind, nb_cnt,
                        lambda nb, lists: R.lar_allocate_SM_and_LV_bcrs(4 * nb, 24, 4, P(nb_cnt), lists))

    def asm(l_bl, K, F):
        R.lar_assemble_SM_and_LV_bcrs(rid, 0, 6, P(l_bl), None, P(K), P(F), None)
    bcrs_assemble(en, em, asm)
    x = np.random.default_rng(1).uniform(-1, 1, 4 * nn)
    v = np.zeros(4 * nn)
    reps, t = 0, 0.0
    while t < 10.0 and reps < 200:
        t0 = time.perf_counter()
        R.lar_compute_residual_bcrs(rid, 1, 0, 4 * nn, P(x), None, P(v))
        t += time.perf_counter() - t0
        reps += 1
    R.lar_free_SM_and_LV_bcrs(rid)
    gomp.omp_set_num_threads(threads)
    return {"value": res_bytes / (t / reps) / 1e9, "unit": "GB/s", "cores": 1, "kind": "reference",
            "sample": "%d calls of the reference's lar_compute_residual_bcrs (las_bcrs_intf.c, built "
                      "from its sources, OpenMP at one thread) on the full system" % reps}


def device_setup_figures(mf, mesh, p, t_host):
    """the problem set-up (DOF numbering, pattern, tables, device order)
    built on the host and on the device; the device maps checked against
    the host's (DOFs and pattern here, everything in the GPU tests)"""
    t0 = time.perf_counter()
    pd = mf.Problem.from_mesh(mesh, device_setup=True)
    t_dev = time.perf_counter() - t0
    same = bool(np.array_equal(pd.dofs(), p.dofs()) and
                all(np.array_equal(a, b) for a, b in zip(pd.pattern(), p.pattern())))
    pd.free()
    return {"host_s": t_host, "device_s": t_dev, "speedup": t_host / t_dev,
            "device_maps_equal_host": same}


def other_workloads(args):
    """the batch (256 problems), served-stream and mixed-degree DG
    configurations, compact, measured in the same run"""
    out = {}
    a = argparse.Namespace(**vars(args))
    a.no_cpu = True
    a.count, a.el_min, a.el_max = 256, 2000, 200000
    print("workload: batch of 256 problems", file=sys.stderr, flush=True)
    r = bench_batch(a)
    out["batch_256"] = {k: r[k] for k in ("value", "unit", "config", "mixed_batch_s", "speedup_vs_loop",
                                          "mixed_all_converged", "mixed_max_fp64_measure",
                                          "fp64_batch_s", "fp64_speedup_vs_loop", "loop_s",
                                          "loop_solves_per_s", "all_converged", "iters",
                                          "max_rel_diff_vs_loop_solution", "tiles")}
    a.count, a.el_min, a.el_max, a.rates = 128, 5000, 500000, [1000.0, 2000.0, 4000.0]
    print("workload: served stream", file=sys.stderr, flush=True)
    a.dg_frac = 0.0
    r = bench_serve(a)
    out["serve"] = {k: r[k] for k in ("value", "unit", "config", "rates",
                                      "max_fp64_measure_of_served_solutions")}
    print("workload: served stream with DG requests", file=sys.stderr, flush=True)
    a.dg_frac, a.dg_el_min, a.dg_el_max, a.rates = 0.25, 2000, 10000, [2000.0]
    r = bench_serve(a)
    out["serve_with_dg"] = {k: r[k] for k in ("value", "unit", "config", "rates",
                                              "max_fp64_measure_of_served_solutions")}
    # config 3 at its size: assembly timings, then GMRES(30) with block
    # ILU(0) to 1e-8 (block Gauss-Seidel diverges on mixed degrees and block
    # Jacobi needs thousands of iterations)
    a.dg_n, a.dg_pdegs, a.dg_solve, a.dg_precons, a.steps, a.warmup = 96, [], ["mixed", "hybrid"], [3], 3, 1
    a.dg_tet_frac = 0.18
    print("workload: DG mixed degrees 96^3 (prisms, and prism/tet layers), assembly and "
          "block-ILU(0) solves", file=sys.stderr, flush=True)
    r = bench_dg(a)
    out["dg_mixed_96"] = {k: r[k] for k in ("value", "unit", "config", "roofline", "roofline_op")}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="hexes per box edge")
    ap.add_argument("--krylbas", type=int, default=50)
    ap.add_argument("--seed", type=int, default=2017)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="the default workload only (no batch, serve and DG figures)")
    ap.add_argument("--cpu-iters", type=int, default=50)
    ap.add_argument("--probe", action="store_true",
                    help="short run for hardware-counter passes: one integrate+assemble, one "
                         "GMRES restart and a 1 GiB stream copy (the counters' yardstick)")
    ap.add_argument("--dg-n", type=int, default=96, help="DG workload: hexes per box edge")
    ap.add_argument("--dg-pdegs", type=int, nargs="*", default=[101, 202, 303],
                    help="DG workload: uniform degrees run beside the mixed-degree case")
    ap.add_argument("--dg-max-iter", type=int, default=6000)
    ap.add_argument("--dg-tet-frac", type=float, default=0.18,
                    help="DG workload: share of tetrahedral z-layers in the hybrid prism/tet case "
                         "(0.18: about 40%% tets by count); 0 skips the hybrid case")
    ap.add_argument("--dg-valu-ns", type=int, default=None,
                    help="DG blocks of at most this many shape functions assembled on the vector "
                         "ALUs (library default: %d)" % DG_VALU_NS)
    ap.add_argument("--dg-precons", type=int, nargs="*", default=[],
                    help="DG solves: preconditioners to run (1 block Jacobi, 2 block GS, 3 block "
                         "ILU(0), 4 p-multigrid V-cycle for 202 / 303); default 1 3 for mixed "
                         "degrees, 2 3 for uniform ones")
    ap.add_argument("--dg-solve", nargs="*", default=[],
                    help="DG cases solved to 1e-8: mixed (block Jacobi, block ILU(0)), hybrid "
                         "(block ILU(0)) and/or the uniform degrees 101 202 303 (block "
                         "Gauss-Seidel with 2 sweeps, block ILU(0))")
    ap.add_argument("--bcrs-n", type=int, default=48, help="BCRS workload: hexes per box edge")
    ap.add_argument("--bcrs-max-iter", type=int, default=3000)
    ap.add_argument("--bcrs-iluk", type=int, default=0,
                    help="BCRS workload: also solve with block ILU(k) for this k > 0")
    ap.add_argument("--workload", choices=["box128", "batch", "serve", "dg", "bcrs"], default="box128",
                    help="box128: the single large problem (default line); batch: many "
                         "independent problems in one pipeline; serve: a request stream; "
                         "dg: the DG NIPG Poisson case, p = 1, 2, 3")
    ap.add_argument("--count", type=int, default=256, help="problems (batch) or requests (serve)")
    ap.add_argument("--el-min", type=int, default=2000)
    ap.add_argument("--el-max", type=int, default=200000)
    ap.add_argument("--rates", type=float, nargs="+", default=[10.0, 100.0, 1000.0, 2000.0])
    ap.add_argument("--slots", type=int, default=256)
    ap.add_argument("--hbm-gb", type=float, default=64.0)
    ap.add_argument("--dg-frac", type=float, default=0.0, help="serve: share of DG requests")
    ap.add_argument("--dg-el-min", type=int, default=2000)
    ap.add_argument("--dg-el-max", type=int, default=10000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "dg":
        print(json.dumps(bench_dg(args)), flush=True)
        return
    if args.workload == "bcrs":
        return bench_bcrs(args)
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    import modfem_amd as mf
    t_setup = time.perf_counter()
    mesh, kappa, b, cc, coef_el = make_workload(args.n, args.seed)
    t_host_setup = time.perf_counter()
    p = mf.Problem.from_mesh(mesh)
    t_host_setup = time.perf_counter() - t_host_setup
    p.set_coefficients(mf.IND_CONV_DIFF, mf.COEF_KAPPA_B, coef_const=cc, coef_el=coef_el)
    p.set_assembly_mode(ASSEMBLY_MODE)
    t_setup = time.perf_counter() - t_setup
    if args.probe:
        p.integrate_assemble(64)
        _, info = p.gmres(precon=1, krylbas=args.krylbas, max_iter=args.krylbas, tol=1e-8)
        ms = mf.bench_copy(COPY_BYTES, reps=3)
        print(json.dumps({"probe": True, "gmres_iters": info["iters"], "copy_bytes": COPY_BYTES,
                          "copy_launches": 4, "copy_GBs": 2 * COPY_BYTES / (ms * 1e-3) / 1e9}))
        return

    def step(fetch=False):
        # inputs are device resident; the solution stays on the device unless
        # fetch (host copy through the C ABI, PCIe + host re-numbering)
        p.integrate_assemble(64)
        return p.gmres(precon=1, krylbas=args.krylbas, max_iter=20000, tol=1e-8, fetch=fetch)

    for _ in range(args.warmup):
        step()
    mf.gmres_kernel_stats(True)
    lat = []
    infos = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        x, info = step()
        lat.append(time.perf_counter() - ts)
        infos.append(info)
    t_total = time.perf_counter() - t0
    ks = mf.gmres_kernel_stats(False)
    x = p.solution()
    true_rel = p.true_residual(x)
    ts = time.perf_counter()
    step(fetch=True)
    t_fetch = time.perf_counter() - ts
    plain_solve_ms = p.last_solve_ms()
    # the same step with ILU(0) in multicolour order (structure built once,
    # factors rebuilt after every assembly, as the step re-assembles)
    p.set_ilu_order(1)
    p.integrate_assemble(64)
    p.gmres(precon=2, krylbas=args.krylbas, max_iter=20000, tol=1e-8, fetch=False)
    t0i = time.perf_counter()
    for _ in range(args.steps):
        p.integrate_assemble(64)
        _, iinfo = p.gmres(precon=2, krylbas=args.krylbas, max_iter=20000, tol=1e-8, fetch=False)
    t_ilu = (time.perf_counter() - t0i) / args.steps
    xi = p.solution()
    _, minfo_ilu = p.gmres_mixed(precon=2, krylbas=args.krylbas, max_iter=20000, tol=1e-8, fetch=False)
    ilu = {"ms_per_step": t_ilu * 1e3, "elements_per_s": p.nel / t_ilu,
           "iters": iinfo["iters"], "res_rel": iinfo["res_rel"],
           "true_rel_residual": p.true_residual(xi),
           "speedup_vs_jacobi_step": (t_total / args.steps) / t_ilu,
           "mixed_precision": {"solve_ms": p.last_solve_ms(), "iters": minfo_ilu["iters"],
                               "refinements": minfo_ilu["refinements"],
                               "res_rel_fp64": minfo_ilu["res_rel"],
                               "rel_diff_vs_fp64_solution":
                                   float(np.linalg.norm(p.solution() - xi) / np.linalg.norm(xi))}}
    # mixed-precision form (FP32-valued operator, FP64 refinement), same stop
    _, minfo = p.gmres_mixed(precon=1, krylbas=args.krylbas, max_iter=20000, tol=1e-8, fetch=False)
    mixed_ms = p.last_solve_ms()
    xm = p.solution()
    mixed = {"solve_ms": mixed_ms, "fp64_solve_ms": plain_solve_ms, "iters": minfo["iters"],
             "refinements": minfo["refinements"], "res_rel_fp64": minfo["res_rel"],
             "true_rel_residual": p.true_residual(xm),
             "rel_diff_vs_fp64_solution": float(np.linalg.norm(xm - x) / np.linalg.norm(x)),
             "speedup_vs_fp64_solve": plain_solve_ms / mixed_ms}
    modes = {}
    for mode, prec in [(0, 64), (1, 64), (2, 64), (2, 32)]:
        p.set_assembly_mode(mode)
        t_e, t_f = p.bench_integrate(prec, reps=5)
        modes["mode%d_fp%d_ms" % (mode, prec)] = t_e + t_f
    p.set_assembly_mode(ASSEMBLY_MODE)
    ms_int, ms_faces = p.bench_integrate(64, reps=5)
    ms_spmv = p.bench_spmv(reps=20)
    ms_copy = mf.bench_copy(COPY_BYTES, reps=5)

    nel, n, nnz = p.nel, p.n, p.nnz
    iters = infos[-1]["iters"]
    n_op, t_op, b_op = ks["op_dots"]
    n_up, t_up, b_up = ks["update"]
    ach_op = b_op / (t_op * 1e-3) / 1e9 if t_op > 0 else 0.0
    # integrate+assemble: R's FLOP model (aps_std_lin_intf.c:1847,2200,2266,2813):
    # G*(366 + 38*ns^2 + 3*ns) with convection, G=6, ns=6
    f_int = nel * 6 * (366 + 38 * 36 + 3 * 6)
    b_int = 8.0 * nnz + 8.0 * n + 24.0 * n + 4.0 * 6 * nel + 8.0 * 4 * nel + 36.0 * nel
    out = {
        "metric": METRIC,
        "value": nel * args.steps / t_total,
        "unit": "elements/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_total / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "std-linear prism convection-diffusion, 128^3 hex box x 2 prisms "
                               "(jitter 0.2h, lognormal kappa sigma=1, random b), FP64, "
                               "GMRES(%d)+Jacobi to 1e-8" % args.krylbas if args.n == 128 else
                               "std-linear prism convection-diffusion %d^3 hex box x 2 prisms" % args.n,
                   "elements": nel, "dofs": n, "nnz": nnz, "boundary_faces": p.nbf},
        "solves_per_s": args.steps / t_total,
        "host_boundary_ms_per_step": t_fetch * 1e3,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "gmres": {"iters": iters, "res_rel": infos[-1]["res_rel"], "true_rel_residual": true_rel},
        "integrate_assemble": {"elements_per_s": nel / ((ms_int + ms_faces) * 1e-3),
                               "ms_elements": ms_int, "ms_faces": ms_faces,
                               "fp64_tflops_R_model": f_int / (ms_int * 1e-3) / 1e12,
                               "frac_fp64_peak": f_int / (ms_int * 1e-3) / 1e12 / FP64_PEAK_TFS,
                               "algorithmic_GBs": b_int / (ms_int * 1e-3) / 1e9,
                               "modes_ms": modes},
        "copy_GBs": 2 * COPY_BYTES / (ms_copy * 1e-3) / 1e9,
        "spmv_GBs": (12.0 * nnz + 4.0 * (n + 1) + 16.0 * n) / (ms_spmv * 1e-3) / 1e9,
        "roofline": {"kernel": "k_op_dots (Jacobi-preconditioned SpMV + CGS dots)",
                     "bound": "hbm", "achieved": ach_op, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_op / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.n, args.krylbas),
                     "launches": int(n_op), "avg_launch_ms": t_op / max(n_op, 1),
                     "bytes_per_launch": b_op / max(n_op, 1)},
        "roofline_update": {"kernel": "k_update (CGS update + norm + Givens)",
                            "achieved": b_up / (t_up * 1e-3) / 1e9 if t_up > 0 else 0.0,
                            "launches": int(n_up), "avg_launch_ms": t_up / max(n_up, 1)},
        "gmres_mixed": mixed,
        "ilu0_multicolour": ilu,
        "setup_s": t_setup,
        "problem_setup": device_setup_figures(mf, mesh, p, t_host_setup),
    }
    if not args.quick:
        out["workloads"] = other_workloads(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, kappa, b, iters, args.krylbas, args.cpu_iters,
                                           args.seed)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
