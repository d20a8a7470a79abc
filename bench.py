#!/usr/bin/env python3
"""Benchmark: pressure-solve time to 1e-8 true relative residual on the
128^3 perturbed, permuted hex-mesh pressure Poisson system (2,097,152 cells,
14,581,760 nonzeros), mixed FP32/FP64 against plain FP64 on one MI355X, plus
SpMV bandwidth against the 8 TB/s HBM peak.

One "step" is one full pressure solve (x0 = 0, FP64 true residual <= 1e-8)
of the mixed-precision solver on a system assembled on the GPU from the SoA
mesh.  Inputs (matrix, rhs) are resident in HBM before the timed region.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_DOUBLES = 1 << 26  # calibration copy: 512 MiB read + 512 MiB written


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, vbytes, ibytes=4):
    # values + column indices + row pointers + x read once + y written once
    return nnz * (vbytes + ibytes) + (n + 1) * 4 + 2 * n * vbytes


def host_info():
    """nproc, the CPUs this process may run on, and the CPU model."""
    aff = sorted(os.sched_getaffinity(0))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpus_allowed": len(aff), "cpu": model}, aff


class Pinned:
    """Runs the enclosed CPU work on one host core (the calling thread's
    affinity; the HIP runtime's own threads keep theirs): the reference's
    solver path is single-threaded."""

    def __init__(self, core):
        self.core = core

    def __enter__(self):
        self.old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.old)


def _rel(a, b):
    import numpy as np

    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _pressure_oracle(oracle, w):
    import numpy as np

    rp, ci = oracle.build_sparsity(w.mesh)
    P = oracle.Matrix(rp, ci, np.zeros(ci.size))
    vals, _, pinned = oracle.assemble_pressure(w.mesh, P, w.bc_of_zone, gamma=w.gamma)
    b = w.b.copy()
    if pinned:
        b[0] = 0.0
    return oracle.Matrix(rp, ci, vals), b, pinned


def _gpu_solve(capi, Ao, b, **kw):
    cfg = dict(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=1e-12,
               max_iterations=50000, true_residual=1, max_outer=30)
    cfg.update(kw)
    s = capi.Solver(**cfg)
    s.setup(capi.Csr(Ao.rowPtr, Ao.colIdx, Ao.values))
    x, st, rc = s.solve(b)
    return x, st.as_dict(), rc


def cpu_leg(capi, wl, lib, A, b, x_mixed, d_b, d_x, sample_iters, core, host):
    """The reference's CPU path timed on one pinned host core, and each
    config's GPU solution against the FP64 oracle (the checker).

    Returns (cpu_baseline, error_vs_oracle).  The headline value is the
    reference algorithm for config 2 -- FP64 Jacobi-PCG (CGSolver::solve,
    src/solver/LinearSolvers.cpp:415-459) to 1e-8 on the 128^3 system: the
    iteration count comes from the GPU FP64 Jacobi-PCG, which runs the same
    recurrence (within +-2 iterations of the oracle), and the time per
    iteration from a bounded sample of the oracle's serial loop."""
    import numpy as np

    import oracle

    t_leg = time.perf_counter()
    rp, ci = A.pattern()
    Ao = oracle.Matrix(rp, ci, A.values())
    n = A.n
    # --- config 2: Jacobi-PCG to 1e-8 at 128^3 (projected)
    sj = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-8,
                     max_iterations=100000)
    sj.setup(A)
    st, rc = sj.solve_device(d_b, d_x)
    lib.mi_device_synchronize()
    t1 = time.perf_counter()
    st, rc = sj.solve_device(d_b, d_x)
    lib.mi_device_synchronize()
    gpu_j_ms = 1e3 * (time.perf_counter() - t1)
    its_j = st.as_dict()["iterations"]
    with Pinned(core):
        t1 = time.perf_counter()
        _, done, _, _ = oracle.cg(Ao, b, oracle.PC_JACOBI, 1e-30, sample_iters)
        dt = time.perf_counter() - t1
    per_it = dt / max(done, 1)
    value_ms = per_it * its_j * 1e3
    log(f"cpu: 128^3 Jacobi-PCG {per_it * 1e3:.2f} ms/iteration x {its_j} iterations "
        f"= {value_ms / 1e3:.1f} s (GPU FP64 Jacobi-PCG {gpu_j_ms:.1f} ms)")
    configs = {"2": {"jacobi_pcg_fp64": {
        "cpu_ms_projected": round(value_ms, 1), "ms_per_iteration": round(1e3 * per_it, 3),
        "iterations_to_1e-8": its_j, "sample_iterations": done, "sample_s": round(dt, 2),
        "gpu_fp64_jacobi_pcg_ms": round(gpu_j_ms, 2)}}}
    errors = {}

    # --- config 1: 128x128 cavity, Jacobi-PCG FP64 to 1e-8, full size
    wc = wl.cavity2d(128)
    Ac, bc, _ = _pressure_oracle(oracle, wc)
    tc = []
    with Pinned(core):
        for _ in range(5):
            t1 = time.perf_counter()
            xo8, ito8, _, _ = oracle.cg(Ac, bc, oracle.PC_JACOBI, 1e-8, 10000)
            tc.append(time.perf_counter() - t1)
    xot, _, _, _ = oracle.cg(Ac, bc, oracle.PC_JACOBI, 1e-13, 20000)
    xg8, st8, _ = _gpu_solve(capi, Ac, bc, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-8,
                             true_residual=0)
    xgt, stt, _ = _gpu_solve(capi, Ac, bc, preconditioner=capi.MI_PC_JACOBI)
    xgm, stm, _ = _gpu_solve(capi, Ac, bc, preconditioner=capi.MI_PC_JACOBI,
                             precision=capi.MI_PRECISION_MIXED)
    configs["1"] = {"workload": "128x128 cavity slab, Jacobi-PCG FP64 to 1e-8 (full size)",
                    "cpu_ms": round(1e3 * float(np.median(tc)), 2), "iterations": ito8}
    errors["1"] = {"size": "128x128x1",
                   "fp64_jacobi_1e-8_vs_oracle_1e-8": _rel(xg8, xo8),
                   "iterations_gpu_oracle": [st8["iterations"], ito8],
                   "fp64_1e-12_vs_oracle_1e-13": _rel(xgt, xot),
                   "mixed_1e-12_vs_oracle_1e-13": _rel(xgm, xot)}

    # --- config 2 at 48^3 (oracle-sized) and the 128^3 mixed solve
    w48 = wl.poisson3d(48)
    A48, b48, _ = _pressure_oracle(oracle, w48)
    xo48, _, _, _ = oracle.cg(A48, b48, oracle.PC_JACOBI, 1e-13, 50000)
    x64t, _, _ = _gpu_solve(capi, A48, b48)
    xmt, _, _ = _gpu_solve(capi, A48, b48, precision=capi.MI_PRECISION_MIXED)
    st_t = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=1e-12,
                       max_iterations=5000, true_residual=1, max_outer=30)
    st_t.setup(A)
    st_t.solve_device(d_b, d_x)
    x128t = d_x.download()
    errors["2"] = {"size": "48^3 (config-2 generator)",
                   "fp64_amg_1e-12_vs_oracle_1e-13": _rel(x64t, xo48),
                   "mixed_amg_1e-12_vs_oracle_1e-13": _rel(xmt, xo48),
                   "128^3_mixed_1e-8_vs_gpu_fp64_1e-12": _rel(x_mixed, x128t),
                   "128^3_true_relative_residual_mixed_host_checked":
                       float(np.linalg.norm(b - oracle.spmv(Ao, x_mixed)) / np.linalg.norm(b))}

    # --- config 2: the reference's RS-AMG-PCG (compiled reference sources)
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libsimall_ref.so")
    if os.path.exists(ref_path):
        import ctypes as C

        ref = C.CDLL(ref_path)
        P_, i_, d_ = C.c_void_p, C.c_int, C.c_double
        ref.ref_rs_amg_pcg.argtypes = [i_, P_, P_, P_, P_, d_, i_, P_, P_, P_]
        ref.ref_rs_amg_pcg.restype = i_
        rs = {}
        for N in (24, 32):
            Ar, br, _ = _pressure_oracle(oracle, wl.poisson3d(N))
            xr = np.zeros(Ar.n)
            last, setup = C.c_double(), C.c_double()
            with Pinned(core):
                t1 = time.perf_counter()
                it = ref.ref_rs_amg_pcg(Ar.n, Ar.rowPtr.ctypes.data, Ar.colIdx.ctypes.data,
                                        Ar.values.ctypes.data, br.ctypes.data, 1e-8, 10000,
                                        xr.ctypes.data, C.byref(last), C.byref(setup))
                tot = time.perf_counter() - t1
            rs[f"{N}^3"] = {"n": Ar.n, "setup_s": round(setup.value, 3),
                            "solve_s": round(tot - setup.value, 3), "iterations": it,
                            "last_residual": last.value}
        n1, n2 = rs["24^3"]["n"], rs["32^3"]["n"]
        p = np.log(rs["32^3"]["setup_s"] / rs["24^3"]["setup_s"]) / np.log(n2 / n1)
        rs["setup_growth_exponent"] = round(float(p), 2)
        rs["projected_128^3_setup_s"] = round(float(rs["32^3"]["setup_s"] * (n / n2) ** p), 0)
        rs["kind"] = "reference"
        configs["2"]["rs_amg_pcg"] = rs
        log(f"cpu: reference RS-AMG-PCG {rs}")

    # --- config 3: ILU(0)-BiCGStab FP64 to 1e-8, 96^3 full size
    w3 = wl.momentum3d(96)
    m3 = w3.mesh
    phi = oracle.face_flux(m3, w3.bc_of_zone, 1.0, w3.U)
    rp3, ci3 = oracle.build_sparsity(m3)
    v3, _ = oracle.assemble_momentum(m3, oracle.Matrix(rp3, ci3, np.zeros(ci3.size)),
                                     w3.bc_of_zone, 0, w3.mu, 1.0, phi, dt=w3.dt)
    A3 = oracle.Matrix(rp3, ci3, v3)
    with Pinned(core):
        t1 = time.perf_counter()
        _, it3, _, _ = oracle.bicgstab(A3, w3.b, oracle.PC_ILU, 1e-8, 2000)
        t3 = time.perf_counter() - t1
    xo3, _, _, _ = oracle.bicgstab(A3, w3.b, oracle.PC_ILU, 1e-13, 5000)
    bi = dict(kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_ILU)
    x3, _, _ = _gpu_solve(capi, A3, w3.b, **bi)
    x3m, _, _ = _gpu_solve(capi, A3, w3.b, precision=capi.MI_PRECISION_MIXED, **bi)
    configs["3"] = {"workload": "96^3 momentum, ILU(0)-BiCGStab FP64 to 1e-8 (full size, "
                                "factorization included)",
                    "cpu_ms": round(1e3 * t3, 1), "iterations": it3}
    errors["3"] = {"size": "96^3 (full)", "fp64_ilu_bicgstab_1e-12_vs_oracle_1e-13": _rel(x3, xo3),
                   "mixed_1e-12_vs_oracle_1e-13": _rel(x3m, xo3)}

    # --- config 4: 32 sequential Jacobi-PCG solves at 100^3 (projected), and
    # the batched solve against the oracle at 32^3
    w4 = wl.transport(100, 32)
    dm4 = capi.DeviceMesh(w4.mesh, w4.bc_of_zone)
    A4 = dm4.create_matrix()
    rhs4 = capi.DeviceArray(A4.n)
    phi4 = capi.DeviceArray.from_host(np.zeros(w4.mesh.n_faces))
    dt4 = capi.DeviceArray.from_host(np.full(A4.n, w4.dt))
    u4 = capi.DeviceArray.from_host(np.zeros(A4.n))
    s4 = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-8,
                     max_iterations=20000)
    its4 = []
    for e in range(32):
        dm4.assemble_momentum(A4, rhs4, 0, float(w4.gammas[e]), 1.0, phi4, d_dt=dt4, d_Uold=u4)
        s4.setup(A4)
        st, rc = s4.solve_device(capi.DeviceArray.from_host(w4.bs[e]), capi.DeviceArray(A4.n))
        its4.append(st.as_dict()["iterations"])
        if e == 0:
            rp4, ci4 = A4.pattern()
            A40 = oracle.Matrix(rp4, ci4, A4.values())
    with Pinned(core):
        t1 = time.perf_counter()
        _, d4, _, _ = oracle.cg(A40, w4.bs[0], oracle.PC_JACOBI, 1e-30, 150)
        t4 = time.perf_counter() - t1
    configs["4"] = {"workload": "32 transport equations on 100^3, one Jacobi-PCG FP64 solve each "
                                "to 1e-8 (iteration counts from the GPU FP64 Jacobi-PCG)",
                    "cpu_ms_projected": round(1e3 * t4 / d4 * sum(its4), 1),
                    "ms_per_iteration": round(1e3 * t4 / d4, 3),
                    "iterations_total": int(sum(its4)),
                    "iterations_min_max": [int(min(its4)), int(max(its4))]}
    w4s = wl.transport(32, 4)
    dm4s = capi.DeviceMesh(w4s.mesh, w4s.bc_of_zone)
    A4s = dm4s.create_matrix()
    rhs4s = capi.DeviceArray(A4s.n)
    phi4s = capi.DeviceArray.from_host(np.zeros(w4s.mesh.n_faces))
    dt4s = capi.DeviceArray.from_host(np.full(A4s.n, w4s.dt))
    u4s = capi.DeviceArray.from_host(np.zeros(A4s.n))
    rp4s, ci4s = A4s.pattern()
    vals4 = []
    for e in range(4):
        dm4s.assemble_momentum(A4s, rhs4s, 0, float(w4s.gammas[e]), 1.0, phi4s, d_dt=dt4s,
                               d_Uold=u4s)
        vals4.append(A4s.values())
    B4 = capi.Batch(A4s, 4)
    B4.set_values(np.stack(vals4))
    xb, _, _, _ = B4.solve(np.stack(w4s.bs[:4]), tolerance=1e-12, max_iterations=5000,
                           max_outer=40)
    e4 = [_rel(xb[e], oracle.cg(oracle.Matrix(rp4s, ci4s, vals4[e]), w4s.bs[e],
                                oracle.PC_JACOBI, 1e-13, 20000)[0]) for e in range(4)]
    errors["4"] = {"size": "32^3, 4 equations (config-4 generator)",
                   "batched_1e-12_vs_oracle_1e-13_max": max(e4)}

    # --- config 5: the reference adjoint's GMRES(30)-Jacobi on A^T
    # (src/adjoint/DiscreteAdjoint.cpp:54-146) on the 128^3 system, a bounded
    # sample; and the GPU adjoint against the oracle at 24^3
    wts = wl.uniform(wl.seed(5, wl.Q_WEIGHTS), 0.0, 1.0, n)
    g = 2.0 * wts * x128t
    with Pinned(core):
        t1 = time.perf_counter()
        _, ok5, it5, res5 = oracle.gmres_transpose(Ao, g, max_iter=60, restart=30, tol=1e-10)
        t5 = time.perf_counter() - t1
    configs["5"] = {"workload": "GMRES(30)-Jacobi on A^T, 128^3 (the reference adjoint), "
                                "a 60-iteration sample",
                    "ms_per_iteration": round(1e3 * t5 / max(it5, 1), 2),
                    "iterations_sampled": it5, "converged": ok5,
                    "residual_after_sample": res5,
                    "projected_s_for_its_500_iteration_cap": round(t5 / max(it5, 1) * 500, 1)}
    w5 = wl.poisson3d(24)
    m5 = w5.mesh
    A5, b5, pinned5 = _pressure_oracle(oracle, w5)
    w5t = wl.uniform(wl.seed(5, wl.Q_WEIGHTS), 0.0, 1.0, m5.n_cells)
    dm5 = capi.DeviceMesh(m5, w5.bc_of_zone)
    A5d = dm5.create_matrix()
    r5 = capi.DeviceArray(m5.n_cells)
    dm5.assemble_pressure(A5d, r5, d_gamma=capi.DeviceArray.from_host(w5.gamma))
    sf = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=1e-12,
                     max_iterations=5000, true_residual=1, max_outer=30)
    sf.setup(A5d)
    dx5 = capi.DeviceArray(m5.n_cells)
    sf.solve_device(capi.DeviceArray.from_host(b5), dx5)
    x5 = dx5.download()
    sa = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=1e-10,
                     max_iterations=5000, true_residual=1, max_outer=30)
    sa.setup(A5d)
    dl5 = capi.DeviceArray(m5.n_cells)
    sa.adjoint_solve_device(capi.DeviceArray.from_host(2.0 * w5t * x5), dl5)
    dg5 = capi.DeviceArray(m5.n_faces)
    dm5.face_gradient(dx5, dl5, dg5, pinned5)
    xo5, _, _, _ = oracle.cg(A5, b5, oracle.PC_JACOBI, 1e-13, 50000)
    # the pinned pressure matrix is symmetric: A^T = A
    lo5, _, _, _ = oracle.cg(A5, 2.0 * w5t * xo5, oracle.PC_JACOBI, 1e-13, 50000)
    go5 = oracle.face_gradient(m5, w5.bc_of_zone, xo5, lo5, pinned5)
    gg5 = dg5.download()
    errors["5"] = {"size": "24^3 (config-5 system)", "x": _rel(x5, xo5),
                   "lambda": _rel(dl5.download(), lo5),
                   "gradient_max_err_over_max": float(np.abs(gg5 - go5).max() / np.abs(go5).max())}

    cpu = {"value": round(value_ms, 1), "unit": "ms", "cores": 1, "kind": "port",
           "sample": f"the oracle's serial FP64 Jacobi-PCG (the reference's CG arithmetic) on the "
                     f"128^3 system, {done} iterations in {dt:.1f} s on one pinned core, times "
                     f"the {its_j} iterations its recurrence needs to reach 1e-8 (counted by the "
                     f"GPU FP64 Jacobi-PCG, which tracks it within +-2)",
           "core": core, "host": host, "configs": configs,
           "leg_s": round(time.perf_counter() - t_leg, 1)}
    log(f"cpu leg: {time.perf_counter() - t_leg:.1f} s")
    return cpu, errors


def measure_traffic(mode="--spmv-only", kernel="k_sell_spmv<double>", grid=None, last=None):
    """HBM bytes per launch of `kernel` from two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE) over `bench.py <mode>`, each scaled by what
    the same counter reads for the plain copy kernel of known size that run
    also launches; `grid` (work-items) keeps only launches of that size,
    `last` only the last `last` of them.  Runs the passes as child processes
    before this process touches the GPU; returns (bytes or None, details)."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if not prof:
        return None, {"source": "rocprofv3 not found"}
    out = {}
    base = tempfile.mkdtemp(prefix="simall_pmc_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = os.path.join(base, counter.lower())
        cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", counter, "--output-format", "csv",
               "-d", d, "-o", counter.lower(), "--", sys.executable, os.path.abspath(__file__),
               mode]
        t1 = time.perf_counter()
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True)
        log(f"traffic pass {counter}: rc {r.returncode} ({time.perf_counter() - t1:.1f}s)")
        if r.returncode != 0:
            return None, {"source": f"rocprofv3 --pmc {counter} exited {r.returncode}"}
        sums = {"spmv": [], "copy": []}
        launches = {}  # (kernel name, grid) -> counter values of the matching launches
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                for row in csv.DictReader(f):
                    name = row.get("Kernel_Name", "")
                    if row.get("Counter_Name") != counter:
                        continue
                    if kernel in name and (grid is None or int(row["Grid_Size"]) == grid):
                        order = row.get("Dispatch_Id") or row.get("Correlation_Id") or 0
                        launches.setdefault((name, row["Grid_Size"]), []).append(
                            (int(order), float(row["Counter_Value"])))
                    elif "k_axpby" in name:
                        sums["copy"].append(float(row["Counter_Value"]))
        if launches:
            # the shape launched most often: with grid=None, the autotuner's
            # pick (timed like every candidate, then used by the solve)
            key = max(launches, key=lambda k: len(launches[k]))
            vals = [v for _, v in sorted(launches[key])]
            if last:
                vals = vals[-last:]
            # launches that found every equation converged read almost
            # nothing: keep those within half of the largest reading
            top = max(vals)
            sums["spmv"] = [v for v in vals if v >= 0.5 * top]
            out.setdefault("picked", {})[counter] = {"kernel": key[0], "grid": int(key[1])}
        if not sums["spmv"] or not sums["copy"]:
            return None, {"source": f"no {counter} rows for the SpMV or the copy kernel"}
        out[counter] = {k: (sum(v) / len(v), len(v)) for k, v in sums.items()}
    shutil.rmtree(base, ignore_errors=True)
    copy_kib = 8 * COPY_DOUBLES / 1024.0
    sf = copy_kib / out["FETCH_SIZE"]["copy"][0]
    sw = copy_kib / out["WRITE_SIZE"]["copy"][0]
    fetch = out["FETCH_SIZE"]["spmv"][0] * sf * 1024.0
    write = out["WRITE_SIZE"]["spmv"][0] * sw * 1024.0
    det = {"source": f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over bench.py {mode}, "
                     "measured in this run",
           "fetch_bytes": round(fetch), "write_bytes": round(write),
           "fetch_scale": round(sf, 4), "write_scale": round(sw, 4),
           "spmv_dispatches": out["FETCH_SIZE"]["spmv"][1]}
    if "picked" in out:
        det["kernel"] = out["picked"]["FETCH_SIZE"]["kernel"]
        det["grid"] = out["picked"]["FETCH_SIZE"]["grid"]
    return fetch + write, det


def bench_momentum(capi, wl, lib, steps, log, N=96):
    """Nonsymmetric momentum system on a 96^3 hex mesh (884,736 cells): upwind
    convection of the Taylor-Green field plus diffusion, pseudo-transient
    diagonal, assembled on the GPU; ILU(0)-BiCGStab to 1e-8 true residual in
    FP64 and in the mixed form (FP32 inner, FP64 refinement); and GMRES(20)-ILU(0)
    as the turbulence transport equations configure it."""
    import numpy as np

    t0 = time.perf_counter()
    w = wl.momentum3d(N)
    m = w.mesh
    dm = capi.DeviceMesh(m, w.bc_of_zone)
    A = dm.create_matrix()
    rhs = capi.DeviceArray(m.n_cells)
    dU = [capi.DeviceArray.from_host(u) for u in w.U]
    phi = capi.DeviceArray(m.n_faces)
    dm.face_flux(dU[0], dU[1], dU[2], phi, rho=w.rho)
    dt = capi.DeviceArray.from_host(w.dt)
    dm.assemble_momentum(A, rhs, 0, w.mu, w.rho, phi, d_dt=dt, d_Uold=dU[0], d_Ucomp=dU[0],
                         urf=w.urf)
    db = capi.DeviceArray.from_host(w.b)
    dx = capi.DeviceArray(m.n_cells)
    log(f"momentum {N}^3: n={A.n} nnz={A.nnz} ({time.perf_counter() - t0:.1f}s)")
    out = {"workload": f"{N}^3 Taylor-Green upwind+diffusion momentum, pseudo-transient diagonal, "
                       f"ILU(0)-BiCGStab to 1e-8 true residual", "n": A.n, "nnz": A.nnz}

    def run(cfg, prec):
        s = capi.Solver(precision=prec, **cfg)
        s.setup(A)
        st, rc = s.solve_device(db, dx)
        times = []
        for _ in range(steps):
            lib.mi_device_synchronize()
            t1 = time.perf_counter()
            st, rc = s.solve_device(db, dx)
            lib.mi_device_synchronize()
            times.append(time.perf_counter() - t1)
        d = st.as_dict()
        return {"ms": round(1e3 * float(np.median(times)), 3), "iterations": d["iterations"],
                "outer": d["outer_iterations"], "relative_residual": d["relative_residual"],
                "true_relative_residual": d["true_relative_residual"],
                "setup_s": round(d["setup_seconds"], 3), "status": rc}

    bicg = dict(kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_ILU, tolerance=1e-8,
                max_iterations=2000, true_residual=1)
    # the transport-equation configuration of the turbulence models
    # (src/turbulence/KOmegaSST.cpp:32-37): GMRES(20), ILU(0), 1e-6, 100 iterations
    gmres = dict(kind=capi.MI_SOLVER_GMRES, preconditioner=capi.MI_PC_ILU, tolerance=1e-6,
                 max_iterations=100, restart=20)
    out["gmres"] = {"workload": "GMRES(20)-ILU(0) to 1e-6 preconditioned residual, "
                                "100 iterations at most"}
    for name, prec in (("fp64", capi.MI_PRECISION_FP64), ("mixed", capi.MI_PRECISION_MIXED)):
        out[name] = run(bicg, prec)
        out["gmres"][name] = run(gmres, prec)
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    out["gmres"]["speedup_mixed_vs_fp64"] = round(out["gmres"]["fp64"]["ms"] /
                                                  out["gmres"]["mixed"]["ms"], 3)
    # The three momentum components: one matrix, three right-hand sides
    # (SimpleAlgorithm.cpp:529-532), batched on one shared matrix with one
    # device ILU(0) factorization, against three mixed ILU(0)-BiCGStab solves.
    n = m.n_cells
    bs3 = [w.b] + [wl.normal(90 + k, n) for k in range(2)]
    db3 = capi.DeviceArray.from_host(np.concatenate(bs3))
    dx3 = capi.DeviceArray(3 * n)
    B3 = capi.Batch(A, 3, flags=capi.MI_BATCH_SHARED_VALUES | capi.MI_BATCH_ILU0)
    B3.set_matrix(A)
    kw = dict(kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_ILU, tolerance=1e-8,
              max_iterations=2000)
    B3.solve_device(db3, dx3, **kw)
    times = []
    for _ in range(steps):
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        its3, rel3, rc3 = B3.solve_device(db3, dx3, **kw)
        times.append(time.perf_counter() - t1)
    s = capi.Solver(precision=capi.MI_PRECISION_MIXED, **bicg)
    s.setup(A)
    one, its1 = 0.0, []
    for k in range(3):
        dbk = capi.DeviceArray.from_host(bs3[k])
        s.solve_device(dbk, dx)
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        st, _ = s.solve_device(dbk, dx)
        lib.mi_device_synchronize()
        one += time.perf_counter() - t1
        its1.append(st.as_dict()["iterations"])
    bms = 1e3 * float(np.median(times))
    out["three_components_batched"] = {
        "workload": "3 right-hand sides on the one momentum matrix: mi_batch with shared values "
                    "and a device ILU(0) (two-colour ordering), BiCGStab FP32 inner / FP64 "
                    "refinement to 1e-8, against 3 mixed ILU(0)-BiCGStab solves one at a time",
        "ms": round(bms, 3), "one_at_a_time_ms": round(1e3 * one, 3),
        "speedup": round(1e3 * one / bms, 3), "iterations": [int(i) for i in its3],
        "one_at_a_time_iterations": its1, "max_true_relative_residual": float(rel3.max()),
        "status": rc3}
    B3.close()
    log(f"momentum: {out}")
    return out


def bench_flow(capi, wl, lib, log, N=96, iters=3, variants=None):
    """Device-resident SIMPLE outer iterations and PISO time steps on a 96^3
    channel (velocity inlet, pressure outlet, walls; rho 1, mu 1e-3): per
    iteration the face fluxes, three momentum assemblies and GMRES(30)-ILU(0)
    solves, the pressure assembly, its AMG-PCG solve and the correction.  With
    the reference's preconditioner rebuild at every solve, and keeping the
    preconditioners across iterations.  `simple` and `piso` set up the AMG
    hierarchy for every pressure solve by the device numeric re-setup
    (amg_resetup = 1, the default); `simple_full_setup` by a full host setup
    each time (amg_resetup = 0, the reference's semantics);
    `simple_keep_pc` keeps the first hierarchy; `simple_batched_momentum` solves
    the three momentum components in one batched BiCGStab-ILU(0) solve on the
    shared matrix (momentum_batch = 1)."""
    import numpy as np

    t0 = time.perf_counter()
    m = wl.Mesh(N, N, N)
    bc = np.array([0, wl.BC_VELOCITY_INLET, wl.BC_PRESSURE] + [wl.BC_WALL] * 4, np.int32)
    bcv = np.zeros((7, 3))
    bcv[1] = [1.0, 0.0, 0.0]
    dm = capi.DeviceMesh(m, bc)
    log(f"flow {N}^3: n={m.n_cells} ({time.perf_counter() - t0:.1f}s)")
    out = {"workload": f"{N}^3 channel, velocity inlet / pressure outlet / walls, rho 1, mu 1e-3; "
                       f"momentum GMRES(30)-ILU(0) 1e-8, pressure CG-AMG 1e-8",
           "n": m.n_cells}
    for name, alg, keep, resetup, mbatch in (("simple", capi.MI_FLOW_SIMPLE, 0, 1, 0),
                                             ("simple_full_setup", capi.MI_FLOW_SIMPLE, 0, 0, 0),
                                             ("simple_keep_pc", capi.MI_FLOW_SIMPLE, 1, 1, 0),
                                             ("simple_batched_momentum", capi.MI_FLOW_SIMPLE, 0, 1,
                                              1),
                                             ("piso", capi.MI_FLOW_PISO, 0, 1, 0)):
        if variants and name not in variants:
            continue
        cfg = capi.flow_cfg(alg, momentum=dict(keep_preconditioner=keep),
                            pressure=dict(keep_preconditioner=keep, amg_resetup=resetup),
                            momentum_batch=mbatch)
        F = capi.Flow(dm, cfg, bcv)
        times, res = [], None
        for it in range(iters + 1):  # the first iteration builds the patterns
            lib.mi_device_synchronize()
            t1 = time.perf_counter()
            res = F.advance_time_step(1) if alg == capi.MI_FLOW_PISO else F.iterate()
            lib.mi_device_synchronize()
            if it > 0:
                times.append(time.perf_counter() - t1)
        out[name] = {"ms_per_iteration": round(1e3 * float(np.median(times)), 1),
                     "iterations_timed": iters, "last_residuals": res}
        F.close()
    log(f"flow: {out}")
    return out


# The batched CG SpMM (per-equation values, SELL-64 order); the counter
# passes keep the shape launched most often, the autotuner's pick.
BSPMM_KERNEL = "k_bspmm<3, true, "
BSPMM_GRID = None


def transport_batch(capi, wl, N=100, n_eq=32):
    """The 100^3 perturbed, permuted transport mesh and a batch of n_eq
    equations A_e = V/dt I + Gamma_e L assembled on the GPU."""
    import numpy as np

    w = wl.transport(N, n_eq)
    m = w.mesh
    dm = capi.DeviceMesh(m, w.bc_of_zone)
    A = dm.create_matrix()
    rhs = capi.DeviceArray(m.n_cells)
    phi = capi.DeviceArray.from_host(np.zeros(m.n_faces))
    dt = capi.DeviceArray.from_host(np.full(m.n_cells, w.dt))
    uold = capi.DeviceArray.from_host(np.zeros(m.n_cells))
    B = capi.Batch(A, n_eq)
    for e in range(n_eq):
        dm.assemble_momentum(A, rhs, 0, float(w.gammas[e]), 1.0, phi, d_dt=dt, d_Uold=uold)
        B.set_equation(e, A)
    db = capi.DeviceArray.from_host(np.concatenate(w.bs))
    return w, dm, A, B, db, (rhs, phi, dt, uold)


def batch_spmm_only(capi, wl, lib):
    """Short run for the batched SpMM's counter passes: the autotuner's
    launches, three batched CG iterations on the real values, then the
    calibration copy."""
    w, dm, A, B, db, keep = transport_batch(capi, wl)
    n = w.mesh.n_cells
    dx = capi.DeviceArray(32 * n)
    # three iterations in one chunk: no launch runs after the iteration limit
    B.solve_device(db, dx, tolerance=1e-8, max_iterations=3, inner_tolerance=1e-5,
                   check_interval=3)
    t = B.tuning()
    src = capi.DeviceArray(COPY_DOUBLES)
    dst = capi.DeviceArray(COPY_DOUBLES)
    for _ in range(3):
        capi.check(lib.mi_copy(src.ptr, dst.ptr, COPY_DOUBLES, None))
    lib.mi_device_synchronize()
    return {"bspmm_us": t["spmm_us"], "bspmm_algorithmic_bytes": t["spmm_bytes"],
            "rows_per_block": t["rows_per_block"], "copy_bytes_read": 8 * COPY_DOUBLES,
            "copy_bytes_written": 8 * COPY_DOUBLES}


def bench_batch(capi, wl, lib, steps, log, N=100, n_eq=32, traffic=None):
    """32 transport equations A_e = V/dt I + Gamma_e L on one 100^3 perturbed,
    permuted mesh (1,000,000 cells): one batched solve of all equations to
    1e-8 true residual, against 32 one-at-a-time solves of the same arithmetic
    (FP32 Jacobi-PCG inner, FP64 refinement) with the single-equation solver."""
    import numpy as np

    t0 = time.perf_counter()
    w, dm, A, B, db, (rhs, phi, dt, uold) = transport_batch(capi, wl, N, n_eq)
    n = w.mesh.n_cells
    dx = capi.DeviceArray(n_eq * n)
    log(f"batched transport: {n_eq} equations, n={n} nnz={A.nnz} ({time.perf_counter() - t0:.1f}s)")
    # Jacobi-PCG on these mass-dominated systems prefers a tighter inner
    # tolerance than the AMG pressure solve; both paths below use the same
    kw = dict(tolerance=1e-8, max_iterations=2000, inner_tolerance=1e-5)
    B.solve_device(db, dx, **kw)
    times = []
    for _ in range(steps):
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        its, rel, rc = B.solve_device(db, dx, **kw)
        times.append(time.perf_counter() - t1)
    batched_ms = 1e3 * float(np.median(times))
    # one equation at a time, same arithmetic
    s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-8,
                    max_iterations=2000, precision=capi.MI_PRECISION_MIXED, true_residual=1,
                    inner_tolerance=1e-5)
    single = 0.0
    its1 = []
    for e in range(n_eq):
        dm.assemble_momentum(A, rhs, 0, float(w.gammas[e]), 1.0, phi, d_dt=dt, d_Uold=uold)
        s.setup(A)
        dbe = capi.DeviceArray.from_host(w.bs[e])
        dxe = capi.DeviceArray(n)
        s.solve_device(dbe, dxe)  # builds the preconditioner for these values
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        st, rc1 = s.solve_device(dbe, dxe)
        lib.mi_device_synchronize()
        single += time.perf_counter() - t1
        its1.append(st.as_dict()["iterations"])
    tune = B.tuning()
    # cost of a grid-wide barrier inside one persistent kernel vs a kernel boundary
    gbar = {}
    for g in (256, 1024, 2048):
        sync_us, launch_us = capi.time_grid_barrier(g, 200)
        gbar[str(g)] = {"grid_sync_us": round(sync_us, 2), "kernel_boundary_us": round(launch_us, 2)}
    out = {"workload": f"{n_eq} transport equations V/dt I + Gamma_e L, Gamma_e = 10^U[-6,-3], "
                       f"{N}^3 perturbed+permuted hex mesh",
           "n": n, "nnz": A.nnz, "equations": n_eq, "ms": round(batched_ms, 3),
           "one_at_a_time_ms": round(1e3 * single, 3),
           "speedup": round(1e3 * single / batched_ms, 3),
           "iterations_min_max": [int(its.min()), int(its.max())],
           "one_at_a_time_iterations_min_max": [int(min(its1)), int(max(its1))],
           "max_true_relative_residual": float(rel.max()), "status": rc,
           "rows_per_block": tune["rows_per_block"],
           "spmm_us": round(tune["spmm_us"], 2),
           "spmm_GBps": round(tune["spmm_bytes"] / tune["spmm_us"] / 1e3, 1),
           "spmm_frac_of_peak": round(tune["spmm_bytes"] / tune["spmm_us"] / 1e3 / HBM_PEAK_GBS, 4),
           "spmm_algorithmic_bytes": round(tune["spmm_bytes"]),
           "tuned_shape": {k: tune[k] for k in ("slots_per_thread_max", "spmm_workgroups",
                                                "banded", "vector_workgroups")},
           "vector_stage_us": round(tune["vector_us"], 2),
           "grid_barrier_vs_kernel_boundary": gbar}
    if traffic is not None:
        out["spmm_traffic"] = round(traffic[0]) if traffic[0] else None
        out["spmm_traffic_detail"] = traffic[1]
    log(f"batched: {out}")
    return out


def bench_batch_convective(capi, wl, lib, steps, log, N=96, n_eq=8):
    """8 convection-diffusion transport equations (Taylor-Green mass flux,
    pseudo-transient diagonal, diffusivity 10^U[-3,-1] per equation) on one
    96^3 mesh: one batched Jacobi-BiCGStab solve of all equations to 1e-8
    true residual against one-at-a-time solves of the same arithmetic (FP32
    Jacobi-BiCGStab inner, FP64 refinement) with the single-equation solver."""
    import numpy as np

    t0 = time.perf_counter()
    L = 2 * np.pi
    m = wl.Mesh(N, N, N, extent=(L, L, L), perturb=0.1, node_seed=3)
    bc = [wl.BC_WALL] * 7
    dm = capi.DeviceMesh(m, bc)
    A = dm.create_matrix()
    rhs = capi.DeviceArray(m.n_cells)
    U = [capi.DeviceArray.from_host(u) for u in wl.taylor_green(m)]
    phi = capi.DeviceArray(m.n_faces)
    dm.face_flux(U[0], U[1], U[2], phi)
    dt = capi.DeviceArray.from_host(wl.loguniform(3, -1, 0, m.n_cells))
    uold = capi.DeviceArray.from_host(np.zeros(m.n_cells))
    gam = wl.loguniform(4, -3, -1, n_eq)
    B = capi.Batch(A, n_eq)
    for e in range(n_eq):
        dm.assemble_momentum(A, rhs, 0, float(gam[e]), 1.0, phi, d_dt=dt, d_Uold=uold)
        B.set_equation(e, A)
    n = m.n_cells
    bs = [wl.normal(20 + e, n) for e in range(n_eq)]
    db = capi.DeviceArray.from_host(np.concatenate(bs))
    dx = capi.DeviceArray(n_eq * n)
    log(f"batched convective transport: {n_eq} equations, n={n} ({time.perf_counter() - t0:.1f}s)")
    kw = dict(kind=capi.MI_SOLVER_BICGSTAB, tolerance=1e-8, max_iterations=2000,
              inner_tolerance=1e-5)
    B.solve_device(db, dx, **kw)
    times = []
    for _ in range(steps):
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        its, rel, rc = B.solve_device(db, dx, **kw)
        times.append(time.perf_counter() - t1)
    batched_ms = 1e3 * float(np.median(times))
    s = capi.Solver(kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-8,
                    max_iterations=2000, precision=capi.MI_PRECISION_MIXED, true_residual=1,
                    inner_tolerance=1e-5)
    single, its1 = 0.0, []
    for e in range(n_eq):
        dm.assemble_momentum(A, rhs, 0, float(gam[e]), 1.0, phi, d_dt=dt, d_Uold=uold)
        s.setup(A)
        dbe = capi.DeviceArray.from_host(bs[e])
        dxe = capi.DeviceArray(n)
        s.solve_device(dbe, dxe)
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        st, rc1 = s.solve_device(dbe, dxe)
        lib.mi_device_synchronize()
        single += time.perf_counter() - t1
        its1.append(st.as_dict()["iterations"])
    out = {"workload": f"{n_eq} convection-diffusion transport equations (Taylor-Green flux), "
                       f"{N}^3 perturbed hex mesh, Jacobi-BiCGStab to 1e-8",
           "n": n, "equations": n_eq, "ms": round(batched_ms, 3),
           "one_at_a_time_ms": round(1e3 * single, 3),
           "speedup": round(1e3 * single / batched_ms, 3),
           "iterations_min_max": [int(its.min()), int(its.max())],
           "one_at_a_time_iterations_min_max": [int(min(its1)), int(max(its1))],
           "max_true_relative_residual": float(rel.max()), "status": rc}
    # the same equations with ILU(0): factorized per equation on the device
    # (two-colour ordering), against one-at-a-time mixed ILU(0)-BiCGStab
    Bi = capi.Batch(A, n_eq, flags=capi.MI_BATCH_ILU0)
    for e in range(n_eq):
        dm.assemble_momentum(A, rhs, 0, float(gam[e]), 1.0, phi, d_dt=dt, d_Uold=uold)
        Bi.set_equation(e, A)
    kwi = dict(kw, preconditioner=capi.MI_PC_ILU)
    t1 = time.perf_counter()
    Bi.solve_device(db, dx, **kwi)  # includes the device factorization of the 8 equations
    first_ms = 1e3 * (time.perf_counter() - t1)
    times = []
    for _ in range(steps):
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        itsi, reli, rci = Bi.solve_device(db, dx, **kwi)
        times.append(time.perf_counter() - t1)
    ilu_ms = 1e3 * float(np.median(times))
    si = capi.Solver(kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_ILU, tolerance=1e-8,
                     max_iterations=2000, precision=capi.MI_PRECISION_MIXED, true_residual=1,
                     inner_tolerance=1e-5)
    single, itsi1 = 0.0, []
    for e in range(n_eq):
        dm.assemble_momentum(A, rhs, 0, float(gam[e]), 1.0, phi, d_dt=dt, d_Uold=uold)
        si.setup(A)
        dbe = capi.DeviceArray.from_host(bs[e])
        dxe = capi.DeviceArray(n)
        si.solve_device(dbe, dxe)
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        st, rc1 = si.solve_device(dbe, dxe)
        lib.mi_device_synchronize()
        single += time.perf_counter() - t1
        itsi1.append(st.as_dict()["iterations"])
    # GMRES(20)-ILU(0), the turbulence models' transport configuration
    # (src/turbulence/KOmegaSST.cpp:32-37), batched, to 1e-8 true residual
    kwg = dict(kind=capi.MI_SOLVER_GMRES, preconditioner=capi.MI_PC_ILU, restart=20,
               tolerance=1e-8, max_iterations=2000, inner_tolerance=1e-5)
    Bi.solve_device(db, dx, **kwg)
    times = []
    for _ in range(steps):
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        itsg, relg, rcg = Bi.solve_device(db, dx, **kwg)
        times.append(time.perf_counter() - t1)
    out["gmres20_ilu0"] = {"workload": "the same equations, GMRES(20) with ILU(0), one cycle per "
                                       "FP64 refinement sweep",
                           "ms": round(1e3 * float(np.median(times)), 3),
                           "arnoldi_steps_min_max": [int(itsg.min()), int(itsg.max())],
                           "max_true_relative_residual": float(relg.max()), "status": rcg}
    out["ilu0"] = {"workload": "the same equations, BiCGStab with ILU(0) factorized per equation "
                               "on the device",
                   "ms": round(ilu_ms, 3), "first_solve_with_factorization_ms": round(first_ms, 3),
                   "one_at_a_time_ms": round(1e3 * single, 3),
                   "speedup": round(1e3 * single / ilu_ms, 3),
                   "iterations_min_max": [int(itsi.min()), int(itsi.max())],
                   "one_at_a_time_iterations_min_max": [int(min(itsi1)), int(max(itsi1))],
                   "max_true_relative_residual": float(reli.max()), "status": rci}
    Bi.close()
    log(f"batched convective: {out}")
    return out


def bench_pressure_batched(capi, wl, lib, A, pinned, single, steps, log, n_eq=8):
    """n_eq right-hand sides on the headline pressure matrix: one batched
    AMG-PCG solve (mi_batch with shared values: one SA-AMG hierarchy, each
    V-cycle launch streams a level's operators once for every right-hand
    side) to 1e-8 true residual, against the headline mixed solver's
    one-at-a-time solves of the same right-hand sides."""
    import numpy as np

    n = A.n
    bs = []
    for e in range(n_eq):
        b = wl.remove_mean(wl.normal(200 + e, n))
        if pinned:
            b[0] = 0.0
        bs.append(b)
    db = capi.DeviceArray.from_host(np.concatenate(bs))
    dx = capi.DeviceArray(n_eq * n)
    B = capi.Batch(A, n_eq, flags=capi.MI_BATCH_SHARED_VALUES)
    B.set_matrix(A)
    kw = dict(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=1e-8,
              max_iterations=500)
    lib.mi_device_synchronize()
    t1 = time.perf_counter()
    B.solve_device(db, dx, **kw)  # includes the host hierarchy setup
    first_ms = 1e3 * (time.perf_counter() - t1)
    times = []
    for _ in range(steps):
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        its, rel, rc = B.solve_device(db, dx, **kw)
        times.append(time.perf_counter() - t1)
    batched_ms = 1e3 * float(np.median(times))
    one, its1 = 0.0, []
    for e in range(n_eq):
        dbe = capi.DeviceArray.from_host(bs[e])
        dxe = capi.DeviceArray(n)
        lib.mi_device_synchronize()
        t1 = time.perf_counter()
        st, rc1 = single.solve_device(dbe, dxe)
        lib.mi_device_synchronize()
        one += time.perf_counter() - t1
        its1.append(st.as_dict()["iterations"])
    B.close()
    out = {"workload": f"{n_eq} right-hand sides on the {n}-row pressure matrix, CG with one "
                       f"shared SA-AMG hierarchy, FP32 inner / FP64 refinement, to 1e-8",
           "equations": n_eq, "ms": round(batched_ms, 3),
           "first_solve_with_setup_ms": round(first_ms, 3),
           "one_at_a_time_ms": round(1e3 * one, 3),
           "speedup": round(1e3 * one / batched_ms, 3),
           "iterations_min_max": [int(its.min()), int(its.max())],
           "one_at_a_time_iterations_min_max": [int(min(its1)), int(max(its1))],
           "max_true_relative_residual": float(rel.max()), "status": rc}
    log(f"batched pressure: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="cells per axis")
    ap.add_argument("--pc", default="amg", choices=["jacobi", "amg"])
    ap.add_argument("--cpu-iters", type=int, default=400,
                    help="oracle Jacobi-PCG iterations timed for the 128^3 CPU baseline")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline leg and the comparisons with the oracle")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure HBM traffic")
    ap.add_argument("--modes", default="fp64,mixed", help="solver precisions to time")
    ap.add_argument("--degree", type=int, default=0, help="AMG smoother degree (0: default)")
    ap.add_argument("--coarse-degree", type=int, default=0,
                    help="AMG smoother degree below the finest level (0: default)")
    ap.add_argument("--inner-tol", type=float, default=0.0)
    ap.add_argument("--strength", type=float, default=-1.0,
                    help="AMG aggregation strength threshold (negative: the solver default)")
    ap.add_argument("--truncation", type=float, default=-1.0,
                    help="AMG prolongator truncation (negative: the solver default)")
    ap.add_argument("--caller-order", action="store_true",
                    help="also time SpMV on the caller's (randomly numbered) matrix")
    ap.add_argument("--no-flow", action="store_true", help="skip the SIMPLE / PISO iteration section")
    ap.add_argument("--no-momentum", action="store_true",
                    help="skip the ILU(0)-BiCGStab momentum measurement (96^3)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched transport measurement (32 equations on a 100^3 mesh)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the discrete-adjoint timing (J = sum w x^2 on the same system)")
    ap.add_argument("--flow-only", action="store_true",
                    help="run the SIMPLE / PISO iteration section alone (short run for profiles)")
    ap.add_argument("--flow-n", type=int, default=96, help="cells per axis of the flow section")
    ap.add_argument("--flow-iters", type=int, default=3, help="timed iterations per flow variant")
    ap.add_argument("--flow-variants", default="",
                    help="comma list of simple, simple_full_setup, simple_keep_pc, "
                         "simple_batched_momentum, piso (default all)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="time the working-matrix SpMV and a calibration copy only "
                         "(short run for hardware-counter passes)")
    ap.add_argument("--pressure-batch-only", action="store_true",
                    help="assemble the pressure matrix, then only the batched AMG-PCG solve of 8 "
                         "right-hand sides against the mixed solver one at a time (for profiling)")
    ap.add_argument("--batch-spmm-only", action="store_true",
                    help="build the 32-equation transport batch, run three batched CG "
                         "iterations and a calibration copy only (hardware-counter passes)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from simall_amd import capi
    from simall_amd import workload as wl

    if args.flow_only:
        lib = capi.load()
        capi.check(lib.mi_init(0))
        out = bench_flow(capi, wl, lib, log, N=args.flow_n, iters=args.flow_iters,
                         variants=[v for v in args.flow_variants.split(",") if v])
        print(json.dumps({"flow": out}), flush=True)
        return

    if args.batch_spmm_only:
        lib = capi.load()
        capi.check(lib.mi_init(0))
        print(json.dumps(batch_spmm_only(capi, wl, lib)), flush=True)
        return

    traffic, traffic_detail = None, {"source": "not measured (--no-traffic)"}
    batch_traffic = None
    if not args.spmv_only and not args.no_traffic and not args.no_batch:
        # the solve's three iterations (the autotuner times the same kernel
        # at the same size with the other row walk before them)
        batch_traffic = measure_traffic("--batch-spmm-only", BSPMM_KERNEL, BSPMM_GRID, last=3)
        log(f"batched SpMM traffic: {batch_traffic}")
    if not args.spmv_only and not args.no_traffic:
        # child processes, before this process initialises the GPU
        traffic, traffic_detail = measure_traffic()
        log(f"traffic: {traffic} {traffic_detail}")

    lib = capi.load()
    capi.check(lib.mi_init(0))
    t0 = time.perf_counter()
    w = wl.poisson3d(args.n)
    m = w.mesh
    log(f"mesh {args.n}^3: {m.n_cells} cells, {m.n_faces} faces ({time.perf_counter() - t0:.1f}s)")
    dm = capi.DeviceMesh(m, w.bc_of_zone)
    A = dm.create_matrix()
    d_rhs = capi.DeviceArray(m.n_cells)
    d_gamma = capi.DeviceArray.from_host(w.gamma)
    pinned = dm.assemble_pressure(A, d_rhs, d_gamma=d_gamma)
    b = w.b.copy()
    if pinned:
        b[0] = 0.0
    n, nnz = A.n, A.nnz
    log(f"assembled on GPU: n={n} nnz={nnz} pinned={pinned} ({time.perf_counter() - t0:.1f}s)")
    d_b = capi.DeviceArray.from_host(b)
    d_x = capi.DeviceArray(n)
    pc = capi.MI_PC_AMG if args.pc == "amg" else capi.MI_PC_JACOBI
    if args.spmv_only:
        s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-8,
                        max_iterations=10, precision=capi.MI_PRECISION_FP64)
        s.setup(A)
        us64, by64 = s.time_spmv(capi.MI_PRECISION_FP64, reps=10)
        src = capi.DeviceArray(COPY_DOUBLES)
        dst = capi.DeviceArray(COPY_DOUBLES)
        for _ in range(3):
            capi.check(lib.mi_copy(src.ptr, dst.ptr, COPY_DOUBLES, None))
        lib.mi_device_synchronize()
        print(json.dumps({"spmv_fp64_us": us64, "spmv_fp64_algorithmic_bytes": by64,
                          "copy_bytes_read": 8 * COPY_DOUBLES,
                          "copy_bytes_written": 8 * COPY_DOUBLES}), flush=True)
        return

    def make_solver(precision):
        kw = {}
        if args.degree:
            kw["amg_sweeps"] = args.degree
        if args.coarse_degree:
            kw["amg_coarse_sweeps"] = args.coarse_degree
        if args.strength >= 0:
            kw["amg_strength"] = args.strength
        if args.truncation >= 0:
            kw["amg_truncation"] = args.truncation
        if args.inner_tol:
            kw["inner_tolerance"] = args.inner_tol
        s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=pc, tolerance=1e-8,
                        max_iterations=20000, precision=precision, true_residual=1, **kw)
        s.setup(A)
        return s

    if args.pressure_batch_only:
        out = bench_pressure_batched(capi, wl, lib, A, pinned,
                                     make_solver(capi.MI_PRECISION_MIXED), args.steps, log)
        print(json.dumps(out), flush=True)
        return

    results, solvers = {}, {}
    modes = [m for m in (("fp64", capi.MI_PRECISION_FP64), ("mixed", capi.MI_PRECISION_MIXED))
             if m[0] in args.modes.split(",")]
    for name, prec in modes:
        s = make_solver(prec)
        solvers[name] = s
        for _ in range(args.warmup):
            st, rc = s.solve_device(d_b, d_x)
        lib.mi_device_synchronize()
        times, stats = [], None
        for _ in range(args.steps):
            lib.mi_device_synchronize()
            t1 = time.perf_counter()
            st, rc = s.solve_device(d_b, d_x)
            lib.mi_device_synchronize()
            times.append(time.perf_counter() - t1)
            stats = st.as_dict()
        x = d_x.download()
        results[name] = dict(ms=1e3 * float(np.median(times)), total_s=float(sum(times)),
                             stats=stats, rc=rc, x=x)
        log(f"{name}: {results[name]['ms']:.2f} ms/solve, its={stats['iterations']} "
            f"outer={stats['outer_iterations']} rel={stats['true_relative_residual']:.3e} "
            f"setup={stats['setup_seconds']:.3f}s rc={rc}")

    # SpMV bandwidth on the solver's working (reordered) matrix, FP64 and the
    # FP32 inner form; and on the caller's (randomly numbered) matrix
    ws = next(iter(solvers.values()))
    us64, by64 = ws.time_spmv(capi.MI_PRECISION_FP64, reps=50)
    us32, by32 = ws.time_spmv(capi.MI_PRECISION_MIXED, reps=50)
    gbs64, gbs32 = by64 / us64 / 1e3, by32 / us32 / 1e3
    # the matrix (217 MB) fits in the 256 MB Infinity Cache, which a back-to-back
    # loop leaves warm: the roofline figure rotates over copies of the matrix
    # (more than 1 GiB in all), so every launch streams it from HBM
    us64c, _ = ws.time_spmv_cold(capi.MI_PRECISION_FP64, reps=20)
    us32c, _ = ws.time_spmv_cold(capi.MI_PRECISION_MIXED, reps=20)
    gbs64c, gbs32c = by64 / us64c / 1e3, by32 / us32c / 1e3
    gbs_raw = None
    if args.caller_order:
        dx = capi.DeviceArray.from_host(np.ones(n))
        dy = capi.DeviceArray(n)
        us_raw = A.time_spmv(dx, dy, capi.MI_PRECISION_FP64, reps=20)
        gbs_raw = round(spmv_bytes(n, nnz, 8) / us_raw / 1e3, 1)
    info = ws.info()
    log(f"SpMV (reordered, bandwidth {info['bandwidth_before']} -> {info['bandwidth_after']}) "
        f"fp64 {us64:.1f} us {gbs64:.0f} GB/s ({100 * gbs64 / HBM_PEAK_GBS:.1f}%), cold {us64c:.1f} us "
        f"{gbs64c:.0f} GB/s ({100 * gbs64c / HBM_PEAK_GBS:.1f}%); fp32 {us32:.1f} us {gbs32:.0f} GB/s, "
        f"cold {us32c:.1f} us; caller ordering fp64 {gbs_raw} GB/s; "
        f"AMG levels {info['levels']} complexity {info['operator_complexity']:.2f}")
    adjoint = None
    if "fp64" in solvers and not args.no_adjoint:
        # discrete adjoint of the FP64 solve: A^T lambda = 2 w x, then dJ/dgamma_f per face
        s64 = solvers["fp64"]
        wts = wl.uniform(wl.seed(5, wl.Q_WEIGHTS), 0.0, 1.0, n)
        d_g = capi.DeviceArray.from_host(2.0 * wts * results["fp64"]["x"])
        d_l = capi.DeviceArray(n)
        d_xs = capi.DeviceArray.from_host(results["fp64"]["x"])
        d_grad = capi.DeviceArray(m.n_faces)
        t1 = time.perf_counter()
        st, rc = s64.adjoint_solve_device(d_g, d_l)
        lib.mi_device_synchronize()
        first = time.perf_counter() - t1
        ta = []
        for _ in range(args.steps):
            t1 = time.perf_counter()
            st, rc = s64.adjoint_solve_device(d_g, d_l)
            lib.mi_device_synchronize()
            ta.append(time.perf_counter() - t1)
        tg = []
        for _ in range(5):
            t1 = time.perf_counter()
            dm.face_gradient(d_xs, d_l, d_grad, pinned)
            tg.append(time.perf_counter() - t1)
        sa = st.as_dict()
        adjoint = {"solve_ms": round(1e3 * float(np.median(ta)), 3),
                   "first_call_s": round(first, 3), "iterations": sa["iterations"],
                   "true_relative_residual": sa["true_relative_residual"],
                   "face_gradient_ms": round(1e3 * float(np.median(tg)), 3),
                   "faces": m.n_faces}
        log(f"adjoint: {adjoint}")

    pressure_batched = None
    if not args.no_batch and "mixed" in solvers:
        pressure_batched = bench_pressure_batched(capi, wl, lib, A, pinned, solvers["mixed"],
                                                  args.steps, log)

    momentum = None
    if not args.no_momentum:
        momentum = bench_momentum(capi, wl, lib, args.steps, log)

    flow = None
    if not args.no_flow:
        flow = bench_flow(capi, wl, lib, log)

    batch = None
    if not args.no_batch:
        batch = bench_batch(capi, wl, lib, args.steps, log, traffic=batch_traffic)
        batch["convective"] = bench_batch_convective(capi, wl, lib, args.steps, log)
        batch["pressure_amg"] = pressure_batched

    err = None
    if "fp64" in results and "mixed" in results:
        x64, xm = results["fp64"]["x"], results["mixed"]["x"]
        err = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))

    cpu, errors = None, None
    if not args.no_cpu and "mixed" in results:
        host, aff = host_info()
        cpu, errors = cpu_leg(capi, wl, lib, A, b, results["mixed"]["x"], d_b, d_x,
                              args.cpu_iters, aff[0], host)
        log(f"cpu baseline: {cpu}")
        log(f"error vs oracle: {errors}")

    mixed_ms = results["mixed"]["ms"] if "mixed" in results else float("nan")
    fp64_ms = results["fp64"]["ms"] if "fp64" in results else float("nan")
    get = lambda m, k: results[m]["stats"][k] if m in results else None
    out = {
        "metric": "pressure-solve time to 1e-8 rel residual, mixed vs FP64 GPU; SpMV HBM GB/s",
        "value": round(mixed_ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed_ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 inner / fp64 refinement",
        "data": "synthetic",
        "config": {
            "workload": f"{args.n}^3 perturbed+permuted hex pressure Poisson, log-uniform "
                        f"conductance, {args.pc}-PCG to 1e-8 true residual",
            "n": n, "nnz": nnz, "preconditioner": args.pc,
        },
        "fp64_ms": round(fp64_ms, 3),
        "speedup_mixed_vs_fp64": round(fp64_ms / mixed_ms, 3),
        "iterations": {"fp64": get("fp64", "iterations"), "mixed_inner": get("mixed", "iterations"),
                       "mixed_outer": get("mixed", "outer_iterations")},
        "true_relative_residual": {"fp64": get("fp64", "true_relative_residual"),
                                   "mixed": get("mixed", "true_relative_residual")},
        "setup_s": {"fp64": get("fp64", "setup_seconds"), "mixed": get("mixed", "setup_seconds")},
        "mixed_vs_fp64_solution_rel_diff": err,
        "amg": {"levels": info["levels"], "operator_complexity": info["operator_complexity"]},
        "spmv": {"fp64_us": round(us64, 2), "fp64_GBps": round(gbs64, 1),
                 "fp32_us": round(us32, 2), "fp32_GBps": round(gbs32, 1),
                 "fp64_cold_us": round(us64c, 2), "fp64_cold_GBps": round(gbs64c, 1),
                 "fp32_cold_us": round(us32c, 2), "fp32_cold_GBps": round(gbs32c, 1),
                 "caller_order_fp64_GBps": gbs_raw,
                 "bandwidth": [info["bandwidth_before"], info["bandwidth_after"]]},
        "roofline": {"bound": "hbm", "achieved": round(gbs64c, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs64c / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic) if traffic else None,
                     "traffic_detail": traffic_detail, "kernel": "k_sell_spmv<double>",
                     "algorithmic_bytes": by64, "avg_us": round(us64c, 2),
                     "cache": "cold: 20 back-to-back launches rotating over copies of the "
                              "matrix (> 1 GiB in all), HIP events",
                     "warm_achieved": round(gbs64, 1), "warm_avg_us": round(us64, 2)},
        "cpu_baseline": cpu,
        "error_vs_oracle": errors,
        "adjoint": adjoint,
        "batched": batch,
        "momentum": momentum,
        "flow": flow,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
