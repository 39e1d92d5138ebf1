"""Headline benchmark: static Newton load steps of a 3D simple-cubic lattice 64^3
(262,144 particles, 1st+2nd neighbours 6+12), E=200 GPa, nu=0.3, uniaxial tension to
0.5% strain in 20 steps, FP32 BSR stiffness / SpMV / CG with FP64 residual, refinement
and positions (precision mode 0), on one MI355X.

One step = SolverStatic::solveProblemStep (reference include/solver_static.h:97-149):
BCs, stiffness assembly, Dirichlet elimination, modified-Newton loop with a Jacobi-PCG
solve per iteration, state and damage update, state store.

Prints one JSON line. `value` = load steps/s; `roofline` = the CG SpMV kernel's achieved
HBM bandwidth (algorithmic bytes / mean kernel time from HIP events recorded on the
library's stream around every SpMV launch of the timed steps) against 8 TB/s;
`cpu_baseline` = the FP64 oracle (same algorithm, same tolerances) timed on the host on a
bounded sample of the same workload and projected to one load step.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "lpm-mac_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12
HOLD_TOL_NR = 1e-8  # Newton test of the fatigue ensemble's holding step before its adjoint
# HBM traffic of one CG SpMV launch (full store, FP32, SC 64^3), measured with two rocprofv3
# --pmc passes (FETCH_SIZE, then WRITE_SIZE) over
#   bench.py --steps 1 --warmup 0 --no-cpu-baseline --pmc-calibrate
# Each reading is scaled by what the same counter reports for k_copy_calib, which moves a known
# 256 MiB in and 256 MiB out: FETCH_SIZE reads 131083.3 KiB for it (x1.9998), WRITE_SIZE
# 262144.0 KiB (x1.0000). Mean over the fused CG SpMV launches that did work (the configuration
# the tuner kept in that run): FETCH_SIZE 317832.7 KiB -> 650.9 MB; WRITE_SIZE 4121.8 KiB -> 4.2 MB
# (the 3 MB output vector, padded to 4 entries). The tuner's other variants fetch 640-653 MB.
# With int16 column offsets (the default where they fit, as at 64^3), the configuration the
# tuner kept in both passes: FETCH_SIZE 296578.8 KiB -> 607.3 MB; WRITE_SIZE 4114.8 KiB -> 4.2 MB
# (calibration 131083.4 KiB fetched for the 256 MiB copy).
# With the matrix and index streams read by non-temporal loads (the default) and 3-entry vectors,
# the configuration the tuner kept in the WRITE_SIZE pass, k_spmv with unroll 4: FETCH_SIZE
# 290359.3 KiB -> 594.6 MB (its 28 tuner launches in the FETCH_SIZE pass; every variant fetches
# 594.1-595.9 MB); WRITE_SIZE 3102.7 KiB -> 3.2 MB (calibration 131082.8 KiB fetched for the copy).
MEASURED_TRAFFIC = {
    (64, "k_spmv (CG SpMV, full store)"): 317832.7 * 1024 * 262144.0 / 131083.3 + 4121.8 * 1024,
    (64, "k_spmv (CG SpMV, full store, int16 column offsets)"): 296578.8 * 1024 * 262144.0 / 131083.4 + 4114.8 * 1024,
    (64, "k_spmv (CG SpMV, full store, int16 column offsets, non-temporal matrix loads)"):
        290359.3 * 1024 * 262144.0 / 131082.8 + 3102.7 * 1024,
}


def sc_workload(n):
    """SC lattice with n particles per side, spacing 1 mm (radius 0.5); the box is shifted so that
    the generator's row offset (utilities.h:174-195) keeps n rows in y and z."""
    import lpmgpu as G

    r = 0.5
    L = 2 * r * (n - 1)
    a = math.sqrt(3 * L * L)
    off = (-a) % (2 * r)
    box = [0.0, L, off, off + L, off, off + L]
    xyz = G.generate_lattice(G.SC3D, box, r)
    assert len(xyz) == n ** 3, (len(xyz), n)
    return xyz, box, r, L


def spmv_bytes(ctx):
    """Algorithmic bytes of one FP32 3x3-block SpMV and the kernel that ran it.

    Full store: every block's 9 values + its column index (4 B, or 2 B when the columns are
    stored as int16 offsets from the row), row pointers, one read of the input vector and one
    write of the output (4 B per DoF; the 16 B per particle padding of the device vectors is
    not counted)."""
    nblk = ctx.size("nconn")
    n = ctx.n
    vec = 8 * ctx.size("ndof")  # input read once + output written once, 4 B per DoF
    ib = int(ctx.get_mode("spmv_index_bytes"))
    name = "k_spmv (CG SpMV, full store)" if ib == 4 else "k_spmv (CG SpMV, full store, int16 column offsets)"
    if ib == 2 and int(ctx.get_mode("spmv_nt")):
        name = name[:-1] + ", non-temporal matrix loads)"
    return (36 + ib) * nblk + 4 * (n + 1) + vec, name


def setup_gpu(n, tol_nr, cg_tol, precond, cg_forcing, coarse_cells=0, predictor=0, coarse_reuse=0):
    import lpmgpu as G

    xyz, box, r, L = sc_workload(n)
    t0 = time.time()
    ctx = G.Context(G.SC3D, r, G.ELASTIC, xyz, bonds="device")  # bond search on the device
    ctx.set_elastic(0.0, False, 200e3, 0.3, 1.0, elmap_literal=0)
    ctx.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, cg_tol=cg_tol, tol_NR=tol_nr,
             max_NR_iter=30, cg_maxit=20000, precond=precond, cg_forcing=cg_forcing,
             coarse_cells=coarse_cells, predictor=predictor, coarse_reuse=coarse_reuse)
    z = xyz[:, 2]
    bottom = np.nonzero(z < z.min() + 1e-6)[0]
    top = np.nonzero(z > z.max() - 1e-6)[0]
    dz = 0.005 * L / 20.0
    ls = ctx.load_step(disp=[(bottom, "x", 0.0), (bottom, "y", 0.0), (bottom, "z", 0.0), (top, "z", dz)])
    ctx.update_geometry()
    ctx.update_force_state()
    ctx.box = box
    return ctx, ls, xyz, time.time() - t0


def spmv_launch_info(ctx):
    """The CG SpMV launch shape in use (the tuner's pick for the current matrix kind)."""
    coop = int(ctx.get_mode("spmv_coop"))
    return {"slots_unrolled": int(ctx.get_mode("spmv_unr")), "wavefronts_per_workgroup": int(ctx.get_mode("spmv_wpb")),
            "xcd_remap": int(ctx.get_mode("spmv_remap")), "tuned": int(ctx.get_mode("spmv_tune")),
            "wavefronts_per_slice": coop or 1, "slices_per_workgroup": int(ctx.get_mode("spmv_spw")) if coop else None,
            "nontemporal_matrix_loads": int(ctx.get_mode("spmv_nt"))}


def apply_modes(ctx, args):
    """The --mode KEY=VALUE settings, applied after a workload's own modes."""
    for kv in args.mode:
        key, val = kv.split("=")
        ctx.mode(**{key: float(val)})


def reset_state(ctx, xyz):
    n = ctx.n
    ctx.set("xyz", xyz)
    ctx.set("xyz_last", xyz)
    ctx.set("Pex", np.zeros(3 * n))
    ctx.set("disp_constraint", np.zeros(3 * n, dtype=np.int32))
    ctx.update_geometry()
    ctx.update_force_state()


def host_cpu():
    """Model name, physical cores and the CPUs this process may run on (its affinity)."""
    model, cores, cur = None, set(), {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k in ("physical id", "core id"):
                cur[k] = v
            elif not line.strip() and cur:
                cores.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "physical_cores": len(cores) or None, "usable_cpus": usable}


def cpu_full_step(n, tol_nr, cg_tol):
    """One whole load step of the reference algorithm on the host, timed end to end: the FP64
    oracle from rest, SolverStatic::solveProblemStep with Jacobi CG (Eigen's preconditioner) at
    the same relative CG tolerance and inexact-Newton forcing as the GPU run, the same Newton test."""
    import oracle as O

    xyz, box, r, L = sc_workload(n)
    s = O.System(O.SC3D, r, O.ELASTIC, xyz, box)
    s.mode(elmap_literal=0, dirichlet_static=1, cg_transpose=0, diag10_literal=0, cg_tol=cg_tol, cg_forcing=1,
           tol_NR=tol_nr, max_NR_iter=30)
    s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    z = xyz[:, 2]
    bottom = np.nonzero(z < z.min() + 1e-6)[0]
    top = np.nonzero(z > z.max() - 1e-6)[0]
    ls = O.LoadStep()
    for ax in "xyz":
        ls.add_disp(bottom, ax, 0.0)
    ls.add_disp(top, "z", 0.005 * L / 20.0)
    s.update_geometry()
    s.update_force_state()
    t0 = time.perf_counter()
    ok = s.solve_step_static(ls, -1)
    t = time.perf_counter() - t0
    return t, bool(ok), s.info("newton"), int(np.sum(s.geti("cg_log")))


def cpu_baseline(n, gpu_counts, max_threads, tol_nr, cg_tol):
    """The FP64 oracle on the same 64^3 system on the host's cores. The value is one whole load
    step timed end to end (cpu_full_step). Beside it, the earlier projection: one assembly, one
    geometry/force/residual pass and a bounded number of Jacobi-CG iterations timed, projected with
    the Newton and CG counts of the GPU's Jacobi run of the same steps."""
    import oracle as O

    xyz, box, r, L = sc_workload(n)
    s = O.System(O.SC3D, r, O.ELASTIC, xyz, box)
    s.mode(elmap_literal=0, dirichlet_static=1, cg_transpose=0, diag10_literal=0)
    s.set_elastic(0.0, False, 200e3, 0.3, 1.0)
    z = xyz[:, 2]
    bottom = np.nonzero(z < z.min() + 1e-6)[0]
    top = np.nonzero(z > z.max() - 1e-6)[0]
    ls = O.LoadStep()
    for ax in "xyz":
        ls.add_disp(bottom, ax, 0.0)
    ls.add_disp(top, "z", 0.005 * L / 20.0)
    s.update_geometry()
    s.update_force_state()
    s.apply_bc(ls)
    t0 = time.perf_counter()
    s.calc_stiffness()
    s.update_stiffness_disp_bc()
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.update_geometry()
    s.update_force_state()
    s.update_rr()
    t_force = time.perf_counter() - t0
    n_cg_sample = 40
    s.mode(cg_tol=1e-30, cg_maxit=n_cg_sample)
    t0 = time.perf_counter()
    s.cg()
    t_cg = (time.perf_counter() - t0) / n_cg_sample
    newton_per_step, cg_per_step, asm_per_step = gpu_counts  # of the GPU run with the same (Jacobi) solver
    t_step = asm_per_step * t_asm + (newton_per_step + asm_per_step) * t_force + cg_per_step * t_cg
    # the reference-faithful variant: assembly and forces serial, as the reference's loops are;
    # only the CG's sparse products threaded (Eigen parallelises them when built with OpenMP)
    ref = None
    try:
        import ctypes

        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        try:
            s.apply_bc(ls)
            t0 = time.perf_counter()
            s.calc_stiffness()
            s.update_stiffness_disp_bc()
            t_asm1 = time.perf_counter() - t0
            t0 = time.perf_counter()
            s.update_geometry()
            s.update_force_state()
            s.update_rr()
            t_force1 = time.perf_counter() - t0
        finally:
            gomp.omp_set_num_threads(max_threads)
        t_step1 = asm_per_step * t_asm1 + (newton_per_step + asm_per_step) * t_force1 + cg_per_step * t_cg
        ref = {"value": 1.0 / t_step1, "unit": "load_steps/s", "cores": max_threads,
               "sample": (f"same projection with the assembly ({t_asm1:.2f} s) and the geometry+force+residual pass "
                          f"({t_force1:.3f} s) on one thread, the CG iterations on {max_threads}")}
    except OSError:
        pass
    del s
    t_full, ok, full_newton, full_cg = cpu_full_step(n, tol_nr, cg_tol)
    return {
        "value": 1.0 / t_full,
        "unit": "load_steps/s",
        "cores": max_threads,
        "kind": "port",
        "sample": (f"FP64 oracle, SC {n}^3, the first load step from rest timed end to end ({t_full:.1f} s, "
                   f"{'converged' if ok else 'NOT converged'}): reference algorithm (Jacobi CG, no predictor) with "
                   f"the GPU run's CG tolerance {cg_tol:g} and forcing, Newton test {tol_nr:g}; {full_newton} Newton "
                   f"iterations, {full_cg} CG iterations; OpenMP over particles and matrix rows on {max_threads} threads"),
        "host": host_cpu(),
        "projected_from_phases": {
            "value": 1.0 / t_step, "unit": "load_steps/s",
            "sample": (f"1 assembly ({t_asm:.2f} s), 1 geometry+force+residual pass ({t_force:.3f} s), "
                       f"{n_cg_sample} Jacobi-CG iterations ({t_cg * 1e3:.1f} ms each) timed; step projected with the "
                       f"per-step counts of the GPU Jacobi run: {asm_per_step:.2f} assemblies, {newton_per_step:.2f} Newton "
                       f"iterations, {cg_per_step:.0f} CG iterations")},
        "reference_faithful_threading": ref,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="particles per side (64 -> 262,144)")
    ap.add_argument("--tol-nr", type=float, default=1e-10)
    ap.add_argument("--cg-tol", type=float, default=1e-3)
    ap.add_argument("--cg-forcing", type=int, default=1, help="1: inexact-Newton CG tolerance per iteration")
    ap.add_argument("--precond", type=int, default=2,
                    help="0 Jacobi (as the reference's Eigen CG), 1 block Jacobi, 2 Jacobi + rigid-body coarse correction")
    ap.add_argument("--coarse-cells", type=int, default=0, help="aggregate edge of the two-level preconditioner (0: auto)")
    ap.add_argument("--mode", action="append", default=[], metavar="KEY=VALUE",
                    help="an extra lpm_set_mode after each workload's own modes (repeatable)")
    ap.add_argument("--coarse-budget", type=int, default=0,
                    help="coarse DoFs per specimen of the automatic aggregate size (0: the library default)")
    ap.add_argument("--predictor", type=int, default=2,
                    help="2: start each static step from the quadratic through the last three converged states, 1: "
                         "from the previous step's converged increment (Newton runs to the same test either way); 0: "
                         "the reference's start. The cracked plate with bond breaking (--config 4) uses at most 1")
    ap.add_argument("--coarse-reuse", type=int, default=1,
                    help="1: keep the two-level coarse space across assemblies while it still works (rebuilt when the "
                         "eliminated DoFs change or the CG slows down); 0: rebuild it with every assembly")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="instead of the headline, time this many uneven hex specimens in one context")
    ap.add_argument("--fatigue", action="store_true", help="with --ensemble: HCF dogbones through fatigue cycles")
    ap.add_argument("--timing-every", type=int, default=16,
                    help="bracket the CG SpMV of every k-th iteration with HIP events")
    ap.add_argument("--config", type=int, default=2,
                    help="2: the headline (SC 64^3); 3: FCC 40^3 J2 crystal; 4: cracked 400 x 400 plate with breaking "
                         "and the adjoint (one JSON line each)")
    ap.add_argument("--cg-graph", type=int, default=0,
                    help="with --config 3|4: 1 replays each batch of CG iterations from a captured HIP graph")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="run the 256 MiB calibration copy (k_copy_calib) only 3 times, for rocprofv3 --pmc passes "
                         "whose FETCH_SIZE / WRITE_SIZE readings it scales")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding of this path is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    import lpmgpu as G

    G.lib()
    if args.config == 3:
        config3_bench(args)
        return
    if args.config == 4:
        config4_bench(args)
        return
    if args.ensemble > 0:
        if args.fatigue:
            fatigue_ensemble_bench(args)
            return
        if args.steps == 20:
            args.steps = 5
        ensemble_bench(args)
        return
    ctx, ls, xyz, t_setup = setup_gpu(args.n, args.tol_nr, args.cg_tol, args.precond, args.cg_forcing,
                                      args.coarse_cells, args.predictor, args.coarse_reuse)
    if args.verbose:
        print(f"setup {t_setup:.2f} s, n={ctx.n} bonds={ctx.size('nbonds')} blocks={ctx.size('nconn')}", flush=True)
    # warmup steps, then back to the initial state so that the timed steps are steps 1..K
    apply_modes(ctx, args)
    for _ in range(args.warmup):
        ok = ctx.solve_step_static(ls, -1)
        assert ok
    reset_state(ctx, xyz)
    ctx.reset_stats()
    ctx.set_timing(args.timing_every)
    # the timed steps start cold: the first assembly rebuilds the coarse space (a warm-up build
    # is not carried into the timed region)
    ctx.mode(coarse_rebuild=1)
    hip_sync(ctx)
    t0 = time.perf_counter()
    for k in range(args.steps):
        ok = ctx.solve_step_static(ls, -1)
        if not ok:
            raise RuntimeError(f"load step {k} did not converge")
        if args.verbose:
            print(f"step {k}: newton {ctx.stat('newton_iters'):.0f} cg {ctx.stat('cg_iters'):.0f}", flush=True)
    hip_sync(ctx)
    t = time.perf_counter() - t0
    ctx.set_timing(0)
    spmv_iso_ms = ctx.spmv_bench(50)
    tuned_ms = ctx.stat("spmv_tuned_ms")
    # practical ceiling: a plain streaming copy of the SpMV's size (read + write)
    copy_bytes = 256 << 20
    copy_ms = ctx.copy_bench(copy_bytes, 3 if args.pmc_calibrate else 20)
    K = args.steps
    spmv_n = ctx.stat("spmv_launches_timed")
    spmv_ms = ctx.stat("spmv_ms") / max(spmv_n, 1)
    B, kname = spmv_bytes(ctx)
    achieved = B / (spmv_ms * 1e-3) / 1e9
    counts = (ctx.stat("newton_iters") / K, ctx.stat("cg_iters") / K, ctx.stat("assemblies") / K)
    cg_solves = ctx.stat("cg_solves")
    coarse_builds = ctx.stat("coarse_builds") / K
    ncoarse = ctx.get_mode("ncoarse")
    # equilibrium check of the last step (size-independent): residual and reaction balance
    nr, nR = ctx.update_rr()
    def rerun(**modes):
        reset_state(ctx, xyz)
        ctx.reset_stats()
        ctx.mode(**modes)
        ctx.mode(coarse_rebuild=1)
        hip_sync(ctx)
        tr = time.perf_counter()
        for _ in range(K):
            if not ctx.solve_step_static(ls, -1):
                raise RuntimeError(f"comparison run {modes} did not converge")
        hip_sync(ctx)
        tr = time.perf_counter() - tr
        return ({"load_steps_per_s": K / tr, "newton_iters_per_step": ctx.stat("newton_iters") / K,
                 "cg_iters_per_step": ctx.stat("cg_iters") / K},
                (ctx.stat("newton_iters") / K, ctx.stat("cg_iters") / K, ctx.stat("assemblies") / K))

    # Per-step output (Assembly::writeDump, about 50 MB of text per step here; excluded from the
    # headline): 3 steps each appending a frame from the asynchronous writer, against one frame
    # written with nothing to overlap it.
    dump = None
    if not args.no_cpu_baseline:
        import tempfile

        path = os.path.join(tempfile.gettempdir(), f"lpm_bench_{os.getpid()}.dump")
        try:
            reset_state(ctx, xyz)
            hip_sync(ctx)
            t3 = time.perf_counter()
            for k in range(3):
                if not ctx.solve_step_static(ls, -1):
                    raise RuntimeError("dump run did not converge")
            hip_sync(ctx)
            t3 = time.perf_counter() - t3
            reset_state(ctx, xyz)
            hip_sync(ctx)
            td = time.perf_counter()
            for k in range(3):
                if not ctx.solve_step_static(ls, -1):
                    raise RuntimeError("dump run did not converge")
                ctx.dump_async(path, k + 1, ctx.box)
            ctx.dump_wait()
            td = time.perf_counter() - td
            frame_bytes = os.path.getsize(path) / 3
            t1 = time.perf_counter()
            ctx.dump_async(path, 4, ctx.box)
            ctx.dump_wait()
            t1 = time.perf_counter() - t1
            dump = {"load_steps_per_s_with_a_frame_each_step": 3 / td, "load_steps_per_s_same_steps_without": 3 / t3,
                    "frame_bytes": frame_bytes,
                    "frame_alone_ms": 1e3 * t1, "steps": 3}
        finally:
            if os.path.exists(path):
                os.remove(path)

    # The same K steps again: without the predictor (the reference's Newton start), and with the
    # reference's algorithm (Jacobi, no predictor), whose iteration counts drive the CPU
    # projection; both rates are reported beside the headline.
    jac = nopred = graph = noreuse = linpred = None
    jac_counts = counts
    if not args.no_cpu_baseline:
        # the same steps with each CG batch replayed from a captured HIP graph (per-SpMV timing
        # is off in that mode, so the headline and its roofline come from direct launches)
        graph, _ = rerun(cg_graph=1)
        ctx.mode(cg_graph=0)
        if args.predictor == 2:
            # the linear predictor (the previous step's increment) on the same steps
            linpred, _ = rerun(predictor=1)
            ctx.mode(predictor=args.predictor)
        if args.coarse_reuse and args.precond == 2:
            # the coarse space rebuilt with every assembly
            noreuse, _ = rerun(coarse_reuse=0)
        if args.predictor:
            # neither the predictor nor coarse reuse: the two-level solver from the reference's start
            nopred, _ = rerun(predictor=0, coarse_reuse=0)
        if args.precond != 0 or args.predictor:
            jac, jac_counts = rerun(precond=0, predictor=0)
    out = {
        "metric": "load steps/s + CG SpMV HBM GB/s (% peak), 3D SC 64^3, FP32-mixed vs FP64",
        "value": K / t,
        "unit": "load_steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (FP32 K/SpMV/CG, FP64 residual/positions)",
        "data": "synthetic",
        "config": {
            "workload": f"SC {args.n}^3 elastic uniaxial tension, {ctx.n} particles, 0.5% strain in 20 steps",
            "particles": ctx.n, "bonds": ctx.size("nbonds"), "bsr_blocks": ctx.size("nconn"),
            "newton_tol": args.tol_nr, "cg_rtol": args.cg_tol, "cg_forcing": args.cg_forcing,
            "predictor": {0: "none (the reference's start)", 1: "previous step's converged increment",
                          2: "quadratic through the last three converged states"}[args.predictor],
            "precond": {0: "jacobi", 1: "block-jacobi", 2: "two-level (jacobi + rigid-body aggregate coarse space)"}[args.precond],
            "coarse_dofs": int(ncoarse), "coarse_reuse": args.coarse_reuse, "coarse_builds_per_step": coarse_builds,
            "cg_batches": {"sized_first_batch": int(ctx.get_mode("cg_batch_auto")),
                           "next_batches": int(ctx.get_mode("cg_batch_next")), "batch": int(ctx.get_mode("cg_batch"))},
            "spmv_launch": spmv_launch_info(ctx),
            "assembly_launch": {"rows_per_workgroup": int(ctx.get_mode("asm_wpb")), "tuned": int(ctx.get_mode("asm_tune")),
                                "tuned_ms": ctx.stat("asm_tuned_ms")},
            "cg_solves": cg_solves,
            "newton_iters_per_step": counts[0], "cg_iters_per_step": counts[1],
            "final_residual_over_reaction": nr / nR if nR > 0 else None,
            "parallelism": "single GPU",
            "setup_seconds": t_setup,
            "async_dump": dump,
            "same_steps_with_cg_graph": graph,
            "same_steps_with_the_linear_predictor": linpred,
            "same_steps_without_coarse_reuse": noreuse,
            "same_steps_without_predictor_or_coarse_reuse": nopred,
            "same_steps_reference_algorithm_jacobi_no_predictor": jac,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved * 1e9 / HBM_PEAK, "traffic": MEASURED_TRAFFIC.get((args.n, kname)),
            "kernel": kname, "bytes_per_launch": B, "mean_kernel_ms": spmv_ms,
            "launches_timed": spmv_n, "isolated_spmv_ms": spmv_iso_ms, "isolated_fused_spmv_ms": tuned_ms,
            "copy_kernel_GBps": 2 * copy_bytes / (copy_ms * 1e-3) / 1e9,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.n, jac_counts, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
                                           args.tol_nr, args.cg_tol)
    print(json.dumps(out), flush=True)


def ensemble_bench(args):
    """Specimen ensemble (the shape of the 64-dogbone calibration workload): 64 hex plates of
    uniform random size in [5k, 60k] particles (seed 5), Ti-like E = 110 GPa, nu = 0.34, plane
    stress, bottom edge clamped, top edge pulled each step by its own strain increment (uniform in
    [1e-4, 3e-4] of the specimen height), `steps` static load steps. Timed once as one context holding all specimens (one launch per kernel) and once
    as one context per specimen run back to back; prints one JSON line."""
    import lpmgpu as G

    rng = np.random.default_rng(5)
    S = args.ensemble
    sizes = rng.integers(5000, 60001, S)
    strains = rng.uniform(1e-4, 3e-4, S)
    parts, pulls = [], []
    for m, e in zip(sizes, strains):
        nx = int(round(math.sqrt(m)))
        box = [0.0, float(nx), 0.0, 0.8660254037844386 * nx, 0.0, 1.0]
        x = G.generate_lattice(G.HEXAGON2D, box, 0.5)
        parts.append(x)
        pulls.append(e * (x[:, 1].max() - x[:, 1].min()))

    def make(xyzs, bonds=None):
        ctx = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, np.vstack(xyzs), bonds=bonds)
        ctx.set_elastic(0.0, True, 110e3, 0.34, 1.0, elmap_literal=0)
        ctx.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=args.precond,
                 cg_tol=args.cg_tol, cg_forcing=args.cg_forcing, tol_NR=args.tol_nr, max_NR_iter=30, cg_maxit=50000,
                 predictor=args.predictor)
        if args.coarse_budget > 0:
            ctx.mode(coarse_budget=args.coarse_budget)
        apply_modes(ctx, args)
        return ctx

    def bcs(xyzs, offs, pulls_):
        disp = []
        for x, off, pull in zip(xyzs, offs, pulls_):
            y = x[:, 1]
            ids = np.arange(off, off + len(x))
            lo, hi = ids[y < y.min() + 1e-6], ids[y > y.max() - 1e-6]
            disp += [(lo, "x", 0.0), (lo, "y", 0.0), (hi, "y", float(pull))]
        return disp

    # one context for all specimens: concatenated positions, bond CSR with shifted ids
    bp, bs, p2, first = [np.zeros(1, np.int64)], [], [], [0]
    for x in parts:
        b_ptr, b_split, b_p2 = (np.asarray(a, dtype=np.int64) for a in G.build_bonds(G.HEXAGON2D, 0.5, x))
        boff = int(bp[-1][-1])
        bp.append(b_ptr[1:] + boff)
        bs.append(b_split + boff)
        p2.append(b_p2 + first[-1])
        first.append(first[-1] + len(x))
    union = tuple(np.concatenate(a).astype(np.int32) for a in (bp, bs, p2))
    ctx = make(parts, union)
    ctx.set_specimens(first)
    ls = ctx.load_step(disp=bcs(parts, first[:-1], pulls))
    ctx.update_geometry()
    ctx.update_force_state()
    ctx.solve_step_static(ls, -1)  # warm-up step (tuner, coarse space)
    hip_sync(ctx)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if not ctx.solve_step_static(ls, -1):
            bad = np.nonzero(ctx.specimen_stat("converged") != 1)[0]
            raise RuntimeError(f"ensemble step did not converge: specimens {bad.tolist()}, Newton iterations "
                               f"{ctx.specimen_stat('newton_iters')[bad].tolist()}")
    hip_sync(ctx)
    t_batch = time.perf_counter() - t0
    newton = ctx.specimen_stat("newton_iters")
    # the next steps of the same load path with each CG batch replayed from a captured graph
    ctx.mode(cg_graph=1)
    ctx.solve_step_static(ls, -1)  # capture
    hip_sync(ctx)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if not ctx.solve_step_static(ls, -1):
            raise RuntimeError("ensemble step did not converge")
    hip_sync(ctx)
    t_graph = time.perf_counter() - t0
    # and direct launches again over the next steps, so that the graph's rate sits between two
    # direct runs of neighbouring steps
    ctx.mode(cg_graph=0)
    hip_sync(ctx)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if not ctx.solve_step_static(ls, -1):
            raise RuntimeError("ensemble step did not converge")
    hip_sync(ctx)
    t_direct2 = time.perf_counter() - t0
    del ctx
    # the same specimens, one context each, back to back
    ctxs = []
    for x, pull in zip(parts, pulls):
        c1 = make([x])
        l1 = c1.load_step(disp=bcs([x], [0], [pull]))
        c1.update_geometry()
        c1.update_force_state()
        c1.solve_step_static(l1, -1)
        ctxs.append((c1, l1))
    for c1, _ in ctxs:
        hip_sync(c1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for c1, l1 in ctxs:
            if not c1.solve_step_static(l1, -1):
                raise RuntimeError("single-specimen step did not converge")
    for c1, _ in ctxs:
        hip_sync(c1)
    t_seq = time.perf_counter() - t0
    for c1, l1 in ctxs:
        c1.mode(cg_graph=1)
        c1.solve_step_static(l1, -1)
        hip_sync(c1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for c1, l1 in ctxs:
            if not c1.solve_step_static(l1, -1):
                raise RuntimeError("single-specimen step did not converge")
    for c1, _ in ctxs:
        hip_sync(c1)
    t_seq_graph = time.perf_counter() - t0
    total = int(first[-1])
    out = {
        "metric": "specimen load steps/s, 64 uneven hex specimens in one context",
        "value": S * args.steps / t_batch,
        "unit": "specimen_load_steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": 1e3 * t_batch / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (FP32 K/SpMV/CG, FP64 residual/positions)",
        "data": "synthetic",
        "config": {"workload": f"{S} hex plates, {total} particles, sizes {int(sizes.min())}-{int(sizes.max())}",
                   "precond": args.precond, "newton_iters_per_step_max": float(newton.max()),
                   "one_context_per_specimen_specimen_steps_per_s": S * args.steps / t_seq,
                   "speedup_vs_one_context_per_specimen": t_seq / t_batch,
                   "with_cg_graph": {"specimen_steps_per_s": S * args.steps / t_graph,
                                     "direct_launches_next_steps_specimen_steps_per_s": S * args.steps / t_direct2,
                                     "one_context_per_specimen_specimen_steps_per_s": S * args.steps / t_seq_graph}},
    }
    print(json.dumps(out), flush=True)


def dogbone(lattice_mod, m):
    """Positions of a 2D hex dogbone of about m particles, loaded along y: grips of width W over
    the top and bottom quarter of the height H = 2W, a gauge of width W/2 over the middle 30%,
    straight tapers between (a mask over a hex plate from the lattice generator)."""
    G = lattice_mod
    W = math.sqrt(m * 0.8660254037844386 / 1.6)
    H = 2.0 * W
    x = G.generate_lattice(G.HEXAGON2D, [0.0, W, 0.0, H, 0.0, 1.0], 0.5)
    y = (x[:, 1] - x[:, 1].min()) / (x[:, 1].max() - x[:, 1].min())
    d = np.abs(y - 0.5)  # 0 at mid-height, 0.5 at the ends
    width = np.where(d <= 0.15, 0.5, np.where(d >= 0.25, 1.0, 0.5 + 0.5 * (d - 0.15) / 0.10)) * W
    keep = np.abs(x[:, 0] - 0.5 * W) <= 0.5 * width + 1e-9
    return np.ascontiguousarray(x[keep])


def fatigue_ensemble_bench(args):
    """HCF fatigue ensemble (the shape of the 64-dogbone calibration workload): 64 2D hex
    dogbones of uniform random size in [5k, 60k] particles (seed 5), Ti-like E = 110 GPa,
    nu = 0.34, plane stress, HCF damage with per-specimen A (uniform in [3e-5, 8e-5]), B = 2.019,
    d = 0.5, nonlocal length 4; bottom row clamped, top row displaced (LoadType dogbone
    displacement) with R = 0.1 cycles of amplitude 1e-3 of each specimen's height, 200 cycles
    with the linear time map (tau = 0.05: 20 cycles per simulated cycle). Timed once as one
    context holding all specimens and once as one context per specimen run back to back."""
    import lpmgpu as G

    rng = np.random.default_rng(5)
    S = args.ensemble
    sizes = rng.integers(5000, 60001, S)
    A = rng.uniform(3e-5, 8e-5, S)
    parts = [dogbone(G, m) for m in sizes]
    heights = [x[:, 1].max() - x[:, 1].min() for x in parts]
    amp = [1e-3 * h for h in heights]
    cycles, tau = 200, 0.05
    spec = [0.1] + [1.0, 0.1] * cycles

    def rows(x, off):
        y = x[:, 1]
        ids = np.arange(off, off + len(x))
        return ids[y > y.max() - 1e-6], ids[y < y.min() + 1e-6]

    def make(xyzs, offs, bonds=None, nspec=False):
        ctx = G.Context(G.HEXAGON2D, 0.5, G.HCF, np.vstack(xyzs), bonds=bonds)
        if nspec:
            ctx.set_specimens(offs + [offs[-1] + len(xyzs[-1])])
        t = np.zeros(sum(len(x) for x in xyzs), dtype=np.int32)
        groups = []
        for x, off in zip(xyzs, offs):
            top, bottom = rows(x, off)
            t[top] = 4
            t[bottom] = 4
            groups.append([top, bottom])
        ctx.set("type", t)
        ctx.set_hcf(4.0, True, 110e3, 0.34, 5e-5, 2.019, 0.5, elmap_literal=0)
        mat = ctx.get("mat").reshape(-1, 10)
        return ctx, groups, mat

    def setup(ctx):
        ctx.search_nonlocal(1.5)
        ctx.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=args.precond,
                 cg_tol=args.cg_tol, cg_forcing=args.cg_forcing, tol_NR=args.tol_nr, max_NR_iter=30, cg_maxit=50000)
        if args.coarse_budget > 0:
            ctx.mode(coarse_budget=args.coarse_budget)
        apply_modes(ctx, args)
        ctx.update_geometry()
        ctx.update_force_state()

    # one context for all specimens
    bp, bs, p2, first = [np.zeros(1, np.int64)], [], [], [0]
    for x in parts:
        b_ptr, b_split, b_p2 = (np.asarray(a, dtype=np.int64) for a in G.build_bonds(G.HEXAGON2D, 0.5, x))
        boff = int(bp[-1][-1])
        bp.append(b_ptr[1:] + boff)
        bs.append(b_split + boff)
        p2.append(b_p2 + first[-1])
        first.append(first[-1] + len(x))
    union = tuple(np.concatenate(a).astype(np.int32) for a in (bp, bs, p2))

    def run_union(graph):
        ctx, groups, mat = make(parts, first[:-1], union, nspec=True)
        for s in range(S):
            mat[first[s]:first[s + 1], 0] = A[s]
        ctx.set("mat", mat)
        setup(ctx)
        ctx.mode(cg_graph=graph)
        disp = []
        for (top, bottom), a in zip(groups, amp):
            disp += [(top, "x", 0.0), (top, "y", 0.1 * a), (top, "z", 0.0), (bottom, "x", 0.0), (bottom, "y", 0.0),
                     (bottom, "z", 0.0)]
        hip_sync(ctx)
        t0 = time.perf_counter()
        ctx.solve_problem_fatigue_static([ctx.load_step(disp=disp)])
        n_done = ctx.solve_problem_cyclic_specimens(G.LOAD_DOGBONE_DISP, 1, spec, groups, amp, tau, 4)
        hip_sync(ctx)
        t = time.perf_counter() - t0
        res = (t, n_done, float(ctx.get("damage").max()), ctx.stat("cg_solves"))
        # batched sensitivity of the summed top-row reactions to every particle's (A, B, d) through
        # the last fatigue update: re-solve at the current loads with the updated damage, then one
        # transposed solve for all specimens (no bond or nonlocal neighbour joins two of them)
        hold = []
        for top, bottom in groups:
            hold += [(top, "x", 0.0), (top, "y", 0.0), (top, "z", 0.0), (bottom, "x", 0.0), (bottom, "y", 0.0),
                     (bottom, "z", 0.0)]
        # Newton test of this step: 1e-8 of each specimen's reaction. A step with no load
        # increment starts from a residual far below the reaction, so the reference's multiplier
        # is the reaction itself, and 1e-10 of it lies under the residual's rounding floor: the
        # FP64 oracle on one of these dogbones stalls at 5e-10 of the reaction after one iteration.
        ctx.mode(max_NR_iter=30, tol_NR=HOLD_TOL_NR)
        ok = ctx.solve_step_fatigue(ctx.load_step(disp=hold))
        ctx.mode(tol_NR=args.tol_nr)
        newton = ctx.specimen_stat("newton_iters") if hasattr(ctx, "specimen_stat") else None
        if ok:
            w = np.zeros(2 * ctx.n)
            for top, _ in groups:
                w[2 * top + 1] = 1.0
            hip_sync(ctx)
            ta = time.perf_counter()
            Gv, dA, dB, dd, nref, rres = ctx.adjoint_fatigue_params(w_pin=w)
            hip_sync(ctx)
            adj = {"seconds": time.perf_counter() - ta, "objective": "sum of the 64 top-row reactions", "G": Gv,
                   "holding_step_newton_tol": HOLD_TOL_NR, "holding_step_newton_iters_max": int(max(newton)),
                   "refinements": nref, "relative_residual": rres,
                   "particles_with_nonzero_dG_dA": int(np.count_nonzero(dA)), "max_abs_dG_dA": float(np.abs(dA).max())}
        else:
            adj = {"skipped": "the step holding the last loads after the final update did not converge",
                   "specimen_newton_iters": None if newton is None else [int(v) for v in newton]}
        ctx.close()
        return res + (adj,)

    # direct launches only: a graph-replayed rerun of this mode crashed on the host inside
    # hipGraphLaunch under rocprofv3's kernel trace, and replay was never faster here
    t_batch, n_done, dmax, steps, adj = run_union(0)
    print(f"ensemble: {S} specimens, {n_done} cycles in one context: {t_batch:.2f} s", flush=True)
    # the same specimens, one context each, back to back
    t_seq = 0.0
    for s, x in enumerate(parts):
        c1, g1, m1 = make([x], [0])
        m1[:, 0] = A[s]
        c1.set("mat", m1)
        setup(c1)
        top, bottom = g1[0]
        d1 = [(top, "x", 0.0), (top, "y", 0.1 * amp[s]), (top, "z", 0.0), (bottom, "x", 0.0), (bottom, "y", 0.0),
              (bottom, "z", 0.0)]
        hip_sync(c1)
        t0 = time.perf_counter()
        c1.solve_problem_fatigue_static([c1.load_step(disp=d1)])
        c1.solve_problem_cyclic(G.LOAD_DOGBONE_DISP, 1, [v * amp[s] for v in spec], [top, bottom], tau, 4)
        hip_sync(c1)
        t_seq += time.perf_counter() - t0
        c1.close()
        if (s + 1) % 8 == 0:
            print(f"one context per specimen: {s + 1}/{S} done, {t_seq:.2f} s", flush=True)
    out = {
        "metric": "specimen fatigue cycles/s, 64 uneven HCF dogbones in one context",
        "value": S * n_done / t_batch,
        "unit": "specimen_cycles/s",
        "n_gpus": 1,
        "steps": int(n_done),
        "warmup": 0,
        "ms_per_step": 1e3 * t_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (FP32 K/SpMV/CG, FP64 residual/positions/damage)",
        "data": "synthetic",
        "config": {"workload": f"{S} hex dogbones, {first[-1]} particles, sizes {int(sizes.min())}-{int(sizes.max())}, "
                               f"R=0.1, {cycles} cycles, tau={tau}",
                   "precond": args.precond, "fatigue_steps": steps, "max_damage": dmax,
                   "one_context_per_specimen_specimen_cycles_per_s": S * n_done / t_seq,
                   "speedup_vs_one_context_per_specimen": t_seq / t_batch,
                   "batched_hcf_parameter_adjoint": adj},
    }
    print(json.dumps(out), flush=True)


def _line(metric, value, unit, steps, t, config, dtype="fp32-mixed (FP32 K/SpMV/CG, FP64 residual/positions/state)"):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": 0,
            "ms_per_step": 1e3 * t / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic", "config": config}


def config3_bench(args):
    """FCC single crystal of 40^3 cubic cells (256,009 particles in the rotated box; rotation
    angles uniform in [0, 2 pi) from numpy's default_rng(7)), J2 plasticity (E = 200 GPa,
    nu = 0.3, sigma_y = 300 MPa, H = 2 GPa, no damage), uniaxial tension to 2% in 50 load steps
    with the headline solver settings; all 50 steps are timed, the device setup apart."""
    import lpmgpu as G

    r = 0.5 / math.sqrt(2)
    rng = np.random.default_rng(7)
    ang = tuple(rng.uniform(0, 2 * math.pi, 3))
    xyz = G.generate_lattice(G.FCC3D, [0, 40, 0, 40, 0, 40], r, 0, ang)
    z = xyz[:, 2]
    lo, hi = np.nonzero(z < z.min() + 0.4)[0], np.nonzero(z > z.max() - 0.4)[0]
    L = z.max() - z.min()
    t0 = time.perf_counter()
    g = G.Context(G.FCC3D, r, G.J2, xyz, bonds="device")
    g.set_j2(0.0, False, 200e3, 0.3, 300.0, 0.0, 2e3, 0.0, 1.0, damage_threshold=1.0, elmap_literal=0)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=args.precond, cg_tol=args.cg_tol,
           cg_forcing=args.cg_forcing, tol_NR=args.tol_nr, max_NR_iter=50, cg_maxit=20000, predictor=args.predictor,
           cg_graph=args.cg_graph, coarse_reuse=args.coarse_reuse)
    apply_modes(g, args)
    steps = 50
    ls = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", 0.0), (lo, "z", 0.0), (hi, "z", 0.02 * L / steps)])
    g.update_geometry()
    g.update_force_state()
    hip_sync(g)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k in range(steps):
        if not g.solve_step_static(ls, -1):
            raise RuntimeError(f"load step {k} did not converge")
    hip_sync(g)
    t = time.perf_counter() - t0
    alpha = g.get("state_var").reshape(-1, 7)[:, 0]
    out = _line("load steps/s, FCC 40^3 cells J2 single crystal to 2% strain", steps / t, "load_steps/s", steps, t,
                {"workload": f"FCC 40^3 cells, {len(xyz)} particles, J2 (E 200 GPa, sigma_y 300 MPa, H 2 GPa), "
                             f"uniaxial tension 2% in {steps} steps",
                 "bonds": g.size("nbonds"), "newton_iters_per_step": g.stat("newton_iters") / steps,
                 "cg_iters_per_step": g.stat("cg_iters") / steps, "newton_tol": args.tol_nr,
                 "predictor": args.predictor, "precond": args.precond, "cg_graph": args.cg_graph,
                 "spmv_launch": spmv_launch_info(g),
                 "coarse_reuse": args.coarse_reuse, "coarse_builds_per_step": g.stat("coarse_builds") / steps,
                 "setup_seconds": t_setup,
                 "yielded_fraction": float(np.mean(alpha > 0))})
    print(json.dumps(out), flush=True)


def config4_bench(args):
    """Hex plate of 400 x 400 (160,000 particles) with an edge crack of a quarter of the width at
    mid-height (bonds crossing the crack line removed), bond breaking at a critical stretch of
    1e-3, mode-I opening of 6e-4 per edge per step over 40 steps (Newton 1e-6, the headline FP32
    solver settings), then the adjoint d(top reaction)/d(Kn, Tv) at the final state with the broken
    set frozen (an FP64 polish step and one transposed solve). Both parts are timed."""
    import lpmgpu as G

    box = [0.0, 400.0, 0.0, 0.8660254037844386 * 400, 0.0, 1.0]
    xyz = G.generate_lattice(G.HEXAGON2D, box, 0.5)
    bp, bs, p2 = (np.asarray(a) for a in G.build_bonds(G.HEXAGON2D, 0.5, xyz))
    n = len(xyz)
    p1 = np.repeat(np.arange(n), np.diff(bp))
    lay = (np.arange(len(p2)) >= bs[p1]).astype(np.int32)
    y0 = 0.5 * (xyz[:, 1].min() + xyz[:, 1].max()) + 0.25
    a = xyz[:, 0].min() + 0.25 * np.ptp(xyz[:, 0])
    ya, yb = xyz[p1, 1], xyz[p2, 1]
    cross = (ya - y0) * (yb - y0) < 0
    tt = np.where(cross, (y0 - ya) / np.where(cross, yb - ya, 1.0), 0.0)
    keep = ~(cross & (xyz[p1, 0] + tt * (xyz[p2, 0] - xyz[p1, 0]) < a))
    lay, p1, p2 = lay[keep], p1[keep], p2[keep]
    bp2 = np.zeros(n + 1, dtype=np.int64)
    np.add.at(bp2, p1 + 1, 1)
    bp2 = np.cumsum(bp2).astype(np.int32)
    bs2 = (bp2[:-1] + np.bincount(p1[lay == 0], minlength=n)).astype(np.int32)
    g = G.Context(G.HEXAGON2D, 0.5, G.ELASTIC, xyz, bonds=(bp2, bs2, p2.astype(np.int32)))
    g.set_elastic(0.0, True, 69e3, 0.3, 1e-3, elmap_literal=0)
    # bonds break along the way, so the load path is not smooth: the quadratic predictor measured
    # 142 load steps/s (2.0 Newton per step) against 151 (1.7) for the linear one here
    pred4 = min(args.predictor, 1)
    g.mode(dirichlet_static=1, cg_transpose=0, diag10_literal=0, precision=0, precond=args.precond, cg_tol=args.cg_tol,
           cg_forcing=args.cg_forcing, tol_NR=1e-6, max_NR_iter=50, cg_maxit=50000, predictor=pred4,
           cg_graph=args.cg_graph, coarse_cells=args.coarse_cells, coarse_reuse=args.coarse_reuse)
    apply_modes(g, args)
    y = xyz[:, 1]
    lo, hi = np.nonzero(y < y.min() + 1e-6)[0], np.nonzero(y > y.max() - 1e-6)[0]
    inc, steps = 0.0006, 40
    ls = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", -inc), (hi, "x", 0.0), (hi, "y", inc)])
    g.update_geometry()
    g.update_force_state()
    hip_sync(g)
    t0 = time.perf_counter()
    for k in range(steps):
        if not g.solve_step_static(ls, -1):
            raise RuntimeError(f"load step {k} did not converge")
    hip_sync(g)
    t = time.perf_counter() - t0
    bd = g.get("bdamage")
    newton, cg = g.stat("newton_iters") / steps, g.stat("cg_iters") / steps
    spmv32 = spmv_launch_info(g)
    t0 = time.perf_counter()
    g.set_elastic(0.0, True, 69e3, 0.3, 1e9, elmap_literal=0)
    g.set("bdamage", bd)
    g.mode(precision=1, cg_tol=1e-12, tol_NR=5e-9, predictor=0)
    still = g.load_step(disp=[(lo, "x", 0.0), (lo, "y", 0.0), (hi, "x", 0.0), (hi, "y", 0.0)])
    if not g.solve_step_static(still, -1):
        raise RuntimeError("FP64 polish step did not converge")
    F, gKn, gTv, nref, rres = g.adjoint_reaction(g.group(hi), "y")
    t_adj = time.perf_counter() - t0
    out = _line("load steps/s, 400 x 400 hex plate with an edge crack, bond breaking, mode I", steps / t, "load_steps/s",
                steps, t,
                {"workload": f"hex 400 x 400, {n} particles, edge crack 0.25 W, critical stretch 1e-3, "
                             f"opening {inc} per edge per step, {steps} steps",
                 "bonds": g.size("nbonds"), "broken_bonds": int((bd > 0.5).sum()), "newton_iters_per_step": newton,
                 "cg_iters_per_step": cg, "newton_tol": 1e-6, "predictor": pred4, "precond": args.precond,
                 "cg_graph": args.cg_graph, "coarse_dofs": int(g.get_mode("ncoarse")), "coarse_reuse": args.coarse_reuse,
                 "spmv_launch": spmv32,
                 "adjoint": {"seconds_polish_and_adjoint": t_adj, "refinements": nref, "relative_residual": rres,
                             "top_reaction": F}})
    print(json.dumps(out), flush=True)


def hip_sync(ctx):
    # every lpm_* call that returns host values synchronises the context stream; a residual
    # update is a cheap way to fence without touching the state
    ctx.update_rr()


if __name__ == "__main__":
    main()
