"""Benchmark of the depth-kernel stage on one MI355X.

Workload (the production configuration of the solves/s metric): plain_4096 — 64x64 = 4,096
columns, 25 depth nodes 0-120 km refined to 49 layers, 30 log-spaced periods
1-60 s; one step = model prep + Rayleigh fundamental-mode root search (batched
persistent launch) + isotropic node kernels dc/dVs, dc/dVp, dc/drho + the TI
dc/dA, dc/dL splice Lsen_Gsc for every column x period, FP64.  Inputs are
resident in HBM before the timed region.

Prints one JSON line with value = column x period solves per second, the
roofline of the dominant kernel (FP64 VALU bound) and a CPU baseline: the
reference's own Fortran (flang build in oracle/_ref/): depthkernel with OpenMP
over columns plus depthkernel_ti_column in a process pool, on a bounded column
sample.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = ("col×period dispersion+kernel solves/s (FP64, |Δc|≤1e-5 km/s); "
          "batched vs plain")
PEAK_FP64 = 78.6e12          # MI355X FP64 vector (and dense matrix) peak, FLOP/s
FLOP_PER_LAYER_EVAL = 172    # FP64 ops per layer per secular-function evaluation
N_EV = 12                    # evaluations per root of the reference's bracket + refinement
N_K_ISO = 3                  # evaluation-equivalents of the isotropic kernel pass
# The reference's tregn96 per refined layer of one (column, period): 10,615
# arithmetic operations (complex*16 as real operations) + 156 transcendentals,
# counted statically over svfunc (up, down) and energy (getmat, six intijr)
# by oracle/count_tregn96_flops.py.  The TI kernel restates the same
# mathematics in far fewer operations (real arithmetic where the eigenvalues
# are real, the compound matrix in factored form), so its tregn96-equivalent
# rate can exceed the FP64 peak.
TI_OPS_PER_LAYER = 10771
N_VJP = 5                    # evaluation-equivalents of the VJP pass
PIPE = "FP64 VALU (vector ALU; no MFMA: nothing here is a dense contraction)"


def roofline_rec(flop, seconds, scope, **extra):
    """Algorithmic FP64 rate of `flop` operations in `seconds` against the
    MI355X FP64 vector peak."""
    a = flop / seconds
    return dict(bound="valu", achieved=a / 1e12, peak=PEAK_FP64 / 1e12, unit="TFLOP/s",
                frac=a / PEAK_FP64, pipe=PIPE, scope=scope, flop=flop, seconds=seconds, **extra)


def host_cpus():
    """(threads, description) of the CPU share this process may use: the CPUs of
    its affinity mask, limited by OMP_NUM_THREADS where the job sets it (the
    share a one-GPU job gets on a node), with nproc and the CPU model."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "").strip()
    n = aff
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:  # noqa: BLE001 - lscpu missing: fall back to /proc/cpuinfo
        pass
    if not model:
        try:
            with open("/proc/cpuinfo") as fh:
                for line in fh:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return n, dict(nproc=os.cpu_count(), affinity_cpus=aff, omp_num_threads=omp or None,
                   model=model)


def _ti_chunk(args):
    from oracle import port as P
    vel, periods, depz, minthk = args
    for v in vel:
        P.ti_column(v, periods, depz, minthk)
    return len(vel)


def _node_sample(name, ncols):
    from dazi_current_release_amd import configs
    if name == "voronoi_100":
        W = configs.voronoi_workload(1)
        return dict(vel=W["vel"][:ncols], depz=W["depz"], periods=W["periods"],
                    minthk=W["minthk"])
    return configs.node_workload(name, ncol=ncols)


def cpu_child_node(name, ncols, threads, ti):
    """The reference stage on the first ncols columns of a node workload:
    depthkernel (FD kernels, OpenMP over columns, CalSurfG.f90:86-107) and, with
    ti, the TI column routine in a pool of `threads` processes (tregn96 keeps
    global state, so it cannot run in threads)."""
    import multiprocessing as mp
    from oracle import port as P
    W = _node_sample(name, ncols)
    kind = "reference" if P.ref_available() else "port"
    lib = "ref" if kind == "reference" else "port"
    # the reference stage derives vp and rho from vs (Brocher); same layer and
    # period counts as the GPU workload
    t0 = time.perf_counter()
    P.depthkernel(W["vel"], W["periods"], W["depz"], W["minthk"], lib=lib, nthreads=threads)
    t_iso = time.perf_counter() - t0
    t_ti = 0.0
    if ti and kind == "reference":
        vel = W["vel"][:, 0, :]
        chunks = [(vel[i::threads], W["periods"], W["depz"], W["minthk"]) for i in range(threads)]
        ctx = mp.get_context("fork")
        with ctx.Pool(threads) as pool:
            t1 = time.perf_counter()
            pool.map(_ti_chunk, chunks)
            t_ti = time.perf_counter() - t1
    dt = t_iso + t_ti
    nsol = ncols * len(W["periods"])
    nz = W["vel"].shape[-1]
    what = (f"reference depthkernel, FD node kernels = {1 + 6 * nz} surfdisp96 curves per "
            f"column, OpenMP {t_iso:.1f} s")
    if ti and kind == "reference":
        what += (f" + reference depthkernel_ti_column (surfdisp96 + tregn96) in a process "
                 f"pool {t_ti:.1f} s")
    return dict(value=nsol / dt, seconds=dt, kind=kind, cols=ncols, threads=threads, solves=nsol,
                sample=f"first {ncols} columns of {name} ({nsol} column-period solves): {what}")


def cpu_child_layered(name, ncols, threads):
    """The reference's own layered forward driver
    (depthkernel_lab/surfdisp96_forward_batch.f90, built from source by
    oracle/Makefile) on the finite-difference model set of each of the first
    ncols columns (the model and 6 perturbed models per layer, as
    DepthKernel1D.f90:146-183 perturbs node models), one driver process per
    column, `threads` at a time: c and every layer's dc/dVs, dc/dVp, dc/drho the
    reference's way."""
    import tempfile
    import numpy as np
    from oracle import port as P
    from dazi_current_release_amd import configs
    W = configs.layered_workload(name, ncol=ncols)
    nl, nk = W["nl"], W["nk"]
    loff = np.concatenate([[0], np.cumsum(nl)[:-1]])
    toff = np.concatenate([[0], np.cumsum(nk)[:-1]])
    with tempfile.TemporaryDirectory() as d:
        files = []
        for j in range(ncols):
            sl = slice(loff[j], loff[j] + nl[j])
            f = os.path.join(d, f"col{j}.in")
            P.write_layered_fd_input(f, *(W[k][sl] for k in ("thk", "vp", "vs", "rho")),
                                     W["periods"][toff[j]:toff[j] + nk[j]])
            files.append(f)
        dt, _ = P.run_forward_batch(files, threads)
    nsol = int(nk.sum())
    curves = int(((1 + 6 * nl) * nk).sum())
    return dict(value=nsol / dt, seconds=dt, kind="reference", cols=ncols, threads=threads,
                solves=nsol,
                sample=(f"first {ncols} columns of {name} ({nsol} column-period solves): the "
                        f"reference's surfdisp96_forward_batch on each column's model and its "
                        f"6 L FD perturbations ({curves} period curves), one process per "
                        f"column, {threads} at a time"))


def cpu_baseline_child(spec: str, threads: int) -> None:
    """Child process of one CPU baseline: spec = node:<workload>:<ncols>:<ti> or
    layered:<workload>:<ncols>."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    parts = spec.split(":")
    if parts[0] == "node":
        rec = cpu_child_node(parts[1], int(parts[2]), threads, parts[3] == "1")
    else:
        rec = cpu_child_layered(parts[1], int(parts[2]), threads)
    sys.stdout.flush()
    print("CPUBASE " + json.dumps(rec), flush=True)


def run_cpu_baseline(spec: str, threads: int, cpuinfo: dict):
    """Runs one CPU baseline in a child process; returns the cpu_baseline record
    (or a record saying why it failed: the GPU figures stand on their own)."""
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", spec,
           "--threads", str(threads)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        for line in r.stdout.splitlines():
            if line.startswith("CPUBASE "):
                cb = json.loads(line[len("CPUBASE "):])
                return dict(value=cb["value"], unit="solves/s", cores=cb["threads"],
                            kind=cb["kind"], sample=cb["sample"], seconds=cb["seconds"],
                            host=cpuinfo)
        raise RuntimeError(f"rc={r.returncode} {r.stderr[-1500:]}")
    except Exception as e:  # noqa: BLE001
        return dict(value=None, unit="solves/s", cores=threads, kind="reference",
                    sample=f"{spec} failed: {e}", host=cpuinfo)


def _timed(torch, fn, n):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / n


def bench_c1(dk, configs, np, torch, cpu, reps=10):
    """plain_1024 (32 x 32 columns, 13 depth nodes 0-60 km, 18 periods 2-40 s, Vp =
    1.73 Vs): c and the isotropic node kernels of every column x period on the
    device, HIP-event-free wall clock per call with inputs resident."""
    W = configs.node_workload("plain_1024")
    nz = W["vel"].shape[-1]
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a.reshape(-1, nz, order="F").T),  # noqa
                                  device="cuda")
    vel, vp, rho = T(W["vel"]), T(W["vp"]), T(W["rho"])
    depz = torch.as_tensor(W["depz"], device="cuda")
    per = torch.as_tensor(W["periods"], device="cuda")
    out = dk.stage_dev(vel, depz, per, W["minthk"], vp=vp, rho=rho, iso=True, ti=False)
    t = _timed(torch, lambda: dk.stage_dev(vel, depz, per, W["minthk"], vp=vp, rho=rho, iso=True,
                                           ti=False, out=out), reps)
    ncol, K, L = vel.shape[1], per.numel(), 13
    ns = ncol * K
    rec = dict(workload="plain_1024", columns=ncol, nodes=nz, layers=L, periods=K,
               outputs="c + dc/dVs, dc/dVp, dc/drho node kernels", ms_per_step=1e3 * t,
               solves=ns, solves_per_s=ns / t,
               roofline=roofline_rec(ns * (N_EV + N_K_ISO) * FLOP_PER_LAYER_EVAL * L, t,
                                     "whole step (prep, root search, kernel pass)",
                                     flop_per_solve=(N_EV + N_K_ISO) * FLOP_PER_LAYER_EVAL * L))
    if cpu:
        rec["cpu_baseline"] = cpu("node:plain_1024:1024:0")
        rec["cpu_baseline"]["note"] = ("the reference's depthkernel derives Vp and rho from Vs "
                                       "(Brocher); same columns, layers and periods")
    return rec


def bench_vjp(dk, configs, np, torch, cpu, reps=5):
    """diff_4096 (4,096 columns x 40 layers x 30 periods, float64 models): the
    forward root search and the backward VJP kernel."""
    W = configs.layered_workload("diff_4096")
    b = dk.LayeredBatch(W["nl"], W["nk"], *(W[k].astype(np.float64)
                                            for k in ("thk", "vp", "vs", "rho")), W["periods"])
    out = dk.layered_dev(b)
    gc = torch.ones(b.nsolves, dtype=torch.float64, device="cuda")
    dk.layered_vjp(b, out["c_unrounded"], gc)
    t_fwd = _timed(torch, lambda: dk.layered_dev(b, out=out), reps)
    t_vjp = _timed(torch, lambda: dk.layered_vjp(b, out["c_unrounded"], gc), reps)
    ns, L = b.nsolves, 40
    step = t_fwd + t_vjp
    rec = dict(workload="diff_4096", columns=b.ncol, layers=L, periods=30,
               forward_ms=1e3 * t_fwd, vjp_ms=1e3 * t_vjp, vjp_over_forward=t_vjp / t_fwd,
               solves=ns, solves_per_s=ns / step,
               roofline=roofline_rec(ns * (N_EV + N_VJP) * FLOP_PER_LAYER_EVAL * L, step,
                                     "forward + VJP",
                                     forward_frac=ns * N_EV * FLOP_PER_LAYER_EVAL * L / t_fwd /
                                     PEAK_FP64,
                                     vjp_frac=ns * N_VJP * FLOP_PER_LAYER_EVAL * L / t_vjp /
                                     PEAK_FP64))
    if cpu:
        rec["cpu_baseline"] = cpu("layered:diff_4096:48")
        rec["cpu_baseline"]["note"] = ("the reference has no VJP: its gradient is the "
                                       "contraction of FD kernels, whose curves are timed")
    return rec


def bench_c3(dk, configs, np, torch, cpu, reps=5):
    """ragged_16384 (16,384 layered columns, 8-64 layers, 10-40 periods, c and
    refined-layer kernels): plain one launch per 64-column tile, one
    non-persistent launch over the whole batch, and the batched launch with
    the autotuned shape and mode (persistent or one launch), eager and
    replayed as a HIP graph.  Times are wall-clock per call with inputs
    resident on the device."""
    from dazi_current_release_amd import batched
    W = configs.layered_workload("ragged_16384")
    b = dk.LayeredBatch(W["nl"], W["nk"], W["thk"], W["vp"], W["vs"], W["rho"], W["periods"])
    out = b.alloc_outputs(True)
    run = lambda t: dk.layered_dev(b, kernels=True, tune=t, out=out)  # noqa: E731
    t_plain = _timed(torch, lambda: run(dk.make_tune(16, 4, 1, 64, 2e-6)), 2)
    t_one = _timed(torch, lambda: run(dk.make_tune(16, 4, 1, b.ncol, 2e-6)), reps)
    tune, table = batched.autotune(run)
    t_batch = _timed(torch, lambda: run(tune), reps)
    g = batched.GraphRunner(lambda: run(tune))
    t_graph = _timed(torch, g.replay, reps)
    ns = b.nsolves
    flop = (N_EV + N_K_ISO) * FLOP_PER_LAYER_EVAL * int((W["nl"].astype(np.int64) *
                                                         W["nk"]).sum())
    rec = dict(workload="ragged_16384", columns=b.ncol, solves=ns,
               solves_per_s=ns / t_batch,
               plain_solves_per_s=ns / t_plain, plain_launches=int(np.ceil(b.ncol / 64)),
               one_launch_solves_per_s=ns / t_one,
               batched_solves_per_s=ns / t_batch, graph_solves_per_s=ns / t_graph,
               batched_over_plain=t_plain / t_batch, graph_over_plain=t_plain / t_graph,
               batched_over_one_launch=t_one / t_batch,
               tuned=dict(lanes_per_scan=tune.lanes_per_scan, cols_per_wg=tune.cols_per_wg,
                          mode="persistent" if tune.mode == 0 else "one launch"),
               autotune_ms={f"{l}x{c}{'p' if m == 0 else 'o'}": round(ms, 4)
                            for l, c, m, ms in table},
               roofline=roofline_rec(flop, t_batch, "whole step, autotuned launch",
                                     flop_per_solve_mean=flop / ns))
    if cpu:
        rec["cpu_baseline"] = cpu("layered:ragged_16384:64")
    return rec


def bench_c5(dk, configs, np, torch, cpu, chunk=40960):
    """voronoi_100: 100 Voronoi realizations (9 layers x 100 Vs cells + 50 ANI cells
    per layer, projected onto 64x64 columns x 25 nodes) = 409,600 node columns x 30
    periods.  Per chunk of `chunk` columns one graph replay computes c, the
    isotropic node kernels and Lsen_Gsc, grad_c = c - c_obs and the
    effective-dVs VJP; the cell gradients are accumulated after each chunk.
    c_obs: the stage's c on an independent realization (the same map for every
    realization)."""
    from dazi_current_release_amd.realizations import RealizationStep, cell_gradient
    W = configs.voronoi_workload(100)
    N, nz = W["vel"].shape[0], W["vel"].shape[-1]
    vel = torch.as_tensor(np.ascontiguousarray(W["vel"][:, 0, :].T), device="cuda")
    depz = torch.as_tensor(W["depz"], device="cuda")
    per = torch.as_tensor(W["periods"], device="cuda")
    T = configs.voronoi_workload(1, seed=6)
    vt = torch.as_tensor(np.ascontiguousarray(T["vel"][:, 0, :].T), device="cuda")
    c_obs = dk.stage_dev(vt, depz, per, W["minthk"], iso=False, ti=False)["pv"].repeat(1, 100)
    lon = torch.as_tensor(W["layer_of_node"], device="cuda")
    cell = torch.as_tensor(W["cell"], dtype=torch.int64, device="cuda")
    acc = torch.zeros((100, 9, 100), dtype=torch.float64, device="cuda")
    torch.cuda.reset_peak_memory_stats()
    step = RealizationStep(nz, depz, per, W["minthk"], chunk)
    misfit = torch.zeros((), dtype=torch.float64, device="cuda")

    def consume(c0, n, s):
        misfit.add_(0.5 * (s.gc[:, :n] ** 2).sum())
        cell_gradient(s.grad["grad_vs"], lon, cell, 100, c0, n, 4096, acc)

    # the first pass uploads the graph and fills the allocator's cache; the
    # second is timed
    step.run(vel, c_obs, consume)
    torch.cuda.synchronize()
    acc.zero_()
    misfit.zero_()
    t0 = time.perf_counter()
    step.run(vel, c_obs, consume)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    pool = dk.lib()
    import ctypes
    nb = ctypes.c_int64(0)
    pool.dazi_dk_pool_bytes(ctypes.byref(nb))
    solves = N * per.numel()
    L = 49
    rec = dict(workload="voronoi_100", realizations=100, columns=N, solves=solves,
               seconds=dt, solves_per_s=solves / dt, chunk_columns=chunk,
               chunks=int(np.ceil(N / chunk)), outputs="c, node kernels, Lsen_Gsc, VJP",
               torch_peak_bytes=int(torch.cuda.max_memory_allocated()),
               library_pool_bytes=int(nb.value), misfit=float(misfit),
               cell_grad_norm=float(acc.norm()),
               roofline=roofline_rec(solves * ((N_EV + N_K_ISO) * FLOP_PER_LAYER_EVAL +
                                              TI_OPS_PER_LAYER) * L,
                                     dt, "whole population: stage + VJP + cell gradients",
                                     flop_per_solve=((N_EV + N_K_ISO) * FLOP_PER_LAYER_EVAL +
                                                     TI_OPS_PER_LAYER) * L))
    if cpu:
        rec["cpu_baseline"] = cpu("node:voronoi_100:192:1")
        rec["cpu_baseline"]["note"] = ("the reference has no VJP: the sample times its "
                                       "depthkernel and depthkernelTI on Voronoi columns")
    return rec


def bench_grows(dk, configs, np, torch, vel, out, nvx, nvz, reps=5):
    """G-row assembly on the plain_4096 stage outputs (64x64 columns: a 62x62
    interior grid, 24 depth rows, 30 periods): 1,000 synthetic straight rays per
    period, all 30,000 rays in one call (count, scan, fill).  Algorithmic HBM
    bytes per call: 12 B per emitted triplet + 16 B per (ray, cell) input (cell
    index and three weights); the stage arrays it gathers from (~140 MB) are
    re-read from cache across rays.  Timed with HIP events on the launch stream."""
    from dazi_current_release_amd import grows
    K = out["pv"].shape[0]
    R = configs.ray_workload(nvx, nvz, K, 1000, seed=7)
    rays = grows.rays_to_device(R)
    args = (rays, nvx, nvz, vel, out["sen_vs"], out["sen_vp"], out["sen_rho"], out["lsen"])
    g = grows.g_rows_dev(*args)
    nar = int(g["nar"].item())
    buf = dict(rw=g["rw"], iw=g["iw"], col=g["col"])
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    grows.g_rows_dev(*args, out=buf)
    e0.record(st)
    for _ in range(reps):
        grows.g_rows_dev(*args, out=buf)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    ncell = int(R["cell_off"][-1])
    nray = len(R["ray_k"])
    alg = 12 * nar + 16 * ncell
    return dict(workload="plain_4096 + 30,000 rays", rays=nray, ray_cells=ncell, triplets=nar,
                ms=ms, rays_per_s=nray / (ms * 1e-3), triplets_per_s=nar / (ms * 1e-3),
                roofline=dict(bound="hbm", achieved=alg / (ms * 1e-3) / 1e9, peak=8000.0,
                              unit="GB/s", frac=alg / (ms * 1e-3) / 8e12, bytes=alg))


def bench_azimuth(dk, np, torch, out, nx, ny, reps=20):
    """Period azimuthal-anisotropy map on the plain_4096 stage outputs (62 x 62
    interior cells, 24 layers, 30 periods) with random Gc, Gs models.
    Algorithmic HBM bytes per call: per (cell, period) 24 x 4 B of Lsen_Gsc, 8 B
    of tRcV and 20 B out, plus 2 x 24 x 4 B of Gc, Gs per cell."""
    from dazi_current_release_amd import azimuth
    nd, K, ncol = out["lsen"].shape
    cells = (nx - 2) * (ny - 2)
    g = torch.Generator(device="cuda").manual_seed(5)
    gcf = 0.03 * torch.randn((nd, ny - 2, nx - 2), device="cuda", generator=g)
    gsf = 0.03 * torch.randn((nd, ny - 2, nx - 2), device="cuda", generator=g)
    tRcV = out["pv"][:, :cells].contiguous()
    m = azimuth.azimuth_map_dev(out["lsen"], gcf, gsf, tRcV, nx, ny)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        azimuth.azimuth_map_dev(out["lsen"], gcf, gsf, tRcV, nx, ny, out=m)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    alg = cells * K * (nd * 4 + 8 + 20) + cells * nd * 8
    return dict(workload="plain_4096 azimuthal map", cells=cells, periods=K, layers=nd, ms=ms,
                roofline=dict(bound="hbm", achieved=alg / (ms * 1e-3) / 1e9, peak=8000.0,
                              unit="GB/s", frac=alg / (ms * 1e-3) / 8e12, bytes=alg))


def bench_travel(dk, configs, np, torch, out, nx, ny, threads, nsta=100, reps=3,
                 cpu_jobs=128):
    """First-arrival traveltimes on the plain_4096 stage's phase velocities
    (64 x 64 columns: a 306 x 306 propagation grid, 30 periods): nsta stations,
    every station a source at every period with the others as receivers, all
    jobs in one call (gridder, refined and coarse fast marching, receiver
    times).  Timed with HIP events on the launch stream.  The CPU figure is the
    oracle's restatement (bit-identical to the reference's routines) with
    OpenMP over jobs on a sample of cpu_jobs jobs, beside the reference's own
    routines run serially on 4 jobs."""
    from dazi_current_release_amd import travel
    K = out["pv"].shape[0]
    G = configs.travel_geometry(nx, ny, K, nsta, seed=11)
    d = travel.desc(nx, ny, 29.0, 98.0, 0.25, 0.25, kmax=K)
    plan = travel.TravelPlan(d, G["job_k"], G["scx"], G["scz"], G["rec_off"], G["rcx"], G["rcz"])
    pv = out["pv"]
    res = plan.run(pv)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        plan.run(pv, out=res)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    status = res["status"].cpu().numpy()
    njob, nrec = plan.njob, plan.nrec
    nnz, nnx = travel.grid_shape(d)
    rec = dict(workload=f"plain_4096 pvRc, {nsta} stations x {K} periods", jobs=njob,
               receivers=nrec, grid=[nnz, nnx], ms=ms,
               jobs_per_s=njob / (ms * 1e-3), receivers_per_s=nrec / (ms * 1e-3),
               node_updates_per_s=njob * nnx * nnz / (ms * 1e-3),
               failed_jobs=int((status != 0).sum()))
    plan.close()
    # the reference example's station layout (tests/golden/yunnan_travel.npz: its data
    # file's 1,469 records and 20,877 receivers on the 38 x 42 Yunnan grid, with the
    # example model's phase velocities)
    yt = os.path.join(HERE, "tests", "golden", "yunnan_travel.npz")
    if os.path.exists(yt):
        with np.load(yt, allow_pickle=False) as f:
            Y = {k: f[k] for k in f.files}
        GY = configs.stations_radians(Y)
        dy = travel.desc(38, 42, 29.0, 98.0, 0.25, 0.25, kmax=Y["pv"].shape[0])
        py = travel.TravelPlan(dy, GY["job_k"], GY["scx"], GY["scz"], GY["rec_off"], GY["rcx"],
                               GY["rcz"])
        pvy = torch.as_tensor(Y["pv"].astype(np.float64), device="cuda")
        ry = py.run(pvy)
        torch.cuda.synchronize()
        e0.record(st)
        for _ in range(reps):
            py.run(pvy, out=ry)
        e1.record(st)
        torch.cuda.synchronize()
        msy = e0.elapsed_time(e1) / reps
        same = bool(np.array_equal(ry["times"].cpu().numpy(), Y["times"]))
        rec["example_layout"] = dict(
            workload="reference example data file (Yunnan, 38 x 42 grid, 176 x 196 nodes)",
            jobs=py.njob, receivers=py.nrec, ms=msy, jobs_per_s=py.njob / (msy * 1e-3),
            receivers_per_s=py.nrec / (msy * 1e-3), identical_to_reference=same)
        py.close()
    try:
        from oracle import port as P
        od = P.fmm_desc(d.nvx, d.nvz, 29.0, 98.0, 0.25, 0.25)
        pvh = pv.cpu().numpy()
        n = min(cpu_jobs, njob)
        sel = np.arange(n)
        off = G["rec_off"][: n + 1]
        t0 = time.perf_counter()
        P.fmm_jobs(od, pvh, G["job_k"][sel], G["scx"][sel], G["scz"][sel], off,
                   G["rcx"][: off[-1]], G["rcz"][: off[-1]], nthreads=threads)
        t_port = time.perf_counter() - t0
        cpu = dict(value=n / t_port, unit="jobs/s", cores=threads, kind="port",
                   sample=f"first {n} jobs, oracle restatement, OpenMP over jobs")
        if P.ref_available():
            t0 = time.perf_counter()
            for q in range(4):
                a, b = G["rec_off"][q], G["rec_off"][q + 1]
                P.fmm_job(od, pvh[G["job_k"][q]], G["scx"][q], G["scz"][q], G["rcx"][a:b],
                          G["rcz"][a:b], lib="ref")
            cpu["reference_serial_jobs_per_s"] = 4 / (time.perf_counter() - t0)
        rec["cpu_baseline"] = cpu
        rec["over_cpu"] = rec["jobs_per_s"] / cpu["value"]
    except Exception as e:  # the GPU figure stands on its own
        rec["cpu_baseline"] = dict(value=None, sample=f"failed: {e}")
    return rec


def bench_host_path(dk, configs, np, reps=10):
    """The Fortran drop-in's host entry points on plain_4096 in the G builder's
    order: depthkernelTI then depthkernel on the same model
    (CalSurfGAniso_Joint.f90:478,486), both with the reference's Brocher vp and
    rho, through the C ABI with host arrays allocated once (as the Fortran
    caller's are): inputs copied in, all outputs copied back (the
    PCIe-inclusive rate)."""
    W = configs.node_workload("plain_4096")
    vel = np.asfortranarray(W["vel"])
    nx, ny, nz = vel.shape
    K = len(W["periods"])
    pv = np.zeros((nx * ny, K), order="F")
    sens = [np.zeros((nx * ny, K, nz), order="F") for _ in range(3)]
    tpv = np.zeros((nx * ny, K), order="F")
    ls = np.zeros((nx * ny, K, nz - 1), np.float32, order="F")

    def step():
        dk.depthkernelTI(nx, ny, nz, vel, tpv, 2, 0, K, W["periods"], W["depz"], W["minthk"], ls,
                         log=None)
        dk.depthkernel(nx, ny, nz, vel, pv, *sens, 2, 0, K, W["periods"], W["depz"],
                       W["minthk"], log=None)

    # two untimed calls (the first allocates the library's pinned staging and
    # device buffers; a single one left the next calls about 15 % slower)
    step()
    step()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    dt = (time.perf_counter() - t0) / reps
    ns = nx * ny * K
    out_bytes = pv.nbytes + tpv.nbytes + ls.nbytes + sum(x.nbytes for x in sens)
    return dict(workload="plain_4096", seconds_per_step=dt, solves_per_s=ns / dt,
                output_bytes=out_bytes,
                note="depthkernelTI then depthkernel on the same model (Brocher), host arrays; "
                     "the second call reuses the first one's root search; outputs copied back "
                     "through pinned staging")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-cols", type=int, default=640,
                    help="columns of the CPU baseline sample (about 10-20 s at 16 threads)")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-vjp", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--no-grows", action="store_true")
    ap.add_argument("--no-travel", action="store_true")
    ap.add_argument("--stations", type=int, default=100)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    ap.add_argument("--cpu-baseline-child", default="")
    ap.add_argument("--lanes", type=int, default=16)
    ap.add_argument("--cpw", type=int, default=16)
    args = ap.parse_args()

    if args.cpu_baseline_child:
        cpu_baseline_child(args.cpu_baseline_child, args.threads or host_cpus()[0])
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this stage shards columns across GPUs with no "
              "exchange, and multi-GPU runs are not part of this build (use --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_probe:
        n = args.copy_probe << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_probe_bytes=n, copies=5)), flush=True)
        return

    import dazi_current_release_amd as dk
    from dazi_current_release_amd import configs, _lib

    threads, cpuinfo = host_cpus()
    if args.threads:
        threads = args.threads
    cpu = None if args.no_cpu else (lambda spec: run_cpu_baseline(spec, threads, cpuinfo))

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    W = configs.node_workload("plain_4096")
    nz = W["vel"].shape[-1]
    ncol = W["nx"] * W["ny"]
    K = len(W["periods"])
    to_dev = lambda a: torch.as_tensor(  # noqa: E731
        np.ascontiguousarray(a.reshape(-1, nz, order="F").T), device=dev)
    vel, vp, rho = to_dev(W["vel"]), to_dev(W["vp"]), to_dev(W["rho"])
    depz = torch.as_tensor(W["depz"], device=dev)
    per = torch.as_tensor(W["periods"], device=dev)
    tune = dk.make_tune(args.lanes, args.cpw, 0, 64, 2e-6)
    stream = torch.cuda.current_stream()

    def step(out=None, t=tune):
        return dk.stage_dev(vel, depz, per, W["minthk"], vp=vp, rho=rho, iso=True, ti=True,
                            tune=t, out=out, stream=stream)

    out = step()
    for _ in range(args.warmup):
        out = step(out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step(out)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    nsolves = ncol * K
    value = nsolves / (dt / args.steps)

    # per-phase kernel durations from HIP events recorded on the launch stream
    _lib.lib().dazi_dk_enable_timing(1)
    ph = []
    for _ in range(args.steps):
        out = step(out)
        ph.append(dk.last_timings())
    _lib.lib().dazi_dk_enable_timing(0)
    ph = np.array(ph)
    prep_ms = float(ph[:, 0].mean())
    fwd_ms = float(ph[:, 1].mean())
    kern_ms = float(ph[:, 2].mean())
    ti_ms = float(ph[:, 3].mean())

    # accuracy guard of the measured output: zero-fill pattern and |dc| against
    # the oracle on a column subset
    status = out["status"].cpu().numpy()
    pv = out["pv"].cpu().numpy()

    L = 49
    # algorithmic FP64 work per solve: 172 ops per layer per evaluation-
    # equivalent of the isotropic path, tregn96's count per layer for TI
    phases = {"k_forward (root search)": (fwd_ms, N_EV * FLOP_PER_LAYER_EVAL * L),
              "k_kernels (isotropic adjoint)": (kern_ms, N_K_ISO * FLOP_PER_LAYER_EVAL * L),
              "k_ti (TI kernels)": (ti_ms, TI_OPS_PER_LAYER * L)}
    dom = max(phases, key=lambda k: phases[k][0])
    dom_ms, dom_fps = phases[dom]
    flop_dom = nsolves * dom_fps
    achieved = flop_dom / (dom_ms * 1e-3)
    flop_step = nsolves * sum(fps for _, fps in phases.values())
    # HBM bytes per launch of the dominant kernel: FETCH_SIZE / WRITE_SIZE
    # counter passes of this workload, scaled by a copy of known size
    traffic, traffic_src = None, None
    tj = os.path.join(HERE, "profiles", "k_forward_traffic.json")
    if dom.startswith("k_forward") and os.path.exists(tj):
        with open(tj) as fh:
            t = json.load(fh)
        if t.get("workload") == "plain_4096":
            traffic, traffic_src = t["bytes_per_launch"], t["source"]
    roofline = dict(bound="valu", achieved=achieved / 1e12, peak=PEAK_FP64 / 1e12,
                    unit="TFLOP/s", frac=achieved / PEAK_FP64, traffic=traffic,
                    traffic_source=traffic_src, kernel=dom, pipe=PIPE,
                    flop_per_launch=flop_dom, avg_launch_ms=dom_ms,
                    step_frac=flop_step / (ms_per_step * 1e-3) / PEAK_FP64,
                    phase_frac={k: nsolves * fps / (ms * 1e-3) / PEAK_FP64
                                for k, (ms, fps) in phases.items()},
                    phase_flop_per_solve={k: fps for k, (_, fps) in phases.items()},
                    phase_note=("k_ti's work is tregn96's operation count per layer "
                                "(oracle/count_tregn96_flops.py), which its restatement "
                                "evaluates with far fewer operations: a fraction above 1 "
                                "is the reference's work done faster than the FP64 peak "
                                "would allow it"))

    plain = None
    if not args.no_plain:
        tp = dk.make_tune(args.lanes, max(1, 64 // args.lanes), 1, 64, 2e-6)
        step(out, tp)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        n_plain = 2
        for _ in range(n_plain):
            step(out, tp)
        torch.cuda.synchronize()
        plain_s = (time.perf_counter() - t1) / n_plain
        plain = dict(solves_per_s=nsolves / plain_s, ms_per_step=1e3 * plain_s,
                     launches_per_step=int(np.ceil(ncol / 64)))

    c1 = None
    if not args.no_c1:
        c1 = bench_c1(dk, configs, np, torch, cpu)
    vjp = None
    if not args.no_vjp:
        vjp = bench_vjp(dk, configs, np, torch, cpu)
    c3 = None
    if not args.no_c3:
        c3 = bench_c3(dk, configs, np, torch, cpu)
    c5 = None
    if not args.no_c5:
        c5 = bench_c5(dk, configs, np, torch, cpu)
    host = None
    if not args.no_host:
        host = bench_host_path(dk, configs, np)
    grow = None
    if not args.no_grows:
        grow = bench_grows(dk, configs, np, torch, vel, out, W["nx"] - 2, W["ny"] - 2)
        grow["azimuth_map"] = bench_azimuth(dk, np, torch, out, W["nx"], W["ny"])
    trav = None
    if not args.no_travel:
        trav = bench_travel(dk, configs, np, torch, out, W["nx"], W["ny"], threads,
                            nsta=args.stations)

    cpu_main = None
    if cpu:
        cpu_main = cpu(f"node:plain_4096:{args.cpu_cols}:1")

    rec = dict(
        metric=METRIC, value=value, unit="solves/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=ms_per_step, higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f64", data="synthetic",
        config=dict(workload="plain_4096", columns=ncol, nodes=nz, layers=L, periods=K,
                    outputs="c + dc/dVs, dc/dVp, dc/drho node kernels + Lsen_Gsc (TI)",
                    mode="batched persistent", lanes_per_scan=args.lanes,
                    cols_per_wg=args.cpw,
                    phase_ms=dict(prep=prep_ms, root_search=fwd_ms, iso_kernels=kern_ms,
                                  ti_kernels=ti_ms),
                    failed_columns=int((status != 0).sum()),
                    zero_periods=int((pv == 0).sum()),
                    plain=plain,
                    batched_over_plain=(value / plain["solves_per_s"]) if plain else None,
                    step_roofline=roofline_rec(flop_step, ms_per_step * 1e-3,
                                               "whole step (prep, root search, kernels, TI)",
                                               flop_per_solve=flop_step // nsolves),
                    plain_1024=c1, vjp=vjp, ragged=c3, realizations=c5, host_path=host,
                    g_rows=grow, travel=trav),
        roofline=roofline, cpu_baseline=cpu_main)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
