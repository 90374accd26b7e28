"""FDTD3D throughput on one MI355X: Gcell-updates/s and HBM-roofline fraction.

``python bench.py [--steps K] [--warmup W] [--workload c3_512|c3_256|c2_128|c1_64|
c1_64_ref|adjoint_c4|closure_c5|closure_net|tm2d|snapshots_c3] [--mode fused|two_pass]``

A step is one full Yee time step (H half-step, then E half-step with CPML
psi, conductivity and the three dipoles) over the whole synthetic grid; the
default workload is the 512^3 Gaussian-random-field grid with a 16-cell CPML
(the roofline size).  Fields, coefficients and psi are resident in HBM before
the timed region.  The K timed steps run back to back on one stream; HIP
events around every H and E launch give each kernel's average duration.
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "FDTD3D Gcell-upd/s & %HBM roofline; fwd+adjoint s; closure speedup@equal err"


def build_workload(name: str, device):
    from jaxphys_amd import workloads as W
    if name == "c3_512":
        g, steps = W.c3_grf(512, device=device)
    elif name == "c3_256":
        g, steps = W.c3_grf(256, device=device)
    elif name == "c2_128":
        g, steps = W.c2_boxes(128)
    elif name == "c1_64":
        g, steps = W.c1_sphere(64)
    elif name == "c1_64_ref":
        g, steps = W.c1_sphere(64, reference_compat=True)
    else:
        raise SystemExit(f"unknown workload {name}")
    return g, steps


def cpu_baseline(solver, budget_s: float):
    """The float32 C/OpenMP restatement of the same scheme (the oracle), timed
    on this host's cores for about ``budget_s`` seconds of steps of the same
    workload (same grid, coefficients, CPML tables and sources)."""
    import numpy as np

    from oracle import c_oracle

    cb = solver.cb.cpu().numpy()
    ca = solver.ca.cpu().numpy() if solver.ca is not None else None
    cp = solver.cpml.cpu().numpy() if solver.cpml is not None else None
    # the source values do not change the cost of a step: the table is
    # repeated so that the sample can run for the whole budget
    tab = np.resize(solver.source_table, (max(4096, solver.source_table.shape[0]),
                                          solver.source_table.shape[1]))
    comps = solver.source_comps
    nx, ny, _ = solver.nx, solver.ny, solver.nz
    sidx = [(k * ny + j) * nx + i for i, j, k in solver.source_positions]
    ch = float(np.float32(solver.dt / (solver.mu0 * solver.dx)))
    cores = c_oracle.host_cores()
    steps, secs = c_oracle.time_arrays((solver.nx, solver.ny, solver.nz), cb, ca, cp,
                                       solver.pml_layers, ch, tab, comps, sidx, budget_s,
                                       threads=cores)
    cells = solver.nx * solver.ny * solver.nz
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(cells * steps / secs / 1e9, 4), "unit": "Gcell-upd/s",
            "cores": cores, "kind": "port",
            "sample": f"{steps} float32 steps of the same {solver.nx}x{solver.ny}x{solver.nz} "
                      f"workload in {secs:.1f} s (C/OpenMP oracle, {cpu_model})"}


def cpu_baseline_numpy(budget_s: float, max_steps: int = 500):
    """The NumPy float32 restatement of the reference's step (the oracle,
    oracle/ref_fdtd.py: one thread, array at a time, the closest analogue of
    the reference's XLA program) on the reference-compatible 64^3 sphere
    workload (10-layer sponge, sine source), timed for about ``budget_s``
    seconds of steps."""
    import numpy as np

    from jaxphys_amd import workloads as W
    from oracle import ref_fdtd as R

    n, dtype = 64, np.float32
    size, c = (n, n, n), n // 2
    dt = W.dt_for(1.0)
    I, J, K = np.indices(size)
    mats = [R.Mat(mask=(I - c) ** 2 + (J - c) ** 2 + (K - c) ** 2 <= 144, epsilon_r=4.0)]
    srcs = [R.Src(R.C0 / 20.0, (c + 8, c, c))]
    ch, ce = dt / (R.MU0 * 1.0), dt / (R.EPS0 * 1.0)
    eps_r = R.build_eps_r(size, mats, dtype)
    damp = R.damping_factor(R.sigma_avg(size, 10, 1.0, dtype), dt, dtype)
    fields = [np.zeros(size, dtype=dtype) for _ in range(6)]
    steps, t0 = 0, time.perf_counter()
    while True:
        fields = R.step(fields, eps_r, damp, "absorbing", ch, ce, dtype)
        vals, _ = R.source_values(srcs, steps, 0.0, dt, dtype)
        fields = R.add_sources(fields, srcs, vals)
        steps += 1
        secs = time.perf_counter() - t0
        if secs >= budget_s or steps >= max_steps:
            break
    return {"value": round(n ** 3 * steps / secs / 1e9, 5), "unit": "Gcell-upd/s", "cores": 1,
            "kind": "port",
            "sample": f"{steps} float32 steps of the reference-compatible 64^3 sphere workload "
                      f"in {secs:.1f} s (NumPy restatement of the reference's step, one thread)"}


def measured_traffic(workload: str, mode: str):
    """HBM bytes per step of this workload and kernel design, measured with
    rocprofv3 PMC counters (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by
    the same counters on a plain copy kernel) and committed as
    profiles/traffic.json; None when no measurement matches."""
    import json as _json

    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            rows = _json.load(fh)
    except (OSError, ValueError):
        return None, None
    for r in rows:
        if r.get("workload") == workload and r.get("mode") == mode:
            return r.get("bytes_per_step"), r
    return None, None


def adjoint_measure(dev, steps: int = 800, checkpoint_every: int = 100, record_box=False):
    """The adjoint workload: 160^3, 12-cell CPML, ``steps`` steps; 96^3
    design region eps_r U[1,4] in vacuum; z dipole 20 cells below it; J =
    time-integrated Ez^2 on a 32x32 patch 20 cells above it.  Times one
    plain forward run and one value + gradient (``record_box=False``:
    checkpoints every ``checkpoint_every`` steps, each segment recomputed
    while recording the design box, then its adjoint steps).  Returns the
    timings, the problem, the gradient and the inputs."""
    import torch

    from jaxphys_amd import workloads as W
    from jaxphys_amd.adjoint import AdjointProblem

    n = 160
    g, box, probes, eps_box = W.c4_design(n, dev)
    prob = AdjointProblem(g, steps, W.dt_for(1.0), box, probes,
                          checkpoint_every=checkpoint_every, record_box=record_box)
    # plain forward (the run being differentiated), timed alone
    plain = g.build_solver(W.dt_for(1.0), steps)
    eps_full = prob.background.clone()
    eps_full[prob.region] = eps_box.permute(2, 1, 0)
    plain.set_eps(eps_full)
    plain.run(0, 5)
    plain.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.run(0, steps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    del plain
    prob.value_and_grad(eps_box)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = prob.forward(eps_box)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    grad = prob.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    mode = ("design-box history recorded in the forward run" if prob.record_box else
            f"checkpoints every {prob.K} steps, segments recomputed")
    summary = {"ratio": round((t2 - t0) / t_fwd, 3), "forward_s": round(t_fwd, 4),
               "value_and_grad_s": round(t2 - t0, 4), "checkpoint_forward_s": round(t1 - t0, 4),
               "backward_s": round(t2 - t1, 4), "mode": mode,
               "checkpoint_every": prob.K, "record_box": prob.record_box,
               "grid": [n, n, n], "steps": steps, "design_box": list(box), "J": J}
    return summary, prob, grad, (g, box, probes, eps_box)


def adjoint_fd_float64(g, box, probes, eps_box, steps, cells):
    """Central differences (h = 1e-3) of the objective of the float64 C
    oracle (the checker) at design cells ``cells`` (box-relative)."""
    import numpy as np

    from jaxphys_amd import workloads as W
    from oracle import c_oracle as CO
    from oracle import ext_fdtd as X

    n = g.size[0]
    x0, y0, z0, bx, by, bz = box
    eps64 = np.ones((n, n, n))
    eps64[x0:x0 + bx, y0:y0 + by, z0:z0 + bz] = eps_box.cpu().numpy().astype(np.float64)
    tab = X.waveform_table("gaussian_sine", steps, t0=90, spread=30, wavelength=20,
                           courant=W.COURANT)
    zmon = probes[0][1][2]
    lo, hi = probes[0][1][0], probes[-1][1][0] + 1
    threads = CO.host_cores()

    def J64(eps):
        o = CO.COracle((n,) * 3, 1.0, W.dt_for(1.0), "cpml", eps, npml=g._pml_layers,
                       dtype=np.float64)
        o.set_sources([(2, g._sources[0].position, tab)])
        tot = 0.0
        for m in range(steps):
            o.run(m, 1, threads=threads)
            patch = o.f[2][zmon, lo:hi, lo:hi]
            tot += float((patch * patch).sum())
        return tot

    h = 1e-3
    out = []
    for cell in cells:
        ijk = (x0 + cell[0], y0 + cell[1], z0 + cell[2])
        ep, em = eps64.copy(), eps64.copy()
        ep[ijk] += h
        em[ijk] -= h
        out.append((J64(ep) - J64(em)) / (2 * h))
    return out


def adjoint_bench(args, dev):
    """The adjoint workload (``adjoint_measure``) in the mode asked for
    (default: checkpoints every 100 steps with recomputation), with central
    finite differences (h=1e-3) of the float64 oracle's objective on
    ``--fd-cells`` seeded design cells, reported per cell."""
    import numpy as np

    steps = args.steps or 800
    rec = {"auto": None, "on": True, "off": False}[args.record_box]
    summary, prob, grad, (g, box, probes, eps_box) = adjoint_measure(
        dev, steps, args.checkpoint_every, rec)
    gnp = grad.cpu().numpy()
    rng = np.random.default_rng(42)
    cells = [tuple(int(rng.integers(0, s)) for s in eps_box.shape) for _ in range(args.fd_cells)]
    fd = adjoint_fd_float64(g, box, probes, eps_box, steps, cells) if cells else []
    samples = [{"cell": c, "adjoint": float(gnp[c]), "fd_float64": f,
                "rel_err": abs(float(gnp[c]) - f) / abs(f) if f != 0 else None}
               for c, f in zip(cells, fd)]
    scale = float(np.abs(gnp).max())
    return {
        "metric": "adjoint dJ/deps_r cost (x one forward run)",
        "value": summary["ratio"], "unit": "x forward", "higher_is_better": False,
        "n_gpus": 1, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "adjoint_c4", "grid": summary["grid"], "steps": steps, "npml": 12,
                   "design_box": list(box), "checkpoint_every": summary["checkpoint_every"],
                   "record_box": summary["record_box"], "probes": len(probes)},
        "adjoint": summary,
        "fd_check": {"h": 1e-3, "oracle": "float64 C oracle", "cells": len(samples),
                     "max_rel_err": max((x["rel_err"] for x in samples
                                         if x["rel_err"] is not None), default=None),
                     "max_abs_err_over_max_grad": max(
                         (abs(x["adjoint"] - x["fd_float64"]) / scale for x in samples),
                         default=None),
                     "samples": samples},
    }


def snapshots_bench(args, dev):
    """The reference API's saved history at the roofline size: the 512^3
    Gaussian-random-field CPML workload through ``EMGrid3D.simulate``, 100
    steps with every 20th step saved (5 snapshots of 6 fields, 16.1 GB),
    kept in HBM (``snapshots="device"``) or streamed into pinned host memory
    while the run continues (``snapshots="host"``), against the same call
    saving only step 0.  Wall time of the whole call (set-up included, the
    same in all three)."""
    import torch

    from jaxphys_amd import workloads as W

    n = 512
    steps, every = args.steps or 100, 20
    grid, _ = W.c3_grf(n, device=dev)
    dt = W.dt_for(1.0)

    def timed(save_every, mode):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hist = grid.simulate(dt=dt, save_every=save_every, n_steps=steps, snapshots=mode)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        nsnap = int(hist.ex.shape[0])
        del hist
        return secs, nsnap

    timed(steps, "device")  # warm-up: kernels, allocator
    bare, _ = timed(steps + 1, "device")
    on_dev, nsnap = timed(every, "device")
    torch.cuda.empty_cache()
    # the first host call pins its history buffers; torch's caching host
    # allocator hands the same pinned blocks to the next call
    first_host, _ = timed(every, "host")
    on_host, nsnap_h = timed(every, "host")
    # the plain device-to-host copy rate of one snapshot-sized buffer
    a = torch.empty(n ** 3, dtype=torch.float32, device=dev)
    b = torch.empty(n ** 3, dtype=torch.float32, pin_memory=True)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(4):
        b.copy_(a, non_blocking=True)
    torch.cuda.synchronize()
    d2h = 4 * 4 * n ** 3 / (time.perf_counter() - t0) / 1e9
    del a, b
    snap_bytes = 6 * 4 * n ** 3
    return {
        "metric": "FDTD3D history streaming (saved steps to pinned host memory)",
        "value": round(nsnap_h * snap_bytes / on_host / 1e9, 2),
        "unit": "GB/s of history delivered to host memory (whole call)",
        "higher_is_better": True, "n_gpus": 1, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "snapshots_c3", "grid": [n, n, n], "steps": steps,
                   "save_every": every, "snapshots": nsnap, "bytes_per_snapshot": snap_bytes},
        "wall_s": {"no_history": round(bare, 3), "device_history": round(on_dev, 3),
                   "host_history": round(on_host, 3),
                   "host_history_first_call": round(first_host, 3)},
        "overhead_vs_no_history": {"device": round(on_dev / bare, 3),
                                   "host": round(on_host / bare, 3)},
        "d2h_copy_GBps": round(d2h, 1),
        "copy_bound_s": round(nsnap_h * snap_bytes / (d2h * 1e9), 3),
    }


def tm2d_bench(args, dev):
    """2D TM grid (the reference's EMGrid, src/neurosim/em/fdtd.py) at
    16384 x 16384: absorbing boundary with 10 damped layers, a conducting
    wall with a 400-cell slit across the middle and two plane-wave rows.
    One fused kernel per step; HIP events around every launch."""
    import numpy as np
    import torch

    from jaxphys_amd import EMGrid, PlaneWave, Wall

    n = args.tm2d_n
    K = args.steps if args.steps is not None else 200
    Wm = args.warmup
    g = EMGrid(size=(n, n), resolution=0.01, boundary="absorbing", pml_layers=10)
    g.add_source(PlaneWave(frequency=3e9, y=n // 8, amplitude=1.0))
    g.add_source(PlaneWave(frequency=5e9, y=n // 4, amplitude=0.5))
    g.add_conductor(Wall(y=n // 2, gap_start=n // 2 - 200, gap_end=n // 2 + 200))
    dt = 0.99 * 0.01 / (2.99792458e8 * float(np.sqrt(np.float32(2.0))))
    t_setup = time.time()
    sol = g.build_solver(dt, K + Wm)
    sol.run(0, Wm)
    torch.cuda.synchronize()
    setup_s = time.time() - t_setup
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(K + 1)]
    ev[0].record(stream)
    for q in range(K):
        sol.run(Wm + q, 1, stream)
        ev[q + 1].record(stream)
    torch.cuda.synchronize()
    total_ms = ev[0].elapsed_time(ev[K])
    launch_ms = total_ms / K
    cells = n * n
    nbytes = sol.bytes_per_step()
    achieved = nbytes / (launch_ms / 1e3) / 1e9
    finite = bool(all(torch.isfinite(f).all().item() for f in sol.fields))
    cpu = None
    if args.cpu_budget > 0:
        cpu = tm2d_cpu_baseline(g, sol, dt, args.cpu_budget)
    traffic, traffic_rec = measured_traffic("tm2d", "fused")
    return {
        "metric": "FDTD2D TM Gcell-upd/s & %HBM roofline", "value": round(cells / launch_ms / 1e6, 3),
        "unit": "Gcell-upd/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": round(launch_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "tm2d", "grid": [n, n], "boundary": "absorbing", "npml": 10,
                   "walls": 1, "sources": 2},
        "roofline": {"bound": "hbm", "kernel": "tm2d_step_kernel<0>", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": None if traffic_rec is None else traffic_rec.get("source"),
                     "bytes_per_launch": nbytes, "ms_per_launch": round(launch_ms, 4)},
        "cpu_baseline": cpu, "fields_finite": finite, "setup_s": round(setup_s, 2),
    }


def tm2d_cpu_baseline(grid, sol, dt, budget_s):
    """The NumPy float32 restatement of the 2D step (the oracle), one
    thread, on the same grid, tables and sources, for about ``budget_s``."""
    import numpy as np

    from oracle import ref_tm2d as R

    nx, ny = grid.size
    f32 = np.float32
    waves = [R.Wave(s.frequency, s.y, s.amplitude) for s in grid._sources]
    walls = [R.Wall(w.y, w.gap_start, w.gap_end) for w in grid._conductors]
    mask = R.conductor_mask(nx, ny, walls)
    damp = R.damping(nx, ny, grid._config.pml_layers, grid._dx, dt, f32)
    vals = R.source_values(waves, 4096, 0.0, dt, f32)
    ez, hx, hy = (np.zeros((nx, ny), f32) for _ in range(3))
    ch, ce = dt / (R.MU0 * grid._dx), dt / (R.EPS0 * grid._dx)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < 4096:
        ez, hx, hy = R.step(ez, hx, hy, steps, boundary="absorbing", ch=ch, ce=ce, damp=damp,
                            mask=mask, src_vals=vals, waves=waves, dtype=f32)
        steps += 1
    secs = time.perf_counter() - t0
    return {"value": round(nx * ny * steps / secs / 1e9, 4), "unit": "Gcell-upd/s", "cores": 1,
            "kind": "port",
            "sample": f"{steps} float32 steps of the same {nx}x{ny} workload in {secs:.1f} s "
                      "(NumPy oracle, one thread)"}


def closure_net_bench(args, dev):
    """The closure net's BF16-MFMA conv kernels at the coarse 128^3 grid:
    one inference (three implicit-GEMM 3^3 convolutions, 7->32->32->6
    channels, GELU) per step, inputs resident in HBM."""
    import torch

    from jaxphys_amd.closure import CIN, ClosureNet

    n = args.closure_n // 2
    shape = (n, n, n)
    cells = n ** 3
    net = ClosureNet(dev, seed=0, out_scale=1.0)
    x = torch.randn(cells, CIN, device=dev).to(torch.bfloat16)
    x[:, 7] = 0
    for _ in range(args.warmup):
        net.forward(x, shape)
    torch.cuda.synchronize()
    K = args.steps if args.steps is not None else 50
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(K):
        net.forward(x, shape)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / K
    flops = 2.0 * cells * 27 * (7 * 32 + 32 * 32 + 32 * 6)
    tflops = flops / (ms / 1e3) / 1e12
    # the float32 far-field operator applied with the net (27 taps x 6
    # fields x 12 outputs = 1944 FMAs per cell)
    from jaxphys_amd import _conv
    fields = [torch.randn(shape, device=dev) for _ in range(6)]
    eps = 1.0 + torch.rand(shape, device=dev)
    sc = torch.ones(6, device=dev)
    w = 0.01 * torch.randn(27 * 72, device=dev)
    y = torch.zeros(cells, 8, device=dev)
    for _ in range(3):
        _conv.stencil_f32((n, n, n), fields, eps, sc, w, y, 8)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(K):
        _conv.stencil_f32((n, n, n), fields, eps, sc, w, y, 8)
    e1.record(stream)
    torch.cuda.synchronize()
    st_ms = e0.elapsed_time(e1) / K
    st_tflops = 2.0 * 1944 * cells / (st_ms / 1e3) / 1e12
    # one training step of the net (the trainer's work per sample): forward
    # keeping the activations, backward (input gradients through the
    # flipped-weight forward kernel, weight gradients), on a 64^3 crop and
    # on the whole 128^3 grid
    train = {}
    for m in (64, n):
        shp = (m, m, m)
        xt = torch.randn(m ** 3, CIN, device=dev).to(torch.bfloat16)
        xt[:, 7] = 0
        gy = torch.randn(m ** 3, 8, device=dev) * 1e-3
        for _ in range(3):
            net.forward(xt, shp, train=True)
            net.backward(gy)
        torch.cuda.synchronize()
        reps = max(5, K // 5)
        e0.record(stream)
        for _ in range(reps):
            net.forward(xt, shp, train=True)
            net.backward(gy)
        e1.record(stream)
        torch.cuda.synchronize()
        t_ms = e0.elapsed_time(e1) / reps
        fl = 3.0 * 2.0 * m ** 3 * 27 * (7 * 32 + 32 * 32 + 32 * 6)
        train[f"{m}^3"] = {"ms": round(t_ms, 4), "TFLOP/s": round(fl / (t_ms / 1e3) / 1e12, 2),
                           "frac_of_bf16_peak": round(fl / (t_ms / 1e3) / 1e12 / 2500.0, 4)}
    return {
        "metric": "closure net inference (3 conv3d layers, 128^3)", "value": round(tflops, 2),
        "unit": "TFLOP/s", "higher_is_better": True, "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic", "scaling": "weak", "vs_baseline": None,
        "config": {"workload": "closure_net", "grid": [n] * 3, "channels": [7, 32, 32, 6],
                   "kernel": "3x3x3"},
        "roofline": {"bound": "mfma", "kernel": "conv_roll8_kernel + 2 conv_slab_kernel launches per inference",
                     "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tflops / 2500.0, 4), "traffic": None},
        "training_step": train,
        "far_field_stencil": {"kernel": "stencil4_f32_kernel", "ms": round(st_ms, 4),
                              "TFLOP/s": round(st_tflops, 2),
                              "frac_of_fp32_vector_peak": round(st_tflops / 78.6, 4)},
    }


CLOSURE_FILE = os.path.join(ROOT, "jaxphys_amd", "data", "closure_c5.safetensors")


def closure_summary(net, dev, seeds, nf: int = 256, every: int = 10, **pkw):
    """Evaluate a trained closure on held-out seeds (fine nf^3 dx=1 vs coarse
    (nf/2)^3 dx=2, 400 coarse steps): mean / max interior field error of E
    at the final step against the restricted fine run, the far-field error
    (cells more than 3 coarse cells from the source), the same for plain
    coarse FDTD and for the linear parts alone, and the wall-time ratio of
    the fine run to the corrected coarse run."""
    import numpy as np

    from jaxphys_amd import closure as CL

    res = [CL.evaluate_closure(net, s, n_fine=nf, n_coarse=nf * 400 // 256, every=every, **pkw)
           for s in seeds]

    def col(k):
        return np.array([r.get(k, float("nan")) for r in res], dtype=np.float64)

    far_c, far_0 = col("err_far_closure"), col("err_far_coarse")
    return {
        "seeds": [int(s) for s in seeds],
        "speedup_vs_fine": round(float(np.mean(col("speedup"))), 3),
        "err": round(float(np.mean(col("err_closure"))), 5),
        "err_max": round(float(np.max(col("err_closure"))), 5),
        "err_far_median": round(float(np.median(far_c)), 5),
        "err_far_max": round(float(np.max(far_c)), 5),
        "coarse_err": round(float(np.mean(col("err_coarse"))), 5),
        "coarse_err_far_median": round(float(np.median(far_0)), 5),
        "seeds_far_worse_than_coarse": int(np.sum(far_c > far_0)),
        "linear_parts_err_far_median": round(float(np.median(col("err_far_far_field"))), 5),
        "net_gain": net.net_gain, "far_gain": net.far_gain,
        "t_fine_s": round(float(np.mean(col("t_fine_s"))), 4),
        "t_closure_s": round(float(np.mean(col("t_closure_s"))), 4),
        "per_seed": [{"seed": r["seed"], "err": round(r["err_closure"], 5),
                      "err_far": round(r["err_far_closure"], 5),
                      "err_far_coarse": round(r["err_far_coarse"], 5),
                      "err_far_linear": round(r.get("err_far_far_field", float("nan")), 5)}
                     for r in res],
        "_res": res,
    }


def closure_bench(args, dev):
    """Learned closure: train on fine-run/coarse pairs, evaluate on held-out
    seeds; reports the corrected coarse error and the wall-time ratio."""
    import numpy as np

    from jaxphys_amd import closure as CL

    nf = args.closure_n
    train_seeds = list(range(args.closure_train))
    test_seeds = list(range(1000, 1000 + args.closure_test))
    t0 = time.time()
    pkw = {"t0": args.closure_t0, "spread": args.closure_spread,
           "waveform": args.closure_waveform}
    if args.closure_load:
        net, hist = CL.ClosureNet.load(args.closure_load, dev), []
        net.source_patch = args.closure_source == "patch"
        net.patch_half = args.closure_patch_half
        if args.closure_patch_tail is not None:
            net.patch_tail = args.closure_patch_tail
    else:
        net, hist = CL.train_closure(dev, train_seeds, n_fine=nf, n_coarse=nf * 400 // 256,
                                 every=args.closure_every, epochs=args.closure_epochs,
                                 crop=min(64, nf // 2 - 16), lr=args.closure_lr,
                                 lr_final=args.closure_lr_final,
                                 onpolicy_from=(None if args.closure_onpolicy_from < 0
                                                else args.closure_onpolicy_from),
                                 onpolicy_with_net=args.closure_onpolicy_with_net,
                                 forcing=not args.closure_no_forcing,
                                 seed=args.closure_seed,
                                 far_field=not args.closure_no_far_field,
                                 val_seeds=list(range(900, 900 + args.closure_val)),
                                 warm_start=args.closure_warm_start,
                                 net_in_box=args.closure_net_in_box,
                                 source_patch=args.closure_source == "patch",
                                 patch_half=args.closure_patch_half,
                                 patch_tail=(CL.PATCH_TAIL if args.closure_patch_tail is None
                                             else args.closure_patch_tail),
                                 log=lambda m: print("[closure]", m, file=sys.stderr, flush=True),
                                 **pkw)
    val_seeds = list(range(900, 900 + args.closure_val))
    rollout = {}
    if args.closure_rollout_epochs > 0:
        # the net trained through the corrected rollouts (backpropagation
        # through the coarse steps), its gain chosen on its own validation seeds
        val_seeds = list(range(900, 900 + args.closure_rollout_val))
        rollout = {"rollout_epochs": args.closure_rollout_epochs,
                   "rollout_lr": args.closure_rollout_lr,
                   "rollout_window": args.closure_rollout_window}
        CL.train_rollout(net, dev, train_seeds, n_fine=nf, n_coarse=nf * 400 // 256,
                         every=args.closure_every, epochs=args.closure_rollout_epochs,
                         lr=args.closure_rollout_lr, lr_final=0.1 * args.closure_rollout_lr,
                         window=args.closure_rollout_window, seed=args.closure_seed,
                         val_seeds=val_seeds,
                         log=lambda m: print("[closure]", m, file=sys.stderr, flush=True), **pkw)
    t_train = time.time() - t0
    if args.closure_save:
        os.makedirs(os.path.dirname(os.path.abspath(args.closure_save)), exist_ok=True)
        info = {"workload": "closure_c5", "n_fine": nf, "train_seeds": train_seeds,
                "epochs": args.closure_epochs, "every": args.closure_every,
                "seed": args.closure_seed, "train_s": round(t_train, 1), **pkw}
        if args.closure_load:
            # the linear parts come from the loaded file: keep its record,
            # with what was trained here added
            info = dict(CL.ClosureNet.info(args.closure_load))
            if rollout:
                info["rollout_train_s"] = round(t_train, 1)
        info.update(rollout)
        info["val_seeds"] = val_seeds
        net.save(args.closure_save, info=info)
    summ = closure_summary(net, dev, test_seeds, nf, args.closure_every, **pkw)
    res = summ.pop("_res")
    err_c = [r["err_closure"] for r in res]
    err_0 = [r["err_coarse"] for r in res]
    trj_c = [r["err_traj_closure"] for r in res]
    trj_0 = [r["err_traj_coarse"] for r in res]
    sp = [r["speedup"] for r in res]
    return {
        "metric": "closure: coarse+net field error vs fine FDTD",
        "value": round(float(np.mean(err_c)), 5), "unit": "rel-L2 (E, interior)",
        "higher_is_better": False, "n_gpus": 1, "dtype": "bf16 net / f32 FDTD",
        "data": "synthetic",
        "config": {"workload": "closure_c5", "fine": [nf] * 3, "coarse": [nf // 2] * 3,
                   "train_seeds": len(train_seeds), "test_seeds": len(test_seeds),
                   "epochs": args.closure_epochs, "correct_every": args.closure_every,
                   "pulse_t0": args.closure_t0, "pulse_spread": args.closure_spread,
                   "pulse": args.closure_waveform, "forcing": not args.closure_no_forcing,
                   "train_seed": args.closure_seed,
                   "far_field": not args.closure_no_far_field, "far_gain": net.far_gain,
                   "val_seeds": args.closure_val, "net_gain": net.net_gain,
                   "warm_start": args.closure_warm_start,
                   "net_in_box": net.net_in_box,
                   "source_correction": "near-source patch" if net.source_patch else
                   "linear source model", "patch_half": net.patch_half,
                   "patch_tail": net.patch_tail,
                   "onpolicy_from_epoch": args.closure_onpolicy_from,
                   "onpolicy_with_net": args.closure_onpolicy_with_net,
                   "lr": [args.closure_lr, args.closure_lr_final]},
        "err_coarse_no_closure": round(float(np.mean(err_0)), 5),
        "err_far_closure": round(float(np.mean([r["err_far_closure"] for r in res])), 5),
        "source_model_only": {k: round(float(np.mean([r.get(f"{k}_source_only", float("nan"))
                                                      for r in res])), 5)
                              for k in ("err", "err_far", "err_traj", "err_traj_far")},
        "source_model_and_far_field": {k: round(float(np.mean([r.get(f"{k}_far_field",
                                                                     float("nan"))
                                                               for r in res])), 5)
                                       for k in ("err", "err_far", "err_traj", "err_traj_far")},
        "median_err_far": {k: round(float(np.median([r.get(f"err_far_{k}", float("nan"))
                                                     for r in res])), 5)
                           for k in ("coarse", "source_only", "far_field", "closure",
                                     "net_gain1")},
        "with_net_gain1": {k: round(float(np.mean([r.get(f"{k}_net_gain1", float("nan"))
                                                   for r in res])), 5)
                           for k in ("err", "err_far", "err_traj", "err_traj_far")},
        "err_far_coarse_no_closure": round(float(np.mean([r["err_far_coarse"] for r in res])), 5),
        "err_traj_far_closure": round(float(np.mean([r["err_traj_far_closure"] for r in res])), 5),
        "err_traj_far_coarse_no_closure": round(float(np.mean([r["err_traj_far_coarse"]
                                                               for r in res])), 5),
        "err_traj_closure": round(float(np.mean(trj_c)), 5),
        "err_traj_coarse_no_closure": round(float(np.mean(trj_0)), 5),
        "err_closure_max": round(float(np.max(err_c)), 5),
        "per_seed": [{"seed": r["seed"], "err": round(r["err_closure"], 5),
                      "err_source_only": round(r.get("err_source_only", float("nan")), 5),
                      "err_far": round(r["err_far_closure"], 5),
                      "err_far_source_only": round(r.get("err_far_source_only", float("nan")), 5),
                      "err_far_far_field": round(r.get("err_far_far_field", float("nan")), 5),
                      "err_far_coarse": round(r["err_far_coarse"], 5)}
                     for r in res],
        "speedup_vs_fine": round(float(np.mean(sp)), 3),
        "t_fine_s": round(float(np.mean([r["t_fine_s"] for r in res])), 4),
        "t_closure_s": round(float(np.mean([r["t_closure_s"] for r in res])), 4),
        "t_coarse_s": round(float(np.mean([r["t_coarse_s"] for r in res])), 4),
        "train_s": round(t_train, 1), "final_loss": float(np.mean(hist[-40:])) if hist else None,
        "summary": summ,
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c3_512")
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU work for the cpu_baseline sample (0 = skip)")
    ap.add_argument("--mode", default="fused", choices=["fused", "two_pass"],
                    help="one fused launch per step (default) or the H and E half-step kernels")
    ap.add_argument("--checkpoint-every", type=int, default=100)
    ap.add_argument("--record-box", default="off", choices=["auto", "on", "off"],
                    help="adjoint: recompute the design-box history from checkpoints (off, "
                         "the configured mode) or record it in the forward run (on; auto: "
                         "when it fits)")
    ap.add_argument("--fd-cells", type=int, default=16)
    ap.add_argument("--closure-n", type=int, default=256)
    ap.add_argument("--closure-train", type=int, default=64)
    ap.add_argument("--closure-test", type=int, default=16)
    ap.add_argument("--closure-epochs", type=int, default=2)
    ap.add_argument("--closure-lr", type=float, default=1e-3)
    ap.add_argument("--closure-lr-final", type=float, default=1e-4,
                    help="learning rate at the end of the cosine schedule")
    ap.add_argument("--closure-onpolicy-with-net", action="store_true",
                    help="on-policy samples come from rollouts corrected by the whole closure, "
                    "the net included")
    ap.add_argument("--closure-onpolicy-from", type=int, default=-1,
                    help="epoch from which training uses the corrected rollout's own states "
                         "(-1: teacher-forced only)")
    ap.add_argument("--closure-seed", type=int, default=0,
                    help="seed of the net's initial weights and of the crop sampling")
    ap.add_argument("--closure-warm-start", action="store_true",
                    help="teacher-forced samples start one corrected interval earlier")
    ap.add_argument("--closure-val", type=int, default=2,
                    help="validation seeds (900, 901, ...) that choose the net's gain")
    ap.add_argument("--closure-rollout-epochs", type=int, default=0,
                    help="closure_c5: epochs of training the net through the corrected "
                    "rollouts (backpropagation through the coarse steps; fresh net weights)")
    ap.add_argument("--closure-rollout-lr", type=float, default=6e-5,
                    help="rollout training: Adam learning rate (cosine to a tenth of it)")
    ap.add_argument("--closure-rollout-window", type=int, default=10,
                    help="rollout training: correction intervals per update")
    ap.add_argument("--closure-rollout-val", type=int, default=8,
                    help="rollout training: validation seeds (900, ...) that choose the gain")
    ap.add_argument("--closure-net-in-box", action="store_true",
                    help="the net also corrects the source boxes (on top of the source model)")
    ap.add_argument("--closure-source", default="patch", choices=["patch", "linear"],
                    help="correction of the source boxes: the near-source patch (local "
                    "fine/coarse grids) or the least-squares linear source model")
    ap.add_argument("--closure-patch-half", type=int, default=16,
                    help="half side (coarse cells) of the near-source patch's local grids")
    ap.add_argument("--closure-patch-tail", type=int, default=None,
                    help="near-source patch: intervals computed after the source's last "
                    "(-1: all; default: the closure's own setting)")
    ap.add_argument("--closure-no-far-field", action="store_true",
                    help="no linear far-field operator (net and source model only)")
    ap.add_argument("--closure-no-forcing", action="store_true",
                    help="no forcing input channel (fields and eps_r only)")
    ap.add_argument("--closure-save", default=None,
                    help="closure_c5: write the trained closure to this safetensors file")
    ap.add_argument("--closure-load", default=None,
                    help="closure_c5: evaluate this saved closure instead of training one")
    ap.add_argument("--closure-every", type=int, default=10,
                    help="coarse steps between closure corrections")
    ap.add_argument("--closure-t0", type=float, default=60.0,
                    help="pulse centre (coarse steps)")
    ap.add_argument("--closure-spread", type=float, default=15.0,
                    help="pulse width (coarse steps)")
    ap.add_argument("--closure-waveform", default="gaussian",
                    choices=["gaussian", "gaussian_derivative"])
    ap.add_argument("--tm2d-n", type=int, default=16384, help="tm2d workload: cells per side")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload: leave out the adjoint and closure objects")
    ap.add_argument("--closure-seeds", type=int, default=16,
                    help="default workload: held-out seeds (1000, 1001, ...) of the closure object")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain copy of the field bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; multi-GPU runs are not implemented "
              "(a z-slab decomposition with one halo exchange per half-step is the design for more)", file=sys.stderr)
        raise SystemExit(2)

    import numpy as np
    import torch

    from jaxphys_amd import _native
    from jaxphys_amd import workloads as W

    dev = _native.require_device()
    torch.cuda.set_device(dev)

    if args.copy_probe:
        n = 512**3
        a = torch.randn(n, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            _native.copy_probe(a, b)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        e0.record()
        for _ in range(reps):
            _native.copy_probe(a, b)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        print(json.dumps({"copy_probe_bytes": 8 * n, "ms": ms, "GBps": 8 * n / ms / 1e6}))
        return

    if args.workload == "adjoint_c4":
        print(json.dumps(adjoint_bench(args, dev)), flush=True)
        return
    if args.workload == "closure_c5":
        print(json.dumps(closure_bench(args, dev)), flush=True)
        return
    if args.workload == "closure_net":
        print(json.dumps(closure_net_bench(args, dev)), flush=True)
        return
    if args.workload == "tm2d":
        print(json.dumps(tm2d_bench(args, dev)), flush=True)
        return
    if args.workload == "snapshots_c3":
        print(json.dumps(snapshots_bench(args, dev)), flush=True)
        return

    t_setup = time.time()
    grid, default_steps = build_workload(args.workload, dev)
    K = args.steps if args.steps is not None else default_steps
    Wm = args.warmup
    solver = grid.build_solver(W.dt_for(1.0), K + Wm)
    if args.mode == "two_pass":
        solver.fused = False
    torch.cuda.synchronize()
    setup_s = time.time() - t_setup
    nx, ny, nz = grid.size
    cells = nx * ny * nz
    lossy = solver.ca is not None
    npml = solver.pml_layers if solver.cpml is not None else 0

    # warm-up (untimed)
    solver.run(0, Wm)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    psi_bytes = W.psi_bytes_per_step(grid.size, npml) if npml else 0
    if args.mode == "fused":
        # one launch per step: HIP events around every launch on this stream
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(K + 1)]
        torch.cuda.synchronize()
        start.record(stream)
        ev[0].record(stream)
        for q in range(K):
            solver.step_fused(Wm + q, stream)
            ev[q + 1].record(stream)
        stop.record(stream)
        torch.cuda.synchronize()
        total_ms = start.elapsed_time(stop)
        launch_ms = sum(ev[q].elapsed_time(ev[q + 1]) for q in range(K)) / K
        # one step = the fused launches (CPML: slab-free interior, then the
        # slab shell) plus the point-source kernel; algorithmic bytes per
        # cell-update are those of the two-pass design (76 B, 80 B with a
        # per-cell Ca) plus 8 B per psi cell, whatever the kernel moves
        dom_name = "yee_fused_v2_kernel<..., 3> (interior + shell launches of one step)"
        dom_bytes = (80 if lossy else 76) * cells + psi_bytes
        dom_ms = launch_ms
        step_bytes = dom_bytes
        h_ms = e_ms = None
        # what the fused design itself must move: E and H read and written
        # once (48 B), the material id (1 B) or Cb (and Ca) per cell, psi
        coef = 1 if solver.mat is not None else (8 if lossy else 4)
        own_min_bytes = (48 + coef) * cells + psi_bytes
    else:
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(K)]
        torch.cuda.synchronize()
        start.record(stream)
        for q in range(K):
            ev[q][0].record(stream)
            solver.step_h(stream)
            ev[q][1].record(stream)
            solver.step_e(Wm + q, stream)
            ev[q][2].record(stream)
        stop.record(stream)
        torch.cuda.synchronize()
        total_ms = start.elapsed_time(stop)
        h_ms = sum(ev[q][0].elapsed_time(ev[q][1]) for q in range(K)) / K
        e_ms = sum(ev[q][1].elapsed_time(ev[q][2]) for q in range(K)) / K
        h_bytes = 36 * cells + psi_bytes // 2
        e_bytes = (44 if lossy else 40) * cells + psi_bytes // 2
        dom_name, dom_bytes, dom_ms = ("yee_e_kernel", e_bytes, e_ms) if e_ms >= h_ms else \
            ("yee_h_kernel", h_bytes, h_ms)
        step_bytes = h_bytes + e_bytes
        own_min_bytes = None
    traffic, traffic_rec = measured_traffic(args.workload, args.mode)
    ms_step = total_ms / K
    gcells = cells * K / (total_ms / 1e3) / 1e9
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    finite = bool(all(torch.isfinite(f).all().item() for f in solver.fields))

    cpu = None
    if args.cpu_budget > 0:
        try:
            cpu = cpu_baseline(solver, args.cpu_budget)
        except Exception as exc:  # reported, never silently replaced
            cpu = {"value": None, "error": f"{type(exc).__name__}: {exc}"}

    cpu_np = None
    if args.cpu_budget > 0 and args.workload == "c1_64_ref":
        try:
            cpu_np = cpu_baseline_numpy(args.cpu_budget)
        except Exception as exc:  # reported, never silently replaced
            cpu_np = {"value": None, "error": f"{type(exc).__name__}: {exc}"}

    out = {
        "metric": METRIC,
        "value": round(gcells, 3),
        "unit": "Gcell-upd/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "grid": [nx, ny, nz],
            "boundary": solver.boundary_name,
            "npml": npml,
            "lossy": lossy,
            "sources": solver.n_sources,
            "courant": W.COURANT,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": None if traffic_rec is None else traffic_rec.get("source"),
            "bytes_per_launch": dom_bytes,
            "ms_per_launch": round(dom_ms, 4),
            "own_min_bytes": own_min_bytes,
            "frac_own_min": None if own_min_bytes is None else
            round(own_min_bytes / (dom_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "step": {
            "mode": args.mode,
            "h_ms": None if h_ms is None else round(h_ms, 4),
            "e_ms": None if e_ms is None else round(e_ms, 4),
            "bytes": step_bytes,
            "GBps": round(step_bytes / (ms_step / 1e3) / 1e9, 1),
            "frac": round(step_bytes / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "cpu_baseline": cpu,
        "fields_finite": finite,
        "setup_s": round(setup_s, 2),
    }
    if cpu_np is not None:
        out["cpu_baseline_numpy"] = cpu_np
    if not args.no_extras and args.workload == "c3_512":
        # the rest of the metric: gradient cost and the closure, each on its
        # own configuration (after the timed region above)
        del solver, grid
        torch.cuda.empty_cache()
        try:
            out["adjoint"] = adjoint_measure(dev)[0]
        except Exception as exc:  # reported, never silently replaced
            out["adjoint"] = {"error": f"{type(exc).__name__}: {exc}"}
        torch.cuda.empty_cache()
        try:
            from jaxphys_amd.closure import ClosureNet
            net = ClosureNet.load(CLOSURE_FILE, dev)
            summ = closure_summary(net, dev, list(range(1000, 1000 + args.closure_seeds)))
            summ.pop("_res")
            summ["model"] = os.path.relpath(CLOSURE_FILE, ROOT)
            out["closure"] = summ
        except Exception as exc:  # reported, never silently replaced
            out["closure"] = {"error": f"{type(exc).__name__}: {exc}"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
