#!/usr/bin/env python3
"""Benchmark of the fused D2Q9 step on one MI355X.

Workload (the configuration the metric is quoted on): lid-driven cavity, D2Q9 MRT,
1024x1024, moving-wall lid u = 0.1, Re = 1000, FP32 populations. A "step" is one pass of
the fused collide-stream kernel over the grid. Also reported on the same line: the same
cavity with FP16 shifted storage (and its velocity difference from FP32), the 8192^2
periodic BGK grid that lives in HBM (the 256 MB Infinity Cache holds the 1024^2 cavity),
the IBM cylinder, forward + adjoint of the 512^2 inverse problem, the Taylor-Green error
per storage format, and the CPU oracle timed on the host.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cuda-lbm_amd"))

METRIC = "D2Q9 MLUPS & % HBM peak: FP32 vs FP16-storage at stated L2 err; adjoint MLUPS"
HBM_PEAK_GBS = 8000.0
Q = 9


def bytes_per_node(storage):
    from lbm_amd import _abi as A

    return 2 * Q * (4 if storage == A.LBM_F32 else 2)


BATCHES = 5


def time_steps(solver, steps, warmup, torch, batches=BATCHES):
    """Warm up, then time `batches` batches of exactly `steps` steps, each with HIP events on the
    solver's stream and a device synchronisation on both sides. Returns (median seconds,
    per-batch seconds)."""
    stream = torch.cuda.Stream()
    solver.set_stream(stream.cuda_stream)
    solver.step(warmup)
    torch.cuda.synchronize()
    secs = []
    for _ in range(batches):
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            start.record(stream)
            solver.step(steps)
            end.record(stream)
        torch.cuda.synchronize()
        secs.append(start.elapsed_time(end) / 1000.0)
    return sorted(secs)[len(secs) // 2], secs


def run_case(case, steps, warmup, torch, batches=BATCHES):
    """MLUPS (median of `batches` batches of `steps` steps), and the bandwidth of the step
    kernel: a launch advances `steps_per_launch` steps and moves every population once in and
    once out (2*Q*s bytes per node), so the algorithmic bytes per node update are
    2*Q*s / steps_per_launch."""
    from lbm_amd import Solver

    s = Solver(case.config)
    case.apply(s)
    secs, all_secs = time_steps(s, steps, warmup, torch, batches)
    spl = s.steps_per_launch
    N = case.config.N
    mlups = N * steps / secs / 1e6
    B = bytes_per_node(case.config.storage)
    launch_s = secs / steps * spl
    gbs = B * N / launch_s / 1e9
    return s, {"mlups": round(mlups, 1), "ms_per_step": round(1000.0 * secs / steps, 5),
               "steps_per_launch": spl, "us_per_launch": round(1e6 * launch_s, 3),
               "GB_s": round(gbs, 1), "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
               "bytes_per_node_per_launch": B, "bytes_per_node_update": B / spl,
               "batches": len(all_secs), "batch_mlups": [round(N * steps / t / 1e6, 1) for t in all_secs]}


def run_adjoint(torch, steps=2000, every=100):
    """Forward + adjoint of the inverse-problem grid (BGK 512^2 periodic, nu = 0.02, FP32,
    checkpoint every 100 steps). Times plain forward steps, the checkpointing forward and
    the backward (segment recompute + adjoint steps), all with HIP events on one stream."""
    from lbm_amd import Solver
    from lbm_amd import scenarios as sc

    case = sc.inverse_problem(512, steps=steps)
    N = case.config.N
    stream = torch.cuda.Stream()
    s = Solver(case.config)
    s.set_stream(stream.cuda_stream)
    case.apply(s)
    dev = torch.device("cuda")
    F = torch.as_tensor(case.force_field, device=dev).contiguous()
    w = torch.randn(N * 2, device=dev, generator=torch.Generator(device=dev).manual_seed(7))
    gF = torch.empty(N * 2, device=dev)
    gU = torch.empty(N * 2, device=dev)

    def timed(fn):
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / 1000.0

    s.step(50)  # warm-up
    t_plain = timed(lambda: s.step(steps))
    spl_plain = s.steps_per_launch
    # the same forward with one step per launch: the kernel the backward's recompute runs
    one = sc.inverse_problem(512, steps=steps)
    one.config.steps_per_launch = 1
    s1 = Solver(one.config)
    s1.set_stream(stream.cuda_stream)
    one.apply(s1)
    s1.step(50)
    t_one = timed(lambda: s1.step(steps))
    s1.close()

    def fwd():
        with torch.cuda.stream(stream):
            s.set_force_field(F)
            s.init_from_macros(torch.as_tensor(case.rho, device=dev), torch.as_tensor(case.u, device=dev))
            s.adjoint_forward(steps, every)

    fwd()
    s.adjoint_backward(w, gF, gU)  # warm-up (allocates the segment buffers)
    t_fwd = timed(fwd)
    t_bwd = timed(lambda: s.adjoint_backward(w, gF, gU))
    s.close()
    ok = bool(torch.isfinite(gF).all())
    return {"workload": "inverse problem BGK 512x512 periodic nu=0.02 FP32, 2000 steps, checkpoint every 100",
            "forward_mlups": round(N * steps / t_plain / 1e6, 1),
            "forward_steps_per_launch": spl_plain,
            "forward_one_step_launches_mlups": round(N * steps / t_one / 1e6, 1),
            "checkpointing_forward_mlups": round(N * steps / t_fwd / 1e6, 1),
            "adjoint_mlups": round(N * steps / t_bwd / 1e6, 1),
            "adjoint_vs_forward_step_cost": round(t_bwd / t_plain, 3),
            "adjoint_vs_one_step_forward_cost": round(t_bwd / t_one, 3),
            "backward_includes": "segment recompute (1 forward step per adjoint step, two per launch with the "
                                 "intermediate state stored) + adjoint step; cost ratio against the default forward "
                                 "(two steps per launch)",
            "gradient_finite": ok}


def low_pass(F, kmax):
    """The Fourier modes of a periodic field [ny, nx, 2] with |kx|, |ky| <= kmax cycles."""
    import torch

    G = torch.fft.fft2(F.double(), dim=(0, 1))
    ny, nx = F.shape[:2]
    ky = torch.fft.fftfreq(ny, 1.0 / ny, device=F.device).abs()[:, None]
    kx = torch.fft.fftfreq(nx, 1.0 / nx, device=F.device).abs()[None, :]
    mask = ((kx <= kmax) & (ky <= kmax)).to(G.dtype)[..., None]
    return torch.fft.ifft2(G * mask, dim=(0, 1)).real


def run_inverse_problem(torch, steps=2000, every=100, directions=16, iters=200, lr=1e-6, fd_scale=1e-5):
    """The differentiable inverse problem as configured: BGK 512^2 periodic, nu = 0.02, FP32,
    2000 steps, checkpoint every 100; the unknown is the 16-mode Fourier body force (amplitude
    1e-5, seed 0xF0F0); the loss is the L2 misfit of the final velocity. Reports the adjoint
    gradient at F = 0 against central finite differences along 16 seeded random directions
    (seed 0xD1D1), and Adam's recovery of the field from F = 0 over 200 iterations."""
    import numpy as np

    from lbm_amd import scenarios as sc
    from lbm_amd.autograd import Simulation, recover_force

    case = sc.inverse_problem(512, steps=steps)
    dev = torch.device("cuda")
    sim = Simulation(case.config, nsteps=steps, checkpoint_every=every,
                     rho0=torch.as_tensor(case.rho, device=dev))
    u0 = torch.as_tensor(case.u, device=dev)
    F_true = torch.as_tensor(case.force_field, device=dev)
    with torch.no_grad():
        u_target = sim(F_true, u0).clone()

    def loss_of(F):
        with torch.no_grad():
            return float(torch.sum((sim(F, u0).double() - u_target.double()) ** 2))

    F0 = torch.zeros_like(F_true, requires_grad=True)
    loss = torch.sum((sim(F0, u0) - u_target) ** 2)
    loss.backward()
    g = F0.grad.detach().double()
    rng = np.random.default_rng(0xD1D1)
    errs = []
    for _ in range(directions):
        d = torch.as_tensor(rng.standard_normal(tuple(F_true.shape)), device=dev)
        d *= fd_scale / float(d.abs().max())  # well above the FP32 forward's round-off in the loss
        fd = (loss_of((F0.detach() + d).float()) - loss_of((F0.detach() - d).float())) / 2.0
        an = float(torch.sum(g * d))
        errs.append(abs(fd - an) / max(abs(an), 1e-300))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    F_rec, losses = recover_force(sim, u_target, u0, iters=iters, lr=lr)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    sim.close()
    ferr = float(torch.linalg.vector_norm(F_rec - F_true) / torch.linalg.vector_norm(F_true))
    # the unknown's own modes (|k| <= 4 cycles per period): the part the final snapshot determines
    lp = low_pass(F_rec, 4)
    ferr_lp = float(torch.linalg.vector_norm(lp - F_true.double()) / torch.linalg.vector_norm(F_true.double()))
    return {"workload": "inverse problem BGK 512x512 periodic nu=0.02 FP32, 2000 steps, checkpoint every 100, "
                        "16-mode Fourier force amplitude 1e-5",
            "fd_directions": directions, "fd_max_rel_err": float(max(errs)), "fd_median_rel_err": float(np.median(errs)),
            "adam_iters": iters, "adam_lr": lr, "loss_first": losses[0], "loss_last": losses[-1],
            "loss_ratio": losses[-1] / losses[0], "force_rel_l2_err": ferr, "force_rel_l2_err_k_le_4": ferr_lp,
            "fd_scale": fd_scale,
            "seconds_per_iteration": round(dt / iters, 4)}


def run_accuracy(torch):
    """Taylor-Green 128^2 (BGK, nu = 0.01, u0 = 0.04, 1000 steps) L2 error against the analytic
    decay for each storage format, the reference's TaylorGreenScenario::compute_error."""
    from lbm_amd import Solver
    from lbm_amd import _abi as A
    from lbm_amd import scenarios as sc

    out = {}
    for name, st in (("f32", A.LBM_F32), ("f16", A.LBM_F16_SHIFTED), ("bf16", A.LBM_BF16_SHIFTED)):
        case = sc.taylor_green(128, nu=0.01, u0=0.04, storage=st)
        s = Solver(case.config)
        case.apply(s)
        s.step(1000)
        _, u = s.get_macros()
        s.close()
        out[name] = round(sc.tg_error(u, 0.01, 0.04, 1000), 5)
    out["unit"] = "% relative L2 vs analytic decay, Taylor-Green 128^2 BGK 1000 steps"
    return out


def run_poiseuille(torch, steps=400, warmup=20, adapter_steps=100):
    """The force-driven Poiseuille channel (4096x1024 CM, periodic x, bounce-back walls,
    nu = 0.05, Guo force for u_max = 0.05), from rest.

    - CM<NoAdapter>: MLUPS per storage (two steps per launch) and the 16-bit storages'
      velocity against FP32 after the same steps.
    - CM<OptimalAdapter>, the configuration as specified: MLUPS per storage over its first
      warmup + adapter_steps steps. Each step is a moments pass, an FP64 finalize of the
      domain means and the fused step (one step per launch). The reference's adapter
      diverges on this channel (non-finite after ~160-280 steps, in the FP64 oracle too), so
      it is timed only over that finite window and its velocities are not compared."""
    import numpy as np

    from lbm_amd import _abi as A
    from lbm_amd import scenarios as sc

    stores = (("f32", A.LBM_F32), ("bf16", A.LBM_BF16_SHIFTED), ("f16", A.LBM_F16_SHIFTED))
    plain, u = {}, {}
    for name, st in stores:
        case = sc.poiseuille(4096, 1024, storage=st, adapter=0)
        s, r = run_case(case, steps, warmup, torch)
        _, u[name] = s.get_macros()
        s.close()
        plain[name] = r
    umax = float(np.abs(u["f32"][..., 0]).max())
    for name in ("bf16", "f16"):
        plain[name]["u_rel_l2_vs_fp32"] = rel_l2(u[name], u["f32"])
        plain[name]["speedup_vs_fp32"] = round(plain[name]["mlups"] / plain["f32"]["mlups"], 3)
    plain["workload"] = (f"CM<NoAdapter>, {warmup + steps} steps from rest (max u_x {umax:.2e}: the flow is "
                         f"still accelerating; the steady parabola needs ~ny^2/nu = 2e7 steps)")
    adapt = {}
    for name, st in stores:
        case = sc.poiseuille(4096, 1024, storage=st)
        s, r = run_case(case, adapter_steps, warmup, torch)
        _, ua = s.get_macros()
        s.close()
        adapt[name] = dict(r, finite=bool(np.isfinite(ua).all()))
    adapt["workload"] = (f"CM<OptimalAdapter> (same-step means), first {warmup + adapter_steps} steps; per step: "
                         f"one fused step that also sums the next step's means (look-ahead) + an FP64 finalize; "
                         f"the F32/BF16 runs diverge here as the reference's adapter does on this channel")
    return {"cm_no_adapter": plain, "cm_optimal_adapter": adapt,
            "workload": "Poiseuille channel 4096x1024, nu=0.05, Guo force for u_max=0.05, periodic x, bounce-back walls"}


def run_ibm_cylinder(torch, steps=200, warmup=20):
    """The IBM cylinder channel (2048x512 CM, 201 markers, 3 multi-direct forcing iterations
    per step): one launch per step, the IBM pass in its first workgroups beside the step
    outside the IBM box."""
    from lbm_amd import scenarios as sc

    s, r = run_case(sc.ibm_cylinder(), steps, warmup, torch)
    s.close()
    return dict(r, workload="IBM cylinder 2048x512 CM, 201 markers, 3 IBM iterations per step "
                            "(one launch per step: the IBM pass beside the step outside its box)")


def run_d3q27(torch, steps=200, warmup=20):
    """The D3Q27 BGK step (one fused pull-scheme launch per step, 27 planes in and out):
    Taylor-Green 3-D at the reference scenario's 128^3, a 256^3 periodic grid whose 2 x 1.8 GB
    FP32 buffers stream from HBM (FP32 and FP16 storage), and a 256x128x96 channel between
    bounce-back walls (the reference wind tunnel's grid)."""
    import numpy as np

    from lbm_amd import _abi as A
    from lbm_amd.solver3d import Config3D, Solver3D

    def tg_init(n):
        k = 2 * np.pi / n
        z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float32),) * 3, indexing="ij")
        u = np.stack([0.05 * np.sin(k * x) * np.cos(k * y) * np.cos(k * z),
                      -0.05 * np.cos(k * x) * np.sin(k * y) * np.cos(k * z), np.zeros_like(x)], -1)
        return np.ones((n, n, n), np.float32), u.astype(np.float32)

    def one(cfg, init, flags=None, n_steps=steps, markers=None):
        with Solver3D(cfg) as s:
            s.init_from_macros(*init)
            if flags is not None:
                s.set_flags(flags)
            if markers is not None:
                s.ibm_set_markers(*markers)
            secs, _ = time_steps(s, n_steps, warmup, torch, batches=3)
            B = 2 * 27 * (4 if cfg.storage == A.LBM_F32 else 2)
            gbs = B * cfg.N / (secs / n_steps) / 1e9
            return {"mlups": round(cfg.N * n_steps / secs / 1e6, 1), "us_per_step": round(1e6 * secs / n_steps, 2),
                    "GB_s": round(gbs, 1), "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_node_update": B}

    per = dict(periodic_x=True, periodic_y=True, periodic_z=True)
    res = {}
    res["tg3d_128_f32"] = one(Config3D(128, 128, 128, omega=1.8, **per), tg_init(128))
    for name, st in (("hbm_256_f32", A.LBM_F32), ("hbm_256_f16", A.LBM_F16_SHIFTED)):
        res[name] = one(Config3D(256, 256, 256, omega=1.8, storage=st, **per), tg_init(256), n_steps=max(20, steps // 2))
    res["speedup_f16_vs_f32"] = round(res["hbm_256_f16"]["mlups"] / res["hbm_256_f32"]["mlups"], 3)
    S = [1.0] * 27
    S[4:9] = [1.8] * 5
    res["hbm_256_mrt_f32"] = one(Config3D(256, 256, 256, collision=A.LBM_MRT, S=tuple(S), **per), tg_init(256),
                                 n_steps=max(20, steps // 2))
    res["hbm_256_cm_f32"] = one(Config3D(256, 256, 256, collision=A.LBM_CM, omega=1.8, **per), tg_init(256),
                                n_steps=max(20, steps // 2))
    nx, ny, nz = 256, 128, 96
    fl = np.zeros((nz, ny, nx), np.uint8)
    fl[:, 0, :] = fl[:, -1, :] = A.BOUNCE_BACK
    fl[0], fl[-1] = A.BOUNCE_BACK, A.BOUNCE_BACK
    ch = Config3D(nx, ny, nz, omega=1.6, periodic_x=True, force=(1e-6, 0.0, 0.0))
    res["channel_256x128x96_f32"] = one(ch, (np.ones((nz, ny, nx), np.float32), np.zeros((nz, ny, nx, 3), np.float32)), fl)
    # the reference's default 3-D case (windTestScenario.cuh) without its IBM body: CM<3>, Guo
    # velocity inlet / pressure outlet on x, bounce-back on y and z, plug flow u = 0.05
    wt = fl.copy()
    wt[:, :, 0] = A.GUO_VELOCITY_INLET
    wt[:, :, -1] = A.GUO_PRESSURE_OUTLET
    u0 = np.zeros((nz, ny, nx, 3), np.float32)
    u0[..., 0] = 0.05
    wcfg = Config3D(nx, ny, nz, collision=A.LBM_CM, omega=1.0 / (3.0 * 0.05 * ny / 1000.0 + 0.5), u_inlet=0.05)
    res["wind_tunnel_256x128x96_cm_f32"] = one(wcfg, (np.ones((nz, ny, nx), np.float32), u0), wt)
    # ... and with its body (windTest/sphere_config.ini): a 64x64 UV-sphere mesh scaled to D = 16
    # with its anchor at (128, 64, 48), 256 markers by weighted sample elimination in Morton
    # order (the host pipeline of include/lbm_body.h), area pi D^2 / 256 each, 3 IBM iterations
    import tempfile

    from lbm_amd import bodies as Bd

    with tempfile.TemporaryDirectory() as td:
        v, f = Bd.uv_sphere(64, 64)
        Bd.Body.from_mesh(v, f).collect_file(os.path.join(td, "sphere_unit_64x64.obj"))
        ini = os.path.join(td, "sphere_config.ini")
        with open(ini, "w") as fh:
            fh.write("[General]\nobj_filepath = sphere_unit_64x64.obj\n[Transformations]\nscale_overall_size = 16.0\n"
                     "position_anchor_x = 128.0\nposition_anchor_y = 64.0\nposition_anchor_z = 48.0\n"
                     "[Sampling]\ntype = target_points\ntarget_points = 256\n")
        t0 = time.perf_counter()
        with Bd.Body.from_config(ini, (nx, ny, nz), base_dir=td) as body:
            xyz = body.markers
        prep = time.perf_counter() - t0
    ds = np.full(len(xyz), np.pi * 16.0 ** 2 / len(xyz), np.float32)
    r = one(wcfg, (np.ones((nz, ny, nx), np.float32), u0), wt, markers=(xyz, ds, 3))
    r.update(markers=int(len(xyz)), body_preparation_s=round(prep, 3),
             ibm_cost_vs_no_body=round(res["wind_tunnel_256x128x96_cm_f32"]["mlups"] / r["mlups"], 3))
    res["wind_tunnel_sphere_256x128x96_cm_f32"] = r
    # the flags of flowPastSquareCylinderScenario.cuh:10-31 without its body: corner-edge
    # bounce-back on edges and corners, regularized inlet at x = 0, Guo pressure outlet at
    # x = NX-1, bounce-back on y and z (the regularized kernel variant)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij", sparse=True)
    faces = (x == 0) * 1 + (x == nx - 1) + (y == 0) + (y == ny - 1) + (z == 0) + (z == nz - 1)
    sq = np.zeros((nz, ny, nx), np.uint8)
    sq[np.broadcast_to((y == 0) | (y == ny - 1) | (z == 0) | (z == nz - 1), sq.shape)] = A.BOUNCE_BACK
    sq[:, :, -1] = A.GUO_PRESSURE_OUTLET
    sq[:, :, 0] = A.REGULARIZED_INLET_LEFT
    sq[faces >= 2] = A.CORNER_EDGE_BOUNCE_BACK
    res["square_duct_256x128x96_cm_f32"] = one(wcfg, (np.ones((nz, ny, nx), np.float32), u0), sq)
    res["workload"] = ("D3Q27 BGK with Guo forcing (and MRT<3>, CM<3> at 256^3), FP32 collision; Taylor-Green 3-D "
                       "(u_max 0.05) periodic, a force-driven channel with bounce-back walls on y and z, and the wind "
                       "tunnel (CM<3>, Guo inlet/outlet, walls) with and without its IBM sphere, and the "
                       "square-cylinder duct's flags (corner-edge bounce-back, regularized inlet, Guo outlet, walls) "
                       "without the body; 216 B "
                       "(FP32) / 108 B (FP16) per node update")
    return res


def rel_l2(a, b):
    import numpy as np

    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def profiled_traffic(kernel_tag):
    """HBM bytes per launch from the committed rocprofv3 PMC summary for this kernel and
    workload (FETCH_SIZE and WRITE_SIZE, each scaled by the copy-kernel calibration)."""
    p = ROOT / "profiles" / f"pmc_{kernel_tag}.txt"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def copy_calibration(torch, mb):
    """One plain device copy of `mb` MB, whose profiler byte counters calibrate the step's."""
    n = mb * 1000000 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    return {"read": 4 * n, "write": 4 * n}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def cpu_baseline(sample_steps=800):
    """The FP32 oracle (OpenMP over rows) on the host cores: a bounded sample of the
    same cavity (1024^2 MRT, `sample_steps` steps)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    from lbm_amd import scenarios as sc

    case = sc.lid_driven_cavity(1024)
    o = orc.Oracle(case.config.to_c(), precision=32)
    case.apply_oracle(o)
    o.step(1)
    t0 = time.perf_counter()
    o.step(sample_steps)
    dt = time.perf_counter() - t0
    return {"value": round(case.config.N * sample_steps / dt / 1e6, 2), "unit": "MLUPS",
            "cores": orc.num_threads(), "cpu_model": cpu_model(), "kind": "port",
            "sample": f"lid-driven cavity 1024x1024 MRT FP32, {sample_steps} steps after 1 warm-up step "
                      f"({dt:.1f} s)"}


# Step counts of the FP64 host runs the configs' precision is stated against (bounded so the
# host part of the bench stays near 20 s on 16 cores).
REF_STEPS = {"ldc": 200, "poiseuille": 20000, "ibm": 200, "tg": 1000}


def cpu_fp64_references():
    """FP64 oracle runs of the configs, the "FP64 run" each config's precision is stated
    against: the cavity, the channel and the IBM cylinder over REF_STEPS steps, Taylor-Green
    over its full 1000 steps. Host only; the GPU runs the same cases in run_config_precision."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np

    import oracle as orc
    from lbm_amd import scenarios as sc

    t0 = time.perf_counter()
    out = {}
    c = sc.lid_driven_cavity(1024)
    o = orc.Oracle(c.config.to_c(), precision=64)
    c.apply_oracle(o)
    o.step(REF_STEPS["ldc"])
    out["ldc_u"] = o.get_macros()[1]
    del o
    # the channel is periodic in x under a uniform force, so its flow does not depend on x: a
    # 4-wide FP64 channel carries the 4096-wide one's profile over the full 20k steps
    c = sc.poiseuille(4, 1024, adapter=0)
    o = orc.Oracle(c.config.to_c(), precision=64)
    c.apply_oracle(o)
    o.step(REF_STEPS["poiseuille"])
    out["poiseuille_profile"] = o.get_macros()[1][..., 0].mean(axis=1)
    del o
    c = sc.ibm_cylinder()
    o = orc.Oracle(c.config.to_c(), precision=64)
    c.apply_oracle(o)
    series = []
    for _ in range(REF_STEPS["ibm"]):
        o.step(1)
        series.append(o.ibm_get_forces()[1])
    out["ibm_series"] = np.array(series)
    del o
    c = sc.taylor_green(128, nu=0.01, u0=0.04)
    o = orc.Oracle(c.config.to_c(), precision=64)
    c.apply_oracle(o)
    o.step(REF_STEPS["tg"])
    out["tg_error"] = sc.tg_error(o.get_macros()[1], 0.01, 0.04, REF_STEPS["tg"])
    out["seconds"] = round(time.perf_counter() - t0, 1)
    return out


def run_config_precision(torch, refs):
    """Each config's stated precision, on the GPU:
    - cavity 1024^2 MRT, Re = 1000: FP32 and FP16 storage against the FP64 host run after
      REF_STEPS steps, and the Ghia centre-line NRMSE at the configured 50k steps;
    - Poiseuille 4096x1024 (CM without the adapter, which diverges on this channel): FP32, BF16
      and FP16 after the configured 20k steps, as the reference's RMS % of u_max against its
      steady parabola (the config's analytic), and the profile's relative L2 error against the
      start-up solution at that step and against the FP64 host run of the same 20k steps;
    - IBM cylinder: the drag and lift series over REF_STEPS steps against the FP64 host run;
    - Taylor-Green: the FP64 host run's own analytic error beside the GPU's (run_accuracy).
    Fields compared are velocities; "rel_l2" = ||u - u_ref|| / ||u_ref||."""
    import numpy as np

    from lbm_amd import Solver
    from lbm_amd import _abi as A
    from lbm_amd import scenarios as sc

    def run(case, steps):
        s = Solver(case.config)
        case.apply(s)
        s.step(steps)
        _, u = s.get_macros()
        s.close()
        return u

    out = {}
    ldc = {"re": 1000, "ref_steps": REF_STEPS["ldc"], "steps": 50000}
    for name, st in (("f32", A.LBM_F32), ("f16", A.LBM_F16_SHIFTED)):
        r = {}
        if refs:
            r["u_rel_l2_vs_fp64"] = rel_l2(run(sc.lid_driven_cavity(1024, storage=st), REF_STEPS["ldc"]), refs["ldc_u"])
        u = run(sc.lid_driven_cavity(1024, storage=st), 50000)
        r["ghia_nrmse_pct"] = round(sc.ghia_error(u, 0.1, 1000), 3)
        ldc[name] = r
    # the 1024^2 cavity is far from steady at 50k steps (the vortex spins up over ~10 L / u_lid
    # = 100k steps and settles later); the validator's Ghia comparison on a cavity that has
    # converged: 257^2, Re = 1000, 300k steps (~117 L / u_lid)
    for name, st in (("f32", A.LBM_F32), ("f16", A.LBM_F16_SHIFTED)):
        ldc[name]["ghia_nrmse_pct_257_converged"] = round(
            sc.ghia_error(run(sc.lid_driven_cavity(257, storage=st), 300000), 0.1, 1000), 3)
    out["lid_driven_cavity_1024_mrt"] = ldc
    nu, umax, T = 0.05, 0.05, 20000
    startup = sc.poiseuille_startup_profile(1024, nu, umax, T)
    poi = {"adapter": "none (CM<OptimalAdapter> diverges on this channel, in the FP64 oracle too)", "steps": T,
           "ref_steps": REF_STEPS["poiseuille"]}
    for name, st in (("f32", A.LBM_F32), ("bf16", A.LBM_BF16_SHIFTED), ("f16", A.LBM_F16_SHIFTED)):
        u = run(sc.poiseuille(4096, 1024, storage=st, adapter=0), T)
        prof = u[..., 0].astype(np.float64).mean(axis=1)
        r = {"rms_pct_umax_vs_parabola": round(sc.poiseuille_error(prof, nu, umax), 4),
             "profile_rel_l2_vs_startup": rel_l2(prof, startup),
             "centre_u": float(prof[512])}
        if refs:
            r["profile_rel_l2_vs_fp64"] = rel_l2(prof, refs["poiseuille_profile"])
        poi[name] = r
    poi["startup_centre_u"] = float(startup[512])
    out["poiseuille_4096x1024_cm"] = poi
    if refs:
        c = sc.ibm_cylinder()
        s = Solver(c.config)
        c.apply(s)
        series = []
        for _ in range(REF_STEPS["ibm"]):
            s.step(1)
            series.append(s.ibm_get_forces()[1])
        s.close()
        g, h = np.array(series), refs["ibm_series"]
        out["ibm_cylinder_drag_lift"] = {
            "steps": REF_STEPS["ibm"],
            "drag_max_rel_err": float(np.max(np.abs(g[:, 0] - h[:, 0])) / np.max(np.abs(h[:, 0]))),
            "lift_max_err_over_max_drag": float(np.max(np.abs(g[:, 1] - h[:, 1])) / np.max(np.abs(h[:, 0]))),
            "drag_last": float(g[-1, 0]), "lift_last": float(g[-1, 1])}
        out["taylor_green_fp64_error_pct"] = round(refs["tg_error"], 5)
    out["storage_validity"] = {
        "f32": "all configs",
        "f16": "Taylor-Green and the cavity; not the force-driven channel from rest, whose per-step "
               "increments fall below half an FP16 ulp of the stored deviations",
        "bf16": "reported only (8-bit mantissa): Taylor-Green error 8x the FP64 run's"}
    return out


_T0 = time.perf_counter()


def progress(what):
    """One line per finished section on stderr (stdout carries the single JSON line)."""
    print(f"bench: {what} ({time.perf_counter() - _T0:.1f} s)", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-extras", action="store_true", help="only the headline workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["hbm_grid", "adjoint", "poiseuille", "ibm_cylinder", "d3q27", "inverse",
                                       "precision"], default=None,
                    help="run one extra section alone (for profiling its kernels); prints that section")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="also run one plain device copy of MB megabytes (for scaling profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "a design note, not part of this build)"}))
        sys.exit(2)

    import torch

    from lbm_amd import _abi as A
    from lbm_amd import scenarios as sc

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        sys.exit(1)

    if args.only == "hbm_grid":
        res = {}
        for name, st in (("f32", A.LBM_F32), ("f16", A.LBM_F16_SHIFTED)):
            s, res[name] = run_case(sc.bandwidth_grid(8192, storage=st), max(20, args.steps // 20), 10, torch)
            s.close()
        if args.calibrate_copy:
            res["copy_calibration_bytes"] = copy_calibration(torch, args.calibrate_copy)
        print(json.dumps({"hbm_grid": res}))
        return
    if args.only == "adjoint":
        print(json.dumps({"adjoint": run_adjoint(torch)}))
        return
    if args.only == "poiseuille":
        print(json.dumps({"poiseuille": run_poiseuille(torch)}))
        return
    if args.only == "inverse":
        print(json.dumps({"inverse_problem": run_inverse_problem(torch)}))
        return
    if args.only == "precision":
        refs = None if args.no_cpu else cpu_fp64_references()
        print(json.dumps({"config_precision": run_config_precision(torch, refs)}))
        return
    if args.only == "ibm_cylinder":
        print(json.dumps({"ibm_cylinder": run_ibm_cylinder(torch)}))
        return
    if args.only == "d3q27":
        res = run_d3q27(torch, steps=max(20, args.steps // 10))
        if args.calibrate_copy:
            res["copy_calibration_bytes"] = copy_calibration(torch, args.calibrate_copy)
        print(json.dumps({"d3q27": res}))
        return

    # headline: LDC 1024^2 MRT FP32
    case = sc.lid_driven_cavity(1024, storage=A.LBM_F32)
    s, head = run_case(case, args.steps, args.warmup, torch)
    _, u32 = s.get_macros()
    s.close()
    progress("headline cavity")
    out = {
        "metric": METRIC,
        "value": head["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "lid-driven cavity D2Q9 MRT 1024x1024, moving-wall lid u=0.1, Re=1000, FP32 "
                               "populations, fluid at rest at t=0",
                   "nodes": case.config.N, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": head["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": head["frac_hbm_peak"],
                     "traffic": profiled_traffic("fused_step2_ldc1024_mrt_f32" if head["steps_per_launch"] == 2
                                                 else "fused_step_ldc1024_mrt_f32"),
                     "kernel": ("fused_step2_kernel<MRT,F32>" if head["steps_per_launch"] == 2
                                else "fused_step_kernel<MRT,F32>"),
                     "steps_per_launch": head["steps_per_launch"],
                     "kernel_avg_us": head["us_per_launch"],
                     "algorithmic_bytes_per_launch": head["bytes_per_node_per_launch"] * case.config.N,
                     "note": "a launch advances steps_per_launch steps and reads + writes every population "
                             "once (72 B/node); launch time = HIP-event time over the timed steps / launches; "
                             "the 75 MB working set stays in the 256 MB Infinity Cache, see hbm_grid for the "
                             "HBM-bound grid"},
    }
    if not args.no_extras:
        c16 = sc.lid_driven_cavity(1024, storage=A.LBM_F16_SHIFTED)
        s, r16 = run_case(c16, args.steps, args.warmup, torch)
        _, u16 = s.get_macros()
        s.close()
        out["fp16_storage"] = dict(r16, workload="same cavity, FP16 shifted storage",
                                   speedup_vs_fp32=round(r16["mlups"] / head["mlups"], 3),
                                   u_rel_l2_vs_fp32=rel_l2(u16, u32))
        grid = {}
        for name, st in (("f32", A.LBM_F32), ("f16", A.LBM_F16_SHIFTED)):
            g = sc.bandwidth_grid(8192, storage=st)
            s, r = run_case(g, max(20, args.steps // 20), 10, torch)
            s.close()
            grid[name] = r
        grid["speedup_f16_vs_f32"] = round(grid["f16"]["mlups"] / grid["f32"]["mlups"], 3)
        grid["workload"] = "periodic BGK 8192x8192 (2 x 2.4 GB FP32 population buffers)"
        out["hbm_grid"] = grid
        progress("hbm grid")
        out["ibm_cylinder"] = run_ibm_cylinder(torch)
        progress("ibm cylinder")
        out["poiseuille"] = run_poiseuille(torch)
        progress("poiseuille")
        out["adjoint"] = run_adjoint(torch)
        progress("adjoint")
        out["inverse_problem"] = run_inverse_problem(torch)
        progress("inverse problem")
        out["accuracy"] = run_accuracy(torch)
        out["d3q27"] = run_d3q27(torch, steps=max(20, args.steps // 10))
        progress("d3q27")
    if args.calibrate_copy:
        out["copy_calibration_bytes"] = copy_calibration(torch, args.calibrate_copy)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        progress("cpu baseline")
        if not args.no_extras:
            refs = cpu_fp64_references()
            progress("fp64 host runs")
            out["config_precision"] = run_config_precision(torch, refs)
            out["config_precision"]["fp64_host_runs_s"] = refs["seconds"]
    elif not args.no_extras:
        out["config_precision"] = run_config_precision(torch, None)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
