#!/usr/bin/env python3
"""Benchmark of the FP64 RK3 flow step on one MI355X.

A step is one TVD-RK3 step (flow_time_advance, code/flow-solver/OCFD_Time_Adv.f90:7-92)
over the whole synthetic Mach 2.25 flat-plate boundary layer, 640 x 120 x 64
cells, hybrid UD7L/WENO7/WENO5 inviscid scheme switched by the pressure
gradient sensor, Steger-Warming splitting, CD6 viscous terms, boundary-layer
BCs.  The state is resident in HBM before the timed region; the step is
replayed from a captured hipGraph.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "Mcell-steps/s per RK3 step (FP64); chem cells/s surrogate vs plain at ≤1% ΔT"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBPS = 8000.0
PEAK_BF16_TFLOPS = 2500.0


def load_pkg():
    if "ocfd_amd" in sys.modules:
        return sys.modules["ocfd_amd"]
    pkg_dir = os.path.join(ROOT, "opencfd-comb_amd")
    spec = importlib.util.spec_from_file_location("ocfd_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ocfd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def setup(engine, case):
    cfg, mesh, span, st, bc, info = case
    engine.set_mesh(*mesh, span)
    engine.set_bc_boundarylayer(bc["inlet1d"], bc["upper0d"], bc["nonreflect_upper"], bc["wall_pertb"])
    engine.set_bc_time(info["ht"](0.0))
    engine.upload_state(st)
    engine.comput_flow_variables()
    engine.OCFD_bc()


def run_steps(engine, info, nsteps, t0_step):
    """nsteps RK3 steps; the reference advances tt by dt after every second
    RK3 step (OCFD_NS_Solver.f90:28-46), which sets the wall forcing ht(tt)."""
    dt = info["dt"]
    for s in range(t0_step, t0_step + nsteps):
        if s % 2 == 0:
            engine.set_bc_time(info["ht"]((s // 2) * dt))
        engine.flow_time_advance(0.5 * dt)
    return t0_step + nsteps


def cpu_baseline(m, nx, ny, nz, budget_s=15.0, max_steps=40):
    """The oracle (the C restatement, OpenMP over every host thread that
    OMP_NUM_THREADS allows) on the same workload: the whole flat plate at
    nx x ny x nz, as many RK3 steps as fit in about budget_s after a first
    (timed, discarded) step."""
    sys.path.insert(0, ROOT)
    import oracle
    from ocfd_amd import cases
    case = cases.flat_plate(nx, ny, nz)
    o = oracle.Oracle(m, case[0])
    threads = o.set_threads(0)
    setup(o, case)
    t = time.perf_counter()
    step = run_steps(o, case[5], 1, 0)
    first = time.perf_counter() - t
    nsteps = int(max(1, min(max_steps, budget_s / max(first, 1e-3))))
    t = time.perf_counter()
    run_steps(o, case[5], nsteps, step)
    el = time.perf_counter() - t
    o.close()
    ncell = nx * ny * nz
    return dict(value=ncell * nsteps / el / 1e6, unit="Mcell-steps/s", cores=threads, kind="port",
                sample=f"oracle (C restatement, OpenMP, {threads} threads) on the whole flat plate "
                       f"{nx}x{ny}x{nz}, {nsteps} RK3 steps after one untimed, {el:.1f} s")


SURR_TOL = 5e-4           # mass-fraction error the increment bound is calibrated to
SURR_CAL_CELLS = 1 << 20  # held-out labelled cells for that calibration


def chem_bench(m, ncell=1 << 21, cpu=True, surrogate_cells=1 << 24, surrogate_epochs=25):
    """Finite-rate chemistry on the jet-flame batch: ncell independent cells
    (jet fuel/oxidiser mixtures, T 900-2500 K, p 0.5-5 atm, taken along their
    constant-volume ignition trajectories on the GPU), one chemistry step dt
    ~ log-U[1e-8, 1e-6] s per cell.  Times the plain FP64 integrators (torch
    events on the stream the kernels run on) and the CPU oracle on a sample."""
    import numpy as np
    import torch
    from ocfd_amd import cases, chemistry, mechanism
    M = mechanism.Mechanism()
    ch = chemistry.ChemBatch(M)
    s = cases.jet_chem_states(ncell, seed=3, mech=M)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda").contiguous()
    rho, E, tau, dt = dev(s["rho"]), dev(s["E"]), dev(s["tau"]), dev(s["dt"])
    ch.adaptive(rho, E, tau, rtol_y=0.05, y_floor=1e-6, rtol_t=0.01, max_sub=100000)
    if ch.unfinished():
        raise RuntimeError("chemistry batch set-up left cells unfinished")
    r0, e0 = rho.clone(), E.clone()

    def timed(fn, reps=3):
        best, out = None, None
        for _ in range(reps):
            rho.copy_(r0)
            E.copy_(e0)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            out = fn()
            b.record()
            torch.cuda.synchronize()
            ms = a.elapsed_time(b)
            best = ms if best is None else min(best, ms)
        return best, out

    from ocfd_amd import flops
    res = {}
    ms, ns = timed(lambda: ch.adaptive(rho, E, dt, scheduler=1))
    k = ns.cpu().numpy()
    cnt = ch.counters()
    # FP64 roofline of the integrator: about 7e3 flops per point-implicit
    # substep, times the substeps the batch took
    chem_tf = flops.CHEM_SUBSTEP * float(k.sum()) / (ms / 1e3) / 1e12
    res["plain_adaptive"] = dict(ms=ms, cells_per_s=ncell / ms * 1e3, unfinished=cnt["unfinished"],
                                 scheduler="persistent work queue (one cell per lane)",
                                 substeps=dict(mean=float(k.mean()), p50=float(np.median(k)),
                                               p99=float(np.percentile(k, 99)), max=int(k.max())),
                                 roofline=dict(bound="fp64", achieved=chem_tf, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                                               frac=chem_tf / PEAK_FP64_TFLOPS, flops_per_substep=flops.CHEM_SUBSTEP,
                                               substeps_total=int(k.sum())))
    ms, _ = timed(lambda: ch.adaptive(rho, E, dt, scheduler=2))
    res["plain_adaptive_queue_plus_cooperative"] = dict(ms=ms, cells_per_s=ncell / ms * 1e3, lane_budget=128,
                                                        cooperative_cells=ch.counters()["cooperative"])
    ms, _ = timed(lambda: ch.adaptive(rho, E, dt, scheduler=0))
    res["plain_adaptive_thread"] = dict(ms=ms, cells_per_s=ncell / ms * 1e3)
    # the plain baseline of the surrogate comparison is the fastest plain scheduler
    best_plain = min(res[k]["ms"] for k in ("plain_adaptive", "plain_adaptive_queue_plus_cooperative",
                                            "plain_adaptive_thread"))
    ms, _ = timed(lambda: ch.chemical_timeAdv(rho, E, dt, mode=1, nsub=5))
    res["fixed_euler_M5"] = dict(ms=ms, cells_per_s=ncell / ms * 1e3)
    ms, _ = timed(lambda: ch.chemical_timeAdv(rho, E, dt, mode=-1, nsub=5))
    res["fixed_point_implicit_M5"] = dict(ms=ms, cells_per_s=ncell / ms * 1e3)
    # plain reference result of the batch (the measure of the surrogate's error)
    rho.copy_(r0)
    E.copy_(e0)
    ch.adaptive(rho, E, dt, scheduler=1)
    rp, Ep = rho.clone(), E.clone()
    out = dict(workload=f"jet-flame chemistry batch, {ncell} cells, 9 species / 19 reactions, "
                        "dt log-U[1e-8,1e-6] s", dtype="f64",
               controller=dict(rtol_y=chemistry.RTOL_Y, y_floor=chemistry.Y_FLOOR, rtol_t=chemistry.RTOL_T),
               **res)
    S = None
    if surrogate_cells:
        from ocfd_amd import surrogate
        t0 = time.perf_counter()
        S = surrogate.Surrogate(ch, seed=0)
        data = S.dataset(surrogate_cells, seed=1)
        S.configure(data)
        x, y = S.training_arrays(data)
        del data
        t1 = time.perf_counter()
        hist = S.train(x, y, epochs=surrogate_epochs, batch=8192, lr=1e-3)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        del x, y
        # increment bound from held-out labelled cells (a different seed)
        cal = S.calibrate(S.dataset(SURR_CAL_CELLS, seed=2), tol=SURR_TOL)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        rho.copy_(r0)
        E.copy_(e0)
        tile_rows, tile_waves, tile_ms = S.autotune(rho, E, dt, reps=3)
        flags = torch.zeros(ncell, dtype=torch.int32, device="cuda")
        Tp = S.temperature(rp, Ep).cpu().numpy()
        Yp = (rp / rp.sum(0)).cpu().numpy()
        major = [M.index(k) for k in ("H2", "O2", "H2O", "N2")]

        def surrogate_run():
            ms_s, _ = timed(lambda: S.step(rho, E, dt, fallback=True, flags=flags))
            nfb = S.fallbacks()
            ms_nofb, _ = timed(lambda: S.step(rho, E, dt, fallback=False))
            # BF16 MFMA roofline of the inference (features, network, checks,
            # no fallback): 535,040 flops per cell
            mlp_tf = flops.MLP_CELL * ncell / (ms_nofb / 1e3) / 1e12
            rho.copy_(r0)
            E.copy_(e0)
            S.step(rho, E, dt, fallback=True, flags=flags)
            Ts = S.temperature(rho, E).cpu().numpy()
            Ys = (rho / rho.sum(0)).cpu().numpy()
            dT = np.abs(Ts - Tp)
            dYm = np.abs(Ys[major] - Yp[major]).max(axis=0)
            dYa = np.abs(Ys - Yp).max(axis=0)
            f = flags.cpu().numpy()
            return dict(
                ms=ms_s, cells_per_s=ncell / ms_s * 1e3, fallbacks=nfb,
                fallback_reasons={k: int(np.sum((f & b) != 0)) for k, b in
                                  (("domain", 1), ("element_balance", 2), ("negative_Y", 4), ("T_bounds", 8),
                                   ("increment_bound", 16))},
                ms_without_fallback=ms_nofb,
                roofline_inference=dict(bound="mfma", pipe="bf16 dense MFMA", achieved=mlp_tf,
                                        peak=PEAK_BF16_TFLOPS, unit="TFLOP/s", frac=mlp_tf / PEAK_BF16_TFLOPS,
                                        flops_per_cell=flops.MLP_CELL),
                speedup_vs_plain=best_plain / ms_s, plain_ms=best_plain,
                error_vs_plain=dict(dT_max_K=float(dT.max()), dT_p99_K=float(np.percentile(dT, 99)),
                                    dT_rel_max=float(np.max(dT / Tp)),
                                    peak_T_rel=float(abs(Ts.max() - Tp.max()) / Tp.max()),
                                    dY_major_max=float(dYm.max()), dY_major_p99=float(np.percentile(dYm, 99)),
                                    dY_all_species_max=float(dYa.max())))

        # headline: the network with the domain, balance, sign and T checks;
        # then the same network with the calibrated increment bound as well
        cfg_b = S.cfg.copy()
        cfg_u = cfg_b.copy()
        cfg_u[S.DY_MAX_INDEX] = np.inf
        S.set_config(cfg_u)
        res_u = surrogate_run()
        S.set_config(cfg_b)
        res_b = surrogate_run()
        keep = ("ms", "cells_per_s", "fallbacks", "fallback_reasons", "speedup_vs_plain", "error_vs_plain",
                "roofline_inference")
        out["surrogate"] = dict(
            network="12 -> 256 x5 -> 9 GELU, BF16 MFMA (FP32 accumulation), features / network / checks kernels",
            tile=dict(cells_per_workgroup=tile_rows, waves=tile_waves,
                      ms_per_shape={"128x4": tile_ms[0], "128x4 (2x2 waves)": tile_ms[1], "64x4": tile_ms[2],
                                    "96x4": tile_ms[3]}),
            train=dict(cells=surrogate_cells, epochs=len(hist), batch=8192, data_s=t1 - t0, train_s=t2 - t1,
                       final_loss=hist[-1]),
            **res_u,
            with_increment_bound=dict(calibration=dict(cal, seconds=t3 - t2), **{k: res_b[k] for k in keep}))
    if cpu:
        import oracle
        orc = oracle.ChemOracle(m, M)
        threads = orc.set_threads(1)
        nsample = 32768
        r = rho.cpu().numpy()
        rs, es, ds = r0[:, :nsample].cpu().numpy().T.copy(), e0[:nsample].cpu().numpy(), s["dt"][:nsample]
        t0 = time.perf_counter()
        orc.adaptive_batch(rs, es, ds, chemistry.RTOL_Y, chemistry.Y_FLOOR, chemistry.RTOL_T, 4000)
        el = time.perf_counter() - t0
        orc.set_threads(0)
        out["cpu_baseline"] = dict(value=nsample / el, unit="cells/s", cores=threads, kind="port",
                                   sample=f"first {nsample} cells of the same batch, adaptive point-implicit "
                                          f"(oracle, C, {threads} thread), {el:.1f} s")
    return out, S


def surrogate_inference_only(tile, n=1 << 21, ndata=1 << 18):
    """The three surrogate inference launches alone on n cells (the C4
    sampler's states, repeated), untrained network, no fallback: the cost of
    the inference does not depend on the weights."""
    import torch
    load_pkg()
    from ocfd_amd import chemistry, mechanism, surrogate
    M = mechanism.Mechanism()
    ch = chemistry.ChemBatch(M)
    S = surrogate.Surrogate(ch, seed=0)
    data = S.dataset(ndata, seed=1)
    S.configure(data)
    idx = torch.arange(n, device="cuda") % ndata
    rho0, E0, dt = data["rho0"][:, idx].contiguous(), data["E0"][idx].contiguous(), data["dt"][idx].contiguous()
    S.tile(tile)
    ms = []
    for it in range(5):
        rho, E = rho0.clone(), E0.clone()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        S.step(rho, E, dt, fallback=False)
        b.record()
        torch.cuda.synchronize()
        if it >= 2:
            ms.append(a.elapsed_time(b))
    return dict(cells=n, tile=S.tile(), ms=ms)


def jet_bench(m, S, nx=256, ny=192, nz=128, steps=10, warmup=2, tau=2e-5, cfl=0.5, serial=False):
    """The reacting planar H2 jet (WENO5, 9 species, real gas, Chemkin/Wilke
    transport, BC_JET) in NS_solver steps: flow dt/2, chemistry over dt,
    comput_T_P, flow dt/2.  The field is first pre-reacted for tau so the
    igniter has lit the mixing layers.  Runs the same steps with the plain
    adaptive FP64 integrator and with the surrogate plus fallback, from the
    same state, and compares the final fields."""
    import numpy as np
    import torch
    from ocfd_amd import cases, chemistry
    case = cases.planar_jet(nx, ny, nz)
    cfg, mesh, span, st, bc, info = case
    mech = info["mech"]
    chemistry.ChemBatch(mech)
    g = m.FlowSolver(cfg)
    g.set_option(m.OPT_CONCURRENT, 0 if serial else 1)
    g.set_mesh(*mesh, span)
    cases.jet_gas(g, mech)
    g.set_bc_jet(bc["inlet2d"])
    g.upload_state(st)
    del st
    g.comput_flow_variables()
    g.OCFD_bc()
    ctl = dict(rtol_y=chemistry.RTOL_Y, y_floor=chemistry.Y_FLOOR, rtol_t=chemistry.RTOL_T, max_sub=100000)
    g.set_chem_method(m.CHEM_ADAPTIVE, scheduler=1, **ctl)
    t = time.perf_counter()
    g.Chemical_time_advance(tau)
    g.comput_T_P()
    g.synchronize()
    t_pre = time.perf_counter() - t
    dt = g.cfl_dt(cfl)
    names = ("F", "DI", "D", "U", "V", "W", "T", "P", "ET")
    saved = {k: g.download(k) for k in names}
    T0max = float(np.max(saved["T"]))
    g.set_option(m.OPT_USE_GRAPH, 1)
    ncell = nx * ny * nz
    ctl["max_sub"] = 4000

    def run(method, reaction=True):
        for k in names:
            g.upload(k, saved[k])
        g.set_chem_method(method, scheduler=1, **ctl)
        for _ in range(warmup):
            g.NS_step(dt, 1, 1, reaction)
        g.synchronize()
        for k in names:
            g.upload(k, saved[k])
        g.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            g.NS_step(dt, 1, 1, reaction)
        g.synchronize()
        el = time.perf_counter() - t
        return el / steps * 1e3

    out = dict(workload=f"reacting planar H2 jet {nx}x{ny}x{nz}, 9 species / 19 reactions, WENO5 + SW, CD6, "
                        "real gas, Chemkin/Wilke transport, mixture-averaged diffusion, BC_JET",
               dt=dt, cfl=cfl, pre_react_s=t_pre, pre_react_tau=tau, steps=steps,
               step="NS_solver: flow RK3 dt/2, chemistry dt, comput_T_P, flow RK3 dt/2",
               streams=1 if serial else 2)
    ms_flow = run(m.CHEM_FIXED, reaction=False)
    out["flow_only_ms_per_step"] = ms_flow
    rate = ncell * 2 / ms_flow / 1e3
    out["flow_Mcell_steps_per_s"] = rate
    # step roofline of the reacting flow (two RK3 steps per NS step):
    # min(HBM / 896 B, FP64 / 2.9e4 flops) per cell-step
    from ocfd_amd import flops
    b_alg, f_alg = flops.jet_step_bytes_per_cell(9), flops.JET_STEP
    bh, bf = PEAK_HBM_GBPS * 1e9 / b_alg / 1e6, PEAK_FP64_TFLOPS * 1e12 / f_alg / 1e6
    out["flow_step_roofline"] = dict(F_alg_per_cell=f_alg, B_alg_per_cell=b_alg, hbm_bound=bh, fp64_bound=bf,
                                     unit="Mcell-steps/s", frac=rate / min(bh, bf), hbm_frac=rate / bh)
    ms_plain = run(m.CHEM_ADAPTIVE)
    Tp, dip = g.download("T"), g.download("DI")
    unfinished = None
    out["plain"] = dict(ms_per_step=ms_plain, chem_ms_per_step=ms_plain - ms_flow,
                        integrator="adaptive point-implicit FP64, persistent work queue")
    if S is not None:
        I = (slice(4, -4),) * 3
        Tp_i = Tp[I]
        Yp = dip[I] / dip[I].sum(axis=3, keepdims=True)
        major = [mech.index(k) for k in ("H2", "O2", "H2O", "N2")]

        def surrogate_run(cfg):
            S.set_config(cfg)
            ms_s = run(m.CHEM_SURROGATE)
            nfb = S.fallbacks()
            Ts, dis = g.download("T"), g.download("DI")
            Ts_i = Ts[I]
            Ys = dis[I] / dis[I].sum(axis=3, keepdims=True)
            dY = np.abs(Ys[..., major] - Yp[..., major]).max()
            # why cells fall back: the fused checks on the final field's
            # interior cells, one chemistry step of dt, no fallback
            rho_c = torch.tensor(np.ascontiguousarray(dis[I].reshape(-1, dis.shape[3]).T), device="cuda")
            E_c = torch.tensor(np.ascontiguousarray(g.download("ET")[I].reshape(-1)), device="cuda")
            fl = torch.zeros(E_c.shape[0], dtype=torch.int32, device="cuda")
            S.step(rho_c, E_c, torch.full_like(E_c, dt), fallback=False, flags=fl)
            fl = fl.cpu().numpy()
            del rho_c, E_c
            reasons = {k: int(np.sum((fl & b) != 0)) for k, b in (("domain", 1), ("element_balance", 2),
                                                                   ("negative_Y", 4), ("T_bounds", 8),
                                                                   ("increment_bound", 16))}
            return dict(ms_per_step=ms_s, chem_ms_per_step=ms_s - ms_flow, fallbacks_last_step=nfb,
                        fallback_reasons_final_field=reasons,
                        fallback_fraction=nfb / ncell, speedup_step_vs_plain=ms_plain / ms_s,
                        speedup_chem_vs_plain=(ms_plain - ms_flow) / max(ms_s - ms_flow, 1e-9),
                        error_vs_plain=dict(peak_T_rel=float(abs(Ts_i.max() - Tp_i.max()) / Tp_i.max()),
                                            dT_max_K=float(np.abs(Ts_i - Tp_i).max()), dY_major_max=float(dY)))

        cfg_b = S.cfg.copy()
        cfg_u = cfg_b.copy()
        cfg_u[S.DY_MAX_INDEX] = np.inf
        out["surrogate"] = surrogate_run(cfg_u)
        if np.isfinite(cfg_b[S.DY_MAX_INDEX]):
            out["surrogate"]["with_increment_bound"] = surrogate_run(cfg_b)
        S.set_config(cfg_b)
    out["T_max_initial"] = T0max
    g.close()
    return out


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "flat_plate_640x120x64_traffic.md")


def traffic_table():
    """{kernel: (HBM bytes per launch, launches per RK3 step)} from the
    committed rocprofv3 counter summary of the flat plate (FETCH_SIZE +
    WRITE_SIZE, each scaled by what it reads for a plain copy)."""
    out = {}
    if not os.path.exists(TRAFFIC_PROFILE):
        return out
    col = per = None
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip() for c in line.split("|")]
        if col is None and "traffic (MiB)" in cells:
            col = cells.index("traffic (MiB)")
            per = cells.index("launches per step") if "launches per step" in cells else None
        elif col is not None and len(cells) > col and cells[1] and not cells[1].startswith("-"):
            try:
                out[cells[1]] = (float(cells[col]) * 2 ** 20, float(cells[per]) if per is not None else None)
            except ValueError:
                pass
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=640)
    ap.add_argument("--ny", type=int, default=120)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--serial", action="store_true",
                    help="one stream throughout (no viscous preparation beside the sweeps), e.g. for a kernel "
                         "trace whose per-kernel durations are solo ones")
    ap.add_argument("--no-chem", action="store_true", help="skip the chemistry batch")
    ap.add_argument("--chem-cells", type=int, default=1 << 21)
    ap.add_argument("--surrogate-train-cells", type=int, default=1 << 24, help="0 skips the surrogate")
    ap.add_argument("--surrogate-epochs", type=int, default=25)
    ap.add_argument("--no-jet", action="store_true", help="skip the reacting planar jet")
    ap.add_argument("--only-jet", action="store_true", help="only the reacting planar jet, plain chemistry")
    ap.add_argument("--jet-steps", type=int, default=10)
    ap.add_argument("--jet-nx", type=int, default=256)
    ap.add_argument("--jet-ny", type=int, default=192)
    ap.add_argument("--jet-nz", type=int, default=128)
    ap.add_argument("--surrogate-inference-only", type=int, default=-1, metavar="TILE",
                    help="only the surrogate inference (untrained network, no fallback) on 2^21 cells with the "
                         "given network tile shape, 3 timed calls: a short program for counter runs")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a plain 512 MiB device copy (3 times): the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_MiB": 512, "repeats": 3}))
        return
    if args.surrogate_inference_only >= 0:
        print(json.dumps({"surrogate_inference": surrogate_inference_only(args.surrogate_inference_only)}))
        return
    if args.only_jet:
        m = load_pkg()
        print(json.dumps({"jet": jet_bench(m, None, args.jet_nx, args.jet_ny, args.jet_nz, steps=args.jet_steps,
                                           serial=args.serial)}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design note, not an implemented path.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    m = load_pkg()
    from ocfd_amd import cases, flops
    case = cases.flat_plate(args.nx, args.ny, args.nz)
    cfg, info = case[0], case[5]
    ncell = cfg.nx * cfg.ny * cfg.nz
    g = m.FlowSolver(cfg)
    setup(g, case)
    g.set_option(m.OPT_HYBRID_SUM_MODE, 1)
    g.set_option(m.OPT_CONCURRENT, 0 if args.serial else 1)
    tiles = g.autotune(3)
    g.set_option(m.OPT_USE_GRAPH, 1)
    step = run_steps(g, info, args.warmup, 0)
    g.synchronize()
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    step = run_steps(g, info, args.steps, step)
    g.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = ncell * args.steps / elapsed / 1e6
    f = g.download("F")
    if not np.all(np.isfinite(f)):
        raise RuntimeError("non-finite state after the timed steps")

    # hybrid mask fractions (interior interfaces) of the last step
    frac = {}
    for d in range(3):
        mk = g.download_mask(d)
        sl = [slice(None)] * 3
        sl[d] = slice(1, -1)
        mk = mk[tuple(sl)]
        frac[d] = {v: float(np.mean(mk == v)) for v in (1, 2, 3)}

    roofline = None
    kstats = {}
    if not args.no_kernel_timing:
        # per-launch HIP events on the library stream, over a second timed
        # region of the same steps (graph replay off while events are on);
        # kernels run one at a time here so each duration is the kernel's own
        g.set_option(m.OPT_CONCURRENT, 0)
        g.set_option(m.OPT_KERNEL_TIMING, 1)
        g.kernel_stats_reset()
        nprof = min(args.steps, 20)
        step = run_steps(g, info, nprof, step)
        g.synchronize()
        kstats = g.kernel_stats()
        g.set_option(m.OPT_KERNEL_TIMING, 0)
        g.set_option(m.OPT_CONCURRENT, 0 if args.serial else 1)
        # the dominant kernel among those with an algorithmic count (the
        # sweeps, the viscous preparation and the update; the sensor, mask and
        # boundary kernels are each a few percent of a step)
        counted = {k: v for k, v in kstats.items()
                   if k in ("k_sweep_x", "k_sweep_y", "k_sweep_z", "k_rk", "k_visc_prep")}
        name, (ms, n) = max(counted.items(), key=lambda kv: kv[1][0])
        avg_s = ms / n / 1e3
        per_cell = {
            "k_sweep_x": flops.sweep_flops_per_cell(0, cfg.scheme_invis, cfg.flux_splitting, frac[0]),
            "k_sweep_y": flops.sweep_flops_per_cell(1, cfg.scheme_invis, cfg.flux_splitting, frac[1]),
            "k_sweep_z": flops.sweep_flops_per_cell(2, cfg.scheme_invis, cfg.flux_splitting, frac[2]),
            "k_rk": flops.VISC_DIV + flops.RK + flops.PRIMS,
            "k_visc_prep": flops.VISC_PREP,
        }.get(name)
        # algorithmic HBM bytes per cell of the kernel (one read of each input
        # field, one write of each output field)
        # algorithmic HBM bytes per cell (each input field read once, each
        # output written once): sweeps read d,u,v,w,p,Et + 3 metrics + mask and
        # write du (y, z also read du, z reads Ajac); k_rk reads f,fn,du,Ev(12),
        # d,Et,Ajac and writes f,d,u,v,w,Et,di,T,p,Et_prev; k_visc_prep reads
        # u,v,w,T,d,di + 18 metrics and writes Amu,Amk,Ev(12)
        bytes_cell = {"k_sweep_x": 73 + 40, "k_sweep_y": 73 + 80, "k_sweep_z": 73 + 88,
                      "k_rk": 240 + 112, "k_visc_prep": 192 + 112}.get(name)
        traffic = traffic_table().get(name, (None, None))[0]
        if per_cell is not None:
            fp64 = per_cell * ncell / avg_s / 1e12
            hbm_own = bytes_cell * ncell / avg_s / 1e9
            memory_bound = name in ("k_rk", "k_visc_prep")
            if memory_bound:
                # the compulsory state bytes of a step (320 B per cell-step for
                # 5 conserved FP64 variables), one stage launch = 1/3 step
                b_stage = flops.step_bytes_per_cell(5) / 3.0
                hbm = b_stage * ncell / avg_s / 1e9
                roofline = dict(bound="hbm", kernel=name, achieved=hbm, peak=PEAK_HBM_GBPS, unit="GB/s",
                                frac=hbm / PEAK_HBM_GBPS, traffic=traffic,
                                bytes_per_cell=b_stage,
                                bytes_rule="compulsory state bytes: 320 B per cell per RK3 step, one stage "
                                           "launch = 1/3 step",
                                kernel_own=dict(bytes_per_cell=bytes_cell, achieved_GBps=hbm_own,
                                                frac=hbm_own / PEAK_HBM_GBPS,
                                                rule="every input field read once, every output written once"),
                                fp64_achieved_TFLOPs=fp64)
            else:
                roofline = dict(bound="mfma", pipe="fp64 vector (VALU); MI355X FP64 peak", kernel=name,
                                achieved=fp64, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s", frac=fp64 / PEAK_FP64_TFLOPS,
                                traffic=traffic, flops_per_cell=per_cell, hbm_achieved_GBps=hbm_own)
            roofline.update(cells_per_launch=ncell, avg_launch_ms=avg_s * 1e3,
                            timing="HIP events around every launch on the library stream, second pass of "
                                   f"{nprof} steps (graph replay off, one kernel at a time)")

    # whole-step roofline: measured rate / min(HBM / B_alg, FP64 / F_alg)
    mean_frac = {v: sum(frac[d][v] for d in range(3)) / 3 for v in (1, 2, 3)}
    f_alg = flops.step_flops_per_cell(cfg.scheme_invis, cfg.flux_splitting, mean_frac)
    b_alg = flops.step_bytes_per_cell(5)
    bound_hbm = PEAK_HBM_GBPS * 1e9 / b_alg / 1e6
    bound_fp64 = PEAK_FP64_TFLOPS * 1e12 / f_alg / 1e6
    step_roofline = dict(F_alg_per_cell=f_alg, B_alg_per_cell=b_alg, hbm_bound=bound_hbm, fp64_bound=bound_fp64,
                         unit="Mcell-steps/s", frac=value / min(bound_hbm, bound_fp64),
                         hbm_frac=value / bound_hbm)
    # HBM bytes of a whole RK3 step from the counter summary, against the
    # compulsory 320 B per cell
    tt = traffic_table()
    if tt and all(v[1] is not None for v in tt.values()):
        step_bytes = sum(b * n for b, n in tt.values())
        step_roofline["step_traffic_ratio"] = step_bytes / (b_alg * ncell)
        step_roofline["step_traffic_bytes"] = step_bytes
        step_roofline["step_traffic_source"] = os.path.relpath(TRAFFIC_PROFILE, ROOT)

    # the type-2 sensor's fast mean is certified: report the certificate of
    # the last step, then time the ordered (single-wavefront) mean
    unc, margin, eta = g.hybrid_certificate()
    g.set_option(m.OPT_HYBRID_SUM_MODE, 0)
    nord = min(args.steps, 10)
    step = run_steps(g, info, 2, step)
    g.synchronize()
    t = time.perf_counter()
    step = run_steps(g, info, nord, step)
    g.synchronize()
    ordered_ms = (time.perf_counter() - t) / nord * 1e3
    g.set_option(m.OPT_HYBRID_SUM_MODE, 1)
    mask = dict(mode="fast tree mean, certified", uncertain_interfaces=unc, margin=margin, eta=eta,
                mask_certified=unc == 0,
                masks_equal_ordered="always: uncertain interfaces trigger the ordered mean on the device",
                ordered_mode_ms_per_step=ordered_ms, ordered_mode_Mcell_steps_per_s=ncell / ordered_ms / 1e3)

    cpu = None if args.no_cpu_baseline else cpu_baseline(m, cfg.nx, cfg.ny, cfg.nz)
    g.close()
    chem, S = (None, None) if args.no_chem else chem_bench(m, args.chem_cells, cpu=not args.no_cpu_baseline,
                                                            surrogate_cells=args.surrogate_train_cells,
                                                            surrogate_epochs=args.surrogate_epochs)
    jet = None
    if not args.no_jet:
        jet = jet_bench(m, S, args.jet_nx, args.jet_ny, args.jet_nz, steps=args.jet_steps, serial=args.serial)
    out = {
        "metric": METRIC, "value": value, "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Mach 2.25 flat-plate boundary layer {cfg.nx}x{cfg.ny}x{cfg.nz}, one TVD-RK3 step",
                   "scheme": "HYBRID UD7L/WENO7/WENO5, pressure-gradient sensor (certified fast mean)",
                   "splitting": "Steger-Warming", "viscous": "CD6", "gas": "perfect gas, Sutherland",
                   "graph": "hipGraph replay", "parallelism": "single GPU",
                   "streams": 1 if args.serial else 2,
                   "sweep_tiles": {"xyz"[d]: {"NP": tiles[d][1], "TS": tiles[d][2],
                                              "ms_per_variant": [round(x, 4) for x in tiles[d][3]]}
                                   for d in range(3)},
                   "mask_fractions": {"x": frac[0], "y": frac[1], "z": frac[2]}},
        "roofline": roofline, "step_roofline": step_roofline, "hybrid_mask": mask, "cpu_baseline": cpu,
        "kernels_ms_per_step": {k: v[0] / max(1, min(args.steps, 20)) for k, v in kstats.items()},
        "kernels_launches_per_step": {k: v[1] / max(1, min(args.steps, 20)) for k, v in kstats.items()},
        "chem": chem,
        "jet": jet,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
