"""Throughput of the lattice-Boltzmann step on one MI355X.

Headline: the plain FP32 collide-stream step (BGK + exact-difference forcing,
halfway + interpolated bounce-back, k-omega off) on the 8192 x 1024 channel
with a seeded 16-Gaussian fouling layer on each wall, in MLUPS (million
fluid-node updates per second), with inputs resident in HBM.  Further forms
(FP16 / BF16 population storage, the fused LB + k-omega step on the
4096 x 512 serpentine channel) are reported beside it under "forms".

Algorithmic bytes per lattice update (compulsory state traffic only):
  P32  9*4 read + 9*4 write + 2 (mask)           = 74 B
  P16  9*2 read + 9*2 write + 2                  = 38 B
  K32  P32 + k, omega read/write (16) + WallD (4) = 94 B

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_LU = {"f32": 74.0, "f16": 38.0, "bf16": 38.0, "f64": 146.0, "f32x": 74.0, "k32": 94.0,
        "k16": 58.0, "t": 68.0}
# t: one temperature call per node: T read and written (16), u (8, FP32),
# nu_t (8), the clean-wall spacings (32), the two masks (4)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


GRAPHS = True  # --no-graphs: every kernel launched on its own (for kernel traces)


def setup(w, storage, komega=False, rtol=1e-4, maxit=1000, tune=True):
    from komegafoulingmodel_amd import lib

    L = lib.Lattice(w.nX, w.nY, storage=storage, komega=komega)
    if not GRAPHS:
        L.set_graphs(False)
    L.set_params(w.tau0, w.Fval, w.nu, kappa_rel_tol=rtol, kappa_max_iters=maxit)
    L.build_geometry(w.C0, w.BL, **w.geom_params())
    L.launch = L.autotune(40) if tune else None
    nType = L.geometry("nType")
    rho, ux, uy = w.initial_velocity(nType)
    L.init_state(rho, ux, uy)
    if komega:
        L.init_komega(w.k_ini, w.omega_ini, w.Re)
    fluid = int(np.count_nonzero((nType.astype(np.int32) & 1) == 0))
    return L, fluid


def measure(L, fluid, steps, warmup, blu, options=0):
    """Times exactly `steps` steps after `warmup` untimed ones: the window is
    the step alone (no events inside it), bracketed by stream
    synchronisation.  A second, separate pass of the same length brackets
    every collide-stream launch with events for the kernel's own time."""
    L.step(warmup, options)
    L.synchronize()
    t0 = time.perf_counter()
    L.step(steps, options)
    L.synchronize()
    wall = time.perf_counter() - t0
    mlups = fluid * steps / wall / 1e6
    ms_step, _ = L.time_steps(steps, options, kernel=False)
    _, ms_kernel = L.time_steps(steps, options, kernel=True)
    # with k-omega the captured step is timed as a whole (ms_kernel < 0): the
    # algorithmic bytes are then set against the whole step
    achieved = blu * fluid / ((ms_kernel if ms_kernel > 0 else ms_step) * 1e-3) / 1e9
    out = dict(mlups=mlups, ms_per_step=wall * 1e3 / steps, ms_per_step_events=ms_step,
               ms_per_kernel=ms_kernel, achieved_gbs=achieved, frac=achieved / HBM_PEAK_GBS,
               fluid_nodes=fluid, steps=steps)
    if getattr(L, "launch", None):
        out["launch"] = {"block_x": L.launch[0], "nodes_per_lane": L.launch[1],
                         "steps_per_pass": L.launch[4], "plane_pad": L.launch[2]}
    return out


def measure_adjoint(theta_scale=1.0, **kw):
    """Wall-shape gradient on the 2048x256 serpentine channel (64 control
    points per wall, T = 4000, J over the last 1000 steps, checkpoints every
    64, k-omega on unless komega=False): plain forward, objective run with
    checkpoints, and the adjoint sweep (recompute + adjoint), each timed end
    to end."""
    from komegafoulingmodel_amd import sensitivity as S

    P = S.WallShapeProblem(**kw)
    th = theta_scale * P.initial_theta()
    L = P.lattice()
    P.build(th)
    launch = L.autotune(20)  # the forward it is compared with runs its best shape
    fluid = int(np.count_nonzero((L.geometry("nType").astype(np.int32) & 1) == 0))
    L.step(64)  # warm-up of the plain path
    L.synchronize()
    P.build(th)
    t0 = time.perf_counter()
    L.step(P.steps)
    L.synchronize()
    t_fwd = time.perf_counter() - t0
    P.build(th)
    t0 = time.perf_counter()
    J = L.objective_run(P.steps, P.interval, P.window, shear_weight=P.shear_weight,
                        frozen_turbulence=P.frozen_turbulence)
    t_obj = time.perf_counter() - t0
    t0 = time.perf_counter()
    dJdy, _ = L.adjoint()
    t_adj = time.perf_counter() - t0
    g = P.dy_dtheta(P._C0).T @ dJdy
    P.close()
    mode = ("k-omega, eddy viscosity frozen in the adjoint" if P.komega and P.frozen_turbulence
            else ("k-omega, exact adjoint" if P.komega else "laminar"))
    shape = "serpentine channel" if P.amp else "flat-wall channel (walls at half-integer rows)"
    return {"workload": f"{P.nX}x{P.nY} {shape}, B-spline fouling x {theta_scale:g}, {mode}, "
                        f"T={P.steps}, window {P.window}, checkpoints every {P.interval}",
            "fluid_nodes": fluid, "J": J, "grad_norm": float(np.linalg.norm(g)),
            "grad_finite": bool(np.all(np.isfinite(g))),
            "launch": {"block_x": launch[0], "nodes_per_lane": launch[1],
                       "steps_per_pass": launch[4], "plane_pad": launch[2]},
            "forward_ms": t_fwd * 1e3, "forward_mlups": fluid * P.steps / t_fwd / 1e6,
            "objective_run_ms": t_obj * 1e3, "adjoint_ms": t_adj * 1e3,
            "adjoint_cost_ratio": t_adj / t_fwd,
            "gradient_cost_ratio": (t_obj + t_adj) / t_fwd}


def measure_ensemble(storage, steps=300):
    """48 channels of uneven size, Re and fouling stepped by the ensemble
    (graph-captured blocks of 32 steps) against the same channels stepped one
    after another, each autotuned and with its own graphs."""
    from komegafoulingmodel_amd import channel, lib

    ws = channel.ensemble_members()
    Ls, fluid = [], 0
    for w in ws:
        L, nf = setup(w, storage, tune=False)
        L.autotune(20)  # each channel alone gets its own best launch shape
        Ls.append(L)
        fluid += nf
    for L in Ls:  # sequential reference, warm (graphs captured)
        L.step(64)
    for L in Ls:
        L.synchronize()
    t0 = time.perf_counter()
    for L in Ls:
        L.step(steps)
    for L in Ls:
        L.synchronize()
    t_seq = time.perf_counter() - t0
    ens = lib.Ensemble(Ls)
    # the ensemble's own autotuner picks the launch shape (one step per launch,
    # or two with 64 x 13 / 64 x 26 tiles) and the cache budget of the groups
    block, spp, budget, _ = ens.autotune(64)
    groups = ens.set_cache_budget(budget)
    ens.step(64)
    ms, _ = ens.time_steps(steps)
    out = {"channels": len(ws), "fluid_nodes": fluid, "steps": steps,
           "sequential_mlups": fluid * steps / t_seq / 1e6,
           "autotuned": {"block_x": block, "steps_per_launch": spp, "cache_budget": budget,
                         "groups": groups},
           "mlups": fluid / (ms * 1e-3) / 1e6, "ms_per_step": ms,
           "speedup_vs_sequential": (t_seq / steps) / (ms * 1e-3)}
    ens.close()
    for L in Ls:
        L.close()
    return out


def measure_precision(name="c3", steps=100_000):
    """A channel run `steps` steps from the same seeded state with FP64, FP32,
    FP16 and BF16 populations, each on its autotuned launch shape. For each
    field, the error of the last step against the FP64 run: the largest
    difference over fluid nodes (wall nodes and shear links for the wall
    shear), relative to the largest FP64 magnitude. u is compared as a
    vector; ux and uy are each relative to max |u|."""
    from komegafoulingmodel_amd import channel, lib

    w = channel.config(name)
    runs = {}
    for st in ("f64", "f32", "f32x", "f16", "bf16"):
        t0 = time.perf_counter()
        L, _ = setup(w, st)
        L.step(steps - 1)
        L.step(1, lib.OPTION_SAVE_MACRO_FIELDS)
        L.wall_shear()
        fluid = (L.geometry("nType").astype(np.int32) & 1) == 0
        runs[st] = {"ux": L.download("ux")[fluid], "uy": L.download("uy")[fluid],
                    "tau_node": L.download("tau_node"), "tau_bl": L.download("tau_bl")}
        L.close()
        log(f"prec {name} {st}: {steps} steps in {time.perf_counter() - t0:.1f} s")
    ref = runs["f64"]
    umax = float(np.max(np.hypot(ref["ux"], ref["uy"])))
    out = {"workload": f"{name}: {w.nX}x{w.nY}, k-omega off, seeded state", "steps": steps,
           "reference": "FP64 populations and arithmetic on the GPU", "max_u_fp64": umax}
    for st in ("f32", "f32x", "f16", "bf16"):
        r = runs[st]
        out[st] = {
            "u": float(np.max(np.hypot(r["ux"] - ref["ux"], r["uy"] - ref["uy"]))) / umax,
            "ux": float(np.max(np.abs(r["ux"] - ref["ux"]))) / umax,
            "uy": float(np.max(np.abs(r["uy"] - ref["uy"]))) / umax,
            "tau_node": float(np.max(np.abs(r["tau_node"] - ref["tau_node"])
                                     / np.max(np.abs(ref["tau_node"])))),
            "tau_bl": float(np.max(np.abs(r["tau_bl"] - ref["tau_bl"])
                                   / np.max(np.abs(ref["tau_bl"])))),
        }
    return out


def measure_ensemble_ko(storage, steps=40):
    """The 48 channels with k-omega on: ensemble steps (one batched
    collide-stream launch and every member's implicit solves captured per
    step, one poll per step) against the same channels stepped one after
    another."""
    from komegafoulingmodel_amd import channel, lib

    ws = channel.ensemble_members(komega=True)
    Ls, fluid = [], 0
    for w in ws:
        L, nf = setup(w, storage, komega=True, tune=False)
        Ls.append(L)
        fluid += nf
    for L in Ls:  # warm (graphs captured, iteration predictions settled)
        L.step(8)
    for L in Ls:
        L.synchronize()
    t0 = time.perf_counter()
    for L in Ls:
        L.step(steps)
    for L in Ls:
        L.synchronize()
    t_seq = time.perf_counter() - t0
    ens = lib.Ensemble(Ls)
    # the autotuner picks the work-group width and the cache budget of the
    # groups (all members every step, or groups whose k-omega working set
    # fits the last-level cache, stepped one group at a time)
    block, spp, budget, _ = ens.autotune(24)
    groups = ens.set_cache_budget(budget)
    ens.step(8)
    ms, _ = ens.time_steps(steps)
    out = {"channels": len(ws), "fluid_nodes": fluid, "steps": steps, "komega": True,
           "sequential_mlups": fluid * steps / t_seq / 1e6,
           "autotuned": {"block_x": block, "steps_per_launch": spp, "cache_budget": budget,
                         "groups": groups},
           "mlups": fluid / (ms * 1e-3) / 1e6, "ms_per_step": ms,
           "speedup_vs_sequential": (t_seq / steps) / (ms * 1e-3)}
    ens.close()
    for L in Ls:
        L.close()
    return out


def measure_reference_step(L, fluid, steps=240, warmup=60):
    """The reference's production step semantics: rho and u written every
    step (OPTION_SAVE_MACRO_FIELDS, LBLS.cpp:348-349) and the wall shear of
    nodes and boundary links every 3 steps (LBLS.cpp:237-241, 322-326),
    timed over `steps` steps with no events inside the window."""
    from komegafoulingmodel_amd import lib

    def run(n):
        for _ in range(n // 3):
            L.step(3, lib.OPTION_SAVE_MACRO_FIELDS)
            L.wall_shear()

    run(warmup)
    L.synchronize()
    t0 = time.perf_counter()
    run(steps)
    L.synchronize()
    wall = time.perf_counter() - t0
    return {"mlups": fluid * steps / wall / 1e6, "ms_per_step": wall * 1e3 / steps,
            "steps": steps, "fluid_nodes": fluid,
            "semantics": "macros every step, wall shear every 3 steps"}


def measure_thermal(calls=50):
    """The temperature solve (clVariablesFD::Solve) on the 4096x512 k-omega
    channel after 200 k-omega steps: the time of one call with the flow
    held (every call assembles the rows and solves to the reference's
    tolerances), of the reference's step, LB + k-omega + temperature, and of
    one Nusselt record read back to the host."""
    from komegafoulingmodel_amd import channel, lib

    wk = channel.config("c2")
    log(f"t32 on {wk.nX}x{wk.nY}")
    L, n = setup(wk, "f32", komega=True)
    rows = int(np.count_nonzero((L.geometry("nType").astype(np.int32) & 0x4) == 0))
    L.thermal_init(alpha_fluid=wk.nu / 0.708, alpha_foul=(0.057 / 35.0 / 866.8) * wk.nu / 4.1e-5,
                   pr_turb=0.9, t_in=2.0, t_initial=0.5)
    opt = lib.OPTION_SAVE_MACRO_FIELDS
    L.step(200, opt)
    for _ in range(20):
        L.step(1, opt)
        L.thermal_step(1)
    L.synchronize()
    t0 = time.perf_counter()
    L.thermal_step(calls)
    L.synchronize()
    ms_t = (time.perf_counter() - t0) * 1e3 / calls
    st = L.solver_stats()
    t0 = time.perf_counter()
    for _ in range(calls):
        L.step(1, opt)
        L.thermal_step(1)
    L.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / calls
    # the Nusselt record (clVariablesFD::updateTimeData) at the period
    # boundaries x = 0, 1024, 2048, 3072, copied to the host each call
    L.nusselt_init(height=2 * int(wk.pipe_radius), wavy_start=0, wavy_period=1024,
                   wavy_periods=3, k_soot=0.057, k_air=0.042)
    cols = L.nusselt_sizes()[0]
    L.nusselt()
    t0 = time.perf_counter()
    for _ in range(calls):
        L.nusselt()
    ms_nu = (time.perf_counter() - t0) * 1e3 / calls
    L.close()
    achieved = B_LU["t"] * rows / (ms_t * 1e-3) / 1e9
    return {"rows": rows, "ms_per_call": ms_t, "mrows_per_s": rows / ms_t / 1e3,
            "achieved_gbs": achieved, "frac": achieved / HBM_PEAK_GBS, "bytes_per_row": B_LU["t"],
            "iterations_last": st.temp_iters,
            "residual_last": st.temp_res / st.temp_res0 if st.temp_res0 else 0.0,
            "ms_per_step_lb_komega_temperature": ms_step,
            "mlups_with_temperature": n / ms_step / 1e3,
            "nusselt_columns": cols, "ms_per_nusselt_record": ms_nu}


def measure_removal(n=1 << 26, calls=50):
    """TR_shear_removal over n particles (the tracer's device arrays) sitting
    on random links of the 8192x1024 fouled channel's shear range, with the
    link shear of its flow after 100 steps, in link order (the order the
    tracer's sort by location leaves them in).  Two passes: every particle
    below its critical shear, so each call reads the flag, size and deposit
    timer and writes the timer back (10 B per particle; the steady state of
    particles on a wall), timed over `calls` calls; and every particle
    re-entrained (critical shear 0) from fresh arrays, reading 8 B and
    writing position, flag, location and timer (36 B per particle), the
    fastest of 3 single calls, each bracketed by synchronisation.  The
    countdown pass is also timed with the particles in random order."""
    import ctypes as C

    from komegafoulingmodel_amd import channel, lib

    w = channel.config("c3")
    L, _ = setup(w, "f32", tune=False)
    L.step(100, lib.OPTION_SAVE_MACRO_FIELDS)
    L.wall_shear()
    b = L.geometry("bounds")
    tau = L.download("tau_bl")
    links = np.r_[np.arange(b[0], b[0] + b[1]), np.arange(b[2], b[2] + b[3])]
    links = links[np.abs(tau) > 0]
    rng = np.random.Generator(np.random.PCG64(3))
    host = dict(pos=np.zeros(2 * n), num_rep=np.ones(n, np.uint32),
                type=rng.integers(0, 4, n).astype(np.int16),
                dep_flag=np.sort(rng.choice(links, n)).astype(np.int32),
                dep_timer=np.full(n, 60000, np.uint16), loc=np.zeros(n, np.int32),
                timer=np.zeros(n, np.int32))
    h = C.CDLL(lib.load()._name)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    dev = {}
    for k, a in host.items():
        ptr = C.c_void_p()
        assert h.hipMalloc(C.byref(ptr), a.nbytes) == 0
        dev[k] = ptr
    dep = C.c_void_p()
    assert h.hipMalloc(C.byref(dep), 4 * 4 * len(w.BL)) == 0

    def upload(keys):
        for k in keys:
            assert h.hipMemcpy(dev[k], host[k].ctypes.data_as(C.c_void_p), host[k].nbytes, 1) == 0

    upload(host)
    addr = {k: v.value for k, v in dev.items()}
    k = dict(drag_coeff=np.full(4, 1e3), lift_coeff=np.full(4, 1e2), mass=np.ones(4),
             loc_stride=8192, timer_start=1000)
    stay = np.full(8, 1e30)
    for _ in range(2):
        L.shear_removal(addr, dep.value, 0, n, tau_crit=stay, **k)
    L.synchronize()
    t0 = time.perf_counter()
    for _ in range(calls):
        L.shear_removal(addr, dep.value, 0, n, tau_crit=stay, **k)
    L.synchronize()
    ms_stay = (time.perf_counter() - t0) * 1e3 / calls
    sorted_flags = host["dep_flag"]
    host["dep_flag"] = rng.permutation(sorted_flags)
    upload(("dep_flag",))
    L.synchronize()
    t0 = time.perf_counter()
    for _ in range(calls):
        L.shear_removal(addr, dep.value, 0, n, tau_crit=stay, **k)
    L.synchronize()
    ms_random = (time.perf_counter() - t0) * 1e3 / calls
    host["dep_flag"] = sorted_flags
    ms_leave = float("inf")
    for _ in range(3):
        upload(("dep_flag", "dep_timer"))
        L.synchronize()
        t0 = time.perf_counter()
        L.shear_removal(addr, dep.value, 0, n, tau_crit=np.zeros(8), **k)
        L.synchronize()
        ms_leave = min(ms_leave, (time.perf_counter() - t0) * 1e3)
    for v in list(dev.values()) + [dep]:
        h.hipFree(v)
    L.close()
    g_stay = 10.0 * n / (ms_stay * 1e-3) / 1e9
    g_leave = 36.0 * n / (ms_leave * 1e-3) / 1e9
    return {"particles": n, "links": int(len(links)), "ms_per_call_countdown": ms_stay,
            "achieved_gbs_countdown": g_stay, "frac_countdown": g_stay / HBM_PEAK_GBS,
            "ms_per_call_countdown_random_order": ms_random,
            "ms_per_call_reentrain": ms_leave, "achieved_gbs_reentrain": g_leave,
            "frac_reentrain": g_leave / HBM_PEAK_GBS, "bytes_per_particle": [10, 36]}


def kernel_name(launch, storage="float"):
    if launch.get("steps_per_pass", 1) == 2:
        tile = {512: (64, 8), 640: (128, 5), 832: (64, 13), 1856: (64, 29), 1024: (64, 16),
                1536: (128, 12), 1664: (64, 26), 3584: (128, 28)}[launch["block_x"]]
        k = ("lbCollideStreamT2P" if launch["block_x"] in (1024, 1536, 1664, 3584)
             else "lbCollideStreamT2")
        return f"{k}<{storage},float,{tile[0]},{tile[1]}>"
    k = "lbCollideStream2" if launch.get("nodes_per_lane", 1) == 2 else "lbCollideStream"
    return f"{k}<float,float,false,{launch['block_x']}>"


def read_traffic(name, kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (committed summary, one `traffic_bytes_per_launch[<kernel>]:` line per
    measured kernel), or None when that kernel was not measured."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.txt")
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch[{kernel}]:"
    for line in open(path):
        if line.startswith(key):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(w, seconds=15.0, max_steps=40):
    """The FP64 oracle (the reference algorithm restated, OpenMP over rows) on
    the host cores, on a bounded number of steps of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    g = oracle.geometry(w.nX, w.nY, w.C0, w.BL, **w.geom_params())
    threads = oracle.load().orc_max_threads()
    orc = oracle.OracleLattice(g, tau0=w.tau0, Fval=w.Fval, nu=w.nu, threads=threads)
    rho, ux, uy = w.initial_velocity(g["nType"])
    orc.init_state(rho, ux, uy)
    fluid = int(np.count_nonzero((g["nType"].astype(np.int32) & 1) == 0))
    orc.step(1)
    t0 = time.perf_counter()
    orc.step(2)
    per = (time.perf_counter() - t0) / 2
    n = int(max(2, min(max_steps, seconds / max(per, 1e-9))))
    t0 = time.perf_counter()
    orc.step(n)
    dt = time.perf_counter() - t0
    return dict(value=fluid * n / dt / 1e6, unit="MLUPS", cores=int(threads), kind="port",
                sample=f"{n} FP64 steps of the {w.nX}x{w.nY} fouled channel (oracle, OpenMP)",
                cpu_model=cpu_model())


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_calibration(nbytes=2 << 30, reps=20):
    """A plain device-to-device copy of known size (torch's copy kernel), the
    reference for scaling the FETCH_SIZE / WRITE_SIZE counters and for the
    achievable copy bandwidth."""
    import torch

    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "ms": ms,
                      "gbs": 2 * n * 4 / (ms * 1e-3) / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--forms", default="p32_ref,f32x,f16,bf16,prec,k32,k16,k32_c3,t32,rem32,adj32,adj32_exact,adj32_exact_smooth,adj32_c3,ens32,ens16,ens32ko,ens16ko")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--launch", default="",
                    help="block_x,nodes_per_lane,steps_per_pass,plane_pad for the headline "
                         "form instead of the autotuner's choice (for profiling one kernel)")
    ap.add_argument("--prec-steps", type=int, default=100_000,
                    help="steps of the precision form (errors against FP64 at 8192x1024)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel on its own instead of replaying captured graphs "
                         "(the same kernels and work; for per-kernel traces)")
    ap.add_argument("--launch-16bit", default="",
                    help="the same for the f16 and bf16 forms")
    args = ap.parse_args()
    global GRAPHS
    GRAPHS = not args.no_graphs
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the lattice step runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    from komegafoulingmodel_amd import channel

    w = channel.config("c3")
    log(f"workload c3 {w.nX}x{w.nY} fouled, FP32, {args.steps} steps")
    L, fluid = setup(w, "f32", tune=not args.launch)
    if args.launch:
        b, n, sp, pad = (int(v) for v in args.launch.split(","))
        L.set_launch(b, n, pad, sp)
        L.launch = (b, n, pad, 0.0, sp)
    r = measure(L, fluid, args.steps, args.warmup, B_LU["f32"])
    log(f"p32: {r}")
    forms = {"p32": r}
    if "p32_ref" in args.forms.split(","):
        forms["p32_ref"] = measure_reference_step(L, fluid)
        forms["p32_ref"]["vs_p32"] = forms["p32_ref"]["mlups"] / r["mlups"]
        log(f"p32_ref: {forms['p32_ref']}")
    L.close()
    for form in [f for f in args.forms.split(",") if f]:
        if form == "f32x":
            # FP32 storage, FP64 collisions, stochastically rounded stores
            Lf, nf = setup(w, "f32x")
            forms["p32x"] = measure(Lf, nf, args.steps, args.warmup, B_LU["f32x"])
            Lf.close()
        elif form in ("f16", "bf16"):
            Lf, nf = setup(w, form, tune=not args.launch_16bit)
            if args.launch_16bit:
                b, n, sp, pad = (int(v) for v in args.launch_16bit.split(","))
                Lf.set_launch(b, n, pad, sp)
                Lf.launch = (b, n, pad, 0.0, sp)
            forms["p" + form[1:] if form == "f16" else "pbf16"] = measure(
                Lf, nf, args.steps, args.warmup, B_LU[form])
            Lf.close()
        elif form in ("k32", "k16", "k32_c3"):
            wk = channel.config("c3" if form == "k32_c3" else "c2")
            if form == "k32_c3":  # the roofline shape with k-omega on
                wk.komega, wk.Re = True, 10000.0
                wk.k_ini = 1.5 * (0.05 * 0.05) ** 2
                wk.omega_ini = math.sqrt(wk.k_ini) / (0.09 ** 0.25 * 0.07 * (wk.nY - 2))
            log(f"{form} on {wk.nX}x{wk.nY}")
            Lk, nk = setup(wk, "f16" if form == "k16" else "f32", komega=True)
            # past the start-up transient (the first steps need tens of
            # Krylov iterations), then a timed window at the reference tolerances
            forms[form] = measure(Lk, nk, 50, 200, B_LU["k16" if form == "k16" else "k32"])
            st = Lk.solver_stats()
            forms[form]["solver_sweeps_last"] = [st.kappa_iters, st.omega_iters]
            # the achieved relative residuals of the last solves (tolerance 1e-4)
            forms[form]["residual_last"] = [st.kappa_res / st.kappa_res0 if st.kappa_res0 else 0.0,
                                            st.omega_res / st.omega_res0 if st.omega_res0 else 0.0]
            Lk.close()
        elif form == "t32":
            forms["t32"] = measure_thermal()
        elif form == "rem32":
            forms["rem32"] = measure_removal()
        elif form in ("ens32", "ens16"):
            forms[form] = measure_ensemble("f32" if form == "ens32" else "f16")
        elif form in ("ens32ko", "ens16ko"):
            forms[form] = measure_ensemble_ko("f32" if form == "ens32ko" else "f16")
        elif form == "prec":
            # the configured lengths: 100,000 steps at 8192x1024 and 20,000 at 1024x128
            forms["prec"] = measure_precision("c3", args.prec_steps)
            forms["prec_c1"] = measure_precision("c1", max(1, args.prec_steps // 5))
        elif form == "adj32":
            # the 2048x256 k-omega sensitivity workload; its k-omega transient
            # takes k and omega through the clamps, where the exact tangent
            # overflows, so the full run uses the frozen-eddy-viscosity adjoint
            forms["adj32"] = measure_adjoint(frozen_turbulence=True)
        elif form == "adj32_exact":
            # the exact k-omega adjoint (transposed solves, row, diffusivity and
            # collision adjoints) on the same channel over a horizon it stays finite
            forms["adj32_exact"] = measure_adjoint(steps=64, window=16, interval=16)
        elif form == "adj32_exact_smooth":
            # the exact k-omega adjoint over the full T = 4000 on the same grid
            # with flat walls and a fouling layer of at most 0.2 nodes, where k
            # and omega stay positive and the clamps never act
            forms["adj32_exact_smooth"] = measure_adjoint(theta_scale=0.05, amp=0.0)
        elif form == "adj32_c3":
            # the laminar gradient at the 8192x1024 roofline shape (bandwidth bound)
            forms["adj32_c3"] = measure_adjoint(nX=8192, nY=1024, H=840.0, amp=84.0,
                                                steps=512, window=128, interval=32, komega=False)
        log(f"{form}: done")
    for k, v in forms.items():
        forms[k] = {a: (round(b, 6) if isinstance(b, float) else b) for a, b in v.items()}
    if "p16" in forms:
        forms["p16"]["speedup_vs_p32"] = round(forms["p16"]["mlups"] / r["mlups"], 3)
    if "p32x" in forms:
        forms["p32x"]["speedup_vs_p32"] = round(forms["p32x"]["mlups"] / r["mlups"], 3)
    if "pbf16" in forms:
        forms["pbf16"]["speedup_vs_p32"] = round(forms["pbf16"]["mlups"] / r["mlups"], 3)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(w, seconds=args.cpu_seconds)

    traffic = read_traffic("c3_p32", kernel_name(r["launch"]))
    # the counters give bytes per launch, a launch makes steps_per_pass steps
    launch_ms = r["ms_per_kernel"] * r["launch"]["steps_per_pass"]
    out = {
        "metric": "MLUPS + % HBM roofline: FP32 vs FP16-store fwd, adjoint cost ratio",
        "value": round(r["mlups"], 3),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fouling layer, seeded +-1e-4 velocity noise)",
        "config": {"workload": "c3: 8192x1024 channel, 16-Gaussian fouling per wall, tau=0.8, "
                               "F=1e-6, k-omega off, FP32 populations",
                   "nX": w.nX, "nY": w.nY, "fluid_nodes": fluid, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_gbs"], 3),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(r["frac"], 4),
                     "traffic": traffic, "bytes_per_lu": B_LU["f32"],
                     # HBM bytes the counters saw per launch over the launch time
                     "frac_counters": (round(traffic / (launch_ms * 1e-3) / 1e9
                                             / HBM_PEAK_GBS, 4) if traffic else None),
                     "kernel": kernel_name(r["launch"]),
                     "kernel_ms": round(r["ms_per_kernel"], 6),
                     "kernel_ms_per_launch": round(launch_ms, 6)},
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
