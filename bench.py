"""Benchmark: FP64 HEX64 explicit RK4 step of the compressible Taylor-Green
vortex on one MI355X, reported as GDOF/s per RK4 stage.

A step is rk_4_main (4 RHS stages + stage updates + entropy viscosity) plus
the CFL time-step evaluation that the reference's time loop runs after it,
on a periodic box of n^3 HEX64 elements (default n = 64: 262,144 elements,
7,077,888 working nodes).  Inputs are resident in HBM before timing starts.

Prints one JSON line (see the fields below).  --gpus N with N > 1 exits: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 vector


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_ms(name):
    from sod2d_refmap_amd import _lib
    tot, cnt = ctypes.c_double(0.0), ctypes.c_int(0)
    _lib.call("sod_profile_get", name.encode(), ctypes.byref(tot), ctypes.byref(cnt))
    return tot.value, cnt.value


def _host_threads() -> int:
    """Host threads the CPU baseline may use: the CPUs this process may run
    on, capped by OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sol, m, gas, state, seconds: float):
    """The oracle (the C restatement of the reference routines, built with
    OpenMP: element and node loops on all host threads, atomic scatters)
    stepping the same workload as the GPU -- TGV on the same mesh, with the
    geometry the GPU run uses -- one rk_4_main + adapt_dt_cfl at a time, as
    the reference's time loop does, until `seconds` have passed (at least
    one step), timed per step like CFDSolverBase.f90:1552-1575."""
    import numpy as np
    import oracle
    from sod2d_refmap_amd.device import to_host
    threads = _host_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    oracle.use_openmp(True)
    t = oracle.hex64_tables()
    ne, npn = m.nelem, m.npoin
    He = to_host(sol.He, (3, 3, 64, ne))
    gpvol = to_host(sol.gpvol, (1, 64, ne))
    Ml = to_host(sol.Ml, (npn,))
    helem = to_host(sol.helem, (ne,))
    helem_l = to_host(sol.helem_l, (ne, 64))
    s = {k: np.asfortranarray(v).copy(order="F") for k, v in state.items()}
    p = oracle.default_params(Rgas_inf=gas.Rgas)
    zero = np.zeros((ne, 64), order="F")
    steps = 0
    t0 = time.perf_counter()
    while True:
        dt = oracle.adapt_dt_cfl(m.connec, helem, s["u"], s["csound"], 0.5, 0.5, s["mu_fluid"], zero, s["rho"])
        oracle.rk_4_main(m.connec, m.lpoin_w, He, gpvol, t, Ml, helem, helem_l, dt, gas.Rgas, gas.gamma_gas,
                         gas.Cp, gas.Prt, s["mu_factor"], s, p)
        steps += 1
        el = time.perf_counter() - t0
        log(f"cpu baseline: step {steps} at {el:.1f}s")
        if el >= seconds:
            break
    oracle.use_openmp(False)
    gdofs = 5.0 * m.npoin_w * 4 * steps / el / 1e9
    return {"value": gdofs, "unit": "GDOF/s per RK4 stage", "cores": threads, "kind": "port",
            "sample": f"oracle rk_4_main + adapt_dt_cfl built with OpenMP, the same {m.n[0]}^3 TGV workload "
                      f"({m.npoin_w} working nodes), {steps} steps in {el:.1f} s on {threads} host threads "
                      f"({_cpu_model()})"}


def measured_traffic(kernel: str, n: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a copy
    kernel of known size), when one exists for this workload."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_tgv{n}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def kernel_symbol(kernel: str) -> str:
    """The HIP kernel behind a profiling range name."""
    return {"stage_rhs": "k_stage_rhs", "stage_update": "k_stage_update",
            "smart_visc_fused": "k_smart_visc_lane"}.get(kernel, kernel)


def measured_flops_per_wave(kernel: str, n: int):
    """FP64 operations per wavefront (one element) of `kernel`: 2 per FMA,
    1 per add or multiply, from the SQ instruction counters of the committed
    rocprofv3 summary for this workload, when one exists."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_tgv{n}_sq_counters.txt")
    if not os.path.exists(path):
        return None
    per_wave = {}
    section = None
    for line in open(path):
        f = line.split()
        if line.startswith("kernel "):
            section = line.split(None, 1)[1].strip()
        elif section is not None and section.startswith(kernel_symbol(kernel)) and len(f) == 3 \
                and f[0].startswith("SQ_INSTS_VALU_"):
            per_wave[f[0]] = float(f[2])
    try:
        return 64 * (2 * per_wave["SQ_INSTS_VALU_FMA_F64"] + per_wave["SQ_INSTS_VALU_ADD_F64"] +
                     per_wave["SQ_INSTS_VALU_MUL_F64"])
    except KeyError:
        return None


def curvilinear_rate(steps: int, n: int):
    """The randomised curvilinear workload: n^3 HEX64 elements in a periodic
    box with the smooth node mapping x_i += 0.05 h sin(2 pi x_j/L)
    sin(2 pi x_k/L), elements numbered by a seeded random permutation,
    Taylor-Green plus seeded Fourier modes (|k| <= 8, amplitude 1e-2),
    Re 3200, Mach 0.2.  Two untimed and `steps` timed eager steps (drop-in
    rk_4_main + adapt_dt_cfl, as for the main workload): GDOF/s per RK4 stage,
    and the stage kernel's and the whole stage's algorithmic bytes against
    the HBM peak."""
    import torch
    from sod2d_refmap_amd import _lib, cases, mesh
    from sod2d_refmap_amd.solver import PeriodicSolver
    t0 = time.perf_counter()
    m = mesh.periodic_box(n, curved_amp=0.05, perm_seed=1)
    gas = cases.GasParams(Re=3200.0, Mach=0.2)
    sol = PeriodicSolver(m, gas, cases.placeholder_state(m, gas), cfl_conv=0.5, cfl_diff=0.5)
    cases.tgv_state_on_device(sol, m, gas, modes_seed=2)
    log(f"curvilinear {n}^3: nelem={m.nelem} npoin_w={m.npoin_w} ({time.perf_counter() - t0:.1f}s)")
    sol.adapt_dt()
    for _ in range(2):
        sol.rk_4_main(sol.dt)
        sol.adapt_dt()
    torch.cuda.synchronize()
    _lib.call("sod_profile_enable", 1)
    _lib.call("sod_profile_reset")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sol.rk_4_main(sol.dt)
        sol.adapt_dt()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    _lib.call("sod_profile_enable", 0)
    ms, cnt = kernel_ms("stage_rhs")
    per_step = {}
    for name in ("stage_rhs", "stage_update", "smart_visc_fused"):
        kms, kc = kernel_ms(name)
        if kc:
            per_step[name] = round(kms / steps, 4)
    Ne, Nn = m.nelem, m.npoin_w
    alg = Ne * 5888 + Nn * 96
    stage_bytes = Ne * 5888 + Nn * 272
    stages = 4 * steps
    kern_gbs = alg / (ms / cnt / 1e3) / 1e9
    stage_gbs = stage_bytes / (el / stages) / 1e9
    log(f"curvilinear {n}^3: {steps} steps in {el:.3f}s")
    res = {"workload": f"curvilinear_permuted_fourier_hex64_{n}^3_re3200_mach0.2", "elements": Ne,
           "working_nodes": Nn, "steps": steps, "ms_per_step": round(el / steps * 1e3, 4),
           "GDOF_s": round(5.0 * Nn * stages / el / 1e9, 4),
           "kernels_ms_per_step": per_step,
           "stage_rhs_ms": round(ms / cnt, 4), "stage_rhs_frac_of_hbm_peak": round(kern_gbs / HBM_PEAK_GBS, 4),
           "stage_frac_of_hbm_peak": round(stage_gbs / HBM_PEAK_GBS, 4)}
    sol.close()
    del sol
    torch.cuda.empty_cache()
    return res


def adjoint_ratio(sol, m, reps: int = 5):
    """Time of one adjoint (VJP) of the full stage RHS over one forward RHS,
    both through the C ABI on the solver's current state (HIP events on the
    library stream)."""
    import torch
    from sod2d_refmap_amd import _lib
    g = sol.gas
    n, ne = m.npoin, m.nelem
    z = lambda k: torch.zeros(k, dtype=torch.float64, device="cuda")
    rho, q, E = sol.level("rho"), sol.level("q"), sol.level("E")
    st = [sol.connec.data_ptr(), sol.lpoin_w.data_ptr(), sol.He.data_ptr(), sol.gpvol.data_ptr(), sol.Ml.data_ptr()]
    common = st + [rho.data_ptr(), q.data_ptr(), E.data_ptr(), sol.mu_factor.data_ptr(), sol.mu_e.data_ptr(),
                   sol.mu_sgs.data_ptr(), g.Rgas, g.gamma_gas, g.Cp, g.Prt]
    outs = [z(n), z(3 * n), z(n), z(n)]
    lam = [torch.randn(k * n, dtype=torch.float64, device="cuda") for k in (1, 3, 1, 1)]
    gout = [z(n), z(3 * n), z(n), z(64 * ne)]
    fwd = lambda: _lib.call("sod_full_rhs", ne, n, m.npoin_w, *common, *[o.data_ptr() for o in outs])
    adj = lambda: _lib.call("sod_full_rhs_vjp", ne, n, m.npoin_w, *common, *[x.data_ptr() for x in lam],
                            *[o.data_ptr() for o in gout])
    fwd()
    adj()
    _lib.call("sod_profile_enable", 1)
    _lib.call("sod_profile_reset")
    for _ in range(reps):
        fwd()
        adj()
    torch.cuda.synchronize()
    f_ms, f_n = kernel_ms("full_rhs")
    a_ms, a_n = kernel_ms("full_rhs_vjp")
    _lib.call("sod_profile_enable", 0)
    return {"forward_rhs_ms": round(f_ms / max(f_n, 1), 4), "vjp_ms": round(a_ms / max(a_n, 1), 4),
            "ratio": round((a_ms / max(a_n, 1)) / (f_ms / max(f_n, 1)), 3) if f_n and a_n else None}


def batch_speedup(count: int, steps: int):
    """The seeded ensemble of `count` TGV boxes (n^3, n in {8,12,16,24,32},
    Re ~ U[400,3200], Mach ~ U[0.05,0.3], 1e-2 Fourier perturbation): `steps`
    steps of the batched device loop against the same steps taken member by
    member through the plain entry points, and the largest relative L2
    difference between the two results."""
    import torch
    from sod2d_refmap_amd.ensemble import Ensemble, ensemble_members, sequential_solvers, sequential_step
    t0 = time.perf_counter()
    members = ensemble_members(count)
    E = Ensemble(members)
    sols = sequential_solvers(members)
    log(f"ensemble of {count}: {E.nelem} elements, {E.npoin_w} working nodes "
        f"(setup {time.perf_counter() - t0:.1f}s)")
    # no time-level copies: the per-instance baseline makes none either
    E.run(1, keep_levels=False)
    sequential_step(sols)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E.run(steps, keep_levels=False)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(steps):
        sequential_step(sols)
    torch.cuda.synchronize()
    ts = time.perf_counter() - t0
    import numpy as np
    err = 0.0
    for i, sol in enumerate(sols):
        a, b = E.member_state(i), sol.state()
        for k in ("rho", "q", "E", "mu_e"):
            err = max(err, float(np.linalg.norm(a[k] - b[k]) / max(np.linalg.norm(b[k]), 1e-300)))
    E.close()
    return {"members": count, "elements": E.nelem, "working_nodes": E.npoin_w, "steps": steps,
            "batched_ms_per_step": round(tb / steps * 1e3, 3), "sequential_ms_per_step": round(ts / steps * 1e3, 3),
            "speedup": round(ts / tb, 3), "batched_GDOF_s": round(5.0 * E.npoin_w * 4 * steps / tb / 1e9, 4),
            "max_rel_l2_vs_sequential": err}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64, help="elements per direction")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--curvilinear-n", type=int, default=96,
                    help="elements per direction of the randomised curvilinear workload (0: skip it)")
    ap.add_argument("--batch", type=int, default=48,
                    help="also time the seeded ensemble of this many boxes (0: skip)")
    ap.add_argument("--stage-config", default="",
                    help="pin the stage kernel configuration as ELEMS_PER_WORKGROUP,ASSEMBLY[,LDS] "
                         "(assembly pull|atomic|colour; 0,pull: persistent wavefronts; LDS swizzled|linear)")
    ap.add_argument("--autotune", type=int, default=0, help="stage configuration mode (sod_set_autotune)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="start with one 1 GiB device-to-device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from sod2d_refmap_amd import _lib, cases, mesh
    from sod2d_refmap_amd.solver import PeriodicSolver

    n = args.n
    from sod2d_refmap_amd.device import init as dev_init
    dev_init(0)
    if args.stage_config:
        nw, asm, *lay = args.stage_config.split(",")
        _lib.set_stage_config(int(nw), asm, lay[0] if lay else "swizzled")
    elif args.autotune:
        _lib.set_autotune(args.autotune)
    if args.copy_probe:
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
        log("copy probe: 1 GiB device-to-device")
    t0 = time.perf_counter()
    m = mesh.periodic_box(n)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas)
    log(f"mesh {n}^3: nelem={m.nelem} npoin_w={m.npoin_w} ({time.perf_counter() - t0:.1f}s)")
    sol = PeriodicSolver(m, gas, s, cfl_conv=0.5, cfl_diff=0.5)
    dt = sol.adapt_dt()
    log(f"setup done, dt0={dt:.4e}")
    for _ in range(args.warmup):
        sol.rk_4_main(sol.dt)
        sol.adapt_dt()
    torch.cuda.synchronize()
    _lib.call("sod_profile_enable", 1)
    _lib.call("sod_profile_reset")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sol.rk_4_main(sol.dt)
        sol.adapt_dt()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    _lib.call("sod_profile_enable", 0)
    K = args.steps
    stages = 4 * K
    gdofs = 5.0 * m.npoin_w * stages / el / 1e9
    ek = sol.volAvg_EK()
    log(f"{K} steps in {el:.3f}s, EK={ek:.6f}, dt={sol.dt:.4e}")

    # dominant kernel and its algorithmic bytes per launch
    kernels = {}
    for name in ("stage_rhs", "stage_update", "smart_visc_fused", "full_convec", "full_diffusion",
                 "scalar_convec", "rk4_step"):
        ms, cnt = kernel_ms(name)
        if cnt:
            kernels[name] = (ms, cnt)
    dom = max((k for k in kernels if k != "rk4_step"), key=lambda k: kernels[k][0])
    ms, cnt = kernels[dom]
    Ne, Nn = m.nelem, m.npoin_w
    # Algorithmic bytes per launch (FP64, int32 ids).  Per element, the
    # element-node data every element kernel streams: connec 64*4 + He 576*8
    # + gpvol 64*8 (+ mu_e 64*8 for the fused RHS and the diffusion).  Per
    # working node: the nodal inputs gathered once and the residuals written
    # once, 8 bytes each.
    per_elem = {"stage_rhs": 5888, "full_diffusion": 5888}.get(dom, 5376)
    per_node = {"stage_rhs": 5 + 1 + 6, "full_convec": 9 + 5, "full_diffusion": 6 + 5,
                "scalar_convec": 7 + 1, "stage_update": 0, "smart_visc_fused": 0}[dom]
    traffic_alg = Ne * per_elem + Nn * 8 * per_node
    avg_s = ms / cnt / 1e3
    achieved = traffic_alg / avg_s / 1e9
    # whole-stage algorithmic bytes (element data + node state, residual,
    # mass, RK accumulators and next-stage state): N_e*5888 + N_n*272
    stage_bytes = Ne * 5888 + Nn * 272
    stage_gbs = stage_bytes / (el / stages) / 1e9
    out = {
        "metric": "GDOF/s per RK4 stage (FP64, HEX64)",
        "value": round(gdofs, 4),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(el / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (compressible Taylor-Green vortex initial condition, Re=1600, Mach 0.1)",
        "config": {"workload": f"tgv_hex64_{n}^3_rk4_cfl0.5", "elements": Ne, "working_nodes": Nn,
                   "porder": 3, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_ms": round(ms / cnt, 4), "alg_bytes_per_launch": traffic_alg},
        "kernels_ms_per_step": {k: round(v[0] / K, 4) for k, v in kernels.items()},
        "stage": {"alg_bytes": stage_bytes, "ms": round(el / stages * 1e3, 4),
                  "alg_GBs": round(stage_gbs, 1), "frac_of_hbm_peak": round(stage_gbs / HBM_PEAK_GBS, 4)},
        "ek_final": ek,
    }
    # the stage-kernel configuration the timed steps ran with (sod_stage_config)
    out["stage_config"] = _lib.stage_config()
    traffic = measured_traffic(dom, n)
    if traffic is not None:
        out["roofline"]["traffic"] = traffic
    # the stage kernel is co-bound by FP64 issue: the same launch time against
    # the FP64 vector peak, with the operation count the SQ counters give
    fpw = measured_flops_per_wave(dom, n)
    if fpw is not None:
        tf = fpw * Ne / avg_s / 1e12
        out["roofline_fp64"] = {"bound": "fp64-valu", "achieved": round(tf, 2), "peak": FP64_PEAK_TFLOPS,
                                "unit": "TFLOP/s", "frac": round(tf / FP64_PEAK_TFLOPS, 4),
                                "flop_per_element": fpw}
    # the same steps through the device time loop (graph-captured step, dt on
    # the device, no time-level copies: the same work as the eager steps):
    # reported beside the eager figure
    try:
        sol.run_device(1, keep_levels=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol.run_device(K, keep_levels=False)
        torch.cuda.synchronize()
        tl = time.perf_counter() - t0
        out["time_loop"] = {"ms_per_step": round(tl / K * 1e3, 4),
                            "GDOF_s": round(5.0 * m.npoin_w * 4 * K / tl / 1e9, 4), "keep_levels": False}
    except Exception as exc:  # reported, not fatal: the eager figure stands
        out["time_loop"] = {"error": str(exc)[:200]}
    if not args.no_adjoint:
        out["adjoint"] = adjoint_ratio(sol, m)
    if args.batch > 0:
        out["batch"] = batch_speedup(args.batch, max(2, K // 2))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sol, m, gas, s, args.cpu_seconds)
    if args.curvilinear_n > 0:
        sol.close()
        out["curvilinear"] = curvilinear_rate(max(3, K // 4), args.curvilinear_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
