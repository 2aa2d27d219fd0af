#!/usr/bin/env python3
"""Throughput of the HEOM hot path on one MI355X: classical RK4 over the full ADO
hierarchy of the 8-site, 16-mode, depth-8 system (735,471 complex128 8x8 ADOs).
Beside it: the same workload with FP32 storage (and its population error against FP64
over 1000 fs), the 24-site system in FP64 and FP32 storage, the batched 512-realisation
static-disorder ensemble in FP64 and FP32 (ensemble ADO-steps/s and the FP32
population error), and the adjoint gradient of the 3003-ADO system (time per gradient,
a central-difference spot check, the CPU oracle's adjoint on a sample).

One "step" is one RK4 step of the whole hierarchy (four fused RHS+RK stage kernels).
The instance is built by the product's host path (config file -> C++ hierarchy graph
and prefactors -> heom_create); the state is resident in HBM before the timed region.
Prints one JSON line with ADO-RK4-steps/s, the roofline of the stage kernel and the
CPU oracle timed on the host cores beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME] [--no-cpu-baseline]
                  [--no-extras | --extras precision,ensemble,adjoint,filtering,field,tss]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dm-heom_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# per-launch memory-side traffic of the stage kernels, measured with rocprofv3 counters
# (profiles/*.md give the commands), one record per workload and storage precision
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "rk4_stage_traffic.json")


def parse_args():
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--config", default="excitonic8_d8")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the FP32-storage run of the same workload and the batched ensemble")
    p.add_argument("--extras", default="precision,ensemble,adjoint,filtering,field,tss",
                   help="comma-separated extra sections to run: precision (FP32 storage, 24 sites), ensemble, "
                        "adjoint, filtering (compaction), field (time-dependent H), tss (thermal state search), "
                        "norms (hierarchy norms; not in the default set)")
    p.add_argument("--ensemble-members", type=int, default=512)
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    p.add_argument("--precision", choices=("fp64", "fp32"), default="fp64",
                   help="state storage of the timed run: the plain FP64 path (the headline figure) or FP32 "
                        "storage with FP64 tier-0/1 ADOs")
    p.add_argument("--xcd-chunk", type=int, default=0,
                   help="after the autotuner, pin this many tiles per XCD chunk in the configuration it kept "
                        "(for profiler runs, whose timings can make the autotuner keep another chunk)")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="after the timed region, copy a buffer of this many MiB once (a plain copy kernel of "
                        "known size, so that profiler byte counters can be calibrated in the same run)")
    return p.parse_args()


def copy_ceiling(gib: float = 1.0, reps: int = 5) -> float:
    """A plain device copy of `gib` GiB (read + write), the best of `reps`, in GB/s moved:
    the HBM rate a streaming kernel can reach on this GPU."""
    import torch

    n = int(gib * (1 << 30)) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    del src, dst
    torch.cuda.empty_cache()
    return 2 * n * 8 / best / 1e9


def host_cpu() -> dict:
    """The host the CPU baseline ran on: lscpu's model name, vendor, family/model/stepping
    and socket/core/thread counts, nproc, and OMP_NUM_THREADS as set in the environment."""
    import subprocess

    info = {"nproc": os.cpu_count(), "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Vendor ID": "vendor", "CPU family": "family", "Model": "model_number",
                "Stepping": "stepping", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus", "CPU max MHz": "max_mhz"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if "model" not in info:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        info["model"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return info


def cpu_baseline(cfg_path: str, target_s: float) -> dict:
    """The FP64 CPU oracle (statement-by-statement restatement of the reference kernels,
    OpenMP over ADOs) on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O

    O.build()
    cfg = O.parse_config(cfg_path)
    inst = O.instance_from_config(cfg)
    y = inst.state_with_top(cfg.rho_init)
    work = np.empty(5 * inst.state_size)
    t0 = time.perf_counter()
    inst.rk4_step(y, cfg.step_size, work)  # warm-up step
    t_one = time.perf_counter() - t0
    n = int(max(1, min(20, round(target_s / max(t_one, 1e-3)))))
    t0 = time.perf_counter()
    for _ in range(n):
        inst.rk4_step(y, cfg.step_size, work)
    dt = time.perf_counter() - t0
    return {
        "value": inst.matrices * n / dt,
        "unit": "ADO-RK4-steps/s",
        "cores": int(O.lib().oracle_num_threads()),
        "kind": "port",
        "sample": f"{n} RK4 steps of the same {inst.matrices}-ADO hierarchy after 1 warm-up step "
                  f"(unfused: 4 RHS + 4 weighted adds per step), {dt:.1f} s",
        "host": host_cpu(),
    }


def precision_run(host, steps_1000fs: int) -> dict:
    """The same workload with complex64 storage and FP32 RHS (FP64 tier-0/1 ADOs):
    throughput, and the population error against FP64 over 1000 fs."""
    import heom_amd

    info = host.info
    res = {}
    pops = {}
    for prec, name in ((heom_amd.HEOM_PRECISION_FP64, "fp64"), (heom_amd.HEOM_PRECISION_FP32, "fp32")):
        h = host.create_handle(precision=prec)
        p, _ = h.propagate_observe(info.step_size, steps_1000fs, 10)
        t = h.timing().last_call_seconds
        pops[name] = p[:, 0]
        res[name] = {"ms_per_step": t / steps_1000fs * 1e3, "value": host.matrices * steps_1000fs / t}
        h.close()
    res["max_population_error_vs_fp64"] = float(np.max(np.abs(pops["fp32"] - pops["fp64"])))
    res["span_fs"] = steps_1000fs * info.step_size * 1e15
    res["unit"] = "ADO-RK4-steps/s"
    return res


def wide_precision_run() -> dict:
    """The 24-site system (24 baths, depth 4, 20,475 ADOs of 24 x 24) in FP64 and with
    FP32 storage over 1000 fs: rates and the FP32 population error."""
    import heom_amd

    host = heom_amd.HostInstance(os.path.join(heom_amd.CONFIG_DIR, "excitonic24_d4.cfg"))
    info = host.info
    steps = int(round(1e-12 / info.step_size))
    res, pops = {}, {}
    for prec, name in ((heom_amd.HEOM_PRECISION_FP64, "fp64"), (heom_amd.HEOM_PRECISION_FP32, "fp32")):
        h = host.create_handle(precision=prec)
        h.step_rk4(info.step_size, 2)  # warm-up
        h.close()
        h = host.create_handle(precision=prec)
        kept = next(r for r in h.autotune(steps=3) if r["chosen"])
        tops = h.propagate(info.step_size, steps, 10)
        t = h.timing().rk4_seconds
        pops[name] = np.real(np.diagonal(tops, axis1=1, axis2=2))
        res[name] = {"ms_per_step": t / steps * 1e3, "value": host.matrices * steps / t,
                     "ados_per_workgroup": kept["ados_per_workgroup"]}
        h.close()
    res["max_population_error_vs_fp64"] = float(np.max(np.abs(pops["fp32"] - pops["fp64"])))
    res["workload"] = f"24 sites, 24 modes, depth {info.depth}, {host.matrices} ADOs, {steps} RK4 steps of 1 fs"
    res["unit"] = "ADO-RK4-steps/s"
    host.close()
    return res


def adjoint_run(cpu: bool) -> dict:
    """Adjoint gradient of J = integral over 1 ps of the population of site 3 (8 sites,
    8 modes, depth 6, 3003 ADOs, 1000 RK4 steps of 1 fs, checkpoints every 50 steps)
    with respect to 8 site energies, 28 couplings and 8 (lambda, gamma) pairs: time of
    one gradient, a spot check against central differences (h = 1e-3 cm^-1) and the
    CPU oracle's adjoint on a sample of the same workload."""
    import heom_amd

    cfg_path = os.path.join(heom_amd.CONFIG_DIR, "excitonic8_d6.cfg")
    host = heom_amd.HostInstance(cfg_path)
    info = host.info
    N, nsteps, dt, site = host.states, int(info.steps), float(info.step_size), 3
    op = np.zeros((N, N))
    op[site, site] = 1.0
    w = np.full(nsteps + 1, dt * 1e15)
    w[0] = w[-1] = 0.5 * dt * 1e15

    def gradient():
        h = host.create_handle()
        h.use_graphs(True)
        t0 = time.perf_counter()
        X = h.adjoint_forward(dt, nsteps, 1, op, 50)[:, 0]
        t_fwd = h.timing().last_call_seconds
        g = h.adjoint_backward(w, state=False)
        t_bwd = h.timing().last_call_seconds
        pg = host.parameter_gradient(g)
        wall = time.perf_counter() - t0
        h.close()
        return float(np.sum(w * X.real)), pg, t_fwd, t_bwd, wall

    gradient()
    J, pg, t_fwd, t_bwd, wall = gradient()
    p0 = host.get_parameters()

    def objective(H, lam, gam):
        host.set_parameters(H, lam, gam)
        h = host.create_handle()
        _, tr = h.propagate_observe(dt, nsteps, 1, op)
        h.close()
        return float(np.sum(w * np.real(tr[:, 0])))

    hstep = 1e-3
    checks = {}
    for name in ("eps_3", "J_2_3", "lambda_3", "gamma_3"):
        vals = []
        for sgn in (1.0, -1.0):
            H, lam, gam = p0["hamiltonian"].copy(), p0["lambda"].copy(), p0["gamma"].copy()
            if name == "eps_3":
                H[3, 3] += sgn * hstep
            elif name == "J_2_3":
                H[2, 3] += sgn * hstep
                H[3, 2] += sgn * hstep
            elif name == "lambda_3":
                lam[3] += sgn * hstep
            else:
                gam[3] += sgn * hstep
            vals.append(objective(H, lam, gam))
        fd = (vals[0] - vals[1]) / (2 * hstep)
        adj = {"eps_3": pg["hamiltonian"][3, 3].real, "J_2_3": pg["hamiltonian"][2, 3].real + pg["hamiltonian"][3, 2].real,
               "lambda_3": pg["lambda"][3], "gamma_3": pg["gamma"][3]}[name]
        checks[name] = {"adjoint": float(adj), "central_difference": fd, "rel": abs(adj - fd) / abs(fd)}
    host.set_parameters(p0["hamiltonian"], p0["lambda"], p0["gamma"])
    host.close()
    out = {
        "workload": "8 sites, 8 modes, depth 6 (3003 ADOs), 1000 RK4 steps of 1 fs, checkpoints every 50 steps; "
                    "J = trapezoidal integral of the site-3 population over 1 ps (fs); gradient w.r.t. 8 eps, "
                    "28 J_ij, 8 lambda, 8 gamma",
        "J_fs": J,
        "parameters": 52,
        "gradient_seconds": wall,
        "forward_device_seconds": t_fwd,
        "backward_device_seconds": t_bwd,
        "spot_check_h_cm-1": hstep,
        "max_rel_vs_central_differences": max(c["rel"] for c in checks.values()),
        "checks": checks,
    }
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        O.build()
        cfg = O.parse_config(cfg_path)
        inst = O.instance_from_config(cfg)
        y0 = inst.state_with_top(cfg.rho_init)
        sample = 100
        ws = np.full(sample + 1, dt * 1e15)
        t0 = time.perf_counter()
        inst.adjoint_gradient(y0, dt, sample, 1, op, ws, checkpoint_steps=50)
        t_cpu = time.perf_counter() - t0
        out["cpu_oracle"] = {"seconds_per_gradient_estimate": t_cpu * nsteps / sample,
                             "sample": f"adjoint gradient over the first {sample} of the {nsteps} steps, "
                                       f"{t_cpu:.1f} s, scaled by {nsteps // sample}",
                             "cores": int(O.lib().oracle_num_threads())}
        out["speedup_vs_cpu_oracle"] = out["cpu_oracle"]["seconds_per_gradient_estimate"] / wall
    return out


def filtering_run(host, quantile: float = 0.5, settle_steps: int = 20, steps: int = 20) -> dict:
    """Filtering with compaction on the same hierarchy: after `settle_steps` RK4 steps,
    the ADOs whose squared norm lies below the `quantile` of all norms are masked
    (heom_hierarchy_norm with a threshold) and removed from the device layout
    (heom_compact).  Reports the step time before and after, and the deviation of the
    compacted run's tops from the masked, uncompacted run."""
    info = host.info
    dt = info.step_size
    handles = [host.create_handle(), host.create_handle()]
    for h in handles:
        h.step_rk4(dt, settle_steps)
    norms = handles[0].norms()
    threshold = float(np.quantile(norms, quantile))
    for h in handles:
        h.norms(threshold)
    t0 = handles[1].timing().rk4_seconds
    handles[1].step_rk4(dt, steps)
    handles[1].sync()
    before = (handles[1].timing().rk4_seconds - t0) / steps
    kept = handles[0].compact()
    handles[0].autotune(steps=2)
    t0 = handles[0].timing().rk4_seconds
    handles[0].step_rk4(dt, steps)
    handles[0].sync()
    after = (handles[0].timing().rk4_seconds - t0) / steps
    # tops of both over the same further steps
    a = handles[0].propagate(dt, steps, steps)
    b = handles[1].propagate(dt, steps, steps)
    dev = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    for h in handles:
        h.close()
    return {
        "workload": f"{host.matrices}-ADO hierarchy after {settle_steps} steps; ADOs with squared norm below the "
                    f"{quantile:.0%} quantile masked and compacted",
        "threshold": threshold,
        "ados_kept": kept,
        "ms_per_step_masked": before * 1e3,
        "ms_per_step_compacted": after * 1e3,
        "speedup": before / after,
        "max_rel_top_deviation_compacted_vs_masked": dev,
    }


def field_run(host, steps: int = 40) -> dict:
    """A time-dependent Hamiltonian (pump/probe field, heom_propagate_timedep) on the same
    hierarchy: step time against the constant-H step time."""
    import heom_amd
    from heom_amd.field import pump_probe, stage_hamiltonians

    info = host.info
    dt = info.step_size
    N = host.states
    rng = np.random.default_rng(7)
    dm = rng.normal(size=(N, N)) * 1e-21
    H = host.arrays()["hamiltonian"].reshape(N, N, 2)
    H0 = H[..., 0] + 1j * H[..., 1]
    field = pump_probe(probe=dict(efield=2.0, center=20e-15, width=10e-15, phase=0.3, frequency=2.3e15),
                       pump=dict(efield=5.0, center=10e-15, width=6e-15, phase=-0.5, frequency=2.25e15))
    sh = stage_hamiltonians(H0, dm, dm.T, field, 0.0, dt, steps)
    h = host.create_handle()
    h.propagate_timedep(dt, 2, 2, sh[:2])
    t0 = h.timing().rk4_seconds
    h.propagate_timedep(dt, steps, steps, sh)
    t_field = (h.timing().rk4_seconds - t0) / steps
    t0 = h.timing().rk4_seconds
    h.propagate(dt, steps, steps)
    t_plain = (h.timing().rk4_seconds - t0) / steps
    h.close()
    return {"workload": f"pump/probe field on the {host.matrices}-ADO hierarchy, {steps} RK4 steps",
            "ms_per_step_field": t_field * 1e3, "ms_per_step_constant_h": t_plain * 1e3}


def norms_run(calls: int = 12) -> dict:
    """heom_hierarchy_norm (the norms of every ADO, read back to the host) on the 8-site
    depth-8 and the 24-site depth-4 hierarchies in FP64 and FP32 storage, after two RK4
    steps: milliseconds per call (kernel, 8 bytes per ADO to the host and the sync) and
    the state bytes each call reads."""
    import heom_amd

    out = {}
    for cfg in ("excitonic8_d8", "excitonic24_d4"):
        host = heom_amd.HostInstance(os.path.join(heom_amd.CONFIG_DIR, f"{cfg}.cfg"))
        for prec, name in ((heom_amd.HEOM_PRECISION_FP64, "fp64"), (heom_amd.HEOM_PRECISION_FP32, "fp32")):
            h = host.create_handle(precision=prec)
            h.step_rk4(host.info.step_size, 2)
            h.norms()
            t0 = time.perf_counter()
            for _ in range(calls):
                h.norms()
            dt = (time.perf_counter() - t0) / calls
            out[f"{cfg}/{name}"] = {"ms_per_call": dt * 1e3, "ados": host.matrices,
                                    "state_bytes": host.matrices * host.states ** 2 * (8 if name == "fp32" else 16)}
            h.close()
        host.close()
    return out


def tss_run(host, iterations: int = 20) -> dict:
    """Thermal state search iterations (heom_tss_step: one f_TSS evaluation by the stage
    kernel plus the per-ADO eigenbasis transform) on the same hierarchy, timed on the
    host around a synchronised run of `iterations` iterations."""
    import time

    h = host.create_tss_handle()
    h.tss_step(2)
    h.sync()
    t0 = time.perf_counter()
    h.tss_step(iterations)
    h.sync()
    per = (time.perf_counter() - t0) / iterations
    top = h.get_top()
    h.close()
    nel = host.states * host.states
    transform_bytes = 2.0 * host.matrices * nel * 16
    return {"workload": f"thermal state search on the {host.matrices}-ADO hierarchy, {iterations} iterations",
            "ms_per_iteration": per * 1e3, "ado_iterations_per_s": host.matrices / per,
            "transform_alg_bytes": transform_bytes,
            "top_trace_minus_1": float(abs(np.trace(top) - 1.0))}


def ensemble_run(members: int) -> dict:
    """The 512-realisation static-disorder ensemble (7 sites, depths 3-6, populations
    plus xx/yy/zz linear absorption, 2048 steps of 2 fs), batched, FP64 and FP32."""
    import heom_amd
    from heom_amd.ensemble import Ensemble

    out = {}
    results = {}
    for prec, name in ((heom_amd.HEOM_PRECISION_FP64, "fp64"), (heom_amd.HEOM_PRECISION_FP32, "fp32")):
        ens = Ensemble(members=members, precision=prec)
        r = ens.run()
        results[name] = r
        out[name] = {"value": r["ado_steps"] / r["device_seconds"], "device_seconds": r["device_seconds"],
                     "ms_per_step": r["device_seconds"] / r["steps"] * 1e3}
        out["hierarchies"] = ens.members * (1 + ens.ncomp)
        out[name]["tuning"] = ens.tuning
        out["ados"] = r["ados"]
        ens.close()
    out["max_population_error_fp32_vs_fp64"] = float(
        np.max(np.abs(results["fp32"]["populations"] - results["fp64"]["populations"])))
    a, b = results["fp64"]["absorption"], results["fp32"]["absorption"]
    out["max_absorption_error_fp32_vs_fp64"] = float(np.max(np.abs(a - b)) / np.max(np.abs(a)))
    out["unit"] = "ensemble ADO-RK4-steps/s"
    out["workload"] = (f"{members} static-disorder realisations of the 7-site system, depths 3-6, populations + "
                       f"linear absorption (3 tensor components, rotating frame), 2048 RK4 steps of 2 fs")
    return out


def pcie_inclusive(h, dt, steps, M, wall_steps) -> dict:
    """The rate with the whole hierarchy handed over at the C ABI: the state uploaded from
    (pageable) host memory with heom_set_state before the steps and downloaded with
    heom_get_state after them, both timed; `value` excludes these copies."""
    import time

    y = h.get_state()
    h.sync()
    t0 = time.perf_counter()
    h.set_state(y)
    h.sync()
    up = time.perf_counter() - t0
    t0 = time.perf_counter()
    y = h.get_state()
    down = time.perf_counter() - t0
    gb = y.nbytes / 1e9
    return {"value": M * steps / (wall_steps + up + down), "unit": "ADO-RK4-steps/s",
            "upload_s": up, "download_s": down, "state_GB": gb,
            "note": f"{steps} steps with the {gb:.2f} GB state uploaded before and downloaded after them"}


def stage_kernel_name(h, fp32: bool) -> str:
    """The stage kernel a handle runs, by its template name."""
    kind = h.stage_kernel()
    if kind == "pipelined":
        return "rk4_stage_pair" if fp32 else "rk4_stage_pipe"
    return "rk4_stage_wide" if kind == "wide" else "rk4_stage"


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the hierarchy would shard by ADO partitions with halo exchange, not built here)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import heom_amd

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    cfg_path = os.path.join(heom_amd.CONFIG_DIR, f"{args.config}.cfg")
    host = heom_amd.HostInstance(cfg_path, task=heom_amd.TASK_POPULATION_DYNAMICS)
    info = host.info
    fp32 = args.precision == "fp32"
    h = host.create_handle(precision=heom_amd.HEOM_PRECISION_FP32 if fp32 else heom_amd.HEOM_PRECISION_FP64)
    M, N = host.matrices, host.states
    dt = info.step_size
    tuning = h.autotune(steps=3)  # the stage-kernel configuration, timed on this hierarchy
    kept = next(r for r in tuning if r["chosen"])
    if args.xcd_chunk > 0:
        kept = {**kept, "xcd_chunk": args.xcd_chunk, "pinned": True}
        h.set_stage_config(kept)

    h.step_rk4(dt, args.warmup)
    h.sync()
    torch.cuda.synchronize()
    t_ev0 = h.timing().rk4_seconds
    t0 = time.perf_counter()
    h.step_rk4(dt, args.steps)
    h.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev = h.timing().rk4_seconds - t_ev0  # HIP events on the library's stream
    if args.copy_mib > 0:
        n = args.copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)  # elementwise kernel: reads and writes copy_mib MiB
        torch.cuda.synchronize()
        del src, dst
    top = h.get_top()
    trace_err = abs(complex(top.trace()) - 1.0)
    copy_gbs = copy_ceiling()
    pcie = pcie_inclusive(h, dt, args.steps, M, wall)

    value = M * args.steps / wall
    bytes_per_ado_step = 16 * N * N * (8 if fp32 else 16)  # 16 state passes of N^2 complex (algorithmic)
    launches = 4 * args.steps
    per_launch_s = ev / launches
    alg_bytes_per_launch = bytes_per_ado_step * M / 4
    achieved = alg_bytes_per_launch / per_launch_s / 1e9
    traffic, traffic_rec = None, None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            traffic_rec = json.load(f).get(f"{args.config}/{args.precision}")
        # the traffic of a stage launch depends on the tiles per XCD chunk the tuner kept:
        # the record holds the chunks the tuner keeps on this hierarchy
        chunk_rec = (traffic_rec or {}).get("by_chunk", {}).get(str(kept["xcd_chunk"]))
        if chunk_rec:
            traffic = chunk_rec["l2_fabric_bytes_per_launch"]

    out = {
        "metric": "ADO\u00b7RK4-steps/s + % HBM peak; batched FP32 vs plain FP64 at \u22641e-6 pop err",
        "value": value,
        "unit": "ADO-RK4-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 tier-0/1 ADOs and reductions)" if fp32 else "f64",
        "data": "synthetic (seeded config file, no external data)",
        "config": {
            "workload": f"{args.config}: {info.sites} sites, {info.width} bath modes, depth {info.depth}, "
                        f"{M} {'complex64' if fp32 else 'complex128'} {N}x{N} ADOs, classical RK4 dt={dt:g} s",
            "ados": M,
            "states": N,
            "modes": info.width,
            "depth": info.depth,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_note": None if traffic is None else (
                "bytes per launch between the L2 and the memory fabric (rocprofv3: reads as "
                f"{traffic_rec.get('read_counter', 'FETCH_SIZE x 2 KiB')}, writes as WRITE_SIZE; every L2 read "
                "request is 128 B on gfx950, so TCC_EA0_RDREQ x 128 B and FETCH_SIZE x 2 KiB count the same bytes); it "
                "includes hits in the 256 MB last-level cache (MALL), which no L2-side counter separates from "
                f"DRAM (TCC_EA0_RDREQ_DRAM equals TCC_EA0_RDREQ); measured on build {traffic_rec.get('build')} with "
                f"{kept['xcd_chunk']} tiles per XCD chunk, {traffic_rec.get('source')}"),
            "kernel": f"{stage_kernel_name(h, fp32)} (RHS + RK4 stage epilogue), 4 launches per step",
            "alg_bytes_per_launch": alg_bytes_per_launch,
            "avg_launch_ms": per_launch_s * 1e3,
            "measured_copy_GBs": copy_gbs,
            "frac_of_measured_copy": achieved / copy_gbs,
        },
        "checks": {"top_trace_minus_1": trace_err},
        "pcie_inclusive": pcie,
        "tuning": {"kept": kept, "candidates": len(tuning)},
    }
    h.close()
    extras = set() if args.no_extras else {x.strip() for x in args.extras.split(",") if x.strip()}
    if "precision" in extras:
        out["precision"] = precision_run(host, int(round(1e-12 / dt)))
        out["precision"]["sites24"] = wide_precision_run()
    if "filtering" in extras:
        out["filtering"] = filtering_run(host)
    if "field" in extras:
        out["field"] = field_run(host)
    if "tss" in extras:
        out["tss"] = tss_run(host)
    if "norms" in extras:
        out["norms"] = norms_run()
    if "ensemble" in extras:
        out["ensemble"] = ensemble_run(args.ensemble_members)
        out["ensemble"]["fp32_ensemble_vs_plain_fp64"] = out["ensemble"]["fp32"]["value"] / value
    if "adjoint" in extras:
        out["adjoint"] = adjoint_run(not args.no_cpu_baseline)
    host.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg_path, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
