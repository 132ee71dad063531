"""Benchmark of the implicit Monte Carlo transport hot path on one MI355X.

Workload (the line-expansion configuration): 1D homologous ejecta, 128
zones to v_max = 2e9 cm/s at t = 20 d, rho ~ exp(-v/3e8) normalised to
1.4 Msun with 56Ni inside 0.6 Msun; 1e5 synthetic fuzz lines, LTE
Saha-Boltzmann for 4 ion stages per element; 4096 log-frequency bins; 1e7
radioactive packets per step, steady state.  A step is one pass of the hot
path with inputs resident in HBM: opacity / emissivity build, emission,
propagation to escape, tally normalisation, radiative-equilibrium gas
temperature.

Reported: packets/s of the batched mixed-precision form (FP32 propagation,
FP64 tallies), beside the plain FP64 drop-in at equal packet count, the NLTE
zone-solves/s of the batched FP32+refinement solve versus the per-system
FP64 drop-in, the roofline of the propagation kernel and the CPU oracle timed
on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--packets N]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the mixed-form propagation kernel the default launch runs (4-wave register
# bound; the last template argument is the grid kind: 1 sphere, 2 cylinder)
PROP_KERNEL = "propagate_kernel<float, true, true, 4, %d, false>"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=int, default=None,
                    help="packets per step (default: the configuration's own count)")
    ap.add_argument("--config", default="c2_line")
    ap.add_argument("--plain-packets", type=int, default=1_000_000,
                    help="packets of the plain FP64 comparison step (its rate is linear in the "
                         "packet count from about 4e5 packets up; 0: same as --packets)")
    ap.add_argument("--traffic-from", default=os.path.join(ROOT, "profiles", "propagate_mixed_c2_line.md"),
                    help="rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the propagation kernel")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (the HBM counter calibration) and exit")
    ap.add_argument("--nlte-only", action="store_true",
                    help="only the batched NLTE measurement (for profiling its kernel)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nlte", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    return ap.parse_args()


S_STATE = 96   # packet record in HBM: 9 FP64 + 3 int32 + uint32 draw + uint64 id


def packet_bytes(c, nn, precision):
    """Algorithmic HBM bytes per packet of the propagation kernel,
    B_pkt = 2 S_state + n_step (2 w_kappa + 8) + n_redist 8 ceil(log2 n_nu) + f_esc 16:
    the packet record read and written once, per event the opacity gather
    (w_kappa = 8 B per FP64 entry, 4 B per FP32 entry) and an 8-byte J_nu
    accumulate, per redistribution the binary search of the FP64 emissivity
    CDF, 16 B of spectrum per escape.  (The FP32 form samples the CDF through
    a guide table, one 4-byte entry and one or two 4-byte CDF probes instead
    of 12 probes: fewer bytes than this figure counts.)"""
    n = max(c["n_packets"], 1)
    n_step = c["n_events"] / n
    n_redist = c["n_redist"] / n
    f_esc = c["n_escaped"] / n
    wk = 8 if precision == 0 else 4
    return 2 * S_STATE + n_step * (2 * wk + 8) + n_redist * 8 * math.ceil(math.log2(nn)) + f_esc * 16


def read_traffic(path, n_packets):
    """HBM bytes per launch of the propagation kernel from the committed
    rocprofv3 summary: its `traffic_json:` line holds FETCH_SIZE and
    WRITE_SIZE per packet, already divided by the copy-kernel calibration
    factors.  Scaled to this run's packet count; None without a summary."""
    if not path or not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_json:"):
                t = json.loads(line.split(":", 1)[1])
                per = t["fetch_bytes_per_packet"] + t["write_bytes_per_packet"]
                return per * n_packets, t
    return None, None


def copy_calibration():
    """One 1 GiB device-to-device copy, three times: FETCH_SIZE and
    WRITE_SIZE of this kernel against its 2^30 bytes give the counters'
    scale factors on this GPU."""
    import torch
    n = (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": 1 << 30, "launches": 3}))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no "
              "multi-GPU path (--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    if a.nlte_only:
        import numpy as np
        import torch
        from pubsed_amd import abi, workloads as W
        from pubsed_amd.runtime import GpuBackend
        gpu = GpuBackend(0)
        gpu.use_torch_stream()
        gpu.enable_timing(True)
        print(json.dumps(bench_nlte(gpu, torch, W, abi, np)))
        return
    import numpy as np
    import torch
    from pubsed_amd import runtime, abi, workloads as W
    from pubsed_amd.runtime import GpuBackend

    dev = "cuda:0"
    gpu = GpuBackend(0)
    gpu.use_torch_stream()
    gpu.enable_timing(True)
    N = a.packets or runtime.CONFIGS[a.config]["n_packets"]
    if not runtime.CONFIGS[a.config]["steady"]:
        print(json.dumps(bench_time_dependent(a, gpu)))
        return
    P = runtime.Problem(runtime.Space(dev), a.config)
    P.alloc_packets(N)
    cnt = torch.zeros(10, dtype=torch.int64, device=dev)

    # every step starts from the same gas temperatures: a step's radiative-
    # equilibrium update is computed (and timed) but not carried into the
    # next step, so that all timed steps (and both forms) run one workload
    T0 = P.a["T"].clone()
    radeq = P.cfg["radiative_eq"]

    def step(precision, n, counters=None):
        P.tparams.precision = precision
        P.a["T"].copy_(T0)
        gpu.wipe(P)
        gpu.opacity_build(P)
        gpu.emit(P, n)
        gpu.propagate(P, n, counters)
        gpu.normalize(P)
        if radeq:
            gpu.solve_eq_temperature(P)   # transport.cpp:126-131

    def timed(precision, n, steps, warm):
        for _ in range(warm):
            step(precision, n)
        torch.cuda.synchronize()
        cnt.zero_()
        prop_ms, op_ms = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            step(precision, n, cnt)
            prop_ms.append(gpu.last_ms("propagate"))
            op_ms.append(gpu.last_ms("opacity"))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        c = dict(zip(abi.COUNTER_FIELDS, cnt.cpu().tolist()))
        return dt, prop_ms, c, op_ms

    print("# bench: config %s, %d packets/step, mixed precision" % (a.config, N), file=sys.stderr, flush=True)
    dt, prop_ms, c, op_ms = timed(abi.PREC_MIXED, N, a.steps, a.warmup)
    value = a.steps * N / dt
    ms_step = 1e3 * dt / a.steps
    prop_avg = float(np.mean(prop_ms))
    bpk = packet_bytes(c, P.n_nu, 1)
    achieved = bpk * N / (prop_avg * 1e-3) / 1e9
    per_pkt = {k: c[k] / max(c["n_packets"], 1) for k in ("n_events", "n_binshift", "n_redist",
                                                             "n_escaped", "n_scatter")}
    print("# mixed: %.3f s/step, propagate %.1f ms, events/packet %.1f" %
          (dt / a.steps, prop_avg, per_pkt["n_events"]), file=sys.stderr, flush=True)

    traffic, tinfo = read_traffic(a.traffic_from, N)
    traffic_src = None
    if tinfo:
        traffic_src = "%s (rocprofv3 --pmc at %d packets, per-packet bytes scaled to %d)" % (
            os.path.relpath(a.traffic_from, ROOT), tinfo["packets"], N)
    op_avg = float(np.mean(op_ms))
    n_lines = P.cfg.get("n_lines", 0) if P.cfg.get("fuzz") else 0
    # K1-K3 algorithmic bytes per zone: abs, scat and emissivity CDF written
    # (3 nu 8 B), the B_nu / means pass over them (nu 8 B), and the zone's
    # full line list streamed once (32 B per line: nu, gf, E_l, ion + bin)
    b_zone = 3 * P.n_nu * 8 + P.n_nu * 8 + 32 * n_lines
    opacity = {"kernel_ms": op_avg, "zones": P.nz, "lines": n_lines,
               "zone_builds_per_s": P.nz / (op_avg * 1e-3) if op_avg > 0 else None,
               "line_exp_per_s": 3.0 * n_lines * P.nz / (op_avg * 1e-3) if op_avg > 0 else None,
               "bytes_per_zone": b_zone,
               "achieved_GBps": b_zone * P.nz / (op_avg * 1e-3) / 1e9 if op_avg > 0 else None}

    # plain FP64 drop-in (its rate is linear in the packet count at these sizes)
    Np = min(a.plain_packets or N, N)
    print("# plain FP64: %d packets" % Np, file=sys.stderr, flush=True)
    dtp, prop_p, cp, _ = timed(abi.PREC_FP64, Np, 1, 0)
    plain = Np / dtp
    esc_m = c["n_escaped"] / max(c["n_packets"], 1)
    esc_p = cp["n_escaped"] / max(cp["n_packets"], 1)

    nlte = None
    if not a.no_nlte:
        nlte = bench_nlte(gpu, torch, W, abi, np)
        nlte["opacity_build_nlte"] = bench_nlte_state(gpu, P, torch, abi, np)

    cpu = None
    if not a.no_cpu:
        cpu = bench_cpu(a.config, a.cpu_seconds)

    out = {
        "metric": "packets/s & NLTE solves/s: batched/mixed vs plain GPU path at matched MC noise",
        "value": value,
        "unit": "packets/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 geometry / f64 tallies",
        "data": "synthetic (seeded ejecta, line list and rates)",
        "config": {"workload": "line-expansion 1D ejecta: 128 zones, 1e5 fuzz lines, 4 ion stages, "
                               "4096 log bins, steady state",
                   "packets_per_step": N, "zones": P.nz, "nu_bins": P.n_nu},
        "plain_fp64_packets_per_s": plain,
        "mixed_over_plain": value / plain,
        "escape_fraction": {"mixed": esc_m, "plain": esc_p},
        "per_packet": per_pkt,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": PROP_KERNEL % 1, "kernel_ms": prop_avg,
                     "bytes_per_packet": bpk, "traffic_source": traffic_src},
        "opacity_build": opacity,
        "nlte": nlte,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def bench_nlte_state(gpu, P, torch, abi, np):
    """The opacity build with every atom in NLTE (sgpu_opacity_build_nlte) on
    the mixed step's normalised J_nu: its time, the n_e evaluation passes,
    and (zone, atom) NLTE solves per second (systems x passes / time), with
    the FP32 + refinement and the FP64 batched LU."""
    J = P.a["J_nu"].clone()
    flags = [1] * P.atoms.n_atoms
    nsys = P.nz * P.atoms.n_atoms
    P.op_params.first_step = 0
    out = {"zones": P.nz, "systems_per_pass": nsys}
    for name, prec in (("fp32_refine", abi.NLTE_FP32_REFINE), ("fp64", abi.NLTE_FP64)):
        gpu.opacity_build_nlte(P, J, flags, lu_precision=prec)     # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, ev = gpu.opacity_build_nlte(P, J, flags, lu_precision=prec)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[name] = {"ms": 1e3 * el, "passes": ev, "solves_per_s": nsys * ev / el,
                     "n_elec_mean": float(P.get("n_elec").mean())}
    return out


def bench_time_dependent(a, gpu):
    """Time-dependent configuration (the 2D ellipsoidal outflow): every step
    is the homologous hydro step, opacity build, wipe, emission of N new
    packets after the census, propagation to t + dt, compaction and
    normalisation (runtime.time_step).  Census packets persist, so the timed
    step (after the warm-up steps) propagates the census plus N.  `value`
    is packets propagated per second of the whole step."""
    import numpy as np
    import torch
    from pubsed_amd import runtime, abi, workloads as W
    cfg = runtime.CONFIGS[a.config]
    dt = cfg.get("dt", 0.1 * W.DAY)
    cnt = torch.zeros(10, dtype=torch.int64, device="cuda:0")

    def run(precision, n_emit, steps, warm):
        P = runtime.Problem(runtime.Space("cuda:0"), a.config)
        cap = n_emit * (steps + warm)
        P.alloc_packets(cap)
        scr = P.alloc_packets(cap, name="scr")
        P.tparams.precision = precision
        for k in range(warm):
            runtime.time_step(gpu, P, dt, n_emit, scr, k)
            print("# warm-up step %d: %d packets on the grid" % (k, P.n_packets), file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        cnt.zero_()
        n_prop, prop_ms, op_ms = 0, [], []
        t0 = time.perf_counter()
        for k in range(warm, warm + steps):
            n_prop += P.n_packets + n_emit
            runtime.time_step(gpu, P, dt, n_emit, scr, k, cnt)
            prop_ms.append(gpu.last_ms("propagate"))
            op_ms.append(gpu.last_ms("opacity"))
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        c = dict(zip(abi.COUNTER_FIELDS, cnt.cpu().tolist()))
        return P, el, n_prop, prop_ms, op_ms, c

    N = a.packets or cfg["n_packets"]
    print("# bench: config %s, %d new packets per step, dt %.3g s, mixed precision" % (a.config, N, dt),
          file=sys.stderr, flush=True)
    P, el, n_prop, prop_ms, op_ms, c = run(abi.PREC_MIXED, N, a.steps, a.warmup)
    value = n_prop / el
    prop_avg = float(np.mean(prop_ms))
    bpk = packet_bytes(c, P.n_nu, 1)
    achieved = bpk * (n_prop / a.steps) / (prop_avg * 1e-3) / 1e9
    print("# mixed: %.3f s/step, propagate %.1f ms, %d packets propagated" % (el / a.steps, prop_avg, n_prop),
          file=sys.stderr, flush=True)
    Np = min(a.plain_packets or N, N)
    print("# plain FP64: %d new packets per step" % Np, file=sys.stderr, flush=True)
    _, elp, n_prop_p, _, _, cp = run(abi.PREC_FP64, Np, 1, a.warmup)
    plain = n_prop_p / elp
    per = {k: c[k] / max(c["n_packets"], 1) for k in ("n_events", "n_binshift", "n_redist", "n_escaped",
                                                       "n_stopped", "n_scatter")}
    return {
        "metric": "packets/s & NLTE solves/s: batched/mixed vs plain GPU path at matched MC noise",
        "value": value, "unit": "packets/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * el / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 geometry / f64 tallies",
        "data": "synthetic (seeded ellipsoidal ejecta and line list)",
        "config": {"workload": "2D axisymmetric ellipsoidal outflow: 128 x 256 (varpi, z) zones, axis ratio "
                               "2, 1e5 fuzz lines, 4096 log bins, time-dependent with census",
                   "new_packets_per_step": N, "dt_s": dt, "zones": P.nz, "nu_bins": P.n_nu,
                   "packets_on_grid_after": P.n_packets},
        "plain_fp64_packets_per_s": plain, "plain_new_packets_per_step": Np,
        "mixed_over_plain": value / plain, "per_packet": per,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": PROP_KERNEL % P.grid.kind, "kernel_ms": prop_avg,
                     "bytes_per_packet": bpk},
        "opacity_build": {"kernel_ms": float(np.mean(op_ms)), "zones": P.nz,
                          "lines": P.cfg.get("n_lines", 0)},
        "nlte": None, "cpu_baseline": None,
    }


def bench_nlte(gpu, torch, W, abi, np):
    """Batched NLTE (256 zones x 4 ions = 1024 systems, n ~ U{20..200}):
    FP32 LU + FP64 refinement in one persistent launch vs the per-system
    FP64 drop-in (one sgpu_nlte_solve per system)."""
    sy = W.nlte_systems(n_sys=1024, seed=20261203)
    d = {k: torch.from_numpy(np.ascontiguousarray(sy[k])).cuda() for k in
         ("n", "rate_off", "vec_off", "rates", "lte")}
    nv = len(sy["lte"])
    b = torch.zeros(nv, dtype=torch.float64, device="cuda:0")
    b64 = torch.zeros(nv, dtype=torch.float64, device="cuda:0")
    it = torch.zeros(sy["n_sys"], dtype=torch.int32, device="cuda:0")
    rep = gpu.nlte_autotune(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], b,
                            None, it)
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        gpu.nlte_batched(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], b, None,
                         abi.NLTE_FP32_REFINE, 1e-13, it)
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / reps
    k_ms = gpu.last_ms("nlte")
    # batched FP64 (same launch, FP64 LU in device memory)
    t0 = time.perf_counter()
    gpu.nlte_batched(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], b64, None,
                     abi.NLTE_FP64, 0.0, None)
    torch.cuda.synchronize()
    t_b64 = time.perf_counter() - t0
    # plain drop-in: one call per system
    work = torch.zeros(200 * 200 + 200, dtype=torch.float64, device="cuda:0")
    bp = torch.zeros(nv, dtype=torch.float64, device="cuda:0")
    ro, vo, ns = sy["rate_off"], sy["vec_off"], sy["n"]
    t0 = time.perf_counter()
    for s in range(sy["n_sys"]):
        gpu.nlte_solve(int(ns[s]), d["rates"][ro[s]:], d["lte"][vo[s]:], bp[vo[s]:], None, work)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    x, r = b.cpu().numpy(), bp.cpu().numpy()
    err = max(np.max(np.abs(x[vo[s]:vo[s + 1]] - r[vo[s]:vo[s + 1]])) / np.max(np.abs(r[vo[s]:vo[s + 1]]))
              for s in range(sy["n_sys"]))
    n2 = (ns.astype(np.float64))
    flops = float(np.sum(2.0 / 3.0 * n2 ** 3))
    itv = it.cpu().numpy()
    return {"systems": int(sy["n_sys"]), "n_range": [int(ns.min()), int(ns.max())],
            "batched_mixed_solves_per_s": sy["n_sys"] / t_mixed,
            "batched_fp64_solves_per_s": sy["n_sys"] / t_b64,
            "plain_fp64_solves_per_s": sy["n_sys"] / t_plain,
            "mixed_over_plain": t_plain / t_mixed,
            "max_rel_err_vs_fp64": float(err),
            "refine_iters": {"min": int(itv.min()), "max": int(itv.max()),
                             "fp64_fallbacks": int((itv < 0).sum())},
            "kernel_ms": k_ms,
            "lu_tflops": flops / (k_ms * 1e-3) / 1e12 if k_ms > 0 else None,
            "autotune": rep.strip().splitlines()[-1] if rep else None}


def bench_cpu(config, seconds):
    """The CPU oracle (FP64 restatement with the reference loop structure:
    OpenMP over packets, atomic tallies) on this host's cores, on a bounded
    packet sample scaled to about `seconds` of work."""
    from pubsed_amd import runtime
    from oracle.oracle_py import OracleBackend
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    ob = OracleBackend(n_threads=cores)
    P = runtime.Problem(runtime.Space(None), config)
    ob.opacity_build(P)
    T0 = P.get("T").copy()
    n = 256
    rate = None
    while True:
        P.alloc_packets(n)
        P.a["T"][:] = T0
        ob.wipe(P)
        t0 = time.perf_counter()
        ob.opacity_build(P)
        ob.emit(P, n)
        ob.propagate(P, n)
        ob.normalize(P)
        if P.cfg["radiative_eq"]:
            ob.solve_eq_temperature(P)
        dt = time.perf_counter() - t0
        rate = n / dt
        if dt > seconds / 4 or n >= 1 << 22:
            break
        n = int(min(n * max(2.0, seconds / max(dt, 1e-3) / 2), 1 << 22))
    return {"value": rate, "unit": "packets/s", "cores": cores, "kind": "port",
            "sample": "%d packets of the same workload, one full step (opacity, emission, "
                      "propagation, normalisation, equilibrium temperature), FP64 oracle, OpenMP "
                      "over packets" % n}


if __name__ == "__main__":
    main()
