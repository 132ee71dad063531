"""Benchmark of the implicit Monte Carlo transport hot path on one MI355X.

Workload (the line-expansion configuration): 1D homologous ejecta, 128
zones to v_max = 2e9 cm/s at t = 20 d, rho ~ exp(-v/3e8) normalised to
1.4 Msun with 56Ni inside 0.6 Msun; 1e5 synthetic fuzz lines, LTE
Saha-Boltzmann for 4 ion stages per element; 4096 log-frequency bins; 1e7
radioactive packets per step, steady state.  A step is one pass of the hot
path with inputs resident in HBM: opacity / emissivity build, emission,
propagation to escape, tally normalisation, radiative-equilibrium gas
temperature.

Reported (one JSON line): packets/s of the batched mixed-precision form
(FP32 propagation, FP64 tallies) as `value`; beside it the plain FP64
drop-in and the batched FP64 form (the FP64 arithmetic with the mixed
form's launch and tally scheme) at the same packet count when the time
budget allows; the propagation kernel's roofline on its own algorithmic
bytes and on the survey formula; the batched NLTE solves (FP32 LU + FP64
refinement and FP64) against the per-system drop-in and against the CPU
oracle; the NLTE-coupled opacity build; the 2D time-dependent outflow
(census, 1e8 packets per step); the differentiable transport's forward /
adjoint times and its gradient against central finite differences; and the
CPU oracle timed on this host's cores (all cores and one thread, per phase).

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
T_START = time.perf_counter()

HBM_PEAK_GBS = 8000.0
# the mixed-form propagation kernel the default launch runs (4-wave register
# bound; the last template argument is the grid kind: 1 sphere, 2 cylinder)
PROP_KERNEL = "propagate_kernel<float, true, true, 4, %d, false>"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=int, default=None,
                    help="packets per step (default: the configuration's own count)")
    ap.add_argument("--config", default="c2_line")
    ap.add_argument("--budget-s", type=float, default=540.0,
                    help="wall-clock budget of the whole run: the FP64 comparison legs run at the "
                         "full packet count when their estimated time fits, else at the largest "
                         "count that does (reported as plain_packets / fp64_batched_packets)")
    ap.add_argument("--traffic-from", default=os.path.join(ROOT, "profiles", "propagate_mixed_c2_line.md"),
                    help="rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the propagation kernel")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (the HBM counter calibration) and exit")
    ap.add_argument("--nlte-only", action="store_true",
                    help="only the batched NLTE measurement (for profiling its kernel)")
    ap.add_argument("--nlte-coupled-only", action="store_true",
                    help="only the coupled-ion-stage NLTE batch (for profiling its kernels)")
    ap.add_argument("--mixed-only", action="store_true",
                    help="only the headline mixed-precision leg (for profiling its kernel)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nlte", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the 2D and differentiable legs")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison legs")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    return ap.parse_args()


def elapsed():
    return time.perf_counter() - T_START


def log(msg):
    print("# [%6.1f s] %s" % (elapsed(), msg), file=sys.stderr, flush=True)


S_STATE = 96   # packet record in HBM: 9 FP64 + 3 int32 + uint32 draw + uint64 id


def packet_bytes(c, nn, precision, survey=False):
    """Algorithmic HBM bytes per packet of the propagation kernel: the packet
    record read and written once (2 S_state), per event the opacity gather
    (2 w_kappa: abs and scat at 8 B in FP64, one packed FP32 (kappa, eps)
    pair in the mixed form) and an 8-byte J_nu accumulate, per
    redistribution the frequency sample, 16 B of spectrum per escape.  The
    FP64 forms sample by binary search over the zone's FP64 emissivity CDF
    (8 ceil(log2 n_nu) B); the mixed form draws one 8-byte Walker alias
    entry.  survey=True charges every form the binary search (the
    B_pkt formula as the survey writes it)."""
    n = max(c["n_packets"], 1)
    n_step = c["n_events"] / n
    n_redist = c["n_redist"] / n
    f_esc = c["n_escaped"] / n
    wk = 4 if precision == 1 else 8
    per_redist = 8 if (precision == 1 and not survey) else 8 * math.ceil(math.log2(nn))
    return 2 * S_STATE + n_step * (2 * wk + 8) + n_redist * per_redist + f_esc * 16


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
    import numpy as np
    import torch
    from pubsed_amd import runtime, abi, workloads as W
    from pubsed_amd.runtime import GpuBackend
    gpu = GpuBackend(0)
    gpu.use_torch_stream()
    gpu.enable_timing(True)
    if a.nlte_only:
        print(json.dumps(bench_nlte(gpu, torch, W, abi, np)))
        return
    if a.nlte_coupled_only:
        print(json.dumps(bench_nlte_coupled(gpu, torch, W, abi, np)))
        return
    if not runtime.CONFIGS[a.config]["steady"]:
        print(json.dumps(bench_time_dependent(a, gpu, a.steps, a.warmup, a.packets)))
        return

    dev = "cuda:0"
    N = a.packets or runtime.CONFIGS[a.config]["n_packets"]
    P = runtime.Problem(runtime.Space(dev), a.config)
    P.alloc_packets(N)
    cnt = torch.zeros(10, dtype=torch.int64, device=dev)

    # every step starts from the same gas temperatures: a step's radiative-
    # equilibrium update is computed (and timed) but not carried into the
    # next step, so that all timed steps (and every form) run one workload
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

    log("config %s, %d packets/step, mixed precision: %d warm-up + %d timed steps" %
        (a.config, N, a.warmup, a.steps))
    dt, prop_ms, c, op_ms = timed(abi.PREC_MIXED, N, a.steps, a.warmup)
    value = a.steps * N / dt
    ms_step = 1e3 * dt / a.steps
    prop_avg = float(np.mean(prop_ms))
    bpk = packet_bytes(c, P.n_nu, 1)
    bpk_survey = packet_bytes(c, P.n_nu, 1, survey=True)
    achieved = bpk * N / (prop_avg * 1e-3) / 1e9
    per_pkt = {k: c[k] / max(c["n_packets"], 1) for k in ("n_events", "n_binshift", "n_redist",
                                                             "n_escaped", "n_scatter")}
    events_per_s = c["n_events"] / (np.sum(prop_ms) * 1e-3)
    log("mixed: %.3f s/step, propagate %.1f ms, events/packet %.1f, %.3g events/s" %
        (dt / a.steps, prop_avg, per_pkt["n_events"], events_per_s))
    # ceiling of the per-event J_nu tally: random FP64 atomics into an array
    # of the J_nu table's size (one per event in the propagation)
    jbuf = torch.zeros(P.nz * P.n_nu, dtype=torch.float64, device=dev)
    adds = 1 << 30
    gpu.atomic_rate(jbuf, adds)                       # warm-up
    jbuf.zero_()
    at_ms = gpu.atomic_rate(jbuf, adds)
    assert float(jbuf.sum().item()) == float(adds)
    atomic_ceiling = adds / (at_ms * 1e-3)
    del jbuf
    traffic, tinfo = read_traffic(a.traffic_from, N)
    traffic_src = None
    if tinfo:
        traffic_src = "%s (rocprofv3 --pmc at %d packets, per-packet bytes scaled to %d)" % (
            os.path.relpath(a.traffic_from, ROOT), tinfo["packets"], N)
    op_avg = float(np.mean(op_ms))
    n_lines = P.cfg.get("n_lines", 0) if P.cfg.get("fuzz") else 0
    # K1-K3 per zone: abs, scat and emissivity CDF written (3 nu 8 B) and the
    # B_nu / means pass over them (nu 8 B); the line list is read through L2
    # (3.2 MB, shared by all zones), so its bytes are not charged per zone.
    # The bound is FP64 exp: 3 per (zone, line)
    b_zone = 4 * P.n_nu * 8
    opacity = {"kernel_ms": op_avg, "zones": P.nz, "lines": n_lines,
               "zone_builds_per_s": P.nz / (op_avg * 1e-3) if op_avg > 0 else None,
               "line_exp_per_s": 3.0 * n_lines * P.nz / (op_avg * 1e-3) if op_avg > 0 else None,
               "table_bytes_per_zone": b_zone}
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_over_algorithmic": (traffic / (bpk * N)) if traffic else None,
                "kernel": PROP_KERNEL % 1, "kernel_ms": prop_avg,
                "bytes_per_packet": bpk,
                "bytes_per_packet_note": "2x96 B record + per event 8 B (kappa, eps) + 8 B J_nu "
                                         "+ 8 B alias entry per redistribution + 16 B per escape",
                "survey_formula_bytes_per_packet": bpk_survey,
                "survey_formula_frac": bpk_survey * N / (prop_avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "events_per_s": events_per_s,
                "jnu_atomic_ceiling_per_s": atomic_ceiling,
                "events_over_atomic_ceiling": events_per_s / atomic_ceiling,
                "atomic_note": "one FP64 atomic J_nu[zone][bin] += E per event (transport.cpp "
                               "J_nu tally); ceiling = sgpu_atomic_rate, 2^30 random FP64 "
                               "atomic adds into an array of the J_nu table's size",
                "traffic_source": traffic_src}
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
        "per_packet": per_pkt,
        "roofline": roofline,
        "opacity_build": opacity,
    }
    if a.mixed_only:
        print(json.dumps(out))
        return

    if not a.no_nlte:
        log("NLTE batched solves (C3)")
        nlte = bench_nlte(gpu, torch, W, abi, np)
        log("NLTE-coupled opacity build")
        nlte["opacity_build_nlte"] = bench_nlte_state(gpu, P, torch, abi, np)
        log("NLTE coupled ion stages (256 per-atom systems, n up to 801)")
        nlte["coupled_ion_stages"] = bench_nlte_coupled(gpu, torch, W, abi, np)
        out["nlte"] = nlte
    if not a.no_extra:
        log("2D time-dependent outflow (C4)")
        out["c4_2d"] = bench_time_dependent(argparse.Namespace(config="c4_2d"), gpu, 2, 1, None, brief=True)
        log("differentiable transport (C5)")
        out["c5_adjoint"] = bench_adjoint(gpu, torch, np)
    if not a.no_cpu:
        log("CPU oracle")
        out["cpu_baseline"] = bench_cpu(a.config, a.cpu_seconds)

    # FP64 comparison legs, the plain drop-in first: each is calibrated on
    # 2e5 packets, then run at the full packet count when its
    # estimated time fits the remaining budget (the plain leg keeps room for
    # the batched FP64 leg at 1e6 packets), else at the largest count that
    # does; both rates are linear in the count from about 4e5 packets up
    if not a.no_fp64:
        esc_m = c["n_escaped"] / max(c["n_packets"], 1)
        n_cal = min(N, 200_000)
        rates = {}
        for key, prec in (("plain_fp64", abi.PREC_FP64), ("fp64_batched", abi.PREC_FP64_BATCHED)):
            dtc, _, _, _ = timed(prec, n_cal, 1, 0)
            rates[key] = n_cal / dtc
        for key, prec in (("plain_fp64", abi.PREC_FP64), ("fp64_batched", abi.PREC_FP64_BATCHED)):
            left = a.budget_s - elapsed() - 15.0
            if key == "plain_fp64":
                left -= min(N, 1_000_000) / rates["fp64_batched"]
            n_leg = N if N / rates[key] <= left else max(int(rates[key] * left), n_cal)
            log("%s: %d packets (estimated %.0f s, %.0f s left)" % (key, n_leg, n_leg / rates[key], left))
            # steps of at most ~60 s each (the rate is linear in the count in
            # this range), each reporting its progress
            chunk = max(n_cal, int(rates[key] * 60.0))
            dtp, prop_sum, cp, done = 0.0, 0.0, None, 0
            while done < n_leg:
                n_c = min(chunk, n_leg - done)
                dtc, prop_c, cc, _ = timed(prec, n_c, 1, 0)
                dtp += dtc
                prop_sum += prop_c[0]
                cp = cc if cp is None else {k: cp[k] + cc[k] for k in cp}
                done += n_c
                log("%s: %d / %d packets, %.1f s" % (key, done, n_leg, dtp))
            out[key] = {"packets": n_leg, "packets_per_s": n_leg / dtp, "propagate_ms": prop_sum,
                        "steps": -(-n_leg // chunk),
                        "escape_fraction": cp["n_escaped"] / max(cp["n_packets"], 1),
                        "events_per_packet": cp["n_events"] / max(cp["n_packets"], 1),
                        "mixed_over_this": value / (n_leg / dtp)}
        out["escape_fraction_mixed"] = esc_m
        out["plain_packets"] = out["plain_fp64"]["packets"]
        out["plain_fp64_packets_per_s"] = out["plain_fp64"]["packets_per_s"]
        out["mixed_over_plain"] = value / out["plain_fp64_packets_per_s"]
    log("done")
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


def bench_time_dependent(a, gpu, steps, warmup, packets, brief=False, graph=True):
    """Time-dependent configuration (the 2D ellipsoidal outflow): every step
    is the homologous hydro step, opacity build, wipe, emission of N new
    packets after the census, propagation to t + dt, compaction and
    normalisation.  Census packets persist, so a timed step (after the
    warm-up steps) propagates the census plus N.  With graph=True the packet
    counts stay on the device and every timed step is captured in a HIP
    graph and launched (runtime.time_step_dev); the warm-up steps run
    uncaptured, the last of them with per-kernel timing for the roofline.
    `value` is packets propagated per second of the whole step."""
    import numpy as np
    import torch
    from pubsed_amd import runtime, abi, workloads as W
    cfg = runtime.CONFIGS[a.config]
    dt = cfg.get("dt", 0.1 * W.DAY)
    cnt = torch.zeros(10, dtype=torch.int64, device="cuda:0")
    cnt_w = torch.zeros(10, dtype=torch.int64, device="cuda:0")

    def run(precision, n_emit, steps, warm, graph):
        P = runtime.Problem(runtime.Space("cuda:0"), a.config)
        cap = n_emit * (steps + warm)
        P.alloc_packets(cap)
        scr = P.alloc_packets(cap, name="scr")
        P.tparams.precision = precision
        d_cnt = torch.zeros(2, dtype=torch.int64, device="cuda:0")
        gpu.use_own_stream()
        try:
            prop_w = op_w = None
            for k in range(warm):
                last = k == warm - 1
                if last:
                    cnt_w.zero_()
                    torch.cuda.synchronize()
                runtime.time_step_dev(gpu, P, dt, n_emit, scr, k, d_cnt, cnt_w if last else None)
                if last:
                    prop_w, op_w = gpu.last_ms("propagate"), gpu.last_ms("opacity")
                log("warm-up step %d: %d packets on the grid" % (k, int(d_cnt[0])))
            torch.cuda.synchronize()
            cnt.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(warm, warm + steps):
                runtime.time_step_dev(gpu, P, dt, n_emit, scr, k, d_cnt, cnt, graph=graph and k > 0)
            gpu.sync()
            el = time.perf_counter() - t0
        finally:
            gpu.use_torch_stream()
        P.n_packets = int(d_cnt[0])
        c = dict(zip(abi.COUNTER_FIELDS, cnt.cpu().tolist()))
        cw = dict(zip(abi.COUNTER_FIELDS, cnt_w.cpu().tolist()))
        return P, el, c["n_packets"], [prop_w], [op_w], c, cw

    N = packets or cfg["n_packets"]
    log("config %s, %d new packets per step, dt %.3g s, mixed precision%s" %
        (a.config, N, dt, ", one HIP graph per step" if graph else ""))
    P, el, n_prop, prop_ms, op_ms, c, cw = run(abi.PREC_MIXED, N, steps, warmup, graph)
    value = n_prop / el
    prop_avg = float(np.mean(prop_ms))
    bpk = packet_bytes(cw, P.n_nu, 1)
    achieved = bpk * cw["n_packets"] / (prop_avg * 1e-3) / 1e9
    log("mixed: %.3f s/step, %d packets propagated; last warm-up step: propagate %.1f ms for %d packets" %
        (el / steps, n_prop, prop_avg, cw["n_packets"]))
    per = {k: c[k] / max(c["n_packets"], 1) for k in ("n_events", "n_binshift", "n_redist", "n_escaped",
                                                       "n_stopped", "n_scatter")}
    spec = P.get("opt_click")
    events_per_s = cw["n_events"] / (prop_avg * 1e-3) if prop_avg else None
    # the per-event J_nu atomic's ceiling at this table's size (1 GB here,
    # far beyond the L2 and MALL)
    jbuf = torch.zeros(P.nz * P.n_nu, dtype=torch.float64, device="cuda:0")
    adds = 1 << 30
    gpu.atomic_rate(jbuf, adds)
    at_ms = gpu.atomic_rate(jbuf, adds)
    del jbuf
    atomic_ceiling = adds / (at_ms * 1e-3)
    res = {
        "metric": "packets/s & NLTE solves/s: batched/mixed vs plain GPU path at matched MC noise",
        "value": value, "unit": "packets/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * el / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 geometry / f64 tallies",
        "data": "synthetic (seeded ellipsoidal ejecta and line list)",
        "config": {"workload": "2D axisymmetric ellipsoidal outflow: 128 x 256 (varpi, z) zones, axis ratio "
                               "2, 1e5 fuzz lines, 4096 log bins, time-dependent with census, escape "
                               "spectrum in %d time x %d nu x %d mu bins" % P.spec_shape[:3],
                   "new_packets_per_step": N, "dt_s": dt, "zones": P.nz, "nu_bins": P.n_nu,
                   "packets_on_grid_after": P.n_packets},
        "per_packet": per, "spectrum_packets": float(spec.sum()),
        "hip_graph_per_step": bool(graph), "graph_nodes": gpu.graph_nodes() if graph else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": PROP_KERNEL % P.grid.kind, "kernel_ms": prop_avg,
                     "bytes_per_packet": bpk,
                     "events_per_s": events_per_s,
                     "jnu_atomic_ceiling_per_s": atomic_ceiling,
                     "events_over_atomic_ceiling": events_per_s / atomic_ceiling if events_per_s else None,
                     "measured_on": "the last warm-up step, run uncaptured with HIP event timing "
                                    "(%d packets)" % cw["n_packets"]},
        "opacity_build": {"kernel_ms": float(np.mean(op_ms)), "zones": P.nz,
                          "lines": P.cfg.get("n_lines", 0)},
    }
    if brief:
        for k in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"):
            res.pop(k)
        return res
    Np = min(N, 10_000_000)
    log("plain FP64: %d new packets per step" % Np)
    _, elp, n_prop_p, _, _, cp, _ = run(abi.PREC_FP64, Np, 1, max(warmup, 1), False)
    res["plain_fp64_packets_per_s"] = n_prop_p / elp
    res["plain_new_packets_per_step"] = Np
    res["mixed_over_plain"] = value / res["plain_fp64_packets_per_s"]
    res["nlte"] = None
    res["cpu_baseline"] = None
    return res


def bench_nlte(gpu, torch, W, abi, np):
    """Batched NLTE (256 zones x 4 ions = 1024 systems, n ~ U{20..200}):
    FP32 LU + FP64 refinement in one persistent launch and the batched FP64
    LU, against the per-system FP64 drop-in (one sgpu_nlte_solve per system)
    and against the CPU oracle's FP64 LU (normwise max |b - b_oracle| /
    max |b_oracle| over the systems)."""
    from oracle.oracle_py import OracleBackend
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
    # batched FP64 (same launch, FP64 LU)
    gpu.nlte_batched(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], b64, None,
                     abi.NLTE_FP64, 0.0, None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        gpu.nlte_batched(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], b64, None,
                         abi.NLTE_FP64, 0.0, None)
    torch.cuda.synchronize()
    t_b64 = (time.perf_counter() - t0) / reps
    k64_ms = gpu.last_ms("nlte")
    # plain drop-in: one call per system
    work = torch.zeros(200 * 200 + 200, dtype=torch.float64, device="cuda:0")
    bp = torch.zeros(nv, dtype=torch.float64, device="cuda:0")
    ro, vo, ns = sy["rate_off"], sy["vec_off"], sy["n"]
    t0 = time.perf_counter()
    for s in range(sy["n_sys"]):
        gpu.nlte_solve(int(ns[s]), d["rates"][ro[s]:], d["lte"][vo[s]:], bp[vo[s]:], None, work)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    # the CPU oracle's FP64 LU of every system
    ob = OracleBackend()
    ref = np.zeros(nv)
    for s in range(sy["n_sys"]):
        m = int(ns[s])
        out = np.zeros(m)
        ob.nlte_solve(m, np.ascontiguousarray(sy["rates"][ro[s]:ro[s + 1]]),
                      np.ascontiguousarray(sy["lte"][vo[s]:vo[s + 1]]), out)
        ref[vo[s]:vo[s + 1]] = out

    def err(x):
        return float(max(np.max(np.abs(x[vo[s]:vo[s + 1]] - ref[vo[s]:vo[s + 1]])) /
                         np.max(np.abs(ref[vo[s]:vo[s + 1]])) for s in range(sy["n_sys"])))

    x, x64, xp = b.cpu().numpy(), b64.cpu().numpy(), bp.cpu().numpy()
    n2 = ns.astype(np.float64)
    flops = float(np.sum(2.0 / 3.0 * n2 ** 3))
    itv = it.cpu().numpy()
    return {"systems": int(sy["n_sys"]), "n_range": [int(ns.min()), int(ns.max())],
            "batched_mixed_solves_per_s": sy["n_sys"] / t_mixed,
            "batched_fp64_solves_per_s": sy["n_sys"] / t_b64,
            "plain_fp64_solves_per_s": sy["n_sys"] / t_plain,
            "mixed_over_plain": t_plain / t_mixed,
            "max_rel_err_vs_oracle": {"batched_mixed": err(x), "batched_fp64": err(x64), "plain_fp64": err(xp)},
            "refine_iters": {"min": int(itv.min()), "max": int(itv.max()),
                             "fp64_fallbacks": int((itv < 0).sum())},
            "kernel_ms": k_ms, "kernel_fp64_ms": k64_ms,
            "lu_tflops": flops / (k_ms * 1e-3) / 1e12 if k_ms > 0 else None,
            "autotune": rep.strip().splitlines()[-1] if rep else None}


def bench_nlte_coupled(gpu, torch, W, abi, np):
    """The reference-shaped NLTE batch: one system per (zone, atom) with the
    atom's four ion stages coupled (AtomicSpecies::solve_nlte over every
    level of the atom, AtomicSpecies_state.cpp:131-196), 256 zones, stages of
    U{20..200} levels, n up to 801 (workloads.nlte_systems_coupled): the
    batched launch in FP64 and with FP32 + refinement (systems above the
    201-level FP32 LDS limit solve in FP64), the per-system FP64 drop-in on a
    sample of 32 systems, and the error against the CPU oracle on a sample
    of 16."""
    from oracle.oracle_py import OracleBackend
    sy = W.nlte_systems_coupled(n_sys=256, seed=20261204)
    d = {k: torch.from_numpy(np.ascontiguousarray(sy[k])).cuda() for k in
         ("n", "rate_off", "vec_off", "rates", "lte")}
    nv = len(sy["lte"])
    ro, vo, ns = sy["rate_off"], sy["vec_off"], sy["n"]
    out = {"systems": int(sy["n_sys"]), "n_range": [int(ns.min()), int(ns.max())],
           "lu_gflop": float(np.sum(2.0 / 3.0 * ns.astype(np.float64) ** 3)) / 1e9}
    bt = torch.zeros(nv, dtype=torch.float64, device="cuda:0")
    itt = torch.zeros(sy["n_sys"], dtype=torch.int32, device="cuda:0")
    rep = gpu.nlte_autotune(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], bt, None, itt)
    out["autotune"] = [l for l in rep.strip().splitlines() if l.startswith("chosen")] if rep else None
    res = {}
    for name, prec in (("batched_fp64", abi.NLTE_FP64), ("batched_mixed", abi.NLTE_FP32_REFINE)):
        b = torch.zeros(nv, dtype=torch.float64, device="cuda:0")
        it = torch.zeros(sy["n_sys"], dtype=torch.int32, device="cuda:0")
        gpu.nlte_batched(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], b, None,
                         prec, 1e-13, it)
        torch.cuda.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            gpu.nlte_batched(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], b, None,
                             prec, 1e-13, it)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        out[name + "_solves_per_s"] = sy["n_sys"] / t
        out[name + "_kernel_ms"] = gpu.last_ms("nlte")
        res[name] = b.cpu().numpy()
        itv = it.cpu().numpy()
        out[name + "_fp64_solved"] = int((itv <= 0).sum())
    # per-system drop-in on a sample
    sample = np.arange(0, sy["n_sys"], sy["n_sys"] // 32)[:32]
    m = int(ns.max())
    work = torch.zeros(m * m + m, dtype=torch.float64, device="cuda:0")
    bp = torch.zeros(nv, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in sample:
        gpu.nlte_solve(int(ns[s]), d["rates"][ro[s]:], d["lte"][vo[s]:], bp[vo[s]:], None, work)
    torch.cuda.synchronize()
    out["plain_fp64_solves_per_s"] = len(sample) / (time.perf_counter() - t0)
    out["batched_fp64_over_plain"] = out["batched_fp64_solves_per_s"] / out["plain_fp64_solves_per_s"]
    ob = OracleBackend()
    errs = {k: 0.0 for k in res}
    for s in sample[:16]:
        n = int(ns[s])
        ref = np.zeros(n)
        ob.nlte_solve(n, np.ascontiguousarray(sy["rates"][ro[s]:ro[s + 1]]),
                      np.ascontiguousarray(sy["lte"][vo[s]:vo[s + 1]]), ref)
        for k, x in res.items():
            errs[k] = max(errs[k], float(np.max(np.abs(x[vo[s]:vo[s + 1]] - ref)) / np.max(np.abs(ref))))
    out["max_rel_err_vs_oracle"] = errs
    return out


def bench_adjoint(gpu, torch, np):
    """Differentiable transport at 1e7 packets on 128 zones x 1024 bins of
    absorption opacity: forward spectrum (recording each packet's escape bin
    and weight) and the adjoint replay, timed; the adjoint gradient of
    L = sum (F - F_target)^2 along two random directions against central
    finite differences with common random numbers (relative step 1e-3)."""
    from pubsed_amd.diff import DiffTransport
    n = 10_000_000
    m = DiffTransport(gpu, n_packets=n)
    kappa = m.kappa_init()
    rng = np.random.default_rng(11)
    noise = torch.from_numpy((1 + 0.1 * rng.uniform(-1, 1, kappa.shape)).astype(np.float32)).cuda()
    F_target = m.forward(kappa * noise, record=False)
    m.forward(kappa)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    F = m.forward(kappa)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    dLdF = 2 * (F - F_target)
    t0 = time.perf_counter()
    g = m.backward(kappa, dLdF)
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    L = lambda k: float(((m.forward(k, record=False) - F_target) ** 2).sum())
    errs = []
    for s in range(2):
        v = torch.from_numpy(np.random.default_rng(100 + s).standard_normal(kappa.shape).astype(np.float32)).cuda()
        h = 1e-3
        kp, km = kappa * (1 + h * v), kappa * (1 - h * v)
        fd = (L(kp) - L(km)) / 2
        ad = float((g * (kappa * v).double()).sum()) * h
        errs.append(abs(ad - fd) / abs(fd))
    return {"packets": n, "params": list(kappa.shape), "forward_ms": 1e3 * t_fwd, "adjoint_ms": 1e3 * t_bwd,
            "forward_packets_per_s": n / t_fwd, "adjoint_packets_per_s": n / t_bwd,
            "grad_rel_err_vs_central_fd": errs}


def bench_cpu(config, seconds):
    """The CPU oracle (FP64 restatement with the reference loop structure:
    OpenMP over packets, atomic tallies) on this host's cores and on one
    thread, phase by phase (opacity build, emission, propagation,
    normalisation, equilibrium temperature) as the reference's step prints
    them, on a bounded packet sample scaled to about `seconds` of
    propagation per thread count."""
    from pubsed_amd import runtime
    from oracle.oracle_py import OracleBackend
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    res = {}
    for label, threads in (("all_cores", cores), ("one_thread", 1)):
        ob = OracleBackend(n_threads=threads)
        P = runtime.Problem(runtime.Space(None), config)
        T0 = P.get("T").copy()
        n = 64
        while True:
            P.alloc_packets(n)
            P.a["T"][:] = T0
            ob.wipe(P)
            ph = {}
            t = time.perf_counter()
            ob.opacity_build(P)
            ph["opacity"] = time.perf_counter() - t
            t = time.perf_counter()
            ob.emit(P, n)
            ph["emission"] = time.perf_counter() - t
            t = time.perf_counter()
            ob.propagate(P, n)
            ph["propagation"] = time.perf_counter() - t
            t = time.perf_counter()
            ob.normalize(P)
            ph["normalisation"] = time.perf_counter() - t
            t = time.perf_counter()
            if P.cfg["radiative_eq"]:
                ob.solve_eq_temperature(P)
            ph["equilibrium_temperature"] = time.perf_counter() - t
            tot = sum(ph.values())
            if ph["propagation"] > seconds / 3 or n >= 1 << 22:
                break
            n = int(min(n * max(2.0, seconds / 2 / max(ph["propagation"], 1e-3)), 1 << 22))
        res[label] = {"threads": threads, "packets": n, "packets_per_s": n / tot,
                      "phase_ms": {k: 1e3 * v for k, v in ph.items()},
                      "propagation_packets_per_s": n / ph["propagation"]}
    a = res["all_cores"]
    return {"value": a["packets_per_s"], "unit": "packets/s", "cores": a["threads"], "kind": "port",
            "sample": "%d packets (all cores) / %d packets (1 thread) of the same workload, one full "
                      "step each (opacity, emission, propagation, normalisation, equilibrium "
                      "temperature), FP64 oracle, OpenMP over packets"
                      % (a["packets"], res["one_thread"]["packets"]),
            "one_thread_packets_per_s": res["one_thread"]["packets_per_s"],
            "phases": res}


if __name__ == "__main__":
    main()
