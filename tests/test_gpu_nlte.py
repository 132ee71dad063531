"""GPU parity of the NLTE solves (K8) against the FP64 oracle LU.

Bars: FP64 forms (single drop-in, batched FP64) within 1e-11 normwise of
the oracle's departure coefficients; the FP32-LU + FP64-refinement form
within 1e-8 normwise.  In that form a system whose FP64 matrix fits the
launch's LDS is factored directly in FP64 (iters 0); every larger one must
converge in FP32 (iters > 0), so the MFMA trailing update is exercised
rather than the FP64 fallback (iters -1).
"""
import numpy as np
import pytest

from pubsed_amd import workloads as W, abi
from oracle.oracle_py import OracleBackend

pytestmark = pytest.mark.gpu


def fp64_direct(n, max_n):
    """True where the FP32 + refinement launch factors directly in FP64: the
    launch's LDS holds the FP32 matrix of min(max_n, 201) levels or the FP64
    one of min(max_n, 143), whichever is larger (row stride n | 1)."""
    c32, c64 = min(max_n, 201), min(max_n, 143)
    lds = max(4 * c32 * (c32 | 1), 32 * c32, 8 * c64 * (c64 | 1), 32 * c64)
    n = np.asarray(n)
    return 8 * n * (n | 1) <= lds


def oracle_solutions(sy):
    ob = OracleBackend()
    out = np.zeros_like(sy["lte"])
    for s in range(sy["n_sys"]):
        n = int(sy["n"][s])
        R = np.ascontiguousarray(sy["rates"][sy["rate_off"][s]:sy["rate_off"][s + 1]])
        l = np.ascontiguousarray(sy["lte"][sy["vec_off"][s]:sy["vec_off"][s + 1]])
        b = np.zeros(n)
        ob.nlte_solve(n, R, l, b)
        out[sy["vec_off"][s]:sy["vec_off"][s + 1]] = b
    return out


def normwise(sy, x, ref):
    errs = []
    for s in range(sy["n_sys"]):
        a, b = sy["vec_off"][s], sy["vec_off"][s + 1]
        errs.append(np.max(np.abs(x[a:b] - ref[a:b])) / np.max(np.abs(ref[a:b])))
    return np.array(errs)


def to_dev(sy):
    import torch
    d = {}
    for k in ("n", "rate_off", "vec_off", "rates", "lte"):
        d[k] = torch.from_numpy(np.ascontiguousarray(sy[k])).to("cuda:0")
    nv = len(sy["lte"])
    d["b"] = torch.zeros(nv, dtype=torch.float64, device="cuda:0")
    d["pop"] = torch.zeros(nv, dtype=torch.float64, device="cuda:0")
    d["iters"] = torch.zeros(sy["n_sys"], dtype=torch.int32, device="cuda:0")
    return d


def test_single_system_drop_in(gpu):
    import torch
    sy = W.nlte_systems(n_sys=6, n_min=20, n_max=200, seed=11)
    ref = oracle_solutions(sy)
    for s in range(sy["n_sys"]):
        n = int(sy["n"][s])
        R = torch.from_numpy(sy["rates"][sy["rate_off"][s]:sy["rate_off"][s + 1]].copy()).cuda()
        l = torch.from_numpy(sy["lte"][sy["vec_off"][s]:sy["vec_off"][s + 1]].copy()).cuda()
        b = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        pop = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        work = torch.zeros(n * n + n, dtype=torch.float64, device="cuda:0")
        print("case nlte single", s, n, flush=True)
        gpu.nlte_solve(n, R, l, b, pop, work)
        gpu.sync()
        r = ref[sy["vec_off"][s]:sy["vec_off"][s + 1]]
        err = np.max(np.abs(b.cpu().numpy() - r)) / np.max(np.abs(r))
        assert err <= 1e-11, (s, n, err)
        p = pop.cpu().numpy()
        assert np.allclose(p, np.maximum(r * l.cpu().numpy(), 1e-30), rtol=1e-10, atol=0)


@pytest.mark.parametrize("precision,bar,n_max", [(abi.NLTE_FP64, 1e-11, 200), (abi.NLTE_FP64, 1e-11, 140),
                                                 (abi.NLTE_FP32_REFINE, 1e-8, 200)])
def test_batched(gpu, precision, bar, n_max):
    """n_max 200: FP64 LU in device memory; 140: FP64 LU in LDS (n <= 143)."""
    sy = W.nlte_systems(n_sys=96, n_max=n_max, seed=5)
    ref = oracle_solutions(sy)
    d = to_dev(sy)
    print("case nlte batched", precision, flush=True)
    gpu.nlte_batched(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], d["b"],
                     d["pop"], precision, 1e-13, d["iters"])
    gpu.sync()
    b = d["b"].cpu().numpy()
    errs = normwise(sy, b, ref)
    print("max err", errs.max(), "iters", np.bincount(d["iters"].cpu().numpy() + 1))
    assert errs.max() <= bar
    if precision == abi.NLTE_FP32_REFINE:
        it = d["iters"].cpu().numpy()
        direct = fp64_direct(sy["n"], int(np.max(sy["n"])))
        assert (it[direct] == 0).all() and direct.any()
        assert (it[~direct] > 0).all(), "some systems fell back to FP64: %s" % it


@pytest.mark.parametrize("precision,n_max,bar", [(abi.NLTE_FP64, 140, 1e-11), (abi.NLTE_FP32_REFINE, 201, 1e-8)])
def test_every_launch_shape(gpu, precision, n_max, bar):
    """Every shape the autotuner may pick (workgroup size x panel width x
    workgroups per CU) solves the same uneven batch to the oracle's answer;
    sizes straddle the panel widths and the MFMA tile edges."""
    sy = W.nlte_systems(n_sys=40, n_min=2, n_max=n_max, seed=11)
    ref = oracle_solutions(sy)
    d = to_dev(sy)
    try:
        for threads in (128, 256, 512):
            for panel in (8, 16):
                for wg in (1, 2):
                    print("case nlte shape", precision, threads, panel, wg, flush=True)
                    gpu.nlte_set_tuning(precision, threads, panel, wg)
                    d["b"].zero_()
                    gpu.nlte_batched(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"],
                                     d["b"], None, precision, 1e-13, d["iters"])
                    gpu.sync()
                    errs = normwise(sy, d["b"].cpu().numpy(), ref)
                    assert errs.max() <= bar, (threads, panel, wg, errs.max())
                    if precision == abi.NLTE_FP32_REFINE:
                        it = d["iters"].cpu().numpy()
                        direct = fp64_direct(sy["n"], int(np.max(sy["n"])))
                        assert (it[direct] == 0).all() and (it[~direct] > 0).all(), (threads, panel, wg, it)
    finally:
        gpu.nlte_set_tuning(precision, 256, 16, 1)


@pytest.mark.parametrize("precision", [abi.NLTE_FP64, abi.NLTE_FP32_REFINE])
def test_coupled_ion_stages(gpu, precision):
    """Per-atom systems as AtomicSpecies::solve_nlte solves them
    (AtomicSpecies_state.cpp:131-196): four ion stages of U{20..200} levels
    coupled by bound-free rates, n up to 801, beside small systems in one
    batch.  Above the FP32 LDS limit (201 levels) the FP32 + refinement
    launch solves in FP64 (iters -1); every system within 1e-11 of the
    oracle."""
    big = W.nlte_systems_coupled(n_sys=10, seed=21)
    small = W.nlte_systems(n_sys=6, n_min=20, n_max=120, seed=22)
    sy = dict(n_sys=16, n=np.concatenate([big["n"], small["n"]]).astype(np.int32))
    sy["rate_off"] = np.concatenate([[0], np.cumsum(sy["n"].astype(np.int64) ** 2)])
    sy["vec_off"] = np.concatenate([[0], np.cumsum(sy["n"].astype(np.int64))])
    sy["rates"] = np.concatenate([big["rates"], small["rates"]])
    sy["lte"] = np.concatenate([big["lte"], small["lte"]])
    assert sy["n"].max() > 600
    ref = oracle_solutions(sy)
    d = to_dev(sy)
    print("case nlte coupled ion stages precision=%d n_max=%d" % (precision, sy["n"].max()), flush=True)
    gpu.nlte_batched(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], d["b"],
                     d["pop"], precision, 1e-13, d["iters"])
    gpu.sync()
    errs = normwise(sy, d["b"].cpu().numpy(), ref)
    print("max err", errs.max())
    assert errs.max() <= 1e-11
    pop = d["pop"].cpu().numpy()
    assert np.allclose(pop, np.maximum(d["b"].cpu().numpy() * sy["lte"], 1e-30), rtol=1e-15, atol=0)
    if precision == abi.NLTE_FP32_REFINE:
        it = d["iters"].cpu().numpy()
        assert (it[sy["n"] > 201] == -1).all()


def test_uneven_sizes_edge(gpu):
    """Tiny, panel-straddling and LDS-limit sizes in one batch (n = 2 ... 201,
    the largest FP32 matrix with an odd row stride in 160 KiB of LDS), plus 202
    and 230, which exceed it and take the FP64 path; up to 143 levels the
    launch factors in FP64 directly."""
    rng = np.random.default_rng(3)
    sizes = [2, 3, 31, 32, 33, 64, 65, 97, 128, 160, 201, 202, 230]
    sy = W.nlte_systems(n_sys=len(sizes), seed=9)
    # rebuild with the chosen sizes
    parts = [W.nlte_systems(n_sys=1, n_min=n, n_max=n, seed=100 + i) for i, n in enumerate(sizes)]
    sy = dict(n_sys=len(sizes), n=np.array(sizes, np.int32))
    sy["rate_off"] = np.concatenate([[0], np.cumsum(np.array(sizes, np.int64) ** 2)])
    sy["vec_off"] = np.concatenate([[0], np.cumsum(np.array(sizes, np.int64))])
    sy["rates"] = np.concatenate([p["rates"] for p in parts])
    sy["lte"] = np.concatenate([p["lte"] for p in parts])
    ref = oracle_solutions(sy)
    d = to_dev(sy)
    print("case nlte edge sizes", flush=True)
    gpu.nlte_batched(sy["n_sys"], d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], d["b"],
                     d["pop"], abi.NLTE_FP32_REFINE, 1e-13, d["iters"])
    gpu.sync()
    errs = normwise(sy, d["b"].cpu().numpy(), ref)
    it = d["iters"].cpu().numpy()
    print("errs", errs, "iters", it)
    assert errs.max() <= 1e-8
    assert (it[-2:] == -1).all()    # n = 202, 230 > LDS limit: FP64 path
    direct = fp64_direct(sy["n"][:-2], 230)
    assert (it[:-2][direct] == 0).all() and (it[:-2][~direct] > 0).all()
    assert (~direct).sum() == 2      # 160 and 201: FP32 + refinement


def rate_case(nz=6, near_lte=False, seed=5):
    """Inputs of set_rates for every (zone, atom) system of a small
    line-expansion problem from its LTE state: seeded J-bar / P_ic / R_ci,
    or (near_lte) J-bar = 0.7 B_nu(T) with P_ic / R_ci balanced at LTE."""
    from pubsed_amd import runtime
    from test_oracle_nlte import lte_inputs
    Ph = runtime.Problem(runtime.Space(None), "c2_line", nz=nz, n_lines=200)
    OracleBackend().opacity_build(Ph)
    A = Ph.atoms_np
    na, nl, nb = Ph.atoms.n_atoms, Ph.atoms.n_levels, Ph.atoms.n_blines
    lev = Ph.get("lev_n").reshape(nz, nl)
    T, ne = Ph.get("T"), Ph.get("n_elec")
    sys_atom = np.array([a for z in range(nz) for a in range(na)], np.int32)
    sys_zone = np.array([z for z in range(nz) for a in range(na)], np.int32)
    ns = len(sys_atom)
    rng = np.random.default_rng(seed)
    if near_lte:
        J, Pic, Rci = np.zeros((ns, nb)), np.zeros((ns, nl)), np.zeros((ns, nl))
        for s in range(ns):
            z = sys_zone[s]
            J[s], Pic[s], Rci[s] = lte_inputs(Ph, lev, z, ne[z], T[z], seed=s)
        J *= 0.7
    else:
        J = 10.0 ** rng.uniform(-8, -3, (ns, nb))
        Pic = 10.0 ** rng.uniform(-4, 0, (ns, nl))
        Rci = 10.0 ** rng.uniform(-14, -10, (ns, nl))
    n = np.diff(A["lev_off"])[sys_atom].astype(np.int32)
    return dict(Ph=Ph, A=A, sys_atom=sys_atom, ne=ne[sys_zone].copy(), T=T[sys_zone].copy(), J=J, Pic=Pic,
                Rci=Rci, lte_in=lev[sys_zone], n=n,
                rate_off=np.concatenate([[0], np.cumsum(n.astype(np.int64) ** 2)]),
                vec_off=np.concatenate([[0], np.cumsum(n.astype(np.int64))]))


def oracle_rates(c):
    import ctypes
    from oracle import oracle_py
    lib = oracle_py.load()
    Ph, A = c["Ph"], c["A"]
    ns = len(c["sys_atom"])
    R_h, l_h, b_h = np.zeros(c["rate_off"][-1]), np.zeros(c["vec_off"][-1]), np.zeros(c["vec_off"][-1])
    ln_h = np.zeros((ns, Ph.atoms.n_levels))
    fr_h = np.zeros((ns, Ph.atoms.n_ions))
    for s in range(ns):
        a, n = int(c["sys_atom"][s]), int(c["n"][s])
        R, l, b = np.zeros(n * n), np.zeros(n), np.zeros(n)
        lib.or_nlte_set_rates(ctypes.byref(Ph.atoms), a, c["ne"][s], c["T"][s],
                              *[np.ascontiguousarray(c[k][s]).ctypes.data for k in ("J", "Pic", "Rci", "lte_in")],
                              1, 1, R.ctypes.data, l.ctypes.data)
        lib.or_nlte_solve(n, R.ctypes.data, l.ctypes.data, b.ctypes.data, None)
        ro, vo = c["rate_off"], c["vec_off"]
        R_h[ro[s]:ro[s + 1]], l_h[vo[s]:vo[s + 1]], b_h[vo[s]:vo[s + 1]] = R, l, b
        l0, l1 = A["lev_off"][a], A["lev_off"][a + 1]
        i0, i1 = A["ion_off"][a], A["ion_off"][a + 1]
        ln, fr = np.zeros(n), np.zeros(i1 - i0)
        lib.or_nlte_populations(ctypes.byref(Ph.atoms), a, b.ctypes.data, l.ctypes.data, ln.ctypes.data,
                                fr.ctypes.data)
        ln_h[s, l0:l1], fr_h[s, i0:i1] = ln, fr
    return R_h, l_h, b_h, ln_h, fr_h


def gpu_rates(gpu, c):
    import ctypes
    import torch
    from pubsed_amd import runtime, check
    Ph = c["Ph"]
    Pg = runtime.Problem(runtime.Space("cuda:0"), "c2_line", nz=Ph.nz, n_lines=200)
    ns = len(c["sys_atom"])
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    d = {k: dev(c[k]) for k in ("sys_atom", "ne", "T", "J", "Pic", "Rci", "lte_in", "n", "rate_off", "vec_off")}
    R_g = torch.zeros(int(c["rate_off"][-1]), dtype=torch.float64, device="cuda:0")
    l_g = torch.zeros(int(c["vec_off"][-1]), dtype=torch.float64, device="cuda:0")
    b_g = torch.zeros_like(l_g)
    ln_g = torch.zeros((ns, Ph.atoms.n_levels), dtype=torch.float64, device="cuda:0")
    fr_g = torch.zeros((ns, Ph.atoms.n_ions), dtype=torch.float64, device="cuda:0")
    p = abi.ptr
    check(gpu.lib.sgpu_nlte_set_rates(gpu.ctx, ctypes.byref(Pg.atoms), ns,
                                      *[p(d[k]) for k in ("sys_atom", "ne", "T", "J", "Pic", "Rci", "lte_in")],
                                      1, 1, p(d["n"]), p(d["rate_off"]), p(d["vec_off"]), p(R_g), p(l_g)),
          "sgpu_nlte_set_rates")
    gpu.nlte_batched(ns, d["n"], d["rate_off"], d["vec_off"], R_g, l_g, b_g, None, abi.NLTE_FP64, 0.0, None)
    check(gpu.lib.sgpu_nlte_populations(gpu.ctx, ctypes.byref(Pg.atoms), ns, p(d["sys_atom"]), p(d["vec_off"]),
                                        p(b_g), p(l_g), p(ln_g), p(fr_g)), "sgpu_nlte_populations")
    gpu.sync()
    return [x.cpu().numpy() for x in (R_g, l_g, b_g, ln_g, fr_g)]


def backward_error(c, R, l, b, s):
    n = int(c["n"][s])
    ro, vo = c["rate_off"], c["vec_off"]
    Rm = R[ro[s]:ro[s + 1]].reshape(n, n)
    lv, bv = l[vo[s]:vo[s + 1]], b[vo[s]:vo[s + 1]]
    M = Rm.T.copy()
    np.fill_diagonal(M, -Rm.sum(1))
    M[n - 1, :] = lv
    rhs = np.zeros(n)
    rhs[-1] = 1
    return np.max(np.abs(M @ bv - rhs)) / (np.max(np.abs(M).sum(1)) * np.max(np.abs(bv)))


def test_set_rates_matches_oracle_and_solves_stably(gpu):
    """Rate assembly (set_rates) for every (zone, atom) system from seeded
    J-bar / P_ic / R_ci: every rate within 1e-13 of the oracle (device pow /
    exp against glibc's), the same sparsity, the same lte.  These seeded
    rates make systems with condition numbers up to 1e40 (rows scaled by LTE
    populations spanning 1e-30..1; the oracle's LU and LAPACK's part by
    5e-3 there), so the batched solve is held to what a backward-stable
    solver guarantees whatever the conditioning: backward error <= 1e-13."""
    c = rate_case()
    print("case set_rates seeded", len(c["sys_atom"]), "systems", flush=True)
    R_h, l_h, _, _, _ = oracle_rates(c)
    R_g, l_g, b_g, _, _ = gpu_rates(gpu, c)
    nzm = R_h != 0
    assert np.array_equal(R_g != 0, nzm)
    assert np.max(np.abs(R_g[nzm] - R_h[nzm]) / np.abs(R_h[nzm])) <= 1e-13
    assert np.array_equal(l_g, l_h)
    for s in range(len(c["sys_atom"])):
        assert backward_error(c, R_g, l_g, b_g, s) <= 1e-13, s


def test_set_rates_solve_and_populations_near_lte_match_oracle(gpu):
    """Near-LTE inputs (J-bar = 0.7 B_nu(T), P_ic / R_ci balanced at LTE):
    populations and ion fractions within 1e-7 of the atom's total, b of
    levels holding >= 1e-6 of the atom within 1e-5 (the oracle's LU and
    LAPACK agree to 2e-8 and 2e-6 on these systems)."""
    c = rate_case(near_lte=True)
    print("case set_rates near LTE", flush=True)
    R_h, l_h, b_h, ln_h, fr_h = oracle_rates(c)
    R_g, l_g, b_g, ln_g, fr_g = gpu_rates(gpu, c)
    nzm = R_h != 0
    assert np.max(np.abs(R_g[nzm] - R_h[nzm]) / np.abs(R_h[nzm])) <= 1e-13
    A, vo = c["A"], c["vec_off"]
    for s in range(len(c["sys_atom"])):
        a = int(c["sys_atom"][s])
        l0, l1 = A["lev_off"][a], A["lev_off"][a + 1]
        i0, i1 = A["ion_off"][a], A["ion_off"][a + 1]
        tot = ln_h[s, l0:l1].sum()
        assert np.max(np.abs(ln_g[s, l0:l1] - ln_h[s, l0:l1])) <= 1e-7 * tot, s
        assert np.max(np.abs(fr_g[s, i0:i1] - fr_h[s, i0:i1])) <= 1e-7 * tot, s
        lv = l_h[vo[s]:vo[s + 1]]
        big = lv >= 1e-6 * lv.sum()
        bh, bg = b_h[vo[s]:vo[s + 1]], b_g[vo[s]:vo[s + 1]]
        assert np.max(np.abs(bg[big] - bh[big]) / np.abs(bh[big])) <= 1e-5, s


def nlte_pass_gpu(gpu, Pg, sys_atom, sys_zone, T_s, ne_s, lte_in, J_nu):
    """Radiative rates -> set_rates -> batched FP64 solve -> populations, on
    the GPU, for the (zone, atom) systems given; returns (Pic, Rci, line_J,
    b, lte) as numpy."""
    import ctypes
    import torch
    from pubsed_amd import check
    at = Pg.atoms
    A = Pg.atoms_np
    ns = len(sys_atom)
    n = np.diff(A["lev_off"])[sys_atom].astype(np.int32)
    rate_off = np.concatenate([[0], np.cumsum(n.astype(np.int64) ** 2)])
    vec_off = np.concatenate([[0], np.cumsum(n.astype(np.int64))])
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    d = {k: dev(v) for k, v in dict(sys_atom=sys_atom, sys_zone=sys_zone, T=T_s, ne=ne_s, lte_in=lte_in,
                                      J_nu=J_nu, elem_A=Pg.ej["elem_A"].astype(np.int32), n=n,
                                      rate_off=rate_off, vec_off=vec_off).items()}
    z = lambda m: torch.zeros(m, dtype=torch.float64, device="cuda:0")
    Pic, Rci, lJ = z(ns * at.n_levels), z(ns * at.n_levels), z(ns * at.n_blines)
    R, l, b = z(int(rate_off[-1])), z(int(vec_off[-1])), z(int(vec_off[-1]))
    p = abi.ptr
    check(gpu.lib.sgpu_nlte_radiative_rates(gpu.ctx, ctypes.byref(at), ctypes.byref(Pg.nu), ns, p(d["sys_atom"]),
                                            p(d["sys_zone"]), p(d["T"]), p(d["elem_A"]), 0.0, p(d["J_nu"]),
                                            p(Pic), p(Rci), p(lJ)), "sgpu_nlte_radiative_rates")
    check(gpu.lib.sgpu_nlte_set_rates(gpu.ctx, ctypes.byref(at), ns, p(d["sys_atom"]), p(d["ne"]), p(d["T"]),
                                      p(lJ), p(Pic), p(Rci), p(d["lte_in"]), 1, 0, p(d["n"]), p(d["rate_off"]),
                                      p(d["vec_off"]), p(R), p(l)), "sgpu_nlte_set_rates")
    gpu.nlte_batched(ns, d["n"], d["rate_off"], d["vec_off"], R, l, b, None, abi.NLTE_FP64, 0.0, None)
    gpu.sync()
    return [x.cpu().numpy() for x in (Pic, Rci, lJ, b, l)], vec_off


def test_radiative_rates_match_oracle(gpu):
    """calculate_radiative_rates from a seeded J_nu (a diluted blackbody with
    noise) for every (zone, atom) system: P_ic, R_ci and line J-bar within
    1e-12 of the oracle (same summation order; device exp / pow within an ulp)."""
    import ctypes
    from pubsed_amd import runtime
    from oracle import oracle_py
    nz = 4
    Ph = runtime.Problem(runtime.Space(None), "c2_line", nz=nz, n_lines=200)
    Pg = runtime.Problem(runtime.Space("cuda:0"), "c2_line", nz=nz, n_lines=200)
    OracleBackend().opacity_build(Ph)
    A = Ph.atoms_np
    na, nl, nb = Ph.atoms.n_atoms, Ph.atoms.n_levels, Ph.atoms.n_blines
    T = Ph.get("T")
    nuc = W.grid_center(Ph.nu_np)
    rng = np.random.default_rng(9)
    H, KB, C = 6.6260755e-27, 1.380658e-16, 2.99792458e10
    with np.errstate(over="ignore"):
        J_nu = np.array([0.5 * 2 * H * nuc ** 3 / C ** 2 / np.expm1(H * nuc / KB / (1.3 * T[z]))
                         * rng.uniform(0.5, 1.5, len(nuc)) for z in range(nz)])
    sys_atom = np.array([a for z in range(nz) for a in range(na)], np.int32)
    sys_zone = np.array([z for z in range(nz) for a in range(na)], np.int32)
    ns = len(sys_atom)
    lib = oracle_py.load()
    Pic_h, Rci_h, lJ_h = np.zeros((ns, nl)), np.zeros((ns, nl)), np.zeros((ns, nb))
    for s in range(ns):
        a, z = int(sys_atom[s]), int(sys_zone[s])
        pi_, rc_, lj_ = np.zeros(nl), np.zeros(nl), np.zeros(nb)
        lib.or_nlte_radiative_rates(ctypes.byref(Ph.atoms), a, ctypes.byref(Ph.nu),
                                    np.ascontiguousarray(J_nu[z]).ctypes.data, T[z], int(Ph.ej["elem_A"][a]), 0.0,
                                    pi_.ctypes.data, rc_.ctypes.data, lj_.ctypes.data)
        l0, l1 = A["lev_off"][a], A["lev_off"][a + 1]
        Pic_h[s, l0:l1], Rci_h[s, l0:l1] = pi_[l0:l1], rc_[l0:l1]
        mine = (A["bline_l"] >= l0) & (A["bline_l"] < l1)
        lJ_h[s, mine] = lj_[mine]
    print("case radiative rates", ns, "systems", flush=True)
    (Pic, Rci, lJ, _, _), _ = nlte_pass_gpu(gpu, Pg, sys_atom, sys_zone, T[sys_zone], Ph.get("n_elec")[sys_zone],
                                            Ph.get("lev_n").reshape(nz, nl)[sys_zone], J_nu)
    for name, g, h in (("Pic", Pic.reshape(ns, nl), Pic_h), ("Rci", Rci.reshape(ns, nl), Rci_h),
                       ("line_J", lJ.reshape(ns, nb), lJ_h)):
        m = h != 0
        assert np.array_equal(g != 0, m), name
        assert m.sum() > 0
        e = np.max(np.abs(g[m] - h[m]) / np.abs(h[m]))
        print(name, e)
        assert e <= 1e-12, (name, e)


def test_nlte_pass_in_blackbody_field_returns_lte(gpu):
    """The whole NLTE pass on the GPU -- radiative rates from J_nu = B_nu(T),
    set_rates with collisions, the batched solve -- gives the LTE populations
    back (b = 1 within 2e-2 for levels holding >= 1e-6 of their atom), on a
    grid that spans every line and photoionisation edge."""
    from pubsed_amd import runtime
    nz = 2
    over = dict(nz=nz, n_lines=200, nu=(1e11, 2e17, 8192))
    Ph = runtime.Problem(runtime.Space(None), "c2_line", **over)
    Pg = runtime.Problem(runtime.Space("cuda:0"), "c2_line", **over)
    OracleBackend().opacity_build(Ph)
    na, nl = Ph.atoms.n_atoms, Ph.atoms.n_levels
    T = Ph.get("T")
    nuc = W.grid_center(Ph.nu_np)
    H, KB, C = 6.6260755e-27, 1.380658e-16, 2.99792458e10
    with np.errstate(over="ignore"):
        J_nu = np.array([2 * H * nuc ** 3 / C ** 2 / np.expm1(H * nuc / KB / T[z]) for z in range(nz)])
    sys_atom = np.array([a for z in range(nz) for a in range(na)], np.int32)
    sys_zone = np.array([z for z in range(nz) for a in range(na)], np.int32)
    print("case NLTE pass in a blackbody field", flush=True)
    (_, _, _, b, l), vo = nlte_pass_gpu(gpu, Pg, sys_atom, sys_zone, T[sys_zone], Ph.get("n_elec")[sys_zone],
                                        Ph.get("lev_n").reshape(nz, nl)[sys_zone], J_nu)
    for s in range(len(sys_atom)):
        bs, ls = b[vo[s]:vo[s + 1]], l[vo[s]:vo[s + 1]]
        big = ls >= 1e-6 * ls.sum()
        assert np.max(np.abs(bs[big] - 1)) < 2e-2, (s, np.max(np.abs(bs[big] - 1)))


@pytest.mark.parametrize("precision,tol", [(abi.NLTE_FP64, 1e-10), (abi.NLTE_FP32_REFINE, 1e-8)])
def test_batched_matches_committed_lapack_fixture(gpu, precision, tol):
    """The batched solve against LAPACK's solutions of the committed fixture
    (tests/golden/nlte_lu_systems.npz); 1e-8 is the NLTE error bound of the
    mixed form."""
    import os
    import torch
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "nlte_lu_systems.npz"))
    ns = len(g["n"])
    d = {k: torch.from_numpy(np.ascontiguousarray(g[k])).cuda() for k in ("n", "rate_off", "vec_off", "rates", "lte")}
    b = torch.zeros(len(g["lte"]), dtype=torch.float64, device="cuda:0")
    it = torch.zeros(ns, dtype=torch.int32, device="cuda:0")
    print("case golden batched", precision, flush=True)
    gpu.nlte_batched(ns, d["n"], d["rate_off"], d["vec_off"], d["rates"], d["lte"], b, None, precision, 1e-13, it)
    gpu.sync()
    bh = b.cpu().numpy()
    for s in range(ns):
        sl = slice(g["vec_off"][s], g["vec_off"][s + 1])
        e = np.max(np.abs(bh[sl] - g["b"][sl])) / np.max(np.abs(g["b"][sl]))
        print(s, int(g["n"][s]), e)
        assert e <= tol, (s, e)
