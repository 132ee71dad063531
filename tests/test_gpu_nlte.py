"""GPU parity of the NLTE solves (K8) against the FP64 oracle LU.

Bars: FP64 forms (single drop-in, batched FP64) within 1e-11 normwise of
the oracle's departure coefficients; the FP32-LU + FP64-refinement form
within 1e-8 normwise, with every system converging in FP32 (iters > 0), so
the MFMA trailing update is exercised rather than the FP64 fallback.
"""
import numpy as np
import pytest

from pubsed_amd import workloads as W, abi
from oracle.oracle_py import OracleBackend

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("precision,bar", [(abi.NLTE_FP64, 1e-11), (abi.NLTE_FP32_REFINE, 1e-8)])
def test_batched(gpu, precision, bar):
    sy = W.nlte_systems(n_sys=96, seed=5)
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
        assert (it > 0).all(), "some systems fell back to FP64: %s" % it


def test_uneven_sizes_edge(gpu):
    """Tiny, panel-straddling and LDS-limit sizes in one batch (n = 2 ... 202,
    plus 230 which exceeds the LDS tile and takes the FP64 path)."""
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
    assert it[-1] == -1            # n = 230 > LDS limit: FP64 path
    assert (it[:-1] > 0).all()


def test_set_rates_solve_and_populations_match_oracle(gpu):
    """Rate assembly (set_rates) for every (zone, atom) system of a small
    line-expansion problem from its LTE state, seeded J-bar / P_ic / R_ci;
    then the batched FP64 solve and the populations.  Rates within 1e-13 of
    the oracle (device pow / exp against glibc's); populations within 1e-10
    of the atom's total (the departure-coefficient systems have condition
    numbers ~1e18, so b of levels populated at 1e-15 is set by rounding on
    either side; levels holding >= 1e-6 of the atom agree to 1e-8)."""
    import ctypes
    import torch
    from pubsed_amd import runtime, check
    from oracle import oracle_py
    nz = 6
    Ph = runtime.Problem(runtime.Space(None), "c2_line", nz=nz, n_lines=200)
    Pg = runtime.Problem(runtime.Space("cuda:0"), "c2_line", nz=nz, n_lines=200)
    OracleBackend().opacity_build(Ph)
    A = Ph.atoms_np
    na, nl, nb = Ph.atoms.n_atoms, Ph.atoms.n_levels, Ph.atoms.n_blines
    lev = Ph.get("lev_n").reshape(nz, nl)
    T, ne = Ph.get("T"), Ph.get("n_elec")
    rng = np.random.default_rng(5)
    sys_atom = np.array([a for z in range(nz) for a in range(na)], np.int32)
    sys_zone = np.array([z for z in range(nz) for a in range(na)], np.int32)
    ns = len(sys_atom)
    n = np.diff(A["lev_off"])[sys_atom].astype(np.int32)
    rate_off = np.concatenate([[0], np.cumsum(n.astype(np.int64) ** 2)])
    vec_off = np.concatenate([[0], np.cumsum(n.astype(np.int64))])
    J = 10.0 ** rng.uniform(-8, -3, (ns, nb))
    Pic = 10.0 ** rng.uniform(-4, 0, (ns, nl))
    Rci = 10.0 ** rng.uniform(-14, -10, (ns, nl))
    lte_in = lev[sys_zone]
    ne_s, T_s = ne[sys_zone].copy(), T[sys_zone].copy()
    # oracle
    lib = oracle_py.load()
    R_h = np.zeros(rate_off[-1])
    l_h = np.zeros(vec_off[-1])
    b_h = np.zeros(vec_off[-1])
    ln_h = np.zeros((ns, nl))
    fr_h = np.zeros((ns, Ph.atoms.n_ions))
    for s in range(ns):
        a = int(sys_atom[s])
        R = np.zeros(n[s] * n[s])
        l = np.zeros(n[s])
        lib.or_nlte_set_rates(ctypes.byref(Ph.atoms), a, ne_s[s], T_s[s], np.ascontiguousarray(J[s]).ctypes.data,
                              np.ascontiguousarray(Pic[s]).ctypes.data, np.ascontiguousarray(Rci[s]).ctypes.data,
                              np.ascontiguousarray(lte_in[s]).ctypes.data, 1, 1, R.ctypes.data, l.ctypes.data)
        R_h[rate_off[s]:rate_off[s + 1]] = R
        l_h[vec_off[s]:vec_off[s + 1]] = l
        b = np.zeros(n[s])
        lib.or_nlte_solve(int(n[s]), R.ctypes.data, l.ctypes.data, b.ctypes.data, None)
        b_h[vec_off[s]:vec_off[s + 1]] = b
        l0, l1 = A["lev_off"][a], A["lev_off"][a + 1]
        i0, i1 = A["ion_off"][a], A["ion_off"][a + 1]
        ln = np.zeros(n[s])
        fr = np.zeros(i1 - i0)
        lib.or_nlte_populations(ctypes.byref(Ph.atoms), a, b.ctypes.data, l.ctypes.data, ln.ctypes.data,
                                fr.ctypes.data)
        ln_h[s, l0:l1] = ln
        fr_h[s, i0:i1] = fr
    # GPU
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    d = {k: dev(v) for k, v in dict(sys_atom=sys_atom, ne=ne_s, T=T_s, J=J, Pic=Pic, Rci=Rci, lte_in=lte_in,
                                      n=n, rate_off=rate_off, vec_off=vec_off).items()}
    R_g = torch.zeros(int(rate_off[-1]), dtype=torch.float64, device="cuda:0")
    l_g = torch.zeros(int(vec_off[-1]), dtype=torch.float64, device="cuda:0")
    b_g = torch.zeros_like(l_g)
    ln_g = torch.zeros((ns, nl), dtype=torch.float64, device="cuda:0")
    fr_g = torch.zeros((ns, Ph.atoms.n_ions), dtype=torch.float64, device="cuda:0")
    p = abi.ptr
    print("case set_rates", ns, "systems", flush=True)
    check(gpu.lib.sgpu_nlte_set_rates(gpu.ctx, ctypes.byref(Pg.atoms), ns, p(d["sys_atom"]), p(d["ne"]),
                                      p(d["T"]), p(d["J"]), p(d["Pic"]), p(d["Rci"]), p(d["lte_in"]), 1, 1,
                                      p(d["n"]), p(d["rate_off"]), p(d["vec_off"]), p(R_g), p(l_g)),
          "sgpu_nlte_set_rates")
    gpu.sync()
    Rg = R_g.cpu().numpy()
    nzm = R_h != 0
    assert np.array_equal(Rg != 0, nzm)
    assert np.max(np.abs(Rg[nzm] - R_h[nzm]) / np.abs(R_h[nzm])) <= 1e-13
    assert np.array_equal(l_g.cpu().numpy(), l_h)
    print("case batched solve + populations", flush=True)
    gpu.nlte_batched(ns, d["n"], d["rate_off"], d["vec_off"], R_g, l_g, b_g, None, abi.NLTE_FP64, 0.0, None)
    check(gpu.lib.sgpu_nlte_populations(gpu.ctx, ctypes.byref(Pg.atoms), ns, p(d["sys_atom"]), p(d["vec_off"]),
                                        p(b_g), p(l_g), p(ln_g), p(fr_g)), "sgpu_nlte_populations")
    gpu.sync()
    lng, frg = ln_g.cpu().numpy(), fr_g.cpu().numpy()
    bg = b_g.cpu().numpy()
    for s in range(ns):
        a = int(sys_atom[s])
        l0, l1 = A["lev_off"][a], A["lev_off"][a + 1]
        tot = ln_h[s, l0:l1].sum()
        assert np.max(np.abs(lng[s, l0:l1] - ln_h[s, l0:l1])) <= 1e-10 * tot, s
        i0, i1 = A["ion_off"][a], A["ion_off"][a + 1]
        assert np.max(np.abs(frg[s, i0:i1] - fr_h[s, i0:i1])) <= 1e-10 * tot, s
        lv = l_h[vec_off[s]:vec_off[s + 1]]
        big = lv >= 1e-6 * lv.sum()
        bh = b_h[vec_off[s]:vec_off[s + 1]]
        assert np.max(np.abs(bg[vec_off[s]:vec_off[s + 1]][big] - bh[big]) / np.abs(bh[big])) <= 1e-8, s
