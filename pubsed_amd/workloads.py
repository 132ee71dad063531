"""Seeded synthetic inputs for the transport hot path (numpy, host side).

Every generator is deterministic in its seed and returns plain numpy arrays
laid out exactly as the C-ABI descriptors of include/sedona_gpu.h expect.
The workloads:

  ejecta_1d     homologous 1D ejecta: nz equal-velocity shells to v_max at
                t, rho ~ exp(-v/v_e) normalised to M_ej, 56Ni inside M_ni
  nu_log_grid   locate_array::log_init replayed bin by bin
  synthetic_atoms  per-element synthetic ion stages / levels and a fuzz-line
                list (lambda log-uniform, log gf uniform, E_l exponential)
  nlte_systems  uneven dense rate matrices for the batched NLTE solve
  ejecta_2d     2D (varpi, z) homologous ellipsoidal outflow
"""
import math
import numpy as np

C_LIGHT = 2.99792458e10
M_SUN = 1.98892e33
M_P = 1.67262158e-24
THOMSON = 0.66523e-24
PI_REF = 3.14159          # the reference's truncated pi (physical_constants.h:9)
DAY = 86400.0
K_EV = 8.6173324e-5


def log_init(start, stop, dl):
    """Right edges of locate_array::log_init (locate_array.cpp:54-80)."""
    x = []
    v = start + start * dl
    while v < stop:
        x.append(v)
        v = v + v * dl
    x.append(stop)
    return np.array(x, dtype=np.float64)


def lin_init_n(start, stop, n):
    """locate_array::init(start, stop, int n) (locate_array.cpp:88-108)."""
    if start == stop or n == 0:
        return dict(n=1, type=3, min=-math.inf, del_=0.0, x=np.array([math.inf]))
    dl = (stop - start) / float(n)
    x = np.array([start + (i + 1) * dl for i in range(n)], dtype=np.float64)
    x[n - 1] = stop
    return dict(n=n, type=1, min=start, del_=dl, x=x)


def lin_init_del(start, stop, dl):
    """locate_array::init(start, stop, double del) (locate_array.cpp:25-46)."""
    if start >= stop:
        return dict(n=1, type=3, min=-math.inf, del_=0.0, x=np.array([math.inf]))
    n = max(int(math.ceil((stop - start) / dl)), 1)
    x = np.array([start + (i + 1) * dl for i in range(n)], dtype=np.float64)
    x[n - 1] = stop
    return dict(n=n, type=1, min=start, del_=dl, x=x)


def nu_log_grid(nu_min, nu_max, n_bins):
    """A log grid with exactly n_bins bins: del = (max/min)^(1/n)-1, nudged up
    by 1e-12 relative so that the repeated-multiplication edges do not leave
    an extra sliver bin before `stop`."""
    dl = (nu_max / nu_min) ** (1.0 / n_bins) - 1.0
    for k in range(0, 40):
        d = dl * (1.0 + 1e-12 * (k + 1))
        x = log_init(nu_min, nu_max, d)
        if len(x) == n_bins:
            return dict(n=n_bins, type=2, min=nu_min, del_=d, x=x)
    raise ValueError("could not build a %d-bin log grid" % n_bins)


def grid_center(g):
    x = g["x"]
    left = np.concatenate([[g["min"]], x[:-1]])
    return 0.5 * (left + x)


def grid_delta(g):
    x = g["x"]
    left = np.concatenate([[g["min"]], x[:-1]])
    return x - left


def ejecta_1d(nz=100, v_max=2e9, t=20 * DAY, v_e=3e8, m_ej=1.4, m_ni=0.6,
              T_in=1.5e4, T_out=4e3, elements=((28, 56), (27, 56), (26, 56), (14, 28))):
    """1D homologous shells.  Composition: X(56Ni)=1 where the mass enclosed
    at the shell centre is <= m_ni Msun, else the last element (inert)."""
    v_out = v_max * np.arange(1, nz + 1) / nz
    v_c = v_out - 0.5 * v_max / nz
    r_out = v_out * t
    r0 = np.concatenate([[0.0], r_out[:-1]])
    vol = 4.0 * PI_REF / 3.0 * (r_out ** 3 - r0 ** 3)
    rho = np.exp(-v_c / v_e)
    rho *= m_ej * M_SUN / np.sum(rho * vol)
    m_enc = np.cumsum(rho * vol) - 0.5 * rho * vol
    ne = len(elements)
    X = np.zeros((nz, ne))
    ni = m_enc <= m_ni * M_SUN
    X[ni, 0] = 1.0
    X[~ni, ne - 1] = 1.0
    T = T_in + (T_out - T_in) * (v_c / v_max)
    return dict(kind=1, nz=nz, t=t, r_out=r_out, r_min=0.0, vol=vol, rho=rho, T=T, X=X,
                elem_Z=np.array([e[0] for e in elements], np.int32),
                elem_A=np.array([e[1] for e in elements], np.int32), v_c=v_c)


def ejecta_2d(nx=128, nzz=256, v_max=2e9, t=20 * DAY, v_e=3e8, m_ej=1.4, m_ni=0.6,
              axis_ratio=2.0, T_in=1.5e4, T_out=4e3, elements=((28, 56), (14, 28))):
    """2D (varpi, z) grid: varpi in [0, v_max t], z in [-v_max t, v_max t];
    density exp(-v_ell/v_e) in the ellipsoidal velocity v_ell =
    sqrt(v_varpi^2 + (v_z/axis_ratio)^2) (elongated along z), zero outside
    v_ell > v_max; 56Ni where v_ell is inside the m_ni mass coordinate."""
    xe = v_max * t * np.arange(1, nx + 1) / nx
    ze = -v_max * t + 2 * v_max * t * np.arange(1, nzz + 1) / nzz
    dx, dz = v_max * t / nx, 2 * v_max * t / nzz
    xc = xe - 0.5 * dx
    zc = ze - 0.5 * dz
    x0 = np.concatenate([[0.0], xe[:-1]])
    vol = (PI_REF * (xe ** 2 - x0 ** 2))[:, None] * dz * np.ones(nzz)[None, :]
    vx = xc[:, None] / t
    vz = zc[None, :] / t
    vell = np.sqrt(vx ** 2 + (vz / axis_ratio) ** 2)
    rho = np.exp(-vell / v_e)
    rho[vell > v_max] = 1e-6 * np.exp(-v_max / v_e)
    rho *= m_ej * M_SUN / np.sum(rho * vol)
    order = np.argsort(vell.ravel())
    m_sorted = np.cumsum((rho * vol).ravel()[order])
    ni_flat = np.zeros(nx * nzz, bool)
    ni_flat[order[m_sorted <= m_ni * M_SUN]] = True
    ne = len(elements)
    X = np.zeros((nx * nzz, ne))
    X[ni_flat, 0] = 1.0
    X[~ni_flat, ne - 1] = 1.0
    T = T_in + (T_out - T_in) * np.clip(vell.ravel() / v_max, 0, 1)
    return dict(kind=2, nz=nx * nzz, nx=nx, nzz=nzz, t=t, x_out=xe, z_out=ze, x_min=0.0,
                z_min=-v_max * t, vol=vol.ravel(), rho=rho.ravel(), T=T, X=X,
                elem_Z=np.array([e[0] for e in elements], np.int32),
                elem_A=np.array([e[1] for e in elements], np.int32))


# ionisation energies (eV) of the first four stages used for the synthetic atoms
ION_CHI = {28: [7.64, 18.17, 35.19, 54.9], 14: [8.15, 16.35, 33.49, 45.14],
           26: [7.90, 16.19, 30.65, 54.8], 27: [7.88, 17.08, 33.50, 51.3]}


def synthetic_atoms(nu_grid, elem_Z, n_lines=100000, n_ions=4, levels_per_ion=12,
                    seed=20261203, lam_min=1000.0, lam_max=10000.0):
    """Synthetic atomic data for each element: n_ions stages with
    levels_per_ion levels (E_0 = 0, others sorted uniform below 0.8 chi),
    g uniform on 1..9, plus the fully ionised stage the reference appends
    (one level, g = 1, chi = 99999; AtomicData.cpp:429-446).  Fuzz lines:
    lambda log-uniform on [lam_min, lam_max] Angstrom, log10 gf ~ U[-4, 0],
    E_l ~ Exp(mean 2 eV), element and ion uniform; lines outside the open
    frequency range of the grid are dropped (AtomicData.cpp:842-844)."""
    rng = np.random.default_rng(seed)
    lev_E, lev_g, lev_ion, ion_chi = [], [], [], []
    lev_off, ion_off = [0], [0]
    for Z in elem_Z:
        chis = ION_CHI.get(int(Z), ION_CHI[26])[:n_ions]
        for s in range(n_ions):
            E = np.sort(rng.uniform(0.0, 0.8 * chis[s], levels_per_ion - 1))
            lev_E += [0.0] + list(E)
            lev_g += list(rng.integers(1, 10, levels_per_ion))
            lev_ion += [s] * levels_per_ion
            ion_chi.append(chis[s])
        lev_E.append(0.0)
        lev_g.append(1)
        lev_ion.append(n_ions)
        ion_chi.append(99999.0)
        lev_off.append(len(lev_E))
        ion_off.append(len(ion_chi))
    n_atoms = len(elem_Z)
    lam = np.exp(rng.uniform(math.log(lam_min), math.log(lam_max), n_lines))
    nu = C_LIGHT / (lam * 1e-8)
    gf = 10.0 ** rng.uniform(-4.0, 0.0, n_lines)
    El = rng.exponential(2.0, n_lines)
    atom = rng.integers(0, n_atoms, n_lines)
    ion = rng.integers(0, n_ions, n_lines)
    keep = (nu > nu_grid["min"]) & (nu < nu_grid["x"][-1])
    nu, gf, El, atom, ion = nu[keep], gf[keep], El[keep], atom[keep], ion[keep]
    gx = nu_grid["x"]
    bins = np.searchsorted(gx, nu, side="right")        # upper_bound == locate (flex)
    bins = np.minimum(bins, len(gx) - 1)
    order = np.lexsort((np.arange(len(nu)), bins, atom))
    nu, gf, El, atom, ion, bins = (a[order] for a in (nu, gf, El, atom, ion, bins))
    nn = nu_grid["n"]
    off = np.zeros((n_atoms, nn + 1), np.int64)
    for a in range(n_atoms):
        sel = atom == a
        counts = np.bincount(bins[sel], minlength=nn)
        start = np.searchsorted(atom, a, side="left")
        off[a, 0] = start
        off[a, 1:] = start + np.cumsum(counts)
    bl = level_lines(nu_grid, np.array(lev_off), np.array(lev_E), np.array(lev_g), np.array(lev_ion),
                     rng)
    return dict(**bl, n_atoms=n_atoms, lev_off=np.array(lev_off, np.int32),
                ion_off=np.array(ion_off, np.int32), lev_E=np.array(lev_E, np.float64),
                lev_g=np.array(lev_g, np.int32), lev_ion=np.array(lev_ion, np.int32),
                ion_chi=np.array(ion_chi, np.float64),
                atom_zero_eps=np.zeros(n_atoms, np.int32),
                max_levels=int(np.max(np.diff(lev_off))),
                line_nu=nu.astype(np.float64), line_gf=gf.astype(np.float64),
                line_El=El.astype(np.float64), line_ion=ion.astype(np.int32),
                line_bin=bins.astype(np.int32), line_bin_off=off.ravel(),
                line_atom=atom.astype(np.int32))


def hydrogen_atom(nu_grid, levels, A, chi=13.6):
    """One hydrogen-like atom: bound levels (E in eV, g) of the neutral stage
    plus the continuum (the ionised stage, one level, g = 1), ionisation
    potential chi eV, Einstein A_ul of every bound-bound pair in (lower,
    upper) order, no fuzz lines.  levels = ((0, 2), (10.20, 8), (12.09, 18))
    with A = (4.696e8, 5.572e7, 4.408e7) is the model atom of the reference's
    three-level NLTE test (tests/3level_nlte/problem_writeup.pdf, "Test
    problem"); the first two levels with A = (4.7e8,) the two-level atom of
    tests/2level_nlte/run_test.py:70-72."""
    lev_E = np.array([e for e, _ in levels] + [0.0], np.float64)
    lev_g = np.array([g for _, g in levels] + [1], np.int32)
    lev_ion = np.array([0] * len(levels) + [1], np.int32)
    lev_off = np.array([0, len(lev_E)], np.int32)
    bl = level_lines(nu_grid, lev_off, lev_E, lev_g, lev_ion, None, A=A)
    nn = nu_grid["n"]
    return dict(**bl, n_atoms=1, lev_off=lev_off, ion_off=np.array([0, 2], np.int32), lev_E=lev_E,
                lev_g=lev_g, lev_ion=lev_ion, ion_chi=np.array([chi, 99999.0]),
                atom_zero_eps=np.zeros(1, np.int32), max_levels=len(lev_E),
                line_nu=np.zeros(0), line_gf=np.zeros(0), line_El=np.zeros(0),
                line_ion=np.zeros(0, np.int32), line_bin=np.zeros(0, np.int32),
                line_bin_off=np.zeros(nn + 1, np.int64), line_atom=np.zeros(0, np.int32))


# physical_constants.h conversion factors used by the derived line data
EV_TO_ERGS = 1.60217646e-12
CM_TO_ANGS = 1.0e8
ANGS_TO_CM = 1.0e-8
H_PLANCK = 6.6260755e-27
SIGMA_TOT = 0.0265400193567


def level_lines(nu_grid, lev_off, lev_E, lev_g, lev_ion, rng, A_min=1e4, A_max=1e8, A=None, pairs=None):
    """Level-resolved lines between every pair of levels of one ion with
    E_u > E_l, Einstein A_ul log-uniform on [A_min, A_max] s^-1 (or A, one
    value per pair in (atom, lower level, upper level) order) -- or the
    lines `pairs` = (lower, upper, atom) given with their A -- and the
    quantities AtomicData.cpp:349-375 derives from them, in its operation
    order: nu = dE ev_to_ergs / h, B_ul = A c^2 / (2 h nu^3),
    B_lu = B_ul g_u / g_l, lam = c / nu * cm_to_angs,
    f_lu = (lam angs_to_cm)^2 A g_u / g_l / (8 pi sigma_tot) with pi = 3.14159,
    bin = locate_within_bounds(nu) (lines off the grid go to its first or
    last bin).  Sorted by (atom, bin) keeping each atom's line order."""
    n_atoms = len(lev_off) - 1
    if pairs is not None:
        ll, lu, at = (np.asarray(x, np.int32) for x in pairs)
    else:
        ll, lu, at = [], [], []
        for a in range(n_atoms):
            for l in range(lev_off[a], lev_off[a + 1]):
                for u in range(l + 1, lev_off[a + 1]):
                    if lev_ion[u] == lev_ion[l] and lev_E[u] > lev_E[l]:
                        ll.append(l); lu.append(u); at.append(a)
        ll, lu, at = np.array(ll, np.int32), np.array(lu, np.int32), np.array(at, np.int32)
    if A is None:
        A = 10.0 ** rng.uniform(math.log10(A_min), math.log10(A_max), len(ll))
    else:
        A = np.asarray(A, np.float64)
        assert A.shape == ll.shape
    dE = lev_E[lu] - lev_E[ll]
    nu = dE * EV_TO_ERGS / H_PLANCK
    gl, gu = lev_g[ll].astype(np.float64), lev_g[lu].astype(np.float64)
    Bul = A * C_LIGHT * C_LIGHT / 2.0 / H_PLANCK / nu / nu / nu
    Blu = Bul * gu / gl
    lam = C_LIGHT / nu * CM_TO_ANGS
    lam_cm = lam * ANGS_TO_CM
    f = lam_cm * lam_cm * A * gu / gl / (8 * PI_REF * SIGMA_TOT)
    gx = nu_grid["x"]
    nn = nu_grid["n"]
    bins = np.minimum(np.searchsorted(gx, nu, side="right"), nn - 1)
    order = np.lexsort((np.arange(len(nu)), bins, at))
    ll, lu, at, A, nu, f, Bul, Blu, bins = (x[order] for x in (ll, lu, at, A, nu, f, Bul, Blu, bins))
    off = np.zeros((n_atoms, nn + 1), np.int64)
    for a in range(n_atoms):
        counts = np.bincount(bins[at == a], minlength=nn)
        start = np.searchsorted(at, a, side="left")
        off[a, 0] = start
        off[a, 1:] = start + np.cumsum(counts)
    return dict(bline_l=ll, bline_u=lu, bline_nu=nu, bline_f=f, bline_A=A, bline_Bul=Bul,
                bline_Blu=Blu, bline_bin=bins.astype(np.int32), bline_bin_off=off.ravel(),
                bline_atom=at)


def nlte_systems(n_sys=1024, n_min=20, n_max=200, T=1.0e4, seed=20261203,
                 coll_median=1.0e2, coll_sigma_dex=2.0, A_min=1e4, A_max=1e8):
    """Uneven dense rate systems in the form AtomicSpecies::set_rates leaves
    them (rates[i][j] from level i to j, multiplied by the LTE population of
    level i, AtomicSpecies_state.cpp:429-431).  Levels: E_0 = 0, the rest
    sorted uniform on [0, 12] eV, g on 1..9; the last level is the
    conservation row's level.  Collisional rates lognormal (median
    coll_median s^-1, sigma coll_sigma_dex dex) on every ordered pair;
    radiative Einstein A uniform on [A_min, A_max] s^-1 on every downward pair.
    Returns sizes, CSR offsets, rates and lte (flattened, FP64)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_min, n_max + 1, n_sys).astype(np.int32)
    rate_off = np.zeros(n_sys + 1, np.int64)
    vec_off = np.zeros(n_sys + 1, np.int64)
    rate_off[1:] = np.cumsum(ns.astype(np.int64) ** 2)
    vec_off[1:] = np.cumsum(ns.astype(np.int64))
    rates = np.empty(rate_off[-1], np.float64)
    lte = np.empty(vec_off[-1], np.float64)
    for s, n in enumerate(ns):
        E = np.concatenate([[0.0], np.sort(rng.uniform(0.0, 12.0, n - 1))])
        g = rng.integers(1, 10, n).astype(np.float64)
        w = g * np.exp(-E / K_EV / T)
        l = np.maximum(w / w.sum(), 1e-30)
        R = 10.0 ** (math.log10(coll_median) + coll_sigma_dex * rng.standard_normal((n, n)))
        A = rng.uniform(A_min, A_max, (n, n))
        R += np.tril(A, -1)          # downward (i > j) radiative decays
        np.fill_diagonal(R, 0.0)
        R *= l[:, None]
        rates[rate_off[s]:rate_off[s + 1]] = R.ravel()
        lte[vec_off[s]:vec_off[s + 1]] = l
    return dict(n_sys=n_sys, n=ns, rate_off=rate_off, vec_off=vec_off, rates=rates, lte=lte)


def nlte_systems_coupled(n_sys=256, n_ions=4, n_min=20, n_max=200, T=1.0e4, seed=20261204,
                         chi=(7.9, 16.2, 30.7, 54.8), coll_median=1.0e2, coll_sigma_dex=2.0,
                         A_min=1e4, A_max=1e8, bf_median=1.0e3):
    """One rate system per atom of a zone, all its ion stages coupled, as
    AtomicSpecies::solve_nlte solves them (AtomicSpecies_state.cpp:131-196
    over the levels of every ion of the atom; set_rates :294-432): n_ions
    stages of U{n_min..n_max} levels each plus the bare nucleus, n up to
    n_ions n_max + 1.  Within a stage the bound-bound rates of
    nlte_systems (lognormal collisions on every ordered pair, Einstein A on
    every downward pair); between stages the bound-free pair of
    set_rates: each level to the ground level of the next stage
    (photoionisation, lognormal about bf_median s^-1) and back
    (recombination, lognormal about 1e-2 bf_median).  Levels: stage k's
    ground at the sum of the first k ionisation energies chi (eV), excited
    levels uniform up to 12 eV above it, g on 1..9; LTE populations
    Boltzmann within each stage times a stage fraction drawn uniform on
    [0.1, 1] and normalised; rates multiplied by the source level's LTE
    population (:429-431).  The stages' excitation energies span far less
    than the chi ladder, so the LTE vector spans the dynamic range the
    stage fractions and the 12 eV excitation give (about 1e-7)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, (n_sys, n_ions))
    ns = (sizes.sum(1) + 1).astype(np.int32)
    rate_off = np.zeros(n_sys + 1, np.int64)
    vec_off = np.zeros(n_sys + 1, np.int64)
    rate_off[1:] = np.cumsum(ns.astype(np.int64) ** 2)
    vec_off[1:] = np.cumsum(ns.astype(np.int64))
    rates = np.empty(rate_off[-1], np.float64)
    lte = np.empty(vec_off[-1], np.float64)
    for s in range(n_sys):
        n = int(ns[s])
        stage = np.concatenate([np.full(m, k) for k, m in enumerate(sizes[s])] + [[n_ions]])
        ground = np.concatenate([[0], np.cumsum(sizes[s])])     # first level of each stage
        E = np.zeros(n)
        g = np.ones(n)
        frac = rng.uniform(0.1, 1.0, n_ions + 1)
        w = np.zeros(n)
        for k in range(n_ions):
            a, b = ground[k], ground[k + 1]
            E[a:b] = np.concatenate([[0.0], np.sort(rng.uniform(0.0, 12.0, b - a - 1))])
            g[a:b] = rng.integers(1, 10, b - a)
            wk = g[a:b] * np.exp(-E[a:b] / K_EV / T)
            w[a:b] = frac[k] * wk / wk.sum()
            E[a:b] += sum(chi[:k])
        E[n - 1] = sum(chi[:n_ions])
        w[n - 1] = frac[n_ions]
        l = np.maximum(w / w.sum(), 1e-30)
        R = np.zeros((n, n))
        for k in range(n_ions):
            a, b = ground[k], ground[k + 1]
            m = b - a
            Rk = 10.0 ** (math.log10(coll_median) + coll_sigma_dex * rng.standard_normal((m, m)))
            Rk += np.tril(rng.uniform(A_min, A_max, (m, m)), -1)
            np.fill_diagonal(Rk, 0.0)
            R[a:b, a:b] = Rk
            c = ground[k + 1]                                    # next stage's ground level
            R[a:b, c] = 10.0 ** (math.log10(bf_median) + coll_sigma_dex * rng.standard_normal(m))
            R[c, a:b] = 10.0 ** (math.log10(1e-2 * bf_median) + coll_sigma_dex * rng.standard_normal(m))
        R *= l[:, None]
        rates[rate_off[s]:rate_off[s + 1]] = R.ravel()
        lte[vec_off[s]:vec_off[s + 1]] = l
    return dict(n_sys=n_sys, n=ns, rate_off=rate_off, vec_off=vec_off, rates=rates, lte=lte,
                stage_sizes=sizes.astype(np.int32))
