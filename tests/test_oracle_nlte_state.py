"""Pins of the oracle's NLTE-coupled gas state inside the opacity build
(GasState::solve_state(J_nu) under opacity_use_nlte, GasState.cpp:208-263,
with AtomicSpecies::solve_nlte at every n_e evaluation; CPU only).

* No atom flagged, or the first step: the LTE build, bit for bit.
* J_nu = B_nu(T) on a grid spanning every line and edge: detailed balance of
  the radiative and collisional rates returns the LTE state: n_e to 2e-4,
  populations to 2e-3 weighted by population and 2e-2 for every level
  holding >= 1e-2 of its atom.  The frequency quadrature of the
  photoionisation integrals limits the Milne relation (as in the NLTE-pass
  pins), most for the far-Wien edges of the coolest zone, where a 3e-4
  ion-stage ground level sits 40% off.
* A diluted field J = W B_nu with W < 1 photoionises less than LTE: n_e
  falls below its LTE value in every zone, more for smaller W.
* The zone keeps the state of its last n_e evaluation, a trial point next
  to the returned root: charge balance sum_atoms n_atom sum_levels n_l Z_l
  = n_e to 1e-5 (a LTE-only zone state evaluated at the root itself would
  balance to rounding).
"""
import numpy as np
import pytest

from pubsed_amd import runtime, workloads as W
from oracle.oracle_py import OracleBackend

H, KB, C, MP = 6.6260755e-27, 1.380658e-16, 2.99792458e10, 1.67262158e-24
NZ = 4
OVER = dict(nz=NZ, n_lines=200, nu=(1e11, 2e17, 8192))


def problem(**extra):
    return runtime.Problem(runtime.Space(None), "c2_line", **dict(OVER, **extra))


def planck_field(P, Wd=1.0):
    T = P.get("T")
    nuc = W.grid_center(P.nu_np)
    with np.errstate(over="ignore"):
        return Wd * np.array([2 * H * nuc ** 3 / C ** 2 / np.expm1(H * nuc / KB / T[z]) for z in range(P.nz)])


@pytest.fixture(scope="module")
def lte():
    P = problem()
    assert OracleBackend().opacity_build(P) == 0
    return P


KEYS = ("abs", "scat", "emis_cdf", "n_elec", "ion_frac", "lev_n", "planck_mean")


def test_no_nlte_atom_is_the_lte_build(lte):
    P = problem()
    OracleBackend().opacity_build_nlte(P, planck_field(P, 0.1), np.zeros(P.atoms.n_atoms, np.int32))
    for k in KEYS:
        assert np.array_equal(P.get(k), lte.get(k)), k


def test_first_step_is_lte(lte):
    P = problem()
    P.op_params.first_step = 1
    L = problem()
    L.op_params.first_step = 1
    ob = OracleBackend()
    ob.opacity_build(L)
    ob.opacity_build_nlte(P, planck_field(P, 0.1), np.ones(P.atoms.n_atoms, np.int32))
    for k in KEYS:
        assert np.array_equal(P.get(k), L.get(k)), k


def test_blackbody_field_returns_lte(lte):
    P = problem()
    st = OracleBackend().opacity_build_nlte(P, planck_field(P), np.ones(P.atoms.n_atoms, np.int32))
    assert st == 0
    assert np.allclose(P.get("n_elec"), lte.get("n_elec"), rtol=2e-4, atol=0)
    nl = P.atoms.n_levels
    ln, l0 = P.get("lev_n").reshape(NZ, nl), lte.get("lev_n").reshape(NZ, nl)
    for z in range(NZ):
        assert np.abs(ln[z] - l0[z]).sum() / l0[z].sum() < 2e-3, z
        big = l0[z] >= 1e-2
        assert np.max(np.abs(ln[z][big] / l0[z][big] - 1)) < 2e-2, z
    a, a0 = P.get("abs"), lte.get("abs")
    assert np.median(np.abs(a / np.maximum(a0, 1e-300) - 1)) < 1e-3


def test_diluted_field_lowers_ionisation(lte):
    ne = {}
    for Wd in (0.1, 0.01):
        P = problem()
        OracleBackend().opacity_build_nlte(P, planck_field(P, Wd), np.ones(P.atoms.n_atoms, np.int32))
        ne[Wd] = P.get("n_elec")
    n0 = lte.get("n_elec")
    assert np.all(ne[0.1] < n0 * (1 - 1e-4))
    assert np.all(ne[0.01] <= ne[0.1] * (1 + 1e-9))
    assert np.any(ne[0.01] < ne[0.1] * (1 - 1e-2))


@pytest.mark.parametrize("flags", [(1, 1), (1, 0), (0, 1)])
def test_final_state_is_charge_balanced(flags):
    P = problem()
    P.op_params.omit_composition_decay = 1
    OracleBackend().opacity_build_nlte(P, planck_field(P, 0.05), np.array(flags, np.int32))
    nl, na = P.atoms.n_levels, P.atoms.n_atoms
    ln = P.get("lev_n").reshape(NZ, nl)
    ion = np.asarray(P.atoms_np["lev_ion"])
    off = np.asarray(P.atoms_np["lev_off"])
    rho = P.get("rho")
    X = P.get("X").reshape(NZ, na)
    Am = np.asarray(P.get("elem_A"), float)
    for z in range(NZ):
        mf = X[z] / X[z].sum()
        f = 0.0
        for a in range(na):
            nd = rho[z] * mf[a] / (Am[a] * MP)
            f += nd * np.sum(ln[z, off[a]:off[a + 1]] * ion[off[a]:off[a + 1]])
        assert abs(f / P.get("n_elec")[z] - 1) < 1e-5, (z, f, P.get("n_elec")[z])


def test_time_step_builds_the_nlte_state_from_the_previous_step():
    """runtime.time_step with nlte_atoms set: the first step is the LTE build
    (every table and tally bit for bit the same as an LTE-configured run);
    the second builds the gas state in NLTE from the J_nu the first step
    tallied, before the wipe (transport_opacity.cpp:109-137).  (CPU: the
    step plumbing, with the oracle as the backend.)"""
    over = dict(steady=0, dt=0.2 * W.DAY, radiative_eq=0)
    flags = (1, 0)
    Pa, Pb = problem(nlte_atoms=flags, **over), problem(**over)
    ob = OracleBackend()
    n_emit, dt = 3000, 0.2 * W.DAY
    for P in (Pa, Pb):
        P.alloc_packets(2 * n_emit)
        runtime.time_step(ob, P, dt, n_emit, None, 0)
    for k in KEYS + ("T",):
        assert np.array_equal(Pa.get(k), Pb.get(k)), k
    # the oracle's threads add tallies in varying order
    for k in ("J_nu", "e_abs"):
        assert np.allclose(Pa.get(k), Pb.get(k), rtol=1e-12, atol=0), k
    J = Pa.get("J_nu").copy()
    assert J.max() > 0
    runtime.time_step(ob, Pa, dt, n_emit, None, 1)
    # the same build by hand on the twin: grid moved, then NLTE from J
    ob.homologous_step(Pb, dt)
    Pb.op_params.dt = dt
    assert ob.opacity_build_nlte(Pb, J, np.array(flags, np.int32)) == 0
    for k in KEYS:
        assert np.array_equal(Pa.get(k), Pb.get(k)), k
    lte = problem(**over)
    lte.alloc_packets(2 * n_emit)
    for s in range(2):
        runtime.time_step(ob, lte, dt, n_emit, None, s)
    assert not np.array_equal(Pa.get("n_elec"), lte.get("n_elec"))
