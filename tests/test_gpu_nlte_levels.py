"""GPU parity of the NLTE-coupled opacity build (sgpu_opacity_build_nlte) on
the reference's few-level hydrogen problems (tests/3level_nlte and
tests/2level_nlte; the problems and their closed forms are described in
test_oracle_nlte_levels.py).  The GPU's level-population ratios match the
closed forms evaluated with the oracle's radiative rates and the oracle's
own ratios to 1e-13 (measured 2.9e-15 and 1.3e-15: the GPU sums the
photoionisation and line integrals in another order), n_e the oracle's to
1e-10 (measured equal), and the committed analytic three-level table within
the same bars as the oracle."""
import numpy as np
import pytest
import torch

from pubsed_amd import runtime, abi
from oracle.oracle_py import OracleBackend
from test_oracle_nlte_levels import (THREE, TWO, R_P, problem, three_level_table, two_level_radii,
                                     closed_form_three, closed_form_two, two_level_formula, level_ratios)

pytestmark = pytest.mark.gpu


def build_both(gpu, spec, r, prec):
    Pg, J = problem(runtime.Space("cuda:0"), spec, r)
    Ph, _ = problem(runtime.Space(None), spec, r)
    st, ev = gpu.opacity_build_nlte(Pg, torch.from_numpy(J).cuda(), (1,), use_collisions=0, lu_precision=prec)
    gpu.sync()
    sh = OracleBackend().opacity_build_nlte(Ph, J, np.ones(1, np.int32), use_collisions=0)
    print("status", st, sh, "evaluations", ev, flush=True)
    assert st == sh == 0 and ev > 2
    e = np.abs(Pg.get("n_elec") / Ph.get("n_elec") - 1).max()
    print("n_elec", e)
    assert e <= 1e-10
    return Pg, Ph, J


@pytest.mark.parametrize("prec", [abi.NLTE_FP64, abi.NLTE_FP32_REFINE])
def test_three_level_atom(gpu, prec):
    tab = three_level_table()
    nz = len(tab)
    print("case three-level NLTE build", prec, nz, "zones", flush=True)
    Pg, Ph, J = build_both(gpu, THREE, tab[:, 0], prec)
    g12, g13 = level_ratios(Pg, nz, (1, 2))
    h12, h13 = level_ratios(Ph, nz, (1, 2))
    worst_o = max(np.abs(g12 / h12 - 1).max(), np.abs(g13 / h13 - 1).max())
    worst_c = 0.0
    for z in range(nz):
        c12, c13 = closed_form_three(Ph, J, z)
        worst_c = max(worst_c, abs(g12[z] / c12 - 1), abs(g13[z] / c13 - 1))
    print("vs oracle", worst_o, "vs closed form", worst_c)
    assert worst_o <= 1e-13 and worst_c <= 1e-13
    outer = tab[:, 0] >= 2 * R_P
    for got, want in ((g12, tab[:, 1]), (g13, tab[:, 2])):
        e = np.abs(got / want - 1)
        print("vs analytic table: max", e.max(), "mean", e.mean())
        assert e.max() <= 0.05 and e.mean() <= 0.01 and e[outer].max() <= 0.01


def test_two_level_atom(gpu):
    r = two_level_radii()
    print("case two-level NLTE build", len(r), "zones", flush=True)
    Pg, Ph, J = build_both(gpu, TWO, r, abi.NLTE_FP64)
    (g12,) = level_ratios(Pg, len(r), (1,))
    (h12,) = level_ratios(Ph, len(r), (1,))
    worst_c = max(abs(g12[z] / closed_form_two(Ph, J, z) - 1) for z in range(len(r)))
    e = np.abs(g12 / two_level_formula(r) - 1).max()
    print("vs oracle", np.abs(g12 / h12 - 1).max(), "vs closed form", worst_c, "vs run_test formula", e)
    assert np.abs(g12 / h12 - 1).max() <= 1e-13 and worst_c <= 1e-13 and e <= 5e-4
