"""Writes tests/golden/nlte_lu_systems.npz: eight seeded synthetic NLTE rate
systems of the batched-solve workload (sizes 20..64, the rate form
AtomicSpecies::set_rates leaves, AtomicSpecies_state.cpp:429-431), the
matrix solve_nlte assembles from them (:146-164: M_ii = -sum_j R_ij,
M_ij = R_ji, last row the LTE populations, right-hand side e_{n-1}) and its
solution by LAPACK's partial-pivot LU (scipy.linalg.lu_factor / lu_solve,
getrf / getrs) -- an independent FP64 solver for the departure coefficients.

Run from the repository root: python tests/golden/make_nlte_golden.py
"""
import os
import sys

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from pubsed_amd import workloads as W   # noqa: E402


def main():
    sy = W.nlte_systems(n_sys=8, n_min=20, n_max=64, seed=424242)
    b = np.zeros_like(sy["lte"])
    cond = np.zeros(sy["n_sys"])
    for s in range(sy["n_sys"]):
        n = int(sy["n"][s])
        R = sy["rates"][sy["rate_off"][s]:sy["rate_off"][s + 1]].reshape(n, n)
        lte = sy["lte"][sy["vec_off"][s]:sy["vec_off"][s + 1]]
        M = R.T.copy()
        np.fill_diagonal(M, -R.sum(axis=1))
        M[n - 1, :] = lte
        e = np.zeros(n); e[n - 1] = 1.0
        b[sy["vec_off"][s]:sy["vec_off"][s + 1]] = scipy.linalg.lu_solve(scipy.linalg.lu_factor(M), e)
        cond[s] = np.linalg.cond(M)
    np.savez_compressed(os.path.join(HERE, "nlte_lu_systems.npz"), n=sy["n"], rate_off=sy["rate_off"],
                        vec_off=sy["vec_off"], rates=sy["rates"], lte=sy["lte"], b=b, cond=cond)
    print("wrote", sy["n"].tolist(), "cond", ["%.2g" % c for c in cond])


if __name__ == "__main__":
    main()
