"""GPU parity of the opacity build (sgpu_opacity_build) on real atoms read
from the reference's data/atom_data.hdf5 (copied under
tests/golden/atomdata/) against the oracle: hydrogen, helium, carbon,
oxygen, silicon, iron and nickel (nickel has levels and no lines), up to
392 levels and 2,376 lines per atom, 9 ion stages, on 16 zones of the
line-expansion ejecta (4,700 - 14,300 K), every term on: Sobolev lines
between the real levels, hydrogenic bound-free from every level, free-free
and electron scattering.  Same bar as the synthetic-atom builds: 1e-10
relative per table entry, per zone mean and per population."""
import os

import numpy as np
import pytest

from pubsed_amd import runtime
from oracle.oracle_py import OracleBackend
from test_gpu_opacity import relerr, KEYS_TABLE, KEYS_ZONE, RTOL

pytestmark = pytest.mark.gpu

ATOM_FILE = os.path.join(os.path.dirname(__file__), "golden", "atomdata", "atom_data.hdf5")
ELEMS = ((1, 1), (2, 4), (6, 12), (8, 16), (14, 28), (26, 56), (28, 59))


@pytest.mark.parametrize("max_ion_stage", [9999, 3])
def test_real_atoms_opacity_build_matches_oracle(gpu, max_ion_stage):
    P = {}
    for dev in ("cuda:0", None):
        P[dev] = runtime.Problem(runtime.Space(dev), "c2_line", nz=16, elements=ELEMS, atom_file=ATOM_FILE,
                                 max_ion_stage=max_ion_stage, fuzz=0, lexp=1, bb=1, ff=1, bf=1)
        X = np.tile(np.array([0.3, 0.2, 0.1, 0.2, 0.1, 0.07, 0.03]), (P[dev].nz, 1))
        P[dev].set("X", X.ravel())
    print("case real atoms", ELEMS, "max_ion_stage", max_ion_stage, "levels", P[None].atoms.n_levels,
          "lines", P[None].atoms.n_blines, flush=True)
    sg = gpu.opacity_build(P["cuda:0"])
    gpu.sync()
    sh = OracleBackend().opacity_build(P[None])
    assert sg == sh == 0
    for k in KEYS_TABLE + KEYS_ZONE + ("n_elec", "ion_frac", "ion_part", "lev_n"):
        e = relerr(P["cuda:0"].get(k), P[None].get(k), floor=1e-300)
        print(k, e)
        assert e <= RTOL, (k, e)
    assert np.array_equal(P["cuda:0"].get("zone_status"), P[None].get("zone_status"))
