"""The FP32 form's frequency sampler: Walker alias tables of the zone
emissivity distributions (sgpu_alias_tables) reproduce the distribution
the reference samples by upper_bound over the normalised emissivity CDF
(transport::sample_photon_frequency, emission.cpp:27-46): p_k = cdf_k -
cdf_{k-1}, the last bin taking the remainder below 1, negative or NaN steps
as zero, a zone without emission sending everything to the last bin.
Reconstruction: column j gives q_j to bin j and 1 - q_j to bin alias_j, so
p_k n = q_k + sum over columns aliasing k of (1 - q_j); exact up to the FP32
thresholds (~6e-8 per column)."""
import ctypes

import numpy as np
import pytest
import torch

from pubsed_amd import load_library, check

pytestmark = pytest.mark.gpu


def cdfs(nn, rng):
    rows = []
    p = rng.uniform(0, 1, nn); rows.append(np.cumsum(p) / p.sum())                 # smooth
    p = np.full(nn, 1e-9); p[nn // 3] = 1.0; rows.append(np.cumsum(p) / p.sum())    # one line
    p = rng.uniform(0, 1, nn) * (rng.uniform(0, 1, nn) < 0.05)
    p[0] += 1e-3; rows.append(np.cumsum(p) / p.sum())                                # sparse
    c = np.cumsum(rng.uniform(0, 1, nn)); rows.append(0.7 * c / c[-1])               # last < 1
    rows.append(np.zeros(nn))                                                        # no emission
    p = np.exp(-np.linspace(0, 60, nn)); rows.append(np.cumsum(p) / p.sum())        # steep spectrum
    return np.array(rows)


def expected(c):
    nn = len(c)
    lo = np.concatenate([[0.0], c[:-1]])
    hi = c.copy()
    hi[-1] = 1.0
    p = hi - lo
    p[~(p > 0)] = 0
    if p.sum() <= 0 or c.max() <= 0:
        e = np.zeros(nn); e[-1] = 1.0
        return e
    return p / p.sum()


@pytest.mark.parametrize("nn", [1, 3, 17, 1000, 4096])
def test_alias_tables_reproduce_the_emissivity_distribution(gpu, nn):
    rng = np.random.default_rng(nn)
    C = cdfs(nn, rng)
    nz = len(C)
    d_cdf = torch.from_numpy(np.ascontiguousarray(C.ravel())).cuda()
    tab = torch.zeros(nz * nn * 2, dtype=torch.int32, device="cuda:0")
    print("case alias tables", nn, flush=True)
    check(load_library().sgpu_alias_tables(gpu.ctx, ctypes.c_void_p(d_cdf.data_ptr()), nz, nn,
                                           ctypes.c_void_p(tab.data_ptr())), "sgpu_alias_tables")
    gpu.sync()
    t = tab.cpu().numpy().reshape(nz, nn, 2)
    q = t[:, :, 0].copy().view(np.float32).astype(np.float64)
    a = t[:, :, 1]
    for z in range(nz):
        assert np.all((a[z] >= 0) & (a[z] < nn))
        assert np.all((q[z] >= 0) & (q[z] <= 1))
        rec = q[z].copy()
        np.add.at(rec, a[z], 1 - q[z])
        rec /= nn
        want = expected(C[z])
        hits = np.bincount(a[z], minlength=nn) + 1
        assert np.all(np.abs(rec - want) * nn <= 2e-7 * hits), (z, np.max(np.abs(rec - want) * nn))
