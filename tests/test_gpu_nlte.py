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
