"""Empty inputs and rejected arguments through the C-ABI.  Every rejected
call is refused on the host before any kernel launches (a negative status
and a message from sgpu_last_error); empty calls return at once and leave
the data untouched."""
import numpy as np
import pytest
import torch

from pubsed_amd import runtime, abi, SgpuError, last_error

pytestmark = pytest.mark.gpu


@pytest.fixture()
def small(gpu):
    P = runtime.Problem(runtime.Space("cuda:0"), "c1_grey", nz=16)
    P.alloc_packets(100)
    scr = P.alloc_packets(100, name="scr")
    gpu.opacity_build(P)
    gpu.wipe(P)
    return P, scr


def test_empty_calls(gpu, small):
    P, scr = small
    print("case empty calls", flush=True)
    before = P.packets_numpy(100)
    assert gpu.emit(P, 0) == 0
    gpu.propagate(P, 0)
    assert gpu.compact(P, 0, scr) == (0, 0)
    gpu.sync()
    after = P.packets_numpy(100)
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    assert np.all(P.get("e_abs") == 0) and np.all(P.get("J_nu") == 0)
    # an empty batch of NLTE systems
    z = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    zd = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    zo = torch.zeros(1, dtype=torch.int64, device="cuda:0")
    gpu.nlte_batched(0, z, zo, zo, zd, zd, zd, zd, abi.NLTE_FP32_REFINE, 1e-10, z)


def test_rejected_arguments(gpu, small):
    P, scr = small
    print("case rejected arguments", flush=True)
    with pytest.raises(SgpuError, match="capacity"):
        gpu.propagate(P, 101)
    with pytest.raises(SgpuError, match="capacity"):
        gpu.emit(P, 60, offset=50)
    with pytest.raises(SgpuError, match="capacity"):
        gpu.compact(P, 101, scr)
    P.tparams.precision = 7
    with pytest.raises(SgpuError, match="precision"):
        gpu.propagate(P, 10)
    P.tparams.precision = abi.PREC_MIXED
    P.tparams.use_ddmc = 1
    with pytest.raises(SgpuError):
        gpu.propagate(P, 10)
    P.tparams.use_ddmc = 0
    assert "" != last_error()
    # the context still works after the refusals
    got = gpu.emit(P, 50)
    gpu.propagate(P, got)
    gpu.sync()
    assert got == 50
    fate = P.packets_numpy(50)["fate"]
    assert np.all(fate != abi.MOVING)
