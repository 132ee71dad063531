"""Differentiable transport: escaping spectrum F_nu as a function of a
per-(zone, frequency bin) absorption opacity, with its adjoint as a
torch.autograd.Function.

The estimator treats absorption continuously (packet weight
exp(-sum alpha dshift ds), alpha = rho kappa) and samples only scattering, so
the packet paths do not depend on kappa; the backward pass replays each path
from its Philox counters on the GPU (sgpu_adjoint_backward) and returns
dL/dkappa exactly for the estimator.  Only the escape bin and final weight of
each packet (12 bytes) are kept between the passes.
"""
import ctypes

import numpy as np

from . import abi, check, load_library, runtime, workloads as W


class DiffTransport:
    """Holds the grid, scattering table, frozen initial packets and records."""

    def __init__(self, backend, n_packets, nz=128, seed=20261203, device="cuda:0"):
        import torch
        self.torch = torch
        self.gpu = backend
        self.lib = load_library()
        # grey Thomson scattering; a tiny absorptive fraction only shapes the
        # emission spectrum (Planck at the zone temperature)
        P = runtime.Problem(runtime.Space(device), "c1_grey", nz=nz, grey=runtime.KAPPA_T,
                            epsilon=1e-6, seed=seed)
        self.P = P
        backend.opacity_build(P)
        P.alloc_packets(n_packets)
        backend.wipe(P)
        got = backend.emit(P, n_packets, force_rproc=1)    # photon-only source
        assert got == n_packets
        self.n = n_packets
        self.nz, self.nn = P.nz, P.n_nu
        self.rec_bin = torch.zeros(n_packets, dtype=torch.int32, device=device)
        self.rec_w = torch.zeros(n_packets, dtype=torch.float64, device=device)
        self.spec = P.tallies.optical.nu_grid
        self.n_spec = P.n_nu
        backend.sync()
        # F is reported as a fraction of the emitted packet energy, so that
        # losses and their gradients stay O(1) (the C-ABI works in erg/s)
        self.scale = 1.0 / float(P.a["pk_e"][:n_packets].sum())

    def kappa_init(self, seed=7):
        rng = np.random.default_rng(seed)
        k = np.exp(rng.uniform(np.log(1e-3), np.log(1.0), (self.nz, self.nn))).astype(np.float32)
        return self.torch.from_numpy(k).to(self.P.space.device)

    def _args(self, kappa):
        P = self.P
        return (self.gpu.ctx, ctypes.byref(P.grid), ctypes.byref(P.nu), ctypes.byref(self.spec),
                abi.ptr(kappa), abi.ptr(P.a["rho"]), abi.ptr(P.a["scat"]), ctypes.byref(P.tparams),
                ctypes.byref(P.packets), self.n)

    def forward(self, kappa, record=True):
        kappa = kappa.contiguous()
        assert kappa.dtype == self.torch.float32 and kappa.numel() == self.nz * self.nn
        F = self.torch.zeros(self.n_spec, dtype=self.torch.float64, device=kappa.device)
        check(self.lib.sgpu_adjoint_forward(*self._args(kappa), abi.ptr(F),
                                            abi.ptr(self.rec_bin) if record else None,
                                            abi.ptr(self.rec_w) if record else None),
              "sgpu_adjoint_forward")
        return F * self.scale

    def backward(self, kappa, dL_dF):
        """dL/dkappa for the upstream gradient dL/dF of the scaled spectrum."""
        g = self.torch.zeros(self.nz * self.nn, dtype=self.torch.float64, device=kappa.device)
        d = (dL_dF.double() * self.scale).contiguous()
        check(self.lib.sgpu_adjoint_backward(*self._args(kappa.contiguous()), abi.ptr(d),
                                             abi.ptr(self.rec_bin), abi.ptr(self.rec_w), abi.ptr(g)),
              "sgpu_adjoint_backward")
        return g.view(self.nz, self.nn)


def spectrum_function():
    """torch.autograd.Function F = spectrum(kappa) with the adjoint backward."""
    import torch

    class Spectrum(torch.autograd.Function):
        @staticmethod
        def forward(ctx, kappa, model):
            ctx.model = model
            ctx.save_for_backward(kappa)
            return model.forward(kappa, record=True)

        @staticmethod
        def backward(ctx, dF):
            (kappa,) = ctx.saved_tensors
            g = ctx.model.backward(kappa, dF)
            return g.to(kappa.dtype).view_as(kappa), None

    return Spectrum
