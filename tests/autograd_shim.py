"""torch.autograd.Function over the library's forward/backward pair (test harness only).

E2(R) = 1/2 sum D G(D; R) at a fixed density D as a function of the atomic
coordinates R (Angstrom): forward builds the system at R and calls
qk_fock_build (G(D), the gpu_get2e_ operation); backward calls qk_fock_grad
(dE2/dR, the gpu_grad_ operation) and scales it from per-bohr to per-Angstrom.
"""
import ctypes as C

import numpy as np
import torch

from quick_amd.basis import build_basis
from quick_amd.fock import FockEngine, QkOpts, _check, lib, make_opts
from quick_amd.geometry import ANGSTROM_TO_BOHR


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class TwoElectronEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz_ang, symbols, D, opts):
        b = build_basis(symbols, xyz_ang.detach().cpu().numpy())
        eng = FockEngine(b)
        eng.compute_schwarz()
        Df = np.asfortranarray(D, dtype=np.float64)
        F = np.zeros_like(Df, order="F")
        _check(lib().qk_fock_build(eng.h, _dptr(Df), _dptr(F), C.byref(opts)), "qk_fock_build")
        ctx.eng, ctx.Df, ctx.opts = eng, Df, opts
        return torch.tensor(0.5 * float(np.sum(Df * F)), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gE):
        eng = ctx.eng
        g = np.zeros(3 * eng.natom)
        _check(lib().qk_fock_grad(eng.h, _dptr(ctx.Df), None, _dptr(g), C.byref(ctx.opts)), "qk_fock_grad")
        eng.close()
        gx = torch.from_numpy(np.ascontiguousarray(g.reshape(3, eng.natom, order="F").T) * ANGSTROM_TO_BOHR)
        return gx * gE, None, None, None


def two_electron_energy(xyz_ang: torch.Tensor, symbols, D, opts: QkOpts | None = None) -> torch.Tensor:
    return TwoElectronEnergy.apply(xyz_ang, symbols, D, opts or make_opts())
