"""The CCSD correlation energy as a differentiable function of a one-electron
perturbation, for torch (ctypes over the C ABI of librelccsd_hip.so).

    E(V) = Re Ec[f + (V + V^H) / 2]

The forward solves CCSD (rcc_cceqn) and the Lambda equations (rcc_lambda) on
the GPU and keeps the response density D (rcc_density); the backward returns
grad_V = grad_E * D, the gradient torch expects for a real function of a
complex tensor (dE = Re sum conj(grad_pq) dV_pq; D is Hermitian and
dE = tr(D dV_h) = Re tr(D dV)).  The Hermitian part is taken explicitly so
that E is defined for every complex V.
"""
from __future__ import annotations

import numpy as np
import torch

from .relccsd import RelCCSD

CLASSES = ("OOOO", "VOOO", "VVOO", "VOVO", "VOVV", "VVVV")


class PerturbedCCSD:
    """One C1 system on the GPU whose Fock matrix is f + V_h: the integrals
    are uploaded once, each evaluation resets the Fock blocks and the orbital
    energies and converges T (from the MP2 guess) and Lambda."""

    def __init__(self, sysd, maxit=80, maxdim=8, ntol=12, frozen=None):
        self.no, self.nv = int(sysd["no"]), int(sysd["nv"])
        self.r = RelCCSD(self.no, self.nv, frozen=frozen, maxdim=maxdim, maxit=maxit,
                         ntol=ntol)

        def host(x):
            return np.asarray(x.detach().cpu().numpy() if hasattr(x, "detach") else x)

        self.foo = host(sysd["FOO"]).astype(np.complex128)
        self.fvo = host(sysd["FVO"]).astype(np.complex128)
        self.fvv = host(sysd["FVV"]).astype(np.complex128)
        self.eps = host(sysd["eps"]).astype(np.float64)
        for c in CLASSES:
            self.r.set_ints(c, sysd[c])
        self.last = None

    def _set(self, Vh):
        o = self.no
        self.r.set_fock(self.foo + Vh[:o, :o].T.reshape(-1), self.fvo + Vh[o:, :o].T.reshape(-1),
                        self.fvv + Vh[o:, o:].T.reshape(-1), self.eps + np.real(np.diag(Vh)))

    def energy(self, V):
        """Converged Re Ec at f + V_h (V any complex n x n array)."""
        V = np.asarray(V, dtype=np.complex128)
        self._set(0.5 * (V + V.conj().T))
        out = self.r.cceqn()
        rms = out["rows"][-1][2]
        if not rms < 10.0 ** (-self.r.ntol):      # also catches NaN
            raise RuntimeError(f"CCSD not converged (rms {rms:.3e} after {out['nit']} iterations)")
        self.last = out
        return out["eccsd"]

    def energy_and_density(self, V):
        e = self.energy(V)
        nit, rms = self.r.lambda_solve()
        if not rms < 10.0 ** (-self.r.ntol):
            raise RuntimeError(f"Lambda not converged (rms {rms:.3e} after {nit} iterations)")
        return e, self.r.density()


class CCSDEnergy(torch.autograd.Function):
    """E = CCSDEnergy.apply(V, problem): V a complex (n x n) tensor."""

    @staticmethod
    def forward(ctx, V, problem):
        e, dens = problem.energy_and_density(V.detach().cpu().numpy())
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(dens)).to(V.device))
        return torch.tensor(e, dtype=torch.float64, device=V.device)

    @staticmethod
    def backward(ctx, grad_e):
        (dens,) = ctx.saved_tensors
        return grad_e * dens, None


def ccsd_energy(V, problem):
    return CCSDEnergy.apply(V, problem)
