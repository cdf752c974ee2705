"""torch.autograd wrapper of the GPU adjoint (axg_gradient), for verification and for use in
torch code: J = the GW spectrum's power summed over the bins with qlo <= qmid <= qhi after
`nsteps` DKD steps from (phi0, pi0), as a differentiable function of the initial fields,
sigma_cr^2 and the N_DW=2 tilt kappa0.

The lattice work runs in libaxgrav (HIP); torch only carries the tensors.  Gradients follow
torch's convention for complex inputs of a real loss, dJ/dRe + i dJ/dIm, which is what
axg_get_adjoint returns.  For V = lambda/4 (|phi|^2 - 1)^2 in these units sigma_cr^2 =
lambda/4, so dJ/dlambda = dJ/dsigma2 / 4; an initial amplitude A enters through phi0 = A phi.
"""
import numpy as np
import torch

from axion_grav_amd import capi


class GWBandEnergy(torch.autograd.Function):
    """J(phi0, pi0, sigma2, kappa0); `run` holds the lattice and schedule:
    dict(N, dspa, t1, dt, nsteps, K, qlo, qhi, and make_params keywords in `params`)."""

    @staticmethod
    def forward(ctx, phi0, pi0, sigma2, kappa0, run):
        kw = dict(run.get("params", {}))
        kw["potential_kappa0"] = float(kappa0)
        prm = capi.make_params(run["N"], run["dspa"], float(sigma2) ** 0.5, **kw)
        with capi.Lattice(prm) as L:
            L.set_fields(phi0.detach().cpu().numpy(), pi0.detach().cpu().numpy())
            g = L.gradient(run["t1"], run["dt"], run["nsteps"], run["K"], run["qlo"], run["qhi"])
            phib, pib = L.adjoint()
        ctx.save_for_backward(torch.from_numpy(phib), torch.from_numpy(pib),
                              torch.tensor(g["dJ_dsigma2"], dtype=torch.float64),
                              torch.tensor(g["dJ_dkappa0"], dtype=torch.float64))
        ctx.timing = (g["forward_ms"], g["adjoint_ms"])
        return torch.tensor(g["J"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        phib, pib, ds2, dk0 = ctx.saved_tensors
        return gJ * phib, gJ * pib, gJ * ds2, gJ * dk0, None


def band_energy(phi0, pi0, sigma2, kappa0, run):
    return GWBandEnergy.apply(phi0, pi0, sigma2, kappa0, run)


def forward_only(phi0, pi0, sigma2, kappa0, run):
    """J without the adjoint (a plain run and the spectrum), for finite differences."""
    kw = dict(run.get("params", {}))
    kw["potential_kappa0"] = float(kappa0)
    prm = capi.make_params(run["N"], run["dspa"], float(sigma2) ** 0.5, **kw)
    with capi.Lattice(prm) as L:
        L.set_fields(np.asarray(phi0), np.asarray(pi0))  # a new lattice's accumulators are zero
        L.advance_steps(run["t1"], run["dt"], run["nsteps"])
        s = L.gw_spectrum()
    band = (s["qmid"] >= run["qlo"]) & (s["qmid"] <= run["qhi"]) & (s["denom"] > 0)
    return float(s["power"][band].sum())
