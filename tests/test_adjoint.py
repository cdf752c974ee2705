"""The adjoint of the DKD + GW step (axg_gradient): the oracle's adjoint (stored trajectory,
numpy backward pass) against central finite differences of the oracle's own FP64 band energy
(relative step 1e-4), then the GPU's adjoint (time reversal from checkpoints) against the
oracle's.  J = the GW spectrum's power summed over the bins with qlo <= qmid <= qhi."""
import numpy as np
import pytest

import oracle as O
from oracle import adjoint as OA
from tests.conftest import SQ_HALF, phys

CASE = dict(t1=1.0, dt=0.05, nsteps=12, nb=12, qlo=0.5, qhi=6.0)


def fields(N, seed=3):
    P = phys(N)
    phi, _ = O.ic_flat(N, seed, 6.0, P["dwav"])
    rng = np.random.default_rng(seed)
    pi = 0.2 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    return phi, pi


def oparams(N, kappa0, tau_cr=9.42, a_now=5.4e30):
    return dict(phys(N), potential_kappa0=kappa0, tau_cr=tau_cr, a_now=a_now)


@pytest.mark.parametrize("fix", [False, True])
def test_oracle_adjoint_matches_finite_differences(fix):
    """Each derivative against the central difference of J with relative step 1e-4 (truncation
    ~1e-8): sigma_cr^2, kappa0, the log-amplitude of (phi0, pi0) and a random direction of phi0."""
    N = 8
    P = oparams(N, 2e-3)
    phi0, pi0 = fields(N)
    c = CASE
    J, ds2, dk0, dlna, phib, _ = OA.gradient(phi0, pi0, P, c["t1"], c["dt"], c["nsteps"], c["nb"], c["qlo"], c["qhi"],
                                             fix)

    def Jf(P2=P, a=1.0, dphi=0.0):
        return OA.band_energy(a * phi0 + dphi, a * pi0, P2, c["t1"], c["dt"], c["nsteps"], c["nb"], c["qlo"], c["qhi"],
                              fix)

    assert J > 0 and abs(J - Jf()) <= 1e-14 * J
    h = 1e-4
    s2 = P["sigma_cr"] ** 2
    fd = (Jf(dict(P, sigma_cr=np.sqrt(s2 * (1 + h)))) - Jf(dict(P, sigma_cr=np.sqrt(s2 * (1 - h))))) / (2 * h * s2)
    assert abs(ds2 / fd - 1) < 1e-6
    k0 = P["potential_kappa0"]
    fd = (Jf(dict(P, potential_kappa0=k0 * (1 + h))) - Jf(dict(P, potential_kappa0=k0 * (1 - h)))) / (2 * h * k0)
    assert abs(dk0 / fd - 1) < 1e-6
    fd = (Jf(a=1 + h) - Jf(a=1 - h)) / (2 * h)
    assert abs(dlna / fd - 1) < 1e-6
    d = np.random.default_rng(9).standard_normal((N, N, N)) + 0j
    fd = (Jf(dphi=h * d) - Jf(dphi=-h * d)) / (2 * h)
    assert abs(np.sum(phib.real * d.real) / fd - 1) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("prec,N,K,fix,tol,units", [("fp64", 16, 4, False, 1e-7, "phys"),
                                                    ("fp64", 16, 5, True, 1e-7, "phys"),
                                                    ("fp64", 24, 4, False, 1e-7, "phys"),
                                                    ("fp64", 36, 5, True, 1e-7, "phys"),
                                                    ("fp32c", 32, 4, False, 2e-4, "phys"),
                                                    ("fp32c", 32, 6, False, 2e-4, "unit")])
def test_gpu_gradient_matches_oracle(prec, N, K, fix, tol, units):
    """axg_gradient against the oracle's adjoint on the same fields: J, the three parameter
    derivatives and the adjoint fields (FP64 within 1e-7: the GPU's time reversal and FFT
    rounding; FP32C storage within 2e-4).  The run leaves the fields and accumulators of a plain
    axg_advance_steps of the same steps.  With tau_cr = a_now = 1 the spectrum's prefactor is
    ~1e120: the FP32 adjoint must not overflow."""
    from axion_grav_amd import capi

    print(f"[case] gradient {prec} N={N} K={K} fix={fix} {units}", flush=True)
    c = CASE
    tc, an = (9.42, 5.4e30) if units == "phys" else (1.0, 1.0)
    P = oparams(N, 1e-3, tc, an)
    prm = capi.make_params(N, P["dspa"], SQ_HALF, num_bin_grav=c["nb"], tau_cr=tc, a_now=an,
                           potential_kappa0=1e-3, gw_fix_swap=fix,
                           precision=capi.FP64 if prec == "fp64" else capi.FP32C)
    phi0, pi0 = fields(N)
    with capi.Lattice(prm) as L:
        L.set_fields(phi0, pi0)
        f_phi, f_pi = L.get_fields()  # the values the GPU holds (FP32C rounds them)
        g = L.gradient(c["t1"], c["dt"], c["nsteps"], K, c["qlo"], c["qhi"])
        phib, pib = L.adjoint()
        e_phi, e_pi = L.get_fields()
        e_A = L.gw_state()
    with capi.Lattice(prm) as L:
        L.set_fields(phi0, pi0)
        L.advance_steps(c["t1"], c["dt"], c["nsteps"])
        a_phi, a_pi = L.get_fields()
        a_A = L.gw_state()
    assert np.array_equal(e_phi, a_phi) and np.array_equal(e_pi, a_pi) and np.array_equal(e_A, a_A)
    J, ds2, dk0, dlna, ophib, opib = OA.gradient(f_phi, f_pi, P, c["t1"], c["dt"], c["nsteps"], c["nb"], c["qlo"],
                                                 c["qhi"], fix)
    print(f"  J {g['J']:.6e} / {J:.6e}; ds2 {g['dJ_dsigma2']:.6e} / {ds2:.6e}; dk0 {g['dJ_dkappa0']:.6e} / "
          f"{dk0:.6e}; dlnA {g['dJ_dlnamp']:.6e} / {dlna:.6e}; fwd {g['forward_ms']:.2f} ms adj "
          f"{g['adjoint_ms']:.2f} ms")
    assert abs(g["J"] / J - 1) < tol
    assert abs(g["dJ_dsigma2"] / ds2 - 1) < tol
    assert abs(g["dJ_dkappa0"] / dk0 - 1) < tol
    assert abs(g["dJ_dlnamp"] / dlna - 1) < tol
    for a, b in ((phib, ophib), (pib, opib)):
        assert np.abs(a - b).max() < tol * np.abs(b).max()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-6), ("fp32c", 2e-3)])
def test_torch_harness_against_finite_differences(prec, tol):
    """The torch.autograd wrapper: J(A phi_seed, A pi_seed, sigma2, kappa0).backward() against
    central differences (relative step 1e-4) of the oracle's FP64 J -- the reference path -- in
    lambda (through sigma_cr^2 = lambda/4), kappa0 and the amplitude A; the FP64 GPU's own
    differences agree as well."""
    import torch

    from axion_grav_amd import capi, gw_grad

    print(f"[case] torch harness {prec}", flush=True)
    N, c = 16, CASE
    P = oparams(N, 1e-3)
    run = dict(N=N, dspa=P["dspa"], t1=c["t1"], dt=c["dt"], nsteps=c["nsteps"], K=4, qlo=c["qlo"], qhi=c["qhi"],
               params=dict(num_bin_grav=c["nb"], tau_cr=9.42, a_now=5.4e30,
                           precision=capi.FP64 if prec == "fp64" else capi.FP32C))
    phi_s, pi_s = fields(N)
    lam = torch.tensor(4 * SQ_HALF ** 2, dtype=torch.float64, requires_grad=True)
    k0 = torch.tensor(1e-3, dtype=torch.float64, requires_grad=True)
    amp = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    J = gw_grad.band_energy(amp * torch.from_numpy(phi_s), amp * torch.from_numpy(pi_s), lam / 4, k0, run)
    J.backward()
    grads = {"lambda": lam.grad.item(), "kappa0": k0.grad.item(), "A": amp.grad.item()}

    def oJ(lmb=2.0, kap=1e-3, a=1.0):
        P2 = dict(P, sigma_cr=(lmb / 4) ** 0.5, potential_kappa0=kap)
        return OA.band_energy(a * phi_s, a * pi_s, P2, c["t1"], c["dt"], c["nsteps"], c["nb"], c["qlo"], c["qhi"])

    h = 1e-4
    fd = {"lambda": (oJ(lmb=2 * (1 + h)) - oJ(lmb=2 * (1 - h))) / (4 * h),
          "kappa0": (oJ(kap=1e-3 * (1 + h)) - oJ(kap=1e-3 * (1 - h))) / (2e-3 * h),
          "A": (oJ(a=1 + h) - oJ(a=1 - h)) / (2 * h)}
    print("  torch", grads, "\n  oracle FD", fd)
    assert abs(J.item() / oJ() - 1) < tol
    for k in grads:
        assert abs(grads[k] / fd[k] - 1) < tol, k
    if prec == "fp64":
        gfd = (gw_grad.forward_only(phi_s, pi_s, 0.5 * (1 + h), 1e-3, run) -
               gw_grad.forward_only(phi_s, pi_s, 0.5 * (1 - h), 1e-3, run)) / (2 * h * 0.5) / 4
        assert abs(grads["lambda"] / gfd - 1) < 1e-5
