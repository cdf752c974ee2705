"""Adjoint of the FP64 oracle's DKD + GW step (TEST INFRASTRUCTURE ONLY).

The gradient of J = the sum of the GW spectrum's power over the bins with qlo <= qmid <= qhi
(grav_waves.f90:257-305) after `nsteps` steps of evolve.f90:238-339 with the accumulation of
grav_waves.f90:193-255 started from zero, with respect to sigma_cr^2, potential_kappa0 and the
log-amplitude of the initial fields.  The forward pass is the C oracle (orc_kick, orc_drift,
orc_gw_integrate); it keeps every kick's phi, so no time reversal is involved.  The backward
pass is numpy on the full spectrum (the oracle stores all N^3 modes of the accumulators):

* accumulator adjoints abar = dJ/dRe A + i dJ/dIm A = 2 w_b A with w_b = fact q_b^3 / denom_b;
* per kick, for each source component c: Dbar_c = conj(P) pdbar + P ndbar with
  pdbar = c abar11 + s abar21', ndbar = c abar12' + s abar22 (P = pol dw, c, s = cos, sin(q tau));
  with the index quirk abar21' = abar21 + abar12, abar12' = abar12 at the run's last update and
  abar21' = 0, abar12' = abar21 + abar12 before it (the update overwrites A(2,1));
* Sbar_c(x) = Re sum_k Dbar_c(k) e^{-ikx} (the adjoint of the BACKWARD transform of a real source);
* gbar_i = 4 Sbar_ii g_i + 2 sum_{j != i} Sbar_ij g_j for S_ij = 2 Re(g_i conj g_j);
* phibar += the adjoint of the central differences, f2 Lap(pibar) and f3 H(phi) pibar (H the
  symmetric Jacobian of dVdphi in (Re, Im)); pibar <- f1 pibar + h phibar for the drift before.
"""
import numpy as np

import oracle as O

THRM_SIGMA = 2.449489742783178
PI_T = 3.14159265359


def _kick_scalars(t, dt, P):
    h = 1 / t
    f0 = 1 + dt * h
    f1 = (1 - dt * h) / f0
    f2 = (dt / (P["dspa"] * P["dspa"])) / f0
    f3 = -(dt * (t * t) * (P["sigma_cr"] * P["sigma_cr"])) / f0
    thrm = 0.0 if P["initial_loop"] else THRM_SIGMA / t
    m2 = O.msquared(thrm)
    ma2 = O.maxion2(thrm, P["kappa"])
    k0 = P.get("potential_kappa0", 0.0)
    kt = k0 * (t * t) * (t * t) / (P["sigma_cr"] * P["sigma_cr"]) if k0 != 0 else 0.0
    return f1, f2, f3, m2, ma2, kt


def _pols(N, dwav):
    n = np.fft.fftfreq(N, 1.0 / N)
    nz, ny, nx = np.meshgrid(n, n, n, indexing="ij")
    qn = np.sqrt(nx * nx + ny * ny + nz * nz)
    with np.errstate(invalid="ignore", divide="ignore"):
        qh = [np.where(qn > 0, v / qn, 0.0) for v in (nx, ny, nz)]
        r = np.sqrt(qh[0] ** 2 + qh[1] ** 2)
        c = np.where(r > 0, qh[0] / np.where(r > 0, r, 1), 1.0)
        s = np.where(r > 0, qh[1] / np.where(r > 0, r, 1), 0.0)
    p = [qh[2] * c - 1j * s, qh[2] * s + 1j * c, -r + 0j]
    pol = []
    for i, j in ((0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)):
        v = p[i] * p[j]
        if i == j:
            v = v / 2
        pol.append(np.where(qn > 0, v, 0))
    return pol, qn * dwav


def _grad(phi, dspa):
    return [(np.roll(phi, -1, axis=a) - np.roll(phi, 1, axis=a)) / (2 * dspa) for a in (2, 1, 0)]


def gradient(phi0, pi0, P, t1, dt, nsteps, nb, qlo, qhi, fix_swap=False):
    """(J, dJ/dsigma_cr^2, dJ/dkappa0, dJ/dlnA, phibar0, pibar0) for the run from (phi0, pi0)."""
    N = phi0.shape[0]
    phi, pi = phi0.copy(), pi0.copy()
    A = O.new_gw_state(N)
    kick_phi, ts = [], []
    O.drift(phi, pi, dt / 2)
    t = t1 + dt / 2
    for s in range(nsteps):
        kick_phi.append(phi.copy())
        ts.append(t)
        O.kick(phi, pi, t, dt, P["dspa"], P["sigma_cr"], P["kappa"], P["initial_loop"],
               P.get("potential_kappa0", 0.0))
        O.gw_integrate(phi, A, t, dt, P["dspa"], P["dwav"], P["dv_spa"], P["sigma_cr"], fix_swap)
        dd = dt / 2 if s == nsteps - 1 else dt
        O.drift(phi, pi, dd)
        t = t + dd
    sp = O.gw_spectrum(A, nb, P["dspa"], P["dwav"], P["tau_cr"], P["a_now"], P["dv_wav"])
    band = (sp["qmid"] >= qlo) & (sp["qmid"] <= qhi) & (sp["denom"] > 0)
    J = float(sp["power"][band].sum())
    w = np.zeros(nb)
    w[band] = sp["power"][band] / np.where(sp["numer"][band] > 0, sp["numer"][band], 1) * \
        (sp["numer"][band] > 0)
    # bins of the full spectrum: the oracle's own bin rule, i = nint(q/dq + 0.5), k != 0
    pol, q = _pols(N, P["dwav"])
    dq = PI_T / P["dspa"] / nb
    b = np.floor(q / dq + 0.5 + 0.5).astype(int)
    wk = np.where((q > 0) & (b >= 1) & (b <= nb), w[np.clip(b - 1, 0, nb - 1)], 0.0)
    abar = [2 * wk * A[m] for m in range(4)]  # m: (1,1) (2,1) (1,2) (2,2)
    s2 = P["sigma_cr"] ** 2
    phibar = np.zeros_like(phi0)
    pibar = np.zeros_like(pi0)
    d_s2 = 0.0
    d_k0 = 0.0
    for s in range(nsteps - 1, -1, -1):
        t = ts[s]
        f = kick_phi[s]
        dw = t * dt * P["dv_spa"] * s2
        c, sn = np.cos(q * t), np.sin(q * t)
        Sbar = []
        for comp in range(6):
            a11, a21, a12, a22 = abar
            if fix_swap:
                pdb = c * a11 + sn * a21
                ndb = c * a12 + sn * a22
            elif s == nsteps - 1 and comp == 5:
                pdb = c * a11 + sn * (a21 + a12)
                ndb = c * a12 + sn * a22
            else:
                pdb = c * a11 + sn * (a21 + a12)
                ndb = c * (a21 + a12) + sn * a22
            Pc = pol[comp] * dw
            Db = np.conj(Pc) * pdb + Pc * ndb
            Sbar.append(np.real(np.fft.fftn(Db)))
        g = _grad(f, P["dspa"])
        S = {(0, 0): Sbar[0], (1, 0): Sbar[1], (1, 1): Sbar[2], (2, 0): Sbar[3], (2, 1): Sbar[4], (2, 2): Sbar[5]}
        gb = []
        for i in range(3):
            acc = 4 * S[(i, i)] * g[i]
            for j in range(3):
                if j != i:
                    acc = acc + 2 * S[(max(i, j), min(i, j))] * g[j]
            gb.append(acc)
        f1, f2, f3, m2, ma2, kt = _kick_scalars(t, dt, P)
        lap = sum(np.roll(pibar, sh, axis=a) for a in (0, 1, 2) for sh in (1, -1)) - 6 * pibar
        a, bb = f.real, f.imag
        gg = 2 * (a * a + bb * bb) + m2
        hrr, hii, hri = gg + 4 * a * a - kt, gg + 4 * bb * bb + kt, 4 * a * bb
        hp = (hrr * pibar.real + hri * pibar.imag) + 1j * (hri * pibar.real + hii * pibar.imag)
        div = sum((np.roll(gb[i], 1, axis=ax) - np.roll(gb[i], -1, axis=ax)) for i, ax in zip(range(3), (2, 1, 0)))
        phibar = phibar + div / (2 * P["dspa"]) + f2 * lap + f3 * hp
        v0 = (gg * a - ma2 / 2) + 1j * gg * bb
        d_s2 += f3 / s2 * float(np.sum(pibar.real * v0.real + pibar.imag * v0.imag))
        d_k0 += f3 * t ** 4 / s2 * float(np.sum(pibar.imag * bb - pibar.real * a))
        hd = dt / 2 if s == 0 else dt
        pibar = f1 * pibar + hd * phibar
    d_lnA = float(np.sum(phibar.real * phi0.real + phibar.imag * phi0.imag) +
                  np.sum(pibar.real * pi0.real + pibar.imag * pi0.imag))
    return J, d_s2 + 2 * J / s2, d_k0, d_lnA, phibar, pibar


def band_energy(phi0, pi0, P, t1, dt, nsteps, nb, qlo, qhi, fix_swap=False):
    """J alone (the forward pass of `gradient`), for finite differences."""
    phi, pi = phi0.copy(), pi0.copy()
    A = O.new_gw_state(phi0.shape[0])
    O.symplectic(phi, pi, A, t1, dt, nsteps, P, fix_swap=fix_swap)
    sp = O.gw_spectrum(A, nb, P["dspa"], P["dwav"], P["tau_cr"], P["a_now"], P["dv_wav"])
    band = (sp["qmid"] >= qlo) & (sp["qmid"] <= qhi) & (sp["denom"] > 0)
    return float(sp["power"][band].sum())
