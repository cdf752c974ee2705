"""Pins for the CPU oracle (no GPU).  Each check is chosen so that a plausible mistake in the
oracle (a dropped term, a wrong sign or index, a transposed operand) fails it: closed forms,
brute force on tiny lattices, numpy cross-checks and invariants.  The reference ships no
golden vectors, fixtures or tests for this path, so these are the pins."""
import os

import numpy as np
import pytest

import oracle as O
from tests.conftest import SQ_HALF, phys

PI_T = 3.14159265359


def brute_dft3(a, sign):
    N = a.shape[0]
    k = np.arange(N)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / N)
    # a[z, y, x]: transform every axis
    return np.einsum("az,by,cx,zyx->abc", W, W, W, a)


# ------------------------------------------------------------------ FFT (grav_waves.f90:225)

def test_fft_matches_committed_golden_vectors():
    """numpy golden vectors (tests/golden/make_golden.py): the BACKWARD transform of seeded
    12^3 (radix 3) and 16^3 inputs."""
    for N in (12, 16):
        z = np.load(os.path.join(os.path.dirname(__file__), "golden", f"fft_{N}.npz"))
        out = O.fft3d(z["input"], +1)
        assert np.abs(out - z["backward"]).max() < 1e-12 * np.abs(z["backward"]).max()


@pytest.mark.parametrize("N", [4, 6, 8, 12])
@pytest.mark.parametrize("sign", [+1, -1])
def test_fft_brute_force(N, sign):
    rng = np.random.default_rng(N)
    a = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    ref = brute_dft3(a, sign)
    assert np.abs(O.fft3d(a, sign) - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("N", [16, 24, 32, 48])
def test_fft_numpy(N):
    rng = np.random.default_rng(N + 1)
    a = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    # FFTW_BACKWARD = unnormalized e^{+i}: numpy's ifftn times N^3
    ref = np.fft.ifftn(a) * N ** 3
    assert np.abs(O.fft3d(a, +1) - ref).max() < 1e-13 * np.abs(ref).max()


def test_fft_delta_plane_wave_parseval():
    N = 8
    d = np.zeros((N, N, N), complex)
    d[0, 0, 0] = 1
    assert np.allclose(O.fft3d(d, +1), 1)
    x = np.arange(N)
    kx, ky, kz = 1, 2, 3
    pw = np.exp(-2j * np.pi * (kx * x[None, None, :] + ky * x[None, :, None] + kz * x[:, None, None]) / N)
    out = O.fft3d(pw, +1)
    assert abs(out[kz, ky, kx] - N ** 3) < 1e-9 and np.abs(out).sum() - N ** 3 < 1e-8
    rng = np.random.default_rng(0)
    a = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    assert abs(np.sum(np.abs(O.fft3d(a)) ** 2) - N ** 3 * np.sum(np.abs(a) ** 2)) < 1e-9 * N ** 3 * np.sum(np.abs(a) ** 2)


# ------------------------------------------------------------------ kick / drift (evolve.f90:259-335)

def kick_scalars(t, dt, dspa, sigma_cr):
    f0 = 1 + dt / t
    return (1 - dt / t) / f0, dt / dspa ** 2 / f0, -dt * t * t * sigma_cr ** 2 / f0


def test_drift_exact():
    N = 4
    rng = np.random.default_rng(0)
    phi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    pi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    ref = phi + 0.3 * pi
    O.drift(phi, pi, 0.3)
    assert np.array_equal(phi, ref)


def test_kick_constant_field_closed_form():
    """Constant field: the Laplacian vanishes, so pi' = f1 pi + f3 V'(phi) exactly."""
    N, dspa, t, dt = 6, 0.25, 2.0, 0.05
    phi = np.full((N, N, N), 0.6 + 0.3j)
    pi = np.full((N, N, N), 0.1 - 0.2j)
    O.kick(phi, pi, t, dt, dspa, SQ_HALF, 1e-30, True)
    f1, f2, f3 = kick_scalars(t, dt, dspa, SQ_HALF)
    m2 = (0.0 - 2.449489742783178 ** 2) / 3  # thrm = 0 (initial_loop)
    c = 0.6 + 0.3j
    dv = (2 * abs(c) ** 2 + m2) * c - O.maxion2(0.0, 1e-30) / 2
    exp = f1 * (0.1 - 0.2j) + f3 * dv
    assert np.abs(pi - exp).max() < 1e-15


def test_kick_plane_wave_lattice_eigenvalue():
    """|phi| = 1 plane wave at zero temperature: V' = 0 (m2 = -2 makes 2|phi|^2+m2 = 0), and the
    7-point Laplacian of exp(i k.x) is exactly -sum_d (2 - 2 cos k_d dspa)/dspa^2 times it."""
    N, dspa, t, dt = 8, 0.25, 1.5, 0.04
    x = np.arange(N)
    k = 2 * np.pi * np.array([1, 2, 3]) / (N * dspa)
    ph = k[0] * dspa * x[None, None, :] + k[1] * dspa * x[None, :, None] + k[2] * dspa * x[:, None, None]
    phi = np.exp(1j * ph)
    pi = np.zeros_like(phi)
    O.kick(phi, pi, t, dt, dspa, SQ_HALF, 1e-30, True)
    f1, f2, f3 = kick_scalars(t, dt, dspa, SQ_HALF)
    m2 = -2.449489742783178 ** 2 / 3
    lam = -np.sum(2 - 2 * np.cos(k * dspa))
    dv = (2 + m2) * phi - O.maxion2(0.0, 1e-30) / 2
    exp = f2 * lam * phi + f3 * dv
    assert np.abs(pi - exp).max() < 1e-12


def test_tilt_vacua_are_static():
    """N_DW=2 tilt -kappa0 t^4 (phi^2+phi*^2)/2 at zero temperature: with the reference's
    normalisation the potential term is (|phi|^2-1)^2 - kt (re^2 - im^2), kt = kappa0 t^4/sigma_cr^2,
    so the uniform field r with r^2 = 1 + kt/2 (minimum, real axis) and i s with s^2 = 1 - kt/2
    (saddle, imaginary axis) feel no force.  A sign or factor error in the tilt moves them."""
    N, dspa, t, dt, kappa0 = 4, 0.25, 3.0, 0.05, 2e-3
    kt = kappa0 * t ** 4 / SQ_HALF ** 2
    ma2 = O.maxion2(0.0, 1e-30)
    for c in (np.sqrt(1 + kt / 2), 1j * np.sqrt(1 - kt / 2)):
        phi = np.full((N, N, N), complex(c))
        pi = np.zeros_like(phi)
        O.kick(phi, pi, t, dt, dspa, SQ_HALF, 1e-30, True, kappa0)
        assert np.abs(pi).max() < 1e-14 + abs(ma2) * dt * t * t
        pi0 = np.zeros_like(phi)
        O.kick(phi, pi0, t, dt, dspa, SQ_HALF, 1e-30, True)  # without the tilt the same field moves
        assert np.abs(pi0).max() > 1e-3


def test_tilt_force_is_gradient_of_energy():
    """The kick's tilt force equals minus the derivative of the energy's tilt term: with
    pi = 0 and a uniform field, the change of pi from the tilt over one kick, divided by f3,
    is dV/dphi*, and E_pot = t^2 sigma_cr^2 dspa^3 N^3 V gives dE/d(re) = 2 Re(dV/dphi*) times
    that prefactor (central differences of the oracle's own energy)."""
    N, dspa, t, dt, kappa0 = 4, 0.25, 2.0, 0.05, 5e-3
    c = 0.8 + 0.45j
    f1, f2, f3 = kick_scalars(t, dt, dspa, SQ_HALF)
    phi = np.full((N, N, N), c)
    on, off = np.zeros_like(phi), np.zeros_like(phi)
    O.kick(phi, on, t, dt, dspa, SQ_HALF, 1e-30, True, kappa0)
    O.kick(phi, off, t, dt, dspa, SQ_HALF, 1e-30, True)
    dvt = (on[0, 0, 0] - off[0, 0, 0]) / f3  # tilt part of dV/dphi*
    pref = t * t * SQ_HALF ** 2 * dspa ** 3 * N ** 3
    h = 1e-6
    for d, part in ((1.0, dvt.real), (1j, dvt.imag)):
        e = []
        for sgn in (1, -1):
            f = np.full((N, N, N), c + sgn * h * d)
            e.append(O.energy(f, np.zeros_like(f), t, dspa, SQ_HALF, 1e-30, True, kappa0)[2]
                     - O.energy(f, np.zeros_like(f), t, dspa, SQ_HALF, 1e-30, True)[2])
        deriv = (e[0] - e[1]) / (2 * h)
        assert abs(deriv - 2 * pref * part) < 1e-7 * abs(2 * pref * part)


def test_massless_mode_converges_second_order():
    """A small-amplitude mode of the linearised equation with H = 1/tau: the lattice solution
    with dispersion w^2 = sum (2-2cos)/dspa^2 is j0-like, sin(w tau)/(w tau); the DKD
    integrator converges to it at O(dt^2)."""
    N, dspa = 8, 0.5
    k = 2 * np.pi / (N * dspa)
    w = np.sqrt((2 - 2 * np.cos(k * dspa))) / dspa
    x = np.arange(N)
    P = dict(phys(N, dspa), sigma_cr=1e-9)  # switch the potential off (sigma_cr -> 0)
    t0, t1 = 1.0, 3.0
    errs = []
    for nsteps in (40, 80, 160):
        dt = (t1 - t0) / nsteps
        amp = np.sin(w * t0) / (w * t0)
        damp = (np.cos(w * t0) * w * t0 - np.sin(w * t0)) / (w * t0 * t0)
        mode = np.cos(k * dspa * x)[None, None, :] * np.ones((N, N, 1))
        phi = (amp * mode).astype(complex)
        pi = (damp * mode).astype(complex)
        O.symplectic(phi, pi, None, t0, dt, nsteps, P)
        exact = np.sin(w * t1) / (w * t1)
        errs.append(abs(phi[0, 0, 0].real - exact))
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5, errs


def test_time_reversal():
    """The Crank-Nicolson friction form is reversible: n steps forward then n steps with -dt
    (times run backwards) return the fields to round-off."""
    N = 8
    P = phys(N)
    phi, pi = O.ic_flat(N, 3, 4.0, P["dwav"])
    p0, q0 = phi.copy(), pi.copy()
    O.symplectic(phi, pi, None, 2.0, 0.02, 25, P)
    O.symplectic(phi, pi, None, 2.5, -0.02, 25, P)
    assert np.abs(phi - p0).max() < 1e-11 and np.abs(pi - q0).max() < 1e-11


# ------------------------------------------------------------------ gradient / source / accumulate

def test_gradient_plane_wave():
    N, dspa = 8, 0.25
    x = np.arange(N)
    k = 2 * np.pi * np.array([1, -2, 3]) / (N * dspa)
    phi = np.exp(1j * (k[0] * dspa * x[None, None, :] + k[1] * dspa * x[None, :, None] + k[2] * dspa * x[:, None, None]))
    g = O.gradient(phi, dspa)
    for d in range(3):
        assert np.abs(g[d] - 1j * np.sin(k[d] * dspa) / dspa * phi).max() < 1e-13


def pol_reference(n, i, j):
    """grav_waves.f90:165-186 restated in numpy for one wavevector."""
    n = np.asarray(n, float)
    q = np.sqrt(np.sum(n ** 2))
    qh = n / q
    r = np.hypot(qh[0], qh[1])
    c, s = (qh[0] / r, qh[1] / r) if r > 0 else (1.0, 0.0)
    p = np.array([qh[2] * c - 1j * s, qh[2] * s + 1j * c, -r + 0j])
    v = p[i] * p[j]
    return v / 2 if i == j else v, p, qh


@pytest.mark.parametrize("n", [(1, 2, 3), (0, 0, 5), (0, 0, -4), (-3, 1, 0), (2, -2, -1)])
def test_polarisation_invariants(n):
    _, p, qh = pol_reference(n, 0, 0)
    assert abs(np.dot(p, qh)) < 1e-15            # transverse
    assert abs(np.dot(p, p)) < 1e-15             # traceless e_ij = p_i p_j
    assert abs(np.vdot(p, p) - 2) < 1e-15        # normalisation
    _, pm, _ = pol_reference(-np.asarray(n), 0, 0)
    if np.hypot(n[0], n[1]) > 0:
        assert np.abs(pm - np.conj(p)).max() < 1e-15   # pol(-k) = conj(pol(k))


def test_gw_single_step_composition():
    """One GW integration on a field whose source has one Fourier mode: checks the helicity
    projection, the dw weight and the swapped accumulator order of grav_waves.f90:239-242
    against a hand composition of the six sequential updates."""
    N, dspa = 8, 0.25
    P = phys(N, dspa)
    x = np.arange(N)
    k = 2 * np.pi * np.array([1, 0, 0]) / (N * dspa)
    # phi = exp(i k x) + 0.5 exp(i 2k y): S_ij has a few modes
    phi = (np.exp(1j * k[0] * dspa * x[None, None, :]) + 0.5 * np.exp(2j * k[0] * dspa * x[None, :, None])) \
        * np.ones((N, 1, 1))
    tau, dtau = 2.0, 0.05
    A0 = np.random.default_rng(1).standard_normal((4, N, N, N)) * (1 + 0j)
    for fix in (False, True):
        A = A0.copy()
        O.gw_integrate(phi, A, tau, dtau, dspa, P["dwav"], P["dv_spa"], P["sigma_cr"], fix_swap=fix)
        g = O.gradient(phi, dspa)
        dw = tau * dtau * P["dv_spa"] * P["sigma_cr"] ** 2
        ref = A0.copy()
        idx = np.arange(N)
        nw = np.where(idx < N // 2, idx, idx - N)
        for (i, j) in [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]:
            S = g[i] * np.conj(g[j]) + np.conj(g[i]) * g[j]
            D = np.fft.ifftn(S) * N ** 3
            for kz in range(N):
                for ky in range(N):
                    for kx in range(N):
                        nn = (nw[kx], nw[ky], nw[kz])
                        if nn == (0, 0, 0):
                            pol = 0
                        else:
                            pol, _, _ = pol_reference(nn, i, j)
                        q = np.sqrt(sum(v * v for v in nn)) * P["dwav"] * tau
                        c, s = np.cos(q), np.sin(q)
                        pp = pol * dw
                        pn = np.conj(pp)
                        d = D[kz, ky, kx]
                        a = ref[:, kz, ky, kx]
                        a[0] += c * pp * d
                        if not fix:
                            a[1] = a[2] + s * pp * d
                            a[2] = a[1] + c * pn * d
                        else:
                            a[1] += s * pp * d
                            a[2] += c * pn * d
                        a[3] += s * pn * d
        assert np.abs(A - ref).max() < 1e-12 * np.abs(ref).max(), fix


def test_gw_half_spectrum_is_exact_for_binned_modes():
    """Every binned mode outside the stored half spectrum (kx > N/2) is the conjugate of its
    mirror inside it, so storing kx = 0..N/2 loses nothing the spectrum reads."""
    N = 16
    P = phys(N)
    phi, pi = O.ic_flat(N, 9, 4.0, P["dwav"])
    A = O.new_gw_state(N)
    O.symplectic(phi, pi, A, 1.0, 0.05, 2, P)
    idx = np.arange(N)
    n = np.where(idx < N // 2, idx, idx - N)
    mirror = (-idx) % N
    for kz in range(N):
        for ky in range(N):
            for kx in range(N // 2 + 1, N):
                if n[kx] ** 2 + n[ky] ** 2 + n[kz] ** 2 > (N // 2) ** 2:
                    continue
                for m in range(4):
                    a = A[m, kz, ky, kx]
                    b = np.conj(A[m, mirror[kz], mirror[ky], mirror[kx]])
                    assert abs(a - b) <= 1e-12 * np.abs(A[m]).max()


# ------------------------------------------------------------------ spectra bins

def nint(x):
    return np.floor(np.abs(x) + 0.5) * np.sign(x)


@pytest.mark.parametrize("N,nb", [(8, 25), (16, 25), (12, 7)])
def test_gw_bin_counts_brute_force(N, nb):
    P = phys(N)
    A = np.ones((4, N, N, N), complex)
    s = O.gw_spectrum(A, nb, P["dspa"], P["dwav"], 1.0, 1.0, P["dv_wav"])
    dq = PI_T / P["dspa"] / nb
    cnt = np.zeros(nb)
    idx = np.arange(N)
    n = np.where(idx < N // 2, idx, idx - N)
    for a in n:
        for b in n:
            for c in n:
                q = np.sqrt(float(a * a + b * b + c * c)) * P["dwav"]
                i = int(nint(q / dq + 0.5))
                if q > 0 and 1 <= i <= nb:
                    cnt[i - 1] += 1
    assert np.array_equal(s["denom"], cnt)
    assert np.allclose(s["numer"], 4 * cnt)


def test_conv_spectrum_single_mode():
    """One Fourier mode of pi lands in one bin with |N^3 dv_spa|^2 dv_wav (q = 0 counted in
    bin 1, spectrum.f90:538-551)."""
    N, nb = 8, 10
    P = phys(N)
    x = np.arange(N)
    pi = np.exp(-2j * np.pi * 2 * x[None, None, :] / N) * np.ones((N, N, 1))
    pw, num = O.conv_spectrum(pi, None, nb, P["dspa"], P["dwav"], P["dv_spa"], P["dv_wav"])
    q = 2 * P["dwav"]
    b = int(nint(q / (PI_T / P["dspa"] / nb) + 0.5))
    exp = np.zeros(nb)
    exp[b - 1] = (N ** 3 * P["dv_spa"]) ** 2 * P["dv_wav"]
    assert np.allclose(pw, exp, rtol=1e-13, atol=1e-30)
    assert num[0] >= 1  # k = 0 is binned


def test_integrate_p_check_shell_overlap():
    """integrate_p_check(li,si,lj,sj,q) = 2 pi int p^2 dp dmu over p in shell j with |p+q| in
    shell i; for q = 0 it is the full shell volume 4/3 pi (l^3 - s^3) of the overlap."""
    v = O.integrate_p_check(200, 2.0, 1.0, 2.0, 1.0, 0.0)
    assert abs(v - 4 / 3 * PI_T * (8 - 1)) < 1e-3 * v
    # Monte Carlo of the overlap volume for q > 0
    rng = np.random.default_rng(0)
    li, si, lj, sj, q = 2.0, 1.5, 1.7, 1.2, 0.6
    pts = rng.uniform(-lj, lj, size=(400000, 3))
    r = np.linalg.norm(pts, axis=1)
    inj = (r >= sj) & (r < lj)
    rq = np.linalg.norm(pts + np.array([0, 0, q]), axis=1)
    ini = (rq >= si) & (rq < li)
    mc = (2 * lj) ** 3 * np.mean(inj & ini)
    assert abs(O.integrate_p_check(200, li, si, lj, sj, q) - mc) < 0.02 * mc


def test_packed_inverse():
    rng = np.random.default_rng(0)
    nb = 7
    M = rng.standard_normal((nb, nb))
    M = M @ M.T + nb * np.eye(nb)
    ap = np.array([M[j - 1, i - 1] for i in range(1, nb + 1) for j in range(1, i + 1)])
    inv, info = O.pp_inverse(nb, ap)
    assert info == 0
    assert np.allclose(O.sym_from_packed(inv, nb), np.linalg.inv(M), rtol=1e-12, atol=1e-14)


def test_deconvolution_without_mask_is_diagonal():
    """No masked cell: W = N^3 dv_spa delta_k0, so each pair's matrix element is
    |W(0)|^2 u_ij(0) / sum u_ij * dv_wav^2, non-zero only where the pair's shells overlap at
    q = 0, i.e. on the diagonal."""
    N, nb = 8, 4
    P = phys(N)
    tab = O.init_spectrum(nb, P["dspa"])
    flags = np.zeros((N, N, N), np.int8)
    mat, inv, info = O.matrix_deconv(flags, nb, P["dwav"], P["dv_spa"], P["dv_wav"], tab)
    M = O.sym_from_packed(mat, nb)
    off = M - np.diag(np.diag(M))
    assert np.abs(off).max() <= 1e-12 * np.abs(np.diag(M)).max()


# ------------------------------------------------------------------ spectrum covariance (spectrum.f90:569-734)

def packed(M):
    nb = M.shape[0]
    return np.array([M[j - 1, i - 1] for i in range(1, nb + 1) for j in range(1, i + 1)])


def bin_index(N, nb, P):
    n = np.fft.fftfreq(N, 1.0 / N)
    n2 = n[:, None, None] ** 2 + n[None, :, None] ** 2 + n[None, None, :] ** 2
    q = np.sqrt(n2) * P["dwav"]
    dq = PI_T / P["dspa"] / nb
    i = np.floor(q / dq + 0.5 + 0.5).astype(int)  # nint of a non-negative argument
    return np.where(i <= nb, i, 0)


def test_cov_without_mask_is_binned_mode_power():
    """No mask: the FORWARD/BACKWARD pair returns N^3 times the modes, so realisation n's
    pow0_b = sum over bin b of |g_k|^2 power_b/2 with the realisation's unit Gaussians g_k;
    with matn = diag(1/num) its pow_b = pow0_b/num_b - power_b."""
    N, nb = 8, 4
    P = phys(N)
    b = bin_index(N, nb, P)
    num = np.array([np.count_nonzero(b == i) for i in range(1, nb + 1)], float)
    power = np.array([3.0, 0.5, 2.0, 1.25])
    matn = packed(np.diag(1 / num))
    flags = np.zeros((N, N, N), np.int8)
    cov, it, pows = O.field_covariance(flags, nb, P["dspa"], P["dwav"], power, matn, 77, 3, 3, 0.0, True)
    assert it == -4  # tol 0 never converges: the counter ends at max_mc + 1
    for n in range(1, 4):
        g2 = np.array([abs(O.cov_gauss(77, n, k)) ** 2 for k in range(N ** 3)]).reshape(N, N, N)
        pow0 = np.array([(g2[b == i] * power[i - 1] / 2).sum() for i in range(1, nb + 1)])
        assert np.allclose(pows[n - 1], pow0 / num - power, rtol=1e-11, atol=1e-12 * power.max())
    S = sum(np.outer(p, p) for p in pows)
    assert np.allclose(cov, packed(S) / 4, rtol=1e-13, atol=0)


def test_cov_fully_masked_is_outer_product():
    """Every cell masked: each realisation's pow is -power, so cov = power power^T; the
    jackknife spread is zero, so the loop stops at min_mc, and without convergence (tol = 0)
    the sum of max_mc products is divided by max_mc + 1."""
    N, nb = 8, 4
    P = phys(N)
    power = np.array([1.0, 2.0, 0.25, 4.0])
    matn = packed(np.diag([0.3, 0.2, 0.1, 0.05]))
    flags = np.full((N, N, N), 4, np.int8)
    cov, it, _ = O.field_covariance(flags, nb, P["dspa"], P["dwav"], power, matn, 5, 6, 20, 0.1)
    assert it == 6
    assert np.allclose(cov, packed(np.outer(power, power)), rtol=1e-14, atol=0)
    cov, it, _ = O.field_covariance(flags, nb, P["dspa"], P["dwav"], power, matn, 5, 6, 20, 0.0)
    assert it == -21
    assert np.allclose(cov, packed(np.outer(power, power)) * 20 / 21, rtol=1e-14, atol=0)


@pytest.mark.parametrize("tol,stop_at", [(0.1, 13), (0.05, 22), (0.01, None)])
def test_cov_jackknife_stop(tol, stop_at):
    """The leave-one-out bookkeeping against a direct restatement: after each realisation
    n >= min_mc, x_k = (S_n - p_k p_k^T)_bb / (n-1) for k <= n; diff = max_b sd_k(x_k) /
    (S_n,bb / n); the loop stops at the first n with diff < tol and returns S_n / n."""
    N, nb = 8, 4
    P = phys(N)
    rng = np.random.default_rng(4)
    flags = np.where(rng.random((N, N, N)) < 0.3, 4, 0).astype(np.int8)
    power = np.array([2.0, 1.0, 0.5, 0.8])
    A = rng.standard_normal((nb, nb))
    matn = packed((A @ A.T + nb * np.eye(nb)) / 50)
    min_mc, max_mc = 4, 60
    cov, it, pows = O.field_covariance(flags, nb, P["dspa"], P["dwav"], power, matn, 9, min_mc, max_mc, tol, True)
    stop = None
    for n in range(min_mc, max_mc + 1):
        p = pows[:n]
        S = p.T @ p
        loo = np.diag(S)[None, :] - p * p
        x = loo / (n - 1)
        sd = np.sqrt((x * x).mean(0) - x.mean(0) ** 2)
        if (sd / (np.diag(S) / n)).max() < tol:
            stop = n
            break
    assert stop == stop_at  # the seeded case stops mid-loop (or never, for the smallest tol)
    if stop is None:
        assert it == -(max_mc + 1)
        assert np.allclose(cov, packed(pows.T @ pows) / (max_mc + 1), rtol=1e-12, atol=0)
    else:
        assert it == stop
        p = pows[:stop]
        assert np.allclose(cov, packed(p.T @ p) / stop, rtol=1e-12, atol=0)


def test_cov_matches_gaussian_variance():
    """Statistics: without a mask and with matn = diag(1/num), pow_b is the mean of num_b
    exponential variables of mean power_b minus power_b, so cov_bb -> power_b^2 / num_b and
    the off-diagonal terms -> 0 (400 realisations: about 7% per bin)."""
    N, nb = 8, 4
    P = phys(N)
    b = bin_index(N, nb, P)
    num = np.array([np.count_nonzero(b == i) for i in range(1, nb + 1)], float)
    power = np.array([1.0, 3.0, 2.0, 0.5])
    matn = packed(np.diag(1 / num))
    flags = np.zeros((N, N, N), np.int8)
    cov, it, _ = O.field_covariance(flags, nb, P["dspa"], P["dwav"], power, matn, 2024, 400, 400, 0.0)
    C = O.sym_from_packed(cov, nb) * 401 / 400
    r = np.diag(C) * num / power ** 2
    assert np.all(np.abs(r - 1) < 0.3) and abs(r.mean() - 1) < 0.12
    corr = C / np.sqrt(np.outer(np.diag(C), np.diag(C)))
    assert np.abs(corr - np.eye(nb)).max() < 0.2


# ------------------------------------------------------------------ networks (defects_3d.f90:1235-2204)

def test_lat_dist_periodic():
    """The nearest periodic image on each axis, also for positions half a cell outside [0, N)."""
    N = 16
    assert O.lat_dist([0.5, 0.5, 0.5], [15.5, 0.5, 0.5], N) == 1.0
    assert abs(O.lat_dist([-0.25, 3.0, 8.0], [16.25, 4.0, 8.0], N) - np.sqrt(0.25 + 1.0)) < 1e-15
    assert abs(O.lat_dist([1.0, 1.0, 1.0], [4.0, 5.0, 13.0], N) - np.sqrt(9 + 16 + 16)) < 1e-14


def components(pos, N, th):
    """scipy's periodic k-d tree pairs (distance <= th) and connected components, as labels =
    the smallest index of each component."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    from scipy.spatial import cKDTree

    n = len(pos)
    pairs = cKDTree(np.mod(pos, N), boxsize=N).query_pairs(th, output_type="ndarray")
    A = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n))
    _, comp = connected_components(A, directed=False)
    first = {}
    for i, c in enumerate(comp):
        first.setdefault(c, i)
    return np.array([first[c] for c in comp], np.int32)


@pytest.mark.parametrize("n,th", [(300, 2.0), (800, 1.3), (1500, 1.05)])
def test_group_points_matches_periodic_kdtree(n, th):
    """Brute-force grouping against scipy's periodic k-d tree and connected components; points
    sit up to 0.45 outside their cell like fitted defect positions, some beyond [0, N)."""
    N = 16
    rng = np.random.default_rng(n)
    cells = rng.choice(N ** 3, size=n, replace=False)
    base = np.stack([cells % N, (cells // N) % N, cells // (N * N)], axis=1) + 0.5
    pos = base + rng.uniform(-0.95, 0.95, size=(n, 3))
    rec = np.concatenate([pos, rng.random((n, 8))], axis=1)
    lab = O.group_points(rec, N, th)
    assert np.array_equal(lab, components(pos, N, th))
    assert 1 < len(set(lab)) < n  # neither all apart nor all joined


def test_group_points_chains_and_threshold():
    """Points 1.5 apart along a line join into one chain through the periodic boundary; a
    parallel line 2.01 away stays apart at th = 2 and joins at th = 2.02."""
    N = 16
    a = np.array([[x + 0.25, 4.0, 4.0] for x in np.arange(0, 16, 1.5)])
    b = a + [0.0, 2.01, 0.0]
    rec = np.concatenate([a, b])
    lab = O.group_points(rec, N, 2.0)
    assert len(set(lab)) == 2 and np.all(lab[: len(a)] == 0) and np.all(lab[len(a):] == len(a))
    assert len(set(O.group_points(rec, N, 2.02))) == 1
    # the chain closes through the boundary: its first and last points are 16 - 15.25 + 0.25 = 1 apart
    assert O.lat_dist(a[0], a[-1], N) == 1.0


def test_group_sums_nested_thresholds():
    """Per-group sums in record order, then totals over all groups and over those above each
    threshold; groups numbered by their first record."""
    label = np.array([0, 1, 0, 3, 1, 3, 3], np.int32)
    val = np.array([1.0, 2.0, 3.0, 0.5, 4.0, 0.25, 0.25])
    ng, ids, s = O.group_sums(label, val, 2.0, 5.0)
    assert ng == 3 and list(ids) == [1, 2, 1, 3, 2, 3, 3]
    assert list(s) == [11.0, 10.0, 6.0]  # groups: 4, 6, 1
    assert O.group_threshold(0.5, 10.0) == 2.0 and abs(O.group_threshold(3.0, 8.0) - 72 ** (1 / 3)) < 1e-14


# ------------------------------------------------------------------ defects (defects_3d.f90:309-349)

def classify_cell(ph, ids=True, idw=True, spec=True):
    p = np.sort(ph)
    t = p[0] - p[7] + 2 * PI_T
    tmax = np.max(np.diff(p))
    fl = 0
    if idw and tmax > t:
        fl += 2
    else:
        tmax = max(t, tmax)
    if ids and tmax < PI_T:
        fl += 1
    if spec and tmax < 7 * PI_T / 4:
        fl += 4
    return fl


def test_classify_uniform_phase():
    N = 6
    phi = np.full((N, N, N), np.exp(0.4j))
    fl, cnt = O.classify(phi)
    assert np.array_equal(cnt, [0, 0, 0])


def test_classify_straight_string():
    """phi = (x - x0) + i (y - y0) with x0, y0 inside one cell: exactly one string cell per
    z plane (the winding cell), every z."""
    N = 8
    x = np.arange(N)
    X = x[None, None, :] - 3.5
    Y = x[None, :, None] - 4.5
    phi = (X + 1j * Y) * np.ones((N, 1, 1))
    fl, cnt = O.classify(phi, True, False, False)
    s = fl & 1
    # the periodic wrap adds the image windings at the box edges: count the cell at (3, 4)
    assert all(s[z, 4, 3] == 1 for z in range(N))
    assert s[:, 1:N - 1, 1:N - 1].sum() == N


def test_classify_wall_plane():
    """An N_DW = 1 wall is where the phase passes through +-pi: phi = -1 + 0.1 i sign, the sign
    flipping across x = 3|4 (and, periodically, 7|0).  The cells straddling the flips hold
    phases on both sides of the branch cut, so their largest gap is the one through 0: walls
    there and nowhere else (defects_3d.f90:332-336)."""
    N = 8
    x = np.arange(N)
    sgn = np.where(x < 4, 1.0, -1.0)
    phi = (-1 + 0.1j * sgn)[None, None, :] * np.ones((N, N, 1))
    fl, cnt = O.classify(phi, True, True, False)
    w = (fl & 2) == 2
    assert np.all(w[:, :, 3]) and np.all(w[:, :, 7])
    assert w.sum() == 2 * N * N and cnt[1] == 2 * N * N
    assert cnt[0] == 0


def test_classify_random_matches_brute_force():
    N = 6
    rng = np.random.default_rng(5)
    phi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    fl, cnt = O.classify(phi)
    exp = np.zeros_like(fl)
    for z in range(N):
        for y in range(N):
            for x in range(N):
                ph = [np.angle(phi[(z + dz) % N, (y + dy) % N, (x + dx) % N])
                      for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)]
                exp[z, y, x] = classify_cell(np.array(ph))
    assert np.array_equal(fl, exp)
    assert cnt[0] == np.sum(exp & 1) and cnt[1] == np.sum((exp >> 1) & 1) and cnt[2] == np.sum((exp >> 2) & 1)


# ------------------------------------------------------------------ energy / IC / settings

def test_energy_constant_field():
    N, dspa, t = 4, 0.25, 2.0
    c = 0.7 + 0.2j
    phi = np.full((N, N, N), c)
    pi = np.full((N, N, N), 0.3 - 0.1j)
    e = O.energy(phi, pi, t, dspa, SQ_HALF, 1e-30, True)
    m2 = -2.449489742783178 ** 2 / 3
    V = (abs(c) ** 2 + m2 / 2) ** 2 - O.maxion2(0.0, 1e-30) * c.real
    dv = dspa ** 3 * N ** 3
    assert abs(e[0] - 0.1 * dv) < 1e-14
    assert e[1] == 0
    assert abs(e[2] - t * t * 0.5 * V * dv) < 1e-13
    assert abs(e[3] - e[0] - e[1] - e[2]) < 1e-14


def test_philox_known_answer():
    # Random123 Philox4x32-10 known-answer vectors (counter, key) -> output
    assert list(O.philox([0, 0, 0, 0], 0, 0)) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(O.philox([0xffffffff] * 4, 0xffffffff, 0xffffffff)) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], 0xa4093822, 0x299f31d0)) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_ic_flat_mode_count_and_norm():
    N = 64
    P = phys(N)
    phi, pi = O.ic_flat(N, 1234, 1.0, P["dwav"])
    assert abs(np.mean(np.abs(phi) ** 2) - 1) < 1e-12
    nz = np.abs(np.fft.fftn(phi)) > 1e-9 * np.abs(np.fft.fftn(phi)).max()
    assert nz.sum() == 81  # |n|^2 <= 6 for |k| < 1 at L = 16
    assert not np.any(pi)


def test_get_dtau_hand_values():
    # omega = sqrt((dwav N)^2 + (T_cr/sigma tau2 sigma_cr)^2), evolve.f90:17-21
    dt, n = O.get_dtau(1.0, 6.0, SQ_HALF, 2 * PI_T / 16, 64, 1.0)
    om = np.sqrt((2 * PI_T / 16 * 64) ** 2 + (2.449489742783178 * 6.0 * SQ_HALF) ** 2)
    assert n == int(np.ceil(5.0 * om)) and abs(dt - 5.0 / n) < 1e-16


# ------------------------------------------------------------------ string fit (defects_3d.f90:384-603, 1027-1056)

def corner_xyz():
    """Corner i = (iz*2+iy)*2+ix of a cell, relative to the cell centre."""
    return np.array([[ix - 0.5, iy - 0.5, iz - 0.5] for iz in (0, 1) for iy in (0, 1) for ix in (0, 1)])


def test_string_length_closed_forms():
    r2 = 2 ** -0.5
    assert O.string_length([0, 0, 0], [0, 0, 1]) == 1.0                      # through the centre along z
    assert abs(O.string_length([0, 0, 0], [r2, r2, 0]) - 2 ** 0.5) < 1e-15   # face diagonal
    assert O.string_length([0.3, -0.2, 0], [0, 0, 1]) == 1.0                 # off centre, still crosses z faces
    assert abs(O.string_length([0, 0.2, 0.1], [1, 0, 0]) - 1.0) < 1e-15
    # chord from (0.5, 0) to (0, -0.5) of a diagonal line through (0.25, -0.25)
    assert abs(O.string_length([0.25, -0.25, 0], [r2, r2, 0]) - r2) < 1e-15
    # the body diagonal exits at sqrt(3)/2, just beyond the reference's cut 0.86602540378378444
    s3 = 3 ** -0.5
    assert O.string_length([0, 0, 0], [s3, s3, s3]) == 0.0


def test_chi2_vanishes_for_straight_strings():
    """Phases of a physical straight string through the cell: phi = (x-x0) + i(y-y0) along z, and
    (y-y0) + i(z-z0) along x.  The reference's model reproduces them up to a constant phase,
    so its chi2 (a variance) is zero at the string's parameters and positive elsewhere."""
    c = corner_xyz()
    x0, y0, z0 = 0.17, -0.08, 0.11
    ph = np.arctan2(c[:, 1] - y0, c[:, 0] - x0)
    assert O.chi2_string(ph, [1.0, 0.0, np.hypot(x0, y0), np.arctan2(y0, x0)]) < 1e-28
    assert O.chi2_string(ph, [1.0, 0.0, np.hypot(x0, y0) + 0.2, np.arctan2(y0, x0)]) > 1e-3
    # along x: dir = (1,0,0) is mu = 0, phi = 0; the model's (du, dv) = (-u0 - z, -v0 + y)
    ph = np.arctan2(c[:, 2] - z0, c[:, 1] - y0)
    u0, v0 = -z0, y0
    assert O.chi2_string(ph, [0.0, 0.0, np.hypot(u0, v0), np.arctan2(v0, u0)]) < 1e-28
    assert O.chi2_string(ph, [0.3, 0.0, np.hypot(u0, v0), np.arctan2(v0, u0)]) > 1e-3


def test_position_string_recovers_straight_strings():
    """The fit (random starts + the reference's simplex, stopping at chi2 <= 0.05) finds strings
    along z through random points: direction within ~20 degrees of z, the chord close to 1."""
    c = corner_xyz()
    rng = np.random.default_rng(5)
    sls, dz = [], []
    for s in range(200):
        x0, y0 = rng.uniform(-0.3, 0.3, 2)
        ph = np.arctan2(c[:, 1] - y0, c[:, 0] - x0)
        pos, dirv, gof, sl = O.position_string(ph, 11, 22, s)
        assert gof >= np.exp(-0.05 / 2) * (1 - 1e-12)           # converged: chi2 <= 0.05
        assert abs(np.linalg.norm(dirv) - 1) < 1e-12
        sls.append(sl)
        dz.append(abs(dirv[2]))
    assert np.median(dz) > 0.97 and np.min(dz) > 0.85
    assert abs(np.median(sls) - 1) < 0.05


def test_velocity_of_moving_string():
    """A straight string moving with velocity u perpendicular to it: phi = (r-q-u tau).e1 +
    i (r-q-u tau).e2 in physical units, so pi = -(u.e1) - i (u.e2); the fitted velocity is u
    exactly (the field is linear, so the least-squares fit is exact)."""
    N, dspa = 6, 0.25
    rng = np.random.default_rng(9)
    for trial in range(5):
        n = rng.standard_normal(3); n /= np.linalg.norm(n)
        e1 = np.cross(n, rng.standard_normal(3)); e1 /= np.linalg.norm(e1)
        e2 = np.cross(n, e1)
        u = 0.4 * e1 - 0.25 * e2
        q = rng.uniform(1, 3, 3) * dspa
        z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
        r = np.stack([x, y, z], -1) * dspa - q
        phi = np.ascontiguousarray(r @ e1 + 1j * (r @ e2))
        pi = np.full((N, N, N), -(u @ e1) - 1j * (u @ e2))
        vel = O.velocity_string(phi, pi, 2, 1, 2, [0.1, -0.2, 0.05], dspa)
        assert np.abs(vel - u).max() < 1e-12, (vel, u)


def test_strings_records_and_sums():
    """orc_strings over a random field: one record per string cell in site order, positions
    within sqrt(3)/2 of their cell centre, sums equal to the records' sums."""
    N = 16
    P = phys(N)
    phi, _ = O.ic_flat(N, 3, 6.0, P["dwav"])
    rng = np.random.default_rng(1)
    pi = 0.2 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    flags, counts = O.classify(phi)
    rec, sums = O.strings(phi, pi, flags, 7, 8, P["dspa"])
    assert rec.shape == (counts[0], 11) and counts[0] > 10
    zz, yy, xx = np.nonzero(flags & 1)
    centre = np.stack([xx, yy, zz], 1) + 0.5
    assert np.abs(rec[:, :3] - centre).max() <= 0.8661
    assert abs(sums[0] - rec[:, 7].sum()) < 1e-12 * max(1.0, sums[0])
    v2 = (rec[:, 8:11] ** 2).sum(1)
    ok = v2 < 1
    assert sums[1] == ok.sum()
    assert abs(sums[2] - np.sqrt(v2[ok]).sum()) < 1e-10 * max(1.0, sums[2])
    rec2, _ = O.strings(phi, pi, flags, 7, 8, P["dspa"])
    assert np.array_equal(rec, rec2)
    rec3, _ = O.strings(phi, pi, flags, 7, 9, P["dspa"])   # another seed: other starting points
    assert not np.array_equal(rec[:, 7], rec3[:, 7])


# ------------------------------------------------------------------ thermal initial state (initial.f90:80-238, 372-390)

def thermal_continuum(tau, sigma):
    """<|phi|^2> = tau^3 int d^3p/(2pi)^3 n(w)/w and <|phidot|^2> = tau^3 sigma^2 int d^3p/(2pi)^3
    n(w) w for the Bose-Einstein occupation n at T = thrm_sigma/tau with the thermal mass
    m^2 = (T^2 - T_cr^2)/3 (at tau = 1, m = 0 these are the reference's own check values
    T^2/12 and (T^2 pi sigma)^2/30, initial.f90:426-433)."""
    from scipy.integrate import quad
    T0 = 2.449489742783178
    T = T0 / tau
    m2 = (T * T - T0 * T0) / 3
    w = lambda p: np.sqrt(m2 + p * p)
    n = lambda p: 1 / np.expm1(w(p) / T)
    i1 = quad(lambda p: p * p * n(p) / w(p), 0, 60 * T, limit=400)[0] / (2 * np.pi ** 2)
    i2 = quad(lambda p: p * p * n(p) * w(p), 0, 60 * T, limit=400)[0] / (2 * np.pi ** 2)
    return tau ** 3 * i1, tau ** 3 * sigma ** 2 * i2


@pytest.mark.parametrize("tau", [0.1, 0.3])
def test_thermal_state_matches_continuum_variances(tau):
    """The lattice thermal state's variances against the continuum integrals (the lattice
    covers the thermal momenta: p_max/T ~ 7 at dspa = 0.25): phi to 3 %, phidot to 5 % (UV
    cutoff), real and imaginary parts alike and uncorrelated."""
    N = 64
    P = phys(N)
    phi, pi = O.ic_thermal(N, 5, tau, SQ_HALF, P["dwav"], P["dv_wav"])
    e1, e2 = thermal_continuum(tau, SQ_HALF)
    assert abs(np.mean(np.abs(phi) ** 2) / e1 - 1) < 0.03
    assert abs(np.mean(np.abs(pi) ** 2) / e2 - 1) < 0.05
    for f in (phi, pi):
        v = np.mean(np.abs(f) ** 2)
        assert abs(2 * np.mean(f.real ** 2) / v - 1) < 0.05
        assert abs(np.mean(f.real * f.imag)) < 0.03 * v


def test_thermal_gaussians_are_unit_normal():
    g = np.array([O.thermal_gauss(9, ip, lin) for ip in (1, 2) for lin in range(20000)])
    for part in (g.real, g.imag):
        assert abs(part.mean()) < 0.02 and abs(part.var() - 1) < 0.03
        assert abs(np.mean(part ** 4) - 3) < 0.15          # Gaussian kurtosis
    assert abs(np.mean(g.real * g.imag)) < 0.02
    assert O.thermal_gauss(9, 1, 5) != O.thermal_gauss(9, 2, 5) != O.thermal_gauss(10, 1, 5)


# ------------------------------------------------------------------ wall fit (defects_3d.f90:1058-1202)

def planar_wall_phases(n, d, eps=1e-3):
    """Corner phases of phi = -1 + i eps (r.n - d): Im(phi) changes sign on the plane r.n = d."""
    n = np.asarray(n, float)
    n = n / np.linalg.norm(n)
    return np.arctan2(eps * (corner_xyz() @ n - d), -np.ones(8))


@pytest.mark.parametrize("n,d,m,area", [([1, 0, 0], 0.1, 4, 1.0),
                                        ([1, 1, 0], 0.0, 4, 2 ** 0.5),
                                        ([1, 1, 1], 0.0, 6, 3 * 3 ** 0.5 / 4),   # central hexagon
                                        ([0.3, -0.2, 1.0], 0.12, 4, np.sqrt(1.13))])
def test_wall_area_of_planar_walls(n, d, m, area):
    """A planar wall through the cell: the edge points are exact (to O(eps^2)), the polygon
    is the plane's cross-section of the cube, whose area is a closed form.  The fitted
    normal comes from a 20 x 20 grid of directions (up to ~9 degrees off), so the area
    agrees to 2 % and the normal to within the grid spacing."""
    mm, pos, dirv, gof, wa = O.position_wall(planar_wall_phases(n, d))
    assert mm == m
    assert abs(wa - area) < 0.02 * area
    nn = np.asarray(n, float) / np.linalg.norm(n)
    assert abs(abs(dirv @ nn) - 1) < 0.02 and abs(np.linalg.norm(dirv) - 1) < 1e-12
    assert gof > 0.95


def test_wall_needs_three_edge_points():
    # a uniform phase has no sign change of Im(phi) on any edge: the reference stops
    assert O.position_wall(np.full(8, 0.3))[0] == 0


def test_walls_records_and_sums():
    N = 16
    P = phys(N)
    phi, _ = O.ic_flat(N, 3, 6.0, P["dwav"])
    flags, counts = O.classify(phi)
    rec, sums = O.walls(phi, flags)
    assert rec.shape == (counts[1], 8) and counts[1] > 10 and sums[1] == 0
    zz, yy, xx = np.nonzero(flags & 2)
    assert np.abs(rec[:, :3] - (np.stack([xx, yy, zz], 1) + 0.5)).max() <= 0.87
    assert abs(sums[0] - rec[:, 7].sum()) < 1e-12 * max(1.0, sums[0])
    # a plane cuts a unit cube in at most ~1.41 (sqrt(2)); for non-planar edge points the triangle
    # fan of the reference can give a slightly negative area
    assert np.all(rec[:, 7] > -0.1) and np.all(rec[:, 7] < 1.5) and np.median(rec[:, 7]) > 0.3


# ------------------------------------------------------------------ GW power prefactor, units, maxion2

# potential.f90:13-19, typed in here so that the oracle's constants are checked against them
BIGH_GEV = 2.1332e-42       # H0/h in GeV
MSTAR_GEV = 2.435e18        # reduced Planck mass in GeV
GEV_IN_K = 1.16045250061657e13
TCMB_K = 2.726
G_REL_NOWS = 3.9091


def _single_mode_state(N, n, amp):
    """Accumulators with one nonzero mode n = (nx, ny, nz) of component (1,1): |A|^2 = amp^2."""
    A = O.new_gw_state(N)
    A[0, n[2] % N, n[1] % N, n[0] % N] = amp
    return A


def test_gw_power_scales_with_tau_cr_a_now_and_dv_wav():
    """power = numer/denom qmid^3 fact with fact = tau_cr^-6 (BigH/M*)^-2 a_now^-4 / 3 / (2 pi)^2
    dv_wav (grav_waves.f90:266, 301): doubling tau_cr divides every bin by 2^6, doubling a_now by
    2^4, doubling dv_wav multiplies by 2.  A wrong exponent fails one of these."""
    N, nb = 16, 8
    P = phys(N)
    A = _single_mode_state(N, (2, 1, 0), 1.5)
    base = O.gw_spectrum(A, nb, P["dspa"], P["dwav"], 1.0, 1.0, P["dv_wav"])["power"]
    b = np.nonzero(base)[0]
    assert len(b) == 1
    for kw, ratio in ((dict(tau_cr=2.0), 2.0 ** -6), (dict(a_now=2.0), 2.0 ** -4), (dict(dv_wav=2 * P["dv_wav"]), 2.0),
                      (dict(tau_cr=3.0, a_now=5.0), 3.0 ** -6 * 5.0 ** -4)):
        args = dict(tau_cr=1.0, a_now=1.0, dv_wav=P["dv_wav"])
        args.update(kw)
        p = O.gw_spectrum(A, nb, P["dspa"], P["dwav"], args["tau_cr"], args["a_now"], args["dv_wav"])["power"]
        assert abs(p[b[0]] / base[b[0]] / ratio - 1) < 1e-14, kw


def test_gw_power_of_single_modes_is_cubic_in_qmid():
    """One unit mode in bin b gives power_b = qmid_b^3 fact / denom_b (denom_b = the lattice
    vectors in the bin, by brute force): two modes in different bins compare as the cube of their
    bin centres, and the absolute value matches fact from the constants of potential.f90:13-19
    (Omega_GW h^2 normalisation: 1/(3 (BigH/M*)^2) for the critical density in units of M*^2
    (H0/h)^2, a_now^-4 redshift, tau_cr^-6 and (2 pi)^-2 of the reference's units)."""
    N, nb = 16, 8
    tau_cr, a_now = 2.5, 3.0e3
    P = phys(N)
    dq = PI_T / P["dspa"] / nb
    idx = np.arange(N)
    n = np.where(idx < N // 2, idx, idx - N)
    nz, ny, nx = np.meshgrid(n, n, n, indexing="ij")
    q = np.sqrt((nx * nx + ny * ny + nz * nz).astype(float)) * P["dwav"]
    b_of = np.floor(q / dq + 0.5 + 0.5).astype(int)
    vals = {}
    for mode in ((1, 0, 0), (3, 2, 1)):
        A = _single_mode_state(N, mode, 1.0)
        s = O.gw_spectrum(A, nb, P["dspa"], P["dwav"], tau_cr, a_now, P["dv_wav"])
        qm = np.sqrt(sum(m * m for m in mode)) * P["dwav"]
        b = int(np.floor(qm / dq + 0.5 + 0.5))
        denom = int(np.count_nonzero((b_of == b) & (q > 0)))
        assert s["denom"][b - 1] == denom and s["numer"][b - 1] == 1.0
        qmid = (b - 0.5) * dq
        assert abs(s["qmid"][b - 1] / qmid - 1) < 1e-14
        vals[b] = (s["power"][b - 1], qmid, denom)
    (p1, q1, d1), (p2, q2, d2) = vals.values()
    assert abs((p1 * d1) / (p2 * d2) / (q1 / q2) ** 3 - 1) < 1e-13
    twopi = 2 * PI_T
    fact = (P["dv_wav"] / (tau_cr ** 6 * a_now ** 4)) / (3 * (BIGH_GEV / MSTAR_GEV) ** 2 * twopi ** 2)
    assert abs(p1 / (q1 ** 3 / d1 * fact) - 1) < 1e-13


def test_derived_units_from_physics():
    """settings.f90:274-276 against the physics they encode: a_now T_cmb = T_cr (g_rel /
    g_s,now)^(1/3) with T_cr = sqrt(6) sigma (entropy conservation, a_cr = 1); tau_cr = 1/H(T_cr)
    with the radiation-era Friedmann rate H^2 = (pi^2 g_rel T^4 / 30)/3 in units of M*;
    sigma_cr = sigma tau_cr.  The reference's truncated pi leaves ~1e-11 of difference against
    numpy's pi in tau_cr."""
    sigma, g_rel = 1.3e-9, 106.75
    d = O.derive(sigma, g_rel, 1e-30, 20.0, 2.0, 64)
    t_cr = np.sqrt(6.0) * sigma
    t_cmb = TCMB_K / GEV_IN_K / MSTAR_GEV
    assert abs(d["a_now"] * t_cmb / (t_cr * (g_rel / G_REL_NOWS) ** (1 / 3)) - 1) < 1e-14
    H = np.sqrt(np.pi ** 2 * g_rel * t_cr ** 4 / 30 / 3)
    assert abs(d["tau_cr"] * H - 1) < 1e-10
    assert d["sigma_cr"] == d["tau_cr"] * sigma
    assert d["dspa"] == 2.0 * 20.0 / 64 and abs(d["dwav"] * d["dspa"] * 64 / (2 * PI_T) - 1) < 1e-15


def test_energy_gradient_term_plane_wave():
    """For phi = a exp(i k.x) with lattice wave numbers k_d = 2 pi n_d / N the forward
    differences give sum_x sum_d |phi(x+e_d) - phi(x)|^2 / dspa^2 = N^3 |a|^2 sum_d 2 (1 - cos k_d)
    / dspa^2, times dv_spa = dspa^3 (a wrong neighbour, a missing axis or the 1/dspa^2 scale
    fails it)."""
    N, dspa = 12, 0.3
    a = 0.4 - 0.7j
    for n in ((1, 0, 0), (0, 2, 0), (1, 3, 5), (6, 0, 2)):
        z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
        k = [2 * np.pi * v / N for v in n]
        phi = np.ascontiguousarray(a * np.exp(1j * (k[0] * x + k[1] * y + k[2] * z)))
        e = O.energy(phi, np.zeros_like(phi), 1.5, dspa, SQ_HALF, 1e-30, True)
        exp = N ** 3 * abs(a) ** 2 * sum(2 * (1 - np.cos(kd)) for kd in k) / dspa ** 2 * dspa ** 3
        assert abs(e[1] / exp - 1) < 1e-12, n


def test_maxion2_branches():
    """potential.f90:62-79: kappa^4 c0 below thrm12 = kappa (cT/c0)^(1/n), kappa^4 cT (kappa/T)^n
    above; the two meet at thrm12 (continuity), and on the power-law side d ln m^2 / d ln T =
    -n = -6.68."""
    for kappa in (0.05, 1.0, 1.37, 3.0):
        k4 = kappa ** 4
        t12 = kappa * 6.26 ** (1 / 6.68)
        assert abs(O.maxion2(0.5 * t12, kappa) / k4 - 1) < 1e-15
        assert O.maxion2(0.0, kappa) == O.maxion2(0.5 * t12, kappa)
        lo, hi = O.maxion2(t12 * (1 - 1e-9), kappa), O.maxion2(t12 * (1 + 1e-9), kappa)
        assert abs(hi / lo - 1) < 1e-7 and lo == O.maxion2(0.0, kappa)
        t = 3.0 * t12
        h = 1e-5
        slope = (np.log(O.maxion2(t * (1 + h), kappa)) - np.log(O.maxion2(t * (1 - h), kappa))) / (
            np.log(1 + h) - np.log(1 - h))
        assert abs(slope + 6.68) < 1e-6
        assert abs(O.maxion2(t, kappa) / (k4 * 6.26 * (kappa / t) ** 6.68) - 1) < 1e-14
    # msquared (potential.f90:34-41): zero at T_cr = sqrt 6, -2 at T = 0
    assert abs(O.msquared(2.449489742783178)) < 1e-15 and abs(O.msquared(0.0) + 2) < 1e-15


# ------------------------------------------------------------------ plaquette windings

def _windings(code):
    """(z, y, x, 3) windings n_o from the oracle's face codes."""
    return np.stack([((code.astype(np.int32) >> (3 * o)) & 7) - 2 for o in range(3)], axis=-1)


def test_plaquettes_uniform_phase_is_empty():
    N = 8
    phi = np.full((N, N, N), np.exp(0.7j))
    code, cnt, rec = O.plaquettes(phi)
    assert len(rec) == 0 and all(np.all(v == 0) for v in cnt.values())
    assert np.all(_windings(code) == 0)


@pytest.mark.parametrize("axis,sign", [(2, 1), (2, -1), (0, 1), (1, -1)])
def test_plaquettes_straight_string(axis, sign):
    """A straight string along `axis` through a non-lattice point: phi = (U - u0) + i sign (V - v0)
    in the plane (U, V) normal to it, with (U, V, axis) cyclic.  Exactly one face per plane
    perpendicular to the string is pierced, all of the orientation whose normal is `axis`, with
    winding `sign` (counterclockwise about the axis is +1), and it is the face containing the
    string; every such face also passes the reference's quadrant test."""
    N = 12
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    X = [x, y, z]
    u_ax, v_ax = (axis + 1) % 3, (axis + 2) % 3
    u0, v0 = 5.3, 6.6
    phi = np.ascontiguousarray((X[u_ax] - u0) + 1j * sign * (X[v_ax] - v0))
    code, cnt, rec = O.plaquettes(phi)
    o_normal = {2: 0, 0: 1, 1: 2}[axis]  # faces (x,y), (y,z), (z,x) have normals z, x, y
    inner = rec[(np.abs(rec[:, u_ax] - 1 - 5) == 0) & (np.abs(rec[:, v_ax] - 1 - 6) == 0)]
    assert len(inner) == N and np.all(inner[:, 3] == o_normal + 1) and np.all(inner[:, 4] == sign)
    # the periodic seams of this non-periodic field carry the compensating windings (net 0)
    assert np.all(cnt["net"] == 0)
    n = _windings(code)
    q = np.stack([(code.astype(np.int32) >> (9 + o)) & 1 for o in range(3)], axis=-1)
    zz, yy, xx = inner[:, 2] - 1, inner[:, 1] - 1, inner[:, 0] - 1
    assert np.all(q[zz, yy, xx, o_normal] == 1)
    assert np.all(n[zz, yy, xx, o_normal] == sign)


def _gauss_residual(n):
    """Windings out of the six faces of every cell: n0(s+z) - n0(s) + n1(s+x) - n1(s) + n2(s+y) - n2(s)."""
    return (np.roll(n[..., 0], -1, axis=0) - n[..., 0] + np.roll(n[..., 1], -1, axis=2) - n[..., 1]
            + np.roll(n[..., 2], -1, axis=1) - n[..., 2])


def test_plaquettes_small_loop_closes():
    """A string loop of radius 3.2 in the plane z = z0: phi = (rho - R) + i (z - z0) with rho the
    distance from the loop's axis (the jump of z - z0 at the periodic seam z = N - 1/2 makes a
    second loop there).  The pierced faces have normals x or y only (the loops lie in xy
    planes), lie within 1.5 cells of one of the two circles, the windings of each orientation
    cancel, and no cell has a net outflow (strings do not end)."""
    N, R = 16, 3.2
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    cx, cy, cz = 7.4, 7.7, 8.5
    rho = np.hypot(x - cx, y - cy)
    phi = np.ascontiguousarray((rho - R) + 1j * (z - cz))
    code, cnt, rec = O.plaquettes(phi)
    assert cnt["n_pierced"][0] == 0 and cnt["n_pierced"][1] > 6 and cnt["n_pierced"][2] > 6
    assert np.all(cnt["net"] == 0)
    # face centres: o=1 (y,z) at (x, y+1/2, z+1/2), o=2 (z,x) at (x+1/2, y, z+1/2)
    off = {2: (0.0, 0.5, 0.5), 3: (0.5, 0.0, 0.5)}
    near = 0
    for r in rec:
        dx, dy, dz = off[int(r[3])]
        px, py, pz = r[0] - 1 + dx, r[1] - 1 + dy, r[2] - 1 + dz
        assert abs(np.hypot(px - cx, py - cy) - R) < 1.5
        assert abs(pz - cz) < 1.0 or abs(pz - (N - 0.5)) < 1.0
        near += abs(pz - cz) < 1.0
    assert near > 12
    assert np.all(_gauss_residual(_windings(code)) == 0)


def test_plaquettes_random_field_matches_phase_winding():
    """On a random field the sign-based count equals the winding from the phases themselves:
    round(sum of the four wrapped differences of np.angle / 2 pi); cells obey Gauss's law; the
    quadrant test matches a direct restatement of defects_3d.f90:856-877."""
    N = 12
    rng = np.random.default_rng(8)
    phi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    code, cnt, rec = O.plaquettes(phi)
    th = np.angle(phi)
    n = _windings(code)
    axes = {"x": 2, "y": 1, "z": 0}
    for o, (u, v) in enumerate((("x", "y"), ("y", "z"), ("z", "x"))):
        c0 = th
        c1 = np.roll(th, -1, axis=axes[u])
        c2 = np.roll(c1, -1, axis=axes[v])
        c3 = np.roll(th, -1, axis=axes[v])
        tot = 0
        for a, b in ((c0, c1), (c1, c2), (c2, c3), (c3, c0)):
            d = b - a
            tot = tot + (d - 2 * np.pi * np.round(d / (2 * np.pi)))
        w = np.round(tot / (2 * np.pi)).astype(int)
        assert np.array_equal(w, n[..., o]), o
        quads = [np.ceil(2 * q / PI_T).astype(int) + 2 for q in (c0, c1, c2, c3)]
        filled = sum(((quads[0] == m) | (quads[1] == m) | (quads[2] == m) | (quads[3] == m)).astype(int)
                     for m in range(1, 5))
        assert cnt["n_quadrant"][o] == np.count_nonzero(filled >= 3)
        assert cnt["n_pierced"][o] == np.count_nonzero(w) > 100
        assert cnt["n_plus"][o] == np.count_nonzero(w == 1) and cnt["n_minus"][o] == np.count_nonzero(w == -1)
    assert np.all(cnt["net"] == 0)
    assert np.all(_gauss_residual(n) == 0)
    assert len(rec) == sum(cnt["n_pierced"])


# ------------------------------------------------------------------ the reference's random stream

def test_mcg59_matches_mkl_golden_vectors():
    """The oracle's MCG59 against MKL's own draws (tests/golden/mcg59_mkl.npz, written by
    make_mcg59.py from the MKL runtime): uniforms equal bit for bit from five seeds (including
    the all-zero seed, which MKL starts at x = 1) and after five skip-ahead distances (up to
    ~4e8 draws); Box-Muller Gaussians from the same positions agree to 5e-8 relative.  MKL
    evaluates ln and sin to about 27 bits there (its draws differ from correctly rounded
    sqrt(-2 ln u1) sin(2 pi u2) by <= 1.2e-8 relative; a swapped pair, cos for sin or another
    pairing differs by O(1))."""
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "mcg59_mkl.npz"))
    for i, (w0, w1) in enumerate(d["seeds"]):
        for j, skip in enumerate(d["skips"]):
            u = O.mcg59_uniforms(int(w0), int(w1), int(skip), d["uniform"].shape[2])
            assert np.array_equal(u, d["uniform"][i, j]), (w0, w1, skip)
            g = O.mcg59_gauss(int(w0), int(w1), int(skip), d["gauss"].shape[2])  # the skip counts draws
            ref = d["gauss"][i, j]
            assert np.all(np.abs(g - ref) <= 5e-8 * np.maximum(np.abs(ref), 1.0)), (w0, w1, skip)


def test_mcg59_jump_is_the_recurrence():
    """Skip-ahead by p equals p steps of x <- 13^13 x mod 2^59 (Python integers), and the seed is
    (w0 + 2^32 w1) mod 2^59 with 0 mapped to 1."""
    a, m = 13 ** 13, 2 ** 59
    for w0, w1, p in ((3, 5, 0), (3, 5, 1), (99, 0, 977), (2 ** 32 - 1, 2 ** 31, 12345)):
        x0 = (w0 + 2 ** 32 * w1) % m or 1
        assert O.mcg59_seed(w0, w1) == x0
        x = x0
        for _ in range(p):
            x = a * x % m
        assert O.mcg59_jump(x0, p) == x
    assert O.mcg59_seed(0, 2 ** 27) == 1  # 2^59 mod 2^59 = 0 -> 1


def test_thermal_state_from_reference_stream_uses_its_draws():
    """With the reference's stream the thermal state's mode lin of field ip holds dv_wav times the
    Gaussians of draws 4 N^3 (ip - 1) + 4 lin .. +3 (initial.f90:102-126) times the Bose-Einstein
    factor: one mode of each field rebuilt from the stream's own Gaussians, through the transform
    of a single mode."""
    N, tau = 8, 0.3
    P = phys(N)
    phi, pi = O.ic_thermal(N, 0, tau, P["sigma_cr"], P["dwav"], P["dv_wav"], stream=(11, 22, 0))
    # back to Fourier space: FORWARD was applied, so the BACKWARD transform / N^3 inverts it
    modes = [O.fft3d(f, +1) / N ** 3 for f in (phi, pi)]
    for ip in (1, 2):
        for lin in (0, 5, 77, N ** 3 - 1):
            x, y, z = lin % N, (lin // N) % N, lin // N ** 2
            nx, ny, nz = [v if v < N // 2 else v - N for v in (x, y, z)]
            g = O.mcg59_gauss(11, 22, 4 * N ** 3 * (ip - 1) + 4 * lin, 2)
            f = O.thermal_factor(ip, nx * nx + ny * ny + nz * nz, tau, P["sigma_cr"], P["dwav"], P["dv_wav"])
            exp = (g[0] * P["dv_wav"]) * f + 1j * (g[1] * P["dv_wav"]) * f
            got = modes[ip - 1][z, y, x]
            assert abs(got - exp) <= 1e-12 * abs(exp), (ip, lin)


# ------------------------------------------------------------- reduced grid (degrade_spec > 0)

def _blocks(a, deg):
    """(N,N,N) -> (Nc, Nc, Nc, deg^3) view of the deg^3 blocks (z, y, x order inside a block)."""
    N = a.shape[0]
    Nc = N // deg
    return a.reshape(Nc, deg, Nc, deg, Nc, deg).transpose(0, 2, 4, 1, 3, 5).reshape(Nc, Nc, Nc, deg ** 3)


@pytest.mark.parametrize("ds", [1, 2])
def test_degrade_field_is_the_block_mean(ds):
    """data2_to_data_spec_degrade is the mean over each 2^ds cube: exact on small integers (every
    partial sum representable), and to 1e-15 on a random field."""
    N, deg = 16, 2 ** ds
    rng = np.random.default_rng(3)
    f = (rng.integers(-50, 50, (N, N, N)) + 1j * rng.integers(-50, 50, (N, N, N))).astype(np.complex128)
    assert np.array_equal(O.degrade_field(f, ds), _blocks(f, deg).sum(-1) / deg ** 3)
    g = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    ref = _blocks(g, deg).mean(-1)
    assert np.max(np.abs(O.degrade_field(g, ds) - ref)) <= 1e-15 * np.max(np.abs(ref))


def test_degrade_mask_marks_blocks_with_a_masked_cell():
    N = 16
    rng = np.random.default_rng(4)
    flags = np.zeros((N, N, N), np.int8)
    flags[rng.random((N, N, N)) < 0.02] = 4
    flags[rng.random((N, N, N)) < 0.1] |= 3  # string / wall bits alone mask nothing
    for ds in (1, 2):
        exp = np.where((_blocks(flags, 2 ** ds) & 4).any(-1), 4, 0).astype(np.int8)
        assert np.array_equal(O.degrade_mask(flags, ds), exp)


def test_degrade_gradient_reduces_to_the_gradient():
    """deg = 1: (0 + a) - b divided by 1 * dspa * 2 is data1_to_arr_dphi's (a - b) / (2 dspa)."""
    N = 12
    rng = np.random.default_rng(5)
    phi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    assert np.array_equal(O.degrade_gradient(phi, 0, 0.3), O.gradient(phi, 0.3))


@pytest.mark.parametrize("ds", [1, 2])
def test_degrade_gradient_plane_wave(ds):
    """phi = e^{i k.x}: phi(s + e_d) - phi(s - e_d) = 2 i sin(k_d dspa) phi(s), so the coarse
    component d is 2 i sin(k_d dspa) / (deg^3 2 dspa) times the block sum of phi, and the block
    sum is phi(block corner) prod_e sum_{j < deg} e^{i k_e j dspa}."""
    N, dspa, deg = 16, 0.35, 2 ** ds
    m = np.array([1, -3, 5])
    k = 2 * np.pi * m / (N * dspa)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    phi = np.exp(1j * dspa * (k[0] * x + k[1] * y + k[2] * z))
    g = O.degrade_gradient(phi, ds, dspa)
    Nc = N // deg
    cz, cy, cx = np.meshgrid(np.arange(Nc), np.arange(Nc), np.arange(Nc), indexing="ij")
    corner = np.exp(1j * dspa * deg * (k[0] * cx + k[1] * cy + k[2] * cz))
    geo = np.prod([np.exp(1j * k[e] * dspa * np.arange(deg)).sum() for e in range(3)])
    for d in range(3):
        exp = 2j * np.sin(k[d] * dspa) / (deg ** 3 * 2 * dspa) * corner * geo
        assert np.max(np.abs(g[d] - exp)) < 1e-12, d


def test_degrade_gradient_of_a_periodic_field_sums_to_zero():
    """Summed over the whole periodic lattice the central differences cancel."""
    N = 16
    rng = np.random.default_rng(6)
    phi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    g = O.degrade_gradient(phi, 1, 0.25)
    assert np.all(np.abs(g.sum(axis=(1, 2, 3))) < 1e-11)
    assert np.all(np.abs(g).max(axis=(1, 2, 3)) > 0.1)


def test_reduced_units():
    """init_grav / init_spectrum's reduced grid: spacing deg * dspa, N / deg sites, so the
    wavenumber unit and dv_wav are the lattice's (to rounding) and dv_spa is deg^3 times it."""
    P = phys(32, 0.3)
    for ds in (1, 2, 3):
        u = O.reduced_units(P, ds)
        deg = 2 ** ds
        assert u["ngrid"] == 32 // deg and u["dspa"] == 0.3 * deg
        assert abs(u["dwav"] / P["dwav"] - 1) < 4e-16
        assert abs(u["dv_wav"] / P["dv_wav"] - 1) < 1e-15
        assert abs(u["dv_spa"] / (deg ** 3 * 0.3 ** 3) - 1) < 1e-15


def test_symplectic_with_degrade_runs_the_reduced_grid():
    """orc_symplectic with degrade_spec = 1: the fields as without it (bit for bit), and the GW
    accumulators those of data1_to_arr_dphi_degrade + the accumulation on the N/2 grid at every
    kick, composed here from the pieces with the reduced units."""
    N, dt, n = 16, 0.05, 3
    P = phys(N)
    P["degrade_spec"] = 1
    phi0, pi0 = O.ic_flat(N, 7, 3.0, P["dwav"])
    phi, pi = phi0.copy(), pi0.copy()
    A = O.new_gw_state(N // 2)
    O.symplectic(phi, pi, A, 1.0, dt, n, P)
    phi1, pi1 = phi0.copy(), pi0.copy()
    O.symplectic(phi1, pi1, None, 1.0, dt, n, P)
    assert np.array_equal(phi, phi1) and np.array_equal(pi, pi1)
    u = O.reduced_units(P, 1)
    f, p = phi0.copy(), pi0.copy()
    B = O.new_gw_state(N // 2)
    t = 1.0
    O.drift(f, p, dt / 2)
    t += dt / 2
    for it in range(n):
        O.kick(f, p, t, dt, P["dspa"], P["sigma_cr"], P["kappa"], P["initial_loop"])
        O.gw_accumulate(O.degrade_gradient(f, 1, P["dspa"]), B, t, dt, u["dwav"], u["dv_spa"], P["sigma_cr"])
        O.drift(f, p, dt / 2 if it == n - 1 else dt)
        t += dt / 2 if it == n - 1 else dt
    assert np.array_equal(A, B)
    assert np.max(np.abs(A)) > 0


# ------------------------------------------- velocity_string_YY (the reference's VEL = 1 build)

def _moving_string(N, axis, vel, dspa, center=(7.3, 8.6), sign=1.0):
    """A straight string along `axis` through `center` of the two other axes (u, v) = the next
    two axes cyclically: phi = (U - u0) + i sign (V - v0) in cell units, and pi = -(vu + i sign vv)
    / dspa, the time derivative of that field moving with velocity (vu, vv) in units of c."""
    u_ax, v_ax = (axis + 1) % 3, (axis + 2) % 3
    idx = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")  # z, y, x
    coord = {0: idx[2], 1: idx[1], 2: idx[0]}
    phi = (coord[u_ax] - center[0]) + 1j * sign * (coord[v_ax] - center[1])
    pi = np.full((N, N, N), -(vel[0] + 1j * sign * vel[1]) / dspa)
    return phi.astype(np.complex128), pi.astype(np.complex128), u_ax, v_ax


@pytest.mark.parametrize("axis,sign", [(2, 1.0), (2, -1.0), (0, 1.0), (1, -1.0)])
def test_velocity_yy_moving_straight_string(axis, sign):
    """A straight string moving with (vu, vv) across its axis: the two faces normal to the axis
    are pierced (the other four see two quadrants) and give the velocity exactly (the field is
    linear, so the weights -- which sum to 1 -- do not matter)."""
    N, dspa, v = 16, 0.3, (0.31, -0.17)
    phi, pi, u_ax, v_ax = _moving_string(N, axis, v, dspa, sign=sign)
    c = [0, 0, 0]
    c[u_ax], c[v_ax], c[axis] = 7, 8, 5
    vel = O.velocity_string_yy(phi, pi, c[0], c[1], c[2], dspa)
    exp = np.zeros(3)
    exp[u_ax], exp[v_ax] = v
    assert np.all(np.abs(vel - exp) < 1e-12), (vel, exp)


def test_velocity_yy_uses_the_last_pierced_face():
    """The reference averages vpen(:, npen) -- the last pierced face's velocity, npen times --
    not the faces' mean: with pi changing along the string between the two pierced faces the
    result is the upper face's velocity."""
    N, dspa = 16, 0.25
    phi, pi, _, _ = _moving_string(N, 2, (0.2, 0.1), dspa)
    z = np.arange(N)[:, None, None]
    pi = pi - (0.05 * (z - 5)) / dspa  # vx grows by 0.05 per plane
    vel = O.velocity_string_yy(phi, pi.astype(np.complex128), 7, 8, 5, dspa)
    assert abs(vel[0] - 0.25) < 1e-12 and abs(vel[1] - 0.1) < 1e-12 and abs(vel[2]) < 1e-12
    assert abs(vel[0] - 0.225) > 1e-3  # not the two faces' mean


def test_velocity_yy_without_a_pierced_face():
    N = 8
    phi = np.full((N, N, N), 0.6 + 0.8j)
    pi = np.zeros((N, N, N), np.complex128)
    assert np.array_equal(O.velocity_string_yy(phi, pi, 3, 3, 3, 0.25), np.full(3, 1e-10))


def test_strings_vel_yy_keeps_the_fits():
    """vel_yy changes only the velocity columns of the records (and the velocity sums)."""
    N = 16
    P = phys(N)
    phi, _ = O.ic_flat(N, 5, 6.0, P["dwav"])
    rng = np.random.default_rng(2)
    pi = 0.2 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    flags, cnt = O.classify(phi)
    assert cnt[0] > 10
    r0, s0 = O.strings(phi, pi, flags, 3, 4, P["dspa"])
    r1, s1 = O.strings(phi, pi, flags, 3, 4, P["dspa"], vel_yy=True)
    assert np.array_equal(r0[:, :8], r1[:, :8]) and s0[0] == s1[0]
    zz, yy, xx = np.nonzero(flags & 1)
    for k in range(0, len(xx), 7):
        assert np.array_equal(r1[k, 8:], O.velocity_string_yy(phi, pi, xx[k], yy[k], zz[k], P["dspa"]))
