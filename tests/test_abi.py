"""The C-ABI library loads, exports every entry point include/axgrav.h declares, and its
host-only entry points (axg_derive, axg_get_dtau) agree with the oracle.  No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
from axion_grav_amd import capi
from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "axgrav.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(axg_[a-z_0-9]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert sorted(capi.EXPORTS) == declared()


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    for name in declared():
        assert hasattr(L, name), name


def test_params_struct_layout():
    # the C struct: int64 + 10 int32 + 11 doubles + degrade_spec + string_vel + 6 int32 reserved
    assert C.sizeof(capi.Params) == 8 + 10 * 4 + 11 * 8 + 8 * 4
    assert capi.Params.dspa.offset == 48
    assert capi.Params.degrade_spec.offset == 8 + 10 * 4 + 11 * 8
    assert capi.Params.string_vel.offset == 8 + 10 * 4 + 11 * 8 + 4


@pytest.mark.parametrize("sigma,g_rel,kappa,tau_fin,boxsize,ngrid", [
    (0.0750264, 90.0, 1e-30, 16.0, 1.0, 64),
    (1e-3, 100.0, 1e-2, 40.0, 2.0, 512),
    (3e-4, 61.75, 5e-3, 25.0, 1.5, 96),
])
def test_derive_matches_oracle_bitwise(sigma, g_rel, kappa, tau_fin, boxsize, ngrid):
    P, extra = capi.derive(sigma, g_rel, kappa, tau_fin, boxsize, ngrid)
    o = O.derive(sigma, g_rel, kappa, tau_fin, boxsize, ngrid)
    for k in ("a_now", "tau_cr", "sigma_cr", "dspa", "dwav", "dv_spa", "dv_wav"):
        assert getattr(P, k) == o[k], k
    for k in ("zeta", "width_cr", "thickness_cr"):
        assert extra[k] == o[k], k


def test_derive_closed_forms():
    # sigma = 1/(6 pi_t sqrt(0.5)) and g_rel = 90 give sigma_cr^2 = 1/2 (lambda = 2)
    pit = 3.14159265359
    sigma = 1 / (6 * pit * np.sqrt(0.5))
    P, _ = capi.derive(sigma, 90.0, 1e-30, 16.0, 1.0, 64)
    assert abs(P.sigma_cr ** 2 - 0.5) < 1e-15
    assert P.dspa == 0.25 and P.dv_spa == 0.25 ** 3
    assert P.dwav == 2 * pit / 0.25 / 64


@pytest.mark.parametrize("t1,t2,eps", [(1.0, 6.0, 1.0), (6.0, 11.0, 0.5), (1.0, 16.0, 0.37)])
def test_get_dtau_matches_oracle(t1, t2, eps):
    P = capi.make_params(64, 0.25, float(np.sqrt(0.5)), eps_dtau=eps)
    dt, n = capi.get_dtau(P, t1, t2)
    odt, on = O.get_dtau(t1, t2, P.sigma_cr, P.dwav, 64, eps)
    assert (dt, n) == (odt, on)
    # evolve.f90:17-21 by hand
    om = np.sqrt((P.dwav * 64) ** 2 + (2.449489742783178 * t2 * P.sigma_cr) ** 2)
    assert n == int(np.ceil((t2 - t1) / (eps / om)))
    assert dt == (t2 - t1) / n


def test_fixed_step_key():
    P = capi.make_params(64, 0.25, float(np.sqrt(0.5)), dtau_fixed=0.05)
    dt, n = capi.get_dtau(P, 1.0, 6.0)
    assert n == 100 and abs(dt - 0.05) < 1e-15


def test_create_fails_loudly_on_bad_input():
    P = capi.make_params(50, 0.25, 0.7)  # 50 = 2 * 5^2 is not 2^a 3^b
    with pytest.raises(capi.AxgError, match="ngrid"):
        capi.Lattice(P)
