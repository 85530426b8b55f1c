"""The FP64 hydrostatic background of the stratified workload
(smaugamd.configs.hydrostatic_profile, used by configs.stratified).

The reference reads its stratified atmosphere from a prepared VAC file; the
rebuilt workload integrates dp/dz = -rho g, p = rho R T for a tanh
chromosphere-corona temperature profile.  The checks pin it against what the
mathematics fixes: the base density, the isothermal closed form
p = p0 exp(-g z / (R T)), an adaptive-quadrature evaluation of
ln p(z) = ln p0 - (g/R) int_0^z dz'/T(z'), and the finite-difference residual
of dp/dz + rho g on a fine grid."""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "smaug-all_amd"))
from smaugamd.configs import RGAS, hydrostatic_profile, stratified  # noqa: E402

G = 274.0
RHO0 = 2.7e-4


def _temp(z, T_lo=6.0e3, T_hi=1.0e6, z_tr=2.0e6, w_tr=0.2e6):
    return T_lo + (T_hi - T_lo) * 0.5 * (1.0 + math.tanh((z - z_tr) / w_tr))


def test_base_density_and_gas_law():
    z = np.linspace(0.0, 6.0e6, 97)
    rho, p, T = hydrostatic_profile(z)
    assert rho[0] == pytest.approx(RHO0, rel=1e-14)
    np.testing.assert_allclose(p, rho * RGAS * T, rtol=1e-14)
    np.testing.assert_allclose(T, [_temp(zz) for zz in z], rtol=1e-14)
    # pressure and density fall monotonically with height
    assert np.all(np.diff(p) < 0)
    assert np.all(np.diff(rho) < 0)


def test_isothermal_closed_form():
    T0 = 1.0e4
    z = np.linspace(0.0, 3.0e6, 41)
    rho, p, T = hydrostatic_profile(z, T_lo=T0, T_hi=T0)
    p0 = RHO0 * RGAS * T0
    np.testing.assert_allclose(p, p0 * np.exp(-G * z / (RGAS * T0)), rtol=1e-12)
    np.testing.assert_allclose(rho, RHO0 * np.exp(-G * z / (RGAS * T0)), rtol=1e-12)


def test_transition_region_against_adaptive_quadrature():
    from scipy.integrate import quad

    z = np.array([0.0, 1.0e5, 1.5e6, 1.9e6, 2.0e6, 2.1e6, 2.5e6, 3.0e6, 6.0e6])
    # unsorted input: the profile is returned in the caller's order
    perm = np.array([4, 0, 8, 2, 6, 1, 3, 7, 5])
    rho, p, T = hydrostatic_profile(z[perm])
    p0 = RHO0 * RGAS * _temp(0.0)
    for zz, pp in zip(z[perm], p):
        I, _ = quad(lambda x: 1.0 / _temp(x), 0.0, zz, epsabs=0.0, epsrel=1e-13, limit=400)
        expect = p0 * math.exp(-G / RGAS * I)
        assert pp == pytest.approx(expect, rel=1e-7), zz


def test_hydrostatic_residual_fine_grid():
    h = 1.0e3
    z = np.arange(0.0, 6.0e6 + h / 2, h)
    rho, p, _ = hydrostatic_profile(z)
    dpdz = (p[:-4] - 8.0 * p[1:-3] + 8.0 * p[3:-1] - p[4:]) / (12.0 * h)
    res = np.abs(dpdz + G * rho[2:-2]) / (G * rho[2:-2])
    assert res.max() < 1e-5


def test_stratified_background_is_the_profile():
    s = stratified(nz=24, nx=8, ny=8)
    n0, n1, n2 = s.n
    rho, p, _ = hydrostatic_profile(s.pos[0])
    bb = s.meta["bz_code"]
    rho_b = s.bg[1].reshape(n2, n1, n0)
    e_b = s.bg[0].reshape(n2, n1, n0)
    for k in (0, n2 // 2, n2 - 1):
        for j in (0, n1 - 1):
            np.testing.assert_allclose(rho_b[k, j], rho, rtol=1e-15)
            np.testing.assert_allclose(e_b[k, j] - 0.5 * bb * bb, p / (s.gamma - 1.0),
                                       rtol=1e-12)
    assert s.g[0] == -G
