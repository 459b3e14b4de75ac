"""The oracle's Coriolis term in rlm space (init_sum_coriolis_rlm,
sum_rot_coriolis_rlm_10, inner_core_rot_z_coriolis_rlm;
MHD_src/sph_MHD/t_coriolis_terms_rlm.f90, interact_coriolis_rlm.f90,
sum_coriolis_terms_rlm.f90), pinned three ways:

* against the reference's own routines compiled from source into
  oracle/_ref: partner addresses equal, integrals and coefficients within a
  few ulp (the reference forms the integrals from prime factorisations and
  products of square roots; the oracle takes their closed forms), the sums
  within round-off of that;
* the Gaunt integrals against Gauss-Legendre quadrature of
  x P_(l-1)^m P_l^m with the Schmidt functions;
* by physics: the poloidal part of the curl equals the toroidal scalar of
  the grid-space Coriolis force -c (z x u) taken through the whole backward
  transform, the grid products and the forward transform (exact weights,
  u band-limited so that z x u is); a solid-body rotation gives none.
"""
import numpy as np
import pytest

import oracle as O
from oracle import ref

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


def _plan(L, nr, nth, nphi, exact=False):
    r = O.chebyshev_radius(nr, 0.35 / 0.65, 1.0 / 0.65)
    return O.OraclePlan(L, nr, nth, nphi, r, exact_weights=exact, full_tables=False), r


@needs_ref
@pytest.mark.parametrize("L,nth", [(15, 24), (31, 48)])
def test_tables_vs_reference(L, nth):
    orc, r = _plan(L, 3, nth, 2 * nth)
    rp = ref.RefPlan(L, 3, nth, r)
    a, b = orc.coriolis_tables(), rp.coriolis_tables()
    np.testing.assert_array_equal(a["jgi"], b["jgi"])
    np.testing.assert_array_equal(a["jei"], b["jei"])
    for k in ("gi", "ei", "sw", "tw"):
        scale = np.abs(b[k]).max()
        err = np.abs(a[k] - b[k]).max()
        assert err <= 8 * np.finfo(float).eps * scale, (k, err, scale)
        # zero where the reference is zero
        assert np.all(a[k][b[k] == 0] == 0), k


@needs_ref
@pytest.mark.parametrize("kr_icb", [0, 1])
def test_sum_vs_reference(kr_icb):
    L, nr, nth = 31, 5, 48
    orc, r = _plan(L, nr, nth, 96)
    rp = ref.RefPlan(L, nr, nth, r)
    ncomp = 9
    rng = np.random.default_rng(7 + kr_icb)
    sp = rng.standard_normal(ncomp * orc.jmax * nr)
    for iv, iw, c in ((3, 0, 2.5), (0, 6, -0.75)):
        got = orc.sum_coriolis_rlm(ncomp, iv, iw, c, sp, kr_icb=kr_icb)
        want = rp.sum_coriolis(ncomp, iv, iw, c, sp, kr_icb=kr_icb)
        scale = np.abs(want).max()
        assert np.abs(got - want).max() <= 1e-14 * scale
        if kr_icb:
            # the inner-core terms replace the degree-one poloidal values at the ICB
            # (reference: dp_rot_cor_rlm(i10c) = 0)
            j10 = orc.table("lstack")[L] + 1
            assert got[0, j10] == 0.0 and want[0, j10] == 0.0


def test_gaunt_integrals_by_quadrature():
    """gi(j,1) = 4 sqrt(l^2-m^2)/((2l-1)(2l+1)) is int x P_(l-1)^m P_l^m dx
    times 2 for m = 0 (normki) with Schmidt P (int (P_l^m)^2 = 2(2-d_m0)/(2l+1));
    gi(j,2) likewise with l+1."""
    L = 20
    orc, _ = _plan(L, 2, 32, 64)
    t = orc.coriolis_tables()
    idx = orc.table("idx_j")
    x, w = O.gauss_points(48, exact=True)
    th = np.arccos(x)
    P = {}
    for m in range(L + 1):
        P[m] = np.array([O.schmidt_legendres_m(L, m, tt)[0] for tt in th])  # (n, L+1)
    for j in range(orc.jmax):
        l, m = idx[1, j], idx[2, j]
        am = abs(m)
        fac = 2.0 if m == 0 else 1.0
        g1 = fac * np.sum(w * x * P[am][:, l - 1] * P[am][:, l]) if (l >= 2 and am <= l - 1) else 0.0
        g2 = fac * np.sum(w * x * P[am][:, l + 1] * P[am][:, l]) if (1 <= l and l + 1 <= L) else 0.0
        assert abs(t["gi"][0, j] - g1) < 1e-14, (l, m)
        assert abs(t["gi"][1, j] - g2) < 1e-14, (l, m)


def _random_velocity(orc, rng, nr):
    jm = orc.jmax
    idx = orc.table("idx_j")
    u = rng.standard_normal((nr, jm, 3))
    u[:, idx[1] >= orc.L, :] = 0.0   # z x u stays within degree L
    u[:, idx[1] == 0, :] = 0.0       # no degree-0 poloidal velocity
    return u


def test_rot_poloidal_equals_toroidal_scalar_of_grid_coriolis():
    L, nr, nth = 15, 4, 24
    orc, _ = _plan(L, nr, nth, 2 * nth, exact=True)
    rng = np.random.default_rng(3)
    u = _random_velocity(orc, rng, nr)
    ncomp = 6
    sp = np.zeros((nr, orc.jmax, ncomp))
    sp[:, :, 0:3] = u
    sp[:, :, 3] = u[:, :, 2]   # vorticity poloidal = velocity toroidal
    coef = 1.7
    grid = orc.backward(1, 0, np.ascontiguousarray(sp[:, :, 0:3]).ravel())
    frc = orc.nonlinear_rtp({"velo": 0, "coriolis": 0}, {"coriolis": coef}, 3, grid, 3)
    F = orc.forward(1, 0, frc).reshape(nr, orc.jmax, 3)
    rot = orc.sum_coriolis_rlm(ncomp, 0, 3, coef, sp.ravel()).reshape(2, nr, orc.jmax)
    scale = np.abs(F[:, :, 2]).max()
    assert scale > 1.0
    assert np.abs(rot[0] - F[:, :, 2]).max() <= 1e-13 * scale


def test_solid_body_rotation_has_no_curl():
    """u = w z x r (toroidal l = 1, m = 0, T = w r^2): the Coriolis force is a
    gradient, its curl vanishes."""
    L, nr = 7, 4
    orc, r = _plan(L, nr, 12, 24)
    ncomp = 6
    sp = np.zeros((nr, orc.jmax, ncomp))
    j10 = orc.table("lstack")[L] + 1
    wr = 0.8
    sp[:, j10, 2] = wr * r ** 2          # velocity toroidal
    sp[:, j10, 3] = wr * r ** 2          # vorticity poloidal
    sp[:, j10, 4] = 2 * wr * r           # its radial derivative
    rot = orc.sum_coriolis_rlm(ncomp, 0, 3, 1.0, sp.ravel())
    assert np.abs(rot).max() < 1e-14
