"""The oracle's backward transform at the poles and the centre, pinned.

The reference holds no vectors for this path, so the oracle is checked
against what the mathematics fixes:

* the pole tables: Schmidt P_l^0 = 1 at theta = 0 and (-1)^l at theta = pi,
  dP_l^1/dtheta = sqrt(l(l+1)/2) at theta = 0 and (-1)^l sqrt(l(l+1)/2)
  at theta = pi (the Schmidt P_l^1 is sqrt(2/(l(l+1))) sin(theta) P_l'(cos));
* the pole rows: the limit of the full-grid backward (itself pinned by the
  oracle tests) at phi = 0 towards theta = 0 and pi, from the Gauss rows
  next to the poles of two grids (Richardson, error O(theta^2)), mapped to
  (x, y, z) = (v_theta, v_phi, v_r) in the north, (-v_theta, v_phi, -v_r)
  in the south: a wrong sign, order, component or scale is an O(1) miss;
* uniform fields: a degree-1 poloidal spectrum S = c r^2 is a uniform field
  of size 2c along z (order 0), x (cos 1) or y (sin 1), the same vector at
  both poles and at every radius.

The centre row (SS/schmidt_b_trans_at_center.f90:45-115) is pinned against
the reference's own routine, compiled from its source into oracle/_ref
(oracle/Makefile.ref): bit for bit, vector and scalar components, with and
without a prior content of the row.  (The reference's code takes x from the
order -1 (sin) coefficient and y from order +1, both times -2 / r_c^2; its
header comment names the cos mode for x.  The oracle follows the code.)
"""
import numpy as np
import pytest

import oracle as O
from oracle import ref
from kemorin_mhd_amd import ShellConfig, gaussian, rlm_l_m

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


def test_pole_tables_closed_forms():
    L = 40
    l = np.arange(L + 1)
    p0, _ = O.schmidt_legendres_m(L, 0, 0.0)
    ps, _ = O.schmidt_legendres_m(L, 0, 4.0 * np.arctan(1.0))
    _, d0 = O.schmidt_legendres_m(L, 1, 0.0)
    _, ds = O.schmidt_legendres_m(L, 1, 4.0 * np.arctan(1.0))
    sign = (-1.0) ** l
    np.testing.assert_allclose(p0, 1.0, rtol=0, atol=1e-13)
    np.testing.assert_allclose(ps, sign, rtol=0, atol=1e-13)
    ref = np.sqrt(l * (l + 1) / 2.0)
    np.testing.assert_allclose(d0[1:], ref[1:], rtol=1e-13)
    np.testing.assert_allclose(ds[1:], sign[1:] * ref[1:], rtol=1e-13)


def _plan(L, nr, nth):
    c = ShellConfig(L, nr, nth, 2 * L + 2)
    return c, O.OraclePlan(L, nr, nth, c.nphi, c.radius(), full_tables=False)


@pytest.mark.parametrize("L", [1, 6])
def test_pole_rows_are_the_limit_of_the_grid(L):
    nr, nv, ns = 3, 2, 1
    ncomp = 3 * nv + ns
    c0, _ = _plan(L, nr, 400)
    sp = gaussian(31 + L, c0.n_rj(ncomp))
    rows = []
    for nth in (400, 800):
        c, orc = _plan(L, nr, nth)
        v = orc.backward(nv, ns, sp).reshape(ncomp, c.nphi, nth, nr)
        col = orc.table("colat")
        # Gauss points ascend in cos(theta): row nth-1 is next to the north
        # pole, row 0 next to the south pole; phi index 0 is phi = 0
        rows.append((col[-1], v[:, 0, -1, :], np.pi - col[0], v[:, 0, 0, :]))
    (ta, fa, sa, ga), (tb, fb, sb, gb) = rows
    north = (tb * fa - ta * fb) / (tb - ta)
    south = (sb * ga - sa * gb) / (sb - sa)
    pole = orc.pole_backward(nv, ns, sp).reshape(ncomp, 2 * nr + 1)
    exp_n = np.empty_like(north)
    exp_s = np.empty_like(south)
    for f in range(nv):
        r, t, p = 3 * f, 3 * f + 1, 3 * f + 2
        exp_n[3 * f], exp_n[3 * f + 1], exp_n[3 * f + 2] = north[t], north[p], north[r]
        exp_s[3 * f], exp_s[3 * f + 1], exp_s[3 * f + 2] = -south[t], south[p], -south[r]
    exp_n[3 * nv:] = north[3 * nv:]
    exp_s[3 * nv:] = south[3 * nv:]
    scale = np.abs(np.concatenate([exp_n, exp_s], axis=1)).max(axis=1, keepdims=True)
    # the residual of the limit is O((L theta_1)^2), about 1e-4 here; a
    # wrong term is off by O(1)
    err = np.abs(np.concatenate([pole[:, :nr] - exp_n, pole[:, nr:2 * nr] - exp_s], 1))
    print(f"L={L}: pole rows within {err.max() / scale.max():.1e} of the grid limit")
    assert err.max() <= 2e-3 * scale.max()


@pytest.mark.parametrize("order,axis", [(0, 2), (1, 0), (-1, 1)])
def test_uniform_field_at_the_poles(order, axis):
    L, nr, nth = 4, 5, 8
    c, orc = _plan(L, nr, nth)
    ls, ms = rlm_l_m(L)
    j = int(np.nonzero((ls == 1) & (ms == order))[0][0])
    r = c.radius()
    cst = 0.75
    sp = np.zeros((nr, c.jmax, 3))
    sp[:, j, 0] = cst * r ** 2          # S
    sp[:, j, 1] = 2.0 * cst * r         # dS/dr
    pole = orc.pole_backward(1, 0, sp.ravel()).reshape(3, 2 * nr + 1)
    want = np.zeros(3)
    want[axis] = 2.0 * cst
    for i in range(2 * nr):
        np.testing.assert_allclose(pole[:, i], want, rtol=0, atol=1e-13)


def test_centre_row_writes_only_vector_components():
    L, nr, nv, ns = 5, 4, 2, 2
    ncomp = 3 * nv + ns
    c, orc = _plan(L, nr, 8)
    sp = gaussian(7, c.n_rj(ncomp))
    init = np.full(ncomp * (2 * nr + 1), 7.0)
    off = orc.pole_backward(nv, ns, sp, center=False, v_pole=init).reshape(ncomp, -1)
    on = orc.pole_backward(nv, ns, sp, center=True, v_pole=init).reshape(ncomp, -1)
    assert (off[:, 2 * nr] == 7.0).all()
    assert (on[3 * nv:, 2 * nr] == 7.0).all()
    assert (on[:3 * nv, 2 * nr] != 7.0).all()
    np.testing.assert_array_equal(on[:, :2 * nr], off[:, :2 * nr])
    # the degree-1 poloidal coefficients at the innermost radius, x 2 / r_c^2
    ls, ms = rlm_l_m(L)
    a = 1.0 / c.radius()[0]
    s = sp.reshape(nr, c.jmax, ncomp)[0]
    for f in range(nv):
        for comp, m, sgn in ((0, -1, -1.0), (1, 1, -1.0), (2, 0, 1.0)):
            j = int(np.nonzero((ls == 1) & (ms == m))[0][0])
            assert on[3 * f + comp, 2 * nr] == pytest.approx(sgn * 2.0 * s[j, 3 * f] * a * a,
                                                             rel=1e-15)


@needs_ref
@pytest.mark.parametrize("L,nr,nv,ns", [(5, 4, 2, 2), (31, 7, 1, 0), (12, 3, 0, 2), (3, 2, 3, 1)])
def test_centre_row_bitexact_vs_reference(L, nr, nv, ns):
    ncomp = 3 * nv + ns
    c, orc = _plan(L, nr, 2 * L + 2)
    R = ref.RefPlan(L, nr, c.nth, c.radius())
    sp = gaussian(11 + L, c.n_rj(ncomp))
    init = gaussian(12 + L, ncomp * (2 * nr + 1))
    got = orc.pole_backward(nv, ns, sp, center=True, v_pole=init).reshape(ncomp, -1)[:, 2 * nr]
    want = R.center(nv, ns, sp, init).reshape(ncomp, -1)[:, 2 * nr]
    np.testing.assert_array_equal(got, want)
