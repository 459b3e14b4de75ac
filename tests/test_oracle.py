"""CPU checks of the oracle (the FP64 restatement of the reference transform in
oracle/sht_oracle.c) against the reference's own routines and against what
the mathematics fixes.

Golden vectors: tests/golden/ref_fixtures.npz, written by
tests/golden/make_golden.py from the reference sources compiled in place
(oracle/Makefile.ref).  Where the compiled reference routines are present the
live tests compare on larger cases too; they skip elsewhere.
"""
import math
import os

import mpmath
import numpy as np
import pytest
from scipy.special import lpmv

import oracle as O
from oracle import ref
from kemorin_mhd_amd.sht import ShellConfig, gaussian, rlm_l_m, synthetic_spectra

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures.npz"))
THETAS = np.array([1e-3, 0.3, 1.0, 1.5707963267948966, 2.2, 3.1])


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def schmidt_closed(l, m, theta):
    """Schmidt quasi-normalised P_l^m(cos theta), no Condon-Shortley phase,
    from scipy's associated Legendre function."""
    p = lpmv(m, l, math.cos(theta))
    if m == 0:
        return p
    return (-1) ** m * p * math.sqrt(2.0 * math.factorial(l - m) / math.factorial(l + m))


# ---------------------------------------------------------------- Gauss points
@pytest.mark.parametrize("n", [48, 384, 768])
def test_gauss_points_bitexact_vs_reference(n):
    """GI/gauss_integration.f90:60-133 Newton loop, reproduced bit for bit."""
    x, w = O.gauss_points(n)
    assert np.array_equal(x, GOLD[f"gauss_x_{n}"])
    assert np.array_equal(w, GOLD[f"gauss_w_{n}"])


def test_gauss_exact_vs_mpmath():
    n = 48
    x, w = O.gauss_points(n, exact=True)
    mpmath.mp.dps = 40
    for i in (0, 1, 7, 23, 24, 40, 47):
        r = mpmath.findroot(lambda t: mpmath.legendre(n, t), mpmath.mpf(x[i]))
        dp = mpmath.diff(lambda t: mpmath.legendre(n, t), r)
        wr = 2 / ((1 - r ** 2) * dp ** 2)
        assert abs(x[i] - float(r)) <= 2e-16
        assert abs(w[i] - float(wr)) <= 2e-16 * float(wr) * 4


@pytest.mark.parametrize("n", [48, 384])
def test_gauss_polynomial_exactness(n):
    x, w = O.gauss_points(n, exact=True)
    assert np.all(np.diff(x) > 0) and x[0] > -1 and x[-1] < 1  # ascending: theta_1 ~ pi
    for k in (0, 1, 2, 10, 2 * n - 2, 2 * n - 1):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-13


def test_reference_weights_deviation_is_bounded():
    """The reference's weights use P' at the pre-update Newton iterate, so they
    differ from exact ones; the difference stays small but non-zero."""
    xr, wr = O.gauss_points(768)
    xe, we = O.gauss_points(768, exact=True)
    d = np.max(np.abs(wr - we) / we)
    assert 1e-12 < d < 1e-5
    assert np.max(np.abs(xr - xe)) < 1e-13


# ------------------------------------------------------------- Schmidt P_l^m
@pytest.mark.parametrize("m", [0, 1, 5, 31])
def test_schmidt_vs_reference(m):
    for i, th in enumerate(THETAS):
        p, dp = O.schmidt_legendres_m(31, m, th)
        g, dg = GOLD[f"plm_{m}"][i], GOLD[f"dplm_{m}"][i]
        scale = max(np.max(np.abs(g)), 1e-300)
        assert np.max(np.abs(p - g)) <= 4e-16 * scale * 31
        dscale = max(np.max(np.abs(dg)), 1e-300)
        assert np.max(np.abs(dp - dg)) <= 4e-16 * dscale * 31


@pytest.mark.parametrize("m", [0, 1, 2, 7, 20])
def test_schmidt_vs_closed_form(m):
    L = 40
    for th in (0.05, 0.7, 1.3, 2.0, 3.0):
        p, _ = O.schmidt_legendres_m(L, m, th)
        for l in range(m, L + 1):
            c = schmidt_closed(l, m, th)
            assert abs(p[l] - c) <= 1e-12 * max(1.0, abs(c)), (l, m, th)
    th = 0.9
    s, c = math.sin(th), math.cos(th)
    p0, _ = O.schmidt_legendres_m(2, 0, th)
    p1, _ = O.schmidt_legendres_m(2, 1, th)
    p2, _ = O.schmidt_legendres_m(2, 2, th)
    assert p0[0] == pytest.approx(1.0) and p0[1] == pytest.approx(c)
    assert p0[2] == pytest.approx(1.5 * c * c - 0.5)
    assert p1[1] == pytest.approx(s) and p1[2] == pytest.approx(math.sqrt(3) * s * c)
    assert p2[2] == pytest.approx(math.sqrt(3) / 2 * s * s)


def test_schmidt_sectoral_closed_form():
    """P_l^l = sqrt(2 / (2l)!) (2l-1)!! sin^l."""
    th = 1.1
    for l in (1, 5, 17, 30):
        p, _ = O.schmidt_legendres_m(l, l, th)
        dfact = math.prod(range(2 * l - 1, 0, -2))
        c = math.sqrt(2.0 / math.factorial(2 * l)) * dfact * math.sin(th) ** l
        assert p[l] == pytest.approx(c, rel=1e-13)


@pytest.mark.parametrize("m", [0, 1, 3, 12])
def test_schmidt_derivative_vs_finite_difference(m):
    L, h = 30, 1e-6
    for th in (0.2, 1.0, 2.5):
        _, dp = O.schmidt_legendres_m(L, m, th)
        pp, _ = O.schmidt_legendres_m(L, m, th + h)
        pm, _ = O.schmidt_legendres_m(L, m, th - h)
        fd = (pp - pm) / (2 * h)
        assert np.max(np.abs(dp[m:] - fd[m:])) <= 1e-7 * max(1.0, np.max(np.abs(fd)))


@pytest.mark.parametrize("m", [0, 1, 9, 31])
def test_schmidt_orthogonality_with_exact_quadrature(m):
    n, L = 48, 31
    x, w = O.gauss_points(n, exact=True)
    P = np.array([O.schmidt_legendres_m(L, m, math.acos(xi))[0] for xi in x])  # (n, L+1)
    G = (P[:, m:] * w[:, None]).T @ P[:, m:]
    l = np.arange(m, L + 1)
    want = np.diag((2.0 if m > 0 else 1.0) * 2.0 / (2 * l + 1))
    assert np.max(np.abs(G - want)) < 1e-13


def test_normalizations():
    for l in range(0, 32, 5):
        for m in (0, 1, l):
            if m > l:
                continue
            g = O.sph_normalizations(l, m)
            assert np.array_equal(g, GOLD[f"g_{l}_{m}"])
            assert g[2] == l * (l + 1)
            g6 = (2 * l + 1) / (2.0 if m == 0 else 4.0)
            assert g[5] == pytest.approx(0.5 if l == 0 and m == 0 else g6)
            if l > 0:
                assert g[6] == pytest.approx(g6 / (l * (l + 1)))


def test_chebyshev_radius():
    r = O.chebyshev_radius(33, 0.35 / 0.65, 1 / 0.65)
    k = np.arange(33)
    ri, ro = 0.35 / 0.65, 1 / 0.65
    assert np.allclose(r, ri + (ro - ri) / 2 * (1 - np.cos(np.pi * k / 32)), rtol=0, atol=1e-15)
    assert r[0] == pytest.approx(ri) and r[-1] == pytest.approx(ro)


# ------------------------------------------------------------------------- FFT
def test_fft_known_answers():
    """The six signals of the reference FFT test program: each has one or two
    non-zero FFTPACK coefficients (c[0]=a0, c[2k-1]=a_k, c[2k]=b_k)."""
    x = GOLD["fft_in"]
    c = O.rfft_fwd(x)
    want = np.zeros_like(x)
    want[0, 0] = 10.0
    want[1, 2] = 2.0
    want[2, 1] = 3.0
    want[3, 6] = -4.0
    want[4, 7] = -5.0
    want[5, 20], want[5, 15] = 6.0, 1.5
    assert np.max(np.abs(c - want)) < 1e-13
    assert np.max(np.abs(c - GOLD["fft_fwd"])) < 1e-14
    assert np.max(np.abs(O.rfft_bwd(c) - GOLD["fft_bwd"])) < 1e-13
    assert np.max(np.abs(O.rfft_bwd(c) - x)) < 1e-13


@pytest.mark.parametrize("n", [16, 96, 250, 1536])
def test_fft_vs_direct_dft(n):
    x = gaussian(5000 + n, 3 * n).reshape(3, n)
    c = O.rfft_fwd(x)
    j = np.arange(n)
    k = np.arange(n // 2 + 1)
    E = np.exp(-2j * np.pi * np.outer(j, k) / n)
    X = x @ E
    want = np.empty_like(x)
    want[:, 0] = X[:, 0].real / n
    want[:, 1:n - 1:2] = 2 * X[:, 1:n // 2].real / n
    want[:, 2:n - 1:2] = -2 * X[:, 1:n // 2].imag / n
    want[:, n - 1] = X[:, n // 2].real / n
    assert np.max(np.abs(c - want)) < 1e-13
    assert np.max(np.abs(c - GOLD[f"fftr_fwd_{n}"])) < 1e-14
    assert np.max(np.abs(O.rfft_bwd(c) - x)) < 1e-13


# -------------------------------------------------------------------- Legendre
def leg_case():
    full = ShellConfig(31, 33, 48, 96)
    r = full.radius()[[3, 17]]
    cfg = ShellConfig(31, 2, 48, 96)
    return cfg, r


def test_legendre_vs_reference_fixture():
    cfg, r = leg_case()
    P = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, r)
    sp = synthetic_spectra(cfg, 2, 2, seed=101)
    assert rel(P.leg_bwd(2, 2, sp), GOLD["leg_b_out"]) < 1e-14
    vr = gaussian(103, 9 * cfg.nth * cfg.nr * (2 * cfg.L + 1))
    assert rel(P.leg_fwd(3, 0, vr), GOLD["leg_f_out"]) < 1e-14


@pytest.mark.parametrize("nv,ns", [(2, 2), (1, 0), (0, 3)])
def test_legendre_symmetric_vs_plain_loops(nv, ns):
    """symmetric_BLAS_big restatement vs the plain-loop definition
    (SS/legendre_bwd_trans_org.f90, SS/legendre_fwd_trans_org.f90)."""
    for nth in (48, 47):
        cfg = ShellConfig(31, 3, nth, 96)
        P = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius())
        sp = synthetic_spectra(cfg, nv, ns, seed=3)
        assert rel(P.leg_bwd(nv, ns, sp), P.leg_bwd(nv, ns, sp, variant="org")) < 1e-13
        vr = gaussian(9, (3 * nv + ns) * cfg.nth * cfg.nr * (2 * cfg.L + 1))
        assert rel(P.leg_fwd(nv, ns, vr), P.leg_fwd(nv, ns, vr, variant="org")) < 1e-13


def single_mode(cfg, nv, ns, nd, l, m, k=None):
    ncomp = 3 * nv + ns
    ls, ms = rlm_l_m(cfg.L)
    j = int(np.nonzero((ls == l) & (ms == m))[0][0])
    sp = np.zeros((cfg.nr, cfg.jmax, ncomp))
    sp[:, j, nd] = 1.0 if k is None else 0.0
    if k is not None:
        sp[k, j, nd] = 1.0
    return sp.ravel()


@pytest.mark.parametrize("l,m", [(0, 0), (3, 0), (5, 2), (5, -2), (12, 7), (12, -12)])
def test_backward_single_scalar_mode(l, m):
    """Scalar spectrum delta_(l,m): grid = P_l^|m| cos(m phi) for m >= 0 and
    P_l^|m| sin(|m| phi) for m < 0 (SH/spherical_harmonics.f90:190-212)."""
    cfg = ShellConfig(15, 2, 24, 40)
    P = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    v = P.backward(0, 1, single_mode(cfg, 0, 1, 0, l, m)).reshape(cfg.nphi, cfg.nth, cfg.nr)
    x, _ = O.gauss_points(cfg.nth)
    th = np.arccos(x)
    phi = 2 * np.pi * np.arange(cfg.nphi) / cfg.nphi
    pl = np.array([schmidt_closed(l, abs(m), t) for t in th])
    ang = np.cos(m * phi) if m >= 0 else np.sin(-m * phi)
    want = ang[:, None] * pl[None, :]
    assert np.max(np.abs(v - want[:, :, None])) < 1e-13


@pytest.mark.parametrize("l,m", [(1, 0), (4, 3), (4, -3), (9, -1)])
def test_backward_single_vector_mode(l, m):
    """Poloidal S = pol P Y and toroidal T = tor P Y with radial field
    vr = l(l+1) pol Y / r^2, vtheta = (dpol dY/dtheta + dY/dphi tor / sin) / r,
    vphi = (dpol dY/dphi / sin - tor dY/dtheta) / r (SS/legendre_bwd_trans_org.f90)."""
    cfg = ShellConfig(15, 3, 24, 40)
    r = cfg.radius()
    P = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, r, full_tables=False)
    x, _ = O.gauss_points(cfg.nth)
    th = np.arccos(x)
    phi = 2 * np.pi * np.arange(cfg.nphi) / cfg.nphi
    M = abs(m)
    p = np.array([O.schmidt_legendres_m(l, M, t)[0][l] for t in th])
    dp = np.array([O.schmidt_legendres_m(l, M, t)[1][l] for t in th])
    ang = np.cos(M * phi) if m >= 0 else np.sin(M * phi)
    dang = -M * np.sin(M * phi) if m >= 0 else M * np.cos(M * phi)
    Y, dYt, dYp = ang[:, None] * p, ang[:, None] * dp, dang[:, None] * p
    s = np.sin(th)[None, :]
    for nd in range(3):
        v = P.backward(1, 0, single_mode(cfg, 1, 0, nd, l, m)).reshape(3, cfg.nphi, cfg.nth, cfg.nr)
        for k in range(cfg.nr):
            a = 1.0 / r[k]
            if nd == 0:
                want = (l * (l + 1) * a * a * Y, 0 * Y, 0 * Y)
            elif nd == 1:
                want = (0 * Y, a * dYt, a * dYp / s)
            else:
                want = (0 * Y, a * dYp / s, -a * dYt)
            for c in range(3):
                scale = max(1.0, np.max(np.abs(want[c])))
                assert np.max(np.abs(v[c, :, :, k] - want[c])) < 1e-13 * scale, (nd, c, k)


@pytest.mark.parametrize("nth", [48, 49])
def test_round_trip_identity_with_exact_weights(nth):
    """forward(backward(x)) = x for band-limited spectra when the Gauss weights
    are exact (vector dpol / tor have no l = 0 part)."""
    cfg = ShellConfig(31, 3, nth, 96)
    P = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), exact_weights=True,
                     full_tables=False)
    sp = synthetic_spectra(cfg, 2, 2, seed=11).reshape(cfg.nr, cfg.jmax, 8)
    l, _ = rlm_l_m(cfg.L)
    sp[:, l == 0, 1] = sp[:, l == 0, 2] = sp[:, l == 0, 4] = sp[:, l == 0, 5] = 0.0
    back = P.forward(2, 2, P.backward(2, 2, sp.ravel()))
    assert rel(back, sp.ravel()) < 1e-13


def test_round_trip_reference_weights_quadrature_error():
    """With the reference's inexact weights the round trip is only good to the
    weight error (about 1e-7 at Ntheta = 48)."""
    cfg = ShellConfig(31, 2, 48, 96)
    P = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    sp = synthetic_spectra(cfg, 0, 1, seed=13)
    e = rel(P.forward(0, 1, P.backward(0, 1, sp)), sp)
    assert e < 1e-5


def test_exchange_permutations_are_inverse():
    cfg = ShellConfig(7, 3, 12, 16)
    P = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    rtm = gaussian(17, P.n_rtm(5))
    rtp = P.rtm_to_rtp(5, rtm)
    assert np.array_equal(P.rtp_to_rtm(5, rtp), rtm)
    # slots of orders above L are the zero dummy slot
    q = rtp.reshape(cfg.nphi, -1)
    assert np.all(q[2 * cfg.L + 2:] == 0.0)


# ------------------------------------------------- live compiled reference
needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


@needs_ref
def test_live_reference_legendre_cfg0():
    cfg = ShellConfig(31, 33, 48, 96)
    r = cfg.radius()
    R = ref.RefPlan(cfg.L, cfg.nr, cfg.nth, r)
    P = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, r, full_tables=False)
    sp = synthetic_spectra(cfg, 2, 2, seed=5)
    assert rel(P.leg_bwd(2, 2, sp), R.leg_bwd(2, 2, sp)) < 1e-14
    vr = gaussian(6, P.n_rtm(9))
    assert rel(P.leg_fwd(3, 0, vr), R.leg_fwd(3, 0, vr)) < 1e-14
    assert np.array_equal(P.table("weight"), R.table("weight"))


@needs_ref
def test_live_reference_odd_grid():
    cfg = ShellConfig(15, 4, 25, 40)
    r = cfg.radius()
    R = ref.RefPlan(cfg.L, cfg.nr, cfg.nth, r)
    P = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, r, full_tables=False)
    sp = synthetic_spectra(cfg, 1, 1, seed=8)
    assert rel(P.leg_bwd(1, 1, sp), R.leg_bwd(1, 1, sp)) < 1e-14
    vr = gaussian(9, P.n_rtm(4))
    assert rel(P.leg_fwd(1, 1, vr), R.leg_fwd(1, 1, vr)) < 1e-14
