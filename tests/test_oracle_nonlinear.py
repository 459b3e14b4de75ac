"""The oracle's grid-space products of the MHD step, pinned.

The reference holds no vectors for nonlinear_terms_in_rtp
(MS/cal_nonlinear_sph_MHD.f90:64-128).  The oracle works in (r, theta, phi)
components; the check goes the other way round: every vector is turned into
Cartesian components with the local frame (e_r, e_theta, e_phi) at the
grid point's colatitude and longitude, the products are taken there with
numpy's cross product (and the rotation axis z for the Coriolis term,
-c z x u), and the result is turned back.  A swapped component, a wrong sign,
a wrong operand order or a wrong colatitude row is an O(1) miss.  The
fluxes (vector times scalar) are checked elementwise, and the field
addressing by a shuffled component layout with untouched outputs.
"""
import numpy as np

import oracle as O
from kemorin_mhd_amd import ShellConfig, gaussian


def _frame(orc, cfg):
    th = orc.table("colat")
    ph = 2.0 * np.pi * np.arange(cfg.nphi) / cfg.nphi
    # grid order: k fastest, then theta, then phi -> arrays of shape (nphi, nth, 1)
    T = th[None, :, None]
    P = ph[:, None, None]
    st, ct, sp, cp = np.sin(T), np.cos(T), np.sin(P), np.cos(P)
    z = np.zeros_like(st * cp)
    er = np.stack([st * cp, st * sp, ct + z])
    et = np.stack([ct * cp, ct * sp, -st + z])
    ef = np.stack([-sp + z * 0, cp + z * 0, z])
    return er, et, ef


def _to_cart(v, fr):
    er, et, ef = fr
    return v[0] * er + v[1] * et + v[2] * ef


def _to_sph(c, fr):
    er, et, ef = fr
    return np.stack([(c * er).sum(0), (c * et).sum(0), (c * ef).sum(0)])


def test_products_against_cartesian():
    cfg = ShellConfig(6, 3, 10, 14)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    shape = (cfg.nphi, cfg.nth, cfg.nr)
    # shuffled input layout: light, magne, temp, velo, current, vort (14 comps)
    idx = {"light": 0, "magne": 1, "temp": 4, "velo": 5, "current": 8, "vort": 11,
           "c_flux": 0, "coriolis": 3, "m_advect": 7, "lorentz": 10, "h_flux": 13,
           "vp_induct": 16}
    coef = {"m_advect": -1.5, "lorentz": 0.7, "vp_induct": 2.25, "h_flux": -0.3,
            "c_flux": 1.1, "coriolis": 3.5}
    nf, no = 14, 21  # outputs occupy 0..18; 19, 20 untouched
    fld = gaussian(11, cfg.n_rtp(nf))
    init = np.full(cfg.n_rtp(no), 9.0)
    out = orc.nonlinear_rtp(idx, coef, nf, fld, no, init).reshape(no, *shape)
    f = fld.reshape(nf, *shape)
    fr = _frame(orc, cfg)

    def vec(name):
        return f[idx[name]:idx[name] + 3]

    cart = {n: _to_cart(vec(n), fr) for n in ("velo", "vort", "magne", "current")}
    want = {
        "m_advect": coef["m_advect"] * np.cross(cart["vort"], cart["velo"], axis=0),
        "lorentz": coef["lorentz"] * np.cross(cart["current"], cart["magne"], axis=0),
        "vp_induct": coef["vp_induct"] * np.cross(cart["velo"], cart["magne"], axis=0),
    }
    zhat = np.zeros_like(cart["velo"])
    zhat[2] = 1.0
    want["coriolis"] = -coef["coriolis"] * np.cross(zhat, cart["velo"], axis=0)
    scale = max(np.abs(v).max() for v in want.values())
    for n, w in want.items():
        got = out[idx[n]:idx[n] + 3]
        err = np.abs(got - _to_sph(w, fr)).max()
        assert err <= 1e-13 * scale, (n, err)
    for n, s in (("h_flux", "temp"), ("c_flux", "light")):
        np.testing.assert_allclose(out[idx[n]:idx[n] + 3], coef[n] * vec("velo") * f[idx[s]],
                                   rtol=1e-15, atol=0)
    assert (out[19:] == 9.0).all()


def test_terms_off_leave_outputs_alone():
    cfg = ShellConfig(3, 2, 4, 8)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    fld = gaussian(3, cfg.n_rtp(3))
    init = np.full(cfg.n_rtp(6), 5.0)
    out = orc.nonlinear_rtp({"velo": 0, "coriolis": 3}, {"coriolis": 2.0}, 3, fld, 6,
                            init).reshape(6, -1)
    assert (out[:3] == 5.0).all()
    assert not (out[3:] == 5.0).all()
