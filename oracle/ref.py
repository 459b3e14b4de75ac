"""ctypes access to oracle/_ref/libkemoref.so: the reference's own numerical
routines compiled from their sources (oracle/Makefile.ref) behind this
project's driver (oracle/ref_harness.f90).  TEST INFRASTRUCTURE ONLY; tests
that use it skip when the library has not been built (it is built only where
the reference tree is present)."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(_HERE, "_ref", "libkemoref.so")
REFERENCE_TREE = "/root/reference/MHD"
_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)


def build() -> bool:
    """Build _ref when the reference tree is present; returns availability."""
    if os.path.isdir(REFERENCE_TREE):
        subprocess.run(["make", "-s", "-f", "Makefile.ref"], cwd=_HERE, check=True)
    return os.path.exists(SO)


def available() -> bool:
    return os.path.exists(SO)


def lib():
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(SO)
        L.ref_gauss_points.argtypes = [ctypes.c_int, _dp, _dp]
        L.ref_schmidt_legendres_m.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.ref_sph_normalizations.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.ref_rfftm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.ref_setup.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.ref_get_table.argtypes = [ctypes.c_int, _dp]
        L.ref_leg_bwd.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.ref_leg_fwd.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.ref_center.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        _ip = ctypes.POINTER(ctypes.c_int)
        L.ref_to_send_buf.argtypes = [ctypes.c_int] * 4 + [_ip] + [ctypes.c_int] * 3 + [_dp, _dp]
        L.ref_from_recv_rev.argtypes = [ctypes.c_int] * 4 + [_ip] + [ctypes.c_int] * 3 + [_dp, _dp]
        L.ref_coriolis_tables.argtypes = [_ip, _ip] + [_dp] * 8
        L.ref_sum_coriolis.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.ref_write_sph_files.argtypes = [ctypes.c_int] * 4 + [_dp, _ip, ctypes.c_char_p]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


def gauss_points(n):
    x = np.empty(n)
    w = np.empty(n)
    lib().ref_gauss_points(n, _p(x), _p(w))
    return x, w


def schmidt_legendres_m(ltr, m, theta):
    p = np.zeros(ltr + 1)
    dp = np.zeros(ltr + 1)
    lib().ref_schmidt_legendres_m(ltr, m, theta, _p(p), _p(dp))
    return p, dp


def sph_normalizations(l, m):
    g = np.empty(17)
    lib().ref_sph_normalizations(l, m, _p(g))
    return g


def rfft_fwd(x):
    """Lines contiguous in, FFTPACK halfcomplex lines out."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[-1]
    lines = x.reshape(-1, n)
    buf = np.ascontiguousarray(lines.T)  # (n, lot): line index fastest
    lib().ref_rfftm(1, n, lines.shape[0], _p(buf))
    return np.ascontiguousarray(buf.T).reshape(x.shape)


def rfft_bwd(c):
    c = np.asarray(c, dtype=np.float64)
    n = c.shape[-1]
    lines = c.reshape(-1, n)
    buf = np.ascontiguousarray(lines.T)
    lib().ref_rfftm(-1, n, lines.shape[0], _p(buf))
    return np.ascontiguousarray(buf.T).reshape(c.shape)


def to_send(fields, ntot_phys, nnod_org, d_org, nb, WS, inod_export):
    """The reference's set_to_send_buf_* for each (i_field, i_trans, ncomp)
    (0-based) in turn; d_org field-major (ntot_phys x nnod_org)."""
    d = np.ascontiguousarray(d_org, dtype=np.float64).ravel()
    ws = np.array(WS, dtype=np.float64).ravel()
    inod = np.ascontiguousarray(inod_export, dtype=np.int32)
    nsend = ws.size // nb
    for i_fld, i_tr, nc in fields:
        lib().ref_to_send_buf(nc, nb, nnod_org, nsend, inod.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              ntot_phys, i_fld + 1, i_tr + 1, _p(d), _p(ws))
    return ws


def from_recv_rev(fields, nb, nnod_recv, WR, ntot_phys, nnod_new, d_new, irev):
    """The reference's set_from_recv_buf_rev_* for each field in turn; WR is
    extended by the dummy record the routines zero."""
    wr = np.zeros(nb * (nnod_recv + 1))
    wr[:nb * nnod_recv] = np.asarray(WR, dtype=np.float64).ravel()[:nb * nnod_recv]
    d = np.array(d_new, dtype=np.float64).ravel()
    rv = np.ascontiguousarray(irev, dtype=np.int32)
    for i_fld, i_tr, nc in fields:
        lib().ref_from_recv_rev(nc, nb, nnod_new, nnod_recv,
                                rv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ntot_phys,
                                i_fld + 1, i_tr + 1, _p(wr), _p(d))
    return d


class RefPlan:
    """Module-level state in the harness: one live plan at a time."""

    def __init__(self, L, nr, nth, radius):
        radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.L, self.nr, self.nth = L, nr, nth
        self.jmax = (L + 1) ** 2
        self.nslot = 2 * L + 1
        self.nth_h = (nth + 1) // 2
        lib().ref_setup(L, nr, nth, _p(radius))

    def table(self, name):
        which = {"gpoint": 1, "weight": 2, "P": 3, "dP": 4, "Ps": 5, "dPs": 6, "g": 7}[name]
        size = {1: self.nth, 2: self.nth, 3: self.nth * self.jmax, 4: self.nth * self.jmax,
                5: self.nth_h * self.jmax, 6: self.nth_h * self.jmax, 7: 17 * self.jmax}[which]
        out = np.empty(size)
        lib().ref_get_table(which, _p(out))
        if which in (3, 4):
            out = out.reshape(self.jmax, self.nth)
        elif which in (5, 6):
            out = out.reshape(self.jmax, self.nth_h)
        elif which == 7:
            out = out.reshape(17, self.jmax)
        return out

    def leg_bwd(self, nv, ns, sp):
        ncomp = 3 * nv + ns
        WR = np.ascontiguousarray(sp, dtype=np.float64).ravel().copy()
        WS = np.zeros(ncomp * self.nth * self.nr * self.nslot)
        lib().ref_leg_bwd(nv, ns, _p(WR), _p(WS))
        return WS

    def leg_fwd(self, nv, ns, vr):
        ncomp = 3 * nv + ns
        WR = np.ascontiguousarray(vr, dtype=np.float64).ravel().copy()
        WS = np.zeros(ncomp * self.jmax * self.nr)
        lib().ref_leg_fwd(nv, ns, _p(WR), _p(WS))
        return WS

    def coriolis_tables(self):
        """cal_gaunt_coriolis_rlm + interact_rot_coriolis_rlm on this plan's
        rlm order, in the layout of oracle.OraclePlan.coriolis_tables."""
        jm = self.jmax
        ip = ctypes.POINTER(ctypes.c_int)
        jgi = np.zeros(2 * jm, dtype=np.int32)
        jei = np.zeros(jm, dtype=np.int32)
        gi, ei = np.zeros(2 * jm), np.zeros(jm)
        sw, tw, sd = np.zeros(6 * jm), np.zeros(8 * jm), np.zeros(4 * jm)
        td, sr, tr = np.zeros(2 * jm), np.zeros(2 * jm), np.zeros(2 * jm)
        lib().ref_coriolis_tables(jgi.ctypes.data_as(ip), jei.ctypes.data_as(ip), _p(gi), _p(ei),
                                  _p(sw), _p(tw), _p(sd), _p(td), _p(sr), _p(tr))
        return dict(jgi=jgi.reshape(2, jm), jei=jei, gi=gi.reshape(2, jm), ei=ei,
                    sw=sw.reshape(jm, 3, 2), tw=tw.reshape(jm, 4, 2))

    def sum_coriolis(self, ncomp, i_velo, i_vort, coef_cor, sp, kr_icb=0):
        """sum_rot_coriolis_rlm_10 (+ inner_core_rot_z_coriolis_rlm when
        kr_icb > 0, 1-based); i_velo / i_vort 0-based.  (2, jmax*nr)."""
        WR = np.ascontiguousarray(sp, dtype=np.float64).ravel().copy()
        out = np.zeros(2 * self.jmax * self.nr)
        lib().ref_sum_coriolis(ncomp, i_velo + 1, i_vort + 1, coef_cor, kr_icb, _p(WR), _p(out))
        return out.reshape(2, -1)

    def center(self, nv, ns, sp, v_pole):
        """The reference's centre row (SS/schmidt_b_trans_at_center.f90)
        applied to a copy of v_pole (ncomp x (2 nr + 1), component-major)."""
        WR = np.ascontiguousarray(sp, dtype=np.float64).ravel().copy()
        out = np.array(v_pole, dtype=np.float64).ravel().copy()
        assert out.size == (3 * nv + ns) * (2 * self.nr + 1)
        lib().ref_center(nv, ns, _p(WR), _p(out))
        return out


def write_sph_files(L, nr, nth, nphi, radius, paths, item_rj=None):
    """A one-rank set of grid / mode files (rtp, rtm, rlm, rj at the four
    paths) written by the reference's own writers with its own ordering
    routines (oracle/ref_sph_files.f90); item_rj: the rj import items
    (1-based rj node per rlm node; None = node order)."""
    jmax = (L + 1) ** 2
    if item_rj is None:
        item_rj = np.arange(1, jmax * nr + 1)
    item = np.ascontiguousarray(item_rj, dtype=np.int32)
    assert item.size == jmax * nr
    buf = b"".join(p.encode().ljust(512, b"\0")[:511] + b"\0" for p in paths)
    assert len(buf) == 4 * 512
    r = np.ascontiguousarray(radius, dtype=np.float64)
    lib().ref_write_sph_files(L, nr, nth, nphi, _p(r),
                              item.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), buf)
