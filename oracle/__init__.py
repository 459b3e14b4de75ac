"""CPU FP64 oracle of the reference spherical-harmonic transform.

TEST INFRASTRUCTURE ONLY.  This package is imported by ``tests/``, by the
oracle leg of ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, always as the checker (or the timed CPU baseline), never as the
thing measured or shipped.  The product package ``kemorin_mhd_amd`` does not
import it.

It wraps ``oracle/liboracle.so`` (built from ``sht_oracle.c`` by
``oracle/Makefile``), a C restatement of the reference routines cited in
``sht_oracle.h``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile; returns its path."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return so


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        L.orc_plan_create.restype = ctypes.c_void_p
        L.orc_plan_create.argtypes = [ctypes.c_int] * 4 + [_dp, ctypes.c_int]
        L.orc_plan_destroy.argtypes = [ctypes.c_void_p]
        L.orc_plan_table.restype = ctypes.c_void_p
        L.orc_plan_table.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.orc_plan_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.orc_gauss_points.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_schmidt_legendres_m.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.orc_sph_normalizations.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_chebyshev_radius.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        for nm in ("orc_rtm_to_rtp", "orc_rtp_to_rtm", "orc_fft_bwd_from_rtp", "orc_fft_fwd_to_rtp"):
            getattr(L, nm).argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        for nm in ("orc_leg_bwd_org", "orc_leg_fwd_org", "orc_leg_bwd_sym", "orc_leg_fwd_sym"):
            getattr(L, nm).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
        for nm in ("orc_sht_backward", "orc_sht_forward"):
            getattr(L, nm).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]
        for nm in ("orc_rfft_fwd", "orc_rfft_bwd"):
            getattr(L, nm).argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_pole_backward.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                        ctypes.c_int]
        L.orc_nonlinear_rtp.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _dp,
                                        ctypes.c_int, _dp, ctypes.c_int, _dp]
        _ip = ctypes.POINTER(ctypes.c_int)
        L.orc_rj_to_send.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_longlong, _dp,
                                     _ip, ctypes.c_longlong, ctypes.c_int, _dp]
        L.orc_rj_from_recv.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_longlong, _dp,
                                       _ip, ctypes.c_int, ctypes.c_longlong, _dp]
        L.orc_coriolis_tables.argtypes = [ctypes.c_void_p, _ip, _ip, _dp, _dp, _dp, _dp]
        L.orc_sum_coriolis_rlm.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, _dp, _ip, ctypes.c_int, _dp]
        L.orc_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _fields_arr(fields):
    a = np.ascontiguousarray(np.asarray(fields, dtype=np.int32).reshape(-1, 3))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _table(t):
    if t is None:
        return None, None
    a = np.ascontiguousarray(np.asarray(t, dtype=np.int32))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def rj_to_send(fields, ntot_phys, nnod_rj, fld, nb, WS, inod_export=None, nnod_send=None):
    """fields: (i_field, i_trans, ncomp) triples, 0-based; WS is updated in
    place (components no field names are kept) and returned."""
    fa, fp = _fields_arr(fields)
    ta, tp = _table(inod_export)
    fld = np.ascontiguousarray(fld, dtype=np.float64)
    WS = np.array(WS, dtype=np.float64)
    n_send = nnod_rj if nnod_send is None else nnod_send
    lib().orc_rj_to_send(len(fa), fp, ntot_phys, nnod_rj, _ptr(fld), tp, n_send, nb, _ptr(WS))
    return WS


def rj_from_recv(fields, nb, nnod_recv, WR, ntot_phys, nnod_rj, fld, irev=None):
    """fld (field-major, ntot_phys x nnod_rj) is updated in place for the
    named components and returned."""
    fa, fp = _fields_arr(fields)
    ta, tp = _table(irev)
    WR = np.ascontiguousarray(WR, dtype=np.float64)
    fld = np.array(fld, dtype=np.float64)
    lib().orc_rj_from_recv(len(fa), fp, nb, nnod_recv, _ptr(WR), tp, ntot_phys, nnod_rj,
                           _ptr(fld))
    return fld


def gauss_points(n: int, exact: bool = False):
    x = np.empty(n)
    w = np.empty(n)
    lib().orc_gauss_points(n, int(exact), _ptr(x), _ptr(w))
    return x, w


def schmidt_legendres_m(ltr: int, m: int, theta: float):
    p = np.zeros(ltr + 3)
    dp = np.zeros(ltr + 3)
    lib().orc_schmidt_legendres_m(ltr, m, theta, _ptr(p), _ptr(dp))
    return p[: ltr + 1], dp[: ltr + 1]


def sph_normalizations(l: int, m: int) -> np.ndarray:
    g = np.empty(17)
    lib().orc_sph_normalizations(l, m, _ptr(g))
    return g


def chebyshev_radius(nr: int, ri: float, ro: float) -> np.ndarray:
    r = np.empty(nr)
    lib().orc_chebyshev_radius(nr, ri, ro, _ptr(r))
    return r


def rfft_fwd(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[-1]
    c = np.empty_like(x)
    lib().orc_rfft_fwd(n, x.size // n, _ptr(x), _ptr(c))
    return c


def rfft_bwd(c: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(c, dtype=np.float64)
    n = c.shape[-1]
    x = np.empty_like(c)
    lib().orc_rfft_bwd(n, c.size // n, _ptr(c), _ptr(x))
    return x


class OraclePlan:
    """Single-rank plan with the reference's index orderings and tables."""

    def __init__(self, L, nr, nth, nphi, radius, exact_weights=False, full_tables=True):
        radius = np.ascontiguousarray(radius, dtype=np.float64)
        assert radius.shape == (nr,)
        flags = (1 if exact_weights else 0) | (2 if full_tables else 0)
        self._h = lib().orc_plan_create(L, nr, nth, nphi, _ptr(radius), flags)
        if not self._h:
            raise ValueError("bad oracle plan dimensions")
        d = (ctypes.c_int * 8)()
        lib().orc_plan_dims(self._h, d)
        self.L, self.nr, self.nth, self.nphi, self.nslot, self.jmax, self.nth_h = list(d)[:7]
        self.full_tables = full_tables

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _LIB is not None:
            _LIB.orc_plan_destroy(h)
            self._h = None

    def table(self, name):
        sizes = {
            "radius": self.nr, "a_r": self.nr, "gpoint": self.nth, "colat": self.nth,
            "weight": self.nth, "asin_t": self.nth, "g3": self.jmax, "g6": self.jmax,
            "g7": self.jmax, "P": self.nth * self.jmax, "dP": self.nth * self.jmax,
            "Ps": self.nth_h * self.jmax, "dPs": self.nth_h * self.jmax,
        }
        isizes = {"lstack": self.nslot + 1, "lstack_even": self.nslot + 1,
                  "idx_j": 3 * self.jmax, "mdx_p": self.jmax, "mdx_n": self.jmax}
        ptr = lib().orc_plan_table(self._h, name.encode())
        if not ptr:
            raise KeyError(name)
        if name in sizes:
            a = np.ctypeslib.as_array(ctypes.cast(ptr, _dp), (sizes[name],)).copy()
            if name in ("P", "dP"):
                a = a.reshape(self.jmax, self.nth)
            if name in ("Ps", "dPs"):
                a = a.reshape(self.jmax, self.nth_h)
            return a
        a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_int)), (isizes[name],)).copy()
        if name == "idx_j":
            a = a.reshape(3, self.jmax)
        return a

    # sizes
    def n_rj(self, ncomp):
        return ncomp * self.jmax * self.nr

    def n_rtm(self, ncomp):
        return ncomp * self.nth * self.nr * self.nslot

    def n_rtp(self, ncomp):
        return ncomp * self.nr * self.nth * self.nphi

    def leg_bwd(self, nv, ns, sp, variant="sym"):
        ncomp = 3 * nv + ns
        sp = np.ascontiguousarray(sp, dtype=np.float64).ravel()
        assert sp.size == self.n_rj(ncomp)
        out = np.empty(self.n_rtm(ncomp))
        fn = lib().orc_leg_bwd_org if variant == "org" else lib().orc_leg_bwd_sym
        if variant == "org" and not self.full_tables:
            raise ValueError("org loops need full tables")
        fn(self._h, nv, ns, _ptr(sp), _ptr(out))
        return out

    def leg_fwd(self, nv, ns, vr, variant="sym"):
        ncomp = 3 * nv + ns
        vr = np.ascontiguousarray(vr, dtype=np.float64).ravel()
        assert vr.size == self.n_rtm(ncomp)
        out = np.empty(self.n_rj(ncomp))
        fn = lib().orc_leg_fwd_org if variant == "org" else lib().orc_leg_fwd_sym
        if variant == "org" and not self.full_tables:
            raise ValueError("org loops need full tables")
        fn(self._h, nv, ns, _ptr(vr), _ptr(out))
        return out

    def rtm_to_rtp(self, ncomp, rtm):
        rtm = np.ascontiguousarray(rtm, dtype=np.float64).ravel()
        out = np.empty(self.n_rtp(ncomp))
        lib().orc_rtm_to_rtp(self._h, ncomp, _ptr(rtm), _ptr(out))
        return out

    def rtp_to_rtm(self, ncomp, rtp):
        rtp = np.ascontiguousarray(rtp, dtype=np.float64).ravel()
        out = np.zeros(self.n_rtm(ncomp))
        lib().orc_rtp_to_rtm(self._h, ncomp, _ptr(rtp), _ptr(out))
        return out

    def fft_bwd_from_rtp(self, ncomp, rtp):
        rtp = np.ascontiguousarray(rtp, dtype=np.float64).ravel()
        out = np.empty(self.n_rtp(ncomp))
        lib().orc_fft_bwd_from_rtp(self._h, ncomp, _ptr(rtp), _ptr(out))
        return out

    def fft_fwd_to_rtp(self, ncomp, v):
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        out = np.empty(self.n_rtp(ncomp))
        lib().orc_fft_fwd_to_rtp(self._h, ncomp, _ptr(v), _ptr(out))
        return out

    def backward(self, nv, ns, WS_rj, variant="sym"):
        ncomp = 3 * nv + ns
        WS_rj = np.ascontiguousarray(WS_rj, dtype=np.float64).ravel()
        assert WS_rj.size == self.n_rj(ncomp)
        if variant == "org" and not self.full_tables:
            raise ValueError("org loops need full tables")
        out = np.empty(self.n_rtp(ncomp))
        lib().orc_sht_backward(self._h, nv, ns, _ptr(WS_rj), _ptr(out), 0 if variant == "org" else 1)
        return out

    def n_pole(self, ncomp):
        return ncomp * (2 * self.nr + 1)

    def pole_backward(self, nv, ns, WS_rj, center=False, v_pole=None):
        """Pole (and centre) values; rows the transform leaves alone keep
        the content of ``v_pole`` (zeros if not given)."""
        ncomp = 3 * nv + ns
        WS_rj = np.ascontiguousarray(WS_rj, dtype=np.float64).ravel()
        assert WS_rj.size == self.n_rj(ncomp)
        out = (np.zeros(self.n_pole(ncomp)) if v_pole is None
               else np.array(v_pole, dtype=np.float64).ravel())
        assert out.size == self.n_pole(ncomp)
        lib().orc_pole_backward(self._h, nv, ns, _ptr(WS_rj), _ptr(out), 1 if center else 0)
        return out

    NL_IN = ("velo", "vort", "magne", "current", "temp", "light")
    NL_OUT = ("m_advect", "lorentz", "vp_induct", "h_flux", "c_flux", "coriolis")

    def nonlinear_rtp(self, idx: dict, coef: dict, ncomp_fld, fld, ncomp_frc, frc=None):
        """Grid-space products; idx maps field names (NL_IN + NL_OUT) to
        their first component, coef maps output names to coefficients.
        Components of frc no term writes keep their values."""
        fld = np.ascontiguousarray(fld, dtype=np.float64).ravel()
        assert fld.size == self.n_rtp(ncomp_fld)
        out = np.zeros(self.n_rtp(ncomp_frc)) if frc is None else np.array(frc, dtype=np.float64).ravel()
        ia = (ctypes.c_int * 12)(*[idx.get(k, -1) for k in self.NL_IN + self.NL_OUT])
        ca = np.array([coef.get(k, 0.0) for k in self.NL_OUT], dtype=np.float64)
        lib().orc_nonlinear_rtp(self._h, ia, _ptr(ca), ncomp_fld, _ptr(fld), ncomp_frc, _ptr(out))
        return out

    def coriolis_tables(self):
        """Gaunt / Elsasser integrals and the Coriolis coefficients in the
        reference's shapes: jgi (2, jmax) 1-based, jei (jmax), gi (2, jmax),
        ei (jmax), sw (jmax, 3, 2), tw (jmax, 4, 2) (C order of sw(2,3,j),
        tw(2,4,j))."""
        jm = self.jmax
        jgi = np.zeros(2 * jm, dtype=np.int32)
        jei = np.zeros(jm, dtype=np.int32)
        gi, ei = np.zeros(2 * jm), np.zeros(jm)
        sw, tw = np.zeros(6 * jm), np.zeros(8 * jm)
        ip = ctypes.POINTER(ctypes.c_int)
        lib().orc_coriolis_tables(self._h, jgi.ctypes.data_as(ip), jei.ctypes.data_as(ip),
                                  _ptr(gi), _ptr(ei), _ptr(sw), _ptr(tw))
        return dict(jgi=jgi.reshape(2, jm), jei=jei, gi=gi.reshape(2, jm), ei=ei,
                    sw=sw.reshape(jm, 3, 2), tw=tw.reshape(jm, 4, 2))

    def sum_coriolis_rlm(self, ncomp, i_velo, i_vort, coef_cor, WR, irev=None, kr_icb=0):
        """Curl of the Coriolis term in rlm space: (2, jmax*nr) array of the
        poloidal and toroidal parts (i_velo, i_vort 0-based components)."""
        WR = np.ascontiguousarray(WR, dtype=np.float64).ravel()
        assert WR.size >= self.n_rj(ncomp)
        ta, tp = _table(irev)
        out = np.zeros(2 * self.jmax * self.nr)
        lib().orc_sum_coriolis_rlm(self._h, ncomp, i_velo, i_vort, coef_cor, _ptr(WR), tp,
                                   kr_icb, _ptr(out))
        return out.reshape(2, -1)

    def forward(self, nv, ns, v_rtp, variant="sym"):
        ncomp = 3 * nv + ns
        v_rtp = np.ascontiguousarray(v_rtp, dtype=np.float64).ravel()
        assert v_rtp.size == self.n_rtp(ncomp)
        if variant == "org" and not self.full_tables:
            raise ValueError("org loops need full tables")
        out = np.empty(self.n_rj(ncomp))
        lib().orc_sht_forward(self._h, nv, ns, _ptr(v_rtp), _ptr(out), 0 if variant == "org" else 1)
        return out
