"""Python host mirror of the reference's sph_trans call surface.

The reference Fortran calls, per time step, ``sph_backward_transforms`` and
``sph_forward_transforms`` (SS/t_sph_transforms.f90:74,250) after
``initialize_sph_trans`` (SS/init_sph_trans.f90:46).  :class:`SphTransPlan`
offers the same operations on device-resident torch tensors through the C ABI
of ``libkemorin_sht.so`` (``include/kemorin_sht.h``); tensors are used only to
carry device memory and the stream, all arithmetic runs in the HIP kernels.
There is no CPU fallback: if the library or a GPU is missing the calls raise.

SS/ = MHD/Fortran_src/PARALLEL_src/SPH_SHELL_src/ in the reference tree.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkemorin_sht.so")

SHT_OK = 0
SHT_PREC_FP64 = 0
SHT_PREC_SPLIT_INT8 = 1

_LIB = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


class ShtError(RuntimeError):
    """Non-zero status from the C ABI (message from sht_get_last_error)."""

    def __init__(self, status: int, where: str):
        msg = last_error()
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


class _Desc(ctypes.Structure):
    _fields_ = [("l_truncation", ctypes.c_int), ("nidx_r", ctypes.c_int),
                ("nidx_theta", ctypes.c_int), ("nidx_phi", ctypes.c_int),
                ("radius", _dp), ("max_ncomp", ctypes.c_int), ("precision", ctypes.c_int),
                ("n_slices", ctypes.c_int), ("flags", ctypes.c_int)]


class _NlDesc(ctypes.Structure):
    _fields_ = ([(f"i_{n}", ctypes.c_int) for n in
                 ("velo", "vort", "magne", "current", "temp", "light", "m_advect", "lorentz",
                  "vp_induct", "h_flux", "c_flux", "coriolis")] +
                [(f"coef_{n}", ctypes.c_double) for n in
                 ("velo", "lor", "induct", "h_advect", "c_advect", "cor")])


NL_INPUTS = ("velo", "vort", "magne", "current", "temp", "light")
NL_OUTPUTS = ("m_advect", "lorentz", "vp_induct", "h_flux", "c_flux", "coriolis")
_NL_COEF = dict(zip(NL_OUTPUTS, ("velo", "lor", "induct", "h_advect", "c_advect", "cor")))


class _FieldCopy(ctypes.Structure):
    _fields_ = [("i_field", ctypes.c_int), ("i_trans", ctypes.c_int), ("ncomp", ctypes.c_int)]


class _ShellCall(ctypes.Structure):
    _fields_ = [("plan", _vp), ("nvector_b", ctypes.c_int), ("nscalar_b", ctypes.c_int),
                ("d_WS_rj", _vp), ("d_v_rtp", _vp), ("nvector_f", ctypes.c_int),
                ("nscalar_f", ctypes.c_int), ("d_v_in", _vp), ("d_WR_rj", _vp)]


def lib() -> ctypes.CDLL:
    """Load the in-tree HIP library (fails loudly when it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run python -m kemorin_mhd_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        L.sht_plan_create.argtypes = [ctypes.POINTER(_Desc), ctypes.POINTER(_vp)]
        L.sht_plan_destroy.argtypes = [_vp]
        L.sht_plan_set_stream.argtypes = [_vp, _vp]
        L.sht_plan_get_table.argtypes = [_vp, ctypes.c_char_p, _vp, ctypes.c_longlong]
        L.sht_plan_device_bytes.argtypes = [_vp]
        L.sht_plan_device_bytes.restype = ctypes.c_longlong
        L.sht_plan_set_variant.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int]
        L.sht_plan_get_variant.argtypes = [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        L.sht_plan_autotune.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_char_p,
                                        ctypes.POINTER(ctypes.c_double)]
        for nm in ("sht_backward", "sht_forward", "sht_backward_host", "sht_forward_host",
                   "sht_legendre_backward", "sht_legendre_forward", "sht_backward_noalloc"):
            getattr(L, nm).argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]
        for nm in ("sht_fft_backward", "sht_fft_forward", "sht_exchange_rtm_to_rtp",
                   "sht_exchange_rtp_to_rtm"):
            getattr(L, nm).argtypes = [_vp, ctypes.c_int, _vp, _vp]
        L.sht_pole_backward.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int]
        L.sht_backward_with_poles.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp,
                                              ctypes.c_int]
        L.sht_nonlinear_rtp.argtypes = [_vp, ctypes.POINTER(_NlDesc), ctypes.c_int, _vp,
                                        ctypes.c_int, _vp]
        L.sht_exchange_gather_rev.argtypes = [_vp, ctypes.c_int, ctypes.c_longlong,
                                              ctypes.c_longlong, _vp, _vp, _vp]
        _ll = ctypes.c_longlong
        L.sht_rj_fields_to_send.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_FieldCopy),
                                            ctypes.c_int, _ll, _vp, _vp, _ll, ctypes.c_int, _vp]
        L.sht_rj_fields_from_recv.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_FieldCopy),
                                              ctypes.c_int, _ll, _vp, _vp, ctypes.c_int, _ll, _vp]
        L.sht_coriolis_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp)]
        L.sht_coriolis_destroy.argtypes = [_vp]
        L.sht_sum_coriolis_rlm.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, _vp, _ll, _vp, _vp]
        L.sht_copy_coriolis_terms_rlm.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _ll,
                                                  _vp]
        L.sht_plan_timer_enable.argtypes = [_vp, ctypes.c_int]
        L.sht_plan_timer_read.argtypes = [_vp, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
        L.sht_batch_create.argtypes = [ctypes.POINTER(_ShellCall), ctypes.c_int,
                                       ctypes.POINTER(_vp)]
        L.sht_batch_launch.argtypes = [_vp, _vp]
        L.sht_batch_synchronize.argtypes = [_vp]
        L.sht_batch_destroy.argtypes = [_vp]
        L.sht_sph_files_read.argtypes = [ctypes.c_char_p] * 4 + [ctypes.POINTER(_vp)]
        L.sht_sph_files_destroy.argtypes = [_vp]
        L.sht_sph_files_destroy.restype = None
        L.sht_sph_files_check.argtypes = [_vp]
        L.sht_sph_files_shape.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), _dp]
        L.sht_sph_files_table.argtypes = [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), _ll]
        L.sht_sph_files_table.restype = _ll
        L.sht_plan_create_from_files.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_uint, ctypes.POINTER(_vp)]
        L.sht_get_last_error.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.sht_version.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def last_error() -> str:
    buf = ctypes.create_string_buffer(512)
    lib().sht_get_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def _check(rc: int, where: str):
    if rc != SHT_OK:
        raise ShtError(rc, where)


HIP_STREAM_LEGACY = 1  # hipStreamLegacy: the legacy default (null) stream
SHT_FLAG_P_ON_THE_FLY = 1  # P tables generated per call, one tile of orders at a time
SHT_FLAG_CHECK_FINITE = 2  # whole transforms count their outputs' NaN / Inf values
SHT_ERR_OOM = 3            # status of a call whose device allocation failed
SHT_ERR_NONFINITE = 6      # status of a checked transform whose output is not finite


def _torch_stream(*bufs):
    """Handle of torch's current stream when any buffer is a torch device
    tensor (0 for the default stream), else None."""
    for t in bufs:
        if hasattr(t, "is_cuda") and t.is_cuda:
            import torch
            return torch.cuda.current_stream(t.device).cuda_stream
    return None


def _ptr(t) -> int:
    """Device (or host numpy) address of a contiguous float64 buffer."""
    if isinstance(t, np.ndarray):
        assert t.dtype == np.float64 and t.flags["C_CONTIGUOUS"]
        return t.ctypes.data
    assert t.is_contiguous(), "buffers must be contiguous"
    return t.data_ptr()


def chebyshev_radius(nr: int, ri: float, ro: float) -> np.ndarray:
    """Chebyshev-Gauss-Lobatto radii of the reference's shell grid
    (MHD/Fortran_src/CORE_src/SPH_SPECTR_src/chebyshev_radial_grid.f90,
    set_chebyshev_distance_shell).  The reference's spacing divides by
    nr - 1; a one-radius grid (a single spherical surface) is taken here as
    the inner radius."""
    if nr == 1:
        return np.array([ri], dtype=np.float64)
    k = np.arange(nr, dtype=np.float64)
    return ri + 0.5 * (ro - ri) * (1.0 - np.cos(np.pi * k / (nr - 1)))


@dataclass
class ShellConfig:
    """Resolution of one shell: truncation L, radii, Gauss and phi points."""
    L: int
    nr: int
    nth: int
    nphi: int
    ri: float = 0.35 / 0.65
    ro: float = 1.0 / 0.65

    @property
    def jmax(self):
        return (self.L + 1) ** 2

    @property
    def nslot(self):
        return 2 * self.L + 1

    def radius(self):
        return chebyshev_radius(self.nr, self.ri, self.ro)

    def n_rj(self, ncomp):
        return ncomp * self.jmax * self.nr

    def n_rtp(self, ncomp):
        return ncomp * self.nr * self.nth * self.nphi

    def n_rtm(self, ncomp):
        return ncomp * self.nth * self.nr * self.nslot

    def n_pole(self, ncomp):
        return ncomp * (2 * self.nr + 1)


class SphTransPlan:
    """One shell's transform plan on the current HIP device.

    Mirrors initialize_sph_trans (SS/init_sph_trans.f90:46-74); the transform
    methods mirror sph_backward_transforms / sph_forward_transforms
    (SS/t_sph_transforms.f90:74-123, 250-297) and the stage selectors of
    SS/t_legendre_trans_select.f90 and SS/t_sph_FFT_selector.F90."""

    def __init__(self, cfg: ShellConfig, max_ncomp: int = 9, precision: int = SHT_PREC_FP64,
                 n_slices: int = 0, radius=None, p_on_the_fly: bool = False,
                 check_finite: bool = False):
        self.cfg = cfg
        r = np.ascontiguousarray(cfg.radius() if radius is None else radius, dtype=np.float64)
        self._radius = r
        flags = ((SHT_FLAG_P_ON_THE_FLY if p_on_the_fly else 0) |
                 (SHT_FLAG_CHECK_FINITE if check_finite else 0))
        d = _Desc(cfg.L, cfg.nr, cfg.nth, cfg.nphi, r.ctypes.data_as(_dp), max_ncomp, precision,
                  n_slices, flags)
        h = _vp()
        _check(lib().sht_plan_create(ctypes.byref(d), ctypes.byref(h)), "sht_plan_create")
        self._h = h
        self.max_ncomp = max_ncomp
        self.precision = precision
        self._user_stream = False

    @classmethod
    def _adopt(cls, h, cfg, radius, max_ncomp, precision):
        """Wrap a plan handle created elsewhere (sht_plan_create_from_files)."""
        self = cls.__new__(cls)
        self.cfg = cfg
        self._radius = np.ascontiguousarray(radius, dtype=np.float64)
        self._h = h
        self.max_ncomp = max_ncomp
        self.precision = precision
        self._user_stream = False
        return self

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().sht_plan_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr: int | None):
        """Run the plan on this HIP stream from now on (None: back to the
        plan's own stream and to following torch's current stream)."""
        _check(lib().sht_plan_set_stream(self._h, stream_ptr), "sht_plan_set_stream")
        self._user_stream = stream_ptr is not None

    def _bind(self, *bufs):
        """Order a call on torch device tensors like a torch operation, unless
        set_stream chose a stream: on torch's current stream, and for torch's
        default stream (handle 0) on HIP's legacy default stream
        (hipStreamLegacy), the same queue.  (A separate blocking stream is
        not enough: an asynchronous host-to-device copy that torch has just
        queued on the default stream was seen still in flight when a kernel
        on such a stream started.)"""
        if self._user_stream:
            return
        s = _torch_stream(*bufs)
        if s is not None:
            _check(lib().sht_plan_set_stream(self._h, s or HIP_STREAM_LEGACY),
                   "sht_plan_set_stream")

    STAGES = ("legendre_backward", "fft_backward", "fft_forward", "legendre_forward", "exchange",
              "legendre_backward_vector_kernel", "legendre_forward_vector_kernel")

    def timer_enable(self, on: bool = True):
        _check(lib().sht_plan_timer_enable(self._h, int(on)), "sht_plan_timer_enable")

    def timer_read(self) -> dict:
        """{stage: (total_ms, launches)} since the last read (synchronises)."""
        n = len(self.STAGES)
        ms = (ctypes.c_double * n)()
        calls = (ctypes.c_longlong * n)()
        _check(lib().sht_plan_timer_read(self._h, ms, calls, n), "sht_plan_timer_read")
        return {s: (ms[i], int(calls[i])) for i, s in enumerate(self.STAGES)}

    def set_variant(self, name: str, value: int):
        """Kernel variant of this plan (sht_plan_set_variant)."""
        _check(lib().sht_plan_set_variant(self._h, name.encode(), int(value)),
               "sht_plan_set_variant")

    def get_variant(self, name: str) -> int:
        v = ctypes.c_int()
        _check(lib().sht_plan_get_variant(self._h, name.encode(), ctypes.byref(v)),
               "sht_plan_get_variant")
        return v.value

    VARIANTS = ("split_fwd_mfma", "fft_form")

    def autotune(self, nvector_b: int, nscalar_b: int, nvector_f: int, nscalar_f: int,
                 cache: str | None = None) -> dict:
        """Time the plan's kernel variants for these field sets and keep the
        fastest (sht_plan_autotune); returns the chosen variants and the
        round-trip ms of the winner and of the default."""
        ms = (ctypes.c_double * 2)()
        _check(lib().sht_plan_autotune(self._h, nvector_b, nscalar_b, nvector_f, nscalar_f,
                                       cache.encode() if cache else None, ms),
               "sht_plan_autotune")
        return {"variants": {v: self.get_variant(v) for v in self.VARIANTS},
                "ms": ms[0], "default_ms": ms[1]}

    def device_bytes(self) -> int:
        return int(lib().sht_plan_device_bytes(self._h))

    def table(self, name: str) -> np.ndarray:
        n = {"gpoint": self.cfg.nth, "weight": self.cfg.nth, "colat": self.cfg.nth,
             "asin_t": self.cfg.nth, "radius": self.cfg.nr}.get(name)
        if name == "lstack":
            out = np.zeros(self.cfg.nslot + 1, dtype=np.int32)
            _check(lib().sht_plan_get_table(self._h, b"lstack", out.ctypes.data, out.size),
                   "sht_plan_get_table")
            return out
        out = np.zeros(n)
        _check(lib().sht_plan_get_table(self._h, name.encode(), out.ctypes.data, n),
               "sht_plan_get_table")
        return out

    # whole transforms -------------------------------------------------------
    def sph_backward_transforms(self, nvector: int, nscalar: int, WS, v_rtp):
        """spectra (rj send buffer) -> grid fields, device buffers."""
        self._bind(WS, v_rtp)
        _check(lib().sht_backward(self._h, nvector, nscalar, _ptr(WS), _ptr(v_rtp)),
               "sht_backward")

    def sph_backward_transforms_noalloc(self, nvector: int, nscalar: int, WS, v_rtp):
        """sph_backward_transforms with the library's fixed device workspace
        (SHT_NOALLOC_SCRATCH_BYTES per device) instead of the plan's; the call
        allocates nothing.  Calls that overlap on one device race."""
        self._bind(WS, v_rtp)
        _check(lib().sht_backward_noalloc(self._h, nvector, nscalar, _ptr(WS), _ptr(v_rtp)),
               "sht_backward_noalloc")

    def sph_forward_transforms(self, nvector: int, nscalar: int, v_rtp, WR):
        """grid fields -> spectra (rj receive buffer), device buffers."""
        self._bind(v_rtp, WR)
        _check(lib().sht_forward(self._h, nvector, nscalar, _ptr(v_rtp), _ptr(WR)),
               "sht_forward")

    def pole_b_transform(self, nvector: int, nscalar: int, WS, v_pole, with_center=False):
        """spectra -> values at the poles (rows 0..Nr-1 north, Nr..2Nr-1
        south) and, for vectors with a centre, row 2Nr; v_pole holds
        ncomp * (2 Nr + 1) doubles, component-major (SS/t_sph_transforms.f90:
        203-245)."""
        self._bind(WS, v_pole)
        _check(lib().sht_pole_backward(self._h, nvector, nscalar, _ptr(WS), _ptr(v_pole),
                                       int(bool(with_center))), "sht_pole_backward")

    def sph_b_trans_w_poles(self, nvector: int, nscalar: int, WS, v_rtp, v_pole,
                            with_center=False):
        """pole_b_transform and sph_backward_transforms on the same spectra
        (SS/t_sph_transforms.f90:127-199)."""
        self._bind(WS, v_rtp, v_pole)
        _check(lib().sht_backward_with_poles(self._h, nvector, nscalar, _ptr(WS), _ptr(v_rtp),
                                             _ptr(v_pole), int(bool(with_center))),
               "sht_backward_with_poles")

    def nonlinear_terms_in_rtp(self, idx: dict, coef: dict, ncomp_fld: int, fld_rtp,
                               ncomp_frc: int, frc_rtp):
        """Grid-space products (MS/cal_nonlinear_sph_MHD.f90:64-128) on device
        buffers in the grid layout.  idx maps input names (velo, vort, magne,
        current, temp, light) and output names (m_advect, lorentz, vp_induct,
        h_flux, c_flux, coriolis) to first components; absent outputs are off.
        coef maps output names to their coefficients."""
        d = _NlDesc()
        for n in NL_INPUTS + NL_OUTPUTS:
            setattr(d, f"i_{n}", int(idx.get(n, -1)))
        for n in NL_OUTPUTS:
            setattr(d, f"coef_{_NL_COEF[n]}", float(coef.get(n, 0.0)))
        self._bind(fld_rtp, frc_rtp)
        _check(lib().sht_nonlinear_rtp(self._h, ctypes.byref(d), ncomp_fld, _ptr(fld_rtp),
                                       ncomp_frc, _ptr(frc_rtp)), "sht_nonlinear_rtp")

    @staticmethod
    def _field_list(fields):
        arr = (_FieldCopy * max(len(fields), 1))()
        for i, (i_field, i_trans, nc) in enumerate(fields):
            arr[i] = _FieldCopy(int(i_field), int(i_trans), int(nc))
        return arr

    def copy_spectr_to_send(self, fields, ntot_phys: int, d_fld, ncomp_send: int, WS,
                            inod_export=None):
        """Field-major rj spectra d_fld[inod + nnod_rj c] -> interleaved send
        buffer WS (copy_mhd_spectr_to_send, MS/copy_sph_MHD_4_send_recv.f90:
        63-136).  fields: (i_field, i_trans, ncomp) triples, 0-based;
        inod_export: device int32 tensor of 1-based nodes or None (identity)."""
        nnod_rj = d_fld.numel() // ntot_phys
        nnod_send = WS.numel() // ncomp_send
        self._bind(d_fld, WS)
        _check(lib().sht_rj_fields_to_send(
            self._h, len(fields), self._field_list(fields), ntot_phys, nnod_rj, _ptr(d_fld),
            inod_export.data_ptr() if inod_export is not None else None, nnod_send, ncomp_send,
            _ptr(WS)), "sht_rj_fields_to_send")

    def copy_spectr_from_recv(self, fields, ncomp_recv: int, WR, ntot_phys: int, d_fld,
                              irev=None):
        """Interleaved receive buffer WR -> field-major rj spectra
        (copy_mhd_spectr_from_recv, MS/copy_sph_MHD_4_send_recv.f90:140-206);
        irev: device int32 tensor of 1-based records (nnod_recv + 1 = zero)
        or None (identity)."""
        nnod_rj = d_fld.numel() // ntot_phys
        nnod_recv = WR.numel() // ncomp_recv
        self._bind(WR, d_fld)
        _check(lib().sht_rj_fields_from_recv(
            self._h, len(fields), self._field_list(fields), ncomp_recv, nnod_recv, _ptr(WR),
            irev.data_ptr() if irev is not None else None, ntot_phys, nnod_rj, _ptr(d_fld)),
            "sht_rj_fields_from_recv")

    def coriolis(self, kr_icb: int = 0) -> "CoriolisRlm":
        """Coriolis term in rlm space on this plan (init_sum_coriolis_rlm)."""
        return CoriolisRlm(self, kr_icb)

    def backward_host(self, nvector, nscalar, WS: np.ndarray) -> np.ndarray:
        ncomp = 3 * nvector + nscalar
        out = np.empty(self.cfg.n_rtp(ncomp))
        WS = np.ascontiguousarray(WS, dtype=np.float64).ravel()
        _check(lib().sht_backward_host(self._h, nvector, nscalar, WS.ctypes.data, out.ctypes.data),
               "sht_backward_host")
        return out

    def forward_host(self, nvector, nscalar, v: np.ndarray) -> np.ndarray:
        ncomp = 3 * nvector + nscalar
        out = np.empty(self.cfg.n_rj(ncomp))
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        _check(lib().sht_forward_host(self._h, nvector, nscalar, v.ctypes.data, out.ctypes.data),
               "sht_forward_host")
        return out

    # stages -------------------------------------------------------------------
    def legendre_backward(self, nvector, nscalar, WR_rlm, WS_rtm):
        self._bind(WR_rlm, WS_rtm)
        _check(lib().sht_legendre_backward(self._h, nvector, nscalar, _ptr(WR_rlm), _ptr(WS_rtm)),
               "sht_legendre_backward")

    def legendre_forward(self, nvector, nscalar, WR_rtm, WS_rlm):
        self._bind(WR_rtm, WS_rlm)
        _check(lib().sht_legendre_forward(self._h, nvector, nscalar, _ptr(WR_rtm), _ptr(WS_rlm)),
               "sht_legendre_forward")

    def fft_backward(self, ncomp, WR_rtp, v_rtp):
        self._bind(WR_rtp, v_rtp)
        _check(lib().sht_fft_backward(self._h, ncomp, _ptr(WR_rtp), _ptr(v_rtp)), "sht_fft_backward")

    def fft_forward(self, ncomp, v_rtp, WS_rtp):
        self._bind(v_rtp, WS_rtp)
        _check(lib().sht_fft_forward(self._h, ncomp, _ptr(v_rtp), _ptr(WS_rtp)), "sht_fft_forward")

    def exchange_rtm_to_rtp(self, ncomp, rtm, rtp):
        self._bind(rtm, rtp)
        _check(lib().sht_exchange_rtm_to_rtp(self._h, ncomp, _ptr(rtm), _ptr(rtp)),
               "sht_exchange_rtm_to_rtp")

    def exchange_rtp_to_rtm(self, ncomp, rtp, rtm):
        self._bind(rtp, rtm)
        _check(lib().sht_exchange_rtp_to_rtm(self._h, ncomp, _ptr(rtp), _ptr(rtm)),
               "sht_exchange_rtp_to_rtm")

    def exchange_gather_rev(self, ncomp, n_dst, n_src, irev, src, dst):
        assert irev.is_contiguous()
        self._bind(src, dst)
        _check(lib().sht_exchange_gather_rev(self._h, ncomp, n_dst, n_src, irev.data_ptr(),
                                             _ptr(src), _ptr(dst)), "sht_exchange_gather_rev")


class CoriolisRlm:
    """Coriolis term in rlm space (MS/t_coriolis_terms_rlm.f90): tables of
    init_sum_coriolis_rlm at creation, then sum_coriolis_rlm after the
    rj -> rlm exchange and copy_coriolis_terms_rlm after the forward
    Legendre transform (MS/sph_trans_w_coriols.f90:146-149, 233-235)."""

    def __init__(self, plan: "SphTransPlan", kr_icb: int = 0):
        self.plan = plan
        h = _vp()
        _check(lib().sht_coriolis_create(plan.handle, kr_icb, ctypes.byref(h)),
               "sht_coriolis_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().sht_coriolis_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sum_coriolis_rlm(self, ncomp_trans: int, i_velo: int, i_vort: int, coef_cor: float, WR,
                         d_cor, irev=None):
        """Curl of -coef_cor (z x u): d_cor[i] poloidal, d_cor[i + nnod_rlm]
        toroidal.  i_velo / i_vort: 0-based first components in WR;
        irev: device int32 tensor of 1-based records or None (identity)."""
        nnod_recv = WR.numel() // ncomp_trans
        self.plan._bind(WR, d_cor)
        _check(lib().sht_sum_coriolis_rlm(
            self._h, ncomp_trans, i_velo, i_vort, float(coef_cor), _ptr(WR), nnod_recv,
            irev.data_ptr() if irev is not None else None, _ptr(d_cor)), "sht_sum_coriolis_rlm")

    def copy_coriolis_terms_rlm(self, ncomp_trans: int, i_rot_coriolis: int, d_cor, WS,
                                item=None):
        """Both parts of d_cor to components i_rot_coriolis and +2 of WS."""
        nnod_send = WS.numel() // ncomp_trans
        self.plan._bind(d_cor, WS)
        _check(lib().sht_copy_coriolis_terms_rlm(
            self._h, ncomp_trans, i_rot_coriolis, _ptr(d_cor),
            item.data_ptr() if item is not None else None, nnod_send, _ptr(WS)),
            "sht_copy_coriolis_terms_rlm")


class SphFiles:
    """A rank's spherical grid / mode files as gen_sph_grids writes them (rtp,
    rtm, rlm, rj), read by sht_sph_files_read: mirrors the reference's
    read_geom_rtp_data / read_geom_rtm_data / read_spectr_modes_rlm_data /
    read_spectr_modes_rj_data (CORE_src/IO/sph_modes_grids_data_IO.f90:53-183)."""

    SPACES = ("rtp", "rtm", "rlm", "rj")

    def __init__(self, rtp: str, rtm: str, rlm: str, rj: str):
        h = _vp()
        _check(lib().sht_sph_files_read(rtp.encode(), rtm.encode(), rlm.encode(), rj.encode(),
                                        ctypes.byref(h)), "sht_sph_files_read")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().sht_sph_files_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self):
        """Raise ShtError unless the set is the single-rank layout."""
        _check(lib().sht_sph_files_check(self._h), "sht_sph_files_check")

    def shape(self):
        """(ShellConfig, radii, nnod_rj) of the set."""
        dims = (ctypes.c_int * 5)()
        _check(lib().sht_sph_files_shape(self._h, dims, None), "sht_sph_files_shape")
        r = np.empty(dims[1])
        _check(lib().sht_sph_files_shape(self._h, dims, r.ctypes.data_as(_dp)),
               "sht_sph_files_shape")
        return ShellConfig(dims[0], dims[1], dims[2], dims[3]), r, dims[4]

    def table(self, name: str) -> np.ndarray:
        """A table of the set (sht_sph_files_table), as int32."""
        n = lib().sht_sph_files_table(self._h, name.encode(), None, 0)
        if n < 0:
            raise ShtError(int(-n), "sht_sph_files_table")
        out = np.empty(n, dtype=np.int32)
        if n:
            m = lib().sht_sph_files_table(self._h, name.encode(),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n)
            if m < 0:
                raise ShtError(int(-m), "sht_sph_files_table")
        return out

    def plan(self, max_ncomp: int = 9, precision: int = SHT_PREC_FP64, n_slices: int = 0,
             p_on_the_fly: bool = False) -> "SphTransPlan":
        """The plan of this set's shape and radii (sht_plan_create_from_files)."""
        cfg, r, _ = self.shape()
        h = _vp()
        _check(lib().sht_plan_create_from_files(self._h, max_ncomp, precision, n_slices,
                                                SHT_FLAG_P_ON_THE_FLY if p_on_the_fly else 0,
                                                ctypes.byref(h)), "sht_plan_create_from_files")
        return SphTransPlan._adopt(h, cfg, r, max_ncomp, precision)


class ShellBatch:
    """Many independent shells replayed as one HIP graph (sht_batch_create).

    ``calls``: sequence of (plan, (nv_b, ns_b), WS, v_rtp, (nv_f, ns_f), v_in,
    WR) with device tensors; a field set of (0, 0) skips that direction (its
    buffers may be None).  The graph keeps the plans and the buffer
    addresses: keep them alive, refill inputs in place between launches.
    No counterpart in the reference, which transforms one shell per call
    (SS/t_sph_transforms.f90:74,250)."""

    def __init__(self, calls):
        arr = (_ShellCall * len(calls))()
        self._keep = []
        for i, (plan, fb, WS, v, ff, vin, WR) in enumerate(calls):
            self._keep.append((plan, WS, v, vin, WR))
            arr[i] = _ShellCall(plan.handle, fb[0], fb[1], _ptr(WS) if WS is not None else None,
                                _ptr(v) if v is not None else None, ff[0], ff[1],
                                _ptr(vin) if vin is not None else None,
                                _ptr(WR) if WR is not None else None)
        h = _vp()
        _check(lib().sht_batch_create(arr, len(calls), ctypes.byref(h)), "sht_batch_create")
        self._h = h

    def launch(self, stream_ptr: int | None = None):
        """One replay of every shell's transforms (None: the batch's own stream)."""
        _check(lib().sht_batch_launch(self._h, stream_ptr), "sht_batch_launch")

    def synchronize(self):
        _check(lib().sht_batch_synchronize(self._h), "sht_batch_synchronize")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().sht_batch_destroy(self._h)
            self._h = _vp()
        self._keep = []  # the plans and buffers the graph used

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def selftest() -> int:
    return int(lib().sht_selftest())


# ---------------------------------------------------------------------------
# Synthetic seeded spectra: splitmix64 -> Box-Muller, sigma(l, r) =
# (l+1)^-p * sin(pi (r - ri)/(ro - ri)); complex c = a + i b per (l, m>=0)
# mapped to the real harmonics as slot (l,+m) <- a, (l,-m) <- b, m = 0 takes a.
def _splitmix64(state: np.ndarray) -> np.ndarray:
    z = state + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _uniform(seed: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        idx = np.arange(n, dtype=np.uint64) + np.uint64(seed) * np.uint64(0x632BE59BD9B4E019)
        bits = _splitmix64(idx)
    return ((bits >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def gaussian(seed: int, n: int) -> np.ndarray:
    u1 = _uniform(2 * seed + 1, n)
    u2 = _uniform(2 * seed + 2, n)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def batched_shell_draw(n: int = 64, seed: int = 20261202):
    """Seeded uneven shells of the batched-shell workload: L drawn from the
    odd values in 63..383, Nr from 33..129, Ntheta = 3(L+1)/2 Gauss points
    and Nphi = 2 Ntheta longitudes (1 vector + 1 scalar per shell each way).
    Returns a list of ShellConfig."""
    rng = np.random.default_rng(seed)
    Ls = [int(x) for x in rng.choice(np.arange(63, 384, 2), size=n)]
    nrs = [int(x) for x in rng.integers(33, 130, size=n)]
    return [ShellConfig(L, nr, 3 * (L + 1) // 2, 3 * (L + 1)) for L, nr in zip(Ls, nrs)]


def rlm_l_m(L: int):
    """(l, m) of each j in the single-rank rlm order (m = -L..L, l = |m|..L)."""
    ls, ms = [], []
    for m in range(-L, L + 1):
        for l in range(abs(m), L + 1):
            ls.append(l)
            ms.append(m)
    return np.array(ls), np.array(ms)


def synthetic_spectra(cfg: ShellConfig, nvector: int, nscalar: int, seed: int = 1,
                      slope: float = 1.5) -> np.ndarray:
    """Comp-interleaved rj buffer sp[nd + ncomp*(j + jmax*k)]."""
    ncomp = 3 * nvector + nscalar
    l, m = rlm_l_m(cfg.L)
    r = cfg.radius()
    prof = np.sin(np.pi * (r - cfg.ri) / (cfg.ro - cfg.ri))
    sig = (l + 1.0) ** (-slope)
    out = np.empty((cfg.nr, cfg.jmax, ncomp))
    for nd in range(ncomp):
        # one complex draw per (l, m >= 0, r): real part to +m, imaginary to -m
        g = gaussian(seed * 1000 + nd, 2 * cfg.jmax * cfg.nr).reshape(2, cfg.nr, cfg.jmax)
        re = g[0]
        # the imaginary part of (l,|m|) is stored at (l,-|m|): index by the partner
        out[:, :, nd] = np.where(m[None, :] >= 0, re, g[1]) * sig[None, :] * prof[:, None]
    return out.ravel()
