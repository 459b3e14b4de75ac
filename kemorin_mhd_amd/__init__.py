"""MI355X (gfx950) spherical-harmonic transform engine for Kemorin/Calypso.

Drop-in for the reference's sph_trans hot path (Legendre rj<->rtm, zonal FFT
rtm<->rtp, exchange reorders) behind the C ABI in include/kemorin_sht.h.
"""
from .sht import (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, CoriolisRlm, ShellBatch, ShellConfig, ShtError,
                  SphFiles, SphTransPlan, batched_shell_draw, chebyshev_radius, gaussian, last_error, lib,
                  rlm_l_m, selftest, synthetic_spectra)

__all__ = ["SHT_PREC_FP64", "SHT_PREC_SPLIT_INT8", "CoriolisRlm", "ShellBatch", "ShellConfig", "ShtError",
           "SphFiles", "SphTransPlan", "batched_shell_draw", "chebyshev_radius", "gaussian", "last_error",
           "lib", "rlm_l_m", "selftest", "synthetic_spectra"]
