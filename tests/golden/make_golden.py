#!/usr/bin/env python3
"""Write tests/golden/ref_fixtures.npz from the reference's own numerical
routines, compiled from the reference sources by oracle/Makefile.ref into
oracle/_ref (built only where the reference tree is present).

Contents (inputs are either stored or regenerated from the seeds below):
  gauss_x_<n>, gauss_w_<n>   Gauss-Legendre points / weights with the
                             reference Newton loop, n in NTH
  plm_<m>, dplm_<m>          Schmidt P_l^m(theta) and dP/dtheta, l = 0..31,
                             at the colatitudes THETAS, m in MS
  g_<l>_<m>                  the 17 normalisation factors of (l, m)
  fft_in, fft_fwd, fft_bwd   the six test signals of the reference's FFT test
                             program (N = 512), their forward transform and
                             the backward transform of that
  fftr_fwd_<n>               forward transform of 3 seeded lines
                             gaussian(FFT_SEED + n, 3 n), n in FFT_N
  leg_b_out                  Legendre backward (2 vectors + 2 scalars) of the
                             seeded spectra at L=31, Ntheta=48, 2 radii
  leg_f_out                  Legendre forward (3 vectors) of the seeded rtm
                             field gaussian(SEED_F, size)
Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from kemorin_mhd_amd.sht import ShellConfig, gaussian, synthetic_spectra  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_fixtures.npz")
NTH = (48, 384, 768)
MS = (0, 1, 5, 31)
THETAS = np.array([1e-3, 0.3, 1.0, 1.5707963267948966, 2.2, 3.1])
FFT_N = (16, 96, 250, 1536)
# Legendre case: L=31, Ntheta=48, radii 3 and 17 of the 33-point shell
LEG = dict(L=31, nth=48, nphi=96, radii=(3, 17))
SEED_B, SEED_F, FFT_SEED = 101, 103, 5000


def fft_test_signals(n=512):
    j = np.arange(n)
    a = 2 * np.pi * j / n
    return np.stack([np.full(n, 10.0), 2.0 * np.sin(a), 3.0 * np.cos(a), -4.0 * np.sin(3 * a),
                     -5.0 * np.cos(4 * a), 6.0 * np.sin(10 * a) + 1.5 * np.cos(8 * a)])


def leg_shell():
    full = ShellConfig(LEG["L"], 33, LEG["nth"], LEG["nphi"])
    r = full.radius()[list(LEG["radii"])]
    return ShellConfig(LEG["L"], len(r), LEG["nth"], LEG["nphi"]), r


def main():
    if not ref.build():
        sys.exit("oracle/_ref is not built (needs the reference tree)")
    d = {}
    for n in NTH:
        d[f"gauss_x_{n}"], d[f"gauss_w_{n}"] = ref.gauss_points(n)
    for m in MS:
        ps, dps = zip(*(ref.schmidt_legendres_m(31, m, th) for th in THETAS))
        d[f"plm_{m}"], d[f"dplm_{m}"] = np.array(ps), np.array(dps)
    for l in range(0, 32, 5):
        for m in (0, 1, l):
            if m <= l:
                d[f"g_{l}_{m}"] = ref.sph_normalizations(l, m)
    x = fft_test_signals()
    d["fft_in"] = x
    d["fft_fwd"] = ref.rfft_fwd(x)
    d["fft_bwd"] = ref.rfft_bwd(d["fft_fwd"])
    for n in FFT_N:
        d[f"fftr_fwd_{n}"] = ref.rfft_fwd(gaussian(FFT_SEED + n, 3 * n).reshape(3, n))
    cfg, r = leg_shell()
    plan = ref.RefPlan(cfg.L, cfg.nr, cfg.nth, r)
    sp = synthetic_spectra(cfg, 2, 2, seed=SEED_B)
    d["leg_b_out"] = plan.leg_bwd(2, 2, sp)
    vr = gaussian(SEED_F, 9 * cfg.nth * cfg.nr * (2 * cfg.L + 1))
    d["leg_f_out"] = plan.leg_fwd(3, 0, vr)
    np.savez_compressed(OUT, **d)
    print(f"wrote {OUT} ({os.path.getsize(OUT) / 1e3:.0f} kB, {len(d)} arrays)")


if __name__ == "__main__":
    main()
