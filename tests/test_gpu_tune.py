"""Tile-shape autotuner (sht_plan_autotune): it picks one of the plan's
variants by timing, caches the winner in a JSON file keyed by shape, and a
tuned plan is held to the same oracle bars as the default (the variants
change only the time)."""
import json

import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellConfig, SphTransPlan, synthetic_spectra

pytestmark = pytest.mark.gpu


def _rel(a, b, ncomp, comp_fastest):
    a = np.asarray(a).reshape(-1, ncomp) if comp_fastest else np.asarray(a).reshape(ncomp, -1).T
    b = np.asarray(b).reshape(-1, ncomp) if comp_fastest else np.asarray(b).reshape(ncomp, -1).T
    return max(np.linalg.norm(a[:, c] - b[:, c]) / np.linalg.norm(b[:, c]) for c in range(ncomp))


@pytest.mark.parametrize("prec,tol", [(SHT_PREC_FP64, 1e-12), (SHT_PREC_SPLIT_INT8, 1e-11)],
                         ids=["fp64", "split"])
def test_autotune_cache_and_parity(torch_cuda, tmp_path, prec, tol):
    t = torch_cuda
    cfg = ShellConfig(95, 7, 144, 384)  # Nphi = 3 * 128 with 128 > L: both FFT forms apply
    cache = str(tmp_path / "tune.json")
    print(f"[gpu] autotune L={cfg.L} precision={prec}", flush=True)
    plan = SphTransPlan(cfg, max_ncomp=9, precision=prec)
    try:
        r = plan.autotune(2, 2, 3, 0, cache=cache)
        print("    ", r, flush=True)
        assert r["ms"] > 0 and r["default_ms"] >= r["ms"]
        assert r["variants"]["split_fwd_mfma"] in (16, 32)
        assert r["variants"]["fft_form"] in (0, 1, 2, 3)
        if prec == SHT_PREC_FP64:
            assert r["variants"]["split_fwd_mfma"] == 16
        entries = json.load(open(cache))["entries"]
        assert len(entries) == 1 and entries[0]["split_fwd_mfma"] == r["variants"]["split_fwd_mfma"]
        orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
        sp = synthetic_spectra(cfg, 2, 2, seed=91)
        grid = orc.backward(3, 0, synthetic_spectra(cfg, 3, 0, seed=93))
        v = t.empty(cfg.n_rtp(8), dtype=t.float64, device="cuda")
        s = t.empty(cfg.n_rj(9), dtype=t.float64, device="cuda")
        plan.sph_backward_transforms(2, 2, t.from_numpy(sp).cuda(), v)
        plan.sph_forward_transforms(3, 0, t.from_numpy(grid).cuda(), s)
        t.cuda.synchronize()
        assert _rel(v.cpu().numpy(), orc.backward(2, 2, sp), 8, False) <= tol
        assert _rel(s.cpu().numpy(), orc.forward(3, 0, grid), 9, True) <= tol
    finally:
        plan.close()
    # a second plan of the same shape takes the cached winner
    plan2 = SphTransPlan(cfg, max_ncomp=9, precision=prec)
    try:
        plan2.set_variant("fft_form", 0 if r["variants"]["fft_form"] else 1)
        r2 = plan2.autotune(2, 2, 3, 0, cache=cache)
        assert r2["variants"] == r["variants"] and r2["ms"] == pytest.approx(r["ms"], abs=1e-3)
        # another field set is another entry
        plan2.autotune(1, 1, 1, 1, cache=cache)
        assert len(json.load(open(cache))["entries"]) == 2
    finally:
        plan2.close()
