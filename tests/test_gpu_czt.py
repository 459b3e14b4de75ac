"""The zonal transform as one chirp-z (Bluestein) convolution per line
(plan variant fft_form = 3) against the oracle: shells whose lengths take the
split form (Nphi = 3 * 64), a prime 23 (Nphi = 2 * 3 * 23), the Bluestein
prime 53 with an odd line count (Nphi = 318, Nr Ntheta = 477), a length whose
Nphi + 2L is exactly the convolution length (342 + 170 = 512), and the largest
shell of the 64-shell draw (L = 361, Nphi = 1086 = 2 * 3 * 181, convolution
length 2048): whole transforms (FP64 engine 1e-12, split 1e-11 per component)
and agreement with the default form to 1e-13; the FFT stage entry points on
the reference's rtp buffers; graph replays equal to eager calls; and the
length limit."""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellBatch, ShellConfig,
                             SphTransPlan, ShtError, synthetic_spectra)

pytestmark = pytest.mark.gpu

SHELLS = {"L63_split_form": ShellConfig(63, 5, 96, 192),
          "L45_prime23": ShellConfig(45, 4, 69, 138),
          "L105_prime53_odd_lines": ShellConfig(105, 3, 159, 318),
          "L85_exact_length": ShellConfig(85, 4, 171, 342),
          "L361_prime181": ShellConfig(361, 2, 543, 1086)}


def _rel(a, b, ncomp, comp_fastest):
    a = np.asarray(a).reshape(-1, ncomp) if comp_fastest else np.asarray(a).reshape(ncomp, -1).T
    b = np.asarray(b).reshape(-1, ncomp) if comp_fastest else np.asarray(b).reshape(ncomp, -1).T
    return max(np.linalg.norm(a[:, c] - b[:, c]) / np.linalg.norm(b[:, c]) for c in range(ncomp))


def _run(t, plan, cfg, sp, grid):
    v = t.full((cfg.n_rtp(8),), np.nan, dtype=t.float64, device="cuda")
    s = t.full((cfg.n_rj(9),), np.nan, dtype=t.float64, device="cuda")
    plan.sph_backward_transforms(2, 2, t.from_numpy(sp).cuda(), v)
    plan.sph_forward_transforms(3, 0, t.from_numpy(grid).cuda(), s)
    t.cuda.synchronize()
    return v.cpu().numpy(), s.cpu().numpy()


@pytest.mark.parametrize("name", list(SHELLS))
@pytest.mark.parametrize("prec,tol", [(SHT_PREC_FP64, 1e-12), (SHT_PREC_SPLIT_INT8, 1e-11)],
                         ids=["fp64", "split"])
def test_czt_form_vs_oracle(torch_cuda, name, prec, tol):
    t = torch_cuda
    cfg = SHELLS[name]
    print(f"[gpu] chirp-z form {name} precision={prec}", flush=True)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    sp = synthetic_spectra(cfg, 2, 2, seed=211)
    grid = orc.backward(3, 0, synthetic_spectra(cfg, 3, 0, seed=213))
    plan = SphTransPlan(cfg, max_ncomp=9, precision=prec)
    try:
        ref = _run(t, plan, cfg, sp, grid)
        plan.set_variant("fft_form", 3)
        assert plan.get_variant("fft_form") == 3
        v, s = _run(t, plan, cfg, sp, grid)
    finally:
        plan.close()
    eb = _rel(v, orc.backward(2, 2, sp), 8, False)
    ef = _rel(s, orc.forward(3, 0, grid), 9, True)
    print(f"    backward {eb:.2e} forward {ef:.2e}", flush=True)
    assert eb <= tol and ef <= tol
    assert _rel(v, ref[0], 8, False) <= 1e-13 and _rel(s, ref[1], 9, True) <= 1e-13


def test_czt_fft_stages_vs_oracle(torch_cuda):
    """sht_fft_backward / sht_fft_forward on the reference's rtp buffers
    (FFTPACK slot addressing) with the chirp-z form."""
    t = torch_cuda
    cfg = SHELLS["L105_prime53_odd_lines"]
    ncomp = 3
    print("[gpu] chirp-z FFT stages L=105 Nphi=318", flush=True)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    plan = SphTransPlan(cfg, max_ncomp=ncomp)
    gen = t.Generator(device="cuda")
    gen.manual_seed(515)
    try:
        plan.set_variant("fft_form", 3)
        rtm = t.randn(cfg.n_rtm(ncomp), generator=gen, dtype=t.float64, device="cuda")
        rtp = t.empty(cfg.n_rtp(ncomp), dtype=t.float64, device="cuda")
        plan.exchange_rtm_to_rtp(ncomp, rtm, rtp)
        v = t.empty(cfg.n_rtp(ncomp), dtype=t.float64, device="cuda")
        plan.fft_backward(ncomp, rtp, v)
        g = t.randn(cfg.n_rtp(ncomp), generator=gen, dtype=t.float64, device="cuda")
        out = t.empty(cfg.n_rtp(ncomp), dtype=t.float64, device="cuda")
        plan.fft_forward(ncomp, g, out)
        t.cuda.synchronize()
        ref = orc.fft_bwd_from_rtp(ncomp, orc.rtm_to_rtp(ncomp, rtm.cpu().numpy()))
        eb = _rel(v.cpu().numpy(), ref, ncomp, comp_fastest=False)
        ref_rtm = orc.rtp_to_rtm(ncomp, orc.fft_fwd_to_rtp(ncomp, g.cpu().numpy()))
        got_rtm = orc.rtp_to_rtm(ncomp, out.cpu().numpy())
        ef = _rel(got_rtm, ref_rtm, ncomp, comp_fastest=True)
    finally:
        plan.close()
    print(f"    backward {eb:.2e} forward {ef:.2e}", flush=True)
    assert eb <= 1e-12 and ef <= 1e-12


def test_czt_form_under_graph_replay(torch_cuda):
    torch = torch_cuda
    cfgs = [SHELLS["L63_split_form"], SHELLS["L45_prime23"], SHELLS["L105_prime53_odd_lines"]]
    plans, bufs = [], []
    for i, c in enumerate(cfgs):
        p = SphTransPlan(c, max_ncomp=4, precision=SHT_PREC_SPLIT_INT8 if i % 2 else SHT_PREC_FP64)
        p.set_variant("fft_form", 3)
        ws = torch.tensor(synthetic_spectra(c, 1, 1, seed=220 + i), device="cuda")
        vin = torch.tensor(np.random.default_rng(230 + i).standard_normal(c.n_rtp(4)), device="cuda")
        vout = torch.zeros(c.n_rtp(4), dtype=torch.float64, device="cuda")
        wr = torch.zeros(c.n_rj(4), dtype=torch.float64, device="cuda")
        p.sph_backward_transforms(1, 1, ws, vout)
        p.sph_forward_transforms(1, 1, vin, wr)
        plans.append(p)
        bufs.append((ws, vout, vin, wr))
    torch.cuda.synchronize()
    ref = [(b[1].clone(), b[3].clone()) for b in bufs]
    batch = ShellBatch([(p, (1, 1), ws, vout, (1, 1), vin, wr)
                        for p, (ws, vout, vin, wr) in zip(plans, bufs)])
    try:
        for rep in range(3):
            for (_, vout, _, wr) in bufs:
                vout.fill_(np.nan)
                wr.fill_(np.nan)
            torch.cuda.synchronize()
            batch.launch()
            batch.synchronize()
            for i, ((_, vout, _, wr), (rv, rw)) in enumerate(zip(bufs, ref)):
                assert torch.equal(vout, rv) and torch.equal(wr, rw), f"replay {rep} shell {i}"
            plans[0].sph_backward_transforms(1, 1, bufs[0][0], bufs[0][1])
            torch.cuda.synchronize()
    finally:
        batch.close()
        for p in plans:
            p.close()


def test_czt_length_limit(torch_cuda):
    """Nphi + 2L above 2048 is refused; the plan keeps its form."""
    cfg = ShellConfig(400, 2, 401, 1250)
    plan = SphTransPlan(cfg, max_ncomp=2)
    try:
        with pytest.raises(ShtError, match="2048"):
            plan.set_variant("fft_form", 3)
        assert plan.get_variant("fft_form") != 3
    finally:
        plan.close()
