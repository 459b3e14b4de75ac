"""Parity at the benchmark configurations.

* Config 1 (L=255, Nr=129, Ntheta=384, Nphi=768), FP64 engine against the
  oracle: whole backward (2 vector + 2 scalar) and forward (3 vector)
  transforms within relative L2 1e-12 per component; the rtm <-> rtp
  exchanges bit-exact.  The GPU transforms all 129 radii; the oracle
  (radii are independent: only r enters) takes 9 of them, both boundaries
  included, so the test stays within seconds on the host.
* Config 2 (L=511, Nr=257, Ntheta=768, Nphi=1536), where the CPU oracle
  would take minutes, through size-independent properties:
  * the split engine (6 slices) within relative L2 1e-11 per component of
    the FP64 engine, for the (l+1)^-1.5 and the steep (l+1)^-3 spectra;
  * the FP64 backward linear to 1e-13.
  The FP64 engine itself is pinned to the oracle at the smaller sizes.
"""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellConfig, SphTransPlan,
                             rlm_l_m)

pytestmark = pytest.mark.gpu

CFG1 = ShellConfig(255, 129, 384, 768)
CFG2 = ShellConfig(511, 257, 768, 1536)


def _say(*a):
    print(*a, flush=True)


def spectra(cfg, ncomp, seed, slope=1.5):
    """Seeded rj buffer sp[nd + ncomp*(j + jmax*k)], sigma = (l+1)^-slope
    sin(pi (r-ri)/(ro-ri)) (numpy generator: fast at these sizes)."""
    l, _ = rlm_l_m(cfg.L)
    r = cfg.radius()
    prof = np.sin(np.pi * (r - cfg.ri) / (cfg.ro - cfg.ri))
    x = np.random.default_rng(seed).standard_normal((cfg.nr, cfg.jmax, ncomp))
    x *= ((l + 1.0) ** -slope)[None, :, None]
    x *= prof[:, None, None]
    return x.ravel()


def spectra_dev(t, cfg, ncomp, seed, slope=1.5):
    """The same distribution drawn on the device (config 2 sizes)."""
    l, _ = rlm_l_m(cfg.L)
    r = cfg.radius()
    gen = t.Generator(device="cuda")
    gen.manual_seed(seed)
    x = t.randn((cfg.nr, cfg.jmax, ncomp), generator=gen, dtype=t.float64, device="cuda")
    x *= t.tensor((l + 1.0) ** -slope, device="cuda")[None, :, None]
    x *= t.tensor(np.sin(np.pi * (r - cfg.ri) / (cfg.ro - cfg.ri)), device="cuda")[:, None, None]
    return x.reshape(-1)


def rel_comp_np(a, b, ncomp, comp_fastest):
    a = np.asarray(a).reshape(-1, ncomp) if comp_fastest else np.asarray(a).reshape(ncomp, -1).T
    b = np.asarray(b).reshape(-1, ncomp) if comp_fastest else np.asarray(b).reshape(ncomp, -1).T
    return max(np.linalg.norm(a[:, c] - b[:, c]) / np.linalg.norm(b[:, c]) for c in range(ncomp))


def rel_comp_dev(t, a, b, ncomp, comp_fastest):
    """Per-component relative L2 of device tensors, computed on the device."""
    a2 = a.view(-1, ncomp).t() if comp_fastest else a.view(ncomp, -1)
    b2 = b.view(-1, ncomp).t() if comp_fastest else b.view(ncomp, -1)
    num = t.linalg.vector_norm(a2 - b2, dim=1)
    den = t.linalg.vector_norm(b2, dim=1)
    return float((num / den).max())


def test_cfg1_fp64_vs_oracle(torch_cuda):
    t = torch_cuda
    cfg = CFG1
    _say("[gpu] config 1 L=255 Nr=129: FP64 engine vs oracle on 9 radii")
    ks = np.linspace(0, cfg.nr - 1, 9).round().astype(int)
    r_s = cfg.radius()[ks]
    ns_r = len(ks)
    orc = O.OraclePlan(cfg.L, ns_r, cfg.nth, cfg.nphi, r_s, full_tables=False)
    plan = SphTransPlan(cfg, max_ncomp=9, precision=SHT_PREC_FP64)
    try:
        # backward, 2 vectors + 2 scalars
        sp = spectra(cfg, 8, 101)
        v = t.empty(cfg.n_rtp(8), dtype=t.float64, device="cuda")
        plan.sph_backward_transforms(2, 2, t.from_numpy(sp).cuda(), v)
        t.cuda.synchronize()
        sp_s = sp.reshape(cfg.nr, cfg.jmax, 8)[ks].ravel()
        ref = orc.backward(2, 2, sp_s).reshape(8, cfg.nphi, cfg.nth, ns_r)
        got = v.view(8, cfg.nphi, cfg.nth, cfg.nr)[:, :, :, t.from_numpy(ks).cuda()].cpu().numpy()
        eb = rel_comp_np(got, ref, 8, comp_fastest=False)
        del v, got, ref, sp
        # forward, 3 vectors, of a band-limited grid made by the GPU backward
        g = t.empty(cfg.n_rtp(9), dtype=t.float64, device="cuda")
        plan.sph_backward_transforms(3, 0, t.from_numpy(spectra(cfg, 9, 103)).cuda(), g)
        s = t.empty(cfg.n_rj(9), dtype=t.float64, device="cuda")
        plan.sph_forward_transforms(3, 0, g, s)
        t.cuda.synchronize()
        g_s = g.view(9, cfg.nphi, cfg.nth, cfg.nr)[:, :, :, t.from_numpy(ks).cuda()]
        ref_s = orc.forward(3, 0, g_s.contiguous().cpu().numpy().ravel())
        got_s = s.view(cfg.nr, cfg.jmax, 9)[t.from_numpy(ks).cuda()].cpu().numpy()
        ef = rel_comp_np(got_s, ref_s, 9, comp_fastest=True)
        _say(f"    backward max rel L2 {eb:.3e}  forward max rel L2 {ef:.3e}")
        assert eb <= 1e-12 and ef <= 1e-12
    finally:
        plan.close()


def test_cfg1_exchanges_bitexact(torch_cuda):
    """All radii through the GPU exchanges; the oracle's reorder on 9 of them
    (the permutation acts on each radius alone); the inverse exchange gives
    the input back bit for bit."""
    t = torch_cuda
    cfg = CFG1
    ncomp = 9
    _say("[gpu] config 1 exchanges")
    ks = np.linspace(0, cfg.nr - 1, 9).round().astype(int)
    kd = t.from_numpy(ks).cuda()
    orc = O.OraclePlan(cfg.L, len(ks), cfg.nth, cfg.nphi, cfg.radius()[ks], full_tables=False)
    plan = SphTransPlan(cfg, max_ncomp=ncomp)
    try:
        gen = t.Generator(device="cuda")
        gen.manual_seed(107)
        rtm = t.randn(cfg.n_rtm(ncomp), generator=gen, dtype=t.float64, device="cuda")
        rtp = t.full((cfg.n_rtp(ncomp),), 7.0, dtype=t.float64, device="cuda")
        plan.exchange_rtm_to_rtp(ncomp, rtm, rtp)
        t.cuda.synchronize()
        nslot = 2 * cfg.L + 1
        rtm_s = rtm.view(nslot, cfg.nr, cfg.nth, ncomp)[:, kd].contiguous().cpu().numpy()
        ref = orc.rtm_to_rtp(ncomp, rtm_s.ravel())
        got = rtp.view(cfg.nphi, cfg.nth, cfg.nr, ncomp)[:, :, kd].contiguous().cpu().numpy()
        assert np.array_equal(got.ravel(), ref)
        back = t.zeros(cfg.n_rtm(ncomp), dtype=t.float64, device="cuda")
        plan.exchange_rtp_to_rtm(ncomp, rtp, back)
        t.cuda.synchronize()
        assert bool(t.equal(back, rtm))
    finally:
        plan.close()


def _round_trip(t, cfg, prec, sp_b, grid_f):
    plan = SphTransPlan(cfg, max_ncomp=9, precision=prec)
    try:
        v = t.empty(cfg.n_rtp(8), dtype=t.float64, device="cuda")
        s = t.empty(cfg.n_rj(9), dtype=t.float64, device="cuda")
        plan.sph_backward_transforms(2, 2, sp_b, v)
        plan.sph_forward_transforms(3, 0, grid_f, s)
        t.cuda.synchronize()
        return v, s
    finally:
        plan.close()


@pytest.mark.parametrize("slope", [1.5, 3.0])
def test_cfg2_split_matches_fp64(torch_cuda, slope):
    t = torch_cuda
    cfg = CFG2
    _say(f"[gpu] config 2 L=511 Nr=257: split vs FP64, slope {slope}")
    sp_b = spectra_dev(t, cfg, 8, 201, slope)
    sp_f = spectra_dev(t, cfg, 9, 203, slope)
    plan = SphTransPlan(cfg, max_ncomp=9, precision=SHT_PREC_FP64)
    try:
        grid_f = t.empty(cfg.n_rtp(9), dtype=t.float64, device="cuda")
        plan.sph_backward_transforms(3, 0, sp_f, grid_f)  # band-limited forward input
        t.cuda.synchronize()
    finally:
        plan.close()
    del sp_f
    v64, s64 = _round_trip(t, cfg, SHT_PREC_FP64, sp_b, grid_f)
    vsp, ssp = _round_trip(t, cfg, SHT_PREC_SPLIT_INT8, sp_b, grid_f)
    eb = rel_comp_dev(t, vsp, v64, 8, comp_fastest=False)
    ef = rel_comp_dev(t, ssp, s64, 9, comp_fastest=True)
    _say(f"    backward max rel L2 {eb:.3e}  forward max rel L2 {ef:.3e}")
    del v64, s64, vsp, ssp, grid_f, sp_b
    t.cuda.empty_cache()
    assert eb <= 1e-11 and ef <= 1e-11


def test_cfg2_fp64_linearity(torch_cuda):
    t = torch_cuda
    cfg = CFG2
    _say("[gpu] config 2 FP64 backward linearity")
    a = spectra_dev(t, cfg, 8, 301)
    b = spectra_dev(t, cfg, 8, 302)
    plan = SphTransPlan(cfg, max_ncomp=9, precision=SHT_PREC_FP64)
    try:
        va = t.empty(cfg.n_rtp(8), dtype=t.float64, device="cuda")
        vb = t.empty_like(va)
        vab = t.empty_like(va)
        plan.sph_backward_transforms(2, 2, a, va)
        plan.sph_backward_transforms(2, 2, b, vb)
        # the plan runs on its own stream: finish before torch modifies a
        t.cuda.synchronize()
        a.add_(b)
        t.cuda.synchronize()
        plan.sph_backward_transforms(2, 2, a, vab)
        t.cuda.synchronize()
        va.add_(vb)
        e = rel_comp_dev(t, vab, va, 8, comp_fastest=False)
        _say(f"    max rel L2 of bwd(a+b) - bwd(a) - bwd(b): {e:.3e}")
        assert e <= 1e-13
    finally:
        plan.close()
        del a, b
        t.cuda.empty_cache()


@pytest.mark.parametrize("nth,nphi", [(1024, 2048), (1536, 3072)])
def test_truncation_L1023(torch_cuda, nth, nphi):
    """Config 4's truncation L=1023, on the smallest grid it allows (Ntheta =
    L+1 = 1024, Nphi = 2L+2 = 2048) and on config 4's own grid (Ntheta =
    1536, Nphi = 3072); both run the split-form FFT with N' = 1024 (a complex
    line across two waves).  2 radii keep the oracle short (the radii are
    independent):
    FP64 engine against the oracle (1e-12 per component), split engine
    against the FP64 engine (1e-11), whole backward (2 vector + 2 scalar) and
    forward (3 vector)."""
    t = torch_cuda
    cfg = ShellConfig(1023, 2, nth, nphi)
    _say(f"[gpu] L=1023 Ntheta={nth} Nphi={nphi}, 2 radii")
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    sp = spectra(cfg, 8, 401)
    spf = spectra(cfg, 9, 403)
    ref_v = orc.backward(2, 2, sp)
    grid = orc.backward(3, 0, spf)
    ref_s = orc.forward(3, 0, grid)
    _say("    oracle done")
    outs = {}
    for prec in (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8):
        plan = SphTransPlan(cfg, max_ncomp=9, precision=prec)
        _say(f"    plan {prec} ready: {plan.device_bytes() / 1e9:.2f} GB on the device")
        try:
            v = t.empty(cfg.n_rtp(8), dtype=t.float64, device="cuda")
            s = t.empty(cfg.n_rj(9), dtype=t.float64, device="cuda")
            plan.sph_backward_transforms(2, 2, t.from_numpy(sp).cuda(), v)
            plan.sph_forward_transforms(3, 0, t.from_numpy(grid).cuda(), s)
            t.cuda.synchronize()
            outs[prec] = (v.cpu().numpy(), s.cpu().numpy())
        finally:
            plan.close()
    eb = rel_comp_np(outs[SHT_PREC_FP64][0], ref_v, 8, comp_fastest=False)
    ef = rel_comp_np(outs[SHT_PREC_FP64][1], ref_s, 9, comp_fastest=True)
    sb = rel_comp_np(outs[SHT_PREC_SPLIT_INT8][0], outs[SHT_PREC_FP64][0], 8, comp_fastest=False)
    sf = rel_comp_np(outs[SHT_PREC_SPLIT_INT8][1], outs[SHT_PREC_FP64][1], 9, comp_fastest=True)
    _say(f"    FP64 vs oracle: backward {eb:.3e} forward {ef:.3e}; "
         f"split vs FP64: backward {sb:.3e} forward {sf:.3e}")
    assert eb <= 1e-12 and ef <= 1e-12
    assert sb <= 1e-11 and sf <= 1e-11
