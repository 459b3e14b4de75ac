"""Coriolis term in rlm space on the device (sht_sum_coriolis_rlm,
sht_copy_coriolis_terms_rlm) against the oracle, which is pinned against the
reference's own routines compiled from source and against the grid-space
Coriolis force.  Same tables, same operation order, contraction off in both:
bit-exact, with the identity receive table the engine uses, with a permuted
table holding dummy records, and with a rotatable inner core; then a device
MHD step with Coriolis: backward of the velocity and vorticity, the rlm sums
from the same send buffer, the forward of grid forces and the copy into its
output, against the oracle chain."""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import ShellConfig, SphTransPlan, synthetic_spectra

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("case", ["identity", "table", "icb"])
def test_sum_coriolis_bitexact_vs_oracle(torch_cuda, case):
    t = torch_cuda
    cfg = ShellConfig(63, 17, 96, 192)
    n = cfg.jmax * cfg.nr
    ncomp, iv, iw, coef = 9, 3, 0, 2.25
    kr_icb = 1 if case == "icb" else 0
    _say(f"[gpu] sum_coriolis_rlm {case}: L={cfg.L} nr={cfg.nr} nodes={n}")
    rng = np.random.default_rng(31)
    sp = rng.standard_normal(ncomp * n)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    irev = None
    wr = sp
    if case == "table":
        # records stored in permuted order; some nodes point at the dummy record
        perm = rng.permutation(n)
        wr = np.empty_like(sp)
        wr.reshape(n, ncomp)[perm] = sp.reshape(n, ncomp)
        irev = (perm + 1).astype(np.int32)
        irev[rng.choice(n, 40, replace=False)] = n + 1
    want = orc.sum_coriolis_rlm(ncomp, iv, iw, coef, wr, irev=irev, kr_icb=kr_icb)
    plan = SphTransPlan(cfg, max_ncomp=9)
    cor = plan.coriolis(kr_icb)
    try:
        d_wr = t.from_numpy(np.ascontiguousarray(wr)).cuda()
        d_cor = t.full((2 * n,), np.nan, dtype=t.float64, device="cuda")
        cor.sum_coriolis_rlm(ncomp, iv, iw, coef, d_wr, d_cor,
                             irev=t.from_numpy(irev).cuda() if irev is not None else None)
        t.cuda.synchronize()
        got = d_cor.cpu().numpy().reshape(2, n)
        bad = np.argwhere(got != want)
        assert len(bad) == 0, (f"{len(bad)} of {got.size} differ, max {np.abs(got - want).max()}, "
                               f"first {bad[:4].tolist()}")
    finally:
        cor.close()
        plan.close()


def test_copy_coriolis_bitexact_vs_oracle(torch_cuda):
    t = torch_cuda
    cfg = ShellConfig(31, 9, 48, 96)
    n = cfg.jmax * cfg.nr
    ncomp, irc = 12, 6
    _say(f"[gpu] copy_coriolis_terms_rlm: nodes={n}")
    rng = np.random.default_rng(5)
    dcor = rng.standard_normal(2 * n)
    ws0 = rng.standard_normal(ncomp * n)
    item = (rng.permutation(n) + 1).astype(np.int32)
    plan = SphTransPlan(cfg, max_ncomp=9)
    cor = plan.coriolis()
    try:
        for tab in (None, item):
            want = O.rj_to_send([(0, irc, 1), (1, irc + 2, 1)], 2, n, dcor, ncomp, ws0,
                                inod_export=tab)
            d_ws = t.from_numpy(ws0.copy()).cuda()
            cor.copy_coriolis_terms_rlm(ncomp, irc, t.from_numpy(dcor).cuda(), d_ws,
                                        item=t.from_numpy(tab).cuda() if tab is not None else None)
            t.cuda.synchronize()
            assert np.array_equal(d_ws.cpu().numpy(), want)
    finally:
        cor.close()
        plan.close()


def test_mhd_step_with_coriolis_vs_oracle(torch_cuda):
    """Backward (velocity, vorticity), rlm Coriolis sums from the same send
    buffer (single rank: the rj and rlm orders coincide), forward of three
    grid forces, the curl of Coriolis copied into the third force's slots."""
    t = torch_cuda
    cfg = ShellConfig(47, 9, 72, 144)
    nv, ns = 2, 0
    ncomp_b = 3 * nv + ns
    n = cfg.jmax * cfg.nr
    _say(f"[gpu] MHD step with rlm Coriolis: L={cfg.L} nr={cfg.nr}")
    sp = synthetic_spectra(cfg, nv, ns, seed=9)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    plan = SphTransPlan(cfg, max_ncomp=9)
    cor = plan.coriolis()
    try:
        d_sp = t.from_numpy(sp).cuda()
        v = t.empty(cfg.n_rtp(ncomp_b), dtype=t.float64, device="cuda")
        plan.sph_backward_transforms(nv, ns, d_sp, v)
        d_cor = t.empty(2 * n, dtype=t.float64, device="cuda")
        cor.sum_coriolis_rlm(ncomp_b, 0, 3, 1.5, d_sp, d_cor)
        # forward of three vector forces: the backward's two vectors and the
        # first again (the slot the rlm Coriolis curl overwrites)
        f = t.cat([v, v[: cfg.n_rtp(3)]])
        out = t.empty(cfg.n_rj(9), dtype=t.float64, device="cuda")
        plan.sph_forward_transforms(3, 0, f, out)
        cor.copy_coriolis_terms_rlm(9, 6, d_cor, out)
        t.cuda.synchronize()
        ref_v = orc.backward(nv, ns, sp)
        ref_c = orc.sum_coriolis_rlm(ncomp_b, 0, 3, 1.5, sp)
        ref_f = np.concatenate([ref_v, ref_v[: cfg.n_rtp(3)]])
        ref_o = orc.forward(3, 0, ref_f).reshape(n, 9)
        ref_o[:, 6] = ref_c[0]
        ref_o[:, 8] = ref_c[1]
        got = out.cpu().numpy().reshape(n, 9)
        assert np.array_equal(d_cor.cpu().numpy(), ref_c.ravel())
        assert np.array_equal(got[:, 6], ref_o[:, 6]) and np.array_equal(got[:, 8], ref_o[:, 8])
        rel = np.linalg.norm(got - ref_o) / np.linalg.norm(ref_o)
        assert rel < 1e-12, rel
    finally:
        cor.close()
        plan.close()
