"""The batched-shell workload against the oracle: the 64 seeded uneven shells
(L odd in 63..383, Nr in 33..129, Ntheta = 3(L+1)/2, Nphi = 2 Ntheta; one
vector + one scalar each way) captured into one HIP graph and replayed.

The draw's longitudes cover every FFT form the engine has: the split form
(Nphi = 3 * 2^k), mixed-radix 2/3/5/7/11/13 lengths, direct-sum passes and
the Bluestein pass for primes from 17 to 181 (Nphi up to 1134).  Every
shell's backward and forward output is compared with the oracle on three of
its radii (first, middle, last; radii are independent).

Bars: FP64 engine within relative L2 1e-12 per component of the oracle,
split engine (6 INT8 slices) within 1e-11; a second replay reproduces the
first bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellBatch, SphTransPlan,
                             batched_shell_draw, rlm_l_m)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _rel(a, b, ncomp, comp_fastest):
    a = np.asarray(a).reshape(-1, ncomp) if comp_fastest else np.asarray(a).reshape(ncomp, -1).T
    b = np.asarray(b).reshape(-1, ncomp) if comp_fastest else np.asarray(b).reshape(ncomp, -1).T
    return max(float(np.linalg.norm(a[:, c] - b[:, c]) / np.linalg.norm(b[:, c]))
               for c in range(ncomp))


@pytest.mark.parametrize("precision,tol", [(SHT_PREC_FP64, 1e-12), (SHT_PREC_SPLIT_INT8, 1e-11)],
                         ids=["fp64", "split"])
def test_shells64_graph_vs_oracle(torch_cuda, precision, tol):
    t = torch_cuda
    shells = batched_shell_draw(64)
    _say(f"[gpu] 64-shell batch, precision {precision}: plans and inputs")
    gen = t.Generator(device="cuda")
    gen.manual_seed(64)
    plans, bufs = [], []
    try:
        for c in shells:
            plans.append(SphTransPlan(c, max_ncomp=4, precision=precision))
            l, _ = rlm_l_m(c.L)
            ws = t.randn((c.nr, c.jmax, 4), generator=gen, dtype=t.float64, device="cuda")
            ws *= t.tensor((l + 1.0) ** -1.5, device="cuda")[None, :, None]
            vin = t.randn(c.n_rtp(4), generator=gen, dtype=t.float64, device="cuda")
            vout = t.empty(c.n_rtp(4), dtype=t.float64, device="cuda")
            wr = t.empty(c.n_rj(4), dtype=t.float64, device="cuda")
            bufs.append((ws.reshape(-1), vout, vin, wr))
        t.cuda.synchronize()
        batch = ShellBatch([(p, (1, 1), ws, vout, (1, 1), vin, wr)
                            for p, (ws, vout, vin, wr) in zip(plans, bufs)])
        try:
            batch.launch()
            batch.synchronize()
            first = [(vout.clone(), wr.clone()) for (_, vout, _, wr) in bufs]
            for (_, vout, _, wr) in bufs:
                vout.fill_(np.nan)
                wr.fill_(np.nan)
            t.cuda.synchronize()
            batch.launch()
            batch.synchronize()
            for i, ((_, vout, _, wr), (v1, w1)) in enumerate(zip(bufs, first)):
                assert t.equal(vout, v1) and t.equal(wr, w1), f"shell {i}: replay differs"
            del first
        finally:
            batch.close()
        _say("    two replays bit-exact; oracle on 3 radii per shell")
        worst = (0.0, -1)
        for i, (c, (ws, vout, vin, wr)) in enumerate(zip(shells, bufs)):
            ks = np.array(sorted({0, c.nr // 2, c.nr - 1}))
            kd = t.from_numpy(ks).cuda()
            orc = O.OraclePlan(c.L, len(ks), c.nth, c.nphi, c.radius()[ks], full_tables=False)
            ws_s = ws.view(c.nr, c.jmax, 4)[kd].contiguous().cpu().numpy().ravel()
            vin_s = vin.view(4, c.nphi, c.nth, c.nr)[:, :, :, kd].contiguous().cpu().numpy()
            got_v = vout.view(4, c.nphi, c.nth, c.nr)[:, :, :, kd].contiguous().cpu().numpy()
            got_s = wr.view(c.nr, c.jmax, 4)[kd].contiguous().cpu().numpy()
            eb = _rel(got_v, orc.backward(1, 1, ws_s), 4, comp_fastest=False)
            ef = _rel(got_s, orc.forward(1, 1, vin_s.ravel()), 4, comp_fastest=True)
            e = max(eb, ef)
            if e > worst[0]:
                worst = (e, i)
            assert e <= tol, (f"shell {i} L={c.L} Nr={c.nr} Nphi={c.nphi}: backward {eb:.2e} "
                              f"forward {ef:.2e}")
        _say(f"    worst rel L2 {worst[0]:.2e} (shell {worst[1]}, "
             f"L={shells[worst[1]].L}, Nphi={shells[worst[1]].nphi})")
    finally:
        for p in plans:
            p.close()
        del bufs
        t.cuda.empty_cache()
