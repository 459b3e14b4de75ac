"""P_l^m generated per call, one tile of orders at a time
(SHT_FLAG_P_ON_THE_FLY), against the stored tables and the oracle.

The tile generator runs the plan generator's recurrence in the same
operation order and the Legendre kernels are the same, so both engines'
outputs equal those of a plan with stored tables bit for bit (every stage,
both MFMA forms of the split forward products, and under graph replay).  On
config 4's own grid (L=1023, Ntheta=1536, Nphi=3072) at 2 radii the plan
with tables per call is held to the oracle: FP64 1e-12 and split 1e-11 per
component, and its device footprint is compared with the stored plan's."""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellBatch, ShellConfig,
                             SphTransPlan, synthetic_spectra)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _rel(a, b, ncomp, comp_fastest):
    a = np.asarray(a).reshape(-1, ncomp) if comp_fastest else np.asarray(a).reshape(ncomp, -1).T
    b = np.asarray(b).reshape(-1, ncomp) if comp_fastest else np.asarray(b).reshape(ncomp, -1).T
    return max(np.linalg.norm(a[:, c] - b[:, c]) / np.linalg.norm(b[:, c]) for c in range(ncomp))


def _round_trip(t, plan, cfg, sp, grid, variant=None):
    if variant:
        plan.set_variant("split_fwd_mfma", variant)
    v = t.full((cfg.n_rtp(8),), np.nan, dtype=t.float64, device="cuda")
    s = t.full((cfg.n_rj(9),), np.nan, dtype=t.float64, device="cuda")
    plan.sph_backward_transforms(2, 2, t.from_numpy(sp).cuda(), v)
    plan.sph_forward_transforms(3, 0, t.from_numpy(grid).cuda(), s)
    t.cuda.synchronize()
    return v.cpu().numpy(), s.cpu().numpy()


@pytest.mark.parametrize("prec", [SHT_PREC_FP64, SHT_PREC_SPLIT_INT8], ids=["fp64", "split"])
@pytest.mark.parametrize("cfg", [ShellConfig(127, 9, 192, 384), ShellConfig(63, 5, 97, 196)],
                         ids=["L127", "L63_odd"])
def test_tables_per_call_bitexact_vs_stored(torch_cuda, prec, cfg):
    t = torch_cuda
    _say(f"[gpu] P per call vs stored: L={cfg.L} nth={cfg.nth} precision={prec}")
    sp = synthetic_spectra(cfg, 2, 2, seed=61)
    grid = np.random.default_rng(63).standard_normal(cfg.n_rtp(9))
    outs = {}
    for otf in (False, True):
        plan = SphTransPlan(cfg, max_ncomp=9, precision=prec, p_on_the_fly=otf)
        try:
            outs[otf] = _round_trip(t, plan, cfg, sp, grid)
            if prec == SHT_PREC_SPLIT_INT8:
                outs[(otf, 32)] = _round_trip(t, plan, cfg, sp, grid, variant=32)
        finally:
            plan.close()
    for k in range(2):
        assert np.array_equal(outs[True][k], outs[False][k])
    if prec == SHT_PREC_SPLIT_INT8:
        for k in range(2):
            assert np.array_equal(outs[(True, 32)][k], outs[False][k])
            assert np.array_equal(outs[(False, 32)][k], outs[False][k])


def test_tables_per_call_under_graph_replay(torch_cuda):
    """A shell batch of plans with tables per call: the tile generation is
    captured with the transforms, and every replay equals the eager calls."""
    torch = torch_cuda
    shells = [ShellConfig(31, 3, 48, 96), ShellConfig(63, 2, 96, 192)]
    plans, bufs, ref = [], [], []
    for i, c in enumerate(shells):
        p = SphTransPlan(c, max_ncomp=4, precision=SHT_PREC_SPLIT_INT8 if i else SHT_PREC_FP64,
                         p_on_the_fly=True)
        ws = torch.tensor(synthetic_spectra(c, 1, 1, seed=70 + i), device="cuda")
        vin = torch.tensor(np.random.default_rng(80 + i).standard_normal(c.n_rtp(4)), device="cuda")
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


def test_cfg4_grid_tables_per_call_vs_oracle(torch_cuda):
    t = torch_cuda
    cfg = ShellConfig(1023, 2, 1536, 3072)
    _say("[gpu] config 4 grid, P per call, 2 radii")
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    sp = synthetic_spectra(cfg, 2, 2, seed=401)
    grid = orc.backward(3, 0, synthetic_spectra(cfg, 3, 0, seed=403))
    ref_v = orc.backward(2, 2, sp)
    ref_s = orc.forward(3, 0, grid)
    _say("    oracle done")
    for prec, tol in ((SHT_PREC_FP64, 1e-12), (SHT_PREC_SPLIT_INT8, 1e-11)):
        plan = SphTransPlan(cfg, max_ncomp=9, precision=prec, p_on_the_fly=True)
        try:
            nbytes = plan.device_bytes()
            v, s = _round_trip(t, plan, cfg, sp, grid)
        finally:
            plan.close()
        eb = _rel(v, ref_v, 8, comp_fastest=False)
        ef = _rel(s, ref_s, 9, comp_fastest=True)
        _say(f"    precision {prec}: plan {nbytes / 1e9:.2f} GB; backward {eb:.2e} forward {ef:.2e}")
        assert eb <= tol and ef <= tol
        # the stored tables alone are 12.9 GB (FP64) and 12.1 GB more (slices)
        assert nbytes < (3e9 if prec == SHT_PREC_FP64 else 8e9)
