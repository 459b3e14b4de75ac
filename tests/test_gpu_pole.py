"""Backward transform at the poles and the centre on the GPU, against the oracle.

Bar: bit-exact with the oracle.  The kernel sums over l in the reference's
order with contraction off, on pole tables that the library builds with the
reference's recurrence on the host, so the operation sequence is the oracle's
own; the FP64 bar of the whole transform (relative L2 1e-12 per component)
is asserted as well, so a failure of the stronger claim names itself.
Covers the small shell (L=31, Nr=33) and the L=511, Nr=257 shell, odd field
sets, with and without a centre, the entries the transform leaves alone,
both precisions of plan (the pole path is FP64 in both), and the combined
entry point against the two calls it stands for.
"""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellConfig, SphTransPlan,
                             gaussian, synthetic_spectra)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _rel(a, b, ncomp):
    a = a.reshape(ncomp, -1)
    b = b.reshape(ncomp, -1)
    out = 0.0
    for c in range(ncomp):
        nb = np.linalg.norm(b[c])
        out = max(out, np.linalg.norm(a[c] - b[c]) / (nb if nb > 0 else 1.0))
    return out


CASES = [
    # (L, Nr, Ntheta, Nphi, nvector, nscalar, center, precision)
    (31, 33, 48, 96, 2, 2, False, SHT_PREC_FP64),
    (31, 33, 48, 96, 2, 2, True, SHT_PREC_FP64),
    (5, 3, 8, 12, 1, 0, True, SHT_PREC_FP64),
    (1, 2, 2, 4, 0, 1, True, SHT_PREC_FP64),
    (47, 9, 72, 144, 3, 0, True, SHT_PREC_SPLIT_INT8),
    (511, 257, 768, 1536, 1, 1, True, SHT_PREC_FP64),
]


@pytest.mark.parametrize("case", CASES, ids=[f"L{c[0]}_nr{c[1]}_v{c[4]}s{c[5]}_c{int(c[6])}_p{c[7]}"
                                             for c in CASES])
def test_pole_backward_matches_oracle(torch_cuda, case):
    t = torch_cuda
    L, nr, nth, nphi, nv, ns, center, prec = case
    ncomp = 3 * nv + ns
    _say(f"[pole] L={L} nr={nr} nv={nv} ns={ns} center={center} precision={prec}")
    cfg = ShellConfig(L, nr, nth, nphi)
    n = cfg.n_rj(ncomp)
    sp = (gaussian(400 + L, n) if n < 10**7
          else np.random.default_rng(400 + L).standard_normal(n))
    init = np.full(cfg.n_pole(ncomp), 7.0)
    orc = O.OraclePlan(L, nr, nth, nphi, cfg.radius(), full_tables=False)
    ref = orc.pole_backward(nv, ns, sp, center=center, v_pole=init)
    plan = SphTransPlan(cfg, max_ncomp=ncomp, precision=prec)
    try:
        vp = t.tensor(init, device="cuda")
        plan.pole_b_transform(nv, ns, t.tensor(sp, device="cuda"), vp, with_center=center)
        t.cuda.synchronize()
        got = vp.cpu().numpy()
    finally:
        plan.close()
    if L == 511:
        # launch time at the largest shell (HIP events on the stream the plan
        # follows, torch's current stream), reported only
        plan = SphTransPlan(cfg, max_ncomp=ncomp, precision=prec)
        try:
            spd = t.tensor(sp, device="cuda")
            vp2 = t.tensor(init, device="cuda")
            plan.pole_b_transform(nv, ns, spd, vp2, with_center=center)
            e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(20):
                plan.pole_b_transform(nv, ns, spd, vp2, with_center=center)
            e1.record()
            e1.synchronize()
            _say(f"    pole transform: {e0.elapsed_time(e1) / 20 * 1e3:.1f} us per launch")
        finally:
            plan.close()
    err = _rel(got, ref, ncomp)
    diff = int(np.count_nonzero(got != ref))
    _say(f"    rel L2 {err:.2e}, entries differing from the oracle: {diff}")
    assert err <= 1e-12
    assert diff == 0


def test_backward_with_poles_is_both_calls(torch_cuda):
    t = torch_cuda
    cfg = ShellConfig(63, 9, 96, 192)
    nv, ns = 2, 2
    ncomp = 3 * nv + ns
    sp = t.tensor(synthetic_spectra(cfg, nv, ns, seed=3), device="cuda")
    plan = SphTransPlan(cfg, max_ncomp=ncomp)
    try:
        v1 = t.zeros(cfg.n_rtp(ncomp), dtype=t.float64, device="cuda")
        p1 = t.full((cfg.n_pole(ncomp),), 7.0, dtype=t.float64, device="cuda")
        v2 = t.zeros_like(v1)
        p2 = t.full_like(p1, 7.0)
        plan.sph_backward_transforms(nv, ns, sp, v1)
        plan.pole_b_transform(nv, ns, sp, p1, with_center=True)
        plan.sph_b_trans_w_poles(nv, ns, sp, v2, p2, with_center=True)
        t.cuda.synchronize()
        assert t.equal(v1, v2)
        assert t.equal(p1, p2)
    finally:
        plan.close()


def test_pole_rejects_bad_arguments(torch_cuda):
    t = torch_cuda
    cfg = ShellConfig(15, 3, 24, 48)
    plan = SphTransPlan(cfg, max_ncomp=4)
    try:
        sp = t.zeros(cfg.n_rj(4), dtype=t.float64, device="cuda")
        vp = t.zeros(cfg.n_pole(4), dtype=t.float64, device="cuda")
        with pytest.raises(Exception, match="ncomp"):
            plan.pole_b_transform(2, 0, sp, vp)
    finally:
        plan.close()
