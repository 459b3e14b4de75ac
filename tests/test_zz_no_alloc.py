"""sht_backward_noalloc: the backward transform with the library's fixed
device workspace instead of the plan's, through the C ABI.

Bar: bit-exact against sht_backward on the same plan and input (the same
kernels run; only the workspace address differs), for both engines, on two
calls in a row where the second is smaller than the first (the fixed array
still holds the first call's contents then), and once more after a call of
the first size on a third input.  A call whose workspace exceeds
SHT_NOALLOC_SCRATCH_BYTES is rejected with SHT_ERR_ARG and a message naming
the size.
"""
import numpy as np
import pytest

from kemorin_mhd_amd import (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellConfig, ShtError,
                             SphTransPlan, synthetic_spectra)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _pair(t, plan, cfg, nv, ns, seed):
    ncomp = 3 * nv + ns
    sp = t.from_numpy(synthetic_spectra(cfg, nv, ns, seed=seed)).cuda()
    ref = t.full((cfg.n_rtp(ncomp),), np.nan, dtype=t.float64, device="cuda")
    got = t.full((cfg.n_rtp(ncomp),), np.nan, dtype=t.float64, device="cuda")
    plan.sph_backward_transforms(nv, ns, sp, ref)
    plan.sph_backward_transforms_noalloc(nv, ns, sp, got)
    t.cuda.synchronize()
    return ref, got


@pytest.mark.parametrize("precision", [SHT_PREC_FP64, SHT_PREC_SPLIT_INT8], ids=["fp64", "split"])
def test_noalloc_matches_backward(torch_cuda, precision):
    t = torch_cuda
    big = ShellConfig(31, 33, 48, 96)
    small = ShellConfig(15, 5, 25, 40)
    _say(f"[gpu] sht_backward_noalloc precision {precision}: L=31 Nr=33 then L=15 Nr=5")
    pb = SphTransPlan(big, max_ncomp=9, precision=precision)
    ps = SphTransPlan(small, max_ncomp=9, precision=precision)
    try:
        # first call: the larger input (2 vectors + 2 scalars)
        ref, got = _pair(t, pb, big, 2, 2, seed=31)
        assert t.equal(got, ref), "first (larger) call differs from sht_backward"
        # second call: a smaller plan and fewer components, over the leftovers
        ref, got = _pair(t, ps, small, 1, 1, seed=32)
        assert t.equal(got, ref), "second (smaller) call differs from sht_backward"
        # and the larger plan again, with a smaller field set
        ref, got = _pair(t, pb, big, 0, 3, seed=33)
        assert t.equal(got, ref), "third call differs from sht_backward"
    finally:
        pb.close()
        ps.close()


def test_noalloc_rejects_oversized_calls(torch_cuda):
    t = torch_cuda
    cfg = ShellConfig(127, 65, 192, 384)  # 8 x 8 x 255 x 65 x 192 bytes > 128 MiB
    _say("[gpu] sht_backward_noalloc size limit")
    plan = SphTransPlan(cfg, max_ncomp=8)
    try:
        sp = t.zeros(cfg.n_rj(8), dtype=t.float64, device="cuda")
        v = t.zeros(cfg.n_rtp(8), dtype=t.float64, device="cuda")
        with pytest.raises(ShtError, match="bytes of workspace") as ei:
            plan.sph_backward_transforms_noalloc(2, 2, sp, v)
        assert ei.value.status == 1  # SHT_ERR_ARG
        # one component fits
        v1 = t.zeros(cfg.n_rtp(1), dtype=t.float64, device="cuda")
        ref = t.zeros(cfg.n_rtp(1), dtype=t.float64, device="cuda")
        s1 = t.from_numpy(synthetic_spectra(cfg, 0, 1, seed=5)).cuda()
        plan.sph_backward_transforms_noalloc(0, 1, s1, v1)
        plan.sph_backward_transforms(0, 1, s1, ref)
        t.cuda.synchronize()
        assert t.equal(v1, ref)
    finally:
        plan.close()
