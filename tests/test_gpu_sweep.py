"""Seeded sweep of small shells through the whole transform, both engines.

Shapes are drawn (fixed seed) so that they cover what the named parity cases
do not: odd and even Ntheta, Ntheta above the minimum, Nphi lengths whose
factors send the FFT to the split form, the mixed-radix form, direct-sum
prime passes and Bluestein passes, and a few radii each.  Bars: the FP64
engine within relative L2 1e-12 per component of the oracle (backward 2
vector + 2 scalar, forward 3 vector); the split engine (6 slices) within
1e-11 of the oracle."""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellConfig, SphTransPlan

pytestmark = pytest.mark.gpu


def _draw(n=8, seed=20261202):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        L = int(rng.integers(6, 90))
        nth = int(L + 1 + rng.integers(0, L // 2 + 2))
        # Nphi even and above 2L+1; the extra factor picks the FFT form
        base = 2 * L + 2
        extra = int(rng.choice([0, 2, 4, 6, 10, 22, 34, 46, 106]))
        nphi = base + extra
        nphi += nphi % 2
        out.append(ShellConfig(L, int(rng.integers(2, 5)), nth, nphi))
    return out


SHELLS = _draw()


def _say(*a):
    print(*a, flush=True)


def _rel_comp(a, b, ncomp, comp_fastest):
    a = np.asarray(a).reshape(-1, ncomp) if comp_fastest else np.asarray(a).reshape(ncomp, -1).T
    b = np.asarray(b).reshape(-1, ncomp) if comp_fastest else np.asarray(b).reshape(ncomp, -1).T
    out = 0.0
    for c in range(ncomp):
        nb = np.linalg.norm(b[:, c])
        out = max(out, np.linalg.norm(a[:, c] - b[:, c]) / (nb if nb > 0 else 1.0))
    return out


@pytest.mark.parametrize("cfg", SHELLS, ids=[f"L{c.L}_nr{c.nr}_nth{c.nth}_nphi{c.nphi}"
                                             for c in SHELLS])
def test_whole_transform_sweep(torch_cuda, cfg):
    _say(f"[sweep] L={cfg.L} nr={cfg.nr} nth={cfg.nth} nphi={cfg.nphi}")
    from kemorin_mhd_amd import synthetic_spectra
    _check(torch_cuda, cfg, synthetic_spectra(cfg, 2, 2, seed=cfg.L + 17),
           synthetic_spectra(cfg, 3, 0, seed=cfg.L + 29))


# The smallest shells the plan accepts: L from 1, one to three radii, the
# minimum Ntheta = max(2, L+1) (and one above it, odd) and Nphi = 2L+2 or a
# little more, whose factors 3, 5 and 7 go to the generic FFT's prime passes.
# One- and two-radius shells are all boundary, where the synthetic spectra's
# radial envelope vanishes, so these take plain seeded Gaussian spectra.
TINY = [ShellConfig(1, 1, 2, 4), ShellConfig(2, 1, 3, 6), ShellConfig(3, 2, 4, 8),
        ShellConfig(4, 1, 5, 10), ShellConfig(5, 3, 9, 14), ShellConfig(1, 2, 3, 6)]


@pytest.mark.parametrize("cfg", TINY, ids=[f"L{c.L}_nr{c.nr}_nth{c.nth}_nphi{c.nphi}"
                                           for c in TINY])
def test_tiny_shells(torch_cuda, cfg):
    _say(f"[tiny] L={cfg.L} nr={cfg.nr} nth={cfg.nth} nphi={cfg.nphi}")
    rng = np.random.default_rng(1000 + 10 * cfg.L + cfg.nr)
    _check(torch_cuda, cfg, rng.standard_normal(cfg.n_rj(8)), rng.standard_normal(cfg.n_rj(9)))


def _check(t, cfg, sp, spf):
    """Backward 2 vector + 2 scalar of sp, and forward 3 vector of the
    oracle's backward of spf, on both engines against the oracle."""
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    ref_v = orc.backward(2, 2, sp)
    grid = orc.backward(3, 0, spf)
    ref_s = orc.forward(3, 0, grid)
    for prec, tol in ((SHT_PREC_FP64, 1e-12), (SHT_PREC_SPLIT_INT8, 1e-11)):
        plan = SphTransPlan(cfg, max_ncomp=9, precision=prec)
        try:
            v = t.empty(cfg.n_rtp(8), dtype=t.float64, device="cuda")
            s = t.empty(cfg.n_rj(9), dtype=t.float64, device="cuda")
            plan.sph_backward_transforms(2, 2, t.from_numpy(sp).cuda(), v)
            plan.sph_forward_transforms(3, 0, t.from_numpy(grid).cuda(), s)
            t.cuda.synchronize()
            eb = _rel_comp(v.cpu().numpy(), ref_v, 8, comp_fastest=False)
            ef = _rel_comp(s.cpu().numpy(), ref_s, 9, comp_fastest=True)
        finally:
            plan.close()
        _say(f"    precision {prec}: backward {eb:.2e} forward {ef:.2e}")
        assert eb <= tol and ef <= tol, (prec, eb, ef)
