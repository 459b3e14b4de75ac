"""Config 4 at full size (L=1023, Nr=257, Ntheta=1536, Nphi=3072) against the
oracle, both engines, P generated per call, with the benchmark's buffer
arrangement (the backward writes its 8 grid components into the first 8 of
the forward's 9).  Every radius is transformed on the device, so the offsets
of the full buffers (9.7 GB per grid component) are exercised; radii are
independent, so the oracle takes two of them (the boundaries).  Bars: FP64
1e-12 and split 1e-11 per component, as on config 4's grid at 2 radii.  The
cases need about 185 GB (FP64) and 235 GB (split) of device memory and are
skipped where less is free."""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellConfig, ShtError,
                             SphTransPlan, rlm_l_m)

pytestmark = pytest.mark.gpu

CFG4 = ShellConfig(1023, 257, 1536, 3072)
KS = np.array([0, 256])
# plan with per-call tables + spectra in / out + 9 grid components
NEED = {SHT_PREC_FP64: 190e9, SHT_PREC_SPLIT_INT8: 240e9}


def _rel(a, b, ncomp, comp_fastest):
    a = np.asarray(a).reshape(-1, ncomp) if comp_fastest else np.asarray(a).reshape(ncomp, -1).T
    b = np.asarray(b).reshape(-1, ncomp) if comp_fastest else np.asarray(b).reshape(ncomp, -1).T
    return max(np.linalg.norm(a[:, c] - b[:, c]) / np.linalg.norm(b[:, c]) for c in range(ncomp))


@pytest.mark.parametrize("prec,tol", [(SHT_PREC_FP64, 1e-12), (SHT_PREC_SPLIT_INT8, 1e-11)],
                         ids=["fp64", "split"])
def test_config4_full_size_vs_oracle(torch_cuda, prec, tol):
    t = torch_cuda
    cfg = CFG4
    t.cuda.empty_cache()
    free, _ = t.cuda.mem_get_info()
    if free < NEED[prec]:
        pytest.skip(f"config 4 at full size needs about {NEED[prec] / 1e9:.0f} GB free, "
                    f"{free / 1e9:.0f} GB is")
    print(f"[gpu] config 4 full size, precision {prec}, P per call", flush=True)
    plan = None
    try:
        try:
            plan = SphTransPlan(cfg, max_ncomp=9, precision=prec, p_on_the_fly=True)
            gen = t.Generator(device="cuda")
            gen.manual_seed(4401)
            l, _ = rlm_l_m(cfg.L)
            r = cfg.radius()
            prof = 0.5 + np.sin(np.pi * (r - cfg.ri) / (cfg.ro - cfg.ri))
            sp = t.randn((cfg.nr, cfg.jmax, 8), generator=gen, dtype=t.float64, device="cuda")
            sp *= t.tensor((l + 1.0) ** -1.5, device="cuda")[None, :, None]
            sp *= t.tensor(prof, device="cuda")[:, None, None]
            sp = sp.reshape(-1)
            grid = t.empty(cfg.n_rtp(9), dtype=t.float64, device="cuda")
            out = t.empty(cfg.n_rj(9), dtype=t.float64, device="cuda")
        except (ShtError, t.cuda.OutOfMemoryError) as e:
            pytest.skip(f"config 4 at full size does not fit next to this process's memory: {e}")
        nnod = cfg.nr * cfg.nth * cfg.nphi
        plan.sph_backward_transforms(2, 2, sp, grid[: 8 * nnod])
        grid[8 * nnod:] = grid[3 * nnod: 4 * nnod]  # the ninth component: a band-limited field
        plan.sph_forward_transforms(3, 0, grid, out)
        t.cuda.synchronize()
        kd = t.from_numpy(KS).cuda()
        sp_s = sp.view(cfg.nr, cfg.jmax * 8)[kd].contiguous().cpu().numpy().ravel()
        grid_s = grid.view(9, cfg.nphi, cfg.nth, cfg.nr)[:, :, :, kd].contiguous().cpu().numpy()
        out_s = out.view(cfg.nr, cfg.jmax * 9)[kd].contiguous().cpu().numpy().ravel()
        del sp, grid, out
    finally:
        if plan is not None:
            plan.close()
        t.cuda.empty_cache()
    orc = O.OraclePlan(cfg.L, len(KS), cfg.nth, cfg.nphi, cfg.radius()[KS], full_tables=False)
    ref_v = orc.backward(2, 2, sp_s)
    ref_s = orc.forward(3, 0, grid_s.ravel())
    eb = _rel(grid_s[:8].ravel(), ref_v, 8, comp_fastest=False)
    ef = _rel(out_s, ref_s, 9, comp_fastest=True)
    print(f"    backward {eb:.2e} forward {ef:.2e}", flush=True)
    assert eb <= tol and ef <= tol
