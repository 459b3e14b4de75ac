"""Plans created from the reference's spherical grid / mode files
(sht_plan_create_from_files) on the committed one-rank set (written by the
reference's own writers): outputs bit for bit equal to a plan described by
hand with the files' shape and radii (both engines), within the oracle's
bars, and the rj send / receive copies driven by the files' rj tables equal
to the single-rank identity."""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, SphFiles, SphTransPlan, synthetic_spectra
from tests.golden.make_sph_files import paths as fixture_paths

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("prec,tol", [(SHT_PREC_FP64, 1e-12), (SHT_PREC_SPLIT_INT8, 1e-11)],
                         ids=["fp64", "split"])
def test_plan_from_files(torch_cuda, prec, tol):
    t = torch_cuda
    files = SphFiles(*fixture_paths())
    cfg, r, _ = files.shape()
    print(f"[gpu] plan from files L={cfg.L} nr={cfg.nr} precision={prec}", flush=True)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, r, full_tables=False)
    sp = synthetic_spectra(cfg, 2, 2, seed=301)
    grid = orc.backward(3, 0, synthetic_spectra(cfg, 3, 0, seed=303))
    pf = files.plan(max_ncomp=9, precision=prec)
    pd = SphTransPlan(cfg, max_ncomp=9, precision=prec, radius=r)
    try:
        out_f = _run(t, pf, cfg, sp, grid)
        out_d = _run(t, pd, cfg, sp, grid)
    finally:
        pf.close()
        pd.close()
        files.close()
    assert np.array_equal(out_f[0], out_d[0]) and np.array_equal(out_f[1], out_d[1])
    assert _rel(out_f[0], orc.backward(2, 2, sp), 8, False) <= tol
    assert _rel(out_f[1], orc.forward(3, 0, grid), 9, True) <= tol


def test_rj_copies_with_file_tables(torch_cuda):
    t = torch_cuda
    files = SphFiles(*fixture_paths())
    cfg, r, nnod_rj = files.shape()
    exp = t.from_numpy(files.table("rj_import")).cuda()
    rev = t.from_numpy(files.table("rj_import_rev")).cuda()
    plan = files.plan(max_ncomp=8)
    try:
        ntot = 8
        d_fld = t.randn(nnod_rj * ntot, dtype=t.float64, device="cuda")
        fields = [(0, 0, 3), (3, 3, 3), (6, 6, 1), (7, 7, 1)]
        ws_a = t.full((cfg.n_rj(8),), np.nan, dtype=t.float64, device="cuda")
        ws_b = ws_a.clone()
        plan.copy_spectr_to_send(fields, ntot, d_fld, 8, ws_a, inod_export=exp)
        plan.copy_spectr_to_send(fields, ntot, d_fld, 8, ws_b)
        back_a = t.full_like(d_fld, np.nan)
        back_b = t.full_like(d_fld, np.nan)
        plan.copy_spectr_from_recv(fields, 8, ws_a, ntot, back_a, irev=rev)
        plan.copy_spectr_from_recv(fields, 8, ws_b, ntot, back_b)
        t.cuda.synchronize()
        assert t.equal(ws_a, ws_b)
        assert t.equal(back_a, back_b) and t.equal(back_a, d_fld)
    finally:
        plan.close()
        files.close()
