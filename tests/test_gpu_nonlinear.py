"""Grid-space products of the MHD step on the GPU, against the oracle.

Bar: bit-exact with the oracle (same products in the same order, contraction
off), for the 16-byte path (aligned buffers) and the 8-byte path (a buffer
starting one double off), every term together and single terms with the
rest of the output left alone.  Then one MHD step on the device: backward
of (u, w, B, J, T, C) -> products -> forward of the six outputs, against the
oracle's chain, relative L2 1e-12 per component (FP64 engine).
"""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import ShellConfig, SphTransPlan, gaussian, synthetic_spectra

pytestmark = pytest.mark.gpu

IDX = {"light": 0, "magne": 1, "temp": 4, "velo": 5, "current": 8, "vort": 11,
       "c_flux": 0, "coriolis": 3, "m_advect": 7, "lorentz": 10, "h_flux": 13, "vp_induct": 16}
COEF = {"m_advect": -1.5, "lorentz": 0.7, "vp_induct": 2.25, "h_flux": -0.3, "c_flux": 1.1,
        "coriolis": 3.5}


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("shift", [0, 1], ids=["aligned", "offset"])
@pytest.mark.parametrize("dims", [(31, 33, 48, 96), (127, 17, 192, 384), (45, 3, 69, 138)])
def test_nonlinear_matches_oracle(torch_cuda, dims, shift):
    t = torch_cuda
    cfg = ShellConfig(*dims)
    _say(f"[nonlinear] dims={dims} shift={shift}")
    nf, no = 14, 21
    fld = gaussian(50 + dims[0], cfg.n_rtp(nf))
    init = np.full(cfg.n_rtp(no), 9.0)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    ref = orc.nonlinear_rtp(IDX, COEF, nf, fld, no, init)
    plan = SphTransPlan(cfg, max_ncomp=4)
    try:
        fbuf = t.zeros(fld.size + shift, dtype=t.float64, device="cuda")
        obuf = t.zeros(init.size + shift, dtype=t.float64, device="cuda")
        fd, od = fbuf[shift:], obuf[shift:]
        fd.copy_(t.from_numpy(fld))
        od.copy_(t.from_numpy(init))
        plan.nonlinear_terms_in_rtp(IDX, COEF, nf, fd, no, od)
        t.cuda.synchronize()
        got = od.cpu().numpy()
    finally:
        plan.close()
    diff = int(np.count_nonzero(got != ref))
    _say(f"    entries differing from the oracle: {diff} of {got.size}")
    assert diff == 0


@pytest.mark.parametrize("term,inputs", [("coriolis", ("velo",)), ("lorentz", ("current", "magne")),
                                         ("h_flux", ("velo", "temp"))])
def test_single_term(torch_cuda, term, inputs):
    t = torch_cuda
    cfg = ShellConfig(15, 5, 24, 48)
    idx = {n: IDX[n] for n in inputs}
    idx[term] = 2
    fld = gaussian(7, cfg.n_rtp(14))
    init = np.full(cfg.n_rtp(6), 4.0)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    ref = orc.nonlinear_rtp(idx, COEF, 14, fld, 6, init)
    plan = SphTransPlan(cfg, max_ncomp=4)
    try:
        od = t.tensor(init, device="cuda")
        plan.nonlinear_terms_in_rtp(idx, COEF, 14, t.tensor(fld, device="cuda"), 6, od)
        got = od.cpu().numpy()
    finally:
        plan.close()
    assert np.array_equal(got, ref)
    assert (got.reshape(6, -1)[:2] == 4.0).all() and (got.reshape(6, -1)[5] == 4.0).all()


def test_nonlinear_rejects_missing_inputs(torch_cuda):
    t = torch_cuda
    cfg = ShellConfig(7, 2, 8, 16)
    plan = SphTransPlan(cfg, max_ncomp=4)
    try:
        f = t.zeros(cfg.n_rtp(3), dtype=t.float64, device="cuda")
        o = t.zeros(cfg.n_rtp(3), dtype=t.float64, device="cuda")
        with pytest.raises(Exception, match="vort"):
            plan.nonlinear_terms_in_rtp({"velo": 0, "m_advect": 0}, COEF, 3, f, 3, o)
        with pytest.raises(Exception, match="outside"):
            plan.nonlinear_terms_in_rtp({"velo": 0, "coriolis": 1}, COEF, 3, f, 3, o)
        with pytest.raises(Exception, match="positive"):
            plan.nonlinear_terms_in_rtp({"velo": 0}, COEF, 0, f, 3, o)
        # every term off: a one-component output is accepted and left alone
        o1 = t.full((cfg.n_rtp(1),), 3.0, dtype=t.float64, device="cuda")
        plan.nonlinear_terms_in_rtp({}, COEF, 3, f, 1, o1)
        assert bool((o1 == 3.0).all())
    finally:
        plan.close()


def _rel(a, b, ncomp):
    a = np.asarray(a).reshape(-1, ncomp)
    b = np.asarray(b).reshape(-1, ncomp)
    return max(np.linalg.norm(a[:, c] - b[:, c]) / max(np.linalg.norm(b[:, c]), 1e-300)
               for c in range(ncomp))


def test_mhd_step_on_device(torch_cuda):
    """backward (4 vectors u, w, B, J + scalars T, C) -> products -> forward
    (6 vectors), all on the device, against the oracle's chain."""
    t = torch_cuda
    cfg = ShellConfig(31, 33, 48, 96)
    nvb, nsb, nvf = 4, 2, 6
    idx = {"velo": 0, "vort": 3, "magne": 6, "current": 9, "temp": 12, "light": 13,
           "m_advect": 0, "lorentz": 3, "vp_induct": 6, "h_flux": 9, "c_flux": 12,
           "coriolis": 15}
    sp = synthetic_spectra(cfg, nvb, nsb, seed=21)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    grid = orc.backward(nvb, nsb, sp)
    frc = orc.nonlinear_rtp(idx, COEF, 14, grid, 18)
    ref = orc.forward(nvf, 0, frc)
    plan = SphTransPlan(cfg, max_ncomp=18)
    try:
        g = t.empty(cfg.n_rtp(14), dtype=t.float64, device="cuda")
        f = t.empty(cfg.n_rtp(18), dtype=t.float64, device="cuda")
        out = t.empty(cfg.n_rj(18), dtype=t.float64, device="cuda")
        plan.sph_backward_transforms(nvb, nsb, t.tensor(sp, device="cuda"), g)
        plan.nonlinear_terms_in_rtp(idx, COEF, 14, g, 18, f)
        plan.sph_forward_transforms(nvf, 0, f, out)
        got = out.cpu().numpy()
    finally:
        plan.close()
    err = _rel(got, ref, 18)
    _say(f"[mhd step] forward of the products: rel L2 {err:.2e}")
    assert err <= 1e-12
