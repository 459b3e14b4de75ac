"""Device copies between field-major rj spectra and the interleaved send /
receive buffers (sht_rj_fields_to_send / sht_rj_fields_from_recv) against
the oracle (itself pinned bit for bit against the reference's
field_to_send_buffer.f90): bit-exact, with the identity tables the engine
uses and with permutation / reverse tables holding dummy records; a field
list of the MHD step's backward set (four vectors, two scalars) plus a
tensor; and the copies round a whole backward transform: rj fields -> send
buffer -> sht_backward equals the oracle's backward of the packed buffer."""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import ShellConfig, SphTransPlan, synthetic_spectra

pytestmark = pytest.mark.gpu

FIELDS = [(3, 0, 3), (9, 3, 3), (0, 6, 3), (15, 9, 3), (6, 12, 1), (21, 13, 1), (22, 14, 6)]
NTOT, NB = 28, 20


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("tables", [False, True], ids=["identity", "tables"])
def test_copies_bitexact_vs_oracle(torch_cuda, tables):
    t = torch_cuda
    cfg = ShellConfig(31, 33, 48, 96)
    n = cfg.jmax * cfg.nr
    nrecv = n - 37
    _say(f"[gpu] rj field copies, tables={tables}, {n} nodes")
    rng = np.random.default_rng(11)
    fld = rng.standard_normal(NTOT * n)
    ws0 = rng.standard_normal(NB * n)
    wr = rng.standard_normal(NB * nrecv)
    d0 = rng.standard_normal(NTOT * n)
    inod = rng.permutation(n).astype(np.int32) + 1 if tables else None
    irev = rng.integers(1, nrecv + 2, size=n).astype(np.int32) if tables else None
    plan = SphTransPlan(cfg, max_ncomp=9)
    try:
        d_fld = t.from_numpy(fld).cuda()
        d_ws = t.from_numpy(ws0).cuda()
        plan.copy_spectr_to_send(FIELDS, NTOT, d_fld, NB, d_ws,
                                 inod_export=t.from_numpy(inod).cuda() if tables else None)
        want = O.rj_to_send(FIELDS, NTOT, n, fld, NB, ws0, inod_export=inod)
        t.cuda.synchronize()
        assert np.array_equal(d_ws.cpu().numpy(), want)
        if tables:
            d_wr = t.from_numpy(wr).cuda()
            d_out = t.from_numpy(d0).cuda()
            plan.copy_spectr_from_recv(FIELDS, NB, d_wr, NTOT, d_out, irev=t.from_numpy(irev).cuda())
            want = O.rj_from_recv(FIELDS, NB, nrecv, wr, NTOT, n, d0, irev=irev)
        else:
            d_wr = t.from_numpy(ws0[:NB * n]).cuda()
            d_out = t.from_numpy(d0).cuda()
            plan.copy_spectr_from_recv(FIELDS, NB, d_wr, NTOT, d_out)
            want = O.rj_from_recv(FIELDS, NB, n, ws0, NTOT, n, d0)
        t.cuda.synchronize()
        assert np.array_equal(d_out.cpu().numpy(), want)
    finally:
        plan.close()


def test_pack_then_backward(torch_cuda):
    """rj fields (velo, magne, temp, light at scattered components) packed on
    the device into the 2 vector + 2 scalar send buffer, then transformed:
    the grid equals the oracle's backward of the oracle's packing."""
    t = torch_cuda
    cfg = ShellConfig(15, 5, 24, 48)
    n = cfg.jmax * cfg.nr
    ntot = 12
    fields = [(4, 0, 3), (0, 3, 3), (9, 6, 1), (11, 7, 1)]
    sp = synthetic_spectra(cfg, 2, 2, seed=21)  # interleaved, as a source of values
    fld = np.zeros((ntot, n))
    for i_fld, i_tr, nc in fields:
        fld[i_fld:i_fld + nc] = sp.reshape(n, 8)[:, i_tr:i_tr + nc].T
    plan = SphTransPlan(cfg, max_ncomp=8)
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    try:
        d_ws = t.zeros(8 * n, dtype=t.float64, device="cuda")
        plan.copy_spectr_to_send(fields, ntot, t.from_numpy(fld.ravel()).cuda(), 8, d_ws)
        v = t.empty(cfg.n_rtp(8), dtype=t.float64, device="cuda")
        plan.sph_backward_transforms(2, 2, d_ws, v)
        t.cuda.synchronize()
        ws = O.rj_to_send(fields, ntot, n, fld.ravel(), 8, np.zeros(8 * n))
        assert np.array_equal(d_ws.cpu().numpy(), ws)
        ref = orc.backward(2, 2, ws)
        got = v.cpu().numpy()
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
    finally:
        plan.close()
