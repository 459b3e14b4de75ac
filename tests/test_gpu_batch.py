"""Shell batch (many uneven shells replayed as one HIP graph) on the GPU.

Bar: every replay of the graph reproduces, bit for bit, the eager per-shell
transforms of the same plans on the same inputs (the kernels are
deterministic; a graph changes only how they are launched).  The graph is
replayed several times with eager work and synchronisations in between, and
each replay is checked.  Shells follow the batched-shell workload: Ntheta =
3(L+1)/2, Nphi = 2 Ntheta, one vector and one scalar per shell in both
directions; the L here are small and uneven, and mix the split-form and the
generic FFT lengths."""
import numpy as np
import pytest

from kemorin_mhd_amd import (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellBatch, ShellConfig,
                             SphTransPlan, gaussian, synthetic_spectra)

pytestmark = pytest.mark.gpu

# (L, Nr); L = 45 gives Nphi = 138 = 2 * 3 * 23 (a direct-sum FFT pass),
# L = 105 gives Nphi = 318 = 2 * 3 * 53 (a Bluestein pass)
SHELLS = [(15, 5), (47, 4), (31, 7), (23, 6), (63, 9), (39, 3), (45, 2), (105, 2)]


def _say(*a):
    print(*a, flush=True)


def _shells():
    return [ShellConfig(L, nr, 3 * (L + 1) // 2, 3 * (L + 1)) for L, nr in SHELLS]


@pytest.mark.parametrize("precision", [SHT_PREC_FP64, SHT_PREC_SPLIT_INT8])
def test_batch_graph_matches_eager(torch_cuda, precision):
    torch = torch_cuda
    dev = "cuda"
    cfgs = _shells()
    _say(f"[batch] precision={precision} shells={SHELLS}")
    plans, bufs = [], []
    for i, c in enumerate(cfgs):
        p = SphTransPlan(c, max_ncomp=4, precision=precision)
        ws = torch.tensor(synthetic_spectra(c, 1, 1, seed=100 + i), device=dev)
        vin = torch.tensor(gaussian(200 + i, c.n_rtp(4)), device=dev)
        vout = torch.zeros(c.n_rtp(4), dtype=torch.float64, device=dev)
        wr = torch.zeros(c.n_rj(4), dtype=torch.float64, device=dev)
        plans.append(p)
        bufs.append((ws, vout, vin, wr))
    # eager, each plan on the current stream
    ref = []
    for p, (ws, vout, vin, wr) in zip(plans, bufs):
        p.sph_backward_transforms(1, 1, ws, vout)
        p.sph_forward_transforms(1, 1, vin, wr)
    torch.cuda.synchronize()
    for (_, vout, _, wr) in bufs:
        ref.append((vout.clone(), wr.clone()))
    batch = ShellBatch([(p, (1, 1), ws, vout, (1, 1), vin, wr)
                        for p, (ws, vout, vin, wr) in zip(plans, bufs)])
    for rep in range(3):
        for (_, vout, _, wr) in bufs:
            vout.fill_(np.nan)
            wr.fill_(np.nan)
        torch.cuda.synchronize()
        batch.launch()
        batch.synchronize()
        for i, ((_, vout, _, wr), (rv, rw)) in enumerate(zip(bufs, ref)):
            assert torch.equal(vout, rv), f"replay {rep} shell {i}: backward differs from eager"
            assert torch.equal(wr, rw), f"replay {rep} shell {i}: forward differs from eager"
        _say(f"[batch] replay {rep} bit-exact")
        # eager work between replays: one more eager transform on shell 0
        plans[0].sph_backward_transforms(1, 1, bufs[0][0], bufs[0][1])
        torch.cuda.synchronize()
    batch.close()
    for p in plans:
        p.close()


def test_batch_calls_sharing_a_plan_run_in_order(torch_cuda):
    # two calls on one plan (one workspace): the second must see the first
    # done, so the graph result equals two eager calls in a row
    torch = torch_cuda
    c = _shells()[1]
    p = SphTransPlan(c, max_ncomp=4)
    ws = [torch.tensor(synthetic_spectra(c, 1, 1, seed=300 + i), device="cuda") for i in range(2)]
    v = [torch.zeros(c.n_rtp(4), dtype=torch.float64, device="cuda") for _ in range(2)]
    for i in range(2):
        p.sph_backward_transforms(1, 1, ws[i], v[i])
    torch.cuda.synchronize()
    ref = [x.clone() for x in v]
    calls = [(p, (1, 1), ws[i], v[i], (0, 0), None, None) for i in range(2)]
    other = SphTransPlan(_shells()[0], max_ncomp=4)
    c0 = _shells()[0]
    wo = torch.tensor(synthetic_spectra(c0, 1, 1, seed=9), device="cuda")
    vo = torch.zeros(c0.n_rtp(4), dtype=torch.float64, device="cuda")
    calls.insert(1, (other, (1, 1), wo, vo, (0, 0), None, None))
    b = ShellBatch(calls)
    for rep in range(2):
        for x in v:
            x.fill_(np.nan)
        torch.cuda.synchronize()
        b.launch()
        b.synchronize()
        for i in range(2):
            assert torch.equal(v[i], ref[i]), f"replay {rep} call {i}"
    b.close()
    other.close()
    p.close()


def test_batch_rejects_timed_plans(torch_cuda):
    torch = torch_cuda
    c = _shells()[0]
    p = SphTransPlan(c, max_ncomp=4)
    ws = torch.zeros(c.n_rj(4), dtype=torch.float64, device="cuda")
    v = torch.zeros(c.n_rtp(4), dtype=torch.float64, device="cuda")
    p.timer_enable(True)
    with pytest.raises(Exception, match="timers"):
        ShellBatch([(p, (1, 1), ws, v, (0, 0), None, None)])
    p.timer_enable(False)
    b = ShellBatch([(p, (1, 1), ws, v, (0, 0), None, None)])
    b.launch()
    b.synchronize()
    b.close()
    p.close()
