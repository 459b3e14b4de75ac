"""Many fields in one call: plans with max_ncomp = 64, the plan's limit, on
both engines.  The reference's MHD step hands its Legendre transforms every
field of a direction at once (velocity, vorticity, magnetic field, current,
temperature, composition and their gradients: twenty-odd components), so
the component-group paths of the kernels (vector groups, scalars fused into
the vector launch of the split backward, scalar-only calls) run here at the
full 64 components: 20 vectors + 4 scalars, 21 vectors + 1 scalar and 64
scalars each way, on a shell whose FFT takes the generic form and one whose
FFT takes the split form, with stored P tables and with P generated per
call (SHT_FLAG_P_ON_THE_FLY).  Inputs are seeded Gaussian spectra and grids.
Bars against the oracle, per component: FP64 relative L2 1e-12, split
(6 slices) 1e-11."""
import numpy as np
import pytest

import oracle as O
from kemorin_mhd_amd import SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellConfig, SphTransPlan

pytestmark = pytest.mark.gpu

SHELLS = {"L15_generic_fft": ShellConfig(15, 3, 24, 48),
          "L31_split_form_fft": ShellConfig(31, 5, 48, 96)}
FIELD_SETS = [(20, 4), (21, 1), (0, 64)]


def _rel_comp(a, b, ncomp, comp_fastest):
    a = np.asarray(a).reshape(-1, ncomp) if comp_fastest else np.asarray(a).reshape(ncomp, -1).T
    b = np.asarray(b).reshape(-1, ncomp) if comp_fastest else np.asarray(b).reshape(ncomp, -1).T
    out = 0.0
    for c in range(ncomp):
        nb = np.linalg.norm(b[:, c])
        out = max(out, np.linalg.norm(a[:, c] - b[:, c]) / (nb if nb > 0 else 1.0))
    return out


@pytest.mark.parametrize("prec,tol", [(SHT_PREC_FP64, 1e-12), (SHT_PREC_SPLIT_INT8, 1e-11)],
                         ids=["fp64", "split"])
@pytest.mark.parametrize("otf", [False, True], ids=["tables", "p_on_the_fly"])
@pytest.mark.parametrize("name", list(SHELLS))
def test_64_components_vs_oracle(torch_cuda, name, otf, prec, tol):
    t = torch_cuda
    cfg = SHELLS[name]
    orc = O.OraclePlan(cfg.L, cfg.nr, cfg.nth, cfg.nphi, cfg.radius(), full_tables=False)
    rng = np.random.default_rng(640 + cfg.L)
    plan = SphTransPlan(cfg, max_ncomp=64, precision=prec, p_on_the_fly=otf)
    try:
        for nv, ns in FIELD_SETS:
            nc = 3 * nv + ns
            print(f"[gpu] {name} otf={otf} precision={prec} nv={nv} ns={ns}", flush=True)
            sp = rng.standard_normal(cfg.n_rj(nc))
            grid = rng.standard_normal(cfg.n_rtp(nc))
            ref_b = orc.backward(nv, ns, sp)
            ref_f = orc.forward(nv, ns, grid)
            v = t.empty(cfg.n_rtp(nc), dtype=t.float64, device="cuda")
            s = t.empty(cfg.n_rj(nc), dtype=t.float64, device="cuda")
            plan.sph_backward_transforms(nv, ns, t.from_numpy(sp).cuda(), v)
            plan.sph_forward_transforms(nv, ns, t.from_numpy(grid).cuda(), s)
            t.cuda.synchronize()
            eb = _rel_comp(v.cpu().numpy(), ref_b, nc, comp_fastest=False)
            ef = _rel_comp(s.cpu().numpy(), ref_f, nc, comp_fastest=True)
            print(f"    backward {eb:.2e} forward {ef:.2e}", flush=True)
            assert eb <= tol and ef <= tol, (nv, ns, eb, ef)
    finally:
        plan.close()
