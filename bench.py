#!/usr/bin/env python3
"""Round-trip throughput of the MI355X spherical-harmonic transform.

One step = one round trip on device-resident synthetic data: a backward
transform of 2 vector + 2 scalar fields (8 components) and a forward
transform of 3 vector fields (9 components), the field set of the
reference's MHD time step (MS/cal_nonlinear.f90:149,164).  The default
workload is L=511, Nr=257, Ntheta=768, Nphi=1536 on one GPU.

Prints ONE JSON line.  `value` is round trips/s of the split-precision
(INT8-sliced) engine when that mode is available, otherwise of the FP64-MFMA
engine; both are reported.  Per-kernel times come from HIP events on the
plan's stream (sht_plan_timer_*), `roofline` is for the kernel with the
largest share of the step, `cpu_baseline` is the oracle's FP64 restatement
of the reference transform (symmetric_BLAS_big, OpenMP) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
INT8_PEAK_TOPS = 5000.0

CONFIGS = {
    "cfg0": dict(L=31, nr=33, nth=48, nphi=96),
    "cfg1": dict(L=255, nr=129, nth=384, nphi=768),
    "cfg2": dict(L=511, nr=257, nth=768, nphi=1536),
    # the large-truncation grid at a quarter of its radii (the full Nr=257
    # round trip with the split plan does not fit in 288 GB with all four
    # field buffers resident)
    "cfg4_nr65": dict(L=1023, nr=65, nth=1536, nphi=3072),
    # the large-truncation grid at all 257 radii: P_l^m generated per call one
    # tile of orders at a time (the stored tables would be 12.9 GB, their
    # split slices 12.1 GB more), and the backward writes its 8 grid
    # components into the first 8 of the forward's 9 (so the forward of a
    # step transforms the backward's output), which keeps the split plan and
    # the field buffers within 288 GB
    "cfg4": dict(L=1023, nr=257, nth=1536, nphi=3072, otf=True, alias=True),
}
METRIC = "SHT round trips/s + % roofline at L=511,Nr=257: split-prec vs FP64 MFMA"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="cfg2", choices=list(CONFIGS))
    ap.add_argument("--precision", default="auto", choices=["auto", "fp64", "split", "both"])
    ap.add_argument("--slices", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-radii", type=int, default=16, help="radii in the CPU sample")
    ap.add_argument("--copy-calib", type=float, default=0.0,
                    help="only run a device copy of this many GiB (PMC calibration)")
    ap.add_argument("--shells", type=int, default=0,
                    help="only run the batched-shell workload: this many seeded uneven shells, "
                         "one HIP graph launch against a loop of per-shell calls")
    ap.add_argument("--autotune", metavar="CACHE", default=None,
                    help="tune every plan's kernel variants (sht_plan_autotune) with this JSON "
                         "cache before timing, and report the tuned against the default")
    ap.add_argument("--variant", action="append", default=[], metavar="NAME=VALUE",
                    help="kernel variant for every plan (sht_plan_set_variant), repeatable")
    ap.add_argument("--nonlinear", action="store_true",
                    help="only time the grid-space products of one MHD step (all six terms) "
                         "on the configuration's grid")
    return ap.parse_args()


def run_nonlinear(args, torch):
    """The grid-space products of one MHD step on the configuration's grid:
    inputs u, w, B, J (vectors) and T, C (scalars), 14 components; outputs
    w x u, J x B, u x B, u T, u C and the Coriolis term, 18 components, all in
    one launch.  Algorithmic traffic: (14 + 18) * 8 bytes per grid point.
    Timed with HIP events on the plan's stream over the timed launches."""
    from kemorin_mhd_amd import ShellConfig, SphTransPlan
    c = CONFIGS[args.config]
    cfg = ShellConfig(c["L"], c["nr"], c["nth"], c["nphi"])
    nnod = cfg.nr * cfg.nth * cfg.nphi
    idx = {"velo": 0, "vort": 3, "magne": 6, "current": 9, "temp": 12, "light": 13,
           "m_advect": 0, "lorentz": 3, "vp_induct": 6, "h_flux": 9, "c_flux": 12,
           "coriolis": 15}
    coef = {"m_advect": 1.0, "lorentz": 0.5, "vp_induct": 1.0, "h_flux": -1.0, "c_flux": -1.0,
            "coriolis": 2.0}
    plan = SphTransPlan(cfg, max_ncomp=1)
    stream = torch.cuda.Stream()
    plan.set_stream(stream.cuda_stream)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    with torch.cuda.stream(stream):
        fld = torch.randn(14 * nnod, generator=gen, dtype=torch.float64, device="cuda")
        frc = torch.empty(18 * nnod, dtype=torch.float64, device="cuda")
    stream.synchronize()
    for _ in range(args.warmup):
        plan.nonlinear_terms_in_rtp(idx, coef, 14, fld, 18, frc)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        plan.nonlinear_terms_in_rtp(idx, coef, 14, fld, 18, frc)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    nbytes = 32 * 8 * nnod
    print(json.dumps({"metric": "grid-space products of one MHD step (6 terms, one launch)",
                      "value": nbytes / ms / 1e6, "unit": "GB/s", "ms_per_launch": ms,
                      "steps": args.steps, "warmup": args.warmup, "dtype": "f64",
                      "data": "synthetic (seeded Gaussian fields)",
                      "config": {"workload": f"{args.config}: Nr={cfg.nr} Ntheta={cfg.nth} "
                                             f"Nphi={cfg.nphi}, 14 input and 18 output components"},
                      "roofline": {"kernel": "ksht::k_nonlinear<2>", "bound": "hbm",
                                   "achieved": nbytes / ms / 1e6, "peak": 8000.0, "unit": "GB/s",
                                   "frac": nbytes / ms / 1e6 / 8000.0,
                                   "algorithmic_bytes": nbytes,
                                   "traffic": (hbm_traffic("products", "ksht::k_nonlinear<2>")
                                               if args.config == "cfg2" else None)}}))
    plan.close()


def run_shells(args, torch):
    """Batched uneven shells: L drawn from the odd values in 63..383 (Ntheta =
    3(L+1)/2, Nphi = 2 Ntheta), Nr from 33..129, one vector + one scalar per
    shell in both directions.  A step is one backward and one forward
    transform of every shell: as a loop of per-shell calls on one stream, and
    as one launch of the shells' captured graph."""
    from kemorin_mhd_amd import (SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, ShellBatch, SphTransPlan,
                                 batched_shell_draw)
    shells = batched_shell_draw(args.shells)
    Ls = [c.L for c in shells]
    nrs = [c.nr for c in shells]
    precs = ["fp64", "split"] if args.precision in ("auto", "both") else [args.precision]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(20261202)
    flops = sum(flops_per_contraction(dict(L=L, nr=nr, nth=3 * (L + 1) // 2)) * (6 + 6)
                for L, nr in zip(Ls, nrs))
    out = {"metric": "batched uneven shells: all-shell round trips/s, one HIP graph vs a loop "
                     "of per-shell calls", "unit": "round_trips/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
           "data": "synthetic (seeded Gaussian spectra and grids)",
           "config": {"workload": f"{args.shells} shells, L odd in 63..383, Nr in 33..129, "
                                  "Ntheta=3(L+1)/2, Nphi=2 Ntheta, 1 vector + 1 scalar each way",
                      "L": Ls, "nr": nrs},
           "fp64_equivalent_flop_per_step": flops, "engines": {}}
    stream = torch.cuda.Stream()
    for prec in precs:
        plans, bufs, tunes = [], [], []
        with torch.cuda.stream(stream):
            for c in shells:
                p = SphTransPlan(c, max_ncomp=4, precision=SHT_PREC_SPLIT_INT8 if prec == "split"
                                 else SHT_PREC_FP64, n_slices=args.slices)
                p.set_stream(stream.cuda_stream)
                apply_variants(p, args)
                if args.autotune:
                    tunes.append(p.autotune(1, 1, 1, 1, cache=args.autotune))
                ws = torch.randn(c.n_rj(4), generator=gen, dtype=torch.float64, device="cuda")
                vin = torch.randn(c.n_rtp(4), generator=gen, dtype=torch.float64, device="cuda")
                vout = torch.empty(c.n_rtp(4), dtype=torch.float64, device="cuda")
                wr = torch.empty(c.n_rj(4), dtype=torch.float64, device="cuda")
                plans.append(p)
                bufs.append((ws, vout, vin, wr))
        stream.synchronize()
        plan_bytes = sum(p.device_bytes() for p in plans)

        def loop():
            for p, (ws, vout, vin, wr) in zip(plans, bufs):
                p.sph_backward_transforms(1, 1, ws, vout)
                p.sph_forward_transforms(1, 1, vin, wr)

        batch = ShellBatch([(p, (1, 1), ws, vout, (1, 1), vin, wr)
                            for p, (ws, vout, vin, wr) in zip(plans, bufs)])
        res = {}
        for name, fn in (("loop", loop), ("graph", lambda: batch.launch(stream.cuda_stream))):
            for _ in range(args.warmup):
                fn()
            stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                fn()
            stream.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / args.steps
            res[name] = {"ms_per_step": ms, "round_trips_per_s": 1e3 / ms,
                         "gflops_fp64_equivalent": flops / ms / 1e6}
        res["graph_speedup_vs_loop"] = res["loop"]["ms_per_step"] / res["graph"]["ms_per_step"]
        res["plan_device_bytes"] = plan_bytes
        if tunes:
            # the tuner's own round-trip times of each shell, tuned against default
            res["autotune"] = {
                "tuned_ms_sum": sum(x["ms"] for x in tunes),
                "default_ms_sum": sum(x["default_ms"] for x in tunes),
                "chosen": [x["variants"] for x in tunes]}
        res["buffer_bytes"] = sum(8 * sum(t.numel() for t in b) for b in bufs)
        out["engines"][prec] = res
        batch.close()
        for p in plans:
            p.close()
        del bufs
        torch.cuda.empty_cache()
    main_prec = precs[-1]
    out["value"] = out["engines"][main_prec]["graph"]["round_trips_per_s"]
    out["engine"] = main_prec
    print(json.dumps(out))


def flops_per_contraction(c):
    return 2.0 * ((c["nth"] + 1) // 2) * (c["L"] + 1) ** 2 * c["nr"]


def apply_variants(plan, args):
    for kv in args.variant:
        name, value = kv.split("=")
        plan.set_variant(name, int(value))


def make_config(cfg_d):
    from kemorin_mhd_amd import ShellConfig
    return ShellConfig(cfg_d["L"], cfg_d["nr"], cfg_d["nth"], cfg_d["nphi"])


def make_inputs(cfg, torch, slope=1.5, seed=20261202, otf=False):
    """Seeded spectra sigma(l, r) = (l+1)^-slope sin(pi (r-ri)/(ro-ri)) for the
    backward (8 comps), and a band-limited grid for the forward (9 comps), made
    by the FP64 engine's backward of such spectra."""
    from kemorin_mhd_amd import SHT_PREC_FP64, SphTransPlan, rlm_l_m
    dev = "cuda"
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    l, _ = rlm_l_m(cfg.L)
    r = cfg.radius()
    prof = torch.tensor(np.sin(np.pi * (r - cfg.ri) / (cfg.ro - cfg.ri)), device=dev)
    sig = torch.tensor((l + 1.0) ** -slope, device=dev)

    def spectra(ncomp):
        x = torch.randn((cfg.nr, cfg.jmax, ncomp), generator=gen, dtype=torch.float64, device=dev)
        x *= sig[None, :, None]
        x *= prof[:, None, None]
        return x.reshape(-1)

    sp_b = spectra(8)
    grid_f = torch.empty(cfg.n_rtp(9), dtype=torch.float64, device=dev)
    tmp = spectra(9)
    plan = SphTransPlan(cfg, max_ncomp=9, precision=SHT_PREC_FP64, p_on_the_fly=otf)
    plan.sph_backward_transforms(3, 0, tmp, grid_f)
    torch.cuda.synchronize()
    plan.close()
    del tmp
    return sp_b, grid_f


def run_engine(cfg, precision, args, torch, sp_b, grid_f, keep=False, otf=False, alias=False):
    from kemorin_mhd_amd import SHT_PREC_FP64, SHT_PREC_SPLIT_INT8, SphTransPlan
    prec = SHT_PREC_SPLIT_INT8 if precision == "split" else SHT_PREC_FP64
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    plan = SphTransPlan(cfg, max_ncomp=9, precision=prec, n_slices=args.slices,
                        p_on_the_fly=otf)
    plan.set_stream(stream.cuda_stream)
    apply_variants(plan, args)
    tuned = None
    if args.autotune:
        from kemorin_mhd_amd import ShtError
        from kemorin_mhd_amd.sht import SHT_ERR_OOM
        try:
            tuned = plan.autotune(2, 2, 3, 0, cache=args.autotune)
        except ShtError as e:
            # the timing runs need field buffers of their own; where they do
            # not fit (config 4 at full size) the plan keeps its default forms
            if e.status != SHT_ERR_OOM:
                raise
            tuned = {"skipped": str(e)}
    dev = "cuda"
    with torch.cuda.stream(stream):
        # aliased: the backward's 8 grid components are the forward's first 8
        grid_b = (grid_f[: cfg.n_rtp(8)] if alias
                  else torch.empty(cfg.n_rtp(8), dtype=torch.float64, device=dev))
        sp_f = torch.empty(cfg.n_rj(9), dtype=torch.float64, device=dev)
    stream.synchronize()
    free_b, total_b = torch.cuda.mem_get_info()

    def step():
        plan.sph_backward_transforms(2, 2, sp_b, grid_b)
        plan.sph_forward_transforms(3, 0, grid_f, sp_f)

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    plan.timer_read()  # reset
    plan.timer_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    stream.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timers = plan.timer_read()
    plan.timer_enable(False)
    dev_bytes = plan.device_bytes()
    # the split backward's vector launches also take one scalar each (5 or 6
    # slices, the default variant): 2 x (5 + 1) contractions in one kernel
    fused = (precision == "split" and (args.slices or 6) <= 6
             and plan.get_variant("split_bwd_fused") == 1)
    plan.close()
    out = (grid_b, sp_f) if keep else None
    if not keep:
        del grid_b, sp_f
    torch.cuda.empty_cache()
    ms = (t1 - t0) * 1e3 / args.steps
    return dict(ms_per_step=ms, round_trips_per_s=1e3 / ms, timers=timers, plan_bytes=dev_bytes,
                outputs=out, slices=(args.slices or 6) if precision == "split" else 0,
                hbm_in_use_bytes=total_b - free_b, hbm_total_bytes=total_b, autotune=tuned,
                bwd_fused=fused)


FIELDS_BWD = (("vector1", 0, 3), ("vector2", 3, 6), ("scalar1", 6, 7), ("scalar2", 7, 8))
FIELDS_FWD = (("vector1", 0, 3), ("vector2", 3, 6), ("vector3", 6, 9))


def field_errors(cfg, torch, out, ref):
    """Per-field relative L2 and max-norm errors of `out` against `ref` for the
    backward grid (comp blocks of nnod) and the forward spectra (comp fastest)."""
    nnod = cfg.nr * cfg.nth * cfg.nphi
    gb, sf = out
    rb, rf = ref
    res = {"backward": {}, "forward": {}}
    for name, c0, c1 in FIELDS_BWD:
        x, y = gb[c0 * nnod:c1 * nnod], rb[c0 * nnod:c1 * nnod]
        d = x - y
        res["backward"][name] = {"rel_l2": float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(y)),
                                 "rel_max": float(d.abs().max() / y.abs().max())}
    xs, ys = sf.view(-1, 9), rf.view(-1, 9)
    for name, c0, c1 in FIELDS_FWD:
        d = xs[:, c0:c1] - ys[:, c0:c1]
        y = ys[:, c0:c1]
        res["forward"][name] = {"rel_l2": float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(y)),
                                "rel_max": float(d.abs().max() / y.abs().max())}
    return res


def split_accuracy(cfg, args, torch, results):
    """Split engine vs the FP64-MFMA engine (itself within 1e-12 of the CPU
    FP64 oracle at parity-test sizes) on the benchmark spectra and on the
    steeper (l+1)^-3 spectra."""
    acc = {"slope_1.5": field_errors(cfg, torch, results["split"]["outputs"],
                                     results["fp64"]["outputs"])}
    for p in results.values():
        p["outputs"] = None
    torch.cuda.empty_cache()
    sp3, gf3 = make_inputs(cfg, torch, slope=3.0, seed=7)
    one = argparse.Namespace(steps=1, warmup=0, slices=args.slices, variant=args.variant,
                             autotune=args.autotune)
    o64 = run_engine(cfg, "fp64", one, torch, sp3, gf3, keep=True)["outputs"]
    osp = run_engine(cfg, "split", one, torch, sp3, gf3, keep=True)["outputs"]
    acc["slope_3"] = field_errors(cfg, torch, osp, o64)
    del o64, osp, sp3, gf3
    torch.cuda.empty_cache()
    return acc


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sht_roundtrip_L511_Nr257_hbm_traffic.txt")


def hbm_traffic(engine, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE / WRITE_SIZE, scaled by a device-copy calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("#"):
                    continue
                parts = [x.strip() for x in line.split("|")]
                if len(parts) == 6 and parts[0] == engine and parts[1] == kernel:
                    return float(parts[5])
    except OSError:
        return None
    return None


def kernel_roofline(cfg_d, res, precision, slices):
    """Roofline of the round trip's dominant kernel.  Legendre kernels: the
    vector-group kernel alone (stage timers 5/6, HIP events on the plan's
    stream), algorithmic work = FP64 flops of its contractions (10 F1 backward:
    2 vectors x 5, or 12 F1 where the split backward's vector launch also
    takes the two scalars; 15 F1 forward: 3 vectors x 5); in the split engine each FP64
    multiply-add is S(S+1)/2 INT8 slice products, so its INT8 work is that
    many times the FP64 flops, against the INT8 MFMA peak.  FFT stages: the
    bytes a launch must move (X in, grid out, or the reverse)."""
    F1 = flops_per_contraction(cfg_d)
    nline = cfg_d["nr"] * cfg_d["nth"]
    nslot = 2 * cfg_d["L"] + 1
    tri = slices * (slices + 1) // 2
    # the plan's FFT form: split form Nphi = R0 NP (NP the smallest power of two
    # above L, 32..512) with 4 (backward) and 8 (forward) complex lines per
    # work-group, else the mixed-radix kernels
    npow = 32
    while npow < cfg_d["L"] + 1:
        npow *= 2
    fast = npow <= 512 and cfg_d["nphi"] % npow == 0
    split = precision == "split"
    kname = {
        "legendre_backward_vector_kernel": (
            f"ksht::k_leg_bwd_split<{slices}, {4 if res.get('bwd_fused') else 3}, true>" if split
            else "ksht::k_leg_bwd<true>"),
        "legendre_forward_vector_kernel": (f"ksht::k_leg_fwd_split<{slices}, 3, true>" if split
                                           else "ksht::k_leg_fwd<true>"),
        "fft_backward": f"ksht::k_fft3_bwd<{npow}, 4>" if fast else "ksht::k_fft_bwd (mixed radix)",
        "fft_forward": f"ksht::k_fft3_fwd<{npow}, 8>" if fast else "ksht::k_fft_fwd (mixed radix)",
    }
    stages = {
        "legendre_backward_vector_kernel": ("mfma", (12 if res.get("bwd_fused") else 10) * F1),
        "legendre_forward_vector_kernel": ("mfma", 15 * F1),
        "fft_backward": ("hbm", 8.0 * 8 * nline * (nslot + cfg_d["nphi"])),
        "fft_forward": ("hbm", 8.0 * 9 * nline * (nslot + cfg_d["nphi"])),
    }
    t = res["timers"]
    avg = {s: t[s][0] / max(t[s][1], 1) for s in stages}
    dom = max(stages, key=lambda s: avg[s])
    bound, work = stages[dom]
    avg_s = avg[dom] / 1e3
    if bound == "mfma":
        if split:
            work = work * tri
            achieved, peak, unit = work / avg_s / 1e12, INT8_PEAK_TOPS, "TOP/s (INT8 slice products)"
        else:
            achieved, peak, unit = work / avg_s / 1e12, FP64_PEAK_TFLOPS, "TFLOP/s"
    else:
        achieved, peak, unit = work / avg_s / 1e9, HBM_PEAK_GBS, "GB/s"
    traffic = hbm_traffic(precision, kname[dom]) if cfg_d == CONFIGS["cfg2"] else None
    per_stage = {s: {"ms_per_launch": t[s][0] / max(t[s][1], 1), "launches": t[s][1]}
                 for s in t}
    for s, (b, w) in stages.items():
        a = avg[s] / 1e3
        if a > 0:
            per_stage[s]["kernel"] = kname[s]
            per_stage[s]["achieved"] = (w / a / 1e12) if b == "mfma" else (w / a / 1e9)
            per_stage[s]["unit"] = "TFLOP/s (FP64-equivalent)" if b == "mfma" else "GB/s"
    return {"kernel": kname[dom], "bound": bound, "achieved": achieved, "peak": peak,
            "unit": unit, "frac": achieved / peak, "traffic": traffic,
            "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, "
                            "copy-calibrated)",
            "work_per_launch": work, "avg_ms_per_launch": avg[dom]}, per_stage


def sample_radii(cfg_d, nr_sample):
    """The radius sample of the CPU baseline (both boundaries included)."""
    nr_s = min(nr_sample, cfg_d["nr"])
    return np.linspace(0, cfg_d["nr"] - 1, nr_s).round().astype(int)


def gather_sample(cfg, ks, sp_b, grid_f, results):
    """Host copies, on the radius sample ks, of the benchmark inputs and of
    each engine's outputs (the oracle's layouts at len(ks) radii)."""
    def grid(v, nc):
        return v.view(nc, cfg.nphi, cfg.nth, cfg.nr)[..., ks].contiguous().cpu().numpy().ravel()

    def spec(s, nc):
        return s.view(cfg.nr, cfg.jmax, nc)[ks].contiguous().cpu().numpy().ravel()

    out = {p: (grid(r["outputs"][0], 8), spec(r["outputs"][1], 9)) for p, r in results.items()}
    return {"ks": ks, "sp_b": spec(sp_b, 8), "grid_f": grid(grid_f, 9), "outputs": out}


def oracle_errors(ref_b, ref_f, got_b, got_f):
    """Per-field relative L2 and max errors of an engine against the oracle."""
    def err(a, b):
        return {"rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                "rel_max": float(np.abs(a - b).max() / np.abs(b).max())}
    gb, rb = got_b.reshape(8, -1), ref_b.reshape(8, -1)
    gf, rf = got_f.reshape(-1, 9), ref_f.reshape(-1, 9)
    bwd = {"vector1": err(gb[0:3], rb[0:3]), "vector2": err(gb[3:6], rb[3:6]),
           "scalar1": err(gb[6], rb[6]), "scalar2": err(gb[7], rb[7])}
    fwd = {f"vector{i + 1}": err(gf[:, 3 * i:3 * i + 3], rf[:, 3 * i:3 * i + 3]) for i in range(3)}
    return {"backward": bwd, "forward": fwd}


def cpu_baseline(cfg_d, nr_sample, sample=None):
    """The oracle timed on a radius sample of the workload; with `sample`
    (the benchmark's own inputs and the engines' outputs on those radii) it
    also reports each engine's error against the oracle there."""
    import oracle as O
    from kemorin_mhd_amd import chebyshev_radius
    ks = sample_radii(cfg_d, nr_sample)
    nr_s = len(ks)
    r = chebyshev_radius(cfg_d["nr"], 0.35 / 0.65, 1 / 0.65)[ks]
    P = O.OraclePlan(cfg_d["L"], nr_s, cfg_d["nth"], cfg_d["nphi"], r, full_tables=False)
    if sample is not None:
        sp_b, grid_f = sample["sp_b"], sample["grid_f"]
    else:
        rng = np.random.default_rng(1)
        sp_b = rng.standard_normal(P.n_rj(8))
        grid_f = rng.standard_normal(P.n_rtp(9))
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        ref_b = P.backward(2, 2, sp_b)
        ref_f = P.forward(3, 0, grid_f)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    full = t * cfg_d["nr"] / nr_s
    errors = None
    if sample is not None:
        errors = {p: oracle_errors(ref_b, ref_f, *o) for p, o in sample["outputs"].items()}
    return errors, {"value": 1.0 / full, "unit": "round_trips/s", "cores": O.lib().orc_num_threads(),
            "kind": "port",
            "sample": (f"L={cfg_d['L']} Ntheta={cfg_d['nth']} Nphi={cfg_d['nphi']}, {nr_s} of "
                       f"{cfg_d['nr']} radii, median of 3 round trips ({t:.2f} s each), "
                       f"scaled linearly to {cfg_d['nr']} radii; the oracle's restatement of "
                       f"symmetric_BLAS_big with unblocked triple loops in place of DGEMM "
                       f"(contraction off), FFTPACK-convention FFT, FP64, OpenMP over radii")}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU (the transform has no coupling "
                          "across radius: run one process per GPU on its share of the radii) "
                          "-- run with --gpus 1"}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    if args.copy_calib > 0:
        n = int(args.copy_calib * (1 << 30)) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes_each_way": 8 * n, "launches": 5}))
        return
    if args.shells > 0:
        run_shells(args, torch)
        return
    if args.nonlinear:
        run_nonlinear(args, torch)
        return
    cfg_d = CONFIGS[args.config]
    from kemorin_mhd_amd import sht as S
    split_ok = True
    if args.precision in ("auto", "both", "split"):
        try:
            S.lib()
            from kemorin_mhd_amd import SHT_PREC_SPLIT_INT8, ShellConfig, SphTransPlan
            probe = SphTransPlan(ShellConfig(7, 2, 8, 16), max_ncomp=9,
                                 precision=SHT_PREC_SPLIT_INT8)
            probe.close()
        except Exception:
            split_ok = False
    precs = []
    if args.precision == "fp64" or args.precision in ("auto", "both"):
        precs.append("fp64")
    if split_ok and args.precision in ("auto", "both", "split"):
        precs.append("split")
    cfg = make_config(cfg_d)
    otf, alias = cfg_d.get("otf", False), cfg_d.get("alias", False)
    sp_b, grid_f = make_inputs(cfg, torch, otf=otf)
    results = {}
    # the split-vs-FP64 comparison keeps both engines' outputs (not with aliasing)
    keep = "split" in precs and "fp64" in precs and not alias
    for p in precs:
        results[p] = run_engine(cfg, p, args, torch, sp_b, grid_f, keep=keep, otf=otf,
                                alias=alias)
        print(f"# {p}: {results[p]['ms_per_step']:.3f} ms/round trip", file=sys.stderr, flush=True)
    accuracy = None
    sample = None
    if keep and not args.no_cpu:
        sample = gather_sample(cfg, sample_radii(cfg_d, args.cpu_radii), sp_b, grid_f, results)
    if keep:
        del sp_b, grid_f
        accuracy = split_accuracy(cfg, args, torch, results)
        print(f"# split accuracy: {json.dumps(accuracy)}", file=sys.stderr, flush=True)
    head = "split" if "split" in results else "fp64"
    res = results[head]
    roof, per_stage = kernel_roofline(cfg_d, res, head, res["slices"])
    F_rt = 27 * flops_per_contraction(cfg_d)
    out = {
        "metric": METRIC if args.config == "cfg2" else
        (f"SHT round trips/s at L={cfg_d['L']},Nr={cfg_d['nr']}: split-prec vs FP64 MFMA"),
        "value": res["round_trips_per_s"], "unit": "round_trips/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if head == "fp64" else "f64 (int8-sliced Legendre)",
        "data": "synthetic (seeded Gaussian spectra, sigma=(l+1)^-1.5 sin(pi(r-ri)/(ro-ri)))",
        "config": {"workload": (f"round trip L={cfg_d['L']} Nr={cfg_d['nr']} "
                                f"Ntheta={cfg_d['nth']} Nphi={cfg_d['nphi']}: backward 2 vector + "
                                f"2 scalar (8 comps), forward 3 vector (9 comps)"),
                   "engine": head, "L": cfg_d["L"], "nr": cfg_d["nr"], "ntheta": cfg_d["nth"],
                   "nphi": cfg_d["nphi"], "variants": args.variant,
                   "p_on_the_fly": otf, "grid_buffers_aliased": alias},
        "roofline": roof,
        "gflops_fp64_equivalent": F_rt / (res["ms_per_step"] / 1e3) / 1e9,
        "stages": per_stage,
        "engines": {p: {"ms_per_step": r["ms_per_step"], "round_trips_per_s": r["round_trips_per_s"],
                        "stage_ms_per_step": {s: v[0] / args.steps for s, v in r["timers"].items()},
                        "plan_device_bytes": r["plan_bytes"],
                        "hbm_in_use_bytes": r["hbm_in_use_bytes"],
                        "hbm_total_bytes": r["hbm_total_bytes"],
                        "autotune": r["autotune"]}
                    for p, r in results.items()},
    }
    if "split" in results and "fp64" in results:
        out["speedup_split_vs_fp64"] = results["fp64"]["ms_per_step"] / results["split"]["ms_per_step"]
        out["split_error_vs_fp64"] = accuracy
    if not args.no_cpu:
        errors, out["cpu_baseline"] = cpu_baseline(cfg_d, args.cpu_radii, sample)
        if errors:
            # each engine against the oracle on the CPU baseline's radius sample
            out["error_vs_oracle"] = {"radii": [int(k) for k in sample["ks"]], **errors}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
