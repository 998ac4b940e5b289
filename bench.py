#!/usr/bin/env python
"""Force-evaluation benchmark of the mapped flare potential on one MI355X.

Workload: 32,000-atom 2-species fcc (20^3 conventional cells, a = 4.05 A,
seeded Gaussian displacements sigma = 0.08 A, 50/50 species by seeded
shuffle), B2 {2 species, n_max 8, l_max 4, rc 5.0 A} -> 680 features, per
species symmetric beta 680 x 680 with A ~ N(0, 1e-3).  One step = one force
evaluation (energy, forces, virial) from positions resident in HBM,
including the neighbor-list build.

Forms timed: plain FP64 (drop-in, materialised dA/dr), adjoint FP64, adjoint
FP32, adjoint + mixed precision (the headline `value`).  Prints one JSON line.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FORMS = ["plain", "adjoint_fp64", "adjoint_fp32", "adjoint_mixed"]
HBM_PEAK_GBS = 8000.0
PEAKS_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--forms", default=",".join(FORMS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def timed_steps(pair, ctx, cell, pos, spc, form, outs, steps, rebuild):
    import torch
    e, f, v = outs
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(steps):
        if rebuild:
            pair.build_neighbors(cell, pos, spc)
        pair.compute_device(spc, form=form, energy=e, forces=f, virial=v)
    ev1.record()
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def roofline(form, stages, n_atoms, n_pairs, F, W):
    """Dominant stage of the timed form and its algorithmic rate."""
    per = {k: t / c for k, (t, c) in stages.items() if c}
    name = max(per, key=per.get)
    sec = per[name] * 1e-3
    fbytes = 8 if form in ("plain", "adjoint_fp64") else 4
    if name == "gemm":
        flops = 2.0 * F * F * n_atoms  # D . beta, 2F^2 per atom
        prec = {"plain": "fp64", "adjoint_fp64": "fp64", "adjoint_fp32": "fp32",
                "adjoint_mixed": "bf16"}[form]
        ach = flops / sec / 1e12
        peak = PEAKS_TFLOPS[prec]
        return name, per[name], {"bound": "mfma", "achieved": ach, "peak": peak,
                                 "unit": "TFLOP/s", "frac": ach / peak,
                                 "traffic": None}
    if name == "pair_dA":
        algo = n_pairs * 3 * (W // 2) * 8.0  # dA/dr written once, FP64
    elif name == "forces_plain":
        algo = n_pairs * 3 * (W // 2) * 8.0 + n_atoms * (W + F) * 8.0 + n_pairs * 24
    elif name == "forces_adjoint":
        algo = n_atoms * (W + F) * fbytes + n_pairs * (32 + 4 + 24) + n_atoms * 72
    elif name == "forward":
        algo = n_pairs * (32 + 4) + n_atoms * (W + F) * fbytes
    elif name == "nlist":
        algo = n_pairs * (8 + 32 + 4 * 4) + n_atoms * 24 * 2
    else:
        algo = n_atoms * F * fbytes * 2
    ach = algo / sec / 1e9
    return name, per[name], {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                             "traffic": None}


def cpu_baseline(c, packed, threads):
    """The oracle restatement of PairFLARE::compute on the host cores."""
    from oracle import pyoracle
    import numpy as np
    nthreads = threads or min(16, os.cpu_count() or 1)
    nl = pyoracle.NeighborList(c["cell"], c["species"], c["positions"], 5.0,
                               nthreads=nthreads)
    beta_full = c["beta"]
    args = (nl, c["species"], 2, 8, 4, 2, np.full((2, 2), 5.0), beta_full)
    pyoracle.pair_compute(*args, packed=False, nthreads=nthreads)  # warm-up
    times = []
    t_end = time.time() + 20.0
    while len(times) < 5 and (time.time() < t_end or len(times) < 2):
        t0 = time.perf_counter()
        pyoracle.pair_compute(*args, packed=False, nthreads=nthreads)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    n = len(c["species"])
    return {"value": n / med, "unit": "atoms/s", "cores": nthreads,
            "kind": "port",
            "sample": f"full 32,000-atom force evaluation (E, F, virial), "
                      f"median of {len(times)} after 1 warm-up, neighbor list "
                      f"prebuilt, FP64, OpenMP over atoms"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU "
                          "sharding is described in DESIGN.md, not run"}))
        sys.exit(2)
    import numpy as np
    import torch

    from flare_pp_amd import PairFLARE, _capi, pack_beta, synth

    forms = [f for f in args.forms.split(",") if f]
    c = synth.config2()
    n = len(c["species"])
    packed = np.stack([pack_beta(b) for b in c["beta"]])
    pair = PairFLARE.from_coefficients(2, 2, 8, 4, np.full((2, 2), 5.0), packed)
    dev = "cuda:0"
    pos = torch.tensor(c["positions"], device=dev)
    spc = torch.tensor(c["species"], device=dev, dtype=torch.int32)
    cell = c["cell"]
    e = torch.empty(1, dtype=torch.float64, device=dev)
    f = torch.empty((n, 3), dtype=torch.float64, device=dev)
    v = torch.empty(6, dtype=torch.float64, device=dev)
    ctx = _capi.context(0)
    F, W = 680, 400

    # reference result: plain FP64
    pair.build_neighbors(cell, pos, spc)
    pair.compute_device(spc, form="plain", energy=e, forces=f, virial=v)
    e_ref, f_ref = e.item(), f.cpu().numpy()
    from flare_pp_amd.structure import d2h  # noqa: F401
    view = _capi.NListView()
    import ctypes
    _capi.check(_capi.lib().flare_nlist_get_view(pair._nlist, ctypes.byref(view)))
    n_pairs = int(view.n_pairs)

    results = {}
    ctx.set_blocking(False)
    for form in forms:
        for _ in range(args.warmup):
            pair.build_neighbors(cell, pos, spc)
            pair.compute_device(spc, form=form, energy=e, forces=f, virial=v)
        ms = timed_steps(pair, ctx, cell, pos, spc, form, (e, f, v), args.steps, True)
        ms_reuse = timed_steps(pair, ctx, cell, pos, spc, form, (e, f, v),
                               args.steps, False)
        ctx.profile(True)
        for _ in range(3):
            pair.build_neighbors(cell, pos, spc)
            pair.compute_device(spc, form=form, energy=e, forces=f, virial=v)
        ctx.synchronize()
        stages = ctx.stage_times()
        ctx.profile(False)
        fd = f.cpu().numpy() - f_ref
        results[form] = {
            "ms_per_step": ms, "atoms_per_s": n / (ms * 1e-3),
            "ms_per_step_reused_nlist": ms_reuse,
            "atoms_per_s_reused_nlist": n / (ms_reuse * 1e-3),
            "stage_ms": {k: t / cnt for k, (t, cnt) in stages.items()},
            "err_energy_meV_per_atom": abs(e.item() - e_ref) / n * 1e3,
            "err_force_rms_meV_per_A": float(np.sqrt(np.mean(fd ** 2)) * 1e3),
            "err_force_max_meV_per_A": float(np.abs(fd).max() * 1e3),
        }
        results[form]["roofline"] = roofline(form, stages, n, n_pairs, F, W)
        print(f"# {form}: {ms:.3f} ms/step ({ms_reuse:.3f} reused nlist) "
              f"stages {results[form]['stage_ms']}", file=sys.stderr, flush=True)
    ctx.set_blocking(True)

    head = "adjoint_mixed" if "adjoint_mixed" in results else forms[-1]
    r = results[head]
    rl_name, rl_ms, rl = r["roofline"]
    rl = dict(rl, kernel=rl_name, kernel_ms=rl_ms)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(c, packed, args.cpu_threads)
    plain = results.get("plain")
    line = {
        "metric": "force-eval atoms/s at 32k atoms: adjoint+mixed-prec vs plain "
                  "FP64 on 1 GPU",
        "value": r["atoms_per_s"],
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3+fp32" if head == "adjoint_mixed" else head,
        "data": "synthetic (seeded fcc structure and random symmetric beta)",
        "config": {"workload": "mapped pair_flare force eval, 32,000-atom 2-species "
                               "fcc, B2 {2,8,4} rc 5.0 (F=680), beta 680^2/species",
                   "n_atoms": n, "n_pairs": n_pairs, "form": head,
                   "includes_neighbor_build": True},
        "speedup_vs_plain_fp64": (r["atoms_per_s"] / plain["atoms_per_s"]
                                  if plain else None),
        "forms": {k: {kk: vv for kk, vv in x.items() if kk != "roofline"}
                  for k, x in results.items()},
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
