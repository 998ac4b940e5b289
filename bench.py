#!/usr/bin/env python
"""Force-evaluation benchmark of the mapped flare potential on one MI355X.

Workload (default, --workload 32k): 32,000-atom 2-species fcc (20^3
conventional cells, a = 4.05 A, seeded Gaussian displacements sigma = 0.08 A,
50/50 species by seeded shuffle), B2 {2 species, n_max 8, l_max 4, rc 5.0 A}
-> 680 features, per species symmetric beta 680 x 680 with A ~ N(0, 1e-3).
--workload 262k: 262,144-atom 3-species disordered box (random sequential
addition at 0.06 atoms/A^3, d_min 2.2 A), B2 {3, 8, 6, rc 5.5 A} -> 2,100
features, beta 2,100 x 2,100 per species, entries of std 1e-4.  One step =
one force evaluation (energy, forces, virial) from positions resident in
HBM, including the neighbor-list build.

Forms timed: plain FP64 (drop-in, materialised dA/dr), adjoint FP64, adjoint
FP32, adjoint + mixed precision (the headline `value`).  Prints one JSON line.

The default run adds, beside the headline figures:
- cpu_baseline: the oracle's restatement of PairFLARE::compute on the host
  cores, in a child process with OMP_PROC_BIND=close and OMP_PLACES=cores
  (one thread per physical core, and one per usable logical CPU where those
  differ), plus the plain form's error against that oracle result;
- extras: SparseGP predict_mean (32,000 atoms, 500 sparse) and
  predict_local_uncertainties (55,296 atoms, 2,000 sparse) in every form, and
  the 262,144-atom scale-up box in every form, each with per-stage rooflines
  (HBM GB/s or MFMA TFLOP/s against the MI355X peaks).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FORMS = ["plain", "adjoint_fp64", "adjoint_fp32", "adjoint_mixed"]

# name -> (generator in flare_pp_amd.synth, S, N, L, rc, description)
WORKLOADS = {
    "32k": ("config2", 2, 8, 4, 5.0,
            "mapped pair_flare force eval, 32,000-atom 2-species fcc, B2 {2,8,4} "
            "rc 5.0 (F=680), beta 680^2/species"),
    "262k": ("config5", 3, 8, 6, 5.5,
             "mapped pair_flare force eval, 262,144-atom 3-species RSA liquid "
             "(0.06/A^3, d_min 2.2 A), B2 {3,8,6} rc 5.5 (F=2100), beta "
             "2100^2/species"),
}


def load_workload(name, path=None):
    """Structure and packed beta of a workload (from `path` when given, as
    written by save_workload)."""
    import numpy as np
    if path:
        z = np.load(path)
        return {k: z[k] for k in z.files}
    from flare_pp_amd import pack_beta, synth
    c = getattr(synth, WORKLOADS[name][0])()
    c["packed"] = np.stack([pack_beta(b) for b in c["beta"]])
    return c


def save_workload(c, path):
    import numpy as np
    np.savez(path, cell=c["cell"], species=c["species"], positions=c["positions"],
             beta=c["beta"], packed=c["packed"])
HBM_PEAK_GBS = 8000.0
PEAKS_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--forms", default=",".join(FORMS))
    ap.add_argument("--workload", default="32k", choices=sorted(WORKLOADS),
                    help="32k: the headline configuration; 262k: the scale-up box")
    ap.add_argument("--workload-file", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes")
    ap.add_argument("--copy-calibration", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--traffic-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the SparseGP and 262k entries of the default run")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-out", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--gp", action="store_true",
                    help="time SparseGP predict_mean (32,000 atoms, 500 sparse) and "
                         "predict_local_uncertainties (55,296 atoms, 2,000 sparse) "
                         "instead of the mapped force evaluation")
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


# single-kernel stages that can be the roofline entry, and the kernel each
# one launches in a given form (name prefix in the rocprof kernel trace)
KERNELS = {
    ("forward", "adjoint_mixed"): "b2_forward_mfma_kernel<float",
    ("forward", "adjoint_fp32"): "b2_forward_mfma_kernel<float",
    ("forward", "adjoint_fp64"): "b2_forward_mfma_kernel<double",
    ("forward", "plain"): "b2_forward_mfma_kernel<double",
    ("gemm", "adjoint_mixed"): "gemm_x3w_kernel",
    ("forces_adjoint", "adjoint_mixed"): "force_kernel<float, 1,",
    ("forces_adjoint", "adjoint_fp32"): "force_kernel<float, 1,",
    ("forces_adjoint", "adjoint_fp64"): "force_kernel<double, 1,",
    ("forces_plain", "plain"): "force_kernel<double, 0,",
    ("pair_dA", "plain"): "pair_dA_kernel",
}


def algorithmic(stage, form, n, P, S, F, W, kp, gp=None):
    """Minimum work of one launch of a stage's kernel: ("hbm", bytes) or
    ("mfma", flops, precision).  gp: per-species (atoms, sparse) counts of a
    SparseGP prediction, whose forward writes d rows (FP32 forms) or B rows
    (FP64) and whose force kernel reads w rows instead of d and Z."""
    fb = 8 if form in ("plain", "adjoint_fp64") else 4
    if stage == "gemm":
        prec = {"adjoint_mixed": "bf16", "adjoint_fp32": "fp32"}.get(form, "fp64")
        return ("mfma", 2.0 * F * F * n, prec)
    if stage in ("gp_K", "gp_G"):  # t = D P^T, G = C P: 2 F M_s per atom
        prec = {"adjoint_mixed": "fp16", "adjoint_fp32": "fp32"}.get(form, "fp64")
        return ("mfma", sum(2.0 * F * m * na for na, m in gp), prec)
    if stage == "gp_Q":  # L^-1 k on the block-diagonal factor: M_s^2 per atom
        return ("mfma", sum(1.0 * m * m * na for na, m in gp), "fp64")
    if stage == "forward":
        if gp is not None or form != "adjoint_mixed":
            rows = F * fb
        else:
            rows = 2 * kp * 2
        return ("hbm", P * 36.0 + n * (16.0 + W * fb + rows + 8))
    if stage == "forces_adjoint":
        if gp is not None:
            return ("hbm", n * (W * fb + F * fb + 36.0) + P * 60.0)
        x = 2 * kp * 2 if form == "adjoint_mixed" else F * fb
        return ("hbm", n * (W * fb + x + F * fb + 36.0) + P * 60.0)
    if stage == "pair_dA":
        return ("hbm", P * (3.0 * (W // S) * 8 + 40))
    if stage == "forces_plain":
        return ("hbm", P * (3.0 * (W // S) * 8 + 60) + n * (W + F) * 8.0)
    return None


def stage_rooflines(form, stages, n, P, S, F, W, kp, gp=None):
    """Achieved rate of every stage with an algorithmic model: GB/s of
    algorithmic bytes or TFLOP/s of algorithmic flops, and its fraction of
    the MI355X peak (per-launch average from HIP events)."""
    out = {}
    for name, (t, cnt) in stages.items():
        a = algorithmic(name, form, n, P, S, F, W, kp, gp)
        if a is None or not cnt:
            continue
        sec = t / cnt * 1e-3
        if a[0] == "mfma":
            ach = a[1] / sec / 1e12
            peak = PEAKS_TFLOPS[a[2]]
            out[name] = {"bound": "mfma", "achieved": ach, "peak": peak,
                         "unit": "TFLOP/s", "frac": ach / peak, "precision": a[2]}
        else:
            ach = a[1] / sec / 1e9
            out[name] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": ach / HBM_PEAK_GBS}
    return out


def roofline(form, stages, n, P, S, F, W, kp):
    """Dominant single-kernel stage of the form and its achieved rate."""
    per = {k: t / c for k, (t, c) in stages.items() if c and (k, form) in KERNELS}
    name = max(per, key=per.get)
    sec = per[name] * 1e-3
    a = algorithmic(name, form, n, P, S, F, W, kp)
    if a[0] == "mfma":
        ach = a[1] / sec / 1e12
        peak = PEAKS_TFLOPS[a[2]]
        rl = {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
              "frac": ach / peak, "traffic": None}
    else:
        ach = a[1] / sec / 1e9
        rl = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
              "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None}
    return name, per[name], rl


# ---------------------------------------------------------------------------
# HBM traffic of one kernel launch from rocprofv3 counters (FETCH_SIZE and
# WRITE_SIZE, KiB per dispatch, one counter per pass), each scaled by what
# the same counter reads for a device copy of known size.

COPY_BYTES = 256 * 1024 * 1024


def pmc_pass(counter, child_args, match):
    """Mean counter value (KiB) per dispatch of kernels whose name contains
    `match`, from one rocprofv3 pass over a child run of this script."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    with tempfile.TemporaryDirectory() as d:
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o",
               "run", "--", sys.executable, os.path.abspath(__file__)] + child_args
        subprocess.run(["timeout", "-s", "KILL", "90"] + cmd, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        vals = []
        for root, _, files in os.walk(d):
            for f in files:
                if f.endswith("counter_collection.csv"):
                    for r in csv.DictReader(open(os.path.join(root, f))):
                        if match in r["Kernel_Name"] and r["Counter_Name"] == counter:
                            vals.append(float(r["Counter_Value"]))
    if not vals:
        raise RuntimeError(f"no {counter} samples for {match}")
    return sum(vals) / len(vals)


def measure_traffic(form, kernel, workload, wfile):
    scale = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        kib = pmc_pass(ctr, ["--copy-calibration"], "copyBuffer")
        scale[ctr] = COPY_BYTES / kib  # bytes per counter unit
    child = ["--traffic-child", form, "--workload", workload, "--workload-file", wfile]
    fetch = pmc_pass("FETCH_SIZE", child, kernel)
    write = pmc_pass("WRITE_SIZE", child, kernel)
    return {"bytes": fetch * scale["FETCH_SIZE"] + write * scale["WRITE_SIZE"],
            "fetch_bytes": fetch * scale["FETCH_SIZE"],
            "write_bytes": write * scale["WRITE_SIZE"],
            "fetch_scale_bytes_per_unit": scale["FETCH_SIZE"],
            "write_scale_bytes_per_unit": scale["WRITE_SIZE"]}


def copy_calibration():
    import torch
    x = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()


def make_pair(name, c):
    import numpy as np
    from flare_pp_amd import PairFLARE
    _, S, N, L, rc, _ = WORKLOADS[name]
    return PairFLARE.from_coefficients(2, S, N, L, np.full((S, S), rc), c["packed"])


def traffic_child(form, workload, wfile):
    import torch
    c = load_workload(workload, wfile)
    n = len(c["species"])
    pair = make_pair(workload, c)
    pos = torch.tensor(c["positions"], device="cuda:0")
    spc = torch.tensor(c["species"], device="cuda:0", dtype=torch.int32)
    e = torch.empty(1, dtype=torch.float64, device="cuda:0")
    f = torch.empty((n, 3), dtype=torch.float64, device="cuda:0")
    v = torch.empty(6, dtype=torch.float64, device="cuda:0")
    for _ in range(3):
        pair.build_neighbors(c["cell"], pos, spc)
        pair.compute_device(spc, form=form, energy=e, forces=f, virial=v)
    torch.cuda.synchronize()


def cpu_topology():
    """Usable logical CPUs of this process and the physical cores among them
    (distinct package/core pairs in sysfs)."""
    try:
        usable = sorted(os.sched_getaffinity(0))
    except AttributeError:
        usable = list(range(os.cpu_count() or 1))
    cores = set()
    for cpu in usable:
        base = f"/sys/devices/system/cpu/cpu{cpu}/topology/"
        try:
            with open(base + "physical_package_id") as a, open(base + "core_id") as b:
                cores.add((a.read().strip(), b.read().strip()))
        except OSError:
            cores.add(("cpu", cpu))
    return usable, len(cores)


def host_cpu():
    """CPU model line, family/model/stepping, logical CPUs, the CPUs this
    process may use and the physical cores among them."""
    info = {}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if ":" not in line:
                    if info:
                        break
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                if k in ("model name", "vendor_id", "cpu family", "model", "stepping",
                         "cpu MHz", "cache size") and k not in info:
                    info[k] = v
    except OSError:
        pass
    usable, phys = cpu_topology()
    return {"model": info.get("model name"),
            "vendor_family_model_stepping": [info.get("vendor_id"), info.get("cpu family"),
                                             info.get("model"), info.get("stepping")],
            "cpu_mhz": info.get("cpu MHz"), "cache": info.get("cache size"),
            "logical_cpus": os.cpu_count(), "usable_cpus": len(usable),
            "physical_cores_usable": phys}


def cpu_child(workload, wfile, out):
    """Child process of cpu_baseline: times the oracle's PairFLARE::compute
    restatement with the OpenMP settings of its environment and writes the
    timings (JSON) and one full result (npz) to `out`."""
    import numpy as np
    from oracle import pyoracle
    _, S, N, L, rc, _ = WORKLOADS[workload]
    c = load_workload(workload, wfile)
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    nl = pyoracle.NeighborList(c["cell"], c["species"], c["positions"], rc,
                               nthreads=nthreads)
    args = (nl, c["species"], S, N, L, 2, np.full((S, S), rc), c["beta"])
    e, f, v, ea = pyoracle.pair_compute(*args, packed=False, nthreads=nthreads,
                                        want_eatom=True)  # also the warm-up
    times = []
    t_end = time.time() + 20.0
    while len(times) < 5 and (time.time() < t_end or len(times) < 2):
        t0 = time.perf_counter()
        pyoracle.pair_compute(*args, packed=False, nthreads=nthreads)
        times.append(time.perf_counter() - t0)
    res = {"threads": nthreads, "median_s": float(np.median(times)), "evals": len(times)}
    if os.environ.get("FLARE_BENCH_SINGLE") == "1":
        # serial figure (pair_flare runs one thread per MPI rank)
        t1 = []
        t_end = time.time() + 12.0
        while len(t1) < 3 and (time.time() < t_end or not t1):
            t0 = time.perf_counter()
            pyoracle.pair_compute(*args, packed=False, nthreads=1)
            t1.append(time.perf_counter() - t0)
        res["single_median_s"] = float(np.median(t1))
        res["single_evals"] = len(t1)
    np.savez(out + ".npz", e=np.array([e]), f=f, v=v, ea=ea)
    with open(out + ".json", "w") as fh:
        json.dump(res, fh)


def cpu_baseline(name, c, wfile):
    """The oracle restatement of PairFLARE::compute on the host cores, timed
    in child processes started with OMP_PROC_BIND=close, OMP_PLACES=cores
    and OMP_NUM_THREADS = the physical cores (and, where hyperthreads add
    usable CPUs, once more with one thread per usable CPU and
    OMP_PLACES=threads).  The faster run is the baseline.  Returns the
    baseline and the oracle's result on the benchmark input."""
    import subprocess
    import tempfile
    import numpy as np
    usable, phys = cpu_topology()
    runs = [(phys, "cores")]
    if len(usable) > phys:
        runs.append((len(usable), "threads"))
    got = []
    ref = None
    with tempfile.TemporaryDirectory() as d:
        for k, (nt, places) in enumerate(runs):
            out = os.path.join(d, f"cpu{k}")
            env = dict(os.environ, OMP_NUM_THREADS=str(nt), OMP_PROC_BIND="close",
                       OMP_PLACES=places, FLARE_BENCH_SINGLE="1" if k == 0 else "0")
            subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child",
                            name, "--workload-file", wfile, "--cpu-out", out],
                           check=True, env=env, timeout=600)
            with open(out + ".json") as fh:
                r = json.load(fh)
            r["omp_places"] = places
            got.append(r)
            if ref is None:
                z = np.load(out + ".npz")
                ref = {k2: z[k2] for k2 in z.files}
    n = len(c["species"])
    best = min(got, key=lambda r: r["median_s"])
    base = {"value": n / best["median_s"], "unit": "atoms/s", "cores": best["threads"],
            "kind": "port",
            "sample": f"full {n:,}-atom force evaluation (E, F, virial), median of "
                      f"{best['evals']} after 1 warm-up, neighbor list prebuilt, FP64, "
                      f"OpenMP over atoms, OMP_PROC_BIND=close OMP_PLACES="
                      f"{best['omp_places']}",
            "runs": [{"threads": r["threads"], "omp_places": r["omp_places"],
                      "atoms_per_s": n / r["median_s"]} for r in got],
            "single_thread": {"value": n / got[0]["single_median_s"], "unit": "atoms/s",
                              "sample": f"median of {got[0]['single_evals']} full "
                                        "evaluations on 1 thread"},
            "host": host_cpu()}
    return base, ref


def gp_model(c, cutoff):
    """Structure + SparseGP of a GP config (the sparse set and alpha it names)."""
    from flare_pp_amd import B2, NormalizedDotProduct, SparseGP, Structure
    b2 = B2("chebyshev", "quadratic", [0, cutoff], [], c["settings"])
    s = Structure(c["cell"], c["species"], c["positions"], cutoff, [b2])
    gp = SparseGP([NormalizedDotProduct(c["sigma"], c["power"])], 0.1, 0.1, 0.1)
    gp.add_specific_environments(s, c["sparse_atoms"])
    gp.set_alpha(c["alpha"])
    return s, gp


def gp_time(gp, s, form, variances, steps, warmup):
    """Milliseconds per predict call (C-ABI on device buffers, HIP events;
    the neighbor list is the structure's)."""
    import ctypes
    import torch
    from flare_pp_amd import _capi
    from flare_pp_amd.sparse_gp import _form
    gp._check_ready(s)
    ctx = _capi.context(0)
    p = s.descriptor_calculators[0].params()
    out = torch.empty(1 + 3 * s.noa + 6, dtype=torch.float64, device="cuda:0")
    var = torch.empty(max(s.noa, 1), dtype=torch.float64, device="cuda:0")
    lib = _capi.lib()

    def call():
        if variances:
            _capi.check(lib.flare_sgp_predict_local_uncertainties(
                ctx.handle, gp._h, ctypes.byref(p), s._nlist, _capi.dptr(s.d_species),
                _form(form), _capi.dptr(out), _capi.dptr(var)))
        else:
            _capi.check(lib.flare_sgp_predict_mean(
                ctx.handle, gp._h, ctypes.byref(p), s._nlist, _capi.dptr(s.d_species),
                _form(form), _capi.dptr(out)))
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        call()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    ctx.profile(True)
    for _ in range(2):
        call()
    ctx.synchronize()
    stages = ctx.stage_times()
    ctx.profile(False)
    return ms, stages


def gp_entry(label, c, cutoff, variances, steps, warmup):
    """One SparseGP call in every form: ms, atoms/s, per-stage ms (summed
    over species) and rooflines, the dominant modelled stage of each form."""
    import ctypes
    from flare_pp_amd import _capi
    s, gp = gp_model(c, cutoff)
    S, N, L = c["settings"]
    nr = S * N
    F, W = nr * (nr + 1) // 2 * (L + 1), nr * (L + 1) ** 2
    view = _capi.NListView()
    _capi.check(_capi.lib().flare_nlist_get_view(s._nlist, ctypes.byref(view)))
    P = int(view.n_pairs)
    n = s.noa
    na = list(s.species_count[:S])
    ms_by = list(gp.sparse_descriptors[0].n_clusters_by_type)
    gpc = list(zip(na, ms_by))
    out = {"workload": label, "n_atoms": n, "n_pairs": P, "n_sparse_by_species": ms_by,
           "forms": {}}
    for form in FORMS:
        ms, stages = gp_time(gp, s, form, variances, steps, warmup)
        # stage totals per call (a stage launched once per species sums)
        per_call = {k: (t / 2, 1) for k, (t, cnt) in stages.items()}
        rls = stage_rooflines(form, per_call, n, P, S, F, W, 0, gpc)
        dom = max(rls, key=lambda k: per_call[k][0]) if rls else None
        out["forms"][form] = {
            "ms": ms, "atoms_per_s": n / (ms * 1e-3),
            "stage_ms": {k: v[0] for k, v in per_call.items()},
            "stage_rooflines": rls,
            "roofline": dict(rls[dom], stage=dom, stage_ms=per_call[dom][0]) if dom else None}
        print(f"# {label} {form}: {ms:.3f} ms", file=sys.stderr, flush=True)
    return out


def gp_cpu_baseline(threads):
    """The oracle's predict_mean and cluster variances (the reference's
    per-pair dB/dr formulation) on a 4,000-atom sample of config 3's
    material: fcc a = 4.05, 2 species, B2 {2, 8, 4}, 500 sparse atoms."""
    import numpy as np
    from oracle import pyoracle as orc
    from flare_pp_amd import synth
    nthreads = threads or cpu_topology()[1]
    cell, species, pos = synth.fcc(10, 4.05, 0.08, [2000, 2000], 3)
    rng = np.random.default_rng(3)
    atoms = np.sort(rng.choice(len(species), size=500, replace=False)).tolist()
    alpha = rng.normal(0.0, 1.0, 500)
    sigma, power, rc, S, N, L = 2.0, 2, 5.0, 2, 8, 4
    nl = orc.NeighborList(cell, species, pos, rc, nthreads=nthreads)
    t0 = time.perf_counter()
    o = orc.B2Result(nl, species, S, N, L, np.full((S, S), rc), dervs=True)
    g = o.grouped()
    off = np.concatenate([[0], np.cumsum(g["n_by_type"])[:-1]])
    sel = [[k for k, a in enumerate(g["atom_indices"][o0:o0 + cnt]) if a in set(atoms)]
           for o0, cnt in zip(off, g["n_by_type"])]
    # sparse rows: the selected atoms' descriptor rows, species by species
    rows = np.concatenate([o0 + np.asarray(k, dtype=int) for o0, k in zip(off, sel)])
    P, Pn, m = g["descriptors"][rows], g["norms"][rows], [len(k) for k in sel]
    orc.predict_mean(o, m, P, Pn, sigma, power, orc.volume(cell), alpha)
    t_mean = time.perf_counter() - t0
    Li = orc.cholesky_inverse(orc.envs_envs(S, m, P, Pn, m, P, Pn, sigma, power), 1e-8)
    t1 = time.perf_counter()
    orc.cluster_uncertainties(o, m, P, Pn, sigma, power, Li)
    t_var = time.perf_counter() - t1
    n = len(species)
    return {"predict_mean": {"value": n / t_mean, "unit": "atoms/s"},
            "variances": {"value": n / t_var, "unit": "atoms/s"},
            "cores": nthreads, "kind": "port",
            "sample": f"{n:,}-atom fcc sample, B2 {{2,8,4}} rc 5, 500 sparse atoms: one "
                      f"predict_mean (descriptors with per-pair dB/dr, as the "
                      f"reference) and one cluster-variance pass, FP64, OpenMP",
            "host": host_cpu()}


def gp_main(args):
    from flare_pp_amd import synth
    steps = max(1, min(args.steps, 10))
    c3 = synth.config3()
    mean = gp_entry("predict_mean: 32,000-atom 2-species fcc, B2 {2,8,4} rc 5, 500 "
                    "sparse, NormalizedDotProduct(2, 2)", c3, 5.0, False, steps, 1)
    c4 = synth.config4()
    var = gp_entry("predict_local_uncertainties: 55,296-atom 3-species fcc, B2 {3,8,3} "
                   "rc 4.5, 2,000 sparse", c4, c4["cutoff"], True, steps, 1)
    line = {"metric": "SparseGP predict_mean atoms/s at 32k atoms (500 sparse), mixed form",
            "value": mean["forms"]["adjoint_mixed"]["atoms_per_s"], "unit": "atoms/s",
            "n_gpus": 1, "steps": steps, "warmup": 1, "higher_is_better": True,
            "data": "synthetic (seeded structures, sparse sets and alpha)",
            "config": {"workload": mean["workload"] + "; " + var["workload"],
                       "includes_neighbor_build": False},
            "predict_mean": mean, "predict_local_uncertainties": var}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = gp_cpu_baseline(args.cpu_threads)
    print(json.dumps(line))


def run_mapped(wl, c, forms, steps, warmup):
    """Time the mapped force evaluation of workload `wl` in `forms`: per form
    ms per step (neighbor build included, and with a reused list), stage ms
    from HIP events on the library stream, errors against the plain FP64
    form, stage rooflines and the dominant modelled stage.  Returns the
    per-form results, the plain form's E/F/virial/eatom, n and n_pairs."""
    import ctypes
    import numpy as np
    import torch
    from flare_pp_amd import _capi
    _, S, N, L, rc, _ = WORKLOADS[wl]
    n = len(c["species"])
    pair = make_pair(wl, c)
    dev = "cuda:0"
    pos = torch.tensor(c["positions"], device=dev)
    spc = torch.tensor(c["species"], device=dev, dtype=torch.int32)
    cell = c["cell"]
    e = torch.empty(1, dtype=torch.float64, device=dev)
    f = torch.empty((n, 3), dtype=torch.float64, device=dev)
    v = torch.empty(6, dtype=torch.float64, device=dev)
    ea = torch.empty(n, dtype=torch.float64, device=dev)
    ctx = _capi.context(0)
    nr = S * N
    F, W = nr * (nr + 1) // 2 * (L + 1), nr * (L + 1) ** 2
    KP = (F + 31) // 32 * 32  # BF16 row length of the GEMM operands

    # reference result: plain FP64
    pair.build_neighbors(cell, pos, spc)
    pair.compute_device(spc, form="plain", energy=e, forces=f, virial=v, eatom=ea)
    ref = {"e": e.item(), "f": f.cpu().numpy(), "v": v.cpu().numpy(),
           "ea": ea.cpu().numpy()}
    f_ref = ref["f"]
    view = _capi.NListView()
    _capi.check(_capi.lib().flare_nlist_get_view(pair._nlist, ctypes.byref(view)))
    n_pairs = int(view.n_pairs)

    results = {}
    ctx.set_blocking(False)
    for form in forms:
        for _ in range(warmup):
            pair.build_neighbors(cell, pos, spc)
            pair.compute_device(spc, form=form, energy=e, forces=f, virial=v)
        ms = timed_steps(pair, ctx, cell, pos, spc, form, (e, f, v), steps, True)
        ms_reuse = timed_steps(pair, ctx, cell, pos, spc, form, (e, f, v), steps, False)
        ctx.profile(True)
        for _ in range(3):
            pair.build_neighbors(cell, pos, spc)
            pair.compute_device(spc, form=form, energy=e, forces=f, virial=v)
        ctx.synchronize()
        stages = ctx.stage_times()
        ctx.profile(False)
        # the algorithmic models count one step's work; a stage that runs in
        # several launches per step (the plain form's bounded atom ranges)
        # is rated on its per-step total
        per_step = {k: (t / 3, 1) for k, (t, cnt) in stages.items()}
        fd = f.cpu().numpy() - f_ref
        results[form] = {
            "ms_per_step": ms, "atoms_per_s": n / (ms * 1e-3),
            "ms_per_step_reused_nlist": ms_reuse,
            "atoms_per_s_reused_nlist": n / (ms_reuse * 1e-3),
            "stage_ms": {k: t / cnt for k, (t, cnt) in stages.items()},
            "stage_ms_per_step": {k: t / 3 for k, (t, cnt) in stages.items()},
            "err_energy_meV_per_atom": abs(e.item() - ref["e"]) / n * 1e3,
            "err_force_rms_meV_per_A": float(np.sqrt(np.mean(fd ** 2)) * 1e3),
            "err_force_max_meV_per_A": float(np.abs(fd).max() * 1e3),
            # the synthetic beta gives small forces: the error relative to
            # the RMS force is the scale-free figure
            "err_force_rms_rel": float(np.sqrt(np.mean(fd ** 2)) /
                                       max(np.sqrt(np.mean(f_ref ** 2)), 1e-300)),
            "force_rms_meV_per_A": float(np.sqrt(np.mean(f_ref ** 2)) * 1e3),
            "stage_rooflines": stage_rooflines(form, per_step, n, n_pairs, S, F, W, KP),
        }
        results[form]["roofline"] = roofline(form, per_step, n, n_pairs, S, F, W, KP)
        print(f"# {wl} {form}: {ms:.3f} ms/step ({ms_reuse:.3f} reused nlist) "
              f"stages per step {results[form]['stage_ms_per_step']}",
              file=sys.stderr, flush=True)
    ctx.set_blocking(True)
    del pair
    return results, ref, n, n_pairs


def parity_vs_oracle(ref, orc):
    """The plain FP64 form against the oracle on the benchmark input:
    relative errors (energy over sum |E_i|, max-norm of forces, virial and
    per-atom energies)."""
    import numpy as np

    def nrel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    return {"energy_rel": float(abs(ref["e"] - orc["e"][0]) / np.abs(orc["ea"]).sum()),
            "force_max_rel": nrel(ref["f"], orc["f"]),
            "virial_max_rel": nrel(ref["v"], orc["v"]),
            "eatom_max_rel": nrel(ref["ea"], orc["ea"]),
            "tolerance": 1e-10}


def main():
    args = parse()
    if args.copy_calibration:
        return copy_calibration()
    if args.traffic_child:
        return traffic_child(args.traffic_child, args.workload, args.workload_file)
    if args.cpu_child:
        return cpu_child(args.cpu_child, args.workload_file, args.cpu_out)
    if args.gpus != 1:
        print(json.dumps({"error": "the force evaluation runs on one GPU; it "
                          "is not sharded across GPUs (--gpus 1 only)"}))
        sys.exit(2)
    if args.gp:
        return gp_main(args)
    import tempfile

    forms = [f for f in args.forms.split(",") if f]
    wl = args.workload
    wdesc = WORKLOADS[wl][5]
    c = load_workload(wl)
    results, ref, n, n_pairs = run_mapped(wl, c, forms, args.steps, args.warmup)

    head = "adjoint_mixed" if "adjoint_mixed" in results else forms[-1]
    r = results[head]
    rl_name, rl_ms, rl = r["roofline"]
    kname = KERNELS[(rl_name, head)]
    rl = dict(rl, stage=rl_name, kernel=kname, kernel_ms=rl_ms)
    cpu = None
    parity = None
    with tempfile.TemporaryDirectory() as d:
        wfile = os.path.join(d, "workload.npz")
        save_workload(c, wfile)
        if not args.no_traffic:
            try:
                t = measure_traffic(head, kname, wl, wfile)
                rl["traffic"] = t["bytes"]
                rl["traffic_detail"] = t
            except Exception as exc:  # profiler absent or refused: no figure
                print(f"# traffic not measured: {exc}", file=sys.stderr, flush=True)
        if not args.no_cpu_baseline:
            cpu, orc_ref = cpu_baseline(wl, c, wfile)
            parity = parity_vs_oracle(ref, orc_ref)
    extras = None
    if not args.no_extras and wl == "32k":
        extras = extras_entries(args)
    plain = results.get("plain")
    line = {
        "metric": f"force-eval atoms/s at {wl} atoms: adjoint+mixed-prec vs plain "
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
        "data": "synthetic (seeded structure and random symmetric beta)",
        "config": {"workload": wdesc, "n_atoms": n, "n_pairs": n_pairs,
                   "form": head, "includes_neighbor_build": True},
        "speedup_vs_plain_fp64": (r["atoms_per_s"] / plain["atoms_per_s"]
                                  if plain else None),
        "forms": {k: {kk: vv for kk, vv in x.items() if kk != "roofline"}
                  for k, x in results.items()},
        "roofline": rl,
        "cpu_baseline": cpu,
        "plain_vs_oracle": parity,
        "extras": extras,
    }
    print(json.dumps(line))


def extras_entries(args):
    """SparseGP predict_mean (config 3) and predict_local_uncertainties
    (config 4) in every form, and the 262,144-atom box in every form (5 steps
    after 1 warm-up), with per-stage rooflines."""
    import gc
    from flare_pp_amd import synth
    out = {}
    c3 = synth.config3()
    out["gp_predict_mean_32k"] = gp_entry(
        "SparseGP predict_mean: 32,000-atom 2-species fcc, B2 {2,8,4} rc 5, 500 "
        "sparse, NormalizedDotProduct(2, 2); descriptors and forces included, "
        "neighbor list prebuilt", c3, 5.0, False, 5, 1)
    del c3
    c4 = synth.config4()
    out["gp_predict_local_uncertainties_55k"] = gp_entry(
        "SparseGP predict_local_uncertainties: 55,296-atom 3-species fcc, B2 {3,8,3} "
        "rc 4.5, 2,000 sparse; mean, forces and variances, neighbor list prebuilt",
        c4, c4["cutoff"], True, 5, 1)
    del c4
    gc.collect()
    c5 = load_workload("262k")
    res, _, n5, p5 = run_mapped("262k", c5, FORMS, 5, 1)
    out["mapped_262k"] = {
        "workload": WORKLOADS["262k"][5], "n_atoms": n5, "n_pairs": p5,
        "forms": {k: {kk: (dict(zip(("stage", "stage_ms", "roofline"), vv))
                           if kk == "roofline" else vv)
                      for kk, vv in x.items()} for k, x in res.items()}}
    return out


if __name__ == "__main__":
    main()
