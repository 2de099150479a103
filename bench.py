"""Benchmark of the pair-counting hot path on one MI355X.

A step is one full correlation call (tpcf_shear_f32 / tpcf_kappa_*) over the
workload with the catalogues already resident in HBM: cell sort, segment and
chunk tables, the pair kernel, the patch-pair reduction and the copy of the
per-bin and per-patch-pair results to the host.  ``value`` is in-bin pairs per
second of wall-clock time over the timed steps.  ``--host-io`` instead passes
host arrays in and out, as a host driver calls the C-ABI (the copies over PCIe
are then inside every step); ``--marks`` adds the per-lens marks (into device
arrays, or host arrays with ``--host-io``).

Default workload: 20M kappa lenses x 20M shear sources on a 3000'x3000'
lognormal-clustered patch, 30 log bins 0.5'-250', 128 patches, FP32 geometry
with FP64 re-binning of pairs within 1e-5 of an edge and FP64 sums.

The roofline line is for the dominant kernel (the pair kernel): achieved
TFLOP/s = in-bin pairs x FLOPs/pair / pair-kernel time (HIP events on the
library stream), against the MI355X FP32 vector peak (FP64 vector peak for
FP64 inputs and the spherical geometry).  The CPU baseline is
the reference's own correlation code (kd-tree, FP32, OpenMP; built from its
sources into oracle/_ref/ by oracle/build_ref.sh), or where that is absent a C
port of the same algorithm, timed on a lens subsample against all sources.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_PAIR = {0: 33, 1: 17}     # shear-scalar, kappa-kappa (algorithmic count)
FP32_PEAK_TFLOPS = 157.3            # MI355X published FP32 vector peak
# Spherical geometry, the reference's per-pair algorithm (distance_sphere,
# angles_sphere, src/utils.f90:512-528, :618-639; src/tpcf.f90:313-345) with
# the per-object sines and cosines of latitude hoisted, counted as the flat
# path is (FMA 2; sqrt, rcp, log and each trigonometric function 1):
# shear: lon difference 1, cos and sin of it 2, cos d 4, acos + bin 4, sin d 1,
# 1/sin d 1, cos theta 2, sin theta 4 + 1, cos / sin 2phi 5, projections 6,
# weights and sums 11 = 42; kappa: lon difference 1, its cos 1, cos d 4,
# acos + bin 4, weights and sums 8 = 18.
FLOPS_PER_PAIR_SPH = {0: 42, 1: 18}
FP64_PEAK_TFLOPS = 78.6             # MI355X published FP64 vector peak
METRIC = "pairs/sec on 1 MI355X vs host-core ref; FP32+FP64-edge vs FP64 at ≤1e-6 rel"
# the sources the pair kernel's profile depends on (its code and the tables it walks)
KERNEL_SOURCES = ("2pcf_amd/csrc/pairs.hip", "2pcf_amd/csrc/prepare.hip", "2pcf_amd/csrc/internal.hpp")


def code_of(text: str) -> str:
    """C++ source with its comments and blank lines removed and each line's
    trailing blanks dropped: the part of a source file the compiler sees (the
    pair kernel's sources hold no "//" or "/*" inside string literals)."""
    import re
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return "\n".join(line.rstrip() for line in text.splitlines() if line.strip())


def kernel_source_hash() -> str:
    """sha256 (16 hex digits) over the code of the pair kernel's sources,
    comments excluded: a profile's traffic figure is reused only for the
    kernel it was measured on, and a change to a comment does not change it."""
    import hashlib
    h = hashlib.sha256()
    for p in KERNEL_SOURCES:
        with open(os.path.join(ROOT, p), encoding="utf-8") as f:
            h.update(code_of(f.read()).encode())
    return h.hexdigest()[:16]


def cpu_threads() -> int:
    """Host threads of the CPU baseline: the CPUs this process may run on,
    capped by OMP_NUM_THREADS where it is set (a one-GPU box grants its GPU a
    share of 16 host CPUs and sets OMP_NUM_THREADS to it)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(n, omp) if omp > 0 else n)


def cpu_info() -> dict:
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "cpus_allowed": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


DEVICE: dict = {}   # the GPU as rocminfo reports it, read before this process touches the GPU


def device_peaks() -> dict:
    """The GPU agent's compute units and maximum engine clock from rocminfo,
    and the vector peaks they give: 16 lanes per SIMD, 2 FLOPs per FMA; packed
    FP32 doubles the FP32 rate, and on gfx950 an FP64 FMA issues at the
    non-packed FP32 rate.  Called before the GPU is initialised (rocminfo runs
    as a child process)."""
    import re
    import subprocess
    try:
        txt = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60).stdout
    except (OSError, subprocess.SubprocessError):
        return {"source": "rocminfo unavailable"}
    for agent in re.split(r"^Agent \d+\s*$", txt, flags=re.M)[1:]:
        f = {}
        for line in agent.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                f.setdefault(k.strip(), v.strip())
        if f.get("Device Type") != "GPU":
            continue
        try:
            cus, simds = int(f["Compute Unit"]), int(f["SIMDs per CU"])
            mhz = int(f["Max Clock Freq. (MHz)"])
        except (KeyError, ValueError):
            return {"source": "rocminfo: GPU agent without compute unit, SIMD or clock fields"}
        fp32 = cus * simds * 16 * 2 * mhz * 1e6 / 1e12
        return {"source": "rocminfo", "name": f.get("Name"), "marketing_name": f.get("Marketing Name"),
                "compute_units": cus, "simds_per_cu": simds, "max_clock_mhz": mhz,
                "fp32_tflops": round(fp32, 2), "fp32_packed_tflops": round(2 * fp32, 2),
                "fp64_tflops": round(fp32, 2)}
    return {"source": "rocminfo: no GPU agent"}


def peak_note() -> dict:
    """Where `peak` comes from, beside the figures rocminfo gives."""
    return {"peak_source": "MI355X published vector peak (FP32 packed, or FP64)", "device": DEVICE}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, pairs_per_lens: float, budget_s: float, threads: int) -> dict:
    """Host-core rate of the reference algorithm on a random lens subsample
    against all sources.  Uses the reference's own tangential_shear_tpcf /
    scalar_kappa_tpcf (kd-tree, FP32, OpenMP; oracle/_ref/libref2pcf.so) when
    it has been built, else the C port of the same algorithm.  The kd-tree
    build over the sources is timed by a call with one lens and reported
    apart, so ``value`` is the pair-search-and-accumulate rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle   # checker family: the CPU baseline only

    rng = np.random.default_rng(12345)
    n_a = len(wl.lens["x"])
    spacing = 1
    use_ref = oracle.have_reference()

    def run(sub):
        t0 = time.perf_counter()
        if use_ref:
            r = oracle.reference(wl.kind, sub, wl.src, wl.x_min, wl.x_max, wl.n_bin, spacing, n_threads=threads,
                                 spherical=wl.spherical, quiet=True)
            r["tree_seconds"] = 0.0
        else:
            r = oracle.ref_port(wl.kind, sub, wl.src, wl.x_min, wl.x_max, wl.n_bin, spacing, n_threads=threads)
        return r, time.perf_counter() - t0

    def sample(n):
        sel = rng.choice(n_a, size=n, replace=False)
        return {k: v[sel] for k, v in wl.lens.items()}

    if use_ref:
        _, t_fixed = run(sample(1))          # kd-tree build over all sources + call overhead
    # calibrate on ~2e8 pairs, then size the sample for ~budget_s of pair work
    probe = min(n_a, max(8, int(2e8 / max(pairs_per_lens, 1.0))))
    r, dt = run(sample(probe))
    fixed = t_fixed if use_ref else r["tree_seconds"]
    rate = max(r["n_pairs"].sum(), 1) / max(dt - fixed, 1e-6)
    n_s = int(min(n_a, max(probe, budget_s * rate / max(pairs_per_lens, 1.0))))
    r, total = run(sample(n_s))
    fixed = t_fixed if use_ref else r["tree_seconds"]
    pair_time = total - fixed
    pairs = int(r["n_pairs"].sum())
    out = {"value": pairs / pair_time, "unit": "pairs/s", "cores": threads,
           "kind": "reference" if use_ref else "port",
           "sample": f"{n_s} random lenses of {n_a} against all {len(wl.src['x'])} sources; "
                     f"{pairs} in-bin pairs in {pair_time:.2f} s on {threads} threads; kd-tree build over the "
                     f"sources {fixed:.2f} s reported apart",
           "tree_build_s": fixed, "pairs": pairs, "seconds": pair_time,
           "code": ("reference tangential_shear_tpcf/scalar_kappa_tpcf (flang -O2 -fopenmp)" if use_ref
                    else "C port of the reference algorithm (gcc -O2 -fopenmp)")}
    out.update(cpu_info())
    return out


def _have_ref() -> bool:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle   # checker family: the CPU baseline only
    return oracle.have_reference()


def profiled_traffic(workload: str):
    """HBM bytes per pair-kernel launch measured for this workload by
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (copy-calibrated), as recorded in
    profiles/k_pairs_<workload>.md, when that profile was taken on the pair
    kernel built now (its kernel_source_hash line equals kernel_source_hash());
    else None."""
    path = os.path.join(ROOT, "profiles", f"k_pairs_{workload}.md")
    if not os.path.exists(path):
        return None, None
    traffic = src_hash = None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            traffic = float(line.split(":")[1])
        elif line.startswith("kernel_source_hash:"):
            src_hash = line.split(":")[1].strip()
    if traffic is None or src_hash != kernel_source_hash():
        return None, None
    return traffic, os.path.relpath(path, ROOT)


def blended_cpu_baseline(parts: list, budget_s: float, threads: int) -> dict:
    """CPU baseline of a mixed workload: each part (a workload-like namespace
    with the pair count the GPU step does for it) timed with cpu_baseline on
    budget_s / len(parts) seconds, combined as the rate at which the host would
    do the step's pairs: sum N / sum (N / rate)."""
    res, t, n = [], 0.0, 0
    for wl, n_pairs, pairs_per_lens in parts:
        r = cpu_baseline(wl, pairs_per_lens, budget_s / len(parts), threads)
        res.append(r)
        t += n_pairs / r["value"]
        n += n_pairs
    out = dict(res[0])
    out["value"] = n / t
    out["sample"] = "; ".join(f"{'shear-scalar' if wl.kind == 0 else 'kappa-kappa'} part ({wl.name}): "
                              f"{r['value']:.3g} pairs/s from {r['sample']}" for (wl, _, _), r in zip(parts, res))
    out["blend"] = "rate at which the host does the step's pairs: sum N / sum (N / rate of its kind)"
    for k in ("pairs", "seconds", "tree_build_s"):
        out.pop(k, None)
    return out


def host_part(name, kind, lens, src, x_min, x_max, n_bin):
    from types import SimpleNamespace
    return SimpleNamespace(name=name, kind=kind, lens=lens, src=src, x_min=x_min, x_max=x_max, n_bin=n_bin,
                           spherical=False)


def bench_mocks(args) -> int:
    """Batched form: 256 mock fields (synth.mock_fields: per field a kappa
    catalogue and a shear catalogue with sizes log-uniform in [5e4, 2e6] on a
    600'x600' patch), kappa-kappa (auto) and kappa x gamma_t per field, 20 log
    bins 1'-100', 16 patches (4 x 4 squares) per field; one tpcf_batch of 512
    catalogues (graph captured on the first run, replayed after).  value =
    in-bin pairs per second of one batch run.  The roofline counts 17 FLOPs per
    kappa-kappa pair and 33 per shear pair over the batch run's wall time
    (every kernel of the graph, on 8 parallel branches)."""
    import torch
    tp = importlib.import_module("2pcf_amd")
    synth = importlib.import_module("2pcf_amd.synth")
    n_f = int(256 * args.scale) or 1
    t0 = time.perf_counter()
    cats = synth.mock_fields(n_f)
    torch.cuda.synchronize()
    log(f"{len(cats)} catalogues generated on the device in {time.perf_counter() - t0:.1f} s")
    bins = tp.create_bins(1.0, 100.0, 20, tp.LOG)
    ctx = tp.Context(0)
    if args.cell_target > 0:
        ctx.set_option(3, args.cell_target)
    if args.split > 0:
        ctx.set_option(4, args.split)
    t0 = time.perf_counter()
    b = tp.Batch([synth.batch_desc(c) for c in cats], bins, autotune=not args.no_autotune, ctx=ctx)
    st = b.stats
    log(f"batch planned in {time.perf_counter() - t0:.1f} s (autotune {st['autotuned']}: cell target "
        f"{st['cell_target']}, split {st['split']}, candidates {[round(t, 1) for t in st['tune_ms']]} ms)")
    for i in range(args.warmup):
        out = b.run()
        log(f"warmup {i}: {b.stats['ms_run']:.1f} ms ({b.stats['graph_nodes']} graph nodes)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pairs, run_ms = 0, []
    kinds = [c["kind"] for c in cats]
    for i in range(args.steps):
        out = b.run()
        pairs += int(sum(int(r["n_pairs"].sum()) for r in out))
        run_ms.append(b.stats["ms_run"])
        log(f"step {i}: {b.stats['ms_run']:.1f} ms")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kk = int(sum(int(r["n_pairs"].sum()) for r, k in zip(out, kinds) if k == 1))
    sk = int(sum(int(r["n_pairs"].sum()) for r, k in zip(out, kinds) if k == 0))
    flops = kk * FLOPS_PER_PAIR[1] + sk * FLOPS_PER_PAIR[0]
    run_s = float(np.mean(run_ms)) * 1e-3
    achieved = flops / run_s / 1e12
    res = {"metric": "pairs/sec, batch of mock catalogues on 1 MI355X", "value": pairs / el, "unit": "pairs/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "mocks_256", "fields": n_f, "catalogues": len(cats),
                      "pairs_per_step": pairs / args.steps, "pairs_kappa_kappa": kk, "pairs_shear_scalar": sk,
                      "cell_target": st["cell_target"], "split": st["split"],
                      "graph_nodes": b.stats["graph_nodes"], "plan_ms": st["ms_plan"]},
           "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        **peak_note(),
                        "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": "k_pairs (graph replay)",
                        "time_base": "batch run wall time (all kernels of the graph)", "run_ms": run_s * 1e3,
                        "flops_per_pair": {"kappa-kappa": FLOPS_PER_PAIR[1], "shear-scalar": FLOPS_PER_PAIR[0]}}}
    if not args.no_cpu:
        # the largest catalogue of each kind, lenses subsampled, against all its sources
        parts = []
        for kind, n_k in ((1, kk), (0, sk)):
            i = max((j for j in range(len(cats)) if kinds[j] == kind),
                    key=lambda j: cats[j]["lens"]["x"].numel() * cats[j]["src"]["x"].numel())
            c = cats[i]
            lens = {k: v.cpu().numpy() for k, v in c["lens"].items()}
            src = {k: v.cpu().numpy() for k, v in c["src"].items()}
            if kind == 0:
                src["m"] = np.zeros(len(src["x"]), np.float32)
            ppl = float(out[i]["n_pairs"].sum()) / len(lens["x"])
            parts.append((host_part(f"mocks_256 catalogue {i}", kind, lens, src, 1.0, 100.0, 20), n_k, ppl))
        res["cpu_baseline"] = blended_cpu_baseline(parts, args.cpu_seconds, cpu_threads())
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    print(json.dumps(res), flush=True)
    b.close()
    ctx.close()
    return 0


def bench_adjoint(args) -> int:
    """Differentiable form: 5M kappa points x 5M sources uniform on a
    1200'x1200' patch, 20 log bins 1'-200': forward xi_kk and gamma_t plus the
    adjoint of chi^2 (diagonal C) with respect to every kappa, e1, e2 and w.
    value = in-bin pairs of the two forward correlations per second of
    forward + backward.  Each correlation's pairs are visited twice per step
    (the lens-side pass of the recorded forward and the source-side pass of
    the backward), so the roofline counts 2 x (17 x kappa-kappa pairs +
    33 x shear pairs) FLOPs over the step's wall time."""
    import torch
    tp = importlib.import_module("2pcf_amd")
    ag = importlib.import_module("2pcf_amd.autograd")
    synth = importlib.import_module("2pcf_amd.synth")
    dev = torch.device("cuda:0")
    n = int(5_000_000 * args.scale)
    side = 1200.0 * np.sqrt(args.scale)
    gen = torch.Generator(device=dev)
    gen.manual_seed(1005)

    def uni(k, lo, hi):
        return torch.rand(k, generator=gen, device=dev) * (hi - lo) + lo

    kx, ky = uni(n, 0, side), uni(n, 0, side)
    kap = (torch.randn(n, generator=gen, device=dev) * 0.02).requires_grad_()
    kw = uni(n, 0.5, 1.5).requires_grad_()
    kz = torch.zeros(n, device=dev)
    gx, gy = uni(n, 0, side), uni(n, 0, side)
    e1 = (torch.randn(n, generator=gen, device=dev) * 0.27).requires_grad_()
    e2 = (torch.randn(n, generator=gen, device=dev) * 0.27).requires_grad_()
    gw = uni(n, 0.5, 1.5).requires_grad_()
    gz, gm = torch.ones(n, device=dev), torch.zeros(n, device=dev)
    kb = torch.randn(n, generator=gen, device=dev) * 0.02
    bins = tp.create_bins(1.0, 200.0, 20, tp.LOG)
    ctx = tp.Context(0)
    if args.cell_target > 0:
        ctx.set_option(3, args.cell_target)

    def step():
        xi, _ = ag.kappa_kappa(dict(x=kx, y=ky, val=kap, w=kw), dict(x=kx, y=ky, kappa_E=kap, kappa_B=kb, w=kw),
                               bins, ctx=ctx)
        gt, _ = ag.shear_scalar(dict(x=kx, y=ky, z=kz, val=kap, w=kw),
                                dict(x=gx, y=gy, z=gz, e1=e1, e2=e2, w=gw, m=gm), bins, ctx=ctx)
        sig = torch.full_like(xi, 1e-4)
        loss = ag.chi2(xi, torch.zeros_like(xi), sig) + ag.chi2(gt, torch.zeros_like(gt), sig)
        for t in (kap, kw, e1, e2, gw):
            t.grad = None
        loss.backward()
        return loss

    for i in range(args.warmup):
        step()
    torch.cuda.synchronize()
    r = tp.correlate(tp.KAPPA, dict(x=kx.data_ptr(), y=ky.data_ptr(), val=kap.data_ptr(), w=kw.data_ptr()),
                     dict(x=kx.data_ptr(), y=ky.data_ptr(), kappa_E=kap.data_ptr(), kappa_B=kb.data_ptr(),
                          w=kw.data_ptr()), bins, device_inputs=True, ctx=ctx, n_A=n, n_B=n)
    r2 = tp.correlate(tp.SHEAR, dict(x=kx.data_ptr(), y=ky.data_ptr(), z=kz.data_ptr(), val=kap.data_ptr(),
                                     w=kw.data_ptr()),
                      dict(x=gx.data_ptr(), y=gy.data_ptr(), z=gz.data_ptr(), e1=e1.data_ptr(), e2=e2.data_ptr(),
                           w=gw.data_ptr(), m=gm.data_ptr()), bins, device_inputs=True, ctx=ctx, n_A=n, n_B=n)
    kk, sk = int(r.n_pairs.sum()), int(r2.n_pairs.sum())
    fwd_pairs = kk + sk
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"step {i} done")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    step_s = el / args.steps
    flops = 2 * (kk * FLOPS_PER_PAIR[1] + sk * FLOPS_PER_PAIR[0])
    achieved = flops / step_s / 1e12
    res = {"metric": "pairs/sec of forward + adjoint on 1 MI355X", "value": fwd_pairs * args.steps / el,
           "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": step_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "adjoint_5Mx5M", "n": n, "forward_pairs_per_step": fwd_pairs,
                      "pairs_kappa_kappa": kk, "pairs_shear_scalar": sk,
                      "passes_per_step": "per correlation: the autograd forward records the lens-side marks "
                                         "pass (its marks give the forward sums), the backward replays it with "
                                         "the source-side marks pass"},
           "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        **peak_note(),
                        "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": "k_pairs (4 marks passes)",
                        "time_base": "step wall time (forward, chi^2 and backward)",
                        "flops_per_step": flops,
                        "flops_per_pair": {"kappa-kappa": FLOPS_PER_PAIR[1], "shear-scalar": FLOPS_PER_PAIR[0],
                                           "passes_per_pair": 2}}}
    if not args.no_cpu:
        # the reference has no derivative: its forward correlations of the same
        # catalogues, blended like the step's pairs
        kx_h, ky_h = kx.detach().cpu().numpy(), ky.detach().cpu().numpy()
        lens = dict(x=kx_h, y=ky_h, val=kap.detach().cpu().numpy(), w=kw.detach().cpu().numpy(),
                    z=np.zeros(n, np.float32))
        ksrc = dict(x=kx_h, y=ky_h, kappa_E=lens["val"], kappa_B=kb.cpu().numpy(), w=lens["w"])
        gsrc = dict(x=gx.cpu().numpy(), y=gy.cpu().numpy(), z=np.ones(n, np.float32), e1=e1.detach().cpu().numpy(),
                    e2=e2.detach().cpu().numpy(), w=gw.detach().cpu().numpy(), m=np.zeros(n, np.float32))
        parts = [(host_part("adjoint_5Mx5M kappa-kappa", 1, lens, ksrc, 1.0, 200.0, 20), kk, kk / n),
                 (host_part("adjoint_5Mx5M shear-scalar", 0, lens, gsrc, 1.0, 200.0, 20), sk, sk / n)]
        res["cpu_baseline"] = blended_cpu_baseline(parts, args.cpu_seconds, cpu_threads())
        res["cpu_baseline"]["note"] = "forward correlations only: the reference has no gradient"
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    print(json.dumps(res), flush=True)
    ctx.close()
    return 0


def bench_kk100k(args) -> int:
    """The survey's first configuration: the kappa-kappa auto-correlation of
    100k points uniform on a 600'x600' patch (FP64, 20 log bins 1'-100', 16
    k-means patches) and 200 bootstrap resamples of the 16 patches with both
    covariances.  A step is the correlation call (host arrays in, as the
    reference's driver holds them) plus block sums and the bootstrap; value =
    in-bin ordered pairs per second of the step.  The roofline is the pair
    kernel's (17 FLOPs per pair, FP64 vector peak); the CPU baseline is the
    reference's correlation on the whole workload."""
    import torch
    tp = importlib.import_module("2pcf_amd")
    synth = importlib.import_module("2pcf_amd.synth")
    wl = synth.config("kk_100k_auto", args.scale)
    bins = tp.create_bins(wl.x_min, wl.x_max, wl.n_bin, tp.LOG)
    ctx = tp.Context(0)
    if args.cell_target > 0:
        ctx.set_option(3, args.cell_target)

    def step():
        g = tp.correlate(tp.KAPPA, wl.lens, wl.src, bins, precision="f64", patch_A=wl.patch_A, patch_B=wl.patch_B,
                         n_patch_A=wl.n_patch, n_patch_B=wl.n_patch, real_bytes=8, ctx=ctx)
        blocks = tp.block_sums(g, wl.patch_A, wl.lens["w"], wl.n_patch)
        bt = tp.bootstrap(blocks, 200, seed=1001, ctx=ctx)
        return g, bt

    for i in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms, pairs = [], 0
    for i in range(args.steps):
        g, bt = step()
        kernel_ms.append(g.stats["ms_pairs"])
        pairs += int(g.n_pairs.sum())
        log(f"step {i}: correlation {g.stats['ms_total']:.1f} ms (pairs kernel {g.stats['ms_pairs']:.2f} ms)")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    per_step = pairs / args.steps
    k_ms = float(np.mean(kernel_ms))
    achieved = per_step * FLOPS_PER_PAIR[1] / (k_ms * 1e-3) / 1e12
    res = {"metric": METRIC, "value": pairs / el, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "kk_100k_auto", "n": len(wl.lens["x"]), "n_bin": wl.n_bin,
                      "theta": [wl.x_min, wl.x_max], "patches": wl.n_patch, "bootstrap_resamples": 200,
                      "pairs_per_step": per_step, "io": "host arrays in and out"},
           "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        **peak_note(),
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_pairs",
                        "kernel_ms": k_ms, "flops_per_pair": FLOPS_PER_PAIR[1]}}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(wl, per_step / len(wl.lens["x"]), args.cpu_seconds, cpu_threads())
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    print(json.dumps(res), flush=True)
    ctx.close()
    return 0


def bench_patches(args) -> int:
    """Patch labels and bootstrap of the bootstrap chain: k-means with 128
    patches over the 20M lenses of sk_20Mx20M_lognormal_f32 (Voronoi form of
    partition_data, fixed starting grid, to convergence or 1000 rounds), then
    the block bootstrap of 128 blocks x 30 bins with 10^4 resamples and the
    covariances.  value = lens assignments per second of k-means (points x
    rounds / time); the reference's partition_data is timed beside it on a
    200k-lens subsample (one thread, as it runs)."""
    import torch
    tp = importlib.import_module("2pcf_amd")
    synth = importlib.import_module("2pcf_amd.synth")
    wl = synth.config("sk_20Mx20M_lognormal_f32", args.scale)
    x, y, w = wl.lens["x"], wl.lens["y"], wl.lens["w"]
    ctx = tp.Context(0)
    tp.kmeans(x[:100000], y[:100000], w[:100000], 128, max_iterations=2, ctx=ctx)   # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    km = tp.kmeans(x, y, w, 128, ctx=ctx)
    t_km = time.perf_counter() - t0
    log(f"k-means: {len(x)} lenses, 128 patches, {km['iterations']} rounds in {t_km:.2f} s")
    r = np.random.default_rng(9)
    blocks = dict(g_t=r.normal(0, 1e-3, (128, 30)), g_x=r.normal(0, 1e-3, (128, 30)), k=r.uniform(1, 2, (128, 30)),
                  weight=r.uniform(1, 2, 128))
    tp.bootstrap(blocks, 100, seed=1, ctx=ctx)
    t0 = time.perf_counter()
    bt = tp.bootstrap(blocks, 10000, seed=2, ctx=ctx)
    t_bt = time.perf_counter() - t0
    log(f"bootstrap: 10^4 resamples of 128 blocks x 30 bins + covariances in {t_bt * 1e3:.1f} ms")
    value = len(x) * km["iterations"] / t_km
    res = {"metric": "k-means lens assignments/sec on 1 MI355X (patch labels of the bootstrap)", "value": value,
           "unit": "assignments/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": t_km * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "patches_20M", "n_lens": len(x), "k": 128, "rounds": km["iterations"],
                      "bootstrap_ms": t_bt * 1e3, "bootstrap": "10^4 resamples x 128 blocks x 30 bins"}}
    if not args.no_cpu and _have_ref():
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle   # checker family: the CPU baseline only
        sel = r.choice(len(x), size=200000, replace=False)
        t0 = time.perf_counter()
        oracle.ref_kmeans(x[sel], y[sel], w[sel], 128, max_iterations=20)
        t_ref = time.perf_counter() - t0
        # the rounds the same algorithm runs on this sample (the reference does not return them)
        rounds = oracle.kmeans(x[sel], y[sel], w[sel], 128, max_iterations=20)["iterations"]
        res["cpu_baseline"] = {"value": 200000 * rounds / t_ref, "unit": "assignments/s", "cores": 1,
                               "kind": "reference",
                               "sample": f"reference partition_data, 200000 random lenses, 128 patches, {rounds} "
                                         f"rounds (max_iterations 20) in {t_ref:.2f} s"}
    print(json.dumps(res), flush=True)
    ctx.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="sk_20Mx20M_lognormal_f32")
    ap.add_argument("--scale", type=float, default=1.0, help="shrink sizes (constant density); 1 = full")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cell", type=float, default=0.0, help="source cell side (0: automatic)")
    ap.add_argument("--cell-target", type=float, default=0.0,
                    help="automatic cell size: mean sources per cell (0: the library's default, 128 for FP32 "
                         "geometry, 64 for FP64)")
    ap.add_argument("--counters", action="store_true",
                    help="after the timed steps, one more call with TPCF_OPT_COUNTERS: segment visits and examined "
                         "pairs per class (fast / edge / multi) in the line's `classes`")
    ap.add_argument("--split", type=float, default=0.0,
                    help="pair-kernel work-item band height in units of x_max (0: the library's default)")
    ap.add_argument("--no-autotune", action="store_true", help="mocks_256: skip the batch autotuner")
    ap.add_argument("--marks", action="store_true",
                    help="compute the per-lens marks as well (the reference's call always does)")
    ap.add_argument("--host-io", action="store_true",
                    help="host arrays in and out (the C-ABI as a host driver calls it: H2D of the catalogues "
                         "and D2H of every output inside each step)")
    ap.add_argument("--e2e-steps", type=int, default=2,
                    help="after the timed steps, time this many calls with host arrays (H2D inside the step) "
                         "and report them as end_to_end (0: skip)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed steps, copy 1 GiB device-to-device (reference kernel for HBM counters)")
    ap.add_argument("--source-hash", action="store_true",
                    help="print the pair kernel's source hash (recorded in the profiles) and exit")
    args = ap.parse_args()
    if args.source_hash:
        print(kernel_source_hash())
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: the pair-counting path runs on a single MI355X; "
              "sharding over several GPUs is not implemented", file=sys.stderr)
        return 2
    DEVICE.update(device_peaks())
    if args.config == "kk_100k_auto":
        return bench_kk100k(args)
    if args.config == "mocks_256":
        return bench_mocks(args)
    if args.config == "adjoint_5Mx5M":
        return bench_adjoint(args)
    if args.config == "patches_20M":
        return bench_patches(args)

    import torch

    tp = importlib.import_module("2pcf_amd")
    synth = importlib.import_module("2pcf_amd.synth")
    t0 = time.perf_counter()
    wl = synth.config(args.config, args.scale)
    log(f"workload {wl.name} generated in {time.perf_counter() - t0:.1f} s: "
        f"{len(wl.lens['x'])} lenses x {len(wl.src['x'])} sources")
    prec = "f64" if wl.lens["x"].dtype == np.float64 else "f32"
    dev = torch.device("cuda:0")
    dl = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in wl.lens.items()}
    ds = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in wl.src.items()}
    pa = torch.from_numpy(wl.patch_A).to(dev) if wl.patch_A is not None else None
    pb = torch.from_numpy(wl.patch_B).to(dev) if wl.patch_B is not None else None
    torch.cuda.synchronize()
    bins = tp.create_bins(wl.x_min, wl.x_max, wl.n_bin, tp.LOG)
    ctx = tp.Context(0)
    if args.cell > 0:
        ctx.set_option(1, args.cell)
    if args.cell_target > 0:
        ctx.set_option(3, args.cell_target)
    if args.split > 0:
        ctx.set_option(4, args.split)

    hl = {k: np.ascontiguousarray(v) for k, v in wl.lens.items()}
    hs = {k: np.ascontiguousarray(v) for k, v in wl.src.items()}

    def host_step():
        return tp.correlate(wl.kind, hl, hs, bins, precision=prec, patch_A=wl.patch_A, patch_B=wl.patch_B,
                            n_patch_A=wl.n_patch, n_patch_B=wl.n_patch, real_bytes=4, ctx=ctx,
                            spherical=wl.spherical, marks=args.marks)

    def device_step():
        return tp.correlate(wl.kind, {k: v.data_ptr() for k, v in dl.items()},
                            {k: v.data_ptr() for k, v in ds.items()}, bins, precision=prec,
                            patch_A=pa.data_ptr() if pa is not None else None,
                            patch_B=pb.data_ptr() if pb is not None else None,
                            n_patch_A=wl.n_patch, n_patch_B=wl.n_patch, real_bytes=4, device_inputs=True,
                            ctx=ctx, n_A=len(wl.lens["x"]), n_B=len(wl.src["x"]), spherical=wl.spherical,
                            marks=args.marks, marks_device=args.marks)

    step = host_step if args.host_io else device_step
    for i in range(args.warmup):
        r = step()
        log(f"warmup {i}: {r.stats['ms_total']:.1f} ms (pairs kernel {r.stats['ms_pairs']:.1f} ms)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms, pairs, stats = [], 0, None
    for i in range(args.steps):
        r = step()
        kernel_ms.append(r.stats["ms_pairs"])
        pairs += int(r.n_pairs.sum())
        stats = r.stats
        log(f"step {i}: {r.stats['ms_total']:.1f} ms (upload {r.stats['ms_upload']:.1f}, prepare "
            f"{r.stats['ms_prepare']:.1f}, pairs kernel {r.stats['ms_pairs']:.1f}, finish {r.stats['ms_finish']:.1f})")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = pairs / elapsed
    per_step_pairs = pairs / args.steps
    k_ms = float(np.mean(kernel_ms))
    fpp = (FLOPS_PER_PAIR_SPH if wl.spherical else FLOPS_PER_PAIR)[wl.kind]
    # FP64 arithmetic (FP64 inputs, or the sphere's FP64 geometry) is priced
    # against the FP64 vector peak
    peak = FP64_PEAK_TFLOPS if (wl.spherical or prec == "f64") else FP32_PEAK_TFLOPS
    achieved = per_step_pairs * fpp / (k_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if wl.spherical else prec, "data": "synthetic",
        "config": {"workload": wl.name, "n_lens": len(wl.lens["x"]), "n_source": len(wl.src["x"]),
                   "n_bin": wl.n_bin, "theta": [wl.x_min, wl.x_max], "patches": wl.n_patch,
                   "kind": "shear-scalar" if wl.kind == 0 else "kappa-kappa", "scale": args.scale,
                   "pairs_per_step": per_step_pairs, "cell_size": stats["cell_size"],
                   "spherical": wl.spherical, "marks": args.marks,
                   "io": "host arrays in and out" if args.host_io else "inputs resident in HBM",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", **peak_note(),
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "k_pairs", "kernel_ms": k_ms, "flops_per_pair": fpp,
                     "kernel_source_hash": kernel_source_hash()},
    }
    if args.scale == 1.0 and args.cell == 0.0 and args.cell_target == 0.0 and args.split == 0.0 and not args.marks:
        traffic, src = profiled_traffic(wl.name)
        if traffic is not None:
            out["roofline"]["traffic"] = traffic
            out["roofline"]["traffic_source"] = (f"{src} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; copy-calibrated; "
                                                 f"profiled on the kernel of this source hash)")
        else:
            out["roofline"]["traffic_source"] = "no profile of this workload on this pair kernel's sources"
    if args.counters:
        # an untimed call that counts the pair kernel's segment visits and
        # examined lens-source pairs per class
        ctx.set_option(2, 1)
        cs = step().stats
        ctx.set_option(2, 0)
        sv, pv = np.array(cs["seg_visits"], float), np.array(cs["pair_visits"], float)
        lanes = stats["n_lens_used"] / max(cs["n_chunks"], 1)
        out["classes"] = {"segment_visits": sv.tolist(), "examined_pairs": pv.tolist(),
                          "examined_share": (pv / pv.sum()).tolist(), "in_bin_share": per_step_pairs / pv.sum(),
                          "sources_per_visit": pv.sum() / sv.sum() / lanes, "lenses_per_chunk": lanes,
                          "order": "fast, edge, multi"}
        log(f"classes: visits {sv.tolist()}, examined pairs {np.round(100 * pv / pv.sum(), 1).tolist()} %")
    if not args.host_io and args.e2e_steps > 0:
        # the same call end to end through the C-ABI as a host driver makes it:
        # host arrays in (H2D of the catalogues) and per-bin outputs out; never
        # `value`, reported beside it
        host_step()
        t1 = time.perf_counter()
        e2e_pairs = 0
        for _ in range(args.e2e_steps):
            e2e_pairs += int(host_step().n_pairs.sum())
        e2e_s = (time.perf_counter() - t1) / args.e2e_steps
        out["end_to_end"] = {"io": "host arrays in and out (H2D of the inputs inside each step)",
                             "steps": args.e2e_steps, "ms_per_step": e2e_s * 1e3,
                             "pairs_per_s": e2e_pairs / args.e2e_steps / e2e_s}
        log(f"end to end with host arrays: {e2e_s * 1e3:.1f} ms per step")
    # the C port has no spherical mode: the spherical CPU baseline needs the reference build
    if not args.no_cpu and (not wl.spherical or _have_ref()):
        threads = cpu_threads()
        ppl = per_step_pairs / max(stats["n_lens_used"], 1)
        out["cpu_baseline"] = cpu_baseline(wl, ppl, args.cpu_seconds, threads)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    if args.calib_copy:
        # a plain copy of known size: 1 GiB read + 1 GiB written, for scaling FETCH_SIZE / WRITE_SIZE
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = 2 * a.numel() * 4
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
