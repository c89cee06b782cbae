"""Benchmark: MC photons propagated per second on one MI355X.

Workload (one step = one pp_propagate of the whole event): cascade_16_strings,
a 10 TeV electron cascade of 1e8 photons inside a 16-string, 320-module array
(4 x 4 strings at 50 m, 20 spheres of r = 0.3 m at 10 m), Cascadia water,
PMT-acceptance-weighted Cherenkov spectrum 300-700 nm, 15 steps. Inputs
(targets, spectrum, medium) are resident on the device before timing; hits
stay on the device (the host copy is not part of a step).

Prints one JSON line. `roofline` is the FP32 VALU roofline of the dominant
kernel (algorithmic FP32 operations per photon step of the reference step,
per sphere test and per grid-cell visit, over the kernel's measured HIP-event
time);
`cpu_baseline` is the oracle (C restatement of the reference kernel, OpenMP)
on a bounded sample of the same workload on the host's cores.

`surrogate` (skip with --no-surrogate): Monte Carlo labels for 2e5 (cascade,
module) samples, bf16 MFMA training of the 8-256x4-10 MLP, the batched
evaluation of 1e4 cascades x 1000 modules, and the full Monte Carlo judging
100 held-out events: pair evaluations/s of both, their ratio, and the mean
relative error of the expected hits on well-measured pairs.

`pmt_detection` (skip with --no-pmt): the POM's PMT-level detection kernel
(pp_pom_pmt_hits) on 2^24 synthetic hits, device time, hits/s and bytes moved.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "photonpropagation.jl_amd"))

# Algorithmic work per photon step of the reference kernel
# (src/photon_prop_cuda.jl:486-514), FMA = 2 ops:
#   step length, move, time, phase function mixture, scattering rotation
W_PHYS_FLOP = 47.0
W_PHYS_TRANSC = 7.4
#   one _intersect_nsphere test (src/photon_prop_cuda.jl:309-343)
W_ISEC_FLOP = 20.0
W_ISEC_TRANSC = 1.0
# Grid search (the uniform grid that replaces the reference's every-target
# scan): per step the segment's end points are compared with the grid box
# (6 min/max, 6 compares: 12 FP32 ops); a segment near the grid is walked:
# clipped to the box (3 reciprocals, 6 slab products, 12 min/max, entry cell
# and boundary parameters: 25 FP32 ops) and per visited cell the walk
# compares and advances (min3, 2 compares, 1 add: 4 FP32 ops).
W_GRID_STEP_FLOP = 12.0
W_GRID_WALK_FLOP = 25.0
W_GRID_WALK_TRANSC = 3.0
W_CELL_FLOP = 4.0
PEAK_FP32_TFLOPS = 157.3          # MI355X published FP32 vector peak
PEAK_FP32_NONPACKED_TFLOPS = 78.6  # without packed-FP32 instructions


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--photons", type=float, default=1e8)
    ap.add_argument("--search", choices=["auto", "brute", "grid"], default="auto")
    ap.add_argument("--cell", type=float, default=0.0, help="grid cell edge in m (0 = auto)")
    ap.add_argument("--cell-z", type=float, default=0.0,
                    help="grid cell edge along z in m (0 = cubic, < 0 = columns)")
    ap.add_argument("--block", type=int, default=0, help="threads per block (0 = library default)")
    ap.add_argument("--wps", type=int, default=0, help="waves per SIMD for grid sizing (0 = auto)")
    ap.add_argument("--workload", choices=["cascade", "uneven"], default="cascade",
                    help="cascade: 1e8-photon 10 TeV cascade in 320 modules; uneven: 256 "
                         "sources of 1e4-1e7 photons in 1000 modules")
    ap.add_argument("--batch", choices=["auto", "per-source", "persistent", "graph"],
                    default="auto", help="source scheduling (pp_batch_mode)")
    ap.add_argument("--chunk", type=int, default=0, help="photons per work chunk (0 = auto)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the autotuner (by default block size, chunk and grid cell are "
                         "tuned on a 2e7-photon sample first, untimed)")
    ap.add_argument("--no-surrogate", dest="surrogate", action="store_false",
                    help="skip the surrogate leg (MC labels, MFMA training, MC judge)")
    ap.add_argument("--sur-samples", type=int, default=200_000)
    ap.add_argument("--sur-steps", type=int, default=24_000)
    ap.add_argument("--sur-events", type=int, default=100)
    ap.add_argument("--cpu-sample", type=float, default=0.0,
                    help="photons for the CPU baseline (0 = calibrate to ~10 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", dest="legs", action="store_false",
                    help="skip the host-copy, README-example and batching legs")
    ap.add_argument("--surrogate-kernels", action="store_true",
                    help="only time the surrogate's kernels without hipGraph (for kernel "
                         "tracing): 3 batched evaluations of 1e4 cascades x 1000 modules and "
                         "50 eager training steps of batch 4096, then exit")
    ap.add_argument("--no-pmt", dest="pmt", action="store_false",
                    help="skip the PMT-detection leg (pp_pom_pmt_hits on 2^24 hits)")
    ap.add_argument("--pmt-kernels", action="store_true",
                    help="only run the PMT-detection leg (for kernel tracing), then exit")
    ap.add_argument("--sur-config", action="store_true",
                    help="only the surrogate's training configuration at its stated size "
                         "(2e6 samples of 1e7-photon labels, 1-300 m, batch 4096): labels, "
                         "training and the MC judge, one JSON record, then exit")
    ap.add_argument("--sur-config-first", type=int, default=0,
                    help="with --sur-config-save: first sample of the slice to label")
    ap.add_argument("--sur-config-count", type=int, default=0,
                    help="with --sur-config-save: samples in the slice (0 = to the end)")
    ap.add_argument("--sur-config-save", default="",
                    help="label only the slice and save its labels to this .npz, then exit")
    ap.add_argument("--sur-config-load", default="",
                    help="directory of .npz label slices that together cover every sample: "
                         "train and judge on them instead of labelling")
    ap.add_argument("--sur-config-steps", type=int, default=48_000)
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a plain 1 GiB device copy twice (reference kernel of known "
                         "size for the FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def calib_copy():
    import torch
    n = 1 << 28  # 1 GiB of float32
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(2):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calib_copy_bytes": a.numel() * 4, "copies": 2}), flush=True)


def total_steps(hits, n_ph_sim, n_steps):
    """Photon steps executed: every photon runs n_steps unless it hits earlier."""
    import numpy as np
    return int(n_ph_sim) * n_steps - int(np.sum(n_steps - hits["n_steps"].astype(np.int64)))


def cpu_model():
    """Host CPU model name (lscpu's "Model name", read from /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_run(wl, threads):
    """One run of the Float32 port of the reference kernel on workload wl:
    (seconds, photons, photon steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import photonprop as P
    spec = wl.spectrum
    kw = dict(n_steps=wl.n_steps, n_threads=threads, capacity=1 << 22)
    if isinstance(spec, P.Monochromatic):
        kw["mono_wl"] = spec.wavelength
    else:
        kw.update(table=spec.table, mapping=spec.mapping)
    t = time.perf_counter()
    _, nsim, steps = O.propagate_f32(P.processing.sources_to_pod(wl.sources), wl.medium.to_pod(),
                                     P.targets_to_pod(wl.targets), wl.seed, **kw)
    return time.perf_counter() - t, nsim, steps


def cpu_baseline(wl_factory, photons, what):
    """The Float32 port of the reference kernel (oracle/pp_ref_f32.c: Float32
    arithmetic, libm float functions, every-target scan, 15 steps; OpenMP over
    photons) on a bounded sample of the same workload, on all host threads:
    median of three runs after one warm-up."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    if photons <= 0:
        dt, _, _ = cpu_run(wl_factory(200_000), threads)
        photons = max(200_000, int(200_000 * 6.0 / max(dt, 1e-3)))
        photons = min(photons, 50_000_000)
    cpu_run(wl_factory(min(photons, 200_000)), threads)  # warm-up
    runs = sorted(cpu_run(wl_factory(photons), threads) for _ in range(3))
    dt, nsim, steps = runs[1]
    return dict(value=nsim / dt, unit="photons/s", cores=threads, kind="port",
                steps_per_s=steps / dt, model=cpu_model(),
                sample=f"median of 3 runs of {photons:.3g} photons of {what} with the Float32 "
                       f"port of the reference kernel (oracle/pp_ref_f32.c: every-target "
                       f"scan, 15 steps, OpenMP {threads} threads)")


def surrogate_kernels():
    import numpy as np

    import photonprop as P
    from photonprop import surrogate as S
    st = S.default_setup()
    targets = P.make_detector_grid(10, 10, 50.0, 10, 10.0, radius=0.3)
    mxyz = np.array([t.shape.position for t in targets], np.float32)
    events = S.array_events(np.random.default_rng(7), 10_000, mxyz, setup=st)
    rng = np.random.default_rng(3)
    x = rng.normal(0, 1, (4096, 8)).astype(np.float32)
    lab = np.zeros((4096, 66), np.float32)
    lab[:, 0] = 1e3
    lab[:, 1] = rng.poisson(50, 4096)
    lab[:, 2:] = rng.poisson(1.0, (4096, 64))
    with S.MLP(seed=1) as mlp:
        ev_ms = [mlp.eval_pairs(events, mxyz, st.domain())[3] for _ in range(3)]
        t0 = time.perf_counter()
        for _ in range(50):
            mlp.loss_grad(x, lab)
            mlp.adam_step(lr=1e-3)
        train_s = (time.perf_counter() - t0) / 50
    print(json.dumps({"eval_ms": ev_ms, "eager_train_step_ms": train_s * 1e3}), flush=True)


def pmt_leg():
    """PMT-level detection of the POM (pp_pom_pmt_hits, the step after the
    path): 2^24 synthetic hits on the module's sphere (isotropic unit
    positions, wavelengths 280-820 nm), device time of the detection kernel
    alone (HIP events; the host copies excluded), median of three runs after
    one warm-up. Bytes per hit: position 12 + wavelength 4 + photon id 8 +
    source 4 read, PMT id 4 written."""
    import numpy as np
    from photonprop import pom
    acc = pom.default_acceptance()
    rng = np.random.default_rng(5)
    n = 1 << 24
    rel = rng.normal(size=(n, 3))
    rel /= np.linalg.norm(rel, axis=1, keepdims=True)
    wl = rng.uniform(280.0, 820.0, n)
    pid = np.arange(n, dtype=np.uint64)
    src = np.zeros(n, np.int32)
    ms, det = [], 0
    for _ in range(4):
        t = {}
        pom._pmt_ids(rel, wl, acc, None, pid, src, 11, 0, timing=t)
        ms.append(t["kernel_ms"])
        det = t["n_detected"]
    k = sorted(ms[1:])[1]
    bytes_per_hit = 32
    gbps = n * bytes_per_hit / (k * 1e-3) / 1e9
    return {"hits": n, "kernel_ms": k, "hits_per_s": n / (k * 1e-3),
            "detected_frac": det / n, "bytes_per_hit": bytes_per_hit,
            "achieved_gbps": gbps, "hbm_frac": gbps / 8000.0,
            "note": "16 PMTs per hit: acos, a square root, exp and a reciprocal each, one "
                    "Philox block; vector-ALU bound, not HBM bound",
            "end_to_end": pmt_end_to_end()}


def pmt_end_to_end():
    """Propagation into POMs, then PMT detection of the hits on the device
    (pp_pom_pmt_hits_ctx reads the context's hit buffer; only the PMT ids come
    back): the cascade event with its 320 modules as POMs, 1e8 photons. Wall
    clock of pp_propagate alone and of pp_propagate + detection, median of
    three after one warm-up."""
    import ctypes

    import numpy as np

    import photonprop as P
    from photonprop import pom
    lib = P._lib.load_library()
    wl = P.workloads.cascade_16_strings()
    cfg = P._lib.pp_config()
    cfg.deterministic_order = 0
    _, dev = P.make_hit_buffers(capacity=1 << 24, config=cfg)
    pods = P.processing.sources_to_pod(wl.sources)
    nh, ns = ctypes.c_int64(), ctypes.c_int64()
    with dev:
        P.processing.stage(dev, wl.targets, wl.medium, wl.spectrum)
        prop, both, kern = [], [], []
        for _ in range(4):
            t0 = time.perf_counter()
            P._lib.check(lib.pp_propagate(dev.handle, pods, 1, ctypes.c_uint64(wl.seed), 15, 1.0,
                                          ctypes.byref(nh), ctypes.byref(ns)), "pp_propagate")
            t1 = time.perf_counter()
            dev.n_hits = nh.value
            t = {}
            ids = pom.pmt_ids_in_context(dev, seed=3, timing=t)
            t2 = time.perf_counter()
            prop.append(t1 - t0)
            both.append(t2 - t0)
            kern.append(t["kernel_ms"])
    med = lambda v: sorted(v[1:])[1]  # noqa: E731
    return {"hits": int(nh.value), "detected": int(np.sum(ids > 0)),
            "propagate_ms": med(prop) * 1e3, "propagate_plus_pmt_ms": med(both) * 1e3,
            "pmt_kernel_ms": med(kern), "photons_per_s_incl_pmt": ns.value / med(both)}


def copy_leg(lib, P, dev, pods, wl):
    """Host-inclusive rate: one pp_propagate plus pp_copy_hits of every column
    (the reference's call returns host columns, src/photon_prop_cuda.jl:
    951-953), for each deterministic_order: 0 buffer order, 1 ordered on the
    device, 2 ordered on the host. Wall clock of propagate + copy, median of
    three after one warm-up (which allocates the pinned staging)."""
    import ctypes

    import numpy as np
    f = P._lib.fptr
    cfg = P._lib.pp_config()
    P._lib.check(lib.pp_get_config(dev.handle, ctypes.byref(cfg)), "pp_get_config")
    saved = cfg.deterministic_order
    n_hits, n_sim = ctypes.c_int64(), ctypes.c_int64()
    out = {}
    for order in (0, 1, 2):
        cfg.deterministic_order = order
        P._lib.check(lib.pp_set_config(dev.handle, ctypes.byref(cfg)), "pp_set_config")
        walls, copies = [], []
        for _ in range(4):
            t0 = time.perf_counter()
            P._lib.check(lib.pp_propagate(dev.handle, pods, len(pods), ctypes.c_uint64(wl.seed),
                                          wl.n_steps, 1.0, ctypes.byref(n_hits),
                                          ctypes.byref(n_sim)), "pp_propagate")
            n = n_hits.value
            cols = [np.empty((n, 3), np.float32), np.empty((n, 3), np.float32),
                    np.empty((n, 3), np.float32), np.empty(n, np.float32),
                    np.empty(n, np.float32), np.empty(n, np.float32), np.empty(n, np.int64),
                    np.empty(n, np.int32), np.empty(n, np.uint64), np.empty(n, np.int32)]
            t1 = time.perf_counter()
            P._lib.check(lib.pp_copy_hits(dev.handle, 0, n, f(cols[0]), f(cols[1]), f(cols[2]),
                                          f(cols[3]), f(cols[4]), 0, f(cols[5]), f(cols[6]),
                                          f(cols[7]), f(cols[8]), f(cols[9])), "pp_copy_hits")
            t2 = time.perf_counter()
            walls.append(t2 - t0)
            copies.append(t2 - t1)
        w = sorted(walls[1:])[1]
        out[f"order_{order}"] = {"propagate_plus_copy_ms": w * 1e3,
                                 "copy_ms": sorted(copies[1:])[1] * 1e3, "hits": n,
                                 "photons_per_s": n_sim.value / w}
    cfg.deterministic_order = saved
    P._lib.check(lib.pp_set_config(dev.handle, ctypes.byref(cfg)), "pp_set_config")
    out["note"] = ("pp_propagate + pp_copy_hits of all 10 columns (position, direction, "
                   "initial direction, wavelength, time, distance, module, steps, photon id, "
                   "source) into host arrays; order 1 sorts on the device, 2 on the host")
    return out


def timed_propagations(lib, P, dev, pods, wl, warmup, steps):
    """(wall ms per propagation, kernel ms, launches, n_ph_sim)."""
    import ctypes

    import numpy as np
    n_hits, n_sim = ctypes.c_int64(), ctypes.c_int64()

    def step():
        P._lib.check(lib.pp_propagate(dev.handle, pods, len(pods), ctypes.c_uint64(wl.seed),
                                      wl.n_steps, 1.0, ctypes.byref(n_hits),
                                      ctypes.byref(n_sim)), "pp_propagate")
    for _ in range(warmup):
        step()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        kms.append(dev.last_timing()[0])
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return wall, float(np.mean(kms)), int(dev.last_timing()[1]), n_sim.value


def readme_leg(lib, P, with_cpu):
    """The reference README's example (one isotropic point source at the
    origin, 450 nm, one r = 0.3 m sphere at 10 m, Cascadia water): GPU at 1e5
    photons (the example's size, launch-bound) and 1e8 photons, and the Float32
    port of the reference kernel on the host at 1e5 photons."""
    import numpy as np
    out = {}
    for photons, reps in ((100_000, 50), (100_000_000, 5)):
        wl = P.workloads.readme_point_source(photons=photons)
        cfg = P._lib.pp_config()
        cfg.deterministic_order = 0
        _, dev = P.make_hit_buffers(capacity=1 << 22, config=cfg)
        with dev:
            P.processing.stage(dev, wl.targets, wl.medium, wl.spectrum)
            pods = P.processing.sources_to_pod(wl.sources)
            wall, kms, _, nsim = timed_propagations(lib, P, dev, pods, wl, 2, reps)
            hits, _ = P.run_photon_prop_no_local_cache(wl.sources, wl.targets, wl.medium,
                                                       wl.spectrum, wl.seed,
                                                       P.HostHitBuffer(np.float32), dev)
            st = total_steps(hits, nsim, wl.n_steps)
        out[f"gpu_{photons:.0e}"] = {"photons_per_s": nsim / (wall * 1e-3),
                                     "steps_per_s": st / (wall * 1e-3), "ms": wall,
                                     "kernel_ms": kms, "hits": int(len(hits["n_steps"]))}
    if with_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        wl = P.workloads.readme_point_source(photons=100_000)
        cpu_run(wl, threads)
        runs = sorted(cpu_run(wl, threads) for _ in range(3))
        dt, nsim, st = runs[1]
        out["cpu_1e+05"] = {"photons_per_s": nsim / dt, "steps_per_s": st / dt, "ms": dt * 1e3,
                            "cores": threads, "kind": "port"}
    return out


def batching_leg(lib, P):
    """The 256-source uneven batch (1000 modules): one launch per source (the
    reference's schedule), one persistent work-stealing launch, and the same
    as a hipGraph; then the autotuner's choice for the persistent launch."""
    wl = P.workloads.uneven_batch_1000()
    pods = P.processing.sources_to_pod(wl.sources)
    out = {"photons": int(sum(s.photons for s in wl.sources)), "sources": len(pods)}
    for mode, name in ((1, "per_source"), (2, "persistent"), (3, "graph")):
        cfg = P._lib.pp_config()
        cfg.batch_mode = mode
        cfg.deterministic_order = 0
        _, dev = P.make_hit_buffers(capacity=1 << 24, config=cfg)
        with dev:
            P.processing.stage(dev, wl.targets, wl.medium, wl.spectrum)
            wall, kms, launches, nsim = timed_propagations(lib, P, dev, pods, wl, 1, 5)
        out[name] = {"ms": wall, "kernel_ms": kms, "launches": launches,
                     "photons_per_s": nsim / (wall * 1e-3)}
    cfg = P._lib.pp_config()
    cfg.deterministic_order = 0
    _, dev = P.make_hit_buffers(capacity=1 << 24, config=cfg)
    with dev:
        P.processing.stage(dev, wl.targets, wl.medium, wl.spectrum)
        tuned = P.autotune.autotune(dev, wl.sources, wl.targets, wl.medium, wl.spectrum, wl.seed,
                                    n_steps=wl.n_steps, sample_photons=4e7, use_cache=False)
        wall, kms, _, nsim = timed_propagations(lib, P, dev, pods, wl, 1, 5)
    out["tuned"] = {"ms": wall, "kernel_ms": kms, "photons_per_s": nsim / (wall * 1e-3),
                    "config": tuned.get("config") if isinstance(tuned, dict) else None}
    return out


def pmt_kernels():
    print(json.dumps({"pmt_detection": pmt_leg()}), flush=True)


def surrogate_leg(args):
    """MC labels -> bf16 MFMA training -> batched pair evaluation (1e4 events x
    1000 modules) -> full-MC judge on held-out events."""
    import numpy as np

    import photonprop as P
    from photonprop import surrogate as S
    st = S.default_setup()
    _, dev = P.make_hit_buffers(capacity=1 << 26)
    with dev:
        c, lab, gs = S.generate_training_data(st, args.sur_samples, seed=1, batch_sources=8192,
                                              dev=dev, log=None)
        mlp, info = S.train_surrogate(st, c, lab, steps=args.sur_steps, batch=4096, lr=2e-3,
                                      lr_stages=5, log=None)
        with mlp:
            model = S.SurrogateModel(mlp, st)
            targets = P.make_detector_grid(10, 10, 50.0, 10, 10.0, radius=0.3)
            mxyz = np.array([t.shape.position for t in targets], np.float32)
            held_out = S.array_events(np.random.default_rng(99), args.sur_events, mxyz, setup=st)
            res, mc, _ = S.judge_events(model, held_out, targets, max_photons=1e9, dev=dev)
            events = S.array_events(np.random.default_rng(7), 10_000, mxyz, setup=st)
            ms = []
            for _ in range(3):
                _, _, fb, t = mlp.eval_pairs(events, mxyz, model.domain)
                ms.append(t)
            t = min(ms)
            n_pairs = events.shape[0] * len(targets)
            dims = mlp.dims
            flop_per_pair = 2.0 * sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1))
            sur_rate = n_pairs / (t * 1e-3)
            mc_rate = held_out.shape[0] * len(targets) / (mc["kernel_ms"] * 1e-3)
    return {
        "pair_evals_per_s": sur_rate,
        "mc_pair_evals_per_s": mc_rate,
        "speedup": sur_rate / mc_rate,
        "mean_rel_err": res["mean_rel_err"],
        "median_rel_err": res["median_rel_err"],
        "mean_signed_err": res["mean_signed_err"],
        "pairs_judged": res["pairs"],
        "mean_mc_stat_err": res["mean_mc_stat_err"],
        "time_ll_per_hit": {"surrogate": res.get("time_ll_surrogate"),
                            "mc_histogram": res.get("time_ll_mc_hist")},
        "strata": res["strata"],
        "judge": f"{args.sur_events} held-out cascades (1-1000 TeV) in the 1000-module array; "
                 "pairs with >= 2500 effective MC hits (<= 2% MC error); MC at min(yield, 1e9) "
                 "photons per event, rate = pairs / MC kernel time",
        "eval": {"pairs": n_pairs, "ms": t, "fallback_frac": float(fb.mean()),
                 "mfma_tflops": flop_per_pair * n_pairs / (t * 1e-3) / 1e12,
                 "peak_bf16_tflops": 2500.0},
        "train": {"samples": args.sur_samples, "label_photons": gs["photons"],
                  "label_photons_per_s": gs["photons"] / (gs["kernel_ms"] * 1e-3),
                  "steps": info["steps"], "batch": info["batch"],
                  "steps_per_s": info["steps_per_s"],
                  "train_flop_per_step": 3.0 * flop_per_pair * info["batch"]},
    }


def measured_counters(workload):
    """VALU utilisation figures of the propagation kernel from the committed
    rocprofv3 counter summary of this workload (lines `key: value` after the
    'Read by bench.py' marker), or None."""
    path = os.path.join(ROOT, "profiles", f"propagate_batch_kernel_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    out, on = {}, False
    for line in open(path):
        t = line.strip()
        if t.startswith("Read by bench.py"):
            on = True
            continue
        if on and ":" in t and not t.startswith("traffic_bytes_per_launch"):
            k, v = t.split(":", 1)
            try:
                out[k.strip()] = float(v)
            except ValueError:
                pass
    if not out:
        return None
    out["source"] = os.path.relpath(path, ROOT)
    return out


def measured_traffic(workload):
    """HBM bytes per launch of the propagation kernel: calibrated FETCH_SIZE +
    WRITE_SIZE from the committed rocprofv3 counter summary of this workload
    (None if there is none)."""
    path = os.path.join(ROOT, "profiles", f"propagate_batch_kernel_{workload}_pmc.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def surrogate_config(args):
    """The surrogate's training configuration at its stated size (see
    photonprop.surrogate.CONFIG_*): MC labels of CONFIG_SAMPLES cascades, each
    from CONFIG_PHOTONS photons (or the label slices of earlier runs, which
    are the same labels: same geometry, batches and random streams), Adam
    training at batch CONFIG_BATCH, and the full-MC judge on held-out events
    in the 1000-module array with the network's domain 5-300 m."""
    import glob

    import numpy as np

    import photonprop as P
    from photonprop import surrogate as S
    st = S.default_setup()
    n = S.CONFIG_SAMPLES
    photons = S.constant_photons(S.CONFIG_PHOTONS)
    t0 = time.perf_counter()
    _, dev = P.make_hit_buffers(capacity=1 << 26)
    with dev:
        if args.sur_config_save:
            count = args.sur_config_count or n - args.sur_config_first
            c, lab, gs = S.generate_training_data(
                st, n, seed=S.CONFIG_SEED, batch_sources=2048, photons_fn=photons, dev=dev,
                r_range=S.CONFIG_R_RANGE, first=args.sur_config_first, count=count,
                log=lambda m: print(m, flush=True))
            a, b = args.sur_config_first, args.sur_config_first + count
            S.save_label_slice(args.sur_config_save, a, lab[a:b], gs)
            print(json.dumps({"slice": [a, b], "label_photons": gs["photons"],
                              "label_kernel_ms": gs["kernel_ms"], "seconds": gs["seconds"],
                              "bytes": os.path.getsize(args.sur_config_save)}), flush=True)
            return
        if args.sur_config_load:
            # geometry and exposures from the seed; labels from the slices
            c, lab, gs = S.generate_training_data(
                st, n, seed=S.CONFIG_SEED, batch_sources=2048, photons_fn=photons, dev=dev,
                r_range=S.CONFIG_R_RANGE, first=n, count=0, log=None)
            covered, gs = S.load_label_slices(
                sorted(glob.glob(os.path.join(args.sur_config_load, "*.npz"))), lab)
            if not covered.all():
                raise SystemExit(f"label slices cover {covered.sum()} of {n} samples")
        else:
            c, lab, gs = S.generate_training_data(
                st, n, seed=S.CONFIG_SEED, batch_sources=2048, photons_fn=photons, dev=dev,
                r_range=S.CONFIG_R_RANGE, log=lambda m: print(m, flush=True))
        print(f"labels ready: {gs['photons']:.4g} photons, {time.perf_counter() - t0:.0f} s",
              flush=True)
        mlp, info = S.train_surrogate(st, c, lab, steps=args.sur_config_steps,
                                      batch=S.CONFIG_BATCH, lr=2e-3, lr_stages=5,
                                      log=lambda m: print(m, flush=True))
        with mlp:
            model = S.SurrogateModel(mlp, st, st.domain(max_dist=S.CONFIG_R_RANGE[1]))
            targets = P.make_detector_grid(10, 10, 50.0, 10, 10.0, radius=0.3)
            mxyz = np.array([t.shape.position for t in targets], np.float32)
            held_out = S.array_events(np.random.default_rng(99), args.sur_events, mxyz, setup=st)
            res, mc, _ = S.judge_events(model, held_out, targets, max_photons=1e9, dev=dev)
    rec = {"configuration": "surrogate training data and training at the stated size",
           "samples": n, "photons_per_label": S.CONFIG_PHOTONS,
           "distance_m": list(S.CONFIG_R_RANGE), "energy_gev": [100.0, 1e6],
           "batch": S.CONFIG_BATCH, "label_photons": gs["photons"],
           "label_kernel_ms": gs["kernel_ms"],
           "label_photons_per_s": gs["photons"] / max(gs["kernel_ms"] * 1e-3, 1e-9),
           "label_slices": gs.get("slices"),
           "train": {k: info[k] for k in ("steps", "batch", "steps_per_s", "samples",
                                          "val_loss", "train_seconds")},
           "judge": res, "seconds": time.perf_counter() - t0}
    print(json.dumps(rec, default=float), flush=True)


def main():
    args = parse()
    if args.sur_config:
        surrogate_config(args)
        return
    if args.calib_copy:
        calib_copy()
        return
    if args.surrogate_kernels:
        surrogate_kernels()
        return
    if args.pmt_kernels:
        pmt_kernels()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design only)"}))
        sys.exit(2)
    import ctypes

    import numpy as np

    import photonprop as P

    lib = P._lib.load_library()
    if lib.pp_device_count() < 1:
        raise SystemExit("no HIP device visible")
    nph = int(args.photons)
    if args.workload == "cascade":
        wl = P.workloads.cascade_16_strings(photons=nph)
    else:
        wl = P.workloads.uneven_batch_1000()
        nph = int(sum(s.photons for s in wl.sources))
    cfg = P._lib.pp_config()
    cfg.batch_mode = {"auto": 0, "per-source": 1, "persistent": 2, "graph": 3}[args.batch]
    cfg.chunk_photons = args.chunk
    cfg.search = {"auto": 0, "brute": 1, "grid": 2}[args.search]
    cfg.grid_cell = args.cell
    cfg.grid_cell_z = args.cell_z
    cfg.block_size = args.block
    cfg.waves_per_simd = args.wps
    cfg.deterministic_order = 0
    _, dev = P.make_hit_buffers(capacity=1 << 24, config=cfg)
    P.processing.stage(dev, wl.targets, wl.medium, wl.spectrum)
    tuned = None
    if args.autotune:
        tuned = P.autotune.autotune(dev, wl.sources, wl.targets, wl.medium, wl.spectrum,
                                    wl.seed, n_steps=wl.n_steps, sample_photons=2e7,
                                    use_cache=False)
    pods = P.processing.sources_to_pod(wl.sources)
    n_hits = ctypes.c_int64()
    n_sim = ctypes.c_int64()

    def step():
        rc = lib.pp_propagate(dev.handle, pods, len(pods), ctypes.c_uint64(wl.seed), wl.n_steps, 1.0,
                              ctypes.byref(n_hits), ctypes.byref(n_sim))
        P._lib.check(rc, "pp_propagate")

    for _ in range(args.warmup):
        step()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        kernel_ms.append(dev.last_timing()[0])
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    photons_per_s = n_sim.value / (ms_per_step / 1e3)

    hbc = P.HostHitBuffer(np.float32)
    hits, nsim = P.run_photon_prop_no_local_cache(wl.sources, wl.targets, wl.medium, wl.spectrum,
                                                  wl.seed, hbc, dev, n_steps=wl.n_steps)
    kms = float(np.mean(kernel_ms))
    search_used = ctypes.c_int32()
    lib.pp_search_info(dev.handle, ctypes.byref(search_used), None, None, None)
    grid = search_used.value == 2
    # work actually executed: one instrumented run (identical hits), untimed
    ccfg = P._lib.pp_config()
    P._lib.check(lib.pp_get_config(dev.handle, ctypes.byref(ccfg)), "pp_get_config")
    ccfg.count_work = 1
    _, cdev = P.make_hit_buffers(capacity=1 << 24, config=ccfg)
    P.processing.stage(cdev, wl.targets, wl.medium, wl.spectrum)
    P._lib.check(lib.pp_propagate(cdev.handle, pods, len(pods), ctypes.c_uint64(wl.seed), wl.n_steps,
                                  1.0, ctypes.byref(n_hits), ctypes.byref(n_sim)),
                 "pp_propagate (count_work)")
    w_steps, w_tests, w_cells, w_walks = (ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(),
                                          ctypes.c_int64())
    P._lib.check(lib.pp_work_counts(cdev.handle, ctypes.byref(w_steps), ctypes.byref(w_tests),
                                    ctypes.byref(w_cells), ctypes.byref(w_walks)),
                 "pp_work_counts")
    cdev.close()
    steps, tests, cells, walks = w_steps.value, w_tests.value, w_cells.value, w_walks.value
    assert steps == total_steps(hits, nsim, wl.n_steps)
    n_t = len(wl.targets)
    # algorithmic work: the reference step plus the sphere tests executed
    flop = steps * W_PHYS_FLOP + tests * W_ISEC_FLOP
    transc = steps * W_PHYS_TRANSC + tests * W_ISEC_TRANSC
    # the grid's own clip and walk, reported apart (work of this build's
    # search, not of the reference algorithm)
    flop_grid = (steps * W_GRID_STEP_FLOP + walks * W_GRID_WALK_FLOP + cells * W_CELL_FLOP
                 if grid else 0.0)
    transc_grid = walks * W_GRID_WALK_TRANSC if grid else 0.0
    achieved = flop / (kms * 1e-3) / 1e12
    traffic, traffic_src = measured_traffic(wl.name)
    counters = measured_counters(wl.name)
    ref_equiv = steps * (W_PHYS_FLOP + W_ISEC_FLOP * n_t) / (kms * 1e-3) / 1e12
    out = {
        "metric": "photons propagated/s on 1 MI355X; surrogate vs MC speedup at \u22645% "
                  "hit-count err",
        "value": photons_per_s,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "photons": nph, "sources": len(pods), "targets": n_t,
                   "n_steps": wl.n_steps, "search": args.search, "cell_m": args.cell,
                   "block": args.block, "waves_per_simd": args.wps, "batch": args.batch,
                   "chunk": args.chunk,
                   "source": ("ExtendedCherenkovEmitter 10 TeV e-, 4x4x20 spheres r=0.3 m"
                              if args.workload == "cascade" else
                              "256 cascades/100 m tracks, log-uniform 1e4-1e7 photons, "
                              "10x10x10 spheres r=0.3 m")},
        "launches_per_step": int(dev.last_timing()[1]),
        "launch": P.autotune.config_to_dict(ccfg),
        "autotune": None if tuned is None else {"ms_on_sample": tuned["ms"],
                                                "sample_photons": tuned["sample_photons"]},
        "photon_steps_per_s": steps / (ms_per_step / 1e3),
        "kernel_ms": kms,
        "hits_per_step": int(len(hits["n_steps"])),
        "sphere_tests_per_step": tests / steps,
        "cell_visits_per_step": cells / steps,
        "grid_walks_per_step": walks / steps,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "peak_nonpacked": PEAK_FP32_NONPACKED_TFLOPS,
                     "transcendental_per_s": transc / (kms * 1e-3),
                     "flop_per_launch": flop,
                     "mean_sphere_tests_per_step": tests / steps,
                     "reference_equivalent_tflops": ref_equiv,
                     "grid_overhead": {"flop_per_launch": flop_grid,
                                       "tflops": flop_grid / (kms * 1e-3) / 1e12,
                                       "transcendental_per_s": transc_grid / (kms * 1e-3),
                                       "basis": "12 FP32 ops per step (box test of the end "
                                                "points), 25 FP32 ops + 3 reciprocals per walk "
                                                "(clip), 4 per cell visit"},
                     "basis": "photon steps x (47 + 20 T) FP32 ops, T = sphere tests executed "
                              "per step, counted by the instrumented kernel (identical hits); "
                              "transcendentals 7.4 + T per step",
                     "counters": counters},
    }
    # issue rates measured on this device (8 chains per lane, 8 waves per SIMD)
    rates = {}
    for key, which, scale in (("fma_f32_tflops", P._lib.MB_FMA_F32, 1e12),
                              ("trans_f32_tops", P._lib.MB_TRANS_F32, 1e12),
                              ("mad_u64_u32_tops", P._lib.MB_MAD_U64, 1e12)):
        r = ctypes.c_double()
        P._lib.check(lib.pp_microbench(0, which, ctypes.byref(r)), "pp_microbench")
        rates[key] = r.value / scale
    out["roofline"]["measured_issue_rates"] = rates
    out["roofline"]["frac_of_measured_fma"] = achieved / rates["fma_f32_tflops"]
    out["host_inclusive"] = copy_leg(lib, P, dev, pods, wl) if args.legs else None
    dev.close()
    out["readme_point_source"] = readme_leg(lib, P, not args.no_cpu) if args.legs else None
    out["batching_uneven_1000"] = batching_leg(lib, P) if args.legs else None
    out["surrogate"] = surrogate_leg(args) if args.surrogate else None
    out["pmt_detection"] = pmt_leg() if args.pmt else None
    if not args.no_cpu and args.workload == "cascade":
        out["cpu_baseline"] = cpu_baseline(
            lambda n: P.workloads.cascade_16_strings(photons=n), args.cpu_sample,
            "the same cascade event (320 spheres)")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
