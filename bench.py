#!/usr/bin/env python3
"""Benchmark of the Mars range x domain x isometry search on one MI355X.

Workload (the quadtree one): a 512x512 synthetic 8-bit plane (seed 2), quadtree
range sizes 32 -> 16 -> 8 -> 4, domain stride 4, 8 isometries, exhaustive.
The searched node set of the RD walk is every aligned node of every size
(independent of lambda), so one step is the whole-plane search table:
  256 x 12769 + 1024 x 14641 + 4096 x 15625 + 16384 x 16129 ranges x positions,
  x 8 isometries = 2.77e9 evals.
A step runs pool build + range prep + search + merge on device-resident data
(plane in HBM at the start, winners table in HBM at the end), replayed from
one captured HIP graph.  `value` is evals/s of the MFMA search (exact integer
correlation: binary16 operands with binary32 accumulation at sizes 4 and 8,
split-INT8 with INT32 accumulation at 16 and 32); the plain FP32-FMA kernel is
timed beside it on the same workload.

Also measured (extra keys of the same line): 64-plane batches, the colour
workload (768x512 RGB -> YCbCr 4:2:0, Fisher classified search with the
exhaustive RD warm-up) and the CPU oracle on a bounded sample.

Prints ONE JSON line.
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

PEAK_INT8_OPS = 5.0e15   # MI355X dense INT8 MFMA, ops/s
PEAK_F16_FLOPS = 2.5e15  # MI355X dense FP16 MFMA, flop/s
PEAK_HBM = 8.0e12        # bytes/s

W = H = 512
SEED = 2
MIN_SIZE, MAX_SIZE, SHIFT = 4, 32, 4
WARMUP_T_RMS = 8.0  # the RD warm-up's threshold walk (encode.rs:1139-1169)


def _gpu_leg(steps: int, warmup: int, cpu_leg=None):
    """Both kernels on the quadtree512 workload.  cpu_leg(evals_by_size,
    gpu_run), when given, runs while the MFMA context is open, so its CPU
    repeats can interleave with GPU runs of the same step."""
    from mars_amd import _native as N
    from mars_amd import synth
    from mars_amd.provider import EncodeParams, HipContext

    img = synth.plane(W, H, SEED)
    params = EncodeParams(min_size=MIN_SIZE, max_size=MAX_SIZE, shift=SHIFT)
    res = {}
    for name, kern in (("mfma_i8", N.KERNEL_MFMA_I8), ("fp32", N.KERNEL_FP32)):
        ctx = HipContext(0, kern)
        ctx.prepare(img, params, [SHIFT])
        tune = ctx.autotune(reps=3)  # chunking shape per size, timed on this workload
        if warmup:
            ctx.run(iters=warmup, use_graph=True)
        # Per-kernel breakdown: average of several eager steps (HIP events on
        # the context stream around each launch).
        br = [ctx.run(iters=1, use_graph=False) for _ in range(5)]
        t = ctx.run(iters=steps, use_graph=True)
        by_size = {str(2 << i): float(np.mean([b.ms_by_size[i] for b in br]))
                   for i in range(6) if br[0].ms_by_size[i] > 0}
        tabs = ctx.tables()
        evals_by_size = {str(s): len(rows) * P * 8 for s, sh, rows, cols, c, P in tabs}
        res[name] = dict(ms_step=t.ms_step, evals=t.evals, ms_pool=float(np.mean([b.ms_pool for b in br])),
                         ms_pool_build=float(np.median([b.ms_pool_build for b in br])),
                         ms_search=float(np.mean([b.ms_search for b in br])),
                         ms_by_size=by_size, evals_by_size=evals_by_size,
                         survivors=int(br[0].survivors),
                         survivors_by_size={str(2 << i): int(br[0].survivors_by_size[i])
                                            for i in range(6) if br[0].ms_by_size[i] > 0},
                         exact_fits_by_size={str(2 << i): int(br[0].exact_fits_by_size[i])
                                             for i in range(6) if br[0].ms_by_size[i] > 0},
                         shapes={str(2 << i): [int(tune.target_blocks[i]), int(tune.min_tiles[i])]
                                 for i in range(6) if tune.ms[i] > 0},
                         winners=[[(c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta)
                                   if c is not None else None for c in tt[4]] for tt in tabs])
        if cpu_leg is not None and kern == N.KERNEL_MFMA_I8:
            res["cpu"] = cpu_leg(evals_by_size,
                                 lambda: ctx.run(iters=steps, use_graph=True).ms_step)
        ctx.close()
    return img, res


def _batch_leg(name: str, specs, steps: int):
    """One batched step over all planes (one search launch per size over every
    plane, one HIP graph) against one step per image (each image its own
    context and graph, the per-image step times summed)."""
    from mars_amd import synth
    from mars_amd.provider import EncodeParams, HipContext

    imgs = [synth.plane(w, h, seed) for w, h, seed, _ in specs]
    params = EncodeParams(min_size=MIN_SIZE, max_size=MAX_SIZE, shift=SHIFT)
    ctx = HipContext(0)
    ctx.prepare_batch(imgs, params, [SHIFT])
    tune = ctx.autotune(reps=1)
    ctx.run(iters=1, use_graph=True)
    t = ctx.run(iters=steps, use_graph=True)
    pool_ms = float(np.median([ctx.run(iters=1, use_graph=False).ms_pool_build for _ in range(3)]))
    ctx.close()
    pool_b = sum(_pool_bytes(w, h, [4, 8, 16, 32], SHIFT) for w, h, _, _ in specs)
    per_ms = 0.0
    for img in imgs:
        c = HipContext(0)
        c.prepare(img, params, [SHIFT])
        per_ms += c.run(iters=steps, use_graph=True).ms_step
        c.close()
    sizes = {}
    for w, h, _, _ in specs:
        sizes[f"{w}x{h}"] = sizes.get(f"{w}x{h}", 0) + 1
    return {"workload": name, "images": len(imgs), "sizes": sizes, "evals_per_batch": t.evals,
            "ms_per_batch": t.ms_step, "images_per_s": len(imgs) * 1e3 / t.ms_step,
            "evals_per_s": t.evals / (t.ms_step * 1e-3),
            "one_step_per_image": {"ms_per_batch": per_ms,
                                   "images_per_s": len(imgs) * 1e3 / per_ms},
            "speedup_vs_one_step_per_image": per_ms / t.ms_step,
            "pool_build": {"ms": pool_ms, "bytes": pool_b, "GB_per_s": pool_b / (pool_ms * 1e-3) / 1e9,
                           "frac_of_hbm_peak": pool_b / (pool_ms * 1e-3) / PEAK_HBM},
            "autotuned_shapes": {str(2 << i): [int(tune.target_blocks[i]), int(tune.min_tiles[i])]
                                 for i in range(6) if tune.ms[i] > 0}}


def _colour_leg(steps: int):
    """The colour workload: 768x512 RGB (seed 3) -> BT.601 YCbCr 4:2:0, Y quadtree
    32->4, Cb/Cr 16->4, stride 4, Fisher classified search as the provider and
    the exhaustive tables for the RD warm-up.  Device steps (both contexts,
    graph replays) plus one full encode to MARC bytes (host walks included)."""
    from mars_amd import _native as N
    from mars_amd import codec, synth
    from mars_amd.codec import EncodeOptions
    from mars_amd.provider import EncodeParams, HipContext
    from tests import _oracle as O

    rgb = synth.rgb(768, 512, synth.WORKLOAD_SEEDS["rgb768x512"])
    y, cb, cr = codec.ycbcr(rgb)
    cb, cr = codec.downsample_box(cb), codec.downsample_box(cr)
    yp = EncodeParams(min_size=4, max_size=32, shift=SHIFT, lambda_=200.0)
    cp = EncodeParams(min_size=4, max_size=16, shift=SHIFT, lambda_=200.0)
    out = {}
    for name, planes, p in (("y", [y], yp), ("chroma", [cb, cr], cp)):
        # As the encoder prepares them: the exhaustive tables only for the
        # nodes the RD warm-up's threshold walk visits; beside, full tables.
        for walk in (True, False):
            ctx = HipContext(0)
            if walk:
                ctx.set_walk_tables(WARMUP_T_RMS)
            ctx.prepare_batch(planes, p, [SHIFT], method=N.METHOD_FISHER)
            ctx.run(iters=2, use_graph=True)
            br = [ctx.run(iters=1, use_graph=False) for _ in range(5)]
            t = ctx.run(iters=steps, use_graph=True)
            ctx.close()
            if walk:
                out[name] = dict(ms_step=t.ms_step, exhaustive_evals=t.evals,
                                 fisher_evals=int(br[0].fisher_evals),
                                 ms_fisher=float(np.mean([b.ms_fisher for b in br])),
                                 ms_exhaustive_search=float(np.mean([b.ms_search for b in br])))
            else:
                out[name]["full_tables"] = dict(ms_step=t.ms_step, exhaustive_evals=t.evals)
    ctx = HipContext(0)
    o = EncodeOptions(allowed_modes=(True, False, True, False))
    t0 = time.perf_counter()
    data, stats, nbytes = codec.hip_encode_rgb(ctx, rgb, yp, cp, o, N.METHOD_FISHER)
    enc_ms = (time.perf_counter() - t0) * 1e3
    ctx.close()
    fe = out["y"]["fisher_evals"] + out["chroma"]["fisher_evals"]
    fms = out["y"]["ms_fisher"] + out["chroma"]["ms_fisher"]
    # the oracle's Fisher tables of the same nodes, all host threads
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    cpu_evals = 0
    for plane, p in ((y, yp), (cb, cp), (cr, cp)):
        h, w = plane.shape
        for size, rc in O.nodes(w, h, p.min_size, p.max_size).items():
            _, each = O.fisher_table(plane, size, SHIFT, [r for r, _ in rc], [c for _, c in rc],
                                     threads=threads)
            cpu_evals += sum(each)
    cpu_s = time.perf_counter() - t0
    return {"workload": "rgb768x512: 768x512 RGB (seed 3) -> YCbCr 4:2:0; Y 32->4, Cb/Cr 16->4, "
                        "stride 4, lambda 200, Fisher provider + exhaustive RD warm-up",
            "ms_device_step": out["y"]["ms_step"] + out["chroma"]["ms_step"],
            "device_step": out,
            "fisher_evals": fe, "ms_fisher": fms, "fisher_evals_per_s": fe / (fms * 1e-3),
            "encode_ms_wall": enc_ms, "marc_bytes": len(data), "plane_bytes": list(nbytes),
            "search_evals": {k: v.search_evals for k, v in stats.items()},
            "warmup_evals": {k: v.warmup_evals for k, v in stats.items()},
            "cpu_fisher": {"value": cpu_evals / cpu_s, "unit": "evals/s", "cores": threads,
                           "kind": "port", "sample": f"the oracle's Fisher tables of every node "
                                                     f"of the three planes, {cpu_evals} evals "
                                                     f"in {cpu_s:.2f} s"}}


def _ranked_leg(k: int, steps: int, warmup: int) -> int:
    """The quadtree512 step with ranked tables (each range's best k,
    search_top_k), against the same step at k = 1; one JSON line."""
    from mars_amd import synth
    from mars_amd.provider import EncodeParams, HipContext

    img = synth.plane(W, H, SEED)
    params = EncodeParams(min_size=MIN_SIZE, max_size=MAX_SIZE, shift=SHIFT)
    res = {}
    for kk in sorted({1, k}):
        ctx = HipContext(0)
        ctx.set_ranked(kk)
        ctx.prepare(img, params, [SHIFT])
        ctx.autotune(reps=1)
        ctx.run(iters=max(warmup, 1), use_graph=True)
        br = [ctx.run(iters=1, use_graph=False) for _ in range(3)]
        t = ctx.run(iters=steps, use_graph=True)
        ctx.close()
        res[kk] = {"ms_per_step": t.ms_step, "evals_per_s": t.evals / (t.ms_step * 1e-3),
                   "search_by_size_ms": {str(2 << i): float(np.mean([b.ms_by_size[i] for b in br]))
                                         for i in range(6) if br[0].ms_by_size[i] > 0},
                   "survivors_by_size": {str(2 << i): int(br[0].survivors_by_size[i])
                                         for i in range(6) if br[0].ms_by_size[i] > 0},
                   "exact_fits_by_size": {str(2 << i): int(br[0].exact_fits_by_size[i])
                                          for i in range(6) if br[0].ms_by_size[i] > 0}}
    print(json.dumps({"metric": f"quadtree512 step with ranked tables, k={k} vs k=1",
                      "unit": "evals/s", "ranked": {str(kk): v for kk, v in res.items()},
                      "slowdown_vs_k1": res[k]["ms_per_step"] / res[1]["ms_per_step"]}),
          flush=True)
    return 0


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _host_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _median_mad(xs):
    med = float(np.median(xs))
    return med, float(np.median([abs(x - med) for x in xs]))


def _cpu_leg(img, evals_by_size, gpu_run, reps: int = 5, per_size_s: float = 0.4):
    """The oracle (C restatement of the reference CPU search) on the host, in
    the reference's protocol (crates/mars-bench/src/gpu_search.rs:350-408):
    runs A/B interleaved with GPU runs, `reps` repeats, median + MAD, one
    thread and all host threads.  Each repeat times a fixed seeded sample of
    every size's ranges; the rate of the workload is the per-size rates
    combined with the workload's own per-size eval counts (a step is 76% s = 4
    evals, the CPU's fastest size per eval): sum(E_s) / sum(E_s / rate_s)."""
    from tests import _oracle as O

    threads = _host_threads()
    nodes = O.nodes(W, H, MIN_SIZE, MAX_SIZE)
    rng = np.random.default_rng(0)
    sizes = sorted(nodes)

    def timed(size, idx, nthreads):
        rows = [nodes[size][i][0] for i in idx]
        cols = [nodes[size][i][1] for i in idx]
        t0 = time.perf_counter()
        _, ev = O.search_table(img, size, SHIFT, rows, cols, threads=nthreads)
        return ev, time.perf_counter() - t0

    # sample sizes: about per_size_s of work per size and mode
    sample = {}
    for size in sizes:
        ev, dt = timed(size, [0], 1)
        n1 = int(max(1, min(len(nodes[size]), round(per_size_s / max(dt, 1e-6)))))
        nall = int(min(len(nodes[size]), n1 * threads))
        sample[size] = (rng.choice(len(nodes[size]), size=n1, replace=False),
                        rng.choice(len(nodes[size]), size=nall, replace=False))
    mix = {1: [], threads: []}
    per = {1: {s_: [] for s_ in sizes}, threads: {s_: [] for s_ in sizes}}
    gpu_ms = []
    for _ in range(reps):
        gpu_ms.append(gpu_run())  # A/B: a GPU run before each CPU repeat
        for mode, k in ((1, 0), (threads, 1)):
            denom = 0.0
            for size in sizes:
                ev, dt = timed(size, sample[size][k], mode)
                rate = ev / dt
                per[mode][size].append(rate)
                denom += evals_by_size[str(size)] / rate
            mix[mode].append(sum(evals_by_size[str(s_)] for s_ in sizes) / denom)
    st_med, st_mad = _median_mad(mix[1])
    mt_med, mt_mad = _median_mad(mix[threads])
    g_med, g_mad = _median_mad(gpu_ms)
    share = {str(s_): evals_by_size[str(s_)] / sum(evals_by_size[str(x)] for x in sizes) for s_ in sizes}
    return dict(
        value=mt_med, unit="evals/s", cores=threads, kind="port",
        sample=(f"{reps} repeats A/B-interleaved with GPU runs; each repeat times seeded samples "
                f"of every size's ranges of the 512x512 quadtree workload "
                f"({', '.join(f'{len(sample[s_][0])}/{len(sample[s_][1])} of size {s_}' for s_ in sizes)} "
                f"ranges, one thread/all threads) and combines the per-size rates with the "
                f"workload's eval shares ({', '.join(f's{k} {v:.1%}' for k, v in share.items())}); "
                f"oracle/liboracle.so, {_cpu_model()}"),
        mad=mt_mad, runs=mix[threads],
        single_thread={"value": st_med, "mad": st_mad, "runs": mix[1], "cores": 1},
        per_size_evals_per_s={"1": {str(s_): float(np.median(v)) for s_, v in per[1].items()},
                              str(threads): {str(s_): float(np.median(v)) for s_, v in per[threads].items()}},
        cpu_model=_cpu_model(),
        gpu_interleaved={"ms_per_step_median": g_med, "mad": g_mad, "runs": gpu_ms},
        gpu_vs_cpu_all_threads=None)


def _precision_study(sample: int, time_steps: bool):
    """Dense-stride precision study: a 2048x2048 plane (seed 7), stride 1, the
    fixed 8x8 and 16x16 grids (2.70e12 evals), searched by the split-INT8
    MFMA kernel and by the plain FP32 kernel.  Per size: the FP32 kernel's
    error against binary64 -- its own binary32 score of the winner it picked,
    and that winner after the binary64 re-score, against the exact winner --
    and both kernels against the oracle on a sample of ranges.  With
    time_steps, each kernel's step is also timed (a graph replay; the FP32
    step alone takes about 14 s)."""
    from mars_amd import _native as N
    from mars_amd import synth
    from mars_amd.provider import EncodeParams, HipContext
    from tests import _oracle as O

    size_w = 2048
    img = synth.plane(size_w, size_w, synth.WORKLOAD_SEEDS["dense2048"])
    params = EncodeParams(min_size=8, max_size=16, shift=1)
    runs = {}
    for name, kern in (("mfma_i8", N.KERNEL_MFMA_I8), ("fp32", N.KERNEL_FP32)):
        ctx = HipContext(0, kern)
        t0 = time.perf_counter()
        ctx.prepare(img, params, [1])
        prep_s = time.perf_counter() - t0
        r = dict(prepare_s=prep_s, tables={(tt[0], tt[1]): tt for tt in ctx.tables()},
                 scores=ctx.table_scores())
        if time_steps:
            t = ctx.run(iters=1, use_graph=True)
            r.update(ms_step=t.ms_step, evals=t.evals,
                     ms_by_size={str(2 << i): t.ms_by_size[i] for i in range(6) if t.ms_by_size[i] > 0})
        runs[name] = r
        ctx.close()
    codes = lambda c: None if c is None else (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta)
    per_size = {}
    rng = np.random.default_rng(0)
    threads = min(_host_threads(), 16)
    for key in sorted(runs["mfma_i8"]["tables"]):
        size, shift, rows, cols, exact, P = runs["mfma_i8"]["tables"][key]
        fp = runs["fp32"]["tables"][key][4]
        fsc = runs["fp32"]["scores"][key]
        msc = runs["mfma_i8"]["scores"][key]
        code_diff = sum(1 for a, b in zip(exact, fp) if codes(a) != codes(b))
        rms_diff = [abs(b.rms - a.rms) for a, b in zip(exact, fp)]
        own = [abs(s_ - b.rms) for s_, b in zip(fsc, fp)]           # binary32 score vs binary64
        own_rel = [o / b.rms for o, b in zip(own, fp) if b.rms > 0]
        mf_own = max(abs(s_ - a.rms) for s_, a in zip(msc, exact))
        pick = np.sort(rng.choice(len(rows), size=min(sample, len(rows)), replace=False))
        t0 = time.perf_counter()
        ref, _ = O.search_table(img, size, shift, [rows[i] for i in pick], [cols[i] for i in pick],
                                threads=threads)
        oracle_s = time.perf_counter() - t0
        okey = lambda c: None if c is None else (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta,
                                                  c.rms, c.m.t1_x4)
        gkey = lambda c: None if c is None else (c.dom_row, c.dom_col, c.isometry, c.qalfa, c.qbeta,
                                                  c.rms, c.t1_x4)
        per_size[str(size)] = {
            "ranges": len(rows), "positions": P, "evals": len(rows) * P * 8,
            "fp32_ranges_with_other_codes": code_diff,
            "fp32_ranges_with_other_rms": sum(1 for d in rms_diff if d != 0.0),
            "fp32_rescored_rms_err_max": max(rms_diff), "fp32_rescored_rms_err_mean": float(np.mean(rms_diff)),
            "fp32_own_score_err_max": max(own), "fp32_own_score_rel_err_max": max(own_rel),
            "fp32_own_score_rel_err_mean": float(np.mean(own_rel)),
            "mfma_own_score_err_max": mf_own,
            "oracle_sample": len(pick), "oracle_s": oracle_s,
            "mfma_vs_oracle_mismatches": sum(1 for j, i in enumerate(pick) if gkey(exact[i]) != okey(ref[j])),
            "fp32_vs_oracle_code_mismatches": sum(1 for j, i in enumerate(pick)
                                                  if codes(fp[i]) != (None if ref[j] is None else
                                                                      (ref[j].dom_row, ref[j].dom_col, ref[j].isometry,
                                                                       ref[j].qalfa, ref[j].qbeta))),
        }
        if time_steps:
            per_size[str(size)].update(mfma_ms=runs["mfma_i8"]["ms_by_size"].get(str(size)),
                                       fp32_ms=runs["fp32"]["ms_by_size"].get(str(size)))
    out = {"workload": "dense2048: 2048x2048 u8 plane (seed 7), stride 1, fixed 8x8 and 16x16 grids",
           "per_size": per_size}
    if time_steps:
        out.update(mfma_ms_per_step=runs["mfma_i8"]["ms_step"], fp32_ms_per_step=runs["fp32"]["ms_step"],
                   evals_per_step=runs["mfma_i8"]["evals"],
                   mfma_evals_per_s=runs["mfma_i8"]["evals"] / (runs["mfma_i8"]["ms_step"] * 1e-3),
                   fp32_evals_per_s=runs["fp32"]["evals"] / (runs["fp32"]["ms_step"] * 1e-3))
    return out


def _precision(sample: int) -> int:
    study = _precision_study(sample, True)
    line = {"metric": "precision study: split-INT8 MFMA vs FP32 kernel, error vs binary64",
            "config": {"workload": study.pop("workload")}}
    line.update(study)
    print(json.dumps(line), flush=True)
    return 0


def _traffic_probe() -> int:
    """Autotuned shapes, one eager step of the MFMA search and one 256 MiB copy
    kernel: the command rocprofv3 --pmc runs to read FETCH_SIZE / WRITE_SIZE
    (the step's launches are the last of each search kernel)."""
    from mars_amd import _native as N
    from mars_amd import synth
    from mars_amd.provider import EncodeParams, HipContext

    ctx = HipContext(0, N.KERNEL_MFMA_I8)
    ctx.prepare(synth.plane(W, H, SEED), EncodeParams(min_size=MIN_SIZE, max_size=MAX_SIZE,
                                                     shift=SHIFT), [SHIFT])
    ctx.autotune(reps=3)  # the shapes the bench's timed steps use
    ctx.run(iters=1, use_graph=False)
    ms = ctx.copy_probe(COPY_PROBE_BYTES)
    print(json.dumps({"copy_probe_bytes": COPY_PROBE_BYTES, "copy_probe_ms": ms}), flush=True)
    return 0


COPY_PROBE_BYTES = 256 << 20
CLOCK_HZ = 2.4e9    # MI355X peak engine clock
N_SIMD = 1024       # 256 CUs x 4 SIMDs


def _pool_bytes(w: int, h: int, sizes, shift: int) -> int:
    """Algorithmic bytes of one plane's pool build: the u8 plane read, the
    contracted plane as u16, and two u32 window sums per domain position of
    every searched size."""
    b = w * h + (w // 2) * (h // 2) * 2
    for s_ in sizes:
        if h >= 2 * s_ and w >= 2 * s_:
            b += ((h - 2 * s_) // shift + 1) * ((w - 2 * s_) // shift + 1) * 8
    return b


def _profiled_counter(kernel: str, counter: str):
    """One counter of `kernel` from the committed instruction-counter summary
    (profiles/quadtree512_search_counters.md), or None."""
    path = os.path.join(ROOT, "profiles", "quadtree512_search_counters.md")
    if not os.path.exists(path):
        return None
    col = None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if cells and cells[0] == "counter":
            col = cells.index(kernel) if kernel in cells else None
        elif col is not None and cells and cells[0] == counter:
            try:
                return float(cells[col])
            except (ValueError, IndexError):
                return None
    return None


def _profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy kernel's reading), or None."""
    import re
    path = os.path.join(ROOT, "profiles", "quadtree512_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    pat = re.compile(r"^\|\s*" + re.escape(kernel) + r"\s*\|.*\|\s*([0-9.eE+]+)\s*\|\s*$")
    for line in open(path):
        m = pat.match(line.strip())
        if m:
            return float(m.group(1))
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-image legs")
    ap.add_argument("--no-colour", action="store_true", help="skip the RGB 4:2:0 Fisher leg")
    ap.add_argument("--ranked", type=int, default=0,
                    help="time the quadtree512 step with ranked tables of this k (1..32) only")
    ap.add_argument("--colour-only", action="store_true",
                    help="only the RGB 4:2:0 Fisher leg (one JSON line of its own)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="one eager MFMA step + a 256 MiB copy kernel (for rocprofv3 --pmc)")
    ap.add_argument("--precision", action="store_true",
                    help="the 2048x2048 dense-stride precision study instead of the benchmark")
    ap.add_argument("--precision-sample", type=int, default=8,
                    help="ranges per size checked against the oracle in the precision study")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 2048x2048 precision key of the default line")
    args = ap.parse_args()
    if args.traffic_probe:
        return _traffic_probe()
    if args.precision:
        return _precision(args.precision_sample)
    if args.ranked:
        return _ranked_leg(args.ranked, args.steps, args.warmup)
    if args.colour_only:
        print(json.dumps({"metric": "RGB 4:2:0 Fisher encode, device step and Fisher evals/s",
                          "colour": _colour_leg(args.steps)}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU's share of the search; --gpus {args.gpus} is not "
              "supported: it times the search on one GPU (the search would shard by range block "
              "with no exchange but a per-range winner gather)", file=sys.stderr)
        return 2

    def cpu_hook(evals_by_size, gpu_run):
        from mars_amd import synth
        return _cpu_leg(synth.plane(W, H, SEED), evals_by_size, gpu_run)

    img, res = _gpu_leg(args.steps, args.warmup, None if args.no_cpu else cpu_hook)
    mf, fp = res["mfma_i8"], res["fp32"]
    evals = mf["evals"]
    value = evals / (mf["ms_step"] * 1e-3)
    fp_value = fp["evals"] / (fp["ms_step"] * 1e-3)
    # Dominant kernel: the search launch with the largest share of the step.
    dom_size = max(mf["ms_by_size"], key=lambda s: mf["ms_by_size"][s])
    dom_ms = mf["ms_by_size"][dom_size]
    s = int(dom_size)
    ops = mf["evals_by_size"][dom_size] * 2 * s * s   # 2*s^2 integer ops per eval
    achieved = ops / (dom_ms * 1e-3)
    diverge = sum(1 for a, b in zip(sum(mf["winners"], []), sum(fp["winners"], [])) if a != b)
    n_ranges = sum(len(t) for t in mf["winners"])

    batches = None
    if not args.no_batch:
        from mars_amd import synth
        b512 = [(W, H, 200 + i, 200.0) for i in range(64)]
        batches = {"batch512": _batch_leg("64 x 512x512 planes (seeds 200..263), quadtree 32->4, "
                                          "stride 4", b512, 3),
                   "batch64": _batch_leg("64 planes of uneven sizes (mars_amd.synth.batch_specs, "
                                         "seeds 100..163), quadtree 32->4, stride 4",
                                         synth.batch_specs(), 3)}
    colour = None if args.no_colour else _colour_leg(args.steps)
    cpu = res.get("cpu")
    if cpu is not None:
        cpu["gpu_vs_cpu_all_threads"] = value / cpu["value"]
    precision = None if args.no_precision else _precision_study(args.precision_sample, False)
    # VALU issue of the dominant launch: its SQ_INSTS_VALU (committed
    # counter summary) at 4 cycles per wave64 instruction on one SIMD,
    # against every SIMD's cycles over the launch measured here.
    valu = _profiled_counter(f"k_search_mfma<{s}>", "SQ_INSTS_VALU")
    valu_frac = valu * 4 / (N_SIMD * CLOCK_HZ * dom_ms * 1e-3) if valu else None
    pool_b = _pool_bytes(W, H, [int(k) for k in mf["evals_by_size"]], SHIFT)
    pool_gbs = pool_b / (mf["ms_pool_build"] * 1e-3) / 1e9
    line = {
        "metric": "R×D×iso evals/s, split-INT8 MFMA vs FP32 kernel; batched img/s @512²",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mf["ms_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "exact integer correlation on MFMA (f16 operands / f32 acc at s<=8, "
                 "split int8 / i32 acc at s>=16) + f64 fit",
        "data": "synthetic (seeded generator, mars_amd/synth.py, seed 2)",
        "config": {"workload": "quadtree512: 512x512 u8 plane, range sizes 32->4, domain stride 4, "
                               "8 isometries, exhaustive, whole RD node set per step",
                   "evals_per_step": evals, "ranges": n_ranges},
        "images_per_s": 1e3 / mf["ms_step"],
        "batched_images_per_s_512": batches["batch512"]["images_per_s"] if batches else None,
        "batches": batches,
        "colour": colour,
        "fp32_kernel": {"value": fp_value, "unit": "evals/s", "ms_per_step": fp["ms_step"],
                        "ranges_differing_from_mfma_winner": diverge},
        "speedup_vs_fp32": value / fp_value,
        "autotuned_shapes": {"mfma_i8": mf["shapes"], "fp32": fp["shapes"]},
        "breakdown_ms": {"pool_and_ranges": mf["ms_pool"], "search": mf["ms_search"],
                         "search_by_size": mf["ms_by_size"],
                         "fp32_search_by_size": fp["ms_by_size"]},
        "survivor_fraction": mf["survivors"] / evals,
        "survivors_by_size": mf["survivors_by_size"],
        "exact_fits_by_size": mf["exact_fits_by_size"],
        "roofline": {"bound": "mfma", "kernel": f"k_search_mfma<{s}>", "achieved": achieved / 1e12,
                     "peak": PEAK_INT8_OPS / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_INT8_OPS,
                     "traffic": _profiled_traffic(f"k_search_mfma<{s}>"),
                     "op": "2*s^2 ops per eval (integer multiply-adds)",
                     "mfma_instruction": ("v_mfma_f32_32x32x16_f16" if s in (4, 8)
                                          else "v_mfma_i32_32x32x32_i8 (two split planes)"),
                     "frac_of_f16_peak": (achieved / PEAK_F16_FLOPS if s in (4, 8) else None),
                     "avg_launch_ms": dom_ms,
                     "valu_frac": valu_frac,
                     "valu_frac_basis": "SQ_INSTS_VALU of the launch (profiles/quadtree512_search_counters.md) "
                                        "x 4 cycles / (1024 SIMDs x 2.4 GHz x avg_launch_ms)",
                     "pool": {"bound": "hbm", "achieved": pool_gbs, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                              "frac": pool_gbs * 1e9 / PEAK_HBM, "bytes": pool_b,
                              "ms": mf["ms_pool_build"],
                              "kernels": "k_contract, k_sat_rows, k_sat_cols, k_possums (serial)",
                              "basis": "W*H read + (W/2)(H/2) u16 contracted plane + 8 B per domain "
                                       "position of every size"}},
        "precision": precision,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
