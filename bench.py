#!/usr/bin/env python3
"""Benchmark: plain GPU tile-fit scoring on MI355X (one GPU).

Workload (default): 1920x1080 synthetic RGB float32 image (seed 0, 600 primitives),
adaptive quadtree tiles of 8-64 px split while the tile's luminance variance > 1e-3,
1024 candidates per tile per step drawn on the device, stack depth 16, float64 R-D
objective (lambda of quality 75).  A step is one greedy round over every tile: score all
candidates against the cached canvases (one launch per tile-size class), take the first
minimum, accept and composite (one launch).  Inputs are resident in HBM before timing.

Prints one JSON line (metric, value, unit, n_gpus, steps, warmup, ms_per_step, ...): value = candidate*pixel evaluations per
second over the K timed steps; roofline = the dominant scoring kernel's algorithmic FP32
rate (FLOP per evaluation x evaluations per launch / mean launch time measured with HIP
events on the launch stream) against MI355X's 157.3 TFLOP/s FP32 vector peak;
cpu_baseline = the CPU oracle (C, float32, OpenMP) on a bounded sample of the same step.

Further objects on the same line (skipped with --no-extras):
- time_to_psnr: on the same image, the plain fit (1024 candidates, 16 rounds, one
  persistent launch) against the refined fit (256 drawn candidates, top 2 refined by 12
  Adam steps, exact re-score), also one persistent launch: wall time until the refined
  fit reaches the plain fit's final PSNR, and both fits' PSNR at the end and at equal
  primitive bytes (the stepwise refined fit gives the PSNR after every round).
- time_to_psnr_4k: the same at 3840x2160 (2500 primitives + value-noise texture),
  depth 24, plain with 2048 candidates, refined with 128 drawn and the best one refined
  by 20 Adam steps.  (Both refined settings are the fastest to the plain PSNR in a sweep
  of drawn count, top k and Adam steps.)
- batched: 64 synthetic images (~88.8 Mpx) in one persistent hipGraph-captured launch,
  512 candidates, 16 rounds, autotuned queue order and workgroup size; throughput and
  per-image p50 / p99 latency.
- refine: 4096 tiles of 32 px from the 256x256 generator, 256 drawn candidates, the top
  64 refined by 20 Adam steps; the refinement's time and candidate-pixel rate (forward
  and backward) and how much the first round improves on plain search.
- drop_in: the reference API gpu_batch_score_prepacked on 32 stacks of 4 primitives per
  tile of the 1920x1080 quadtree: device-resident and host-inclusive rates.
- decode: gpu_render_tiles_to_image of the 3840x2160 plain fit's primitives.
- entropy: chunked rANS encode and decode (64 tiles per chunk, the reference default) of that fit's PRIM bytes.
- candidates: the template candidate pools (encoder._generate_candidates) of every tile
  of the 1920x1080 quadtree on the device, and the NumPy restatement on one CPU core for a
  sample of the same tiles; under "residual" the same for the residual pools
  (encoder._generate_residual_candidates) against a seeded noisy prediction, and under
  "edge" for the edge-driven pools (edge_analysis.generate_edge_driven_candidates) and the
  merged pools of encoder._worker_gen_cands.
- adaptive: the reference encoder's fit phase (encoder._fit_adaptive_state) on the
  1920x1080 image: its quadtree, the merged pools of every tile and the batched greedy
  rounds with re-seeding, all on the device, with the NumPy restatement of pools and loop
  on one CPU core for a sample of the tiles.
- select: the selection helpers (argmin, exclusive scan, compaction, per-segment top-k)
  on 2^22 device elements.
- template_fit: the stage-3 template fit (fit_tiles_gpu_constant) of the 3840x2160 image at
  16-px tiles: device rate, HBM rate against the 8 TB/s peak, the host API's time and the
  single-threaded CPU oracle on the same image.
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

METRIC = "candidate·pixel evals/s; time-to-PSNR, refined+batched vs plain GPU search"
PEAK_FP32 = 157.3e12
# Algorithmic FLOP per candidate-pixel evaluation (add/sub/mul/div = 1, FMA-shaped = 2,
# compares 0, per-candidate constants hoisted): coverage + composite (10) + SSE (9).
FLOP_PER_EVAL = {"triangle": 32, "ellipse": 33, "rect": 28}
# Soft-edge forward + backward per candidate-pixel (same rules; sigmoid = exp + add + rcp):
# ellipse 80, rectangle 114, triangle 96, in the drawn family mix (a third each).
SOFT_FLOP_MIX = (80 + 114 + 96) / 3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flop_per_eval_mix() -> float:
    # families are drawn in groups of 64 candidates, (r + group) mod 3 with r uniform per
    # tile and step: each family has an expected share of exactly one third
    return (FLOP_PER_EVAL["ellipse"] + FLOP_PER_EVAL["rect"] + FLOP_PER_EVAL["triangle"]) / 3.0


def build_workload(args):
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.synth import CONFIG_C2, image_checksum, synth_image

    c = dict(CONFIG_C2)
    img = synth_image(c["height"], c["width"], c["n_prims"], seed=c["seed"])
    qt = decompose_quadtree_variance(img, threshold=c["var_threshold"], min_tile=c["min_tile"],
                                     max_tile=c["max_tile"])
    tiles = TileSet.from_images([img], [qt])
    return c, img, tiles, image_checksum(img)


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch (read + written) from profiles/traffic.txt, measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and calibrated against a device copy."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == kernel:
                return float(parts[1]) + float(parts[2])
    return None


def cpu_baseline(tiles, c, lam, seed, budget_s, gpu_rec0):
    """Time the C oracle on a bounded tile sample of step 0 and check its winners."""
    from oracle import coracle

    coracle.build()
    cores = len(os.sched_getaffinity(0))  # every host CPU this process may run on
    order = np.random.default_rng(1).permutation(tiles.n_tiles)

    def run(n, threads):
        ids = np.sort(order[:n]).astype(np.int32)
        ts = tiles.ts[ids]
        offs = np.concatenate(([0], np.cumsum(ts.astype(np.int64) ** 2)[:-1]))
        tgt = np.concatenate([tiles.tile_pixels(int(t)).ravel() for t in ids])
        fit = coracle.PlainFit(tgt, offs, ts, c["n_cand"], c["depth"], seed, lam, tile_ids=ids)
        t0 = time.perf_counter()
        res = fit.run(1, n_threads=threads)
        return ids, ts, res, time.perf_counter() - t0

    def sample(threads, budget):
        n = max(4, tiles.n_tiles // 500)
        while True:
            ids, ts, res, dt = run(n, threads)
            if dt > 0.3 * budget or n >= tiles.n_tiles:
                return ids, ts, res, dt
            n = min(tiles.n_tiles, int(n * min(20.0, max(2.0, 0.6 * budget / max(dt, 1e-3)))))

    ids, ts, res, dt = sample(cores, budget_s)
    evals = int((ts.astype(np.int64) ** 2).sum()) * c["n_cand"]
    agree = bool(np.array_equal(res["win_idx"][0], gpu_rec0["win_idx"][ids])
                 and np.array_equal(res["win_acc"][0], gpu_rec0["win_acc"][ids])
                 and res["win_obj"][0].tobytes() == np.ascontiguousarray(gpu_rec0["win_obj"][ids]).tobytes())
    ids1, ts1, _, dt1 = sample(1, budget_s / 3)
    evals1 = int((ts1.astype(np.int64) ** 2).sum()) * c["n_cand"]
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": evals / dt, "unit": "candidate·pixel evals/s", "cores": cores, "kind": "port",
        "sample": f"step 0 of the same fit on {len(ids)} of {tiles.n_tiles} tiles "
                  f"({evals / 1e9:.2f} G evals in {dt:.1f} s; C float32 oracle, OpenMP {cores} threads)",
        "cpu_model": model, "affinity_cpus": cores,
        "single_core_value": evals1 / dt1,
        "single_core_sample": f"{len(ids1)} tiles, {evals1 / 1e9:.3f} G evals in {dt1:.1f} s, one thread",
        "winners_match_gpu": agree,
    }


def valid_mask(tiles, heights, widths):
    """Packed-layout mask (3 floats per pixel) of pixels inside their image (not padding)."""
    import torch

    valid = np.zeros(tiles.n_pixels, np.float32)
    for t in range(tiles.n_tiles):
        s = int(tiles.ts[t])
        o = int(tiles.pix_off[t])
        k = int(tiles.image[t])
        yy, xx = np.mgrid[0:s, 0:s]
        valid[o:o + s * s] = ((tiles.y[t] + yy < heights[k]) & (tiles.x[t] + xx < widths[k])).ravel()
    return torch.from_numpy(np.repeat(valid, 3)).cuda(), float(valid.sum())


def psnr_db(fitter, mask) -> float:
    d_valid, npx = mask
    e = (fitter.d_canvas.clamp(0, 1) - fitter.d_target) ** 2
    mse = float((e * d_valid).sum()) / (3 * npx)
    return float(10 * np.log10(1 / mse))


def refined_round_report(c, img, tiles, reps: int = 4):
    """Round one of the 1080p refined fit at its time-to-PSNR setting (256 drawn, top 2, 12
    Adam steps, fixed work split), `reps` times from a reset: wall time and PSNR."""
    import torch
    from weft_amd.rd import tile_lambda
    from weft_amd.refine import AdamConfig, RefinedFitter

    rf = RefinedFitter(tiles, n_cand=256, depth=c["depth"], k=2, adam=AdamConfig(n_iter=12), seed=0,
                       lam=tile_lambda(75), max_steps=c["depth"], cpw={8: 8, 16: 2, 32: 1, 64: 1})
    walls = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rf.reset()
        rf.step(1)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
    mask = valid_mask(tiles, [img.shape[0]], [img.shape[1]])
    return dict(n_cand=256, top_k=2, adam_steps=12, cpw={8: 8, 16: 2, 32: 1, 64: 1}, reps=reps,
                wall_ms=[round(w * 1e3, 3) for w in walls], psnr_db=psnr_db(rf, mask))


def time_to_psnr(tiles, mask, n_cand_plain, depth, seed, lam, refined_cfg):
    """Plain fit (persistent launch) vs refined fit round by round."""
    import torch
    from weft_amd.batch import BatchFitter
    from weft_amd.refine import AdamConfig, RefinedFitter

    bf = BatchFitter(tiles, n_cand=n_cand_plain, depth=depth, n_steps=depth, seed=seed, lam=lam)
    times = bf.time_runs(3, warmup=1)
    plain_records = bf.records()
    plain = dict(n_cand=n_cand_plain, rounds=depth, wall_s=min(times), psnr_db=psnr_db(bf, mask),
                 prims=int(plain_records["win_acc"].sum()))
    plain["prim_bytes"] = 17 * plain["prims"]
    del bf
    n_cand, k, iters = refined_cfg
    adam = AdamConfig(n_iter=iters)
    warm = RefinedFitter(tiles, n_cand=n_cand, depth=depth, k=k, adam=adam, seed=seed, lam=lam, max_steps=depth)
    cpw = warm.autotune()  # one-time choice of the refinement's work split, before timing
    warm.step(1)
    del warm
    rf = RefinedFitter(tiles, n_cand=n_cand, depth=depth, k=k, adam=adam, seed=seed, lam=lam, max_steps=depth,
                       cpw=cpw)
    torch.cuda.synchronize()
    # from an empty canvas, the initial objectives included, as in the plain fit's graph
    t0 = time.perf_counter()
    rf.reset()
    torch.cuda.synchronize()
    elapsed, reach, curve, psnr_at_bytes = time.perf_counter() - t0, None, [], None
    for r in range(depth):
        t0 = time.perf_counter()
        rf.step(1)
        torch.cuda.synchronize()
        elapsed += time.perf_counter() - t0
        p = psnr_db(rf, mask)
        prims = int(rf.records(0, r + 1)["win_acc"].sum())
        curve.append([round(elapsed, 5), round(p, 3), prims])
        if reach is None and p >= plain["psnr_db"]:
            reach = dict(rounds=r + 1, wall_s=elapsed, psnr_db=p, prims=prims)
        if psnr_at_bytes is None and prims >= plain["prims"]:
            psnr_at_bytes = p
    refined = dict(n_cand=n_cand, top_k=k, adam_steps=iters, cands_per_work_item=cpw, rounds=depth, wall_s=elapsed,
                   psnr_db=curve[-1][1], prims=curve[-1][2], prim_bytes=17 * curve[-1][2], reached_plain_psnr=reach,
                   psnr_at_plain_bytes=psnr_at_bytes, curve=curve,
                   timing="from an empty canvas (initial objectives included), round by round, five launches "
                          "per round, a device synchronisation after each")
    del rf
    # refined + batched: the refined fit in the persistent kernel, r rounds per launch, the
    # whole fit a graph replay.  It takes its own work split and workgroup size, the fastest
    # of a few (the split decides how many small tiles a workgroup refines together, and
    # cpw = -2 refines each 32/64-px candidate on two waves, which shortens the long queue
    # entries).  (With the stepwise split it equals the stepwise fit bit for bit; another split
    # orders the soft sums differently.)
    batched = None
    if reach:
        r = reach["rounds"]
        tr, block, bcpw = None, None, None
        splits = [cpw] + [m for m in ({8: 8, 16: 4, 32: 2, 64: 1}, {8: 4, 16: 2, 32: 1, 64: 1},
                                      {8: 16, 16: 4, 32: 2, 64: 1}, {8: 16, 16: 4, 32: -2, 64: -2},
                                      {8: 8, 16: 4, 32: -2, 64: -2}) if m != cpw]
        for sp in splits:
            for b in (128, 256):
                if 64 * max(-min(sp.values()), 1) > b:  # cpw = -W takes W waves of the workgroup
                    continue
                bfr = BatchFitter(tiles, n_cand=n_cand, depth=depth, n_steps=r, seed=seed, lam=lam, k=k,
                                  refine=adam, cpw=sp, block=b)
                t = min(bfr.time_runs(3, warmup=1))
                p = psnr_db(bfr, mask)
                if p >= plain["psnr_db"] and (tr is None or t < tr):  # it must reach the plain PSNR too
                    tr, block, bcpw, pr = t, b, sp, p
                del bfr
    if reach and tr is not None:
        bfd = BatchFitter(tiles, n_cand=n_cand, depth=depth, n_steps=depth, seed=seed, lam=lam, k=k, refine=adam,
                          cpw=bcpw, block=block)
        td = min(bfd.time_runs(2, warmup=1))
        pd = psnr_db(bfd, mask)
        del bfd
        batched = dict(kernel="weft_batch_refine_kernel", block=block, cands_per_work_item=bcpw,
                       rounds_to_plain_psnr=r, wall_s_to_plain_psnr=tr, psnr_db_after_those_rounds=pr,
                       whole_fit_wall_s=td, whole_fit_psnr_db=pd,
                       timing="fastest of 3 graph replays of the whole fit")
    speed_batched = plain["wall_s"] / batched["wall_s_to_plain_psnr"] if batched else None
    speed_stepwise = plain["wall_s"] / reach["wall_s"] if reach else None
    best = max((v, n) for v, n in ((speed_batched, "persistent"), (speed_stepwise, "stepwise")) if v is not None) \
        if (speed_batched or speed_stepwise) else (None, None)
    return dict(plain=plain, refined_batched=batched, refined_stepwise=refined, speedup_to_plain_psnr=best[0],
                fastest_refined_form=best[1], speedup_persistent=speed_batched,
                speedup_stepwise=speed_stepwise), plain_records


def drop_in_report(tiles, seed):
    """The reference API (gpu_batch_score_prepacked): 32 stacks of 4 drawn primitives per
    tile of the 1920x1080 quadtree (one separate record array per tile), host arrays in
    and out."""
    import torch
    from weft_amd import _lib
    from weft_amd.fit import draw_codes
    from weft_amd.gpu_render import gpu_batch_score_prepacked, score_evals_split_device, split_order
    from weft_amd.records import EVAL_DTYPE

    T, P, D = tiles.n_tiles, 32, 4
    _, prims = draw_codes(np.arange(T, dtype=np.int32), tiles.ts, P * D, seed=seed, step=1000)
    # one separate array per tile, as the reference encoder builds its trial stacks
    arrays = [prims[t * P * D:(t + 1) * P * D].copy() for t in range(T)]
    counts = [np.full(P, D, np.int32)] * T
    offsets = {t: int(tiles.pix_off[t]) * 3 for t in range(T)}
    args = (tiles.target, offsets, list(range(T)), [int(v) for v in tiles.ts], arrays, counts)
    gpu_batch_score_prepacked(*args)  # warm
    host = []
    for _ in range(5):
        t0 = time.perf_counter()
        out = gpu_batch_score_prepacked(*args)
        host.append(time.perf_counter() - t0)
    assert out is not None and len(out) == T
    # device-resident: the same evals with inputs already in HBM, HIP events around the launch
    ev = np.zeros(T * P, EVAL_DTYPE)
    ev["prim_offset"] = np.arange(T * P) * D
    ev["prim_count"] = D
    ev["tile_offset"] = np.repeat(tiles.pix_off.astype(np.int64) * 3, P)
    ev["tile_size"] = np.repeat(tiles.ts, P)
    d_ev = torch.from_numpy(ev.view(np.uint8)).cuda()
    d_pr = torch.from_numpy(np.ascontiguousarray(prims).view(np.uint8)).cuda()
    d_t = torch.from_numpy(tiles.target).cuda()
    d_out = torch.empty(T * P, dtype=torch.float32, device="cuda")
    L = _lib.library()
    stream = torch.cuda.current_stream()

    def timed(fn):
        times = []
        for _ in range(4):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            times.append(a.elapsed_time(b) * 1e-3)
        return float(np.median(times[1:]))

    one_s = timed(lambda: _lib.check(L.weft_score_evals(d_ev.data_ptr(), T * P, d_pr.data_ptr(), d_t.data_ptr(),
                                                        d_out.data_ptr(), None, _lib.stream_ptr()), "weft_score_evals"))
    one_lane = d_out.cpu().numpy()
    order, groups = split_order(ev["tile_size"])
    d_order = torch.from_numpy(order).cuda()
    dev_s = timed(lambda: score_evals_split_device(d_ev, d_order, groups, d_pr, d_t, d_out))
    same = bool(np.array_equal(d_out.cpu().numpy(), np.concatenate(out)))
    same_one = bool(d_out.cpu().numpy().tobytes() == one_lane.tobytes())
    px = int((tiles.ts.astype(np.int64) ** 2).sum()) * P
    return dict(api="gpu_batch_score_prepacked", tiles=T, evals=T * P, stack_depth=D,
                kernel="weft_score_evals_split_kernel", device_s=dev_s, evals_per_s=T * P / dev_s,
                eval_px_per_s=px / dev_s, eval_px_prim_per_s=px * D / dev_s, host_inclusive_s=min(host),
                host_inclusive_evals_per_s=T * P / min(host), host_inclusive_all_s=[round(h, 6) for h in host],
                device_equals_api=same,
                one_lane_kernel_s=one_s, one_lane_evals_per_s=T * P / one_s, split_equals_one_lane=same_one,
                split_speedup=one_s / dev_s)


def rounds_report(c, img, tiles, rounds: int = 40):
    """The plain fit's scoring launch, round by round from an empty canvas (every tile stays
    active): its duration per round, timed with HIP events on the launch stream."""
    import torch
    from weft_amd import _lib
    from weft_amd.fit import PlainFitter
    from weft_amd.rd import tile_lambda

    f = PlainFitter(tiles, n_cand=c["n_cand"], depth=c["depth"], seed=0, lam=tile_lambda(75), max_steps=rounds)
    stream = torch.cuda.current_stream()
    sp = _lib.stream_ptr()
    evs = []
    for s in range(rounds):
        cands = f._cands(s)
        f.order(s, sp)
        for w, start in f._works:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            f.score_class(w, start, cands, None, None, sp)
            b.record(stream)
            evs.append((a, b))
        f.select()
        f.step_idx += 1
    torch.cuda.synchronize()
    rec = f.records(0, rounds)
    acc = rec["win_acc"].astype(np.int64)
    active = ((np.cumsum(acc, axis=0) - acc) < c["depth"]).mean(axis=1)
    return dict(kernel="weft_fit_score_mixed", rounds=rounds,
                launch_ms=[round(a.elapsed_time(b), 4) for a, b in evs],
                active_tile_frac=[round(float(x), 4) for x in active])


def decode_report(tiles, records, height, width):
    """Decode walk (gpu_render_tiles_to_image) of a fitted image's accepted primitives."""
    import torch
    from weft_amd import _lib
    from weft_amd.gpu_render import gpu_render_tiles_to_image
    from weft_amd.records import PRIM_DTYPE, TILE_DECODE_DTYPE, codes_to_records

    acc, codes = records["win_acc"].astype(bool), records["win_code"]
    per_tile = [codes[acc[:, t], t] for t in range(tiles.n_tiles)]
    counts = np.array([len(c) for c in per_tile], np.int32)
    prims = codes_to_records(np.concatenate(per_tile)) if counts.sum() else np.zeros(1, PRIM_DTYPE)
    descs = np.zeros(tiles.n_tiles, TILE_DECODE_DTYPE)
    descs["prim_offset"] = np.concatenate(([0], np.cumsum(counts)[:-1]))
    descs["prim_count"] = counts
    descs["target_x0"], descs["target_y0"] = tiles.x, tiles.y
    descs["target_w"] = descs["target_h"] = descs["source_size"] = tiles.ts
    gpu_render_tiles_to_image(descs, prims, height, width)  # warm
    t0 = time.perf_counter()
    img = gpu_render_tiles_to_image(descs, prims, height, width)
    host = time.perf_counter() - t0
    d_d = torch.from_numpy(descs.view(np.uint8)).cuda()
    d_p = torch.from_numpy(np.ascontiguousarray(prims).view(np.uint8)).cuda()
    d_o = torch.zeros(height * width * 3, dtype=torch.float32, device="cuda")
    L = _lib.library()
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(4):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        _lib.check(L.weft_render_tiles(d_d.data_ptr(), tiles.n_tiles, d_p.data_ptr(), d_o.data_ptr(), width, height,
                                       _lib.stream_ptr()), "weft_render_tiles")
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b) * 1e-3)
    dev_s = float(np.median(times[1:]))
    same = bool(np.array_equal(d_o.cpu().numpy().reshape(height, width, 3), img))
    # algorithmic HBM bytes: every output pixel written once (12 B), descriptors and records read once
    alg = height * width * 12 + descs.nbytes + prims.nbytes
    return dict(api="gpu_render_tiles_to_image", width=width, height=height, tiles=int(tiles.n_tiles),
                prims=int(counts.sum()), kernel="weft_render_tiles_kernel", device_s=dev_s,
                mpx_per_s=height * width / dev_s / 1e6, algorithmic_bytes=int(alg),
                hbm_gb_per_s=alg / dev_s / 1e9, hbm_frac=alg / dev_s / 8e12, host_inclusive_s=host,
                device_equals_api=same)


def template_report(image, ts=16):
    """Stage-3 template fit (18 templates per tile) of one image, device-resident and via the API."""
    import torch
    from oracle.coracle import template_fit
    from weft_amd.gpu_encoder import MODEL_COUNT, TemplateFitter, fit_tiles_gpu_constant

    h, w, _ = image.shape
    res = fit_tiles_gpu_constant(image, ts)  # warm (module, masks)
    t0 = time.perf_counter()
    res = fit_tiles_gpu_constant(image, ts)
    host = time.perf_counter() - t0
    f = TemplateFitter(w, h, ts)
    d_img = torch.from_numpy(np.ascontiguousarray(image)).cuda()
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(6):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        f.launch(d_img)
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b) * 1e-3)
    dev_s = float(np.median(times[1:]))
    T = f.n_tiles
    # algorithmic bytes: the image once, the per-tile outputs (model, 2 colours, 18 MSE, 2 x 18 colours)
    nbytes = h * w * 12 + T * (4 + 24 + 4 * MODEL_COUNT + 24 * MODEL_COUNT)
    t0 = time.perf_counter()
    o = template_fit(image, ts)
    cpu_s = time.perf_counter() - t0
    same_model = float(np.mean(o["model"] == res.model_ids))
    return dict(api="fit_tiles_gpu_constant", width=w, height=h, tile_size=ts, tiles=T,
                kernel="weft_template_fit_kernel", device_s=dev_s, mpx_per_s=h * w / dev_s / 1e6,
                hbm_gb_per_s=nbytes / dev_s / 1e9, hbm_frac=nbytes / dev_s / 8e12, algorithmic_bytes=nbytes,
                host_inclusive_s=host, cpu_oracle_s=cpu_s, cpu_oracle_threads=1,
                speedup_vs_cpu_oracle=cpu_s / dev_s, models_equal_oracle_frac=same_model)


def candidates_report(tiles, quality=75, sample=48):
    """Template candidate pools (encoder._generate_candidates) for every tile of a quadtree."""
    import torch
    from oracle import candgen_np
    from weft_amd.candidates import MAX_CANDS, generate_pools
    from weft_amd.records import PRIM_DTYPE

    d_px = torch.from_numpy(tiles.target).cuda()
    d_off = torch.from_numpy(np.asarray(tiles.pix_off, np.int64)).cuda()
    generate_pools(tiles, quality, d_pixels=d_px, d_tile_pix=d_off)  # warm
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(3):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        recs, counts, _ = generate_pools(tiles, quality, d_pixels=d_px, d_tile_pix=d_off)
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b) * 1e-3)
    dev_s = float(np.median(times))
    cnt = counts.cpu().numpy()
    # CPU: the restatement (NumPy, one core) on a sample of the same tiles
    rng = np.random.default_rng(0)
    pick = rng.choice(tiles.n_tiles, min(sample, tiles.n_tiles), replace=False)
    t0 = time.perf_counter()
    same = 0
    host = recs.cpu().numpy().view(PRIM_DTYPE).reshape(-1, MAX_CANDS)
    for t in pick:
        s = int(tiles.ts[t])
        o = 3 * int(tiles.pix_off[t])
        ref = candgen_np.generate(tiles.target[o:o + 3 * s * s].reshape(s, s, 3), quality)
        same += int(cnt[t] == len(ref) and host[t, :cnt[t]].tobytes()
                    == candgen_np.to_records(ref, PRIM_DTYPE).tobytes())
    cpu_per_tile = (time.perf_counter() - t0) / len(pick)
    # algorithmic HBM bytes: every tile's pixels read once (12 B per pixel), every slot written (60 B)
    nbytes = 12 * int(np.sum(np.asarray(tiles.ts, np.int64) ** 2)) + 60 * MAX_CANDS * int(tiles.n_tiles)
    rep = dict(api="generate_pools", tiles=int(tiles.n_tiles), quality=quality, kernel="weft_gen_candidates_kernel",
               device_s=dev_s, tiles_per_s=tiles.n_tiles / dev_s, mean_pool=float(cnt.mean()),
               algorithmic_bytes=nbytes, hbm_gb_per_s=nbytes / dev_s / 1e9, hbm_frac=nbytes / dev_s / 8e12,
               candidates_per_s=float(cnt.sum()) / dev_s, cpu_numpy_s_per_tile=cpu_per_tile,
               cpu_numpy_tiles=len(pick), speedup_vs_cpu_numpy_one_core=cpu_per_tile * tiles.n_tiles / dev_s,
               sample_equal_restatement=f"{same}/{len(pick)}")
    rep["residual"] = residual_candidates_report(tiles, d_px, d_off, pick, quality)
    rep["edge"] = edge_candidates_report(tiles, d_px, d_off, pick, quality)
    return rep


def edge_candidates_report(tiles, d_px, d_off, pick, quality=75):
    """Edge-driven pools (edge_analysis.generate_edge_driven_candidates) and the merged pools
    of encoder._worker_gen_cands (edge-driven + template, de-duplicated) for every tile."""
    import torch
    from oracle import candgen_np, edge_np
    from weft_amd.candidates import (EDGE_MAX_CANDS, MAX_CANDS, MERGED_MAX_CANDS, generate_edge_pools,
                                     generate_merged_pools)
    from weft_amd.records import PRIM_DTYPE

    out = {}
    for name, fn, cap, ref_fn in (("edge", generate_edge_pools, EDGE_MAX_CANDS, edge_np.generate),
                                  ("merged", generate_merged_pools, MERGED_MAX_CANDS, edge_np.merged)):
        fn(tiles, quality, d_pixels=d_px, d_tile_pix=d_off)  # warm, and no tile with edge points
        stream = torch.cuda.current_stream()
        times = []
        for _ in range(3):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            recs, counts, _ = fn(tiles, quality, d_pixels=d_px, d_tile_pix=d_off, check=False)
            b.record(stream)
            torch.cuda.synchronize()
            times.append(a.elapsed_time(b) * 1e-3)
        dev_s = float(np.median(times))
        cnt = counts.cpu().numpy()
        host = recs.cpu().numpy().view(PRIM_DTYPE).reshape(-1, cap)
        t0 = time.perf_counter()
        same = 0
        for t in pick:
            s = int(tiles.ts[t])
            o = 3 * int(tiles.pix_off[t])
            ref = ref_fn(tiles.target[o:o + 3 * s * s].reshape(s, s, 3), quality)
            same += int(cnt[t] == len(ref) and host[t, :cnt[t]].tobytes()
                        == candgen_np.to_records(ref, PRIM_DTYPE).tobytes())
        cpu_per_tile = (time.perf_counter() - t0) / len(pick)
        px_bytes = 12 * int(np.sum(np.asarray(tiles.ts, np.int64) ** 2))
        nbytes = px_bytes + 60 * cap * int(tiles.n_tiles)
        if name == "merged":  # both generators' pixels, records and float64 values, read back by the merge
            nbytes = 2 * px_bytes + 2 * (60 + 112) * (EDGE_MAX_CANDS + MAX_CANDS) * int(tiles.n_tiles) \
                + 60 * cap * int(tiles.n_tiles)
        out[name] = dict(api=fn.__name__, device_s=dev_s, tiles_per_s=tiles.n_tiles / dev_s,
                         mean_pool=float(cnt.mean()), algorithmic_bytes=nbytes, hbm_gb_per_s=nbytes / dev_s / 1e9,
                         hbm_frac=nbytes / dev_s / 8e12, cpu_numpy_s_per_tile=cpu_per_tile, cpu_numpy_tiles=len(pick),
                         speedup_vs_cpu_numpy_one_core=cpu_per_tile * tiles.n_tiles / dev_s,
                         sample_equal_restatement=f"{same}/{len(pick)}")
    out["edge"]["kernel"] = "weft_gen_edge_candidates_kernel"
    out["merged"]["kernels"] = ["weft_gen_edge_candidates_kernel", "weft_gen_candidates_kernel",
                                "weft_merge_pools_kernel"]
    return out


def adaptive_report(img, quality=75, sample=24, reps=3):
    """The reference's fit phase (encoder._fit_adaptive_state, encoder.py:2603-2831) on the
    1920x1080 image: its quadtree (complexity criterion, split threshold 0.08 at quality 75,
    8-32 px), the merged candidate pools of _worker_gen_cands for every tile, and the batched
    greedy rounds with winner removal, stalled-tile drop-out and residual re-seeding -- all on
    the device (generate_merged_pools, AdaptiveFitter), timed on the host clock with the
    device synchronised.  The CPU figure is the NumPy restatement of the same loop
    (oracle/adaptive_np.py, pools from oracle/edge_np.py) on a sample of the tiles, one core."""
    import torch
    from oracle import adaptive_np, candgen_np, edge_np
    from oracle.coracle import PRIM_DTYPE as ORC_PRIM
    from weft_amd.adaptive import AdaptiveFitter
    from weft_amd.candidates import MERGED_MAX_CANDS, generate_merged_pools
    from weft_amd.quadtree import TileSet, decompose_quadtree

    qt = decompose_quadtree(img, split_threshold=0.08, min_tile=8, max_tile=32)
    tiles = TileSet.from_images([img], [qt], pad_multiple=32)
    d_px = torch.from_numpy(tiles.target).cuda()
    d_off = torch.from_numpy(np.asarray(tiles.pix_off, np.int64)).cuda()

    def once():
        recs, counts, _ = generate_merged_pools(tiles, quality, d_pixels=d_px, d_tile_pix=d_off)
        f = AdaptiveFitter(tiles, (recs, counts, MERGED_MAX_CANDS), quality).run()
        torch.cuda.synchronize()
        return f, counts

    once()  # warm
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f, counts = once()
        times.append(time.perf_counter() - t0)
    wall = float(np.median(times))
    canvas = np.clip(f.canvas(), 0.0, 1.0)
    mse = float(np.mean((canvas.astype(np.float64) - tiles.target) ** 2))
    nsel = f.d_nsel.cpu().numpy()
    # CPU: pools and the loop restated in NumPy, on a sample of the tiles
    rng = np.random.default_rng(1)
    pick = np.sort(rng.choice(tiles.n_tiles, min(sample, tiles.n_tiles), replace=False))
    patches = [tiles.tile_pixels(int(t)) for t in pick]
    t0 = time.perf_counter()
    packed = [candgen_np.to_records(edge_np.merged(p, quality), ORC_PRIM) for p in patches]
    adaptive_np.fit(patches, packed, quality, f.lam)
    cpu_per_tile = (time.perf_counter() - t0) / len(pick)
    return dict(api="generate_merged_pools + AdaptiveFitter", width=int(img.shape[1]), height=int(img.shape[0]),
                quality=quality, tiles=int(tiles.n_tiles), tile_sizes={int(s): int((tiles.ts == s).sum())
                                                                    for s in np.unique(tiles.ts)},
                mean_pool=float(counts[:tiles.n_tiles].float().mean().item()), rounds=f.rounds_run,
                primitives=int(nsel.sum()), psnr_db=10.0 * np.log10(1.0 / mse), wall_s=wall,
                tiles_per_s=tiles.n_tiles / wall, cpu_numpy_s_per_tile=cpu_per_tile, cpu_numpy_tiles=len(pick),
                speedup_vs_cpu_numpy_one_core=cpu_per_tile * tiles.n_tiles / wall)


def residual_candidates_report(tiles, d_px, d_off, pick, quality=75):
    """Residual pools (encoder._generate_residual_candidates) of every tile against a prediction.

    The prediction is the target with a seeded smooth error added (an unclipped canvas,
    clipped by the kernel), the shape of a fit's canvas after a few stages.
    """
    import torch
    from oracle import candgen_np
    from weft_amd.candidates import RES_MAX_CANDS, generate_residual_pools
    from weft_amd.records import PRIM_DTYPE

    g = torch.Generator(device="cuda").manual_seed(5)
    d_pred = d_px + 0.15 * (torch.rand(d_px.shape, generator=g, device="cuda") - 0.5)
    generate_residual_pools(tiles, d_pred, d_pixels=d_px, d_tile_pix=d_off)  # warm
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(3):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        recs, counts, _ = generate_residual_pools(tiles, d_pred, d_pixels=d_px, d_tile_pix=d_off)
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b) * 1e-3)
    dev_s = float(np.median(times))
    cnt = counts.cpu().numpy()
    pred = np.clip(d_pred.cpu().numpy(), 0, 1)
    host = recs.cpu().numpy().view(PRIM_DTYPE).reshape(-1, RES_MAX_CANDS)
    t0 = time.perf_counter()
    same = 0
    for t in pick:
        s = int(tiles.ts[t])
        o = 3 * int(tiles.pix_off[t])
        ref = candgen_np.generate_residual(tiles.target[o:o + 3 * s * s].reshape(s, s, 3),
                                           pred[o:o + 3 * s * s].reshape(s, s, 3), quality)
        same += int(cnt[t] == len(ref) and host[t, :cnt[t]].tobytes()
                    == candgen_np.to_records(ref, PRIM_DTYPE).tobytes())
    cpu_per_tile = (time.perf_counter() - t0) / len(pick)
    nbytes = 24 * int(np.sum(np.asarray(tiles.ts, np.int64) ** 2)) + 60 * RES_MAX_CANDS * int(tiles.n_tiles)
    return dict(api="generate_residual_pools", kernel="weft_gen_residual_candidates_kernel", device_s=dev_s,
                tiles_per_s=tiles.n_tiles / dev_s, mean_pool=float(cnt.mean()), algorithmic_bytes=nbytes,
                hbm_gb_per_s=nbytes / dev_s / 1e9, hbm_frac=nbytes / dev_s / 8e12,
                cpu_numpy_s_per_tile=cpu_per_tile, cpu_numpy_tiles=len(pick),
                speedup_vs_cpu_numpy_one_core=cpu_per_tile * tiles.n_tiles / dev_s,
                sample_equal_restatement=f"{same}/{len(pick)}")


def select_report(n=1 << 22, seed=0):
    """The selection helpers (argmin, exclusive scan, compaction, segmented top-k) on n device elements."""
    import torch
    from oracle import select_np
    from weft_amd.cuda_backend import argmin_device, compact_device, exclusive_scan_device, segmented_topk_device

    rng = np.random.default_rng(seed)
    v = torch.from_numpy(rng.normal(size=n)).cuda()
    x = torch.from_numpy(rng.integers(0, 1000, n)).cuda()
    f = torch.from_numpy((rng.random(n) < 0.5).astype(np.uint8)).cuda()
    seg = torch.from_numpy(rng.integers(0, n // 64, n)).cuda()
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        times = []
        for _ in range(5):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            times.append(a.elapsed_time(b) * 1e-3)
        return float(np.median(times))

    out = dict(n=n)
    for name, fn, nbytes in (("argmin", lambda: argmin_device(v), 8 * n),
                             ("exclusive_scan", lambda: exclusive_scan_device(x), 16 * n),
                             ("compact", lambda: compact_device(f), n + 8 * n // 2),
                             ("segmented_topk_k4", lambda: segmented_topk_device(v, seg, 4), 16 * n)):
        t = timed(fn)
        out[name] = dict(device_s=t, melem_per_s=n / t / 1e6, gb_per_s=nbytes / t / 1e9)
    o, c = segmented_topk_device(v, seg, 4)
    m = 1 << 16  # correctness spot check on a prefix against the restatement
    o2, c2 = segmented_topk_device(v[:m], seg[:m], 4)
    out["topk_prefix_equals_oracle"] = (o2[: int(c2[0])].cpu().numpy().tolist()
                                        == select_np.segmented_topk(v[:m].cpu().numpy(), seg[:m].cpu().numpy(), 4))
    return out


def entropy_report(tiles, records, chunk_tiles=64):
    """Chunked rANS of a fitted image's PRIM payload, encode and decode on the device."""
    import torch
    from weft_amd import _lib
    from weft_amd.entropy import decode_chunks_device, encode_chunks_device
    from weft_amd.primitives import encode_code_tiles

    acc, codes = records["win_acc"].astype(bool), records["win_code"]
    raw, toc = encode_code_tiles([list(codes[acc[:, t], t]) for t in range(tiles.n_tiles)])
    T = len(toc) - 1
    starts = np.arange(0, T, chunk_tiles)
    ends = np.minimum(T, starts + chunk_tiles)
    toc_a = np.asarray(toc, np.int64)
    descs = np.zeros((len(starts), 6), np.int32)
    descs[:, 0], descs[:, 1] = starts, ends - starts
    descs[:, 2], descs[:, 3] = toc_a[starts], toc_a[ends] - toc_a[starts]
    d_raw = torch.from_numpy(np.frombuffer(raw, np.uint8).copy()).cuda()
    stream = torch.cuda.current_stream()
    enc_t, dec_t = [], []
    for _ in range(4):
        a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a.record(stream)
        d_coded, out, total = encode_chunks_device(d_raw, descs)
        b.record(stream)
        back = decode_chunks_device(d_coded, total, out, len(raw))
        c.record(stream)
        torch.cuda.synchronize()
        enc_t.append(a.elapsed_time(b) * 1e-3)
        dec_t.append(b.elapsed_time(c) * 1e-3)
    same = back[:len(raw)].cpu().numpy().tobytes() == raw
    e, dd = float(np.median(enc_t[1:])), float(np.median(dec_t[1:]))
    return dict(payload="PRIM bytes of the 3840x2160 plain fit", raw_bytes=len(raw), coded_bytes=int(total),
                ratio=len(raw) / max(1, int(total)), chunks=len(descs), chunk_tiles=chunk_tiles,
                encode_s=e, decode_s=dd, encode_mb_per_s=len(raw) / e / 1e6, decode_mb_per_s=len(raw) / dd / 1e6,
                roundtrip_exact=bool(same),
                note="encode/decode times include the host-side launches and index transfers of the calls")


def batched_report(seed, lam):
    """64-image batch, one persistent launch per fit, autotuned."""
    from weft_amd.batch import BatchFitter, autotune
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.synth import config_c3_images

    t0 = time.perf_counter()
    imgs = config_c3_images(64, seed=0)
    qts = [decompose_quadtree_variance(im, threshold=1e-3, min_tile=8, max_tile=64) for im in imgs]
    tiles = TileSet.from_images(imgs, qts)
    log(f"[bench] batched: {tiles.n_tiles} tiles of 64 images built in {time.perf_counter() - t0:.1f} s")
    bf = BatchFitter(tiles, n_cand=512, depth=16, n_steps=16, seed=seed, lam=lam)
    tune = autotune(bf, orders=("sjf", "lpt"), blocks=(128, 256), trials=1)
    times = bf.time_runs(3, warmup=0)
    lat = bf.image_latency_s()
    ev = bf.evals()
    wall = float(np.median(times))
    return dict(images=64, tiles=int(tiles.n_tiles), pixels=int(tiles.n_pixels), n_cand=512, rounds=16,
                graph=True, order=bf.order, block=bf.block, grid=bf.grid, wall_s=wall, evals=ev,
                evals_per_s=ev / wall, p50_latency_ms=float(np.percentile(lat, 50)) * 1e3,
                p99_latency_ms=float(np.percentile(lat, 99)) * 1e3, max_latency_ms=float(lat.max()) * 1e3,
                autotune=[{k: (round(v, 5) if isinstance(v, float) else v) for k, v in r.items()}
                          for r in tune["table"]])


def refine_report(seed, lam):
    """4096 tiles of 32 px, top 64 of 256 drawn candidates refined by 20 Adam steps."""
    import torch
    from weft_amd.fit import PlainFitter
    from weft_amd.quadtree import TileSet, uniform_tiles
    from weft_amd.refine import AdamConfig, RefinedFitter, refine, topk
    from weft_amd.synth import CONFIG_C1, synth_image

    c = CONFIG_C1
    imgs = [synth_image(c["height"], c["width"], c["n_prims"], seed=s) for s in range(64)]
    tiles = TileSet.from_images(imgs, [uniform_tiles(c["height"], c["width"], 32)] * 64, pad_multiple=32)
    k, iters = 64, 20
    rf = RefinedFitter(tiles, n_cand=256, depth=8, k=k, adam=AdamConfig(n_iter=iters), seed=seed, lam=lam,
                       max_steps=8)
    pf = PlainFitter(tiles, n_cand=256, depth=8, seed=seed, lam=lam, max_steps=8)
    cpw = rf.autotune()
    rf.step(1)
    pf.step(1)
    o_r, o_p = rf.state()["obj"], pf.state()["obj"]
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(5):
        s = rf.step_idx
        rf.score(s)
        topk(rf, rf.d_mse, rf.n_cand, k, rf.d_topk)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        refine(rf, rf._cands(s), rf.d_topk, k, rf.adam, rf.d_rcodes, rf.d_rprims, work=rf.refine_work)
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b) * 1e-3)
    t = float(np.median(times))
    cand_px = tiles.n_pixels * k * iters
    rate = cand_px / t
    return dict(flop_per_cand_px=SOFT_FLOP_MIX, achieved_tflops=rate * SOFT_FLOP_MIX / 1e12,
                frac_of_fp32_peak=rate * SOFT_FLOP_MIX / PEAK_FP32,
                tiles=int(tiles.n_tiles), tile_size=32, n_cand=256, top_k=k, adam_steps=iters, cands_per_work_item=cpw[32],
                kernel="weft_refine_kernel", refine_ms=t * 1e3, cand_px_fwd_bwd_per_s=rate,
                round1_refined_better_frac=float((o_r < o_p).mean()),
                round1_objective_ratio_mean=float((o_r / o_p).mean()))


def class_probe_report(tiles, c, lam, reps: int = 3):
    """The 1080p scoring launch split by tile size and by primitive family.

    Each row is `reps` launches of weft_fit_score_mixed over one subset of the step's work
    items (one tile-size class, or all of them), timed with HIP events on the launch stream;
    the rows are launched in the order listed, so the dispatches of a counter run map onto
    them one to one.  Sources: the drawn candidates of step 0 (families stratified in groups
    of 64), and three explicit pools of 1024 candidates per tile holding one family each (the
    16 groups of that family among 3072 drawn candidates; explicit pools add a class-order
    launch, weft_fit_order_kernel, before their rows)."""
    import torch
    from weft_amd import _lib
    from weft_amd.fit import PlainFitter, draw_codes

    n_cand, depth = c["n_cand"], c["depth"]
    stream = torch.cuda.current_stream()
    sp = _lib.stream_ptr()
    px = tiles.ts.astype(np.int64) ** 2
    rows = []

    def probe(source, fitter):
        cands = fitter._cands(0)
        fitter.order(0, sp)
        subsets = [(str(cl["ts"]), cl["start"], cl["n"], tiles.ts == cl["ts"]) for cl in fitter.classes]
        subsets.append(("all", 0, fitter.n_work, np.ones(tiles.n_tiles, bool)))
        blk = fitter.classes[0]["block"]
        for label, start, n, ids in subsets:
            w = _lib.Work(fitter.d_work_tile.data_ptr() + 4 * start, fitter.d_work_c0.data_ptr() + 4 * start,
                          n, 0, blk, 0)
            times = []
            for _ in range(reps):
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record(stream)
                fitter.score_class(w, start, cands, None, None, sp)
                b.record(stream)
                torch.cuda.synchronize()
                times.append(a.elapsed_time(b) * 1e-3)
            evals = int(px[ids].sum()) * n_cand
            t = float(np.mean(times))
            rows.append(dict(source=source, tile_size=label, work_items=int(n), launches=reps, mean_ms=t * 1e3,
                             evals=evals, t_evals_per_s=evals / t / 1e12))
            log(f"[probe] {source:9s} ts {label:>3s}: {n:6d} items, {t * 1e3:7.3f} ms, {evals / t / 1e12:.3f} T evals/s")

    probe("drawn", PlainFitter(tiles, n_cand=n_cand, depth=depth, seed=0, lam=lam, max_steps=1))
    T = tiles.n_tiles
    codes, prims = draw_codes(np.arange(T, dtype=np.int32), tiles.ts, 3 * n_cand, seed=0, step=0)
    fam = codes["family"].reshape(T, 3 * n_cand)
    prims = prims.reshape(T, 3 * n_cand)
    for f, name in enumerate(("ellipse", "rect", "triangle")):
        sel = fam == f
        if not (sel.sum(axis=1) == n_cand).all():
            raise RuntimeError("a family does not hold exactly a third of 3 * n_cand drawn candidates")
        pool = np.ascontiguousarray(prims[sel].reshape(T * n_cand))
        probe(name, PlainFitter(tiles, n_cand=n_cand, depth=depth, seed=0, lam=lam, max_steps=1, cand_prims=pool))
        del pool
    return dict(workload="1920x1080 plain-fit step 0, 1024 candidates per tile", kernel="weft_fit_score_mixed",
                reps=reps, rows=rows)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the time-to-PSNR, batched and refine reports")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES of device memory once (HBM counter calibration) and exit")
    ap.add_argument("--only", choices=["template_fit", "select", "candidates", "refined_round", "class_probe",
                                       "drop_in", "decode", "rounds", "adaptive"],
                    default=None,
                    help="run only that report (for profiling) and print it as JSON")
    args = ap.parse_args()
    if args.only:
        from weft_amd.synth import synth_image

        if args.only == "template_fit":
            rep = template_report(synth_image(2160, 3840, 2500, seed=0, texture=True))
        elif args.only == "candidates":
            rep = candidates_report(build_workload(args)[2])
        elif args.only == "adaptive":
            rep = adaptive_report(build_workload(args)[1])
        elif args.only == "drop_in":
            c, _, tiles, _ = build_workload(args)
            rep = drop_in_report(tiles, c["seed"])
        elif args.only == "rounds":
            rep = rounds_report(*build_workload(args)[:3])
        elif args.only == "decode":
            from weft_amd.batch import BatchFitter
            from weft_amd.quadtree import TileSet, decompose_quadtree_variance
            from weft_amd.rd import tile_lambda

            img4 = synth_image(2160, 3840, 2500, seed=0, texture=True)
            t4 = TileSet.from_images([img4], [decompose_quadtree_variance(img4, threshold=1e-3, min_tile=8,
                                                                           max_tile=64)])
            bf = BatchFitter(t4, n_cand=2048, depth=24, n_steps=24, seed=0, lam=tile_lambda(75))
            bf.run()
            rep = decode_report(t4, bf.records(), 2160, 3840)
        elif args.only == "refined_round":
            rep = refined_round_report(*build_workload(args)[:3])
        elif args.only == "class_probe":
            from weft_amd.rd import tile_lambda

            c, _, tiles, _ = build_workload(args)
            rep = class_probe_report(tiles, c, tile_lambda(75))
        else:
            rep = select_report()
        print(json.dumps({args.only: rep}), flush=True)
        return
    if args.copy_probe:
        import torch

        src = torch.empty(args.copy_probe // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        torch.cuda.synchronize()
        dst = src.clone()
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": int(dst.numel() * 4)}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the tile-fit path "
              "shards by tiles (each GPU fits its own tiles, no collective), and one GPU is measured here", file=sys.stderr)
        sys.exit(2)

    import torch
    from weft_amd import _lib
    from weft_amd.fit import PlainFitter
    from weft_amd.rd import tile_lambda

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    c, img, tiles, checksum = build_workload(args)
    seed = 0
    lam = tile_lambda(75)
    total_steps = args.warmup + args.steps
    log(f"[bench] {tiles.n_tiles} tiles, sizes {dict(zip(*np.unique(tiles.ts, return_counts=True)))}, "
        f"{tiles.n_pixels} px, {c['n_cand']} candidates, steps {args.warmup}+{args.steps}")
    fitter = PlainFitter(tiles, n_cand=c["n_cand"], depth=c["depth"], seed=seed, lam=lam,
                         max_steps=max(total_steps, 1))
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        fitter.step()
    torch.cuda.synchronize()
    log("[bench] warmup done")

    # HIP events on the launch stream around every scoring launch.
    stream = torch.cuda.current_stream()
    ev = {i: [] for i in range(len(fitter._works))}
    t0 = time.perf_counter()
    torch.cuda.synchronize()
    start_ev = torch.cuda.Event(enable_timing=True)
    end_ev = torch.cuda.Event(enable_timing=True)
    start_ev.record(stream)
    sp = _lib.stream_ptr()
    for _ in range(args.steps):
        s = fitter.step_idx
        cands = fitter._cands(s)
        fitter.order(s, sp)
        for i, (w, start) in enumerate(fitter._works):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fitter.score_class(w, start, cands, None, None, sp)
            b.record(stream)
            ev[i].append((a, b))
        fitter.select()
        fitter.step_idx += 1
    end_ev.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start_ev.elapsed_time(end_ev)
    log(f"[bench] timed {args.steps} steps: wall {wall * 1e3:.2f} ms, events {gpu_ms:.2f} ms")

    # Evaluations actually performed: a tile is scored in step s iff its stack was not full.
    rec = fitter.records(0, total_steps)
    acc = rec["win_acc"].astype(np.int64)
    nsel_before = np.cumsum(acc, axis=0) - acc
    px = tiles.ts.astype(np.int64) ** 2
    active = nsel_before < c["depth"]
    evals_steps = (active * px[None, :]).sum(axis=1) * c["n_cand"]
    evals_timed = int(evals_steps[args.warmup:].sum())
    value = evals_timed / wall

    fpe = flop_per_eval_mix()
    per_class = {}
    for i, (w, start) in enumerate(fitter._works):
        ts = int(w.tile_size)
        times = [a.elapsed_time(b) * 1e-3 for a, b in ev[i]]
        ids = np.ones(tiles.n_tiles, bool) if ts == 0 else (tiles.ts == ts)
        ev_launch = float(np.mean([(active[s][ids] * px[ids]).sum() * c["n_cand"]
                                   for s in range(args.warmup, total_steps)]))
        per_class[ts] = dict(mean_s=float(np.mean(times)), total_s=float(np.sum(times)),
                             evals_per_launch=ev_launch,
                             tflops=ev_launch * fpe / float(np.mean(times)) / 1e12)
    dom = max(per_class, key=lambda k: per_class[k]["total_s"])
    d = per_class[dom]
    achieved = d["evals_per_launch"] * fpe / d["mean_s"]
    kname = "weft_fit_score_mixed" if dom == 0 else f"weft_fit_score_ts{dom}"
    roofline = {"bound": "valu", "kernel": kname, "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": traffic_from_profiles(kname),
                "flop_per_eval": fpe, "evals_per_launch": d["evals_per_launch"],
                "mean_launch_ms": d["mean_s"] * 1e3,
                "per_class": {str(k): {"mean_launch_ms": v["mean_s"] * 1e3, "tflops": v["tflops"]}
                              for k, v in per_class.items()}}
    log(f"[bench] {value / 1e12:.3f} T evals/s; dominant {kname}: {achieved / 1e12:.1f} TFLOP/s")

    cpu = None
    if not args.no_cpu:
        rec0 = fitter.records(0, 1)
        rec0 = {k: v[0] for k, v in rec0.items()}
        cpu = cpu_baseline(tiles, c, lam, seed, args.cpu_seconds, rec0)
        log(f"[bench] cpu baseline {cpu['value'] / 1e9:.3f} G evals/s on {cpu['cores']} cores; "
            f"winners match: {cpu['winners_match_gpu']}")

    extras = {}
    if not args.no_extras:
        mask = valid_mask(tiles, [c["height"]], [c["width"]])
        extras["time_to_psnr"], _ = time_to_psnr(tiles, mask, c["n_cand"], c["depth"], seed, lam, (256, 2, 12))
        t = extras["time_to_psnr"]
        log(f"[bench] time-to-PSNR 1080p: plain {t['plain']['psnr_db']:.2f} dB in {t['plain']['wall_s']:.4f} s; "
            f"speedup {t['speedup_to_plain_psnr']:.1f} ({t['fastest_refined_form']}; persistent "
            f"{t['speedup_persistent']:.1f}, stepwise {t['speedup_stepwise']:.1f})")
        del fitter
        from weft_amd.quadtree import TileSet, decompose_quadtree_variance
        from weft_amd.synth import synth_image

        img4 = synth_image(2160, 3840, 2500, seed=0, texture=True)
        t4 = TileSet.from_images([img4], [decompose_quadtree_variance(img4, threshold=1e-3, min_tile=8, max_tile=64)])
        m4 = valid_mask(t4, [2160], [3840])
        extras["time_to_psnr_4k"], rec4 = time_to_psnr(t4, m4, 2048, 24, seed, lam, (128, 1, 20))
        extras["time_to_psnr_4k"]["tiles"] = int(t4.n_tiles)
        t = extras["time_to_psnr_4k"]
        log(f"[bench] time-to-PSNR 4k: plain {t['plain']['psnr_db']:.2f} dB in {t['plain']['wall_s']:.4f} s; "
            f"speedup {t['speedup_to_plain_psnr']:.1f} ({t['fastest_refined_form']}; persistent "
            f"{t['speedup_persistent']:.1f}, stepwise {t['speedup_stepwise']:.1f})")
        extras["decode"] = decode_report(t4, rec4, 2160, 3840)
        extras["entropy"] = entropy_report(t4, rec4)
        log(f"[bench] entropy 4k PRIM: ratio {extras['entropy']['ratio']:.2f}, encode "
            f"{extras['entropy']['encode_mb_per_s']:.0f} MB/s, decode {extras['entropy']['decode_mb_per_s']:.0f} MB/s")
        log(f"[bench] decode 4k: {extras['decode']['mpx_per_s']:.0f} Mpx/s on the device")
        extras["template_fit"] = template_report(img4)
        log(f"[bench] template fit 4k: {extras['template_fit']['mpx_per_s']:.0f} Mpx/s, "
            f"{extras['template_fit']['hbm_gb_per_s']:.0f} GB/s on the device; CPU oracle "
            f"{extras['template_fit']['speedup_vs_cpu_oracle']:.0f}x slower")
        del t4, m4, rec4
        extras["candidates"] = candidates_report(tiles)
        log(f"[bench] candidate pools 1080p: {extras['candidates']['device_s'] * 1e3:.2f} ms for "
            f"{extras['candidates']['tiles']} tiles ({extras['candidates']['speedup_vs_cpu_numpy_one_core']:.0f}x "
            f"one CPU core); sample equal: {extras['candidates']['sample_equal_restatement']}")
        extras["adaptive"] = adaptive_report(img)
        log(f"[bench] adaptive fit phase 1080p: {extras['adaptive']['wall_s'] * 1e3:.1f} ms for "
            f"{extras['adaptive']['tiles']} tiles (pools + {extras['adaptive']['rounds']} rounds, "
            f"{extras['adaptive']['psnr_db']:.2f} dB; {extras['adaptive']['speedup_vs_cpu_numpy_one_core']:.0f}x "
            f"one CPU core)")
        extras["select"] = select_report()
        log(f"[bench] select 4M: argmin {extras['select']['argmin']['gb_per_s']:.0f} GB/s, scan "
            f"{extras['select']['exclusive_scan']['gb_per_s']:.0f} GB/s, segmented top-k "
            f"{extras['select']['segmented_topk_k4']['melem_per_s']:.0f} Melem/s")
        extras["drop_in"] = drop_in_report(tiles, seed)
        log(f"[bench] drop-in scorer: {extras['drop_in']['evals_per_s'] / 1e6:.1f} M evals/s on the device, "
            f"{extras['drop_in']['host_inclusive_evals_per_s'] / 1e6:.1f} M/s through the host API")
        extras["batched"] = batched_report(seed, lam)
        log(f"[bench] batched: {extras['batched']['evals_per_s'] / 1e12:.3f} T evals/s, "
            f"p50 {extras['batched']['p50_latency_ms']:.1f} ms, p99 {extras['batched']['p99_latency_ms']:.1f} ms")
        extras["refine"] = refine_report(seed, lam)
        log(f"[bench] refine: {extras['refine']['refine_ms']:.2f} ms, "
            f"{extras['refine']['cand_px_fwd_bwd_per_s'] / 1e9:.1f} G cand*px/s")

    out = {
        "metric": METRIC, "value": value, "unit": "candidate·pixel evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain GPU tile-fit scoring, 1920x1080 synthetic image (seed 0, 600 "
                               "primitives), quadtree 8-64 px (luma variance > 1e-3), 1024 drawn "
                               "candidates/tile/step, depth 16",
                   "tiles": int(tiles.n_tiles), "pixels": int(tiles.n_pixels),
                   "tile_sizes": {str(int(k)): int(v) for k, v in zip(*np.unique(tiles.ts, return_counts=True))},
                   "n_cand": c["n_cand"], "depth": c["depth"], "image_sha256": checksum,
                   "evals_per_step": int(evals_steps[args.warmup]), "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
