"""Differentiable refinement on the GPU: soft SSE, its hand-written gradient, top-k,
the fused Adam refinement and the refined greedy fit.

Tolerances (float32 on the device against float64 on the CPU):
- soft SSE: 1e-4 relative (float32 sums of up to 12288 terms, hardware exp).
- gradient: against central finite differences of the float64 forward with h = 1e-3
  (pixels for the geometry, theta stepped h / max(rx, ry)), 1e-3 relative for every
  component at least 1 % of the candidate's largest; the others 1e-4 of the largest.
- w -> 0: the soft SSE equals the exact hard-edge SSE to 1e-5 relative for candidates
  with no pixel centre within 0.02 px of an edge.
- top-k, the refinement's codes and parameters: bit-exact (stable order; the kernel's
  own gradient followed by a float32 Adam step mirrored on the CPU).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _random_tiles(sizes, seed=0):
    from weft_amd.quadtree import TileSet

    rng = np.random.default_rng(seed)
    sizes = np.asarray(sizes, np.int32)
    off = np.concatenate([[0], np.cumsum(sizes.astype(np.int64) ** 2)[:-1]]).astype(np.int64)
    n = int((sizes.astype(np.int64) ** 2).sum())
    target = rng.random(n * 3, dtype=np.float32)
    tiles = TileSet(np.zeros(len(sizes)), np.zeros(len(sizes)), sizes, np.zeros(len(sizes)), target, off)
    canvas = rng.random(n * 3, dtype=np.float32)
    return tiles, canvas


def _fitter(tiles, canvas, n_cand=64, depth=4, seed=0):
    from weft_amd.fit import PlainFitter

    f = PlainFitter(tiles, n_cand=n_cand, depth=depth, seed=seed, lam=1e-5, max_steps=4)
    f.set_canvas(canvas)
    return f


def _tile(flat, tiles, t):
    ts = int(tiles.ts[t])
    o = int(tiles.pix_off[t]) * 3
    return flat[o:o + ts * ts * 3].reshape(ts, ts, 3)


def _random_params(rng, tiles, k):
    G = tiles.n_tiles
    kinds = rng.choice([6, 7, 8], G * k).astype(np.int32)
    prm = np.zeros((G * k, 9), np.float32)
    for g in range(G):
        ts = float(tiles.ts[g])
        for j in range(k):
            prm[g * k + j] = [rng.uniform(0, ts), rng.uniform(0, ts), rng.uniform(1.5, ts), rng.uniform(1.5, ts),
                              rng.uniform(0, np.pi), *rng.uniform(0.05, 0.95, 3), rng.uniform(0.3, 1.0)]
    return kinds, prm


@pytest.mark.parametrize("k,waves", [(1, 1), (8, 1), (64, 1), (1, 2), (1, 4), (1, 8)])
def test_soft_sse_and_gradient_vs_finite_differences(gpu, k, waves):
    import torch
    from oracle import soft_np
    from weft_amd.refine import soft_eval

    tiles, canvas = _random_tiles([8, 16, 32, 64, 16, 32], seed=k)
    f = _fitter(tiles, canvas)
    rng = np.random.default_rng(100 + k)
    kinds, prm = _random_params(rng, tiles, k)
    _say(f"soft_eval: {tiles.n_tiles} tiles x {k} candidates, {waves} waves each, w 0.5")
    gt = torch.arange(tiles.n_tiles, dtype=torch.int32, device="cuda")
    sse, g = soft_eval(f, gt, torch.from_numpy(kinds).cuda(), torch.from_numpy(prm).cuda(), k, 0.5, waves=waves)
    sse, g = sse.cpu().numpy(), g.cpu().numpy()
    worst = 0.0
    for i in range(len(kinds)):
        t = i // k
        cv, tg = _tile(canvas, tiles, t), _tile(tiles.target, tiles, t)
        ref = soft_np.soft_sse(int(kinds[i]), prm[i], cv, tg, 0.5)
        assert abs(sse[i] - ref) <= 1e-4 * ref, (i, sse[i], ref)
        if i % max(1, k // 8):
            continue
        fd = soft_np.fd_grad(int(kinds[i]), prm[i].astype(np.float64), cv, tg, 0.5, h=1e-3)
        sc = np.max(np.abs(fd))
        big = np.abs(fd) >= 1e-2 * sc
        err = np.abs(g[i] - fd)
        rel = float(np.max(err[big] / np.abs(fd[big])))
        worst = max(worst, rel)
        assert rel <= 1e-3, (i, int(kinds[i]), g[i], fd)
        assert np.max(err[~big], initial=0.0) <= 1e-4 * sc, (i, int(kinds[i]), g[i], fd)
    _say(f"worst relative gradient error {worst:.2e}")


def test_soft_limit_is_hard_edge_render(gpu, oracle_lib):
    import torch
    from oracle import soft_np
    from weft_amd.refine import soft_eval

    sizes = [8, 16, 32, 64]
    tiles, canvas = _random_tiles(sizes, seed=4)
    f = _fitter(tiles, canvas)
    k = 64
    kinds = np.zeros(len(sizes) * k, np.int32)
    prm = np.zeros((len(sizes) * k, 9), np.float32)
    keep = np.zeros(len(sizes) * k, bool)
    codes = []
    for t, ts in enumerate(sizes):
        cs = oracle_lib.draw_codes(5, 2, t, 400, ts)
        picked = []
        for c in cs:
            kd, p = soft_np.from_code(c)
            if np.min(np.abs(soft_np.edge_distances(kd, p, ts))) >= 0.02:
                picked.append((c, kd, p))
            if len(picked) == k:
                break
        assert len(picked) >= 16
        for j, (c, kd, p) in enumerate(picked):
            kinds[t * k + j], prm[t * k + j], keep[t * k + j] = kd, p, True
            codes.append((t, j, c))
        for j in range(len(picked), k):  # padding: a copy of the first
            kinds[t * k + j], prm[t * k + j] = picked[0][1], picked[0][2]
    gt = torch.arange(len(sizes), dtype=torch.int32, device="cuda")
    sse, _ = soft_eval(f, gt, torch.from_numpy(kinds).cuda(), torch.from_numpy(prm).cuda(), k, 1e-4, grad=False)
    sse = sse.cpu().numpy()
    for t, j, c in codes:
        ts = sizes[t]
        prim = oracle_lib.codes_to_prims(np.array([c]))
        hard, _ = oracle_lib.score_candidates(_tile(tiles.target, tiles, t), _tile(canvas, tiles, t), prim)
        assert abs(sse[t * k + j] - hard[0]) <= 1e-5 * hard[0], (t, j, sse[t * k + j], hard[0])
    _say(f"hard limit: {len(codes)} candidates equal the exact SSE")


def test_autograd_function_backward(gpu):
    import torch
    from weft_amd.refine import soft_eval, soft_tile_sse

    tiles, canvas = _random_tiles([16, 32, 8, 64], seed=9)
    f = _fitter(tiles, canvas)
    kinds, prm = _random_params(np.random.default_rng(3), tiles, 8)
    gt = torch.arange(tiles.n_tiles, dtype=torch.int32, device="cuda")
    kd = torch.from_numpy(kinds).cuda()
    p = torch.from_numpy(prm).cuda().requires_grad_(True)
    wts = torch.rand(len(kinds), device="cuda")
    y = soft_tile_sse(p, f, gt, kd, 8, 0.5)
    (y * wts).sum().backward()
    sse, g = soft_eval(f, gt, kd, p.detach(), 8, 0.5)
    assert bits_equal(y.detach().cpu().numpy(), sse.cpu().numpy())
    assert bits_equal(p.grad.cpu().numpy(), (wts[:, None] * g).cpu().numpy())


def test_topk_matches_stable_sort(gpu):
    import torch
    from weft_amd.refine import topk

    tiles, canvas = _random_tiles([8] * 5 + [16] * 3, seed=2)
    f = _fitter(tiles, canvas)
    rng = np.random.default_rng(8)
    for n in (1, 77, 300, 1024, 2500, 8192):
        tab = rng.random((tiles.n_tiles, n)).astype(np.float32)
        if n > 50:
            tab[:, 5:40] = tab[:, 5:6]
            tab[0, 7] = np.nan
            tab[1, 20] = 0.0
            tab[1, 30] = -0.0
        for k in sorted({1, min(n, 2), min(n, 4), min(n, 8), min(n, 16), min(n, 64), min(n, 256)}):
            _say(f"topk n {n} k {k}")
            out = topk(f, torch.from_numpy(tab).cuda(), n, k).cpu().numpy()
            ref = np.stack([np.argsort(np.where(np.isnan(r), np.inf, r), kind="stable")[:k] for r in tab])
            assert np.array_equal(out, ref.astype(np.int32)), (n, k)


def _mirror_grads(f, tiles, kinds, prm, k, cpw, w):
    """Kernel gradients with the refinement's lane split: per tile size, 64 / cpw[ts] lanes
    per candidate (groups of cpw[ts] candidates of one tile; when cpw[ts] > k, a tile's k
    candidates padded with copies of themselves), or 64 W lanes over W waves for cpw = -W."""
    import torch
    from weft_amd.refine import soft_eval

    g = np.zeros_like(prm)
    for ts in sorted(set(tiles.ts.tolist())):
        c = cpw[ts]
        ids = np.nonzero(tiles.ts == ts)[0]
        rows = ids[:, None] * k + np.arange(k)[None, :]
        if c < 0:
            gt = torch.from_numpy(np.repeat(ids, k).astype(np.int32)).cuda()
            r = rows.ravel()
            _, gg = soft_eval(f, gt, torch.from_numpy(kinds[r]).cuda(), torch.from_numpy(prm[r]).cuda(), 1, w,
                              waves=-c)
            g[r] = gg.cpu().numpy()
        elif c <= k:
            gt = torch.from_numpy(np.repeat(ids, k // c).astype(np.int32)).cuda()
            r = rows.ravel()
            _, gg = soft_eval(f, gt, torch.from_numpy(kinds[r]).cuda(), torch.from_numpy(prm[r]).cuda(), c, w)
            g[r] = gg.cpu().numpy()
        else:
            r = np.tile(rows, (1, c // k)).ravel()
            gt = torch.from_numpy(ids.astype(np.int32)).cuda()
            _, gg = soft_eval(f, gt, torch.from_numpy(kinds[r]).cuda(), torch.from_numpy(prm[r]).cuda(), c, w)
            g[rows.ravel()] = gg.cpu().numpy().reshape(len(ids), c, 9)[:, :k].reshape(-1, 9)
    return g


@pytest.mark.parametrize("k,n_iter,cpw", [(8, 0, None), (8, 3, None), (64, 2, None), (2, 5, None),
                                          (16, 3, {8: 16, 16: 2, 32: 1, 64: 4}), (64, 2, {8: 64, 16: 64, 32: 64, 64: 64}),
                                          (4, 3, {8: 64, 16: 16, 32: 8, 64: 2}), (1, 2, {8: 32, 16: 4, 32: 1, 64: 2}),
                                          (2, 3, {8: 8, 16: 2, 32: 1, 64: -4}), (1, 3, {8: 8, 16: -2, 32: -2, 64: -2}),
                                          (2, 2, {8: -8, 16: 2, 32: -8, 64: -8})])
def test_refine_matches_kernel_gradient_and_cpu_adam(gpu, oracle_lib, k, n_iter, cpw):
    import torch
    from oracle import soft_np
    from weft_amd import _lib
    from weft_amd.primitives import CODE_DTYPE
    from weft_amd.records import PRIM_DTYPE
    from weft_amd.refine import DEFAULT_CPW, AdamConfig, refine, refine_work

    tiles, canvas = _random_tiles([8, 16, 32, 64, 32, 16, 8, 64], seed=k + n_iter)
    f = _fitter(tiles, canvas)
    T, n_cand, seed, step = tiles.n_tiles, 100, 77, 5
    rng = np.random.default_rng(k * 10 + n_iter)
    tk = np.stack([rng.permutation(n_cand)[:k] for _ in range(T)]).astype(np.int32)
    cfg = AdamConfig(n_iter=n_iter, w0=0.8, w1=0.3)
    cpw = cpw or DEFAULT_CPW
    wt, wj, cpwl = refine_work(tiles.ts, k, cpw)
    work = (torch.from_numpy(wt).cuda(), torch.from_numpy(wj).cuda(), cpwl)
    _say(f"refine k {k} n_iter {n_iter} cpw {cpwl}, {len(wj)} work items")
    codes, prims = refine(f, _lib.Cands(None, None, seed, step, n_cand, 0, 0), torch.from_numpy(tk).cuda(), k, cfg,
                          out_params=(pout := torch.zeros((T * k, 9), dtype=torch.float32, device="cuda")),
                          work=work)
    codes = codes.cpu().numpy().view(CODE_DTYPE)
    prims = prims.cpu().numpy().view(PRIM_DTYPE)
    dcodes = [oracle_lib.draw_codes(seed, step, t, n_cand, int(tiles.ts[t])) for t in range(T)]
    kinds = np.zeros(T * k, np.int32)
    prm = np.zeros((T * k, 9), np.float32)
    for t in range(T):
        for j in range(k):
            kinds[t * k + j], prm[t * k + j] = soft_np.from_code(dcodes[t][tk[t, j]])
    m1, m2 = np.zeros_like(prm), np.zeros_like(prm)
    sched = cfg.schedule()
    lr = np.array(cfg.lr(), np.float32)
    for i in range(n_iter):
        g = _mirror_grads(f, tiles, kinds, prm, k, cpw, float(sched[i, 0]))
        for r in range(T * k):
            prm[r], m1[r], m2[r] = soft_np.adam_step(prm[r], m1[r], m2[r], g[r], lr, cfg.beta1,
                                                      np.float32(1) - np.float32(cfg.beta1), cfg.beta2,
                                                      np.float32(1) - np.float32(cfg.beta2), cfg.eps, sched[i, 1],
                                                      sched[i, 2])
            prm[r] = soft_np.project(int(kinds[r]), prm[r], int(tiles.ts[r // k]))
    assert bits_equal(pout.cpu().numpy(), prm)
    for t in range(T):
        assert codes[t * (k + 1)] == dcodes[t][tk[t, 0]]
        for j in range(k):
            q = soft_np.to_code(int(kinds[t * k + j]), prm[t * k + j], int(tiles.ts[t]), CODE_DTYPE)
            assert codes[t * (k + 1) + 1 + j] == q
            if n_iter == 0:
                assert q == dcodes[t][tk[t, j]]
    assert bits_equal(prims.view(np.uint8), oracle_lib.codes_to_prims(codes).view(np.uint8))


def test_refine_lowers_soft_sse(gpu):
    """Twenty Adam steps lower the soft SSE of nearly every refined candidate."""
    import torch
    from weft_amd import _lib
    from weft_amd.refine import AdamConfig, refine, soft_eval, topk

    from weft_amd.quadtree import TileSet, uniform_tiles
    from weft_amd.synth import CONFIG_C1, synth_image

    c = CONFIG_C1
    img = synth_image(c["height"], c["width"], c["n_prims"], seed=c["seed"])
    tiles = TileSet.from_images([img], [uniform_tiles(c["height"], c["width"], 32)], pad_multiple=32)
    from weft_amd.fit import PlainFitter

    f = PlainFitter(tiles, n_cand=256, depth=8, seed=0, keep_sse=True, max_steps=2)
    f.score(0)
    k = 16
    tk = topk(f, f.d_mse, 256, k)
    cfg = AdamConfig(n_iter=20)
    pout = torch.zeros((tiles.n_tiles * k, 9), dtype=torch.float32, device="cuda")
    refine(f, f._cands(0), tk, k, cfg, out_params=pout)
    # soft SSE before (the drawn parameters) and after
    from oracle import soft_np

    T = tiles.n_tiles
    kinds = np.zeros(T * k, np.int32)
    p0 = np.zeros((T * k, 9), np.float32)
    from weft_amd.fit import draw_codes

    dc, _ = draw_codes(np.arange(T, dtype=np.int32), tiles.ts, 256, seed=0, step=0, want_prims=False)
    tkn = tk.cpu().numpy()
    for t in range(T):
        for j in range(k):
            kinds[t * k + j], p0[t * k + j] = soft_np.from_code(dc[t * 256 + tkn[t, j]])
    gt = torch.arange(T, dtype=torch.int32, device="cuda")
    kd = torch.from_numpy(kinds).cuda()
    before, _ = soft_eval(f, gt, kd, torch.from_numpy(p0).cuda(), k, 0.5, grad=False)
    after, _ = soft_eval(f, gt, kd, pout, k, 0.5, grad=False)
    better = (after < before).float().mean().item()
    _say(f"refined soft SSE lower for {better:.3f} of candidates; mean ratio "
         f"{(after / before).mean().item():.3f}")
    assert better > 0.9


def test_refined_fit_is_exact_and_no_worse(gpu, oracle_lib):
    """Accepted refined winners are re-scored exactly: canvas and objectives equal the
    oracle's composite of the accepted codes; round one is never worse than plain."""
    from weft_amd.fit import PlainFitter
    from weft_amd.primitives import CODE_DTYPE
    from weft_amd.quadtree import TileSet, decompose_quadtree_variance
    from weft_amd.refine import AdamConfig, RefinedFitter
    from weft_amd.synth import synth_image

    img = synth_image(200, 136, 30, seed=3)
    tiles = TileSet.from_images([img], [decompose_quadtree_variance(img, threshold=1e-3, min_tile=8, max_tile=64)])
    lam = 1e-5
    steps = 4
    rf = RefinedFitter(tiles, n_cand=128, depth=3, k=8, adam=AdamConfig(n_iter=10), seed=4, lam=lam,
                       max_steps=steps)
    pf = PlainFitter(tiles, n_cand=128, depth=3, seed=4, lam=lam, max_steps=steps)
    _say(f"refined fit: {tiles.n_tiles} tiles, 128 candidates, k 8, {steps} rounds")
    rf.step(1)
    pf.step(1)
    o_r, o_p = rf.state()["obj"], pf.state()["obj"]
    assert (o_r <= o_p).all()
    _say(f"round 1: refined better on {(o_r < o_p).mean():.2f} of tiles")
    rf.step(steps - 1)
    rec = rf.records()
    canvas = np.zeros_like(tiles.target)
    for t in range(tiles.n_tiles):
        ts = int(tiles.ts[t])
        cv = np.zeros((ts, ts, 3), np.float32)
        bits = 8
        for s in range(steps):
            if rec["win_acc"][s, t]:
                prim = oracle_lib.codes_to_prims(np.array([rec["win_code"][s, t]]).view(CODE_DTYPE))
                cv = oracle_lib.composite(prim[0], cv)
                bits += 17 * 8
                _, mse = oracle_lib.sse_mse(_tile(tiles.target, tiles, t), np.clip(cv, 0, 1))
                assert rec["win_obj"][s, t] == float(mse) + lam * bits, (t, s)
        o = int(tiles.pix_off[t]) * 3
        canvas[o:o + ts * ts * 3] = cv.ravel()
    assert bits_equal(rf.canvas(), canvas)


@pytest.mark.parametrize("cpw,n", [(None, 9), (None, 3), (None, 1), ({8: 64, 16: 64, 32: 64, 64: 64}, 9),
                                   ({8: 32, 16: 32, 32: 4, 64: 8}, 9)])
def test_split_scorer_matches_oracle(gpu, oracle_lib, cpw, n):
    """Lane-split exact scores equal the oracle's and the sequential kernel's bit for bit,
    with one tile per workgroup and with several small tiles sharing one."""
    import ctypes

    import torch
    from weft_amd import _lib
    from weft_amd.fit import PlainFitter
    from weft_amd.records import PRIM_DTYPE
    from weft_amd.refine import split_work

    tiles, canvas = _random_tiles([8, 16, 32, 64, 8, 16, 32, 64, 64, 32] + [8] * 19 + [16] * 5, seed=12)
    T = tiles.n_tiles
    prims = np.zeros(T * n, PRIM_DTYPE)
    for t in range(T):
        prims[t * n:(t + 1) * n] = oracle_lib.codes_to_prims(oracle_lib.draw_codes(3, 1, t, n, int(tiles.ts[t])))
    prims[2]["kind"] = 0
    prims[3]["kind"] = 4
    prims[3]["geom"][:6] = [1, 1, 6, 2, 3, 7]
    f = PlainFitter(tiles, n_cand=n, depth=4, lam=1e-5, cand_prims=prims, keep_sse=True, max_steps=2)
    f.set_canvas(canvas)
    f.score(0)
    seq = f.sse_table()
    wt, wc, w0, nw, st, cp, tp = split_work(tiles.ts, n, cpw)
    d = [torch.from_numpy(a).cuda() for a in (wt, wc)]
    w = _lib.SplitWork(d[0].data_ptr(), d[1].data_ptr(), len(wc), (ctypes.c_int32 * 4)(*cp),
                       (ctypes.c_int32 * 4)(*tp), 0)
    sse = torch.zeros(T * n, dtype=torch.float32, device="cuda")
    po = torch.zeros(len(wt), dtype=torch.float64, device="cuda")
    pi = torch.zeros(len(wt), dtype=torch.int32, device="cuda")
    _say(f"split scorer cpw {cp}, tiles per item {tp}, {len(wc)} work items")
    _lib.check(f.L.weft_fit_score_split(ctypes.byref(f._state), ctypes.byref(f._cands(0)), ctypes.byref(w),
                                        sse.data_ptr(), None, po.data_ptr(), pi.data_ptr(), _lib.stream_ptr()),
               "weft_fit_score_split")
    sse = sse.cpu().numpy().reshape(T, n)
    assert bits_equal(sse, seq)
    for t in range(T):
        ts = int(tiles.ts[t])
        ref, _ = oracle_lib.score_candidates(_tile(tiles.target, tiles, t), _tile(canvas, tiles, t),
                                             prims[t * n:(t + 1) * n])
        assert bits_equal(sse[t], ref), t
    # per-item first minima reduce to the sequential kernel's per-tile winner
    po, pi = po.cpu().numpy(), pi.cpu().numpy()
    for t in range(T):
        items = [int(w0[t]) + j * int(st[t]) for j in range(int(nw[t]))]
        best = min((po[i], pi[i]) for i in items)
        mse = sse[t] / np.float32(3 * int(tiles.ts[t]) ** 2)
        nbytes = np.array([{0: 7, 4: 19}.get(int(k), 17) for k in prims["kind"][t * n:(t + 1) * n]])
        obj = mse.astype(np.float64) + 1e-5 * (8 + 8 * nbytes)
        assert best[1] == int(np.argmin(obj)) and best[0] == obj.min()
