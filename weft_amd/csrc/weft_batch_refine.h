// Persistent refined fit: the batched fit of weft_batch_fit with, in every round, the
// tile's top-k drawn candidates refined by Adam on the soft SSE and the refined codes
// re-scored exactly before anything is accepted.  Included by weft_kernels.hip inside
// its anonymous namespace, after the batched fit, the soft surrogate and the split scorer.
//
// Per tile and round (one 256-thread workgroup): class-sort and score the n_cand drawn
// candidates (as weft_batch_fit), keep each candidate's (MSE bits, index) key in LDS, take
// the top k (k workgroup-wide minima for k <= 8, else a bitonic sort), refine them (cpw candidates per wave, the
// arithmetic of weft_refine_kernel), score the k refined codes and the unrefined best
// exactly (each candidate's pixels split over lanes along the pairwise tree), take the
// first minimum and accept it if it beats the tile's objective.  The result equals
// that of the stepwise refined round (weft_fit_score + weft_topk + weft_refine +
// weft_fit_score_split + weft_fit_select) with the same work split.

#ifndef WEFT_BATCH_REFINE_UNROLL
#define WEFT_BATCH_REFINE_UNROLL 1  // pixels in flight in the refinement (registers are shared with scoring)
#endif

struct RefineShared {
    int topk[64];
    WeftCode codes[65];
    unsigned long long wmin[16];  // per-wave minima of the top-k rounds (up to 1024 threads)
    unsigned long long last;
};

// Adam on the tile's top k (64 / cpw lanes per candidate, min(cpw, k) candidates per
// wave; with cpw > k the other lanes idle, so the soft sums run in the order of the
// stepwise refinement's shared workgroups); writes rs.codes[1 + j] and the unrefined best
// as rs.codes[0].
template <int TS>
__device__ __forceinline__ void refine_phase(const WeftBatchRefineArgs& R, uint64_t dbase, const float* T,
                                          const float* cvp, RefineShared& rs) {
    constexpr int CLS = TS == 8 ? 0 : (TS == 16 ? 1 : (TS == 32 ? 2 : 3));
    const int tid = threadIdx.x, wave = tid >> 6, nwaves = blockDim.x >> 6, lane = tid & 63;
    const int k = R.o.k;
    const int cpw = R.o.cpw_by_size[CLS];
    const int cpt = cpw < k ? cpw : k;
    const int Ln = 64 / cpw;
    const int groups = k / cpt;
#pragma unroll 1
    for (int g = wave; g < groups; g += nwaves) {
        const int q = lane / Ln, slice = lane - q * Ln;
        if (q >= cpt) continue;  // idle lanes (whole candidate groups)
        const int j = g * cpt + q;
        SoftCand C = soft_from_code(draw_code(dbase, (uint32_t)rs.topk[j], TS));
        refine_candidate<WEFT_BATCH_REFINE_UNROLL, false>(R.o, T, cvp, TS, slice, Ln, C);
        if (slice == 0) rs.codes[1 + j] = soft_to_code(C, TS);
    }
    if (tid == 0) rs.codes[0] = draw_code(dbase, (uint32_t)rs.topk[0], TS);
}

// Exact scores of rs.codes[0 .. k] (pixels split over lanes); this lane's first minimum
// of (objective, slot) in (bo, bi).
template <int TS>
__device__ __forceinline__ void rescore_phase(const WeftBatchRefineArgs& R, int t, const float* T, const float* cvp,
                                           const RefineShared& rs, double& bo, int& bi) {
    constexpr int CLS = TS == 8 ? 0 : (TS == 16 ? 1 : (TS == 32 ? 2 : 3));
    const WeftFitState& st = R.b.st;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int K1 = R.o.k + 1;
    const int Ls = R.split_lanes[CLS];
    const int per = nthr / Ls;
#pragma unroll 1
    for (int c0 = 0; c0 < K1; c0 += per) {
        const int c = c0 + tid / Ls, sl = tid - (tid / Ls) * Ls;
        const bool live = c < K1;
        WeftPrim P;
        P.kind = -1;
        P.alpha = 0.0f;
        P.color0[0] = P.color0[1] = P.color0[2] = 0.0f;
        if (live) P = code_to_prim(rs.codes[c]);
        const float sse = split_sse<TS, true>(P, cvp, T, Ls, sl);
        if (live) {
            const float mse = sse / (float)(3 * TS * TS);
            argmin_pair(bo, bi, objective(mse, st.d_tile_bits[t] + prim_bytes(P.kind) * 8, st.lam, st.obj_mode), c);
        }
    }
}

// Inlined into the kernel on purpose: a call would pass the kernel arguments by address,
// which copies the whole argument block into scratch memory.
template <int TS>
__device__ __forceinline__ void batch_refine_tile(const WeftBatchRefineArgs& R, int t, float4* LA, float* LT,
                                  unsigned long long* keys, BatchShared& sh, RefineShared& rs) {
    const WeftBatchArgs& A = R.b;
    constexpr int ROWS = Band<TS>::ROWS;
    constexpr int NB = TS / ROWS;
    constexpr int BPX = ROWS * TS;
    constexpr int CLS = TS == 8 ? 0 : (TS == 16 ? 1 : (TS == 32 ? 2 : 3));
    const WeftFitState& st = A.st;
    const int tid = threadIdx.x, nthr = blockDim.x, wave = tid >> 6, nwaves = nthr >> 6, lane = tid & 63;
    const int64_t base = st.d_tile_pix[t] * 3;
    float* cvp = st.d_canvas + base;
    const float* T = st.d_target + base;
    const int n = A.cd.n_cand;
    const int k = R.o.k;
    const int K1 = k + 1;
    const int np2 = R.n_pow2;
    float* LAf = reinterpret_cast<float*>(LA);
    const LdsBand L{LA, LT};
    const int64_t T_all = st.n_tiles;
#pragma unroll 1
    for (int s = 0; s < A.n_steps; ++s) {
        const int64_t rec = (int64_t)s * T_all + t;
        if (st.d_tile_nsel[t] >= st.depth) {  // uniform: written by this workgroup only
            if (tid == 0) {
                A.win_idx[rec] = -1;
                A.win_obj[rec] = __builtin_huge_val();
                A.win_acc[rec] = 0;
            }
            continue;
        }
        const uint32_t step = A.step0 + (uint32_t)s;
        const uint64_t dbase = draw_base(A.cd.seed, step, (uint32_t)t);
        // ---- exact scores of the drawn candidates -> keys ----
        bool staged = false;
#pragma unroll 1
        for (int c0 = 0; c0 < n; c0 += nthr) {
            const int slot = c0 + tid;
            const int ks = slot < n ? drawn_sorted_index(dbase, n, slot) : -1;
            WeftPrim P;
            P.kind = -1;
            P.alpha = 0.0f;
            P.color0[0] = P.color0[1] = P.color0[2] = 0.0f;
            int mycls = CL_NONE;
            if (ks >= 0) {
                P = code_to_prim(draw_code(dbase, (uint32_t)ks, TS));
                mycls = class_of_kind(P.kind);
            }
            CandConsts cc;
            mycls = cand_consts(P, mycls, cc);
            PwState pw;
            pw_init(pw);
#pragma unroll 1
            for (int y0 = 0; y0 < TS; y0 += ROWS) {
                if (NB > 1 || !staged) {
                    __syncthreads();
                    const int e0 = 3 * y0 * TS;
                    for (int i = tid; i < 3 * BPX; i += nthr) {
                        const int p = i / 3, ch = i - 3 * (i / 3);
                        const float cv = cvp[e0 + i];
                        const float tv = T[e0 + i];
                        LAf[8 * p + ch] = cv;
                        LAf[8 * p + 4 + ch] = clip01(cv);
                        if (ch < 2) LAf[8 * p + 3 + 4 * ch] = tv;
                        else LT[p] = tv;
                    }
                    __syncthreads();
                    staged = true;
                }
                score_band<TS, false>(mycls, cc, P, L, pw, y0, y0 + ROWS);
            }
            if (ks >= 0) {
                const float mse = pw.template result<TS>() / (float)(3 * TS * TS);
                keys[ks] = ((unsigned long long)__float_as_uint(mse) << 32) | (unsigned int)ks;
            }
        }
        for (int i = n + tid; i < np2; i += nthr) keys[i] = ~0ull;
        __syncthreads();
        // ---- top k by (MSE, index) (keys are distinct): for k <= 8, k workgroup-wide
        // minima, each above the last; above that, a bitonic sort of the keys ----
        if (k <= 8) {
            unsigned long long last = 0ull;
            for (int r = 0; r < k; ++r) {
                unsigned long long m = ~0ull;
                for (int i = tid; i < n; i += nthr) {
                    const unsigned long long kk = keys[i];
                    if ((r == 0 || kk > last) && kk < m) m = kk;
                }
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                    const unsigned long long o = __shfl_xor(m, off, 64);
                    m = o < m ? o : m;
                }
                if (lane == 0) rs.wmin[wave] = m;
                __syncthreads();
                if (tid == 0) {
                    unsigned long long mm = rs.wmin[0];
                    for (int ww = 1; ww < nwaves; ++ww) mm = rs.wmin[ww] < mm ? rs.wmin[ww] : mm;
                    rs.topk[r] = (int)(mm & 0xFFFFFFFFull);
                    rs.last = mm;
                }
                __syncthreads();
                last = rs.last;
            }
        } else {
            for (int size = 2; size <= np2; size <<= 1) {
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    for (int i = tid; i < (np2 >> 1); i += nthr) {
                        const int lo = 2 * i - (i & (stride - 1));
                        const int hi = lo + stride;
                        const bool asc = (lo & size) == 0;
                        const unsigned long long a = keys[lo], b = keys[hi];
                        if ((a > b) == asc) {
                            keys[lo] = b;
                            keys[hi] = a;
                        }
                    }
                    __syncthreads();
                }
            }
            if (tid < k) rs.topk[tid] = (int)(keys[tid] & 0xFFFFFFFFull);
            __syncthreads();
        }
        // ---- Adam on the top k; exact scores of the best drawn and the refined ----
        refine_phase<TS>(R, dbase, T, cvp, rs);
        __syncthreads();
        double bo = __builtin_huge_val();
        int bi = 0x7fffffff;
        rescore_phase<TS>(R, t, T, cvp, rs, bo, bi);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double o2 = __shfl_xor(bo, off, 64);
            const int i2 = __shfl_xor(bi, off, 64);
            argmin_pair(bo, bi, o2, i2);
        }
        if (lane == 0) {
            sh.red_o[wave] = bo;
            sh.red_i[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double o = sh.red_o[0];
            int i = sh.red_i[0];
            for (int ww = 1; ww < nwaves; ++ww) argmin_pair(o, i, sh.red_o[ww], sh.red_i[ww]);
            const bool found = i != 0x7fffffff;
            sh.best = found ? i : -1;
            sh.best_obj = o;
            sh.acc = found && (o < st.d_tile_obj[t]);
        }
        __syncthreads();
        const int win = sh.best;
        const int acc = sh.acc;
        if (acc) {
            const WeftCode code = rs.codes[win];
            const WeftPrim Pw = code_to_prim(code);
            for (int p = tid; p < TS * TS; p += nthr) {
                const int y = p / TS, x = p - y * TS;
                float r = cvp[3 * p], g = cvp[3 * p + 1], b = cvp[3 * p + 2];
                composite_px(Pw, (float)x + 0.5f, (float)y + 0.5f, r, g, b);
                cvp[3 * p] = r;
                cvp[3 * p + 1] = g;
                cvp[3 * p + 2] = b;
            }
            if (tid == 0) {
                if (A.win_code) A.win_code[rec] = code;
                st.d_tile_nsel[t] += 1;
                st.d_tile_bits[t] += prim_bytes(Pw.kind) * 8;
                st.d_tile_obj[t] = sh.best_obj;
            }
        }
        if (tid == 0) {
            A.win_idx[rec] = win;
            A.win_obj[rec] = sh.best_obj;
            A.win_acc[rec] = acc;
        }
        __threadfence_block();
        __syncthreads();  // canvas and state are settled before the next round reads them
    }
}

__device__ __forceinline__ void batch_refine_body(const WeftBatchRefineArgs& R) {
    extern __shared__ float4 weft_smem[];
    __shared__ BatchShared sh;
    __shared__ RefineShared rs;
    const WeftBatchArgs& A = R.b;
    float4* LA = weft_smem;                                             // [2 * 512]
    float* LT = reinterpret_cast<float*>(LA + 2 * 512);                 // [512]
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(LT + 512);  // [n_pow2]
    const int tid = threadIdx.x;
    if (tid == 0 && A.t_start) atomicMin(A.t_start, (unsigned long long)__builtin_amdgcn_s_memrealtime());
#pragma unroll 1
    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(A.queue, 1);
            sh.tile = q < A.n_queue ? A.tile_order[q] : -1;
        }
        __syncthreads();
        const int t = sh.tile;
        __syncthreads();
        if (t < 0) break;
        switch (A.st.d_tile_ts[t]) {
        case 8: batch_refine_tile<8>(R, t, LA, LT, keys, sh, rs); break;
        case 16: batch_refine_tile<16>(R, t, LA, LT, keys, sh, rs); break;
        case 32: batch_refine_tile<32>(R, t, LA, LT, keys, sh, rs); break;
        case 64: batch_refine_tile<64>(R, t, LA, LT, keys, sh, rs); break;
        default: break;
        }
        if (tid == 0 && A.image_left) {
            const int img = A.tile_image[t];
            if (atomicSub(&A.image_left[img], 1) == 1)
                A.image_done[img] = (unsigned long long)__builtin_amdgcn_s_memrealtime();
        }
    }
}
