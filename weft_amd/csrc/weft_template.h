// Stage-3 template fit: per tile, 18 two-bucket pixel templates (constant, 8 angle
// splits, 3 line bands, 3 corner triangles, 3 quadratic-curve bands), the mean colour of
// each bucket and the MSE of the two-colour prediction; the best template per tile.
// Included by weft_kernels.hip inside its anonymous namespace.
//
// Reference: src/weft/gpu_encoder.py:42-251 (weft_model_bucket, weft_tile_fit_candidates).
// The bucket of a pixel depends only on its position in the tile, so the 18 buckets of
// every position are computed once per tile size into a bit table (bit m = bucket of
// template m).  The fit reads a tile into LDS and forms the bucket sums and squared errors
// with float64 accumulators in a fixed order: the result does not depend on scheduling
// (the reference's float32 shared-memory atomics do).  The colour sums are exact when each
// channel value of a tile is 0 or in [2^-17, 1] (a sum of at most 4096 such values needs
// no more than 53 significant bits).

constexpr int TMPL_MODELS = 18;
constexpr int TMPL_MAX_TS = 64;

__device__ __forceinline__ float tmpl_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float tmpl_max(float a, float b) { return a > b ? a : b; }

__device__ __forceinline__ float tmpl_seg_dist(float px, float py, float x0, float y0, float x1, float y1) {
    const float vx = x1 - x0;
    const float vy = y1 - y0;
    const float den = vx * vx + vy * vy + 1e-8f;
    float t = ((px - x0) * vx + (py - y0) * vy) / den;
    t = tmpl_max(0.0f, tmpl_min(1.0f, t));
    const float cx = x0 + t * vx;
    const float cy = y0 + t * vy;
    const float dx = px - cx;
    const float dy = py - cy;
    return sqrtf(dx * dx + dy * dy);
}

__device__ __forceinline__ bool tmpl_inside_tri(float px, float py, float x0, float y0, float x1, float y1, float x2,
                                                float y2) {
    const float den = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2);
    if (fabsf(den) < 1e-6f) return false;
    const float a = ((y1 - y2) * (px - x2) + (x2 - x1) * (py - y2)) / den;
    const float b = ((y2 - y0) * (px - x2) + (x0 - x2) * (py - y2)) / den;
    const float c = 1.0f - a - b;
    return a >= 0.0f && b >= 0.0f && c >= 0.0f;
}

// Distance to the quadratic Bezier (x0,y0)-(cx,cy)-(x1,y1) sampled as a 12-segment polyline.
__device__ __forceinline__ float tmpl_curve_dist(float px, float py, float x0, float y0, float cx, float cy, float x1,
                                                 float y1) {
    float best = 1e9f;
    float ax = x0, ay = y0;
    for (int i = 1; i <= 12; ++i) {
        const float t = (float)i / 12.0f;
        const float u = 1.0f - t;
        const float qx = u * u * x0 + 2.0f * u * t * cx + t * t * x1;
        const float qy = u * u * y0 + 2.0f * u * t * cy + t * t * y1;
        best = tmpl_min(best, tmpl_seg_dist(px, py, ax, ay, qx, qy));
        ax = qx;
        ay = qy;
    }
    return best;
}

// Bucket (0 or 1) of integer tile position (lx, ly) under template m.  The pixel is sampled
// at its integer corner, as the reference kernel does.  The angle-split normal is the
// correctly rounded float32 cos / sin of the float32 angle.
__device__ int tmpl_bucket(int m, int lx, int ly, int tile_size) {
    const float x = (float)lx, y = (float)ly;
    const float ts = (float)(tile_size - 1);
    const float cx = ts * 0.5f, cy = ts * 0.5f;
    if (m == 0) return 0;
    if (m <= 8) {
        const float angle = (float)(m - 1) * (3.14159265358979323846f / 8.0f);
        const float nx = (float)cos((double)angle);
        const float ny = (float)sin((double)angle);
        const float proj = (x - cx) * nx + (y - cy) * ny;
        return proj < 0.0f ? 0 : 1;
    }
    if (m == 9) return fabsf(x - cx) <= 1.5f ? 0 : 1;
    if (m == 10) return fabsf(y - cy) <= 1.5f ? 0 : 1;
    if (m == 11) return fabsf(y - x) <= 1.5f ? 0 : 1;
    if (m == 12) return tmpl_inside_tri(x, y, 0.0f, 0.0f, ts, 0.0f, 0.0f, ts) ? 0 : 1;
    if (m == 13) return tmpl_inside_tri(x, y, ts, 0.0f, ts, ts, 0.0f, 0.0f) ? 0 : 1;
    if (m == 14) return tmpl_inside_tri(x, y, 0.0f, ts, ts, ts, 0.0f, 0.0f) ? 0 : 1;
    float d;
    if (m == 15) d = tmpl_curve_dist(x, y, 0.0f, cy, cx, 0.0f, ts, cy);
    else if (m == 16) d = tmpl_curve_dist(x, y, 0.0f, cy, cx, ts, ts, cy);
    else d = tmpl_curve_dist(x, y, cx, 0.0f, 0.0f, cy, cx, ts);
    return d <= 1.25f ? 0 : 1;
}

// One 256-thread workgroup: the bit table, then the bucket-0 count of every template.
__device__ void tmpl_masks_body(const WeftTemplateMaskArgs& A) {
    __shared__ int cnt[TMPL_MODELS][256];
    const int ts = A.tile_size;
    const int n = ts * ts;
    const int tid = threadIdx.x;
    int c0[TMPL_MODELS];
#pragma unroll
    for (int m = 0; m < TMPL_MODELS; ++m) c0[m] = 0;
    for (int p = tid; p < n; p += blockDim.x) {
        const int ly = p / ts, lx = p - ly * ts;
        uint32_t bits = 0;
        for (int m = 1; m < TMPL_MODELS; ++m) bits |= (uint32_t)tmpl_bucket(m, lx, ly, ts) << m;
        A.mask[p] = bits;
#pragma unroll
        for (int m = 0; m < TMPL_MODELS; ++m) c0[m] += ((bits >> m) & 1u) ? 0 : 1;
    }
#pragma unroll
    for (int m = 0; m < TMPL_MODELS; ++m) cnt[m][tid] = c0[m];
    __syncthreads();
    if (tid < TMPL_MODELS) {
        int c = 0;
        for (int k = 0; k < (int)blockDim.x; ++k) c += cnt[tid][k];
        A.mask[n + tid] = (uint32_t)c;
    }
}

// Dynamic LDS: three channel planes of n = ts^2 floats (each n + 16 apart) and n mask words.
__host__ __device__ constexpr int tmpl_plane_stride(int n) { return n + 16; }

// Both passes: thread = (pixel segment, group of 6 templates), 84 segments x 3 groups.  A
// pixel read (three values and a mask word) feeds 36 float64 sums in the sum pass and 6 in
// the error pass.
constexpr int TMPL_GROUP = 6;
constexpr int TMPL_GROUPS = TMPL_MODELS / TMPL_GROUP;
constexpr int TMPL_SEGS = 84;
constexpr int TMPL_THREADS = TMPL_SEGS * TMPL_GROUPS;   // 252
constexpr int TMPL_CH_SUMS = TMPL_MODELS * 2;           // (template, bucket) sums of one channel
constexpr int TMPL_L1 = TMPL_THREADS / TMPL_CH_SUMS;    // 7 first-level partial sums
constexpr int TMPL_L1_SPAN = TMPL_SEGS / TMPL_L1;        // of 12 segments each

struct TmplShared {
    double part[TMPL_SEGS * TMPL_CH_SUMS];  // per-segment partials (one channel at a time)
    double part2[TMPL_L1][TMPL_CH_SUMS];
    double sums[3][TMPL_CH_SUMS];
    float col[TMPL_MODELS][2][3];
    float mse[TMPL_MODELS];
};
static_assert(TMPL_L1 * TMPL_CH_SUMS == TMPL_THREADS && TMPL_L1 * TMPL_L1_SPAN == TMPL_SEGS, "combine shape");
static_assert(TMPL_SEGS * TMPL_MODELS <= TMPL_SEGS * TMPL_CH_SUMS, "error partials reuse the buffer");

// Sum pass of one thread.  FMA: every value is finite, so b * v + s (b = 0 or 1) is the
// exact b ? s + v : s with one instruction; otherwise the selects keep an infinite or NaN
// value out of the other bucket.
template <bool FMA>
__device__ __forceinline__ void tmpl_sum_pass(const float* px0, int PS, const uint32_t* mask, int p0, int p1, int g,
                                              double (&s)[TMPL_GROUP][2][3]) {
#pragma unroll
    for (int j = 0; j < TMPL_GROUP; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) s[j][0][c] = s[j][1][c] = 0.0;
    for (int p = p0; p < p1; ++p) {
        const double v[3] = {(double)px0[p], (double)px0[PS + p], (double)px0[2 * PS + p]};
        const uint32_t mk = mask[p] >> (g * TMPL_GROUP);
#pragma unroll
        for (int j = 0; j < TMPL_GROUP; ++j) {
            const uint32_t bit = (mk >> j) & 1u;
            if constexpr (FMA) {
                const double b1 = (double)bit, b0 = 1.0 - b1;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    s[j][0][c] = __builtin_fma(b0, v[c], s[j][0][c]);
                    s[j][1][c] = __builtin_fma(b1, v[c], s[j][1][c]);
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    s[j][0][c] += bit ? 0.0 : v[c];
                    s[j][1][c] += bit ? v[c] : 0.0;
                }
            }
        }
    }
}

// One 256-thread workgroup per tile.  Every float64 sum is taken in a fixed order.
__device__ void tmpl_fit_body(const WeftTemplateFitArgs& A) {
    __shared__ TmplShared S;
    extern __shared__ float4 weft_smem[];
    const int tile = blockIdx.x;
    const int tid = threadIdx.x;
    const int ts = A.tile_size;
    const int n = ts * ts;
    float* const px0 = reinterpret_cast<float*>(weft_smem);
    const int PS = tmpl_plane_stride(n);
    uint32_t* const mask = reinterpret_cast<uint32_t*>(px0 + 3 * PS);
    const int x0 = (tile % A.tile_cols) * ts;
    const int y0 = (tile / A.tile_cols) * ts;
    int nonfinite = 0;
    for (int i = tid; i < n; i += blockDim.x) {
        const int ly = i / ts, lx = i - ly * ts;
        const int x = min(x0 + lx, A.width - 1);
        const int y = min(y0 + ly, A.height - 1);
        const float* src = A.image + ((int64_t)y * A.width + x) * 3;
        const float r = src[0], g = src[1], b = src[2];
        nonfinite |= !(isfinite(r) && isfinite(g) && isfinite(b));
        px0[i] = r;
        px0[PS + i] = g;
        px0[2 * PS + i] = b;
        mask[i] = A.mask[i];
    }
    const bool all_finite = !__syncthreads_or(nonfinite);

    const bool worker = tid < TMPL_THREADS;
    const int seg = tid / TMPL_GROUPS;
    const int g = tid - seg * TMPL_GROUPS;
    const int p0 = (seg * n) / TMPL_SEGS, p1 = ((seg + 1) * n) / TMPL_SEGS;

    // bucket sums, then per channel: segment partials -> 7 partials -> total
    double s[TMPL_GROUP][2][3];
    if (worker) {
        if (all_finite) tmpl_sum_pass<true>(px0, PS, mask, p0, p1, g, s);
        else tmpl_sum_pass<false>(px0, PS, mask, p0, p1, g, s);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (worker) {
#pragma unroll
            for (int j = 0; j < TMPL_GROUP; ++j)
#pragma unroll
                for (int b = 0; b < 2; ++b) S.part[seg * TMPL_CH_SUMS + (g * TMPL_GROUP + j) * 2 + b] = s[j][b][c];
        }
        __syncthreads();
        if (worker) {
            const int q = tid % TMPL_CH_SUMS, k = tid / TMPL_CH_SUMS;
            double v = 0.0;
            for (int i = 0; i < TMPL_L1_SPAN; ++i) v += S.part[(k * TMPL_L1_SPAN + i) * TMPL_CH_SUMS + q];
            S.part2[k][q] = v;
        }
        __syncthreads();
        if (tid < TMPL_CH_SUMS) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < TMPL_L1; ++k) v += S.part2[k][tid];
            S.sums[c][tid] = v;
        }
    }
    __syncthreads();
    if (tid < TMPL_MODELS * 3) {
        const int m = tid / 3, c = tid - m * 3;
        const int c0 = (int)A.mask[n + m], c1 = n - c0;
        const float v0 = (float)S.sums[c][m * 2] * (1.0f / (float)max(c0, 1));
        S.col[m][0][c] = v0;
        S.col[m][1][c] = m == 0 ? v0 : (float)S.sums[c][m * 2 + 1] * (1.0f / (float)max(c1, 1));
    }
    __syncthreads();

    // squared error of each template's two-colour prediction
    if (worker) {
        float a[TMPL_GROUP][3], b[TMPL_GROUP][3];
        double acc[TMPL_GROUP];
#pragma unroll
        for (int j = 0; j < TMPL_GROUP; ++j) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                a[j][c] = S.col[g * TMPL_GROUP + j][0][c];
                b[j][c] = S.col[g * TMPL_GROUP + j][1][c];
            }
            acc[j] = 0.0;
        }
        for (int p = p0; p < p1; ++p) {
            const float r = px0[p], gg = px0[PS + p], bb = px0[2 * PS + p];
            const uint32_t mk = mask[p] >> (g * TMPL_GROUP);
#pragma unroll
            for (int j = 0; j < TMPL_GROUP; ++j) {
                const bool o = (mk >> j) & 1u;
                const float dr = r - (o ? b[j][0] : a[j][0]);
                const float dg = gg - (o ? b[j][1] : a[j][1]);
                const float db = bb - (o ? b[j][2] : a[j][2]);
                acc[j] += (double)(dr * dr + dg * dg + db * db);
            }
        }
#pragma unroll
        for (int j = 0; j < TMPL_GROUP; ++j) S.part[seg * TMPL_MODELS + g * TMPL_GROUP + j] = acc[j];
    }
    __syncthreads();
    if (tid < TMPL_L1 * TMPL_MODELS) {  // 7 x 18: 12 segments each
        const int k = tid / TMPL_MODELS, m = tid - k * TMPL_MODELS;
        double e = 0.0;
        for (int i = 0; i < TMPL_L1_SPAN; ++i) e += S.part[(k * TMPL_L1_SPAN + i) * TMPL_MODELS + m];
        S.part2[k][m] = e;
    }
    __syncthreads();
    if (tid < TMPL_MODELS) {
        double e = 0.0;
#pragma unroll
        for (int k = 0; k < TMPL_L1; ++k) e += S.part2[k][tid];
        const float mse = (float)e * (1.0f / (float)n);
        S.mse[tid] = mse;
        const int64_t o = (int64_t)tile * TMPL_MODELS + tid;
        A.out_mse[o] = mse;
        for (int c = 0; c < 3; ++c) {
            A.out_c0_all[o * 3 + c] = S.col[tid][0][c];
            A.out_c1_all[o * 3 + c] = S.col[tid][1][c];
        }
    }
    __syncthreads();
    if (tid == 0) {
        int best = 0;
        float be = S.mse[0];
        for (int m = 0; m < TMPL_MODELS; ++m)
            if (S.mse[m] < be) {
                be = S.mse[m];
                best = m;
            }
        A.out_model[tile] = best;
        for (int c = 0; c < 3; ++c) {
            A.out_c0[(int64_t)tile * 3 + c] = S.col[best][0][c];
            A.out_c1[(int64_t)tile * 3 + c] = S.col[best][1][c];
        }
    }
}
