// Stage-3 template fit: per tile, 18 two-bucket pixel templates (constant, 8 angle
// splits, 3 line bands, 3 corner triangles, 3 quadratic-curve bands), the mean colour of
// each bucket and the MSE of the two-colour prediction; the best template per tile.
// Included by weft_kernels.hip inside its anonymous namespace.
//
// Reference: src/weft/gpu_encoder.py:42-251 (weft_model_bucket, weft_tile_fit_candidates).
// The bucket of a pixel depends only on its position in the tile, so the 18 buckets of
// every position are computed once per tile size into a bit table (bit m = bucket of
// template m).  The fit gives each (tile, template) one thread that forms the bucket sums
// and the squared errors with float64 accumulators in position order: the result does not
// depend on scheduling (the reference's float32 shared-memory atomics do).

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

// The fit: one thread per (tile, template), 14 tiles of 18 templates per 256-thread
// workgroup.  Every thread walks its tile's positions in order 0 .. n-1 twice: the bucket
// sums (float64), then the squared errors of the two-colour prediction (float32 per pixel,
// accumulated in float64) -- the sequential order of gpu_encoder.py's per-template loops as
// the oracle restates them, so sums, means and MSE do not depend on scheduling.  Positions
// are staged through LDS in bands of 64 for all 14 tiles (float64 for the sums, float32 for
// the errors) with the band's mask words beside them.
constexpr int TMPL_WG = 256;
constexpr int TMPL_SLOTS = 14;                       // tiles per workgroup
constexpr int TMPL_BAND = 64;                        // positions per staged band
constexpr int TMPL_ST64 = TMPL_BAND * 3 + 1;         // slot stride of the float64 band (odd: no bank conflicts)
constexpr int TMPL_ST32 = TMPL_BAND * 3 + 1;         // slot stride of the float32 band
static_assert(TMPL_SLOTS * TMPL_MODELS <= TMPL_WG, "one thread per (tile, template)");

struct TmplShared {
    union {
        double v64[TMPL_SLOTS * TMPL_ST64];
        float v32[TMPL_SLOTS * TMPL_ST32];
    };
    uint32_t mask[TMPL_BAND];
    float mse[TMPL_SLOTS][TMPL_MODELS];
    int x0[TMPL_SLOTS], y0[TMPL_SLOTS];
};

// A band of positions [p0, p0 + cnt) of the workgroup's tiles, fetched into registers (so
// that the next band's loads are in flight while the current one is scored) and then
// written to LDS.  Edge pixels repeat the last row / column, as the reference kernel reads
// them.
constexpr int TMPL_PER = (TMPL_SLOTS * TMPL_BAND + TMPL_WG - 1) / TMPL_WG;  // (slot, position) pairs per thread

struct TmplBand {
    float v[TMPL_PER][3];
};

__device__ __forceinline__ void tmpl_fetch(const WeftTemplateFitArgs& A, const TmplShared& S, int tile0, int n_tiles,
                                           int p0, int cnt, float inv_ts, TmplBand& B) {
    const int ts = A.tile_size;
#pragma unroll
    for (int j = 0; j < TMPL_PER; ++j) {
        const int e = threadIdx.x + j * TMPL_WG;
        const int sl = e / TMPL_BAND, k = e - sl * TMPL_BAND;
        if (sl < TMPL_SLOTS && k < cnt && tile0 + sl < n_tiles) {
            const int i = p0 + k;
            // i / ts exactly: (i + 1/2) / ts lies at least 1/128 from an integer for ts <= 64,
            // and the float32 quotient is within 5e-4 of it for i < 4096
            const int ly = (int)(((float)i + 0.5f) * inv_ts);
            const int lx = i - ly * ts;
            const int x = min(S.x0[sl] + lx, A.width - 1);
            const int y = min(S.y0[sl] + ly, A.height - 1);
            const float* src = A.image + ((int64_t)y * A.width + x) * 3;
            B.v[j][0] = src[0];
            B.v[j][1] = src[1];
            B.v[j][2] = src[2];
        }
    }
}

// Write a fetched band (float64 for the sums, float32 for the errors) and the band's mask
// words; returns whether this thread wrote a non-finite value.
template <bool F64>
__device__ __forceinline__ int tmpl_store(const WeftTemplateFitArgs& A, TmplShared& S, int tile0, int n_tiles, int p0,
                                          int cnt, const TmplBand& B) {
    int nonfinite = 0;
#pragma unroll
    for (int j = 0; j < TMPL_PER; ++j) {
        const int e = threadIdx.x + j * TMPL_WG;
        const int sl = e / TMPL_BAND, k = e - sl * TMPL_BAND;
        if (sl < TMPL_SLOTS && k < cnt && tile0 + sl < n_tiles) {
            const float r = B.v[j][0], g = B.v[j][1], b = B.v[j][2];
            nonfinite |= !(isfinite(r) && isfinite(g) && isfinite(b));
            if constexpr (F64) {
                double* d = S.v64 + sl * TMPL_ST64 + k * 3;
                d[0] = (double)r;
                d[1] = (double)g;
                d[2] = (double)b;
            } else {
                float* d = S.v32 + sl * TMPL_ST32 + k * 3;
                d[0] = r;
                d[1] = g;
                d[2] = b;
            }
        }
    }
    for (int k = threadIdx.x; k < cnt; k += TMPL_WG) S.mask[k] = A.mask[p0 + k];
    return nonfinite;
}

__device__ __forceinline__ double tmpl_hi(uint32_t hi) { return __longlong_as_double((long long)((uint64_t)hi << 32)); }

__device__ void tmpl_fit_body(const WeftTemplateFitArgs& A) {
    __shared__ TmplShared S;
    const int tid = threadIdx.x;
    const int ts = A.tile_size;
    const int n = ts * ts;
    const int n_tiles = A.tile_cols * A.tile_rows;
    const int tile0 = blockIdx.x * TMPL_SLOTS;
    const int slot = tid / TMPL_MODELS;
    const int m = tid - slot * TMPL_MODELS;
    const int tile = tile0 + slot;
    const bool valid = tid < TMPL_SLOTS * TMPL_MODELS && tile < n_tiles;
    const float inv_ts = 1.0f / (float)ts;
    if (tid < TMPL_SLOTS) {
        const int t = tile0 + tid;
        S.x0[tid] = (t % A.tile_cols) * ts;
        S.y0[tid] = (t / A.tile_cols) * ts;
    }
    __syncthreads();

    // bucket sums in position order; b * v + s with b = 0 or 1 is the exact conditional add
    // for finite v, and bands holding inf / NaN take selects so that a non-finite value
    // stays in its own bucket
    double s0[3] = {0.0, 0.0, 0.0}, s1[3] = {0.0, 0.0, 0.0};
    TmplBand B;
    tmpl_fetch(A, S, tile0, n_tiles, 0, min(TMPL_BAND, n), inv_ts, B);
    for (int p0 = 0; p0 < n; p0 += TMPL_BAND) {
        const int cnt = min(TMPL_BAND, n - p0);
        const int nf = __syncthreads_or(tmpl_store<true>(A, S, tile0, n_tiles, p0, cnt, B));
        // the next band (or the error pass's first) is fetched while this one is summed
        {
            const int q0 = p0 + TMPL_BAND < n ? p0 + TMPL_BAND : 0;
            tmpl_fetch(A, S, tile0, n_tiles, q0, min(TMPL_BAND, n - q0), inv_ts, B);
        }
        if (valid) {
            const double* v = S.v64 + slot * TMPL_ST64;
            if (!nf) {
#pragma unroll 8
                for (int k = 0; k < cnt; ++k) {
                    // bit m of the mask word sign-extended: 0 or all ones
                    const uint32_t hi1 = (uint32_t)__builtin_amdgcn_sbfe((int)S.mask[k], (unsigned)m, 1u) & 0x3FF00000u;
                    const double b1 = tmpl_hi(hi1), b0 = tmpl_hi(hi1 ^ 0x3FF00000u);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        s0[c] = __builtin_fma(b0, v[3 * k + c], s0[c]);
                        s1[c] = __builtin_fma(b1, v[3 * k + c], s1[c]);
                    }
                }
            } else {
                for (int k = 0; k < cnt; ++k) {
                    const bool o = (S.mask[k] >> m) & 1u;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        s0[c] += o ? 0.0 : v[3 * k + c];
                        s1[c] += o ? v[3 * k + c] : 0.0;
                    }
                }
            }
        }
        __syncthreads();
    }

    // bucket means: (float)sum * (1 / max(count, 1)), the constant template's buckets alike
    float a[3], b[3];
    {
        const int c0 = valid ? (int)A.mask[n + m] : 1, c1 = n - c0;
        const float i0 = 1.0f / (float)max(c0, 1), i1 = 1.0f / (float)max(c1, 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a[c] = (float)s0[c] * i0;
            b[c] = m == 0 ? a[c] : (float)s1[c] * i1;
        }
    }

    // squared errors in position order: e = dr^2, e += dg^2, e += db^2 in float32, summed in float64
    const uint32_t ua[3] = {__float_as_uint(a[0]), __float_as_uint(a[1]), __float_as_uint(a[2])};
    const uint32_t ub[3] = {__float_as_uint(b[0]), __float_as_uint(b[1]), __float_as_uint(b[2])};
    double err = 0.0;
    for (int p0 = 0; p0 < n; p0 += TMPL_BAND) {
        const int cnt = min(TMPL_BAND, n - p0);
        tmpl_store<false>(A, S, tile0, n_tiles, p0, cnt, B);
        __syncthreads();
        if (p0 + TMPL_BAND < n)
            tmpl_fetch(A, S, tile0, n_tiles, p0 + TMPL_BAND, min(TMPL_BAND, n - p0 - TMPL_BAND), inv_ts, B);
        if (valid) {
            const float* v = S.v32 + slot * TMPL_ST32;
#pragma unroll 4
            for (int k = 0; k < cnt; ++k) {
                // the bucket's colour by a bitfield select on the sign-extended bucket bit
                const uint32_t o = (uint32_t)__builtin_amdgcn_sbfe((int)S.mask[k], (unsigned)m, 1u);
                const float dr = v[3 * k] - __uint_as_float((o & ub[0]) | (~o & ua[0]));
                const float dg = v[3 * k + 1] - __uint_as_float((o & ub[1]) | (~o & ua[1]));
                const float db = v[3 * k + 2] - __uint_as_float((o & ub[2]) | (~o & ua[2]));
                float e = dr * dr;
                e = e + dg * dg;
                e = e + db * db;
                err += (double)e;
            }
        }
        __syncthreads();
    }

    const float mse = (float)err * (1.0f / (float)n);
    if (valid) {
        S.mse[slot][m] = mse;
        const int64_t o = (int64_t)tile * TMPL_MODELS + m;
        A.out_mse[o] = mse;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            A.out_c0_all[o * 3 + c] = a[c];
            A.out_c1_all[o * 3 + c] = b[c];
        }
    }
    __syncthreads();
    if (valid) {
        // first least MSE (a NaN never replaces the current best)
        int best = 0;
        for (int q = 1; q < TMPL_MODELS; ++q)
            if (S.mse[slot][q] < S.mse[slot][best]) best = q;
        if (best == m) {
            A.out_model[tile] = m;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                A.out_c0[(int64_t)tile * 3 + c] = a[c];
                A.out_c1[(int64_t)tile * 3 + c] = b[c];
            }
        }
    }
}
