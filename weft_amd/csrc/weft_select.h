// Device-wide selection primitives behind the reference's CCCL helpers
// (src/weft/cuda_backend.py:80-181): first-index argmin of float64 values, exclusive scan
// of int64 values, compaction of flags to indices, and per-segment top-k.
// Included by weft_kernels.hip inside its anonymous namespace.
//
// Every result is deterministic.  Orders follow NumPy: NaN sorts after +inf and compares
// as the minimum in argmin (np.argmin returns the first NaN), -0.0 equals +0.0, and equal
// keys keep index order.

constexpr int SEL_BLOCK = 256;
constexpr int SEL_TILE = 2048;       // elements per workgroup in scan / count passes (8 per thread)
constexpr int SEL_SORT_BLOCK = 512;  // elements sorted in LDS by one 256-thread workgroup

// ---------------------------------------------------------------- argmin (first index)
// (value, index) pairs ordered as np.argmin sees them: NaN first, then by value, then index.
__device__ __forceinline__ bool sel_before(double a, int64_t ia, double b, int64_t ib) {
    const bool na = a != a, nb = b != b;
    if (na || nb) return na && (!nb || ia < ib);
    return a < b || (a == b && ia < ib);
}

__device__ __forceinline__ void sel_argmin_wave(double& v, int64_t& i) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double v2 = __shfl_xor(v, off, 64);
        const int64_t i2 = __shfl_xor(i, off, 64);
        if (sel_before(v2, i2, v, i)) {
            v = v2;
            i = i2;
        }
    }
}

// One workgroup reduces values [lo, hi) into (*out_v, *out_i); an empty range gives index -1.
__device__ void sel_argmin_range(const double* vals, const int64_t* idx_in, int64_t lo, int64_t hi, double* out_v,
                                 int64_t* out_i) {
    __shared__ double wv[SEL_BLOCK / 64];
    __shared__ int64_t wi[SEL_BLOCK / 64];
    double v = __builtin_huge_val();
    int64_t i = -1;
    for (int64_t p = lo + threadIdx.x; p < hi; p += blockDim.x) {
        const double x = vals[p];
        const int64_t ix = idx_in ? idx_in[p] : p;
        if (ix >= 0 && (i < 0 || sel_before(x, ix, v, i))) {
            v = x;
            i = ix;
        }
    }
    // empty lanes (index -1) must lose every comparison
    if (i < 0) {
        v = __builtin_huge_val();
        i = INT64_MAX;
    }
    sel_argmin_wave(v, i);
    if ((threadIdx.x & 63) == 0) {
        wv[threadIdx.x >> 6] = v;
        wi[threadIdx.x >> 6] = i;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            if (sel_before(wv[w], wi[w], v, i)) {
                v = wv[w];
                i = wi[w];
            }
        *out_v = v;
        *out_i = i == INT64_MAX ? -1 : i;
    }
}

__device__ void sel_argmin_partial_body(const WeftSelectArgs& A) {
    const int64_t per = (A.n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * per;
    const int64_t hi = lo + per < A.n ? lo + per : A.n;
    sel_argmin_range(A.f64_in, nullptr, lo, hi, A.work_f64 + blockIdx.x, A.work_i64 + blockIdx.x);
}

__device__ void sel_argmin_final_body(const WeftSelectArgs& A) {
    double v;
    int64_t i;
    sel_argmin_range(A.work_f64, A.work_i64, 0, A.n_blocks, &v, &i);
    if (threadIdx.x == 0) A.out_i64[0] = i;
}

// ---------------------------------------------------------------- exclusive scan (int64)
// Inclusive scan of one value per thread across the workgroup (Hillis-Steele in LDS).
__device__ __forceinline__ int64_t sel_block_inclusive(int64_t x, int64_t* buf) {
    const int t = threadIdx.x;
    buf[t] = x;
    __syncthreads();
    for (int off = 1; off < SEL_BLOCK; off <<= 1) {
        const int64_t y = t >= off ? buf[t - off] : 0;
        __syncthreads();
        buf[t] += y;
        __syncthreads();
    }
    return buf[t];
}

// Value of element p for the scan: int64 input, or 0/1 flags (compaction).
__device__ __forceinline__ int64_t sel_load(const WeftSelectArgs& A, int64_t p) {
    if (p >= A.n) return 0;
    return A.i64_in ? A.i64_in[p] : (int64_t)(A.u8_in[p] != 0);
}

// Pass 1: sum of every tile of SEL_TILE elements -> work_i64[tile].
__device__ void sel_tile_sums_body(const WeftSelectArgs& A) {
    __shared__ int64_t buf[SEL_BLOCK];
    const int64_t base = (int64_t)blockIdx.x * SEL_TILE + threadIdx.x * (SEL_TILE / SEL_BLOCK);
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SEL_TILE / SEL_BLOCK; ++k) s += sel_load(A, base + k);
    const int64_t tot = sel_block_inclusive(s, buf);
    if (threadIdx.x == SEL_BLOCK - 1) A.work_i64[blockIdx.x] = tot;
}

// Pass 2 (one workgroup): exclusive scan of the n_blocks tile sums in place; the grand
// total goes to *out_total.
__device__ void sel_scan_tiles_body(const WeftSelectArgs& A) {
    __shared__ int64_t buf[SEL_BLOCK];
    int64_t carry = 0;
    for (int64_t c0 = 0; c0 < A.n_blocks; c0 += SEL_BLOCK) {
        const int64_t p = c0 + threadIdx.x;
        const int64_t x = p < A.n_blocks ? A.work_i64[p] : 0;
        const int64_t inc = sel_block_inclusive(x, buf);
        if (p < A.n_blocks) A.work_i64[p] = carry + inc - x;
        carry += buf[SEL_BLOCK - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0 && A.out_total) *A.out_total = carry;
}

// Pass 3: each tile rescans its elements from its offset.  Scan: out[p] = exclusive
// prefix.  Compaction (u8_in): out[prefix] = p for every set flag.
__device__ void sel_scan_apply_body(const WeftSelectArgs& A) {
    __shared__ int64_t buf[SEL_BLOCK];
    constexpr int PER = SEL_TILE / SEL_BLOCK;
    const int64_t base = (int64_t)blockIdx.x * SEL_TILE + threadIdx.x * PER;
    int64_t x[PER];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        x[k] = sel_load(A, base + k);
        s += x[k];
    }
    int64_t run = A.work_i64[blockIdx.x] + sel_block_inclusive(s, buf) - s;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int64_t p = base + k;
        if (p < A.n) {
            if (A.i64_in) A.out_i64[p] = run;
            else if (x[k]) A.out_i64[run] = p;
        }
        run += x[k];
    }
}

// ---------------------------------------------------------------- segmented top-k
// Sort records (WeftSelRec: a, b, idx), compared on (a, b, idx) as unsigned integers.
using SelRec = WeftSelRec;

__device__ __forceinline__ bool sel_rec_less(const SelRec& x, const SelRec& y) {
    if (x.a != y.a) return x.a < y.a;
    if (x.b != y.b) return x.b < y.b;
    return x.idx < y.idx;
}

// float64 -> uint64 whose unsigned order is NumPy's ascending order (-0.0 = +0.0, NaN last).
__device__ __forceinline__ uint64_t sel_f64_key(double v) {
    if (v != v) return ~0ull;
    if (v == 0.0) v = 0.0;
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}

// Records for the first sort: (segment, score key, index); `largest` reverses both the
// score and the index order, so that ascending order equals NumPy's stable ascending order
// read backwards.  Slots [n, n_pad) get the largest record.
__device__ void sel_topk_pack_body(const WeftSelectArgs& A) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n_pad) return;
    SelRec r;
    r._pad = 0;
    if (p < A.n) {
        r.a = (uint64_t)A.i64_in[p] ^ (1ull << 63);
        const uint64_t key = sel_f64_key(A.f64_in[p]);
        r.b = A.largest ? ~key : key;
        r.idx = A.largest ? ~(uint32_t)p : (uint32_t)p;
    } else {
        r.a = ~0ull;
        r.b = ~0ull;
        r.idx = ~0u;
    }
    A.recs[p] = r;
}

// Bitonic sort of A.n_pad records (a power of two).  Local: every block of SEL_SORT_BLOCK
// records in LDS, either sorted from scratch (k_stage == 0) or merged for stage k_stage
// with every distance j < SEL_SORT_BLOCK.
__device__ void sel_bitonic_local_body(const WeftSelectArgs& A) {
    __shared__ SelRec s[SEL_SORT_BLOCK];
    const int64_t base = (int64_t)blockIdx.x * SEL_SORT_BLOCK;
    const int t = threadIdx.x;
    s[t] = A.recs[base + t];
    s[t + SEL_BLOCK] = A.recs[base + t + SEL_BLOCK];
    __syncthreads();
    const int64_t k0 = A.k_stage == 0 ? 2 : A.k_stage;
    const int64_t k1 = A.k_stage == 0 ? SEL_SORT_BLOCK : A.k_stage;
    for (int64_t k = k0; k <= k1; k <<= 1) {
        for (int j = (int)((k >> 1) < SEL_SORT_BLOCK / 2 ? (k >> 1) : SEL_SORT_BLOCK / 2); j >= 1; j >>= 1) {
            const int lo = ((t / j) * 2 * j) + (t % j);  // pair (lo, lo + j)
            const int64_t gi = base + lo;
            const bool up = (gi & k) == 0;
            SelRec x = s[lo], y = s[lo + j];
            if (sel_rec_less(y, x) == up) {
                s[lo] = y;
                s[lo + j] = x;
            }
            __syncthreads();
        }
    }
    A.recs[base + t] = s[t];
    A.recs[base + t + SEL_BLOCK] = s[t + SEL_BLOCK];
}

// One global compare-exchange step of stage k at distance j (j >= SEL_SORT_BLOCK).
__device__ void sel_bitonic_global_body(const WeftSelectArgs& A) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.n_pad / 2) return;
    const int64_t j = A.j_step, k = A.k_stage;
    const int64_t lo = (t / j) * 2 * j + (t % j);
    const bool up = (lo & k) == 0;
    const SelRec x = A.recs[lo], y = A.recs[lo + j];
    if (sel_rec_less(y, x) == up) {
        A.recs[lo] = y;
        A.recs[lo + j] = x;
    }
}

// After the first sort: keep sorted position p if fewer than k earlier positions share its
// segment (the record k places back belongs to another segment), as a 0/1 flag.
__device__ void sel_topk_keep_body(const WeftSelectArgs& A) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n) return;
    const bool keep = p < A.k || A.recs[p - A.k].a != A.recs[p].a;
    A.flags[p] = keep ? 1 : 0;
}

// Gather the *count_in kept records (sorted positions from the compaction) as
// second-sort records keyed by (score key, index); pad to n_pad.
__device__ void sel_topk_gather_body(const WeftSelectArgs& A) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n_pad) return;
    SelRec r;
    r._pad = 0;
    if (p < *A.count_in) {
        const SelRec& s = A.recs[A.positions[p]];
        r.a = s.b;
        r.b = 0;
        r.idx = s.idx;
    } else {
        r.a = ~0ull;
        r.b = ~0ull;
        r.idx = ~0u;
    }
    A.recs2[p] = r;
}

__device__ void sel_topk_emit_body(const WeftSelectArgs& A) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *A.count_in) return;
    const uint32_t i = A.recs[p].idx;
    A.out_i64[p] = (int64_t)(A.largest ? ~i : i);
}
