// Chunked static order-0 rANS for PRIM payloads (format of src/weft/entropy.py):
//   raw:  [0][u32 n][n bytes]
//   rANS: [1][u32 n][u16 n_pairs][(u8 sym, u16 freq) x n_pairs][u32 state][u32 n_emit][emitted]
// 12-bit frequencies summing to 4096, state in [2^23, 2^31), bytes emitted low first while
// encoding backwards; the decoder reads them from the end.  A chunk is the PRIM bytes
// of consecutive tiles (src/weft/prim_chunks.py), coded independently.
// Included by weft_kernels.hip inside its anonymous namespace.

constexpr int RANS_PROB_BITS = 12;
constexpr int RANS_TOT = 1 << RANS_PROB_BITS;
constexpr uint32_t RANS_LOW = 1u << 23;
constexpr int RANS_SEG = 8192;      // input bytes staged per encoder pass
constexpr int RANS_WIN = 8192;      // emitted bytes staged per decoder window

__device__ __forceinline__ uint32_t rans_rd32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
__device__ __forceinline__ void rans_wr32(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}

// ---- encoder: one 256-thread workgroup per chunk ---------------------------------
__device__ void rans_encode_body(const WeftRansEncodeArgs& A) {
    __shared__ int cnt[256];
    __shared__ int fr[256];
    __shared__ int ord[256];
    __shared__ int cum[257];
    __shared__ uint8_t seg[RANS_SEG];
    __shared__ int s_npairs, s_ne, s_err, s_rans;
    __shared__ uint32_t s_state;
    const int cid = blockIdx.x;
    const int tid = threadIdx.x;
    const WeftChunkDesc d = A.descs[cid];
    const int n = d.raw_length;
    const uint8_t* src = A.raw + d.raw_offset;
    uint8_t* dst = A.stage + A.stage_off[cid];
    const int raw_len = 5 + n;

    cnt[tid] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += 256) atomicAdd(&cnt[src[i]], 1);
    __syncthreads();
    // _build_model: round((c / total) * 4096) half to even, at least 1 (entropy.py:41-45)
    const int c = cnt[tid];
    int f = 0;
    if (c > 0) {
        const double r = rint(((double)c / (double)n) * (double)RANS_TOT);
        f = r < 1.0 ? 1 : (int)r;
    }
    fr[tid] = f;
    // position of this symbol in sorted(range(256), key=counts, reverse=True) (stable)
    int rank = 0;
    for (int s = 0; s < 256; ++s) {
        const int cs = cnt[s];
        rank += (cs > c) || (cs == c && s < tid);
    }
    ord[rank] = tid;
    __syncthreads();
    if (tid == 0) {
        int err = 0;
        if (n > 0) {
            int sum = 0;
            for (int s = 0; s < 256; ++s) sum += fr[s];
            if (sum == 0) {
                fr[0] = RANS_TOT;
                sum = RANS_TOT;
            }
            int diff = RANS_TOT - sum;
            for (int k = 0; diff != 0 && k < 100000; ++k) {
                int idx = ord[k & 255];
                if (cnt[idx] == 0) idx = 0;
                if (diff > 0) {
                    fr[idx] += 1;
                    diff -= 1;
                } else if (fr[idx] > 1) {
                    fr[idx] -= 1;
                    diff += 1;
                }
            }
            sum = 0;
            for (int s = 0; s < 256; ++s) sum += fr[s];
            if (sum != RANS_TOT) fr[0] += RANS_TOT - sum;
        }
        int run = 0, np_ = 0;
        for (int s = 0; s < 256; ++s) {
            cum[s] = run;
            run += fr[s];
            np_ += fr[s] > 0;
        }
        cum[256] = run;
        s_npairs = np_;
        s_ne = 0;
        s_err = err;
        s_state = RANS_LOW;
        s_rans = n > 0;
    }
    __syncthreads();
    const int hdr = 7 + 3 * s_npairs + 8;
    // encode backwards, RANS_SEG input bytes staged at a time; give up as soon as the
    // payload cannot be shorter than the raw form (the raw form is then chosen, as
    // encode_bytes would choose it)
    for (int hi = n; hi > 0 && s_rans; hi -= RANS_SEG) {
        const int lo = hi > RANS_SEG ? hi - RANS_SEG : 0;
        for (int i = lo + tid; i < hi; i += 256) seg[i - lo] = src[i];
        __syncthreads();
        if (tid == 0) {
            uint32_t state = s_state;
            int ne = s_ne;
            bool go = true;
            for (int i = hi - 1; i >= lo && go; --i) {
                const int sym = seg[i - lo];
                const int freq = fr[sym];
                if (freq <= 0) {
                    s_err = 12;  // symbol frequency is zero (entropy.py:99-100)
                    go = false;
                    break;
                }
                const uint32_t x_max = ((RANS_LOW >> RANS_PROB_BITS) << 8) * (uint32_t)freq;
                while (state >= x_max) {
                    if (hdr + ne + 1 >= raw_len) {
                        go = false;
                        break;
                    }
                    dst[hdr + ne] = (uint8_t)(state & 0xFF);
                    ++ne;
                    state >>= 8;
                }
                if (!go) break;
                state = ((state / (uint32_t)freq) << RANS_PROB_BITS) + (state % (uint32_t)freq) + (uint32_t)cum[sym];
            }
            s_state = state;
            s_ne = ne;
            if (!go) s_rans = 0;
        }
        __syncthreads();
    }
    const bool use_rans = s_rans && hdr + s_ne < raw_len && s_err == 0;
    if (use_rans) {
        if (tid == 0) {
            dst[0] = 1;
            rans_wr32(dst + 1, (uint32_t)n);
            dst[5] = (uint8_t)s_npairs;
            dst[6] = (uint8_t)(s_npairs >> 8);
            int p = 7;
            for (int s = 0; s < 256; ++s) {
                if (fr[s] > 0) {
                    dst[p] = (uint8_t)s;
                    dst[p + 1] = (uint8_t)fr[s];
                    dst[p + 2] = (uint8_t)(fr[s] >> 8);
                    p += 3;
                }
            }
            rans_wr32(dst + p, s_state);
            rans_wr32(dst + p + 4, (uint32_t)s_ne);
            A.descs[cid].coded_length = hdr + s_ne;
        }
    } else {
        if (tid == 0) {
            dst[0] = 0;
            rans_wr32(dst + 1, (uint32_t)n);
            A.descs[cid].coded_length = raw_len;
        }
        for (int i = tid; i < n; i += 256) dst[5 + i] = src[i];
    }
    if (tid == 0) A.status[cid] = s_err;
}

// Exclusive scan of the chunks' coded lengths into their offsets (one workgroup).
__device__ void rans_scan_body(const WeftRansEncodeArgs& A) {
    __shared__ int64_t part[1024];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = A.n_chunks;
    const int per = (n + nt - 1) / nt;
    const int a = tid * per, b = min(n, a + per);
    int64_t s = 0;
    for (int i = a; i < b; ++i) s += A.descs[i].coded_length;
    part[tid] = s;
    __syncthreads();
    if (tid == 0) {
        int64_t run = 0;
        for (int i = 0; i < nt; ++i) {
            const int64_t v = part[i];
            part[i] = run;
            run += v;
        }
        A.coded_total[0] = run;
    }
    __syncthreads();
    int64_t run = part[tid];
    for (int i = a; i < b; ++i) {
        A.descs[i].coded_offset = (int32_t)run;
        run += A.descs[i].coded_length;
    }
}

// Copy each chunk's payload from its staging slot to its place in the packed stream.
__device__ void rans_pack_body(const WeftRansEncodeArgs& A) {
    const int cid = blockIdx.x;
    const WeftChunkDesc d = A.descs[cid];
    const uint8_t* s = A.stage + A.stage_off[cid];
    uint8_t* o = A.coded + d.coded_offset;
    for (int i = threadIdx.x; i < d.coded_length; i += blockDim.x) o[i] = s[i];
}

// ---- decoder: one 64-lane workgroup per chunk ------------------------------------
// Status codes are those of the reference decoder (gpu_entropy.py:96-211).
__device__ void rans_decode_body(const WeftRansDecodeArgs& A) {
    __shared__ uint32_t tab[RANS_TOT];   // sym | (freq - 1) << 8 | (slot - cum) << 20
    __shared__ int fr[256];
    __shared__ int cum[257];
    __shared__ uint8_t win[RANS_WIN];
    __shared__ int s_status, s_off;
    const int cid = blockIdx.x;
    const int lane = threadIdx.x;
    const WeftChunkDesc d = A.descs[cid];
    const uint8_t* ptr = A.coded + d.coded_offset;
    uint8_t* out = A.out_raw + d.raw_offset;
    for (int s = lane; s < 256; s += 64) fr[s] = 0;
    __syncthreads();
    if (lane == 0) {
        int st = 0;
        if (d.coded_length <= 0 || d.raw_length < 0) {
            st = 1;
        } else if (ptr[0] == 0) {
            if (d.coded_length < 5) st = 2;
            else if ((int)rans_rd32(ptr + 1) != d.raw_length || d.raw_length + 5 != d.coded_length) st = 3;
            else st = -1;  // raw copy
        } else if (ptr[0] != 1) {
            st = 4;
        } else if (d.coded_length < 7) {
            st = 5;
        } else if ((int)rans_rd32(ptr + 1) != d.raw_length) {
            st = 6;
        } else {
            const int n_syms = ptr[5] | (ptr[6] << 8);
            int off = 7;
            if (off + n_syms * 3 + 8 > d.coded_length) {
                st = 7;
            } else {
                for (int i = 0; i < n_syms && st == 0; ++i) {
                    const int sym = ptr[off];
                    const int freq = ptr[off + 1] | (ptr[off + 2] << 8);
                    off += 3;
                    if (freq <= 0) st = 8;
                    else fr[sym] = freq;
                }
                if (st == 0) {
                    int run = 0;
                    for (int s = 0; s < 256; ++s) {
                        cum[s] = run;
                        run += fr[s];
                    }
                    cum[256] = run;
                    if (run != RANS_TOT) st = 9;
                    else if (off + 8 + (int)rans_rd32(ptr + off + 4) != d.coded_length) st = 10;
                }
            }
            s_off = off;
        }
        s_status = st;
    }
    __syncthreads();
    const int st = s_status;
    if (st == -1) {
        for (int i = lane; i < d.raw_length; i += 64) out[i] = ptr[5 + i];
        if (lane == 0) A.status[cid] = 0;
        return;
    }
    if (st != 0) {
        if (lane == 0) A.status[cid] = st;
        return;
    }
    // slot table: binary search of the slot's symbol in cum (symbols of frequency 0 own no slot)
    for (int q = lane; q < RANS_TOT; q += 64) {
        int lo = 0, hi = 256;  // cum[lo] <= q < cum[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (cum[mid] <= q) lo = mid;
            else hi = mid;
        }
        tab[q] = (uint32_t)lo | ((uint32_t)(fr[lo] - 1) << 8) | ((uint32_t)(q - cum[lo]) << 20);
    }
    const int off = s_off;
    uint32_t state = rans_rd32(ptr + off);
    const int ne = (int)rans_rd32(ptr + off + 4);
    const uint8_t* em = ptr + off + 8;
    int eptr = ne - 1;
    int win_lo = ne > RANS_WIN ? ne - RANS_WIN : 0;
    for (int i = win_lo + lane; i < ne; i += 64) win[i - win_lo] = em[i];
    __syncthreads();
    // every lane runs the same serial decode; lane j keeps output byte j of each 64
    uint8_t mine = 0;
    const int n = d.raw_length;
    for (int i = 0; i < n; ++i) {
        const uint32_t e = tab[state & (RANS_TOT - 1)];
        const uint32_t sym = e & 0xFF;
        if ((i & 63) == lane) mine = (uint8_t)sym;
        if ((i & 63) == 63 || i == n - 1) {
            if (lane <= (i & 63)) out[(i & ~63) + lane] = mine;
        }
        state = (((e >> 8) & 0xFFF) + 1) * (state >> RANS_PROB_BITS) + (e >> 20);
        while (state < RANS_LOW && eptr >= 0) {
            if (eptr < win_lo) {  // uniform across the wave: refill the window
                __syncthreads();
                const int nlo = eptr + 1 > RANS_WIN ? eptr + 1 - RANS_WIN : 0;
                for (int j = nlo + lane; j <= eptr; j += 64) win[j - nlo] = em[j];
                win_lo = nlo;
                __syncthreads();
            }
            state = (state << 8) | win[eptr - win_lo];
            --eptr;
        }
    }
    if (lane == 0) A.status[cid] = 0;
}
