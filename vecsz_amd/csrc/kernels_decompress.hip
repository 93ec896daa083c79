// gfx950 kernels of the decompress path.
//
//   Huffman decode (decode_withTree/decode, src/huffman.hh:343-372, :679-691)
//   as a chunk-parallel, self-synchronising decode of the archive's single
//   MSB-first bitstream:
//     k_decode_sync   each thread decodes kSubBits bits from an assumed
//                     codeword start and records its first codeword starts
//     k_decode_check  a subsequence is synchronised when the previous one's end
//                     is among its recorded starts (then every later codeword
//                     agrees); counts are corrected, the rest marked dirty
//     k_decode_redo   dirty subsequences re-decode from the previous end until
//                     no end moves
//     k_decode_write  final decode: quant codes + outlier scatter
//                     (src/vecsz_interface.hh:313-317)
//   Lorenzo reconstruction (x_lorenzo_{1,2,3}d1l, src/dualquant.hh:915-1043):
//     k_recon_*       integer segmented scans per Lorenzo block; exact, and
//                     therefore bit-identical to the reference's sequential
//                     fp32 recurrence, whenever every |s| < 2^21 and every
//                     outlier is an integer below 2^21 (all fp32 partial sums
//                     then stay below 2^24); other blocks are listed
//     k_recon_fallback one thread per listed block replays the reference's
//                     fp32 recurrence in its own association
#include "common.hpp"
#include "kernels.hpp"

namespace vsz {

__device__ __forceinline__ uint32_t lane_id2() { return __lane_id(); }

struct BitReader {
    const uint32_t* base4;
    uint64_t shift0;
    uint64_t last_word;
    uint64_t buf;
    int nb;
    uint64_t next;

    __device__ __forceinline__ void setup(const uint8_t* stream, uint64_t stream_bits) {
        uintptr_t p = reinterpret_cast<uintptr_t>(stream);
        base4 = reinterpret_cast<const uint32_t*>(p & ~(uintptr_t)3);
        shift0 = 8 * (p & 3);
        last_word = stream_bits ? (shift0 + stream_bits - 1) >> 5 : 0;
    }
    __device__ __forceinline__ uint32_t load(uint64_t w) const {
        return w <= last_word ? __builtin_bswap32(base4[w]) : 0u;
    }
    __device__ __forceinline__ void refill() {
        if (nb <= 32) {
            buf |= (uint64_t)load(next++) << (32 - nb);
            nb += 32;
        }
    }
    __device__ __forceinline__ void init(uint64_t pos) {
        const uint64_t a = pos + shift0;
        const uint64_t w = a >> 5;
        const int o = (int)(a & 31);
        buf = ((uint64_t)load(w) << 32) | load(w + 1);
        buf <<= o;
        nb = 64 - o;
        next = w + 2;
        refill();
    }
    __device__ __forceinline__ void consume(int l) {
        buf <<= l;
        nb -= l;
        refill();
    }
};

// One codeword: LDS table for the first kLutBits bits, bit-serial walk of the
// node table beyond.  The tree was checked to be full on the host, so every
// walk ends at a leaf; `maxwalk` bounds it anyway.
__device__ __forceinline__ uint32_t decode_one(BitReader& r, const uint32_t* __restrict__ lut,
                                               const uint32_t* __restrict__ child, uint32_t& len, bool& bad) {
    const uint32_t e = lut[r.buf >> (64 - kLutBits)];
    if (!(e & kLutSlow)) {
        len = (e >> 16) & 0xFF;
        if (len == 0) {
            bad = true;
            len = 1;
            r.consume(1);
            return 0;
        }
        r.consume((int)len);
        return e & 0xFFFF;
    }
    uint32_t node = e & 0x7FFFFFFFu;
    r.consume(kLutBits);
    len = kLutBits;
    for (int guard = 0; guard < 65536; guard++) {
        const uint32_t bit = (uint32_t)(r.buf >> 63);
        r.consume(1);
        len++;
        const uint32_t nx = child[2 * node + bit];
        if (nx == kNodeAbsent) break;
        if (nx & kNodeLeaf) return nx & 0xFFFF;
        node = nx;
    }
    bad = true;
    return 0;
}

__device__ __forceinline__ void load_lut(uint32_t* slut, const uint32_t* lut) {
    for (int i = threadIdx.x; i < (1 << kLutBits); i += blockDim.x) slut[i] = lut[i];
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_decode_sync(DecodeArgs a, SubState s, DecSummary* sum) {
    __shared__ uint32_t slut[1 << kLutBits];
    __shared__ uint16_t sb[256 * kSyncRecord];
    load_lut(slut, a.lut);
    const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= a.nsub) return;
    BitReader r;
    r.setup(a.stream, a.stream_bits);
    const uint64_t start = j * kSubBits;
    const uint64_t target = min(start + (uint64_t)kSubBits, a.stream_bits);
    r.init(start);
    uint64_t pos = start;
    uint32_t cnt = 0, z = 0, zm = 0;
    bool bad = false;
    uint16_t* mine = sb + threadIdx.x * kSyncRecord;
    while (pos < target) {
        if (cnt < kSyncRecord) mine[cnt] = (uint16_t)(pos - start);
        uint32_t len;
        const uint32_t sym = decode_one(r, slut, a.child, len, bad);
        if (sym == 0) {
            z++;
            if (cnt < kSyncRecord) zm |= 1u << cnt;
        }
        pos += len;
        cnt++;
    }
    if (bad) atomicOr(&sum->error, 1u);
    s.start[j] = start;
    s.end[j] = pos;
    s.cnt[j] = cnt;
    s.zeros[j] = z;
    s.zmask[j] = (uint16_t)zm;
    const uint32_t nb = min(cnt, (uint32_t)kSyncRecord);
    s.nbnd[j] = (uint8_t)nb;
    for (uint32_t i = 0; i < nb; i++) s.bnd[j * kSyncRecord + i] = mine[i];
}

__global__ void __launch_bounds__(256) k_decode_check(DecodeArgs a, SubState s, DecSummary* sum) {
    const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= a.nsub) return;
    s.dirty[j] = 0;
    s.dirty[a.nsub + j] = 0;
    if (j == 0) return;
    const uint64_t e = s.end[j - 1];
    const uint64_t st = s.start[j];
    if (e == st) return;
    const uint64_t off = e - st;
    const uint32_t nb = s.nbnd[j];
    for (uint32_t i = 1; i < nb; i++) {
        if ((uint64_t)s.bnd[j * kSyncRecord + i] == off) {
            s.start[j] = e;
            s.cnt[j] -= i;
            s.zeros[j] -= __popc((uint32_t)s.zmask[j] & ((1u << i) - 1u));
            return;
        }
    }
    s.dirty[j] = 1;
    atomicAdd(&sum->n_changed, 1u);
}

// One round of re-decoding: dirty_cur -> dirty_nxt.
__global__ void __launch_bounds__(256) k_decode_redo(DecodeArgs a, SubState s, const uint64_t* end_in, uint64_t* end_out,
                                                     uint8_t* dirty_cur, uint8_t* dirty_nxt, DecSummary* sum) {
    __shared__ uint32_t slut[1 << kLutBits];
    load_lut(slut, a.lut);
    const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= a.nsub) return;
    if (!dirty_cur[j]) {
        end_out[j] = end_in[j];
        return;
    }
    dirty_cur[j] = 0;
    BitReader r;
    r.setup(a.stream, a.stream_bits);
    const uint64_t st = end_in[j - 1];
    const uint64_t target = min((j + 1) * (uint64_t)kSubBits, a.stream_bits);
    r.init(st);
    uint64_t pos = st;
    uint32_t cnt = 0, z = 0;
    bool bad = false;
    while (pos < target) {
        uint32_t len;
        const uint32_t sym = decode_one(r, slut, a.child, len, bad);
        z += sym == 0;
        pos += len;
        cnt++;
    }
    if (bad) atomicOr(&sum->error, 1u);
    s.start[j] = st;
    s.cnt[j] = cnt;
    s.zeros[j] = z;
    end_out[j] = pos;
    if (pos != end_in[j] && j + 1 < a.nsub) {
        dirty_nxt[j + 1] = 1;
        atomicAdd(&sum->n_changed, 1u);
    }
}

// Device-wide exclusive scan of two u32 arrays into u64 offsets.
constexpr int kScanPer = 4;
constexpr int kScanBlock = 1024 * kScanPer;

__global__ void __launch_bounds__(1024) k_scan_local(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                                     uint64_t n, uint64_t* __restrict__ oa, uint64_t* __restrict__ ob,
                                                     uint64_t* __restrict__ bsum) {
    __shared__ uint64_t sa[16], sbb[16];
    const uint64_t base = blockIdx.x * (uint64_t)kScanBlock + threadIdx.x * kScanPer;
    uint64_t va[kScanPer], vb[kScanPer], ta = 0, tb = 0;
#pragma unroll
    for (int i = 0; i < kScanPer; i++) {
        va[i] = base + i < n ? a[base + i] : 0;
        vb[i] = base + i < n ? b[base + i] : 0;
        ta += va[i];
        tb += vb[i];
    }
    uint64_t ia = ta, ib = tb;
    const uint32_t lane = lane_id2();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t xa = __shfl_up(ia, d), xb = __shfl_up(ib, d);
        if (lane >= (uint32_t)d) {
            ia += xa;
            ib += xb;
        }
    }
    const int w = threadIdx.x >> 6;
    if (lane == 63) {
        sa[w] = ia;
        sbb[w] = ib;
    }
    __syncthreads();
    uint64_t pa = 0, pb = 0, tot_a = 0, tot_b = 0;
    for (int i = 0; i < 16; i++) {
        if (i < w) {
            pa += sa[i];
            pb += sbb[i];
        }
        tot_a += sa[i];
        tot_b += sbb[i];
    }
    pa += ia - ta;
    pb += ib - tb;
#pragma unroll
    for (int i = 0; i < kScanPer; i++) {
        if (base + i < n) {
            oa[base + i] = pa;
            ob[base + i] = pb;
        }
        pa += va[i];
        pb += vb[i];
    }
    if (threadIdx.x == 0) {
        bsum[2 * blockIdx.x] = tot_a;
        bsum[2 * blockIdx.x + 1] = tot_b;
    }
}

__global__ void __launch_bounds__(1024) k_scan_blocks(uint64_t* __restrict__ bsum, uint64_t nb, uint64_t* __restrict__ totals) {
    if (threadIdx.x == 0) {
        uint64_t ra = 0, rb = 0;
        for (uint64_t i = 0; i < nb; i++) {
            uint64_t x = bsum[2 * i], y = bsum[2 * i + 1];
            bsum[2 * i] = ra;
            bsum[2 * i + 1] = rb;
            ra += x;
            rb += y;
        }
        totals[0] = ra;
        totals[1] = rb;
    }
}

__global__ void __launch_bounds__(1024) k_scan_add(uint64_t* __restrict__ oa, uint64_t* __restrict__ ob, uint64_t n,
                                                   const uint64_t* __restrict__ bsum) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) {
        const uint64_t blk = i / kScanBlock;
        oa[i] += bsum[2 * blk];
        ob[i] += bsum[2 * blk + 1];
    }
}

__device__ __forceinline__ float read_f32_bytes(const uint8_t* p) {
    uint32_t v = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
    return __uint_as_float(v);
}

__global__ void __launch_bounds__(256) k_decode_write(DecodeArgs a, SubState s, const uint64_t* __restrict__ sym_off,
                                                      const uint64_t* __restrict__ zero_off, uint16_t* __restrict__ codes,
                                                      float* __restrict__ sparse, DecSummary* sum) {
    __shared__ uint32_t slut[1 << kLutBits];
    load_lut(slut, a.lut);
    const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= a.nsub) return;
    BitReader r;
    r.setup(a.stream, a.stream_bits);
    r.init(s.start[j]);
    const uint32_t c = s.cnt[j];
    uint64_t idx = sym_off[j];
    uint64_t rank = zero_off[j];
    bool bad = false;
    for (uint32_t i = 0; i < c && idx < a.n; i++, idx++) {
        uint32_t len;
        const uint32_t sym = decode_one(r, slut, a.child, len, bad);
        codes[idx] = (uint16_t)sym;
        if (sym == 0) {
            if (rank < a.n_outlier)
                sparse[idx] = read_f32_bytes(a.outliers + 4 * rank);
            else
                bad = true;
            rank++;
        }
    }
    if (bad) atomicOr(&sum->error, 1u);
}

__global__ void k_fill_u16(uint16_t* p, uint64_t n, uint16_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_scatter_all(const uint8_t* __restrict__ outl, uint64_t n, float* __restrict__ sparse) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        sparse[i] = read_f32_bytes(outl + 4 * i);
}

// ------------------------------------------------------------------ reconstruction

constexpr int32_t kExact = 1 << 21;

__device__ __forceinline__ float scale_out(int32_t s, double twoeb) { return (float)((double)(float)s * twoeb); }

// An outlier's s is its stored value itself, sign of zero included (a
// non-outlier s is never -0.0: pred + (float)(code - R) with a +0.0 addend).
__device__ __forceinline__ float scale_outlier(float v, double twoeb) { return (float)((double)v * twoeb); }

// Inclusive segmented scan over segments of B lanes: a lane with f set starts
// a new run.
template <int B>
__device__ __forceinline__ int32_t seg_scan(int32_t v, bool f) {
    const uint32_t lib = lane_id2() % B;
#pragma unroll
    for (int d = 1; d < B; d <<= 1) {
        int32_t nv = __shfl_up(v, d, B);
        int nf = __shfl_up((int)f, d, B);
        if (lib >= (uint32_t)d) {
            if (!f) v += nv;
            f = f || nf;
        }
    }
    return v;
}

__device__ __forceinline__ bool exact_outlier(float v) { return v == __builtin_rintf(v) && __builtin_fabsf(v) < (float)kExact; }

__device__ __forceinline__ void flag_blocks(bool bad, int B, uint64_t blk_x0, uint64_t by, uint64_t bz, uint64_t nbx,
                                            uint64_t nby, uint32_t* flagged, uint32_t* nflagged) {
    const uint64_t m = __ballot(bad);
    if (!m) return;
    const uint32_t lane = lane_id2();
    if (lane % B == 0) {
        const uint64_t seg = (B == 64) ? ~0ull : (((1ull << B) - 1) << lane);
        if (m & seg) {
            const uint64_t bx = blk_x0 + lane / B;
            const uint64_t id = bx + nbx * (by + nby * bz);
            const uint32_t slot = atomicAdd(nflagged, 1u);
            flagged[slot] = (uint32_t)id;
        }
    }
}

template <int B>
__global__ void __launch_bounds__(256) k_recon_3d(const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, ReconArgs a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    const uint32_t lane = lane_id2();
    const uint64_t ntiles = a.ntiles_x * a.ntiles_y * a.ntiles_z;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    const uint64_t sy = a.d0, sz = a.d0 * a.d1;
    const uint64_t nbx = (a.d0 + B - 1) / B, nby = (a.d1 + B - 1) / B;
    for (uint64_t t = wave; t < ntiles; t += nwaves) {
        const uint64_t tx = t % a.ntiles_x;
        const uint64_t ty = (t / a.ntiles_x) % a.ntiles_y;
        const uint64_t tz = t / (a.ntiles_x * a.ntiles_y);
        const uint64_t x = tx * 64 + lane;
        const uint64_t y0 = ty * B, z0 = tz * B;
        const bool xin = x < a.d0;
        const bool bstart = (lane % B) == 0;
        bool bad = false;
        int32_t prevplane[B];
#pragma unroll
        for (int j = 0; j < B; j++) prevplane[j] = 0;
        for (int zz = 0; zz < B; zz++) {
            const uint64_t z = z0 + zz;
            if (z >= a.d2) break;
            int32_t srow = 0, zrow = 0;
#pragma unroll
            for (int yy = 0; yy < B; yy++) {
                const uint64_t y = y0 + yy;
                if (y < a.d1) {
                    const uint64_t idx = x + y * sy + z * sz;
                    const uint32_t c = xin ? codes[idx] : (uint32_t)a.radius;
                    const bool isout = xin && c == 0;
                    const int32_t Y = yy > 0 ? srow : 0;
                    const int32_t Z = zz > 0 ? prevplane[yy] : 0;
                    const int32_t YZ = (yy > 0 && zz > 0) ? zrow : 0;
                    int32_t g0;
                    float v = 0.f;
                    if (isout) {
                        v = sparse[idx];
                        if (!exact_outlier(v)) bad = true;
                        g0 = (int32_t)v - Y - Z + YZ;
                    } else {
                        g0 = (int32_t)c - a.radius;
                    }
                    const int32_t G = seg_scan<B>(g0, isout || bstart);
                    const int32_t sv = G + Y + Z - YZ;
                    if (sv >= kExact || sv <= -kExact) bad = true;
                    if (xin) out[idx] = isout ? scale_outlier(v, a.twoeb) : scale_out(sv, a.twoeb);
                    zrow = Z;
                    prevplane[yy] = sv;
                    srow = sv;
                }
            }
        }
        flag_blocks(bad, B, tx * (64 / B), ty, tz, nbx, nby, flagged, nflagged);
    }
}

template <int B>
__global__ void __launch_bounds__(256) k_recon_2d(const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, ReconArgs a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    const uint32_t lane = lane_id2();
    const uint64_t ntiles = a.ntiles_x * a.ntiles_y;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    const uint64_t nbx = (a.d0 + B - 1) / B, nby = (a.d1 + B - 1) / B;
    for (uint64_t t = wave; t < ntiles; t += nwaves) {
        const uint64_t tx = t % a.ntiles_x;
        const uint64_t ty = t / a.ntiles_x;
        const uint64_t x = tx * 64 + lane;
        const uint64_t y0 = ty * B;
        const bool xin = x < a.d0;
        const bool bstart = (lane % B) == 0;
        bool bad = false;
        int32_t srow = 0;
#pragma unroll
        for (int yy = 0; yy < B; yy++) {
            const uint64_t y = y0 + yy;
            if (y < a.d1) {
                const uint64_t idx = x + y * a.d0;
                const uint32_t c = xin ? codes[idx] : (uint32_t)a.radius;
                const bool isout = xin && c == 0;
                const int32_t N = yy > 0 ? srow : 0;
                int32_t g0;
                float v = 0.f;
                if (isout) {
                    v = sparse[idx];
                    if (!exact_outlier(v)) bad = true;
                    g0 = (int32_t)v - N;
                } else {
                    g0 = (int32_t)c - a.radius;
                }
                const int32_t G = seg_scan<B>(g0, isout || bstart);
                const int32_t sv = G + N;
                if (sv >= kExact || sv <= -kExact) bad = true;
                if (xin) out[idx] = isout ? scale_outlier(v, a.twoeb) : scale_out(sv, a.twoeb);
                srow = sv;
            }
        }
        flag_blocks(bad, B, tx * (64 / B), ty, 0, nbx, nby, flagged, nflagged);
    }
}

template <int B>
__global__ void __launch_bounds__(256) k_recon_1d(const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, ReconArgs a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    const uint32_t lane = lane_id2();
    const uint64_t ntiles = a.ntiles_x;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    const uint64_t nbx = (a.d0 + B - 1) / B;
    for (uint64_t t = wave; t < ntiles; t += nwaves) {
        const uint64_t x = t * 64 + lane;
        const bool xin = x < a.d0;
        const bool bstart = (lane % B) == 0;
        bool bad = false;
        const uint32_t c = xin ? codes[x] : (uint32_t)a.radius;
        const bool isout = xin && c == 0;
        int32_t g0;
        float v = 0.f;
        if (isout) {
            v = sparse[x];
            if (!exact_outlier(v)) bad = true;
            g0 = (int32_t)v;
        } else {
            g0 = (int32_t)c - a.radius;
        }
        const int32_t sv = seg_scan<B>(g0, isout || bstart);
        if (sv >= kExact || sv <= -kExact) bad = true;
        if (xin) out[x] = isout ? scale_outlier(v, a.twoeb) : scale_out(sv, a.twoeb);
        flag_blocks(bad, B, t * (64 / B), 0, 0, nbx, 1, flagged, nflagged);
    }
}

// One thread per listed block: the reference's sequential fp32 recurrence in
// its own association, using the block's region of `out` for the s values
// before scaling them in place.
__global__ void __launch_bounds__(64) k_recon_fallback(int ndim, int B, const uint16_t* __restrict__ codes,
                                                       const float* __restrict__ sparse, float* __restrict__ out,
                                                       ReconArgs a, const uint32_t* __restrict__ flagged,
                                                       const uint32_t* __restrict__ nflagged) {
    const uint32_t nf = *nflagged;
    const uint64_t nbx = (a.d0 + B - 1) / B, nby = (a.d1 + B - 1) / B;
    const uint64_t sy = a.d0, sz = a.d0 * a.d1;
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t id = flagged[f];
        const uint64_t bx = id % nbx, by = (id / nbx) % nby, bz = id / (nbx * nby);
        const uint64_t x0 = bx * B, y0 = by * B, z0 = bz * B;
        const uint64_t xe = min(x0 + B, a.d0), ye = min(y0 + B, a.d1), ze = min(z0 + B, a.d2);
        for (uint64_t z = z0; z < ze; z++)
            for (uint64_t y = y0; y < ye; y++)
                for (uint64_t x = x0; x < xe; x++) {
                    const uint64_t idx = x + y * sy + z * sz;
                    const bool hx = x > x0, hy = y > y0, hz = z > z0;
                    float pred;
                    if (ndim == 1) {
                        pred = hx ? out[idx - 1] : 0.f;
                    } else if (ndim == 2) {
                        const float W = hx ? out[idx - 1] : 0.f;
                        const float N = hy ? out[idx - sy] : 0.f;
                        const float NW = (hx && hy) ? out[idx - 1 - sy] : 0.f;
                        pred = N + W;
                        pred = pred - NW;
                    } else {
                        const float X = hx ? out[idx - 1] : 0.f;
                        const float Y = hy ? out[idx - sy] : 0.f;
                        const float Z = hz ? out[idx - sz] : 0.f;
                        const float XY = (hx && hy) ? out[idx - 1 - sy] : 0.f;
                        const float XZ = (hx && hz) ? out[idx - 1 - sz] : 0.f;
                        const float YZ = (hy && hz) ? out[idx - sy - sz] : 0.f;
                        const float XYZ = (hx && hy && hz) ? out[idx - 1 - sy - sz] : 0.f;
                        pred = XYZ - XY;
                        pred = pred - XZ;
                        pred = pred - YZ;
                        pred = pred + X;
                        pred = pred + Y;
                        pred = pred + Z;
                    }
                    const uint32_t c = codes[idx];
                    out[idx] = c == 0 ? sparse[idx] : pred + (float)((int32_t)c - a.radius);
                }
        for (uint64_t z = z0; z < ze; z++)
            for (uint64_t y = y0; y < ye; y++)
                for (uint64_t x = x0; x < xe; x++) {
                    const uint64_t idx = x + y * sy + z * sz;
                    out[idx] = (float)((double)out[idx] * a.twoeb);
                }
    }
}

// ------------------------------------------------------------------ launchers

static unsigned grid_for(uint64_t threads, unsigned block) { return (unsigned)((threads + block - 1) / block); }

void launch_decode_sync(const DecodeArgs& a, SubState s, DecSummary* sum, hipStream_t st) {
    hipLaunchKernelGGL(k_decode_sync, dim3(grid_for(a.nsub, 256)), dim3(256), 0, st, a, s, sum);
}
void launch_decode_check(const DecodeArgs& a, SubState s, DecSummary* sum, hipStream_t st) {
    hipLaunchKernelGGL(k_decode_check, dim3(grid_for(a.nsub, 256)), dim3(256), 0, st, a, s, sum);
}
void launch_decode_redo(const DecodeArgs& a, SubState s, const uint64_t* end_in, uint64_t* end_out, uint8_t* dirty_cur,
                        uint8_t* dirty_nxt, DecSummary* sum, hipStream_t st) {
    hipLaunchKernelGGL(k_decode_redo, dim3(grid_for(a.nsub, 256)), dim3(256), 0, st, a, s, end_in, end_out, dirty_cur,
                       dirty_nxt, sum);
}
void launch_scan_u32x2(const uint32_t* a, const uint32_t* b, uint64_t n, uint64_t* oa, uint64_t* ob, uint64_t* block_tmp,
                       uint64_t* totals, hipStream_t st) {
    const uint64_t nb = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(1024), 0, st, a, b, n, oa, ob, block_tmp);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(64), 0, st, block_tmp, nb, totals);
    hipLaunchKernelGGL(k_scan_add, dim3(grid_for(n, 1024)), dim3(1024), 0, st, oa, ob, n, block_tmp);
}
void launch_decode_write(const DecodeArgs& a, SubState s, const uint64_t* sym_off, const uint64_t* zero_off,
                         uint16_t* codes, float* sparse, DecSummary* sum, hipStream_t st) {
    hipLaunchKernelGGL(k_decode_write, dim3(grid_for(a.nsub, 256)), dim3(256), 0, st, a, s, sym_off, zero_off, codes,
                       sparse, sum);
}
void launch_fill_u16(uint16_t* p, uint64_t n, uint16_t v, hipStream_t st) {
    unsigned g = (unsigned)std::min<uint64_t>(grid_for(n, 256), 8192);
    hipLaunchKernelGGL(k_fill_u16, dim3(std::max(g, 1u)), dim3(256), 0, st, p, n, v);
}
void launch_scatter_all_outliers(const uint8_t* outliers, uint64_t n, float* sparse, hipStream_t st) {
    unsigned g = (unsigned)std::min<uint64_t>(grid_for(n, 256), 8192);
    hipLaunchKernelGGL(k_scatter_all, dim3(std::max(g, 1u)), dim3(256), 0, st, outliers, n, sparse);
}

void launch_reconstruct(int ndim, int B, const uint16_t* codes, const float* sparse, float* out, const ReconArgs& a,
                        uint32_t* flagged, uint32_t* nflagged, int grid, hipStream_t st) {
#define RC_CASE(ND, KERN, BB)                                                                               \
    if (ndim == ND && B == BB) {                                                                            \
        hipLaunchKernelGGL(KERN<BB>, dim3(grid), dim3(256), 0, st, codes, sparse, out, a, flagged, nflagged); \
        return;                                                                                             \
    }
    RC_CASE(3, k_recon_3d, 8)
    RC_CASE(3, k_recon_3d, 16)
    RC_CASE(3, k_recon_3d, 32)
    RC_CASE(3, k_recon_3d, 64)
    RC_CASE(2, k_recon_2d, 8)
    RC_CASE(2, k_recon_2d, 16)
    RC_CASE(2, k_recon_2d, 32)
    RC_CASE(2, k_recon_2d, 64)
    RC_CASE(1, k_recon_1d, 8)
    RC_CASE(1, k_recon_1d, 16)
    RC_CASE(1, k_recon_1d, 32)
    RC_CASE(1, k_recon_1d, 64)
#undef RC_CASE
}

void launch_reconstruct_fallback(int ndim, int B, const uint16_t* codes, const float* sparse, float* out,
                                 const ReconArgs& a, const uint32_t* flagged, const uint32_t* nflagged, hipStream_t st) {
    hipLaunchKernelGGL(k_recon_fallback, dim3(256), dim3(64), 0, st, ndim, B, codes, sparse, out, a, flagged, nflagged);
}

}  // namespace vsz
