// gfx950 kernels of the compress path.
//
//   k_range        value range for r2r (GetDatumValueRange, src/types.cc:27-35)
//   k_dualquant_*  prequant + blocked Lorenzo postquant + outlier capture +
//                  quant-code histogram (c_lorenzo_{1,2,3}d1l, src/dualquant.hh:18-913;
//                  initHuffman's frequency loop, src/huffman.hh:239-242)
//   k_chunk_bits   bits and outliers per 1024-symbol chunk (HuffmanEncode sizes, src/huffman.hh:268-341)
//   k_enc_summary  archive sizes from the scanned chunk totals (src/vecsz_interface.hh:189-202)
//   k_pack         MSB-first bit packing into the archive + outlier compaction in index order
//                  (HuffmanEncode; outlier gather, src/vecsz_interface.hh:166-170)
//
// Arithmetic of the dual quantization is per element and mirrors the
// reference's per-element form exactly; the build uses -ffp-contract=off and
// IEEE denormals so no product is fused or flushed.
#include "common.hpp"
#include "kernels.hpp"

namespace vsz {

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Value held by lane-1 of the wavefront; lane 0 receives 0.0f.  DPP
// wave_shr:1 (GFX9 family) -- every lane of the wave must be active.
__device__ __forceinline__ float wave_shr1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, false));
}

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// AVX prequant: _mm256_round_ps(_mm256_mul_ps(x, r_f), NEAREST_INT).
__device__ __forceinline__ float prequant_vec(float x, float r_f) { return __builtin_rintf(x * r_f); }

// Scalar prequant: (float) round((double)x * r_d), round half away from zero.
__device__ __forceinline__ float prequant_scalar(float x, double r_d) {
    return (float)__builtin_round((double)x * r_d);
}

// ------------------------------------------------------------------ range

// Min/max ignoring NaN; the host applies the first-element rule of
// std::max_element/min_element (a NaN x[0] makes the range NaN).
// Per-workgroup min/max (ordered-integer form) into out[2 + 2 b], out[3 + 2 b];
// k_ebfam reduces them.  Partials rather than atomics on one address: a
// thousand workgroups contending for one L2 line cost more than the reads.
// (Plain loads: the dual quantization walks the field backwards next and
// finds its tail still in the caches.)
__global__ void __launch_bounds__(256) k_range(const float* __restrict__ x, uint64_t n, uint32_t* __restrict__ out) {
    float mn = __builtin_inff(), mx = -__builtin_inff();
    // float4 loads need 16-byte alignment; an unaligned field takes the scalar tail path
    const uint64_t n4 = ((reinterpret_cast<uintptr_t>(x) & 15) == 0) ? n / 4 : 0;
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    // four independent 16-byte loads in flight per thread
    for (; i + 3 * stride < n4; i += 4 * stride) {
        float4 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = x4[i + k * stride];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            mn = fminf(mn, fminf(fminf(v[k].x, v[k].y), fminf(v[k].z, v[k].w)));
            mx = fmaxf(mx, fmaxf(fmaxf(v[k].x, v[k].y), fmaxf(v[k].z, v[k].w)));
        }
    }
    for (; i < n4; i += stride) {
        float4 v = x4[i];
        mn = fminf(mn, fminf(fminf(v.x, v.y), fminf(v.z, v.w)));
        mx = fmaxf(mx, fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
    }
    for (uint64_t i = n4 * 4 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        float v = x[i];
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    for (int d = 32; d > 0; d >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, d));
        mx = fmaxf(mx, __shfl_xor(mx, d));
    }
    __shared__ float smn[4], smx[4];
    int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        smn[w] = mn;
        smx[w] = mx;
    }
    lds_sync();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
            mn = fminf(mn, smn[i]);
            mx = fmaxf(mx, smx[i]);
        }
        // fminf/fmaxf skip NaNs, so a block of NaNs leaves +inf / -inf,
        // which every other value overrides
        out[2 + 2 * blockIdx.x] = f2ord(mn);
        out[3 + 2 * blockIdx.x] = f2ord(mx);
    }
}

// r2r error-bound family from the device-side min/max.  std::max_element /
// std::min_element keep a leading NaN (then the range is NaN); otherwise
// they ignore NaNs, which is what the NaN-skipping reduction found.
__device__ __forceinline__ void ebfam(const uint32_t* __restrict__ mm, const float* __restrict__ x, double eb,
                                      EbDev* __restrict__ out) {
    const float x0 = x[0];
    float mn, mx;
    if (x0 != x0) {
        mn = mx = x0;
    } else {
        mn = ord2f(mm[0]);
        mx = ord2f(mm[1]);
    }
    const float r = mx - mn;
    EbDev e;
    e.range = (double)r;
    e.zero_range = e.range == 0;
    e.eb_final = eb * e.range;
    e.r_d = 1 / (2 * e.eb_final);
    e.twoeb = 2 * e.eb_final;
    e.r_f = (float)e.r_d;
    *out = e;
}

// Reduces k_range's nparts partials into mm[0], mm[1], then the eb family.
__global__ void __launch_bounds__(256) k_ebfam(uint32_t* __restrict__ mm, uint32_t nparts, const float* __restrict__ x,
                                               double eb, EbDev* __restrict__ out) {
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
        lo = min(lo, mm[2 + 2 * i]);
        hi = max(hi, mm[3 + 2 * i]);
    }
    for (int d = 32; d > 0; d >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, d));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, d));
    }
    __shared__ uint32_t slo[4], shi[4];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        slo[w] = lo;
        shi[w] = hi;
    }
    lds_sync();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
            lo = min(lo, slo[i]);
            hi = max(hi, shi[i]);
        }
        mm[0] = lo;
        mm[1] = hi;
        ebfam(mm, x, eb, out);
    }
}

// Replaces a's r_f / r_d by the device-computed family when there is one.
__device__ __forceinline__ void dq_resolve(DQArgs& a) {
    if (a.ebd) {
        a.r_f = a.ebd->r_f;
        a.r_d = a.ebd->r_d;
    }
}

// ------------------------------------------------------------------ histogram

// LDS histogram: a private copy of every bin plus 64 lane-private copies of
// the 64 bins around the radius, where a peaked code distribution would
// otherwise serialise a wavefront's atomics on a few addresses.
constexpr int kHotHalf = 32;
constexpr int kHotBins = 2 * kHotHalf;

struct LdsHist {
    uint32_t* cold;  // dict bins
    uint32_t* hot;   // kHotBins * 64
    int dict, lo;
    bool lds_cold;
    unsigned long long* global;

    __device__ __forceinline__ void add(int bin) {
        const int h = bin - lo;
        const bool is_hot = (unsigned)h < (unsigned)kHotBins;
        if (lds_cold) {
            // hot and cold bins are contiguous in LDS: one atomic, no branch
            atomicAdd(&hot[is_hot ? h * 64 + (int)lane_id() : kHotBins * 64 + bin], 1u);
        } else if (is_hot) {
            atomicAdd(&hot[h * 64 + lane_id()], 1u);
        } else {
            atomicAdd(&global[bin], 1ull);
        }
    }
};

// The same when every bin is in LDS (dict <= kMaxLdsHistBins): the add is
// one atomic with no branch.  The kernels run their tile loop with this
// form when it applies, so the choice is made once per kernel, not per code.
struct LdsHistAll {
    uint32_t* hot;  // kHotBins * 64 lane-private hot bins, then the dict bins
    int lo;
    __device__ __forceinline__ void add(int bin) {
        const int h = bin - lo;
        atomicAdd(&hot[(unsigned)h < (unsigned)kHotBins ? h * 64 + (int)lane_id() : kHotBins * 64 + bin], 1u);
    }
};

__device__ void hist_init(LdsHist& H, uint32_t* lds, int dict, int radius, unsigned long long* global) {
    H.dict = dict;
    H.lo = radius - kHotHalf;
    H.hot = lds;
    H.lds_cold = dict <= kMaxLdsHistBins;
    H.cold = lds + kHotBins * 64;
    H.global = global;
    int total = kHotBins * 64 + (H.lds_cold ? dict : 0);
    for (int i = threadIdx.x; i < total; i += blockDim.x) lds[i] = 0;
    lds_sync();
}

// Folds the lane-private hot copies into the bins, then writes this
// workgroup's histogram densely to partials[blockIdx.x][*] (k_hist_reduce sums
// them) -- or, for dictionaries too large for LDS, adds to the global bins.
__device__ void hist_flush_row(LdsHist& H, uint32_t* __restrict__ dst) {
    lds_sync();
    for (int h = threadIdx.x; h < kHotBins; h += blockDim.x) {
        uint32_t s = 0;
        for (int c = 0; c < 64; c++) s += H.hot[h * 64 + c];
        const int bin = H.lo + h;
        if (s && bin >= 0 && bin < H.dict) H.cold[bin] += s;
    }
    lds_sync();
    for (int b = threadIdx.x; b < H.dict; b += blockDim.x) dst[b] = H.cold[b];
}

__device__ void hist_flush(LdsHist& H, uint32_t* __restrict__ partials) {
    lds_sync();
    for (int h = threadIdx.x; h < kHotBins; h += blockDim.x) {
        uint32_t s = 0;
        for (int c = 0; c < 64; c++) s += H.hot[h * 64 + c];
        int bin = H.lo + h;
        if (s && bin >= 0 && bin < H.dict) {
            if (H.lds_cold)
                H.cold[bin] += s;
            else
                atomicAdd(&H.global[bin], (unsigned long long)s);
        }
    }
    lds_sync();
    if (H.lds_cold) {
        uint32_t* dst = partials + (uint64_t)blockIdx.x * H.dict;
        for (int b = threadIdx.x; b < H.dict; b += blockDim.x) dst[b] = H.cold[b];
    }
}

// Sums the per-workgroup histograms: blockIdx.x picks 256 bins, blockIdx.y a
// group of kHistRows partial rows; one 64-bit atomic per bin and group.
constexpr int kHistRows = 32;

__global__ void __launch_bounds__(256) k_hist_reduce(const uint32_t* __restrict__ partials, int nparts, int dict,
                                                     unsigned long long* __restrict__ hist) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= dict) return;
    const int g0 = blockIdx.y * kHistRows, g1 = min(nparts, g0 + kHistRows);
    uint32_t s = 0;
#pragma unroll 8
    for (int g = g0; g < g1; g++) s += partials[(uint64_t)g * dict + b];
    if (s) atomicAdd(&hist[b], (unsigned long long)s);
}

// Postquant of one element given its Lorenzo prediction (src/dualquant.hh:169-187):
// quantizable = |e| < thr; code = e + R; outlier keeps q.
__device__ __forceinline__ uint32_t postquant(float q, float pred, float thr, float radius_f) {
    float e = q - pred;
    bool quant = __builtin_fabsf(e) < thr;
    return quant ? (uint32_t)(int)(e + radius_f) : 0u;
}

// ------------------------------------------------------------------ 3D

// One wavefront owns a tile of 64 (x) x B (y) x B (z) elements aligned to the
// Lorenzo blocks, so every neighbour outside the tile is the zero halo.  Lane =
// x; the x-1 neighbour comes from DPP, y-1 from the previous row's registers
// and z-1 from the previous plane's registers.  A workgroup is 4 wavefronts
// working on consecutive tiles of a grid-stride loop.
// Halo values {z, y, x face} of the block holding x (PAD kernels only).
struct Halo {
    float z, y, x;
};
__device__ __forceinline__ Halo load_halo(const DQArgs& a, bool xin, uint64_t x, int B, uint64_t by, uint64_t bz) {
    Halo h{0.f, 0.f, 0.f};
    if (xin) {
        const float* p = a.halo + a.halo_stride * ((x / B) + a.hb0 * (by + a.hb1 * bz));
        h.z = p[0];
        h.y = p[1];
        h.x = p[2];
    }
    return h;
}

// PAD: the halo cells hold pad values instead of +0.0 (src/dualquant.hh:
// 226-246, :553-573): a neighbour across the block's z face reads h.z, else
// across its y face h.y, else h.x (src/utils/padding.hh:245, :290).
template <int B, bool PAD, class Hist>
__device__ __forceinline__ void dq3_tile(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                         float* __restrict__ sparse, Hist& H, const DQArgs& a, uint64_t t,
                                         uint32_t lane) {
    const uint64_t sy = a.d0, sz = a.d0 * a.d1;
    {
        const uint64_t tx = t % a.ntiles_x;
        const uint64_t ty = (t / a.ntiles_x) % a.ntiles_y;
        const uint64_t tz = t / (a.ntiles_x * a.ntiles_y);
        const uint64_t x = tx * 64 + lane;
        const uint64_t y0 = ty * B, z0 = tz * B;
        const bool xin = x < a.d0;
        const bool scal = x >= a.x8;  // partial x-group: scalar path
        const bool xstart = (x % B) == 0;
        const float thr = scal ? a.thr_scalar : a.thr_vec;
        Halo h{0.f, 0.f, 0.f};
        if constexpr (PAD) h = load_halo(a, xin, x, B, ty, tz);
        float prevplane[B];
#pragma unroll
        for (int j = 0; j < B; j++) prevplane[j] = 0.f;
        for (int zz = 0; zz < B; zz++) {
            const uint64_t z = z0 + zz;
            if (z >= a.d2) break;
            float qrow = 0.f, xrow = 0.f;      // q and X of row y-1 (this plane)
            float zrow = 0.f, xzrow = 0.f;     // Z and XZ of row y-1
            const float hyz = zz > 0 ? h.y : h.z;  // the y-face halo of this plane's first row
#pragma unroll
            for (int yy = 0; yy < B; yy++) {
                const uint64_t y = y0 + yy;
                if (y < a.d1) {
                    const uint64_t idx = x + y * sy + z * sz;
                    const bool ok = xin;
                    float v = ok ? __builtin_nontemporal_load(&in[idx]) : 0.f;
                    float q = scal ? prequant_scalar(v, a.r_d) : prequant_vec(v, a.r_f);
                    float X = wave_shr1(q);
                    float Z = zz > 0 ? prevplane[yy] : h.z;
                    float XZ = wave_shr1(Z);
                    if (xstart) {
                        X = h.x;
                        XZ = zz > 0 ? h.x : h.z;
                    }
                    float Y = yy > 0 ? qrow : h.y;
                    float XY = yy > 0 ? xrow : h.y;
                    float YZ = yy > 0 ? zrow : hyz;
                    float XYZ = yy > 0 ? xzrow : hyz;
                    float pred;
                    if (!scal) {
                        float dist1 = X + (Y + Z);
                        float dist2 = XY + (XZ + YZ);
                        pred = (XYZ - dist2) + dist1;
                    } else {
                        pred = XYZ - XY - XZ - YZ + X + Y + Z;
                    }
                    uint32_t code = postquant(q, pred, thr, a.radius_f);
                    if (ok) {
                        codes[idx] = (uint16_t)code;
                        if (code == 0) sparse[idx] = q;
                        H.add((int)code);
                    }
                    prevplane[yy] = q;
                    qrow = q;
                    xrow = X;
                    zrow = Z;
                    xzrow = XZ;
                }
            }
        }
    }
}

template <int B, bool PAD = false>
__global__ void __launch_bounds__(256) k_dualquant_3d(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                                      float* __restrict__ sparse, unsigned long long* __restrict__ hist,
                                                      uint32_t* __restrict__ partials, DQArgs a) {
    extern __shared__ uint32_t lds[];
    LdsHist H;
    hist_init(H, lds, a.dict, a.radius, hist);
    dq_resolve(a);
    const uint32_t lane = lane_id();
    const uint64_t ntiles = a.ntiles_x * a.ntiles_y * a.ntiles_z;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    if (H.lds_cold) {
        LdsHistAll L{H.hot, H.lo};
        for (uint64_t tt = wave; tt < ntiles; tt += nwaves)
            dq3_tile<B, PAD>(in, codes, sparse, L, a, a.reverse ? ntiles - 1 - tt : tt, lane);
    } else {
        for (uint64_t tt = wave; tt < ntiles; tt += nwaves)
            dq3_tile<B, PAD>(in, codes, sparse, H, a, a.reverse ? ntiles - 1 - tt : tt, lane);
    }
    hist_flush(H, partials);
}

// ------------------------------------------------------------------ 2D

// Tile of 64 (x) x B (y) per wavefront.  Elements with x >= 8*floor(DIM0/8)
// are never coded by the reference (the `continue` at src/dualquant.hh:494-495):
// code 0, outlier +0.0.
// PAD: a neighbour across the block's x face reads h.x, else across its
// y face h.y (src/utils/padding.hh:226, :275).
template <int B, bool PAD, class Hist>
__device__ __forceinline__ void dq2_tile(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                         float* __restrict__ sparse, Hist& H, const DQArgs& a, uint64_t t,
                                         uint32_t lane) {
    {
        const uint64_t tx = t % a.ntiles_x;
        const uint64_t ty = t / a.ntiles_x;
        const uint64_t x = tx * 64 + lane;
        const uint64_t y0 = ty * B;
        const bool xin = x < a.d0;
        const bool lost = x >= a.x8;
        const bool xstart = (x % B) == 0;
        Halo h{0.f, 0.f, 0.f};
        if constexpr (PAD) h = load_halo(a, xin, x, B, ty, 0);
        float qrow = 0.f, xrow = 0.f;
#pragma unroll
        for (int yy = 0; yy < B; yy++) {
            const uint64_t y = y0 + yy;
            if (y < a.d1) {
                const uint64_t idx = x + y * a.d0;
                float v = xin ? __builtin_nontemporal_load(&in[idx]) : 0.f;
                float q = prequant_vec(v, a.r_f);
                float W = wave_shr1(q);
                if (xstart) W = h.x;
                float N = yy > 0 ? qrow : h.y;
                float NW = yy > 0 ? xrow : (xstart ? h.x : h.y);
                float pred = W + (N - NW);
                uint32_t code = postquant(q, pred, a.thr_vec, a.radius_f);
                if (lost) code = 0;
                if (xin) {
                    codes[idx] = (uint16_t)code;
                    if (code == 0) sparse[idx] = lost ? 0.f : q;
                    H.add((int)code);
                }
                qrow = q;
                xrow = W;
            }
        }
    }
}

template <int B, bool PAD = false>
__global__ void __launch_bounds__(256) k_dualquant_2d(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                                      float* __restrict__ sparse, unsigned long long* __restrict__ hist,
                                                      uint32_t* __restrict__ partials, DQArgs a) {
    extern __shared__ uint32_t lds[];
    LdsHist H;
    hist_init(H, lds, a.dict, a.radius, hist);
    dq_resolve(a);
    const uint32_t lane = lane_id();
    const uint64_t ntiles = a.ntiles_x * a.ntiles_y;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    if (H.lds_cold) {
        LdsHistAll L{H.hot, H.lo};
        for (uint64_t tt = wave; tt < ntiles; tt += nwaves)
            dq2_tile<B, PAD>(in, codes, sparse, L, a, a.reverse ? ntiles - 1 - tt : tt, lane);
    } else {
        for (uint64_t tt = wave; tt < ntiles; tt += nwaves)
            dq2_tile<B, PAD>(in, codes, sparse, H, a, a.reverse ? ntiles - 1 - tt : tt, lane);
    }
    hist_flush(H, partials);
}

// ------------------------------------------------------------------ 4-wide tiles

// Same per-element arithmetic as the kernels above, with each lane owning 4
// consecutive x, so a wavefront covers 256 x and moves 1 KiB per row.  ALN:
// DIM0 % 4 == 0 and the field is 16-byte aligned, so a lane's 4 elements are
// one float4 load and one 8-byte code store; otherwise each element is loaded
// and stored on its own and the lane holding x = DIM0 - 1 has fewer than 4 in
// the field (the rest read 0 and are neither stored nor counted; they are no
// element's x-1 neighbour).  Since 8*floor(DIM0/8) is a multiple of 4, the 4
// elements of a lane all take the same prequant/postquant form, and only
// element 0 can be a block start.
// Element-load forms, UV (the 2D tiles): a lane with all four elements in
// the row moves them as one 16-byte load and one 8-byte store at their
// natural (4- and 2-byte) alignment, which gfx950's global memory accesses
// allow (one instruction instead of four); the row's last lane, with fewer,
// moves them one by one.  The 3D tiles keep element accesses: with their
// next plane's eight rows in flight the unaligned 16-byte loads ran the
// batched 3D form 1.47 -> 2.02 ms, where the 2D form gained 0.69 -> 0.63 ms.
template <bool ALN, bool UV = false>
__device__ __forceinline__ float4 load_x4(const float* __restrict__ p, uint32_t nv) {
    if constexpr (ALN) {
        return *reinterpret_cast<const float4*>(p);
    } else {
        float4 v;
        if (UV && nv == 4) {
            __builtin_memcpy(&v, p, 16);
        } else {
            v = make_float4(p[0], 0.f, 0.f, 0.f);
            if (nv > 1) v.y = p[1];
            if (nv > 2) v.z = p[2];
            if (nv > 3) v.w = p[3];
        }
        return v;
    }
}
template <bool ALN, bool UV = false>
__device__ __forceinline__ void store_x4(uint16_t* __restrict__ p, const uint32_t* c, uint32_t nv) {
    ushort4 cv;
    cv.x = (uint16_t)c[0];
    cv.y = (uint16_t)c[1];
    cv.z = (uint16_t)c[2];
    cv.w = (uint16_t)c[3];
    if constexpr (ALN) {
        *reinterpret_cast<ushort4*>(p) = cv;
    } else if (UV && nv == 4) {
        __builtin_memcpy(p, &cv, 8);
    } else {
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
            if (k < nv) p[k] = (uint16_t)c[k];
    }
}

// One 256 x B x B tile.  SCAL is true only for tiles that contain lanes of
// the partial x-group (x >= 8*floor(DIM0/8)); every other tile runs the AVX
// forms alone, without evaluating the scalar ones.
template <int B, bool SCAL, bool ALN, class Hist>
__device__ __forceinline__ void dq3_x4_tile(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                            float* __restrict__ sparse, Hist& H, const DQArgs& a, uint64_t x,
                                            uint64_t y0, uint64_t z0) {
    const uint64_t sy = a.d0, sz = a.d0 * a.d1;
    const bool xin = x < a.d0;
    const uint32_t nv = ALN ? 4u : (xin ? (uint32_t)min<uint64_t>(4, a.d0 - x) : 0u);
    const bool scal = SCAL && x >= a.x8;
    const bool xstart = (x % B) == 0;
    const float thr = scal ? a.thr_scalar : a.thr_vec;
    const int ny = (int)min((uint64_t)B, a.d1 - y0);
    const int nz = (int)min((uint64_t)B, a.d2 - z0);
    float pp[B][4];
#pragma unroll
    for (int j = 0; j < B; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) pp[j][k] = 0.f;
    // rows of the next plane are loaded while the current one is processed
    float4 vnext[B];
#pragma unroll
    for (int yy = 0; yy < B; yy++)
        vnext[yy] = (xin && yy < ny) ? load_x4<ALN>(in + x + (y0 + yy) * sy + z0 * sz, nv)
                                     : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int zz = 0; zz < nz; zz++) {
        const uint64_t z = z0 + zz;
        float4 vrow[B];
#pragma unroll
        for (int yy = 0; yy < B; yy++) vrow[yy] = vnext[yy];
        if (zz + 1 < nz) {
#pragma unroll
            for (int yy = 0; yy < B; yy++)
                vnext[yy] = (xin && yy < ny) ? load_x4<ALN>(in + x + (y0 + yy) * sy + (z + 1) * sz, nv)
                                             : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        float qr[4] = {0.f, 0.f, 0.f, 0.f}, xr[4] = {0.f, 0.f, 0.f, 0.f};
        float zr[4] = {0.f, 0.f, 0.f, 0.f}, xzr[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int yy = 0; yy < B; yy++) {
            if (yy < ny) {
                const uint64_t idx = x + (y0 + yy) * sy + z * sz;
                const float4 v = vrow[yy];
                float q[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int k = 0; k < 4; k++)
                    q[k] = (SCAL && scal) ? prequant_scalar(q[k], a.r_d) : prequant_vec(q[k], a.r_f);
                float Z[4], X[4], XZ[4];
#pragma unroll
                for (int k = 0; k < 4; k++) Z[k] = zz > 0 ? pp[yy][k] : 0.f;
                X[0] = wave_shr1(q[3]);
                XZ[0] = wave_shr1(Z[3]);
                if (xstart) {
                    X[0] = 0.f;
                    XZ[0] = 0.f;
                }
#pragma unroll
                for (int k = 1; k < 4; k++) {
                    X[k] = q[k - 1];
                    XZ[k] = Z[k - 1];
                }
                uint32_t c[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float Y = yy > 0 ? qr[k] : 0.f;
                    const float XY = yy > 0 ? xr[k] : 0.f;
                    const float YZ = yy > 0 ? zr[k] : 0.f;
                    const float XYZ = yy > 0 ? xzr[k] : 0.f;
                    float pred;
                    if (SCAL && scal) {
                        pred = XYZ - XY - XZ[k] - YZ + X[k] + Y + Z[k];
                    } else {
                        const float dist1 = X[k] + (Y + Z[k]);
                        const float dist2 = XY + (XZ[k] + YZ);
                        pred = (XYZ - dist2) + dist1;
                    }
                    c[k] = postquant(q[k], pred, thr, a.radius_f);
                }
                if (xin) {
                    store_x4<ALN>(codes + idx, c, nv);
                    // outliers are rare: one test for the lane's four codes
                    if ((c[0] == 0) | (c[1] == 0) | (c[2] == 0) | (c[3] == 0)) {
#pragma unroll
                        for (int k = 0; k < 4; k++)
                            if ((ALN || (uint32_t)k < nv) && c[k] == 0) sparse[idx + k] = q[k];
                    }
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (ALN || (uint32_t)k < nv) H.add((int)c[k]);
                }
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    pp[yy][k] = q[k];
                    qr[k] = q[k];
                    xr[k] = X[k];
                    zr[k] = Z[k];
                    xzr[k] = XZ[k];
                }
            }
        }
    }
}

// Tile t of the 256 x B x B grid.
template <int B, bool ALN, class Hist>
__device__ __forceinline__ void dq3_x4_at(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                          float* __restrict__ sparse, Hist& H, const DQArgs& a, uint64_t t,
                                          uint32_t lane) {
    const uint64_t tx = t % a.ntiles_x;
    const uint64_t ty = (t / a.ntiles_x) % a.ntiles_y;
    const uint64_t tz = t / (a.ntiles_x * a.ntiles_y);
    const uint64_t x = tx * 256 + 4 * lane;
    // the partial x-group only exists in the last tile column
    if (tx * 256 + 256 > a.x8 && a.x8 < a.d0)
        dq3_x4_tile<B, true, ALN>(in, codes, sparse, H, a, x, ty * B, tz * B);
    else
        dq3_x4_tile<B, false, ALN>(in, codes, sparse, H, a, x, ty * B, tz * B);
}

template <int B, bool ALN, bool COLD>
__global__ void __launch_bounds__(256) k_dualquant_3d_x4(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                                         float* __restrict__ sparse,
                                                         unsigned long long* __restrict__ hist,
                                                         uint32_t* __restrict__ partials, DQArgs a) {
    extern __shared__ uint32_t lds[];
    LdsHist H;
    hist_init(H, lds, a.dict, a.radius, hist);
    dq_resolve(a);
    const uint32_t lane = lane_id();
    const uint64_t ntiles = a.ntiles_x * a.ntiles_y * a.ntiles_z;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    if constexpr (COLD) {  // every bin in LDS (dict <= kMaxLdsHistBins)
        LdsHistAll L{H.hot, H.lo};
        for (uint64_t tt = wave; tt < ntiles; tt += nwaves)
            dq3_x4_at<B, ALN>(in, codes, sparse, L, a, a.reverse ? ntiles - 1 - tt : tt, lane);
    } else {
        for (uint64_t tt = wave; tt < ntiles; tt += nwaves)
            dq3_x4_at<B, ALN>(in, codes, sparse, H, a, a.reverse ? ntiles - 1 - tt : tt, lane);
    }
    hist_flush(H, partials);
}

template <int B, bool ALN, class Hist>
__device__ __forceinline__ void dq2_x4_tile(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                            float* __restrict__ sparse, Hist& H, const DQArgs& a, uint64_t t,
                                            uint32_t lane) {
    {
        const uint64_t tx = t % a.ntiles_x;
        const uint64_t ty = t / a.ntiles_x;
        const uint64_t x = tx * 256 + 4 * lane;
        const uint64_t y0 = ty * B;
        const bool xin = x < a.d0;
        const uint32_t nv = ALN ? 4u : (xin ? (uint32_t)min<uint64_t>(4, a.d0 - x) : 0u);
        const bool lost = x >= a.x8;
        const bool xstart = (x % B) == 0;
        float qr[4] = {0.f, 0.f, 0.f, 0.f}, xr[4] = {0.f, 0.f, 0.f, 0.f};
        for (int yy = 0; yy < B; yy++) {
            const uint64_t y = y0 + yy;
            if (y >= a.d1) break;
            const uint64_t idx = x + y * a.d0;
            float4 v = xin ? load_x4<ALN, true>(in + idx, nv) : make_float4(0.f, 0.f, 0.f, 0.f);
            float q[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 4; k++) q[k] = prequant_vec(q[k], a.r_f);
            float W[4];
            W[0] = wave_shr1(q[3]);
            if (xstart) W[0] = 0.f;
#pragma unroll
            for (int k = 1; k < 4; k++) W[k] = q[k - 1];
            uint32_t c[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float N = yy > 0 ? qr[k] : 0.f;
                const float NW = yy > 0 ? xr[k] : 0.f;
                const float pred = W[k] + (N - NW);
                c[k] = lost ? 0u : postquant(q[k], pred, a.thr_vec, a.radius_f);
            }
            if (xin) {
                store_x4<ALN, true>(codes + idx, c, nv);
                // outliers are rare: one test for the lane's four codes
                if ((c[0] == 0) | (c[1] == 0) | (c[2] == 0) | (c[3] == 0)) {
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if ((ALN || (uint32_t)k < nv) && c[k] == 0) sparse[idx + k] = lost ? 0.f : q[k];
                }
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (ALN || (uint32_t)k < nv) H.add((int)c[k]);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                qr[k] = q[k];
                xr[k] = W[k];
            }
        }
    }
}

template <int B, bool ALN, bool COLD>
__global__ void __launch_bounds__(256) k_dualquant_2d_x4(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                                         float* __restrict__ sparse,
                                                         unsigned long long* __restrict__ hist,
                                                         uint32_t* __restrict__ partials, DQArgs a) {
    extern __shared__ uint32_t lds[];
    LdsHist H;
    hist_init(H, lds, a.dict, a.radius, hist);
    dq_resolve(a);
    const uint32_t lane = lane_id();
    const uint64_t ntiles = a.ntiles_x * a.ntiles_y;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    if constexpr (COLD) {  // every bin in LDS (dict <= kMaxLdsHistBins)
        LdsHistAll L{H.hot, H.lo};
        for (uint64_t tt = wave; tt < ntiles; tt += nwaves)
            dq2_x4_tile<B, ALN>(in, codes, sparse, L, a, a.reverse ? ntiles - 1 - tt : tt, lane);
    } else {
        for (uint64_t tt = wave; tt < ntiles; tt += nwaves)
            dq2_x4_tile<B, ALN>(in, codes, sparse, H, a, a.reverse ? ntiles - 1 - tt : tt, lane);
    }
    hist_flush(H, partials);
}

// ------------------------------------------------------------------ 1D

// One thread per run of 8 consecutive elements (the AVX group).  The element
// before the run comes from lane-1 (DPP) or, for lane 0, from memory.  All
// elements take the AVX prequant (src/dualquant.hh:83-93); the postquant test
// is the AVX one in blocks ending before DIM0 and the scalar one in the last
// block (src/dualquant.hh:114, :190-205).
// PAD: a block's first element predicts from the pad value, which 1D block
// and edge padding keep raw (src/dualquant.hh:37-42), so the prediction error
// may be non-integral: then the AVX forms' sqrt(e*e) < R test and
// round-to-nearest conversion (_mm256_cvtps_epi32, :172) and the scalar
// forms' |e| < R and truncation (:200) are evaluated as written.
// Groups wb .. wb + 63 (wave tile wb / 64).
template <int B, bool PAD, class Hist>
__device__ __forceinline__ void dq1_tile(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                         float* __restrict__ sparse, Hist& H, const DQArgs& a, uint64_t wb,
                                         uint32_t lane, bool aligned) {
    const uint64_t ngroups = (a.n + 7) / 8;
    {
        const uint64_t g = wb + lane;
        const uint64_t x0 = g * 8;
        float q[8];
        if (x0 + 8 <= a.n && aligned) {
            const float4* p = reinterpret_cast<const float4*>(in + x0);
            float4 v0 = p[0], v1 = p[1];
            q[0] = v0.x; q[1] = v0.y; q[2] = v0.z; q[3] = v0.w;
            q[4] = v1.x; q[5] = v1.y; q[6] = v1.z; q[7] = v1.w;
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) q[j] = (x0 + j < a.n) ? in[x0 + j] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) q[j] = prequant_vec(q[j], a.r_f);
        float prev = wave_shr1(q[7]);
        if (lane == 0 && x0 > 0 && g < ngroups) prev = prequant_vec(in[x0 - 1], a.r_f);
        if ((x0 % B) == 0) {
            prev = 0.f;
            if constexpr (PAD)
                if (g < ngroups) prev = a.halo[a.halo_stride * (x0 / B) + 2];
        }
        if (g < ngroups) {
            const uint64_t bstart = x0 - (x0 % B);
            const bool vec = bstart + B < a.d0;
            const float thr = vec ? a.thr_vec : a.thr_scalar;
            const bool whole = x0 + 8 <= a.n;  // the group's 8 codes go out as one 16-byte store
            uint32_t cw[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint64_t idx = x0 + j;
                float pred = j == 0 ? prev : q[j - 1];
                uint32_t code;
                if constexpr (PAD) {
                    const float e = q[j] - pred;
                    const bool quant = vec ? __builtin_sqrtf(e * e) < a.radius_f : __builtin_fabsf(e) < a.radius_f;
                    code = quant ? (uint32_t)(int)(vec ? __builtin_rintf(e + a.radius_f) : e + a.radius_f) : 0u;
                } else {
                    code = postquant(q[j], pred, thr, a.radius_f);
                }
                cw[j >> 1] |= (code & 0xFFFFu) << (16 * (j & 1));
                if (idx < a.n) {
                    if (!whole) codes[idx] = (uint16_t)code;
                    if (code == 0) sparse[idx] = q[j];
                    H.add((int)code);
                }
            }
            // codes is 16-byte aligned and x0 a multiple of 8
            if (whole) *reinterpret_cast<uint4*>(codes + x0) = make_uint4(cw[0], cw[1], cw[2], cw[3]);
        }
    }
}

template <int B, bool PAD = false>
__global__ void __launch_bounds__(256) k_dualquant_1d(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                                      float* __restrict__ sparse, unsigned long long* __restrict__ hist,
                                                      uint32_t* __restrict__ partials, DQArgs a) {
    extern __shared__ uint32_t lds[];
    LdsHist H;
    hist_init(H, lds, a.dict, a.radius, hist);
    dq_resolve(a);
    const uint32_t lane = lane_id();
    const uint64_t ngroups = (a.n + 7) / 8;
    const bool aligned = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    if (H.lds_cold) {
        LdsHistAll L{H.hot, H.lo};
        for (uint64_t wb = wave * 64; wb < ngroups; wb += nwaves * 64)
            dq1_tile<B, PAD>(in, codes, sparse, L, a, wb, lane, aligned);
    } else {
        for (uint64_t wb = wave * 64; wb < ngroups; wb += nwaves * 64)
            dq1_tile<B, PAD>(in, codes, sparse, H, a, wb, lane, aligned);
    }
    hist_flush(H, partials);
}

// ------------------------------------------------------------------ encode

__device__ __forceinline__ uint32_t load_len(const uint8_t* len_tab, uint32_t s) { return len_tab[s]; }

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T n = __shfl_up(v, d);
        if (lane >= (uint32_t)d) v += n;
    }
    return v;
}

// DPP move (GFX9 encodings): lanes without a source, or in rows outside
// ROWS, read 0.  Every lane of the wavefront must be active.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ uint32_t dpp_move(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false);
}

// Inclusive scan of 32-bit values over the wavefront in six DPP steps and no
// LDS traffic: row_shr 1, 2, 4, 8 scan each row of 16 lanes, row_bcast:15
// adds row 0's total to row 1 and row 2's to row 3, row_bcast:31 adds the
// total of rows 0-1 to rows 2-3.  OR = true gives the OR-scan.
template <bool OR = false>
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t v) {
    auto op = [](uint32_t a, uint32_t b) { return OR ? (a | b) : (a + b); };
    v = op(v, dpp_move<0x111>(v));
    v = op(v, dpp_move<0x112>(v));
    v = op(v, dpp_move<0x114>(v));
    v = op(v, dpp_move<0x118>(v));
    v = op(v, dpp_move<0x142, 0xA>(v));
    v = op(v, dpp_move<0x143, 0xC>(v));
    return v;
}

__device__ __forceinline__ uint32_t lane63(uint32_t v) { return (uint32_t)__builtin_amdgcn_readlane((int)v, 63); }

// Exclusive scan across a 256-thread workgroup; returns the thread's prefix and
// the workgroup total through *total.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* total, T* smem /* >= 4 */) {
    T inc = wave_incl_scan(v);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 63) smem[w] = inc;
    lds_sync();
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        T s = smem[i];
        if (i < w) off += s;
        tot += s;
    }
    lds_sync();
    *total = tot;
    return off + inc - v;
}

// Bits and outliers of each 1024-symbol chunk; one wavefront per chunk, 16
// consecutive symbols per lane.  Persistent workgroups load the code-length
// table once; LDS_TAB fixes its address space at compile time (ds_read, not
// FLAT).
__device__ __forceinline__ void load_syms(const uint16_t* __restrict__ codes, uint64_t n, uint64_t base,
                                          uint32_t* sym) {
    if (base + kSymPerThread <= n) {
        const uint4* p = reinterpret_cast<const uint4*>(codes + base);
        const uint4 u0 = p[0], u1 = p[1];
        const uint32_t w[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
        for (int j = 0; j < 8; j++) {
            sym[2 * j] = w[j] & 0xFFFF;
            sym[2 * j + 1] = w[j] >> 16;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kSymPerThread; j++) sym[j] = (base + j < n) ? codes[base + j] : 0xFFFFFFFFu;
    }
}

template <bool LDS_TAB>
__global__ void __launch_bounds__(256) k_chunk_bits(const uint16_t* __restrict__ codes, uint64_t n, uint64_t nchunks,
                                                    const uint8_t* __restrict__ len_tab, int dict,
                                                    uint32_t* __restrict__ chunk_bits, uint32_t* __restrict__ chunk_zeros) {
    __shared__ uint8_t slen[LDS_TAB ? kMaxLdsLenTab : 4];
    if (LDS_TAB)
        for (int i = threadIdx.x; i < dict; i += blockDim.x) slen[i] = len_tab[i];
    lds_sync();
    const uint32_t lane = lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t k = blockIdx.x * (uint64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); k < nchunks; k += nw) {
        uint32_t sym[kSymPerThread];
        load_syms(codes, n, k * kChunk + lane * kSymPerThread, sym);
        uint32_t bits = 0, zeros = 0;
#pragma unroll
        for (int j = 0; j < kSymPerThread; j++) {
            const uint32_t sj = sym[j];
            if (sj != 0xFFFFFFFFu) {
                bits += LDS_TAB ? slen[sj] : len_tab[sj];
                zeros += sj == 0;
            }
        }
        // bits (< 2^17) and zeros (< 2^11) summed in one wavefront scan
        const uint32_t tot = lane63(wave_scan_dpp(bits | (zeros << 17)));
        if (lane == 0) {
            chunk_bits[k] = tot & 0x1FFFF;
            chunk_zeros[k] = tot >> 17;
        }
    }
}

// Archive sizes of src/vecsz_interface.hh:189-202 from the scanned chunk
// totals, and the two length fields of the payload header
// (src/utils/datapack.hh:288) written in place.
__device__ __forceinline__ void enc_summary(uint64_t tb, uint64_t tz, EncSummary* __restrict__ sum,
                                            uint8_t* __restrict__ archive, uint64_t meta_bytes, uint64_t pad_bytes,
                                            uint64_t tree_bytes, uint64_t out_cap, uint32_t* __restrict__ err,
                                            uint64_t* __restrict__ index = nullptr) {
    if (index) {  // decode-index header: stream bits and outliers
        index[3] = tb;
        index[4] = tz;
        // the per-segment arrays' padding slots (past the last segment, up to
        // a whole word), so that an index is the same byte for byte on every
        // run; the packer writes only the slots of existing segments
        const IdxLayout L = idx_layout(index[1]);
        uint32_t* rb = reinterpret_cast<uint32_t*>(index + L.rel_bits);
        for (uint64_t g = L.nseg; g < (L.nseg + 1) / 2 * 2; g++) rb[g] = 0;
        uint16_t* ro = reinterpret_cast<uint16_t*>(index + L.rel_outs);
        for (uint64_t g = L.nseg; g < (L.nseg + 3) / 4 * 4; g++) ro[g] = 0;
    }
    const uint64_t enc = (tb + 7) / 8;
    const uint64_t huff = 8 + tree_bytes + enc;
    const uint64_t stream_byte = meta_bytes + pad_bytes + 8 + tree_bytes;
    EncSummary s;
    s.total_bits = tb;
    s.n_outlier = tz;
    s.huff_size = huff;
    s.archive_size = meta_bytes + pad_bytes + huff + 4 * tz;
    s.outlier_byte_off = meta_bytes + pad_bytes + huff;
    s.stream_bit_off = stream_byte * 8;
    s.stream_end_byte = stream_byte + enc;
    *sum = s;
    if (s.archive_size > out_cap) {
        *err = 1;
    } else {
        uint8_t* f = archive + meta_bytes - 16;
        for (int i = 0; i < 8; i++) f[i] = (uint8_t)(huff >> (8 * i));
        for (int i = 0; i < 8; i++) f[8 + i] = (uint8_t)(tz >> (8 * i));
    }
}

__global__ void k_enc_summary(const uint64_t* __restrict__ totals, EncSummary* __restrict__ sum,
                              uint8_t* __restrict__ archive, uint64_t meta_bytes, uint64_t pad_bytes,
                              uint64_t tree_bytes, uint64_t out_cap, uint32_t* __restrict__ err,
                              uint64_t* __restrict__ index) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        enc_summary(totals[0], totals[1], sum, archive, meta_bytes, pad_bytes, tree_bytes, out_cap, err, index);
}

__global__ void k_enc_summary_dev(const uint64_t* __restrict__ totals, EncSummary* __restrict__ sum,
                                  uint8_t* __restrict__ archive, uint64_t meta_bytes, uint64_t pad_bytes,
                                  const CodebookInfo* __restrict__ info, uint64_t out_cap, uint32_t* __restrict__ err,
                                  uint64_t* __restrict__ index) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (info->status != 0) {
        *err = 2;
        return;
    }
    enc_summary(totals[0], totals[1], sum, archive, meta_bytes, pad_bytes, info->tree_bytes, out_cap, err, index);
}

// Bit packing, one wavefront per 1024-symbol chunk, no workgroup barriers or
// global atomics.  Lane L holds 16 consecutive symbols; a wavefront scan gives
// its first bit inside the chunk.  The lanes OR their codewords into a
// wave-private LDS image of the 32-bit words the chunk touches; the wavefront
// completes the chunk's last word with the leading bits of the next chunk
// (fewer than 32 bits, always inside the next chunk's first 32 codewords since
// a codeword has at least one bit) and stores the image with coalesced dword
// writes.  A word shared by two chunks belongs to the earlier one; the stream's
// first word keeps the tree bytes in front of it.
struct Code {
    uint32_t lo, hi, len;  // codeword = hi:lo, right-aligned, len bits
};

// LDS_BOOK: LDS table entry (codeword << 6) | length for codewords of at most
// 26 bits; longer ones (rare symbols) carry kLongCode and are read from the
// global tables.
constexpr uint32_t kLongCode = 63;

__device__ __forceinline__ Code lookup_global(uint32_t s, const uint32_t* __restrict__ cw_hi,
                                              const uint32_t* __restrict__ cw_lo, const uint8_t* __restrict__ len_tab) {
    Code c;
    c.len = len_tab[s];
    c.lo = cw_lo[s];
    c.hi = c.len > 32 ? cw_hi[s] : 0;
    return c;
}

template <bool LDS_BOOK>
__device__ __forceinline__ Code lookup(uint32_t s, const uint32_t* tab, const uint32_t* __restrict__ cw_hi,
                                       const uint32_t* __restrict__ cw_lo, const uint8_t* __restrict__ len_tab) {
    if (LDS_BOOK) {
        const uint32_t e = tab[s];
        if ((e & 63) != kLongCode) return Code{e >> 6, 0, e & 63};
    }
    return lookup_global(s, cw_hi, cw_lo, len_tab);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Raw codes of one chunk as lane L sees them: its 16 symbols (two 16-byte
// loads), how many of them exist, and, for lanes 0..31, symbol L of the next
// chunk (the next chunk's leading bits complete this chunk's last word).
struct ChunkRaw {
    uint32_t w[8];
    uint32_t nvalid;
    uint32_t next;  // 0xFFFFFFFF: none
    uint64_t bit_off, out_off, nz;  // nz: outliers in the chunk
};

// Symbols past the field read as `fill` (an LDS table slot that holds 0).
__device__ __forceinline__ void load_chunk(const uint16_t* __restrict__ codes, uint64_t n, uint64_t nchunks,
                                           const uint64_t* __restrict__ bit_off,
                                           const uint64_t* __restrict__ out_off, uint64_t n_outlier, uint64_t k,
                                           uint32_t lane, uint32_t fill, ChunkRaw& r) {
    const uint64_t base = k * kChunk + lane * kSymPerThread;
    r.nvalid = base >= n ? 0u : (uint32_t)min<uint64_t>(kSymPerThread, n - base);
    if (r.nvalid == kSymPerThread) {
        const uint4* p = reinterpret_cast<const uint4*>(codes + base);
        const uint4 u0 = p[0], u1 = p[1];
        r.w[0] = u0.x, r.w[1] = u0.y, r.w[2] = u0.z, r.w[3] = u0.w;
        r.w[4] = u1.x, r.w[5] = u1.y, r.w[6] = u1.z, r.w[7] = u1.w;
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) r.w[j] = 0;
#pragma unroll
        for (int j = 0; j < kSymPerThread; j++)
            r.w[j >> 1] |= ((uint32_t)j < r.nvalid ? (uint32_t)codes[base + j] : fill) << (16 * (j & 1));
    }
    const uint64_t i = (k + 1) * kChunk + lane;
    r.next = (lane < 32 && k + 1 < nchunks && i < n) ? (uint32_t)codes[i] : 0xFFFFFFFFu;
    r.bit_off = bit_off[k];
    r.out_off = out_off[k];
    r.nz = (k + 1 < nchunks ? out_off[k + 1] : n_outlier) - r.out_off;
}

// Codebook layouts of k_pack: the global tables; LDS entries (codeword << 6)
// | length with kLongCode for codewords over 26 bits; or, when no codeword
// exceeds 32 bits, an LDS codeword array and an LDS length array.  The LDS
// forms hold one more slot, `dict`, that is 0: symbols past the field.
enum { kBookGlobal = 0, kBookPacked = 1, kBookSplit = 2 };

__host__ __device__ constexpr uint32_t book_words(int book, int dict) {
    return book == kBookPacked ? (uint32_t)dict + 1 : book == kBookSplit ? (uint32_t)dict + 1 + ((uint32_t)dict + 4) / 4 : 0;
}

// Chunks of one decode-index group (its segments' bit offsets are stored
// relative to the group's first chunk).
constexpr uint64_t kChunksPerGroup = kIdxGroup * kIdxSeg / kChunk;
static_assert(kChunksPerGroup >= 1 && kIdxGroup * kIdxSeg % kChunk == 0, "an index group spans whole chunks");

// Where pack_chunk writes: outlier r goes to outl + 4 r (little-endian
// bytes); the bitstream starts at archive bit stream_bit_off; bytes at or past
// stream_end_byte are not the stream's (~0: the end of the field's last chunk,
// known only to that chunk); words that end past cap are not written (err is
// set instead); no_stream: a one-symbol tree, no bitstream.
struct PackTarget {
    uint8_t* outl;
    uint64_t stream_bit_off;
    uint64_t stream_end_byte;
    uint64_t cap;
    uint32_t* err;
    bool no_stream;
};

__device__ __forceinline__ PackTarget pack_target(const EncSummary& S, uint8_t* archive) {
    PackTarget T;
    T.outl = archive + S.outlier_byte_off;
    T.stream_bit_off = S.stream_bit_off;
    T.stream_end_byte = S.stream_end_byte;
    T.cap = ~0ull;
    T.err = nullptr;
    T.no_stream = S.total_bits == 0;
    return T;
}

// Packs chunk k (codes already in `cur`) into the archive; see k_pack.
// index: the decode index (layout L), or null; gbits / gouts: the first bit
// and outlier rank of the chunk's index group (its first chunk's offsets).
template <int BOOK>
__device__ __forceinline__ void pack_chunk(const ChunkRaw& cur, uint64_t k, uint64_t nchunks, uint32_t lane,
                                           const uint32_t* tab, const uint8_t* tlen,
                                           const uint32_t* __restrict__ cw_hi, const uint32_t* __restrict__ cw_lo,
                                           const uint8_t* __restrict__ len_tab, const PackTarget& T,
                                           const float* __restrict__ sparse, uint8_t* __restrict__ archive,
                                           uint32_t* stage, uint64_t* __restrict__ index = nullptr,
                                           IdxLayout L = IdxLayout{}, uint64_t gbits = 0, uint64_t gouts = 0) {
    constexpr bool LDS_BOOK = BOOK != kBookGlobal;
    uint32_t* out32 = reinterpret_cast<uint32_t*>(archive);
    const uint64_t base = k * kChunk + lane * kSymPerThread;

    // the LDS books keep the raw entries (one or two registers per symbol)
    uint32_t ent[LDS_BOOK ? kSymPerThread : 1];
    uint32_t elen[BOOK == kBookSplit ? kSymPerThread : 1];
    Code c[LDS_BOOK ? 1 : kSymPerThread];
    uint32_t tbits = 0, tz = 0, long_mask = 0;
#pragma unroll
    for (int j = 0; j < kSymPerThread; j++) {
        const uint32_t sj = (cur.w[j >> 1] >> (16 * (j & 1))) & 0xFFFF;
        if constexpr (BOOK == kBookPacked) {
            ent[j] = tab[sj];
            const uint32_t l = ent[j] & 63;
            tbits += l;
            long_mask |= (uint32_t)(l == kLongCode) << j;
        } else if constexpr (BOOK == kBookSplit) {
            ent[j] = tab[sj];
            elen[j] = tlen[sj];
            tbits += elen[j];
        } else {
            const bool valid = (uint32_t)j < cur.nvalid;
            c[j] = Code{0, 0, 0};
            if (valid) c[j] = lookup_global(sj, cw_hi, cw_lo, len_tab);
            tbits += c[j].len;
        }
    }
    if (cur.nz) {  // outliers are rare: count them per lane only in chunks that hold some
#pragma unroll
        for (int j = 0; j < kSymPerThread; j++)
            tz += (uint32_t)j < cur.nvalid && ((cur.w[j >> 1] >> (16 * (j & 1))) & 0xFFFF) == 0;
    }
    if (BOOK == kBookPacked && long_mask) {
        // codewords of more than 26 bits: true lengths from the global table
#pragma unroll
        for (int j = 0; j < kSymPerThread; j++)
            if ((long_mask >> j) & 1) tbits += (uint32_t)len_tab[(cur.w[j >> 1] >> (16 * (j & 1))) & 0xFFFF] - kLongCode;
    }
    // bits (low 17) and outliers (high 15) in one wavefront scan
    const uint32_t v = tbits | (tz << 17);
    const uint32_t incl = wave_scan_dpp(v);
    const uint32_t excl = incl - v;
    const uint32_t cbits = lane63(incl) & 0x1FFFF;

    // decode index: lanes holding a segment's first symbol record its first
    // codeword bit and outlier rank, relative to its group's (the first
    // segment of a group also records the group's)
    constexpr uint32_t kLanesPerSeg = kIdxSeg / kSymPerThread;
    if (index && (lane % kLanesPerSeg) == 0) {
        const uint64_t g = k * (kChunk / kIdxSeg) + lane / kLanesPerSeg;
        if (g < L.nseg) {
            const uint64_t G = g / kIdxGroup;
            const uint64_t bits = cur.bit_off + (excl & 0x1FFFF), outs = cur.out_off + (excl >> 17);
            if (g % kIdxGroup == 0) {
                index[L.grp + 2 * G] = bits;
                index[L.grp + 2 * G + 1] = outs;
            }
            reinterpret_cast<uint32_t*>(index + L.rel_bits)[g] = (uint32_t)(bits - gbits);
            reinterpret_cast<uint16_t*>(index + L.rel_outs)[g] = (uint16_t)(outs - gouts);
        }
    }

    // outlier compaction, index order
    if (tz) {
        uint64_t r = cur.out_off + (excl >> 17);
#pragma unroll
        for (int j = 0; j < kSymPerThread; j++) {
            const uint32_t sj = (cur.w[j >> 1] >> (16 * (j & 1))) & 0xFFFF;
            if ((uint32_t)j < cur.nvalid && sj == 0) {
                const uint32_t bv = __float_as_uint(sparse[base + j]);
                uint8_t* dst = T.outl + 4 * r;
                dst[0] = (uint8_t)bv;
                dst[1] = (uint8_t)(bv >> 8);
                dst[2] = (uint8_t)(bv >> 16);
                dst[3] = (uint8_t)(bv >> 24);
                r++;
            }
        }
    }
    if (T.no_stream) return;  // one-symbol tree: no bitstream

    const uint64_t Ak = T.stream_bit_off + cur.bit_off;
    const uint64_t end_byte =
        T.stream_end_byte != ~0ull ? T.stream_end_byte : k + 1 == nchunks ? (Ak + cbits + 7) / 8 : ~0ull;
    const uint64_t wbase = Ak >> 5;
    const uint32_t a0 = (uint32_t)(Ak & 31);
    const uint32_t nwords = (a0 + cbits + 31) >> 5;
    for (uint32_t i = lane; i < nwords; i += kWave) stage[i] = 0;
    wave_lds_sync();

    // Each codeword is ORed into the one or two stage words it overlaps
    // (codeword at bit offset o of word w, MSB first).  The second OR may
    // add zero to the word after the chunk's last; the stage has room.
    uint32_t pos = a0 + (excl & 0x1FFFF);
    auto place = [&](uint32_t cw, uint32_t l) {  // l <= 32; l = 0 places nothing (cw = 0)
        const uint32_t o = pos & 31;
        const uint64_t v = (uint64_t)cw << ((64 - o - l) & 63);
        uint32_t* w = stage + (pos >> 5);
        atomicOr(w, (uint32_t)(v >> 32));
        atomicOr(w + 1, (uint32_t)v);
        pos += l;
    };
    auto place_code = [&](const Code& c) {
        if (c.len > 32) {
            place(c.hi, c.len - 32);
            place(c.lo, 32);
        } else if (c.len) {
            place(c.lo, c.len);
        }
    };
    if constexpr (BOOK == kBookSplit) {
#pragma unroll
        for (int j = 0; j < kSymPerThread; j++) place(ent[j], elen[j]);  // absent symbols: length 0
    } else if constexpr (BOOK == kBookPacked) {
        if (!__any(long_mask != 0)) {
#pragma unroll
            for (int j = 0; j < kSymPerThread; j++) place(ent[j] >> 6, ent[j] & 63);  // absent symbols: ent 0
        } else {
#pragma unroll
            for (int j = 0; j < kSymPerThread; j++) {
                const uint32_t l = ent[j] & 63;
                if (l == kLongCode)
                    place_code(lookup_global((cur.w[j >> 1] >> (16 * (j & 1))) & 0xFFFF, cw_hi, cw_lo, len_tab));
                else if (l)
                    place(ent[j] >> 6, l);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < kSymPerThread; j++) place_code(c[j]);
    }

    // complete the last word with the next chunk's leading bits
    const uint32_t e = (a0 + cbits) & 31;
    if (e != 0 && k + 1 < nchunks) {
        Code d{0, 0, 0};
        if (cur.next != 0xFFFFFFFFu) {
            if constexpr (BOOK == kBookSplit)
                d = Code{tab[cur.next], 0, tlen[cur.next]};
            else
                d = lookup<LDS_BOOK>(cur.next, tab, cw_hi, cw_lo, len_tab);
        }
        const uint32_t p = wave_scan_dpp(d.len) - d.len;
        uint32_t t = 0;
        if (d.len && p < 32) {
            const uint32_t kk = min(d.len, 32u - p);
            const uint64_t cw = ((uint64_t)d.hi << 32) | d.lo;
            t = (uint32_t)((cw >> (d.len - kk)) << (32 - p - kk));
        }
        t = lane63(wave_scan_dpp<true>(t));
        if (lane == 0) atomicOr(&stage[nwords - 1], t >> e);
    }
    wave_lds_sync();

    for (uint32_t i = ((a0 != 0 && k != 0) ? 1u : 0u) + lane; i < nwords; i += kWave) {
        const uint64_t Wg = wbase + i;
        uint32_t wv = stage[i];
        if (k == 0 && i == 0 && a0) {
            // keep the tree bytes that precede the stream in this word
            const uint32_t mask = ~0u << (32 - a0);
            wv = (__builtin_bswap32(out32[Wg]) & mask) | (wv & ~mask);
        }
        // bytes at or past the stream end are not the stream's; bytes at or
        // past cap are outside the buffer (a stream byte there: too small)
        const uint64_t lim = end_byte < T.cap ? end_byte : T.cap;
        if ((Wg + 1) * 4 > lim) {
            uint8_t* bp = archive + Wg * 4;
            for (int j = 0; j < 4; j++)
                if (Wg * 4 + j < lim) bp[j] = (uint8_t)(wv >> (24 - 8 * j));
            if (T.cap < min(end_byte, (Wg + 1) * 4)) atomicOr(T.err, 1u);
        } else {
            out32[Wg] = __builtin_bswap32(wv);
        }
    }
    wave_lds_sync();
}

template <int BOOK>
__global__ void __launch_bounds__(256) k_pack(const uint16_t* __restrict__ codes, uint64_t n,
                                              const uint32_t* __restrict__ cw_hi, const uint32_t* __restrict__ cw_lo,
                                              const uint8_t* __restrict__ len_tab, int dict,
                                              const uint64_t* __restrict__ bit_off, const uint64_t* __restrict__ out_off,
                                              uint64_t nchunks, const EncSummary* __restrict__ sum,
                                              const float* __restrict__ sparse, uint8_t* __restrict__ archive,
                                              const uint32_t* __restrict__ err, uint32_t stage_words,
                                              uint64_t* __restrict__ index, IdxLayout L) {
    constexpr bool LDS_BOOK = BOOK != kBookGlobal;
    if (*err) return;
    extern __shared__ uint32_t lds[];
    uint32_t* tab = lds;                                                        // packed entries, or split codewords
    uint8_t* tlen = reinterpret_cast<uint8_t*>(lds + (uint32_t)dict + 1);       // split lengths
    if (BOOK == kBookPacked) {
        for (int i = threadIdx.x; i <= dict; i += blockDim.x) {
            const uint32_t l = i < dict ? len_tab[i] : 0u;
            tab[i] = l == 0 ? 0u : l <= 26 ? (cw_lo[i] << 6) | l : kLongCode;
        }
    } else if (BOOK == kBookSplit) {
        for (int i = threadIdx.x; i <= dict; i += blockDim.x) {
            tab[i] = i < dict ? cw_lo[i] : 0u;
            tlen[i] = i < dict ? len_tab[i] : 0u;
        }
    }
    lds_sync();
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t* stage = lds + book_words(BOOK, dict) + wave * stage_words;
    const EncSummary S = *sum;
    const uint32_t lane = lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x >> 6);
    uint64_t k = blockIdx.x * (uint64_t)(blockDim.x >> 6) + wave;
    if (k >= nchunks) return;  // no workgroup barrier follows

    // software pipeline: the next chunk's loads are in flight while this one is packed
    const uint32_t fill = LDS_BOOK ? (uint32_t)dict : 0u;
    ChunkRaw nxt;
    load_chunk(codes, n, nchunks, bit_off, out_off, S.n_outlier, k, lane, fill, nxt);
    const PackTarget T = pack_target(S, archive);
    for (; k < nchunks; k += nw) {
        const ChunkRaw cur = nxt;
        if (k + nw < nchunks) load_chunk(codes, n, nchunks, bit_off, out_off, S.n_outlier, k + nw, lane, fill, nxt);
        const uint64_t k0 = k / kChunksPerGroup * kChunksPerGroup;  // the index group's first chunk
        pack_chunk<BOOK>(cur, k, nchunks, lane, tab, tlen, cw_hi, cw_lo, len_tab, T, sparse, archive, stage, index, L,
                         index ? bit_off[k0] : 0, index ? out_off[k0] : 0);
    }
}

// ------------------------------------------------------------------ fused single-pass encode
//
// k_encode does the work of k_chunk_bits, the chunk scan, k_enc_summary and
// k_pack in one pass, so the field's codes are read from HBM once (a
// wavefront keeps its chunks' codes in registers until it packs them).  The
// chunks are cut into partitions of kEncPartChunks (kEncWaveChunks per
// wavefront, four wavefronts per workgroup).  A workgroup takes the next
// partition from a counter, sums every chunk's
// bits and outliers, and finds the partition's first bit and outlier rank by
// a decoupled look-back over the partitions before it: a partition publishes
// its own totals (aggregate) as soon as it has summed them, and its
// inclusive prefix once it knows it.  Each published value shares one 64-bit
// word with its 2-bit flag and is written and read with agent-scope atomics,
// so a value never arrives without its flag and no cache maintenance is
// needed between the XCDs.  Partitions are handed out in counter order, so
// every partition a workgroup waits for was taken earlier by a running
// workgroup, which publishes its aggregate without waiting on anyone: the
// look-back always ends.  Then each chunk is packed exactly as k_pack packs
// it.  The outlier section's offset depends on the total stream length, so
// outliers go to a compact scratch array by rank, and k_enc_finish writes
// the sizes and moves them into the archive.
// WC chunks per wavefront (a divisor or multiple of an index group)
constexpr uint64_t kStAgg = 1ull << 62, kStIncl = 2ull << 62, kStVal = (1ull << 62) - 1;

__device__ __forceinline__ uint64_t ld_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

// Exclusive prefixes (bits, outliers) of partition `part` from the status
// words of the partitions before it, found by the whole workgroup (every
// thread must call it; control flow is uniform, from LDS values all threads
// read alike): in each round thread t looks at partition base - t, so 256
// predecessors are examined per round trip to memory.  Partition -1 counts
// as an inclusive prefix of 0.  sh: 24 words of LDS.
__device__ void lookback_wg(const uint64_t* st_b, const uint64_t* st_z, uint64_t part, uint64_t* sh, uint64_t& pre_b,
                            uint64_t& pre_z) {
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint64_t* s_mask = sh;     // 4 waves x {incl_b, ready_b, incl_z, ready_z}
    uint64_t* s_sum = sh + 16; // 4 waves x {bits, zeros}
    uint64_t sum_b = 0, sum_z = 0;
    bool done_b = false, done_z = false;
    int64_t base = (int64_t)part - 1;
    for (;;) {
        const int64_t q = base - (int64_t)t;
        const uint64_t wb = done_b ? 0 : q >= 0 ? ld_status(st_b + q) : kStIncl;
        const uint64_t wz = done_z ? 0 : q >= 0 ? ld_status(st_z + q) : kStIncl;
        const uint64_t ib = __ballot((wb >> 62) == 2), rb = __ballot((wb >> 62) != 0);
        const uint64_t iz = __ballot((wz >> 62) == 2), rz = __ballot((wz >> 62) != 0);
        if (lane == 0) {
            s_mask[4 * wave] = ib;
            s_mask[4 * wave + 1] = rb;
            s_mask[4 * wave + 2] = iz;
            s_mask[4 * wave + 3] = rz;
        }
        lds_sync();
        // nearest inclusive prefix (j = 256: none in this window) and whether
        // every partition up to it has published
        int jb = 256, jz = 256;
        bool okb = true, okz = true;
#pragma unroll
        for (int w = 3; w >= 0; w--) {
            if (s_mask[4 * w]) jb = 64 * w + __builtin_ctzll(s_mask[4 * w]);
            if (s_mask[4 * w + 2]) jz = 64 * w + __builtin_ctzll(s_mask[4 * w + 2]);
        }
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const int lo = 64 * w;
            const uint64_t needb = jb >= lo + 63 ? ~0ull : jb < lo ? 0ull : (2ull << (jb - lo)) - 1;
            const uint64_t needz = jz >= lo + 63 ? ~0ull : jz < lo ? 0ull : (2ull << (jz - lo)) - 1;
            okb = okb && (s_mask[4 * w + 1] & needb) == needb;
            okz = okz && (s_mask[4 * w + 3] & needz) == needz;
        }
        const bool ok = (done_b || okb) && (done_z || okz);
        if (ok) {
            const uint64_t vb = (!done_b && (int)t <= jb) ? (wb & kStVal) : 0ull;
            const uint64_t vz = (!done_z && (int)t <= jz) ? (wz & kStVal) : 0ull;
            const uint64_t sb = wave_sum_u64(vb), sz = wave_sum_u64(vz);
            if (lane == 0) {
                s_sum[2 * wave] = sb;
                s_sum[2 * wave + 1] = sz;
            }
        }
        lds_sync();
        if (ok) {
#pragma unroll
            for (int w = 0; w < 4; w++) {
                sum_b += s_sum[2 * w];
                sum_z += s_sum[2 * w + 1];
            }
            done_b = done_b || jb < 256;
            done_z = done_z || jz < 256;
            base -= 256;
        } else {
            __builtin_amdgcn_s_sleep(1);
        }
        lds_sync();  // the masks and sums are read before the next round writes them
        if (done_b && done_z) break;
    }
    pre_b = sum_b;
    pre_z = sum_z;
}

// Codes of chunk k as lane `lane` holds them (16 symbols, two per word);
// symbols past the field read as `fill`.
__device__ __forceinline__ void enc_load_raw(const uint16_t* __restrict__ codes, uint64_t n, uint64_t k, uint32_t lane,
                                             uint32_t fill, uint32_t* raw) {
    const uint64_t base = k * kChunk + lane * kSymPerThread;
    if (base + kSymPerThread <= n) {
        const uint4* p = reinterpret_cast<const uint4*>(codes + base);
        const uint4 u0 = p[0], u1 = p[1];
        raw[0] = u0.x, raw[1] = u0.y, raw[2] = u0.z, raw[3] = u0.w;
        raw[4] = u1.x, raw[5] = u1.y, raw[6] = u1.z, raw[7] = u1.w;
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t a = base + 2 * j < n ? (uint32_t)codes[base + 2 * j] : fill;
            const uint32_t b = base + 2 * j + 1 < n ? (uint32_t)codes[base + 2 * j + 1] : fill;
            raw[j] = a | (b << 16);
        }
    }
}

template <int BOOK, int WC>
__global__ void __launch_bounds__(256) k_encode(const uint16_t* __restrict__ codes, uint64_t n, uint64_t nchunks,
                                                uint64_t nparts, const uint32_t* __restrict__ cw_hi,
                                                const uint32_t* __restrict__ cw_lo,
                                                const uint8_t* __restrict__ len_tab, int dict,
                                                const float* __restrict__ sparse, uint8_t* __restrict__ archive,
                                                PackTarget T, uint64_t* __restrict__ st_bits,
                                                uint64_t* __restrict__ st_zeros, uint32_t* __restrict__ part_ctr,
                                                uint32_t stage_words, uint64_t* __restrict__ index, IdxLayout L) {
    static_assert(BOOK != kBookGlobal, "the fused encode keeps the codebook in LDS");
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wb[4], s_wz[4];
    __shared__ uint64_t s_lb[24];  // look-back masks and sums
    uint32_t* tab = lds;
    uint8_t* tlen = reinterpret_cast<uint8_t*>(lds + (uint32_t)dict + 1);
    if (BOOK == kBookPacked) {
        for (int i = threadIdx.x; i <= dict; i += blockDim.x) {
            const uint32_t l = i < dict ? len_tab[i] : 0u;
            tab[i] = l == 0 ? 0u : l <= 26 ? (cw_lo[i] << 6) | l : kLongCode;
        }
    } else {
        for (int i = threadIdx.x; i <= dict; i += blockDim.x) {
            tab[i] = i < dict ? cw_lo[i] : 0u;
            tlen[i] = i < dict ? len_tab[i] : 0u;
        }
    }
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t* stage = lds + book_words(BOOK, dict) + wave * stage_words;
    const uint32_t lane = lane_id();
    const uint32_t fill = (uint32_t)dict;  // symbols past the field: an LDS slot holding 0
    for (;;) {
        if (threadIdx.x == 0) s_part = atomicAdd(part_ctr, 1u);
        lds_sync();  // (also orders the codebook fill before the first lookup)
        const uint64_t part = s_part;
        if (part >= nparts) break;
        const uint64_t kw = part * (4 * WC) + (uint64_t)wave * WC;

        // pass 1: bits and outliers of the wavefront's chunks, whose codes stay
        // in registers for pass 2 (a partition's codes do not survive the
        // look-back in the L2, which other workgroups' streams flush)
        uint32_t ctot[WC];  // chunk bits (low 17) and outliers (high 15), uniform
        uint32_t keep[WC][8];
        uint32_t wb = 0, wz = 0;
#pragma unroll
        for (int i = 0; i < WC; i++) enc_load_raw(codes, n, kw + i, lane, fill, keep[i]);
#pragma unroll
        for (int i = 0; i < WC; i++) {
            const uint32_t* raw = keep[i];
            uint32_t tbits = 0, tz = 0;
#pragma unroll
            for (int j = 0; j < kSymPerThread; j++) {
                const uint32_t sj = (raw[j >> 1] >> (16 * (j & 1))) & 0xFFFF;
                uint32_t l;
                if constexpr (BOOK == kBookSplit) {
                    l = tlen[sj];
                } else {
                    l = tab[sj] & 63;
                    if (l == kLongCode) l = len_tab[sj];
                }
                tbits += l;
                tz += sj == 0;
            }
            ctot[i] = lane63(wave_scan_dpp(tbits | (tz << 17)));
            wb += ctot[i] & 0x1FFFF;
            wz += ctot[i] >> 17;
        }
        if (lane == 0) {
            s_wb[wave] = wb;
            s_wz[wave] = wz;
        }
        lds_sync();

        // the partition's prefix: publish the aggregate, look back (the whole
        // workgroup), publish the inclusive prefix
        uint64_t pb = 0, pz = 0;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            pb += s_wb[w];
            pz += s_wz[w];
        }
        if (threadIdx.x == 0) {
            st_status(st_bits + part, (part == 0 ? kStIncl : kStAgg) | pb);
            st_status(st_zeros + part, (part == 0 ? kStIncl : kStAgg) | pz);
        }
        uint64_t eb = 0, ez = 0;
        if (part > 0) {
            lookback_wg(st_bits, st_zeros, part, s_lb, eb, ez);
            if (threadIdx.x == 0) {
                st_status(st_bits + part, kStIncl | (eb + pb));
                st_status(st_zeros + part, kStIncl | (ez + pz));
            }
        }

        // pass 2: pack each chunk from its first bit and outlier rank
        uint64_t cb = eb, cz = ez;
        for (uint32_t w = 0; w < wave; w++) {
            cb += s_wb[w];
            cz += s_wz[w];
        }
        // the index group's first bit and outlier rank (a group is
        // kChunksPerGroup chunks: this wavefront's, or it and the one before)
        uint64_t gbits = cb, gouts = cz;
        if constexpr (WC < (int)kChunksPerGroup) {
            static_assert((int)kChunksPerGroup == 2 * WC, "two wavefronts per index group");
            if (wave & 1) {
                gbits -= s_wb[wave - 1];
                gouts -= s_wz[wave - 1];
            }
        }
#pragma unroll 1
        for (int i = 0; i < WC; i++) {
            const uint64_t k = kw + i;
            if (k >= nchunks) break;
            ChunkRaw cur;
#pragma unroll
            for (int j = 0; j < 8; j++) cur.w[j] = keep[0][j];
            const uint64_t base = k * kChunk + lane * kSymPerThread;
            cur.nvalid = base >= n ? 0u : (uint32_t)min<uint64_t>(kSymPerThread, n - base);
            const uint64_t nx = (k + 1) * kChunk + lane;
            cur.next = (lane < 32 && k + 1 < nchunks && nx < n) ? (uint32_t)codes[nx] : 0xFFFFFFFFu;
            cur.bit_off = cb;
            cur.out_off = cz;
            cur.nz = ctot[0] >> 17;
            pack_chunk<BOOK>(cur, k, nchunks, lane, tab, tlen, cw_hi, cw_lo, len_tab, T, sparse, archive, stage,
                             index, L, gbits, gouts);
            cb += ctot[0] & 0x1FFFF;
            cz += ctot[0] >> 17;
            // the next chunk's totals and codes move to the front (registers,
            // not an indexed array)
#pragma unroll
            for (int q = 0; q + 1 < WC; q++) {
                ctot[q] = ctot[q + 1];
#pragma unroll
                for (int j = 0; j < 8; j++) keep[q][j] = keep[q + 1][j];
            }
        }
    }
}

// Sizes of the fused encode (the last partition's inclusive prefix is the
// stream's total bits and outliers): the summary and the two length fields of
// the payload header, then the outliers from the compact scratch into the
// archive's outlier section, which follows the stream.
__global__ void __launch_bounds__(256) k_enc_finish(const uint64_t* __restrict__ st_bits,
                                                    const uint64_t* __restrict__ st_zeros, uint64_t nparts,
                                                    const uint32_t* __restrict__ outl, EncSummary* __restrict__ sum,
                                                    uint8_t* __restrict__ archive, uint64_t meta_bytes,
                                                    uint64_t pad_bytes, uint64_t tree_bytes, uint64_t out_cap,
                                                    uint32_t* __restrict__ err, uint64_t* __restrict__ index) {
    const uint64_t tb = nparts ? st_bits[nparts - 1] & kStVal : 0;
    const uint64_t tz = nparts ? st_zeros[nparts - 1] & kStVal : 0;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        enc_summary(tb, tz, sum, archive, meta_bytes, pad_bytes, tree_bytes, out_cap, err, index);
    const uint64_t off = meta_bytes + pad_bytes + 8 + tree_bytes + (tb + 7) / 8;
    if (off + 4 * tz > out_cap) return;  // enc_summary reports it
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < tz; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = outl[r];
        uint8_t* d = archive + off + 4 * r;
        d[0] = (uint8_t)v;
        d[1] = (uint8_t)(v >> 8);
        d[2] = (uint8_t)(v >> 16);
        d[3] = (uint8_t)(v >> 24);
    }
}

// ------------------------------------------------------------------ batched fields

// Min/max words of every field whose range is found on the device.
__global__ void k_range_init_multi(const BatchField* __restrict__ fd, int nf) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nf && fd[f].range) {
        fd[f].range[0] = 0xFFFFFFFFu;
        fd[f].range[1] = 0u;
    }
}

// One workgroup per unit of kRangeUnit floats of one field (same reduction
// and NaN rule as k_range).
constexpr uint64_t kRangeUnit = 65536;

__global__ void __launch_bounds__(256) k_range_multi(const BatchField* __restrict__ fd,
                                                     const uint64_t* __restrict__ start, int nf, uint64_t nunits) {
    __shared__ float smn[4], smx[4];
    for (uint64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int f = field_of(start, nf, u);
        const float* x = fd[f].in;
        const uint64_t n = fd[f].a.n;
        const uint64_t b = (u - start[f]) * kRangeUnit, e = min(n, b + kRangeUnit);
        float mn = __builtin_inff(), mx = -__builtin_inff();
        uint64_t tail = b;
        if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
            const float4* x4 = reinterpret_cast<const float4*>(x);
            const uint64_t i0 = b / 4, i1 = e / 4;
            uint64_t i = i0 + threadIdx.x;
            for (; i + 768 < i1; i += 1024) {
                float4 v[4];
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = x4[i + 256 * k];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    mn = fminf(mn, fminf(fminf(v[k].x, v[k].y), fminf(v[k].z, v[k].w)));
                    mx = fmaxf(mx, fmaxf(fmaxf(v[k].x, v[k].y), fmaxf(v[k].z, v[k].w)));
                }
            }
            for (; i < i1; i += 256) {
                const float4 v = x4[i];
                mn = fminf(mn, fminf(fminf(v.x, v.y), fminf(v.z, v.w)));
                mx = fmaxf(mx, fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
            }
            tail = i1 * 4;
        }
        for (uint64_t i = tail + threadIdx.x; i < e; i += 256) {
            const float v = x[i];
            mn = fminf(mn, v);
            mx = fmaxf(mx, v);
        }
        for (int d = 32; d > 0; d >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, d));
            mx = fmaxf(mx, __shfl_xor(mx, d));
        }
        const int w = threadIdx.x >> 6;
        if (lane_id() == 0) {
            smn[w] = mn;
            smx[w] = mx;
        }
        lds_sync();
        if (threadIdx.x == 0) {
            for (int i = 1; i < 4; i++) {
                mn = fminf(mn, smn[i]);
                mx = fmaxf(mx, smx[i]);
            }
            if (mn == mn) atomicMin(&fd[f].range[0], f2ord(mn));
            if (mx == mx) atomicMax(&fd[f].range[1], f2ord(mx));
        }
        lds_sync();
    }
}

__global__ void k_ebfam_multi(const BatchField* __restrict__ fd, int nf) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nf && fd[f].range) ebfam(fd[f].range, fd[f].in, fd[f].eb, fd[f].ebd);
}

// No histogram: the batched dual quantization counts nothing (the codes are
// counted by k_hist_codes_multi).
struct NoHist {
    __device__ __forceinline__ void add(int) {}
};

// Dual quantization of every field of a batch whose tiles take form KIND
// (one launch per form, so each keeps its own register budget).  The work
// units (four wave tiles of one field each) are taken grid-stride, so the
// waves in flight work on neighbouring tiles of the same planes as in the
// single-field kernels; fields of other forms have no units.  The tile bodies
// are the single-field kernels'.
template <int B, int KIND>
__global__ void __launch_bounds__(256) k_dualquant_multi(const BatchField* __restrict__ fd,
                                                         const uint64_t* __restrict__ start, int nf,
                                                         uint64_t nunits) {
    NoHist H;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    DQArgs a;
    const float* in = nullptr;
    uint16_t* codes = nullptr;
    float* sparse = nullptr;
    uint64_t ntiles = 0, ub = 0, ue = 0;
    bool aligned = false;
    for (uint64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
        if (u >= ue) {  // next field (uniform)
            const int f = field_of(start, nf, u);
            a = fd[f].a;
            dq_resolve(a);
            in = fd[f].in;
            codes = fd[f].codes;
            sparse = fd[f].sparse;
            ntiles = fd[f].dq_tiles;
            ub = start[f];
            ue = start[f + 1];
            aligned = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
        }
        const uint64_t t = (u - ub) * 4 + wave;
        if (t >= ntiles) continue;
        if constexpr (KIND == kDq3x4 || KIND == kDq3x4u) {
            if constexpr (B <= 16) dq3_x4_at<B, KIND == kDq3x4>(in, codes, sparse, H, a, t, lane);
        } else if constexpr (KIND == kDq3) {
            dq3_tile<B, false>(in, codes, sparse, H, a, t, lane);
        } else if constexpr (KIND == kDq2x4 || KIND == kDq2x4u) {
            dq2_x4_tile<B, KIND == kDq2x4>(in, codes, sparse, H, a, t, lane);
        } else if constexpr (KIND == kDq2) {
            dq2_tile<B, false>(in, codes, sparse, H, a, t, lane);
        } else {
            dq1_tile<B, false>(in, codes, sparse, H, a, t * 64, lane, aligned);
        }
    }
}

// Histogram of every field's codes.  A unit is kHistUnit codes of one field;
// workgroup g takes the contiguous units [U g / G, U (g + 1) / G) (streaming
// reads, one field at a time) and writes its LDS histogram of each field it
// visits as one dense row of `partials`, rows row0[g], row0[g] + 1, ... in
// field order, so a field's rows are contiguous (k_hist_reduce_multi sums
// them).
constexpr uint64_t kHistUnit = 16384;

__global__ void __launch_bounds__(256) k_hist_codes_multi(const BatchField* __restrict__ fd,
                                                          const uint64_t* __restrict__ start, int nf, uint64_t nunits,
                                                          int dict, int radius, uint32_t* __restrict__ partials,
                                                          const uint32_t* __restrict__ row0) {
    extern __shared__ uint32_t lds[];
    LdsHist H;
    uint64_t u = nunits * blockIdx.x / gridDim.x;
    const uint64_t u1 = nunits * (blockIdx.x + 1) / gridDim.x;
    if (u >= u1) return;
    uint32_t row = row0[blockIdx.x];
    for (int f = field_of(start, nf, u); u < u1; f++) {
        const uint64_t ue = min(u1, start[f + 1]);
        if (u >= ue) continue;
        lds_sync();  // the previous field's row has been read out of LDS
        hist_init(H, lds, dict, radius, nullptr);
        const uint16_t* codes = fd[f].codes;  // 128-byte aligned
        const uint64_t n = fd[f].a.n;
        const uint64_t b = (u - start[f]) * kHistUnit, e = min(n, (ue - start[f]) * kHistUnit);
        const uint64_t e8 = e / 8 * 8;  // whole groups of 8 codes
        const uint4* c8 = reinterpret_cast<const uint4*>(codes);
        for (uint64_t i = b / 8 + threadIdx.x; i < e8 / 8; i += 256) {
            const uint4 v = c8[i];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                H.add((int)(w[j] & 0xFFFF));
                H.add((int)(w[j] >> 16));
            }
        }
        for (uint64_t i = max(b, e8) + threadIdx.x; i < e; i += 256) H.add((int)codes[i]);
        u = ue;
        hist_flush_row(H, partials + (uint64_t)row * dict);
        row++;
    }
}

// Sums each field's partial rows: blockIdx.x picks 256 bins, blockIdx.y a
// group {field, first row, rows (<= kHistRows)}; one 64-bit atomic per bin
// and group.
__global__ void __launch_bounds__(256) k_hist_reduce_multi(const uint32_t* __restrict__ partials, int dict,
                                                           const uint32_t* __restrict__ groups,
                                                           unsigned long long* __restrict__ hist) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= dict) return;
    const uint32_t f = groups[3 * blockIdx.y], r0 = groups[3 * blockIdx.y + 1], nr = groups[3 * blockIdx.y + 2];
    uint32_t s = 0;
#pragma unroll 8
    for (uint32_t r = 0; r < nr; r++) s += partials[(uint64_t)(r0 + r) * dict + b];
    if (s) atomicAdd(&hist[(uint64_t)f * dict + b], (unsigned long long)s);
}

// Copies each field's staged header (metadata, Huffman header, tree) to the
// front of its archive; one workgroup per field.
__global__ void k_place_heads(const BatchField* __restrict__ fd, int nf, const uint32_t* __restrict__ errs) {
    const int f = blockIdx.x;
    if (f >= nf || errs[f]) return;
    const uint8_t* src = fd[f].head;
    uint8_t* dst = fd[f].archive;
    for (uint64_t i = threadIdx.x; i < fd[f].head_bytes; i += blockDim.x) dst[i] = src[i];
}

// Bits and outliers of every chunk of every field (as k_chunk_bits); a unit
// is four chunks of one field, whose length table the workgroup holds in LDS.
__global__ void __launch_bounds__(256) k_chunk_bits_multi(const BatchField* __restrict__ fd,
                                                          const uint64_t* __restrict__ start, int nf, uint64_t nunits,
                                                          int dict, uint32_t* __restrict__ chunk_bits,
                                                          uint32_t* __restrict__ chunk_zeros) {
    __shared__ uint8_t slen[kMaxLdsBook];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    uint64_t u = nunits * blockIdx.x / gridDim.x;
    const uint64_t u1 = nunits * (blockIdx.x + 1) / gridDim.x;
    if (u >= u1) return;
    for (int f = field_of(start, nf, u); u < u1; f++) {
        const uint64_t ue = min(u1, start[f + 1]);
        if (u >= ue) continue;
        const uint8_t* len_tab = reinterpret_cast<const uint8_t*>(fd[f].book + 2 * dict);
        lds_sync();
        for (int i = threadIdx.x; i < dict; i += blockDim.x) slen[i] = len_tab[i];
        lds_sync();
        const uint16_t* codes = fd[f].codes;
        const uint64_t n = fd[f].a.n, nch = fd[f].nchunks, c0 = fd[f].chunk0, ub = start[f];
        // the next unit's codes are in flight while this one is summed
        uint32_t nxt[kSymPerThread];
        uint64_t k = (u - ub) * 4 + wave;
        if (k < nch) load_syms(codes, n, k * kChunk + lane * kSymPerThread, nxt);
        for (; u < ue; u++, k += 4) {
            uint32_t sym[kSymPerThread];
#pragma unroll
            for (int j = 0; j < kSymPerThread; j++) sym[j] = nxt[j];
            if (u + 1 < ue && k + 4 < nch) load_syms(codes, n, (k + 4) * kChunk + lane * kSymPerThread, nxt);
            if (k >= nch) continue;
            uint32_t bits = 0, zeros = 0;
#pragma unroll
            for (int j = 0; j < kSymPerThread; j++) {
                const uint32_t sj = sym[j];
                if (sj != 0xFFFFFFFFu) {
                    bits += slen[sj];
                    zeros += sj == 0;
                }
            }
            const uint32_t tot = lane63(wave_scan_dpp(bits | (zeros << 17)));
            if (lane == 0) {
                chunk_bits[c0 + k] = tot & 0x1FFFF;
                chunk_zeros[c0 + k] = tot >> 17;
            }
        }
    }
}

// Per-field sizes from the batch-wide exclusive scan of chunk bits and
// outliers (a field's totals are the differences of the scan at its first
// chunk and at the next field's); one thread per field.
__global__ void k_enc_summary_multi(const BatchField* __restrict__ fd, int nf, const uint64_t* __restrict__ bit_off,
                                    const uint64_t* __restrict__ out_off, const uint64_t* __restrict__ totals,
                                    uint64_t ctot, EncSummary* __restrict__ sums, uint32_t* __restrict__ errs) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf || errs[f]) return;
    const uint64_t c0 = fd[f].chunk0, c1 = c0 + fd[f].nchunks;
    const uint64_t tb = (c1 < ctot ? bit_off[c1] : totals[0]) - bit_off[c0];
    const uint64_t tz = (c1 < ctot ? out_off[c1] : totals[1]) - out_off[c0];
    enc_summary(tb, tz, sums + f, fd[f].archive, fd[f].meta_bytes, 0, fd[f].tree_bytes, fd[f].out_cap, errs + f);
}

// Bit packing of every field (as k_pack, without the load pipeline); a unit
// is four chunks of one field, whose codebook the workgroup holds in LDS.
template <int BOOK>
__global__ void __launch_bounds__(256) k_pack_multi(const BatchField* __restrict__ fd,
                                                    const uint64_t* __restrict__ start, int nf, uint64_t nunits,
                                                    int dict, const uint64_t* __restrict__ bit_off,
                                                    const uint64_t* __restrict__ out_off,
                                                    const EncSummary* __restrict__ sums,
                                                    const uint32_t* __restrict__ errs, uint32_t stage_words) {
    static_assert(BOOK != kBookGlobal, "batched pack keeps the codebook in LDS");
    extern __shared__ uint32_t lds[];
    uint32_t* tab = lds;
    uint8_t* tlen = reinterpret_cast<uint8_t*>(lds + (uint32_t)dict + 1);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t* stage = lds + book_words(BOOK, dict) + wave * stage_words;
    const uint32_t lane = lane_id();
    const uint32_t fill = (uint32_t)dict;
    uint64_t u = nunits * blockIdx.x / gridDim.x;
    const uint64_t u1 = nunits * (blockIdx.x + 1) / gridDim.x;
    if (u >= u1) return;
    for (int f = field_of(start, nf, u); u < u1; f++) {
        const uint64_t ue = min(u1, start[f + 1]);
        if (u >= ue) continue;
        const uint32_t* cw_hi = fd[f].book;
        const uint32_t* cw_lo = cw_hi + dict;
        const uint8_t* len_tab = reinterpret_cast<const uint8_t*>(cw_hi + 2 * dict);
        lds_sync();
        if (BOOK == kBookPacked) {
            for (int i = threadIdx.x; i <= dict; i += blockDim.x) {
                const uint32_t l = i < dict ? len_tab[i] : 0u;
                tab[i] = l == 0 ? 0u : l <= 26 ? (cw_lo[i] << 6) | l : kLongCode;
            }
        } else {
            for (int i = threadIdx.x; i <= dict; i += blockDim.x) {
                tab[i] = i < dict ? cw_lo[i] : 0u;
                tlen[i] = i < dict ? len_tab[i] : 0u;
            }
        }
        lds_sync();
        if (!errs[f]) {
            const EncSummary S = sums[f];
            const uint16_t* codes = fd[f].codes;
            const float* sparse = fd[f].sparse;
            uint8_t* archive = fd[f].archive;
            const uint64_t n = fd[f].a.n, nch = fd[f].nchunks, c0 = fd[f].chunk0, ub = start[f];
            const uint64_t bb = bit_off[c0], zb = out_off[c0];
            // software pipeline: the next unit's chunk is in flight while this one is packed
            ChunkRaw nxt;
            uint64_t k = (u - ub) * 4 + wave;
            if (k < nch) load_chunk(codes, n, nch, bit_off + c0, out_off + c0, zb + S.n_outlier, k, lane, fill, nxt);
            for (uint64_t v = u; v < ue; v++, k += 4) {
                const ChunkRaw cur = nxt;
                if (v + 1 < ue && k + 4 < nch)
                    load_chunk(codes, n, nch, bit_off + c0, out_off + c0, zb + S.n_outlier, k + 4, lane, fill, nxt);
                if (k >= nch) continue;
                ChunkRaw c = cur;
                c.bit_off -= bb;
                c.out_off -= zb;
                pack_chunk<BOOK>(c, k, nch, lane, tab, tlen, cw_hi, cw_lo, len_tab, pack_target(S, archive), sparse,
                                 archive, stage);
            }
        }
        u = ue;
    }
}

// ------------------------------------------------------------------ launchers

int dq_lds_bytes(int dict) {
    return (kHotBins * 64 + (dict <= kMaxLdsHistBins ? dict : 0)) * 4;
}

// Clears n words with stores from a kernel (not hipMemsetAsync, so that a
// caller's graph capture records only kernel nodes).
__global__ void __launch_bounds__(256) k_zero_u32(uint32_t* __restrict__ p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

void launch_zero_u32(uint32_t* p, uint64_t n, hipStream_t st) {
    const unsigned g = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 1024));
    hipLaunchKernelGGL(k_zero_u32, dim3(g), dim3(256), 0, st, p, n);
}

uint32_t range_parts(uint64_t n) {
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>((n / 4 + 255) / 256, 1), kRangeParts);
}

void launch_range(const float* x, uint64_t n, uint32_t* out, hipStream_t st) {
    hipLaunchKernelGGL(k_range, dim3(range_parts(n)), dim3(256), 0, st, x, n, out);
}

void launch_ebfam(uint32_t* mm, uint64_t n, const float* x, double eb, EbDev* out, hipStream_t st) {
    hipLaunchKernelGGL(k_ebfam, dim3(1), dim3(256), 0, st, mm, range_parts(n), x, eb, out);
}

bool dualquant_x4_aligned(const float* in, uint64_t d0) {
    return (d0 % 4) == 0 && (reinterpret_cast<uintptr_t>(in) & 15) == 0;
}

bool dualquant_use_x4(int ndim, int B, const float* in, uint64_t d0, uint64_t d1, uint64_t d2, bool batched) {
    if (ndim == 1 || d0 < 128) return false;
    if (ndim == 3 && B > 16) return false;
    if (batched || dualquant_x4_aligned(in, d0)) return true;
    // element loads: only when the 4-wide tiles alone keep the chip busy
    const uint64_t tiles = (d0 + 255) / 256 * ((d1 + B - 1) / B) * (ndim == 3 ? (d2 + B - 1) / B : 1);
    return tiles >= 2048;
}


void launch_dualquant(int ndim, int B, const float* in, uint16_t* codes, float* sparse, unsigned long long* hist,
                      uint32_t* partials, const DQArgs& a0, int grid, hipStream_t st) {
    const int lds = dq_lds_bytes(a0.dict);
    DQArgs a = a0;
    const bool pad = a.halo != nullptr;
    const bool x4 = !pad && dualquant_use_x4(ndim, B, in, a.d0, a.d1, a.d2, false);
    const bool aln = dualquant_x4_aligned(in, a.d0);
    const bool cold = a.dict <= kMaxLdsHistBins;  // every histogram bin in LDS
    if (x4) a.ntiles_x = (a.d0 + 255) / 256;
#define DQ_CASE(ND, KERN, BB)                                                                              \
    if (ndim == ND && B == BB) {                                                                           \
        hipLaunchKernelGGL(KERN<BB>, dim3(grid), dim3(256), lds, st, in, codes, sparse, hist, partials, a); \
        goto launched;                                                                                     \
    }
#define DQ_X4(ND, KERN, BB)                                                                                    \
    if (ndim == ND && B == BB) {                                                                               \
        if (aln && cold)                                                                                       \
            hipLaunchKernelGGL((KERN<BB, true, true>), dim3(grid), dim3(256), lds, st, in, codes, sparse, hist,  \
                               partials, a);                                                                   \
        else if (aln)                                                                                          \
            hipLaunchKernelGGL((KERN<BB, true, false>), dim3(grid), dim3(256), lds, st, in, codes, sparse, hist, \
                               partials, a);                                                                   \
        else if (cold)                                                                                         \
            hipLaunchKernelGGL((KERN<BB, false, true>), dim3(grid), dim3(256), lds, st, in, codes, sparse, hist, \
                               partials, a);                                                                   \
        else                                                                                                   \
            hipLaunchKernelGGL((KERN<BB, false, false>), dim3(grid), dim3(256), lds, st, in, codes, sparse,      \
                               hist, partials, a);                                                             \
        goto launched;                                                                                         \
    }
    if (x4) {
        DQ_X4(3, k_dualquant_3d_x4, 8)
        DQ_X4(3, k_dualquant_3d_x4, 16)
        DQ_X4(2, k_dualquant_2d_x4, 8)
        DQ_X4(2, k_dualquant_2d_x4, 16)
        DQ_X4(2, k_dualquant_2d_x4, 32)
        DQ_X4(2, k_dualquant_2d_x4, 64)
    }
#undef DQ_X4
    if (pad) {
#define DQ_PAD(ND, KERN, BB)                                                                                       \
    if (ndim == ND && B == BB) {                                                                                   \
        hipLaunchKernelGGL((KERN<BB, true>), dim3(grid), dim3(256), lds, st, in, codes, sparse, hist, partials, a); \
        goto launched;                                                                                             \
    }
        DQ_PAD(3, k_dualquant_3d, 8)
        DQ_PAD(3, k_dualquant_3d, 16)
        DQ_PAD(3, k_dualquant_3d, 32)
        DQ_PAD(3, k_dualquant_3d, 64)
        DQ_PAD(2, k_dualquant_2d, 8)
        DQ_PAD(2, k_dualquant_2d, 16)
        DQ_PAD(2, k_dualquant_2d, 32)
        DQ_PAD(2, k_dualquant_2d, 64)
        DQ_PAD(1, k_dualquant_1d, 8)
        DQ_PAD(1, k_dualquant_1d, 16)
        DQ_PAD(1, k_dualquant_1d, 32)
        DQ_PAD(1, k_dualquant_1d, 64)
#undef DQ_PAD
        return;
    }
    DQ_CASE(3, k_dualquant_3d, 8)
    DQ_CASE(3, k_dualquant_3d, 16)
    DQ_CASE(3, k_dualquant_3d, 32)
    DQ_CASE(3, k_dualquant_3d, 64)
    DQ_CASE(2, k_dualquant_2d, 8)
    DQ_CASE(2, k_dualquant_2d, 16)
    DQ_CASE(2, k_dualquant_2d, 32)
    DQ_CASE(2, k_dualquant_2d, 64)
    DQ_CASE(1, k_dualquant_1d, 8)
    DQ_CASE(1, k_dualquant_1d, 16)
    DQ_CASE(1, k_dualquant_1d, 32)
    DQ_CASE(1, k_dualquant_1d, 64)
#undef DQ_CASE
    return;
launched:
    if (a.dict <= kMaxLdsHistBins)
        hipLaunchKernelGGL(k_hist_reduce, dim3((a.dict + 255) / 256, (grid + kHistRows - 1) / kHistRows), dim3(256), 0,
                           st, partials, grid, a.dict, hist);
}

void launch_chunk_bits(const uint16_t* codes, uint64_t n, const uint8_t* len_tab, int dict, uint32_t* cb, uint32_t* cz,
                       uint64_t nchunks, int grid_cap, hipStream_t st) {
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((nchunks + 3) / 4, grid_cap));
    if (dict <= kMaxLdsLenTab)
        hipLaunchKernelGGL(k_chunk_bits<true>, dim3(grid), dim3(256), 0, st, codes, n, nchunks, len_tab, dict, cb, cz);
    else
        hipLaunchKernelGGL(k_chunk_bits<false>, dim3(grid), dim3(256), 0, st, codes, n, nchunks, len_tab, dict, cb, cz);
}

void launch_enc_summary(const uint64_t* totals, EncSummary* sum, uint8_t* archive, uint64_t meta_bytes,
                        uint64_t pad_bytes, uint64_t tree_bytes, uint64_t out_cap, uint32_t* err, uint64_t* index,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_enc_summary, dim3(1), dim3(64), 0, st, totals, sum, archive, meta_bytes, pad_bytes,
                       tree_bytes, out_cap, err, index);
}

void launch_enc_summary_dev(const uint64_t* totals, EncSummary* sum, uint8_t* archive, uint64_t meta_bytes,
                            uint64_t pad_bytes, const CodebookInfo* info, uint64_t out_cap, uint32_t* err,
                            uint64_t* index, hipStream_t st) {
    hipLaunchKernelGGL(k_enc_summary_dev, dim3(1), dim3(64), 0, st, totals, sum, archive, meta_bytes, pad_bytes, info,
                       out_cap, err, index);
}

void launch_pack(const uint16_t* codes, uint64_t n, const uint32_t* cw_hi, const uint32_t* cw_lo, const uint8_t* len_tab,
                 int dict, const uint64_t* bit_off, const uint64_t* out_off, uint64_t nchunks, const EncSummary* sum,
                 const float* sparse, uint8_t* archive, const uint32_t* err, int max_len, int grid_cap,
                 uint64_t* index, hipStream_t st) {
    const IdxLayout L = idx_layout(index ? n : 0);
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((nchunks + 3) / 4, grid_cap));
    // words one chunk can touch: 1024 codewords of at most max_len bits, plus
    // the partial words at either end
    const uint32_t stage_words = 32u * (uint32_t)std::max(max_len, 1) + 2;
    const int book = dict > kMaxLdsBook ? kBookGlobal : max_len <= 32 ? kBookSplit : kBookPacked;
    const size_t lds = ((size_t)book_words(book, dict) + 4 * (size_t)stage_words) * 4;
#define PACK_ARGS \
    codes, n, cw_hi, cw_lo, len_tab, dict, bit_off, out_off, nchunks, sum, sparse, archive, err, stage_words, index, L
    if (book == kBookSplit)
        hipLaunchKernelGGL(k_pack<kBookSplit>, dim3(grid), dim3(256), lds, st, PACK_ARGS);
    else if (book == kBookPacked)
        hipLaunchKernelGGL(k_pack<kBookPacked>, dim3(grid), dim3(256), lds, st, PACK_ARGS);
    else
        hipLaunchKernelGGL(k_pack<kBookGlobal>, dim3(grid), dim3(256), lds, st, PACK_ARGS);
#undef PACK_ARGS
}

uint64_t encode_parts(uint64_t nchunks) { return (nchunks + kEncPartChunks - 1) / kEncPartChunks; }

bool encode_fused_ok(int dict) { return dict <= kMaxLdsBook; }

void launch_encode(const uint16_t* codes, uint64_t n, const uint32_t* cw_hi, const uint32_t* cw_lo,
                   const uint8_t* len_tab, int dict, int max_len, const float* sparse, uint8_t* archive,
                   uint64_t stream_bit_off, uint64_t out_cap, bool no_stream, uint32_t* outl, uint64_t* status,
                   uint32_t* err, uint64_t* index, int grid_cap, hipStream_t st) {
    const uint64_t nchunks = (n + kChunk - 1) / kChunk;
    const uint64_t nparts = encode_parts(nchunks);
    const IdxLayout L = idx_layout(index ? n : 0);
    PackTarget T;
    T.outl = reinterpret_cast<uint8_t*>(outl);
    T.stream_bit_off = stream_bit_off;
    T.stream_end_byte = ~0ull;
    T.cap = out_cap;
    T.err = err;
    T.no_stream = no_stream;
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(nparts, (uint64_t)grid_cap));
    const uint32_t stage_words = 32u * (uint32_t)std::max(max_len, 1) + 2;
    const int book = max_len <= 32 ? kBookSplit : kBookPacked;
    const size_t lds = ((size_t)book_words(book, dict) + 4 * (size_t)stage_words) * 4;
    uint64_t* sb = status;
    uint64_t* sz = status + nparts;
    uint32_t* ctr = reinterpret_cast<uint32_t*>(status + 2 * nparts);
#define ENC_ARGS codes, n, nchunks, nparts, cw_hi, cw_lo, len_tab, dict, sparse, archive, T, sb, sz, ctr, stage_words, index, L
    if (book == kBookSplit)
        hipLaunchKernelGGL((k_encode<kBookSplit, kEncWaveChunks>), dim3(grid), dim3(256), lds, st, ENC_ARGS);
    else
        hipLaunchKernelGGL((k_encode<kBookPacked, kEncWaveChunks>), dim3(grid), dim3(256), lds, st, ENC_ARGS);
#undef ENC_ARGS
}

void launch_enc_finish(const uint64_t* status, uint64_t n, const uint32_t* outl, uint64_t n_outlier_max,
                       EncSummary* sum, uint8_t* archive, uint64_t meta_bytes, uint64_t pad_bytes, uint64_t tree_bytes,
                       uint64_t out_cap, uint32_t* err, uint64_t* index, hipStream_t st) {
    const uint64_t nparts = encode_parts((n + kChunk - 1) / kChunk);
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n_outlier_max + 255) / 256, 1024));
    hipLaunchKernelGGL(k_enc_finish, dim3(grid), dim3(256), 0, st, status, status + nparts, nparts, outl, sum, archive,
                       meta_bytes, pad_bytes, tree_bytes, out_cap, err, index);
}

// Batched launches (one per stage for all fields).
void launch_range_multi(const BatchField* fd, int nf, const uint64_t* start, uint64_t nunits, hipStream_t st) {
    hipLaunchKernelGGL(k_range_init_multi, dim3((nf + 63) / 64), dim3(64), 0, st, fd, nf);
    if (nunits)
        hipLaunchKernelGGL(k_range_multi, dim3((unsigned)std::min<uint64_t>(nunits, 2048)), dim3(256), 0, st, fd, start,
                           nf, nunits);
    hipLaunchKernelGGL(k_ebfam_multi, dim3((nf + 63) / 64), dim3(64), 0, st, fd, nf);
}

uint64_t range_units(uint64_t n) { return (n + kRangeUnit - 1) / kRangeUnit; }

void launch_dualquant_multi(int B, int kind, const BatchField* fd, int nf, const uint64_t* start, uint64_t nunits,
                            int grid_cap, hipStream_t st) {
    if (!nunits) return;
    const unsigned grid = (unsigned)std::min<uint64_t>(nunits, grid_cap);
#define DQM(BB, KK)                                                                                              \
    if (B == BB && kind == KK)                                                                                   \
        hipLaunchKernelGGL((k_dualquant_multi<BB, KK>), dim3(grid), dim3(256), 0, st, fd, start, nf, nunits);
#define DQM_B(BB) DQM(BB, kDq1) DQM(BB, kDq2) DQM(BB, kDq2x4) DQM(BB, kDq2x4u) DQM(BB, kDq3)
    DQM_B(8) DQM_B(16) DQM_B(32) DQM_B(64)
    DQM(8, kDq3x4) DQM(16, kDq3x4) DQM(8, kDq3x4u) DQM(16, kDq3x4u)
#undef DQM_B
#undef DQM
}

uint64_t hist_units(uint64_t n) { return (n + kHistUnit - 1) / kHistUnit; }
unsigned hist_multi_grid(uint64_t nunits) { return (unsigned)std::min<uint64_t>(nunits, 1024); }

void launch_hist_multi(const BatchField* fd, int nf, const uint64_t* start, uint64_t nunits, int dict,
                       uint32_t* partials, const uint32_t* row0, const uint32_t* groups, uint32_t ngroups,
                       unsigned long long* hist, hipStream_t st) {
    if (!nunits) return;
    hipLaunchKernelGGL(k_hist_codes_multi, dim3(hist_multi_grid(nunits)), dim3(256), dq_lds_bytes(dict), st, fd, start,
                       nf, nunits, dict, dict / 2, partials, row0);
    hipLaunchKernelGGL(k_hist_reduce_multi, dim3((dict + 255) / 256, ngroups), dim3(256), 0, st, partials, dict, groups,
                       hist);
}

void launch_place_heads(const BatchField* fd, int nf, const uint32_t* errs, hipStream_t st) {
    if (nf) hipLaunchKernelGGL(k_place_heads, dim3(nf), dim3(256), 0, st, fd, nf, errs);
}

void launch_chunk_bits_multi(const BatchField* fd, int nf, const uint64_t* start, uint64_t nunits, int dict,
                             uint32_t* cb, uint32_t* cz, int grid_cap, hipStream_t st) {
    if (!nunits) return;
    const unsigned grid = (unsigned)std::min<uint64_t>(nunits, grid_cap);
    hipLaunchKernelGGL(k_chunk_bits_multi, dim3(grid), dim3(256), 0, st, fd, start, nf, nunits, dict, cb, cz);
}

void launch_enc_summary_multi(const BatchField* fd, int nf, const uint64_t* bit_off, const uint64_t* out_off,
                              const uint64_t* totals, uint64_t ctot, EncSummary* sums, uint32_t* errs,
                              hipStream_t st) {
    hipLaunchKernelGGL(k_enc_summary_multi, dim3((nf + 63) / 64), dim3(64), 0, st, fd, nf, bit_off, out_off, totals,
                       ctot, sums, errs);
}

void launch_pack_multi(const BatchField* fd, int nf, const uint64_t* start, uint64_t nunits, int dict, int max_len,
                       const uint64_t* bit_off, const uint64_t* out_off, const EncSummary* sums, const uint32_t* errs,
                       int grid_cap, hipStream_t st) {
    if (!nunits) return;
    const unsigned grid = (unsigned)std::min<uint64_t>(nunits, grid_cap);
    const uint32_t stage_words = 32u * (uint32_t)std::max(max_len, 1) + 2;
    const int book = max_len <= 32 ? kBookSplit : kBookPacked;
    const size_t lds = ((size_t)book_words(book, dict) + 4 * (size_t)stage_words) * 4;
    if (book == kBookSplit)
        hipLaunchKernelGGL(k_pack_multi<kBookSplit>, dim3(grid), dim3(256), lds, st, fd, start, nf, nunits, dict,
                           bit_off, out_off, sums, errs, stage_words);
    else
        hipLaunchKernelGGL(k_pack_multi<kBookPacked>, dim3(grid), dim3(256), lds, st, fd, start, nf, nunits, dict,
                           bit_off, out_off, sums, errs, stage_words);
}

}  // namespace vsz
