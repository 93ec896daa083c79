// gfx950 kernels of the compress path.
//
//   k_range        value range for r2r (GetDatumValueRange, src/types.cc:27-35)
//   k_dualquant_*  prequant + blocked Lorenzo postquant + outlier capture +
//                  quant-code histogram (c_lorenzo_{1,2,3}d1l, src/dualquant.hh:18-913;
//                  initHuffman's frequency loop, src/huffman.hh:239-242)
//   k_chunk_bits   bits and outliers per 4096-symbol chunk (HuffmanEncode sizes, src/huffman.hh:268-341)
//   k_scan_chunks  exclusive scan of chunk totals + archive sizes (src/vecsz_interface.hh:189-202)
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
__global__ void __launch_bounds__(256) k_range(const float* __restrict__ x, uint64_t n, uint32_t* __restrict__ out) {
    float mn = __builtin_inff(), mx = -__builtin_inff();
    // float4 loads need 16-byte alignment; an unaligned field takes the scalar tail path
    const uint64_t n4 = ((reinterpret_cast<uintptr_t>(x) & 15) == 0) ? n / 4 : 0;
    const float4* x4 = reinterpret_cast<const float4*>(x);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x) {
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
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
            mn = fminf(mn, smn[i]);
            mx = fmaxf(mx, smx[i]);
        }
        if (mn == mn) atomicMin(&out[0], f2ord(mn));
        if (mx == mx) atomicMax(&out[1], f2ord(mx));
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
        int h = bin - lo;
        if ((unsigned)h < (unsigned)kHotBins) {
            atomicAdd(&hot[h * 64 + lane_id()], 1u);
        } else if (lds_cold) {
            atomicAdd(&cold[bin], 1u);
        } else {
            atomicAdd(&global[bin], 1ull);
        }
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
    __syncthreads();
}

__device__ void hist_flush(LdsHist& H) {
    __syncthreads();
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
    __syncthreads();
    if (H.lds_cold) {
        for (int b = threadIdx.x; b < H.dict; b += blockDim.x) {
            uint32_t v = H.cold[b];
            if (v) atomicAdd(&H.global[b], (unsigned long long)v);
        }
    }
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
template <int B>
__global__ void __launch_bounds__(256) k_dualquant_3d(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                                      float* __restrict__ sparse, unsigned long long* __restrict__ hist,
                                                      DQArgs a) {
    extern __shared__ uint32_t lds[];
    LdsHist H;
    hist_init(H, lds, a.dict, a.radius, hist);
    const uint32_t lane = lane_id();
    const uint64_t ntiles = a.ntiles_x * a.ntiles_y * a.ntiles_z;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    const uint64_t sy = a.d0, sz = a.d0 * a.d1;
    for (uint64_t t = wave; t < ntiles; t += nwaves) {
        const uint64_t tx = t % a.ntiles_x;
        const uint64_t ty = (t / a.ntiles_x) % a.ntiles_y;
        const uint64_t tz = t / (a.ntiles_x * a.ntiles_y);
        const uint64_t x = tx * 64 + lane;
        const uint64_t y0 = ty * B, z0 = tz * B;
        const bool xin = x < a.d0;
        const bool scal = x >= a.x8;  // partial x-group: scalar path
        const bool xstart = (x % B) == 0;
        const float thr = scal ? a.thr_scalar : a.thr_vec;
        float prevplane[B];
#pragma unroll
        for (int j = 0; j < B; j++) prevplane[j] = 0.f;
        for (int zz = 0; zz < B; zz++) {
            const uint64_t z = z0 + zz;
            if (z >= a.d2) break;
            float qrow = 0.f, xrow = 0.f;      // q and X of row y-1 (this plane)
            float zrow = 0.f, xzrow = 0.f;     // Z and XZ of row y-1
#pragma unroll
            for (int yy = 0; yy < B; yy++) {
                const uint64_t y = y0 + yy;
                if (y < a.d1) {
                    const uint64_t idx = x + y * sy + z * sz;
                    const bool ok = xin;
                    float v = ok ? __builtin_nontemporal_load(&in[idx]) : 0.f;
                    float q = scal ? prequant_scalar(v, a.r_d) : prequant_vec(v, a.r_f);
                    float X = wave_shr1(q);
                    float Z = zz > 0 ? prevplane[yy] : 0.f;
                    float XZ = wave_shr1(Z);
                    if (xstart) {
                        X = 0.f;
                        XZ = 0.f;
                    }
                    float Y = yy > 0 ? qrow : 0.f;
                    float XY = yy > 0 ? xrow : 0.f;
                    float YZ = yy > 0 ? zrow : 0.f;
                    float XYZ = yy > 0 ? xzrow : 0.f;
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
    hist_flush(H);
}

// ------------------------------------------------------------------ 2D

// Tile of 64 (x) x B (y) per wavefront.  Elements with x >= 8*floor(DIM0/8)
// are never coded by the reference (the `continue` at src/dualquant.hh:494-495):
// code 0, outlier +0.0.
template <int B>
__global__ void __launch_bounds__(256) k_dualquant_2d(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                                      float* __restrict__ sparse, unsigned long long* __restrict__ hist,
                                                      DQArgs a) {
    extern __shared__ uint32_t lds[];
    LdsHist H;
    hist_init(H, lds, a.dict, a.radius, hist);
    const uint32_t lane = lane_id();
    const uint64_t ntiles = a.ntiles_x * a.ntiles_y;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    for (uint64_t t = wave; t < ntiles; t += nwaves) {
        const uint64_t tx = t % a.ntiles_x;
        const uint64_t ty = t / a.ntiles_x;
        const uint64_t x = tx * 64 + lane;
        const uint64_t y0 = ty * B;
        const bool xin = x < a.d0;
        const bool lost = x >= a.x8;
        const bool xstart = (x % B) == 0;
        float qrow = 0.f, xrow = 0.f;
#pragma unroll
        for (int yy = 0; yy < B; yy++) {
            const uint64_t y = y0 + yy;
            if (y < a.d1) {
                const uint64_t idx = x + y * a.d0;
                float v = xin ? __builtin_nontemporal_load(&in[idx]) : 0.f;
                float q = prequant_vec(v, a.r_f);
                float W = wave_shr1(q);
                if (xstart) W = 0.f;
                float N = yy > 0 ? qrow : 0.f;
                float NW = yy > 0 ? xrow : 0.f;
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
    hist_flush(H);
}

// ------------------------------------------------------------------ 1D

// One thread per run of 8 consecutive elements (the AVX group).  The element
// before the run comes from lane-1 (DPP) or, for lane 0, from memory.  All
// elements take the AVX prequant (src/dualquant.hh:83-93); the postquant test
// is the AVX one in blocks ending before DIM0 and the scalar one in the last
// block (src/dualquant.hh:114, :190-205).
template <int B>
__global__ void __launch_bounds__(256) k_dualquant_1d(const float* __restrict__ in, uint16_t* __restrict__ codes,
                                                      float* __restrict__ sparse, unsigned long long* __restrict__ hist,
                                                      DQArgs a) {
    extern __shared__ uint32_t lds[];
    LdsHist H;
    hist_init(H, lds, a.dict, a.radius, hist);
    const uint32_t lane = lane_id();
    const uint64_t ngroups = (a.n + 7) / 8;
    const bool aligned = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    for (uint64_t wb = wave * 64; wb < ngroups; wb += nwaves * 64) {
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
        if ((x0 % B) == 0) prev = 0.f;
        if (g < ngroups) {
            const uint64_t bstart = x0 - (x0 % B);
            const float thr = (bstart + B < a.d0) ? a.thr_vec : a.thr_scalar;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint64_t idx = x0 + j;
                float pred = j == 0 ? prev : q[j - 1];
                uint32_t code = postquant(q[j], pred, thr, a.radius_f);
                if (idx < a.n) {
                    codes[idx] = (uint16_t)code;
                    if (code == 0) sparse[idx] = q[j];
                    H.add((int)code);
                }
            }
        }
    }
    hist_flush(H);
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

// Exclusive scan across a 256-thread workgroup; returns the thread's prefix and
// the workgroup total through *total.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* total, T* smem /* >= 4 */) {
    T inc = wave_incl_scan(v);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 63) smem[w] = inc;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        T s = smem[i];
        if (i < w) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ void __launch_bounds__(256) k_chunk_bits(const uint16_t* __restrict__ codes, uint64_t n,
                                                    const uint8_t* __restrict__ len_tab, int dict,
                                                    uint32_t* __restrict__ chunk_bits, uint32_t* __restrict__ chunk_zeros) {
    __shared__ uint8_t slen[kMaxLdsLenTab];
    __shared__ uint32_t sred[8];
    const bool lds_tab = dict <= kMaxLdsLenTab;
    if (lds_tab)
        for (int i = threadIdx.x; i < dict; i += blockDim.x) slen[i] = len_tab[i];
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kChunk + threadIdx.x * kSymPerThread;
    uint32_t bits = 0, zeros = 0;
    if (base + kSymPerThread <= n) {
        const uint4* p = reinterpret_cast<const uint4*>(codes + base);
        uint4 u[2] = {p[0], p[1]};
        const uint32_t* w = reinterpret_cast<const uint32_t*>(u);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint32_t s0 = w[j] & 0xFFFF, s1 = w[j] >> 16;
            bits += lds_tab ? slen[s0] : len_tab[s0];
            bits += lds_tab ? slen[s1] : len_tab[s1];
            zeros += (s0 == 0) + (s1 == 0);
        }
    } else {
        for (int j = 0; j < kSymPerThread; j++) {
            uint64_t i = base + j;
            if (i < n) {
                uint32_t s = codes[i];
                bits += lds_tab ? slen[s] : len_tab[s];
                zeros += s == 0;
            }
        }
    }
    for (int d = 32; d > 0; d >>= 1) {
        bits += __shfl_xor(bits, d);
        zeros += __shfl_xor(zeros, d);
    }
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        sred[w] = bits;
        sred[4 + w] = zeros;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        chunk_bits[blockIdx.x] = sred[0] + sred[1] + sred[2] + sred[3];
        chunk_zeros[blockIdx.x] = sred[4] + sred[5] + sred[6] + sred[7];
    }
}

// Single workgroup of 1024 threads: exclusive scan of per-chunk bits and
// outliers, then the archive sizes of src/vecsz_interface.hh:189-202 and the
// two length fields of the payload header (src/utils/datapack.hh:288).
__global__ void __launch_bounds__(1024) k_scan_chunks(const uint32_t* __restrict__ chunk_bits,
                                                      const uint32_t* __restrict__ chunk_zeros, uint64_t nchunks,
                                                      uint64_t* __restrict__ bit_off, uint64_t* __restrict__ out_off,
                                                      EncSummary* __restrict__ sum, uint8_t* __restrict__ archive,
                                                      uint64_t meta_bytes, uint64_t tree_bytes, uint64_t out_cap,
                                                      uint32_t* __restrict__ err) {
    __shared__ uint64_t sb[16], sz[16];
    const uint64_t per = (nchunks + blockDim.x - 1) / blockDim.x;
    const uint64_t lo = threadIdx.x * per, hi = min(nchunks, lo + per);
    uint64_t b = 0, z = 0;
    for (uint64_t i = lo; i < hi; i++) {
        b += chunk_bits[i];
        z += chunk_zeros[i];
    }
    uint64_t ib = wave_incl_scan(b), iz = wave_incl_scan(z);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 63) {
        sb[w] = ib;
        sz[w] = iz;
    }
    __syncthreads();
    uint64_t ob = 0, oz = 0, tb = 0, tz = 0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; i++) {
        if (i < w) {
            ob += sb[i];
            oz += sz[i];
        }
        tb += sb[i];
        tz += sz[i];
    }
    ob += ib - b;
    oz += iz - z;
    for (uint64_t i = lo; i < hi; i++) {
        bit_off[i] = ob;
        out_off[i] = oz;
        ob += chunk_bits[i];
        oz += chunk_zeros[i];
    }
    if (threadIdx.x == 0) {
        const uint64_t enc = (tb + 7) / 8;
        const uint64_t huff = 8 + tree_bytes + enc;
        const uint64_t stream_byte = meta_bytes + 8 + tree_bytes;
        EncSummary s;
        s.total_bits = tb;
        s.n_outlier = tz;
        s.huff_size = huff;
        s.archive_size = meta_bytes + huff + 4 * tz;
        s.outlier_byte_off = meta_bytes + huff;
        s.stream_bit_off = stream_byte * 8;
        s.stream_end_byte = stream_byte + enc;
        *sum = s;
        if (s.archive_size > out_cap) {
            *err = 1;
        } else {
            // huff_size and n_outlier fields: the last two u64 of the metadata
            uint8_t* f = archive + meta_bytes - 16;
            for (int i = 0; i < 8; i++) f[i] = (uint8_t)(huff >> (8 * i));
            for (int i = 0; i < 8; i++) f[8 + i] = (uint8_t)(tz >> (8 * i));
        }
    }
}

// OR `l` (1..32) right-aligned bits `v` into the big-endian bit buffer at bit p.
__device__ __forceinline__ void emit_bits(uint32_t* words, uint64_t p, uint32_t v, uint32_t l, uint64_t wlimit) {
    const uint64_t w = p >> 5;
    const uint32_t o = (uint32_t)(p & 31);
    if (w >= wlimit) return;
    if (o + l <= 32) {
        atomicOr(&words[w], v << (32 - o - l));
    } else {
        const uint32_t spill = o + l - 32;
        atomicOr(&words[w], v >> spill);
        if (w + 1 < wlimit) atomicOr(&words[w + 1], v << (32 - spill));
    }
}

__device__ __forceinline__ uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

// One workgroup per chunk.  A chunk's bits start at absolute archive bit A_k;
// it writes the 32-bit words whose first bit lies in [A_k, A_k+1), filling the
// tail of its last word with the first codewords of the next chunk, so no word
// is written by two workgroups and no atomics touch global memory.
__global__ void __launch_bounds__(256) k_pack(const uint16_t* __restrict__ codes, uint64_t n,
                                              const uint32_t* __restrict__ cw_hi, const uint32_t* __restrict__ cw_lo,
                                              const uint8_t* __restrict__ len_tab, int dict,
                                              const uint64_t* __restrict__ bit_off, const uint64_t* __restrict__ out_off,
                                              uint64_t nchunks, const EncSummary* __restrict__ sum,
                                              const float* __restrict__ sparse, uint8_t* __restrict__ archive,
                                              const uint32_t* __restrict__ err, uint32_t lds_words) {
    if (*err) return;
    extern __shared__ uint32_t words[];
    __shared__ uint64_t sscan[8];
    const EncSummary S = *sum;
    const uint64_t k = blockIdx.x;
    const uint64_t A0 = S.stream_bit_off;
    const uint64_t Ak = A0 + bit_off[k];
    const uint64_t An = A0 + (k + 1 < nchunks ? bit_off[k + 1] : S.total_bits);
    const uint64_t Wb = Ak >> 5;
    const uint64_t base_rel = Ak & ~31ull;  // bit position of LDS word 0
    for (uint32_t i = threadIdx.x; i < lds_words; i += blockDim.x) words[i] = 0;
    __syncthreads();

    // This thread's symbols.
    const uint64_t base = k * kChunk + threadIdx.x * kSymPerThread;
    uint16_t sym[kSymPerThread];
    if (base + kSymPerThread <= n) {
        const uint4* p = reinterpret_cast<const uint4*>(codes + base);
        uint4 u[2] = {p[0], p[1]};
        const uint16_t* s16 = reinterpret_cast<const uint16_t*>(u);
#pragma unroll
        for (int j = 0; j < kSymPerThread; j++) sym[j] = s16[j];
    } else {
#pragma unroll
        for (int j = 0; j < kSymPerThread; j++) sym[j] = (base + j < n) ? codes[base + j] : 0xFFFF;
    }
    uint32_t tbits = 0, tz = 0;
#pragma unroll
    for (int j = 0; j < kSymPerThread; j++) {
        if (base + j < n) {
            tbits += len_tab[sym[j]];
            tz += sym[j] == 0;
        }
    }
    uint64_t wg_bits, wg_z;
    uint64_t my_off = block_excl_scan<uint64_t>(tbits, &wg_bits, sscan);
    uint64_t my_z = block_excl_scan<uint64_t>(tz, &wg_z, sscan + 4);

    // Emit this thread's codewords.
    uint64_t p = (Ak - base_rel) + my_off;
    uint64_t r = out_off[k] + my_z;
#pragma unroll
    for (int j = 0; j < kSymPerThread; j++) {
        const uint64_t i = base + j;
        if (i >= n) break;
        const uint32_t s = sym[j];
        const uint32_t l = len_tab[s];
        if (l > 32) {
            emit_bits(words, p, cw_hi[s], l - 32, lds_words);
            emit_bits(words, p + (l - 32), cw_lo[s], 32, lds_words);
        } else if (l > 0) {
            emit_bits(words, p, cw_lo[s], l, lds_words);
        }
        p += l;
        if (s == 0) {
            // outlier compaction, index order
            float v = sparse[i];
            uint32_t bitsv = __float_as_uint(v);
            uint8_t* dst = archive + S.outlier_byte_off + 4 * r;
            dst[0] = (uint8_t)bitsv;
            dst[1] = (uint8_t)(bitsv >> 8);
            dst[2] = (uint8_t)(bitsv >> 16);
            dst[3] = (uint8_t)(bitsv >> 24);
            r++;
        }
    }
    if (S.total_bits == 0) return;

    // Owned words: first bit in [Ak, An) (chunk 0 also owns the word holding A0).
    const uint64_t w_first = (k == 0) ? (A0 >> 5) : ((Ak + 31) >> 5);
    const uint64_t w_end = (An + 31) >> 5;  // exclusive
    if (w_first >= w_end) return;
    // Tail of the last owned word: codewords of the following chunk.
    const uint64_t tail_bits = w_end * 32 - An;
    if (tail_bits > 0 && k + 1 < nchunks && threadIdx.x < 64) {
        const uint64_t nb = (k + 1) * kChunk;
        const uint64_t i = nb + threadIdx.x;
        uint32_t s = 0, l = 0;
        if (threadIdx.x < 32 && i < n) {
            s = codes[i];
            l = len_tab[s];
        }
        uint32_t incl = wave_incl_scan(l);
        uint64_t q = (An - base_rel) + (incl - l);
        if (threadIdx.x < 32 && i < n && (incl - l) < tail_bits) {
            if (l > 32) {
                emit_bits(words, q, cw_hi[s], l - 32, lds_words);
                emit_bits(words, q + (l - 32), cw_lo[s], 32, lds_words);
            } else if (l > 0) {
                emit_bits(words, q, cw_lo[s], l, lds_words);
            }
        }
    }
    __syncthreads();
    uint32_t* out32 = reinterpret_cast<uint32_t*>(archive);
    for (uint64_t W = w_first + threadIdx.x; W < w_end; W += blockDim.x) {
        uint32_t v = words[W - Wb];
        if (W == (A0 >> 5) && (A0 & 31)) {
            // merge with the tree bytes that precede the stream in this word
            const uint32_t keep = (uint32_t)(A0 & 31);
            const uint32_t old = bswap32(out32[W]);
            const uint32_t mask = ~0u << (32 - keep);
            v = (old & mask) | (v & ~mask);
        }
        if ((W + 1) * 4 > S.stream_end_byte) {
            uint8_t* b = archive + W * 4;
            for (int j = 0; j < 4; j++)
                if (W * 4 + j < S.stream_end_byte) b[j] = (uint8_t)(v >> (24 - 8 * j));
        } else {
            out32[W] = bswap32(v);
        }
    }
}

// ------------------------------------------------------------------ launchers

int dq_lds_bytes(int dict) {
    return (kHotBins * 64 + (dict <= kMaxLdsHistBins ? dict : 0)) * 4;
}

void launch_range(const float* x, uint64_t n, uint32_t* out, hipStream_t st) {
    uint64_t n4 = (n / 4 + 255) / 256;
    int grid = (int)std::min<uint64_t>(std::max<uint64_t>(n4, 1), 2048);
    hipLaunchKernelGGL(k_range, dim3(grid), dim3(256), 0, st, x, n, out);
}

void launch_dualquant(int ndim, int B, const float* in, uint16_t* codes, float* sparse, unsigned long long* hist,
                      const DQArgs& a, int grid, hipStream_t st) {
    const int lds = dq_lds_bytes(a.dict);
#define DQ_CASE(ND, KERN, BB)                                                                    \
    if (ndim == ND && B == BB) {                                                                 \
        hipLaunchKernelGGL(KERN<BB>, dim3(grid), dim3(256), lds, st, in, codes, sparse, hist, a); \
        return;                                                                                  \
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
}

void launch_chunk_bits(const uint16_t* codes, uint64_t n, const uint8_t* len_tab, int dict, uint32_t* cb, uint32_t* cz,
                       uint64_t nchunks, hipStream_t st) {
    hipLaunchKernelGGL(k_chunk_bits, dim3((unsigned)nchunks), dim3(256), 0, st, codes, n, len_tab, dict, cb, cz);
}

void launch_scan_chunks(const uint32_t* cb, const uint32_t* cz, uint64_t nchunks, uint64_t* bit_off, uint64_t* out_off,
                        EncSummary* sum, uint8_t* archive, uint64_t meta_bytes, uint64_t tree_bytes, uint64_t out_cap,
                        uint32_t* err, hipStream_t st) {
    hipLaunchKernelGGL(k_scan_chunks, dim3(1), dim3(1024), 0, st, cb, cz, nchunks, bit_off, out_off, sum, archive,
                       meta_bytes, tree_bytes, out_cap, err);
}

void launch_pack(const uint16_t* codes, uint64_t n, const uint32_t* cw_hi, const uint32_t* cw_lo, const uint8_t* len_tab,
                 int dict, const uint64_t* bit_off, const uint64_t* out_off, uint64_t nchunks, const EncSummary* sum,
                 const float* sparse, uint8_t* archive, const uint32_t* err, int max_len, hipStream_t st) {
    uint32_t lds_words = (uint32_t)(((uint64_t)kChunk * (uint64_t)max_len + 64) / 32 + 4);
    hipLaunchKernelGGL(k_pack, dim3((unsigned)nchunks), dim3(256), lds_words * 4, st, codes, n, cw_hi, cw_lo, len_tab,
                       dict, bit_off, out_off, nchunks, sum, sparse, archive, err, lds_words);
}

}  // namespace vsz
