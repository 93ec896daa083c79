// Pad values of the padding modes (src/utils/padding.hh; src/dualquant.hh:36-46,
// :226-246, :553-573; src/vecsz_interface.hh:70-90, :287-298).
//
//   k_pad_blocks     one thread per block of the reference's block grid
//                    ((d + B + 1) / B blocks per dimension, src/dimensions.hh:31):
//                    find_pad_value over the block (block mode) or over its
//                    first row / column / planes (edge mode), in the
//                    reference's loop order, then the block_pad / edge_pad
//                    prequantization; 1D keeps the raw value
//   k_pad_global_mm  max / min of the field ignoring NaN, and the first index
//                    holding a zero (the sign of a zero extreme)
//   k_pad_global_sum find_avg's float running sum: one wavefront adds the
//                    field in index order, since float addition is not
//                    associative and the reference's sum is sequential
//   k_pad_global_fin the global pad constant, stored and used values
//   k_pad_halo       decompress: halo values from an archive's pad section
//
// Elements outside the field read as +0.0 (the reference reads its
// uninitialised block buffer there).
#include "common.hpp"
#include "kernels.hpp"

namespace vsz {

namespace {

// find_max / find_min / find_avg (src/utils/padding.hh:65-119) as a running
// fold in the reference's order: max / min keep the first value and replace
// it by a strictly larger / smaller one; avg sums in float.
struct Find {
    int type;
    float m, sum;
    bool first;
    __device__ explicit Find(int t) : type(t), m(0.f), sum(0.f), first(true) {}
    __device__ __forceinline__ void add(float v) {
        if (first) {
            m = v;
            first = false;
        }
        if (type == kPadMax) m = (v > m) ? v : m;
        else if (type == kPadMin) m = (v < m) ? v : m;
        else sum = sum + v;
    }
    __device__ __forceinline__ float result(uint64_t count, float constant) const {
        if (type == kPadConst) return constant;
        if (type == kPadAvg) return sum / (float)count;
        return m;
    }
};

// block_pad / edge_pad prequantization (src/utils/padding.hh:219-220,
// :236-237, :259-263): scaled in double, stored to float, rounded half away.
__device__ __forceinline__ float pad_q_block(float p, double r_d) {
    const float t = (float)((double)p * r_d);
    return (float)__builtin_round((double)t);
}

__device__ __forceinline__ double pad_rd(const PadArgs& a) { return a.ebd ? a.ebd->r_d : a.r_d; }

}  // namespace

__global__ void __launch_bounds__(256) k_pad_blocks(PadArgs a, float* __restrict__ pad_out, float* __restrict__ halo) {
    const uint64_t nb = a.hb0 * a.hb1 * a.hb2;
    const uint64_t id = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (id >= nb) return;
    const uint64_t b0 = id % a.hb0, b1 = (id / a.hb0) % a.hb1, b2 = id / (a.hb0 * a.hb1);
    const uint64_t B = (uint64_t)a.B;
    const uint64_t x0 = b0 * B, y0 = b1 * B, z0 = b2 * B;
    const uint64_t sy = a.d0, sz = a.d0 * a.d1;
    auto val = [&](uint64_t i, uint64_t j, uint64_t k) -> float {
        const uint64_t x = x0 + i, y = y0 + j, z = z0 + k;
        return (x < a.d0 && y < a.d1 && z < a.d2) ? a.x[x + y * sy + z * sz] : 0.f;
    };
    const double rd = pad_rd(a);
    const bool cst = a.type == kPadConst;
    float hz, hy, hx;
    if (a.ndim == 1) {
        // 1D block and edge padding keep the raw value (src/dualquant.hh:37-42)
        Find f(a.type);
        if (!cst)
            for (uint64_t i = 0; i < B; i++) f.add(val(i, 0, 0));
        const float p = f.result(B, a.constant);
        pad_out[id] = p;
        hz = hy = hx = p;
    } else if (a.mode == kPadBlock) {
        Find f(a.type);
        const uint64_t nz = a.ndim >= 3 ? B : 1;
        if (!cst)
            for (uint64_t k = 0; k < nz; k++)
                for (uint64_t j = 0; j < B; j++)
                    for (uint64_t i = 0; i < B; i++) f.add(val(i, j, k));
        const float p = pad_q_block(f.result(B * B * nz, a.constant), rd);
        pad_out[id] = p;
        hz = hy = hx = p;
    } else if (a.ndim == 2) {
        // find_edge_pad_value (src/utils/padding.hh:145-206): the first row,
        // then the first column; halo (i == 0) ? row value : column value
        Find fr(a.type), fc(a.type);
        if (!cst) {
            for (uint64_t i = 0; i < B; i++) fr.add(val(i, 0, 0));
            for (uint64_t j = 0; j < B; j++) fc.add(val(0, j, 0));
        }
        const float p0 = pad_q_block(fr.result(B, a.constant), rd);
        const float p1 = pad_q_block(fc.result(B, a.constant), rd);
        pad_out[2 * id] = p0;
        pad_out[2 * id + 1] = p1;
        hz = 0.f;
        hx = p0;
        hy = p1;
    } else {
        // 3D: the z = 0 plane (x fastest), the x = 0 plane (y fastest) and
        // the y = 0 plane (x fastest); halo (k == 0) ? p0 : (j == 0) ? p1 : p2
        Find f0(a.type), f1(a.type), f2(a.type);
        if (!cst) {
            for (uint64_t j = 0; j < B; j++)
                for (uint64_t i = 0; i < B; i++) f0.add(val(i, j, 0));
            for (uint64_t j = 0; j < B; j++)
                for (uint64_t i = 0; i < B; i++) f1.add(val(0, i, j));
            for (uint64_t j = 0; j < B; j++)
                for (uint64_t i = 0; i < B; i++) f2.add(val(i, 0, j));
        }
        const float p0 = pad_q_block(f0.result(B * B, a.constant), rd);
        const float p1 = pad_q_block(f1.result(B * B, a.constant), rd);
        const float p2 = pad_q_block(f2.result(B * B, a.constant), rd);
        pad_out[3 * id] = p0;
        pad_out[3 * id + 1] = p1;
        pad_out[3 * id + 2] = p2;
        hz = p0;
        hy = p1;
        hx = p2;
    }
    halo[3 * id] = hz;
    halo[3 * id + 1] = hy;
    halo[3 * id + 2] = hx;
}

// scratch: [0] ordered max, [1] ordered min (u32), [2..3] first zero index (u64)
__device__ __forceinline__ uint32_t f2ord_p(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f_p(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__global__ void k_pad_global_init(uint32_t* s) {
    if (threadIdx.x == 0) {
        s[0] = 0u;
        s[1] = 0xFFFFFFFFu;
        reinterpret_cast<unsigned long long*>(s)[1] = ~0ull;
        reinterpret_cast<float*>(s)[4] = 0.f;
    }
}

__global__ void __launch_bounds__(256) k_pad_global_mm(const float* __restrict__ x, uint64_t n, uint32_t* s) {
    uint32_t mx = 0u, mn = 0xFFFFFFFFu;
    unsigned long long z = ~0ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        const float v = x[i];
        if (v == v) {
            const uint32_t o = f2ord_p(v);
            mx = o > mx ? o : mx;
            mn = o < mn ? o : mn;
        }
        if (v == 0.f && i < z) z = i;
    }
    for (int d = 32; d > 0; d >>= 1) {
        const uint32_t a = __shfl_xor(mx, d), b = __shfl_xor(mn, d);
        const unsigned long long c = __shfl_xor(z, d);
        mx = a > mx ? a : mx;
        mn = b < mn ? b : mn;
        z = c < z ? c : z;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&s[0], mx);
        atomicMin(&s[1], mn);
        atomicMin(reinterpret_cast<unsigned long long*>(s) + 1, z);
    }
}

// find_avg over the whole field (src/utils/padding.hh:101-119): the float
// running sum in index order, on one wavefront.  Lanes load 64 consecutive
// values at a time; every lane then adds them in order (the same scalar
// chain in each lane), the next batch's loads already in flight.
__global__ void __launch_bounds__(64) k_pad_global_sum(const float* __restrict__ x, uint64_t n, uint32_t* s) {
    constexpr int U = 16;
    const uint32_t lane = threadIdx.x;
    float sum = 0.f;
    float cur[U], nxt[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint64_t i = (uint64_t)u * 64 + lane;
        cur[u] = i < n ? x[i] : 0.f;
    }
    for (uint64_t base = 0; base < n; base += 64 * U) {
        const uint64_t nb = base + 64 * U;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t i = nb + (uint64_t)u * 64 + lane;
            nxt[u] = i < n ? x[i] : 0.f;
        }
        const uint64_t left = n - base;
        if (left >= 64 * U) {
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int k = 0; k < 64; k++)
                    sum = sum + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur[u]), k));
        } else {
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int k = 0; k < 64; k++) {
                    const float v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur[u]), k));
                    if ((uint64_t)u * 64 + k < left) sum = sum + v;
                }
        }
#pragma unroll
        for (int u = 0; u < U; u++) cur[u] = nxt[u];
    }
    if (lane == 0) reinterpret_cast<float*>(s)[4] = sum;
}

// The global constant (src/vecsz_interface.hh:73-80): find_pad_value over the
// field; stored as round(c * ebx2_r) in double, and used by the 2D/3D blocks
// through block_pad with CONST_VAL, by 1D as the stored value
// (src/dualquant.hh:43-46).
__global__ void k_pad_global_fin(PadArgs a, const uint32_t* s, float* pad_out, float* halo) {
    if (threadIdx.x != 0) return;
    float c;
    if (a.type == kPadConst) {
        c = a.constant;
    } else if (a.type == kPadAvg) {
        c = reinterpret_cast<const float*>(s)[4] / (float)a.n;
    } else {
        const float x0 = a.x[0];
        if (x0 != x0) {
            c = x0;  // nothing compares greater or smaller than a leading NaN
        } else {
            c = ord2f_p(a.type == kPadMax ? s[0] : s[1]);
            // equal values keep the first one: only zeros differ in bits
            if (c == 0.f) c = a.x[reinterpret_cast<const unsigned long long*>(s)[1]];
        }
    }
    const double rd = pad_rd(a);
    const float stored = (float)__builtin_round((double)c * rd);
    const float used = a.ndim == 1 ? stored : pad_q_block(c, rd);
    pad_out[0] = stored;
    halo[0] = halo[1] = halo[2] = used;
}

// Halo values from an archive's pad section (x_block_pad / x_edge_pad,
// src/utils/padding.hh:302-372; src/dualquant.hh:918-927, :951-964,
// :1003-1016; mode from the count, src/vecsz_interface.hh:287-298).
__global__ void __launch_bounds__(256) k_pad_halo(int ndim, int mode, const float* __restrict__ pad, uint64_t pad_len,
                                                  uint64_t nblocks, float* __restrict__ halo) {
    const uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    float hz = 0.f, hy = 0.f, hx = 0.f;
    if (mode == kPadGlobal) {
        hz = hy = hx = pad[0];
    } else if (mode == kPadBlock) {
        hz = hy = hx = pad[b];
    } else if (mode == kPadEdge) {
        if (ndim == 2) {
            hx = pad[2 * b];
            hy = pad[2 * b + 1];
        } else {
            hz = pad[3 * b];
            hy = pad[3 * b + 1];
            hx = pad[3 * b + 2];
        }
    } else if (ndim == 1 && pad_len > 0) {
        hz = hy = hx = pad[0];  // x_lorenzo_1d1l's pad_vals[0] without a mode
    }
    halo[3 * b] = hz;
    halo[3 * b + 1] = hy;
    halo[3 * b + 2] = hx;
}

void launch_pad_blocks(const PadArgs& a, float* pad_out, float* halo, hipStream_t st) {
    const uint64_t nb = a.hb0 * a.hb1 * a.hb2;
    hipLaunchKernelGGL(k_pad_blocks, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, a, pad_out, halo);
}

void launch_pad_global(const PadArgs& a, void* scratch, float* pad_out, float* halo, hipStream_t st) {
    uint32_t* s = static_cast<uint32_t*>(scratch);
    hipLaunchKernelGGL(k_pad_global_init, dim3(1), dim3(64), 0, st, s);
    if (a.type == kPadMax || a.type == kPadMin) {
        const unsigned g = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((a.n + 1023) / 1024, 2048));
        hipLaunchKernelGGL(k_pad_global_mm, dim3(g), dim3(256), 0, st, a.x, a.n, s);
    } else if (a.type == kPadAvg) {
        hipLaunchKernelGGL(k_pad_global_sum, dim3(1), dim3(64), 0, st, a.x, a.n, s);
    }
    hipLaunchKernelGGL(k_pad_global_fin, dim3(1), dim3(64), 0, st, a, s, pad_out, halo);
}

void launch_pad_halo(int ndim, int mode, const float* pad, uint64_t pad_len, uint64_t nblocks, float* halo,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_pad_halo, dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0, st, ndim, mode, pad,
                       pad_len, nblocks, halo);
}

}  // namespace vsz
