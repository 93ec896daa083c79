// Quality metrics of a reconstruction on the device (VerifyData,
// src/utils/verify.hh:30-74): two passes of per-workgroup partials, summed
// on the host in workgroup order.  Per element the arithmetic is the
// reference's (abserr and the relative error in float, sums in double); the
// sums are taken in a different order, so they agree with the reference's
// serial loop to rounding, while min, max, max abserr and its (first) index
// are exact.
#include <float.h>

#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace vsz {

namespace {

constexpr int kVerifyWg = 256;

__device__ __forceinline__ bool better(float e, uint64_t i, float be, uint64_t bi) {
    return e > be || (e == be && i < bi);
}

}  // namespace

__global__ void __launch_bounds__(kVerifyWg) k_verify_pass1(const float* __restrict__ o, const float* __restrict__ x,
                                                            uint64_t n, VerifyPart1* __restrict__ part) {
    __shared__ double s_so[kVerifyWg], s_sx[kVerifyWg], s_sq[kVerifyWg];
    __shared__ float s_mn[kVerifyWg], s_mx[kVerifyWg], s_err[kVerifyWg], s_rel[kVerifyWg];
    __shared__ uint64_t s_idx[kVerifyWg];
    double so = 0, sx = 0, sq = 0;
    float mn = FLT_MAX, mx = -FLT_MAX, err = -1.f, rel = 0.f;
    uint64_t idx = ~0ull;
    const uint64_t stride = (uint64_t)gridDim.x * kVerifyWg;
    for (uint64_t i = blockIdx.x * (uint64_t)kVerifyWg + threadIdx.x; i < n; i += stride) {
        const float ov = o[i], xv = x[i];
        so += ov;
        sx += xv;
        mn = fminf(mn, ov);
        mx = fmaxf(mx, ov);
        const float ae = fabsf(xv - ov);
        if (ov != 0.f) rel = fmaxf(rel, ae / fabsf(ov));
        if (ae > err) {  // indices only grow within a thread: the first maximum wins
            err = ae;
            idx = i;
        }
        sq += (double)(ae * ae);
    }
    const int t = threadIdx.x;
    s_so[t] = so, s_sx[t] = sx, s_sq[t] = sq, s_mn[t] = mn, s_mx[t] = mx, s_err[t] = err, s_rel[t] = rel, s_idx[t] = idx;
    lds_sync();
    for (int h = kVerifyWg / 2; h > 0; h >>= 1) {
        if (t < h) {
            s_so[t] += s_so[t + h];
            s_sx[t] += s_sx[t + h];
            s_sq[t] += s_sq[t + h];
            s_mn[t] = fminf(s_mn[t], s_mn[t + h]);
            s_mx[t] = fmaxf(s_mx[t], s_mx[t + h]);
            s_rel[t] = fmaxf(s_rel[t], s_rel[t + h]);
            if (better(s_err[t + h], s_idx[t + h], s_err[t], s_idx[t])) {
                s_err[t] = s_err[t + h];
                s_idx[t] = s_idx[t + h];
            }
        }
        lds_sync();
    }
    if (t == 0) {
        VerifyPart1 p;
        p.sum_o = s_so[0], p.sum_x = s_sx[0], p.sum_sq = s_sq[0];
        p.min_o = s_mn[0], p.max_o = s_mx[0], p.max_err = s_err[0], p.max_rel = s_rel[0], p.err_idx = s_idx[0];
        part[blockIdx.x] = p;
    }
}

__global__ void __launch_bounds__(kVerifyWg) k_verify_pass2(const float* __restrict__ o, const float* __restrict__ x,
                                                            uint64_t n, double mean_o, double mean_x,
                                                            VerifyPart2* __restrict__ part) {
    __shared__ double s_c[kVerifyWg], s_vo[kVerifyWg], s_vx[kVerifyWg];
    double c = 0, vo = 0, vx = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kVerifyWg;
    for (uint64_t i = blockIdx.x * (uint64_t)kVerifyWg + threadIdx.x; i < n; i += stride) {
        const double dov = (double)o[i] - mean_o, dxv = (double)x[i] - mean_x;
        c += dov * dxv;
        vo += dov * dov;
        vx += dxv * dxv;
    }
    const int t = threadIdx.x;
    s_c[t] = c, s_vo[t] = vo, s_vx[t] = vx;
    lds_sync();
    for (int h = kVerifyWg / 2; h > 0; h >>= 1) {
        if (t < h) {
            s_c[t] += s_c[t + h];
            s_vo[t] += s_vo[t + h];
            s_vx[t] += s_vx[t + h];
        }
        lds_sync();
    }
    if (t == 0) part[blockIdx.x] = VerifyPart2{s_c[0], s_vo[0], s_vx[0]};
}

int verify_grid(uint64_t n) {
    return (int)std::max<uint64_t>(1, std::min<uint64_t>((n + 4 * kVerifyWg - 1) / (4 * kVerifyWg), 2048));
}

void launch_verify_pass1(const float* o, const float* x, uint64_t n, VerifyPart1* part, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_verify_pass1, dim3(grid), dim3(kVerifyWg), 0, st, o, x, n, part);
}

void launch_verify_pass2(const float* o, const float* x, uint64_t n, double mean_o, double mean_x, VerifyPart2* part,
                         int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_verify_pass2, dim3(grid), dim3(kVerifyWg), 0, st, o, x, n, mean_o, mean_x, part);
}

}  // namespace vsz
