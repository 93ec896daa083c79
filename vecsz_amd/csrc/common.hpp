// Internal definitions shared by the host orchestration and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace vsz {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kChunk = 1024;       // symbols per encode chunk (one wavefront)
constexpr int kSymPerThread = kChunk / kWave;  // 16 consecutive symbols per lane
constexpr int kLutBits = 12;       // first-level decode table width
constexpr int kSubTabBits = 4;     // width of each deeper decode table
constexpr int kSubBits = 1024;     // bits per decode subsequence
constexpr int kCkptBits = 128;     // decoder checkpoint window
constexpr int kCkpts = kSubBits / kCkptBits;

// Error-bound family computed on the device from the value range
// (ChangeToRelativeMode + InitializeErrorBoundFamily, src/types.cc:49-60,
// :84-93), so the compress pipeline does not stop for the range.
struct EbDev {
    double range;     // (double)(float max - float min)
    double eb_final;
    double r_d;       // 1 / (2 eb_final)
    double twoeb;     // 2 eb_final
    float r_f;        // (float) r_d
    uint32_t zero_range;
};

// Scalars of one dual-quantization launch.  Names follow the reference's
// dims_L16 / ebs_L4 (src/constants.hh:14-31).
struct DQArgs {
    uint64_t d0, d1, d2;   // DIM0 (x), DIM1, DIM2
    uint64_t n;            // LEN
    uint64_t x8;           // 8 * floor(DIM0 / 8): first x of the partial AVX group
    float r_f;             // (float) 1/(2 eb_final)
    double r_d;            // 1/(2 eb_final)
    float radius_f;        // (float) radius
    float thr_vec;         // |e| < thr_vec  <=>  sqrt(e*e) < radius  for integral e
    float thr_scalar;      // |e| < thr_scalar (= radius)
    int radius;
    int dict;              // histogram bins
    int reverse;           // visit tiles last-to-first (reuses what a preceding pass left in cache)
    const EbDev* ebd;      // when set, r_f / r_d come from here (r2r range computed on the device)
    uint64_t ntiles_x, ntiles_y, ntiles_z;
    // Padding: halo values {z face, y face, x face} of block (bx, by, bz) at
    // halo + halo_stride * (bx + hb0 * (by + hb1 * bz)), blocks counted on the
    // reference's grid ((d + B + 1) / B per dimension); halo_stride 0 = one
    // set for every block; halo == nullptr = the +0.0 halo.
    const float* halo;
    uint64_t halo_stride, hb0, hb1;
};

// Padding modes and pad value types (szwf.*_padding; MAX_VALUE..CONST_VAL,
// src/utils/padding.hh:10-13).
enum PadMode { kPadNone = 0, kPadGlobal = 1, kPadBlock = 2, kPadEdge = 3 };
enum PadType { kPadMax = 0, kPadMin = 1, kPadAvg = 2, kPadConst = 3 };

// Blocks along a dimension as the reference counts them (src/dimensions.hh:31).
__host__ __device__ inline uint64_t ref_nblk(uint64_t d, int B) { return (d + (uint64_t)B + 1) / (uint64_t)B; }

// Per-field encode summary written by the scan kernel.
struct EncSummary {
    uint64_t total_bits;
    uint64_t n_outlier;
    uint64_t huff_size;
    uint64_t archive_size;
    uint64_t outlier_byte_off;  // byte offset of the outlier section in the archive
    uint64_t stream_bit_off;    // absolute bit offset of the bitstream in the archive
    uint64_t stream_end_byte;   // first byte after the bitstream
};

// One field of a fused batch compress.  A device array of these drives the
// batched kernels; each kernel walks a queue of work units that spans every
// field (unit u belongs to field f when start[f] <= u < start[f + 1]).
// x4 forms: 4 elements per lane, 16-byte loads (x4) or element loads (x4u).
enum DqKind { kDq1 = 0, kDq2 = 1, kDq2x4 = 2, kDq3 = 3, kDq3x4 = 4, kDq2x4u = 5, kDq3x4u = 6, kDqKinds = 7 };
struct BatchField {
    DQArgs a;                   // dual-quant scalars (a.ebd = ebd when the range is found on the device)
    const float* in;
    uint16_t* codes;            // this field's quant codes and outlier scratch
    float* sparse;
    unsigned long long* hist;   // dict bins
    uint32_t* range;            // min/max words (range found on the device), else null
    EbDev* ebd;
    double eb;                  // the caller's error bound (relative for r2r)
    int kind;                   // DqKind
    uint64_t dq_tiles;          // wave tiles of that form
    const uint32_t* book;       // cw_hi[dict], cw_lo[dict], then dict u8 lengths
    const uint8_t* head;        // metadata + Huffman header + tree, staged on the device
    uint64_t head_bytes, meta_bytes, tree_bytes;
    uint8_t* archive;
    uint64_t out_cap;
    uint64_t nchunks, chunk0;   // 1024-symbol chunks, and the first one's index in the batch
};

// Decode index: an optional side product of compress (never part of the
// archive) that lets decompress start every kIdxSeg-symbol segment at its
// first codeword instead of synchronising.  u64 words: [0] kIdxMagic, [1] n,
// [2] segments, [3] stream bits, [4] outliers, [5] kIdxSeg | kIdxGroup << 32;
// then, per group of kIdxGroup segments, the first codeword bit (from the
// stream start) and the outlier rank of its first segment (two u64); then
// per segment its first bit relative to its group's (u32, < 2^19), then its
// outlier rank relative to its group's (u16, < 2^13), each array padded to
// whole words.
constexpr int kIdxSeg = 128;
constexpr int kIdxGroup = 64;
constexpr int kIdxHead = 6;
constexpr uint64_t kIdxMagic = 0x3258444956535A56ull;  // "VZSVIDX2" little-endian
struct IdxLayout {
    uint64_t nseg, ngrp, grp, rel_bits, rel_outs, words;  // counts and word offsets
};
__host__ __device__ inline uint64_t idx_segments(uint64_t n) { return (n + kIdxSeg - 1) / kIdxSeg; }
__host__ __device__ inline IdxLayout idx_layout(uint64_t n) {
    IdxLayout L;
    L.nseg = n ? idx_segments(n) : 0;
    L.ngrp = (L.nseg + kIdxGroup - 1) / kIdxGroup;
    L.grp = kIdxHead;
    L.rel_bits = L.grp + 2 * L.ngrp;
    L.rel_outs = L.rel_bits + (L.nseg + 1) / 2;
    L.words = L.rel_outs + (L.nseg + 3) / 4;
    return L;
}

// Workgroup barriers.  __syncthreads alone does not wait for the calling
// wave's own LDS operations, and on gfx950 a wave past the barrier has read
// LDS before another wave's atomic without return (issued before the
// barrier) had landed: histogram counts were lost, a different few on every
// run.  So every barrier here first waits for the wave's LDS operations
// (lds_sync), or for its global memory operations as well where waves
// exchange data through global memory (mem_sync).
__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __syncthreads();
}
__device__ __forceinline__ void mem_sync() {
    __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0)
    __syncthreads();
}

// Field of work unit u in a batch: the f with start[f] <= u < start[f + 1]
// (start is nondecreasing, start[0] = 0, start[nf] = total units > u).
// Where u is uniform across the workgroup the search runs on scalar loads.
__device__ __forceinline__ int field_of(const uint64_t* __restrict__ start, int nf, uint64_t u) {
    int lo = 0, hi = nf;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= u)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Decoder summary.
struct DecSummary {
    uint32_t error;       // nonzero: corrupt stream
    uint32_t pad_;
    uint64_t total_syms;  // symbols found by the sync pass
};

// Decode table entry.  The first-level table is indexed by the next kLutBits
// bits; an internal node reached after them points to a table of
// 2^kSubTabBits entries indexed by the bits that follow, and so on.
//  bit 31 clear: leaf; bits 16..22 = total code length (1..64), bits 0..15 = symbol
//  bit 31 set  : bits 0..30 = first entry of the next-level table
//  0           : invalid path (corrupt tree)
constexpr uint32_t kLutSlow = 0x80000000u;
// Compact tables for symbols below 4096: 16-bit entries, bits 0..11 a symbol
// or a deeper table's number, bits 12..15 the bits consumed by a leaf in this
// table (first level: total length 1..12; deeper: 1..kSubTabBits), 0 for a
// pointer to the next table.
constexpr uint32_t kMaxSub16 = 12288;  // deeper-table entries kept in LDS
// Node table entry for the bit-serial slow path: child[2*i + bit]
//  bit 31 set: leaf, bits 0..15 symbol; otherwise internal node index; 0xFFFFFFFF = absent.
constexpr uint32_t kNodeLeaf = 0x80000000u;
constexpr uint32_t kNodeAbsent = 0xFFFFFFFFu;

}  // namespace vsz
