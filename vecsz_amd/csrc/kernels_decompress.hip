// gfx950 kernels of the decompress path.
//
//   Huffman decode (decode_withTree/decode, src/huffman.hh:343-372, :679-691)
//   as a chunk-parallel, self-synchronising decode of the archive's single
//   MSB-first bitstream, cut into subsequences of kSubBits bits:
//     k_decode_sync   a workgroup stages 256 subsequences in LDS; each thread
//                     decodes its own from the assumed start j*kSubBits and
//                     records checkpoints (first codeword start in every
//                     kCkptBits window, with the symbol and zero counts before
//                     it); then it continues from its own end into the next
//                     subsequence until a codeword start meets a checkpoint,
//                     after which both decodes agree, and writes the next
//                     subsequence's true start, end and counts
//                     (after kContBits bits without a meeting point it
//                     gives up and lists the subsequence)
//     k_decode_pass   listed subsequences re-decode in full from their true
//                     starts, from LDS-staged rows; a moved end lists the
//                     successor for the next pass, until no list is left
//     k_decode_write  final decode from the true starts: quant codes in
//                     16-byte stores + outlier scatter
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

// The bitstream in "aligned space": word w is the big-endian u32 at byte 4w
// of the 16-byte aligned address at or before the stream start; stream bit p
// is aligned-space bit p + shift0.  Bytes before the stream belong to the
// archive header; bytes at or past the stream end are never read and count
// as zero.
struct StreamView {
    const uint8_t* base;
    uint32_t shift0;
    uint64_t nbytes;  // readable bytes from base (through the stream end)

    __device__ __forceinline__ void setup(const uint8_t* stream, uint64_t stream_bits) {
        // pointer arithmetic (not an integer round trip) keeps the global
        // address space visible to the compiler: global_load, not flat_load
        const uint32_t mis = (uint32_t)(reinterpret_cast<uintptr_t>(stream) & 15);
        base = stream - mis;
        shift0 = 8 * mis;
        nbytes = mis + stream_bits / 8;
    }
    __device__ __forceinline__ uint32_t word(uint64_t w) const {
        const uint64_t b = 4 * w;
        if (b + 4 <= nbytes) return __builtin_bswap32(reinterpret_cast<const uint32_t*>(base)[w]);
        uint32_t v = 0;
        for (int i = 0; i < 4; i++) v = (v << 8) | (b + i < nbytes ? base[b + i] : 0u);
        return v;
    }
};

// Staged words are stored in rows of 33: the 32 words of one 1024-bit
// subsequence followed by a copy of the next row's first word.  Threads of a
// wavefront, each at its own subsequence, then read different LDS banks, and
// any two consecutive words sit side by side (one ds_read2).
__device__ __forceinline__ uint32_t pad_idx(uint32_t w) { return w + (w >> 5); }

// 32-bit window whose MSB is bit `lp` of the staged words.
__device__ __forceinline__ uint32_t win32(const uint32_t* st, uint32_t lp) {
    const uint32_t i0 = pad_idx(lp >> 5);
    return __builtin_elementwise_fshl(st[i0], st[i0 + 1], lp & 31);
}
struct PadWin {  // the rows-of-33 staging above
    const uint32_t* st;
    __device__ __forceinline__ uint32_t operator()(uint32_t lp) const { return win32(st, lp); }
};
struct RowWin {  // plain consecutive words
    const uint32_t* row;
    __device__ __forceinline__ uint32_t operator()(uint32_t lp) const {
        const uint32_t i0 = lp >> 5;
        return __builtin_elementwise_fshl(row[i0], row[i0 + 1], lp & 31);
    }
};

// The same from global memory, at stream bit p.
struct GlobalWin {
    StreamView sv;
    __device__ __forceinline__ uint64_t operator()(uint64_t p) const {
        const uint64_t a = p + sv.shift0;
        const uint64_t w = a >> 5;
        return ((((uint64_t)sv.word(w)) << 32) | sv.word(w + 1)) << (a & 31);
    }
};

// One codeword at position p, 32-bit tables: first level in LDS, deeper
// levels (codes longer than kLutBits) from the global sub-tables.  len = 0
// marks an invalid path.
template <typename Win, typename P>
__device__ __forceinline__ uint32_t decode_sym(const Win& win, P p, const uint32_t* lut,
                                               const uint32_t* __restrict__ sub, uint32_t& len) {
    uint32_t e = lut[(uint32_t)(win(p) >> (64 - kLutBits))];
    if (e & kLutSlow) {
        P q = p + kLutBits;
        for (int lv = 0; lv < 16 && (e & kLutSlow); lv++) {
            e = sub[(e & 0x7FFFFFFFu) + (uint32_t)(win(q) >> (64 - kSubTabBits))];
            q += kSubTabBits;
        }
    }
    len = (e & kLutSlow) ? 0u : (e >> 16) & 0x7F;
    return e & 0xFFFF;
}

// The same with the compact 16-bit tables, every level in LDS.
template <typename Win, typename P>
__device__ __forceinline__ uint32_t decode_sym16(const Win& win, P p, const uint16_t* lut, const uint16_t* sub,
                                                 uint32_t& len) {
    uint32_t e = lut[(uint32_t)(win(p) >> (64 - kLutBits))];
    len = e >> 12;
    if (len == 0) {
        P q = p + kLutBits;
        for (int lv = 0; lv < 16; lv++) {
            e = sub[((e & 0xFFF) << kSubTabBits) + (uint32_t)(win(q) >> (64 - kSubTabBits))];
            if (e >> 12) {
                len = (uint32_t)(q - p) + (e >> 12);
                break;
            }
            q += kSubTabBits;
        }
    }
    return e & 0xFFF;
}

// Decode tables as a kernel sees them; load() copies what lives in LDS to
// the front of the dynamic LDS area and returns the words it used.
template <bool T16>
struct Decoder;

template <>
struct Decoder<true> {
    const uint16_t* lut;
    const uint16_t* sub;
    __device__ __forceinline__ uint32_t load(uint32_t* lds, const DecodeArgs& a) {
        uint16_t* l = reinterpret_cast<uint16_t*>(lds);
        const uint32_t nl = 1u << kLutBits, ns = a.n_sub16;
        for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) l[i] = a.lut16[i];
        for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) l[nl + i] = a.sub16[i];
        lut = l;
        sub = ns ? l + nl : a.sub16;  // n_sub16 0: the deeper tables stay in global memory
        return (nl + ns + 1) / 2;
    }
    template <typename Win, typename P>
    __device__ __forceinline__ uint32_t operator()(const Win& w, P p, uint32_t& len) const {
        return decode_sym16(w, p, lut, sub, len);
    }
    // Decodes the codeword at staged bit lp; len = 0: invalid path.  The
    // window covers codewords of up to 32 bits; longer ones re-read.
    __device__ __forceinline__ uint32_t take_at(const uint32_t* st, uint32_t lp, uint32_t& len) const {
        return take_w(PadWin{st}, lp, len);
    }
    template <typename W>
    __device__ __forceinline__ uint32_t take_w(const W& win, uint32_t lp, uint32_t& len) const {
        return take_v(win, win(lp), lp, len);
    }
    // w: the 32 bits from lp (the window is read again only past them)
    template <typename W>
    __device__ __forceinline__ uint32_t take_v(const W& win, uint32_t w, uint32_t lp, uint32_t& len) const {
        uint32_t e = lut[w >> (32 - kLutBits)];
        uint32_t l = e >> 12;
        if (l == 0) {
            uint32_t used = kLutBits;
            for (int lv = 0; lv < 16; lv++) {
                const uint32_t idx = used + kSubTabBits <= 32 ? (w << used) >> (32 - kSubTabBits)
                                                              : win(lp + used) >> (32 - kSubTabBits);
                e = sub[((e & 0xFFF) << kSubTabBits) + idx];
                if (e >> 12) {
                    l = used + (e >> 12);
                    break;
                }
                used += kSubTabBits;
            }
        }
        len = l;
        return e & 0xFFF;
    }
};

template <>
struct Decoder<false> {
    const uint32_t* lut;
    const uint32_t* sub;
    __device__ __forceinline__ uint32_t load(uint32_t* lds, const DecodeArgs& a) {
        const uint32_t nl = 1u << kLutBits;
        for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) lds[i] = a.lut[i];
        lut = lds;
        sub = a.sub;
        return nl;
    }
    template <typename Win, typename P>
    __device__ __forceinline__ uint32_t operator()(const Win& w, P p, uint32_t& len) const {
        return decode_sym(w, p, lut, sub, len);
    }
    __device__ __forceinline__ uint32_t take_at(const uint32_t* st, uint32_t lp, uint32_t& len) const {
        return take_w(PadWin{st}, lp, len);
    }
    template <typename W>
    __device__ __forceinline__ uint32_t take_w(const W& win, uint32_t lp, uint32_t& len) const {
        return take_v(win, win(lp), lp, len);
    }
    template <typename W>
    __device__ __forceinline__ uint32_t take_v(const W& win, uint32_t w, uint32_t lp, uint32_t& len) const {
        uint32_t e = lut[w >> (32 - kLutBits)];
        if (e & kLutSlow) {
            uint32_t used = kLutBits;
            for (int lv = 0; lv < 16 && (e & kLutSlow); lv++) {
                const uint32_t idx = used + kSubTabBits <= 32 ? (w << used) >> (32 - kSubTabBits)
                                                              : win(lp + used) >> (32 - kSubTabBits);
                e = sub[(e & 0x7FFFFFFFu) + idx];
                used += kSubTabBits;
            }
        }
        len = (e & kLutSlow) ? 0u : (e >> 16) & 0x7F;
        return e & 0xFFFF;
    }
};

constexpr int kSyncWg = 256;                                     // subsequences decoded per workgroup
constexpr int kSyncStride = kSyncWg - 1;                         // ... whose results it writes
constexpr uint32_t kStageWords = kSyncWg * (kSubBits / 32) + 12;  // + shift0 and the tail read by the last codewords
constexpr uint32_t kStageLds = kStageWords + kStageWords / 32 + 1;
constexpr int kCkStride = kCkpts + 1;                            // odd row stride: one bank per thread
constexpr uint32_t kNoCkpt = 0xFF;                               // window without a codeword start
// Bits a continuation decodes before it gives up and lists the subsequence
// for a full re-decode: most continuations meet a checkpoint within a few
// windows (a meeting shows at the first codeword start of the window after
// the two decodes agree, so the budget must span more than one window
// whatever the code lengths), and a wavefront otherwise waits for its
// slowest lane.
constexpr uint32_t kContBits = 288;

// Stages NW aligned-space words starting at w0 (a multiple of 4) into the
// padded layout.  Every thread issues all of its 16-byte loads before the
// first LDS write, so the workgroup waits for one memory latency, not one per
// word.
template <uint32_t NW>
__device__ __forceinline__ void stage_words(uint32_t* st, const StreamView& sv, uint64_t w0) {
    constexpr uint32_t kVec = (NW + 4 * kSyncWg - 1) / (4 * kSyncWg);  // uint4 per thread
    uint4 v[kVec];
    const uint4* src = reinterpret_cast<const uint4*>(sv.base) + w0 / 4;
    if (4 * (w0 + kVec * 4 * kSyncWg) <= sv.nbytes) {
        // whole range inside the stream: unconditional loads, all in flight together
#pragma unroll
        for (uint32_t m = 0; m < kVec; m++) v[m] = src[threadIdx.x + m * kSyncWg];
#pragma unroll
        for (uint32_t m = 0; m < kVec; m++)
            v[m] = make_uint4(__builtin_bswap32(v[m].x), __builtin_bswap32(v[m].y), __builtin_bswap32(v[m].z),
                              __builtin_bswap32(v[m].w));
    } else {
#pragma unroll
        for (uint32_t m = 0; m < kVec; m++) {
            const uint64_t w = w0 + 4 * (threadIdx.x + m * kSyncWg);
            v[m] = make_uint4(sv.word(w), sv.word(w + 1), sv.word(w + 2), sv.word(w + 3));
        }
    }
#pragma unroll
    for (uint32_t m = 0; m < kVec; m++) {
        const uint32_t k = 4 * (threadIdx.x + m * kSyncWg);
        if (k < NW) {
            // four consecutive words never straddle a 32-word row
            uint32_t* d = st + pad_idx(k);
            d[0] = v[m].x;
            d[1] = v[m].y;
            d[2] = v[m].z;
            d[3] = v[m].w;
            if ((k & 31) == 0 && k > 0) d[-1] = v[m].x;  // the previous row's copy of its successor
        }
    }
}

// Checkpoint word: bits 0..7 offset of the first codeword start in the
// window (kNoCkpt: none), 8..18 codewords before it, 19..29 zeros before it.
__device__ __forceinline__ uint32_t ckpt(uint32_t off, uint32_t cnt, uint32_t z) {
    return off | (cnt << 8) | (z << 19);
}

// Dynamic LDS words of the sync kernel after the tables.
constexpr uint32_t kSyncLdsWords = kStageLds + kSyncWg * kCkStride + 2 * kSyncWg;

// Workgroup `unit` of the sync pass: subsequences unit * kSyncStride on.
template <bool T16>
__device__ __forceinline__ void decode_sync_unit(const DecodeArgs& a, const SubState& s, DecSummary* sum,
                                                 uint64_t unit) {
    extern __shared__ uint32_t lds[];
    Decoder<T16> dec;
    uint32_t* st = lds + dec.load(lds, a);
    uint32_t* ck = st + kStageLds;
    uint32_t* own_cz = ck + kSyncWg * kCkStride;  // own codeword count | zero count << 16 (both <= kSubBits)
    uint32_t* own_e = own_cz + kSyncWg;
    StreamView sv;
    sv.setup(a.stream, a.stream_bits);
    const uint64_t J0 = unit * kSyncStride;
    stage_words<kStageWords>(st, sv, J0 * (kSubBits / 32));
    lds_sync();
    const uint32_t t = threadIdx.x;
    const uint64_t j = J0 + t;
    // local position of stream bit p: p - J0*kSubBits + shift0
    auto local = [&](uint64_t p) -> uint32_t { return (uint32_t)(p - J0 * kSubBits) + sv.shift0; };
    bool bad = false;

    // 1. own decode from the assumed start, with checkpoints
    if (j < a.nsub) {
        const uint32_t s0 = t * kSubBits + sv.shift0;
        const uint32_t e0 = local(min((j + 1) * (uint64_t)kSubBits, a.stream_bits));
#pragma unroll
        for (int k = 0; k < kCkpts; k++) ck[t * kCkStride + k] = kNoCkpt;
        // minl: the shortest codeword length seen (0: an invalid path), a
        // per-lane minimum rather than a flag the loop merges across lanes
        uint32_t lp = s0, cnt = 0, z = 0, last_k = 0xFFFFFFFFu, minl = ~0u;
        while (lp < e0) {
            // the first codeword start in a window is its checkpoint; the
            // other codewords write the row's pad slot (never read), so the
            // record needs no branch
            const uint32_t rel = lp - s0;
            const uint32_t kk = rel / kCkptBits;
            ck[t * kCkStride + (kk != last_k ? kk : (uint32_t)kCkpts)] = ckpt(rel % kCkptBits, cnt, z);
            last_k = kk;
            uint32_t len;
            const uint32_t sym = dec.take_at(st, lp, len);
            minl = min(minl, len);
            lp += max(len, 1u);
            cnt++;
            z += sym == 0;
        }
        bad |= minl == 0;
        own_cz[t] = cnt | (z << 16);
        own_e[t] = lp;
        if (j == 0) {  // the stream's first subsequence starts at a true boundary
            s.b[0] = 0;
            s.e[0] = (uint64_t)lp - sv.shift0;
            s.cnt[0] = cnt;
            s.zer[0] = z;
        }
    }
    lds_sync();

    // 2. continuation into subsequence j + 1 from this thread's own end
    if (t < kSyncStride && j + 1 < a.nsub) {
        const uint32_t u = t + 1;
        const uint32_t su = u * kSubBits + sv.shift0;
        const uint32_t eu = local(min((j + 2) * (uint64_t)kSubBits, a.stream_bits));
        const uint32_t q0 = own_e[t];
        uint32_t cc = 0, cz = 0, minl = ~0u, e = 0;
        uint32_t q = q0;
        bool hit = false;
        // One exit, taken when the budget is spent, past u's end, or at a
        // codeword start that meets u's checkpoint; which one is worked out
        // after the loop (exits that each set their own results cost the
        // divergent loop a set of lane masks per exit and iteration).  Past
        // the budget or u's end the read goes to the row's pad slot and is
        // not used.
        for (;;) {
            const uint32_t rel = q - su;
            e = ck[u * kCkStride + min(rel / kCkptBits, (uint32_t)kCkpts)];
            const bool stop = (q - q0 >= kContBits) | (q >= eu);
            hit = !stop && (e & 0xFF) == rel % kCkptBits;
            if (stop | hit) break;
            uint32_t len;
            const uint32_t sym = dec.take_at(st, q, len);
            minl = min(minl, len);
            q += max(len, 1u);
            cc++;
            cz += sym == 0;
        }
        bad |= minl == 0;
        const bool capped = !hit && q - q0 >= kContBits;
        bool met = hit;
        uint32_t before_c = 0, before_z = 0;
        if (hit) {
            before_c = (e >> 8) & 0x7FF;
            before_z = e >> 19;
        } else if (!capped && q == own_e[u]) {  // past u's end, at u's own end
            met = true;
            before_c = own_cz[u] & 0xFFFF;
            before_z = own_cz[u] >> 16;
        }
        const uint64_t base = J0 * kSubBits - sv.shift0;
        s.b[j + 1] = base + q0;
        if (capped) {
            // re-decoded in full from q0 by the first pass, which then compares
            // its end with the own end that j + 2's continuation started from
            s.e[j + 1] = base + own_e[u];
            const uint32_t k = atomicAdd(&s.list_n[0], 1u);
            s.list_j[0][k] = s.jbase + (uint32_t)(j + 1);
            s.list_start[0][k] = base + q0;
        } else if (met) {
            s.cnt[j + 1] = cc + (own_cz[u] & 0xFFFF) - before_c;
            s.zer[j + 1] = cz + (own_cz[u] >> 16) - before_z;
            s.e[j + 1] = base + own_e[u];
        } else {
            // the continuation decoded all of j + 1 from its true start; the
            // check of j + 2 started from j + 1's own (different) end
            s.cnt[j + 1] = cc;
            s.zer[j + 1] = cz;
            s.e[j + 1] = base + q;
            if (j + 2 < a.nsub) {
                const uint32_t k = atomicAdd(&s.list_n[0], 1u);
                s.list_j[0][k] = s.jbase + (uint32_t)(j + 2);
                s.list_start[0][k] = base + q;
            }
        }
    }
    if (bad) atomicOr(&sum->error, 1u);
}

template <bool T16>
__global__ void __launch_bounds__(256) k_decode_sync(DecodeArgs a, SubState s, DecSummary* sum) {
    decode_sync_unit<T16>(a, s, sum, blockIdx.x);
}

// Re-decode pass: each listed subsequence j is decoded in full from its
// listed start (the true end of j - 1); this gives its start, counts and
// end.  An end other than the one j + 1's data was computed from lists j + 1,
// from that end, for the next pass.  Chains shorten by at least one per
// pass.  Workgroups take 256 entries at a time, one per thread; a thread
// stages its entry's row (the subsequence plus the tail its last codeword
// may read) in LDS with independent 16-byte loads, all in flight together,
// then decodes it.  (Staging rows with coalesced word loads, one dependent
// load of the entry number per word, put a memory latency per word on the
// round.)
//
// A row is aligned-space words 32 j .. 32 j + kRowWords - 1, 16-byte aligned
// like the stream base: the decode starts at local bit start + shift0 - 1024 j
// (>= 0, since a subsequence's true start is at or past its first bit), the
// last codeword starts before local bit 1024 + shift0 (<= 1144), and a
// decode reads at most 72 + 32 bits past a codeword start (the deepest
// table walk of an invalid path), so the last word read is word 38.
constexpr uint32_t kRowWords = 40;   // ten 16-byte loads
constexpr uint32_t kRowStride = 41;  // odd: the lanes' reads of one word index fall in different banks
constexpr uint32_t kPassGrid = 1024;  // persistent workgroups of a pass (at most ~3 resident per CU)
__device__ __forceinline__ void stage_row(uint32_t* __restrict__ row, const StreamView& sv, uint64_t j) {
    const uint64_t w0 = j * (kSubBits / 32);
    if (4 * (w0 + kRowWords) <= sv.nbytes) {
        const uint4* src = reinterpret_cast<const uint4*>(sv.base) + w0 / 4;
        uint4 v[kRowWords / 4];
#pragma unroll
        for (uint32_t m = 0; m < kRowWords / 4; m++) v[m] = src[m];
#pragma unroll
        for (uint32_t m = 0; m < kRowWords / 4; m++) {
            row[4 * m] = __builtin_bswap32(v[m].x);
            row[4 * m + 1] = __builtin_bswap32(v[m].y);
            row[4 * m + 2] = __builtin_bswap32(v[m].z);
            row[4 * m + 3] = __builtin_bswap32(v[m].w);
        }
    } else {  // the stream's last rows: bytes past its end read as zero
        for (uint32_t w = 0; w < kRowWords; w++) row[w] = sv.word(w0 + w);
    }
}
// Entries e0 .. e0 + 255 (those below m) of pass `pass`'s list: one entry
// per thread, staged in the thread's own LDS row.
template <bool T16>
__device__ __forceinline__ void redecode_entries(const DecodeArgs& a, const SubState& s, uint32_t pass,
                                                 const Decoder<T16>& dec, uint32_t* rows, const StreamView& sv,
                                                 uint64_t e0, uint32_t m, bool& bad) {
    const uint32_t* lj = s.list_j[pass & 1];
    const uint64_t* ls = s.list_start[pass & 1];
    uint32_t* oj = s.list_j[(pass + 1) & 1];
    uint64_t* os = s.list_start[(pass + 1) & 1];
    uint32_t* on = &s.list_n[(pass + 1) % 3];
    const uint32_t ne = (uint32_t)min<uint64_t>(256, m - e0);
    if (threadIdx.x < ne) {
        const uint64_t j = lj[e0 + threadIdx.x];
        const uint64_t start = ls[e0 + threadIdx.x];
        stage_row(rows + threadIdx.x * kRowStride, sv, j);  // the thread's own row: no barrier
        const uint64_t rowbit = j * kSubBits;  // aligned-space bit of word 0
        const RowWin win{rows + threadIdx.x * kRowStride};
        const uint64_t target = min((j + 1) * (uint64_t)kSubBits, a.stream_bits);
        const uint32_t tl = (uint32_t)(target + sv.shift0 - rowbit);
        uint32_t lp = (uint32_t)(start + sv.shift0 - rowbit);
        uint32_t cnt = 0, z = 0, minl = ~0u;
        while (lp < tl) {
            uint32_t len;
            const uint32_t sym = dec.take_w(win, lp, len);
            minl = min(minl, len);
            lp += max(len, 1u);
            cnt++;
            z += sym == 0;
        }
        bad |= minl == 0;
        const uint64_t pos = rowbit + lp - sv.shift0;
        s.b[j] = start;
        s.cnt[j] = cnt;
        s.zer[j] = z;
        if (pos != s.e[j]) {
            s.e[j] = pos;
            if (j + 1 < a.nsub) {
                const uint32_t k = atomicAdd(on, 1u);
                oj[k] = (uint32_t)(j + 1);
                os[k] = pos;
            }
        }
    }
}

template <bool T16>
__global__ void __launch_bounds__(256) k_decode_pass(DecodeArgs a, SubState s, uint32_t pass, DecSummary* sum) {
    const uint32_t m = s.list_n[pass % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) s.list_n[(pass + 2) % 3] = 0;  // the list after next
    if ((uint64_t)blockIdx.x * 256 >= m) return;  // whole workgroup
    extern __shared__ uint32_t lds[];
    Decoder<T16> dec;
    uint32_t* rows = lds + dec.load(lds, a);
    StreamView sv;
    sv.setup(a.stream, a.stream_bits);
    bool bad = false;
    lds_sync();  // the tables
    for (uint64_t e0 = blockIdx.x * 256ull; e0 < m; e0 += (uint64_t)gridDim.x * 256)
        redecode_entries<T16>(a, s, pass, dec, rows, sv, e0, m, bad);
    if (bad) atomicOr(&sum->error, 1u);
}

// Device-wide exclusive scan of two u32 arrays into u64 offsets.
constexpr int kScanPer = 4;
constexpr int kScanBlock = 1024 * kScanPer;

// guard (nullable): the scan does nothing while *guard is nonzero.
__global__ void __launch_bounds__(1024) k_scan_local(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                                     uint64_t n, uint64_t* __restrict__ oa, uint64_t* __restrict__ ob,
                                                     uint64_t* __restrict__ bsum, const uint32_t* guard) {
    if (guard && *guard) return;
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
    lds_sync();
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

// Exclusive scan of the per-block sums, 1024 at a time on one workgroup
// (an LDS Hillis-Steele scan per pass, carried across passes).
__global__ void __launch_bounds__(1024) k_scan_blocks(uint64_t* __restrict__ bsum, uint64_t nb, uint64_t* __restrict__ totals,
                                                      const uint32_t* guard) {
    if (guard && *guard) return;
    __shared__ uint64_t sa[1024], sb[1024];
    const uint32_t t = threadIdx.x;
    uint64_t ca = 0, cb = 0;
    for (uint64_t base = 0; base < nb; base += 1024) {
        const uint64_t i = base + t;
        const uint64_t x = i < nb ? bsum[2 * i] : 0, y = i < nb ? bsum[2 * i + 1] : 0;
        sa[t] = x;
        sb[t] = y;
        lds_sync();
        for (uint32_t d = 1; d < 1024; d <<= 1) {
            const uint64_t ax = t >= d ? sa[t - d] : 0, bx = t >= d ? sb[t - d] : 0;
            lds_sync();
            sa[t] += ax;
            sb[t] += bx;
            lds_sync();
        }
        if (i < nb) {
            bsum[2 * i] = ca + sa[t] - x;
            bsum[2 * i + 1] = cb + sb[t] - y;
        }
        ca += sa[1023];
        cb += sb[1023];
        lds_sync();
    }
    if (t == 0) {
        totals[0] = ca;
        totals[1] = cb;
    }
}

__global__ void __launch_bounds__(1024) k_scan_add(uint64_t* __restrict__ oa, uint64_t* __restrict__ ob, uint64_t n,
                                                   const uint64_t* __restrict__ bsum, const uint32_t* guard) {
    if (guard && *guard) return;
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

// The stream span of workgroup w's segments [256 w, 256 w + 256) in the
// indexed decode: from the first segment's first codeword to the next
// workgroup's (the stream end for the last), as aligned-space words w_lo ..
// w_lo + nwords (plus kIdxSpanPad words the last codewords' windows read).
// fits: the span is staged in LDS by k_decode_indexed_staged (a span that
// runs backwards, from a damaged index, never fits).
constexpr uint32_t kIdxSpanWords = 9984;  // stage capacity: ~9.75 bits per symbol over 256 segments
constexpr uint32_t kIdxSpanPad = 6;
constexpr uint32_t kIdxSpanOver = 264;    // a damaged stream: 128 codewords of up to 64 bits past the span, and the windows past those
struct IdxSpan {
    uint64_t bs, be;  // stream bits of the span
    uint64_t w_lo, nwords;
    bool fits;
};
__device__ __forceinline__ IdxSpan idx_span(const uint64_t* __restrict__ index, const IdxLayout& L,
                                            uint64_t total_bits, const StreamView& sv, uint64_t w) {
    const uint64_t* grp = index + L.grp;
    const uint32_t* rel_bits = reinterpret_cast<const uint32_t*>(index + L.rel_bits);
    const uint64_t g0 = w * 256, g1 = min<uint64_t>(g0 + 256, L.nseg);
    const uint64_t bs = grp[2 * (g0 / kIdxGroup)] + rel_bits[g0];
    const uint64_t be = g1 < L.nseg ? grp[2 * (g1 / kIdxGroup)] + rel_bits[g1] : total_bits;
    IdxSpan sp;
    sp.bs = bs;
    sp.be = be;
    sp.w_lo = (bs + sv.shift0) >> 5;
    sp.nwords = be >= bs ? ((be + sv.shift0) >> 5) - sp.w_lo + kIdxSpanPad : 0;
    sp.fits = be >= bs && sp.nwords <= kIdxSpanWords;
    return sp;
}

// Window over the stream for the indexed decode: 64 bits held in a register
// (relative bit b0 at its MSB, refilled a word at a time by the caller);
// windows beyond it (the deeper levels of codewords over 32 bits) are read
// from memory.  Relative bit 0 is aligned-space bit 32 * w0.
struct CursorWin {
    StreamView sv;
    uint64_t w0;
    uint64_t buf;
    uint32_t b0;
    __device__ __forceinline__ uint32_t operator()(uint32_t lp) const {
        const uint32_t o = lp - b0;
        if (o <= 32) return (uint32_t)((buf << o) >> 32);
        const uint64_t ab = 32 * w0 + lp;
        const uint64_t wi = ab >> 5;
        return __builtin_elementwise_fshl(sv.word(wi), sv.word(wi + 1), (uint32_t)(ab & 31));
    }
};

// Decode with a decode index, for the workgroups whose span does not fit the
// LDS stage of k_decode_indexed_staged (it returns at once for the others):
// thread g decodes segment g (kIdxSeg symbols)
// from the segment's first codeword, reading the stream through a 64-bit
// register window refilled from 16-byte groups loaded ahead, so no LDS
// staging of the bits and no synchronisation is needed.  Codes leave through
// LDS: every lane puts 32 symbols (four 16-byte units) in its wavefront's
// stage, then the wavefront writes the stage so that each store instruction
// covers 16 whole 64-byte pieces instead of 64 scattered 16-byte units.
// Outliers are written from the segment's outlier rank on.  A segment that
// does not end where the next one starts (an index that does not belong to
// this stream) is an error.
constexpr uint32_t kIdxFlush = 32;                       // symbols per lane between stage flushes
constexpr uint32_t kIdxLaneWords = 16 + 1;               // u32 per lane in the stage (odd: no bank conflicts)
constexpr uint32_t kIdxStageWords = 64 * kIdxLaneWords;  // per wavefront
// R1D: the 1D Lorenzo reconstruction done in the same pass (a segment holds
// whole blocks for every B), writing floats instead of codes.
template <bool T16, bool R1D>
__global__ void __launch_bounds__(256) k_decode_indexed(DecodeArgs a, const uint64_t* __restrict__ index,
                                                        IdxLayout L, uint64_t total_bits, uint16_t* __restrict__ codes,
                                                        float* __restrict__ sparse, Recon1D r, DecSummary* sum) {
    {  // workgroups k_decode_indexed_staged has decoded (the whole workgroup)
        StreamView sv0;
        sv0.setup(a.stream, a.stream_bits);
        if (idx_span(index, L, total_bits, sv0, blockIdx.x).fits) return;
    }
    // symbols per lane between flushes (16 words of stage either way) and
    // per 16-byte piece of output
    constexpr uint32_t kFlush = R1D ? 16 : kIdxFlush;
    constexpr uint32_t kPer16 = R1D ? 4 : 8;
    extern __shared__ uint32_t lds[];
    Decoder<T16> dec;
    const uint32_t tw = dec.load(lds, a);
    const uint32_t lane = threadIdx.x & 63;
    uint32_t* stage = lds + tw + (threadIdx.x >> 6) * kIdxStageWords;
    lds_sync();
    const uint64_t g = blockIdx.x * 256ull + threadIdx.x;
    const uint64_t gw = g - lane;  // the wavefront's first segment
    const uint64_t nseg = L.nseg;
    const bool live = g < nseg;
    StreamView sv;
    sv.setup(a.stream, a.stream_bits);
    const uint64_t s0 = g * kIdxSeg;
    const uint32_t ns = live ? (uint32_t)min<uint64_t>(kIdxSeg, a.n - s0) : 0u;
    const uint64_t* grp = index + L.grp;
    const uint32_t* rel_bits = reinterpret_cast<const uint32_t*>(index + L.rel_bits);
    const uint16_t* rel_outs = reinterpret_cast<const uint16_t*>(index + L.rel_outs);
    auto first_bit = [&](uint64_t h) { return grp[2 * (h / kIdxGroup)] + rel_bits[h]; };
    const uint64_t start = live ? first_bit(g) : 0;
    const uint64_t end = !live ? 0 : g + 1 < nseg ? first_bit(g + 1) : total_bits;
    uint64_t rank = live ? grp[2 * (g / kIdxGroup) + 1] + rel_outs[g] : 0;
    bool bad = start > end || end > a.stream_bits;
    const uint64_t abit = start + sv.shift0;
    CursorWin cw;
    cw.sv = sv;
    cw.w0 = abit >> 5;
    cw.b0 = 0;
    cw.buf = live ? ((uint64_t)sv.word(cw.w0) << 32) | sv.word(cw.w0 + 1) : 0;
    // refill words come from 16-byte groups: the current one and the next,
    // already loaded (about 14 codewords of decode ahead of their use)
    auto load4 = [&](uint64_t q) -> uint4 {
        if (16 * (q + 1) <= sv.nbytes) {
            const uint4 v = reinterpret_cast<const uint4*>(sv.base)[q];
            return make_uint4(__builtin_bswap32(v.x), __builtin_bswap32(v.y), __builtin_bswap32(v.z),
                              __builtin_bswap32(v.w));
        }
        return make_uint4(sv.word(4 * q), sv.word(4 * q + 1), sv.word(4 * q + 2), sv.word(4 * q + 3));
    };
    uint64_t q4 = (cw.w0 + 2) >> 2;
    uint32_t qi = (uint32_t)((cw.w0 + 2) & 3);
    uint4 cur4 = make_uint4(0, 0, 0, 0), nxt4 = make_uint4(0, 0, 0, 0);
    if (live) {
        cur4 = load4(q4);
        nxt4 = load4(q4 + 1);
    }
    q4 += 2;
    uint32_t lp = (uint32_t)(abit & 31);
    float sprev = 0.f;  // R1D: the previous element's value in the block
    // the wavefront's segments lie kIdxSeg symbols apart; the last field
    // segment may be short, and only its whole pieces go through the stage
    for (uint32_t f0 = 0; f0 < kIdxSeg; f0 += kFlush) {
#pragma unroll
        for (uint32_t u = f0; u < f0 + kFlush; u += 8) {
            // Every lane decodes 8 codewords; a lane whose (last, short)
            // segment has ended keeps decoding at its end position (nothing
            // kept): its errors and outliers count only up to ns.  Outliers
            // wait for the end of the 8.
            uint32_t c[8];
            float o[8];
            uint32_t omask = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                // one refill covers codewords of at most 32 bits (selects, no
                // branch); the 16-byte reload follows every fourth
                {
                    const bool need = lp - cw.b0 >= 32;
                    const uint32_t wv = qi == 0 ? cur4.x : qi == 1 ? cur4.y : qi == 2 ? cur4.z : cur4.w;
                    cw.buf = need ? (cw.buf << 32) | wv : cw.buf;
                    cw.b0 += need ? 32u : 0u;
                    qi += need ? 1u : 0u;
                    if (qi == 4) {
                        cur4 = nxt4;
                        nxt4 = load4(q4++);
                        qi = 0;
                    }
                }
                // after a longer codeword, more
                while (lp - cw.b0 >= 32) {
                    const uint32_t wv = qi == 0 ? cur4.x : qi == 1 ? cur4.y : qi == 2 ? cur4.z : cur4.w;
                    cw.buf = (cw.buf << 32) | wv;
                    cw.b0 += 32;
                    if (++qi == 4) {
                        cur4 = nxt4;
                        nxt4 = load4(q4++);
                        qi = 0;
                    }
                }
                // lp - cw.b0 < 32 here: the codeword's first 32 bits are in buf
                const uint32_t w = (uint32_t)((cw.buf << (lp - cw.b0)) >> 32);
                uint32_t len;
                const uint32_t sym = dec.take_v(cw, w, lp, len);
                // past ns (or at an invalid codeword, an error) the lane
                // stays where it is: lp ends at the segment's end
                const bool in = u + k < ns;
                bad |= in & (len == 0);
                lp += in ? len : 0u;
                c[k] = sym;
                omask |= (uint32_t)(in & (sym == 0)) << k;
            }
            float ov[8];
#pragma unroll
            for (int k = 0; k < 8; k++) ov[k] = 0.f;
            if (omask) {
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    if ((omask >> k) & 1) {
                        float v = 0.f;
                        if (rank < a.n_outlier)
                            v = read_f32_bytes(a.outliers + 4 * rank);
                        else
                            bad = true;
                        rank++;
                        ov[k] = v;
                        if constexpr (!R1D) sparse[s0 + u + k] = v;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                o[k] = 0.f;
                if constexpr (R1D) {
                    // x_lorenzo_1d1l (src/dualquant.hh:915-931): a block's
                    // first element predicts from the halo (+0.0 or the
                    // block's pad value), the others from their predecessor
                    // (only for elements of the field: lanes past its end,
                    // whose values are dropped, must not index the halo,
                    // which has one entry per block of the field)
                    const uint64_t x = s0 + u + k;
                    if ((x % (uint64_t)r.B) == 0 && u + k < ns)
                        sprev = r.halo ? r.halo[r.halo_stride * (x / r.B) + 2] : 0.f;
                    const float sx = c[k] == 0 ? ov[k] : sprev + (float)((int32_t)c[k] - r.radius);
                    sprev = sx;
                    o[k] = (float)((double)sx * r.twoeb);
                }
            }
            uint32_t* d = stage + lane * kIdxLaneWords + (u - f0) * 16 / kFlush;
            if constexpr (R1D) {
#pragma unroll
                for (int k = 0; k < 8; k++) d[k] = __float_as_uint(o[k]);
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) d[k] = c[2 * k] | (c[2 * k + 1] << 16);
            }
            // pieces that end past a short last segment: straight out
            if (ns != kIdxSeg) {
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const uint32_t i = u + k;
                    if (i < ns && (i / kPer16 + 1) * kPer16 > ns) {
                        if constexpr (R1D)
                            r.out[s0 + i] = o[k];
                        else
                            codes[s0 + i] = (uint16_t)c[k];
                    }
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // piece i of this lane's store: lane q's piece j, slot lane + 64 i
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            const uint32_t slot = lane + 64 * i, q = slot / 4, j = slot % 4;
            const uint64_t gr = gw + q;
            const uint32_t nr = gr < nseg ? (uint32_t)min<uint64_t>(kIdxSeg, a.n - gr * kIdxSeg) : 0u;
            if (f0 + kPer16 * (j + 1) <= nr) {
                const uint32_t* w = stage + q * kIdxLaneWords + 4 * j;
                const uint4 val = make_uint4(w[0], w[1], w[2], w[3]);
                // codes and the 1D output are 16-byte aligned; gr * kIdxSeg + f0 + kPer16 * j is a
                // multiple of kPer16
                if constexpr (R1D)
                    *reinterpret_cast<uint4*>(r.out + gr * kIdxSeg + f0 + kPer16 * j) = val;
                else
                    *reinterpret_cast<uint4*>(codes + gr * kIdxSeg + f0 + kPer16 * j) = val;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // every segment ends where the next starts, the last at the stream's
    // last codeword bit
    if (live) bad |= 32 * cw.w0 + lp - sv.shift0 != end;
    if (bad) atomicOr(&sum->error, 1u);
}

// Decode with a decode index, the bits staged in LDS: the workgroup's 256
// segments are one contiguous span of the stream, staged with coalesced
// loads, and thread g decodes segment g out of LDS as the write pass decodes
// a subsequence.  A segment holds kIdxSeg = 128 symbols, 16 whole 16-byte
// units, so every thread writes its own units, in groups of four held in
// registers and stored back to back (R1D: the 1D reconstruction in the
// same pass, a segment holding whole blocks for every B, writing floats in
// groups of 16).  Workgroups whose span does not fit the stage return at
// once and leave their segments to k_decode_indexed.
template <bool T16, bool R1D>
__global__ void __launch_bounds__(256) k_decode_indexed_staged(DecodeArgs a, const uint64_t* __restrict__ index,
                                                               IdxLayout L, uint64_t total_bits,
                                                               uint16_t* __restrict__ codes,
                                                               float* __restrict__ sparse, Recon1D r,
                                                               DecSummary* sum) {
    StreamView sv;
    sv.setup(a.stream, a.stream_bits);
    const IdxSpan sp = idx_span(index, L, total_bits, sv, blockIdx.x);
    if (!sp.fits) return;  // (the whole workgroup)
    extern __shared__ uint32_t lds[];
    Decoder<T16> dec;
    uint32_t* stage = lds + dec.load(lds, a);
    for (uint32_t i = threadIdx.x; i < (uint32_t)sp.nwords; i += 256) stage[i] = sv.word(sp.w_lo + i);
    lds_sync();
    const uint64_t g = blockIdx.x * 256ull + threadIdx.x;
    if (g >= L.nseg) return;  // no barrier follows
    const uint64_t* grp = index + L.grp;
    const uint32_t* rel_bits = reinterpret_cast<const uint32_t*>(index + L.rel_bits);
    const uint16_t* rel_outs = reinterpret_cast<const uint16_t*>(index + L.rel_outs);
    const uint64_t start = grp[2 * (g / kIdxGroup)] + rel_bits[g];
    const uint64_t end = g + 1 < L.nseg ? grp[2 * ((g + 1) / kIdxGroup)] + rel_bits[g + 1] : total_bits;
    uint64_t rank = grp[2 * (g / kIdxGroup) + 1] + rel_outs[g];
    const uint64_t s0 = g * kIdxSeg;
    const uint64_t span_lo = 32 * sp.w_lo;  // aligned-space bit of stage word 0
    // a segment outside the span (a damaged index) is an error, not decoded
    bool bad = start > end || end > a.stream_bits || start < sp.bs || end > sp.be;
    const uint32_t ns = bad ? 0u : (uint32_t)min<uint64_t>(kIdxSeg, a.n - s0);
    const RowWin win{stage};
    uint32_t lp = bad ? 0u : (uint32_t)(start + sv.shift0 - span_lo);
    // the loops count symbols, so an invalid codeword (len 0, an error)
    // needs no forced progress; a damaged stream stays within kIdxSpanOver
    // words past the span
    uint32_t minl = ~0u;  // shortest codeword length seen (0: an invalid path)
    auto next = [&]() -> uint32_t {
        uint32_t len;
        const uint32_t sym = dec.take_w(win, lp, len);
        minl = min(minl, len);
        lp += len;
        return sym;
    };
    auto outlier = [&](uint64_t i) {
        if (rank < a.n_outlier)
            sparse[i] = read_f32_bytes(a.outliers + 4 * rank);
        else
            bad = true;
        rank++;
    };
    if constexpr (R1D) {
        // x_lorenzo_1d1l (src/dualquant.hh:915-931), as in k_decode_indexed
        float sprev = 0.f;
        auto value = [&](uint32_t c, float ov, uint64_t x) -> float {
            if ((x % (uint64_t)r.B) == 0) sprev = r.halo ? r.halo[r.halo_stride * (x / r.B) + 2] : 0.f;
            const float sx = c == 0 ? ov : sprev + (float)((int32_t)c - r.radius);
            sprev = sx;
            return (float)((double)sx * r.twoeb);
        };
        const uint32_t nfull16 = ns & ~15u;
        for (uint32_t G = 0; G < nfull16; G += 16) {
            uint32_t c[16], omask = 0;
#pragma unroll
            for (int k = 0; k < 16; k++) {
                c[k] = next();
                omask |= (uint32_t)(c[k] == 0) << k;
            }
            float ov[16];
#pragma unroll
            for (int k = 0; k < 16; k++) ov[k] = 0.f;
            if (omask) {
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    if ((omask >> k) & 1) {
                        if (rank < a.n_outlier)
                            ov[k] = read_f32_bytes(a.outliers + 4 * rank);
                        else
                            bad = true;
                        rank++;
                    }
                }
            }
            float o[16];
#pragma unroll
            for (int k = 0; k < 16; k++) o[k] = value(c[k], ov[k], s0 + G + k);
#pragma unroll
            for (int q = 0; q < 4; q++)
                *reinterpret_cast<float4*>(r.out + s0 + G + 4 * q) =
                    make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
        }
        for (uint32_t i = nfull16; i < ns; i++) {  // the field's last, partial group
            const uint32_t c = next();
            float ov = 0.f;
            if (c == 0) {
                if (rank < a.n_outlier)
                    ov = read_f32_bytes(a.outliers + 4 * rank);
                else
                    bad = true;
                rank++;
            }
            r.out[s0 + i] = value(c, ov, s0 + i);
        }
    } else {
        const uint32_t nfull = ns & ~7u;  // whole units
        for (uint32_t G = 0; G < nfull; G += 32) {
            // outliers are rare: one test per group, then in symbol order
            uint32_t w[16], omask = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const bool mine = G + 8 * u < nfull;  // differs across lanes only for the field's last segment
#pragma unroll
                for (int m = 0; m < 4; m++) w[4 * u + m] = 0;
                if (mine) {
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        const uint32_t lo = next();
                        const uint32_t hi = next();
                        omask |= ((uint32_t)(lo == 0) | ((uint32_t)(hi == 0) << 1)) << (8 * u + 2 * m);
                        w[4 * u + m] = lo | (hi << 16);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (G + 8 * u < nfull)
                    *reinterpret_cast<uint4*>(codes + s0 + G + 8 * u) =
                        make_uint4(w[4 * u], w[4 * u + 1], w[4 * u + 2], w[4 * u + 3]);
            if (omask) {
                for (int k = 0; k < 32; k++)
                    if ((omask >> k) & 1) outlier(s0 + G + k);
            }
        }
        for (uint32_t i = nfull; i < ns; i++) {  // the field's last, partial unit
            const uint32_t sym = next();
            codes[s0 + i] = (uint16_t)sym;
            if (sym == 0) outlier(s0 + i);
        }
    }
    bad |= minl == 0;
    // the segment ends where the next one starts
    if (!bad) bad = span_lo + lp - sv.shift0 != end;
    if (bad) atomicOr(&sum->error, 1u);
}

// Final decode.  Thread j decodes cnt[j] codewords from its true start b[j]
// out of the workgroup's staged bits: its symbols are [O, E), O = sym_off[j].
// Output goes out in whole 16-byte units of 8 symbols: the thread whose range
// holds a unit's first symbol writes the unit, decoding on past E into the
// following subsequences where the unit ends beyond it (the decode is on the
// true path there), and skips the symbols before its first unit (written by
// a predecessor).  Only the field's last unit may be partial.
//
// A workgroup (256 threads) takes kWriteWg = 240 subsequences, so its LDS
// (the 8 KiB first-level table and ~31 KiB of staged bits) stays under
// 40 KiB and four workgroups fit a CU instead of three: the decode is
// latency-bound, and the resident lanes count for more than the 16 that
// only help stage.
constexpr uint32_t kWriteWg = 240;
constexpr uint32_t kWriteStageWords = kWriteWg * (kSubBits / 32) + 32;  // + up to 7 codewords past the last range
constexpr uint32_t kWriteStageLds = kWriteStageWords + kWriteStageWords / 32 + 3;
// sym_off / zero_off: exclusive prefixes of the subsequences' codeword and
// zero counts, counted from the field's subsequence 0 (entry 0 is the base
// subtracted: 0 for one field, the earlier fields' totals in a batch).
template <bool T16>
__device__ __forceinline__ void decode_write_unit(const DecodeArgs& a, const SubState& s,
                                                  const uint64_t* __restrict__ sym_off,
                                                  const uint64_t* __restrict__ zero_off, uint16_t* __restrict__ codes,
                                                  float* __restrict__ sparse, DecSummary* sum, uint64_t unit,
                                                  const uint32_t* guard = nullptr) {
    if (guard && *guard) return;  // (the whole workgroup)
    extern __shared__ uint32_t lds[];
    Decoder<T16> dec;
    uint32_t* st = lds + dec.load(lds, a);
    StreamView sv;
    sv.setup(a.stream, a.stream_bits);
    const uint64_t J0 = unit * kWriteWg;
    stage_words<kWriteStageWords>(st, sv, J0 * (kSubBits / 32));
    lds_sync();
    const uint64_t j = J0 + threadIdx.x;
    if (threadIdx.x >= kWriteWg || j >= a.nsub) return;
    uint32_t lp = (uint32_t)(s.b[j] - J0 * kSubBits) + sv.shift0;
    const uint64_t O = sym_off[j] - sym_off[0];
    const uint64_t E = min(O + s.cnt[j], a.n);
    uint64_t rank = zero_off[j] - zero_off[0];
    bool bad = false;
    // the loops below count symbols, so an invalid codeword (len 0, an
    // error) needs no forced progress
    uint32_t minl = ~0u;  // shortest codeword length seen (0: an invalid path)
    auto next = [&]() -> uint32_t {
        uint32_t len;
        const uint32_t sym = dec.take_at(st, lp, len);
        minl = min(minl, len);
        lp += len;
        return sym;
    };
    auto outlier = [&](uint64_t i) {
        if (rank < a.n_outlier)
            sparse[i] = read_f32_bytes(a.outliers + 4 * rank);
        else
            bad = true;
        rank++;
    };
    const uint64_t W0 = (O + 7) & ~7ull;  // this thread's first unit
    uint64_t i = O;
    for (; i < E && i < W0; i++)          // a predecessor's unit: decode, count zeros, skip
        rank += next() == 0;
    if (W0 < E) {
        const uint64_t W1 = min((E + 7) & ~7ull, a.n);
        const uint64_t Wf = W0 + ((W1 - W0) & ~7ull);  // end of the whole units
        // Whole units go out in groups of four (64 bytes) held in registers
        // and stored back to back: the L2 then sees long runs of each lane's
        // output, where units stored as they were decoded left partially
        // written lines that were written back several times.  Groups start
        // at the lane's first unit, so only its last group is partial.
        for (uint64_t G = W0; G < Wf; G += 32) {
            // outliers are rare: one test per group, then in symbol order
            uint32_t w[16], omask = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint64_t iu = G + 8 * u;
                const bool mine = iu < Wf;  // differs across lanes only in a last group
#pragma unroll
                for (int m = 0; m < 4; m++) w[4 * u + m] = 0;
                if (mine) {
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        const uint32_t lo = next();
                        const uint32_t hi = next();
                        omask |= ((uint32_t)(lo == 0) | ((uint32_t)(hi == 0) << 1)) << (8 * u + 2 * m);
                        w[4 * u + m] = lo | (hi << 16);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint64_t iu = G + 8 * u;
                if (iu < Wf)
                    *reinterpret_cast<uint4*>(codes + iu) =
                        make_uint4(w[4 * u], w[4 * u + 1], w[4 * u + 2], w[4 * u + 3]);
            }
            if (omask) {
                for (int k = 0; k < 32; k++)
                    if ((omask >> k) & 1) outlier(G + k);
            }
        }
        i = Wf;
        for (; i < W1; i++) {  // the field's last, partial unit
            const uint32_t sym = next();
            codes[i] = (uint16_t)sym;
            if (sym == 0) outlier(i);
        }
    }
    bad |= minl == 0;
    if (bad) atomicOr(&sum->error, 1u);
}

template <bool T16>
__global__ void __launch_bounds__(256) k_decode_write(DecodeArgs a, SubState s, const uint64_t* __restrict__ sym_off,
                                                      const uint64_t* __restrict__ zero_off, uint16_t* __restrict__ codes,
                                                      float* __restrict__ sparse, DecSummary* sum, const uint32_t* guard) {
    decode_write_unit<T16>(a, s, sym_off, zero_off, codes, sparse, sum, blockIdx.x, guard);
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

// Halo values of the block holding x, for the integer recurrences: exact
// when every one is an integer below kExact (else the block is listed).
struct IHalo {
    int32_t z, y, x;
    bool exact;
};
__device__ __forceinline__ IHalo load_ihalo(const ReconArgs& a, bool xin, uint64_t x, int B, uint64_t by,
                                            uint64_t bz) {
    IHalo h{0, 0, 0, true};
    if (xin) {
        const float* p = a.halo + a.halo_stride * ((x / B) + a.hb0 * (by + a.hb1 * bz));
        h.exact = exact_outlier(p[0]) && exact_outlier(p[1]) && exact_outlier(p[2]);
        h.z = h.exact ? (int32_t)p[0] : 0;
        h.y = h.exact ? (int32_t)p[1] : 0;
        h.x = h.exact ? (int32_t)p[2] : 0;
    }
    return h;
}

// PAD: with halo values h (x_block_pad / x_edge_pad), r = s - Y - Z + YZ
// is still a running sum along x, with Y, Z, YZ read from the halo faces in
// the first row / plane of the block and a start of h.x - h.y in the block's
// first row of its first plane (the only row where X - XY - XZ + XYZ of the
// halo does not cancel).
template <int B, bool PAD>
__device__ __forceinline__ void recon_3d_tile(uint64_t t, const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, const ReconArgs& a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    const uint32_t lane = lane_id2();
    const uint64_t sy = a.d0, sz = a.d0 * a.d1;
    const uint64_t nbx = (a.d0 + B - 1) / B, nby = (a.d1 + B - 1) / B;
    {
        const uint64_t tx = t % a.ntiles_x;
        const uint64_t ty = (t / a.ntiles_x) % a.ntiles_y;
        const uint64_t tz = t / (a.ntiles_x * a.ntiles_y);
        const uint64_t x = tx * 64 + lane;
        const uint64_t y0 = ty * B, z0 = tz * B;
        const bool xin = x < a.d0;
        const bool bstart = (lane % B) == 0;
        bool bad = false;
        IHalo h{0, 0, 0, true};
        if constexpr (PAD) {
            h = load_ihalo(a, xin, x, B, ty, tz);
            bad = !h.exact;
        }
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
                    const int32_t Y = yy > 0 ? srow : h.y;
                    const int32_t Z = zz > 0 ? prevplane[yy] : h.z;
                    const int32_t YZ = (yy > 0 && zz > 0) ? zrow : (zz == 0 ? h.z : h.y);
                    int32_t g0;
                    float v = 0.f;
                    if (isout) {
                        v = sparse[idx];
                        if (!exact_outlier(v)) bad = true;
                        g0 = (int32_t)v - Y - Z + YZ;
                    } else {
                        g0 = (int32_t)c - a.radius;
                        if (PAD && bstart && zz == 0 && yy == 0) g0 += h.x - h.y;
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

template <int B, bool PAD = false>
__global__ void __launch_bounds__(256) k_recon_3d(const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, ReconArgs a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    if (a.guard && *a.guard) return;  // decode output not final: a later launch redoes this
    const uint64_t ntiles = a.ntiles_x * a.ntiles_y * a.ntiles_z;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    for (uint64_t t = wave; t < ntiles; t += nwaves) recon_3d_tile<B, PAD>(t, codes, sparse, out, a, flagged, nflagged);
}

// PAD: r = s - N runs along x from 0 at the block start (W - NW of the
// x-face halo cancels), with N = h.y in the block's first row.
template <int B, bool PAD>
__device__ __forceinline__ void recon_2d_tile(uint64_t t, const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, const ReconArgs& a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    const uint32_t lane = lane_id2();
    const uint64_t nbx = (a.d0 + B - 1) / B, nby = (a.d1 + B - 1) / B;
    {
        const uint64_t tx = t % a.ntiles_x;
        const uint64_t ty = t / a.ntiles_x;
        const uint64_t x = tx * 64 + lane;
        const uint64_t y0 = ty * B;
        const bool xin = x < a.d0;
        const bool bstart = (lane % B) == 0;
        bool bad = false;
        IHalo h{0, 0, 0, true};
        if constexpr (PAD) {
            h = load_ihalo(a, xin, x, B, ty, 0);
            bad = !h.exact;
        }
        int32_t srow = 0;
#pragma unroll
        for (int yy = 0; yy < B; yy++) {
            const uint64_t y = y0 + yy;
            if (y < a.d1) {
                const uint64_t idx = x + y * a.d0;
                const uint32_t c = xin ? codes[idx] : (uint32_t)a.radius;
                const bool isout = xin && c == 0;
                const int32_t N = yy > 0 ? srow : h.y;
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

template <int B, bool PAD = false>
__global__ void __launch_bounds__(256) k_recon_2d(const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, ReconArgs a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    if (a.guard && *a.guard) return;  // decode output not final: a later launch redoes this
    const uint64_t ntiles = a.ntiles_x * a.ntiles_y;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    for (uint64_t t = wave; t < ntiles; t += nwaves) recon_2d_tile<B, PAD>(t, codes, sparse, out, a, flagged, nflagged);
}

// PAD: the running sum starts from the block's pad value (a raw, usually
// non-integral value for 1D block / edge padding: such blocks are listed).
template <int B, bool PAD>
__device__ __forceinline__ void recon_1d_tile(uint64_t t, const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, const ReconArgs& a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    const uint32_t lane = lane_id2();
    const uint64_t nbx = (a.d0 + B - 1) / B;
    {
        const uint64_t x = t * 64 + lane;
        const bool xin = x < a.d0;
        const bool bstart = (lane % B) == 0;
        bool bad = false;
        IHalo h{0, 0, 0, true};
        if constexpr (PAD) {
            h = load_ihalo(a, xin, x, B, 0, 0);
            bad = !h.exact;
        }
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
            if (PAD && bstart) g0 += h.x;
        }
        const int32_t sv = seg_scan<B>(g0, isout || bstart);
        if (sv >= kExact || sv <= -kExact) bad = true;
        if (xin) out[x] = isout ? scale_outlier(v, a.twoeb) : scale_out(sv, a.twoeb);
        flag_blocks(bad, B, t * (64 / B), 0, 0, nbx, 1, flagged, nflagged);
    }
}

template <int B, bool PAD = false>
__global__ void __launch_bounds__(256) k_recon_1d(const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, ReconArgs a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    if (a.guard && *a.guard) return;  // decode output not final: a later launch redoes this
    const uint64_t ntiles = a.ntiles_x;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    for (uint64_t t = wave; t < ntiles; t += nwaves) recon_1d_tile<B, PAD>(t, codes, sparse, out, a, flagged, nflagged);
}

// One thread per listed block: the reference's sequential fp32 recurrence in
// its own association, using the block's region of `out` for the s values
// before scaling them in place.
// Listed blocks i0, i0 + di, ... of one field.
__device__ __forceinline__ void recon_fallback_blocks(int ndim, int B, const uint16_t* __restrict__ codes,
                                                      const float* __restrict__ sparse, float* __restrict__ out,
                                                      const ReconArgs& a, const uint32_t* __restrict__ flagged,
                                                      const uint32_t* __restrict__ nflagged, uint64_t i0, uint64_t di) {
    const uint32_t nl = *nflagged;
    const uint64_t nbx = (a.d0 + B - 1) / B, nby = (a.d1 + B - 1) / B;
    const uint64_t sy = a.d0, sz = a.d0 * a.d1;
    for (uint64_t i = i0; i < nl; i += di) {
        const uint64_t id = flagged[i];
        const uint64_t bx = id % nbx, by = (id / nbx) % nby, bz = id / (nbx * nby);
        const uint64_t x0 = bx * B, y0 = by * B, z0 = bz * B;
        // halo cells: z face, else y face, else x face in 3D; x face, else y in 2D
        float hz = 0.f, hy = 0.f, hx = 0.f;
        if (a.halo) {
            const float* hp = a.halo + a.halo_stride * (bx + a.hb0 * (by + a.hb1 * bz));
            hz = hp[0];
            hy = hp[1];
            hx = hp[2];
        }
        const uint64_t xe = min(x0 + B, a.d0), ye = min(y0 + B, a.d1), ze = min(z0 + B, a.d2);
        for (uint64_t z = z0; z < ze; z++)
            for (uint64_t y = y0; y < ye; y++)
                for (uint64_t x = x0; x < xe; x++) {
                    const uint64_t idx = x + y * sy + z * sz;
                    const bool ix = x > x0, iy = y > y0, iz = z > z0;  // neighbour inside the block
                    float pred;
                    if (ndim == 1) {
                        pred = ix ? out[idx - 1] : hx;
                    } else if (ndim == 2) {
                        const float W = ix ? out[idx - 1] : hx;
                        const float N = iy ? out[idx - sy] : hy;
                        const float NW = (ix && iy) ? out[idx - 1 - sy] : (!ix ? hx : hy);
                        pred = N + W;
                        pred = pred - NW;
                    } else {
                        const float X = ix ? out[idx - 1] : hx;
                        const float Y = iy ? out[idx - sy] : hy;
                        const float Z = iz ? out[idx - sz] : hz;
                        const float XY = (ix && iy) ? out[idx - 1 - sy] : (!iy ? hy : hx);
                        const float XZ = (ix && iz) ? out[idx - 1 - sz] : (!iz ? hz : hx);
                        const float YZ = (iy && iz) ? out[idx - sy - sz] : (!iz ? hz : hy);
                        const float XYZ = (ix && iy && iz) ? out[idx - 1 - sy - sz] : (!iz ? hz : !iy ? hy : hx);
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

__global__ void __launch_bounds__(64) k_recon_fallback(int ndim, int B, const uint16_t* __restrict__ codes,
                                                       const float* __restrict__ sparse, float* __restrict__ out,
                                                       ReconArgs a, const uint32_t* __restrict__ flagged,
                                                       const uint32_t* __restrict__ nflagged) {
    if (a.guard && *a.guard) return;  // decode output not final: a later launch redoes this
    recon_fallback_blocks(ndim, B, codes, sparse, out, a, flagged, nflagged,
                          blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, (uint64_t)gridDim.x * blockDim.x);
}

// Four consecutive x per lane (8-byte code loads, 16-byte stores).  A block
// row of B elements spans B/4 lanes: each lane runs the Lorenzo prefix over
// its four elements sequentially, then a segmented scan over the lanes of the
// row carries the running value in, up to the lane's first restart (an
// outlier or the block start).
template <int B>
__device__ __forceinline__ void row_scan4(int32_t (&D)[4], const int32_t (&g)[4], const bool (&rs)[4]) {
    constexpr int L = B / 4;
    D[0] = g[0];
#pragma unroll
    for (int i = 1; i < 4; i++) D[i] = rs[i] ? g[i] : D[i - 1] + g[i];
    if constexpr (L > 1) {
        const uint32_t lib = lane_id2() % L;
        int32_t T = D[3];
        int f = rs[0] | rs[1] | rs[2] | rs[3];
#pragma unroll
        for (int d = 1; d < L; d <<= 1) {
            const int32_t t2 = __shfl_up(T, d, L);
            const int f2 = __shfl_up(f, d, L);
            if (lib >= (uint32_t)d) {
                if (!f) T += t2;
                f |= f2;
            }
        }
        int32_t carry = __shfl_up(T, 1, L);
        if (lib == 0) carry = 0;
        bool open = true;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            open = open && !rs[i];
            if (open) D[i] += carry;
        }
    }
}

// ALN: the x extent is a multiple of 4 and codes / output are 8- / 16-byte
// aligned, so a lane's four codes and outputs move as one uint2 / float4;
// otherwise element by element, elements past the row's end reading as the
// radius code (no change) and storing nothing.
template <int ND, int B, bool ALN>
__device__ __forceinline__ void recon_x4_tile(uint64_t t, const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, const ReconArgs& a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    constexpr int L = B / 4;
    constexpr int YB = ND >= 2 ? B : 1;
    constexpr int ZB = ND >= 3 ? B : 1;
    const uint32_t lane = lane_id2();
    const uint64_t ntx = (a.d0 + 255) / 256;
    const uint64_t nty = ND >= 2 ? a.ntiles_y : 1;
    const uint64_t sy = a.d0, sz = a.d0 * a.d1;
    const uint64_t nbx = (a.d0 + B - 1) / B, nby = ND >= 2 ? (a.d1 + B - 1) / B : 1;
    const uint32_t R = (uint32_t)a.radius;
    {
        const uint64_t tx = t % ntx;
        const uint64_t ty = (t / ntx) % nty;
        const uint64_t tz = t / (ntx * nty);
        const uint64_t x0 = tx * 256 + 4 * lane;
        const bool xin = x0 < a.d0;  // ALN: all four or none
        bool ein[4];                 // element x0 + i inside the row
#pragma unroll
        for (int i = 0; i < 4; i++) ein[i] = ALN ? xin : x0 + i < a.d0;
        const bool rowstart = (lane % L) == 0;
        bool bad = false;
        int32_t prev[ZB > 1 ? B : 1][4];
#pragma unroll
        for (int j = 0; j < (ZB > 1 ? B : 1); j++)
#pragma unroll
            for (int i = 0; i < 4; i++) prev[j][i] = 0;
        for (int zz = 0; zz < ZB; zz++) {
            const uint64_t z = ND >= 3 ? tz * B + zz : 0;
            if (z >= a.d2) break;
            int32_t srow[4] = {0, 0, 0, 0}, zrow[4] = {0, 0, 0, 0};
#pragma unroll
            for (int yy = 0; yy < YB; yy++) {
                const uint64_t y = ND >= 2 ? ty * B + yy : 0;
                if (y < a.d1) {
                    const uint64_t idx = x0 + y * sy + z * sz;
                    uint32_t c[4];
                    if constexpr (ALN) {
                        uint2 cw = make_uint2(R | (R << 16), R | (R << 16));
                        if (xin) cw = *reinterpret_cast<const uint2*>(codes + idx);
                        c[0] = cw.x & 0xFFFF;
                        c[1] = cw.x >> 16;
                        c[2] = cw.y & 0xFFFF;
                        c[3] = cw.y >> 16;
                    } else {
#pragma unroll
                        for (int i = 0; i < 4; i++) c[i] = ein[i] ? (uint32_t)codes[idx + i] : R;
                    }
                    int32_t g[4], base[4], Zv[4], D[4];
                    bool rs[4], isout[4];
                    float v[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int32_t Y = (ND >= 2 && yy > 0) ? srow[i] : 0;
                        Zv[i] = (ND >= 3 && zz > 0) ? prev[ZB > 1 ? yy : 0][i] : 0;
                        const int32_t YZ = (ND >= 3 && yy > 0 && zz > 0) ? zrow[i] : 0;
                        base[i] = Y + Zv[i] - YZ;
                        isout[i] = ein[i] && c[i] == 0;
                        v[i] = 0.f;
                        g[i] = (int32_t)c[i] - a.radius;
                    }
                    // outliers are rare: one test for the lane's four elements
                    const bool anyout = isout[0] | isout[1] | isout[2] | isout[3];
                    if (anyout) {
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            if (isout[i]) {
                                v[i] = sparse[idx + i];
                                if (!exact_outlier(v[i])) bad = true;
                                g[i] = (int32_t)v[i] - base[i];
                            }
                        }
                    }
#pragma unroll
                    for (int i = 0; i < 4; i++) rs[i] = isout[i] || (i == 0 && rowstart);
                    row_scan4<B>(D, g, rs);
                    float o[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int32_t sv = D[i] + base[i];
                        if (sv >= kExact || sv <= -kExact) bad = true;
                        o[i] = scale_out(sv, a.twoeb);
                        if (ND >= 3) {
                            zrow[i] = Zv[i];
                            prev[ZB > 1 ? yy : 0][i] = sv;
                        }
                        srow[i] = sv;
                    }
                    if (anyout) {
#pragma unroll
                        for (int i = 0; i < 4; i++)
                            if (isout[i]) o[i] = scale_outlier(v[i], a.twoeb);
                    }
                    if constexpr (ALN) {
                        if (xin) *reinterpret_cast<float4*>(out + idx) = make_float4(o[0], o[1], o[2], o[3]);
                    } else {
#pragma unroll
                        for (int i = 0; i < 4; i++)
                            if (ein[i]) out[idx + i] = o[i];
                    }
                }
            }
        }
        flag_blocks(bad, L, tx * (256 / B), ty, tz, nbx, nby, flagged, nflagged);
    }
}

template <int ND, int B, bool ALN>
__global__ void __launch_bounds__(256) k_recon_x4(const uint16_t* __restrict__ codes, const float* __restrict__ sparse,
                                                  float* __restrict__ out, ReconArgs a, uint32_t* flagged,
                                                  uint32_t* nflagged) {
    if (a.guard && *a.guard) return;  // decode output not final: a later launch redoes this
    const uint64_t ntx = (a.d0 + 255) / 256;
    const uint64_t nty = ND >= 2 ? a.ntiles_y : 1, ntz = ND >= 3 ? a.ntiles_z : 1;
    const uint64_t ntiles = ntx * nty * ntz;
    const uint64_t wave = blockIdx.x * 4ull + (threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    for (uint64_t t = wave; t < ntiles; t += nwaves)
        recon_x4_tile<ND, B, ALN>(t, codes, sparse, out, a, flagged, nflagged);
}

// ------------------------------------------------------------------ batched fields
//
// A batch decompress runs every stage once for all of its fields: each
// launch walks a queue of units that spans the fields (unit u belongs to
// field f when start[f] <= u < start[f + 1]) and runs the single-field body
// on the field's descriptor.  A unit is uniform across its workgroup (sync,
// write) or its wavefront (reconstruction), so the field lookup and the
// descriptor reads are scalar.

template <bool T16>
__global__ void __launch_bounds__(256) k_decode_sync_multi(const DecField* __restrict__ fd,
                                                           const uint64_t* __restrict__ start, int nf) {
    const uint64_t u = blockIdx.x;
    const int f = field_of(start, nf, u);
    const DecodeArgs a = fd[f].a;
    const SubState s = fd[f].s;
    decode_sync_unit<T16>(a, s, fd[f].sum, u - start[f]);
}

template <bool T16>
__global__ void __launch_bounds__(256) k_decode_write_multi(const DecField* __restrict__ fd,
                                                            const uint64_t* __restrict__ start, int nf) {
    const uint64_t u = blockIdx.x;
    const int f = field_of(start, nf, u);
    const DecodeArgs a = fd[f].a;
    const SubState s = fd[f].s;
    decode_write_unit<T16>(a, s, fd[f].sym_off, fd[f].zero_off, fd[f].codes, fd[f].sparse, fd[f].sum, u - start[f]);
}

// The decode tables of a field read from global memory (entries of one
// re-decode batch may belong to different fields).
template <bool T16>
__device__ __forceinline__ Decoder<T16> global_decoder(const DecodeArgs& a);
template <>
__device__ __forceinline__ Decoder<true> global_decoder<true>(const DecodeArgs& a) {
    Decoder<true> d;
    d.lut = a.lut16;
    d.sub = a.sub16;
    return d;
}
template <>
__device__ __forceinline__ Decoder<false> global_decoder<false>(const DecodeArgs& a) {
    Decoder<false> d;
    d.lut = a.lut;
    d.sub = a.sub;
    return d;
}

// Re-decode pass over the batch-wide list: as k_decode_pass, but entry J
// (a batch-wide subsequence number) belongs to field field_of(sub_start, J),
// whose stream its row is staged from and whose tables decode it.
template <bool T16>
__global__ void __launch_bounds__(256) k_decode_pass_multi(const DecField* __restrict__ fd,
                                                           const uint64_t* __restrict__ sub_start, int nf, SubState S,
                                                           uint32_t pass) {
    const uint32_t m = S.list_n[pass % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) S.list_n[(pass + 2) % 3] = 0;  // the list after next
    if ((uint64_t)blockIdx.x * 256 >= m) return;                            // whole workgroup
    __shared__ uint32_t rows[256 * kRowStride];
    const uint32_t* lj = S.list_j[pass & 1];
    const uint64_t* ls = S.list_start[pass & 1];
    uint32_t* oj = S.list_j[(pass + 1) & 1];
    uint64_t* os = S.list_start[(pass + 1) & 1];
    uint32_t* on = &S.list_n[(pass + 1) % 3];
    for (uint64_t e0 = blockIdx.x * 256ull; e0 < m; e0 += (uint64_t)gridDim.x * 256) {
        const uint32_t ne = (uint32_t)min<uint64_t>(256, m - e0);
        if (threadIdx.x < ne) {
            const uint32_t J = lj[e0 + threadIdx.x];
            const int f = field_of(sub_start, nf, J);
            const DecodeArgs& a = fd[f].a;
            const SubState& s = fd[f].s;
            const Decoder<T16> dec = global_decoder<T16>(a);
            StreamView sv;
            sv.setup(a.stream, a.stream_bits);
            const uint64_t j = J - sub_start[f];
            const uint64_t start = ls[e0 + threadIdx.x];
            stage_row(rows + threadIdx.x * kRowStride, sv, j);  // the thread's own row: no barrier
            const uint64_t rowbit = j * kSubBits;  // aligned-space bit of word 0
            const RowWin win{rows + threadIdx.x * kRowStride};
            const uint64_t target = min((j + 1) * (uint64_t)kSubBits, a.stream_bits);
            const uint32_t tl = (uint32_t)(target + sv.shift0 - rowbit);
            uint32_t lp = (uint32_t)(start + sv.shift0 - rowbit);
            uint32_t cnt = 0, z = 0, minl = ~0u;
            while (lp < tl) {
                uint32_t len;
                const uint32_t sym = dec.take_w(win, lp, len);
                minl = min(minl, len);
                lp += max(len, 1u);
                cnt++;
                z += sym == 0;
            }
            const bool bad = minl == 0;
            const uint64_t pos = rowbit + lp - sv.shift0;
            s.b[j] = start;
            s.cnt[j] = cnt;
            s.zer[j] = z;
            if (pos != s.e[j]) {
                s.e[j] = pos;
                if (j + 1 < a.nsub) {
                    const uint32_t k = atomicAdd(on, 1u);
                    oj[k] = J + 1;
                    os[k] = pos;
                }
            }
            if (bad) atomicOr(&fd[f].sum->error, 1u);
        }
    }
}

__global__ void k_dec_field_totals(const DecField* __restrict__ fd, int nf, uint64_t* __restrict__ tot) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const uint64_t ns = fd[f].a.nsub;
    tot[f] = ns ? fd[f].sym_off[ns - 1] + fd[f].s.cnt[ns - 1] - fd[f].sym_off[0] : 0;
}

template <int KIND, int B>
__global__ void __launch_bounds__(256) k_recon_multi(const DecField* __restrict__ fd,
                                                     const uint64_t* __restrict__ start, int nf, uint64_t ntot) {
    const uint64_t wave = blockIdx.x * 4ull + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t nwaves = gridDim.x * 4ull;
    for (uint64_t T = wave; T < ntot; T += nwaves) {
        const int f = field_of(start, nf, T);
        const uint64_t t = T - start[f];
        const ReconArgs ra = fd[f].ra;
        const uint16_t* codes = fd[f].codes;
        const float* sparse = fd[f].sparse;
        float* out = fd[f].out;
        uint32_t* fl = fd[f].flagged;
        uint32_t* nfl = fd[f].nflagged;
        if constexpr (KIND == kRc1x4)
            recon_x4_tile<1, B, true>(t, codes, sparse, out, ra, fl, nfl);
        else if constexpr (KIND == kRc2x4)
            recon_x4_tile<2, B, true>(t, codes, sparse, out, ra, fl, nfl);
        else if constexpr (KIND == kRc3x4)
            recon_x4_tile<3, B, true>(t, codes, sparse, out, ra, fl, nfl);
        else if constexpr (KIND == kRc1x4u)
            recon_x4_tile<1, B, false>(t, codes, sparse, out, ra, fl, nfl);
        else if constexpr (KIND == kRc2x4u)
            recon_x4_tile<2, B, false>(t, codes, sparse, out, ra, fl, nfl);
        else if constexpr (KIND == kRc3x4u)
            recon_x4_tile<3, B, false>(t, codes, sparse, out, ra, fl, nfl);
        else if constexpr (KIND == kRc3)
            recon_3d_tile<B, false>(t, codes, sparse, out, ra, fl, nfl);
        else if constexpr (KIND == kRc2)
            recon_2d_tile<B, false>(t, codes, sparse, out, ra, fl, nfl);
        else
            recon_1d_tile<B, false>(t, codes, sparse, out, ra, fl, nfl);
    }
}

__global__ void __launch_bounds__(256) k_gather_heads(const HeadCopy* __restrict__ hc, uint8_t* __restrict__ dst) {
    const HeadCopy h = hc[blockIdx.x];
    for (uint64_t i = threadIdx.x; i < h.len; i += blockDim.x) dst[h.off + i] = h.src[i];
}

// blockIdx.y: the field.
__global__ void __launch_bounds__(64) k_recon_fallback_multi(int B, const DecField* __restrict__ fd) {
    const DecField& F = fd[blockIdx.y];
    if (F.kind < 0) return;
    recon_fallback_blocks(F.ndim, B, F.codes, F.sparse, F.out, F.ra, F.flagged, F.nflagged,
                          blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, (uint64_t)gridDim.x * blockDim.x);
}

// ------------------------------------------------------------------ launchers

static unsigned grid_for(uint64_t threads, unsigned block) { return (unsigned)((threads + block - 1) / block); }

// Dynamic LDS bytes of the decode tables a kernel keeps in LDS.
static size_t table_lds_bytes(const DecodeArgs& a, bool t16) {
    return t16 ? (((size_t)(1u << kLutBits) + a.n_sub16 + 1) / 2) * 4 : (size_t)(1u << kLutBits) * 4;
}

// The deeper 16-bit tables serve only the codewords over kLutBits bits (a
// few percent of them at eb 1e-4), so a kernel copies them to LDS only when
// that costs no resident workgroup per CU, given its other LDS (own_bytes
// besides the first-level table); otherwise it reads them in global memory
// (n_sub16 0).  Measured on the 1 GiB field, whose deeper tables took 22 KB:
// the write pass 0.75 -> 0.53 ms, the sync pass 1.08 -> 1.04 ms in global
// memory.
static DecodeArgs with_lds_sub16(DecodeArgs a, size_t own_bytes) {
    if (!a.lut16 || !a.n_sub16) return a;
    constexpr size_t kCuLds = 160 * 1024, kGranule = 1024;  // (workgroup LDS counted in whole KiB, conservatively)
    DecodeArgs b = a;
    b.n_sub16 = 0;
    auto granules = [](size_t bytes) { return (bytes + kGranule - 1) / kGranule * kGranule; };
    const size_t without = granules(table_lds_bytes(b, true) + own_bytes);
    const size_t with = granules(table_lds_bytes(a, true) + own_bytes);
    if (kCuLds / with < kCuLds / without) a.n_sub16 = 0;
    return a;
}

// Lets a kernel take more than the default 64 KiB of dynamic LDS (gfx950
// has 160 KiB per workgroup; the kernels here have no static LDS).  Set once
// per kernel: each call site keeps its own once-flag (kernels of one
// signature share a function type, so a flag per type would skip all but
// the first of them).
static bool set_big_lds(const void* kernel) {
    if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
        (void)hipGetLastError();  // the launch then reports what does not fit
    return true;
}
#define allow_big_lds(K)                                                                 \
    do {                                                                                 \
        static const bool once_ = set_big_lds(reinterpret_cast<const void*>(&(K)));      \
        (void)once_;                                                                     \
    } while (0)

void launch_decode_sync(const DecodeArgs& a0, SubState s, DecSummary* sum, hipStream_t st) {
    allow_big_lds(k_decode_sync<true>);
    allow_big_lds(k_decode_sync<false>);
    const bool t16 = a0.lut16 != nullptr;
    const DecodeArgs a = with_lds_sub16(a0, (size_t)kSyncLdsWords * 4);
    const size_t lds = table_lds_bytes(a, t16) + (size_t)kSyncLdsWords * 4;
    const dim3 grid(grid_for(a.nsub, kSyncStride));
    if (t16)
        hipLaunchKernelGGL(k_decode_sync<true>, grid, dim3(kSyncWg), lds, st, a, s, sum);
    else
        hipLaunchKernelGGL(k_decode_sync<false>, grid, dim3(kSyncWg), lds, st, a, s, sum);
}
void launch_decode_pass(const DecodeArgs& a0, SubState s, uint32_t pass, DecSummary* sum, hipStream_t st) {
    allow_big_lds(k_decode_pass<true>);
    allow_big_lds(k_decode_pass<false>);
    const bool t16 = a0.lut16 != nullptr;
    const DecodeArgs a = with_lds_sub16(a0, (size_t)256 * kRowStride * 4);
    const size_t lds = table_lds_bytes(a, t16) + (size_t)256 * kRowStride * 4;
    // persistent workgroups; those past the listed count return at once
    const dim3 grid((unsigned)std::min<uint64_t>(kPassGrid, std::max<uint64_t>(1, grid_for(a.nsub, 256))));
    if (t16)
        hipLaunchKernelGGL(k_decode_pass<true>, grid, dim3(256), lds, st, a, s, pass, sum);
    else
        hipLaunchKernelGGL(k_decode_pass<false>, grid, dim3(256), lds, st, a, s, pass, sum);
}
void launch_decode_indexed(const DecodeArgs& a0, const uint64_t* index, uint64_t total_bits, uint16_t* codes,
                           float* sparse, const Recon1D* r1d, DecSummary* sum, hipStream_t st) {
    allow_big_lds((k_decode_indexed<true, false>));
    allow_big_lds((k_decode_indexed<false, false>));
    allow_big_lds((k_decode_indexed<true, true>));
    allow_big_lds((k_decode_indexed<false, true>));
    allow_big_lds((k_decode_indexed_staged<true, false>));
    allow_big_lds((k_decode_indexed_staged<false, false>));
    allow_big_lds((k_decode_indexed_staged<true, true>));
    allow_big_lds((k_decode_indexed_staged<false, true>));
    const bool t16 = a0.lut16 != nullptr;
    const DecodeArgs a = with_lds_sub16(a0, (size_t)4 * kIdxStageWords * 4);
    const size_t lds = table_lds_bytes(a, t16) + (size_t)4 * kIdxStageWords * 4;
    const IdxLayout L = idx_layout(a.n);
    const dim3 grid((unsigned)std::max<uint64_t>(1, (L.nseg + 255) / 256));
    const Recon1D r = r1d ? *r1d : Recon1D{};
    // workgroups whose span fits the LDS stage first, then the others through
    // the register-window kernel (which skips the staged ones)
    {
        const DecodeArgs as = with_lds_sub16(a0, (size_t)(kIdxSpanWords + kIdxSpanOver) * 4);
        const size_t lds_s = table_lds_bytes(as, t16) + (size_t)(kIdxSpanWords + kIdxSpanOver) * 4;
#define DS(T, R)                                                                                                    \
    hipLaunchKernelGGL((k_decode_indexed_staged<T, R>), grid, dim3(256), lds_s, st, as, index, L, total_bits, codes, \
                       sparse, r, sum)
        if (r1d) {
            if (t16)
                DS(true, true);
            else
                DS(false, true);
        } else {
            if (t16)
                DS(true, false);
            else
                DS(false, false);
        }
#undef DS
    }
#define DI(T, R) \
    hipLaunchKernelGGL((k_decode_indexed<T, R>), grid, dim3(256), lds, st, a, index, L, total_bits, codes, sparse, r, sum)
    if (r1d) {
        if (t16)
            DI(true, true);
        else
            DI(false, true);
    } else {
        if (t16)
            DI(true, false);
        else
            DI(false, false);
    }
#undef DI
}
void launch_scan_u32x2(const uint32_t* a, const uint32_t* b, uint64_t n, uint64_t* oa, uint64_t* ob, uint64_t* block_tmp,
                       uint64_t* totals, hipStream_t st, const uint32_t* guard) {
    const uint64_t nb = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(1024), 0, st, a, b, n, oa, ob, block_tmp, guard);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, block_tmp, nb, totals, guard);
    hipLaunchKernelGGL(k_scan_add, dim3(grid_for(n, 1024)), dim3(1024), 0, st, oa, ob, n, block_tmp, guard);
}
void launch_decode_write(const DecodeArgs& a0, SubState s, const uint64_t* sym_off, const uint64_t* zero_off,
                         uint16_t* codes, float* sparse, DecSummary* sum, hipStream_t st,
                         const uint32_t* guard) {
    allow_big_lds(k_decode_write<true>);
    allow_big_lds(k_decode_write<false>);
    const bool t16 = a0.lut16 != nullptr;
    const DecodeArgs a = with_lds_sub16(a0, (size_t)kWriteStageLds * 4);
    const size_t lds = table_lds_bytes(a, t16) + (size_t)kWriteStageLds * 4;
    const dim3 grid((unsigned)decode_write_units(a.nsub));
    if (t16)
        hipLaunchKernelGGL(k_decode_write<true>, grid, dim3(kSyncWg), lds, st, a, s, sym_off, zero_off, codes, sparse,
                           sum, guard);
    else
        hipLaunchKernelGGL(k_decode_write<false>, grid, dim3(kSyncWg), lds, st, a, s, sym_off, zero_off, codes, sparse,
                           sum, guard);
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
    // four x per lane when rows are whole multiples of 4 and the output is
    // 16-byte aligned (3D keeps B x 4 running values per lane: B <= 16)
    if (a.halo) {
#define RC_PAD(ND, KERN, BB)                                                                                      \
    if (ndim == ND && B == BB) {                                                                                  \
        hipLaunchKernelGGL((KERN<BB, true>), dim3(grid), dim3(256), 0, st, codes, sparse, out, a, flagged, nflagged); \
        return;                                                                                                   \
    }
        RC_PAD(3, k_recon_3d, 8)
        RC_PAD(3, k_recon_3d, 16)
        RC_PAD(3, k_recon_3d, 32)
        RC_PAD(3, k_recon_3d, 64)
        RC_PAD(2, k_recon_2d, 8)
        RC_PAD(2, k_recon_2d, 16)
        RC_PAD(2, k_recon_2d, 32)
        RC_PAD(2, k_recon_2d, 64)
        RC_PAD(1, k_recon_1d, 8)
        RC_PAD(1, k_recon_1d, 16)
        RC_PAD(1, k_recon_1d, 32)
        RC_PAD(1, k_recon_1d, 64)
#undef RC_PAD
        return;
    }
    const int kind = recon_kind(ndim, B, a.d0, codes, out);
    if (kind != kRc1 && kind != kRc2 && kind != kRc3) {
        const uint64_t work = recon_tiles(kind, a);
        const dim3 g4((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((work + 3) / 4, 2048)));
        const bool aln = kind == kRc1x4 || kind == kRc2x4 || kind == kRc3x4;
#define RX4(ND, BB)                                                                                              \
    if (ndim == ND && B == BB) {                                                                                 \
        if (aln)                                                                                                 \
            hipLaunchKernelGGL((k_recon_x4<ND, BB, true>), g4, dim3(256), 0, st, codes, sparse, out, a, flagged,  \
                               nflagged);                                                                        \
        else                                                                                                     \
            hipLaunchKernelGGL((k_recon_x4<ND, BB, false>), g4, dim3(256), 0, st, codes, sparse, out, a, flagged, \
                               nflagged);                                                                        \
        return;                                                                                                  \
    }
        RX4(3, 8)
        RX4(3, 16)
        RX4(2, 8)
        RX4(2, 16)
        RX4(2, 32)
        RX4(2, 64)
        RX4(1, 8)
        RX4(1, 16)
        RX4(1, 32)
        RX4(1, 64)
#undef RX4
    }
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

uint32_t batch_sub16_in_lds(uint32_t max_sub16) {
    DecodeArgs probe{};
    probe.lut16 = reinterpret_cast<const uint16_t*>(&probe);  // (only tested for null)
    probe.n_sub16 = max_sub16;
    return with_lds_sub16(probe, (size_t)kSyncLdsWords * 4).n_sub16;
}

void launch_decode_sync_multi(bool t16, uint32_t max_sub16, const DecField* fd, const uint64_t* start, int nf,
                              uint64_t nunits, hipStream_t st) {
    allow_big_lds(k_decode_sync_multi<true>);
    allow_big_lds(k_decode_sync_multi<false>);
    if (!nunits) return;
    DecodeArgs probe{};
    probe.n_sub16 = max_sub16;
    const size_t lds = table_lds_bytes(probe, t16) + (size_t)kSyncLdsWords * 4;
    if (t16)
        hipLaunchKernelGGL(k_decode_sync_multi<true>, dim3((unsigned)nunits), dim3(kSyncWg), lds, st, fd, start, nf);
    else
        hipLaunchKernelGGL(k_decode_sync_multi<false>, dim3((unsigned)nunits), dim3(kSyncWg), lds, st, fd, start, nf);
}
void launch_decode_pass_multi(bool t16, const DecField* fd, const uint64_t* sub_start, int nf, SubState S,
                              uint32_t pass, uint64_t nsub_tot, hipStream_t st) {
    // persistent workgroups; those past the listed count return at once
    const dim3 grid((unsigned)std::min<uint64_t>(kPassGrid, std::max<uint64_t>(1, grid_for(nsub_tot, 256))));
    if (t16)
        hipLaunchKernelGGL(k_decode_pass_multi<true>, grid, dim3(256), 0, st, fd, sub_start, nf, S, pass);
    else
        hipLaunchKernelGGL(k_decode_pass_multi<false>, grid, dim3(256), 0, st, fd, sub_start, nf, S, pass);
}
void launch_decode_write_multi(bool t16, uint32_t max_sub16, const DecField* fd, const uint64_t* start, int nf,
                               uint64_t nunits, hipStream_t st) {
    allow_big_lds(k_decode_write_multi<true>);
    allow_big_lds(k_decode_write_multi<false>);
    if (!nunits) return;
    DecodeArgs probe{};
    probe.n_sub16 = max_sub16;
    const size_t lds = table_lds_bytes(probe, t16) + (size_t)kWriteStageLds * 4;
    if (t16)
        hipLaunchKernelGGL(k_decode_write_multi<true>, dim3((unsigned)nunits), dim3(kSyncWg), lds, st, fd, start, nf);
    else
        hipLaunchKernelGGL(k_decode_write_multi<false>, dim3((unsigned)nunits), dim3(kSyncWg), lds, st, fd, start, nf);
}
void launch_dec_field_totals(const DecField* fd, int nf, uint64_t* tot, hipStream_t st) {
    if (nf > 0) hipLaunchKernelGGL(k_dec_field_totals, dim3(grid_for((uint64_t)nf, 256)), dim3(256), 0, st, fd, nf, tot);
}
int recon_kind_ndim(int kind) {
    switch (kind) {
        case kRc1: case kRc1x4: case kRc1x4u: return 1;
        case kRc2: case kRc2x4: case kRc2x4u: return 2;
        default: return 3;
    }
}
int recon_kind(int ndim, int B, uint64_t d0, const void* codes, const void* out) {
    // four x per lane unless a 3D tile's B x 4 running values per lane would
    // not fit (B > 16); vector loads and stores when rows are whole multiples
    // of 4 and codes / output are 8- / 16-byte aligned
    if (ndim == 3 && B > 16) return kRc3;
    const bool aln = d0 % 4 == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(codes) & 7) == 0;
    if (ndim == 1) return aln ? kRc1x4 : kRc1x4u;
    if (ndim == 2) return aln ? kRc2x4 : kRc2x4u;
    return aln ? kRc3x4 : kRc3x4u;
}
uint64_t recon_tiles(int kind, const ReconArgs& a) {
    const bool x4 = kind != kRc1 && kind != kRc2 && kind != kRc3;
    const int nd = recon_kind_ndim(kind);
    const uint64_t tx = x4 ? (a.d0 + 255) / 256 : a.ntiles_x;
    return tx * (nd >= 2 ? a.ntiles_y : 1) * (nd >= 3 ? a.ntiles_z : 1);
}
void launch_reconstruct_multi(int kind, int B, const DecField* fd, const uint64_t* start, int nf, uint64_t ntiles,
                              hipStream_t st) {
    if (!ntiles) return;
    const dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((ntiles + 3) / 4, 2048)));
#define RM(K, BB)                                                                                \
    if (kind == K && B == BB) {                                                                  \
        hipLaunchKernelGGL((k_recon_multi<K, BB>), grid, dim3(256), 0, st, fd, start, nf, ntiles); \
        return;                                                                                  \
    }
    RM(kRc3x4, 8)
    RM(kRc3x4, 16)
    RM(kRc2x4, 8)
    RM(kRc2x4, 16)
    RM(kRc2x4, 32)
    RM(kRc2x4, 64)
    RM(kRc1x4, 8)
    RM(kRc1x4, 16)
    RM(kRc1x4, 32)
    RM(kRc1x4, 64)
    RM(kRc3, 8)
    RM(kRc3, 16)
    RM(kRc3, 32)
    RM(kRc3, 64)
    RM(kRc2, 8)
    RM(kRc2, 16)
    RM(kRc2, 32)
    RM(kRc2, 64)
    RM(kRc1, 8)
    RM(kRc1, 16)
    RM(kRc1, 32)
    RM(kRc1, 64)
    RM(kRc3x4u, 8)
    RM(kRc3x4u, 16)
    RM(kRc2x4u, 8)
    RM(kRc2x4u, 16)
    RM(kRc2x4u, 32)
    RM(kRc2x4u, 64)
    RM(kRc1x4u, 8)
    RM(kRc1x4u, 16)
    RM(kRc1x4u, 32)
    RM(kRc1x4u, 64)
#undef RM
}
uint64_t decode_sync_units(uint64_t nsub) { return (nsub + kSyncStride - 1) / kSyncStride; }
uint64_t decode_write_units(uint64_t nsub) { return (nsub + kWriteWg - 1) / kWriteWg; }
void launch_gather_heads(const HeadCopy* hc, int n, uint8_t* dst, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_gather_heads, dim3((unsigned)n), dim3(256), 0, st, hc, dst);
}
void launch_reconstruct_fallback_multi(int B, const DecField* fd, int nf, hipStream_t st) {
    if (nf > 0) hipLaunchKernelGGL(k_recon_fallback_multi, dim3(16, (unsigned)nf), dim3(64), 0, st, B, fd);
}

}  // namespace vsz
