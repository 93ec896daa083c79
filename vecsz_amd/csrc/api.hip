// C ABI and host orchestration of the compress / decompress pipelines.
//
// Compress (vecsz::interface::Compress, src/vecsz_interface.hh:31-229):
//   [r2r] k_range -> k_ebfam (eb family on the device, src/types.cc:84-93)
//   [padding] pad values and halos
//   k_dualquant -> histogram to host -> codebook + tree bytes (codebook.cpp)
//   header + tree H2D -> k_chunk_bits -> chunk scan -> k_enc_summary -> k_pack
//   Batches: the same stages once for all fields (k_*_multi launches).
// Decompress (vecsz::interface::Decompress, src/vecsz_interface.hh:232-375):
//   header + tree D2H -> decode tables -> k_decode_sync -> k_decode_pass
//   (re-decode of listed subsequences, until no list remains) -> scan
//   -> k_decode_write -> k_recon_* [-> fallback]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <initializer_list>
#include <mutex>
#include <unordered_map>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vecsz_hip.h"
#include "codebook.hpp"
#include "common.hpp"
#include "kernels.hpp"

using namespace vsz;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    // Grows by at least half again (and to 64 KiB at least) so that buffers
    // sized by varying quantities (trees, tables) settle after a few calls:
    // hipFree synchronises the whole device, which would stall other streams.
    bool ensure(size_t bytes) {
        if (bytes <= cap && p) return true;
        const size_t want = std::max<size_t>({bytes, cap + cap / 2, (size_t)64 << 10});
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return false;
        }
        cap = want;
        return true;
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

struct HostBuf {
    void* p = nullptr;
    size_t cap = 0;
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
    bool ensure(size_t bytes) {
        if (bytes <= cap && p) return true;
        const size_t want = std::max<size_t>({bytes, 2 * cap, (size_t)1 << 20});
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return false;
        }
        cap = want;
        return true;
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

struct Stage {
    const char* name;
    hipEvent_t a, b;
};

}  // namespace

// A captured sequence of stream work and the arguments it was captured with
// (key); seen: the last key that ran eagerly.
struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> key, seen;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
        seen.clear();
    }
};

struct vecsz_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string last_error;
    // batch workers: each a context of its own (stream + scratch), created on first use
    std::vector<vecsz_ctx*> workers;
    int nworkers = 8;
    bool fused_batch = true;  // batches as one launch per stage when they qualify
    bool fused_batch_always = false;  // decompress batches of large fields fused as well
    bool fused_encode = false; // single-field encode as one pass (k_encode): measured slower, off by default
    bool device_codebook = false; // build the codebook on the device (no stop; measured slower, off by default)
    // launch shapes (workgroup caps of the grid-stride compress kernels), set
    // by vecsz_autotune; the output does not depend on them
    int dq_grid = 768, enc_grid = 2048;
    // compress scratch
    // small: histogram, eb family, codebook tables, error word, summary (small_layout)
    DevBuf codes, sparse, small, partials, chunk_bits, chunk_zeros, bit_off, out_off, range;
    DevBuf enc_status, outl;  // fused encode: partition status words + counter, compact outliers
    DevBuf cbinfo;            // device-built codebook: node count, tree bytes, longest codeword, status
    DevBuf padv, halo, pad_scratch;  // padding: pad values, per-block halo values, reduction scratch
    // fused batch compress: field descriptors, unit starts, per-field histograms,
    // range words, eb families, summaries, status words, codebooks, headers
    DevBuf b_desc, b_start, b_hist, b_range, b_ebd, b_sums, b_errs, b_book, b_head, b_rows;
    HostBuf b_pin, b_pin_hist;
    std::vector<Codebook> b_cb;
    // the fused compress batch's work before and after its codebook stop,
    // each captured as one HIP graph when a batch layout repeats
    bool batch_graph = true;
    GraphSlot b_graph[2];
    // fused batch decompress: field descriptors, unit starts, head copies,
    // gathered heads, summaries, symbol totals, fallback counts; pinned
    // staging of the same; decode tables per field
    DevBuf db_desc, db_start, db_hc, db_heads, db_sums, db_tot, db_nfl;
    HostBuf db_pin, db_hpin, db_back;
    std::vector<DecodeTables> db_dt;
    Codebook cb;
    DecodeTables dt;  // kept between calls: its vectors keep their capacity
    // metadata of the last successful compress
    vecsz_sidecar sidecar{};
    bool have_sidecar = false;
    double sc_range = 0, sc_eb_final = 0;
    // verify partials
    DevBuf vparts;
    HostBuf vpin;
    // decompress scratch
    // dec_small: decoder summary, symbol and zero totals, list counts (kDecSmall*)
    DevBuf tabs, sub_b, sub_e, sub_cnt, sub_zeros, sub_lj, sub_ls, sub_ln, sym_off, zero_off, scan_tmp, scan_tot,
        dec_small, flagged, nflagged;
    // host staging
    DevBuf stage_in, stage_out;
    HostBuf pin, pin2, pin_tab;
    std::vector<uint64_t> hist_host;
    // timing
    bool timing = false;
    std::vector<Stage> stages;
    std::vector<hipEvent_t> event_pool;
    size_t event_next = 0;
    std::vector<std::string> tnames;
    std::vector<double> tms;
    std::vector<std::string> host_names;
    std::vector<double> host_ms;
};

namespace {

int fail(vecsz_ctx* c, int code, const std::string& msg) {
    if (c) c->last_error = msg;
    return code;
}
// (a literal message is copied without building a std::string: with the
// capacity a context reserves at creation, reporting it allocates nothing)
int fail(vecsz_ctx* c, int code, const char* msg) {
    if (c) c->last_error.assign(msg);
    return code;
}

#define HIPCHK(ctx, expr)                                                                       \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            return fail(ctx, VECSZ_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));    \
    } while (0)

hipEvent_t take_event(vecsz_ctx* c) {
    if (c->event_next == c->event_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->event_pool.push_back(e);
    }
    return c->event_pool[c->event_next++];
}

struct StageScope {
    vecsz_ctx* c;
    const char* name;
    hipEvent_t a = nullptr;
    StageScope(vecsz_ctx* ctx, const char* n) : c(ctx), name(n) {
        if (c->timing) {
            a = take_event(c);
            if (a) (void)hipEventRecord(a, c->stream);
        }
    }
    ~StageScope() {
        if (c->timing && a) {
            hipEvent_t b = take_event(c);
            if (b) {
                (void)hipEventRecord(b, c->stream);
                c->stages.push_back({name, a, b});
            }
        }
    }
};

void timing_begin(vecsz_ctx* c) {
    c->stages.clear();
    c->event_next = 0;
    c->host_names.clear();
    c->host_ms.clear();
}

void timing_end(vecsz_ctx* c) {
    if (!c->timing) return;
    c->tnames.clear();
    c->tms.clear();
    (void)hipStreamSynchronize(c->stream);
    for (auto& s : c->stages) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, s.a, s.b);
        c->tnames.push_back(s.name);
        c->tms.push_back(ms);
    }
    for (size_t i = 0; i < c->host_names.size(); i++) {
        c->tnames.push_back(c->host_names[i]);
        c->tms.push_back(c->host_ms[i]);
    }
}

// |e| < thr  <=>  sqrtf(e*e) < R for every integral float e: the AVX
// quantizability test of src/dualquant.hh:170-171.
int vec_threshold(int R) {
    const float Rf = (float)R;
    for (int k = std::max(0, R - 64);; k++) {
        volatile float e = (float)k;
        float sq = e * e;
        if (!(std::sqrt(sq) < Rf)) return k;
    }
}

bool valid_block(int B) { return B == 8 || B == 16 || B == 32 || B == 64; }

// Largest field the library accepts: 2^37 elements (512 GiB of f32, beyond
// any one GPU's memory), so that every per-element size it computes (at most
// 16 bytes per element) stays far below 2^64.
constexpr uint64_t kMaxElems = 1ull << 37;
// Decompress batches whose fields average at most this many elements run
// fused by default (a 16 MiB f32 field; larger ones take the worker contexts).
constexpr uint64_t kDBatchFuseElems = 4ull << 20;

// Element count of an ndim field with the given extents (x first), each in
// [1, 2^31); false when an extent is out of range or the product exceeds
// kMaxElems (it cannot wrap: every partial product is checked).
bool field_elems(int ndim, const uint64_t* dims, uint64_t* n) {
    if (ndim < 1 || ndim > 3) return false;
    uint64_t m = 1;
    for (int i = 0; i < ndim; i++) {
        const uint64_t d = dims[i];
        if (d == 0 || d > 0x7FFFFFFFull) return false;
        if (__builtin_mul_overflow(m, d, &m) || m > kMaxElems) return false;
    }
    *n = m;
    return true;
}

int check_params(vecsz_ctx* c, const vecsz_params* p, bool need_dims) {
    if (!p) return fail(c, VECSZ_EINVAL, "params is NULL");
    if (!valid_block(p->block)) return fail(c, VECSZ_EINVAL, "block must be 8, 16, 32 or 64");
    if (p->dict_size < 2 || p->dict_size > 65536 || (p->dict_size & 1))
        return fail(c, VECSZ_EINVAL, "dict_size must be even and in [2, 65536]");
    if (p->mode != VECSZ_MODE_ABS && p->mode != VECSZ_MODE_R2R) return fail(c, VECSZ_EINVAL, "mode");
    if (need_dims) {
        if (p->ndim < 1 || p->ndim > 3) return fail(c, VECSZ_EINVAL, "ndim must be 1, 2 or 3");
        for (int i = 0; i < 3; i++) {
            uint64_t d = i < p->ndim ? p->dims[i] : 1;
            if (d == 0 || d > 0x7FFFFFFFull) return fail(c, VECSZ_EINVAL, "dims must be in [1, 2^31)");
        }
        uint64_t n;
        if (!field_elems(p->ndim, p->dims, &n)) return fail(c, VECSZ_EINVAL, "field has more than 2^37 elements");
    }
    int R = p->dict_size / 2;
    if (vec_threshold(R) != R) return fail(c, VECSZ_EINVAL, "dict_size not supported (AVX test differs from |e|<R)");
    if (p->pad_mode < VECSZ_PAD_NONE || p->pad_mode > VECSZ_PAD_EDGE) return fail(c, VECSZ_EINVAL, "pad_mode");
    if (p->pad_type < VECSZ_PADV_MAX || p->pad_type > VECSZ_PADV_CONST) return fail(c, VECSZ_EINVAL, "pad_type");
    return VECSZ_OK;
}

uint64_t meta_bytes(int ndim) { return 4ull * (ndim + 1) + 8 + 24; }

// Blocks of the reference's grid and the pad values a compress stores:
// 1 (global), one per block (block) or ndim per block (edge)
// (src/vecsz_interface.hh:73-90, src/dimensions.hh:31).
uint64_t ref_blocks(int ndim, const uint64_t* dims, int B) {
    uint64_t nb = 1;
    for (int i = 0; i < ndim; i++) nb *= ref_nblk(dims[i], B);
    return nb;
}
uint64_t pad_len_for(int ndim, const uint64_t* dims, int B, int mode) {
    const uint64_t nb = ref_blocks(ndim, dims, B);
    return mode == VECSZ_PAD_NONE ? 0 : mode == VECSZ_PAD_GLOBAL ? 1 : mode == VECSZ_PAD_BLOCK ? nb : (uint64_t)ndim * nb;
}

struct EbFamily {
    double eb_final, r_d, twoeb;
    float r_f;
};

EbFamily eb_family(double eb, bool r2r, double range) {
    EbFamily f;
    f.eb_final = eb;
    if (r2r) f.eb_final *= range;
    f.r_d = 1 / (2 * f.eb_final);
    f.twoeb = 2 * f.eb_final;
    f.r_f = (float)f.r_d;
    return f;
}

void put_u32be(uint8_t* b, uint32_t v) {
    b[0] = (uint8_t)(v >> 24);
    b[1] = (uint8_t)(v >> 16);
    b[2] = (uint8_t)(v >> 8);
    b[3] = (uint8_t)v;
}

struct HostTimer {
    vecsz_ctx* c;
    const char* name;
    std::chrono::steady_clock::time_point t0;
    HostTimer(vecsz_ctx* ctx, const char* n) : c(ctx), name(n), t0(std::chrono::steady_clock::now()) {}
    ~HostTimer() {
        if (!c->timing) return;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        c->host_names.push_back(name);
        c->host_ms.push_back(ms);
    }
};

// Device scratch of one compress: the context's buffers (grown on demand) or
// slices of the allocation-free entry point's static array.
struct CScratch {
    uint16_t* codes = nullptr;
    float* sparse = nullptr;
    unsigned long long* hist = nullptr;
    uint32_t* partials = nullptr;
    uint32_t *chunk_bits = nullptr, *chunk_zeros = nullptr;
    uint64_t *bit_off = nullptr, *out_off = nullptr, *scan_tmp = nullptr, *scan_tot = nullptr;
    EncSummary* summary = nullptr;
    uint32_t* err = nullptr;
    uint32_t* range = nullptr;
    EbDev* ebdev = nullptr;
    uint32_t* book = nullptr;
    uint64_t* enc_status = nullptr;  // fused encode (null: that stage takes the three-kernel path)
    uint32_t* outl = nullptr;
    uint64_t outl_cap = 0;           // outliers outl holds
};

// The small per-call scratch of a compress in one block, so that each host
// round trip moves one copy: [histogram | eb family] goes to the host at the
// codebook stop, [codebook tables | error word] comes back after it (the
// error word's zeros travel with the tables), and [error word | summary]
// goes to the host at the end.
struct SmallLayout {
    size_t hist, ebdev, book, err, summary, bytes;
};
constexpr size_t align_up(size_t b, size_t a) { return (b + a - 1) / a * a; }
constexpr SmallLayout small_layout(int dict) {
    SmallLayout L{};
    L.hist = 0;
    L.ebdev = (size_t)dict * 8;
    L.book = align_up(L.ebdev + sizeof(EbDev), 256);
    L.err = L.book + align_up((size_t)dict * 9, 16);
    L.summary = L.err + 16;
    L.bytes = L.summary + align_up(sizeof(EncSummary), 16) + 64;
    return L;
}
CScratch small_scratch(CScratch s, uint8_t* block, int dict) {
    const SmallLayout L = small_layout(dict);
    s.hist = reinterpret_cast<unsigned long long*>(block + L.hist);
    s.ebdev = reinterpret_cast<EbDev*>(block + L.ebdev);
    s.book = reinterpret_cast<uint32_t*>(block + L.book);
    s.err = reinterpret_cast<uint32_t*>(block + L.err);
    s.summary = reinterpret_cast<EncSummary*>(block + L.summary);
    return s;
}

// The small state of a single-field decompress in one block, so that each
// host check is one copy: [decoder summary | symbol and zero totals | the
// three re-decode list counts].
constexpr size_t kDecSmallTotals = 64;
constexpr size_t kDecSmallLists = 96;
constexpr size_t kDecSmallBytes = 112;
static_assert(sizeof(DecSummary) <= kDecSmallTotals, "decoder summary fits its slot");

// ---- allocation-free compress (vecsz_compress_noalloc)
//
// Its device scratch is the file-scope array `scratch`: one copy per GPU,
// loaded with the code object and held until the process ends.  Host code
// never passes the array's name to a kernel (on the host that name is only a
// stand-in); it passes the device address hipGetSymbolAddress returned for
// the current device, kept in g_na_ptr at context creation.
constexpr uint64_t kNaMaxElems = 1ull << 22;  // 4 Mi elements: a 16 MiB field
constexpr int kNaMaxDict = 4096;
constexpr int kNaDqGrid = 256;                 // dual-quant workgroups (partial histogram rows)
constexpr uint64_t kNaChunks = kNaMaxElems / kChunk;
constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNaCodes = 0;
constexpr size_t kNaSparse = kNaCodes + na_align(kNaMaxElems * 2 + 64);
constexpr size_t kNaSmall = kNaSparse + na_align(kNaMaxElems * 4 + 64);
constexpr size_t kNaPartials = kNaSmall + na_align(small_layout(kNaMaxDict).bytes);
constexpr size_t kNaChunkBits = kNaPartials + na_align((size_t)kNaDqGrid * kNaMaxDict * 4);
constexpr size_t kNaChunkZeros = kNaChunkBits + na_align(kNaChunks * 4);
constexpr size_t kNaBitOff = kNaChunkZeros + na_align(kNaChunks * 4);
constexpr size_t kNaOutOff = kNaBitOff + na_align(kNaChunks * 8);
constexpr size_t kNaScanTmp = kNaOutOff + na_align(kNaChunks * 8);
constexpr size_t kNaScanTot = kNaScanTmp + na_align(((kNaChunks + 4095) / 4096 + 1) * 16);
constexpr size_t kNaRange = kNaScanTot + 256;
constexpr uint64_t kNaParts = (kNaChunks + kEncPartChunks - 1) / kEncPartChunks;  // fused-encode partitions
constexpr uint64_t kNaOutl = 1ull << 20;              // outliers of the fused encode; more: three-kernel path
constexpr size_t kNaEncStatus = kNaRange + na_align(8 + 8 * (size_t)kRangeParts);
constexpr size_t kNaOutlOff = kNaEncStatus + na_align((2 * kNaParts + 1) * 8);
constexpr size_t kNaScratchBytes = kNaOutlOff + na_align(kNaOutl * 4);  // about 32.3 MiB
static_assert(kNaParts >= kNaChunks / kEncPartChunks, "status words for every partition");

}  // namespace

__device__ __align__(256) uint8_t scratch[kNaScratchBytes];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;                  // guards g_na_ptr / g_na_state and g_na_host
void* g_na_ptr[kNaMaxDevices] = {};     // scratch's device address on each GPU
int g_na_state[kNaMaxDevices] = {};     // 0 not looked up, 1 ok, -1 lookup failed
// Host side of an allocation-free call: the codebook and its work arrays,
// reserved once for kNaMaxDict bins.  Allocation-free calls take g_na_mutex
// for their whole run, so they never overlap in one process.
struct NaHost {
    bool reserved = false;
    Codebook cb;
    CodebookWork work;
} g_na_host;

// Looks up scratch's address on `device` (current) once per device.
void na_setup(int device) {
    if (device < 0 || device >= kNaMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_na_mutex);
    if (!g_na_host.reserved) {
        codebook_reserve(g_na_host.cb, g_na_host.work, kNaMaxDict);
        g_na_host.reserved = true;
    }
    if (g_na_state[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        g_na_state[device] = -1;
        return;
    }
    g_na_ptr[device] = ptr;
    g_na_state[device] = 1;
}

CScratch na_scratch(void* base, int dict) {
    uint8_t* b = static_cast<uint8_t*>(base);
    CScratch s = small_scratch(CScratch{}, b + kNaSmall, dict);
    s.codes = reinterpret_cast<uint16_t*>(b + kNaCodes);
    s.sparse = reinterpret_cast<float*>(b + kNaSparse);
    s.partials = reinterpret_cast<uint32_t*>(b + kNaPartials);
    s.chunk_bits = reinterpret_cast<uint32_t*>(b + kNaChunkBits);
    s.chunk_zeros = reinterpret_cast<uint32_t*>(b + kNaChunkZeros);
    s.bit_off = reinterpret_cast<uint64_t*>(b + kNaBitOff);
    s.out_off = reinterpret_cast<uint64_t*>(b + kNaOutOff);
    s.scan_tmp = reinterpret_cast<uint64_t*>(b + kNaScanTmp);
    s.scan_tot = reinterpret_cast<uint64_t*>(b + kNaScanTot);
    s.range = reinterpret_cast<uint32_t*>(b + kNaRange);
    s.enc_status = reinterpret_cast<uint64_t*>(b + kNaEncStatus);
    s.outl = reinterpret_cast<uint32_t*>(b + kNaOutlOff);
    s.outl_cap = kNaOutl;
    return s;
}

// Failure without a heap allocation: the message is copied into the
// context's error string, whose capacity was reserved at creation.
int fail_fixed(vecsz_ctx* c, int code, const char* msg, const char* detail = "") {
    c->last_error.assign(msg);
    c->last_error.append(detail);
    return code;
}

// The compress pipeline.  NA (allocation-free): scratch is `fixed` (slices of
// the static array), nothing is allocated on the device or the host, errors
// are reported through fail_fixed, stage timing is off, and padding and the
// decode index are not available.  Otherwise scratch comes from the context.
// d_index (nullable): decode index of kIdxHead + 2 * idx_segments(n) words,
// written beside the archive.
template <bool NA>
int compress_run(vecsz_ctx* c, const vecsz_params* p, const float* d_in, uint8_t* d_out, uint64_t out_cap,
                 uint64_t* out_bytes, uint64_t* n_outlier, uint64_t* d_index, const CScratch* fixed) {
#define CFAIL(code, msg) return NA ? fail_fixed(c, code, msg) : fail(c, code, msg)
#define CHIP(expr)                                                                                          \
    do {                                                                                                    \
        const hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess)                                                                               \
            return NA ? fail_fixed(c, VECSZ_EHIP, #expr ": ", hipGetErrorString(_e))                        \
                      : fail(c, VECSZ_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));             \
    } while (0)
    int rc = check_params(c, p, true);
    if (rc) return rc;
    CHIP(hipSetDevice(c->device));
    timing_begin(c);
    const int ndim = p->ndim;
    const uint64_t d0 = p->dims[0], d1 = ndim >= 2 ? p->dims[1] : 1, d2 = ndim >= 3 ? p->dims[2] : 1;
    const uint64_t n = d0 * d1 * d2;
    const int dict = p->dict_size, R = dict / 2, B = p->block;
    const bool r2r = p->mode == VECSZ_MODE_R2R;
    const bool device_range = r2r && !(p->value_range > 0);
    const uint64_t nchunks = (n + kChunk - 1) / kChunk;

    CScratch S;
    if (NA) {
        S = *fixed;
    } else {
        if (!c->codes.ensure(n * 2 + 64) || !c->sparse.ensure(n * 4 + 64) ||
            !c->small.ensure(small_layout(dict).bytes) || !c->range.ensure(8 + 8 * (size_t)kRangeParts))
            return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
        if (!c->pin.ensure((size_t)dict * 8 + 4096)) return fail(c, VECSZ_ENOMEM, "pinned buffer");
        S = small_scratch(S, c->small.as<uint8_t>(), dict);
        S.codes = c->codes.as<uint16_t>();
        S.sparse = c->sparse.as<float>();
        S.range = c->range.as<uint32_t>();
    }

    DQArgs a;
    a.d0 = d0;
    a.d1 = d1;
    a.d2 = d2;
    a.n = n;
    a.x8 = (d0 / 8) * 8;
    a.radius = R;
    a.radius_f = (float)R;
    a.thr_vec = (float)vec_threshold(R);
    a.thr_scalar = (float)R;
    a.dict = dict;
    a.ntiles_x = (d0 + 63) / 64;
    a.ntiles_y = ndim >= 2 ? (d1 + B - 1) / B : 1;
    a.ntiles_z = ndim >= 3 ? (d2 + B - 1) / B : 1;
    a.reverse = device_range ? 1 : 0;  // the range pass just streamed the field forward
    a.ebd = nullptr;
    EbFamily eb{};
    if (device_range) {
        // min/max and the eb family stay on the device: no stop before the dual quant
        StageScope s(c, "range");
        launch_range(d_in, n, S.range, c->stream);
        launch_ebfam(S.range, n, d_in, p->eb, S.ebdev, c->stream);
        a.ebd = S.ebdev;
        a.r_f = 0;
        a.r_d = 0;
    } else {
        const double range = r2r ? p->value_range : 0.0;
        eb = eb_family(p->eb, r2r, range);
        a.r_f = eb.r_f;
        a.r_d = eb.r_d;
    }
    // padding: pad values and the halo the dual quantization reads
    a.halo = nullptr;
    a.halo_stride = 0;
    a.hb0 = ref_nblk(d0, B);
    a.hb1 = ndim >= 2 ? ref_nblk(d1, B) : 1;
    const uint64_t hb2 = ndim >= 3 ? ref_nblk(d2, B) : 1;
    const int pmode = p->pad_mode;
    const uint64_t pad_len = pad_len_for(ndim, p->dims, B, pmode);
    if (!NA && pmode != VECSZ_PAD_NONE) {
        const uint64_t nbref = a.hb0 * a.hb1 * hb2;
        if (!c->padv.ensure(pad_len * 4 + 16) || !c->halo.ensure((pmode == VECSZ_PAD_GLOBAL ? 1 : nbref) * 12 + 16) ||
            !c->pad_scratch.ensure(64))
            return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
        PadArgs pa;
        pa.x = d_in;
        pa.d0 = d0;
        pa.d1 = d1;
        pa.d2 = d2;
        pa.n = n;
        pa.ndim = ndim;
        pa.B = B;
        pa.mode = pmode;
        pa.type = p->pad_type;
        pa.constant = p->pad_constant;
        pa.hb0 = a.hb0;
        pa.hb1 = a.hb1;
        pa.hb2 = hb2;
        pa.ebd = device_range ? S.ebdev : nullptr;
        pa.r_d = device_range ? 0.0 : eb.r_d;
        StageScope s(c, "pad");
        if (pmode == VECSZ_PAD_GLOBAL) {
            launch_pad_global(pa, c->pad_scratch.p, c->padv.as<float>(), c->halo.as<float>(), c->stream);
            a.halo_stride = 0;
        } else {
            launch_pad_blocks(pa, c->padv.as<float>(), c->halo.as<float>(), c->stream);
            a.halo_stride = 3;
        }
        a.halo = c->halo.as<float>();
    }
    const bool x4 = pmode == VECSZ_PAD_NONE && dualquant_use_x4(ndim, B, d_in, d0, d1, d2, false);
    const uint64_t tiles_x = x4 ? (d0 + 255) / 256 : a.ntiles_x;
    const uint64_t work = ndim == 1 ? ((n + 7) / 8 + 63) / 64 : tiles_x * a.ntiles_y * a.ntiles_z;
    const int grid_cap = NA ? std::min(c->dq_grid, kNaDqGrid) : c->dq_grid;
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((work + 3) / 4, (uint64_t)grid_cap));
    if (!NA) {
        if (!c->partials.ensure((size_t)grid * dict * 4)) return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
        S.partials = c->partials.as<uint32_t>();
    }

    launch_zero_u32(reinterpret_cast<uint32_t*>(S.hist), (uint64_t)dict * 2, c->stream);
    {
        StageScope s(c, "dualquant");
        launch_dualquant(ndim, B, d_in, S.codes, S.sparse, S.hist, S.partials, a, grid, c->stream);
    }
    CHIP(hipGetLastError());
    // Codebook: on the host from the histogram (one stop: histogram down,
    // header, tree and tables up), or, with device_codebook, by k_codebook on
    // the device, so the compress runs to its end without a stop.
    const uint64_t mb = meta_bytes(ndim);
    const uint64_t pad_bytes = 4 * pad_len;
    const bool dev_cb = !NA && c->device_codebook && dict <= kMaxLdsBook;
    const size_t book_bytes = (size_t)dict * 9;
    const SmallLayout SL = small_layout(dict);
    const size_t book_up = SL.err + 16 - SL.book;  // the tables' upload: through the error word
    Codebook& cb = NA ? g_na_host.cb : c->cb;
    uint64_t hist0 = 0, tb = 0, head = 0;
    int max_len = 64;  // the longest a device-built codebook may hold
    auto put_meta = [&](uint8_t* hdr) {
        memset(hdr, 0, mb);
        int32_t nd = ndim;
        memcpy(&hdr[0], &nd, 4);
        for (int i = 0; i < ndim; i++) {
            int32_t d = (int32_t)p->dims[i];
            memcpy(&hdr[4 + 4 * i], &d, 4);
        }
        memcpy(&hdr[4 * (ndim + 1)], &p->eb, 8);
        memcpy(&hdr[4 * (ndim + 1) + 8], &pad_len, 8);
    };
    if (!dev_cb) {
        uint8_t* hp = c->pin.as<uint8_t>();
        // the histogram and, right behind it, the eb family found on the device
        CHIP(hipMemcpyAsync(hp, S.hist, (size_t)dict * 8 + (device_range ? sizeof(EbDev) : 0), hipMemcpyDeviceToHost,
                            c->stream));
        CHIP(hipStreamSynchronize(c->stream));
        hist0 = reinterpret_cast<const uint64_t*>(hp)[0];  // code 0: the outliers
        if (device_range) {
            EbDev e;
            memcpy(&e, hp + (size_t)dict * 8, sizeof(e));
            if (e.zero_range) CFAIL(VECSZ_ERANGE0, "r2r value range is zero");
            c->sc_range = e.range;
            c->sc_eb_final = e.eb_final;
        } else {
            c->sc_range = r2r ? p->value_range : 0.0;
            c->sc_eb_final = eb.eb_final;
        }
        {
            HostTimer ht(c, "host_codebook");
            rc = build_codebook(reinterpret_cast<const uint64_t*>(hp), dict, cb, NA ? &g_na_host.work : nullptr);
        }
        if (rc == VECSZ_ECODELEN) CFAIL(rc, "Huffman code longer than 64 bits");
        if (rc) CFAIL(rc, "codebook construction failed");
        max_len = cb.max_len;

        // Header: datapack metadata (huff_size / n_outlier filled by k_enc_summary),
        // Huffman header [u32be nodeCount][u32be dict] (src/huffman.hh:649-653) and
        // tree; then the codebook tables.  Both staged in pinned memory.
        tb = cb.tree.size();
        head = mb + 8 + tb;
        if (head + pad_bytes > out_cap) {
            *out_bytes = head + pad_bytes;
            CFAIL(VECSZ_ENOMEM, "output buffer too small");
        }
        if (NA ? c->pin.cap < head + book_bytes + 64 : !c->pin.ensure(head + book_bytes + 64))
            CFAIL(VECSZ_ENOMEM, "pinned buffer");
        uint8_t* hdr = c->pin.as<uint8_t>();
        {
            put_meta(hdr);
            put_u32be(&hdr[mb], (uint32_t)cb.node_count);
            put_u32be(&hdr[mb + 4], (uint32_t)dict);  // stateNum / 2
            memcpy(&hdr[mb + 8], cb.tree.data(), tb);
            uint8_t* bk = hdr + head;
            memcpy(bk, cb.cw_hi.data(), (size_t)dict * 4);
            memcpy(bk + (size_t)dict * 4, cb.cw_lo.data(), (size_t)dict * 4);
            memcpy(bk + (size_t)dict * 8, cb.len.data(), (size_t)dict);
            memset(bk + book_bytes, 0, book_up - book_bytes);  // through the error word: zeros
        }
        if (pad_bytes == 0) {
            CHIP(hipMemcpyAsync(d_out, hdr, head, hipMemcpyHostToDevice, c->stream));
        } else {
            // [metadata][pad values][Huffman header + tree] (src/utils/datapack.hh:17-29)
            CHIP(hipMemcpyAsync(d_out, hdr, mb, hipMemcpyHostToDevice, c->stream));
            CHIP(hipMemcpyAsync(d_out + mb, c->padv.p, pad_bytes, hipMemcpyDeviceToDevice, c->stream));
            CHIP(hipMemcpyAsync(d_out + mb + pad_bytes, hdr + mb, head - mb, hipMemcpyHostToDevice, c->stream));
        }
        CHIP(hipMemcpyAsync(S.book, hdr + head, book_up, hipMemcpyHostToDevice, c->stream));
    } else {
        if (mb + pad_bytes + 8 > out_cap) {
            *out_bytes = mb + pad_bytes + 8;
            return fail(c, VECSZ_ENOMEM, "output buffer too small");
        }
        if (!c->cbinfo.ensure(sizeof(CodebookInfo))) return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
        // the metadata goes up from pinned memory (no stop); pad values follow it
        uint8_t* hdr = c->pin.as<uint8_t>();
        put_meta(hdr);
        CHIP(hipMemcpyAsync(d_out, hdr, mb, hipMemcpyHostToDevice, c->stream));
        if (pad_bytes) CHIP(hipMemcpyAsync(d_out + mb, c->padv.p, pad_bytes, hipMemcpyDeviceToDevice, c->stream));
        StageScope s(c, "codebook");
        launch_codebook(S.hist, dict, S.book, d_out + mb + pad_bytes, out_cap - mb - pad_bytes,
                        c->cbinfo.as<CodebookInfo>(), c->stream);
    }
    const uint32_t* cw_hi = S.book;
    const uint32_t* cw_lo = cw_hi + dict;
    const uint8_t* cw_len = reinterpret_cast<const uint8_t*>(cw_hi + 2 * dict);
    if (dev_cb) launch_zero_u32(S.err, 4, c->stream);  // (the host path's upload cleared it)
    if (NA ? c->pin2.cap < 512 : !c->pin2.ensure(512)) CFAIL(VECSZ_ENOMEM, "pinned buffer");
    if (d_index) {
        // index header words 0-2 and 5; k_enc_summary writes words 3 and 4
        uint64_t* ih = reinterpret_cast<uint64_t*>(c->pin2.as<uint8_t>() + 192);
        ih[0] = kIdxMagic;
        ih[1] = n;
        ih[2] = idx_segments(n);
        ih[5] = (uint64_t)kIdxSeg | ((uint64_t)kIdxGroup << 32);
        CHIP(hipMemcpyAsync(d_index, ih, 24, hipMemcpyHostToDevice, c->stream));
        CHIP(hipMemcpyAsync(d_index + 5, ih + 5, 8, hipMemcpyHostToDevice, c->stream));
    }
    // Encode: one pass (k_encode + k_enc_finish) when the codebook fits in LDS,
    // else chunk bits, a scan of them, the sizes and the packing.
    const uint64_t nparts = encode_parts(nchunks);
    bool fused = c->fused_encode && encode_fused_ok(dict) && !dev_cb;  // (needs the outlier count on the host)
    if (fused && !NA) {
        if (!c->enc_status.ensure((2 * nparts + 1) * 8) || !c->outl.ensure(hist0 * 4 + 64))
            return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
        S.enc_status = c->enc_status.as<uint64_t>();
        S.outl = c->outl.as<uint32_t>();
        S.outl_cap = hist0;
    }
    fused = fused && hist0 <= S.outl_cap;
    if (fused) {
        StageScope s(c, "encode");
        launch_zero_u32(reinterpret_cast<uint32_t*>(S.enc_status), (2 * nparts + 1) * 2, c->stream);
        launch_encode(S.codes, n, cw_hi, cw_lo, cw_len, dict, max_len, S.sparse, d_out, (head + pad_bytes) * 8,
                      out_cap, max_len == 0, S.outl, S.enc_status, S.err, d_index, c->enc_grid, c->stream);
        launch_enc_finish(S.enc_status, n, S.outl, hist0, S.summary, d_out, mb, pad_bytes, tb, out_cap, S.err, d_index,
                          c->stream);
    } else {
        if (!NA) {
            if (!c->chunk_bits.ensure(nchunks * 4) || !c->chunk_zeros.ensure(nchunks * 4) ||
                !c->bit_off.ensure(nchunks * 8) || !c->out_off.ensure(nchunks * 8) ||
                !c->scan_tmp.ensure(((nchunks + 4095) / 4096 + 1) * 16) || !c->scan_tot.ensure(16))
                return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
            S.chunk_bits = c->chunk_bits.as<uint32_t>();
            S.chunk_zeros = c->chunk_zeros.as<uint32_t>();
            S.bit_off = c->bit_off.as<uint64_t>();
            S.out_off = c->out_off.as<uint64_t>();
            S.scan_tmp = c->scan_tmp.as<uint64_t>();
            S.scan_tot = c->scan_tot.as<uint64_t>();
        }
        {
            StageScope s(c, "chunk_bits");
            launch_chunk_bits(S.codes, n, cw_len, dict, S.chunk_bits, S.chunk_zeros, nchunks, c->enc_grid, c->stream);
        }
        {
            StageScope s(c, "scan");
            launch_scan_u32x2(S.chunk_bits, S.chunk_zeros, nchunks, S.bit_off, S.out_off, S.scan_tmp, S.scan_tot,
                              c->stream);
            if (dev_cb)
                launch_enc_summary_dev(S.scan_tot, S.summary, d_out, mb, pad_bytes, c->cbinfo.as<CodebookInfo>(),
                                       out_cap, S.err, d_index, c->stream);
            else
                launch_enc_summary(S.scan_tot, S.summary, d_out, mb, pad_bytes, tb, out_cap, S.err, d_index,
                                   c->stream);
        }
        {
            StageScope s(c, "pack");
            launch_pack(S.codes, n, cw_hi, cw_lo, cw_len, dict, S.bit_off, S.out_off, nchunks, S.summary, S.sparse,
                        d_out, S.err, std::max(max_len, 1), c->enc_grid, d_index, c->stream);
        }
    }
    CHIP(hipGetLastError());
    uint8_t* h = c->pin2.as<uint8_t>();
    // the error word and, behind it, the summary
    CHIP(hipMemcpyAsync(h, S.err, SL.summary - SL.err + sizeof(EncSummary), hipMemcpyDeviceToHost, c->stream));
    if (dev_cb) {
        CHIP(hipMemcpyAsync(h + 256, c->cbinfo.p, sizeof(CodebookInfo), hipMemcpyDeviceToHost, c->stream));
        if (device_range) CHIP(hipMemcpyAsync(h + 320, S.ebdev, sizeof(EbDev), hipMemcpyDeviceToHost, c->stream));
    }
    CHIP(hipStreamSynchronize(c->stream));
    EncSummary ES;
    memcpy(&ES, h + (SL.summary - SL.err), sizeof(ES));
    uint32_t e;
    memcpy(&e, h, 4);
    timing_end(c);
    if (dev_cb) {
        // what the host path checks at its stop: the range, then the codebook
        if (device_range) {
            EbDev ed;
            memcpy(&ed, h + 320, sizeof(ed));
            if (ed.zero_range) return fail(c, VECSZ_ERANGE0, "r2r value range is zero");
            c->sc_range = ed.range;
            c->sc_eb_final = ed.eb_final;
        } else {
            c->sc_range = r2r ? p->value_range : 0.0;
            c->sc_eb_final = eb.eb_final;
        }
        CodebookInfo ci;
        memcpy(&ci, h + 256, sizeof(ci));
        if (ci.status == VECSZ_ECODELEN) return fail(c, VECSZ_ECODELEN, "Huffman code longer than 64 bits");
        if (ci.status == VECSZ_ENOMEM) {
            *out_bytes = mb + pad_bytes + 8 + ci.tree_bytes;
            return fail(c, VECSZ_ENOMEM, "output buffer too small");
        }
        if (ci.status) return fail(c, ci.status, "codebook construction failed");
    }
    *out_bytes = ES.archive_size;
    if (n_outlier) *n_outlier = ES.n_outlier;
    if (e) CFAIL(VECSZ_ENOMEM, "output buffer too small");
    vecsz_sidecar& sc = c->sidecar;
    sc.magic = VECSZ_SIDECAR_MAGIC;
    sc.version = 1;
    sc.ndim = ndim;
    sc.mode = p->mode;
    sc.block = B;
    sc.dict_size = dict;
    sc.eb = p->eb;
    sc.value_range = c->sc_range;
    sc.eb_final = c->sc_eb_final;
    c->have_sidecar = true;
    return VECSZ_OK;
#undef CHIP
#undef CFAIL
}

int compress_device(vecsz_ctx* c, const vecsz_params* p, const float* d_in, uint8_t* d_out, uint64_t out_cap,
                    uint64_t* out_bytes, uint64_t* n_outlier, uint64_t* d_index = nullptr) {
    return compress_run<false>(c, p, d_in, d_out, out_cap, out_bytes, n_outlier, d_index, nullptr);
}

int parse_meta(const uint8_t* h, uint64_t avail, uint64_t total, int* ndim, uint64_t dims[3], double* eb,
               uint64_t* pad_len, uint64_t* huff_size, uint64_t* n_outlier) {
    if (avail < 4) return VECSZ_ECORRUPT;
    int32_t nd;
    memcpy(&nd, h, 4);
    if (nd < 1 || nd > 3) return VECSZ_ECORRUPT;
    const uint64_t mb = meta_bytes(nd);
    if (avail < mb) return VECSZ_ECORRUPT;
    *ndim = nd;
    dims[0] = dims[1] = dims[2] = 1;
    for (int i = 0; i < nd; i++) {
        int32_t d;
        memcpy(&d, h + 4 + 4 * i, 4);
        if (d <= 0) return VECSZ_ECORRUPT;
        dims[i] = (uint64_t)d;
    }
    // the element count must not wrap (every later size is derived from it)
    uint64_t n;
    if (!field_elems(nd, dims, &n)) return VECSZ_ECORRUPT;
    memcpy(eb, h + 4 * (nd + 1), 8);
    uint64_t lens[3];
    memcpy(lens, h + 4 * (nd + 1) + 8, 24);
    *pad_len = lens[0];
    *huff_size = lens[1];
    *n_outlier = lens[2];
    if (lens[0] > total || lens[1] > total || lens[2] > total) return VECSZ_ECORRUPT;
    if (mb + 4 * lens[0] + lens[1] + 4 * lens[2] > total) return VECSZ_ECORRUPT;
    // at most one outlier per element
    if (lens[2] > n) return VECSZ_ECORRUPT;
    return VECSZ_OK;
}

bool tree_is_full(const DecodeTables& dt) {
    // every internal node reachable from the root has both children
    std::vector<uint32_t> stk{0};
    std::vector<uint8_t> seen(dt.node_count, 0);
    while (!stk.empty()) {
        uint32_t nd = stk.back();
        stk.pop_back();
        if (seen[nd]) return false;  // cycle or shared node
        seen[nd] = 1;
        for (int b = 0; b < 2; b++) {
            uint32_t ch = dt.child[2 * nd + b];
            if (ch == kNodeAbsent) return false;
            if (!(ch & kNodeLeaf)) stk.push_back(ch);
        }
    }
    return true;
}

// d_index (nullable): a decode index written by compress_device for this
// archive (index_words words); the decode then starts every segment at its
// first codeword instead of synchronising.
int decompress_device(vecsz_ctx* c, const vecsz_params* p, const uint8_t* d_in, uint64_t in_bytes, float* d_out,
                      uint64_t out_cap, const uint64_t* d_index = nullptr, uint64_t index_words = 0) {
    int rc = check_params(c, p, false);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    timing_begin(c);
    // One round trip for the metadata, the Huffman header and (for trees of
    // up to ~14,000 nodes) the tree itself; a larger tree takes a second copy.
    constexpr uint64_t kHeadRead = 1 << 17;
    if (!c->pin.ensure(kHeadRead + 64)) return fail(c, VECSZ_ENOMEM, "pinned buffer");
    uint8_t* h = c->pin.as<uint8_t>();
    const uint64_t head_avail = std::min<uint64_t>(in_bytes, kHeadRead);
    HIPCHK(c, hipMemcpyAsync(h, d_in, head_avail, hipMemcpyDeviceToHost, c->stream));
    uint64_t ih[kIdxHead] = {0, 0, 0, 0, 0, 0};
    if (d_index) {
        if (index_words < kIdxHead) return fail(c, VECSZ_EINVAL, "decode index too short");
        HIPCHK(c, hipMemcpyAsync(h + kHeadRead, d_index, 8 * kIdxHead, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (d_index) memcpy(ih, h + kHeadRead, sizeof(ih));
    int ndim;
    uint64_t dims[3], pad_len, huff_size, nout;
    double ebu;
    rc = parse_meta(h, head_avail, in_bytes, &ndim, dims, &ebu, &pad_len, &huff_size, &nout);
    if (rc) return fail(c, rc, "archive header is corrupt");
    const uint64_t n = dims[0] * dims[1] * dims[2];
    if (n > out_cap) return fail(c, VECSZ_ENOMEM, "output buffer too small");
    const bool r2r = p->mode == VECSZ_MODE_R2R;
    if (r2r && !(p->value_range != 0)) return fail(c, VECSZ_EINVAL, "r2r decompress needs params.value_range");
    const EbFamily eb = eb_family(ebu, r2r, p->value_range);
    const int B = p->block, R = p->dict_size / 2;

    const uint64_t mb = meta_bytes(ndim);
    const uint8_t* d_huff = d_in + mb + 4 * pad_len;
    if (huff_size < 8) return fail(c, VECSZ_ECORRUPT, "Huffman section too short");
    DecodeTables& dt = c->dt;
    uint64_t tb = 0;
    {
        HostTimer ht(c, "host_tables");
        const uint64_t hoff = mb + 4 * pad_len;
        const uint8_t* hp = h + hoff;
        if (hoff + 8 > head_avail) {
            // a pad section longer than the first read: fetch the Huffman header
            if (!c->pin2.ensure(64)) return fail(c, VECSZ_ENOMEM, "pinned buffer");
            hp = c->pin2.as<uint8_t>();
            HIPCHK(c, hipMemcpyAsync(c->pin2.p, d_huff, 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        int32_t ncs = (int32_t)(((uint32_t)hp[0] << 24) | ((uint32_t)hp[1] << 16) | ((uint32_t)hp[2] << 8) | hp[3]);
        if (ncs <= 0) return fail(c, VECSZ_ECORRUPT, "Huffman node count");
        tb = tree_byte_size((uint64_t)ncs);
        if (8 + tb > huff_size) return fail(c, VECSZ_ECORRUPT, "Huffman tree truncated");
        if (hoff + 8 + tb > head_avail) {
            if (!c->pin2.ensure(8 + tb)) return fail(c, VECSZ_ENOMEM, "pinned buffer");
            hp = c->pin2.as<uint8_t>();
            HIPCHK(c, hipMemcpyAsync(c->pin2.p, d_huff, 8 + tb, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        rc = build_decode_tables(hp, huff_size, dt);
        if (rc) return fail(c, rc, "Huffman tree is corrupt");
        if (!dt.root_leaf && !tree_is_full(dt)) return fail(c, VECSZ_ECORRUPT, "Huffman tree is not full");
    }

    if (!c->codes.ensure(n * 2 + 64) || !c->sparse.ensure(n * 4 + 64))
        return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
    const uint8_t* d_stream = d_huff + 8 + tb;
    const uint64_t stream_bytes = huff_size - 8 - tb;
    if (d_index) {
        const IdxLayout L = idx_layout(n);
        if (ih[0] != kIdxMagic || ih[1] != n || ih[2] != L.nseg ||
            ih[5] != ((uint64_t)kIdxSeg | ((uint64_t)kIdxGroup << 32)) || index_words < L.words ||
            (ih[3] + 7) / 8 != stream_bytes || ih[4] != nout)
            return fail(c, VECSZ_EINVAL, "decode index does not belong to this archive");
    }
    const uint8_t* d_outl = d_huff + huff_size;
    uint16_t* codes = c->codes.as<uint16_t>();
    float* sparse = c->sparse.as<float>();
    bool indexed = false;  // decode errors then show at the end
    // Pad halo of the reconstruction.
    const uint64_t d0 = dims[0], d1 = dims[1], d2 = dims[2];
    ReconArgs ra;
    ra.halo = nullptr;
    ra.halo_stride = 0;
    ra.hb0 = ref_nblk(d0, B);
    ra.hb1 = ndim >= 2 ? ref_nblk(d1, B) : 1;
    if (pad_len > 0) {
        // padding mode from the pad value count (src/vecsz_interface.hh:287-298);
        // 1D without a mode still reads pad_vals[0] (src/dualquant.hh:925-927)
        const uint64_t nbref = ref_blocks(ndim, dims, B);
        const int mode = pad_len == 1 ? VECSZ_PAD_GLOBAL
                         : pad_len == nbref ? VECSZ_PAD_BLOCK
                         : pad_len == (uint64_t)ndim * nbref ? VECSZ_PAD_EDGE : VECSZ_PAD_NONE;
        if (mode != VECSZ_PAD_NONE || ndim == 1) {
            if (!c->padv.ensure(pad_len * 4 + 16) || !c->halo.ensure(nbref * 12 + 16))
                return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
            HIPCHK(c, hipMemcpyAsync(c->padv.p, d_in + mb, 4 * pad_len, hipMemcpyDeviceToDevice, c->stream));
            launch_pad_halo(ndim, mode, c->padv.as<float>(), pad_len, nbref, c->halo.as<float>(), c->stream);
            ra.halo = c->halo.as<float>();
            ra.halo_stride = 3;
        }
    }
    // 1D with a decode index: the reconstruction runs inside the decode
    const bool fused_1d = d_index && ndim == 1 && !dt.root_leaf && (reinterpret_cast<uintptr_t>(d_out) & 15) == 0;
    ra.d0 = d0;
    ra.d1 = d1;
    ra.d2 = d2;
    ra.n = n;
    ra.radius = R;
    ra.twoeb = eb.twoeb;
    ra.ntiles_x = (d0 + 63) / 64;
    ra.ntiles_y = ndim >= 2 ? (d1 + B - 1) / B : 1;
    ra.ntiles_z = ndim >= 3 ? (d2 + B - 1) / B : 1;
    const uint64_t nblocks = ((d0 + B - 1) / B) * (ndim >= 2 ? (d1 + B - 1) / B : 1) * (ndim >= 3 ? (d2 + B - 1) / B : 1);
    const uint64_t rwork = ra.ntiles_x * ra.ntiles_y * ra.ntiles_z;
    const int rgrid = (int)std::max<uint64_t>(1, std::min<uint64_t>((rwork + 3) / 4, 2048));
    if (!fused_1d && (!c->flagged.ensure(nblocks * 4 + 16) || !c->nflagged.ensure(16)))
        return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
    // Lorenzo reconstruction; guard: as the write pass's (nonzero: nothing
    // runs, the decode is not final yet)
    bool recon_issued = false;
    auto reconstruct = [&](const uint32_t* guard) {
        launch_zero_u32(c->nflagged.as<uint32_t>(), 1, c->stream);
        StageScope s(c, "reconstruct");
        ra.guard = guard;
        launch_reconstruct(ndim, B, codes, sparse, d_out, ra, c->flagged.as<uint32_t>(), c->nflagged.as<uint32_t>(),
                           rgrid, c->stream);
        launch_reconstruct_fallback(ndim, B, codes, sparse, d_out, ra, c->flagged.as<uint32_t>(),
                                    c->nflagged.as<uint32_t>(), c->stream);
        recon_issued = true;
    };

    if (dt.root_leaf) {
        StageScope s(c, "decode");
        launch_fill_u16(codes, n, (uint16_t)dt.root_sym, c->stream);
        if (dt.root_sym == 0) {
            if (nout < n) return fail(c, VECSZ_ECORRUPT, "outlier section shorter than the field");
            launch_scatter_all_outliers(d_outl, n, sparse, c->stream);
        }
    } else {
        const uint64_t T = stream_bytes * 8;
        const uint64_t nsub = std::max<uint64_t>(1, (T + kSubBits - 1) / kSubBits);
        if (!c->sub_b.ensure(nsub * 8) || !c->sub_e.ensure(nsub * 8) || !c->sub_cnt.ensure(nsub * 4) ||
            !c->sub_zeros.ensure(nsub * 4) || !c->sub_lj.ensure(nsub * 8) || !c->sub_ls.ensure(nsub * 16) ||
            !c->sym_off.ensure(nsub * 8) || !c->zero_off.ensure(nsub * 8) ||
            !c->scan_tmp.ensure(((nsub + 4095) / 4096 + 1) * 16) || !c->dec_small.ensure(kDecSmallBytes))
            return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
        // Decode tables in one pinned staging copy: the compact form when the
        // symbols fit 12 bits (the kernels keep its deeper tables in LDS when
        // they fit beside the first-level table, else read them in global
        // memory), else 32-bit.
        const bool t16 = !dt.lut16.empty();
        const size_t tab_bytes = t16 ? (dt.lut16.size() + dt.sub16.size()) * 2 : (dt.lut.size() + dt.sub.size()) * 4;
        if (!c->tabs.ensure(tab_bytes + 16) || !c->pin_tab.ensure(tab_bytes + 16))
            return fail(c, VECSZ_ENOMEM, "decode table allocation failed");
        // (the previous call's copy out of pin_tab finished before that call's
        // mid-call synchronisation)
        uint8_t* tp = c->pin_tab.as<uint8_t>();
        if (t16) {
            memcpy(tp, dt.lut16.data(), dt.lut16.size() * 2);
            memcpy(tp + dt.lut16.size() * 2, dt.sub16.data(), dt.sub16.size() * 2);
        } else {
            memcpy(tp, dt.lut.data(), dt.lut.size() * 4);
            memcpy(tp + dt.lut.size() * 4, dt.sub.data(), dt.sub.size() * 4);
        }
        HIPCHK(c, hipMemcpyAsync(c->tabs.p, tp, tab_bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemsetAsync(c->dec_small.p, 0, kDecSmallBytes, c->stream));
        DecodeArgs da;
        da.stream = d_stream;
        da.stream_bits = T;
        da.n = n;
        da.lut = t16 ? nullptr : c->tabs.as<uint32_t>();
        da.sub = t16 ? nullptr : c->tabs.as<uint32_t>() + dt.lut.size();
        da.lut16 = t16 ? c->tabs.as<uint16_t>() : nullptr;
        da.sub16 = t16 ? c->tabs.as<uint16_t>() + dt.lut16.size() : nullptr;
        da.n_sub16 = t16 && dt.sub16.size() <= kMaxSub16 ? (uint32_t)dt.sub16.size() : 0;
        da.nsub = nsub;
        da.outliers = d_outl;
        da.n_outlier = nout;
        SubState ss{};  // jbase 0: list entries are this field's subsequence numbers
        ss.b = c->sub_b.as<uint64_t>();
        ss.e = c->sub_e.as<uint64_t>();
        ss.cnt = c->sub_cnt.as<uint32_t>();
        ss.zer = c->sub_zeros.as<uint32_t>();
        ss.list_j[0] = c->sub_lj.as<uint32_t>();
        ss.list_j[1] = c->sub_lj.as<uint32_t>() + nsub;
        ss.list_start[0] = c->sub_ls.as<uint64_t>();
        ss.list_start[1] = c->sub_ls.as<uint64_t>() + nsub;
        ss.list_n = reinterpret_cast<uint32_t*>(c->dec_small.as<uint8_t>() + kDecSmallLists);
        DecSummary* dsum = c->dec_small.as<DecSummary>();
        uint64_t* stot = reinterpret_cast<uint64_t*>(c->dec_small.as<uint8_t>() + kDecSmallTotals);
        if (d_index) {
            StageScope s(c, "decode_indexed");
            Recon1D r1{d_out, B, R, eb.twoeb, ra.halo, ra.halo_stride};
            launch_decode_indexed(da, d_index, ih[3], codes, sparse, fused_1d ? &r1 : nullptr, dsum, c->stream);
            HIPCHK(c, hipGetLastError());
            indexed = true;
        } else {
        {
            StageScope s(c, "decode_sync");
            launch_decode_sync(da, ss, dsum, c->stream);
        }
        // Re-decode passes, two without a host check (most streams need one or
        // none), then the scan and the write pass, which do nothing unless the
        // list the third pass would read is empty: in the common case the
        // decode needs one host check, after the write.  Otherwise passes run
        // in growing batches between host checks while the lists ask for more
        // (each pass resolves at least the first subsequence of every chain,
        // so nsub passes suffice), and then the scan and the write again.
        constexpr uint32_t kPassesAhead = 2;
        uint32_t pass = 0;
        {
            StageScope s(c, "decode_redo");
            for (; pass < kPassesAhead; pass++) launch_decode_pass(da, ss, pass, dsum, c->stream);
        }
        auto scan_write = [&](const uint32_t* guard) {
            {
                StageScope s(c, "decode_scan");
                launch_scan_u32x2(ss.cnt, ss.zer, nsub, c->sym_off.as<uint64_t>(), c->zero_off.as<uint64_t>(),
                                  c->scan_tmp.as<uint64_t>(), stot, c->stream, guard);
            }
            StageScope s(c, "decode_write");
            launch_decode_write(da, ss, c->sym_off.as<uint64_t>(), c->zero_off.as<uint64_t>(), codes, sparse, dsum,
                                c->stream, guard);
        };
        scan_write(ss.list_n + pass % 3);
        // the reconstruction goes in behind the write pass, under the same
        // guard, before the host check: in the common case the check then
        // waits on the whole decompress
        if (!fused_1d) reconstruct(ss.list_n + pass % 3);
        HIPCHK(c, hipGetLastError());
        DecSummary ds;
        uint64_t tot[2];
        uint32_t left = 0;
        auto check = [&]() -> int {  // summary, symbol totals and the pending list count
            HIPCHK(c, hipMemcpyAsync(h, c->dec_small.p, kDecSmallBytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            memcpy(&ds, h, sizeof(ds));
            memcpy(tot, h + kDecSmallTotals, 16);
            memcpy(&left, h + kDecSmallLists + 4 * (pass % 3), 4);
            return ds.error ? fail(c, VECSZ_ECORRUPT, "bitstream does not decode") : VECSZ_OK;
        };
        if (int rc2 = check()) return rc2;
        if (left) {
            for (uint32_t batch = 2; left; batch = std::min<uint32_t>(2 * batch, 16)) {
                if (pass > nsub + 2) return fail(c, VECSZ_ECORRUPT, "subsequence boundaries do not converge");
                // (a pass over an empty list does nothing but clear the list after next)
                {
                    StageScope s(c, "decode_redo");
                    for (uint32_t b = 0; b < batch; b++) launch_decode_pass(da, ss, pass++, dsum, c->stream);
                }
                if (int rc2 = check()) return rc2;
            }
            scan_write(nullptr);
            if (!fused_1d) reconstruct(nullptr);
            HIPCHK(c, hipGetLastError());
            if (int rc2 = check()) return rc2;
        }
        if (tot[0] < n) return fail(c, VECSZ_ECORRUPT, "bitstream holds fewer symbols than the field");
        }
    }

    if (!fused_1d && !recon_issued) reconstruct(nullptr);
    HIPCHK(c, hipGetLastError());
    if (indexed) HIPCHK(c, hipMemcpyAsync(h, c->dec_small.p, sizeof(DecSummary), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    timing_end(c);
    if (indexed) {
        DecSummary ds;
        memcpy(&ds, h, sizeof(ds));
        if (ds.error) return fail(c, VECSZ_ECORRUPT, "bitstream does not decode with its decode index");
    }
    return VECSZ_OK;
}

}  // namespace

extern "C" {

const char* vecsz_version(void) { return "vecsz_amd 0.1 gfx950"; }

const char* vecsz_strerror(int s) {
    switch (s) {
        case VECSZ_OK: return "ok";
        case VECSZ_EINVAL: return "invalid argument";
        case VECSZ_ENOMEM: return "out of memory or buffer too small";
        case VECSZ_ECORRUPT: return "corrupt archive";
        case VECSZ_ECODELEN: return "Huffman code longer than 64 bits";
        case VECSZ_EHIP: return "HIP runtime error";
        case VECSZ_ERANGE0: return "zero value range in r2r mode";
        case VECSZ_ENODEV: return "no usable GPU";
        default: return "unknown status";
    }
}

void vecsz_params_default(vecsz_params* p) {
    memset(p, 0, sizeof(*p));
    p->ndim = 1;
    p->dims[0] = p->dims[1] = p->dims[2] = 1;
    p->mode = VECSZ_MODE_ABS;
    p->eb = 1e-4;
    p->value_range = 0;
    p->block = 8;
    p->dict_size = 4096;
    p->pad_mode = VECSZ_PAD_NONE;
    p->pad_type = VECSZ_PADV_CONST;  // pad_type{3} (src/argument_parser/argparse.hh:77)
    p->pad_constant = 0.f;
}

int vecsz_ctx_create(vecsz_ctx** out, int device, void* stream) {
    if (!out) return VECSZ_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return VECSZ_ENODEV;
    }
    if (device < 0 || device >= ndev) return VECSZ_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return VECSZ_EHIP;
    vecsz_ctx* c = new vecsz_ctx();
    c->device = device;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            return VECSZ_EHIP;
        }
        c->own_stream = true;
    }
    // set-up for vecsz_compress_noalloc: error text capacity, the pinned
    // staging it uses, and (once per device, this device current) the
    // address of the static scratch array
    c->last_error.reserve(512);
    if (!c->pin.ensure(1 << 20) || !c->pin2.ensure(256)) {
        vecsz_ctx_destroy(c);
        return VECSZ_ENOMEM;
    }
    na_setup(device);
    *out = c;
    return VECSZ_OK;
}

int vecsz_ctx_destroy(vecsz_ctx* c) {
    if (!c) return VECSZ_OK;
    for (vecsz_ctx* w : c->workers) vecsz_ctx_destroy(w);
    c->workers.clear();
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (GraphSlot& g : c->b_graph) g.reset();
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return VECSZ_OK;
}

const char* vecsz_ctx_last_error(vecsz_ctx* c) { return c ? c->last_error.c_str() : "no context"; }

int vecsz_compress_bound(const vecsz_params* p, uint64_t* bytes) {
    if (!p || !bytes || p->ndim < 1 || p->ndim > 3 || p->dict_size < 2 || p->dict_size > 65536) return VECSZ_EINVAL;
    uint64_t n;
    if (!field_elems(p->ndim, p->dims, &n)) return VECSZ_EINVAL;
    const uint64_t nsym = std::min<uint64_t>((uint64_t)p->dict_size, n);
    const uint64_t tree = tree_byte_size(2 * nsym - 1);
    // every symbol at most 64 bits, every element possibly an outlier
    const int B = valid_block(p->block) ? p->block : 8;
    *bytes = meta_bytes(p->ndim) + 4 * pad_len_for(p->ndim, p->dims, B, p->pad_mode) + 8 + tree + n * 8 + n * 4 + 8;
    return VECSZ_OK;
}

int vecsz_compress(vecsz_ctx* c, const vecsz_params* p, const float* d_in, uint8_t* d_out, uint64_t out_cap,
                   uint64_t* out_bytes, uint64_t* n_outlier) {
    if (!c || !d_in || !d_out || !out_bytes) return fail(c, VECSZ_EINVAL, "NULL argument");
    return compress_device(c, p, d_in, d_out, out_cap, out_bytes, n_outlier);
}

int vecsz_compress_noalloc(vecsz_ctx* c, const vecsz_params* p, const float* d_in, uint8_t* d_out, uint64_t out_cap,
                           uint64_t* out_bytes, uint64_t* n_outlier) {
    if (!c) return VECSZ_EINVAL;
    if (!p || !d_in || !d_out || !out_bytes) return fail(c, VECSZ_EINVAL, "NULL argument");
    int rc = check_params(c, p, true);
    if (rc) return rc;
    uint64_t n = 0;
    (void)field_elems(p->ndim, p->dims, &n);
    if (n > kNaMaxElems) return fail(c, VECSZ_EINVAL, "allocation-free compress: field above 2^22 elements");
    if (p->dict_size > kNaMaxDict) return fail(c, VECSZ_EINVAL, "allocation-free compress: dict_size above 4096");
    if (p->pad_mode != VECSZ_PAD_NONE) return fail(c, VECSZ_EINVAL, "allocation-free compress: no padding");
    // one allocation-free call at a time in the process: they share the
    // scratch array of their GPU and the host codebook arrays
    std::lock_guard<std::mutex> lk(g_na_mutex);
    if (c->device < 0 || c->device >= kNaMaxDevices || g_na_state[c->device] != 1)
        return fail(c, VECSZ_EHIP, "allocation-free compress: no scratch address on this device");
    const CScratch S = na_scratch(g_na_ptr[c->device], p->dict_size);
    const bool was_timing = c->timing;
    c->timing = false;  // stage timing would create events
    rc = compress_run<true>(c, p, d_in, d_out, out_cap, out_bytes, n_outlier, nullptr, &S);
    c->timing = was_timing;
    return rc;
}

int vecsz_index_words(const vecsz_params* p, uint64_t* words) {
    if (!p || !words) return VECSZ_EINVAL;
    uint64_t n;
    if (!field_elems(p->ndim, p->dims, &n)) return VECSZ_EINVAL;
    *words = idx_layout(n).words;
    return VECSZ_OK;
}

int vecsz_compress_indexed(vecsz_ctx* c, const vecsz_params* p, const float* d_in, uint8_t* d_out, uint64_t out_cap,
                           uint64_t* out_bytes, uint64_t* n_outlier, uint64_t* d_index, uint64_t index_words) {
    if (!c) return VECSZ_EINVAL;
    if (!d_in || !d_out || !out_bytes || !d_index) return fail(c, VECSZ_EINVAL, "NULL argument");
    uint64_t need = 0;
    if (vecsz_index_words(p, &need)) return fail(c, VECSZ_EINVAL, "params");
    if (index_words < need) return fail(c, VECSZ_ENOMEM, "decode index buffer too small");
    return compress_device(c, p, d_in, d_out, out_cap, out_bytes, n_outlier, d_index);
}

int vecsz_decompress_indexed(vecsz_ctx* c, const vecsz_params* p, const uint8_t* d_in, uint64_t in_bytes,
                             const uint64_t* d_index, uint64_t index_words, float* d_out, uint64_t out_cap) {
    if (!c) return VECSZ_EINVAL;
    if (!d_in || !d_out || !d_index) return fail(c, VECSZ_EINVAL, "NULL argument");
    return decompress_device(c, p, d_in, in_bytes, d_out, out_cap, d_index, index_words);
}

int vecsz_decompress(vecsz_ctx* c, const vecsz_params* p, const uint8_t* d_in, uint64_t in_bytes, float* d_out,
                     uint64_t out_cap) {
    if (!c || !d_in || !d_out) return fail(c, VECSZ_EINVAL, "NULL argument");
    return decompress_device(c, p, d_in, in_bytes, d_out, out_cap);
}

int vecsz_compress_host(vecsz_ctx* c, const vecsz_params* p, const float* h_in, uint8_t* h_out, uint64_t out_cap,
                        uint64_t* out_bytes, uint64_t* n_outlier) {
    if (!c || !p || !h_in || !h_out || !out_bytes) return fail(c, VECSZ_EINVAL, "NULL argument");
    int rc = check_params(c, p, true);
    if (rc) return rc;
    uint64_t n = 1;
    for (int i = 0; i < p->ndim; i++) n *= p->dims[i];
    uint64_t bound;
    vecsz_compress_bound(p, &bound);
    if (!c->stage_in.ensure(n * 4) || !c->stage_out.ensure(bound)) return fail(c, VECSZ_ENOMEM, "staging buffers");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->stage_in.p, h_in, n * 4, hipMemcpyHostToDevice, c->stream));
    uint64_t sz = 0;
    rc = compress_device(c, p, c->stage_in.as<float>(), c->stage_out.as<uint8_t>(), bound, &sz, n_outlier);
    *out_bytes = sz;
    if (rc) return rc;
    if (sz > out_cap) return fail(c, VECSZ_ENOMEM, "output buffer too small");
    HIPCHK(c, hipMemcpyAsync(h_out, c->stage_out.p, sz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VECSZ_OK;
}

int vecsz_decompress_host(vecsz_ctx* c, const vecsz_params* p, const uint8_t* h_in, uint64_t in_bytes, float* h_out,
                          uint64_t out_cap) {
    if (!c || !p || !h_in || !h_out) return fail(c, VECSZ_EINVAL, "NULL argument");
    int ndim;
    uint64_t dims[3], pad, huff, nout;
    double eb;
    int rc = parse_meta(h_in, in_bytes, in_bytes, &ndim, dims, &eb, &pad, &huff, &nout);
    if (rc) return fail(c, rc, "archive header is corrupt");
    const uint64_t n = dims[0] * dims[1] * dims[2];
    if (n > out_cap) return fail(c, VECSZ_ENOMEM, "output buffer too small");
    if (!c->stage_in.ensure(in_bytes + 64) || !c->stage_out.ensure(n * 4)) return fail(c, VECSZ_ENOMEM, "staging buffers");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->stage_in.p, h_in, in_bytes, hipMemcpyHostToDevice, c->stream));
    rc = decompress_device(c, p, c->stage_in.as<uint8_t>(), in_bytes, c->stage_out.as<float>(), n);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(h_out, c->stage_out.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VECSZ_OK;
}

int vecsz_archive_info(const uint8_t* h, uint64_t bytes, int* ndim, uint64_t dims[3], double* eb, uint64_t* huff_size,
                       uint64_t* n_outlier) {
    if (!h || !ndim || !dims || !eb || !huff_size || !n_outlier) return VECSZ_EINVAL;
    uint64_t pad;
    return parse_meta(h, bytes, bytes, ndim, dims, eb, &pad, huff_size, n_outlier);
}

int vecsz_huffman_codebook(const uint64_t* freq, int dict, uint64_t* code_bits, uint8_t* code_len, uint8_t* tree,
                           uint64_t tree_cap, uint64_t* tree_bytes, int* node_count) {
    if (!freq || dict < 2 || dict > 65536 || !tree_bytes || !node_count) return VECSZ_EINVAL;
    Codebook cb;
    const int rc = build_codebook(freq, dict, cb);
    if (rc) return rc;
    *tree_bytes = cb.tree.size();
    *node_count = cb.node_count;
    for (int s = 0; s < dict; s++) {
        if (code_bits) code_bits[s] = ((uint64_t)cb.cw_hi[s] << 32) | cb.cw_lo[s];
        if (code_len) code_len[s] = cb.len[s];
    }
    if (tree) {
        if (tree_cap < cb.tree.size()) return VECSZ_ENOMEM;
        memcpy(tree, cb.tree.data(), cb.tree.size());
    }
    return VECSZ_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches

// Runs job(worker, k) for every field k, largest first, on up to
// c->nworkers host threads.  Each worker is a context with its own stream and
// scratch, so small fields overlap on the GPU and one field's host codebook
// overlaps the others' kernels.  Returns after every worker's stream is idle.
template <typename Prep, typename Job>
static int run_batch(vecsz_ctx* c, int nf, const uint64_t* weight, int* status, Prep&& prep, Job&& job) {
    if (nf <= 0) return VECSZ_OK;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // inputs produced on the caller's stream
    const int nw = std::max(1, std::min(c->nworkers, nf));
    while ((int)c->workers.size() < nw) {
        vecsz_ctx* w = nullptr;
        const int rc = vecsz_ctx_create(&w, c->device, nullptr);
        if (rc) return fail(c, rc, "batch worker context");
        c->workers.push_back(w);
    }
    // size every worker's field-proportional scratch for the largest field
    // up front: growing it later means hipFree, which stalls all streams
    for (int t = 0; t < nw; t++) {
        c->workers[t]->dq_grid = c->dq_grid;
        c->workers[t]->enc_grid = c->enc_grid;
        if (!prep(c->workers[t])) return fail(c, VECSZ_ENOMEM, "batch worker scratch");
    }
    std::vector<int> order(nf);
    for (int k = 0; k < nf; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return weight[a] > weight[b]; });
    std::atomic<int> next{0};
    auto body = [&](vecsz_ctx* w) {
        (void)hipSetDevice(w->device);
        for (int i; (i = next.fetch_add(1)) < nf;) {
            const int k = order[i];
            status[k] = job(w, k);
        }
        (void)hipStreamSynchronize(w->stream);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nw; t++) th.emplace_back(body, c->workers[t]);
    body(c->workers[0]);
    for (auto& t : th) t.join();
    for (int k = 0; k < nf; k++)
        if (status[k]) return fail(c, status[k], "batch field " + std::to_string(k) + " failed");
    return VECSZ_OK;
}

// Fused batch compress: every field shares the block size and dictionary, has
// no padding and a dictionary small enough for the LDS codebooks.  Each stage
// is one launch over a queue of work units that spans every field, so the
// launch count does not grow with the number of fields: range + eb family,
// dual quantization (per-field histograms), one copy of all histograms to
// the host, the codebooks built on up to c->nworkers host threads, one copy
// of all codebooks and headers back, then chunk bits, one scan over every
// chunk of the batch, per-field sizes and the bit packing.  The archives are
// the ones the single-field path writes.
// Issues `enqueue` (asynchronous copies and launches on c->stream only)
// eagerly, or through the graph in slot g.  `key` lists every argument the
// enqueued calls take.  A key seen for the first time runs eagerly; seen
// again right after, the calls are captured (relaxed mode: other threads'
// calls are not affected) and the graph launched; from then on the graph is
// replayed while the key stays the same.  Batches that change from call to
// call therefore never pay for a capture.  Timed calls, and a stream that
// cannot be captured, run eagerly.
template <class F>
static hipError_t run_captured(vecsz_ctx* c, GraphSlot& g, const std::vector<uint64_t>& key, F&& enqueue) {
    if (!c->batch_graph || c->timing) return enqueue();
    if (g.exec && key == g.key) return hipGraphLaunch(g.exec, c->stream);
    if (key != g.seen) {
        g.seen = key;
        return enqueue();
    }
    if (g.exec) {
        (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        g.key.clear();
    }
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
        (void)hipGetLastError();
        return enqueue();
    }
    const hipError_t e = enqueue();
    hipGraph_t gr = nullptr;
    const hipError_t e2 = hipStreamEndCapture(c->stream, &gr);
    if (e != hipSuccess || e2 != hipSuccess || !gr) {
        if (gr) (void)hipGraphDestroy(gr);
        return e != hipSuccess ? e : (e2 != hipSuccess ? e2 : hipErrorUnknown);
    }
    hipGraphExec_t x = nullptr;
    const hipError_t e3 = hipGraphInstantiate(&x, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e3 != hipSuccess) return e3;
    g.exec = x;
    g.key = key;
    return hipGraphLaunch(x, c->stream);
}

static bool batch_fusable(const vecsz_ctx* c, int nf, const vecsz_params* p, const float* const* d_in,
                          uint8_t* const* d_out) {
    if (nf < 2 || nf > 4096) return false;
    for (int k = 0; k < nf; k++) {
        const vecsz_params& q = p[k];
        if (!d_in[k] || !d_out[k]) return false;
        if (q.block != p[0].block || q.dict_size != p[0].dict_size || q.pad_mode != VECSZ_PAD_NONE) return false;
        if (check_params(const_cast<vecsz_ctx*>(c), &q, true)) return false;
    }
    return p[0].dict_size <= kMaxLdsBook;
}

// *reported is set once every field's own status is final; a failure before
// that point (scratch allocation, a HIP error) belongs to the whole batch.
static int compress_batch_fused_run(vecsz_ctx* c, int nf, const vecsz_params* p, const float* const* d_in,
                                    uint8_t* const* d_out, const uint64_t* out_cap, uint64_t* out_bytes,
                                    uint64_t* n_outlier, int* status, bool* reported) {
    HIPCHK(c, hipSetDevice(c->device));
    timing_begin(c);
    const int B = p[0].block, dict = p[0].dict_size, R = dict / 2;
    const float thr_vec = (float)vec_threshold(R);
    // code / outlier-scratch regions: each field starts on a 64-element boundary
    std::vector<uint64_t> off(nf), nel(nf);
    uint64_t ntot = 0;
    for (int k = 0; k < nf; k++) {
        uint64_t n = 0;
        (void)field_elems(p[k].ndim, p[k].dims, &n);  // batch_fusable checked every field
        nel[k] = n;
        off[k] = ntot;
        ntot += (n + 63) / 64 * 64;
    }
    // unit starts: range, chunks (the histogram pass borrows them first), then
    // one dual-quant array per tile form
    const size_t start_bytes = (2 + kDqKinds) * (size_t)(nf + 1) * 8;
    if (!c->codes.ensure(ntot * 2 + 64) || !c->sparse.ensure(ntot * 4 + 64) ||
        !c->b_hist.ensure((size_t)nf * dict * 8) || !c->b_range.ensure((size_t)nf * 8) ||
        !c->b_ebd.ensure((size_t)nf * sizeof(EbDev)) || !c->b_desc.ensure((size_t)nf * sizeof(BatchField)) ||
        !c->b_start.ensure(start_bytes) || !c->b_sums.ensure((size_t)nf * sizeof(EncSummary)) ||
        !c->b_errs.ensure((size_t)nf * 4))
        return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
    const size_t desc_bytes = (size_t)nf * sizeof(BatchField);
    const size_t pin_fixed = desc_bytes + start_bytes + (size_t)nf * 4 + 256;
    if (!c->b_pin.ensure(pin_fixed)) return fail(c, VECSZ_ENOMEM, "pinned buffer");
    // descriptors and unit starts are built here and copied into pinned memory
    // for each upload (the pinned block may be reallocated in between)
    std::vector<BatchField> fd(nf);
    std::vector<uint64_t> st((2 + kDqKinds) * (size_t)(nf + 1), 0);
    uint64_t* rg0 = st.data();
    uint64_t* ck0 = rg0 + (nf + 1);
    uint64_t* dqk = rg0 + 2 * (nf + 1);  // dqk + kind * (nf + 1) for each DqKind
    std::vector<char> dev_range(nf, 0);
    BatchField* fdd = c->b_desc.as<BatchField>();
    for (int k = 0; k < nf; k++) {
        const vecsz_params& q = p[k];
        const int ndim = q.ndim;
        const uint64_t d0 = q.dims[0], d1 = ndim >= 2 ? q.dims[1] : 1, d2 = ndim >= 3 ? q.dims[2] : 1;
        const bool r2r = q.mode == VECSZ_MODE_R2R;
        dev_range[k] = r2r && !(q.value_range > 0);
        BatchField& f = fd[k];
        memset(&f, 0, sizeof(f));
        DQArgs& a = f.a;
        a.d0 = d0;
        a.d1 = d1;
        a.d2 = d2;
        a.n = nel[k];
        a.x8 = (d0 / 8) * 8;
        a.radius = R;
        a.radius_f = (float)R;
        a.thr_vec = thr_vec;
        a.thr_scalar = (float)R;
        a.dict = dict;
        a.reverse = 0;
        a.halo = nullptr;
        a.hb0 = ref_nblk(d0, B);
        a.hb1 = ndim >= 2 ? ref_nblk(d1, B) : 1;
        const bool x4 = dualquant_use_x4(ndim, B, d_in[k], d0, d1, d2, true);
        a.ntiles_x = x4 ? (d0 + 255) / 256 : (d0 + 63) / 64;
        a.ntiles_y = ndim >= 2 ? (d1 + B - 1) / B : 1;
        a.ntiles_z = ndim >= 3 ? (d2 + B - 1) / B : 1;
        if (dev_range[k]) {
            f.range = c->b_range.as<uint32_t>() + 2 * k;
            f.ebd = c->b_ebd.as<EbDev>() + k;
            a.ebd = f.ebd;
        } else {
            const EbFamily eb = eb_family(q.eb, r2r, r2r ? q.value_range : 0.0);
            a.ebd = nullptr;
            a.r_f = eb.r_f;
            a.r_d = eb.r_d;
        }
        f.in = d_in[k];
        f.codes = c->codes.as<uint16_t>() + off[k];
        f.sparse = c->sparse.as<float>() + off[k];
        f.hist = c->b_hist.as<unsigned long long>() + (size_t)k * dict;
        f.eb = q.eb;
        const bool aln = dualquant_x4_aligned(d_in[k], d0);
        f.kind = ndim == 1 ? kDq1
                 : ndim == 2 ? (x4 ? (aln ? kDq2x4 : kDq2x4u) : kDq2)
                             : (x4 ? (aln ? kDq3x4 : kDq3x4u) : kDq3);
        f.dq_tiles = ndim == 1 ? ((nel[k] + 7) / 8 + 63) / 64 : a.ntiles_x * a.ntiles_y * a.ntiles_z;
        f.archive = d_out[k];
        f.out_cap = out_cap[k];
        f.meta_bytes = meta_bytes(ndim);
        rg0[k + 1] = rg0[k] + (dev_range[k] ? range_units(nel[k]) : 0);
        for (int kind = 0; kind < kDqKinds; kind++) {
            uint64_t* s0 = dqk + (size_t)kind * (nf + 1);
            s0[k + 1] = s0[k] + (kind == f.kind ? (f.dq_tiles + 3) / 4 : 0);
        }
        status[k] = VECSZ_OK;
        out_bytes[k] = 0;
        n_outlier[k] = 0;
    }
    // Partial histogram rows of the code histogram pass: workgroup g writes
    // one row per field its unit range [U g / G, U (g + 1) / G) touches, so
    // each field's rows are contiguous; they are summed in groups of up to 32.
    uint64_t* hs0 = ck0;  // the chunk starts are filled later: borrow their slots
    hs0[0] = 0;
    for (int k = 0; k < nf; k++) hs0[k + 1] = hs0[k] + hist_units(nel[k]);
    const uint64_t U = hs0[nf];
    const unsigned G = hist_multi_grid(U);
    std::vector<uint32_t> rowtab(G);
    std::vector<uint32_t> frow(nf, 0), fcnt(nf, 0);
    uint32_t nrows = 0;
    for (unsigned g = 0; g < G; g++) {
        const uint64_t u0 = U * g / G, u1 = U * (g + 1) / G;
        rowtab[g] = nrows;
        if (u0 >= u1) continue;
        const int f0 = (int)(std::upper_bound(hs0, hs0 + nf + 1, u0) - hs0) - 1;
        const int f1 = (int)(std::upper_bound(hs0, hs0 + nf + 1, u1 - 1) - hs0) - 1;
        for (int f = f0; f <= f1; f++) {
            if (fcnt[f]++ == 0) frow[f] = nrows;
            nrows++;
        }
    }
    for (int f = 0; f < nf; f++)
        for (uint32_t r = 0; r < fcnt[f]; r += 32) {
            rowtab.push_back((uint32_t)f);
            rowtab.push_back(frow[f] + r);
            rowtab.push_back(std::min<uint32_t>(32, fcnt[f] - r));
        }
    const uint32_t ngroups = (uint32_t)((rowtab.size() - G) / 3);
    if (!c->partials.ensure((size_t)nrows * dict * 4 + 64) || !c->b_rows.ensure(rowtab.size() * 4 + 64))
        return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
    if (!c->b_pin.ensure(pin_fixed + rowtab.size() * 4 + 64)) return fail(c, VECSZ_ENOMEM, "pinned buffer");
    memcpy(c->b_pin.p, fd.data(), desc_bytes);
    memcpy(c->b_pin.as<uint8_t>() + desc_bytes, st.data(), start_bytes);
    memcpy(c->b_pin.as<uint8_t>() + pin_fixed, rowtab.data(), rowtab.size() * 4);
    const uint64_t* rg0_d = c->b_start.as<uint64_t>();
    const uint64_t* ck0_d = rg0_d + (nf + 1);
    const uint64_t* dqk_d = rg0_d + 2 * (nf + 1);
    const size_t hist_bytes = (size_t)nf * dict * 8;
    if (!c->b_pin_hist.ensure(hist_bytes + (size_t)nf * sizeof(EbDev) + 64))
        return fail(c, VECSZ_ENOMEM, "pinned buffer");
    uint8_t* hp = c->b_pin_hist.as<uint8_t>();
    // uploads, range, dual quantization, histograms and the copy down: what
    // the graph before the stop holds
    auto enqueue_pre = [&]() -> hipError_t {
        hipError_t e;
#define ENQ(expr) \
    if ((e = (expr)) != hipSuccess) return e
        ENQ(hipMemcpyAsync(c->b_desc.p, c->b_pin.p, desc_bytes, hipMemcpyHostToDevice, c->stream));
        ENQ(hipMemcpyAsync(c->b_start.p, c->b_pin.as<uint8_t>() + desc_bytes, start_bytes, hipMemcpyHostToDevice,
                           c->stream));
        ENQ(hipMemcpyAsync(c->b_rows.p, c->b_pin.as<uint8_t>() + pin_fixed, rowtab.size() * 4, hipMemcpyHostToDevice,
                           c->stream));
        {
            StageScope s(c, "range");
            launch_range_multi(fdd, nf, rg0_d, rg0[nf], c->stream);
        }
        // (a kernel, not hipMemsetAsync: this may be captured)
        launch_zero_u32(reinterpret_cast<uint32_t*>(c->b_hist.p), (uint64_t)nf * dict * 2, c->stream);
        static const char* const kDqStage[kDqKinds] = {"dualquant_1d", "dualquant_2d",     "dualquant_2d_x4",
                                                       "dualquant_3d", "dualquant_3d_x4", "dualquant_2d_x4u",
                                                       "dualquant_3d_x4u"};
        for (int kind = 0; kind < kDqKinds; kind++) {
            const uint64_t nu = dqk[(size_t)kind * (nf + 1) + nf];
            if (!nu) continue;
            StageScope s(c, kDqStage[kind]);
            launch_dualquant_multi(B, kind, fdd, nf, dqk_d + (size_t)kind * (nf + 1), nu, c->dq_grid, c->stream);
        }
        {
            StageScope s(c, "histogram");
            const uint32_t* rows_d = c->b_rows.as<uint32_t>();
            launch_hist_multi(fdd, nf, ck0_d, U, dict, c->partials.as<uint32_t>(), rows_d, rows_d + G, ngroups,
                              c->b_hist.as<unsigned long long>(), c->stream);
        }
        ENQ(hipGetLastError());
        ENQ(hipMemcpyAsync(hp, c->b_hist.p, hist_bytes, hipMemcpyDeviceToHost, c->stream));
        ENQ(hipMemcpyAsync(hp + hist_bytes, c->b_ebd.p, (size_t)nf * sizeof(EbDev), hipMemcpyDeviceToHost,
                           c->stream));
#undef ENQ
        return hipSuccess;
    };
    {
        auto ptr = [](const void* q) { return (uint64_t)(uintptr_t)q; };
        std::vector<uint64_t> key = {ptr(c->stream),  ptr(c->b_desc.p),   ptr(c->b_start.p), ptr(c->b_rows.p),
                                     ptr(c->b_hist.p), ptr(c->b_ebd.p),    ptr(c->partials.p), ptr(c->b_pin.p),
                                     ptr(c->b_pin_hist.p), (uint64_t)nf,  (uint64_t)dict,    (uint64_t)B,
                                     desc_bytes,       start_bytes,        pin_fixed,         rowtab.size(),
                                     rg0[nf],          U,                  G,                 ngroups,
                                     (uint64_t)c->dq_grid};
        for (int kind = 0; kind < kDqKinds; kind++) key.push_back(dqk[(size_t)kind * (nf + 1) + nf]);
        HIPCHK(c, run_captured(c, c->b_graph[0], key, enqueue_pre));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));

    // codebooks on host threads
    if ((int)c->b_cb.size() < nf) c->b_cb.resize(nf);
    for (int k = 0; k < nf; k++) {
        if (!dev_range[k]) continue;
        EbDev e;
        memcpy(&e, hp + hist_bytes + (size_t)k * sizeof(EbDev), sizeof(e));
        if (e.zero_range) status[k] = VECSZ_ERANGE0;
    }
    {
        HostTimer ht(c, "host_codebook");
        std::atomic<int> next{0};
        auto body = [&]() {
            for (int k; (k = next.fetch_add(1)) < nf;)
                if (status[k] == VECSZ_OK)
                    status[k] = build_codebook(reinterpret_cast<const uint64_t*>(hp) + (size_t)k * dict, dict,
                                               c->b_cb[k]);
        };
        // host threads: at least the worker count, up to 16 where the host has them
        const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
        const int nt = std::max(1, std::min(nf, std::max(c->nworkers, std::min(16, hw))));
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(body);
        body();
        for (auto& t : th) t.join();
    }

    // headers and codebooks, staged in one pinned block
    const size_t book_stride = ((size_t)dict * 9 + 15) / 16 * 16;
    std::vector<uint64_t> hoff(nf, 0);
    uint64_t head_tot = 0;
    for (int k = 0; k < nf; k++) {
        if (status[k]) continue;
        const uint64_t hb = fd[k].meta_bytes + 8 + c->b_cb[k].tree.size();
        if (hb > out_cap[k]) {
            out_bytes[k] = hb;
            status[k] = VECSZ_ENOMEM;
            continue;
        }
        hoff[k] = head_tot;
        head_tot += (hb + 15) / 16 * 16;
    }
    const size_t books_bytes = book_stride * nf;
    if (!c->b_pin.ensure(pin_fixed + books_bytes + head_tot + 64)) return fail(c, VECSZ_ENOMEM, "pinned buffer");
    BatchField* fd_h = c->b_pin.as<BatchField>();
    uint64_t* ck0_h = reinterpret_cast<uint64_t*>(c->b_pin.as<uint8_t>() + desc_bytes);
    uint32_t* errs_h = reinterpret_cast<uint32_t*>(ck0_h + (nf + 1));
    uint8_t* books_h = c->b_pin.as<uint8_t>() + pin_fixed;
    uint8_t* heads_h = books_h + books_bytes;
    if (!c->b_book.ensure(books_bytes + 64) || !c->b_head.ensure(head_tot + 64))
        return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
    uint64_t ctot = 0;
    int max_len = 1;
    ck0[0] = 0;
    for (int k = 0; k < nf; k++) {
        BatchField& f = fd[k];
        errs_h[k] = status[k] ? 2u : 0u;
        f.nchunks = 0;
        f.chunk0 = ctot;
        if (!status[k]) {
            const Codebook& cb = c->b_cb[k];
            const vecsz_params& q = p[k];
            const uint64_t mb = f.meta_bytes, tb = cb.tree.size();
            uint8_t* hdr = heads_h + hoff[k];
            memset(hdr, 0, mb);
            int32_t nd = q.ndim;
            memcpy(&hdr[0], &nd, 4);
            for (int i = 0; i < q.ndim; i++) {
                int32_t d = (int32_t)q.dims[i];
                memcpy(&hdr[4 + 4 * i], &d, 4);
            }
            memcpy(&hdr[4 * (q.ndim + 1)], &q.eb, 8);  // pad_len (next 8 bytes) stays 0
            put_u32be(&hdr[mb], (uint32_t)cb.node_count);
            put_u32be(&hdr[mb + 4], (uint32_t)dict);
            memcpy(&hdr[mb + 8], cb.tree.data(), tb);
            uint8_t* bk = books_h + book_stride * k;
            memcpy(bk, cb.cw_hi.data(), (size_t)dict * 4);
            memcpy(bk + (size_t)dict * 4, cb.cw_lo.data(), (size_t)dict * 4);
            memcpy(bk + (size_t)dict * 8, cb.len.data(), (size_t)dict);
            f.book = reinterpret_cast<const uint32_t*>(c->b_book.as<uint8_t>() + book_stride * k);
            f.head = c->b_head.as<uint8_t>() + hoff[k];
            f.head_bytes = mb + 8 + tb;
            f.tree_bytes = tb;
            f.nchunks = (nel[k] + kChunk - 1) / kChunk;
            max_len = std::max(max_len, cb.max_len);
        }
        ctot += f.nchunks;
        ck0[k + 1] = ck0[k] + (f.nchunks + 3) / 4;
    }
    if (ctot) {
        if (!c->chunk_bits.ensure(ctot * 4) || !c->chunk_zeros.ensure(ctot * 4) || !c->bit_off.ensure(ctot * 8) ||
            !c->out_off.ensure(ctot * 8) || !c->scan_tmp.ensure(((ctot + 4095) / 4096 + 1) * 16) ||
            !c->scan_tot.ensure(16))
            return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
        memcpy(fd_h, fd.data(), desc_bytes);
        memcpy(ck0_h, ck0, (size_t)(nf + 1) * 8);
        uint32_t* errs_d = c->b_errs.as<uint32_t>();
        EncSummary* sums_h = reinterpret_cast<EncSummary*>(c->b_pin_hist.as<uint8_t>());
        uint32_t* errs_back = reinterpret_cast<uint32_t*>(c->b_pin_hist.as<uint8_t>() + (size_t)nf * sizeof(EncSummary));
        // uploads, launches and the copies back: what the graph holds
        auto enqueue = [&]() -> hipError_t {
            hipError_t e;
#define ENQ(expr) \
    if ((e = (expr)) != hipSuccess) return e
            ENQ(hipMemcpyAsync(c->b_book.p, books_h, books_bytes, hipMemcpyHostToDevice, c->stream));
            ENQ(hipMemcpyAsync(c->b_head.p, heads_h, head_tot, hipMemcpyHostToDevice, c->stream));
            ENQ(hipMemcpyAsync(c->b_desc.p, fd_h, desc_bytes, hipMemcpyHostToDevice, c->stream));
            ENQ(hipMemcpyAsync((void*)ck0_d, ck0_h, (size_t)(nf + 1) * 8, hipMemcpyHostToDevice, c->stream));
            ENQ(hipMemcpyAsync(errs_d, errs_h, (size_t)nf * 4, hipMemcpyHostToDevice, c->stream));
            launch_place_heads(fdd, nf, errs_d, c->stream);
            {
                StageScope s(c, "chunk_bits");
                launch_chunk_bits_multi(fdd, nf, ck0_d, ck0[nf], dict, c->chunk_bits.as<uint32_t>(),
                                        c->chunk_zeros.as<uint32_t>(), c->enc_grid, c->stream);
            }
            {
                StageScope s(c, "scan");
                launch_scan_u32x2(c->chunk_bits.as<uint32_t>(), c->chunk_zeros.as<uint32_t>(), ctot,
                                  c->bit_off.as<uint64_t>(), c->out_off.as<uint64_t>(), c->scan_tmp.as<uint64_t>(),
                                  c->scan_tot.as<uint64_t>(), c->stream);
                launch_enc_summary_multi(fdd, nf, c->bit_off.as<uint64_t>(), c->out_off.as<uint64_t>(),
                                         c->scan_tot.as<uint64_t>(), ctot, c->b_sums.as<EncSummary>(), errs_d,
                                         c->stream);
            }
            {
                StageScope s(c, "pack");
                launch_pack_multi(fdd, nf, ck0_d, ck0[nf], dict, max_len, c->bit_off.as<uint64_t>(),
                                  c->out_off.as<uint64_t>(), c->b_sums.as<EncSummary>(), errs_d, c->enc_grid,
                                  c->stream);
            }
            ENQ(hipGetLastError());
            ENQ(hipMemcpyAsync(sums_h, c->b_sums.p, (size_t)nf * sizeof(EncSummary), hipMemcpyDeviceToHost,
                               c->stream));
            ENQ(hipMemcpyAsync(errs_back, errs_d, (size_t)nf * 4, hipMemcpyDeviceToHost, c->stream));
#undef ENQ
            return hipSuccess;
        };
        auto ptr = [](const void* q) { return (uint64_t)(uintptr_t)q; };
        const std::vector<uint64_t> key = {
            ptr(c->stream),        ptr(c->b_book.p),     ptr(c->b_head.p),  ptr(c->b_desc.p),  ptr(c->b_start.p),
            ptr(c->b_errs.p),      ptr(c->chunk_bits.p), ptr(c->chunk_zeros.p), ptr(c->bit_off.p), ptr(c->out_off.p),
            ptr(c->scan_tmp.p),    ptr(c->scan_tot.p),   ptr(c->b_sums.p),  ptr(c->b_pin.p),   ptr(c->b_pin_hist.p),
            (uint64_t)nf,          (uint64_t)dict,       (uint64_t)max_len, pin_fixed,         desc_bytes,
            books_bytes,           head_tot,             ck0[nf],           ctot,              (uint64_t)c->enc_grid};
        HIPCHK(c, run_captured(c, c->b_graph[1], key, enqueue));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < nf; k++) {
            if (status[k]) continue;
            out_bytes[k] = sums_h[k].archive_size;
            n_outlier[k] = sums_h[k].n_outlier;
            if (errs_back[k]) status[k] = VECSZ_ENOMEM;
        }
    }
    timing_end(c);
    *reported = true;
    for (int k = 0; k < nf; k++)
        if (status[k]) {
            const char* why = status[k] == VECSZ_ERANGE0   ? "r2r value range is zero"
                              : status[k] == VECSZ_ENOMEM  ? "output buffer too small"
                              : status[k] == VECSZ_ECODELEN ? "Huffman code longer than 64 bits"
                                                            : "codebook construction failed";
            return fail(c, status[k], "batch field " + std::to_string(k) + " failed: " + why);
        }
    return VECSZ_OK;
}

// Every field gets a status: on a failure of the whole batch, the fields
// not already failed on their own take the batch's status.
static int compress_batch_fused(vecsz_ctx* c, int nf, const vecsz_params* p, const float* const* d_in,
                                uint8_t* const* d_out, const uint64_t* out_cap, uint64_t* out_bytes,
                                uint64_t* n_outlier, int* status) {
    for (int k = 0; k < nf; k++) status[k] = VECSZ_OK;
    bool reported = false;
    const int rc = compress_batch_fused_run(c, nf, p, d_in, d_out, out_cap, out_bytes, n_outlier, status, &reported);
    if (rc && !reported)
        for (int k = 0; k < nf; k++)
            if (status[k] == VECSZ_OK) status[k] = rc;
    return rc;
}

// Fused batch decompress: every field shares the block size and a dictionary
// of at most kMaxLdsBook, and has no pad section.  One copy brings every
// archive's metadata, Huffman header and tree to the host (the tree of such
// a dictionary is at most tree_byte_size(2 dict - 1) bytes), the decode
// tables are built on host threads and go back in one copy, and then each
// stage is one launch over all fields: the sync pass, the re-decode passes
// over one batch-wide list, one scan over every subsequence of the batch,
// the write pass, and the reconstruction (one launch per tile form) with its
// fp32 fallback.  Outputs are the single-field path's.
static bool dbatch_fusable(const vecsz_ctx* c, int nf, const vecsz_params* p, const uint8_t* const* d_in,
                           float* const* d_out) {
    if (nf < 2 || nf > 4096) return false;
    for (int k = 0; k < nf; k++) {
        if (!d_in[k] || !d_out[k]) return false;
        if (p[k].block != p[0].block || p[k].dict_size != p[0].dict_size) return false;
        if (check_params(const_cast<vecsz_ctx*>(c), &p[k], false)) return false;
    }
    return p[0].dict_size <= kMaxLdsBook;
}

// *fallback: an archive has a pad section; nothing was written and the
// caller takes the per-field path instead.
static int decompress_batch_fused_run(vecsz_ctx* c, int nf, const vecsz_params* p, const uint8_t* const* d_in,
                                      const uint64_t* in_bytes, float* const* d_out, const uint64_t* out_cap,
                                      int* status, bool* reported, bool* fallback) {
    HIPCHK(c, hipSetDevice(c->device));
    timing_begin(c);
    const int B = p[0].block, dict = p[0].dict_size, R = dict / 2;

    // 1. the head of every archive (metadata, Huffman header, tree) in one copy
    const uint64_t head_max = meta_bytes(3) + 8 + tree_byte_size(2 * (uint64_t)dict - 1);
    std::vector<HeadCopy> hc(nf);
    uint64_t htot = 0;
    for (int k = 0; k < nf; k++) {
        hc[k].src = d_in[k];
        hc[k].len = std::min<uint64_t>(in_bytes[k], head_max);
        hc[k].off = htot;
        htot += (hc[k].len + 15) / 16 * 16;
    }
    const size_t hc_bytes = (size_t)nf * sizeof(HeadCopy);
    if (!c->db_hc.ensure(hc_bytes) || !c->db_heads.ensure(htot + 64))
        return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
    if (!c->db_pin.ensure(hc_bytes + 64) || !c->db_hpin.ensure(htot + 64)) return fail(c, VECSZ_ENOMEM, "pinned buffer");
    memcpy(c->db_pin.p, hc.data(), hc_bytes);
    HIPCHK(c, hipMemcpyAsync(c->db_hc.p, c->db_pin.p, hc_bytes, hipMemcpyHostToDevice, c->stream));
    launch_gather_heads(c->db_hc.as<HeadCopy>(), nf, c->db_heads.as<uint8_t>(), c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->db_hpin.p, c->db_heads.p, htot, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));

    // 2. headers, then the decode tables on host threads
    struct Hdr {
        int ndim;
        uint64_t dims[3], n, pad_len, huff_size, nout, tb;
        double ebu;
    };
    std::vector<Hdr> hd(nf);
    const uint8_t* heads = c->db_hpin.as<uint8_t>();
    for (int k = 0; k < nf; k++) {
        Hdr& h = hd[k];
        const uint8_t* hp = heads + hc[k].off;
        if (parse_meta(hp, hc[k].len, in_bytes[k], &h.ndim, h.dims, &h.ebu, &h.pad_len, &h.huff_size, &h.nout)) {
            status[k] = VECSZ_ECORRUPT;
            continue;
        }
        if (h.pad_len) {
            *fallback = true;
            return VECSZ_OK;
        }
        h.n = h.dims[0] * h.dims[1] * h.dims[2];
        const uint64_t mb = meta_bytes(h.ndim);
        if (h.n > out_cap[k]) {
            status[k] = VECSZ_ENOMEM;
        } else if (p[k].mode == VECSZ_MODE_R2R && !(p[k].value_range != 0)) {
            status[k] = VECSZ_EINVAL;
        } else if (h.huff_size < 8 || mb + 8 > hc[k].len) {
            status[k] = VECSZ_ECORRUPT;
        } else {
            const uint8_t* q = hp + mb;
            const int32_t ncs = (int32_t)(((uint32_t)q[0] << 24) | ((uint32_t)q[1] << 16) | ((uint32_t)q[2] << 8) | q[3]);
            h.tb = ncs > 0 ? tree_byte_size((uint64_t)ncs) : 0;
            // a tree beyond what the dictionary allows cannot be in the head read
            if (ncs <= 0 || 8 + h.tb > h.huff_size || mb + 8 + h.tb > hc[k].len) status[k] = VECSZ_ECORRUPT;
        }
    }
    if ((int)c->db_dt.size() < nf) c->db_dt.resize(nf);
    {
        HostTimer ht(c, "host_tables");
        std::atomic<int> next{0};
        auto body = [&]() {
            for (int k; (k = next.fetch_add(1)) < nf;) {
                if (status[k]) continue;
                DecodeTables& dt = c->db_dt[k];
                if (build_decode_tables(heads + hc[k].off + meta_bytes(hd[k].ndim), hd[k].huff_size, dt) ||
                    (!dt.root_leaf && !tree_is_full(dt)))
                    status[k] = VECSZ_ECORRUPT;
            }
        };
        const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
        const int nt = std::max(1, std::min(nf, std::max(c->nworkers, std::min(16, hw))));
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(body);
        body();
        for (auto& t : th) t.join();
    }

    // 3. layout: code / outlier regions (64-element aligned), tables,
    // subsequences, fallback block lists, unit and tile starts
    bool t16 = true;
    uint32_t max_sub16 = 0;
    for (int k = 0; k < nf; k++) {
        if (status[k] || c->db_dt[k].root_leaf) continue;
        const DecodeTables& dt = c->db_dt[k];
        if (dt.lut16.empty()) t16 = false;
        if (dt.sub16.size() <= kMaxSub16) max_sub16 = std::max<uint32_t>(max_sub16, (uint32_t)dt.sub16.size());
    }
    max_sub16 = batch_sub16_in_lds(max_sub16);
    std::vector<uint64_t> off(nf, 0), toff(nf, 0), fl_off(nf, 0);
    uint64_t ntot = 0, tab_tot = 0, nsub_tot = 0, fl_tot = 0;
    const size_t ns = (size_t)nf + 1;
    // starts: sync units, write units, subsequences, then one array per tile form
    std::vector<uint64_t> st((3 + kRcKinds) * ns, 0);
    uint64_t* su = st.data();
    uint64_t* wu = su + ns;
    uint64_t* sb = wu + ns;
    uint64_t* rk = sb + ns;  // rk + kind * ns
    std::vector<DecField> fd(nf);
    for (int k = 0; k < nf; k++) {
        DecField& F = fd[k];
        memset(&F, 0, sizeof(F));
        F.kind = -1;
        su[k + 1] = su[k];
        wu[k + 1] = wu[k];
        sb[k + 1] = sb[k];
        for (int kind = 0; kind < kRcKinds; kind++) rk[kind * ns + k + 1] = rk[kind * ns + k];
        if (status[k]) continue;
        const Hdr& h = hd[k];
        const DecodeTables& dt = c->db_dt[k];
        off[k] = ntot;
        ntot += (h.n + 63) / 64 * 64;
        F.ndim = h.ndim;
        F.a.n = h.n;
        F.a.n_outlier = h.nout;
        const uint8_t* d_huff = d_in[k] + meta_bytes(h.ndim);
        F.a.outliers = d_huff + h.huff_size;
        if (!dt.root_leaf) {
            const uint64_t stream_bytes = h.huff_size - 8 - h.tb;
            F.a.stream = d_huff + 8 + h.tb;
            F.a.stream_bits = stream_bytes * 8;
            F.a.nsub = std::max<uint64_t>(1, (F.a.stream_bits + kSubBits - 1) / kSubBits);
            toff[k] = tab_tot;
            tab_tot += t16 ? ((dt.lut16.size() + dt.sub16.size()) * 2 + 15) / 16 * 16
                           : ((dt.lut.size() + dt.sub.size()) * 4 + 15) / 16 * 16;
            F.s.jbase = (uint32_t)sb[k];
            su[k + 1] = su[k] + decode_sync_units(F.a.nsub);
            wu[k + 1] = wu[k] + decode_write_units(F.a.nsub);
            sb[k + 1] = sb[k] + F.a.nsub;
        } else if (dt.root_sym == 0 && h.nout < h.n) {
            status[k] = VECSZ_ECORRUPT;  // every element an outlier, but fewer outliers than elements
            continue;
        }
        ReconArgs& ra = F.ra;
        ra.d0 = h.dims[0];
        ra.d1 = h.dims[1];
        ra.d2 = h.dims[2];
        ra.n = h.n;
        ra.radius = R;
        ra.twoeb = eb_family(h.ebu, p[k].mode == VECSZ_MODE_R2R, p[k].value_range).twoeb;
        ra.ntiles_x = (ra.d0 + 63) / 64;
        ra.ntiles_y = h.ndim >= 2 ? (ra.d1 + B - 1) / B : 1;
        ra.ntiles_z = h.ndim >= 3 ? (ra.d2 + B - 1) / B : 1;
        ra.halo = nullptr;
        ra.halo_stride = 0;
        ra.hb0 = ref_nblk(ra.d0, B);
        ra.hb1 = h.ndim >= 2 ? ref_nblk(ra.d1, B) : 1;
        // (code regions start on 64-element boundaries: aligned like a null pointer)
        F.kind = recon_kind(h.ndim, B, ra.d0, nullptr, d_out[k]);
        rk[F.kind * ns + k + 1] = rk[F.kind * ns + k] + recon_tiles(F.kind, ra);
        fl_off[k] = fl_tot;
        fl_tot += ((ra.d0 + B - 1) / B) * (h.ndim >= 2 ? (ra.d1 + B - 1) / B : 1) * (h.ndim >= 3 ? (ra.d2 + B - 1) / B : 1);
        F.out = d_out[k];
    }
    nsub_tot = sb[nf];
    if (nsub_tot > 0xFFFFFFF0ull) return fail(c, VECSZ_EINVAL, "batch too large");  // u32 list entries
    const size_t st_bytes = st.size() * 8, fd_bytes = (size_t)nf * sizeof(DecField);
    const uint64_t nsub_alloc = std::max<uint64_t>(nsub_tot, 1);
    if (!c->codes.ensure(ntot * 2 + 64) || !c->sparse.ensure(ntot * 4 + 64) || !c->tabs.ensure(tab_tot + 64) ||
        !c->sub_b.ensure(nsub_alloc * 8) || !c->sub_e.ensure(nsub_alloc * 8) || !c->sub_cnt.ensure(nsub_alloc * 4) ||
        !c->sub_zeros.ensure(nsub_alloc * 4) || !c->sub_lj.ensure(nsub_alloc * 8) ||
        !c->sub_ls.ensure(nsub_alloc * 16) || !c->sub_ln.ensure(16) || !c->sym_off.ensure(nsub_alloc * 8) ||
        !c->zero_off.ensure(nsub_alloc * 8) || !c->scan_tmp.ensure(((nsub_alloc + 4095) / 4096 + 1) * 16) ||
        !c->scan_tot.ensure(16) || !c->flagged.ensure(fl_tot * 4 + 64) || !c->db_nfl.ensure((size_t)nf * 4) ||
        !c->db_sums.ensure((size_t)nf * sizeof(DecSummary)) || !c->db_tot.ensure((size_t)nf * 8) ||
        !c->db_desc.ensure(fd_bytes) || !c->db_start.ensure(st_bytes))
        return fail(c, VECSZ_ENOMEM, "device scratch allocation failed");
    if (!c->db_pin.ensure(fd_bytes + st_bytes + tab_tot + 64) ||
        !c->db_back.ensure((size_t)nf * (sizeof(DecSummary) + 8) + 64))
        return fail(c, VECSZ_ENOMEM, "pinned buffer");
    SubState S{};
    S.b = c->sub_b.as<uint64_t>();
    S.e = c->sub_e.as<uint64_t>();
    S.cnt = c->sub_cnt.as<uint32_t>();
    S.zer = c->sub_zeros.as<uint32_t>();
    S.list_j[0] = c->sub_lj.as<uint32_t>();
    S.list_j[1] = c->sub_lj.as<uint32_t>() + nsub_alloc;
    S.list_start[0] = c->sub_ls.as<uint64_t>();
    S.list_start[1] = c->sub_ls.as<uint64_t>() + nsub_alloc;
    S.list_n = c->sub_ln.as<uint32_t>();
    S.jbase = 0;
    uint8_t* tab_h = c->db_pin.as<uint8_t>() + fd_bytes + st_bytes;
    for (int k = 0; k < nf; k++) {
        DecField& F = fd[k];
        if (status[k]) continue;
        const DecodeTables& dt = c->db_dt[k];
        F.codes = c->codes.as<uint16_t>() + off[k];
        F.sparse = c->sparse.as<float>() + off[k];
        F.flagged = c->flagged.as<uint32_t>() + fl_off[k];
        F.nflagged = c->db_nfl.as<uint32_t>() + k;
        F.sum = c->db_sums.as<DecSummary>() + k;
        const uint64_t s0 = sb[k];
        F.s.b = S.b + s0;
        F.s.e = S.e + s0;
        F.s.cnt = S.cnt + s0;
        F.s.zer = S.zer + s0;
        for (int q = 0; q < 2; q++) {
            F.s.list_j[q] = S.list_j[q];
            F.s.list_start[q] = S.list_start[q];
        }
        F.s.list_n = S.list_n;
        F.sym_off = c->sym_off.as<uint64_t>() + s0;
        F.zero_off = c->zero_off.as<uint64_t>() + s0;
        if (dt.root_leaf) continue;
        uint8_t* tp = tab_h + toff[k];
        uint8_t* tdev = c->tabs.as<uint8_t>() + toff[k];
        if (t16) {
            memcpy(tp, dt.lut16.data(), dt.lut16.size() * 2);
            memcpy(tp + dt.lut16.size() * 2, dt.sub16.data(), dt.sub16.size() * 2);
            F.a.lut16 = reinterpret_cast<const uint16_t*>(tdev);
            F.a.sub16 = reinterpret_cast<const uint16_t*>(tdev) + dt.lut16.size();
            F.a.n_sub16 = max_sub16 && dt.sub16.size() <= max_sub16 ? (uint32_t)dt.sub16.size() : 0;
        } else {
            memcpy(tp, dt.lut.data(), dt.lut.size() * 4);
            memcpy(tp + dt.lut.size() * 4, dt.sub.data(), dt.sub.size() * 4);
            F.a.lut = reinterpret_cast<const uint32_t*>(tdev);
            F.a.sub = reinterpret_cast<const uint32_t*>(tdev) + dt.lut.size();
        }
    }
    DecField* fdd = c->db_desc.as<DecField>();
    uint64_t* std_ = c->db_start.as<uint64_t>();
    const uint64_t* su_d = std_;
    const uint64_t* wu_d = std_ + ns;
    const uint64_t* sb_d = std_ + 2 * ns;
    const uint64_t* rk_d = std_ + 3 * ns;
    memcpy(c->db_pin.p, fd.data(), fd_bytes);
    memcpy(c->db_pin.as<uint8_t>() + fd_bytes, st.data(), st_bytes);
    HIPCHK(c, hipMemcpyAsync(c->db_desc.p, c->db_pin.p, fd_bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->db_start.p, c->db_pin.as<uint8_t>() + fd_bytes, st_bytes, hipMemcpyHostToDevice,
                             c->stream));
    if (tab_tot) HIPCHK(c, hipMemcpyAsync(c->tabs.p, tab_h, tab_tot, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->db_sums.p, 0, (size_t)nf * sizeof(DecSummary), c->stream));
    HIPCHK(c, hipMemsetAsync(c->db_nfl.p, 0, (size_t)nf * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->sub_ln.p, 0, 16, c->stream));

    // 4. decode: one-symbol trees by a fill; the others by the batch passes
    {
        StageScope s(c, "decode_sync");
        for (int k = 0; k < nf; k++) {
            if (status[k] || !c->db_dt[k].root_leaf) continue;
            launch_fill_u16(fd[k].codes, hd[k].n, (uint16_t)c->db_dt[k].root_sym, c->stream);
            if (c->db_dt[k].root_sym == 0) launch_scatter_all_outliers(fd[k].a.outliers, hd[k].n, fd[k].sparse, c->stream);
        }
        launch_decode_sync_multi(t16, max_sub16, fdd, su_d, nf, su[nf], c->stream);
    }
    DecSummary* sums_h = c->db_back.as<DecSummary>();
    uint64_t* tot_h = reinterpret_cast<uint64_t*>(c->db_back.as<uint8_t>() + (size_t)nf * sizeof(DecSummary));
    uint32_t* left_h = reinterpret_cast<uint32_t*>(tot_h + nf);
    uint64_t max_nsub = 0;
    for (int k = 0; k < nf; k++) max_nsub = std::max<uint64_t>(max_nsub, fd[k].a.nsub);
    if (nsub_tot) {
        constexpr uint32_t kPassesAhead = 2;
        uint32_t pass = 0;
        {
            StageScope s(c, "decode_redo");
            for (; pass < kPassesAhead; pass++) launch_decode_pass_multi(t16, fdd, sb_d, nf, S, pass, nsub_tot, c->stream);
        }
        HIPCHK(c, hipGetLastError());
        uint32_t batch = 2;
        for (;;) {
            HIPCHK(c, hipMemcpyAsync(left_h, S.list_n + pass % 3, 4, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (*left_h == 0) break;
            if (pass > max_nsub + 2) return fail(c, VECSZ_ECORRUPT, "subsequence boundaries do not converge");
            StageScope s(c, "decode_redo");
            for (uint32_t b = 0; b < batch; b++) launch_decode_pass_multi(t16, fdd, sb_d, nf, S, pass++, nsub_tot, c->stream);
            batch = std::min<uint32_t>(2 * batch, 16);
        }
        // fields whose stream does not decode are left out of the write pass
        // and the reconstruction (their subsequence state is not trusted)
        HIPCHK(c, hipMemcpyAsync(sums_h, c->db_sums.p, (size_t)nf * sizeof(DecSummary), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        bool changed = false;
        for (int k = 0; k < nf; k++) {
            if (status[k] || !sums_h[k].error) continue;
            status[k] = VECSZ_ECORRUPT;
            changed = true;
        }
        if (changed) {
            for (int k = 0; k < nf; k++) {
                wu[k + 1] = wu[k] + (status[k] ? 0 : decode_write_units(fd[k].a.nsub));
                for (int kind = 0; kind < kRcKinds; kind++)
                    rk[kind * ns + k + 1] = rk[kind * ns + k] + (!status[k] && fd[k].kind == kind ? recon_tiles(kind, fd[k].ra) : 0);
                if (status[k]) fd[k].kind = -1;
            }
            memcpy(c->db_pin.p, fd.data(), fd_bytes);
            memcpy(c->db_pin.as<uint8_t>() + fd_bytes, st.data(), st_bytes);
            HIPCHK(c, hipMemcpyAsync(c->db_desc.p, c->db_pin.p, fd_bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->db_start.p, c->db_pin.as<uint8_t>() + fd_bytes, st_bytes,
                                     hipMemcpyHostToDevice, c->stream));
        }
        {
            StageScope s(c, "decode_scan");
            launch_scan_u32x2(S.cnt, S.zer, nsub_tot, c->sym_off.as<uint64_t>(), c->zero_off.as<uint64_t>(),
                              c->scan_tmp.as<uint64_t>(), c->scan_tot.as<uint64_t>(), c->stream);
        }
        {
            StageScope s(c, "decode_write");
            launch_decode_write_multi(t16, max_sub16, fdd, wu_d, nf, wu[nf], c->stream);
        }
        launch_dec_field_totals(fdd, nf, c->db_tot.as<uint64_t>(), c->stream);
        HIPCHK(c, hipGetLastError());
    }
    {
        StageScope s(c, "reconstruct");
        for (int kind = 0; kind < kRcKinds; kind++)
            launch_reconstruct_multi(kind, B, fdd, rk_d + (size_t)kind * ns, nf, rk[kind * ns + nf], c->stream);
        launch_reconstruct_fallback_multi(B, fdd, nf, c->stream);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(sums_h, c->db_sums.p, (size_t)nf * sizeof(DecSummary), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(tot_h, c->db_tot.p, (size_t)nf * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    timing_end(c);
    *reported = true;
    for (int k = 0; k < nf; k++) {
        if (status[k] || c->db_dt[k].root_leaf) continue;
        if (sums_h[k].error || tot_h[k] < hd[k].n) status[k] = VECSZ_ECORRUPT;
    }
    for (int k = 0; k < nf; k++)
        if (status[k]) {
            const char* why = status[k] == VECSZ_ENOMEM   ? "output buffer too small"
                              : status[k] == VECSZ_EINVAL ? "r2r decompress needs params.value_range"
                                                          : "archive is corrupt";
            return fail(c, status[k], "batch field " + std::to_string(k) + " failed: " + why);
        }
    return VECSZ_OK;
}

static int decompress_batch_fused(vecsz_ctx* c, int nf, const vecsz_params* p, const uint8_t* const* d_in,
                                  const uint64_t* in_bytes, float* const* d_out, const uint64_t* out_cap, int* status,
                                  bool* fallback) {
    for (int k = 0; k < nf; k++) status[k] = VECSZ_OK;
    bool reported = false;
    const int rc = decompress_batch_fused_run(c, nf, p, d_in, in_bytes, d_out, out_cap, status, &reported, fallback);
    if (rc && !reported)
        for (int k = 0; k < nf; k++)
            if (status[k] == VECSZ_OK) status[k] = rc;
    return rc;
}

extern "C" {

int vecsz_ctx_set_workers(vecsz_ctx* c, int nworkers) {
    if (!c || nworkers < 1 || nworkers > 16) return VECSZ_EINVAL;
    c->nworkers = nworkers;
    return VECSZ_OK;
}

int vecsz_ctx_set_device_codebook(vecsz_ctx* c, int on) {
    if (!c) return VECSZ_EINVAL;
    c->device_codebook = on != 0;
    return VECSZ_OK;
}

int vecsz_ctx_set_fused_encode(vecsz_ctx* c, int on) {
    if (!c) return VECSZ_EINVAL;
    c->fused_encode = on != 0;
    return VECSZ_OK;
}

int vecsz_ctx_set_batch_graph(vecsz_ctx* c, int on) {
    if (!c) return VECSZ_EINVAL;
    c->batch_graph = on != 0;
    if (!on) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (GraphSlot& g : c->b_graph) g.reset();
    }
    return VECSZ_OK;
}

int vecsz_ctx_set_batch_fused(vecsz_ctx* c, int on) {
    if (!c || on < 0 || on > 2) return VECSZ_EINVAL;
    c->fused_batch = on != 0;
    c->fused_batch_always = on == 2;
    return VECSZ_OK;
}

int vecsz_ctx_set_launch_shape(vecsz_ctx* c, int dq_grid, int enc_grid) {
    if (!c) return VECSZ_EINVAL;
    if (dq_grid < 1 || dq_grid > 65536 || enc_grid < 1 || enc_grid > 65536)
        return fail(c, VECSZ_EINVAL, "launch shape out of range");
    c->dq_grid = dq_grid;
    c->enc_grid = enc_grid;
    return VECSZ_OK;
}

int vecsz_autotune(vecsz_ctx* c, const vecsz_params* p, const float* d_in, uint8_t* d_out, uint64_t out_cap,
                   vecsz_tune* t) {
    if (!c) return VECSZ_EINVAL;
    if (!p || !d_in || !d_out || !t) return fail(c, VECSZ_EINVAL, "NULL argument");
    int rc = check_params(c, p, true);
    if (rc) return rc;
    memset(t, 0, sizeof(*t));
    const bool was_timing = c->timing;
    c->timing = true;
    uint64_t bytes = 0, nout = 0;
    // best (smallest) time of two runs of the named stages
    auto stage_ms = [&](const vecsz_params* q, std::initializer_list<const char*> names, double* ms) -> int {
        *ms = 1e30;
        for (int rep = 0; rep < 2; rep++) {
            const int r = compress_device(c, q, d_in, d_out, out_cap, &bytes, &nout);
            if (r) return r;
            double sum = 0;
            for (size_t i = 0; i < c->tnames.size(); i++)
                for (const char* nm : names)
                    if (c->tnames[i] == nm) sum += c->tms[i];
            *ms = std::min(*ms, sum);
        }
        return VECSZ_OK;
    };
    do {
        if ((rc = compress_device(c, p, d_in, d_out, out_cap, &bytes, &nout))) break;  // warm-up
        static const int kDqGrid[] = {256, 512, 768, 1024, 1536, 2048, 4096};
        static const int kEncGrid[] = {256, 512, 768, 1024, 2048, 4096, 8192};
        double best = 1e30, ms = 0;
        int pick = c->dq_grid;
        for (int g : kDqGrid) {
            c->dq_grid = g;
            if ((rc = stage_ms(p, {"dualquant"}, &ms))) break;
            if (ms < best) best = ms, pick = g;
        }
        if (rc) break;
        c->dq_grid = pick;
        t->dq_grid = pick;
        t->dq_ms = best;
        best = 1e30;
        pick = c->enc_grid;
        for (int g : kEncGrid) {
            c->enc_grid = g;
            // the one-pass encode, or chunk bits and packing
            if ((rc = stage_ms(p, {"encode", "chunk_bits", "pack"}, &ms))) break;
            if (ms < best) best = ms, pick = g;
        }
        if (rc) break;
        c->enc_grid = pick;
        t->enc_grid = pick;
        t->enc_ms = best;
        // block sizes: whole-compress wall time (best of two) and archive size
        vecsz_params q = *p;
        static const int kBlocks[4] = {8, 16, 32, 64};
        double bt = 1e30;
        for (int i = 0; i < 4 && !rc; i++) {
            q.block = kBlocks[i];
            double m = 1e30;
            for (int rep = 0; rep < 2 && !rc; rep++) {
                const auto t0 = std::chrono::steady_clock::now();
                rc = compress_device(c, &q, d_in, d_out, out_cap, &bytes, &nout);
                m = std::min(m, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
            }
            t->ms_by_block[i] = m;
            t->bytes_by_block[i] = bytes;
            if (m < bt) bt = m, t->block = kBlocks[i];
        }
    } while (false);
    c->timing = was_timing;
    return rc;
}

int vecsz_compress_batch(vecsz_ctx* c, int nfields, const vecsz_params* p, const float* const* d_in,
                         uint8_t* const* d_out, const uint64_t* out_cap, uint64_t* out_bytes, uint64_t* n_outlier,
                         int* status) {
    if (!c || nfields < 0 || (nfields && (!p || !d_in || !d_out || !out_cap || !out_bytes || !n_outlier || !status)))
        return fail(c, VECSZ_EINVAL, "NULL argument");
    std::vector<uint64_t> weight(nfields);
    for (int k = 0; k < nfields; k++) {
        uint64_t n = 0;  // a field with invalid extents fails on its own in compress_device
        weight[k] = field_elems(p[k].ndim, p[k].dims, &n) ? n : 0;
    }
    if (c->fused_batch && batch_fusable(c, nfields, p, d_in, d_out))
        return compress_batch_fused(c, nfields, p, d_in, d_out, out_cap, out_bytes, n_outlier, status);
    const uint64_t nmax = nfields ? *std::max_element(weight.begin(), weight.end()) : 0;
    auto prep = [&](vecsz_ctx* w) {
        (void)hipSetDevice(w->device);
        const uint64_t nch = (nmax + kChunk - 1) / kChunk;
        return w->codes.ensure(nmax * 2 + 64) && w->sparse.ensure(nmax * 4 + 64) && w->chunk_bits.ensure(nch * 4) &&
               w->chunk_zeros.ensure(nch * 4) && w->bit_off.ensure(nch * 8) && w->out_off.ensure(nch * 8);
    };
    return run_batch(c, nfields, weight.data(), status, prep, [&](vecsz_ctx* w, int k) {
        if (!d_in[k] || !d_out[k]) return fail(w, VECSZ_EINVAL, "NULL field");
        return compress_device(w, &p[k], d_in[k], d_out[k], out_cap[k], &out_bytes[k], &n_outlier[k]);
    });
}

int vecsz_decompress_batch(vecsz_ctx* c, int nfields, const vecsz_params* p, const uint8_t* const* d_in,
                           const uint64_t* in_bytes, float* const* d_out, const uint64_t* out_cap, int* status) {
    if (!c || nfields < 0 || (nfields && (!p || !d_in || !in_bytes || !d_out || !out_cap || !status)))
        return fail(c, VECSZ_EINVAL, "NULL argument");
    // Fused when the batch qualifies and its fields are small on average
    // (per-field launches and host round trips then dominate).  Batches of
    // large fields go to the worker contexts, whose independent pipelines
    // overlap one field's latency-bound decode with another's streaming
    // reconstruction and host work, where the fused stages run one after
    // another (measured faster on the ~8 GiB mix of ~85 MB fields).
    uint64_t cap_sum = 0;
    for (int k = 0; k < nfields; k++) cap_sum += std::min<uint64_t>(out_cap[k], kMaxElems);
    const bool small = nfields > 0 && cap_sum / (uint64_t)nfields <= kDBatchFuseElems;
    if (c->fused_batch && (small || c->fused_batch_always) && dbatch_fusable(c, nfields, p, d_in, d_out)) {
        bool fallback = false;  // an archive with a pad section: the per-field path below
        const int rc = decompress_batch_fused(c, nfields, p, d_in, in_bytes, d_out, out_cap, status, &fallback);
        if (!fallback) return rc;
    }
    std::vector<uint64_t> weight(nfields);
    for (int k = 0; k < nfields; k++) weight[k] = out_cap[k];
    const uint64_t nmax = nfields ? *std::max_element(weight.begin(), weight.end()) : 0;
    uint64_t bmax = 0;
    for (int k = 0; k < nfields; k++) bmax = std::max(bmax, in_bytes[k]);
    auto prep = [&](vecsz_ctx* w) {
        (void)hipSetDevice(w->device);
        const uint64_t nsub = bmax * 8 / kSubBits + 2;
        return w->codes.ensure(nmax * 2 + 64) && w->sparse.ensure(nmax * 4 + 64) && w->sub_b.ensure(nsub * 8) &&
               w->sub_e.ensure(nsub * 8) && w->sub_cnt.ensure(nsub * 4) && w->sub_zeros.ensure(nsub * 4) &&
               w->sub_lj.ensure(nsub * 8) && w->sub_ls.ensure(nsub * 16) && w->sub_ln.ensure(16) &&
               w->sym_off.ensure(nsub * 8) && w->zero_off.ensure(nsub * 8);
    };
    return run_batch(c, nfields, weight.data(), status, prep, [&](vecsz_ctx* w, int k) {
        if (!d_in[k] || !d_out[k]) return fail(w, VECSZ_EINVAL, "NULL field");
        return decompress_device(w, &p[k], d_in[k], in_bytes[k], d_out[k], out_cap[k]);
    });
}

}  // extern "C"

// ---------------------------------------------------------------- server

namespace {

// Device bytes a request holds while it runs: its input and output buffers
// plus the scratch its slot's context needs for it.
uint64_t request_bytes(int kind, const vecsz_params* p, uint64_t in_bytes, uint64_t out_cap) {
    // a buffer larger than any device counts as 2^48 bytes (the sums below cannot wrap)
    in_bytes = std::min<uint64_t>(in_bytes, 1ull << 48);
    out_cap = std::min<uint64_t>(out_cap, 1ull << 48);
    if (kind == 0) {
        uint64_t n = 0;
        if (!field_elems(p->ndim, p->dims, &n)) return ~0ull;  // refused at submit
        const uint64_t nch = (n + kChunk - 1) / kChunk;
        return 4 * n + out_cap + 6 * n + 24 * nch + (uint64_t)768 * p->dict_size * 4 + (1 << 20);
    }
    const uint64_t nsub = in_bytes * 8 / kSubBits + 2;
    return in_bytes + 4 * out_cap + 6 * out_cap + 50 * nsub + (1 << 20);
}

struct Request {
    int kind = 0;  // 0 compress, 1 decompress
    vecsz_params p{};
    const void* in = nullptr;
    void* out = nullptr;
    uint64_t in_bytes = 0, out_cap = 0, need = 0;
    int status = 0;
    uint64_t out_bytes = 0, n_outlier = 0;
    double t_submit = 0, t_start = 0, t_end = 0;
    bool done = false;
};

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct vecsz_server {
    int device = 0;
    uint64_t budget = 0, in_use = 0, next_ticket = 1;
    bool stop = false;
    std::vector<vecsz_ctx*> slots;
    std::vector<std::thread> threads;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Request> reqs;

    // Small requests at the head of the queue that can share one fused batch
    // launch: all compress or all decompress, the same block size and
    // dictionary (at most 4096), at most kFoldElems elements each (a
    // decompress request counts its output capacity); compress requests
    // without padding.  A decompress archive with a pad section sends its
    // folded batch to single calls.
    static constexpr uint64_t kFoldElems = 4ull << 20;   // 16 MiB of f32 per request
    static constexpr uint64_t kFoldTotal = 32ull << 20;  // elements of one folded batch
    static constexpr int kFoldMax = 64;
    // (submit has checked the extents: the product does not wrap)
    static uint64_t elems(const vecsz_params& p) {
        uint64_t n = 0;
        return field_elems(p.ndim, p.dims, &n) ? n : 0;
    }
    static uint64_t fold_elems(const Request& r) { return r.kind == 0 ? elems(r.p) : r.out_cap; }
    static bool foldable(const Request& r) {
        if (r.p.dict_size > kMaxLdsBook || fold_elems(r) > kFoldElems) return false;
        return r.kind == 1 || r.p.pad_mode == VECSZ_PAD_NONE;
    }

    void slot_loop(vecsz_ctx* c) {
        (void)hipSetDevice(c->device);
        for (;;) {
            uint64_t t;
            Request r;
            std::vector<uint64_t> fold;  // further requests folded into r's launch
            {
                std::unique_lock<std::mutex> lk(m);
                // FIFO admission: the oldest request starts once a slot is
                // free and its bytes fit in what the budget has left
                cv_work.wait(lk, [&] { return stop || (!queue.empty() && reqs[queue.front()].need <= budget - in_use); });
                if (stop && queue.empty()) return;
                if (queue.empty() || reqs[queue.front()].need > budget - in_use) continue;
                t = queue.front();
                queue.pop_front();
                Request& q = reqs[t];
                in_use += q.need;
                q.t_start = now_ms();
                r = q;
                // continuous batching: requests of the same kind waiting right
                // behind it join its launch while they fit the budget
                if (foldable(r)) {
                    uint64_t total = fold_elems(r);
                    while (!queue.empty() && (int)fold.size() + 1 < kFoldMax) {
                        Request& n = reqs[queue.front()];
                        if (!foldable(n) || n.kind != r.kind || n.p.block != r.p.block ||
                            n.p.dict_size != r.p.dict_size || total + fold_elems(n) > kFoldTotal ||
                            n.need > budget - in_use)
                            break;
                        total += fold_elems(n);
                        in_use += n.need;
                        n.t_start = q.t_start;
                        fold.push_back(queue.front());
                        queue.pop_front();
                    }
                }
            }
            if (!fold.empty()) {
                run_folded(c, t, r, fold);
                continue;
            }
            int rc;
            if (r.kind == 0) {
                rc = compress_device(c, &r.p, static_cast<const float*>(r.in), static_cast<uint8_t*>(r.out), r.out_cap,
                                     &r.out_bytes, &r.n_outlier);
            } else {
                rc = decompress_device(c, &r.p, static_cast<const uint8_t*>(r.in), r.in_bytes, static_cast<float*>(r.out),
                                       r.out_cap);
                if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = VECSZ_EHIP;
            }
            {
                std::lock_guard<std::mutex> lk(m);
                Request& q = reqs[t];
                q.status = rc;
                q.out_bytes = r.out_bytes;
                q.n_outlier = r.n_outlier;
                q.t_end = now_ms();
                q.done = true;
                in_use -= q.need;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }

    // Runs request t (copied in r) and the folded requests as one fused
    // batch; each keeps its own status and sizes.
    void run_folded(vecsz_ctx* c, uint64_t t, const Request& r, const std::vector<uint64_t>& fold) {
        const int nf = 1 + (int)fold.size();
        std::vector<vecsz_params> ps(nf);
        std::vector<const float*> ins(nf);
        std::vector<uint8_t*> outs(nf);
        std::vector<uint64_t> caps(nf), sizes(nf, 0), nouts(nf, 0), in_bytes(nf, 0);
        std::vector<int> st(nf, 0);
        {
            std::lock_guard<std::mutex> lk(m);
            for (int k = 0; k < nf; k++) {
                const Request& q = k == 0 ? r : reqs[fold[k - 1]];
                ps[k] = q.p;
                ins[k] = static_cast<const float*>(q.in);
                outs[k] = static_cast<uint8_t*>(q.out);
                caps[k] = q.out_cap;
                in_bytes[k] = q.in_bytes;
            }
        }
        if (r.kind == 1) {
            run_folded_decompress(c, nf, ps, ins, in_bytes, outs, caps, st);
        } else if (batch_fusable(c, nf, ps.data(), ins.data(), outs.data())) {
            // every request gets a status: the fused call gives a failure of
            // the whole launch to each field not already failed on its own
            const int rc = compress_batch_fused(c, nf, ps.data(), ins.data(), outs.data(), caps.data(), sizes.data(),
                                                nouts.data(), st.data());
            if (rc && std::all_of(st.begin(), st.end(), [](int v) { return v == VECSZ_OK; }))
                std::fill(st.begin(), st.end(), rc);
        } else {  // a request with invalid parameters: each alone, for its own status
            for (int k = 0; k < nf; k++)
                st[k] = compress_device(c, &ps[k], ins[k], outs[k], caps[k], &sizes[k], &nouts[k]);
        }
        finish_folded(t, fold, st, sizes, nouts);
    }

    void run_folded_decompress(vecsz_ctx* c, int nf, const std::vector<vecsz_params>& ps,
                               const std::vector<const float*>& ins_f, const std::vector<uint64_t>& in_bytes,
                               const std::vector<uint8_t*>& outs_u8, const std::vector<uint64_t>& caps,
                               std::vector<int>& st) {
        std::vector<const uint8_t*> ins(nf);
        std::vector<float*> outs(nf);
        for (int k = 0; k < nf; k++) {
            ins[k] = reinterpret_cast<const uint8_t*>(ins_f[k]);
            outs[k] = reinterpret_cast<float*>(outs_u8[k]);
        }
        bool fallback = true;
        if (dbatch_fusable(c, nf, ps.data(), ins.data(), outs.data())) {
            fallback = false;
            const int rc = decompress_batch_fused(c, nf, ps.data(), ins.data(), in_bytes.data(), outs.data(),
                                                  caps.data(), st.data(), &fallback);
            if (!fallback && rc && std::all_of(st.begin(), st.end(), [](int v) { return v == VECSZ_OK; }))
                std::fill(st.begin(), st.end(), rc);
        }
        if (fallback)  // invalid parameters or a pad section: each alone, for its own status
            for (int k = 0; k < nf; k++) {
                st[k] = decompress_device(c, &ps[k], ins[k], in_bytes[k], outs[k], caps[k]);
                if (!st[k] && hipStreamSynchronize(c->stream) != hipSuccess) st[k] = VECSZ_EHIP;
            }
    }

    void finish_folded(uint64_t t, const std::vector<uint64_t>& fold, const std::vector<int>& st,
                       const std::vector<uint64_t>& sizes, const std::vector<uint64_t>& nouts) {
        const int nf = 1 + (int)fold.size();
        {
            std::lock_guard<std::mutex> lk(m);
            const double te = now_ms();
            for (int k = 0; k < nf; k++) {
                Request& q = reqs[k == 0 ? t : fold[k - 1]];
                q.status = st[k];
                q.out_bytes = sizes[k];
                q.n_outlier = nouts[k];
                q.t_end = te;
                q.done = true;
                in_use -= q.need;
            }
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
};

extern "C" {

int vecsz_server_create(vecsz_server** out, int device, int nslots, uint64_t hbm_budget) {
    if (!out || nslots < 1 || nslots > 16 || hbm_budget == 0) return VECSZ_EINVAL;
    *out = nullptr;
    vecsz_server* s = new vecsz_server();
    s->device = device;
    s->budget = hbm_budget;
    for (int i = 0; i < nslots; i++) {
        vecsz_ctx* c = nullptr;
        const int rc = vecsz_ctx_create(&c, device, nullptr);
        if (rc) {
            for (vecsz_ctx* x : s->slots) vecsz_ctx_destroy(x);
            delete s;
            return rc;
        }
        s->slots.push_back(c);
    }
    for (vecsz_ctx* c : s->slots) s->threads.emplace_back([s, c] { s->slot_loop(c); });
    *out = s;
    return VECSZ_OK;
}

int vecsz_server_destroy(vecsz_server* s) {
    if (!s) return VECSZ_OK;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->threads) t.join();
    for (vecsz_ctx* c : s->slots) vecsz_ctx_destroy(c);
    delete s;
    return VECSZ_OK;
}

static int server_submit(vecsz_server* s, int kind, const vecsz_params* p, const void* in, uint64_t in_bytes, void* out,
                         uint64_t out_cap, uint64_t* ticket) {
    if (!s || !p || !in || !out || !ticket) return VECSZ_EINVAL;
    Request r;
    r.kind = kind;
    r.p = *p;
    r.in = in;
    r.out = out;
    r.in_bytes = in_bytes;
    r.out_cap = out_cap;
    uint64_t n = 0;
    if (kind == 0 && !field_elems(p->ndim, p->dims, &n)) return VECSZ_EINVAL;
    if (kind == 1 && (out_cap > kMaxElems || in_bytes > 16 * kMaxElems)) return VECSZ_EINVAL;
    r.need = request_bytes(kind, p, in_bytes, out_cap);
    if (r.need > s->budget) return VECSZ_ENOMEM;  // could never be admitted
    r.t_submit = now_ms();
    {
        std::lock_guard<std::mutex> lk(s->m);
        *ticket = s->next_ticket++;
        s->reqs[*ticket] = r;
        s->queue.push_back(*ticket);
    }
    s->cv_work.notify_one();
    return VECSZ_OK;
}

int vecsz_server_submit_compress(vecsz_server* s, const vecsz_params* p, const float* d_in, uint8_t* d_out,
                                 uint64_t out_cap, uint64_t* ticket) {
    return server_submit(s, 0, p, d_in, 0, d_out, out_cap, ticket);
}

int vecsz_server_submit_decompress(vecsz_server* s, const vecsz_params* p, const uint8_t* d_in, uint64_t in_bytes,
                                   float* d_out, uint64_t out_cap, uint64_t* ticket) {
    return server_submit(s, 1, p, d_in, in_bytes, d_out, out_cap, ticket);
}

int vecsz_server_wait(vecsz_server* s, uint64_t ticket, vecsz_server_result* res) {
    if (!s || !res) return VECSZ_EINVAL;
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return VECSZ_EINVAL;
    s->cv_done.wait(lk, [&] { return s->reqs[ticket].done; });
    const Request& r = s->reqs[ticket];
    res->status = r.status;
    res->out_bytes = r.out_bytes;
    res->n_outlier = r.n_outlier;
    res->queue_ms = r.t_start - r.t_submit;
    res->run_ms = r.t_end - r.t_start;
    res->latency_ms = r.t_end - r.t_submit;
    s->reqs.erase(ticket);
    return VECSZ_OK;
}

int vecsz_ctx_last_sidecar(vecsz_ctx* c, vecsz_sidecar* out) {
    if (!c || !out) return VECSZ_EINVAL;
    if (!c->have_sidecar) return fail(c, VECSZ_EINVAL, "no compress has completed on this context");
    *out = c->sidecar;
    return VECSZ_OK;
}

int vecsz_params_from_sidecar(const vecsz_sidecar* sc, vecsz_params* p) {
    if (!sc || !p || sc->magic != VECSZ_SIDECAR_MAGIC || sc->version != 1) return VECSZ_EINVAL;
    vecsz_params_default(p);
    p->ndim = sc->ndim;
    p->mode = sc->mode;
    p->eb = sc->eb;
    p->value_range = sc->value_range;
    p->block = sc->block;
    p->dict_size = sc->dict_size;
    return VECSZ_OK;
}

int vecsz_verify(vecsz_ctx* c, const float* d_orig, const float* d_recon, uint64_t n, vecsz_stat* out) {
    if (!c || !d_orig || !d_recon || !out || n == 0) return fail(c, VECSZ_EINVAL, "bad verify arguments");
    HIPCHK(c, hipSetDevice(c->device));
    const int grid = verify_grid(n);
    const size_t bytes = (size_t)grid * std::max(sizeof(VerifyPart1), sizeof(VerifyPart2));
    if (!c->vparts.ensure(bytes) || !c->vpin.ensure(bytes)) return fail(c, VECSZ_ENOMEM, "verify scratch");
    launch_verify_pass1(d_orig, d_recon, n, c->vparts.as<VerifyPart1>(), grid, c->stream);
    HIPCHK(c, hipMemcpyAsync(c->vpin.p, c->vparts.p, grid * sizeof(VerifyPart1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const VerifyPart1* p1 = c->vpin.as<VerifyPart1>();
    double so = 0, sx = 0, sq = 0;
    float mn = p1[0].min_o, mx = p1[0].max_o, err = p1[0].max_err, rel = 0;
    uint64_t idx = p1[0].err_idx;
    for (int g = 0; g < grid; g++) {
        so += p1[g].sum_o;
        sx += p1[g].sum_x;
        sq += p1[g].sum_sq;
        mn = std::min(mn, p1[g].min_o);
        mx = std::max(mx, p1[g].max_o);
        rel = std::max(rel, p1[g].max_rel);
        if (p1[g].max_err > err || (p1[g].max_err == err && p1[g].err_idx < idx)) {
            err = p1[g].max_err;
            idx = p1[g].err_idx;
        }
    }
    const double mean_o = so / (double)n, mean_x = sx / (double)n;
    launch_verify_pass2(d_orig, d_recon, n, mean_o, mean_x, c->vparts.as<VerifyPart2>(), grid, c->stream);
    HIPCHK(c, hipMemcpyAsync(c->vpin.p, c->vparts.p, grid * sizeof(VerifyPart2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const VerifyPart2* p2 = c->vpin.as<VerifyPart2>();
    double corr = 0, vo = 0, vx = 0;
    for (int g = 0; g < grid; g++) {
        corr += p2[g].corr;
        vo += p2[g].var_o;
        vx += p2[g].var_x;
    }
    // the closing formulas of VerifyData (src/utils/verify.hh:60-73)
    const double std_o = std::sqrt(vo / (double)n), std_x = std::sqrt(vx / (double)n);
    out->len = n;
    out->coeff = corr / (double)n / std_o / std_x;
    out->maximum = mx;
    out->minimum = mn;
    out->range = (double)mx - (double)mn;
    out->max_abserr_index = idx;
    out->max_abserr = err;
    out->max_abserr_vs_range = err / out->range;
    out->max_pwr_rel_abserr = rel;
    out->MSE = sq / (double)n;
    out->NRMSE = std::sqrt(out->MSE) / out->range;
    out->PSNR = 20 * std::log10(out->range) - 10 * std::log10(out->MSE);
    return VECSZ_OK;
}

int vecsz_ctx_enable_timing(vecsz_ctx* c, int on) {
    if (!c) return VECSZ_EINVAL;
    c->timing = on != 0;
    return VECSZ_OK;
}

int vecsz_ctx_timings(vecsz_ctx* c, const char** names, double* ms, int cap) {
    if (!c) return 0;
    int n = (int)std::min<size_t>(c->tnames.size(), (size_t)cap);
    for (int i = 0; i < n; i++) {
        if (names) names[i] = c->tnames[i].c_str();
        if (ms) ms[i] = c->tms[i];
    }
    return n;
}

}  // extern "C"
