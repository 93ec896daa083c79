// Launch wrappers of the gfx950 kernels (host-callable).
#pragma once

#include "common.hpp"

namespace vsz {

constexpr int kMaxLdsHistBins = 8192;   // LDS-private histogram up to this dict size
constexpr int kMaxLdsLenTab = 16384;    // LDS code-length table up to this dict size
constexpr int kMaxLdsBook = 4096;       // LDS codebook (k_pack) up to this dict size

int dq_lds_bytes(int dict);
// Zeroes n 32-bit words with a kernel.
void launch_zero_u32(uint32_t* p, uint64_t n, hipStream_t st);
// out: 2 + 2 * kRangeParts words (min/max, then one pair per workgroup).
constexpr uint32_t kRangeParts = 1024;
uint32_t range_parts(uint64_t n);
void launch_range(const float* x, uint64_t n, uint32_t* out, hipStream_t st);
void launch_ebfam(uint32_t* mm, uint64_t n, const float* x, double eb, EbDev* out, hipStream_t st);
void launch_dualquant(int ndim, int B, const float* in, uint16_t* codes, float* sparse, unsigned long long* hist,
                      uint32_t* partials, const DQArgs& a, int grid, hipStream_t st);
// 4-wide tiles (256 x per wavefront) for 2D/3D fields at least 128 wide;
// aligned: DIM0 % 4 == 0 and a 16-byte aligned field (vector loads/stores).
// A single field takes the element-load form only when it has enough
// 4-wide tiles to fill the chip; a batch (tiles of many fields) always does.
bool dualquant_use_x4(int ndim, int B, const float* in, uint64_t d0, uint64_t d1, uint64_t d2, bool batched);
bool dualquant_x4_aligned(const float* in, uint64_t d0);
// grid_cap: most workgroups of a launch (a tuned launch shape; the output
// does not depend on it).
void launch_chunk_bits(const uint16_t* codes, uint64_t n, const uint8_t* len_tab, int dict, uint32_t* cb, uint32_t* cz,
                       uint64_t nchunks, int grid_cap, hipStream_t st);
// index (nullable): decode index whose header words 3 and 4 get the totals.
void launch_enc_summary(const uint64_t* totals, EncSummary* sum, uint8_t* archive, uint64_t meta_bytes,
                        uint64_t pad_bytes, uint64_t tree_bytes, uint64_t out_cap, uint32_t* err, uint64_t* index,
                        hipStream_t st);
void launch_pack(const uint16_t* codes, uint64_t n, const uint32_t* cw_hi, const uint32_t* cw_lo, const uint8_t* len_tab,
                 int dict, const uint64_t* bit_off, const uint64_t* out_off, uint64_t nchunks, const EncSummary* sum,
                 const float* sparse, uint8_t* archive, const uint32_t* err, int max_len, int grid_cap,
                 uint64_t* index, hipStream_t st);

// Huffman codebook built on the device (kernels_codebook.hip), one
// workgroup, dictionaries up to kMaxLdsBook: book = cw_hi[dict], cw_lo[dict],
// u8 len[dict]; huff = the archive's Huffman section ([u32be node count]
// [u32be dict][tree]), huff_cap bytes available there.
struct CodebookInfo {
    uint32_t node_count, tree_bytes, max_len;
    int32_t status;  // VECSZ_OK, VECSZ_ECODELEN, VECSZ_EINVAL (empty histogram), VECSZ_ENOMEM (no room)
};
size_t codebook_lds_bytes(int dict);
void launch_codebook(const unsigned long long* hist, int dict, uint32_t* book, uint8_t* huff, uint64_t huff_cap,
                     CodebookInfo* info, hipStream_t st);
// k_enc_summary with the tree size from a device-built codebook (a failed
// build sets err to 2 and writes nothing).
void launch_enc_summary_dev(const uint64_t* totals, EncSummary* sum, uint8_t* archive, uint64_t meta_bytes,
                            uint64_t pad_bytes, const CodebookInfo* info, uint64_t out_cap, uint32_t* err,
                            uint64_t* index, hipStream_t st);

// Fused single-pass encode (bits, prefix by decoupled look-back, packing) for
// dictionaries whose codebook fits in LDS (encode_fused_ok).  status: 2 *
// encode_parts(nchunks) + 1 words, zeroed before the launch; outl: 4 bytes per
// outlier (the histogram's bin 0); the header must be in place.  Then
// launch_enc_finish writes the sizes and the outlier section.
constexpr int kEncWaveChunks = 8;                   // chunks per wavefront
constexpr int kEncPartChunks = 4 * kEncWaveChunks;  // chunks per partition (one workgroup)
uint64_t encode_parts(uint64_t nchunks);
bool encode_fused_ok(int dict);
void launch_encode(const uint16_t* codes, uint64_t n, const uint32_t* cw_hi, const uint32_t* cw_lo,
                   const uint8_t* len_tab, int dict, int max_len, const float* sparse, uint8_t* archive,
                   uint64_t stream_bit_off, uint64_t out_cap, bool no_stream, uint32_t* outl, uint64_t* status,
                   uint32_t* err, uint64_t* index, int grid_cap, hipStream_t st);
void launch_enc_finish(const uint64_t* status, uint64_t n, const uint32_t* outl, uint64_t n_outlier_max,
                       EncSummary* sum, uint8_t* archive, uint64_t meta_bytes, uint64_t pad_bytes, uint64_t tree_bytes,
                       uint64_t out_cap, uint32_t* err, uint64_t* index, hipStream_t st);

// Batched compress (kernels_compress.hip): every field of a batch per launch.
// start arrays hold nf + 1 unit offsets (see BatchField).
uint64_t range_units(uint64_t n);
void launch_range_multi(const BatchField* fd, int nf, const uint64_t* start, uint64_t nunits, hipStream_t st);
// Dual quantization of the fields whose tile form is `kind` (DqKind); start
// gives units only to those fields.
void launch_dualquant_multi(int B, int kind, const BatchField* fd, int nf, const uint64_t* start, uint64_t nunits,
                            int grid_cap, hipStream_t st);
// Histograms of the codes: workgroup g of hist_multi_grid(U) takes units
// [U g / G, U (g + 1) / G) and writes one partial row per field it visits,
// from row0[g] on; groups: {field, first row, rows <= 32} triples.
uint64_t hist_units(uint64_t n);
unsigned hist_multi_grid(uint64_t nunits);
void launch_hist_multi(const BatchField* fd, int nf, const uint64_t* start, uint64_t nunits, int dict,
                       uint32_t* partials, const uint32_t* row0, const uint32_t* groups, uint32_t ngroups,
                       unsigned long long* hist, hipStream_t st);
void launch_place_heads(const BatchField* fd, int nf, const uint32_t* errs, hipStream_t st);
void launch_chunk_bits_multi(const BatchField* fd, int nf, const uint64_t* start, uint64_t nunits, int dict,
                             uint32_t* cb, uint32_t* cz, int grid_cap, hipStream_t st);
void launch_enc_summary_multi(const BatchField* fd, int nf, const uint64_t* bit_off, const uint64_t* out_off,
                              const uint64_t* totals, uint64_t ctot, EncSummary* sums, uint32_t* errs, hipStream_t st);
void launch_pack_multi(const BatchField* fd, int nf, const uint64_t* start, uint64_t nunits, int dict, int max_len,
                       const uint64_t* bit_off, const uint64_t* out_off, const EncSummary* sums, const uint32_t* errs,
                       int grid_cap, hipStream_t st);

// Decoder (kernels_decompress.hip)
struct DecodeArgs {
    const uint8_t* stream;   // first byte of the bitstream (any alignment)
    uint64_t stream_bits;    // 8 * stream bytes
    uint64_t n;              // symbols to decode
    const uint32_t* lut;     // first-level table, 1 << kLutBits entries
    const uint32_t* sub;     // deeper tables
    const uint16_t* lut16;   // compact tables (nullptr: use lut / sub)
    const uint16_t* sub16;
    uint32_t n_sub16;        // entries of sub16 the kernels copy to LDS; 0: read in global memory
    uint64_t nsub;           // subsequences of kSubBits bits
    const uint8_t* outliers; // outlier section of the archive (any alignment)
    uint64_t n_outlier;
};

struct SubState {
    uint64_t* b;      // true start bit of each subsequence (first codeword start >= j * kSubBits)
    uint64_t* e;      // its end: first codeword start >= (j + 1) * kSubBits
    uint32_t* cnt;    // codewords starting in [b, e)
    uint32_t* zer;    // of which decode to symbol 0
    // Re-decode lists, used in alternation by successive passes: subsequence
    // list_j[.][k] to re-decode from bit list_start[.][k] (nsub entries each),
    // and entry counts in a ring of three (pass p reads list_n[p % 3]).
    uint32_t* list_j[2];
    uint64_t* list_start[2];
    uint32_t* list_n;
    // List entries hold jbase + j: 0 for one field; in a batch the arrays
    // above are views of one field's part of batch-wide arrays, and jbase is
    // the batch-wide number of its subsequence 0 (the lists are batch-wide).
    uint32_t jbase = 0;
};

void launch_decode_sync(const DecodeArgs& a, SubState s, DecSummary* sum, hipStream_t st);
// Re-decode pass `pass` over the list the previous pass (or the sync pass) built.
void launch_decode_pass(const DecodeArgs& a, SubState s, uint32_t pass, DecSummary* sum, hipStream_t st);
// guard (nullable): every kernel of the scan does nothing while *guard is nonzero.
void launch_scan_u32x2(const uint32_t* a, const uint32_t* b, uint64_t n, uint64_t* oa, uint64_t* ob, uint64_t* block_tmp,
                       uint64_t* totals, hipStream_t st, const uint32_t* guard = nullptr);
void launch_decode_write(const DecodeArgs& a, SubState s, const uint64_t* sym_off, const uint64_t* zero_off,
                         uint16_t* codes, float* sparse, DecSummary* sum, hipStream_t st,
                         const uint32_t* guard = nullptr);
void launch_fill_u16(uint16_t* p, uint64_t n, uint16_t v, hipStream_t st);
// Decode with a decode index (index = its kIdxHead header words first): codes
// and outliers as k_decode_write leaves them.
// 1D reconstruction fused into the indexed decode: output (16-byte
// aligned), block size, radius, 2 eb, and the halo (3 floats per block, x
// face last) or null for +0.0.
struct Recon1D {
    float* out;
    int B, radius;
    double twoeb;
    const float* halo;
    uint64_t halo_stride;
};
// total_bits: the stream's bit count from the index header.  r1d null: codes
// and outliers for the reconstruction kernels; else the 1D field itself.
void launch_decode_indexed(const DecodeArgs& a, const uint64_t* index, uint64_t total_bits, uint16_t* codes,
                           float* sparse, const Recon1D* r1d, DecSummary* sum, hipStream_t st);
void launch_scatter_all_outliers(const uint8_t* outliers, uint64_t n, float* sparse, hipStream_t st);

struct ReconArgs {
    uint64_t d0, d1, d2, n;
    int radius;
    double twoeb;
    uint64_t ntiles_x, ntiles_y, ntiles_z;
    const float* halo;  // as DQArgs::halo; nullptr = the +0.0 halo
    uint64_t halo_stride, hb0, hb1;
    const uint32_t* guard = nullptr;  // nonzero *guard: the launch does nothing
};
void launch_reconstruct(int ndim, int B, const uint16_t* codes, const float* sparse, float* out, const ReconArgs& a,
                        uint32_t* flagged, uint32_t* nflagged, int grid, hipStream_t st);
void launch_reconstruct_fallback(int ndim, int B, const uint16_t* codes, const float* sparse, float* out,
                                 const ReconArgs& a, const uint32_t* flagged, const uint32_t* nflagged, hipStream_t st);

// Batched decompress: one field as the batch kernels see it.  The
// subsequence state, the code-count prefixes and the re-decode lists are
// batch-wide arrays (a field's subsequences are numbered from sub0); s,
// sym_off and zero_off are views of its part.
// 64 x per wavefront (kRcN), four x per lane with vector (kRcNx4) or element
// (kRcNx4u) loads and stores.
enum ReconKind { kRc1 = 0, kRc1x4, kRc2, kRc2x4, kRc3, kRc3x4, kRc1x4u, kRc2x4u, kRc3x4u, kRcKinds };
int recon_kind(int ndim, int B, uint64_t d0, const void* codes, const void* out);
int recon_kind_ndim(int kind);
struct DecField {
    DecodeArgs a;              // nsub = 0: nothing to decode on the device
    SubState s;
    const uint64_t* sym_off;
    const uint64_t* zero_off;
    uint16_t* codes;
    float* sparse;
    float* out;
    ReconArgs ra;
    uint32_t* flagged;         // blocks for the fp32 fallback
    uint32_t* nflagged;
    DecSummary* sum;
    int ndim;
    int kind;                  // ReconKind; -1: not reconstructed
};
// t16: every field has compact tables; max_sub16: the largest n_sub16.
// start: per-field unit starts (nf + 1), nunits = start[nf].
// The largest deeper-table size of a batch's fields (entries), or 0 when
// keeping tables that size in LDS would cost the batch's sync pass a
// resident workgroup per CU (the fields then read them in global memory).
uint32_t batch_sub16_in_lds(uint32_t max_sub16);
void launch_decode_sync_multi(bool t16, uint32_t max_sub16, const DecField* fd, const uint64_t* start, int nf,
                              uint64_t nunits, hipStream_t st);
// sub_start: batch-wide number of each field's subsequence 0 (nf + 1 entries);
// S: the batch-wide lists (list_j, list_start, list_n).
void launch_decode_pass_multi(bool t16, const DecField* fd, const uint64_t* sub_start, int nf, SubState S,
                              uint32_t pass, uint64_t nsub_tot, hipStream_t st);
void launch_decode_write_multi(bool t16, uint32_t max_sub16, const DecField* fd, const uint64_t* start, int nf,
                               uint64_t nunits, hipStream_t st);
// Symbols each field's subsequences hold (tot[f]).
void launch_dec_field_totals(const DecField* fd, int nf, uint64_t* tot, hipStream_t st);
// One reconstruction form over its fields' tiles (start: per-field tile starts).
void launch_reconstruct_multi(int kind, int B, const DecField* fd, const uint64_t* start, int nf, uint64_t ntiles,
                              hipStream_t st);
void launch_reconstruct_fallback_multi(int B, const DecField* fd, int nf, hipStream_t st);
// Tiles a field of form `kind` has for the reconstruction kernels.
uint64_t recon_tiles(int kind, const ReconArgs& a);
// Workgroups of the sync and write passes for a field of nsub subsequences.
uint64_t decode_sync_units(uint64_t nsub);
uint64_t decode_write_units(uint64_t nsub);
// Copies the first len bytes of each archive to dst + off (one workgroup each).
struct HeadCopy {
    const uint8_t* src;
    uint64_t len, off;
};
void launch_gather_heads(const HeadCopy* hc, int n, uint8_t* dst, hipStream_t st);

// Padding (kernels_pad.hip)
struct PadArgs {
    const float* x;
    uint64_t d0, d1, d2, n;
    int ndim, B, mode, type;
    float constant;
    uint64_t hb0, hb1, hb2;  // reference block grid
    const EbDev* ebd;        // r_d from here when set, else r_d below
    double r_d;
};
// Per-block pad values (block / edge modes): pad_out in archive order, halo 3 per block.
void launch_pad_blocks(const PadArgs& a, float* pad_out, float* halo, hipStream_t st);
// Global mode: pad_out[0] as stored, halo[0..2] as the dual quantization uses it.
// scratch: 32 bytes of device memory.
void launch_pad_global(const PadArgs& a, void* scratch, float* pad_out, float* halo, hipStream_t st);
// Decompress: halo (3 per block of nblocks) from an archive's pad section.
void launch_pad_halo(int ndim, int mode, const float* pad, uint64_t pad_len, uint64_t nblocks, float* halo,
                     hipStream_t st);

// Reconstruction quality (kernels_verify.hip)
struct VerifyPart1 {
    double sum_o, sum_x, sum_sq;
    float min_o, max_o, max_err, max_rel;
    uint64_t err_idx;
};
struct VerifyPart2 {
    double corr, var_o, var_x;
};
int verify_grid(uint64_t n);
void launch_verify_pass1(const float* o, const float* x, uint64_t n, VerifyPart1* part, int grid, hipStream_t st);
void launch_verify_pass2(const float* o, const float* x, uint64_t n, double mean_o, double mean_x, VerifyPart2* part,
                         int grid, hipStream_t st);

}  // namespace vsz
