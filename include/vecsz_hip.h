/*
 * vecsz_hip.h -- C ABI of the MI355X (gfx950) implementation of vecSZ's
 * dual-quantization + Huffman compression path.
 *
 * Every entry point replaces a piece of the reference's header-only
 * interface; the reference function each one stands in for is cited beside
 * it.  Archives are byte-identical to the reference's datapack::pack payload
 * (src/utils/datapack.hh:17-29) and decompressed floats are bit-identical to
 * the reference's Decompress output.
 *
 * Conventions
 *  - Plain pointers and sizes only; no C++ or torch types.
 *  - Device pointers (d_*) live on the context's GPU; host pointers (h_*)
 *    are ordinary host memory.
 *  - All calls are stream-ordered on the context's stream and return when
 *    their results (sizes, status) are known on the host.
 *  - Functions return a VECSZ_* status and never exit the process (the
 *    reference calls exit(1) on IO errors and a zero r2r range,
 *    src/types.cc:51-54, src/utils/io.hh:26-30).
 *  - A context is not thread-safe; use one context per host thread.
 */
#ifndef VECSZ_HIP_H
#define VECSZ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VECSZ_OK 0
#define VECSZ_EINVAL 1     /* bad argument / unsupported parameter */
#define VECSZ_ENOMEM 2     /* device or host allocation failed, or out_cap too small */
#define VECSZ_ECORRUPT 3   /* malformed or truncated archive */
#define VECSZ_ECODELEN 4   /* a Huffman code longer than 64 bits (reference's >64-bit branch) */
#define VECSZ_EHIP 5       /* HIP runtime error */
#define VECSZ_ERANGE0 6    /* r2r with zero value range (reference exits, src/types.cc:51-54) */
#define VECSZ_ENODEV 7     /* no usable GPU / library built without device code */

#define VECSZ_MODE_ABS 0
#define VECSZ_MODE_R2R 1

/* Parameters that the reference carries in ArgParse
 * (src/argument_parser/argparse.hh:45-98) and config_t (src/types.hh:126-140).
 * Decompress reads ndim, dims and eb from the archive and takes mode,
 * value_range, block and dict_size from here, because the archive does not
 * record them (the reference takes them from its command line). */
typedef struct vecsz_params {
    int ndim;              /* 1, 2 or 3 */
    uint64_t dims[3];      /* x (fastest) first; unused dims = 1 */
    int mode;              /* VECSZ_MODE_ABS or VECSZ_MODE_R2R */
    double eb;             /* user error bound, stored in the archive */
    double value_range;    /* r2r: value range to use; <= 0 on compress = compute max-min of the input */
    int block;             /* Lorenzo block size B: 8, 16, 32 or 64 (reference default 8) */
    int dict_size;         /* quant-code dictionary (reference default 4096); radius = dict_size / 2 */
    /* Compress only: padding (--padding, src/argument_parser/argparse.cc:399-415;
     * src/utils/padding.hh).  Decompress reads the mode from the archive's pad
     * section (src/vecsz_interface.hh:287-298). */
    int pad_mode;          /* VECSZ_PAD_NONE (default), _GLOBAL, _BLOCK or _EDGE */
    int pad_type;          /* VECSZ_PADV_MAX, _MIN, _AVG or _CONST (default) */
    float pad_constant;    /* value of VECSZ_PADV_CONST */
} vecsz_params;

/* Padding modes (szwf.global_padding / block_padding / edge_padding) and pad
 * value types (MAX_VALUE..CONST_VAL, src/utils/padding.hh:10-13). */
enum { VECSZ_PAD_NONE = 0, VECSZ_PAD_GLOBAL = 1, VECSZ_PAD_BLOCK = 2, VECSZ_PAD_EDGE = 3 };
enum { VECSZ_PADV_MAX = 0, VECSZ_PADV_MIN = 1, VECSZ_PADV_AVG = 2, VECSZ_PADV_CONST = 3 };

typedef struct vecsz_ctx vecsz_ctx;

/* Create a context bound to `device`; `stream` is a hipStream_t or NULL for a
 * context-owned stream (default priority). */
int vecsz_ctx_create(vecsz_ctx** ctx, int device, void* stream);
int vecsz_ctx_destroy(vecsz_ctx* ctx);
const char* vecsz_strerror(int status);
/* Human-readable detail of the last failure on this context. */
const char* vecsz_ctx_last_error(vecsz_ctx* ctx);

/* Default parameters: abs, eb 1e-4, B 8, dict 4096
 * (src/argument_parser/argparse.hh:62-77). */
void vecsz_params_default(vecsz_params* p);

/* Upper bound of the archive payload for `p` (no allocation). */
int vecsz_compress_bound(const vecsz_params* p, uint64_t* bytes);

/* Replaces vecsz::interface::Compress<float,int> (src/vecsz_interface.hh:31-229),
 * lossless pass excluded: field d_in (device) -> archive payload d_out
 * (device, out_cap bytes).  *out_bytes receives the payload size
 * (Compress's *data_size) and *n_outlier the outlier count.  The input is
 * not modified (the reference frees it and mutates it in 1D). */
int vecsz_compress(vecsz_ctx* ctx, const vecsz_params* p, const float* d_in, uint8_t* d_out,
                   uint64_t out_cap, uint64_t* out_bytes, uint64_t* n_outlier);

/* Replaces vecsz::interface::Decompress<float,int> (src/vecsz_interface.hh:232-375),
 * lossless pass excluded: archive payload d_in (device) -> field d_out
 * (device, out_cap floats). */
int vecsz_decompress(vecsz_ctx* ctx, const vecsz_params* p, const uint8_t* d_in, uint64_t in_bytes,
                     float* d_out, uint64_t out_cap);

/* Allocation-free compress: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same archive as
 * vecsz_compress, byte for byte, with every device scratch buffer inside a
 * static array of the library (about 32 MiB per GPU, held until the process
 * ends) and no host allocation during the call (the context reserves its
 * staging when it is created).  Limits: at most 2^22 elements (a 16 MiB
 * field), dict_size at most 4096, no padding; larger requests return
 * VECSZ_EINVAL.  Calls in one process run one at a time (they share their
 * GPU's array and the host codebook arrays, so overlapping calls would
 * race; the library serialises them).  The call still stops once on the
 * host for the codebook, so it cannot itself be captured in a graph. */
int vecsz_compress_noalloc(vecsz_ctx* ctx, const vecsz_params* p, const float* d_in, uint8_t* d_out,
                           uint64_t out_cap, uint64_t* out_bytes, uint64_t* n_outlier);

/* Decode index (no counterpart in the reference; its decoder walks the
 * bitstream serially, src/huffman.hh:343-372): a side product of compress,
 * kept beside the archive, never inside it.  For every 128 symbols it records
 * the bit where their first codeword starts and the rank of their first
 * outlier, so decompress can decode each 128-symbol segment independently
 * instead of synchronising on the bitstream.
 * vecsz_index_words: size of the index for a field, in u64 words.
 * vecsz_compress_indexed: vecsz_compress (the same archive, byte for byte)
 *   that also writes the index to d_index (device, index_words words).
 * vecsz_decompress_indexed: vecsz_decompress (the same output, bit for bit)
 *   taking the index; VECSZ_EINVAL when the index was not made for this
 *   archive (element count, segment count, stream bits or outlier count
 *   differ), VECSZ_ECORRUPT when a segment does not end where the next one
 *   starts. */
int vecsz_index_words(const vecsz_params* p, uint64_t* words);
int vecsz_compress_indexed(vecsz_ctx* ctx, const vecsz_params* p, const float* d_in, uint8_t* d_out,
                           uint64_t out_cap, uint64_t* out_bytes, uint64_t* n_outlier, uint64_t* d_index,
                           uint64_t index_words);
int vecsz_decompress_indexed(vecsz_ctx* ctx, const vecsz_params* p, const uint8_t* d_in, uint64_t in_bytes,
                             const uint64_t* d_index, uint64_t index_words, float* d_out, uint64_t out_cap);

/* Host-memory conveniences: stage through device buffers owned by the
 * context.  Same semantics as the device entry points. */
int vecsz_compress_host(vecsz_ctx* ctx, const vecsz_params* p, const float* h_in, uint8_t* h_out,
                        uint64_t out_cap, uint64_t* out_bytes, uint64_t* n_outlier);
int vecsz_decompress_host(vecsz_ctx* ctx, const vecsz_params* p, const uint8_t* h_in, uint64_t in_bytes,
                          float* h_out, uint64_t out_cap);

/* Parse an archive payload header on the host (datapack::unpack,
 * src/utils/datapack.hh:33-56). */
int vecsz_archive_info(const uint8_t* h_archive, uint64_t bytes, int* ndim, uint64_t dims[3], double* eb,
                       uint64_t* huff_size, uint64_t* n_outlier);

/* Batched form: many independent fields per call, each compressed exactly as
 * vecsz_compress would (identical archives).  No counterpart in the
 * reference, whose main compresses one file per run (src/vecsz.cc:37).
 * Compress of two or more fields that share block size and dict_size (at
 * most 4096) and use no padding runs fused: one launch per stage covers every
 * field (a work queue of Lorenzo-block tiles and encode chunks spanning all
 * fields, per-field histograms and codebooks, the codebooks built on
 * set_workers host threads).  Decompress of two or more archives with one
 * block size and dict_size (at most 4096), no pad section and output
 * capacities of at most 4M elements on average runs fused the same way (see
 * vecsz_ctx_set_batch_fused): one copy of every archive's header and tree to the host, the
 * decode tables built on host threads, then one launch per stage (sync pass,
 * re-decode passes over one batch-wide list, scan, write pass, one
 * reconstruction launch per tile form).  Other batches are spread, largest
 * first, over worker contexts (own HIP stream and scratch each, set_workers
 * of them), so small fields overlap on the GPU and host codebook work
 * overlaps kernels.  status[k] receives each field's VECSZ_*
 * status; the call returns the first failure (or VECSZ_OK) after all fields
 * are done.  Inputs must be complete on the context's stream. */
int vecsz_compress_batch(vecsz_ctx* ctx, int nfields, const vecsz_params* params, const float* const* d_in,
                         uint8_t* const* d_out, const uint64_t* out_cap, uint64_t* out_bytes, uint64_t* n_outlier,
                         int* status);
int vecsz_decompress_batch(vecsz_ctx* ctx, int nfields, const vecsz_params* params, const uint8_t* const* d_in,
                           const uint64_t* in_bytes, float* const* d_out, const uint64_t* out_cap, int* status);
/* Worker contexts used by the batched calls (1..16, default 8). */
int vecsz_ctx_set_workers(vecsz_ctx* ctx, int nworkers);
/* 0: batches always take the worker-context path.  1 (default): compress
 * batches take the fused path when they qualify; decompress batches when
 * they qualify and their output capacities average at most 4M elements
 * (larger fields decompress faster as independent per-field pipelines on
 * the worker contexts).  2: the fused path whenever a batch qualifies.  All
 * write identical archives and identical decompressed fields. */
int vecsz_ctx_set_batch_fused(vecsz_ctx* ctx, int on);
/* 1 (default): a fused compress batch is captured as two HIP graphs, the
 * uploads, launches and copies before its host codebook stop and those
 * after it, when the same batch layout comes twice in a row (same field
 * count, buffers, sizes and launch shapes; the field data may differ), and
 * the graphs are replayed for as long as the layout stays the same; other
 * batches run eagerly.  0: always eagerly.  No counterpart in the
 * reference.  Identical archives. */
int vecsz_ctx_set_batch_graph(vecsz_ctx* ctx, int on);

/* 1: a single-field compress encodes in one pass (bit counts, prefix by a
 * decoupled look-back over partitions, and packing in one kernel; the codes
 * are read from HBM once) when dict_size is at most 4096; 0 (default): chunk
 * bit counts, a scan of them and the packing as three kernels, which is
 * faster on an MI355X (0.221 against 0.243 ms at 512^3).  Both write
 * identical archives. */
int vecsz_ctx_set_fused_encode(vecsz_ctx* ctx, int on);

/* 1: a single-field compress with dict_size at most 4096 builds its Huffman
 * codebook on the device (one workgroup replays the reference's queue, then
 * codes and the tree bytes in parallel), so it runs to its end without a stop
 * on the host; 0 (default): the histogram goes to the host, which builds the
 * codebook (faster: the queue is a sequential chain of dependent steps that a
 * host core runs many times faster than one GPU lane).  Identical archives. */
int vecsz_ctx_set_device_codebook(vecsz_ctx* ctx, int on);

/* GPU launch-shape autotuner; replaces autotune_vector_len /
 * autotune_block_sizes (src/autotune.hh:80-127, called from
 * src/vecsz_interface.hh:60-68).  Compresses the field repeatedly (d_out is
 * scratch of out_cap bytes; it ends holding the B = 64 archive) and:
 *  - keeps on the context the workgroup caps of the dual quantization and of
 *    the encode kernels (chunk bits, pack) that ran fastest; every later
 *    compress on the context, single or batched, uses them, and the archives
 *    do not depend on them;
 *  - like the reference's search over B in {8, 16, 32, 64}, times the whole
 *    compress for each B and reports the fastest as `block`, with each B's
 *    archive size.  B is a format parameter (decompress needs it), so it is
 *    reported, not applied: the caller sets params->block and keeps it with
 *    the archive (the sidecar records it). */
typedef struct vecsz_tune {
    int block;                  /* fastest B of {8, 16, 32, 64} */
    int dq_grid, enc_grid;      /* launch shapes now set on the context */
    int pad_;
    double ms_by_block[4];      /* whole compress, B = 8, 16, 32, 64 */
    uint64_t bytes_by_block[4]; /* archive bytes for each B */
    double dq_ms, enc_ms;       /* stage times at the chosen shapes */
} vecsz_tune;
int vecsz_autotune(vecsz_ctx* ctx, const vecsz_params* params, const float* d_in, uint8_t* d_out, uint64_t out_cap,
                   vecsz_tune* out);
/* Sets the launch shapes vecsz_autotune chooses (e.g. from a stored result):
 * workgroup caps of the dual quantization and of the encode kernels, each in
 * [1, 65536] (defaults 768 and 2048). */
int vecsz_ctx_set_launch_shape(vecsz_ctx* ctx, int dq_grid, int enc_grid);

/* Served form: a request server on one GPU (no counterpart in the
 * reference).  nslots slot threads, each with its own context (stream and
 * scratch kept resident between requests), take requests from one FIFO
 * queue; a request is admitted when a slot is free and the device bytes it
 * holds while running (its buffers plus its scratch) fit in what is left of
 * hbm_budget, so requests that arrive while others run start as soon as room
 * frees up and overlap on the GPU.  A slot that starts a small request (at
 * most 4M elements) also takes the compatible requests of the same kind
 * queued right behind it (same block size and dictionary, at most 4096; no
 * padding for compress) and runs them as one fused batch launch, compress or
 * decompress.  Results are those of vecsz_compress / vecsz_decompress on the
 * same request.  Buffers are the caller's device memory and must stay valid
 * until the request is waited for. */
typedef struct vecsz_server vecsz_server;
typedef struct vecsz_server_result {
    int status;          /* VECSZ_* of the request */
    uint64_t out_bytes;  /* compress: archive size */
    uint64_t n_outlier;  /* compress: outlier count */
    double queue_ms;     /* submit -> start */
    double run_ms;       /* start -> done */
    double latency_ms;   /* submit -> done */
} vecsz_server_result;
int vecsz_server_create(vecsz_server** server, int device, int nslots, uint64_t hbm_budget);
int vecsz_server_destroy(vecsz_server* server);
/* VECSZ_ENOMEM when the request could never fit the budget. */
int vecsz_server_submit_compress(vecsz_server* server, const vecsz_params* p, const float* d_in, uint8_t* d_out,
                                 uint64_t out_cap, uint64_t* ticket);
int vecsz_server_submit_decompress(vecsz_server* server, const vecsz_params* p, const uint8_t* d_in,
                                   uint64_t in_bytes, float* d_out, uint64_t out_cap, uint64_t* ticket);
/* Blocks until the request is done; each ticket is waited for once. */
int vecsz_server_wait(vecsz_server* server, uint64_t ticket, vecsz_server_result* result);

/* Huffman codebook of a quant-code histogram, built on the host exactly as the
 * reference builds it: HuffmanTree init / build_tree / build_code
 * (src/huffman.hh:239-262, :200-227) and the tree serialization written into
 * the archive (encode_withTree, src/huffman.hh:425-502).  Part of every
 * compress call; exported so the codebook can be checked and reused on its
 * own.  h_code_bits / h_code_len receive `dict` entries (right-aligned
 * codeword, length; length 0 for absent symbols and for a one-symbol tree).
 * With h_tree NULL only *tree_bytes and *node_count are set. */
int vecsz_huffman_codebook(const uint64_t* h_freq, int dict, uint64_t* h_code_bits, uint8_t* h_code_len,
                           uint8_t* h_tree, uint64_t tree_cap, uint64_t* tree_bytes, int* node_count);

/* Metadata the archive does not carry (mode, B, dict, the r2r range and
 * eb_final: reference defect D7), as a small fixed-size record to keep beside
 * an archive so that decompress needs nothing else.  No counterpart in the
 * reference (it takes these from its command line).  The archive itself is
 * unchanged. */
#define VECSZ_SIDECAR_MAGIC 0x4D5A5356u /* "VSZM" */
typedef struct vecsz_sidecar {
    uint32_t magic;      /* VECSZ_SIDECAR_MAGIC */
    uint32_t version;    /* 1 */
    int32_t ndim;
    int32_t mode;        /* VECSZ_MODE_ABS or VECSZ_MODE_R2R */
    int32_t block;
    int32_t dict_size;
    double eb;           /* user eb, as stored in the archive */
    double value_range;  /* r2r: the range compress used; abs: 0 */
    double eb_final;
} vecsz_sidecar;
/* The record of this context's last successful compress. */
int vecsz_ctx_last_sidecar(vecsz_ctx* ctx, vecsz_sidecar* out);
/* Decompress parameters from a record (EINVAL on a bad magic or version). */
int vecsz_params_from_sidecar(const vecsz_sidecar* s, vecsz_params* p);

/* Quality of a reconstruction, computed on the device: VerifyData
 * (src/utils/verify.hh:30-74) into the reference's stat_t fields
 * (src/types.hh:52-63).  Per-element arithmetic is the reference's; its sums
 * are taken in parallel order, so they match the reference's serial sums to
 * rounding; min, max, max_abserr and its first index are exact. */
typedef struct vecsz_stat {
    double minimum, maximum, range;
    double PSNR, MSE, NRMSE;
    double coeff;
    double max_abserr_vs_range, max_pwr_rel_abserr;
    uint64_t len, max_abserr_index;
    double max_abserr;
} vecsz_stat;
int vecsz_verify(vecsz_ctx* ctx, const float* d_orig, const float* d_recon, uint64_t n, vecsz_stat* out);

/* Lossless pass over an archive payload, on the host (src/lossless.hh:179-219,
 * called at level 3 from src/vecsz_interface.hh:206-217).  GZIP uses zlib
 * 1.2.11, the reference's release, so the bytes match the reference's; ZSTD
 * uses the system zstd runtime, whose frames decode both ways but whose
 * compressed bytes differ from the reference's 1.3.5.  VECSZ_EINVAL when the
 * zstd runtime is absent. */
#define VECSZ_LOSSLESS_NONE 0
#define VECSZ_LOSSLESS_GZIP 1
#define VECSZ_LOSSLESS_ZSTD 2
int vecsz_lossless_bound(int kind, uint64_t bytes, uint64_t* bound);
int vecsz_lossless_compress(int kind, const uint8_t* h_in, uint64_t bytes, uint8_t* h_out, uint64_t out_cap,
                            uint64_t* out_bytes);
/* Kind of a lossless stream by its first bytes, as
 * is_lossless_compressed_data (src/lossless.hh:60-75); 0 when neither. */
int vecsz_lossless_kind(const uint8_t* h_in, uint64_t bytes);
/* Inverse of the pass: exactly `target` bytes (the payload size) or ECORRUPT. */
int vecsz_lossless_decompress(const uint8_t* h_in, uint64_t bytes, uint8_t* h_out, uint64_t target,
                              uint64_t* out_bytes);

/* .sz file container (src/vecsz.cc:43-56): [u64 lossless_size][u64 data_size]
 * then the lossless stream (lossless_size > 0) or the payload.  The
 * reference writes data_size bytes after the header even when the stream is
 * lossless_size long (D8); pack writes the exact body, unpack accepts both. */
int vecsz_sz_pack(const uint8_t* h_body, uint64_t body_bytes, uint64_t lossless_size, uint64_t data_size,
                  uint8_t* h_out, uint64_t out_cap, uint64_t* out_bytes);
int vecsz_sz_unpack(const uint8_t* h_file, uint64_t bytes, uint64_t* lossless_size, uint64_t* data_size,
                    uint64_t* body_offset, uint64_t* body_bytes);

/* Per-stage device timing of the last compress/decompress call, measured
 * with HIP events on the context's stream when enabled (off by default).
 * names/ms receive up to `cap` entries; returns the entry count. */
int vecsz_ctx_enable_timing(vecsz_ctx* ctx, int on);
int vecsz_ctx_timings(vecsz_ctx* ctx, const char** names, double* ms, int cap);

/* Library build identification (gfx target and version string). */
const char* vecsz_version(void);

#ifdef __cplusplus
}
#endif
#endif /* VECSZ_HIP_H */
