/* TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's dual-quantization + Huffman compression
 * path (vecSZ: src/dualquant.hh, src/huffman.hh, src/utils/datapack.hh,
 * src/types.cc, src/vecsz_interface.hh).  It is the correctness oracle for the
 * HIP implementation and may be timed as a CPU baseline; it is never linked
 * into, called by, or a fallback of the product library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every function states the reference lines it restates in oracle/vecsz_oracle.c.
 */
#ifndef VECSZ_ORACLE_H
#define VECSZ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_EINVAL = 1,
    ORC_ECORRUPT = 2,
    ORC_ECODELEN = 3,
    ORC_ENOMEM = 4,
    ORC_ERANGE0 = 5
};

/* Error-bound family (src/types.cc:49-60, :84-93). */
typedef struct {
    double eb_final;  /* eb (abs) or eb * range (r2r) */
    double r_d;       /* 1 / (2 eb_final) */
    double twoeb_d;   /* 2 eb_final */
    float r_f;        /* (float) r_d, as _mm256_set1_ps receives it */
} orc_ebs;

int orc_value_range(const float* x, size_t n, double* range);
int orc_eb_family(int r2r, double eb, double range, orc_ebs* out);

/* Padding (src/utils/padding.hh, src/dualquant.hh:36-46, :226-246,
 * :553-573, :920-927, :951-964, :1003-1016, src/vecsz_interface.hh:70-90,
 * :287-298).  Modes as szwf.{global,block,edge}_padding; value types as
 * MAX_VALUE..CONST_VAL (src/utils/padding.hh:10-13). */
enum { ORC_PAD_NONE = 0, ORC_PAD_GLOBAL = 1, ORC_PAD_BLOCK = 2, ORC_PAD_EDGE = 3 };
enum { ORC_PADV_MAX = 0, ORC_PADV_MIN = 1, ORC_PADV_AVG = 2, ORC_PADV_CONST = 3 };

/* Number of Lorenzo blocks along a dimension as the reference counts them:
 * (d + B + 1) / B (src/dimensions.hh:31), one more than ceil(d / B) when
 * d % B is 0 or B - 1. */
size_t orc_nblk(size_t d, int block);

/* Pad values of a compress: pad_vals[] as the archive stores them (*pad_len
 * of them; pad_vals may be NULL to query the count) and, per block of the
 * reference's block grid, the three halo values (z face, y face, x face) its
 * dual quantization reads (halo may be NULL).  Elements outside the field
 * read as +0.0. */
int orc_pad_values(int ndim, const size_t* dims, int block, int mode, int type, float constant,
                   const orc_ebs* ebs, const float* x, float* pad_vals, size_t* pad_len, float* halo);

/* Halo values decompress derives from an archive's pad section
 * (src/vecsz_interface.hh:287-298, x_lorenzo_*): 3 per block of the
 * reference's block grid. */
int orc_pad_halo_from_archive(int ndim, const size_t* dims, int block, const float* pad_vals,
                              uint64_t pad_len, float* halo);

/* Dual quantization with halo values (3 per block, NULL = the +0.0 halo). */
int orc_dualquant_halo(int ndim, const size_t* dims, int block, int radius, const orc_ebs* ebs,
                       const float* x, const float* halo, int32_t* code, float* outlier);

/* Dual quantization of a whole field (dims x fastest, unused dims = 1).
 * code[] receives the int quant codes, outlier[] the per-element outlier
 * scratch exactly as Compress leaves them (src/vecsz_interface.hh:45-46,
 * :106-140). */
int orc_dualquant(int ndim, const size_t* dims, int block, int radius, const orc_ebs* ebs,
                  const float* x, int32_t* code, float* outlier);

/* Outlier gather in index order (src/vecsz_interface.hh:166-170). */
size_t orc_gather_outliers(const int32_t* code, const float* outlier, size_t n, float* out);

/* Huffman codebook (src/huffman.hh:116-266): returns the per-symbol code
 * (right-aligned bits) and length for symbols [0, state_num), plus the
 * serialized tree (src/huffman.hh:425-502) and node count.  tree_out may be
 * NULL to query tree_bytes. */
int orc_huffman_codebook(const uint64_t* freq, int state_num, uint64_t* code_bits,
                         uint8_t* code_len, int* node_count, uint8_t* tree_out,
                         size_t* tree_bytes);

/* Bit packing (src/huffman.hh:268-341). Returns the byte count written. */
size_t orc_huffman_pack(const int32_t* sym, size_t n, const uint64_t* code_bits,
                        const uint8_t* code_len, uint8_t* out);

/* Whole compress: returns a malloc'd archive payload (src/utils/datapack.hh:17-29). */
int orc_compress(int ndim, const size_t* dims, int r2r, double eb, int dict, int block,
                 const float* x, uint8_t** archive, size_t* archive_size, size_t* n_outlier,
                 double* eb_final_out);

/* Compress with padding (pad_mode ORC_PAD_*, pad_type ORC_PADV_*). */
int orc_compress_pad(int ndim, const size_t* dims, int r2r, double eb, int dict, int block, int pad_mode,
                     int pad_type, float pad_constant, const float* x, uint8_t** archive, size_t* archive_size,
                     size_t* n_outlier, double* eb_final_out);

/* Header fields of an archive payload (src/utils/datapack.hh:33-56). */
int orc_archive_info(const uint8_t* archive, size_t size, int* ndim, size_t* dims, double* eb,
                     uint64_t* pad_len, uint64_t* huff_size, uint64_t* n_outlier);

/* Huffman decode of an archive's huff section (src/huffman.hh:561-691). */
int orc_huffman_decode(const uint8_t* huff, size_t huff_size, size_t n, int32_t* out);

/* Whole decompress (src/vecsz_interface.hh:232-375).  `range` is the r2r value
 * range the reference recomputes from the original file (ignored for abs). */
int orc_decompress(const uint8_t* archive, size_t size, int r2r, double range, int dict,
                   int block, float* out, size_t out_cap);

/* Lorenzo reconstruction from codes + scattered outliers (src/dualquant.hh:915-1043). */
int orc_reconstruct(int ndim, const size_t* dims, int block, int radius, double twoeb,
                    const int32_t* code, const float* outlier, float* out);
/* The same with halo values (3 per block, NULL = the +0.0 halo). */
int orc_reconstruct_halo(int ndim, const size_t* dims, int block, int radius, double twoeb,
                         const int32_t* code, const float* outlier, const float* halo, float* out);

void orc_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
