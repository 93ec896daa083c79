/* TEST INFRASTRUCTURE ONLY -- the correctness oracle, never the product.
 *
 * A plain-C restatement of vecSZ's compression path, written from the
 * reference's behaviour (file:line cited per function).  It is loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, always as
 * the checker.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off).
 *
 * Floating-point evaluation follows the reference exactly: every expression
 * below is evaluated in the same type and association as the reference's
 * AVX2 or scalar code for that element, with no contraction and no
 * reassociation (FLT_EVAL_METHOD == 0 on x86-64).
 *
 * Each part is pinned by the tests under tests/ (the compiled reference's
 * outputs, golden fixtures, and closed-form and brute-force checks).
 */
#include "vecsz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MIN(a, b) ((a) < (b) ? (a) : (b))

/* ---------------------------------------------------------------- eb family */

/* GetDatumValueRange<float> (src/types.cc:27-35): std::max_element and
 * std::min_element keep the first extreme they meet, and the difference is
 * taken in float before it widens to double. */
int orc_value_range(const float* x, size_t n, double* range) {
    if (n == 0) return ORC_EINVAL;
    float mx = x[0], mn = x[0];
    for (size_t i = 1; i < n; i++) {
        if (mx < x[i]) mx = x[i];
        if (x[i] < mn) mn = x[i];
    }
    float r = mx - mn;
    *range = (double)r;
    return ORC_OK;
}

/* ErrorBoundConfigurator (src/types.cc:40-60) + InitializeErrorBoundFamily
 * (src/types.cc:84-93); r_f is what _mm256_set1_ps(ebs_L4[EBx2_r]) holds
 * (src/dualquant.hh:56, :264, :592). */
int orc_eb_family(int r2r, double eb, double range, orc_ebs* out) {
    double ebf = eb;
    if (r2r) {
        if (range == 0) return ORC_ERANGE0; /* src/types.cc:51-54 exits here */
        ebf *= range;
    }
    out->eb_final = ebf;
    out->r_d = 1 / (2 * ebf);
    out->twoeb_d = 2 * ebf;
    out->r_f = (float)out->r_d;
    return ORC_OK;
}

/* ------------------------------------------------------------- dual quant */

/* AVX prequant: _mm256_round_ps(_mm256_mul_ps(x, r_f), NEAREST_INT)
 * (src/dualquant.hh:89, :290, :422, :622, :784). */
static inline float prequant_vec(float x, float r_f) { return rintf(x * r_f); }

/* Scalar prequant: round(data[id] * ebs_L4[EBx2_r]) stored to float
 * (src/dualquant.hh:109, :298, :432, :631, :796). */
static inline float prequant_scalar(float x, double r_d) { return (float)round((double)x * r_d); }

/* c_lorenzo_{1,2,3}d1l (src/dualquant.hh:18-913), restated per element.
 *
 * Which form each element takes (from the branches cited below):
 *  - 1D: AVX prequant everywhere (:65, :83-93).  Postquant uses the AVX test
 *    sqrt(e*e) < R in blocks with start + B < DIM0 (:114, :157-178), else the
 *    scalar test |e| < R (:190-205).
 *  - 2D: every element with x < 8*floor(DIM0/8) is AVX prequant + AVX
 *    postquant with pred = W + (N - NW) (:287-292, :350-368, :419-424,
 *    :490-512); elements with x >= 8*floor(DIM0/8) are skipped by the
 *    `continue` at :494-495 and keep code 0 / outlier +0.0.
 *  - 3D: x < 8*floor(DIM0/8): AVX prequant, pred = (XYZ - (XY + (XZ + YZ))) +
 *    (X + (Y + Z)) (:619-624, :691-714, :776-787, :863-891); otherwise the
 *    scalar prequant and pred = ((((((XYZ-XY)-XZ)-YZ)+X)+Y)+Z) (:788-797,
 *    :892-908).
 *  Neighbours in another Lorenzo block read the zero halo (:224, :246, :551,
 *  :573).  code = quantizable ? (int)(e + R) : 0; outlier = q where code == 0.
 */
/* Value of a neighbour at offsets (dz, dy, dx) in {0, 1} of an element at
 * in-block position (iz, iy, ix), given the element's own block halo
 * h = {z face, y face, x face}.  A neighbour inside the block reads v.
 * Halo cells take the face that the reference's padded block gives them
 * first: 3D (k == 0) ? z : (j == 0) ? y : x (src/utils/padding.hh:245, :290,
 * :330, :365); 2D (i == 0) ? x : y (:226, :275, :316, :353). */
static inline float halo_nb(int ndim, const float* h, int dz, int dy, int dx, size_t iz, size_t iy, size_t ix,
                            const float* v, size_t off) {
    static const float zero3[3] = {0.0f, 0.0f, 0.0f};
    if (!h) h = zero3;
    int oz = dz && iz == 0, oy = dy && iy == 0, ox = dx && ix == 0;
    if (ndim == 3) {
        if (oz) return h[0];
        if (oy) return h[1];
        if (ox) return h[2];
    } else if (ndim == 2) {
        if (ox) return h[2];
        if (oy) return h[1];
    } else if (ox) {
        return h[2];
    }
    return v[off];
}

int orc_dualquant(int ndim, const size_t* dims, int block, int radius, const orc_ebs* ebs,
                  const float* x, int32_t* code, float* outlier) {
    return orc_dualquant_halo(ndim, dims, block, radius, ebs, x, NULL, code, outlier);
}

int orc_dualquant_halo(int ndim, const size_t* dims, int block, int radius, const orc_ebs* ebs,
                       const float* x, const float* halo, int32_t* code, float* outlier) {
    if (ndim < 1 || ndim > 3 || block <= 0 || (block % 8) != 0) return ORC_EINVAL;
    size_t d0 = dims[0], d1 = ndim >= 2 ? dims[1] : 1, d2 = ndim >= 3 ? dims[2] : 1;
    size_t n = d0 * d1 * d2;
    size_t B = (size_t)block;
    size_t x8 = (d0 / 8) * 8;
    const float R = (float)radius;
    const float rf = ebs->r_f;
    const double rd = ebs->r_d;
    const size_t nb0 = orc_nblk(d0, block), nb1 = ndim >= 2 ? orc_nblk(d1, block) : 1;
    float* q = (float*)malloc(n * sizeof(float) + 4);
    if (!q) return ORC_ENOMEM;

    for (size_t z = 0; z < d2; z++)
        for (size_t y = 0; y < d1; y++)
            for (size_t i = 0; i < d0; i++) {
                size_t id = i + d0 * (y + d1 * z);
                q[id] = (ndim == 3 && i >= x8) ? prequant_scalar(x[id], rd) : prequant_vec(x[id], rf);
            }

    memset(code, 0, n * sizeof(int32_t));
    memset(outlier, 0, n * sizeof(float));

    const size_t sy = d0, sz = d0 * d1;
    for (size_t z = 0; z < d2; z++)
        for (size_t y = 0; y < d1; y++)
            for (size_t i = 0; i < d0; i++) {
                size_t id = i + d0 * (y + d1 * z);
                size_t ix = i % B, iy = y % B, iz = z % B;
                const float* h = halo ? halo + 3 * ((i / B) + nb0 * ((y / B) + nb1 * (z / B))) : NULL;
                float cur = q[id], pred;
                int vec_test;
#define NB(dz, dy, dx) halo_nb(ndim, h, dz, dy, dx, iz, iy, ix, q, id - (dx) - (dy) * sy - (dz) * sz)
                if (ndim == 1) {
                    pred = NB(0, 0, 1);
                    size_t bstart = i - ix;
                    vec_test = (bstart + B) < d0;
                } else if (ndim == 2) {
                    if (i >= x8) continue; /* D2: lost partial x-group */
                    float W = NB(0, 0, 1);
                    float N = NB(0, 1, 0);
                    float NW = NB(0, 1, 1);
                    pred = W + (N - NW);
                    vec_test = 1;
                } else {
                    float X = NB(0, 0, 1);
                    float Y = NB(0, 1, 0);
                    float Z = NB(1, 0, 0);
                    float XY = NB(0, 1, 1);
                    float XZ = NB(1, 0, 1);
                    float YZ = NB(1, 1, 0);
                    float XYZ = NB(1, 1, 1);
                    if (i < x8) {
                        float dist1 = X + (Y + Z);
                        float dist2 = XY + (XZ + YZ);
                        pred = (XYZ - dist2) + dist1;
                        vec_test = 1;
                    } else {
                        pred = XYZ - XY - XZ - YZ + X + Y + Z;
                        vec_test = 0;
                    }
                }
#undef NB
                float e = cur - pred;
                int quant = vec_test ? (sqrtf(e * e) < R) : (fabsf(e) < R);
                if (quant) {
                    /* the AVX forms convert with _mm256_cvtps_epi32 (round to
                     * nearest even, src/dualquant.hh:172, :362, :506, :708, :885),
                     * the scalar forms with static_cast (truncation, :200, :378,
                     * :524, :726, :905); they differ only for a non-integral e, which
                     * arises in 1D block/edge padding (raw pad value) */
                    code[id] = vec_test ? (int32_t)rintf(e + R) : (int32_t)(e + R);
                } else {
                    code[id] = 0;
                    outlier[id] = cur;
                }
            }
    free(q);
    return ORC_OK;
}

/* ------------------------------------------------------------- padding */

size_t orc_nblk(size_t d, int block) { return (d + (size_t)block + 1) / (size_t)block; }

/* find_pad_value (src/utils/padding.hh:121-143) over a linear array in
 * index order: find_max / find_min keep the first value and replace it by a
 * strictly larger / smaller one (:65-99); find_avg sums in float and divides
 * by the float element count (:101-119). */
static float pad_find(const float* a, size_t count, int type, float constant) {
    if (type == ORC_PADV_CONST) return constant;
    if (type == ORC_PADV_AVG) {
        float sum = 0.0f;
        for (size_t i = 0; i < count; i++) sum += a[i];
        return sum / (float)count;
    }
    float m = a[0];
    for (size_t i = 0; i < count; i++) {
        if (type == ORC_PADV_MAX) m = (a[i] > m) ? a[i] : m;
        else m = (a[i] < m) ? a[i] : m;
    }
    return m;
}

/* block_pad / edge_pad prequantization of a pad value (src/utils/padding.hh:
 * 219-220, :236-237, :259-263): the float is scaled in double, stored back to
 * float, then rounded half away from zero. */
static float pad_prequant_block(float p, double r_d) {
    float t = (float)((double)p * r_d);
    return (float)round((double)t);
}

/* Global pad value as Compress stores it: round(pad_constant * ebx2_r) in
 * double, stored to float (src/vecsz_interface.hh:79). */
static float pad_prequant_global(float c, double r_d) { return (float)round((double)c * r_d); }

/* fill_2d_block / fill_3d_block (src/utils/padding.hh:17-63): the block's
 * elements in x-fastest order; elements outside the field read +0.0. */
static void pad_fill(int ndim, const float* x, size_t d0, size_t d1, size_t d2, size_t B, size_t x0, size_t y0,
                     size_t z0, float* blk) {
    size_t nz = ndim >= 3 ? B : 1, ny = ndim >= 2 ? B : 1;
    for (size_t k = 0; k < nz; k++)
        for (size_t j = 0; j < ny; j++)
            for (size_t i = 0; i < B; i++) {
                size_t gx = x0 + i, gy = y0 + j, gz = z0 + k;
                int in = gx < d0 && gy < d1 && gz < d2;
                blk[i + B * (j + ny * k)] = in ? x[gx + d0 * (gy + d1 * gz)] : 0.0f;
            }
}

int orc_pad_values(int ndim, const size_t* dims, int block, int mode, int type, float constant,
                   const orc_ebs* ebs, const float* x, float* pad_vals, size_t* pad_len, float* halo) {
    if (ndim < 1 || ndim > 3 || block <= 0 || mode < ORC_PAD_NONE || mode > ORC_PAD_EDGE || type < 0 ||
        type > ORC_PADV_CONST)
        return ORC_EINVAL;
    size_t d0 = dims[0], d1 = ndim >= 2 ? dims[1] : 1, d2 = ndim >= 3 ? dims[2] : 1;
    size_t B = (size_t)block;
    size_t nb0 = orc_nblk(d0, block), nb1 = ndim >= 2 ? orc_nblk(d1, block) : 1, nb2 = ndim >= 3 ? orc_nblk(d2, block) : 1;
    size_t nblocks = nb0 * nb1 * nb2;
    size_t len = mode == ORC_PAD_NONE ? 0 : mode == ORC_PAD_GLOBAL ? 1 : mode == ORC_PAD_BLOCK ? nblocks : (size_t)ndim * nblocks;
    *pad_len = len;
    if (!pad_vals && !halo) return ORC_OK;
    const double rd = ebs->r_d;
    if (halo) memset(halo, 0, 3 * nblocks * sizeof(float));
    if (mode == ORC_PAD_NONE) return ORC_OK;
    if (mode == ORC_PAD_GLOBAL) {
        /* the whole field (src/vecsz_interface.hh:73-80), then every block pads
         * with that constant (block_pad with CONST_VAL, :77) */
        float c = pad_find(x, d0 * d1 * d2, type, constant);
        float stored = pad_prequant_global(c, rd);
        float used = ndim == 1 ? stored : pad_prequant_block(c, rd); /* 1D reads pad_vals[0] (src/dualquant.hh:43-46) */
        if (pad_vals) pad_vals[0] = stored;
        if (halo)
            for (size_t b = 0; b < nblocks; b++) halo[3 * b] = halo[3 * b + 1] = halo[3 * b + 2] = used;
        return ORC_OK;
    }
    size_t bsz = B * (ndim >= 2 ? B : 1) * (ndim >= 3 ? B : 1);
    float* blk = (float*)malloc(bsz * sizeof(float));
    float* edge = (float*)malloc(B * B * sizeof(float));
    if (!blk || !edge) { free(blk); free(edge); return ORC_ENOMEM; }
    size_t pi = 0;
    for (size_t b2 = 0; b2 < nb2; b2++)
        for (size_t b1 = 0; b1 < nb1; b1++)
            for (size_t b0 = 0; b0 < nb0; b0++) {
                size_t id = b0 + nb0 * (b1 + nb1 * b2);
                float hz, hy, hx;
                pad_fill(ndim, x, d0, d1, d2, B, b0 * B, b1 * B, b2 * B, blk);
                if (ndim == 1) {
                    /* 1D block and edge padding: the raw pad value of the block's
                     * data, not prequantized (src/dualquant.hh:37-42) */
                    float p = pad_find(blk, B, type, constant);
                    if (pad_vals) pad_vals[pi] = p;
                    pi++;
                    hz = hy = hx = p;
                } else if (mode == ORC_PAD_BLOCK) {
                    float p = pad_prequant_block(pad_find(blk, bsz, type, constant), rd);
                    if (pad_vals) pad_vals[pi] = p;
                    pi++;
                    hz = hy = hx = p;
                } else if (ndim == 2) {
                    /* find_edge_pad_value (src/utils/padding.hh:145-206): pv0 from
                     * the block's first row, pv1 from its first column */
                    float pv[2];
                    for (size_t i = 0; i < B; i++) edge[i] = blk[i];
                    pv[0] = pad_find(edge, B, type, constant);
                    for (size_t i = 0; i < B; i++) edge[i] = blk[i * B];
                    pv[1] = pad_find(edge, B, type, constant);
                    for (int f = 0; f < 2; f++) {
                        pv[f] = pad_prequant_block(pv[f], rd);
                        if (pad_vals) pad_vals[pi] = pv[f];
                        pi++;
                    }
                    hz = 0.0f;
                    hx = pv[0];
                    hy = pv[1];
                } else {
                    /* 3D: pv0 from the z = 0 plane, pv1 from the x = 0 plane,
                     * pv2 from the y = 0 plane */
                    float pv[3];
                    for (size_t j = 0; j < B; j++)
                        for (size_t i = 0; i < B; i++) edge[i + j * B] = blk[i + j * B];
                    pv[0] = pad_find(edge, B * B, type, constant);
                    for (size_t j = 0; j < B; j++)
                        for (size_t i = 0; i < B; i++) edge[i + j * B] = blk[i * B + j * B * B];
                    pv[1] = pad_find(edge, B * B, type, constant);
                    for (size_t j = 0; j < B; j++)
                        for (size_t i = 0; i < B; i++) edge[i + j * B] = blk[i + j * B * B];
                    pv[2] = pad_find(edge, B * B, type, constant);
                    for (int f = 0; f < 3; f++) {
                        pv[f] = pad_prequant_block(pv[f], rd);
                        if (pad_vals) pad_vals[pi] = pv[f];
                        pi++;
                    }
                    hz = pv[0];
                    hy = pv[1];
                    hx = pv[2];
                }
                if (halo) {
                    halo[3 * id] = hz;
                    halo[3 * id + 1] = hy;
                    halo[3 * id + 2] = hx;
                }
            }
    free(blk);
    free(edge);
    return ORC_OK;
}

int orc_pad_halo_from_archive(int ndim, const size_t* dims, int block, const float* pad_vals, uint64_t pad_len,
                              float* halo) {
    size_t d0 = dims[0], d1 = ndim >= 2 ? dims[1] : 1, d2 = ndim >= 3 ? dims[2] : 1;
    size_t nb0 = orc_nblk(d0, block), nb1 = ndim >= 2 ? orc_nblk(d1, block) : 1, nb2 = ndim >= 3 ? orc_nblk(d2, block) : 1;
    size_t nblocks = nb0 * nb1 * nb2;
    int mode = pad_len == 1 ? ORC_PAD_GLOBAL
             : pad_len == nblocks ? ORC_PAD_BLOCK
             : pad_len == (uint64_t)ndim * nblocks ? ORC_PAD_EDGE : ORC_PAD_NONE;
    for (size_t b = 0; b < nblocks; b++) {
        float hz = 0.0f, hy = 0.0f, hx = 0.0f;
        if (mode == ORC_PAD_GLOBAL) {
            hz = hy = hx = pad_vals[0];
        } else if (mode == ORC_PAD_BLOCK) {
            hz = hy = hx = pad_vals[b];
        } else if (mode == ORC_PAD_EDGE) {
            if (ndim == 2) {
                hx = pad_vals[2 * b];
                hy = pad_vals[2 * b + 1];
            } else {
                hz = pad_vals[3 * b];
                hy = pad_vals[3 * b + 1];
                hx = pad_vals[3 * b + 2];
            }
        } else if (ndim == 1) {
            /* x_lorenzo_1d1l reads pad_vals[0] without a padding mode
             * (src/dualquant.hh:925-927); with no pad section that is D3's
             * read of new float[0], taken as +0.0 */
            hz = hy = hx = pad_len > 0 ? pad_vals[0] : 0.0f;
        }
        halo[3 * b] = hz;
        halo[3 * b + 1] = hy;
        halo[3 * b + 2] = hx;
    }
    return ORC_OK;
}

/* src/vecsz_interface.hh:166-170 */
size_t orc_gather_outliers(const int32_t* code, const float* outlier, size_t n, float* out) {
    size_t k = 0;
    for (size_t i = 0; i < n; i++)
        if (code[i] == 0) out[k++] = outlier[i];
    return k;
}

/* ---------------------------------------------------------------- Huffman */

typedef struct {
    int32_t left, right; /* -1: none */
    uint64_t freq;
    uint8_t t;  /* 1: leaf */
    uint32_t c; /* symbol; 0 for internal nodes (pool is zeroed, :127) */
} onode;

typedef struct {
    onode* pool;
    int n_nodes;
    int32_t* qq; /* 1-indexed heap of pool indices */
    int qend;
} oheap;

/* qinsert (src/huffman.hh:169-177): sift up, stop at a parent with freq <= n. */
static void qinsert(oheap* h, int32_t n) {
    int j, i = h->qend++;
    while ((j = (i >> 1))) {
        if (h->pool[h->qq[j]].freq <= h->pool[n].freq) break;
        h->qq[i] = h->qq[j];
        i = j;
    }
    h->qq[i] = n;
}

/* qremove (src/huffman.hh:179-192): the hole walks down to a leaf taking the
 * strictly smaller child, then the last element drops into it with no
 * sift-up. */
static int32_t qremove(oheap* h) {
    int i = 1, l;
    int32_t n = h->qq[1];
    if (h->qend < 2) return -1;
    h->qend--;
    while ((l = (i << 1)) < h->qend) {
        if (l + 1 < h->qend && h->pool[h->qq[l + 1]].freq < h->pool[h->qq[l]].freq) l++;
        h->qq[i] = h->qq[l];
        i = l;
    }
    h->qq[i] = h->qq[h->qend];
    return n;
}

/* initHuffman + build_code (src/huffman.hh:200-266), the tree serialization
 * convert_HuffTree_to_bytes_anyStates / pad_tree_* (:374-502), and the node
 * count of encode_withTree (:643-645).
 *
 * Merge order: `new_node(.., qremove(), qremove())` (:262) is evaluated right
 * to left by g++, so the first node removed becomes the RIGHT child. */
int orc_huffman_codebook(const uint64_t* freq, int state_num, uint64_t* code_bits,
                         uint8_t* code_len, int* node_count, uint8_t* tree_out,
                         size_t* tree_bytes) {
    int all = 2 * state_num;
    oheap h;
    h.pool = (onode*)calloc((size_t)all * 2, sizeof(onode));
    h.qq = (int32_t*)calloc((size_t)all * 2 + 2, sizeof(int32_t));
    h.n_nodes = 0;
    h.qend = 1;
    if (!h.pool || !h.qq) return ORC_ENOMEM;
    int nsym = 0;
    for (int s = 0; s < state_num; s++) {
        code_len[s] = 0;
        code_bits[s] = 0;
        if (freq[s]) {
            int32_t n = h.n_nodes++;
            h.pool[n].left = h.pool[n].right = -1;
            h.pool[n].freq = freq[s];
            h.pool[n].t = 1;
            h.pool[n].c = (uint32_t)s;
            qinsert(&h, n);
            nsym++;
        }
    }
    if (nsym == 0) {
        free(h.pool);
        free(h.qq);
        return ORC_EINVAL;
    }
    while (h.qend > 2) {
        int32_t b = qremove(&h); /* first removed -> right */
        int32_t a = qremove(&h); /* second removed -> left */
        int32_t n = h.n_nodes++;
        h.pool[n].left = a;
        h.pool[n].right = b;
        h.pool[n].freq = h.pool[a].freq + h.pool[b].freq;
        h.pool[n].t = 0;
        h.pool[n].c = 0;
        qinsert(&h, n);
    }
    int32_t root = h.qq[1];

    /* build_code: left appends 0, right appends 1, MSB first. */
    int rc = ORC_OK;
    int32_t* stk = (int32_t*)malloc(sizeof(int32_t) * (size_t)(h.n_nodes + 1));
    int* sl = (int*)malloc(sizeof(int) * (size_t)(h.n_nodes + 1));
    uint64_t* sb = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(h.n_nodes + 1));
    int sp = 0;
    stk[sp] = root; sl[sp] = 0; sb[sp] = 0; sp++;
    while (sp > 0) {
        sp--;
        int32_t n = stk[sp];
        int len = sl[sp];
        uint64_t bits = sb[sp];
        if (h.pool[n].t) {
            if (len > 64) rc = ORC_ECODELEN;
            code_bits[h.pool[n].c] = len > 64 ? 0 : bits;
            code_len[h.pool[n].c] = (uint8_t)(len > 255 ? 255 : len);
            continue;
        }
        stk[sp] = h.pool[n].right; sl[sp] = len + 1; sb[sp] = (bits << 1) | 1; sp++;
        stk[sp] = h.pool[n].left;  sl[sp] = len + 1; sb[sp] = (bits << 1);     sp++;
    }

    int nc = 2 * nsym - 1;
    *node_count = nc;
    int w = nc <= 256 ? 1 : (nc <= 65536 ? 2 : 4);
    size_t total = 1 + (size_t)nc * (2 * w + 4 + 1);
    *tree_bytes = total;
    if (tree_out) {
        uint32_t* L = (uint32_t*)calloc((size_t)nc, sizeof(uint32_t));
        uint32_t* Rr = (uint32_t*)calloc((size_t)nc, sizeof(uint32_t));
        uint32_t* C = (uint32_t*)calloc((size_t)nc, sizeof(uint32_t));
        uint8_t* T = (uint8_t*)calloc((size_t)nc, 1);
        /* preorder numbering: root 0, each child takes ++n_inode when visited */
        int32_t* pidx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(h.n_nodes + 1));
        int* pside = (int*)malloc(sizeof(int) * (size_t)(h.n_nodes + 1));
        uint32_t n_inode = 0;
        int first = 1;
        sp = 0;
        stk[sp] = root; pidx[sp] = -1; pside[sp] = 0; sp++;
        while (sp > 0) {
            sp--;
            int32_t n = stk[sp];
            int32_t par = pidx[sp];
            int side = pside[sp];
            uint32_t idx = first ? 0 : ++n_inode;
            first = 0;
            if (par >= 0) {
                if (side == 0) L[par] = idx; else Rr[par] = idx;
            }
            C[idx] = h.pool[n].c;
            T[idx] = h.pool[n].t;
            if (h.pool[n].right >= 0) { stk[sp] = h.pool[n].right; pidx[sp] = (int32_t)idx; pside[sp] = 1; sp++; }
            if (h.pool[n].left >= 0)  { stk[sp] = h.pool[n].left;  pidx[sp] = (int32_t)idx; pside[sp] = 0; sp++; }
        }
        uint8_t* o = tree_out;
        *o++ = 0; /* CPU_sysEndianType (src/huffman.hh:72, :440) */
        for (int pass = 0; pass < 2; pass++) {
            const uint32_t* A = pass == 0 ? L : Rr;
            for (int i = 0; i < nc; i++) {
                uint32_t v = A[i];
                for (int b = 0; b < w; b++) *o++ = (uint8_t)(v >> (8 * b));
            }
        }
        for (int i = 0; i < nc; i++)
            for (int b = 0; b < 4; b++) *o++ = (uint8_t)(C[i] >> (8 * b));
        for (int i = 0; i < nc; i++) *o++ = T[i];
        free(L); free(Rr); free(C); free(T); free(pidx); free(pside);
    }
    free(stk); free(sl); free(sb);
    free(h.pool);
    free(h.qq);
    return rc;
}

/* HuffmanEncode (src/huffman.hh:268-341): codewords concatenated MSB first,
 * the last byte zero-padded; outSize = ceil(total bits / 8). */
size_t orc_huffman_pack(const int32_t* sym, size_t n, const uint64_t* code_bits,
                        const uint8_t* code_len, uint8_t* out) {
    uint64_t acc = 0; /* pending bits, right-aligned */
    int nacc = 0;
    size_t o = 0;
    for (size_t i = 0; i < n; i++) {
        int len = code_len[sym[i]];
        uint64_t bits = code_bits[sym[i]];
        while (len > 0) {
            int take = MIN(len, 56 - nacc);
            uint64_t part = (bits >> (len - take)) & ((take == 64) ? ~0ull : ((1ull << take) - 1));
            acc = (acc << take) | part;
            nacc += take;
            len -= take;
            while (nacc >= 8) {
                out[o++] = (uint8_t)(acc >> (nacc - 8));
                nacc -= 8;
            }
        }
    }
    if (nacc > 0) out[o++] = (uint8_t)(acc << (8 - nacc));
    return o;
}

static void put_u32be(uint8_t* b, uint32_t v) {
    b[0] = (uint8_t)(v >> 24); b[1] = (uint8_t)(v >> 16); b[2] = (uint8_t)(v >> 8); b[3] = (uint8_t)v;
}

/* Compress (src/vecsz_interface.hh:31-229) without the optional lossless pass:
 * dual quant, Huffman (encode_withTree, src/huffman.hh:637-662), outlier
 * gather, sizes (:189-202) and datapack::pack (src/utils/datapack.hh:17-29). */
int orc_compress(int ndim, const size_t* dims, int r2r, double eb, int dict, int block,
                 const float* x, uint8_t** archive, size_t* archive_size, size_t* n_outlier,
                 double* eb_final_out) {
    return orc_compress_pad(ndim, dims, r2r, eb, dict, block, ORC_PAD_NONE, ORC_PADV_CONST, 0.0f, x, archive,
                            archive_size, n_outlier, eb_final_out);
}

int orc_compress_pad(int ndim, const size_t* dims, int r2r, double eb, int dict, int block, int pad_mode,
                     int pad_type, float pad_constant, const float* x, uint8_t** archive, size_t* archive_size,
                     size_t* n_outlier, double* eb_final_out) {
    if (ndim < 1 || ndim > 3 || dict < 2 || (dict % 2) != 0 || dict > 65536) return ORC_EINVAL;
    size_t n = 1;
    size_t dd[3] = {1, 1, 1};
    for (int i = 0; i < ndim; i++) { dd[i] = dims[i]; n *= dims[i]; }
    if (n == 0) return ORC_EINVAL;
    double range = 0;
    if (r2r) orc_value_range(x, n, &range);
    orc_ebs ebs;
    int rc = orc_eb_family(r2r, eb, range, &ebs);
    if (rc) return rc;
    if (eb_final_out) *eb_final_out = ebs.eb_final;
    int radius = dict / 2;
    int32_t* code = (int32_t*)malloc(n * sizeof(int32_t));
    float* outl = (float*)malloc(n * sizeof(float));
    if (!code || !outl) return ORC_ENOMEM;
    size_t pad_len = 0;
    rc = orc_pad_values(ndim, dd, block, pad_mode, pad_type, pad_constant, &ebs, x, NULL, &pad_len, NULL);
    if (rc) return rc;
    size_t nblocks = orc_nblk(dd[0], block) * (ndim >= 2 ? orc_nblk(dd[1], block) : 1) *
                     (ndim >= 3 ? orc_nblk(dd[2], block) : 1);
    float* pvals = (float*)malloc(4 * pad_len + 4);
    float* halo = (float*)malloc(3 * nblocks * sizeof(float) + 4);
    if (!pvals || !halo) return ORC_ENOMEM;
    rc = orc_pad_values(ndim, dd, block, pad_mode, pad_type, pad_constant, &ebs, x, pvals, &pad_len, halo);
    if (rc) return rc;
    rc = orc_dualquant_halo(ndim, dd, block, radius, &ebs, x, pad_mode == ORC_PAD_NONE ? NULL : halo, code, outl);
    if (rc) return rc;

    int state_num = 2 * dict; /* createHuffmanTree(dict_size * 2) (src/vecsz_interface.hh:160) */
    uint64_t* freq = (uint64_t*)calloc((size_t)state_num, sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) freq[code[i]]++;
    uint64_t* cb = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)state_num);
    uint8_t* cl = (uint8_t*)malloc((size_t)state_num);
    int nc;
    size_t tb;
    rc = orc_huffman_codebook(freq, state_num, cb, cl, &nc, NULL, &tb);
    if (rc) return rc;
    uint64_t total_bits = 0;
    for (int s = 0; s < state_num; s++) total_bits += freq[s] * cl[s];
    size_t enc = (size_t)((total_bits + 7) / 8);
    size_t huff = 8 + tb + enc;
    size_t nout = 0;
    for (size_t i = 0; i < n; i++) nout += code[i] == 0;
    size_t meta = sizeof(int32_t) * (size_t)(ndim + 1) + sizeof(double) + 3 * sizeof(uint64_t);
    size_t total = meta + 4 * pad_len + huff + 4 * nout;
    uint8_t* a = (uint8_t*)malloc(total + 8);
    if (!a) return ORC_ENOMEM;
    uint8_t* p = a;
    int32_t nd = ndim;
    memcpy(p, &nd, 4); p += 4;
    for (int i = 0; i < ndim; i++) { int32_t d = (int32_t)dd[i]; memcpy(p, &d, 4); p += 4; }
    memcpy(p, &eb, 8); p += 8; /* the user eb, not eb_final (src/vecsz_interface.hh:204) */
    uint64_t lens[3] = {(uint64_t)pad_len, (uint64_t)huff, (uint64_t)nout};
    memcpy(p, lens, 24); p += 24;
    memcpy(p, pvals, 4 * pad_len); p += 4 * pad_len;
    put_u32be(p, (uint32_t)nc);
    put_u32be(p + 4, (uint32_t)(state_num / 2));
    p += 8;
    orc_huffman_codebook(freq, state_num, cb, cl, &nc, p, &tb);
    p += tb;
    size_t w = orc_huffman_pack(code, n, cb, cl, p);
    (void)w;
    p += enc;
    size_t k = 0;
    for (size_t i = 0; i < n; i++)
        if (code[i] == 0) { memcpy(p + 4 * k, &outl[i], 4); k++; }
    *archive = a;
    *archive_size = total;
    *n_outlier = nout;
    free(code); free(outl); free(freq); free(cb); free(cl); free(pvals); free(halo);
    return ORC_OK;
}

/* datapack::unpack (src/utils/datapack.hh:33-56), with bounds checks. */
int orc_archive_info(const uint8_t* a, size_t size, int* ndim, size_t* dims, double* eb,
                     uint64_t* pad_len, uint64_t* huff_size, uint64_t* n_outlier) {
    if (size < 4) return ORC_ECORRUPT;
    int32_t nd;
    memcpy(&nd, a, 4);
    if (nd < 1 || nd > 3) return ORC_ECORRUPT;
    size_t meta = 4 * (size_t)(nd + 1) + 8 + 24;
    if (size < meta) return ORC_ECORRUPT;
    *ndim = nd;
    dims[0] = dims[1] = dims[2] = 1;
    for (int i = 0; i < nd; i++) {
        int32_t d;
        memcpy(&d, a + 4 + 4 * i, 4);
        if (d <= 0) return ORC_ECORRUPT;
        dims[i] = (size_t)d;
    }
    memcpy(eb, a + 4 * (nd + 1), 8);
    uint64_t lens[3];
    memcpy(lens, a + 4 * (nd + 1) + 8, 24);
    *pad_len = lens[0];
    *huff_size = lens[1];
    *n_outlier = lens[2];
    if (lens[0] > size || lens[1] > size || lens[2] > size ||
        meta + 4 * lens[0] + lens[1] + 4 * lens[2] > size)
        return ORC_ECORRUPT;
    return ORC_OK;
}

static uint32_t get_u32be(const uint8_t* b) {
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
}

/* decode_withTree + reconstruct_HuffTree_from_bytes_anyStates + decode
 * (src/huffman.hh:343-372, :504-635, :679-691).  The serialized arrays are
 * walked directly: node i's children are L[i] / R[i] (0 = absent); only nodes
 * with t == 0 are expanded (:506, :525, :544).  The root is (C[0], t[0]) when
 * nodeCount <= 256 and a forced internal node otherwise (:577, :602, :627). */
int orc_huffman_decode(const uint8_t* huff, size_t huff_size, size_t n, int32_t* out) {
    if (huff_size < 8) return ORC_ECORRUPT;
    int32_t ncs = (int32_t)get_u32be(huff);
    if (ncs <= 0) return ORC_ECORRUPT;
    size_t nc = (size_t)ncs;
    size_t w = nc <= 256 ? 1 : (nc <= 65536 ? 2 : 4);
    size_t tb = 1 + nc * (2 * w + 4 + 1);
    if (8 + tb > huff_size) return ORC_ECORRUPT;
    const uint8_t* t0 = huff + 8 + 1;
    const uint8_t* Lp = t0;
    const uint8_t* Rp = t0 + nc * w;
    const uint8_t* Cp = t0 + 2 * nc * w;
    const uint8_t* Tp = Cp + 4 * nc;
#define RD(P, i) (w == 1 ? (uint32_t)(P)[i] : (w == 2 ? (uint32_t)((P)[2*(i)] | ((P)[2*(i)+1] << 8)) : \
        ((uint32_t)(P)[4*(i)] | ((uint32_t)(P)[4*(i)+1] << 8) | ((uint32_t)(P)[4*(i)+2] << 16) | ((uint32_t)(P)[4*(i)+3] << 24))))
#define RC(i) ((uint32_t)Cp[4*(i)] | ((uint32_t)Cp[4*(i)+1] << 8) | ((uint32_t)Cp[4*(i)+2] << 16) | ((uint32_t)Cp[4*(i)+3] << 24))
    int root_leaf = nc <= 256 ? Tp[0] != 0 : 0;
    if (root_leaf) {
        int32_t c = (int32_t)RC(0);
        for (size_t i = 0; i < n; i++) out[i] = c;
        return ORC_OK;
    }
    const uint8_t* bits = huff + 8 + tb;
    size_t nbits = (huff_size - 8 - tb) * 8;
    size_t cnt = 0, bi = 0;
    uint32_t node = 0;
    while (cnt < n) {
        if (bi >= nbits) return ORC_ECORRUPT;
        int b = (bits[bi >> 3] >> (7 - (bi & 7))) & 1;
        bi++;
        uint32_t nx = b ? RD(Rp, node) : RD(Lp, node);
        if (nx == 0 || nx >= nc) return ORC_ECORRUPT;
        node = nx;
        if (Tp[node]) {
            out[cnt++] = (int32_t)RC(node);
            node = 0;
        }
    }
#undef RD
#undef RC
    return ORC_OK;
}

/* x_lorenzo_{1,2,3}d1l (src/dualquant.hh:915-1043):
 *   s = code == 0 ? outlier : (float)(pred + (float)(code - R))
 *   1D pred = block start ? pad(+0.0) : s[x-1]                 (:933-934)
 *   2D pred = (N + W) - NW                                       (:978)
 *   3D pred = ((((((XYZ-XY)-XZ)-YZ)+X)+Y)+Z)                     (:1034-1036)
 *   out = (float)((double)s * 2eb)                               (:941, :980, :1038)
 * Each element depends only on earlier neighbours inside its block, so a
 * global raster walk computes the same values as the per-block loops. */
int orc_reconstruct(int ndim, const size_t* dims, int block, int radius, double twoeb,
                    const int32_t* code, const float* outlier, float* out) {
    return orc_reconstruct_halo(ndim, dims, block, radius, twoeb, code, outlier, NULL, out);
}

/* With padding the halo cells of _s hold the pad values (x_block_pad /
 * x_edge_pad, src/utils/padding.hh:302-372; src/dualquant.hh:951-964,
 * :1003-1016) and the 1D block start predicts from the pad value (:918-933). */
int orc_reconstruct_halo(int ndim, const size_t* dims, int block, int radius, double twoeb,
                         const int32_t* code, const float* outlier, const float* halo, float* out) {
    size_t d0 = dims[0], d1 = ndim >= 2 ? dims[1] : 1, d2 = ndim >= 3 ? dims[2] : 1;
    size_t n = d0 * d1 * d2, B = (size_t)block;
    const size_t nb0 = orc_nblk(d0, block), nb1 = ndim >= 2 ? orc_nblk(d1, block) : 1;
    float* s = (float*)malloc(n * sizeof(float) + 4);
    if (!s) return ORC_ENOMEM;
    const size_t sy = d0, sz = d0 * d1;
    for (size_t z = 0; z < d2; z++)
        for (size_t y = 0; y < d1; y++)
            for (size_t i = 0; i < d0; i++) {
                size_t id = i + d0 * (y + d1 * z);
                size_t ix = i % B, iy = y % B, iz = z % B;
                const float* h = halo ? halo + 3 * ((i / B) + nb0 * ((y / B) + nb1 * (z / B))) : NULL;
                float pred;
#define NB(dz, dy, dx) halo_nb(ndim, h, dz, dy, dx, iz, iy, ix, s, id - (dx) - (dy) * sy - (dz) * sz)
                if (ndim == 1) {
                    pred = NB(0, 0, 1);
                } else if (ndim == 2) {
                    float W = NB(0, 0, 1);
                    float N = NB(0, 1, 0);
                    float NW = NB(0, 1, 1);
                    pred = N + W - NW;
                } else {
                    float X = NB(0, 0, 1);
                    float Y = NB(0, 1, 0);
                    float Z = NB(1, 0, 0);
                    float XY = NB(0, 1, 1);
                    float XZ = NB(1, 0, 1);
                    float YZ = NB(1, 1, 0);
                    float XYZ = NB(1, 1, 1);
                    pred = XYZ - XY - XZ - YZ + X + Y + Z;
                }
#undef NB
                int32_t c = code[id];
                s[id] = c == 0 ? outlier[id] : (float)(pred + (float)(c - radius));
                out[id] = (float)((double)s[id] * twoeb);
            }
    free(s);
    return ORC_OK;
}

/* Decompress (src/vecsz_interface.hh:232-375) without the lossless pass. */
int orc_decompress(const uint8_t* a, size_t size, int r2r, double range, int dict, int block,
                   float* out, size_t out_cap) {
    int nd;
    size_t dims[3];
    double eb;
    uint64_t pad_len, huff_size, nout;
    int rc = orc_archive_info(a, size, &nd, dims, &eb, &pad_len, &huff_size, &nout);
    if (rc) return rc;
    size_t n = dims[0] * dims[1] * dims[2];
    if (n > out_cap) return ORC_EINVAL;
    orc_ebs ebs;
    rc = orc_eb_family(r2r, eb, range, &ebs);
    if (rc) return rc;
    size_t meta = 4 * (size_t)(nd + 1) + 8 + 24;
    const uint8_t* huff = a + meta + 4 * pad_len;
    const uint8_t* outl = huff + huff_size;
    int32_t* code = (int32_t*)malloc(n * sizeof(int32_t));
    float* o = (float*)malloc(n * sizeof(float));
    if (!code || !o) return ORC_ENOMEM;
    rc = orc_huffman_decode(huff, (size_t)huff_size, n, code);
    if (rc) { free(code); free(o); return rc; }
    /* outlier scatter (src/vecsz_interface.hh:313-317) */
    size_t k = 0;
    for (size_t i = 0; i < n; i++) {
        if (code[i] == 0) {
            if (k >= nout) { free(code); free(o); return ORC_ECORRUPT; }
            memcpy(&o[i], outl + 4 * k, 4);
            k++;
        } else {
            o[i] = 0;
        }
    }
    size_t nblocks = orc_nblk(dims[0], block) * (nd >= 2 ? orc_nblk(dims[1], block) : 1) *
                     (nd >= 3 ? orc_nblk(dims[2], block) : 1);
    float* halo = (float*)malloc(3 * nblocks * sizeof(float));
    float* pv = (float*)malloc(4 * (size_t)pad_len + 4);
    if (!halo || !pv) { free(code); free(o); free(halo); free(pv); return ORC_ENOMEM; }
    memcpy(pv, a + meta, 4 * (size_t)pad_len);
    orc_pad_halo_from_archive(nd, dims, block, pv, pad_len, halo);
    rc = orc_reconstruct_halo(nd, dims, block, dict / 2, ebs.twoeb_d, code, o, halo, out);
    free(code);
    free(o);
    free(halo);
    free(pv);
    return rc;
}

void orc_free(void* p) { free(p); }
