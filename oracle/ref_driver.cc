// TEST INFRASTRUCTURE ONLY -- never shipped, never on the product path.
//
// In-process driver around the reference's own Compress/Decompress
// (src/vecsz_interface.hh:31-229, :232-375), compiled from the reference
// sources where they lie by oracle/Makefile.ref into oracle/_ref/.  Tests and
// tests/golden/make_golden.py use it to pin the C restatement (oracle/) to the
// real reference; bench.py may time it as the CPU baseline.
//
// The driver fills an ArgParse the way src/argument_parser/argparse.cc:259-546
// would for `-l x,y,z -m abs|r2r -e eb -d dict -b block`, keeps the
// reference's stdout chatter away from the caller, and exposes plain C entry
// points.  r2r mode in the reference re-reads the input file
// (src/vecsz_interface.hh:49-50, :280-284), so the caller passes the path of a
// file that holds the same floats.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "vecsz_interface.hh"
#include "dimensions.hh"

namespace {

struct Mute {
    int saved = -1;
    explicit Mute(bool on) {
        if (!on) return;
        fflush(stdout);
        std::cout.flush();
        saved = dup(1);
        int devnull = open("/dev/null", O_WRONLY);
        dup2(devnull, 1);
        close(devnull);
    }
    ~Mute() {
        if (saved < 0) return;
        fflush(stdout);
        std::cout.flush();
        dup2(saved, 1);
        close(saved);
    }
};

void fill_args(ArgParse* ap, int ndim, const size_t* dims, int r2r, double eb, int dict, int block,
               const char* path) {
    ap->ndim = ndim;
    ap->dim4 = {1, 1, 1, 1};
    if (ndim >= 1) ap->dim4._0 = (int)dims[0];
    if (ndim >= 2) ap->dim4._1 = (int)dims[1];
    if (ndim >= 3) ap->dim4._2 = (int)dims[2];
    ap->mode = r2r ? "r2r" : "abs";
    ap->eb = eb;
    ap->dict_size = dict;
    ap->radius = dict / 2;
    ap->block_size = block;
    ap->files.input_file = path ? path : "";
    ap->dtype = "f32";
    ap->verbose = false;
}

}  // namespace

extern "C" {

int ref_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// Runs vecsz::interface::Compress<float,int> (src/vecsz_interface.hh:31).
// Returns the archive payload (the bytes datapack::pack produced) in *out;
// free it with ref_free.
int ref_compress(int ndim, const size_t* dims, int r2r, double eb, int dict, int block,
                 const char* path, const float* data, uint8_t** out, size_t* out_size,
                 size_t* n_outlier, int quiet) {
    ArgParse* ap = new ArgParse();
    fill_args(ap, ndim, dims, r2r, eb, dict, block, path);
    size_t len = 1;
    for (int i = 0; i < ndim; i++) len *= dims[i];
    // Compress deletes its input (src/vecsz_interface.hh:222) and mutates it in
    // 1D (src/dualquant.hh:92): hand it a private copy.  It also writes up to
    // 7 floats past the end in 1D prequant (src/dualquant.hh:83-93), so pad.
    float* copy = new float[len + 64]();
    memcpy(copy, data, len * sizeof(float));
    size_t data_size = 0, nnz = 0;
    unsigned long lossless_size = 0;
    void* res;
    {
        Mute m(quiet != 0);
        res = vecsz::interface::Compress<float, int>(ap, copy, &nnz, &data_size, &lossless_size);
    }
    *out = (uint8_t*)res;
    *out_size = data_size;
    *n_outlier = nnz;
    delete ap;
    return 0;
}

// Compress with padding: pad_mode 1 global, 2 block, 3 edge (the flags
// --padding sets, src/argument_parser/argparse.cc:399-415), pad_type 0 max,
// 1 min, 2 avg, 3 constant.  The reference appends pad values through a
// counter shared by its OpenMP threads (src/dualquant.hh:41, :232, :243), so
// callers run it with one thread (ref_set_threads(1)) for the serial order.
int ref_compress_pad(int ndim, const size_t* dims, int r2r, double eb, int dict, int block, int pad_mode,
                     int pad_type, float pad_constant, const char* path, const float* data, uint8_t** out,
                     size_t* out_size, size_t* n_outlier, int quiet) {
    ArgParse* ap = new ArgParse();
    fill_args(ap, ndim, dims, r2r, eb, dict, block, path);
    ap->szwf.global_padding = pad_mode == 1;
    ap->szwf.block_padding = pad_mode == 2;
    ap->szwf.edge_padding = pad_mode == 3;
    ap->pad_type = pad_type;
    ap->pad_constant = pad_constant;
    size_t len = 1;
    for (int i = 0; i < ndim; i++) len *= dims[i];
    float* copy = new float[len + 64]();
    memcpy(copy, data, len * sizeof(float));
    size_t data_size = 0, nnz = 0;
    unsigned long lossless_size = 0;
    void* res;
    {
        Mute m(quiet != 0);
        res = vecsz::interface::Compress<float, int>(ap, copy, &nnz, &data_size, &lossless_size);
    }
    *out = (uint8_t*)res;
    *out_size = data_size;
    *n_outlier = nnz;
    delete ap;
    return 0;
}

// Runs vecsz::interface::Decompress<float,int> (src/vecsz_interface.hh:232).
int ref_decompress(int r2r, int dict, int block, const char* path, const uint8_t* archive,
                   size_t archive_size, float** out, size_t* out_len, int quiet) {
    ArgParse* ap = new ArgParse();
    size_t dims0[3] = {1, 1, 1};
    fill_args(ap, 1, dims0, r2r, 1e-4, dict, block, path);
    uint8_t* buf = new uint8_t[archive_size + 64]();
    memcpy(buf, archive, archive_size);
    void* res;
    {
        Mute m(quiet != 0);
        res = vecsz::interface::Decompress<float, int>(ap, buf, archive_size, 0);
    }
    *out = (float*)res;
    *out_len = (size_t)ap->dim4._0 * ap->dim4._1 * ap->dim4._2 * ap->dim4._3;
    delete[] buf;
    delete ap;
    return 0;
}

// Same sequence as Decompress (src/vecsz_interface.hh:259-357) built from the
// reference's own pieces, except that the no-padding pad value is the +0.0 the
// compressor used (src/dualquant.hh:36) instead of the unallocated
// pad_vals[0] Decompress reads in 1D (src/dualquant.hh:925-927).
int ref_decompress_pad0(int r2r, int dict, int block, const char* path, const uint8_t* archive,
                        float** out, size_t* out_len, int quiet) {
    Mute m(quiet != 0);
    ArgParse* ap = new ArgParse();
    size_t dims0[3] = {1, 1, 1};
    fill_args(ap, 1, dims0, r2r, 1e-4, dict, block, path);
    int ndims;
    double eb;
    int* dims;
    size_t pad_length;
    uint8_t* huffman;
    float* outliers;
    float* pad_vals;
    datapack::unpack<uint8_t, float>(const_cast<uint8_t*>(archive), &ndims, &dims, &eb, &pad_vals,
                                     &pad_length, &huffman, &outliers);
    ap->ndim = ndims;
    ap->dim4 = {1, 1, 1, 1};
    if (ndims >= 1) ap->dim4._0 = dims[0];
    if (ndims >= 2) ap->dim4._1 = dims[1];
    if (ndims >= 3) ap->dim4._2 = dims[2];
    ap->eb = eb;
    size_t const* const dims_L16 = InitializeDims(ap);
    auto eb_config = new config_t(ap->dict_size, ap->eb);
    size_t len = dims_L16[LEN];
    if (r2r) {
        double value_range = GetDatumValueRange<float>(ap->files.input_file, len);
        eb_config->ChangeToRelativeMode(value_range);
    }
    double const* const ebs_L4 = InitializeErrorBoundFamily(eb_config);
    float* xdata = new float[len];
    float* outlier = new float[len];
    int* code = new int[len];
    float pad0[1] = {0.0f};
    size_t pad_idx = 0;
    auto tree = DesignVerification::createHuffmanTree(ap->dict_size * 2);
    DesignVerification::decode_withTree(tree, huffman, len, code);
    size_t oi = 0;
    for (size_t i = 0; i < len; i++) outlier[i] = (code[i] == 0) ? outliers[oi++] : 0;
    namespace pq = vecsz::predictor_quantizer;
    size_t blksz = ap->block_size;
    if (ndims == 1) {
        for (size_t b0 = 0; b0 < dims_L16[nBLK0]; b0++)
            pq::x_lorenzo_1d1l<float, int>(xdata, outlier, code, dims_L16, ebs_L4[EBx2], b0, blksz,
                                           ap->szwf, &pad_idx, pad0);
    } else if (ndims == 2) {
        for (size_t b1 = 0; b1 < dims_L16[nBLK1]; b1++)
            for (size_t b0 = 0; b0 < dims_L16[nBLK0]; b0++)
                pq::x_lorenzo_2d1l<float, int>(xdata, outlier, code, dims_L16, ebs_L4[EBx2], b0, b1,
                                               blksz, ap->szwf, &pad_idx, pad0);
    } else {
        for (size_t b2 = 0; b2 < dims_L16[nBLK2]; b2++)
            for (size_t b1 = 0; b1 < dims_L16[nBLK1]; b1++)
                for (size_t b0 = 0; b0 < dims_L16[nBLK0]; b0++)
                    pq::x_lorenzo_3d1l<float, int>(xdata, outlier, code, dims_L16, ebs_L4[EBx2], b0,
                                                   b1, b2, blksz, ap->szwf, &pad_idx, pad0);
    }
    *out = xdata;
    *out_len = len;
    delete[] outlier;
    delete[] code;
    delete[] dims;
    delete[] huffman;
    delete[] outliers;
    delete ap;
    return 0;
}

void ref_free(void* p) { delete[] (uint8_t*)p; }
void ref_free_f32(float* p) { delete[] p; }

// The reference's lossless pass (src/lossless.hh:179-198) at its level 3;
// the result is malloc'ed (free with ref_free_malloc).
unsigned long ref_lossless_compress(int zstd, const uint8_t* in, unsigned long n, uint8_t** out) {
    return vecsz::lossless::sz_lossless_compress(zstd ? ZSTD_COMPRESSOR : GZIP_COMPRESSOR, 3,
                                                 const_cast<unsigned char*>(in), n, out);
}

// Detection + inverse (src/lossless.hh:60-75, :200-218); -1 when undetected.
long ref_lossless_decompress(const uint8_t* in, unsigned long n, uint8_t** out, unsigned long target) {
    const int kind = vecsz::lossless::is_lossless_compressed_data(const_cast<unsigned char*>(in), n);
    if (kind < 0) return -1;
    return (long)vecsz::lossless::sz_lossless_decompress(kind, const_cast<unsigned char*>(in), n, out, target);
}

void ref_free_malloc(void* p) { free(p); }

}  // extern "C"
