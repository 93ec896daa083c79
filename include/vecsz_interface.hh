// vecsz_interface.hh -- drop-in replacement of the reference's header-only
// interface (src/vecsz_interface.hh) that runs the path on an MI355X through
// the C ABI of libvecsz_hip.so (include/vecsz_hip.h).
//
// Same names, template parameters and argument meaning as the reference:
//   void* vecsz::interface::Compress<T,Q>(argparse* ap, T* data_in, size_t* num_outlier,
//                                         size_t* data_size, unsigned long* lossless_size)
//   void* vecsz::interface::Decompress<T,Q>(argparse* ap, unsigned char* tmp_data,
//                                           size_t target_size, unsigned long lossless_size)
// The argparse type is a template parameter so this header does not depend on
// the reference's own headers; the reference's ArgParse binds to it unchanged.
//
// Contract kept from the reference (src/vecsz_interface.hh):
//  - Compress returns a `new unsigned char[*data_size]` archive payload
//    (datapack::pack layout), increments *num_outlier by the outlier count,
//    and delete[]s data_in (src/vecsz_interface.hh:222).
//  - Decompress sets ap->ndim, ap->dim4 and ap->eb from the archive
//    (:262-268), then ap->len, ap->nblk4 and ap->stride4 as InitializeDims
//    does (:270, src/dimensions.hh:31-45) and the padding flags from the
//    archive's pad count (:287-298), and returns `new T[len]`.
//  - r2r: Compress takes the value range from data_in itself (the reference
//    re-reads ap->files.input_file, which holds the same floats, :49-50);
//    Decompress re-reads ap->files.input_file like the reference (:280-284).
//  - with ap->szwf.lossless_pass, Compress runs gzip (szwf.lossless_gzip) or
//    zstd at level 3 over the payload and returns that malloc'ed stream with
//    *lossless_size set (:206-217); Decompress inflates a lossless_size > 0
//    stream to the target_size payload first (:237-244).
//  - padding (ap->szwf.{global,block,edge}_padding, ap->pad_type,
//    ap->pad_constant) is applied when the argparse type has those members;
//    Decompress takes it from the archive, as the reference does (:287-298).
// Differences: errors throw std::runtime_error instead of exit(1); only
// T = float is supported (the only type the reference instantiates,
// src/vecsz.cc:37); the zstd pass uses the system zstd runtime (its frames
// decode both ways, its bytes differ from the reference's zstd 1.3.5).
#ifndef VECSZ_AMD_INTERFACE_HH
#define VECSZ_AMD_INTERFACE_HH

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "vecsz_hip.h"

// The reference's header also brings its analysis helpers into every file
// that includes it (src/vecsz_interface.hh:13); its main calls
// analysis::histogram (src/vecsz.cc:112-113).  Keep that when the
// reference's source tree is on the include path, so main compiles unchanged.
#if __has_include("utils/analysis.hh")
#include "utils/analysis.hh"
#endif

namespace vecsz {
namespace interface {

namespace detail {

inline vecsz_ctx* context() {
    static vecsz_ctx* ctx = nullptr;
    if (!ctx) {
        int rc = vecsz_ctx_create(&ctx, 0, nullptr);
        if (rc != VECSZ_OK) throw std::runtime_error(std::string("vecsz: ") + vecsz_strerror(rc));
    }
    return ctx;
}

inline void check(int rc) {
    if (rc != VECSZ_OK)
        throw std::runtime_error(std::string("vecsz: ") + vecsz_strerror(rc) + ": " +
                                 vecsz_ctx_last_error(context()));
}

// GetDatumValueRange<float> (src/types.cc:27-35) on memory.
inline double value_range(const float* d, size_t n) {
    float mx = d[0], mn = d[0];
    for (size_t i = 1; i < n; i++) {
        if (mx < d[i]) mx = d[i];
        if (d[i] < mn) mn = d[i];
    }
    float r = mx - mn;
    return (double)r;
}

inline double value_range_of_file(const std::string& path, size_t n) {
    std::vector<float> buf(n);
    std::ifstream ifs(path.c_str(), std::ios::binary | std::ios::in);
    if (!ifs.is_open()) throw std::runtime_error("vecsz: cannot open " + path);
    ifs.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(n * sizeof(float)));
    return value_range(buf.data(), n);
}

// The reference's ArgParse carries the padding options (szwf.*_padding,
// pad_type, pad_constant; src/argument_parser/argparse.hh:34-36, :75-77);
// an argparse type without them compresses without padding.
template <typename AP, typename = void>
struct has_padding : std::false_type {};
template <typename AP>
struct has_padding<AP, std::void_t<decltype(std::declval<const AP&>().szwf.global_padding),
                                   decltype(std::declval<const AP&>().szwf.block_padding),
                                   decltype(std::declval<const AP&>().szwf.edge_padding),
                                   decltype(std::declval<const AP&>().pad_type),
                                   decltype(std::declval<const AP&>().pad_constant)>> : std::true_type {};

// InitializeDims (src/dimensions.hh:11-64) also fills ap->len, ap->nblk4 and
// ap->stride4; the reference's main reads ap->len after Decompress
// (src/vecsz.cc:62).  Set where the argparse type has them.
template <typename AP, typename = void>
struct has_len : std::false_type {};
template <typename AP>
struct has_len<AP, std::void_t<decltype(std::declval<AP&>().len = size_t(1))>> : std::true_type {};
template <typename AP, typename = void>
struct has_nblk4 : std::false_type {};
template <typename AP>
struct has_nblk4<AP, std::void_t<decltype(std::declval<AP&>().nblk4._0), decltype(std::declval<AP&>().nblk4._3)>>
    : std::true_type {};
template <typename AP, typename = void>
struct has_stride4 : std::false_type {};
template <typename AP>
struct has_stride4<AP, std::void_t<decltype(std::declval<AP&>().stride4._0), decltype(std::declval<AP&>().stride4._3)>>
    : std::true_type {};

template <typename AP>
void initialize_dims(AP* ap) {
    const int d0 = ap->dim4._0, d1 = ap->dim4._1, d2 = ap->dim4._2, d3 = ap->dim4._3;
    if constexpr (has_nblk4<AP>::value) {
        // the reference's block count, (d + B + 1) / B (src/dimensions.hh:31)
        const int B = ap->block_size;
        ap->nblk4._0 = (d0 + B + 1) / B;
        ap->nblk4._1 = (d1 + B + 1) / B;
        ap->nblk4._2 = (d2 + B + 1) / B;
        ap->nblk4._3 = (d3 + B + 1) / B;
    }
    if constexpr (has_len<AP>::value) ap->len = (size_t)d0 * (size_t)d1 * (size_t)d2 * (size_t)d3;
    if constexpr (has_stride4<AP>::value) {
        ap->stride4._0 = 1;
        ap->stride4._1 = d0;
        ap->stride4._2 = d0 * d1;
        ap->stride4._3 = d0 * d1 * d2;
    }
}

template <typename AP>
vecsz_params params_of(const AP* ap) {
    vecsz_params p;
    vecsz_params_default(&p);
    p.ndim = ap->ndim;
    p.dims[0] = (uint64_t)ap->dim4._0;
    p.dims[1] = (uint64_t)ap->dim4._1;
    p.dims[2] = (uint64_t)ap->dim4._2;
    p.mode = ap->mode == "r2r" ? VECSZ_MODE_R2R : VECSZ_MODE_ABS;
    p.eb = ap->eb;
    p.block = ap->block_size;
    p.dict_size = ap->dict_size;
    if constexpr (has_padding<AP>::value) {
        // one mode at a time, global over block over edge (src/argument_parser/argparse.cc:156-169)
        p.pad_mode = ap->szwf.global_padding ? VECSZ_PAD_GLOBAL
                     : ap->szwf.block_padding ? VECSZ_PAD_BLOCK
                     : ap->szwf.edge_padding  ? VECSZ_PAD_EDGE
                                              : VECSZ_PAD_NONE;
        p.pad_type = ap->pad_type;
        p.pad_constant = ap->pad_constant;
    }
    return p;
}

}  // namespace detail

template <typename T, typename Q, typename AP>
void* Compress(AP* ap, T* data_in, size_t* num_outlier, size_t* data_size, unsigned long* lossless_size) {
    static_assert(std::is_same<T, float>::value, "vecSZ's path is float32");
    vecsz_params p = detail::params_of(ap);
    size_t len = 1;
    for (int i = 0; i < p.ndim; i++) len *= (size_t)p.dims[i];
    if (p.mode == VECSZ_MODE_R2R) p.value_range = detail::value_range(data_in, len);
    uint64_t bound = 0;
    detail::check(vecsz_compress_bound(&p, &bound));
    std::vector<uint8_t> tmp(bound);
    uint64_t size = 0, nout = 0;
    detail::check(vecsz_compress_host(detail::context(), &p, data_in, tmp.data(), bound, &size, &nout));
    *num_outlier += (size_t)nout;
    *data_size = (size_t)size;
    if (lossless_size) *lossless_size = 0;
    if (ap->szwf.lossless_pass) {
        const int kind = ap->szwf.lossless_gzip ? VECSZ_LOSSLESS_GZIP : VECSZ_LOSSLESS_ZSTD;
        uint64_t lbound = 0, lsize = 0;
        detail::check(vecsz_lossless_bound(kind, size, &lbound));
        unsigned char* lout = static_cast<unsigned char*>(std::malloc(lbound));
        if (!lout) throw std::bad_alloc();
        const int rc = vecsz_lossless_compress(kind, tmp.data(), size, lout, lbound, &lsize);
        if (rc != VECSZ_OK) {
            std::free(lout);
            detail::check(rc);
        }
        if (lossless_size) *lossless_size = (unsigned long)lsize;
        return lout;  // the reference keeps data_in in this branch (:217)
    }
    unsigned char* out = new unsigned char[size];
    std::memcpy(out, tmp.data(), size);
    delete[] data_in;
    return out;
}

template <typename T, typename Q, typename AP>
void* Decompress(AP* ap, unsigned char* tmp_data, size_t target_size, unsigned long lossless_size) {
    static_assert(std::is_same<T, float>::value, "vecSZ's path is float32");
    std::vector<uint8_t> inflated;
    if (lossless_size) {
        inflated.resize(target_size);
        uint64_t got = 0;
        detail::check(vecsz_lossless_decompress(tmp_data, lossless_size, inflated.data(), target_size, &got));
        tmp_data = inflated.data();
    }
    int ndim = 0;
    uint64_t dims[3];
    double eb = 0;
    uint64_t huff = 0, nout = 0;
    detail::check(vecsz_archive_info(tmp_data, target_size, &ndim, dims, &eb, &huff, &nout));
    ap->ndim = ndim;
    ap->dim4._0 = (int)dims[0];
    ap->dim4._1 = ndim >= 2 ? (int)dims[1] : 1;
    ap->dim4._2 = ndim >= 3 ? (int)dims[2] : 1;
    ap->dim4._3 = 1;
    ap->eb = eb;
    detail::initialize_dims(ap);
    if constexpr (detail::has_padding<AP>::value) {
        // padding mode from the pad value count, as the reference decides it
        // (src/vecsz_interface.hh:287-298): 1 global, one per block, ndim per block
        uint64_t pad_len = 0;
        std::memcpy(&pad_len, tmp_data + 4 * (ndim + 1) + 8, 8);
        uint64_t nblk = 1;
        for (int i = 0; i < ndim; i++) nblk *= (dims[i] + (uint64_t)ap->block_size + 1) / (uint64_t)ap->block_size;
        ap->szwf.global_padding = pad_len == 1;
        ap->szwf.block_padding = pad_len != 1 && pad_len == nblk;
        ap->szwf.edge_padding = pad_len != 1 && pad_len != nblk && pad_len == (uint64_t)ndim * nblk;
    }
    vecsz_params p = detail::params_of(ap);
    const size_t len = (size_t)dims[0] * (size_t)dims[1] * (size_t)dims[2];
    if (p.mode == VECSZ_MODE_R2R) p.value_range = detail::value_range_of_file(ap->files.input_file, len);
    T* xdata = new T[len];
    int rc = vecsz_decompress_host(detail::context(), &p, tmp_data, target_size, xdata, len);
    if (rc != VECSZ_OK) {
        delete[] xdata;
        detail::check(rc);
    }
    return xdata;
}

}  // namespace interface
}  // namespace vecsz

#endif
