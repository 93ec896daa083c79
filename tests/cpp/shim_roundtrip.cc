// Drives include/vecsz_interface.hh the way the reference's main does
// (src/vecsz.cc:37, 59): Compress<float,int>, then Decompress<float,int> on
// the payload, with an ArgParse carrying the fields the interface reads.
// Prints one line: "nodevice" when no GPU is usable, else
// "ok <data_size> <lossless_size> <max_abs_err> <eb> <ndim> <dim0> <dim2> <len> <nblk0> <stride2>"
// (ap2's fields after Decompress).  With a third argument <dir>, the field is
// read from <dir>/in.f32 and the payload (after the lossless pass, if any) and
// the decompressed field are written to <dir>/out.bin and <dir>/out.f32.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "vecsz_interface.hh"

struct Dim4 {
    int _0 = 1, _1 = 1, _2 = 1, _3 = 1;
};
struct Files {
    std::string input_file;
};
struct Szwf {
    bool lossless_pass = false, lossless_gzip = false;
    bool global_padding = false, block_padding = false, edge_padding = false;
};
struct ArgParse {
    int ndim = 3;
    size_t len = 1;  // set by InitializeDims in the reference (src/dimensions.hh:37)
    Dim4 dim4;
    Dim4 nblk4;
    Dim4 stride4;
    std::string mode = "abs";
    double eb = 1e-4;
    int block_size = 8;
    int dict_size = 4096;
    Szwf szwf;
    Files files;
    float pad_constant = 0;
    int pad_type = 3;
};

int main(int argc, char** argv) {
    const std::string lossless = argc > 1 ? argv[1] : "none";
    const std::string padding = argc > 2 ? argv[2] : "none";  // "none" or "<mode>,<type id>"
    ArgParse ap;
    ap.dim4 = {40, 30, 20, 1};
    ap.eb = 1e-3;
    ap.szwf.lossless_pass = lossless != "none";
    ap.szwf.lossless_gzip = lossless == "gzip";
    if (padding != "none") {
        ap.szwf.global_padding = padding.find("global") != std::string::npos;
        ap.szwf.block_padding = padding.find("block") != std::string::npos;
        ap.szwf.edge_padding = padding.find("edge") != std::string::npos;
        ap.pad_type = padding.back() - '0';
        ap.pad_constant = 0.75f;
    }
    const std::string dir = argc > 3 ? argv[3] : "";
    const size_t n = 40 * 30 * 20;
    float* x = new float[n];
    float* orig = new float[n];
    if (!dir.empty()) {
        FILE* f = std::fopen((dir + "/in.f32").c_str(), "rb");
        if (!f || std::fread(x, 4, n, f) != n) {
            std::printf("error reading input\n");
            return 1;
        }
        std::fclose(f);
        std::memcpy(orig, x, n * 4);
    } else {
        for (size_t i = 0; i < n; i++) orig[i] = x[i] = (float)(std::sin(0.01 * (double)i) * 3.0 + 0.001 * (double)(i % 7));
    }
    size_t nout = 0, size = 0;
    unsigned long lsize = 0;
    void* arch = nullptr;
    try {
        arch = vecsz::interface::Compress<float, int>(&ap, x, &nout, &size, &lsize);
    } catch (const std::runtime_error& e) {
        if (std::strstr(e.what(), "no usable GPU")) {
            std::printf("nodevice\n");
            return 0;
        }
        std::printf("error %s\n", e.what());
        return 1;
    }
    if (!lsize) x = nullptr;  // Compress deleted it, like the reference
    ArgParse ap2;
    ap2.mode = "abs";
    ap2.szwf = ap.szwf;
    ap2.szwf.global_padding = ap2.szwf.block_padding = ap2.szwf.edge_padding = false;  // taken from the archive
    float* y = (float*)vecsz::interface::Decompress<float, int>(&ap2, (unsigned char*)arch, size, lsize);
    if (!dir.empty()) {
        FILE* f = std::fopen((dir + "/out.bin").c_str(), "wb");
        std::fwrite(arch, 1, lsize ? lsize : size, f);
        std::fclose(f);
        f = std::fopen((dir + "/out.f32").c_str(), "wb");
        std::fwrite(y, 4, ap2.len, f);
        std::fclose(f);
    }
    double maxerr = 0;
    for (size_t i = 0; i < n; i++) maxerr = std::fmax(maxerr, std::fabs((double)y[i] - (double)orig[i]));
    const int pad_flags = (ap2.szwf.global_padding ? 1 : 0) | (ap2.szwf.block_padding ? 2 : 0) |
                          (ap2.szwf.edge_padding ? 4 : 0);
    std::printf("ok %zu %lu %.9g %.9g %d %d %d %zu %d %d %d\n", size, lsize, maxerr, ap.eb, ap2.ndim, ap2.dim4._0,
                ap2.dim4._2, ap2.len, ap2.nblk4._0, ap2.stride4._2, pad_flags);
    if (lsize) {
        std::free(arch);
        delete[] x;
    } else {
        delete[] (unsigned char*)arch;
    }
    delete[] y;
    delete[] orig;
    return 0;
}
