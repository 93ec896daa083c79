// Host-side Huffman codebook construction and decode-table preparation.
#pragma once

#include <stdint.h>
#include <vector>

namespace vsz {

struct Codebook {
    int dict = 0;        // symbols live in [0, dict)
    int nsym = 0;        // symbols with nonzero frequency
    int node_count = 0;  // 2 * nsym - 1 (src/huffman.hh:643-645)
    int max_len = 0;
    std::vector<uint32_t> cw_hi, cw_lo;  // codeword bits, right-aligned, split at 32
    std::vector<uint8_t> len;            // codeword length (0 for a one-symbol tree)
    std::vector<uint8_t> tree;           // serialized tree (src/huffman.hh:425-502)
};

// Work arrays of a codebook build.  build_codebook keeps one set per host
// thread; a caller may pass its own, reserved beforehand with
// codebook_reserve, so that the build allocates nothing.
struct CodebookWork {
    struct Pair {
        uint64_t f;
        int32_t n;
    };
    std::vector<int32_t> sym, kid, len, size, idx;
    std::vector<uint64_t> bits;
    std::vector<uint64_t> heap_packed;
    std::vector<Pair> heap_pair;
};

// Capacity for any histogram of at most `dict` bins in cb and w.
void codebook_reserve(Codebook& cb, CodebookWork& w, int dict);

// Builds the reference's tree from a symbol histogram of `dict` bins.
// Returns a VECSZ_* status.  With cb and *work reserved for dict bins
// (codebook_reserve), it allocates nothing.
int build_codebook(const uint64_t* freq, int dict, Codebook& cb, CodebookWork* work = nullptr);

struct DecodeTables {
    int node_count = 0;
    uint64_t tree_bytes = 0;
    bool root_leaf = false;
    uint32_t root_sym = 0;
    std::vector<uint32_t> lut;    // 1 << kLutBits entries
    std::vector<uint32_t> sub;    // deeper tables, 1 << kSubTabBits entries each
    std::vector<uint32_t> child;  // 2 * node_count entries (host-side walks)
    // compact form of lut/sub (empty when a symbol or table number needs
    // more than 12 bits or the tree has an invalid path)
    std::vector<uint16_t> lut16, sub16;
};

// Parses the Huffman section of an archive (header + serialized tree) into
// decoder tables.  `huff` points at the section, `huff_size` is its length.
int build_decode_tables(const uint8_t* huff, uint64_t huff_size, DecodeTables& dt);

// Byte size of a serialized tree with `node_count` nodes.
uint64_t tree_byte_size(uint64_t node_count);

}  // namespace vsz
