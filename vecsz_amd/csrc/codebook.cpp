// Huffman codebook of the reference's SZ-1.4-style coder, built on the host.
//
// The tree must be the reference's tree bit for bit, and that tree is defined
// by its priority queue, not by Huffman optimality:
//  - leaves enter in ascending symbol order (src/huffman.hh:248-254);
//  - qinsert sifts up and stops at a parent with freq <= new (:169-177);
//  - qremove walks the hole down to a leaf taking the strictly smaller child
//    and drops the last element there without sifting it up (:179-192);
//  - each merge removes b then a and inserts node(left=a, right=b) (:262; g++
//    evaluates the two qremove() arguments right to left);
//  - codes: left = 0, right = 1, MSB first (build_code, :200-227).
// The heap keeps (freq, node) pairs inline so the sift loops touch one array.
#include "codebook.hpp"

#include <string.h>

#include "common.hpp"
#include "../../include/vecsz_hip.h"

namespace vsz {

namespace {

struct HeapEntry {
    uint64_t f;
    int32_t n;
};

}  // namespace

uint64_t tree_byte_size(uint64_t nc) {
    uint64_t w = nc <= 256 ? 1 : (nc <= 65536 ? 2 : 4);
    return 1 + nc * (2 * w + 4 + 1);
}

int build_codebook(const uint64_t* freq, int dict, Codebook& cb) {
    cb.dict = dict;
    cb.cw_hi.assign(dict, 0);
    cb.cw_lo.assign(dict, 0);
    cb.len.assign(dict, 0);
    int nsym = 0;
    for (int s = 0; s < dict; s++) nsym += freq[s] != 0;
    cb.nsym = nsym;
    if (nsym == 0) return VECSZ_EINVAL;

    const int cap = 2 * nsym;
    std::vector<int32_t> left(cap, -1), right(cap, -1), sym(cap, -1);
    std::vector<HeapEntry> qq(nsym + 2);
    int n_nodes = 0;
    int qend = 1;

    auto qinsert = [&](HeapEntry e) {
        int i = qend++, j;
        while ((j = i >> 1)) {
            if (qq[j].f <= e.f) break;
            qq[i] = qq[j];
            i = j;
        }
        qq[i] = e;
    };
    auto qremove = [&]() -> HeapEntry {
        HeapEntry top = qq[1];
        qend--;
        int i = 1, l;
        while ((l = i << 1) < qend) {
            if (l + 1 < qend && qq[l + 1].f < qq[l].f) l++;
            qq[i] = qq[l];
            i = l;
        }
        qq[i] = qq[qend];
        return top;
    };

    for (int s = 0; s < dict; s++) {
        if (!freq[s]) continue;
        int n = n_nodes++;
        sym[n] = s;
        qinsert({freq[s], n});
    }
    while (qend > 2) {
        HeapEntry b = qremove();  // first removed: right child
        HeapEntry a = qremove();  // second removed: left child
        int n = n_nodes++;
        left[n] = a.n;
        right[n] = b.n;
        qinsert({a.f + b.f, n});
    }
    const int32_t root = qq[1].n;

    // Codewords (DFS, left first).
    struct Item {
        int32_t n;
        int len;
        uint64_t bits;
    };
    std::vector<Item> stk;
    stk.reserve(cap);
    stk.push_back({root, 0, 0});
    int max_len = 0;
    bool too_long = false;
    while (!stk.empty()) {
        Item it = stk.back();
        stk.pop_back();
        if (sym[it.n] >= 0) {
            int s = sym[it.n];
            if (it.len > 64) {
                too_long = true;
                continue;
            }
            cb.len[s] = (uint8_t)it.len;
            if (it.len > 32) {
                cb.cw_hi[s] = (uint32_t)(it.bits >> 32);
                cb.cw_lo[s] = (uint32_t)it.bits;
            } else {
                cb.cw_lo[s] = (uint32_t)it.bits;
            }
            if (it.len > max_len) max_len = it.len;
            continue;
        }
        stk.push_back({right[it.n], it.len + 1, (it.bits << 1) | 1});
        stk.push_back({left[it.n], it.len + 1, it.bits << 1});
    }
    if (too_long) return VECSZ_ECODELEN;
    cb.max_len = max_len;

    // Serialization: preorder numbering, root 0, each child takes ++n_inode
    // when it is visited (pad_tree_*, src/huffman.hh:374-423).
    const int nc = 2 * nsym - 1;
    cb.node_count = nc;
    const int w = nc <= 256 ? 1 : (nc <= 65536 ? 2 : 4);
    std::vector<uint32_t> L(nc, 0), R(nc, 0), C(nc, 0);
    std::vector<uint8_t> T(nc, 0);
    struct Visit {
        int32_t n;
        int32_t parent;
        int side;
    };
    std::vector<Visit> vs;
    vs.reserve(cap);
    vs.push_back({root, -1, 0});
    uint32_t n_inode = 0;
    bool first = true;
    while (!vs.empty()) {
        Visit v = vs.back();
        vs.pop_back();
        uint32_t idx = first ? 0 : ++n_inode;
        first = false;
        if (v.parent >= 0) (v.side ? R : L)[v.parent] = idx;
        bool leaf = sym[v.n] >= 0;
        C[idx] = leaf ? (uint32_t)sym[v.n] : 0;
        T[idx] = leaf ? 1 : 0;
        if (!leaf) {
            vs.push_back({right[v.n], (int32_t)idx, 1});
            vs.push_back({left[v.n], (int32_t)idx, 0});
        }
    }
    cb.tree.assign(tree_byte_size(nc), 0);
    uint8_t* o = cb.tree.data();
    *o++ = 0;  // CPU_sysEndianType (src/huffman.hh:72)
    for (int pass = 0; pass < 2; pass++) {
        const std::vector<uint32_t>& A = pass == 0 ? L : R;
        for (int i = 0; i < nc; i++)
            for (int b = 0; b < w; b++) *o++ = (uint8_t)(A[i] >> (8 * b));
    }
    for (int i = 0; i < nc; i++)
        for (int b = 0; b < 4; b++) *o++ = (uint8_t)(C[i] >> (8 * b));
    for (int i = 0; i < nc; i++) *o++ = T[i];
    return VECSZ_OK;
}

int build_decode_tables(const uint8_t* huff, uint64_t huff_size, DecodeTables& dt) {
    if (huff_size < 8) return VECSZ_ECORRUPT;
    int32_t ncs = (int32_t)(((uint32_t)huff[0] << 24) | ((uint32_t)huff[1] << 16) |
                            ((uint32_t)huff[2] << 8) | huff[3]);
    if (ncs <= 0) return VECSZ_ECORRUPT;
    uint64_t nc = (uint64_t)ncs;
    uint64_t tb = tree_byte_size(nc);
    if (8 + tb > huff_size) return VECSZ_ECORRUPT;
    dt.node_count = (int)nc;
    dt.tree_bytes = tb;
    const int w = nc <= 256 ? 1 : (nc <= 65536 ? 2 : 4);
    const uint8_t* base = huff + 9;
    auto rd = [&](const uint8_t* p, uint64_t i) -> uint32_t {
        uint32_t v = 0;
        for (int b = 0; b < w; b++) v |= (uint32_t)p[w * i + b] << (8 * b);
        return v;
    };
    const uint8_t* Lp = base;
    const uint8_t* Rp = base + nc * w;
    const uint8_t* Cp = base + 2 * nc * w;
    const uint8_t* Tp = Cp + 4 * nc;
    auto rc = [&](uint64_t i) -> uint32_t {
        return (uint32_t)Cp[4 * i] | ((uint32_t)Cp[4 * i + 1] << 8) | ((uint32_t)Cp[4 * i + 2] << 16) |
               ((uint32_t)Cp[4 * i + 3] << 24);
    };
    // Root: (C[0], t[0]) for <= 256 nodes, forced internal otherwise
    // (src/huffman.hh:577, :602, :627).
    dt.root_leaf = nc <= 256 ? Tp[0] != 0 : false;
    dt.root_sym = nc <= 256 ? rc(0) : 0;
    if (dt.root_leaf) {
        if (dt.root_sym > 0xFFFF) return VECSZ_ECORRUPT;
        dt.lut.assign((size_t)1 << kLutBits, 0);
        dt.child.assign(2, kNodeAbsent);
        return VECSZ_OK;
    }
    dt.child.assign(2 * nc, kNodeAbsent);
    for (uint64_t i = 0; i < nc; i++) {
        bool internal = (i == 0) ? true : Tp[i] == 0;
        if (!internal) continue;
        uint32_t ch[2] = {rd(Lp, i), rd(Rp, i)};
        for (int b = 0; b < 2; b++) {
            uint32_t c = ch[b];
            if (c == 0 || c >= nc) continue;  // absent (src/huffman.hh:509, :515)
            if (Tp[c]) {
                uint32_t s = rc(c);
                if (s > 0xFFFF) return VECSZ_ECORRUPT;
                dt.child[2 * i + b] = kNodeLeaf | s;
            } else {
                dt.child[2 * i + b] = c;
            }
        }
    }
    const uint32_t K = kLutBits;
    dt.lut.assign((size_t)1 << K, 0);
    for (uint32_t p = 0; p < (1u << K); p++) {
        uint32_t node = 0, entry = 0;
        for (uint32_t b = 0; b < K; b++) {
            uint32_t bit = (p >> (K - 1 - b)) & 1;
            uint32_t nx = dt.child[2 * node + bit];
            if (nx == kNodeAbsent) {
                entry = 0;
                break;
            }
            if (nx & kNodeLeaf) {
                entry = ((b + 1) << 16) | (nx & 0xFFFF);
                break;
            }
            node = nx;
            if (b == K - 1) entry = kLutSlow | node;
        }
        dt.lut[p] = entry;
    }
    return VECSZ_OK;
}

}  // namespace vsz
