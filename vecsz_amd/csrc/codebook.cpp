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
// The queue keeps its entries inline (one word when the sums allow) so the
// sift loops touch one array.
#include "codebook.hpp"

#include <string.h>

#include "common.hpp"
#include "../../include/vecsz_hip.h"

namespace vsz {

namespace {

// Queue entries.  PackedQueue keeps (freq << kNodeBits | node) in one word,
// usable when every sum of frequencies stays below 2^(64 - kNodeBits); it
// compares the frequency bits only, as the reference compares freq alone.
// PairQueue keeps (freq, node) pairs for larger totals.
constexpr int kNodeBits = 17;  // nodes < 2 * 65536
struct PackedQueue {
    using E = uint64_t;
    static E make(uint64_t f, int32_t n) { return (f << kNodeBits) | (uint64_t)n; }
    static uint64_t f(E e) { return e >> kNodeBits; }
    static int32_t node(E e) { return (int32_t)(e & ((1u << kNodeBits) - 1)); }
};
struct PairQueue {
    using E = CodebookWork::Pair;
    static E make(uint64_t f, int32_t n) { return E{f, n}; }
    static uint64_t f(const E& e) { return e.f; }
    static int32_t node(const E& e) { return e.n; }
};

// Per-thread work arrays, kept between calls so a codebook build allocates
// nothing once the first field of a given symbol count has been seen.
thread_local CodebookWork t_work_default;

template <typename Q>
std::vector<typename Q::E>& heap_of(CodebookWork& w);
template <>
std::vector<PackedQueue::E>& heap_of<PackedQueue>(CodebookWork& w) { return w.heap_packed; }
template <>
std::vector<PairQueue::E>& heap_of<PairQueue>(CodebookWork& w) { return w.heap_pair; }

// The reference's queue (src/huffman.hh:146-192) and merge loop (:244-262):
// leaves enter in ascending symbol order; each merge removes b, then a, and
// inserts node(left = a, right = b) with freq a + b.
template <typename Q>
void merge_all(const uint64_t* freq, int dict, int nsym, int32_t* sym, int32_t* kid, CodebookWork& w) {
    using E = typename Q::E;
    std::vector<E>& qq = heap_of<Q>(w);
    qq.resize(nsym + 2);
    int qend = 1;
    auto qinsert = [&](E e) {
        int i = qend++, j;
        const uint64_t fe = Q::f(e);
        while ((j = i >> 1)) {
            if (Q::f(qq[j]) <= fe) break;
            qq[i] = qq[j];
            i = j;
        }
        qq[i] = e;
    };
    // The hole walk is the hot loop.  Its child choices are data dependent
    // and unpredictable, so they are index arithmetic, not branches, and two
    // levels are resolved per step: the grandchildren's comparison is made
    // for both children before the children's comparison is known.
    // qq[l + 1] with l + 1 == qend is still inside the array.
    auto qremove = [&]() -> E {
        E top = qq[1];
        const int end = --qend;
        int i = 1, l;
        while (4 * i + 3 < end) {
            const int c = 2 * i + (int)(Q::f(qq[2 * i + 1]) < Q::f(qq[2 * i]));
            const int g0 = 4 * i + (int)(Q::f(qq[4 * i + 1]) < Q::f(qq[4 * i]));
            const int g1 = 4 * i + 2 + (int)(Q::f(qq[4 * i + 3]) < Q::f(qq[4 * i + 2]));
            const int g = g0 + ((g1 - g0) & -(c & 1));
            qq[i] = qq[c];
            qq[c] = qq[g];
            i = g;
        }
        while ((l = i << 1) < end) {
            l += (int)((l + 1 < end) & (Q::f(qq[l + 1]) < Q::f(qq[l])));
            qq[i] = qq[l];
            i = l;
        }
        qq[i] = qq[end];
        return top;
    };
    for (int s = 0, n = 0; s < dict; s++) {
        if (!freq[s]) continue;
        sym[n] = s;
        qinsert(Q::make(freq[s], n));
        n++;
    }
    for (int n = nsym; qend > 2; n++) {
        E b = qremove();  // first removed: right child
        E a = qremove();  // second removed: left child
        kid[2 * (n - nsym)] = Q::node(a);
        kid[2 * (n - nsym) + 1] = Q::node(b);
        qinsert(Q::make(Q::f(a) + Q::f(b), n));
    }
}

}  // namespace

uint64_t tree_byte_size(uint64_t nc) {
    uint64_t w = nc <= 256 ? 1 : (nc <= 65536 ? 2 : 4);
    return 1 + nc * (2 * w + 4 + 1);
}

void codebook_reserve(Codebook& cb, CodebookWork& w, int dict) {
    const size_t nc = 2 * (size_t)dict;
    cb.cw_hi.reserve(dict);
    cb.cw_lo.reserve(dict);
    cb.len.reserve(dict);
    cb.tree.reserve(tree_byte_size(nc));
    w.sym.reserve(dict);
    w.kid.reserve(nc);
    w.len.reserve(nc);
    w.size.reserve(nc);
    w.idx.reserve(nc);
    w.bits.reserve(nc);
    w.heap_packed.reserve(dict + 2);
    w.heap_pair.reserve(dict + 2);
}

int build_codebook(const uint64_t* freq, int dict, Codebook& cb, CodebookWork* work) {
    CodebookWork& t_work = work ? *work : t_work_default;
    cb.dict = dict;
    cb.cw_hi.assign(dict, 0);
    cb.cw_lo.assign(dict, 0);
    cb.len.assign(dict, 0);
    int nsym = 0;
    for (int s = 0; s < dict; s++) nsym += freq[s] != 0;
    cb.nsym = nsym;
    if (nsym == 0) return VECSZ_EINVAL;

    // Node numbering: leaves 0..nsym-1 in ascending symbol order, then one
    // internal node per merge, so a parent always outnumbers its children and
    // the root is the last node.
    const int nc = 2 * nsym - 1;
    std::vector<int32_t>& sym = t_work.sym;
    std::vector<int32_t>& kid = t_work.kid;  // kid[2(n - nsym)] = left, +1 = right
    sym.resize(nsym);
    kid.resize(2 * (size_t)nsym);
    uint64_t total = 0;
    for (int s2 = 0; s2 < dict; s2++) total += freq[s2];
    if (total < (1ull << (64 - kNodeBits)))
        merge_all<PackedQueue>(freq, dict, nsym, sym.data(), kid.data(), t_work);
    else
        merge_all<PairQueue>(freq, dict, nsym, sym.data(), kid.data(), t_work);

    // Codewords top-down in descending node order (parents before children).
    std::vector<uint64_t>& bits = t_work.bits;
    std::vector<int32_t>& len = t_work.len;
    bits.resize(nc);
    len.resize(nc);
    bits[nc - 1] = 0;
    len[nc - 1] = 0;
    for (int n = nc - 1; n >= nsym; n--) {
        const int l = kid[2 * (n - nsym)], r = kid[2 * (n - nsym) + 1];
        const int ln = len[n] + 1;
        len[l] = ln;
        len[r] = ln;
        bits[l] = bits[n] << 1;
        bits[r] = (bits[n] << 1) | 1;
    }
    int max_len = 0;
    for (int n = 0; n < nsym; n++) {
        const int L = len[n];
        if (L > 64) return VECSZ_ECODELEN;
        const int s = sym[n];
        cb.len[s] = (uint8_t)L;
        cb.cw_lo[s] = (uint32_t)bits[n];
        if (L > 32) cb.cw_hi[s] = (uint32_t)(bits[n] >> 32);
        max_len = L > max_len ? L : max_len;
    }
    cb.max_len = max_len;

    // Serialization (pad_tree_*, src/huffman.hh:374-423): nodes are numbered
    // in preorder, left subtree first, root 0.  With subtree sizes that is
    // idx(left) = idx(parent) + 1, idx(right) = idx(left) + size(left).
    std::vector<int32_t>& size = t_work.size;
    std::vector<int32_t>& idx = t_work.idx;
    size.resize(nc);
    idx.resize(nc);
    for (int n = 0; n < nsym; n++) size[n] = 1;
    for (int n = nsym; n < nc; n++) size[n] = 1 + size[kid[2 * (n - nsym)]] + size[kid[2 * (n - nsym) + 1]];
    idx[nc - 1] = 0;
    for (int n = nc - 1; n >= nsym; n--) {
        const int l = kid[2 * (n - nsym)], r = kid[2 * (n - nsym) + 1];
        idx[l] = idx[n] + 1;
        idx[r] = idx[n] + 1 + size[l];
    }
    cb.node_count = nc;
    const int w = nc <= 256 ? 1 : (nc <= 65536 ? 2 : 4);
    cb.tree.assign(tree_byte_size(nc), 0);  // byte 0: CPU_sysEndianType (src/huffman.hh:72)
    uint8_t* Lp = cb.tree.data() + 1;
    uint8_t* Rp = Lp + (size_t)nc * w;
    uint8_t* Cp = Rp + (size_t)nc * w;
    uint8_t* Tp = Cp + (size_t)nc * 4;
    auto put = [&](uint8_t* base, int i, uint32_t v) {
        for (int b = 0; b < w; b++) base[(size_t)w * i + b] = (uint8_t)(v >> (8 * b));
    };
    for (int n = 0; n < nsym; n++) {
        const int i = idx[n];
        const uint32_t s = (uint32_t)sym[n];
        Cp[4 * (size_t)i] = (uint8_t)s;
        Cp[4 * (size_t)i + 1] = (uint8_t)(s >> 8);
        Cp[4 * (size_t)i + 2] = (uint8_t)(s >> 16);
        Cp[4 * (size_t)i + 3] = (uint8_t)(s >> 24);
        Tp[i] = 1;
    }
    for (int n = nsym; n < nc; n++) {
        const int i = idx[n];
        put(Lp, i, (uint32_t)idx[kid[2 * (n - nsym)]]);
        put(Rp, i, (uint32_t)idx[kid[2 * (n - nsym) + 1]]);
    }
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
    // Multi-level tables.  fill() walks `bits` bits from `node` (`depth` bits
    // already consumed) for every bit pattern: a leaf reached early is
    // replicated over the unused low bits, a node still internal after all
    // `bits` bits is queued and its entry patched once its own table exists.
    // A full tree gives each internal node at most one table; the depth and
    // size limits stop corrupt (cyclic or shared) node tables.
    struct Pending {
        uint32_t node, depth;
        bool in_lut;
        size_t slot;
    };
    std::vector<Pending> todo;
    std::vector<uint32_t> tab_depth;  // bits consumed before each deeper table
    // Depth-first over the subtree of `node`: a leaf b bits below it fills
    // its 2^(bits - b) entries at once, so a table costs its size plus the
    // nodes it covers.
    struct Walk {
        uint32_t node, b, prefix;
    };
    std::vector<Walk> stk;
    auto fill = [&](uint32_t node, uint32_t depth, uint32_t bits, bool in_lut, size_t out_base) -> int {
        uint32_t* out = (in_lut ? dt.lut : dt.sub).data() + out_base;
        for (uint32_t p = 0; p < (1u << bits); p++) out[p] = 0;  // invalid unless reached
        stk.clear();
        stk.push_back({node, 0, 0});
        while (!stk.empty()) {
            const Walk w = stk.back();
            stk.pop_back();
            for (uint32_t bit = 0; bit < 2; bit++) {
                const uint32_t nx = dt.child[2 * w.node + bit];
                if (nx == kNodeAbsent) continue;
                const uint32_t b = w.b + 1, prefix = (w.prefix << 1) | bit;
                if (nx & kNodeLeaf) {
                    const uint32_t len = depth + b;
                    if (len > 64) return VECSZ_ECORRUPT;
                    const uint32_t e = (len << 16) | (nx & 0xFFFF);
                    const uint32_t lo = prefix << (bits - b), hi = (prefix + 1) << (bits - b);
                    for (uint32_t q = lo; q < hi; q++) out[q] = e;
                } else if (b == bits) {
                    if (depth + bits >= 64) return VECSZ_ECORRUPT;
                    out[prefix] = kLutSlow;  // patched with the table index below
                    todo.push_back({nx, depth + bits, in_lut, out_base + prefix});
                } else {
                    if (stk.size() > 4096) return VECSZ_ECORRUPT;
                    stk.push_back({nx, b, prefix});
                }
            }
        }
        return VECSZ_OK;
    };
    dt.lut.assign((size_t)1 << kLutBits, 0);
    dt.sub.clear();
    int st = fill(0, 0, kLutBits, true, 0);
    if (st) return st;
    for (size_t i = 0; i < todo.size(); i++) {
        const Pending pd = todo[i];
        const size_t base = dt.sub.size();
        if (base + (1u << kSubTabBits) > ((size_t)1 << 24)) return VECSZ_ECORRUPT;
        dt.sub.resize(base + (1u << kSubTabBits));
        (pd.in_lut ? dt.lut : dt.sub)[pd.slot] = kLutSlow | (uint32_t)base;
        tab_depth.push_back(pd.depth);
        st = fill(pd.node, pd.depth, kSubTabBits, false, base);
        if (st) return st;
    }

    // compact 16-bit form
    dt.lut16.clear();
    dt.sub16.clear();
    bool ok16 = tab_depth.size() <= 4096;
    auto to16 = [&](uint32_t e, uint32_t depth, uint16_t& out) {
        if (e == 0) return false;  // invalid path
        if (e & kLutSlow) {
            out = (uint16_t)((e & 0x7FFFFFFFu) >> kSubTabBits);
            return true;
        }
        const uint32_t sym = e & 0xFFFF, used = ((e >> 16) & 0x7F) - depth;
        if (sym >= 4096 || used == 0 || used > 15) return false;
        out = (uint16_t)((used << 12) | sym);
        return true;
    };
    if (ok16) {
        dt.lut16.resize(dt.lut.size());
        for (size_t i = 0; ok16 && i < dt.lut.size(); i++) ok16 = to16(dt.lut[i], 0, dt.lut16[i]);
        dt.sub16.resize(dt.sub.size());
        for (size_t i = 0; ok16 && i < dt.sub.size(); i++)
            ok16 = to16(dt.sub[i], tab_depth[i >> kSubTabBits], dt.sub16[i]);
    }
    if (!ok16) {
        dt.lut16.clear();
        dt.sub16.clear();
    }
    return VECSZ_OK;
}

}  // namespace vsz
