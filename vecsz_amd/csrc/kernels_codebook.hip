// gfx950 kernel of the Huffman codebook, built on the device from the
// device histogram (the reference builds it on the host: initHuffman's
// queue and merge loop, src/huffman.hh:169-192, :244-262; build_code,
// :200-227; the tree serialization, :374-502, written by encode_withTree,
// :643-653).  One workgroup:
//
//   leaves     the nonzero bins in ascending symbol order (a workgroup scan)
//   queue      the reference's priority queue and merge loop, replayed by one
//              lane in LDS: every operation's path depends on the one before,
//              so this part is sequential by nature
//   tree       parents from the merge record; each leaf then walks to the
//              root counting leaves per subtree (LDS atomics), and every node
//              walks to the root again for its depth, its codeword (leaves)
//              and its preorder number (1 per edge, plus the left sibling's
//              subtree size where the walk leaves a right child)
//   output     codebook tables (codeword high / low words, lengths), the
//              Huffman header and the serialized tree in the archive, and the
//              node count, tree bytes and longest codeword
//
// The tree is the reference's, node for node, so the archive bytes are the
// host build's.
#include "../../include/vecsz_hip.h"
#include "common.hpp"
#include "kernels.hpp"

namespace vsz {

namespace {

constexpr int kCbNodeBits = 17;  // queue entries: (freq << 17) | node, as the host's packed queue
__device__ __forceinline__ uint64_t qf(uint64_t e) { return e >> kCbNodeBits; }
__device__ __forceinline__ uint32_t qn(uint64_t e) { return (uint32_t)(e & ((1u << kCbNodeBits) - 1)); }

}  // namespace

// LDS layout (bytes): queue u64[dict + 2] (later the per-node leaf counts,
// u32), leaf frequencies u64[dict], then u16 arrays: symbols [dict], merge
// record [2 dict], parent | side [2 dict], preorder numbers [2 dict].
size_t codebook_lds_bytes(int dict) {
    const size_t d = (size_t)dict;
    return (d + 2) * 8 + d * 8 + 7 * d * 2;
}

__global__ void __launch_bounds__(256) k_codebook(const unsigned long long* __restrict__ hist, int dict,
                                                  uint32_t* __restrict__ book, uint8_t* __restrict__ huff,
                                                  uint64_t huff_cap, CodebookInfo* __restrict__ info) {
    extern __shared__ uint64_t cb_lds[];
    const uint32_t D = (uint32_t)dict;
    uint64_t* qq = cb_lds;                                          // queue, 1-based
    uint32_t* leaves = reinterpret_cast<uint32_t*>(cb_lds);         // after the merges: leaves per subtree
    uint64_t* lfreq = cb_lds + D + 2;                               // leaf n's frequency
    uint16_t* sym = reinterpret_cast<uint16_t*>(lfreq + D);         // leaf n's symbol
    uint16_t* kid = sym + D;                                        // internal node nsym + k: kid[2k] left, kid[2k+1] right
    uint16_t* par = kid + 2 * D;                                    // parent << 1 | side; 0xFFFF: the root
    uint16_t* pre = par + 2 * D;                                    // preorder number
    __shared__ uint32_t s_scan[256], s_nsym, s_maxlen, s_status;
    const uint32_t t = threadIdx.x;

    // leaves: the nonzero bins in ascending symbol order
    const uint32_t per = (D + 255) / 256, b0 = t * per, b1 = min(b0 + per, D);
    uint32_t cnt = 0;
    for (uint32_t s = b0; s < b1; s++) cnt += hist[s] != 0;
    s_scan[t] = cnt;
    if (t == 0) {
        s_maxlen = 0;
        s_status = VECSZ_OK;
    }
    mem_sync();
    uint32_t off = 0;
    for (uint32_t i = 0; i < t; i++) off += s_scan[i];  // 256 entries: a short serial prefix per thread
    for (uint32_t s = b0; s < b1; s++) {
        const uint64_t f = hist[s];
        if (f != 0) {
            lfreq[off] = f;
            sym[off++] = (uint16_t)s;
        }
    }
    if (t == 255) s_nsym = off;
    mem_sync();
    const uint32_t nsym = s_nsym;
    if (nsym == 0) {
        if (t == 0) info->status = VECSZ_EINVAL;
        return;
    }
    const uint32_t nc = 2 * nsym - 1;

    // the reference's queue and merge loop, on one lane
    if (t == 0) {
        uint32_t qend = 1;
        auto qinsert = [&](uint64_t e) {
            uint32_t i = qend++, j;
            const uint64_t fe = qf(e);
            while ((j = i >> 1)) {
                const uint64_t pj = qq[j];
                if (qf(pj) <= fe) break;
                qq[i] = pj;
                i = j;
            }
            qq[i] = e;
        };
        auto qremove = [&]() -> uint64_t {
            const uint64_t top = qq[1];
            const uint32_t end = --qend;
            uint32_t i = 1, l;
            while ((l = i << 1) < end) {
                uint64_t el = qq[l];
                if (l + 1 < end) {
                    const uint64_t er = qq[l + 1];
                    if (qf(er) < qf(el)) {
                        el = er;
                        l++;
                    }
                }
                qq[i] = el;
                i = l;
            }
            qq[i] = qq[end];
            return top;
        };
        for (uint32_t n = 0; n < nsym; n++) qinsert((lfreq[n] << kCbNodeBits) | n);
        for (uint32_t n = nsym; qend > 2; n++) {
            const uint64_t b = qremove();  // first removed: the right child
            const uint64_t a = qremove();
            kid[2 * (n - nsym)] = (uint16_t)qn(a);
            kid[2 * (n - nsym) + 1] = (uint16_t)qn(b);
            qinsert(((qf(a) + qf(b)) << kCbNodeBits) | n);
        }
    }
    mem_sync();

    // parents, and the leaf count of every subtree
    for (uint32_t v = t; v < nc; v += 256) {
        leaves[v] = v < nsym ? 1u : 0u;
        if (v == nc - 1) par[v] = 0xFFFF;
    }
    for (uint32_t k = t; k + 1 < nsym; k += 256) {
        const uint32_t m = nsym + k;
        par[kid[2 * k]] = (uint16_t)(m << 1);
        par[kid[2 * k + 1]] = (uint16_t)((m << 1) | 1);
    }
    mem_sync();
    // (every walk is bounded by the node count, so a damaged record cannot loop)
    for (uint32_t v = t; v < nsym; v += 256) {
        uint32_t steps = 0;
        for (uint32_t p = par[v]; p != 0xFFFF && steps < nc; p = par[p >> 1], steps++) atomicAdd(&leaves[p >> 1], 1u);
    }
    mem_sync();

    // depth, codeword and preorder number of every node; leaves fill the codebook
    for (uint32_t s = t; s < D; s += 256) {
        book[s] = 0;
        book[D + s] = 0;
        reinterpret_cast<uint8_t*>(book + 2 * D)[s] = 0;
    }
    mem_sync();
    for (uint32_t v = t; v < nc; v += 256) {
        uint32_t depth = 0, idx = 0;
        uint64_t bits = 0;
        for (uint32_t c = v, p = par[v]; p != 0xFFFF && depth < nc; c = p >> 1, p = par[c]) {
            const uint32_t side = p & 1;
            if (depth < 64) bits |= (uint64_t)side << depth;
            depth++;
            idx += 1;
            if (side) {  // past the left sibling's subtree (2 leaves - 1 nodes)
                const uint32_t ls = kid[2 * ((p >> 1) - nsym)];
                idx += 2 * leaves[ls] - 1;
            }
        }
        pre[v] = (uint16_t)idx;
        if (v < nsym) {
            if (depth > 64) {
                s_status = VECSZ_ECODELEN;
            } else {
                const uint32_t s = sym[v];
                book[D + s] = (uint32_t)bits;
                book[s] = depth > 32 ? (uint32_t)(bits >> 32) : 0u;
                reinterpret_cast<uint8_t*>(book + 2 * D)[s] = (uint8_t)depth;
                atomicMax(&s_maxlen, depth);
            }
        }
    }
    mem_sync();

    // Huffman header and tree: [u32be node count][u32be dict][u8 0][L][R][C u32 LE][t u8],
    // L/R 1 byte wide up to 256 nodes, else 2 (src/huffman.hh:374-423)
    const uint32_t w = nc <= 256 ? 1u : 2u;
    const uint64_t tb = 1 + (uint64_t)nc * (2 * w + 4 + 1);
    if (8 + tb > huff_cap && t == 0) s_status = s_status ? s_status : VECSZ_ENOMEM;
    mem_sync();
    if (s_status == VECSZ_OK) {
        uint8_t* tree = huff + 8;
        uint8_t* Lp = tree + 1;
        uint8_t* Rp = Lp + (size_t)nc * w;
        uint8_t* Cp = Rp + (size_t)nc * w;
        uint8_t* Tp = Cp + (size_t)nc * 4;
        if (t == 0) {
            const uint32_t hv[2] = {nc, D};
            for (int q = 0; q < 2; q++)
                for (int b = 0; b < 4; b++) huff[4 * q + b] = (uint8_t)(hv[q] >> (24 - 8 * b));
            tree[0] = 0;  // CPU_sysEndianType (src/huffman.hh:72)
        }
        for (uint32_t v = t; v < nc; v += 256) {
            const uint32_t i = pre[v];
            const bool leaf = v < nsym;
            const uint32_t l = leaf ? 0u : pre[kid[2 * (v - nsym)]], r = leaf ? 0u : pre[kid[2 * (v - nsym) + 1]];
            const uint32_t c = leaf ? sym[v] : 0u;
            for (uint32_t b = 0; b < w; b++) {
                Lp[(size_t)w * i + b] = (uint8_t)(l >> (8 * b));
                Rp[(size_t)w * i + b] = (uint8_t)(r >> (8 * b));
            }
            for (int b = 0; b < 4; b++) Cp[4 * (size_t)i + b] = (uint8_t)(c >> (8 * b));
            Tp[i] = leaf ? 1 : 0;
        }
    }
    if (t == 0) {
        info->node_count = nc;
        info->tree_bytes = (uint32_t)tb;
        info->max_len = s_maxlen;
        info->status = s_status;
    }
}

void launch_codebook(const unsigned long long* hist, int dict, uint32_t* book, uint8_t* huff, uint64_t huff_cap,
                     CodebookInfo* info, hipStream_t st) {
    const size_t lds = codebook_lds_bytes(dict);
    // more than the default 64 KiB of dynamic LDS (the kernel's static LDS,
    // about 1 KiB, counts against the 160 KiB too)
    static const bool big = [] {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_codebook), hipFuncAttributeMaxDynamicSharedMemorySize,
                                156 * 1024) != hipSuccess)
            (void)hipGetLastError();  // the launch then reports what does not fit
        return true;
    }();
    (void)big;
    hipLaunchKernelGGL(k_codebook, dim3(1), dim3(256), lds, st, hist, dict, book, huff, huff_cap, info);
}

}  // namespace vsz
