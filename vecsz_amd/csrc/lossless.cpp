// Host-side lossless pass over archive payloads and the .sz file container.
//
// The reference runs gzip (zlib 1.2.11) or zstd (1.3.5) at level 3 over the
// datapack payload after the GPU-relevant path (src/lossless.hh:179-219,
// src/vecsz_interface.hh:206-217), and recognises the result by its first
// bytes (src/lossless.hh:40-75).  gzip goes through the system zlib, which is
// the same 1.2.11 release, so its output is the reference's byte for byte.
// zstd goes through the system libzstd runtime (loaded on first use): frames
// are interchangeable with the reference's, but the compressed bytes of a
// newer release are not the 1.3.5 bytes.
#include <dlfcn.h>
#include <stdint.h>
#include <string.h>
#include <zlib.h>

#include <mutex>

#include "../../include/vecsz_hip.h"

namespace {

constexpr int kLevel = 3;          // the reference's level for both compressors (src/vecsz_interface.hh:213-214)
constexpr uInt kZChunk = 65536;    // SZ_ZLIB_BUFFER_SIZE (src/lossless.hh:31)

struct Zstd {
    size_t (*compress)(void*, size_t, const void*, size_t, int) = nullptr;
    size_t (*decompress)(void*, size_t, const void*, size_t) = nullptr;
    size_t (*bound)(size_t) = nullptr;
    unsigned (*is_error)(size_t) = nullptr;
    unsigned long long (*content_size)(const void*, size_t) = nullptr;
    bool ok = false;
};

const Zstd& zstd() {
    static Zstd z;
    static std::once_flag once;
    std::call_once(once, [] {
        void* h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) return;
        z.compress = reinterpret_cast<decltype(z.compress)>(dlsym(h, "ZSTD_compress"));
        z.decompress = reinterpret_cast<decltype(z.decompress)>(dlsym(h, "ZSTD_decompress"));
        z.bound = reinterpret_cast<decltype(z.bound)>(dlsym(h, "ZSTD_compressBound"));
        z.is_error = reinterpret_cast<decltype(z.is_error)>(dlsym(h, "ZSTD_isError"));
        z.content_size = reinterpret_cast<decltype(z.content_size)>(dlsym(h, "ZSTD_getFrameContentSize"));
        z.ok = z.compress && z.decompress && z.bound && z.is_error && z.content_size;
    });
    return z;
}

// zlib stream header pairs the reference accepts (isZlibFormat, src/lossless.hh:40-58).
bool zlib_magic(uint8_t a, uint8_t b) {
    static const uint8_t pairs[][2] = {{104, 5}, {104, 129}, {104, 222}, {120, 1}, {120, 94}, {120, 156}, {120, 218}};
    for (const auto& p : pairs)
        if (a == p[0] && b == p[1]) return true;
    return false;
}

bool zstd_frame(const uint8_t* in, uint64_t n) {
    // frame magic 0xFD2FB528, little endian (the reference asks
    // ZSTD_getFrameContentSize, src/lossless.hh:62-65, which fails on anything else)
    if (n < 4 || !(in[0] == 0x28 && in[1] == 0xB5 && in[2] == 0x2F && in[3] == 0xFD)) return false;
    const Zstd& z = zstd();
    if (!z.ok) return true;  // a zstd frame, even if it cannot be decoded here
    return z.content_size(in, n) != (unsigned long long)-2;  // ZSTD_CONTENTSIZE_ERROR
}

}  // namespace

extern "C" {

int vecsz_lossless_bound(int kind, uint64_t bytes, uint64_t* bound) {
    if (!bound) return VECSZ_EINVAL;
    if (kind == VECSZ_LOSSLESS_GZIP) {
        *bound = compressBound((uLong)bytes) + 64;
        return VECSZ_OK;
    }
    if (kind == VECSZ_LOSSLESS_ZSTD) {
        const Zstd& z = zstd();
        if (!z.ok) return VECSZ_EINVAL;
        *bound = z.bound((size_t)bytes);
        return VECSZ_OK;
    }
    return VECSZ_EINVAL;
}

int vecsz_lossless_compress(int kind, const uint8_t* in, uint64_t bytes, uint8_t* out, uint64_t out_cap,
                            uint64_t* out_bytes) {
    if (!in || !out || !out_bytes) return VECSZ_EINVAL;
    if (kind == VECSZ_LOSSLESS_GZIP) {
        // deflate at level 3, input fed in 64 KiB pieces as the reference does
        // (zlib_compress, src/lossless.hh:78-133)
        z_stream s;
        memset(&s, 0, sizeof(s));
        if (deflateInit(&s, kLevel) != Z_OK) return VECSZ_EINVAL;
        uint64_t done = 0;
        int flush;
        do {
            const uint64_t take = bytes - done > kZChunk ? kZChunk : bytes - done;
            flush = done + take >= bytes ? Z_FINISH : Z_NO_FLUSH;
            s.next_in = const_cast<Bytef*>(in + done);
            s.avail_in = (uInt)take;
            do {
                const uint64_t room = out_cap - s.total_out;
                if (room == 0) {
                    deflateEnd(&s);
                    return VECSZ_ENOMEM;
                }
                s.next_out = out + s.total_out;
                s.avail_out = (uInt)(room > kZChunk ? kZChunk : room);
                const int rc = deflate(&s, flush);
                if (rc == Z_STREAM_ERROR) {
                    deflateEnd(&s);
                    return VECSZ_EINVAL;
                }
            } while (s.avail_out == 0);
            done += take;
        } while (flush != Z_FINISH);
        *out_bytes = s.total_out;
        deflateEnd(&s);
        return VECSZ_OK;
    }
    if (kind == VECSZ_LOSSLESS_ZSTD) {
        const Zstd& z = zstd();
        if (!z.ok) return VECSZ_EINVAL;
        const size_t r = z.compress(out, (size_t)out_cap, in, (size_t)bytes, kLevel);
        if (z.is_error(r)) return VECSZ_ENOMEM;
        *out_bytes = r;
        return VECSZ_OK;
    }
    return VECSZ_EINVAL;
}

int vecsz_lossless_kind(const uint8_t* in, uint64_t bytes) {
    if (!in || bytes < 2) return 0;
    if (zstd_frame(in, bytes)) return VECSZ_LOSSLESS_ZSTD;
    if (zlib_magic(in[0], in[1])) return VECSZ_LOSSLESS_GZIP;
    return 0;
}

int vecsz_lossless_decompress(const uint8_t* in, uint64_t bytes, uint8_t* out, uint64_t target, uint64_t* out_bytes) {
    if (!in || !out || !out_bytes) return VECSZ_EINVAL;
    const int kind = vecsz_lossless_kind(in, bytes);
    if (kind == VECSZ_LOSSLESS_GZIP) {
        z_stream s;
        memset(&s, 0, sizeof(s));
        if (inflateInit(&s) != Z_OK) return VECSZ_ECORRUPT;
        s.next_in = const_cast<Bytef*>(in);
        s.next_out = out;
        int rc = Z_OK;
        while (s.total_out < target && s.total_in < bytes) {
            const uint64_t left_in = bytes - s.total_in, left_out = target - s.total_out;
            s.avail_in = (uInt)(left_in > kZChunk ? kZChunk : left_in);
            s.avail_out = (uInt)(left_out > kZChunk ? kZChunk : left_out);
            rc = inflate(&s, Z_SYNC_FLUSH);
            if (rc == Z_STREAM_END) break;
            if (rc != Z_OK) break;
        }
        *out_bytes = s.total_out;
        inflateEnd(&s);
        return (rc == Z_OK || rc == Z_STREAM_END) && s.total_out == target ? VECSZ_OK : VECSZ_ECORRUPT;
    }
    if (kind == VECSZ_LOSSLESS_ZSTD) {
        const Zstd& z = zstd();
        if (!z.ok) return VECSZ_EINVAL;
        const size_t r = z.decompress(out, (size_t)target, in, (size_t)bytes);
        if (z.is_error(r) || r != target) return VECSZ_ECORRUPT;
        *out_bytes = r;
        return VECSZ_OK;
    }
    return VECSZ_ECORRUPT;
}

int vecsz_sz_pack(const uint8_t* body, uint64_t body_bytes, uint64_t lossless_size, uint64_t data_size, uint8_t* out,
                  uint64_t out_cap, uint64_t* out_bytes) {
    if (!body || !out || !out_bytes) return VECSZ_EINVAL;
    if (body_bytes != (lossless_size ? lossless_size : data_size)) return VECSZ_EINVAL;
    if (out_cap < 16 || body_bytes > out_cap - 16) return VECSZ_ENOMEM;
    memcpy(out, &lossless_size, 8);
    memcpy(out + 8, &data_size, 8);
    memcpy(out + 16, body, body_bytes);
    *out_bytes = 16 + body_bytes;
    return VECSZ_OK;
}

int vecsz_sz_unpack(const uint8_t* file, uint64_t bytes, uint64_t* lossless_size, uint64_t* data_size,
                    uint64_t* body_offset, uint64_t* body_bytes) {
    if (!file || !lossless_size || !data_size || !body_offset || !body_bytes) return VECSZ_EINVAL;
    if (bytes < 16) return VECSZ_ECORRUPT;
    memcpy(lossless_size, file, 8);
    memcpy(data_size, file + 8, 8);
    *body_offset = 16;
    *body_bytes = *lossless_size ? *lossless_size : *data_size;
    // the reference's lossless files carry data_size bytes after the header
    // (D8: written from a lossless_size-byte buffer); the body is the first
    // lossless_size of them
    if (*body_bytes > bytes - 16) return VECSZ_ECORRUPT;  // (bytes >= 16: no wrap)
    return VECSZ_OK;
}

}  // extern "C"
