// usct_h5.cpp -- reader and writer for the HDF5 subset of the acquisition
// dataset (software/schema/README.md:13-41; simulation/ringfwi/dataset.py:112-174).
//
// Host code, built with g++ -O2 -fopenmp, linked with zlib.  The structures
// follow the published HDF5 File Format Specification (superblock versions
// 0-3, object header versions 1-2, symbol-table and compact-link groups,
// data layout versions 3-4, version-1 B-trees, local and global heaps,
// filter pipeline versions 1-2).  Every length and address read from a file
// is bounds-checked against the file or the message that holds it, so a
// damaged file ends in USCT_H5_EFORMAT, never in an out-of-bounds access.

#include "../../include/usct_h5.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>
#include <omp.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct H5Err {
    int code;
    std::string msg;
};

[[noreturn]] void raise(int code, const std::string &msg) { throw H5Err{code, msg}; }

constexpr uint64_t UNDEF = ~0ull;
constexpr int MAX_DEPTH = 64;           // B-tree levels / group nesting followed
constexpr uint64_t MAX_NODES = 1u << 20;  // B-tree nodes visited for one listing
constexpr uint64_t MAX_META = 1ull << 30;  // largest metadata block read at once

uint64_t le(const uint8_t *p, int n) {
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}

uint64_t align8(uint64_t n) { return (n + 7) & ~uint64_t(7); }

// Bounds-checked little-endian cursor over a byte range.
struct Cur {
    const uint8_t *b, *p, *e;
    const char *what;
    Cur(const uint8_t *data, size_t n, const char *w) : b(data), p(data), e(data + n), what(w) {}
    void need(size_t n) const {
        if (size_t(e - p) < n) raise(USCT_H5_EFORMAT, std::string(what) + ": truncated");
    }
    uint64_t u(int n) {
        need(n);
        uint64_t v = le(p, n);
        p += n;
        return v;
    }
    uint8_t u8() { return uint8_t(u(1)); }
    void skip(size_t n) {
        need(n);
        p += n;
    }
    const uint8_t *take(size_t n) {
        need(n);
        const uint8_t *q = p;
        p += n;
        return q;
    }
    size_t left() const { return size_t(e - p); }
    size_t pos() const { return size_t(p - b); }
};

// ---------------------------------------------------------------------------
// Element types and conversion.

struct DType {
    int cls = -1;  // 0 fixed-point, 1 float, 3 string, 9 variable-length, other: unsupported
    uint32_t size = 0;
    bool be = false, sgn = false;
    bool vlen_str = false;
    std::shared_ptr<DType> base;
};

int type_code(const DType &t) {
    if (t.cls == 0) {
        switch (t.size) {
        case 1: return t.sgn ? USCT_H5_I8 : USCT_H5_U8;
        case 2: return t.sgn ? USCT_H5_I16 : USCT_H5_U16;
        case 4: return t.sgn ? USCT_H5_I32 : USCT_H5_U32;
        case 8: return t.sgn ? USCT_H5_I64 : USCT_H5_U64;
        }
        return USCT_H5_OTHER;
    }
    if (t.cls == 1) {
        if (t.size == 2) return USCT_H5_F16;
        if (t.size == 4) return USCT_H5_F32;
        if (t.size == 8) return USCT_H5_F64;
        return USCT_H5_OTHER;
    }
    if (t.cls == 3 || (t.cls == 9 && t.vlen_str)) return USCT_H5_STR;
    return USCT_H5_OTHER;
}

int code_size(int code) {
    switch (code) {
    case USCT_H5_I8: case USCT_H5_U8: return 1;
    case USCT_H5_I16: case USCT_H5_U16: case USCT_H5_F16: return 2;
    case USCT_H5_I32: case USCT_H5_U32: case USCT_H5_F32: return 4;
    case USCT_H5_I64: case USCT_H5_U64: case USCT_H5_F64: return 8;
    }
    return 0;
}

// IEEE binary16 -> binary32, exact (including subnormals, infinities and NaN).
float half_to_float(uint16_t h) {
    uint32_t s = uint32_t(h >> 15) << 31, ex = (h >> 10) & 31, m = h & 1023, bits;
    if (ex == 0) {
        if (m == 0) {
            bits = s;
        } else {  // subnormal: normalise
            int k = 0;
            while (!(m & 1024)) { m <<= 1; ++k; }
            bits = s | uint32_t(127 - 15 - k + 1) << 23 | (m & 1023) << 13;
        }
    } else if (ex == 31) {
        bits = s | 0x7f800000u | m << 13;
    } else {
        bits = s | (ex + 127 - 15) << 23 | m << 13;
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// Decode one element of a numeric file type.
struct Num {
    bool fl;
    double d;
    int64_t i;
    uint64_t u;
};

Num load_num(const uint8_t *p, const DType &t) {
    uint8_t tmp[8];
    const uint8_t *q = p;
    if (t.be) {
        for (uint32_t k = 0; k < t.size; ++k) tmp[k] = p[t.size - 1 - k];
        q = tmp;
    }
    Num n{false, 0.0, 0, 0};
    if (t.cls == 1) {
        n.fl = true;
        if (t.size == 2) n.d = half_to_float(uint16_t(le(q, 2)));
        else if (t.size == 4) { float f; std::memcpy(&f, q, 4); n.d = f; }
        else { std::memcpy(&n.d, q, 8); }
        return n;
    }
    uint64_t v = le(q, int(t.size));
    if (t.sgn && t.size < 8 && (v >> (8 * t.size - 1) & 1)) v |= ~0ull << (8 * t.size);
    n.u = v;
    n.i = int64_t(v);
    if (!t.sgn) n.i = int64_t(v);
    return n;
}

template <class T>
T num_as(const Num &n, bool src_signed) {
    if (n.fl) return T(n.d);
    return src_signed ? T(n.i) : T(n.u);
}

void store_num(const Num &n, bool src_signed, int code, uint8_t *dst) {
    switch (code) {
    case USCT_H5_F32: { float v = num_as<float>(n, src_signed); std::memcpy(dst, &v, 4); break; }
    case USCT_H5_F64: { double v = num_as<double>(n, src_signed); std::memcpy(dst, &v, 8); break; }
    case USCT_H5_I8: { int8_t v = num_as<int8_t>(n, src_signed); std::memcpy(dst, &v, 1); break; }
    case USCT_H5_I16: { int16_t v = num_as<int16_t>(n, src_signed); std::memcpy(dst, &v, 2); break; }
    case USCT_H5_I32: { int32_t v = num_as<int32_t>(n, src_signed); std::memcpy(dst, &v, 4); break; }
    case USCT_H5_I64: { int64_t v = num_as<int64_t>(n, src_signed); std::memcpy(dst, &v, 8); break; }
    case USCT_H5_U8: { uint8_t v = num_as<uint8_t>(n, src_signed); std::memcpy(dst, &v, 1); break; }
    case USCT_H5_U16: { uint16_t v = num_as<uint16_t>(n, src_signed); std::memcpy(dst, &v, 2); break; }
    case USCT_H5_U32: { uint32_t v = num_as<uint32_t>(n, src_signed); std::memcpy(dst, &v, 4); break; }
    case USCT_H5_U64: { uint64_t v = num_as<uint64_t>(n, src_signed); std::memcpy(dst, &v, 8); break; }
    }
}

// Convert n elements from file type t to output code.  Same-type little-endian
// runs are copied; float32 -> float64 (the channel data into FP64) has its own loop.
void convert(const uint8_t *src, const DType &t, int code, uint8_t *dst, size_t n) {
    int sc = type_code(t);
    if (sc == code && !t.be) {
        std::memcpy(dst, src, n * t.size);
        return;
    }
    if (sc == USCT_H5_F32 && code == USCT_H5_F64 && !t.be) {
        const float *s = reinterpret_cast<const float *>(src);
        double *d = reinterpret_cast<double *>(dst);
        for (size_t k = 0; k < n; ++k) d[k] = double(s[k]);
        return;
    }
    int os = code_size(code);
    bool sg = t.sgn;
    for (size_t k = 0; k < n; ++k) store_num(load_num(src + k * t.size, t), sg, code, dst + k * os);
}

// ---------------------------------------------------------------------------
// Messages, dataspaces, layouts.

struct Msg {
    int type;
    std::vector<uint8_t> data;
};

struct DSpace {
    bool null = false;
    std::vector<uint64_t> dims;  // empty: scalar
    uint64_t count() const {
        if (null) return 0;
        uint64_t c = 1;
        for (uint64_t d : dims)
            if (__builtin_mul_overflow(c, d, &c)) raise(USCT_H5_EFORMAT, "dataspace: size overflow");
        return c;
    }
};

struct Filter {
    int id;
    std::vector<uint32_t> cd;
};

struct Chunk {
    std::vector<uint64_t> off;
    uint64_t addr;
    uint64_t size;
    uint32_t mask;
};

struct Desc {
    DSpace space;
    DType type;
    int layout = -1;  // 0 compact, 1 contiguous, 2 chunked
    uint64_t addr = UNDEF, csize = 0;
    std::vector<uint8_t> compact;
    std::vector<uint64_t> cdims;  // chunk extents (rank entries)
    int index = 0;                // chunked: 0 v1 B-tree, 1 single chunk, 2 implicit
    uint64_t single_size = 0;
    uint32_t single_mask = 0;
    bool single_filtered = false;
    std::vector<Filter> filters;
    std::vector<uint8_t> fill;
    std::vector<Chunk> chunks;
    bool chunks_loaded = false;
};

struct Attr {
    std::string name;
    DType type;
    DSpace space;
    std::vector<uint8_t> data;
};

DType parse_dtype(Cur &c, int depth = 0) {
    if (depth > 8) raise(USCT_H5_ENOTSUP, "datatype: nesting too deep");
    uint8_t cv = c.u8();
    int cls = cv & 15;
    uint32_t bits = uint32_t(c.u(3));
    DType t;
    t.cls = cls;
    t.size = uint32_t(c.u(4));
    switch (cls) {
    case 0:
        t.be = bits & 1;
        t.sgn = bits & 8;
        c.skip(4);  // bit offset, precision
        if (t.size != 1 && t.size != 2 && t.size != 4 && t.size != 8)
            raise(USCT_H5_ENOTSUP, "integer datatype of " + std::to_string(t.size) + " bytes");
        break;
    case 1: {
        t.be = bits & 1;
        if (bits & 0x40) raise(USCT_H5_ENOTSUP, "VAX-order floating point");
        c.skip(2);
        unsigned prec = unsigned(c.u(2));
        unsigned eloc = c.u8(), esz = c.u8(), mloc = c.u8(), msz = c.u8();
        uint32_t bias = uint32_t(c.u(4));
        bool ieee = (t.size == 2 && prec == 16 && eloc == 10 && esz == 5 && mloc == 0 && msz == 10 && bias == 15) ||
                    (t.size == 4 && prec == 32 && eloc == 23 && esz == 8 && mloc == 0 && msz == 23 && bias == 127) ||
                    (t.size == 8 && prec == 64 && eloc == 52 && esz == 11 && mloc == 0 && msz == 52 && bias == 1023);
        if (!ieee) raise(USCT_H5_ENOTSUP, "non-IEEE floating-point datatype");
        break;
    }
    case 3:
        break;
    case 9:
        t.vlen_str = (bits & 15) == 1;
        t.base = std::make_shared<DType>(parse_dtype(c, depth + 1));
        break;
    default:
        break;  // reported as USCT_H5_OTHER; its properties are not needed
    }
    return t;
}

DSpace parse_space(Cur &c, int sl) {
    int ver = c.u8();
    int rank = c.u8();
    int flags = c.u8();
    DSpace s;
    if (ver == 1) {
        c.skip(5);
    } else if (ver == 2) {
        int kind = c.u8();
        if (kind == 2) s.null = true;
    } else {
        raise(USCT_H5_ENOTSUP, "dataspace message version " + std::to_string(ver));
    }
    if (rank > USCT_H5_MAX_RANK) raise(USCT_H5_ENOTSUP, "dataspace rank above 32");
    for (int i = 0; i < rank; ++i) s.dims.push_back(c.u(sl));
    (void)flags;  // maximum dimensions follow; not needed to read
    return s;
}

// ---------------------------------------------------------------------------
// The open file.

}  // namespace

struct usct_h5_file {
    int fd = -1;
    uint64_t fsize = 0, base = 0;
    int so = 8, sl = 8;
    uint64_t root = UNDEF;
    std::mutex mu;  // guards the caches below
    std::map<std::string, uint64_t> paths;
    std::map<uint64_t, std::shared_ptr<Desc>> descs;

    ~usct_h5_file() {
        if (fd >= 0) ::close(fd);
    }

    void readat(uint64_t addr, void *dst, uint64_t n) const {
        if (addr == UNDEF) raise(USCT_H5_EFORMAT, "read at an undefined address");
        uint64_t a = base + addr;
        if (a < base || a > fsize || n > fsize - a)
            raise(USCT_H5_EFORMAT, "read past the end of the file (truncated?)");
        uint64_t done = 0;
        while (done < n) {
            ssize_t r = ::pread(fd, static_cast<char *>(dst) + done, size_t(n - done), off_t(a + done));
            if (r < 0 && errno == EINTR) continue;
            if (r <= 0) raise(USCT_H5_EIO, std::string("read failed: ") + std::strerror(errno));
            done += uint64_t(r);
        }
    }

    std::vector<uint8_t> bytes(uint64_t addr, uint64_t n) const {
        if (addr == UNDEF || base + addr > fsize || n > fsize - (base + addr) || n > MAX_META)
            raise(USCT_H5_EFORMAT, "metadata block past the end of the file (truncated?)");
        std::vector<uint8_t> v(n);
        if (n) readat(addr, v.data(), n);
        return v;
    }

    uint64_t off(Cur &c) const {
        uint64_t v = c.u(so);
        if (so < 8 && v == (1ull << (8 * so)) - 1) return UNDEF;
        return v;
    }
    uint64_t len(Cur &c) const { return c.u(sl); }

    void open_super() {
        static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
        for (uint64_t at = 0; at + 8 <= fsize; at = at ? at * 2 : 512) {
            uint8_t s[8];
            base = 0;
            readat(at, s, 8);
            if (std::memcmp(s, sig, 8) != 0) continue;
            uint8_t hd[16];
            readat(at + 8, hd, 1);
            int ver = hd[0];
            if (ver == 0 || ver == 1) {
                std::vector<uint8_t> v = bytes(at + 8, std::min<uint64_t>(fsize - at - 8, 256));
                Cur c(v.data(), v.size(), "superblock");
                c.skip(5);
                so = c.u8();
                sl = c.u8();
                if ((so != 2 && so != 4 && so != 8) || (sl != 2 && sl != 4 && sl != 8))
                    raise(USCT_H5_EFORMAT, "superblock: bad offset/length sizes");
                c.skip(1 + 2 + 2 + 4);
                if (ver == 1) c.skip(4);
                uint64_t b = off(c);
                off(c);  // free-space info
                off(c);  // end of file
                off(c);  // driver info
                off(c);  // root entry: link name offset
                root = off(c);
                (void)b;
                base = at;  // addresses count from the superblock, as libhdf5 resolves them
            } else if (ver == 2 || ver == 3) {
                std::vector<uint8_t> v = bytes(at + 8, std::min<uint64_t>(fsize - at - 8, 64));
                Cur c(v.data(), v.size(), "superblock");
                c.skip(1);
                so = c.u8();
                sl = c.u8();
                if ((so != 2 && so != 4 && so != 8) || (sl != 2 && sl != 4 && sl != 8))
                    raise(USCT_H5_EFORMAT, "superblock: bad offset/length sizes");
                c.skip(1);
                uint64_t b = off(c);
                off(c);  // extension
                off(c);  // end of file
                root = off(c);
                (void)b;
                base = at;
            } else {
                raise(USCT_H5_ENOTSUP, "superblock version " + std::to_string(ver));
            }
            if (root == UNDEF) raise(USCT_H5_EFORMAT, "superblock: no root group");
            return;
        }
        raise(USCT_H5_EFORMAT, "not an HDF5 file (no superblock signature)");
    }

    // All messages of an object header (versions 1 and 2), continuations followed.
    std::vector<Msg> header(uint64_t addr) const {
        std::vector<Msg> out;
        uint8_t pre[16];
        readat(addr, pre, 4);
        std::vector<std::pair<uint64_t, uint64_t>> todo;
        bool v2 = std::memcmp(pre, "OHDR", 4) == 0;
        if (v2) {
            uint64_t avail = fsize > base + addr ? fsize - (base + addr) : 0;
            std::vector<uint8_t> h = bytes(addr, std::min<uint64_t>(avail, 48));
            Cur c(h.data(), h.size(), "object header");
            c.skip(4);
            if (c.u8() != 2) raise(USCT_H5_EFORMAT, "object header: bad version");
            int flags = c.u8();
            if (flags & 0x20) c.skip(16);
            if (flags & 0x10) c.skip(4);
            uint64_t s0 = c.u(1 << (flags & 3));
            todo.push_back({addr + c.pos(), s0});
            size_t nchunk = 0;
            while (!todo.empty()) {
                if (++nchunk > 4096) raise(USCT_H5_EFORMAT, "object header: too many continuations");
                auto [a, n] = todo.back();
                todo.pop_back();
                std::vector<uint8_t> d = bytes(a, n);
                Cur m(d.data(), d.size(), "object header message");
                if (nchunk > 1) {
                    if (n < 8 || std::memcmp(d.data(), "OCHK", 4) != 0)
                        raise(USCT_H5_EFORMAT, "object header: bad continuation block");
                    m.skip(4);
                    m.e -= 4;  // checksum
                }
                size_t mh = 4 + ((flags & 0x04) ? 2 : 0);
                while (m.left() >= mh) {
                    int type = m.u8();
                    uint64_t sz = m.u(2);
                    int mf = m.u8();
                    (void)mf;
                    if (flags & 0x04) m.skip(2);
                    const uint8_t *p = m.take(sz);
                    if (type == 0x10) {
                        Cur cc(p, sz, "continuation message");
                        uint64_t ca = off(cc), cn = len(cc);
                        todo.push_back({ca, cn});
                    } else if (type != 0) {
                        out.push_back({type, std::vector<uint8_t>(p, p + sz)});
                    }
                }
            }
            return out;
        }
        readat(addr, pre, 16);
        if (pre[0] != 1) raise(USCT_H5_EFORMAT, "object header: unknown version " + std::to_string(pre[0]));
        uint64_t nmsg = le(pre + 2, 2), hsize = le(pre + 8, 4);
        todo.push_back({addr + 16, hsize});
        uint64_t seen = 0;
        size_t nchunk = 0;
        while (!todo.empty() && seen < nmsg) {
            if (++nchunk > 4096) raise(USCT_H5_EFORMAT, "object header: too many continuations");
            auto [a, n] = todo.front();
            todo.erase(todo.begin());
            std::vector<uint8_t> d = bytes(a, n);
            Cur m(d.data(), d.size(), "object header message");
            while (m.left() >= 8 && seen < nmsg) {
                int type = int(m.u(2));
                uint64_t sz = m.u(2);
                m.skip(4);  // flags, reserved
                const uint8_t *p = m.take(sz);
                ++seen;
                if (type == 0x10) {
                    Cur cc(p, sz, "continuation message");
                    uint64_t ca = off(cc), cn = len(cc);
                    todo.push_back({ca, cn});
                } else if (type != 0) {
                    out.push_back({type, std::vector<uint8_t>(p, p + sz)});
                }
            }
        }
        return out;
    }

    static const Msg *find(const std::vector<Msg> &ms, int type) {
        for (const Msg &m : ms)
            if (m.type == type) return &m;
        return nullptr;
    }

    // Members of a group: (name, object header address), in storage order.
    std::vector<std::pair<std::string, uint64_t>> links(const std::vector<Msg> &ms) const {
        std::vector<std::pair<std::string, uint64_t>> out;
        if (const Msg *st = find(ms, 0x11)) {
            Cur c(st->data.data(), st->data.size(), "symbol table message");
            uint64_t bt = off(c), hp = off(c);
            std::vector<uint8_t> hh = bytes(hp, 8 + 2 * sl + so);
            Cur h(hh.data(), hh.size(), "local heap");
            if (std::memcmp(h.take(4), "HEAP", 4) != 0) raise(USCT_H5_EFORMAT, "local heap: bad signature");
            h.skip(4);
            uint64_t dsz = len(h);
            len(h);
            uint64_t da = off(h);
            std::vector<uint8_t> heap = bytes(da, dsz);
            uint64_t visits = 0;
            group_btree(bt, heap, out, -1, visits);
        } else {
            for (const Msg &m : ms) {
                if (m.type != 0x06) continue;
                Cur c(m.data.data(), m.data.size(), "link message");
                c.u8();
                int fl = c.u8();
                int ltype = (fl & 0x08) ? c.u8() : 0;
                if (fl & 0x04) c.skip(8);
                if (fl & 0x10) c.skip(1);
                uint64_t nl = c.u(1 << (fl & 3));
                const uint8_t *nm = c.take(nl);
                if (ltype != 0) continue;  // soft / external links are not followed
                out.push_back({std::string(reinterpret_cast<const char *>(nm), nl), off(c)});
            }
            if (out.empty())
                if (const Msg *li = find(ms, 0x02)) {
                    Cur c(li->data.data(), li->data.size(), "link info message");
                    c.u8();
                    int fl = c.u8();
                    if (fl & 1) c.skip(8);
                    if (off(c) != UNDEF)
                        raise(USCT_H5_ENOTSUP, "group with dense link storage (fractal heap)");
                }
        }
        return out;
    }

    // expect: the level the parent says this node has (-1 at the root); levels
    // fall by one per step down, so a damaged tree cannot recurse forever.
    void group_btree(uint64_t node, const std::vector<uint8_t> &heap,
                     std::vector<std::pair<std::string, uint64_t>> &out, int expect, uint64_t &visits) const {
        if (++visits > MAX_NODES) raise(USCT_H5_EFORMAT, "group B-tree: too many nodes");
        std::vector<uint8_t> hd = bytes(node, 8 + 2 * so);
        Cur h(hd.data(), hd.size(), "group B-tree node");
        if (std::memcmp(h.take(4), "TREE", 4) != 0) raise(USCT_H5_EFORMAT, "group B-tree: bad signature");
        if (h.u8() != 0) raise(USCT_H5_EFORMAT, "group B-tree: wrong node type");
        int level = h.u8();
        if ((expect >= 0 && level != expect) || level > MAX_DEPTH)
            raise(USCT_H5_EFORMAT, "group B-tree: inconsistent node levels");
        uint64_t n = h.u(2);
        std::vector<uint8_t> body = bytes(node + 8 + 2 * so, n * so + (n + 1) * sl);
        Cur c(body.data(), body.size(), "group B-tree node");
        for (uint64_t i = 0; i < n; ++i) {
            c.skip(sl);
            uint64_t child = off(c);
            if (level > 0) {
                group_btree(child, heap, out, level - 1, visits);
                continue;
            }
            uint8_t sh[8];
            readat(child, sh, 8);
            if (std::memcmp(sh, "SNOD", 4) != 0) raise(USCT_H5_EFORMAT, "symbol node: bad signature");
            uint64_t ns = le(sh + 6, 2);
            uint64_t esz = 2 * uint64_t(so) + 24;
            std::vector<uint8_t> ents = bytes(child + 8, ns * esz);
            Cur e(ents.data(), ents.size(), "symbol node");
            for (uint64_t k = 0; k < ns; ++k) {
                uint64_t no = off(e), oh = off(e);
                e.skip(24);
                if (no >= heap.size()) raise(USCT_H5_EFORMAT, "symbol node: name outside the heap");
                const char *s = reinterpret_cast<const char *>(heap.data() + no);
                size_t l = strnlen(s, heap.size() - no);
                if (l == heap.size() - no) raise(USCT_H5_EFORMAT, "local heap: unterminated name");
                out.push_back({std::string(s, l), oh});
            }
        }
    }

    uint64_t resolve(const std::string &path) {
        {
            std::lock_guard<std::mutex> g(mu);
            auto it = paths.find(path);
            if (it != paths.end()) return it->second;
        }
        uint64_t cur = root;
        size_t i = 0;
        int depth = 0;
        while (i < path.size()) {
            while (i < path.size() && path[i] == '/') ++i;
            if (i >= path.size()) break;
            size_t j = path.find('/', i);
            if (j == std::string::npos) j = path.size();
            std::string comp = path.substr(i, j - i);
            i = j;
            if (++depth > MAX_DEPTH) raise(USCT_H5_EINVAL, "path too deep");
            auto ls = links(header(cur));
            uint64_t next = UNDEF;
            for (auto &l : ls)
                if (l.first == comp) next = l.second;
            if (next == UNDEF) return UNDEF;
            cur = next;
        }
        std::lock_guard<std::mutex> g(mu);
        paths[path] = cur;
        return cur;
    }

    uint64_t need(const std::string &path) {
        uint64_t a = resolve(path);
        if (a == UNDEF) raise(USCT_H5_ENOENT, "no object '" + path + "'");
        return a;
    }

    static int kind_of(const std::vector<Msg> &ms) {
        if (find(ms, 0x08)) return USCT_H5_DATASET;
        return USCT_H5_GROUP;
    }

    std::shared_ptr<Desc> desc(const std::string &path) {
        uint64_t a = need(path);
        {
            std::lock_guard<std::mutex> g(mu);
            auto it = descs.find(a);
            if (it != descs.end()) return it->second;
        }
        std::vector<Msg> ms = header(a);
        auto d = std::make_shared<Desc>();
        const Msg *sp = find(ms, 0x01), *ty = find(ms, 0x03), *lay = find(ms, 0x08);
        if (!lay) raise(USCT_H5_EINVAL, "'" + path + "' is not a dataset");
        if (!sp || !ty) raise(USCT_H5_EFORMAT, "dataset without dataspace or datatype");
        {
            Cur c(sp->data.data(), sp->data.size(), "dataspace");
            d->space = parse_space(c, sl);
        }
        {
            Cur c(ty->data.data(), ty->data.size(), "datatype");
            d->type = parse_dtype(c);
        }
        if (d->type.size == 0) raise(USCT_H5_EFORMAT, "datatype of size 0");
        parse_layout(*lay, *d);
        if (const Msg *f = find(ms, 0x0B)) parse_filters(*f, *d);
        if (const Msg *f = find(ms, 0x05)) {
            Cur c(f->data.data(), f->data.size(), "fill value");
            int ver = c.u8();
            if (ver == 1 || ver == 2) {
                c.skip(2);
                int defined = c.u8();
                if (ver == 1 || defined) {
                    uint64_t n = c.u(4);
                    const uint8_t *v = c.take(n);
                    if (n == d->type.size) d->fill.assign(v, v + n);
                }
            } else if (ver == 3) {
                int fl = c.u8();
                if (fl & 0x20) {
                    uint64_t n = c.u(4);
                    const uint8_t *v = c.take(n);
                    if (n == d->type.size) d->fill.assign(v, v + n);
                }
            }
        } else if (const Msg *f = find(ms, 0x04)) {
            Cur c(f->data.data(), f->data.size(), "fill value");
            uint64_t n = c.u(4);
            const uint8_t *v = c.take(n);
            if (n == d->type.size) d->fill.assign(v, v + n);
        }
        std::lock_guard<std::mutex> g(mu);
        descs[a] = d;
        return d;
    }

    void parse_layout(const Msg &m, Desc &d) const {
        Cur c(m.data.data(), m.data.size(), "data layout");
        int ver = c.u8();
        if (ver != 3 && ver != 4) raise(USCT_H5_ENOTSUP, "data layout message version " + std::to_string(ver));
        int cls = c.u8();
        d.layout = cls;
        size_t rank = d.space.dims.size();
        if (cls == 0) {
            uint64_t n = c.u(2);
            const uint8_t *p = c.take(n);
            d.compact.assign(p, p + n);
        } else if (cls == 1) {
            d.addr = off(c);
            d.csize = len(c);
        } else if (cls == 2) {
            std::vector<uint64_t> dims;
            if (ver == 3) {
                int nd = c.u8();
                d.addr = off(c);
                for (int i = 0; i < nd; ++i) dims.push_back(c.u(4));
                d.index = 0;
            } else {
                int fl = c.u8();
                int nd = c.u8();
                int enc = c.u8();
                if (enc < 1 || enc > 8) raise(USCT_H5_EFORMAT, "data layout: bad dimension encoding");
                for (int i = 0; i < nd; ++i) dims.push_back(c.u(enc));
                int idx = c.u8();
                if (idx == 1) {
                    d.index = 1;
                    if (fl & 2) {
                        d.single_filtered = true;
                        d.single_size = len(c);
                        d.single_mask = uint32_t(c.u(4));
                    }
                } else if (idx == 2) {
                    d.index = 2;
                } else {
                    raise(USCT_H5_ENOTSUP, "chunk index type " + std::to_string(idx) +
                                               " (fixed / extensible array, version-2 B-tree)");
                }
                d.addr = off(c);
            }
            if (dims.size() != rank + 1) raise(USCT_H5_EFORMAT, "data layout: chunk rank mismatch");
            for (size_t i = 0; i < rank; ++i) {
                if (dims[i] == 0) raise(USCT_H5_EFORMAT, "data layout: zero chunk extent");
                d.cdims.push_back(dims[i]);
            }
            if (dims[rank] != d.type.size) raise(USCT_H5_EFORMAT, "data layout: chunk element size mismatch");
            if (chunk_bytes(d) > 0xffffffffull) raise(USCT_H5_EFORMAT, "data layout: chunk above 4 GiB");
        } else {
            raise(USCT_H5_ENOTSUP, "virtual dataset layout");
        }
    }

    void parse_filters(const Msg &m, Desc &d) const {
        Cur c(m.data.data(), m.data.size(), "filter pipeline");
        int ver = c.u8();
        int n = c.u8();
        if (ver == 1) c.skip(6);
        else if (ver != 2) raise(USCT_H5_ENOTSUP, "filter pipeline version " + std::to_string(ver));
        for (int i = 0; i < n; ++i) {
            Filter f;
            f.id = int(c.u(2));
            uint64_t nl = (ver == 1 || f.id >= 256) ? c.u(2) : 0;
            c.u(2);  // flags
            uint64_t ncd = c.u(2);
            if (nl) c.skip(ver == 1 ? align8(nl) : nl);
            for (uint64_t k = 0; k < ncd; ++k) f.cd.push_back(uint32_t(c.u(4)));
            if (ver == 1 && (ncd & 1)) c.skip(4);
            if (f.id != 1 && f.id != 2 && f.id != 3)
                raise(USCT_H5_ENOTSUP, "filter " + std::to_string(f.id) +
                                           " (only deflate, shuffle and fletcher32 are read)");
            d.filters.push_back(f);
        }
    }

    // The chunk list of a dataset, built once; on failure nothing is kept, so a
    // later call starts afresh.  Every chunk offset must be distinct: a damaged
    // index could otherwise leave parts of the output unwritten.
    void load_chunks(Desc &d) const {
        if (d.chunks_loaded) return;
        std::vector<Chunk> list;
        std::swap(list, d.chunks);
        try {
            load_chunk_list(d);
        } catch (...) {
            d.chunks.clear();
            throw;
        }
        std::vector<const std::vector<uint64_t> *> offs;
        for (const Chunk &c : d.chunks) offs.push_back(&c.off);
        std::sort(offs.begin(), offs.end(), [](auto *a, auto *b) { return *a < *b; });
        for (size_t i = 1; i < offs.size(); ++i)
            if (*offs[i] == *offs[i - 1]) {
                d.chunks.clear();
                raise(USCT_H5_EFORMAT, "chunk index lists a chunk twice");
            }
        d.chunks_loaded = true;
    }

    void load_chunk_list(Desc &d) const {
        if (d.addr != UNDEF) {
            if (d.index == 0) {
                uint64_t visits = 0;
                chunk_btree(d.addr, d, -1, visits);
            } else if (d.index == 1) {
                Chunk ch{std::vector<uint64_t>(d.cdims.size(), 0), d.addr, 0, d.single_mask};
                ch.size = d.single_filtered ? d.single_size : chunk_bytes(d);
                d.chunks.push_back(ch);
            } else {
                // implicit index: every chunk stored unfiltered, in row-major chunk order
                size_t r = d.cdims.size();
                std::vector<uint64_t> nch(r);
                uint64_t total = 1;
                for (size_t i = 0; i < r; ++i) {
                    nch[i] = (d.space.dims[i] + d.cdims[i] - 1) / d.cdims[i];
                    if (__builtin_mul_overflow(total, nch[i], &total)) raise(USCT_H5_EFORMAT, "too many chunks");
                }
                uint64_t cb = chunk_bytes(d);
                for (uint64_t k = 0; k < total; ++k) {
                    Chunk ch{std::vector<uint64_t>(r), d.addr + k * cb, cb, 0};
                    uint64_t q = k;
                    for (size_t i = r; i-- > 0;) {
                        ch.off[i] = (q % nch[i]) * d.cdims[i];
                        q /= nch[i];
                    }
                    d.chunks.push_back(ch);
                }
            }
        }
    }

    static uint64_t chunk_bytes(const Desc &d) {
        uint64_t n = d.type.size;
        for (uint64_t c : d.cdims)
            if (__builtin_mul_overflow(n, c, &n)) raise(USCT_H5_EFORMAT, "chunk size overflow");
        return n;
    }

    void chunk_btree(uint64_t node, Desc &d, int expect, uint64_t &visits) const {
        if (++visits > MAX_NODES) raise(USCT_H5_EFORMAT, "chunk B-tree: too many nodes");
        size_t r = d.cdims.size();
        std::vector<uint8_t> hd = bytes(node, 8 + 2 * so);
        Cur h(hd.data(), hd.size(), "chunk B-tree node");
        if (std::memcmp(h.take(4), "TREE", 4) != 0) raise(USCT_H5_EFORMAT, "chunk B-tree: bad signature");
        if (h.u8() != 1) raise(USCT_H5_EFORMAT, "chunk B-tree: wrong node type");
        int level = h.u8();
        if ((expect >= 0 && level != expect) || level > MAX_DEPTH)
            raise(USCT_H5_EFORMAT, "chunk B-tree: inconsistent node levels");
        uint64_t n = h.u(2);
        uint64_t ksz = 8 + 8 * (r + 1);
        std::vector<uint8_t> body = bytes(node + 8 + 2 * so, n * so + (n + 1) * ksz);
        Cur c(body.data(), body.size(), "chunk B-tree node");
        for (uint64_t i = 0; i < n; ++i) {
            Chunk ch;
            ch.size = c.u(4);
            ch.mask = uint32_t(c.u(4));
            for (size_t k = 0; k <= r; ++k) {
                uint64_t o = c.u(8);
                if (k < r) ch.off.push_back(o);
            }
            ch.addr = off(c);
            if (level > 0) {
                chunk_btree(ch.addr, d, level - 1, visits);
                continue;
            }
            for (size_t k = 0; k < r; ++k)
                if (ch.off[k] % d.cdims[k] != 0 || ch.off[k] >= std::max<uint64_t>(d.space.dims[k], 1))
                    raise(USCT_H5_EFORMAT, "chunk B-tree: chunk offset outside the dataset");
            d.chunks.push_back(std::move(ch));
        }
    }

    std::vector<Attr> attrs(uint64_t oh) const {
        std::vector<Attr> out;
        std::vector<Msg> ms = header(oh);
        for (const Msg &m : ms) {
            if (m.type != 0x0C) continue;
            Cur c(m.data.data(), m.data.size(), "attribute message");
            int ver = c.u8();
            int fl = 0;
            if (ver == 1) c.skip(1);
            else if (ver == 2 || ver == 3) fl = c.u8();
            else raise(USCT_H5_ENOTSUP, "attribute message version " + std::to_string(ver));
            if (fl & 3) raise(USCT_H5_ENOTSUP, "attribute with a shared datatype or dataspace");
            uint64_t nl = c.u(2), tl = c.u(2), sln = c.u(2);
            if (ver == 3) c.skip(1);
            Attr a;
            const uint8_t *nm = c.take(nl);
            a.name.assign(reinterpret_cast<const char *>(nm), strnlen(reinterpret_cast<const char *>(nm), nl));
            if (ver == 1) c.skip(align8(nl) - nl);
            {
                const uint8_t *tp = c.take(tl);
                Cur t(tp, tl, "attribute datatype");
                a.type = parse_dtype(t);
            }
            if (ver == 1) c.skip(align8(tl) - tl);
            {
                const uint8_t *sp = c.take(sln);
                Cur s(sp, sln, "attribute dataspace");
                a.space = parse_space(s, sl);
            }
            if (ver == 1) c.skip(align8(sln) - sln);
            uint64_t n = a.space.count(), nb;
            if (__builtin_mul_overflow(n, uint64_t(a.type.size), &nb)) raise(USCT_H5_EFORMAT, "attribute size");
            const uint8_t *dp = c.take(nb);
            a.data.assign(dp, dp + nb);
            out.push_back(std::move(a));
        }
        return out;
    }

    Attr attr(const std::string &path, const std::string &name) {
        for (Attr &a : attrs(need(path)))
            if (a.name == name) return a;
        raise(USCT_H5_ENOENT, "no attribute '" + name + "' on '" + path + "'");
    }

    // Bytes of a variable-length element (length, collection address, index).
    std::string vlen_bytes(const uint8_t *p, const DType &t) const {
        Cur c(p, t.size, "variable-length element");
        uint64_t n = c.u(4);
        uint64_t ga = off(c);
        uint64_t idx = c.u(4);
        if (ga == UNDEF || ga == 0 || n == 0) return std::string();
        uint8_t hh[16];
        readat(ga, hh, 8 + sl);
        if (std::memcmp(hh, "GCOL", 4) != 0) raise(USCT_H5_EFORMAT, "global heap: bad signature");
        uint64_t csz = le(hh + 8, sl);
        std::vector<uint8_t> col = bytes(ga, csz);
        Cur g(col.data(), col.size(), "global heap");
        g.skip(8 + sl);
        while (g.left() >= 8 + uint64_t(sl)) {
            uint64_t id = g.u(2);
            g.skip(6);
            uint64_t osz = g.u(sl);
            if (id == 0) break;
            const uint8_t *d = g.take(osz);
            if (id == idx) {
                uint64_t elem = t.base ? t.base->size : 1;
                if (elem == 0 || n > osz / elem) raise(USCT_H5_EFORMAT, "global heap: object too small");
                return std::string(reinterpret_cast<const char *>(d), n * elem);
            }
            g.skip(std::min<uint64_t>(align8(osz) - osz, g.left()));
        }
        raise(USCT_H5_EFORMAT, "global heap: object not found");
    }

    // ---- dataset read -------------------------------------------------------

    void read(const std::string &path, int64_t row0, int64_t nrows, int code, void *out,
              int64_t out_elems, int nthreads) {
        auto dp = desc(path);
        Desc &d = *dp;
        if (code_size(code) == 0) raise(USCT_H5_EINVAL, "unsupported output type");
        if (d.type.cls != 0 && d.type.cls != 1)
            raise(USCT_H5_ENOTSUP, "dataset '" + path + "' is not numeric");
        size_t r = d.space.dims.size();
        uint64_t rows = r ? d.space.dims[0] : 1;
        if (d.space.null) rows = 0;
        if (row0 < 0 || uint64_t(row0) > rows) raise(USCT_H5_EINVAL, "row0 out of range");
        uint64_t nr = nrows < 0 ? rows - uint64_t(row0) : uint64_t(nrows);
        if (nr > rows - uint64_t(row0)) raise(USCT_H5_EINVAL, "rows out of range");
        uint64_t inner = 1;
        for (size_t i = 1; i < r; ++i)
            if (__builtin_mul_overflow(inner, d.space.dims[i], &inner)) raise(USCT_H5_EFORMAT, "size overflow");
        uint64_t want;
        if (__builtin_mul_overflow(nr, inner, &want)) raise(USCT_H5_EFORMAT, "size overflow");
        if (out_elems < 0 || uint64_t(out_elems) != want)
            raise(USCT_H5_EINVAL, "output holds " + std::to_string(out_elems) + " elements, the selection " +
                                      std::to_string(want));
        if (want == 0) return;
        uint8_t *o = static_cast<uint8_t *>(out);
        const int os = code_size(code);
        const uint64_t es = d.type.size;
        if (nthreads <= 0) nthreads = omp_get_max_threads();

        if (d.layout == 0 || d.layout == 1) {
            uint64_t total = d.space.count(), tb;
            if (__builtin_mul_overflow(total, es, &tb)) raise(USCT_H5_EFORMAT, "size overflow");
            if (d.layout == 0 && d.compact.size() < tb) raise(USCT_H5_EFORMAT, "compact data too short");
            if (d.layout == 1 && d.addr != UNDEF && d.csize < tb) raise(USCT_H5_EFORMAT, "contiguous data too short");
            uint64_t first = uint64_t(row0) * inner;
            if (d.layout == 0) {
                convert(d.compact.data() + first * es, d.type, code, o, want);
                return;
            }
            if (d.addr == UNDEF) {
                fill(d, code, o, want);
                return;
            }
            // Blocks of 8 MiB read and converted in parallel.
            const uint64_t blk = std::max<uint64_t>(1, (8ull << 20) / es);
            const uint64_t nb = (want + blk - 1) / blk;
            std::string err;
            int ecode = 0;
#pragma omp parallel num_threads(nthreads)
            {
                std::vector<uint8_t> buf;
#pragma omp for schedule(dynamic)
                for (uint64_t b = 0; b < nb; ++b) {
                    uint64_t k0 = b * blk, k1 = std::min(want, k0 + blk);
                    try {
                        if (type_code(d.type) == code && !d.type.be) {
                            readat(d.addr + (first + k0) * es, o + k0 * os, (k1 - k0) * es);
                        } else {
                            buf.resize((k1 - k0) * es);
                            readat(d.addr + (first + k0) * es, buf.data(), buf.size());
                            convert(buf.data(), d.type, code, o + k0 * os, k1 - k0);
                        }
                    } catch (const H5Err &e) {
#pragma omp critical(usct_h5_err)
                        if (!ecode) { ecode = e.code; err = e.msg; }
                    } catch (const std::bad_alloc &) {
#pragma omp critical(usct_h5_err)
                        if (!ecode) { ecode = USCT_H5_ENOMEM; err = "out of host memory"; }
                    }
                }
            }
            if (ecode) raise(ecode, err);
            return;
        }

        // Chunked.
        {
            std::lock_guard<std::mutex> g(mu);
            load_chunks(d);
        }
        const uint64_t cb = chunk_bytes(d);
        std::vector<uint64_t> nch(r);
        uint64_t expect = 1;
        for (size_t i = 0; i < r; ++i) {
            nch[i] = (d.space.dims[i] + d.cdims[i] - 1) / d.cdims[i];
            if (__builtin_mul_overflow(expect, nch[i], &expect)) raise(USCT_H5_EFORMAT, "too many chunks");
        }
        std::vector<const Chunk *> sel;
        for (const Chunk &ch : d.chunks) {
            if (r && (ch.off[0] >= uint64_t(row0) + nr || ch.off[0] + d.cdims[0] <= uint64_t(row0))) continue;
            sel.push_back(&ch);
        }
        if (d.chunks.size() < expect) fill(d, code, o, want);  // unwritten chunks read as the fill value
        std::string err;
        int ecode = 0;
#pragma omp parallel num_threads(nthreads)
        {
            std::vector<uint8_t> raw, dec, tmp;
#pragma omp for schedule(dynamic)
            for (size_t k = 0; k < sel.size(); ++k) {
                try {
                    const Chunk &ch = *sel[k];
                    if (ch.size > fsize) raise(USCT_H5_EFORMAT, "chunk larger than the file");
                    raw.resize(ch.size);
                    readat(ch.addr, raw.data(), ch.size);
                    const uint8_t *data = raw.data();
                    uint64_t n = ch.size;
                    for (size_t fi = d.filters.size(); fi-- > 0;) {
                        if (ch.mask & (1u << fi)) continue;
                        const Filter &f = d.filters[fi];
                        if (f.id == 1) {
                            // deflate expands at most 1032:1; a chunk that claims more is damaged
                            const uint64_t cap = cb + 16;
                            if (cap > n * 1100 + 4096) raise(USCT_H5_EFORMAT, "deflate: chunk cannot expand to its extent");
                            dec.resize(cap);
                            uLongf dl = uLongf(cap);
                            int zr = uncompress(dec.data(), &dl, data, uLong(n));
                            if (zr != Z_OK) raise(USCT_H5_EFORMAT, "deflate: corrupt chunk (zlib " + std::to_string(zr) + ")");
                            std::swap(raw, dec);
                            data = raw.data();
                            n = dl;
                        } else if (f.id == 2) {
                            uint64_t s = f.cd.empty() ? es : f.cd[0];
                            if (s > 1 && n >= s) {
                                tmp.resize(n);
                                uint64_t ne = n / s;
                                for (uint64_t j = 0; j < s; ++j)
                                    for (uint64_t e = 0; e < ne; ++e) tmp[e * s + j] = data[j * ne + e];
                                std::memcpy(tmp.data() + ne * s, data + ne * s, n - ne * s);
                                std::swap(raw, tmp);
                                data = raw.data();
                            }
                        } else if (f.id == 3) {
                            if (n < 4) raise(USCT_H5_EFORMAT, "fletcher32: chunk too short");
                            n -= 4;
                            uint32_t sum1 = 0, sum2 = 0;
                            uint64_t len = n / 2;
                            const uint8_t *q = data;
                            while (len) {
                                uint64_t tl = len > 360 ? 360 : len;
                                len -= tl;
                                do {
                                    sum1 += (uint32_t(q[0]) << 8) | q[1];
                                    q += 2;
                                    sum2 += sum1;
                                } while (--tl);
                                sum1 = (sum1 & 0xffff) + (sum1 >> 16);
                                sum2 = (sum2 & 0xffff) + (sum2 >> 16);
                            }
                            if (n & 1) {
                                sum1 += uint32_t(*q) << 8;
                                sum2 += sum1;
                                sum1 = (sum1 & 0xffff) + (sum1 >> 16);
                                sum2 = (sum2 & 0xffff) + (sum2 >> 16);
                            }
                            sum1 = (sum1 & 0xffff) + (sum1 >> 16);
                            sum2 = (sum2 & 0xffff) + (sum2 >> 16);
                            uint32_t want_sum = (sum2 << 16) | sum1;
                            uint32_t stored = uint32_t(le(data + n, 4));
                            if (stored != want_sum) raise(USCT_H5_EFORMAT, "fletcher32: checksum mismatch");
                        }
                    }
                    if (n < cb) raise(USCT_H5_EFORMAT, "chunk shorter than its extent");
                    scatter(d, ch, data, uint64_t(row0), nr, code, o);
                } catch (const H5Err &e) {
#pragma omp critical(usct_h5_err)
                    if (!ecode) { ecode = e.code; err = e.msg; }
                } catch (const std::bad_alloc &) {
#pragma omp critical(usct_h5_err)
                    if (!ecode) { ecode = USCT_H5_ENOMEM; err = "out of host memory"; }
                }
            }
        }
        if (ecode) raise(ecode, err);
    }

    static void fill(const Desc &d, int code, uint8_t *o, uint64_t n) {
        const int os = code_size(code);
        if (d.fill.empty()) {
            std::memset(o, 0, n * os);
            return;
        }
        uint8_t one[8];
        convert(d.fill.data(), d.type, code, one, 1);
        for (uint64_t k = 0; k < n; ++k) std::memcpy(o + k * os, one, os);
    }

    // Copy the part of one decoded chunk that lies inside the dataset and the
    // selected rows [row0, row0 + nr) into the output (shape nr x dims[1:]).
    static void scatter(const Desc &d, const Chunk &ch, const uint8_t *data, uint64_t row0, uint64_t nr,
                        int code, uint8_t *o) {
        const size_t r = d.cdims.size();
        const uint64_t es = d.type.size;
        const int os = code_size(code);
        if (r == 0) {
            convert(data, d.type, code, o, 1);
            return;
        }
        const std::vector<uint64_t> &D = d.space.dims, &C = d.cdims;
        bool whole = true;  // the chunk spans every trailing dimension: its rows are contiguous
        uint64_t inner = 1;
        for (size_t i = 1; i < r; ++i) {
            if (C[i] != D[i]) whole = false;
            inner *= D[i];
        }
        if (whole) {
            uint64_t a = std::max(ch.off[0], row0);
            uint64_t b = std::min(std::min(ch.off[0] + C[0], D[0]), row0 + nr);
            if (a < b)
                convert(data + (a - ch.off[0]) * inner * es, d.type, code, o + (a - row0) * inner * os,
                        (b - a) * inner);
            return;
        }
        const uint64_t run = std::min(C[r - 1], D[r - 1] - ch.off[r - 1]);
        std::vector<uint64_t> idx(r - 1, 0);  // position inside the chunk over the leading dims
        while (true) {
            bool inside = true;
            for (size_t i = 0; i + 1 < r; ++i)
                if (ch.off[i] + idx[i] >= D[i]) inside = false;
            const uint64_t g0 = ch.off[0] + idx[0];
            if (g0 < row0 || g0 >= row0 + nr) inside = false;
            if (inside) {
                uint64_t ci = 0, go = 0;
                for (size_t i = 0; i + 1 < r; ++i) {
                    ci = ci * C[i] + idx[i];
                    go = go * (i == 0 ? nr : D[i]) + (ch.off[i] + idx[i] - (i == 0 ? row0 : 0));
                }
                ci *= C[r - 1];
                go = go * D[r - 1] + ch.off[r - 1];
                convert(data + ci * es, d.type, code, o + go * os, run);
            }
            size_t i = r - 1;  // advance the leading-dims index, last one fastest
            while (i-- > 0) {
                if (++idx[i] < C[i]) break;
                idx[i] = 0;
            }
            if (i == size_t(-1)) break;
        }
    }
};

// ---------------------------------------------------------------------------
// Writer.

namespace {

struct WAttr {
    std::string name;
    int type;
    int64_t n;  // < 0: scalar
    std::vector<uint8_t> raw;
    std::string str;
    uint64_t gaddr = UNDEF;
    uint32_t gidx = 0;
};

struct WChunk {
    std::vector<uint64_t> off;
    uint64_t addr, size;
};

struct WNode {
    std::string name;
    bool group = true;
    int parent = -1;
    std::vector<int> kids;
    std::vector<WAttr> attrs;
    int type = 0;
    std::vector<uint64_t> dims, cdims;
    int level = -1;
    bool chunked = false;
    uint64_t addr = UNDEF, size = 0;
    std::vector<WChunk> chunks;
    uint64_t ohdr = UNDEF, btree = UNDEF, heap = UNDEF;
};

void put(std::vector<uint8_t> &b, uint64_t v, int n) {
    for (int i = 0; i < n; ++i) b.push_back(uint8_t(v >> (8 * i)));
}
void pad8(std::vector<uint8_t> &b) {
    while (b.size() % 8) b.push_back(0);
}

std::vector<uint8_t> enc_dtype(int code) {
    std::vector<uint8_t> b;
    switch (code) {
    case USCT_H5_F32:
    case USCT_H5_F64: {
        bool d = code == USCT_H5_F64;
        put(b, 0x11, 1);
        put(b, 0x20, 1);
        put(b, d ? 63 : 31, 1);
        put(b, 0, 1);
        put(b, d ? 8 : 4, 4);
        put(b, 0, 2);
        put(b, d ? 64 : 32, 2);
        put(b, d ? 52 : 23, 1);
        put(b, d ? 11 : 8, 1);
        put(b, 0, 1);
        put(b, d ? 52 : 23, 1);
        put(b, d ? 1023 : 127, 4);
        break;
    }
    case USCT_H5_I32:
    case USCT_H5_I64:
    case USCT_H5_U8: {
        int sz = code_size(code);
        put(b, 0x10, 1);
        put(b, code == USCT_H5_U8 ? 0x00 : 0x08, 1);
        put(b, 0, 2);
        put(b, sz, 4);
        put(b, 0, 2);
        put(b, 8 * sz, 2);
        break;
    }
    case USCT_H5_STR: {
        // variable-length UTF-8 string of unsigned bytes, null-terminated padding
        put(b, 0x19, 1);
        put(b, 0x01, 1);
        put(b, 0x01, 1);
        put(b, 0, 1);
        put(b, 16, 4);
        std::vector<uint8_t> base = enc_dtype(USCT_H5_U8);
        b.insert(b.end(), base.begin(), base.end());
        break;
    }
    }
    return b;
}

std::vector<uint8_t> enc_space(const std::vector<uint64_t> &dims, bool scalar) {
    std::vector<uint8_t> b;
    put(b, 1, 1);
    put(b, scalar ? 0 : dims.size(), 1);
    put(b, scalar ? 0 : 1, 1);  // maximum dimensions present (equal to the dimensions)
    put(b, 0, 5);
    if (!scalar) {
        for (uint64_t d : dims) put(b, d, 8);
        for (uint64_t d : dims) put(b, d, 8);
    }
    return b;
}

void add_msg(std::vector<uint8_t> &h, int type, const std::vector<uint8_t> &data, int flags, int &count) {
    uint64_t n = align8(data.size());
    if (n > 0xffff) raise(USCT_H5_EINVAL, "object header message above 64 KiB (attribute too large)");
    put(h, type, 2);
    put(h, n, 2);
    put(h, flags, 1);
    put(h, 0, 3);
    h.insert(h.end(), data.begin(), data.end());
    pad8(h);
    ++count;
}

}  // namespace

struct usct_h5_writer {
    int fd = -1;
    std::string path;
    uint64_t eof = 96;  // the version-0 superblock occupies [0, 96)
    std::vector<WNode> nodes;

    usct_h5_writer() {
        WNode root;
        root.name = "";
        nodes.push_back(root);
    }
    ~usct_h5_writer() {
        if (fd >= 0) ::close(fd);
    }

    void writeat(uint64_t at, const void *p, uint64_t n) {
        uint64_t done = 0;
        while (done < n) {
            ssize_t w = ::pwrite(fd, static_cast<const char *>(p) + done, size_t(n - done), off_t(at + done));
            if (w < 0 && errno == EINTR) continue;
            if (w <= 0) raise(USCT_H5_EIO, std::string("write failed: ") + std::strerror(errno));
            done += uint64_t(w);
        }
    }
    uint64_t append(const void *p, uint64_t n) {
        uint64_t at = eof;
        writeat(at, p, n);
        eof += n;
        return at;
    }
    uint64_t append(const std::vector<uint8_t> &b) { return append(b.data(), b.size()); }

    int lookup(const std::string &path) const {
        int cur = 0;
        size_t i = 0;
        while (i < path.size()) {
            while (i < path.size() && path[i] == '/') ++i;
            if (i >= path.size()) break;
            size_t j = path.find('/', i);
            if (j == std::string::npos) j = path.size();
            std::string comp = path.substr(i, j - i);
            i = j;
            int nx = -1;
            for (int k : nodes[size_t(cur)].kids)
                if (nodes[size_t(k)].name == comp) nx = k;
            if (nx < 0) return -1;
            cur = nx;
        }
        return cur;
    }

    int new_node(const std::string &path, bool group) {
        if (path.empty() || path[0] != '/') raise(USCT_H5_EINVAL, "paths are absolute ('/name')");
        std::string p = path;
        while (p.size() > 1 && p.back() == '/') p.pop_back();
        size_t s = p.rfind('/');
        std::string parent = p.substr(0, s), name = p.substr(s + 1);
        if (name.empty()) raise(USCT_H5_EINVAL, "empty object name");
        if (lookup(p) >= 0) raise(USCT_H5_EINVAL, "'" + p + "' exists");
        int pi = lookup(parent.empty() ? "/" : parent);
        if (pi < 0) raise(USCT_H5_ENOENT, "parent group of '" + p + "' does not exist");
        if (!nodes[size_t(pi)].group) raise(USCT_H5_EINVAL, "parent of '" + p + "' is a dataset");
        WNode n;
        n.name = name;
        n.group = group;
        n.parent = pi;
        nodes.push_back(n);
        int id = int(nodes.size() - 1);
        nodes[size_t(pi)].kids.push_back(id);
        return id;
    }

    void dataset(const std::string &path, int type, int ndim, const int64_t *dims, const void *data,
                 const int64_t *chunk, int level, int nthreads) {
        if (type != USCT_H5_F32 && type != USCT_H5_F64 && type != USCT_H5_I32 && type != USCT_H5_I64 &&
            type != USCT_H5_U8)
            raise(USCT_H5_EINVAL, "dataset type must be F32, F64, I32, I64 or U8");
        if (ndim < 0 || ndim > USCT_H5_MAX_RANK) raise(USCT_H5_EINVAL, "rank out of range");
        if (level < -1 || level > 9) raise(USCT_H5_EINVAL, "gzip level must be -1..9");
        std::vector<uint64_t> D;
        uint64_t n = 1;
        for (int i = 0; i < ndim; ++i) {
            if (dims[i] < 0) raise(USCT_H5_EINVAL, "negative dimension");
            D.push_back(uint64_t(dims[i]));
            if (__builtin_mul_overflow(n, D.back(), &n)) raise(USCT_H5_EINVAL, "size overflow");
        }
        if (n && !data) raise(USCT_H5_EINVAL, "data is NULL");
        const uint64_t es = uint64_t(code_size(type));
        int id = new_node(path, false);
        WNode &w = nodes[size_t(id)];
        w.type = type;
        w.dims = D;
        if (!chunk || ndim == 0 || n == 0) {
            w.size = n * es;
            if (n) w.addr = append(data, n * es);
            return;
        }
        w.chunked = true;
        w.level = level;
        uint64_t cbytes = es;
        for (int i = 0; i < ndim; ++i) {
            if (chunk[i] <= 0) raise(USCT_H5_EINVAL, "chunk extents must be positive");
            w.cdims.push_back(uint64_t(std::min<int64_t>(chunk[i], std::max<int64_t>(dims[i], 1))));
            if (__builtin_mul_overflow(cbytes, w.cdims.back(), &cbytes)) raise(USCT_H5_EINVAL, "chunk too large");
        }
        if (cbytes > 0xffffffffull) raise(USCT_H5_EINVAL, "chunk above 4 GiB");
        std::vector<uint64_t> nch(static_cast<size_t>(ndim));
        uint64_t total = 1;
        for (int i = 0; i < ndim; ++i) {
            nch[size_t(i)] = (D[size_t(i)] + w.cdims[size_t(i)] - 1) / w.cdims[size_t(i)];
            total *= nch[size_t(i)];
        }
        if (nthreads <= 0) nthreads = omp_get_max_threads();
        const uint8_t *src = static_cast<const uint8_t *>(data);
        const uint64_t batch = uint64_t(std::max(1, nthreads)) * 4;
        std::vector<std::vector<uint8_t>> out(size_t(std::min(batch, total)));
        std::vector<std::vector<uint64_t>> offs(out.size());
        for (uint64_t c0 = 0; c0 < total; c0 += batch) {
            uint64_t c1 = std::min(total, c0 + batch);
            std::string err;
            int ecode = 0;
#pragma omp parallel num_threads(nthreads)
            {
                std::vector<uint8_t> buf;
#pragma omp for schedule(dynamic)
                for (uint64_t c = c0; c < c1; ++c) {
                  try {
                    std::vector<uint64_t> off(static_cast<size_t>(ndim));
                    uint64_t q = c;
                    for (int i = ndim; i-- > 0;) {
                        off[size_t(i)] = (q % nch[size_t(i)]) * w.cdims[size_t(i)];
                        q /= nch[size_t(i)];
                    }
                    // gather: full chunk extent, zero beyond the dataset's edge
                    buf.assign(cbytes, 0);
                    const size_t r = size_t(ndim);
                    std::vector<uint64_t> idx(r - 1, 0);
                    uint64_t run = std::min(w.cdims[r - 1], D[r - 1] - off[r - 1]);
                    while (true) {
                        bool inside = true;
                        for (size_t i = 0; i + 1 < r; ++i)
                            if (off[i] + idx[i] >= D[i]) inside = false;
                        if (inside) {
                            uint64_t ci = 0, gi = 0;
                            for (size_t i = 0; i + 1 < r; ++i) {
                                ci = ci * w.cdims[i] + idx[i];
                                gi = gi * D[i] + off[i] + idx[i];
                            }
                            ci *= w.cdims[r - 1];
                            gi = gi * D[r - 1] + off[r - 1];
                            std::memcpy(buf.data() + ci * es, src + gi * es, run * es);
                        }
                        size_t i = r - 1;
                        while (i-- > 0) {
                            if (++idx[i] < w.cdims[i]) break;
                            idx[i] = 0;
                        }
                        if (i == size_t(-1)) break;
                    }
                    std::vector<uint8_t> &o = out[size_t(c - c0)];
                    if (level < 0) {
                        o = buf;
                    } else {
                        uLongf dl = compressBound(uLong(cbytes));
                        o.resize(dl);
                        int zr = compress2(o.data(), &dl, buf.data(), uLong(cbytes), level);
                        if (zr != Z_OK) {
#pragma omp critical(usct_h5_err)
                            if (!ecode) { ecode = USCT_H5_ENOMEM; err = "deflate failed"; }
                        }
                        o.resize(dl);
                    }
                    offs[size_t(c - c0)] = off;
                  } catch (const std::bad_alloc &) {
#pragma omp critical(usct_h5_err)
                    if (!ecode) { ecode = USCT_H5_ENOMEM; err = "out of host memory"; }
                  }
                }
            }
            if (ecode) raise(ecode, err);
            for (uint64_t c = c0; c < c1; ++c) {
                std::vector<uint8_t> &o = out[size_t(c - c0)];
                WChunk ch{offs[size_t(c - c0)], append(o), o.size()};
                nodes[size_t(id)].chunks.push_back(ch);
                std::vector<uint8_t>().swap(o);
            }
        }
    }

    void attribute(const std::string &path, const std::string &name, int type, int64_t n, const void *v) {
        int id = lookup(path);
        if (id < 0) raise(USCT_H5_ENOENT, "no object '" + path + "'");
        if (name.empty()) raise(USCT_H5_EINVAL, "empty attribute name");
        for (const WAttr &a : nodes[size_t(id)].attrs)
            if (a.name == name) raise(USCT_H5_EINVAL, "attribute '" + name + "' exists");
        WAttr a;
        a.name = name;
        a.type = type;
        a.n = n;
        if (type == USCT_H5_STR) {
            if (!v) raise(USCT_H5_EINVAL, "string is NULL");
            a.str = static_cast<const char *>(v);
            a.n = -1;
        } else if (type == USCT_H5_F32 || type == USCT_H5_F64 || type == USCT_H5_I32 || type == USCT_H5_I64 ||
                   type == USCT_H5_U8) {
            uint64_t cnt = n < 0 ? 1 : uint64_t(n);
            if (cnt && !v) raise(USCT_H5_EINVAL, "value is NULL");
            const uint8_t *p = static_cast<const uint8_t *>(v);
            a.raw.assign(p, p + cnt * uint64_t(code_size(type)));
        } else {
            raise(USCT_H5_EINVAL, "attribute type must be F32, F64, I32, I64, U8 or STR");
        }
        nodes[size_t(id)].attrs.push_back(a);
    }

    // Version-1 B-tree over a dataset's chunks (node type 1, 2K = 64 entries).
    uint64_t chunk_tree(const WNode &w) {
        const size_t r = w.dims.size();
        const uint64_t K2 = 64, ksz = 8 + 8 * (r + 1);
        const uint64_t node_size = 24 + K2 * 8 + (K2 + 1) * ksz;
        struct Ent {
            uint64_t size;
            std::vector<uint64_t> off;
            uint64_t child;
        };
        std::vector<uint64_t> right(r);
        const WChunk &last = w.chunks.back();
        for (size_t i = 0; i < r; ++i) right[i] = last.off[i] + w.cdims[i];
        std::vector<Ent> level;
        for (const WChunk &c : w.chunks) level.push_back({c.size, c.off, c.addr});
        int lv = 0;
        while (true) {
            std::vector<Ent> up;
            for (size_t s = 0; s < level.size(); s += K2) {
                size_t e = std::min(level.size(), s + K2);
                std::vector<uint8_t> b;
                b.insert(b.end(), {'T', 'R', 'E', 'E'});
                put(b, 1, 1);
                put(b, uint64_t(lv), 1);
                put(b, e - s, 2);
                put(b, UNDEF, 8);
                put(b, UNDEF, 8);
                for (size_t k = s; k < e; ++k) {
                    put(b, lv ? 0 : level[k].size, 4);
                    put(b, 0, 4);
                    for (uint64_t o : level[k].off) put(b, o, 8);
                    put(b, 0, 8);
                    put(b, level[k].child, 8);
                }
                put(b, 0, 8);  // right key: one past the last chunk
                for (uint64_t o : right) put(b, o, 8);
                put(b, 0, 8);
                b.resize(node_size, 0);
                uint64_t at = append(b);
                up.push_back({0, level[s].off, at});
            }
            if (up.size() == 1) return up[0].child;
            level.swap(up);
            ++lv;
        }
    }

    void group_meta(WNode &g) {
        // local heap: "" at 0, then the member names (null-terminated, 8-aligned)
        std::vector<int> kids = g.kids;
        std::sort(kids.begin(), kids.end(),
                  [&](int a, int b) { return std::strcmp(nodes[size_t(a)].name.c_str(), nodes[size_t(b)].name.c_str()) < 0; });
        std::vector<uint8_t> heap(8, 0);
        std::vector<uint64_t> noff;
        for (int k : kids) {
            noff.push_back(heap.size());
            const std::string &s = nodes[size_t(k)].name;
            heap.insert(heap.end(), s.begin(), s.end());
            heap.push_back(0);
            pad8(heap);
        }
        std::vector<uint8_t> hh;
        hh.insert(hh.end(), {'H', 'E', 'A', 'P'});
        put(hh, 0, 4);
        put(hh, heap.size(), 8);
        put(hh, 1, 8);  // no free block
        put(hh, eof + 32, 8);
        g.heap = append(hh);
        append(heap);
        // symbol nodes of 2 * 4 entries
        const size_t per = 8;
        std::vector<uint64_t> snods, lastname;
        for (size_t s = 0; s < kids.size(); s += per) {
            size_t e = std::min(kids.size(), s + per);
            std::vector<uint8_t> b;
            b.insert(b.end(), {'S', 'N', 'O', 'D'});
            put(b, 1, 1);
            put(b, 0, 1);
            put(b, e - s, 2);
            for (size_t k = s; k < e; ++k) {
                const WNode &c = nodes[size_t(kids[k])];
                put(b, noff[k], 8);
                put(b, c.ohdr, 8);
                put(b, c.group ? 1 : 0, 4);
                put(b, 0, 4);
                put(b, c.group ? c.btree : 0, 8);
                put(b, c.group ? c.heap : 0, 8);
            }
            b.resize(8 + per * 40, 0);
            snods.push_back(append(b));
            lastname.push_back(noff[e - 1]);
        }
        if (snods.size() > 32) raise(USCT_H5_ENOTSUP, "group with more than 256 members");
        std::vector<uint8_t> t;
        t.insert(t.end(), {'T', 'R', 'E', 'E'});
        put(t, 0, 1);
        put(t, 0, 1);
        put(t, snods.size(), 2);
        put(t, UNDEF, 8);
        put(t, UNDEF, 8);
        put(t, 0, 8);
        for (size_t i = 0; i < snods.size(); ++i) {
            put(t, snods[i], 8);
            put(t, lastname[i], 8);
        }
        t.resize(24 + 32 * 8 + 33 * 8, 0);
        g.btree = append(t);
    }

    void object_header(WNode &w) {
        std::vector<uint8_t> h;
        int count = 0;
        if (w.group) {
            std::vector<uint8_t> st;
            put(st, w.btree, 8);
            put(st, w.heap, 8);
            add_msg(h, 0x11, st, 0, count);
        } else {
            add_msg(h, 0x01, enc_space(w.dims, w.dims.empty()), 0, count);
            add_msg(h, 0x03, enc_dtype(w.type), 1, count);
            std::vector<uint8_t> fv;
            put(fv, 2, 1);
            put(fv, w.chunked ? 3 : 2, 1);  // allocation: incremental (chunked) / late
            put(fv, 2, 1);                  // fill written if set
            put(fv, 0, 1);                  // no fill value defined: zeros
            add_msg(h, 0x05, fv, 1, count);
            std::vector<uint8_t> lay;
            put(lay, 3, 1);
            if (w.chunked) {
                put(lay, 2, 1);
                put(lay, w.dims.size() + 1, 1);
                put(lay, w.btree, 8);
                for (uint64_t c : w.cdims) put(lay, c, 4);
                put(lay, uint64_t(code_size(w.type)), 4);
            } else {
                put(lay, 1, 1);
                put(lay, w.addr, 8);
                put(lay, w.size, 8);
            }
            add_msg(h, 0x08, lay, 0, count);
            if (w.chunked && w.level >= 0) {
                std::vector<uint8_t> fp;
                put(fp, 1, 1);
                put(fp, 1, 1);
                put(fp, 0, 6);
                put(fp, 1, 2);  // deflate
                put(fp, 8, 2);
                put(fp, 1, 2);  // optional, as H5Pset_deflate sets it
                put(fp, 1, 2);
                const char nm[8] = {'d', 'e', 'f', 'l', 'a', 't', 'e', 0};
                fp.insert(fp.end(), nm, nm + 8);
                put(fp, uint64_t(w.level), 4);
                put(fp, 0, 4);
                add_msg(h, 0x0B, fp, 1, count);
            }
        }
        for (const WAttr &a : w.attrs) {
            std::vector<uint8_t> m, t = enc_dtype(a.type),
                                    s = enc_space({uint64_t(std::max<int64_t>(a.n, 0))}, a.n < 0);
            uint64_t nl = a.name.size() + 1;
            put(m, 1, 1);
            put(m, 0, 1);
            put(m, nl, 2);
            put(m, t.size(), 2);
            put(m, s.size(), 2);
            m.insert(m.end(), a.name.begin(), a.name.end());
            m.push_back(0);
            pad8(m);
            m.insert(m.end(), t.begin(), t.end());
            pad8(m);
            m.insert(m.end(), s.begin(), s.end());
            pad8(m);
            if (a.type == USCT_H5_STR) {
                put(m, a.str.size(), 4);
                put(m, a.gaddr, 8);
                put(m, a.gidx, 4);
            } else {
                m.insert(m.end(), a.raw.begin(), a.raw.end());
            }
            add_msg(h, 0x0C, m, 0, count);
        }
        std::vector<uint8_t> pre;
        put(pre, 1, 1);
        put(pre, 0, 1);
        put(pre, uint64_t(count), 2);
        put(pre, 1, 4);
        put(pre, h.size(), 4);
        put(pre, 0, 4);
        pre.insert(pre.end(), h.begin(), h.end());
        w.ohdr = append(pre);
    }

    void finish() {
        // one global heap collection holds every string attribute
        std::vector<WAttr *> strs;
        for (WNode &n : nodes)
            for (WAttr &a : n.attrs)
                if (a.type == USCT_H5_STR) strs.push_back(&a);
        if (!strs.empty()) {
            std::vector<uint8_t> g;
            g.insert(g.end(), {'G', 'C', 'O', 'L'});
            put(g, 1, 1);
            put(g, 0, 3);
            put(g, 0, 8);  // size, set below
            uint64_t at = eof;
            uint32_t idx = 0;
            for (WAttr *a : strs) {
                a->gaddr = at;
                a->gidx = ++idx;
                if (idx > 0xffff) raise(USCT_H5_ENOTSUP, "more than 65535 string attributes");
                put(g, idx, 2);
                put(g, 0, 2);
                put(g, 0, 4);
                put(g, a->str.size(), 8);
                g.insert(g.end(), a->str.begin(), a->str.end());
                pad8(g);
            }
            uint64_t total = std::max<uint64_t>(4096, align8(g.size() + 16));
            uint64_t freeb = total - g.size();
            put(g, 0, 2);
            put(g, 0, 2);
            put(g, 0, 4);
            put(g, freeb, 8);
            g.resize(total, 0);
            for (int i = 0; i < 8; ++i) g[8 + size_t(i)] = uint8_t(total >> (8 * i));
            append(g);
        }
        for (WNode &n : nodes)
            if (!n.group && n.chunked && !n.chunks.empty()) n.btree = chunk_tree(n);
        // post-order: members before the groups that list them
        std::vector<int> order;
        std::vector<std::pair<int, size_t>> st{{0, 0}};
        while (!st.empty()) {
            auto &[id, k] = st.back();
            if (k < nodes[size_t(id)].kids.size()) {
                int c = nodes[size_t(id)].kids[k++];
                st.push_back({c, 0});
            } else {
                order.push_back(id);
                st.pop_back();
            }
        }
        for (int id : order) {
            WNode &n = nodes[size_t(id)];
            if (n.group) group_meta(n);
            object_header(n);
        }
        const WNode &root = nodes[0];
        std::vector<uint8_t> sb = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n', 0, 0, 0, 0, 0, 8, 8, 0};
        put(sb, 4, 2);
        put(sb, 16, 2);
        put(sb, 0, 4);
        put(sb, 0, 8);
        put(sb, UNDEF, 8);
        put(sb, eof, 8);
        put(sb, UNDEF, 8);
        put(sb, 0, 8);
        put(sb, root.ohdr, 8);
        put(sb, 1, 4);
        put(sb, 0, 4);
        put(sb, root.btree, 8);
        put(sb, root.heap, 8);
        writeat(0, sb.data(), sb.size());
    }
};

// ---------------------------------------------------------------------------
// C ABI.

namespace {

template <class F>
int guard(F &&f) {
    try {
        f();
        return USCT_H5_OK;
    } catch (const H5Err &e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "out of host memory";
        return USCT_H5_ENOMEM;
    } catch (const std::exception &e) {
        g_err = e.what();
        return USCT_H5_EINVAL;
    }
}

void names_out(const std::vector<std::string> &names, char *buf, int64_t cap, int64_t *needed) {
    int64_t n = 0;
    for (const std::string &s : names) n += int64_t(s.size()) + 1;
    if (needed) *needed = n;
    if (!buf) return;
    if (cap < n) raise(USCT_H5_EINVAL, "name buffer too small");
    for (const std::string &s : names) {
        std::memcpy(buf, s.data(), s.size());
        buf += s.size();
        *buf++ = 0;
    }
}

}  // namespace

extern "C" {

int usct_h5_open(usct_h5_file **file, const char *path) {
    return guard([&] {
        if (!file || !path) raise(USCT_H5_EINVAL, "NULL argument");
        *file = nullptr;
        auto f = std::make_unique<usct_h5_file>();
        f->fd = ::open(path, O_RDONLY | O_CLOEXEC);
        if (f->fd < 0)
            raise(errno == ENOENT ? USCT_H5_ENOENT : USCT_H5_EIO,
                  std::string("cannot open '") + path + "': " + std::strerror(errno));
        struct stat st;
        if (fstat(f->fd, &st) != 0) raise(USCT_H5_EIO, std::strerror(errno));
        f->fsize = uint64_t(st.st_size);
        f->open_super();
        *file = f.release();
    });
}

int usct_h5_close(usct_h5_file *file) {
    delete file;
    return USCT_H5_OK;
}

int usct_h5_kind(usct_h5_file *file, const char *path, int *kind) {
    return guard([&] {
        if (!file || !path || !kind) raise(USCT_H5_EINVAL, "NULL argument");
        uint64_t a = file->resolve(path);
        *kind = a == UNDEF ? USCT_H5_ABSENT : usct_h5_file::kind_of(file->header(a));
    });
}

int usct_h5_list(usct_h5_file *file, const char *path, char *buf, int64_t cap, int64_t *needed) {
    return guard([&] {
        if (!file || !path) raise(USCT_H5_EINVAL, "NULL argument");
        std::vector<Msg> ms = file->header(file->need(path));
        if (usct_h5_file::kind_of(ms) != USCT_H5_GROUP) raise(USCT_H5_EINVAL, "'" + std::string(path) + "' is not a group");
        std::vector<std::string> names;
        for (auto &l : file->links(ms)) names.push_back(l.first);
        names_out(names, buf, cap, needed);
    });
}

int usct_h5_dataset_info(usct_h5_file *file, const char *path, int *ndim, int64_t *dims, int *type,
                         int *chunked) {
    return guard([&] {
        if (!file || !path || !ndim || !dims || !type) raise(USCT_H5_EINVAL, "NULL argument");
        auto d = file->desc(path);
        *ndim = int(d->space.dims.size());
        for (size_t i = 0; i < d->space.dims.size(); ++i) dims[i] = int64_t(d->space.dims[i]);
        *type = type_code(d->type);
        if (chunked) *chunked = d->layout == 2;
    });
}

int usct_h5_read(usct_h5_file *file, const char *path, int64_t row0, int64_t nrows, int out_type, void *out,
                 int64_t out_elems, int nthreads) {
    return guard([&] {
        if (!file || !path || (!out && out_elems)) raise(USCT_H5_EINVAL, "NULL argument");
        file->read(path, row0, nrows, out_type, out, out_elems, nthreads);
    });
}

int usct_h5_attr_list(usct_h5_file *file, const char *path, char *buf, int64_t cap, int64_t *needed) {
    return guard([&] {
        if (!file || !path) raise(USCT_H5_EINVAL, "NULL argument");
        std::vector<std::string> names;
        for (const Attr &a : file->attrs(file->need(path))) names.push_back(a.name);
        names_out(names, buf, cap, needed);
    });
}

int usct_h5_attr_info(usct_h5_file *file, const char *path, const char *name, int *type, int64_t *n,
                      int *rank) {
    return guard([&] {
        if (!file || !path || !name || !type || !n) raise(USCT_H5_EINVAL, "NULL argument");
        Attr a = file->attr(path, name);
        *type = type_code(a.type);
        if (rank) *rank = int(a.space.dims.size());
        if (*type == USCT_H5_STR) {
            if (a.space.count() == 0) { *n = 0; return; }
            if (a.type.cls == 3) {
                *n = int64_t(strnlen(reinterpret_cast<const char *>(a.data.data()), a.type.size));
            } else {
                *n = int64_t(file->vlen_bytes(a.data.data(), a.type).size());
            }
        } else {
            *n = int64_t(a.space.count());
        }
    });
}

int usct_h5_attr_read(usct_h5_file *file, const char *path, const char *name, int out_type, void *out,
                      int64_t cap) {
    return guard([&] {
        if (!file || !path || !name || !out) raise(USCT_H5_EINVAL, "NULL argument");
        Attr a = file->attr(path, name);
        int tc = type_code(a.type);
        if (out_type == USCT_H5_STR) {
            if (tc != USCT_H5_STR) raise(USCT_H5_EINVAL, "attribute '" + std::string(name) + "' is not a string");
            std::string s;
            if (a.space.count() > 0) {
                if (a.type.cls == 3) {
                    s.assign(reinterpret_cast<const char *>(a.data.data()),
                             strnlen(reinterpret_cast<const char *>(a.data.data()), a.type.size));
                } else {
                    s = file->vlen_bytes(a.data.data(), a.type);
                }
            }
            if (cap < int64_t(s.size()) + 1) raise(USCT_H5_EINVAL, "string buffer too small");
            std::memcpy(out, s.data(), s.size());
            static_cast<char *>(out)[s.size()] = 0;
            return;
        }
        if (code_size(out_type) == 0) raise(USCT_H5_EINVAL, "unsupported output type");
        if (a.type.cls != 0 && a.type.cls != 1)
            raise(USCT_H5_EINVAL, "attribute '" + std::string(name) + "' is not numeric");
        uint64_t n = a.space.count();
        if (cap < int64_t(n)) raise(USCT_H5_EINVAL, "attribute buffer too small");
        convert(a.data.data(), a.type, out_type, static_cast<uint8_t *>(out), n);
    });
}

int usct_h5_create(usct_h5_writer **writer, const char *path) {
    return guard([&] {
        if (!writer || !path) raise(USCT_H5_EINVAL, "NULL argument");
        *writer = nullptr;
        auto w = std::make_unique<usct_h5_writer>();
        w->path = path;
        w->fd = ::open(path, O_RDWR | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
        if (w->fd < 0) raise(USCT_H5_EIO, std::string("cannot create '") + path + "': " + std::strerror(errno));
        std::vector<uint8_t> zero(96, 0);
        w->writeat(0, zero.data(), zero.size());
        *writer = w.release();
    });
}

int usct_h5_add_group(usct_h5_writer *writer, const char *path) {
    return guard([&] {
        if (!writer || !path) raise(USCT_H5_EINVAL, "NULL argument");
        writer->new_node(path, true);
    });
}

int usct_h5_add_dataset(usct_h5_writer *writer, const char *path, int type, int ndim, const int64_t *dims,
                        const void *data, const int64_t *chunk, int gzip_level, int nthreads) {
    return guard([&] {
        if (!writer || !path || (ndim > 0 && !dims)) raise(USCT_H5_EINVAL, "NULL argument");
        writer->dataset(path, type, ndim, dims, data, chunk, gzip_level, nthreads);
    });
}

int usct_h5_add_attr(usct_h5_writer *writer, const char *path, const char *name, int type, int64_t n,
                     const void *value) {
    return guard([&] {
        if (!writer || !path || !name) raise(USCT_H5_EINVAL, "NULL argument");
        writer->attribute(path, name, type, n, value);
    });
}

int usct_h5_finish(usct_h5_writer *writer, int commit) {
    if (!writer) return USCT_H5_OK;
    std::unique_ptr<usct_h5_writer> w(writer);
    int rc = USCT_H5_OK;
    if (commit) rc = guard([&] { w->finish(); });
    if (w->fd >= 0) {
        if (::close(w->fd) != 0 && rc == USCT_H5_OK) {
            g_err = std::string("close failed: ") + std::strerror(errno);
            rc = USCT_H5_EIO;
        }
        w->fd = -1;
    }
    if (!commit || rc != USCT_H5_OK) ::unlink(w->path.c_str());
    return rc;
}

const char *usct_h5_last_error(void) { return g_err.c_str(); }

}  // extern "C"
