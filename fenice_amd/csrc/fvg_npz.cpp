// .npz containers: see fvg_npz.hpp. The zip records follow the PKWARE application note (local file header,
// central directory, end of central directory; zip64 extra field 0x0001, zip64 end record and locator); the .npy
// entries follow numpy's format description (magic, version, little-endian header length, a Python-literal
// dictionary with descr, fortran_order and shape, padded with spaces to a multiple of 64 bytes).
#include "fvg_npz.hpp"

#include <array>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>

namespace fvg {
namespace npz {

namespace {

constexpr uint32_t kLocalSig = 0x04034b50u, kCentralSig = 0x02014b50u, kEndSig = 0x06054b50u;
constexpr uint32_t kEnd64Sig = 0x06064b50u, kLocator64Sig = 0x07064b50u;
constexpr uint16_t kDate19800101 = (0 << 9) | (1 << 5) | 1;   // MS-DOS date: years since 1980, month, day

struct Out {   // little-endian byte sink
    std::string b;
    void u16(uint32_t v) {
        b += char(v & 0xFF);
        b += char((v >> 8) & 0xFF);
    }
    void u32(uint32_t v) {
        u16(v & 0xFFFF);
        u16(v >> 16);
    }
    void u64(uint64_t v) {
        u32(uint32_t(v));
        u32(uint32_t(v >> 32));
    }
};

uint64_t le(const uint8_t* p, int n) {
    uint64_t v = 0;
    for (int k = n - 1; k >= 0; --k) v = (v << 8) | p[k];
    return v;
}

std::string shape_text(const std::vector<int64_t>& s) {
    std::string t = "(";
    for (size_t k = 0; k < s.size(); ++k) {
        if (k) t += ", ";
        t += std::to_string(s[k]);
    }
    if (s.size() == 1) t += ",";
    return t + ")";
}

// magic, version 1.0, header length and the padded header dictionary of an array
std::string npy_preamble(const Array& a) {
    std::string dict = "{'descr': '" + a.descr + "', 'fortran_order': False, 'shape': " + shape_text(a.shape) + ", }";
    const size_t used = 10 + dict.size() + 1;   // magic 6, version 2, length 2, newline 1
    dict.append((64 - used % 64) % 64, ' ');
    dict += '\n';
    if (dict.size() > 0xFFFF) throw std::runtime_error("npz: .npy header too long");
    std::string p("\x93NUMPY\x01\x00", 8);
    p += char(dict.size() & 0xFF);
    p += char(dict.size() >> 8);
    return p + dict;
}

struct File {   // std::FILE with closing and checked I/O
    std::FILE* f = nullptr;
    std::string path;
    File(const std::string& p, const char* mode) : f(std::fopen(p.c_str(), mode)), path(p) {
        if (!f) throw std::runtime_error("npz: cannot open " + p);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void put(const void* d, uint64_t n) {
        const uint8_t* p = static_cast<const uint8_t*>(d);
        while (n) {
            const size_t c = size_t(n < (1u << 30) ? n : (1u << 30));
            if (std::fwrite(p, 1, c, f) != c) throw std::runtime_error("npz: write failed: " + path);
            p += c;
            n -= c;
        }
    }
    void get(void* d, uint64_t n) {
        uint8_t* p = static_cast<uint8_t*>(d);
        while (n) {
            const size_t c = size_t(n < (1u << 30) ? n : (1u << 30));
            if (std::fread(p, 1, c, f) != c) throw std::runtime_error("npz: truncated file: " + path);
            p += c;
            n -= c;
        }
    }
    void seek(uint64_t at) {
        if (fseeko(f, off_t(at), SEEK_SET) != 0) throw std::runtime_error("npz: seek failed: " + path);
    }
    uint64_t length() {
        if (fseeko(f, 0, SEEK_END) != 0) throw std::runtime_error("npz: seek failed: " + path);
        return uint64_t(ftello(f));
    }
    void close() {
        const int rc = std::fclose(f);
        f = nullptr;
        if (rc != 0) throw std::runtime_error("npz: write failed: " + path);
    }
};

size_t dtype_size(const std::string& descr) {
    if (descr == "<f8" || descr == "<i8") return 8;
    if (descr.size() > 2 && descr.compare(0, 2, "<U") == 0) return 4 * size_t(std::stoul(descr.substr(2)));
    throw std::runtime_error("npz: unsupported dtype " + descr);
}

// the array of one .npy entry
Array parse_npy(std::vector<uint8_t>&& bytes, const std::string& name) {
    const uint8_t* p = bytes.data();
    const uint64_t n = bytes.size();
    if (n < 10 || std::memcmp(p, "\x93NUMPY", 6) != 0) throw std::runtime_error("npz: entry " + name + " is not a .npy array");
    const int major = p[6];
    const uint64_t hl = major == 1 ? le(p + 8, 2) : le(p + 8, 4), h0 = major == 1 ? 10 : 12;
    if (major < 1 || major > 3 || h0 + hl > n) throw std::runtime_error("npz: bad .npy header in " + name);
    const std::string h(reinterpret_cast<const char*>(p + h0), size_t(hl));
    auto after = [&](const char* key) {
        const size_t k = h.find(key);
        if (k == std::string::npos) throw std::runtime_error("npz: .npy header of " + name + " lacks " + key);
        return k + std::strlen(key);
    };
    Array a;
    const size_t d0 = h.find('\'', after("'descr':"));
    a.descr = h.substr(d0 + 1, h.find('\'', d0 + 1) - d0 - 1);
    const size_t fo = h.find_first_not_of(' ', after("'fortran_order':"));
    if (h.compare(fo, 4, "True") == 0) throw std::runtime_error("npz: Fortran-ordered array " + name);
    const size_t s0 = h.find('(', after("'shape':")), s1 = h.find(')', s0);
    std::string dims = h.substr(s0 + 1, s1 - s0 - 1);
    size_t k = 0;
    while (k < dims.size()) {
        const size_t e = dims.find(',', k);
        const std::string tok = dims.substr(k, e == std::string::npos ? std::string::npos : e - k);
        if (tok.find_first_not_of(' ') != std::string::npos) a.shape.push_back(std::stoll(tok));
        if (e == std::string::npos) break;
        k = e + 1;
    }
    const uint64_t want = a.elements() * dtype_size(a.descr);
    if (n - h0 - hl != want) throw std::runtime_error("npz: entry " + name + " holds " + std::to_string(n - h0 - hl) +
                                                      " data bytes for shape " + shape_text(a.shape) + " of " + a.descr);
    a.store.assign(bytes.begin() + long(h0 + hl), bytes.end());
    a.data = a.store.data();
    a.size = a.store.size();
    return a;
}

}   // namespace

uint64_t Array::elements() const {
    uint64_t e = 1;
    for (const int64_t d : shape) e *= uint64_t(d);
    return e;
}

namespace {
Array owned(const std::string& descr, std::vector<int64_t> shape, const void* d, size_t n) {
    Array a;
    a.descr = descr;
    a.shape = std::move(shape);
    a.store.resize(n);
    if (n) std::memcpy(a.store.data(), d, n);
    a.data = a.store.data();
    a.size = n;
    return a;
}
}   // namespace

Array scalar_f8(double v) { return owned("<f8", {}, &v, 8); }
Array scalar_i8(int64_t v) { return owned("<i8", {}, &v, 8); }
Array vector_f8(const std::vector<double>& v) { return owned("<f8", {int64_t(v.size())}, v.data(), 8 * v.size()); }

Array view_f8(const double* v, const std::vector<int64_t>& shape) {
    Array a;
    a.descr = "<f8";
    a.shape = shape;
    a.data = reinterpret_cast<const uint8_t*>(v);
    a.size = a.elements() * 8;
    return a;
}

Array strings(const std::vector<std::string>& v) {
    size_t w = 1;
    for (const auto& s : v) w = s.size() > w ? s.size() : w;
    std::vector<uint8_t> d(v.size() * w * 4, 0);
    for (size_t i = 0; i < v.size(); ++i)
        for (size_t c = 0; c < v[i].size(); ++c) d[(i * w + c) * 4] = uint8_t(v[i][c]);   // ASCII into UTF-32LE
    return owned("<U" + std::to_string(w), {int64_t(v.size())}, d.data(), d.size());
}

double as_double(const Array& a) {
    if (a.elements() != 1 || a.size != 8) throw std::runtime_error("npz: expected a one-element array");
    if (a.descr == "<f8") {
        double v;
        std::memcpy(&v, a.data, 8);
        return v;
    }
    if (a.descr == "<i8") {
        int64_t v;
        std::memcpy(&v, a.data, 8);
        return double(v);
    }
    throw std::runtime_error("npz: expected an <f8 or <i8 value, found " + a.descr);
}

uint32_t crc32(const uint8_t* p, uint64_t n, uint32_t crc) {
    static const std::array<uint32_t, 256> table = [] {
        std::array<uint32_t, 256> t{};
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t r = i;
            for (int b = 0; b < 8; ++b) r = (r >> 1) ^ (0xEDB88320u & (0u - (r & 1u)));
            t[i] = r;
        }
        return t;
    }();
    crc = ~crc;
    for (uint64_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

void write(const std::string& path, const std::map<std::string, Array>& arrays, uint64_t zip64_from) {
    const std::string tmp = path + ".tmp";
    {
        File f(tmp, "wb");
        Out central;
        uint64_t at = 0;
        for (const auto& kv : arrays) {
            const Array& a = kv.second;
            if (a.size != a.elements() * dtype_size(a.descr)) throw std::runtime_error("npz: size of " + kv.first + " disagrees with its shape");
            const std::string name = kv.first + ".npy", pre = npy_preamble(a);
            const uint64_t size = pre.size() + a.size;
            const uint32_t crc = crc32(a.data, a.size, crc32(reinterpret_cast<const uint8_t*>(pre.data()), pre.size()));
            const bool big = size >= zip64_from, far = at >= zip64_from;
            Out lh;
            lh.u32(kLocalSig);
            lh.u16(big ? 45 : 20);   // version needed: 2.0, 4.5 with zip64
            lh.u16(0);               // flags
            lh.u16(0);               // method: stored
            lh.u16(0);               // time
            lh.u16(kDate19800101);
            lh.u32(crc);
            lh.u32(big ? 0xFFFFFFFFu : uint32_t(size));
            lh.u32(big ? 0xFFFFFFFFu : uint32_t(size));
            lh.u16(uint32_t(name.size()));
            lh.u16(big ? 20 : 0);
            lh.b += name;
            if (big) {
                lh.u16(0x0001);
                lh.u16(16);
                lh.u64(size);
                lh.u64(size);
            }
            f.put(lh.b.data(), lh.b.size());
            f.put(pre.data(), pre.size());
            f.put(a.data, a.size);

            central.u32(kCentralSig);
            central.u16(big || far ? 45 : 20);   // made by
            central.u16(big || far ? 45 : 20);   // needed
            central.u16(0);
            central.u16(0);
            central.u16(0);
            central.u16(kDate19800101);
            central.u32(crc);
            central.u32(big ? 0xFFFFFFFFu : uint32_t(size));
            central.u32(big ? 0xFFFFFFFFu : uint32_t(size));
            central.u16(uint32_t(name.size()));
            const uint32_t extra = (big ? 16u : 0u) + (far ? 8u : 0u);
            central.u16(extra ? extra + 4 : 0);
            central.u16(0);   // comment
            central.u16(0);   // disk
            central.u16(0);   // internal attributes
            central.u32(0);   // external attributes
            central.u32(far ? 0xFFFFFFFFu : uint32_t(at));
            central.b += name;
            if (extra) {
                central.u16(0x0001);
                central.u16(extra);
                if (big) {
                    central.u64(size);
                    central.u64(size);
                }
                if (far) central.u64(at);
            }
            at += lh.b.size() + size;
        }
        const uint64_t count = arrays.size(), cd_at = at, cd_size = central.b.size();
        const bool z64 = count >= 0xFFFF || cd_at >= zip64_from || cd_size >= zip64_from;
        Out end;
        if (z64) {
            end.u32(kEnd64Sig);
            end.u64(44);   // size of the rest of this record
            end.u16(45);
            end.u16(45);
            end.u32(0);
            end.u32(0);
            end.u64(count);
            end.u64(count);
            end.u64(cd_size);
            end.u64(cd_at);
            end.u32(kLocator64Sig);
            end.u32(0);
            end.u64(cd_at + cd_size);   // where the zip64 end record starts
            end.u32(1);
        }
        end.u32(kEndSig);
        end.u16(0);
        end.u16(0);
        end.u16(z64 ? 0xFFFF : uint32_t(count));
        end.u16(z64 ? 0xFFFF : uint32_t(count));
        end.u32(z64 ? 0xFFFFFFFFu : uint32_t(cd_size));
        end.u32(z64 ? 0xFFFFFFFFu : uint32_t(cd_at));
        end.u16(0);
        f.put(central.b.data(), central.b.size());
        f.put(end.b.data(), end.b.size());
        f.close();
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        throw std::runtime_error("npz: cannot rename " + tmp + " to " + path);
    }
}

std::map<std::string, Array> read(const std::string& path) {
    File f(path, "rb");
    const uint64_t len = f.length();
    // the end of central directory record: 22 bytes plus a comment of at most 65535
    const uint64_t tail = len < 65557 ? len : 65557;
    std::vector<uint8_t> t(size_t(tail), 0);
    f.seek(len - tail);
    f.get(t.data(), tail);
    int64_t e = -1;
    for (int64_t k = int64_t(tail) - 22; k >= 0; --k)
        if (le(&t[size_t(k)], 4) == kEndSig) {
            e = k;
            break;
        }
    if (e < 0) throw std::runtime_error("npz: no zip end record in " + path);
    uint64_t count = le(&t[size_t(e) + 10], 2), cd_size = le(&t[size_t(e) + 12], 4), cd_at = le(&t[size_t(e) + 16], 4);
    if (e >= 20 && le(&t[size_t(e) - 20], 4) == kLocator64Sig) {
        const uint64_t rec = le(&t[size_t(e) - 20 + 8], 8);
        uint8_t z[56];
        f.seek(rec);
        f.get(z, sizeof z);
        if (le(z, 4) != kEnd64Sig) throw std::runtime_error("npz: bad zip64 end record in " + path);
        count = le(z + 32, 8);
        cd_size = le(z + 40, 8);
        cd_at = le(z + 48, 8);
    }
    if (cd_at + cd_size > len) throw std::runtime_error("npz: central directory outside the file " + path);
    std::vector<uint8_t> cd(size_t(cd_size), 0);
    f.seek(cd_at);
    f.get(cd.data(), cd_size);
    std::map<std::string, Array> out;
    uint64_t k = 0;
    for (uint64_t i = 0; i < count; ++i) {
        if (k + 46 > cd_size || le(&cd[size_t(k)], 4) != kCentralSig) throw std::runtime_error("npz: bad central directory in " + path);
        const uint8_t* c = &cd[size_t(k)];
        const uint32_t method = uint32_t(le(c + 10, 2)), crc = uint32_t(le(c + 16, 4));
        uint64_t csize = le(c + 20, 4), usize = le(c + 24, 4), at = le(c + 42, 4);
        const uint64_t nl = le(c + 28, 2), xl = le(c + 30, 2), cl = le(c + 32, 2);
        // the entry's name, extra field and comment must lie inside the central directory
        if (k + 46 + nl + xl + cl > cd_size) throw std::runtime_error("npz: bad central directory in " + path);
        const std::string name(reinterpret_cast<const char*>(c + 46), size_t(nl));
        for (uint64_t x = 46 + nl; x + 4 <= 46 + nl + xl;) {   // zip64 extra: the fields whose 32-bit value is all ones
            const uint64_t id = le(c + x, 2), sz = le(c + x + 2, 2);
            const uint64_t end = x + 4 + sz;
            if (end > 46 + nl + xl) throw std::runtime_error("npz: bad extra field of " + name + " in " + path);
            if (id == 0x0001) {
                uint64_t y = x + 4;
                auto take = [&](uint64_t& v) {
                    if (y + 8 > end) throw std::runtime_error("npz: short zip64 field of " + name + " in " + path);
                    v = le(c + y, 8);
                    y += 8;
                };
                if (usize == 0xFFFFFFFFu) take(usize);
                if (csize == 0xFFFFFFFFu) take(csize);
                if (at == 0xFFFFFFFFu) take(at);
            }
            x = end;
        }
        k += 46 + nl + xl + cl;
        if (method != 0 || csize != usize) throw std::runtime_error("npz: entry " + name + " is compressed (only stored entries are read)");
        uint8_t lh[30];
        f.seek(at);
        f.get(lh, 30);
        if (le(lh, 4) != kLocalSig) throw std::runtime_error("npz: bad local header of " + name);
        const uint64_t data = at + 30 + le(lh + 26, 2) + le(lh + 28, 2);
        if (data + usize > len) throw std::runtime_error("npz: entry " + name + " runs past the end of " + path);
        std::vector<uint8_t> bytes(size_t(usize), 0);
        f.seek(data);
        f.get(bytes.data(), usize);
        if (crc32(bytes.data(), usize) != crc) throw std::runtime_error("npz: CRC mismatch in entry " + name + " of " + path);
        std::string key = name;
        if (key.size() > 4 && key.compare(key.size() - 4, 4, ".npy") == 0) key.resize(key.size() - 4);
        out[key] = parse_npy(std::move(bytes), name);
    }
    return out;
}

}   // namespace npz
}   // namespace fvg
