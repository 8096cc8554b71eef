// npy_io.cpp -- see npy_io.h.  The .npy layout (version 1.0): the magic "\x93NUMPY", two
// version bytes, a little-endian uint16 header length, then an ASCII dict
// {'descr': ..., 'fortran_order': False, 'shape': (...), } padded with spaces and a newline
// so the data starts at a multiple of 64 bytes.
#include "npy_io.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>

namespace npy {

static const char kMagic[] = "\x93NUMPY";

static size_t type_size(Type t) {
    switch (t) {
        case Type::f8: return 8;
        case Type::i4: return 4;
        case Type::i1: return 1;
        case Type::u8: return 8;
        case Type::i8: return 8;
    }
    return 0;
}

static const char *type_descr(Type t) {
    switch (t) {
        case Type::f8: return "<f8";
        case Type::i4: return "<i4";
        case Type::i1: return "|i1";
        case Type::u8: return "<u8";
        case Type::i8: return "<i8";
    }
    return "";
}

int64_t Array::count() const {
    int64_t c = 1;
    for (int64_t d : shape) c *= d;
    return c;
}

bool exists(const std::string &path) {
    struct stat st;
    return stat(path.c_str(), &st) == 0;
}

std::string make_dir(const std::string &path) {
    if (mkdir(path.c_str(), 0755) == 0 || errno == EEXIST) return "";
    return path + ": " + strerror(errno);
}

// the value of key 'name' in the header dict, up to the next top-level comma or brace
static bool header_field(const std::string &h, const std::string &name, std::string &val) {
    const size_t k = h.find("'" + name + "'");
    if (k == std::string::npos) return false;
    size_t p = h.find(':', k);
    if (p == std::string::npos) return false;
    p++;
    while (p < h.size() && h[p] == ' ') p++;
    size_t e = p;
    int depth = 0;
    while (e < h.size()) {
        const char ch = h[e];
        if (ch == '(') depth++;
        if (ch == ')') {
            if (depth == 0) break;
            depth--;
            if (depth == 0) {
                e++;
                break;
            }
        }
        if (depth == 0 && (ch == ',' || ch == '}')) break;
        e++;
    }
    val = h.substr(p, e - p);
    return true;
}

std::string read(const std::string &path, Array &out) {
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) return path + ": " + strerror(errno);
    unsigned char pre[10];
    if (fread(pre, 1, 10, fp) != 10 || memcmp(pre, kMagic, 6) != 0) {
        fclose(fp);
        return path + ": not a .npy file";
    }
    size_t hlen = 0, skip = 0;
    if (pre[6] == 1) {
        hlen = (size_t)pre[8] | ((size_t)pre[9] << 8);
    } else {
        unsigned char b[2];
        if (fread(b, 1, 2, fp) != 2) {
            fclose(fp);
            return path + ": truncated header";
        }
        hlen = (size_t)pre[8] | ((size_t)pre[9] << 8) | ((size_t)b[0] << 16) | ((size_t)b[1] << 24);
        skip = 2;
    }
    (void)skip;
    std::string h(hlen, '\0');
    if (fread(&h[0], 1, hlen, fp) != hlen) {
        fclose(fp);
        return path + ": truncated header";
    }
    std::string descr, fortran, shape;
    if (!header_field(h, "descr", descr) || !header_field(h, "fortran_order", fortran) ||
        !header_field(h, "shape", shape)) {
        fclose(fp);
        return path + ": header without descr / fortran_order / shape";
    }
    if (fortran.find("False") == std::string::npos) {
        fclose(fp);
        return path + ": Fortran-ordered arrays are not supported";
    }
    const std::pair<const char *, Type> types[] = {{"'<f8'", Type::f8}, {"'<i4'", Type::i4}, {"'|i1'", Type::i1},
                                                   {"'<u8'", Type::u8}, {"'<i8'", Type::i8}};
    bool known = false;
    for (const auto &t : types)
        if (descr.find(t.first) != std::string::npos) {
            out.type = t.second;
            known = true;
        }
    if (!known) {
        fclose(fp);
        return path + ": element type " + descr + " is not supported (f8, i4, i1, u8, i8)";
    }
    out.shape.clear();
    const char *s = shape.c_str();
    while (*s) {
        if (*s >= '0' && *s <= '9') {
            char *e = nullptr;
            out.shape.push_back(strtoll(s, &e, 10));
            s = e;
        } else {
            s++;
        }
    }
    const size_t bytes = (size_t)out.count() * type_size(out.type);
    out.data.resize(bytes);
    const size_t got = bytes ? fread(out.data.data(), 1, bytes, fp) : 0;
    fclose(fp);
    if (got != bytes) return path + ": truncated data";
    return "";
}

std::string write(const std::string &path, Type type, const std::vector<int64_t> &shape, const void *data) {
    std::string dims;
    int64_t count = 1;
    for (size_t k = 0; k < shape.size(); k++) {
        dims += std::to_string(shape[k]) + (shape.size() == 1 ? "," : (k + 1 < shape.size() ? ", " : ""));
        count *= shape[k];
    }
    std::string h = std::string("{'descr': '") + type_descr(type) + "', 'fortran_order': False, 'shape': (" + dims +
                    "), }";
    const size_t total = 10 + h.size() + 1;
    h.append((64 - total % 64) % 64, ' ');
    h.push_back('\n');
    FILE *fp = fopen(path.c_str(), "wb");
    if (!fp) return path + ": " + strerror(errno);
    unsigned char pre[10];
    memcpy(pre, kMagic, 6);
    pre[6] = 1;
    pre[7] = 0;
    pre[8] = (unsigned char)(h.size() & 0xff);
    pre[9] = (unsigned char)(h.size() >> 8);
    const size_t bytes = (size_t)count * type_size(type);
    bool ok = fwrite(pre, 1, 10, fp) == 10 && fwrite(h.data(), 1, h.size(), fp) == h.size() &&
              (bytes == 0 || fwrite(data, 1, bytes, fp) == bytes);
    ok = (fclose(fp) == 0) && ok;
    return ok ? "" : path + ": write failed";
}

double as_double(const Array &a, int64_t i) {
    const unsigned char *p = a.data.data() + (size_t)i * type_size(a.type);
    switch (a.type) {
        case Type::f8: { double v; memcpy(&v, p, 8); return v; }
        case Type::i4: { int32_t v; memcpy(&v, p, 4); return v; }
        case Type::i1: { int8_t v; memcpy(&v, p, 1); return v; }
        case Type::u8: { uint64_t v; memcpy(&v, p, 8); return (double)v; }
        case Type::i8: { int64_t v; memcpy(&v, p, 8); return (double)v; }
    }
    return 0.0;
}

int64_t as_int(const Array &a, int64_t i) {
    const unsigned char *p = a.data.data() + (size_t)i * type_size(a.type);
    switch (a.type) {
        case Type::f8: { double v; memcpy(&v, p, 8); return (int64_t)v; }
        case Type::i4: { int32_t v; memcpy(&v, p, 4); return v; }
        case Type::i1: { int8_t v; memcpy(&v, p, 1); return v; }
        case Type::u8: { uint64_t v; memcpy(&v, p, 8); return (int64_t)v; }
        case Type::i8: { int64_t v; memcpy(&v, p, 8); return v; }
    }
    return 0;
}

}  // namespace npy
