// npy.h — reader/writer of the NumPy .npy format (versions 1.0-3.0), the file format
// of every input and output of the reference pipeline (it uses the external cnpy
// library: src/build_model.cu:8, src/spvi_2.cu:39). Little-endian, C order only.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace npy {

struct Array {
    std::string descr;           // e.g. "<f4", "<i4", "<i8", "<u4"
    std::vector<int64_t> shape;
    std::vector<char> data;
    int64_t size() const {
        int64_t n = 1;
        for (auto d : shape) n *= d;
        return n;
    }
    size_t itemsize() const { return (size_t)std::stoi(descr.substr(2)); }
    char kind() const { return descr.size() > 1 ? descr[1] : '?'; }
    template <typename T>
    const T* as() const { return reinterpret_cast<const T*>(data.data()); }
};

inline std::string header_value(const std::string& h, const std::string& key) {
    size_t k = h.find("'" + key + "'");
    if (k == std::string::npos) throw std::runtime_error("npy header lacks " + key);
    size_t c = h.find(':', k);
    size_t b = h.find_first_not_of(" ", c + 1);
    if (h[b] == '(') return h.substr(b, h.find(')', b) - b + 1);
    if (h[b] == '\'') return h.substr(b + 1, h.find('\'', b + 1) - b - 1);
    size_t e = h.find_first_of(",}", b);
    return h.substr(b, e - b);
}

inline Array load(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    unsigned char magic[8];
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "\x93NUMPY", 6) != 0)
        throw std::runtime_error(path + ": not a .npy file");
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        unsigned char b[2];
        if (std::fread(b, 1, 2, f) != 2) throw std::runtime_error(path + ": short header");
        hlen = b[0] | (b[1] << 8);
    } else if (magic[6] == 2 || magic[6] == 3) {
        unsigned char b[4];
        if (std::fread(b, 1, 4, f) != 4) throw std::runtime_error(path + ": short header");
        hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    } else {
        throw std::runtime_error(path + ": unsupported npy version");
    }
    std::string h(hlen, ' ');
    if (std::fread(&h[0], 1, hlen, f) != hlen) throw std::runtime_error(path + ": short header");
    Array a;
    a.descr = header_value(h, "descr");
    if (a.descr.size() < 3 || (a.descr[0] != '<' && a.descr[0] != '|'))
        throw std::runtime_error(path + ": only little-endian arrays are supported (" + a.descr + ")");
    if (header_value(h, "fortran_order").find("True") != std::string::npos)
        throw std::runtime_error(path + ": Fortran-order arrays are not supported");
    std::string sh = header_value(h, "shape");
    for (size_t i = 1; i < sh.size();) {
        size_t j = sh.find_first_of(",)", i);
        std::string tok = sh.substr(i, j - i);
        if (tok.find_first_of("0123456789") != std::string::npos) a.shape.push_back(std::stoll(tok));
        i = j + 1;
    }
    size_t bytes = (size_t)a.size() * a.itemsize();
    a.data.resize(bytes);
    if (bytes && std::fread(a.data.data(), 1, bytes, f) != bytes)
        throw std::runtime_error(path + ": truncated data (" + std::to_string(bytes) + " bytes expected)");
    return a;
}

inline void save(const std::string& path, const void* data, const std::string& descr,
                 const std::vector<int64_t>& shape) {
    std::string shp = "(";
    int64_t n = 1;
    for (size_t i = 0; i < shape.size(); i++) {
        shp += std::to_string(shape[i]);
        if (shape.size() == 1) shp += ",";
        else if (i + 1 < shape.size()) shp += ", ";
        n *= shape[i];
    }
    shp += ")";
    std::string h = "{'descr': '" + descr + "', 'fortran_order': False, 'shape': " + shp + ", }";
    size_t total = 10 + h.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    h += std::string(pad, ' ');
    h += '\n';
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, (unsigned char)(h.size() & 0xff),
                             (unsigned char)(h.size() >> 8)};
    size_t item = (size_t)std::stoi(descr.substr(2));
    bool ok = std::fwrite(pre, 1, 10, f) == 10 && std::fwrite(h.data(), 1, h.size(), f) == h.size() &&
              std::fwrite(data, item, (size_t)n, f) == (size_t)n;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw std::runtime_error("write failed: " + path);
}

}  // namespace npy
