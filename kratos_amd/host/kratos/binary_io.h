// binary_io.h -- snapshot container of the host API.
//
// The reference's binary format (io/binary/binary_io_base.cpp:101-157,
// utilities/data_transfer/serialize.h), written from its description:
//   byte 0              : (little endian ? 1 : 0) | sizeof(size_t)
//   size_t header_size  : bytes before the bulk data
//   serialized tag map  : size_t n; n x { size_t len, key bytes,
//                         size_t size, char unit size, size_t offset }
//                         in key order (std::map)
//   bulk                : every tag's bytes at header_size + offset
// Readers: this class, and the reference's visual/binary_io.py.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace binary_io {

class base_t {
public:
    virtual ~base_t() {}

    // ---- writing: collect tags, then save() lays them out and writes
    void write(const void *data, size_t count, size_t unit, const std::string &tag) {
        item_t &it = items_[tag];
        it.unit = (char)unit;
        it.bytes.assign((const char *)data, (const char *)data + count * unit);
    }
    template <class T>
    void write_value(const T &v, const std::string &tag) {
        write(&v, 1, sizeof(T), tag);
    }
    // serialize::write of a std::map<std::string, std::string> (unit size 1)
    void write_map(const std::map<std::string, std::string> &m, const std::string &tag) {
        std::string s;
        put_size(s, m.size());
        for (auto &kv : m) {
            put_str(s, kv.first);
            put_str(s, kv.second);
        }
        write(s.data(), s.size(), 1, tag);
    }
    void save(const std::string &file) const {
        std::string map;
        put_size(map, items_.size());
        size_t offset = 0;
        for (auto &kv : items_) {
            put_str(map, kv.first);
            put_size(map, kv.second.bytes.size());
            map.push_back(kv.second.unit);
            put_size(map, offset);
            offset += kv.second.bytes.size();
        }
        const char head = (char)(little_endian() | sizeof(size_t));
        const size_t header_size = 1 + sizeof(size_t) + map.size();
        std::ofstream out(file, std::ios::binary | std::ios::trunc);
        if (!out) throw std::runtime_error("Cannot open file " + file);
        out.write(&head, 1);
        out.write((const char *)&header_size, sizeof(size_t));
        out.write(map.data(), map.size());
        for (auto &kv : items_) out.write(kv.second.bytes.data(), kv.second.bytes.size());
        if (!out) throw std::runtime_error("write failed: " + file);
    }

    // ---- reading
    void load(const std::string &file) {
        std::ifstream in(file, std::ios::binary);
        if (!in) throw std::runtime_error("Cannot open file " + file);
        std::vector<char> all((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        if (all.size() < 1 + sizeof(size_t)) throw std::runtime_error("truncated snapshot " + file);
        const char head = all[0];
        if ((size_t)(head & ~1) != sizeof(size_t)) throw std::runtime_error("Inconsistent size_t ");
        if ((head & 1) != little_endian()) throw std::runtime_error("Inconsistent endian ");
        size_t pos = 1, header_size = get_size(all, pos);
        size_t n = get_size(all, pos);
        items_.clear();
        for (size_t k = 0; k < n; ++k) {
            const std::string key = get_str(all, pos);
            const size_t size = get_size(all, pos);
            if (pos >= all.size()) throw std::runtime_error("truncated snapshot header");
            const char unit = all[pos++];
            const size_t off = get_size(all, pos);
            if (header_size + off + size > all.size()) throw std::runtime_error("truncated snapshot data");
            item_t &it = items_[key];
            it.unit = unit;
            it.bytes.assign(all.begin() + header_size + off, all.begin() + header_size + off + size);
        }
    }
    bool has_tag(const std::string &tag) const { return items_.count(tag) > 0; }
    std::vector<std::string> tags() const {
        std::vector<std::string> t;
        for (auto &kv : items_) t.push_back(kv.first);
        return t;
    }
    const std::vector<char> &bytes(const std::string &tag) const {
        auto it = items_.find(tag);
        if (it == items_.end()) throw std::runtime_error("snapshot has no tag " + tag);
        return it->second.bytes;
    }
    template <class T>
    void read(T *dst, size_t count, const std::string &tag) const {
        const auto &b = bytes(tag);
        if (b.size() != count * sizeof(T)) throw std::runtime_error("Invalid read size: " + tag);
        std::memcpy(dst, b.data(), b.size());
    }
    std::map<std::string, std::string> read_map(const std::string &tag) const {
        const auto &b = bytes(tag);
        std::map<std::string, std::string> m;
        size_t pos = 0;
        const size_t n = get_size(b, pos);
        for (size_t k = 0; k < n; ++k) {
            std::string key = get_str(b, pos);
            m[key] = get_str(b, pos);
        }
        return m;
    }

private:
    struct item_t {
        char unit = 1;
        std::vector<char> bytes;
    };
    std::map<std::string, item_t> items_;

    static int little_endian() {
        const uint16_t one = 1;
        return *(const char *)&one == 1 ? 1 : 0;
    }
    static void put_size(std::string &s, size_t v) { s.append((const char *)&v, sizeof(size_t)); }
    static void put_str(std::string &s, const std::string &v) {
        put_size(s, v.size());
        s += v;
    }
    static size_t get_size(const std::vector<char> &b, size_t &pos) {
        if (pos + sizeof(size_t) > b.size()) throw std::runtime_error("truncated snapshot");
        size_t v;
        std::memcpy(&v, b.data() + pos, sizeof(size_t));
        pos += sizeof(size_t);
        return v;
    }
    static std::string get_str(const std::vector<char> &b, size_t &pos) {
        const size_t n = get_size(b, pos);
        if (pos + n > b.size()) throw std::runtime_error("truncated snapshot");
        std::string s(b.data() + pos, b.data() + pos + n);
        pos += n;
        return s;
    }
};

typedef base_t default_t;

}  // namespace binary_io
