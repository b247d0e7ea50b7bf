// input.h -- parameter-file reader of the host API.
//
// Same file format and access rules as the reference's class input
// (io/args/input.h:28-197, input.cpp:59-201): "[section]" lines set the
// prefix, "key = value" lines define "section|key", '#' starts a comment;
// get<T>(sec, key) throws std::runtime_error when the key is missing,
// get<T>(sec, key, default) falls back, operator()(sec, key, t) leaves t
// untouched and returns false.  3-vectors and sequences read their items
// separated by blanks.  A binary file (first byte not printable) is a
// snapshot: its stored parameters are read and [file] binary names it, the
// restart rule of input.cpp:107-160.
#pragma once

#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "binary_io.h"
#include "types.h"

namespace comm {
struct base_t;
}

class input {
public:
    input() : file_("par.par"), silent_(true) {}
    explicit input(const std::string &file, bool silent = true) { set_file(file, silent); }

    void set_file(const std::string &file, bool silent = true) {
        file_ = file;
        silent_ = silent;
    }
    void merge(const input &src) {
        set("args_merge", "source_" + std::to_string(++merges_), src.file_);
        items_.insert(src.items_.begin(), src.items_.end());
    }
    void set_comm(std::shared_ptr<comm::base_t> p) { com_ = p; }

    // Reads the file named by set_file(); false if it cannot be opened.
    bool read() {
        std::ifstream in(file_, std::ios::in | std::ios::binary);
        if (!in) {
            if (!silent_) std::cerr << "Unable to open input file: " << file_ << std::endl;
            return false;
        }
        const int c0 = in.peek();
        if (!(c0 == EOF || (c0 >= 32 && c0 <= 126))) {
            std::cout << "Reading input parameters from binary: " << file_ << std::endl;
            binary_io::base_t bio;
            bio.load(file_);
            for (auto &kv : bio.read_map("input_args")) {
                items_[kv.first] = kv.second;
                const auto bar = kv.first.find('|');
                if (bar != std::string::npos) prefixes_.insert(kv.first.substr(0, bar));
            }
            set("file", "binary", file_);
            return true;
        }
        std::cout << "Reading input parameters from ascii : " << file_ << std::endl;
        std::string line, section;
        prefixes_.clear();
        while (std::getline(in, line)) {
            line = trim(line.substr(0, line.find('#')));
            if (line.empty()) continue;
            if (line[0] == '[') {
                const auto e = line.find(']');
                if (e == std::string::npos) {
                    std::cerr << "Incorrect input section.\n";
                    throw std::runtime_error("input.cpp");
                }
                section = line.substr(1, e - 1);
                prefixes_.insert(section);
                continue;
            }
            const auto eq = line.find('=');
            const std::string key = trim(line.substr(0, eq));
            std::string val;
            if (eq != std::string::npos) {
                val = line.substr(eq + 1);
                val = val.substr(0, val.find('='));
            }
            items_[expand(section, key)] = trim(val);
        }
        return true;
    }

    template <class T>
    bool operator()(const std::string &pre, const std::string &key, T &res) const {
        const auto it = items_.find(expand(pre, key));
        if (it == items_.end()) return false;
        std::stringstream ss(it->second);
        assign(ss, res);
        return true;
    }
    template <class T>
    T get(const std::string &pre, const std::string &key) const {
        T res{};
        if (!(*this)(pre, key, res)) {
            std::cerr << "Entry \"[" + pre + "]: " + key + "\" not found; using default values.\n";
            throw std::runtime_error("input::get");
        }
        return res;
    }
    template <class T, class D>
    T get(const std::string &pre, const std::string &key, const D &def) const {
        T res{};
        if (!(*this)(pre, key, res)) res = T(def);
        return res;
    }
    template <class T, class D>
    T get_set(const std::string &pre, const std::string &key, const D &def) const {
        T res{};
        if (!(*this)(pre, key, res)) {
            res = T(def);
            const_cast<input *>(this)->set(pre, key, res);
        }
        return res;
    }
    template <class T>
    void set(const std::string &pre, const std::string &key, const T &val) {
        std::stringstream ss;
        ss << val;
        items_[expand(pre, key)] = ss.str();
        if (!pre.empty()) prefixes_.insert(pre);
    }
    template <class T>
    void set(const std::string &pre, const std::string &key, const type::v3_t<T> &val) {
        std::stringstream ss;
        ss << val[0] << ' ' << val[1] << ' ' << val[2];
        items_[expand(pre, key)] = ss.str();
        if (!pre.empty()) prefixes_.insert(pre);
    }
    bool found(const std::string &pre, const std::string &key) const {
        return items_.count(expand(pre, key)) > 0;
    }
    const std::map<std::string, std::string> &get_item_map() const { return items_; }
    const std::set<std::string> &get_prefixes() const { return prefixes_; }

private:
    static std::string trim(const std::string &s) {
        const auto b = s.find_first_not_of(" \t\r\n");
        if (b == std::string::npos) return "";
        const auto e = s.find_last_not_of(" \t\r\n");
        return s.substr(b, e - b + 1);
    }
    static std::string expand(const std::string &pre, const std::string &key) {
        return (pre.empty() ? "" : pre + "|") + key;
    }
    static void assign(std::stringstream &ss, std::string &t) { ss >> t; }
    template <class T, std::enable_if_t<std::is_fundamental<T>::value, int> = 0>
    static void assign(std::stringstream &ss, T &t) { ss >> t; }
    template <class T>
    static void assign(std::stringstream &ss, type::v3_t<T> &t) {
        for (int a = 0; a < 3; ++a) ss >> t[a];
    }
    template <class T>
    static void assign(std::stringstream &ss, std::vector<T> &t) {
        t.clear();
        T v;
        while (ss >> v) t.push_back(v);
    }

    std::string file_;
    bool silent_ = true;
    int merges_ = 0;
    std::map<std::string, std::string> items_;
    std::set<std::string> prefixes_;
    std::shared_ptr<comm::base_t> com_;
};
