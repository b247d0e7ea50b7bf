// ref_snapshot_reader.cpp -- test infrastructure, not product code.
//
// Reads a snapshot with the reference's own reader: the reference's
// binary_io::base_t (io/binary/binary_io_base.cpp, open / load / read by
// tag) and its tree_t deserializer (mesh/tree/tree.cpp:22-31, the path a
// reference restart takes at meshgen.cpp:96-104).  Built from the
// reference's sources where they lie by oracle/Makefile.ref into
// oracle/_ref/ (never committed); the tests run it on snapshots the host
// layer wrote and compare what it prints with what the snapshot should
// hold.
//
// Output, one record per line:
//   tag <name> <size> <unit size>          every tag, in the file's order
//   hash <name> <fnv1a-64 of the bytes>    every tag
//   arg <key> <value>                      input_args as the reader parses it
//   tree <n_dim> <reg_max x y z> <periodic x y z> <nodes> <unmasked>
//   node <level> <x> <y> <z> <id_l> <id_g> <rank> <mask>   in tree order
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "io/binary/binary_io_base.h"
#include "mesh/tree/tree.h"

static uint64_t fnv1a(const unsigned char *p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s snapshot.bin\n", argv[0]);
        return 2;
    }
    binary_io::base_t bio;
    bio.open(argv[1], "r");
    bio.load();
    for (const auto &kv : bio.b_info)
        std::printf("tag %s %zu %d\n", kv.first.c_str(), kv.second.size, (int)kv.second.u_size);
    for (const auto &kv : bio.b_info) {
        std::vector<unsigned char> buf(kv.second.size);
        if (!buf.empty()) bio.read((void *)buf.data(), kv.first);
        std::printf("hash %s %016llx\n", kv.first.c_str(),
                    (unsigned long long)fnv1a(buf.data(), buf.size()));
    }
    if (bio.b_info.count("input_args")) {
        std::map<std::string, std::string> args;
        bio.read(args, "input_args");
        for (const auto &kv : args) std::printf("arg %s %s\n", kv.first.c_str(), kv.second.c_str());
    }
    if (bio.b_info.count("mesh_tree")) {
        mesh::tree_t tree;
        bio.read(tree, "mesh_tree");
        std::printf("tree %d %d %d %d %d %d %d %zu %zu\n", tree.n_dim, tree.reg_max[0],
                    tree.reg_max[1], tree.reg_max[2], (int)tree.periodic[0], (int)tree.periodic[1],
                    (int)tree.periodic[2], tree.size(), tree.size_eff());
        for (const auto &kv : tree)
            std::printf("node %d %d %d %d %d %d %d %d\n", kv.first.level, kv.first[0], kv.first[1],
                        kv.first[2], kv.second.id_l, kv.second.id_g, kv.second.rank,
                        (int)kv.second.mask);
    }
    return 0;
}
