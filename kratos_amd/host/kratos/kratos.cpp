// kratos.cpp -- mesh_t, cycle and block generator of the host layer, main()
// and prob::cmd.  See mesh.h for what each piece restates.
#include <cfloat>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "dynamics.h"
#include "probgen.h"

namespace mesh {

void check(int rc, kr_ctx *ctx, const char *what) {
    if (rc == KR_OK) return;
    const char *m = kr_last_error(ctx);     // NULL: the last kr_create's reason
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                             "): " + (m ? m : ""));
}

// ------------------------------------------------------------------ block
namespace block {

dual_t::dual_t(block_t &b, kr_ctx *ctx, bool mhd)
    : b_w(b), h_(new block_data_t()), ctx_(ctx), mhd_(mhd) {
    auto &d = *h_;
    d.geo = b.geo;
    d.hyd.n = b.geo.n_ceff;
    d.hyd.n_var = mhd ? 8 : 5;
    d.hyd.u_data.assign((size_t)b.geo.n_ceff.prod() * d.hyd.n_var, 0.0);
    if (mhd) {
        // n + (n > 1) faces plus two ghost layers each side (fld_data.cpp:18-26)
        for (int a = 0; a < 3; ++a) d.fld.N[a] = b.geo.n_ceff[a] + (b.geo.n_ceff[a] > 1) + 4;
        d.fld.bf_data.assign((size_t)d.fld.N.prod() * 3, 0.0);
    }
}

void dual_t::copy_h2d() {
    static_assert(sizeof(type::float2_t) == sizeof(double), "host state is FP64");
    check(kr_upload(ctx_, b_w.get().id, h_->hyd.u_data.data(),
                    mhd_ ? h_->fld.bf_data.data() : nullptr),
          ctx_, "kr_upload");
}

void dual_t::copy_d2h() {
    check(kr_download(ctx_, b_w.get().id, h_->hyd.u_data.data(),
                      mhd_ ? h_->fld.bf_data.data() : nullptr),
          ctx_, "kr_download");
}

}  // namespace block

// ------------------------------------------------------------------ cycle
namespace {
// SIGINT: cycle::base_t::init installs the handler (cycle.cpp:32), as the
// reference does.  The reference throws from inside its handler
// (cycle.cpp:199-205); here the handler only records the signal, and the
// evolve loop turns it into the same exception between kr_advance batches,
// so the device work in flight finishes and the final snapshot sees a
// consistent state.
volatile std::sig_atomic_t g_signal = 0;
extern "C" void on_signal(int signum) { g_signal = signum; }
}  // namespace

void cycle::base_t::sighdl(int signum) { on_signal(signum); }

void cycle::base_t::poll_signal() {
    const int s = g_signal;
    if (s == 0) return;
    g_signal = 0;
    std::stringstream ss;
    ss << "\nSignal : " << s << '\n';
    throw std::runtime_error(ss.str());
}

void cycle::base_t::init(const input &args) {
    std::signal(SIGINT, on_signal);
    args("cycle", "t_0", t);
    args("cycle", "t_lim", t_lim);
    args("cycle", "dt_init", dt_init);
    args("cycle", "dt_expand", dt_expand);
    args("cycle", "n_cycle_lim", n_cycle_lim);
    top_up = args.get<bool>("cycle", "t_lim_top_up", false);
    args("cycle", "t_display_prec", t_display_prec);
    args("cycle", "n_display_freq", n_display_freq);
    if (n_display_freq < 1) n_display_freq = 1;
    prefix = args.get<std::string>("cycle", "prefix_output", "test");
    t_output = args.get<double>("cycle", "t_output_next", 0.0);
    dt_output = args.get<double>("cycle", "dt_output", 1e32);
    final_output = args.get<bool>("cycle", "final_output", false);
}

bool cycle::base_t::update_tout() {
    const double delta_t = t - t_output;
    if (delta_t < 0) return false;
    const size_t n_dt = size_t(delta_t / dt_output) + 1;
    t_output += n_dt * dt_output;
    return true;
}

bool cycle::base_t::output_enabled() const {
    return !std::isnan(t_output + dt_output) && !std::isinf(t_output + dt_output);
}

std::string cycle::base_t::file_name() {
    std::stringstream ss;
    ss << prefix << "_" << std::setfill('0') << std::setw(5) << i_output << ".bin";
    std::cout << "Making output at cycle = " << i_cycle << ", t = " << t
              << "; file name: " << ss.str() << std::endl;
    return ss.str();
}

// cycle.cpp:273-297: "time", "dt" (the device dt seed), "cycle", "i_out",
// "t_out"; saving counts the output first
void cycle::base_t::save(binary_io::base_t &bio, double dt_seed) {
    ++i_output;
    const type::float_t dt = (type::float_t)dt_seed;
    bio.write_value(t, "time");
    bio.write_value(dt, "dt");
    bio.write_value(i_cycle, "cycle");
    bio.write_value(i_output, "i_out");
    bio.write_value(t_output, "t_out");
}

void cycle::base_t::load(const binary_io::base_t &bio, double &dt_seed) {
    type::float_t dt = 0;
    bio.read(&t, 1, "time");
    bio.read(&dt, 1, "dt");
    bio.read(&i_cycle, 1, "cycle");
    bio.read(&i_output, 1, "i_out");
    bio.read(&t_output, 1, "t_out");
    dt_seed = dt;
}

// ------------------------------------------------------------------ meshgen
void meshgen::base_t::read(const input &args) {
    x_min = args.get<type::coord2_t>("mesh", "x_min");
    x_max = args.get<type::coord2_t>("mesh", "x_max");
    n_global = args.get<type::idx_t>("mesh", "n_cell_global");
    n_block = args.get<type::idx_t>("mesh", "n_cell_block", n_global);
    n_dim = 0;
    for (int a = 0; a < 3; ++a)
        if (n_global[a] > 1) n_dim = a + 1;
    for (int a = 0; a < 3; ++a) {
        if (n_global[a] < 1 || n_block[a] < 1) throw std::runtime_error("Zero mesh size");
        if (n_global[a] % n_block[a] != 0)
            throw std::runtime_error("Mesh size indivisible by sub-mesh.");
        n_logic[a] = n_global[a] / n_block[a];
    }
    bnd_kinds.clear();
    args("boundary", "kinds", bnd_kinds);
    if (bnd_kinds.size() < 6) throw std::runtime_error("[boundary] kinds needs 6 entries");
    // meshgen.cpp:111-135 load_refinement
    ref_regions.clear();
    for (const auto &p : args.get_prefixes()) {
        if (p.find("refine_region") == std::string::npos) continue;
        ref_region_t r;
        r.level = args.get<int>(p, "level", 0);
        type::coord2_t x[2];
        for (int a = 0; a < 3; ++a) x[0][a] = 0;
        args(p, "loc", x[0]);
        args(p, "x_min", x[0]);
        x[1] = x[0];
        args(p, "x_max", x[1]);
        for (int l = 0; l < 2; ++l) {
            for (int a = 0; a < n_dim; ++a) {
                r.i[l][a] = search_blk_idx(x[l][a], r.level, a);
                if (r.i[l][a] < 0) throw std::runtime_error("Incorrect SMR region");
            }
            for (int a = n_dim; a < 3; ++a) r.i[l][a] = 0;
        }
        ref_regions.push_back(r);
    }
}

int meshgen::base_t::search_blk_idx(double loc, int lvl, int axis) const {
    if (loc > x_max[axis] || loc < x_min[axis]) return -1;
    const int i_tot = n_logic[axis] << lvl;
    int i_r = i_tot, i_l = 0;
    const double i_ftot(i_tot);
    while (i_r - i_l > 1) {
        const int i_m = (i_l + i_r) / 2;
        if (location(i_m / i_ftot, axis) < loc) i_l = i_m;
        else i_r = i_m;
    }
    return i_l;
}

// ------------------------------------------------------------------ mesh_t
mesh_t::mesh_t() {}

mesh_t::~mesh_t() {
    for (auto &m : mods_) m->finalize(*this);
    duals_.clear();
    if (ctx_) kr_destroy(ctx_);
}

void mesh_t::init(const input &args) {
    if (!p_dev) throw std::runtime_error("device unspecified");
    if (!p_bio) enroll_binary_io<binary_io::default_t>();
    if (!p_mgn) enroll_mgen<meshgen::base_t>();
    if (!p_cyc) enroll_cycle<cycle::base_t>();
    if (!p_com) enroll_comm<comm::base_t>();
    for (auto &r : reads_) r(args);
    input_args_ = args.get_item_map();
    // "fixd" sides: the ambient [init] medium in primitives
    // (hydro/boundary/keeper.cpp:60-72, mhd/boundary/keeper.h:126-146)
    {
        const auto v0 = args.get<type::coord2_t>("init", "vel0", 0);
        const auto b0 = args.get<type::coord2_t>("init", "b0", 0);
        fix_w_[0] = args.get<double>("init", "rho0", 0);
        fix_w_[1] = args.get<double>("init", "pre0", 0);
        for (int a = 0; a < 3; ++a) {
            fix_w_[2 + a] = v0[a];
            fix_b_[a] = b0[a];
        }
    }
    restart_ = args.found("file", "binary");
    p_cyc->init(args);
    if (restart_) {
        restart_file_ = args.get<std::string>("file", "binary");
        binary_io::base_t bio;
        bio.load(restart_file_);
        double seed = 0;
        p_cyc->load(bio, seed);
        p_cyc->dt_init = seed;                 // the device seed the run stopped with
    }
    p_mgn->read(args);
    if (restart_) blocks_from_snapshot();
    else build_blocks();
    for (auto &p : inits_) p.second(args);
    if (!ctx_) throw std::runtime_error("no dynamics module enrolled");
}

// Blocks of a same-level Cartesian mesh: order, geometry and the 26
// neighbours with periodic wrap (meshgen.cpp:286-321, tree.cpp:77-84,
// 177-197).  A thin axis (one cell) has no neighbours along it.  Block ids
// follow the tree's lexicographic order (x slowest, geometry/region.h
// :17-33), which is the reference's [mesh] dist_mode = 0 numbering
// (meshgen.cpp:167-187): a restart of the reference from this build's
// snapshot finds every block under the id it gives it.
void mesh_t::build_blocks() {
    const auto &g = *p_mgn;
    bool periodic[3];
    for (int a = 0; a < 3; ++a) periodic[a] = g.bnd_kinds[2 * a].find("per") != std::string::npos;
    blocks_.clear();
    tree_.clear();
    set_tree_meta(periodic);
    bool refined = false;
    for (auto &r : g.ref_regions) refined |= r.level > 0;
    if (refined) return build_refined(periodic);
    const int nb = g.n_logic.prod();
    auto lex_id = [&](const int *q) { return (q[0] * g.n_logic[1] + q[1]) * g.n_logic[2] + q[2]; };
    blocks_.resize(nb);
    for (int b = 0; b < nb; ++b) {
        block_t &k = blocks_[b];
        k.id = b;
        k.reg.reg = type::idx_t(b / (g.n_logic[1] * g.n_logic[2]), (b / g.n_logic[2]) % g.n_logic[1],
                                b % g.n_logic[2]);
        tree_[tree_key_t{0, k.reg.reg[0], k.reg.reg[1], k.reg.reg[2]}] = tree_node_t{b, false};
        k.geo.n_ceff = g.n_block;
        for (int a = 0; a < 3; ++a) {
            const int n = g.n_block[a];
            auto loc = [&](type::float2_t i) -> type::float2_t {
                if (n == 1) return i;
                i = (k.reg.reg[a] + i / n) / (type::float2_t)g.n_logic[a];
                return g.x_min[a] + i * (g.x_max[a] - g.x_min[a]);
            };
            k.geo.xf0[a] = (type::float_t)loc(0);
            k.geo.dx0[a] = (type::float_t)(loc(1) - k.geo.xf0[a]);
        }
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int d[3] = {dx, dy, dz};
                    int q[3];
                    bool ok = true;
                    for (int a = 0; a < 3 && ok; ++a) {
                        int v = k.reg.reg[a] + d[a];
                        if (g.n_block[a] == 1 && d[a] != 0) ok = false;
                        if (v < 0 || v >= g.n_logic[a]) {
                            if (!periodic[a]) ok = false;
                            v = (v % g.n_logic[a] + g.n_logic[a]) % g.n_logic[a];
                        }
                        q[a] = v;
                    }
                    k.nbr[((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1)] = ok ? lex_id(q) : -1;
                }
    }
}

// Tree-wide data of mesh/meshgen/meshgen.cpp:44-56: n_dim, the base-level
// block counts (0 along an axis past n_dim, where the reference leaves
// i_logic_lim unset) and the periodic flag per axis from [boundary] kinds.
void mesh_t::set_tree_meta(const bool *periodic) {
    const auto &g = *p_mgn;
    tree_n_dim_ = g.n_dim;
    for (int a = 0; a < 3; ++a) {
        tree_reg_max_[a] = a < g.n_dim ? g.n_logic[a] : 0;
        tree_periodic_[a] = periodic[a];
    }
}

// A statically refined mesh: the reference's tree (tree.cpp:185-218
// insert / fill_base_level, meshgen.cpp:137-158 generate_tree).  Inserting
// a block masks its parent, inserts the parent's other children and the
// parent's 26 same-level neighbours, so the leaves are 2:1 balanced.
// Blocks are the leaves in the tree's order (level, then x, y, z), ids
// 0..n-1 in that order (the reference's [mesh] dist_mode = 0 numbering);
// the library derives the neighbours from level and logical index.
void mesh_t::build_refined(const bool *periodic) {
    const auto &g = *p_mgn;
    typedef std::array<int, 4> key_t;            // level, x, y, z
    std::map<key_t, bool> tree;                  // value: masked
    auto lim = [&](int L, int a) { return a < g.n_dim ? (g.n_logic[a] << L) : 1; };
    std::function<void(int, const int *)> insert = [&](int L, const int *r) {
        const key_t k{L, r[0], r[1], r[2]};
        if (tree.count(k)) return;
        if (L > 0) {
            int rc[3];
            for (int a = 0; a < 3; ++a) rc[a] = a < g.n_dim ? (r[a] >= 0 ? r[a] / 2 : (r[a] - 1) / 2) : 0;
            insert(L - 1, rc);
            tree[key_t{L - 1, rc[0], rc[1], rc[2]}] = true;
            for (int c = 0; c < (1 << g.n_dim); ++c) {
                key_t ch{L, 0, 0, 0};
                for (int a = 0; a < 3; ++a) ch[1 + a] = a < g.n_dim ? 2 * rc[a] + ((c >> a) & 1) : 0;
                tree.emplace(ch, false);
            }
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int d[3] = {dx, dy, dz};
                        bool zero = true, skip = false;
                        int q[3];
                        for (int a = 0; a < 3; ++a) {
                            zero &= d[a] == 0;
                            if (a >= g.n_dim && d[a] != 0) skip = true;
                            q[a] = rc[a] + d[a];
                            const int m = lim(L - 1, a);
                            if (a < g.n_dim && periodic[a]) q[a] = (q[a] + m) % m;
                            if (q[a] < 0 || q[a] >= m) skip = true;
                        }
                        if (!zero && !skip) insert(L - 1, q);
                    }
        }
        tree.emplace(k, false);
    };
    for (auto &rr : g.ref_regions) {
        if (rr.level <= 0) continue;
        for (int z = rr.i[0][2]; z <= rr.i[1][2]; ++z)
            for (int y = rr.i[0][1]; y <= rr.i[1][1]; ++y)
                for (int x = rr.i[0][0]; x <= rr.i[1][0]; ++x) {
                    const int r[3] = {x, y, z};
                    insert(rr.level, r);
                }
    }
    for (int z = 0; z < g.n_logic[2]; ++z)
        for (int y = 0; y < g.n_logic[1]; ++y)
            for (int x = 0; x < g.n_logic[0]; ++x) tree.emplace(key_t{0, x, y, z}, false);
    int id = 0;
    for (auto &kv : tree) {
        tree_[kv.first] = tree_node_t{kv.second ? -1 : id, kv.second};
        if (kv.second) continue;
        block_t k;
        k.id = id++;
        k.reg.level = kv.first[0];
        k.reg.reg = type::idx_t(kv.first[1], kv.first[2], kv.first[3]);
        k.geo.n_ceff = g.n_block;
        for (int a = 0; a < 3; ++a) {
            const int n = g.n_block[a];
            auto loc = [&](type::float2_t i) -> type::float2_t {
                if (n == 1) return i;
                i = (k.reg.reg[a] + i / n) / (type::float2_t)(g.n_logic[a] << k.reg.level);
                return g.location(i, a);
            };
            k.geo.xf0[a] = (type::float_t)loc(0);
            k.geo.dx0[a] = (type::float_t)(loc(1) - k.geo.xf0[a]);
        }
        for (int j = 0; j < 27; ++j) k.nbr[j] = -1;
        blocks_.push_back(k);
    }
}

// A restart takes its blocks from the snapshot's mesh_tree, as the
// reference reads its tree back from the file (meshgen.cpp:96-104): the
// unmasked nodes in the tree's order, numbered 0..n-1 as dist_default does
// (meshgen.cpp:167-187), each reading its state from the tags of the id the
// file gives it.  A snapshot without mesh_tree gives its blocks through the
// per-block level and i_logic tags instead.
void mesh_t::blocks_from_snapshot() {
    binary_io::base_t bio;
    bio.load(restart_file_);
    if (bio.has_tag("mesh_tree")) {
        const auto &b = bio.bytes("mesh_tree");
        return blocks_from_tree(std::string(b.begin(), b.end()));
    }
    std::map<int, block_t> found;
    for (const auto &t : bio.tags()) {
        const std::string suf = "|level";
        if (t.rfind("block_", 0) != 0 || t.size() <= suf.size() ||
            t.compare(t.size() - suf.size(), suf.size(), suf) != 0)
            continue;
        const int id = std::stoi(t.substr(6, t.size() - suf.size() - 6));
        const std::string pre = "block_" + std::to_string(id) + "|";
        block_t k;
        k.id = id;
        bio.read(&k.reg.level, 1, pre + "level");
        bio.read(k.reg.reg.x, 3, pre + "i_logic");
        found[id] = k;
    }
    if (found.empty()) throw std::runtime_error("snapshot holds no blocks: " + restart_file_);
    std::string bytes;
    auto put = [&](const void *p, size_t n) { bytes.append((const char *)p, n); };
    const size_t n = found.size();
    put(&n, sizeof(n));
    std::map<tree_key_t, int> leaves;
    for (auto &kv : found) leaves[tree_key_t{kv.second.reg.level, kv.second.reg.reg[0],
                                             kv.second.reg.reg[1], kv.second.reg.reg[2]}] = kv.first;
    for (auto &kv : leaves) {
        const int key[4] = {kv.first[1], kv.first[2], kv.first[3], kv.first[0]};
        const int node[3] = {kv.second, kv.second, 0};
        const char mask[4] = {0, 0, 0, 0};
        put(key, sizeof(key));
        put(node, sizeof(node));
        put(mask, sizeof(mask));
    }
    blocks_from_tree(bytes);
}

// Parses the serialized tree (mesh/tree/tree.cpp:11-31 with
// utilities/data_transfer/serialize.h: size_t node count, then per node
// the region key {int x[3], int level} and the node {int id_l, id_g, rank;
// bool mask} as raw structs, then n_dim, reg_max[3] and periodic[3]).
// The trailing tree-wide fields are optional here (the tag-built tree above
// has none); this build takes them from the parameters.
void mesh_t::blocks_from_tree(const std::string &bytes) {
    const auto &g = *p_mgn;
    size_t off = 0;
    auto get = [&](void *p, size_t n) {
        if (off + n > bytes.size()) throw std::runtime_error("mesh_tree truncated");
        std::memcpy(p, bytes.data() + off, n);
        off += n;
    };
    size_t n = 0;
    get(&n, sizeof(n));
    std::map<tree_key_t, tree_node_t> nodes;
    for (size_t i = 0; i < n; ++i) {
        int key[4], node[3];
        char mask[4];
        get(key, sizeof(key));
        get(node, sizeof(node));
        get(mask, sizeof(mask));
        nodes[tree_key_t{key[3], key[0], key[1], key[2]}] = tree_node_t{node[1], mask[0] != 0};
    }
    bool periodic[3];
    for (int a = 0; a < 3; ++a) periodic[a] = g.bnd_kinds[2 * a].find("per") != std::string::npos;
    bool refined = false;
    for (auto &kv : nodes) refined |= kv.first[0] != 0 || kv.second.mask;
    if (!refined) {
        // same-level: the Cartesian blocks, each reading the tags of the id
        // the file's tree gives its logical index
        build_blocks();
        for (auto &k : blocks_) {
            auto it = nodes.find(tree_key_t{0, k.reg.reg[0], k.reg.reg[1], k.reg.reg[2]});
            if (it == nodes.end()) throw std::runtime_error("mesh_tree lacks a block of the mesh");
            k.file_id = it->second.id_g;
        }
        return;
    }
    blocks_.clear();
    tree_.clear();
    set_tree_meta(periodic);
    int id = 0;
    for (auto &kv : nodes) {
        tree_[kv.first] = tree_node_t{kv.second.mask ? -1 : id, kv.second.mask};
        if (kv.second.mask) continue;
        block_t k;
        k.id = id++;
        k.file_id = kv.second.id_g;
        k.reg.level = kv.first[0];
        k.reg.reg = type::idx_t(kv.first[1], kv.first[2], kv.first[3]);
        k.geo.n_ceff = g.n_block;
        for (int a = 0; a < 3; ++a) {
            const int nc = g.n_block[a];
            auto loc = [&](type::float2_t i) -> type::float2_t {
                if (nc == 1) return i;
                i = (k.reg.reg[a] + i / nc) / (type::float2_t)(g.n_logic[a] << k.reg.level);
                return g.location(i, a);
            };
            k.geo.xf0[a] = (type::float_t)loc(0);
            k.geo.dx0[a] = (type::float_t)(loc(1) - k.geo.xf0[a]);
        }
        for (int j = 0; j < 27; ++j) k.nbr[j] = -1;
        blocks_.push_back(k);
    }
}

// mesh_tree as the reference writes it (meshgen.cpp:459-466 ->
// tree.cpp:11-20): the nodes in the tree's order with id_l = id_g (one
// rank) and rank 0 for blocks, -1 for masked parents; padding bytes zero.
void mesh_t::write_tree(binary_io::base_t &bio) const {
    std::string bytes;
    auto put = [&](const void *p, size_t n) { bytes.append((const char *)p, n); };
    const size_t n = tree_.size();
    put(&n, sizeof(n));
    for (auto &kv : tree_) {
        const int key[4] = {kv.first[1], kv.first[2], kv.first[3], kv.first[0]};
        const int id = kv.second.id_g;
        const int node[3] = {id, id, id < 0 ? -1 : 0};
        const char mask[4] = {(char)kv.second.mask, 0, 0, 0};
        put(key, sizeof(key));
        put(node, sizeof(node));
        put(mask, sizeof(mask));
    }
    put(&tree_n_dim_, sizeof(int));
    put(tree_reg_max_, 3 * sizeof(int));
    put(tree_periodic_, 3 * sizeof(bool));
    bio.write(bytes.data(), bytes.size(), 1, "mesh_tree");
}

static int bnd_code(const std::string &k) {
    if (k.find("per") != std::string::npos) return KR_PER;
    if (k == "out") return KR_OUT;
    if (k == "fre") return KR_FRE;
    if (k == "ref") return KR_REF;
    if (k == "fixd") return KR_FIX;
    throw std::runtime_error("Undef bnd " + k);
}

void mesh_t::init_dynamics(int model, int precision, double gamma, double cfl, double rho_floor,
                           double tmp_floor) {
    if (ctx_) throw std::runtime_error("one dynamics module per mesh");
    kr_config c{};
    c.model = model;
    c.precision = precision;
    c.gamma = gamma;
    c.cfl = cfl;
    c.rho_floor = rho_floor;
    c.tmp_floor = tmp_floor;
    for (int s = 0; s < 6; ++s) c.bnd[s] = bnd_code(p_mgn->bnd_kinds[s]);
    for (int k = 0; k < 5; ++k) c.fix_w[k] = fix_w_[k];
    for (int k = 0; k < 3; ++k) c.fix_b[k] = fix_b_[k];
    c.t_0 = p_cyc->t;
    c.t_lim = p_cyc->t_lim;
    c.dt_init = p_cyc->dt_init;
    c.dt_expand = p_cyc->dt_expand;
    c.top_up = p_cyc->top_up;
    for (int a = 0; a < 3; ++a) c.n_cell_block[a] = p_mgn->n_block[a];
    c.kernel_path = KR_PATH_AUTO;
    check(kr_create(&c, &ctx_), nullptr, "kr_create");
    mhd_ = model == KR_MHD;
    std::vector<kr_block_desc> ds(blocks_.size());
    for (size_t b = 0; b < blocks_.size(); ++b) {
        auto &k = blocks_[b];
        auto &d = ds[b];
        d.id = k.id;
        d.level = k.reg.level;
        for (int a = 0; a < 3; ++a) {
            d.reg[a] = k.reg.reg[a];
            d.n_ceff[a] = k.geo.n_ceff[a];
            d.xf0[a] = k.geo.xf0[a];
            d.dx0[a] = k.geo.dx0[a];
        }
        for (int j = 0; j < 27; ++j) d.nbr[j] = k.nbr[j];
    }
    check(kr_add_blocks(ctx_, (int)ds.size(), ds.data()), ctx_, "kr_add_blocks");
    duals_.clear();
    for (auto &k : blocks_) duals_.emplace_back(new block::dual_t(k, ctx_, mhd_));
}

void mesh_t::pull_time() {
    double t, dtl, seed;
    check(kr_get_time(ctx_, &t, &dtl, &seed), ctx_, "kr_get_time");
    p_cyc->t = t;
    dt_last_ = dtl;
}

double mesh_t::dt() const { return dt_last_; }

double mesh_t::divb_max() const {
    double d = 0;
    check(kr_divb_max(ctx_, &d), ctx_, "kr_divb_max");
    return d;
}

void mesh_t::copy_d2h() {
    for (auto &d : duals_) d->copy_d2h();
}

void mesh_t::step() {
    double dt = 0;
    check(kr_step(ctx_, &dt), ctx_, "kr_step");
    pull_time();
}

// mesh.cpp:99-111 with the per-block tags of block.cpp:18-38,
// geometry.cpp:161-178, hyd_data.cpp:49-57 (hydro_cons_*) and
// fld_data.cpp:87-95 (field_bf_*); slot 0 of the conserved state and the
// face field with its two ghost layers (zero here: ghosts are not stored).
void mesh_t::save(const std::string &file_name) {
    binary_io::base_t bio;
    // block ids follow the tree's order, the reference's dist_mode = 0: a
    // restart of the reference from this file must number them that way
    auto args = input_args_;
    args["mesh|dist_mode"] = "0";
    bio.write_map(args, "input_args");
    write_tree(bio);
    double t, dtl, seed;
    check(kr_get_time(ctx_, &t, &dtl, &seed), ctx_, "kr_get_time");
    p_cyc->t = t;
    p_cyc->save(bio, seed);
    const int n_gh = 2;
    for (size_t b = 0; b < blocks_.size(); ++b) {
        const block_t &k = blocks_[b];
        auto &d = *duals_[b];
        d.copy_d2h();
        const std::string pre = "block_" + std::to_string(k.id) + "|";
        bio.write(k.reg.reg.x, 3, sizeof(int), pre + "i_logic");
        bio.write_value(k.reg.level, pre + "level");
        type::float_t sfc0[3], vol0 = 1;
        for (int a = 0; a < 3; ++a) {
            type::float_t s = 1;
            for (int m = 1; m < 3; ++m) s *= k.geo.dx0[(a + m) % 3];
            sfc0[a] = s;
            vol0 *= k.geo.dx0[a];
        }
        const bool uniform[3] = {true, true, true};
        bio.write(k.geo.xf0.x, 3, sizeof(type::float_t), pre + "xf0");
        bio.write(k.geo.dx0.x, 3, sizeof(type::float_t), pre + "dx0");
        bio.write(sfc0, 3, sizeof(type::float_t), pre + "sfc0");
        bio.write_value(vol0, pre + "vol0");
        bio.write_value(n_gh, pre + "n_gh");
        bio.write(k.geo.n_ceff.x, 3, sizeof(int), pre + "n_ceff");
        bio.write(uniform, 3, sizeof(bool), pre + "uniform");
        const auto &h = d.h();
        const int n_step = 2, zero[3] = {0, 0, 0};
        bio.write_value(h.hyd.n_var, pre + "hydro_cons_n_int");
        bio.write_value(n_step, pre + "hydro_cons_n_step");
        bio.write(h.hyd.n.x, 3, sizeof(int), pre + "hydro_cons_n_cell");
        bio.write(zero, 3, sizeof(int), pre + "hydro_cons_n_gh");
        bio.write(h.hyd.u_data.data(), h.hyd.u_data.size(), sizeof(double), pre + "hydro_cons_field");
        if (mhd_) {
            const int three = 3, gh[3] = {n_gh, n_gh, n_gh};
            bio.write_value(three, pre + "field_bf_n_int");
            bio.write_value(n_step, pre + "field_bf_n_step");
            bio.write(h.fld.N.x, 3, sizeof(int), pre + "field_bf_n_cell");
            bio.write(gh, 3, sizeof(int), pre + "field_bf_n_gh");
            bio.write(h.fld.bf_data.data(), h.fld.bf_data.size(), sizeof(double),
                      pre + "field_bf_field");
        }
    }
    bio.save(file_name);
}

void mesh_t::load_state() {
    binary_io::base_t bio;
    bio.load(restart_file_);
    for (size_t b = 0; b < blocks_.size(); ++b) {
        auto &d = *duals_[b];
        auto &h = d.h();
        const int fid = blocks_[b].file_id >= 0 ? blocks_[b].file_id : blocks_[b].id;
        const std::string pre = "block_" + std::to_string(fid) + "|";
        int nv = 0;
        bio.read(&nv, 1, pre + "hydro_cons_n_int");
        if (nv != h.hyd.n_var) throw std::runtime_error("snapshot model differs: " + pre);
        bio.read(h.hyd.u_data.data(), h.hyd.u_data.size(), pre + "hydro_cons_field");
        if (mhd_) bio.read(h.fld.bf_data.data(), h.fld.bf_data.size(), pre + "field_bf_field");
        d.copy_h2d();
    }
}

// cycle::evolve (cycle.cpp:152-194): the same loop, stop rule and Speed
// line; cycles run in batches of n_display_freq through kr_advance, which
// replays one captured hipGraph per cycle and stops once t >= t_lim.
void mesh_t::evolve() {
    auto &c = *p_cyc;
    c.i_cycle_0 = c.i_cycle;
    double n_cell_tot = 0;
    for (auto &b : blocks_) n_cell_tot += (double)b.geo.n_ceff.prod();
    using clock = std::chrono::steady_clock;
    const auto start = clock::now();
    auto prev = start;
    check(kr_sync(ctx_), ctx_, "kr_sync");
    // snapshot due before a cycle (cycle.cpp:124-150): with outputs on,
    // cycles run one per kr_advance so that t is checked before each
    const bool outputs = c.output_enabled();
    bool to_stop = false;
    try {
        // the reference's loop runs n_cycle_lim - i_cycle + 1 steps unless
        // t reaches t_lim first; i_cycle stops at n_cycle_lim
        long long left = (long long)c.n_cycle_lim - c.i_cycle + 1;
        while (left > 0) {
            if (outputs) {
                pull_time();
                if (c.update_tout()) save(c.file_name());
            }
            c.poll_signal();
            const int batch = outputs ? 1 : (int)std::min<long long>(left, c.n_display_freq);
            int done = 0;
            check(kr_advance(ctx_, batch, &done), ctx_, "kr_advance");
            pull_time();
            left -= done;
            const bool last = left == 0;
            c.i_cycle += last ? done - 1 : done;
            if (done > 0 && p_com->is_root()) {
                const auto now = clock::now();
                const double wt = std::chrono::duration<double>(now - prev).count();
                prev = now;
                std::cout << "cycle = " << std::setfill(' ') << std::setw(5) << c.i_cycle
                          << std::scientific << std::setprecision(c.t_display_prec)
                          << ", t = " << c.t << ", dt = " << dt_last_
                          << ", Speed = " << n_cell_tot / wt * done << std::endl;
            }
            if (done < batch) break;
        }
    } catch (const std::exception &e) {
        std::cerr << e.what();
        to_stop = true;
    }
    if (to_stop && c.final_output) return save(c.file_name());
    check(kr_sync(ctx_), ctx_, "kr_sync");
    const double dwt = std::chrono::duration<double>(clock::now() - start).count();
    const double speed = n_cell_tot / dwt * (c.i_cycle - c.i_cycle_0 + 1);
    if (p_com->is_root())
        std::cout << "Duration = " << dwt << " s, " << "Speed = " << speed << " cell/s,\n"
                  << "Pace = " << 1 / speed << " s/cell.\n";
    if (c.final_output && !std::isnan(c.t_output)) save(c.file_name());
}

}  // namespace mesh

// ------------------------------------------------------------------ prob
namespace prob {

void cmd(int argc, char *argv[], input &args) {
    for (int i = 1; i < argc; ++i) {
        args.set_file(argv[i]);
        args.read();
    }
}

void __attribute__((weak)) run(int, char *[]) { std::cerr << "Default run function is empty!\n"; }

}  // namespace prob

int main(int argc, char *argv[]) {
    if (argc <= 1) throw std::runtime_error("No input file.");
    prob::run(argc, argv);
    return 0;
}
