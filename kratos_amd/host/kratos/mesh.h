// mesh.h -- mesh_t, blocks and the enrolment API of the host layer.
//
// A C++ restatement of the part of the reference's mesh layer that a
// problem generator touches on the Godunov path: mesh::mesh_t's enrol /
// init / step / evolve (mesh/mesh.h:26-129, mesh.cpp:31-80,
// mesh_enroll.h:77-109), the Cartesian block generator for a same-level
// mesh (meshgen/meshgen.cpp:13-42, 286-321), the cycle loop
// (cycle/cycle.cpp:30-194) and the host view of one block
// (block/dual_data.h:11-45, geometry/geometry.h:73-105).  Device memory,
// halos, fluxes and the time step live behind the C-ABI of kratos_amd.h;
// this layer only describes blocks, fills initial states on the host and
// drives the cycle.  Errors of the C-ABI come back as std::runtime_error,
// the reference's error convention.
#pragma once

#include <array>
#include <chrono>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "binary_io.h"
#include "input.h"
#include "kratos_amd.h"
#include "types.h"

namespace comm {
// Single-rank communicator: this build drives one GPU per process.
struct base_t {
    virtual ~base_t() {}
    virtual bool is_root() const { return true; }
    virtual int n_rank() const { return 1; }
    virtual int l_rank() const { return 0; }
};
}  // namespace comm

namespace device {
struct base_t {
    virtual ~base_t() {}
};
}  // namespace device
// The device type a generator enrols (device/general.h): the MI355X
// context behind the C-ABI.
struct device_t : device::base_t {};

namespace mesh {

class mesh_t;

namespace block {

struct geo_t {
    type::idx_t n_ceff;
    type::coord_t xf0, dx0;
    int n_gh = 2;
    type::float_t x_cc(int a, int i) const { return xf0[a] + (i + 0.5f) * dx0[a]; }
    type::float_t x_fc(int a, int i) const { return xf0[a] + i * dx0[a]; }
    template <class I>
    void x_cc(type::coord_t &x, const I &i) const {
        for (int a = 0; a < 3; ++a) x[a] = x_cc(a, i[a]);
    }
    template <class I>
    type::coord_t x_cc(const I &i) const {
        type::coord_t x;
        x_cc(x, i);
        return x;
    }
};

struct region_t {
    type::idx_t reg;
    int level = 0;
};

struct block_t {
    int id = 0;
    int file_id = -1;      // id of the block's tags in the snapshot restarted from
    region_t reg;
    geo_t geo;
    int nbr[27];
};

// Host copy of the conserved state of one block, RK slot 0, in the
// reference's AoS layout (mesh/types/dat_3d.h:153-156): n_var values per
// interior cell, x fastest.
struct hyd_host_t {
    type::idx_t n;
    int n_var = 5;
    std::vector<type::float2_t> u_data;
    type::float2_t *at(const type::idx_t &i) {
        return u_data.data() + ((size_t)(i[2] * n[1] + i[1]) * n[0] + i[0]) * n_var;
    }
    struct view_t {
        hyd_host_t *p;
        type::float2_t *at(const type::idx_t &i) const { return p->at(i); }
    } u{this};
};

// Host copy of the face field of one block: 3 components per face index
// with two ghost layers on each side (mhd/types/fld_data.cpp:18-26).
struct fld_host_t {
    type::idx_t N;
    std::vector<type::float2_t> bf_data;
    type::float2_t *at(const type::idx_t &i) {
        return bf_data.data() + ((size_t)((i[2] + 2) * N[1] + (i[1] + 2)) * N[0] + (i[0] + 2)) * 3;
    }
    struct view_t {
        fld_host_t *p;
        type::float2_t *at(const type::idx_t &i) const { return p->at(i); }
    } bf{this};
};

struct block_data_t {
    geo_t geo;
    hyd_host_t hyd;
    fld_host_t fld;
    block_data_t() = default;
    block_data_t(const block_data_t &) = delete;
    block_data_t &operator=(const block_data_t &) = delete;
};

// Host/device pair of one block.  h() is the host payload an init_cond
// fills; copy_h2d / copy_d2h move slot 0 through the C-ABI.
class dual_t {
public:
    dual_t(block_t &b, kr_ctx *ctx, bool mhd);
    std::reference_wrapper<block_t> b_w;
    block_data_t &h() { return *h_; }
    void copy_h2d();
    void copy_d2h();

private:
    std::unique_ptr<block_data_t> h_;
    kr_ctx *ctx_;
    bool mhd_;
};

}  // namespace block

using block::block_t;

namespace cycle {
// Time loop state and keys of mesh/cycle/cycle.cpp:20-65, with the
// snapshot schedule (:53-60, :228-237, :256-297).
struct base_t {
    virtual ~base_t() {}
    double t = 0, t_lim = 3.40282346638528859812e+38, dt_init = 0, dt_expand = 2;
    int i_cycle = 0, i_cycle_0 = 0, n_cycle_lim = 0, n_display_freq = 1, t_display_prec = 6;
    bool top_up = false;
    std::string prefix = "test";
    double t_output = 0, dt_output = 1e32;
    bool final_output = false;
    int i_output = 0;
    virtual void init(const input &args);
    // output due at the current t (advances t_output past t), cycle.cpp:228-237
    bool update_tout();
    bool output_enabled() const;
    std::string file_name();
    void save(binary_io::base_t &bio, double dt_seed);
    void load(const binary_io::base_t &bio, double &dt_seed);
    // SIGINT (cycle.cpp:32, :199-205): the handler init installs records
    // the signal; poll_signal, called by evolve between batches of cycles,
    // throws the reference's "Signal : <n>" exception, which evolve turns
    // into the final snapshot when final_output is set
    static void sighdl(int signum);
    void poll_signal();
};
}  // namespace cycle

namespace meshgen {
// Keys of the Cartesian generator (meshgen/meshgen.cpp:13-42).
// Static refinement: [refine_region_*] sections with level and x_min /
// x_max (or loc), meshgen.cpp:111-158.
struct base_t {
    virtual ~base_t() {}
    type::coord2_t x_min, x_max;
    type::idx_t n_global, n_block, n_logic;
    int n_dim = 0;
    std::vector<std::string> bnd_kinds;
    struct ref_region_t {
        int level = 0;
        int i[2][3];           // inclusive block-index box at that level
    };
    std::vector<ref_region_t> ref_regions;
    virtual void read(const input &args);
    double location(double x_logic, int axis) const {
        return x_min[axis] + x_logic * (x_max[axis] - x_min[axis]);
    }
    int search_blk_idx(double loc, int lvl, int axis) const;   // meshgen.cpp:355-374
};
}  // namespace meshgen

class mod_base_t {
public:
    mesh_t *p_mesh = nullptr;
    virtual ~mod_base_t() {}
    virtual void read(const input &) {}
    virtual void init(const input &, mesh_t &) {}
    virtual void init_cond(block::dual_t &) {}
    virtual void step(mesh_t &) {}
    virtual void finalize(mesh_t &) {}
    virtual int space_order() const { return 2; }
};

class mesh_t {
public:
    mesh_t();
    virtual ~mesh_t();
    mesh_t(const mesh_t &) = delete;
    mesh_t &operator=(const mesh_t &) = delete;

    template <class T>
    std::shared_ptr<T> enroll_device() {
        auto p = std::make_shared<T>();
        p_dev = p;
        return p;
    }
    template <class T>
    std::shared_ptr<T> enroll_comm() {
        auto p = std::make_shared<T>();
        p_com = p;
        return p;
    }
    template <class T>
    std::shared_ptr<T> enroll_binary_io() {
        auto p = std::make_shared<T>();
        p_bio = p;
        return p;
    }
    template <class T>
    std::shared_ptr<T> enroll_cycle() {
        auto p = std::make_shared<T>();
        p_cyc = p;
        return p;
    }
    template <class T>
    std::shared_ptr<T> enroll_mgen() {
        auto p = std::make_shared<T>();
        p_mgn = p;
        return p;
    }
    template <class T>
    std::shared_ptr<T> enroll_module(const int &i_init, const int &i_step) {
        auto p = std::make_shared<T>();
        // called through the base: generators may keep their overrides private
        std::shared_ptr<mod_base_t> m = p;
        m->p_mesh = this;
        reads_.push_back([m](const input &a) { m->read(a); });
        inits_[i_init] = [this, m](const input &a) { m->init(a, *this); };
        steps_[i_step] = [this, m]() { m->step(*this); };
        mods_.push_back(m);
        return p;
    }
    template <class T>
    std::shared_ptr<T> enroll_module() {
        const int i_i = inits_.empty() ? 0 : inits_.rbegin()->first + 1;
        const int i_s = steps_.empty() ? 0 : steps_.rbegin()->first + 1;
        return enroll_module<T>(i_i, i_s);
    }

    virtual void init(const input &args);
    virtual void step();
    virtual void evolve();
    // Snapshot of every block's slot 0, the time and the parameters in the
    // reference's binary container (mesh/mesh.cpp:99-111); restart by
    // passing the snapshot as a parameter file ([file] binary).
    virtual void save(const std::string &file_name);
    bool is_restart() const { return restart_; }
    void load_state();                     // blocks' state from the restart snapshot

    // blocks of this process, in the generator's order
    std::vector<block_t>::iterator begin() { return blocks_.begin(); }
    std::vector<block_t>::iterator end() { return blocks_.end(); }
    size_t n_blocks() const { return blocks_.size(); }
    block::dual_t &dual(size_t i) { return *duals_.at(i); }

    // device side
    kr_ctx *ctx() const { return ctx_; }
    void copy_d2h();                       // every block's slot 0 to its host copy
    double time() const { return p_cyc ? p_cyc->t : 0; }
    double dt() const;
    double divb_max() const;

    std::shared_ptr<comm::base_t> p_com;
    std::shared_ptr<cycle::base_t> p_cyc;
    std::shared_ptr<device::base_t> p_dev;
    std::shared_ptr<meshgen::base_t> p_mgn;
    std::shared_ptr<binary_io::base_t> p_bio;

    void init_dynamics(int model, int precision, double gamma, double cfl, double rho_floor,
                       double tmp_floor);

    // The block tree (mesh/tree/tree.h:28-67): every node by (level, x, y,
    // z) in the tree's lexicographic order (geometry/region.h:17-33), with
    // the node's global id (-1 for a masked, refined-away parent) and mask.
    struct tree_node_t {
        int id_g = -1;
        bool mask = false;
    };
    typedef std::array<int, 4> tree_key_t;
    const std::map<tree_key_t, tree_node_t> &tree() const { return tree_; }

private:
    void build_blocks();
    void build_refined(const bool *periodic);
    void blocks_from_snapshot();
    void blocks_from_tree(const std::string &bytes);
    void set_tree_meta(const bool *periodic);
    void write_tree(binary_io::base_t &bio) const;
    void pull_time();
    std::vector<std::function<void(const input &)>> reads_;
    std::map<int, std::function<void(const input &)>> inits_;
    std::map<int, std::function<void()>> steps_;
    std::vector<std::shared_ptr<mod_base_t>> mods_;
    std::vector<block_t> blocks_;
    std::vector<std::unique_ptr<block::dual_t>> duals_;
    kr_ctx *ctx_ = nullptr;
    bool mhd_ = false;
    double dt_last_ = 0;
    bool restart_ = false;
    std::string restart_file_;
    std::map<std::string, std::string> input_args_;
    std::map<tree_key_t, tree_node_t> tree_;
    int tree_n_dim_ = 0;
    int tree_reg_max_[3] = {0, 0, 0};
    bool tree_periodic_[3] = {false, false, false};
    double fix_w_[5] = {0, 0, 0, 0, 0}, fix_b_[3] = {0, 0, 0};   // "fixd" ghost state
};

// Throws std::runtime_error carrying the C-ABI message when rc != KR_OK.
void check(int rc, kr_ctx *ctx, const char *what);

}  // namespace mesh
