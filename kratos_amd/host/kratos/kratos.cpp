// kratos.cpp -- mesh_t, cycle and block generator of the host layer, main()
// and prob::cmd.  See mesh.h for what each piece restates.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>

#include "dynamics.h"
#include "probgen.h"

namespace mesh {

void check(int rc, kr_ctx *ctx, const char *what) {
    if (rc == KR_OK) return;
    const char *m = ctx ? kr_last_error(ctx) : nullptr;
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
        for (int a = 0; a < 3; ++a) d.fld.N[a] = b.geo.n_ceff[a] + 5;
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
void cycle::base_t::init(const input &args) {
    args("cycle", "t_0", t);
    args("cycle", "t_lim", t_lim);
    args("cycle", "dt_init", dt_init);
    args("cycle", "dt_expand", dt_expand);
    args("cycle", "n_cycle_lim", n_cycle_lim);
    top_up = args.get<bool>("cycle", "t_lim_top_up", false);
    args("cycle", "t_display_prec", t_display_prec);
    args("cycle", "n_display_freq", n_display_freq);
    if (n_display_freq < 1) n_display_freq = 1;
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
    if (args.found("file", "binary"))
        throw std::runtime_error("restart from a binary snapshot is not supported by this build");
    p_cyc->init(args);
    p_mgn->read(args);
    build_blocks();
    for (auto &p : inits_) p.second(args);
    if (!ctx_) throw std::runtime_error("no dynamics module enrolled");
}

// Blocks of a same-level Cartesian mesh: order, geometry and the 26
// neighbours with periodic wrap (meshgen.cpp:286-321, tree.cpp:77-84,
// 177-197).  A thin axis (one cell) has no neighbours along it.
void mesh_t::build_blocks() {
    const auto &g = *p_mgn;
    bool periodic[3];
    for (int a = 0; a < 3; ++a) periodic[a] = g.bnd_kinds[2 * a].find("per") != std::string::npos;
    blocks_.clear();
    const int nb = g.n_logic.prod();
    blocks_.resize(nb);
    for (int b = 0; b < nb; ++b) {
        block_t &k = blocks_[b];
        k.id = b;
        k.reg.reg = type::idx_t(b % g.n_logic[0], (b / g.n_logic[0]) % g.n_logic[1],
                                b / (g.n_logic[0] * g.n_logic[1]));
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
                    k.nbr[((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1)] =
                        ok ? (q[2] * g.n_logic[1] + q[1]) * g.n_logic[0] + q[0] : -1;
                }
    }
}

static int bnd_code(const std::string &k) {
    if (k.find("per") != std::string::npos) return KR_PER;
    if (k == "out") return KR_OUT;
    if (k == "fre") return KR_FRE;
    if (k == "ref") return KR_REF;
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

void mesh_t::save(const std::string &file_name) {
    throw std::runtime_error("snapshot output is not supported by this build: " + file_name);
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
    try {
        // the reference's loop runs n_cycle_lim - i_cycle + 1 steps unless
        // t reaches t_lim first; i_cycle stops at n_cycle_lim
        long long left = (long long)c.n_cycle_lim - c.i_cycle + 1;
        while (left > 0) {
            const int batch = (int)std::min<long long>(left, c.n_display_freq);
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
    }
    check(kr_sync(ctx_), ctx_, "kr_sync");
    const double dwt = std::chrono::duration<double>(clock::now() - start).count();
    const double speed = n_cell_tot / dwt * (c.i_cycle - c.i_cycle_0 + 1);
    if (p_com->is_root())
        std::cout << "Duration = " << dwt << " s, " << "Speed = " << speed << " cell/s,\n"
                  << "Pace = " << 1 / speed << " s/cell.\n";
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
