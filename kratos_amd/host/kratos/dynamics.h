// dynamics.h -- the hydro and MHD modules a problem generator derives from.
//
// dynamics::hydro_t (modules/dynamics/hydro/hydro.h:28-57, hydro.cpp:10-45)
// is PLM + HLLC + Heun RK2; dynamics::mhd_t (mhd/mhd.h:25-58, mhd.cpp:10-54)
// is PLM + HLLD + CT + Heun RK2.  enroll<>() selects the default policies;
// this build provides exactly those, so a generator that names other
// policies fails to compile with a message.  The equation-of-state keys
// are the reference's [dynamics] gamma / cfl / rho_floor / tmp_floor with
// its defaults (prototypes/eos.h:43-56).
#pragma once

#include <cmath>

#include "mesh.h"

namespace dynamics {

namespace hydro {
struct proxy_t {
    static mesh::block::block_data_t &ref(mesh::block::block_data_t &d) { return d; }
};
}  // namespace hydro
namespace mhd {
typedef hydro::proxy_t proxy_t;
}  // namespace mhd

struct eos_par_t {
    double gamma = 1.66667, cfl = 0.3, rho_floor = 1e-9, tmp_floor = 1e-9;
};

class base_t : public mesh::mod_base_t {
public:
    std::shared_ptr<eos_par_t> p_eos = std::make_shared<eos_par_t>();

    void read(const input &args) override {
        p_eos->gamma = args.get<double>("dynamics", "gamma", 1.66667);
        p_eos->cfl = args.get<double>("dynamics", "cfl", 0.3);
        p_eos->rho_floor = args.get<double>("dynamics", "rho_floor", 1e-9);
        p_eos->tmp_floor = args.get<double>("dynamics", "tmp_floor", 1e-9);
    }
    void init(const input &, mesh::mesh_t &mesh) override {
        if (!enrolled_) enroll_default();
        mesh.init_dynamics(model(), PRECISION == 1 ? KR_MIXED : KR_FP64, p_eos->gamma,
                           p_eos->cfl, p_eos->rho_floor, p_eos->tmp_floor);
        if (mesh.is_restart()) {
            mesh.load_state();
            return;
        }
        for (size_t b = 0; b < mesh.n_blocks(); ++b) {
            auto &d = mesh.dual(b);
            init_cond(d);
            d.copy_h2d();
        }
    }
    virtual void enroll_default() { enrolled_ = true; }
    virtual int model() const = 0;

protected:
    bool enrolled_ = false;
};

// Default policy tags of enroll<> (the only ones this build provides).
struct default_policy_t {};

class hydro_t : public base_t {
public:
    void read(const input &args) override {
        rho0 = args.get<type::float2_t>("init", "rho0", 0);
        pre0 = args.get<type::float2_t>("init", "pre0", 0);
        vel0 = args.get<type::coord2_t>("init", "vel0", 0);
        base_t::read(args);
    }
    template <class... policy_T>
    void enroll() {
        static_assert(sizeof...(policy_T) == 0,
                      "only the default policies (PLM, HLLC/HLLD, CT, RK2, adiabatic EOS) are built");
        enrolled_ = true;
    }
    void init_cond(mesh::block::dual_t &d) override {
        auto &b_h = hydro::proxy_t::ref(d.h());
        const auto &geo = d.b_w.get().geo;
        for (auto i : utils::loop(geo.n_ceff)) {
            auto *u = b_h.hyd.u.at(i);
            u[idx::rho] = rho0;
            u[idx::ene] = pre0 / (p_eos->gamma - 1);
            for (int n = 0; n < 3; ++n) {
                u[idx::ene] += std::pow(vel0[n], 2) * rho0 / 2;
                u[idx::mom + n] = vel0[n] * rho0;
            }
        }
    }
    int model() const override { return KR_HYDRO; }

protected:
    type::float2_t rho0 = 0, pre0 = 0;
    type::coord2_t vel0;
};

class mhd_t : public hydro_t {
public:
    void read(const input &args) override {
        b0 = args.get<type::coord2_t>("init", "b0", 0);
        hydro_t::read(args);
    }
    void init_cond(mesh::block::dual_t &d) override {
        hydro_t::init_cond(d);
        auto &b_h = mhd::proxy_t::ref(d.h());
        const auto &geo = d.b_w.get().geo;
        type::idx_t n = geo.n_ceff;
        for (int a = 0; a < 3; ++a) n[a] += n[a] > 1;
        for (auto i : utils::loop(n)) {
            auto *bf = b_h.fld.bf.at(i);
            for (int a = 0; a < 3; ++a) bf[a] = b0[a];
            bool cc = true;
            for (int a = 0; a < 3; ++a) cc &= i[a] < geo.n_ceff[a];
            if (cc) {
                auto *u = b_h.hyd.u.at(i);
                for (int a = 0; a < 3; ++a) u[idx::ene] += std::pow(b0[a], 2) / 2;
            }
        }
    }
    int model() const override { return KR_MHD; }

protected:
    type::coord2_t b0;
};

}  // namespace dynamics
