// Sod shock tube as a problem generator of the host API: a pgen written
// the way the reference's generators are (derive from dynamics::hydro_t,
// fill init_cond on the host, enrol and evolve), with the reference's
// include paths.  [prob] dump names an optional raw file that receives the
// initial and the final state of every block (slot 0, AoS doubles).
#include "../../src/device/general.h"
#include "../../src/mesh/mesh_enroll.h"
#include "../../src/modules/dynamics/hydro/hydro.h"
#include "../../src/types.h"
#include "../../src/user/probgen.h"
#include "../../src/utilities/mapping/loop.h"

#include <cstdio>

namespace prob {
class hydro_t : public dynamics::hydro_t {
    type::float2_t rho_l, pre_l, rho_r, pre_r, x_0, gam;

    void read(const input &args) override {
        rho_l = args.get<type::float2_t>("prob", "rho_l", 1.0);
        pre_l = args.get<type::float2_t>("prob", "pre_l", 1.0);
        rho_r = args.get<type::float2_t>("prob", "rho_r", 0.125);
        pre_r = args.get<type::float2_t>("prob", "pre_r", 0.1);
        x_0 = args.get<type::float2_t>("prob", "x_0", 0.5);
        gam = args.get<type::float2_t>("dynamics", "gamma", 1.4);
        dynamics::hydro_t::read(args);
        return enroll<>();
    }

    void init_cond(mesh::block::dual_t &d) override {
        using prx_t = dynamics::hydro::proxy_t;
        auto &b_h = prx_t::ref(d.h());
        const auto &geo = d.b_w.get().geo;
        for (auto i : utils::loop(geo.n_ceff)) {
            auto *u = b_h.hyd.u.at(i);
            const bool left = geo.x_cc(0, i[0]) < x_0;
            u[idx::rho] = left ? rho_l : rho_r;
            u[idx::ene] = (left ? pre_l : pre_r) / (gam - 1);
            u[idx::mom] = u[idx::mom + 1] = u[idx::mom + 2] = 0;
        }
    }
};

static void dump(mesh::mesh_t &mesh, FILE *f) {
    mesh.copy_d2h();
    for (size_t b = 0; b < mesh.n_blocks(); ++b) {
        auto &h = mesh.dual(b).h();
        fwrite(h.hyd.u_data.data(), sizeof(double), h.hyd.u_data.size(), f);
    }
}

void run(int argc, char *argv[]) {
    mesh::mesh_t mesh;
    mesh.enroll_device<device_t>();
    mesh.enroll_module<hydro_t>();
    input args;
    args.set_comm(mesh.p_com);
    cmd(argc, argv, args);
    mesh.init(args);
    const std::string out = args.get<std::string>("prob", "dump", "");
    FILE *f = out.empty() ? nullptr : fopen(out.c_str(), "wb");
    if (f) dump(mesh, f);
    mesh.evolve();
    if (f) {
        dump(mesh, f);
        fclose(f);
    }
}
}  // namespace prob
