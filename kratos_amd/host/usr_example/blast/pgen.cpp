// 3-D MHD blast on a statically refined mesh, as a problem generator of the
// host API (derives from dynamics::mhd_t, reference include paths).  The
// ambient medium and the uniform field come from [init] (rho0, pre0, b0) via
// the base class; cells whose centre lies within [prob] r_in of the origin
// get the pressure [prob] p_in.  The refinement comes from the parameter
// file's [refine_region_*] sections.  [prob] dump names an optional raw file
// that receives the initial and the final state of every block (slot 0: u
// AoS doubles, then the face field with its ghost layers).
#include "../../src/device/general.h"
#include "../../src/mesh/mesh_enroll.h"
#include "../../src/modules/dynamics/mhd/mhd.h"
#include "../../src/types.h"
#include "../../src/user/probgen.h"
#include "../../src/utilities/mapping/loop.h"

#include <cstdio>

namespace prob {
class mhd_t : public dynamics::mhd_t {
    type::float2_t gam, p_in, r_in;

    void read(const input &args) override {
        gam = args.get<type::float2_t>("dynamics", "gamma", 5.0 / 3.0);
        p_in = args.get<type::float2_t>("prob", "p_in", 10.0);
        r_in = args.get<type::float2_t>("prob", "r_in", 0.1);
        dynamics::mhd_t::read(args);
        return enroll<>();
    }

    void init_cond(mesh::block::dual_t &d) override {
        dynamics::mhd_t::init_cond(d);
        auto &b_h = dynamics::mhd::proxy_t::ref(d.h());
        const auto &geo = d.b_w.get().geo;
        for (auto i : utils::loop(geo.n_ceff)) {
            const auto x = geo.x_cc(i);
            if (x[0] * x[0] + x[1] * x[1] + x[2] * x[2] < r_in * r_in)
                b_h.hyd.u.at(i)[idx::ene] += (p_in - pre0) / (gam - 1);
        }
    }
};

static void dump(mesh::mesh_t &mesh, FILE *f) {
    mesh.copy_d2h();
    for (size_t b = 0; b < mesh.n_blocks(); ++b) {
        auto &h = mesh.dual(b).h();
        fwrite(h.hyd.u_data.data(), sizeof(double), h.hyd.u_data.size(), f);
        fwrite(h.fld.bf_data.data(), sizeof(double), h.fld.bf_data.size(), f);
    }
}

void run(int argc, char *argv[]) {
    mesh::mesh_t mesh;
    mesh.enroll_device<device_t>();
    mesh.enroll_module<mhd_t>();
    input args;
    args.set_comm(mesh.p_com);
    cmd(argc, argv, args);
    mesh.init(args);
    std::printf("blocks = %zu\n", mesh.n_blocks());
    const std::string out = args.get<std::string>("prob", "dump", "");
    FILE *f = out.empty() ? nullptr : fopen(out.c_str(), "wb");
    if (f) dump(mesh, f);
    mesh.evolve();
    std::printf("max |div B| = %.3e\n", mesh.divb_max());
    if (f) {
        dump(mesh, f);
        fclose(f);
    }
}
}  // namespace prob
