// 3-D Orszag-Tang-type vortex as a problem generator of the host API
// (derives from dynamics::mhd_t, reference include paths).  The face field
// is the discrete curl of an edge vector potential A_z(x, y), so div B = 0
// to roundoff from the start.  [prob] dump names an optional raw file that
// receives the initial and the final state of every block (slot 0: u AoS
// doubles, then the face field with its ghost layers).
#include "../../src/device/general.h"
#include "../../src/mesh/mesh_enroll.h"
#include "../../src/modules/dynamics/mhd/mhd.h"
#include "../../src/types.h"
#include "../../src/user/probgen.h"
#include "../../src/utilities/mapping/loop.h"

#include <cmath>
#include <cstdio>

namespace prob {
class mhd_t : public dynamics::mhd_t {
    type::float2_t gam, v_z;

    void read(const input &args) override {
        gam = args.get<type::float2_t>("dynamics", "gamma", 5.0 / 3.0);
        v_z = args.get<type::float2_t>("prob", "v_z", 0.1);
        dynamics::mhd_t::read(args);
        return enroll<>();
    }

    void init_cond(mesh::block::dual_t &d) override {
        using prx_t = dynamics::mhd::proxy_t;
        auto &b_h = prx_t::ref(d.h());
        const auto &geo = d.b_w.get().geo;
        const double pi = std::acos(-1.0), tp = 2 * pi;
        const double b0 = 1 / std::sqrt(4 * pi);
        const double rho = 25 / (36 * pi), pre = 5 / (12 * pi);
        auto a_z = [&](double x, double y) {
            return b0 / (4 * pi) * std::cos(2 * tp * x) + b0 / tp * std::cos(tp * y);
        };
        const auto n = geo.n_ceff;
        // faces: bx on (i, j, k) for i <= n0, by for j <= n1, bz = 0
        for (auto i : utils::loop(type::idx_t(n[0] + 1, n[1] + 1, n[2] + 1))) {
            auto *bf = b_h.fld.bf.at(i);
            const double xf = geo.x_fc(0, i[0]), yf = geo.x_fc(1, i[1]);
            const double xf1 = geo.x_fc(0, i[0] + 1), yf1 = geo.x_fc(1, i[1] + 1);
            bf[0] = (a_z(xf, yf1) - a_z(xf, yf)) / geo.dx0[1];
            bf[1] = -(a_z(xf1, yf) - a_z(xf, yf)) / geo.dx0[0];
            bf[2] = 0;
        }
        for (auto i : utils::loop(n)) {
            auto *u = b_h.hyd.u.at(i);
            const double x = geo.x_cc(0, i[0]), y = geo.x_cc(1, i[1]), z = geo.x_cc(2, i[2]);
            const double v[3] = {-std::sin(tp * y), std::sin(tp * x), v_z * std::sin(tp * z)};
            type::idx_t ix = i, iy = i;
            ++ix[0];
            ++iy[1];
            const double bx = 0.5 * (b_h.fld.bf.at(i)[0] + b_h.fld.bf.at(ix)[0]);
            const double by = 0.5 * (b_h.fld.bf.at(i)[1] + b_h.fld.bf.at(iy)[1]);
            u[idx::rho] = rho;
            u[idx::ene] = pre / (gam - 1) +
                          0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) +
                          0.5 * (bx * bx + by * by);
            for (int a = 0; a < 3; ++a) u[idx::mom + a] = rho * v[a];
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
