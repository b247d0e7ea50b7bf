/*
 * kratos_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A plain C++17 FP64 restatement of the reference's Godunov hydro/MHD stage
 * with PRECISION=2 semantics, written from the reference's source as read
 * (not copied).  It mirrors the reference's block structure: per-block AoS
 * arrays with ghost shells, a same-level ghost exchange, per-axis pencil
 * sweeps (with the MHD transverse ghost lines), the corner-upwind edge EMF,
 * constrained transport and the Heun RK2 slots.  Operation order follows
 * the reference expression by expression so that the FP64 GPU form can be
 * compared at round-off level.  Build with -ffp-contract=off.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this file, and only as the checker.
 *
 * Scope: same-level (uniform) block meshes; hydro in 1-3D with periodic /
 * outflow / free / reflect sides; MHD in 3D with periodic sides.
 */
#include "kratos_oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

enum { RHO = 0, ENE = 1, MOM = 2, BCC = 5 };   // dynamics/indices.h:7-13

thread_local std::string g_err;

inline int mod3(int i) { return i < 3 ? i : i - 3; }     // mapping.h:219-223
inline double dmax(double a, double b) { return a > b ? a : b; } // functions.h:51-54
inline double dmin(double a, double b) { return a < b ? a : b; } // functions.h:61-64
template <class T> inline T tmin(T a, T b) { return a < b ? a : b; } // functions.h:26-30
template <class T> inline T tmax(T a, T b) { return a > b ? a : b; } // functions.h:20-24
inline double tabs(double a) { return a > 0 ? a : -a; }  // functions.h:32-36
inline double minmod(double a, double b) {                // functions.h:159-170
    return ((a < 0) ^ (b < 0)) ? 0 : (std::fabs(a) < std::fabs(b) ? a : b);
}
inline int sgn(double a) { return a > 0 ? 1 : (a < 0 ? -1 : 0); } // functions.h:194-198

// utils::idx_swap (utilities/mapping/mapping.h:225-237)
inline void idx_swap(double *x, int axis) {
    if (mod3(axis) != 0) {
        double s[3];
        for (int i = 0; i < 3; ++i) s[i] = x[mod3(i + axis)];
        for (int i = 0; i < 3; ++i) x[i] = s[i];
    }
}

/* ------------------------------------------------------------------ */
/* AoS 3-D array with ghosts and RK slots (mesh/types/dat_3d.h)        */
struct Arr {
    std::vector<double> d;
    int nint = 0, nstep = 0, N[3] = {0, 0, 0}, g[3] = {0, 0, 0};
    size_t nctot = 0;
    void init(const int n[3], const int gh[3], int nint_, int nstep_) {
        nint = nint_; nstep = nstep_;
        nctot = (size_t)nint;
        for (int a = 0; a < 3; ++a) {
            g[a] = gh[a];
            N[a] = n[a] + 2 * gh[a];
            nctot *= (size_t)N[a];
        }
        d.assign(nctot * nstep, 0.0);
    }
    double *at(int i, int j, int k, int s = 0) {
        return d.data() + ((size_t)((k + g[2]) * N[1] + (j + g[1])) * N[0]
                           + (i + g[0])) * nint + (size_t)s * nctot;
    }
    const double *at(int i, int j, int k, int s = 0) const {
        return d.data() + ((size_t)((k + g[2]) * N[1] + (j + g[1])) * N[0]
                           + (i + g[0])) * nint + (size_t)s * nctot;
    }
    double *at(const int i[3], int s = 0) { return at(i[0], i[1], i[2], s); }
    const double *at(const int i[3], int s = 0) const {
        return at(i[0], i[1], i[2], s);
    }
};

struct Cfg {
    int mhd;
    double gamma, cfl, d_floor, t_floor;
    int ng[3], nb[3];
    double xmin[3], xmax[3];
    int bnd[6];
    double dt_init, dt_expand, t_lim;
    bool top_up;
    int nthreads;
    double fix_w[5], fix_b[3];     // KRO_FIX ghost state (keeper.cpp:60-72)
};

struct Block {
    int reg[3];
    int level = 0;
    int n[3];
    int ndim;
    int gw[3];                 // ghosts of w along each axis
    double xf0[3], dx0[3], sfc0[3], vol0;
    bool phy[3][2];
    // SMR: level of the leaf covering the neighbour region in direction
    // ((dz+1)*3 + (dy+1))*3 + (dx+1), relative to this block (-1 coarser,
    // 0 same, +1 finer; kOut outside a non-periodic domain).
    int lev_nb[27];
    Arr u, w, f[3], bf, ec, ee;
};
constexpr int kOut = 9;
inline int dir_index(const int d[3]) { return ((d[2] + 1) * 3 + (d[1] + 1)) * 3 + (d[0] + 1); }

/* ------------------------------------------------------------------ */
/* Equation of state (hydro/eos/adiabatic.h)                           */
struct Eos {
    double gamma, cfl, d_floor, t_floor;
    int nhyd;                  // 5 hydro, 8 MHD (mhd/eos/adiabatic.h:24)

    // w_prho_d / w_erho_d (hydro/eos/adiabatic.h:89-105)
    void coefs(double &cv_kb, double &gam) const {
        gam = gamma;
        cv_kb = 1 / (gam - 1);
    }
    // et_cs (hydro/eos/adiabatic.h:62-76)
    double et_cs(double &e_tot, double &c_s, const double *w) const {
        const double pre = w[ENE], rho = w[RHO];
        double cv_kb, gam;
        coefs(cv_kb, gam);
        double v2 = 0;
        for (int a = 0; a < 3; ++a) v2 += w[MOM + a] * w[MOM + a];
        e_tot = pre * cv_kb + rho * v2 / 2;
        c_s = std::sqrt(gam * dmax(pre / rho, 0));
        return cv_kb;
    }
    // pre_cs2 (hydro/eos/adiabatic.h:77-85)
    double pre_cs2(double *w, double e_g) const {
        double cv_kb, gam;
        coefs(cv_kb, gam);
        w[ENE] = e_g / cv_kb;
        return gam * w[ENE] / w[RHO];
    }
    // c2p (hydro/eos/adiabatic.h:159-206)
    double c2p(double *prim, double *cons, double *p_src, double *c_src) const {
        for (int n = 0; n < nhyd; ++n) cons[n] = c_src[n];
        double &e_t = cons[ENE];
        double &rho = cons[RHO];
        if (std::isnan(rho) || rho < d_floor) {
            rho = d_floor;
            c_src[RHO] = d_floor;
        }
        prim[RHO] = rho;
        const double rho_inv = 1 / rho;
        double e_k = 0.;
        for (int n = MOM; n < MOM + 3; ++n) {
            prim[n] = rho_inv * cons[n];
            e_k += cons[n] * cons[n];
        }
        e_k *= rho_inv / 2;
        const double e_g = e_t - e_k;
        double cs2 = pre_cs2(prim, e_g);
        double &pre = prim[ENE];
        const double pre_floor = t_floor * prim[RHO];
        if (std::isnan(pre) || pre < pre_floor) {
            pre = pre_floor;
            c_src[ENE] = e_k + pre / (gamma - 1);
            cs2 = gamma * pre / rho;
        }
        for (int n = 0; n < nhyd; ++n) p_src[n] = prim[n];
        return cs2;
    }
    // hydro dt (hydro/eos/adiabatic.h:144-157)
    double dt_hydro(const double *v_p, double cs2, const Block &b) const {
        double dt = (double)FLT_MAX;
        const double cs = std::sqrt(cs2);
        for (int a = 0; a < b.ndim; ++a)
            dt = dmin(dt, b.dx0[a] / (cs + std::fabs(v_p[MOM + a])));
        return dt * cfl;
    }
    // c_fms (mhd/eos/adiabatic.h:82-98)
    double c_fms(const double *w, const double *b_2, double gam, int axis) const {
        const double rho = w[RHO];
        const double cs_2 = w[ENE] * gam / rho;
        const double v_ax2 = b_2[axis] / rho;
        const double v_t2 = (b_2[mod3(axis + 1)] + b_2[mod3(axis + 2)]) / rho;
        const double q_2 = v_ax2 + v_t2 + cs_2;
        const double r_2 = v_ax2 + v_t2 - cs_2;
        return std::sqrt(0.5f * (q_2 + std::sqrt(r_2 * r_2 + 4 * cs_2 * v_t2)));
    }
    // MHD dt (mhd/eos/adiabatic.h:103-122)
    double dt_mhd(const double *v_p, const double *bc_2, const double *b_m,
                  const Block &b) const {
        double dt = (double)FLT_MAX;
        double bm_2[3] = {bc_2[0], bc_2[1], bc_2[2]};
        for (int a = 0; a < b.ndim; ++a) {
            bm_2[a] = b_m[a] * b_m[a];
            const double c = c_fms(v_p, bm_2, gamma, a);
            dt = dmin(dt, b.dx0[a] / (c + std::fabs(v_p[MOM + a])));
            bm_2[a] = bc_2[a];
        }
        return dt * cfl;
    }
};

/* ------------------------------------------------------------------ */
/* PLM, uniform mesh (hydro/reconstruct/plm.h:116-153), FP64: norm = 1 */
inline void plm_uniform(double q_ll, double q_l, double q_r, double q_rr,
                        double &ql, double &qr) {
    const double dql = (q_l - q_ll);
    const double dqc = (q_r - q_l);
    const double dqr = (q_rr - q_r);
    const double norm = 1;
    const double dqr2 = norm * dqr * dqr;
    const double dql2 = norm * dql * dql;
    const double dqc2 = norm * dqc * dqc;
    const double dq2_l = norm * dql * dqc;
    const double dq2_r = norm * dqr * dqc;
    const double ql_coef = dq2_l <= 0 ? 0 : dq2_l / (dql2 + dqc2 + dq2_l * 2);
    const double qr_coef = dq2_r <= 0 ? 0 : -dq2_r / (dqr2 + dqc2 + dq2_r * 2);
    ql = q_l + (dq2_l <= 0 ? 0 : ql_coef * (dql + dqc));
    qr = q_r + (dq2_r <= 0 ? 0 : qr_coef * (dqc + dqr));
}

/* ------------------------------------------------------------------ */
/* HLLC (hydro/riemann/hllc.h:53-64, 76-193), rotated frame            */
inline void f_lr_hyd(double *f, const double *w, double v, double e) {
    const double rho = w[RHO];
    const double pre = w[ENE];
    f[RHO] = rho * v;
    f[ENE] = pre * w[MOM] + e * v;
    for (int a = MOM; a < MOM + 3; ++a) f[a] = rho * w[a] * v;
    f[MOM] += pre;
}

void hllc_eval(double *flx, const double *wli, const double *wri, const Eos &eos) {
    double el, er, cl, cr;
    const double gm1i_l = eos.et_cs(el, cl, wli);
    const double gm1i_r = eos.et_cs(er, cr, wri);
    const double sqrtdl = std::sqrt(wli[RHO]);
    const double sqrtdr = std::sqrt(wri[RHO]);
    const double isdlpdr = (double)1 / (sqrtdl + sqrtdr);
    double wroe[5];
    wroe[RHO] = sqrtdl * sqrtdr;
    for (int n = MOM; n < MOM + 3; ++n)
        wroe[n] = (sqrtdl * wli[n] + sqrtdr * wri[n]) * isdlpdr;
    const double hroe = isdlpdr * ((el + wli[ENE]) / sqrtdl +
                                   (er + wri[ENE]) / sqrtdr);
    const double gam1_roe = (double)1 /
        (isdlpdr * (sqrtdl * gm1i_l + sqrtdr * gm1i_r));
    const double v2roe = wroe[MOM] * wroe[MOM] + wroe[MOM + 1] * wroe[MOM + 1]
                       + wroe[MOM + 2] * wroe[MOM + 2];
    const double q = hroe - 0.5f * v2roe;
    const double a = (q <= 0) ? 0 : std::sqrt(gam1_roe * q);
    const double al = dmin(wroe[MOM] - a, wli[MOM] - cl);
    const double ar = dmax(wroe[MOM] + a, wri[MOM] + cr);
    const double bp = ar > 0 ? ar : 0;
    const double bm = al < 0 ? al : 0;
    const double vxl = wli[MOM] - al;
    const double vxr = wri[MOM] - ar;
    const double tl = wli[ENE] + vxl * wli[RHO] * wli[MOM];
    const double tr = wri[ENE] + vxr * wri[RHO] * wri[MOM];
    const double ml = wli[RHO] * vxl;
    const double mr = -wri[RHO] * vxr;
    double cp, am;
    const double f_max = (double)(FLT_MAX / 1e1f);
    const double f_min = (double)(FLT_MIN * 1e1f);
    if (std::fabs(ml + mr) < f_min) {
        cp = 0;
        am = 0;
    } else if ((tr != 0 && std::fabs(ml) > f_max / std::fabs(tr)) ||
               (tl != 0 && std::fabs(mr) > f_max / std::fabs(tl))) {
        const double mli = sgn(ml) / dmax(std::fabs(ml), f_min);
        const double mri = sgn(mr) / dmax(std::fabs(mr), f_min);
        am = (tl - tr) / (ml + mr);
        cp = dmax((mri * tr + mli * tl) / (mli + mri), (double)0);
    } else {
        am = (tl - tr) / (ml + mr);
        cp = dmax((ml * tr + mr * tl) / (ml + mr), (double)0);
    }
    const double d_am_bm = dmax(std::fabs(am - bm), f_min) * (am > bm ? 1 : -1);
    const double d_am_bp = dmax(std::fabs(am - bp), f_min) * (am > bp ? 1 : -1);
    const double sl = (am >= 0 ? am / d_am_bm : 0);
    const double sr = (am < 0 ? am / d_am_bp : 0);
    const double sm = (am > 0 ? -bm / d_am_bm : -bp / d_am_bp);
    double fl[5], fr[5];
    f_lr_hyd(fl, wli, wli[MOM] - bm, el);
    f_lr_hyd(fr, wri, wri[MOM] - bp, er);
    for (int n = RHO; n < MOM + 3; ++n) flx[n] = sl * fl[n] + sr * fr[n];
    flx[MOM] += sm * cp;
    flx[ENE] += sm * cp * am;
}

/* ------------------------------------------------------------------ */
/* HLLD (mhd/riemann/hlld.h:49-69, 81-298), rotated frame              */
inline void f_flux(double *flux, const double *u, const double *w,
                   double p_tot, double b_x) {
    const double bx_2 = b_x * b_x;
    flux[RHO] = u[MOM];
    flux[ENE] = w[MOM] * (p_tot + u[ENE] - bx_2) - b_x *
                (w[MOM + 1] * u[BCC + 1] + w[MOM + 2] * u[BCC + 2]);
    flux[MOM] = p_tot - bx_2 + u[MOM] * w[MOM];
    flux[MOM + 1] = -b_x * u[BCC + 1] + u[MOM + 1] * w[MOM];
    flux[MOM + 2] = -b_x * u[BCC + 2] + u[MOM + 2] * w[MOM];
    flux[BCC + 1] = -b_x * w[MOM + 1] + u[BCC + 1] * w[MOM];
    flux[BCC + 2] = -b_x * w[MOM + 2] + u[BCC + 2] * w[MOM];
}

void hlld_eval(double *flx, const double *wli, const double *wri, const Eos &eos) {
    flx[BCC] = 0;
    double spd_[5];
    double *spd = spd_ + 2;
    const int n_wave = 8;
    double u_l[8], u_r[8];
    const double b_x = wli[BCC];
    const double bx_2 = b_x * b_x;
    auto f_cons = [&](double *u, const double *w, double &gam1) {
        double p_m = bx_2 / 2, cs;
        for (int a = 1; a <= 2; ++a) {
            const int n = BCC + a;
            p_m += w[n] * w[n] / 2;
            u[n] = w[n];
        }
        gam1 = 1 / eos.et_cs(u[ENE], cs, w);
        for (int n = MOM; n < MOM + 3; ++n) u[n] = w[RHO] * w[n];
        u[RHO] = w[RHO];
        u[ENE] += p_m;
        u[BCC] = b_x;
        return p_m;
    };
    double gam1_l, gam1_r;
    const double pm_l = f_cons(u_l, wli, gam1_l);
    const double pm_r = f_cons(u_r, wri, gam1_r);
    const double ptot_l = wli[ENE] + pm_l;
    const double ptot_r = wri[ENE] + pm_r;
    double b2_eff_l[3] = {b_x, wli[BCC + 1], wli[BCC + 2]};
    double b2_eff_r[3] = {b_x, wri[BCC + 1], wri[BCC + 2]};
    for (int a = 0; a < 3; ++a) {
        b2_eff_l[a] *= b2_eff_l[a];
        b2_eff_r[a] *= b2_eff_r[a];
    }
    const double cfm_l = eos.c_fms(wli, b2_eff_l, gam1_l + 1, 0);
    const double cfm_r = eos.c_fms(wri, b2_eff_r, gam1_r + 1, 0);
    spd[-2] = tmin(wli[MOM] - cfm_l, wri[MOM] - cfm_r);
    spd[2] = tmax(wli[MOM] + cfm_l, wri[MOM] + cfm_r);
    if (spd[-2] >= 0 || spd[2] <= 0) {
        const bool is_left(spd[-2] >= 0);
        f_flux(flx, is_left ? u_l : u_r, is_left ? wli : wri,
               is_left ? ptot_l : ptot_r, b_x);
        return;
    }
    const double sd_l = spd[-2] - wli[MOM];
    const double sd_r = spd[2] - wri[MOM];
    spd[0] = (sd_r * u_r[MOM] - ptot_r - sd_l * u_l[MOM] + ptot_l)
           / (sd_r * u_r[RHO] - sd_l * u_l[RHO]);
    const double sdm_l = spd[-2] - spd[0];
    const double sdm_r = spd[2] - spd[0];
    const double rhos_l = u_l[RHO] * sd_l / sdm_l;
    const double rhos_r = u_r[RHO] * sd_r / sdm_r;
    const double rhos_l_sqrt = std::sqrt(rhos_l);
    const double rhos_r_sqrt = std::sqrt(rhos_r);
    spd[-1] = spd[0] - tabs(b_x) / rhos_l_sqrt;
    spd[1] = spd[0] + tabs(b_x) / rhos_r_sqrt;

    const double f_tiny = (double)(1.00e-4f);
    const double ptot_s = ptot_l + u_l[RHO] * sd_l * (sd_l - sdm_l);
    auto f_star = [&](double *u_star, const double *u, const double *w,
                      double rhos, double sd, double sdm, double p_tot) {
        u_star[RHO] = rhos;
        u_star[MOM] = rhos * spd[0];
        u_star[BCC] = b_x;
        const double denom = u[RHO] * sd * sdm - bx_2;
        if (tabs(denom) < f_tiny * ptot_s) {
            for (int a = 1; a <= 2; ++a) {
                u_star[MOM + a] = rhos * w[MOM + a];
                u_star[BCC + a] = u[BCC + a];
            }
        } else {
            const double tmp_v = b_x * (sd - sdm) / denom;
            const double tmp_b = (u[RHO] * sd * sd - bx_2) / denom;
            for (int a = 1; a <= 2; ++a) {
                u_star[MOM + a] = rhos * (w[MOM + a] - u[BCC + a] * tmp_v);
                u_star[BCC + a] = u[BCC + a] * tmp_b;
            }
        }
        double u_dot_b = 0, w_dot_b = 0;
        for (int a = 0; a < 3; ++a) {
            u_dot_b += u_star[MOM + a] * u_star[BCC + a];
            w_dot_b += w[MOM + a] * u[BCC + a];
        }
        u_dot_b /= rhos;
        u_star[ENE] = b_x * (w_dot_b - u_dot_b) - p_tot * w[MOM]
                    + sd * u[ENE] + ptot_s * spd[0];
        u_star[ENE] /= sdm;
        return u_dot_b;
    };
    double us_l[8], us_r[8];
    const double u_dot_b_l = f_star(us_l, u_l, wli, rhos_l, sd_l, sdm_l, ptot_l);
    const double u_dot_b_r = f_star(us_r, u_r, wri, rhos_r, sd_r, sdm_r, ptot_r);
    if (spd[-1] >= 0) {
        f_flux(flx, u_l, wli, ptot_l, b_x);
        for (int n = 0; n < n_wave; ++n) flx[n] += spd[-2] * (us_l[n] - u_l[n]);
        return;
    } else if (spd[1] <= 0) {
        f_flux(flx, u_r, wri, ptot_r, b_x);
        for (int n = 0; n < n_wave; ++n) flx[n] += spd[2] * (us_r[n] - u_r[n]);
        return;
    }
    double uss_l[8], uss_r[8];
    for (int n = 0; n < n_wave; ++n) {
        uss_l[n] = us_l[n];
        uss_r[n] = us_r[n];
    }
    if (bx_2 > ptot_s * f_tiny) {
        const int bx_sgn(b_x > 0 ? 1 : -1);
        const double rhos_norm_inv = 1.f / (rhos_l_sqrt + rhos_r_sqrt);
        for (int a = 1; a <= 2; ++a) {
            double tmp = rhos_norm_inv *
                (us_l[MOM + a] / rhos_l_sqrt + us_r[MOM + a] / rhos_r_sqrt +
                 bx_sgn * (us_r[BCC + a] - us_l[BCC + a]));
            uss_l[MOM + a] = uss_l[RHO] * tmp;
            uss_r[MOM + a] = uss_r[RHO] * tmp;
            tmp = rhos_norm_inv *
                (rhos_l_sqrt * us_r[BCC + a] + rhos_r_sqrt * us_l[BCC + a] +
                 bx_sgn * rhos_l_sqrt * rhos_r_sqrt *
                     (us_r[MOM + a] / rhos_r - us_l[MOM + a] / rhos_l));
            uss_l[BCC + a] = tmp;
            uss_r[BCC + a] = tmp;
        }
        const double tmp = spd[0] * b_x +
            (uss_l[MOM + 1] * uss_l[BCC + 1] + uss_l[MOM + 2] * uss_l[BCC + 2])
            / uss_l[RHO];
        uss_l[ENE] -= rhos_l_sqrt * bx_sgn * (u_dot_b_l - tmp);
        uss_r[ENE] += rhos_r_sqrt * bx_sgn * (u_dot_b_r - tmp);
    }
    if (spd[0] >= 0) {
        f_flux(flx, u_l, wli, ptot_l, b_x);
        const double d_v = spd[-1] - spd[-2];
        for (int n = 0; n < n_wave; ++n)
            flx[n] += -spd[-2] * u_l[n] - d_v * us_l[n] + spd[-1] * uss_l[n];
    } else {
        f_flux(flx, u_r, wri, ptot_r, b_x);
        const double d_v = spd[1] - spd[2];
        for (int n = 0; n < n_wave; ++n)
            flx[n] += -spd[2] * u_r[n] - d_v * us_r[n] + spd[1] * uss_r[n];
    }
}

/* Face flux in the mesh frame: hllc.h:32-43 (calc) / hlld.h:29-47 (calc).
 * ql / qr are overwritten (rotated), as the reference does in LDS.      */
void face_flux(bool mhd, int axis, double *ql, double *qr, double *flx,
               const Eos &eos) {
    idx_swap(ql + MOM, axis);
    idx_swap(qr + MOM, axis);
    if (mhd) {
        idx_swap(ql + BCC, axis);
        idx_swap(qr + BCC, axis);
        hlld_eval(flx, ql, qr, eos);
        idx_swap(flx + MOM, 3 - axis);
        const double ef[2] = {flx[BCC + 2], -flx[BCC + 1]};
        flx[BCC + mod3(axis + 1)] = ef[0];
        flx[BCC + mod3(axis + 2)] = ef[1];
        flx[BCC + axis] = 0;
    } else {
        hllc_eval(flx, ql, qr, eos);
        idx_swap(flx + MOM, 3 - axis);
    }
}

/* ------------------------------------------------------------------ */
/* Mesh                                                                */
struct Mesh {
    Cfg cfg;
    Eos eos;
    int nblk[3];
    int ndim;
    std::vector<Block> blocks;
    double t = 0, p_dt = 0, dt_cur = 0;
    bool dt_reduced = false;

    int nvar() const { return cfg.mhd ? 8 : 5; }
    int bid(const int r[3]) const { return (r[2] * nblk[1] + r[1]) * nblk[0] + r[0]; }

    template <class F> void par_blocks(F &&f) {
        const int nb = (int)blocks.size();
        int nt = cfg.nthreads > 0 ? cfg.nthreads
                                  : (int)std::max(1u, std::thread::hardware_concurrency());
        nt = std::min(nt, nb);
        if (nt <= 1) {
            for (int b = 0; b < nb; ++b) f(b);
            return;
        }
        std::atomic<int> next(0);
        std::vector<std::thread> th;
        for (int t_ = 0; t_ < nt; ++t_)
            th.emplace_back([&] {
                for (int b = next++; b < nb; b = next++) f(b);
            });
        for (auto &x : th) x.join();
    }

    void build(const std::vector<int> *levels = nullptr, const std::vector<int> *regs = nullptr) {
        ndim = 0;                                  // meshgen.cpp:23-27
        for (int a = 0; a < 3; ++a)
            if (cfg.ng[a] > 1) ndim = a + 1;
        for (int a = 0; a < 3; ++a) {
            if (cfg.nb[a] <= 0 || cfg.ng[a] % cfg.nb[a] != 0)
                throw std::runtime_error("Mesh size indivisible by sub-mesh.");
            nblk[a] = cfg.ng[a] / cfg.nb[a];
        }
        // MHD thin axes (one cell, mhd/integrator/integrator.h:221-227 "kept
        // last by convention") must be the trailing ones and one cell wide
        // over the whole mesh: they have no neighbour (keeper.h:59-99)
        // (the sweeps of both models run over the first b.ndim axes, which
        // counts the axes with more than one cell, geometry.cpp:46)
        for (int a = 1; a < 3; ++a)
            if (cfg.nb[a] >= 2 && cfg.nb[a - 1] < 2)
                throw std::runtime_error("thin axes must be the trailing ones");
        if (cfg.mhd) {
            if (cfg.nb[0] < 2) throw std::runtime_error("MHD needs at least two cells along x");
            for (int a = 1; a < 3; ++a)
                if (cfg.nb[a] < 2 && cfg.ng[a] != 1)
                    throw std::runtime_error("an MHD thin axis must be one cell wide");
        }
        if (!levels) {
            blocks.resize((size_t)nblk[0] * nblk[1] * nblk[2]);
            for (int bz = 0; bz < nblk[2]; ++bz)
                for (int by = 0; by < nblk[1]; ++by)
                    for (int bx = 0; bx < nblk[0]; ++bx) {
                        const int r[3] = {bx, by, bz};
                        init_block(blocks[bid(r)], 0, r);
                    }
        } else {
            blocks.resize(levels->size());
            for (size_t b = 0; b < levels->size(); ++b)
                init_block(blocks[b], (*levels)[b], &(*regs)[3 * b]);
            build_smr();
        }
        // cycle.cpp:163-166: seed of the device dt
        p_dt = cfg.dt_init < FLT_MIN ? (double)(FLT_MAX / 10) : cfg.dt_init;
        t = 0;
        dt_cur = cfg.dt_init;
    }

    void init_block(Block &b, int level, const int r[3]) {
        const int nv = nvar();
        b.ndim = 0;
        b.level = level;
        for (int a = 0; a < 3; ++a) {
            b.reg[a] = r[a];
            b.n[a] = cfg.nb[a];
            b.ndim += (b.n[a] > 1 ? 1 : 0);     // geometry.cpp:46
            b.gw[a] = b.n[a] > 1 ? 2 : 0;       // hyd_data.cpp:21-24
        }
        // meshgen.cpp:286-321 (uniform: loc(0), loc(1) - loc(0))
        for (int a = 0; a < 3; ++a) {
            auto loc = [&](double i) -> double {
                if (b.n[a] == 1) return i;
                i = (b.reg[a] + i / b.n[a]) / (double)(nblk[a] << level);
                return cfg.xmin[a] + i * (cfg.xmax[a] - cfg.xmin[a]);
            };
            b.xf0[a] = loc(0);
            b.dx0[a] = loc(1) - b.xf0[a];
        }
        // meshgen.cpp:263-283, 339-353
        for (int a = 0; a < 3; ++a) {
            double s = 1.;
            for (int m = 1; m < 3; ++m) s *= b.dx0[(a + m) % 3];
            b.sfc0[a] = s;
        }
        double v = 1.;
        for (int a = 0; a < 3; ++a) v *= b.dx0[a];
        b.vol0 = v;
        // meshgen.cpp:406-423 (physical-boundary flags)
        for (int a = 0; a < 3; ++a) {
            const bool per = cfg.bnd[2 * a] == KRO_PER;
            b.phy[a][0] = !per && b.reg[a] == 0;
            b.phy[a][1] = !per && b.reg[a] == (nblk[a] << level) - 1;
        }
        for (int d = 0; d < 27; ++d) b.lev_nb[d] = 0;
        const int zero[3] = {0, 0, 0}, one[3] = {1, 1, 1};
        b.u.init(b.n, zero, nv, 2);
        b.w.init(b.n, b.gw, nv, 1);
        for (int a = 0; a < b.ndim; ++a) {
            int nf[3] = {b.n[0], b.n[1], b.n[2]};
            nf[a] += 1;
            b.f[a].init(nf, one, nv, 1);
        }
        if (cfg.mhd) {                  // fld_data.cpp:18-26
            int nf[3];
            for (int a = 0; a < 3; ++a) nf[a] = b.n[a] + (b.n[a] > 1);
            const int g2[3] = {2, 2, 2};
            b.bf.init(nf, g2, 3, 2);
            b.ec.init(nf, g2, 3, 1);
            b.ee.init(nf, one, 3, 1);
        }
    }

    /* -------- static mesh refinement (SMR) ----------------------------
     * Blocks of one size at several levels, 2:1 balanced (tree.cpp:185-205
     * insert()).  The leaf covering any position is found through one
     * dense table per level (the tree_t lookup of tree.cpp:54-58).      */
    bool smr = false;
    int nlev = 1;
    std::vector<std::vector<int>> table;   // [level][logical index] -> block or -1

    size_t lidx(int L, const int r[3]) const {
        return ((size_t)r[2] * (size_t)(nblk[1] << L) + (size_t)r[1]) * (size_t)(nblk[0] << L)
               + (size_t)r[0];
    }
    int leaf(int L, const int r[3]) const {
        if (L < 0 || L >= nlev) return -1;
        for (int a = 0; a < 3; ++a)
            if (r[a] < 0 || r[a] >= (nblk[a] << L)) return -1;
        return table[L][lidx(L, r)];
    }
    // Wrap a global cell index of level L into the domain along periodic
    // axes (tree.cpp:77-84); false outside a non-periodic side.
    bool wrap(int L, int G[3]) const {
        for (int a = 0; a < 3; ++a) {
            const int N = (nblk[a] * cfg.nb[a]) << L;
            if (G[a] >= 0 && G[a] < N) continue;
            if (cfg.bnd[2 * a] != KRO_PER) return false;
            G[a] = ((G[a] % N) + N) % N;
        }
        return true;
    }
    // Leaf covering global cell G of level L (in the domain): the relative
    // level of that leaf (0, -1 coarser, +1 finer) and its block index.
    int cover(int L, const int G[3], int &blk) const {
        int r[3];
        for (int a = 0; a < 3; ++a) r[a] = G[a] / cfg.nb[a];
        if ((blk = leaf(L, r)) >= 0) return 0;
        for (int a = 0; a < 3; ++a) r[a] = (G[a] >> 1) / cfg.nb[a];
        if ((blk = leaf(L - 1, r)) >= 0) return -1;
        for (int a = 0; a < 3; ++a) r[a] = (2 * G[a]) / cfg.nb[a];
        if ((blk = leaf(L + 1, r)) >= 0) return 1;
        throw std::runtime_error("SMR: a ghost cell has no leaf within one level");
    }

    void build_smr() {
        smr = true;
        if (ndim != 3) throw std::runtime_error("SMR meshes are 3-D");
        for (int a = 0; a < 3; ++a)
            if (cfg.nb[a] < 4 || cfg.nb[a] % 2)
                throw std::runtime_error("SMR blocks need an even size >= 4 per axis");
        nlev = 0;
        for (auto &b : blocks) nlev = std::max(nlev, b.level + 1);
        table.assign(nlev, std::vector<int>());
        for (int L = 0; L < nlev; ++L)
            table[L].assign((size_t)(nblk[0] << L) * (nblk[1] << L) * (nblk[2] << L), -1);
        for (size_t bi = 0; bi < blocks.size(); ++bi) {
            const Block &b = blocks[bi];
            if (b.level < 0 || leaf(b.level, b.reg) != -1) {
                for (int a = 0; a < 3; ++a)
                    if (b.reg[a] < 0 || b.reg[a] >= (nblk[a] << b.level))
                        throw std::runtime_error("SMR: block outside the domain");
                if (b.level < 0 || table[b.level][lidx(b.level, b.reg)] != -1)
                    throw std::runtime_error("SMR: duplicate block");
            }
            table[b.level][lidx(b.level, b.reg)] = (int)bi;
        }
        // neighbour levels per direction (tree.cpp:177-197 neighbors())
        for (auto &b : blocks)
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int d[3] = {dx, dy, dz};
                        if (!dx && !dy && !dz) continue;
                        // the ghost cells in direction d (two deep across, the
                        // block's extent along) must lie on one level
                        int &lv = b.lev_nb[dir_index(d)];
                        for (int c = 0; c < 8; ++c) {
                            int H[3];
                            for (int a = 0; a < 3; ++a) {
                                const int o = (c >> a) & 1;
                                H[a] = b.reg[a] * b.n[a] + (d[a] < 0 ? -2 + o
                                                           : d[a] > 0 ? b.n[a] + o
                                                                      : o * (b.n[a] - 1));
                            }
                            if (!wrap(b.level, H)) { lv = kOut; break; }
                            int s2;
                            const int l2 = cover(b.level, H, s2);
                            if (c == 0) lv = l2;
                            else if (l2 != lv)
                                throw std::runtime_error("SMR: mesh is not 2:1 balanced");
                        }
                    }
    }
    bool cross_face(const Block &b, int a, int s) const {
        int d[3] = {0, 0, 0};
        d[a] = s ? 1 : -1;
        const int lv = b.lev_nb[dir_index(d)];
        return lv == 1 || lv == -1;
    }
    bool refine_face(const Block &b, int a, int s) const {   // geometry refine_flag
        int d[3] = {0, 0, 0};
        d[a] = s ? 1 : -1;
        return smr && b.lev_nb[dir_index(d)] == 1;
    }

    /* Cross-level ghost cells of w.  Restriction: volume-weighted mean of
     * the 2x2x2 finer cells (hydro/boundary/comm_functions.h:329-380).
     * Prolongation: the covering coarser cell plus minmod-limited slopes,
     * dq = minmod(q_r - q_c, q_c - q_l) / 4 per axis, signed by the fine
     * cell's half (comm_functions.h:382-452, mhd :470-571).  Both sum in
     * x, y, z order, so every block sees the same value at one position.
     * Same-level and restricted ghosts come first (act_bnd's q_ghst
     * stage), prolongations last (q_plgt, hydro/boundary/keeper.cpp
     * :166-179), so the coarse slopes read filled ghosts.               */
    void smr_ghost(int bi, int phase) {
        Block &b = blocks[bi];
        const int nv = nvar();
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { lo[a] = -b.gw[a]; hi[a] = b.n[a] + b.gw[a]; }
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    const int loc[3] = {i, j, k};
                    int nout = 0;
                    for (int a = 0; a < 3; ++a) nout += (loc[a] < 0 || loc[a] >= b.n[a]);
                    if (nout == 0) continue;
                    if (!cfg.mhd && nout > 1) continue;
                    int G[3];
                    for (int a = 0; a < 3; ++a) G[a] = b.reg[a] * b.n[a] + loc[a];
                    if (!wrap(b.level, G)) continue;
                    int src;
                    const int rel = cover(b.level, G, src);
                    const Block &s = blocks[src];
                    double *d = b.w.at(loc);
                    if (phase == 1 && rel == 0) {
                        int l[3];
                        for (int a = 0; a < 3; ++a) l[a] = G[a] - s.reg[a] * s.n[a];
                        const double *p = s.w.at(l);
                        for (int v = 0; v < nv; ++v) d[v] = p[v];
                    } else if (phase == 1 && rel == 1) {
                        int l[3];
                        for (int a = 0; a < 3; ++a) l[a] = 2 * G[a] - s.reg[a] * s.n[a];
                        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}, norm = 0;
                        const double dv = s.vol0;
                        for (int k0 = 0; k0 < 2; ++k0)
                            for (int d1 = 0; d1 < 2; ++d1)
                                for (int d2 = 0; d2 < 2; ++d2) {
                                    const double *p = s.w.at(l[0] + k0, l[1] + d1, l[2] + d2);
                                    for (int v = 0; v < nv; ++v) acc[v] += p[v] * dv;
                                    norm += dv;
                                }
                        for (int v = 0; v < nv; ++v) d[v] = acc[v] / norm;
                    } else if (phase == 2 && rel == -1) {
                        int l[3];
                        for (int a = 0; a < 3; ++a) l[a] = (G[a] >> 1) - s.reg[a] * s.n[a];
                        const double *qc = s.w.at(l);
                        double dq[3][8];
                        for (int a = 0; a < 3; ++a) {
                            int lr[3] = {l[0], l[1], l[2]}, ll[3] = {l[0], l[1], l[2]};
                            ++lr[a];
                            --ll[a];
                            for (const int *q : {lr, ll}) {
                                int dd[3], out = 0;
                                for (int x = 0; x < 3; ++x) {
                                    dd[x] = q[x] < 0 ? -1 : (q[x] >= s.n[x] ? 1 : 0);
                                    out |= dd[x];
                                }
                                if (out && s.lev_nb[dir_index(dd)] == -1)
                                    throw std::runtime_error("SMR: prolongation stencil "
                                                             "reaches a coarser ghost");
                            }
                            const double *qr = s.w.at(lr), *ql = s.w.at(ll);
                            for (int v = 0; v < nv; ++v)
                                dq[a][v] = minmod(qr[v] - qc[v], qc[v] - ql[v]) * 0.25f;
                        }
                        int sg[3];
                        for (int a = 0; a < 3; ++a) sg[a] = 2 * (G[a] & 1) - 1;
                        for (int v = 0; v < nv; ++v)
                            d[v] = qc[v] + dq[0][v] * sg[0] + dq[1][v] * sg[1] + dq[2][v] * sg[2];
                    }
                }
    }

    // Tangential face field of same-level face neighbours, one layer
    // (mhd/boundary/comm_functions.h:433-465); cross-level faces carry none.
    void smr_bf_ghost(int bi, int step) {
        Block &b = blocks[bi];
        for (int ax = 0; ax < 3; ++ax)
            for (int side = 0; side < 2; ++side) {
                int d[3] = {0, 0, 0};
                d[ax] = side ? 1 : -1;
                if (b.lev_nb[dir_index(d)] != 0) continue;
                int r[3];
                for (int a = 0; a < 3; ++a) r[a] = b.reg[a] + d[a];
                int G[3];
                for (int a = 0; a < 3; ++a) G[a] = r[a] * b.n[a];
                wrap(b.level, G);
                for (int a = 0; a < 3; ++a) r[a] = G[a] / b.n[a];
                const Block &s = blocks[leaf(b.level, r)];
                const int a1 = mod3(ax + 1), a2 = mod3(ax + 2);
                for (int t2 = 0; t2 <= b.n[a2]; ++t2)
                    for (int t1 = 0; t1 <= b.n[a1]; ++t1) {
                        int loc[3], sl[3];
                        loc[ax] = side ? b.n[ax] : -1;
                        sl[ax] = side ? 0 : b.n[ax] - 1;
                        loc[a1] = sl[a1] = t1;
                        loc[a2] = sl[a2] = t2;
                        const double *sp = s.bf.at(sl, step);
                        double *dp = b.bf.at(loc, step);
                        dp[a1] = sp[a1];
                        dp[a2] = sp[a2];
                    }
            }
    }

    /* Flux correction at faces whose neighbour is finer: the coarse face
     * flux becomes the area-weighted mean of the four finer face fluxes
     * (hydro/boundary/comm_functions.h:457-486 flux_to_buf, :489-505
     * flux_from_buf), summed along axis+1 then axis+2.                  */
    void smr_flux_restrict(Block &b) {
        const int nv = nvar();
        for (int a = 0; a < 3; ++a)
            for (int sd = 0; sd < 2; ++sd) {
                if (!refine_face(b, a, sd)) continue;
                const int a1 = mod3(a + 1), a2 = mod3(a + 2);
                for (int j2 = 0; j2 < b.n[a2]; ++j2)
                    for (int j1 = 0; j1 < b.n[a1]; ++j1) {
                        int G[3];               // coarse cell just outside the face
                        G[a] = b.reg[a] * b.n[a] + (sd ? b.n[a] : -1);
                        G[a1] = b.reg[a1] * b.n[a1] + j1;
                        G[a2] = b.reg[a2] * b.n[a2] + j2;
                        wrap(b.level, G);
                        int src;
                        if (cover(b.level, G, src) != 1)
                            throw std::runtime_error("SMR: refined face without finer leaf");
                        const Block &f = blocks[src];
                        int l[3];
                        for (int x = 0; x < 3; ++x) l[x] = 2 * G[x] - f.reg[x] * f.n[x];
                        l[a] = sd ? 0 : f.n[a];            // the finer block's face on the interface
                        double fb[8] = {0, 0, 0, 0, 0, 0, 0, 0}, norm = 0;
                        const double ds = f.sfc0[a];
                        for (int d1 = 0; d1 < 2; ++d1)
                            for (int d2 = 0; d2 < 2; ++d2) {
                                int p[3] = {l[0], l[1], l[2]};
                                p[a1] += d1;
                                p[a2] += d2;
                                const double *fm = f.f[a].at(p);
                                for (int v = 0; v < nv; ++v) fb[v] += fm[v] * ds;
                                norm += ds;
                            }
                        int q[3];
                        q[a] = sd ? b.n[a] : 0;
                        q[a1] = j1;
                        q[a2] = j2;
                        double *fc = b.f[a].at(q);
                        for (int v = 0; v < nv; ++v) fc[v] = fb[v] / norm;
                    }
            }
    }

    /* intg_ker (hydro/integrator/integrator.h:313-382): the corrected flux
     * of each refined face is added to the adjacent cell of the target
     * slot after the regular update skipped it (load_flux_sfc's null
     * faces, :156-182); z sides first, then y, then x.                  */
    void smr_intg_faces(Block &b, int step, double dt) {
        const int nv = nvar();
        for (int a = b.ndim - 1; a >= 0; --a)
            for (int sd = 0; sd <= 1; ++sd) {
                if (!refine_face(b, a, sd)) continue;
                const int a1 = mod3(a + 1), a2 = mod3(a + 2);
                for (int j2 = 0; j2 < b.n[a2]; ++j2)
                    for (int j1 = 0; j1 < b.n[a1]; ++j1) {
                        int i[3];
                        i[a] = sd * b.n[a];
                        i[a1] = j1;
                        i[a2] = j2;
                        const double *f = b.f[a].at(i);
                        const double ds = b.sfc0[a];
                        i[a] -= sd;
                        double *u = b.u.at(i, 1 - step);
                        const double dv = b.vol0 * (sd ? -1 : 1);
                        const double conv = dt * ds / dv;
                        for (int v = 0; v < nv; ++v) u[v] += conv * f[v];
                    }
            }
    }

    /* Edge EMFs where levels meet (mhd/boundary/comm_functions.h:795-864,
     * :919-954).  An edge of a coarser block that a finer leaf shares takes
     * the length-weighted mean of the two finer edges along it; afterwards
     * the finer blocks' edge pairs on such lines are overwritten with that
     * mean (the write-back), so both sides see one edge value.          */
    int edge_levels(const Block &b, int a, const int i[3], int &fine_o1, int &fine_o2) const {
        const int a1 = mod3(a + 1), a2 = mod3(a + 2);
        int has = 0;
        fine_o1 = fine_o2 = 0;
        for (int o2 = -1; o2 <= 0; ++o2)
            for (int o1 = -1; o1 <= 0; ++o1) {
                int c[3] = {i[0], i[1], i[2]};
                c[a1] += o1;
                c[a2] += o2;
                int d[3];
                for (int x = 0; x < 3; ++x) d[x] = c[x] < 0 ? -1 : (c[x] >= b.n[x] ? 1 : 0);
                const int lv = (d[0] || d[1] || d[2]) ? b.lev_nb[dir_index(d)] : 0;
                if (lv == 1 && !(has & 1)) {
                    has |= 1;
                    fine_o1 = o1;
                    fine_o2 = o2;
                }
                if (lv == -1) has |= 2;
            }
        return has;
    }
    void smr_emf_coarse(Block &b) {
        for (int i2 = 0; i2 <= b.n[2]; ++i2)
            for (int i1 = 0; i1 <= b.n[1]; ++i1)
                for (int i0 = 0; i0 <= b.n[0]; ++i0) {
                    const int i[3] = {i0, i1, i2};
                    bool bnd = false;
                    for (int x = 0; x < 3; ++x) bnd |= i[x] == 0 || i[x] == b.n[x];
                    if (!bnd) continue;
                    for (int a = 0; a < 3; ++a) {
                        if (i[a] >= b.n[a]) continue;
                        int o1, o2;
                        if (!(edge_levels(b, a, i, o1, o2) & 1)) continue;
                        const int a1 = mod3(a + 1), a2 = mod3(a + 2);
                        int G[3];
                        for (int x = 0; x < 3; ++x) G[x] = b.reg[x] * b.n[x] + i[x];
                        G[a1] += o1;
                        G[a2] += o2;
                        wrap(b.level, G);
                        int src;
                        cover(b.level, G, src);
                        const Block &f = blocks[src];
                        int e[3];
                        for (int x = 0; x < 3; ++x) e[x] = 2 * G[x] - f.reg[x] * f.n[x];
                        e[a1] += o1 < 0 ? 2 : 0;           // the finer edge on the line
                        e[a2] += o2 < 0 ? 2 : 0;
                        const double dl = f.dx0[a];
                        const double el = f.ee.at(e)[a] * dl;
                        ++e[a];
                        const double er = f.ee.at(e)[a] * dl;
                        b.ee.at(i)[a] = (el + er) / (dl + dl);
                    }
                }
    }
    void smr_emf_writeback(Block &b) {
        for (int i2 = 0; i2 <= b.n[2]; ++i2)
            for (int i1 = 0; i1 <= b.n[1]; ++i1)
                for (int i0 = 0; i0 <= b.n[0]; ++i0) {
                    const int i[3] = {i0, i1, i2};
                    bool bnd = false;
                    for (int x = 0; x < 3; ++x) bnd |= i[x] == 0 || i[x] == b.n[x];
                    if (!bnd) continue;
                    for (int a = 0; a < 3; ++a) {
                        if (i[a] >= b.n[a] || (i[a] & 1)) continue;
                        const int a1 = mod3(a + 1), a2 = mod3(a + 2);
                        if ((i[a1] & 1) || (i[a2] & 1)) continue;   // not on a coarser edge line
                        int o1, o2;
                        if (!(edge_levels(b, a, i, o1, o2) & 2)) continue;
                        int j[3] = {i[0], i[1], i[2]};
                        ++j[a];
                        const double dl = b.dx0[a];
                        double *pl = b.ee.at(i) + a, *pr = b.ee.at(j) + a;
                        const double avg = (*pl * dl + *pr * dl) / (dl + dl);
                        *pl = avg;
                        *pr = avg;
                    }
                }
    }

    /* -------- c2p + CFL (hydro/eos/adiabatic.h:211-259,
     *          mhd/eos/adiabatic.h:127-235) ------------------------- */
    double c2p_block(Block &b, int step) {
        double dtmin = (double)FLT_MAX;
        double v_c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        double v_p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < b.n[2]; ++k)
            for (int j = 0; j < b.n[1]; ++j)
                for (int i = 0; i < b.n[0]; ++i) {
                    double *p = b.w.at(i, j, k);
                    double *c = b.u.at(i, j, k, step);
                    if (!cfg.mhd) {
                        const double cs2 = eos.c2p(v_p, v_c, p, c);
                        if (step == 0) dtmin = dmin(dtmin, eos.dt_hydro(v_p, cs2, b));
                        continue;
                    }
                    int ii[3] = {i, j, k};
                    const double *bf = b.bf.at(ii, step);
                    double bc_2[3] = {0, 0, 0}, b_max[3] = {0, 0, 0}, bc[3];
                    double e_m(0);
                    for (int a = 0; a < 3; ++a) {
                        // mhd/eos/adiabatic.h:164-177: a thin axis has one
                        // face per cell, both sides of the cell
                        const bool thin = b.n[a] < 2;
                        const double b_l = bf[a];
                        ++ii[a];
                        const double b_r = thin ? b_l : b.bf.at(ii, step)[a];
                        --ii[a];
                        const double bc_m = (b_l + b_r) / 2;
                        bc[a] = bc_m;
                        const double bcs2 = bc_m * bc_m;
                        e_m += bcs2;
                        if (step != 0) continue;
                        bc_2[a] = bcs2;
                        b_max[a] = dmax(std::fabs(b_l), std::fabs(b_r));
                    }
                    e_m /= 2;
                    c[ENE] -= e_m;
                    eos.c2p(v_p, v_c, p, c);
                    c[ENE] += e_m;
                    for (int n = 0; n < 5; ++n) p[n] = v_p[n];
                    for (int a = 0; a < 3; ++a) {
                        p[BCC + a] = bc[a];
                        c[BCC + a] = bc[a];
                    }
                    const double *v = v_p + MOM;
                    double *ec = b.ec.at(ii);
                    ec[0] = bc[1] * v[2] - bc[2] * v[1];
                    ec[1] = bc[2] * v[0] - bc[0] * v[2];
                    ec[2] = bc[0] * v[1] - bc[1] * v[0];
                    if (step == 0) dtmin = dmin(dtmin, eos.dt_mhd(v_p, bc_2, b_max, b));
                }
        return dtmin;
    }

    /* -------- same-level ghost exchange (hydro/boundary/comm_functions.h
     * :304-327, mhd/boundary/comm_functions.h:433-465): ghosts receive the
     * neighbour's interior at the same global position (periodic wrap,
     * tree.cpp:77-84).  Hydro fills face ghosts only (keeper.cpp:81-84),
     * MHD all 26 neighbour directions plus one layer of tangential face-B
     * per face neighbour.                                                */
    bool global_src(const Block &b, const int loc[3], int &src, int lsrc[3]) const {
        int r[3];
        for (int a = 0; a < 3; ++a) {
            int gidx = b.reg[a] * b.n[a] + loc[a];
            const int ng = nblk[a] * b.n[a];
            if (gidx < 0 || gidx >= ng) {
                if (cfg.bnd[2 * a] != KRO_PER) return false;
                gidx = (gidx % ng + ng) % ng;
            }
            r[a] = gidx / b.n[a];
            lsrc[a] = gidx - r[a] * b.n[a];
        }
        src = bid(r);
        return true;
    }

    void exchange_block(int bi, int step) {
        Block &b = blocks[bi];
        const int nv = nvar();
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { lo[a] = -b.gw[a]; hi[a] = b.n[a] + b.gw[a]; }
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    const int loc[3] = {i, j, k};
                    int nout = 0;
                    for (int a = 0; a < 3; ++a) nout += (loc[a] < 0 || loc[a] >= b.n[a]);
                    if (nout == 0) continue;
                    if (!cfg.mhd && nout > 1) continue;
                    int src, ls[3];
                    if (!global_src(b, loc, src, ls)) continue;
                    const double *s = blocks[src].w.at(ls);
                    double *d = b.w.at(loc);
                    for (int n = 0; n < nv; ++n) d[n] = s[n];
                }
        if (!cfg.mhd) return;
        for (int ax = 0; ax < b.ndim; ++ax)       // a thin axis has no neighbour
            for (int side = 0; side < 2; ++side) {
                const int a1 = mod3(ax + 1), a2 = mod3(ax + 2);
                for (int t2 = 0; t2 < b.bf.N[a2] - 4; ++t2)
                    for (int t1 = 0; t1 < b.bf.N[a1] - 4; ++t1) {
                        int loc[3];
                        loc[ax] = side ? b.n[ax] : -1;
                        loc[a1] = t1;
                        loc[a2] = t2;
                        // the source block: the face neighbour along ax
                        int probe[3] = {0, 0, 0};
                        probe[ax] = loc[ax];
                        int src, ls[3];
                        if (!global_src(b, probe, src, ls)) continue;
                        int sl[3];
                        sl[ax] = ls[ax];
                        sl[a1] = t1;
                        sl[a2] = t2;
                        const double *s = blocks[src].bf.at(sl, step);
                        double *d = b.bf.at(loc, step);
                        d[a1] = s[a1];
                        d[a2] = s[a2];
                    }
            }
    }

    /* -------- physical boundaries on w (hydro/boundary/phys.h:184-229,
     *          phys_functors.h:10-89) --------------------------------- */
    void phys_block(Block &b) {
        const int nv = nvar();
        for (int side = 0; side < 6; ++side) {
            const int axis = side >> 1;
            const bool right = side & 1;
            if (!b.phy[axis][right]) continue;
            const int kind = cfg.bnd[side];
            const int ap1 = mod3(axis + 1), ap2 = mod3(axis + 2);
            const int n_gh = 2;
            if (b.n[axis] < 2) continue;
            for (int t2 = 0; t2 < b.n[ap2]; ++t2)
                for (int t1 = 0; t1 < b.n[ap1]; ++t1) {
                    double buf[4][8];           // positions -2..1 -> rows 0..3
                    auto B = [&](int pos) { return buf[pos + n_gh]; };
                    int i3[3];
                    i3[ap1] = t1;
                    i3[ap2] = t2;
                    for (int i = 0; i < n_gh; ++i) {
                        i3[axis] = right ? b.n[axis] - i - 1 : i;
                        const double *p = b.w.at(i3);
                        for (int n = 0; n < nv; ++n) B(i)[n] = p[n];
                    }
                    if (kind == KRO_FRE) {
                        for (int i = -n_gh; i < 0; ++i)
                            for (int n = 0; n < nv; ++n) B(i)[n] = B(0)[n];
                    } else if (kind == KRO_OUT) {
                        const int i_mom = side / 2 + MOM;
                        if (B(0)[i_mom] * (right ? -1 : 1) > 0) B(0)[i_mom] = 0;
                        for (int i = -n_gh; i < 0; ++i)
                            for (int n = 0; n < nv; ++n) B(i)[n] = B(0)[n];
                    } else if (kind == KRO_REF) {
                        const int i_mom = MOM + side / 2;
                        for (int i = -n_gh; i < 0; ++i)
                            for (int n = 0; n < nv; ++n)
                                B(i)[n] = B(-i - 1)[n] * (n == i_mom ? -1 : 1);
                    } else if (kind == KRO_FIX) {       // phys_functors.h:25-48
                        for (int i = -n_gh; i < 0; ++i)
                            for (int n = 0; n < nv; ++n) B(i)[n] = cfg.fix_w[n];
                    }
                    for (int i = -n_gh; i < 0; ++i) {
                        i3[axis] = right ? b.n[axis] - i - 1 : i;
                        double *p = b.w.at(i3);
                        for (int n = 0; n < nv; ++n) p[n] = B(i)[n];
                    }
                }
        }
    }

    /* -------- MHD physical boundaries (mhd/boundary/phys.h:144-251,
     * phys_functors.h:10-137).  One kernel per physical side, in the order
     * x-, x+, y-, y+, z-, z+ (keeper.cpp:91-117); each covers its two ghost
     * layers over the transverse range including the transverse ghost
     * bands, reading sources clamped into the block along transverse axes
     * that are physical on that end (f_i) and taken from the exchanged
     * ghosts otherwise.  The primitives (with B_cc) and the face field get
     * the functor's values: free / outflow copy the boundary cell (outflow
     * zeroes an inflowing normal velocity first); reflect mirrors the cells
     * with the normal velocity negated, B_cc's tangential components
     * negated, and the face field mirrored at -i (tangential negated).
     * The reference runs this once with the face field before the ghost
     * exchange and once more (primitives only) after it; the face-field
     * entries whose sources are exchanged ghosts are never read, so one
     * pass after the exchange gives every value that is used.            */
    void phys_block_mhd(Block &b, int step) {
        const int nv = nvar();
        const int n_gh = 2;
        for (int side = 0; side < 6; ++side) {
            const int axis = side >> 1;
            const bool right = side & 1;
            // mesh/boundary/keeper.cpp:100-103: physical sides of the
            // active axes only
            if (axis >= b.ndim || !b.phy[axis][right]) continue;
            const int kind = cfg.bnd[side];
            const int ap1 = mod3(axis + 1), ap2 = mod3(axis + 2);
            const int n_ceff = b.n[axis];
            // a thin transverse axis has no ghost band (mhd/boundary/phys.h
            // :80-90, :154-157)
            const int g1 = b.n[ap1] > 1 ? n_gh : 0, g2 = b.n[ap2] > 1 ? n_gh : 0;
            auto f_i = [&](int a, int v, int i_m) {
                if (b.phy[a][0]) v = std::max(v, 0);
                if (b.phy[a][1]) v = std::min(v, i_m);
                return v;
            };
            for (int t2 = -g2; t2 < b.n[ap2] + g2; ++t2)
                for (int t1 = -g1; t1 < b.n[ap1] + g1; ++t1) {
                    int i_tgt[3], i_src[3], i_sfc[3];
                    i_tgt[ap1] = t1;
                    i_tgt[ap2] = t2;
                    i_src[ap1] = f_i(ap1, t1, b.n[ap1] - 1);
                    i_src[ap2] = f_i(ap2, t2, b.n[ap2] - 1);
                    i_sfc[ap1] = f_i(ap1, t1, b.n[ap1]);
                    i_sfc[ap2] = f_i(ap2, t2, b.n[ap2]);
                    double hyd[5][8], bc[5][3], bfb[5][3];   // buffer positions -2..2
                    auto H = [&](int i) { return hyd[i + n_gh]; };
                    auto C = [&](int i) { return bc[i + n_gh]; };
                    auto F = [&](int i) { return bfb[i + n_gh]; };
                    for (int i = 0; i <= n_gh; ++i) {
                        i_src[axis] = right ? n_ceff - i - 1 : i;
                        i_sfc[axis] = right ? n_ceff - i : i;
                        const double *hs = b.w.at(i_src);
                        const double *fs = b.bf.at(i_sfc, step);
                        const double *fc = b.bf.at(i_src, step);
                        for (int v = 0; v < nv; ++v) H(i)[v] = hs[v];
                        for (int a = 0; a < 3; ++a) {
                            C(i)[a] = hs[BCC + a];
                            F(i)[a] = a == axis ? fs[a] : fc[a];
                        }
                    }
                    const int i_mom = MOM + axis;
                    if (kind == KRO_OUT || kind == KRO_FRE) {
                        if (kind == KRO_OUT && H(0)[i_mom] * (right ? -1 : 1) > 0) H(0)[i_mom] = 0;
                        for (int i = -n_gh; i < 0; ++i) {
                            for (int v = 0; v < nv; ++v) H(i)[v] = H(0)[v];
                            for (int a = 0; a < 3; ++a) {
                                C(i)[a] = C(0)[a];
                                F(i)[a] = F(0)[a];
                            }
                        }
                    } else if (kind == KRO_FIX) {       // phys_functors.h:38-70 (MHD)
                        for (int i = -n_gh; i < 0; ++i) {
                            for (int v = 0; v < 5; ++v) H(i)[v] = cfg.fix_w[v];
                            for (int a = 0; a < 3; ++a) {
                                H(i)[BCC + a] = cfg.fix_b[a];
                                C(i)[a] = cfg.fix_b[a];
                                F(i)[a] = cfg.fix_b[a];
                            }
                        }
                    } else if (kind == KRO_REF) {
                        for (int i = -n_gh; i < 0; ++i) {
                            for (int v = 0; v < nv; ++v)
                                H(i)[v] = H(-i - 1)[v] * (v == i_mom ? -1 : 1);
                            for (int a = 0; a < 3; ++a) {
                                if (a != axis) {
                                    F(i)[a] = -F(-i)[a];
                                    C(i)[a] = -C(-i - 1)[a];
                                } else {
                                    F(i)[a] = F(-i)[a];
                                    C(i)[a] = C(-i - 1)[a];
                                }
                            }
                        }
                    }
                    for (int i = -n_gh; i < 0; ++i) {
                        i_tgt[axis] = right ? n_ceff - i - 1 : i;
                        int i_fbc[3] = {i_tgt[0], i_tgt[1], i_tgt[2]};
                        i_fbc[axis] = right ? n_ceff - i : i;
                        double *ht = b.w.at(i_tgt);
                        double *ff = b.bf.at(i_fbc, step);
                        double *fc = b.bf.at(i_tgt, step);
                        for (int v = 0; v < nv; ++v) ht[v] = H(i)[v];
                        for (int a = 0; a < 3; ++a) {
                            ht[BCC + a] = C(i)[a];
                            if (a == axis) ff[a] = F(i)[a];
                            else fc[a] = F(i)[a];
                        }
                    }
                }
        }
    }

    /* -------- one pencil sweep along `axis` (prototypes/integrator.h
     * :14-32; hydro/integrator/{data,integrator}.h; mhd/integrator/
     * {data,integrator}.h).  Lines cover the interior and, for MHD, the
     * four transverse ghost lines bands (mhd/integrator/data.h:53-91). */
    void sweep_block(Block &b, int axis, int step, double dt) {
        const int nv = nvar();
        const bool mhd = cfg.mhd;
        const int a1 = mod3(axis + 1), a2 = mod3(axis + 2);
        const int n = b.n[axis], g = b.gw[axis];
        std::vector<double> q((size_t)(n + 2 * g) * nv);
        std::vector<double> flx((size_t)(n + 1) * nv);
        auto Q = [&](int i) { return &q[(size_t)(i + g) * nv]; };
        auto do_line = [&](int t1, int t2, bool interior) {
            int i3[3];
            i3[a1] = t1;
            i3[a2] = t2;
            for (int i = -g; i < n + g; ++i) {
                i3[axis] = i;
                const double *src = b.w.at(i3);
                for (int v = 0; v < nv; ++v)
                    if (!(mhd && v == BCC + axis)) Q(i)[v] = src[v];
            }
            // mhd/integrator/data.h:108-140: on a transverse ghost line
            // next to a cross-level face the normal B is the cell value of w
            bool cross = false;
            if (mhd && smr && !interior) {
                if (t1 < 0 || t1 >= b.n[a1]) cross = cross_face(b, a1, t1 >= 0);
                else cross = cross_face(b, a2, t2 >= 0);
            }
            if (mhd)
                for (int i = 0; i <= n; ++i) {
                    i3[axis] = i;
                    Q(i)[BCC + axis] = cross ? b.w.at(i3)[BCC + axis] : b.bf.at(i3, step)[axis];
                }
            double ql[8], qr[8];
            for (int i = 0; i <= n; ++i) {
                for (int v = 0; v < nv; ++v) {
                    if (mhd && v == BCC + axis) {
                        ql[v] = Q(i)[v];
                        qr[v] = ql[v];
                        continue;
                    }
                    plm_uniform(Q(i - 2)[v], Q(i - 1)[v], Q(i)[v], Q(i + 1)[v],
                                ql[v], qr[v]);
                }
                face_flux(mhd, axis, ql, qr, &flx[(size_t)i * nv], eos);
            }
            if (mhd || axis != 0 || smr)                // save_flux / slice_flux
                for (int i = 0; i <= n; ++i) {
                    i3[axis] = i;
                    double *dst = b.f[axis].at(i3);
                    for (int v = 0; v < nv; ++v) dst[v] = flx[(size_t)i * nv + v];
                }
            if (axis != 0 || !interior) return;
            // add_flux (hydro/integrator/integrator.h:241-265) with
            // rk2_t::load (hydro/integrator/rk2.h:20-35)
            for (int i = 0; i < n; ++i) {
                i3[0] = i;
                double uu[8];
                const double *u_src = b.u.at(i3, step);
                double *u_tgt = b.u.at(i3, 1 - step);
                if (step == 0)
                    for (int v = 0; v < nv; ++v) uu[v] = u_src[v];
                else
                    for (int v = 0; v < nv; ++v) uu[v] = (u_src[v] + u_tgt[v]) * 0.5f;
                const double dt_dv = dt / b.vol0;
                for (int ax = 0; ax < b.ndim; ++ax) {
                    // refined faces are skipped here and added by
                    // smr_intg_faces (integrator.h:156-182 null_l / null_r)
                    const bool null_l = refine_face(b, ax, 0) && i3[ax] == 0;
                    const bool null_r = refine_face(b, ax, 1) && i3[ax] + 1 == b.n[ax];
                    const double fa0 = null_l ? -1 : b.sfc0[ax] * dt_dv;
                    const double fa1 = null_r ? -1 : b.sfc0[ax] * dt_dv;
                    const double *F0, *F1;
                    if (ax == 0) {
                        F0 = &flx[(size_t)i * nv];
                        F1 = &flx[(size_t)(i + 1) * nv];
                    } else {
                        F0 = b.f[ax].at(i3);
                        ++i3[ax];
                        F1 = b.f[ax].at(i3);
                        --i3[ax];
                    }
                    for (int v = 0; v < nv; ++v) {
                        double res(0);
                        if (fa0 > 0) res += fa0 * F0[v];
                        if (fa1 > 0) res -= fa1 * F1[v];
                        uu[v] += res;
                    }
                }
                for (int v = 0; v < nv; ++v) u_tgt[v] = uu[v];
            }
        };
        const int n1 = b.n[a1], n2 = b.n[a2];
        for (int t2 = 0; t2 < n2; ++t2)
            for (int t1 = 0; t1 < n1; ++t1) do_line(t1, t2, true);
        if (!mhd) return;
        if (n1 > 1)
            for (int s = 0; s < 2; ++s)
                for (int t2 = 0; t2 < n2; ++t2) do_line(s ? n1 : -1, t2, false);
        if (n2 > 1)
            for (int s = 0; s < 2; ++s)
                for (int t1 = 0; t1 < n1; ++t1) do_line(t1, s ? n2 : -1, false);
    }

    /* -------- edge EMF (mhd/integrator/integrator.h:189-327) ---------- */
    void calc_field_block(Block &b) {
        auto f_ec = [&](const int j[3], const int ax[3]) -> double {
            bool is_in = true;
            for (int a = 0; a < 3; ++a) is_in &= j[a] >= 0 && j[a] < b.n[a];
            if (is_in) return b.ec.at(j)[ax[0]];
            const double *bb = b.w.at(j) + BCC;
            const double *v = b.w.at(j) + MOM;
            return bb[ax[1]] * v[ax[2]] - bb[ax[2]] * v[ax[1]];
        };
        const int n_dm = b.ndim;
        for (int i2 = 0; i2 <= b.n[2]; ++i2)
            for (int i1 = 0; i1 <= b.n[1]; ++i1)
                for (int i0 = 0; i0 <= b.n[0]; ++i0) {
                    int i[3] = {i0, i1, i2};
                    // thin axes carry one edge layer (integrator.h:221-230)
                    bool th_skip = false;
                    for (int a = n_dm; a < 3; ++a) th_skip |= i[a] > 0;
                    if (th_skip) continue;
                    double ee[3] = {0, 0, 0};
                    double ec[2][2];
                    int axis[3];
                    int &ax = axis[0], &ax_1 = axis[1], &ax_2 = axis[2];
                    for (ax = 0; ax < 3; ++ax) {
                        if (i[ax] >= b.n[ax]) continue;
                        ax_1 = mod3(1 + ax);
                        ax_2 = mod3(2 + ax);
                        // no shift along a thin axis (:236-252)
                        const bool th_1 = ax_1 >= n_dm, th_2 = ax_2 >= n_dm;
                        ec[1][1] = f_ec(i, axis);
                        if (!th_1) --i[ax_1];
                        ec[1][0] = f_ec(i, axis);
                        if (!th_2) --i[ax_2];
                        ec[0][0] = f_ec(i, axis);
                        if (!th_1) ++i[ax_1];
                        ec[0][1] = f_ec(i, axis);
                        if (!th_2) ++i[ax_2];
                        ee[ax] = 0;
                        // the surviving sweep takes the full weight when a
                        // thin axis is one of the two (:258-266)
                        const double w_e = (th_1 || th_2) ? 0.5f : 0.25f;
                        const int paras[2] = {ax_1, ax_2};
                        for (int pi = 0; pi < 2; ++pi) {
                            const int a_para = paras[pi];
                            const int a_perp = (a_para == ax_1 ? ax_2 : ax_1);
                            const bool thn_p = a_perp >= n_dm;
                            // a thin sweep direction has no flux (:272-276)
                            if (a_para < n_dm) {
                                if (!thn_p) --i[a_perp];
                                const double ef_m = b.f[a_para].at(i)[BCC + ax];
                                if (!thn_p) ++i[a_perp];
                                const double ef_p = b.f[a_para].at(i)[BCC + ax];
                                ee[ax] += w_e * (ef_m + ef_p);
                                if (!thn_p) {       // corner upwinding
                                    const double frho_p = b.f[a_perp].at(i)[RHO];
                                    --i[a_para];
                                    const double frho_m = b.f[a_perp].at(i)[RHO];
                                    ++i[a_para];
                                    const double de_p = (frho_p > 0 ? (ec[0][1] - ef_m)
                                                                    : (ec[1][1] - ef_p));
                                    const double de_m = (frho_m > 0 ? -(ec[0][0] - ef_m)
                                                                    : -(ec[1][0] - ef_p));
                                    ee[ax] += w_e * (de_m - de_p);
                                }
                            }
                            if (a_para != ax_1) continue;
                            const double swp = ec[0][1];
                            ec[0][1] = ec[1][0];
                            ec[1][0] = swp;
                        }
                    }
                    double *ee_g = b.ee.at(i);
                    for (int a = 0; a < 3; ++a)
                        if (i[a] < b.n[a]) ee_g[a] = ee[a];
                    // boundary-plane advective edge EMF on a high out / fre
                    // side (integrator.h:328-381): the interior cell's
                    // -(v x B) with v = w[mom] / rho as written there
                    for (int a = 0; a < 3; ++a) {
                        if (i[a] != b.n[a] || b.n[a] < 2) continue;
                        if (!b.phy[a][1]) continue;
                        const int k = cfg.bnd[2 * a + 1];
                        if (k != KRO_OUT && k != KRO_FRE) continue;
                        int jc[3] = {i[0], i[1], i[2]};
                        --jc[a];
                        const double *wc = b.w.at(jc);
                        const double irho = 1.0 / std::fmax(wc[RHO], 1e-30);
                        double v_c[3], b_c[3];
                        for (int q = 0; q < 3; ++q) {
                            v_c[q] = wc[MOM + q] * irho;
                            b_c[q] = wc[BCC + q];
                        }
                        const int a1 = mod3(a + 1), a2 = mod3(a + 2);
                        ee_g[a1] = v_c[a] * b_c[a2] - v_c[a2] * b_c[a];
                        ee_g[a2] = v_c[a1] * b_c[a] - v_c[a] * b_c[a1];
                    }
                }
    }

    /* -------- CT update (mhd/integrator/integrator.h:385-462,
     *          mhd/integrator/rk2.h:191-205) --------------------------- */
    void intg_field_block(Block &b, int step, double dt) {
        const int n_dm = b.ndim;
        for (int i2 = 0; i2 <= b.n[2]; ++i2)
            for (int i1 = 0; i1 <= b.n[1]; ++i1)
                for (int i0 = 0; i0 <= b.n[0]; ++i0) {
                    int i[3] = {i0, i1, i2};
                    // the wrap slot of a thin axis (index 1) is a copy of
                    // slot 0 made by the thin-axis sync below; the reference
                    // computes it from an edge layer it never fills and
                    // then overwrites it (integrator.h:598-632)
                    bool th_skip = false;
                    for (int a = n_dm; a < 3; ++a) th_skip |= i[a] > 0;
                    if (th_skip) continue;
                    double db_v[3] = {0, 0, 0};
                    unsigned msk = 0;
                    for (int a = 0; a < 3; ++a) {
                        const int ax_1 = mod3(a + 1), ax_2 = mod3(a + 2);
                        if (i[ax_1] >= b.n[ax_1] || i[ax_2] >= b.n[ax_2]) continue;
                        msk |= 1u << a;
                        // pairs spanning a thin axis vanish (integrator.h:419-447)
                        const bool th_1 = ax_1 >= n_dm, th_2 = ax_2 >= n_dm;
                        double &db = db_v[a];
                        if (!th_2) db = (double)(b.ee.at(i)[ax_1] * b.dx0[ax_1]);
                        else db = (double)(0);
                        if (!th_1) db -= (double)(b.ee.at(i)[ax_2] * b.dx0[ax_2]);
                        ++i[ax_1];
                        if (!th_1) db += (double)(b.ee.at(i)[ax_2] * b.dx0[ax_2]);
                        --i[ax_1];
                        ++i[ax_2];
                        if (!th_2) db -= (double)(b.ee.at(i)[ax_1] * b.dx0[ax_1]);
                        --i[ax_2];
                        // the thin face family's area is dx_f(a1) * dx_f(a2)
                        // (sfc_thin, integrator.h:74-81)
                        const double sfc = a >= n_dm ? b.dx0[ax_1] * b.dx0[ax_2] : b.sfc0[a];
                        db *= (sfc > 0 ? (double)(-dt / sfc) : (double)(0));
                    }
                    double *b_src = b.bf.at(i, step);
                    double *b_tgt = b.bf.at(i, 1 - step);
                    for (int a = 0; a < 3; ++a)
                        if (msk & (1u << a))
                            b_tgt[a] = db_v[a] + (step == 0 ? b_src[a]
                                                            : (b_src[a] + b_tgt[a]) / 2);
                }
        thin_axis_sync(b);
    }

    /* -------- thin-axis face sync (mhd/integrator/integrator.h:520-560,
     * launched after every CT update at :585-600): along a periodic
     * one-cell axis the wrap slots -2..2 of slot 0 are the slot-0 face,
     * over the other two axes' interior and ghost band.                  */
    void thin_axis_sync(Block &b) {
        for (int a = 0; a < 3; ++a) {
            if (b.n[a] != 1 || b.phy[a][0] || b.phy[a][1]) continue;
            const int a1 = mod3(a + 1), a2 = mod3(a + 2);
            const int g = 2;
            for (int t2 = -g; t2 < b.n[a2] + g; ++t2)
                for (int t1 = -g; t1 < b.n[a1] + g; ++t1) {
                    int i[3];
                    i[a] = 0;
                    i[a1] = t1;
                    i[a2] = t2;
                    const double *d0 = b.bf.at(i);
                    for (int k = -g; k <= g; ++k) {
                        if (k == 0) continue;
                        int j[3] = {i[0], i[1], i[2]};
                        j[a] = k;
                        double *d1 = b.bf.at(j);
                        for (int c = 0; c < 3; ++c) d1[c] = d0[c];
                    }
                }
        }
    }

    /* -------- one RK substep (dynamics/dynamics.h:60-100) ------------ */
    void integrate(int step) {
        const int nb = (int)blocks.size();
        std::vector<double> dtb(nb, (double)FLT_MAX);
        par_blocks([&](int bi) { dtb[bi] = c2p_block(blocks[bi], step); });
        if (step == 0)
            for (int bi = 0; bi < nb; ++bi) p_dt = std::fmin(dtb[bi], p_dt);
        if (!dt_reduced) {                       // cycle.cpp:77-115
            const double dt_h = p_dt;
            const double dt_use = cfg.top_up ? tmin(dt_h, cfg.t_lim - t) : dt_h;
            dt_cur = dt_use;
            if (cfg.top_up) t = tmin(dt_use + t, cfg.t_lim);
            else            t = dt_use + t;
            p_dt = dt_use * cfg.dt_expand;
            dt_reduced = true;
        }
        if (!smr) {
            par_blocks([&](int bi) {
                exchange_block(bi, step);
            });
            if (!cfg.mhd) par_blocks([&](int bi) { phys_block(blocks[bi]); });
            else par_blocks([&](int bi) { phys_block_mhd(blocks[bi], step); });
        } else {
            par_blocks([&](int bi) {
                smr_ghost(bi, 1);
                if (cfg.mhd) smr_bf_ghost(bi, step);
            });
            if (!cfg.mhd) par_blocks([&](int bi) { phys_block(blocks[bi]); });
            else par_blocks([&](int bi) { phys_block_mhd(blocks[bi], step); });
            par_blocks([&](int bi) { smr_ghost(bi, 2); });
            // MHD: the physical ghosts once more after the prolongation
            // (mhd/boundary/keeper.h:288-293 runs the side kernels again
            // after act_bnd), so that a ghost corner whose transverse source
            // is a prolonged halo reads this substep's value; every other
            // value is the same as in the first pass
            if (cfg.mhd) par_blocks([&](int bi) { phys_block_mhd(blocks[bi], step); });
        }
        const double dt = dt_cur * (step == 0 ? 1.0 : 0.5);
        par_blocks([&](int bi) {
            Block &b = blocks[bi];
            for (int ax = b.ndim - 1; ax >= 0; --ax) sweep_block(b, ax, step, dt);
            if (cfg.mhd) calc_field_block(b);
        });
        if (smr) {                               // act_flx, then intg_flx
            par_blocks([&](int bi) { smr_flux_restrict(blocks[bi]); });
            if (cfg.mhd) {
                par_blocks([&](int bi) { smr_emf_coarse(blocks[bi]); });
                par_blocks([&](int bi) { smr_emf_writeback(blocks[bi]); });
            }
            par_blocks([&](int bi) { smr_intg_faces(blocks[bi], step, dt); });
        }
        if (cfg.mhd)
            par_blocks([&](int bi) { intg_field_block(blocks[bi], step, dt); });
    }

    bool cycle() {                               // cycle.cpp:124-150
        if (t >= cfg.t_lim) return false;
        integrate(0);
        integrate(1);
        dt_reduced = false;
        return true;
    }
};

}  // namespace

/* ==================================================================== */
/* C interface                                                          */
struct kro_mesh { Mesh m; };

extern "C" {

const char *kro_last_error(void) { return g_err.c_str(); }

static kro_mesh *kro_create_impl(const kro_config *c, bool do_build) {
    try {
        auto *km = new kro_mesh();
        Mesh &m = km->m;
        m.cfg.mhd = c->mhd;
        m.cfg.gamma = c->gamma; m.cfg.cfl = c->cfl;
        m.cfg.d_floor = c->rho_floor; m.cfg.t_floor = c->tmp_floor;
        for (int a = 0; a < 3; ++a) {
            m.cfg.ng[a] = c->n_cell_global[a];
            m.cfg.nb[a] = c->n_cell_block[a];
            m.cfg.xmin[a] = c->x_min[a];
            m.cfg.xmax[a] = c->x_max[a];
        }
        for (int s = 0; s < 6; ++s) m.cfg.bnd[s] = c->bnd[s];
        for (int a = 0; a < 3; ++a)
            if ((c->bnd[2 * a] == KRO_PER) != (c->bnd[2 * a + 1] == KRO_PER)) {
                delete km;
                g_err = "periodic must be set on both sides";
                return nullptr;
            }
        m.cfg.dt_init = c->dt_init; m.cfg.dt_expand = c->dt_expand;
        m.cfg.t_lim = c->t_lim; m.cfg.top_up = c->top_up != 0;
        m.cfg.nthreads = c->n_threads;
        for (int k = 0; k < 5; ++k) m.cfg.fix_w[k] = c->fix_w[k];
        for (int k = 0; k < 3; ++k) m.cfg.fix_b[k] = c->fix_b[k];
        m.eos.gamma = c->gamma; m.eos.cfl = c->cfl;
        m.eos.d_floor = c->rho_floor; m.eos.t_floor = c->tmp_floor;
        m.eos.nhyd = c->mhd ? 8 : 5;
        if (do_build) m.build();
        return km;
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

kro_mesh *kro_create_blocks(const kro_config *c, int n, const int *levels, const int *regs) {
    try {
        if (n <= 0 || !levels || !regs) throw std::runtime_error("no blocks");
        kro_config c0 = *c;
        auto *km = kro_create_impl(&c0, false);
        if (!km) return nullptr;
        std::vector<int> lv(levels, levels + n), rg(regs, regs + 3 * (size_t)n);
        try {
            km->m.build(&lv, &rg);
        } catch (...) {
            delete km;
            throw;
        }
        return km;
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

kro_mesh *kro_create(const kro_config *c) { return kro_create_impl(c, true); }

void kro_destroy(kro_mesh *m) { delete m; }

int kro_n_blocks(const kro_mesh *m) { return (int)m->m.blocks.size(); }

void kro_block_geom(const kro_mesh *km, int b, int reg[3], int n[3],
                    double xf0[3], double dx0[3]) {
    const Block &bl = km->m.blocks[b];
    for (int a = 0; a < 3; ++a) {
        reg[a] = bl.reg[a]; n[a] = bl.n[a];
        xf0[a] = bl.xf0[a]; dx0[a] = bl.dx0[a];
    }
}

int kro_set_block(kro_mesh *km, int b, const double *u_aos, const double *bf_aos) {
    Mesh &m = km->m;
    if (b < 0 || b >= (int)m.blocks.size()) return -1;
    Block &bl = m.blocks[b];
    const int nv = m.nvar();
    size_t c = 0;
    for (int k = 0; k < bl.n[2]; ++k)
        for (int j = 0; j < bl.n[1]; ++j)
            for (int i = 0; i < bl.n[0]; ++i, ++c) {
                double *d = bl.u.at(i, j, k, 0);
                for (int v = 0; v < nv; ++v) d[v] = u_aos[c * nv + v];
            }
    if (m.cfg.mhd && bf_aos)
        std::memcpy(bl.bf.d.data(), bf_aos, bl.bf.nctot * sizeof(double));
    return 0;
}

int kro_get_block(const kro_mesh *km, int b, double *u_aos, double *bf_aos) {
    const Mesh &m = km->m;
    if (b < 0 || b >= (int)m.blocks.size()) return -1;
    const Block &bl = m.blocks[b];
    const int nv = m.nvar();
    size_t c = 0;
    for (int k = 0; k < bl.n[2]; ++k)
        for (int j = 0; j < bl.n[1]; ++j)
            for (int i = 0; i < bl.n[0]; ++i, ++c) {
                const double *d = bl.u.at(i, j, k, 0);
                for (int v = 0; v < nv; ++v) u_aos[c * nv + v] = d[v];
            }
    if (m.cfg.mhd && bf_aos)
        std::memcpy(bf_aos, bl.bf.d.data(), bl.bf.nctot * sizeof(double));
    return 0;
}

int kro_advance(kro_mesh *km, int n_cycles) {
    try {
        int done = 0;
        while (done < n_cycles && km->m.cycle()) ++done;
        return done;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

double kro_time(const kro_mesh *km) { return km->m.t; }
double kro_dt(const kro_mesh *km) { return km->m.dt_cur; }
double kro_dt_next(const kro_mesh *km) { return km->m.p_dt; }

void kro_plm(const double q[4], double *ql, double *qr) {
    plm_uniform(q[0], q[1], q[2], q[3], *ql, *qr);
}

void kro_hllc(const double wl[5], const double wr[5], double gamma, double flx[5]) {
    Eos e{gamma, 0.3, 1e-9, 1e-9, 5};
    hllc_eval(flx, wl, wr, e);
}

void kro_hlld(const double wl[8], const double wr[8], double gamma, double flx[8]) {
    Eos e{gamma, 0.3, 1e-9, 1e-9, 8};
    hlld_eval(flx, wl, wr, e);
}

void kro_face_flux(int mhd, int axis, const double *wl, const double *wr,
                   double gamma, double *flx) {
    Eos e{gamma, 0.3, 1e-9, 1e-9, mhd ? 8 : 5};
    double ql[8], qr[8];
    const int nv = mhd ? 8 : 5;
    for (int v = 0; v < nv; ++v) { ql[v] = wl[v]; qr[v] = wr[v]; }
    face_flux(mhd != 0, axis, ql, qr, flx, e);
}

double kro_c2p_hydro(double *u, double *w, double gamma, double rho_floor,
                     double tmp_floor) {
    Eos e{gamma, 0.3, rho_floor, tmp_floor, 5};
    double prim[5], cons[5];
    return e.c2p(prim, cons, w, u);
}

}  // extern "C"
