// kr_physics.h -- device-side physics of the Godunov hydro/MHD stage.
//
// Every function restates one reference functor with the same expression
// order, so that the FP64 instantiation (compiled with -ffp-contract=off)
// rounds exactly like the reference's PRECISION=2 build.  The flux-side
// functions are templated on the flux type FT: double for the plain form,
// float for the mixed form (the reference's PRECISION=1 semantics: FP32
// reconstruction and Riemann solve, FP64 conserved state).
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <type_traits>

namespace kr {

enum { RHO = 0, ENE = 1, MOM = 2, BCC = 5 };   // dynamics/indices.h:7-13

struct EosPar {
    double gamma, cfl, d_floor, t_floor;
};

__device__ __forceinline__ int mod3(int i) { return i < 3 ? i : i - 3; }

template <class T> __device__ __forceinline__ T vmax(T a, T b) { return a > b ? a : b; }
template <class T> __device__ __forceinline__ T vmin(T a, T b) { return a < b ? a : b; }
template <class T> __device__ __forceinline__ T vabs(T a) { return a > 0 ? a : -a; }
template <class T> __device__ __forceinline__ int vsgn(T a) { return a > 0 ? 1 : (a < 0 ? -1 : 0); }

__device__ __forceinline__ double ksqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float ksqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double kfabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float kfabs(float x) { return __builtin_fabsf(x); }

// Reference PRECISION-dependent max/min helpers (utilities/functions.h:
// 51-64 for FP64, fmaxf/fminf for FP32 at :84-85).
__device__ __forceinline__ double pmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double pmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ float pmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float pmin(float a, float b) { return __builtin_fminf(a, b); }

// utils::idx_swap (utilities/mapping/mapping.h:225-237)
template <class T> __device__ __forceinline__ void idx_swap(T *x, int axis) {
    if (mod3(axis) != 0) {
        T s[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) s[i] = x[mod3(i + axis)];
#pragma unroll
        for (int i = 0; i < 3; ++i) x[i] = s[i];
    }
}

// ---------------------------------------------------------------------
// Conserved -> primitive with floors, FP64.  hydro/eos/adiabatic.h:159-206
// with pre_cs2 (:77-85).  cons = {rho, E, m0, m1, m2} (E without the
// magnetic part for MHD).  On return cons holds the healed values, prim the
// primitives {rho, p, v0, v1, v2}; the return value is c_s^2.  heal_rho /
// heal_e report which conserved slots were rewritten.
__device__ __forceinline__ double c2p(double *cons, double *prim, const EosPar &e,
                                      bool &heal_rho, bool &heal_e) {
    double rho = cons[RHO];
    heal_rho = false;
    heal_e = false;
    if (__builtin_isnan(rho) || rho < e.d_floor) {
        rho = e.d_floor;
        cons[RHO] = e.d_floor;
        heal_rho = true;
    }
    prim[RHO] = rho;
    const double rho_inv = 1 / rho;
    double e_k = 0.;
#pragma unroll
    for (int n = MOM; n < MOM + 3; ++n) {
        prim[n] = rho_inv * cons[n];
        e_k += cons[n] * cons[n];
    }
    e_k *= rho_inv / 2;
    const double e_g = cons[ENE] - e_k;
    const double gam = e.gamma;
    const double cv_kb = 1 / (gam - 1);
    prim[ENE] = e_g / cv_kb;
    double cs2 = gam * prim[ENE] / prim[RHO];
    const double pre_floor = e.t_floor * prim[RHO];
    if (__builtin_isnan(prim[ENE]) || prim[ENE] < pre_floor) {
        prim[ENE] = pre_floor;
        cons[ENE] = e_k + pre_floor / (e.gamma - 1);
        cs2 = e.gamma * pre_floor / rho;
        heal_e = true;
    }
    return cs2;
}

// Hydro CFL dt of one cell.  hydro/eos/adiabatic.h:144-157
__device__ __forceinline__ double dt_hydro(const double *v_p, double cs2, const double *dx,
                                           int ndim, double cfl) {
    double dt = (double)FLT_MAX;
    const double cs = ksqrt(cs2);
    for (int a = 0; a < ndim; ++a)
        dt = pmin(dt, dx[a] / (cs + kfabs(v_p[MOM + a])));
    return dt * cfl;
}

// Fast magnetosonic speed.  mhd/eos/adiabatic.h:82-98
template <class FT>
__device__ __forceinline__ FT c_fms(const FT *w, const FT *b_2, FT gam, int axis) {
    const FT rho = w[RHO];
    const FT cs_2 = w[ENE] * gam / rho;
    const FT v_ax2 = b_2[axis] / rho;
    const FT v_t2 = (b_2[mod3(axis + 1)] + b_2[mod3(axis + 2)]) / rho;
    const FT q_2 = v_ax2 + v_t2 + cs_2;
    const FT r_2 = v_ax2 + v_t2 - cs_2;
    return ksqrt((FT)0.5f * (q_2 + ksqrt(r_2 * r_2 + 4 * cs_2 * v_t2)));
}

// MHD CFL dt of one cell, FP64.  mhd/eos/adiabatic.h:103-122
__device__ __forceinline__ double dt_mhd(const double *v_p, const double *bc_2,
                                         const double *b_m, const double *dx, int ndim,
                                         double gamma, double cfl) {
    double dt = (double)FLT_MAX;
    double bm_2[3] = {bc_2[0], bc_2[1], bc_2[2]};
    for (int a = 0; a < ndim; ++a) {
        bm_2[a] = b_m[a] * b_m[a];
        const double c = c_fms<double>(v_p, bm_2, gamma, a);
        dt = pmin(dt, dx[a] / (c + kfabs(v_p[MOM + a])));
        bm_2[a] = bc_2[a];
    }
    return dt * cfl;
}

// The same dt for the mixed form, still in FP64: one reciprocal of rho in
// place of the seven divisions by it (the values differ from dt_mhd's by
// rounding; the FP64 form keeps dt_mhd's expressions for bit parity).
__device__ __forceinline__ double dt_mhd_mixed(const double *v_p, const double *bc_2,
                                               const double *b_m, const double *dx, int ndim,
                                               double gamma, double cfl) {
    const double ri = 1 / v_p[RHO];
    const double cs_2 = v_p[ENE] * gamma * ri;
    double dt = (double)FLT_MAX;
    for (int a = 0; a < ndim; ++a) {
        const double v_ax2 = b_m[a] * b_m[a] * ri;
        const double v_t2 = (bc_2[mod3(a + 1)] + bc_2[mod3(a + 2)]) * ri;
        const double q_2 = v_ax2 + v_t2 + cs_2;
        const double r_2 = v_ax2 + v_t2 - cs_2;
        const double c = ksqrt(0.5 * (q_2 + ksqrt(r_2 * r_2 + 4 * cs_2 * v_t2)));
        dt = pmin(dt, dx[a] / (c + kfabs(v_p[MOM + a])));
    }
    return dt * cfl;
}

// ---------------------------------------------------------------------
// PLM on a uniform mesh.  hydro/reconstruct/plm.h:116-153.  The FP32 form
// rescales the slopes by their largest magnitude (:128-137) to keep the
// products in range; the FP64 form uses norm = 1.
template <class FT>
__device__ __forceinline__ void plm(FT q_ll, FT q_l, FT q_r, FT q_rr, FT &ql, FT &qr) {
    const FT dql = (q_l - q_ll);
    const FT dqc = (q_r - q_l);
    const FT dqr = (q_rr - q_r);
    FT norm(1);
    if constexpr (std::is_same_v<FT, float>) {
        const FT fmin_c = (FT)1e-24f;   // type::float_min at PRECISION 0/1
        if (kfabs(dqc) > fmin_c || kfabs(dqr) > fmin_c || kfabs(dql) > fmin_c) {
            norm = pmax(kfabs(dqr), kfabs(dql));
            norm = 1 / pmax(kfabs(dqc), norm);
        }
    }
    const FT dqr2 = norm * dqr * dqr;
    const FT dql2 = norm * dql * dql;
    const FT dqc2 = norm * dqc * dqc;
    const FT dq2_l = norm * dql * dqc;
    const FT dq2_r = norm * dqr * dqc;
    const FT ql_coef = dq2_l <= 0 ? (FT)0 : dq2_l / (dql2 + dqc2 + dq2_l * 2);
    const FT qr_coef = dq2_r <= 0 ? (FT)0 : -dq2_r / (dqr2 + dqc2 + dq2_r * 2);
    ql = q_l + (dq2_l <= 0 ? (FT)0 : ql_coef * (dql + dqc));
    qr = q_r + (dq2_r <= 0 ? (FT)0 : qr_coef * (dqc + dqr));
}

// ---------------------------------------------------------------------
// et_cs (hydro/eos/adiabatic.h:62-76) in the flux precision.
template <class FT>
__device__ __forceinline__ FT et_cs(FT &e_tot, FT &c_s, const FT *w, FT gam) {
    const FT pre = w[ENE], rho = w[RHO];
    const FT cv_kb = 1 / (gam - 1);
    FT v2(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) v2 += w[MOM + a] * w[MOM + a];
    e_tot = pre * cv_kb + rho * v2 / 2;
    c_s = ksqrt(gam * pmax(pre / rho, (FT)0));
    return cv_kb;
}

template <class FT>
__device__ __forceinline__ void f_lr_hyd(FT *f, const FT *w, FT v, FT e) {
    const FT rho = w[RHO];
    const FT pre = w[ENE];
    f[RHO] = rho * v;
    f[ENE] = pre * w[MOM] + e * v;
#pragma unroll
    for (int a = MOM; a < MOM + 3; ++a) f[a] = rho * w[a] * v;
    f[MOM] += pre;
}

// HLLC in the rotated frame.  hydro/riemann/hllc.h:76-193
template <class FT>
__device__ __forceinline__ void hllc(FT *flx, const FT *wli, const FT *wri, FT gam) {
    FT el, er, cl, cr;
    const FT gm1i_l = et_cs<FT>(el, cl, wli, gam);
    const FT gm1i_r = et_cs<FT>(er, cr, wri, gam);
    const FT sqrtdl = ksqrt(wli[RHO]);
    const FT sqrtdr = ksqrt(wri[RHO]);
    const FT isdlpdr = (FT)1 / (sqrtdl + sqrtdr);
    FT wroe[5];
    wroe[RHO] = sqrtdl * sqrtdr;
#pragma unroll
    for (int n = MOM; n < MOM + 3; ++n) wroe[n] = (sqrtdl * wli[n] + sqrtdr * wri[n]) * isdlpdr;
    const FT hroe = isdlpdr * ((el + wli[ENE]) / sqrtdl + (er + wri[ENE]) / sqrtdr);
    const FT gam1_roe = (FT)1 / (isdlpdr * (sqrtdl * gm1i_l + sqrtdr * gm1i_r));
    const FT v2roe = wroe[MOM] * wroe[MOM] + wroe[MOM + 1] * wroe[MOM + 1] +
                     wroe[MOM + 2] * wroe[MOM + 2];
    const FT q = hroe - (FT)0.5f * v2roe;
    const FT a = (q <= 0) ? (FT)0 : ksqrt(gam1_roe * q);
    const FT al = pmin(wroe[MOM] - a, wli[MOM] - cl);
    const FT ar = pmax(wroe[MOM] + a, wri[MOM] + cr);
    const FT bp = ar > 0 ? ar : (FT)0;
    const FT bm = al < 0 ? al : (FT)0;
    const FT vxl = wli[MOM] - al;
    const FT vxr = wri[MOM] - ar;
    const FT tl = wli[ENE] + vxl * wli[RHO] * wli[MOM];
    const FT tr = wri[ENE] + vxr * wri[RHO] * wri[MOM];
    const FT ml = wli[RHO] * vxl;
    const FT mr = -wri[RHO] * vxr;
    FT cp, am;
    const FT f_max = (FT)(FLT_MAX / 1e1f);
    const FT f_min = (FT)(FLT_MIN * 1e1f);
    if (kfabs(ml + mr) < f_min) {
        cp = 0;
        am = 0;
    } else if ((tr != 0 && kfabs(ml) > f_max / kfabs(tr)) ||
               (tl != 0 && kfabs(mr) > f_max / kfabs(tl))) {
        const FT mli = vsgn(ml) / pmax(kfabs(ml), f_min);
        const FT mri = vsgn(mr) / pmax(kfabs(mr), f_min);
        am = (tl - tr) / (ml + mr);
        cp = pmax((mri * tr + mli * tl) / (mli + mri), (FT)0);
    } else {
        am = (tl - tr) / (ml + mr);
        cp = pmax((ml * tr + mr * tl) / (ml + mr), (FT)0);
    }
    const FT d_am_bm = pmax(kfabs(am - bm), f_min) * (am > bm ? 1 : -1);
    const FT d_am_bp = pmax(kfabs(am - bp), f_min) * (am > bp ? 1 : -1);
    const FT sl = (am >= 0 ? am / d_am_bm : (FT)0);
    const FT sr = (am < 0 ? am / d_am_bp : (FT)0);
    const FT sm = (am > 0 ? -bm / d_am_bm : -bp / d_am_bp);
    FT fl[5], fr[5];
    f_lr_hyd<FT>(fl, wli, wli[MOM] - bm, el);
    f_lr_hyd<FT>(fr, wri, wri[MOM] - bp, er);
#pragma unroll
    for (int n = RHO; n < MOM + 3; ++n) flx[n] = sl * fl[n] + sr * fr[n];
    flx[MOM] += sm * cp;
    flx[ENE] += sm * cp * am;
}

// ---------------------------------------------------------------------
// HLLD in the rotated frame.  mhd/riemann/hlld.h:49-69 (f_flux), 81-298.
template <class FT>
__device__ __forceinline__ void f_flux(FT *flux, const FT *u, const FT *w, FT p_tot, FT b_x) {
    const FT bx_2 = b_x * b_x;
    flux[RHO] = u[MOM];
    flux[ENE] = w[MOM] * (p_tot + u[ENE] - bx_2) -
                b_x * (w[MOM + 1] * u[BCC + 1] + w[MOM + 2] * u[BCC + 2]);
    flux[MOM] = p_tot - bx_2 + u[MOM] * w[MOM];
    flux[MOM + 1] = -b_x * u[BCC + 1] + u[MOM + 1] * w[MOM];
    flux[MOM + 2] = -b_x * u[BCC + 2] + u[MOM + 2] * w[MOM];
    flux[BCC + 1] = -b_x * w[MOM + 1] + u[BCC + 1] * w[MOM];
    flux[BCC + 2] = -b_x * w[MOM + 2] + u[BCC + 2] * w[MOM];
}

template <class FT>
__device__ __forceinline__ FT hlld_cons(FT *u, const FT *w, FT b_x, FT bx_2, FT gam,
                                        FT &gam1) {
    FT p_m = bx_2 / 2, cs;
#pragma unroll
    for (int a = 1; a <= 2; ++a) {
        const int n = BCC + a;
        p_m += w[n] * w[n] / 2;
        u[n] = w[n];
    }
    gam1 = 1 / et_cs<FT>(u[ENE], cs, w, gam);
#pragma unroll
    for (int n = MOM; n < MOM + 3; ++n) u[n] = w[RHO] * w[n];
    u[RHO] = w[RHO];
    u[ENE] += p_m;
    u[BCC] = b_x;
    return p_m;
}

template <class FT>
__device__ __forceinline__ FT hlld_star(FT *u_star, const FT *u, const FT *w, FT rhos, FT sd,
                                        FT sdm, FT p_tot, FT spd0, FT b_x, FT bx_2, FT ptot_s,
                                        FT f_tiny) {
    u_star[RHO] = rhos;
    u_star[MOM] = rhos * spd0;
    u_star[BCC] = b_x;
    const FT denom = u[RHO] * sd * sdm - bx_2;
    if (vabs(denom) < f_tiny * ptot_s) {
#pragma unroll
        for (int a = 1; a <= 2; ++a) {
            u_star[MOM + a] = rhos * w[MOM + a];
            u_star[BCC + a] = u[BCC + a];
        }
    } else {
        const FT tmp_v = b_x * (sd - sdm) / denom;
        const FT tmp_b = (u[RHO] * sd * sd - bx_2) / denom;
#pragma unroll
        for (int a = 1; a <= 2; ++a) {
            u_star[MOM + a] = rhos * (w[MOM + a] - u[BCC + a] * tmp_v);
            u_star[BCC + a] = u[BCC + a] * tmp_b;
        }
    }
    FT u_dot_b(0), w_dot_b(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        u_dot_b += u_star[MOM + a] * u_star[BCC + a];
        w_dot_b += w[MOM + a] * u[BCC + a];
    }
    u_dot_b /= rhos;
    u_star[ENE] = b_x * (w_dot_b - u_dot_b) - p_tot * w[MOM] + sd * u[ENE] + ptot_s * spd0;
    u_star[ENE] /= sdm;
    return u_dot_b;
}

template <class FT>
__device__ __forceinline__ void hlld(FT *flx, const FT *wli, const FT *wri, FT gam) {
    flx[BCC] = 0;
    const FT b_x = wli[BCC];
    const FT bx_2 = b_x * b_x;
    FT u_l[8], u_r[8];
    FT gam1_l, gam1_r;
    const FT pm_l = hlld_cons<FT>(u_l, wli, b_x, bx_2, gam, gam1_l);
    const FT pm_r = hlld_cons<FT>(u_r, wri, b_x, bx_2, gam, gam1_r);
    const FT ptot_l = wli[ENE] + pm_l;
    const FT ptot_r = wri[ENE] + pm_r;
    FT b2_eff_l[3] = {b_x, wli[BCC + 1], wli[BCC + 2]};
    FT b2_eff_r[3] = {b_x, wri[BCC + 1], wri[BCC + 2]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b2_eff_l[a] *= b2_eff_l[a];
        b2_eff_r[a] *= b2_eff_r[a];
    }
    const FT cfm_l = c_fms<FT>(wli, b2_eff_l, gam1_l + 1, 0);
    const FT cfm_r = c_fms<FT>(wri, b2_eff_r, gam1_r + 1, 0);
    const FT spd_m2 = vmin(wli[MOM] - cfm_l, wri[MOM] - cfm_r);
    const FT spd_p2 = vmax(wli[MOM] + cfm_l, wri[MOM] + cfm_r);
    if (spd_m2 >= 0 || spd_p2 <= 0) {
        const bool is_left(spd_m2 >= 0);
        f_flux<FT>(flx, is_left ? u_l : u_r, is_left ? wli : wri, is_left ? ptot_l : ptot_r,
                   b_x);
        return;
    }
    const FT sd_l = spd_m2 - wli[MOM];
    const FT sd_r = spd_p2 - wri[MOM];
    const FT spd0 = (sd_r * u_r[MOM] - ptot_r - sd_l * u_l[MOM] + ptot_l) /
                    (sd_r * u_r[RHO] - sd_l * u_l[RHO]);
    const FT sdm_l = spd_m2 - spd0;
    const FT sdm_r = spd_p2 - spd0;
    const FT rhos_l = u_l[RHO] * sd_l / sdm_l;
    const FT rhos_r = u_r[RHO] * sd_r / sdm_r;
    const FT rhos_l_sqrt = ksqrt(rhos_l);
    const FT rhos_r_sqrt = ksqrt(rhos_r);
    const FT spd_m1 = spd0 - vabs(b_x) / rhos_l_sqrt;
    const FT spd_p1 = spd0 + vabs(b_x) / rhos_r_sqrt;
    const FT f_tiny = (FT)(1.00e-4f);
    const FT ptot_s = ptot_l + u_l[RHO] * sd_l * (sd_l - sdm_l);
    FT us_l[8], us_r[8];
    const FT u_dot_b_l = hlld_star<FT>(us_l, u_l, wli, rhos_l, sd_l, sdm_l, ptot_l, spd0, b_x,
                                       bx_2, ptot_s, f_tiny);
    const FT u_dot_b_r = hlld_star<FT>(us_r, u_r, wri, rhos_r, sd_r, sdm_r, ptot_r, spd0, b_x,
                                       bx_2, ptot_s, f_tiny);
    if (spd_m1 >= 0) {
        f_flux<FT>(flx, u_l, wli, ptot_l, b_x);
#pragma unroll
        for (int n = 0; n < 8; ++n) flx[n] += spd_m2 * (us_l[n] - u_l[n]);
        return;
    } else if (spd_p1 <= 0) {
        f_flux<FT>(flx, u_r, wri, ptot_r, b_x);
#pragma unroll
        for (int n = 0; n < 8; ++n) flx[n] += spd_p2 * (us_r[n] - u_r[n]);
        return;
    }
    FT uss_l[8], uss_r[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        uss_l[n] = us_l[n];
        uss_r[n] = us_r[n];
    }
    if (bx_2 > ptot_s * f_tiny) {
        const int bx_sgn(b_x > 0 ? 1 : -1);
        const FT rhos_norm_inv = 1.f / (rhos_l_sqrt + rhos_r_sqrt);
#pragma unroll
        for (int a = 1; a <= 2; ++a) {
            FT tmp = rhos_norm_inv * (us_l[MOM + a] / rhos_l_sqrt + us_r[MOM + a] / rhos_r_sqrt +
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
        const FT tmp = spd0 * b_x +
                       (uss_l[MOM + 1] * uss_l[BCC + 1] + uss_l[MOM + 2] * uss_l[BCC + 2]) /
                           uss_l[RHO];
        uss_l[ENE] -= rhos_l_sqrt * bx_sgn * (u_dot_b_l - tmp);
        uss_r[ENE] += rhos_r_sqrt * bx_sgn * (u_dot_b_r - tmp);
    }
    if (spd0 >= 0) {
        f_flux<FT>(flx, u_l, wli, ptot_l, b_x);
        const FT d_v = spd_m1 - spd_m2;
#pragma unroll
        for (int n = 0; n < 8; ++n) flx[n] += -spd_m2 * u_l[n] - d_v * us_l[n] + spd_m1 * uss_l[n];
    } else {
        f_flux<FT>(flx, u_r, wri, ptot_r, b_x);
        const FT d_v = spd_p1 - spd_p2;
#pragma unroll
        for (int n = 0; n < 8; ++n) flx[n] += -spd_p2 * u_r[n] - d_v * us_r[n] + spd_p1 * uss_r[n];
    }
}

}  // namespace kr

#include "kr_fast32.h"

namespace kr {

// Face flux in the mesh frame for a sweep along `axis`: hllc.h:32-43 and
// hlld.h:29-47 (the B-flux slots are rewritten as edge EMFs, normal slot 0).
// ql / qr are rotated in place.
// Rotation with a compile-time axis, so that every register array index is
// static (no scratch spills).
template <int AXIS, class T> __device__ __forceinline__ void rot(T *x) {
    if constexpr (AXIS % 3 != 0) {
        const T s0 = x[AXIS % 3], s1 = x[(AXIS + 1) % 3], s2 = x[(AXIS + 2) % 3];
        x[0] = s0;
        x[1] = s1;
        x[2] = s2;
    }
}

template <class FT, bool MHD, int AXIS>
__device__ __forceinline__ void face_flux(FT *ql, FT *qr, FT *flx, FT gam) {
    rot<AXIS>(ql + MOM);
    rot<AXIS>(qr + MOM);
    if constexpr (MHD) {
        rot<AXIS>(ql + BCC);
        rot<AXIS>(qr + BCC);
        if constexpr (std::is_same_v<FT, float>)
            hlld_fast(flx, ql, qr, gam);
        else
            hlld<FT>(flx, ql, qr, gam);
        rot<3 - AXIS>(flx + MOM);
        const FT e0 = flx[BCC + 2], e1 = -flx[BCC + 1];
        flx[BCC + (AXIS + 1) % 3] = e0;
        flx[BCC + (AXIS + 2) % 3] = e1;
        flx[BCC + AXIS] = 0;
    } else {
        if constexpr (std::is_same_v<FT, float>)
            hllc_fast(flx, ql, qr, gam);
        else
            hllc<FT>(flx, ql, qr, gam);
        rot<3 - AXIS>(flx + MOM);
    }
}

}  // namespace kr
