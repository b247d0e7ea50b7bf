// kr_fast32.h -- FP32 flux functions of the mixed precision form.
//
// The mixed form is not bound to the reference's rounding (its bar is an L1
// distance to the FP64 oracle), so the FP32 Riemann solve and PLM use the
// hardware reciprocal / square root (v_rcp_f32, v_sqrt_f32, 1 ulp) instead
// of the IEEE-exact division sequences, and the HLLD solution is evaluated
// branch-free: every face computes the L/R star states and the double-star
// state of its side, and the region (supersonic, star, double star) only
// selects the weights.  Wave speeds and intermediate states are those of
// hlld<FT> (mhd/riemann/hlld.h:81-298, Miyoshi & Kusano 2005).  The library
// is built with -ffp-contract=off for the FP64 form's bit parity; these FP32
// functions opt back in to fused multiply-add contraction.
// Included by kr_physics.h after its variable indices (RHO, ENE, MOM, BCC).
#pragma once

namespace kr {

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsq(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Harmonic (van Leer) PLM increment of one cell: dl*dr/(dl+dr) where the
// differences agree in sign, else 0 (hydro/reconstruct/plm.h:116-153).
__device__ __forceinline__ float plm_inc_fast(float qm, float q0, float qp) {
#pragma clang fp contract(fast)
    const float dl = q0 - qm, dr = qp - q0;
    const float pr = dl * dr;
    return pr > 0.f ? pr * frcp(dl + dr) : 0.f;
}

// Face states of the face between cells l and r from the four cells
// ll, l, r, rr (the per-face form of plm<float>).
__device__ __forceinline__ void plm_fast(float q_ll, float q_l, float q_r, float q_rr, float &ql,
                                         float &qr) {
#pragma clang fp contract(fast)
    ql = q_l + plm_inc_fast(q_ll, q_l, q_r);
    qr = q_r - plm_inc_fast(q_l, q_r, q_rr);
}

struct Side32 {
    float r, p, u, v, w, by, bz;   // primitives (rotated frame)
    float e, pt;                   // total energy, total pressure
    float cf;                      // fast magnetosonic speed
};

__device__ __forceinline__ Side32 side32(const float *q, float bx2, float gam, float gm1i) {
#pragma clang fp contract(fast)
    Side32 s;
    s.r = q[RHO];
    s.p = q[ENE];
    s.u = q[MOM];
    s.v = q[MOM + 1];
    s.w = q[MOM + 2];
    s.by = q[BCC + 1];
    s.bz = q[BCC + 2];
    const float ir = frcp(s.r);
    const float bt2 = s.by * s.by + s.bz * s.bz;
    const float pm = 0.5f * (bx2 + bt2);
    s.e = s.p * gm1i + 0.5f * s.r * (s.u * s.u + s.v * s.v + s.w * s.w) + pm;
    s.pt = s.p + pm;
    const float cs2 = gam * s.p * ir, va2 = bx2 * ir, vt2 = bt2 * ir;
    const float q2 = va2 + vt2 + cs2, r2 = va2 + vt2 - cs2;
    s.cf = fsq(0.5f * (q2 + fsq(r2 * r2 + 4.f * cs2 * vt2)));
    return s;
}

// Transverse momentum and field of one side's star state (hlld.h:160-190):
// the Alfven-degenerate case (|denom| < f_tiny p*_T) keeps v_t and B_t.
struct Star32 {
    float my, mz, by, bz;
};

__device__ __forceinline__ Star32 star32(const Side32 &s, float bx, float bx2, float sd,
                                         float smd, float rs, float pts, float tiny) {
#pragma clang fp contract(fast)
    Star32 o;
    const float rsd = s.r * sd;
    const float denom = rsd * smd - bx2;
    const bool deg = fabsf(denom) < tiny * pts;
    const float idn = frcp(denom);
    const float tv = bx * (sd - smd) * idn;
    const float tb = (rsd * sd - bx2) * idn;
    o.my = deg ? rs * s.v : rs * (s.v - s.by * tv);
    o.mz = deg ? rs * s.w : rs * (s.w - s.bz * tv);
    o.by = deg ? s.by : s.by * tb;
    o.bz = deg ? s.bz : s.bz * tb;
    return o;
}

// FP32 HLLD flux in the rotated frame (wl / wr: rho, p, v_n, v_t1, v_t2,
// B_n, B_t1, B_t2).  flx slots as in hlld<FT>: mass, energy, momentum,
// B_n (0), B_t1, B_t2.  The energy of the star state and the physical flux
// are formed for the upwind side X of the contact only (the flux is
// F_X + a (U*_X - U_X) + c (U**_X - U*_X)).
__device__ __forceinline__ void hlld_fast(float *flx, const float *wl, const float *wr, float gam) {
#pragma clang fp contract(fast)
    const float gm1i = frcp(gam - 1.f);
    const float bx = wl[BCC], bx2 = bx * bx;
    const Side32 L = side32(wl, bx2, gam, gm1i);
    const Side32 R = side32(wr, bx2, gam, gm1i);
    const float sl = fminf(L.u - L.cf, R.u - R.cf);
    const float sr = fmaxf(L.u + L.cf, R.u + R.cf);
    const float sdl = sl - L.u, sdr = sr - R.u;
    const float rsdl = L.r * sdl, rsdr = R.r * sdr;
    const float sm = (rsdr * R.u - R.pt - rsdl * L.u + L.pt) * frcp(rsdr - rsdl);
    const float sml = sl - sm, smr = sr - sm;
    const float rsl = rsdl * frcp(sml), rsr = rsdr * frcp(smr);
    const float isql = frsq(rsl), isqr = frsq(rsr);
    const float sql = rsl * isql, sqr = rsr * isqr;
    const float irsl = isql * isql, irsr = isqr * isqr;       // 1 / rho*
    const float abx = fabsf(bx);
    const float ssl = sm - abx * isql, ssr = sm + abx * isqr;
    const float tiny = 1.0e-4f;
    const float pts = L.pt + rsdl * (sdl - sml);
    const Star32 SL = star32(L, bx, bx2, sdl, sml, rsl, pts, tiny);
    const Star32 SR = star32(R, bx, bx2, sdr, smr, rsr, pts, tiny);

    // double-star transverse velocity and field (hlld.h:246-276)
    const float sgn = bx > 0.f ? 1.f : -1.f;
    const float inv = frcp(sql + sqr);
    const float vy = inv * (SL.my * isql + SR.my * isqr + sgn * (SR.by - SL.by));
    const float vz = inv * (SL.mz * isql + SR.mz * isqr + sgn * (SR.bz - SL.bz));
    const float sgq = sgn * sql * sqr;
    const float byss = inv * (sql * SR.by + sqr * SL.by + sgq * (SR.my * irsr - SL.my * irsl));
    const float bzss = inv * (sql * SR.bz + sqr * SL.bz + sgq * (SR.mz * irsr - SL.mz * irsl));
    const float tmp = sm * bx + (vy * byss + vz * bzss);
    const bool dstar = bx2 > pts * tiny;

    // upwind side of the contact
    const bool left = sm >= 0.f;
    const float xr = left ? L.r : R.r, xu = left ? L.u : R.u, xv = left ? L.v : R.v;
    const float xw = left ? L.w : R.w, xby = left ? L.by : R.by, xbz = left ? L.bz : R.bz;
    const float xe = left ? L.e : R.e, xpt = left ? L.pt : R.pt;
    const float sdx = left ? sdl : sdr, smdx = left ? sml : smr;
    const float rsx = left ? rsl : rsr, irsx = left ? irsl : irsr, sqx = left ? sql : sqr;
    const float mysx = left ? SL.my : SR.my, mzsx = left ? SL.mz : SR.mz;
    const float bysx = left ? SL.by : SR.by, bzsx = left ? SL.bz : SR.bz;
    const float s1 = left ? sl : sr;
    const float s2 = left ? ssl : ssr;
    // star energy of side X (hlld.h:192-200)
    const float vbx = xu * bx + xv * xby + xw * xbz;
    const float udb = (rsx * sm * bx + mysx * bysx + mzsx * bzsx) * irsx;
    const float esx = (bx * (vbx - udb) - xpt * xu + sdx * xe + pts * sm) * frcp(smdx);
    const float sx = left ? -sgn : sgn;
    const float ess = esx + sqx * sx * (udb - tmp);
    // double-star conserved (falls back to the star state)
    const float mys = dstar ? rsx * vy : mysx;
    const float mzs = dstar ? rsx * vz : mzsx;
    const float bys = dstar ? byss : bysx;
    const float bzs = dstar ? bzss : bzsx;
    const float es = dstar ? ess : esx;
    const bool super_ = sl >= 0.f || sr <= 0.f;
    const float a = left ? (s1 < 0.f ? s1 : 0.f) : (s1 > 0.f ? s1 : 0.f);
    const float c = left ? (s2 < 0.f ? s2 : 0.f) : (s2 > 0.f ? s2 : 0.f);
    // physical flux of side X
    const float mx = xr * xu;
    float f[7];
    f[0] = mx;
    f[1] = xu * (xe + xpt) - bx * vbx;
    f[2] = mx * xu + xpt - bx2;
    f[3] = mx * xv - bx * xby;
    f[4] = mx * xw - bx * xbz;
    f[5] = xby * xu - bx * xv;
    f[6] = xbz * xu - bx * xw;
    // U_X, U*_X, U**_X
    const float ux[7] = {xr, xe, mx, xr * xv, xr * xw, xby, xbz};
    const float usx[7] = {rsx, esx, rsx * sm, mysx, mzsx, bysx, bzsx};
    const float ussx[7] = {rsx, es, rsx * sm, mys, mzs, bys, bzs};
    float g[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const float v = f[k] + a * (usx[k] - ux[k]) + c * (ussx[k] - usx[k]);
        g[k] = super_ ? f[k] : v;
    }
    flx[RHO] = g[0];
    flx[ENE] = g[1];
    flx[MOM] = g[2];
    flx[MOM + 1] = g[3];
    flx[MOM + 2] = g[4];
    flx[BCC] = 0.f;
    flx[BCC + 1] = g[5];
    flx[BCC + 2] = g[6];
}

// HLLC of the mixed form (hydro/riemann/hllc.h:76-193, the wave speeds,
// contact speed and FLT-constant guards of hllc<FT>) with the hardware
// reciprocal and square root in place of the IEEE division / square-root
// sequences, 1/(gamma-1) and 1/rho formed once per side.  No multiply-add
// contraction here: every call site (the fused kernel's z and in-plane
// faces, the staged flux kernel) rounds the same way, so a face shared by
// two of them gets one flux and the update stays conservative.
__device__ __forceinline__ void hllc_fast(float *flx, const float *wl, const float *wr, float gam) {
#pragma clang fp contract(off)
    const float cv_kb = frcp(gam - 1.f);
    const float irl = frcp(wl[RHO]), irr = frcp(wr[RHO]);
    const float v2l = wl[MOM] * wl[MOM] + wl[MOM + 1] * wl[MOM + 1] + wl[MOM + 2] * wl[MOM + 2];
    const float v2r = wr[MOM] * wr[MOM] + wr[MOM + 1] * wr[MOM + 1] + wr[MOM + 2] * wr[MOM + 2];
    const float el = wl[ENE] * cv_kb + wl[RHO] * v2l * 0.5f;
    const float er = wr[ENE] * cv_kb + wr[RHO] * v2r * 0.5f;
    const float cl = fsq(gam * pmax(wl[ENE] * irl, 0.f));
    const float cr = fsq(gam * pmax(wr[ENE] * irr, 0.f));
    const float sdl = fsq(wl[RHO]), sdr = fsq(wr[RHO]);
    const float isd = frcp(sdl + sdr);
    float wroe[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) wroe[n] = (sdl * wl[MOM + n] + sdr * wr[MOM + n]) * isd;
    // (e + p) / sqrt(rho) = (e + p) * sqrt(rho) / rho
    const float hroe = isd * ((el + wl[ENE]) * sdl * irl + (er + wr[ENE]) * sdr * irr);
    // gamma - 1 of the Roe state: both sides share gamma, so the weighted
    // mean of 1 / (gamma - 1) is cv_kb itself
    const float gam1_roe = frcp(isd * (sdl * cv_kb + sdr * cv_kb));
    const float v2roe = wroe[0] * wroe[0] + wroe[1] * wroe[1] + wroe[2] * wroe[2];
    const float q = hroe - 0.5f * v2roe;
    const float a = q <= 0.f ? 0.f : fsq(gam1_roe * q);
    const float al = pmin(wroe[0] - a, wl[MOM] - cl);
    const float ar = pmax(wroe[0] + a, wr[MOM] + cr);
    const float bp = ar > 0.f ? ar : 0.f;
    const float bm = al < 0.f ? al : 0.f;
    const float vxl = wl[MOM] - al, vxr = wr[MOM] - ar;
    const float tl = wl[ENE] + vxl * wl[RHO] * wl[MOM];
    const float tr = wr[ENE] + vxr * wr[RHO] * wr[MOM];
    const float ml = wl[RHO] * vxl, mr = -wr[RHO] * vxr;
    const float f_max = FLT_MAX / 1e1f, f_min = FLT_MIN * 1e1f;
    float cp, am;
    const float msum = ml + mr;
    if (kfabs(msum) < f_min) {
        cp = 0.f;
        am = 0.f;
    } else {
        const float ims = frcp(msum);
        am = (tl - tr) * ims;
        if ((tr != 0.f && kfabs(ml) > f_max * frcp(kfabs(tr))) ||
            (tl != 0.f && kfabs(mr) > f_max * frcp(kfabs(tl)))) {
            const float mli = (float)vsgn(ml) * frcp(pmax(kfabs(ml), f_min));
            const float mri = (float)vsgn(mr) * frcp(pmax(kfabs(mr), f_min));
            cp = pmax((mri * tr + mli * tl) * frcp(mli + mri), 0.f);
        } else {
            cp = pmax((ml * tr + mr * tl) * ims, 0.f);
        }
    }
    const float d_am_bm = pmax(kfabs(am - bm), f_min) * (am > bm ? 1.f : -1.f);
    const float d_am_bp = pmax(kfabs(am - bp), f_min) * (am > bp ? 1.f : -1.f);
    const float i_bm = frcp(d_am_bm), i_bp = frcp(d_am_bp);
    const float sl = am >= 0.f ? am * i_bm : 0.f;
    const float sr = am < 0.f ? am * i_bp : 0.f;
    const float sm = am > 0.f ? -bm * i_bm : -bp * i_bp;
    const float vl = wl[MOM] - bm, vr = wr[MOM] - bp;
    flx[RHO] = sl * (wl[RHO] * vl) + sr * (wr[RHO] * vr);
    flx[ENE] = sl * (wl[ENE] * wl[MOM] + el * vl) + sr * (wr[ENE] * wr[MOM] + er * vr);
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        float fl = wl[RHO] * wl[MOM + n] * vl, fr = wr[RHO] * wr[MOM + n] * vr;
        if (n == 0) {
            fl += wl[ENE];
            fr += wr[ENE];
        }
        flx[MOM + n] = sl * fl + sr * fr;
    }
    flx[MOM] += sm * cp;
    flx[ENE] += sm * cp * am;
}

}  // namespace kr
