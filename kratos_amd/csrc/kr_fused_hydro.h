// kr_fused_hydro.h -- the fused hydro stage kernel: PLM + HLLC fluxes along
// x, y and z and the flux-divergence Heun update in one launch.
//
// One workgroup owns a TX x TY column tile of one block and marches through
// the block along z, as the MHD stage kernel does, with a 3-plane LDS ring
// of primitives (tile plus a 2-cell halo, brought in by global_load_lds).
// Hydro needs no edge EMFs, so only the tile's own faces are solved: the
// (TX+1) x TY x faces and TX x (TY+1) y faces of each plane, and one z face
// per owned column (the lane of that column keeps its pencil state in
// registers).  Per plane iteration L (leading plane):
//   1. W(L) enters the ring (prefetched during the previous iteration);
//   2. fluxes: z face L-1 of every owned column, which finishes cell L-2
//      (its x, y part was accumulated one iteration earlier), and the
//      in-plane faces of plane L-1 into LDS;
//   3. x + y part of the update of the cells of plane L-1.
// Each face is solved once per tile with one code path; a face on a tile
// seam is solved by both tiles from the same inputs with the same
// instructions, so every face flux has one value and the update conserves
// exactly.  The per-face arithmetic and the order of the update sums are
// those of the staged kernels (k_flux, k_update), so the FP64 form equals
// them, and the oracle, bit for bit, and the mixed form (hllc_fast,
// plm_fast) equals the staged mixed form bit for bit.  hydro/integrator/integrator.h:56-68,
// 156-207, 241-265; hydro/reconstruct/plm.h:116-153; hydro/riemann/
// hllc.h:32-193; rk2.h:20-41.
#pragma once

#include "kr_fused.h"

namespace kr {

template <class FT, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_sweep_hyd(MeshView<FT> m, int step) {
    using S = TileShape<TX, TY>;          // halo columns: HX x HY, col()
    constexpr int OWN = TX * TY;
    constexpr int OWNP = (OWN + 63) / 64 * 64;                 // pencil lanes, padded to waves
    constexpr int NFX = (TX + 1) * TY, NFY = TX * (TY + 1), NF = NFX + NFY;
    static_assert(OWN <= NT, "one owned column per lane");
    if (m.ts[m.geo[row_block(m)].grp].skip) return;
    const int tid = threadIdx.x;
    const int b = row_block(m);
    const int ntx = m.n[0] / TX;
    const int x0 = (blockIdx.x % ntx) * TX, y0 = (blockIdx.x / ntx) * TY;
    const int nz = m.n[2];
    const BlockGeo &g = m.geo[b];
    const double dt = m.ts[g.grp].dt_use * (step == 0 ? 1.0 : 0.5);
    const FT gam = (FT)m.gamma;
    const size_t nc = m.ncell;

    constexpr int DPC = (int)sizeof(FT) / 4;                   // dwords per column
    constexpr int NDW = S::NCOL * DPC;                         // dwords per row
    constexpr int NCHUNK = (NDW + 63) / 64;
    __shared__ FT sW[3][5][S::NCOL];
    __shared__ FT sF[5][NF];            // in-plane face fluxes of one plane, mesh frame
    __shared__ int sNbr[27];
    if (tid < 27) sNbr[tid] = g.wnbr[tid];
    // flux factors sfc_a * dt / vol of the FP32 form, formed once per launch
    // in LDS (the expression of each use, so the same values)
    constexpr bool F32 = std::is_same_v<FT, float>;
    __shared__ double sFa[3];
    if (tid < 3) sFa[tid] = g.sfc[tid] * (dt / g.vol);
    __syncthreads();

    auto fxi = [](int x, int y) { return y * (TX + 1) + x; };
    auto fyi = [](int x, int y) { return NFX + y * TX + x; };

    // primitive source of halo column (cx, cy) at plane z; out-of-block
    // positions resolve through the neighbour table (a halo block at a
    // physical side, filled by k_phys_fill)
    auto wsrc = [&](int cx, int cy, int z) -> const FT * {
        int p[3] = {x0 + cx, y0 + cy, z};
        int d[3];
        bool out = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            d[a] = p[a] < 0 ? -1 : (p[a] >= m.n[a] ? 1 : 0);
            out |= d[a] != 0;
        }
        int blk = b;
        if (out) {
            blk = sNbr[((d[2] + 1) * 3 + (d[1] + 1)) * 3 + (d[0] + 1)];
#pragma unroll
            for (int a = 0; a < 3; ++a) p[a] -= d[a] * m.n[a];
        }
        return m.W + (size_t)blk * 5 * nc + cidx(m.n, p[0], p[1], p[2]);
    };
    // plane z of W into ring slot `slot`: lane l of chunk k carries dword
    // k*64 + l of a variable row; lanes past the row issue no load
    auto issue_w = [&](int z, int slot) {
        const int wave = tid >> 6, lane = tid & 63;
        for (int k = wave; k < NCHUNK; k += NT / 64) {
            const int d = k * 64 + lane;
            if (d >= NDW) continue;
            const int col = d / DPC;
            const int half = d - col * DPC;
            const FT *src = wsrc(col % S::HX - 2, col / S::HX - 2, z);
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const char *gp = (const char *)(src + (size_t)v * nc) + 4 * half;
                __builtin_amdgcn_global_load_lds(gp, &sW[slot][v][0] + k * (64 / DPC), 4, 0, 0);
            }
        }
    };

    // owned column of a pencil lane
    const bool own = tid < OWN;
    const int cx = own ? tid % TX : 0, cy = own ? tid / TX : 0;
    FT qL[5];                 // left state of the next z face
    FT fzp[5];                // z flux of the previous face
    double part[5];           // x + y partial update of the previous plane
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        qL[v] = 0;
        fzp[v] = 0;
        part[v] = 0;
    }
    const double dt_dv = dt / g.vol;

    issue_w(-2, ring3(-2));
    for (int L = -2; L <= nz + 1; ++L) {
        // ------------------------------------------------ 1. plane L arrives
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const int P = L - 2;

        // ------------------------------------------------ 2. fluxes
        if (L >= 0) {
            const int p = L - 1;
            const int s0 = ring3(p), sm = ring3(p - 1), sp = ring3(p + 1);
            if (own) {
                const int c = S::col(cx, cy);
                FT qR[5], qLn[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    const FT q0 = sW[s0][v][c];
                    if constexpr (std::is_same_v<FT, float>) {
                        const float t = plm_inc_fast(sW[sm][v][c], q0, sW[sp][v][c]);
                        qR[v] = q0 - t;
                        qLn[v] = q0 + t;
                    } else {
                        bool flat;
                        const FT t = plm_inc<FT>(sW[sm][v][c], q0, sW[sp][v][c], flat);
                        qR[v] = q0 + (flat ? (FT)0 : -t);
                        qLn[v] = q0 + (flat ? (FT)0 : t);
                    }
                }
                if (p >= 0 && p <= nz) {
                    FT wl[5], wr[5], flx[5];
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        wl[v] = qL[v];
                        wr[v] = qR[v];
                    }
                    rot<2>(wl + MOM);
                    rot<2>(wr + MOM);
                    if constexpr (std::is_same_v<FT, float>)
                        hllc_fast(flx, wl, wr, gam);
                    else
                        hllc<FT>(flx, wl, wr, gam);
                    rot<1>(flx + MOM);
                    if (P >= 0) {
                        // z part of cell P: faces P (previous) and p = P + 1
                        const FT fa = F32 ? (FT)sFa[2] : (FT)(g.sfc[2] * dt_dv);
                        const size_t cc = cidx(m.n, x0 + cx, y0 + cy, P);
                        double *Ut = m.U + ((size_t)b * 2 + (1 - step)) * 5 * nc + cc;
#pragma unroll
                        for (int v = 0; v < 5; ++v) {
                            double res(0);
                            if constexpr (std::is_same_v<FT, double>) {
                                if (fa > 0) res += fa * fzp[v];
                                if (fa > 0) res -= fa * flx[v];
                            } else {
                                if (fa > 0) res += (double)(fa * fzp[v]);
                                if (fa > 0) res -= (double)(fa * flx[v]);
                            }
                            Ut[(size_t)v * nc] = part[v] + res;
                        }
                    }
#pragma unroll
                    for (int v = 0; v < 5; ++v) fzp[v] = flx[v];
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) qL[v] = qLn[v];
            }
            // in-plane faces of plane p; items start at the first lane past
            // the pencil lanes (a wave boundary) and wrap around
            if (p >= 0 && p < nz)
                for (int it = tid - OWNP + (tid < OWNP ? NT : 0); it < NF; it += NT) {
                    if (it < 0) continue;
                    const bool xf = it < NFX;
                    int fx, fy;
                    if (xf) {
                        fx = it % (TX + 1);
                        fy = it / (TX + 1);
                    } else {
                        const int k = it - NFX;
                        fx = k % TX;
                        fy = k / TX;
                    }
                    const int a = xf ? 0 : 1;
                    const int stride = xf ? 1 : S::HX;
                    const int cR = S::col(fx, fy);      // cell on the right of the face
                    FT wl[5], wr[5], flx[5];
                    // rotated frame: slot k holds mesh variable perm(k)
#pragma unroll
                    for (int k = 0; k < 5; ++k) {
                        int v = k;
                        if (k >= MOM) v = MOM + (a + k - MOM) % 3;
                        const FT *q = &sW[s0][v][0];
                        const FT qm2 = q[cR - 2 * stride], qm1 = q[cR - stride];
                        const FT q00 = q[cR], qp1 = q[cR + stride];
                        if constexpr (std::is_same_v<FT, float>) {
                            plm_fast(qm2, qm1, q00, qp1, wl[k], wr[k]);
                        } else {
                            bool fl, fr;
                            const FT tl = plm_inc<FT>(qm2, qm1, q00, fl);
                            const FT tr = plm_inc<FT>(qm1, q00, qp1, fr);
                            wl[k] = qm1 + (fl ? (FT)0 : tl);
                            wr[k] = q00 + (fr ? (FT)0 : -tr);
                        }
                    }
                    if constexpr (std::is_same_v<FT, float>)
                        hllc_fast(flx, wl, wr, gam);
                    else
                        hllc<FT>(flx, wl, wr, gam);
                    sF[RHO][it] = flx[RHO];
                    sF[ENE][it] = flx[ENE];
#pragma unroll
                    for (int k = 0; k < 3; ++k) sF[MOM + (a + k) % 3][it] = flx[MOM + k];
                }
        }
        __syncthreads();

        // plane L+1 into the slot of plane L-2, which phase 2 of this
        // iteration was the last to read; it lands while the update runs
        if (L + 1 <= nz + 1) issue_w(L + 1, ring3(L + 1));

        // ------------------------------------------------ 3. x + y update of plane L-1
        const int q = L - 1;
        if (own && q >= 0 && q < nz) {
            const size_t c = cidx(m.n, x0 + cx, y0 + cy, q);
            const double *Us = m.U + ((size_t)b * 2 + step) * 5 * nc + c;
            const double *Ut = m.U + ((size_t)b * 2 + (1 - step)) * 5 * nc + c;
            double uu[5];
            if (step == 0) {
#pragma unroll
                for (int v = 0; v < 5; ++v) uu[v] = Us[(size_t)v * nc];
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) uu[v] = (Us[(size_t)v * nc] + Ut[(size_t)v * nc]) * 0.5f;
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int i0 = a == 0 ? fxi(cx, cy) : fyi(cx, cy);
                const int i1 = a == 0 ? fxi(cx + 1, cy) : fyi(cx, cy + 1);
                if constexpr (std::is_same_v<FT, double>) {
                    const double fa = g.sfc[a] * dt_dv;
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        double res(0);
                        if (fa > 0) res += fa * sF[v][i0];
                        if (fa > 0) res -= fa * sF[v][i1];
                        uu[v] += res;
                    }
                } else {
                    const float fa = (float)sFa[a];
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        double res(0);
                        if (fa > 0) res += (double)(fa * sF[v][i0]);
                        if (fa > 0) res -= (double)(fa * sF[v][i1]);
                        uu[v] += res;
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) part[v] = uu[v];
        }
    }
}

}  // namespace kr
