// kr_sweep_mx.h -- fused MHD stage kernel of the mixed form (FP32 fluxes,
// FP64 state) on meshes of one level without physical sides: uniform
// periodic meshes and ensembles of periodic boxes.
//
// The tile march, the LDS ring of primitive planes and the update / CT of
// k_sweep (kr_fused.h) are kept; the flux and EMF phases are reorganised so
// that each face or edge costs fewer VALU instructions:
//   * a cell pass, run in the update phase of the iteration in which plane
//     L arrives, forms once per cell of the tile region (columns -1..TX,
//     rows -1..TY) the harmonic PLM increments along x and y of the seven
//     transverse variables, and loads the three normal face fields of the
//     cell's low faces; the next iteration's in-plane faces then read their
//     two face states as q + t / q - t from LDS instead of forming two
//     increments per variable per face, and issue no global load;
//   * x faces, y faces and z pencils run in whole waves of one kind, and so
//     do the three edge kinds of the EMF phase, so that every branch on the
//     kind is uniform across the wave and the rotation of the frame is a
//     compile-time permutation.
// The arithmetic per face is the mixed form's (plm_inc_fast, hlld_fast,
// the corner-upwind edge EMF with contraction, FP64 update and CT), so a
// face shared by two tiles (or two blocks) goes through the same
// instructions in both and gets one flux and one EMF.  Every block of a
// mesh runs this kernel or none does.
#pragma once

#include "kr_fused.h"

namespace kr {

template <int TX, int TY>
struct MxTile {
    using S = TileShape<TX, TY>;
    static constexpr int RX = TX + 2, NR = (TX + 2) * (TY + 2);   // region R: cols -1..TX, rows -1..TY
    static constexpr int WZ = (S::NZI + 63) / 64;                 // waves of z pencils
    static constexpr int WFX = (S::NFX + 63) / 64, WFY = (S::NFY + 63) / 64;
    static constexpr int WEZ = (S::NEZ + 63) / 64, WEX = (S::NEX + 63) / 64, WEY = (S::NEY + 63) / 64;
    static constexpr int WR = (NR + 63) / 64;
    __device__ static int rg(int x, int y) { return (y + 1) * RX + (x + 1); }
};

// Grid launch (LIST = false): one workgroup per (tile, block row), marching
// the whole block, planes [0, n_z).  Work-list launch (LIST = true,
// ensembles): a persistent grid walks the items (block, tile, z0, z1) that
// k_sweep_list wrote for the boxes still running, item i on workgroup
// i mod gridDim.x.  A partial range [z0, z1) repeats the march's lead-in
// planes z0 - 2 .. z0 and computes the faces and edges at its ends as the
// neighbouring range does, through the same code and from the same inputs,
// and updates only its own cells and faces, so a split changes no result.
// The list form's item loop costs registers; its 256- and 192-thread tiles
// are held to 3 waves per SIMD (168 VGPRs), the occupancy their LDS allows.
template <int TX, int TY, int NT, bool LIST>
__global__ void __launch_bounds__(NT, (LIST && NT <= 256) ? 3 : 1)
    k_sweep_mx(MeshView<float> m, int step, const int4 *items, const int *n_items) {
    using FT = float;
    using S = TileShape<TX, TY>;
    using M = MxTile<TX, TY>;
    constexpr int NW = NT / 64;
    static_assert(S::NCOL <= 4 * NT, "the plane load covers the halo in at most four chunks");
    static_assert(M::WZ <= NW, "z pencils on the first waves, one pass");
    static_assert(TX * TY + TX + TY <= NT, "one owned cell / edge face per thread");
    static_assert(M::WR <= NW, "one cell-pass item per thread at most");
    for (int item = (int)blockIdx.x;; item += (int)gridDim.x) {
        unsigned row, tile;
        int z0, z1;
        if constexpr (LIST) {
            if (item >= *n_items) break;
            const int4 it = items[item];
            row = (unsigned)it.x;
            tile = (unsigned)it.y;
            z0 = it.z;
            z1 = it.w;
        } else {
            row = blockIdx.y;
            tile = blockIdx.x;
            z0 = 0;
            z1 = m.n[2];
        }
        if (m.ts[m.geo[m.blist ? m.blist[row] : (int)row].grp].skip) {
            if constexpr (LIST) continue;
            else return;
        }
        const int tid = threadIdx.x, lane = tid & 63;
        // the wave's role in a scalar register: the kind branches below are
        // uniform; rtid is the thread's index within the roles
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        const int rtid = wave * 64 + lane;
        const int b = m.blist ? m.blist[row] : (int)row;
        const int ntx = m.n[0] / TX;
        const int x0 = (tile % ntx) * TX, y0 = (tile / ntx) * TY;
        const int nz = m.n[2];
        const bool xlast = x0 + TX == m.n[0], ylast = y0 + TY == m.n[1];
        const BlockGeo &g = m.geo[b];
        const double dt = m.ts[g.grp].dt_use * (step == 0 ? 1.0 : 0.5);
        const FT gam = (FT)m.gamma;
        const size_t nc = m.ncell;

        constexpr int NDW = S::NCOL;                    // dwords per W row (FP32)
        constexpr int NCHUNK = (NDW + 63) / 64;
        __shared__ FT sW[3][8][S::NCOL];
        __shared__ FT sFe[2][2][S::NF];
        __shared__ FT sFez[S::NF];
        __shared__ FT sFh[4][S::NF];
        __shared__ FT sFz[3][S::NZG];
        __shared__ FT sEz[2][S::NEZ];
        __shared__ FT sEx[2][S::NEX];
        __shared__ FT sEy[2][S::NEY];
        // cell pass output: PLM increments along x (rho, p, vx, vy, vz, By, Bz)
        // and along y (rho, p, vy, vz, vx, Bz, Bx -- the y frame's order), and
        // the normal fields Bx, By, Bz of the cell's low faces
        __shared__ FT sTx[7][M::NR];
        __shared__ FT sTy[7][M::NR];
        __shared__ FT sBn[3][M::NR];
        __shared__ int sNbr[27];
        __shared__ double sGeo[9];
        if (tid < 27) sNbr[tid] = g.wnbr[tid];
        if (tid < 3) {
            const double sfc = g.sfc[tid];
            sGeo[tid] = sfc > 0 ? (double)(-dt / sfc) : (double)(0);
            sGeo[3 + tid] = sfc * (dt / g.vol);
            sGeo[6 + tid] = g.dx[tid];
        }
        __syncthreads();

        // W source of halo column (cx, cy) at plane z: the block or a neighbour
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
            return m.W + (size_t)blk * 8 * nc + cidx(m.n, p[0], p[1], p[2]);
        };
        // asynchronous load of plane z of W into ring slot `slot`
        auto issue_w = [&](int z, int slot) {
            for (int k = wave; k < NCHUNK; k += NW) {
                const int d = k * 64 + lane;
                if (d >= NDW) continue;                 // past the row: no load, no write
                const FT *src = wsrc(d % S::HX - 2, d / S::HX - 2, z);
#pragma unroll
                for (int v = 0; v < 8; ++v)
                    __builtin_amdgcn_global_load_lds(src + (size_t)v * nc, &sW[slot][v][0] + k * 64, 4, 0,
                                                     0);
            }
        };
        // cell pass of plane L (see the header), one region cell per thread at
        // most, on the waves from the last: its three normal face fields (slot
        // `step` of BF, which the launch only reads) are requested one
        // iteration ahead (cp_load) and written to LDS with the increments
        // (cp_finish), so that no wave waits for them
        // The item's coordinates are formed in the update phase from an opaque
        // copy of the thread index, so that they are not kept in registers
        // across the flux phase, where the kernel's register use peaks.
        struct CpItem {
            int r, x, y;
            bool corner;
        };
        auto cp_item = [&]() {
            int t = rtid;
            asm volatile("" : "+v"(t));
            CpItem it;
            it.r = (NW - 1 - (t >> 6)) * 64 + (t & 63);
            it.x = it.r % M::RX - 1;
            it.y = it.r / M::RX - 1;
            it.corner = (it.x < 0 || it.x >= TX) && (it.y < 0 || it.y >= TY);
            return it;
        };
        // address of a normal face field: component c of face (fx, fy, fz), the
        // transverse indices possibly in a neighbour block
        auto bf_ptr = [&](int c, int fx, int fy, int fz) -> const double * {
            int p[3] = {x0 + fx, y0 + fy, fz};
            int d[3] = {0, 0, 0};
            bool out = false;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (a == c) continue;
                d[a] = p[a] < 0 ? -1 : (p[a] >= m.n[a] ? 1 : 0);
                out |= d[a] != 0;
            }
            int blk = b;
            if (out) {
                blk = sNbr[((d[2] + 1) * 3 + (d[1] + 1)) * 3 + (d[0] + 1)];
#pragma unroll
                for (int a = 0; a < 3; ++a) p[a] -= d[a] * m.n[a];
            }
            return m.BF + (((size_t)blk * 2 + step) * 3 + c) * m.nface + fidx(m.n, p[0], p[1], p[2]);
        };
        // the loads are unconditional, from clamped (valid) faces where the
        // value is not used, so that all three are in flight together
        auto cp_load = [&](const CpItem &it, int L, double *bv) {
            bv[0] = *bf_ptr(0, it.x < 0 ? 0 : it.x, it.y, L);
            bv[1] = *bf_ptr(1, it.x, it.y < 0 ? 0 : it.y, L);
            bv[2] = *bf_ptr(2, it.x, it.y, L < 0 ? 0 : (L > nz ? nz : L));
        };
        auto cp_finish = [&](const CpItem &it, int L, const double *bv) {
            const int sl = ring3(L);
            const int r = it.r;
            const int c = S::col(it.x, it.y);
            constexpr int vx_[7] = {RHO, ENE, MOM, MOM + 1, MOM + 2, BCC + 1, BCC + 2};
            constexpr int vy_[7] = {RHO, ENE, MOM + 1, MOM + 2, MOM, BCC + 2, BCC};
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const FT *q = &sW[sl][vx_[k]][0];
                sTx[k][r] = plm_inc_fast(q[c - 1], q[c], q[c + 1]);
            }
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const FT *q = &sW[sl][vy_[k]][0];
                sTy[k][r] = plm_inc_fast(q[c - S::HX], q[c], q[c + S::HX]);
            }
            sBn[0][r] = it.x >= 0 ? (FT)bv[0] : (FT)0;
            sBn[1][r] = it.y >= 0 ? (FT)bv[1] : (FT)0;
            sBn[2][r] = (L >= 0 && L <= nz && !it.corner) ? (FT)bv[2] : (FT)0;
        };

        // ---- per-thread persistent state (z pencil owners, rtid < NZI)
        int zx = 0, zy = 0;
        const bool zown = rtid < S::NZI;
        if (zown) S::ring_item(rtid, zx, zy);
        const bool zcorner = (zx < 0 || zx >= TX) && (zy < 0 || zy >= TY);
        const bool zint = zown && zx >= 0 && zx < TX && zy >= 0 && zy < TY;
        FT qL[8];
        FT fzp[5];
        double part[5];
#pragma unroll
        for (int v = 0; v < 8; ++v) qL[v] = 0;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            fzp[v] = 0;
            part[v] = 0;
        }
        const int cx_ = zx, cy_ = zy;
        const int et = rtid - TX * TY;
        const bool xedge = xlast && et >= 0 && et < TY;
        const bool yedge = ylast && et >= TY && et < TY + TX;

        double bvn[3] = {0, 0, 0};       // normal face fields of the next cell pass
        issue_w(z0 - 2, ring3(z0 - 2));
        for (int L = z0 - 2; L <= z1 + 1; ++L) {
            // ------------------------------------------------ 1. plane L arrives
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            const int P = L - 2;

            // ------------------------------------------------ 2. fluxes
            if (L >= z0) {
                const int p = L - 1;
                const int s0 = ring3(p), sm = ring3(p - 1), sp = ring3(p + 1);
                const int fe = ring2(p);
                const bool ghost_plane = p < 0 || p >= nz;
                for (int wi = wave; wi < M::WZ + M::WFX + M::WFY; wi += NW) {
                    if (wi < M::WZ) {
                        // z pencil of this thread (wi == wave): increment of cell
                        // p, face states, flux at face p
                        if (!zown || zcorner) continue;
                        const int c = S::col(zx, zy);
                        FT qR[8], qLn[8];
#pragma unroll
                        for (int v = 0; v < 8; ++v) {
                            if (v == BCC + 2) {
                                qR[v] = 0;
                                qLn[v] = 0;
                                continue;
                            }
                            const FT q0 = sW[s0][v][c];
                            const float t = plm_inc_fast(sW[sm][v][c], q0, sW[sp][v][c]);
                            qR[v] = q0 - t;
                            qLn[v] = q0 + t;
                        }
                        if (p >= z0 && p <= nz) {
                            FT wl[8], wr[8], flx[8];
#pragma unroll
                            for (int v = 0; v < 8; ++v) {
                                wl[v] = qL[v];
                                wr[v] = qR[v];
                            }
                            const FT bn = sBn[2][M::rg(zx, zy)];
                            wl[BCC + 2] = bn;
                            wr[BCC + 2] = bn;
                            rot<2>(wl + MOM);
                            rot<2>(wr + MOM);
                            rot<2>(wl + BCC);
                            rot<2>(wr + BCC);
                            hlld_fast(flx, wl, wr, gam);
                            rot<1>(flx + MOM);
                            const int gz = S::zg(zx, zy);
                            sFz[0][gz] = flx[RHO];
                            sFz[1][gz] = flx[BCC + 2];
                            sFz[2][gz] = -flx[BCC + 1];
                            if (zint) {
                                if (P >= z0) {
                                    const FT fa = (FT)sGeo[3 + 2];
                                    const size_t cc = cidx(m.n, x0 + cx_, y0 + cy_, P);
                                    double *Ut = m.U + ((size_t)b * 2 + (1 - step)) * 5 * nc + cc;
#pragma unroll
                                    for (int v = 0; v < 5; ++v) {
                                        double res(0);
                                        if (fa > 0) res += (double)(fa * fzp[v]);
                                        if (fa > 0) res -= (double)(fa * flx[v]);
                                        Ut[(size_t)v * nc] = part[v] + res;
                                    }
                                }
#pragma unroll
                                for (int v = 0; v < 5; ++v) fzp[v] = flx[v];
                            }
                        }
#pragma unroll
                        for (int v = 0; v < 8; ++v) qL[v] = qLn[v];
                    } else if (wi < M::WZ + M::WFX) {
                        // x face (fx, fy) of plane p, fx = 0..TX, fy = -1..TY:
                        // the x frame is the mesh frame
                        const int it = (wi - M::WZ) * 64 + lane;
                        if (it >= S::NFX) continue;
                        const int fx = it % (TX + 1), fy = it / (TX + 1) - 1;
                        if (ghost_plane && (fy < 0 || fy >= TY)) continue;
                        const int cl = S::col(fx - 1, fy), cr = S::col(fx, fy);
                        const int rl = M::rg(fx - 1, fy), rr = M::rg(fx, fy);
                        constexpr int vx_[7] = {RHO, ENE, MOM, MOM + 1, MOM + 2, BCC + 1, BCC + 2};
                        FT wl[8], wr[8], flx[8];
#pragma unroll
                        for (int k = 0; k < 7; ++k) {
                            const int v = vx_[k];
                            wl[v] = sW[s0][v][cl] + sTx[k][rl];
                            wr[v] = sW[s0][v][cr] - sTx[k][rr];
                        }
                        const FT bn = sBn[0][rr];
                        wl[BCC] = bn;
                        wr[BCC] = bn;
                        hlld_fast(flx, wl, wr, gam);
                        sFh[0][it] = flx[ENE];
                        sFh[1][it] = flx[MOM];
                        sFh[2][it] = flx[MOM + 1];
                        sFh[3][it] = flx[MOM + 2];
                        sFe[fe][0][it] = flx[RHO];
                        sFe[fe][1][it] = flx[BCC + 2];     // E_y
                        sFez[it] = -flx[BCC + 1];          // E_z
                    } else {
                        // y face (fx, fy) of plane p, fx = -1..TX, fy = 0..TY: the
                        // y frame (n, t1, t2) = (y, z, x)
                        const int k0 = (wi - M::WZ - M::WFX) * 64 + lane;
                        if (k0 >= S::NFY) continue;
                        const int fx = k0 % (TX + 2) - 1, fy = k0 / (TX + 2);
                        if (ghost_plane && (fx < 0 || fx >= TX)) continue;
                        const int cl = S::col(fx, fy - 1), cr = S::col(fx, fy);
                        const int rl = M::rg(fx, fy - 1), rr = M::rg(fx, fy);
                        // rotated slot k <- mesh variable vy_[k]
                        constexpr int sl_[7] = {RHO, ENE, MOM, MOM + 1, MOM + 2, BCC + 1, BCC + 2};
                        constexpr int vy_[7] = {RHO, ENE, MOM + 1, MOM + 2, MOM, BCC + 2, BCC};
                        FT wl[8], wr[8], flx[8];
#pragma unroll
                        for (int k = 0; k < 7; ++k) {
                            wl[sl_[k]] = sW[s0][vy_[k]][cl] + sTy[k][rl];
                            wr[sl_[k]] = sW[s0][vy_[k]][cr] - sTy[k][rr];
                        }
                        const FT bn = sBn[1][rr];
                        wl[BCC] = bn;
                        wr[BCC] = bn;
                        hlld_fast(flx, wl, wr, gam);
                        const int it = S::NFX + k0;
                        sFh[0][it] = flx[ENE];
                        sFh[2][it] = flx[MOM];             // y momentum (normal)
                        sFh[3][it] = flx[MOM + 1];         // z momentum
                        sFh[1][it] = flx[MOM + 2];         // x momentum
                        sFe[fe][0][it] = flx[RHO];
                        sFe[fe][1][it] = -flx[BCC + 1];    // E_x
                        sFez[it] = flx[BCC + 2];           // E_z
                    }
                }
            }
            __syncthreads();

            // ------------------------------------------------ 3. EMFs
            if (L >= z0 + 1) {
                const int p = L - 1;
                const int s0 = ring3(p), sm = ring3(p - 1);
                const int fe = ring2(p), fm = ring2(p - 1);
                auto ec = [&](int s, int x, int y, int c) -> FT {
                    const int cc = S::col(x, y);
                    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                    return cross_f32(sW[s][BCC + c1][cc], sW[s][MOM + c2][cc], sW[s][BCC + c2][cc],
                                     sW[s][MOM + c1][cc]);
                };
                for (int wi = wave; wi < M::WEZ + M::WEX + M::WEY; wi += NW) {
                    if (wi < M::WEZ) {
                        const int it = wi * 64 + lane;
                        if (it >= S::NEZ || p >= nz) continue;
                        const int x = it % (TX + 1), y = it / (TX + 1);
                        // ee_z on plane p: a1 = x, a2 = y
                        sEz[fe][S::ezi(x, y)] = edge_emf<FT>(
                            ec(s0, x, y, 2), ec(s0, x - 1, y, 2), ec(s0, x - 1, y - 1, 2), ec(s0, x, y - 1, 2),
                            sFez[S::fxi(x, y - 1)], sFez[S::fxi(x, y)], sFe[fe][0][S::fyi(x, y)],
                            sFe[fe][0][S::fyi(x - 1, y)], sFez[S::fyi(x - 1, y)], sFez[S::fyi(x, y)],
                            sFe[fe][0][S::fxi(x, y)], sFe[fe][0][S::fxi(x, y - 1)]);
                    } else if (wi < M::WEZ + M::WEX) {
                        const int k = (wi - M::WEZ) * 64 + lane;
                        if (k >= S::NEX) continue;
                        const int x = k % TX, y = k / TX;
                        // ee_x on z face p: a1 = y, a2 = z
                        sEx[fe][S::exi(x, y)] = edge_emf<FT>(
                            ec(s0, x, y, 0), ec(s0, x, y - 1, 0), ec(sm, x, y - 1, 0), ec(sm, x, y, 0),
                            sFe[fm][1][S::fyi(x, y)], sFe[fe][1][S::fyi(x, y)], sFz[0][S::zg(x, y)],
                            sFz[0][S::zg(x, y - 1)], sFz[1][S::zg(x, y - 1)], sFz[1][S::zg(x, y)],
                            sFe[fe][0][S::fyi(x, y)], sFe[fm][0][S::fyi(x, y)]);
                    } else {
                        const int k = (wi - M::WEZ - M::WEX) * 64 + lane;
                        if (k >= S::NEY) continue;
                        const int x = k % (TX + 1), y = k / (TX + 1);
                        // ee_y on z face p: a1 = z, a2 = x
                        sEy[fe][S::eyi(x, y)] = edge_emf<FT>(
                            ec(s0, x, y, 1), ec(sm, x, y, 1), ec(sm, x - 1, y, 1), ec(s0, x - 1, y, 1),
                            sFz[2][S::zg(x - 1, y)], sFz[2][S::zg(x, y)], sFe[fe][0][S::fxi(x, y)],
                            sFe[fm][0][S::fxi(x, y)], sFe[fm][1][S::fxi(x, y)], sFe[fe][1][S::fxi(x, y)],
                            sFz[0][S::zg(x, y)], sFz[0][S::zg(x - 1, y)]);
                    }
                }
            }
            __syncthreads();

            // ------------------------------------------------ 4. cell pass, next plane, update
            // the cell pass of plane L (read by the next iteration's faces) and
            // plane L+1 into the slot of plane L-2, which phases 2 and 3 of this
            // iteration were the last to read
            const CpItem cpi = cp_item();
            const bool cp_own = cpi.r < M::NR;
            const bool cp_now = cp_own && L >= z0 - 1 && L <= z1;
            const double bv[3] = {bvn[0], bvn[1], bvn[2]};
            if (cp_own && L + 1 >= z0 - 1 && L + 1 <= z1) cp_load(cpi, L + 1, bvn);
            if (L + 1 <= z1 + 1) issue_w(L + 1, ring3(L + 1));

            if (L >= z0 + 1) {
                const int q = L - 1;
                const int fq = ring2(q), fpp = ring2(P);
                // Face-field update of one face (mhd/integrator/rk2.h:48-62): its
                // two slots are read first (face_load) for every face a thread
                // updates, together with the conserved state, and written last
                // (face_store), so that the loads of one iteration are in
                // flight together rather than one face after the other.
                struct FaceUpd {
                    double *bt;
                    double s, t;
                };
                auto face_load = [&](int c, int fx, int fy, int fz) {
                    const size_t f = fidx(m.n, x0 + fx, y0 + fy, fz);
                    FaceUpd u;
                    u.bt = m.BF + (((size_t)b * 2 + (1 - step)) * 3 + c) * m.nface + f;
                    u.s = m.BF[(((size_t)b * 2 + step) * 3 + c) * m.nface + f];
                    u.t = step == 0 ? 0.0 : *u.bt;
                    return u;
                };
                auto face_store = [&](const FaceUpd &u, double db) {
                    *u.bt = db + (step == 0 ? u.s : (u.s + u.t) / 2);
                };
                // CT circulation of one face (mhd/integrator/integrator.h:429-457)
                auto ct = [&](double e1, double d1, double e2, double d2, double e3, double e4, int a) {
                    double db = e1 * d1;
                    db -= e2 * d2;
                    db += e3 * d2;
                    db -= e4 * d1;
                    db *= sGeo[a];
                    return db;
                };
                const double dx = sGeo[6], dy = sGeo[7], dz = sGeo[8];
                if (zint) {
                    const int x = cx_, y = cy_;
                    const FaceUpd uz = face_load(2, x, y, q);
                    const FaceUpd ux = face_load(0, x, y, P < 0 ? 0 : P);
                    const FaceUpd uy = face_load(1, x, y, P < 0 ? 0 : P);
                    if (q < z1) {
                        const size_t c = cidx(m.n, x0 + x, y0 + y, q);
                        const double *Us = m.U + ((size_t)b * 2 + step) * 5 * nc + c;
                        const double *Ut = m.U + ((size_t)b * 2 + (1 - step)) * 5 * nc + c;
                        double uu[5];
                        if (step == 0) {
#pragma unroll
                            for (int v = 0; v < 5; ++v) uu[v] = Us[(size_t)v * nc];
                        } else {
#pragma unroll
                            for (int v = 0; v < 5; ++v)
                                uu[v] = (Us[(size_t)v * nc] + Ut[(size_t)v * nc]) * 0.5f;
                        }
#pragma unroll
                        for (int a = 0; a < 2; ++a) {
                            const int i0 = a == 0 ? S::fxi(x, y) : S::fyi(x, y);
                            const int i1 = a == 0 ? S::fxi(x + 1, y) : S::fyi(x, y + 1);
                            const float fa = (float)sGeo[3 + a];
#pragma unroll
                            for (int v = 0; v < 5; ++v) {
                                double res(0);
                                const FT f0 = v == 0 ? sFe[fq][0][i0] : sFh[v - 1][i0];
                                const FT f1 = v == 0 ? sFe[fq][0][i1] : sFh[v - 1][i1];
                                if (fa > 0) res += (double)(fa * f0);
                                if (fa > 0) res -= (double)(fa * f1);
                                uu[v] += res;
                            }
                        }
#pragma unroll
                        for (int v = 0; v < 5; ++v) part[v] = uu[v];
                    }
                    // z face q: the range's own faces, and the block's top face
                    // in the last range
                    if (q < z1 || z1 == nz)
                        face_store(uz, ct(sEx[fq][S::exi(x, y)], dx, sEy[fq][S::eyi(x, y)], dy,
                                          sEy[fq][S::eyi(x + 1, y)], sEx[fq][S::exi(x, y + 1)], 2));
                    if (P >= z0) {
                        face_store(ux, ct(sEy[fpp][S::eyi(x, y)], dy, sEz[fpp][S::ezi(x, y)], dz,
                                          sEz[fpp][S::ezi(x, y + 1)], sEy[fq][S::eyi(x, y)], 0));
                        face_store(uy, ct(sEz[fpp][S::ezi(x, y)], dz, sEx[fpp][S::exi(x, y)], dx,
                                          sEx[fq][S::exi(x, y)], sEz[fpp][S::ezi(x + 1, y)], 1));
                    }
                }
                if (P >= z0 && xedge) {
                    const int x = TX, y = et;
                    const FaceUpd ux = face_load(0, x, y, P);
                    face_store(ux, ct(sEy[fpp][S::eyi(x, y)], dy, sEz[fpp][S::ezi(x, y)], dz,
                                      sEz[fpp][S::ezi(x, y + 1)], sEy[fq][S::eyi(x, y)], 0));
                }
                if (P >= z0 && yedge) {
                    const int x = et - TY, y = TY;
                    const FaceUpd uy = face_load(1, x, y, P);
                    face_store(uy, ct(sEz[fpp][S::ezi(x, y)], dz, sEx[fpp][S::exi(x, y)], dx,
                                      sEx[fq][S::exi(x, y)], sEz[fpp][S::ezi(x + 1, y)], 1));
                }
            }
            if (cp_now) cp_finish(cpi, L, bv);
        }
        if constexpr (!LIST) break;
        __syncthreads();          // the next item reuses the LDS
    }
}

// Work list of a stage's k_sweep_mx launch: every tile of the blocks whose
// time group still runs, in block, tile, range order.  With T such tiles
// and S resident workgroup slots, the first T - T mod S tiles march their
// whole column and the last T mod S are split into kz z-ranges (at least 8
// planes each, kz about S / (T mod S)), so that the persistent grid's
// workgroups, taking items i, i + S, ..., all finish about together (when
// T < S, every tile is split: more ranges as fewer boxes remain).  One
// workgroup.
__global__ void __launch_bounds__(1024) k_sweep_list(const TimeState *ts, const BlockGeo *geo, int nb,
                                                     int ntile, int nz, int slots, int4 *items,
                                                     int *n_items) {
    __shared__ int part[1024];
    __shared__ int s_kz, s_full;
    const int t = threadIdx.x, nt = blockDim.x;
    const int per = (nb + nt - 1) / nt, b0 = t * per, b1 = b0 + per < nb ? b0 + per : nb;
    int cnt = 0;
    for (int b = b0; b < b1; ++b) cnt += ts[geo[b].grp].skip ? 0 : 1;
    part[t] = cnt;
    __syncthreads();
    if (t == 0) {
        int act = 0;
        const int used = (nb + per - 1) / per;   // threads that own blocks
        for (int i = 0; i < used; ++i) {         // exclusive prefix of the active counts
            const int c = part[i];
            part[i] = act;
            act += c;
        }
        const int tiles = act * ntile, rem = tiles % slots;
        const int kz_max = nz / 8 > 1 ? nz / 8 : 1;
        int kz = rem ? (slots + rem / 2) / rem : 1;
        kz = kz < 1 ? 1 : (kz > kz_max ? kz_max : kz);
        s_kz = kz;
        s_full = tiles - rem;
        *n_items = s_full + rem * kz;
    }
    __syncthreads();
    const int kz = s_kz, full = s_full;
    int r = part[t];
    for (int b = b0; b < b1; ++b) {
        if (ts[geo[b].grp].skip) continue;
        for (int tl = 0; tl < ntile; ++tl) {
            const int k0 = r * ntile + tl;         // the tile's rank among the running tiles
            if (k0 < full) {
                items[k0] = make_int4(b, tl, 0, nz);
            } else {
                for (int k = 0; k < kz; ++k)
                    items[full + (k0 - full) * kz + k] = make_int4(b, tl, k * nz / kz, (k + 1) * nz / kz);
            }
        }
        ++r;
    }
}

}  // namespace kr
