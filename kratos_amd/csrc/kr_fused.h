// kr_fused.h -- the fused MHD stage kernel: fluxes along x, y and z, edge
// EMFs, flux-divergence update and constrained transport in one launch.
//
// One workgroup owns a TX x TY column tile of one block and marches through
// the block along z.  Primitives W (written by k_c2p) are staged plane by
// plane into a 3-plane LDS ring over the tile plus a 2-cell halo; HBM holds
// no fluxes and no EMFs.  Per plane iteration L (leading plane):
//   1. W(L) enters the ring (prefetched during the previous iteration);
//   2. fluxes: in-plane x/y faces of plane L-1 and z faces at L-1 (one
//      lane per column keeps its z pencil state in registers);
//   3. EMFs: ee_z on plane L-1, ee_x / ee_y on z-face L-1;
//   4. update: Heun slot of the cells of plane L-2 (x, y parts were
//      accumulated one iteration earlier, z part now) and CT of the faces of
//      plane L-2 and z-face L-1.
// Arithmetic is the same per-face restatement as the staged kernels
// (kr_physics.h), in the same order, so both paths agree bit for bit in the
// FP64 form.  PLM is evaluated per cell: the reference's two face states of
// a cell share one limited increment t, q_l(face i+1) = q_i + t and
// q_r(face i) = q_i + (-t) exactly (hydro/reconstruct/plm.h:116-153).
#pragma once

#include "kr_staged.h"

namespace kr {

template <int TX, int TY>
struct TileShape {
    static constexpr int HX = TX + 4, HY = TY + 4, NCOL = HX * HY;    // halo columns
    static constexpr int NFX = (TX + 1) * (TY + 2);                   // x faces, rows -1..TY
    static constexpr int NFY = (TX + 2) * (TY + 1);                   // y faces, cols -1..TX
    static constexpr int NF = NFX + NFY;
    static constexpr int RX = TX + 2, NZG = (TX + 2) * (TY + 2);      // z-face ring grid
    static constexpr int NZI = TX * TY + 2 * (TX + 2) + 2 * TY;       // ring items
    static constexpr int NZIP = (NZI + 63) / 64 * 64;                 // ... padded to waves
    static constexpr int NEZ = (TX + 1) * (TY + 1), NEX = TX * (TY + 1), NEY = (TX + 1) * TY;
    __device__ static int col(int x, int y) { return (y + 2) * HX + (x + 2); }
    __device__ static int zg(int x, int y) { return (y + 1) * RX + (x + 1); }
    __device__ static int fxi(int x, int y) { return (y + 1) * (TX + 1) + x; }
    __device__ static int fyi(int x, int y) { return NFX + y * (TX + 2) + (x + 1); }
    __device__ static int ezi(int x, int y) { return y * (TX + 1) + x; }
    __device__ static int exi(int x, int y) { return y * TX + x; }
    __device__ static int eyi(int x, int y) { return y * (TX + 1) + x; }
    // ring item r -> column (interior first, then rows -1 / TY, then cols -1 / TX)
    __device__ static void ring_item(int r, int &x, int &y) {
        if (r < TX * TY) {
            x = r % TX;
            y = r / TX;
            return;
        }
        r -= TX * TY;
        if (r < 2 * (TX + 2)) {
            x = r % (TX + 2) - 1;
            y = r < TX + 2 ? -1 : TY;
            return;
        }
        r -= 2 * (TX + 2);
        x = r < TY ? -1 : TX;
        y = r % TY;
    }
};

__device__ __forceinline__ int ring3(int p) { return ((p % 3) + 3) % 3; }
__device__ __forceinline__ int ring2(int p) { return (p + 4) & 1; }

// Per-cell PLM increment of one variable (see header): returns t with the
// reference's expression order; `flat` is the reference's dq2 <= 0 branch.
template <class FT>
__device__ __forceinline__ FT plm_inc(FT qm, FT q0, FT qp, bool &flat) {
    const FT dql = q0 - qm;
    const FT dqc = qp - q0;
    FT norm(1);
    if constexpr (std::is_same_v<FT, float>) {
        // FP32: the slope ratio is scale free; rescale by the larger
        // difference so the squares stay in range (plm.h:128-137 in spirit)
        const FT s = pmax(kfabs(dql), kfabs(dqc));
        if (s > (FT)1e-24f) norm = 1 / s;
    }
    const FT dql2 = norm * dql * dql;
    const FT dqc2 = norm * dqc * dqc;
    const FT dq2 = norm * dql * dqc;
    flat = dq2 <= 0;
    const FT coef = flat ? (FT)0 : dq2 / (dql2 + dqc2 + dq2 * 2);
    return coef * (dql + dqc);
}

// a * b - c * d in FP32 with multiply-add contraction (mixed form)
__device__ __forceinline__ float cross_f32(float a, float b, float c, float d) {
#pragma clang fp contract(fast)
    return a * b - c * d;
}

// Edge EMF from the four surrounding cell E values and the two pairs of
// face EMFs / upwind mass fluxes (mhd/integrator/integrator.h:244-317).
// ec[s2][s1] naming of the reference: ec11 = E(i), ec10 = E(i - e1),
// ec00 = E(i - e1 - e2), ec01 = E(i - e2).
template <class FT>
__device__ __forceinline__ FT edge_emf(FT ec11, FT ec10, FT ec00, FT ec01, FT efm1, FT efp1,
                                       FT rp1, FT rm1, FT efm2, FT efp2, FT rp2, FT rm2);

// FP32 (mixed form): the same expression with multiply-add contraction.
template <>
__device__ __forceinline__ float edge_emf<float>(float ec11, float ec10, float ec00, float ec01,
                                                 float efm1, float efp1, float rp1, float rm1,
                                                 float efm2, float efp2, float rp2, float rm2) {
#pragma clang fp contract(fast)
    const float w_e = 0.25f;
    float ee = 0;
    ee += w_e * (efm1 + efp1);
    {
        const float de_p = (rp1 > 0 ? (ec01 - efm1) : (ec11 - efp1));
        const float de_m = (rm1 > 0 ? -(ec00 - efm1) : -(ec10 - efp1));
        ee += w_e * (de_m - de_p);
    }
    ee += w_e * (efm2 + efp2);
    {
        const float de_p = (rp2 > 0 ? (ec10 - efm2) : (ec11 - efp2));
        const float de_m = (rm2 > 0 ? -(ec00 - efm2) : -(ec01 - efp2));
        ee += w_e * (de_m - de_p);
    }
    return ee;
}

template <>
__device__ __forceinline__ double edge_emf<double>(double ec11, double ec10, double ec00,
                                                   double ec01, double efm1, double efp1,
                                                   double rp1, double rm1, double efm2,
                                                   double efp2, double rp2, double rm2) {
    using FT = double;
    const FT w_e = (FT)0.25f;
    FT ee = 0;
    ee += w_e * (efm1 + efp1);
    {
        const FT de_p = (rp1 > 0 ? (ec01 - efm1) : (ec11 - efp1));
        const FT de_m = (rm1 > 0 ? -(ec00 - efm1) : -(ec10 - efp1));
        ee += w_e * (de_m - de_p);
    }
    ee += w_e * (efm2 + efp2);
    {
        const FT de_p = (rp2 > 0 ? (ec10 - efm2) : (ec11 - efp2));
        const FT de_m = (rm2 > 0 ? -(ec00 - efm2) : -(ec01 - efp2));
        ee += w_e * (de_m - de_p);
    }
    return ee;
}

// WAVES > 1 asks the compiler for at least that many waves per SIMD (fewer
// VGPRs, possibly spills); the tile autotuner times both forms.
//
// PHY: the mesh has physical sides, and the blocks next to them apply the
// boundary rules of the face field and of the boundary-plane EMF (the ghost
// primitives come from halo blocks, as a neighbour's would).  Periodic
// meshes take the build without those branches, which keeps their
// registers and code out of the common path.
//
// PROD (blocks of a refined mesh that touch another level): the kernel
// only produces what the staged update needs -- the hydro fluxes of the
// block's own faces into F and the edge EMFs of its own edges into EE, in
// the staged kernels' layout and with the same values -- and stops before
// the update, so that the cross-level corrections (restricted fluxes and
// EMFs) can act on them before k_update.  Where a transverse ghost line
// runs next to a face shared with another level, the normal field is the
// cell value of W there (mhd/integrator/data.h:108-140), as in load_bf.
//
// PROD = 2 (refined meshes): one launch for both kinds of block, the grid
// rows below m.nfull updating (the full form), the rows from m.nfull on
// producing (the producer form); a workgroup takes its row's form.
template <class FT, int TX, int TY, int NT, int WAVES, int PROD = 0, bool PHY = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_sweep(MeshView<FT> m, int step) {
    using S = TileShape<TX, TY>;
    static_assert(S::NCOL <= NT, "one halo column per thread");
    static_assert(S::NZIP <= NT, "one z pencil per thread, in-plane items from a wave boundary");
    static_assert(TX * TY + TX + TY <= NT, "one owned cell / edge face per thread");
    if (m.ts[m.geo[row_block(m)].grp].skip) return;
    const int tid = threadIdx.x;
    const int b = row_block(m);
    const bool prod = PROD == 1 || (PROD == 2 && (int)blockIdx.y >= m.nfull);
    const int ntx = m.n[0] / TX;
    const int x0 = (blockIdx.x % ntx) * TX, y0 = (blockIdx.x / ntx) * TY;
    const int nz = m.n[2];
    const bool xlast = x0 + TX == m.n[0], ylast = y0 + TY == m.n[1];
    const BlockGeo &g = m.geo[b];
    const double dt = m.ts[g.grp].dt_use * (step == 0 ? 1.0 : 0.5);
    const FT gam = (FT)m.gamma;
    const size_t nc = m.ncell;

    // W ring: the planes are brought in by global_load_lds in 64-lane dword
    // chunks (lane i of a chunk writes dword i of the chunk); lanes past the
    // end of a row issue no load, so nothing is written outside the row
    constexpr int DPC = (int)sizeof(FT) / 4;                       // dwords per column
    constexpr int NDW = S::NCOL * DPC;                             // dwords per row
    constexpr int NCHUNK = (NDW + 63) / 64;                        // chunks per row
    __shared__ FT sW[3][8][S::NCOL];
    // in-plane faces: mass flux and the EMF slot the next plane's edges
    // need (E_y of x faces, E_x of y faces) for two planes; E_z and the
    // energy / momentum fluxes of the current plane only
    __shared__ FT sFe[2][2][S::NF];
    __shared__ FT sFez[S::NF];
    __shared__ FT sFh[4][S::NF];
    __shared__ FT sFz[3][S::NZG];       // rho, E_x, E_y of the current z face
    __shared__ FT sEz[2][S::NEZ];
    __shared__ FT sEx[2][S::NEX];
    __shared__ FT sEy[2][S::NEY];

    // The block's 27 neighbour entries, read per lane from LDS rather than
    // from the block table in HBM (they resolve the halo rows and columns
    // of every plane).
    __shared__ int sNbr[27];
    if (tid < 27) sNbr[tid] = g.wnbr[tid];
    // Per-block factors of the update and the CT, formed once per launch
    // (the same expressions as at each use, so the same values): CT factor
    // -dt / sfc_a, flux factor sfc_a * dt / vol, edge length dx_a.  Kept in
    // LDS, not registers, and read where used by the FP32 flux form; the
    // FP64 form forms them at each use (its code is faster that way).
    __shared__ double sGeo[9];
    if (tid < 3) {
        const double sfc = g.sfc[tid];
        sGeo[tid] = sfc > 0 ? (double)(-dt / sfc) : (double)(0);
        sGeo[3 + tid] = sfc * (dt / g.vol);
        sGeo[6 + tid] = g.dx[tid];
    }
    __syncthreads();

    // ---- primitive source of halo column (cx, cy) at plane z
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
    // normal face field of face (fx, fy, fz), component c: transverse
    // indices may leave the block (same-level neighbour), the component's
    // own index stays in 0 .. n.  `slot` is the ring slot of plane fz.
    auto bfn = [&](int c, int fx, int fy, int fz, int slot) -> FT {
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
        if (PHY && out && g.phy) {
            // physical side (mhd/boundary/phys.h:178-229, phys_functors.h
            // :10-137), as load_bf: the tangential field of the ghost cell
            // next to the side -- free / outflow the boundary cell's, reflect
            // minus the one at -i (two cells in), fixd the fixed field; along
            // the component's own axis the row is clamped into the block where
            // that axis is physical at the top
            int a = 0;
#pragma unroll
            for (int x = 0; x < 3; ++x)
                if (d[x]) a = x;
            if ((g.phy >> (2 * a + (d[a] > 0))) & 1) {
                const int side = d[a] > 0;
                const int kind = m.bnd[2 * a + side];
                if (kind == 4) return (FT)m.fix_b[c];
                double sgn = 1;
                if (kind == 3) {
                    p[a] = side ? m.n[a] - 2 : 1;
                    sgn = -1;
                } else {
                    p[a] = side ? m.n[a] - 1 : 0;
                }
                if (((g.phy >> (2 * c + 1)) & 1) && p[c] >= m.n[c]) p[c] = m.n[c] - 1;
                return (FT)(sgn * m.BF[(((size_t)b * 2 + step) * 3 + c) * m.nface +
                                        fidx(m.n, p[0], p[1], p[2])]);
            }
        }
        if (prod) {
            if (out && g.xface) {
                int bit = 0;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    if (d[a]) bit = 1 << (2 * a + (d[a] > 0));
                if (g.xface & bit) return sW[slot][BCC + c][S::col(fx, fy)];
            }
        }
        if (out) {
            blk = sNbr[((d[2] + 1) * 3 + (d[1] + 1)) * 3 + (d[0] + 1)];
#pragma unroll
            for (int a = 0; a < 3; ++a) p[a] -= d[a] * m.n[a];
        }
        return (FT)m.BF[(((size_t)blk * 2 + step) * 3 + c) * m.nface + fidx(m.n, p[0], p[1], p[2])];
    };

    // Asynchronous load of plane z of W into ring slot `slot` (no staging
    // registers).  Wave w fills chunks w, w + NT/64, ... of every variable
    // row; lane l of chunk k carries dword k*64 + l of the row, i.e. dword
    // `half` of column col.
    auto issue_w = [&](int z, int slot) {
        const int wave = tid >> 6, lane = tid & 63;
        for (int k = wave; k < NCHUNK; k += NT / 64) {
            const int d = k * 64 + lane;
            if (d >= NDW) continue;                 // past the row: no load, no write
            const int col = d / DPC;
            const int half = d - col * DPC;
            const FT *src = wsrc(col % S::HX - 2, col / S::HX - 2, z);
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const char *gp = (const char *)(src + (size_t)v * nc) + 4 * half;
                __builtin_amdgcn_global_load_lds(gp, &sW[slot][v][0] + k * (64 / DPC), 4, 0, 0);
            }
        }
    };

    // ---- per-thread persistent state
    // z pencil owner (tid < NZI): column (zx, zy)
    int zx = 0, zy = 0;
    const bool zown = tid < S::NZI;
    if (zown) S::ring_item(tid, zx, zy);
    const bool zcorner = (zx < 0 || zx >= TX) && (zy < 0 || zy >= TY);
    // owned cell: only pencil threads own one (the others keep zx = zy = 0)
    const bool zint = zown && zx >= 0 && zx < TX && zy >= 0 && zy < TY;
    FT qL[8];                 // left state of the next z face (mesh frame)
    FT fzp[5];                // hydro z flux of the previous face (owned cells)
    double part[5];           // x + y partial update of the previous plane
#pragma unroll
    for (int v = 0; v < 8; ++v) qL[v] = 0;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        fzp[v] = 0;
        part[v] = 0;
    }
    const int cx_ = zx, cy_ = zy;                 // owned cell (when zint)
    // edge faces of the tile's high sides: threads TX*TY .. +TY-1 (x faces at
    // x = TX) and +TY .. +TY+TX-1 (y faces at y = TY)
    const int et = tid - TX * TY;
    const bool xedge = xlast && et >= 0 && et < TY;
    const bool yedge = ylast && et >= TY && et < TY + TX;

    constexpr bool F32 = std::is_same_v<FT, float>;
    const double dt_dv = dt / g.vol;


    issue_w(-2, ring3(-2));
    for (int L = -2; L <= nz + 1; ++L) {
        // ------------------------------------------------ 1. plane L arrives
        __builtin_amdgcn_s_waitcnt(0);   // this wave's loads into the ring
        __syncthreads();                 // ... and every other wave's

        const int P = L - 2;          // plane finished in this iteration

        // ------------------------------------------------ 2. fluxes
        if (L >= 0) {
            const int p = L - 1;                 // plane of the in-plane faces / z face index
            const int s0 = ring3(p), sm = ring3(p - 1), sp = ring3(p + 1);
            const int fe = ring2(p);
            // z pencil: increment of cell p, face states, flux at face p
            if (zown && !zcorner) {
                const int c = S::col(zx, zy);
                FT qR[8];
                FT qLn[8];
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    if (v == BCC + 2) {
                        qR[v] = 0;
                        qLn[v] = 0;
                        continue;
                    }
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
                    FT wl[8], wr[8], flx[8];
#pragma unroll
                    for (int v = 0; v < 8; ++v) {
                        wl[v] = qL[v];
                        wr[v] = qR[v];
                    }
                    const FT bn = bfn(2, zx, zy, p, s0);
                    wl[BCC + 2] = bn;
                    wr[BCC + 2] = bn;
                    rot<2>(wl + MOM);
                    rot<2>(wr + MOM);
                    rot<2>(wl + BCC);
                    rot<2>(wr + BCC);
                    if constexpr (std::is_same_v<FT, float>)
                        hlld_fast(flx, wl, wr, gam);
                    else
                        hlld<FT>(flx, wl, wr, gam);
                    rot<1>(flx + MOM);
                    // EMF rewrite for axis 2: E_x = flx[BCC+2], E_y = -flx[BCC+1]
                    const int gz = S::zg(zx, zy);
                    sFz[0][gz] = flx[RHO];
                    sFz[1][gz] = flx[BCC + 2];
                    sFz[2][gz] = -flx[BCC + 1];
                    if (prod && zint) {
                        // the block's z face (x, y, p): hydro flux for the staged update
                        const int q3[3] = {x0 + cx_, y0 + cy_, p};
                        FT *F = m.F[2] + (size_t)b * 8 * m.nfa[2] + xidx(m.n, 2, q3);
#pragma unroll
                        for (int v = 0; v < 5; ++v) F[(size_t)v * m.nfa[2]] = flx[v];
                    } else if (zint) {
                        // z part of the cell finished now (plane p-1 = P)
                        if (P >= 0) {
                            const FT fa = F32 ? (FT)sGeo[3 + 2] : (FT)(g.sfc[2] * dt_dv);
                            const size_t cc = cidx(m.n, x0 + cx_, y0 + cy_, P);
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
                }
#pragma unroll
                for (int v = 0; v < 8; ++v) qL[v] = qLn[v];
            }
            // in-plane faces of plane p.  The z pencils occupy threads
            // 0 .. NZI-1; the in-plane items start at the next wave boundary
            // NZIP, so that every wave runs the Riemann solver the same
            // number of times (thread t starts at item (t - NZIP) mod NT).
            for (int it = tid - S::NZIP + (tid < S::NZIP ? NT : 0); it < S::NF; it += NT) {
                if (it < 0) continue;
                const bool xf = it < S::NFX;
                int fx, fy;
                if (xf) {
                    fx = it % (TX + 1);
                    fy = it / (TX + 1) - 1;
                } else {
                    const int k = it - S::NFX;
                    fx = k % (TX + 2) - 1;
                    fy = k / (TX + 2);
                }
                const bool ghost_plane = p < 0 || p >= nz;
                if (ghost_plane && (xf ? (fy < 0 || fy >= TY) : (fx < 0 || fx >= TX))) continue;
                const int a = xf ? 0 : 1;
                const int stride = xf ? 1 : S::HX;
                const int cR = S::col(fx, fy);          // cell on the right of the face
                FT wl[8], wr[8], flx[8];
                // rotated frame: slot k holds mesh variable perm(k)
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    int v = k;
                    if (k >= MOM && k < MOM + 3) v = MOM + (a + k - MOM) % 3;
                    if (k >= BCC) v = BCC + (a + k - BCC) % 3;
                    if (k == BCC) continue;            // normal field: face value
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
                const FT bn = bfn(a, fx, fy, p, s0);
                wl[BCC] = bn;
                wr[BCC] = bn;
                if constexpr (std::is_same_v<FT, float>)
                    hlld_fast(flx, wl, wr, gam);
                else
                    hlld<FT>(flx, wl, wr, gam);
                // back to the mesh frame
                sFh[0][it] = flx[ENE];
#pragma unroll
                for (int k = 0; k < 3; ++k) sFh[1 + (a + k) % 3][it] = flx[MOM + k];
                if (prod) {
                    // a face of the block (not a transverse ghost line, not a
                    // ghost plane): hydro flux for the staged update
                    if (!ghost_plane && (xf ? (fy >= 0 && fy < TY) : (fx >= 0 && fx < TX))) {
                        const int q3[3] = {x0 + fx, y0 + fy, p};
                        FT *F = m.F[a] + (size_t)b * 8 * m.nfa[a] + xidx(m.n, a, q3);
                        F[(size_t)RHO * m.nfa[a]] = flx[RHO];
                        F[(size_t)ENE * m.nfa[a]] = flx[ENE];
#pragma unroll
                        for (int k = 0; k < 3; ++k)
                            F[(size_t)(MOM + (a + k) % 3) * m.nfa[a]] = flx[MOM + k];
                    }
                }
                sFe[fe][0][it] = flx[RHO];
                if (xf) {                        // x faces: E_y, E_z
                    sFe[fe][1][it] = flx[BCC + 2];
                    sFez[it] = -flx[BCC + 1];
                } else {                         // y faces: E_x, E_z
                    sFe[fe][1][it] = -flx[BCC + 1];
                    sFez[it] = flx[BCC + 2];
                }
            }
        }
        __syncthreads();

        // ------------------------------------------------ 3. EMFs
        if (L >= 1) {
            const int p = L - 1;
            const int s0 = ring3(p), sm = ring3(p - 1);
            const int fe = ring2(p), fm = ring2(p - 1);
            auto ec = [&](int s, int x, int y, int c) -> FT {
                const int cc = S::col(x, y);
                const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                const FT b1 = sW[s][BCC + c1][cc], v2 = sW[s][MOM + c2][cc];
                const FT b2 = sW[s][BCC + c2][cc], v1 = sW[s][MOM + c1][cc];
                if constexpr (std::is_same_v<FT, float>)
                    return cross_f32(b1, v2, b2, v1);
                else
                    return b1 * v2 - b2 * v1;
            };
            // boundary-plane EMF of a high out / fre side
            // (mhd/integrator/integrator.h:328-381), as k_emf: edge c at
            // tile position (x, y) and z index p takes the interior cell's
            // value, v = w[mom] / rho as written there, the last such side
            // (x, y, z order) winning
            auto bplane = [&](int c, int x, int y, FT e) -> FT {
                const int gi[3] = {x0 + x, y0 + y, p};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (a == c || gi[a] != m.n[a] || m.n[a] < 2 || !((g.phy >> (2 * a + 1)) & 1)) continue;
                    const int k = m.bnd[2 * a + 1];
                    if (k != 1 && k != 2) continue;
                    const int cx = x - (a == 0), cy = y - (a == 1);
                    const int sl = a == 2 ? sm : s0;
                    const int cc = S::col(cx, cy);
                    FT w[8];
#pragma unroll
                    for (int v = 0; v < 8; ++v) w[v] = sW[sl][v][cc];
                    const FT irho = (FT)1 / pmax(w[RHO], (FT)1e-30);
                    FT v_c[3], b_c[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        v_c[q] = w[MOM + q] * irho;
                        b_c[q] = w[BCC + q];
                    }
                    const int a1 = mod3(a + 1), a2 = mod3(a + 2);
                    if (c == a1) e = v_c[a] * b_c[a2] - v_c[a2] * b_c[a];
                    else e = v_c[a1] * b_c[a] - v_c[a] * b_c[a1];
                }
                return e;
            };
            for (int it = tid; it < S::NEZ + S::NEX + S::NEY; it += NT) {
                if (it < S::NEZ) {
                    if (p >= nz) continue;
                    const int x = it % (TX + 1), y = it / (TX + 1);
                    // ee_z on plane p: a1 = x, a2 = y
                    FT e = edge_emf<FT>(
                        ec(s0, x, y, 2), ec(s0, x - 1, y, 2), ec(s0, x - 1, y - 1, 2), ec(s0, x, y - 1, 2),
                        sFez[S::fxi(x, y - 1)], sFez[S::fxi(x, y)],
                        sFe[fe][0][S::fyi(x, y)], sFe[fe][0][S::fyi(x - 1, y)],
                        sFez[S::fyi(x - 1, y)], sFez[S::fyi(x, y)],
                        sFe[fe][0][S::fxi(x, y)], sFe[fe][0][S::fxi(x, y - 1)]);
                    if (PHY && g.phy) e = bplane(2, x, y, e);
                    sEz[fe][S::ezi(x, y)] = e;
                    if (prod)
                        m.EE[((size_t)b * 3 + 2) * m.nface + fidx(m.n, x0 + x, y0 + y, p)] = e;
                } else if (it < S::NEZ + S::NEX) {
                    const int k = it - S::NEZ;
                    const int x = k % TX, y = k / TX;
                    // ee_x on z face p: a1 = y, a2 = z
                    FT e = edge_emf<FT>(
                        ec(s0, x, y, 0), ec(s0, x, y - 1, 0), ec(sm, x, y - 1, 0), ec(sm, x, y, 0),
                        sFe[fm][1][S::fyi(x, y)], sFe[fe][1][S::fyi(x, y)],
                        sFz[0][S::zg(x, y)], sFz[0][S::zg(x, y - 1)],
                        sFz[1][S::zg(x, y - 1)], sFz[1][S::zg(x, y)],
                        sFe[fe][0][S::fyi(x, y)], sFe[fm][0][S::fyi(x, y)]);
                    if (PHY && g.phy) e = bplane(0, x, y, e);
                    sEx[fe][S::exi(x, y)] = e;
                    if (prod)
                        m.EE[((size_t)b * 3 + 0) * m.nface + fidx(m.n, x0 + x, y0 + y, p)] = e;
                } else {
                    const int k = it - S::NEZ - S::NEX;
                    const int x = k % (TX + 1), y = k / (TX + 1);
                    // ee_y on z face p: a1 = z, a2 = x
                    FT e = edge_emf<FT>(
                        ec(s0, x, y, 1), ec(sm, x, y, 1), ec(sm, x - 1, y, 1), ec(s0, x - 1, y, 1),
                        sFz[2][S::zg(x - 1, y)], sFz[2][S::zg(x, y)],
                        sFe[fe][0][S::fxi(x, y)], sFe[fm][0][S::fxi(x, y)],
                        sFe[fm][1][S::fxi(x, y)], sFe[fe][1][S::fxi(x, y)],
                        sFz[0][S::zg(x, y)], sFz[0][S::zg(x - 1, y)]);
                    if (PHY && g.phy) e = bplane(1, x, y, e);
                    sEy[fe][S::eyi(x, y)] = e;
                    if (prod)
                        m.EE[((size_t)b * 3 + 1) * m.nface + fidx(m.n, x0 + x, y0 + y, p)] = e;
                }
            }
        }
        __syncthreads();

        // plane L+1 into the slot of plane L-2, which phases 2 and 3 of this
        // iteration were the last to read; it lands while the update runs
        if (L + 1 <= nz + 1) issue_w(L + 1, ring3(L + 1));

        // ------------------------------------------------ 4. update
        if (!prod && L >= 1) {
            const int q = L - 1;                     // z face updated now
            const int fq = ring2(q), fpp = ring2(P);
            // Face-field update of one face (mhd/integrator/rk2.h:48-62): the
            // two slots of every face a thread updates are read first
            // (face_load), with the conserved state, and written last
            // (face_store), so that the loads are in flight together.
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
            auto set_face = [&](int c, int fx, int fy, int fz, double db) {
                face_store(face_load(c, fx, fy, fz), db);
            };
            // CT circulation of one face (mhd/integrator/integrator.h:429-457)
            auto ct = [&](double e1, double d1, double e2, double d2, double e3, double e4,
                          int a) {
                double db = e1 * d1;
                db -= e2 * d2;
                db += e3 * d2;
                db -= e4 * d1;
                if constexpr (F32) {
                    db *= sGeo[a];              // -dt / sfc, or 0 for a zero-area face
                } else {
                    const double sfc = g.sfc[a];
                    db *= (sfc > 0 ? (double)(-dt / sfc) : (double)(0));
                }
                return db;
            };
            const double dx = F32 ? sGeo[6] : g.dx[0], dy = F32 ? sGeo[7] : g.dx[1],
                         dz = F32 ? sGeo[8] : g.dx[2];
            if (zint) {
                const int x = cx_, y = cy_;
                const FaceUpd uz = face_load(2, x, y, q);
                const FaceUpd ux = face_load(0, x, y, P < 0 ? 0 : P);
                const FaceUpd uy = face_load(1, x, y, P < 0 ? 0 : P);
                // x + y partial of plane q (cells), finished next iteration
                if (q < nz) {
                    // conserved state of plane q, loaded here so that it is not
                    // live across the flux phase
                    const size_t c = cidx(m.n, x0 + x, y0 + y, q);
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
                        const int i0 = a == 0 ? S::fxi(x, y) : S::fyi(x, y);
                        const int i1 = a == 0 ? S::fxi(x + 1, y) : S::fyi(x, y + 1);
                        if constexpr (std::is_same_v<FT, double>) {
                            const double fa = g.sfc[a] * dt_dv;
#pragma unroll
                            for (int v = 0; v < 5; ++v) {
                                double res(0);
                                const FT f0 = v == 0 ? sFe[fq][0][i0] : sFh[v - 1][i0];
                                const FT f1 = v == 0 ? sFe[fq][0][i1] : sFh[v - 1][i1];
                                if (fa > 0) res += fa * f0;
                                if (fa > 0) res -= fa * f1;
                                uu[v] += res;
                            }
                        } else {
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
                    }
#pragma unroll
                    for (int v = 0; v < 5; ++v) part[v] = uu[v];
                }
                // z face q
                face_store(uz, ct(sEx[fq][S::exi(x, y)], dx, sEy[fq][S::eyi(x, y)], dy,
                                  sEy[fq][S::eyi(x + 1, y)], sEx[fq][S::exi(x, y + 1)], 2));
                if (P >= 0) {
                    // x face (x, y, P): a1 = y, a2 = z
                    face_store(ux, ct(sEy[fpp][S::eyi(x, y)], dy, sEz[fpp][S::ezi(x, y)], dz,
                                      sEz[fpp][S::ezi(x, y + 1)], sEy[fq][S::eyi(x, y)], 0));
                    // y face (x, y, P): a1 = z, a2 = x
                    face_store(uy, ct(sEz[fpp][S::ezi(x, y)], dz, sEx[fpp][S::exi(x, y)], dx,
                                      sEx[fq][S::exi(x, y)], sEz[fpp][S::ezi(x + 1, y)], 1));
                }
            }
            if (P >= 0 && xedge) {
                const int x = TX, y = et;
                set_face(0, x, y, P,
                         ct(sEy[fpp][S::eyi(x, y)], dy, sEz[fpp][S::ezi(x, y)], dz,
                            sEz[fpp][S::ezi(x, y + 1)], sEy[fq][S::eyi(x, y)], 0));
            }
            if (P >= 0 && yedge) {
                const int x = et - TY, y = TY;
                set_face(1, x, y, P,
                         ct(sEz[fpp][S::ezi(x, y)], dz, sEx[fpp][S::exi(x, y)], dx,
                            sEx[fq][S::exi(x, y)], sEz[fpp][S::ezi(x + 1, y)], 1));
            }
        }
    }
}

// Physical-side halo of the fused kernel.  One task {halo block, owner,
// direction, 0} per grid row: the halo block's cells at the positions a
// same-level neighbour in that direction would hold get the ghost
// primitives that load_cell gives there -- the boundary functors
// (hydro / mhd boundary/phys.h, phys_functors.h) the staged kernels apply
// on the fly.  Runs after c2p, before the fused kernel reads the halo.
// NV: 8 for MHD, 5 for hydro.
template <class FT, int NV>
__global__ void __launch_bounds__(256) k_phys_fill(MeshView<FT> m, const int4 *task) {
    const int4 tk = task[blockIdx.y];
    const int hb = tk.x, b = tk.y, dir = tk.z;
    if (m.ts[m.geo[b].grp].skip) return;
    const int d[3] = {dir % 3 - 1, (dir / 3) % 3 - 1, dir / 9 - 1};
    int ext[3], lo[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ext[a] = d[a] == 0 ? m.n[a] : 2;
        lo[a] = d[a] < 0 ? m.n[a] - 2 : 0;          // halo-block local origin
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ext[0] * ext[1] * ext[2]) return;
    const int hl[3] = {lo[0] + t % ext[0], lo[1] + (t / ext[0]) % ext[1], lo[2] + t / (ext[0] * ext[1])};
    int p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = hl[a] + d[a] * m.n[a];
    FT w[NV];
    load_cell<FT, NV>(m, b, p, w);
    FT *Wh = m.W + (size_t)hb * NV * m.ncell + cidx(m.n, hl[0], hl[1], hl[2]);
#pragma unroll
    for (int v = 0; v < NV; ++v) Wh[(size_t)v * m.ncell] = w[v];
}

}  // namespace kr
