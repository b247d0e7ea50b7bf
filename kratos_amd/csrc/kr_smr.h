// kr_smr.h -- the cross-level kernels of a statically refined mesh: the
// halo fill where levels meet, the flux correction of faces whose
// neighbour is finer, and the edge-EMF restriction with write-back.
//
// Semantics (the reference's act_bnd / act_flx / intg_flx where d_lvl != 0,
// hydro/boundary/comm_functions.h:329-544, mhd/boundary/comm_functions.h
// :467-954, hydro/integrator/integrator.h:313-382):
//   * a halo cell covered by a finer leaf holds the volume-weighted mean of
//     the 2x2x2 finer cells; one covered by a coarser leaf holds the coarser
//     cell plus minmod-limited slopes (dq = minmod(q_r - q_c, q_c - q_l)/4
//     per axis, signed by the half of the coarser cell it sits in).  Both
//     sum in x, y, z order, so every block sees the same value at a given
//     position; restricted halos are filled first and the prolongation's
//     slopes read them (act_bnd's q_ghst stage before q_plgt);
//   * a face whose neighbour is finer takes the area-weighted mean of the
//     four finer face fluxes (k_smr_restrict); the regular update skips it
//     and k_smr_update adds it to the adjacent cell (z sides, then y, then
//     x) before the cell is stored;
//   * an edge of a coarser block that a finer leaf shares takes the
//     length-weighted mean of the two finer edges along it
//     (k_smr_restrict); the finer blocks' edge pairs on such lines take
//     that mean too, which k_smr_update applies where it reads them.
#pragma once

#include "kr_staged.h"

namespace kr {

constexpr int kSmrThreads = 256;

template <class FT>
__device__ __forceinline__ FT minmod_(FT a, FT b) {   // utilities/functions.h:159-170
    return ((a < 0) ^ (b < 0)) ? (FT)0 : (kfabs(a) < kfabs(b) ? a : b);
}

// Wrap a global cell index of level L along the periodic axes.
__device__ __forceinline__ void smr_wrap(const SmrView &s, const int *n, int L, int *G) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int N = (s.R[a] * n[a]) << L;
        if (s.per[a] && (G[a] < 0 || G[a] >= N)) G[a] = ((G[a] % N) + N) % N;
    }
}

// One task per grid.y: fill the halo block's cells of the owner's
// neighbour region in direction `dir`.  prolong = 0: finer source
// (restriction), 1: coarser source (prolongation).
template <class FT, int NV>
__global__ void __launch_bounds__(kSmrThreads) k_smr_fill(MeshView<FT> m, SmrView s,
                                                         const int4 *task, int prolong) {
    const int4 tk = task[blockIdx.y];
    const int hb = tk.x, b = tk.y, dir = tk.z;
    const BlockGeo &g = m.geo[b];
    if (m.ts[g.grp].skip) return;
    int d[3] = {dir % 3 - 1, (dir / 3) % 3 - 1, dir / 9 - 1};
    int ext[3], lo[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ext[a] = d[a] == 0 ? m.n[a] : 2;
        lo[a] = d[a] < 0 ? m.n[a] - 2 : 0;          // halo-block local origin
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ext[0] * ext[1] * ext[2]) return;
    int hl[3] = {lo[0] + t % ext[0], lo[1] + (t / ext[0]) % ext[1], lo[2] + t / (ext[0] * ext[1])};
    int G[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) G[a] = g.reg[a] * m.n[a] + hl[a] + d[a] * m.n[a];
    smr_wrap(s, m.n, g.level, G);
    FT out[NV];
    if (!prolong) {
        int r[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) r[a] = (2 * G[a]) / m.n[a];
        const int f = smr_leaf(s, g.level + 1, r);
        if (f < 0) return;                    // excluded by the host's topology check
        const BlockGeo &gf = m.geo[f];
        int l[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) l[a] = 2 * G[a] - gf.reg[a] * m.n[a];
        const FT dv = (FT)gf.vol;
        FT acc[NV], norm = 0;
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0;
        const FT *W = m.W + (size_t)f * NV * m.ncell;
#pragma unroll
        for (int k0 = 0; k0 < 2; ++k0)
#pragma unroll
            for (int d1 = 0; d1 < 2; ++d1)
#pragma unroll
                for (int d2 = 0; d2 < 2; ++d2) {
                    const size_t c = cidx(m.n, l[0] + k0, l[1] + d1, l[2] + d2);
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[v] += W[(size_t)v * m.ncell + c] * dv;
                    norm += dv;
                }
#pragma unroll
        for (int v = 0; v < NV; ++v) out[v] = acc[v] / norm;
    } else {
        int r[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) r[a] = (G[a] >> 1) / m.n[a];
        const int cb = smr_leaf(s, g.level - 1, r);
        if (cb < 0) return;
        const BlockGeo &gc = m.geo[cb];
        int l[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) l[a] = (G[a] >> 1) - gc.reg[a] * m.n[a];
        FT qc[NV], dq[3][NV];
        load_cell<FT, NV>(m, cb, l, qc);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            FT qr[NV], ql[NV];
            int p[3] = {l[0], l[1], l[2]};
            ++p[a];
            load_cell<FT, NV>(m, cb, p, qr);
            p[a] -= 2;
            load_cell<FT, NV>(m, cb, p, ql);
#pragma unroll
            for (int v = 0; v < NV; ++v) dq[a][v] = minmod_<FT>(qr[v] - qc[v], qc[v] - ql[v]) * (FT)0.25f;
        }
        int sg[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) sg[a] = 2 * (G[a] & 1) - 1;
#pragma unroll
        for (int v = 0; v < NV; ++v)
            out[v] = qc[v] + dq[0][v] * (FT)sg[0] + dq[1][v] * (FT)sg[1] + dq[2][v] * (FT)sg[2];
    }
    FT *Wh = m.W + (size_t)hb * NV * m.ncell + cidx(m.n, hl[0], hl[1], hl[2]);
#pragma unroll
    for (int v = 0; v < NV; ++v) Wh[(size_t)v * m.ncell] = out[v];
}

// Flux restriction onto a face whose neighbour is finer.  Task {block,
// 2*axis + side}; thread t takes coarse face t of the task's side.
template <class FT, int NV>
__device__ __forceinline__ void smr_flux_at(const MeshView<FT> &m, const SmrView &s, int2 tk,
                                            int t) {
    const int b = tk.x, a = tk.y >> 1, sd = tk.y & 1;
    const BlockGeo &g = m.geo[b];
    if (m.ts[g.grp].skip) return;
    const int a1 = mod3(a + 1), a2 = mod3(a + 2);
    if (t >= m.n[a1] * m.n[a2]) return;
    const int j1 = t % m.n[a1], j2 = t / m.n[a1];
    int G[3];
    G[a] = g.reg[a] * m.n[a] + (sd ? m.n[a] : -1);
    G[a1] = g.reg[a1] * m.n[a1] + j1;
    G[a2] = g.reg[a2] * m.n[a2] + j2;
    smr_wrap(s, m.n, g.level, G);
    int r[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) r[x] = (2 * G[x]) / m.n[x];
    const int f = smr_leaf(s, g.level + 1, r);
    if (f < 0) return;
    const BlockGeo &gf = m.geo[f];
    int l[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) l[x] = 2 * G[x] - gf.reg[x] * m.n[x];
    l[a] = sd ? 0 : m.n[a];
    // the hydro slots only: nothing reads the EMF slots of F after the EMF
    // stage (and the fused kernel's producer form does not write them)
    constexpr int NH = 5;
    const FT ds = (FT)gf.sfc[a];
    FT fb[NH], norm = 0;
#pragma unroll
    for (int v = 0; v < NH; ++v) fb[v] = 0;
    const FT *Ff = m.F[a] + (size_t)f * NV * m.nfa[a];
#pragma unroll
    for (int d1 = 0; d1 < 2; ++d1)
#pragma unroll
        for (int d2 = 0; d2 < 2; ++d2) {
            int p[3] = {l[0], l[1], l[2]};
            p[a1] += d1;
            p[a2] += d2;
            const size_t x = xidx(m.n, a, p);
#pragma unroll
            for (int v = 0; v < NH; ++v) fb[v] += Ff[(size_t)v * m.nfa[a] + x] * ds;
            norm += ds;
        }
    int q[3];
    q[a] = sd ? m.n[a] : 0;
    q[a1] = j1;
    q[a2] = j2;
    FT *Fc = m.F[a] + (size_t)b * NV * m.nfa[a] + xidx(m.n, a, q);
#pragma unroll
    for (int v = 0; v < NH; ++v) Fc[(size_t)v * m.nfa[a]] = fb[v] / norm;
}

// Levels of the four cells around edge (component a) at local position i:
// bit 0 set if a finer leaf shares it (first such cell in fo1/fo2), bit 1 if
// a coarser one does.
template <class FT>
__device__ __forceinline__ int edge_levels(const MeshView<FT> &m, const BlockGeo &g, int a,
                                           const int *i, int &fo1, int &fo2) {
    const int a1 = mod3(a + 1), a2 = mod3(a + 2);
    int has = 0;
    fo1 = fo2 = 0;
    for (int o2 = -1; o2 <= 0; ++o2)
        for (int o1 = -1; o1 <= 0; ++o1) {
            int cc[3] = {i[0], i[1], i[2]};
            cc[a1] += o1;
            cc[a2] += o2;
            int d[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) d[x] = cc[x] < 0 ? -1 : (cc[x] >= m.n[x] ? 1 : 0);
            const int lv = (d[0] | d[1] | d[2]) ? g.lev[((d[2] + 1) * 3 + (d[1] + 1)) * 3 + (d[0] + 1)] : 0;
            if (lv == 1 && !(has & 1)) {
                has |= 1;
                fo1 = o1;
                fo2 = o2;
            }
            if (lv == -1) has |= 2;
        }
    return has;
}

template <class FT>
__device__ __forceinline__ bool on_block_surface(const MeshView<FT> &m, const int *i) {
    bool s = false;
#pragma unroll
    for (int x = 0; x < 3; ++x) s |= i[x] == 0 || i[x] == m.n[x];
    return s;
}

// Coarser side: edges shared with a finer leaf take the length-weighted
// mean of the two finer edges (mhd/boundary/comm_functions.h:795-864 and
// :919-954, the packed value).  Reads only finer blocks' edges, which
// nothing writes in the stage (the finer side's write-back is applied where
// the update reads them, k_smr_update), and writes only this block's.
template <class FT>
__device__ __forceinline__ void smr_emf_coarse_at(const MeshView<FT> &m, const SmrView &s, int b,
                                                  size_t t) {
    const BlockGeo &g = m.geo[b];
    if (m.ts[g.grp].skip) return;
    if (t >= m.nface) return;
    const int i[3] = {(int)(t % (m.n[0] + 1)), (int)((t / (m.n[0] + 1)) % (m.n[1] + 1)),
                      (int)(t / ((size_t)(m.n[0] + 1) * (m.n[1] + 1)))};
    if (!on_block_surface(m, i)) return;
    for (int a = 0; a < 3; ++a) {
        if (i[a] >= m.n[a]) continue;
        int o1, o2;
        if (!(edge_levels(m, g, a, i, o1, o2) & 1)) continue;
        const int a1 = mod3(a + 1), a2 = mod3(a + 2);
        int G[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) G[x] = g.reg[x] * m.n[x] + i[x];
        G[a1] += o1;
        G[a2] += o2;
        smr_wrap(s, m.n, g.level, G);
        int r[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) r[x] = (2 * G[x]) / m.n[x];
        const int f = smr_leaf(s, g.level + 1, r);
        if (f < 0) continue;
        const BlockGeo &gf = m.geo[f];
        int e[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) e[x] = 2 * G[x] - gf.reg[x] * m.n[x];
        e[a1] += o1 < 0 ? 2 : 0;
        e[a2] += o2 < 0 ? 2 : 0;
        const FT *EEf = m.EE + ((size_t)f * 3 + a) * m.nface;
        const FT dl = (FT)gf.dx[a];
        const FT el = EEf[fidx(m.n, e[0], e[1], e[2])] * dl;
        ++e[a];
        const FT er = EEf[fidx(m.n, e[0], e[1], e[2])] * dl;
        m.EE[((size_t)b * 3 + a) * m.nface + t] = (el + er) / (dl + dl);
    }
}

// Finer side: an edge pair on a coarser edge line shared with a coarser
// leaf takes its length-weighted mean (the write-back of comm_functions.h
// :855-861, 950-952).  Evaluated where the update reads the edge: the value
// of edge (component c, edge-box index idx, coordinates j) of block b.
// Pairs start at even indices along c and lie on even transverse lines;
// whether the line touches a coarser leaf depends only on the surface
// regions of its two transverse coordinates (BlockGeo::wbmask).
template <class FT>
__device__ __forceinline__ FT smr_edge_read(const MeshView<FT> &m, const BlockGeo &g, const FT *EE,
                                            int c, size_t idx, const int *j) {
    const FT *E = EE + (size_t)c * m.nface;
    const int c1 = mod3(c + 1), c2 = mod3(c + 2);
    if (((j[c1] | j[c2]) & 1) || j[c] >= m.n[c] + (j[c] & 1)) return E[idx];
    const int r1 = j[c1] == 0 ? 0 : (j[c1] == m.n[c1] ? 2 : 1);
    const int r2 = j[c2] == 0 ? 0 : (j[c2] == m.n[c2] ? 2 : 1);
    if (!((g.wbmask >> ((c * 3 + r1) * 3 + r2)) & 1)) return E[idx];
    const size_t sh = c == 0 ? 1 : (c == 1 ? (size_t)(m.n[0] + 1) : (size_t)(m.n[0] + 1) * (m.n[1] + 1));
    const size_t i0 = idx - (j[c] & 1) * sh;
    const FT dl = (FT)g.dx[c];
    return (E[i0] * dl + E[i0 + sh] * dl) / (dl + dl);
}

// act_flx of a stage in one launch: rows below nx restrict the EMFs of the
// coarser blocks' shared edges (block xblk[row], one thread per edge-box
// point), the others restrict the fluxes of the faces whose neighbour is
// finer, tpr tasks per row (ftask[(row - nx) * tpr + k], fmax threads per
// task).  Both read only what the sweep wrote and write disjoint coarser
// data.
template <class FT, int NV>
__global__ void __launch_bounds__(kSmrThreads) k_smr_restrict(MeshView<FT> m, SmrView s,
                                                             const int *xblk, int nx,
                                                             const int2 *ftask, int ntask,
                                                             int fmax, int tpr) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ((int)blockIdx.y < nx) {
        if constexpr (NV == 8) smr_emf_coarse_at(m, s, xblk[blockIdx.y], t);
    } else {
        const int k = (int)(t / fmax);
        const int task = ((int)blockIdx.y - nx) * tpr + k;
        if (k >= tpr || task >= ntask) return;
        smr_flux_at<FT, NV>(m, s, ftask[task], (int)(t - (size_t)k * fmax));
    }
}

// The update of the blocks of a refined mesh that are not updated by the
// fused kernel (k_update's arithmetic), with the cross-level corrections
// of intg_flx folded in: the edges of a finer block on a coarser edge line
// read as their pair's mean (smr_edge_read), and a cell next to a face
// whose neighbour is finer adds dt A / V times the restricted flux after
// the regular update, z sides first, then y, then x
// (hydro/integrator/integrator.h:313-382), before the state is stored.
template <class FT, bool MHD>
__global__ void __launch_bounds__(kThreads) k_smr_update(MeshView<FT> m, int step) {
    if (m.ts[m.geo[row_block(m)].grp].skip) return;
    const int b = row_block(m);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (MHD ? m.nface : m.ncell)) return;
    const BlockGeo &g = m.geo[b];
    update_point<FT, MHD>(
        m, b, t, step,
        [&](const FT *EE, int c, size_t idx, const int *i, int sa) -> FT {
            int j[3] = {i[0], i[1], i[2]};
            if (sa >= 0) ++j[sa];
            return smr_edge_read(m, g, EE, c, idx, j);
        },
        [&](const int *i, double *uu) {
            if (!g.rface) return;
            const double dt = m.ts[g.grp].dt_use * (step == 0 ? 1.0 : 0.5);
            for (int a = 2; a >= 0; --a)
                for (int sd = 0; sd <= 1; ++sd) {
                    if (!((g.rface >> (2 * a + sd)) & 1)) continue;
                    if (i[a] != (sd ? m.n[a] - 1 : 0)) continue;
                    int q[3] = {i[0], i[1], i[2]};
                    q[a] += sd;
                    const FT *F = m.F[a] + (size_t)b * m.nv * m.nfa[a] + xidx(m.n, a, q);
                    const FT ds = (FT)g.sfc[a];
                    const FT dv = (FT)g.vol * (FT)(sd ? -1 : 1);
                    const double conv = dt * (double)ds / (double)dv;
#pragma unroll
                    for (int v = 0; v < 5; ++v) uu[v] += conv * (double)F[(size_t)v * m.nfa[a]];
                }
        });
}

}  // namespace kr
