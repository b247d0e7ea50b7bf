// kr_staged.h -- the staged Godunov kernels: one launch per stage step, each
// launch covering every block of the mesh (grid.y = block), SoA state in
// HBM, neighbour-table reads instead of ghost copies.
//
// Per RK substep: k_c2p -> (k_dt_finalize at substep 0) -> k_flux x ndim
// -> k_emf (MHD) -> k_update.  This is the reference's integrate() pipeline
// (dynamics/dynamics.h:60-100) without per-block launches, streams, host
// syncs or halo-copy kernels.
#pragma once

#include "kr_mesh.h"
#include "kr_physics.h"

namespace kr {

constexpr int kThreads = 256;

// ------------------------------------------------------------------ index
__device__ __forceinline__ size_t cidx(const int *n, int i, int j, int k) {
    return ((size_t)k * n[1] + j) * n[0] + i;
}
__device__ __forceinline__ size_t fidx(const int *n, int i, int j, int k) {
    return ((size_t)k * (n[1] + 1) + j) * (n[0] + 1) + i;
}
// flux box of sweep axis `ax`: n+1 faces along ax, n+2 (one ghost line on
// each side) along the transverse axes.
__device__ __forceinline__ size_t xidx(const int *n, int ax, const int *p) {
    const int d0 = n[0] + (ax == 0 ? 1 : 2), d1 = n[1] + (ax == 1 ? 1 : 2);
    const int o0 = ax == 0 ? 0 : 1, o1 = ax == 1 ? 0 : 1, o2 = ax == 2 ? 0 : 1;
    return ((size_t)(p[2] + o2) * d1 + (p[1] + o1)) * d0 + (p[0] + o0);
}

// Block handled by grid row blockIdx.y of a stage kernel.
template <class FT>
__device__ __forceinline__ int row_block(const MeshView<FT> &m) {
    return m.blist ? m.blist[blockIdx.y] : (int)blockIdx.y;
}

template <class FT>
__device__ __forceinline__ int nbr_dir(const MeshView<FT> &m, const int *d) {
    return ((d[2] + 1) * 3 + (d[1] + 1)) * 3 + (d[0] + 1);
}

// Load the primitive vector of cell p of block b.  Cells outside the block
// come from the same-level neighbour (the reference fills them by the
// ghost exchange, hydro/boundary/comm_functions.h:304-327); on a physical
// side the reference's boundary functor is applied on the fly
// (hydro/boundary/phys_functors.h:10-89).
template <class FT, int NV>
__device__ __forceinline__ void load_cell(const MeshView<FT> &m, int b, const int *pin, FT *w) {
    int p[3] = {pin[0], pin[1], pin[2]};
    int d[3];
    bool out = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        d[a] = p[a] < 0 ? -1 : (p[a] >= m.n[a] ? 1 : 0);
        out |= d[a] != 0;
    }
    int blk = b;
    int flip = -1;      // reflect: negate v_ax;  outflow: clamp inflow of v_ax
    int clamp = -1;     // 2*ax + side of an outflow side
    bool phys8 = false; // MHD cell behind a physical side
    bool fixd = false;  // Dirichlet side: the ghost is the fixed state
    if (out) {
        const int nb = m.geo[b].nbr[nbr_dir(m, d)];
        if (nb >= 0) {
            blk = nb;
#pragma unroll
            for (int a = 0; a < 3; ++a) p[a] -= d[a] * m.n[a];
        } else if constexpr (NV == 5) {
            // physical side: one axis out of range (hydro face ghosts);
            // every index below is resolved with unrolled selects so the
            // register arrays keep static indices.
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (d[a] == 0) continue;
                const int side = d[a] > 0;
                const int kind = m.bnd[2 * a + side];
                const int g = side ? p[a] - m.n[a] : -1 - p[a];
                if (kind == 3) {                       // reflect
                    p[a] = side ? m.n[a] - 1 - g : g;
                    flip = a;
                } else {                               // outflow / free / fixd
                    p[a] = side ? m.n[a] - 1 : 0;
                    if (kind == 1) clamp = 2 * a + side;
                    if (kind == 4) fixd = true;
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) p[a] = p[a] < 0 ? 0 : (p[a] >= m.n[a] ? m.n[a] - 1 : p[a]);
        } else {
            phys8 = true;
        }
    }
    int mflip = -1, mclamp = -1;       // MHD physical side: reflect axis / outflow side
    if constexpr (NV == 8) {
        if (phys8) {
            // MHD physical side (mhd/boundary/phys.h:144-251): the cell is
            // written last by the side of the highest out-of-range axis that
            // is physical; its source is mirrored (reflect) or the boundary
            // cell (free / outflow) along that axis, clamped into the block
            // along other axes that are physical on that end, and taken from
            // the neighbour along the remaining (periodic / internal) ones.
            const BlockGeo &g = m.geo[b];
            int al = 0;
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if (d[a] != 0 && ((g.phy >> (2 * a + (d[a] > 0))) & 1)) al = a;
            int q[3], e[3];
            bool out2 = false;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int side = d[a] > 0;
                q[a] = p[a];
                if (a == al) {
                    const int kind = m.bnd[2 * a + side];
                    const int gd = side ? p[a] - m.n[a] : -1 - p[a];
                    if (kind == 3) {
                        q[a] = side ? m.n[a] - 1 - gd : gd;
                        mflip = a;
                    } else {
                        q[a] = side ? m.n[a] - 1 : 0;
                        if (kind == 1) mclamp = 2 * a + side;
                        if (kind == 4) fixd = true;
                    }
                } else if (d[a] != 0 && ((g.phy >> (2 * a + side)) & 1)) {
                    q[a] = side ? m.n[a] - 1 : 0;
                }
                e[a] = q[a] < 0 ? -1 : (q[a] >= m.n[a] ? 1 : 0);
                out2 |= e[a] != 0;
            }
            if (out2) {
                blk = g.nbr[nbr_dir(m, e)];
#pragma unroll
                for (int a = 0; a < 3; ++a) q[a] -= e[a] * m.n[a];
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) p[a] = q[a];
        }
    }
    const FT *src = m.W + (size_t)blk * NV * m.ncell + cidx(m.n, p[0], p[1], p[2]);
#pragma unroll
    for (int v = 0; v < NV; ++v) w[v] = src[(size_t)v * m.ncell];
    if constexpr (NV == 8) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (mflip == a) {
                w[MOM + a] = w[MOM + a] * -1;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (c != a) w[BCC + c] = -w[BCC + c];
            }
            if (mclamp >= 0 && (mclamp >> 1) == a) {
                const int side = mclamp & 1;
                if (w[MOM + a] * (side ? -1 : 1) > 0) w[MOM + a] = 0;
            }
        }
    }
    if constexpr (NV == 5) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (flip == a) w[MOM + a] = w[MOM + a] * -1;
            if ((clamp >> 1) == a && clamp >= 0) {
                const int side = clamp & 1;
                if (w[MOM + a] * (side ? -1 : 1) > 0) w[MOM + a] = 0;
            }
        }
    }
    if (fixd) {             // hydro / mhd phys_functors.h fixed_t
#pragma unroll
        for (int v = 0; v < 5; ++v) w[v] = (FT)m.fix_w[v];
        if constexpr (NV == 8) {
#pragma unroll
            for (int c = 0; c < 3; ++c) w[BCC + c] = (FT)m.fix_b[c];
        }
    }
}

// Normal face field of block b at face p (component c); transverse indices
// may sit one row outside the block (MHD transverse ghost lines), which the
// reference fills by the tangential face-field exchange
// (mhd/boundary/comm_functions.h:433-465).
template <class FT>
__device__ __forceinline__ double load_bf(const MeshView<FT> &m, int b, int slot, int c,
                                          const int *pin) {
    int p[3] = {pin[0], pin[1], pin[2]};
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
        const BlockGeo &g = m.geo[b];
        if (g.xface) {
            // mhd/integrator/data.h:108-140: a transverse ghost line next to a
            // cross-level face takes the normal field from the cell value of w
            int bit = 0;
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if (d[a]) bit = 1 << (2 * a + (d[a] > 0));
            if (g.xface & bit) {
                FT w[8];
                load_cell<FT, 8>(m, b, pin, w);
                return (double)w[BCC + c];
            }
        }
        blk = g.nbr[nbr_dir(m, d)];
        if (blk < 0) {
            // physical side (mhd/boundary/phys.h:178-229, phys_functors.h
            // :10-137): the tangential face field of the ghost cell next to
            // the side; free / outflow copy the boundary cell's, reflect
            // takes minus the one at -i (two cells in).  Along the
            // component's own axis the source row is clamped into the block
            // where that axis is physical at the top.
            int a = 0;
#pragma unroll
            for (int x = 0; x < 3; ++x)
                if (d[x]) a = x;
            const int side = d[a] > 0;
            const int kind = m.bnd[2 * a + side];
            if (kind == 4) return m.fix_b[c];
            double sgn = 1;
            if (kind == 3) {
                p[a] = side ? m.n[a] - 2 : 1;
                sgn = -1;
            } else {
                p[a] = side ? m.n[a] - 1 : 0;
            }
            if (((g.phy >> (2 * c + 1)) & 1) && p[c] >= m.n[c]) p[c] = m.n[c] - 1;
            return sgn * m.BF[(((size_t)b * 2 + slot) * 3 + c) * m.nface + fidx(m.n, p[0], p[1], p[2])];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) p[a] -= d[a] * m.n[a];
    }
    return m.BF[(((size_t)blk * 2 + slot) * 3 + c) * m.nface + fidx(m.n, p[0], p[1], p[2])];
}

// Wave64 + workgroup min reduction of a double; lane 0 of the workgroup
// returns the result.
__device__ __forceinline__ double block_min(double v) {
    __shared__ double red[kThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = pmin(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) v = pmin(v, red[i]);
    }
    return v;
}

// ------------------------------------------------------------------- c2p
// hydro/eos/adiabatic.h:211-259 and mhd/eos/adiabatic.h:127-235: primitives
// with floors written back to the conserved slot, cell-centred B from the
// face average, and at substep 0 the per-cell CFL dt reduced per workgroup.
template <class FT, bool MHD>
__global__ void __launch_bounds__(kThreads) k_c2p(MeshView<FT> m, int step) {
    if (m.ts[m.geo[blockIdx.y].grp].skip) return;
    constexpr int NV = MHD ? 8 : 5;
    const int b = blockIdx.y;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dtc = (double)FLT_MAX;
    if (c < m.ncell) {
        const int i = (int)(c % m.n[0]);
        const int j = (int)((c / m.n[0]) % m.n[1]);
        const int k = (int)(c / ((size_t)m.n[0] * m.n[1]));
        double *U = m.U + ((size_t)b * 2 + step) * 5 * m.ncell + c;
        double cons[5], prim[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) cons[v] = U[(size_t)v * m.ncell];
        const EosPar e{m.gamma, m.cfl, m.d_floor, m.t_floor};
        const BlockGeo &g = m.geo[b];
        FT *W = m.W + (size_t)b * NV * m.ncell + c;
        bool hr, he;
        if constexpr (MHD) {
            const double *BF = m.BF + ((size_t)b * 2 + step) * 3 * m.nface;
            double bc[3], bc_2[3], b_max[3];
            double e_m(0);
            const size_t f0 = fidx(m.n, i, j, k);
            const size_t sh[3] = {1, (size_t)(m.n[0] + 1), (size_t)(m.n[0] + 1) * (m.n[1] + 1)};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                // a thin (one-cell) axis has one face per cell, both sides
                // of it (mhd/eos/adiabatic.h:164-177)
                const double b_l = BF[(size_t)a * m.nface + f0];
                const double b_r = m.n[a] < 2 ? b_l : BF[(size_t)a * m.nface + f0 + sh[a]];
                const double bc_m = (b_l + b_r) / 2;
                bc[a] = bc_m;
                const double bcs2 = bc_m * bc_m;
                e_m += bcs2;
                bc_2[a] = bcs2;
                b_max[a] = pmax(kfabs(b_l), kfabs(b_r));
            }
            e_m /= 2;
            const double e_in = cons[ENE];
            cons[ENE] -= e_m;
            c2p(cons, prim, e, hr, he);
            cons[ENE] += e_m;
            // the subtract / add round trip (and a healed pressure) can move
            // the stored energy; where it did not, the slot already holds it
            if (cons[ENE] != e_in) U[(size_t)ENE * m.ncell] = cons[ENE];
            if (hr) U[0] = cons[RHO];
#pragma unroll
            for (int v = 0; v < 5; ++v) W[(size_t)v * m.ncell] = (FT)prim[v];
#pragma unroll
            for (int a = 0; a < 3; ++a) W[(size_t)(BCC + a) * m.ncell] = (FT)bc[a];
            if (m.EC) {     // mhd/eos/adiabatic.h:214-219, from the stored values
                const FT B[3] = {(FT)bc[0], (FT)bc[1], (FT)bc[2]};
                const FT V[3] = {(FT)prim[MOM], (FT)prim[MOM + 1], (FT)prim[MOM + 2]};
                FT *E = m.EC + (size_t)b * 3 * m.ncell + c;
                E[0] = B[1] * V[2] - B[2] * V[1];
                E[m.ncell] = B[2] * V[0] - B[0] * V[2];
                E[2 * m.ncell] = B[0] * V[1] - B[1] * V[0];
            }
            if (step == 0)
                dtc = std::is_same_v<FT, float> ? dt_mhd_mixed(prim, bc_2, b_max, g.dx, m.ndim, m.gamma, m.cfl)
                                                : dt_mhd(prim, bc_2, b_max, g.dx, m.ndim, m.gamma, m.cfl);
        } else {
            const double cs2 = c2p(cons, prim, e, hr, he);
            if (hr) U[0] = cons[RHO];
            if (he) U[(size_t)ENE * m.ncell] = cons[ENE];
#pragma unroll
            for (int v = 0; v < 5; ++v) W[(size_t)v * m.ncell] = (FT)prim[v];
            if (step == 0) dtc = dt_hydro(prim, cs2, g.dx, m.ndim, m.cfl);
        }
    }
    if (step == 0) {
        const double r = block_min(dtc);
        if (threadIdx.x == 0) m.dt_part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = r;
    }
}

// ------------------------------------------------------------------ flux
// One thread per face of the sweep along `axis` (PLM + HLLC / HLLD), the
// flux_kernel of prototypes/integrator.h:14-32 with hydro/integrator/data.h
// :334-352 (hydro) or mhd/integrator/data.h:53-141 (MHD lines, including the
// transverse ghost lines) for the pencil, PLM per face
// (hydro/reconstruct/plm.h:116-153, mhd/reconstruct/plm.h:18-50).
template <class FT, bool MHD, int AXIS>
__global__ void __launch_bounds__(kThreads) k_flux(MeshView<FT> m, int step) {
    constexpr int axis = AXIS;
    if (m.ts[m.geo[row_block(m)].grp].skip) return;
    constexpr int NV = MHD ? 8 : 5;
    const int b = row_block(m);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.nfa[axis]) return;
    int D[3], o[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        D[a] = m.n[a] + (a == axis ? 1 : 2);
        o[a] = a == axis ? 0 : 1;
    }
    int p[3];
    p[0] = (int)(t % D[0]) - o[0];
    p[1] = (int)((t / D[0]) % D[1]) - o[1];
    p[2] = (int)(t / ((size_t)D[0] * D[1])) - o[2];
    constexpr int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    const bool in1 = p[a1] >= 0 && p[a1] < m.n[a1];
    const bool in2 = p[a2] >= 0 && p[a2] < m.n[a2];
    if constexpr (MHD) {
        // transverse ghost lines exist only along active axes
        // (mhd/integrator/integrator.h:152-155)
        if (!in1 && !in2) return;
        if ((!in1 && m.n[a1] < 2) || (!in2 && m.n[a2] < 2)) return;
    } else {
        if (!in1 || !in2) return;
    }
    FT q[4][NV];
    const int f = p[axis];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        int pc[3] = {p[0], p[1], p[2]};
        pc[axis] = f - 2 + s;
        load_cell<FT, NV>(m, b, pc, q[s]);
    }
    FT ql[NV], qr[NV], flx[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        if (MHD && v == BCC + axis) continue;
        if constexpr (std::is_same_v<FT, float>)
            plm_fast(q[0][v], q[1][v], q[2][v], q[3][v], ql[v], qr[v]);
        else
            plm<FT>(q[0][v], q[1][v], q[2][v], q[3][v], ql[v], qr[v]);
    }
    if constexpr (MHD) {
        const FT bn = (FT)load_bf(m, b, step, axis, p);
        ql[BCC + axis] = bn;
        qr[BCC + axis] = bn;
    }
    face_flux<FT, MHD, AXIS>(ql, qr, flx, (FT)m.gamma);
    FT *F = m.F[axis] + (size_t)b * NV * m.nfa[axis] + t;
#pragma unroll
    for (int v = 0; v < NV; ++v) F[(size_t)v * m.nfa[axis]] = flx[v];
}

// ------------------------------------------------------------------- EMF
template <class FT>
__device__ __forceinline__ FT ec_at(const MeshView<FT> &m, int b, const int *p, const int *ax) {
    bool in = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) in &= p[a] >= 0 && p[a] < m.n[a];
    if (in) {   // the block's own cell: the c2p value (equal to B x v of its w)
        const size_t c = cidx(m.n, p[0], p[1], p[2]);
        if (m.EC) return m.EC[((size_t)b * 3 + ax[0]) * m.ncell + c];
        const FT *W = m.W + (size_t)b * 8 * m.ncell + c;
        const size_t s = m.ncell;
        return W[(BCC + ax[1]) * s] * W[(MOM + ax[2]) * s] - W[(BCC + ax[2]) * s] * W[(MOM + ax[1]) * s];
    }
    FT w[8];
    load_cell<FT, 8>(m, b, p, w);
    return w[BCC + ax[1]] * w[MOM + ax[2]] - w[BCC + ax[2]] * w[MOM + ax[1]];
}

template <class FT>
__device__ __forceinline__ FT flux_at(const MeshView<FT> &m, int b, int a, const int *p, int v) {
    return m.F[a][((size_t)b * 8 + v) * m.nfa[a] + xidx(m.n, a, p)];
}

// Edge EMF with the corner-upwind correction, mhd/integrator/integrator.h
// :189-327.  Cell-centred E outside the block is B x v from the neighbour's
// primitives (:201-213); inside it equals the c2p ec.  Thin (one-cell) axes
// come last (m.ndim counts the active ones): they carry one edge layer, no
// shift, no flux and no corner upwinding, and an edge next to one takes the
// surviving sweep at weight 0.5 (:221-300).
template <class FT>
__global__ void __launch_bounds__(kThreads) k_emf(MeshView<FT> m) {
    if (m.ts[m.geo[row_block(m)].grp].skip) return;
    const int b = row_block(m);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.nface) return;
    int i[3];
    i[0] = (int)(t % (m.n[0] + 1));
    i[1] = (int)((t / (m.n[0] + 1)) % (m.n[1] + 1));
    i[2] = (int)(t / ((size_t)(m.n[0] + 1) * (m.n[1] + 1)));
    const int n_dm = m.ndim;
    for (int a = n_dm; a < 3; ++a)
        if (i[a] > 0) return;
    for (int ax = 0; ax < 3; ++ax) {
        if (i[ax] >= m.n[ax]) continue;
        const int ax_1 = mod3(1 + ax), ax_2 = mod3(2 + ax);
        const bool th_1 = ax_1 >= n_dm, th_2 = ax_2 >= n_dm;
        const int axis[3] = {ax, ax_1, ax_2};
        FT ec[2][2];
        ec[1][1] = ec_at(m, b, i, axis);
        if (!th_1) --i[ax_1];
        ec[1][0] = ec_at(m, b, i, axis);
        if (!th_2) --i[ax_2];
        ec[0][0] = ec_at(m, b, i, axis);
        if (!th_1) ++i[ax_1];
        ec[0][1] = ec_at(m, b, i, axis);
        if (!th_2) ++i[ax_2];
        FT ee = 0;
        const FT w_e = (th_1 || th_2) ? (FT)0.5f : (FT)0.25f;
#pragma unroll
        for (int pi = 0; pi < 2; ++pi) {
            const int a_para = pi == 0 ? ax_1 : ax_2;
            const int a_perp = pi == 0 ? ax_2 : ax_1;
            const bool thn_p = a_perp >= n_dm;
            if (a_para < n_dm) {
                if (!thn_p) --i[a_perp];
                const FT ef_m = flux_at(m, b, a_para, i, BCC + ax);
                if (!thn_p) ++i[a_perp];
                const FT ef_p = flux_at(m, b, a_para, i, BCC + ax);
                ee += w_e * (ef_m + ef_p);
                if (!thn_p) {
                    const FT frho_p = flux_at(m, b, a_perp, i, RHO);
                    --i[a_para];
                    const FT frho_m = flux_at(m, b, a_perp, i, RHO);
                    ++i[a_para];
                    const FT de_p = (frho_p > 0 ? (ec[0][1] - ef_m) : (ec[1][1] - ef_p));
                    const FT de_m = (frho_m > 0 ? -(ec[0][0] - ef_m) : -(ec[1][0] - ef_p));
                    ee += w_e * (de_m - de_p);
                }
            }
            if (pi == 0) {
                const FT swp = ec[0][1];
                ec[0][1] = ec[1][0];
                ec[1][0] = swp;
            }
        }
        m.EE[((size_t)b * 3 + ax) * m.nface + t] = ee;
    }
    // boundary-plane advective edge EMF of a high out / fre side
    // (mhd/integrator/integrator.h:328-381): the interior cell's value with
    // v = w[mom] / rho as written there; later axes overwrite earlier ones
    const BlockGeo &g = m.geo[b];
    if (g.phy)
        for (int a = 0; a < 3; ++a) {
            if (i[a] != m.n[a] || m.n[a] < 2 || !((g.phy >> (2 * a + 1)) & 1)) continue;
            const int k = m.bnd[2 * a + 1];
            if (k != 1 && k != 2) continue;
            int jc[3] = {i[0], i[1], i[2]};
            --jc[a];
            FT w[8];
            load_cell<FT, 8>(m, b, jc, w);
            const FT irho = (FT)1 / pmax(w[RHO], (FT)1e-30);
            FT v_c[3], b_c[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                v_c[q] = w[MOM + q] * irho;
                b_c[q] = w[BCC + q];
            }
            const int a1 = mod3(a + 1), a2 = mod3(a + 2);
            m.EE[((size_t)b * 3 + a1) * m.nface + t] = v_c[a] * b_c[a2] - v_c[a2] * b_c[a];
            m.EE[((size_t)b * 3 + a2) * m.nface + t] = v_c[a1] * b_c[a] - v_c[a] * b_c[a1];
        }
}

// ---------------------------------------------------------------- update
// Flux divergence + Heun slot (hydro/integrator/integrator.h:56-68, 156-182,
// 241-265; rk2.h:20-41) and, for MHD, the CT update of the face field
// (mhd/integrator/integrator.h:385-462 with rk2.h:191-205).
//
// One point of the update: the cell at face-box point t of block b (when
// it is a cell) and, for MHD, the faces of that point.  edge(EE, c, idx,
// i, s) reads the edge EMF of component c at edge-box index idx, EE being
// block b's edges; the edge sits at the point's coordinates i, one further
// along axis s when s >= 0.  finish(i, uu) may add to the cell's new state
// before it is stored.
template <class FT, bool MHD, class EdgeRead, class CellFinish>
__device__ __forceinline__ void update_point(const MeshView<FT> &m, int b, size_t t, int step,
                                             EdgeRead edge, CellFinish finish) {
    constexpr int NV = MHD ? 8 : 5;
    int i[3];
    const int e0 = m.n[0] + (MHD ? 1 : 0), e1 = m.n[1] + (MHD ? 1 : 0);
    i[0] = (int)(t % e0);
    i[1] = (int)((t / e0) % e1);
    i[2] = (int)(t / ((size_t)e0 * e1));
    const double dt = m.ts[m.geo[b].grp].dt_use * (step == 0 ? 1.0 : 0.5);
    const BlockGeo &g = m.geo[b];
    if (i[0] < m.n[0] && i[1] < m.n[1] && i[2] < m.n[2]) {
        const size_t c = cidx(m.n, i[0], i[1], i[2]);
        const double *Us = m.U + ((size_t)b * 2 + step) * 5 * m.ncell + c;
        double *Ut = m.U + ((size_t)b * 2 + (1 - step)) * 5 * m.ncell + c;
        double uu[5];
        if (step == 0) {
#pragma unroll
            for (int v = 0; v < 5; ++v) uu[v] = Us[(size_t)v * m.ncell];
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v)
                uu[v] = (Us[(size_t)v * m.ncell] + Ut[(size_t)v * m.ncell]) * 0.5f;
        }
        const double dt_dv = dt / g.vol;
        for (int ax = 0; ax < m.ndim; ++ax) {
            int p1[3] = {i[0], i[1], i[2]};
            ++p1[ax];
            const size_t x0 = xidx(m.n, ax, i), x1 = xidx(m.n, ax, p1);
            const FT *F = m.F[ax] + (size_t)b * NV * m.nfa[ax];
            // faces whose neighbour is finer are skipped here and added with
            // the restricted flux afterwards (integrator.h:156-182 null_l/r)
            const bool n0 = ((g.rface >> (2 * ax)) & 1) && i[ax] == 0;
            const bool n1 = ((g.rface >> (2 * ax + 1)) & 1) && i[ax] + 1 == m.n[ax];
            if constexpr (std::is_same_v<FT, double>) {
                const double fa = g.sfc[ax] * dt_dv;
                const double fa0 = n0 ? -1 : fa, fa1 = n1 ? -1 : fa;
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    double res(0);
                    if (fa0 > 0) res += fa0 * F[(size_t)v * m.nfa[ax] + x0];
                    if (fa1 > 0) res -= fa1 * F[(size_t)v * m.nfa[ax] + x1];
                    uu[v] += res;
                }
            } else {
                const float fa = (float)(g.sfc[ax] * dt_dv);
                const float fa0 = n0 ? -1.f : fa, fa1 = n1 ? -1.f : fa;
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    double res(0);
                    if (fa0 > 0) res += (double)(fa0 * F[(size_t)v * m.nfa[ax] + x0]);
                    if (fa1 > 0) res -= (double)(fa1 * F[(size_t)v * m.nfa[ax] + x1]);
                    uu[v] += res;
                }
            }
        }
        finish(i, uu);
#pragma unroll
        for (int v = 0; v < 5; ++v) Ut[(size_t)v * m.ncell] = uu[v];
    }
    if constexpr (MHD) {
        // the index-1 slot of a thin axis is not a face of the block: it
        // is never computed, and every reader takes slot 0 for both sides
        for (int a = m.ndim; a < 3; ++a)
            if (i[a] > 0) return;
        const FT *EE = m.EE + (size_t)b * 3 * m.nface;
        double *Bs = m.BF + ((size_t)b * 2 + step) * 3 * m.nface + t;
        double *Bt = m.BF + ((size_t)b * 2 + (1 - step)) * 3 * m.nface + t;
        const size_t sh[3] = {1, (size_t)(m.n[0] + 1), (size_t)(m.n[0] + 1) * (m.n[1] + 1)};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int ax_1 = mod3(a + 1), ax_2 = mod3(a + 2);
            if (i[ax_1] >= m.n[ax_1] || i[ax_2] >= m.n[ax_2]) continue;
            // an edge pair spanning a thin axis vanishes
            // (mhd/integrator/integrator.h:419-447)
            const bool th_1 = ax_1 >= m.ndim, th_2 = ax_2 >= m.ndim;
            // FP64 edge circulation (also in the mixed form: the EMF is
            // widened before the product so CT keeps div B at round-off).
            double db = th_2 ? 0.0 : (double)edge(EE, ax_1, t, i, -1) * g.dx[ax_1];
            if (!th_1) {
                db -= (double)edge(EE, ax_2, t, i, -1) * g.dx[ax_2];
                db += (double)edge(EE, ax_2, t + sh[ax_1], i, ax_1) * g.dx[ax_2];
            }
            if (!th_2) db -= (double)edge(EE, ax_1, t + sh[ax_2], i, ax_2) * g.dx[ax_1];
            // the thin face family's area (sfc_thin, integrator.h:74-81)
            const double sfc = a >= m.ndim ? g.dx[ax_1] * g.dx[ax_2] : g.sfc[a];
            db *= (sfc > 0 ? (double)(-dt / sfc) : (double)(0));
            const double bs = Bs[(size_t)a * m.nface];
            Bt[(size_t)a * m.nface] =
                db + (step == 0 ? bs : (bs + Bt[(size_t)a * m.nface]) / 2);
        }
    }
}

template <class FT, bool MHD>
__global__ void __launch_bounds__(kThreads) k_update(MeshView<FT> m, int step) {
    if (m.ts[m.geo[row_block(m)].grp].skip) return;
    const int b = row_block(m);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (MHD ? m.nface : m.ncell)) return;
    update_point<FT, MHD>(
        m, b, t, step,
        [&](const FT *EE, int c, size_t idx, const int *, int) -> FT {
            return EE[(size_t)c * m.nface + idx];
        },
        [](const int *, double *) {});
}

// Max |div B| per workgroup (diagnostic): sum over axes of the face
// difference divided by the cell width, slot 0.
template <class FT>
__global__ void __launch_bounds__(kThreads) k_divb(MeshView<FT> m, double *part) {
    const int b = blockIdx.y;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0;
    if (c < m.ncell) {
        const int i = (int)(c % m.n[0]);
        const int j = (int)((c / m.n[0]) % m.n[1]);
        const int k = (int)(c / ((size_t)m.n[0] * m.n[1]));
        const double *BF = m.BF + (size_t)b * 2 * 3 * m.nface;
        const size_t f0 = fidx(m.n, i, j, k);
        const size_t sh[3] = {1, (size_t)(m.n[0] + 1), (size_t)(m.n[0] + 1) * (m.n[1] + 1)};
        double s = 0;
        for (int a = 0; a < m.ndim; ++a)      // no variation along a thin axis
            s += (BF[(size_t)a * m.nface + f0 + sh[a]] - BF[(size_t)a * m.nface + f0]) / m.geo[b].dx[a];
        v = kfabs(s);
    }
    v = -v;
    const double r = block_min(v);
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = -r;
}

}  // namespace kr
