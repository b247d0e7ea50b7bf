// kr_api.hip -- host side of the C-ABI (include/kratos_amd.h): context,
// block registry, device pools, AoS <-> SoA packing, launch sequencing and
// the hipGraph replay of whole cycles.
#include "../../include/kratos_amd.h"
#include "kr_fused.h"
#include "kr_fused_hydro.h"
#include "kr_smr.h"
#include "kr_sweep_mx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace kr {

__global__ void k_cycle_begin(TimeState *tsv, int ngrp) {
    // cycle::step stop rule (mesh/cycle/cycle.cpp:143-144), per time group
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrp) return;
    TimeState *ts = tsv + g;
    ts->skip = (ts->t >= ts->t_lim) ? 1 : 0;
    if (!ts->skip) ts->cycles += 1;
}

// Global CFL reduction and the dt / t bookkeeping of redc_dt_start/finish
// (mesh/cycle/cycle.cpp:77-115): p_dt = min(p_dt, CFL), dt_use clamped to
// t_lim when topping up, t advanced, next seed = dt_use * dt_expand.  One
// workgroup per time group; the group's blocks are contiguous, so its CFL
// partials are part[grp_first[g] * ppb, grp_first[g + 1] * ppb).
__global__ void __launch_bounds__(1024) k_dt_finalize(TimeState *tsv, const double *part,
                                                      const int *grp_first, int ppb) {
    TimeState *ts = tsv + blockIdx.x;
    if (ts->skip) return;
    __shared__ double red[16];
    double v = (double)FLT_MAX;
    const size_t i0 = (size_t)grp_first[blockIdx.x] * ppb;
    const size_t i1 = (size_t)grp_first[blockIdx.x + 1] * ppb;
    for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) v = pmin(v, part[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = pmin(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) v = pmin(v, red[i]);
        const double p = __builtin_fmin(v, ts->p_dt);
        const double dt_use = ts->top_up ? (p < ts->t_lim - ts->t ? p : ts->t_lim - ts->t) : p;
        if (ts->top_up) {
            const double tn = dt_use + ts->t;
            ts->t = tn < ts->t_lim ? tn : ts->t_lim;
        } else {
            ts->t = dt_use + ts->t;
        }
        ts->dt_use = dt_use;
        ts->p_dt = dt_use * ts->dt_expand;
    }
}

__global__ void __launch_bounds__(1024) k_max_reduce(const double *part, int npart, double *out) {
    __shared__ double red[16];
    double v = 0;
    for (int i = threadIdx.x; i < npart; i += blockDim.x) v = pmax(v, part[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = pmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) v = pmax(v, red[i]);
        *out = v;
    }
}


// Partial maxima of |div B| over blocks whose face field is in the AoS host
// layout of kr_upload (one slot, two ghost layers per side, a thin axis
// with a single face), block after block in bf.  Same expression as
// k_divb: s = sum over the active axes of (B_a(i + e_a) - B_a(i)) / dx_a.
// Workgroup (x, b) covers cells x, x + wpb, ... of block b (grid stride).
__global__ void __launch_bounds__(kThreads) k_divb_aos(const double *bf, int n0, int n1, int n2,
                                                       const double *dx, int wpb, double *part) {
    const int b = blockIdx.y;
    const int n[3] = {n0, n1, n2};
    int N[3], ndim = 0;
    for (int a = 0; a < 3; ++a) {
        N[a] = n[a] + (n[a] > 1) + 4;
        if (n[a] > 1) ndim = a + 1;
    }
    const size_t ncell = (size_t)n0 * n1 * n2;
    const double *B = bf + (size_t)b * N[0] * N[1] * N[2] * 3;
    const size_t sh[3] = {3, (size_t)N[0] * 3, (size_t)N[0] * N[1] * 3};
    double v = 0;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (size_t)wpb * blockDim.x) {
        const int i = (int)(c % n0);
        const int j = (int)((c / n0) % n1);
        const int k = (int)(c / ((size_t)n0 * n1));
        const size_t f0 = (((size_t)(k + 2) * N[1] + (j + 2)) * N[0] + (i + 2)) * 3;
        double s = 0;
        for (int a = 0; a < ndim; ++a) s += (B[f0 + sh[a] + a] - B[f0 + a]) / dx[3 * b + a];
        v = pmax(v, kfabs(s));
    }
    const double r = block_min(-v);
    if (threadIdx.x == 0) part[(size_t)b * wpb + blockIdx.x] = -r;
}

// Device scratch of kr_divb_max_noalloc: the per-workgroup partial maxima.
// One copy per GPU, for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.
constexpr int kNoAllocParts = 65536;
static __device__ __align__(256) double scratch[kNoAllocParts];

}  // namespace kr

using namespace kr;

namespace {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define KR_HIP(call)                                                               \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            throw HipError(std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// graph replays between host checks of the stop rule in kr_advance
constexpr int kAdvanceChunk = 64;

// Fused stage-kernel variants: column tile, workgroup size, the minimum
// waves per SIMD asked of the compiler (1 = no constraint), and the kernel:
// k_sweep (mx = 0) or k_sweep_mx (mx = 1, the mixed form's kernel for MHD
// meshes of one level without physical sides).  The autotuner times every
// variant that fits the block shape and keeps the fastest; on a mesh that
// k_sweep_mx serves it times only those (the two kernels round the FP32
// fluxes differently, and the tuning cycles must leave the state any
// single variant would), k_sweep's 4-wave builds being reachable there
// only when kr_config.tile forces one.
struct SweepVariant {
    int tx, ty, nt, waves, mx;
};
constexpr SweepVariant kSweepVariants[] = {
    {32, 8, 512, 1, 0}, {32, 8, 512, 4, 0}, {16, 16, 512, 1, 0}, {16, 16, 512, 4, 0},
    {16, 8, 256, 1, 0}, {8, 8, 192, 1, 0},  {16, 8, 256, 4, 0},  {16, 8, 256, 1, 1},
    {32, 8, 512, 1, 1}, {16, 16, 512, 1, 1}, {8, 8, 192, 1, 1},
};
constexpr int kNumSweepVariants = sizeof(kSweepVariants) / sizeof(kSweepVariants[0]);

}  // namespace

struct kr_ctx {
    kr_config cfg{};
    std::string err;
    std::vector<kr_block_desc> descs;
    std::unordered_map<int, int> id2idx;
    // A caller's block of twice n_cell_block on every axis is kept as its
    // eight octants, blocks of n_cell_block one level-L index apart: part[b]
    // is -1 for a block the caller gave, 8 m + o for octant o of merged
    // block m (o's bits: x, y, z half)
    std::vector<int> part;
    std::vector<int> merged_id;             // m -> caller's id
    std::unordered_map<int, int> mid2m;     // caller's id -> m
    std::vector<std::array<int, 8>> oct;    // m -> internal index of each octant
    bool committed = false;
    bool mhd = false, mixed = false;
    bool fused = false;
    int variant = -1;               // kSweepVariants index of the fused kernel
    int prod = -1;                  // refined mesh: producer tile (0: 16x8, 1: 8x8), -1 none
    bool mx_list = false;           // k_sweep_mx walks a device work list (ensembles)
    int4 *d_items = nullptr;        // its items (block, tile, z0, z1) and their count
    int *d_nitems = nullptr;
    int slots[16] = {0};            // resident workgroups on the GPU per variant (0: not asked yet)
    std::vector<int> candidates;    // variants the autotuner times
    int tune_next = 0;              // next candidate to time
    bool tuned = false;
    int n[3] = {0, 0, 0};
    int ndim = 0, nv = 5;
    size_t ncell = 0, nface = 0, nfa[3] = {0, 0, 0};
    // device buffers
    double *U = nullptr, *BF = nullptr;
    void *W = nullptr, *F[3] = {nullptr, nullptr, nullptr}, *EE = nullptr, *EC = nullptr;
    BlockGeo *geo = nullptr;
    TimeState *ts = nullptr;        // [ngrp]
    int ngrp = 1;
    int *grp_first = nullptr;       // [ngrp + 1] first internal block of each group
    int ppb = 0;                    // CFL partials per block
    double *dt_part = nullptr, *red_part = nullptr, *red_out = nullptr;
    size_t n_dt_part = 0, n_red_part = 0;
    // refined mesh (blocks at several levels)
    bool smr = false;
    int nhalo = 0;                  // halo blocks after the real ones in W
    SmrView sv{};
    int *d_table = nullptr;
    int4 *d_fill[2] = {nullptr, nullptr};   // restriction, prolongation tasks
    int4 *d_pfill = nullptr;        // physical-side halo tasks {halo, block, direction, 0}
    int npfill = 0;
    int nfill[2] = {0, 0};
    int2 *d_ftask = nullptr;        // faces whose neighbour is finer
    int nftask = 0;
    int *d_xblk = nullptr;          // blocks with cross-level sides
    int nxblk = 0;
    int *d_pure = nullptr;          // blocks with no cross-level direction (fused kernel)
    int npure = 0;
    int *d_sweep = nullptr;         // the pure blocks, then the cross-level ones: the rows
                                    // of the one stage launch of the fused kernel (MHD)
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    kr_stats stats{};
    float tune_ms[16] = {0};
    std::vector<double> stage;   // host staging for packing

    template <class FT> MeshView<FT> view() const {
        MeshView<FT> m{};
        for (int a = 0; a < 3; ++a) m.n[a] = n[a];
        m.ndim = ndim;
        m.nv = nv;
        for (int s = 0; s < 6; ++s) m.bnd[s] = cfg.bnd[s];
        m.gamma = cfg.gamma;
        for (int k = 0; k < 5; ++k) m.fix_w[k] = cfg.fix_w[k];
        for (int k = 0; k < 3; ++k) m.fix_b[k] = cfg.fix_b[k];
        m.cfl = cfg.cfl;
        m.d_floor = cfg.rho_floor;
        m.t_floor = cfg.tmp_floor;
        m.ncell = ncell;
        m.nface = nface;
        for (int a = 0; a < 3; ++a) m.nfa[a] = nfa[a];
        m.U = U;
        m.BF = BF;
        m.W = (FT *)W;
        for (int a = 0; a < 3; ++a) m.F[a] = (FT *)F[a];
        m.EE = (FT *)EE;
        m.EC = (FT *)EC;
        m.geo = geo;
        m.ts = ts;
        m.dt_part = dt_part;
        return m;
    }
};

namespace {

void release(kr_ctx *c) {
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
    void *bufs[] = {c->U,  c->BF, c->W,       c->F[0],  c->F[1],   c->F[2],  c->EE, c->EC,
                    c->geo, c->ts, c->grp_first, c->dt_part, c->red_part, c->red_out,
                    c->d_table, c->d_fill[0], c->d_fill[1], c->d_ftask, c->d_xblk,
                    c->d_pure, c->d_pfill, c->d_sweep, c->d_items, c->d_nitems};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    c->d_table = nullptr;
    c->d_fill[0] = c->d_fill[1] = nullptr;
    c->d_pfill = nullptr;
    c->d_ftask = nullptr;
    c->d_xblk = c->d_pure = c->d_sweep = nullptr;
    c->d_items = nullptr;
    c->d_nitems = nullptr;
    c->U = c->BF = nullptr;
    c->W = c->EE = c->EC = nullptr;
    c->F[0] = c->F[1] = c->F[2] = nullptr;
    c->geo = nullptr;
    c->ts = nullptr;
    c->grp_first = nullptr;
    c->dt_part = c->red_part = c->red_out = nullptr;
}

template <class T> T *dalloc(size_t count) {
    void *p = nullptr;
    if (count == 0) count = 1;
    KR_HIP(hipMalloc(&p, count * sizeof(T)));
    return (T *)p;
}

// Topology of a refined mesh, derived from each block's level and logical
// index as tree_t::neighbors does (mesh/tree/tree.cpp:177-197): for every
// direction the leaf covering the neighbour region is at the same level
// (neighbour index), one level coarser or finer (a halo block and a fill
// task), or outside a non-periodic side (-1).  Validates the 2:1 balance
// and that no prolongation reads a halo that is itself prolonged.
struct SmrBuild {
    int R[3], nlev = 0, n[3];
    bool per[3];
    std::vector<std::vector<int>> table;
    size_t lidx(int L, const int *r) const {
        return ((size_t)r[2] * (size_t)(R[1] << L) + (size_t)r[1]) * (size_t)(R[0] << L) + (size_t)r[0];
    }
    int leaf(int L, const int *r) const {
        if (L < 0 || L >= nlev) return -1;
        for (int a = 0; a < 3; ++a)
            if (r[a] < 0 || r[a] >= (R[a] << L)) return -1;
        return table[L][lidx(L, r)];
    }
    bool wrap(int L, int *G) const {
        for (int a = 0; a < 3; ++a) {
            const int N = (R[a] * n[a]) << L;
            if (G[a] >= 0 && G[a] < N) continue;
            if (!per[a]) return false;
            G[a] = ((G[a] % N) + N) % N;
        }
        return true;
    }
    int cover(int L, const int *G, int &blk) const {
        int r[3];
        for (int a = 0; a < 3; ++a) r[a] = G[a] / n[a];
        if ((blk = leaf(L, r)) >= 0) return 0;
        for (int a = 0; a < 3; ++a) r[a] = (G[a] >> 1) / n[a];
        if ((blk = leaf(L - 1, r)) >= 0) return -1;
        for (int a = 0; a < 3; ++a) r[a] = (2 * G[a]) / n[a];
        if ((blk = leaf(L + 1, r)) >= 0) return 1;
        throw std::invalid_argument("refined mesh: a neighbour region has no leaf within one level");
    }
};

void build_smr(kr_ctx *c, std::vector<BlockGeo> &g) {
    const size_t nb = c->descs.size();
    if (c->ngrp != 1) throw std::invalid_argument("refined meshes use one time group");
    if (c->ndim != 3) throw std::invalid_argument("refined meshes are 3-D");
    SmrBuild S;
    for (int a = 0; a < 3; ++a) {
        S.n[a] = c->n[a];
        S.per[a] = c->cfg.bnd[2 * a] == KR_PER;
        if (c->n[a] < 4 || (c->n[a] & 1))
            throw std::invalid_argument("refined meshes need an even block size >= 4");
        S.R[a] = 0;
    }
    for (auto &d : c->descs) {
        if (d.level < 0 || d.level > 7) throw std::invalid_argument("refinement level out of 0..7");
        S.nlev = std::max(S.nlev, d.level + 1);
        for (int a = 0; a < 3; ++a)
            S.R[a] = std::max(S.R[a], (d.reg[a] + 1 + (1 << d.level) - 1) >> d.level);
    }
    S.table.assign(S.nlev, {});
    for (int L = 0; L < S.nlev; ++L)
        S.table[L].assign((size_t)(S.R[0] << L) * (S.R[1] << L) * (S.R[2] << L), -1);
    for (size_t b = 0; b < nb; ++b) {
        const kr_block_desc &d = c->descs[b];
        for (int a = 0; a < 3; ++a)
            if (d.reg[a] < 0) throw std::invalid_argument("negative logical block index");
        int &slot = S.table[d.level][S.lidx(d.level, d.reg)];
        if (slot >= 0) throw std::invalid_argument("two blocks at one position");
        slot = (int)b;
    }
    std::vector<int4> fill[2];
    std::vector<int2> ftask;
    std::vector<int> xblk, pure;
    int nhalo = 0;
    for (size_t b = 0; b < nb; ++b) {
        const kr_block_desc &d = c->descs[b];
        BlockGeo &G = g[b];
        G.level = d.level;
        for (int a = 0; a < 3; ++a) G.reg[a] = d.reg[a];
        G.rface = G.xface = 0;
        for (int k = 0; k < 27; ++k) {
            G.lev[k] = 0;
            G.nbr[k] = k == 13 ? (int)b : -1;
            if (k == 13) continue;
            const int dd[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
            int lv = 9, src = -1;
            for (int cnr = 0; cnr < 8; ++cnr) {     // the ghost cells must lie on one level
                int H[3];
                for (int a = 0; a < 3; ++a) {
                    const int o = (cnr >> a) & 1;
                    H[a] = d.reg[a] * S.n[a] + (dd[a] < 0 ? -2 + o : dd[a] > 0 ? S.n[a] + o
                                                                               : o * (S.n[a] - 1));
                }
                if (!S.wrap(d.level, H)) { lv = 9; break; }
                int s2;
                const int l2 = S.cover(d.level, H, s2);
                if (cnr == 0) { lv = l2; src = s2; }
                else if (l2 != lv) throw std::invalid_argument("refined mesh is not 2:1 balanced");
            }
            G.lev[k] = (signed char)lv;
            if (lv == 0) G.nbr[k] = src;
            else if (lv == 1 || lv == -1) {
                const int hb = (int)nb + nhalo++;
                G.nbr[k] = hb;
                fill[lv == 1 ? 0 : 1].push_back(make_int4(hb, (int)b, k, 0));
            }
        }
        for (int a = 0; a < 3; ++a)
            for (int sd = 0; sd < 2; ++sd) {
                int dir[3] = {0, 0, 0};
                dir[a] = sd ? 1 : -1;
                const int lv = G.lev[((dir[2] + 1) * 3 + (dir[1] + 1)) * 3 + (dir[0] + 1)];
                if (lv == 1) {
                    G.rface |= 1 << (2 * a + sd);
                    ftask.push_back(make_int2((int)b, 2 * a + sd));
                }
                if (lv == 1 || lv == -1) G.xface |= 1 << (2 * a + sd);
            }
        // the four cells around an edge of component c in regions (r1, r2):
        // offsets -1 / 0 along c1 and c2, outside the block on a surface side
        for (int c = 0; c < 3; ++c)
            for (int r1 = 0; r1 < 3; ++r1)
                for (int r2 = 0; r2 < 3; ++r2) {
                    bool coarser = false;
                    for (int o1 = -1; o1 <= 0; ++o1)
                        for (int o2 = -1; o2 <= 0; ++o2) {
                            int d[3] = {0, 0, 0};
                            d[(c + 1) % 3] = (r1 == 0 && o1 < 0) ? -1 : (r1 == 2 && o1 == 0 ? 1 : 0);
                            d[(c + 2) % 3] = (r2 == 0 && o2 < 0) ? -1 : (r2 == 2 && o2 == 0 ? 1 : 0);
                            coarser |= G.lev[((d[2] + 1) * 3 + (d[1] + 1)) * 3 + (d[0] + 1)] == -1;
                        }
                    if (coarser) G.wbmask |= 1 << ((c * 3 + r1) * 3 + r2);
                }
        bool any = false;
        for (int k = 0; k < 27; ++k) any |= G.lev[k] == 1 || G.lev[k] == -1;
        if (any) xblk.push_back((int)b);
        else pure.push_back((int)b);
    }
    // a prolongation's slopes read the coarser block's cells one step past
    // the covering cell; those must never be prolonged themselves
    for (const int4 &t : fill[1]) {
        const BlockGeo &G = g[t.y];
        const int dd[3] = {t.z % 3 - 1, (t.z / 3) % 3 - 1, t.z / 9 - 1};
        int ext[3], lo[3];
        for (int a = 0; a < 3; ++a) {
            ext[a] = dd[a] == 0 ? S.n[a] : 2;
            lo[a] = dd[a] < 0 ? -2 : (dd[a] > 0 ? S.n[a] : 0);
        }
        for (int z = 0; z < ext[2]; ++z)
            for (int y = 0; y < ext[1]; ++y)
                for (int x = 0; x < ext[0]; ++x) {
                    int Gc[3] = {G.reg[0] * S.n[0] + lo[0] + x, G.reg[1] * S.n[1] + lo[1] + y,
                                 G.reg[2] * S.n[2] + lo[2] + z};
                    S.wrap(G.level, Gc);
                    int cb;
                    S.cover(G.level, Gc, cb);
                    const BlockGeo &C = g[cb];
                    int l[3];
                    for (int a = 0; a < 3; ++a) l[a] = (Gc[a] >> 1) - C.reg[a] * S.n[a];
                    for (int a = 0; a < 3; ++a)
                        for (int sgn = -1; sgn <= 1; sgn += 2) {
                            int q[3] = {l[0], l[1], l[2]};
                            q[a] += sgn;
                            int e[3];
                            for (int x2 = 0; x2 < 3; ++x2) e[x2] = q[x2] < 0 ? -1 : (q[x2] >= S.n[x2] ? 1 : 0);
                            const int k = ((e[2] + 1) * 3 + (e[1] + 1)) * 3 + (e[0] + 1);
                            if (k != 13 && C.lev[k] == -1)
                                throw std::invalid_argument("refined mesh: a prolongation needs a "
                                                            "prolonged halo (refine more widely)");
                        }
                }
    }
    c->smr = true;
    c->nhalo = nhalo;
    for (int a = 0; a < 3; ++a) {
        c->sv.R[a] = S.R[a];
        c->sv.per[a] = S.per[a] ? 1 : 0;
    }
    c->sv.nlev = S.nlev;
    std::vector<int> flat;
    for (int L = 0; L < S.nlev; ++L) {
        c->sv.toff[L] = (long long)flat.size();
        flat.insert(flat.end(), S.table[L].begin(), S.table[L].end());
    }
    c->d_table = dalloc<int>(flat.size());
    KR_HIP(hipMemcpy(c->d_table, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice));
    c->sv.table = c->d_table;
    for (int k = 0; k < 2; ++k) {
        c->nfill[k] = (int)fill[k].size();
        c->d_fill[k] = dalloc<int4>(fill[k].size());
        if (!fill[k].empty())
            KR_HIP(hipMemcpy(c->d_fill[k], fill[k].data(), fill[k].size() * sizeof(int4),
                             hipMemcpyHostToDevice));
    }
    c->nftask = (int)ftask.size();
    c->d_ftask = dalloc<int2>(ftask.size());
    if (!ftask.empty())
        KR_HIP(hipMemcpy(c->d_ftask, ftask.data(), ftask.size() * sizeof(int2), hipMemcpyHostToDevice));
    c->nxblk = (int)xblk.size();
    c->d_xblk = dalloc<int>(xblk.size());
    if (!xblk.empty())
        KR_HIP(hipMemcpy(c->d_xblk, xblk.data(), xblk.size() * sizeof(int), hipMemcpyHostToDevice));
    c->npure = (int)pure.size();
    c->d_pure = dalloc<int>(pure.size());
    if (!pure.empty())
        KR_HIP(hipMemcpy(c->d_pure, pure.data(), pure.size() * sizeof(int), hipMemcpyHostToDevice));
    std::vector<int> rows(pure);
    rows.insert(rows.end(), xblk.begin(), xblk.end());
    c->d_sweep = dalloc<int>(rows.size());
    if (!rows.empty())
        KR_HIP(hipMemcpy(c->d_sweep, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice));
}

// Lay out the pools once every block is known (the reference allocates per
// block in dual_t::setup; here each field is one pool, block-major).
void commit(kr_ctx *c) {
    if (c->committed) return;
    if (c->descs.empty()) throw std::invalid_argument("no blocks registered");
    const size_t nb = c->descs.size();
    // internal order: by time group (stable), so each group's blocks and
    // CFL partials are contiguous
    {
        std::vector<int> perm(nb);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(),
                         [&](int a, int b) { return c->descs[a].group < c->descs[b].group; });
        std::vector<kr_block_desc> d(nb);
        std::vector<int> p(nb);
        for (size_t b = 0; b < nb; ++b) {
            d[b] = c->descs[perm[b]];
            p[b] = c->part[perm[b]];
        }
        c->descs.swap(d);
        c->part.swap(p);
    }
    c->id2idx.clear();
    c->oct.assign(c->merged_id.size(), {});
    for (size_t b = 0; b < nb; ++b) {
        if (c->part[b] < 0) c->id2idx[c->descs[b].id] = (int)b;
        else c->oct[c->part[b] / 8][c->part[b] % 8] = (int)b;
    }
    c->ngrp = c->descs.back().group + 1;
    std::vector<int> first(c->ngrp + 1, 0);
    for (size_t b = 0; b < nb; ++b) first[c->descs[b].group + 1] += 1;
    for (int g = 0; g < c->ngrp; ++g) {
        if (first[g + 1] == 0) throw std::invalid_argument("time group " + std::to_string(g) + " has no blocks");
        first[g + 1] += first[g];
    }
    for (int a = 0; a < 3; ++a) c->n[a] = c->cfg.n_cell_block[a];
    c->ndim = 0;
    for (int a = 0; a < 3; ++a) c->ndim += c->n[a] > 1 ? 1 : 0;
    c->ncell = (size_t)c->n[0] * c->n[1] * c->n[2];
    c->nface = (size_t)(c->n[0] + 1) * (c->n[1] + 1) * (c->n[2] + 1);
    for (int ax = 0; ax < 3; ++ax) {
        size_t s = 1;
        for (int a = 0; a < 3; ++a) s *= (size_t)(c->n[a] + (a == ax ? 1 : 2));
        c->nfa[ax] = s;
    }

    // refined meshes, and meshes with merged blocks, take their topology
    // from level and logical index
    bool multi_level = !c->merged_id.empty();
    for (auto &d : c->descs) multi_level |= d.level != 0;
    std::vector<BlockGeo> g(nb);
    for (size_t b = 0; b < nb; ++b) {
        const kr_block_desc &d = c->descs[b];
        BlockGeo &G = g[b];
        std::memset(&G, 0, sizeof(G));
        for (int a = 0; a < 3; ++a) G.dx[a] = d.dx0[a];
        // meshgen.cpp:263-283: products in the reference's order
        for (int a = 0; a < 3; ++a) {
            double s = 1.;
            for (int m = 1; m < 3; ++m) s *= G.dx[(a + m) % 3];
            G.sfc[a] = s;
        }
        double v = 1.;
        for (int a = 0; a < 3; ++a) v *= G.dx[a];
        G.vol = v;
        G.phy = 0;
        G.grp = d.group;
        for (int k = 0; k < 27; ++k) {
            const int id = multi_level ? -1 : d.nbr[k];   // refined: derived by build_smr
            if (id < 0) {
                G.nbr[k] = -1;
                continue;
            }
            auto it = c->id2idx.find(id);
            if (it == c->id2idx.end())
                throw std::invalid_argument("neighbour id " + std::to_string(id) + " not registered");
            if (c->descs[it->second].group != d.group)
                throw std::invalid_argument("neighbouring blocks in different time groups");
            G.nbr[k] = it->second;
        }
        for (int a = 0; a < 3; ++a)
            for (int s = 0; s < 2; ++s) {
                int dir[3] = {0, 0, 0};
                dir[a] = s ? 1 : -1;
                const int k = ((dir[2] + 1) * 3 + (dir[1] + 1)) * 3 + (dir[0] + 1);
                if (c->n[a] < 2) {
                    // an MHD thin axis is one cell wide over the whole mesh: one
                    // block layer, no neighbour but the block itself
                    // (mhd/boundary/keeper.h:59-99)
                    if (c->mhd && (d.reg[a] != 0 || (G.nbr[k] >= 0 && G.nbr[k] != (int)b)))
                        throw std::invalid_argument("an MHD thin axis must be one block (one cell) wide");
                    continue;
                }
                if (G.nbr[k] < 0) G.phy |= 1 << (2 * a + s);
            }
    }
    if (multi_level) {
        build_smr(c, g);
        for (size_t b = 0; b < nb; ++b) {
            BlockGeo &G = g[b];
            G.phy = 0;
            for (int a = 0; a < 3; ++a)
                for (int sd = 0; sd < 2; ++sd) {
                    int dir[3] = {0, 0, 0};
                    dir[a] = sd ? 1 : -1;
                    if (G.lev[((dir[2] + 1) * 3 + (dir[1] + 1)) * 3 + (dir[0] + 1)] == 9)
                        G.phy |= 1 << (2 * a + sd);
                }
        }
    }
    // fused stage kernel: 3-D blocks divisible by a supported tile
    // (kr_config.tile forces a shape); the variant is picked by timing at
    // the first step.  On a refined mesh the fused kernel takes the blocks
    // with no cross-level direction; the others run, for MHD, its producer
    // form and the staged update, for hydro the staged kernels.  Refined
    // meshes with physical sides (hydro and MHD) run the staged kernels
    // throughout: they apply the side functors where cells are read, while
    // the fused kernels read physical-side ghosts from halo blocks, which
    // are built on one-level meshes only.
    bool phys_sides = false;
    for (int s = 0; s < 6; ++s) phys_sides |= c->cfg.bnd[s] != KR_PER;
    if (c->smr && phys_sides && c->cfg.kernel_path == KR_PATH_FUSED)
        throw std::invalid_argument("the fused kernel does not run on refined meshes with "
                                    "physical sides");
    if (!(c->smr && phys_sides) && c->cfg.kernel_path != KR_PATH_STAGED) {
        const int tx = c->cfg.tile[0], ty = c->cfg.tile[1], tw = c->cfg.tile[2];
        const bool mxmesh = c->mixed && c->mhd && !c->smr && !phys_sides;
        c->candidates.clear();
        for (int v = 0; v < kNumSweepVariants; ++v) {
            const SweepVariant &sv = kSweepVariants[v];
            if (sv.mx && !mxmesh) continue;
            if (mxmesh && !sv.mx && !(tx && tw > 1)) continue;
            if (tx && (sv.tx != tx || sv.ty != ty)) continue;
            if (tw && sv.waves != tw) continue;
            if (sv.waves > 1 && (!c->mixed || !c->mhd)) continue;
            if (c->n[0] % sv.tx == 0 && c->n[1] % sv.ty == 0 && c->n[2] >= 4)
                c->candidates.push_back(v);
        }
        if (c->candidates.empty() && c->cfg.kernel_path == KR_PATH_FUSED)
            throw std::invalid_argument("no fused tile fits the block shape");
        c->fused = !c->candidates.empty();
        if (c->fused) c->variant = c->candidates[0];
        c->tune_next = 0;
        c->tuned = c->candidates.size() <= 1;
        // an ensemble in the mixed form: k_sweep_mx walks a device work list
        // of the boxes still running (room for the most tiles, 8x8, in up to
        // n_z / 8 z-ranges each)
        if (c->fused && mxmesh && c->ngrp > 1) {
            c->mx_list = true;
            const size_t ntile_max = (size_t)(c->n[0] / 8) * (size_t)(c->n[1] / 8);
            const size_t kz_max = (size_t)std::max(1, c->n[2] / 8);
            c->d_items = dalloc<int4>(nb * ntile_max * kz_max);
            c->d_nitems = dalloc<int>(1);
        }
    }
    // The fused kernel reads W through wnbr.  With physical sides, every
    // direction of a block that crosses one gets a halo block in W: after
    // c2p, k_phys_fill writes there the ghost primitives the boundary
    // functors give (load_cell), at the positions a neighbour would hold.
    for (auto &G : g) std::memcpy(G.wnbr, G.nbr, sizeof(G.nbr));
    c->npfill = 0;
    std::vector<int4> pfill;
    if (c->fused && phys_sides && !c->smr) {
        for (size_t b = 0; b < nb; ++b)
            for (int k = 0; k < 27; ++k) {
                if (k == 13 || g[b].nbr[k] >= 0) continue;
                const int hb = (int)nb + c->nhalo++;
                g[b].wnbr[k] = hb;
                pfill.push_back(make_int4(hb, (int)b, k, 0));
            }
        c->npfill = (int)pfill.size();
    }
    // a direction with neither a neighbour nor a halo block (an edge or
    // corner of a hydro mesh given face neighbours only) reads the block
    // itself: those halo columns are staged by the fused kernel but enter
    // no hydro flux, and every staged address stays inside W
    for (size_t b = 0; b < nb; ++b)
        for (int k = 0; k < 27; ++k)
            if (g[b].wnbr[k] < 0) g[b].wnbr[k] = (int)b;
    // pools: W also holds the halo blocks (refined mesh, physical sides)
    const size_t fs = c->mixed ? sizeof(float) : sizeof(double);
    const size_t nw = nb + (size_t)c->nhalo;
    c->U = dalloc<double>(nb * 2 * 5 * c->ncell);
    if (c->mhd) c->BF = dalloc<double>(nb * 2 * 3 * c->nface);
    KR_HIP(hipMalloc(&c->W, nw * c->nv * c->ncell * fs));
    for (int a = 0; a < c->ndim; ++a) KR_HIP(hipMalloc(&c->F[a], nb * c->nv * c->nfa[a] * fs));
    if (c->mhd) KR_HIP(hipMalloc(&c->EE, nb * 3 * c->nface * fs));
    KR_HIP(hipMemset(c->U, 0, nb * 2 * 5 * c->ncell * sizeof(double)));
    if (c->mhd) KR_HIP(hipMemset(c->BF, 0, nb * 2 * 3 * c->nface * sizeof(double)));
    KR_HIP(hipMemset(c->W, 0, nw * c->nv * c->ncell * fs));
    c->geo = dalloc<BlockGeo>(nb);
    KR_HIP(hipMemcpy(c->geo, g.data(), nb * sizeof(BlockGeo), hipMemcpyHostToDevice));
    if (c->npfill) {
        c->d_pfill = dalloc<int4>(pfill.size());
        KR_HIP(hipMemcpy(c->d_pfill, pfill.data(), pfill.size() * sizeof(int4), hipMemcpyHostToDevice));
    }

    TimeState ts{};
    ts.t = c->cfg.t_0;
    ts.t_lim = c->cfg.t_lim;
    ts.p_dt = c->cfg.dt_init < FLT_MIN ? (double)(FLT_MAX / 10) : c->cfg.dt_init;
    ts.dt_use = c->cfg.dt_init;
    ts.dt_expand = c->cfg.dt_expand;
    ts.top_up = c->cfg.top_up;
    std::vector<TimeState> tsv(c->ngrp, ts);
    c->ts = dalloc<TimeState>(c->ngrp);
    KR_HIP(hipMemcpy(c->ts, tsv.data(), c->ngrp * sizeof(ts), hipMemcpyHostToDevice));
    c->grp_first = dalloc<int>(c->ngrp + 1);
    KR_HIP(hipMemcpy(c->grp_first, first.data(), (c->ngrp + 1) * sizeof(int), hipMemcpyHostToDevice));

    c->ppb = (int)ceil_div(c->ncell, kThreads);
    c->n_dt_part = nb * c->ppb;
    c->dt_part = dalloc<double>(c->n_dt_part);
    c->n_red_part = c->n_dt_part;
    c->red_part = dalloc<double>(c->n_red_part);
    c->red_out = dalloc<double>(1);
    // cell-centred E for the staged EMF (c2p writes it only when allocated)
    // producer form of the fused kernel for the blocks of a refined mesh
    // that touch another level: 16x8 or 8x8 column tiles
    c->prod = -1;
    if (c->mhd && c->fused && c->smr && c->n[2] >= 4) {
        if (c->n[0] % 16 == 0 && c->n[1] % 8 == 0) c->prod = 0;
        else if (c->n[0] % 8 == 0 && c->n[1] % 8 == 0) c->prod = 1;
    }
    // mixed form: one stage launch runs both forms (k_sweep<..., PROD = 2>),
    // so the pure blocks take the producer's tile; no tuning.  The FP64 form
    // keeps two launches (the combined kernel would run at one wave per SIMD)
    if (c->prod >= 0 && c->mixed) {
        const int want = c->prod == 0 ? 4 : 5;      // 16x8/256, 8x8/192
        c->candidates.assign(1, want);
        c->variant = want;
        c->tuned = true;
        c->tune_next = 0;
    }
    // cell-centred E for the staged EMF kernel, when it runs
    if (c->mhd && (!c->fused || (c->smr && c->prod < 0)))
        KR_HIP(hipMalloc(&c->EC, nb * 3 * c->ncell * fs));
    c->stats.n_blocks = (int)nb;
    c->stats.n_cells = (long long)(nb * c->ncell);
    // 5 state passes per Heun cycle of (5 conserved + 3 face-B) FP64 values
    c->stats.alg_bytes_per_cell_update = 5.0 * (5 + (c->mhd ? 3 : 0)) * 8.0;
    c->committed = true;
}

// -------------------------------------------------------------- launches
struct Launcher {
    kr_ctx *c;
    bool prof = false;   // events around every launch
    bool tune = false;   // events around the fused stage-kernel launches only
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evs;

    template <class F> void run(int kind, F &&f, bool sweep_launch = false) {
        if (!prof && !(tune && sweep_launch)) {
            f();
            return;
        }
        hipEvent_t a, b;
        KR_HIP(hipEventCreate(&a));
        KR_HIP(hipEventCreate(&b));
        KR_HIP(hipEventRecord(a, c->stream));
        f();
        KR_HIP(hipEventRecord(b, c->stream));
        evs.push_back({kind, {a, b}});
    }

    template <class FT, int TX, int TY, int NT, int W, int PROD = 0>
    void sweep_t(const MeshView<FT> &m, int step, unsigned nb) {
        const unsigned nt = (unsigned)((c->n[0] / TX) * (c->n[1] / TY));
        if (c->npfill)                      // physical sides (one-level meshes only)
            hipLaunchKernelGGL((k_sweep<FT, TX, TY, NT, W, PROD, true>), dim3(nt, nb), dim3(NT), 0,
                               c->stream, m, step);
        else
            hipLaunchKernelGGL((k_sweep<FT, TX, TY, NT, W, PROD, false>), dim3(nt, nb), dim3(NT), 0,
                               c->stream, m, step);
    }
    // fluxes and edge EMFs of the listed blocks into F / EE (no update)
    template <class FT> void sweep_prod(const MeshView<FT> &m, int step, unsigned nb) {
        if (c->prod == 0) sweep_t<FT, 16, 8, 256, 1, 1>(m, step, nb);
        else sweep_t<FT, 8, 8, 192, 1, 1>(m, step, nb);
    }
    template <class FT> void sweep_both(const MeshView<FT> &m, int step, unsigned nb) {
        static_assert(std::is_same_v<FT, float>, "the combined launch is the mixed form's");
        if (c->prod == 0) sweep_t<FT, 16, 8, 256, 1, 2>(m, step, nb);
        else sweep_t<FT, 8, 8, 192, 1, 2>(m, step, nb);
    }

    template <class FT, int TX, int TY, int NT>
    void sweep_hyd_t(const MeshView<FT> &m, int step, unsigned nb) {
        const unsigned nt = (unsigned)((c->n[0] / TX) * (c->n[1] / TY));
        hipLaunchKernelGGL((k_sweep_hyd<FT, TX, TY, NT>), dim3(nt, nb), dim3(NT), 0, c->stream, m,
                           step);
    }
    // hydro: the one-wave-per-SIMD-minimum variants of the tile table
    template <class FT> void sweep_hyd(const MeshView<FT> &m, int step, int variant, unsigned nb) {
        switch (variant) {
        case 0: sweep_hyd_t<FT, 32, 8, 512>(m, step, nb); break;
        case 2: sweep_hyd_t<FT, 16, 16, 512>(m, step, nb); break;
        case 5: sweep_hyd_t<FT, 8, 8, 192>(m, step, nb); break;
        default: sweep_hyd_t<FT, 16, 8, 256>(m, step, nb); break;
        }
    }

    // grid rows: nb blocks (m.blist maps a row to its block when set)
    template <class FT> void sweep(const MeshView<FT> &m, int step, int variant, unsigned nb) {
        constexpr bool f32 = std::is_same_v<FT, float>;
        switch (variant) {
        case 0: sweep_t<FT, 32, 8, 512, 1>(m, step, nb); break;
        case 1: sweep_t<FT, 32, 8, 512, f32 ? 4 : 1>(m, step, nb); break;
        case 2: sweep_t<FT, 16, 16, 512, 1>(m, step, nb); break;
        case 3: sweep_t<FT, 16, 16, 512, f32 ? 4 : 1>(m, step, nb); break;
        case 4: sweep_t<FT, 16, 8, 256, 1>(m, step, nb); break;
        case 5: sweep_t<FT, 8, 8, 192, 1>(m, step, nb); break;
        case 6: sweep_t<FT, 16, 8, 256, f32 ? 4 : 1>(m, step, nb); break;
        default:
            if constexpr (f32) {
                switch (variant) {
                case 7: sweep_mx_t<16, 8, 256>(m, step, nb); break;
                case 8: sweep_mx_t<32, 8, 512>(m, step, nb); break;
                case 9: sweep_mx_t<16, 16, 512>(m, step, nb); break;
                default: sweep_mx_t<8, 8, 192>(m, step, nb); break;
                }
            }
            break;
        }
    }
    // Ensembles: the stage's work list (the tiles of the boxes still
    // running, split into more z-ranges as fewer remain), then a persistent
    // grid of as many workgroups as the GPU holds at once walking it.
    template <int TX, int TY, int NT>
    void sweep_mx_t(const MeshView<float> &m, int step, unsigned nb) {
        const unsigned nt = (unsigned)((c->n[0] / TX) * (c->n[1] / TY));
        if (!c->mx_list) {
            hipLaunchKernelGGL((k_sweep_mx<TX, TY, NT, false>), dim3(nt, nb), dim3(NT), 0, c->stream,
                               m, step, (const int4 *)nullptr, (const int *)nullptr);
            return;
        }
        int &slots = c->slots[c->variant & 15];
        if (slots == 0) {
            int dev = 0, ncu = 0, per_cu = 0;
            KR_HIP(hipGetDevice(&dev));
            KR_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            KR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void *>(k_sweep_mx<TX, TY, NT, true>), NT, 0));
            slots = std::max(1, ncu * per_cu);
        }
        const unsigned lt = (unsigned)std::min<size_t>(1024, ceil_div((size_t)nb, 64) * 64);
        hipLaunchKernelGGL(k_sweep_list, dim3(1), dim3(lt), 0, c->stream, (const TimeState *)c->ts,
                           (const BlockGeo *)c->geo, (int)nb, (int)nt, c->n[2], slots, c->d_items,
                           c->d_nitems);
        hipLaunchKernelGGL((k_sweep_mx<TX, TY, NT, true>), dim3(slots), dim3(NT), 0, c->stream, m,
                           step, (const int4 *)c->d_items, (const int *)c->d_nitems);
    }

    template <class FT, bool MHD> void stage(int step) {
        const MeshView<FT> m = c->view<FT>();
        const unsigned nb = (unsigned)c->descs.size();
        const dim3 blk(kThreads);
        run(KR_K_C2P, [&] {
            hipLaunchKernelGGL((k_c2p<FT, MHD>), dim3(ceil_div(c->ncell, kThreads), nb), blk, 0,
                               c->stream, m, step);
        });
        if (step == 0)
            run(KR_K_DTFIN, [&] {
                hipLaunchKernelGGL(k_dt_finalize, dim3(c->ngrp), dim3(1024), 0, c->stream, c->ts,
                                   (const double *)c->dt_part, (const int *)c->grp_first,
                                   c->ppb);
            });
        constexpr int NV = MHD ? 8 : 5;
        if (c->fused && !c->smr) {
            if (c->npfill)                  // physical-side halos (accounted as c2p)
                run(KR_K_C2P, [&] {
                    const size_t mx = 4 * (size_t)std::max(c->n[0] * c->n[1],
                                                           std::max(c->n[1] * c->n[2], c->n[0] * c->n[2]));
                    hipLaunchKernelGGL((k_phys_fill<FT, NV>), dim3(ceil_div(mx, 256), c->npfill),
                                       dim3(256), 0, c->stream, m, (const int4 *)c->d_pfill);
                });
            if constexpr (MHD)
                run(KR_K_UPDATE, [&] { sweep<FT>(m, step, c->variant, nb); }, true);
            else
                run(KR_K_UPDATE, [&] { sweep_hyd<FT>(m, step, c->variant, nb); }, true);
            return;
        }
        if (c->smr) {                       // halo where levels meet (accounted as c2p)
            size_t mx = 0;                  // largest face region: 2 x n x n
            for (int a = 0; a < 3; ++a) {
                size_t e = 2;
                for (int x = 0; x < 3; ++x)
                    if (x != a) e *= (size_t)c->n[x];
                mx = std::max(mx, e);
            }
            for (int k = 0; k < 2; ++k)
                if (c->nfill[k])
                    run(KR_K_C2P, [&] {
                        hipLaunchKernelGGL((k_smr_fill<FT, NV>), dim3(ceil_div(mx, kSmrThreads), c->nfill[k]),
                                           dim3(kSmrThreads), 0, c->stream, m, c->sv,
                                           (const int4 *)c->d_fill[k], k);
                    });
        }
        // staged kernels: every block, or on a refined mesh with the fused
        // kernel the blocks that touch another level (the rest run fused)
        MeshView<FT> ms = m;
        unsigned ns = nb;
        if (c->smr && c->fused) {
            bool both = false;
            if constexpr (MHD && std::is_same_v<FT, float>)
                if (c->prod >= 0) {
                    // one launch: the full form on the pure blocks' rows, the
                    // producer form (fluxes and EMFs into F / EE) on the others
                    MeshView<FT> mb = m;
                    mb.blist = c->d_sweep;
                    mb.nfull = c->npure;
                    run(KR_K_UPDATE,
                        [&] { sweep_both<FT>(mb, step, (unsigned)(c->npure + c->nxblk)); }, true);
                    both = true;
                }
            if (!both && c->npure) {
                MeshView<FT> mp = m;
                mp.blist = c->d_pure;
                if constexpr (MHD)
                    run(KR_K_UPDATE, [&] { sweep<FT>(mp, step, c->variant, (unsigned)c->npure); }, true);
                else
                    run(KR_K_UPDATE, [&] { sweep_hyd<FT>(mp, step, c->variant, (unsigned)c->npure); },
                        true);
            }
            ms.blist = c->d_xblk;
            ns = (unsigned)c->nxblk;
        }
        if (ns == 0) return;
        if (MHD && c->smr && c->prod >= 0) {
            // producer form: fluxes and EMFs into F / EE (the mixed form ran
            // it in the launch above)
            if constexpr (!std::is_same_v<FT, float>)
                run(KR_K_FLUX_X, [&] { sweep_prod<FT>(ms, step, ns); });
        } else {
        if (c->ndim > 2)
            run(KR_K_FLUX_Z, [&] {
                hipLaunchKernelGGL((k_flux<FT, MHD, 2>), dim3(ceil_div(c->nfa[2], kThreads), ns),
                                   blk, 0, c->stream, ms, step);
            });
        if (c->ndim > 1)
            run(KR_K_FLUX_Y, [&] {
                hipLaunchKernelGGL((k_flux<FT, MHD, 1>), dim3(ceil_div(c->nfa[1], kThreads), ns),
                                   blk, 0, c->stream, ms, step);
            });
        run(KR_K_FLUX_X, [&] {
            hipLaunchKernelGGL((k_flux<FT, MHD, 0>), dim3(ceil_div(c->nfa[0], kThreads), ns), blk,
                               0, c->stream, ms, step);
        });
        if constexpr (MHD)
            run(KR_K_EMF, [&] {
                hipLaunchKernelGGL(k_emf<FT>, dim3(ceil_div(c->nface, kThreads), ns), blk, 0,
                                   c->stream, ms);
            });
        }
        if (c->smr) {
            // act_flx in one launch (EMF restriction on the coarser blocks'
            // shared edges, flux restriction on the faces whose neighbour is
            // finer), then the update with the finer side's EMF write-back and
            // intg_flx folded in
            const int nx = MHD ? c->nxblk : 0;
            if (nx + c->nftask) {
                // largest face side; one row covers an edge box (MHD) or at
                // least one face side, tpr face tasks per row
                const int fmax = c->n[0] * c->n[1] * c->n[2] /
                                 std::min(c->n[0], std::min(c->n[1], c->n[2]));
                const size_t w = ceil_div(std::max(MHD ? c->nface : (size_t)0, (size_t)fmax),
                                          kSmrThreads) * kSmrThreads;
                const int tpr = (int)(w / fmax);
                const int frows = (int)ceil_div((size_t)c->nftask, (size_t)tpr);
                run(MHD ? KR_K_EMF : KR_K_UPDATE, [&] {
                    hipLaunchKernelGGL((k_smr_restrict<FT, NV>), dim3(w / kSmrThreads, nx + frows),
                                       dim3(kSmrThreads), 0, c->stream, m, c->sv,
                                       (const int *)c->d_xblk, nx, (const int2 *)c->d_ftask,
                                       c->nftask, fmax, tpr);
                });
            }
            run(KR_K_UPDATE, [&] {
                hipLaunchKernelGGL((k_smr_update<FT, MHD>),
                                   dim3(ceil_div(MHD ? c->nface : c->ncell, kThreads), ns), blk, 0,
                                   c->stream, ms, step);
            });
            return;
        }
        run(KR_K_UPDATE, [&] {
            hipLaunchKernelGGL((k_update<FT, MHD>),
                               dim3(ceil_div(MHD ? c->nface : c->ncell, kThreads), ns), blk, 0,
                               c->stream, ms, step);
        });
    }

    template <class FT, bool MHD> void cycle_t() {
        run(KR_K_BEGIN, [&] {
            hipLaunchKernelGGL(k_cycle_begin, dim3(ceil_div(c->ngrp, 256)), dim3(256), 0,
                               c->stream, c->ts, c->ngrp);
        });
        stage<FT, MHD>(0);
        stage<FT, MHD>(1);
    }

    void cycle() {
        if (c->mixed) {
            if (c->mhd) cycle_t<float, true>();
            else cycle_t<float, false>();
        } else {
            if (c->mhd) cycle_t<double, true>();
            else cycle_t<double, false>();
        }
        KR_HIP(hipGetLastError());
    }
};

void fill_tile_stats(kr_ctx *c) {
    for (int k = 0; k < 4; ++k) c->stats.sweep_tile[k] = 0;
    c->stats.sweep_kernel = 0;
    c->stats.tune_pending = (c->fused && !c->tuned) ? (int)c->candidates.size() - c->tune_next : 0;
    c->stats.graph_ready = c->graph != nullptr;
    if (!c->fused) return;
    const SweepVariant &sv = kSweepVariants[c->variant];
    c->stats.sweep_tile[0] = sv.tx;
    c->stats.sweep_tile[1] = sv.ty;
    c->stats.sweep_tile[2] = sv.nt;
    c->stats.sweep_tile[3] = sv.waves;
    c->stats.sweep_kernel = sv.mx;
    for (int v = 0; v < kNumSweepVariants && v < 16; ++v) c->stats.tune_ms[v] = c->tune_ms[v];
}

// Tile autotuner on real cycles: while candidates are left, each cycle runs
// the next candidate with events around its two stage-kernel launches; when
// all are timed the fastest is kept.  Every variant performs the same
// arithmetic per face and per cell in the same order, so tuning cycles leave
// exactly the state any fixed choice would.
bool tuning(const kr_ctx *c) { return c->fused && !c->tuned; }

void tune_cycle(kr_ctx *c) {
    const int v = c->candidates[c->tune_next];
    c->variant = v;
    Launcher L{c, false, true};
    L.cycle();
    KR_HIP(hipStreamSynchronize(c->stream));
    float ms = 0;
    for (auto &e : L.evs) {
        float t = 0;
        KR_HIP(hipEventElapsedTime(&t, e.second.first, e.second.second));
        ms += t;
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    c->tune_ms[v] = L.evs.empty() ? 0.f : ms / (float)L.evs.size();
    if (++c->tune_next == (int)c->candidates.size()) {
        int pick = c->candidates[0];
        for (int w : c->candidates)
            if (c->tune_ms[w] < c->tune_ms[pick]) pick = w;
        c->variant = pick;
        c->tuned = true;
    }
}

// Message of the last kr_create that returned no context on this thread
// (kr_last_error(NULL) returns it).
thread_local std::string g_create_err;

int fail(kr_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

template <class F> int guarded(kr_ctx *c, F &&f) {
    if (!c) return KR_EINVAL;
    try {
        return f();
    } catch (const HipError &e) {
        return fail(c, KR_EHIP, e.what());
    } catch (const std::invalid_argument &e) {
        return fail(c, KR_EINVAL, e.what());
    } catch (const std::bad_alloc &e) {
        return fail(c, KR_ENOMEM, e.what());
    } catch (const std::exception &e) {
        return fail(c, KR_ESTATE, e.what());
    }
}

TimeState read_ts(kr_ctx *c, int g = 0) {
    TimeState ts{};
    KR_HIP(hipMemcpy(&ts, c->ts + g, sizeof(ts), hipMemcpyDeviceToHost));
    return ts;
}

std::vector<TimeState> read_all_ts(kr_ctx *c) {
    std::vector<TimeState> v(c->ngrp);
    KR_HIP(hipMemcpy(v.data(), c->ts, c->ngrp * sizeof(TimeState), hipMemcpyDeviceToHost));
    return v;
}

// cycles run by the slowest-stopping group between two snapshots, and
// whether every group has reached t_lim
int cycles_between(const std::vector<TimeState> &a, const std::vector<TimeState> &b) {
    int n = 0;
    for (size_t g = 0; g < a.size(); ++g) n = std::max(n, b[g].cycles - a[g].cycles);
    return n;
}
bool all_done(const std::vector<TimeState> &v) {
    for (auto &t : v)
        if (t.t < t.t_lim) return false;
    return true;
}

// Address of the calling thread's current device's copy of the no-alloc
// scratch, looked up once per device (with that device current) and kept.
double *noalloc_scratch() {
    static std::mutex mu;
    static std::unordered_map<int, double *> addr;
    int dev = 0;
    KR_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    double *p = nullptr;
    KR_HIP(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(kr::scratch)));
    addr[dev] = p;
    return p;
}

}  // namespace

namespace {

// One block of n_cell_block cells from the reference's host layout into
// slot 0 of the pools: u_aos interior cells, bf_aos the ghosted face array.
void upload_one(kr_ctx *c, size_t b, const double *u_aos, const double *bf_aos) {
    const int nvar = c->mhd ? 8 : 5;
    std::vector<double> &s = c->stage;
    s.assign(5 * c->ncell, 0.0);
    for (size_t cc = 0; cc < c->ncell; ++cc)
        for (int v = 0; v < 5; ++v) s[v * c->ncell + cc] = u_aos[cc * nvar + v];
    KR_HIP(hipMemcpy(c->U + b * 2 * 5 * c->ncell, s.data(), 5 * c->ncell * sizeof(double),
                     hipMemcpyHostToDevice));
    if (c->mhd) {
        // host face array: n + (n > 1) faces per axis plus two ghost layers
        // each side (mhd/types/fld_data.cpp:18-26); a thin axis has one
        const int nf[3] = {c->n[0] + (c->n[0] > 1), c->n[1] + (c->n[1] > 1),
                           c->n[2] + (c->n[2] > 1)};
        const int N0 = nf[0] + 4, N1 = nf[1] + 4;
        s.assign(3 * c->nface, 0.0);
        for (int k = 0; k < nf[2]; ++k)
            for (int j = 0; j < nf[1]; ++j)
                for (int i = 0; i < nf[0]; ++i) {
                    const size_t src = ((size_t)((k + 2) * N1 + (j + 2)) * N0 + (i + 2)) * 3;
                    const size_t f = ((size_t)k * (c->n[1] + 1) + j) * (c->n[0] + 1) + i;
                    for (int a = 0; a < 3; ++a) s[a * c->nface + f] = bf_aos[src + a];
                }
        KR_HIP(hipMemcpy(c->BF + b * 2 * 3 * c->nface, s.data(), 3 * c->nface * sizeof(double),
                         hipMemcpyHostToDevice));
    }
}

// One block of n_cell_block cells back into the reference's host layout
// (cell-centred B from the face average; faces with zero ghost layers).
void download_one(kr_ctx *c, size_t b, double *u_aos, double *bf_aos) {
    const int nvar = c->mhd ? 8 : 5;
    std::vector<double> s(5 * c->ncell);
    KR_HIP(hipMemcpy(s.data(), c->U + b * 2 * 5 * c->ncell, 5 * c->ncell * sizeof(double),
                     hipMemcpyDeviceToHost));
    for (size_t cc = 0; cc < c->ncell; ++cc)
        for (int v = 0; v < 5; ++v) u_aos[cc * nvar + v] = s[v * c->ncell + cc];
    if (c->mhd) {
        std::vector<double> f(3 * c->nface);
        KR_HIP(hipMemcpy(f.data(), c->BF + b * 2 * 3 * c->nface, 3 * c->nface * sizeof(double),
                         hipMemcpyDeviceToHost));
        // a thin axis has one face, both sides of the cell
        const size_t sh[3] = {c->n[0] > 1 ? (size_t)1 : 0,
                              c->n[1] > 1 ? (size_t)(c->n[0] + 1) : 0,
                              c->n[2] > 1 ? (size_t)(c->n[0] + 1) * (c->n[1] + 1) : 0};
        for (int k = 0; k < c->n[2]; ++k)
            for (int j = 0; j < c->n[1]; ++j)
                for (int i = 0; i < c->n[0]; ++i) {
                    const size_t cc = ((size_t)k * c->n[1] + j) * c->n[0] + i;
                    const size_t f0 = ((size_t)k * (c->n[1] + 1) + j) * (c->n[0] + 1) + i;
                    for (int a = 0; a < 3; ++a)
                        u_aos[cc * nvar + BCC + a] =
                            (f[a * c->nface + f0] + f[a * c->nface + f0 + sh[a]]) / 2;
                }
        if (bf_aos) {
            const int nf[3] = {c->n[0] + (c->n[0] > 1), c->n[1] + (c->n[1] > 1),
                               c->n[2] + (c->n[2] > 1)};
            const int N0 = nf[0] + 4, N1 = nf[1] + 4, N2 = nf[2] + 4;
            std::memset(bf_aos, 0, (size_t)N0 * N1 * N2 * 3 * sizeof(double));
            for (int k = 0; k < nf[2]; ++k)
                for (int j = 0; j < nf[1]; ++j)
                    for (int i = 0; i < nf[0]; ++i) {
                        const size_t dst = ((size_t)((k + 2) * N1 + (j + 2)) * N0 + (i + 2)) * 3;
                        const size_t f0 = ((size_t)k * (c->n[1] + 1) + j) * (c->n[0] + 1) + i;
                        for (int a = 0; a < 3; ++a) bf_aos[dst + a] = f[a * c->nface + f0];
                    }
        }
    }
}

// Host arrays of a block of 2n cells per axis <-> its octant o's arrays of
// n cells: cells [h n, h n + n) and faces [h n, h n + n] per axis, h the
// octant's bit for that axis.  `to_oct`: from the block's arrays (src) into
// the octant's (dst); otherwise from the octant's (src) into the block's.
void octant_copy(const kr_ctx *c, int o, bool to_oct, const double *src_u, const double *src_bf,
                 double *dst_u, double *dst_bf) {
    const int nvar = c->mhd ? 8 : 5;
    const int *n = c->n;
    int h[3];
    for (int a = 0; a < 3; ++a) h[a] = ((o >> a) & 1) * n[a];
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const size_t so = (((size_t)k * n[1] + j) * n[0] + i) * nvar;
                const size_t sb = (((size_t)(k + h[2]) * 2 * n[1] + (j + h[1])) * 2 * n[0] + (i + h[0])) * nvar;
                const size_t from = to_oct ? sb : so, to = to_oct ? so : sb;
                for (int v = 0; v < nvar; ++v) dst_u[to + v] = src_u[from + v];
            }
    if (!c->mhd || !src_bf || !dst_bf) return;
    // ghosted face arrays: n + 1 faces per axis plus two ghost layers a side
    const int M0 = n[0] + 5, M1 = n[1] + 5, B0 = 2 * n[0] + 5, B1 = 2 * n[1] + 5;
    for (int k = 0; k <= n[2]; ++k)
        for (int j = 0; j <= n[1]; ++j)
            for (int i = 0; i <= n[0]; ++i) {
                const size_t so = (((size_t)(k + 2) * M1 + (j + 2)) * M0 + (i + 2)) * 3;
                const size_t sb = (((size_t)(k + h[2] + 2) * B1 + (j + h[1] + 2)) * B0 + (i + h[0] + 2)) * 3;
                const size_t from = to_oct ? sb : so, to = to_oct ? so : sb;
                for (int a = 0; a < 3; ++a) dst_bf[to + a] = src_bf[from + a];
            }
}

}  // namespace

extern "C" {

int kr_create(const kr_config *cfg, kr_ctx **out) {
    g_create_err.clear();
    if (!cfg || !out) {
        g_create_err = "kr_create: null config or output pointer";
        return KR_EINVAL;
    }
    *out = nullptr;
    auto *c = new (std::nothrow) kr_ctx();
    if (!c) return KR_ENOMEM;
    c->cfg = *cfg;
    c->mhd = cfg->model == KR_MHD;
    c->mixed = cfg->precision == KR_MIXED;
    c->nv = c->mhd ? 8 : 5;
    std::string why;
    if (cfg->model != KR_HYDRO && cfg->model != KR_MHD) why = "unknown model";
    else if (cfg->precision != KR_FP64 && cfg->precision != KR_MIXED) why = "unknown precision";
    else if (!(cfg->gamma > 1)) why = "gamma must exceed 1";
    else if (!(cfg->cfl > 0)) why = "cfl must be positive";
    // thin (one-cell) axes are the trailing ones: the sweeps run over the
    // first n_dim axes, n_dim counting the axes with more than one cell
    // (geometry.cpp:46, dynamics calc_flx), and the MHD EMF and CT assume
    // the same (mhd/integrator/integrator.h:221-227)
    if (why.empty()) {
        if (c->mhd && cfg->n_cell_block[0] < 2) why = "MHD needs at least two cells along x";
        else if ((cfg->n_cell_block[0] < 2 && (cfg->n_cell_block[1] >= 2 || cfg->n_cell_block[2] >= 2)) ||
                 (cfg->n_cell_block[1] < 2 && cfg->n_cell_block[2] >= 2))
            why = "thin axes must be the trailing ones";
    }
    for (int a = 0; a < 3 && why.empty(); ++a) {
        if (cfg->n_cell_block[a] < 1) why = "n_cell_block must be >= 1";
        if ((cfg->bnd[2 * a] == KR_PER) != (cfg->bnd[2 * a + 1] == KR_PER))
            why = "periodic must be set on both sides of an axis";
    }
    for (int s = 0; s < 6 && why.empty(); ++s)
        if (cfg->bnd[s] < KR_PER || cfg->bnd[s] > KR_FIX) why = "unknown boundary kind";
    if (!why.empty()) {
        g_create_err = "kr_create: " + why;
        delete c;
        return KR_EINVAL;
    }
    int rc = guarded(c, [&] {
        int ndev = 0;
        KR_HIP(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw HipError("no HIP device");
        KR_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        KR_HIP(hipEventCreate(&c->ev0));
        KR_HIP(hipEventCreate(&c->ev1));
        return KR_OK;
    });
    if (rc != KR_OK) {
        g_create_err = "kr_create: " + c->err;
        delete c;
        return rc;
    }
    *out = c;
    return KR_OK;
}

int kr_add_blocks(kr_ctx *c, int n, const kr_block_desc *blocks) {
    return guarded(c, [&] {
        if (c->committed) throw std::runtime_error("blocks must be added before the first upload/step");
        if (n < 0 || (n > 0 && !blocks)) throw std::invalid_argument("bad block list");
        for (int i = 0; i < n; ++i) {
            const kr_block_desc &d = blocks[i];
            bool same = true, twice = true;
            for (int a = 0; a < 3; ++a) {
                same = same && d.n_ceff[a] == c->cfg.n_cell_block[a];
                twice = twice && d.n_ceff[a] == 2 * c->cfg.n_cell_block[a];
            }
            if (!same && !twice)
                throw std::invalid_argument("block size is neither n_cell_block nor twice it on every axis");
            if (d.level < 0 || d.level > 7)
                throw std::invalid_argument("refinement level out of 0..7");
            if (d.group < 0 || d.group >= (1 << 20)) throw std::invalid_argument("bad time group");
            if (c->id2idx.count(d.id) || c->mid2m.count(d.id))
                throw std::invalid_argument("duplicate block id");
            if (same) {
                c->id2idx[d.id] = (int)c->descs.size();
                c->descs.push_back(d);
                c->part.push_back(-1);
                continue;
            }
            // twice n_cell_block: its eight octants at the block's level,
            // reg the index of the lowest one; neighbours from level and reg
            const int m = (int)c->merged_id.size();
            c->merged_id.push_back(d.id);
            c->mid2m[d.id] = m;
            for (int o = 0; o < 8; ++o) {
                kr_block_desc od = d;
                for (int a = 0; a < 3; ++a) {
                    const int h = (o >> a) & 1;
                    od.n_ceff[a] = c->cfg.n_cell_block[a];
                    od.reg[a] = d.reg[a] + h;
                    od.xf0[a] = d.xf0[a] + h * (c->cfg.n_cell_block[a] * d.dx0[a]);
                }
                for (int k = 0; k < 27; ++k) od.nbr[k] = -1;
                c->descs.push_back(od);
                c->part.push_back(8 * m + o);
            }
        }
        return KR_OK;
    });
}

int kr_upload(kr_ctx *c, int id, const double *u_aos, const double *bf_aos) {
    return guarded(c, [&] {
        commit(c);
        if (!u_aos || (c->mhd && !bf_aos)) throw std::invalid_argument("null state pointer");
        auto mi = c->mid2m.find(id);
        if (mi != c->mid2m.end()) {
            // a block of twice n_cell_block: each octant from its part
            const int nvar = c->mhd ? 8 : 5;
            std::vector<double> u(c->ncell * nvar), bf;
            if (c->mhd) bf.assign((size_t)(c->n[0] + 5) * (c->n[1] + 5) * (c->n[2] + 5) * 3, 0.0);
            for (int o = 0; o < 8; ++o) {
                octant_copy(c, o, true, u_aos, bf_aos, u.data(), c->mhd ? bf.data() : nullptr);
                upload_one(c, (size_t)c->oct[mi->second][o], u.data(), c->mhd ? bf.data() : nullptr);
            }
        } else {
            auto it = c->id2idx.find(id);
            if (it == c->id2idx.end()) throw std::invalid_argument("unknown block id");
            upload_one(c, (size_t)it->second, u_aos, bf_aos);
        }
        return KR_OK;
    });
}

int kr_download(kr_ctx *c, int id, double *u_aos, double *bf_aos) {
    return guarded(c, [&] {
        commit(c);
        if (!u_aos) throw std::invalid_argument("null state pointer");
        auto mi = c->mid2m.find(id);
        auto it = c->id2idx.find(id);
        if (mi == c->mid2m.end() && it == c->id2idx.end())
            throw std::invalid_argument("unknown block id");
        KR_HIP(hipStreamSynchronize(c->stream));
        if (mi != c->mid2m.end()) {
            // a block of twice n_cell_block, gathered from its octants; a face
            // two octants share is taken from the upper one (both hold it
            // bit for bit)
            const int nvar = c->mhd ? 8 : 5;
            std::vector<double> u(c->ncell * nvar), bf;
            if (c->mhd && bf_aos) {
                bf.assign((size_t)(c->n[0] + 5) * (c->n[1] + 5) * (c->n[2] + 5) * 3, 0.0);
                std::memset(bf_aos, 0,
                            (size_t)(2 * c->n[0] + 5) * (2 * c->n[1] + 5) * (2 * c->n[2] + 5) * 3 * sizeof(double));
            }
            for (int o = 0; o < 8; ++o) {
                download_one(c, (size_t)c->oct[mi->second][o], u.data(), bf.empty() ? nullptr : bf.data());
                octant_copy(c, o, false, u.data(), bf.empty() ? nullptr : bf.data(), u_aos,
                            bf.empty() ? nullptr : bf_aos);
            }
        } else {
            download_one(c, (size_t)it->second, u_aos, bf_aos);
        }
        return KR_OK;
    });
}

int kr_step(kr_ctx *c, double *dt_taken) {
    return guarded(c, [&] {
        commit(c);
        const auto before = read_all_ts(c);
        KR_HIP(hipEventRecord(c->ev0, c->stream));
        if (tuning(c)) {
            tune_cycle(c);
        } else {
            Launcher L{c};
            L.cycle();
        }
        KR_HIP(hipEventRecord(c->ev1, c->stream));
        KR_HIP(hipStreamSynchronize(c->stream));
        float ms = 0;
        KR_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        const auto after = read_all_ts(c);
        c->stats.gpu_ms = ms;
        fill_tile_stats(c);
        c->stats.cycles = cycles_between(before, after);
        if (dt_taken)
            *dt_taken = after[0].cycles != before[0].cycles ? after[0].dt_use : 0.0;
        return KR_OK;
    });
}

int kr_advance(kr_ctx *c, int n_cycles, int *n_done) {
    return guarded(c, [&] {
        commit(c);
        if (n_cycles < 0) throw std::invalid_argument("n_cycles < 0");
        const auto before = read_all_ts(c);
        double ms_tot = 0;
        int left = n_cycles;
        bool done = all_done(before);
        // tuning cycles first (eager launches), then the captured graph
        while (left > 0 && !done && tuning(c)) {
            KR_HIP(hipEventRecord(c->ev0, c->stream));
            tune_cycle(c);
            KR_HIP(hipEventRecord(c->ev1, c->stream));
            KR_HIP(hipEventSynchronize(c->ev1));
            float ms = 0;
            KR_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            ms_tot += ms;
            --left;
            done = all_done(read_all_ts(c));
        }
        if (left > 0 && !done && !c->graph) {
            hipGraph_t g;
            KR_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            Launcher L{c};
            L.cycle();
            KR_HIP(hipStreamEndCapture(c->stream, &g));
            KR_HIP(hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0));
            KR_HIP(hipGraphDestroy(g));
        }
        // replays in chunks; between chunks the host stops early once every
        // group has reached t_lim (later replays would skip all work)
        while (left > 0 && !done) {
            const int chunk = left < kAdvanceChunk ? left : kAdvanceChunk;
            KR_HIP(hipEventRecord(c->ev0, c->stream));
            for (int i = 0; i < chunk; ++i) KR_HIP(hipGraphLaunch(c->graph, c->stream));
            KR_HIP(hipEventRecord(c->ev1, c->stream));
            KR_HIP(hipStreamSynchronize(c->stream));
            float ms = 0;
            KR_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            ms_tot += ms;
            left -= chunk;
            if (left > 0) done = all_done(read_all_ts(c));
        }
        const auto after = read_all_ts(c);
        c->stats.gpu_ms = ms_tot;
        fill_tile_stats(c);
        c->stats.cycles = cycles_between(before, after);
        if (n_done) *n_done = c->stats.cycles;
        return KR_OK;
    });
}

int kr_profile(kr_ctx *c, int n_cycles) {
    return guarded(c, [&] {
        commit(c);
        Launcher L{c, true};
        for (int i = 0; i < n_cycles; ++i) L.cycle();
        KR_HIP(hipStreamSynchronize(c->stream));
        double sum[8] = {0};
        long long cnt[8] = {0};
        for (auto &e : L.evs) {
            float ms = 0;
            KR_HIP(hipEventElapsedTime(&ms, e.second.first, e.second.second));
            sum[e.first] += ms;
            cnt[e.first] += 1;
            (void)hipEventDestroy(e.second.first);
            (void)hipEventDestroy(e.second.second);
        }
        for (int k = 0; k < 8; ++k) {
            c->stats.kernel_ms[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
            c->stats.kernel_calls[k] = cnt[k];
        }
        return KR_OK;
    });
}

int kr_get_time(kr_ctx *c, double *t, double *dt_last, double *dt_seed) {
    return guarded(c, [&] {
        commit(c);
        KR_HIP(hipStreamSynchronize(c->stream));
        const TimeState ts = read_ts(c);
        if (t) *t = ts.t;
        if (dt_last) *dt_last = ts.dt_use;
        if (dt_seed) *dt_seed = ts.p_dt;
        return KR_OK;
    });
}

int kr_set_time(kr_ctx *c, double t, double dt_seed) {
    return guarded(c, [&] {
        commit(c);
        KR_HIP(hipStreamSynchronize(c->stream));
        auto v = read_all_ts(c);
        for (auto &ts : v) {
            ts.t = t;
            ts.p_dt = dt_seed < FLT_MIN ? (double)(FLT_MAX / 10) : dt_seed;
        }
        KR_HIP(hipMemcpy(c->ts, v.data(), v.size() * sizeof(TimeState), hipMemcpyHostToDevice));
        return KR_OK;
    });
}

int kr_get_group_time(kr_ctx *c, int group, double *t, double *dt_last, int *cycles) {
    return guarded(c, [&] {
        commit(c);
        if (group < 0 || group >= c->ngrp) throw std::invalid_argument("no such time group");
        KR_HIP(hipStreamSynchronize(c->stream));
        const TimeState ts = read_ts(c, group);
        if (t) *t = ts.t;
        if (dt_last) *dt_last = ts.dt_use;
        if (cycles) *cycles = ts.cycles;
        return KR_OK;
    });
}

int kr_divb_max(kr_ctx *c, double *divb) {
    return guarded(c, [&] {
        commit(c);
        if (!divb) throw std::invalid_argument("null output");
        if (!c->mhd) {
            *divb = 0;
            return KR_OK;
        }
        const unsigned nb = (unsigned)c->descs.size();
        const unsigned gx = ceil_div(c->ncell, kThreads);
        if (c->mixed)
            hipLaunchKernelGGL(k_divb<float>, dim3(gx, nb), dim3(kThreads), 0, c->stream,
                               c->view<float>(), c->red_part);
        else
            hipLaunchKernelGGL(k_divb<double>, dim3(gx, nb), dim3(kThreads), 0, c->stream,
                               c->view<double>(), c->red_part);
        hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(1024), 0, c->stream,
                           (const double *)c->red_part, (int)(gx * nb), c->red_out);
        KR_HIP(hipGetLastError());
        KR_HIP(hipMemcpyAsync(divb, c->red_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        KR_HIP(hipStreamSynchronize(c->stream));
        return KR_OK;
    });
}


int kr_divb_max_noalloc(const double *bf_dev, int n_blocks, const int n_cell_block[3],
                        const double *dx_dev, double *divb_dev, void *stream) {
    g_create_err.clear();
    try {
        if (!bf_dev || !n_cell_block || !dx_dev || !divb_dev || n_blocks < 1)
            throw std::invalid_argument("kr_divb_max_noalloc: null pointer or no block");
        for (int a = 0; a < 3; ++a)
            if (n_cell_block[a] < 1)
                throw std::invalid_argument("kr_divb_max_noalloc: cells per block < 1");
        if (n_blocks > kr::kNoAllocParts)
            throw std::invalid_argument("kr_divb_max_noalloc: " + std::to_string(n_blocks) +
                                        " blocks exceed the scratch limit of " +
                                        std::to_string(kr::kNoAllocParts));
        double *part = noalloc_scratch();
        const size_t ncell = (size_t)n_cell_block[0] * n_cell_block[1] * n_cell_block[2];
        const int wpb = (int)std::max<size_t>(
            1, std::min<size_t>(ceil_div(ncell, kr::kThreads), kr::kNoAllocParts / n_blocks));
        hipStream_t st = (hipStream_t)stream;
        hipLaunchKernelGGL(kr::k_divb_aos, dim3(wpb, n_blocks), dim3(kr::kThreads), 0, st, bf_dev,
                           n_cell_block[0], n_cell_block[1], n_cell_block[2], dx_dev, wpb, part);
        hipLaunchKernelGGL(kr::k_max_reduce, dim3(1), dim3(1024), 0, st, (const double *)part,
                           wpb * n_blocks, divb_dev);
        KR_HIP(hipGetLastError());
        return KR_OK;
    } catch (const HipError &e) {
        g_create_err = e.what();
        return KR_EHIP;
    } catch (const std::invalid_argument &e) {
        g_create_err = e.what();
        return KR_EINVAL;
    } catch (const std::exception &e) {
        g_create_err = e.what();
        return KR_ESTATE;
    }
}

int kr_get_stats(kr_ctx *c, kr_stats *out) {
    if (!c || !out) return KR_EINVAL;
    *out = c->stats;
    return KR_OK;
}

int kr_sync(kr_ctx *c) {
    return guarded(c, [&] {
        KR_HIP(hipStreamSynchronize(c->stream));
        return KR_OK;
    });
}

const char *kr_last_error(kr_ctx *c) {
    if (c) return c->err.c_str();
    return g_create_err.empty() ? "null context" : g_create_err.c_str();
}

void kr_destroy(kr_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"
