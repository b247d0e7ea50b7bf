// kr_mesh.h -- device-side mesh description shared by the kernels and the
// host library.  Plain structs only: they are copied by value into kernel
// arguments.
#pragma once

#include <cstddef>
#include <cstdint>

namespace kr {

// Per-block geometry and topology, resident in HBM (one entry per block).
// Uniform Cartesian blocks: the reference's geo_t with dx0 / sfc0 / vol0
// (mesh/geometry/geometry.h:236-245), plus the 27-entry same-level
// neighbour table that replaces the reference's ghost-exchange queues.
struct BlockGeo {
    double dx[3];      // dx0
    double sfc[3];     // sfc0 (face areas, meshgen.cpp:263-272 product order)
    double vol;        // vol0
    int nbr[27];       // block index at ((dz+1)*3 + (dy+1))*3 + (dx+1); -1 = physical side.
                       // On a refined mesh a direction whose neighbour region lies at
                       // another level points at a halo block (see SmrView).
    int wnbr[27];      // W source of each direction for the fused kernel: nbr, except that a
                       // direction crossing a physical side points at a halo block filled
                       // with the boundary functors' ghost primitives (k_phys_fill)
    int phy;           // bit (2*axis + side) set on a physical side
    int grp;           // time group: index of this block's TimeState
    // refined meshes (all zero on a same-level mesh)
    int level;         // refinement level
    int reg[3];        // logical index of the block at its level
    int rface;         // bit (2*axis + side): the face neighbour is finer (refine_flag)
    int xface;         // bit (2*axis + side): the face neighbour is at another level
    int wbmask;        // bit (3*c + r1)*3 + r2: an edge of component c whose two other
                       // coordinates lie in regions r1, r2 (0 low surface, 1 inside, 2 high
                       // surface) touches a coarser leaf (the finer side's write-back)
    signed char lev[27];  // level of the leaf covering each direction, relative
                          // (-1 coarser, 0 same, +1 finer; 9 outside the domain)
};

// Halo fill of a refined mesh.  Every direction of a block whose neighbour
// region lies at another level gets a halo block: a block-sized slot of W
// (index >= the number of real blocks) whose cells at the same local
// positions a same-level neighbour would have hold the restricted (finer
// source) or prolonged (coarser source) primitives.  Tasks are {halo
// block, owner block, direction, 0}.
struct SmrView {
    int R[3];          // root blocks per axis
    int nlev;
    const int *table;  // leaf block per (level, logical index), -1 if none
    long long toff[8]; // offset of each level's table
    int per[3];        // periodic axes
};

__host__ __device__ inline int smr_leaf(const SmrView &s, int L, const int *r) {
    if (L < 0 || L >= s.nlev) return -1;
    for (int a = 0; a < 3; ++a)
        if (r[a] < 0 || r[a] >= (s.R[a] << L)) return -1;
    return s.table[s.toff[L] + ((long long)r[2] * (s.R[1] << L) + r[1]) * (s.R[0] << L) + r[0]];
}

// Time / dt bookkeeping kept on the device so that a whole cycle (and a
// multi-cycle hipGraph) runs without host round trips.  Mirrors
// mesh/cycle/cycle.cpp:77-115 and :124-150.  One per time group: a group is
// one mesh with its own t and dt (a single mesh is group 0; an ensemble of
// independent boxes is one group per box, all advanced by the same launches).
struct TimeState {
    double t;          // simulation time
    double t_lim;
    double p_dt;       // device dt seed: min(prev * dt_expand, CFL)
    double dt_use;     // dt actually taken in the current cycle
    double dt_expand;
    int top_up;
    int skip;          // 1 once t >= t_lim at the start of a cycle
    int cycles;        // cycles actually performed
    int dt_stale;      // 1 when the CFL partials were not produced yet
};

// Device view of one mesh (all blocks share one block size).
template <class FT>
struct MeshView {
    int n[3];          // cells per block along each axis
    int ndim;          // axes with more than one cell
    int nfull;         // k_sweep<..., PROD = 2>: grid rows below run the full form, the
                       // rest the producer form
    int nv;            // variables in W and F: 5 hydro, 8 MHD
    int bnd[6];        // boundary kind per side (0 per, 1 out, 2 fre, 3 ref, 4 fixd)
    double gamma, cfl, d_floor, t_floor;
    double fix_w[5], fix_b[3];   // fixd ghost primitives and B
    size_t ncell;      // n0*n1*n2
    size_t nface;      // (n0+1)(n1+1)(n2+1): face / edge box
    size_t nfa[3];     // flux-box size for each sweep axis
    double *U;         // [block][slot 2][var 5][ncell]
    double *BF;        // [block][slot 2][comp 3][nface]          (MHD)
    FT *W;             // [block][var nv][ncell]
    FT *F[3];          // [block][var nv][nfa[a]]
    FT *EE;            // [block][comp 3][nface]                   (MHD)
    FT *EC;            // [block][comp 3][cell] cell-centred E = B x v written by c2p
                       // for the staged EMF (MHD; nullptr when only the fused kernel runs)
    const BlockGeo *geo;
    const int *blist;  // stage kernels: block of grid row y (nullptr: row y is block y)
    TimeState *ts;     // [group]
    double *dt_part;   // one CFL partial per c2p workgroup, block-major
};

}  // namespace kr
