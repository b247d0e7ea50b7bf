/*
 * kratos_amd.h -- C-ABI of the MI355X-native Godunov hydro/MHD stage.
 *
 * Plain C: pointers, sizes and POD structs only.  Every call returns 0 on
 * success or a negative kr_status; kr_last_error() gives the message.  No
 * exception crosses this boundary.  One context per host thread; the
 * library owns all device memory and its HIP stream.
 *
 * Each entry point names the reference interface it replaces for this path
 * (paths relative to the reference's src/).
 *
 * Host state layout (the reference's dat_3d_t AoS, one RK slot,
 * mesh/types/dat_3d.h:153-156):
 *   u_aos : n0*n1*n2*nvar doubles, idx = ((k*n1 + j)*n0 + i)*nvar + v,
 *           nvar = 5 (hydro: rho, E, m0, m1, m2) or 8 (MHD: + Bcc0..2)
 *   bf_aos: N0*N1*N2*3 doubles, N[a] = n[a] + (n[a] > 1) + 4 (the faces
 *           plus two ghost layers on each side, mhd/types/fld_data.cpp
 *           :18-26: a thin, one-cell axis has a single face, both sides
 *           of the cell), idx = (((k+2)*N1 + (j+2))*N0 + (i+2))*3 + c
 *
 * Blocks may be 2-D or 1-D: the one-cell (thin) axes must be the trailing
 * ones (z, or y and z; the sweeps run over the first n_dim axes,
 * geometry.cpp:46), and for MHD one cell wide over the whole mesh
 * (mhd/integrator/integrator.h:221-227, 419-456).
 */
#ifndef KRATOS_AMD_H
#define KRATOS_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum kr_status {
    KR_OK = 0,
    KR_EINVAL = -1,     /* bad argument / configuration            */
    KR_EHIP = -2,       /* HIP runtime error                        */
    KR_ESTATE = -3,     /* call not valid in the current state      */
    KR_ENOMEM = -4
};

enum kr_model { KR_HYDRO = 0, KR_MHD = 1 };
enum kr_precision { KR_FP64 = 0, KR_MIXED = 1 };
enum kr_bnd { KR_PER = 0, KR_OUT = 1, KR_FRE = 2, KR_REF = 3, KR_FIX = 4 };

typedef struct kr_config {
    int    model;          /* kr_model: PLM+HLLC+RK2 or PLM+HLLD+CT+RK2      */
    int    precision;      /* kr_precision                                    */
    double gamma, cfl, rho_floor, tmp_floor;   /* [dynamics] keys          */
    int    bnd[6];         /* [boundary] kinds, sides x-, x+, y-, y+, z-, z+ */
    double t_0, t_lim, dt_init, dt_expand;     /* [cycle] keys             */
    int    top_up;         /* [cycle] t_lim_top_up                            */
    int    n_cell_block[3];
    int    kernel_path;    /* kr_path: KR_PATH_AUTO picks the fused stage
                              kernel (MHD; hydro on one-level meshes) when
                              the block shape allows it                     */
    int    tile[3];        /* fused-kernel column tile (x, y) and minimum
                              waves per SIMD; 0 entries are left to the
                              autotuner, which runs the first cycles with
                              one fitting variant each and keeps the
                              fastest (every variant gives the same state) */
    double fix_w[5];       /* KR_FIX (the reference's "fixd") ghost state, in
                              primitives: rho, p, v0, v1, v2 -- the [init]
                              rho0 / pre0 / vel0 of hydro/boundary/
                              keeper.cpp:60-72 */
    double fix_b[3];       /* KR_FIX face and cell-centred B ([init] b0,
                              mhd/boundary/keeper.h:126-146)              */
} kr_config;

enum kr_path { KR_PATH_AUTO = 0, KR_PATH_STAGED = 1, KR_PATH_FUSED = 2 };

typedef struct kr_block_desc {
    int    id;             /* caller's block id                              */
    int    level;          /* refinement level.  When any block has level > 0
                              the mesh is statically refined: neighbours are
                              derived from level + reg as tree_t::neighbors
                              does (mesh/tree/tree.cpp:177-197), nbr[] is not
                              read, and the mesh must be 3-D, periodic or
                              physical per axis, one time group, 2:1 balanced,
                              blocks of one even size >= 4 per axis.  Where
                              levels meet, the library fills the ghost
                              primitives by restriction / minmod prolongation,
                              corrects the coarser face fluxes and the shared
                              edge EMFs (hydro/boundary/comm_functions.h
                              :329-544, mhd/boundary/comm_functions.h
                              :467-954, hydro/integrator/integrator.h:313-382) */
    int    group;          /* time group: one mesh with its own t and dt.
                              0 for a single mesh; an ensemble of
                              independent boxes uses 0 .. n_box-1, and blocks
                              of different groups never neighbour.         */
    int    reg[3];         /* logical block index at the block's level       */
    int    n_ceff[3];      /* cells per axis: kr_config.n_cell_block, or
                              twice it on every axis -- a block that covers
                              the eight blocks reg + {0,1}^3 of its level
                              (reg is the lowest; xf0 its corner, dx0 its
                              cell).  The library runs it as those eight, so
                              a mesh with such blocks takes its topology from
                              level and reg (as a refined mesh does) and
                              gives the same cells as the mesh of eight.
                              The reference has one block size per mesh
                              (mesh/meshgen/meshgen.cpp:446-448); this is an
                              extension of the descriptor, not a counterpart. */
    double xf0[3], dx0[3]; /* geometry, as meshgen computes it               */
    int    nbr[27];        /* neighbour block id per direction
                              ((dz+1)*3 + (dy+1))*3 + (dx+1); -1 = physical  */
} kr_block_desc;

typedef struct kr_stats {
    double gpu_ms;         /* device time of the last kr_step / kr_advance   */
    int    cycles;         /* cycles performed by that call                  */
    int    n_blocks;       /* blocks on the device (a block of twice
                              n_cell_block counts as its eight)             */
    long long n_cells;     /* cells per cycle (all blocks)                   */
    double kernel_ms[8];   /* mean launch time per kernel kind (kr_profile)  */
    long long kernel_calls[8];
    double alg_bytes_per_cell_update;  /* algorithmic HBM bytes            */
    int    sweep_tile[4];  /* fused kernel in use: tile x, y, threads, waves
                              per SIMD asked (zeros: staged path)           */
    float  tune_ms[16];    /* autotuner: mean stage-kernel time per variant
                              over its tuning cycle (0: not timed)          */
    int    tune_pending;   /* variants the autotuner has still to time: the
                              next that many cycles run eagerly, one
                              variant each                                  */
    int    graph_ready;    /* 1 once the cycle's hipGraph is instantiated
                              (kr_advance then only replays it)            */
    int    sweep_kernel;   /* the fused MHD kernel in use: 0 k_sweep, 1 the
                              mixed form's k_sweep_mx                       */
} kr_stats;

enum kr_kernel_kind {
    KR_K_C2P = 0, KR_K_DTFIN = 1, KR_K_FLUX_X = 2, KR_K_FLUX_Y = 3,
    KR_K_FLUX_Z = 4, KR_K_EMF = 5, KR_K_UPDATE = 6, KR_K_BEGIN = 7
};
/* In the fused path the whole flux / EMF / update / CT stage is one launch,
 * reported under KR_K_UPDATE.                                             */

typedef struct kr_ctx kr_ctx;

/* Create a context.  Replaces the policy enrolment and reads of
 * dynamics::hydro_t / mhd_t::enroll<prx,itg,eos,rec,rie,com>()
 * (modules/dynamics/hydro/hydro.h:28-57, mhd/mhd.h:25-58,
 * dynamics/dynamics.h:51-103) and the eos / cycle key reads
 * (prototypes/eos.h:43-56, mesh/cycle/cycle.cpp:30-65).                  */
int kr_create(const kr_config *cfg, kr_ctx **out);

/* Register blocks.  Replaces mod_base_t::update's per-block block_yield +
 * dual_t::setup (mesh/module_base/module_base.cpp:89-125,
 * mesh/block/dual_data.cpp:30-41) and the neighbour discovery that feeds
 * the boundary keeper (mesh/tree/tree.cpp:177-197).                      */
int kr_add_blocks(kr_ctx *ctx, int n, const kr_block_desc *blocks);

/* Copy one block's state (slot 0) host -> device.  Replaces
 * dual_t::copy_h2d (mesh/block/dual_data.cpp:46-49 ->
 * hydro/types/hyd_data.cpp:59-63, mhd/types/fld_data.cpp:97-101).
 * bf_aos is ignored for hydro.  Arrays are of the block's own size
 * (n_ceff), also for a block of twice n_cell_block.                      */
int kr_upload(kr_ctx *ctx, int id, const double *u_aos, const double *bf_aos);

/* Copy one block's state (slot 0) device -> host.  Replaces the D2H of
 * dual_t::write (mesh/block/dual_data.cpp:68-79).  For MHD, u's Bcc slots
 * are returned as the face average of bf (the value the next c2p writes,
 * mhd/eos/adiabatic.h:204-212); bf ghost layers are returned as zero.    */
int kr_download(kr_ctx *ctx, int id, double *u_aos, double *bf_aos);

/* One Heun cycle if t < t_lim: c2p + CFL dt, halo, fluxes, EMF/CT, RK2
 * slots for both substeps.  Replaces cycle::step -> mesh_t::step ->
 * dynamics::base_t::step (mesh/cycle/cycle.cpp:124-150,
 * modules/dynamics/dynamics.cpp:37-42, dynamics.h:60-100).
 * *dt_taken (may be NULL) receives the dt used, 0 if no cycle ran.       */
int kr_step(kr_ctx *ctx, double *dt_taken);

/* Up to n_cycles cycles, stopping once t >= t_lim, as the loop of
 * cycle::evolve (mesh/cycle/cycle.cpp:152-194), replayed from one
 * captured hipGraph.  *n_done (may be NULL) receives the cycles run.     */
int kr_advance(kr_ctx *ctx, int n_cycles, int *n_done);

/* Time and dt: cycle::dt() / t (mesh/cycle/cycle.cpp:117-122).           */
int kr_get_time(kr_ctx *ctx, double *t, double *dt_last, double *dt_seed);
/* Set t and the device dt seed of every time group (restart /
 * cycle::init's dt_init).                                                 */
int kr_set_time(kr_ctx *ctx, double t, double dt_seed);

/* t, the last dt taken and the cycles run by one time group: the per-mesh
 * cycle::t / dt() (mesh/cycle/cycle.cpp:117-122) of one ensemble member,
 * each member having its own global-dt reduction (cycle.cpp:91-115).
 * kr_get_time reports group 0; kr_step / kr_advance count a cycle when
 * any group ran it, and kr_advance stops once every group reached t_lim.  */
int kr_get_group_time(kr_ctx *ctx, int group, double *t, double *dt_last, int *cycles);

/* Max |div B| over all cells, slot 0 (the CT identity probe of
 * mhd/integrator/ct_id_dbg.h:10-40, as a diagnostic).                    */
int kr_divb_max(kr_ctx *ctx, double *divb);

/* Run n_cycles with HIP events around every launch and record the mean
 * launch time per kernel kind into the stats (instrumented, not timed
 * by kr_advance).                                                        */
int kr_profile(kr_ctx *ctx, int n_cycles);

/* Statistics of the last call; timing of the last kr_step/kr_advance.
 * Replaces the Speed print of cycle::evolve (mesh/cycle/cycle.cpp:184-190). */
int kr_get_stats(kr_ctx *ctx, kr_stats *out);

/* Block until all queued device work of the context has finished.       */
int kr_sync(kr_ctx *ctx);

/* Max |div B| of n_blocks blocks whose face fields the caller holds in
 * device memory, in kr_upload's bf_aos layout (one slot with its ghost
 * layers, block after block); dx_dev[3*b + a] (device memory) is block b's
 * cell size along a.  The result is written to *divb_dev (device memory),
 * in stream order on `stream` (a hipStream_t; NULL: the default stream),
 * with the expression of kr_divb_max, so both give the same value.
 * No counterpart in the reference: an allocation-free variant of
 * kr_divb_max for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops).  It calls no hipMalloc / hipFree: its
 * per-workgroup partial maxima live in a static device array of 65,536
 * doubles (512 KiB per GPU, held until the process ends), so n_blocks is
 * at most 65,536 (KR_EINVAL beyond, reason in kr_last_error(NULL)).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (two streams, two threads) race and must be ordered by the caller.    */
int kr_divb_max_noalloc(const double *bf_dev, int n_blocks, const int n_cell_block[3],
                        const double *dx_dev, double *divb_dev, void *stream);

/* Message of the last failed call on ctx; with ctx == NULL, the reason the
 * last kr_create or kr_divb_max_noalloc on this thread failed (the
 * reference throws std::runtime_error with such a text,
 * device/hip/hip.cpp:8-16).                                              */
const char *kr_last_error(kr_ctx *ctx);
void kr_destroy(kr_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* KRATOS_AMD_H */
