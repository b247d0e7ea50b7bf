/*
 * kratos_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * C interface of the CPU FP64 restatement of the reference's Godunov
 * hydro/MHD stage (PRECISION=2 semantics).  Only tests/, the smoke check
 * and the cpu_baseline leg of bench.py may load this library, and only as
 * the checker.  Nothing in the kratos_amd product links or calls it.
 *
 * State layout (identical to the reference's host dat_3d_t arrays, one RK
 * slot; mesh/types/dat_3d.h:153-156):
 *   u_aos : n0*n1*n2*nvar doubles, idx = ((k*n1 + j)*n0 + i)*nvar + v,
 *           nvar = 5 (hydro: rho, E, m0, m1, m2) or 8 (MHD: + Bcc0..2)
 *   bf_aos: N0*N1*N2*3 doubles with N[a] = n[a] + (n[a] > 1) + 4 (two
 *           ghost layers on every axis, mhd/types/fld_data.cpp:18-26),
 *           idx = (((k+2)*N1 + (j+2))*N0 + (i+2))*3 + c
 */
#ifndef KRATOS_ORACLE_H
#define KRATOS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum kro_bnd { KRO_PER = 0, KRO_OUT = 1, KRO_FRE = 2, KRO_REF = 3, KRO_FIX = 4 };

typedef struct kro_config {
    int    mhd;               /* 0: hydro PLM+HLLC, 1: MHD PLM+HLLD+CT   */
    double gamma, cfl, rho_floor, tmp_floor;
    int    n_cell_global[3];
    int    n_cell_block[3];
    double x_min[3], x_max[3];
    int    bnd[6];            /* kro_bnd per side: x-, x+, y-, y+, z-, z+ */
    double dt_init, dt_expand, t_lim;
    int    top_up;
    int    n_threads;         /* <= 0: all hardware threads               */
    double fix_w[5];          /* KRO_FIX ghost primitives rho, p, v0..2    */
    double fix_b[3];          /* KRO_FIX face / cell-centred B             */
} kro_config;

typedef struct kro_mesh kro_mesh;

kro_mesh *kro_create(const kro_config *cfg);
/* A statically refined mesh: n blocks of n_cell_block cells, block b at
 * refinement level levels[b] with logical index regs[3b..3b+2] of that
 * level (meshgen.cpp:111-158 load_refinement / generate_tree); the root
 * grid is n_cell_global / n_cell_block blocks.  3-D, 2:1 balanced.     */
kro_mesh *kro_create_blocks(const kro_config *cfg, int n, const int *levels,
                            const int *regs);
void      kro_destroy(kro_mesh *m);
const char *kro_last_error(void);

int  kro_n_blocks(const kro_mesh *m);
/* reg: logical block index (of the block's level), n: cells, xf0/dx0:
 * per-block geometry */
void kro_block_geom(const kro_mesh *m, int b, int reg[3], int n[3],
                    double xf0[3], double dx0[3]);
int  kro_set_block(kro_mesh *m, int b, const double *u_aos,
                   const double *bf_aos);
int  kro_get_block(const kro_mesh *m, int b, double *u_aos, double *bf_aos);

/* Run up to n_cycles Heun cycles, stopping early once t >= t_lim
 * (mesh/cycle/cycle.cpp:124-150, 152-194).  Returns cycles done or <0.  */
int    kro_advance(kro_mesh *m, int n_cycles);
double kro_time(const kro_mesh *m);
double kro_dt(const kro_mesh *m);          /* last dt actually taken   */
double kro_dt_next(const kro_mesh *m);     /* device-side p_dt seed    */

/* Per-face functors, exposed for unit tests (rotated frame, axis 0). */
void kro_plm(const double q[4], double *ql, double *qr);
void kro_hllc(const double wl[5], const double wr[5], double gamma,
              double flx[5]);
void kro_hlld(const double wl[8], const double wr[8], double gamma,
              double flx[8]);
/* Face flux in the mesh frame along `axis` (0..2), with the reference's
 * momentum / field rotation and the HLLD EMF slot rewrite.             */
void kro_face_flux(int mhd, int axis, const double *wl, const double *wr,
                   double gamma, double *flx);
/* Hydro c2p on one cell (hydro/eos/adiabatic.h:159-206): u is healed in
 * place, w receives the primitives; returns cs^2.                      */
double kro_c2p_hydro(double *u, double *w, double gamma, double rho_floor,
                     double tmp_floor);

#ifdef __cplusplus
}
#endif
#endif
