/* hmhd_oracle.h -- C restatement of the hosono-2DMHD step (test oracle).
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg; never by the product library. */
#ifndef HMHD_ORACLE_H
#define HMHD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_PERIODIC = 0, ORC_BC_OUTFLOW = 1 };

typedef struct {
    int nx, ny;                 /* grid_Nx, grid_Ny */
    double xbeg, xend, ybeg, yend;
    int R;                      /* GP_maxRadius (1..3) */
    double ell_over_delta;      /* GP_ellOverDelta */
    double gamma;               /* sim_gamma */
    double small_p, small_rho, small_B;
    int force_hydro;            /* sim_forceHydro */
    int bc;                     /* ORC_BC_* */
    double cfl;                 /* sim_cfl */
} orc_cfg;

typedef struct orc orc_t;

orc_t *orc_create(const orc_cfg *cfg);
void orc_destroy(orc_t *o);
void orc_dims(const orc_t *o, int *ngc, int *mx, int *my);
void orc_gp_tables(const orc_t *o, double *pv, double *qw, int *st);
void orc_set_state(orc_t *o, const double *U, const double *V);
void orc_get_state(const orc_t *o, double *U, double *V);
double orc_cfl_dt(const orc_t *o);
void orc_advance(orc_t *o, double dt);
void orc_get_stage_schemes(const orc_t *o, int *out);
void orc_get_stats(const orc_t *o, int *passes, int *inner);
void orc_forward_euler(orc_t *o, double dt, const double *U, double *Unew, int *scheme_out);
void orc_reconstruct_all(orc_t *o, const double *U, int radius, double *faceVals_out);
void orc_bc_fill(const orc_t *o, double *A, int nv);
void orc_hllc(const orc_t *o, const double *uL, const double *uR, int dir, double *flux);
void orc_cons2prim(const orc_t *o, const double *U, double *V);
void orc_prim2cons(const orc_t *o, const double *V, double *U);
void orc_cons2flux(const orc_t *o, const double *U, int dir, double *F);
double orc_cons2pres(const orc_t *o, const double *U);
double orc_eigen_cfs(double gam, double rho, double p, double BB, double BN, double pm);

#ifdef __cplusplus
}
#endif
#endif
