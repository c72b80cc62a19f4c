/* CPU FP64 oracle of the reference DUGKS step (test infrastructure only).
 * See dugks_oracle.c for the restatement and its reference citations. */
#ifndef DUGKS_ORACLE_H
#define DUGKS_ORACLE_H
#include "../include/mpdugks.h"
#ifdef __cplusplus
extern "C" {
#endif
typedef struct ora ora;
ora *ora_create(const mpd_config *cfg);
void ora_destroy(ora *o);
int ora_set_state(ora *o, const double *h, const double *f, const double *m,
                  const double *face_fx, const double *face_fy);
int ora_get_state(const ora *o, double *h, double *f, double *m, double *face_fx,
                  double *face_fy);
int ora_step(ora *o, int nsteps);
double ora_wall_flux_residual(const ora *o);
int ora_cell_phases(ora *o);
/* pseudopotential D2Q9 (periodic, CD): 1 = visit faces in the reference's
 * one-thread order, 0 = Jacobi order (default) */
int ora_set_face_order(ora *o, int ref_order);
int ora_set_one_copy(ora *o, int one);  /* test knob: UW without the second face copies */
/* face Psi of the last step [2N]; forces of the second copies of the periodic
 * faces [nx + ny] (x-faces by row, then y-faces by column) */
int ora_set_face_extra(ora *o, const double *psi_old, const double *copy_fx, const double *copy_fy);
int ora_get_face_extra(const ora *o, double *psi, double *copy_fx, double *copy_fy);
void ora_d2v16_tables(double *a, double *xi_x, double *xi_y, double *e);
void ora_d2v16_feq(double rho, double U, double V, double rt, double *feq);
#ifdef __cplusplus
}
#endif
#endif
