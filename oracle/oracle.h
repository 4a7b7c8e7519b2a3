/*
 * oracle.h -- CPU FP64 restatement of Aether's FDTD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP engine
 * (include/aether_fdtd.h) is compared against, and the CPU baseline timed
 * beside it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product path never links or calls it.
 *
 * Every aeoracle_* function has the signature and meaning of the aefdtd_*
 * function of the same suffix in include/aether_fdtd.h; only the prefix and
 * the handle type differ.  aeoracle_set_threads selects the worker count of
 * the reference's thread decomposition (0 = std::thread::hardware_concurrency,
 * src/common/thread_utils.cpp:325-331).
 */
#ifndef AETHER_FDTD_ORACLE_H
#define AETHER_FDTD_ORACLE_H

#include "../include/aether_fdtd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aeoracle_grid aeoracle_grid;

void aeoracle_set_threads(int n);
int  aeoracle_get_threads(void);
/* Plants one known mistake in the oracle (code as in fdtd_oracle.cpp; 0 =
 * none), so that a test can show its pin fails on it. */
void aeoracle_set_slip(int code);
/* The full-grid material indices after set_matsgrid's extension (x fastest). */
int aeoracle_read_matsgrid(const aeoracle_grid *g, uint32_t *out);

int aeoracle_create(aeoracle_grid **out, int Nx, int Ny, int Nz, int Nt,
                    double Dx, double Dy, double Dz, double Dt, int mode,
                    const int pml[6], const int pad[6], int precision);
int aeoracle_destroy(aeoracle_grid *g);
const char *aeoracle_last_error(const aeoracle_grid *g);
int aeoracle_layout(const aeoracle_grid *g, int out[16]);
int aeoracle_set_pml(aeoracle_grid *g, int face, double kappa, double sigma, double alpha);
int aeoracle_set_tapering(aeoracle_grid *g, int Ntap);
int aeoracle_set_matsgrid(aeoracle_grid *g, const uint32_t *ids, int nx, int ny, int nz);
int aeoracle_set_material(aeoracle_grid *g, unsigned int index, const aefdtd_material *m);
int aeoracle_disable_fields(aeoracle_grid *g, const int *fields, int n);
int aeoracle_bootstrap(aeoracle_grid *g);
int aeoracle_update_E(aeoracle_grid *g);
int aeoracle_update_H(aeoracle_grid *g);
int aeoracle_reset_fields(aeoracle_grid *g);
int aeoracle_read_field(aeoracle_grid *g, int field, double *out);
int aeoracle_write_field(aeoracle_grid *g, int field, const double *in);
int aeoracle_set_plane_injection(aeoracle_grid *g, int k_plane, int nsteps,
                                 const double *dEx, const double *dEy,
                                 const double *dHx, const double *dHy);
int aeoracle_set_tfsf(aeoracle_grid *g, const aeoracle_grid *aux,
                      int i1, int i2, int j1, int j2, int k1, int k2,
                      int periodic_x, int periodic_y);
int aeoracle_add_spect_poynting(aeoracle_grid *g, int orientation,
                                int x1, int x2, int y1, int y2, int z1, int z2,
                                int Nl, const double *lambda, int *id);
int aeoracle_add_oscillator(aeoracle_grid *g, int orientation, int x1, int x2, int y1, int y2,
                            int z1, int z2, double lambda_min, double lambda_max, int *id);
int aeoracle_feed_sensors(aeoracle_grid *g);
int aeoracle_add_field_block(aeoracle_grid *g, int x1, int x2, int y1, int y2, int z1, int z2,
                             double lambda, int *id);
int aeoracle_add_field_map(aeoracle_grid *g, int orientation, int x1, int x2, int y1, int y2,
                           int z1, int z2, double lambda, int cumulative, int magnetic, int *id);
int aeoracle_field_sensor_read(aeoracle_grid *g, int id, double *out, int dims[4]);
int aeoracle_field_sensor_mats(aeoracle_grid *g, int id, uint32_t *out);
int aeoracle_sensor_poynting(aeoracle_grid *g, int id, double *rsp);
int aeoracle_sensor_spectra(aeoracle_grid *g, int id, double *out, int dims[3]);
int aeoracle_add_plane_sum(aeoracle_grid *g, int k_plane, const aefdtd_pulse *inc,
                           double z_pos, int subtract, const double *tap, int *id);
int aeoracle_read_plane_sum(aeoracle_grid *g, int id, int nsteps, double *out);
int aeoracle_pulse_init(aefdtd_pulse *p, double lambda_min, double lambda_max,
                        double inc_deg, double pol_deg, double dir_deg,
                        double startx, double starty, double startz, double n_ind);
double aeoracle_pulse_eval(const aefdtd_pulse *p, int comp,
                           double x, double y, double z, double t);

/* Material-level pins (no grid): pole time exponentials (a,b) as
 * Dielec_model::get_time_exp, RC constants and C1..C4 as RC_dielec_treat /
 * RC_recalc / const_recalc.  out: per pole (a.re,a.im,b.re,b.im,Crec.re,
 * Crec.im,chi.re,chi.im,Dchi.re,Dchi.im) then C1,C2x,C2y,C2z,C3,C4,ei,sig. */
int aeoracle_material_coeffs(const aefdtd_material *m, double Dx, double Dy,
                             double Dz, double Dt, double *out, int *n_poles);
/* Dielec_model::eval(w) -- continuous permittivity (dielec_models.cpp:52-64).*/
int aeoracle_material_eval(const aefdtd_material *m, double w, double *re, double *im);
/* CPML profile: kappa,b,c of pml_{x,y,z}_coeffs at fractional index ind on a
 * grid of N cells with pml (pm,pp) and set_pml-scaled sigma (fdtd_pml.cpp:345-447). */
int aeoracle_pml_profile(int N, int pm, int pp, double D, double Dt,
                         const double kap[2], const double sig[2], const double alp[2],
                         double ind, double out[3]);
double aeoracle_s_curve(double x, double xmin, double xmax);
int aeoracle_nearest_integer(double x);

int aeoracle_run_normal_incidence(const aefdtd_mode *m, int n_sensors,
                                  const aefdtd_sensor_req *sensors,
                                  aefdtd_normal_result *res, double *sensor_out,
                                  double *fields_out, const int *end_computation);
int aeoracle_mode_finalize(aefdtd_mode *m, int kind, int n_sensors, const aefdtd_sensor_req *sensors);
int aeoracle_run_custom(const aefdtd_mode *m, int n_sensors, const aefdtd_sensor_req *sensors,
                        double *sensor_out, double *fields_out, const int *end_computation);
int aeoracle_run_single_particle(const aefdtd_mode *m, int n_sensors,
                                 const aefdtd_sensor_req *sensors,
                                 double *sensor_out, double *fields_out, const int *end_computation);
/* Steps the last driver call ran (early stop by the energy completion check
 * included). */
int aeoracle_last_steps(void);
/* Structure::discretize, serially on the CPU (structure_oracle.cpp). */
int aeoracle_discretize(const aefdtd_structure *s, int Nx, int Ny, int Nz, double Dx, double Dy, double Dz,
                        uint32_t *matsgrid);
/* The energy sums of the last driver call with the energy completion check,
 * one per step; copies up to n into out and returns how many there are. */
int aeoracle_energy_trace(double *out, int n);
/* The spectral completion criterion of the last driver call that used it: the
 * ring length, cell count and number of checks; then the cells (i, j, k), their
 * wavelengths, the last ring [6][np][navg] and per check (step count, settled
 * cells). */
int aeoracle_ft_info(int *navg, int *np, int *nchecks);
int aeoracle_ft_read(int *pts, double *lambda, double *ring, int *checks);
/* Steps a configured grid n full steps (update_E, update_H) with no driver
 * actions: the timed CPU-baseline loop. */
int aeoracle_step(aeoracle_grid *g, int n);

#ifdef __cplusplus
}
#endif
#endif
