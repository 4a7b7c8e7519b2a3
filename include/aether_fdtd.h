/*
 * aether_fdtd.h -- C ABI of the MI355X (gfx950) FDTD engine.
 *
 * The engine is a drop-in for Aether's 3D/2D Yee solver
 * (src/finite_differences/FDTD).  Every entry point below replaces one
 * member of Aether's C++ `class FDTD` (src/finite_differences/FDTD/core/
 * fdtd_core.h:26-285), one per-step action that Aether's mode drivers perform
 * by touching `fdtd.Ex(i,j,k)` directly between `update_E()` and `update_H()`
 * (those actions move onto the device so that the fields stay resident in
 * HBM), or one mode driver (modes/lua_fdtd.h:91-108).
 *
 * Conventions
 *  - Plain C types only; no exceptions cross the ABI.  Every function returns
 *    an int status (AEFDTD_OK == 0).  aefdtd_last_error() gives the text of
 *    the last failure on a handle (or of the last failed create when called
 *    with NULL).
 *  - Grids are x-fastest: value (i,j,k) lives at i + j*Nx + k*Nx*Ny, as in
 *    Aether's Grid3 (src/common/grid.h:868-877).
 *  - Indices of faces: 0 xm, 1 xp, 2 ym, 3 yp, 4 zm, 5 zp.
 *  - Field ids: 0 Ex, 1 Ey, 2 Ez, 3 Hx, 4 Hy, 5 Hz (Aether's EX_FIELD..HZ_FIELD
 *    order, src/common/enum_constants.h).
 *  - A handle is single-threaded: the caller serialises calls on one handle.
 *    Distinct handles may be driven from distinct host threads.
 *  - The engine owns all device memory.  Inputs are copied; outputs are
 *    written to caller buffers.
 */
#ifndef AETHER_FDTD_H
#define AETHER_FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ----------------------------------------------------- */
#define AEFDTD_OK          0
#define AEFDTD_EINVAL      1   /* bad argument or dimension                   */
#define AEFDTD_ENOMEM      2   /* device allocation failed / over HBM budget  */
#define AEFDTD_EDEVICE     3   /* HIP runtime error, or no usable gfx950 GPU  */
#define AEFDTD_ESTATE      4   /* call out of order (e.g. step before setup)  */
#define AEFDTD_ECANCELLED  5   /* run stopped by its cancel flag              */
#define AEFDTD_ERANGE      6   /* a field left the FP16 storage range         */

/* ---- grid modes (the `smod` string of FDTD::FDTD, fdtd_core.cpp:198-204) */
#define AEFDTD_MODE_NORM    0  /* "NORM"   : normal-incidence slab driver     */
#define AEFDTD_MODE_CUSTOM  5  /* "CUSTOM" : single-particle / custom drivers */

/* ---- storage precision of the fields ------------------------------------ */
#define AEFDTD_FP64  0     /* FP64 fields and arithmetic (parity form)         */
#define AEFDTD_FP32  1     /* FP32 fields/arithmetic; FP64 RC poles and DFT    */
#define AEFDTD_FP16  2     /* FP16 field storage with per-component 2^e scale, */
                           /* FP32 arithmetic and CPML state, FP64 poles/DFT  */

/* ---- sensor orientations (Aether's NORMAL_X.. enum order) --------------- */
#define AEFDTD_NORMAL_X   0
#define AEFDTD_NORMAL_XM  1
#define AEFDTD_NORMAL_Y   2
#define AEFDTD_NORMAL_YM  3
#define AEFDTD_NORMAL_Z   4
#define AEFDTD_NORMAL_ZM  5
#define AEFDTD_NORMAL_UNDEF 6

/* A material as Aether's FDTD sees it: eps_inf plus pole lists.
 * Mirrors the FDTD-compatible part of `Material` (src/materials/material.h)
 * and the setters of Dielec_model (src/materials/dielec_models.cpp:233,297,
 * 365,430).  Arrays hold n_* consecutive tuples. */
typedef struct aefdtd_material {
    double        eps_inf;
    int           n_debye;     const double *debye;      /* (ds, t0)       */
    int           n_drude;     const double *drude;      /* (wd, g)        */
    int           n_lorentz;   const double *lorentz;    /* (A, O, G)      */
    int           n_critpoint; const double *critpoint;  /* (A, O, P, G)   */
} aefdtd_material;

/* Analytic Gaussian plane-wave pulse; the state of Aether's ChpIn after
 * init_pulse (core/chpin.cpp:479-545). Filled by aefdtd_pulse_init. */
typedef struct aefdtd_pulse {
    double lambda, omega, k, kx, ky, kz, n_ind, dz, C1;
    double Ex0, Ey0, Ez0, Hx0, Hy0, Hz0;
} aefdtd_pulse;

typedef struct aefdtd_grid aefdtd_grid;   /* opaque handle == one `FDTD` */

/* ======================================================================== */
/*  Grid object: replaces `class FDTD` (core/fdtd_core.h:26-285)            */
/* ======================================================================== */

/* FDTD::FDTD(Nx,..,pad_zp) -- core/fdtd_core.cpp:158-285.  Applies Aether's
 * padding rules (:212-243) and grid extension (:245-251).  pml[6]/pad[6] in
 * face order.  `precision` is AEFDTD_FP64, AEFDTD_FP32 or AEFDTD_FP16. */
int aefdtd_create(aefdtd_grid **out, int Nx, int Ny, int Nz, int Nt,
                  double Dx, double Dy, double Dz, double Dt, int mode,
                  const int pml[6], const int pad[6], int precision);
int aefdtd_destroy(aefdtd_grid *g);
const char *aefdtd_last_error(const aefdtd_grid *g);

/* Layout after extension: out[0..2]=Nx,Ny,Nz (full grid); out[3..8]=xs_s,xs_e,
 * ys_s,ys_e,zs_s,zs_e; out[9..14]=pad_xm..pad_zp; out[15]=tstep.
 * Mirrors the public members of FDTD (fdtd_core.h:35-44,75-77). */
int aefdtd_layout(const aefdtd_grid *g, int out[16]);

/* FDTD::set_pml_xm..zp -- core/fdtd_pml.cpp:457-497 (sigma scaled there). */
int aefdtd_set_pml(aefdtd_grid *g, int face, double kappa, double sigma, double alpha);
/* FDTD::set_tapering -- core/fdtd_utils.cpp:915-918. */
int aefdtd_set_tapering(aefdtd_grid *g, int Ntap);
/* FDTD::set_matsgrid -- core/fdtd_utils.cpp:572-796: copies the structure
 * grid (nx,ny,nz == the Nx,Ny,Nz given to create) and extends its edge
 * planes into pads and PML. */
int aefdtd_set_matsgrid(aefdtd_grid *g, const uint32_t *ids, int nx, int ny, int nz);
/* FDTD::set_material -- core/fdtd_utils.cpp:810-820 (no-op for an index
 * absent from the grid, as mats_in_grid). Derives C1..C4 and the recursive-
 * convolution pole constants (core/mats_RC.cpp:126-184, mats_const.cpp:18-80).*/
int aefdtd_set_material(aefdtd_grid *g, unsigned int index, const aefdtd_material *m);
/* FDTD::disable_fields -- core/fdtd_core.cpp:1150-1161. */
int aefdtd_disable_fields(aefdtd_grid *g, const int *fields, int n);
/* FDTD::bootstrap -- core/fdtd_core.cpp:1095-1135 (reallocates CPML state,
 * recomputes profiles). Also uploads everything set so far to the device. */
int aefdtd_bootstrap(aefdtd_grid *g);
/* FDTD::update_E / update_H -- core/fdtd_core.cpp:1163-1218, 1316-1335.
 * One fused device pass each: RC-ante, E stencil, [TFSF-E], RC-apply, CPML-E,
 * PEC, [plane injection]  /  H stencil, CPML-H, PEC, [plane injection],
 * [TFSF-H].  Bracketed actions run only when registered below, in the place
 * where Aether's drivers perform them. update_H advances tstep. */
int aefdtd_update_E(aefdtd_grid *g);
int aefdtd_update_H(aefdtd_grid *g);
/* n full steps: update_E then update_H, as the drivers' loops call them
 * (normal_incidence.cpp:330-357, default_fdtd.cpp:148-160), with the
 * registered per-step actions.  On a 3D grid up to 1024 cells wide (and not
 * fed by a TFSF column, which the drivers step with it) one step is one pass
 * over the grid that updates E and then H of the planes behind it, with the
 * same per-cell arithmetic as the two calls (bit-identical results); other
 * grids make the two calls.  No counterpart in the reference, whose callers
 * always make the two calls. */
int aefdtd_step(aefdtd_grid *g, int n);
/* FDTD::reset_fields -- core/fdtd_utils.cpp:855-881. */
int aefdtd_reset_fields(aefdtd_grid *g);

/* Host access to the public Grid3<double> Ex..Hz (fdtd_core.h:62). Full
 * extended grid, x-fastest, converted to/from double. Synchronous. */
int aefdtd_read_field(aefdtd_grid *g, int field, double *out);
int aefdtd_write_field(aefdtd_grid *g, int field, const double *in);

/* ---- per-step actions of the mode drivers, moved onto the device ------- */

/* Plane injection of normal_incidence.cpp:334-357 / single_particle.cpp:
 * 270-275,343-348: after each update_E, Ex(:,:,k)+=dE[0][t], Ey(:,:,k)+=dE[1][t];
 * after each update_H, Hx(:,:,k)+=dH[0][t], Hy(:,:,k)+=dH[1][t], t = tstep.
 * Tables hold nsteps values each and are copied to the device. */
int aefdtd_set_plane_injection(aefdtd_grid *g, int k_plane, int nsteps,
                               const double *dEx, const double *dEy,
                               const double *dHx, const double *dHy);

/* Total-field/scattered-field box of single_particle.cpp:277-332 (E) and
 * :350-378 (H), fed by the 1D auxiliary grid `aux` advanced in lockstep
 * (the caller steps aux.update_E before g.update_E and aux.update_H before
 * g.update_H, as the reference does). Box in full-grid indices. */
int aefdtd_set_tfsf(aefdtd_grid *g, const aefdtd_grid *aux,
                    int i1, int i2, int j1, int j2, int k1, int k2,
                    int periodic_x, int periodic_y);

/* Oscillator source: the object generate_fdtd_source builds for
 * SOURCE_GEN_OSCILLATOR (sources/sources.cpp:323-348), Oscillator
 * (sources/sc_oscillator.cpp:20-84).  After every update_E, the field
 * normal to `orientation` (NORMAL_X/XM -> Ex, Y -> Ey, Z -> Ez; NORMAL_UNDEF
 * -> nothing) gets Re(e^{-i w0 s Dt} e^{-(dw/2 (s-tshift) Dt)^2})*50 added on
 * [x1,x2)x[y1,y2)x[z1,z2), s = update_H calls since this call (Source::step,
 * sources.cpp:240-249).  Location in structure cells, offset and clamped as
 * Source::link (sources.cpp:251-283); w0, dw, tshift from lambda_min/max as
 * Oscillator::initialize.  At most 8 per grid.  Writes the source id. */
int aefdtd_add_oscillator(aefdtd_grid *g, int orientation, int x1, int x2, int y1, int y2,
                          int z1, int z2, double lambda_min, double lambda_max, int *id);

/* Planar spectral Poynting sensor, Spect_Poynting (sensors/s_spec_power.cpp:
 * 32-119) over SensorFieldHolder with interpolation (sensors/sensors.cpp:
 * 423-644). Location in structure cells, offset and clamped as Sensor::link
 * (sensors.cpp:196-230). lambda[Nl] copied. Writes the sensor id. */
int aefdtd_add_spect_poynting(aefdtd_grid *g, int orientation,
                              int x1, int x2, int y1, int y2, int z1, int z2,
                              int Nl, const double *lambda, int *id);
/* Sensor::feed for every registered sensor (sensors.cpp:232-243): one
 * interpolated plane sample + running DFT, tapering as Aether. */
int aefdtd_feed_sensors(aefdtd_grid *g);
/* Spect_Poynting::treat -- rsp_result[l] for l < Nl (s_spec_power.cpp:44-119).*/
int aefdtd_sensor_poynting(aefdtd_grid *g, int id, double *rsp);
/* aefdtd_sensor_poynting without device allocations.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-wavelength staging buffer is a static device array of
 * AEFDTD_NOALLOC_MAX_NL doubles per GPU instead of a hipMalloc per call; a
 * sensor with more wavelengths is refused (AEFDTD_EINVAL), and so is a call
 * whose pending DFT flush would need a larger coefficient table (AEFDTD_ESTATE:
 * stepping a grid past the Nt it was created with).  Every call on one GPU
 * shares that array: calls that overlap on one GPU race (serialise them). */
#define AEFDTD_NOALLOC_MAX_NL 4096
int aefdtd_sensor_poynting_noalloc(aefdtd_grid *g, int id, double *rsp);
/* Single-wavelength field sensors (sensors/s_fieldmap.cpp), location in
 * structure cells, offset and clamped as Sensor::link.  Every feed adds the
 * interpolated fields (local_E*, local_H*) times the DFT phase to FP64 complex
 * accumulators, without tapering:
 *  FieldBlock (:89-230): the box; E at e^{i w step Dt}, H at e^{i w (step+1/2) Dt}.
 *  FieldMap (:321-470): the first plane of the box normal to `orientation`
 *  (the - orientations fold to +); E (H with `magnetic`) at e^{i w step Dt};
 *  with `cumulative`, the values summed along the normal over the box depth.
 * Each writes its id (ids count field sensors separately from the spectral
 * ones). */
int aefdtd_add_field_block(aefdtd_grid *g, int x1, int x2, int y1, int y2, int z1, int z2,
                           double lambda, int *id);
int aefdtd_add_field_map(aefdtd_grid *g, int orientation, int x1, int x2, int y1, int y2,
                         int z1, int z2, double lambda, int cumulative, int magnetic, int *id);
/* Accumulators of field sensor id: out[c][q] complex as (re,im), c over the
 * ncomp = dims[0] components (block: Ex..Hz; map: Ex,Ey,Ez or Hx,Hy,Hz), q =
 * a + s1 (b + s2 c) over dims[1..3] = s1, s2, s3 (block: the spans along x, y,
 * z; map: the in-plane spans, x then y for a z map, y then z for an x map, x
 * then z for a y map, and 1).  out may be NULL to read dims only. */
int aefdtd_field_sensor_read(aefdtd_grid *g, int id, double *out, int dims[4]);
/* The material indices under those cells (FieldBlock::mats, FieldMap::mats). */
int aefdtd_field_sensor_mats(aefdtd_grid *g, int id, uint32_t *out);
/* Raw DFT accumulators sp_Ex..sp_Hz of one sensor: out[c][l][j][i] complex
 * as (re,im), c in Ex,Ey,Ez,Hx,Hy,Hz; span1 x span2 written to dims[0..1]. */
int aefdtd_sensor_spectra(aefdtd_grid *g, int id, double *out, int dims[3]);

/* Reflection/transmission plane sums of normal_incidence.cpp:378-415: each
 * feed appends sum_ij (F(i,j,k)-s*inc)*tap and sum_ij inc*tap for F=Ex,Ey,Ez,
 * where inc is the analytic pulse at (i*Dx(+Dx/2), j*Dy(+Dy/2), z_pos(+Dz/2))
 * and time (tstep-1+0.5)*Dt, s = subtract (1 for the reflection plane, 0 for
 * the transmission plane). tap[Nt] is copied. */
int aefdtd_add_plane_sum(aefdtd_grid *g, int k_plane, const aefdtd_pulse *inc,
                         double z_pos, int subtract, const double *tap, int *id);
/* Reads nsteps x 6 doubles: (sumEx-s*inc, sumEy-s*inc, sumEz-s*inc, incX, incY, incZ). */
int aefdtd_read_plane_sum(aefdtd_grid *g, int id, int nsteps, double *out);

/* ChpIn::init_pulse + setfields (core/chpin.cpp:479-545). Host only. */
int aefdtd_pulse_init(aefdtd_pulse *p, double lambda_min, double lambda_max,
                      double inc_deg, double pol_deg, double dir_deg,
                      double startx, double starty, double startz, double n_ind);
/* ChpIn::PulseEx..Hz (core/chpin.cpp:44-90); comp 0..5 = Ex..Hz. Host only. */
double aefdtd_pulse_eval(const aefdtd_pulse *p, int comp,
                         double x, double y, double z, double t);

/* ======================================================================== */
/*  Mode drivers: replace `void FDTD_normal_incidence(FDTD_Mode const&,..)` */
/*  and `FDTD_single_particle` (modes/lua_fdtd.h:91-108)                    */
/* ======================================================================== */

/* The FDTD_Mode fields the drivers read (modes/lua_fdtd.h:22-89,
 * core/lua_fd.h:51-150) plus the discretised structure. */
typedef struct aefdtd_mode {
    int     Nt;
    double  Dx, Dy, Dz;
    double  time_mod;
    int     tapering;
    int     pml[6];              /* after FDTD_Mode::finalize              */
    int     pad[6];
    double  kappa[6], sigma[6], alpha[6];
    int     Nl;
    double  lambda_min, lambda_max;
    int     polarization;        /* 0 "TE", 1 "TM", 2 "mix"                */
    double  polar_angle;         /* degrees, used by "mix"                 */
    int     n_disable;  const int *disable_fields;
    int     nx, ny, nz;  const uint32_t *matsgrid;   /* Structure::discretize */
    int     n_materials; const aefdtd_material *materials;
    int     precision;           /* AEFDTD_FP64 / AEFDTD_FP32 / AEFDTD_FP16 */
    int     n_sources;  const struct aefdtd_source_req *sources;   /* custom mode */
    /* auto_tsteps (FDTD_Mode::set_auto_tsteps, modes/lua_fdtd.cpp:271-294):
     * time_type 0 runs all Nt steps; 1 stops early with the energy
     * CompletionSensor, 2 with its spectral (FT) criterion
     * (sensors/s_completion.cpp:26-36,39-130,132-163,232-398), checked every
     * cc_step steps.  The FT criterion watches cc_quant points placed on the
     * structure region's faces by the 3-letter cc_layout (per axis n none,
     * d low face, u high face, b both; "v" the volume; unknown or "nnn" become
     * "nnb"), at wavelengths drawn in [cc_lmin, cc_lmax]; placement and
     * wavelengths are drawn from std::mt19937 seeded with cc_seed as the
     * reference draws them from its generator after rng_seed(cc_seed). */
    int     time_type;
    int     cc_step;
    double  cc_coeff;
    double  cc_lmin, cc_lmax;
    int     cc_quant;
    char    cc_layout[4];
    unsigned int cc_seed;
} aefdtd_mode;

/* A source request (Source_generator, core/sources.h:29-59, already
 * to_discrete'd: structure cells).  type 0 = oscillator. */
typedef struct aefdtd_source_req {
    int    type;
    int    orientation;
    int    x1, x2, y1, y2, z1, z2;
    double lambda_min, lambda_max;
} aefdtd_source_req;


/* A box or planar spectral Poynting sensor request (Sensor_generator,
 * sensors/sensors.cpp:27-161, already to_discrete'd: structure cells). */
typedef struct aefdtd_sensor_req {
    int    box;                  /* 1: box_spectral_poynting, 0: planar   */
    int    orientation;          /* planar only                           */
    int    x1, x2, y1, y2, z1, z2;
    int    disable[6];           /* box faces to skip                     */
    int    Nl;
    double lambda_min, lambda_max;
    int    kind;                 /* AEFDTD_SENSOR_*                       */
} aefdtd_sensor_req;

/* Sensor request kinds (Sensor_type, sensors/lua_sensors.cpp:126-147) and
 * what aefdtd_run_finish writes to sensor_out for each, in request order:
 *   POYNTING  box_spectral_poynting (box=1): Nl totals, then 6 x Nl per face;
 *             planar_spectral_poynting (box=0): Nl values
 *   FIELDBLOCK  fieldblock at lambda_min: 6 x s1 s2 s3 complex (re,im), the
 *             layout of aefdtd_field_sensor_read (s = the box extents)
 *   FIELDMAP  fieldmap at lambda_min, normal to `orientation`: 3 x s1 s2
 *             complex, E (as the reference's scripts create it)
 *   FIELDMAP2 fieldmap2: the spectral accumulators of a planar sensor with Nl
 *             wavelengths, 6 x Nl x s1 s2 complex (aefdtd_sensor_spectra)
 * Field sensors must lie inside the grid (their extents are not clamped). */
/* Doubles aefdtd_run_finish and the one-call drivers write to sensor_out for
 * request q (the sizes above; field sensors after their extents are linked to
 * the grid, as given). */
size_t aefdtd_sensor_out_len(const aefdtd_sensor_req *q);
#define AEFDTD_SENSOR_POYNTING   0
#define AEFDTD_SENSOR_FIELDBLOCK 1
#define AEFDTD_SENSOR_FIELDMAP   2
#define AEFDTD_SENSOR_FIELDMAP2  3

/* The sizing part of FDTD_Mode::finalize (modes/lua_fdtd.cpp:74-203) for
 * mode type `kind` (0 normal incidence, 1 single particle, 2 custom): normal
 * incidence drops the x/y PML and defaults the z PML to 25 cells; on every
 * axis longer than one cell a PML face gets a pad of at least 5 cells, more
 * when a sensor (5) or an oscillator source (50) comes near it, and a face
 * without PML gets no pad.  Updates m->pml and m->pad in place. */
int aefdtd_mode_finalize(aefdtd_mode *m, int kind, int n_sensors, const aefdtd_sensor_req *sensors);

/* ---------------------------------------------------------------------------
 * Structures (src/structure).  A structure is the operation list that a
 * structure script builds through the Lua functions of
 * structure/lua_structure.cpp, plus its default material, nominal size,
 * flip and loop (periodic) modifiers.  Parameters are as the script passes
 * them, in metres:
 *   BLOCK     p = x1 x2 y1 y2 z1 z2            add_block (lua_structure.cpp:292-310)
 *   SPHERE    p = x y z r                      add_sphere (:531-547)
 *   CYLINDER  p = x y z dx dy dz r             add_cylinder (:376-397): axis from
 *                                              (x,y,z) to (x+dx, y+dy, z+dz)
 *   CONE      p = x y z dx dy dz r             add_cone (:334-354): radius r at
 *                                              (x,y,z), 0 at the tip
 *   ELLIPSOID p = x y z rx ry rz               add_ellipsoid (:399-418)
 *   LAYER     p = axis z1 z2                   add_layer (:420-440): axis 0 "X",
 *                                              1 "Y", 2 "Z"
 *   COATING   p = theta phi thickness delta origin
 *                                              add_coating (:314-331,
 *             structure_primitives.cpp:45-96): a point of material `origin`
 *             (as the operations before this one make it) takes mat_index when,
 *             stepping by delta along the deposition direction (theta, phi in
 *             degrees; dz = -cos theta), a point of another material lies
 *             within thickness, and stepping the other way up to lz meets only
 *             `origin`.  Needs delta > 0 and theta below 90 degrees (the
 *             reference's upward march never ends otherwise); the operations
 *             before a coating may not include another coating here.
 * Conformal coatings, meshes, sinusoidal layers, vector blocks/triangles and
 * Lua defined operations are not provided.
 * ------------------------------------------------------------------------- */
enum { AEFDTD_OP_BLOCK = 0, AEFDTD_OP_SPHERE = 1, AEFDTD_OP_CYLINDER = 2, AEFDTD_OP_CONE = 3,
       AEFDTD_OP_ELLIPSOID = 4, AEFDTD_OP_LAYER = 5, AEFDTD_OP_COATING = 6 };

typedef struct aefdtd_structure_op {
    int    type;        /* AEFDTD_OP_* */
    int    mat_index;
    double p[7];
} aefdtd_structure_op;

typedef struct aefdtd_structure {
    int n_ops;  const aefdtd_structure_op *ops;   /* script order: later ones win */
    int    default_material;                      /* default_material(index) */
    double lx, ly, lz;                            /* the script's lx, ly, lz */
    int    flip[3];                               /* flip(x, y, z) */
    int    loop[3];                               /* loop(x, y, z) */
} aefdtd_structure;

/* Structure::discretize (structure/structure.cpp:94-113, with
 * Structure::index, :223-262): cell (i,j,k) of the Nx x Ny x Nz grid takes
 * the material of the last operation containing the point (i Dx, j Dy, k Dz)
 * (flipped, then folded into [0, l) on looped axes and tested against the
 * operation and its +-l images), else the default material.  Evaluated on
 * the device; matsgrid (Nx*Ny*Nz values, x fastest) is a host buffer.
 * Results are identical to the reference's: the point tests use the same
 * FP64 operations in the same order, without fused multiply-adds. */
int aefdtd_discretize(const aefdtd_structure *s, int Nx, int Ny, int Nz, double Dx, double Dy, double Dz,
                      uint32_t *matsgrid);

/* Results of the normal-incidence driver (normal_incidence.cpp:442-570):
 * arrays of Nl. R,T are the `spectdata_norm` columns; refx..transy the
 * complex amplitudes written to `spectdata_norm2` (as re,im pairs). */
typedef struct aefdtd_normal_result {
    double *lambda, *R, *T;
    double *refx, *refy, *transx, *transy;   /* 2*Nl each */
    double  index_sup, index_sub;
    int     Nt_run;
} aefdtd_normal_result;

/* FDTD_normal_incidence (modes/normal_incidence.cpp:161-607) on the device.
 * sensor_out: per sensor Nl doubles of rsp_result (box: total), may be NULL.
 * fields_out: if non-NULL, 6 x (full grid) doubles of the final fields.
 * end_computation (may be NULL): the driver's `std::atomic<bool>
 * *end_computation` (lua_fdtd.h:91-108), here an int the caller may set
 * non-zero from another thread; it is read after every step, and a set flag
 * ends the run after that step (Nt = t, normal_incidence.cpp:431-435), which
 * is then post-processed as usual; the call returns AEFDTD_ECANCELLED. */
int aefdtd_run_normal_incidence(const aefdtd_mode *m, int n_sensors,
                                const aefdtd_sensor_req *sensors,
                                aefdtd_normal_result *res, double *sensor_out,
                                double *fields_out, const int *end_computation);

/* FDTD_single_particle (modes/single_particle.cpp:29-415) on the device:
 * TFSF box fed by the 1D auxiliary grid. sensor_out as above; for a box
 * sensor it holds Nl totals followed by 6*Nl per-face values (xm..zp).
 * end_computation as above (single_particle.cpp:403: the sensors keep the
 * steps made, including the one during which the flag was set). */
int aefdtd_run_single_particle(const aefdtd_mode *m, int n_sensors,
                               const aefdtd_sensor_req *sensors,
                               double *sensor_out, double *fields_out,
                               const int *end_computation);

/* mode_default_fdtd (modes/default_fdtd.cpp:24-210), the custom mode, on the
 * device: m->sources (oscillators) and the sensor requests, loop update_E,
 * sources' inject_E, update_H, sensors' feed, sources' inject_H for Nt steps
 * (fixed-time runs).  sensor_out and end_computation as above
 * (default_fdtd.cpp:201). */
int aefdtd_run_custom(const aefdtd_mode *m, int n_sensors, const aefdtd_sensor_req *sensors,
                      double *sensor_out, double *fields_out, const int *end_computation);

/* The same drivers split into create / step / finish, so that a caller can
 * interleave several runs (serving) or time the stepping alone. kind 0 =
 * normal incidence, 1 = single particle, 2 = custom. run_step advances up to n steps
 * (never past Nt); run_finish performs the driver's post-processing (res is
 * used by kind 0 only and may be NULL for kind 1). */
typedef struct aefdtd_run aefdtd_run;
int aefdtd_run_create(aefdtd_run **out, int kind, const aefdtd_mode *m,
                      int n_sensors, const aefdtd_sensor_req *sensors);
int aefdtd_run_step(aefdtd_run *r, int n);
/* Cancels this run (and no other): the step in progress, or the next one if
 * none is, is its last (the end_computation poll of normal_incidence.cpp:
 * 431-435 on this run only); aefdtd_run_step then returns AEFDTD_ECANCELLED
 * and aefdtd_run_finish post-processes the steps made.  Safe to call from
 * another host thread. */
int aefdtd_run_cancel(aefdtd_run *r);
int aefdtd_run_progress(const aefdtd_run *r, int *t, int *Nt);
aefdtd_grid *aefdtd_run_grid(aefdtd_run *r);
const char *aefdtd_run_error(const aefdtd_run *r);
int aefdtd_run_finish(aefdtd_run *r, aefdtd_normal_result *res,
                      double *sensor_out, double *fields_out);
int aefdtd_run_destroy(aefdtd_run *r);
/* The result files of a finished run, in the reference's formats, into
 * `directory` (call after aefdtd_run_finish, before aefdtd_run_destroy, with
 * what finish returned):
 *   planar sensor q:  <names[q]>_pspdft  "lambda flux flux/area", 15 significant
 *                     digits (Spect_Poynting::treat, sensors/s_spec_power.cpp:44-118);
 *   box sensor q:     <names[q]>_bspdft  "lambda total xm xp ym yp zm zp" (enabled
 *                     faces only), 15 digits, and the Octave reader <names[q]>.m
 *                     (Box_Spect_Poynting::treat, s_spec_power.cpp:174-232);
 *   normal incidence: <prefix>spectdata_norm "lambda R T" and <prefix>spectdata_norm2
 *                     "lambda |rx| arg(rx) Prx |ry| arg(ry) Pry |tx| arg(tx) Ptx |ty|
 *                     arg(ty) Pty", 6 significant digits
 *                     (modes/normal_incidence.cpp:490-570).
 * names: one per sensor request (NULL: "sensor<q>"); sensor_out and res may be
 * NULL to skip the sensor or the normal-incidence files. */
int aefdtd_run_write(aefdtd_run *r, const char *directory, const char *prefix, const char *const *names,
                     const double *sensor_out, const aefdtd_normal_result *res);
/* Waits for all work queued on the grid's stream. */
int aefdtd_synchronize(aefdtd_grid *g);

/* ---- batches: many independent runs stepped together -------------------- */
/* No counterpart in the reference, which steps one FDTD object at a time (its
 * only multi-grid loop is the lockstep pair of modes/oblique_biphase.cpp:
 * 251-292).  Each member advances exactly as aefdtd_run_step would advance
 * it alone; the batch issues one launch per half step for all members (after
 * one for the RC pre-pass of the dispersive ones) and replays chunks of 64
 * steps from a captured HIP graph.
 *  - Members are custom-mode (kind 2) and single-particle (kind 1) runs of
 *    the batch's precision, dispersive materials included, added at step 0
 *    (or any multiple of 64); a single-particle run's TFSF source column
 *    joins with it.  Normal-incidence runs (plane sums) and runs with a
 *    completion check are refused.
 *  - batch_step(b, c) advances every member by up to 64*c steps (never past
 *    its Nt); runs that reach Nt idle until removed.  A member must not be
 *    stepped on its own while in the batch; remove it, then finish it.
 *  - Membership changes take effect at the next batch_step (the graph is
 *    captured again). */
typedef struct aefdtd_batch aefdtd_batch;
int aefdtd_batch_create(aefdtd_batch **out, int precision);
int aefdtd_batch_destroy(aefdtd_batch *b);
const char *aefdtd_batch_error(const aefdtd_batch *b);
int aefdtd_batch_add(aefdtd_batch *b, aefdtd_run *r);
int aefdtd_batch_remove(aefdtd_batch *b, aefdtd_run *r);
int aefdtd_batch_step(aefdtd_batch *b, int chunks);
/* Members with steps left. */
int aefdtd_batch_running(const aefdtd_batch *b);
int aefdtd_batch_synchronize(aefdtd_batch *b);
/* Tile configurations (codes as aefdtd_set_tiles) of the batched passes. */
int aefdtd_batch_set_tiles(aefdtd_batch *b, int tile_e, int tile_h);
/* Time-skewed passes for 2D TE/TM members (one cell thick in z, no
 * dispersive material, TFSF or plane injection, FP64/FP32 storage): a
 * workgroup advances a tile `depth` steps in LDS before writing it back.
 * code = tile height index (0: 32 rows FP32 / 16 FP64, 1: 48 / 32) | depth
 * index (0: 2, 1: 4, 2: 8 steps) << 2; -1 steps every member one step at a
 * time.  Same arithmetic per cell either way. */
int aefdtd_batch_set_skew(aefdtd_batch *b, int code);
/* out[0..3]: members, members with steps left, members stepped by the
 * time-skewed passes, current skew code. */
int aefdtd_batch_info(const aefdtd_batch *b, int out[4]);
/* Online tuner: advances the batch by timed chunks, one per candidate
 * time-skew setting (off, large tiles at depth 4 and 8) and then per tile
 * shape of the per-step passes (BY 2..8, KZ 1..8, XCD fold on); all
 * candidates compute the same steps.  Keeps the fastest; out[3] = chosen E
 * tile, H tile and skew codes. */
int aefdtd_batch_autotune(aefdtd_batch *b, int out[3]);
/* Device time of aefdtd_batch_step (HIP events on the batch stream): on=1
 * resets and starts, 0 stops, -1 only reads. */
int aefdtd_batch_profile(aefdtd_batch *b, int on, double *ms, long *chunks);

/* ---- request server: continuous batching under an HBM budget ------------ */
/* No counterpart in the reference (it runs one mode per process call,
 * modes/lua_fdtd.cpp:296-320).  A worker thread owns the GPU work: requests
 * (a mode driver kind + its FDTD_Mode fields + sensor requests, all copied at
 * submit) wait in FIFO order until their estimated device bytes fit in the
 * budget next to the runs already resident, then stay resident until their
 * spectra are out.  Every round advances each resident run by 64 steps:
 * custom-mode runs without dispersive materials share one batch per
 * precision (aefdtd_batch_*), the others step on their own streams, so the
 * work of all resident runs overlaps on the device.  Requests that arrive
 * while others run join at the next round.  A request is served with the
 * same arithmetic alone or in a mix, so its result does not depend on the
 * mix. */
typedef struct aefdtd_server aefdtd_server;
#define AEFDTD_REQ_QUEUED    0
#define AEFDTD_REQ_RUNNING   1
#define AEFDTD_REQ_DONE      2
#define AEFDTD_REQ_FAILED    3
#define AEFDTD_REQ_CANCELLED 4
int aefdtd_server_create(aefdtd_server **out, size_t budget_bytes);
/* Cancels what is queued or resident, then stops the worker. */
int aefdtd_server_destroy(aefdtd_server *s);
/* kind as aefdtd_run_create (mode already finalized). Thread-safe. */
int aefdtd_server_submit(aefdtd_server *s, int kind, const aefdtd_mode *m, int n_sensors,
                         const aefdtd_sensor_req *sensors, long *id);
int aefdtd_server_cancel(aefdtd_server *s, long id);
/* Waits up to timeout_s (<0: no limit) for the request to end; *state gets
 * its AEFDTD_REQ_* state either way. */
int aefdtd_server_wait(aefdtd_server *s, long id, double timeout_s, int *state);
/* Output of a finished request: the sensor outputs as aefdtd_run_finish
 * writes them (n doubles at most), then for kind 0 the R and T spectra (Nl
 * each).  times (may be NULL): [0] queue wait, [1] admission to spectrum,
 * [2] submit to spectrum (time-to-spectrum), [3] resident device bytes, all
 * in seconds / bytes.  A failed request returns its error code. */
int aefdtd_server_result(aefdtd_server *s, long id, double *out, size_t n, double times[4]);
const char *aefdtd_server_error(aefdtd_server *s, long id);
/* Forgets an ended request. */
int aefdtd_server_release(aefdtd_server *s, long id);
/* out[0..7]: queued, resident, done, failed, bytes resident, peak bytes
 * resident, budget, rounds. */
int aefdtd_server_stats(aefdtd_server *s, double out[8]);
/* Device bytes a driver run of `kind` will hold for mode m (the admission
 * estimate; an upper bound of aefdtd_device_bytes over its grids). */
int aefdtd_estimate_bytes(int kind, const aefdtd_mode *m, int n_sensors,
                          const aefdtd_sensor_req *sensors, size_t *bytes);

/* ---- introspection used by the benchmark -------------------------------- */
/* Per-kernel device time accumulated on the handle's stream with HIP events
 * while profiling is enabled: names "E","H","sensor","dft","rc" (the RC pole
 * pre-pass of the E update). */
int aefdtd_profile_enable(aefdtd_grid *g, int on);
int aefdtd_profile_read(aefdtd_grid *g, const char *kernel, double *total_ms, long *launches);
/* Which kernels profiling brackets with events (bit 0 E, 1 H, 2 sensor, 3 dft,
 * 4 rc; default all): fewer events keep a timed region closer to unprofiled. */
int aefdtd_profile_kinds(aefdtd_grid *g, int mask);
/* Tile autotuner: times every compiled tile shape of the E and H passes on
 * this grid (reps launches each), then every tile of the one-pass step
 * (3D grids) against the best pair, keeps the fastest and restores the grid's state. out (may be NULL) receives the chosen codes (BY index | KZ index<<2 |
 * XCD-remap<<4; BY in 1,2,4,8 rows per workgroup, KZ in 1,2,4,8 planes per
 * lane, or rows per lane on grids one cell thick in z).  No counterpart in the reference (a CPU thread-count choice). */
int aefdtd_autotune(aefdtd_grid *g, int reps, int out[2]);
int aefdtd_set_tiles(aefdtd_grid *g, int tile_e, int tile_h);
/* Tile of the one-pass step of 3D grids (aefdtd_step and the drivers; the
 * fields of step n are read from one set of arrays and those of step n+1
 * written to a second set, which holds the state afterwards): code = NW index
 * (4, 8, 16 wavefronts per workgroup; a tile owns 64 columns and NW-2 rows) |
 * KZ index (8, 16, 32, 64 planes) << 2 | XCD fold << 4; NW index 3 is
 * refused; -1 makes every step the two passes.  The second set is
 * allocated at the first one-pass step (six field arrays and the E CPML psi).
 * No counterpart in the reference. */
int aefdtd_set_step_tiles(aefdtd_grid *g, int code);
/* The step tile code in use (-1: the two passes), e.g. the autotuner's choice. */
int aefdtd_get_step_tiles(aefdtd_grid *g, int *code);
/* Power-of-two storage scales of the six field components (stored value =
 * field * 2^exps[c]); FP16 storage only (defaults E: 2^8, H: 2^16, so that
 * unit-amplitude plane waves sit near the middle of the FP16 range; the custom
 * driver uses E: 2^0, H: 2^8 when it has oscillator sources, whose amplitude
 * is 50). Before the first update. */
int aefdtd_set_storage_scale(aefdtd_grid *g, const int exps[6]);
/* The storage exponents in use.  Runs with FP16 storage (the drivers, run
 * objects, the server and batches) check every 64 steps the largest stored
 * value of each component: above 2^13 the component's exponent is lowered so
 * that it is at most 2^12 (its stored values rescaled, exactly), and a stored
 * infinity or NaN ends the run with AEFDTD_ERANGE.  No counterpart in the
 * reference (which stores FP64). */
int aefdtd_get_storage_scale(aefdtd_grid *g, int exps[6]);
/* Reference point for the roofline, no counterpart in the reference: the
 * best bandwidth of a streaming device copy of `bytes` (read + write counted)
 * over grid sizes 1024..8192 x 256 lanes, 16-byte accesses; out[0] GB/s with
 * plain accesses, out[1] GB/s with non-temporal ones, out[2] ms of the best
 * plain copy. */
int aefdtd_stream_bandwidth(size_t bytes, int reps, double out[3]);
/* Reference point for the roofline, no counterpart in the reference: the
 * best bandwidth of a streaming kernel with the E and H passes' read:write
 * mix, six arrays of `bytes` read and three written (4-byte lanes,
 * non-temporal accesses), over grid sizes 2048..65536 x 256 lanes; out[0]
 * GB/s (all nine arrays counted), out[1] ms of the best launch. */
int aefdtd_stream_mix_bandwidth(size_t bytes, int reps, double out[2]);
/* Bytes of device memory held by the handle. */
int aefdtd_device_bytes(const aefdtd_grid *g, size_t *bytes);
/* out[0]=cells of the full grid, out[1]=dispersive (RC) cells, out[2]=max
 * poles per dispersive material, out[3]=bytes per stored field value. */
int aefdtd_stats(const aefdtd_grid *g, long out[4]);

#ifdef __cplusplus
}
#endif
#endif /* AETHER_FDTD_H */
