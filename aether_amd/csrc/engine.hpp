// engine.hpp -- host-side state of one FDTD grid on the device, shared by the
// HIP translation unit (kernels + grid C ABI) and the host mode drivers.
#pragma once

#include "../../include/aether_fdtd.h"

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace aefdtd {

using cplx = std::complex<double>;

// SI constants of Aether (src/common/phys_constants.h:19-21) and Pi (mathUT.h:34).
constexpr double kE0 = 8.854187817e-12;
constexpr double kMu0 = 12.566370614e-7;
constexpr double kC = 299792458.0;
constexpr double kPi = 3.141592653589793238462643383279502884;

// Update coefficients of one material index (FDTD_Material, core/fdtd_material.h:38-268).
struct HostMat {
    bool set = false;      // set_material was applied (else Aether's default object)
    bool rc = false;       // recursive-convolution poles present (m_type == MAT_RC)
    double C1 = 1, C2x = 0, C2y = 0, C2z = 0, C3 = 0, C4 = 0, ei = 1, sig = 0;
    std::vector<cplx> Crec, chi, Dchi;
};

// Host side of the pole conversion and the per-material constants; defined in
// host_setup.cpp.
void material_coefficients(const aefdtd_material &m, double Dx, double Dy, double Dz,
                           double Dt, HostMat &out);
void pml_profile(double ind, int N, int pm, int pp, const double kap[2], const double sig[2],
                 const double alp[2], double Dt, double &kappa, double &b, double &c);
double pml_sigma_scale(double D);
void pulse_init(aefdtd_pulse &p, double li, double lf, double inc, double pol, double dir,
                double sx, double sy, double sz, double n_ind);
double pulse_eval(const aefdtd_pulse &p, int comp, double x, double y, double z, double t);
double s_curve(double x, double xmin, double xmax);
// Text returned by aefdtd_last_error(NULL) (engine.hip).
void set_global_error(const std::string &msg);

// Oscillator source (sources/sc_oscillator.cpp:34-84): after every E update
// comp += val[step] over the box [x1,x2) x [y1,y2) x [z1,z2) (full-grid
// indices, after Source::link), step = grid step - step0 counting the
// source's own time steps.  val is a device table of the injected values,
// filled on the host by oscillator_value for steps [0, nval).
struct OscDev {
    int comp;                  // 0 Ex, 1 Ey, 2 Ez
    int x1, x2, y1, y2, z1, z2;
    int step0, nval;
    const double *val;
};
constexpr int MAX_OSC = 8;
// Oscillator::initialize (sources/sc_oscillator.cpp:70-84) for time step Dt.
struct OscParams {
    double w0 = 0, dw = 0, Dt = 0;
    int tshift = 0;            // an int in Aether (core/sources.h:222)
};
void oscillator_init(double lambda_min, double lambda_max, double Dt, OscParams &o);
// Re(e^{-i w0 step Dt} e^{-(dw/2 (step-tshift) Dt)^2}) * 50 (sc_oscillator.cpp:37-38).
double oscillator_value(const OscParams &o, int step);

// Batches of grids stepped together (engine.hip): one launch per half step
// over every member, chunks of 64 steps captured in a HIP graph.
struct Batch;
int batch_create(Batch **out, int precision);
void batch_destroy(Batch *b);
const char *batch_error(const Batch *b);
int batch_add(Batch *b, aefdtd_grid *g);
int batch_remove(Batch *b, aefdtd_grid *g);
int batch_members(const Batch *b);
int batch_step(Batch *b, int chunks);          // advances every member by up to chunks x 64 steps
int batch_synchronize(Batch *b);
// Completion energy of the current E (sum of E^2 over the grid): energy_feed
// queues it and updates the running maximum; energy_read waits and reads both.
int energy_feed(aefdtd_grid *g);
int energy_read(aefdtd_grid *g, double *last, double *max);

// Device state of the spectral completion criterion: np watched cells, their
// running DFT of local_E (FP64 complex, 3 components) and the last navg values
// of its real and imaginary parts ([6][np][navg], Ex Ey Ez real then imaginary).
struct FTState {
    int np = 0, navg = 0;
    int *d_pts = nullptr;             // [3 * np] i, j, k
    double *d_w = nullptr;            // [np] angular frequencies
    double *d_acc = nullptr;          // [6 * np]
    double *d_ring = nullptr;         // [6 * np * navg]
};
int ft_setup(aefdtd_grid *g, FTState &f, const std::vector<int> &pts, const std::vector<double> &w, int navg);
int ft_feed(aefdtd_grid *g, FTState &f, int step, double Dt);
int ft_read(aefdtd_grid *g, const FTState &f, std::vector<double> &ring);
void ft_free(FTState &f);
int ft_place(const aefdtd_grid *g, double Dt, double lmin, double lmax, int Np, const char layout[4], unsigned seed,
             std::vector<int> &pts, std::vector<double> &w, int &navg);
bool ft_complete(const std::vector<double> &ring, int np, int navg, double coeff, int steps, int nmax);
// FP16 storage range control of a grid (no-op for FP32/FP64): the largest
// stored value of each enabled component; above 2^13 the component's storage
// exponent is lowered so that it is at most 2^12 and its values are rescaled
// (*changed set); a stored infinity or NaN fails with AEFDTD_ERANGE.
int fp16_range(aefdtd_grid *g, bool *changed);
// Plane extent (cells) of sensor id after its link to the grid.
int sensor_spans(const aefdtd_grid *g, int id, int *span1, int *span2);
// Linked (full-grid) location x1, x2, y1, y2, z1, z2 of sensor id.
int sensor_box(const aefdtd_grid *g, int id, int box[6]);
// Field sensor id: block, normal axis, x1, y1, z1 (full grid), s1, s2, s3, and its wavelength.
int field_sensor_geom(const aefdtd_grid *g, int id, int out[8], double *lambda);
// A grid and its attached TFSF column stepped with one launch per half step.
int update_E_pair(aefdtd_grid *g, aefdtd_grid *aux);
int update_H_pair(aefdtd_grid *g, aefdtd_grid *aux);
// One full step (E then H) of a grid and its TFSF column (aux may be null):
// the one-pass kernel where the grid allows it, else the two passes.
int step_pair(aefdtd_grid *g, aefdtd_grid *aux);
hipStream_t batch_stream(const Batch *b);
int batch_set_tiles(Batch *b, int tile_e, int tile_h);
// Time-skewed passes of 2D TE/TM members: code = tile height index (0 small,
// 1 large) | depth index (2, 4, 8 steps) << 2; -1 steps them one step at a time.
int batch_set_skew(Batch *b, int code);
int batch_skew(const Batch *b);
int batch_skewed_members(const Batch *b);
// Device time of batch_step calls while on (HIP events): on=1 resets and
// starts, 0 stops, -1 only reads.
void batch_profile(Batch *b, int on, double *ms, long *chunks);

// DFT staging depth of a grid stepped on its own (steps between accumulator
// flushes): the matrix-core flushes stream their samples, so the depth only
// sets how often the FP64 accumulators are read and written (and the stage's
// size: depth x 6 x plane cells x the arithmetic type).  FP32 arithmetic
// stages 512 steps; the FP64 flush is bound by the FP64 matrix rate, so a
// deeper stage would only cost memory there.  Batches flush every 64-step
// chunk.
constexpr int stage_depth(int precision) { return precision == AEFDTD_FP64 ? 256 : 512; }

struct Sensor;      // planar spectral Poynting sensor (engine.hip)
struct FieldSensor; // single-wavelength FieldBlock / FieldMap (engine.hip)
struct PlaneSum;    // normal-incidence R/T plane sum (engine.hip)

}  // namespace aefdtd

struct aefdtd_grid {
    // ---- parameters (as the members of Aether's FDTD, core/fdtd_core.h:33-77)
    int precision = AEFDTD_FP64;
    int mode = AEFDTD_MODE_CUSTOM;
    int tstep = 0;
    int Nx = 0, Ny = 0, Nz = 0, Nt = 0, Ntap = 0;
    int Nx_s = 0, Ny_s = 0, Nz_s = 0;
    int xs_s = 0, xs_e = 0, ys_s = 0, ys_e = 0, zs_s = 0, zs_e = 0;
    int pad[6] = {0, 0, 0, 0, 0, 0};
    int pml[6] = {0, 0, 0, 0, 0, 0};
    double Dx = 0, Dy = 0, Dz = 0, Dt = 0;
    double dtm = 0, dtdex = 0, dtdey = 0, dtdez = 0, dtdmx = 0, dtdmy = 0, dtdmz = 0;
    double pml_kappa[6], pml_sigma[6], pml_alpha[6];
    bool enable[6] = {true, true, true, true, true, true};

    // ---- host copies
    std::vector<uint32_t> matsgrid;
    unsigned Nmat = 1;
    std::vector<aefdtd::HostMat> mats;
    std::vector<double> kE[3], kH[3], bE[3], bH[3], cE[3], cH[3];

    // ---- device state
    int device = 0;                   // the HIP device current at creation
    hipStream_t stream = nullptr;
    bool own_stream = true;
    void *F[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};   // T* fields
    uint8_t *d_mat = nullptr;
    void *d_mtab = nullptr;           // per-material coefficient table (MatDev<T>)
    double2 *d_poles = nullptr;       // per pole: Crec, Dchi
    int32_t *d_slot = nullptr;        // dispersive slot per cell, -1 elsewhere
    int2 *d_cells = nullptr;          // cell and material of each dispersive slot
    void *d_corr = nullptr;           // per slot: RC increment of the current E update (T[3])
    double2 *d_psi = nullptr;         // RC accumulators [3*Npmax][Ndisp]
    long n_disp = 0;
    int np_max = 0;
    void *d_prof = nullptr;           // 1D profile arrays (T): ikE/ikH/bE/bH/cE/cH per axis
    void *d_psipml[12] = {};          // CPML slab state, T
    size_t psipml_n[12] = {};
    // driver actions
    int inj_k = -1;
    std::vector<double> inj[4];
    double *d_inj = nullptr;          // the same values on the device, [step][dEx, dEy, dHx, dHy] (batches)
    void *tb_alt[7] = {};             // second live-state set of the time-skewed batch passes
    int tb_alt_pol = -1;              // (3 field components, 4 CPML psi), for this polarisation
    size_t tb_alt_bytes = 0;
    std::vector<aefdtd::OscDev> osc;  // oscillator sources, registration order
    std::vector<aefdtd::OscParams> osc_par;
    void *d_osc = nullptr;            // OscDev[MAX_OSC]
    aefdtd_grid *tfsf_aux = nullptr;
    int tb[6] = {0, 0, 0, 0, 0, 0};
    bool tf_per_x = false, tf_per_y = false;
    // sensors
    std::vector<aefdtd::Sensor *> sensors;
    std::vector<aefdtd::FieldSensor *> fsensors;
    std::vector<aefdtd::PlaneSum *> sums;
    int sensor_step = 0;              // Sensor::step shared by every registered sensor
    int stage_fill = 0;
    void *d_stage = nullptr;          // [KS][6][Ptot] T
    void *d_sdesc = nullptr;          // sensor descriptors for the sampling kernel
    long stage_P = 0, stage_maxP = 0;
    // x columns whose cells the H pass copies to the sensor strip (3D grids with
    // sensor planes normal to x), and whether the strip holds the current fields
    std::vector<int> strip_cols;      // x1 of each strip plane (columns x1, x1+1)
    void *d_strip = nullptr;          // [plane][Nz][Ny][2 columns][6] T
    bool strip_valid = false;
    unsigned *d_absmax = nullptr;     // FP16 range control: largest |stored| per component (float bits)
    double *d_energy = nullptr;       // completion energy: [EN_BLOCKS partials][last, max]
    bool en_fuse = false;             // the 3D E pass also sums the energy (FP32/FP64 runs with the check)
    double *d_en_part = nullptr;      // its per-workgroup partials
    long en_part_cap = 0;
    int en_parts = 0;                 // partials the last E launch wrote (0: none)
    // tile configuration of the 3D passes (BY index | KZ index << 2 | remap << 4)
    int tile_e = 1 | (2 << 2) | (1 << 4), tile_h = 1 | (2 << 2) | (1 << 4);
    // one-pass step of 3D grids (k_step): tile code (BY index | KZ index << 2 |
    // XCD fold << 4), -1 = the E and H passes (the default; the autotuner picks
    // the one-pass step where it is faster)
    int step_cfg = -1;
    // the set the one-pass step writes (fields, E CPML psi at the d_psipml
    // slots 0 1 4 5 8 9); swapped with F / d_psipml after each such step
    void *F_alt[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    void *psi_alt[12] = {};
    // bookkeeping
    bool ready = false;               // device copies up to date
    bool rc_ahead = false;            // the last sensor feed also ran the next E pass's RC pre-pass
    bool pml_dirty = true;
    size_t dev_bytes = 0;
    std::string err;
    // profiling
    bool prof_on = false;
    int prof_mask = 31;               // kinds bracketed with events (bit k: kind k)
    struct Ev { int kind; hipEvent_t a, b; };
    std::vector<Ev> events;
    double prof_ms[5] = {0, 0, 0, 0, 0};
    long prof_n[5] = {0, 0, 0, 0, 0};

    // power-of-two storage scale per field component (FP16 storage only)
    int scale_exp[6] = {0, 0, 0, 0, 0, 0};

    size_t ncell() const { return (size_t)Nx * Ny * Nz; }
    // device cell arrays: row pitch px >= Nx (3D grids pad rows to 64 cells)
    int px = 0;
    size_t nalloc() const { return (size_t)px * Ny * Nz; }
    size_t dcell(size_t c) const { return c % Nx + (size_t)px * (c / Nx); }   // dense -> device index
    size_t ssize() const { return precision == AEFDTD_FP64 ? 8 : (precision == AEFDTD_FP32 ? 4 : 2); }   // field storage
    size_t csize() const { return precision == AEFDTD_FP64 ? 8 : 4; }   // arithmetic, CPML psi, tables
};
