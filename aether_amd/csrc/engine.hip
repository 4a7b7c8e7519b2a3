// engine.hip -- grid-level C ABI of the FDTD engine: device state, uploads,
// launches, sensors and the tile autotuner.  The kernels are in kernels.hpp.
#include "kernels.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

using namespace aefdtd;
using namespace aefdtd::dev;

namespace aefdtd {

struct Sensor {                   // Spect_Poynting with interpolated samples
    int type, x1, x2, y1, y2, z1, z2, span1, span2, Nl;
    long P, off;                  // cells, offset in the stage/accumulator pools
    std::vector<double> lambda;
    double2 *d_acc = nullptr;     // [6][Nl][P]
    double2 *d_coef = nullptr;    // [ncoef steps][2][Nl]: E then H coefficient
    int ncoef = 0;
    void *d_coefmx = nullptr;     // [ncoef steps][E|H][LP] for the matrix-core flush: float (FP32 arithmetic) or double
    int LP = 0;                   // 2 Nl rounded up to 32
};

// FieldBlock / FieldMap (sensors/s_fieldmap.cpp:89-230, 321-470): the device
// descriptor, the wavelength and the sensor's own feed count (Sensor::step).
struct FieldSensor {
    dev::FieldDesc d;
    double lambda = 0;
    int step = 0;
    int ncomp = 0;
    int type = 0;                 // orientation of a FieldMap (folded to X/Y/Z)
};

struct PlaneSum {
    int k;
    bool has_inc, subtract;
    aefdtd_pulse inc;
    double zpos;
    double *d_tap = nullptr;      // [Nt]
    double *d_out = nullptr;      // [cap][6]
    int cap = 0, n = 0;
};

}  // namespace aefdtd

namespace {

// --------------------------------------------------------------- helpers
int fail(aefdtd_grid *g, int code, const std::string &msg) {
    if (g) g->err = msg;
    return code;
}

thread_local std::string g_create_err;   // the calling thread's last creation error
std::mutex g_create_mu;

}  // namespace

void aefdtd::set_global_error(const std::string &msg) {
    std::lock_guard<std::mutex> lk(g_create_mu);
    g_create_err = msg;
}

namespace {

#define HIPOK(g, call)                                                                                  \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return fail((g), e_ == hipErrorOutOfMemory ? AEFDTD_ENOMEM : AEFDTD_EDEVICE, \
                                          std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

int dalloc(aefdtd_grid *g, void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return AEFDTD_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(g, AEFDTD_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    g->dev_bytes += bytes;
    return AEFDTD_OK;
}

void dfree(void *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename P>
void dfree_t(P *&p) {
    void *v = (void *)p;
    dfree(v);
    p = nullptr;
}

bool is_fp64(const aefdtd_grid *g) { return g->precision == AEFDTD_FP64; }

// Calls f.template operator()<S, T>() with the storage and arithmetic types of
// the grid: (double, double), (float, float) or (_Float16, float).
template <class F>
int with_types(const aefdtd_grid *g, F &&f) {
    if (g->precision == AEFDTD_FP64) return f.template operator()<double, double>();
    if (g->precision == AEFDTD_FP32) return f.template operator()<float, float>();
    return f.template operator()<_Float16, float>();
}

// Calls f.template operator()<T>() with the arithmetic type of the grid.
template <class F>
int with_ctype(const aefdtd_grid *g, F &&f) {
    if (g->precision == AEFDTD_FP64) return f.template operator()<double>();
    return f.template operator()<float>();
}

// Host copy of a double array into a device buffer of the arithmetic type.
int upload_T(aefdtd_grid *g, void *dst, const double *src, size_t n) {
    if (n == 0) return AEFDTD_OK;
    if (is_fp64(g)) {
        HIPOK(g, hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, g->stream));
    } else {
        std::vector<float> t(n);
        for (size_t i = 0; i < n; i++) t[i] = (float)src[i];
        HIPOK(g, hipMemcpyAsync(dst, t.data(), n * 4, hipMemcpyHostToDevice, g->stream));
        HIPOK(g, hipStreamSynchronize(g->stream));
    }
    return AEFDTD_OK;
}

// Host <-> device copies and fills ordered on the grid's stream (never the
// legacy null stream, which would tie the grid to every other stream of the
// process, including one another thread is capturing into a graph).
hipError_t copy_sync(aefdtd_grid *g, void *dst, const void *src, size_t n, hipMemcpyKind k) {
    hipError_t e = hipMemcpyAsync(dst, src, n, k, g->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(g->stream);
}
hipError_t fill_sync(aefdtd_grid *g, void *dst, int v, size_t n) {
    hipError_t e = hipMemsetAsync(dst, v, n, g->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(g->stream);
}

// Dense host cell array (x fastest, Nx per row) <-> pitched device cell array.
// A strided copy from pageable host memory goes row by row (tens of thousands
// of small transfers on a 3D grid), so the rows are staged in a pitched host
// buffer and moved as one contiguous copy; the row padding is written as zero.
hipError_t cells_to_device(aefdtd_grid *g, void *dst, const void *src, size_t es) {
    const size_t rows = (size_t)g->Ny * g->Nz, w = (size_t)g->Nx * es, pw = (size_t)g->px * es;
    if (pw == w) return copy_sync(g, dst, src, rows * w, hipMemcpyHostToDevice);
    std::vector<char> st(rows * pw, 0);
    for (size_t r = 0; r < rows; r++) std::memcpy(st.data() + r * pw, (const char *)src + r * w, w);
    return copy_sync(g, dst, st.data(), rows * pw, hipMemcpyHostToDevice);
}
hipError_t cells_to_host(aefdtd_grid *g, void *dst, const void *src, size_t es) {
    const size_t rows = (size_t)g->Ny * g->Nz, w = (size_t)g->Nx * es, pw = (size_t)g->px * es;
    if (pw == w) return copy_sync(g, dst, src, rows * w, hipMemcpyDeviceToHost);
    std::vector<char> st(rows * pw);
    hipError_t e = copy_sync(g, st.data(), src, rows * pw, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (size_t r = 0; r < rows; r++) std::memcpy((char *)dst + r * w, st.data() + r * pw, w);
    return hipSuccess;
}

void pml_coeff_calc(aefdtd_grid *g) {          // core/fdtd_pml.cpp:305-343
    int N[3] = {g->Nx, g->Ny, g->Nz};
    for (int a = 0; a < 3; a++) {
        g->kE[a].assign(N[a], 1.0); g->kH[a].assign(N[a], 1.0);
        g->bE[a].assign(N[a], 0.0); g->bH[a].assign(N[a], 0.0);
        g->cE[a].assign(N[a], 0.0); g->cH[a].assign(N[a], 0.0);
        if (g->pml[2 * a] || g->pml[2 * a + 1]) {
            const double kap[2] = {g->pml_kappa[2 * a], g->pml_kappa[2 * a + 1]};
            const double sig[2] = {g->pml_sigma[2 * a], g->pml_sigma[2 * a + 1]};
            const double alp[2] = {g->pml_alpha[2 * a], g->pml_alpha[2 * a + 1]};
            for (int i = 0; i < N[a]; i++) {
                pml_profile(i, N[a], g->pml[2 * a], g->pml[2 * a + 1], kap, sig, alp, g->Dt, g->kE[a][i], g->bE[a][i], g->cE[a][i]);
                pml_profile(i + 0.5, N[a], g->pml[2 * a], g->pml[2 * a + 1], kap, sig, alp, g->Dt, g->kH[a][i], g->bH[a][i], g->cH[a][i]);
            }
        }
    }
}

// Uploads the 1D profile tables: inverse kappa as the stencils use it (x axis
// gated by pml_xm||pml_xp, y axis by pml_ym||pml_xp as core/fdtd_core.cpp:399,
// z by pml_zm||pml_zp), b, and c pre-multiplied by 1/D (fdtd_pml.cpp:523 etc.).
int upload_profiles(aefdtd_grid *g) {
    pml_coeff_calc(g);
    int N[3] = {g->Nx, g->Ny, g->Nz};
    const bool gate[3] = {g->pml[0] || g->pml[1], g->pml[2] || g->pml[1], g->pml[4] || g->pml[5]};
    const double invD[3] = {1.0 / g->Dx, 1.0 / g->Dy, 1.0 / g->Dz};
    size_t tot = 6 * (size_t)(N[0] + N[1] + N[2]);
    std::vector<double> h(tot);
    size_t o = 0;
    for (int a = 0; a < 3; a++) {
        for (int i = 0; i < N[a]; i++) h[o + i] = 1.0 / (gate[a] ? g->kE[a][i] : 1.0);
        o += N[a];
        for (int i = 0; i < N[a]; i++) h[o + i] = 1.0 / (gate[a] ? g->kH[a][i] : 1.0);
        o += N[a];
        for (int i = 0; i < N[a]; i++) h[o + i] = g->bE[a][i];
        o += N[a];
        for (int i = 0; i < N[a]; i++) h[o + i] = g->cE[a][i] * invD[a];
        o += N[a];
        for (int i = 0; i < N[a]; i++) h[o + i] = g->bH[a][i];
        o += N[a];
        for (int i = 0; i < N[a]; i++) h[o + i] = g->cH[a][i] * invD[a];
        o += N[a];
    }
    if (!g->d_prof) {
        int rc = dalloc(g, &g->d_prof, tot * g->csize());
        if (rc) return rc;
    }
    return upload_T(g, g->d_prof, h.data(), tot);
}

template <typename T>
Prof<T> prof_of(const aefdtd_grid *g) {
    Prof<T> p;
    const T *b = (const T *)g->d_prof;
    int N[3] = {g->Nx, g->Ny, g->Nz};
    for (int a = 0; a < 3; a++) {
        p.ikE[a] = b; b += N[a];
        p.ikH[a] = b; b += N[a];
        p.bE[a] = b; b += N[a];
        p.cE[a] = b; b += N[a];
        p.bH[a] = b; b += N[a];
        p.cH[a] = b; b += N[a];
    }
    return p;
}

// CPML slab state (core/fdtd_pml.cpp:26-85): order EYX,EZX,HYX,HZX (x slabs),
// EXY,EZY,HXY,HZY (y slabs), EXZ,EYZ,HXZ,HYZ (z slabs).
int alloc_pml_state(aefdtd_grid *g) {
    for (int s = 0; s < 12; s++) {
        if (g->psi_alt[s]) g->dev_bytes -= g->psipml_n[s] * g->csize();
        dfree(g->psi_alt[s]);           // the one-pass step's second set is made again from these
        if (g->d_psipml[s]) g->dev_bytes -= g->psipml_n[s] * g->csize();
        dfree(g->d_psipml[s]);
        g->psipml_n[s] = 0;
    }
    size_t n[3] = {0, 0, 0};
    if (g->pml[0] || g->pml[1]) n[0] = (size_t)(g->pml[0] + g->pml[1]) * g->Ny * g->Nz;
    if (g->pml[2] || g->pml[3]) n[1] = (size_t)g->Nx * (g->pml[2] + g->pml[3]) * g->Nz;
    if (g->pml[4] || g->pml[5]) n[2] = (size_t)g->Nx * g->Ny * (g->pml[4] + g->pml[5]);
    for (int s = 0; s < 12; s++) {
        size_t m = n[s / 4];
        if (!m) continue;
        int rc = dalloc(g, &g->d_psipml[s], m * g->csize());
        if (rc) return rc;
        HIPOK(g, hipMemsetAsync(g->d_psipml[s], 0, m * g->csize(), g->stream));
        g->psipml_n[s] = m;
    }
    return AEFDTD_OK;
}

template <typename T>
int upload_materials_T(aefdtd_grid *g) {
    std::vector<MatDev<T>> tab(g->Nmat);
    std::vector<double2> poles;
    int npmax = 0;
    for (unsigned m = 0; m < g->Nmat; m++) {
        const HostMat &M = g->mats[m];
        MatDev<T> d;
        d.C1 = (T)M.C1; d.C2x = (T)M.C2x; d.C2y = (T)M.C2y; d.C2z = (T)M.C2z; d.C4 = (T)M.C4;
        d.C3 = M.C3;
        d.rc = M.rc ? 1 : 0;
        d.np = (int)M.Crec.size();
        d.pole0 = (int)(poles.size() / 2);
        for (int p = 0; p < d.np; p++) {
            poles.push_back(make_double2(M.Crec[p].real(), M.Crec[p].imag()));
            poles.push_back(make_double2(M.Dchi[p].real(), M.Dchi[p].imag()));
        }
        if (M.rc) npmax = std::max(npmax, d.np);
        tab[m] = d;
    }
    dfree(g->d_mtab);
    dfree_t(g->d_poles);
    int rc = dalloc(g, &g->d_mtab, tab.size() * sizeof(MatDev<T>));
    if (rc) return rc;
    HIPOK(g, hipMemcpyAsync(g->d_mtab, tab.data(), tab.size() * sizeof(MatDev<T>), hipMemcpyHostToDevice, g->stream));
    if (!poles.empty()) {
        void *p;
        if ((rc = dalloc(g, &p, poles.size() * sizeof(double2)))) return rc;
        g->d_poles = (double2 *)p;
        HIPOK(g, hipMemcpyAsync(g->d_poles, poles.data(), poles.size() * sizeof(double2), hipMemcpyHostToDevice, g->stream));
    }
    g->np_max = npmax;
    HIPOK(g, hipStreamSynchronize(g->stream));
    return AEFDTD_OK;
}

// Material ids (uint8), dispersive slots and RC accumulators.  The reference
// keeps one psi block per material over its bounding box (core/mats.cpp:394-496);
// here only dispersive cells get a slot, numbered in x-fastest order so that
// neighbouring lanes touch neighbouring slots.
int upload_grid(aefdtd_grid *g) {
    if (g->Nmat > 256) return fail(g, AEFDTD_EINVAL, "more than 256 material indices");
    g->rc_ahead = false;                // the pole states are rebuilt below
    const size_t NA = g->nalloc(), rows = (size_t)g->Ny * g->Nz;
    // one pass over the cells: uint8 ids in the pitched layout, and the device
    // indices of the dispersive cells in x-fastest order
    bool disp[256] = {};
    bool any = false;
    for (unsigned m = 0; m < g->Nmat; m++) {
        disp[m] = g->mats[m].rc;
        any = any || disp[m];
    }
    std::vector<uint8_t> m8(NA, 0);
    std::vector<int2> cells;
    for (size_t r = 0; r < rows; r++) {
        const uint32_t *src = g->matsgrid.data() + r * g->Nx;
        uint8_t *dst = m8.data() + r * g->px;
        for (int i = 0; i < g->Nx; i++) {
            dst[i] = (uint8_t)src[i];
            if (disp[src[i]]) cells.push_back(make_int2((int)(r * g->px + i), (int)src[i]));
        }
    }
    if (!g->d_mat) {
        void *p;
        int rc = dalloc(g, &p, NA);
        if (rc) return rc;
        g->d_mat = (uint8_t *)p;
    }
    HIPOK(g, copy_sync(g, g->d_mat, m8.data(), NA, hipMemcpyHostToDevice));
    int rc = with_ctype(g, [&]<typename T>() { return upload_materials_T<T>(g); });
    if (rc) return rc;
    dfree_t(g->d_slot);
    dfree_t(g->d_psi);
    dfree_t(g->d_cells);
    dfree(g->d_corr);
    g->n_disp = 0;
    if (any && !cells.empty()) {
        const long nd = (long)cells.size();
        void *p;
        if ((rc = dalloc(g, &p, NA * sizeof(int32_t)))) return rc;
        g->d_slot = (int32_t *)p;
        HIPOK(g, hipMemsetAsync(g->d_slot, 0xFF, NA * sizeof(int32_t), g->stream));     // -1: no slot
        if ((rc = dalloc(g, &p, cells.size() * sizeof(int2)))) return rc;
        g->d_cells = (int2 *)p;
        HIPOK(g, hipMemcpyAsync(g->d_cells, cells.data(), cells.size() * sizeof(int2), hipMemcpyHostToDevice,
                                g->stream));
        hipLaunchKernelGGL(k_slot_scatter, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, g->stream,
                           (const int2 *)g->d_cells, nd, g->d_slot);
        HIPOK(g, hipGetLastError());
        if ((rc = dalloc(g, &g->d_corr, (size_t)3 * nd * g->csize()))) return rc;
        size_t np = (size_t)3 * g->np_max * nd;
        if ((rc = dalloc(g, &p, np * sizeof(double2)))) return rc;
        g->d_psi = (double2 *)p;
        HIPOK(g, hipMemsetAsync(g->d_psi, 0, np * sizeof(double2), g->stream));
        g->n_disp = nd;
    }
    HIPOK(g, hipStreamSynchronize(g->stream));
    return AEFDTD_OK;
}

int ensure_ready(aefdtd_grid *g) {
    if (g->ready) return AEFDTD_OK;
    int rc = upload_grid(g);
    if (rc) return rc;
    if ((rc = upload_profiles(g))) return rc;
    g->pml_dirty = false;
    g->ready = true;
    return AEFDTD_OK;
}

// Events used only to time device work: no system-scope fence (and hence no
// cache writeback and invalidation) when they are recorded, so that timing a
// pass does not slow the passes around it.
hipError_t timing_event(hipEvent_t *e) { return hipEventCreateWithFlags(e, hipEventDisableSystemFence); }

void prof_begin(aefdtd_grid *g, int kind, hipEvent_t *b) {
    *b = nullptr;
    if (!g->prof_on || !((g->prof_mask >> kind) & 1)) return;
    hipEvent_t a;
    if (timing_event(&a) != hipSuccess || timing_event(b) != hipSuccess) { *b = nullptr; return; }
    (void)hipEventRecord(a, g->stream);
    g->events.push_back({kind, a, *b});
}

void prof_end(aefdtd_grid *g, hipEvent_t b) {
    if (b) (void)hipEventRecord(b, g->stream);
}

void prof_collect(aefdtd_grid *g) {
    if (g->events.empty()) return;
    (void)hipStreamSynchronize(g->stream);
    for (auto &e : g->events) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            g->prof_ms[e.kind] += ms;
            g->prof_n[e.kind] += 1;
        }
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    g->events.clear();
}

// ---------------------------------------------------------------- launches
// Tile shapes compiled for the 3D passes: (BY, KZ) with BY in {1,2,4,8} and
// KZ in {1,2,4,8}; tile_cfg packs (BY index, KZ index, remap).
constexpr int kBY[4] = {1, 2, 4, 8};
constexpr int kKZ[4] = {1, 2, 4, 8};

// A 3D launch of nb tile workgroups; `aux` (a 1 x 1 x Nz column, or nullptr)
// is updated by extra workgroups of the same launch.
template <typename S, typename T, int BY, int KZ>
void launch_E_cfg(const EArgs<S, T> &a, int nblocks, hipStream_t st, const EArgs<S, T> *aux) {
    if (a.Nz == 1) {
        hipLaunchKernelGGL((k_update_E2<S, T, BY, KZ>), dim3(nblocks), dim3(BX, BY, 1), a.nmat * sizeof(MatL<T>), st, a);
    } else {
        const int naux = aux ? (aux->Nz + BX * BY - 1) / (BX * BY) : 0;
        const int nmat = aux ? std::max(a.nmat, aux->nmat) : a.nmat;
        if constexpr (std::is_same_v<S, T>) {
            if (a.en_part) {
                hipLaunchKernelGGL((k_update_E<S, T, BY, KZ, true>), dim3(nblocks + naux), dim3(BX, BY, 1),
                                   nmat * sizeof(MatL<T>), st, a, aux ? *aux : a, nblocks);
                return;
            }
        }
        hipLaunchKernelGGL((k_update_E<S, T, BY, KZ>), dim3(nblocks + naux), dim3(BX, BY, 1), nmat * sizeof(MatL<T>), st,
                           a, aux ? *aux : a, nblocks);
    }
}
template <typename S, typename T, int BY, int KZ>
void launch_H_cfg(const HArgs<S, T> &a, int nblocks, hipStream_t st, const HArgs<S, T> *aux) {
    if (a.Nz == 1) {
        hipLaunchKernelGGL((k_update_H2<S, T, BY, KZ>), dim3(nblocks), dim3(BX, BY, 1), 0, st, a);
    } else {
        const int naux = aux ? (aux->Nz + BX * BY - 1) / (BX * BY) : 0;
        hipLaunchKernelGGL((k_update_H<S, T, BY, KZ>), dim3(nblocks + naux), dim3(BX, BY, 1), 0, st, a, aux ? *aux : a,
                           nblocks);
    }
}

template <typename S, typename T, int BY>
void launch_E_by(const EArgs<S, T> &a, int kzi, int nb, hipStream_t st, const EArgs<S, T> *aux) {
    switch (kzi) {
        case 0: launch_E_cfg<S, T, BY, 1>(a, nb, st, aux); break;
        case 1: launch_E_cfg<S, T, BY, 2>(a, nb, st, aux); break;
        case 2: launch_E_cfg<S, T, BY, 4>(a, nb, st, aux); break;
        default: launch_E_cfg<S, T, BY, 8>(a, nb, st, aux); break;
    }
}
template <typename S, typename T, int BY>
void launch_H_by(const HArgs<S, T> &a, int kzi, int nb, hipStream_t st, const HArgs<S, T> *aux) {
    switch (kzi) {
        case 0: launch_H_cfg<S, T, BY, 1>(a, nb, st, aux); break;
        case 1: launch_H_cfg<S, T, BY, 2>(a, nb, st, aux); break;
        case 2: launch_H_cfg<S, T, BY, 4>(a, nb, st, aux); break;
        default: launch_H_cfg<S, T, BY, 8>(a, nb, st, aux); break;
    }
}

// Fills the tile geometry of an argument block for configuration cfg and
// returns the number of workgroups.
// A grid one cell thick in z runs the 2D passes, which march the KZ depth
// along y instead.
template <typename A>
int tile_geometry(A &a, int cfg) {
    const int by = kBY[(cfg >> 0) & 3], kz = kKZ[(cfg >> 2) & 3];
    a.nbx = (a.Nx + BX - 1) / BX;
    if (a.Nz == 1) {
        a.nby = (a.Ny + by * kz - 1) / (by * kz);
        a.nbz = 1;
    } else {
        a.nby = (a.Ny + by - 1) / by;
        a.nbz = (a.Nz + kz - 1) / kz;
    }
    a.remap = (cfg >> 4) & 1;
    return a.nbx * a.nby * a.nbz;
}

bool is_column(const aefdtd_grid *g) {
    return g->Nx == 1 && g->Ny == 1 && !(g->pml[0] || g->pml[1] || g->pml[2] || g->pml[3]);
}

// Field component c of grid g as seen by the kernels.
template <typename S, typename T>
Fld<S, T> fld(const aefdtd_grid *g, int c) {
    Fld<S, T> f;
    f.p = (S *)g->F[c];
    f.ld = (T)std::ldexp(1.0, -g->scale_exp[c]);
    f.st = (T)std::ldexp(1.0, g->scale_exp[c]);
    return f;
}

// Value tables of the oscillators up to grid step `upto` (exclusive), grown
// by doubling; the descriptors are re-uploaded when a table moves.
int ensure_osc_tables(aefdtd_grid *g, int upto) {
    bool moved = false;
    for (size_t q = 0; q < g->osc.size(); q++) {
        OscDev &o = g->osc[q];
        const int need = upto - o.step0;
        if (need <= o.nval) continue;
        const int n = std::max(need, 2 * o.nval);
        std::vector<double> v(n);
        for (int st = 0; st < n; st++) v[st] = oscillator_value(g->osc_par[q], st);
        void *p;
        int rc = dalloc(g, &p, n * sizeof(double));
        if (rc) return rc;
        HIPOK(g, hipStreamSynchronize(g->stream));
        HIPOK(g, copy_sync(g, p, v.data(), n * sizeof(double), hipMemcpyHostToDevice));
        if (o.val) {
            g->dev_bytes -= (size_t)o.nval * sizeof(double);
            void *old = (void *)o.val;
            dfree(old);
        }
        o.val = (const double *)p;
        o.nval = n;
        moved = true;
    }
    if (moved) HIPOK(g, copy_sync(g, g->d_osc, g->osc.data(), g->osc.size() * sizeof(OscDev), hipMemcpyHostToDevice));
    return AEFDTD_OK;
}

// Load masks of the passes.  An E component is loaded and stored when it is
// updated (enable_*, core/fdtd_threads.cpp:60-82) or written by an injection;
// an H component is loaded when an updated E component takes a difference of
// it along an axis longer than one cell (along a one-cell axis the periodic
// difference is exactly zero, core/fdtd_core.cpp:376-379).  The H pass mirrors
// this with E and H exchanged.  Fields outside the masks stay zero.
int load_mask_E(const aefdtd_grid *g) {
    const bool ex = g->enable[0], ey = g->enable[1], ez = g->enable[2];
    const bool lx = g->Nx > 1, ly = g->Ny > 1, lz = g->Nz > 1;
    int m = (ex ? 1 : 0) | (ey ? 2 : 0) | (ez ? 4 : 0);
    if (g->inj_k >= 0) m |= 1 | 2;
    if (g->tfsf_aux) m |= 7;
    for (const OscDev &o : g->osc) m |= 1 << o.comp;
    if ((ey && lz) || (ez && ly)) m |= 8;     // Hx: dHx/dz in Ey, dHx/dy in Ez
    if ((ex && lz) || (ez && lx)) m |= 16;    // Hy: dHy/dz in Ex, dHy/dx in Ez
    if ((ex && ly) || (ey && lx)) m |= 32;    // Hz: dHz/dy in Ex, dHz/dx in Ey
    return m;
}

int load_mask_H(const aefdtd_grid *g) {
    const bool hx = g->enable[3], hy = g->enable[4], hz = g->enable[5];
    const bool lx = g->Nx > 1, ly = g->Ny > 1, lz = g->Nz > 1;
    int m = (hx ? 8 : 0) | (hy ? 16 : 0) | (hz ? 32 : 0);
    if (g->inj_k >= 0) m |= 8 | 16;
    if (g->tfsf_aux) m |= 8 | 16 | 32;
    if ((hy && lz) || (hz && ly)) m |= 1;     // Ex: dEx/dz in Hy, dEx/dy in Hz
    if ((hx && lz) || (hz && lx)) m |= 2;     // Ey: dEy/dz in Hx, dEy/dx in Hz
    if ((hx && ly) || (hy && lx)) m |= 4;     // Ez: dEz/dy in Hx, dEz/dx in Hy
    return m;
}

template <typename S, typename T>
EArgs<S, T> e_args(aefdtd_grid *g) {
    EArgs<S, T> a;
    std::memset(&a, 0, sizeof(a));
    a.Nx = g->Nx; a.Ny = g->Ny; a.Nz = g->Nz; a.P = g->px;
    a.nmat = (int)g->Nmat;
    a.Ex = fld<S, T>(g, 0); a.Ey = fld<S, T>(g, 1); a.Ez = fld<S, T>(g, 2);
    a.Hx = fld<S, T>(g, 3); a.Hy = fld<S, T>(g, 4); a.Hz = fld<S, T>(g, 5);
    a.mat = g->d_mat;
    a.mt = (const MatDev<T> *)g->d_mtab;
    a.slot = g->d_slot;
    a.corr = (const T *)g->d_corr;
    a.pf = prof_of<T>(g);
    for (int f = 0; f < 6; f++) a.pml[f] = g->pml[f];
    a.pEyx = (T *)g->d_psipml[0]; a.pEzx = (T *)g->d_psipml[1];
    a.pExy = (T *)g->d_psipml[4]; a.pEzy = (T *)g->d_psipml[5];
    a.pExz = (T *)g->d_psipml[8]; a.pEyz = (T *)g->d_psipml[9];
    a.en = (g->enable[0] ? 1 : 0) | (g->enable[1] ? 2 : 0) | (g->enable[2] ? 4 : 0);
    a.inj_k = -1;
    if (g->inj_k >= 0 && g->tstep < (int)g->inj[0].size()) {
        a.inj_k = g->inj_k;
        a.injx = (T)g->inj[0][g->tstep];
        a.injy = (T)g->inj[1][g->tstep];
    }
    if (g->tfsf_aux) {
        a.tfsf = 1;
        a.tb0 = g->tb[0]; a.tb1 = g->tb[1]; a.tb2 = g->tb[2]; a.tb3 = g->tb[3]; a.tb4 = g->tb[4]; a.tb5 = g->tb[5];
        a.tper_x = g->tf_per_x; a.tper_y = g->tf_per_y;
        a.aHx = fld<S, T>(g->tfsf_aux, 3); a.aHy = fld<S, T>(g->tfsf_aux, 4); a.aHz = fld<S, T>(g->tfsf_aux, 5);
    }
    a.lm = load_mask_E(g);
    a.nosc = (int)g->osc.size();
    a.osc = (const OscDev *)g->d_osc;
    a.ostep = g->tstep;
    return a;
}

template <typename S, typename T>
HArgs<S, T> h_args(aefdtd_grid *g) {
    HArgs<S, T> a;
    std::memset(&a, 0, sizeof(a));
    a.Nx = g->Nx; a.Ny = g->Ny; a.Nz = g->Nz; a.P = g->px;
    a.Ex = fld<S, T>(g, 0); a.Ey = fld<S, T>(g, 1); a.Ez = fld<S, T>(g, 2);
    a.Hx = fld<S, T>(g, 3); a.Hy = fld<S, T>(g, 4); a.Hz = fld<S, T>(g, 5);
    a.pf = prof_of<T>(g);
    for (int f = 0; f < 6; f++) a.pml[f] = g->pml[f];
    a.pHyx = (T *)g->d_psipml[2]; a.pHzx = (T *)g->d_psipml[3];
    a.pHxy = (T *)g->d_psipml[6]; a.pHzy = (T *)g->d_psipml[7];
    a.pHxz = (T *)g->d_psipml[10]; a.pHyz = (T *)g->d_psipml[11];
    a.en = (g->enable[3] ? 8 : 0) | (g->enable[4] ? 16 : 0) | (g->enable[5] ? 32 : 0);
    a.dtm = (T)g->dtm; a.dtdmx = (T)g->dtdmx; a.dtdmy = (T)g->dtdmy; a.dtdmz = (T)g->dtdmz;
    a.inj_k = -1;
    if (g->inj_k >= 0 && g->tstep < (int)g->inj[2].size()) {
        a.inj_k = g->inj_k;
        a.injx = (T)g->inj[2][g->tstep];
        a.injy = (T)g->inj[3][g->tstep];
    }
    if (g->tfsf_aux) {
        a.tfsf = 1;
        a.tb0 = g->tb[0]; a.tb1 = g->tb[1]; a.tb2 = g->tb[2]; a.tb3 = g->tb[3]; a.tb4 = g->tb[4]; a.tb5 = g->tb[5];
        a.tper_x = g->tf_per_x; a.tper_y = g->tf_per_y;
        a.aEx = fld<S, T>(g->tfsf_aux, 0); a.aEy = fld<S, T>(g->tfsf_aux, 1); a.aEz = fld<S, T>(g->tfsf_aux, 2);
    }
    a.lm = load_mask_H(g);
    if (g->d_strip) {
        a.nstrip = (int)g->strip_cols.size();
        for (int q = 0; q < a.nstrip; q++) a.scol[q] = g->strip_cols[q];
        a.strip = (T *)g->d_strip;
    }
    return a;
}

template <typename S, typename T>
void dispatch_E(aefdtd_grid *g, EArgs<S, T> &a, int cfg, const EArgs<S, T> *aux = nullptr) {
    if (is_column(g)) {
        hipLaunchKernelGGL((k_update_E_col<S, T>), dim3((g->Nz + 255) / 256), dim3(256), a.nmat * sizeof(MatL<T>),
                           g->stream, a);
        return;
    }
    const int nb = tile_geometry(a, cfg);
    switch (cfg & 3) {
        case 0: launch_E_by<S, T, 1>(a, (cfg >> 2) & 3, nb, g->stream, aux); break;
        case 1: launch_E_by<S, T, 2>(a, (cfg >> 2) & 3, nb, g->stream, aux); break;
        case 2: launch_E_by<S, T, 4>(a, (cfg >> 2) & 3, nb, g->stream, aux); break;
        default: launch_E_by<S, T, 8>(a, (cfg >> 2) & 3, nb, g->stream, aux); break;
    }
}

template <typename S, typename T>
void dispatch_H(aefdtd_grid *g, HArgs<S, T> &a, int cfg, const HArgs<S, T> *aux = nullptr) {
    if (is_column(g)) {
        hipLaunchKernelGGL((k_update_H_col<S, T>), dim3((g->Nz + 255) / 256), dim3(256), 0, g->stream, a);
        return;
    }
    const int nb = tile_geometry(a, cfg);
    switch (cfg & 3) {
        case 0: launch_H_by<S, T, 1>(a, (cfg >> 2) & 3, nb, g->stream, aux); break;
        case 1: launch_H_by<S, T, 2>(a, (cfg >> 2) & 3, nb, g->stream, aux); break;
        case 2: launch_H_by<S, T, 4>(a, (cfg >> 2) & 3, nb, g->stream, aux); break;
        default: launch_H_by<S, T, 8>(a, (cfg >> 2) & 3, nb, g->stream, aux); break;
    }
}

template <typename S, typename T>
RcArgs<S, T> rc_args(aefdtd_grid *g) {
    RcArgs<S, T> r;
    r.ndisp = (int)g->n_disp;
    r.np_max = g->np_max;
    r.cells = g->d_cells;
    r.mt = (const MatDev<T> *)g->d_mtab;
    r.poles = g->d_poles;
    r.psi = g->d_psi;
    r.Ex = fld<S, T>(g, 0); r.Ey = fld<S, T>(g, 1); r.Ez = fld<S, T>(g, 2);
    r.corr = (T *)g->d_corr;
    return r;
}

template <typename S, typename T>
int launch_rc(aefdtd_grid *g) {
    if (g->n_disp == 0) return AEFDTD_OK;
    const RcArgs<S, T> r = rc_args<S, T>(g);
    hipEvent_t ev;
    prof_begin(g, 4, &ev);
    hipLaunchKernelGGL((k_rc_ante<S, T>), dim3((unsigned)((g->n_disp + 255) / 256), 3), dim3(256), 0, g->stream, r);
    prof_end(g, ev);
    HIPOK(g, hipGetLastError());
    return AEFDTD_OK;
}

// E update of g (and of its TFSF column aux in the same launch, when given).
template <typename S, typename T>
int launch_E(aefdtd_grid *g, aefdtd_grid *aux = nullptr) {
    int rc = ensure_osc_tables(g, g->tstep + 1);
    if (rc) return rc;
    if (aux && (rc = ensure_osc_tables(aux, aux->tstep + 1))) return rc;
    if (aux && (rc = launch_rc<S, T>(aux))) return rc;
    if (!g->rc_ahead && (rc = launch_rc<S, T>(g))) return rc;      // else the sensor feed ran it
    g->rc_ahead = false;
    g->strip_valid = false;             // E changes: the strip is stale until the next H pass
    EArgs<S, T> a = e_args<S, T>(g), x{};
    if (aux) x = e_args<S, T>(aux);
    g->en_parts = 0;
    if constexpr (std::is_same_v<S, T>) {
        if (g->en_fuse && g->Nz > 1 && !is_column(g)) {
            EArgs<S, T> t = a;
            const int nb = tile_geometry(t, g->tile_e);
            if (nb > g->en_part_cap) {
                dfree_t(g->d_en_part);
                g->en_part_cap = 0;
                void *p;
                if ((rc = dalloc(g, &p, (size_t)nb * sizeof(double)))) return rc;
                g->d_en_part = (double *)p;
                g->en_part_cap = nb;
            }
            a.en_part = g->d_en_part;
            g->en_parts = nb;
        }
    }
    hipEvent_t ev;
    prof_begin(g, 0, &ev);
    dispatch_E<S, T>(g, a, g->tile_e, aux ? &x : nullptr);
    prof_end(g, ev);
    HIPOK(g, hipGetLastError());
    return AEFDTD_OK;
}

template <typename S, typename T>
int launch_H(aefdtd_grid *g, aefdtd_grid *aux = nullptr) {
    HArgs<S, T> a = h_args<S, T>(g), x{};
    if (aux) x = h_args<S, T>(aux);
    hipEvent_t ev;
    prof_begin(g, 1, &ev);
    dispatch_H<S, T>(g, a, g->tile_h, aux ? &x : nullptr);
    prof_end(g, ev);
    HIPOK(g, hipGetLastError());
    // the 3D pass wrote the strip from the fields it left (the column kernel
    // and the 2D pass do not write it)
    g->strip_valid = a.nstrip > 0 && g->Nz > 1 && !is_column(g);
    return AEFDTD_OK;
}

// ---- one-pass step (k_step) of 3D grids
constexpr int kStepNW[3] = {4, 8, 16};                // wavefronts per workgroup (NW - 2 rows owned)
constexpr int kStepKZ[4] = {8, 16, 32, 64};
constexpr int kStepPsiE[6] = {0, 1, 4, 5, 8, 9};     // E slabs in d_psipml (EYX EZX EXY EZY EXZ EYZ)

// Tile plan of step code cfg on grid g (NW index | KZ index << 2 | XCD fold
// << 4; tiles of 64 columns, NW - 2 rows and KZ planes); ok = false where the
// one-pass kernel does not apply (2D grids, columns) or the code names no tile.
struct StepPlan {
    int NW = 0, KZ = 0, ntx = 0, nty = 0, ntz = 0, remap = 0;
    bool ok = false;
};

StepPlan step_plan(const aefdtd_grid *g, int cfg) {
    StepPlan p;
    if (cfg < 0 || cfg > 31 || (cfg & 3) == 3 || g->Nz < 2 || is_column(g)) return p;
    p.NW = kStepNW[cfg & 3];
    p.KZ = kStepKZ[(cfg >> 2) & 3];
    p.remap = (cfg >> 4) & 1;
    p.ntx = (g->Nx + BX - 1) / BX;
    p.nty = (g->Ny + p.NW - 3) / (p.NW - 2);
    p.ntz = (g->Nz + p.KZ - 1) / p.KZ;
    p.ok = true;
    return p;
}

// The second set of fields and E CPML psi the one-pass step writes into,
// allocated at the first one-pass step and filled with the current state (so
// that entries no step updates, such as the psi of a disabled component, hold
// the same values in both sets).
int ensure_step_set(aefdtd_grid *g) {
    const size_t nf = g->nalloc() * g->ssize();
    for (int c = 0; c < 6; c++) {
        if (g->F_alt[c]) continue;
        int rc = dalloc(g, &g->F_alt[c], nf);
        if (rc) return rc;
        HIPOK(g, hipMemcpyAsync(g->F_alt[c], g->F[c], nf, hipMemcpyDeviceToDevice, g->stream));
    }
    for (int s : kStepPsiE) {
        if (g->psi_alt[s] || !g->d_psipml[s]) continue;
        const size_t n = g->psipml_n[s] * g->csize();
        int rc = dalloc(g, &g->psi_alt[s], n);
        if (rc) return rc;
        HIPOK(g, hipMemcpyAsync(g->psi_alt[s], g->d_psipml[s], n, hipMemcpyDeviceToDevice, g->stream));
    }
    return AEFDTD_OK;
}

void free_step_set(aefdtd_grid *g) {
    for (void *&p : g->F_alt) {
        if (p) g->dev_bytes -= g->nalloc() * g->ssize();
        dfree(p);
    }
    for (int s = 0; s < 12; s++) {
        if (g->psi_alt[s]) g->dev_bytes -= g->psipml_n[s] * g->csize();
        dfree(g->psi_alt[s]);
    }
}

// After a one-pass launch the second set holds the current state.
void swap_step_set(aefdtd_grid *g) {
    for (int c = 0; c < 6; c++) std::swap(g->F[c], g->F_alt[c]);
    for (int s : kStepPsiE) std::swap(g->d_psipml[s], g->psi_alt[s]);
}

template <typename S, typename T>
StepOut<S, T> step_out(const aefdtd_grid *g) {
    StepOut<S, T> o;
    Fld<S, T> *f[6] = {&o.Ex, &o.Ey, &o.Ez, &o.Hx, &o.Hy, &o.Hz};
    for (int c = 0; c < 6; c++) {
        *f[c] = fld<S, T>(g, c);
        f[c]->p = (S *)g->F_alt[c];
    }
    o.pEyx = (T *)g->psi_alt[0]; o.pEzx = (T *)g->psi_alt[1];
    o.pExy = (T *)g->psi_alt[4]; o.pEzy = (T *)g->psi_alt[5];
    o.pExz = (T *)g->psi_alt[8]; o.pEyz = (T *)g->psi_alt[9];
    return o;
}

template <typename S, typename T, int NW>
void launch_step_nw(const StepPlan &p, const EArgs<S, T> &a, const HArgs<S, T> &h, const StepOut<S, T> &o, bool en,
                    hipStream_t st) {
    const StepGeom sg{p.ntx, p.nty, p.ntz, p.KZ, p.remap};
    const dim3 grid((unsigned)(p.ntx * p.nty * p.ntz)), blk(BX, NW);
    // LDS: rows [2][NW-1][2][64] and halo column [2][2][64] of T, then the materials
    const size_t lds = (size_t)(4 * (NW - 1) + 4) * BX * sizeof(T) + (size_t)a.nmat * sizeof(MatL<T>);
    if constexpr (std::is_same_v<S, T>) {
        if (en) {
            hipLaunchKernelGGL((k_step<S, T, NW, true>), grid, blk, lds, st, a, h, o, sg);
            return;
        }
    }
    hipLaunchKernelGGL((k_step<S, T, NW, false>), grid, blk, lds, st, a, h, o, sg);
}

// One one-pass launch from the current set into the second, then the swap.
template <typename S, typename T>
void launch_step(aefdtd_grid *g, const StepPlan &p, const EArgs<S, T> &a, const HArgs<S, T> &h, bool en) {
    const StepOut<S, T> o = step_out<S, T>(g);
    hipEvent_t ev;
    prof_begin(g, 0, &ev);
    switch (p.NW) {
        case 4: launch_step_nw<S, T, 4>(p, a, h, o, en, g->stream); break;
        case 8: launch_step_nw<S, T, 8>(p, a, h, o, en, g->stream); break;
        default: launch_step_nw<S, T, 16>(p, a, h, o, en, g->stream); break;
    }
    prof_end(g, ev);
    swap_step_set(g);
}

// One full step of g (and of its TFSF column aux) by the one-pass kernel:
// the column's E update first (the main grid's H update reads its new E),
// the RC pre-pass (unless the last sensor feed ran it), the one-pass launch,
// then the column's H update (the main grid's E update read its old H).
template <typename S, typename T>
int step_T(aefdtd_grid *g, aefdtd_grid *aux, const StepPlan &p) {
    int rc = ensure_osc_tables(g, g->tstep + 1);
    if (rc || (rc = ensure_step_set(g))) return rc;
    if (aux) {
        if ((rc = ensure_osc_tables(aux, aux->tstep + 1)) || (rc = launch_rc<S, T>(aux))) return rc;
        const EArgs<S, T> x = e_args<S, T>(aux);
        hipLaunchKernelGGL((k_update_E_col<S, T>), dim3((aux->Nz + 255) / 256), dim3(256), x.nmat * sizeof(MatL<T>),
                           g->stream, x);
    }
    if (!g->rc_ahead && (rc = launch_rc<S, T>(g))) return rc;
    g->rc_ahead = false;
    EArgs<S, T> a = e_args<S, T>(g);
    const HArgs<S, T> h = h_args<S, T>(g);
    g->en_parts = 0;
    bool en = false;
    if constexpr (std::is_same_v<S, T>) {
        if (g->en_fuse) {
            const int nt = p.ntx * p.nty * p.ntz;
            if (nt > g->en_part_cap) {
                dfree_t(g->d_en_part);
                g->en_part_cap = 0;
                void *q;
                if ((rc = dalloc(g, &q, (size_t)nt * sizeof(double)))) return rc;
                g->d_en_part = (double *)q;
                g->en_part_cap = nt;
            }
            a.en_part = g->d_en_part;
            g->en_parts = nt;
            en = true;
        }
    }
    launch_step<S, T>(g, p, a, h, en);
    if (aux) {
        const HArgs<S, T> x = h_args<S, T>(aux);
        hipLaunchKernelGGL((k_update_H_col<S, T>), dim3((aux->Nz + 255) / 256), dim3(256), 0, g->stream, x);
    }
    HIPOK(g, hipGetLastError());
    g->strip_valid = h.nstrip > 0;
    return AEFDTD_OK;
}

// Times one tile configuration of the E or H pass: reps launches between events.
template <typename S, typename T>
float time_cfg(aefdtd_grid *g, bool e_pass, int cfg, int reps, hipEvent_t ea, hipEvent_t eb) {
    (void)hipEventRecord(ea, g->stream);
    for (int r = 0; r < reps; r++) {
        if (e_pass) {
            EArgs<S, T> a = e_args<S, T>(g);
            dispatch_E<S, T>(g, a, cfg);
        } else {
            HArgs<S, T> a = h_args<S, T>(g);
            dispatch_H<S, T>(g, a, cfg);
        }
    }
    (void)hipEventRecord(eb, g->stream);
    (void)hipEventSynchronize(eb);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ea, eb);
    return ms;
}

// Times one tile plan of the one-pass step: 2 reps launches between events
// (an even number, so that the current set is where it was), per step.
template <typename S, typename T>
float time_step(aefdtd_grid *g, const StepPlan &p, int reps, hipEvent_t ea, hipEvent_t eb) {
    (void)hipEventRecord(ea, g->stream);
    for (int r = 0; r < 2 * reps; r++) {
        const EArgs<S, T> a = e_args<S, T>(g);
        const HArgs<S, T> h = h_args<S, T>(g);
        launch_step<S, T>(g, p, a, h, false);
    }
    (void)hipEventRecord(eb, g->stream);
    (void)hipEventSynchronize(eb);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ea, eb);
    return ms / 2;
}

// Project autotuner: runs every compiled (BY, KZ, remap) tile of the E and H
// passes on the grid itself, keeps the fastest, and restores the grid's state
// (fields, CPML and RC accumulators) from a device snapshot afterwards.
template <typename S, typename T>
int autotune_T(aefdtd_grid *g, int reps, int out[2]) {
    int rc0 = ensure_osc_tables(g, g->tstep + 1);
    if (rc0) return rc0;
    std::vector<std::pair<void *, size_t>> state;
    for (int c = 0; c < 6; c++) state.push_back({g->F[c], g->nalloc() * g->ssize()});
    for (int s = 0; s < 12; s++)
        if (g->d_psipml[s]) state.push_back({g->d_psipml[s], g->psipml_n[s] * g->csize()});
    if (g->d_psi) state.push_back({g->d_psi, (size_t)3 * g->np_max * g->n_disp * sizeof(double2)});
    std::vector<void *> snap(state.size(), nullptr);
    for (size_t q = 0; q < state.size(); q++) {
        int rc = dalloc(g, &snap[q], state[q].second);
        if (rc) {
            for (void *&p : snap) dfree(p);
            return rc;
        }
        HIPOK(g, hipMemcpyAsync(snap[q], state[q].first, state[q].second, hipMemcpyDeviceToDevice, g->stream));
    }
    hipEvent_t ea, eb;
    HIPOK(g, timing_event(&ea));
    HIPOK(g, timing_event(&eb));
    int best[2] = {g->tile_e, g->tile_h};
    float bsum = 0;
    for (int pass = 0; pass < 2; pass++) {
        float bt = 1e30f;
        for (int cfg = 0; cfg < 32; cfg++) {
            time_cfg<S, T>(g, pass == 0, cfg, 1, ea, eb);          // warm the instruction cache
            float ms = time_cfg<S, T>(g, pass == 0, cfg, reps, ea, eb);
            if (ms < bt) { bt = ms; best[pass] = cfg; }
        }
        bsum += bt;
    }
    // every tile of the one-pass step (3D grids) against the best pair of passes
    int best_step = -1;
    if (step_plan(g, 0).ok) {
        int rc = ensure_step_set(g);
        if (rc) {
            for (void *&p : snap) dfree(p);
            return rc;
        }
        for (int code = 0; code < 32; code++) {
            const StepPlan p = step_plan(g, code);
            if (!p.ok) continue;
            time_step<S, T>(g, p, 1, ea, eb);
            const float ms = time_step<S, T>(g, p, reps, ea, eb);
            if (ms < bsum) { bsum = ms; best_step = code; }
        }
    }
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    for (size_t q = 0; q < state.size(); q++) {
        HIPOK(g, hipMemcpyAsync(state[q].first, snap[q], state[q].second, hipMemcpyDeviceToDevice, g->stream));
    }
    HIPOK(g, hipStreamSynchronize(g->stream));
    for (size_t q = 0; q < state.size(); q++) {
        g->dev_bytes -= state[q].second;
        dfree(snap[q]);
    }
    HIPOK(g, hipGetLastError());
    g->strip_valid = false;             // the timing passes rewrote the strip; the fields were restored
    g->tile_e = best[0];
    g->tile_h = best[1];
    g->step_cfg = best_step;
    if (best_step < 0) free_step_set(g);         // the second set only serves the one-pass step
    if (out) { out[0] = best[0]; out[1] = best[1]; }
    return AEFDTD_OK;
}

// DFT coefficients of a sensor for steps [0, n): tapering_E/H * e^{i w step dt}
// with tapering taken from the previous feed (sensors/sensors.cpp:232-243,464-466).
void sensor_coefs(const aefdtd_grid *g, const Sensor &s, int n, std::vector<double2> &out) {
    out.resize((size_t)n * 2 * s.Nl);
    const std::complex<double> I(0.0, 1.0);
    double tapE = 1.0, tapH = 1.0;
    for (int st = 0; st < n; st++) {
        if (st > 0 && st - 1 >= g->Nt - g->Ntap) {
            tapE = s_curve(st - 1, g->Nt - 1.0, g->Nt - g->Ntap);
            tapH = s_curve(st - 1 + 0.5, g->Nt - 1.0, g->Nt - g->Ntap);
        }
        for (int l = 0; l < s.Nl; l++) {
            double w = 2.0 * kPi * kC / s.lambda[l];
            std::complex<double> cE = tapE * std::exp(w * st * g->Dt * I);
            std::complex<double> cH = tapH * std::exp(w * (st + 0.5) * g->Dt * I);
            out[((size_t)st * 2 + 0) * s.Nl + l] = make_double2(cE.real(), cE.imag());
            out[((size_t)st * 2 + 1) * s.Nl + l] = make_double2(cH.real(), cH.imag());
        }
    }
}

int ensure_coefs(aefdtd_grid *g, Sensor &s, int upto) {
    if (upto <= s.ncoef) return AEFDTD_OK;
    int n = std::max(upto, std::max(g->Nt, 2 * s.ncoef));
    std::vector<double2> h;
    sensor_coefs(g, s, n, h);
    void *p;
    int rc = dalloc(g, &p, h.size() * sizeof(double2));
    if (rc) return rc;
    HIPOK(g, hipStreamSynchronize(g->stream));
    dfree_t(s.d_coef);
    s.d_coef = (double2 *)p;
    HIPOK(g, copy_sync(g, s.d_coef, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
    // the matrix-core flush's table: the same coefficients in the arithmetic
    // type, (re, im) interleaved per wavelength and each step's row padded
    // with zeros to a multiple of 32 columns
    const int LP = (2 * s.Nl + 31) / 32 * 32;
    auto table = [&]<typename V>() {
        std::vector<V> m((size_t)n * 2 * LP, (V)0);
        for (int st = 0; st < n; st++)
            for (int e = 0; e < 2; e++)
                for (int l = 0; l < s.Nl; l++) {
                    const double2 v = h[((size_t)st * 2 + e) * s.Nl + l];
                    m[((size_t)st * 2 + e) * LP + 2 * l] = (V)v.x;
                    m[((size_t)st * 2 + e) * LP + 2 * l + 1] = (V)v.y;
                }
        void *q;
        int rq = dalloc(g, &q, m.size() * sizeof(V));
        if (rq) return rq;
        dfree_t(s.d_coefmx);
        s.d_coefmx = q;
        s.LP = LP;
        HIPOK(g, copy_sync(g, s.d_coefmx, m.data(), m.size() * sizeof(V), hipMemcpyHostToDevice));
        return AEFDTD_OK;
    };
    if ((rc = g->precision == AEFDTD_FP64 ? table.template operator()<double>() : table.template operator()<float>()))
        return rc;
    s.ncoef = n;
    return AEFDTD_OK;
}

template <typename T>
int flush_dft(aefdtd_grid *g) {
    if (g->stage_fill == 0) return AEFDTD_OK;
    const int s0 = g->sensor_step - g->stage_fill;
    // one launch per DFT_MAX_SENS sensors: every plane, component and wavelength
    for (size_t q0 = 0; q0 < g->sensors.size(); q0 += DFT_MAX_SENS) {
        const size_t n = std::min<size_t>(DFT_MAX_SENS, g->sensors.size() - q0);
        if constexpr (sizeof(T) == 4) {
            // FP32 arithmetic: the matrix-core flush (4 wavefronts of 32 cells per workgroup)
            DftSetMx set{};
            long maxP = 0;
            for (size_t q = 0; q < n; q++) {
                Sensor *s = g->sensors[q0 + q];
                int rc = ensure_coefs(g, *s, g->sensor_step);
                if (rc) return rc;
                set.d[q] = DftDescMx{s->off, s->P, s->Nl, s->LP, s->ncoef, (const float *)s->d_coefmx, s->d_acc};
                maxP = std::max(maxP, s->P);
            }
            hipEvent_t ev;
            prof_begin(g, 3, &ev);
            hipLaunchKernelGGL(k_dft_mx, dim3((unsigned)((maxP + 127) / 128), (unsigned)(6 * n)), dim3(256), 0,
                               g->stream, (const float *)g->d_stage, g->stage_P, set, s0, g->stage_fill);
            prof_end(g, ev);
            HIPOK(g, hipGetLastError());
            continue;
        }
        // FP64: the FP64 matrix-core flush (8 wavefronts of 16 cells per workgroup)
        DftSetMx64 set{};
        long maxP = 0;
        for (size_t q = 0; q < n; q++) {
            Sensor *s = g->sensors[q0 + q];
            int rc = ensure_coefs(g, *s, g->sensor_step);
            if (rc) return rc;
            set.d[q] = DftDescMx64{s->off, s->P, s->Nl, s->LP, s->ncoef, (const double *)s->d_coefmx, s->d_acc};
            maxP = std::max(maxP, s->P);
        }
        hipEvent_t ev;
        prof_begin(g, 3, &ev);
        hipLaunchKernelGGL(k_dft_mx64, dim3((unsigned)((maxP + 127) / 128), (unsigned)(6 * n)), dim3(512), 0,
                           g->stream, (const double *)g->d_stage, g->stage_P, set, s0, g->stage_fill);
        prof_end(g, ev);
        HIPOK(g, hipGetLastError());
    }
    g->stage_fill = 0;
    return AEFDTD_OK;
}

// First feed: lays the sensors out in one stage pool and uploads their descriptors.
template <typename T>
int prepare_stage(aefdtd_grid *g) {
    if (g->d_stage || g->sensors.empty()) return AEFDTD_OK;
    long Ptot = 0;
    std::vector<SensDesc> sd;
    // planes normal to x get a strip (columns x1 and x1+1) written by the 3D
    // H pass; planes beyond MAX_STRIP distinct columns sample the fields
    const bool strips = g->Nz > 1 && g->Nx > 1;
    for (Sensor *s : g->sensors) {
        s->off = Ptot;
        Ptot += s->P;
        int sp = -1;
        if (strips && (s->type == AEFDTD_NORMAL_X || s->type == AEFDTD_NORMAL_XM)) {
            for (size_t q = 0; q < g->strip_cols.size(); q++)
                if (g->strip_cols[q] == s->x1) sp = (int)q;
            if (sp < 0 && g->strip_cols.size() < (size_t)MAX_STRIP) {
                sp = (int)g->strip_cols.size();
                g->strip_cols.push_back(s->x1);
            }
        }
        sd.push_back({s->type, s->x1, s->y1, s->z1, s->span1, s->span2, s->P, s->off, sp});
        g->stage_maxP = std::max(g->stage_maxP, s->P);
    }
    g->stage_P = Ptot;
    int rc;
    if (!g->strip_cols.empty() &&
        (rc = dalloc(g, &g->d_strip, g->strip_cols.size() * 12 * (size_t)g->Ny * g->Nz * sizeof(T))))
        return rc;
    if ((rc = dalloc(g, &g->d_stage, (size_t)stage_depth(g->precision) * 6 * Ptot * sizeof(T)))) return rc;
    if ((rc = dalloc(g, &g->d_sdesc, sd.size() * sizeof(SensDesc)))) return rc;
    HIPOK(g, copy_sync(g, g->d_sdesc, sd.data(), sd.size() * sizeof(SensDesc), hipMemcpyHostToDevice));
    // the DFT coefficient tables for the whole run now, so that no flush waits
    // for host work and a stream synchronisation
    for (Sensor *s : g->sensors)
        if ((rc = ensure_coefs(g, *s, std::max(g->Nt, 1)))) return rc;
    return AEFDTD_OK;
}

template <typename S, typename T>
int feed_T(aefdtd_grid *g) {
    if (!g->sensors.empty()) {
        int rc;
        if ((rc = prepare_stage<T>(g))) return rc;
        const int nsx = (int)((g->stage_maxP + 255) / 256);
        hipEvent_t ev;
        prof_begin(g, 2, &ev);
        if (g->n_disp > 0 && !g->rc_ahead) {
            // the next E pass's RC pre-pass rides in this launch (fields are those of the H pass)
            const int nrx = (int)((g->n_disp + 255) / 256), ns = nsx * (int)g->sensors.size();
            hipLaunchKernelGGL((k_sample_rc<S, T>), dim3((unsigned)(ns + 3 * nrx)), dim3(256), 0, g->stream,
                               (const SensDesc *)g->d_sdesc, g->stage_fill, g->stage_P, fld<S, T>(g, 0),
                               fld<S, T>(g, 1), fld<S, T>(g, 2), fld<S, T>(g, 3), fld<S, T>(g, 4), fld<S, T>(g, 5),
                               g->Nx, g->Ny, g->Nz, g->px, (T *)g->d_stage,
                               g->strip_valid ? (const T *)g->d_strip : nullptr, nsx, ns, rc_args<S, T>(g), nrx);
            g->rc_ahead = true;
        } else {
            dim3 grid((unsigned)nsx, (unsigned)g->sensors.size(), 1);
            hipLaunchKernelGGL((k_sample<S, T>), grid, dim3(256), 0, g->stream, (const SensDesc *)g->d_sdesc,
                               g->stage_fill, g->stage_P, fld<S, T>(g, 0), fld<S, T>(g, 1), fld<S, T>(g, 2),
                               fld<S, T>(g, 3), fld<S, T>(g, 4), fld<S, T>(g, 5), g->Nx, g->Ny, g->Nz, g->px,
                               (T *)g->d_stage, g->strip_valid ? (const T *)g->d_strip : nullptr);
        }
        prof_end(g, ev);
        HIPOK(g, hipGetLastError());
        g->stage_fill++;
        g->sensor_step++;
        if (g->stage_fill == stage_depth(g->precision)) {
            if ((rc = flush_dft<T>(g))) return rc;
        }
    }
    // single-wavelength field sensors: one accumulation per feed
    // (FieldBlock::threaded_computation, FieldMap::FT_compute)
    for (FieldSensor *s : g->fsensors) {
        const double w = 2.0 * kPi * kC / s->lambda;
        const std::complex<double> I(0.0, 1.0);
        const std::complex<double> cE = std::exp(w * s->step * g->Dt * I);
        const std::complex<double> cH = s->d.block ? std::exp(w * (s->step + 0.5) * g->Dt * I) : cE;
        hipLaunchKernelGGL((k_field_feed<S, T>), dim3((unsigned)((s->d.n + 255) / 256)), dim3(256), 0, g->stream, s->d,
                           fld<S, T>(g, 0), fld<S, T>(g, 1), fld<S, T>(g, 2), fld<S, T>(g, 3), fld<S, T>(g, 4),
                           fld<S, T>(g, 5), g->Nx, g->Ny, g->Nz, g->px, make_double2(cE.real(), cE.imag()),
                           make_double2(cH.real(), cH.imag()));
        HIPOK(g, hipGetLastError());
        s->step++;
    }
    const int t = g->tstep - 1;
    if (t >= 0 && t < g->Nt) {
        // two sums per launch (blockIdx.y)
        for (size_t q = 0; q < g->sums.size(); q += 2) {
            PlaneSumArgs a[2];
            const int n = std::min<int>(2, (int)(g->sums.size() - q));
            for (int u = 0; u < n; u++) {
                const PlaneSum *s = g->sums[q + u];
                a[u] = PlaneSumArgs{s->k, s->inc, (int)s->has_inc, (int)s->subtract, s->zpos,
                                    (const double *)s->d_tap, s->d_out};
            }
            if (n == 1) a[1] = a[0];
            hipLaunchKernelGGL((k_plane_sum<S, T>), dim3(PS_BLOCKS, n), dim3(256), 0, g->stream, fld<S, T>(g, 0),
                               fld<S, T>(g, 1), fld<S, T>(g, 2), g->Nx, g->Ny, g->px, a[0], a[1], g->Dx, g->Dy, g->Dz,
                               t, g->Dt);
            HIPOK(g, hipGetLastError());
            for (int u = 0; u < n; u++) g->sums[q + u]->n = std::max(g->sums[q + u]->n, t + 1);
        }
    }
    return AEFDTD_OK;
}


// ================================================================ batches
// Many independent grids advanced together (the batching form): one launch
// per half step covers every member grid, the per-step sensor samples and
// the end-of-chunk DFT of all members are one launch each, and a chunk of KS
// steps is captured once in a HIP graph and replayed while the membership is
// unchanged.  Members join with an empty DFT stage (a multiple of KS steps
// done) and leave between chunks; a member whose step count ends inside a
// chunk idles for the rest of it.

#define BHIPOK(b, call)                                                                                 \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) {                                                                         \
            (b)->err = std::string(#call) + ": " + hipGetErrorString(e_);                               \
            return e_ == hipErrorOutOfMemory ? AEFDTD_ENOMEM : AEFDTD_EDEVICE;                          \
        }                                                                                               \
    } while (0)

template <typename S, typename T, int BY, int KZ>
void launch_Eb_cfg(int nwg, size_t lds, hipStream_t st, const void *d, const int *wg0, int n, const int *tc,
                   const int *te, int s) {
    hipLaunchKernelGGL((k_update_E_batch<S, T, BY, KZ>), dim3(nwg), dim3(BX, BY, 1), lds, st,
                       (const EArgs<S, T> *)d, wg0, n, tc, te, s);
}
template <typename S, typename T, int BY, int KZ>
void launch_Hb_cfg(int nwg, hipStream_t st, const void *d, const int *wg0, int n, const int *tc, const int *te,
                   int s) {
    hipLaunchKernelGGL((k_update_H_batch<S, T, BY, KZ>), dim3(nwg), dim3(BX, BY, 1), 0, st, (const HArgs<S, T> *)d,
                       wg0, n, tc, te, s);
}

template <typename S, typename T, int BY>
void launch_Eb_by(int kzi, int nwg, size_t lds, hipStream_t st, const void *d, const int *wg0, int n, const int *tc,
                  const int *te, int s) {
    switch (kzi) {
        case 0: launch_Eb_cfg<S, T, BY, 1>(nwg, lds, st, d, wg0, n, tc, te, s); break;
        case 1: launch_Eb_cfg<S, T, BY, 2>(nwg, lds, st, d, wg0, n, tc, te, s); break;
        case 2: launch_Eb_cfg<S, T, BY, 4>(nwg, lds, st, d, wg0, n, tc, te, s); break;
        default: launch_Eb_cfg<S, T, BY, 8>(nwg, lds, st, d, wg0, n, tc, te, s); break;
    }
}
template <typename S, typename T, int BY>
void launch_Hb_by(int kzi, int nwg, hipStream_t st, const void *d, const int *wg0, int n, const int *tc,
                  const int *te, int s) {
    switch (kzi) {
        case 0: launch_Hb_cfg<S, T, BY, 1>(nwg, st, d, wg0, n, tc, te, s); break;
        case 1: launch_Hb_cfg<S, T, BY, 2>(nwg, st, d, wg0, n, tc, te, s); break;
        case 2: launch_Hb_cfg<S, T, BY, 4>(nwg, st, d, wg0, n, tc, te, s); break;
        default: launch_Hb_cfg<S, T, BY, 8>(nwg, st, d, wg0, n, tc, te, s); break;
    }
}

// Time-skewed 2D passes: tile heights (small, large) per arithmetic type and
// depths 2, 4, 8 steps.
template <typename T>
constexpr int tb_ly(int lyi) {
    return sizeof(T) == 8 ? (lyi ? 32 : 16) : (lyi ? 48 : 32);
}
constexpr int kTS[3] = {2, 4, 8};

bool tb_code_valid(int precision, int code) {
    if (code == -1) return true;
    if (code < 0 || code > 11) return false;
    const int lyi = code & 3, tsi = code >> 2;
    if (lyi > 1 || tsi > 2) return false;
    const int ly = precision == AEFDTD_FP64 ? (lyi ? 32 : 16) : (lyi ? 48 : 32);
    return ly - 2 * kTS[tsi] - 1 >= 4;
}

// 0 = TM, 1 = TE, -1 = the grid is stepped by the per-step passes.
int tb_polarisation(const aefdtd_grid *g) {
    if (g->Nz != 1 || g->n_disp || g->tfsf_aux || g->inj_k >= 0 || g->precision == AEFDTD_FP16) return -1;
    const bool *e = g->enable;
    int pol = -1;
    if (!e[0] && !e[1] && e[2] && e[3] && e[4] && !e[5]) pol = 0;
    else if (e[0] && e[1] && !e[2] && !e[3] && !e[4] && e[5]) pol = 1;
    if (pol < 0) return -1;
    for (const OscDev &o : g->osc)
        if ((pol == 0) != (o.comp == 2)) return -1;
    return pol;
}

template <typename S, typename T, int POL, int LY, int TS, bool PML>
void launch_tb_k(int nwg, size_t lds, hipStream_t st, const void *ea, const void *ha, const int *mem, const int *wg0,
                 int nm, const int *tc, const int *te, const void *samp, const int *sbeg, const void *alt,
                 const int4 *rect, int s0) {
    hipLaunchKernelGGL((k_tb2d_batch<S, T, POL, LY, TS, PML>), dim3(nwg), dim3(64, TB_WAVES, 1), lds, st,
                       (const EArgs<S, T> *)ea, (const HArgs<S, T> *)ha, mem, wg0, nm, tc, te,
                       (const SampB<S, T> *)samp, sbeg, (const TbAlt<S, T> *)alt, rect, s0, (s0 / TS) & 1);
}

// The interior tiles of a time-skewed grid: [bx0, bx1) x [by0, by1) of its
// tile grid, the tiles none of whose cells (halo included) lies in a CPML
// slab or on a PEC wall; an empty rectangle when there are none.  E cells of
// the slabs lie in 1..pm-1 and N-pp+1..N-1 (wall at 0), H cells in 0..pm-1
// and N-pp..N-2 (wall at N-1), on axes with any CPML.
int4 tb_interior(const aefdtd_grid *g, int IX, int IY, int LY, int TS) {
    auto range = [&](int N, int pm, int pp, int step, int width, int &b0, int &b1) {
        const int nt = (N + step - 1) / step;
        b0 = 0;
        b1 = nt;
        if (!(pm || pp)) return;
        const int lo = std::max(pm, 1), hi = N - std::max(pp, 1) - 1;
        b0 = nt;
        b1 = 0;
        for (int t = 0; t < nt; t++) {
            const int c0 = t * step - TS, c1 = c0 + width - 1;
            if (c0 >= lo && c1 <= hi) {
                b0 = std::min(b0, t);
                b1 = std::max(b1, t + 1);
            }
        }
        if (b0 >= b1) b0 = b1 = 0;
    };
    int bx0, bx1, by0, by1;
    range(g->Nx, g->pml[0], g->pml[1], IX, TBX, bx0, bx1);
    range(g->Ny, g->pml[2], g->pml[3], IY, LY, by0, by1);
    if (bx0 >= bx1 || by0 >= by1) return make_int4(0, 0, 0, 0);
    return make_int4(bx0, bx1, by0, by1);
}

// The second live-state set of a time-skewed grid (engine-owned, kept until
// the grid is destroyed): field components and CPML psi arrays of `pol`.
const int kTbPsi[2][4] = {{1, 5, 6, 2}, {4, 0, 3, 7}};  // TM: Ezx Ezy Hxy Hyx; TE: Exy Eyx Hzx Hzy

int ensure_tb_alt(aefdtd_grid *g, int pol) {
    if (g->tb_alt_pol == pol) return AEFDTD_OK;
    for (void *&p : g->tb_alt) dfree(p);
    g->dev_bytes -= g->tb_alt_bytes;
    g->tb_alt_bytes = 0;
    g->tb_alt_pol = -1;
    const size_t before = g->dev_bytes;
    for (int c = 0; c < 3; c++) {
        int rc = dalloc(g, &g->tb_alt[c], g->nalloc() * g->ssize());
        if (rc) return rc;
    }
    for (int c = 0; c < 4; c++) {
        int rc = dalloc(g, &g->tb_alt[3 + c], g->psipml_n[kTbPsi[pol][c]] * g->csize());
        if (rc) return rc;
    }
    g->tb_alt_bytes = g->dev_bytes - before;
    g->tb_alt_pol = pol;
    return AEFDTD_OK;
}

template <typename S, typename T, int POL, int LY, int TS, bool PML>
hipError_t tb_attr(size_t lds) {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_tb2d_batch<S, T, POL, LY, TS, PML>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// Calls f.template operator()<POL, LY, TS>() for a polarisation and a tile code.
template <typename T, class F>
void with_tb(int pol, int code, F &&f) {
    constexpr int L0 = tb_ly<T>(0), L1 = tb_ly<T>(1);
    const int lyi = code & 3, tsi = (code >> 2) & 3;
    auto by_ts = [&]<int P, int L>() {
        if (tsi == 0) f.template operator()<P, L, 2>();
        else if (tsi == 1) f.template operator()<P, L, 4>();
        else if constexpr (L - 2 * 8 - 1 >= 4) f.template operator()<P, L, 8>();
    };
    auto by_ly = [&]<int P>() {
        if (lyi) by_ts.template operator()<P, L1>();
        else by_ts.template operator()<P, L0>();
    };
    if (pol == 0) by_ly.template operator()<0>();
    else by_ly.template operator()<1>();
}

}  // namespace

struct aefdtd::Batch {
    int precision = AEFDTD_FP32;
    hipStream_t stream = nullptr;
    std::vector<aefdtd_grid *> members;
    int tile_e = 2 | (2 << 2), tile_h = 2 | (2 << 2);
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // device tables of the current membership
    void *d_ea = nullptr, *d_ha = nullptr, *d_samp = nullptr, *d_dft = nullptr;
    int *d_wgE = nullptr, *d_wgH = nullptr, *d_tcur = nullptr, *d_tend = nullptr;
    int nwgE = 0, nwgH = 0, nsamp = 0, ndft = 0;
    long maxP = 0;
    size_t ldsE = 0;
    // time-skewed 2D TE/TM members: tile code (small/large tile height | depth
    // index << 2), -1 = off; member lists and workgroup prefixes per polarisation
    int tb_code = 1 | (2 << 2);
    // [polarisation][tile class: 0 interior, 1 boundary]
    int *d_mem[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}}, *d_wgT[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int *d_sbeg = nullptr;
    void *d_alt = nullptr;                 // TbAlt per member
    int4 *d_rect = nullptr;                // interior tile rectangle per member
    int nmem[2][2] = {{0, 0}, {0, 0}}, nwgT[2][2] = {{0, 0}, {0, 0}};
    int nstep = 0;                         // members stepped by the per-step passes
    // dispersive members' RC pre-pass (RcArgs per entry, workgroup prefix,
    // member index, skipped step) and members with a plane injection (InjB)
    void *d_rc = nullptr, *d_inj = nullptr;
    int *d_wgR = nullptr, *d_rmem = nullptr, *d_rskip = nullptr;
    int nrc = 0, nwgR = 0, ninj = 0;
    int nsamp_step = 0;                    // leading sample entries that belong to those members
    size_t ldsT = 0;
    std::vector<int> tcur;                 // host mirror of d_tcur
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool prof_on = false;
    double prof_ms = 0;
    long prof_chunks = 0;
    std::string err;
};

namespace {

void batch_free_tables(Batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (void **p : {&b->d_ea, &b->d_ha, &b->d_samp, &b->d_dft, (void **)&b->d_wgE, (void **)&b->d_wgH,
                     (void **)&b->d_tcur, (void **)&b->d_tend, (void **)&b->d_sbeg, &b->d_alt, (void **)&b->d_rect,
                     &b->d_rc, &b->d_inj, (void **)&b->d_wgR, (void **)&b->d_rmem, (void **)&b->d_rskip})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    for (int p = 0; p < 2; p++)
        for (int c = 0; c < 2; c++)
            for (int **q : {&b->d_mem[p][c], &b->d_wgT[p][c]})
                if (*q) { (void)hipFree(*q); *q = nullptr; }
}

// The plane injection values of g on the device (batches read the step's
// values there), [step][dEx, dEy, dHx, dHy].
int ensure_inj_table(aefdtd_grid *g) {
    if (g->d_inj || g->inj[0].empty()) return AEFDTD_OK;
    const size_t n = g->inj[0].size();
    std::vector<double> t(4 * n);
    for (size_t s = 0; s < n; s++)
        for (int c = 0; c < 4; c++) t[4 * s + c] = g->inj[c][s];
    void *p;
    int rc = dalloc(g, &p, t.size() * sizeof(double));
    if (rc) return rc;
    g->d_inj = (double *)p;
    HIPOK(g, copy_sync(g, g->d_inj, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    return AEFDTD_OK;
}

template <typename V>
int batch_upload(Batch *b, void **dst, const std::vector<V> &v) {
    if (v.empty()) return AEFDTD_OK;
    BHIPOK(b, hipMalloc(dst, v.size() * sizeof(V)));
    BHIPOK(b, hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice, b->stream));
    BHIPOK(b, hipStreamSynchronize(b->stream));
    return AEFDTD_OK;
}

// Lays out the device tables of the membership and captures one chunk.
template <typename S, typename T>
int batch_build(Batch *b) {
    BHIPOK(b, hipStreamSynchronize(b->stream));      // the previous graph may still read the tables
    batch_free_tables(b);
    const int n = (int)b->members.size();
    b->tcur.assign(n, 0);
    if (n == 0) {
        b->dirty = false;
        return AEFDTD_OK;
    }
    const bool tb_on = b->tb_code >= 0 && std::is_same_v<S, T>;
    // tile height, depth and inner tile of the time-skewed passes (codes are
    // validated by tb_code_valid; -1 = off, then none of these is used)
    const int ly = tb_on ? tb_ly<T>(b->tb_code & 3) : 0, TS = tb_on ? kTS[(b->tb_code >> 2) & 3] : 0;
    const int IX = TBX - 2 * TS - 1, IY = ly - 2 * TS - 1;
    std::vector<EArgs<S, T>> ea;
    std::vector<HArgs<S, T>> ha;
    std::vector<int> wgE, wgH, tend(n), pol(n, -1), mem[2][2], wgT[2][2], sbeg(2 * n, 0);
    std::vector<int4> rect(n, make_int4(0, 0, 0, 0));
    std::vector<SampB<S, T>> samp;
    std::vector<DftB<T>> dft;
    std::vector<TbAlt<S, T>> alt(n);
    std::vector<RcArgs<S, T>> rca;
    std::vector<int> wgR, rmem, rskip;
    std::vector<aefdtd_grid *> ahead;   // members whose next pre-pass already ran
    std::vector<InjB> inj;
    int totE = 0, totH = 0, totR = 0, totT[2][2] = {{0, 0}, {0, 0}}, maxmat = 1;
    b->maxP = 0;
    for (int r = 0; r < n; r++) {
        aefdtd_grid *g = b->members[r];
        int rc;
        if ((rc = ensure_ready(g)) || (rc = ensure_osc_tables(g, g->Nt)) || (rc = prepare_stage<T>(g))) {
            b->err = g->err;
            return rc;
        }
        if (g->pml_dirty) {
            if ((rc = upload_profiles(g))) { b->err = g->err; return rc; }
            g->pml_dirty = false;
        }
        // a TFSF grid's source column must be stepped in the same launches
        if (g->tfsf_aux && std::find(b->members.begin(), b->members.end(), g->tfsf_aux) == b->members.end()) {
            b->err = "a TFSF grid's source column is not in the batch";
            return AEFDTD_ESTATE;
        }
        if (g->n_disp) {
            rca.push_back(rc_args<S, T>(g));
            wgR.push_back(totR);
            totR += (int)((g->n_disp + 255) / 256);
            rmem.push_back(r);
            // the last sensor feed on its own ran the pre-pass of this step
            // (the pre-pass advances the pole recursions, so the flag is
            // cleared only once the chunk that skips it exists)
            rskip.push_back(g->rc_ahead ? g->tstep : -1);
            if (g->rc_ahead) ahead.push_back(g);
        }
        if (g->inj_k >= 0) {
            if ((rc = ensure_inj_table(g))) { b->err = g->err; return rc; }
            inj.push_back(InjB{r, g->inj_k, (int)g->inj[0].size(), g->d_inj});
        }
        for (Sensor *s : g->sensors)
            if ((rc = ensure_coefs(g, *s, g->Nt))) { b->err = g->err; return rc; }
        if (tb_on) pol[r] = tb_polarisation(g);
        EArgs<S, T> a = e_args<S, T>(g);
        HArgs<S, T> h = h_args<S, T>(g);
        h.nstrip = 0;                       // batch sampling reads the fields
        const int nb = tile_geometry(a, b->tile_e), nh = tile_geometry(h, b->tile_h);
        wgE.push_back(totE);
        wgH.push_back(totH);
        if (pol[r] < 0) {
            totE += (nb + 7) / 8 * 8;
            totH += (nh + 7) / 8 * 8;
        } else {
            const int p = pol[r];
            if ((rc = ensure_tb_alt(g, p))) {
                b->err = g->err;
                return rc;
            }
            for (int c = 0; c < 3; c++) alt[r].f[c] = (S *)g->tb_alt[c];
            for (int c = 0; c < 4; c++) alt[r].p[c] = (T *)g->tb_alt[3 + c];
            rect[r] = tb_interior(g, IX, IY, ly, TS);
            const int nall = ((g->Nx + IX - 1) / IX) * ((g->Ny + IY - 1) / IY);
            const int nin = (rect[r].y - rect[r].x) * (rect[r].w - rect[r].z);
            const int cnt[2] = {nin, nall - nin};
            for (int c = 0; c < 2; c++)
                if (cnt[c] > 0) {
                    mem[p][c].push_back(r);
                    wgT[p][c].push_back(totT[p][c]);
                    totT[p][c] += cnt[c];
                }
        }
        maxmat = std::max(maxmat, a.nmat);
        ea.push_back(a);
        ha.push_back(h);
        b->tcur[r] = g->tstep;
        tend[r] = g->Nt;
    }
    // sample entries: members of the per-step passes first, then the time-skewed ones
    for (int pass = 0; pass < 2; pass++)
        for (int r = 0; r < n; r++) {
            if ((pass == 0) != (pol[r] < 0)) continue;
            aefdtd_grid *g = b->members[r];
            sbeg[2 * r] = (int)samp.size();
            for (Sensor *s : g->sensors) {
                SampB<S, T> q;
                q.d = SensDesc{s->type, s->x1, s->y1, s->z1, s->span1, s->span2, s->P, s->off, -1};
                for (int c = 0; c < 6; c++) q.F[c] = fld<S, T>(g, c);
                q.Nx = g->Nx; q.Ny = g->Ny; q.Nz = g->Nz; q.P = g->px; q.member = r;
                q.Ptot = g->stage_P;
                q.stage = (T *)g->d_stage;
                samp.push_back(q);
                dft.push_back(DftB<T>{(const T *)g->d_stage, g->stage_P, s->off, s->P, s->Nl, r,
                                      (const double2 *)s->d_coef, s->d_acc});
                b->maxP = std::max(b->maxP, s->P);
            }
            sbeg[2 * r + 1] = (int)samp.size();
            if (pass == 0) b->nsamp_step = (int)samp.size();
        }
    wgE.push_back(totE);
    wgH.push_back(totH);
    for (int p = 0; p < 2; p++)
        for (int c = 0; c < 2; c++) {
            wgT[p][c].push_back(totT[p][c]);
            b->nmem[p][c] = (int)mem[p][c].size();
            b->nwgT[p][c] = totT[p][c];
        }
    b->nwgE = totE;
    b->nwgH = totH;
    b->nsamp = (int)samp.size();
    b->ndft = (int)dft.size();
    b->ldsE = maxmat * sizeof(MatL<T>);
    b->ldsT = 3 * (size_t)TBX * ly * sizeof(T);
    int rc;
    if ((rc = batch_upload(b, &b->d_ea, ea)) || (rc = batch_upload(b, &b->d_ha, ha)) ||
        (rc = batch_upload(b, (void **)&b->d_wgE, wgE)) || (rc = batch_upload(b, (void **)&b->d_wgH, wgH)) ||
        (rc = batch_upload(b, (void **)&b->d_tcur, b->tcur)) || (rc = batch_upload(b, (void **)&b->d_tend, tend)) ||
        (rc = batch_upload(b, &b->d_samp, samp)) || (rc = batch_upload(b, &b->d_dft, dft)) ||
        (rc = batch_upload(b, (void **)&b->d_sbeg, sbeg)) || (rc = batch_upload(b, &b->d_alt, alt)) ||
        (rc = batch_upload(b, (void **)&b->d_rect, rect)))
        return rc;
    wgR.push_back(totR);
    b->nrc = (int)rca.size();
    b->nwgR = totR;
    b->ninj = (int)inj.size();
    if (b->nrc && ((rc = batch_upload(b, &b->d_rc, rca)) || (rc = batch_upload(b, (void **)&b->d_wgR, wgR)) ||
                   (rc = batch_upload(b, (void **)&b->d_rmem, rmem)) ||
                   (rc = batch_upload(b, (void **)&b->d_rskip, rskip))))
        return rc;
    if (b->ninj && (rc = batch_upload(b, &b->d_inj, inj))) return rc;
    for (int p = 0; p < 2; p++)
        for (int c = 0; c < 2; c++)
            if (b->nmem[p][c]) {
                if ((rc = batch_upload(b, (void **)&b->d_mem[p][c], mem[p][c])) ||
                    (rc = batch_upload(b, (void **)&b->d_wgT[p][c], wgT[p][c])))
                    return rc;
                hipError_t ae = hipSuccess;
                if constexpr (std::is_same_v<S, T>)
                    with_tb<T>(p, b->tb_code, [&]<int P, int L, int TSK>() {
                        ae = c ? tb_attr<S, T, P, L, TSK, true>(b->ldsT) : tb_attr<S, T, P, L, TSK, false>(b->ldsT);
                    });
                BHIPOK(b, ae);
            }
    // one chunk: the time-skewed blocks, KS x (E, H, samples) of the other
    // members, the DFT flush, the step advance
    // relaxed: other threads (the request server's builders) keep making
    // allocation and copy calls on their own streams while this one captures
    BHIPOK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    if constexpr (std::is_same_v<S, T>) {
        for (int s0 = 0; tb_on && s0 < KS; s0 += TS)
            for (int p = 0; p < 2; p++)
                for (int c = 0; c < 2; c++)
                    if (b->nmem[p][c])
                        with_tb<T>(p, b->tb_code, [&]<int P, int L, int TSK>() {
                            auto go = [&]<bool PML>() {
                                launch_tb_k<S, T, P, L, TSK, PML>(b->nwgT[p][c], b->ldsT, b->stream, b->d_ea, b->d_ha,
                                                                  b->d_mem[p][c], b->d_wgT[p][c], b->nmem[p][c],
                                                                  b->d_tcur, b->d_tend, b->d_samp, b->d_sbeg,
                                                                  b->d_alt, b->d_rect, s0);
                            };
                            if (c) go.template operator()<true>();
                            else go.template operator()<false>();
                        });
    }
    const int ebi = b->tile_e & 3, ekz = (b->tile_e >> 2) & 3, hbi = b->tile_h & 3, hkz = (b->tile_h >> 2) & 3;
    for (int s = 0; s < KS && totE > 0; s++) {
        if (b->ninj)
            hipLaunchKernelGGL((k_batch_inj<S, T>), dim3((unsigned)((b->ninj + 63) / 64)), dim3(64), 0, b->stream,
                               (EArgs<S, T> *)b->d_ea, (HArgs<S, T> *)b->d_ha, (const InjB *)b->d_inj, b->ninj,
                               b->d_tcur, b->d_tend, s);
        if (b->nrc)
            hipLaunchKernelGGL((k_rc_batch<S, T>), dim3((unsigned)b->nwgR, 3), dim3(256), 0, b->stream,
                               (const RcArgs<S, T> *)b->d_rc, b->d_wgR, b->nrc, b->d_rmem, b->d_rskip, b->d_tcur,
                               b->d_tend, s);
        switch (ebi) {
            case 0: launch_Eb_by<S, T, 1>(ekz, totE, b->ldsE, b->stream, b->d_ea, b->d_wgE, n, b->d_tcur, b->d_tend, s); break;
            case 1: launch_Eb_by<S, T, 2>(ekz, totE, b->ldsE, b->stream, b->d_ea, b->d_wgE, n, b->d_tcur, b->d_tend, s); break;
            case 2: launch_Eb_by<S, T, 4>(ekz, totE, b->ldsE, b->stream, b->d_ea, b->d_wgE, n, b->d_tcur, b->d_tend, s); break;
            default: launch_Eb_by<S, T, 8>(ekz, totE, b->ldsE, b->stream, b->d_ea, b->d_wgE, n, b->d_tcur, b->d_tend, s); break;
        }
        switch (hbi) {
            case 0: launch_Hb_by<S, T, 1>(hkz, totH, b->stream, b->d_ha, b->d_wgH, n, b->d_tcur, b->d_tend, s); break;
            case 1: launch_Hb_by<S, T, 2>(hkz, totH, b->stream, b->d_ha, b->d_wgH, n, b->d_tcur, b->d_tend, s); break;
            case 2: launch_Hb_by<S, T, 4>(hkz, totH, b->stream, b->d_ha, b->d_wgH, n, b->d_tcur, b->d_tend, s); break;
            default: launch_Hb_by<S, T, 8>(hkz, totH, b->stream, b->d_ha, b->d_wgH, n, b->d_tcur, b->d_tend, s); break;
        }
        if (b->nsamp_step)
            hipLaunchKernelGGL((k_sample_batch<S, T>), dim3((unsigned)((b->maxP + 255) / 256), (unsigned)b->nsamp_step),
                               dim3(256), 0, b->stream, (const SampB<S, T> *)b->d_samp, b->d_tcur, b->d_tend, s);
    }
    if (b->ndft)
        hipLaunchKernelGGL((k_dft_batch<T>), dim3((unsigned)((b->maxP + 255) / 256), 6, (unsigned)b->ndft), dim3(256),
                           0, b->stream, (const DftB<T> *)b->d_dft, b->d_tcur, b->d_tend);
    hipLaunchKernelGGL(k_batch_advance, dim3((n + 255) / 256), dim3(256), 0, b->stream, b->d_tcur, b->d_tend, n);
    hipError_t le = hipGetLastError();
    hipError_t ce = hipStreamEndCapture(b->stream, &b->graph);
    BHIPOK(b, le);
    BHIPOK(b, ce);
    BHIPOK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    for (aefdtd_grid *g : ahead) g->rc_ahead = false;
    b->dirty = false;
    return AEFDTD_OK;
}

template <class F>
int with_types_p(int precision, F &&f) {
    if (precision == AEFDTD_FP64) return f.template operator()<double, double>();
    if (precision == AEFDTD_FP32) return f.template operator()<float, float>();
    return f.template operator()<_Float16, float>();
}

}  // namespace

namespace aefdtd {

int batch_create(Batch **out, int precision) {
    *out = nullptr;
    if (precision < AEFDTD_FP64 || precision > AEFDTD_FP16) return AEFDTD_EINVAL;
    auto *b = new Batch();
    b->precision = precision;
    b->tb_code = precision == AEFDTD_FP16 ? -1 : (1 | (1 << 2));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        timing_event(&b->ev0) != hipSuccess || timing_event(&b->ev1) != hipSuccess) {
        set_global_error("batch stream/event creation failed");
        batch_destroy(b);
        return AEFDTD_EDEVICE;
    }
    *out = b;
    return AEFDTD_OK;
}

void batch_destroy(Batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free_tables(b);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *batch_error(const Batch *b) { return b->err.c_str(); }

int batch_add(Batch *b, aefdtd_grid *g) {
    if (!b || !g) return AEFDTD_EINVAL;
    auto bad = [&](const char *why) {
        b->err = why;
        return AEFDTD_ESTATE;
    };
    if (g->precision != b->precision) return bad("grid precision differs from the batch's");
    if (std::find(b->members.begin(), b->members.end(), g) != b->members.end()) return bad("grid already in the batch");
    if (!g->sums.empty()) return bad("plane sums (normal incidence) are not batched");
    if (g->tstep >= g->Nt) return bad("grid has no steps left");
    if (g->tstep % KS != 0) return bad("a grid joins a batch after a multiple of 64 steps");
    int rc = ensure_ready(g);
    if (rc) { b->err = g->err; return rc; }
    // the batch stages 64 steps per chunk from an empty stage: flush what the
    // grid staged on its own (its stage is deeper than a chunk)
    if ((rc = with_ctype(g, [&]<typename T>() { return flush_dft<T>(g); }))) { b->err = g->err; return rc; }
    BHIPOK(b, hipStreamSynchronize(g->stream));
    BHIPOK(b, hipStreamSynchronize(b->stream));
    g->strip_valid = false;             // batch sampling reads the fields, not the strip
    b->members.push_back(g);
    b->dirty = true;
    return AEFDTD_OK;
}

int batch_remove(Batch *b, aefdtd_grid *g) {
    if (!b || !g) return AEFDTD_EINVAL;
    auto it = std::find(b->members.begin(), b->members.end(), g);
    if (it == b->members.end()) {
        b->err = "grid is not in the batch";
        return AEFDTD_EINVAL;
    }
    BHIPOK(b, hipStreamSynchronize(b->stream));
    b->members.erase(it);
    b->dirty = true;
    return AEFDTD_OK;
}

int batch_members(const Batch *b) { return (int)b->members.size(); }

hipStream_t batch_stream(const Batch *b) { return b->stream; }

int batch_step(Batch *b, int chunks) {
    if (!b || chunks < 0) return AEFDTD_EINVAL;
    if (b->members.empty() || chunks == 0) return AEFDTD_OK;
    if (b->prof_on) BHIPOK(b, hipEventRecord(b->ev0, b->stream));
    for (int c = 0; c < chunks; c++) {
        if (b->dirty) {
            int rc = with_types_p(b->precision, [&]<typename S, typename T>() { return batch_build<S, T>(b); });
            if (rc) return rc;
        }
        BHIPOK(b, hipGraphLaunch(b->exec, b->stream));
        for (size_t r = 0; r < b->members.size(); r++) {
            aefdtd_grid *g = b->members[r];
            b->tcur[r] = std::min(b->tcur[r] + KS, g->Nt);
        }
        if (b->precision == AEFDTD_FP16) {
            // FP16 range control after every chunk (every 64 steps, as a run
            // stepped on its own checks): a changed exponent is baked into the
            // chunk's tables, so the next chunk is built again
            BHIPOK(b, hipStreamSynchronize(b->stream));
            for (size_t r = 0; r < b->members.size(); r++) {
                aefdtd_grid *g = b->members[r];
                g->tstep = b->tcur[r];
                if (g->tstep % KS != 0) continue;         // a run alone checks at multiples of 64 steps
                bool ch = false;
                int rc = fp16_range(g, &ch);
                if (rc) { b->err = g->err; return rc; }
                if (ch) {
                    BHIPOK(b, hipStreamSynchronize(g->stream));
                    b->dirty = true;
                }
            }
        }
    }
    if (b->prof_on) {
        BHIPOK(b, hipEventRecord(b->ev1, b->stream));
        BHIPOK(b, hipEventSynchronize(b->ev1));
        float ms = 0;
        BHIPOK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->prof_ms += ms;
        b->prof_chunks += chunks;
    }
    // host state of the members: steps, feeds and an empty stage
    for (size_t r = 0; r < b->members.size(); r++) {
        aefdtd_grid *g = b->members[r];
        g->tstep = b->tcur[r];
        g->sensor_step = b->tcur[r];
        g->stage_fill = 0;
    }
    return AEFDTD_OK;
}

int batch_synchronize(Batch *b) {
    BHIPOK(b, hipStreamSynchronize(b->stream));
    return AEFDTD_OK;
}

int batch_set_tiles(Batch *b, int tile_e, int tile_h) {
    if (tile_e < 0 || tile_e > 31 || tile_h < 0 || tile_h > 31) return AEFDTD_EINVAL;
    b->tile_e = tile_e;
    b->tile_h = tile_h;
    b->dirty = true;
    return AEFDTD_OK;
}

int batch_set_skew(Batch *b, int code) {
    if (!tb_code_valid(b->precision, code)) return AEFDTD_EINVAL;
    b->tb_code = code;
    b->dirty = true;
    return AEFDTD_OK;
}

int batch_skew(const Batch *b) { return b->tb_code; }

int batch_skewed_members(const Batch *b) {
    int n = 0;
    for (aefdtd_grid *g : b->members) n += tb_polarisation(g) >= 0;
    return b->tb_code >= 0 && b->precision != AEFDTD_FP16 ? n : 0;
}

void batch_profile(Batch *b, int on, double *ms, long *chunks) {
    if (ms) *ms = b->prof_ms;
    if (chunks) *chunks = b->prof_chunks;
    if (on >= 0) {
        b->prof_on = on != 0;
        if (on) { b->prof_ms = 0; b->prof_chunks = 0; }
    }
}

}  // namespace aefdtd

namespace {
}  // namespace

// ================================================================= C ABI
extern "C" {

const char *aefdtd_last_error(const aefdtd_grid *g) {
    if (g) return g->err.c_str();
    return g_create_err.c_str();
}

int aefdtd_create(aefdtd_grid **out, int Nx, int Ny, int Nz, int Nt, double Dx, double Dy, double Dz, double Dt,
                  int mode, const int pml[6], const int pad[6], int precision) {
    auto cfail = [&](int code, const std::string &m) {
        std::lock_guard<std::mutex> lk(g_create_mu);
        g_create_err = m;
        return code;
    };
    if (!out) return cfail(AEFDTD_EINVAL, "null output handle");
    *out = nullptr;
    if (Nx < 1 || Ny < 1 || Nz < 1 || Nt < 0) return cfail(AEFDTD_EINVAL, "bad grid dimensions");
    if (!(Dx > 0) || !(Dy > 0) || !(Dz > 0) || !(Dt > 0)) return cfail(AEFDTD_EINVAL, "bad discretisation");
    if (precision != AEFDTD_FP64 && precision != AEFDTD_FP32 && precision != AEFDTD_FP16)
        return cfail(AEFDTD_EINVAL, "bad precision");
    if (mode != AEFDTD_MODE_NORM && mode != AEFDTD_MODE_CUSTOM) return cfail(AEFDTD_EINVAL, "unsupported grid mode");
    for (int f = 0; f < 6; f++)
        if (pml[f] < 0 || pad[f] < 0) return cfail(AEFDTD_EINVAL, "negative pml/pad");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return cfail(AEFDTD_EDEVICE, "no HIP device visible");
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return cfail(AEFDTD_EDEVICE, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return cfail(AEFDTD_EDEVICE, std::string("engine is built for gfx950, device is ") + prop.gcnArchName);
    auto *g = new aefdtd_grid();
    g->device = dev;
    g->precision = precision;
    if (precision == AEFDTD_FP16)
        for (int c = 0; c < 6; c++) g->scale_exp[c] = c < 3 ? 8 : 16;
    g->mode = mode;
    g->Nt = Nt;
    g->Dx = Dx; g->Dy = Dy; g->Dz = Dz; g->Dt = Dt;
    for (int f = 0; f < 6; f++) {
        g->pml[f] = pml[f]; g->pad[f] = pad[f];
        g->pml_kappa[f] = 1.0; g->pml_sigma[f] = 0.0; g->pml_alpha[f] = 0.0;
    }
    // padding rules of core/fdtd_core.cpp:212-243 and extension of :245-251
    for (int a = 0; a < 3; a++)
        if (pml[2 * a] || pml[2 * a + 1]) {
            g->pad[2 * a] = std::max(5, g->pad[2 * a]);
            g->pad[2 * a + 1] = std::max(5, g->pad[2 * a + 1]);
        }
    g->Nx_s = Nx; g->Ny_s = Ny; g->Nz_s = Nz;
    g->Nx = Nx + pml[0] + g->pad[0] + g->pad[1] + pml[1];
    g->Ny = Ny + pml[2] + g->pad[2] + g->pad[3] + pml[3];
    g->Nz = Nz + pml[4] + g->pad[4] + g->pad[5] + pml[5];
    g->xs_s = pml[0] + g->pad[0]; g->xs_e = g->xs_s + Nx;
    g->ys_s = pml[2] + g->pad[2]; g->ys_e = g->ys_s + Ny;
    g->zs_s = pml[4] + g->pad[4]; g->zs_e = g->zs_s + Nz;
    // row pitch of the device cell arrays: rows of 3D grids start on 64-cell
    // (256-byte or wider) boundaries
    g->px = (g->Nz > 1 && g->Nx > 1) ? (g->Nx + 63) / 64 * 64 : g->Nx;
    if (g->nalloc() >= ((size_t)1 << 31)) {     // kernels index cells with 32-bit offsets
        delete g;
        return cfail(AEFDTD_EINVAL, "grid too large (2^31 cells or more)");
    }
    // basic_differentials_compute (core/fdtd_core.cpp:1137-1148)
    g->dtm = Dt / kMu0;
    g->dtdex = Dt / (Dx * kE0); g->dtdey = Dt / (Dy * kE0); g->dtdez = Dt / (Dz * kE0);
    g->dtdmx = Dt / (Dx * kMu0); g->dtdmy = Dt / (Dy * kMu0); g->dtdmz = Dt / (Dz * kMu0);
    size_t N = g->nalloc();
    g->matsgrid.assign(g->ncell(), 0);
    g->mats.assign(1, HostMat());
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return cfail(AEFDTD_EDEVICE, "hipStreamCreate failed");
    }
    for (int c = 0; c < 6; c++) {
        if (dalloc(g, &g->F[c], N * g->ssize()) || hipMemsetAsync(g->F[c], 0, N * g->ssize(), g->stream) != hipSuccess) {
            std::string e = g->err.empty() ? "field allocation failed" : g->err;
            aefdtd_destroy(g);
            return cfail(AEFDTD_ENOMEM, e);
        }
    }
    if (alloc_pml_state(g)) {
        std::string e = g->err;
        aefdtd_destroy(g);
        return cfail(AEFDTD_ENOMEM, e);
    }
    *out = g;
    return AEFDTD_OK;
}

int aefdtd_destroy(aefdtd_grid *g) {
    if (!g) return AEFDTD_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    prof_collect(g);
    for (int c = 0; c < 6; c++) dfree(g->F[c]);
    for (void *&p : g->F_alt) dfree(p);
    for (void *&p : g->psi_alt) dfree(p);
    dfree_t(g->d_mat);
    dfree(g->d_mtab);
    dfree_t(g->d_poles);
    dfree_t(g->d_slot);
    dfree_t(g->d_psi);
    dfree_t(g->d_cells);
    dfree(g->d_corr);
    dfree(g->d_prof);
    for (int s = 0; s < 12; s++) dfree(g->d_psipml[s]);
    dfree(g->d_stage);
    dfree(g->d_sdesc);
    dfree(g->d_strip);
    dfree_t(g->d_energy);
    dfree_t(g->d_en_part);
    dfree_t(g->d_absmax);
    dfree_t(g->d_inj);
    dfree(g->d_osc);
    for (void *&p : g->tb_alt) dfree(p);
    for (OscDev &o : g->osc) {
        void *p = (void *)o.val;
        dfree(p);
    }
    for (Sensor *s : g->sensors) {
        dfree_t(s->d_acc);
        dfree_t(s->d_coef);
        dfree_t(s->d_coefmx);
        delete s;
    }
    for (FieldSensor *s : g->fsensors) {
        dfree_t(s->d.acc);
        delete s;
    }
    for (PlaneSum *s : g->sums) {
        dfree_t(s->d_tap);
        dfree_t(s->d_out);
        delete s;
    }
    if (g->stream && g->own_stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return AEFDTD_OK;
}

int aefdtd_layout(const aefdtd_grid *g, int out[16]) {
    if (!g || !out) return AEFDTD_EINVAL;
    int v[16] = {g->Nx, g->Ny, g->Nz, g->xs_s, g->xs_e, g->ys_s, g->ys_e, g->zs_s, g->zs_e,
                 g->pad[0], g->pad[1], g->pad[2], g->pad[3], g->pad[4], g->pad[5], g->tstep};
    std::memcpy(out, v, sizeof(v));
    return AEFDTD_OK;
}

int aefdtd_set_pml(aefdtd_grid *g, int face, double kap, double sig, double alp) {   // fdtd_pml.cpp:457-497
    if (!g) return AEFDTD_EINVAL;
    if (face < 0 || face > 5) return fail(g, AEFDTD_EINVAL, "bad face");
    double D = face < 2 ? g->Dx : (face < 4 ? g->Dy : g->Dz);
    g->pml_kappa[face] = kap;
    g->pml_sigma[face] = pml_sigma_scale(D) * sig;
    g->pml_alpha[face] = alp;
    g->pml_dirty = true;
    return AEFDTD_OK;
}

int aefdtd_set_tapering(aefdtd_grid *g, int Ntap) {
    if (!g) return AEFDTD_EINVAL;
    if (Ntap < 0) return fail(g, AEFDTD_EINVAL, "negative tapering");
    g->Ntap = Ntap;
    return AEFDTD_OK;
}

int aefdtd_set_matsgrid(aefdtd_grid *g, const uint32_t *ids, int nx, int ny, int nz) {   // fdtd_utils.cpp:572-796
    if (!g || !ids) return AEFDTD_EINVAL;
    if (nx != g->Nx_s || ny != g->Ny_s || nz != g->Nz_s) return fail(g, AEFDTD_EINVAL, "matsgrid size mismatch");
    aefdtd_grid &G = *g;
    auto M = [&](int i, int j, int k) -> uint32_t & {
        return G.matsgrid[(size_t)i + (size_t)G.Nx * ((size_t)j + (size_t)G.Ny * k)];
    };
    for (int k = G.zs_s; k < G.zs_e; k++)
        for (int j = G.ys_s; j < G.ys_e; j++)
            for (int i = G.xs_s; i < G.xs_e; i++)
                M(i, j, k) = ids[(size_t)(i - G.xs_s) + (size_t)nx * ((size_t)(j - G.ys_s) + (size_t)ny * (k - G.zs_s))];
    for (int k = G.zs_s; k < G.zs_e; k++)
        for (int j = G.ys_s; j < G.ys_e; j++) {
            for (int i = 0; i < G.xs_s; i++) M(i, j, k) = M(G.xs_s, j, k);
            for (int i = G.xs_e; i < G.Nx; i++) M(i, j, k) = M(G.xs_e - 1, j, k);
        }
    for (int k = G.zs_s; k < G.zs_e; k++) {
        for (int j = 0; j < G.ys_s; j++)
            for (int i = 0; i < G.Nx; i++) M(i, j, k) = M(i, G.ys_s, k);
        for (int j = G.ys_e; j < G.Ny; j++)
            for (int i = 0; i < G.Nx; i++) M(i, j, k) = M(i, G.ys_e - 1, k);
    }
    for (int k = 0; k < G.zs_s; k++)
        for (int j = 0; j < G.Ny; j++)
            for (int i = 0; i < G.Nx; i++) M(i, j, k) = M(i, j, G.zs_s);
    for (int k = G.zs_e; k < G.Nz; k++)
        for (int j = 0; j < G.Ny; j++)
            for (int i = 0; i < G.Nx; i++) M(i, j, k) = M(i, j, G.zs_e - 1);
    unsigned nmat = G.Nmat;
    for (uint32_t v : G.matsgrid) nmat = std::max(nmat, v + 1);
    if (nmat > 256) return fail(g, AEFDTD_EINVAL, "more than 256 material indices");
    G.Nmat = nmat;
    G.mats.assign(nmat, HostMat());
    G.ready = false;
    return AEFDTD_OK;
}

int aefdtd_set_material(aefdtd_grid *g, unsigned int index, const aefdtd_material *m) {   // fdtd_utils.cpp:810-820
    if (!g || !m) return AEFDTD_EINVAL;
    if (index >= g->Nmat) return AEFDTD_OK;
    bool present = false;
    for (uint32_t v : g->matsgrid)
        if (v == index) { present = true; break; }
    if (!present) return AEFDTD_OK;
    material_coefficients(*m, g->Dx, g->Dy, g->Dz, g->Dt, g->mats[index]);
    g->ready = false;
    return AEFDTD_OK;
}

int aefdtd_disable_fields(aefdtd_grid *g, const int *fields, int n) {
    if (!g) return AEFDTD_EINVAL;
    for (int i = 0; i < n; i++)
        if (fields[i] >= 0 && fields[i] < 6) g->enable[fields[i]] = false;
    return AEFDTD_OK;
}

int aefdtd_bootstrap(aefdtd_grid *g) {     // core/fdtd_core.cpp:1095-1135
    if (!g) return AEFDTD_EINVAL;
    int rc = alloc_pml_state(g);
    if (rc) return rc;
    g->ready = false;
    return ensure_ready(g);
}

int aefdtd_update_E(aefdtd_grid *g) {
    if (!g) return AEFDTD_EINVAL;
    int rc = ensure_ready(g);
    if (rc) return rc;
    if (g->tstep == 0 || g->pml_dirty) {     // pml_coeff_calc at tstep==0 (core/fdtd_core.cpp:1165)
        if ((rc = upload_profiles(g))) return rc;
        g->pml_dirty = false;
    }
    if (g->tfsf_aux && !g->tfsf_aux->ready) return fail(g, AEFDTD_ESTATE, "TFSF auxiliary grid not set up");
    return with_types(g, [&]<typename S, typename T>() { return launch_E<S, T>(g); });
}

int aefdtd_update_H(aefdtd_grid *g) {
    if (!g) return AEFDTD_EINVAL;
    int rc = ensure_ready(g);
    if (rc) return rc;
    rc = with_types(g, [&]<typename S, typename T>() { return launch_H<S, T>(g); });
    if (rc) return rc;
    g->tstep += 1;
    return AEFDTD_OK;
}

}  // extern "C"

namespace aefdtd {

// A grid and its TFSF column stepped together: the column's E and H updates
// ride in the grid's E and H launches (two launches per step instead of four).
// The column must be g's attached TFSF grid on g's stream; otherwise the two
// are stepped one after the other as through the C ABI.
bool pairable(const aefdtd_grid *g, const aefdtd_grid *aux) {
    return aux && g->tfsf_aux == aux && is_column(aux) && !is_column(g) && g->Nz > 1 && aux->stream == g->stream &&
           aux->precision == g->precision;
}

int update_E_pair(aefdtd_grid *g, aefdtd_grid *aux) {
    if (!g || !aux) return AEFDTD_EINVAL;
    if (!pairable(g, aux)) {
        int rc = aefdtd_update_E(aux);
        return rc ? rc : aefdtd_update_E(g);
    }
    for (aefdtd_grid *q : {aux, g}) {
        int rc = ensure_ready(q);
        if (rc) return rc;
        if (q->tstep == 0 || q->pml_dirty) {
            if ((rc = upload_profiles(q))) return rc;
            q->pml_dirty = false;
        }
    }
    return with_types(g, [&]<typename S, typename T>() { return launch_E<S, T>(g, aux); });
}

int fp16_range(aefdtd_grid *g, bool *changed) {
    if (changed) *changed = false;
    if (!g || g->precision != AEFDTD_FP16) return AEFDTD_OK;
    int rc = ensure_ready(g);
    if (rc) return rc;
    if (!g->d_absmax) {
        void *p;
        if ((rc = dalloc(g, &p, 6 * sizeof(unsigned)))) return rc;
        g->d_absmax = (unsigned *)p;
    }
    FieldPtrs fp;
    for (int c = 0; c < 6; c++) fp.f[c] = g->F[c];
    const size_t n = g->nalloc();
    const unsigned nb = (unsigned)std::min<size_t>(1024, (n + 255) / 256);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, g->stream, g->d_absmax, 6);
    hipLaunchKernelGGL((k_absmax<_Float16>), dim3(nb, 6), dim3(256), 0, g->stream, fp, n, g->d_absmax);
    HIPOK(g, hipGetLastError());
    unsigned bits[6];
    HIPOK(g, copy_sync(g, bits, g->d_absmax, sizeof(bits), hipMemcpyDeviceToHost));
    static const char *names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
    for (int c = 0; c < 6; c++) {
        if (!g->enable[c]) continue;
        if (bits[c] >= 0x7f800000u)
            return fail(g, AEFDTD_ERANGE, std::string("FP16 storage overflow: ") + names[c] + " holds a non-finite "
                                          "value at step " + std::to_string(g->tstep) + " (storage exponent " +
                                          std::to_string(g->scale_exp[c]) + ")");
        float m;
        std::memcpy(&m, &bits[c], sizeof(m));
        if (m > 8192.0f) {
            const int d = 12 - (int)std::ceil(std::log2((double)m));
            hipLaunchKernelGGL((k_rescale<_Float16>), dim3(nb), dim3(256), 0, g->stream, (_Float16 *)g->F[c], n,
                               (float)std::ldexp(1.0, d));
            HIPOK(g, hipGetLastError());
            g->scale_exp[c] += d;
            if (changed) *changed = true;
        }
    }
    return AEFDTD_OK;
}

int update_H_pair(aefdtd_grid *g, aefdtd_grid *aux) {
    if (!g || !aux) return AEFDTD_EINVAL;
    if (!pairable(g, aux)) {
        int rc = aefdtd_update_H(aux);
        return rc ? rc : aefdtd_update_H(g);
    }
    int rc = ensure_ready(aux);
    if (rc || (rc = ensure_ready(g))) return rc;
    rc = with_types(g, [&]<typename S, typename T>() { return launch_H<S, T>(g, aux); });
    if (rc) return rc;
    aux->tstep += 1;
    g->tstep += 1;
    return AEFDTD_OK;
}

int step_pair(aefdtd_grid *g, aefdtd_grid *aux) {
    if (!g) return AEFDTD_EINVAL;
    const StepPlan p = step_plan(g, g->step_cfg);
    // the one-pass kernel needs the TFSF column stepped with the grid
    if (!p.ok || (aux ? !pairable(g, aux) : g->tfsf_aux != nullptr)) {
        int rc;
        if (aux) {
            if ((rc = update_E_pair(g, aux))) return rc;
            return update_H_pair(g, aux);
        }
        if ((rc = aefdtd_update_E(g))) return rc;
        return aefdtd_update_H(g);
    }
    for (aefdtd_grid *q : {aux, g}) {
        if (!q) continue;
        int rc = ensure_ready(q);
        if (rc) return rc;
        if (q->tstep == 0 || q->pml_dirty) {     // pml_coeff_calc at tstep==0 (core/fdtd_core.cpp:1165)
            if ((rc = upload_profiles(q))) return rc;
            q->pml_dirty = false;
        }
    }
    int rc = with_types(g, [&]<typename S, typename T>() { return step_T<S, T>(g, aux, p); });
    if (rc) return rc;
    if (aux) aux->tstep += 1;
    g->tstep += 1;
    return AEFDTD_OK;
}

}  // namespace aefdtd

extern "C" int aefdtd_step(aefdtd_grid *g, int n) {
    if (!g || n < 0) return AEFDTD_EINVAL;
    for (int s = 0; s < n; s++) {
        int rc = aefdtd::step_pair(g, nullptr);
        if (rc) return rc;
    }
    return AEFDTD_OK;
}

extern "C" int aefdtd_get_step_tiles(aefdtd_grid *g, int *cfg) {
    if (!g || !cfg) return AEFDTD_EINVAL;
    *cfg = g->step_cfg;
    return AEFDTD_OK;
}

extern "C" int aefdtd_set_step_tiles(aefdtd_grid *g, int cfg) {
    if (!g || cfg < -1 || cfg > 31 || (cfg >= 0 && (cfg & 3) == 3)) return AEFDTD_EINVAL;
    g->step_cfg = cfg;
    return AEFDTD_OK;
}

extern "C" {

int aefdtd_reset_fields(aefdtd_grid *g) {   // core/fdtd_utils.cpp:855-881
    if (!g) return AEFDTD_EINVAL;
    g->strip_valid = false;
    g->rc_ahead = false;
    HIPOK(g, hipStreamSynchronize(g->stream));
    free_step_set(g);                   // made again from the zeroed state at the next one-pass step
    size_t N = g->nalloc();
    for (int c = 0; c < 6; c++) HIPOK(g, hipMemsetAsync(g->F[c], 0, N * g->ssize(), g->stream));
    for (int s = 0; s < 12; s++)
        if (g->d_psipml[s]) HIPOK(g, hipMemsetAsync(g->d_psipml[s], 0, g->psipml_n[s] * g->csize(), g->stream));
    if (g->d_psi) HIPOK(g, hipMemsetAsync(g->d_psi, 0, (size_t)3 * g->np_max * g->n_disp * sizeof(double2), g->stream));
    HIPOK(g, hipStreamSynchronize(g->stream));
    return AEFDTD_OK;
}

int aefdtd_read_field(aefdtd_grid *g, int f, double *out) {
    if (!g || !out) return AEFDTD_EINVAL;
    if (f < 0 || f > 5) return fail(g, AEFDTD_EINVAL, "bad field id");
    size_t N = g->ncell();
    HIPOK(g, hipStreamSynchronize(g->stream));
    if (g->precision == AEFDTD_FP64) {
        HIPOK(g, cells_to_host(g, out, g->F[f], 8));
    } else if (g->precision == AEFDTD_FP32) {
        std::vector<float> t(N);
        HIPOK(g, cells_to_host(g, t.data(), g->F[f], 4));
        for (size_t i = 0; i < N; i++) out[i] = t[i];
    } else {
        std::vector<_Float16> t(N);
        HIPOK(g, cells_to_host(g, t.data(), g->F[f], 2));
        const double inv = std::ldexp(1.0, -g->scale_exp[f]);
        for (size_t i = 0; i < N; i++) out[i] = (double)(float)t[i] * inv;
    }
    return AEFDTD_OK;
}

int aefdtd_write_field(aefdtd_grid *g, int f, const double *in) {
    if (!g || !in) return AEFDTD_EINVAL;
    g->strip_valid = false;
    g->rc_ahead = false;
    if (f < 0 || f > 5) return fail(g, AEFDTD_EINVAL, "bad field id");
    HIPOK(g, hipStreamSynchronize(g->stream));
    if (g->precision == AEFDTD_FP16) {
        size_t N = g->ncell();
        std::vector<_Float16> t(N);
        const double sc = std::ldexp(1.0, g->scale_exp[f]);
        for (size_t i = 0; i < N; i++) t[i] = (_Float16)(float)(in[i] * sc);
        HIPOK(g, cells_to_device(g, g->F[f], t.data(), 2));
        return AEFDTD_OK;
    }
    if (is_fp64(g)) {
        HIPOK(g, cells_to_device(g, g->F[f], in, 8));
    } else {
        const size_t N = g->ncell();
        std::vector<float> t(N);
        for (size_t i = 0; i < N; i++) t[i] = (float)in[i];
        HIPOK(g, cells_to_device(g, g->F[f], t.data(), 4));
    }
    return AEFDTD_OK;
}

int aefdtd_get_storage_scale(aefdtd_grid *g, int exps[6]) {
    if (!g || !exps) return AEFDTD_EINVAL;
    for (int c = 0; c < 6; c++) exps[c] = g->scale_exp[c];
    return AEFDTD_OK;
}

int aefdtd_set_storage_scale(aefdtd_grid *g, const int exps[6]) {
    if (!g || !exps) return AEFDTD_EINVAL;
    if (g->precision != AEFDTD_FP16) return fail(g, AEFDTD_EINVAL, "storage scales apply to FP16 storage only");
    if (g->tstep != 0) return fail(g, AEFDTD_ESTATE, "storage scales must be set before the first update");
    for (int c = 0; c < 6; c++)
        if (exps[c] < -24 || exps[c] > 24) return fail(g, AEFDTD_EINVAL, "storage scale exponent out of range");
    for (int c = 0; c < 6; c++) g->scale_exp[c] = exps[c];
    return AEFDTD_OK;
}

int aefdtd_set_plane_injection(aefdtd_grid *g, int k, int n, const double *dEx, const double *dEy, const double *dHx,
                               const double *dHy) {
    if (!g) return AEFDTD_EINVAL;
    if (k < 0 || k >= g->Nz || n < 0 || (n > 0 && (!dEx || !dEy || !dHx || !dHy)))
        return fail(g, AEFDTD_EINVAL, "bad injection plane");
    g->inj_k = k;
    if (g->d_inj) {
        HIPOK(g, hipStreamSynchronize(g->stream));
        g->dev_bytes -= g->inj[0].size() * 4 * sizeof(double);
        dfree_t(g->d_inj);
    }
    g->inj[0].assign(dEx, dEx + n); g->inj[1].assign(dEy, dEy + n);
    g->inj[2].assign(dHx, dHx + n); g->inj[3].assign(dHy, dHy + n);
    return AEFDTD_OK;
}

int aefdtd_set_tfsf(aefdtd_grid *g, const aefdtd_grid *aux_c, int i1, int i2, int j1, int j2, int k1, int k2, int px,
                    int py) {
    if (!g || !aux_c) return AEFDTD_EINVAL;
    aefdtd_grid *aux = const_cast<aefdtd_grid *>(aux_c);
    if ((!px && (i1 < 1 || i2 >= g->Nx)) || (!py && (j1 < 1 || j2 >= g->Ny)) || i1 < 0 || j1 < 0 || i2 > g->Nx ||
        j2 > g->Ny || i1 >= i2 || j1 >= j2 || k1 < 1 || k2 >= g->Nz || k1 >= k2)
        return fail(g, AEFDTD_EINVAL, "bad TFSF box");
    if (aux->Nx != 1 || aux->Ny != 1 || aux->Nz != g->Nz || aux->precision != g->precision)
        return fail(g, AEFDTD_EINVAL, "TFSF auxiliary grid must be 1x1xNz with the same precision");
    // the auxiliary grid is stepped on the main grid's stream so that both
    // follow one order of launches
    if (aux->stream != g->stream) {
        HIPOK(g, hipStreamSynchronize(aux->stream));
        if (aux->own_stream) (void)hipStreamDestroy(aux->stream);
        aux->stream = g->stream;
        aux->own_stream = false;
    }
    g->tfsf_aux = aux;
    int b[6] = {i1, i2, j1, j2, k1, k2};
    std::memcpy(g->tb, b, sizeof(b));
    g->tf_per_x = px != 0;
    g->tf_per_y = py != 0;
    return AEFDTD_OK;
}

int aefdtd_add_oscillator(aefdtd_grid *g, int type, int x1, int x2, int y1, int y2, int z1, int z2,
                          double lambda_min, double lambda_max, int *id) {
    if (!g || !id) return AEFDTD_EINVAL;
    *id = -1;
    if (type < 0 || type > 6) return fail(g, AEFDTD_EINVAL, "bad oscillator orientation");
    if (!(lambda_min > 0) || !(lambda_max > 0) || lambda_min == lambda_max)
        return fail(g, AEFDTD_EINVAL, "bad oscillator spectrum");
    if ((int)g->osc.size() >= MAX_OSC) return fail(g, AEFDTD_EINVAL, "too many oscillators");
    if (type == 6) return AEFDTD_OK;    // NORMAL_UNDEF injects nothing (sc_oscillator.cpp:63-65)
    OscDev o;
    o.comp = type / 2;                  // NORMAL_XM -> NORMAL_X etc. (sc_oscillator.cpp:72-74)
    o.x1 = std::clamp(x1 + g->xs_s, 0, g->Nx); o.x2 = std::clamp(x2 + g->xs_s, 0, g->Nx);   // Source::link
    o.y1 = std::clamp(y1 + g->ys_s, 0, g->Ny); o.y2 = std::clamp(y2 + g->ys_s, 0, g->Ny);
    o.z1 = std::clamp(z1 + g->zs_s, 0, g->Nz); o.z2 = std::clamp(z2 + g->zs_s, 0, g->Nz);
    o.step0 = g->tstep;
    o.nval = 0;
    o.val = nullptr;
    OscParams par;
    oscillator_init(lambda_min, lambda_max, g->Dt, par);
    if (!g->d_osc) {
        int rc = dalloc(g, &g->d_osc, MAX_OSC * sizeof(OscDev));
        if (rc) return rc;
    }
    g->osc.push_back(o);
    g->osc_par.push_back(par);
    int rc = ensure_osc_tables(g, g->tstep + std::max(g->Nt, 1));
    if (rc) return rc;
    *id = (int)g->osc.size() - 1;
    return AEFDTD_OK;
}

int aefdtd_add_spect_poynting(aefdtd_grid *g, int type, int x1, int x2, int y1, int y2, int z1, int z2, int Nl,
                              const double *lambda, int *id) {
    if (!g || !lambda || !id) return AEFDTD_EINVAL;
    if (type < 0 || type > 5) return fail(g, AEFDTD_EINVAL, "bad orientation");
    if (g->sensor_step > 0 || g->d_stage) return fail(g, AEFDTD_ESTATE, "sensors must be added before the first feed");
    auto *s = new Sensor();
    s->type = type;
    s->x1 = std::clamp(x1 + g->xs_s, 0, g->Nx); s->x2 = std::clamp(x2 + g->xs_s, 0, g->Nx);
    s->y1 = std::clamp(y1 + g->ys_s, 0, g->Ny); s->y2 = std::clamp(y2 + g->ys_s, 0, g->Ny);
    s->z1 = std::clamp(z1 + g->zs_s, 0, g->Nz); s->z2 = std::clamp(z2 + g->zs_s, 0, g->Nz);
    if (type == AEFDTD_NORMAL_X || type == AEFDTD_NORMAL_XM) { s->span1 = s->y2 - s->y1; s->span2 = s->z2 - s->z1; }
    else if (type == AEFDTD_NORMAL_Y || type == AEFDTD_NORMAL_YM) { s->span1 = s->x2 - s->x1; s->span2 = s->z2 - s->z1; }
    else { s->span1 = s->x2 - s->x1; s->span2 = s->y2 - s->y1; }
    bool plane_ok = (type <= 1) ? s->x1 < g->Nx : (type <= 3 ? s->y1 < g->Ny : s->z1 < g->Nz);
    if (s->span1 <= 0 || s->span2 <= 0 || Nl < 1 || !plane_ok) {
        delete s;
        return fail(g, AEFDTD_EINVAL, "empty sensor");
    }
    s->Nl = Nl;
    s->lambda.assign(lambda, lambda + Nl);
    s->P = (long)s->span1 * s->span2;
    void *p;
    int rc = dalloc(g, &p, (size_t)6 * Nl * s->P * sizeof(double2));
    if (rc) { delete s; return rc; }
    s->d_acc = (double2 *)p;
    HIPOK(g, hipMemsetAsync(s->d_acc, 0, (size_t)6 * Nl * s->P * sizeof(double2), g->stream));
    g->sensors.push_back(s);
    *id = (int)g->sensors.size() - 1;
    return AEFDTD_OK;
}

namespace {

// Sensor::link's offset and clamp (sensors.cpp:196-230) of a region given in
// structure cells; false when it is empty.
bool link_region(const aefdtd_grid *g, int &x1, int &x2, int &y1, int &y2, int &z1, int &z2) {
    x1 = std::clamp(x1 + g->xs_s, 0, g->Nx); x2 = std::clamp(x2 + g->xs_s, 0, g->Nx);
    y1 = std::clamp(y1 + g->ys_s, 0, g->Ny); y2 = std::clamp(y2 + g->ys_s, 0, g->Ny);
    z1 = std::clamp(z1 + g->zs_s, 0, g->Nz); z2 = std::clamp(z2 + g->zs_s, 0, g->Nz);
    return x2 > x1 && y2 > y1 && z2 > z1;
}

int add_field_sensor(aefdtd_grid *g, FieldSensor *s, int *id) {
    void *p;
    const size_t bytes = (size_t)s->ncomp * s->d.n * sizeof(double2);
    int rc = dalloc(g, &p, bytes);
    if (rc) { delete s; return rc; }
    s->d.acc = (double2 *)p;
    HIPOK(g, hipMemsetAsync(s->d.acc, 0, bytes, g->stream));
    g->fsensors.push_back(s);
    *id = (int)g->fsensors.size() - 1;
    return AEFDTD_OK;
}

}  // namespace

int aefdtd_add_field_block(aefdtd_grid *g, int x1, int x2, int y1, int y2, int z1, int z2, double lambda, int *id) {
    if (!g || !id) return AEFDTD_EINVAL;
    if (!(lambda > 0)) return fail(g, AEFDTD_EINVAL, "bad wavelength");
    if (!link_region(g, x1, x2, y1, y2, z1, z2)) return fail(g, AEFDTD_EINVAL, "empty sensor");
    auto *s = new FieldSensor();
    s->lambda = lambda;
    s->ncomp = 6;
    s->d = dev::FieldDesc{1, 0, 0, 0, x1, y1, z1, x2 - x1, y2 - y1, z2 - z1,
                          (long)(x2 - x1) * (y2 - y1) * (z2 - z1), nullptr};
    s->step = g->sensor_step;
    return add_field_sensor(g, s, id);
}

int aefdtd_add_field_map(aefdtd_grid *g, int orientation, int x1, int x2, int y1, int y2, int z1, int z2,
                         double lambda, int cumulative, int magnetic, int *id) {
    if (!g || !id) return AEFDTD_EINVAL;
    if (orientation < 0 || orientation > 5) return fail(g, AEFDTD_EINVAL, "bad orientation");
    if (!(lambda > 0)) return fail(g, AEFDTD_EINVAL, "bad wavelength");
    if (!link_region(g, x1, x2, y1, y2, z1, z2)) return fail(g, AEFDTD_EINVAL, "empty sensor");
    // FieldMap::initialize folds the - orientations (s_fieldmap.cpp:433-440)
    const int normal = orientation / 2;
    auto *s = new FieldSensor();
    s->lambda = lambda;
    s->ncomp = 3;
    s->type = 2 * normal;
    const int ex[3] = {x2 - x1, y2 - y1, z2 - z1};
    const int s1 = normal == 0 ? ex[1] : ex[0], s2 = normal == 2 ? ex[1] : ex[2];
    s->d = dev::FieldDesc{0, normal, cumulative ? 1 : 0, magnetic ? 1 : 0, x1, y1, z1, s1, s2, ex[normal],
                          (long)s1 * s2, nullptr};
    s->step = g->sensor_step;
    return add_field_sensor(g, s, id);
}

int aefdtd_field_sensor_read(aefdtd_grid *g, int id, double *out, int dims[4]) {
    if (!g || !dims) return AEFDTD_EINVAL;
    if (id < 0 || id >= (int)g->fsensors.size()) return fail(g, AEFDTD_EINVAL, "bad field sensor id");
    const FieldSensor &s = *g->fsensors[id];
    dims[0] = s.ncomp;
    dims[1] = s.d.s1;
    dims[2] = s.d.s2;
    dims[3] = s.d.block ? s.d.s3 : 1;
    if (!out) return AEFDTD_OK;
    HIPOK(g, copy_sync(g, out, s.d.acc, (size_t)s.ncomp * s.d.n * sizeof(double2), hipMemcpyDeviceToHost));
    return AEFDTD_OK;
}

int aefdtd_field_sensor_mats(aefdtd_grid *g, int id, uint32_t *out) {
    if (!g || !out) return AEFDTD_EINVAL;
    if (id < 0 || id >= (int)g->fsensors.size()) return fail(g, AEFDTD_EINVAL, "bad field sensor id");
    if (g->matsgrid.empty()) return fail(g, AEFDTD_ESTATE, "no material grid");
    const dev::FieldDesc &d = g->fsensors[id]->d;
    auto mg = [&](int i, int j, int k) { return g->matsgrid[(size_t)i + (size_t)g->Nx * ((size_t)j + (size_t)g->Ny * k)]; };
    for (long q = 0; q < d.n; q++) {
        const int a = (int)(q % d.s1), b = (int)((q / d.s1) % d.s2), c = d.block ? (int)(q / ((long)d.s1 * d.s2)) : 0;
        if (d.block) out[q] = mg(d.x1 + a, d.y1 + b, d.z1 + c);
        else if (d.normal == 0) out[q] = mg(d.x1, d.y1 + a, d.z1 + b);
        else if (d.normal == 1) out[q] = mg(d.x1 + a, d.y1, d.z1 + b);
        else out[q] = mg(d.x1 + a, d.y1 + b, d.z1);
    }
    return AEFDTD_OK;
}

int aefdtd_feed_sensors(aefdtd_grid *g) {
    if (!g) return AEFDTD_EINVAL;
    int rc = ensure_ready(g);
    if (rc) return rc;
    return with_types(g, [&]<typename S, typename T>() { return feed_T<S, T>(g); });
}

}  // extern "C"

int aefdtd::energy_feed(aefdtd_grid *g) {
    if (!g) return AEFDTD_EINVAL;
    if (!g->d_energy) {
        void *p;
        int rc = dalloc(g, &p, (EN_BLOCKS + 2) * sizeof(double));
        if (rc) return rc;
        g->d_energy = (double *)p;
        HIPOK(g, fill_sync(g, g->d_energy, 0, (EN_BLOCKS + 2) * sizeof(double)));
    }
    if (g->en_parts > 0) {
        // the E pass of this step left one partial per tile workgroup: summed
        // in 256 chunks, then those in one workgroup
        hipLaunchKernelGGL(k_energy_sum, dim3(256), dim3(256), 0, g->stream, (const double *)g->d_en_part,
                           (long)g->en_parts, g->d_energy);
        hipLaunchKernelGGL(k_energy_fin, dim3(1), dim3(256), 0, g->stream, (const double *)g->d_energy, 256,
                           g->d_energy + EN_BLOCKS);
        g->en_parts = 0;
    } else {
        with_types(g, [&]<typename S, typename T>() {
            hipLaunchKernelGGL((k_energy_part<S, T>), dim3(EN_BLOCKS), dim3(256), 0, g->stream, fld<S, T>(g, 0),
                               fld<S, T>(g, 1), fld<S, T>(g, 2), g->Nx, g->Ny * g->Nz, g->px, g->d_energy);
            return AEFDTD_OK;
        });
        hipLaunchKernelGGL(k_energy_fin, dim3(1), dim3(256), 0, g->stream, (const double *)g->d_energy, EN_BLOCKS,
                           g->d_energy + EN_BLOCKS);
    }
    HIPOK(g, hipGetLastError());
    return AEFDTD_OK;
}

int aefdtd::energy_read(aefdtd_grid *g, double *last, double *max) {
    if (!g || !g->d_energy) return AEFDTD_EINVAL;
    double em[2];
    HIPOK(g, copy_sync(g, em, g->d_energy + EN_BLOCKS, sizeof(em), hipMemcpyDeviceToHost));
    *last = em[0];
    *max = em[1];
    return AEFDTD_OK;
}

int aefdtd::ft_setup(aefdtd_grid *g, FTState &f, const std::vector<int> &pts, const std::vector<double> &w,
                     int navg) {
    ft_free(f);
    const int np = (int)w.size();
    if (np < 1 || navg < 1 || pts.size() != 3 * w.size()) return fail(g, AEFDTD_EINVAL, "bad completion points");
    for (int p = 0; p < np; p++)
        if (pts[3 * p] < 0 || pts[3 * p] >= g->Nx || pts[3 * p + 1] < 0 || pts[3 * p + 1] >= g->Ny ||
            pts[3 * p + 2] < 0 || pts[3 * p + 2] >= g->Nz)
            return fail(g, AEFDTD_EINVAL, "completion point outside the grid");
    void *p0, *p1, *p2, *p3;
    int rc;
    if ((rc = dalloc(g, &p0, pts.size() * sizeof(int))) || (rc = dalloc(g, &p1, w.size() * sizeof(double))) ||
        (rc = dalloc(g, &p2, (size_t)6 * np * sizeof(double))) ||
        (rc = dalloc(g, &p3, (size_t)6 * np * navg * sizeof(double))))
        return rc;
    f.d_pts = (int *)p0; f.d_w = (double *)p1; f.d_acc = (double *)p2; f.d_ring = (double *)p3;
    f.np = np;
    f.navg = navg;
    HIPOK(g, copy_sync(g, f.d_pts, pts.data(), pts.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPOK(g, copy_sync(g, f.d_w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPOK(g, fill_sync(g, f.d_acc, 0, (size_t)6 * np * sizeof(double)));
    HIPOK(g, fill_sync(g, f.d_ring, 0, (size_t)6 * np * navg * sizeof(double)));
    return AEFDTD_OK;
}

int aefdtd::ft_feed(aefdtd_grid *g, FTState &f, int step, double Dt) {
    if (!f.np) return AEFDTD_OK;
    with_types(g, [&]<typename S, typename T>() {
        hipLaunchKernelGGL((k_ft_feed<S, T>), dim3((f.np + 255) / 256), dim3(256), 0, g->stream, (const int *)f.d_pts,
                           (const double *)f.d_w, f.np, step, Dt, f.navg, fld<S, T>(g, 0), fld<S, T>(g, 1),
                           fld<S, T>(g, 2), g->Nx, g->Ny, g->Nz, g->px, f.d_acc, f.d_ring);
        return AEFDTD_OK;
    });
    HIPOK(g, hipGetLastError());
    return AEFDTD_OK;
}

int aefdtd::ft_read(aefdtd_grid *g, const FTState &f, std::vector<double> &ring) {
    ring.resize((size_t)6 * f.np * f.navg);
    HIPOK(g, copy_sync(g, ring.data(), f.d_ring, ring.size() * sizeof(double), hipMemcpyDeviceToHost));
    return AEFDTD_OK;
}

void aefdtd::ft_free(FTState &f) {
    dfree_t(f.d_pts);
    dfree_t(f.d_w);
    dfree_t(f.d_acc);
    dfree_t(f.d_ring);
    f.np = f.navg = 0;
}

int aefdtd::sensor_spans(const aefdtd_grid *g, int id, int *span1, int *span2) {
    if (!g || id < 0 || id >= (int)g->sensors.size()) return AEFDTD_EINVAL;
    *span1 = g->sensors[id]->span1;
    *span2 = g->sensors[id]->span2;
    return AEFDTD_OK;
}

int aefdtd::sensor_box(const aefdtd_grid *g, int id, int box[6]) {
    if (!g || id < 0 || id >= (int)g->sensors.size()) return AEFDTD_EINVAL;
    const Sensor &s = *g->sensors[id];
    const int b[6] = {s.x1, s.x2, s.y1, s.y2, s.z1, s.z2};
    for (int q = 0; q < 6; q++) box[q] = b[q];
    return AEFDTD_OK;
}

int aefdtd::field_sensor_geom(const aefdtd_grid *g, int id, int out[8], double *lambda) {
    if (!g || id < 0 || id >= (int)g->fsensors.size()) return AEFDTD_EINVAL;
    const dev::FieldDesc &d = g->fsensors[id]->d;
    const int v[8] = {d.block, d.normal, d.x1, d.y1, d.z1, d.s1, d.s2, d.s3};
    for (int q = 0; q < 8; q++) out[q] = v[q];
    if (lambda) *lambda = g->fsensors[id]->lambda;
    return AEFDTD_OK;
}

extern "C" {

int aefdtd_sensor_poynting(aefdtd_grid *g, int id, double *rsp) {
    if (!g || !rsp) return AEFDTD_EINVAL;
    if (id < 0 || id >= (int)g->sensors.size()) return fail(g, AEFDTD_EINVAL, "bad sensor id");
    int rc = with_ctype(g, [&]<typename T>() { return flush_dft<T>(g); });
    if (rc) return rc;
    Sensor &s = *g->sensors[id];
    double f;
    if (s.type == AEFDTD_NORMAL_X || s.type == AEFDTD_NORMAL_XM) f = g->Dy * g->Dz;
    else if (s.type == AEFDTD_NORMAL_Y || s.type == AEFDTD_NORMAL_YM) f = g->Dx * g->Dz;
    else f = g->Dx * g->Dy;
    double *d;
    if ((rc = dalloc(g, (void **)&d, s.Nl * sizeof(double)))) return rc;
    hipLaunchKernelGGL(k_poynting, dim3(s.Nl), dim3(256), 0, g->stream, (const double2 *)s.d_acc, s.P, s.Nl, f, s.type, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e == hipSuccess) e = copy_sync(g, rsp, d, s.Nl * sizeof(double), hipMemcpyDeviceToHost);
    void *v = d;
    g->dev_bytes -= s.Nl * sizeof(double);
    dfree(v);
    if (e != hipSuccess) return fail(g, AEFDTD_EDEVICE, std::string("poynting: ") + hipGetErrorString(e));
    return AEFDTD_OK;
}

}  // extern "C"

// ---- the sensor power without allocations
// k_poynting writes one value per wavelength to a staging buffer before the
// copy to the caller's host array.  The allocating entry point takes that
// buffer from hipMalloc on every call; the one below uses `scratch`, a
// file-scope device array of AEFDTD_NOALLOC_MAX_NL doubles.  Each GPU holds
// its own copy at its own address, found once per device with
// hipGetSymbolAddress; host code uses only that address.  Every call on a GPU
// shares that GPU's array (k_poynting overwrites the Nl values it reads back,
// so nothing needs clearing): calls that overlap on one GPU race.
namespace {

static __device__ __align__(256) double scratch[AEFDTD_NOALLOC_MAX_NL];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (dev), or null.
double *scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double *p = nullptr;
        if (hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int aefdtd_sensor_poynting_noalloc(aefdtd_grid *g, int id, double *rsp) {
    if (!g || !rsp) return AEFDTD_EINVAL;
    if (id < 0 || id >= (int)g->sensors.size()) return fail(g, AEFDTD_EINVAL, "bad sensor id");
    Sensor &s = *g->sensors[id];
    if (s.Nl > AEFDTD_NOALLOC_MAX_NL)
        return fail(g, AEFDTD_EINVAL, "sensor has " + std::to_string(s.Nl) + " wavelengths; the allocation-free "
                                      "entry point takes at most " + std::to_string(AEFDTD_NOALLOC_MAX_NL));
    // the pending DFT flush must not need a larger coefficient table (that would allocate)
    if (g->stage_fill > 0)
        for (const Sensor *q : g->sensors)
            if (g->sensor_step > q->ncoef) return fail(g, AEFDTD_ESTATE, "the DFT flush would allocate coefficients");
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return fail(g, AEFDTD_EDEVICE, "hipGetDevice failed");
    if (cur != g->device && hipSetDevice(g->device) != hipSuccess) return fail(g, AEFDTD_EDEVICE, "hipSetDevice failed");
    double *d = scratch_on(g->device);
    int rc = d ? with_ctype(g, [&]<typename T>() { return flush_dft<T>(g); }) : AEFDTD_EDEVICE;
    hipError_t e = hipSuccess;
    if (rc == AEFDTD_OK) {
        double f;
        if (s.type == AEFDTD_NORMAL_X || s.type == AEFDTD_NORMAL_XM) f = g->Dy * g->Dz;
        else if (s.type == AEFDTD_NORMAL_Y || s.type == AEFDTD_NORMAL_YM) f = g->Dx * g->Dz;
        else f = g->Dx * g->Dy;
        hipLaunchKernelGGL(k_poynting, dim3(s.Nl), dim3(256), 0, g->stream, (const double2 *)s.d_acc, s.P, s.Nl, f,
                           s.type, d);
        e = hipGetLastError();
        if (e == hipSuccess) e = copy_sync(g, rsp, d, s.Nl * sizeof(double), hipMemcpyDeviceToHost);
    }
    if (cur != g->device) (void)hipSetDevice(cur);
    if (!d) return fail(g, AEFDTD_EDEVICE, "hipGetSymbolAddress of the scratch array failed");
    if (rc) return rc;
    if (e != hipSuccess) return fail(g, AEFDTD_EDEVICE, std::string("poynting: ") + hipGetErrorString(e));
    return AEFDTD_OK;
}

int aefdtd_sensor_spectra(aefdtd_grid *g, int id, double *out, int dims[3]) {
    if (!g || !dims) return AEFDTD_EINVAL;
    if (id < 0 || id >= (int)g->sensors.size()) return fail(g, AEFDTD_EINVAL, "bad sensor id");
    int rc = with_ctype(g, [&]<typename T>() { return flush_dft<T>(g); });
    if (rc) return rc;
    Sensor &s = *g->sensors[id];
    dims[0] = s.span1; dims[1] = s.span2; dims[2] = s.Nl;
    if (!out) return AEFDTD_OK;
    HIPOK(g, hipStreamSynchronize(g->stream));
    HIPOK(g, copy_sync(g, out, s.d_acc, (size_t)6 * s.Nl * s.P * sizeof(double2), hipMemcpyDeviceToHost));
    return AEFDTD_OK;
}

int aefdtd_add_plane_sum(aefdtd_grid *g, int k, const aefdtd_pulse *inc, double zpos, int subtract, const double *tap,
                         int *id) {
    if (!g || !tap || !id) return AEFDTD_EINVAL;
    if (k < 0 || k >= g->Nz) return fail(g, AEFDTD_EINVAL, "bad plane");
    if (g->Nt < 1) return fail(g, AEFDTD_EINVAL, "plane sums need Nt >= 1");
    auto *s = new PlaneSum();
    s->k = k;
    s->has_inc = inc != nullptr;
    s->subtract = subtract != 0 && inc != nullptr;
    if (inc) s->inc = *inc;
    else std::memset(&s->inc, 0, sizeof(s->inc));
    s->zpos = zpos;
    s->cap = g->Nt;
    void *p;
    int rc = dalloc(g, &p, g->Nt * sizeof(double));
    if (rc) { delete s; return rc; }
    s->d_tap = (double *)p;
    if ((rc = dalloc(g, &p, (size_t)g->Nt * PS_BLOCKS * 6 * sizeof(double)))) { dfree_t(s->d_tap); delete s; return rc; }
    s->d_out = (double *)p;
    HIPOK(g, copy_sync(g, s->d_tap, tap, g->Nt * sizeof(double), hipMemcpyHostToDevice));
    HIPOK(g, fill_sync(g, s->d_out, 0, (size_t)g->Nt * PS_BLOCKS * 6 * sizeof(double)));
    g->sums.push_back(s);
    *id = (int)g->sums.size() - 1;
    return AEFDTD_OK;
}

int aefdtd_read_plane_sum(aefdtd_grid *g, int id, int n, double *out) {
    if (!g || !out) return AEFDTD_EINVAL;
    if (id < 0 || id >= (int)g->sums.size()) return fail(g, AEFDTD_EINVAL, "bad sum id");
    PlaneSum &s = *g->sums[id];
    int m = std::min(n, s.cap);
    HIPOK(g, hipStreamSynchronize(g->stream));
    std::vector<double> part((size_t)m * PS_BLOCKS * 6);
    HIPOK(g, copy_sync(g, part.data(), s.d_out, part.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int t = 0; t < m; t++)
        for (int c = 0; c < 6; c++) {
            double v = 0;
            for (int b = 0; b < PS_BLOCKS; b++) v += part[((size_t)t * PS_BLOCKS + b) * 6 + c];
            out[(size_t)t * 6 + c] = v;
        }
    return AEFDTD_OK;
}

int aefdtd_autotune(aefdtd_grid *g, int reps, int out[2]) {
    if (!g) return AEFDTD_EINVAL;
    int rc = ensure_ready(g);
    if (rc) return rc;
    if (g->pml_dirty || g->tstep == 0) {
        if ((rc = upload_profiles(g))) return rc;
        g->pml_dirty = false;
    }
    if (is_column(g)) {
        if (out) { out[0] = g->tile_e; out[1] = g->tile_h; }
        return AEFDTD_OK;
    }
    return with_types(g, [&]<typename S, typename T>() { return autotune_T<S, T>(g, std::max(1, reps), out); });
}

int aefdtd_set_tiles(aefdtd_grid *g, int tile_e, int tile_h) {
    if (!g) return AEFDTD_EINVAL;
    if (tile_e < 0 || tile_e > 31 || tile_h < 0 || tile_h > 31) return fail(g, AEFDTD_EINVAL, "tile configuration out of range");
    g->tile_e = tile_e;
    g->tile_h = tile_h;
    return AEFDTD_OK;
}

int aefdtd_profile_enable(aefdtd_grid *g, int on) {
    if (!g) return AEFDTD_EINVAL;
    prof_collect(g);
    g->prof_on = on != 0;
    if (on) {
        for (int k = 0; k < 5; k++) { g->prof_ms[k] = 0; g->prof_n[k] = 0; }
    }
    return AEFDTD_OK;
}

int aefdtd_profile_kinds(aefdtd_grid *g, int mask) {
    if (!g || mask < 0 || mask > 31) return AEFDTD_EINVAL;
    g->prof_mask = mask;
    return AEFDTD_OK;
}

int aefdtd_profile_read(aefdtd_grid *g, const char *kernel, double *total_ms, long *launches) {
    if (!g || !kernel) return AEFDTD_EINVAL;
    prof_collect(g);
    int k = -1;
    if (!std::strcmp(kernel, "E")) k = 0;
    else if (!std::strcmp(kernel, "H")) k = 1;
    else if (!std::strcmp(kernel, "sensor")) k = 2;
    else if (!std::strcmp(kernel, "dft")) k = 3;
    else if (!std::strcmp(kernel, "rc")) k = 4;
    if (k < 0) return fail(g, AEFDTD_EINVAL, "unknown kernel name");
    if (total_ms) *total_ms = g->prof_ms[k];
    if (launches) *launches = g->prof_n[k];
    return AEFDTD_OK;
}

int aefdtd_stream_bandwidth(size_t bytes, int reps, double out[3]) {
    if (!out || bytes < (1u << 20) || reps < 1) return AEFDTD_EINVAL;
    const size_t n = bytes / sizeof(float4);
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = AEFDTD_OK;
    if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess || hipMalloc(&b, n * sizeof(float4)) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || timing_event(&e0) != hipSuccess ||
        timing_event(&e1) != hipSuccess) {
        rc = AEFDTD_EDEVICE;
    } else {
        (void)hipMemsetAsync(a, 0, n * sizeof(float4), st);
        (void)hipMemsetAsync(b, 0, n * sizeof(float4), st);
        double best = 0, best_nt = 0, best_ms = 0;
        const unsigned grids[4] = {1024, 2048, 4096, 8192};
        for (int nt = 0; nt < 2; nt++)
            for (unsigned gdim : grids) {
                for (int r = 0; r < reps + 1; r++) {
                    (void)hipEventRecord(e0, st);
                    if (nt) hipLaunchKernelGGL(k_stream_copy<true>, dim3(gdim), dim3(256), 0, st, (const f32x4 *)a, (f32x4 *)b, n);
                    else hipLaunchKernelGGL(k_stream_copy<false>, dim3(gdim), dim3(256), 0, st, (const f32x4 *)a, (f32x4 *)b, n);
                    (void)hipEventRecord(e1, st);
                    (void)hipEventSynchronize(e1);
                    float ms = 0;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (r == 0 || ms <= 0) continue;            // first launch warms up
                    const double gbs = 2.0 * n * sizeof(float4) / (ms * 1e-3) / 1e9;
                    if (nt) best_nt = std::max(best_nt, gbs);
                    else if (gbs > best) { best = gbs; best_ms = ms; }
                }
            }
        if (hipGetLastError() != hipSuccess) rc = AEFDTD_EDEVICE;
        out[0] = best;
        out[1] = best_nt;
        out[2] = best_ms;
    }
    if (st) (void)hipStreamSynchronize(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

int aefdtd_stream_mix_bandwidth(size_t bytes, int reps, double out[2]) {
    if (!out || bytes < (1u << 20) || reps < 1) return AEFDTD_EINVAL;
    const size_t n = bytes / sizeof(float);
    float *buf[9] = {};
    float **ptrs = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = AEFDTD_OK;
    bool ok = hipMalloc(&ptrs, 9 * sizeof(float *)) == hipSuccess;
    for (int i = 0; i < 9 && ok; i++) ok = hipMalloc(&buf[i], n * sizeof(float)) == hipSuccess;
    if (!ok || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || timing_event(&e0) != hipSuccess ||
        timing_event(&e1) != hipSuccess) {
        rc = AEFDTD_EDEVICE;
    } else {
        for (int i = 0; i < 9; i++) (void)hipMemsetAsync(buf[i], 0, n * sizeof(float), st);
        (void)hipMemcpyAsync(ptrs, buf, 9 * sizeof(float *), hipMemcpyHostToDevice, st);
        double best = 0, best_ms = 0;
        const unsigned grids[3] = {2048, 8192, 65536};
        for (unsigned gdim : grids)
            for (int r = 0; r < reps + 1; r++) {
                (void)hipEventRecord(e0, st);
                hipLaunchKernelGGL(k_stream_mix, dim3(gdim), dim3(256), 0, st, (const float *const *)ptrs,
                                   (float *const *)(ptrs + 6), n);
                (void)hipEventRecord(e1, st);
                (void)hipEventSynchronize(e1);
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (r == 0 || ms <= 0) continue;            // first launch warms up
                const double gbs = 9.0 * n * sizeof(float) / (ms * 1e-3) / 1e9;
                if (gbs > best) { best = gbs; best_ms = ms; }
            }
        if (hipGetLastError() != hipSuccess) rc = AEFDTD_EDEVICE;
        out[0] = best;
        out[1] = best_ms;
    }
    if (st) (void)hipStreamSynchronize(st);
    for (float *p : buf)
        if (p) (void)hipFree(p);
    if (ptrs) (void)hipFree(ptrs);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

int aefdtd_stats(const aefdtd_grid *g, long out[4]) {
    if (!g || !out) return AEFDTD_EINVAL;
    out[0] = (long)g->ncell();
    out[1] = g->n_disp;
    out[2] = g->np_max;
    out[3] = (long)g->ssize();
    return AEFDTD_OK;
}

int aefdtd_device_bytes(const aefdtd_grid *g, size_t *bytes) {
    if (!g || !bytes) return AEFDTD_EINVAL;
    *bytes = g->dev_bytes;
    return AEFDTD_OK;
}

}  // extern "C"
