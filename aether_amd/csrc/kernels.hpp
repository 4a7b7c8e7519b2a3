// kernels.hpp -- gfx950 device code of the FDTD engine (included by engine.hip).
//
// One fused pass per half step keeps the leapfrog at its HBM floor:
//   E pass : read E(3) H(3) mat -> RC-ante (FP64 poles) -> E stencil -> TFSF-E
//            -> RC-apply -> CPML-E (slab psi) -> PEC -> plane injection -> write E(3)
//   H pass : read H(3) E(3)     -> H stencil -> CPML-H -> PEC -> plane injection
//            -> TFSF-H -> write H(3)
// This is the phase order of Aether's threaded_process_E/H (core/fdtd_threads.
// cpp:32-235) and of its drivers (modes/normal_incidence.cpp:330-357,
// modes/single_particle.cpp:256-378); every phase is cell-local, so fusing them
// keeps each cell's sequence of floating-point updates.
//
// Thread mapping of the 3D passes: a workgroup is 64 x BY lanes over (x, y);
// x is the contiguous axis, so one wavefront reads 64 consecutive values of a
// field.  Each lane marches KZ planes in z and carries the previous plane's H
// (E pass) / next plane's E (H pass) in registers.  Workgroups are launched
// as a 1D grid; with `remap` set, the hardware workgroup id is folded so that
// each of the 8 XCDs takes chunks of a few consecutive tile rows (x fastest,
// then y, then z), which keeps most halo rows of neighbouring tiles in one L2.
// The per-material coefficients are staged once per workgroup in LDS.
#pragma once

#include "engine.hpp"

#include <type_traits>

namespace aefdtd {
namespace dev {

constexpr int BX = 64;
constexpr int KS = 64;             // steps per batch chunk (and its DFT staging depth)
constexpr int DFT_LCH = 32;        // wavelengths per LDS chunk in the DFT pass
constexpr int MAX_MATS = 256;

// A field array in storage type S read and written in compute type T.  With
// S == T this is a plain array; with FP16 storage each component carries a
// power-of-two scale (stored = value * 2^e) so that E (~1) and H (~E/377) both
// sit well inside the FP16 range.
template <typename S, typename T>
struct Fld {
    S *p;
    T ld, st;        // 2^-e and 2^e
    __device__ __forceinline__ T operator[](int i) const {
        if constexpr (std::is_same_v<S, T>) return p[i];
        else return (T)p[i] * ld;
    }
    // streaming load (not kept in the caches): for values a pass reads once
    __device__ __forceinline__ T nt(int i) const {
        if constexpr (std::is_same_v<S, T>) return __builtin_nontemporal_load(p + i);
        else return (T)__builtin_nontemporal_load(p + i) * ld;
    }
    __device__ __forceinline__ void set(int i, T v) const {
        if constexpr (std::is_same_v<S, T>) p[i] = v;
        else p[i] = (S)(v * st);
    }
    // streaming store (not kept in the caches): for passes whose output is
    // next read a whole pass later
    __device__ __forceinline__ void set_nt(int i, T v) const {
        if constexpr (std::is_same_v<S, T>) __builtin_nontemporal_store(v, p + i);
        else __builtin_nontemporal_store((S)(v * st), p + i);
    }
};

// Store-mask bit asking for streaming stores.
constexpr int STORE_NT = 64;

// Per-material constants (FDTD_Material, core/fdtd_material.h:46-47).
template <typename T>
struct MatDev {
    T C1, C2x, C2y, C2z, C4;
    int rc;          // recursive-convolution poles present
    double C3;
    int np, pole0;
};

// 1D profile arrays per axis a (length N_a): ikE, ikH (inverse kappa with the
// stencil's gating), bE, cE (c * 1/D), bH, cH.
template <typename T>
struct Prof {
    const T *ikE[3], *ikH[3], *bE[3], *cE[3], *bH[3], *cH[3];
};

// Cell arrays (fields, material ids, dispersive slots) are stored x fastest
// with a row pitch P >= Nx: cell (i, j, k) at i + P (j + Ny k).  3D grids pad
// P to a multiple of 64 cells so that every row starts on a 256-byte boundary
// and a wavefront's 64-cell row segment is a whole number of cache lines;
// the pad cells are never read.  The CPML slab arrays are dense in Nx.
template <typename S, typename T>
struct EArgs {
    int Nx, Ny, Nz, nmat, P;
    int nbx, nby, nbz, remap;
    Fld<S, T> Ex, Ey, Ez, Hx, Hy, Hz;
    const uint8_t *mat;
    const MatDev<T> *mt;
    const double2 *poles;       // [pole]: Crec at 2p, Dchi at 2p+1
    const int32_t *slot;
    const T *corr;              // per dispersive slot: (T)(C3 sum_p Re psi) for x, y, z
    Prof<T> pf;
    int pml[6];
    T *pEyx, *pEzx, *pExy, *pEzy, *pExz, *pEyz;   // slab psi, x:(pxm+pxp,Ny,Nz) y:(Nx,pym+pyp,Nz) z:(Nx,Ny,pzm+pzp)
    int en;                    // bit c set: component c enabled
    int inj_k;
    T injx, injy;
    int tfsf, tb0, tb1, tb2, tb3, tb4, tb5, tper_x, tper_y;
    Fld<S, T> aHx, aHy, aHz;   // aux 1D grid H columns
    int lm;                    // bits 0-2: E components loaded and stored, bits 3-5: H components loaded
    double *en_part;           // non-null: per-workgroup sums of Ex^2+Ey^2+Ez^2 after the update
    int nosc, ostep;           // oscillators and the step their value is taken at
    const OscDev *osc;
};

constexpr int MAX_STRIP = 4;

template <typename S, typename T>
struct HArgs {
    int Nx, Ny, Nz, P;
    int nbx, nby, nbz, remap;
    Fld<S, T> Hx, Hy, Hz, Ex, Ey, Ez;
    Prof<T> pf;
    int pml[6];
    T *pHyx, *pHzx, *pHxy, *pHzy, *pHxz, *pHyz;
    int en;
    T dtm, dtdmx, dtdmy, dtdmz;
    int inj_k;
    T injx, injy;
    int tfsf, tb0, tb1, tb2, tb3, tb4, tb5, tper_x, tper_y;
    Fld<S, T> aEx, aEy, aEz;
    int lm;                    // bits 0-2: E components loaded, bits 3-5: H components loaded and stored
    // sensor strip: for each sensor plane normal to x at column x = scol[p]
    // (p < nstrip), the H pass also writes E and the new H of the cells of
    // columns x and x+1 (wrapping) to strip[((p Nz + k) Ny + j) 12 + 6 col + c],
    // so that the samples of those planes read contiguous memory instead of one
    // cache line per row and component
    int nstrip;
    int scol[MAX_STRIP];
    T *strip;
};

// Hardware workgroup id -> logical tile (bx, by, bz).  Workgroups are dealt to
// the 8 XCDs round-robin.  With the fold, the tiles are cut into chunks of
// XCD_ROWS rows of tiles (x fastest) and chunk c goes to XCD c mod 8, so that a
// tile's y-1 halo row was mostly loaded by the same XCD's L2 a moment before,
// while the eight XCDs still stream neighbouring chunks at the same time (one
// contiguous range per XCD keeps the halos too, but measured slower).  Tiles
// past the last full round of 8 chunks keep the round-robin order.
constexpr int XCD_ROWS = 4;

__device__ __forceinline__ void tile_of(int w, int nbx, int nby, int nbz, int remap, int &bx, int &by, int &bz) {
    if (remap) {
        const int CH = nbx * XCD_ROWS;
        const int full = nbx * nby * nbz / (8 * CH);
        if (w < full * 8 * CH) {
            const int x = w % 8, i = w / 8;
            w = ((i / CH) * 8 + x) * CH + i % CH;
        }
    }
    bx = w % nbx;
    const int t = w / nbx;
    by = t % nby;
    bz = t / nby;
}

template <typename T>
struct MatL {
    T C1, C2x, C2y, C2z, C4;
    int rc;
};

// A load whose address is the same in every lane of the wavefront (a profile
// value or an auxiliary-grid value of the current plane): through the
// constant address space it becomes a scalar load, which neither occupies
// vector registers nor counts against the vector-memory counter that the
// software pipelines of the passes wait on.  Only for data that the kernel
// itself does not write.
template <typename T>
__device__ __forceinline__ T uload(const T *p, int n) {
    return ((const __attribute__((address_space(4))) T *)p)[n];
}
template <typename S, typename T>
__device__ __forceinline__ T uload(const Fld<S, T> &f, int n) {
    if constexpr (std::is_same_v<S, T>) return uload(f.p, n);
    else return (T)uload(f.p, n) * f.ld;
}

// A CPML psi held in a register and written through to its slab array.
template <typename T>
struct PsiRef {
    T &r;
    T *g;
    __device__ __forceinline__ operator T() const { return r; }
    __device__ __forceinline__ void operator=(T v) const {
        r = v;
        *g = v;
    }
};

// Which boundary treatments a set of cells can need (the side mask): CPML
// slabs or PEC walls along x, y, z, and TFSF surface corrections.  The cell
// updates test their per-cell conditions only under the bits set, so a tile
// that meets no boundary runs the bare stencil.
constexpr int SIDE_X = 1, SIDE_Y = 2, SIDE_Z = 4, SIDE_TFSF = 8;

template <class A>
__device__ __forceinline__ int side_all(const A &a) {
    return ((a.pml[0] || a.pml[1]) ? SIDE_X : 0) | ((a.pml[2] || a.pml[3]) ? SIDE_Y : 0) |
           ((a.pml[4] || a.pml[5]) ? SIDE_Z : 0) | (a.tfsf ? SIDE_TFSF : 0);
}

// [lo, hi) meets [p, q]
__device__ __forceinline__ bool meets(int lo, int hi, int p, int q) { return lo <= q && hi > p; }

// Side mask of the cell block [ilo,ihi) x [jlo,jhi) x [klo,khi) in the E pass
// (h = 0) or the H pass (h = 1).  E cells in a CPML slab lie in 1..pm-1 and
// N-pp+1..N-1 with the PEC wall at 0; H cells in 0..pm-1 and N-pp..N-2 with
// the wall at N-1.  TFSF corrections touch E cells on the box faces and H
// cells one cell outside them; the test here is a superset of both.
template <class A>
__device__ __forceinline__ int side_block(const A &a, int h, int ilo, int ihi, int jlo, int jhi, int klo, int khi) {
    const int N[3] = {a.Nx, a.Ny, a.Nz};
    const int lo[3] = {ilo, jlo, klo}, hi[3] = {ihi, jhi, khi};
    int m = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int pm = a.pml[2 * d], pp = a.pml[2 * d + 1];
        if (!(pm || pp)) continue;
        const bool near = h ? lo[d] < pm : lo[d] < max(pm, 1);
        const bool far = h ? hi[d] > N[d] - max(pp, 1) : hi[d] > N[d] - pp + 1;
        if (near || far) m |= 1 << d;
    }
    if (a.tfsf && meets(ilo, ihi, a.tb0 - 1, a.tb1) && meets(jlo, jhi, a.tb2 - 1, a.tb3) &&
        meets(klo, khi, a.tb4 - 1, a.tb5)) {
        const bool fx = !a.tper_x && (meets(ilo, ihi, a.tb0 - 1, a.tb0) || meets(ilo, ihi, a.tb1 - 1, a.tb1));
        const bool fy = !a.tper_y && (meets(jlo, jhi, a.tb2 - 1, a.tb2) || meets(jlo, jhi, a.tb3 - 1, a.tb3));
        const bool fz = meets(klo, khi, a.tb4 - 1, a.tb4) || meets(klo, khi, a.tb5 - 1, a.tb5);
        if (fx || fy || fz) m |= SIDE_TFSF;
    }
    return m;
}

// ---------------------------------------------------------------- E cell
// CPML psi of an E cell in the grid's slab arrays (x slab: (pxm+pxp, Ny, Nz),
// y slab: (Nx, pym+pyp, Nz), z slab: (Nx, Ny, pzm+pzp)); the argument is the
// cell's slab index along the PML axis.
template <typename S, typename T>
struct PsiAddrE {
    const EArgs<S, T> &a;
    int i, j, k;
    __device__ __forceinline__ T *xy(int n) const { return a.pExy + i + a.Nx * (n + (a.pml[2] + a.pml[3]) * k); }
    __device__ __forceinline__ T *xz(int n) const { return a.pExz + i + a.Nx * j + a.Nx * a.Ny * n; }
    __device__ __forceinline__ T *yx(int n) const { return a.pEyx + n + (a.pml[0] + a.pml[1]) * (j + a.Ny * k); }
    __device__ __forceinline__ T *yz(int n) const { return a.pEyz + i + a.Nx * j + a.Nx * a.Ny * n; }
    __device__ __forceinline__ T *zx(int n) const { return a.pEzx + n + (a.pml[0] + a.pml[1]) * (j + a.Ny * k); }
    __device__ __forceinline__ T *zy(int n) const { return a.pEzy + i + a.Nx * (n + (a.pml[2] + a.pml[3]) * k); }
};

// CPML z-slab index of an E cell at plane k (near 1..pm-1, far N-pp+1..N-1).
template <typename S, typename T>
__device__ __forceinline__ int slabEz(const EArgs<S, T> &a, int k) {
    if (k >= 1 && k < a.pml[4]) return k;
    if (k >= a.Nz - a.pml[5] + 1) return k - (a.Nz - a.pml[5]) + a.pml[4];
    return -1;
}

// Boundary data of an E cell read where they lie: psi in the slab arrays,
// x and y CPML profile values in the profile arrays.
template <typename S, typename T>
struct GlobalPsiE {
    const EArgs<S, T> &a;
    int i, j, k;
    __device__ __forceinline__ T &Exy(int n) const { return *PsiAddrE<S, T>{a, i, j, k}.xy(n); }
    __device__ __forceinline__ T &Exz(int n) const { return *PsiAddrE<S, T>{a, i, j, k}.xz(n); }
    __device__ __forceinline__ T &Eyx(int n) const { return *PsiAddrE<S, T>{a, i, j, k}.yx(n); }
    __device__ __forceinline__ T &Eyz(int n) const { return *PsiAddrE<S, T>{a, i, j, k}.yz(n); }
    __device__ __forceinline__ T &Ezx(int n) const { return *PsiAddrE<S, T>{a, i, j, k}.zx(n); }
    __device__ __forceinline__ T &Ezy(int n) const { return *PsiAddrE<S, T>{a, i, j, k}.zy(n); }
    __device__ __forceinline__ T bx() const { return a.pf.bE[0][i]; }
    __device__ __forceinline__ T cx() const { return a.pf.cE[0][i]; }
    __device__ __forceinline__ T by() const { return a.pf.bE[1][j]; }
    __device__ __forceinline__ T cy() const { return a.pf.cE[1][j]; }
};

// The psi of one E cell loaded ahead of their use (only those of the slabs
// the cell lies in are meaningful).
template <typename T>
struct PsiSetE {
    T xy, xz, yx, yz, zx, zy;
};

// Boundary data of an E cell from registers: psi loaded ahead (written
// through on update) and the lane's x and y profile values.
template <typename S, typename T>
struct PrefPsiE {
    const EArgs<S, T> &a;
    int i, j, k;
    PsiSetE<T> &q;
    T bx_, cx_, by_, cy_;
    __device__ __forceinline__ PsiRef<T> Exy(int n) const { return {q.xy, PsiAddrE<S, T>{a, i, j, k}.xy(n)}; }
    __device__ __forceinline__ PsiRef<T> Exz(int n) const { return {q.xz, PsiAddrE<S, T>{a, i, j, k}.xz(n)}; }
    __device__ __forceinline__ PsiRef<T> Eyx(int n) const { return {q.yx, PsiAddrE<S, T>{a, i, j, k}.yx(n)}; }
    __device__ __forceinline__ PsiRef<T> Eyz(int n) const { return {q.yz, PsiAddrE<S, T>{a, i, j, k}.yz(n)}; }
    __device__ __forceinline__ PsiRef<T> Ezx(int n) const { return {q.zx, PsiAddrE<S, T>{a, i, j, k}.zx(n)}; }
    __device__ __forceinline__ PsiRef<T> Ezy(int n) const { return {q.zy, PsiAddrE<S, T>{a, i, j, k}.zy(n)}; }
    __device__ __forceinline__ T bx() const { return bx_; }
    __device__ __forceinline__ T cx() const { return cx_; }
    __device__ __forceinline__ T by() const { return by_; }
    __device__ __forceinline__ T cy() const { return cy_; }
};

template <typename T>
struct V3 {
    T x, y, z;
};

// The new E of one cell from its old E and the curl differences dH*; the
// CPML psi live where P says.
template <typename S, typename T, class P>
__device__ __forceinline__ V3<T> e_cell_vals(const EArgs<S, T> &a, const MatL<T> &M, int i, int j, int k, int id,
                                             T ex0, T ey0, T ez0, T dHz_y, T dHy_z, T dHx_z, T dHz_x, T dHy_x,
                                             T dHx_y, T ikx, T iky, T ikz, int nx_, int ny_, int ostep, int side,
                                             P ps) {
    // stencil (core/fdtd_core.cpp:414-415, 528-529, 658-659)
    T ex = (a.en & 1) ? M.C1 * ex0 + M.C2y * iky * dHz_y - M.C2z * ikz * dHy_z : ex0;
    T ey = (a.en & 2) ? M.C1 * ey0 + M.C2z * ikz * dHx_z - M.C2x * ikx * dHz_x : ey0;
    T ez = (a.en & 4) ? M.C1 * ez0 + M.C2x * ikx * dHy_x - M.C2y * iky * dHx_y : ez0;
    // TFSF-E (modes/single_particle.cpp:286-332): z faces, then y, then x
    if (side & SIDE_TFSF) {
        const bool in_i = i >= a.tb0 && i < a.tb1, in_j = j >= a.tb2 && j < a.tb3, in_k = k >= a.tb4 && k < a.tb5;
        if (in_i && in_j) {
            if (k == a.tb4) { ex += M.C2z * uload(a.aHy, a.tb4 - 1); ey -= M.C2z * uload(a.aHx, a.tb4 - 1); }
            if (k == a.tb5) { ex -= M.C2z * uload(a.aHy, a.tb5 - 1); ey += M.C2z * uload(a.aHx, a.tb5 - 1); }
        }
        if (!a.tper_y && in_i && in_k) {
            if (j == a.tb2) { ex -= M.C2y * uload(a.aHz, k); ez += M.C2y * uload(a.aHx, k); }
            if (j == a.tb3) { ex += M.C2y * uload(a.aHz, k); ez -= M.C2y * uload(a.aHx, k); }
        }
        if (!a.tper_x && in_k && in_j) {
            if (i == a.tb0) { ey += M.C2x * uload(a.aHz, k); ez -= M.C2x * uload(a.aHy, k); }
            if (i == a.tb1) { ey -= M.C2x * uload(a.aHz, k); ez += M.C2x * uload(a.aHy, k); }
        }
    }
    // RC-apply: E += C3 sum_p Re psi (core/mats_RC.cpp:38-50); the sum was
    // formed from E^n by the RC pre-pass (k_rc_ante) of this step
    if (M.rc) {
        const T *cr = a.corr + 3 * a.slot[id];
        ex += cr[0];
        ey += cr[1];
        ez += cr[2];
    }
    // CPML-E + PEC walls (core/fdtd_pml.cpp:504-735)
    const bool px = side & SIDE_X, py = side & SIDE_Y, pz = side & SIDE_Z;
    const int nz_ = pz ? slabEz(a, k) : -1;
    if (a.en & 1) {
        if (py) {
            if (ny_ >= 0) {
                auto &&q = ps.Exy(ny_);
                q = ps.by() * q + ps.cy() * dHz_y;
                ex += M.C4 * q;
            }
            if (j == 0) ex = 0;
        }
        if (pz) {
            if (nz_ >= 0) {
                auto &&q = ps.Exz(nz_);
                q = uload(a.pf.bE[2], k) * q + uload(a.pf.cE[2], k) * dHy_z;
                ex -= M.C4 * q;
            }
            if (k == 0) ex = 0;
        }
    }
    if (a.en & 2) {
        if (px) {
            if (nx_ >= 0) {
                auto &&q = ps.Eyx(nx_);
                q = ps.bx() * q + ps.cx() * dHz_x;
                ey -= M.C4 * q;
            }
            if (i == 0) ey = 0;
        }
        if (pz) {
            if (nz_ >= 0) {
                auto &&q = ps.Eyz(nz_);
                q = uload(a.pf.bE[2], k) * q + uload(a.pf.cE[2], k) * dHx_z;
                ey += M.C4 * q;
            }
            if (k == 0) ey = 0;
        }
    }
    if (a.en & 4) {
        if (px) {
            if (nx_ >= 0) {
                auto &&q = ps.Ezx(nx_);
                q = ps.bx() * q + ps.cx() * dHy_x;
                ez += M.C4 * q;
            }
            if (i == 0) ez = 0;
        }
        if (py) {
            if (ny_ >= 0) {
                auto &&q = ps.Ezy(ny_);
                q = ps.by() * q + ps.cy() * dHx_y;
                ez -= M.C4 * q;
            }
            if (j == 0) ez = 0;
        }
    }
    // plane injection after the whole E update (modes/normal_incidence.cpp:338-342)
    if (k == a.inj_k) { ex += a.injx; ey += a.injy; }
    // oscillator sources, in registration order (modes/default_fdtd.cpp:158-159)
    for (int q = 0; q < a.nosc; q++) {
        const OscDev &o = a.osc[q];
        if (i >= o.x1 && i < o.x2 && j >= o.y1 && j < o.y2 && k >= o.z1 && k < o.z2) {
            const T v = (T)uload(o.val, ostep - o.step0);
            if (o.comp == 0) ex += v;
            else if (o.comp == 1) ey += v;
            else ez += v;
        }
    }
    return V3<T>{ex, ey, ez};
}

// Everything after the loads of one E cell, with the grid's CPML state; the
// components in store mask `smask` (bit c: component c) are written back.
template <typename S, typename T, class P>
__device__ __forceinline__ V3<T> e_cell(const EArgs<S, T> &a, const MatL<T> &M, int m, int i, int j, int k, int id,
                                       T ex0, T ey0, T ez0, T dHz_y, T dHy_z, T dHx_z, T dHz_x, T dHy_x, T dHx_y,
                                       T ikx, T iky, T ikz, int nx_, int ny_, int ostep, int side, int smask, P ps) {
    const V3<T> e = e_cell_vals(a, M, i, j, k, id, ex0, ey0, ez0, dHz_y, dHy_z, dHx_z, dHz_x, dHy_x, dHx_y, ikx, iky,
                                ikz, nx_, ny_, ostep, side, ps);
    if (smask & STORE_NT) {
        a.Ex.set_nt(id, e.x);
        a.Ey.set_nt(id, e.y);
        a.Ez.set_nt(id, e.z);
        return e;
    }
    if (smask & 1) a.Ex.set(id, e.x);
    if (smask & 2) a.Ey.set(id, e.y);
    if (smask & 4) a.Ez.set(id, e.z);
    return e;
}

// CPML slab index of an E cell along an axis (near 1..pm-1, far N-pp+1..N-1).
__device__ __forceinline__ int slabE(int v, int N, int pm, int pp) {
    if (!(pm || pp)) return -1;
    if (v >= 1 && v < pm) return v;
    if (v >= N - pp + 1) return v - (N - pp) + pm;
    return -1;
}
// ... of an H cell (near 0..pm-1, far N-pp..N-2).
__device__ __forceinline__ int slabH(int v, int N, int pm, int pp) {
    if (!(pm || pp)) return -1;
    if (v < pm) return v;
    if (v >= N - pp && v < N - 1) return v - (N - pp) + pm;
    return -1;
}

// RC pre-pass over the dispersive cells (slot order): psi = Crec psi + Dchi E^n
// (core/mats_RC.cpp:21-36) in FP64, then the E increment (T)(C3 sum_p Re psi)
// of RC_apply_E (:38-50), which the E pass adds after its stencil and TFSF.
template <typename S, typename T>
struct RcArgs {
    int ndisp, np_max;              // np_max: pole states allocated per slot and component
    const int2 *cells;              // per slot: its cell and material
    const MatDev<T> *mt;
    const double2 *poles;       // [pole]: Crec at 2p, Dchi at 2p+1
    double2 *psi;               // [3 * pole + component][slot]
    Fld<S, T> Ex, Ey, Ez;
    T *corr;
};

// One lane per (dispersive slot, component).  The lane's pole states depend
// only on the slot, so the first RC_PRELOAD of them (up to the grid's largest
// pole count) are requested at once with the slot's (cell, material) pair;
// the cell's field and the material's constants follow, and the pole table
// after them: two memory latencies in a row instead of the four of reading
// the cell, then its material id, then the material, then its poles.
constexpr int RC_PRELOAD = 8;

template <typename S, typename T>
__device__ __forceinline__ void rc_lane(const RcArgs<S, T> &a, int s, int c) {
    double2 *q = a.psi + (size_t)c * a.ndisp + s;               // pole p at q[3 p ndisp]
    const size_t stride = (size_t)3 * a.ndisp;
    double2 v[RC_PRELOAD];
#pragma unroll
    for (int p = 0; p < RC_PRELOAD; p++)
        if (p < a.np_max) v[p] = q[p * stride];
    const int2 cm = a.cells[s];
    const int id = cm.x;
    const MatDev<T> &G = a.mt[cm.y];
    const Fld<S, T> &E = c == 0 ? a.Ex : (c == 1 ? a.Ey : a.Ez);
    const double e = (double)E[id];
    const int np = G.np;
    double sum = 0;
    for (int p = 0; p < np; p++) {
        const double2 cr = a.poles[2 * (G.pole0 + p)], dc = a.poles[2 * (G.pole0 + p) + 1];
        double2 x = make_double2(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < RC_PRELOAD; u++)
            if (u == p) x = v[u];
        if (p >= RC_PRELOAD) x = q[p * stride];
        double2 r;
        r.x = (cr.x * x.x - cr.y * x.y) + dc.x * e;
        r.y = (cr.x * x.y + cr.y * x.x) + dc.y * e;
        q[p * stride] = r;
        sum += r.x;
    }
    a.corr[3 * s + c] = (T)(G.C3 * sum);
}

template <typename S, typename T>
__global__ __launch_bounds__(256) void k_rc_ante(const RcArgs<S, T> a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < a.ndisp) rc_lane(a, s, blockIdx.y);
}

template <typename S, typename T>
__device__ __forceinline__ void stage_mats(const EArgs<S, T> &a, MatL<T> *sm) {
    for (int m = threadIdx.x + blockDim.x * threadIdx.y; m < a.nmat; m += blockDim.x * blockDim.y) {
        const MatDev<T> &G = a.mt[m];
        sm[m] = MatL<T>{G.C1, G.C2x, G.C2y, G.C2z, G.C4, G.rc};
    }
    __syncthreads();
}

// 3D E pass over tile w of the grid.
// Every component is loaded and stored here, whatever the load mask says: a
// component outside the mask is either never written (it keeps the value it
// is loaded with) or only differenced along a one-cell axis, where the
// periodic difference of a finite value is exactly zero, so the results are
// those of the masked loads.  Unconditional loads keep the number of loads
// per plane fixed, which lets the compiler wait for exactly one plane while
// the next plane's loads stay in flight (software pipeline: plane k+1 is
// requested before plane k is computed and stored; the last plane of the
// tile is peeled off so that the loop body never skips its loads).
// With EN the lane also returns the sum of Ex^2+Ey^2+Ez^2 of the values it
// stored (all three are stored in 3D), for the energy completion check; the
// lane sums its KZ cells in the field type (one register more than without
// the sum, which keeps the kernel at its occupancy), the workgroup in FP64.
template <typename S, typename T, int BY, int KZ, bool EN = false>
__device__ __forceinline__ double e_tile(const EArgs<S, T> &a, const MatL<T> *sm, int w, int ostep) {
    int bx, by, bz;
    tile_of(w, a.nbx, a.nby, a.nbz, a.remap, bx, by, bz);
    const int i = bx * BX + threadIdx.x;
    const int j = by * BY + threadIdx.y;
    if (i >= a.Nx || j >= a.Ny) return 0.0;
    const int k0 = bz * KZ;
    const int k1 = min(k0 + KZ, a.Nz);
    const int NxNy = a.P * a.Ny;                                 // plane stride
    T en = 0;
    const int base = i + a.P * j;
    const int oim = (i == 0) ? a.Nx - 1 : -1;                    // (i-1) with wrap
    const int ojm = (j == 0) ? (a.Ny - 1) * a.P : -a.P;          // (j-1) with wrap
    const int km0 = (k0 == 0) ? a.Nz - 1 : k0 - 1;
    T hx_km = a.Hx[base + km0 * NxNy];
    T hy_km = a.Hy[base + km0 * NxNy];
    const T ikx = a.pf.ikE[0][i], iky = a.pf.ikE[1][j];
    const int nx_ = slabE(i, a.Nx, a.pml[0], a.pml[1]);
    const int ny_ = slabE(j, a.Ny, a.pml[2], a.pml[3]);
    const int side = side_block(a, 0, bx * BX, min(bx * BX + BX, a.Nx), by * BY, min(by * BY + BY, a.Ny), k0, k1);
    // the lane's x and y CPML profile values (loop invariant)
    T bx_ = 0, cx_ = 0, by_ = 0, cy_ = 0;
    if ((side & SIDE_X) && nx_ >= 0) { bx_ = a.pf.bE[0][i]; cx_ = a.pf.cE[0][i]; }
    if ((side & SIDE_Y) && ny_ >= 0) { by_ = a.pf.bE[1][j]; cy_ = a.pf.cE[1][j]; }
    struct Ld {
        T hx, hy, hz, hz_jm, hz_im, hy_im, hx_jm, ex, ey, ez;
        int m;
    };
    auto load = [&](int k) {
        const int id = base + k * NxNy;
        Ld v;
        v.hx = a.Hx[id];
        v.hx_jm = a.Hx[id + ojm];
        v.hy = a.Hy[id];
        v.hy_im = a.Hy[id + oim];
        v.hz = a.Hz[id];
        v.hz_jm = a.Hz[id + ojm];
        v.hz_im = a.Hz[id + oim];
        v.ex = a.Ex.nt(id);          // read once per pass: streamed
        v.ey = a.Ey.nt(id);
        v.ez = a.Ez.nt(id);
        v.m = __builtin_nontemporal_load(a.mat + id);
        return v;
    };
    // The psi of a plane are requested with the plane's fields, one plane
    // ahead of their use.
    auto load_psi = [&](int k) {
        PsiSetE<T> q{};
        const PsiAddrE<S, T> ad{a, i, j, k};
        if ((side & SIDE_X) && nx_ >= 0) {
            if (a.en & 2) q.yx = *ad.yx(nx_);
            if (a.en & 4) q.zx = *ad.zx(nx_);
        }
        if ((side & SIDE_Y) && ny_ >= 0) {
            if (a.en & 1) q.xy = *ad.xy(ny_);
            if (a.en & 4) q.zy = *ad.zy(ny_);
        }
        if (side & SIDE_Z) {
            const int nz_ = slabEz(a, k);
            if (nz_ >= 0) {
                if (a.en & 1) q.xz = *ad.xz(nz_);
                if (a.en & 2) q.yz = *ad.yz(nz_);
            }
        }
        return q;
    };
    auto cell = [&](const Ld &c, PsiSetE<T> &q, int k) {
        const T dHy_z = c.hy - hy_km;
        const T dHx_z = c.hx - hx_km;
        hx_km = c.hx;
        hy_km = c.hy;
        const V3<T> e = e_cell(a, sm[c.m], c.m, i, j, k, base + k * NxNy, c.ex, c.ey, c.ez, c.hz - c.hz_jm, dHy_z,
                               dHx_z, c.hz - c.hz_im, c.hy - c.hy_im, c.hx - c.hx_jm, ikx, iky, uload(a.pf.ikE[2], k),
                               nx_, ny_, ostep, side, STORE_NT, PrefPsiE<S, T>{a, i, j, k, q, bx_, cx_, by_, cy_});
        if constexpr (EN) en += e.x * e.x + e.y * e.y + e.z * e.z;
    };
    Ld cur = load(k0);
    PsiSetE<T> q = load_psi(k0);
    for (int k = k0; k < k1 - 1; k++) {
        PsiSetE<T> qn = load_psi(k + 1);
        const Ld nxt = load(k + 1);
        cell(cur, q, k);
        cur = nxt;
        q = qn;
    }
    cell(cur, q, k1 - 1);
    return (double)en;
}

// 2D E pass (Nz == 1) over tile w: each lane marches KY consecutive rows in y
// and carries the previous row's Hx, Hz in registers, so the per-lane set-up
// is paid once per KY cells and every row load is a contiguous 64-wide read.
// Differences along the one-cell z axis are exactly zero.
template <typename S, typename T, int BY, int KY>
__device__ __forceinline__ void e_tile2d(const EArgs<S, T> &a, const MatL<T> *sm, int w, int ostep) {
    int bx, by, bz;
    tile_of(w, a.nbx, a.nby, a.nbz, a.remap, bx, by, bz);
    const int i = bx * BX + threadIdx.x;
    const int j0 = (by * BY + threadIdx.y) * KY;
    if (i >= a.Nx || j0 >= a.Ny) return;
    const int j1 = min(j0 + KY, a.Ny);
    const int oim = (i == 0) ? a.Nx - 1 : -1;
    const int lm = a.lm;
    const T zero = 0;
    const int jm0 = (j0 == 0) ? a.Ny - 1 : j0 - 1;
    T hx_jm = (lm & 8) ? a.Hx[i + a.P * jm0] : zero;
    T hz_jm = (lm & 32) ? a.Hz[i + a.P * jm0] : zero;
    const T ikx = a.pf.ikE[0][i];
    const int nx_ = slabE(i, a.Nx, a.pml[0], a.pml[1]);
    const T ikz = a.pf.ikE[2][0];
    struct Ld {
        T hx, hy, hz, hz_im, hy_im, ex, ey, ez;
        int m;
    };
    auto load = [&](int id) {
        Ld v;
        v.hx = zero; v.hy = zero; v.hy_im = zero; v.hz = zero; v.hz_im = zero;
        v.ex = zero; v.ey = zero; v.ez = zero;
        if (lm & 8) v.hx = a.Hx[id];
        if (lm & 16) { v.hy = a.Hy[id]; v.hy_im = a.Hy[id + oim]; }
        if (lm & 32) { v.hz = a.Hz[id]; v.hz_im = a.Hz[id + oim]; }
        if (lm & 1) v.ex = a.Ex[id];
        if (lm & 2) v.ey = a.Ey[id];
        if (lm & 4) v.ez = a.Ez[id];
        v.m = a.mat[id];
        return v;
    };
    Ld cur = load(i + a.P * j0);
    for (int j = j0; j < j1; j++) {
        const int id = i + a.P * j;
        Ld nxt = cur;
        if (j + 1 < j1) nxt = load(id + a.P);
        const T iky = a.pf.ikE[1][j];
        const int ny_ = slabE(j, a.Ny, a.pml[2], a.pml[3]);
        e_cell(a, sm[cur.m], cur.m, i, j, 0, id, cur.ex, cur.ey, cur.ez, cur.hz - hz_jm, zero, zero, cur.hz - cur.hz_im,
               cur.hy - cur.hy_im, cur.hx - hx_jm, ikx, iky, ikz, nx_, ny_, ostep, side_all(a), lm,
               GlobalPsiE<S, T>{a, i, j, 0});
        hx_jm = cur.hx;
        hz_jm = cur.hz;
        cur = nxt;
    }
}

template <typename S, typename T, int BY, int KY>
__global__ __launch_bounds__(BX *BY) void k_update_E2(const EArgs<S, T> a) {
    extern __shared__ char lds_raw[];
    MatL<T> *sm = reinterpret_cast<MatL<T> *>(lds_raw);
    stage_mats(a, sm);
    e_tile2d<S, T, BY, KY>(a, sm, blockIdx.x, a.ostep);
}

template <typename S, typename T>
__device__ __forceinline__ void e_col_cell(const EArgs<S, T> &a, const MatL<T> *sm, int k);

// 3D E pass; workgroups from nb on update the 1 x 1 x Nz column `aux` instead
// (the TFSF source grid, stepped in the same launch: it neither reads nor
// writes anything the 3D pass touches).
// With EN, tile workgroup w also writes en_part[w], the sum of
// Ex^2+Ey^2+Ez^2 over its cells after the update (lanes in a fixed order, so
// the partials are the same on every run).
template <typename S, typename T, int BY, int KZ, bool EN = false>
__global__ __launch_bounds__(BX *BY) void k_update_E(const EArgs<S, T> a, const EArgs<S, T> aux, int nb) {
    extern __shared__ char lds_raw[];          // max(nmat) entries, sized at launch
    MatL<T> *sm = reinterpret_cast<MatL<T> *>(lds_raw);
    if ((int)blockIdx.x >= nb) {
        stage_mats(aux, sm);
        const int k = ((int)blockIdx.x - nb) * (BX * BY) + threadIdx.x + BX * threadIdx.y;
        if (k < aux.Nz) e_col_cell(aux, sm, k);
        return;
    }
    stage_mats(a, sm);
    double en = e_tile<S, T, BY, KZ, EN>(a, sm, blockIdx.x, a.ostep);
    if constexpr (EN) {
        __shared__ double red[BY];
        for (int o = 32; o > 0; o >>= 1) en += __shfl_xor(en, o, 64);
        if (threadIdx.x == 0) red[threadIdx.y] = en;
        __syncthreads();
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            double t = 0;
            for (int y = 0; y < BY; y++) t += red[y];
            a.en_part[blockIdx.x] = t;
        }
    }
}

// Workgroup -> grid of a batch: grid r owns workgroups [wg0[r], wg0[r+1]),
// each range starting at a multiple of 8 so that the XCD fold of tile_of
// holds inside every grid.
__device__ __forceinline__ int batch_member(const int *wg0, int n) {
    const int w = blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (wg0[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// E pass of many independent grids in one launch (the batching form): step s
// of a chunk that started at step tcur[r] of grid r; grids already at their
// last step (tend) do nothing.
template <typename S, typename T, int BY, int KZ>
__global__ __launch_bounds__(BX *BY) void k_update_E_batch(const EArgs<S, T> *__restrict__ d,
                                                           const int *__restrict__ wg0, int n,
                                                           const int *__restrict__ tcur,
                                                           const int *__restrict__ tend, int s) {
    extern __shared__ char lds_raw[];
    MatL<T> *sm = reinterpret_cast<MatL<T> *>(lds_raw);
    const int r = batch_member(wg0, n);
    const EArgs<S, T> &a = d[r];
    const int t = tcur[r] + s;
    const int w = blockIdx.x - wg0[r];
    if (t >= tend[r] || w >= a.nbx * a.nby * a.nbz) return;
    stage_mats(a, sm);
    if (a.Nz == 1) e_tile2d<S, T, BY, KZ>(a, sm, w, t);
    else e_tile<S, T, BY, KZ>(a, sm, w, t);
}

// E pass of a 1 x 1 x Nz column (the auxiliary grid of the TFSF source):
// one lane per cell.
template <typename S, typename T>
__device__ __forceinline__ void e_col_cell(const EArgs<S, T> &a, const MatL<T> *sm, int k) {
    const int km = (k == 0) ? a.Nz - 1 : k - 1;
    const T hx = a.Hx[k], hy = a.Hy[k];
    const int m = a.mat[k];
    const T zero = 0;
    e_cell(a, sm[m], m, 0, 0, k, k, a.Ex[k], a.Ey[k], a.Ez[k], zero, hy - a.Hy[km], hx - a.Hx[km], zero, zero,
           zero, a.pf.ikE[0][0], a.pf.ikE[1][0], a.pf.ikE[2][k], -1, -1, a.ostep, side_all(a), a.lm,
           GlobalPsiE<S, T>{a, 0, 0, k});
}

template <typename S, typename T>
__global__ __launch_bounds__(256) void k_update_E_col(const EArgs<S, T> a) {
    extern __shared__ char lds_raw[];
    MatL<T> *sm = reinterpret_cast<MatL<T> *>(lds_raw);
    stage_mats(a, sm);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < a.Nz) e_col_cell(a, sm, k);
}

// ---------------------------------------------------------------- H cell
// CPML psi of an H cell in the grid's slab arrays (layout as for E).
template <typename S, typename T>
struct PsiAddrH {
    const HArgs<S, T> &a;
    int i, j, k;
    __device__ __forceinline__ T *xy(int n) const { return a.pHxy + i + a.Nx * (n + (a.pml[2] + a.pml[3]) * k); }
    __device__ __forceinline__ T *xz(int n) const { return a.pHxz + i + a.Nx * j + a.Nx * a.Ny * n; }
    __device__ __forceinline__ T *yx(int n) const { return a.pHyx + n + (a.pml[0] + a.pml[1]) * (j + a.Ny * k); }
    __device__ __forceinline__ T *yz(int n) const { return a.pHyz + i + a.Nx * j + a.Nx * a.Ny * n; }
    __device__ __forceinline__ T *zx(int n) const { return a.pHzx + n + (a.pml[0] + a.pml[1]) * (j + a.Ny * k); }
    __device__ __forceinline__ T *zy(int n) const { return a.pHzy + i + a.Nx * (n + (a.pml[2] + a.pml[3]) * k); }
};

template <typename S, typename T>
struct GlobalPsiH {
    const HArgs<S, T> &a;
    int i, j, k;
    __device__ __forceinline__ T &Hxy(int n) const { return *PsiAddrH<S, T>{a, i, j, k}.xy(n); }
    __device__ __forceinline__ T &Hxz(int n) const { return *PsiAddrH<S, T>{a, i, j, k}.xz(n); }
    __device__ __forceinline__ T &Hyx(int n) const { return *PsiAddrH<S, T>{a, i, j, k}.yx(n); }
    __device__ __forceinline__ T &Hyz(int n) const { return *PsiAddrH<S, T>{a, i, j, k}.yz(n); }
    __device__ __forceinline__ T &Hzx(int n) const { return *PsiAddrH<S, T>{a, i, j, k}.zx(n); }
    __device__ __forceinline__ T &Hzy(int n) const { return *PsiAddrH<S, T>{a, i, j, k}.zy(n); }
    __device__ __forceinline__ T bx() const { return a.pf.bH[0][i]; }
    __device__ __forceinline__ T cx() const { return a.pf.cH[0][i]; }
    __device__ __forceinline__ T by() const { return a.pf.bH[1][j]; }
    __device__ __forceinline__ T cy() const { return a.pf.cH[1][j]; }
};

template <typename T>
struct PsiSetH {
    T xy, xz, yx, yz, zx, zy;
};

template <typename S, typename T>
struct PrefPsiH {
    const HArgs<S, T> &a;
    int i, j, k;
    PsiSetH<T> &q;
    T bx_, cx_, by_, cy_;
    __device__ __forceinline__ PsiRef<T> Hxy(int n) const { return {q.xy, PsiAddrH<S, T>{a, i, j, k}.xy(n)}; }
    __device__ __forceinline__ PsiRef<T> Hxz(int n) const { return {q.xz, PsiAddrH<S, T>{a, i, j, k}.xz(n)}; }
    __device__ __forceinline__ PsiRef<T> Hyx(int n) const { return {q.yx, PsiAddrH<S, T>{a, i, j, k}.yx(n)}; }
    __device__ __forceinline__ PsiRef<T> Hyz(int n) const { return {q.yz, PsiAddrH<S, T>{a, i, j, k}.yz(n)}; }
    __device__ __forceinline__ PsiRef<T> Hzx(int n) const { return {q.zx, PsiAddrH<S, T>{a, i, j, k}.zx(n)}; }
    __device__ __forceinline__ PsiRef<T> Hzy(int n) const { return {q.zy, PsiAddrH<S, T>{a, i, j, k}.zy(n)}; }
    __device__ __forceinline__ T bx() const { return bx_; }
    __device__ __forceinline__ T cx() const { return cx_; }
    __device__ __forceinline__ T by() const { return by_; }
    __device__ __forceinline__ T cy() const { return cy_; }
};

// The new H of one cell (in place in hx, hy, hz) from the curl differences dE*.
template <typename S, typename T, class P>
__device__ __forceinline__ void h_cell_vals(const HArgs<S, T> &a, int i, int j, int k, T &hx, T &hy, T &hz, T dEy_z,
                                            T dEx_z, T dEz_y, T dEz_x, T dEx_y, T dEy_x, T ikx, T iky, T ikz, int nx_,
                                            int ny_, int side, P ps) {
    // stencil (core/fdtd_core.cpp:756-757, 852-853, 964-965)
    if (a.en & 8) hx += a.dtdmz * ikz * dEy_z - a.dtdmy * iky * dEz_y;
    if (a.en & 16) hy += a.dtdmx * ikx * dEz_x - a.dtdmz * ikz * dEx_z;
    if (a.en & 32) hz += a.dtdmy * iky * dEx_y - a.dtdmx * ikx * dEy_x;
    // CPML-H + PEC walls (core/fdtd_pml.cpp:737-929)
    const bool px = side & SIDE_X, py = side & SIDE_Y, pz = side & SIDE_Z;
    const int nz_ = pz ? slabH(k, a.Nz, a.pml[4], a.pml[5]) : -1;
    if (a.en & 8) {
        if (py) {
            if (ny_ >= 0) {
                auto &&q = ps.Hxy(ny_);
                q = ps.by() * q + ps.cy() * dEz_y;
                hx -= a.dtm * q;
            }
            if (j == a.Ny - 1) hx = 0;
        }
        if (pz) {
            if (nz_ >= 0) {
                auto &&q = ps.Hxz(nz_);
                q = uload(a.pf.bH[2], k) * q + uload(a.pf.cH[2], k) * dEy_z;
                hx += a.dtm * q;
            }
            if (k == a.Nz - 1) hx = 0;
        }
    }
    if (a.en & 16) {
        if (px) {
            if (nx_ >= 0) {
                auto &&q = ps.Hyx(nx_);
                q = ps.bx() * q + ps.cx() * dEz_x;
                hy += a.dtm * q;
            }
            if (i == a.Nx - 1) hy = 0;
        }
        if (pz && nz_ >= 0) {
            auto &&q = ps.Hyz(nz_);
            q = uload(a.pf.bH[2], k) * q + uload(a.pf.cH[2], k) * dEx_z;
            hy -= a.dtm * q;
        }
    }
    // top z wall of Hy: zeroed by the Hx pass and by the Hy pass (fdtd_pml.cpp:799,862-865)
    if (pz && k == a.Nz - 1 && (a.en & (8 | 16))) hy = 0;
    if (a.en & 32) {
        if (px) {
            if (nx_ >= 0) {
                auto &&q = ps.Hzx(nx_);
                q = ps.bx() * q + ps.cx() * dEy_x;
                hz -= a.dtm * q;
            }
            if (i == a.Nx - 1) hz = 0;
        }
        if (py) {
            if (ny_ >= 0) {
                auto &&q = ps.Hzy(ny_);
                q = ps.by() * q + ps.cy() * dEx_y;
                hz += a.dtm * q;
            }
            if (j == a.Ny - 1) hz = 0;
        }
    }
    // plane injection (modes/normal_incidence.cpp:353-357)
    if (k == a.inj_k) { hx += a.injx; hy += a.injy; }
    // TFSF-H (modes/single_particle.cpp:351-378): z faces, then y, then x
    if (side & SIDE_TFSF) {
        const bool in_i = i >= a.tb0 && i < a.tb1, in_j = j >= a.tb2 && j < a.tb3, in_k = k >= a.tb4 && k < a.tb5;
        if (in_i && in_j) {
            if (k == a.tb4 - 1) { hx -= a.dtdmz * uload(a.aEy, a.tb4); hy += a.dtdmz * uload(a.aEx, a.tb4); }
            if (k == a.tb5 - 1) { hx += a.dtdmz * uload(a.aEy, a.tb5); hy -= a.dtdmz * uload(a.aEx, a.tb5); }
        }
        if (!a.tper_y && in_i && in_k) {
            if (j == a.tb2 - 1) { hx += a.dtdmy * uload(a.aEz, k); hz -= a.dtdmy * uload(a.aEx, k); }
            if (j == a.tb3 - 1) { hx -= a.dtdmy * uload(a.aEz, k); hz += a.dtdmy * uload(a.aEx, k); }
        }
        if (!a.tper_x && in_k && in_j) {
            if (i == a.tb0 - 1) { hy -= a.dtdmx * uload(a.aEz, k); hz += a.dtdmx * uload(a.aEy, k); }
            if (i == a.tb1 - 1) { hy += a.dtdmx * uload(a.aEz, k); hz -= a.dtdmx * uload(a.aEy, k); }
        }
    }
}

// The new H of one cell written back for the components in store mask
// `smask` (bit 3 + c: component c).
template <typename S, typename T, class P>
__device__ __forceinline__ V3<T> h_cell(const HArgs<S, T> &a, int i, int j, int k, int id, T hx, T hy, T hz, T dEy_z,
                                        T dEx_z, T dEz_y, T dEz_x, T dEx_y, T dEy_x, T ikx, T iky, T ikz, int nx_,
                                        int ny_, int side, int smask, P ps) {
    h_cell_vals(a, i, j, k, hx, hy, hz, dEy_z, dEx_z, dEz_y, dEz_x, dEx_y, dEy_x, ikx, iky, ikz, nx_, ny_, side, ps);
    if (smask & STORE_NT) {
        a.Hx.set_nt(id, hx);
        a.Hy.set_nt(id, hy);
        a.Hz.set_nt(id, hz);
    } else {
        if (smask & 8) a.Hx.set(id, hx);
        if (smask & 16) a.Hy.set(id, hy);
        if (smask & 32) a.Hz.set(id, hz);
    }
    return V3<T>{hx, hy, hz};
}

// The value a field component reads back after storing v (FP16 storage
// rounds; FP32/FP64 storage is exact).
template <typename S, typename T>
__device__ __forceinline__ T stored(const Fld<S, T> &f, T v) {
    if constexpr (std::is_same_v<S, T>) return v;
    else return (T)(S)(v * f.st) * f.ld;
}

// Six values of one cell into strip slot s = 2 p + col (three 2-wide stores).
template <typename S, typename T>
__device__ __forceinline__ void strip_put(const HArgs<S, T> &a, int s, int j, int k, const T v[6]) {
    using T2 = std::conditional_t<std::is_same_v<T, double>, double2, float2>;
    T2 *sp = reinterpret_cast<T2 *>(a.strip + (((size_t)(s >> 1) * a.Nz + k) * a.Ny + j) * 12 + 6 * (s & 1));
    sp[0] = T2{v[0], v[1]};
    sp[1] = T2{v[2], v[3]};
    sp[2] = T2{v[4], v[5]};
}

// 3D H pass over tile w of the grid: the mirror of e_tile (every component
// loaded and stored, plane k+1 requested before plane k is computed, last
// plane peeled).  E(k+1) of plane k is E(k) of plane k+1 and is carried.
template <typename S, typename T, int BY, int KZ>
__device__ __forceinline__ void h_tile(const HArgs<S, T> &a, int w) {
    int bx, by, bz;
    tile_of(w, a.nbx, a.nby, a.nbz, a.remap, bx, by, bz);
    const int i = bx * BX + threadIdx.x;
    const int j = by * BY + threadIdx.y;
    if (i >= a.Nx || j >= a.Ny) return;
    const int k0 = bz * KZ;
    const int k1 = min(k0 + KZ, a.Nz);
    const int NxNy = a.P * a.Ny;
    const int base = i + a.P * j;
    const int oip = (i == a.Nx - 1) ? -(a.Nx - 1) : 1;
    const int ojp = (j == a.Ny - 1) ? -(a.Ny - 1) * a.P : a.P;
    const T ikx = a.pf.ikH[0][i], iky = a.pf.ikH[1][j];
    const int nx_ = slabH(i, a.Nx, a.pml[0], a.pml[1]);
    const int ny_ = slabH(j, a.Ny, a.pml[2], a.pml[3]);
    const int side = side_block(a, 1, bx * BX, min(bx * BX + BX, a.Nx), by * BY, min(by * BY + BY, a.Ny), k0, k1);
    T bx_ = 0, cx_ = 0, by_ = 0, cy_ = 0;
    if ((side & SIDE_X) && nx_ >= 0) { bx_ = a.pf.bH[0][i]; cx_ = a.pf.cH[0][i]; }
    if ((side & SIDE_Y) && ny_ >= 0) { by_ = a.pf.bH[1][j]; cy_ = a.pf.cH[1][j]; }
    // strip slots (2 p + col) of the lane's column; a column can serve two planes
    int sq0 = -1, sq1 = -1;
    for (int q = 0; q < a.nstrip; q++) {
        const int x0 = a.scol[q], x1 = (x0 + 1 == a.Nx) ? 0 : x0 + 1;
        const int s = (i == x0) ? 2 * q : (i == x1 ? 2 * q + 1 : -1);
        if (s >= 0) {
            if (sq0 < 0) sq0 = s;
            else sq1 = s;
        }
    }
    T ex_k = a.Ex[base + k0 * NxNy];
    T ey_k = a.Ey[base + k0 * NxNy];
    struct Ld {
        T ex_kp, ey_kp, ex_jp, ey_ip, ez, ez_jp, ez_ip, hx, hy, hz;
    };
    auto load = [&](int k) {
        const int id = base + k * NxNy;
        const int idkp = base + ((k == a.Nz - 1) ? 0 : k + 1) * NxNy;
        Ld v;
        v.ex_kp = a.Ex[idkp];
        v.ey_kp = a.Ey[idkp];
        v.ex_jp = a.Ex[id + ojp];
        v.ey_ip = a.Ey[id + oip];
        v.ez = a.Ez[id];
        v.ez_jp = a.Ez[id + ojp];
        v.ez_ip = a.Ez[id + oip];
        v.hx = a.Hx.nt(id);          // read once per pass: streamed
        v.hy = a.Hy.nt(id);
        v.hz = a.Hz.nt(id);
        return v;
    };
    auto load_psi = [&](int k) {
        PsiSetH<T> q{};
        const PsiAddrH<S, T> ad{a, i, j, k};
        if ((side & SIDE_X) && nx_ >= 0) {
            if (a.en & 16) q.yx = *ad.yx(nx_);
            if (a.en & 32) q.zx = *ad.zx(nx_);
        }
        if ((side & SIDE_Y) && ny_ >= 0) {
            if (a.en & 8) q.xy = *ad.xy(ny_);
            if (a.en & 32) q.zy = *ad.zy(ny_);
        }
        if (side & SIDE_Z) {
            const int nz_ = slabH(k, a.Nz, a.pml[4], a.pml[5]);
            if (nz_ >= 0) {
                if (a.en & 8) q.xz = *ad.xz(nz_);
                if (a.en & 16) q.yz = *ad.yz(nz_);
            }
        }
        return q;
    };
    auto cell = [&](const Ld &c, PsiSetH<T> &q, int k) {
        const T ex = ex_k, ey = ey_k;
        ex_k = c.ex_kp;
        ey_k = c.ey_kp;
        const V3<T> h = h_cell(a, i, j, k, base + k * NxNy, c.hx, c.hy, c.hz, c.ey_kp - ey, c.ex_kp - ex,
                               c.ez_jp - c.ez, c.ez_ip - c.ez, c.ex_jp - ex, c.ey_ip - ey, ikx, iky,
                               uload(a.pf.ikH[2], k), nx_, ny_, side, STORE_NT,
                               PrefPsiH<S, T>{a, i, j, k, q, bx_, cx_, by_, cy_});
        if (sq0 >= 0) {
            const T v[6] = {ex, ey, c.ez, stored(a.Hx, h.x), stored(a.Hy, h.y), stored(a.Hz, h.z)};
            strip_put(a, sq0, j, k, v);
            if (sq1 >= 0) strip_put(a, sq1, j, k, v);
        }
    };
    Ld cur = load(k0);
    PsiSetH<T> q = load_psi(k0);
    for (int k = k0; k < k1 - 1; k++) {
        PsiSetH<T> qn = load_psi(k + 1);
        const Ld nxt = load(k + 1);
        cell(cur, q, k);
        cur = nxt;
        q = qn;
    }
    cell(cur, q, k1 - 1);
}

// 2D H pass (Nz == 1): rows ascending; E(j+1) of row j is E(j) of row j+1.
template <typename S, typename T, int BY, int KY>
__device__ __forceinline__ void h_tile2d(const HArgs<S, T> &a, int w) {
    int bx, by, bz;
    tile_of(w, a.nbx, a.nby, a.nbz, a.remap, bx, by, bz);
    const int i = bx * BX + threadIdx.x;
    const int j0 = (by * BY + threadIdx.y) * KY;
    if (i >= a.Nx || j0 >= a.Ny) return;
    const int j1 = min(j0 + KY, a.Ny);
    const int oip = (i == a.Nx - 1) ? -(a.Nx - 1) : 1;
    const T ikx = a.pf.ikH[0][i];
    const int nx_ = slabH(i, a.Nx, a.pml[0], a.pml[1]);
    const T ikz = a.pf.ikH[2][0];
    const int lm = a.lm;
    const T zero = 0;
    T ex_j = (lm & 1) ? a.Ex[i + a.P * j0] : zero;
    T ez_j = (lm & 4) ? a.Ez[i + a.P * j0] : zero;
    for (int j = j0; j < j1; j++) {
        const int id = i + a.P * j;
        const int idp = (j == a.Ny - 1) ? i : id + a.P;      // (j+1) with wrap
        T ex_jp = zero, ez_jp = zero, ey = zero, ey_ip = zero, ez_ip = zero;
        if (lm & 1) ex_jp = a.Ex[idp];
        if (lm & 2) { ey = a.Ey[id]; ey_ip = a.Ey[id + oip]; }
        if (lm & 4) { ez_jp = a.Ez[idp]; ez_ip = a.Ez[id + oip]; }
        const T hx = (lm & 8) ? a.Hx[id] : zero;
        const T hy = (lm & 16) ? a.Hy[id] : zero;
        const T hz = (lm & 32) ? a.Hz[id] : zero;
        const T iky = a.pf.ikH[1][j];
        const int ny_ = slabH(j, a.Ny, a.pml[2], a.pml[3]);
        h_cell(a, i, j, 0, id, hx, hy, hz, zero, zero, ez_jp - ez_j, ez_ip - ez_j, ex_jp - ex_j, ey_ip - ey, ikx, iky,
               ikz, nx_, ny_, side_all(a), lm, GlobalPsiH<S, T>{a, i, j, 0});
        ex_j = ex_jp;
        ez_j = ez_jp;
    }
}

template <typename S, typename T, int BY, int KY>
__global__ __launch_bounds__(BX *BY) void k_update_H2(const HArgs<S, T> a) {
    h_tile2d<S, T, BY, KY>(a, blockIdx.x);
}

template <typename S, typename T>
__device__ __forceinline__ void h_col_cell(const HArgs<S, T> &a, int k);

// 3D H pass; workgroups from nb on update the column `aux` (see k_update_E).
template <typename S, typename T, int BY, int KZ>
__global__ __launch_bounds__(BX *BY) void k_update_H(const HArgs<S, T> a, const HArgs<S, T> aux, int nb) {
    if ((int)blockIdx.x >= nb) {
        const int k = ((int)blockIdx.x - nb) * (BX * BY) + threadIdx.x + BX * threadIdx.y;
        if (k < aux.Nz) h_col_cell(aux, k);
        return;
    }
    h_tile<S, T, BY, KZ>(a, blockIdx.x);
}

template <typename S, typename T, int BY, int KZ>
__global__ __launch_bounds__(BX *BY) void k_update_H_batch(const HArgs<S, T> *__restrict__ d,
                                                           const int *__restrict__ wg0, int n,
                                                           const int *__restrict__ tcur,
                                                           const int *__restrict__ tend, int s) {
    const int r = batch_member(wg0, n);
    const HArgs<S, T> &a = d[r];
    const int w = blockIdx.x - wg0[r];
    if (tcur[r] + s >= tend[r] || w >= a.nbx * a.nby * a.nbz) return;
    if (a.Nz == 1) h_tile2d<S, T, BY, KZ>(a, w);
    else h_tile<S, T, BY, KZ>(a, w);
}

template <typename S, typename T>
__device__ __forceinline__ void h_col_cell(const HArgs<S, T> &a, int k) {
    const int kp = (k == a.Nz - 1) ? 0 : k + 1;
    const T zero = 0;
    h_cell(a, 0, 0, k, k, a.Hx[k], a.Hy[k], a.Hz[k], a.Ey[kp] - a.Ey[k], a.Ex[kp] - a.Ex[k], zero, zero, zero,
           zero, a.pf.ikH[0][0], a.pf.ikH[1][0], a.pf.ikH[2][k], -1, -1, side_all(a), a.lm, GlobalPsiH<S, T>{a, 0, 0, k});
}

template <typename S, typename T>
__global__ __launch_bounds__(256) void k_update_H_col(const HArgs<S, T> a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < a.Nz) h_col_cell(a, k);
}

// ================================================================ sensors
struct SensDesc {
    int type, x1, y1, z1, span1, span2;
    long P, off;
    int sp;          // planes normal to x: strip plane index (-1: sample the fields)
};

// local_Ex..local_Hz (core/fdtd_utils.cpp:457-507) at plane cell p of a
// sensor, written to stage slot `slot`.
template <typename S, typename T>
__device__ __forceinline__ void sample_point(const SensDesc &d, long p, int slot, long Ptot, const Fld<S, T> &Ex,
                                             const Fld<S, T> &Ey, const Fld<S, T> &Ez, const Fld<S, T> &Hx,
                                             const Fld<S, T> &Hy, const Fld<S, T> &Hz, int Nx, int Ny, int Nz,
                                             int P, T *stage, const T *strip = nullptr) {
    const int a = (int)(p % d.span1), b = (int)(p / d.span1);
    int i, j, k;
    if (d.type == AEFDTD_NORMAL_X || d.type == AEFDTD_NORMAL_XM) { i = d.x1; j = d.y1 + a; k = d.z1 + b; }
    else if (d.type == AEFDTD_NORMAL_Y || d.type == AEFDTD_NORMAL_YM) { i = d.x1 + a; j = d.y1; k = d.z1 + b; }
    else { i = d.x1 + a; j = d.y1 + b; k = d.z1; }
    const int ip = (i + 1 == Nx) ? 0 : i + 1, jp = (j + 1 == Ny) ? 0 : j + 1, kp = (k + 1 == Nz) ? 0 : k + 1;
    auto at = [&](const Fld<S, T> &f, int x, int y, int z) { return f[x + P * (y + Ny * z)]; };
    T v[6];
    if (strip && d.sp >= 0) {
        // the same sums from the strip the H pass wrote (columns i and ip)
        auto st = [&](int col, int c, int y, int z) {
            return strip[(((size_t)d.sp * Nz + z) * Ny + y) * 12 + 6 * col + c];
        };
        v[0] = (st(0, 0, j, k) + st(0, 0, jp, k) + st(0, 0, j, kp) + st(0, 0, jp, kp)) / (T)4.0;
        v[1] = (st(0, 1, j, k) + st(1, 1, j, k) + st(0, 1, j, kp) + st(1, 1, j, kp)) / (T)4.0;
        v[2] = (st(0, 2, j, k) + st(1, 2, j, k) + st(0, 2, jp, k) + st(1, 2, jp, k)) / (T)4.0;
        v[3] = (st(0, 3, j, k) + st(1, 3, j, k)) / (T)2.0;
        v[4] = (st(0, 4, j, k) + st(0, 4, jp, k)) / (T)2.0;
        v[5] = (st(0, 5, j, k) + st(0, 5, j, kp)) / (T)2.0;
#pragma unroll
        for (int c = 0; c < 6; c++) stage[((size_t)slot * 6 + c) * Ptot + d.off + p] = v[c];
        return;
    }
    v[0] = (at(Ex, i, j, k) + at(Ex, i, jp, k) + at(Ex, i, j, kp) + at(Ex, i, jp, kp)) / (T)4.0;
    v[1] = (at(Ey, i, j, k) + at(Ey, ip, j, k) + at(Ey, i, j, kp) + at(Ey, ip, j, kp)) / (T)4.0;
    v[2] = (at(Ez, i, j, k) + at(Ez, ip, j, k) + at(Ez, i, jp, k) + at(Ez, ip, jp, k)) / (T)4.0;
    v[3] = (at(Hx, i, j, k) + at(Hx, ip, j, k)) / (T)2.0;
    v[4] = (at(Hy, i, j, k) + at(Hy, i, jp, k)) / (T)2.0;
    v[5] = (at(Hz, i, j, k) + at(Hz, i, j, kp)) / (T)2.0;
#pragma unroll
    for (int c = 0; c < 6; c++) stage[((size_t)slot * 6 + c) * Ptot + d.off + p] = v[c];
}

template <typename S, typename T>
__global__ __launch_bounds__(256) void k_sample(const SensDesc *sd, int slot, long Ptot, Fld<S, T> Ex, Fld<S, T> Ey,
                                                Fld<S, T> Ez, Fld<S, T> Hx, Fld<S, T> Hy, Fld<S, T> Hz, int Nx, int Ny,
                                                int Nz, int P, T *stage, const T *strip) {
    const SensDesc d = sd[blockIdx.y];
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    sample_point(d, p, slot, Ptot, Ex, Ey, Ez, Hx, Hy, Hz, Nx, Ny, Nz, P, stage, strip);
}

// The sensor samples of step n and the RC pre-pass of step n+1 in one launch:
// both only read the fields the H pass of step n left (the RC pass writes its
// pole states and increments, which nothing else touches), so they share the
// launch instead of two: workgroups [0, 3 nrx) run the RC pre-pass
// (component blockIdx / nrx), the rest sample (sensor (blockIdx - 3 nrx) /
// nsx).  The RC lanes go first: each waits on a chain of dependent loads, and
// dispatched first they wait while the sampling workgroups run.
template <typename S, typename T>
__global__ __launch_bounds__(256) void k_sample_rc(const SensDesc *sd, int slot, long Ptot, Fld<S, T> Ex, Fld<S, T> Ey,
                                                   Fld<S, T> Ez, Fld<S, T> Hx, Fld<S, T> Hy, Fld<S, T> Hz, int Nx,
                                                   int Ny, int Nz, int P, T *stage, const T *strip, int nsx, int ns,
                                                   const RcArgs<S, T> r, int nrx) {
    const int b = blockIdx.x;
    if (b < 3 * nrx) {
        const int s = (b % nrx) * 256 + threadIdx.x;
        if (s < r.ndisp) rc_lane(r, s, b / nrx);
        return;
    }
    const int q = b - 3 * nrx;
    const SensDesc d = sd[q / nsx];
    const long p = (long)(q % nsx) * 256 + threadIdx.x;
    if (p < d.P) sample_point(d, p, slot, Ptot, Ex, Ey, Ez, Hx, Hy, Hz, Nx, Ny, Nz, P, stage, strip);
}

// ---- single-wavelength field sensors (sensors/s_fieldmap.cpp)
// FieldBlock (:89-230): over a box, E and H interpolated to the cell
// (local_E*/local_H*) times e^{i w step Dt} (E) and e^{i w (step+1/2) Dt} (H),
// added to FP64 complex accumulators every step.  FieldMap (:321-470): over a
// plane normal to x, y or z, E (or H with `mag`) times e^{i w step Dt}; with
// `cum` each plane cell first sums the values along the normal over the
// sensor's depth (in FP64).  No tapering: the reference applies none.
struct FieldDesc {
    int block;                  // 1 FieldBlock, 0 FieldMap
    int normal;                 // FieldMap: 0 x, 1 y, 2 z
    int cum, mag;
    int x1, y1, z1;             // first cell (full grid)
    int s1, s2, s3;             // block: spans along x, y, z; map: in-plane spans and the depth
    long n;                     // accumulator cells: s1 s2 s3 (block) or s1 s2 (map)
    double2 *acc;               // [6 or 3][n]
};

// local_Ex..local_Hz (core/fdtd_utils.cpp:457-507) of cell (i, j, k), the
// arithmetic of sample_point.
template <typename S, typename T>
__device__ __forceinline__ void local_fields(const Fld<S, T> &Ex, const Fld<S, T> &Ey, const Fld<S, T> &Ez,
                                             const Fld<S, T> &Hx, const Fld<S, T> &Hy, const Fld<S, T> &Hz, int Nx,
                                             int Ny, int Nz, int P, int i, int j, int k, T v[6]) {
    const int ip = (i + 1 == Nx) ? 0 : i + 1, jp = (j + 1 == Ny) ? 0 : j + 1, kp = (k + 1 == Nz) ? 0 : k + 1;
    auto at = [&](const Fld<S, T> &f, int x, int y, int z) { return f[x + P * (y + Ny * z)]; };
    v[0] = (at(Ex, i, j, k) + at(Ex, i, jp, k) + at(Ex, i, j, kp) + at(Ex, i, jp, kp)) / (T)4.0;
    v[1] = (at(Ey, i, j, k) + at(Ey, ip, j, k) + at(Ey, i, j, kp) + at(Ey, ip, j, kp)) / (T)4.0;
    v[2] = (at(Ez, i, j, k) + at(Ez, ip, j, k) + at(Ez, i, jp, k) + at(Ez, ip, jp, k)) / (T)4.0;
    v[3] = (at(Hx, i, j, k) + at(Hx, ip, j, k)) / (T)2.0;
    v[4] = (at(Hy, i, j, k) + at(Hy, i, jp, k)) / (T)2.0;
    v[5] = (at(Hz, i, j, k) + at(Hz, i, j, kp)) / (T)2.0;
}

// One feed of a field sensor: lane q updates accumulator cell q.
template <typename S, typename T>
__global__ __launch_bounds__(256) void k_field_feed(const FieldDesc d, Fld<S, T> Ex, Fld<S, T> Ey, Fld<S, T> Ez,
                                                    Fld<S, T> Hx, Fld<S, T> Hy, Fld<S, T> Hz, int Nx, int Ny,
                                                    int Nz, int P, double2 cE, double2 cH) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.n) return;
    auto add = [&](int c, double v, double2 w) {
        double2 &A = d.acc[(size_t)c * d.n + q];
        A.x += v * w.x;
        A.y += v * w.y;
    };
    T v[6];
    if (d.block) {
        const int i = (int)(q % d.s1), j = (int)((q / d.s1) % d.s2), k = (int)(q / ((long)d.s1 * d.s2));
        local_fields(Ex, Ey, Ez, Hx, Hy, Hz, Nx, Ny, Nz, P, d.x1 + i, d.y1 + j, d.z1 + k, v);
        for (int c = 0; c < 6; c++) add(c, (double)v[c], c < 3 ? cE : cH);
        return;
    }
    const int a = (int)(q % d.s1), b = (int)(q / d.s1);
    // plane cell (a, b) -> grid offsets: normal x: (., a, b), y: (a, ., b), z: (a, b, .)
    int ox = 0, oy = 0, oz = 0;
    if (d.normal == 0) { oy = a; oz = b; }
    else if (d.normal == 1) { ox = a; oz = b; }
    else { ox = a; oy = b; }
    const int o = d.mag ? 3 : 0;
    if (!d.cum) {
        local_fields(Ex, Ey, Ez, Hx, Hy, Hz, Nx, Ny, Nz, P, d.x1 + ox, d.y1 + oy, d.z1 + oz, v);
        for (int c = 0; c < 3; c++) add(c, (double)v[o + c], cE);
        return;
    }
    double sum[3] = {0, 0, 0};
    for (int t = 0; t < d.s3; t++) {
        const int dx = d.normal == 0 ? t : 0, dy = d.normal == 1 ? t : 0, dz = d.normal == 2 ? t : 0;
        local_fields(Ex, Ey, Ez, Hx, Hy, Hz, Nx, Ny, Nz, P, d.x1 + ox + dx, d.y1 + oy + dy, d.z1 + oz + dz, v);
        for (int c = 0; c < 3; c++) sum[c] += (double)v[o + c];
    }
    for (int c = 0; c < 3; c++) add(c, sum[c], cE);
}

// One sensor of one grid of a batch: the grid's fields, its stage pool and
// the batch member index (for the step window).
template <typename S, typename T>
struct SampB {
    SensDesc d;
    Fld<S, T> F[6];
    int Nx, Ny, Nz, P, member;
    long Ptot;
    T *stage;
};

// Sensor sampling of every grid of a batch at chunk step s (stage slot s:
// members join a batch with an empty stage at a multiple of KS steps).
template <typename S, typename T>
__global__ __launch_bounds__(256) void k_sample_batch(const SampB<S, T> *__restrict__ e, const int *__restrict__ tcur,
                                                      const int *__restrict__ tend, int s) {
    const SampB<S, T> &q = e[blockIdx.y];
    if (tcur[q.member] + s >= tend[q.member]) return;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= q.d.P) return;
    sample_point(q.d, p, s, q.Ptot, q.F[0], q.F[1], q.F[2], q.F[3], q.F[4], q.F[5], q.Nx, q.Ny, q.Nz, q.P, q.stage);
}

// Running DFT over a block of nfill <= KD staged steps (sensors/sensors.cpp:
// 456-479): acc[c][l][p] += sum_s coef[s0+s][E|H][l] * stage[s][c][p].  The
// accumulators are FP64; the block's KD-term sums are formed in the field
// arithmetic type T (FP64 for FP64 runs, so FP64 runs keep the reference's
// double sums; FP32 for FP32/FP16 runs, whose samples are FP32 already).
// One lane per (cell, component) keeps its KD samples in registers; the
// coefficients of DFT_LCH wavelengths at a time are staged in LDS and read as
// broadcasts, so each accumulator is read and written once per KD steps.  The
// next wavelength's accumulator is requested before the current one's sums
// are formed, so its latency hides behind the 2 KD multiply-adds.
typedef double dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_nt(const double2 *p) {
    const dvec2 v = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(double2 v, double2 *p) {
    const dvec2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<dvec2 *>(p));
}

template <typename T, int KD>
__device__ __forceinline__ void dft_block(const T *stage, long Ptot, long off, long P, int Nl, const double2 *coef,
                                          int s0, int nfill, double2 *acc, int c) {
    using T2 = std::conditional_t<std::is_same_v<T, double>, double2, float2>;
    __shared__ T2 W[KD * DFT_LCH];
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = c >= 3 ? 1 : 0;
    const bool live = p < P;
    T x[KD];
#pragma unroll
    for (int s = 0; s < KD; s++)
        x[s] = (live && s < nfill) ? __builtin_nontemporal_load(stage + ((size_t)s * 6 + c) * Ptot + off + p) : (T)0;
    double2 *a = acc + (size_t)c * Nl * P + p;
    for (int l0 = 0; l0 < Nl; l0 += DFT_LCH) {
        const int nl = min(DFT_LCH, Nl - l0);
        __syncthreads();
        for (int q = threadIdx.x; q < KD * DFT_LCH; q += blockDim.x) {
            const int s = q / DFT_LCH, l = q % DFT_LCH;
            const double2 w = (s < nfill && l < nl) ? coef[((size_t)(s0 + s) * 2 + h) * Nl + l0 + l] : make_double2(0.0, 0.0);
            W[q] = T2{(T)w.x, (T)w.y};
        }
        __syncthreads();
        if (live) {
            double2 A = ld_nt(a + (size_t)l0 * P);
            for (int l = 0; l < nl; l++) {
                double2 An = A;
                if (l + 1 < nl) An = ld_nt(a + (size_t)(l0 + l + 1) * P);
                T re = 0, im = 0;
#pragma unroll
                for (int s = 0; s < KD; s++) {
                    const T2 w = W[s * DFT_LCH + l];
                    re += w.x * x[s];
                    im += w.y * x[s];
                }
                A.x += (double)re;
                A.y += (double)im;
                st_nt(A, a + (size_t)(l0 + l) * P);
                A = An;
            }
        }
    }
}

// Sensors of one grid flushed by one launch.
constexpr int DFT_MAX_SENS = 16;

// The DFT flush of FP32-arithmetic grids on the matrix cores.  For one
// component c of a sensor plane, the block's sums are a product
//   D[l'][p] = sum_s W[s][l'] X[s][p],   l' = 2 l + (0 re, 1 im),
// of the coefficient table W (steps x 2 Nl, in FP32) and the staged samples X
// (steps x cells).  The product runs on v_mfma_f32_32x32x16_bf16 with both
// operands split into two bf16 parts, w = wh + wl and x = xh + xl (wh, xh
// rounded to bf16, wl, xl the rounded remainders), as
//   w x ~ wh xh + wh xl + wl xh,
// three bf16 products whose dropped term wl xl and the rounding of the
// remainders leave about 2^-17 of each term (about FP32's product error over a
// block of steps); at 16 times the rate of the FP32 matrix op, the three cost
// a fifth of one FP32 product, which would otherwise take as long as the
// flush's HBM traffic.  A = W^T (32 rows l'), B = X (32 cells), 16 steps per
// product: lane L supplies steps 8 (L / 32) + e, e = 0..7, of row or cell
// L % 32 in both operands.  A wavefront owns 32 cells of one component and up
// to MXB blocks of 32 l' (D in 16 registers per block); the four wavefronts
// of a workgroup (128 cells) share each batch's coefficients, split once into
// LDS.  The sums are then added to the FP64 accumulators, lane = cell, so that
// each accumulator row is read and written as 32 consecutive double2.  Output
// layout of a 32x32 block: lane L holds column L % 32 and rows
// 8 (r / 4) + 4 (L / 32) + r % 4 in its register r.
constexpr int MXB = 4;                     // l' blocks per pass of a wavefront
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct DftDescMx {
    long off, P;
    int Nl, LP;                            // LP: 2 Nl rounded up to 32
    int rows;                              // steps in W
    const float *W;                        // [rows][E|H][LP], zero past 2 Nl
    double2 *acc;                          // [6][Nl][P]
};
struct DftSetMx {
    DftDescMx d[DFT_MAX_SENS];
};

__device__ __forceinline__ void bf16_split(const f32x8 &v, bf16x8 &hi, bf16x8 &lo) {
    hi = __builtin_convertvector(v, bf16x8);
    lo = __builtin_convertvector(v - __builtin_convertvector(hi, f32x8), bf16x8);
}

// Steps go in batches of MX_K, one product each.  A batch's coefficients are
// loaded one batch ahead (thread t: row l' = 32 b0 + t % 128, steps
// 8 (t / 128) + e) and split into LDS at the end of the batch before; each
// lane requests its samples two batches ahead, after the coefficients, so that
// waiting for the coefficients never waits for those samples.  Steps past
// nfill read the block's last staged row and a zero coefficient.
constexpr int MX_K = 16;

__global__ __launch_bounds__(256, 2) void k_dft_mx(const float *stage, long Ptot, const DftSetMx set, int s0, int nfill) {
    __shared__ bf16x8 Lh[2][MXB * 64], Ll[2][MXB * 64];
    const DftDescMx &d = set.d[blockIdx.y / 6];
    if ((long)blockIdx.x * 128 >= d.P) return;              // the whole workgroup: no cell of this sensor
    const int c = blockIdx.y % 6, h = c >= 3 ? 1 : 0;
    const int tid = threadIdx.x, lane = tid & 63, half = lane >> 5;
    const long p = ((long)blockIdx.x * 4 + (tid >> 6)) * 32 + (lane & 31);
    const bool live = p < d.P;                              // lanes past the sensor compute and discard
    const float *xs = stage + (size_t)c * Ptot + d.off + (live ? p : 0);
    const int nb = (2 * d.Nl + 31) / 32, nbatch = (nfill + MX_K - 1) / MX_K, last = max(nfill - 1, 0);
    const int wr = tid & 127, wh = tid >> 7, wslot = (wr >> 5) * 64 + wh * 32 + (wr & 31);
    for (int b0 = 0; b0 < nb; b0 += MXB) {
        const float *wcol = d.W + (size_t)h * d.LP + min(b0 + (wr >> 5), nb - 1) * 32 + (wr & 31);
        auto loadw = [&](int b, f32x8 &w) {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int rel = b * MX_K + 8 * wh + e;
                w[e] = wcol[(size_t)min(s0 + rel, d.rows - 1) * 2 * d.LP] * (rel < nfill ? 1.0f : 0.0f);
            }
        };
        auto storew = [&](int buf, const f32x8 &w) {
            bf16x8 hi, lo;
            bf16_split(w, hi, lo);
            Lh[buf][wslot] = hi;
            Ll[buf][wslot] = lo;
        };
        auto loadx = [&](int b, f32x8 &x) {
#pragma unroll
            for (int e = 0; e < 8; e++)
                x[e] = __builtin_nontemporal_load(xs + (size_t)min(b * MX_K + 8 * half + e, last) * 6 * Ptot);
        };
        f32x16 D[MXB];
#pragma unroll
        for (int q = 0; q < MXB; q++)
            for (int r = 0; r < 16; r++) D[q][r] = 0.0f;
        auto batch = [&](int b, const f32x8 &x, f32x8 &xnext, f32x8 &w) {
            loadw(b + 1, w);
            loadx(b + 2, xnext);
            bf16x8 xh, xl, ah[MXB], al[MXB];
            bf16_split(x, xh, xl);
#pragma unroll
            for (int q = 0; q < MXB; q++) {
                ah[q] = Lh[b & 1][q * 64 + lane];
                al[q] = Ll[b & 1][q * 64 + lane];
            }
#pragma unroll
            for (int q = 0; q < MXB; q++) D[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[q], xh, D[q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < MXB; q++) D[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[q], xl, D[q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < MXB; q++) D[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[q], xh, D[q], 0, 0, 0);
            storew((b + 1) & 1, w);
            __syncthreads();
        };
        f32x8 x0, x1, x2, w;
        loadw(0, w);
        loadx(0, x0);
        loadx(1, x1);
        storew(0, w);
        __syncthreads();
        for (int b = 0; b < nbatch; b += 3) {                  // three sample buffers in rotation
            batch(b, x0, x2, w);
            if (b + 1 < nbatch) batch(b + 1, x1, x0, w);
            if (b + 2 < nbatch) batch(b + 2, x2, x1, w);
        }
        if (!live) continue;
        if ((b0 + MXB) * 32 <= 2 * d.Nl) {
            // every row of the MXB blocks exists: the accumulators of two
            // blocks requested before the first of them is written
#pragma unroll
            for (int q0 = 0; q0 < MXB; q0 += 2) {
                double2 v[2][8];
#pragma unroll
                for (int q = 0; q < 2; q++)
#pragma unroll
                    for (int r = 0; r < 16; r += 2) {
                        const int l = ((b0 + q0 + q) * 32 + 8 * (r >> 2) + 4 * half + (r & 3)) >> 1;
                        v[q][r / 2] = ld_nt(d.acc + ((size_t)c * d.Nl + l) * d.P + p);
                    }
#pragma unroll
                for (int q = 0; q < 2; q++)
#pragma unroll
                    for (int r = 0; r < 16; r += 2) {
                        const int l = ((b0 + q0 + q) * 32 + 8 * (r >> 2) + 4 * half + (r & 3)) >> 1;
                        v[q][r / 2].x += (double)D[q0 + q][r];
                        v[q][r / 2].y += (double)D[q0 + q][r + 1];
                        st_nt(v[q][r / 2], d.acc + ((size_t)c * d.Nl + l) * d.P + p);
                    }
            }
            continue;
        }
        // the eight accumulators of a block are all requested before the
        // first is written (rows l >= Nl read row Nl - 1 and write nothing)
#pragma unroll
        for (int q = 0; q < MXB; q++) {
            if (b0 + q >= nb) break;
            double2 v[8];
#pragma unroll
            for (int r = 0; r < 16; r += 2) {
                const int l = ((b0 + q) * 32 + 8 * (r >> 2) + 4 * half + (r & 3)) >> 1;
                v[r / 2] = ld_nt(d.acc + ((size_t)c * d.Nl + min(l, d.Nl - 1)) * d.P + p);
            }
#pragma unroll
            for (int r = 0; r < 16; r += 2) {
                const int l = ((b0 + q) * 32 + 8 * (r >> 2) + 4 * half + (r & 3)) >> 1;
                if (l < d.Nl) {
                    v[r / 2].x += (double)D[q][r];
                    v[r / 2].y += (double)D[q][r + 1];
                    st_nt(v[r / 2], d.acc + ((size_t)c * d.Nl + l) * d.P + p);
                }
            }
        }
    }
}

// The DFT flush of FP64 grids on the matrix cores: the same product
//   D[l'][p] = sum_s W[s][l'] X[s][p]
// in FP64 with v_mfma_f64_16x16x4f64: A = W^T (16 rows l'), B = X (16 cells),
// 4 steps per product, lane L supplying step L / 16 of row or cell L % 16 in
// both operands.  The matrix rate of FP64 equals its vector rate on this part;
// what the matrix form buys is room: the block sums of 16 cells x 128 l' are
// spread over the wavefront (4 registers per block), so a flush can take 256
// staged steps instead of the 64 samples a lane could hold, and the FP64
// accumulators are read and written once per 256 steps.  A workgroup is 8
// wavefronts (128 cells of one component) sharing each 16-step batch of
// coefficients through LDS (thread t: l' = 16 b0 + t % 128, steps
// 4 (t / 128) + e), one batch ahead, with the samples two batches ahead.
// Output layout of a 16x16 block: lane L holds column L % 16 and rows
// 4 r + L / 16 in its register r.  Row i of A takes l' = 4 (i % 4) + i / 4,
// so that lane L's four registers hold l' = 4 (L / 16) + r: the (re, im)
// pairs of wavelengths 2 (L / 16) and 2 (L / 16) + 1 of the block.  Steps
// past nfill read the block's last staged row and a zero coefficient.
constexpr int M64B = 8;                    // l' blocks of 16 per pass (128 l')
constexpr int M64_K = 16;                  // steps per batch: four products
typedef double f64x4 __attribute__((ext_vector_type(4)));

struct DftDescMx64 {
    long off, P;
    int Nl, LP;                            // LP: 2 Nl rounded up to 32
    int rows;                              // steps in W
    const double *W;                       // [rows][E|H][LP], zero past 2 Nl
    double2 *acc;                          // [6][Nl][P]
};
struct DftSetMx64 {
    DftDescMx64 d[DFT_MAX_SENS];
};

__global__ __launch_bounds__(512) void k_dft_mx64(const double *stage, long Ptot, const DftSetMx64 set, int s0,
                                                  int nfill) {
    __shared__ double Ws[2][M64_K][M64B * 16];
    const DftDescMx64 &d = set.d[blockIdx.y / 6];
    if ((long)blockIdx.x * 128 >= d.P) return;              // the whole workgroup: no cell of this sensor
    const int c = blockIdx.y % 6, h = c >= 3 ? 1 : 0;
    const int tid = threadIdx.x, lane = tid & 63, g4 = lane >> 4, col = lane & 15;
    const int arow = 4 * (col & 3) + (col >> 2);            // l' of the A row this lane supplies
    const long p = ((long)blockIdx.x * 8 + (tid >> 6)) * 16 + col;
    const bool live = p < d.P;                              // lanes past the sensor compute and discard
    const double *xs = stage + (size_t)c * Ptot + d.off + (live ? p : 0);
    const int nb = (2 * d.Nl + 15) / 16, nbatch = (nfill + M64_K - 1) / M64_K, last = max(nfill - 1, 0);
    const int wr = tid & 127, wh = tid >> 7;
    for (int b0 = 0; b0 < nb; b0 += M64B) {
        const double *wcol = d.W + (size_t)h * d.LP + min(b0 * 16 + wr, d.LP - 1);
        auto loadw = [&](int b, double (&w)[4]) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int rel = b * M64_K + 4 * wh + e;
                w[e] = wcol[(size_t)min(s0 + rel, d.rows - 1) * 2 * d.LP] * (rel < nfill ? 1.0 : 0.0);
            }
        };
        auto storew = [&](int buf, const double (&w)[4]) {
#pragma unroll
            for (int e = 0; e < 4; e++) Ws[buf][4 * wh + e][wr] = w[e];
        };
        auto loadx = [&](int b, double (&x)[4]) {
#pragma unroll
            for (int u = 0; u < 4; u++)
                x[u] = __builtin_nontemporal_load(xs + (size_t)min(b * M64_K + 4 * u + g4, last) * 6 * Ptot);
        };
        f64x4 D[M64B];
#pragma unroll
        for (int q = 0; q < M64B; q++)
            for (int r = 0; r < 4; r++) D[q][r] = 0.0;
        auto batch = [&](int b, const double (&x)[4], double (&xnext)[4], double (&w)[4]) {
            loadw(b + 1, w);
            loadx(b + 2, xnext);
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int q = 0; q < M64B; q++)
                    D[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ws[b & 1][4 * u + g4][q * 16 + arow], x[u], D[q],
                                                               0, 0, 0);
            storew((b + 1) & 1, w);
            __syncthreads();
        };
        double x0[4], x1[4], x2[4], w[4];
        loadw(0, w);
        loadx(0, x0);
        loadx(1, x1);
        storew(0, w);
        __syncthreads();
        for (int b = 0; b < nbatch; b += 3) {                  // three sample buffers in rotation
            batch(b, x0, x2, w);
            if (b + 1 < nbatch) batch(b + 1, x1, x0, w);
            if (b + 2 < nbatch) batch(b + 2, x2, x1, w);
        }
        if (!live) continue;
        // register r of block q: l' = 16 (b0 + q) + 4 g4 + r, i.e. l = 8 (b0 + q) + 2 g4 + r / 2,
        // (re, im) = (r even, r odd);
        // the accumulators of four blocks requested before the first is written
#pragma unroll
        for (int q0 = 0; q0 < M64B; q0 += 4) {
            double2 v[4][2];
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const int l = 8 * (b0 + q0 + q) + 2 * g4 + t;
                    v[q][t] = ld_nt(d.acc + ((size_t)c * d.Nl + min(l, d.Nl - 1)) * d.P + p);
                }
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const int l = 8 * (b0 + q0 + q) + 2 * g4 + t;
                    if (l < d.Nl) {
                        v[q][t].x += D[q0 + q][2 * t];
                        v[q][t].y += D[q0 + q][2 * t + 1];
                        st_nt(v[q][t], d.acc + ((size_t)c * d.Nl + l) * d.P + p);
                    }
                }
        }
    }
}

// One sensor of one grid of a batch for the DFT flush at the end of a chunk.
template <typename T>
struct DftB {
    const T *stage;
    long Ptot, off, P;
    int Nl, member;
    const double2 *coef;
    double2 *acc;
};

// DFT flush of every sensor of a batch after a KS-step chunk that started at
// tcur (each member staged min(KS, tend - tcur) steps).
template <typename T>
__global__ __launch_bounds__(256) void k_dft_batch(const DftB<T> *__restrict__ e, const int *__restrict__ tcur,
                                                   const int *__restrict__ tend) {
    const DftB<T> &q = e[blockIdx.z];
    const int t0 = tcur[q.member];
    const int nfill = min(KS, tend[q.member] - t0);
    if (nfill <= 0 || (long)blockIdx.x * blockDim.x >= q.P) return;
    dft_block<T, KS>(q.stage, q.Ptot, q.off, q.P, q.Nl, q.coef, t0, nfill, q.acc, blockIdx.y);
}

// Plane injection of batch members (the TFSF source column of a
// single-particle run): before the passes of step s of a chunk, member r's E
// and H arguments take the injected values of its step tcur + s from its
// table ([step][dEx, dEy, dHx, dHy], as aefdtd_set_plane_injection gave them),
// or no injection past the table's end (as e_args / h_args do for a grid
// stepped on its own).
struct InjB {
    int member, k, n;
    const double *tab;
};

template <typename S, typename T>
__global__ void k_batch_inj(EArgs<S, T> *ea, HArgs<S, T> *ha, const InjB *__restrict__ q, int nq,
                            const int *__restrict__ tcur, const int *__restrict__ tend, int s) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nq) return;
    const InjB &b = q[e];
    const int t = tcur[b.member] + s;
    if (t >= tend[b.member]) return;
    EArgs<S, T> &a = ea[b.member];
    HArgs<S, T> &h = ha[b.member];
    if (t < b.n) {
        a.inj_k = b.k;
        h.inj_k = b.k;
        a.injx = (T)b.tab[4 * (size_t)t];
        a.injy = (T)b.tab[4 * (size_t)t + 1];
        h.injx = (T)b.tab[4 * (size_t)t + 2];
        h.injy = (T)b.tab[4 * (size_t)t + 3];
    } else {
        a.inj_k = -1;
        h.inj_k = -1;
    }
}

// RC pre-pass of the dispersive batch members at step s of a chunk: entry r
// of the table (member mem[r]) owns workgroups [wg0[r], wg0[r+1]), one lane per
// dispersive slot, blockIdx.y the component.  skip[r]: a step whose pre-pass
// the member's last sensor feed on its own already ran (-1: none).
template <typename S, typename T>
__global__ __launch_bounds__(256) void k_rc_batch(const RcArgs<S, T> *__restrict__ rc, const int *__restrict__ wg0,
                                                  int n, const int *__restrict__ mem, const int *__restrict__ skip,
                                                  const int *__restrict__ tcur, const int *__restrict__ tend, int s) {
    const int r = batch_member(wg0, n);
    const int t = tcur[mem[r]] + s;
    if (t >= tend[mem[r]] || t == skip[r]) return;
    const int slot = (blockIdx.x - wg0[r]) * blockDim.x + threadIdx.x;
    if (slot < rc[r].ndisp) rc_lane(rc[r], slot, blockIdx.y);
}

// End of a batch chunk: every member advances KS steps, stopping at its end.
__global__ void k_batch_advance(int *tcur, const int *tend, int n) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) tcur[r] = min(tcur[r] + KS, tend[r]);
}

// Spect_Poynting::treat (sensors/s_spec_power.cpp:61-111): one workgroup per
// wavelength, fixed-order tree reduction so the result is deterministic.
__global__ __launch_bounds__(256) void k_poynting(const double2 *acc, long P, int Nl, double f, int type, double *rsp) {
    const int l = blockIdx.x;
    __shared__ double sh[3][256];
    double S[3] = {0, 0, 0};
    for (long p = threadIdx.x; p < P; p += blockDim.x) {
        double2 e[3], h[3];
        for (int c = 0; c < 3; c++) {
            e[c] = acc[((size_t)c * Nl + l) * P + p];
            h[c] = acc[((size_t)(3 + c) * Nl + l) * P + p];
        }
        // Re(a conj(b)) = a.x b.x + a.y b.y
        S[0] += (e[1].x * h[2].x + e[1].y * h[2].y) - (e[2].x * h[1].x + e[2].y * h[1].y);
        S[1] += (e[2].x * h[0].x + e[2].y * h[0].y) - (e[0].x * h[2].x + e[0].y * h[2].y);
        S[2] += (e[0].x * h[1].x + e[0].y * h[1].y) - (e[1].x * h[0].x + e[1].y * h[0].y);
    }
    for (int c = 0; c < 3; c++) sh[c][threadIdx.x] = S[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double Sx = sh[0][0] * f, Sy = sh[1][0] * f, Sz = sh[2][0] * f;
        double r;
        switch (type) {
            case AEFDTD_NORMAL_X: r = Sx; break;
            case AEFDTD_NORMAL_Y: r = Sy; break;
            case AEFDTD_NORMAL_Z: r = Sz; break;
            case AEFDTD_NORMAL_XM: r = -Sx; break;
            case AEFDTD_NORMAL_YM: r = -Sy; break;
            default: r = -Sz; break;
        }
        rsp[l] = r;
    }
}

// Analytic pulse on the device (core/chpin.cpp:44-90).
__device__ inline double d_pulse(const aefdtd_pulse &p, double amp, double x, double y, double z, double t) {
    const double ph = p.kx * x + p.ky * y + p.kz * z;
    const double a = ph - p.omega * t;
    const double b = (1.0 / p.k * ph - kC * t / p.n_ind + p.C1) / p.dz;
    return amp * exp(-b * b) * cos(a);
}

// R/T plane sums of modes/normal_incidence.cpp:384-415 for step t.  Each of
// the PS_BLOCKS workgroups reduces a fixed strided subset of the plane in a
// fixed order and writes its partial; the partials are summed in block order
// when they are read back, so the result does not depend on scheduling.
constexpr int PS_BLOCKS = 64;

// One plane sum of a launch: plane kpl, incident pulse (if has_inc), whether
// the incident field is subtracted, its z, the tapering table and the output.
struct PlaneSumArgs {
    int kpl;
    aefdtd_pulse inc;
    int has_inc, subtract;
    double zpos;
    const double *tap;
    double *out;
};

// The plane sums of a grid for step t, blockIdx.y selecting the sum (the R
// and T sums of a normal-incidence run share one launch).
template <typename S, typename T>
__global__ __launch_bounds__(256) void k_plane_sum(Fld<S, T> Ex, Fld<S, T> Ey, Fld<S, T> Ez, int Nx, int Ny, int P,
                                                   PlaneSumArgs s0, PlaneSumArgs s1, double Dx, double Dy, double Dz,
                                                   int t, double Dt) {
    const PlaneSumArgs &ps = blockIdx.y == 0 ? s0 : s1;
    __shared__ double sh[6][256];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const double tb = (t + 0.5) * Dt;
    const double tp = ps.tap[t];
    const long NxNy = (long)Nx * Ny;
    const long stride = (long)blockDim.x * gridDim.x;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < NxNy; q += stride) {
        const int i = (int)(q % Nx), j = (int)(q / Nx);
        const double x = i * Dx, y = j * Dy;
        double in[3] = {0, 0, 0};
        if (ps.has_inc) {
            in[0] = d_pulse(ps.inc, ps.inc.Ex0, x + Dx / 2.0, y, ps.zpos, tb);
            in[1] = d_pulse(ps.inc, ps.inc.Ey0, x, y + Dy / 2.0, ps.zpos, tb);
            in[2] = d_pulse(ps.inc, ps.inc.Ez0, x, y, ps.zpos + Dz / 2.0, tb);
        }
        const int id = i + P * (j + Ny * ps.kpl);
        const double f[3] = {(double)Ex[id], (double)Ey[id], (double)Ez[id]};
        for (int c = 0; c < 3; c++) {
            acc[3 + c] += in[c] * tp;
            acc[c] += (ps.subtract ? (f[c] - in[c]) : f[c]) * tp;
        }
    }
    for (int c = 0; c < 6; c++) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 6; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) ps.out[((size_t)t * gridDim.x + blockIdx.x) * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

// ================================================================ 2D time skewing
// Temporally blocked passes for 2D TE/TM grids of a batch: a workgroup loads
// a tile of TBX x LY cells (three live field components into LDS, material
// ids and CPML psi of its own cells into registers), advances it TS full
// steps (E update, oscillators, H update, sensor samples) and writes back the
// inner (TBX-2TS-1) x (LY-2TS-1) cells, whose dependency cones stay inside the
// tile (overlapped tiling: the halo is recomputed by the neighbours).  Loads
// wrap around the grid like the stencil does, so boundary tiles see exactly
// the neighbours the per-step passes see.  Each cell's update is e_cell_vals /
// h_cell_vals, the same arithmetic as the per-step passes.
constexpr int TBX = 64;           // tile width: one column per lane
constexpr int TB_WAVES = 8;       // waves per workgroup

// CPML psi of one 2D cell in registers: TM (Ez,Hx,Hy) uses Ezx, Ezy, Hxy, Hyx;
// TE (Ex,Ey,Hz) uses Exy, Eyx, Hzx, Hzy; z-slab terms never occur in 2D.
template <typename T>
struct RegPsiE {
    T &a, &b;
    __device__ __forceinline__ T &Exy(int) const { return a; }
    __device__ __forceinline__ T &Exz(int) const { return a; }
    __device__ __forceinline__ T &Eyx(int) const { return b; }
    __device__ __forceinline__ T &Eyz(int) const { return b; }
    __device__ __forceinline__ T &Ezx(int) const { return a; }
    __device__ __forceinline__ T &Ezy(int) const { return b; }
};
template <typename T>
struct RegPsiH {
    T &c, &d;
    __device__ __forceinline__ T &Hxy(int) const { return c; }
    __device__ __forceinline__ T &Hxz(int) const { return c; }
    __device__ __forceinline__ T &Hyx(int) const { return d; }
    __device__ __forceinline__ T &Hyz(int) const { return d; }
    __device__ __forceinline__ T &Hzx(int) const { return c; }
    __device__ __forceinline__ T &Hzy(int) const { return d; }
};

// The cell updates of the time-skewed passes: e_cell_vals / h_cell_vals
// restricted to what a time-skewed grid has (2D, TM or TE, CPML on x/y, no
// TFSF, RC or plane injection, oscillators), the same operations in the same
// order, with every per-grid constant held in registers for the whole launch.
template <typename T>
struct TbCtx {
    bool px, py;
    int Nx, Ny, nosc;
    const OscDev *osc;
    const T *bEx, *cEx, *bEy, *cEy, *bHx, *cHx, *bHy, *cHy;
    T ikzE, ikzH, dtm, dtdmx, dtdmy, dtdmz;
};

// Per-cell update coefficients held in registers (FDTD_Material C1, C2x, C2y, C4).
template <typename T>
struct TbCoef {
    T C1, C2x, C2y, C4;
};

// CPML profile values of a cell's column (x) and row (y), E and H positions.
template <typename T>
struct TbProf {
    T bE, cE, bH, cH;
};

template <int POL, bool PML, typename T>
__device__ __forceinline__ void tb_e_cell(const TbCtx<T> &x, const TbCoef<T> &M, int i, int j, int t, T &e0, T &e1,
                                          T dHa, T dHb, T ikx, T iky, int nx_, int ny_, const TbProf<T> &PX,
                                          const TbProf<T> &PY, T &pa, T &pb, int om) {
    // terms along the one-cell z axis are products with an exact zero and are left out
    if (POL == 0) {        // Ez from dHy_x = dHa, dHx_y = dHb
        T ez = M.C1 * e0 + M.C2x * ikx * dHa - M.C2y * iky * dHb;
        if (PML && x.px) {
            if (nx_ >= 0) {
                pa = PX.bE * pa + PX.cE * dHa;
                ez += M.C4 * pa;
            }
            if (i == 0) ez = 0;
        }
        if (PML && x.py) {
            if (ny_ >= 0) {
                pb = PY.bE * pb + PY.cE * dHb;
                ez -= M.C4 * pb;
            }
            if (j == 0) ez = 0;
        }
        // om: the oscillators whose box holds this cell, in registration order
        for (int q = 0; om; q++, om >>= 1)
            if (om & 1) ez += (T)x.osc[q].val[t - x.osc[q].step0];
        e0 = ez;
    } else {               // Ex, Ey from dHz_y = dHa, dHz_x = dHb
        T ex = M.C1 * e0 + M.C2y * iky * dHa;
        T ey = M.C1 * e1 - M.C2x * ikx * dHb;
        if (PML && x.py) {
            if (ny_ >= 0) {
                pa = PY.bE * pa + PY.cE * dHa;
                ex += M.C4 * pa;
            }
            if (j == 0) ex = 0;
        }
        if (PML && x.px) {
            if (nx_ >= 0) {
                pb = PX.bE * pb + PX.cE * dHb;
                ey -= M.C4 * pb;
            }
            if (i == 0) ey = 0;
        }
        for (int q = 0; om; q++, om >>= 1)
            if (om & 1) {
                const T v = (T)x.osc[q].val[t - x.osc[q].step0];
                if (x.osc[q].comp == 0) ex += v;
                else ey += v;
            }
        e0 = ex;
        e1 = ey;
    }
}

template <int POL, bool PML, typename T>
__device__ __forceinline__ void tb_h_cell(const TbCtx<T> &x, int i, int j, T &h0, T &h1, T dEa, T dEb, T ikx, T iky,
                                          int nx_, int ny_, const TbProf<T> &PX, const TbProf<T> &PY, T &pc, T &pd) {
    if (POL == 0) {        // Hx, Hy from dEz_y = dEa, dEz_x = dEb
        T hx = h0, hy = h1;
        hx += -(x.dtdmy * iky * dEa);
        hy += x.dtdmx * ikx * dEb;
        if (PML && x.py) {
            if (ny_ >= 0) {
                pc = PY.bH * pc + PY.cH * dEa;
                hx -= x.dtm * pc;
            }
            if (j == x.Ny - 1) hx = 0;
        }
        if (PML && x.px) {
            if (nx_ >= 0) {
                pd = PX.bH * pd + PX.cH * dEb;
                hy += x.dtm * pd;
            }
            if (i == x.Nx - 1) hy = 0;
        }
        h0 = hx;
        h1 = hy;
    } else {               // Hz from dEx_y = dEa, dEy_x = dEb
        T hz = h0;
        hz += x.dtdmy * iky * dEa - x.dtdmx * ikx * dEb;
        if (PML && x.px) {
            if (nx_ >= 0) {
                pc = PX.bH * pc + PX.cH * dEb;
                hz -= x.dtm * pc;
            }
            if (i == x.Nx - 1) hz = 0;
        }
        if (PML && x.py) {
            if (ny_ >= 0) {
                pd = PY.bH * pd + PY.cH * dEa;
                hz += x.dtm * pd;
            }
            if (j == x.Ny - 1) hz = 0;
        }
        h0 = hz;
    }
}

// Second copy of a time-skewed grid's live state: the three live field
// components (TM: Ez Hx Hy, TE: Ex Ey Hz) and its four CPML psi arrays (TM:
// Ezx Ezy Hxy Hyx, TE: Exy Eyx Hzx Hzy).  Launches alternate between reading
// the grid's own arrays and writing these and the reverse, so no workgroup
// ever loads a halo cell that another one has already advanced.
template <typename S, typename T>
struct TbAlt {
    S *f[3];
    T *p[4];
};

// Tiles of a time-skewed grid fall in two classes, launched as two kernels:
// interior tiles (rect = [bx0, bx1) x [by0, by1) of the tile grid), none of
// whose cells, halo included, lies in a CPML slab or on a PEC wall, and the
// frame of boundary tiles around them.  The interior kernel carries no CPML
// state at all, which roughly halves its registers (two workgroups per CU
// instead of one) for the bulk of the tiles.
__device__ __forceinline__ void tb_tile(bool pml, int4 rc, int ntx, int w, int &bx, int &by) {
    const int nix = rc.y - rc.x;
    if (!pml) {
        bx = rc.x + w % nix;
        by = rc.z + w / nix;
        return;
    }
    const int per = ntx - nix;                   // boundary tiles per row of the rect's rows
    if (w < rc.z * ntx) {
        by = w / ntx;
        bx = w % ntx;
        return;
    }
    w -= rc.z * ntx;
    if (w < (rc.w - rc.z) * per) {
        by = rc.z + w / per;
        const int c = w % per;
        bx = c < rc.x ? c : rc.y + (c - rc.x);
        return;
    }
    w -= (rc.w - rc.z) * per;
    by = rc.w + w / ntx;
    bx = w % ntx;
}

// POL 0 = TM (LDS planes: Ez, Hx, Hy), 1 = TE (Ex, Ey, Hz).  PML: boundary tiles.
template <typename S, typename T, int POL, int LY, int TS, bool PML>
__global__ __launch_bounds__(64 * TB_WAVES) void k_tb2d_batch(const EArgs<S, T> *__restrict__ ead,
                                                             const HArgs<S, T> *__restrict__ had,
                                                             const int *__restrict__ members,
                                                             const int *__restrict__ wg0, int nm,
                                                             const int *__restrict__ tcur,
                                                             const int *__restrict__ tend,
                                                             const SampB<S, T> *__restrict__ samp,
                                                             const int *__restrict__ sbeg,
                                                             const TbAlt<S, T> *__restrict__ altd,
                                                             const int4 *__restrict__ rect, int s0, int odd) {
    constexpr int CX = TBX / 64, RY = LY / TB_WAVES;
    constexpr int IX = TBX - 2 * TS - 1, IY = LY - 2 * TS - 1;
    extern __shared__ char lds_raw[];
    const int q = batch_member(wg0, nm);
    const int r = members[q];
    const EArgs<S, T> &ea = ead[r];
    const HArgs<S, T> &ha = had[r];
    const int tbeg = tcur[r] + s0;
    // a grid at its end still passes its state to the other buffer set
    const int nsteps = max(0, min(TS, tend[r] - tbeg));
    const int w = blockIdx.x - wg0[q];
    const int Nx = ea.Nx, Ny = ea.Ny;
    const int ntx = (Nx + IX - 1) / IX, nty = (Ny + IY - 1) / IY;
    const int4 rc = rect[r];
    const int nin = (rc.y - rc.x) * (rc.w - rc.z);
    if (w >= (PML ? ntx * nty - nin : nin)) return;
    T *F = reinterpret_cast<T *>(lds_raw);
    T *F0 = F, *F1 = F + TBX * LY, *F2 = F + 2 * TBX * LY;      // TM: Ez Hx Hy; TE: Ex Ey Hz
    int bx, by;
    tb_tile(PML, rc, ntx, w, bx, by);
    const int ox = bx * IX, oy = by * IY;           // inner origin (global)
    const int tx = threadIdx.x, wv = threadIdx.y;
    const int wxs = ea.pml[0] + ea.pml[1];
    const Fld<S, T> &G0 = POL == 0 ? ea.Ez : ea.Ex, &G1 = POL == 0 ? ea.Hx : ea.Ey, &G2 = POL == 0 ? ea.Hy : ea.Hz;
    // source and destination of this launch (the grid's own arrays or the second set)
    const TbAlt<S, T> &alt = altd[r];
    T *const own_p[4] = {POL == 0 ? ea.pEzx : ea.pExy, POL == 0 ? ea.pEzy : ea.pEyx,
                         POL == 0 ? ha.pHxy : ha.pHzx, POL == 0 ? ha.pHyx : ha.pHzy};
    Fld<S, T> src[3] = {G0, G1, G2}, dst[3] = {G0, G1, G2};
    T *psrc[4], *pdst[4];
#pragma unroll
    for (int c = 0; c < 3; c++) (odd ? src[c] : dst[c]).p = alt.f[c];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        psrc[c] = odd ? alt.p[c] : own_p[c];
        pdst[c] = odd ? own_p[c] : alt.p[c];
    }
    TbCtx<T> x;
    x.px = ea.pml[0] || ea.pml[1];
    x.py = ea.pml[2] || ea.pml[3];
    x.Nx = Nx;
    x.Ny = Ny;
    x.nosc = ea.nosc;
    x.osc = ea.osc;
    x.ikzE = ea.pf.ikE[2][0];
    x.ikzH = ha.pf.ikH[2][0];
    x.dtm = ha.dtm; x.dtdmx = ha.dtdmx; x.dtdmy = ha.dtdmy; x.dtdmz = ha.dtdmz;
    x.bEx = ea.pf.bE[0]; x.cEx = ea.pf.cE[0]; x.bEy = ea.pf.bE[1]; x.cEy = ea.pf.cE[1];
    x.bHx = ha.pf.bH[0]; x.cHx = ha.pf.cH[0]; x.bHy = ha.pf.bH[1]; x.cHy = ha.pf.cH[1];
    // per-column / per-row data of the lane's cells (profiles b, c/D of E and H)
    int gi[CX], gj[RY], nxE[CX], nyE[RY], nxH[CX], nyH[RY];
    T ikxE[CX], ikyE[RY], ikxH[CX], ikyH[RY];
    TbProf<T> PX[CX], PY[RY];
#pragma unroll
    for (int cx = 0; cx < CX; cx++) {
        int g = (ox - TS + tx + 64 * cx) % Nx;
        if (g < 0) g += Nx;
        gi[cx] = g;
        nxE[cx] = slabE(g, Nx, ea.pml[0], ea.pml[1]);
        nxH[cx] = slabH(g, Nx, ea.pml[0], ea.pml[1]);
        ikxE[cx] = ea.pf.ikE[0][g];
        ikxH[cx] = ha.pf.ikH[0][g];
        const T z = 0;
        PX[cx] = PML && x.px ? TbProf<T>{x.bEx[g], x.cEx[g], x.bHx[g], x.cHx[g]} : TbProf<T>{z, z, z, z};
    }
#pragma unroll
    for (int ry = 0; ry < RY; ry++) {
        int g = (oy - TS + wv + TB_WAVES * ry) % Ny;
        if (g < 0) g += Ny;
        gj[ry] = g;
        nyE[ry] = slabE(g, Ny, ea.pml[2], ea.pml[3]);
        nyH[ry] = slabH(g, Ny, ea.pml[2], ea.pml[3]);
        ikyE[ry] = ea.pf.ikE[1][g];
        ikyH[ry] = ha.pf.ikH[1][g];
        const T z = 0;
        PY[ry] = PML && x.py ? TbProf<T>{x.bEy[g], x.cEy[g], x.bHy[g], x.cHy[g]} : TbProf<T>{z, z, z, z};
    }
    // load: fields into LDS, material ids and psi into registers
    TbCoef<T> cf[CX * RY];
    T pa[CX * RY], pb[CX * RY], pc[CX * RY], pd[CX * RY];
#pragma unroll
    for (int ry = 0; ry < RY; ry++)
#pragma unroll
        for (int cx = 0; cx < CX; cx++) {
            const int c = ry * CX + cx, li = tx + 64 * cx, lj = wv + TB_WAVES * ry, l = li + TBX * lj;
            const int id = gi[cx] + Nx * gj[ry];
            F0[l] = src[0][id];
            F1[l] = src[1][id];
            F2[l] = src[2][id];
            const MatDev<T> &G = ea.mt[ea.mat[id]];
            cf[c] = TbCoef<T>{G.C1, G.C2x, G.C2y, G.C4};
            const T z = 0;
            if (!PML) {
                pa[c] = pb[c] = pc[c] = pd[c] = z;
            } else if (POL == 0) {
                pa[c] = nxE[cx] >= 0 ? psrc[0][nxE[cx] + wxs * gj[ry]] : z;
                pb[c] = nyE[ry] >= 0 ? psrc[1][gi[cx] + Nx * nyE[ry]] : z;
                pc[c] = nyH[ry] >= 0 ? psrc[2][gi[cx] + Nx * nyH[ry]] : z;
                pd[c] = nxH[cx] >= 0 ? psrc[3][nxH[cx] + wxs * gj[ry]] : z;
            } else {
                pa[c] = nyE[ry] >= 0 ? psrc[0][gi[cx] + Nx * nyE[ry]] : z;
                pb[c] = nxE[cx] >= 0 ? psrc[1][nxE[cx] + wxs * gj[ry]] : z;
                pc[c] = nxH[cx] >= 0 ? psrc[2][nxH[cx] + wxs * gj[ry]] : z;
                pd[c] = nyH[ry] >= 0 ? psrc[3][gi[cx] + Nx * nyH[ry]] : z;
            }
        }
    // sensors of this grid
    const int sb = sbeg[2 * r], se = sbeg[2 * r + 1];
    // per cell: the oscillators covering it (bits 0-7) and the sensors it is a
    // sample point of (bits 8-31, sensors beyond the 24th are checked in full)
    int hit[CX * RY];
#pragma unroll
    for (int c = 0; c < CX * RY; c++) hit[c] = 0;
    for (int q = 0; q < x.nosc; q++) {
        const int x1 = x.osc[q].x1, x2 = x.osc[q].x2, y1 = x.osc[q].y1, y2 = x.osc[q].y2;
        const bool zin = x.osc[q].z1 <= 0 && x.osc[q].z2 > 0;
#pragma unroll
        for (int ry = 0; ry < RY; ry++)
#pragma unroll
            for (int cx = 0; cx < CX; cx++)
                if (zin && gi[cx] >= x1 && gi[cx] < x2 && gj[ry] >= y1 && gj[ry] < y2) hit[ry * CX + cx] |= 1 << q;
    }
    for (int e = sb; e < se; e++) {
        const SensDesc d = samp[e].d;
        const int bit = e - sb < 24 ? 1 << (8 + e - sb) : 1 << 31;
#pragma unroll
        for (int ry = 0; ry < RY; ry++)
#pragma unroll
            for (int cx = 0; cx < CX; cx++) {
                const int i = gi[cx], j = gj[ry];
                bool on;
                if (d.type == AEFDTD_NORMAL_X || d.type == AEFDTD_NORMAL_XM)
                    on = i == d.x1 && j >= d.y1 && j < d.y1 + d.span1 && d.z1 == 0;
                else if (d.type == AEFDTD_NORMAL_Y || d.type == AEFDTD_NORMAL_YM)
                    on = j == d.y1 && i >= d.x1 && i < d.x1 + d.span1 && d.z1 == 0;
                else
                    on = d.z1 == 0 && i >= d.x1 && i < d.x1 + d.span1 && j >= d.y1 && j < d.y1 + d.span2;
                if (on) hit[ry * CX + cx] |= bit;
            }
    }
    // a tile none of whose inner cells is a sample point skips the sampling and its barrier
    int any_sample = 0;
#pragma unroll
    for (int ry = 0; ry < RY; ry++)
#pragma unroll
        for (int cx = 0; cx < CX; cx++) {
            const int ui = tx + 64 * cx - TS, uj = wv + TB_WAVES * ry - TS;
            if (ui >= 0 && ui < IX && uj >= 0 && uj < IY && ox + ui < Nx && oy + uj < Ny)
                any_sample |= hit[ry * CX + cx] >> 8;
        }
    any_sample = __syncthreads_or(any_sample);
    const T zero = 0;
#pragma unroll 1
    for (int st = 0; st < nsteps; st++) {
        const int t = tbeg + st;
        // E update (cells with an i-1 and a j-1 neighbour in the tile)
#pragma unroll
        for (int ry = 0; ry < RY; ry++)
#pragma unroll
            for (int cx = 0; cx < CX; cx++) {
                const int c = ry * CX + cx, li = tx + 64 * cx, lj = wv + TB_WAVES * ry, l = li + TBX * lj;
                if (li < 1 || lj < 1) continue;
                if (POL == 0) {
                    T e0 = F0[l], e1 = zero;
                    tb_e_cell<0, PML>(x, cf[c], gi[cx], gj[ry], t, e0, e1, F2[l] - F2[l - 1], F1[l] - F1[l - TBX],
                                 ikxE[cx], ikyE[ry], nxE[cx], nyE[ry], PX[cx], PY[ry], pa[c], pb[c], hit[c] & 0xff);
                    F0[l] = e0;
                } else {
                    T e0 = F0[l], e1 = F1[l];
                    tb_e_cell<1, PML>(x, cf[c], gi[cx], gj[ry], t, e0, e1, F2[l] - F2[l - TBX], F2[l] - F2[l - 1],
                                 ikxE[cx], ikyE[ry], nxE[cx], nyE[ry], PX[cx], PY[ry], pa[c], pb[c], hit[c] & 0xff);
                    F0[l] = e0;
                    F1[l] = e1;
                }
            }
        __syncthreads();
        // H update (cells with an i+1 and a j+1 neighbour in the tile)
#pragma unroll
        for (int ry = 0; ry < RY; ry++)
#pragma unroll
            for (int cx = 0; cx < CX; cx++) {
                const int c = ry * CX + cx, li = tx + 64 * cx, lj = wv + TB_WAVES * ry, l = li + TBX * lj;
                if (li > TBX - 2 || lj > LY - 2) continue;
                if (POL == 0) {
                    T h0 = F1[l], h1 = F2[l];
                    const T ez = F0[l];
                    tb_h_cell<0, PML>(x, gi[cx], gj[ry], h0, h1, F0[l + TBX] - ez, F0[l + 1] - ez, ikxH[cx], ikyH[ry],
                                 nxH[cx], nyH[ry], PX[cx], PY[ry], pc[c], pd[c]);
                    F1[l] = h0;
                    F2[l] = h1;
                } else {
                    T h0 = F2[l], h1 = zero;
                    tb_h_cell<1, PML>(x, gi[cx], gj[ry], h0, h1, F0[l + TBX] - F0[l], F1[l + 1] - F1[l], ikxH[cx],
                                 ikyH[ry], nxH[cx], nyH[ry], PX[cx], PY[ry], pc[c], pd[c]);
                    F2[l] = h0;
                }
            }
        __syncthreads();
        // sensor samples of the inner cells (local_E/H, core/fdtd_utils.cpp:457-507)
        if (any_sample) {
#pragma unroll
            for (int ry = 0; ry < RY; ry++)
#pragma unroll
                for (int cx = 0; cx < CX; cx++) {
                    const int c = ry * CX + cx, li = tx + 64 * cx, lj = wv + TB_WAVES * ry, l = li + TBX * lj;
                    const int ui = li - TS, uj = lj - TS;
                    if ((hit[c] >> 8) == 0 || ui < 0 || ui >= IX || uj < 0 || uj >= IY || ox + ui >= Nx ||
                        oy + uj >= Ny)
                        continue;
                    const int i = gi[cx], j = gj[ry];
#pragma unroll 1
                    for (int e = sb; e < se; e++) {
                        const int bit = e - sb < 24 ? 1 << (8 + e - sb) : 1 << 31;
                        if (!(hit[c] & bit)) continue;
                        const SensDesc &d = samp[e].d;
                        long p;
                        if (d.type == AEFDTD_NORMAL_X || d.type == AEFDTD_NORMAL_XM) {
                            if (i != d.x1 || j < d.y1 || j >= d.y1 + d.span1 || d.z1 != 0) continue;
                            p = j - d.y1;
                        } else if (d.type == AEFDTD_NORMAL_Y || d.type == AEFDTD_NORMAL_YM) {
                            if (j != d.y1 || i < d.x1 || i >= d.x1 + d.span1 || d.z1 != 0) continue;
                            p = i - d.x1;
                        } else {
                            if (d.z1 != 0 || i < d.x1 || i >= d.x1 + d.span1 || j < d.y1 || j >= d.y1 + d.span2)
                                continue;
                            p = (i - d.x1) + (long)d.span1 * (j - d.y1);
                        }
                        T v[6];
                        if (POL == 0) {
                            v[0] = zero; v[1] = zero; v[5] = zero;
                            v[2] = (F0[l] + F0[l + 1] + F0[l + TBX] + F0[l + 1 + TBX]) / (T)4.0;
                            v[3] = (F1[l] + F1[l + 1]) / (T)2.0;
                            v[4] = (F2[l] + F2[l + TBX]) / (T)2.0;
                        } else {
                            v[2] = zero; v[3] = zero; v[4] = zero;
                            v[0] = (F0[l] + F0[l + TBX] + F0[l] + F0[l + TBX]) / (T)4.0;
                            v[1] = (F1[l] + F1[l + 1] + F1[l] + F1[l + 1]) / (T)4.0;
                            v[5] = (F2[l] + F2[l]) / (T)2.0;
                        }
                        const SampB<S, T> &qs = samp[e];
                        const int slot = s0 + st;
#pragma unroll
                        for (int cc = 0; cc < 6; cc++) qs.stage[((size_t)slot * 6 + cc) * qs.Ptot + d.off + p] = v[cc];
                    }
                }
            __syncthreads();
        }
    }
    // write back the inner cells
#pragma unroll
    for (int ry = 0; ry < RY; ry++)
#pragma unroll
        for (int cx = 0; cx < CX; cx++) {
            const int c = ry * CX + cx, li = tx + 64 * cx, lj = wv + TB_WAVES * ry, l = li + TBX * lj;
            const int ui = li - TS, uj = lj - TS;
            if (ui < 0 || ui >= IX || uj < 0 || uj >= IY || ox + ui >= Nx || oy + uj >= Ny) continue;
            const int id = gi[cx] + Nx * gj[ry];
            dst[0].set(id, F0[l]);
            dst[1].set(id, F1[l]);
            dst[2].set(id, F2[l]);
            if (!PML) continue;
            if (POL == 0) {
                if (nxE[cx] >= 0) pdst[0][nxE[cx] + wxs * gj[ry]] = pa[c];
                if (nyE[ry] >= 0) pdst[1][gi[cx] + Nx * nyE[ry]] = pb[c];
                if (nyH[ry] >= 0) pdst[2][gi[cx] + Nx * nyH[ry]] = pc[c];
                if (nxH[cx] >= 0) pdst[3][nxH[cx] + wxs * gj[ry]] = pd[c];
            } else {
                if (nyE[ry] >= 0) pdst[0][gi[cx] + Nx * nyE[ry]] = pa[c];
                if (nxE[cx] >= 0) pdst[1][nxE[cx] + wxs * gj[ry]] = pb[c];
                if (nxH[cx] >= 0) pdst[2][nxH[cx] + wxs * gj[ry]] = pc[c];
                if (nyH[ry] >= 0) pdst[3][gi[cx] + Nx * nyH[ry]] = pd[c];
            }
        }
}

// Spectral completion criterion, one lane per watched cell
// (CompletionSensor::deep_feed, sensors/s_completion.cpp:132-158): the running
// DFT acc += local_E * exp(i w step Dt) in FP64 (local_E as local_Ex..Ez,
// core/fdtd_utils.cpp:457-479), and its real and imaginary parts stored at
// slot step % navg of the ring.
template <typename S, typename T>
__global__ __launch_bounds__(256) void k_ft_feed(const int *__restrict__ pts, const double *__restrict__ w, int np,
                                                 int step, double Dt, int navg, Fld<S, T> Ex, Fld<S, T> Ey,
                                                 Fld<S, T> Ez, int Nx, int Ny, int Nz, int P,
                                                 double *__restrict__ acc, double *__restrict__ ring) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const int i = pts[3 * p], j = pts[3 * p + 1], k = pts[3 * p + 2];
    const int ip = (i + 1 == Nx) ? 0 : i + 1, jp = (j + 1 == Ny) ? 0 : j + 1, kp = (k + 1 == Nz) ? 0 : k + 1;
    auto at = [&](const Fld<S, T> &f, int x, int y, int z) { return f[x + P * (y + Ny * z)]; };
    const double e[3] = {
        (double)((at(Ex, i, j, k) + at(Ex, i, jp, k) + at(Ex, i, j, kp) + at(Ex, i, jp, kp)) / (T)4.0),
        (double)((at(Ey, i, j, k) + at(Ey, ip, j, k) + at(Ey, i, j, kp) + at(Ey, ip, j, kp)) / (T)4.0),
        (double)((at(Ez, i, j, k) + at(Ez, ip, j, k) + at(Ez, i, jp, k) + at(Ez, ip, jp, k)) / (T)4.0)};
    double sn, cs;
    sincos(w[p] * step * Dt, &sn, &cs);
    const int off = step % navg;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double re = acc[(2 * c) * np + p] + e[c] * cs, im = acc[(2 * c + 1) * np + p] + e[c] * sn;
        acc[(2 * c) * np + p] = re;
        acc[(2 * c + 1) * np + p] = im;
        ring[((size_t)c * np + p) * navg + off] = re;
        ring[((size_t)(3 + c) * np + p) * navg + off] = im;
    }
}

// Slot table of the dispersive cells: slot[cells[i]] = i (the table is -1
// elsewhere); cells holds device indices, all below the table's length.
__global__ __launch_bounds__(256) void k_slot_scatter(const int2 *__restrict__ cells, long n,
                                                      int32_t *__restrict__ slot) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) slot[cells[i].x] = (int32_t)i;
}

// CompletionSensor energy (sensors/s_completion.cpp:160-180): the sum of
// Ex^2 + Ey^2 + Ez^2 over every cell of the grid in FP64, as fixed per-block
// partials reduced in a fixed order (reproducible run to run); em[0] is the
// last sum, em[1] the running maximum.
constexpr int EN_BLOCKS = 2048;

template <typename S, typename T>
__global__ __launch_bounds__(256) void k_energy_part(Fld<S, T> Ex, Fld<S, T> Ey, Fld<S, T> Ez, int Nx, int rows, int P,
                                                     double *part) {
    // wavefronts take groups of four grid rows (j, k), lanes along x, the four
    // rows' loads issued together; groups are strided over the launch
    __shared__ double sh[256];
    double acc = 0;
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    for (int r0 = wave * 4; r0 < rows; r0 += EN_BLOCKS * 16) {
        for (int i = lane; i < Nx; i += 64) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (r0 + q < rows) {
                    const int c = P * (r0 + q) + i;
                    const double ex = Ex[c], ey = Ey[c], ez = Ez[c];
                    acc += ex * ex + ey * ey + ez * ez;
                }
            }
        }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Sum of in[0..n) in gridDim.x contiguous chunks, one per workgroup, into
// out[0..gridDim.x) (fixed order: the same sums on every run).
__global__ __launch_bounds__(256) void k_energy_sum(const double *__restrict__ in, long n, double *__restrict__ out) {
    __shared__ double sh[256];
    const long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long lo = (long)blockIdx.x * chunk, hi = min(n, lo + chunk);
    double acc = 0;
    for (long b = lo + threadIdx.x; b < hi; b += 256) acc += in[b];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_energy_fin(const double *part, int n, double *em) {
    __shared__ double sh[256];
    double acc = 0;
    for (int b = threadIdx.x; b < n; b += 256) acc += part[b];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        em[0] = sh[0];
        em[1] = fmax(em[1], sh[0]);
    }
}

// ================================================================ one-pass step
// A full time step of a 3D grid in one pass (E, then H from the new E) that
// reads the fields of step n from one set of arrays and writes those of step
// n+1 to a second set, which the host then takes as the current one.  No
// workgroup ever reads a value that another has already advanced, so the
// tiles need no ordering.
//
// A workgroup of NW wavefronts owns a tile of 64 columns (aligned to the
// 64-cell row pitch, so that every row access is whole cache lines), R = NW-2
// rows and planes k0 .. k1-1.  Wavefronts 0..R-1 take its rows, wavefront R
// the row after them (the +y halo) and wavefront NW-1 the column after them
// (the +x halo, one lane per row); every wavefront marches planes k0 .. k1,
// the last being the +z halo.  The H update of an owned cell needs E^{n+1} at
// i+1, j+1 and k+1, so the tile computes E^{n+1} of its halo cells as well
// (their owners compute them too, from the same inputs with the same
// operations, so both agree bit for bit) and discards it.  Per plane k:
//
//   E^{n+1}(k) of every cell of the tile and its halo: E^n, H^n of the cell,
//     H^n at i-1 and j-1 from memory, H^n at k-1 carried from plane k-1
//   Ex, Ez of E^{n+1}(k) of each row, Ey, Ez of the halo column, to LDS
//   H^{n+1}(k-1) of the owned cells: E^{n+1}(k-1) and E^{n+1}(k) of the cell
//     from registers, i+1 by a lane shuffle (the last owned lane: the halo
//     column in LDS), j+1 from the LDS copy of plane k-1
//   one workgroup barrier (LDS is double buffered by plane parity)
//
// HBM sees each field read once and written once per step, 12 values and the
// material id per cell against 18 values and the id of the E and H passes;
// the halo column, row and plane are read again by their tiles, mostly from
// the caches.  E's CPML psi are also read from one set and written to the
// other (a halo cell's psi belongs to its owner); H's are updated in place,
// since only an H cell's owner updates it.  Every cell sees the inputs and
// operations of k_update_E + k_update_H, so the results are bit-identical to
// the two passes.

// Fields and E CPML psi of step n+1 (the second set; same layouts).
template <typename S, typename T>
struct StepOut {
    Fld<S, T> Ex, Ey, Ez, Hx, Hy, Hz;
    T *pEyx, *pEzx, *pExy, *pEzy, *pExz, *pEyz;
};

// Tiles of the one-pass step: ntx x nty x ntz, KZ planes per tile, XCD fold.
struct StepGeom {
    int ntx, nty, ntz, KZ, remap;
};

// A CPML psi of step n held in a register; its update goes to the step n+1
// set, and only where the tile owns the cell (g null: a halo cell).
template <typename T>
struct PsiNext {
    T &r;
    T *g;
    __device__ __forceinline__ operator T() const { return r; }
    __device__ __forceinline__ void operator=(T v) const {
        r = v;
        if (g) *g = v;
    }
};

// Boundary data of an E cell in the one-pass step: psi loaded ahead from the
// step n set, updates written to the step n+1 set at the same offset.
template <typename S, typename T>
struct NextPsiE {
    const EArgs<S, T> &a;
    const StepOut<S, T> &o;
    int i, j, k;
    bool own;
    PsiSetE<T> &q;
    T bx_, cx_, by_, cy_;
    __device__ __forceinline__ PsiAddrE<S, T> ad() const { return PsiAddrE<S, T>{a, i, j, k}; }
    __device__ __forceinline__ PsiNext<T> Exy(int n) const { return {q.xy, own ? o.pExy + (ad().xy(n) - a.pExy) : nullptr}; }
    __device__ __forceinline__ PsiNext<T> Exz(int n) const { return {q.xz, own ? o.pExz + (ad().xz(n) - a.pExz) : nullptr}; }
    __device__ __forceinline__ PsiNext<T> Eyx(int n) const { return {q.yx, own ? o.pEyx + (ad().yx(n) - a.pEyx) : nullptr}; }
    __device__ __forceinline__ PsiNext<T> Eyz(int n) const { return {q.yz, own ? o.pEyz + (ad().yz(n) - a.pEyz) : nullptr}; }
    __device__ __forceinline__ PsiNext<T> Ezx(int n) const { return {q.zx, own ? o.pEzx + (ad().zx(n) - a.pEzx) : nullptr}; }
    __device__ __forceinline__ PsiNext<T> Ezy(int n) const { return {q.zy, own ? o.pEzy + (ad().zy(n) - a.pEzy) : nullptr}; }
    __device__ __forceinline__ T bx() const { return bx_; }
    __device__ __forceinline__ T cx() const { return cx_; }
    __device__ __forceinline__ T by() const { return by_; }
    __device__ __forceinline__ T cy() const { return cy_; }
};

template <typename S, typename T, int NW, bool EN>
__global__ __launch_bounds__(BX *NW) void k_step(const EArgs<S, T> a, const HArgs<S, T> h, const StepOut<S, T> o,
                                                 const StepGeom g) {
    constexpr int R = NW - 2;                                  // rows a tile owns
    extern __shared__ char lds_raw[];
    T *xz = reinterpret_cast<T *>(lds_raw);                    // [2 planes][R+1 rows][Ex, Ez][64]
    T *yzc = xz + 2 * (R + 1) * 2 * BX;                        // [2 planes][Ey, Ez][64]: the halo column
    MatL<T> *sm = reinterpret_cast<MatL<T> *>(yzc + 2 * 2 * BX);
    stage_mats(a, sm);
    const int lane = threadIdx.x, w = threadIdx.y;
    int bx, by, bz;
    tile_of(blockIdx.x, g.ntx, g.nty, g.ntz, g.remap, bx, by, bz);
    const int x0 = bx * BX, y0 = by * R, k0 = bz * g.KZ;
    const int xe = min(x0 + BX, a.Nx), ye = min(y0 + R, a.Ny), ke = min(k0 + g.KZ, a.Nz);
    const int xh = (xe == a.Nx) ? 0 : xe, yh = (ye == a.Ny) ? 0 : ye;   // halo column and row (wrapping)
    const int nr = ye - y0, lastx = xe - 1 - x0;              // owned rows, last owned lane
    // the lane's cell: owned, halo or idle (an idle lane computes on an
    // in-bounds cell and stores nothing)
    const bool colw = w == NW - 1;                             // the halo-column wavefront
    int i, j;
    bool own;
    if (colw) {
        i = xh;
        j = lane < nr ? y0 + lane : y0;
        own = false;
    } else {
        i = lane <= lastx ? x0 + lane : x0;
        j = w < nr ? y0 + w : (w == nr ? yh : y0);
        own = lane <= lastx && w < nr;
    }
    const int NxNy = a.P * a.Ny;
    const int oim = (i == 0) ? a.Nx - 1 : -1;                  // (i-1) with wrap
    const int ojm = (j == 0) ? (a.Ny - 1) * a.P : -a.P;        // (j-1) with wrap
    // side masks: E over the tile and its halo (every treatment where a halo
    // wraps around the grid), H over the owned cells
    const bool wraps = xe == a.Nx || ye == a.Ny || ke == a.Nz;
    const int sideE = wraps ? side_all(a) : side_block(a, 0, x0, xe + 1, y0, ye + 1, k0, ke + 1);
    const int sideH = side_block(h, 1, x0, xe, y0, ye, k0, ke);
    const T ikxE = a.pf.ikE[0][i], ikyE = a.pf.ikE[1][j], ikxH = h.pf.ikH[0][i], ikyH = h.pf.ikH[1][j];
    const int nxE = slabE(i, a.Nx, a.pml[0], a.pml[1]), nyE = slabE(j, a.Ny, a.pml[2], a.pml[3]);
    const int nxH = slabH(i, a.Nx, h.pml[0], h.pml[1]), nyH = slabH(j, a.Ny, h.pml[2], h.pml[3]);
    T bxE = 0, cxE = 0, byE = 0, cyE = 0, bxH = 0, cxH = 0, byH = 0, cyH = 0;
    if ((sideE & SIDE_X) && nxE >= 0) { bxE = a.pf.bE[0][i]; cxE = a.pf.cE[0][i]; }
    if ((sideE & SIDE_Y) && nyE >= 0) { byE = a.pf.bE[1][j]; cyE = a.pf.cE[1][j]; }
    if ((sideH & SIDE_X) && nxH >= 0) { bxH = h.pf.bH[0][i]; cxH = h.pf.cH[0][i]; }
    if ((sideH & SIDE_Y) && nyH >= 0) { byH = h.pf.bH[1][j]; cyH = h.pf.cH[1][j]; }
    // strip slots (sensor planes normal to x) of an owned lane's column
    int sq0 = -1, sq1 = -1;
    for (int q = 0; own && q < h.nstrip; q++) {
        const int c0 = h.scol[q], c1 = (c0 + 1 == a.Nx) ? 0 : c0 + 1;
        const int s = (i == c0) ? 2 * q : (i == c1 ? 2 * q + 1 : -1);
        if (s >= 0) {
            if (sq0 < 0) sq0 = s;
            else sq1 = s;
        }
    }
    // one plane's loads: the E update's, and the CPML psi of the H update of
    // the plane before (requested one plane ahead of their use)
    struct Ld {
        T hx, hy, hz, hy_im, hz_im, hx_jm, hz_jm, ex, ey, ez;
        int m;
        PsiSetE<T> qe;
        PsiSetH<T> qh;
    };
    auto load = [&](int k) {
        const int kk = (k == a.Nz) ? 0 : k;
        const int id = i + a.P * j + kk * NxNy;
        Ld v;
        v.hx = a.Hx[id];
        v.hy = a.Hy[id];
        v.hz = a.Hz[id];
        v.hy_im = a.Hy[id + oim];
        v.hz_im = a.Hz[id + oim];
        v.hx_jm = a.Hx[id + ojm];
        v.hz_jm = a.Hz[id + ojm];
        v.ex = a.Ex.nt(id);
        v.ey = a.Ey.nt(id);
        v.ez = a.Ez.nt(id);
        v.m = __builtin_nontemporal_load(a.mat + id);
        v.qe = PsiSetE<T>{};
        const PsiAddrE<S, T> ae{a, i, j, kk};
        if ((sideE & SIDE_X) && nxE >= 0) {
            if (a.en & 2) v.qe.yx = *ae.yx(nxE);
            if (a.en & 4) v.qe.zx = *ae.zx(nxE);
        }
        if ((sideE & SIDE_Y) && nyE >= 0) {
            if (a.en & 1) v.qe.xy = *ae.xy(nyE);
            if (a.en & 4) v.qe.zy = *ae.zy(nyE);
        }
        if (sideE & SIDE_Z) {
            const int nz_ = slabEz(a, kk);
            if (nz_ >= 0) {
                if (a.en & 1) v.qe.xz = *ae.xz(nz_);
                if (a.en & 2) v.qe.yz = *ae.yz(nz_);
            }
        }
        v.qh = PsiSetH<T>{};
        if (k > k0) {
            const PsiAddrH<S, T> ah{h, i, j, k - 1};
            if ((sideH & SIDE_X) && nxH >= 0) {
                if (h.en & 16) v.qh.yx = *ah.yx(nxH);
                if (h.en & 32) v.qh.zx = *ah.zx(nxH);
            }
            if ((sideH & SIDE_Y) && nyH >= 0) {
                if (h.en & 8) v.qh.xy = *ah.xy(nyH);
                if (h.en & 32) v.qh.zy = *ah.zy(nyH);
            }
            if (sideH & SIDE_Z) {
                const int nz_ = slabH(k - 1, h.Nz, h.pml[4], h.pml[5]);
                if (nz_ >= 0) {
                    if (h.en & 8) v.qh.xz = *ah.xz(nz_);
                    if (h.en & 16) v.qh.yz = *ah.yz(nz_);
                }
            }
        }
        return v;
    };
    // H^n at k0-1 (z neighbour of the first plane), then carried
    const int km0 = (k0 == 0) ? a.Nz - 1 : k0 - 1;
    T hkx = a.Hx[i + a.P * j + km0 * NxNy], hky = a.Hy[i + a.P * j + km0 * NxNy];
    // E^{n+1} and H^n of the cell at the plane before
    T pex = 0, pey = 0, pez = 0, phx = 0, phy = 0, phz = 0;
    T en = 0;
    auto body = [&](Ld &c, int k) {
        const int kk = (k == a.Nz) ? 0 : k;
        const int id = i + a.P * j + kk * NxNy;
        const bool ownE = own && k < ke;
        V3<T> e = e_cell_vals(a, sm[c.m], i, j, kk, id, c.ex, c.ey, c.ez, c.hz - c.hz_jm, c.hy - hky, c.hx - hkx,
                              c.hz - c.hz_im, c.hy - c.hy_im, c.hx - c.hx_jm, ikxE, ikyE, uload(a.pf.ikE[2], kk), nxE,
                              nyE, a.ostep, sideE, NextPsiE<S, T>{a, o, i, j, kk, ownE, c.qe, bxE, cxE, byE, cyE});
        hkx = c.hx;
        hky = c.hy;
        if (ownE) {
            o.Ex.set_nt(id, e.x);
            o.Ey.set_nt(id, e.y);
            o.Ez.set_nt(id, e.z);
            if constexpr (EN) en += e.x * e.x + e.y * e.y + e.z * e.z;
        }
        // the H update reads E as stored
        e.x = stored(o.Ex, e.x);
        e.y = stored(o.Ey, e.y);
        e.z = stored(o.Ez, e.z);
        if (colw) {
            T *col = yzc + (k & 1) * 2 * BX;
            col[lane] = e.y;
            col[BX + lane] = e.z;
        } else if (w <= R) {
            T *row = xz + ((k & 1) * (R + 1) + w) * 2 * BX;
            row[lane] = e.x;
            row[BX + lane] = e.z;
        }
        if (k > k0) {
            T ey_ip = __shfl_down(pey, 1, 64), ez_ip = __shfl_down(pez, 1, 64);
            if (own) {
                const int kh = k - 1;
                if (lane == lastx) {
                    const T *col = yzc + ((k - 1) & 1) * 2 * BX;
                    ey_ip = col[w];
                    ez_ip = col[BX + w];
                }
                const T *up = xz + (((k - 1) & 1) * (R + 1) + w + 1) * 2 * BX;
                const T ex_jp = up[lane], ez_jp = up[BX + lane];
                T hx = phx, hy = phy, hz = phz;
                h_cell_vals(h, i, j, kh, hx, hy, hz, e.y - pey, e.x - pex, ez_jp - pez, ez_ip - pez, ex_jp - pex,
                            ey_ip - pey, ikxH, ikyH, uload(h.pf.ikH[2], kh), nxH, nyH, sideH,
                            PrefPsiH<S, T>{h, i, j, kh, c.qh, bxH, cxH, byH, cyH});
                const int idh = i + a.P * j + kh * NxNy;
                o.Hx.set_nt(idh, hx);
                o.Hy.set_nt(idh, hy);
                o.Hz.set_nt(idh, hz);
                if (sq0 >= 0) {
                    const T v[6] = {pex, pey, pez, stored(o.Hx, hx), stored(o.Hy, hy), stored(o.Hz, hz)};
                    strip_put(h, sq0, j, kh, v);
                    if (sq1 >= 0) strip_put(h, sq1, j, kh, v);
                }
            }
        }
        pex = e.x;
        pey = e.y;
        pez = e.z;
        phx = c.hx;
        phy = c.hy;
        phz = c.hz;
        __syncthreads();
    };
    // software pipeline: plane k+1 requested before plane k is computed; the
    // last (halo) plane peeled off
    Ld cur = load(k0);
    for (int k = k0; k < ke; k++) {
        Ld nxt = load(k + 1);
        body(cur, k);
        cur = nxt;
    }
    body(cur, ke);
    if constexpr (EN) {
        __shared__ double red[NW];
        double s = (double)en;
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) red[w] = s;
        __syncthreads();
        if (lane == 0 && w == 0) {
            double t = 0;
            for (int y = 0; y < NW; y++) t += red[y];
            a.en_part[blockIdx.x] = t;
        }
    }
}

// ================================================================ FP16 storage range
// Largest |stored value| of each field component, as the bit pattern of a
// float (non-negative floats order as their bits; infinity and NaN above every
// finite value), so that one atomicMax per workgroup and component suffices.
struct FieldPtrs {
    const void *f[6];
};

template <typename S>
__global__ __launch_bounds__(256) void k_absmax(const FieldPtrs fp, size_t n, unsigned *out) {
    const S *f = static_cast<const S *>(fp.f[blockIdx.y]);
    unsigned m = 0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        m = max(m, __float_as_uint(fabsf((float)f[q])));
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
    __shared__ unsigned red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out + blockIdx.y, max(max(red[0], red[1]), max(red[2], red[3])));
}

__global__ void k_zero_u32(unsigned *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Stored values of one component times 2^d (exact for normal FP16 values).
template <typename S>
__global__ __launch_bounds__(256) void k_rescale(S *f, size_t n, float fac) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        f[q] = (S)((float)f[q] * fac);
}

// Streaming copy used to measure the HBM bandwidth a kernel can reach on
// this device (16-byte accesses, grid-stride).
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__global__ __launch_bounds__(256) void k_stream_copy(const f32x4 *__restrict__ a, f32x4 *__restrict__ b, size_t n) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        if constexpr (NT) __builtin_nontemporal_store(__builtin_nontemporal_load(a + q), b + q);
        else b[q] = a[q];
    }
}

// Streaming reference with the read:write mix of the E and H passes: six
// arrays read and three written, 4 bytes per lane, non-temporal accesses.
__global__ __launch_bounds__(256) void k_stream_mix(const float *const *__restrict__ x, float *const *__restrict__ y,
                                                    size_t n) {
    const float *x0 = x[0], *x1 = x[1], *x2 = x[2], *x3 = x[3], *x4 = x[4], *x5 = x[5];
    float *y0 = y[0], *y1 = y[1], *y2 = y[2];
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const float a0 = __builtin_nontemporal_load(x0 + q), a1 = __builtin_nontemporal_load(x1 + q);
        const float a2 = __builtin_nontemporal_load(x2 + q), a3 = __builtin_nontemporal_load(x3 + q);
        const float a4 = __builtin_nontemporal_load(x4 + q), a5 = __builtin_nontemporal_load(x5 + q);
        __builtin_nontemporal_store(a0 + a3, y0 + q);
        __builtin_nontemporal_store(a1 + a4, y1 + q);
        __builtin_nontemporal_store(a2 + a5, y2 + q);
    }
}

}  // namespace dev
}  // namespace aefdtd
