// structure.hip -- Structure::discretize on the device.
//
// Aether builds a structure as a list of operations (structure/
// structure_primitives.cpp) and discretises it with a serial triple loop that
// asks Structure::index for the material at every cell corner
// (structure/structure.cpp:94-113,223-262).  Here one lane evaluates one
// corner: a 64 x 4 workgroup covers 64 x-consecutive cells of four rows, so
// the uint32 output rows are written coalesced, and the operation table is
// read with wave-uniform addresses (scalar loads, shared by the wave).
//
// Identical results: every point test below performs the reference's FP64
// operations in the reference's order, and FP contraction is off for the
// whole file, so no multiply-add is fused where the reference rounds twice.
// Derived operation constants (block corner order, cylinder/cone axis and
// length, bounding boxes, a coating's deposition direction) are computed on
// the host exactly as the reference's constructors compute them.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numbers>
#include <string>
#include <vector>

#include "engine.hpp"

namespace {

struct OpDev {
    int type, mat;
    double a[14];
};

// Host-side constants of one operation, as the reference's constructors set
// them up from the script's arguments.
int prepare(const aefdtd_structure_op &q, OpDev &o) {
    const double *p = q.p;
    o.type = q.type;
    o.mat = q.mat_index;
    std::fill(o.a, o.a + 14, 0.0);
    switch (q.type) {
    case AEFDTD_OP_BLOCK: {   // Add_Block ctor swaps each pair into order (structure_primitives.cpp:24-32)
        for (int d = 0; d < 3; d++) {
            double lo = p[2 * d], hi = p[2 * d + 1];
            if (lo > hi) std::swap(lo, hi);
            o.a[2 * d] = lo;
            o.a[2 * d + 1] = hi;
        }
        return AEFDTD_OK;
    }
    case AEFDTD_OP_SPHERE:    // centre, radius (:484-488)
        if (!(p[3] >= 0)) return AEFDTD_EINVAL;
        for (int d = 0; d < 4; d++) o.a[d] = p[d];
        return AEFDTD_OK;
    case AEFDTD_OP_CYLINDER:
    case AEFDTD_OP_CONE: {
        // the Lua binding passes the far end as x1 + dx (lua_structure.cpp:376-397,
        // 334-354); the ctor takes the axis as far end minus start, its length
        // as sqrt(x^2+y^2+z^2) and divides each component by it (:255-264,
        // common/vector23.cpp:266-283)
        const double x1 = p[0], y1 = p[1], z1 = p[2], r = p[6];
        const double x2 = x1 + p[3], y2 = y1 + p[4], z2 = z1 + p[5];
        double bx = x2 - x1, by = y2 - y1, bz = z2 - z1;
        const double len = std::sqrt(bx * bx + by * by + bz * bz);
        if (!(len > 0) || !(r >= 0)) return AEFDTD_EINVAL;
        bx = bx / len;
        by = by / len;
        bz = bz / len;
        const double v[14] = {x1, y1, z1, bx, by, bz, len, r,
                              std::min(x1, x2) - r, std::max(x1, x2) + r,
                              std::min(y1, y2) - r, std::max(y1, y2) + r,
                              std::min(z1, z2) - r, std::max(z1, z2) + r};
        std::copy(v, v + 14, o.a);
        return AEFDTD_OK;
    }
    case AEFDTD_OP_ELLIPSOID: {   // centre, arm lengths; box centre -+ arm (:291-310)
        const double v[12] = {p[0], p[1], p[2], p[3], p[4], p[5],
                              p[0] - p[3], p[0] + p[3], p[1] - p[4], p[1] + p[4], p[2] - p[5], p[2] + p[5]};
        std::copy(v, v + 12, o.a);
        return AEFDTD_OK;
    }
    case AEFDTD_OP_LAYER: {   // axis 0/1/2, half-open [z1, z2) (:317-338)
        const int axis = (int)p[0];
        if (axis < 0 || axis > 2 || (double)axis != p[0]) return AEFDTD_EINVAL;
        o.a[0] = axis;
        o.a[1] = p[1];
        o.a[2] = p[2];
        return AEFDTD_OK;
    }
    case AEFDTD_OP_COATING: {   // Add_Coating ctor (:45-53), angles as Degree() converts them
        double theta = p[0], phi = p[1];
        theta *= std::numbers::pi / 180.0;
        phi *= std::numbers::pi / 180.0;
        const double v[6] = {std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta), -std::cos(theta),
                             p[2], p[3], p[4]};
        // the support march ends past `thickness` only with delta > 0, the
        // march back up to lz only when it climbs (theta below 90 degrees)
        if (!(p[3] > 0) || !(p[2] >= 0) || !std::isfinite(p[2]) || !(v[2] < 0) || p[4] < 0 ||
            (double)(int)p[4] != p[4])
            return AEFDTD_EINVAL;
        std::copy(v, v + 6, o.a);
        return AEFDTD_OK;
    }
    }
    return AEFDTD_EINVAL;
}

// One operation's index() at (x, y, z): its material, or -1 (every operation
// but a coating).
__device__ __forceinline__ int op_index(const OpDev *__restrict__ ops, int l, double x, double y, double z) {
    const OpDev &o = ops[l];
    const double *a = o.a;
    switch (o.type) {
    case AEFDTD_OP_BLOCK:     // Add_Block::index (:34-39)
        if (x < a[0] || x >= a[1] || y < a[2] || y >= a[3] || z < a[4] || z >= a[5]) return -1;
        return o.mat;
    case AEFDTD_OP_SPHERE: {  // Add_Sphere::index (:490-500)
        const double dx = x - a[0], dy = y - a[1], dz = z - a[2];
        return dx * dx + dy * dy + dz * dz <= a[3] * a[3] ? o.mat : -1;
    }
    case AEFDTD_OP_CYLINDER:  // Add_Cylinder::index (:266-286)
    case AEFDTD_OP_CONE: {    // Add_Cone::index (:112-135)
        if (x < a[8] || x > a[9] || y < a[10] || y > a[11] || z < a[12] || z > a[13]) return -1;
        double vx = x - a[0], vy = y - a[1], vz = z - a[2];
        const double proj = a[3] * vx + a[4] * vy + a[5] * vz;
        if (!(proj >= 0 && proj <= a[6])) return -1;
        vx = vx - proj * a[3];
        vy = vy - proj * a[4];
        vz = vz - proj * a[5];
        const double rt = vx * vx + vy * vy + vz * vz;
        if (o.type == AEFDTD_OP_CYLINDER) return rt <= a[7] * a[7] ? o.mat : -1;
        double r2 = a[7] * (1.0 - proj / a[6]);
        r2 = r2 * r2;
        return rt <= r2 ? o.mat : -1;
    }
    case AEFDTD_OP_ELLIPSOID: {   // Add_Ellipsoid::index (:298-313)
        if (x < a[6] || x > a[7] || y < a[8] || y > a[9] || z < a[10] || z > a[11]) return -1;
        const double ux = (x - a[0]) / a[3], uy = (y - a[1]) / a[4], uz = (z - a[2]) / a[5];
        return ux * ux + uy * uy + uz * uz <= 1.0 ? o.mat : -1;
    }
    case AEFDTD_OP_LAYER: {   // Add_Layer::index (:323-341)
        const int axis = (int)a[0];
        const double v = axis == 0 ? x : (axis == 1 ? y : z);
        return v >= a[1] && v < a[2] ? o.mat : -1;
    }
    }
    return -1;
}

// modulus() of common/mathUT.h:605-611: repeated add / subtract.
__device__ __forceinline__ double fold(double v, double l) {
    while (v < 0) v += l;
    while (v >= l) v -= l;
    return v;
}

struct StructDev {
    int n_ops, def;
    double lx, ly, lz;
    int flip[3], loop[3];
};

// Structure::index(x, y, z, level) (structure.cpp:228-262) over the
// operations below `level`, none of them a coating: flips, folding on looped
// axes, then the operations last first, each with its images.
__device__ int stack_index(const OpDev *__restrict__ ops, const StructDev &s, int level, double x, double y,
                           double z) {
    if (s.flip[0]) x = s.lx - x;
    if (s.flip[1]) y = s.ly - y;
    if (s.flip[2]) z = s.lz - z;
    if (s.loop[0]) x = fold(x, s.lx);
    if (s.loop[1]) y = fold(y, s.ly);
    if (s.loop[2]) z = fold(z, s.lz);
    for (int l = level - 1; l >= 0; l--)
        for (int a = -s.loop[0]; a <= s.loop[0]; a++)
            for (int b = -s.loop[1]; b <= s.loop[1]; b++)
                for (int c = -s.loop[2]; c <= s.loop[2]; c++) {
                    const int v = op_index(ops, l, x + a * s.lx, y + b * s.ly, z + c * s.lz);
                    if (v != -1) return v;
                }
    return s.def;
}

// Add_Coating::index (structure_primitives.cpp:55-96) of operation l: the
// point must be of the origin material under the operations below l; a
// support of another material within `thickness` along dir (the loop tests
// one point past it, as the reference's does); nothing but the origin
// material from the point back up to lz.  The host bounds both marches
// (cap1, cap2 points) from thickness / delta and the grid's z extent, so
// that the caps are never what ends them.
__device__ int coating_index(const OpDev *__restrict__ ops, const StructDev &s, int l, double x, double y, double z) {
    const OpDev &o = ops[l];
    const double dx = o.a[0], dy = o.a[1], dz = o.a[2], th = o.a[3], delta = o.a[4];
    const int origin = (int)o.a[5], cap1 = (int)o.a[6], cap2 = (int)o.a[7];
    if (stack_index(ops, s, l, x, y, z) != origin) return -1;
    double px = x, py = y, pz = z;
    bool support = false;
    for (int i = 0; i < cap1; i++) {
        const double ex = px - x, ey = py - y, ez = pz - z;
        if (!(ex * ex + ey * ey + ez * ez <= th * th && pz > 0)) break;
        const double a = i * delta;
        px = x + a * dx;
        py = y + a * dy;
        pz = z + a * dz;
        if (stack_index(ops, s, l, px, py, pz) != origin) {
            support = true;
            break;
        }
    }
    if (!support) return -1;
    pz = z;
    for (int i = 0; i < cap2 && pz <= s.lz; i++) {
        const double a = i * delta;
        px = x - a * dx;
        py = y - a * dy;
        pz = z - a * dz;
        if (stack_index(ops, s, l, px, py, pz) != origin) return -1;
    }
    return o.mat;
}

__global__ __launch_bounds__(256) void k_discretize(const OpDev *__restrict__ ops, StructDev s, int Nx, int Ny,
                                                    int Nz, double Dx, double Dy, double Dz,
                                                    uint32_t *__restrict__ out) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= Nx || j >= Ny) return;
    double x = i * Dx, y = j * Dy, z = k * Dz;
    if (s.flip[0]) x = s.lx - x;
    if (s.flip[1]) y = s.ly - y;
    if (s.flip[2]) z = s.lz - z;
    if (s.loop[0]) x = fold(x, s.lx);
    if (s.loop[1]) y = fold(y, s.ly);
    if (s.loop[2]) z = fold(z, s.lz);
    int m = s.def;
    for (int l = s.n_ops - 1; l >= 0; l--) {
        int hit = -1;
        const bool coat = ops[l].type == AEFDTD_OP_COATING;
        for (int a = -s.loop[0]; a <= s.loop[0] && hit == -1; a++)
            for (int b = -s.loop[1]; b <= s.loop[1] && hit == -1; b++)
                for (int c = -s.loop[2]; c <= s.loop[2] && hit == -1; c++)
                    hit = coat ? coating_index(ops, s, l, x + a * s.lx, y + b * s.ly, z + c * s.lz)
                               : op_index(ops, l, x + a * s.lx, y + b * s.ly, z + c * s.lz);
        if (hit != -1) {
            m = hit;
            break;
        }
    }
    out[(size_t)i + (size_t)Nx * ((size_t)j + (size_t)Ny * k)] = (uint32_t)m;
}

int sfail(int code, const std::string &msg) {
    aefdtd::set_global_error(msg);
    return code;
}

}  // namespace

extern "C" int aefdtd_discretize(const aefdtd_structure *s, int Nx, int Ny, int Nz, double Dx, double Dy,
                                 double Dz, uint32_t *matsgrid) {
    if (!s || !matsgrid || (s->n_ops > 0 && !s->ops)) return sfail(AEFDTD_EINVAL, "discretize: null argument");
    if (Nx < 1 || Ny < 1 || Nz < 1 || Nz > 65535 || (double)Nx * Ny * Nz >= 2147483648.0)
        return sfail(AEFDTD_EINVAL, "discretize: grid size out of range");
    if (!(Dx > 0) || !(Dy > 0) || !(Dz > 0)) return sfail(AEFDTD_EINVAL, "discretize: cell sizes must be positive");
    StructDev sd{};
    sd.n_ops = s->n_ops;
    sd.def = s->default_material;
    sd.lx = s->lx; sd.ly = s->ly; sd.lz = s->lz;
    const double L[3] = {s->lx, s->ly, s->lz};
    for (int d = 0; d < 3; d++) {
        sd.flip[d] = s->flip[d] != 0;
        sd.loop[d] = s->loop[d] != 0;
        if (sd.loop[d] && !(L[d] > 0)) return sfail(AEFDTD_EINVAL, "discretize: a looped axis needs a positive length");
    }
    std::vector<OpDev> ops(std::max(1, s->n_ops));
    bool coated = false;
    for (int l = 0; l < s->n_ops; l++) {
        if (prepare(s->ops[l], ops[l]))
            return sfail(AEFDTD_EINVAL, "discretize: operation " + std::to_string(l) + " is malformed");
        if (ops[l].type != AEFDTD_OP_COATING) continue;
        if (coated)
            return sfail(AEFDTD_EINVAL, "discretize: a coating over another coating is not provided");
        coated = true;
        // points the marches test: the support march at most thickness / delta
        // + 2; the climb from the lowest point a lane can start at (a grid
        // corner, flipped, or its image one period below) to above lz
        const double *a = ops[l].a;
        const double zr = std::max(Nz * Dz, std::fabs(s->lz)), z0 = -(2 * zr + std::fabs(s->lz));
        const double n1 = a[3] / a[4] + 2, n2 = (s->lz - z0) / (a[4] * -a[2]) + 2;
        if (!(n1 <= 1e6) || !(n2 <= 1e7))
            return sfail(AEFDTD_EINVAL, "discretize: coating " + std::to_string(l) +
                                            " needs too many steps (thickness / delta, or lz / (delta cos theta))");
        ops[l].a[6] = std::ceil(n1) + 1;
        ops[l].a[7] = std::ceil(n2) + 1;
    }
    const size_t N = (size_t)Nx * Ny * Nz;
    hipStream_t st = nullptr;
    OpDev *d_ops = nullptr;
    uint32_t *d_out = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&d_ops, ops.size() * sizeof(OpDev));
    if (e == hipSuccess) e = hipMalloc(&d_out, N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpyAsync(d_ops, ops.data(), ops.size() * sizeof(OpDev), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_discretize, dim3((Nx + 63) / 64, (Ny + 3) / 4, Nz), dim3(64, 4), 0, st, d_ops, sd, Nx, Ny,
                           Nz, Dx, Dy, Dz, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(matsgrid, d_out, N * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    const hipError_t e_run = e;
    if (d_out) (void)hipFree(d_out);
    if (d_ops) (void)hipFree(d_ops);
    if (st) (void)hipStreamDestroy(st);
    if (e_run != hipSuccess)
        return sfail(e_run == hipErrorOutOfMemory ? AEFDTD_ENOMEM : AEFDTD_EDEVICE,
                     std::string("discretize: ") + hipGetErrorString(e_run));
    return AEFDTD_OK;
}
