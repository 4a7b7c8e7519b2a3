// structure_oracle.cpp -- CPU restatement of Aether's Structure::discretize
// for the operations of aefdtd_structure (TEST INFRASTRUCTURE ONLY: loaded by
// tests/ as the checker of aefdtd_discretize, never by the product).
//
// A literal serial reading of structure/structure.cpp:94-113 (discretize) and
// :223-262 (index), with each operation's constructor and index() from
// structure/structure_primitives.cpp and the argument handling of the Lua
// bindings in structure/lua_structure.cpp.  Plain double arithmetic in the
// reference's order; built without -march, so x86-64 has no fused
// multiply-add to contract into.
//
// Pins (tests/test_structure.py): integer lattice-point counts of closed balls
// and discs, half-open blocks and layers, later-wins and the default material,
// loop images against a rolled grid, flip against a mirrored grid, coatings on
// a flat substrate in closed form (vertical and oblique deposition, the point
// past the thickness, an overhang's shadow).  Parity unpinned: the axis
// normalisation in the cylinder and cone constructors for axes that are not
// along x, y or z (pinned only through axis-aligned cases, where it is
// exact), and coatings over curved surfaces and on flipped structures beyond
// what the engine-equals-oracle tests show.
#include <algorithm>
#include <cmath>
#include <memory>
#include <numbers>
#include <vector>

#include "oracle.h"

namespace {

struct Op {
    virtual ~Op() = default;
    virtual int index(double x, double y, double z) const = 0;
};

struct Block : Op {   // Add_Block (structure_primitives.cpp:24-39)
    double x1, x2, y1, y2, z1, z2;
    int mat;
    Block(const double *p, int m) : x1(p[0]), x2(p[1]), y1(p[2]), y2(p[3]), z1(p[4]), z2(p[5]), mat(m) {
        if (x1 > x2) std::swap(x1, x2);
        if (y1 > y2) std::swap(y1, y2);
        if (z1 > z2) std::swap(z1, z2);
    }
    int index(double x, double y, double z) const override {
        if (x < x1 || x >= x2 || y < y1 || y >= y2 || z < z1 || z >= z2) return -1;
        return mat;
    }
};

struct Sphere : Op {   // Add_Sphere (:484-500)
    double x1, y1, z1, r;
    int mat;
    Sphere(const double *p, int m) : x1(p[0]), y1(p[1]), z1(p[2]), r(p[3]), mat(m) {}
    int index(double x, double y, double z) const override {
        x = x - x1;
        y = y - y1;
        z = z - z1;
        if (x * x + y * y + z * z <= r * r) return mat;
        return -1;
    }
};

struct V3 {
    double x, y, z;
};

// Add_Cylinder (:255-286) and Add_Cone (:101-135); the binding hands over the
// far end as start + (dx, dy, dz) (lua_structure.cpp:334-397).
struct Axial : Op {
    bool cone;
    double x1, y1, z1, x2, y2, z2, r, length;
    V3 b;
    int mat;
    Axial(const double *p, int m, bool cone_)
        : cone(cone_), x1(p[0]), y1(p[1]), z1(p[2]), x2(p[0] + p[3]), y2(p[1] + p[4]), z2(p[2] + p[5]), r(p[6]),
          mat(m) {
        b = {x2 - x1, y2 - y1, z2 - z1};
        length = std::sqrt(b.x * b.x + b.y * b.y + b.z * b.z);   // Vector3::norm (vector23.cpp:266-269)
        const double n = std::sqrt(b.x * b.x + b.y * b.y + b.z * b.z);   // Vector3::normalize (:276-283)
        b.x = b.x / n;
        b.y = b.y / n;
        b.z = b.z / n;
    }
    int index(double x, double y, double z) const override {
        if (x < std::min(x1, x2) - r || x > std::max(x1, x2) + r || y < std::min(y1, y2) - r ||
            y > std::max(y1, y2) + r || z < std::min(z1, z2) - r || z > std::max(z1, z2) + r)
            return -1;
        V3 dv{x - x1, y - y1, z - z1};
        const double proj = b.x * dv.x + b.y * dv.y + b.z * dv.z;   // scalar_prod (vector23.cpp:702-705)
        if (proj >= 0 && proj <= length) {
            dv = {dv.x - proj * b.x, dv.y - proj * b.y, dv.z - proj * b.z};
            const double rt = dv.x * dv.x + dv.y * dv.y + dv.z * dv.z;
            if (!cone) {
                if (rt <= r * r) return mat;
            } else {
                double r2 = r * (1.0 - proj / length);
                r2 = r2 * r2;
                if (rt <= r2) return mat;
            }
        }
        return -1;
    }
};

struct Ellipsoid : Op {   // Add_Ellipsoid (:291-313)
    double x1, y1, z1, rx, ry, rz;
    int mat;
    Ellipsoid(const double *p, int m) : x1(p[0]), y1(p[1]), z1(p[2]), rx(p[3]), ry(p[4]), rz(p[5]), mat(m) {}
    int index(double x, double y, double z) const override {
        if (x < x1 - rx || x > x1 + rx || y < y1 - ry || y > y1 + ry || z < z1 - rz || z > z1 + rz) return -1;
        x = (x - x1) / rx;
        y = (y - y1) / ry;
        z = (z - z1) / rz;
        if (x * x + y * y + z * z <= 1.0) return mat;
        return -1;
    }
};

struct Layer : Op {   // Add_Layer (:317-341)
    int type;
    double z1, z2;
    int mat;
    Layer(const double *p, int m) : type((int)p[0]), z1(p[1]), z2(p[2]), mat(m) {}
    int index(double x, double y, double z) const override {
        const double v = type == 0 ? x : (type == 1 ? y : z);
        if (v >= z1 && v < z2) return mat;
        return -1;
    }
};

double modulus(double v, double m) {   // common/mathUT.h:605-611
    while (v < 0) v += m;
    while (v >= m) v -= m;
    return v;
}

// The structure's operation stack with Structure::index(x, y, z,
// restrict_level) (structure.cpp:228-262): flips, folding on looped axes, then
// the operations below restrict_level, last first, each with its images.
struct Stack {
    const aefdtd_structure *s;
    std::vector<std::unique_ptr<Op>> ops;
    int index(double x, double y, double z, int level) const {
        const int px = s->loop[0] != 0, py = s->loop[1] != 0, pz = s->loop[2] != 0;
        if (s->flip[0]) x = s->lx - x;
        if (s->flip[1]) y = s->ly - y;
        if (s->flip[2]) z = s->lz - z;
        if (px) x = modulus(x, s->lx);
        if (py) y = modulus(y, s->ly);
        if (pz) z = modulus(z, s->lz);
        for (int l = level - 1; l >= 0; l--)
            for (int i = -px; i <= px; i++)
                for (int j = -py; j <= py; j++)
                    for (int k = -pz; k <= pz; k++) {
                        const int v = ops[l]->index(x + i * s->lx, y + j * s->ly, z + k * s->lz);
                        if (v != -1) return v;
                    }
        return s->default_material;
    }
};

// Add_Coating (structure_primitives.cpp:45-96, angles as Degree() converts
// them, common/mathUT.cpp:84-88): a point of the origin material (under the
// operations before this one) with a support of another material within
// `thickness` along the deposition direction, and nothing but the origin
// material from it back up to lz.  The loops test their points exactly as
// the reference's do (the first one point past `thickness`).
struct Coating : Op {
    const Stack *parent;
    int level;
    V3 dir;
    double thickness, delta;
    int origin, mat;
    Coating(const double *p, int m, const Stack *parent_, int level_)
        : parent(parent_), level(level_), thickness(p[2]), delta(p[3]), origin((int)p[4]), mat(m) {
        double theta = p[0], phi = p[1];
        theta *= std::numbers::pi / 180.0;
        phi *= std::numbers::pi / 180.0;
        dir = {std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta), -std::cos(theta)};
    }
    int index(double x, double y, double z) const override {
        if (parent->index(x, y, z, level) != origin) return -1;
        const V3 O{x, y, z};
        int i = 0;
        V3 disp = O;
        const double lz = parent->s->lz;
        bool support = false;
        auto nsq = [&](const V3 &d) {
            const double dx = d.x - O.x, dy = d.y - O.y, dz = d.z - O.z;
            return dx * dx + dy * dy + dz * dz;
        };
        while (nsq(disp) <= thickness * thickness && disp.z > 0) {
            const double a = i * delta;
            disp = {O.x + a * dir.x, O.y + a * dir.y, O.z + a * dir.z};
            if (parent->index(disp.x, disp.y, disp.z, level) != origin) {
                support = true;
                break;
            }
            i++;
        }
        if (!support) return -1;
        i = 0;
        disp = O;
        while (disp.z <= lz) {
            const double a = i * delta;
            disp = {O.x - a * dir.x, O.y - a * dir.y, O.z - a * dir.z};
            if (parent->index(disp.x, disp.y, disp.z, level) != origin) return -1;
            i++;
        }
        return mat;
    }
};

}  // namespace

extern "C" int aeoracle_discretize(const aefdtd_structure *s, int Nx, int Ny, int Nz, double Dx, double Dy, double Dz,
                                   uint32_t *matsgrid) {
    if (!s || !matsgrid || Nx < 1 || Ny < 1 || Nz < 1) return AEFDTD_EINVAL;
    Stack st{s, {}};
    auto &ops = st.ops;
    for (int l = 0; l < s->n_ops; l++) {
        const aefdtd_structure_op &q = s->ops[l];
        switch (q.type) {
        case AEFDTD_OP_BLOCK: ops.emplace_back(new Block(q.p, q.mat_index)); break;
        case AEFDTD_OP_SPHERE: ops.emplace_back(new Sphere(q.p, q.mat_index)); break;
        case AEFDTD_OP_CYLINDER: ops.emplace_back(new Axial(q.p, q.mat_index, false)); break;
        case AEFDTD_OP_CONE: ops.emplace_back(new Axial(q.p, q.mat_index, true)); break;
        case AEFDTD_OP_ELLIPSOID: ops.emplace_back(new Ellipsoid(q.p, q.mat_index)); break;
        case AEFDTD_OP_LAYER: ops.emplace_back(new Layer(q.p, q.mat_index)); break;
        case AEFDTD_OP_COATING:
            if (!(q.p[3] > 0) || !(std::cos(q.p[0] * (std::numbers::pi / 180.0)) > 0)) return AEFDTD_EINVAL;
            ops.emplace_back(new Coating(q.p, q.mat_index, &st, l));
            break;
        default: return AEFDTD_EINVAL;
        }
    }
    auto index = [&](double x, double y, double z) { return st.index(x, y, z, (int)ops.size()); };
    for (int i = 0; i < Nx; i++)      // Structure::discretize (structure.cpp:94-113)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++)
                matsgrid[(size_t)i + (size_t)Nx * ((size_t)j + (size_t)Ny * k)] =
                    (uint32_t)index(i * Dx, j * Dy, k * Dz);
    return AEFDTD_OK;
}
