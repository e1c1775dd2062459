// The island build: a fragment of a fractured lattice made into a lattice of
// its own again, to run through the lattice lane once more (re-fracture of a
// piece that was handed to the rigid world). The reference's
// buildFragmentLattice (src/fastlattice/Refracture.cpp:267-393). Host code,
// run once per re-entry; its result is an asset the schedule and state
// builders of this library take, and the device backend then runs.
//
// Every value is computed in the reference's operation order, so the
// sub-lattice (and so everything the lattice lane does with it) is the
// reference's bit for bit.

#include "banjo_lattice.h"
#include "lattice_errors.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace blx {
namespace {

struct D3 {
    double x, y, z;
};
D3 add(const D3 &a, const D3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 sub(const D3 &a, const D3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 scale(const D3 &a, double s) { return {a.x * s, a.y * s, a.z * s}; }
D3 divide(const D3 &a, double s) { return {a.x / s, a.y / s, a.z / s}; }
D3 cross(const D3 &a, const D3 &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(const D3 &a, const D3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 at3(const double *p, size_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
void put3(double *p, size_t i, const D3 &v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

struct Quat {
    double w, x, y, z;
    // v + w t + q x t with t = 2 q x v: a unit quaternion applied to v.
    D3 rotate(const D3 &v) const {
        const D3 q{x, y, z};
        const D3 t = scale(cross(q, v), 2.0);
        return add(add(v, scale(t, w)), cross(q, t));
    }
};
Quat product(const Quat &a, const Quat &b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}

// Row-major 3x3.
struct M3 {
    double m[3][3];
    D3 col(int c) const { return {m[0][c], m[1][c], m[2][c]}; }
    D3 apply(const D3 &v) const {
        return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
    }
};
M3 matrixOf(const Quat &q) {
    M3 r{};
    const D3 e[3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    for (int c = 0; c < 3; ++c) {
        const D3 v = q.rotate(e[c]);
        r.m[0][c] = v.x;
        r.m[1][c] = v.y;
        r.m[2][c] = v.z;
    }
    return r;
}

// The rotation nearest to a (possibly rank-deficient) correlation matrix, by
// the quaternion iteration of Mueller, Bender, Chentanez and Macklin, "A robust
// method to extract the rotational part of deformations" (MIG 2016): turn the
// current rotation about sum_c (R_c x A_c) / (sum_c |R_c . A_c| + 1e-12) until
// that turn is below 1e-14 rad, at most 64 times.
M3 nearestRotation(const M3 &a) {
    Quat q{1.0, 0.0, 0.0, 0.0};
    for (int it = 0; it < 64; ++it) {
        const M3 r = matrixOf(q);
        D3 turn{0.0, 0.0, 0.0};
        double denom = 1.0e-12;
        for (int c = 0; c < 3; ++c) {
            turn = add(turn, cross(r.col(c), a.col(c)));
            denom += std::abs(dot(r.col(c), a.col(c)));
        }
        turn = divide(turn, denom);
        const double angle = std::sqrt(dot(turn, turn));
        if (!(angle > 1.0e-14)) break;
        const D3 axis = divide(turn, angle);
        const double half = 0.5 * angle;
        const double s = std::sin(half);
        q = product(Quat{std::cos(half), axis.x * s, axis.y * s, axis.z * s}, q);
        const double norm = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
        if (!(norm > 0.0)) return matrixOf(Quat{1.0, 0.0, 0.0, 0.0});
        q = Quat{q.w / norm, q.x / norm, q.y / norm, q.z / norm};
    }
    return matrixOf(q);
}

} // namespace
} // namespace blx

using namespace blx;

extern "C" int blx_fragment_lattice(const blx_asset_view *parent, const uint32_t *nodes, uint32_t count,
                                    const double *cell_offset, const blx_fragment_pose *pose,
                                    const double *plastic_extension, const double *plastic_strain,
                                    blx_fragment_lattice_view *out) {
    return guard([&] {
        if (!parent || !out || !pose) throw Invalid("a fragment lattice needs a parent asset, a pose and an output");
        if (count == 0 || !nodes) throw Invalid("a fragment lattice needs at least one cell");
        if (!cell_offset || !parent->reference_position || !parent->mass) throw Invalid("a fragment lattice needs the parent's cells");
        const uint32_t N = parent->node_count, B = parent->bond_count;
        std::vector<uint32_t> cells(nodes, nodes + count);
        for (uint32_t p : cells)
            if (p >= N) throw Invalid("fragment node index out of range");
        // Sorted, so the sub-lattice depends on the fragment's membership, not
        // on the order its cells were found in.
        std::sort(cells.begin(), cells.end());
        cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
        const uint32_t n = uint32_t(cells.size());
        constexpr uint32_t kOut = std::numeric_limits<uint32_t>::max();
        std::vector<uint32_t> local(N, kOut);
        for (uint32_t i = 0; i < n; ++i) local[cells[i]] = i;
        std::vector<uint32_t> bonds;
        for (uint32_t o = 0; o < B; ++o)
            if (local[parent->bond_a[o]] != kOut && local[parent->bond_b[o]] != kOut) bonds.push_back(o);
        const uint32_t k = uint32_t(bonds.size());
        out->node_count = n;
        out->bond_count = k;
        if (out->node_capacity < n || out->bond_capacity < k) return;  // sizing call

        // The rest shape's mass centre and its rotation inside the frozen fragment.
        D3 centroid{0.0, 0.0, 0.0};
        double total = 0.0;
        for (uint32_t p : cells) {
            total += parent->mass[p];
            centroid = add(centroid, scale(at3(parent->reference_position, p), parent->mass[p]));
        }
        if (!(total > 0.0)) throw Runtime("a fragment lattice has no mass");
        centroid = divide(centroid, total);
        M3 corr{};
        for (uint32_t p : cells) {
            const D3 rest = sub(at3(parent->reference_position, p), centroid);
            const D3 frozen = at3(cell_offset, p);
            const double m = parent->mass[p];
            const double f[3] = {frozen.x, frozen.y, frozen.z}, r[3] = {rest.x, rest.y, rest.z};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) corr.m[a][b] += m * f[a] * r[b];
        }
        const M3 turn = nearestRotation(corr);

        const Quat q{pose->orientation_wxyz[0], pose->orientation_wxyz[1], pose->orientation_wxyz[2],
                     pose->orientation_wxyz[3]};
        const D3 com{pose->center_of_mass[0], pose->center_of_mass[1], pose->center_of_mass[2]};
        const D3 vel{pose->linear_velocity[0], pose->linear_velocity[1], pose->linear_velocity[2]};
        const D3 omega{pose->angular_velocity[0], pose->angular_velocity[1], pose->angular_velocity[2]};
        put3(out->origin, 0, com);
        put3(out->carried_spin, 0, omega);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t p = cells[i];
            out->parent_node[i] = p;
            const D3 pos = add(com, q.rotate(at3(cell_offset, p)));
            const D3 arm = sub(pos, com);
            put3(out->reference_position, i,
                 add(com, q.rotate(turn.apply(sub(at3(parent->reference_position, p), centroid)))));
            put3(out->position, i, pos);
            put3(out->velocity, i, add(vel, cross(omega, arm)));
            out->mass[i] = parent->mass[p];
        }
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t o = bonds[j];
            out->parent_bond[j] = o;
            out->bond_a[j] = local[parent->bond_a[o]];
            out->bond_b[j] = local[parent->bond_b[o]];
            out->rest_length[j] = parent->rest_length[o];
            out->compliance[j] = parent->compliance[o];
            for (int t = 0; t < 6; ++t) out->threshold[6 * size_t(j) + t] = parent->threshold[6 * size_t(o) + t];
            out->alive[j] = parent->alive ? parent->alive[o] : 1;
            out->damage[j] = parent->damage ? parent->damage[o] : 0.0;
            out->failure_mode[j] = parent->failure_mode ? parent->failure_mode[o] : 0;
            out->plastic_extension[j] = plastic_extension ? plastic_extension[o] : 0.0;
            out->plastic_strain[j] = plastic_strain ? plastic_strain[o] : 0.0;
        }
    });
}
