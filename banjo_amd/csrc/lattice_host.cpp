// Host-side lattice build for the MI355X lattice step: the schedule, the
// schedule-order state, the Moore box stencil, slab splits and host fragment
// labelling. Exported through include/banjo_lattice.h.
//
//   blx_schedule_build   buildLatticeSchedule  (src/fastlattice/LatticeSchedule.cpp:18-104)
//   blx_state_build      buildLatticeState     (src/fastlattice/FastLattice.cpp:47-169)
//   blx_box26_bonds      generateBoxTileLattice conventions, Moore stencil
//                        (src/matter/BoxLattice.cpp:49-171)
//   blx_box_slabs        boxLatticeSlabs       (src/matter/BoxLattice.cpp:175-195)
//   blx_components_host  findConnectedComponents (src/fracture/ConnectedComponents.cpp:51-87)
//
// The colouring is greedy in asset bond order with a 64-bit used-colour mask
// per node and the lowest free colour, so the sweep order (and therefore every
// bit of the result) is the reference's. The stable sort by (owner slab,
// boundary, colour) is done as a counting sort over that composite key, which
// is stable by construction and linear in the bond count.

#include "banjo_lattice.h"
#include "lattice_errors.hpp"

#include <algorithm>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace blx {

namespace {

struct ScheduleBuild {
    std::uint32_t colors = 0, boundary = 0;
    std::vector<std::uint8_t> color, is_boundary;
    std::vector<std::uint32_t> owner;
};

ScheduleBuild colour(const blx_asset_view &a, const std::uint32_t *nbb, std::uint32_t blocks) {
    const std::uint32_t N = a.node_count, B = a.bond_count;
    if (blocks < 1) throw Invalid("lattice schedule needs at least one block");
    if (nbb[0] != 0 || nbb[blocks] != N) throw Invalid("lattice schedule needs node block offsets covering every node");
    for (std::uint32_t k = 0; k < blocks; ++k)
        if (nbb[k] > nbb[k + 1]) throw Invalid("lattice schedule needs sorted node block offsets");
    std::vector<std::uint32_t> block_of(N);
    for (std::uint32_t k = 0; k < blocks; ++k)
        std::fill(block_of.begin() + nbb[k], block_of.begin() + nbb[k + 1], k);
    ScheduleBuild s;
    s.color.resize(B);
    s.is_boundary.resize(B);
    s.owner.resize(B);
    std::vector<std::uint64_t> used(N, 0ULL);
    for (std::uint32_t j = 0; j < B; ++j) {
        const std::uint32_t na = a.bond_a[j], nb = a.bond_b[j];
        if (na >= N || nb >= N || na == nb) throw Invalid("lattice bond joins invalid nodes");
        const std::uint64_t taken = used[na] | used[nb];
        if (taken == ~0ULL) throw Runtime("lattice needs more than 64 bond colours");
        const unsigned c = static_cast<unsigned>(std::countr_one(taken));
        used[na] |= 1ULL << c;
        used[nb] |= 1ULL << c;
        s.color[j] = static_cast<std::uint8_t>(c);
        s.colors = std::max(s.colors, c + 1U);
        const std::uint32_t ba = block_of[na], bb = block_of[nb];
        const std::uint32_t lo = ba < bb ? ba : bb, hi = ba < bb ? bb : ba;
        if (hi - lo > 1) throw Invalid("slab decomposition needs slabs at least one neighbour horizon thick");
        s.owner[j] = lo;
        s.is_boundary[j] = hi != lo;
        s.boundary += hi != lo;
    }
    return s;
}

} // namespace

void scheduleBuild(const blx_asset_view &a, const std::uint32_t *nbb, std::uint32_t blocks, blx_schedule_view &out) {
    const std::uint32_t B = a.bond_count;
    const ScheduleBuild s = colour(a, nbb, blocks);
    const std::uint32_t cc = std::max(s.colors, 1U);
    const std::size_t keys = std::size_t(blocks) * 2 * cc;
    // Counting sort on key = (owner * 2 + boundary) * colours + colour.
    std::vector<std::uint32_t> count(keys + 1, 0);
    auto key_of = [&](std::uint32_t j) { return (std::size_t(s.owner[j]) * 2 + s.is_boundary[j]) * cc + s.color[j]; };
    for (std::uint32_t j = 0; j < B; ++j) ++count[key_of(j) + 1];
    for (std::size_t k = 0; k < keys; ++k) count[k + 1] += count[k];
    std::vector<std::uint32_t> cursor(count.begin(), count.end() - 1);
    for (std::uint32_t j = 0; j < B; ++j) {
        const std::uint32_t k = cursor[key_of(j)]++;
        out.bond_order[k] = j;
        out.bond_schedule_index[j] = k;
        out.bond_color[k] = s.color[j];
    }
    // Non-empty ranges are [count[k], count[k+1]); empty ones are [0, 0).
    for (std::size_t k = 0; k < keys; ++k) {
        const bool empty = count[k] == count[k + 1];
        out.range_begin[k] = empty ? 0U : count[k];
        out.range_end[k] = empty ? 0U : count[k + 1];
    }
    // bond_block_begin[b + 1] is the end of owner b's bonds.
    out.bond_block_begin[0] = 0;
    for (std::uint32_t b = 0; b < blocks; ++b) out.bond_block_begin[b + 1] = count[std::size_t(b + 1) * 2 * cc];
    std::copy(nbb, nbb + blocks + 1, out.node_block_begin);
    out.node_count = a.node_count;
    out.bond_count = B;
    out.block_count = blocks;
    out.color_count = s.colors;
    out.boundary_bond_count = s.boundary;
}

std::uint32_t maxDegree(const blx_asset_view &a) {
    std::vector<std::uint32_t> deg(a.node_count, 0);
    for (std::uint32_t j = 0; j < a.bond_count; ++j) {
        if (a.bond_a[j] >= a.node_count || a.bond_b[j] >= a.node_count) throw Invalid("lattice bond joins invalid nodes");
        ++deg[a.bond_a[j]];
        ++deg[a.bond_b[j]];
    }
    std::uint32_t m = 0;
    for (const std::uint32_t d : deg) m = std::max(m, d);
    return m;
}

void stateBuild(const blx_asset_view &a, const blx_schedule_view &sch, const double origin[3], blx_state_view &o) {
    const std::uint32_t N = a.node_count, B = a.bond_count;
    if (sch.bond_count != B || sch.node_block_begin[sch.block_count] != N)
        throw Invalid("lattice schedule does not match the lattice");
    for (int k = 0; k < 3; ++k) o.origin[k] = origin[k];
    o.node_count = N;
    o.bond_count = B;
    for (std::uint32_t i = 0; i < N; ++i) {
        for (int k = 0; k < 3; ++k) {
            const double ref = a.reference_position[3 * std::size_t(i) + k];
            const double u = a.position[3 * std::size_t(i) + k] - ref;
            o.x0[3 * std::size_t(i) + k] = ref - origin[k];
            o.u[3 * std::size_t(i) + k] = u;
            o.u_prev[3 * std::size_t(i) + k] = u;
            o.v[3 * std::size_t(i) + k] = a.velocity[3 * std::size_t(i) + k];
        }
        o.mass[i] = a.mass[i];
        o.inv_mass[i] = a.mass[i] > 0.0 ? 1.0 / a.mass[i] : 0.0;
    }
    for (std::uint32_t k = 0; k < B; ++k) {
        const std::uint32_t j = sch.bond_order[k];
        const std::uint32_t na = a.bond_a[j], nb = a.bond_b[j];
        o.bond_a[k] = na;
        o.bond_b[k] = nb;
        double e[3];
        for (int q = 0; q < 3; ++q) {
            e[q] = a.reference_position[3 * std::size_t(nb) + q] - a.reference_position[3 * std::size_t(na) + q];
            o.rest_edge[3 * std::size_t(k) + q] = e[q];
        }
        const double L = a.rest_length[j];
        const double e2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        o.rest_length[k] = L;
        o.rest_length_sq_minus[k] = e2 - L * L;
        o.weight[k] = e2 > 1.0e-18 ? 1.0 / e2 : 0.0;
        o.compliance[k] = a.compliance[j];
        for (int q = 0; q < 6; ++q) o.threshold[6 * std::size_t(k) + q] = a.threshold[6 * std::size_t(j) + q];
        o.alive[k] = a.alive ? (a.alive[j] ? 1 : 0) : 1;
        o.failure_mode[k] = a.failure_mode ? a.failure_mode[j] : 0;
        o.damage[k] = a.damage ? a.damage[j] : 0.0;
        o.plastic_extension[k] = 0.0;
        o.plastic_strain[k] = 0.0;
        o.prev_tensile[k] = 0.0;
        o.prev_compressive[k] = 0.0;
        o.prev_shear[k] = 0.0;
    }
    // Adjacency in asset bond order, holding schedule indices.
    std::fill(o.adj_offsets, o.adj_offsets + N + 1, 0U);
    for (std::uint32_t j = 0; j < B; ++j) {
        ++o.adj_offsets[a.bond_a[j] + 1];
        ++o.adj_offsets[a.bond_b[j] + 1];
    }
    for (std::uint32_t i = 0; i < N; ++i) o.adj_offsets[i + 1] += o.adj_offsets[i];
    std::vector<std::uint32_t> fill(o.adj_offsets, o.adj_offsets + N);
    for (std::uint32_t j = 0; j < B; ++j) {
        const std::uint32_t k = sch.bond_schedule_index[j];
        o.adj_bonds[fill[a.bond_a[j]]++] = k;
        o.adj_bonds[fill[a.bond_b[j]]++] = k;
    }
    std::uint32_t D = 0;
    for (std::uint32_t i = 0; i < N; ++i) D = std::max(D, o.adj_offsets[i + 1] - o.adj_offsets[i]);
    o.max_degree = D;
    const std::size_t P = std::size_t(D) * N;
    std::fill(o.nbr_bond, o.nbr_bond + P, BLX_NO_BOND);
    std::fill(o.nbr_other, o.nbr_other + P, 0U);
    std::fill(o.nbr_rest, o.nbr_rest + 3 * P, 0.0);
    std::fill(o.nbr_weight, o.nbr_weight + P, 0.0);
    std::fill(o.nbr_alive, o.nbr_alive + P, std::uint8_t(0));
    for (std::uint32_t i = 0; i < N; ++i) {
        const std::uint32_t first = o.adj_offsets[i], deg = o.adj_offsets[i + 1] - first;
        for (std::uint32_t k = 0; k < deg; ++k) {
            const std::uint32_t bond = o.adj_bonds[first + k];
            const std::size_t slot = std::size_t(k) * N + i;
            const bool is_a = o.bond_a[bond] == i;
            o.nbr_bond[slot] = bond;
            o.nbr_alive[slot] = o.alive[bond];
            o.nbr_other[slot] = is_a ? o.bond_b[bond] : o.bond_a[bond];
            if (is_a) o.bond_slot_a[bond] = static_cast<std::uint32_t>(slot);
            else o.bond_slot_b[bond] = static_cast<std::uint32_t>(slot);
            const double sign = is_a ? 1.0 : -1.0;
            for (int q = 0; q < 3; ++q) o.nbr_rest[3 * slot + q] = sign * o.rest_edge[3 * std::size_t(bond) + q];
            o.nbr_weight[slot] = o.weight[bond];
        }
    }
}

std::uint64_t box26(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, double h, double *pos, std::uint32_t *ba,
                    std::uint32_t *bb) {
    if (nx == 0 || ny == 0 || nz == 0 || !(h > 0.0)) throw Invalid("box lattice needs positive cell counts and size");
    const std::uint64_t total = std::uint64_t(nx) * ny * nz;
    if (total > 0xffffffffULL) throw Invalid("box lattice exceeds 32-bit node indices");
    static const int off[13][3] = {{1, 0, 0},  {-1, 1, 0}, {0, 1, 0},  {1, 1, 0},  {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
                                   {-1, 0, 1}, {0, 0, 1},  {1, 0, 1},  {-1, 1, 1}, {0, 1, 1},   {1, 1, 1}};
    std::uint64_t n = 0;
    for (std::uint32_t z = 0; z < nz; ++z)
        for (std::uint32_t y = 0; y < ny; ++y)
            for (std::uint32_t x = 0; x < nx; ++x) {
                const std::uint32_t i = x + nx * (y + ny * z);
                if (pos) {
                    pos[3 * std::size_t(i)] = (double(x) + 0.5 - 0.5 * nx) * h;
                    pos[3 * std::size_t(i) + 1] = (double(y) + 0.5 - 0.5 * ny) * h;
                    pos[3 * std::size_t(i) + 2] = (double(z) + 0.5 - 0.5 * nz) * h;
                }
                for (const auto &d : off) {
                    const long X = long(x) + d[0], Y = long(y) + d[1], Z = long(z) + d[2];
                    if (X < 0 || Y < 0 || Z < 0 || X >= long(nx) || Y >= long(ny) || Z >= long(nz)) continue;
                    if (ba) {
                        ba[n] = i;
                        bb[n] = static_cast<std::uint32_t>(X + long(nx) * (Y + long(ny) * Z));
                    }
                    ++n;
                }
            }
    return n;
}

std::uint32_t boxSlabs(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, std::uint32_t horizon,
                       std::uint32_t requested, std::uint32_t *out) {
    if (nx == 0 || ny == 0 || nz == 0 || horizon == 0) throw Invalid("slab split needs a non-empty layout and a positive horizon");
    std::uint32_t blocks = std::max(1U, requested);
    blocks = std::min(blocks, std::max(1U, nz / horizon));
    if (out) {
        const std::uint32_t layer = nx * ny;
        for (std::uint32_t b = 0; b <= blocks; ++b)
            out[b] = static_cast<std::uint32_t>((std::uint64_t(nz) * b) / blocks) * layer;
        out[blocks] = layer * nz;
    }
    return blocks;
}

std::uint32_t componentsHost(std::uint32_t N, std::uint32_t B, const std::uint32_t *ba, const std::uint32_t *bb,
                             const std::uint8_t *alive, std::uint32_t *label) {
    // Union by linking the larger root under the smaller: each root ends as its
    // component's smallest node, which is the tie-break of the canonical order.
    std::vector<std::uint32_t> parent(N);
    std::iota(parent.begin(), parent.end(), 0U);
    auto find = [&](std::uint32_t x) {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    };
    for (std::uint32_t j = 0; j < B; ++j) {
        if (!alive[j]) continue;
        if (ba[j] >= N || bb[j] >= N) throw Invalid("lattice bond joins invalid nodes");
        const std::uint32_t ra = find(ba[j]), rb = find(bb[j]);
        if (ra == rb) continue;
        if (ra < rb) parent[rb] = ra;
        else parent[ra] = rb;
    }
    std::vector<std::uint32_t> size(N, 0);
    for (std::uint32_t i = 0; i < N; ++i) ++size[label[i] = find(i)];
    std::vector<std::uint32_t> roots;
    for (std::uint32_t i = 0; i < N; ++i) if (label[i] == i) roots.push_back(i);
    std::stable_sort(roots.begin(), roots.end(), [&](std::uint32_t l, std::uint32_t r) { return size[l] > size[r]; });
    std::vector<std::uint32_t> rank(N, 0);
    for (std::uint32_t k = 0; k < roots.size(); ++k) rank[roots[k]] = k;
    for (std::uint32_t i = 0; i < N; ++i) label[i] = rank[label[i]];
    return static_cast<std::uint32_t>(roots.size());
}

// The fastest nodal frequency bounds the stable substep at 2 / omega
// (FastLattice.cpp:30-45); stiffnesses sum in adjacency order.
double substepLimit(const blx_state_view &s) {
    if (!s.adj_offsets) return 0.0;
    double omega = 0.0;
    for (std::uint32_t i = 0; i < s.node_count; ++i) {
        const double m = s.mass[i];
        if (!(m > 0.0)) continue;
        double k = 0.0;
        for (std::uint32_t q = s.adj_offsets[i]; q < s.adj_offsets[i + 1]; ++q) {
            const std::uint32_t j = s.adj_bonds[q];
            if (j >= s.bond_count) throw Invalid("adjacency names a bond out of range");
            if (s.alive[j] && s.compliance[j] > 0.0) k += 1.0 / s.compliance[j];
        }
        if (k > 0.0) omega = std::max(omega, std::sqrt(k / m));
    }
    return omega > 0.0 ? 2.0 / omega : 0.0;
}

// Stored elastic energy of the live bonds in state order (FastLattice.cpp:199-214):
// the bond vector is formed as ((x0_b + u_b) - x0_a) - u_a per axis.
double elasticEnergy(const blx_state_view &s) {
    double sum = 0.0;
    for (std::uint32_t j = 0; j < s.bond_count; ++j) {
        const double c = s.compliance[j];
        if (!s.alive[j] || !(c > 0.0)) continue;
        const std::uint32_t a = s.bond_a[j], b = s.bond_b[j];
        if (a >= s.node_count || b >= s.node_count) throw Invalid("bond end out of range");
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = s.x0[3 * b + k] + s.u[3 * b + k] - s.x0[3 * a + k] - s.u[3 * a + k];
        const double stretch = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - s.rest_length[j] -
                               s.plastic_extension[j];
        sum += 0.5 * stretch * stretch / c;
    }
    return sum;
}

// Kinetic energy in node order (FastLattice.cpp:216-224).
double kineticEnergy(const blx_state_view &s) {
    double sum = 0.0;
    for (std::uint32_t i = 0; i < s.node_count; ++i) {
        const double *v = s.v + 3 * size_t(i);
        sum += 0.5 * s.mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    return sum;
}

} // namespace blx

extern "C" {

int blx_schedule_build(const blx_asset_view *asset, const uint32_t *nbb, uint32_t blocks, blx_schedule_view *out) {
    return blx::guard([&] {
        if (!asset || !nbb || !out) throw blx::Invalid("null argument");
        blx::scheduleBuild(*asset, nbb, blocks, *out);
    });
}

int blx_state_max_degree(const blx_asset_view *asset, uint32_t *d) {
    return blx::guard([&] {
        if (!asset || !d) throw blx::Invalid("null argument");
        *d = blx::maxDegree(*asset);
    });
}

int blx_state_build(const blx_asset_view *asset, const blx_schedule_view *schedule, const double origin[3],
                    blx_state_view *out) {
    return blx::guard([&] {
        if (!asset || !schedule || !origin || !out) throw blx::Invalid("null argument");
        blx::stateBuild(*asset, *schedule, origin, *out);
    });
}

int blx_box26_bonds(uint32_t nx, uint32_t ny, uint32_t nz, double cell, double *pos, uint32_t *ba, uint32_t *bb,
                    uint64_t *count) {
    return blx::guard([&] { *count = blx::box26(nx, ny, nz, cell, pos, ba, bb); });
}

int blx_box_slabs(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t horizon, uint32_t requested, uint32_t *out,
                  uint32_t *block_count) {
    return blx::guard([&] { *block_count = blx::boxSlabs(nx, ny, nz, horizon, requested, out); });
}

int blx_components_host(uint32_t N, uint32_t B, const uint32_t *ba, const uint32_t *bb, const uint8_t *alive,
                        uint32_t *label, uint32_t *count) {
    return blx::guard([&] { *count = blx::componentsHost(N, B, ba, bb, alive, label); });
}

int blx_state_substep_limit(const blx_state_view *state, double *seconds) {
    return blx::guard([&] {
        if (!state || !seconds) throw blx::Invalid("null argument");
        *seconds = blx::substepLimit(*state);
    });
}

int blx_state_elastic_energy(const blx_state_view *state, double *joules) {
    return blx::guard([&] {
        if (!state || !joules) throw blx::Invalid("null argument");
        *joules = blx::elasticEnergy(*state);
    });
}

int blx_state_kinetic_energy(const blx_state_view *state, double *joules) {
    return blx::guard([&] {
        if (!state || !joules) throw blx::Invalid("null argument");
        *joules = blx::kineticEnergy(*state);
    });
}

} // extern "C"
