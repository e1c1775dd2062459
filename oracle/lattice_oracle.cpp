// ORACLE -- TEST INFRASTRUCTURE ONLY. Never linked into, loaded by or called
// from the product library. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it, and only as the checker.
//
// A serial CPU restatement of the fast explicit lattice lane of the reference
// (src/fastlattice), written from the reference's published behaviour, in both
// precisions. Every element function keeps the reference's expression trees
// operation for operation (no FMA contraction: built with -ffp-contract=off),
// so the float path is the bit-exact yardstick for the HIP kernels and the
// double path is the FP64 oracle positions are reported against.
//
// Reference map (file:line in the reference tree):
//   schedule          LatticeSchedule.cpp:18-104
//   state build       FastLattice.cpp:47-169
//   working copy      LatticeWorking.hpp:53-142, 236-288
//   element physics   LatticePhysics.hpp (per function below)
//   substep           CpuLatticeBackend.cpp:357-488
//   run loop          CpuLatticeBackend.cpp:255-304
//   external forces   ExternalLoads.hpp:28-50, 79-148
//   fragments         fracture/ConnectedComponents.cpp:14-87
//   fragment mass     fracture/FragmentMassProperties.cpp:7-80, Mat3::inverse
//                     (core/Math.hpp:110-127)
//   node contact      LatticePhysics.hpp:1430-1667, CpuLatticeBackend.cpp:424-437
//   box stencil       matter/BoxLattice.cpp:49-171 (Moore stencil in place of
//                     the Euclidean horizon)
//
// Not restated (the entry points refuse them): plate bending, the Verlet
// integrator, wrenches, reversible trials.

#include "../include/banjo_lattice.h"

#include <algorithm>
#include <bit>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using u8 = std::uint8_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;
constexpr u32 kNoBond = 0xffffffffU;
constexpr u32 kCap = BLX_MAX_CANDIDATES_PER_BLOCK;

struct invalid : std::invalid_argument { using std::invalid_argument::invalid_argument; };
struct overflow : std::overflow_error { using std::overflow_error::overflow_error; };
struct logic : std::logic_error { using std::logic_error::logic_error; };

// ---------------------------------------------------------------------------
// Vector algebra in the reference's expression order (LatticePhysics.hpp:39-98).
template <class R> struct V3 { R x, y, z; };
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R length2(V3<R> a) { return dot(a, a); }
template <class R> inline R length(V3<R> a) { return std::sqrt(length2(a)); }
template <class R> inline R absR(R x) { return x < R(0) ? -x : x; }
template <class R> inline R maxR(R a, R b) { return (a < b) ? b : a; }
template <class R> inline R minR(R a, R b) { return (b < a) ? b : a; }
template <class R> inline V3<R> normalized(V3<R> a) {
    const R m = length(a);
    return m > R(1.0e-12) ? a / m : V3<R>{R(1), R(0), R(0)};
}
template <class R> inline V3<R> ld3(const std::vector<R> &p, std::size_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
template <class R> inline void st3(std::vector<R> &p, std::size_t i, V3<R> v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
template <class R> inline V3<R> cv(const double *p) { return {R(p[0]), R(p[1]), R(p[2])}; }
using V3d = V3<double>;
template <class R> inline bool fin3(V3<R> v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
template <class R> inline V3d widen(V3<R> v) { return {double(v.x), double(v.y), double(v.z)}; }

// ---------------------------------------------------------------------------
// Schedule (LatticeSchedule.cpp:18-104): greedy edge colouring in asset bond
// order with 64-bit colour masks, then a stable sort by (owner slab, boundary,
// colour) and the range table.
struct Schedule {
    u32 node_count{}, bond_count{}, block_count{1}, color_count{}, boundary_bond_count{};
    std::vector<u32> node_block_begin, bond_order, bond_schedule_index, bond_block_begin, range_begin, range_end;
    std::vector<u8> bond_color;
};

inline Schedule buildSchedule(u32 N, u32 B, const u32 *ba, const u32 *bb, const u32 *nbb, u32 blocks) {
    if (blocks < 1 || nbb[0] != 0 || nbb[blocks] != N) throw invalid("schedule needs node block offsets covering every node");
    for (u32 k = 0; k < blocks; ++k) if (nbb[k] > nbb[k + 1]) throw invalid("node block offsets must be sorted");
    Schedule s;
    s.node_count = N; s.bond_count = B; s.block_count = blocks;
    s.node_block_begin.assign(nbb, nbb + blocks + 1);
    std::vector<u32> block_of(N);
    for (u32 k = 0; k < blocks; ++k) for (u32 n = nbb[k]; n < nbb[k + 1]; ++n) block_of[n] = k;
    std::vector<u64> used(N, 0);
    std::vector<u8> color(B), boundary(B);
    std::vector<u32> owner(B);
    u32 colors = 0;
    for (u32 j = 0; j < B; ++j) {
        const u32 a = ba[j], b = bb[j];
        if (a >= N || b >= N || a == b) throw invalid("lattice bond joins invalid nodes");
        const u64 free = ~(used[a] | used[b]);
        if (free == 0) throw std::runtime_error("lattice needs more than 64 bond colours");
        const unsigned c = unsigned(std::countr_zero(free));
        used[a] |= 1ULL << c; used[b] |= 1ULL << c;
        color[j] = u8(c);
        colors = std::max(colors, c + 1U);
        const u32 lo = std::min(block_of[a], block_of[b]), hi = std::max(block_of[a], block_of[b]);
        if (hi - lo > 1) throw invalid("slab decomposition needs slabs at least one neighbour horizon thick");
        owner[j] = lo; boundary[j] = hi != lo;
        s.boundary_bond_count += boundary[j];
    }
    s.color_count = colors;
    s.bond_order.resize(B);
    std::iota(s.bond_order.begin(), s.bond_order.end(), 0U);
    std::stable_sort(s.bond_order.begin(), s.bond_order.end(), [&](u32 l, u32 r) {
        if (owner[l] != owner[r]) return owner[l] < owner[r];
        if (boundary[l] != boundary[r]) return boundary[l] < boundary[r];
        return color[l] < color[r];
    });
    s.bond_schedule_index.resize(B); s.bond_color.resize(B);
    for (u32 k = 0; k < B; ++k) { s.bond_schedule_index[s.bond_order[k]] = k; s.bond_color[k] = color[s.bond_order[k]]; }
    const u32 cc = std::max(colors, 1U);
    const std::size_t ranges = std::size_t(blocks) * 2 * cc;
    s.range_begin.assign(ranges, 0); s.range_end.assign(ranges, 0);
    s.bond_block_begin.assign(blocks + 1, 0);
    for (u32 k = 0; k < B; ++k) {
        const u32 o = s.bond_order[k];
        const std::size_t key = (std::size_t(owner[o]) * 2 + boundary[o]) * cc + color[o];
        if (s.range_end[key] == 0 && s.range_begin[key] == 0) s.range_begin[key] = k;
        s.range_end[key] = k + 1;
        s.bond_block_begin[owner[o] + 1] = k + 1;
    }
    for (std::size_t key = 0; key < ranges; ++key) if (s.range_end[key] == 0) s.range_begin[key] = 0;
    for (u32 k = 1; k <= blocks; ++k) s.bond_block_begin[k] = std::max(s.bond_block_begin[k], s.bond_block_begin[k - 1]);
    return s;
}

// ---------------------------------------------------------------------------
// State (FastLattice.cpp:47-169), double, schedule order.
struct State {
    double origin[3]{};
    u32 N{}, B{}, D{};
    std::vector<double> x0, u, u_prev, v, inv_mass, mass;
    std::vector<u32> adj_offsets, adj_bonds, nbr_bond, nbr_other, bond_slot_a, bond_slot_b, bond_a, bond_b;
    std::vector<double> nbr_rest, nbr_weight, rest_edge, rest_length, rsm, weight, compliance, threshold;
    std::vector<u8> nbr_alive, alive, failure_mode;
    std::vector<double> damage, prev_t, prev_c, prev_s, plastic_ext, plastic_strain;
};

inline u32 maxDegree(const blx_asset_view &a) {
    std::vector<u32> deg(a.node_count, 0);
    for (u32 j = 0; j < a.bond_count; ++j) { ++deg[a.bond_a[j]]; ++deg[a.bond_b[j]]; }
    u32 m = 0;
    for (u32 d : deg) m = std::max(m, d);
    return m;
}

inline State buildState(const blx_asset_view &a, const Schedule &sch, const double origin[3]) {
    const u32 N = a.node_count, B = a.bond_count;
    if (sch.bond_count != B || sch.node_block_begin.back() != N) throw invalid("lattice schedule does not match the lattice");
    State s;
    for (int k = 0; k < 3; ++k) s.origin[k] = origin[k];
    s.N = N; s.B = B;
    s.x0.resize(3 * std::size_t(N)); s.u.resize(3 * std::size_t(N)); s.u_prev.resize(3 * std::size_t(N)); s.v.resize(3 * std::size_t(N));
    s.inv_mass.resize(N); s.mass.resize(N);
    for (u32 i = 0; i < N; ++i) {
        for (int k = 0; k < 3; ++k) {
            const double ref = a.reference_position[3 * i + k];
            s.x0[3 * i + k] = ref - origin[k];
            s.u[3 * i + k] = a.position[3 * i + k] - ref;
            s.u_prev[3 * i + k] = s.u[3 * i + k];
            s.v[3 * i + k] = a.velocity[3 * i + k];
        }
        s.mass[i] = a.mass[i];
        s.inv_mass[i] = a.mass[i] > 0.0 ? 1.0 / a.mass[i] : 0.0;
    }
    s.bond_a.resize(B); s.bond_b.resize(B); s.rest_edge.resize(3 * std::size_t(B));
    s.rest_length.resize(B); s.rsm.resize(B); s.weight.resize(B); s.compliance.resize(B);
    s.threshold.resize(6 * std::size_t(B)); s.alive.resize(B); s.failure_mode.resize(B); s.damage.resize(B);
    s.plastic_ext.assign(B, 0.0); s.plastic_strain.assign(B, 0.0);
    s.prev_t.assign(B, 0.0); s.prev_c.assign(B, 0.0); s.prev_s.assign(B, 0.0);
    for (u32 k = 0; k < B; ++k) {
        const u32 o = sch.bond_order[k];
        const u32 na = a.bond_a[o], nb = a.bond_b[o];
        s.bond_a[k] = na; s.bond_b[k] = nb;
        const double e[3] = {a.reference_position[3 * nb] - a.reference_position[3 * na],
                             a.reference_position[3 * nb + 1] - a.reference_position[3 * na + 1],
                             a.reference_position[3 * nb + 2] - a.reference_position[3 * na + 2]};
        for (int q = 0; q < 3; ++q) s.rest_edge[3 * k + q] = e[q];
        s.rest_length[k] = a.rest_length[o];
        const double esq = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        s.rsm[k] = esq - a.rest_length[o] * a.rest_length[o];
        s.weight[k] = esq > 1.0e-18 ? 1.0 / esq : 0.0;
        s.compliance[k] = a.compliance[o];
        for (int q = 0; q < 6; ++q) s.threshold[6 * k + q] = a.threshold[6 * std::size_t(o) + q];
        s.alive[k] = a.alive ? (a.alive[o] ? 1 : 0) : 1;
        s.failure_mode[k] = a.failure_mode ? a.failure_mode[o] : 0;
        s.damage[k] = a.damage ? a.damage[o] : 0.0;
    }
    std::vector<u32> deg(N, 0);
    for (u32 o = 0; o < B; ++o) { ++deg[a.bond_a[o]]; ++deg[a.bond_b[o]]; }
    s.adj_offsets.assign(N + 1, 0);
    for (u32 i = 0; i < N; ++i) s.adj_offsets[i + 1] = s.adj_offsets[i] + deg[i];
    s.adj_bonds.resize(s.adj_offsets.back());
    std::vector<u32> cur(s.adj_offsets.begin(), s.adj_offsets.end());
    for (u32 o = 0; o < B; ++o) {
        const u32 k = sch.bond_schedule_index[o];
        s.adj_bonds[cur[a.bond_a[o]]++] = k;
        s.adj_bonds[cur[a.bond_b[o]]++] = k;
    }
    s.D = 0;
    for (u32 i = 0; i < N; ++i) s.D = std::max(s.D, s.adj_offsets[i + 1] - s.adj_offsets[i]);
    const std::size_t P = std::size_t(s.D) * N;
    s.nbr_bond.assign(P, kNoBond); s.nbr_other.assign(P, 0); s.nbr_rest.assign(3 * P, 0.0);
    s.nbr_weight.assign(P, 0.0); s.nbr_alive.assign(P, 0);
    s.bond_slot_a.assign(B, 0); s.bond_slot_b.assign(B, 0);
    for (u32 i = 0; i < N; ++i) {
        for (u32 k = 0; k < s.adj_offsets[i + 1] - s.adj_offsets[i]; ++k) {
            const u32 bond = s.adj_bonds[s.adj_offsets[i] + k];
            const std::size_t slot = std::size_t(k) * N + i;
            s.nbr_bond[slot] = bond;
            s.nbr_alive[slot] = s.alive[bond];
            const bool is_a = s.bond_a[bond] == i;
            (is_a ? s.bond_slot_a : s.bond_slot_b)[bond] = u32(slot);
            s.nbr_other[slot] = is_a ? s.bond_b[bond] : s.bond_a[bond];
            const double sign = is_a ? 1.0 : -1.0;
            for (int q = 0; q < 3; ++q) s.nbr_rest[3 * slot + q] = sign * s.rest_edge[3 * bond + q];
            s.nbr_weight[slot] = s.weight[bond];
        }
    }
    return s;
}

// ---------------------------------------------------------------------------
// Settings in backend precision (LatticeWorking.hpp:236-288).
template <class R> struct Plane {
    V3<R> point, normal, tangent, bitangent;
    R restitution, static_friction, dynamic_friction, node_radius;
    u8 reach_capped; u32 footprint_count;
    R fp[BLX_MAX_FOOTPRINTS][4];
};
template <class R> struct Settings {
    R dt; V3<R> gravity; u32 iterations; R damping; u8 sphere_enabled, direct;
    R yield, hardening;
    R c_sf, c_df, c_rest, c_rest_thr, c_node_radius, c_margin, c_slack;
    u32 plane_count; Plane<R> planes[BLX_MAX_SUPPORT_PLANES];
};
template <class R> struct Sphere { V3<R> center, velocity, omega; R radius, mass, inertia; };
// The kinematic blade of the HIP library's extension (include/banjo_lattice.h,
// blx_blade): no reference counterpart, restated here so that its use is
// checked against a serial lane.
template <class R> struct Blade { bool enabled; V3<R> c0, f, m; R speed, hw, r; };

template <class R> Settings<R> convertSettings(const blx_settings &s) {
    Settings<R> o{};
    o.dt = R(s.dt); o.gravity = cv<R>(s.gravity); o.iterations = s.constraint_iterations;
    o.damping = R(s.damping_fraction); o.sphere_enabled = s.sphere_enabled; o.direct = s.direct_arithmetic;
    o.yield = R(s.plastic_yield_stretch); o.hardening = R(s.plastic_hardening);
    o.c_sf = R(s.contact.static_friction); o.c_df = R(s.contact.dynamic_friction);
    o.c_rest = R(s.contact.restitution); o.c_rest_thr = R(s.contact.restitution_speed_threshold);
    o.c_node_radius = R(s.contact.node_contact_radius); o.c_margin = R(s.contact.contact_margin);
    o.c_slack = R(s.contact.prefilter_slack);
    o.plane_count = s.plane_count;
    for (unsigned p = 0; p < BLX_MAX_SUPPORT_PLANES; ++p) {
        const blx_support_plane &in = s.planes[p];
        Plane<R> &pl = o.planes[p];
        pl.point = cv<R>(in.point); pl.normal = cv<R>(in.normal); pl.tangent = cv<R>(in.tangent); pl.bitangent = cv<R>(in.bitangent);
        pl.restitution = R(in.restitution); pl.static_friction = R(in.static_friction); pl.dynamic_friction = R(in.dynamic_friction);
        pl.node_radius = R(in.node_radius); pl.reach_capped = in.reach_capped; pl.footprint_count = in.footprint_count;
        for (unsigned f = 0; f < BLX_MAX_FOOTPRINTS; ++f) {
            pl.fp[f][0] = R(in.footprints[f].center_t); pl.fp[f][1] = R(in.footprints[f].center_b);
            pl.fp[f][2] = R(in.footprints[f].half_t); pl.fp[f][3] = R(in.footprints[f].half_b);
        }
    }
    return o;
}

// ---------------------------------------------------------------------------
// The serial backend in precision R.
template <class R>
class Backend {
public:
    explicit Backend(Schedule s) : sch_(std::move(s)) {}

    void upload(const blx_state_view &st, const blx_settings &settings, const blx_sphere &sp) {
        if (settings.bond_integrator != BLX_BOND_XPBD) throw invalid("the oracle restates the XPBD integrator only");
        if (settings.node_contact.mode > BLX_NODE_CONTACT_ON) throw invalid("unknown node contact mode");
        if (settings.node_contact.mode != BLX_NODE_CONTACT_OFF &&
            (settings.node_contact.bucket_mask & (settings.node_contact.bucket_mask + 1U)) != 0U)
            throw invalid("node contact bucket_mask must be a power of two minus one");
        if (settings.plate_half_thickness != 0.0) throw invalid("the oracle does not restate plate bending");
        if (settings.plane_count > BLX_MAX_SUPPORT_PLANES) throw invalid("too many support planes");
        if (st.node_count != sch_.node_count || st.bond_count != sch_.bond_count) throw invalid("state does not match the schedule");
        N = st.node_count; B = st.bond_count; D = st.max_degree;
        const std::size_t n3 = 3 * std::size_t(N), P = std::size_t(D) * N;
        auto cast = [](const double *p, std::size_t n) { std::vector<R> o(n); for (std::size_t i = 0; i < n; ++i) o[i] = R(p[i]); return o; };
        x0 = cast(st.x0, n3); u = cast(st.u, n3); u_prev = cast(st.u_prev, n3); v = cast(st.v, n3);
        inv_mass = cast(st.inv_mass, N); mass = cast(st.mass, N);
        rinv.assign(9 * std::size_t(N), R(0)); unmeasured.assign(n3, R(0)); strain.assign(6 * std::size_t(N), R(0));
        approach.assign(2 * std::size_t(N), R(0));
        node_valid.assign(N, 0); node_dirty.assign(N, 1); engaged.assign(2 * std::size_t(N), 0); candidate.assign(N, 0);
        nbr_bond.assign(st.nbr_bond, st.nbr_bond + P); nbr_other.assign(st.nbr_other, st.nbr_other + P);
        nbr_rest = cast(st.nbr_rest, 3 * P); nbr_weight = cast(st.nbr_weight, P); nbr_alive.assign(st.nbr_alive, st.nbr_alive + P);
        slot_a.assign(st.bond_slot_a, st.bond_slot_a + B); slot_b.assign(st.bond_slot_b, st.bond_slot_b + B);
        bond_a.assign(st.bond_a, st.bond_a + B); bond_b.assign(st.bond_b, st.bond_b + B);
        rest_edge = cast(st.rest_edge, 3 * std::size_t(B)); rest_length = cast(st.rest_length, B);
        rsm = cast(st.rest_length_sq_minus, B); compliance = cast(st.compliance, B); threshold = cast(st.threshold, 6 * std::size_t(B));
        alive.assign(st.alive, st.alive + B); failure_mode.assign(st.failure_mode, st.failure_mode + B);
        damage = cast(st.damage, B); lambda.assign(B, R(0));
        plastic_ext = cast(st.plastic_extension, B); plastic_strain = cast(st.plastic_strain, B);
        prev_t = cast(st.prev_tensile, B); prev_c = cast(st.prev_compressive, B); prev_s = cast(st.prev_shear, B);
        cand_list.assign(std::size_t(sch_.block_count) * kCap, 0); cand_count.assign(sch_.block_count, 0);
        rank_deficient = 0;
        S = convertSettings<R>(settings);
        sphere = {cv<R>(sp.center), cv<R>(sp.velocity), cv<R>(sp.angular_velocity), R(sp.radius), R(sp.mass), R(sp.inertia)};
        std::memset(&status, 0, sizeof(status));
        status.first_failure_step = status.last_failure_step = status.last_energy_gain_step = std::numeric_limits<u64>::max();
        for (int k = 0; k < 3; ++k) origin[k] = st.origin[k];
        loads.clear(); load_remaining = 0;
        dirty_start = true;
        const blx_node_contact_settings &nc = settings.node_contact;
        NC = {nc.mode, R(nc.radius), R(nc.skin), R(nc.margin), R(nc.restitution), R(nc.restitution_speed_threshold),
              R(nc.static_friction), R(nc.dynamic_friction), nc.bucket_mask};
        if (NC.mode != BLX_NODE_CONTACT_OFF) {
            const std::size_t buckets = std::size_t(NC.mask) + 1;
            node_cell.assign(n3, 0); build_u.assign(n3, R(0));
            bucket_begin.assign(buckets + 1, 0U); bucket_cursor.assign(buckets, 0U); bucket_nodes.assign(N, 0U);
            pair_other.assign(std::size_t(kPairs) * N, 0U); pair_bond.assign(std::size_t(kPairs) * N, kNoBond);
            pair_fill.assign(N, 0U); pair_node_list.assign(N, 0U); pair_node_count = 0;
        }
        contact_rebuild = true;
        first_failure.clear(); frames.clear();
        uploaded = true;
    }

    // ExternalLoads.hpp:28-50: validate everything, then replace.
    void setBlade(const blx_blade &b) {
        Blade<R> o{};
        o.enabled = b.enabled != 0;
        if (o.enabled) {
            o.c0 = cv<R>(b.edge_point);
            o.f = cv<R>(b.flat_normal);
            o.m = cv<R>(b.direction);
            o.speed = R(b.speed);
            o.hw = R(b.half_width);
            o.r = R(b.node_radius);
        }
        blade = o;
    }

    void setForces(const double *f, u64 substeps) {
        if (!uploaded) throw logic("external forces before upload");
        if (f == nullptr) {
            if (substeps != 0) throw invalid("empty external load needs zero substeps");
            loads.clear(); load_remaining = 0; return;
        }
        if (substeps == 0) throw invalid("external load needs one force per node and a finite nonzero span");
        std::vector<Load> staged;
        for (u32 i = 0; i < N; ++i) {
            const V3d fd{f[3 * i], f[3 * i + 1], f[3 * i + 2]};
            const V3<R> fr{R(fd.x), R(fd.y), R(fd.z)};
            if (!fin3(fd) || !std::isfinite(std::hypot(fd.x, fd.y, fd.z)) || !fin3(fr))
                throw invalid("external force is not finite in backend precision");
            if (fd.x == 0 && fd.y == 0 && fd.z == 0) continue;
            if (!(mass[i] > 0) || !(inv_mass[i] > 0) || !std::isfinite(mass[i]) || !std::isfinite(inv_mass[i]))
                throw invalid("external force needs a finite positive movable node mass");
            staged.push_back({i, fd, fr});
        }
        loads = std::move(staged);
        load_remaining = loads.empty() ? 0 : substeps;
    }

    void run(const blx_run_control &c) {
        if (!uploaded) throw logic("run before upload");
        const auto start = std::chrono::steady_clock::now();
        flat_fraction = c.energy_flat_fraction;
        status.exit_reason = 0;
        u64 done = 0;
        while (done < c.max_steps) {
            const u64 step = status.total_steps;
            if (c.capture_stride != 0 && step % c.capture_stride == 0 && status.frames_captured < c.max_frames) {
                captureFrame(step);
                ++status.frames_captured;
            }
            const bool failed = advance();
            status.total_steps = step + 1;
            ++done;
            if (failed) {
                if (status.first_failure_step == std::numeric_limits<u64>::max()) status.first_failure_step = step;
                status.last_failure_step = step;
                ++status.failure_rounds;
            }
            const double available = c.removable_energy_j + status.external_load.work_j;
            status.exit_reason = c.removable_energy_j > 0 && available <= 0 ? 6U :
                exitReason(status.total_steps, status.broken_bonds, status.last_failure_step, c.quiet_steps, c.min_steps,
                           c.no_failure_steps, c.energy_flat_steps, status.last_energy_gain_step, c.calm_steps,
                           status.last_damage_gain_step, status.max_damage, c.calm_damage_margin, status.removed_energy_j,
                           c.removable_energy_j > 0 ? available : 0.0);
            if (status.exit_reason != 0) break;
        }
        if (status.exit_reason == 0 && done >= c.max_steps) status.exit_reason = 3;
        ++status.launches;
        status.wall_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    }

    void download(blx_state_view &st, blx_sphere &sp) const {
        const std::size_t n3 = 3 * std::size_t(N);
        for (std::size_t k = 0; k < n3; ++k) { st.u[k] = double(u[k]); st.u_prev[k] = double(u_prev[k]); st.v[k] = double(v[k]); }
        for (u32 j = 0; j < B; ++j) {
            st.alive[j] = alive[j]; st.failure_mode[j] = failure_mode[j]; st.damage[j] = double(damage[j]);
            st.plastic_extension[j] = double(plastic_ext[j]); st.plastic_strain[j] = double(plastic_strain[j]);
            st.prev_tensile[j] = double(prev_t[j]); st.prev_compressive[j] = double(prev_c[j]); st.prev_shear[j] = double(prev_s[j]);
        }
        const V3d c = widen(sphere.center), vel = widen(sphere.velocity), w = widen(sphere.omega);
        sp.center[0] = c.x; sp.center[1] = c.y; sp.center[2] = c.z;
        sp.velocity[0] = vel.x; sp.velocity[1] = vel.y; sp.velocity[2] = vel.z;
        sp.angular_velocity[0] = w.x; sp.angular_velocity[1] = w.y; sp.angular_velocity[2] = w.z;
        sp.radius = double(sphere.radius); sp.mass = double(sphere.mass); sp.inertia = double(sphere.inertia);
    }

    blx_run_status status{};
    std::vector<u32> first_failure;
    struct Frame { u64 step; std::vector<float> u; std::vector<u8> alive; std::vector<float> damage; Sphere<R> sphere; };
    std::vector<Frame> frames;

private:
    struct Load { u32 index; V3d requested; V3<R> applied; };

    Schedule sch_;
    u32 N{}, B{}, D{};
    std::vector<R> x0, u, u_prev, v, inv_mass, mass, rinv, unmeasured, strain, approach;
    std::vector<u8> node_valid, node_dirty, engaged, candidate;
    std::vector<u32> nbr_bond, nbr_other, slot_a, slot_b, bond_a, bond_b, cand_list, cand_count;
    std::vector<R> nbr_rest, nbr_weight, rest_edge, rest_length, rsm, compliance, threshold;
    std::vector<u8> nbr_alive, alive, failure_mode;
    std::vector<R> damage, lambda, plastic_ext, plastic_strain, prev_t, prev_c, prev_s;
    u32 rank_deficient{};
    Settings<R> S{};
    Blade<R> blade{};
    Sphere<R> sphere{};
    double origin[3]{};
    std::vector<Load> loads;
    u64 load_remaining{};
    double flat_fraction{1.0e-3};
    bool dirty_start{true};
    bool uploaded{false};

    // Node-node contact (LatticePhysics.hpp:1430-1650): settings in backend
    // precision, the broad phase's cells, hash and pair lists, and the flag that
    // forces a rebuild after upload and after every failure round.
    static constexpr u32 kPairs = 24;  // kMaxPairsPerNode (LatticePhysics.hpp:179)
    struct NodeContact { u8 mode; R radius, skin, margin, restitution, rest_thr, sf, df; u32 mask; } NC{};
    std::vector<std::int32_t> node_cell;
    std::vector<R> build_u;
    std::vector<u32> bucket_begin, bucket_cursor, bucket_nodes, pair_other, pair_bond, pair_fill, pair_node_list;
    u32 pair_node_count{};
    bool contact_rebuild{true};

    R ncCutoff() const { return R(2) * NC.radius + NC.skin; }
    // Teschner's spatial hash (LatticePhysics.hpp:1447-1456).
    static u32 ncBucket(std::int32_t ix, std::int32_t iy, std::int32_t iz, u32 mask) {
        return ((u32(ix) * 73856093U) ^ (u32(iy) * 19349663U) ^ (u32(iz) * 83492791U)) & mask;
    }
    // Step 1 (LatticePhysics.hpp:1460-1468).
    void ncStoreCell(u32 i) {
        const R inv = R(1) / ncCutoff();
        const V3<R> p = position(i);
        node_cell[3 * std::size_t(i)] = std::int32_t(std::floor(p.x * inv));
        node_cell[3 * std::size_t(i) + 1] = std::int32_t(std::floor(p.y * inv));
        node_cell[3 * std::size_t(i) + 2] = std::int32_t(std::floor(p.z * inv));
        st3(build_u, i, ld3(u, i));
    }
    u32 ncBucketOf(u32 i) const {
        return ncBucket(node_cell[3 * std::size_t(i)], node_cell[3 * std::size_t(i) + 1], node_cell[3 * std::size_t(i) + 2],
                        NC.mask);
    }
    // Step 2 (LatticePhysics.hpp:1471-1489): counting sort into the buckets, in node order.
    void ncHashBuild() {
        const u32 buckets = NC.mask + 1U;
        std::fill(bucket_begin.begin(), bucket_begin.end(), 0U);
        for (u32 i = 0; i < N; ++i) if (mass[i] > R(0)) ++bucket_begin[ncBucketOf(i) + 1U];
        for (u32 b = 0; b < buckets; ++b) bucket_begin[b + 1U] += bucket_begin[b];
        for (u32 b = 0; b < buckets; ++b) bucket_cursor[b] = bucket_begin[b];
        for (u32 i = 0; i < N; ++i) if (mass[i] > R(0)) bucket_nodes[bucket_cursor[ncBucketOf(i)]++] = i;
    }
    // Step 3 (LatticePhysics.hpp:1494-1546): the pairs node i owns, over the 27
    // cells around it, skipping pairs a live bond holds. Returns the overflow.
    u32 ncGather(u32 i) {
        pair_fill[i] = 0U;
        if (mass[i] <= R(0)) return 0U;
        const R cutoff = ncCutoff(), cutoff2 = cutoff * cutoff;
        const V3<R> p = position(i);
        const std::int32_t cx = node_cell[3 * std::size_t(i)], cy = node_cell[3 * std::size_t(i) + 1],
                           cz = node_cell[3 * std::size_t(i) + 2];
        u32 fill = 0, over = 0;
        for (std::int32_t dz = -1; dz <= 1; ++dz)
            for (std::int32_t dy = -1; dy <= 1; ++dy)
                for (std::int32_t dx = -1; dx <= 1; ++dx) {
                    const std::int32_t gx = cx + dx, gy = cy + dy, gz = cz + dz;
                    const u32 b = ncBucket(gx, gy, gz, NC.mask);
                    for (u32 k = bucket_begin[b]; k < bucket_begin[b + 1U]; ++k) {
                        const u32 j = bucket_nodes[k];
                        if (j <= i) continue;
                        if (node_cell[3 * std::size_t(j)] != gx || node_cell[3 * std::size_t(j) + 1] != gy ||
                            node_cell[3 * std::size_t(j) + 2] != gz)
                            continue;
                        if (length2(position(j) - p) > cutoff2) continue;
                        u32 bond = kNoBond;
                        bool held = false;
                        for (u32 q = 0; q < D; ++q) {
                            const std::size_t slot = std::size_t(q) * N + i;
                            if (nbr_bond[slot] == kNoBond) break;
                            if (nbr_other[slot] != j) continue;
                            bond = nbr_bond[slot];
                            held = nbr_alive[slot] != 0;
                            break;
                        }
                        if (held) continue;
                        if (fill >= kPairs) { ++over; continue; }
                        pair_other[std::size_t(kPairs) * i + fill] = j;
                        pair_bond[std::size_t(kPairs) * i + fill] = bond;
                        ++fill;
                    }
                }
        pair_fill[i] = fill;
        return over;
    }
    // Step 4 (LatticePhysics.hpp:1551-1561): the owners in node order; returns the pairs.
    u32 ncActiveList() {
        u32 count = 0, pairs = 0;
        for (u32 i = 0; i < N; ++i) {
            if (pair_fill[i] == 0U) continue;
            pair_node_list[count++] = i;
            pairs += pair_fill[i];
        }
        pair_node_count = count;
        return pairs;
    }
    // LatticePhysics.hpp:1566-1574.
    bool ncStale() const {
        const R half = R(0.5) * NC.skin, limit = half * half;
        for (u32 i = 0; i < N; ++i)
            if (length2(ld3(u, i) - ld3(build_u, i)) > limit) return true;
        return false;
    }
    // One pair's response (LatticePhysics.hpp:1585-1648).
    void ncPair(u32 a, u32 b, u32 bond, blx_node_contact_accumulators &acc) {
        if (bond != kNoBond && alive[bond]) return;
        const R ma = mass[a], mb = mass[b];
        if (ma <= R(0) || mb <= R(0)) return;
        ++acc.pair_tests;
        const V3<R> r = position(b) - position(a);
        const R distance = length(r);
        const V3<R> normal = normalized(r);
        const R gap = distance - R(2) * NC.radius;
        if (gap < R(0) && -gap > R(acc.maximum_overlap_m)) acc.maximum_overlap_m = double(-gap);
        if (NC.mode != BLX_NODE_CONTACT_ON) return;
        V3<R> va = ld3(v, a), vb = ld3(v, b);
        const R vn = dot(vb - va, normal);
        if (gap > maxR(NC.margin, -vn * S.dt)) return;
        const R inverse_a = R(1) / ma, inverse_b = R(1) / mb;
        const R inverse_pair = inverse_a + inverse_b;
        const R restitution = -vn > NC.rest_thr ? NC.restitution : R(0);
        const R desired = gap > NC.margin ? -gap / S.dt : -restitution * minR(vn, R(0));
        const R normal_impulse = maxR(R(0), (desired - vn) / inverse_pair);
        if (normal_impulse > R(0)) {
            const V3<R> va0 = va, vb0 = vb;
            const R before = R(0.5) * ma * length2(va) + R(0.5) * mb * length2(vb);
            auto apply = [&](V3<R> impulse) {
                vb = vb + inverse_b * impulse;
                va = va - inverse_a * impulse;
            };
            apply(normal_impulse * normal);
            const V3<R> relative = vb - va;
            const V3<R> tangent = relative - dot(relative, normal) * normal;
            const R tangent_speed = length(tangent);
            if (tangent_speed > R(1.0e-12)) {
                const V3<R> direction = tangent / tangent_speed;
                const R sticking = tangent_speed / inverse_pair;
                const R friction = sticking <= NC.sf * normal_impulse ? sticking : minR(sticking, NC.df * normal_impulse);
                apply(-friction * direction);
            }
            st3(v, a, va);
            st3(v, b, vb);
            acc.dissipated_kinetic_energy_j += double(before - (R(0.5) * ma * length2(va) + R(0.5) * mb * length2(vb)));
            acc.momentum_residual[0] += double(ma * (va.x - va0.x) + mb * (vb.x - vb0.x));
            acc.momentum_residual[1] += double(ma * (va.y - va0.y) + mb * (vb.y - vb0.y));
            acc.momentum_residual[2] += double(ma * (va.z - va0.z) + mb * (vb.z - vb0.z));
            ++acc.contacts;
        }
        if (gap < -NC.margin) {
            const R correction = minR(-gap - NC.margin, R(0.2) * NC.radius);
            const V3<R> shift = (correction / inverse_pair) * normal;
            st3(u, a, ld3(u, a) - inverse_a * shift);
            st3(u, b, ld3(u, b) + inverse_b * shift);
            if (correction > R(acc.maximum_position_correction_m)) acc.maximum_position_correction_m = double(correction);
        }
    }
    // CpuLatticeBackend.cpp:424-437: rebuild when forced or stale, then the
    // serial narrow phase in pair order (LatticePhysics.hpp:1655-1667).
    void nodeContact() {
        blx_node_contact_accumulators &acc = status.node_contact;
        if (contact_rebuild || ncStale()) {
            for (u32 i = 0; i < N; ++i) ncStoreCell(i);
            ncHashBuild();
            for (u32 i = 0; i < N; ++i) acc.pair_overflow += ncGather(i);
            acc.pairs_listed += ncActiveList();
            ++acc.rebuilds;
            contact_rebuild = false;
        }
        for (u32 k = 0; k < pair_node_count; ++k) {
            const u32 a = pair_node_list[k];
            for (u32 q = 0; q < pair_fill[a]; ++q)
                ncPair(a, pair_other[std::size_t(kPairs) * a + q], pair_bond[std::size_t(kPairs) * a + q], acc);
        }
    }

    V3<R> position(u32 i) const { return ld3(x0, i) + ld3(u, i); }
    V3<R> bondVector(u32 j) const {
        const u32 a = bond_a[j], b = bond_b[j];
        if (S.direct) return position(b) - position(a);
        return ld3(rest_edge, j) + (ld3(u, b) - ld3(u, a));
    }

    void captureFrame(u64 step) {
        Frame f;
        f.step = step;
        f.u.resize(u.size());
        for (std::size_t k = 0; k < u.size(); ++k) f.u[k] = float(u[k]);
        f.alive = alive;
        f.damage.resize(damage.size());
        for (std::size_t k = 0; k < damage.size(); ++k) f.damage[k] = float(damage[k]);
        f.sphere = sphere;
        frames.push_back(std::move(f));
    }

    // LatticePhysics.hpp:456-493.
    static void eigen3(const R m[9], R d[3], R vv[9]) {
        R a[9];
        for (int k = 0; k < 9; ++k) a[k] = m[k];
        for (int k = 0; k < 9; ++k) vv[k] = (k % 4 == 0) ? R(1) : R(0);
        for (int sweep = 0; sweep < 12; ++sweep) {
            const R off = absR(a[1]) + absR(a[2]) + absR(a[5]);
            if (!(off > R(0))) break;
            for (int p = 0; p < 2; ++p) {
                for (int q = p + 1; q < 3; ++q) {
                    const R apq = a[p * 3 + q];
                    if (!(absR(apq) > R(0))) continue;
                    const R app = a[p * 3 + p], aqq = a[q * 3 + q];
                    const R theta = (aqq - app) / (R(2) * apq);
                    const R sign = theta >= R(0) ? R(1) : R(-1);
                    const R t = sign / (absR(theta) + std::sqrt(theta * theta + R(1)));
                    const R c = R(1) / std::sqrt(t * t + R(1));
                    const R sn = t * c;
                    for (int k = 0; k < 3; ++k) {
                        const R akp = a[k * 3 + p], akq = a[k * 3 + q];
                        a[k * 3 + p] = c * akp - sn * akq;
                        a[k * 3 + q] = sn * akp + c * akq;
                    }
                    for (int k = 0; k < 3; ++k) {
                        const R apk = a[p * 3 + k], aqk = a[q * 3 + k];
                        a[p * 3 + k] = c * apk - sn * aqk;
                        a[q * 3 + k] = sn * apk + c * aqk;
                    }
                    for (int k = 0; k < 3; ++k) {
                        const R vkp = vv[k * 3 + p], vkq = vv[k * 3 + q];
                        vv[k * 3 + p] = c * vkp - sn * vkq;
                        vv[k * 3 + q] = sn * vkp + c * vkq;
                    }
                }
            }
        }
        d[0] = a[0]; d[1] = a[4]; d[2] = a[8];
    }

    // LatticePhysics.hpp:446-449, 506-531.
    static int pinv3(const R m[9], R out[9], R un[3]) {
        R d[3], vv[9];
        eigen3(m, d, vv);
        R largest = absR(d[0]);
        if (absR(d[1]) > largest) largest = absR(d[1]);
        if (absR(d[2]) > largest) largest = absR(d[2]);
        const R floor_value = (sizeof(R) == 8 ? R(1.0e-9) : R(1.0e-5)) * largest;
        for (int k = 0; k < 9; ++k) out[k] = R(0);
        for (int k = 0; k < 3; ++k) un[k] = R(0);
        int rank = 0, dropped = -1;
        for (int c = 0; c < 3; ++c) {
            if (!(d[c] > floor_value)) { dropped = c; continue; }
            ++rank;
            const R inv = R(1) / d[c];
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) out[r * 3 + q] += inv * vv[r * 3 + c] * vv[q * 3 + c];
        }
        if (rank == 2 && dropped >= 0) for (int k = 0; k < 3; ++k) un[k] = vv[k * 3 + dropped];
        return rank;
    }

    // LatticePhysics.hpp:792-907 (plate bending excluded: refused at upload).
    void nodeStrain(u32 i) {
        if (node_dirty[i]) {
            R cov[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            u32 live = 0;
            for (u32 k = 0; k < D; ++k) {
                const std::size_t slot = std::size_t(k) * N + i;
                if (nbr_bond[slot] == kNoBond) break;
                if (!nbr_alive[slot]) continue;
                const R w = nbr_weight[slot];
                if (w <= R(0)) continue;
                const R rr[3] = {nbr_rest[3 * slot], nbr_rest[3 * slot + 1], nbr_rest[3 * slot + 2]};
                for (int r = 0; r < 3; ++r) for (int q = 0; q < 3; ++q) cov[r * 3 + q] += w * rr[r] * rr[q];
                ++live;
            }
            u8 valid = 0;
            if (live >= 1) {
                R inv[9], un[3];
                const int rank = pinv3(cov, inv, un);
                if (rank >= 2) {
                    for (int k = 0; k < 9; ++k) rinv[9 * std::size_t(i) + k] = inv[k];
                    for (int k = 0; k < 3; ++k) unmeasured[3 * std::size_t(i) + k] = un[k];
                    valid = 1;
                }
                if (rank < 3) ++rank_deficient;
            }
            node_valid[i] = valid;
            node_dirty[i] = 0;
        }
        if (!node_valid[i]) return;
        R acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const V3<R> xi = S.direct ? position(i) : ld3(u, i);
        for (u32 k = 0; k < D; ++k) {
            const std::size_t slot = std::size_t(k) * N + i;
            const bool present = nbr_bond[slot] != kNoBond;
            const u32 other = present ? nbr_other[slot] : i;
            const R w = (present && nbr_alive[slot]) ? nbr_weight[slot] : R(0);
            const V3<R> rest = present ? V3<R>{nbr_rest[3 * slot], nbr_rest[3 * slot + 1], nbr_rest[3 * slot + 2]} : V3<R>{0, 0, 0};
            const V3<R> xo = S.direct ? position(other) : ld3(u, other);
            const V3<R> cur = S.direct ? (xo - xi) - rest : xo - xi;
            const R cc[3] = {cur.x, cur.y, cur.z}, rr[3] = {rest.x, rest.y, rest.z};
            for (int r = 0; r < 3; ++r) for (int q = 0; q < 3; ++q) acc[r * 3 + q] += w * cc[r] * rr[q];
        }
        const R *inv = &rinv[9 * std::size_t(i)];
        R f[9];
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) {
                R s = R(0);
                for (int k = 0; k < 3; ++k) s += acc[r * 3 + k] * inv[k * 3 + q];
                f[r * 3 + q] = s;
            }
        static const int rows[6] = {0, 1, 2, 0, 0, 1}, cols[6] = {0, 1, 2, 1, 2, 2};
        R e[6];
        for (int q = 0; q < 6; ++q) {
            R c = R(0);
            for (int k = 0; k < 3; ++k) c += f[k * 3 + rows[q]] * f[k * 3 + cols[q]];
            e[q] = R(0.5) * (f[rows[q] * 3 + cols[q]] + f[cols[q] * 3 + rows[q]] + c);
        }
        const R un[3] = {unmeasured[3 * std::size_t(i)], unmeasured[3 * std::size_t(i) + 1], unmeasured[3 * std::size_t(i) + 2]};
        if (un[0] * un[0] + un[1] * un[1] + un[2] * un[2] > R(0)) {
            const R a[3] = {e[0] * un[0] + e[3] * un[1] + e[4] * un[2], e[3] * un[0] + e[1] * un[1] + e[5] * un[2],
                            e[4] * un[0] + e[5] * un[1] + e[2] * un[2]};
            const R s = a[0] * un[0] + a[1] * un[1] + a[2] * un[2];
            for (int q = 0; q < 6; ++q) e[q] += -un[rows[q]] * a[cols[q]] - a[rows[q]] * un[cols[q]] + s * un[rows[q]] * un[cols[q]];
        }
        for (int q = 0; q < 6; ++q) strain[6 * std::size_t(i) + q] = e[q];
    }

    // LatticePhysics.hpp:911-919.
    static void resolve(const R *e, V3<R> d, R &normal, R &shear) {
        const V3<R> tr{e[0] * d.x + e[3] * d.y + e[4] * d.z, e[3] * d.x + e[1] * d.y + e[5] * d.z,
                       e[4] * d.x + e[5] * d.y + e[2] * d.z};
        normal = dot(d, tr);
        shear = length(tr - normal * d);
    }

    // LatticePhysics.hpp:924-934.
    R bondStretch(u32 j) const {
        const V3<R> delta = bondVector(j);
        const R len = length(delta);
        const R rest = rest_length[j];
        if (S.direct) return len / rest - R(1);
        const V3<R> du = ld3(u, bond_b[j]) - ld3(u, bond_a[j]);
        const V3<R> r = ld3(rest_edge, j);
        const R l2 = rsm[j] + R(2) * dot(r, du) + dot(du, du);
        return l2 / ((len + rest) * rest);
    }

    // LatticePhysics.hpp:938-959.
    void sample(u32 j, R &t, R &c, R &s) const {
        const R stretch = bondStretch(j);
        t = stretch; c = -stretch; s = R(0);
        const V3<R> e = ld3(rest_edge, j);
        const R rl = length(e);
        if (rl > R(1.0e-9)) {
            const V3<R> dir = e * (R(1) / rl);
            const u32 ends[2] = {bond_a[j], bond_b[j]};
            for (int q = 0; q < 2; ++q) {
                const u32 n = ends[q];
                if (!node_valid[n]) continue;
                R nn, sh;
                resolve(&strain[6 * std::size_t(n)], dir, nn, sh);
                t = maxR(t, nn); c = maxR(c, -nn); s = maxR(s, sh);
            }
        }
    }

    // LatticePhysics.hpp:963-968.
    static R progress(R value, R start, R end) {
        if (!std::isfinite(start) || value <= start) return R(0);
        if (!std::isfinite(end) || end <= start) return value > start ? R(1) : R(0);
        const R p = (value - start) / (end - start);
        return p < R(0) ? R(0) : (p > R(1) ? R(1) : p);
    }

    // LatticePhysics.hpp:979-991.
    R elasticExtension(u32 j) const { return length(bondVector(j)) - rest_length[j] - plastic_ext[j]; }
    R storedEnergy(u32 j) const {
        const R c = compliance[j];
        if (c <= R(0)) return R(0);
        const R e = elasticExtension(j);
        return R(0.5) * e * e / c;
    }

    // LatticePhysics.hpp:1018-1039.
    R plasticReturn(u32 j) {
        if (!(S.yield > R(0))) return R(0);
        const R elastic = elasticExtension(j);
        const R mag = absR(elastic);
        const R h = S.hardening;
        const R ye = S.yield * rest_length[j] + h * plastic_strain[j];
        if (!(mag > ye)) return R(0);
        const R inc = (mag - ye) / (R(1) + h);
        plastic_ext[j] = plastic_ext[j] + (elastic < R(0) ? -inc : inc);
        plastic_strain[j] = plastic_strain[j] + inc;
        const R c = compliance[j];
        if (!(c > R(0))) return R(0);
        const R yend = ye + h * inc;
        return R(0.5) * (ye + yend) * inc / c;
    }

    struct Outcome { bool broke; double removed; float pt, pc, ps; double plastic; float pstretch; float dmg; };

    // LatticePhysics.hpp:1068-1114.
    Outcome endSample(u32 j) {
        Outcome o{false, 0.0, 0.f, 0.f, 0.f, 0.0, 0.f, 0.f};
        if (S.yield > R(0)) {
            o.plastic = double(plasticReturn(j));
            o.pstretch = float(absR(plastic_ext[j]) / rest_length[j]);
        }
        R t, c, s;
        sample(j, t, c, s);
        const R pt = maxR(R(0), maxR(prev_t[j], t)), pc = maxR(R(0), maxR(prev_c[j], c)), ps = maxR(R(0), maxR(prev_s[j], s));
        o.pt = float(pt); o.pc = float(pc); o.ps = float(ps);
        prev_t[j] = maxR(R(0), t); prev_c[j] = maxR(R(0), c); prev_s[j] = maxR(R(0), s);
        const R *th = &threshold[6 * std::size_t(j)];
        const R td = progress(pt, th[0], th[1]), cd = progress(pc, th[2], th[3]), sd = progress(ps, th[4], th[5]);
        R dmg = td; u8 mode = 1;
        if (cd > dmg) { dmg = cd; mode = 2; }
        if (sd > dmg) { dmg = sd; mode = 3; }
        if (dmg > damage[j]) { damage[j] = dmg; failure_mode[j] = mode; }
        o.dmg = float(damage[j]);
        if (damage[j] >= R(1)) {
            o.removed = double(storedEnergy(j));
            o.broke = true;
            alive[j] = 0;
            nbr_alive[slot_a[j]] = 0;
            nbr_alive[slot_b[j]] = 0;
        }
        return o;
    }

    // LatticePhysics.hpp:1118-1126.
    void startSample(u32 j) {
        R t, c, s;
        sample(j, t, c, s);
        prev_t[j] = maxR(R(0), t); prev_c[j] = maxR(R(0), c); prev_s[j] = maxR(R(0), s);
    }

    // LatticePhysics.hpp:1132-1168.
    void bondSolve(u32 j, bool first) {
        if (!alive[j]) return;
        const u32 a = bond_a[j], b = bond_b[j];
        const V3<R> delta = bondVector(j);
        const R len = length(delta);
        if (len <= R(1.0e-12)) return;
        const R orig = rest_length[j];
        const R plastic = plastic_ext[j];
        const R rest = orig + plastic;
        R C;
        if (S.direct) {
            C = len - rest;
        } else {
            const V3<R> du = ld3(u, b) - ld3(u, a);
            const V3<R> r = ld3(rest_edge, j);
            C = (rsm[j] + R(2) * dot(r, du) + dot(du, du) - plastic * (rest + orig)) / (len + rest);
        }
        const V3<R> dir = delta / len;
        const R wa = inv_mass[a], wb = inv_mass[b];
        const R alpha = compliance[j] / (S.dt * S.dt);
        const R lam = first ? R(0) : lambda[j];
        const R dl = (-C - alpha * lam) / (wa + wb + alpha);
        lambda[j] = lam + dl;
        st3(u, a, ld3(u, a) - wa * dl * dir);
        st3(u, b, ld3(u, b) + wb * dl * dir);
    }

    // LatticePhysics.hpp:1171-1184.
    void bondDamp(u32 j) {
        if (!alive[j]) return;
        const u32 a = bond_a[j], b = bond_b[j];
        const R ma = mass[a], mb = mass[b];
        if (ma <= R(0) || mb <= R(0)) return;
        const V3<R> n = normalized(bondVector(j));
        const V3<R> va = ld3(v, a), vb = ld3(v, b);
        const R rs = dot(vb - va, n);
        const R im = R(1) / ma + R(1) / mb;
        const R imp = -S.damping * rs / im;
        st3(v, a, va - (imp / ma) * n);
        st3(v, b, vb + (imp / mb) * n);
    }

    // LatticePhysics.hpp:1189-1198.
    static bool inside(const Plane<R> &pl, V3<R> p) {
        const V3<R> rel = p - pl.point;
        const R t = dot(rel, pl.tangent), b = dot(rel, pl.bitangent);
        for (u32 k = 0; k < pl.footprint_count; ++k)
            if (absR(t - pl.fp[k][0]) <= pl.fp[k][2] && absR(b - pl.fp[k][1]) <= pl.fp[k][3]) return true;
        return false;
    }

    // LatticePhysics.hpp:1204-1227.
    R project(const Plane<R> &pl, u32 i, R appr) {
        const V3<R> p = position(i);
        if (!inside(pl, p)) return R(0);
        const R dist = dot(p - pl.point, pl.normal) - pl.node_radius;
        if (dist > R(1.0e-8)) return R(0);
        if (pl.reach_capped) {
            const R reach = R(8) * maxR(R(0), -appr) * S.dt + R(1.0e-5);
            if (-dist > reach) return R(0);
        }
        st3(u, i, ld3(u, i) - dist * pl.normal);
        return -dist;
    }

    // LatticePhysics.hpp:1230-1252.
    void supportVelocity(const Plane<R> &pl, u32 i, R appr) {
        V3<R> vel = ld3(v, i);
        const R rest = -appr > R(0.5) ? (pl.restitution < R(0) ? R(0) : (pl.restitution > R(1) ? R(1) : pl.restitution)) : R(0);
        const R target = appr < R(0) ? -rest * appr : R(0);
        const R ns = dot(vel, pl.normal);
        R corr = ns;
        if (ns > target) corr = target;
        else if (ns < R(0)) corr = R(0);
        vel = vel + (corr - ns) * pl.normal;
        const R nd = corr - minR(R(0), appr);
        if (nd <= R(0)) { st3(v, i, vel); return; }
        const V3<R> tan = vel - dot(vel, pl.normal) * pl.normal;
        const R speed = length(tan);
        if (speed <= R(1.0e-12)) { st3(v, i, vel); return; }
        const R fd = speed <= pl.static_friction * nd ? speed : minR(speed, pl.dynamic_friction * nd);
        vel = vel - (fd / speed) * tan;
        st3(v, i, vel);
    }

    void nodeSupportProject(u32 i) {
        for (u32 p = 0; p < S.plane_count; ++p)
            if (project(S.planes[p], i, approach[std::size_t(p) * N + i]) > R(0)) engaged[std::size_t(p) * N + i] = 1;
        if (blade.enabled) projectBlade(i);
    }
    // The shortest move of a node out of the blade, across a flat or ahead of
    // the edge; the edge at substep k is at c0 + m * speed * dt * (k + 1).
    void projectBlade(u32 i) {
        const Blade<R> &bl = blade;
        const R travel = bl.speed * (S.dt * R(status.total_steps + 1));
        const V3<R> c = bl.c0 + travel * bl.m;
        const V3<R> rel = position(i) - c;
        const R a = dot(rel, bl.f), e = dot(rel, bl.m);
        const R reach = bl.hw + bl.r;
        if (!(absR(a) < reach) || !(e < bl.r)) return;
        const R flat = reach - absR(a), edge = bl.r - e;
        const V3<R> d = edge < flat ? edge * bl.m : (a < R(0) ? -flat : flat) * bl.f;
        st3(u, i, ld3(u, i) + d);
    }
    void nodeSupportVelocity(u32 i) {
        for (u32 p = 0; p < S.plane_count; ++p)
            if (engaged[std::size_t(p) * N + i]) supportVelocity(S.planes[p], i, approach[std::size_t(p) * N + i]);
    }
    void recordApproach(u32 i) {
        const V3<R> vel = ld3(v, i);
        for (u32 p = 0; p < S.plane_count; ++p) approach[std::size_t(p) * N + i] = dot(vel, S.planes[p].normal);
    }

    // LatticePhysics.hpp:1284-1295.
    bool sphereCandidate(const Sphere<R> &sp, u32 i) const {
        if (mass[i] <= R(0)) return false;
        const V3<R> r = position(i) - sp.center;
        const R dist = length(r);
        const R gap = dist - (sp.radius + S.c_node_radius);
        const R sb = length(ld3(v, i) - sp.velocity) + length(sp.omega) * dist;
        const R reach = maxR(S.c_margin, sb * S.dt) + S.c_slack;
        return gap <= reach;
    }

    R nodeKE(u32 i) const { return R(0.5) * mass[i] * length2(ld3(v, i)); }
    static R sphereKE(const Sphere<R> &s) { return R(0.5) * s.mass * length2(s.velocity) + R(0.5) * s.inertia * length2(s.omega); }

    // LatticePhysics.hpp:1324-1387.
    void contactNode(u32 i, bool correct, blx_contact_accumulators &acc) {
        if (mass[i] <= R(0)) return;
        const R dt = S.dt;
        const R ism = R(1) / sphere.mass;
        const R iin = R(1) / sphere.inertia;
        const R cr = sphere.radius + S.c_node_radius;
        const V3<R> r = position(i) - sphere.center;
        const R dist = length(r);
        const V3<R> n = normalized(r);
        const R gap = dist - cr;
        V3<R> nv = ld3(v, i);
        const V3<R> rel = nv - sphere.velocity - cross(sphere.omega, r);
        const R vn = dot(rel, n);
        if (gap > maxR(S.c_margin, -vn * dt)) return;
        if (-gap > R(acc.maximum_penetration_m)) acc.maximum_penetration_m = double(-gap);
        const R inm = R(1) / mass[i];
        auto eff = [&](V3<R> d) { return inm + ism + iin * length2(cross(r, d)); };
        const R rest = -vn > S.c_rest_thr ? S.c_rest : R(0);
        const R dvn = gap > S.c_margin ? -gap / dt : -rest * minR(vn, R(0));
        const R ni = maxR(R(0), (dvn - vn) / eff(n));
        if (ni > R(0)) {
            const R before = nodeKE(i);
            auto apply = [&](V3<R> imp) {
                nv = nv + inm * imp;
                st3(v, i, nv);
                sphere.velocity = sphere.velocity - ism * imp;
                const V3<R> ti = -cross(r, imp);
                sphere.omega = sphere.omega + iin * ti;
                acc.impulse_to_material[0] += double(imp.x);
                acc.impulse_to_material[1] += double(imp.y);
                acc.impulse_to_material[2] += double(imp.z);
            };
            apply(ni * n);
            const V3<R> anr = nv - sphere.velocity - cross(sphere.omega, r);
            const V3<R> tan = anr - dot(anr, n) * n;
            const R ts = length(tan);
            if (ts > R(1.0e-12)) {
                const V3<R> dir = tan / ts;
                const R stick = ts / eff(dir);
                const R fi = stick <= S.c_sf * ni ? stick : minR(stick, S.c_df * ni);
                apply(-fi * dir);
            }
            acc.dissipated_kinetic_energy_j += double(before - nodeKE(i));
            ++acc.impulse_contacts;
        }
        if (correct && gap < -S.c_margin) {
            const R corr = minR(-gap - S.c_margin, R(0.2) * maxR(S.c_node_radius, sphere.radius * R(0.01)));
            const R sum = inm + ism;
            const V3<R> ns = (corr * inm / sum) * n;
            const V3<R> ss = -((corr * ism / sum) * n);
            st3(u, i, ld3(u, i) + ns);
            sphere.center = sphere.center + ss;
            if (corr > R(acc.maximum_position_correction_m)) acc.maximum_position_correction_m = double(corr);
        }
    }

    // LatticePhysics.hpp:1395-1428.
    void contactPass(int pass, blx_contact_accumulators &acc) {
        const V3<R> cb = sphere.center;
        const R eb = sphereKE(sphere);
        for (u32 blk = 0; blk < sch_.block_count; ++blk) {
            const u32 count = cand_count[blk];
            for (u32 k = 0; k < count; ++k) {
                const u32 i = cand_list[std::size_t(blk) * kCap + k];
                contactNode(i, pass == 2, acc);
                if (pass == 1) {
                    recordApproach(i);
                    st3(u, i, ld3(u, i) + S.dt * ld3(v, i));
                }
            }
        }
        acc.dissipated_kinetic_energy_j += double(eb - sphereKE(sphere));
        if (pass == 2)
            for (u32 blk = 0; blk < sch_.block_count; ++blk)
                for (u32 k = 0; k < cand_count[blk]; ++k) nodeSupportVelocity(cand_list[std::size_t(blk) * kCap + k]);
        const R shift = length(sphere.center - cb);
        if (shift > R(acc.maximum_center_shift_m)) acc.maximum_center_shift_m = double(shift);
    }

    // LatticePhysics.hpp:1706-1719.
    bool kickClassify(const Sphere<R> &kicked, u32 i) {
        const V3<R> vel = ld3(v, i) + S.dt * S.gravity;
        st3(v, i, vel);
        st3(u_prev, i, ld3(u, i));
        engaged[i] = 0; engaged[std::size_t(N) + i] = 0;
        const bool cand = S.sphere_enabled && sphereCandidate(kicked, i);
        candidate[i] = cand ? 1 : 0;
        if (!cand) {
            recordApproach(i);
            st3(u, i, ld3(u, i) + S.dt * vel);
        }
        return cand;
    }

    // LatticePhysics.hpp:1726-1730.
    void velocityUpdate(u32 i) {
        const V3<R> vel = (ld3(u, i) - ld3(u_prev, i)) / S.dt;
        st3(v, i, vel);
        if (S.damping <= R(0) && !candidate[i]) nodeSupportVelocity(i);
    }

    // LatticePhysics.hpp:1797-1825.
    void sphereSupport(blx_contact_accumulators &acc) {
        for (u32 p = 0; p < S.plane_count; ++p) {
            const Plane<R> &pl = S.planes[p];
            const V3<R> foot = sphere.center - sphere.radius * pl.normal;
            if (!inside(pl, foot)) continue;
            const R dist = dot(foot - pl.point, pl.normal);
            if (dist > R(0)) continue;
            const R vn0 = dot(sphere.velocity, pl.normal);
            if (-dist > R(8) * maxR(R(0), -vn0) * S.dt + R(1.0e-5)) continue;
            const R vn = dot(sphere.velocity, pl.normal);
            if (vn < R(0)) {
                const R rest = -vn > R(0.5) ? pl.restitution : R(0);
                const R nd = -(R(1) + rest) * vn;
                sphere.velocity = sphere.velocity + nd * pl.normal;
                const V3<R> tan = sphere.velocity - dot(sphere.velocity, pl.normal) * pl.normal;
                const R speed = length(tan);
                if (speed > R(1.0e-12)) {
                    const R fd = speed <= pl.static_friction * nd ? speed : minR(speed, pl.dynamic_friction * nd);
                    sphere.velocity = sphere.velocity - (fd / speed) * tan;
                }
                ++acc.ball_support_events;
            }
            sphere.center = sphere.center - dist * pl.normal;
        }
    }

    // ExternalLoads.hpp:79-148, the force-only form (no wrenches, no clamps).
    void kickLoads() {
        if (!load_remaining) return;
        if (!(S.dt > 0) || !std::isfinite(S.dt)) throw invalid("external load needs a finite positive timestep");
        blx_load_ledger ch{};
        ch.steps = 1; ch.elapsed_s = double(S.dt);
        std::vector<V3<R>> after(loads.size());
        for (std::size_t k = 0; k < loads.size(); ++k) {
            const Load &ld = loads[k];
            const V3<R> before = ld3(v, ld.index);
            after[k] = before + (S.dt * inv_mass[ld.index]) * ld.applied;
            if (!fin3(after[k])) throw overflow("external force kick exceeds backend precision");
            const V3d old = widen(before), next = widen(after[k]);
            const V3d imp = double(mass[ld.index]) * (next - old);
            const V3d o{origin[0], origin[1], origin[2]};
            const V3d where = o + widen(position(ld.index));
            const V3d rq = double(S.dt) * ld.requested;
            const V3d arq = cross(where, rq), aimp = cross(where, imp);
            const double w = dot(imp, 0.5 * (old + next));
            ch.requested_impulse_n_s[0] += rq.x; ch.requested_impulse_n_s[1] += rq.y; ch.requested_impulse_n_s[2] += rq.z;
            ch.requested_angular_impulse_kg_m2_s[0] += arq.x; ch.requested_angular_impulse_kg_m2_s[1] += arq.y; ch.requested_angular_impulse_kg_m2_s[2] += arq.z;
            ch.impulse_n_s[0] += imp.x; ch.impulse_n_s[1] += imp.y; ch.impulse_n_s[2] += imp.z;
            ch.angular_impulse_kg_m2_s[0] += aimp.x; ch.angular_impulse_kg_m2_s[1] += aimp.y; ch.angular_impulse_kg_m2_s[2] += aimp.z;
            ch.work_j += w;
        }
        blx_load_ledger &L = status.external_load;
        L.steps += ch.steps; L.elapsed_s += ch.elapsed_s;
        for (int q = 0; q < 3; ++q) {
            L.requested_impulse_n_s[q] += ch.requested_impulse_n_s[q];
            L.impulse_n_s[q] += ch.impulse_n_s[q];
            L.requested_angular_impulse_kg_m2_s[q] += ch.requested_angular_impulse_kg_m2_s[q];
            L.angular_impulse_kg_m2_s[q] += ch.angular_impulse_kg_m2_s[q];
        }
        L.work_j += ch.work_j;
        if (!std::isfinite(L.work_j) || !std::isfinite(L.impulse_n_s[0]) || !std::isfinite(L.impulse_n_s[1]) || !std::isfinite(L.impulse_n_s[2]))
            throw overflow("external load ledger exceeds finite range");
        for (std::size_t k = 0; k < loads.size(); ++k) st3(v, loads[k].index, after[k]);
        --load_remaining;
    }

    template <class F> void sweep(u32 blk, u32 boundary, F &&f) {
        const u32 cc = std::max(sch_.color_count, 1U);
        for (u32 c = 0; c < sch_.color_count; ++c) {
            const std::size_t key = (std::size_t(blk) * 2 + boundary) * cc + c;
            for (u32 j = sch_.range_begin[key]; j < sch_.range_end[key]; ++j) f(j);
        }
    }
    template <class F> void sweepAll(F &&f) {
        for (u32 b = 0; b < sch_.block_count; ++b) sweep(b, 0, f);
        for (u32 b = 0; b < sch_.block_count; b += 2) sweep(b, 1, f);
        for (u32 b = 1; b < sch_.block_count; b += 2) sweep(b, 1, f);
    }

    // CpuLatticeBackend.cpp:357-488.
    bool advance() {
        if (dirty_start) {
            for (u32 i = 0; i < N; ++i) nodeStrain(i);
            for (u32 j = 0; j < B; ++j) if (alive[j]) startSample(j);
            dirty_start = false;
        }
        kickLoads();
        Sphere<R> kicked = sphere;
        kicked.velocity = kicked.velocity + S.dt * S.gravity;
        for (u32 blk = 0; blk < sch_.block_count; ++blk) {
            u32 count = 0;
            for (u32 i = sch_.node_block_begin[blk]; i < sch_.node_block_begin[blk + 1]; ++i) {
                if (!kickClassify(kicked, i)) continue;
                if (count < kCap) cand_list[std::size_t(blk) * kCap + count] = i;
                else ++status.contact.candidate_overflow;
                ++count;
            }
            cand_count[blk] = count < kCap ? count : kCap;
        }
        sphere = kicked;
        if (S.sphere_enabled) contactPass(1, status.contact);
        for (u32 it = 0; it < S.iterations; ++it) {
            const bool first = it == 0;
            for (u32 b = 0; b < sch_.block_count; ++b) sweep(b, 0, [&](u32 j) { bondSolve(j, first); });
            for (u32 b = 0; b < sch_.block_count; b += 2) sweep(b, 1, [&](u32 j) { bondSolve(j, first); });
            for (u32 b = 1; b < sch_.block_count; b += 2) sweep(b, 1, [&](u32 j) { bondSolve(j, first); });
            for (u32 i = 0; i < N; ++i) nodeSupportProject(i);
        }
        for (u32 i = 0; i < N; ++i) velocityUpdate(i);
        if (S.damping > R(0)) {
            sweepAll([&](u32 j) { bondDamp(j); });
            for (u32 i = 0; i < N; ++i) if (!candidate[i]) nodeSupportVelocity(i);
        }
        if (S.sphere_enabled) contactPass(2, status.contact);
        if (NC.mode != BLX_NODE_CONTACT_OFF) nodeContact();
        return finish();
    }

    bool finish() {
        for (u32 i = 0; i < N; ++i) nodeStrain(i);
        bool any = false;
        const bool first_round = status.broken_bonds == 0;
        for (u32 j = 0; j < B; ++j) {
            if (!alive[j]) continue;
            const Outcome o = endSample(j);
            status.max_tensile_stretch = std::max(status.max_tensile_stretch, o.pt);
            status.max_compressive_strain = std::max(status.max_compressive_strain, o.pc);
            status.max_shear_strain = std::max(status.max_shear_strain, o.ps);
            status.plastic_work_j += o.plastic;
            status.max_plastic_stretch = std::max(status.max_plastic_stretch, o.pstretch);
            if (o.dmg > status.max_damage + 1.0e-6) {
                status.max_damage = o.dmg;
                status.last_damage_gain_step = status.total_steps;
            }
            if (!o.broke) continue;
            any = true;
            if (first_round) first_failure.push_back(j);
            status.removed_energy_j += o.removed;
            if (o.removed > flat_fraction * std::max(status.removed_energy_j, 1.0e-12))
                status.last_energy_gain_step = status.total_steps;
            ++status.broken_bonds;
            node_dirty[bond_a[j]] = 1;
            node_dirty[bond_b[j]] = 1;
        }
        status.rank_deficient_nodes = rank_deficient;
        sphere.center = sphere.center + S.dt * sphere.velocity;
        sphereSupport(status.contact);
        if (any) dirty_start = contact_rebuild = true;
        return any;
    }

    // LatticePhysics.hpp:1766-1791.
    static unsigned exitReason(u64 completed, unsigned broken, u64 last_failure, u64 quiet, u64 min_steps, u64 no_failure,
                               u64 energy_flat, u64 last_energy_gain, u64 calm, u64 last_damage_gain, double max_damage,
                               double calm_margin, double removed, double removable) {
        if (removable > 0.0 && removed >= removable) return 6U;
        if (broken > 0U && quiet > 0 && completed >= min_steps && completed > last_failure && completed - last_failure > quiet) return 1U;
        if (broken == 0U && no_failure > 0 && completed >= no_failure) return 2U;
        if (broken == 0U && calm > 0 && completed >= calm && max_damage < calm_margin && completed > last_damage_gain &&
            completed - last_damage_gain > calm)
            return 5U;
        if (broken > 0U && energy_flat > 0 && completed >= min_steps && completed > last_energy_gain &&
            completed - last_energy_gain > energy_flat)
            return 4U;
        return 0U;
    }
};

// fracture/ConnectedComponents.cpp:14-87: union by rank with path compression
// over the live bonds, then ids ranked by (size descending, smallest node).
inline u32 components(u32 N, u32 B, const u32 *ba, const u32 *bb, const u8 *alive, u32 *label) {
    std::vector<u32> parent(N);
    std::vector<u8> rank(N, 0);
    std::iota(parent.begin(), parent.end(), 0U);
    auto find = [&](u32 x) {
        u32 r = x;
        while (parent[r] != r) r = parent[r];
        while (parent[x] != r) { const u32 n = parent[x]; parent[x] = r; x = n; }
        return r;
    };
    for (u32 j = 0; j < B; ++j) {
        if (!alive[j]) continue;
        u32 a = find(ba[j]), b = find(bb[j]);
        if (a == b) continue;
        if (rank[a] < rank[b]) std::swap(a, b);
        parent[b] = a;
        if (rank[a] == rank[b]) ++rank[a];
    }
    std::vector<u32> root_comp(N, kNoBond), size, first;
    std::vector<u32> comp_of(N);
    for (u32 i = 0; i < N; ++i) {
        const u32 r = find(i);
        if (root_comp[r] == kNoBond) { root_comp[r] = u32(size.size()); size.push_back(0); first.push_back(i); }
        comp_of[i] = root_comp[r];
        ++size[root_comp[r]];
    }
    std::vector<u32> order(size.size());
    std::iota(order.begin(), order.end(), 0U);
    std::sort(order.begin(), order.end(), [&](u32 l, u32 r) {
        if (size[l] != size[r]) return size[l] > size[r];
        return first[l] < first[r];
    });
    std::vector<u32> id(size.size());
    for (u32 k = 0; k < order.size(); ++k) id[order[k]] = k;
    for (u32 i = 0; i < N; ++i) label[i] = id[comp_of[i]];
    return u32(size.size());
}

struct Handle {
    int precision;
    std::unique_ptr<Backend<float>> f;
    std::unique_ptr<Backend<double>> d;
};

thread_local std::string g_error;

template <class F> int guard(F &&f) {
    try { f(); g_error.clear(); return BLX_OK; }
    catch (const overflow &e) { g_error = e.what(); return BLX_ERR_OVERFLOW; }
    catch (const logic &e) { g_error = e.what(); return BLX_ERR_LOGIC; }
    catch (const std::invalid_argument &e) { g_error = e.what(); return BLX_ERR_INVALID; }
    catch (const std::overflow_error &e) { g_error = e.what(); return BLX_ERR_OVERFLOW; }
    catch (const std::logic_error &e) { g_error = e.what(); return BLX_ERR_LOGIC; }
    catch (const std::exception &e) { g_error = e.what(); return BLX_ERR_RUNTIME; }
}

inline Schedule fromView(const blx_schedule_view &v) {
    Schedule s;
    s.node_count = v.node_count; s.bond_count = v.bond_count; s.block_count = v.block_count;
    s.color_count = v.color_count; s.boundary_bond_count = v.boundary_bond_count;
    s.node_block_begin.assign(v.node_block_begin, v.node_block_begin + v.block_count + 1);
    s.bond_order.assign(v.bond_order, v.bond_order + v.bond_count);
    s.bond_schedule_index.assign(v.bond_schedule_index, v.bond_schedule_index + v.bond_count);
    s.bond_block_begin.assign(v.bond_block_begin, v.bond_block_begin + v.block_count + 1);
    const std::size_t ranges = std::size_t(v.block_count) * 2 * std::max(v.color_count, 1U);
    s.range_begin.assign(v.range_begin, v.range_begin + ranges);
    s.range_end.assign(v.range_end, v.range_end + ranges);
    s.bond_color.assign(v.bond_color, v.bond_color + v.bond_count);
    return s;
}

// calculateFragmentMassProperties (fracture/FragmentMassProperties.cpp:7-80)
// of one fragment: serial sums in node order, the reference's accumulators
// and expressions; Mat3::inverse with epsilon 0 (core/Math.hpp:110-127).
inline void fragmentMass(const double *pos, const double *vel, const double *spin, const double *mass, double cell,
                         const u32 *nodes, u32 count, u32 n_nodes, blx_fragment_mass &out) {
    if (count == 0) throw invalid("fragment requires at least one material node");
    out = blx_fragment_mass{};
    double m_sum = 0.0, wp[3] = {0.0, 0.0, 0.0}, lm[3] = {0.0, 0.0, 0.0};
    for (u32 k = 0; k < count; ++k) {
        const u32 i = nodes[k];
        if (i >= n_nodes) throw std::out_of_range("fragment node index is invalid");
        const double m = mass[i];
        m_sum += m;
        for (int c = 0; c < 3; ++c) wp[c] += pos[3 * size_t(i) + c] * m;
        for (int c = 0; c < 3; ++c) lm[c] += vel[3 * size_t(i) + c] * m;
    }
    if (m_sum <= 0.0) throw std::runtime_error("fragment has no mass");
    double com[3], vl[3];
    for (int c = 0; c < 3; ++c) {
        com[c] = wp[c] / m_sum;
        vl[c] = lm[c] / m_sum;
    }
    double I[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}}, L[3] = {0.0, 0.0, 0.0}, src = 0.0;
    for (u32 k = 0; k < count; ++k) {
        const u32 i = nodes[k];
        const double m = mass[i];
        const double *p = pos + 3 * size_t(i), *v = vel + 3 * size_t(i);
        const double s0 = spin ? spin[3 * size_t(i)] : 0.0, s1 = spin ? spin[3 * size_t(i) + 1] : 0.0,
                     s2 = spin ? spin[3 * size_t(i) + 2] : 0.0;
        const double rx = p[0] - com[0], ry = p[1] - com[1], rz = p[2] - com[2];
        const double cdi = m * cell * cell / 6.0;
        I[0][0] += m * (ry * ry + rz * rz) + cdi;
        I[1][1] += m * (rx * rx + rz * rz) + cdi;
        I[2][2] += m * (rx * rx + ry * ry) + cdi;
        I[0][1] -= m * rx * ry;
        I[1][0] = I[0][1];
        I[0][2] -= m * rx * rz;
        I[2][0] = I[0][2];
        I[1][2] -= m * ry * rz;
        I[2][1] = I[1][2];
        const double wx = (v[0] - vl[0]) * m, wy = (v[1] - vl[1]) * m, wz = (v[2] - vl[2]) * m;
        L[0] += (ry * wz - rz * wy) + s0 * cdi;
        L[1] += (rz * wx - rx * wz) + s1 * cdi;
        L[2] += (rx * wy - ry * wx) + s2 * cdi;
        src += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 0.5 * cdi * (s0 * s0 + s1 * s1 + s2 * s2);
    }
    const double det = I[0][0] * (I[1][1] * I[2][2] - I[1][2] * I[2][1]) -
                       I[0][1] * (I[1][0] * I[2][2] - I[1][2] * I[2][0]) +
                       I[0][2] * (I[1][0] * I[2][1] - I[1][1] * I[2][0]);
    if (std::abs(det) <= 0.0) throw std::runtime_error("fragment inertia is singular");
    double J[3][3];
    J[0][0] = (I[1][1] * I[2][2] - I[1][2] * I[2][1]) / det;
    J[0][1] = (I[0][2] * I[2][1] - I[0][1] * I[2][2]) / det;
    J[0][2] = (I[0][1] * I[1][2] - I[0][2] * I[1][1]) / det;
    J[1][0] = (I[1][2] * I[2][0] - I[1][0] * I[2][2]) / det;
    J[1][1] = (I[0][0] * I[2][2] - I[0][2] * I[2][0]) / det;
    J[1][2] = (I[0][2] * I[1][0] - I[0][0] * I[1][2]) / det;
    J[2][0] = (I[1][0] * I[2][1] - I[1][1] * I[2][0]) / det;
    J[2][1] = (I[0][1] * I[2][0] - I[0][0] * I[2][1]) / det;
    J[2][2] = (I[0][0] * I[1][1] - I[0][1] * I[1][0]) / det;
    out.mass_kg = m_sum;
    for (int c = 0; c < 3; ++c) {
        out.center_of_mass_world_m[c] = com[c];
        out.linear_velocity_m_s[c] = vl[c];
        out.angular_momentum_kg_m2_s[c] = L[c];
        out.angular_velocity_rad_s[c] = J[c][0] * L[0] + J[c][1] * L[1] + J[c][2] * L[2];
        for (int r = 0; r < 3; ++r) out.inertia_world_kg_m2[3 * c + r] = I[c][r];
    }
    const double *w = out.angular_velocity_rad_s;
    out.source_kinetic_energy_j = src;
    out.rigid_kinetic_energy_j = 0.5 * m_sum * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]) +
                                 0.5 * (w[0] * L[0] + w[1] * L[1] + w[2] * L[2]);
    out.coarsening_kinetic_loss_j = out.source_kinetic_energy_j - out.rigid_kinetic_energy_j;
    out.node_count = count;
    out.first_node = nodes[0];
}

} // namespace orc

using namespace orc;

extern "C" {

const char *orc_last_error(void) { return orc::g_error.c_str(); }

int orc_schedule_build(const blx_asset_view *a, const uint32_t *nbb, uint32_t blocks, blx_schedule_view *out) {
    return guard([&] {
        const Schedule s = buildSchedule(a->node_count, a->bond_count, a->bond_a, a->bond_b, nbb, blocks);
        out->node_count = s.node_count; out->bond_count = s.bond_count; out->block_count = s.block_count;
        out->color_count = s.color_count; out->boundary_bond_count = s.boundary_bond_count;
        std::copy(s.node_block_begin.begin(), s.node_block_begin.end(), out->node_block_begin);
        std::copy(s.bond_order.begin(), s.bond_order.end(), out->bond_order);
        std::copy(s.bond_schedule_index.begin(), s.bond_schedule_index.end(), out->bond_schedule_index);
        std::copy(s.bond_block_begin.begin(), s.bond_block_begin.end(), out->bond_block_begin);
        std::copy(s.range_begin.begin(), s.range_begin.end(), out->range_begin);
        std::copy(s.range_end.begin(), s.range_end.end(), out->range_end);
        std::copy(s.bond_color.begin(), s.bond_color.end(), out->bond_color);
    });
}

int orc_state_max_degree(const blx_asset_view *a, uint32_t *d) { return guard([&] { *d = maxDegree(*a); }); }

int orc_state_build(const blx_asset_view *a, const blx_schedule_view *sv, const double origin[3], blx_state_view *o) {
    return guard([&] {
        const State s = buildState(*a, fromView(*sv), origin);
        for (int k = 0; k < 3; ++k) o->origin[k] = s.origin[k];
        o->node_count = s.N; o->bond_count = s.B; o->max_degree = s.D;
        auto cp = [](const auto &src, auto *dst) { std::copy(src.begin(), src.end(), dst); };
        cp(s.x0, o->x0); cp(s.u, o->u); cp(s.u_prev, o->u_prev); cp(s.v, o->v); cp(s.inv_mass, o->inv_mass); cp(s.mass, o->mass);
        cp(s.adj_offsets, o->adj_offsets); cp(s.adj_bonds, o->adj_bonds); cp(s.nbr_bond, o->nbr_bond); cp(s.nbr_other, o->nbr_other);
        cp(s.nbr_rest, o->nbr_rest); cp(s.nbr_weight, o->nbr_weight); cp(s.nbr_alive, o->nbr_alive);
        cp(s.bond_slot_a, o->bond_slot_a); cp(s.bond_slot_b, o->bond_slot_b); cp(s.bond_a, o->bond_a); cp(s.bond_b, o->bond_b);
        cp(s.rest_edge, o->rest_edge); cp(s.rest_length, o->rest_length); cp(s.rsm, o->rest_length_sq_minus); cp(s.weight, o->weight);
        cp(s.compliance, o->compliance); cp(s.threshold, o->threshold); cp(s.alive, o->alive); cp(s.failure_mode, o->failure_mode);
        cp(s.damage, o->damage); cp(s.prev_t, o->prev_tensile); cp(s.prev_c, o->prev_compressive); cp(s.prev_s, o->prev_shear);
        cp(s.plastic_ext, o->plastic_extension); cp(s.plastic_strain, o->plastic_strain);
    });
}

// Moore-stencil box: node-major, the 13 positive half offsets in the
// reference's dz, dy, dx loop order (matter/BoxLattice.cpp:141-168).
int orc_box26_bonds(uint32_t nx, uint32_t ny, uint32_t nz, double cell, double *pos, uint32_t *ba, uint32_t *bb, uint64_t *count) {
    return guard([&] {
        u64 n = 0;
        for (u32 z = 0; z < nz; ++z) for (u32 y = 0; y < ny; ++y) for (u32 x = 0; x < nx; ++x) {
            const u32 i = x + nx * (y + ny * z);
            if (pos) {
                pos[3 * std::size_t(i)] = (double(x) + 0.5 - 0.5 * nx) * cell;
                pos[3 * std::size_t(i) + 1] = (double(y) + 0.5 - 0.5 * ny) * cell;
                pos[3 * std::size_t(i) + 2] = (double(z) + 0.5 - 0.5 * nz) * cell;
            }
            for (int dz = -1; dz <= 1; ++dz) for (int dy = -1; dy <= 1; ++dy) for (int dx = -1; dx <= 1; ++dx) {
                if (!(dz > 0 || (dz == 0 && dy > 0) || (dz == 0 && dy == 0 && dx > 0))) continue;
                const long X = long(x) + dx, Y = long(y) + dy, Z = long(z) + dz;
                if (X < 0 || Y < 0 || Z < 0 || X >= long(nx) || Y >= long(ny) || Z >= long(nz)) continue;
                if (ba) { ba[n] = i; bb[n] = u32(X + long(nx) * (Y + long(ny) * Z)); }
                ++n;
            }
        }
        *count = n;
    });
}

int orc_components(uint32_t N, uint32_t B, const uint32_t *ba, const uint32_t *bb, const uint8_t *alive, uint32_t *label,
                   uint32_t *count) {
    return guard([&] { *count = components(N, B, ba, bb, alive, label); });
}

// Mass properties of fragments given as node lists: fragment f is
// nodes[offsets[f] .. offsets[f+1]); pos, vel, spin (NULL: zero) 3N, mass N.
int orc_fragment_mass(uint32_t N, const double *pos, const double *vel, const double *spin, const double *mass,
                      double cell, uint32_t K, const uint32_t *offsets, const uint32_t *nodes, blx_fragment_mass *out) {
    return guard([&] {
        for (u32 f = 0; f < K; ++f)
            fragmentMass(pos, vel, spin, mass, cell, nodes + offsets[f], offsets[f + 1] - offsets[f], N, out[f]);
    });
}

int orc_backend_create(int precision, const blx_schedule_view *sv, void **out) {
    return guard([&] {
        auto h = std::make_unique<Handle>();
        h->precision = precision;
        if (precision == BLX_PRECISION_FLOAT) h->f = std::make_unique<Backend<float>>(fromView(*sv));
        else if (precision == BLX_PRECISION_DOUBLE) h->d = std::make_unique<Backend<double>>(fromView(*sv));
        else throw invalid("unknown precision");
        *out = h.release();
    });
}

int orc_backend_destroy(void *h) { return guard([&] { delete static_cast<Handle *>(h); }); }

#define ORC_DISPATCH(h, expr)                                                  \
    do {                                                                       \
        Handle *hh = static_cast<Handle *>(h);                                 \
        if (hh->f) { auto &b = *hh->f; expr; } else { auto &b = *hh->d; expr; } \
    } while (0)

int orc_backend_upload(void *h, const blx_state_view *st, const blx_settings *s, const blx_sphere *sp) {
    return guard([&] { ORC_DISPATCH(h, b.upload(*st, *s, *sp)); });
}
int orc_backend_set_blade(void *h, const blx_blade *bl) {
    return guard([&] { ORC_DISPATCH(h, b.setBlade(*bl)); });
}
int orc_backend_set_external_forces(void *h, const double *f, uint64_t substeps) {
    return guard([&] { ORC_DISPATCH(h, b.setForces(f, substeps)); });
}
int orc_backend_run(void *h, const blx_run_control *c, blx_run_status *st) {
    return guard([&] { ORC_DISPATCH(h, { b.run(*c); *st = b.status; }); });
}
int orc_backend_status(void *h, blx_run_status *st) { return guard([&] { ORC_DISPATCH(h, *st = b.status); }); }
int orc_backend_download(void *h, blx_state_view *st, blx_sphere *sp) {
    return guard([&] { ORC_DISPATCH(h, b.download(*st, *sp)); });
}
int orc_backend_first_failure_bonds(void *h, uint32_t *out, uint64_t cap, uint64_t *count) {
    return guard([&] {
        ORC_DISPATCH(h, {
            *count = b.first_failure.size();
            for (std::size_t k = 0; k < b.first_failure.size() && k < cap; ++k) out[k] = b.first_failure[k];
        });
    });
}
int orc_backend_frame_count(void *h, uint32_t *count) { return guard([&] { ORC_DISPATCH(h, *count = u32(b.frames.size())); }); }
int orc_backend_take_frame(void *h, uint32_t index, uint64_t *step, float *u, uint8_t *alive, float *damage, blx_sphere *sp) {
    return guard([&] {
        ORC_DISPATCH(h, {
            if (index >= b.frames.size()) throw invalid("frame index out of range");
            const auto &f = b.frames[index];
            *step = f.step;
            std::copy(f.u.begin(), f.u.end(), u);
            std::copy(f.alive.begin(), f.alive.end(), alive);
            std::copy(f.damage.begin(), f.damage.end(), damage);
            if (sp) {
                sp->center[0] = double(f.sphere.center.x); sp->center[1] = double(f.sphere.center.y); sp->center[2] = double(f.sphere.center.z);
                sp->velocity[0] = double(f.sphere.velocity.x); sp->velocity[1] = double(f.sphere.velocity.y); sp->velocity[2] = double(f.sphere.velocity.z);
                sp->angular_velocity[0] = double(f.sphere.omega.x); sp->angular_velocity[1] = double(f.sphere.omega.y); sp->angular_velocity[2] = double(f.sphere.omega.z);
                sp->radius = double(f.sphere.radius); sp->mass = double(f.sphere.mass); sp->inertia = double(f.sphere.inertia);
            }
        });
    });
}
int orc_backend_clear_frames(void *h) { return guard([&] { ORC_DISPATCH(h, { b.frames.clear(); b.status.frames_captured = 0; }); }); }

} // extern "C"
