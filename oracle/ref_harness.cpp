// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A C-ABI harness around the reference's own lattice lane, compiled from the
// reference's sources where they lie (see oracle/Makefile; nothing from the
// reference is copied into this repository). It exposes the same function
// shapes as the product's C ABI under a ref_ prefix so tests can put the real
// reference, the oracle restatement and the HIP backend side by side:
//   ref_schedule_build   -> buildLatticeSchedule   (LatticeSchedule.cpp:18-104)
//   ref_state_build      -> buildLatticeState      (FastLattice.cpp:47-169)
//   ref_backend_*        -> makeCpuLatticeBackend / makeParallelCpuLatticeBackend
//                           (FastLattice.hpp:371-381) and LatticeBackend methods
//   ref_components       -> findConnectedComponents (ConnectedComponents.cpp:51-87)
//   ref_fragment_mass    -> calculateFragmentMassProperties
//                           (FragmentMassProperties.cpp:7-80)
//   ref_fragment_build   -> buildFragmentRepresentations (FragmentGeometry.cpp:203-291)
//   ref_merge_cells_into_boxes -> mergeCellsIntoBoxes (PieceGeometry.cpp:9-46)
//   ref_fragment_lattice -> buildFragmentLattice     (Refracture.cpp:267-393)
//   ref_box_slabs        -> boxLatticeSlabs         (BoxLattice.cpp:175-195)
//   ref_state_substep_limit, ref_state_elastic_energy, ref_state_kinetic_energy
//                        -> latticeStateSubstepLimit / ElasticEnergy / KineticEnergy
//                           (FastLattice.cpp:30-45,199-224)

#include "../include/banjo_lattice.h"

#include "fastlattice/FastLattice.hpp"
#include "fastlattice/LatticeSchedule.hpp"
#include "fastlattice/Refracture.hpp"
#include "fracture/ConnectedComponents.hpp"
#include "fracture/FragmentGeometry.hpp"
#include "fracture/FragmentMassProperties.hpp"
#include "fracture/PieceGeometry.hpp"
#include "matter/BoxLattice.hpp"

#include <atomic>
#include <cstring>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

using namespace banjo;
using namespace banjo::fastlattice;

namespace {

thread_local std::string g_error;

template <class F> int guard(F &&f) {
    try { f(); g_error.clear(); return BLX_OK; }
    catch (const std::invalid_argument &e) { g_error = e.what(); return BLX_ERR_INVALID; }
    catch (const std::overflow_error &e) { g_error = e.what(); return BLX_ERR_OVERFLOW; }
    catch (const std::logic_error &e) { g_error = e.what(); return BLX_ERR_LOGIC; }
    catch (const std::exception &e) { g_error = e.what(); return BLX_ERR_RUNTIME; }
}

LatticeAsset makeAsset(const blx_asset_view &a) {
    LatticeAsset asset;
    asset.nodes.resize(a.node_count);
    for (std::uint32_t i = 0; i < a.node_count; ++i)
        if (a.reference_position)
            asset.nodes[i].local_position_m = {a.reference_position[3 * i], a.reference_position[3 * i + 1], a.reference_position[3 * i + 2]};
    asset.bonds.resize(a.bond_count);
    for (std::uint32_t j = 0; j < a.bond_count; ++j) {
        BondRest &b = asset.bonds[j];
        b.node_a = a.bond_a[j];
        b.node_b = a.bond_b[j];
        if (a.rest_length) b.rest_length_m = a.rest_length[j];
        if (a.compliance) b.compliance = a.compliance[j];
        if (a.threshold) {
            const double *t = a.threshold + 6 * std::size_t(j);
            b.damage_start_stretch = t[0]; b.damage_end_stretch = t[1];
            b.compression_damage_start_strain = t[2]; b.compression_damage_end_strain = t[3];
            b.shear_damage_start_strain = t[4]; b.shear_damage_end_strain = t[5];
        }
    }
    return asset;
}

void fillMatter(ActiveMatter &m, const LatticeAsset &asset, const blx_asset_view &a) {
    m.asset = &asset;
    m.nodes.resize(a.node_count);
    m.reference_positions_world_m.resize(a.node_count);
    for (std::uint32_t i = 0; i < a.node_count; ++i) {
        m.reference_positions_world_m[i] = {a.reference_position[3 * i], a.reference_position[3 * i + 1], a.reference_position[3 * i + 2]};
        m.nodes[i].position_world_m = {a.position[3 * i], a.position[3 * i + 1], a.position[3 * i + 2]};
        m.nodes[i].previous_position_world_m = m.nodes[i].position_world_m;
        m.nodes[i].velocity_m_s = {a.velocity[3 * i], a.velocity[3 * i + 1], a.velocity[3 * i + 2]};
        m.nodes[i].mass_kg = a.mass[i];
    }
    m.bonds.resize(a.bond_count);
    for (std::uint32_t j = 0; j < a.bond_count; ++j) {
        m.bonds[j].alive = a.alive ? a.alive[j] != 0 : true;
        m.bonds[j].damage = a.damage ? a.damage[j] : 0.0;
        m.bonds[j].failure_mode = a.failure_mode ? static_cast<BondFailureMode>(a.failure_mode[j]) : BondFailureMode::None;
    }
}

LatticeSchedule scheduleFromView(const blx_schedule_view &v) {
    LatticeSchedule s;
    s.block_count = v.block_count;
    s.color_count = v.color_count;
    s.boundary_bond_count = v.boundary_bond_count;
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

template <class T> void put(const std::vector<T> &src, T *dst) { if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(T)); }
template <class T> std::vector<T> take(const T *p, std::size_t n) { return std::vector<T>(p, p + n); }

LatticeState stateFromView(const blx_state_view &v) {
    LatticeState s;
    s.origin = {v.origin[0], v.origin[1], v.origin[2]};
    s.node_count = v.node_count; s.bond_count = v.bond_count; s.max_degree = v.max_degree;
    const std::size_t N = v.node_count, B = v.bond_count, P = std::size_t(v.max_degree) * N;
    s.x0 = take(v.x0, 3 * N); s.u = take(v.u, 3 * N); s.u_prev = take(v.u_prev, 3 * N); s.v = take(v.v, 3 * N);
    s.inv_mass = take(v.inv_mass, N); s.mass = take(v.mass, N);
    s.adj_offsets = take(v.adj_offsets, N + 1); s.adj_bonds = take(v.adj_bonds, s.adj_offsets.back());
    s.nbr_bond = take(v.nbr_bond, P); s.nbr_other = take(v.nbr_other, P); s.nbr_rest = take(v.nbr_rest, 3 * P);
    s.nbr_weight = take(v.nbr_weight, P); s.nbr_alive = take(v.nbr_alive, P);
    s.bond_slot_a = take(v.bond_slot_a, B); s.bond_slot_b = take(v.bond_slot_b, B);
    s.bond_a = take(v.bond_a, B); s.bond_b = take(v.bond_b, B);
    s.rest_edge = take(v.rest_edge, 3 * B); s.rest_length = take(v.rest_length, B);
    s.rest_length_sq_minus = take(v.rest_length_sq_minus, B); s.weight = take(v.weight, B);
    s.compliance = take(v.compliance, B); s.threshold = take(v.threshold, 6 * B);
    s.alive = take(v.alive, B); s.failure_mode = take(v.failure_mode, B);
    s.damage = take(v.damage, B); s.prev_tensile = take(v.prev_tensile, B);
    s.prev_compressive = take(v.prev_compressive, B); s.prev_shear = take(v.prev_shear, B);
    s.plastic_extension = take(v.plastic_extension, B); s.plastic_strain = take(v.plastic_strain, B);
    return s;
}

V3<double> v3of(const double *p) { return {p[0], p[1], p[2]}; }
void of3(const V3<double> &v, double *p) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

StepSettings<double> settingsFromC(const blx_settings &c) {
    StepSettings<double> s{};
    s.dt = c.dt; s.gravity = v3of(c.gravity); s.constraint_iterations = c.constraint_iterations;
    s.bond_integrator = c.bond_integrator; s.damping_fraction = c.damping_fraction;
    s.sphere_enabled = c.sphere_enabled; s.direct_arithmetic = c.direct_arithmetic; s.audit_energy = c.audit_energy;
    s.plastic_yield_stretch = c.plastic_yield_stretch; s.plastic_hardening = c.plastic_hardening;
    s.plate_half_thickness = c.plate_half_thickness;
    s.contact = {c.contact.static_friction, c.contact.dynamic_friction, c.contact.restitution,
                 c.contact.restitution_speed_threshold, c.contact.node_contact_radius, c.contact.contact_margin,
                 c.contact.prefilter_slack};
    s.node_contact.mode = c.node_contact.mode; s.node_contact.radius = c.node_contact.radius;
    s.node_contact.skin = c.node_contact.skin; s.node_contact.margin = c.node_contact.margin;
    s.node_contact.restitution = c.node_contact.restitution;
    s.node_contact.restitution_speed_threshold = c.node_contact.restitution_speed_threshold;
    s.node_contact.static_friction = c.node_contact.static_friction;
    s.node_contact.dynamic_friction = c.node_contact.dynamic_friction;
    s.node_contact.bucket_mask = c.node_contact.bucket_mask;
    s.support.plane_count = c.plane_count;
    for (unsigned p = 0; p < kMaxSupportPlanes; ++p) {
        const blx_support_plane &in = c.planes[p];
        SupportPlane<double> &o = s.support.planes[p];
        o.point = v3of(in.point); o.normal = v3of(in.normal); o.tangent = v3of(in.tangent); o.bitangent = v3of(in.bitangent);
        o.restitution = in.restitution; o.static_friction = in.static_friction; o.dynamic_friction = in.dynamic_friction;
        o.node_radius = in.node_radius; o.reach_capped = in.reach_capped; o.footprint_count = in.footprint_count;
        for (unsigned f = 0; f < kMaxFootprints; ++f)
            o.footprints[f] = {in.footprints[f].center_t, in.footprints[f].center_b, in.footprints[f].half_t, in.footprints[f].half_b};
    }
    return s;
}

SphereState<double> sphereFromC(const blx_sphere &c) {
    return {v3of(c.center), v3of(c.velocity), v3of(c.angular_velocity), c.radius, c.mass, c.inertia};
}
void sphereToC(const SphereState<double> &s, blx_sphere &c) {
    of3(s.center, c.center); of3(s.velocity, c.velocity); of3(s.angular_velocity, c.angular_velocity);
    c.radius = s.radius; c.mass = s.mass; c.inertia = s.inertia;
}

void statusToC(const RunStatus &s, blx_run_status &c) {
    std::memset(&c, 0, sizeof(c));
    c.total_steps = s.total_steps; c.first_failure_step = s.first_failure_step; c.last_failure_step = s.last_failure_step;
    c.failure_rounds = s.failure_rounds; c.broken_bonds = s.broken_bonds; c.removed_energy_j = s.removed_energy_j;
    c.last_energy_gain_step = s.last_energy_gain_step; c.max_damage = s.max_damage; c.last_damage_gain_step = s.last_damage_gain_step;
    c.contact.impulse_contacts = s.contact.impulse_contacts;
    c.contact.impulse_to_material[0] = s.contact.impulse_to_material_x;
    c.contact.impulse_to_material[1] = s.contact.impulse_to_material_y;
    c.contact.impulse_to_material[2] = s.contact.impulse_to_material_z;
    c.contact.dissipated_kinetic_energy_j = s.contact.dissipated_kinetic_energy_j;
    c.contact.maximum_penetration_m = s.contact.maximum_penetration_m;
    c.contact.maximum_position_correction_m = s.contact.maximum_position_correction_m;
    c.contact.maximum_center_shift_m = s.contact.maximum_center_shift_m;
    c.contact.candidate_overflow = s.contact.candidate_overflow;
    c.contact.ball_support_events = s.contact.ball_support_events;
    const ExternalLoadLedger &L = s.external_load;
    c.external_load.steps = L.steps; c.external_load.elapsed_s = L.elapsed_s;
    c.external_load.requested_impulse_n_s[0] = L.requested_impulse_n_s.x;
    c.external_load.requested_impulse_n_s[1] = L.requested_impulse_n_s.y;
    c.external_load.requested_impulse_n_s[2] = L.requested_impulse_n_s.z;
    c.external_load.impulse_n_s[0] = L.impulse_n_s.x; c.external_load.impulse_n_s[1] = L.impulse_n_s.y;
    c.external_load.impulse_n_s[2] = L.impulse_n_s.z;
    c.external_load.requested_angular_impulse_kg_m2_s[0] = L.requested_angular_impulse_kg_m2_s.x;
    c.external_load.requested_angular_impulse_kg_m2_s[1] = L.requested_angular_impulse_kg_m2_s.y;
    c.external_load.requested_angular_impulse_kg_m2_s[2] = L.requested_angular_impulse_kg_m2_s.z;
    c.external_load.angular_impulse_kg_m2_s[0] = L.angular_impulse_kg_m2_s.x;
    c.external_load.angular_impulse_kg_m2_s[1] = L.angular_impulse_kg_m2_s.y;
    c.external_load.angular_impulse_kg_m2_s[2] = L.angular_impulse_kg_m2_s.z;
    c.external_load.work_j = L.work_j;
    c.damping_dissipated_j = s.damping_dissipated_j; c.striker_dissipated_j = s.striker_dissipated_j;
    c.energy_audited = s.energy_audited; c.exit_reason = s.exit_reason; c.launches = s.launches;
    c.kernel_seconds = s.kernel_seconds; c.wall_seconds = s.wall_seconds; c.frames_captured = s.frames_captured;
    for (unsigned k = 0; k < BLX_PHASE_COUNT; ++k) c.phase_seconds[k] = s.phase_seconds[k];
    c.max_tensile_stretch = s.max_tensile_stretch; c.max_compressive_strain = s.max_compressive_strain;
    c.max_shear_strain = s.max_shear_strain; c.rank_deficient_nodes = s.rank_deficient_nodes;
    c.plastic_work_j = s.plastic_work_j; c.max_plastic_stretch = s.max_plastic_stretch;
    const NodeContactAccumulators &n = s.node_contact;
    c.node_contact.contacts = n.contacts; c.node_contact.pair_tests = n.pair_tests; c.node_contact.rebuilds = n.rebuilds;
    c.node_contact.pairs_listed = n.pairs_listed; c.node_contact.pair_overflow = n.pair_overflow;
    c.node_contact.dissipated_kinetic_energy_j = n.dissipated_kinetic_energy_j;
    c.node_contact.maximum_overlap_m = n.maximum_overlap_m;
    c.node_contact.maximum_position_correction_m = n.maximum_position_correction_m;
    c.node_contact.momentum_residual[0] = n.momentum_residual_x;
    c.node_contact.momentum_residual[1] = n.momentum_residual_y;
    c.node_contact.momentum_residual[2] = n.momentum_residual_z;
    c.shared_memory_positions = s.shared_memory_positions ? 1 : 0;
    c.shared_memory_bytes = s.shared_memory_bytes;
}

struct Handle {
    std::unique_ptr<LatticeBackend> backend;
    std::uint32_t node_count{};
    std::vector<FrameCapture> frames;
};

} // namespace

extern "C" {

const char *ref_last_error(void) { return g_error.c_str(); }

int ref_schedule_build(const blx_asset_view *a, const uint32_t *nbb, uint32_t blocks, blx_schedule_view *out) {
    return guard([&] {
        const LatticeAsset asset = makeAsset(*a);
        const LatticeSchedule s = buildLatticeSchedule(asset, std::span<const std::uint32_t>(nbb, blocks + 1));
        out->node_count = a->node_count; out->bond_count = a->bond_count; out->block_count = s.block_count;
        out->color_count = s.color_count; out->boundary_bond_count = s.boundary_bond_count;
        put(s.node_block_begin, out->node_block_begin); put(s.bond_order, out->bond_order);
        put(s.bond_schedule_index, out->bond_schedule_index); put(s.bond_block_begin, out->bond_block_begin);
        put(s.range_begin, out->range_begin); put(s.range_end, out->range_end); put(s.bond_color, out->bond_color);
    });
}

// Sizing helper only (the reference sizes its own vectors): the largest node degree.
int ref_state_max_degree(const blx_asset_view *a, uint32_t *d) {
    return guard([&] {
        std::vector<std::uint32_t> deg(a->node_count, 0);
        for (std::uint32_t j = 0; j < a->bond_count; ++j) { ++deg[a->bond_a[j]]; ++deg[a->bond_b[j]]; }
        std::uint32_t m = 0;
        for (const auto x : deg) m = std::max(m, x);
        *d = m;
    });
}

int ref_state_build(const blx_asset_view *a, const blx_schedule_view *sv, const double origin[3], blx_state_view *o) {
    return guard([&] {
        const LatticeAsset asset = makeAsset(*a);
        ActiveMatter matter;
        fillMatter(matter, asset, *a);
        const LatticeState s = buildLatticeState(matter, scheduleFromView(*sv), {origin[0], origin[1], origin[2]});
        of3({s.origin.x, s.origin.y, s.origin.z}, o->origin);
        o->node_count = s.node_count; o->bond_count = s.bond_count; o->max_degree = s.max_degree;
        put(s.x0, o->x0); put(s.u, o->u); put(s.u_prev, o->u_prev); put(s.v, o->v); put(s.inv_mass, o->inv_mass);
        put(s.mass, o->mass); put(s.adj_offsets, o->adj_offsets); put(s.adj_bonds, o->adj_bonds);
        put(s.nbr_bond, o->nbr_bond); put(s.nbr_other, o->nbr_other); put(s.nbr_rest, o->nbr_rest);
        put(s.nbr_weight, o->nbr_weight); put(s.nbr_alive, o->nbr_alive); put(s.bond_slot_a, o->bond_slot_a);
        put(s.bond_slot_b, o->bond_slot_b); put(s.bond_a, o->bond_a); put(s.bond_b, o->bond_b);
        put(s.rest_edge, o->rest_edge); put(s.rest_length, o->rest_length);
        put(s.rest_length_sq_minus, o->rest_length_sq_minus); put(s.weight, o->weight);
        put(s.compliance, o->compliance); put(s.threshold, o->threshold); put(s.alive, o->alive);
        put(s.failure_mode, o->failure_mode); put(s.damage, o->damage); put(s.prev_tensile, o->prev_tensile);
        put(s.prev_compressive, o->prev_compressive); put(s.prev_shear, o->prev_shear);
        put(s.plastic_extension, o->plastic_extension); put(s.plastic_strain, o->plastic_strain);
    });
}

int ref_box_slabs(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t horizon, uint32_t requested, uint32_t *out, uint32_t *blocks) {
    return guard([&] {
        const auto s = boxLatticeSlabs({nx, ny, nz}, horizon, requested);
        *blocks = std::uint32_t(s.size() - 1);
        if (out) put(s, out);
    });
}

int ref_state_substep_limit(const blx_state_view *v, double *out) {
    return guard([&] { *out = latticeStateSubstepLimit(stateFromView(*v)); });
}
int ref_state_elastic_energy(const blx_state_view *v, double *out) {
    return guard([&] { *out = latticeStateElasticEnergy(stateFromView(*v)); });
}
int ref_state_kinetic_energy(const blx_state_view *v, double *out) {
    return guard([&] { *out = latticeStateKineticEnergy(stateFromView(*v)); });
}
int ref_components(uint32_t N, uint32_t B, const uint32_t *ba, const uint32_t *bb, const uint8_t *alive, uint32_t *label, uint32_t *count) {
    return guard([&] {
        LatticeAsset asset;
        asset.nodes.resize(N);
        asset.bonds.resize(B);
        for (std::uint32_t j = 0; j < B; ++j) { asset.bonds[j].node_a = ba[j]; asset.bonds[j].node_b = bb[j]; }
        ActiveMatter m;
        m.asset = &asset;
        m.nodes.resize(N);
        m.bonds.resize(B);
        for (std::uint32_t j = 0; j < B; ++j) m.bonds[j].alive = alive[j] != 0;
        const auto comps = findConnectedComponents(m);
        for (const auto &c : comps) for (std::uint32_t n : c.node_indices) label[n] = c.id;
        *count = std::uint32_t(comps.size());
    });
}

// buildFragmentRepresentations over the components of `label` (canonical ids
// 0 .. components-1, each component's nodes ascending), on a matter of N cells
// with grid coordinates, world positions, velocities, masses and spins (NULL:
// none), the asset's voxel size and the material's density. settings:
// first_body_id, maximum_rigid_fragments, minimum_nodes_per_rigid_fragment,
// maximum_collision_points, friction, restitution, rolling_resistance.
// The result is flattened into doubles (written up to capacity; *size is the
// total):
//   rigid count, debris count, total mass, rigid mass, debris mass,
//   total linear momentum (3), total angular momentum about the origin (3),
//   coarsening kinetic loss;
//   per rigid fragment: body id, source node count, mass, centre of mass (3),
//   face count, vertex count, index count, collision point count, voxel centre
//   count, vertices (position 3, normal 3 each), indices, points (3 each),
//   voxel centres (3 each), friction, restitution, rolling resistance, voxel size;
//   per debris particle: position (3), velocity (3), angular velocity (3),
//   mass, radius, inertia (9).
int ref_fragment_build(uint32_t N, const int32_t *grid, const double *pos, const double *vel, const double *mass,
                       const double *spin, double cell, double density, const uint32_t *label, uint32_t components,
                       const double *settings, double *out, uint64_t capacity, uint64_t *size) {
    return guard([&] {
        LatticeAsset asset;
        asset.recipe.voxel_size_m = cell;
        asset.nodes.resize(N);
        ActiveMatter m;
        m.asset = &asset;
        m.material.density_kg_m3 = density;
        m.nodes.resize(N);
        for (std::uint32_t i = 0; i < N; ++i) {
            asset.nodes[i].grid = {grid[3 * i], grid[3 * i + 1], grid[3 * i + 2]};
            ActiveNodeState &n = m.nodes[i];
            n.position_world_m = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
            n.velocity_m_s = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
            n.mass_kg = mass[i];
            if (spin) n.spin_angular_velocity_rad_s = {spin[3 * i], spin[3 * i + 1], spin[3 * i + 2]};
        }
        std::vector<FragmentComponent> comps(components);
        for (std::uint32_t f = 0; f < components; ++f) comps[f].id = f;
        for (std::uint32_t i = 0; i < N; ++i) {
            if (label[i] >= components) throw std::invalid_argument("label out of range");
            comps[label[i]].node_indices.push_back(i);
        }
        FragmentBuildSettings fs;
        fs.first_body_id = MatterBodyId(settings[0]);
        fs.maximum_rigid_fragments = std::size_t(settings[1]);
        fs.minimum_nodes_per_rigid_fragment = std::size_t(settings[2]);
        fs.maximum_collision_points = std::size_t(settings[3]);
        fs.friction = settings[4];
        fs.restitution = settings[5];
        fs.rolling_resistance = settings[6];
        const FragmentBuildResult r = buildFragmentRepresentations(m, comps, fs);
        std::vector<double> o;
        auto v3 = [&](const Vec3 &v) { o.insert(o.end(), {v.x, v.y, v.z}); };
        o.push_back(double(r.rigid_fragments.size()));
        o.push_back(double(r.debris_particles.size()));
        o.insert(o.end(), {r.total_mass_kg, r.rigid_fragment_mass_kg, r.debris_mass_kg});
        v3(r.total_linear_momentum_kg_m_s);
        v3(r.total_angular_momentum_about_origin_kg_m2_s);
        o.push_back(r.coarsening_kinetic_loss_j);
        for (const RigidFragmentDescription &f : r.rigid_fragments) {
            o.insert(o.end(), {double(f.body_id), double(f.source_node_count), f.mass_properties.mass_kg});
            v3(f.mass_properties.center_of_mass_world_m);
            o.insert(o.end(), {double(f.surface_mesh.exposed_face_count), double(f.surface_mesh.vertices.size()),
                               double(f.surface_mesh.indices.size()), double(f.collision_points_local_m.size()),
                               double(f.voxel_centers_local_m.size())});
            for (const SurfaceVertex &v : f.surface_mesh.vertices) {
                v3(v.position_local_m);
                v3(v.normal_local);
            }
            for (std::uint32_t k : f.surface_mesh.indices) o.push_back(double(k));
            for (const Vec3 &p : f.collision_points_local_m) v3(p);
            for (const Vec3 &p : f.voxel_centers_local_m) v3(p);
            o.insert(o.end(), {f.friction, f.restitution, f.rolling_resistance, f.voxel_size_m});
        }
        for (const DebrisParticleDescription &d : r.debris_particles) {
            v3(d.position_world_m);
            v3(d.velocity_m_s);
            v3(d.angular_velocity_rad_s);
            o.insert(o.end(), {d.mass_kg, d.radius_m});
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) o.push_back(d.inertia_world_kg_m2.m[a][b]);
        }
        *size = o.size();
        if (out) std::memcpy(out, o.data(), std::min<std::uint64_t>(capacity, o.size()) * sizeof(double));
    });
}

// buildFragmentLattice on a parent matter built from the asset view (reference
// positions as the asset's and the matter's reference positions, masses, bonds
// and bond state); fills the same view the product's blx_fragment_lattice does.
int ref_fragment_lattice(const blx_asset_view *a, const uint32_t *nodes, uint32_t count, const double *offset,
                         const blx_fragment_pose *pose, const double *pext, const double *pstr,
                         blx_fragment_lattice_view *out) {
    return guard([&] {
        const LatticeAsset asset = makeAsset(*a);
        ActiveMatter m;
        fillMatter(m, asset, *a);
        std::vector<Vec3> off(a->node_count);
        for (std::uint32_t i = 0; i < a->node_count; ++i) off[i] = {offset[3 * i], offset[3 * i + 1], offset[3 * i + 2]};
        FragmentPose p;
        p.center_of_mass_world_m = {pose->center_of_mass[0], pose->center_of_mass[1], pose->center_of_mass[2]};
        p.orientation_world = {pose->orientation_wxyz[0], pose->orientation_wxyz[1], pose->orientation_wxyz[2],
                               pose->orientation_wxyz[3]};
        p.linear_velocity_m_s = {pose->linear_velocity[0], pose->linear_velocity[1], pose->linear_velocity[2]};
        p.angular_velocity_rad_s = {pose->angular_velocity[0], pose->angular_velocity[1], pose->angular_velocity[2]};
        const std::vector<double> ext = pext ? std::vector<double>(pext, pext + a->bond_count) : std::vector<double>{};
        const std::vector<double> str = pstr ? std::vector<double>(pstr, pstr + a->bond_count) : std::vector<double>{};
        const FragmentLattice f = buildFragmentLattice(m, std::span<const std::uint32_t>(nodes, count), off, p, ext, str);
        const std::uint32_t n = std::uint32_t(f.parent_node.size()), k = std::uint32_t(f.parent_bond.size());
        out->node_count = n;
        out->bond_count = k;
        if (out->node_capacity < n || out->bond_capacity < k) return;
        out->origin[0] = f.origin.x; out->origin[1] = f.origin.y; out->origin[2] = f.origin.z;
        out->carried_spin[0] = f.carried_spin_rad_s.x; out->carried_spin[1] = f.carried_spin_rad_s.y;
        out->carried_spin[2] = f.carried_spin_rad_s.z;
        for (std::uint32_t i = 0; i < n; ++i) {
            out->parent_node[i] = f.parent_node[i];
            const Vec3 r = f.matter.reference_positions_world_m[i], x = f.matter.nodes[i].position_world_m,
                       v = f.matter.nodes[i].velocity_m_s;
            out->reference_position[3 * i] = r.x; out->reference_position[3 * i + 1] = r.y; out->reference_position[3 * i + 2] = r.z;
            out->position[3 * i] = x.x; out->position[3 * i + 1] = x.y; out->position[3 * i + 2] = x.z;
            out->velocity[3 * i] = v.x; out->velocity[3 * i + 1] = v.y; out->velocity[3 * i + 2] = v.z;
            out->mass[i] = f.matter.nodes[i].mass_kg;
        }
        for (std::uint32_t j = 0; j < k; ++j) {
            const BondRest &b = f.asset->bonds[j];
            out->parent_bond[j] = f.parent_bond[j];
            out->bond_a[j] = b.node_a;
            out->bond_b[j] = b.node_b;
            out->rest_length[j] = b.rest_length_m;
            out->compliance[j] = b.compliance;
            const double t[6] = {b.damage_start_stretch, b.damage_end_stretch, b.compression_damage_start_strain,
                                 b.compression_damage_end_strain, b.shear_damage_start_strain, b.shear_damage_end_strain};
            std::memcpy(out->threshold + 6 * std::size_t(j), t, sizeof t);
            out->alive[j] = f.matter.bonds[j].alive ? 1 : 0;
            out->damage[j] = f.matter.bonds[j].damage;
            out->failure_mode[j] = std::uint8_t(f.matter.bonds[j].failure_mode);
            out->plastic_extension[j] = f.plastic_extension_m[j];
            out->plastic_strain[j] = f.plastic_strain_m[j];
        }
    });
}

int ref_merge_cells_into_boxes(const int32_t *cells, uint64_t count, int32_t *boxes, uint64_t capacity,
                               uint64_t *box_count) {
    return guard([&] {
        std::vector<GridCoord> c(count);
        for (uint64_t k = 0; k < count; ++k) c[k] = {cells[3 * k], cells[3 * k + 1], cells[3 * k + 2]};
        const std::vector<VoxelBox> b = mergeCellsIntoBoxes(c);
        *box_count = b.size();
        for (uint64_t k = 0; k < std::min<uint64_t>(capacity, b.size()); ++k) {
            const int32_t v[6] = {b[k].minimum.x, b[k].minimum.y, b[k].minimum.z, b[k].maximum.x, b[k].maximum.y,
                                  b[k].maximum.z};
            std::memcpy(boxes + 6 * k, v, sizeof v);
        }
    });
}

int ref_fragment_mass(uint32_t N, const double *pos, const double *vel, const double *spin, const double *mass,
                      double cell, uint32_t K, const uint32_t *offsets, const uint32_t *nodes, blx_fragment_mass *out) {
    return guard([&] {
        LatticeAsset asset;
        asset.recipe.voxel_size_m = cell;
        asset.nodes.resize(N);
        ActiveMatter m;
        m.asset = &asset;
        m.nodes.resize(N);
        for (std::uint32_t i = 0; i < N; ++i) {
            ActiveNodeState &n = m.nodes[i];
            n.position_world_m = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
            n.velocity_m_s = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
            n.mass_kg = mass[i];
            if (spin) n.spin_angular_velocity_rad_s = {spin[3 * i], spin[3 * i + 1], spin[3 * i + 2]};
        }
        for (std::uint32_t f = 0; f < K; ++f) {
            const std::span<const std::uint32_t> idx(nodes + offsets[f], offsets[f + 1] - offsets[f]);
            const FragmentMassProperties p = calculateFragmentMassProperties(m, idx);
            blx_fragment_mass &o = out[f];
            o = blx_fragment_mass{};
            o.mass_kg = p.mass_kg;
            const Vec3 *vs[5] = {&p.center_of_mass_world_m, &p.linear_velocity_m_s, &p.angular_momentum_kg_m2_s,
                                 &p.angular_velocity_rad_s, nullptr};
            double *dst[4] = {o.center_of_mass_world_m, o.linear_velocity_m_s, o.angular_momentum_kg_m2_s,
                              o.angular_velocity_rad_s};
            for (int q = 0; q < 4; ++q) {
                dst[q][0] = vs[q]->x;
                dst[q][1] = vs[q]->y;
                dst[q][2] = vs[q]->z;
            }
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) o.inertia_world_kg_m2[3 * r + c] = p.inertia_world_kg_m2.m[r][c];
            o.source_kinetic_energy_j = p.source_kinetic_energy_j;
            o.rigid_kinetic_energy_j = p.rigid_kinetic_energy_j;
            o.coarsening_kinetic_loss_j = p.coarsening_kinetic_loss_j;
            o.node_count = std::uint32_t(idx.size());
            o.first_node = idx.empty() ? 0U : idx[0];
        }
    });
}

// kind 0: serial CpuLatticeBackend; kind 1: ParallelCpuLatticeBackend with `threads`.
int ref_backend_create(int precision, const blx_schedule_view *sv, int kind, unsigned threads, void **out) {
    return guard([&] {
        auto h = std::make_unique<Handle>();
        const Precision p = precision == BLX_PRECISION_FLOAT ? Precision::Float : Precision::Double;
        const LatticeSchedule s = scheduleFromView(*sv);
        h->backend = kind == 0 ? makeCpuLatticeBackend(s, p) : makeParallelCpuLatticeBackend(s, p, threads);
        h->node_count = sv->node_count;
        *out = h.release();
    });
}
int ref_backend_destroy(void *h) { return guard([&] { delete static_cast<Handle *>(h); }); }
int ref_backend_upload(void *h, const blx_state_view *st, const blx_settings *s, const blx_sphere *sp) {
    return guard([&] { static_cast<Handle *>(h)->backend->upload(stateFromView(*st), settingsFromC(*s), sphereFromC(*sp)); });
}
int ref_backend_set_external_forces(void *h, const double *f, uint64_t substeps) {
    return guard([&] {
        Handle *hh = static_cast<Handle *>(h);
        std::vector<Vec3> forces;
        if (f) {
            forces.resize(hh->node_count);
            for (std::uint32_t i = 0; i < hh->node_count; ++i) forces[i] = {f[3 * i], f[3 * i + 1], f[3 * i + 2]};
        }
        hh->backend->setExternalForces(forces, substeps);
    });
}
int ref_backend_run(void *h, const blx_run_control *c, blx_run_status *st) {
    return guard([&] {
        RunControl rc{};
        rc.max_steps = c->max_steps; rc.quiet_steps = c->quiet_steps; rc.min_steps = c->min_steps;
        rc.no_failure_steps = c->no_failure_steps; rc.energy_flat_steps = c->energy_flat_steps;
        rc.energy_flat_fraction = c->energy_flat_fraction; rc.calm_steps = c->calm_steps;
        rc.calm_damage_margin = c->calm_damage_margin; rc.removable_energy_j = c->removable_energy_j;
        rc.capture_stride = c->capture_stride; rc.max_frames = c->max_frames; rc.steps_per_launch = c->steps_per_launch;
        const RunStatus s = static_cast<Handle *>(h)->backend->run(rc);
        statusToC(s, *st);
    });
}
int ref_backend_status(void *h, blx_run_status *st) { return guard([&] { statusToC(static_cast<Handle *>(h)->backend->status(), *st); }); }
int ref_backend_download(void *h, blx_state_view *o, blx_sphere *sp) {
    return guard([&] {
        LatticeState s;
        SphereState<double> sphere{};
        static_cast<Handle *>(h)->backend->download(s, sphere);
        put(s.u, o->u); put(s.u_prev, o->u_prev); put(s.v, o->v); put(s.alive, o->alive);
        put(s.failure_mode, o->failure_mode); put(s.damage, o->damage); put(s.prev_tensile, o->prev_tensile);
        put(s.prev_compressive, o->prev_compressive); put(s.prev_shear, o->prev_shear);
        put(s.plastic_extension, o->plastic_extension); put(s.plastic_strain, o->plastic_strain);
        sphereToC(sphere, *sp);
    });
}
int ref_backend_first_failure_bonds(void *h, uint32_t *out, uint64_t cap, uint64_t *count) {
    return guard([&] {
        const auto v = static_cast<Handle *>(h)->backend->firstFailureBonds();
        *count = v.size();
        for (std::size_t k = 0; k < v.size() && k < cap; ++k) out[k] = v[k];
    });
}
int ref_backend_frame_count(void *h, uint32_t *count) {
    return guard([&] {
        Handle *hh = static_cast<Handle *>(h);
        auto fresh = hh->backend->takeFrames();
        for (auto &f : fresh) hh->frames.push_back(std::move(f));
        *count = std::uint32_t(hh->frames.size());
    });
}
int ref_backend_take_frame(void *h, uint32_t index, uint64_t *step, float *u, uint8_t *alive, float *damage, blx_sphere *sp) {
    return guard([&] {
        Handle *hh = static_cast<Handle *>(h);
        if (index >= hh->frames.size()) throw std::invalid_argument("frame index out of range");
        const FrameCapture &f = hh->frames[index];
        *step = f.step;
        put(f.u, u); put(f.alive, alive); put(f.damage, damage);
        if (sp) sphereToC(f.sphere, *sp);
    });
}
int ref_backend_clear_frames(void *h) { return guard([&] { static_cast<Handle *>(h)->frames.clear(); }); }

} // extern "C"
