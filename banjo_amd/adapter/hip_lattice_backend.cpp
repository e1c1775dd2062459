// The GPU lane of the reference's lattice API on MI355X, through the C ABI of
// include/banjo_lattice.h. A maintainer replaces
// src/fastlattice/CudaLatticeUnavailable.cpp (which defines the four functions
// of FastLattice.hpp:401-406 as "no device") with this file, adds
// banjo_amd/adapter and include to the include path, and links
// libbanjo_lattice.so. No caller changes: makeBackend for BackendKind::Cuda
// (TileImpactScene.cpp:345-351), LiveWorld (LiveWorld.cpp:12637) and the tests
// (fast_lattice_tests.cpp:410-412) reach the device through these functions.
//
// The backend implements LatticeBackend (FastLattice.hpp:255-369) by copying
// the reference's value types into the C views and calling blx_* entry points.
// The virtuals it does not override keep the base class's throwing defaults,
// as the reference's GPU lane did: wrenches, reversible trials, external
// point contact and the coupled contact steps are the serial double CPU
// lane's. Errors come back as BLX_ERR_* codes and are rethrown as the
// exception types the callers already catch.

#include "hip_lattice_backend.hpp"

#include <atomic>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace banjo::fastlattice {

namespace hip_adapter {

namespace {

void copy3(double *out, const Vec3 &v) {
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}
void copy3(double *out, const V3<double> &v) {
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}
V3<double> v3(const double *in) { return V3<double>{in[0], in[1], in[2]}; }
Vec3 vec3(const double *in) { return Vec3{in[0], in[1], in[2]}; }

int pollAtomic(const void *context) {
    return static_cast<const std::atomic<bool> *>(context)->load(std::memory_order_acquire) ? 1 : 0;
}

void ledgerBack(const blx_load_ledger &in, ExternalLoadLedger &out) {
    out.steps = in.steps;
    out.elapsed_s = in.elapsed_s;
    out.requested_impulse_n_s = vec3(in.requested_impulse_n_s);
    out.impulse_n_s = vec3(in.impulse_n_s);
    out.requested_angular_impulse_kg_m2_s = vec3(in.requested_angular_impulse_kg_m2_s);
    out.angular_impulse_kg_m2_s = vec3(in.angular_impulse_kg_m2_s);
    out.work_j = in.work_j;
}

} // namespace

blx_schedule_view scheduleView(const LatticeSchedule &s) {
    blx_schedule_view v{};
    v.node_count = s.node_block_begin.empty() ? 0U : s.node_block_begin.back();
    v.bond_count = std::uint32_t(s.bond_order.size());
    v.block_count = s.block_count;
    v.color_count = s.color_count;
    v.boundary_bond_count = s.boundary_bond_count;
    auto p = [](const std::vector<std::uint32_t> &vec) { return const_cast<std::uint32_t *>(vec.data()); };
    v.node_block_begin = p(s.node_block_begin);
    v.bond_order = p(s.bond_order);
    v.bond_schedule_index = p(s.bond_schedule_index);
    v.bond_block_begin = p(s.bond_block_begin);
    v.range_begin = p(s.range_begin);
    v.range_end = p(s.range_end);
    v.bond_color = const_cast<std::uint8_t *>(s.bond_color.data());
    return v;
}

blx_state_view stateView(LatticeState &st) {
    blx_state_view v{};
    copy3(v.origin, st.origin);
    v.node_count = st.node_count;
    v.bond_count = st.bond_count;
    v.max_degree = st.max_degree;
    v.x0 = st.x0.data();
    v.u = st.u.data();
    v.u_prev = st.u_prev.data();
    v.v = st.v.data();
    v.inv_mass = st.inv_mass.data();
    v.mass = st.mass.data();
    v.adj_offsets = st.adj_offsets.data();
    v.adj_bonds = st.adj_bonds.data();
    v.nbr_bond = st.nbr_bond.data();
    v.nbr_other = st.nbr_other.data();
    v.nbr_rest = st.nbr_rest.data();
    v.nbr_weight = st.nbr_weight.data();
    v.nbr_alive = st.nbr_alive.data();
    v.bond_slot_a = st.bond_slot_a.data();
    v.bond_slot_b = st.bond_slot_b.data();
    v.bond_a = st.bond_a.data();
    v.bond_b = st.bond_b.data();
    v.rest_edge = st.rest_edge.data();
    v.rest_length = st.rest_length.data();
    v.rest_length_sq_minus = st.rest_length_sq_minus.data();
    v.weight = st.weight.data();
    v.compliance = st.compliance.data();
    v.threshold = st.threshold.data();
    v.alive = st.alive.data();
    v.failure_mode = st.failure_mode.data();
    v.damage = st.damage.data();
    v.prev_tensile = st.prev_tensile.data();
    v.prev_compressive = st.prev_compressive.data();
    v.prev_shear = st.prev_shear.data();
    v.plastic_extension = st.plastic_extension.data();
    v.plastic_strain = st.plastic_strain.data();
    return v;
}

blx_settings settingsView(const StepSettings<double> &s) {
    blx_settings o{};
    o.dt = s.dt;
    copy3(o.gravity, s.gravity);
    o.constraint_iterations = s.constraint_iterations;
    o.bond_integrator = s.bond_integrator;
    o.damping_fraction = s.damping_fraction;
    o.sphere_enabled = s.sphere_enabled;
    o.direct_arithmetic = s.direct_arithmetic;
    o.audit_energy = s.audit_energy;
    o.plastic_yield_stretch = s.plastic_yield_stretch;
    o.plastic_hardening = s.plastic_hardening;
    o.plate_half_thickness = s.plate_half_thickness;
    o.contact.static_friction = s.contact.static_friction;
    o.contact.dynamic_friction = s.contact.dynamic_friction;
    o.contact.restitution = s.contact.restitution;
    o.contact.restitution_speed_threshold = s.contact.restitution_speed_threshold;
    o.contact.node_contact_radius = s.contact.node_contact_radius;
    o.contact.contact_margin = s.contact.contact_margin;
    o.contact.prefilter_slack = s.contact.prefilter_slack;
    o.node_contact.mode = s.node_contact.mode;
    o.node_contact.radius = s.node_contact.radius;
    o.node_contact.skin = s.node_contact.skin;
    o.node_contact.margin = s.node_contact.margin;
    o.node_contact.restitution = s.node_contact.restitution;
    o.node_contact.restitution_speed_threshold = s.node_contact.restitution_speed_threshold;
    o.node_contact.static_friction = s.node_contact.static_friction;
    o.node_contact.dynamic_friction = s.node_contact.dynamic_friction;
    o.node_contact.bucket_mask = s.node_contact.bucket_mask;
    o.plane_count = s.support.plane_count;
    for (unsigned p = 0; p < kMaxSupportPlanes; ++p) {
        const SupportPlane<double> &in = s.support.planes[p];
        blx_support_plane &pl = o.planes[p];
        copy3(pl.point, in.point);
        copy3(pl.normal, in.normal);
        copy3(pl.tangent, in.tangent);
        copy3(pl.bitangent, in.bitangent);
        pl.restitution = in.restitution;
        pl.static_friction = in.static_friction;
        pl.dynamic_friction = in.dynamic_friction;
        pl.node_radius = in.node_radius;
        pl.reach_capped = in.reach_capped;
        pl.footprint_count = in.footprint_count;
        for (unsigned f = 0; f < kMaxFootprints; ++f) {
            pl.footprints[f].center_t = in.footprints[f].center_t;
            pl.footprints[f].center_b = in.footprints[f].center_b;
            pl.footprints[f].half_t = in.footprints[f].half_t;
            pl.footprints[f].half_b = in.footprints[f].half_b;
        }
    }
    return o;
}

blx_sphere sphereView(const SphereState<double> &s) {
    blx_sphere o{};
    copy3(o.center, s.center);
    copy3(o.velocity, s.velocity);
    copy3(o.angular_velocity, s.angular_velocity);
    o.radius = s.radius;
    o.mass = s.mass;
    o.inertia = s.inertia;
    return o;
}

void sphereBack(const blx_sphere &in, SphereState<double> &out) {
    out.center = v3(in.center);
    out.velocity = v3(in.velocity);
    out.angular_velocity = v3(in.angular_velocity);
    out.radius = in.radius;
    out.mass = in.mass;
    out.inertia = in.inertia;
}

blx_run_control controlView(const RunControl &c) {
    blx_run_control o{};
    o.max_steps = c.max_steps;
    o.quiet_steps = c.quiet_steps;
    o.min_steps = c.min_steps;
    o.no_failure_steps = c.no_failure_steps;
    o.energy_flat_steps = c.energy_flat_steps;
    o.energy_flat_fraction = c.energy_flat_fraction;
    o.calm_steps = c.calm_steps;
    o.calm_damage_margin = c.calm_damage_margin;
    o.removable_energy_j = c.removable_energy_j;
    o.capture_stride = c.capture_stride;
    o.max_frames = c.max_frames;
    o.steps_per_launch = c.steps_per_launch;
    o.abandon = nullptr;
    if (c.abandon) {
        o.abandon_poll = &pollAtomic;
        o.abandon_context = c.abandon;
    }
    return o;
}

void statusBack(const blx_run_status &in, RunStatus &out) {
    out = RunStatus{};
    out.bond_integrator = kBondXpbd;
    ledgerBack(in.external_load, out.external_load);
    out.total_steps = in.total_steps;
    out.first_failure_step = in.first_failure_step;
    out.last_failure_step = in.last_failure_step;
    out.failure_rounds = in.failure_rounds;
    out.broken_bonds = in.broken_bonds;
    out.removed_energy_j = in.removed_energy_j;
    out.last_energy_gain_step = in.last_energy_gain_step;
    out.max_damage = in.max_damage;
    out.last_damage_gain_step = in.last_damage_gain_step;
    out.contact.impulse_contacts = in.contact.impulse_contacts;
    out.contact.impulse_to_material_x = in.contact.impulse_to_material[0];
    out.contact.impulse_to_material_y = in.contact.impulse_to_material[1];
    out.contact.impulse_to_material_z = in.contact.impulse_to_material[2];
    out.contact.dissipated_kinetic_energy_j = in.contact.dissipated_kinetic_energy_j;
    out.contact.maximum_penetration_m = in.contact.maximum_penetration_m;
    out.contact.maximum_position_correction_m = in.contact.maximum_position_correction_m;
    out.contact.maximum_center_shift_m = in.contact.maximum_center_shift_m;
    out.contact.candidate_overflow = in.contact.candidate_overflow;
    out.contact.ball_support_events = in.contact.ball_support_events;
    out.node_contact.contacts = in.node_contact.contacts;
    out.node_contact.pair_tests = in.node_contact.pair_tests;
    out.node_contact.rebuilds = in.node_contact.rebuilds;
    out.node_contact.pairs_listed = in.node_contact.pairs_listed;
    out.node_contact.pair_overflow = in.node_contact.pair_overflow;
    out.node_contact.dissipated_kinetic_energy_j = in.node_contact.dissipated_kinetic_energy_j;
    out.node_contact.maximum_overlap_m = in.node_contact.maximum_overlap_m;
    out.node_contact.maximum_position_correction_m = in.node_contact.maximum_position_correction_m;
    out.node_contact.momentum_residual_x = in.node_contact.momentum_residual[0];
    out.node_contact.momentum_residual_y = in.node_contact.momentum_residual[1];
    out.node_contact.momentum_residual_z = in.node_contact.momentum_residual[2];
    out.damping_dissipated_j = in.damping_dissipated_j;
    out.striker_dissipated_j = in.striker_dissipated_j;
    out.energy_audited = in.energy_audited != 0;
    out.exit_reason = in.exit_reason;
    out.launches = in.launches;
    out.kernel_seconds = in.kernel_seconds;
    out.wall_seconds = in.wall_seconds;
    out.frames_captured = in.frames_captured;
    for (unsigned p = 0; p < kPhaseCount && p < BLX_PHASE_COUNT; ++p) out.phase_seconds[p] = in.phase_seconds[p];
    out.max_tensile_stretch = in.max_tensile_stretch;
    out.max_compressive_strain = in.max_compressive_strain;
    out.max_shear_strain = in.max_shear_strain;
    out.rank_deficient_nodes = in.rank_deficient_nodes;
    out.plastic_work_j = in.plastic_work_j;
    out.max_plastic_stretch = in.max_plastic_stretch;
    out.shared_memory_positions = in.shared_memory_positions != 0;
    out.shared_memory_bytes = std::size_t(in.shared_memory_bytes);
}

} // namespace hip_adapter

namespace {

using namespace hip_adapter;

// BLX_ERR_* -> the exception types the reference's callers catch.
void check(int rc) {
    if (rc == BLX_OK) return;
    const std::string msg = blx_last_error();
    switch (rc) {
    case BLX_ERR_INVALID: throw std::invalid_argument(msg);
    case BLX_ERR_OVERFLOW: throw std::overflow_error(msg);
    case BLX_ERR_LOGIC: throw std::logic_error(msg);
    default: throw std::runtime_error(msg);
    }
}

class HipLatticeBackend final : public LatticeBackend {
public:
    HipLatticeBackend(const LatticeSchedule &schedule, Precision precision) {
        const blx_schedule_view v = scheduleView(schedule);
        nodes_ = v.node_count;
        bonds_ = v.bond_count;
        check(blx_backend_create(0, precision == Precision::Float ? BLX_PRECISION_FLOAT : BLX_PRECISION_DOUBLE, &v, &h_));
    }
    ~HipLatticeBackend() override { (void)blx_backend_destroy(h_); }
    HipLatticeBackend(const HipLatticeBackend &) = delete;
    HipLatticeBackend &operator=(const HipLatticeBackend &) = delete;

    [[nodiscard]] std::string name() const override {
        char buf[64] = {};
        check(blx_backend_name(h_, buf, sizeof buf));
        return buf;
    }
    void upload(const LatticeState &state, const StepSettings<double> &settings,
                const SphereState<double> &sphere) override {
        // The view's pointers are only read by an upload.
        blx_state_view v = stateView(const_cast<LatticeState &>(state));
        const blx_settings s = settingsView(settings);
        const blx_sphere sp = sphereView(sphere);
        check(blx_backend_upload(h_, &v, &s, &sp));
        status_ = RunStatus{};
    }
    void setExternalForces(const std::vector<Vec3> &forces_world_n, std::uint64_t substeps) override {
        std::vector<double> flat(3 * forces_world_n.size());
        for (std::size_t i = 0; i < forces_world_n.size(); ++i) copy3(&flat[3 * i], forces_world_n[i]);
        check(blx_backend_set_external_forces(h_, forces_world_n.empty() ? nullptr : flat.data(), substeps));
    }
    RunStatus run(const RunControl &control) override {
        const blx_run_control c = controlView(control);
        blx_run_status st{};
        check(blx_backend_run(h_, &c, &st));
        statusBack(st, status_);
        return status_;
    }
    void download(LatticeState &state, SphereState<double> &sphere) override {
        blx_state_view v = stateView(state);
        blx_sphere sp{};
        check(blx_backend_download(h_, &v, &sp));
        sphereBack(sp, sphere);
    }
    std::vector<FrameCapture> takeFrames() override {
        std::uint32_t n = 0;
        check(blx_backend_frame_count(h_, &n));
        std::vector<FrameCapture> out(n);
        for (std::uint32_t k = 0; k < n; ++k) {
            FrameCapture &f = out[k];
            f.u.resize(3 * std::size_t(nodes_));
            f.alive.resize(bonds_);
            f.damage.resize(bonds_);
            blx_sphere sp{};
            check(blx_backend_take_frame(h_, k, &f.step, f.u.data(), f.alive.data(), f.damage.data(), &sp));
            sphereBack(sp, f.sphere);
        }
        check(blx_backend_clear_frames(h_));
        return out;
    }
    [[nodiscard]] const RunStatus &status() const override { return status_; }
    [[nodiscard]] std::vector<std::uint32_t> firstFailureBonds() const override {
        std::uint64_t n = 0;
        check(blx_backend_first_failure_bonds(h_, nullptr, 0, &n));
        std::vector<std::uint32_t> out(n);
        check(blx_backend_first_failure_bonds(h_, out.data(), n, &n));
        out.resize(n);
        return out;
    }

private:
    std::uint32_t nodes_ = 0, bonds_ = 0;
    blx_backend *h_ = nullptr;
    RunStatus status_{};
};

} // namespace

// FastLattice.hpp:401-406. threads_per_block has no meaning here (each kernel
// has its own fixed launch shape) and is ignored. No device: std::runtime_error,
// as the stub this replaces throws (CudaLatticeUnavailable.cpp:19-25).
std::unique_ptr<LatticeBackend> makeCudaLatticeBackend(const LatticeSchedule &schedule, Precision precision,
                                                       unsigned /*threads_per_block*/) {
    if (!blx_available()) throw std::runtime_error("no MI355X (gfx950) device for the lattice backend");
    return std::make_unique<HipLatticeBackend>(schedule, precision);
}

bool cudaLatticeAvailable() { return blx_available() != 0; }

std::string cudaLatticeDescription() {
    char buf[256] = {};
    return blx_description(buf, sizeof buf) == BLX_OK ? std::string(buf) : std::string("no MI355X (gfx950) device");
}

unsigned cudaLatticeMultiprocessorCount() { return blx_multiprocessor_count(); }

} // namespace banjo::fastlattice
