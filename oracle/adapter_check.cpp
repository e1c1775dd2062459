// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// Checks the reference-side adapter (banjo_amd/adapter/hip_lattice_backend.cpp)
// compiled in the reference's own terms: its headers, its value types, its
// LatticeBackend interface and its CPU backend, all linked from the reference's
// sources (oracle/Makefile; nothing from the reference is copied here).
//
//   adapter_check product  linked with the product library. On a machine with
//                          no GPU the four factory functions answer as the stub
//                          they replace does (CudaLatticeUnavailable.cpp:11-25);
//                          with an MI355X visible, the HIP backend driven
//                          through LatticeBackend& gives the reference CPU
//                          backend's answers bit for bit, as below. Either way
//                          every value type crosses the C views field for field
//                          in both directions.
//   adapter_check oracle   linked with adapter_shim.cpp (the same entry points
//                          over the oracle restatement): the adapter driven
//                          through LatticeBackend& -- upload, external loads,
//                          runs continued across calls, frames, download,
//                          first failures, status -- gives the reference CPU
//                          backend's answers bit for bit, in float and double.
//
// Prints one line per check and "adapter_check ok" at the end; any failure
// prints what differed and exits 1.

#include "hip_lattice_backend.hpp"

#include "fastlattice/FastLattice.hpp"
#include "fastlattice/LatticeSchedule.hpp"
#include "fracture/ActiveMatter.hpp"
#include "matter/Lattice.hpp"

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

using namespace banjo;
using namespace banjo::fastlattice;

namespace {

int g_failures = 0;

void expect(bool ok, const std::string &what) {
    if (!ok) {
        std::printf("FAIL %s\n", what.c_str());
        ++g_failures;
    }
}

template <class T> bool sameBits(const T &a, const T &b) { return std::memcmp(&a, &b, sizeof(T)) == 0; }
template <class T> bool sameVec(const std::vector<T> &a, const std::vector<T> &b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

// A 26-neighbour box (node index x + nx (y + ny z)), bonds node-major over the
// 13 positive half-offsets with node_a < node_b, as the box workloads use.
struct Box {
    LatticeAsset asset;
    ActiveMatter matter;
};

std::unique_ptr<Box> makeBox(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz, double h) {
    auto box = std::make_unique<Box>();
    LatticeAsset &a = box->asset;
    const std::uint32_t N = nx * ny * nz;
    a.nodes.resize(N);
    auto id = [&](std::uint32_t x, std::uint32_t y, std::uint32_t z) { return x + nx * (y + ny * z); };
    for (std::uint32_t z = 0; z < nz; ++z)
        for (std::uint32_t y = 0; y < ny; ++y)
            for (std::uint32_t x = 0; x < nx; ++x)
                a.nodes[id(x, y, z)].local_position_m = {(x - 0.5 * (nx - 1)) * h, (y - 0.5 * (ny - 1)) * h,
                                                         (z - 0.5 * (nz - 1)) * h};
    const int off[13][3] = {{1, 0, 0},  {-1, 1, 0}, {0, 1, 0},   {1, 1, 0},   {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
                            {-1, 0, 1}, {0, 0, 1},  {1, 0, 1},   {-1, 1, 1},  {0, 1, 1},   {1, 1, 1}};
    std::uint32_t k = 0;
    for (std::uint32_t i = 0; i < N; ++i) {
        const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / (nx * ny));
        for (const auto &o : off) {
            const int X = x + o[0], Y = y + o[1], Z = z + o[2];
            if (X < 0 || Y < 0 || Z < 0 || X >= int(nx) || Y >= int(ny) || Z >= int(nz)) continue;
            BondRest b;
            b.node_a = i;
            b.node_b = id(std::uint32_t(X), std::uint32_t(Y), std::uint32_t(Z));
            b.rest_length_m = std::sqrt(double(o[0] * o[0] + o[1] * o[1] + o[2] * o[2])) * h;
            // Compliance and break strain vary per bond, so a mis-copied array shows.
            b.compliance = 1e-7 * (1.0 + double((k * 7919U) % 97U));
            b.damage_start_stretch = b.damage_end_stretch = 0.004 + 0.00005 * double((k * 104729U) % 61U);
            b.compression_damage_start_strain = b.compression_damage_end_strain = std::numeric_limits<double>::infinity();
            b.shear_damage_start_strain = b.shear_damage_end_strain = std::numeric_limits<double>::infinity();
            a.bonds.push_back(b);
            ++k;
        }
    }
    ActiveMatter &m = box->matter;
    m.asset = &box->asset;
    m.nodes.resize(N);
    m.reference_positions_world_m.resize(N);
    for (std::uint32_t i = 0; i < N; ++i) {
        m.reference_positions_world_m[i] = a.nodes[i].local_position_m;
        m.nodes[i].position_world_m = a.nodes[i].local_position_m;
        m.nodes[i].previous_position_world_m = a.nodes[i].local_position_m;
        m.nodes[i].mass_kg = 1e-3;
    }
    m.bonds.resize(a.bonds.size());
    for (auto &b : m.bonds) b.alive = true;
    return box;
}

// A ground plane under the box and a sphere striking its top face.
StepSettings<double> impactSettings(double h, double bottom) {
    StepSettings<double> s{};
    s.dt = 1.0 / 2400.0;
    s.gravity = {0.0, -9.81, 0.0};
    s.constraint_iterations = 1;
    s.bond_integrator = kBondXpbd;
    s.sphere_enabled = 1;
    s.contact = {0.5, 0.4, 0.1, 0.05, 0.5 * h, 1e-5, 0.05 * 0.04};
    s.node_contact.mode = kNodeContactOff;
    s.support.plane_count = 1;
    SupportPlane<double> &p = s.support.planes[0];
    p.point = {0.0, bottom, 0.0};
    p.normal = {0.0, 1.0, 0.0};
    p.tangent = {1.0, 0.0, 0.0};
    p.bitangent = {0.0, 0.0, 1.0};
    p.restitution = 0.1;
    p.static_friction = 0.6;
    p.dynamic_friction = 0.5;
    p.node_radius = 0.5 * h;
    p.reach_capped = 0;
    p.footprint_count = 0;
    return s;
}

// Every value type through the C views and back, with a distinct value in
// every field, so a dropped or swapped field shows.
void checkValueTypes() {
    StepSettings<double> s{};
    double x = 1.0;
    auto next = [&] { return x += 0.125; };
    s.dt = next();
    s.gravity = {next(), next(), next()};
    s.constraint_iterations = 3;
    s.bond_integrator = kBondXpbd;
    s.damping_fraction = next();
    s.sphere_enabled = 1;
    s.direct_arithmetic = 1;
    s.audit_energy = 0;
    s.plastic_yield_stretch = next();
    s.plastic_hardening = next();
    s.plate_half_thickness = next();
    s.contact = {next(), next(), next(), next(), next(), next(), next()};
    s.node_contact = {kNodeContactMeasure, next(), next(), next(), next(), next(), next(), next(), 1023U};
    s.support.plane_count = 2;
    for (auto &p : s.support.planes) {
        p.point = {next(), next(), next()};
        p.normal = {next(), next(), next()};
        p.tangent = {next(), next(), next()};
        p.bitangent = {next(), next(), next()};
        p.restitution = next();
        p.static_friction = next();
        p.dynamic_friction = next();
        p.node_radius = next();
        p.reach_capped = 1;
        p.footprint_count = 3;
        for (auto &f : p.footprints) f = {next(), next(), next(), next()};
    }
    const blx_settings o = hip_adapter::settingsView(s);
    bool ok = o.dt == s.dt && o.gravity[0] == s.gravity.x && o.gravity[1] == s.gravity.y && o.gravity[2] == s.gravity.z &&
              o.constraint_iterations == 3 && o.bond_integrator == s.bond_integrator &&
              o.damping_fraction == s.damping_fraction && o.sphere_enabled == 1 && o.direct_arithmetic == 1 &&
              o.audit_energy == 0 && o.plastic_yield_stretch == s.plastic_yield_stretch &&
              o.plastic_hardening == s.plastic_hardening && o.plate_half_thickness == s.plate_half_thickness &&
              o.contact.static_friction == s.contact.static_friction &&
              o.contact.dynamic_friction == s.contact.dynamic_friction && o.contact.restitution == s.contact.restitution &&
              o.contact.restitution_speed_threshold == s.contact.restitution_speed_threshold &&
              o.contact.node_contact_radius == s.contact.node_contact_radius &&
              o.contact.contact_margin == s.contact.contact_margin &&
              o.contact.prefilter_slack == s.contact.prefilter_slack && o.node_contact.mode == s.node_contact.mode &&
              o.node_contact.radius == s.node_contact.radius && o.node_contact.skin == s.node_contact.skin &&
              o.node_contact.margin == s.node_contact.margin &&
              o.node_contact.restitution == s.node_contact.restitution &&
              o.node_contact.restitution_speed_threshold == s.node_contact.restitution_speed_threshold &&
              o.node_contact.static_friction == s.node_contact.static_friction &&
              o.node_contact.dynamic_friction == s.node_contact.dynamic_friction &&
              o.node_contact.bucket_mask == 1023U && o.plane_count == 2;
    for (unsigned k = 0; k < 2; ++k) {
        const auto &p = s.support.planes[k];
        const auto &q = o.planes[k];
        ok = ok && q.point[0] == p.point.x && q.point[1] == p.point.y && q.point[2] == p.point.z &&
             q.normal[0] == p.normal.x && q.normal[1] == p.normal.y && q.normal[2] == p.normal.z &&
             q.tangent[0] == p.tangent.x && q.tangent[1] == p.tangent.y && q.tangent[2] == p.tangent.z &&
             q.bitangent[0] == p.bitangent.x && q.bitangent[1] == p.bitangent.y && q.bitangent[2] == p.bitangent.z &&
             q.restitution == p.restitution && q.static_friction == p.static_friction &&
             q.dynamic_friction == p.dynamic_friction && q.node_radius == p.node_radius && q.reach_capped == 1 &&
             q.footprint_count == 3;
        for (unsigned f = 0; f < 4; ++f)
            ok = ok && q.footprints[f].center_t == p.footprints[f].center_t &&
                 q.footprints[f].center_b == p.footprints[f].center_b &&
                 q.footprints[f].half_t == p.footprints[f].half_t && q.footprints[f].half_b == p.footprints[f].half_b;
    }
    expect(ok, "StepSettings<double> -> blx_settings field for field");

    SphereState<double> sp{{next(), next(), next()}, {next(), next(), next()}, {next(), next(), next()}, next(), next(), next()};
    SphereState<double> back{};
    hip_adapter::sphereBack(hip_adapter::sphereView(sp), back);
    expect(sameBits(sp, back), "SphereState<double> -> blx_sphere -> SphereState<double>");

    std::atomic<bool> flag{false};
    RunControl c{};
    c.max_steps = 101;
    c.quiet_steps = 102;
    c.min_steps = 103;
    c.no_failure_steps = 104;
    c.energy_flat_steps = 105;
    c.energy_flat_fraction = next();
    c.calm_steps = 106;
    c.calm_damage_margin = next();
    c.removable_energy_j = next();
    c.capture_stride = 107;
    c.max_frames = 108;
    c.steps_per_launch = 0;
    c.abandon = &flag;
    const blx_run_control rc = hip_adapter::controlView(c);
    expect(rc.max_steps == 101 && rc.quiet_steps == 102 && rc.min_steps == 103 && rc.no_failure_steps == 104 &&
               rc.energy_flat_steps == 105 && rc.energy_flat_fraction == c.energy_flat_fraction && rc.calm_steps == 106 &&
               rc.calm_damage_margin == c.calm_damage_margin && rc.removable_energy_j == c.removable_energy_j &&
               rc.capture_stride == 107 && rc.max_frames == 108 && rc.steps_per_launch == 0 && rc.abandon == nullptr,
           "RunControl -> blx_run_control field for field");
    expect(rc.abandon_poll && rc.abandon_context == &flag && rc.abandon_poll(rc.abandon_context) == 0,
           "RunControl::abandon polls the atomic (false)");
    flag.store(true);
    expect(rc.abandon_poll(rc.abandon_context) == 1, "RunControl::abandon polls the atomic (true)");
    RunControl none{};
    expect(hip_adapter::controlView(none).abandon_poll == nullptr, "no abandon, no poll");

    blx_run_status st{};
    std::uint64_t u = 1000;
    auto nu = [&] { return ++u; };
    st.total_steps = nu();
    st.first_failure_step = nu();
    st.last_failure_step = nu();
    st.failure_rounds = std::uint32_t(nu());
    st.broken_bonds = std::uint32_t(nu());
    st.removed_energy_j = next();
    st.last_energy_gain_step = nu();
    st.max_damage = next();
    st.last_damage_gain_step = nu();
    st.contact = {nu(), {next(), next(), next()}, next(), next(), next(), next(), nu(), nu()};
    st.external_load = {nu(), next(), {next(), next(), next()}, {next(), next(), next()}, {next(), next(), next()},
                        {next(), next(), next()}, next()};
    st.damping_dissipated_j = next();
    st.striker_dissipated_j = next();
    st.energy_audited = 1;
    st.exit_reason = 5;
    st.launches = std::uint32_t(nu());
    st.kernel_seconds = next();
    st.wall_seconds = next();
    st.frames_captured = std::uint32_t(nu());
    for (double &p : st.phase_seconds) p = next();
    st.max_tensile_stretch = float(next());
    st.max_compressive_strain = float(next());
    st.max_shear_strain = float(next());
    st.rank_deficient_nodes = std::uint32_t(nu());
    st.plastic_work_j = next();
    st.max_plastic_stretch = float(next());
    st.node_contact = {nu(), nu(), nu(), nu(), nu(), next(), next(), next(), {next(), next(), next()}};
    st.shared_memory_positions = 1;
    st.shared_memory_bytes = nu();
    RunStatus r;
    hip_adapter::statusBack(st, r);
    bool sok = r.total_steps == st.total_steps && r.first_failure_step == st.first_failure_step &&
               r.last_failure_step == st.last_failure_step && r.failure_rounds == st.failure_rounds &&
               r.broken_bonds == st.broken_bonds && r.removed_energy_j == st.removed_energy_j &&
               r.last_energy_gain_step == st.last_energy_gain_step && r.max_damage == st.max_damage &&
               r.last_damage_gain_step == st.last_damage_gain_step &&
               r.contact.impulse_contacts == st.contact.impulse_contacts &&
               r.contact.impulse_to_material_x == st.contact.impulse_to_material[0] &&
               r.contact.impulse_to_material_y == st.contact.impulse_to_material[1] &&
               r.contact.impulse_to_material_z == st.contact.impulse_to_material[2] &&
               r.contact.dissipated_kinetic_energy_j == st.contact.dissipated_kinetic_energy_j &&
               r.contact.maximum_penetration_m == st.contact.maximum_penetration_m &&
               r.contact.maximum_position_correction_m == st.contact.maximum_position_correction_m &&
               r.contact.maximum_center_shift_m == st.contact.maximum_center_shift_m &&
               r.contact.candidate_overflow == st.contact.candidate_overflow &&
               r.contact.ball_support_events == st.contact.ball_support_events &&
               r.external_load.steps == st.external_load.steps &&
               r.external_load.elapsed_s == st.external_load.elapsed_s &&
               r.external_load.requested_impulse_n_s.y == st.external_load.requested_impulse_n_s[1] &&
               r.external_load.impulse_n_s.z == st.external_load.impulse_n_s[2] &&
               r.external_load.requested_angular_impulse_kg_m2_s.x == st.external_load.requested_angular_impulse_kg_m2_s[0] &&
               r.external_load.angular_impulse_kg_m2_s.y == st.external_load.angular_impulse_kg_m2_s[1] &&
               r.external_load.work_j == st.external_load.work_j &&
               r.damping_dissipated_j == st.damping_dissipated_j && r.striker_dissipated_j == st.striker_dissipated_j &&
               r.energy_audited && r.exit_reason == 5 && r.launches == st.launches &&
               r.kernel_seconds == st.kernel_seconds && r.wall_seconds == st.wall_seconds &&
               r.frames_captured == st.frames_captured && r.max_tensile_stretch == st.max_tensile_stretch &&
               r.max_compressive_strain == st.max_compressive_strain && r.max_shear_strain == st.max_shear_strain &&
               r.rank_deficient_nodes == st.rank_deficient_nodes && r.plastic_work_j == st.plastic_work_j &&
               r.max_plastic_stretch == st.max_plastic_stretch && r.node_contact.contacts == st.node_contact.contacts &&
               r.node_contact.pair_tests == st.node_contact.pair_tests &&
               r.node_contact.rebuilds == st.node_contact.rebuilds &&
               r.node_contact.pairs_listed == st.node_contact.pairs_listed &&
               r.node_contact.pair_overflow == st.node_contact.pair_overflow &&
               r.node_contact.dissipated_kinetic_energy_j == st.node_contact.dissipated_kinetic_energy_j &&
               r.node_contact.maximum_overlap_m == st.node_contact.maximum_overlap_m &&
               r.node_contact.maximum_position_correction_m == st.node_contact.maximum_position_correction_m &&
               r.node_contact.momentum_residual_x == st.node_contact.momentum_residual[0] &&
               r.node_contact.momentum_residual_z == st.node_contact.momentum_residual[2] &&
               r.shared_memory_positions && r.shared_memory_bytes == st.shared_memory_bytes &&
               r.bond_integrator == kBondXpbd;
    for (unsigned p = 0; p < kPhaseCount; ++p) sok = sok && r.phase_seconds[p] == st.phase_seconds[p];
    expect(sok, "blx_run_status -> RunStatus field for field");
    std::printf("value types: settings, sphere, run control, run status\n");
}

// No device: the factory functions answer as the stub does.
void checkNoDevice(const LatticeSchedule &schedule) {
    expect(!cudaLatticeAvailable(), "cudaLatticeAvailable() is false without a device");
    expect(cudaLatticeMultiprocessorCount() == 0, "no multiprocessors without a device");
    expect(!cudaLatticeDescription().empty(), "a description without a device");
    bool threw = false;
    try {
        (void)makeCudaLatticeBackend(schedule, Precision::Float, 256);
    } catch (const std::runtime_error &) {
        threw = true;
    } catch (...) {
    }
    expect(threw, "makeCudaLatticeBackend throws std::runtime_error without a device");
    std::printf("no device: factory throws std::runtime_error, available false\n");
}

struct Result {
    LatticeState state;
    SphereState<double> sphere{};
    RunStatus status;
    std::vector<FrameCapture> frames;
    std::vector<std::uint32_t> first;
};

Result drive(LatticeBackend &be, const LatticeState &state0, const StepSettings<double> &s, const SphereState<double> &sp,
             bool loads) {
    Result r;
    be.upload(state0, s, sp);
    if (loads) {
        std::vector<Vec3> f(state0.node_count);
        for (std::uint32_t i = 0; i < state0.node_count; ++i)
            if (i % 7 == 3) f[i] = {0.01 * double(i % 5), -0.02, 0.005};
        be.setExternalForces(f, 40);
    }
    RunControl c{};
    c.capture_stride = 9;
    c.max_frames = 4;
    c.max_steps = 17;
    (void)be.run(c);
    c.max_steps = 33;
    r.status = be.run(c);
    r.state = state0;
    be.download(r.state, r.sphere);
    r.frames = be.takeFrames();
    r.first = be.firstFailureBonds();
    return r;
}

bool close(double x, double y, double rtol) { return std::abs(x - y) <= rtol * std::max(std::abs(x), std::abs(y)); }

// ledger_rtol: 0 for an exact external-load ledger; the device sums it in a
// fixed tree order rather than node order and is held to 1e-12 relative.
void compare(const Result &a, const Result &b, const std::string &tag, double ledger_rtol) {
    const LatticeState &x = a.state, &y = b.state;
    expect(sameVec(x.u, y.u) && sameVec(x.u_prev, y.u_prev) && sameVec(x.v, y.v), tag + ": u, u_prev, v");
    expect(sameVec(x.alive, y.alive) && sameVec(x.failure_mode, y.failure_mode) && sameVec(x.damage, y.damage),
           tag + ": alive, failure mode, damage");
    expect(sameVec(x.prev_tensile, y.prev_tensile) && sameVec(x.prev_compressive, y.prev_compressive) &&
               sameVec(x.prev_shear, y.prev_shear),
           tag + ": previous samples");
    expect(sameVec(x.plastic_extension, y.plastic_extension) && sameVec(x.plastic_strain, y.plastic_strain),
           tag + ": plastic state");
    expect(sameBits(a.sphere, b.sphere), tag + ": sphere");
    const RunStatus &p = a.status, &q = b.status;
    expect(p.total_steps == q.total_steps && p.first_failure_step == q.first_failure_step &&
               p.last_failure_step == q.last_failure_step && p.failure_rounds == q.failure_rounds &&
               p.broken_bonds == q.broken_bonds && p.removed_energy_j == q.removed_energy_j &&
               p.last_energy_gain_step == q.last_energy_gain_step && p.max_damage == q.max_damage &&
               p.last_damage_gain_step == q.last_damage_gain_step && p.exit_reason == q.exit_reason &&
               p.frames_captured == q.frames_captured && p.max_tensile_stretch == q.max_tensile_stretch &&
               p.max_compressive_strain == q.max_compressive_strain && p.max_shear_strain == q.max_shear_strain &&
               p.rank_deficient_nodes == q.rank_deficient_nodes,
           tag + ": run status counters");
    expect(sameBits(p.contact, q.contact), tag + ": striker contact ledger");
    const ExternalLoadLedger &la = p.external_load, &lb = q.external_load;
    expect(la.steps == lb.steps && close(la.work_j, lb.work_j, ledger_rtol) &&
               close(la.impulse_n_s.x, lb.impulse_n_s.x, ledger_rtol) &&
               close(la.impulse_n_s.y, lb.impulse_n_s.y, ledger_rtol) &&
               close(la.impulse_n_s.z, lb.impulse_n_s.z, ledger_rtol) &&
               close(la.requested_impulse_n_s.y, lb.requested_impulse_n_s.y, ledger_rtol),
           tag + ": external load ledger");
    bool frames = a.frames.size() == b.frames.size();
    for (std::size_t k = 0; frames && k < a.frames.size(); ++k)
        frames = a.frames[k].step == b.frames[k].step && sameVec(a.frames[k].u, b.frames[k].u) &&
                 sameVec(a.frames[k].alive, b.frames[k].alive) && sameVec(a.frames[k].damage, b.frames[k].damage) &&
                 sameBits(a.frames[k].sphere, b.frames[k].sphere);
    expect(frames, tag + ": frames");
    expect(a.first == b.first, tag + ": first failure bonds");
}

void checkDriven(const Box &box, const LatticeSchedule &schedule, const char *tag, bool loads, double ledger_rtol) {
    const double h = 0.01;
    double bottom = 1e9, top = -1e9;
    for (const Vec3 &p : box.matter.reference_positions_world_m) {
        bottom = std::min(bottom, p.y - 0.5 * h);
        top = std::max(top, p.y + 0.5 * h);
    }
    // The state holds positions relative to its origin, and the sphere and the
    // support plane are given in the same frame (fast_lattice_tests.cpp:402).
    const Vec3 origin{0.25, -0.5, 1.0};
    StepSettings<double> s = impactSettings(h, bottom - origin.y);
    if (loads) {
        s.sphere_enabled = 0;
        s.gravity = {0.0, 0.0, 0.0};
    }
    const SphereState<double> sp{{-origin.x, top + 0.04 + 0.002 - origin.y, -origin.z}, {0.0, -6.0, 0.0}, {0.0, 0.0, 0.0},
                                 0.04, 1.0, 0.00064};
    const LatticeState state0 = buildLatticeState(box.matter, schedule, origin);
    for (Precision prec : {Precision::Float, Precision::Double}) {
        auto gpu = makeCudaLatticeBackend(schedule, prec, 256);
        auto cpu = makeCpuLatticeBackend(schedule, prec);
        const Result a = drive(*gpu, state0, s, sp, loads);
        const Result b = drive(*cpu, state0, s, sp, loads);
        const std::string t = std::string(tag) + (prec == Precision::Float ? " float" : " double");
        expect(a.status.broken_bonds > 0 || loads, t + ": the run breaks bonds");
        compare(a, b, t, ledger_rtol);
        std::printf("%s: %u broken, %llu contacts, %zu frames, identical to makeCpuLatticeBackend\n", t.c_str(),
                    a.status.broken_bonds, (unsigned long long)a.status.contact.impulse_contacts, a.frames.size());
    }
}

} // namespace

int main(int argc, char **argv) {
    const std::string mode = argc > 1 ? argv[1] : "product";
    checkValueTypes();
    auto box = makeBox(12, 6, 12, 0.01);
    const LatticeSchedule one = buildLatticeSchedule(box->asset);
    if (mode == "product" && !cudaLatticeAvailable()) {
        checkNoDevice(one);
    } else if (mode == "product" || mode == "oracle") {
        std::printf("backend: %s\n", cudaLatticeDescription().c_str());
        // The oracle sums the load ledger in node order, as the reference does;
        // the device in a fixed tree order.
        const double ledger_rtol = mode == "oracle" ? 0.0 : 1e-12;
        checkDriven(*box, one, "impact, one block", false, ledger_rtol);
        const std::uint32_t layer = 12 * 6;
        const std::vector<std::uint32_t> slabs{0, 4 * layer, 8 * layer, 12 * layer};
        const LatticeSchedule three = buildLatticeSchedule(box->asset, slabs);
        checkDriven(*box, three, "impact, three slabs", false, ledger_rtol);
        checkDriven(*box, one, "external loads", true, ledger_rtol);
    } else {
        std::printf("usage: adapter_check product|oracle\n");
        return 2;
    }
    if (g_failures) {
        std::printf("adapter_check: %d failure(s)\n", g_failures);
        return 1;
    }
    std::printf("adapter_check ok\n");
    return 0;
}
