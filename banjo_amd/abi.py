"""ctypes mirror of include/banjo_lattice.h (the C ABI of the lattice step).

Field order and types follow the header exactly; ctypes applies the same
natural-alignment rules as the C compiler, so the layouts agree. Each struct
cites the reference type it mirrors (paths in the reference tree).
"""
import ctypes as C

import numpy as np

BLX_OK = 0
BLX_ERR_RUNTIME = -1
BLX_ERR_INVALID = -2
BLX_ERR_OVERFLOW = -3
BLX_ERR_LOGIC = -4

PRECISION_FLOAT = 0
PRECISION_DOUBLE = 1
MAX_SUPPORT_PLANES = 2
MAX_FOOTPRINTS = 4
MAX_CANDIDATES_PER_BLOCK = 1024
NO_BOND = 0xFFFFFFFF
PHASE_COUNT = 16
NO_STEP = 0xFFFFFFFFFFFFFFFF

u8, u32, u64, f32, f64 = C.c_uint8, C.c_uint32, C.c_uint64, C.c_float, C.c_double
P = C.POINTER


class Footprint(C.Structure):
    """Footprint<double>, src/fastlattice/LatticePhysics.hpp:114-120."""
    _fields_ = [("center_t", f64), ("center_b", f64), ("half_t", f64), ("half_b", f64)]


class SupportPlane(C.Structure):
    """SupportPlane<double>, src/fastlattice/LatticePhysics.hpp:122-136."""
    _fields_ = [("point", f64 * 3), ("normal", f64 * 3), ("tangent", f64 * 3), ("bitangent", f64 * 3),
                ("restitution", f64), ("static_friction", f64), ("dynamic_friction", f64),
                ("node_radius", f64), ("reach_capped", u8), ("footprint_count", u32),
                ("footprints", Footprint * MAX_FOOTPRINTS)]


class ContactSettings(C.Structure):
    """ContactSettings<double>, src/fastlattice/LatticePhysics.hpp:150-158."""
    _fields_ = [("static_friction", f64), ("dynamic_friction", f64), ("restitution", f64),
                ("restitution_speed_threshold", f64), ("node_contact_radius", f64),
                ("contact_margin", f64), ("prefilter_slack", f64)]


class NodeContactSettings(C.Structure):
    """NodeContactSettings<double>, src/fastlattice/LatticePhysics.hpp:188-202."""
    _fields_ = [("mode", u8), ("radius", f64), ("skin", f64), ("margin", f64), ("restitution", f64),
                ("restitution_speed_threshold", f64), ("static_friction", f64),
                ("dynamic_friction", f64), ("bucket_mask", u32)]


class Settings(C.Structure):
    """StepSettings<double>, src/fastlattice/LatticePhysics.hpp:248-284."""
    _fields_ = [("dt", f64), ("gravity", f64 * 3), ("constraint_iterations", u32),
                ("bond_integrator", u8), ("damping_fraction", f64), ("sphere_enabled", u8),
                ("direct_arithmetic", u8), ("audit_energy", u8), ("plastic_yield_stretch", f64),
                ("plastic_hardening", f64), ("plate_half_thickness", f64),
                ("contact", ContactSettings), ("node_contact", NodeContactSettings),
                ("plane_count", u32), ("planes", SupportPlane * MAX_SUPPORT_PLANES)]


class Sphere(C.Structure):
    """SphereState<double>, src/fastlattice/LatticePhysics.hpp:144-148."""
    _fields_ = [("center", f64 * 3), ("velocity", f64 * 3), ("angular_velocity", f64 * 3),
                ("radius", f64), ("mass", f64), ("inertia", f64)]


class Blade(C.Structure):
    """blx_blade: the kinematic blade extension (no reference counterpart)."""
    _fields_ = [("enabled", u8), ("edge_point", f64 * 3), ("flat_normal", f64 * 3), ("direction", f64 * 3),
                ("speed", f64), ("half_width", f64), ("node_radius", f64)]


class RunControl(C.Structure):
    """RunControl, src/fastlattice/FastLattice.hpp:79-128."""
    _fields_ = [("max_steps", u64), ("quiet_steps", u64), ("min_steps", u64),
                ("no_failure_steps", u64), ("energy_flat_steps", u64),
                ("energy_flat_fraction", f64), ("calm_steps", u64), ("calm_damage_margin", f64),
                ("removable_energy_j", f64), ("capture_stride", u64), ("max_frames", u32),
                ("steps_per_launch", u64), ("abandon", P(u8)),
                ("abandon_poll", C.CFUNCTYPE(C.c_int, C.c_void_p)), ("abandon_context", C.c_void_p)]


class ContactAccumulators(C.Structure):
    """ContactAccumulators, src/fastlattice/LatticePhysics.hpp:379-388."""
    _fields_ = [("impulse_contacts", u64), ("impulse_to_material", f64 * 3),
                ("dissipated_kinetic_energy_j", f64), ("maximum_penetration_m", f64),
                ("maximum_position_correction_m", f64), ("maximum_center_shift_m", f64),
                ("candidate_overflow", u64), ("ball_support_events", u64)]


class LoadLedger(C.Structure):
    """ExternalLoadLedger, src/fastlattice/FastLattice.hpp:140-147."""
    _fields_ = [("steps", u64), ("elapsed_s", f64), ("requested_impulse_n_s", f64 * 3),
                ("impulse_n_s", f64 * 3), ("requested_angular_impulse_kg_m2_s", f64 * 3),
                ("angular_impulse_kg_m2_s", f64 * 3), ("work_j", f64)]


class NodeContactAccumulators(C.Structure):
    """NodeContactAccumulators, src/fastlattice/LatticePhysics.hpp:204-222."""
    _fields_ = [("contacts", u64), ("pair_tests", u64), ("rebuilds", u64), ("pairs_listed", u64),
                ("pair_overflow", u64), ("dissipated_kinetic_energy_j", f64), ("maximum_overlap_m", f64),
                ("maximum_position_correction_m", f64), ("momentum_residual", f64 * 3)]


class FragmentMass(C.Structure):
    """FragmentMassProperties, src/fracture/FragmentMassProperties.hpp:11-21,
    with the fragment's node count and smallest node."""
    _fields_ = [("mass_kg", f64), ("center_of_mass_world_m", f64 * 3), ("linear_velocity_m_s", f64 * 3),
                ("inertia_world_kg_m2", f64 * 9), ("angular_momentum_kg_m2_s", f64 * 3),
                ("angular_velocity_rad_s", f64 * 3), ("source_kinetic_energy_j", f64),
                ("rigid_kinetic_energy_j", f64), ("coarsening_kinetic_loss_j", f64),
                ("node_count", u32), ("first_node", u32)]


# The same record as a numpy structured dtype (one row per fragment).
FRAGMENT_MASS_DTYPE = np.dtype([
    ("mass_kg", "<f8"), ("center_of_mass_world_m", "<f8", (3,)), ("linear_velocity_m_s", "<f8", (3,)),
    ("inertia_world_kg_m2", "<f8", (3, 3)), ("angular_momentum_kg_m2_s", "<f8", (3,)),
    ("angular_velocity_rad_s", "<f8", (3,)), ("source_kinetic_energy_j", "<f8"),
    ("rigid_kinetic_energy_j", "<f8"), ("coarsening_kinetic_loss_j", "<f8"),
    ("node_count", "<u4"), ("first_node", "<u4")])


class RunStatus(C.Structure):
    """RunStatus, src/fastlattice/FastLattice.hpp:182-245 (lattice-lane fields)."""
    _fields_ = [("total_steps", u64), ("first_failure_step", u64), ("last_failure_step", u64),
                ("failure_rounds", u32), ("broken_bonds", u32), ("removed_energy_j", f64),
                ("last_energy_gain_step", u64), ("max_damage", f64), ("last_damage_gain_step", u64),
                ("contact", ContactAccumulators), ("external_load", LoadLedger),
                ("damping_dissipated_j", f64), ("striker_dissipated_j", f64),
                ("energy_audited", u8), ("exit_reason", u32), ("launches", u32),
                ("kernel_seconds", f64), ("wall_seconds", f64), ("frames_captured", u32),
                ("phase_seconds", f64 * PHASE_COUNT), ("max_tensile_stretch", f32),
                ("max_compressive_strain", f32), ("max_shear_strain", f32),
                ("rank_deficient_nodes", u32), ("plastic_work_j", f64), ("max_plastic_stretch", f32),
                ("node_contact", NodeContactAccumulators), ("shared_memory_positions", u8),
                ("shared_memory_bytes", u64)]


class Tuning(C.Structure):
    """blx_tuning: launch tuning of a backend or batch (same bits for every setting)."""
    _fields_ = [("colours_per_launch", u32), ("lds_sweep", u32), ("lds_threads", u32), ("graph_substeps", u32),
                ("streams", u32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


TUNE_MAX = 32


class AutotuneReport(C.Structure):
    """blx_autotune_report: the candidates tried, their host seconds per substep, the one kept."""
    _fields_ = [("chosen", Tuning), ("candidates", u32), ("tried", Tuning * TUNE_MAX),
                ("seconds_per_substep", f64 * TUNE_MAX)]


class FragmentPose(C.Structure):
    """FragmentPose, src/fastlattice/Refracture.hpp:262-268."""
    _fields_ = [("center_of_mass", f64 * 3), ("orientation_wxyz", f64 * 4), ("linear_velocity", f64 * 3),
                ("angular_velocity", f64 * 3)]


class FragmentLatticeView(C.Structure):
    """FragmentLattice, src/fastlattice/Refracture.hpp:241-260, as caller-owned arrays."""
    _fields_ = [("node_count", u32), ("bond_count", u32), ("node_capacity", u32), ("bond_capacity", u32),
                ("origin", f64 * 3), ("carried_spin", f64 * 3), ("parent_node", P(u32)), ("parent_bond", P(u32)),
                ("reference_position", P(f64)), ("position", P(f64)), ("velocity", P(f64)), ("mass", P(f64)),
                ("bond_a", P(u32)), ("bond_b", P(u32)), ("rest_length", P(f64)), ("compliance", P(f64)),
                ("threshold", P(f64)), ("alive", P(u8)), ("damage", P(f64)), ("failure_mode", P(u8)),
                ("plastic_extension", P(f64)), ("plastic_strain", P(f64))]


class AssetView(C.Structure):
    """LatticeAsset bonds + ActiveMatter state, src/matter/Lattice.hpp:26-44,82-93."""
    _fields_ = [("node_count", u32), ("bond_count", u32), ("reference_position", P(f64)),
                ("position", P(f64)), ("velocity", P(f64)), ("mass", P(f64)),
                ("bond_a", P(u32)), ("bond_b", P(u32)), ("rest_length", P(f64)),
                ("compliance", P(f64)), ("threshold", P(f64)), ("alive", P(u8)),
                ("damage", P(f64)), ("failure_mode", P(u8))]


class ScheduleView(C.Structure):
    """LatticeSchedule, src/fastlattice/LatticeSchedule.hpp:28-46."""
    _fields_ = [("node_count", u32), ("bond_count", u32), ("block_count", u32),
                ("color_count", u32), ("boundary_bond_count", u32),
                ("node_block_begin", P(u32)), ("bond_order", P(u32)),
                ("bond_schedule_index", P(u32)), ("bond_block_begin", P(u32)),
                ("range_begin", P(u32)), ("range_end", P(u32)), ("bond_color", P(u8))]


class StateView(C.Structure):
    """LatticeState, src/fastlattice/FastLattice.hpp:37-55."""
    _fields_ = [("origin", f64 * 3), ("node_count", u32), ("bond_count", u32), ("max_degree", u32),
                ("x0", P(f64)), ("u", P(f64)), ("u_prev", P(f64)), ("v", P(f64)),
                ("inv_mass", P(f64)), ("mass", P(f64)), ("adj_offsets", P(u32)),
                ("adj_bonds", P(u32)), ("nbr_bond", P(u32)), ("nbr_other", P(u32)),
                ("nbr_rest", P(f64)), ("nbr_weight", P(f64)), ("nbr_alive", P(u8)),
                ("bond_slot_a", P(u32)), ("bond_slot_b", P(u32)), ("bond_a", P(u32)),
                ("bond_b", P(u32)), ("rest_edge", P(f64)), ("rest_length", P(f64)),
                ("rest_length_sq_minus", P(f64)), ("weight", P(f64)), ("compliance", P(f64)),
                ("threshold", P(f64)), ("alive", P(u8)), ("failure_mode", P(u8)),
                ("damage", P(f64)), ("prev_tensile", P(f64)), ("prev_compressive", P(f64)),
                ("prev_shear", P(f64)), ("plastic_extension", P(f64)), ("plastic_strain", P(f64))]


def status_dict(s):
    """A RunStatus as a plain dict (nested structs flattened one level)."""
    out = {}
    for name, _ in RunStatus._fields_:
        val = getattr(s, name)
        if isinstance(val, (ContactAccumulators, LoadLedger)):
            out[name] = {k: (list(getattr(val, k)) if hasattr(getattr(val, k), "__len__") else getattr(val, k))
                         for k, _ in val._fields_}
        elif hasattr(val, "__len__"):
            out[name] = list(val)
        else:
            out[name] = val
    return out
