"""The on-disk side of a lattice run: the reference's measurement report and
its playback recording, filled from this library's lattice phase.

* ``measurements_json`` writes the object ``measurementsJson`` writes
  (src/fastlattice/TileImpactScene.cpp:2195-2309, with ``refractureJson``
  :2124-2192): the same key tree, serialised the way nlohmann::json's default
  ``dump()`` serialises it -- keys in lexicographic order (the object is a
  std::map), no whitespace, doubles in their shortest round-trip form,
  integers as integers, non-finite numbers as ``null``.
* ``write_playback`` writes a ``banjo.playback.v1`` recording
  (``writePlayback``, :2311-2580) of the lattice frames: one box body per
  cell, the striker sphere, the ground plane, and per frame the cell poses,
  their fragment ids and (while they fit the 48 MiB budget) the bond lines.
* ``measure_lattice_run`` fills the lattice-phase fields of
  ``TileImpactMeasurements`` (TileImpactScene.hpp:485-590) from a run's
  status, downloaded state, first-failure set and labels, as
  TileImpactScene.cpp:785-895 and :1040-1075 compute them.
* ``record_run`` runs a scene on a library with frame capture and returns
  both.

The rigid phase, re-fracture and the unloaded-shape probe belong to the Jolt
world and the scene harness, which this library does not run; their fields
keep the reference's not-run values (zero, -1 for times, false).
"""
from __future__ import annotations

import json
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi

# latticePhaseName for the XPBD integrator (src/fastlattice/FastLattice.cpp:22-26).
PHASE_NAMES = ("capture", "start_sample_nodes", "start_sample_bonds", "kick_classify", "contact_pass_1",
               "sweep_interior", "sweep_boundary", "support_project", "velocity_update", "damping",
               "contact_pass_2", "end_sample_nodes", "end_sample_bonds", "sphere_exit",
               "node_contact_broad", "node_contact_narrow")
assert len(PHASE_NAMES) == abi.PHASE_COUNT

PLAYBACK_BUDGET_BYTES = 48 * 1024 * 1024   # TileImpactScene.cpp:2334
PLAYBACK_LIMIT_BYTES = 64 * 1024 * 1024    # :2574
CELL_RGBA = 0x9FD3FFFF
BALL_RGBA = 0x8A8F99FF
LEDGE_RGBA = 0x7A7A7AFF


@dataclass
class Measurements:
    """TileImpactMeasurements (TileImpactScene.hpp:485-590), lattice phase."""
    backend_name: str = ""
    cells: int = 0
    bonds: int = 0
    colors: int = 0
    blocks: int = 0
    boundary_bonds: int = 0
    cell_size_m: float = 0.0
    tile_mass_kg: float = 0.0
    ball_mass_kg: float = 0.0
    dt_s: float = 0.0
    substep_limit_s: float = 0.0
    fastest_period_s: float = 0.0
    lattice_steps: int = 0
    lattice_simulated_s: float = 0.0
    lattice_wall_s: float = 0.0
    lattice_kernel_s: float = 0.0
    launches: int = 0
    exit_reason: int = 0
    failure_rounds: int = 0
    broken_bonds: int = 0
    first_failure_s: float = -1.0
    last_failure_s: float = -1.0
    removed_energy_j: float = 0.0
    bond_updates_per_s: float = 0.0
    failure_law: str = ""
    critical_stretch: float = 0.0
    energy_scaled_stretch: float = 0.0
    strength_stretch: float = 0.0
    strength_bound_active: bool = False
    lattice_crack_energy_j_m2: float = 0.0
    tensile_failures: int = 0
    compressive_failures: int = 0
    shear_failures: int = 0
    first_failure_bonds: int = 0
    first_failure_centroid_m: tuple = (0.0, 0.0, 0.0)
    first_failure_depth_m: float = -1.0
    first_failure_radius_m: float = -1.0
    pieces_over_1pct: int = 0
    pieces_over_5pct: int = 0
    mass_fraction_under_1pct: float = 0.0
    phase_seconds: tuple = (0.0,) * abi.PHASE_COUNT
    shared_memory_positions: bool = False
    shared_memory_bytes: int = 0
    max_tensile_stretch: float = 0.0
    max_compressive_strain: float = 0.0
    max_shear_strain: float = 0.0
    max_damage: float = 0.0
    rank_deficient_nodes: int = 0
    contact: dict = field(default_factory=dict)
    node_contact: dict = field(default_factory=dict)
    damping_dissipated_j: float = 0.0
    striker_dissipated_j: float = 0.0
    energy_audited: bool = False
    plasticity_enabled: bool = False
    yield_stretch: float = 0.0
    plastic_hardening_ratio: float = 0.0
    yield_strength_pa: float = 0.0
    plastic_work_j: float = 0.0
    max_plastic_stretch: float = 0.0
    plastic_bonds: int = 0
    plastic_extension_total_m: float = 0.0
    elastic_energy_j: float = 0.0
    lattice_kinetic_j: float = 0.0
    initial_kinetic_j: float = 0.0
    strike_deflection_mean_m: float = 0.0
    strike_deflection_amplitude_m: float = 0.0
    plate_deflection_mean_m: float = 0.0
    dent_depth_m: float = 0.0
    dent_frames: int = 0
    relax_steps: int = 0
    relax_broken_bonds: int = 0
    relax_plastic_work_j: float = 0.0
    relax_elastic_energy_j: float = 0.0
    relax_kinetic_j: float = 0.0
    permanent_dent_m: float = 0.0
    permanent_max_dip_m: float = 0.0
    loaded_dent_m: float = 0.0
    components: int = 0
    rigid_fragments: int = 0
    debris_particles: int = 0
    largest_piece_mass_kg: float = 0.0
    largest_piece_cells: int = 0
    pieces_at_end: int = 0
    broken_bonds_total: int = 0
    second_strike_time_s: float = -1.0
    second_ball_mass_kg: float = 0.0
    second_ball_speed_m_s: float = 0.0
    second_ball_start_y_m: float = 0.0
    handoff_wall_s: float = 0.0
    rigid_simulated_s: float = 0.0
    rigid_wall_s: float = 0.0
    rigid_steps: int = 0
    came_to_rest: bool = False
    rest_time_s: float = -1.0
    ball_speed_at_end_m_s: float = 0.0
    ball_height_at_end_m: float = 0.0
    simulated_total_s: float = 0.0
    wall_total_s: float = 0.0
    realtime_ratio: float = 0.0
    rule_met: bool = False
    recording_wall_s: float = 0.0


def _contact(c) -> dict:
    imp = list(c.impulse_to_material)
    return dict(impulse_contacts=int(c.impulse_contacts), impulse_to_material_n_s=[float(x) for x in imp],
                dissipated_kinetic_energy_j=float(c.dissipated_kinetic_energy_j),
                maximum_penetration_m=float(c.maximum_penetration_m),
                maximum_position_correction_m=float(c.maximum_position_correction_m),
                maximum_center_shift_m=float(c.maximum_center_shift_m),
                ball_support_events=int(c.ball_support_events))


def _node_contact(n) -> dict:
    return dict(contacts=int(n.contacts), pair_tests=int(n.pair_tests), rebuilds=int(n.rebuilds),
                pairs_listed=int(n.pairs_listed), pair_overflow=int(n.pair_overflow),
                dissipated_kinetic_energy_j=float(n.dissipated_kinetic_energy_j),
                maximum_overlap_m=float(n.maximum_overlap_m),
                maximum_position_correction_m=float(n.maximum_position_correction_m),
                momentum_residual_n_s=[float(x) for x in n.momentum_residual])


def _metrics_lib(lib):
    """The library that computes the state metrics: `lib` when it exports them
    (the product and the reference build), else the product's host code."""
    if lib.has("state_substep_limit"):
        return lib
    from .lattice import product
    return product()


def _by_asset(state, schedule, name):
    arr = getattr(state, name)
    out = np.empty_like(arr)
    out[schedule.bond_order] = arr
    return out


def measure_lattice_run(lib, scene, schedule, state, status, first_failure, labels, wall_s, backend_name,
                        origin=(0.0, 0.0, 0.0), tile_top_y=None, strike_xz=None, material=None):
    """The lattice-phase measurements of a finished run (TileImpactScene.cpp:785-895, 1040-1075).

    lib: the library whose host functions compute the state metrics
    (latticeStateSubstepLimit / ElasticEnergy / KineticEnergy). state: the
    downloaded state (schedule order). first_failure: asset indices of the
    first failure round's bonds. labels: canonical fragment ids per node.
    tile_top_y / strike_xz: the tile's top face and the strike axis the first
    failure's depth and radius are measured from (default: the top cell face
    and the sphere's initial centre). material: optional failure-law fields
    (failure_law, critical_stretch, energy_scaled_stretch, strength_stretch,
    strength_bound_active, lattice_crack_energy_j_m2, yield_strength_pa).
    """
    m = Measurements()
    asset = scene.asset
    s = scene.settings
    m.backend_name = backend_name
    m.cells = asset.node_count
    m.bonds = asset.bond_count
    m.colors = int(schedule.color_count)
    m.blocks = int(schedule.block_count)
    m.boundary_bonds = int(schedule.boundary_bond_count)
    m.cell_size_m = float(scene.cell)
    m.tile_mass_kg = float(np.sum(asset.mass))
    m.ball_mass_kg = float(scene.sphere.mass) if s.sphere_enabled else 0.0
    m.dt_s = float(s.dt)
    # The explicit limit 2 / omega_max and the period 2 pi / omega_max of the
    # as-built lattice (matter/Lattice.cpp:399-427); at upload every bond is
    # alive, so the state's limit is the asset's.
    # The period 2 pi / omega is taken as pi times the limit 2 / omega (the
    # state metrics give the limit), which agrees with it to rounding.
    m.substep_limit_s = float(getattr(scene, "initial_substep_limit", 0.0) or 0.0)
    m.fastest_period_s = m.substep_limit_s * math.acos(-1.0) if m.substep_limit_s > 0.0 else 0.0
    st = status
    m.lattice_steps = int(st.total_steps)
    m.lattice_simulated_s = float(st.total_steps) * m.dt_s
    m.lattice_wall_s = float(wall_s)
    m.lattice_kernel_s = float(st.kernel_seconds)
    m.launches = int(st.launches)
    m.exit_reason = int(st.exit_reason)
    m.failure_rounds = int(st.failure_rounds)
    m.broken_bonds = int(st.broken_bonds)
    m.removed_energy_j = float(st.removed_energy_j)
    m.contact = _contact(st.contact)
    m.node_contact = _node_contact(st.node_contact)
    m.damping_dissipated_j = float(st.damping_dissipated_j)
    m.striker_dissipated_j = float(st.striker_dissipated_j)
    m.energy_audited = bool(st.energy_audited)
    never = (1 << 64) - 1
    if st.last_failure_step != never:
        m.last_failure_s = float(st.last_failure_step) * m.dt_s
    if st.first_failure_step != never:
        m.first_failure_s = float(st.first_failure_step) * m.dt_s
    m.phase_seconds = tuple(float(x) for x in st.phase_seconds)
    m.max_tensile_stretch = float(st.max_tensile_stretch)
    m.max_damage = float(st.max_damage)
    m.max_compressive_strain = float(st.max_compressive_strain)
    m.max_shear_strain = float(st.max_shear_strain)
    m.rank_deficient_nodes = int(st.rank_deficient_nodes)
    m.plasticity_enabled = s.plastic_yield_stretch > 0.0
    m.yield_stretch = float(s.plastic_yield_stretch)
    m.plastic_hardening_ratio = float(s.plastic_hardening)
    m.plastic_work_j = float(st.plastic_work_j)
    m.max_plastic_stretch = float(st.max_plastic_stretch)
    pext = state.plastic_extension
    m.plastic_bonds = int(np.count_nonzero(pext != 0.0))
    m.plastic_extension_total_m = float(sum(abs(float(p)) for p in pext if p != 0.0))
    _, m.elastic_energy_j, m.lattice_kinetic_j = _metrics_lib(lib).state_metrics(state)
    sv = scene.sphere.velocity
    m.initial_kinetic_j = (0.5 * float(scene.sphere.mass) * (sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2])
                           if s.sphere_enabled else 0.0)
    m.bond_updates_per_s = float(m.bonds) * float(st.total_steps) / m.lattice_wall_s if m.lattice_wall_s > 0 else 0.0
    for k, v in (material or {}).items():
        setattr(m, k, v)
    # The first failure round: centroid of the rest midpoints (TileImpactScene.cpp:881-895).
    ref = np.asarray(asset.reference_position, np.float64).reshape(-1, 3)
    first = np.asarray(first_failure, np.int64)
    m.first_failure_bonds = int(first.size)
    if first.size:
        c = np.zeros(3)
        for k in first:
            c = c + 0.5 * (ref[asset.bond_a[k]] + ref[asset.bond_b[k]])
        c = c / float(first.size)
        top = float(ref[:, 1].max() + 0.5 * scene.cell) if tile_top_y is None else float(tile_top_y)
        sx, sz = (scene.sphere.center[0], scene.sphere.center[2]) if strike_xz is None else strike_xz
        m.first_failure_centroid_m = (float(c[0]), float(c[1]), float(c[2]))
        m.first_failure_depth_m = top - float(c[1])
        m.first_failure_radius_m = math.hypot(float(c[0]) - sx, float(c[2]) - sz)
    # Failure modes of the removed bonds, pieces (TileImpactScene.cpp:1055-1075).
    alive = _by_asset(state, schedule, "alive")
    mode = _by_asset(state, schedule, "failure_mode")
    dead = alive == 0
    m.tensile_failures = int(np.count_nonzero(dead & (mode == 1)))
    m.compressive_failures = int(np.count_nonzero(dead & (mode == 2)))
    m.shear_failures = int(np.count_nonzero(dead & (mode == 3)))
    labels = np.asarray(labels, np.int64)
    mass = np.asarray(asset.mass, np.float64)
    count = int(labels.max()) + 1 if labels.size else 0
    m.components = count
    small = 0.0
    for f in range(count):
        nodes = np.nonzero(labels == f)[0]
        piece = 0.0
        for i in nodes:  # the reference's serial sum in node order
            piece += float(mass[i])
        if f == 0:
            m.largest_piece_cells = int(nodes.size)
            m.largest_piece_mass_kg = piece
        if piece >= 0.01 * m.tile_mass_kg:
            m.pieces_over_1pct += 1
        else:
            small += piece
        if piece >= 0.05 * m.tile_mass_kg:
            m.pieces_over_5pct += 1
    m.mass_fraction_under_1pct = small / m.tile_mass_kg if m.tile_mass_kg > 0.0 else 0.0
    m.pieces_at_end = count
    m.broken_bonds_total = m.broken_bonds
    m.simulated_total_s = m.lattice_simulated_s
    m.wall_total_s = m.lattice_wall_s
    m.realtime_ratio = m.simulated_total_s / m.wall_total_s if m.wall_total_s > 0 else 0.0
    return m


def _refracture_tree():
    """refractureJson of a run with re-fracture off (TileImpactScene.cpp:2165-2191)."""
    return dict(enabled=False, contacts_tested=0, admitted=0,
                rejected=dict(no_live_bond=0, below_stress_bound=0, below_energy_bound=0, below_crack_energy=0,
                              max_closing_speed_m_s=0.0, max_closing_speed_any_m_s=0.0, max_margin=0.0),
                refused=dict(budget_events=0, budget_steps=0, unsupported_partner=0, too_many_cells=0,
                             support_penetration=0, striker_overlap=0, no_rollback=0),
                rollbacks=0, trial_steps=0, substeps=0, simulated_s=0.0, wall_s=0.0, broken_bonds=0,
                pieces_created=0, removed_energy_j=0.0, plastic_work_j=0.0,
                worst_residual=dict(entry_momentum_kg_m_s=0.0, entry_energy_j=0.0, exit_momentum_kg_m_s=0.0,
                                    exit_energy_j=0.0),
                events=[])


def measurements_tree(m: Measurements) -> dict:
    """The object measurementsJson builds (TileImpactScene.cpp:2196-2307)."""
    f = float
    lattice = dict(
        steps=int(m.lattice_steps), simulated_s=f(m.lattice_simulated_s), wall_s=f(m.lattice_wall_s),
        kernel_s=f(m.lattice_kernel_s), launches=int(m.launches), exit_reason=int(m.exit_reason),
        failure_rounds=int(m.failure_rounds), broken_bonds=int(m.broken_bonds),
        first_failure_s=f(m.first_failure_s), last_failure_s=f(m.last_failure_s),
        removed_energy_j=f(m.removed_energy_j), bond_updates_per_s=f(m.bond_updates_per_s),
        failure_law=str(m.failure_law), critical_stretch=f(m.critical_stretch),
        energy_scaled_stretch=f(m.energy_scaled_stretch), strength_stretch=f(m.strength_stretch),
        strength_bound_active=bool(m.strength_bound_active),
        lattice_crack_energy_j_m2=f(m.lattice_crack_energy_j_m2),
        tensile_failures=int(m.tensile_failures), compressive_failures=int(m.compressive_failures),
        shear_failures=int(m.shear_failures),
        first_failure=dict(bonds=int(m.first_failure_bonds), centroid_m=[f(x) for x in m.first_failure_centroid_m],
                           depth_below_top_m=f(m.first_failure_depth_m),
                           radius_from_strike_m=f(m.first_failure_radius_m)),
        contact=dict(m.contact) if m.contact else _contact(abi.ContactAccumulators()),
        node_contact=dict(m.node_contact) if m.node_contact else _node_contact(abi.NodeContactAccumulators()),
        dissipated_kinetic_energy_j=dict(
            audited=bool(m.energy_audited),
            node_contact=f((m.node_contact or {}).get("dissipated_kinetic_energy_j", 0.0)),
            striker_contact=f(m.striker_dissipated_j if m.energy_audited
                              else (m.contact or {}).get("dissipated_kinetic_energy_j", 0.0)),
            bond_damping=f(m.damping_dissipated_j), plastic_work=f(m.plastic_work_j),
            fracture=f(m.removed_energy_j)),
        plasticity=dict(enabled=bool(m.plasticity_enabled), yield_strength_pa=f(m.yield_strength_pa),
                        yield_stretch=f(m.yield_stretch), hardening_ratio=f(m.plastic_hardening_ratio),
                        plastic_work_j=f(m.plastic_work_j), max_plastic_stretch=f(m.max_plastic_stretch),
                        plastic_bonds=int(m.plastic_bonds),
                        plastic_extension_total_m=f(m.plastic_extension_total_m)),
        energy=dict(initial_kinetic_j=f(m.initial_kinetic_j), elastic_stored_j=f(m.elastic_energy_j),
                    lattice_kinetic_j=f(m.lattice_kinetic_j)),
        deformation=dict(relax_steps=int(m.relax_steps), relax_broken_bonds=int(m.relax_broken_bonds),
                         relax_plastic_work_j=f(m.relax_plastic_work_j),
                         relax_elastic_energy_j=f(m.relax_elastic_energy_j), relax_kinetic_j=f(m.relax_kinetic_j),
                         permanent_dent_m=f(m.permanent_dent_m), permanent_max_dip_m=f(m.permanent_max_dip_m),
                         loaded_dent_m=f(m.loaded_dent_m), frames=int(m.dent_frames),
                         strike_deflection_mean_m=f(m.strike_deflection_mean_m),
                         strike_deflection_amplitude_m=f(m.strike_deflection_amplitude_m),
                         plate_deflection_mean_m=f(m.plate_deflection_mean_m), dent_depth_m=f(m.dent_depth_m)))
    return dict(
        phase_seconds={PHASE_NAMES[p]: f(m.phase_seconds[p]) for p in range(abi.PHASE_COUNT)},
        shared_memory_positions=bool(m.shared_memory_positions), shared_memory_bytes=int(m.shared_memory_bytes),
        max_tensile_stretch=f(m.max_tensile_stretch), max_damage=f(m.max_damage),
        max_compressive_strain=f(m.max_compressive_strain), max_shear_strain=f(m.max_shear_strain),
        rank_deficient_nodes=int(m.rank_deficient_nodes), backend=str(m.backend_name),
        cells=int(m.cells), bonds=int(m.bonds), colors=int(m.colors), blocks=int(m.blocks),
        boundary_bonds=int(m.boundary_bonds), cell_size_m=f(m.cell_size_m), tile_mass_kg=f(m.tile_mass_kg),
        ball_mass_kg=f(m.ball_mass_kg), dt_s=f(m.dt_s), substep_limit_s=f(m.substep_limit_s),
        fastest_period_s=f(m.fastest_period_s),
        lattice=lattice,
        handoff=dict(components=int(m.components), rigid_fragments=int(m.rigid_fragments),
                     debris_particles=int(m.debris_particles), largest_piece_mass_kg=f(m.largest_piece_mass_kg),
                     largest_piece_cells=int(m.largest_piece_cells), wall_s=f(m.handoff_wall_s),
                     pieces_over_1pct=int(m.pieces_over_1pct), pieces_over_5pct=int(m.pieces_over_5pct),
                     mass_fraction_under_1pct=f(m.mass_fraction_under_1pct)),
        rigid=dict(simulated_s=f(m.rigid_simulated_s), wall_s=f(m.rigid_wall_s), steps=int(m.rigid_steps),
                   came_to_rest=bool(m.came_to_rest), rest_time_s=f(m.rest_time_s),
                   ball_speed_at_end_m_s=f(m.ball_speed_at_end_m_s),
                   ball_height_at_end_m=f(m.ball_height_at_end_m), pieces_at_end=int(m.pieces_at_end),
                   broken_bonds_total=int(m.broken_bonds_total),
                   second_strike_time_s=f(m.second_strike_time_s), second_ball_mass_kg=f(m.second_ball_mass_kg),
                   second_ball_speed_m_s=f(m.second_ball_speed_m_s),
                   second_ball_start_y_m=f(m.second_ball_start_y_m)),
        refracture=_refracture_tree(),
        simulated_total_s=f(m.simulated_total_s), wall_total_s=f(m.wall_total_s),
        realtime_ratio=f(m.realtime_ratio), rule_met=bool(m.rule_met), recording_wall_s=f(m.recording_wall_s))


def _finite(x):
    """nlohmann::json writes a non-finite number as null."""
    if isinstance(x, float):
        return x if math.isfinite(x) else None
    if isinstance(x, dict):
        return {k: _finite(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_finite(v) for v in x]
    return x


def dumps(obj) -> str:
    """nlohmann::json::dump() of an object built from std::map nodes: sorted
    keys, no whitespace, shortest round-trip doubles, null for non-finite."""
    return json.dumps(_finite(obj), sort_keys=True, separators=(",", ":"), ensure_ascii=False, allow_nan=False)


def measurements_json(m: Measurements) -> str:
    return dumps(measurements_tree(m))


def _vec(v):
    return [float(v[0]), float(v[1]), float(v[2])]


def write_playback(path, measurements: Measurements, frames, bond_nodes, cell_size, ball_radius, ground_y,
                   tile_material="tile", ball_material="ball", ground_material="ground", ledges=(),
                   report_json=None, striker=True):
    """A banjo.playback.v1 recording of lattice frames (TileImpactScene.cpp:2311-2580).

    frames: dicts with time_s, positions (N, 3) world, ball_center (3), alive
    and damage (asset bond order), component_ids (N), fracture_count.
    bond_nodes: (B, 2) asset bond ends. ledges: (center (3), dimensions (3)) boxes.
    """
    t0 = time.perf_counter()
    N = len(frames[0]["positions"]) if frames else 0
    B = len(bond_nodes)
    ident = [1.0, 0.0, 0.0, 0.0]

    def poses_for(fr):
        pos = fr["positions"]
        comp = fr.get("component_ids")
        poses = [{"id": f"cell:{i}", "position_m": _vec(pos[i]), "orientation_wxyz": ident,
                  "component_id": int(comp[i]) if comp is not None and len(comp) else 0} for i in range(N)]
        if striker:
            poses.append({"id": "ball", "position_m": _vec(fr["ball_center"]), "orientation_wxyz": ident,
                          "component_id": 0})
        for k, (center, _) in enumerate(ledges):
            poses.append({"id": f"ledge:{k}", "position_m": _vec(center), "orientation_wxyz": [1, 0, 0, 0],
                          "component_id": 0})
        return poses

    def bonds_for(fr):
        if len(fr["alive"]) != B:
            return []
        pos = fr["positions"]
        return [{"a_m": _vec(pos[a]), "b_m": _vec(pos[b]), "live": bool(fr["alive"][o] != 0),
                 "damage": float(np.float32(fr["damage"][o]))} for o, (a, b) in enumerate(bond_nodes)]

    # The frame budget, priced on frames sampled across the run (:2390-2430).
    pose_bytes, bond_bytes = 1, 0
    if frames:
        last = len(frames) - 1
        for f in sorted({0, last // 3, (2 * last) // 3, last}):
            pose_bytes = max(pose_bytes, len(dumps(poses_for(frames[f]))))
            if len(frames[f]["alive"]) == B:
                bond_bytes = max(bond_bytes, len(dumps(bonds_for(frames[f]))))
    frame_bytes = pose_bytes + 96
    fixed_bytes = 160 * (N + 1 + len(ledges)) + 64 * 1024
    spare = PLAYBACK_BUDGET_BYTES - fixed_bytes if PLAYBACK_BUDGET_BYTES > fixed_bytes else frame_bytes * 8
    budget_frames = max(8, spare // frame_bytes)
    total = len(frames)
    stride = 1 if total <= budget_frames else (total + budget_frames - 1) // budget_frames
    last_lattice = total - 1 if total else total
    kept = [f for f in range(total) if f % stride == 0 or f + 1 == total or f == last_lattice]
    record_bonds = bond_bytes * len(kept) + len(kept) * frame_bytes + fixed_bytes <= PLAYBACK_BUDGET_BYTES

    h = float(cell_size)
    bodies = [{"id": f"cell:{i}", "object_id": 2, "element_id": i, "material_id": tile_material,
               "color_rgba": CELL_RGBA, "shape": "box", "dimensions_m": [h, h, h]} for i in range(N)]
    if striker:
        d = 2.0 * float(ball_radius)
        bodies.append({"id": "ball", "object_id": 1, "element_id": 0, "material_id": ball_material,
                       "color_rgba": BALL_RGBA, "shape": "sphere", "dimensions_m": [d, d, d]})
    for k, (_, dims) in enumerate(ledges):
        bodies.append({"id": f"ledge:{k}", "object_id": 900 + k, "element_id": 0, "material_id": ground_material,
                       "color_rgba": LEDGE_RGBA, "shape": "box", "dimensions_m": _vec(dims)})
    g, gy = 1.0, float(ground_y)
    supports = [[[-g, gy, -g], [g, gy, -g], [g, gy, g]], [[-g, gy, -g], [g, gy, g], [-g, gy, g]]]
    out_frames = []
    for f in kept:
        fr = frames[f]
        out_frames.append({"time_s": float(fr["time_s"]), "phase": "lattice", "poses": poses_for(fr),
                           "bonds": bonds_for(fr) if record_bonds else [],
                           "fracture_count": int(fr["fracture_count"])})
    steps = int(measurements.lattice_steps) + int(measurements.rigid_steps)
    measurements.recording_wall_s = time.perf_counter() - t0
    artifact = {"schema": "banjo.playback.v1", "mode": "network", "units": "SI", "bodies": bodies,
                "supports": supports, "contacts": [], "frames": out_frames, "requested_steps": steps,
                "completed_steps": steps,
                "sampling": {"stride_steps": 0, "maximum_frames": len(kept), "interpolation": "none",
                             "captured_frames": total, "bond_lines": bool(record_bonds)},
                "physical_response_validated": False, "status": "complete", "error": "",
                "report": json.loads(report_json) if report_json is not None else measurements_tree(measurements)}
    text = dumps(artifact)
    if len(text.encode()) > PLAYBACK_LIMIT_BYTES:
        raise RuntimeError("playback exceeds 64 MiB")
    with open(path, "w", encoding="utf-8") as fh:
        fh.write(text)
    return artifact


def lattice_frames(lib, scene, schedule, state, raw_frames, origin=(0.0, 0.0, 0.0)):
    """Recorded frames from captures (TileImpactScene.cpp:898-929): world
    positions origin + (x0 + u), the ball centre, bond state in asset order
    and the fragment ids of the captured aliveness."""
    x0 = np.asarray(state.x0, np.float64).reshape(-1, 3)
    org = np.asarray(origin, np.float64)
    order = schedule.bond_order
    out = []
    for cap in raw_frames:
        u = np.asarray(cap["u"], np.float64).reshape(-1, 3)
        alive = np.empty(len(order), np.uint8)
        alive[order] = cap["alive"]
        damage = np.empty(len(order), np.float32)
        damage[order] = cap["damage"]
        labels, _ = lib.components(scene.node_count, scene.asset.bond_a, scene.asset.bond_b, alive)
        sp = cap["sphere"]
        out.append(dict(time_s=float(cap["step"]) * float(scene.settings.dt), positions=org + (x0 + u),
                        ball_center=org + np.array(sp.center[:3]), alive=alive, damage=damage,
                        component_ids=labels, fracture_count=int(np.count_nonzero(alive == 0))))
    return out


def record_run(lib, scene, precision=abi.PRECISION_FLOAT, capture_stride=0, max_frames=0, playback_path=None,
               backend_name=None, **measure_kw):
    """Runs `scene` on `lib` (HIP library or a checker library) with frame
    capture and returns (Measurements, frames, final), final holding the end
    state, its bond aliveness in asset order and its fragment labels; writes
    the playback when a path is given."""
    from .lattice import control
    from .runner import prepare

    sched, state = prepare(lib, scene)
    limit, _, _ = _metrics_lib(lib).state_metrics(state)
    scene.initial_substep_limit = limit
    be = lib.backend(precision, sched, *((0,) if lib.prefix == "blx_" else ()))
    try:
        be.upload(state, scene.settings, scene.sphere)
        if scene.info.get("blade") is not None:
            be.set_blade(scene.info["blade"])
        if scene.forces is not None:
            be.set_external_forces(scene.forces, scene.force_substeps)
        t0 = time.perf_counter()
        st = be.run(control(scene.substeps, capture_stride=capture_stride, max_frames=max_frames))
        wall = time.perf_counter() - t0
        out = state.copy()
        be.download(out)
        first = np.sort(sched.bond_order[be.first_failure_bonds()])
        raw = be.frames()
        alive = np.empty(scene.bond_count, np.uint8)
        alive[sched.bond_order] = out.alive
        labels, _ = lib.components(scene.node_count, scene.asset.bond_a, scene.asset.bond_b, alive)
    finally:
        be.close()
    m = measure_lattice_run(lib, scene, sched, out, st, first, labels, wall, backend_name or lib.name,
                            **measure_kw)
    frames = lattice_frames(lib, scene, sched, state, raw)
    if playback_path is not None:
        ground = scene.settings.planes[0].point[1] if scene.settings.plane_count else float(
            np.asarray(scene.asset.reference_position)[:, 1].min() - 0.5 * scene.cell)
        bond_nodes = np.stack([scene.asset.bond_a, scene.asset.bond_b], 1)
        write_playback(playback_path, m, frames, bond_nodes, scene.cell, scene.sphere.radius, ground,
                       striker=bool(scene.settings.sphere_enabled))
    return m, frames, dict(state=out, alive=alive, labels=labels, schedule=sched)
