"""The on-disk formats: the measurement report (measurementsJson,
src/fastlattice/TileImpactScene.cpp:2195-2309) and the banjo.playback.v1
recording (writePlayback, :2311-2580), written by banjo_amd/recording.py.

The report's key tree is the reference's, node for node (the expected tree
below is read off TileImpactScene.cpp:2196-2307 and refractureJson
:2165-2191). It is serialised the way nlohmann::json::dump() serialises a
std::map-backed object. The fields derived from a run are checked against
their definitions (TileImpactScene.cpp:785-895, 1055-1075) on an oracle run.
"""
import json
import math

import numpy as np
import pytest

from banjo_amd import abi, recording, scenes

CONTACT = {"impulse_contacts", "impulse_to_material_n_s", "dissipated_kinetic_energy_j", "maximum_penetration_m",
           "maximum_position_correction_m", "maximum_center_shift_m", "ball_support_events"}
NODE_CONTACT = {"contacts", "pair_tests", "rebuilds", "pairs_listed", "pair_overflow", "dissipated_kinetic_energy_j",
                "maximum_overlap_m", "maximum_position_correction_m", "momentum_residual_n_s"}
EXPECTED = {
    "phase_seconds": set(recording.PHASE_NAMES),
    "shared_memory_positions": None, "shared_memory_bytes": None, "max_tensile_stretch": None, "max_damage": None,
    "max_compressive_strain": None, "max_shear_strain": None, "rank_deficient_nodes": None, "backend": None,
    "cells": None, "bonds": None, "colors": None, "blocks": None, "boundary_bonds": None, "cell_size_m": None,
    "tile_mass_kg": None, "ball_mass_kg": None, "dt_s": None, "substep_limit_s": None, "fastest_period_s": None,
    "lattice": {
        "steps": None, "simulated_s": None, "wall_s": None, "kernel_s": None, "launches": None, "exit_reason": None,
        "failure_rounds": None, "broken_bonds": None, "first_failure_s": None, "last_failure_s": None,
        "removed_energy_j": None, "bond_updates_per_s": None, "failure_law": None, "critical_stretch": None,
        "energy_scaled_stretch": None, "strength_stretch": None, "strength_bound_active": None,
        "lattice_crack_energy_j_m2": None, "tensile_failures": None, "compressive_failures": None,
        "shear_failures": None,
        "first_failure": {"bonds": None, "centroid_m": None, "depth_below_top_m": None, "radius_from_strike_m": None},
        "contact": CONTACT, "node_contact": NODE_CONTACT,
        "dissipated_kinetic_energy_j": {"audited", "node_contact", "striker_contact", "bond_damping",
                                        "plastic_work", "fracture"},
        "plasticity": {"enabled", "yield_strength_pa", "yield_stretch", "hardening_ratio", "plastic_work_j",
                       "max_plastic_stretch", "plastic_bonds", "plastic_extension_total_m"},
        "energy": {"initial_kinetic_j", "elastic_stored_j", "lattice_kinetic_j"},
        "deformation": {"relax_steps", "relax_broken_bonds", "relax_plastic_work_j", "relax_elastic_energy_j",
                        "relax_kinetic_j", "permanent_dent_m", "permanent_max_dip_m", "loaded_dent_m", "frames",
                        "strike_deflection_mean_m", "strike_deflection_amplitude_m", "plate_deflection_mean_m",
                        "dent_depth_m"},
    },
    "handoff": {"components", "rigid_fragments", "debris_particles", "largest_piece_mass_kg", "largest_piece_cells",
                "wall_s", "pieces_over_1pct", "pieces_over_5pct", "mass_fraction_under_1pct"},
    "rigid": {"simulated_s", "wall_s", "steps", "came_to_rest", "rest_time_s", "ball_speed_at_end_m_s",
              "ball_height_at_end_m", "pieces_at_end", "broken_bonds_total", "second_strike_time_s",
              "second_ball_mass_kg", "second_ball_speed_m_s", "second_ball_start_y_m"},
    "refracture": {"enabled": None, "contacts_tested": None, "admitted": None,
                   "rejected": {"no_live_bond", "below_stress_bound", "below_energy_bound", "below_crack_energy",
                                "max_closing_speed_m_s", "max_closing_speed_any_m_s", "max_margin"},
                   "refused": {"budget_events", "budget_steps", "unsupported_partner", "too_many_cells",
                               "support_penetration", "striker_overlap", "no_rollback"},
                   "rollbacks": None, "trial_steps": None, "substeps": None, "simulated_s": None, "wall_s": None,
                   "broken_bonds": None, "pieces_created": None, "removed_energy_j": None, "plastic_work_j": None,
                   "worst_residual": {"entry_momentum_kg_m_s", "entry_energy_j", "exit_momentum_kg_m_s",
                                      "exit_energy_j"},
                   "events": None},
    "simulated_total_s": None, "wall_total_s": None, "realtime_ratio": None, "rule_met": None,
    "recording_wall_s": None,
}


def check_tree(got, expected, where="report"):
    keys = set(expected) if isinstance(expected, (set, dict)) else set()
    assert set(got) == keys, f"{where}: {sorted(set(got) ^ keys)}"
    if isinstance(expected, dict):
        for k, sub in expected.items():
            if sub is not None:
                assert isinstance(got[k], dict), f"{where}.{k}"
                check_tree(got[k], sub, f"{where}.{k}")


def test_report_key_tree_is_the_references():
    check_tree(json.loads(recording.measurements_json(recording.Measurements())), EXPECTED)


def test_serialised_as_nlohmann_dumps():
    m = recording.Measurements(backend_name="x", cells=3, removed_energy_j=float("nan"), dt_s=1e-7,
                               tile_mass_kg=2.0, max_tensile_stretch=float(np.float32(0.1)))
    text = recording.measurements_json(m)
    assert " " not in text.replace('"failure_law":""', "")  # compact: no whitespace anywhere
    d = json.loads(text)
    assert list(d) == sorted(d) and list(d["lattice"]) == sorted(d["lattice"])
    assert d["lattice"]["removed_energy_j"] is None                # non-finite -> null
    assert '"dt_s":1e-07' in text and '"tile_mass_kg":2.0' in text  # shortest round trip, doubles keep ".0"
    assert '"cells":3,' in text and '"first_failure_s":-1.0' in text
    assert d["max_tensile_stretch"] == float(np.float32(0.1))      # a float widened to double, as the C++ does


@pytest.fixture(scope="module")
def recorded(oracle, tmp_path_factory):
    sc = scenes.small_impact(nx=12, ny=6, nz=12, substeps=60)
    path = tmp_path_factory.mktemp("rec") / "impact.json"
    m, frames, final = recording.record_run(oracle, sc, capture_stride=10, max_frames=20, playback_path=str(path))
    return sc, m, frames, json.loads(path.read_text()), final


def test_measured_fields_follow_their_definitions(oracle, hip, recorded):
    sc, m, frames, _, final = recorded
    from banjo_amd.runner import prepare
    sched, state = prepare(oracle, sc)
    limit, _, _ = hip.state_metrics(state)
    assert m.cells == sc.node_count and m.bonds == sc.bond_count and m.colors == sched.color_count
    omega = 2.0 / limit
    assert m.substep_limit_s == limit and m.fastest_period_s == pytest.approx(2.0 * math.acos(-1.0) / omega, rel=1e-15)
    assert m.lattice_simulated_s == m.lattice_steps * sc.settings.dt
    assert m.bond_updates_per_s == pytest.approx(m.bonds * m.lattice_steps / m.lattice_wall_s)
    assert m.broken_bonds > 0 and m.tensile_failures + m.compressive_failures + m.shear_failures == m.broken_bonds
    labels = final["labels"]
    masses = np.bincount(labels, weights=sc.asset.mass)
    assert m.components == len(masses) and m.largest_piece_cells == int(np.sum(labels == 0))
    assert m.pieces_over_1pct == int(np.sum(masses >= 0.01 * m.tile_mass_kg))
    assert m.pieces_over_5pct == int(np.sum(masses >= 0.05 * m.tile_mass_kg))
    assert m.mass_fraction_under_1pct == pytest.approx(masses[masses < 0.01 * m.tile_mass_kg].sum() / m.tile_mass_kg)
    assert m.initial_kinetic_j == pytest.approx(0.5 * sc.sphere.mass * np.dot(sc.sphere.velocity[:], sc.sphere.velocity[:]))
    assert m.first_failure_bonds > 0 and m.first_failure_depth_m > 0.0


def test_playback_recording(recorded):
    sc, m, frames, d, _ = recorded
    N, B = sc.node_count, sc.bond_count
    assert d["schema"] == "banjo.playback.v1" and d["units"] == "SI" and d["status"] == "complete"
    assert len(d["bodies"]) == N + 1 and d["bodies"][-1]["id"] == "ball" and d["bodies"][0]["shape"] == "box"
    assert len(d["supports"]) == 2 and d["contacts"] == []
    assert d["sampling"]["captured_frames"] == len(frames) == len(d["frames"]) and d["sampling"]["bond_lines"]
    assert d["requested_steps"] == d["completed_steps"] == m.lattice_steps
    check_tree(d["report"], EXPECTED)
    for fr, src in zip(d["frames"], frames):
        assert fr["phase"] == "lattice" and fr["time_s"] == src["time_s"]
        pos = np.array([p["position_m"] for p in fr["poses"][:N]])
        assert np.array_equal(pos, src["positions"])
        assert [p["component_id"] for p in fr["poses"][:N]] == list(map(int, src["component_ids"]))
        assert len(fr["bonds"]) == B
        assert [b["live"] for b in fr["bonds"]] == [bool(a) for a in src["alive"]]
        assert fr["fracture_count"] == int(np.sum(src["alive"] == 0))
    first = d["frames"][0]["bonds"][0]
    a, b = int(sc.asset.bond_a[0]), int(sc.asset.bond_b[0])
    assert first["a_m"] == list(frames[0]["positions"][a]) and first["b_m"] == list(frames[0]["positions"][b])


def test_playback_thins_frames_to_its_budget(recorded, tmp_path, monkeypatch):
    sc, m, frames, _, _ = recorded
    many = [dict(f) for f in frames for _ in range(8)]  # 48 frames
    monkeypatch.setattr(recording, "PLAYBACK_BUDGET_BYTES", 3 * 1024 * 1024)
    bond_nodes = np.stack([sc.asset.bond_a, sc.asset.bond_b], 1)
    d = recording.write_playback(tmp_path / "thin.json", m, many, bond_nodes, sc.cell, 0.05, 0.0)
    kept = d["sampling"]["maximum_frames"]
    assert d["sampling"]["captured_frames"] == 48 and kept < 48
    assert d["frames"][-1]["time_s"] == many[-1]["time_s"]  # the last frame is always kept
    assert not d["sampling"]["bond_lines"] and all(fr["bonds"] == [] for fr in d["frames"])
