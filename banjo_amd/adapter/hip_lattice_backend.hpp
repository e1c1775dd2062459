// The reference-side adapter of the MI355X lattice backend: the value-type
// conversions between the reference's lattice API (src/fastlattice/FastLattice.hpp,
// src/fastlattice/LatticePhysics.hpp) and the C views of include/banjo_lattice.h.
//
// hip_lattice_backend.cpp implements LatticeBackend (FastLattice.hpp:255-369)
// on these and defines the four GPU-lane functions the reference declares at
// FastLattice.hpp:401-406 (today defined by the stub
// src/fastlattice/CudaLatticeUnavailable.cpp:11-25, which this file replaces).
// It is compiled in the reference's tree, against its headers; this header is
// its only interface besides those four functions.
#pragma once

#include "banjo_lattice.h"
#include "fastlattice/FastLattice.hpp"
#include "fastlattice/LatticeSchedule.hpp"

namespace banjo::fastlattice::hip_adapter {

// LatticeSchedule -> blx_schedule_view (pointers into the schedule's arrays; read only).
blx_schedule_view scheduleView(const LatticeSchedule &schedule);
// LatticeState <-> blx_state_view: the same arrays, by pointer (upload reads
// them, download writes them).
blx_state_view stateView(LatticeState &state);
// StepSettings<double> -> blx_settings, field for field (LatticePhysics.hpp:249-264).
blx_settings settingsView(const StepSettings<double> &settings);
// SphereState<double> <-> blx_sphere (LatticePhysics.hpp:145-148).
blx_sphere sphereView(const SphereState<double> &sphere);
void sphereBack(const blx_sphere &in, SphereState<double> &out);
// RunControl -> blx_run_control (FastLattice.hpp:79-128). The abandon flag
// travels as a poll function over the std::atomic<bool>, loaded between device
// batches.
blx_run_control controlView(const RunControl &control);
// blx_run_status -> RunStatus, field for field (FastLattice.hpp:182-245). The
// fields this backend never fills (the Verlet lane's ledgers, the single-block
// shared-memory mode) keep their defaults.
void statusBack(const blx_run_status &in, RunStatus &out);

} // namespace banjo::fastlattice::hip_adapter
