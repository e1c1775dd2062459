"""Degenerate and ragged lattices under the striker, on the device against the
oracle, bit for bit:

- the smallest block, 2x2x2 nodes (one node tile, one bond tile, every node a
  candidate's neighbour);
- a block whose bonds are all broken when it is uploaded: every node is free,
  no node has a strain sample, and every node is its own fragment;
- a block with a seeded half of its bonds broken at upload: ragged neighbour
  lists, rank-deficient and invalid nodes, bonds whose ends have no strain.

A state with bonds already broken is what a caller uploads to continue a run
(LatticeState::alive and nbr_alive, FastLattice.hpp:37-54, with nbr_alive the
alive flag of each slot's bond, FastLattice.cpp:139-168).
"""
import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.runner import prepare, run_scene
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu
F, D = abi.PRECISION_FLOAT, abi.PRECISION_DOUBLE
NO_BOND = 0xFFFFFFFF


def broken_at_upload(lib, sc, fraction, seed):
    """The scene's schedule and state with a seeded fraction of its bonds
    (chosen in asset order) broken, the slot flags agreeing."""
    sched, state = prepare(lib, sc)
    dead = np.random.default_rng(seed).random(sc.bond_count) < fraction
    state.alive[dead[sched.bond_order]] = 0
    slots = state.nbr_bond
    used = slots != NO_BOND
    state.nbr_alive[used] = state.alive[slots[used]]
    return sched, state


@pytest.mark.parametrize("precision", [F, D], ids=["float", "double"])
def test_smallest_block_under_the_striker(gpu, oracle, precision):
    sc = scenes.small_impact(nx=2, ny=2, nz=2, ball_radius=0.01, ball_mass=0.05, substeps=120)
    print(f"impact 2x2x2, N {sc.node_count} B {sc.bond_count}", flush=True)
    h = run_scene(gpu, sc, precision, labels="device")
    o = run_scene(oracle, sc, precision)
    assert h["status"].contact.impulse_contacts > 0
    assert_same(h, o)


@pytest.mark.parametrize("fraction", [1.0, 0.5], ids=["all", "half"])
def test_bonds_broken_at_upload(gpu, oracle, fraction):
    sc = scenes.small_impact(nx=10, ny=6, nz=10, substeps=90)
    print(f"impact 10x6x10 with {fraction:.0%} of its bonds broken at upload", flush=True)
    h = run_scene(gpu, sc, F, labels="device", prepared=broken_at_upload(gpu, sc, fraction, 5))
    o = run_scene(oracle, sc, F, prepared=broken_at_upload(oracle, sc, fraction, 5))
    assert_same(h, o)
    if fraction == 1.0:
        assert h["components"] == sc.node_count and h["status"].broken_bonds == 0
        assert not h["alive"].any()
    else:
        assert 1 < h["components"] < sc.node_count
