"""Small hand-built lattices whose answers the mathematics fixes.

Used to pin the oracle (and, on the GPU, the HIP backend) where no reference
fixture exists: a single XPBD spring against backward Euler, the plastic
return's closed form, an affine stretch against the Green-Lagrange strain, a
rigid rotation against zero strain.
"""
import math

import numpy as np

from banjo_amd import abi, scenes
from banjo_amd.lattice import Asset


def two_node(length=0.01, compliance=1e-6, mass=1e-3, thresholds=(math.inf,) * 6):
    """Two nodes on the x axis joined by one bond."""
    pos = np.array([[0.0, 0.0, 0.0], [length, 0.0, 0.0]])
    return Asset(reference_position=pos, mass=np.full(2, mass), bond_a=np.array([0], np.uint32),
                 bond_b=np.array([1], np.uint32), rest_length=np.array([length]),
                 compliance=np.array([compliance]), threshold=np.array([thresholds], dtype=np.float64))


def scene_of(asset, dt, substeps=1, clamped=None, velocity=None, position=None, direct=False, yield_stretch=0.0,
             hardening=0.0, gravity=(0.0, 0.0, 0.0), iterations=1):
    if velocity is not None:
        asset.velocity = velocity
    if position is not None:
        asset.position = position
    st = scenes.settings(dt, gravity=gravity, iterations=iterations, direct=direct, yield_stretch=yield_stretch,
                         hardening=hardening)
    return scenes.Scene(name="case", dims=(0, 0, 0), cell=0.01, asset=asset,
                        slabs=np.array([0, asset.node_count], np.uint32), settings=st,
                        sphere=scenes.sphere_state(center=(0.0, 100.0, 0.0)), substeps=substeps, clamped=clamped)


def affine_box(n=5, cell=0.01, F=None, thresholds=None):
    """A Moore box deformed by x -> F x (about the box centre), at rest."""
    pos, a, b, L = scenes.box26(n, n, n, cell)
    F = np.eye(3) if F is None else np.asarray(F)
    B = a.shape[0]
    thr = np.full((B, 6), math.inf) if thresholds is None else thresholds
    asset = Asset(reference_position=pos, mass=np.full(pos.shape[0], 1e-3), bond_a=a, bond_b=b, rest_length=L,
                  compliance=np.full(B, 1e-6), threshold=thr)
    asset.position = pos @ F.T
    return asset


def green_lagrange(F):
    F = np.asarray(F)
    return 0.5 * (F.T @ F - np.eye(3))
