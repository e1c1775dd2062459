"""The fragment build: a fractured lattice's pieces as rigid fragments and debris.

The host mirror of the reference's buildFragmentRepresentations
(src/fracture/FragmentGeometry.cpp:203-291) over this library's parts:

- the canonical fragment labels (blx_backend_fragment_labels: size descending,
  then smallest node, the order the reference's build sorts its components in);
- each fragment's mass properties (blx_backend_fragment_mass_properties on the
  device, or any table of the same fields in canonical order);
- its exposed cell surface and collision proxy (blx_exposed_voxel_surface,
  blx_collision_points), and its cells merged into boxes
  (blx_merge_cells_into_boxes, src/fracture/PieceGeometry.cpp:9-46).

A fragment is rigid while fewer than maximum_rigid_fragments have been made and
it has at least minimum_nodes_per_rigid_fragment cells (the first, largest one
always is); the rest become debris particles of the fragment's mass, at its
centre of mass, with the radius of a sphere of its volume at the material's
density (at least a fifth of a cell).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .lattice import LatticeLibrary, _ptr

# The C library's cube root, the function the reference's std::cbrt calls
# (numpy's cbrt rounds differently in the last place).
_LIBM = C.CDLL("libm.so.6")
_LIBM.cbrt.restype = C.c_double
_LIBM.cbrt.argtypes = [C.c_double]


@dataclass
class FragmentBuildSettings:
    """FragmentBuildSettings (src/fracture/FragmentGeometry.hpp:71-79)."""
    first_body_id: int = 1000
    maximum_rigid_fragments: int = 96
    minimum_nodes_per_rigid_fragment: int = 3
    maximum_collision_points: int = 192
    friction: float = 0.35
    restitution: float = 0.08
    rolling_resistance: float = 0.001

    def as_array(self):
        return np.array([self.first_body_id, self.maximum_rigid_fragments, self.minimum_nodes_per_rigid_fragment,
                         self.maximum_collision_points, self.friction, self.restitution, self.rolling_resistance],
                        np.float64)


@dataclass
class RigidFragment:
    body_id: int
    mass: np.void                 # the fragment's blx_fragment_mass record
    vertex_position: np.ndarray   # (4F, 3) relative to the centre of mass
    vertex_normal: np.ndarray     # (4F, 3)
    indices: np.ndarray           # (6F,)
    collision_points: np.ndarray  # (P, 3) relative to the centre of mass
    voxel_centers: np.ndarray     # (n, 3) relative to the centre of mass
    source_node_count: int
    voxel_size: float
    friction: float
    restitution: float
    rolling_resistance: float


@dataclass
class DebrisParticle:
    position: np.ndarray
    velocity: np.ndarray
    angular_velocity: np.ndarray
    mass: float
    radius: float
    inertia: np.ndarray           # (3, 3)


@dataclass
class FragmentBuild:
    rigid: list = field(default_factory=list)
    debris: list = field(default_factory=list)
    total_mass: float = 0.0
    rigid_mass: float = 0.0
    debris_mass: float = 0.0
    linear_momentum: tuple = (0.0, 0.0, 0.0)
    angular_momentum_about_origin: tuple = (0.0, 0.0, 0.0)
    coarsening_kinetic_loss: float = 0.0

    def flatten(self):
        """The build as the flat record the reference harness writes
        (oracle/ref_harness.cpp ref_fragment_build), for comparison."""
        o = [float(len(self.rigid)), float(len(self.debris)), self.total_mass, self.rigid_mass, self.debris_mass,
             *self.linear_momentum, *self.angular_momentum_about_origin, self.coarsening_kinetic_loss]
        for f in self.rigid:
            o += [float(f.body_id), float(f.source_node_count), float(f.mass["mass_kg"]),
                  *map(float, f.mass["center_of_mass_world_m"])]
            o += [len(f.indices) / 6.0, float(len(f.vertex_position)), float(len(f.indices)),
                  float(len(f.collision_points)), float(len(f.voxel_centers))]
            o += np.concatenate([f.vertex_position, f.vertex_normal], axis=1).reshape(-1).tolist()
            o += f.indices.astype(np.float64).tolist()
            o += f.collision_points.reshape(-1).tolist()
            o += f.voxel_centers.reshape(-1).tolist()
            o += [f.friction, f.restitution, f.rolling_resistance, f.voxel_size]
        for d in self.debris:
            o += [*d.position, *d.velocity, *d.angular_velocity, d.mass, d.radius, *d.inertia.reshape(-1)]
        return np.array(o, np.float64)


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def exposed_surface(lib: LatticeLibrary, grid, position, center_of_mass, voxel_size):
    """blx_exposed_voxel_surface of one fragment's cells (in the order given)."""
    g = np.ascontiguousarray(grid, np.int32).reshape(-1)
    p = np.ascontiguousarray(position, np.float64).reshape(-1)
    com = np.ascontiguousarray(center_of_mass, np.float64)
    n = len(g) // 3
    faces = C.c_uint64()
    fn = lib.fn("exposed_voxel_surface")
    lib.check(fn(_ptr(g, C.c_int32), _ptr(p, C.c_double), C.c_uint32(n), _ptr(com, C.c_double),
                 C.c_double(voxel_size), None, None, None, C.c_uint64(0), C.byref(faces)), "exposed_voxel_surface")
    F = faces.value
    vp = np.zeros((4 * F, 3))
    vn = np.zeros((4 * F, 3))
    idx = np.zeros(6 * F, np.uint32)
    lib.check(fn(_ptr(g, C.c_int32), _ptr(p, C.c_double), C.c_uint32(n), _ptr(com, C.c_double),
                 C.c_double(voxel_size), _ptr(vp, C.c_double), _ptr(vn, C.c_double), _ptr(idx, C.c_uint32),
                 C.c_uint64(F), C.byref(faces)), "exposed_voxel_surface")
    return vp, vn, idx


def collision_points(lib: LatticeLibrary, vertex_position, first_voxel_center, voxel_size, maximum_points):
    """blx_collision_points of a fragment surface."""
    vp = np.ascontiguousarray(vertex_position, np.float64).reshape(-1)
    c0 = np.ascontiguousarray(first_voxel_center, np.float64)
    cap = max(int(maximum_points), 8) + 8
    out = np.zeros((cap, 3))
    count = C.c_uint64()
    lib.check(lib.fn("collision_points")(_ptr(vp, C.c_double), C.c_uint64(len(vp) // 3), _ptr(c0, C.c_double),
                                         C.c_double(voxel_size), C.c_uint64(int(maximum_points)),
                                         _ptr(out, C.c_double), C.c_uint64(cap), C.byref(count)),
              "collision_points")
    return out[:count.value].copy()


def merge_cells_into_boxes(lib: LatticeLibrary, cells):
    """blx_merge_cells_into_boxes: (K, 6) inclusive boxes (min xyz, max xyz)."""
    c = np.ascontiguousarray(cells, np.int32).reshape(-1)
    count = C.c_uint64()
    fn = lib.fn("merge_cells_into_boxes")
    lib.check(fn(_ptr(c, C.c_int32), C.c_uint64(len(c) // 3), None, C.c_uint64(0), C.byref(count)),
              "merge_cells_into_boxes")
    out = np.zeros((count.value, 6), np.int32)
    lib.check(fn(_ptr(c, C.c_int32), C.c_uint64(len(c) // 3), _ptr(out, C.c_int32), C.c_uint64(count.value),
                 C.byref(count)), "merge_cells_into_boxes")
    return out


def build_fragments(lib: LatticeLibrary, labels, mass_properties, grid, position, voxel_size, density,
                    settings: FragmentBuildSettings | None = None) -> FragmentBuild:
    """buildFragmentRepresentations (FragmentGeometry.cpp:203-291).

    labels: canonical fragment id per node; mass_properties: one
    blx_fragment_mass record per fragment in canonical order; grid: (N, 3)
    cell coordinates; position: (N, 3) world positions; voxel_size: the
    asset's cell size; density: the material's, for the debris radius.
    """
    s = settings or FragmentBuildSettings()
    if s.first_body_id == 0 or s.maximum_rigid_fragments == 0 or s.maximum_collision_points < 8:
        raise ValueError("fragment build settings are invalid")
    labels = np.asarray(labels)
    position = np.asarray(position, np.float64).reshape(-1, 3)
    grid = np.asarray(grid, np.int32).reshape(-1, 3)
    if len(mass_properties) == 0:
        raise ValueError("fragment build requires active material components")
    order = np.argsort(labels, kind="stable")  # nodes by fragment, ascending within each
    bounds = np.searchsorted(labels[order], np.arange(len(mass_properties) + 1))
    out = FragmentBuild()
    p_tot = [0.0, 0.0, 0.0]
    l_tot = [0.0, 0.0, 0.0]
    body = s.first_body_id
    rigid_count = 0
    for f, mp in enumerate(mass_properties):
        nodes = order[bounds[f]:bounds[f + 1]]
        m = float(mp["mass_kg"])
        com = [float(x) for x in mp["center_of_mass_world_m"]]
        v = [float(x) for x in mp["linear_velocity_m_s"]]
        out.total_mass += m
        out.coarsening_kinetic_loss += float(mp["coarsening_kinetic_loss_j"])
        p = (v[0] * m, v[1] * m, v[2] * m)
        p_tot = [p_tot[k] + p[k] for k in range(3)]
        c = _cross(com, p)
        L = [float(x) for x in mp["angular_momentum_kg_m2_s"]]
        l_tot = [l_tot[k] + (L[k] + c[k]) for k in range(3)]
        if not (rigid_count < s.maximum_rigid_fragments
                and (len(nodes) >= s.minimum_nodes_per_rigid_fragment or rigid_count == 0)):
            radius = 0.0
            if m > 0.0 and density > 0.0:
                radius = _LIBM.cbrt(3.0 * (m / density) / (4.0 * np.pi))
            out.debris.append(DebrisParticle(position=np.array(com), velocity=np.array(v),
                                             angular_velocity=np.array(mp["angular_velocity_rad_s"], np.float64),
                                             mass=m, radius=max(0.2 * voxel_size, radius),
                                             inertia=np.array(mp["inertia_world_kg_m2"], np.float64).reshape(3, 3)))
            out.debris_mass += m
            continue
        centers = position[nodes] - np.array(com)
        vp, vn, idx = exposed_surface(lib, grid[nodes], position[nodes], com, voxel_size)
        pts = collision_points(lib, vp, centers[0], voxel_size, s.maximum_collision_points)
        out.rigid.append(RigidFragment(body_id=body, mass=mp, vertex_position=vp, vertex_normal=vn, indices=idx,
                                       collision_points=pts, voxel_centers=centers, source_node_count=len(nodes),
                                       voxel_size=voxel_size, friction=s.friction, restitution=s.restitution,
                                       rolling_resistance=s.rolling_resistance))
        out.rigid_mass += m
        body += 1
        rigid_count += 1
    out.linear_momentum = tuple(p_tot)
    out.angular_momentum_about_origin = tuple(l_tot)
    return out
