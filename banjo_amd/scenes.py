"""Synthetic workloads for the lattice step, seeded and reproducible.

Lattices are simple-cubic boxes with all 26 Moore neighbours. The reference's
box generator (src/matter/BoxLattice.cpp:49-171) builds bonds from a Euclidean
horizon, which cannot express the 26-neighbour stencil (horizon 1 gives 6
neighbours, horizon 2 gives 32), so the stencil here keeps every one of its
conventions with the Moore offsets instead: node index x + nx*(y + ny*z),
reference positions centred on the lattice origin, bonds emitted node-major
over the 13 positive half offsets in the generator's dz, dy, dx loop order
(so node_a < node_b), rest length |offset| * cell computed in double.

Per-bond randomness is keyed by (seed, node_a, node_b) with the reference's
splitmix64 mix (src/matter/Lattice.cpp:28-47): value = seed ^ (a << 32) ^ b,
then splitmix64, then U = (h >> 11) * 2**-53. Each random field uses its own
stream seed.

Clamps follow the recipe expressible in the reference's XPBD lane: a clamped
node keeps its mass but gets inv_mass 0 and v 0, StepSettings gravity is zero,
and gravity is applied as external forces m*g on movable nodes only
(src/fastlattice/ExternalLoads.hpp:42-46 skips zero forces, so fixed nodes are
never loaded).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .lattice import Asset

# The 13 positive half offsets of the Moore stencil in the generator's
# dz-outer, dy, dx-inner loop order (src/matter/BoxLattice.cpp:141-147 with
# positiveHalfOffset, src/matter/Lattice.cpp:49-52).
HALF_OFFSETS = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                if dz > 0 or (dz == 0 and dy > 0) or (dz == 0 and dy == 0 and dx > 0)]

INF = float("inf")


def splitmix64(x):
    """splitmix64 on a uint64 array (src/matter/Lattice.cpp:28-33)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def bond_uniform(seed, a, b):
    """U in [0, 1) per bond, keyed like signedVariation (src/matter/Lattice.cpp:35-47)."""
    v = np.uint64(seed) ^ (np.asarray(a, np.uint64) << np.uint64(32)) ^ np.asarray(b, np.uint64)
    h = splitmix64(v)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def box26(nx, ny, nz, cell):
    """Reference positions (N,3) and bond ends/rest lengths of a Moore box."""
    x = np.arange(nx, dtype=np.float64)
    y = np.arange(ny, dtype=np.float64)
    z = np.arange(nz, dtype=np.float64)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    pos = np.stack([(X + 0.5 - 0.5 * nx) * cell, (Y + 0.5 - 0.5 * ny) * cell,
                    (Z + 0.5 - 0.5 * nz) * cell], axis=-1).reshape(-1, 3)
    ix = np.arange(nx, dtype=np.int64)
    iy = np.arange(ny, dtype=np.int64)
    iz = np.arange(nz, dtype=np.int64)
    GZ, GY, GX = np.meshgrid(iz, iy, ix, indexing="ij")
    GX, GY, GZ = GX.reshape(-1), GY.reshape(-1), GZ.reshape(-1)
    node = GX + nx * (GY + ny * GZ)
    others, valid, lengths = [], [], []
    for dx, dy, dz in HALF_OFFSETS:
        X2, Y2, Z2 = GX + dx, GY + dy, GZ + dz
        ok = (X2 >= 0) & (Y2 >= 0) & (Z2 >= 0) & (X2 < nx) & (Y2 < ny) & (Z2 < nz)
        others.append(X2 + nx * (Y2 + ny * Z2))
        valid.append(ok)
        lengths.append(math.sqrt(dx * dx + dy * dy + dz * dz) * cell)
    other = np.stack(others, axis=1)
    mask = np.stack(valid, axis=1)
    L = np.broadcast_to(np.array(lengths), other.shape)
    a = np.broadcast_to(node[:, None], other.shape)[mask].astype(np.uint32)
    b = other[mask].astype(np.uint32)
    return pos, a, b, L[mask].astype(np.float64)


def box_slabs(nx, ny, nz, horizon, requested):
    """boxLatticeSlabs (src/matter/BoxLattice.cpp:175-195): z-slab node offsets."""
    blocks = max(1, requested)
    blocks = min(blocks, max(1, nz // horizon))
    layer = nx * ny
    out = [((nz * b) // blocks) * layer for b in range(blocks + 1)]
    out[-1] = layer * nz
    return np.array(out, np.uint32)


def ground_plane(y, node_radius, restitution=0.1, static_friction=0.6, dynamic_friction=0.5):
    """An infinite support plane with normal +y at height y (TileImpactScene.cpp:83-96,317-321)."""
    p = abi.SupportPlane()
    p.point[:] = (0.0, y, 0.0)
    p.normal[:] = (0.0, 1.0, 0.0)
    p.tangent[:] = (1.0, 0.0, 0.0)
    p.bitangent[:] = (0.0, 0.0, 1.0)
    p.restitution, p.static_friction, p.dynamic_friction = restitution, static_friction, dynamic_friction
    p.node_radius = node_radius
    p.reach_capped = 0
    p.footprint_count = 1
    p.footprints[0].center_t = 0.0
    p.footprints[0].center_b = 0.0
    p.footprints[0].half_t = INF
    p.footprints[0].half_b = INF
    return p


def settings(dt, gravity=(0.0, 0.0, 0.0), iterations=1, planes=(), sphere=False, cell=0.01,
             ball_radius=0.0, direct=False, damping=0.0, yield_stretch=0.0, hardening=0.0):
    """StepSettings<double> as TileImpactScene buildSettings fills it (TileImpactScene.cpp:259-336)."""
    s = abi.Settings()
    s.dt = dt
    s.gravity[:] = gravity
    s.constraint_iterations = iterations
    s.bond_integrator = 0
    s.damping_fraction = damping
    s.sphere_enabled = 1 if sphere else 0
    s.direct_arithmetic = 1 if direct else 0
    s.plastic_yield_stretch = yield_stretch
    s.plastic_hardening = hardening
    s.contact.static_friction = 0.5
    s.contact.dynamic_friction = 0.4
    s.contact.restitution = 0.1
    s.contact.restitution_speed_threshold = 0.5
    s.contact.node_contact_radius = 0.5 * cell
    s.contact.contact_margin = 1.0e-5
    s.contact.prefilter_slack = 0.05 * ball_radius
    s.node_contact.mode = 0
    s.plane_count = len(planes)
    for k, p in enumerate(planes):
        s.planes[k] = p
    return s


def contact_bucket_mask(node_count):
    """latticeContactBucketMask (LatticePhysics.hpp:239-243): the smallest power of
    two at least twice the node count (64 at least, 2^22 at most), minus one."""
    size = 64
    while size < 2 * node_count and size < (1 << 22):
        size <<= 1
    return size - 1


def with_node_contact(s, cell, node_count, mode=2):
    """Node-node contact settings as the reference's contact tests set them
    (fast_lattice_tests.cpp:745-753): radius half a cell, skin a quarter."""
    s.node_contact.mode = mode
    s.node_contact.radius = 0.5 * cell
    s.node_contact.skin = 0.25 * cell
    s.node_contact.margin = 1.0e-5
    s.node_contact.restitution = 0.0
    s.node_contact.restitution_speed_threshold = 0.5
    s.node_contact.static_friction = 0.5
    s.node_contact.dynamic_friction = 0.4
    s.node_contact.bucket_mask = contact_bucket_mask(node_count)
    return s


def two_cells(cell=0.01, separation=0.02, bonded=False, mode=2, mass_ratio=1.0, v=(4.0, -1.0), substeps=4000,
              compliance=1.0e-6, node_mass=1.0e-3):
    """Two free cells on the x axis closing on each other, optionally joined by
    a bond at rest; dt 1e-6 s, no gravity, absolute-position arithmetic
    (the pair of fast_lattice_tests.cpp:690-755)."""
    pos = np.array([[-0.5 * separation, 0.0, 0.0], [0.5 * separation, 0.0, 0.0]])
    m = np.array([node_mass, node_mass * mass_ratio])
    if bonded:
        a, b = np.array([0], np.uint32), np.array([1], np.uint32)
        L, c = np.array([separation]), np.array([compliance])
        thr = np.array([[1.0, 2.0, 1.0, 2.0, 1.0, 2.0]])
    else:
        a = b = np.zeros(0, np.uint32)
        L = c = np.zeros(0)
        thr = np.zeros((0, 6))
    vel = np.array([[v[0], 0.0, 0.0], [v[1], 0.0, 0.0]])
    asset = Asset(reference_position=pos, mass=m, bond_a=a, bond_b=b, rest_length=L, compliance=c, threshold=thr,
                  velocity=vel)
    st = with_node_contact(settings(1.0e-6, direct=True, cell=cell), cell, 2, mode)
    return Scene(name=f"two_cells_{'bonded' if bonded else 'free'}", dims=(2, 1, 1), cell=cell, asset=asset,
                 slabs=np.array([0, 2], np.uint32), settings=st, sphere=sphere_state(center=(0.0, 100.0, 0.0)),
                 substeps=substeps)


def sphere_state(center=(0, 0, 0), velocity=(0, 0, 0), radius=1.0, mass=1.0):
    sp = abi.Sphere()
    sp.center[:] = center
    sp.velocity[:] = velocity
    sp.radius = radius
    sp.mass = mass
    sp.inertia = 0.4 * mass * radius * radius
    return sp


@dataclass
class Scene:
    """One lattice workload: asset, slabs, settings, sphere, loads and a run length."""
    name: str
    dims: tuple
    cell: float
    asset: Asset
    slabs: np.ndarray
    settings: abi.Settings
    sphere: abi.Sphere
    substeps: int
    clamped: np.ndarray | None = None        # bool (N,)
    forces: np.ndarray | None = None         # (N,3) world-space external forces
    force_substeps: int = 0
    info: dict = field(default_factory=dict)

    @property
    def node_count(self):
        return self.asset.node_count

    @property
    def bond_count(self):
        return self.asset.bond_count


def _thresholds(start, end):
    B = start.shape[0]
    t = np.empty((B, 6))
    t[:, 0] = start
    t[:, 1] = end
    t[:, 2:] = INF
    return t


def cantilever(nx, ny, nz, cell=0.01, node_mass=1.0e-3, compliance=None, strain=None, tip_load=5.0,
               tip_node=None, gravity=9.81, dt=1.0 / 1920.0, substeps=1920, blocks=1, seed=971, name="cantilever"):
    """A bar clamped on its x=0 face, gravity and a downward point load at the far end.

    compliance/strain: floats (uniform) or callables (a, b) -> per-bond values.
    """
    pos, a, b, L = box26(nx, ny, nz, cell)
    N, B = pos.shape[0], a.shape[0]
    c = np.full(B, compliance if compliance is not None else 1.0e-6) if not callable(compliance) else compliance(a, b)
    e = np.full(B, strain if strain is not None else 0.01) if not callable(strain) else strain(a, b)
    mass = np.full(N, node_mass)
    asset = Asset(reference_position=pos, mass=mass, bond_a=a, bond_b=b, rest_length=L, compliance=c,
                  threshold=_thresholds(e, e))
    gx = np.arange(N) % nx
    clamped = gx == 0
    forces = np.zeros((N, 3))
    forces[~clamped, 1] = -node_mass * gravity
    if tip_node is None:
        tip_node = (nx - 1) + nx * ((ny // 2) + ny * (nz // 2))
    forces[tip_node, 1] -= tip_load
    st = settings(dt, gravity=(0.0, 0.0, 0.0), cell=cell)
    return Scene(name=name, dims=(nx, ny, nz), cell=cell, asset=asset, slabs=box_slabs(nx, ny, nz, 1, blocks),
                 settings=st, sphere=sphere_state(center=(0.0, 100.0, 0.0)), substeps=substeps,
                 clamped=clamped, forces=forces, force_substeps=substeps,
                 info=dict(tip_node=int(tip_node), tip_load_n=tip_load))


def config1(substeps=1920, blocks=1):
    """16x16x16 cantilever, 1 cm, 1 g, c 1e-6, break strain 1%, 5 N tip load, dt 1/1920 s."""
    return cantilever(16, 16, 16, substeps=substeps, blocks=blocks, name="c1_cantilever_16^3")


def log_uniform_compliance(seed, lo=1.0e-7, hi=1.0e-5):
    def f(a, b):
        U = bond_uniform(seed, a, b)
        return np.exp(U * math.log(hi / lo)) * lo
    return f


def weibull_strain(seed, k, scale):
    def f(a, b):
        U = bond_uniform(seed, a, b)
        return scale * (-np.log1p(-U)) ** (1.0 / k)
    return f


def impact(nx, ny, nz, cell=0.01, node_mass=1.0e-3, ball_radius=0.2, ball_mass=10.0, ball_speed=8.0,
           gap=0.002, dt=1.0 / 2400.0, substeps=1200, blocks=1, seed=971, weibull_k=6.0, weibull_scale=0.015,
           name="impact"):
    """A block on a ground plane struck on its top-face centre by a rigid sphere.

    Compliance log-uniform in [1e-7, 1e-5]; tensile break strain Weibull(k, scale)
    with start == end (a step); compression and shear disabled (+inf).
    """
    pos, a, b, L = box26(nx, ny, nz, cell)
    N, B = pos.shape[0], a.shape[0]
    c = log_uniform_compliance(seed)(a, b)
    e = weibull_strain(seed + 1, weibull_k, weibull_scale)(a, b)
    asset = Asset(reference_position=pos, mass=np.full(N, node_mass), bond_a=a, bond_b=b, rest_length=L,
                  compliance=c, threshold=_thresholds(e, e))
    top = pos[:, 1].max() + 0.5 * cell
    bottom = pos[:, 1].min() - 0.5 * cell
    st = settings(dt, gravity=(0.0, -9.81, 0.0), planes=[ground_plane(bottom, 0.5 * cell)], sphere=True,
                  cell=cell, ball_radius=ball_radius)
    sp = sphere_state(center=(0.0, top + ball_radius + gap, 0.0), velocity=(0.0, -ball_speed, 0.0),
                      radius=ball_radius, mass=ball_mass)
    return Scene(name=name, dims=(nx, ny, nz), cell=cell, asset=asset, slabs=box_slabs(nx, ny, nz, 1, blocks),
                 settings=st, sphere=sp, substeps=substeps)


def config2(substeps=1200, blocks=1):
    """128 (x) x 64 (y, vertical) x 128 (z) block, 10 kg r=0.2 m sphere at 8 m/s, dt 1/2400 s."""
    return impact(128, 64, 128, substeps=substeps, blocks=blocks, name="c2_impact_128x64x128")


def small_impact(nx=24, ny=12, nz=24, ball_radius=0.05, ball_mass=2.0, ball_speed=8.0, substeps=200, blocks=1,
                 seed=971):
    """A small block of the config-2 family for parity tests."""
    return impact(nx, ny, nz, ball_radius=ball_radius, ball_mass=ball_mass, ball_speed=ball_speed,
                  substeps=substeps, blocks=blocks, seed=seed, name=f"impact_{nx}x{ny}x{nz}")


def overflow_impact(substeps=30, blocks=1, iterations=1):
    """A wide, thin slab under the config-2 sphere (r = 0.2 m, 10 kg, 8 m/s).

    The sphere's footprint puts thousands of nodes on the striker's candidate
    test per substep, more than the 1,024 a block lists
    (LatticePhysics.hpp:105): the rest are neither contacted nor predicted
    that substep (CpuLatticeBackend.cpp:377-385), the regime the 128x64x128
    impact runs in with one block.
    """
    sc = impact(48, 10, 48, ball_radius=0.2, ball_mass=10.0, substeps=substeps, blocks=blocks,
                name=f"overflow_impact_48x10x48_{blocks}blocks")
    sc.settings.constraint_iterations = iterations
    return sc


def ledge_plane(y, node_radius, footprints, restitution=0.1, static_friction=0.6, dynamic_friction=0.5):
    """A reach-capped support plane with finite footprints (TileImpactScene.cpp:322-332)."""
    p = ground_plane(y, node_radius, restitution, static_friction, dynamic_friction)
    p.reach_capped = 1
    p.footprint_count = len(footprints)
    for k, (ct, cb, ht, hb) in enumerate(footprints):
        p.footprints[k].center_t, p.footprints[k].center_b = ct, cb
        p.footprints[k].half_t, p.footprints[k].half_b = ht, hb
    return p


def bridge_impact(nx=20, ny=6, nz=10, cell=0.01, substeps=150, blocks=1, seed=971, ball_speed=6.0):
    """A block resting on two ledges under its x ends, struck at mid-span; ground below."""
    sc = impact(nx, ny, nz, cell=cell, ball_radius=0.04, ball_mass=1.0, ball_speed=ball_speed, substeps=substeps,
                blocks=blocks, seed=seed, name=f"bridge_{nx}x{ny}x{nz}")
    bottom = sc.asset.reference_position[:, 1].min() - 0.5 * cell
    half_x = 0.5 * nx * cell
    ledge = ledge_plane(bottom, 0.5 * cell, [(-half_x, 0.0, 3 * cell, INF), (half_x, 0.0, 3 * cell, INF)])
    sc.settings.plane_count = 2
    sc.settings.planes[0] = ledge
    sc.settings.planes[1] = ground_plane(bottom - 0.05, 0.5 * cell)
    return sc


def plastic_impact(nx=16, ny=8, nz=16, substeps=120, blocks=1, seed=975):
    """The config-5 material on an impact block: yield 0.4 %, hardening 0.1, Weibull(k=4, 2 %)."""
    sc = impact(nx, ny, nz, ball_radius=0.04, ball_mass=1.5, ball_speed=6.0, substeps=substeps, blocks=blocks,
                seed=seed, weibull_k=4.0, weibull_scale=0.02, name=f"plastic_{nx}x{ny}x{nz}")
    sc.settings.plastic_yield_stretch = 0.004
    sc.settings.plastic_hardening = 0.1
    return sc


def batch_edges(seed, count, lo=8, hi=48):
    """Edge lengths round(exp(U(ln lo, ln hi))) per axis for `count` lattices."""
    out = []
    for t in range(count):
        u = [(splitmix64((seed * 1000003 + t * 3 + ax) & 0xFFFFFFFFFFFFFFFF) >> 11) * 2.0 ** -53 for ax in range(3)]
        out.append(tuple(int(round(math.exp(math.log(lo) + x * (math.log(hi) - math.log(lo))))) for x in u))
    return out


def config3_trial(t, seed=971, edges=None, substeps=8 * 240):
    """One lattice of the batched form: a 26-neighbour cantilever with edges
    log-uniform in [8, 48] per axis, per-bond compliance log-uniform in
    [1e-7, 1e-5] and Weibull(k=6, 1.5%) break strain (as in the impact
    workload), clamped on x=0, gravity as loads and a tip load uniform in
    [1, 50] N; dt = 1/1920 s (8 substeps of a 1/240 s step)."""
    nx, ny, nz = edges if edges is not None else batch_edges(seed, t + 1)[t]
    u = (splitmix64((seed * 7919 + t) & 0xFFFFFFFFFFFFFFFF) >> 11) * 2.0 ** -53
    tip = 1.0 + 49.0 * u
    s = seed + 17 * t
    return cantilever(nx, ny, nz, compliance=log_uniform_compliance(s), strain=weibull_strain(s + 1, 6.0, 0.015),
                      tip_load=tip, dt=1.0 / 1920.0, substeps=substeps, name=f"c3_{t}_{nx}x{ny}x{nz}")


SERVE_DT = 1.0 / 1920.0          # 8 substeps of a 1/240 s step
SERVE_SUBSTEPS = 8 * 720          # 3 s of simulated time
MARBLE_RADIUS, MARBLE_MASS = 0.05, 0.05


def served_lattice(edges, seed=971, cell=0.01, node_mass=1.0e-3, name=None):
    """A lattice of the served trial stream: the batched form's material on a
    ground plane at y = 0 (its bottom layer resting on it), so every trial
    shares the step settings whatever its size. The sphere is set per trial
    (served_marble)."""
    nx, ny, nz = edges
    pos, a, b, L = box26(nx, ny, nz, cell)
    pos = pos.copy()
    pos[:, 1] += 0.5 * (ny - 1) * cell + 0.5 * cell
    N = pos.shape[0]
    c = log_uniform_compliance(seed)(a, b)
    e = weibull_strain(seed + 1, 6.0, 0.015)(a, b)
    asset = Asset(reference_position=pos, mass=np.full(N, node_mass), bond_a=a, bond_b=b, rest_length=L,
                  compliance=c, threshold=_thresholds(e, e))
    st = settings(SERVE_DT, gravity=(0.0, -9.81, 0.0), planes=[ground_plane(0.0, 0.5 * cell)], sphere=True,
                  cell=cell, ball_radius=MARBLE_RADIUS)
    return Scene(name=name or f"served_{nx}x{ny}x{nz}", dims=(nx, ny, nz), cell=cell, asset=asset,
                 slabs=box_slabs(nx, ny, nz, 1, 1), settings=st, sphere=served_marble(edges, 0.0, 0.0, 0.0),
                 substeps=SERVE_SUBSTEPS)


def served_marble(edges, speed, tilt, azimuth, cell=0.01, gap=0.002):
    """The trial's marble: 5 cm, 50 g, just above the top-face centre, moving
    at `speed` within `tilt` radians of straight down."""
    ny = edges[1]
    top = ny * cell
    v = (speed * math.sin(tilt) * math.cos(azimuth), -speed * math.cos(tilt), speed * math.sin(tilt) * math.sin(azimuth))
    return sphere_state(center=(0.0, top + MARBLE_RADIUS + gap, 0.0), velocity=v, radius=MARBLE_RADIUS,
                        mass=MARBLE_MASS)


def served_stream(seed, rate, seconds, templates):
    """Poisson arrivals at `rate` per second for `seconds`: (arrival time,
    template index, speed U(0.5, 3) m/s, tilt up to 0.3 rad, azimuth)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return out
        out.append((t, int(rng.integers(templates)), float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.0, 0.3)),
                    float(rng.uniform(0.0, 2.0 * math.pi))))


def blade(edge_point, speed, half_width=0.0005, node_radius=0.005, flat_normal=(1.0, 0.0, 0.0),
          direction=(0.0, -1.0, 0.0)):
    """A kinematic blade (blx_blade): 1 mm thick by default, moving down."""
    b = abi.Blade()
    b.enabled = 1
    b.edge_point[:] = edge_point
    b.flat_normal[:] = flat_normal
    b.direction[:] = direction
    b.speed = speed
    b.half_width = half_width
    b.node_radius = node_radius
    return b


def plank_cut(nx, ny, nz, cell=0.01, node_mass=1.0e-3, speed=2.0, dt=1.0 / 2880.0, substeps=12 * 600, seed=971,
              yield_stretch=0.004, hardening=0.1, weibull_k=4.0, weibull_scale=0.02, ledge_cells=16,
              name="plank_cut"):
    """A plank on two ledges under its ends, cut at mid-span by a 1 mm blade
    moving down at `speed`: plastic rest-length update (yield, hardening)
    before breaking at Weibull(k, scale) tensile strain; gravity; dt 1/2880 s
    (12 substeps of a 1/240 s step)."""
    pos, a, b, L = box26(nx, ny, nz, cell)
    N = pos.shape[0]
    c = log_uniform_compliance(seed)(a, b)
    e = weibull_strain(seed + 1, weibull_k, weibull_scale)(a, b)
    asset = Asset(reference_position=pos, mass=np.full(N, node_mass), bond_a=a, bond_b=b, rest_length=L,
                  compliance=c, threshold=_thresholds(e, e))
    bottom = pos[:, 1].min() - 0.5 * cell
    top = pos[:, 1].max() + 0.5 * cell
    half_x = 0.5 * nx * cell
    span = ledge_cells * cell
    # Two ledge footprints under the plank's ends (tangent x, bitangent z).
    feet = [(-half_x + 0.5 * span, 0.0, 0.5 * span, INF), (half_x - 0.5 * span, 0.0, 0.5 * span, INF)]
    st = settings(dt, gravity=(0.0, -9.81, 0.0), planes=[ledge_plane(bottom, 0.5 * cell, feet)], cell=cell,
                  yield_stretch=yield_stretch, hardening=hardening)
    sc = Scene(name=name, dims=(nx, ny, nz), cell=cell, asset=asset, slabs=box_slabs(nx, ny, nz, 1, 1),
               settings=st, sphere=sphere_state(center=(0.0, 100.0, 0.0)), substeps=substeps)
    sc.info = dict(blade=blade((0.0, top + 0.002, 0.0), speed, node_radius=0.5 * cell))
    return sc


def config5(substeps=12 * 600):
    """256 x 32 x 32 plank, plastic (0.4%, H 0.1), Weibull(4, 2%), blade at 2 m/s."""
    return plank_cut(256, 32, 32, substeps=substeps, name="c5_plank_cut_256x32x32")
