"""Python face of the lattice step's C ABI (include/banjo_lattice.h).

`LatticeLibrary` binds one shared library that exports the lattice ABI under a
name prefix. The product is the HIP library (prefix ``blx_``); the same binder
serves any library with the same function shapes, which is how the parity
tests put implementations side by side. The objects mirror the reference's C++
lattice API (src/fastlattice/FastLattice.hpp): `Schedule` is LatticeSchedule,
`State` is LatticeState, `Backend` is LatticeBackend.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
PRODUCT_LIBRARY = os.path.join(_HERE, "lib", "libbanjo_lattice.so")

_ERRORS = {
    abi.BLX_ERR_RUNTIME: RuntimeError,
    abi.BLX_ERR_INVALID: ValueError,
    abi.BLX_ERR_OVERFLOW: OverflowError,
    abi.BLX_ERR_LOGIC: AssertionError,
}


class LatticeError(Exception):
    pass


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class Asset:
    """LatticeAsset + ActiveMatter in asset order (src/matter/Lattice.hpp:26-44,82-93)."""
    reference_position: np.ndarray  # (N, 3) float64, world
    mass: np.ndarray                # (N,)
    bond_a: np.ndarray              # (B,) uint32, a < b
    bond_b: np.ndarray
    rest_length: np.ndarray         # (B,)
    compliance: np.ndarray          # (B,)
    threshold: np.ndarray           # (B, 6): ds, de, cs, ce, ss, se
    position: np.ndarray | None = None
    velocity: np.ndarray | None = None
    alive: np.ndarray | None = None
    damage: np.ndarray | None = None
    failure_mode: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def node_count(self):
        return int(self.mass.shape[0])

    @property
    def bond_count(self):
        return int(self.bond_a.shape[0])

    def view(self):
        n = self.node_count
        ref = np.ascontiguousarray(self.reference_position, dtype=np.float64).reshape(n, 3)
        pos = ref if self.position is None else np.ascontiguousarray(self.position, dtype=np.float64).reshape(n, 3)
        vel = np.zeros((n, 3)) if self.velocity is None else np.ascontiguousarray(self.velocity, dtype=np.float64).reshape(n, 3)
        arrays = dict(
            mass=np.ascontiguousarray(self.mass, dtype=np.float64),
            a=np.ascontiguousarray(self.bond_a, dtype=np.uint32),
            b=np.ascontiguousarray(self.bond_b, dtype=np.uint32),
            L=np.ascontiguousarray(self.rest_length, dtype=np.float64),
            c=np.ascontiguousarray(self.compliance, dtype=np.float64),
            t=np.ascontiguousarray(self.threshold, dtype=np.float64).reshape(-1, 6),
            alive=None if self.alive is None else np.ascontiguousarray(self.alive, dtype=np.uint8),
            damage=None if self.damage is None else np.ascontiguousarray(self.damage, dtype=np.float64),
            mode=None if self.failure_mode is None else np.ascontiguousarray(self.failure_mode, dtype=np.uint8),
            ref=ref, pos=pos, vel=vel)
        v = abi.AssetView()
        v.node_count = n
        v.bond_count = self.bond_count
        v.reference_position = _ptr(ref, C.c_double)
        v.position = _ptr(pos, C.c_double)
        v.velocity = _ptr(vel, C.c_double)
        v.mass = _ptr(arrays["mass"], C.c_double)
        v.bond_a = _ptr(arrays["a"], C.c_uint32)
        v.bond_b = _ptr(arrays["b"], C.c_uint32)
        v.rest_length = _ptr(arrays["L"], C.c_double)
        v.compliance = _ptr(arrays["c"], C.c_double)
        v.threshold = _ptr(arrays["t"], C.c_double)
        v.alive = _ptr(arrays["alive"], C.c_uint8)
        v.damage = _ptr(arrays["damage"], C.c_double)
        v.failure_mode = _ptr(arrays["mode"], C.c_uint8)
        return v, arrays


class Schedule:
    """LatticeSchedule (src/fastlattice/LatticeSchedule.hpp:28-46) as numpy arrays."""

    def __init__(self, node_count, bond_count, block_count):
        self.node_count, self.bond_count, self.block_count = node_count, bond_count, block_count
        self.color_count = 0
        self.boundary_bond_count = 0
        B = bond_count
        self.node_block_begin = np.zeros(block_count + 1, np.uint32)
        self.bond_order = np.zeros(B, np.uint32)
        self.bond_schedule_index = np.zeros(B, np.uint32)
        self.bond_block_begin = np.zeros(block_count + 1, np.uint32)
        self.range_begin = np.zeros(block_count * 2 * 64, np.uint32)
        self.range_end = np.zeros(block_count * 2 * 64, np.uint32)
        self.bond_color = np.zeros(B, np.uint8)

    def view(self):
        v = abi.ScheduleView()
        v.node_count, v.bond_count, v.block_count = self.node_count, self.bond_count, self.block_count
        v.color_count, v.boundary_bond_count = self.color_count, self.boundary_bond_count
        for name, ct in (("node_block_begin", C.c_uint32), ("bond_order", C.c_uint32),
                         ("bond_schedule_index", C.c_uint32), ("bond_block_begin", C.c_uint32),
                         ("range_begin", C.c_uint32), ("range_end", C.c_uint32), ("bond_color", C.c_uint8)):
            setattr(v, name, _ptr(getattr(self, name), ct))
        return v

    def trim(self):
        """Shrink the range table to block_count * 2 * max(color_count, 1) entries."""
        n = self.block_count * 2 * max(self.color_count, 1)
        self.range_begin = np.ascontiguousarray(self.range_begin[:n])
        self.range_end = np.ascontiguousarray(self.range_end[:n])
        return self


_STATE_F64_NODE3 = ("x0", "u", "u_prev", "v")
_STATE_F64_NODE = ("inv_mass", "mass")
_STATE_F64_BOND = ("rest_length", "rest_length_sq_minus", "weight", "compliance", "damage",
                   "prev_tensile", "prev_compressive", "prev_shear", "plastic_extension", "plastic_strain")
_STATE_U32_BOND = ("bond_slot_a", "bond_slot_b", "bond_a", "bond_b")
_STATE_U8_BOND = ("alive", "failure_mode")
DYNAMIC_FIELDS = ("u", "u_prev", "v", "alive", "failure_mode", "damage", "prev_tensile",
                  "prev_compressive", "prev_shear", "plastic_extension", "plastic_strain")


class State:
    """LatticeState (src/fastlattice/FastLattice.hpp:37-55) as numpy arrays."""

    def __init__(self, N, B, D):
        self.origin = np.zeros(3)
        self.node_count, self.bond_count, self.max_degree = N, B, D
        for n in _STATE_F64_NODE3:
            setattr(self, n, np.zeros(3 * N))
        for n in _STATE_F64_NODE:
            setattr(self, n, np.zeros(N))
        for n in _STATE_F64_BOND:
            setattr(self, n, np.zeros(B))
        for n in _STATE_U32_BOND:
            setattr(self, n, np.zeros(B, np.uint32))
        for n in _STATE_U8_BOND:
            setattr(self, n, np.zeros(B, np.uint8))
        self.rest_edge = np.zeros(3 * B)
        self.threshold = np.zeros(6 * B)
        self.adj_offsets = np.zeros(N + 1, np.uint32)
        self.adj_bonds = np.zeros(2 * B, np.uint32)
        P_ = D * N
        self.nbr_bond = np.zeros(P_, np.uint32)
        self.nbr_other = np.zeros(P_, np.uint32)
        self.nbr_rest = np.zeros(3 * P_)
        self.nbr_weight = np.zeros(P_)
        self.nbr_alive = np.zeros(P_, np.uint8)

    def view(self):
        v = abi.StateView()
        for k in range(3):
            v.origin[k] = float(self.origin[k])
        v.node_count, v.bond_count, v.max_degree = self.node_count, self.bond_count, self.max_degree
        for name, ct in abi.StateView._fields_:
            if name in ("origin", "node_count", "bond_count", "max_degree"):
                continue
            elem = ct._type_
            setattr(v, name, _ptr(getattr(self, name), elem))
        return v

    def copy(self):
        s = State.__new__(State)
        for k, val in self.__dict__.items():
            setattr(s, k, val.copy() if isinstance(val, np.ndarray) else val)
        return s


class LatticeLibrary:
    """One library exporting the lattice C ABI under `prefix`."""

    def __init__(self, path=PRODUCT_LIBRARY, prefix="blx_", name=None):
        if not os.path.exists(path):
            raise LatticeError(f"lattice library not built: {path}")
        self.path = path
        self.prefix = prefix
        self.name = name or prefix.rstrip("_")
        self.lib = C.CDLL(path)
        err = getattr(self.lib, prefix + "last_error")
        err.restype = C.c_char_p
        self._err = err

    def fn(self, name):
        f = getattr(self.lib, self.prefix + name)
        if f.restype is not C.c_int:
            f.restype = C.c_int
        return f

    def has(self, name):
        return hasattr(self.lib, self.prefix + name)

    def description(self) -> str:
        """The visible devices (blx_description): name, arch, CUs, memory, LDS per workgroup."""
        buf = C.create_string_buffer(4096)
        self.check(self.fn("description")(buf, C.c_size_t(len(buf))), "description")
        return buf.value.decode()

    def check(self, rc, what):
        if rc != abi.BLX_OK:
            msg = self._err().decode(errors="replace")
            raise _ERRORS.get(rc, LatticeError)(f"{self.name}.{what}: {msg}")

    # -- host builders ------------------------------------------------------
    def schedule_build(self, asset: Asset, node_block_begin) -> Schedule:
        nbb = np.ascontiguousarray(node_block_begin, dtype=np.uint32)
        blocks = len(nbb) - 1
        s = Schedule(asset.node_count, asset.bond_count, blocks)
        av, keep = asset.view()
        sv = s.view()
        self.check(self.fn("schedule_build")(C.byref(av), _ptr(nbb, C.c_uint32), C.c_uint32(blocks), C.byref(sv)),
                   "schedule_build")
        s.color_count, s.boundary_bond_count, s.block_count = sv.color_count, sv.boundary_bond_count, sv.block_count
        return s.trim()

    def state_build(self, asset: Asset, schedule: Schedule, origin=(0.0, 0.0, 0.0)) -> State:
        av, keep = asset.view()
        d = C.c_uint32()
        self.check(self.fn("state_max_degree")(C.byref(av), C.byref(d)), "state_max_degree")
        st = State(asset.node_count, asset.bond_count, d.value)
        org = (C.c_double * 3)(*origin)
        stv = st.view()
        sv = schedule.view()
        self.check(self.fn("state_build")(C.byref(av), C.byref(sv), org, C.byref(stv)), "state_build")
        st.origin = np.array(stv.origin[:])
        st.max_degree = stv.max_degree
        return st

    def components(self, node_count, bond_a, bond_b, alive):
        a = np.ascontiguousarray(bond_a, np.uint32)
        b = np.ascontiguousarray(bond_b, np.uint32)
        al = np.ascontiguousarray(alive, np.uint8)
        label = np.zeros(node_count, np.uint32)
        count = C.c_uint32()
        name = "components" if self.has("components") else "components_host"
        self.check(self.fn(name)(C.c_uint32(node_count), C.c_uint32(len(a)), _ptr(a, C.c_uint32),
                                 _ptr(b, C.c_uint32), _ptr(al, C.c_uint8), _ptr(label, C.c_uint32),
                                 C.byref(count)), name)
        return label, count.value

    def fragment_mass(self, position, velocity, mass, cell_size, offsets, nodes, spin=None):
        """calculateFragmentMassProperties (src/fracture/FragmentMassProperties.cpp:7-80)
        of fragments given as node lists, fragment f = nodes[offsets[f]:offsets[f+1]]
        (oracle and reference libraries)."""
        pos = np.ascontiguousarray(position, np.float64).reshape(-1)
        vel = np.ascontiguousarray(velocity, np.float64).reshape(-1)
        m = np.ascontiguousarray(mass, np.float64)
        off = np.ascontiguousarray(offsets, np.uint32)
        nd = np.ascontiguousarray(nodes, np.uint32)
        sp = None if spin is None else np.ascontiguousarray(spin, np.float64).reshape(-1)
        k = len(off) - 1
        out = np.zeros(k, abi.FRAGMENT_MASS_DTYPE)
        self.check(self.fn("fragment_mass")(
            C.c_uint32(len(m)), _ptr(pos, C.c_double), _ptr(vel, C.c_double),
            None if sp is None else _ptr(sp, C.c_double), _ptr(m, C.c_double), C.c_double(cell_size),
            C.c_uint32(k), _ptr(off, C.c_uint32), _ptr(nd, C.c_uint32),
            out.ctypes.data_as(C.POINTER(abi.FragmentMass))), "fragment_mass")
        return out

    def fragment_lattice(self, asset: Asset, nodes, cell_offset, center_of_mass, orientation_wxyz=(1.0, 0.0, 0.0, 0.0),
                         linear_velocity=(0.0, 0.0, 0.0), angular_velocity=(0.0, 0.0, 0.0), plastic_extension=None,
                         plastic_strain=None):
        """buildFragmentLattice (src/fastlattice/Refracture.cpp:267-393): the
        sub-lattice of the parent's cells `nodes`, placed on the pose. Returns
        (Asset of the sub-lattice, dict with parent_node, parent_bond, origin,
        carried_spin, plastic_extension, plastic_strain)."""
        av, _keep = asset.view()
        nd = np.ascontiguousarray(nodes, np.uint32)
        off = np.ascontiguousarray(cell_offset, np.float64).reshape(-1)
        if len(off) != 3 * asset.node_count:
            raise ValueError("cell_offset holds three components per parent node")
        pose = abi.FragmentPose()
        pose.center_of_mass[:] = center_of_mass
        pose.orientation_wxyz[:] = orientation_wxyz
        pose.linear_velocity[:] = linear_velocity
        pose.angular_velocity[:] = angular_velocity
        pe = None if plastic_extension is None else np.ascontiguousarray(plastic_extension, np.float64)
        ps = None if plastic_strain is None else np.ascontiguousarray(plastic_strain, np.float64)
        fn = self.fn("fragment_lattice")
        out = abi.FragmentLatticeView()
        args = (C.byref(av), _ptr(nd, C.c_uint32), C.c_uint32(len(nd)), _ptr(off, C.c_double), C.byref(pose),
                None if pe is None else _ptr(pe, C.c_double), None if ps is None else _ptr(ps, C.c_double))
        self.check(fn(*args, C.byref(out)), "fragment_lattice")
        n, k = out.node_count, out.bond_count
        arr = dict(parent_node=np.zeros(n, np.uint32), parent_bond=np.zeros(k, np.uint32),
                   reference_position=np.zeros((n, 3)), position=np.zeros((n, 3)), velocity=np.zeros((n, 3)),
                   mass=np.zeros(n), bond_a=np.zeros(k, np.uint32), bond_b=np.zeros(k, np.uint32),
                   rest_length=np.zeros(k), compliance=np.zeros(k), threshold=np.zeros((k, 6)),
                   alive=np.zeros(k, np.uint8), damage=np.zeros(k), failure_mode=np.zeros(k, np.uint8),
                   plastic_extension=np.zeros(k), plastic_strain=np.zeros(k))
        out.node_capacity, out.bond_capacity = n, k
        ctype = {np.dtype(np.uint32): C.c_uint32, np.dtype(np.float64): C.c_double, np.dtype(np.uint8): C.c_uint8}
        for name, a in arr.items():
            setattr(out, name, _ptr(a, ctype[a.dtype]))
        self.check(fn(*args, C.byref(out)), "fragment_lattice")
        sub = Asset(reference_position=arr["reference_position"], mass=arr["mass"], bond_a=arr["bond_a"],
                    bond_b=arr["bond_b"], rest_length=arr["rest_length"], compliance=arr["compliance"],
                    threshold=arr["threshold"], position=arr["position"], velocity=arr["velocity"],
                    alive=arr["alive"], damage=arr["damage"], failure_mode=arr["failure_mode"])
        info = dict(parent_node=arr["parent_node"], parent_bond=arr["parent_bond"], origin=tuple(out.origin),
                    carried_spin=tuple(out.carried_spin), plastic_extension=arr["plastic_extension"],
                    plastic_strain=arr["plastic_strain"])
        return sub, info

    def state_metrics(self, state: "State"):
        """(substep limit s, elastic energy J, kinetic energy J) of a state:
        latticeStateSubstepLimit / ElasticEnergy / KineticEnergy
        (src/fastlattice/FastLattice.cpp:30-45,199-224)."""
        sv = state.view()
        out = []
        for name in ("state_substep_limit", "state_elastic_energy", "state_kinetic_energy"):
            x = C.c_double()
            self.check(self.fn(name)(C.byref(sv), C.byref(x)), name)
            out.append(x.value)
        return tuple(out)

    def backend(self, precision, schedule: Schedule, *extra):
        return Backend(self, precision, schedule, *extra)


def _mass_properties(library, name, head, node_count, cell_size, spin):
    sp = None if spin is None else np.ascontiguousarray(spin, np.float64).reshape(-1)
    if sp is not None and len(sp) != 3 * node_count:
        raise ValueError("spin holds three components per node")
    out = np.empty(max(node_count, 1), abi.FRAGMENT_MASS_DTYPE)
    count = C.c_uint32()
    library.check(library.fn(name)(*head, C.c_double(cell_size), None if sp is None else _ptr(sp, C.c_double),
                                   out.ctypes.data_as(C.POINTER(abi.FragmentMass)), C.c_uint32(len(out)),
                                   C.byref(count)), name)
    return out[:count.value].copy()


class _Tunable:
    """Launch tuning and the autotuner (blx_*_tuning, blx_*_autotune); _kind is
    "backend" or "batch"."""
    _kind = "backend"

    def tuning(self) -> dict:
        t = abi.Tuning()
        self.L.check(self.L.fn(f"{self._kind}_get_tuning")(self.h, C.byref(t)), "get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = abi.Tuning(**self.tuning())
        for k, v in kw.items():
            if not hasattr(t, k):
                raise TypeError(f"unknown tuning field {k}")
            setattr(t, k, int(v))
        self.L.check(self.L.fn(f"{self._kind}_set_tuning")(self.h, C.byref(t)), "set_tuning")

    def autotune(self, warm_substeps=32, probe_substeps=64) -> dict:
        """Runs the autotuner; returns the tuning it kept and what it measured."""
        rep = abi.AutotuneReport()
        self.L.check(self.L.fn(f"{self._kind}_autotune")(self.h, C.c_uint64(warm_substeps),
                                                         C.c_uint64(probe_substeps), C.byref(rep)), "autotune")
        tried = [dict(rep.tried[i].as_dict(), us_per_substep=rep.seconds_per_substep[i] * 1e6)
                 for i in range(rep.candidates)]
        return {"chosen": rep.chosen.as_dict(), "tried": tried}


class Backend(_Tunable):
    """LatticeBackend (src/fastlattice/FastLattice.hpp:255-369) over the C ABI."""

    def __init__(self, library: LatticeLibrary, precision, schedule: Schedule, *extra):
        self.L = library
        self.schedule = schedule
        self.node_count = schedule.node_count
        self.bond_count = schedule.bond_count
        self._sv = schedule.view()
        h = C.c_void_p()
        create = self.L.fn("backend_create")
        if library.prefix == "blx_":
            device = extra[0] if extra else 0
            rc = create(C.c_int(device), C.c_int(precision), C.byref(self._sv), C.byref(h))
        elif extra:
            rc = create(C.c_int(precision), C.byref(self._sv), *extra, C.byref(h))
        else:
            rc = create(C.c_int(precision), C.byref(self._sv), C.byref(h))
        self.L.check(rc, "backend_create")
        self.h = h
        self.precision = precision

    def close(self):
        if self.h:
            self.L.fn("backend_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, state: State, settings: abi.Settings, sphere: abi.Sphere):
        sv = state.view()
        self.L.check(self.L.fn("backend_upload")(self.h, C.byref(sv), C.byref(settings), C.byref(sphere)), "upload")

    def set_blade(self, blade):
        self.L.check(self.L.fn("backend_set_blade")(self.h, C.byref(blade)), "set_blade")

    def set_external_forces(self, forces, substeps):
        if forces is None:
            self.L.check(self.L.fn("backend_set_external_forces")(self.h, None, C.c_uint64(substeps)),
                         "set_external_forces")
            return
        f = np.ascontiguousarray(forces, np.float64).reshape(-1)
        self.L.check(self.L.fn("backend_set_external_forces")(self.h, _ptr(f, C.c_double), C.c_uint64(substeps)),
                     "set_external_forces")

    def run(self, control: abi.RunControl) -> abi.RunStatus:
        st = abi.RunStatus()
        self.L.check(self.L.fn("backend_run")(self.h, C.byref(control), C.byref(st)), "run")
        return st

    def status(self) -> abi.RunStatus:
        st = abi.RunStatus()
        self.L.check(self.L.fn("backend_status")(self.h, C.byref(st)), "status")
        return st

    def download(self, state: State):
        sphere = abi.Sphere()
        sv = state.view()
        self.L.check(self.L.fn("backend_download")(self.h, C.byref(sv), C.byref(sphere)), "download")
        return sphere

    def first_failure_bonds(self):
        count = C.c_uint64()
        f = self.L.fn("backend_first_failure_bonds")
        self.L.check(f(self.h, None, C.c_uint64(0), C.byref(count)), "first_failure_bonds")
        out = np.zeros(count.value, np.uint32)
        self.L.check(f(self.h, _ptr(out, C.c_uint32), C.c_uint64(count.value), C.byref(count)),
                     "first_failure_bonds")
        return out

    def frames(self):
        n = C.c_uint32()
        self.L.check(self.L.fn("backend_frame_count")(self.h, C.byref(n)), "frame_count")
        out = []
        for k in range(n.value):
            step = C.c_uint64()
            u = np.zeros(3 * self.node_count, np.float32)
            alive = np.zeros(self.bond_count, np.uint8)
            damage = np.zeros(self.bond_count, np.float32)
            sp = abi.Sphere()
            self.L.check(self.L.fn("backend_take_frame")(self.h, C.c_uint32(k), C.byref(step), _ptr(u, C.c_float),
                                                         _ptr(alive, C.c_uint8), _ptr(damage, C.c_float),
                                                         C.byref(sp)), "take_frame")
            out.append(dict(step=step.value, u=u, alive=alive, damage=damage, sphere=sp))
        self.L.check(self.L.fn("backend_clear_frames")(self.h), "clear_frames")
        return out

    def fragment_labels(self):
        label = np.zeros(self.node_count, np.uint32)
        count = C.c_uint32()
        self.L.check(self.L.fn("backend_fragment_labels")(self.h, _ptr(label, C.c_uint32), C.byref(count)),
                     "fragment_labels")
        return label, count.value

    def fragment_mass_properties(self, cell_size, spin=None):
        """Mass properties of every fragment of the current state, in canonical
        label order (blx_backend_fragment_mass_properties)."""
        return _mass_properties(self.L, "backend_fragment_mass_properties", (self.h,), self.node_count,
                                cell_size, spin)


class Batch(_Tunable):
    """A batch of independent lattices advanced together (blx_batch_* of
    include/banjo_lattice.h). Slots share the step settings; each slot has
    its own state, sphere, loads, run control and status."""
    _kind = "batch"

    def __init__(self, library: LatticeLibrary, precision, capacity, device=0):
        self.L = library
        self.capacity = int(capacity)
        self.precision = precision
        h = C.c_void_p()
        self.L.check(self.L.fn("batch_create")(C.c_int(device), C.c_int(precision), C.c_uint32(self.capacity),
                                               C.byref(h)), "batch_create")
        self.h = h
        self.sizes = {}

    def close(self):
        if self.h:
            self.L.fn("batch_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, slot, schedule: Schedule, state: State, settings: abi.Settings, sphere: abi.Sphere):
        sv, st = schedule.view(), state.view()
        self.L.check(self.L.fn("batch_upload")(self.h, C.c_uint32(slot), C.byref(sv), C.byref(st), C.byref(settings),
                                               C.byref(sphere)), "batch_upload")
        self.sizes[slot] = (schedule.node_count, schedule.bond_count)

    def clone(self, dst, src, sphere: abi.Sphere):
        self.L.check(self.L.fn("batch_clone")(self.h, C.c_uint32(dst), C.c_uint32(src), C.byref(sphere)),
                     "batch_clone")
        self.sizes[dst] = self.sizes[src]

    def set_external_forces(self, slot, forces, substeps):
        f = None if forces is None else np.ascontiguousarray(forces, np.float64).reshape(-1)
        self.L.check(self.L.fn("batch_set_external_forces")(self.h, C.c_uint32(slot),
                                                            None if f is None else _ptr(f, C.c_double),
                                                            C.c_uint64(substeps)), "batch_set_external_forces")

    def set_blade(self, blade):
        self.L.check(self.L.fn("batch_set_blade")(self.h, C.byref(blade)), "batch_set_blade")

    def arm(self, slot, control: abi.RunControl):
        self.L.check(self.L.fn("batch_arm")(self.h, C.c_uint32(slot), C.byref(control)), "batch_arm")

    def advance(self, substeps, steps_per_launch=0):
        running = C.c_uint32()
        self.L.check(self.L.fn("batch_advance")(self.h, C.c_uint64(substeps), C.c_uint64(steps_per_launch), None,
                                                C.byref(running)), "batch_advance")
        return running.value

    def status(self, slot) -> abi.RunStatus:
        st = abi.RunStatus()
        self.L.check(self.L.fn("batch_status")(self.h, C.c_uint32(slot), C.byref(st)), "batch_status")
        return st

    def exit_reasons(self, count):
        """The exit reason of slots 0 .. count - 1 (0 while a slot runs), in one call."""
        out = np.zeros(int(count), dtype=np.uint32)
        self.L.check(self.L.fn("batch_exit_reasons")(self.h, out.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                      C.c_uint32(int(count))), "batch_exit_reasons")
        return out

    def download(self, slot, state: State):
        sphere = abi.Sphere()
        sv = state.view()
        self.L.check(self.L.fn("batch_download")(self.h, C.c_uint32(slot), C.byref(sv), C.byref(sphere)),
                     "batch_download")
        return sphere

    def sphere(self, slot) -> abi.Sphere:
        sp = abi.Sphere()
        self.L.check(self.L.fn("batch_sphere")(self.h, C.c_uint32(slot), C.byref(sp)), "batch_sphere")
        return sp

    def first_failure_bonds(self, slot):
        count = C.c_uint64()
        f = self.L.fn("batch_first_failure_bonds")
        self.L.check(f(self.h, C.c_uint32(slot), None, C.c_uint64(0), C.byref(count)), "batch_first_failure_bonds")
        out = np.zeros(count.value, np.uint32)
        self.L.check(f(self.h, C.c_uint32(slot), _ptr(out, C.c_uint32), C.c_uint64(count.value), C.byref(count)),
                     "batch_first_failure_bonds")
        return out

    def fragment_labels(self, slot):
        label = np.zeros(self.sizes[slot][0], np.uint32)
        count = C.c_uint32()
        self.L.check(self.L.fn("batch_fragment_labels")(self.h, C.c_uint32(slot), _ptr(label, C.c_uint32),
                                                        C.byref(count)), "batch_fragment_labels")
        return label, count.value

    def fragment_mass_properties(self, slot, cell_size, spin=None):
        return _mass_properties(self.L, "batch_fragment_mass_properties", (self.h, C.c_uint32(slot)),
                                self.sizes[slot][0], cell_size, spin)

    def release(self, slot):
        self.L.check(self.L.fn("batch_release")(self.h, C.c_uint32(slot)), "batch_release")
        self.sizes.pop(slot, None)

    def slot_bytes(self, slot):
        b = C.c_uint64()
        self.L.check(self.L.fn("batch_slot_bytes")(self.h, C.c_uint32(slot), C.byref(b)), "batch_slot_bytes")
        return b.value


def control(max_steps, **kw) -> abi.RunControl:
    """RunControl with the reference's defaults (src/fastlattice/FastLattice.hpp:79-128)."""
    c = abi.RunControl()
    c.max_steps = int(max_steps)
    c.energy_flat_fraction = 1.0e-3
    c.calm_damage_margin = 0.5
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def product() -> LatticeLibrary:
    """The HIP library. Raises if it is not built: there is no fallback path."""
    return LatticeLibrary(PRODUCT_LIBRARY, "blx_", "hip")
