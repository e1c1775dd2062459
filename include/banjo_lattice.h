/*
 * banjo_lattice.h -- C ABI of the MI355X bonded-lattice fracture step.
 *
 * This is the drop-in boundary of the fast explicit lattice lane. Every entry
 * point below replaces a piece of the reference's C++ lattice API
 * (src/fastlattice/FastLattice.hpp), cited per function. The structs mirror
 * the reference's value types field for field, widened to plain C:
 *
 *   blx_settings      <- StepSettings<double>     (src/fastlattice/LatticePhysics.hpp:249-284)
 *   blx_sphere        <- SphereState<double>      (src/fastlattice/LatticePhysics.hpp:145-148)
 *   blx_support_plane <- SupportPlane<double>     (src/fastlattice/LatticePhysics.hpp:122-136)
 *   blx_run_control   <- RunControl               (src/fastlattice/FastLattice.hpp:79-128)
 *   blx_run_status    <- RunStatus                (src/fastlattice/FastLattice.hpp:182-245)
 *   blx_schedule_view <- LatticeSchedule          (src/fastlattice/LatticeSchedule.hpp:28-46)
 *   blx_state_view    <- LatticeState             (src/fastlattice/FastLattice.hpp:37-55)
 *   blx_asset_view    <- LatticeAsset bonds + ActiveMatter nodes/bonds
 *                        (src/matter/Lattice.hpp:26-44,82-93; src/fracture/ActiveMatter.hpp:18-57)
 *
 * No C++ and no HIP types cross this boundary: plain pointers, sizes and
 * integer status codes. Every function returns BLX_OK or a negative code; the
 * message of the last failure on the calling thread is blx_last_error(). The
 * codes map onto the reference's exception types (FastLattice.hpp:400-406,
 * CudaLatticeUnavailable.cpp:19-25):
 *   BLX_ERR_RUNTIME   std::runtime_error   (no device, device failure)
 *   BLX_ERR_INVALID   std::invalid_argument (bad input, unsupported feature)
 *   BLX_ERR_OVERFLOW  std::overflow_error  (ledger or counter overflow)
 *   BLX_ERR_LOGIC     std::logic_error     (misuse: run before upload, ...)
 *
 * Threading: one host thread drives a given backend handle; many handles may
 * coexist, on one device or several.
 */
#ifndef BANJO_LATTICE_H
#define BANJO_LATTICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLX_OK 0
#define BLX_ERR_RUNTIME (-1)
#define BLX_ERR_INVALID (-2)
#define BLX_ERR_OVERFLOW (-3)
#define BLX_ERR_LOGIC (-4)

#define BLX_PRECISION_FLOAT 0
#define BLX_PRECISION_DOUBLE 1

#define BLX_MAX_SUPPORT_PLANES 2          /* LatticePhysics.hpp:103 */
#define BLX_MAX_FOOTPRINTS 4              /* LatticePhysics.hpp:104 */
#define BLX_MAX_CANDIDATES_PER_BLOCK 1024 /* LatticePhysics.hpp:105 */
#define BLX_NO_BOND 0xffffffffu           /* LatticePhysics.hpp:106 */
#define BLX_PHASE_COUNT 16                /* FastLattice.hpp:130 */

#define BLX_NODE_CONTACT_OFF 0
#define BLX_NODE_CONTACT_MEASURE 1
#define BLX_NODE_CONTACT_ON 2

#define BLX_BOND_XPBD 0
#define BLX_BOND_VELOCITY_VERLET 1

typedef struct blx_footprint {
    double center_t, center_b, half_t, half_b;
} blx_footprint;

typedef struct blx_support_plane {
    double point[3], normal[3], tangent[3], bitangent[3];
    double restitution, static_friction, dynamic_friction;
    double node_radius;
    uint8_t reach_capped;
    uint32_t footprint_count;
    blx_footprint footprints[BLX_MAX_FOOTPRINTS];
} blx_support_plane;

typedef struct blx_contact_settings {
    double static_friction, dynamic_friction, restitution, restitution_speed_threshold;
    double node_contact_radius, contact_margin, prefilter_slack;
} blx_contact_settings;

typedef struct blx_node_contact_settings {
    uint8_t mode;
    double radius, skin, margin;
    double restitution, restitution_speed_threshold;
    double static_friction, dynamic_friction;
    uint32_t bucket_mask;
} blx_node_contact_settings;

typedef struct blx_settings {
    double dt;
    double gravity[3];
    uint32_t constraint_iterations;
    uint8_t bond_integrator;
    double damping_fraction;
    uint8_t sphere_enabled;
    uint8_t direct_arithmetic;
    uint8_t audit_energy;
    double plastic_yield_stretch;
    double plastic_hardening;
    double plate_half_thickness;
    blx_contact_settings contact;
    blx_node_contact_settings node_contact;
    uint32_t plane_count;
    blx_support_plane planes[BLX_MAX_SUPPORT_PLANES];
} blx_settings;

typedef struct blx_sphere {
    double center[3], velocity[3], angular_velocity[3];
    double radius, mass, inertia;
} blx_sphere;

typedef struct blx_run_control {
    uint64_t max_steps;
    uint64_t quiet_steps;
    uint64_t min_steps;
    uint64_t no_failure_steps;
    uint64_t energy_flat_steps;
    double energy_flat_fraction; /* reference default 1e-3 */
    uint64_t calm_steps;
    double calm_damage_margin;   /* reference default 0.5 */
    double removable_energy_j;
    uint64_t capture_stride;
    uint32_t max_frames;
    uint64_t steps_per_launch;   /* substeps per device batch; 0 = the whole run in one batch */
    const volatile uint8_t *abandon; /* polled between batches; NULL never stops */
    /* RunControl::abandon as a callback, polled between batches when set:
     * nonzero stops the run (exit reason 7). A caller holding the reference's
     * std::atomic<bool> passes a function that loads it, with the atomic as
     * the context. */
    int (*abandon_poll)(const void *context);
    const void *abandon_context;
} blx_run_control;

typedef struct blx_contact_accumulators {
    uint64_t impulse_contacts;
    double impulse_to_material[3];
    double dissipated_kinetic_energy_j;
    double maximum_penetration_m;
    double maximum_position_correction_m;
    double maximum_center_shift_m;
    uint64_t candidate_overflow;
    uint64_t ball_support_events;
} blx_contact_accumulators;

typedef struct blx_load_ledger {
    uint64_t steps;
    double elapsed_s;
    double requested_impulse_n_s[3], impulse_n_s[3];
    double requested_angular_impulse_kg_m2_s[3], angular_impulse_kg_m2_s[3];
    double work_j;
} blx_load_ledger;

/* NodeContactAccumulators (LatticePhysics.hpp:204-222): the node-node
 * contact ledger, summed in pair order. */
typedef struct blx_node_contact_accumulators {
    uint64_t contacts, pair_tests, rebuilds, pairs_listed, pair_overflow;
    double dissipated_kinetic_energy_j;
    double maximum_overlap_m;
    double maximum_position_correction_m;
    double momentum_residual[3];
} blx_node_contact_accumulators;

typedef struct blx_run_status {
    uint64_t total_steps;
    uint64_t first_failure_step;   /* UINT64_MAX: none */
    uint64_t last_failure_step;    /* UINT64_MAX: none */
    uint32_t failure_rounds;
    uint32_t broken_bonds;
    double removed_energy_j;
    uint64_t last_energy_gain_step; /* UINT64_MAX: none */
    double max_damage;
    uint64_t last_damage_gain_step;
    blx_contact_accumulators contact;
    blx_load_ledger external_load;
    double damping_dissipated_j;
    double striker_dissipated_j;
    uint8_t energy_audited;
    uint32_t exit_reason;
    uint32_t launches;
    double kernel_seconds;
    double wall_seconds;
    uint32_t frames_captured;
    double phase_seconds[BLX_PHASE_COUNT];
    float max_tensile_stretch, max_compressive_strain, max_shear_strain;
    uint32_t rank_deficient_nodes;
    double plastic_work_j;
    float max_plastic_stretch;
    blx_node_contact_accumulators node_contact;
    /* RunStatus::shared_memory_positions / shared_memory_bytes
     * (FastLattice.hpp:229-231): the sweep kept this lattice's nodes in LDS,
     * and the bytes they took (0: none). */
    uint8_t shared_memory_positions;
    uint64_t shared_memory_bytes;
} blx_run_status;

/* LatticeAsset bonds + ActiveMatter nodes and bond state, asset order. */
typedef struct blx_asset_view {
    uint32_t node_count, bond_count;
    const double *reference_position; /* 3N world */
    const double *position;           /* 3N world, current */
    const double *velocity;           /* 3N */
    const double *mass;               /* N */
    const uint32_t *bond_a, *bond_b;  /* B */
    const double *rest_length;        /* B */
    const double *compliance;         /* B */
    const double *threshold;          /* 6B: ds, de, cs, ce, ss, se */
    const uint8_t *alive;             /* B, NULL = all alive */
    const double *damage;             /* B, NULL = zero */
    const uint8_t *failure_mode;      /* B, NULL = none */
} blx_asset_view;

/* LatticeSchedule: sizes plus caller-owned arrays. */
typedef struct blx_schedule_view {
    uint32_t node_count, bond_count;
    uint32_t block_count, color_count, boundary_bond_count;
    uint32_t *node_block_begin;    /* block_count + 1 */
    uint32_t *bond_order;          /* B: schedule index -> asset bond */
    uint32_t *bond_schedule_index; /* B: asset bond -> schedule index */
    uint32_t *bond_block_begin;    /* block_count + 1 */
    uint32_t *range_begin;         /* block_count * 2 * max(color_count, 1) */
    uint32_t *range_end;           /* same */
    uint8_t *bond_color;           /* B, schedule order */
} blx_schedule_view;

/* LatticeState: schedule order, double precision, caller-owned arrays. */
typedef struct blx_state_view {
    double origin[3];
    uint32_t node_count, bond_count, max_degree;
    double *x0, *u, *u_prev, *v;          /* 3N */
    double *inv_mass, *mass;              /* N */
    uint32_t *adj_offsets;                /* N + 1 */
    uint32_t *adj_bonds;                  /* 2B */
    uint32_t *nbr_bond, *nbr_other;       /* max_degree * N, k-major */
    double *nbr_rest;                     /* 3 * max_degree * N */
    double *nbr_weight;                   /* max_degree * N */
    uint8_t *nbr_alive;                   /* max_degree * N */
    uint32_t *bond_slot_a, *bond_slot_b;  /* B */
    uint32_t *bond_a, *bond_b;            /* B */
    double *rest_edge;                    /* 3B */
    double *rest_length, *rest_length_sq_minus, *weight, *compliance; /* B */
    double *threshold;                    /* 6B */
    uint8_t *alive, *failure_mode;        /* B */
    double *damage, *prev_tensile, *prev_compressive, *prev_shear;    /* B */
    double *plastic_extension, *plastic_strain;                       /* B */
} blx_state_view;

/* FragmentPose (fastlattice/Refracture.hpp:262-268): where a fragment that
 * re-enters the lattice sits now. Cells go to center_of_mass + orientation *
 * offset; velocities are linear_velocity + angular_velocity x r. */
typedef struct blx_fragment_pose {
    double center_of_mass[3];
    double orientation_wxyz[4];
    double linear_velocity[3];
    double angular_velocity[3];
} blx_fragment_pose;

/* The sub-lattice of one fragment (FragmentLattice, Refracture.hpp:241-260) as
 * the arrays of a blx_asset_view plus the maps back to the parent. The caller
 * owns the arrays and states their capacities; the call sets the counts. */
typedef struct blx_fragment_lattice_view {
    uint32_t node_count, bond_count;        /* set by the call */
    uint32_t node_capacity, bond_capacity;  /* of the arrays below */
    double origin[3];                       /* the fragment's centre of mass now */
    double carried_spin[3];                 /* the pose's angular velocity, carried as cell spin */
    uint32_t *parent_node;                  /* n: local -> parent node (ascending) */
    uint32_t *parent_bond;                  /* k: local -> parent bond (ascending) */
    double *reference_position;             /* 3n world: the rest shape placed on the pose */
    double *position;                       /* 3n world */
    double *velocity;                       /* 3n */
    double *mass;                           /* n */
    uint32_t *bond_a, *bond_b;              /* k, local node indices */
    double *rest_length, *compliance;       /* k */
    double *threshold;                      /* 6k */
    uint8_t *alive;                         /* k */
    double *damage;                         /* k */
    uint8_t *failure_mode;                  /* k */
    double *plastic_extension, *plastic_strain; /* k */
} blx_fragment_lattice_view;

/* FragmentMassProperties (fracture/FragmentMassProperties.hpp:11-21) of one
 * fragment, with the fragment's node count and smallest node. */
typedef struct blx_fragment_mass {
    double mass_kg;
    double center_of_mass_world_m[3];
    double linear_velocity_m_s[3];
    double inertia_world_kg_m2[9]; /* row-major, symmetric */
    double angular_momentum_kg_m2_s[3];
    double angular_velocity_rad_s[3];
    double source_kinetic_energy_j;
    double rigid_kinetic_energy_j;
    double coarsening_kinetic_loss_j;
    uint32_t node_count, first_node;
} blx_fragment_mass;

/* ---- errors and probes ------------------------------------------------- */

/* Message of the last failed call on this thread ("" after success). */
const char *blx_last_error(void);

/* cudaLatticeAvailable (FastLattice.hpp:403): a gfx950 device is present. */
int blx_available(void);
/* cudaLatticeDescription (FastLattice.hpp:404): writes a NUL-terminated text. */
int blx_description(char *out, size_t capacity);
/* cudaLatticeMultiprocessorCount (FastLattice.hpp:405): compute units. */
uint32_t blx_multiprocessor_count(void);

/* ---- host builders (lattice build) -------------------------------------- */

/* Sizes of a schedule for this asset and slab split, for the caller to
 * allocate blx_schedule_view arrays: colour count is only known after the
 * colouring, so range arrays are sized for 64 colours (the reference cap). */
int blx_schedule_build(const blx_asset_view *asset, const uint32_t *node_block_begin,
                       uint32_t block_count, blx_schedule_view *out);
/* buildLatticeSchedule (LatticeSchedule.cpp:18-104). out arrays must hold:
 * node_block_begin block_count+1, bond_order B, bond_schedule_index B,
 * bond_block_begin block_count+1, range_begin/range_end block_count*2*64,
 * bond_color B. On return color_count/boundary_bond_count are set. */

/* Max degree of the asset's adjacency, to size the state's slot arrays. */
int blx_state_max_degree(const blx_asset_view *asset, uint32_t *max_degree);
/* buildLatticeState (FastLattice.cpp:47-169). Every out array must be sized
 * from node_count, bond_count and max_degree. */
int blx_state_build(const blx_asset_view *asset, const blx_schedule_view *schedule,
                    const double origin[3], blx_state_view *out);

/* Box lattice with all 26 Moore neighbours (13 positive half offsets, bond
 * node_a < node_b, node index x + nx*(y + ny*z)), the conventions of
 * generateBoxTileLattice (matter/BoxLattice.cpp:49-171) with the Moore
 * stencil in place of the Euclidean horizon. Writes 3N reference positions
 * centred on the origin and the bond ends, returns the bond count in
 * *bond_count (call once with bond_a == NULL to size). */
int blx_box26_bonds(uint32_t nx, uint32_t ny, uint32_t nz, double cell,
                    double *reference_position, uint32_t *bond_a, uint32_t *bond_b,
                    uint64_t *bond_count);
/* boxLatticeSlabs (matter/BoxLattice.cpp:175-195): block_count+1 offsets. */
int blx_box_slabs(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t horizon,
                  uint32_t requested_blocks, uint32_t *out, uint32_t *block_count);

/* findConnectedComponents (fracture/ConnectedComponents.cpp:51-87), host:
 * label[i] = canonical fragment id (size descending, then smallest node). */
int blx_components_host(uint32_t node_count, uint32_t bond_count, const uint32_t *bond_a,
                        const uint32_t *bond_b, const uint8_t *alive, uint32_t *label,
                        uint32_t *component_count);

/* latticeStateSubstepLimit (FastLattice.cpp:30-45): 2 / the fastest nodal
 * frequency sqrt(sum of 1/compliance over a node's live bonds / mass), over
 * nodes of positive mass; 0 when no node has a stiffness (or no adjacency). */
int blx_state_substep_limit(const blx_state_view *state, double *seconds);
/* latticeStateElasticEnergy (FastLattice.cpp:199-214): sum over live bonds of
 * positive compliance of (|x_b - x_a| - L - plastic extension)^2 / (2 c). */
int blx_state_elastic_energy(const blx_state_view *state, double *joules);
/* latticeStateKineticEnergy (FastLattice.cpp:216-224): sum of m |v|^2 / 2. */
int blx_state_kinetic_energy(const blx_state_view *state, double *joules);

/* ---- device backend (LatticeBackend, FastLattice.hpp:255-369) ----------- */

typedef struct blx_backend blx_backend;

/* makeCudaLatticeBackend (FastLattice.hpp:401-402): the schedule is copied. */
int blx_backend_create(int device, int precision, const blx_schedule_view *schedule,
                       blx_backend **out);
int blx_backend_destroy(blx_backend *backend);
/* LatticeBackend::name. */
int blx_backend_name(const blx_backend *backend, char *out, size_t capacity);
/* LatticeBackend::upload (FastLattice.hpp:267-268): deep copy to the device,
 * resets status, loads, lambda and the start sample. */
int blx_backend_upload(blx_backend *backend, const blx_state_view *state,
                       const blx_settings *settings, const blx_sphere *sphere);
/* LatticeBackend::setExternalForces (FastLattice.hpp:274-277): 3N world-space
 * forces in node order held for `substeps` substeps; NULL with 0 cancels. */
int blx_backend_set_external_forces(blx_backend *backend, const double *forces_3n,
                                    uint64_t substeps);
/* LatticeBackend::run (FastLattice.hpp:360): continue from the current state;
 * the status accumulates across calls. */
int blx_backend_run(blx_backend *backend, const blx_run_control *control,
                    blx_run_status *status);
/* LatticeBackend::status (FastLattice.hpp:364). */
int blx_backend_status(const blx_backend *backend, blx_run_status *status);
/* LatticeBackend::download (FastLattice.hpp:361): the dynamic fields of the
 * state (u, u_prev, v, alive, failure_mode, damage, prev_*, plastic_*) and
 * the sphere, widened to double. Static fields are left untouched. */
int blx_backend_download(blx_backend *backend, blx_state_view *state, blx_sphere *sphere);
/* LatticeBackend::firstFailureBonds (FastLattice.hpp:368): schedule indices
 * removed in the first failure round; *count is set, at most capacity are
 * written (call with capacity 0 to size). */
int blx_backend_first_failure_bonds(const blx_backend *backend, uint32_t *out,
                                    uint64_t capacity, uint64_t *count);
/* LatticeBackend::takeFrames (FastLattice.hpp:363): frames captured since the
 * last call. Query with NULL buffers to get frame count; each frame holds
 * step, 3N float u, B alive, B float damage and the sphere. */
int blx_backend_frame_count(const blx_backend *backend, uint32_t *count);
int blx_backend_take_frame(blx_backend *backend, uint32_t index, uint64_t *step, float *u,
                           uint8_t *alive, float *damage, blx_sphere *sphere);
int blx_backend_clear_frames(blx_backend *backend);

/* Fragments on the device: integer union-find over the live bonds and the
 * canonical relabel of findConnectedComponents (ConnectedComponents.cpp:51-87),
 * label[i] for node i; *component_count pieces. */
int blx_backend_fragment_labels(blx_backend *backend, uint32_t *label,
                                uint32_t *component_count);

/* ---- fragment geometry for the rigid handoff (host) --------------------
 * The geometric half of the reference's fragment build
 * (buildFragmentRepresentations, fracture/FragmentGeometry.cpp:203-291); the
 * mass half is blx_backend_fragment_mass_properties, over the canonical labels
 * of blx_backend_fragment_labels. */

/* buildExposedVoxelSurface (fracture/FragmentGeometry.cpp:145-199) of one
 * fragment: for each of its `count` cells in the order given, each face of
 * the cell (+x, -x, +y, -y, +z, -z) whose neighbouring cell is not in the
 * fragment becomes four vertices at (position - center_of_mass) + voxel_size
 * * corner with the face's normal, and two triangles. grid and position are
 * the cells' grid coordinates and world positions (3 per cell). Writes up to
 * face_capacity faces (12 doubles of vertex_position and of vertex_normal and
 * 6 indices each) and sets *face_count to the total (capacity 0 to size). */
int blx_exposed_voxel_surface(const int32_t *grid, const double *position, uint32_t count,
                              const double center_of_mass[3], double voxel_size_m,
                              double *vertex_position, double *vertex_normal, uint32_t *indices,
                              uint64_t face_capacity, uint64_t *face_count);

/* The collision proxy of a fragment (FragmentGeometry.cpp:54-131): its
 * surface's support points in the 26 lattice directions, then an even sample
 * of the vertices up to max(maximum_points, 8) distinct points, and the first
 * cell's eight corners if fewer than four were found. vertex_position is
 * 3 * vertex_count (the surface's vertices in order); first_voxel_center the
 * first cell's centre relative to the centre of mass. Writes up to capacity
 * points (3 doubles each) and sets *count (at most max(maximum_points, 8) + 8). */
int blx_collision_points(const double *vertex_position, uint64_t vertex_count,
                         const double first_voxel_center[3], double voxel_size_m,
                         uint64_t maximum_points, double *points, uint64_t capacity,
                         uint64_t *count);

/* mergeCellsIntoBoxes (fracture/PieceGeometry.cpp:9-46): the cells (3 grid
 * coordinates each) as inclusive boxes (min x, y, z, max x, y, z), greedily:
 * the smallest free cell in (z, y, x) order grows along x, then by rows
 * along y, then by slabs along z. Writes up to capacity boxes and sets
 * *box_count. */
int blx_merge_cells_into_boxes(const int32_t *cells, uint64_t count, int32_t *boxes, uint64_t capacity,
                               uint64_t *box_count);

/* buildFragmentLattice (fastlattice/Refracture.cpp:267-393): the sub-lattice
 * of a fragment re-entering the lattice lane. Its cells are `nodes` of the
 * parent (any order, sorted and made unique here), its bonds every parent
 * bond with both ends among them (dead ones included), in parent order, with
 * their state (parent alive / damage / failure_mode, NULL for alive, 0, none).
 * The reference configuration is the parent's rest shape, turned by the
 * rotation closest to the frozen shape's (cell_offset, parent-indexed, 3 per
 * parent node) and placed on the pose; positions are the pose applied to the
 * frozen offsets. Plastic state comes from the parent's per-bond arrays (NULL
 * for zero). Call with capacities 0 to learn the counts; the sub-lattice's
 * schedule is blx_schedule_build of the result as one block. */
int blx_fragment_lattice(const blx_asset_view *parent, const uint32_t *nodes, uint32_t count,
                         const double *cell_offset, const blx_fragment_pose *pose,
                         const double *plastic_extension, const double *plastic_strain,
                         blx_fragment_lattice_view *out);

/* blx_backend_fragment_labels without allocating. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * same kernels and the same labels, but every buffer they use comes from one
 * fixed scratch array in device memory that each GPU holds until the process
 * ends (48 MiB), so the lattice may have at most 1,048,576 (2^20) nodes; a
 * larger one is refused with BLX_ERR_INVALID. label_device is the caller's device buffer of
 * node_count entries, written on the backend's stream (complete on return).
 * Calls on one GPU share that GPU's scratch array: calls that overlap on one
 * GPU, from several threads or several backends, race. */
int blx_backend_fragment_labels_noalloc(blx_backend *backend, uint32_t *label_device,
                                        uint32_t *component_count);

/* calculateFragmentMassProperties (fracture/FragmentMassProperties.cpp:7-80)
 * of every fragment of the current state, on the device, as the fragment
 * build calls it per component (FragmentGeometry.cpp:231-234): out[f] for
 * canonical fragment f of blx_backend_fragment_labels, f < capacity;
 * *component_count fragments. A node's world position is origin + x0 + u and
 * its mass the state's, both as uploaded in double; its velocity v; its
 * spin spin[3i..3i+2] (NULL: zero, as the lattice lane carries none).
 * cell_size_m is the asset's voxel size (the finite-cell inertia m h^2 / 6).
 * Sums run in a fixed order (deterministic run to run), not the reference's
 * serial order, so they agree with it to rounding. A fragment of no mass or
 * of singular inertia is BLX_ERR_RUNTIME, as the reference throws. */
int blx_backend_fragment_mass_properties(blx_backend *backend, double cell_size_m,
                                         const double *spin, blx_fragment_mass *out,
                                         uint32_t capacity, uint32_t *component_count);

/* ---- kinematic blade (no counterpart in the reference's lattice API) ------
 *
 * A thin blade moving at constant velocity through the lattice, for cutting
 * workloads. The reference's blade lives in its rigid-body world
 * (docs/cutting-model.md) and has no lattice-lane form, so this is an
 * extension of the support planes (LatticePhysics.hpp:1190-1276): at the same
 * point of the substep (after each constraint iteration, before the velocity
 * update) every node whose contact sphere overlaps the blade is moved out of
 * it by the shortest way, across a flat or ahead of the edge. The edge at
 * substep k is at edge_point + direction * speed * dt * (k + 1). */
typedef struct blx_blade {
    uint8_t enabled;
    double edge_point[3];   /* a point on the edge at time 0 */
    double flat_normal[3];  /* unit normal of the blade's flats */
    double direction[3];    /* unit direction of travel, perpendicular to flat_normal */
    double speed;           /* m/s along direction */
    double half_width;      /* half the thickness across the flats, m */
    double node_radius;     /* contact radius of a node against the blade, m */
} blx_blade;

/* Sets or clears (enabled = 0) the blade of a backend between runs. */
int blx_backend_set_blade(blx_backend *backend, const blx_blade *blade);

/* ---- batches of lattices (no counterpart in the reference's API) ---------
 *
 * A batch advances many independent lattices of any sizes together: each
 * phase of the substep is one launch over every running lattice ("slot").
 * It is for callers that run many trials at once -- the reference runs one
 * LatticeBackend per island or trial (TileImpactScene.cpp:345-351) -- and each
 * slot's results are those of the same lattice on its own backend, bit for
 * bit. The slots share the step settings (blx_settings: substep, gravity,
 * supports, contact, iterations); the state, the sphere, external loads, the
 * run control and the status are per slot. A slot keeps its device memory
 * after blx_batch_release and reuses it for the next upload that fits.
 * Frames are not captured in batches (capture_stride must be 0). */
typedef struct blx_batch blx_batch;

int blx_batch_create(int device, int precision, uint32_t capacity, blx_batch **out);
int blx_batch_destroy(blx_batch *batch);
int blx_batch_capacity(const blx_batch *batch, uint32_t *capacity);
/* makeCudaLatticeBackend + LatticeBackend::upload for one slot. The first
 * upload into an empty batch fixes the shared settings; a later upload with
 * different settings fails with BLX_ERR_INVALID. */
int blx_batch_upload(blx_batch *batch, uint32_t slot, const blx_schedule_view *schedule,
                     const blx_state_view *state, const blx_settings *settings, const blx_sphere *sphere);
/* Slot dst takes the lattice of slot src as it was uploaded (src must not
 * have run yet), with its own sphere: a device-to-device copy, for trials of
 * lattices kept resident as templates. */
int blx_batch_clone(blx_batch *batch, uint32_t dst, uint32_t src, const blx_sphere *sphere);
int blx_batch_set_external_forces(blx_batch *batch, uint32_t slot, const double *forces_3n,
                                  uint64_t substeps);
/* The slot's run control for the following advances: its exit rules and a
 * budget of max_steps substeps, after which it stops with exit reason 3. */
int blx_batch_arm(blx_batch *batch, uint32_t slot, const blx_run_control *control);
/* Advances every armed slot until it stops or `substeps` substeps have run,
 * in device batches of steps_per_launch (0: 240); abandon is polled between
 * device batches. *running is the number of slots still running. */
int blx_batch_advance(blx_batch *batch, uint64_t substeps, uint64_t steps_per_launch,
                      const volatile uint8_t *abandon, uint32_t *running);
int blx_batch_status(const blx_batch *batch, uint32_t slot, blx_run_status *status);
/* The exit reason blx_batch_status reports, for slots 0 .. count - 1 at once
 * (0 while a slot runs or holds no lattice). No counterpart in the reference
 * (whose backends run one lattice): it is for callers that scan many slots
 * after each advance, such as a server retiring the trials that stopped. */
int blx_batch_exit_reasons(const blx_batch *batch, uint32_t *reasons, uint32_t count);
int blx_batch_download(blx_batch *batch, uint32_t slot, blx_state_view *state, blx_sphere *sphere);
int blx_batch_fragment_labels(blx_batch *batch, uint32_t slot, uint32_t *label,
                              uint32_t *component_count);
/* blx_backend_fragment_mass_properties for one slot. */
int blx_batch_fragment_mass_properties(blx_batch *batch, uint32_t slot, double cell_size_m,
                                       const double *spin, blx_fragment_mass *out,
                                       uint32_t capacity, uint32_t *component_count);
/* The slot's sphere without a state download. */
int blx_batch_sphere(blx_batch *batch, uint32_t slot, blx_sphere *sphere);
/* LatticeBackend::firstFailureBonds for one slot (see blx_backend_first_failure_bonds). */
int blx_batch_first_failure_bonds(const blx_batch *batch, uint32_t slot, uint32_t *out, uint64_t capacity,
                                  uint64_t *count);
/* The blade of every slot of the batch (shared, like the step settings). */
int blx_batch_set_blade(blx_batch *batch, const blx_blade *blade);
/* The slot leaves the batch; its device memory is kept for reuse. */
int blx_batch_release(blx_batch *batch, uint32_t slot);
/* Device bytes the slot holds. */
int blx_batch_slot_bytes(const blx_batch *batch, uint32_t slot, uint64_t *bytes);

/* ---- measurement (no counterpart in the reference's API) ---------------- */

/* Per-launch HIP-event timing of every kernel (off by default: it shortens
 * device batches to 4 substeps). blx_run_status.phase_seconds, with the
 * reference's phase ids (FastLattice.cpp:22-26, RunStatus::phase_seconds at
 * FastLattice.hpp:226-228), is filled on every run: with this on, each phase
 * gets its kernels' event time; with it off, each device batch times one of
 * its substeps phase by phase and splits the batch's device time in those
 * proportions. */
int blx_backend_set_profiling(blx_backend *backend, int on);
/* Accumulated event time and launch count of kernel kind `index`; *count is
 * the number of kinds (call with index 0 to learn it). */
int blx_backend_kernel_stats(const blx_backend *backend, uint32_t index, char *name, size_t capacity,
                             double *seconds, uint64_t *launches, uint32_t *count);

/* Launch tuning of a backend or a batch. No counterpart in the reference,
 * whose CUDA lane fixed its launch shape when it was built. Every setting
 * gives the same bits; they differ in speed, which depends on the lattices'
 * sizes. Defaults: a backend 2, 1, 1024, 0, 0; a batch 1, 0, 1024, 0, 0.
 *  colours_per_launch  colour stages per sweep launch in a one-iteration run:
 *                      1, or 2 (stage pairs: each workgroup holds a tile of
 *                      nodes and the halo its two stages reach in LDS, solves
 *                      both stages there and writes the tile back once);
 *  lds_sweep           nonzero: in a one-iteration run, a lattice whose nodes
 *                      fit in one workgroup's LDS (RunStatus
 *                      shared_memory_bytes) runs its whole sweep in one
 *                      workgroup, the nodes held in LDS;
 *  lds_threads         that workgroup's size: 256, 512 or 1024;
 *  graph_substeps      substeps per captured graph launch, 0..256 (0 enqueues
 *                      every substep; node contact always enqueues);
 *  streams             a striker substep without plasticity or node contact:
 *                      0 splits across streams (striker pass 2 beside the end
 *                      sample of what it cannot reach) when a running lattice
 *                      has at least 131,072 nodes, else runs on one stream;
 *                      1 always splits; 2 never does. */
typedef struct blx_tuning {
    uint32_t colours_per_launch;
    uint32_t lds_sweep;
    uint32_t lds_threads;
    uint32_t graph_substeps;
    uint32_t streams;
} blx_tuning;

#define BLX_TUNE_MAX 32
/* What the autotuner tried and measured: host seconds per substep of each
 * candidate (the better of two rounds), and the one it kept. */
typedef struct blx_autotune_report {
    blx_tuning chosen;
    uint32_t candidates;
    blx_tuning tried[BLX_TUNE_MAX];
    double seconds_per_substep[BLX_TUNE_MAX];
} blx_autotune_report;

int blx_backend_get_tuning(const blx_backend *backend, blx_tuning *tuning);
int blx_backend_set_tuning(blx_backend *backend, const blx_tuning *tuning);
int blx_batch_get_tuning(const blx_batch *batch, blx_tuning *tuning);
int blx_batch_set_tuning(blx_batch *batch, const blx_tuning *tuning);
/* The autotuner (no counterpart in the reference). Each candidate tuning
 * (colour stages per launch, the LDS sweep and its workgroup size where a
 * lattice fits, substeps per graph launch) advances the lattice from its
 * current state by warm substeps and, again from that state, by warm + probe
 * substeps; the difference of the two host times is its cost. The fastest is
 * set, and the state (device and host, run status included) is put back as it
 * was: autotuning changes no result. The backend's lattice must be uploaded;
 * a batch tunes over its armed slots, whose budgets should cover
 * warm + probe. Takes a device copy of every uploaded lattice while it runs. */
int blx_backend_autotune(blx_backend *backend, uint64_t warm_substeps, uint64_t probe_substeps,
                         blx_autotune_report *report);
int blx_batch_autotune(blx_batch *batch, uint64_t warm_substeps, uint64_t probe_substeps,
                       blx_autotune_report *report);

/* A streaming device-to-device copy of `bytes` (read once, written once),
 * repeated; the mean seconds per copy. Used to calibrate the hardware
 * FETCH_SIZE / WRITE_SIZE counters and as the achievable-bandwidth yardstick. */
int blx_hbm_copy_probe(int device, uint64_t bytes, uint32_t repeats, double *seconds_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* BANJO_LATTICE_H */
