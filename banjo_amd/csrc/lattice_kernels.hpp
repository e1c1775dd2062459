// HIP kernels of the lattice substep for gfx950.
//
// One substep is the reference's phase order (src/fastlattice/CpuLatticeBackend.cpp:357-488)
// laid out as stream-ordered kernels over SoA arrays in HBM:
//
//   k_strain (gated)      start sample, nodes      LatticePhysics.hpp:792-907
//   k_bond_start (gated)  start sample, bonds      LatticePhysics.hpp:1118-1126
//   k_predict             external kick + gravity kick, predict, classify
//                                                  ExternalLoads.hpp:79-148, LatticePhysics.hpp:1706-1719
//   k_pass (1 WG)         ordered candidate lists and the serial striker pass
//                                                  CpuLatticeBackend.cpp:373-387, LatticePhysics.hpp:1395-1428
//   k_solve x stages      XPBD bond projection, one launch per colour stage
//                                                  LatticePhysics.hpp:1132-1168
//   k_project             support projection      LatticePhysics.hpp:1204-1261
//   k_velocity            velocity from positions LatticePhysics.hpp:1726-1730
//   k_damp x stages       radial bond damping     LatticePhysics.hpp:1171-1184
//   k_strain              end-of-substep nodal strain
//   k_bond_end            plastic return, sample, damage, break, tile event bits
//                                                  LatticePhysics.hpp:1068-1114
//   k_finish (1 WG)       bond-order folds of removed energy and max damage,
//                         sphere drift and supports, exit rule
//                                                  CpuLatticeBackend.cpp:441-488, LatticePhysics.hpp:1766-1825
//
// No floating-point atomics anywhere: reductions that the reference does in
// bond order are done in bond order (a one-wave fold over the ordered event
// list), maxima of non-negative floats use integer atomicMax on their bits,
// and the remaining double sums (plastic work, load ledgers) use a fixed-order
// tree, so every run is bit-reproducible.
#pragma once

#include "banjo_lattice.h"
#include "lattice_device.hpp"

namespace blx {

constexpr int kWG = 256;
constexpr u32 kEndSpan = 1;  // bond tiles per end-sample workgroup (2 and 4 measured no faster)
constexpr u32 kCap = 1024;        // striker candidates per block (LatticePhysics.hpp:105)
constexpr int kFinishThreads = 1024;
// Striker pass 2 stages its candidates with four waves and runs the chain in
// one: launched as the end sample's wide launches become ready on the other
// stream, it fits beside their workgroups on a partly busy CU, where sixteen
// waves would wait for a CU to drain.
constexpr int kPass2Threads = 256;

// Device-resident run state. Everything that changes from one substep to the
// next lives here, so kernel arguments are constant and a batch of substeps
// can be enqueued (or captured) without the host in the loop.
struct DevCtl {
    u64 step;            // RunStatus::total_steps
    u32 stop;            // nonzero: the remaining substeps of the batch are no-ops
    u32 error;           // bit 0: external kick overflowed backend precision
    u32 dirty_start[2];  // by substep parity: recompute the start sample of the next substep
    u32 loads_active;
    u64 load_remaining;
    u64 first_failure_step, last_failure_step;
    u32 failure_rounds, broken_bonds;
    double removed_energy_j;
    u64 last_energy_gain_step;
    double max_damage;
    u64 last_damage_gain_step;
    u64 impulse_contacts;
    double impulse[3];
    double dissipated, max_penetration, max_correction, max_shift;
    u64 candidate_overflow, ball_support_events;
    u64 ledger_steps;
    double ledger_elapsed, ledger_req[3], ledger_imp[3], ledger_areq[3], ledger_aimp[3], ledger_work;
    u32 peak_bits[3];
    u32 pstretch_bits;
    u32 rank_deficient;
    u32 exit_reason;
    u32 dirty_count[2];  // by substep parity: entries in that half of dirty_list
    u32 rank_slot;       // parity slot k_rinv counts into (set by k_bond_end)
    u32 rank_pending[2]; // rank-deficient recomputations not yet booked
    u32 rinv_stamp;      // break round the dirty list belongs to
    double plastic_work;
    u64 first_failure_count;
    u32 affected_count;  // entries in affected_list
    u32 pad0;
    u64 budget;          // substeps left in the current run; the run stops at 0
    // RunControl of the current run() call.
    u64 quiet, min_steps, no_failure, energy_flat, calm;
    double flat_fraction, calm_margin, removable;
    // Node-node contact: the rebuild flag (upload and every failure round set
    // it, the staleness test sets it, a rebuild clears it), the listed pair
    // count, a cell-range error flag and the ledger (NodeContactAccumulators).
    u32 nc_rebuild, nc_pairs, nc_error, pad1;
    u64 nc_contacts, nc_tests, nc_rebuilds, nc_listed, nc_overflow;
    double nc_dissipated, nc_overlap, nc_correction, nc_momentum[3];
};

// Workgroup w runs on XCD w % 8, and every XCD has its own L2. This maps the
// workgroups of one launch so that each XCD walks a contiguous eighth of the
// tiles, which keeps the node data that neighbouring tiles share in one L2.
// Launch xcdGrid(tiles) workgroups; a result >= tiles is an idle workgroup.
constexpr u32 kXcds = 8;
constexpr u32 kSolveFirst = 1U;      // k_solve: first constraint iteration, the multiplier starts at 0
constexpr u32 kSolveKeepLambda = 2U; // k_solve: store the multiplier (runs of more than one iteration)
__host__ __device__ __forceinline__ u32 xcdGrid(u32 tiles) { return (tiles + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ u32 xcdTile(u32 w, u32 tiles) {
    const u32 per = (tiles + kXcds - 1) / kXcds;
    return (w % kXcds) * per + w / kXcds;
}

template <class R> struct DevLattice {
    u32 N, B, D, blocks, node_tiles, bond_tiles, group_words;
    u32 npal;         // edge palette entries
    u32 tensile_only; // every bond's compression and shear ramps are off: thr2 replaces thr
    R4<R> *u4;        // displacement, w = inv_mass
    R4<R> *u4b;       // the paired sweep's second node buffer
    // The paired sweep's node lists (per workgroup its owned nodes, then the
    // halo) and stage bond lists.
    const u32 *pnodes;
    const PairBond<R> *pbonds;
    const uint2 *stage_ranges;  // the colour stages as bond ranges [x, y), in sweep order
    const u32 *stage_end;       // stage s is stage_ranges[stage_end[s - 1] .. stage_end[s])
    u32 stage_count;
    u32 lds_fast;               // every stage one range and the edge palette within kLdsPal
    u32 stage_max;              // bonds in the largest stage
    R4<R> *up4;       // displacement at substep start
    R4<R> *v4;        // velocity
    const R4<R> *x04; // reference position, w = mass
    R *rinv;          // 9 planes of N: cached inverse rest covariance
    R *unm;           // 3 planes of N: unmeasured normal
    R4<R> *sa, *sb;   // strain (xx, yy, zz, xy), (xz, yz, valid, 0)
    u8 *valid;
    R *approach;      // 2 planes of N
    u8 *engaged;      // 2 planes of N
    u8 *cand;         // N
    // Neighbour slots, D planes of N (k-major): slot palette index | kDeadSlot, kNone where absent.
    u32 *nbr_key;
    const R4<R> *spal;   // slot palette: rest vector to the neighbour, weight
    const int *sdelta;   // slot palette: neighbour index minus node index
    const u32 *nbr_bond; // D planes of N: node-major index of the slot's bond (start sample)
    const EdgePal<R> *epal;
    // Bonds in schedule order (the colour sweep).
    const BondRec<R> *rec;  // c holds compliance / dt^2
    const R *cmpl;          // compliance per bond, schedule order
    u8 *alive;
    R *lambda, *pext;
    // Bonds in node-major order (the end sample).
    const BondNode *bnd;
    u8 *alive_n;      // alive again, in node-major order (written together with alive)
    const Thresholds<R> *thr;
    const Tensile<R> *thr2;
    R2<R> *pv;        // prev tensile sample; damage
    // Prev compressive and shear samples. With compression and shear ramps
    // on, read and written every substep. On a tensile-only lattice they feed
    // no ramp, so they are written only at upload and at a break and rebuilt
    // for the live bonds from the end strains at download.
    R2<R> *pyz;
    u8 *mode;
    R *pstr;
    const uint2 *slots;
    // Substep events by schedule index: a bit per bond, a summary bit per 64 words.
    u64 *ev_bits;     // the bond broke, or its damage may raise the running maximum
    u64 *brk_bits;    // the bond broke
    u64 *grp_bits;    // words of ev_bits that hold a set bit
    float *ev_d;      // the event bond's damage as the float the fold compares
    R *ev_e;          // stored energy removed by a break
    u32 *flag_list;   // scratch: flagged groups in order
    u32 *tile_cand;   // candidates per node tile
    // Nodes a break made dirty, each once per substep (unordered), in two halves
    // of N by substep parity: the end sample of substep t fills half t & 1, and
    // the start sample of substep t + 1 reads it while substep t + 1's end
    // sample fills the other half.
    u32 *dirty_list;
    u32 *dirty_stamp; // substep stamp of a node's entry in dirty_list
    u32 *rank_stamp;  // break round a node was last counted rank-deficient in
    u32 *affected;    // substep stamp: striker pass 2 can move the node or a neighbour
    u32 *affected_list; // the affected nodes of this substep (unordered)
    double *tile_plastic; // plastic work per end-sample tile
    double *load_part;    // ledger partial sums per node tile (13 each)
    R4<R> *fext;      // applied external force in backend precision, w = 1 where loaded
    const double *freq; // requested external force (3N double)
    u32 *cand_all;    // ordered candidates, all blocks
    u32 *cand_tile;   // per node tile, its candidates in node order
    u32 *cand_lo, *cand_cnt;
    const u32 *nbb;   // node block begin
    u32 *first_fail;  // schedule indices broken in the first failure round
    // Node-node contact, allocated when it is on: u at the last rebuild, the
    // nodes sorted by cell, a cell table of runs in that order, each node's
    // pairs (kPairsPerNode slots) and the pair list in order (flat_*).
    R4<R> *nc_build;
    u64 *nc_key;
    u32 *nc_bucket, *nc_bucket_sorted, *nc_node, *nc_node_sorted;
    uint2 *nc_table_val;
    u32 nc_table_mask;
    u32 *pair_fill, *pair_other, *pair_bond, *pair_offset, *flat_a, *flat_b, *flat_bond;
    // The narrow phase's schedule: per node, how many pairs name it as partner
    // and how many of its pairs are done; per pair, its rank among the pairs
    // with the same partner; the sort by partner; and each pair's ledger terms.
    u32 *nc_partners, *nc_done, *nc_rank_b, *nc_sort_key, *nc_sort_in, *nc_sort_out;
    R4<R> *nc_terms;
    u8 *nc_mover;
    DevCtl *ctl;
    Sphere<R> *sphere;
    double origin[3];
};

// A launch over a batch of lattices. Every workgroup finds its lattice and
// its share of that lattice's work in the launch's map, so lattices of any
// sizes advance together in one launch per phase; a single lattice is a batch
// of one. The meaning of a, b depends on the kernel: a tile index, a slice of
// a list (slice, slice count), or a bond range (begin, end).
struct WgMap {
    u32 lat, a, b, pad;
};
template <class R> struct Batch {
    const DevLattice<R> *L;
    const WgMap *map;
    u32 count;  // map entries; a launch has xcdGrid(count) workgroups (tile maps) or count
};
// The map entry of this workgroup for tile launches (XCD-aware), or nullptr.
template <class R> __device__ __forceinline__ const WgMap *tileEntry(const Batch<R> &B) {
    const u32 w = xcdTile(blockIdx.x, B.count);
    return w < B.count ? &B.map[w] : nullptr;
}

// ---------------------------------------------------------------------------
// Nodes.

template <class R, bool DIRECT> __device__ __forceinline__ V3<R> nodePos(const DevLattice<R> &L, u32 i, const R4<R> &u) {
    if (DIRECT) return xyz(ldg(&L.x04[i])) + xyz(u);
    return xyz(u);
}

// LatticePhysics.hpp:792-907 (plate bending is refused at upload).
// Dead and absent neighbour slots contribute an exact zero to every sum in the
// reference, so they are skipped; the result is the same bits.
// The cached inverse rest covariance of a node whose live neighbour set
// changed (the dirty branch of LatticePhysics.hpp:796-832).
template <class R> __device__ void rinvNode(const DevLattice<R> &L, u32 i, u32 *rank_counter, u32 stamp) {
    const u32 N = L.N, D = L.D;
    R cov[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    u32 live = 0;
    // Slots in chunks of 8: the chunk's keys, then its palette entries, are
    // loaded before the sums, which stay in slot order.
    constexpr u32 kChunk = 8;
    for (u32 k0 = 0; k0 < D; k0 += kChunk) {
        u32 key[kChunk];
        for (u32 q = 0; q < kChunk; ++q) key[q] = k0 + q < D ? L.nbr_key[size_t(k0 + q) * N + i] : kNone;
        R4<R> rw[kChunk];
        for (u32 q = 0; q < kChunk; ++q) rw[q] = key[q] & kDeadSlot ? R4<R>{R(0), R(0), R(0), R(0)} : L.spal[key[q]];
        for (u32 q = 0; q < kChunk; ++q) {
            if (key[q] & kDeadSlot) continue;  // dead or absent
            const R w = rw[q].w;
            if (w <= R(0)) continue;
            const R rr[3] = {rw[q].x, rw[q].y, rw[q].z};
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) cov[r * 3 + c] += w * rr[r] * rr[c];
            ++live;
        }
        if (key[kChunk - 1] == kNone) break;
    }
    u8 valid = 0;
    if (live >= 1) {
        R inv[9], un[3];
        const int rank = pinv3(cov, inv, un);
        if (rank >= 2) {
            for (int k = 0; k < 9; ++k) L.rinv[size_t(k) * N + i] = inv[k];
            for (int k = 0; k < 3; ++k) L.unm[size_t(k) * N + i] = un[k];
            valid = 1;
        }
        // Counted once per node and break round, however often the node is recomputed.
        if (rank < 3 && atomicExch(&L.rank_stamp[i], stamp) != stamp) atomicAdd(rank_counter, 1U);
    }
    // The strain pass clears an invalid node's sample flag itself (sb.z), so
    // the strains of the last end sample stay as it read them until the next
    // strain pass.
    L.valid[i] = valid;
}

// The nodal strain from the cached inverse (LatticePhysics.hpp:833-906).
// PREV takes the positions from up4, the displacement at substep start: the
// start sample runs after k_predict, which copies u to up4 before it moves u.
template <class R, bool DIRECT, bool PREV> __device__ void strainNode(const DevLattice<R> &L, u32 i) {
    const u32 N = L.N, D = L.D;
#ifndef BLX_STRAIN_CHUNK
#define BLX_STRAIN_CHUNK 4
#endif
#ifndef BLX_STRAIN_PREFETCH
#define BLX_STRAIN_PREFETCH 1
#endif
    // Slots in chunks: every load of a chunk is issued before the sums, which
    // stay in slot order, and the next chunk's keys are in flight meanwhile.
    // Four slots a chunk keep the kernel at 54 VGPRs, eight waves per SIMD;
    // eight slots took 91 and five waves, and ran slower.
    constexpr u32 kChunk = BLX_STRAIN_CHUNK;
    // The node's flag, own displacement and first keys go out together.
    const u8 valid = ldg(&L.valid[i]);
    const R4<R> *pos = PREV ? L.up4 : L.u4;
    const R4<R> ui = ldg(&pos[i]);
    u32 next[kChunk];
    for (u32 q = 0; q < kChunk; ++q) next[q] = q < D ? ldStream(&L.nbr_key[size_t(q) * N + i]) : kNone;
    if (!valid) {  // no strain sample at this node (LatticePhysics.hpp:833-836)
        stg(&L.sb[i], R4<R>{R(0), R(0), R(0), R(0)});
        return;
    }
    R acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const V3<R> xi = nodePos<R, DIRECT>(L, i, ui);
    for (u32 k0 = 0; k0 < D; k0 += kChunk) {
        u32 key[kChunk];
        for (u32 q = 0; q < kChunk; ++q) key[q] = next[q];
        if (BLX_STRAIN_PREFETCH) {
            if (key[kChunk - 1] != kNone)
                for (u32 q = 0; q < kChunk; ++q)
                    next[q] = k0 + kChunk + q < D ? ldStream(&L.nbr_key[size_t(k0 + kChunk + q) * N + i]) : kNone;
        }
        R4<R> rw[kChunk];
        u32 other[kChunk];
        for (u32 q = 0; q < kChunk; ++q) {
            const bool live = !(key[q] & kDeadSlot);
            rw[q] = live ? ldg(&L.spal[key[q]]) : R4<R>{R(0), R(0), R(0), R(0)};
            other[q] = live ? u32(int(i) + ldg(&L.sdelta[key[q]])) : i;
        }
        R4<R> xo4[kChunk];
        for (u32 q = 0; q < kChunk; ++q)
            xo4[q] = !(key[q] & kDeadSlot) ? ldg(&pos[other[q]]) : R4<R>{R(0), R(0), R(0), R(0)};
        for (u32 q = 0; q < kChunk; ++q) {
            if (key[q] & kDeadSlot) continue;
            const R w = rw[q].w;
            const V3<R> rest = xyz(rw[q]);
            const V3<R> xo = nodePos<R, DIRECT>(L, other[q], xo4[q]);
            const V3<R> cur = DIRECT ? (xo - xi) - rest : xo - xi;
            const R cc[3] = {cur.x, cur.y, cur.z}, rr[3] = {rest.x, rest.y, rest.z};
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) acc[r * 3 + c] += w * cc[r] * rr[c];
        }
        if (key[kChunk - 1] == kNone) break;
        if (!BLX_STRAIN_PREFETCH)
            for (u32 q = 0; q < kChunk; ++q)
                next[q] = k0 + kChunk + q < D ? ldStream(&L.nbr_key[size_t(k0 + kChunk + q) * N + i]) : kNone;
    }
    R inv[9], un[3], e[6];
    for (int k = 0; k < 9; ++k) inv[k] = ldg(&L.rinv[size_t(k) * N + i]);
    for (int k = 0; k < 3; ++k) un[k] = ldg(&L.unm[size_t(k) * N + i]);
    greenStrain(acc, inv, un, e);
    stg(&L.sa[i], R4<R>{e[0], e[1], e[2], e[3]});
    stg(&L.sb[i], R4<R>{e[4], e[5], R(1), R(0)});
}

// Which nodes a k_strain launch covers.
enum StrainCover : int {
    kStrainAll = 0,      // every node
    kStrainDirty = 1,    // the start sample after a failure round
    kStrainUnaffected = 2, // the end sample of nodes striker pass 2 cannot reach
    kStrainAffected = 3  // the end sample of the rest
};

// kStrainDirty is the start sample after a failure round (CpuLatticeBackend.cpp:363-370):
// only nodes a break made dirty can have a different strain from the one the
// previous substep ended with (their neighbours and positions are unchanged),
// so only they are recomputed, and they stay marked for k_bond_start.
// The end sample splits in two when a striker is present: striker pass 2 moves
// only its candidates, so the strain of a node with no candidate among itself
// and its neighbours can be taken while pass 2 runs.
template <class R, bool DIRECT, int COVER>
__global__ __launch_bounds__(kWG) void k_strain(Batch<R> B) {
    if (COVER == kStrainDirty || COVER == kStrainAffected) {
        const WgMap m = B.map[blockIdx.x];
        const DevLattice<R> L = B.L[m.lat];
        if (L.ctl->stop) return;
        const u32 half = u32(L.ctl->step & 1U) ^ 1U;  // the previous substep's half
        if (COVER == kStrainDirty && !L.ctl->dirty_start[half]) return;
        const u32 count = COVER == kStrainDirty ? L.ctl->dirty_count[half] : L.ctl->affected_count;
        const u32 *list = COVER == kStrainDirty ? L.dirty_list + size_t(half) * L.N : L.affected_list;
        for (u32 k = m.a * kWG + threadIdx.x; k < count; k += m.b * kWG)
            strainNode<R, DIRECT, COVER == kStrainDirty>(L, list[k]);
        return;
    }
    const WgMap *m = tileEntry(B);
    if (!m) return;
    const DevLattice<R> L = B.L[m->lat];
    if (L.ctl->stop) return;
    const u32 i = m->a * kWG + threadIdx.x;
    if (i >= L.N) return;
    if (COVER == kStrainUnaffected && L.affected[i] == u32(L.ctl->step) + 1U) return;
    strainNode<R, DIRECT, false>(L, i);
}

// The nodes striker pass 2 can reach this substep: its candidates (the
// staged lists of pass 1) and their neighbours, each listed once. Runs on a
// side stream while the colour sweep runs.
// Map entry `entry` of B: (lattice, slice a of b slices of the marking work).
// Every lane of the wave calls it; each wave appends its claims with one
// atomic, since thousands of appends to one counter would queue at it.
template <class R> __device__ __forceinline__ void markAffected(const Batch<R> &B, u32 entry) {
    const WgMap m = B.map[entry];
    const DevLattice<R> L = B.L[m.lat];
    DevCtl *ctl = L.ctl;
    if (ctl->stop) return;
    const u32 stamp = u32(ctl->step) + 1U;
    const u32 per = L.D + 1, lane = threadIdx.x & 63U;
    // Block by block: a block's work (its listed candidates, each with its
    // slots) is below kCap * per entries, a 32-bit index.
    for (u32 b = 0; b < L.blocks; ++b) {
        const u32 total = L.cand_cnt[b] * per;
        const u32 *cands = L.cand_all + size_t(b) * kCap;
        for (u32 base = m.a * kWG; base < total; base += m.b * kWG) {
            const u32 q = base + threadIdx.x;
            u32 n = kNone;
            if (q < total) {
                const u32 c = cands[q / per], k = q % per;
                if (k == 0) {
                    n = c;
                } else {
                    const u32 key = L.nbr_key[size_t(k - 1) * L.N + c];
                    if (key != kNone) n = u32(int(c) + L.sdelta[key & ~kDeadSlot]);
                }
            }
            const bool claim = n != kNone && atomicExch(&L.affected[n], stamp) != stamp;
            const u64 mask = __ballot(claim);
            if (mask) {
                const u32 leader = u32(__ffsll((long long)mask) - 1);
                u32 at = 0;
                if (lane == leader) at = atomicAdd(&ctl->affected_count, u32(__popcll(mask)));
                at = u32(__builtin_amdgcn_readlane(int(at), int(leader)));
                if (claim) L.affected_list[at + u32(__popcll(mask & ((1ULL << lane) - 1ULL)))] = n;
            }
        }
    }
}
template <class R> __global__ __launch_bounds__(kWG) void k_mark_affected(Batch<R> B) { markAffected(B, blockIdx.x); }

// Inverse rest covariances of the nodes the last breaks made dirty. It needs
// only the live set k_bond_end left, so it runs on a side stream while
// k_finish folds the same substep. The reference recomputes these at the start
// of the next substep and counts rank-deficient ones there, so the count goes
// to a per-parity slot that the next substep's k_finish books.
// It ignores the stop flag, which k_finish may raise while it runs: the
// recomputation is idempotent and the count is stamped per node.
template <class R> __global__ __launch_bounds__(kWG) void k_rinv(Batch<R> B) {
    const WgMap m = B.map[blockIdx.x];
    const DevLattice<R> L = B.L[m.lat];
    // rank_slot is the parity of the substep whose end sample filled the list.
    const u32 half = L.ctl->rank_slot;
    const u32 count = L.ctl->dirty_count[half];
    const u32 *list = L.dirty_list + size_t(half) * L.N;
    u32 *counter = &L.ctl->rank_pending[half];
    const u32 stamp = L.ctl->rinv_stamp;
    for (u32 k = m.a * kWG + threadIdx.x; k < count; k += m.b * kWG) rinvNode<R>(L, list[k], counter, stamp);
}

// ---------------------------------------------------------------------------
// Bonds: the strain sample (LatticePhysics.hpp:924-959).

template <class R, bool DIRECT>
__device__ __forceinline__ V3<R> bondVec(const DevLattice<R> &L, u32 a, u32 b, const R4<R> &e, const R4<R> &ua,
                                         const R4<R> &ub) {
    if (DIRECT) return (xyz(L.x04[b]) + xyz(ub)) - (xyz(L.x04[a]) + xyz(ua));
    return xyz(e) + (xyz(ub) - xyz(ua));
}

// The strain sample of one bond (LatticePhysics.hpp:924-959). The unit rest
// direction comes from the edge palette, where it was computed as
// e * (1 / |e|) by the same expression.
template <class R, bool DIRECT>
__device__ __forceinline__ void sampleBondWith(const DevLattice<R> &L, u32 a, u32 b, u32 p, const EdgePal<R> &g,
                                               const R4<R> &ua, const R4<R> &ub, const R4<R> (&st)[4], R &t, R &c,
                                               R &s, R shear_floor2 = R(0)) {
    const V3<R> delta = bondVec<R, DIRECT>(L, a, b, g.e, ua, ub);
    const R len = length(delta);
    const R rest = g.e.w;
    R stretch;
    if (DIRECT) {
        stretch = len / rest - R(1);
    } else {
        const V3<R> du = xyz(ub) - xyz(ua);
        const R l2 = g.d.x + R(2) * dot(xyz(g.e), du) + dot(du, du);
        stretch = l2 / ((len + rest) * rest);
    }
    t = stretch;
    c = -stretch;
    s = R(0);
    if (!(p & kNoDir)) {
        const V3<R> dir{g.d.y, g.d.z, g.d.w};
        // st holds (sa, sb) of end a, then of end b; sb.z is the valid flag.
        // resolve (LatticePhysics.hpp:911-919) with the shear's square root
        // taken apart: the normals first, then the roots of the shears that
        // are not below shear_floor2 (0: all of them).
        R sq[2] = {R(0), R(0)};
        bool on[2];
        for (int q = 0; q < 2; ++q) {
            on[q] = st[2 * q + 1].z != R(0);
            if (!on[q]) continue;
            const R4<R> &ea = st[2 * q], &eb = st[2 * q + 1];
            const V3<R> tr{ea.x * dir.x + ea.w * dir.y + eb.x * dir.z, ea.w * dir.x + ea.y * dir.y + eb.y * dir.z,
                           eb.x * dir.x + eb.y * dir.y + ea.z * dir.z};
            const R nn = dot(dir, tr);
            sq[q] = length2(tr - nn * dir);
            t = maxR(t, nn);
            c = maxR(c, -nn);
        }
        const bool root[2] = {on[0] && !(sq[0] < shear_floor2), on[1] && !(sq[1] < shear_floor2)};
        if (root[0] || root[1]) {
            // A real branch: a wave none of whose shears can raise the running
            // maximum skips both square roots (an empty volatile asm keeps the
            // compiler from evaluating them anyway and selecting).
            asm volatile("");
            for (int q = 0; q < 2; ++q)
                if (on[q]) s = maxR(s, root[q] ? sqrtR(sq[q]) : R(0));
        } else {
            for (int q = 0; q < 2; ++q)
                if (on[q]) s = maxR(s, R(0));
        }
    }
}

// The nodal strains a bond sample reads: (sa, sb) of end a, then of end b.
template <class R> __device__ __forceinline__ void loadEndStrains(const DevLattice<R> &L, u32 a, u32 b, u32 p,
                                                                  R4<R> (&st)[4]) {
    if (p & kNoDir) return;
    st[0] = ldg(&L.sa[a]);
    st[1] = ldg(&L.sb[a]);
    st[2] = ldg(&L.sa[b]);
    st[3] = ldg(&L.sb[b]);
}

template <class R, bool DIRECT>
__device__ void sampleBond(const DevLattice<R> &L, u32 a, u32 b, u32 p, const EdgePal<R> &g, const R4<R> &ua,
                           const R4<R> &ub, R &t, R &c, R &s) {
    R4<R> st[4] = {};
    loadEndStrains(L, a, b, p, st);
    sampleBondWith<R, DIRECT>(L, a, b, p, g, ua, ub, st, t, c, s);
}

// The edge palette's derived terms: |e| and the unit direction
// (LatticePhysics.hpp:944-946), one entry per thread at upload.
template <class R> __global__ void k_edge_palette(EdgePal<R> *pal, u32 *no_dir, u32 n) {
    for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        EdgePal<R> g = pal[k];
        const V3<R> e = xyz(g.e);
        const R rl = length(e);
        const bool has = rl > R(1.0e-9);
        const V3<R> dir = has ? e * (R(1) / rl) : V3<R>{R(0), R(0), R(0)};
        g.d = R4<R>{g.d.x, dir.x, dir.y, dir.z};
        pal[k] = g;
        no_dir[k] = has ? 0U : 1U;
    }
}

template <bool TENSILE, bool PLASTIC>
__device__ __forceinline__ void publishMaxima(DevCtl *ctl, float mt, float mc, float ms, float mp,
                                              u32 (*s_max)[kWG / 64]);

// The start sample of the live bonds at dirty nodes (LatticePhysics.hpp:1118-1126).
// A bond with no dirty end samples exactly what its end sample stored, so
// only the bonds in dirty nodes' slots are resampled; a bond with both ends
// dirty is written twice with the same bits. On a tensile-only lattice the
// compressive and shear samples are not stored; their only other use, the
// running maxima of the end sample (which take the larger of a bond's
// previous and current sample), gets them here instead. A bond that was not
// resampled holds the last end sample's values, which that sample already
// counted; upload marks every node dirty, so the first substep resamples all.
template <class R, bool DIRECT> __global__ __launch_bounds__(kWG) void k_bond_start(Batch<R> B) {
    const WgMap m = B.map[blockIdx.x];
    const DevLattice<R> L = B.L[m.lat];
    const u32 half = u32(L.ctl->step & 1U) ^ 1U;  // the previous substep's half
    if (L.ctl->stop || !L.ctl->dirty_start[half]) return;
    __shared__ u32 s_max[4][kWG / 64];
    float mc = 0.f, ms = 0.f;
    // Node slots D * N are below 2^32 (refused at upload otherwise).
    const u32 pairs = L.ctl->dirty_count[half] * L.D;
    const u32 *list = L.dirty_list + size_t(half) * L.N;
    for (u32 q = m.a * kWG + threadIdx.x; q < pairs; q += m.b * kWG) {
        const u32 i = list[q / L.D], k = q % L.D;
        const u32 o = L.nbr_bond[size_t(k) * L.N + i];
        if (o == kNone || !L.alive_n[o]) continue;
        const BondNode r = L.bnd[o];
        R t, c, s;
        // Positions at substep start (up4): this runs after k_predict.
        sampleBond<R, DIRECT>(L, r.a, r.b, r.p, L.epal[r.p & ~kNoDir], L.up4[r.a], L.up4[r.b], t, c, s);
        L.pv[o].x = maxR(R(0), t);
        const R pc = maxR(R(0), c), ps = maxR(R(0), s);
        if (L.tensile_only) {
            mc = fmaxf(mc, float(pc));
            ms = fmaxf(ms, float(ps));
        } else {
            L.pyz[o] = R2<R>{pc, ps};
        }
    }
    if (L.tensile_only) publishMaxima<false, false>(L.ctl, 0.f, mc, ms, 0.f, s_max);
}

// The previous compressive and shear samples of the live bonds of a
// tensile-only lattice, rebuilt from the last end sample's inputs (the end
// positions and end strains, which stay in place until the next substep), for
// a download: the values that end sample would have stored.
template <class R, bool DIRECT> __global__ __launch_bounds__(kWG) void k_prev_sample(DevLattice<R> L) {
    for (u32 o = blockIdx.x * kWG + threadIdx.x; o < L.B; o += gridDim.x * kWG) {
        if (!L.alive_n[o]) continue;
        const BondNode r = L.bnd[o];
        R t, c, s;
        sampleBond<R, DIRECT>(L, r.a, r.b, r.p, L.epal[r.p & ~kNoDir], L.u4[r.a], L.u4[r.b], t, c, s);
        L.pyz[o] = R2<R>{maxR(R(0), c), maxR(R(0), s)};
    }
}

// ---------------------------------------------------------------------------
// Kick, predict, classify.


// Sum of a double over the 64 lanes of a wave in a fixed order, in lane 63:
// DPP row shifts by 1, 2, 4 and 8 leave each row's sum in its last lane, row
// broadcasts add row 0 into row 1 and row 2 into row 3, then lane 31 into the
// upper half. No LDS and no barrier; every lane of the wave must be active.
template <int CTRL, int ROWS> __device__ __forceinline__ double dppShift(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ double waveSumLast(double v) {
    v += dppShift<0x111, 0xf>(v);  // row_shr:1
    v += dppShift<0x112, 0xf>(v);  // row_shr:2
    v += dppShift<0x114, 0xf>(v);  // row_shr:4
    v += dppShift<0x118, 0xf>(v);  // row_shr:8
    v += dppShift<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dppShift<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Sums of 13 per-thread values over a workgroup of WAVES waves, in a fixed
// order: each wave folds its lanes (waveSumLast), then thread k < 13 adds the
// waves' sums in wave order into sh[WAVES][k], which every thread can read
// after the call. Every thread of the workgroup calls it.
template <u32 WAVES> __device__ void groupSum13(double (*sh)[13], double v[13]) {
    const u32 lane = threadIdx.x & 63U, wave = threadIdx.x >> 6;
    for (int k = 0; k < 13; ++k) v[k] = waveSumLast(v[k]);
    if (lane == 63)
        for (int k = 0; k < 13; ++k) sh[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 13) {
        double t = sh[0][threadIdx.x];
        for (u32 w = 1; w < WAVES; ++w) t += sh[w][threadIdx.x];
        sh[WAVES][threadIdx.x] = t;
    }
    __syncthreads();
}

template <class R, bool SPHERE, bool LOADS, bool DIRECT>
__global__ __launch_bounds__(kWG) void k_predict(Batch<R> B, Step<R> S) {
    const WgMap *m = tileEntry(B);
    if (!m) return;
    const DevLattice<R> L = B.L[m->lat];
    DevCtl *ctl = L.ctl;
    if (ctl->stop) return;
    const u32 tile = m->a;
    if (tile == 0 && threadIdx.x == 0) {
        const u32 half = u32(ctl->step & 1U);  // this substep's half, which its end sample fills
        ctl->dirty_start[half] = 0;
        ctl->dirty_count[half] = 0;
        ctl->affected_count = 0;
    }
    const bool loads = LOADS && ctl->load_remaining > 0;
    const u32 i = tile * kWG + threadIdx.x;
    bool cand = false;
    double led[13];
    if (LOADS)
        for (int k = 0; k < 13; ++k) led[k] = 0.0;
    if (i < L.N) {
        const R4<R> u = L.u4[i];
        V3<R> vel = xyz(L.v4[i]);
        if (LOADS && loads) {
            const R4<R> f = L.fext[i];
            if (f.w != R(0)) {
                const V3<R> before = vel;
                const V3<R> after = before + (S.dt * u.w) * xyz(f);
                if (!finiteR(after.x) || !finiteR(after.y) || !finiteR(after.z)) atomicOr(&ctl->error, 1U);
                vel = after;
                // ExternalLoads.hpp:111-131: the ledger in double.
                const double m = double(L.x04[i].w);
                const double ox = double(before.x), oy = double(before.y), oz = double(before.z);
                const double nx = double(after.x), ny = double(after.y), nz = double(after.z);
                const double ix = m * (nx - ox), iy = m * (ny - oy), iz = m * (nz - oz);
                const V3<R> p = xyz(L.x04[i]) + xyz(u);
                const double wx = L.origin[0] + double(p.x), wy = L.origin[1] + double(p.y), wz = L.origin[2] + double(p.z);
                const double dt = double(S.dt);
                const double rx = dt * L.freq[3 * size_t(i)], ry = dt * L.freq[3 * size_t(i) + 1], rz = dt * L.freq[3 * size_t(i) + 2];
                led[0] = rx; led[1] = ry; led[2] = rz;
                led[3] = ix; led[4] = iy; led[5] = iz;
                led[6] = wy * rz - wz * ry; led[7] = wz * rx - wx * rz; led[8] = wx * ry - wy * rx;
                led[9] = wy * iz - wz * iy; led[10] = wz * ix - wx * iz; led[11] = wx * iy - wy * ix;
                led[12] = ix * ((ox + nx) * 0.5) + iy * ((oy + ny) * 0.5) + iz * ((oz + nz) * 0.5);
            }
        }
        vel = vel + S.dt * S.gravity;
        L.v4[i] = with(vel, R(0));
        L.up4[i] = u;
        L.engaged[i] = 0;
        L.engaged[size_t(L.N) + i] = 0;
        if (SPHERE) {
            const Sphere<R> sp = *L.sphere;
            const V3<R> kv = sp.velocity + S.dt * S.gravity;
            const R4<R> x0 = L.x04[i];
            if (x0.w > R(0)) {
                const V3<R> r = (xyz(x0) + xyz(u)) - sp.center;
                const R dist = length(r);
                const R gap = dist - (sp.radius + S.c_node_radius);
                const R sb = length(vel - kv) + length(sp.omega) * dist;
                const R reach = maxR(S.c_margin, sb * S.dt) + S.c_slack;
                cand = gap <= reach;
            }
            L.cand[i] = cand ? 1 : 0;
        } else {
            L.cand[i] = 0;
        }
        if (!cand) {
            for (u32 p = 0; p < S.plane_count; ++p) L.approach[size_t(p) * L.N + i] = dot(vel, S.planes[p].normal);
            L.u4[i] = with(xyz(u) + S.dt * vel, u.w);
        }
    }
    if (SPHERE) {
        // The tile's candidates in node order (ballot per wave, waves in order),
        // so the striker pass only concatenates tile lists.
        __shared__ u32 s_wc[kWG / 64];
        const u32 lane = threadIdx.x & 63U, wave = threadIdx.x >> 6;
        const u64 ballot = __ballot(cand);
        if (lane == 0) s_wc[wave] = u32(__popcll(ballot));
        __syncthreads();
        u32 before = 0, total = 0;
        for (u32 w = 0; w < kWG / 64; ++w) {
            before += w < wave ? s_wc[w] : 0U;
            total += s_wc[w];
        }
        if (cand) L.cand_tile[tile * kWG + before + u32(__popcll(ballot & ((1ULL << lane) - 1ULL)))] = i;
        if (threadIdx.x == 0) L.tile_cand[tile] = total;
    }
    if (LOADS && loads) {
        __shared__ double sh[kWG / 64 + 1][13];
        groupSum13<kWG / 64>(sh, led);
        if (threadIdx.x == 0)
            for (int k = 0; k < 13; ++k) L.load_part[size_t(tile) * 13 + k] = sh[kWG / 64][k];
    }
}

// Exclusive prefix sum over the 1024 threads of a workgroup (wave shuffles,
// then the 16 wave totals); returns the sum before this thread, sets *total.
__device__ __forceinline__ u32 blockScan1024(u32 v, u32 *s_waves, u32 *total) {
    const u32 lane = threadIdx.x & 63U, wave = threadIdx.x >> 6;
    u32 x = v;
    for (u32 o = 1; o < 64; o <<= 1) {
        const u32 y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_waves[wave] = x;
    __syncthreads();
    u32 base = 0, all = 0;
    for (u32 w = 0; w < kFinishThreads / 64; ++w) {
        base += w < wave ? s_waves[w] : 0U;
        all += s_waves[w];
    }
    __syncthreads();
    *total = all;
    return base + x - v;
}

// ---------------------------------------------------------------------------
// Striker contact (LatticePhysics.hpp:1324-1387), one node, serial caller.
//
// The node's offset from the sphere centre, its distance, the contact normal,
// the gap and |r x n|^2 depend on the centre only, and the centre is constant
// through pass 1 and moves in pass 2 only when a position correction is
// applied. They are computed for all candidates in parallel ("speculated")
// and the serial chain uses them while the centre is the one they were
// computed at; every value is the same expression on the same inputs, so the
// bits are the reference's.

template <class R> struct ContactAcc {
    u64 contacts;
    double imp[3], dissipated, max_pen, max_corr;
};

template <class R> struct Spec {
    V3<R> r, n;
    R gap, crn;  // crn = |r x n|^2
};

template <class R> __device__ __forceinline__ Spec<R> speculate(const Step<R> &S, const Sphere<R> &sp, V3<R> x0, V3<R> u) {
    Spec<R> o;
    o.r = (x0 + u) - sp.center;
    const R dist = length(o.r);
    o.n = normalized(o.r);
    o.gap = dist - (sp.radius + S.c_node_radius);
    o.crn = length2(cross(o.r, o.n));
    return o;
}

// The contact of a candidate that passed the contact test at the current
// sphere state (a node of positive mass whose gap is within the margin or
// the closing distance), in two parts that read nothing the other writes: the
// impulses (LatticePhysics.hpp:1350-1376), which change the node's velocity
// and the sphere's velocity and spin, and pass 2's position correction
// (:1377-1386), which moves the node and the sphere centre by the gap alone.
// The normal impulse ni = max(0, (dvn - vn) / (1/m + 1/M + crn / I)) was
// evaluated at the same sphere state by the caller, and inm, ism, iin are 1/m,
// 1/M and 1/I: the reference's in-line expressions on the same inputs, taken
// off the serial chain. The impulse is written without branches (every value
// is computed and the applied ones selected) so that, on the chain, it can
// overlap the re-speculation a centre move calls for.
template <class R>
__device__ __forceinline__ void impulseNode(const Step<R> &S, Sphere<R> &sp, const Spec<R> &q, V3<R> &v, R mass,
                                            R inm, R ism, R iin, R ni, ContactAcc<R> &acc) {
    const V3<R> r = q.r, n = q.n;
    const bool on = ni > R(0);
    const R before = R(0.5) * mass * length2(v);
    const V3<R> imp = ni * n;
    const V3<R> nv1 = v + inm * imp;
    const V3<R> vel1 = sp.velocity - ism * imp;
    const V3<R> om1 = sp.omega + iin * (-cross(r, imp));
    const double i1[3] = {acc.imp[0] + double(imp.x), acc.imp[1] + double(imp.y), acc.imp[2] + double(imp.z)};
    const V3<R> anr = nv1 - vel1 - cross(om1, r);
    const V3<R> tan = anr - dot(anr, n) * n;
    const R ts = length(tan);
    const bool fr = on && ts > R(1.0e-12);
    const V3<R> dir = tan / ts;
    const R stick = ts / (inm + ism + iin * length2(cross(r, dir)));
    const R fi = stick <= S.c_sf * ni ? stick : minR(stick, S.c_df * ni);
    const V3<R> imp2 = -fi * dir;
    const V3<R> nv2 = nv1 + inm * imp2;
    const V3<R> vel2 = vel1 - ism * imp2;
    const V3<R> om2 = om1 + iin * (-cross(r, imp2));
    const double i2[3] = {i1[0] + double(imp2.x), i1[1] + double(imp2.y), i1[2] + double(imp2.z)};
    const V3<R> nv = fr ? nv2 : nv1;
    const double gone = double(before - R(0.5) * mass * length2(nv));
    if (on) {
        sp.velocity = fr ? vel2 : vel1;
        sp.omega = fr ? om2 : om1;
        for (int k = 0; k < 3; ++k) acc.imp[k] = fr ? i2[k] : i1[k];
        v = nv;
        acc.dissipated += gone;
        ++acc.contacts;
    }
}

// Pass 2's position correction of the node and the sphere centre.
template <class R>
__device__ __forceinline__ void correctNode(const Step<R> &S, Sphere<R> &sp, const Spec<R> &q, V3<R> &u, R inm, R ism,
                                            ContactAcc<R> &acc) {
    const R gap = q.gap;
    const V3<R> n = q.n;
    if (!(gap < -S.c_margin)) return;
    const R corr = minR(-gap - S.c_margin, R(0.2) * maxR(S.c_node_radius, sp.radius * R(0.01)));
    const R sum = inm + ism;
    const V3<R> ns = (corr * inm / sum) * n;
    const V3<R> ss = -((corr * ism / sum) * n);
    u = u + ns;
    sp.center = sp.center + ss;
    if (corr > R(acc.max_corr)) acc.max_corr = double(corr);
}

// Broadcast of one lane's value to the wave (the lane index is uniform).
__device__ __forceinline__ float laneValue(float x, u32 lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), int(lane)));
}
__device__ __forceinline__ double laneValue(double x, u32 lane) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffffLL), int(lane));
    const int hi = __builtin_amdgcn_readlane(int(b >> 32), int(lane));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
template <class R> __device__ __forceinline__ V3<R> laneValue(V3<R> v, u32 lane) {
    return {laneValue(v.x, lane), laneValue(v.y, lane), laneValue(v.z, lane)};
}

constexpr u32 kMaxNodeTiles = 16384;  // candidate tile prefix held in LDS: up to 4,194,304 nodes

// The serial striker pass of one substep (pass 1 or 2) on one workgroup.
// Pass 1 first forms the ordered per-block candidate lists from the per-tile
// lists k_predict left (CpuLatticeBackend.cpp:375-386): only the first 1024
// of each block are ever contacted, so only those are staged. The chain then
// runs in wave 0: each lane holds one candidate of a 64-wide window in
// registers (node data and the centre-only terms), and the whole wave walks
// the window in order with the walked candidate's values broadcast, so the
// sphere state is uniform and nothing on the chain waits on memory. Candidates
// whose contact test fails at the current sphere state are skipped by a ballot
// over the window. A centre move (pass 2) re-speculates the window's
// remaining candidates.
template <class R>
__device__ __forceinline__ void projectNode(const DevLattice<R> &L, const Step<R> &S, u32 i, R4<R> &u, bool &moved);

// OWN_VELOCITY (pass 2): the pass takes its listed candidates' last support
// projection and velocity update itself, the expressions of k_velocity, which
// then skips them; so the pass does not wait for k_velocity.
template <class R, int PASS, bool OWN_VELOCITY = false>
__global__ __launch_bounds__(PASS == 1 ? kFinishThreads : kPass2Threads) void k_pass(Batch<R> B, Step<R> S) {
    constexpr u32 kThreads = PASS == 1 ? kFinishThreads : kPass2Threads;
    const DevLattice<R> L = B.L[B.map[blockIdx.x].lat];
    DevCtl *ctl = L.ctl;
    if (ctl->stop) return;
    // Pass 1's tile prefix is dead once the candidate lists are staged, so the
    // candidates' reference positions reuse its memory.
    constexpr size_t kPrefixBytes = PASS == 1 ? kMaxNodeTiles * sizeof(u32) : 0;
    constexpr size_t kX0Bytes = kCap * sizeof(R4<R>);
    __shared__ __attribute__((aligned(16))) unsigned char s_raw[kPrefixBytes > kX0Bytes ? kPrefixBytes : kX0Bytes];
    u32 *s_prefix = reinterpret_cast<u32 *>(s_raw);
    R4<R> *s_x0 = reinterpret_cast<R4<R> *>(s_raw);
    __shared__ u32 s_waves[kFinishThreads / 64];
    __shared__ u32 s_ids[kCap];
    __shared__ R4<R> s_u[kCap], s_v[kCap];
    __shared__ u32 s_total;
    const u32 tid = threadIdx.x, lane = tid & 63U;
    if (PASS == 1) {
        u32 running = 0;
        for (u32 base = 0; base < L.node_tiles; base += kFinishThreads) {
            const u32 t = base + tid;
            const u32 cnt = t < L.node_tiles ? L.tile_cand[t] : 0U;
            u32 total = 0;
            const u32 at = running + blockScan1024(cnt, s_waves, &total);
            if (t < L.node_tiles) s_prefix[t] = at;
            running += total;
        }
        if (tid == 0) s_total = running;
        __syncthreads();
        const u32 K = s_total;
        // Candidate g (in node order over the whole lattice) is entry g - prefix[t] of tile t.
        auto tileOf = [&](u32 g) {
            u32 lo = 0, hi = L.node_tiles;
            while (hi - lo > 1) {
                const u32 mid = (lo + hi) >> 1;
                if (s_prefix[mid] <= g) lo = mid;
                else hi = mid;
            }
            return lo;
        };
        auto idOf = [&](u32 g) {
            const u32 t = tileOf(g);
            return L.cand_tile[t * kWG + (g - s_prefix[t])];
        };
        for (u32 b = tid; b < L.blocks; b += kFinishThreads) {
            auto lower = [&](u32 key) {
                u32 lo = 0, hi = K;
                while (lo < hi) {
                    const u32 mid = (lo + hi) >> 1;
                    if (idOf(mid) < key) lo = mid + 1;
                    else hi = mid;
                }
                return lo;
            };
            const u32 lo = b == 0 ? 0U : lower(L.nbb[b]);
            const u32 hi = b + 1 == L.blocks ? K : lower(L.nbb[b + 1]);
            const u32 count = hi - lo;
            L.cand_lo[b] = lo;
            L.cand_cnt[b] = count < kCap ? count : kCap;
            if (count > kCap) atomicAdd((unsigned long long *)&ctl->candidate_overflow, (unsigned long long)(count - kCap));
        }
        __syncthreads();
        for (u32 b = 0; b < L.blocks; ++b) {
            const u32 lo = L.cand_lo[b], cnt = L.cand_cnt[b];
            for (u32 e = tid; e < cnt; e += kFinishThreads) L.cand_all[size_t(b) * kCap + e] = idOf(lo + e);
        }
        __syncthreads();
    }
    ContactAcc<R> acc;
    Sphere<R> sp = *L.sphere;
    if (PASS == 1) sp.velocity = sp.velocity + S.dt * S.gravity;
    // The sphere's mass and inertia are constant through the pass.
    const R ism = R(1) / sp.mass, iin = R(1) / sp.inertia;
    acc.contacts = ctl->impulse_contacts;
    for (int k = 0; k < 3; ++k) acc.imp[k] = ctl->impulse[k];
    acc.dissipated = ctl->dissipated;
    acc.max_pen = ctl->max_penetration;
    acc.max_corr = ctl->max_correction;
    const V3<R> center_before = sp.center;
    const R energy_before = R(0.5) * sp.mass * length2(sp.velocity) + R(0.5) * sp.inertia * length2(sp.omega);
    for (u32 b = 0; b < L.blocks; ++b) {
        const u32 n = L.cand_cnt[b];
        for (u32 k = tid; k < n; k += kThreads) {
            const u32 i = L.cand_all[size_t(b) * kCap + k];
            s_ids[k] = i;
            if (PASS == 2 && OWN_VELOCITY) {
                R4<R> u = L.u4[i];
                bool moved = false;
                projectNode(L, S, i, u, moved);
                s_u[k] = u;
                s_v[k] = with((xyz(u) - xyz(L.up4[i])) / S.dt, R(0));
            } else {
                s_u[k] = L.u4[i];
                s_v[k] = L.v4[i];
            }
            s_x0[k] = L.x04[i];
        }
        __syncthreads();
        if (tid < 64) {
            for (u32 k0 = 0; k0 < n;) {
                // This lane's candidate of the window, in registers.
                const u32 k = k0 + lane;
                const bool have = k < n;
                V3<R> x0{R(0), R(0), R(0)}, mu{R(0), R(0), R(0)}, mv{R(0), R(0), R(0)};
                R mass = R(0);
                if (have) {
                    const R4<R> x = s_x0[k];
                    x0 = xyz(x);
                    mass = x.w;
                    mu = xyz(s_u[k]);
                    mv = xyz(s_v[k]);
                }
                Spec<R> q = speculate(S, sp, x0, mu);
                const u32 span = min(64U, n - k0);
                // The impulse denominator (LatticePhysics.hpp:1344-1346), constant while the centre is.
                const R inm = R(1) / mass;
                R den = inm + ism + iin * q.crn;
                R closing = -q.gap / S.dt;  // the desired normal speed of a gap beyond the margin
                // A candidate changes nothing but the maximum penetration unless
                // it passes the contact test (LatticePhysics.hpp:1341) and then
                // takes an impulse or, in pass 2, a position correction. Both
                // read only its own terms and the sphere's velocity and spin, so
                // every lane evaluates them at the current sphere state and the
                // chain jumps to the first lane that acts; the lanes before it
                // fold their penetration into the running maximum in order. An
                // impulse changes the sphere state, so the remaining lanes are
                // evaluated again; a centre move (pass 2) re-speculates them.
                u32 pos = 0;
                while (pos < span) {
                    const V3<R> rel = mv - sp.velocity - cross(sp.omega, q.r);
                    const R vn = dot(rel, q.n);
                    const bool touch = lane >= pos && lane < span && !(mass <= R(0)) &&
                                       !(q.gap > maxR(S.c_margin, -vn * S.dt));
                    const R rest = -vn > S.c_rest_thr ? S.c_rest : R(0);
                    const R dvn = q.gap > S.c_margin ? closing : -rest * minR(vn, R(0));
                    const R ni = maxR(R(0), (dvn - vn) / den);
                    const bool acts = touch && (ni > R(0) || (PASS == 2 && q.gap < -S.c_margin));
                    const u64 am = __ballot(acts);
                    const u32 w = am ? u32(__ffsll((long long)am) - 1) : span;
                    u64 pen = __ballot(touch && lane < w && -q.gap > R(acc.max_pen));
                    while (pen) {
                        const u32 l = u32(__ffsll((long long)pen) - 1);
                        acc.max_pen = double(laneValue(-q.gap, l));
                        pen = __ballot(touch && lane > l && lane < w && -q.gap > R(acc.max_pen));
                    }
                    if (am == 0) break;
                    Spec<R> qw;
                    qw.r = laneValue(q.r, w);
                    qw.n = laneValue(q.n, w);
                    qw.gap = laneValue(q.gap, w);
                    qw.crn = laneValue(q.crn, w);
                    V3<R> u = laneValue(mu, w), v = laneValue(mv, w);
                    const R mw = laneValue(mass, w), iw = laneValue(inm, w), nw = laneValue(ni, w);
                    if (-qw.gap > R(acc.max_pen)) acc.max_pen = double(-qw.gap);
                    if (PASS == 2) {
                        // The correction reads only the gap and the inverse
                        // masses, so it goes first and the window's
                        // re-speculation at the new centre overlaps the
                        // impulse; a centre that did not move gives the same
                        // bits again.
                        correctNode<R>(S, sp, qw, u, iw, ism, acc);
                        q = speculate(S, sp, x0, mu);
                        den = inm + ism + iin * q.crn;
                        closing = -q.gap / S.dt;
                    }
                    impulseNode<R>(S, sp, qw, v, mw, iw, ism, iin, nw, acc);
                    if (lane == w) {
                        mu = u;
                        mv = v;
                    }
                    pos = w + 1;
                }
                if (have) {
                    s_u[k] = with(mu, s_u[k].w);
                    s_v[k] = with(mv, R(0));
                }
                k0 += span;
            }
        }
        __syncthreads();
        for (u32 k = tid; k < n; k += kThreads) {
            const u32 i = s_ids[k];
            V3<R> v = xyz(s_v[k]);
            R4<R> u = s_u[k];
            if (PASS == 1) {
                // nodeRecordApproach and the prediction (LatticePhysics.hpp:1408-1411).
                for (u32 p = 0; p < S.plane_count; ++p) L.approach[size_t(p) * L.N + i] = dot(v, S.planes[p].normal);
                u = with(xyz(u) + S.dt * v, u.w);
                L.cand[i] = 2;  // listed: striker pass 2 contacts it
            } else {
                // Candidates take their support velocity response here (LatticePhysics.hpp:1419-1425).
                for (u32 p = 0; p < S.plane_count; ++p)
                    if (L.engaged[size_t(p) * L.N + i]) v = supportVelocity(S.planes[p], v, L.approach[size_t(p) * L.N + i]);
            }
            L.u4[i] = u;
            L.v4[i] = with(v, R(0));
        }
        __syncthreads();
    }
    if (tid == 0) {
        acc.dissipated += double(energy_before - (R(0.5) * sp.mass * length2(sp.velocity) + R(0.5) * sp.inertia * length2(sp.omega)));
        const R shift = length(sp.center - center_before);
        if (shift > R(ctl->max_shift)) ctl->max_shift = double(shift);
        ctl->impulse_contacts = acc.contacts;
        for (int k = 0; k < 3; ++k) ctl->impulse[k] = acc.imp[k];
        ctl->dissipated = acc.dissipated;
        ctl->max_penetration = acc.max_pen;
        ctl->max_correction = acc.max_corr;
        *L.sphere = sp;
    }
}

// ---------------------------------------------------------------------------
// XPBD bond projection, one colour stage (LatticePhysics.hpp:1132-1168). The
// stage's bonds share no node, so every thread owns its two ends.

// The arithmetic of one bond solve (LatticePhysics.hpp:1133-1168) on node
// values held in registers (or LDS): ua and ub move when the bond solves. g is
// the bond's edge palette entry, lam the multiplier it reads (0 in the first
// iteration); a solve stores lam + dl in lam_out and returns true, a bond
// shorter than 1e-12 returns false. r.a and r.b (the global node indices) are
// read only with DIRECT, c always and p only through g. alpha is the record's
// compliance / dt^2, computed once at upload by the same float operations.
template <class R, bool DIRECT, bool PLASTIC>
__device__ __forceinline__ bool solveBondWith(const DevLattice<R> &L, const Step<R> &S, const EdgePal<R> &g, u32 a,
                                              u32 b, R alpha, u32 j, R lam, R4<R> &ua, R4<R> &ub, R &lam_out) {
    const R4<R> ge = g.e;
    const V3<R> du = xyz(ub) - xyz(ua);
    const V3<R> e = xyz(ge);
    const V3<R> delta = DIRECT ? (xyz(ldg(&L.x04[b])) + xyz(ub)) - (xyz(ldg(&L.x04[a])) + xyz(ua)) : e + du;
    const R len = length(delta);
    if (len <= R(1.0e-12)) return false;
    const R orig = ge.w;
    R C;
    if (PLASTIC) {
        const R plastic = ldg(&L.pext[j]);
        const R rest = orig + plastic;
        C = DIRECT ? len - rest : (g.d.x + R(2) * dot(e, du) + dot(du, du) - plastic * (rest + orig)) / (len + rest);
    } else {
        // plastic == 0: rest == orig and the subtracted term is an exact +0.
        C = DIRECT ? len - orig : (g.d.x + R(2) * dot(e, du) + dot(du, du)) / (len + orig);
    }
    const V3<R> dir = delta / len;
    const R wa = ua.w, wb = ub.w;
    const R dl = (-C - alpha * lam) / (wa + wb + alpha);
    lam_out = lam + dl;
    ua = with(xyz(ua) - wa * dl * dir, ua.w);
    ub = with(xyz(ub) + wb * dl * dir, ub.w);
    return true;
}

template <class R, bool DIRECT, bool PLASTIC>
__device__ __forceinline__ bool solveBond(const DevLattice<R> &L, const Step<R> &S, const BondRec<R> &r, u32 j, R lam,
                                          R4<R> &ua, R4<R> &ub, R &lam_out) {
    const EdgePal<R> g = ldg(&L.epal[r.p & ~kNoDir]);
    return solveBondWith<R, DIRECT, PLASTIC>(L, S, g, r.a, r.b, r.c, j, lam, ua, ub, lam_out);
}

template <class R, bool DIRECT, bool PLASTIC>
__global__ __launch_bounds__(kWG) void k_solve(Batch<R> B, Step<R> S, u32 flags) {
    const WgMap *m = tileEntry(B);
    if (!m) return;
    const DevLattice<R> L = B.L[m->lat];
    const u32 j = m->a + threadIdx.x;
    if (j >= m->b) return;
    const bool first = flags & kSolveFirst;
    // The record is loaded with the alive flag and the run's stop flag, so the
    // node gathers wait on one memory round trip, not two or three.
    const BondRec<R> r = ldStream(&L.rec[j]);
    const u8 live = ldStream(&L.alive[j]);
    const R lam = first ? R(0) : ldg(&L.lambda[j]);
    if (L.ctl->stop || !live) return;
    R4<R> ua = ldg(&L.u4[r.a]), ub = ldg(&L.u4[r.b]);
    R lam_out;
    if (!solveBond<R, DIRECT, PLASTIC>(L, S, r, j, lam, ua, ub, lam_out)) return;
    // The multiplier is read only by a later iteration, so a one-iteration run
    // never stores it.
    if (flags & kSolveKeepLambda) stg(&L.lambda[j], lam_out);
    stStream(&L.u4[r.a], ua);
    stStream(&L.u4[r.b], ub);
}

// The grouped colour sweep of a one-iteration run: K = 2 (a pair) or 3 (a
// triple) consecutive colour stages s .. s + K - 1 in one launch, each
// workgroup on a tile of nodes staged in LDS. A workgroup owns a set of nodes
// and loads, besides them, every node the owned nodes' final values depend
// on. Going back from the last stage: the owned nodes need their stage
// s + K - 1 bonds solved, those bonds' other ends need their values after
// stage s + K - 2, so their stage s + K - 2 bonds are solved too, and so on
// down to stage s, whose bonds' ends are loaded as they stand before the
// launch. For a pair and an owned node x with stage-s partner p1(x) and
// stage-(s + 1) partner p2(x), that is p1(x), p2(x) and p1(p2(x)). The
// workgroup solves its stage lists in order on its LDS copies (a stage shares
// no node, so no two threads touch one slot) and writes back only the owned
// nodes. A bond solved by two workgroups is solved from the same inputs in
// both, so each owned node gets the bits the stage-by-stage sweep gives.
// Workgroups read one node buffer and write the other (u4 <-> u4b), so none
// reads a node another has already written; every node is owned by one
// workgroup, so the written buffer is whole. Groups alternate direction and
// a sweep has an even number of them, so it ends in u4; the stages left over
// run in place on u4 as single entries. Each workgroup has two map entries,
// read together:
//   single: {lat, a, b, pad}: a, b = bond range of one stage (at most
//           kPairThreads bonds); the second entry unused;
//   group:  {lat, node list offset, bond list offset, pad},
//           {owned | all nodes << 16, stage-s | stage-(s + 1) bonds << 16,
//            stage-(s + 2) bonds (0 in a pair), 0};
// pad = kind (bits 16-19) | direction (bit 20: u4b -> u4). A launch has
// xcdGrid(count / 2) workgroups; a triple launch also runs pair entries (a
// batch merges its lattices' groups by position).
constexpr u32 kSweepSingle = 0U, kSweepPair = 1U, kSweepTriple = 2U;
constexpr u32 kSweepReverse = 1U << 20;
constexpr u32 kPairThreads = 512;
constexpr u32 kPairNodeRegs = 4;  // node slots per thread: a workgroup holds at most 2048 nodes
constexpr u32 kPairBondRegs = 2;  // bonds per thread per stage: at most 1024 per stage
constexpr u32 kPairMaxNodes = kPairThreads * kPairNodeRegs;
// A workgroup's tile in cells of the reference grid (x, y, z): 1,024 owned
// nodes of a full lattice, rows of 16 nodes along the fastest axis of the box
// lattices' node order; lattices below kPairWideNodes take 512-node tiles, so
// that a launch still has several workgroups per CU.
constexpr u32 kPairTile[3] = {16, 8, 8};
constexpr u32 kPairTileSmall[3] = {8, 8, 8};
constexpr u32 kPairWideNodes = 1U << 19;
constexpr u32 kPairMaxBonds = kPairThreads * kPairBondRegs;

template <class R, bool DIRECT, bool PLASTIC>
__device__ __forceinline__ void pairStage(const DevLattice<R> &L, const Step<R> &S, const u32 *nodes, R4<R> *u,
                                          const PairBond<R> (&e)[kPairBondRegs], const u8 (&live)[kPairBondRegs],
                                          u32 n) {
#pragma unroll
    for (u32 q = 0; q < kPairBondRegs; ++q) {
        if (threadIdx.x + q * kPairThreads >= n || !live[q]) continue;
        const EdgePal<R> g = ldg(&L.epal[e[q].p & ~kNoDir]);
        const u32 sa = e[q].slots & 0xffffU, sb = e[q].slots >> 16;
        R4<R> ua = u[sa], ub = u[sb];
        const u32 a = DIRECT ? ldg(&nodes[sa]) : 0U, b = DIRECT ? ldg(&nodes[sb]) : 0U;
        R lam_out;
        if (!solveBondWith<R, DIRECT, PLASTIC>(L, S, g, a, b, e[q].c, e[q].j, R(0), ua, ub, lam_out)) continue;
        u[sa] = ua;
        u[sb] = ub;
    }
}

template <class R, bool DIRECT, bool PLASTIC, int K>
__device__ __forceinline__ void sweepGroup(const Batch<R> &B, const Step<R> &S) {
    extern __shared__ __align__(16) unsigned char pair_lds[];
    R4<R> *u = reinterpret_cast<R4<R> *>(pair_lds);
    const u32 wgs = B.count / 2, w = xcdTile(blockIdx.x, wgs);
    if (w >= wgs) return;
    const WgMap m0 = ldg(&B.map[2 * w]), m1 = ldg(&B.map[2 * w + 1]);
    const WgMap *m = &m0;
    const DevLattice<R> L = B.L[m->lat];
    const u32 t = threadIdx.x;
    if (((m->pad >> 16) & 0xfU) == kSweepSingle) {
        const u32 j = m->a + t;
        if (j >= m->b) return;
        const BondRec<R> r = ldStream(&L.rec[j]);
        const u8 live = ldStream(&L.alive[j]);
        if (L.ctl->stop || !live) return;
        R4<R> ua = ldg(&L.u4[r.a]), ub = ldg(&L.u4[r.b]);
        R lam_out;
        if (!solveBond<R, DIRECT, PLASTIC>(L, S, r, j, R(0), ua, ub, lam_out)) return;
        stStream(&L.u4[r.a], ua);
        stStream(&L.u4[r.b], ub);
        return;
    }
    const bool rev = m->pad & kSweepReverse;
    const R4<R> *src = rev ? L.u4b : L.u4;
    R4<R> *dst = rev ? L.u4 : L.u4b;
    // The node ids and bond entries go out with the stop flag.
    const u32 n_own = m1.lat & 0xffffU, n_all = m1.lat >> 16;
    const u32 nb[3] = {m1.a & 0xffffU, m1.a >> 16, K > 2 ? m1.b : 0U};
    const u32 *nodes = L.pnodes + m->a;
    const PairBond<R> *bonds = L.pbonds + m->b;
    u32 id[kPairNodeRegs];
#pragma unroll
    for (u32 q = 0; q < kPairNodeRegs; ++q) {
        const u32 k = t + q * kPairThreads;
        id[q] = k < n_all ? ldStream(&nodes[k]) : 0U;
    }
    PairBond<R> e[K][kPairBondRegs] = {};
    u8 l[K][kPairBondRegs] = {};
#pragma unroll
    for (int g = 0, at = 0; g < K; at += nb[g], ++g)
#pragma unroll
        for (u32 q = 0; q < kPairBondRegs; ++q) {
            const u32 k = t + q * kPairThreads;
            if (k < nb[g]) e[g][q] = ldStream(&bonds[at + k]);
        }
    // Only the end of a substep (k_finish) sets stop, so it holds for the launch.
    if (L.ctl->stop) return;
#pragma unroll
    for (u32 q = 0; q < kPairNodeRegs; ++q) {
        const u32 k = t + q * kPairThreads;
        if (k < n_all) u[k] = ldg(&src[id[q]]);
    }
#pragma unroll
    for (int g = 0; g < K; ++g)
#pragma unroll
        for (u32 q = 0; q < kPairBondRegs; ++q) {
            const u32 k = t + q * kPairThreads;
            if (k < nb[g]) l[g][q] = ldg(&L.alive[e[g][q].j]);
        }
#pragma unroll
    for (int g = 0; g < K; ++g) {
        __syncthreads();
        pairStage<R, DIRECT, PLASTIC>(L, S, nodes, u, e[g], l[g], nb[g]);
    }
    __syncthreads();
#pragma unroll
    for (u32 q = 0; q < kPairNodeRegs; ++q) {
        const u32 k = t + q * kPairThreads;
        if (k < n_own) stStream(&dst[id[q]], u[k]);
    }
}

// Eight waves per SIMD in float: the launch's workgroups of a full lattice
// are all resident at once, so none waits for a second round.
template <class R, bool DIRECT, bool PLASTIC>
__global__ __launch_bounds__(kPairThreads, sizeof(R) == 4 ? 8 : 4) void k_sweep_pair(Batch<R> B, Step<R> S) {
    sweepGroup<R, DIRECT, PLASTIC, 2>(B, S);
}
template <class R, bool DIRECT, bool PLASTIC>
__global__ __launch_bounds__(kPairThreads, sizeof(R) == 4 ? 8 : 4) void k_sweep_triple(Batch<R> B, Step<R> S) {
    sweepGroup<R, DIRECT, PLASTIC, 3>(B, S);
}

// The whole colour sweep of a one-iteration run for a lattice whose nodes fit
// in LDS (FastLattice.hpp:229-231, the single-block shared-memory mode): one
// workgroup per lattice loads every node's displacement into LDS, runs the
// stages in order with a workgroup barrier between them, and writes the nodes
// back. A stage's bonds share no node, so the threads of a stage never touch
// the same LDS entry, and each solve is solveBond on the same values the
// stage-by-stage launches read from memory: the same bits. Launch one
// workgroup per map entry (lat = the lattice) with kLdsPal * sizeof(EdgePal)
// + N * sizeof(R4<R>) bytes of dynamic LDS and 256, 512 or 1024 threads.
//
// When every stage is one bond range of at most kLdsBondRegs bonds per thread
// and the edge palette fits its LDS slots (lds_fast), the palette is copied
// into LDS and each stage's bond records and alive flags are loaded while the
// stage before it is solved, so a stage waits on no memory round trip.
constexpr int kSweepLdsThreads = 1024;
constexpr u32 kLdsPal = 128;       // edge palette entries held in the LDS sweep's LDS
constexpr u32 kLdsBondRegs = 4;    // bonds per thread per stage on the prefetching path
template <class R, bool DIRECT, bool PLASTIC>
__global__ __launch_bounds__(kSweepLdsThreads) void k_sweep_lds(Batch<R> B, Step<R> S) {
    extern __shared__ __align__(16) unsigned char sweep_lds[];
    EdgePal<R> *pal = reinterpret_cast<EdgePal<R> *>(sweep_lds);
    R4<R> *u = reinterpret_cast<R4<R> *>(sweep_lds + kLdsPal * sizeof(EdgePal<R>));
    const DevLattice<R> L = B.L[B.map[blockIdx.x].lat];
    // Only the end of a substep (k_finish) sets stop, so it holds for the sweep.
    if (ldg(&L.ctl->stop)) return;
    const u32 N = L.N, T = blockDim.x, t = threadIdx.x;
    for (u32 i = t; i < N; i += T) u[i] = ldg(&L.u4[i]);
    if (L.lds_fast && L.stage_max <= kLdsBondRegs * T) {
        for (u32 k = t; k < L.npal; k += T) pal[k] = ldg(&L.epal[k]);
        BondRec<R> ra[kLdsBondRegs], rb[kLdsBondRegs];
        u32 ja[kLdsBondRegs], jb[kLdsBondRegs];
        u8 la[kLdsBondRegs], lb[kLdsBondRegs];
        // Stage st's records: bond j = lo + t + k * T for slot k.
        auto fetch = [&](u32 st, BondRec<R> (&r)[kLdsBondRegs], u32 (&jj)[kLdsBondRegs], u8 (&l)[kLdsBondRegs]) {
            const uint2 rg = ldg(&L.stage_ranges[st]);
#pragma unroll
            for (u32 k = 0; k < kLdsBondRegs; ++k) {
                const u32 j = rg.x + t + k * T;
                jj[k] = j;
                l[k] = 0;
                if (j < rg.y) {
                    r[k] = ldg(&L.rec[j]);
                    l[k] = ldg(&L.alive[j]);
                }
            }
        };
        if (L.stage_count > 0) fetch(0, ra, ja, la);
        __syncthreads();
        for (u32 st = 0; st < L.stage_count; ++st) {
            if (st + 1 < L.stage_count) fetch(st + 1, rb, jb, lb);
#pragma unroll
            for (u32 k = 0; k < kLdsBondRegs; ++k) {
                if (!la[k]) continue;
                const BondRec<R> &r = ra[k];
                R4<R> ua = u[r.a], ub = u[r.b];
                R lam_out;
                if (!solveBondWith<R, DIRECT, PLASTIC>(L, S, pal[r.p & ~kNoDir], r.a, r.b, r.c, ja[k], R(0), ua, ub,
                                                       lam_out))
                    continue;
                u[r.a] = ua;
                u[r.b] = ub;
            }
            __syncthreads();
#pragma unroll
            for (u32 k = 0; k < kLdsBondRegs; ++k) {
                ra[k] = rb[k];
                ja[k] = jb[k];
                la[k] = lb[k];
            }
        }
    } else {
        __syncthreads();
        u32 q0 = 0;
        for (u32 st = 0; st < L.stage_count; ++st) {
            const u32 q1 = ldg(&L.stage_end[st]);
            for (u32 q = q0; q < q1; ++q) {
                const uint2 rg = ldg(&L.stage_ranges[q]);
                for (u32 j = rg.x + t; j < rg.y; j += T) {
                    const BondRec<R> r = ldg(&L.rec[j]);
                    if (!ldg(&L.alive[j])) continue;
                    R4<R> ua = u[r.a], ub = u[r.b];
                    R lam_out;
                    if (!solveBond<R, DIRECT, PLASTIC>(L, S, r, j, R(0), ua, ub, lam_out)) continue;
                    u[r.a] = ua;
                    u[r.b] = ub;
                }
            }
            q0 = q1;
            __syncthreads();
        }
    }
    for (u32 i = t; i < N; i += T) stg(&L.u4[i], u[i]);
}

// Radial pair damping, one colour stage (LatticePhysics.hpp:1171-1184).
template <class R, bool DIRECT>
__global__ __launch_bounds__(kWG) void k_damp(Batch<R> B, Step<R> S) {
    const WgMap *m = tileEntry(B);
    if (!m) return;
    const DevLattice<R> L = B.L[m->lat];
    if (L.ctl->stop) return;
    const u32 j = m->a + threadIdx.x;
    if (j >= m->b) return;
    if (!L.alive[j]) return;
    const BondRec<R> r = L.rec[j];
    const R ma = L.x04[r.a].w, mb = L.x04[r.b].w;
    if (ma <= R(0) || mb <= R(0)) return;
    const V3<R> n = normalized(bondVec<R, DIRECT>(L, r.a, r.b, L.epal[r.p & ~kNoDir].e, L.u4[r.a], L.u4[r.b]));
    const V3<R> va = xyz(L.v4[r.a]), vb = xyz(L.v4[r.b]);
    const R rs = dot(vb - va, n);
    const R im = R(1) / ma + R(1) / mb;
    const R imp = -S.damping * rs / im;
    L.v4[r.a] = with(va - (imp / ma) * n, R(0));
    L.v4[r.b] = with(vb + (imp / mb) * n, R(0));
}

// Support projection of one node (LatticePhysics.hpp:1204-1227, 1256-1261).
template <class R> __device__ __forceinline__ void projectNode(const DevLattice<R> &L, const Step<R> &S, u32 i, R4<R> &u, bool &moved) {
    const V3<R> x0 = xyz(L.x04[i]);
    for (u32 p = 0; p < S.plane_count; ++p) {
        const Plane<R> &pl = S.planes[p];
        const V3<R> pos = x0 + xyz(u);
        if (!insideFootprints(pl, pos)) continue;
        const R dist = dot(pos - pl.point, pl.normal) - pl.node_radius;
        if (dist > R(1.0e-8)) continue;
        if (pl.reach_capped) {
            const R appr = L.approach[size_t(p) * L.N + i];
            const R reach = R(8) * maxR(R(0), -appr) * S.dt + R(1.0e-5);
            if (-dist > reach) continue;
        }
        u = with(xyz(u) - dist * pl.normal, u.w);
        moved = true;
        if (-dist > R(0)) L.engaged[size_t(p) * L.N + i] = 1;
    }
    // The kinematic blade after the planes (an extension; see bladePush).
    V3<R> d;
    if (S.blade.enabled && bladePush(S.blade, S.dt, L.ctl->step, x0 + xyz(u), d)) {
        u = with(xyz(u) + d, u.w);
        moved = true;
    }
}

template <class R> __global__ __launch_bounds__(kWG) void k_project(Batch<R> B, Step<R> S) {
    const WgMap *m = tileEntry(B);
    if (!m) return;
    const DevLattice<R> L = B.L[m->lat];
    if (L.ctl->stop) return;
    const u32 i = m->a * kWG + threadIdx.x;
    if (i >= L.N) return;
    R4<R> u = L.u4[i];
    bool moved = false;
    projectNode(L, S, i, u, moved);
    if (moved) L.u4[i] = u;
}

// Last iteration's support projection fused with nodeVelocityUpdate
// (LatticePhysics.hpp:1726-1730): both only touch node i. SKIP_LISTED leaves
// the candidates on striker pass 2's lists to that pass, and the first
// workgroups first mark the nodes pass 2 can reach (k_mark_affected's work over
// the entries of M), which the end sample after this kernel reads.
template <class R, bool PROJECT, bool SKIP_LISTED = false>
__global__ __launch_bounds__(kWG) void k_velocity(Batch<R> B, Step<R> S, Batch<R> M) {
    if (SKIP_LISTED)
        for (u32 e = blockIdx.x; e < M.count; e += gridDim.x) markAffected(M, e);
    const WgMap *m = tileEntry(B);
    if (!m) return;
    const DevLattice<R> L = B.L[m->lat];
    if (L.ctl->stop) return;
    {
        const u32 i = m->a * kWG + threadIdx.x;
        if (i >= L.N) return;
        if (SKIP_LISTED && L.cand[i] == 2) return;
        R4<R> u = L.u4[i];
        if (PROJECT) {
            bool moved = false;
            projectNode(L, S, i, u, moved);
            if (moved) L.u4[i] = u;
        }
        V3<R> vel = (xyz(u) - xyz(L.up4[i])) / S.dt;
        if (S.damping <= R(0) && !L.cand[i])
            for (u32 p = 0; p < S.plane_count; ++p)
                if (L.engaged[size_t(p) * L.N + i]) vel = supportVelocity(S.planes[p], vel, L.approach[size_t(p) * L.N + i]);
        L.v4[i] = with(vel, R(0));
    }
}

// Support velocity of non-candidates after the damping sweep (CpuLatticeBackend.cpp:414-415).
template <class R> __global__ __launch_bounds__(kWG) void k_support_velocity(Batch<R> B, Step<R> S) {
    const WgMap *m = tileEntry(B);
    if (!m) return;
    const DevLattice<R> L = B.L[m->lat];
    if (L.ctl->stop) return;
    const u32 i = m->a * kWG + threadIdx.x;
    if (i >= L.N || L.cand[i]) return;
    {
        V3<R> vel = xyz(L.v4[i]);
        bool any = false;
        for (u32 p = 0; p < S.plane_count; ++p)
            if (L.engaged[size_t(p) * L.N + i]) {
                vel = supportVelocity(S.planes[p], vel, L.approach[size_t(p) * L.N + i]);
                any = true;
            }
        if (any) L.v4[i] = with(vel, R(0));
    }
}

// ---------------------------------------------------------------------------
// End-of-substep sample, damage and break (LatticePhysics.hpp:1068-1114).
// The bonds are walked in node-major order, so a workgroup's bonds share their
// nodes and the node gathers hit in cache (schedule order is colour-major and
// would sweep all node data once per colour). A bond whose break or damage
// could change the ordered folds sets its bit in ev_bits, by schedule index,
// so k_finish reads the events back in bond order, whatever order the bonds
// were sampled in.

// Which bonds a k_bond_end launch covers: all of them, or, while striker pass
// 2 runs, those with neither end affected, and after it the rest.
enum BondCover : int { kBondAll = 0, kBondUnaffected = 2, kBondAffected = 3 };

// The streamed part of a bond's end sample: its record, previous peaks and
// damage, and alive flag. A workgroup loads these for every bond of its span
// before it starts on the first, so the span's first memory round trips overlap.
template <class R> struct EndLoad {
    BondNode r;
    R2<R> pv;
    u32 live;
};
template <class R> __device__ __forceinline__ EndLoad<R> loadEnd(const DevLattice<R> &L, u32 o) {
    return {ldStream(&L.bnd[o]), ldStream(&L.pv[o]), u32(ldStream(&L.alive_n[o]))};
}

// One bond's end sample, damage and break. Returns through the references the
// bond's contributions to the running maxima, its plastic work and the nodes
// it claims for the dirty list.
template <class R, bool DIRECT, bool PLASTIC, int COVER>
__device__ __forceinline__ void endSampleBond(const DevLattice<R> &L, const Step<R> &S, u32 o, const EndLoad<R> &in,
                                              double m0, u32 stamp, R shear_floor2, float &mt, float &mc, float &ms,
                                              float &mp, double &pw, u32 &claim_a, u32 &claim_b) {
    const BondNode r = in.r;
    R2<R> pv = in.pv;
    if (!in.live) return;
    const u32 j = r.j;
    // Every gather the bond needs goes out in one batch, before any test on
    // what it returns: one memory round trip after the record, not three.
    // A tensile-only lattice reads only ds and de (its compression and shear
    // ramps are 0 without them), and its previous compressive and shear
    // samples are 0.
    Thresholds<R> th;
    R2<R> pyz{R(0), R(0)};
    if (L.tensile_only) {
        const Tensile<R> t2 = ldStream(&L.thr2[o]);
        th.ds = t2.ds;
        th.de = t2.de;
    } else {
        th = L.thr[o];
        pyz = ldStream(&L.pyz[o]);
    }
    u32 aff_a = 0, aff_b = 0;
    if (COVER == kBondUnaffected) {
        aff_a = ldg(&L.affected[r.a]);
        aff_b = ldg(&L.affected[r.b]);
    }
    const R4<R> ua = ldg(&L.u4[r.a]), ub = ldg(&L.u4[r.b]);
    const EdgePal<R> g = ldg(&L.epal[r.p & ~kNoDir]);
    R4<R> st[4];  // read only when loaded (a bond with a rest direction)
    loadEndStrains(L, r.a, r.b, r.p, st);
    R plastic = R(0), ps = R(0);
    if (PLASTIC) {
        plastic = ldg(&L.pext[j]);
        ps = ldg(&L.pstr[o]);
    }
    if (COVER == kBondUnaffected && (aff_a == stamp || aff_b == stamp)) return;
    const R restL = g.e.w;
    if (PLASTIC) {
        // bondPlasticReturn (LatticePhysics.hpp:1018-1039).
        const R elastic = length(bondVec<R, DIRECT>(L, r.a, r.b, g.e, ua, ub)) - restL - plastic;
        const R mag = absR(elastic);
        const R h = S.hardening;
        const R ye = S.yield * restL + h * ps;
        if (mag > ye) {
            const R inc = (mag - ye) / (R(1) + h);
            plastic = plastic + (elastic < R(0) ? -inc : inc);
            ps = ps + inc;
            L.pext[j] = plastic;
            L.pstr[o] = ps;
            const R c = L.cmpl[j];
            if (c > R(0)) {
                const R yend = ye + h * inc;
                pw = double(R(0.5) * (ye + yend) * inc / c);
            }
        }
        mp = fmaxf(mp, float(absR(plastic) / restL));
    }
    R t, c, s;
    // On a tensile-only lattice the shear only feeds the running maximum (and
    // the stored sample at a break, recomputed there), so a shear that cannot
    // raise the maximum skips its square root.
    sampleBondWith<R, DIRECT>(L, r.a, r.b, r.p, g, ua, ub, st, t, c, s, L.tensile_only ? shear_floor2 : R(0));
    // The previous compressive and shear samples of a tensile-only lattice
    // were counted in the maxima when they were taken (k_bond_start).
    const R pt = maxR(R(0), maxR(pv.x, t)), pc = maxR(R(0), maxR(pyz.x, c)), pss = maxR(R(0), maxR(pyz.y, s));
    mt = fmaxf(mt, float(pt));
    mc = fmaxf(mc, float(pc));
    ms = fmaxf(ms, float(pss));
    pv.x = maxR(R(0), t);
    pyz.x = maxR(R(0), c);
    pyz.y = maxR(R(0), s);
    // With every compression and shear start non-finite their ramps are an
    // exact 0 (damageProgress), which never wins below.
    const R td = progress(pt, th.ds, th.de);
    const R cd = L.tensile_only ? R(0) : progress(pc, th.cs, th.ce);
    const R sd = L.tensile_only ? R(0) : progress(pss, th.ss, th.se);
    R dmg = td;
    u8 md = 1;
    if (cd > dmg) { dmg = cd; md = 2; }
    if (sd > dmg) { dmg = sd; md = 3; }
    if (dmg > pv.y) {
        pv.y = dmg;
        L.mode[o] = md;
    }
    stStream(&L.pv[o], pv);
    if (!L.tensile_only) stStream(&L.pyz[o], pyz);
    const float event_damage = float(pv.y);
    bool event = double(event_damage) > m0 + 1.0e-6;
    R event_energy = R(0);
    const bool broke = pv.y >= R(1);
    if (broke) {
        if (L.tensile_only) {  // the samples a dead bond keeps, with the exact shear
            R t2, c2, s2;
            sampleBondWith<R, DIRECT>(L, r.a, r.b, r.p, g, ua, ub, st, t2, c2, s2);
            L.pyz[o] = R2<R>{maxR(R(0), c2), maxR(R(0), s2)};
        }
        const R ext = length(bondVec<R, DIRECT>(L, r.a, r.b, g.e, ua, ub)) - restL - plastic;
        const R cc = L.cmpl[j];
        event_energy = cc <= R(0) ? R(0) : R(0.5) * ext * ext / cc;
        L.alive[j] = 0;
        L.alive_n[o] = 0;
        const uint2 sl = L.slots[o];
        L.nbr_key[sl.x] |= kDeadSlot;
        L.nbr_key[sl.y] |= kDeadSlot;
        claim_a = atomicExch(&L.dirty_stamp[r.a], stamp) != stamp ? r.a : kNone;
        claim_b = atomicExch(&L.dirty_stamp[r.b], stamp) != stamp ? r.b : kNone;
        event = true;
    }
    if (event) {
        L.ev_d[j] = event_damage;
        L.ev_e[j] = event_energy;
        const unsigned long long bit = 1ULL << (j & 63U);
        atomicOr((unsigned long long *)&L.ev_bits[j >> 6], bit);
        if (broke) atomicOr((unsigned long long *)&L.brk_bits[j >> 6], bit);
        atomicOr((unsigned long long *)&L.grp_bits[j >> 12], 1ULL << ((j >> 6) & 63U));
    }
}

// Appends the wave's dirty-node claims to dirty_list (one atomic per wave).
__device__ __forceinline__ void appendDirty(u32 *list, u32 *count, u32 claim_a, u32 claim_b) {
    const u32 lane = threadIdx.x & 63U;
    const u32 mine = (claim_a != kNone) + (claim_b != kNone);
    if (!__ballot(mine != 0)) return;  // the common case: no break in this wave
    u32 x = mine;
    for (u32 q = 1; q < 64; q <<= 1) {
        const u32 y = __shfl_up(x, q);
        if (lane >= q) x += y;
    }
    const u32 wave_total = __shfl(x, 63);
    u32 base = 0;
    if (lane == 63 && wave_total) base = atomicAdd(count, wave_total);
    base = __shfl(base, 63);
    u32 at = base + x - mine;
    if (claim_a != kNone) list[at++] = claim_a;
    if (claim_b != kNone) list[at] = claim_b;
}

// The largest positive value of a wave, as its bits (0 when there is none):
// positive floats order as their bit patterns, so the wave folds unsigned
// maxima by DPP row shifts and row broadcasts into lane 63 (every lane active).
__device__ __forceinline__ u32 wavePeakBits(float x) {
    u32 v = x > 0.f ? __float_as_uint(x) : 0U;
    v = max(v, u32(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xf, 0xf, false)));  // row_shr:1
    v = max(v, u32(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xf, 0xf, false)));  // row_shr:2
    v = max(v, u32(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xf, 0xf, false)));  // row_shr:4
    v = max(v, u32(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xf, 0xf, false)));  // row_shr:8
    v = max(v, u32(__builtin_amdgcn_update_dpp(0, int(v), 0x142, 0xa, 0xf, false)));  // row_bcast:15
    v = max(v, u32(__builtin_amdgcn_update_dpp(0, int(v), 0x143, 0xc, 0xf, false)));  // row_bcast:31
    return u32(__builtin_amdgcn_readlane(int(v), 63));
}

// Running maxima of non-negative floats: integer atomicMax on the bits of
// the workgroup's largest positive value. Every thread calls it. A caller
// whose tensile (TENSILE false) or plastic (PLASTIC false) values are all 0
// skips that fold.
template <bool TENSILE, bool PLASTIC>
__device__ __forceinline__ void publishMaxima(DevCtl *ctl, float mt, float mc, float ms, float mp, u32 (*s_max)[kWG / 64]) {
    const u32 lane = threadIdx.x & 63U, wave = threadIdx.x >> 6;
    const u32 bt = TENSILE ? wavePeakBits(mt) : 0U, bc = wavePeakBits(mc), bs = wavePeakBits(ms),
              bp = PLASTIC ? wavePeakBits(mp) : 0U;
    if (lane == 0) {
        s_max[0][wave] = bt;
        s_max[1][wave] = bc;
        s_max[2][wave] = bs;
        s_max[3][wave] = bp;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        u32 m = 0;
        for (int w = 0; w < kWG / 64; ++w) m = max(m, s_max[threadIdx.x][w]);
        u32 *dst = threadIdx.x < 3 ? &ctl->peak_bits[threadIdx.x] : &ctl->pstretch_bits;
        // The running maxima rarely move: read first, so thousands of
        // workgroups do not queue atomics on the same four words. The read goes
        // through the caches: a maximum only grows, so a stale value is never
        // above the current one and never skips a needed update.
        if (m > 0U && m > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT)) atomicMax(dst, m);
    }
}

template <class R, bool DIRECT, bool PLASTIC, int COVER>
__global__ __launch_bounds__(kWG) void k_bond_end(Batch<R> B, Step<R> S) {
    static_assert(!PLASTIC || COVER == kBondAll, "plastic work is summed per end-sample tile in a fixed order");
    const WgMap *m = COVER == kBondAffected ? &B.map[blockIdx.x] : tileEntry(B);
    if (!m) return;
    const DevLattice<R> L = B.L[m->lat];
    DevCtl *ctl = L.ctl;
    // The span's streamed loads go out with the run state's, before the stop test.
    EndLoad<R> in[kEndSpan];
    if (COVER != kBondAffected) {
#pragma unroll
        for (u32 q = 0; q < kEndSpan; ++q) {
            const u32 o = (m->a + q) * kWG + threadIdx.x;
            in[q] = q < m->b && o < L.B ? loadEnd(L, o) : EndLoad<R>{};
        }
    }
    if (ctl->stop) return;
    __shared__ u32 s_max[4][kWG / 64];
    __shared__ double s_pw[kWG];
    const double m0 = ctl->max_damage;
    const u32 stamp = u32(ctl->step) + 1U;
    const u32 half = u32(ctl->step & 1U);
    // A shear below the running maximum read here (which only grows) cannot
    // raise it; the margin keeps the rounded square root at or below it.
    const R shear_max = R(__uint_as_float(ctl->peak_bits[2]));
    const R shear_floor2 = shear_max * shear_max * R(1.0 - 1.0 / 1048576.0);
    if (COVER != kBondAffected && m->a == 0 && threadIdx.x == 0) {
        ctl->rank_slot = u32(ctl->step & 1U);
        ctl->rinv_stamp = stamp;
    }
    float mt = 0.f, mc = 0.f, ms = 0.f, mp = 0.f;
    if (COVER == kBondAffected) {
        // Bonds with an affected end, each taken once: from its first end when
        // that end is affected, else from its second.
        // Node slots D * N are below 2^32 (refused at upload otherwise).
        const u32 pairs = ctl->affected_count * L.D;
        for (u32 base = m->a * kWG; base < pairs; base += m->b * kWG) {
            const u32 q = base + threadIdx.x;
            u32 claim_a = kNone, claim_b = kNone;
            double pw = 0.0;
            if (q < pairs) {
                const u32 i = L.affected_list[q / L.D], k = q % L.D;
                const u32 o = L.nbr_bond[size_t(k) * L.N + i];
                if (o != kNone) {
                    const u32 a = L.bnd[o].a;
                    if (a == i || L.affected[a] != stamp)
                        endSampleBond<R, DIRECT, PLASTIC, COVER>(L, S, o, loadEnd(L, o), m0, stamp, shear_floor2, mt,
                                                                 mc, ms, mp, pw, claim_a, claim_b);
                }
            }
            appendDirty(L.dirty_list + size_t(half) * L.N, &ctl->dirty_count[half], claim_a, claim_b);
        }
        publishMaxima<true, PLASTIC>(ctl, mt, mc, ms, mp, s_max);
        return;
    }
    // A span of up to kEndSpan consecutive tiles of kWG bonds (m->a, m->b).
#pragma unroll
    for (u32 q = 0; q < kEndSpan; ++q) {
        if (q >= m->b) break;
        const u32 tile = m->a + q, o = tile * kWG + threadIdx.x;
        u32 claim_a = kNone, claim_b = kNone;
        double pw = 0.0;
        if (o < L.B)
            endSampleBond<R, DIRECT, PLASTIC, COVER>(L, S, o, in[q], m0, stamp, shear_floor2, mt, mc, ms, mp, pw,
                                                     claim_a, claim_b);
        appendDirty(L.dirty_list + size_t(half) * L.N, &ctl->dirty_count[half], claim_a, claim_b);
        if (PLASTIC) {
            s_pw[threadIdx.x] = pw;
            __syncthreads();
            for (u32 w = kWG / 2; w > 0; w >>= 1) {
                if (threadIdx.x < w) s_pw[threadIdx.x] = s_pw[threadIdx.x] + s_pw[threadIdx.x + w];
                __syncthreads();
            }
            if (threadIdx.x == 0) L.tile_plastic[tile] = s_pw[0];
            __syncthreads();
        }
    }
    publishMaxima<true, PLASTIC>(ctl, mt, mc, ms, mp, s_max);
}

// ---------------------------------------------------------------------------
// The substep's bookkeeping (CpuLatticeBackend.cpp:441-488 and 279-298).

__device__ __forceinline__ unsigned exitReason(u64 completed, unsigned broken, u64 last_failure, u64 quiet, u64 min_steps,
                                               u64 no_failure, u64 energy_flat, u64 last_energy_gain, u64 calm,
                                               u64 last_damage_gain, double max_damage, double calm_margin,
                                               double removed, double removable) {
    // LatticePhysics.hpp:1766-1791.
    if (removable > 0.0 && removed >= removable) return 6U;
    if (broken > 0U && quiet > 0 && completed >= min_steps && completed > last_failure && completed - last_failure > quiet)
        return 1U;
    if (broken == 0U && no_failure > 0 && completed >= no_failure) return 2U;
    if (broken == 0U && calm > 0 && completed >= calm && max_damage < calm_margin && completed > last_damage_gain &&
        completed - last_damage_gain > calm)
        return 5U;
    if (broken > 0U && energy_flat > 0 && completed >= min_steps && completed > last_energy_gain &&
        completed - last_energy_gain > energy_flat)
        return 4U;
    return 0U;
}

template <class R> __device__ void sphereSupport(const Step<R> &S, Sphere<R> &sp, u64 &events) {
    // LatticePhysics.hpp:1797-1825.
    for (u32 p = 0; p < S.plane_count; ++p) {
        const Plane<R> &pl = S.planes[p];
        const V3<R> foot = sp.center - sp.radius * pl.normal;
        if (!insideFootprints(pl, foot)) continue;
        const R dist = dot(foot - pl.point, pl.normal);
        if (dist > R(0)) continue;
        const R vn0 = dot(sp.velocity, pl.normal);
        if (-dist > R(8) * maxR(R(0), -vn0) * S.dt + R(1.0e-5)) continue;
        const R vn = dot(sp.velocity, pl.normal);
        if (vn < R(0)) {
            const R rest = -vn > R(0.5) ? pl.restitution : R(0);
            const R nd = -(R(1) + rest) * vn;
            sp.velocity = sp.velocity + nd * pl.normal;
            const V3<R> tan = sp.velocity - dot(sp.velocity, pl.normal) * pl.normal;
            const R speed = length(tan);
            if (speed > R(1.0e-12)) {
                const R fd = speed <= pl.static_friction * nd ? speed : minR(speed, pl.dynamic_friction * nd);
                sp.velocity = sp.velocity - (fd / speed) * tan;
            }
            ++events;
        }
        sp.center = sp.center - dist * pl.normal;
    }
}

template <class R>
__device__ double fixedOrderSum(const double *part, u32 count, u32 stride, u32 offset, double *sh) {
    const u32 tid = threadIdx.x;
    double s = 0.0;
    for (u32 t = tid; t < count; t += kFinishThreads) s += part[size_t(t) * stride + offset];
    sh[tid] = s;
    __syncthreads();
    for (u32 w = kFinishThreads / 2; w > 0; w >>= 1) {
        if (tid < w) sh[tid] = sh[tid] + sh[tid + w];
        __syncthreads();
    }
    const double total = sh[0];
    __syncthreads();
    return total;
}

constexpr u32 kEventBatch = 4096;
constexpr u32 kGroupsPerThread = 4;  // summary words a finish thread reads at once
constexpr u32 kFlagsInLds = 2048;    // flagged event words kept in LDS
constexpr u32 kEnergyAhead = 24;     // removed energies the finish's sum reads ahead
constexpr u32 kWordsPerThread = 2;    // flagged event words a finish thread stages per batch
constexpr u32 kBatchWords = kFinishThreads * kWordsPerThread;
constexpr unsigned kBudgetSpent = 3U;  // the reference's max-steps exit reason

// Position of the k-th (from 0) set bit of m; k < popcount(m).
__device__ __forceinline__ u32 selectBit(u64 m, u32 k) {
    u32 pos = 0;
    for (u32 half = 32; half > 0; half >>= 1) {
        const u32 c = u32(__popcll(m & ((1ULL << half) - 1ULL)));
        if (k >= c) {
            k -= c;
            m >>= half;
            pos += half;
        }
    }
    return pos;
}

template <class R, bool LOADS, bool PLASTIC>
__global__ __launch_bounds__(kFinishThreads) void k_finish(Batch<R> B, Step<R> S) {
    const DevLattice<R> L = B.L[B.map[blockIdx.x].lat];
    DevCtl *ctl = L.ctl;
    if (ctl->stop) return;
    __shared__ double s_red[kFinishThreads];
    __shared__ u32 s_waves[kFinishThreads / 64];
    __shared__ u32 s_word[kBatchWords], s_incl[kBatchWords];
    __shared__ u64 s_evw[kBatchWords], s_bkw[kBatchWords];
    // A staged batch's bond ids and damages; once wave 0 has read them, the
    // same bytes hold the running removed-energy sums of its broken bonds.
    __shared__ double s_jd[kEventBatch];
    u32 *const s_j = reinterpret_cast<u32 *>(s_jd);
    float *const s_d = reinterpret_cast<float *>(s_jd) + kEventBatch;
    double *const s_sum = s_jd;
    // Removed energies, widened as the ledger sums them; the tail pads the
    // energy sum's reads ahead (kEnergyAhead).
    __shared__ double s_e[kEventBatch + kEnergyAhead];
    __shared__ u32 s_take;
    __shared__ u32 s_flags[kFlagsInLds];
    const u32 tid = threadIdx.x, lane = tid & 63U;
    // The control block and the sphere as the substep left them, copied to LDS
    // while the events are folded: the serial tail then works on the copy
    // instead of a chain of dependent loads. k_rinv, which runs beside this
    // kernel, adds to the current substep's rank_pending slot and touches
    // nothing else here, so that word alone is not written back.
    __shared__ DevCtl s_c;
    __shared__ Sphere<R> s_sp;
    {
        const u32 *src = reinterpret_cast<const u32 *>(ctl);
        u32 *dst = reinterpret_cast<u32 *>(&s_c);
        for (u32 w = tid; w < u32(sizeof(DevCtl) / 4); w += kFinishThreads) dst[w] = src[w];
        const u32 *ssrc = reinterpret_cast<const u32 *>(L.sphere);
        u32 *sdst = reinterpret_cast<u32 *>(&s_sp);
        for (u32 w = tid; w < u32(sizeof(Sphere<R>) / 4); w += kFinishThreads) sdst[w] = ssrc[w];
    }
    const u64 step = ctl->step;
    // Wave 0 carries the folds: its lanes all hold the same maximum damage and
    // counts; lane 0 holds the removed-energy sum.
    double m = ctl->max_damage, removed = ctl->removed_energy_j;
    bool damage_gain = false, energy_gain = false;
    u32 broken_now = 0;
    const bool first_round = ctl->broken_bonds == 0;
    u64 ff = ctl->first_failure_count;
    const double flat = ctl->flat_fraction;
    for (u32 gbase = 0; gbase < L.group_words; gbase += kFinishThreads * kGroupsPerThread) {
        // Event words flagged in these summary words, in order, into the flag
        // list: in LDS when they fit (the usual case), else in flag_list. A
        // thread reads its consecutive summary words in one round trip.
        u64 bits[kGroupsPerThread];
        u32 cnt = 0;
#pragma unroll
        for (u32 q = 0; q < kGroupsPerThread; ++q) {
            const u32 g = gbase + tid * kGroupsPerThread + q;
            bits[q] = g < L.group_words ? L.grp_bits[g] : 0ULL;
        }
#pragma unroll
        for (u32 q = 0; q < kGroupsPerThread; ++q) {
            if (bits[q]) L.grp_bits[gbase + tid * kGroupsPerThread + q] = 0;
            cnt += u32(__popcll(bits[q]));
        }
        u32 nflag = 0;
        u32 at = blockScan1024(cnt, s_waves, &nflag);
        u32 *flags = nflag <= kFlagsInLds ? s_flags : L.flag_list;
#pragma unroll
        for (u32 q = 0; q < kGroupsPerThread; ++q) {
            const u32 g = gbase + tid * kGroupsPerThread + q;
            while (bits[q]) {
                flags[at++] = g * 64U + u32(__ffsll((long long)bits[q]) - 1);
                bits[q] &= bits[q] - 1;
            }
        }
        __syncthreads();
        for (u32 fb = 0; fb < nflag;) {
            // As many whole event words as fit in the event batch; a word holds
            // at most 64 events, so at least one always fits. A thread takes
            // kWordsPerThread consecutive flagged words.
            const u32 mcount = min(kBatchWords, nflag - fb);
            u32 wd[kWordsPerThread], incl[kWordsPerThread], c = 0;
            u64 evw[kWordsPerThread], bkw[kWordsPerThread];
#pragma unroll
            for (u32 h = 0; h < kWordsPerThread; ++h) {
                const u32 k = tid * kWordsPerThread + h;
                wd[h] = k < mcount ? flags[fb + k] : 0U;
                evw[h] = k < mcount ? L.ev_bits[wd[h]] : 0ULL;
                bkw[h] = k < mcount ? L.brk_bits[wd[h]] : 0ULL;
            }
#pragma unroll
            for (u32 h = 0; h < kWordsPerThread; ++h) c += u32(__popcll(evw[h]));
            u32 all = 0;
            u32 run = blockScan1024(c, s_waves, &all);
#pragma unroll
            for (u32 h = 0; h < kWordsPerThread; ++h) {
                const u32 k = tid * kWordsPerThread + h;
                run += u32(__popcll(evw[h]));
                incl[h] = run;
                s_word[k] = wd[h];
                s_incl[k] = run;
                s_evw[k] = evw[h];
                s_bkw[k] = bkw[h];
            }
            __syncthreads();
#pragma unroll
            for (u32 h = 0; h < kWordsPerThread; ++h) {
                const u32 k = tid * kWordsPerThread + h;
                if (k < mcount && incl[h] <= kEventBatch && (k + 1 == mcount || s_incl[k + 1] > kEventBatch))
                    s_take = k + 1;
            }
            __syncthreads();
            const u32 take = s_take;
#pragma unroll
            for (u32 h = 0; h < kWordsPerThread; ++h) {
                if (tid * kWordsPerThread + h < take) {
                    L.ev_bits[wd[h]] = 0;
                    L.brk_bits[wd[h]] = 0;
                }
            }
            const u32 nev = s_incl[take - 1];
            for (u32 e = tid; e < nev; e += kFinishThreads) {
                u32 lo = 0, hi = take;
                while (lo < hi) {
                    const u32 mid = (lo + hi) >> 1;
                    if (s_incl[mid] <= e) lo = mid + 1;
                    else hi = mid;
                }
                const u32 pos = selectBit(s_evw[lo], e - (lo ? s_incl[lo - 1] : 0U));
                const u32 j = s_word[lo] * 64U + pos;
                s_j[e] = j | (((s_bkw[lo] >> pos) & 1ULL) ? 0x80000000U : 0U);
                s_d[e] = L.ev_d[j];
                s_e[e] = double(L.ev_e[j]);
            }
            __syncthreads();
            // Wave 0 folds the staged batch in bond order, 64 events at a time;
            // the sequential steps broadcast one lane's values with readlane.
            if (tid < 64) {
                u32 nb = 0;  // broken bonds of this batch, their energies compacted into s_e[0, nb)
                for (u32 c0 = 0; c0 < nev; c0 += 64) {
                    const u32 e = c0 + lane;
                    const bool ok = e < nev;
                    const u32 jr = ok ? s_j[e] : 0U;
                    const float dmg = ok ? s_d[e] : 0.f;
                    const double en = ok ? s_e[e] : 0.0;
                    // max_damage rises only by more than 1e-6 (CpuLatticeBackend.cpp:462-465).
                    u64 cand = __ballot(ok && double(dmg) > m + 1.0e-6);
                    while (cand) {
                        const u32 l = u32(__ffsll((long long)cand) - 1);
                        m = double(laneValue(dmg, l));
                        damage_gain = true;
                        cand = __ballot(ok && lane > l && double(dmg) > m + 1.0e-6);
                    }
                    // Broken bonds: first-failure list and counts in bond order
                    // (CpuLatticeBackend.cpp:466-477).
                    const u64 bmask = __ballot(ok && (jr >> 31));
                    if (first_round && ok && (jr >> 31))
                        L.first_fail[ff + __popcll(bmask & ((1ULL << lane) - 1ULL))] = jr & 0x7fffffffU;
                    ff += __popcll(bmask);
                    broken_now += u32(__popcll(bmask));
                    // In place: every lane has read its entry, and the write
                    // index never passes the read index.
                    if (ok && (jr >> 31)) s_e[nb + __popcll(bmask & ((1ULL << lane) - 1ULL))] = en;
                    nb += u32(__popcll(bmask));
                }
                // Removed energy summed in bond order by one lane: the only
                // dependence from one bond to the next is the addition, so the
                // lane keeps only the additions and stores each running sum
                // (over the ids and damages, which every lane has read); the
                // energy-gain test on each sum then runs on all 64 lanes.
                // The lane works on three register groups of energies in turn:
                // it adds one group and stores its sums while the reads of the
                // next two are in flight, so no addition waits on LDS (the
                // empty asm statements keep each group's reads ahead of the
                // next group's additions). Reads past nb land in the padding
                // and are not used.
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                if (lane == 0) {
                    constexpr u32 G = kEnergyAhead / 3;
                    double sum = removed, g[3][G];
#pragma unroll
                    for (u32 h = 0; h < 3; ++h)
#pragma unroll
                        for (u32 q = 0; q < G; ++q) g[h][q] = s_e[h * G + q];
                    u32 k = 0;
                    for (; k + 3 * G <= nb; k += 3 * G) {
#pragma unroll
                        for (u32 h = 0; h < 3; ++h) {
#pragma unroll
                            for (u32 q = 0; q < G; ++q) {
                                sum += g[h][q];
                                s_sum[k + h * G + q] = sum;
                            }
#pragma unroll
                            for (u32 q = 0; q < G; ++q) g[h][q] = s_e[k + (h + 3) * G + q];
                            asm volatile("" ::: "memory");
                        }
                    }
                    for (; k < nb; ++k) {
                        sum += s_e[k];
                        s_sum[k] = sum;
                    }
                    removed = sum;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                bool gain = false;
                for (u32 k = lane; k < nb; k += 64) {
                    const double el = s_e[k], sum = s_sum[k];
                    gain |= el > flat * (sum > 1.0e-12 ? sum : 1.0e-12);
                }
                energy_gain |= __ballot(gain) != 0ULL;
            }
            __syncthreads();
            fb += take;
        }
    }
    double ledger[13];
    const bool loads = LOADS && ctl->load_remaining > 0;
    if (LOADS && loads) {
        // The 13 sums over the node tiles in one pass on wave 0: each lane adds
        // its tiles in order, then waveSumLast's fixed order.
        __shared__ double s_led[13];
        if (tid < 64) {
            double v[13] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (u32 t = tid; t < L.node_tiles; t += 64)
                for (int k = 0; k < 13; ++k) v[k] += L.load_part[size_t(t) * 13 + k];
            for (int k = 0; k < 13; ++k) v[k] = waveSumLast(v[k]);
            if (tid == 63)
                for (int k = 0; k < 13; ++k) s_led[k] = v[k];
        }
        __syncthreads();
        for (int k = 0; k < 13; ++k) ledger[k] = s_led[k];
    }
    double plastic = 0.0;
    if (PLASTIC) plastic = fixedOrderSum<R>(L.tile_plastic, L.bond_tiles, 1, 0, s_red);
    // The serial tail, on the LDS copy; then every thread writes the copy back.
    __syncthreads();
    if (tid == 0) {
        {
            const u32 slot = u32((step + 1U) & 1U);  // (step - 1) & 1: the slot the previous substep's k_rinv used
            s_c.rank_deficient += s_c.rank_pending[slot];
            s_c.rank_pending[slot] = 0;
        }
        if (damage_gain) {
            s_c.max_damage = m;
            s_c.last_damage_gain_step = step;
        }
        const bool any = broken_now > 0;
        if (any) {
            s_c.removed_energy_j = removed;
            if (energy_gain) s_c.last_energy_gain_step = step;
            s_c.broken_bonds += broken_now;
            if (first_round) s_c.first_failure_count = ff;
        }
        if (PLASTIC) s_c.plastic_work += plastic;
        if (LOADS && loads) {
            s_c.ledger_steps += 1;
            s_c.ledger_elapsed += double(S.dt);
            for (int k = 0; k < 3; ++k) {
                s_c.ledger_req[k] += ledger[k];
                s_c.ledger_imp[k] += ledger[3 + k];
                s_c.ledger_areq[k] += ledger[6 + k];
                s_c.ledger_aimp[k] += ledger[9 + k];
            }
            s_c.ledger_work += ledger[12];
            s_c.load_remaining -= 1;
        }
        // Sphere drift and sphere against the supports (CpuLatticeBackend.cpp:481-482).
        Sphere<R> sp = s_sp;
        if (!S.sphere_enabled) sp.velocity = sp.velocity + S.dt * S.gravity;
        sp.center = sp.center + S.dt * sp.velocity;
        u64 ev = s_c.ball_support_events;
        sphereSupport(S, sp, ev);
        s_c.ball_support_events = ev;
        s_sp = sp;
        if (any) {
            s_c.dirty_start[step & 1U] = 1;
            s_c.nc_rebuild = 1;  // a failure round forces the contact pair list's rebuild (CpuLatticeBackend.cpp:485)
        }
        // Run loop bookkeeping (CpuLatticeBackend.cpp:279-298).
        const u64 total = step + 1;
        s_c.step = total;
        if (any) {
            if (s_c.first_failure_step == ~0ULL) s_c.first_failure_step = step;
            s_c.last_failure_step = step;
            s_c.failure_rounds += 1;
        }
        const double available = s_c.removable + s_c.ledger_work;
        const unsigned reason = s_c.removable > 0 && available <= 0
            ? 6U
            : exitReason(total, s_c.broken_bonds, s_c.last_failure_step, s_c.quiet, s_c.min_steps, s_c.no_failure,
                         s_c.energy_flat, s_c.last_energy_gain_step, s_c.calm, s_c.last_damage_gain_step, s_c.max_damage,
                         s_c.calm_margin, s_c.removed_energy_j, s_c.removable > 0 ? available : 0.0);
        s_c.exit_reason = reason;
        if (reason != 0 || s_c.error) s_c.stop = 1;
        // The run's step budget (RunControl::max_steps, reason 3).
        const u64 budget = s_c.budget > 0 ? s_c.budget - 1 : 0;
        s_c.budget = budget;
        if (reason == 0 && budget == 0) {
            s_c.exit_reason = kBudgetSpent;
            s_c.stop = 1;
        }
    }
    __syncthreads();
    {
        const u32 *src = reinterpret_cast<const u32 *>(&s_c);
        u32 *dst = reinterpret_cast<u32 *>(ctl);
        const u32 busy = u32(offsetof(DevCtl, rank_pending) / 4) + u32(step & 1U);
        for (u32 w = tid; w < u32(sizeof(DevCtl) / 4); w += kFinishThreads)
            if (w != busy) dst[w] = src[w];
        const u32 *ssrc = reinterpret_cast<const u32 *>(&s_sp);
        u32 *sdst = reinterpret_cast<u32 *>(L.sphere);
        for (u32 w = tid; w < u32(sizeof(Sphere<R>) / 4); w += kFinishThreads) sdst[w] = ssrc[w];
    }
}

// ---------------------------------------------------------------------------
// Fragments: integer union-find over the live bonds, then the canonical
// relabel of findConnectedComponents (ConnectedComponents.cpp:77-85). Roots are
// linked larger index under smaller with a compare-and-swap, so every root
// ends as its component's smallest node: the tie-break of the canonical order.
//
// Finds read parents through the caches. A parent only ever moves to a smaller
// ancestor and trees only merge, so a stale parent is still an ancestor and a
// find over cached values ends at an ancestor of its node, never past the
// root. Only the link needs the current value of a root's entry, and the
// compare-and-swap returns it: a failed link climbs from that value, so the
// larger end strictly decreases and every link attempt ends.

__device__ __forceinline__ u32 cachedParent(u32 *parent, u32 x) {
    return __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

__device__ __forceinline__ u32 findRoot(u32 *parent, u32 x) {
    u32 p = cachedParent(parent, x);
    while (p != x) {
        const u32 g = cachedParent(parent, p);
        // Path halving: g is an ancestor of x, so the store only shortens the path.
        if (g != p) __hip_atomic_store(&parent[x], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
        p = g;
    }
    return x;
}

// Joins the trees of a and b. A root is linked under any node of a smaller
// index in the other tree, not only under its root: the merged tree's root is
// then still its smallest node, and no cycle can form, since a root is the
// smallest node of its own tree. Linking a root under the other end itself
// saves the find of that end whenever it has the smaller index.
__device__ __forceinline__ void unite(u32 *parent, u32 a, u32 b) {
    a = findRoot(parent, a);
    for (;;) {
        if (a == b) return;
        if (b < a) {
            const u32 old = atomicCAS(&parent[a], a, b);
            if (old == a) return;
            a = findRoot(parent, old);
            continue;
        }
        const u32 rb = findRoot(parent, b);
        if (rb == a) return;
        if (rb < a) {
            b = rb;
            continue;
        }
        const u32 old = atomicCAS(&parent[rb], rb, a);
        if (old == rb) return;
        b = old;
    }
}

// Linking is two-level. A workgroup owns a range of kCcRange consecutive
// nodes and joins the live bonds inside it with a union-find in LDS; every
// node's parent becomes its range-local root. A second pass reads the bonds
// that leave the range (from their smaller end), reduces them to the distinct
// pairs of local roots in an LDS hash set, and links only those pairs in
// global memory. Trees keep the smallest node as root at both levels.
constexpr u32 kCcRange = 1024;  // nodes per range: one per thread of a 1024-lane workgroup
constexpr u32 kCcPairs = 1024;
constexpr u32 kCcBatch = 8;     // slot keys loaded together before they are used
constexpr u64 kCcEmpty = ~0ULL;

__device__ __forceinline__ u32 ldsParent(u32 *lp, u32 x) {
    return __hip_atomic_load(&lp[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ u32 ldsFind(u32 *lp, u32 x) {
    u32 p = ldsParent(lp, x);
    while (p != x) {
        const u32 g = ldsParent(lp, p);
        if (g != p) __hip_atomic_store(&lp[x], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        x = p;
        p = g;
    }
    return x;
}
// The same linking rule as unite(), on range-local indices in LDS.
__device__ __forceinline__ void ldsUnite(u32 *lp, u32 a, u32 b) {
    a = ldsFind(lp, a);
    for (;;) {
        if (a == b) return;
        if (b < a) {
            u32 expect = a;
            if (__hip_atomic_compare_exchange_strong(&lp[a], &expect, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP))
                return;
            a = ldsFind(lp, expect);
            continue;
        }
        const u32 rb = ldsFind(lp, b);
        if (rb == a) return;
        if (rb < a) {
            b = rb;
            continue;
        }
        u32 expect = rb;
        if (__hip_atomic_compare_exchange_strong(&lp[rb], &expect, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP))
            return;
        b = expect;
    }
}

// Range-local union-find: parent[i] = the smallest node of i's component
// within its range; sizes are cleared for the final count.
template <class R> __global__ __launch_bounds__(kCcRange) void k_cc_local(DevLattice<R> L, u32 *parent, u32 *size) {
    __shared__ u32 lp[kCcRange];
    const u32 base = blockIdx.x * kCcRange, k = threadIdx.x;
    const u32 n = min(kCcRange, L.N - base), N = L.N, i = base + k;
    const bool in = k < n;
    lp[k] = k;
    __syncthreads();
    for (u32 q0 = 0; q0 < L.D; q0 += kCcBatch) {
        u32 key[kCcBatch];
#pragma unroll
        for (u32 q = 0; q < kCcBatch; ++q) key[q] = in && q0 + q < L.D ? L.nbr_key[size_t(q0 + q) * N + i] : kNone;
#pragma unroll
        for (u32 q = 0; q < kCcBatch; ++q) {
            if (key[q] & kDeadSlot) continue;  // dead, absent, or out of range
            const u32 j = u32(int(i) + L.sdelta[key[q]]);
            if (j > i && j < base + n) ldsUnite(lp, k, j - base);
        }
    }
    __syncthreads();
    if (in) {
        parent[i] = base + ldsFind(lp, k);
        size[i] = 0;
    }
}

// Bonds that leave a range, from their smaller end: the distinct pairs of
// local roots are gathered in LDS, then each is linked once. A pair that finds
// the hash set full is linked at once instead.
template <class R> __global__ __launch_bounds__(kCcRange) void k_cc_cross(DevLattice<R> L, u32 *parent) {
    __shared__ u64 pairs[kCcPairs];
    const u32 base = blockIdx.x * kCcRange, k = threadIdx.x;
    const u32 n = min(kCcRange, L.N - base), N = L.N, i = base + k;
    const bool in = k < n;
    for (u32 e = k; e < kCcPairs; e += kCcRange) pairs[e] = kCcEmpty;
    __syncthreads();
    const u32 ri = in ? cachedParent(parent, i) : 0U;
    u64 last = kCcEmpty;
    for (u32 q0 = 0; q0 < L.D; q0 += kCcBatch) {
        u32 other[kCcBatch];
#pragma unroll
        for (u32 q = 0; q < kCcBatch; ++q) {
            const u32 key = in && q0 + q < L.D ? L.nbr_key[size_t(q0 + q) * N + i] : kNone;
            // Inside the range, or the bond's smaller end is the other node: not this pass's bond.
            const bool live = !(key & kDeadSlot);
            const u32 j = live ? u32(int(i) + L.sdelta[key]) : 0U;
            other[q] = live && j >= base + n ? j : kNone;
        }
#pragma unroll
        for (u32 q = 0; q < kCcBatch; ++q) other[q] = other[q] != kNone ? cachedParent(parent, other[q]) : kNone;
#pragma unroll
        for (u32 q = 0; q < kCcBatch; ++q) {
            if (other[q] == kNone) continue;
            const u64 pr = (u64(ri) << 32) | u64(other[q]);
            if (pr == last) continue;
            last = pr;
            u32 h = (ri * 0x9E3779B1U ^ other[q] * 0x85EBCA77U) & (kCcPairs - 1);
            bool placed = false;
            for (u32 probe = 0; probe < 32 && !placed; ++probe, h = (h + 1) & (kCcPairs - 1)) {
                const u64 seen = atomicCAS(reinterpret_cast<unsigned long long *>(&pairs[h]), kCcEmpty, pr);
                placed = seen == kCcEmpty || seen == pr;
            }
            if (!placed) unite(parent, ri, other[q]);
        }
    }
    __syncthreads();
    for (u32 e = k; e < kCcPairs; e += kCcRange) {
        const u64 pr = pairs[e];
        if (pr != kCcEmpty) unite(parent, u32(pr >> 32), u32(pr));
    }
}

// Flattens every tree to depth one (into parent[], or into root[] for the
// final labels) and, with COUNT, counts the component sizes. No link runs in
// this launch, so the finds end at the true roots. A wave counts its lanes
// once per distinct root, and the counts of the root of the workgroup's first
// node gather in LDS and land with one atomic per workgroup: the largest
// component is the root of nearly every node, and per-wave atomics on its one
// counter would serialize.
template <bool COUNT>
__global__ __launch_bounds__(kWG) void k_cc_flatten(u32 *parent, u32 *out, u32 *size, u32 n) {
    __shared__ u32 s_common, s_sum;
    const u32 stride = gridDim.x * kWG;
    u32 common = kNone, mine = 0;
    for (u32 base = blockIdx.x * kWG; base < n; base += stride) {
        const u32 i = base + threadIdx.x;
        u32 r = kNone;
        if (i < n) {
            r = findRoot(parent, i);
            __hip_atomic_store(&out[i], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (!COUNT) continue;
        if (common == kNone) {
            if (threadIdx.x == 0) {
                s_common = r;
                s_sum = 0;
            }
            __syncthreads();
            common = s_common;
        }
        mine += r == common;
        u64 pending = __ballot(r != common && r != kNone);
        while (pending) {
            const u32 leader = u32(__ffsll((long long)pending) - 1);
            const u32 lr = u32(__builtin_amdgcn_readlane(int(r), int(leader)));
            const u64 same = __ballot(r == lr) & pending;
            if ((threadIdx.x & 63U) == leader) atomicAdd(&size[lr], u32(__popcll(same)));
            pending &= ~same;
        }
    }
    if (!COUNT || common == kNone) return;
    for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o);
    if ((threadIdx.x & 63U) == 0) atomicAdd(&s_sum, mine);
    __syncthreads();
    if (threadIdx.x == 0 && s_sum) atomicAdd(&size[common], s_sum);
}

// One sort key per component: (n - size) << 32 | root. Ascending keys are the
// canonical order, size descending then smallest node; the append order of
// the keys does not matter because the sort is on the whole key.
__global__ void k_cc_keys(const u32 *root, const u32 *size, u64 *keys, u32 *count, u32 n) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (root[i] == i) keys[atomicAdd(count, 1U)] = (u64(n - size[i]) << 32) | u64(i);
}

// Clears the component counter with a kernel rather than a memset, so the
// label launches hold no memset node when a caller captures them in a graph.
__global__ void k_cc_zero(u32 *count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0U;
}

__global__ void k_cc_rank(const u64 *sorted, u32 *rank, u32 c) {
    for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < c; k += gridDim.x * blockDim.x)
        rank[u32(sorted[k] & 0xffffffffULL)] = k;
}

__global__ void k_cc_label(const u32 *root, const u32 *rank, u32 *label, u32 n) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) label[i] = rank[root[i]];
}

// ---------------------------------------------------------------------------
// Node-node contact (LatticePhysics.hpp:1430-1667, CpuLatticeBackend.cpp:424-437).
//
// Broad phase. The reference hashes cells into buckets and scans, for node i,
// the 27 cells around it in (dz, dy, dx) order, taking in each cell the nodes
// j > i in ascending index order. Here the nodes are sorted by cell key with
// a stable sort (so ascending index within a cell), a table maps each
// occupied cell to its run in that order, and the per-node gather walks the
// same 27 cells: the emitted pair order is the reference's. The host runs a
// rebuild when the device's flag asks for one, which it reads once a substep.
//
// Narrow phase. The pairs are resolved in list order. One wave evaluates 64
// pairs at once against the current state; the pairs before the first one
// that moves a node only count a test and an overlap (order-free), and the
// moving pair is applied alone, after which the wave evaluates again from the
// next pair. So every pair sees exactly the state the serial pass gives it,
// and the ledger's sums run in pair order.

constexpr u32 kPairsPerNode = 24;  // kMaxPairsPerNode (LatticePhysics.hpp:179)
constexpr int kCellBias = 1 << 20; // cell coordinates in [-2^20, 2^20) pack in 21 bits
constexpr u64 kNoCell = ~0ULL;

template <class R> __device__ __forceinline__ R ncCutoff(const NodeContactStep<R> &c) { return R(2) * c.radius + c.skin; }

__device__ __forceinline__ u64 cellKey(int x, int y, int z) {
    return (u64(u32(z + kCellBias)) << 42) | (u64(u32(y + kCellBias)) << 21) | u64(u32(x + kCellBias));
}
__device__ __forceinline__ int cellCoord(u64 key, int shift) { return int(u32(key >> shift) & 0x1fffffU) - kCellBias; }
__device__ __forceinline__ u32 cellSlot(u64 key, u32 mask) {
    key ^= key >> 33;
    key *= 0xff51afd7ed558ccdULL;
    key ^= key >> 33;
    return u32(key) & mask;
}

// Staleness (LatticePhysics.hpp:1566-1574): a node moved more than half the
// skin since the last rebuild.
template <class R> __global__ __launch_bounds__(kWG) void k_nc_stale(Batch<R> B, Step<R> S) {
    const WgMap *m = tileEntry(B);
    if (!m) return;
    const DevLattice<R> L = B.L[m->lat];
    if (L.ctl->stop || L.ctl->nc_rebuild) return;
    const u32 i = m->a * kWG + threadIdx.x;
    if (i >= L.N) return;
    const R half = R(0.5) * S.nc.skin, limit = half * half;
    if (length2(xyz(L.u4[i]) - xyz(L.nc_build[i])) > limit) atomicOr(&L.ctl->nc_rebuild, 1U);
}

// Step 1 (LatticePhysics.hpp:1460-1468): each node's cell and the displacement
// the list is built at, and its bucket, the sort key: a hash of the cell
// masked to the table (massless nodes sort last, as the reference leaves them
// out of the buckets). A bucket can hold several cells; the gather takes only
// the nodes of the cell it visits, by the exact cell key.
template <class R> __global__ void k_nc_cells(DevLattice<R> L, Step<R> S) {
    const R inv = R(1) / ncCutoff(S.nc);
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < L.N; i += gridDim.x * blockDim.x) {
        const R4<R> u = L.u4[i], x0 = L.x04[i];
        const V3<R> p = xyz(x0) + xyz(u);
        const R fx = floor(p.x * inv), fy = floor(p.y * inv), fz = floor(p.z * inv);
        const R lo = R(-kCellBias), hi = R(kCellBias);
        u64 key = kNoCell;
        u32 bucket = L.nc_table_mask + 1U;  // past every bucket: sorts last
        if (x0.w > R(0)) {
            if (fx >= lo && fx < hi && fy >= lo && fy < hi && fz >= lo && fz < hi) {
                key = cellKey(int(fx), int(fy), int(fz));
                bucket = cellSlot(key, L.nc_table_mask);
            } else {
                atomicOr(&L.ctl->nc_error, 1U);
            }
        }
        L.nc_build[i] = u;
        L.nc_key[i] = key;
        L.nc_bucket[i] = bucket;
        L.nc_node[i] = i;
    }
}

__global__ void k_nc_table_clear(uint2 *table, u32 size) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < size; t += gridDim.x * blockDim.x) table[t] = make_uint2(0, 0);
}

// Each bucket's run (start, count) in the sorted order.
template <class R> __global__ void k_nc_table_insert(DevLattice<R> L) {
    for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < L.N; k += gridDim.x * blockDim.x) {
        const u32 bucket = L.nc_bucket_sorted[k];
        if (bucket > L.nc_table_mask || (k > 0 && L.nc_bucket_sorted[k - 1] == bucket)) continue;
        u32 count = 1;
        while (k + count < L.N && L.nc_bucket_sorted[k + count] == bucket) ++count;
        L.nc_table_val[bucket] = make_uint2(k, count);
    }
}

constexpr u32 kSlotRegs = 32;  // a node's neighbour slots the gather keeps in registers

// Step 3 (LatticePhysics.hpp:1494-1546): the pairs node i owns, skipping a
// pair a live bond holds; the bond found in i's slots is kept with the pair.
// The node's slots are read once into registers: the first slot naming j
// decides, as in the reference's scan.
template <class R> __global__ void k_nc_gather(DevLattice<R> L, Step<R> S) {
    const R cutoff = ncCutoff(S.nc), cutoff2 = cutoff * cutoff;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < L.N; i += gridDim.x * blockDim.x) {
        const u64 own = L.nc_key[i];
        u32 fill = 0, over = 0;
        if (own != kNoCell) {
            const V3<R> p = xyz(L.x04[i]) + xyz(L.u4[i]);
            const int cx = cellCoord(own, 0), cy = cellCoord(own, 21), cz = cellCoord(own, 42);
            // Neighbour slots: the other end, with bit 31 set where the bond is
            // dead; kNone for no slot.
            u32 other[kSlotRegs];
            const bool regs = L.D <= kSlotRegs;
#pragma unroll
            for (u32 q = 0; q < kSlotRegs; ++q) {
                u32 o = kNone;
                if (regs && q < L.D) {
                    const u32 key = L.nbr_key[size_t(q) * L.N + i];
                    if (key != kNone) o = u32(int(i) + L.sdelta[key & ~kDeadSlot]) | (key & kDeadSlot);
                }
                other[q] = o;
            }
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int gx = cx + dx, gy = cy + dy, gz = cz + dz;
                        if (gx < -kCellBias || gx >= kCellBias || gy < -kCellBias || gy >= kCellBias ||
                            gz < -kCellBias || gz >= kCellBias)
                            continue;  // no node has a cell out there
                        const u64 cell = cellKey(gx, gy, gz);
                        const uint2 run = L.nc_table_val[cellSlot(cell, L.nc_table_mask)];
                        for (u32 k = run.x; k < run.x + run.y; ++k) {
                            const u32 j = L.nc_node_sorted[k];
                            if (j <= i || L.nc_key[j] != cell) continue;
                            if (length2((xyz(L.x04[j]) + xyz(L.u4[j])) - p) > cutoff2) continue;
                            u32 bond = kNone;
                            bool held = false;
                            if (regs) {
                                u32 found = kSlotRegs, fo = kNone;
#pragma unroll
                                for (u32 q = 0; q < kSlotRegs; ++q)
                                    if (found == kSlotRegs && other[q] != kNone && (other[q] & ~kDeadSlot) == j) {
                                        found = q;
                                        fo = other[q];
                                    }
                                if (found < kSlotRegs) {
                                    held = !(fo & kDeadSlot);
                                    if (!held) bond = L.bnd[L.nbr_bond[size_t(found) * L.N + i]].j;
                                }
                            } else {
                                for (u32 q = 0; q < L.D; ++q) {
                                    const u32 key = L.nbr_key[size_t(q) * L.N + i];
                                    if (key == kNone) break;
                                    if (u32(int(i) + L.sdelta[key & ~kDeadSlot]) != j) continue;
                                    bond = L.bnd[L.nbr_bond[size_t(q) * L.N + i]].j;
                                    held = !(key & kDeadSlot);
                                    break;
                                }
                            }
                            if (held) continue;
                            if (fill >= kPairsPerNode) {
                                ++over;
                                continue;
                            }
                            L.pair_other[size_t(kPairsPerNode) * i + fill] = j;
                            L.pair_bond[size_t(kPairsPerNode) * i + fill] = bond;
                            ++fill;
                        }
                    }
        }
        L.pair_fill[i] = fill;
        if (over) atomicAdd(reinterpret_cast<unsigned long long *>(&L.ctl->nc_overflow), (unsigned long long)over);
    }
}

// Step 4 (LatticePhysics.hpp:1551-1561): the pairs in list order (owner
// ascending, then slot), from the exclusive scan of the fills; the last node
// books the rebuild.
template <class R> __global__ void k_nc_flatten(DevLattice<R> L) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < L.N; i += gridDim.x * blockDim.x) {
        const u32 off = L.pair_offset[i], fill = L.pair_fill[i];
        for (u32 q = 0; q < fill; ++q) {
            L.flat_a[off + q] = i;
            L.flat_b[off + q] = L.pair_other[size_t(kPairsPerNode) * i + q];
            L.flat_bond[off + q] = L.pair_bond[size_t(kPairsPerNode) * i + q];
        }
        if (i + 1 == L.N) {
            DevCtl *ctl = L.ctl;
            ctl->nc_pairs = off + fill;
            ctl->nc_listed += off + fill;
            ctl->nc_rebuilds += 1;
            ctl->nc_rebuild = 0;
        }
    }
}

// One pair (LatticePhysics.hpp:1585-1648), evaluated on loaded values. Returns
// whether the pair moves a node (an impulse or a correction), and fills in
// what it would write.
template <class R> struct PairEval {
    bool valid, acts, impulse, correct;
    R gap;
    V3<R> va, vb, ua, ub;
    R before_after, momentum[3], correction;
};

template <class R>
__device__ __forceinline__ PairEval<R> evalPair(const Step<R> &S, const R4<R> &xa, const R4<R> &xb, R4<R> ua4, R4<R> ub4,
                                                R4<R> va4, R4<R> vb4, bool bond_live) {
    const NodeContactStep<R> &c = S.nc;
    PairEval<R> e{};
    const R ma = xa.w, mb = xb.w;
    e.valid = !bond_live && ma > R(0) && mb > R(0);
    if (!e.valid) return e;
    const V3<R> r = (xyz(xb) + xyz(ub4)) - (xyz(xa) + xyz(ua4));
    const R distance = length(r);
    const V3<R> normal = normalized(r);
    e.gap = distance - R(2) * c.radius;
    if (c.mode != 2U) return e;
    V3<R> va = xyz(va4), vb = xyz(vb4);
    const R vn = dot(vb - va, normal);
    if (e.gap > maxR(c.margin, -vn * S.dt)) return e;
    const R inverse_a = R(1) / ma, inverse_b = R(1) / mb;
    const R inverse_pair = inverse_a + inverse_b;
    const R restitution = -vn > c.rest_thr ? c.restitution : R(0);
    const R desired = e.gap > c.margin ? -e.gap / S.dt : -restitution * minR(vn, R(0));
    const R normal_impulse = maxR(R(0), (desired - vn) / inverse_pair);
    if (normal_impulse > R(0)) {
        const V3<R> va0 = va, vb0 = vb;
        const R before = R(0.5) * ma * length2(va) + R(0.5) * mb * length2(vb);
        V3<R> imp = normal_impulse * normal;
        vb = vb + inverse_b * imp;
        va = va - inverse_a * imp;
        const V3<R> relative = vb - va;
        const V3<R> tangent = relative - dot(relative, normal) * normal;
        const R tangent_speed = length(tangent);
        if (tangent_speed > R(1.0e-12)) {
            const V3<R> direction = tangent / tangent_speed;
            const R sticking = tangent_speed / inverse_pair;
            const R friction = sticking <= c.sf * normal_impulse ? sticking : minR(sticking, c.df * normal_impulse);
            imp = -friction * direction;
            vb = vb + inverse_b * imp;
            va = va - inverse_a * imp;
        }
        e.impulse = true;
        e.va = va;
        e.vb = vb;
        e.before_after = before - (R(0.5) * ma * length2(va) + R(0.5) * mb * length2(vb));
        e.momentum[0] = ma * (va.x - va0.x) + mb * (vb.x - vb0.x);
        e.momentum[1] = ma * (va.y - va0.y) + mb * (vb.y - vb0.y);
        e.momentum[2] = ma * (va.z - va0.z) + mb * (vb.z - vb0.z);
    }
    if (e.gap < -c.margin) {
        e.correction = minR(-e.gap - c.margin, R(0.2) * c.radius);
        const V3<R> shift = (e.correction / inverse_pair) * normal;
        e.ua = xyz(ua4) - inverse_a * shift;
        e.ub = xyz(ub4) + inverse_b * shift;
        e.correct = true;
    }
    e.acts = e.impulse || e.correct;
    return e;
}

__device__ __forceinline__ double waveMaxD(double x) {
    for (int o = 32; o > 0; o >>= 1) {
        const double y = __shfl_xor(x, o);
        x = y > x ? y : x;
    }
    return x;
}

// The narrow phase's schedule. A pair must see the state every earlier pair
// sharing one of its nodes leaves, and nothing else orders it. A node's pairs
// in list order are first those that name it as partner (their owners come
// earlier), then its own block; so a pair's rank at its owner a is
// partners(a) + its slot in a's block, and its rank at its partner b is its
// position among the pairs naming b, from a stable sort by partner.
__global__ void k_nc_iota(u32 *out, u32 n) {
    for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) out[k] = k;
}

template <class R> __global__ void k_nc_ranks(DevLattice<R> L, u32 P) {
    for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < P; k += gridDim.x * blockDim.x) {
        const u32 b = L.nc_sort_key[k];
        u32 s0 = k;
        while (s0 > 0 && L.nc_sort_key[s0 - 1] == b) --s0;
        L.nc_rank_b[L.nc_sort_out[k]] = k - s0;
        if (s0 == k) {
            u32 e = k + 1;
            while (e < P && L.nc_sort_key[e] == b) ++e;
            L.nc_partners[b] = e - k;
        }
    }
}

__global__ void k_nc_clear(u32 *a, u32 n) {
    for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) a[k] = 0;
}

constexpr u32 kNarrowThreads = 1024;
constexpr u32 kNarrowSpinLimit = 1U << 22;  // polls of one pair before the phase gives up with an error

// The narrow phase of one lattice per workgroup (LatticePhysics.hpp:1655-1667).
// Every lane takes a pair and applies it once the pairs before it at both
// its nodes are done (per-node counters, workgroup-scope acquire/release).
// Pairs that run at the same time share no node, so each sees the serial
// pass's state. The ledger's sums then run in pair order over the recorded
// terms; the counts and maxima do not depend on the order.
template <class R> __global__ __launch_bounds__(kNarrowThreads) void k_nc_narrow(Batch<R> B, Step<R> S) {
    const DevLattice<R> L = B.L[B.map[blockIdx.x].lat];
    DevCtl *ctl = L.ctl;
    if (ctl->stop) return;
    __shared__ u32 s_abort, s_waves[kNarrowThreads / 64];
    __shared__ R4<R> s_terms[kNarrowThreads];
    __shared__ double s_max[2][kNarrowThreads / 64];
    __shared__ unsigned long long s_count[2];
    const u32 P = ctl->nc_pairs, tid = threadIdx.x;
    if (tid == 0) {
        s_abort = 0;
        s_count[0] = s_count[1] = 0;
    }
    for (u32 p = tid; p < P; p += kNarrowThreads) {
        L.nc_done[L.flat_a[p]] = 0;
        L.nc_done[L.flat_b[p]] = 0;
    }
    __syncthreads();
    u64 tests = 0, contacts = 0;
    double overlap = 0.0, corr = 0.0;
    for (u32 base = 0; base < P; base += kNarrowThreads) {
        const u32 p = base + tid;
        bool pending = p < P;
        u32 a = 0, b = 0, ra = 0, rb = 0;
        bool live = false;
        R4<R> xa{}, xb{};
        if (pending) {
            a = L.flat_a[p];
            b = L.flat_b[p];
            const u32 bond = L.flat_bond[p];
            live = bond != kNone && L.alive[bond];
            ra = L.nc_partners[a] + (p - L.pair_offset[a]);
            rb = L.nc_rank_b[p];
            xa = L.x04[a];
            xb = L.x04[b];
        }
        u32 spins = 0;
        while (__ballot(pending)) {
            if (!pending) continue;
            const u32 da = __hip_atomic_load(&L.nc_done[a], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            const u32 db = __hip_atomic_load(&L.nc_done[b], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (da == ra && db == rb) {
                const R4<R> ua = L.u4[a], ub = L.u4[b], va = L.v4[a], vb = L.v4[b];
                const PairEval<R> e = evalPair(S, xa, xb, ua, ub, va, vb, live);
                u8 mover = 0;
                if (e.valid) {
                    ++tests;
                    if (e.gap < R(0) && double(-e.gap) > overlap) overlap = double(-e.gap);
                }
                if (e.impulse) {
                    L.v4[a] = with(e.va, va.w);
                    L.v4[b] = with(e.vb, vb.w);
                    L.nc_terms[p] = R4<R>{e.before_after, e.momentum[0], e.momentum[1], e.momentum[2]};
                    mover = 1;
                    ++contacts;
                }
                if (e.correct) {
                    L.u4[a] = with(e.ua, ua.w);
                    L.u4[b] = with(e.ub, ub.w);
                    if (double(e.correction) > corr) corr = double(e.correction);
                }
                L.nc_mover[p] = mover;
                __hip_atomic_fetch_add(&L.nc_done[a], 1U, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(&L.nc_done[b], 1U, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                pending = false;
            } else if (++spins > kNarrowSpinLimit ||
                       __hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                // A pair that never became ready is a schedule error: stop every
                // lane, and report it (the host raises it).
                __hip_atomic_store(&s_abort, 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                atomicOr(&ctl->nc_error, 2U);
                pending = false;
            }
        }
    }
    __syncthreads();
    // Counts and maxima, in any order.
    const u32 lane = tid & 63U, wave = tid >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        tests += __shfl_xor(tests, o);
        contacts += __shfl_xor(contacts, o);
        const double ov = __shfl_xor(overlap, o), co = __shfl_xor(corr, o);
        overlap = ov > overlap ? ov : overlap;
        corr = co > corr ? co : corr;
    }
    if (lane == 0) {
        atomicAdd(&s_count[0], (unsigned long long)tests);
        atomicAdd(&s_count[1], (unsigned long long)contacts);
        s_max[0][wave] = overlap;
        s_max[1][wave] = corr;
    }
    // The ledger's sums in pair order, a chunk of pairs at a time: the movers'
    // terms are compacted in order in LDS and one lane adds them.
    double diss = ctl->nc_dissipated, mom[3] = {ctl->nc_momentum[0], ctl->nc_momentum[1], ctl->nc_momentum[2]};
    for (u32 base = 0; base < P; base += kNarrowThreads) {
        const u32 p = base + tid;
        const u32 flag = p < P && L.nc_mover[p] ? 1U : 0U;
        u32 total = 0;
        const u32 at = blockScan1024(flag, s_waves, &total);
        if (flag) s_terms[at] = L.nc_terms[p];
        __syncthreads();
        if (tid == 0)
            for (u32 k = 0; k < total; ++k) {
                const R4<R> t = s_terms[k];
                diss += double(t.x);
                mom[0] += double(t.y);
                mom[1] += double(t.z);
                mom[2] += double(t.w);
            }
        __syncthreads();
    }
    __syncthreads();  // every wave's maxima and counts are in LDS, also when no pair is listed
    if (tid == 0) {
        double ov = ctl->nc_overlap, co = ctl->nc_correction;
        for (u32 w = 0; w < kNarrowThreads / 64; ++w) {
            ov = s_max[0][w] > ov ? s_max[0][w] : ov;
            co = s_max[1][w] > co ? s_max[1][w] : co;
        }
        ctl->nc_tests += s_count[0];
        ctl->nc_contacts += s_count[1];
        ctl->nc_overlap = ov;
        ctl->nc_correction = co;
        ctl->nc_dissipated = diss;
        for (int k = 0; k < 3; ++k) ctl->nc_momentum[k] = mom[k];
    }
}

// ---------------------------------------------------------------------------
// Fragment mass properties (fracture/FragmentMassProperties.cpp:7-80).
//
// The nodes, stably sorted by canonical label (so each fragment's nodes are in
// ascending order, as its node_indices), are cut into tiles of kMpTile. A tile
// builds each node's terms, runs a segmented inclusive scan in LDS over the
// fragment pieces it holds, and writes the sums of the fragments that lie
// inside it. A fragment that crosses tiles leaves one piece sum per tile; the
// tile it starts in owns it, and k_mp_combine adds its pieces in a fixed tree.
// Pass 1 sums mass, m p and m v; pass 2, with the centre of mass and linear
// velocity, the inertia, angular momentum and kinetic energy terms. Every sum
// has a fixed association, so results repeat bit for bit; the reference adds
// serially in node order, so the two agree to rounding. Each node's terms keep
// the reference's expressions.
constexpr u32 kMpTile = 256;

template <class R> struct MpArgs {
    const R4<R> *u4, *v4;
    const double4 *x0m;      // reference position and mass, as uploaded in double
    const double *spin;      // 3N, or null for no spin
    const u32 *key, *order;  // label and node at each sorted position
    const uint2 *bounds;     // [begin, end) of each fragment in the sorted order
    double *sum;             // fragments x terms
    double *first, *last;    // tiles x terms: piece sums of a tile's first and last fragment
    u32 *cross;              // per tile: the fragment it starts and that continues past it, or kNone
    blx_fragment_mass *frag;
    u32 *bad;                // least 2 f + kind over failing fragments (0 no mass, 1 singular inertia)
    double origin[3], cell;
    u32 n, k;
};

template <int PASS> constexpr int kMpTerms = PASS == 1 ? 7 : 10;

template <class R, int PASS> __device__ __forceinline__ void mpTerms(const MpArgs<R> &a, u32 i, u32 f, double *t) {
    const double4 x = a.x0m[i];
    const R4<R> u = a.u4[i], v = a.v4[i];
    const double m = x.w;
    // position_world_m = origin + x0 + u (FastLattice.cpp:179)
    const double px = (a.origin[0] + x.x) + double(u.x), py = (a.origin[1] + x.y) + double(u.y),
                 pz = (a.origin[2] + x.z) + double(u.z);
    const double vx = double(v.x), vy = double(v.y), vz = double(v.z);
    if (PASS == 1) {  // FragmentMassProperties.cpp:20-26
        t[0] = m;
        t[1] = px * m;
        t[2] = py * m;
        t[3] = pz * m;
        t[4] = vx * m;
        t[5] = vy * m;
        t[6] = vz * m;
        return;
    }
    // FragmentMassProperties.cpp:35-58; the off-diagonal terms are summed and subtracted from zero at the end.
    const blx_fragment_mass &g = a.frag[f];
    const double rx = px - g.center_of_mass_world_m[0], ry = py - g.center_of_mass_world_m[1],
                 rz = pz - g.center_of_mass_world_m[2];
    const double cdi = m * a.cell * a.cell / 6.0;
    t[0] = m * (ry * ry + rz * rz) + cdi;
    t[1] = m * (rx * rx + rz * rz) + cdi;
    t[2] = m * (rx * rx + ry * ry) + cdi;
    t[3] = m * rx * ry;
    t[4] = m * rx * rz;
    t[5] = m * ry * rz;
    const double wx = (vx - g.linear_velocity_m_s[0]) * m, wy = (vy - g.linear_velocity_m_s[1]) * m,
                 wz = (vz - g.linear_velocity_m_s[2]) * m;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (a.spin) {
        sx = a.spin[3 * size_t(i)];
        sy = a.spin[3 * size_t(i) + 1];
        sz = a.spin[3 * size_t(i) + 2];
    }
    t[6] = (ry * wz - rz * wy) + sx * cdi;
    t[7] = (rz * wx - rx * wz) + sy * cdi;
    t[8] = (rx * wy - ry * wx) + sz * cdi;
    t[9] = 0.5 * m * (vx * vx + vy * vy + vz * vz) + 0.5 * cdi * (sx * sx + sy * sy + sz * sz);
}

// Fragment bounds in the sorted order (every fragment has a node).
__global__ void k_mp_bounds(const u32 *key, uint2 *bounds, u32 n) {
    for (u32 q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const u32 f = key[q];
        if (q == 0 || key[q - 1] != f) bounds[f].x = q;
        if (q + 1 == n || key[q + 1] != f) bounds[f].y = q + 1;
    }
}

template <class R, int PASS> __global__ __launch_bounds__(kMpTile) void k_mp_tile(MpArgs<R> a) {
    constexpr int M = kMpTerms<PASS>;
    __shared__ double s[M][kMpTile];
    const u32 t = threadIdx.x, base = blockIdx.x * kMpTile, q = base + t;
    const u32 top = min(base + kMpTile, a.n);
    const bool in = q < a.n;
    double v[M];
    u32 f = kNone;
    uint2 b = make_uint2(0U, 0U);
    if (in) {
        f = a.key[q];
        b = a.bounds[f];
        mpTerms<R, PASS>(a, a.order[q], f, v);
    } else {
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = 0.0;
    }
#pragma unroll
    for (int m = 0; m < M; ++m) s[m][t] = v[m];
    const u32 lo = max(b.x, base);  // first position of this node's piece in the tile
    __syncthreads();
    for (u32 d = 1; d < kMpTile; d <<= 1) {
        const bool take = in && q >= lo + d;
        double w[M];
        if (take) {
#pragma unroll
            for (int m = 0; m < M; ++m) w[m] = s[m][t - d];
        }
        __syncthreads();
        if (take) {
#pragma unroll
            for (int m = 0; m < M; ++m) {
                v[m] = w[m] + v[m];
                s[m][t] = v[m];
            }
        }
        __syncthreads();
    }
    if (!in || q + 1 != min(b.y, top)) return;  // not the last node of its piece
    const bool starts = b.x >= base, ends = b.y <= top;
    if (starts && ends) {
#pragma unroll
        for (int m = 0; m < M; ++m) a.sum[size_t(f) * M + m] = v[m];
    } else {
        if (!starts) {
#pragma unroll
            for (int m = 0; m < M; ++m) a.first[size_t(blockIdx.x) * M + m] = v[m];
        }
        if (!ends) {
#pragma unroll
            for (int m = 0; m < M; ++m) a.last[size_t(blockIdx.x) * M + m] = v[m];
        }
    }
    if (q + 1 == top) a.cross[blockIdx.x] = starts && !ends ? f : kNone;
}

// One workgroup per tile; the tile a crossing fragment starts in adds its
// pieces: each lane serially over a stride of tiles, then a tree in LDS.
template <class R, int PASS> __global__ __launch_bounds__(kMpTile) void k_mp_combine(MpArgs<R> a) {
    constexpr int M = kMpTerms<PASS>;
    __shared__ double s[M][kMpTile];
    const u32 t0 = blockIdx.x, f = a.cross[t0];
    if (f == kNone) return;
    const u32 t1 = (a.bounds[f].y - 1) / kMpTile, parts = t1 - t0 + 1, t = threadIdx.x;
    double acc[M];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] = 0.0;
    for (u32 j = t; j < parts; j += kMpTile) {
        const double *p = j == 0 ? a.last + size_t(t0) * M : a.first + size_t(t0 + j) * M;
#pragma unroll
        for (int m = 0; m < M; ++m) acc[m] = acc[m] + p[m];
    }
#pragma unroll
    for (int m = 0; m < M; ++m) s[m][t] = acc[m];
    __syncthreads();
    for (u32 w = kMpTile / 2; w > 0; w >>= 1) {
        if (t < w) {
#pragma unroll
            for (int m = 0; m < M; ++m) s[m][t] = s[m][t] + s[m][t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int m = 0; m < M; ++m) a.sum[size_t(f) * M + m] = s[m][0];
    }
}

// Mass, centre of mass and linear velocity (FragmentMassProperties.cpp:27-33).
// Each sum starts from zero as the reference's accumulators do (0 + -0 = +0).
template <class R> __global__ void k_mp_final1(MpArgs<R> a) {
    for (u32 f = blockIdx.x * blockDim.x + threadIdx.x; f < a.k; f += gridDim.x * blockDim.x) {
        const double *s = a.sum + size_t(f) * 7;
        blx_fragment_mass g{};
        g.mass_kg = 0.0 + s[0];
        if (g.mass_kg <= 0.0) atomicMin(a.bad, 2 * f);
        for (int c = 0; c < 3; ++c) {
            g.center_of_mass_world_m[c] = (0.0 + s[1 + c]) / g.mass_kg;
            g.linear_velocity_m_s[c] = (0.0 + s[4 + c]) / g.mass_kg;
        }
        const uint2 b = a.bounds[f];
        g.node_count = b.y - b.x;
        g.first_node = a.order[b.x];
        a.frag[f] = g;
    }
}

// Inertia, angular momentum and velocity, energies (FragmentMassProperties.cpp:35-78; Mat3::inverse, core/Math.hpp).
template <class R> __global__ void k_mp_final2(MpArgs<R> a) {
    for (u32 f = blockIdx.x * blockDim.x + threadIdx.x; f < a.k; f += gridDim.x * blockDim.x) {
        const double *s = a.sum + size_t(f) * 10;
        blx_fragment_mass g = a.frag[f];
        double I[3][3];
        I[0][0] = 0.0 + s[0];
        I[1][1] = 0.0 + s[1];
        I[2][2] = 0.0 + s[2];
        I[0][1] = I[1][0] = 0.0 - s[3];
        I[0][2] = I[2][0] = 0.0 - s[4];
        I[1][2] = I[2][1] = 0.0 - s[5];
        const double L[3] = {0.0 + s[6], 0.0 + s[7], 0.0 + s[8]};
        g.source_kinetic_energy_j = 0.0 + s[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) g.inertia_world_kg_m2[3 * r + c] = I[r][c];
        for (int c = 0; c < 3; ++c) g.angular_momentum_kg_m2_s[c] = L[c];
        const double det = I[0][0] * (I[1][1] * I[2][2] - I[1][2] * I[2][1]) -
                           I[0][1] * (I[1][0] * I[2][2] - I[1][2] * I[2][0]) +
                           I[0][2] * (I[1][0] * I[2][1] - I[1][1] * I[2][0]);
        if (fabs(det) <= 0.0) atomicMin(a.bad, 2 * f + 1);
        double J[3][3];
        J[0][0] = (I[1][1] * I[2][2] - I[1][2] * I[2][1]) / det;
        J[0][1] = (I[0][2] * I[2][1] - I[0][1] * I[2][2]) / det;
        J[0][2] = (I[0][1] * I[1][2] - I[0][2] * I[1][1]) / det;
        J[1][0] = (I[1][2] * I[2][0] - I[1][0] * I[2][2]) / det;
        J[1][1] = (I[0][0] * I[2][2] - I[0][2] * I[2][0]) / det;
        J[1][2] = (I[0][2] * I[1][0] - I[0][0] * I[1][2]) / det;
        J[2][0] = (I[1][0] * I[2][1] - I[1][1] * I[2][0]) / det;
        J[2][1] = (I[0][1] * I[2][0] - I[0][0] * I[2][1]) / det;
        J[2][2] = (I[0][0] * I[1][1] - I[0][1] * I[1][0]) / det;
        for (int r = 0; r < 3; ++r) g.angular_velocity_rad_s[r] = J[r][0] * L[0] + J[r][1] * L[1] + J[r][2] * L[2];
        const double *vl = g.linear_velocity_m_s, *w = g.angular_velocity_rad_s;
        g.rigid_kinetic_energy_j = 0.5 * g.mass_kg * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]) +
                                   0.5 * (w[0] * L[0] + w[1] * L[1] + w[2] * L[2]);
        g.coarsening_kinetic_loss_j = g.source_kinetic_energy_j - g.rigid_kinetic_energy_j;
        a.frag[f] = g;
    }
}

// Streaming copy, one 16-byte element per thread: the achievable-bandwidth
// reference, and the known traffic that scales the HBM byte counters.
__global__ __launch_bounds__(kWG) void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    const size_t i = size_t(blockIdx.x) * kWG + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

} // namespace blx
