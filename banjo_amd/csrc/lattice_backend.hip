// The MI355X lattice backend: device memory, the substep launch sequence and
// the C ABI of include/banjo_lattice.h. It is the HIP counterpart of the
// reference's LatticeBackend (src/fastlattice/FastLattice.hpp:255-369), whose
// CUDA implementation the reference no longer carries
// (src/fastlattice/CudaLatticeUnavailable.cpp:1-27).
//
// One engine advances a batch of independent lattices ("slots") that share
// step settings: every phase of the substep is one launch over all running
// slots, each workgroup finding its lattice in the launch's map. The plain
// single-lattice backend is an engine with one slot. Each slot's results are
// those of the same lattice run alone, bit for bit: lattices share no data,
// and every slot keeps its own colour order, folds and exit rule.

#include "banjo_lattice.h"
#include "lattice_errors.hpp"
#include "lattice_kernels.hpp"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <bit>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace blx {

std::string &lastError() {
    thread_local std::string message;
    return message;
}

#define HIP_OK(expr)                                                                                  \
    do {                                                                                              \
        const hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) throw Runtime(std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

namespace {

// One device allocation, freed with its owner.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t n) {
        if (p && bytes >= n) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (n == 0) n = 16;
        HIP_OK(hipMalloc(&p, n));
        bytes = n;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// A slot's arrays carved from one allocation, which a later upload reuses
// when it fits: admitting a lattice into a used slot allocates nothing.
class Arena {
public:
    void reset() { total_ = 0; }
    template <class T> size_t reserve(size_t count) {
        total_ = (total_ + 255) & ~size_t(255);
        const size_t at = total_;
        total_ += std::max<size_t>(count, 1) * sizeof(T);
        return at;
    }
    void commit() { buf_.alloc(total_); }
    template <class T> T *at(size_t offset) const { return reinterpret_cast<T *>(static_cast<char *>(buf_.p) + offset); }
    size_t bytes() const { return buf_.bytes; }
    size_t used() const { return total_; }

private:
    DevBuf buf_;
    size_t total_ = 0;
};

// Distinct keys, compared bit for bit, numbered in first-seen order: an
// open-addressing table that stays in cache while the distinct count is small.
template <size_t W> class Dedup {
public:
    using Key = std::array<uint32_t, W>;
    Dedup() : slots_(1024, kEmpty) {}
    uint32_t index(const Key &k) {
        if ((keys_.size() + 1) * 2 > slots_.size()) grow();
        const size_t mask = slots_.size() - 1;
        for (size_t h = hash(k) & mask;; h = (h + 1) & mask) {
            const uint32_t s = slots_[h];
            if (s == kEmpty) {
                slots_[h] = uint32_t(keys_.size());
                keys_.push_back(k);
                return slots_[h];
            }
            if (keys_[s] == k) return s;
        }
    }
    const std::vector<Key> &keys() const { return keys_; }

private:
    static constexpr uint32_t kEmpty = 0xffffffffU;
    std::vector<uint32_t> slots_;
    std::vector<Key> keys_;
    static size_t hash(const Key &k) {
        uint64_t h = 0x243f6a8885a308d3ULL;
        for (uint32_t w : k) h = (h ^ w) * 0x9e3779b97f4a7c15ULL;
        return size_t(h ^ (h >> 29));
    }
    void grow() {
        std::vector<uint32_t> next(slots_.size() * 2, kEmpty);
        const size_t mask = next.size() - 1;
        for (uint32_t s = 0; s < keys_.size(); ++s) {
            size_t h = hash(keys_[s]) & mask;
            while (next[h] != kEmpty) h = (h + 1) & mask;
            next[h] = s;
        }
        slots_.swap(next);
    }
};

// Scratch of the allocation-free fragment labels
// (blx_backend_fragment_labels_noalloc): every buffer of one label computation
// of at most kNoAllocMaxNodes nodes, carved from one file-scope device array
// that each GPU holds from module load to process end. In host code `scratch`
// names a host-side stand-in, not device memory: its device address on each
// device comes from hipGetSymbolAddress (scratchOn), and only that address is
// passed to kernels and copies. Calls on one GPU share its array, so calls
// that overlap on one GPU race.
constexpr uint32_t kNoAllocMaxNodes = 1U << 20;
constexpr size_t kNoAllocTempBytes = size_t(16) << 20;  // radix sort temporary storage
constexpr size_t kNoAllocBytes = size_t(kNoAllocMaxNodes) * (4 * 4 + 2 * 8) + 256 + kNoAllocTempBytes;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, looked up once per device with
// that device current.
void *scratchOn(int device) {
    if (device < 0 || device >= kMaxDevices) throw Invalid("device index out of range for the label scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int current = 0;
        HIP_OK(hipGetDevice(&current));
        HIP_OK(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(current);
        if (e != hipSuccess || !p)
            throw Runtime(std::string("hipGetSymbolAddress of the label scratch: ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

int deviceCount() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

struct Schedule {
    uint32_t N = 0, B = 0, blocks = 0, colors = 0;
    std::vector<uint32_t> nbb, range_begin, range_end;
};

Schedule scheduleOf(const blx_schedule_view &sv) {
    Schedule s;
    s.N = sv.node_count;
    s.B = sv.bond_count;
    s.blocks = sv.block_count;
    s.colors = sv.color_count;
    if (s.blocks == 0) throw Invalid("a schedule has at least one block");
    s.nbb.assign(sv.node_block_begin, sv.node_block_begin + sv.block_count + 1);
    const size_t ranges = size_t(sv.block_count) * 2 * std::max(sv.color_count, 1U);
    s.range_begin.assign(sv.range_begin, sv.range_begin + ranges);
    s.range_end.assign(sv.range_end, sv.range_end + ranges);
    return s;
}

template <class R> struct HostFrame {
    uint64_t step;
    std::vector<float> u;
    std::vector<uint8_t> alive;
    std::vector<float> damage;
    Sphere<R> sphere;
};

template <class R> void sphereOut(const Sphere<R> &s, blx_sphere &o) {
    o.center[0] = double(s.center.x);
    o.center[1] = double(s.center.y);
    o.center[2] = double(s.center.z);
    o.velocity[0] = double(s.velocity.x);
    o.velocity[1] = double(s.velocity.y);
    o.velocity[2] = double(s.velocity.z);
    o.angular_velocity[0] = double(s.omega.x);
    o.angular_velocity[1] = double(s.omega.y);
    o.angular_velocity[2] = double(s.omega.z);
    o.radius = double(s.radius);
    o.mass = double(s.mass);
    o.inertia = double(s.inertia);
}

template <class R> Step<R> convertStep(const blx_settings &s) {
    Step<R> o{};
    o.dt = R(s.dt);
    o.gravity = {R(s.gravity[0]), R(s.gravity[1]), R(s.gravity[2])};
    o.damping = R(s.damping_fraction);
    o.yield = R(s.plastic_yield_stretch);
    o.hardening = R(s.plastic_hardening);
    o.c_sf = R(s.contact.static_friction);
    o.c_df = R(s.contact.dynamic_friction);
    o.c_rest = R(s.contact.restitution);
    o.c_rest_thr = R(s.contact.restitution_speed_threshold);
    o.c_node_radius = R(s.contact.node_contact_radius);
    o.c_margin = R(s.contact.contact_margin);
    o.c_slack = R(s.contact.prefilter_slack);
    o.plane_count = s.plane_count;
    o.sphere_enabled = s.sphere_enabled;
    o.nc = {s.node_contact.mode, s.node_contact.bucket_mask, R(s.node_contact.radius), R(s.node_contact.skin),
            R(s.node_contact.margin), R(s.node_contact.restitution), R(s.node_contact.restitution_speed_threshold),
            R(s.node_contact.static_friction), R(s.node_contact.dynamic_friction)};
    for (int p = 0; p < 2; ++p) {
        const blx_support_plane &in = s.planes[p];
        Plane<R> &pl = o.planes[p];
        pl.point = {R(in.point[0]), R(in.point[1]), R(in.point[2])};
        pl.normal = {R(in.normal[0]), R(in.normal[1]), R(in.normal[2])};
        pl.tangent = {R(in.tangent[0]), R(in.tangent[1]), R(in.tangent[2])};
        pl.bitangent = {R(in.bitangent[0]), R(in.bitangent[1]), R(in.bitangent[2])};
        pl.restitution = R(in.restitution);
        pl.static_friction = R(in.static_friction);
        pl.dynamic_friction = R(in.dynamic_friction);
        pl.node_radius = R(in.node_radius);
        pl.reach_capped = in.reach_capped;
        pl.footprint_count = in.footprint_count;
        for (int f = 0; f < 4; ++f) {
            pl.fp[f][0] = R(in.footprints[f].center_t);
            pl.fp[f][1] = R(in.footprints[f].center_b);
            pl.fp[f][2] = R(in.footprints[f].half_t);
            pl.fp[f][3] = R(in.footprints[f].half_b);
        }
    }
    return o;
}

void validateSettings(const blx_settings &s) {
    if (s.bond_integrator == BLX_BOND_VELOCITY_VERLET)
        throw Invalid("the Verlet bond integrator is the serial double CPU reference only");
    if (s.bond_integrator != BLX_BOND_XPBD) throw Invalid("unknown bond integrator");
    if (s.node_contact.mode > BLX_NODE_CONTACT_ON) throw Invalid("unknown node contact mode");
    if (s.node_contact.mode != BLX_NODE_CONTACT_OFF && (s.node_contact.bucket_mask & (s.node_contact.bucket_mask + 1U)) != 0U)
        throw Invalid("node contact bucket_mask must be a power of two minus one");
    if (s.plate_half_thickness != 0.0) throw Invalid("plate bending is not available on the HIP backend");
    if (s.audit_energy) throw Invalid("the kinetic energy audit is not available on the HIP backend");
    if (s.plane_count > BLX_MAX_SUPPORT_PLANES) throw Invalid("at most two support planes");
    for (uint32_t p = 0; p < s.plane_count; ++p)
        if (s.planes[p].footprint_count > BLX_MAX_FOOTPRINTS) throw Invalid("at most four footprints per plane");
    if (s.constraint_iterations == 0) throw Invalid("at least one constraint iteration");
    if (!(s.dt > 0.0) || !std::isfinite(s.dt)) throw Invalid("the substep must be finite and positive");
}

// The settings every slot of a batch shares, compared as the values the
// kernels see.
struct Shared {
    bool set = false;
    bool direct = false, plastic = false;
    uint32_t iterations = 1;
    std::vector<unsigned char> step_bytes;
};

// ---------------------------------------------------------------------------
// One lattice of a batch: its device arrays (one arena), its colour stages,
// its run status and frames.

template <class R> class Slot {
public:
    Schedule sch;
    bool scheduled = false;
    bool uploaded = false;
    bool armed = false;  // takes part in the next advance
    bool pristine = false;  // uploaded and never run: a template for clones
    bool loads_set = false;
    uint32_t N = 0, B = 0, D = 0, node_tiles = 0, bond_tiles = 0, group_words = 0;
    uint32_t slot_palette = 0, edge_palette = 0;
    bool tensile_only = false;
    bool contact = false;        // node-node contact arrays allocated
    uint32_t nc_table_mask = 0;  // cell table size minus one (a power of two >= 2N)
    DevBuf nc_temp;              // rocPRIM scratch of the contact rebuild
    DevLattice<R> L{};
    DevCtl ctl_host{};
    blx_run_status status{};
    std::vector<HostFrame<R>> frames;
    std::vector<R> mass_host, inv_mass_host;
    std::vector<uint32_t> node_j_host;  // node-major bond position -> schedule index
    // Colour stages in sweep order (CpuLatticeBackend.cpp:395-404): interior
    // bonds of every block colour by colour, then the boundary bonds of even
    // blocks, then of odd blocks. Interior bonds of different blocks share no
    // node, nor do boundary bonds of blocks of one parity, so the ranges of
    // one colour across those blocks form one node-disjoint stage. Each stage
    // is a list of bond ranges of at most one workgroup each.
    std::vector<std::vector<uint2>> stages;
    size_t interior_stages = 0;
    // The paired sweep (k_sweep_pair) of a one-iteration run: per launch group, its
    // map entries (lat filled in when the maps are built) and its first stage.
    std::vector<std::vector<WgMap>> sweep_groups;
    // The stages as merged bond ranges for the one-workgroup sweep
    // (k_sweep_lds): stage s is sweep_ranges[stage_end[s - 1] .. stage_end[s]).
    std::vector<uint2> sweep_ranges;
    std::vector<uint32_t> stage_end;
    std::vector<uint32_t> sweep_first_stage;
    std::vector<uint32_t> sweep_lds_nodes;  // per launch group: the most nodes one of its workgroups holds in LDS
    // Offsets of the device arrays in the arena.
    struct Off {
        size_t x04, u4, u4b, pnodes, pbonds, stage_ranges, stage_end, up4, v4, rinv, unm, sa, sb, valid, approach, engaged, cand;
        size_t nbr_key, spal, sdelta, nbr_bond, epal, nodir, rec, cmpl, alive, lambda, pext;
        size_t bnd, alive_n, thr, thr2, pv, pyz, mode, pstr, slots;
        size_t ev_bits, brk_bits, grp_bits, ev_d, ev_e, flag_list, tile_cand, dirty_list, dirty_stamp, rank_stamp;
        size_t affected, affected_list, tile_plastic, load_part, fext, freq, cand_all, cand_tile, cand_lo, cand_cnt;
        size_t nbb, first_fail;
        size_t nc_build, nc_key, nc_bucket, nc_bucket_sorted, nc_node, nc_node_sorted, nc_table_val;
        size_t pair_fill, pair_other, pair_bond, pair_offset, flat_a, flat_b, flat_bond;
        size_t nc_partners, nc_done, nc_rank_b, nc_sort_key, nc_sort_in, nc_sort_out, nc_terms, nc_mover;
    } off{};
    size_t used = 0;  // arena bytes the arrays occupy
    Arena arena;
    DevBuf cc_parent, cc_root, cc_size, cc_rank, cc_label, cc_keys, cc_sorted, cc_count, cc_temp;
    // Reference positions and masses as uploaded, in double (x, y, z, mass per
    // node), for fragment mass properties; copied to mp_x0m on first use.
    std::shared_ptr<const std::vector<double>> x0m_host;
    bool mp_ready = false;
    DevBuf mp_x0m, mp_spin, mp_key, mp_order, mp_iota, mp_bounds, mp_sum, mp_first, mp_last, mp_cross, mp_frag, mp_bad,
        mp_temp;

    void setSchedule(Schedule s) {
        sch = std::move(s);
        stages.clear();
        const uint32_t cc = std::max(sch.colors, 1U);
        auto add = [&](uint32_t boundary, uint32_t parity_first, uint32_t step) {
            for (uint32_t c = 0; c < sch.colors; ++c) {
                std::vector<uint2> st;
                for (uint32_t b = parity_first; b < sch.blocks; b += step) {
                    const size_t key = (size_t(b) * 2 + boundary) * cc + c;
                    const uint32_t lo = sch.range_begin[key], hi = sch.range_end[key];
                    for (uint32_t j = lo; j < hi; j += kWG) st.push_back(make_uint2(j, std::min(hi, j + kWG)));
                }
                if (!st.empty()) stages.push_back(std::move(st));
            }
        };
        add(0, 0, 1);
        interior_stages = stages.size();
        add(1, 0, 2);
        add(1, 1, 2);
        sweep_ranges.clear();
        stage_end.clear();
        for (const auto &st : stages) {
            for (const uint2 &r : st) {
                if (!sweep_ranges.empty() && sweep_ranges.size() > (stage_end.empty() ? 0 : stage_end.back()) &&
                    sweep_ranges.back().y == r.x)
                    sweep_ranges.back().y = r.y;
                else
                    sweep_ranges.push_back(r);
            }
            stage_end.push_back(uint32_t(sweep_ranges.size()));
        }
        scheduled = true;
        uploaded = false;
        armed = false;
    }

    // The device view of the arrays in this slot's arena.
    void bind(DevCtl *ctl_dev, Sphere<R> *sphere_dev, const double *origin) {
        const Off &o = off;
        DevLattice<R> &l = L;
        l = DevLattice<R>{};
        l.N = N;
        l.B = B;
        l.D = D;
        l.blocks = sch.blocks;
        l.node_tiles = node_tiles;
        l.bond_tiles = bond_tiles;
        l.group_words = group_words;
        l.npal = edge_palette;
        l.tensile_only = tensile_only ? 1U : 0U;
        l.u4 = arena.at<R4<R>>(o.u4);
        l.u4b = arena.at<R4<R>>(o.u4b);
        l.stage_ranges = arena.at<uint2>(o.stage_ranges);
        l.stage_end = arena.at<uint32_t>(o.stage_end);
        l.stage_count = uint32_t(stage_end.size());
        {
            // The LDS sweep's prefetching path: one bond range per stage and
            // a palette that fits its LDS slots.
            bool one = true;
            uint32_t widest = 0;
            for (size_t st = 0; st < stage_end.size(); ++st) {
                const uint32_t q0 = st == 0 ? 0U : stage_end[st - 1], q1 = stage_end[st];
                one = one && q1 == q0 + 1;
                uint32_t n = 0;
                for (uint32_t q = q0; q < q1; ++q) n += sweep_ranges[q].y - sweep_ranges[q].x;
                widest = std::max(widest, n);
            }
            l.lds_fast = one && edge_palette <= kLdsPal ? 1U : 0U;
            l.stage_max = widest;
        }
        l.pnodes = arena.at<uint32_t>(o.pnodes);
        l.pbonds = arena.at<PairBond<R>>(o.pbonds);
        l.up4 = arena.at<R4<R>>(o.up4);
        l.v4 = arena.at<R4<R>>(o.v4);
        l.x04 = arena.at<R4<R>>(o.x04);
        l.rinv = arena.at<R>(o.rinv);
        l.unm = arena.at<R>(o.unm);
        l.sa = arena.at<R4<R>>(o.sa);
        l.sb = arena.at<R4<R>>(o.sb);
        l.valid = arena.at<uint8_t>(o.valid);
        l.approach = arena.at<R>(o.approach);
        l.engaged = arena.at<uint8_t>(o.engaged);
        l.cand = arena.at<uint8_t>(o.cand);
        l.nbr_key = arena.at<uint32_t>(o.nbr_key);
        l.spal = arena.at<R4<R>>(o.spal);
        l.sdelta = arena.at<int>(o.sdelta);
        l.nbr_bond = arena.at<uint32_t>(o.nbr_bond);
        l.epal = arena.at<EdgePal<R>>(o.epal);
        l.rec = arena.at<BondRec<R>>(o.rec);
        l.cmpl = arena.at<R>(o.cmpl);
        l.alive = arena.at<uint8_t>(o.alive);
        l.lambda = arena.at<R>(o.lambda);
        l.pext = arena.at<R>(o.pext);
        l.bnd = arena.at<BondNode>(o.bnd);
        l.alive_n = arena.at<uint8_t>(o.alive_n);
        l.thr = arena.at<Thresholds<R>>(o.thr);
        l.thr2 = arena.at<Tensile<R>>(o.thr2);
        l.pv = arena.at<R2<R>>(o.pv);
        l.pyz = arena.at<R2<R>>(o.pyz);
        l.mode = arena.at<uint8_t>(o.mode);
        l.pstr = arena.at<R>(o.pstr);
        l.slots = arena.at<uint2>(o.slots);
        l.ev_bits = arena.at<uint64_t>(o.ev_bits);
        l.brk_bits = arena.at<uint64_t>(o.brk_bits);
        l.grp_bits = arena.at<uint64_t>(o.grp_bits);
        l.ev_d = arena.at<float>(o.ev_d);
        l.ev_e = arena.at<R>(o.ev_e);
        l.flag_list = arena.at<uint32_t>(o.flag_list);
        l.tile_cand = arena.at<uint32_t>(o.tile_cand);
        l.dirty_list = arena.at<uint32_t>(o.dirty_list);
        l.dirty_stamp = arena.at<uint32_t>(o.dirty_stamp);
        l.rank_stamp = arena.at<uint32_t>(o.rank_stamp);
        l.affected = arena.at<uint32_t>(o.affected);
        l.affected_list = arena.at<uint32_t>(o.affected_list);
        l.tile_plastic = arena.at<double>(o.tile_plastic);
        l.load_part = arena.at<double>(o.load_part);
        l.fext = arena.at<R4<R>>(o.fext);
        l.freq = arena.at<double>(o.freq);
        l.cand_all = arena.at<uint32_t>(o.cand_all);
        l.cand_tile = arena.at<uint32_t>(o.cand_tile);
        l.cand_lo = arena.at<uint32_t>(o.cand_lo);
        l.cand_cnt = arena.at<uint32_t>(o.cand_cnt);
        l.nbb = arena.at<uint32_t>(o.nbb);
        l.first_fail = arena.at<uint32_t>(o.first_fail);
        l.nc_build = arena.at<R4<R>>(o.nc_build);
        l.nc_key = arena.at<uint64_t>(o.nc_key);
        l.nc_bucket = arena.at<uint32_t>(o.nc_bucket);
        l.nc_bucket_sorted = arena.at<uint32_t>(o.nc_bucket_sorted);
        l.nc_node = arena.at<uint32_t>(o.nc_node);
        l.nc_node_sorted = arena.at<uint32_t>(o.nc_node_sorted);
        l.nc_table_val = arena.at<uint2>(o.nc_table_val);
        l.nc_table_mask = nc_table_mask;
        l.pair_fill = arena.at<uint32_t>(o.pair_fill);
        l.pair_other = arena.at<uint32_t>(o.pair_other);
        l.pair_bond = arena.at<uint32_t>(o.pair_bond);
        l.pair_offset = arena.at<uint32_t>(o.pair_offset);
        l.flat_a = arena.at<uint32_t>(o.flat_a);
        l.flat_b = arena.at<uint32_t>(o.flat_b);
        l.flat_bond = arena.at<uint32_t>(o.flat_bond);
        l.nc_partners = arena.at<uint32_t>(o.nc_partners);
        l.nc_done = arena.at<uint32_t>(o.nc_done);
        l.nc_rank_b = arena.at<uint32_t>(o.nc_rank_b);
        l.nc_sort_key = arena.at<uint32_t>(o.nc_sort_key);
        l.nc_sort_in = arena.at<uint32_t>(o.nc_sort_in);
        l.nc_sort_out = arena.at<uint32_t>(o.nc_sort_out);
        l.nc_terms = arena.at<R4<R>>(o.nc_terms);
        l.nc_mover = arena.at<uint8_t>(o.nc_mover);
        l.ctl = ctl_dev;
        l.sphere = sphere_dev;
        for (int k = 0; k < 3; ++k) l.origin[k] = origin[k];
    }

    // Takes the lattice of `src` as uploaded (its arrays copied on the
    // device, its host-side tables copied); the caller places the control
    // block and the sphere.
    void cloneFrom(const Slot &src, hipStream_t stream, DevCtl *ctl_dev, Sphere<R> *sphere_dev) {
        if (!src.uploaded) throw Logic("clone of a slot with no lattice");
        sch = src.sch;
        stages = src.stages;
        interior_stages = src.interior_stages;
        sweep_groups = src.sweep_groups;
        sweep_first_stage = src.sweep_first_stage;
        sweep_lds_nodes = src.sweep_lds_nodes;
        sweep_ranges = src.sweep_ranges;
        stage_end = src.stage_end;
        scheduled = true;
        N = src.N;
        B = src.B;
        D = src.D;
        node_tiles = src.node_tiles;
        bond_tiles = src.bond_tiles;
        group_words = src.group_words;
        slot_palette = src.slot_palette;
        edge_palette = src.edge_palette;
        tensile_only = src.tensile_only;
        contact = src.contact;
        nc_table_mask = src.nc_table_mask;
        if (contact) nc_temp.alloc(src.nc_temp.bytes);
        mass_host = src.mass_host;
        inv_mass_host = src.inv_mass_host;
        node_j_host = src.node_j_host;
        x0m_host = src.x0m_host;
        mp_ready = false;
        off = src.off;
        used = src.used;
        arena.reset();
        arena.template reserve<char>(used);
        arena.commit();
        HIP_OK(hipMemcpyAsync(arena.template at<char>(0), src.arena.template at<char>(0), used, hipMemcpyDeviceToDevice,
                              stream));
        HIP_OK(hipStreamSynchronize(stream));
        bind(ctl_dev, sphere_dev, src.L.origin);
        loads_set = false;
        frames.clear();
        status = blx_run_status{};
        status.first_failure_step = status.last_failure_step = status.last_energy_gain_step = ~0ULL;
        uploaded = true;
        armed = false;
    }

    // The grouped sweep's plan (k_sweep_pair, k_sweep_triple) of a
    // one-iteration run. Nodes are grouped in tiles of kPairTile
    // (kPairTileSmall below kPairWideNodes nodes) cells of the reference grid
    // (the cell is the shortest rest length), ordered z, y, x, so that a
    // workgroup's owned nodes fill whole cache lines and the workgroups next
    // to it in the launch (which run on the same XCD) share its halo. Per
    // group of stages, the workgroup of each tile lists its owned nodes, the
    // halo they depend on and each stage's bonds to solve; a tile whose lists
    // exceed one workgroup is split in halves by node order until they fit
    // (one owned node needs at most 2^K nodes and K bonds). The stages are
    // grouped as a run of single stages, each in place, then an even number
    // of pairs and triples, so that the groups' ping-pong ends in u4. A
    // triple is used, on lattices of at least kPairWideNodes nodes, only where
    // no tile of it splits: a split adds workgroups past the 1,024 a launch
    // holds resident at once, and the launch then runs a second round. Among
    // those choices the plan takes the one of least launch cost (single 11,
    // pair 16, triple 22: the measured microseconds of each kind on the
    // 128 x 64 x 128 impact). Bond entries get their palette index and
    // compliance when the record is built.
    void buildSweep(const blx_state_view &st, std::vector<uint32_t> &nodes, std::vector<PairBond<R>> &bonds) {
        sweep_groups.clear();
        sweep_first_stage.clear();
        sweep_lds_nodes.clear();
        nodes.clear();
        bonds.clear();
        const size_t S = stages.size();
        const uint32_t *end_a = st.bond_a, *end_b = st.bond_b;
        auto single = [&](size_t s) {
            std::vector<uint2> runs;
            for (const uint2 &r : stages[s]) {
                if (!runs.empty() && runs.back().y == r.x) runs.back().y = r.y;
                else runs.push_back(r);
            }
            std::vector<WgMap> g;
            for (const uint2 &r : runs)
                for (uint32_t j = r.x; j < r.y; j += kPairThreads) {
                    g.push_back(WgMap{0, j, std::min(r.y, j + kPairThreads), kSweepSingle << 16});
                    g.push_back(WgMap{0, 0, 0, 0});
                }
            sweep_groups.push_back(std::move(g));
            sweep_first_stage.push_back(uint32_t(s));
            sweep_lds_nodes.push_back(0);
        };
        if (S < 4) {
            for (size_t s = 0; s < S; ++s) single(s);
            return;
        }
        // Nodes in tile order (ties in node order).
        double cell = std::numeric_limits<double>::infinity();
        for (uint32_t k = 0; k < B; ++k)
            if (st.rest_length[k] > 0.0) cell = std::min(cell, st.rest_length[k]);
        if (!(cell < std::numeric_limits<double>::infinity())) cell = 1.0;
        double lo[3] = {0.0, 0.0, 0.0};
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::numeric_limits<double>::infinity();
            for (uint32_t i = 0; i < N; ++i) lo[c] = std::min(lo[c], st.x0[3 * size_t(i) + c]);
        }
        std::vector<uint64_t> key(N);
        const uint32_t *tile = N >= kPairWideNodes ? kPairTile : kPairTileSmall;
        for (uint32_t i = 0; i < N; ++i) {
            uint64_t k = 0;
            for (int c = 2; c >= 0; --c) {
                const double g = std::floor((st.x0[3 * size_t(i) + c] - lo[c]) / cell + 0.5);
                const uint64_t q = g > 0.0 ? std::min<uint64_t>(uint64_t(g) / tile[c], (1ULL << 21) - 1) : 0;
                k = (k << 21) | q;
            }
            key[i] = k;
        }
        std::vector<uint32_t> order(N);
        for (uint32_t i = 0; i < N; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return key[x] < key[y]; });
        std::vector<uint32_t> tile_begin;
        for (uint32_t q = 0; q < N; ++q)
            if (q == 0 || key[order[q]] != key[order[q - 1]]) tile_begin.push_back(q);
        tile_begin.push_back(N);

        // partner[k][v]: the bond of stage s + k at node v, if any.
        std::vector<uint32_t> partner[3], nstamp(N, 0), nslot(N, 0), bstamp(B, 0);
        for (auto &p : partner) p.assign(N, kNone);
        uint32_t stamp = 0;
        std::vector<uint32_t> halo, level, list[3];
        auto other = [&](uint32_t j, uint32_t v) { return end_a[j] == v ? end_b[j] : end_a[j]; };
        auto setPartners = [&](size_t s, size_t K, bool on) {
            for (size_t k = 0; k < K; ++k)
                for (const uint2 &r : stages[s + k])
                    for (uint32_t j = r.x; j < r.y; ++j) partner[k][end_a[j]] = partner[k][end_b[j]] = on ? j : kNone;
        };
        // The lists of one workgroup owning own[0 .. n) for a group of K
        // stages; whether they fit one workgroup.
        auto gather = [&](const uint32_t *own, uint32_t n, size_t K) {
            if (++stamp == 0) throw Logic("grouped sweep plan stamp wrapped");
            halo.clear();
            for (size_t k = 0; k < K; ++k) list[k].clear();
            for (uint32_t q = 0; q < n; ++q) {
                nstamp[own[q]] = stamp;
                nslot[own[q]] = q;
            }
            // level: the nodes whose values must be right after stage s + k,
            // from the owned ones at the last stage back to stage s; every
            // node a listed bond reaches joins it (and the halo).
            level.assign(own, own + n);
            for (size_t k = K; k-- > 0;) {
                const size_t reach = level.size();
                for (size_t q = 0; q < reach; ++q) {
                    const uint32_t v = level[q], j = partner[k][v];
                    if (j == kNone || bstamp[j] == stamp) continue;
                    bstamp[j] = stamp;
                    list[k].push_back(j);
                    const uint32_t y = other(j, v);
                    if (nstamp[y] != stamp) {
                        nstamp[y] = stamp;
                        halo.push_back(y);
                        level.push_back(y);
                    }
                }
            }
            bool fits = n + halo.size() <= kPairMaxNodes;
            for (size_t k = 0; k < K; ++k) fits = fits && list[k].size() <= kPairMaxBonds;
            return fits;
        };
        uint32_t dir = 0, widest = 0;
        std::vector<WgMap> g;
        // One workgroup's lists, or two halves' when they do not fit.
        std::function<void(const uint32_t *, uint32_t, size_t)> emit = [&](const uint32_t *own, uint32_t n, size_t K) {
            if (!gather(own, n, K)) {
                if (n < 2) throw Logic("a grouped sweep tile of one node does not fit a workgroup");
                const uint32_t h = n / 2;
                emit(own, h, K);
                emit(own + h, n - h, K);
                return;
            }
            std::sort(halo.begin(), halo.end());
            for (size_t q = 0; q < halo.size(); ++q) nslot[halo[q]] = uint32_t(n + q);
            const uint32_t all = uint32_t(n + halo.size());
            const uint32_t kind = K == 3 ? kSweepTriple : kSweepPair;
            g.push_back(WgMap{0, uint32_t(nodes.size()), uint32_t(bonds.size()), (kind << 16) | dir});
            g.push_back(WgMap{n | (all << 16), uint32_t(list[0].size()) | (uint32_t(list[1].size()) << 16),
                              K == 3 ? uint32_t(list[2].size()) : 0U, 0});
            nodes.insert(nodes.end(), own, own + n);
            nodes.insert(nodes.end(), halo.begin(), halo.end());
            for (size_t k = 0; k < K; ++k) {
                std::sort(list[k].begin(), list[k].end());
                for (uint32_t j : list[k]) {
                    PairBond<R> e{};
                    e.j = j;
                    e.slots = nslot[end_a[j]] | (nslot[end_b[j]] << 16);
                    bonds.push_back(e);
                }
            }
            if (nodes.size() >= 0xffffffffULL || bonds.size() >= 0xffffffffULL)
                throw Invalid("the grouped sweep holds at most 2^32 - 1 node and bond entries");
            widest = std::max(widest, all);
        };

        // Where a triple fits without a split.
        std::vector<char> triple(S, 0);
        if (N >= kPairWideNodes)
            for (size_t s = 0; s + 3 <= S; ++s) {
                setPartners(s, 3, true);
                bool ok = true;
                for (size_t t = 0; ok && t + 1 < tile_begin.size(); ++t)
                    ok = gather(order.data() + tile_begin[t], tile_begin[t + 1] - tile_begin[t], 3);
                setPartners(s, 3, false);
                triple[s] = ok;
            }
        // Least cost of the groups covering stages s .. S with `odd` groups
        // before them, ending on an even count; then the run of singles.
        constexpr double kSingle = 11.0, kPair = 16.0, kTriple = 22.0, kNever = 1e30;
        std::vector<double> cost(2 * (S + 1), kNever);
        std::vector<uint8_t> take(2 * (S + 1), 0);
        cost[2 * S] = 0.0;
        for (size_t s = S; s-- > 0;)
            for (int odd = 0; odd < 2; ++odd) {
                double &c = cost[2 * s + odd];
                if (s + 2 <= S && kPair + cost[2 * (s + 2) + (odd ^ 1)] < c) {
                    c = kPair + cost[2 * (s + 2) + (odd ^ 1)];
                    take[2 * s + odd] = 2;
                }
                if (s + 3 <= S && triple[s] && kTriple + cost[2 * (s + 3) + (odd ^ 1)] < c) {
                    c = kTriple + cost[2 * (s + 3) + (odd ^ 1)];
                    take[2 * s + odd] = 3;
                }
            }
        size_t singles = S;
        for (size_t f = 0; f <= S; ++f)
            if (kSingle * f + cost[2 * f] < kSingle * singles + cost[2 * singles]) singles = f;
        size_t s = 0;
        while (s < singles) single(s++);
        for (size_t p = 0; s < S; ++p) {
            const size_t K = take[2 * s + (p & 1)];
            if (K != 2 && K != 3) throw Logic("grouped sweep plan without a cover");
            setPartners(s, K, true);
            dir = (p & 1) ? kSweepReverse : 0U;
            widest = 0;
            g.clear();
            for (size_t t = 0; t + 1 < tile_begin.size(); ++t)
                emit(order.data() + tile_begin[t], tile_begin[t + 1] - tile_begin[t], K);
            setPartners(s, K, false);
            sweep_groups.push_back(g);
            sweep_first_stage.push_back(uint32_t(s));
            sweep_lds_nodes.push_back(widest);
            s += K;
        }
    }

    // Builds every device array of this lattice on `stream` (and waits for
    // the copies); returns the initial control block and sphere for the
    // engine to place at ctl_dev and sphere_dev.
    void upload(const blx_state_view &st, const Step<R> &S, const blx_sphere &sp, hipStream_t stream, DevCtl *ctl_dev,
                Sphere<R> *sphere_dev, DevCtl &ctl0, Sphere<R> &sph0) {
        if (st.node_count != sch.N || st.bond_count != sch.B) throw Invalid("state does not match the schedule");
        if (size_t(st.max_degree) * st.node_count > 0xffffffffULL) throw Invalid("neighbour slots exceed 32-bit indices");
        if (st.bond_count >= 0x80000000U) throw Invalid("at most 2^31 - 1 bonds");
        if (uint64_t(st.max_degree) * st.node_count >= 0x100000000ULL)
            throw Invalid("the node slots (max_degree x nodes) must be fewer than 2^32");
        if (S.sphere_enabled && (st.node_count + kWG - 1) / kWG > kMaxNodeTiles)
            throw Invalid("the striker pass holds at most 4,194,304 nodes");
        N = st.node_count;
        B = st.bond_count;
        D = st.max_degree;
        const size_t P = size_t(D) * N;
        node_tiles = (N + kWG - 1) / kWG;
        bond_tiles = (B + kWG - 1) / kWG;
        const uint32_t event_words = (B + 63) / 64;
        group_words = (event_words + 63) / 64;

        // Host arrays in device layout.
        std::vector<R4<R>> x04(N), u4(N), up4(N), v4(N);
        for (uint32_t i = 0; i < N; ++i) {
            const size_t k = 3 * size_t(i);
            x04[i] = {R(st.x0[k]), R(st.x0[k + 1]), R(st.x0[k + 2]), R(st.mass[i])};
            u4[i] = {R(st.u[k]), R(st.u[k + 1]), R(st.u[k + 2]), R(st.inv_mass[i])};
            up4[i] = {R(st.u_prev[k]), R(st.u_prev[k + 1]), R(st.u_prev[k + 2]), R(0)};
            v4[i] = {R(st.v[k]), R(st.v[k + 1]), R(st.v[k + 2]), R(0)};
        }
        mass_host.resize(N);
        inv_mass_host.resize(N);
        for (uint32_t i = 0; i < N; ++i) {
            mass_host[i] = R(st.mass[i]);
            inv_mass_host[i] = R(st.inv_mass[i]);
        }
        {
            auto x0m = std::make_shared<std::vector<double>>(size_t(N) * 4);
            for (uint32_t i = 0; i < N; ++i) {
                for (int c = 0; c < 3; ++c) (*x0m)[4 * size_t(i) + c] = st.x0[3 * size_t(i) + c];
                (*x0m)[4 * size_t(i) + 3] = st.mass[i];
            }
            x0m_host = std::move(x0m);
            mp_ready = false;
        }
        // Bonds in node-major order for the end sample: a counting sort by the
        // first end, stable in schedule index.
        std::vector<uint32_t> node_of(B), pos_of(B);
        {
            std::vector<uint32_t> start(size_t(N) + 1, 0U);
            for (uint32_t k = 0; k < B; ++k) {
                if (st.bond_a[k] >= N || st.bond_b[k] >= N) throw Invalid("bond end out of range");
                ++start[st.bond_a[k] + 1];
            }
            for (uint32_t i = 0; i < N; ++i) start[i + 1] += start[i];
            for (uint32_t k = 0; k < B; ++k) {
                const uint32_t o = start[st.bond_a[k]]++;
                node_of[o] = k;
                pos_of[k] = o;
            }
        }
        node_j_host = node_of;
        // Neighbour slots as slot palette indices.
        std::vector<uint32_t> key(P), nbond(P);
        std::vector<R4<R>> spal;
        std::vector<int32_t> sdelta;
        {
            constexpr size_t W = 4 * sizeof(R) / 4 + 1;
            Dedup<W> table;
            for (size_t k = 0; k < P; ++k) {
                if (st.nbr_bond[k] == BLX_NO_BOND) {
                    key[k] = kNone;
                    nbond[k] = kNone;
                    continue;
                }
                if (st.nbr_bond[k] >= B || st.nbr_other[k] >= N) throw Invalid("neighbour slot out of range");
                const uint32_t i = uint32_t(k % N);
                const R rw[4] = {R(st.nbr_rest[3 * k]), R(st.nbr_rest[3 * k + 1]), R(st.nbr_rest[3 * k + 2]),
                                 R(st.nbr_weight[k])};
                const int32_t delta = int32_t(int64_t(st.nbr_other[k]) - int64_t(i));
                typename Dedup<W>::Key kk{};
                std::memcpy(kk.data(), rw, sizeof(rw));
                std::memcpy(kk.data() + 4 * sizeof(R) / 4, &delta, 4);
                const uint32_t id = table.index(kk);
                if (id >= kDeadSlot) throw Overflow("more than 2^31 distinct neighbour slots");
                key[k] = id | (st.nbr_alive[k] ? 0U : kDeadSlot);
                nbond[k] = pos_of[st.nbr_bond[k]];
            }
            const auto &keys = table.keys();
            spal.resize(keys.size());
            sdelta.resize(keys.size());
            for (size_t q = 0; q < keys.size(); ++q) {
                std::memcpy(&spal[q], keys[q].data(), 4 * sizeof(R));
                std::memcpy(&sdelta[q], keys[q].data() + 4 * sizeof(R) / 4, 4);
            }
            slot_palette = uint32_t(keys.size());
        }
        // Edge palette: the distinct (rest edge, rest length, |e|^2 - L^2) in
        // backend precision; the direction terms are filled on the device.
        std::vector<uint32_t> pal_of(B);
        std::vector<EdgePal<R>> epal;
        {
            constexpr size_t W = 5 * sizeof(R) / 4;
            Dedup<W> table;
            for (uint32_t k = 0; k < B; ++k) {
                const R v[5] = {R(st.rest_edge[3 * size_t(k)]), R(st.rest_edge[3 * size_t(k) + 1]),
                                R(st.rest_edge[3 * size_t(k) + 2]), R(st.rest_length[k]), R(st.rest_length_sq_minus[k])};
                typename Dedup<W>::Key kk{};
                std::memcpy(kk.data(), v, sizeof(v));
                pal_of[k] = table.index(kk);
                if (pal_of[k] >= kNoDir) throw Overflow("more than 2^31 distinct bond edges");
            }
            const auto &keys = table.keys();
            epal.resize(keys.size());
            for (size_t q = 0; q < keys.size(); ++q) {
                R v[5];
                std::memcpy(v, keys[q].data(), sizeof(v));
                epal[q].e = {v[0], v[1], v[2], v[3]};
                epal[q].d = {v[4], R(0), R(0), R(0)};
            }
            edge_palette = uint32_t(keys.size());
        }
        bool tensile = true;
        for (uint32_t k = 0; k < B && tensile; ++k) {
            const double *t = st.threshold + 6 * size_t(k);
            tensile = !std::isfinite(R(t[2])) && !std::isfinite(R(t[4]));
        }
        tensile_only = tensile;

        std::vector<uint32_t> sweep_nodes;
        std::vector<PairBond<R>> sweep_bonds;
        buildSweep(st, sweep_nodes, sweep_bonds);

        // Layout in the arena.
        arena.reset();
        const uint32_t np = edge_palette;
        Off &o = off;
        o = Off{};
        o.x04 = arena.reserve<R4<R>>(N);
        o.u4 = arena.reserve<R4<R>>(N);
        o.u4b = arena.reserve<R4<R>>(N);
        o.stage_ranges = arena.reserve<uint2>(sweep_ranges.size());
        o.stage_end = arena.reserve<uint32_t>(stage_end.size());
        o.pnodes = arena.reserve<uint32_t>(sweep_nodes.size());
        o.pbonds = arena.reserve<PairBond<R>>(sweep_bonds.size());
        o.up4 = arena.reserve<R4<R>>(N);
        o.v4 = arena.reserve<R4<R>>(N);
        o.rinv = arena.reserve<R>(9 * size_t(N));
        o.unm = arena.reserve<R>(3 * size_t(N));
        o.sa = arena.reserve<R4<R>>(N);
        o.sb = arena.reserve<R4<R>>(N);
        o.valid = arena.reserve<uint8_t>(N);
        o.approach = arena.reserve<R>(2 * size_t(N));
        o.engaged = arena.reserve<uint8_t>(2 * size_t(N));
        o.cand = arena.reserve<uint8_t>(N);
        o.nbr_key = arena.reserve<uint32_t>(P);
        o.spal = arena.reserve<R4<R>>(spal.size());
        o.sdelta = arena.reserve<int32_t>(sdelta.size());
        o.nbr_bond = arena.reserve<uint32_t>(P);
        o.epal = arena.reserve<EdgePal<R>>(np);
        o.nodir = arena.reserve<uint32_t>(np);
        o.rec = arena.reserve<BondRec<R>>(B);
        o.cmpl = arena.reserve<R>(B);
        o.alive = arena.reserve<uint8_t>(B);
        o.lambda = arena.reserve<R>(B);
        o.pext = arena.reserve<R>(B);
        o.bnd = arena.reserve<BondNode>(B);
        o.alive_n = arena.reserve<uint8_t>(B);
        o.thr = arena.reserve<Thresholds<R>>(tensile ? 0 : B);
        o.thr2 = arena.reserve<Tensile<R>>(tensile ? B : 0);
        o.pv = arena.reserve<R2<R>>(B);
        o.pyz = arena.reserve<R2<R>>(B);
        o.mode = arena.reserve<uint8_t>(B);
        o.pstr = arena.reserve<R>(B);
        o.slots = arena.reserve<uint2>(B);
        o.ev_bits = arena.reserve<uint64_t>(event_words);
        o.brk_bits = arena.reserve<uint64_t>(event_words);
        o.grp_bits = arena.reserve<uint64_t>(group_words);
        o.ev_d = arena.reserve<float>(B);
        o.ev_e = arena.reserve<R>(B);
        o.flag_list = arena.reserve<uint32_t>(event_words);
        o.tile_cand = arena.reserve<uint32_t>(node_tiles);
        o.dirty_list = arena.reserve<uint32_t>(2 * size_t(N));  // two halves, by substep parity
        o.dirty_stamp = arena.reserve<uint32_t>(N);
        o.rank_stamp = arena.reserve<uint32_t>(N);
        o.affected = arena.reserve<uint32_t>(N);
        o.affected_list = arena.reserve<uint32_t>(N);
        o.tile_plastic = arena.reserve<double>(bond_tiles);
        o.load_part = arena.reserve<double>(13 * size_t(node_tiles));
        o.fext = arena.reserve<R4<R>>(N);
        o.freq = arena.reserve<double>(3 * size_t(N));
        o.cand_all = arena.reserve<uint32_t>(std::max<size_t>(N, size_t(sch.blocks) * kCap));
        o.cand_tile = arena.reserve<uint32_t>(size_t(node_tiles) * kWG);
        o.cand_lo = arena.reserve<uint32_t>(sch.blocks);
        o.cand_cnt = arena.reserve<uint32_t>(sch.blocks);
        o.nbb = arena.reserve<uint32_t>(sch.nbb.size());
        o.first_fail = arena.reserve<uint32_t>(B);
        // Node-node contact: the broad phase's cells, sort, cell table and pair lists.
        contact = S.nc.mode != 0U;
        const size_t nN = contact ? N : 0, nP = contact ? size_t(kPairsPerNode) * N : 0;
        uint32_t table = 64;
        while (contact && table < 2U * N) table <<= 1;
        nc_table_mask = table - 1;
        o.nc_build = arena.reserve<R4<R>>(nN);
        o.nc_key = arena.reserve<uint64_t>(nN);
        o.nc_bucket = arena.reserve<uint32_t>(nN);
        o.nc_bucket_sorted = arena.reserve<uint32_t>(nN);
        o.nc_node = arena.reserve<uint32_t>(nN);
        o.nc_node_sorted = arena.reserve<uint32_t>(nN);
        o.nc_table_val = arena.reserve<uint2>(contact ? table : 0);
        o.pair_fill = arena.reserve<uint32_t>(nN);
        o.pair_other = arena.reserve<uint32_t>(nP);
        o.pair_bond = arena.reserve<uint32_t>(nP);
        o.pair_offset = arena.reserve<uint32_t>(nN);
        o.flat_a = arena.reserve<uint32_t>(nP);
        o.flat_b = arena.reserve<uint32_t>(nP);
        o.flat_bond = arena.reserve<uint32_t>(nP);
        o.nc_partners = arena.reserve<uint32_t>(nN);
        o.nc_done = arena.reserve<uint32_t>(nN);
        o.nc_rank_b = arena.reserve<uint32_t>(nP);
        o.nc_sort_key = arena.reserve<uint32_t>(nP);
        o.nc_sort_in = arena.reserve<uint32_t>(nP);
        o.nc_sort_out = arena.reserve<uint32_t>(nP);
        o.nc_terms = arena.reserve<R4<R>>(nP);
        o.nc_mover = arena.reserve<uint8_t>(nP);
        arena.commit();
        if (contact) {
            size_t sort_bytes = 0, scan_bytes = 0;
            HIP_OK(rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<uint32_t *>(nullptr),
                                             static_cast<uint32_t *>(nullptr), static_cast<uint32_t *>(nullptr),
                                             static_cast<uint32_t *>(nullptr), N, 0, 32, stream));
            HIP_OK(rocprim::exclusive_scan(nullptr, scan_bytes, static_cast<uint32_t *>(nullptr),
                                           static_cast<uint32_t *>(nullptr), 0U, N, rocprim::plus<uint32_t>(), stream));
            size_t pair_bytes = 0;
            HIP_OK(rocprim::radix_sort_pairs(nullptr, pair_bytes, static_cast<uint32_t *>(nullptr),
                                             static_cast<uint32_t *>(nullptr), static_cast<uint32_t *>(nullptr),
                                             static_cast<uint32_t *>(nullptr), size_t(kPairsPerNode) * N, 0, 32, stream));
            nc_temp.alloc(std::max({sort_bytes, scan_bytes, pair_bytes}));
        }

        // Zero everything, then copy the data in.
        HIP_OK(hipMemsetAsync(arena.at<char>(0), 0, arena.used(), stream));
        auto put = [&](size_t off, const void *src, size_t bytes) {
            if (bytes) HIP_OK(hipMemcpyAsync(arena.at<char>(off), src, bytes, hipMemcpyHostToDevice, stream));
        };
        auto putv = [&](size_t off, const auto &v) { put(off, v.data(), v.size() * sizeof(v[0])); };
        std::vector<uint32_t> all(N);
        for (uint32_t i = 0; i < N; ++i) all[i] = i;
        putv(o.x04, x04);
        putv(o.u4, u4);
        putv(o.pnodes, sweep_nodes);
        putv(o.stage_ranges, sweep_ranges);
        putv(o.stage_end, stage_end);
        putv(o.up4, up4);
        putv(o.v4, v4);
        putv(o.nbr_key, key);
        putv(o.nbr_bond, nbond);
        putv(o.spal, spal);
        putv(o.sdelta, sdelta);
        putv(o.epal, epal);
        putv(o.nbb, sch.nbb);
        // Upload marks every node dirty, in the half substep 0's start sample reads.
        put(o.dirty_list + size_t(N) * sizeof(uint32_t), all.data(), all.size() * sizeof(uint32_t));
        // Palette directions on the device, then the bond arrays that carry
        // the no-direction flag.
        hipLaunchKernelGGL((k_edge_palette<R>), dim3(std::max(1U, std::min(4096U, (np + kWG - 1) / kWG))), dim3(kWG), 0,
                           stream, arena.at<EdgePal<R>>(o.epal), arena.at<uint32_t>(o.nodir), np);
        HIP_OK(hipGetLastError());
        std::vector<uint32_t> no_dir(np);
        HIP_OK(hipMemcpyAsync(no_dir.data(), arena.at<uint32_t>(o.nodir), size_t(np) * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for (uint32_t k = 0; k < B; ++k) pal_of[k] |= no_dir[pal_of[k]] ? kNoDir : 0U;
        {
            std::vector<BondRec<R>> rec(B);
            std::vector<R> cmpl(B);
            std::vector<BondNode> bnd(B);
            std::vector<R2<R>> pv(B), pyz(B);
            std::vector<R> pext(B), pstr(B);
            std::vector<uint2> slots(B);
            std::vector<uint8_t> alive_n(B), mode(B);
            std::vector<Thresholds<R>> thr(tensile ? 0 : B);
            std::vector<Tensile<R>> thr2(tensile ? B : 0);
            for (uint32_t k = 0; k < B; ++k) {
                BondRec<R> r{};
                r.a = st.bond_a[k];
                r.b = st.bond_b[k];
                r.p = pal_of[k];
                // The solve's compliance term alpha = compliance / dt^2
                // (LatticePhysics.hpp:1132-1168), the same float operations
                // the solve would repeat each substep.
                cmpl[k] = R(st.compliance[k]);
                r.c = cmpl[k] / (S.dt * S.dt);
                rec[k] = r;
                pext[k] = R(st.plastic_extension[k]);
            }
            for (uint32_t q = 0; q < B; ++q) {
                const uint32_t k = node_of[q];
                bnd[q] = BondNode{st.bond_a[k], st.bond_b[k], k, pal_of[k]};
                const double *t = st.threshold + 6 * size_t(k);
                if (tensile) thr2[q] = {R(t[0]), R(t[1])};
                else thr[q] = {R(t[0]), R(t[1]), R(t[2]), R(t[3]), R(t[4]), R(t[5])};
                pv[q] = {R(st.prev_tensile[k]), R(st.damage[k])};
                pyz[q] = {R(st.prev_compressive[k]), R(st.prev_shear[k])};
                pstr[q] = R(st.plastic_strain[k]);
                slots[q] = make_uint2(st.bond_slot_a[k], st.bond_slot_b[k]);
                alive_n[q] = st.alive[k];
                mode[q] = st.failure_mode[k];
            }
            putv(o.rec, rec);
            putv(o.cmpl, cmpl);
            for (PairBond<R> &e : sweep_bonds) {
                e.p = pal_of[e.j];
                e.c = rec[e.j].c;
            }
            putv(o.pbonds, sweep_bonds);
            putv(o.bnd, bnd);
            putv(o.thr, thr);
            putv(o.thr2, thr2);
            putv(o.pv, pv);
            putv(o.pyz, pyz);
            putv(o.pext, pext);
            putv(o.pstr, pstr);
            putv(o.slots, slots);
            put(o.alive, st.alive, B);
            putv(o.alive_n, alive_n);
            putv(o.mode, mode);
            HIP_OK(hipStreamSynchronize(stream));
        }

        used = arena.used();
        bind(ctl_dev, sphere_dev, st.origin);

        // Control block and sphere. A slot stays stopped until it is armed.
        DevCtl c{};
        c.first_failure_step = c.last_failure_step = c.last_energy_gain_step = ~0ULL;
        c.dirty_start[1] = 1;  // the half of parity 1: read by substep 0 and the upload's k_rinv
        c.dirty_count[1] = N;
        c.rank_slot = 1;  // booked by the first substep's k_finish
        c.nc_rebuild = 1;  // the first substep with contact on builds the pair list
        c.rinv_stamp = 0xffffffffU;
        c.flat_fraction = 1.0e-3;
        c.calm_margin = 0.5;
        c.stop = 1;
        ctl_host = c;
        ctl0 = c;
        sph0 = Sphere<R>{{R(sp.center[0]), R(sp.center[1]), R(sp.center[2])},
                         {R(sp.velocity[0]), R(sp.velocity[1]), R(sp.velocity[2])},
                         {R(sp.angular_velocity[0]), R(sp.angular_velocity[1]), R(sp.angular_velocity[2])},
                         R(sp.radius), R(sp.mass), R(sp.inertia)};
        loads_set = false;
        frames.clear();
        status = blx_run_status{};
        status.first_failure_step = status.last_failure_step = status.last_energy_gain_step = ~0ULL;
        uploaded = true;
        armed = false;
        pristine = true;
    }
};

// ---------------------------------------------------------------------------
// The engine: shared settings, streams, launch maps and the run loop.

template <class R> class Engine {
public:
    Engine(int device, uint32_t capacity) : device_(device), slots_(capacity) {
        if (capacity == 0) throw Invalid("a batch holds at least one lattice");
        HIP_OK(hipSetDevice(device_));
        HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&early_, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&lds_, hipStreamNonBlocking));
        // Only other streams of this device wait on these events, so their
        // release is device-scope; the default system-scope fence makes every
        // signal also write the caches back for the host (about 1% of the step).
        for (hipEvent_t *e : {&ev_bond_end_, &ev_rinv_, &ev_pass1_, &ev_velocity_, &ev_early_, &ev_predict_, &ev_start_,
                              &ev_sweep_, &ev_strain_, &ev_lds_})
            HIP_OK(hipEventCreateWithFlags(e, hipEventDisableTiming | hipEventDisableSystemFence));
        HIP_OK(hipEventCreate(&ev0_));
        HIP_OK(hipEventCreate(&ev1_));
        hipDeviceProp_t prop{};
        HIP_OK(hipGetDeviceProperties(&prop, device_));
        cus_ = uint32_t(prop.multiProcessorCount);
        lds_limit_ = prop.sharedMemPerBlock;
        for (auto &s : slots_) s = std::make_unique<Slot<R>>();
        ls_.alloc(size_t(capacity) * sizeof(DevLattice<R>));
        ctls_.alloc(size_t(capacity) * sizeof(DevCtl));
        spheres_.alloc(size_t(capacity) * sizeof(Sphere<R>));
        HIP_OK(hipMemset(ls_.p, 0, ls_.bytes));
        HIP_OK(hipMemset(ctls_.p, 0, ctls_.bytes));
        HIP_OK(hipMemset(spheres_.p, 0, spheres_.bytes));
    }
    ~Engine() {
        (void)hipSetDevice(device_);
        dropGraph();
        for (hipStream_t s : {stream_, side_, early_, lds_})
            if (s) (void)hipStreamSynchronize(s);
        for (hipEvent_t e : {ev_bond_end_, ev_rinv_, ev_pass1_, ev_velocity_, ev_early_, ev_predict_, ev_start_, ev_sweep_,
                             ev_strain_, ev_lds_, ev0_, ev1_})
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
        for (hipStream_t s : {lds_, early_, side_, stream_})
            if (s) (void)hipStreamDestroy(s);
    }

    uint32_t capacity() const { return uint32_t(slots_.size()); }
    int device() const { return device_; }
    std::string name() const { return std::string("hip-gfx950-") + (sizeof(R) == 4 ? "float" : "double"); }

    Slot<R> &slot(uint32_t k) {
        if (k >= slots_.size()) throw Invalid("slot index out of range");
        return *slots_[k];
    }
    const Slot<R> &slot(uint32_t k) const {
        if (k >= slots_.size()) throw Invalid("slot index out of range");
        return *slots_[k];
    }

    void setSchedule(uint32_t k, Schedule s) {
        slot(k).setSchedule(std::move(s));
        maps_dirty_ = true;
        updateShared();
    }

    // LatticeBackend::upload for one slot. The first upload into an empty
    // batch sets the shared settings; every later one must match them.
    void upload(uint32_t k, const blx_state_view &st, const blx_settings &s, const blx_sphere &sp) {
        HIP_OK(hipSetDevice(device_));
        Slot<R> &sl = slot(k);
        if (!sl.scheduled) throw Logic("slot has no schedule");
        validateSettings(s);
        const Step<R> S = convertStep<R>(s);
        std::vector<unsigned char> bytes(sizeof(Step<R>));
        std::memcpy(bytes.data(), &S, sizeof(S));
        bool others = false;
        for (uint32_t q = 0; q < slots_.size(); ++q)
            if (q != k && slots_[q]->uploaded) others = true;
        if (others && shared_.set) {
            if (bytes != shared_.step_bytes || (s.direct_arithmetic != 0) != shared_.direct ||
                (s.plastic_yield_stretch > 0.0) != shared_.plastic || s.constraint_iterations != shared_.iterations)
                throw Invalid("every lattice of a batch shares the step settings");
        } else {
            shared_.set = true;
            shared_.step_bytes = bytes;
            shared_.direct = s.direct_arithmetic != 0;
            shared_.plastic = s.plastic_yield_stretch > 0.0;
            shared_.iterations = s.constraint_iterations;
            const Blade<R> blade = S_.blade;
            S_ = S;
            S_.blade = blade;
        }
        HIP_OK(hipStreamSynchronize(stream_));
        sl.uploaded = false;
        DevCtl c0{};
        Sphere<R> s0{};
        sl.upload(st, S_, sp, stream_, ctls_.as<DevCtl>() + k, spheres_.as<Sphere<R>>() + k, c0, s0);
        HIP_OK(hipMemcpy(ctls_.as<DevCtl>() + k, &c0, sizeof(DevCtl), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(spheres_.as<Sphere<R>>() + k, &s0, sizeof(s0), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(ls_.as<DevLattice<R>>() + k, &sl.L, sizeof(DevLattice<R>), hipMemcpyHostToDevice));
        // Every node starts dirty: its inverse rest covariance before the first substep.
        {
            std::vector<WgMap> m;
            const uint32_t slices = std::max(1U, std::min(4096U, (sl.N + kWG - 1) / kWG));
            for (uint32_t q = 0; q < slices; ++q) m.push_back(WgMap{k, q, slices, 0});
            upload_map_.alloc(m.size() * sizeof(WgMap));
            HIP_OK(hipMemcpy(upload_map_.p, m.data(), m.size() * sizeof(WgMap), hipMemcpyHostToDevice));
            const Batch<R> b{ls_.as<DevLattice<R>>(), upload_map_.as<WgMap>(), uint32_t(m.size())};
            hipLaunchKernelGGL((k_rinv<R>), dim3(uint32_t(m.size())), dim3(kWG), 0, stream_, b);
            HIP_OK(hipGetLastError());
            HIP_OK(hipStreamSynchronize(stream_));
        }
        maps_dirty_ = true;
    }

    // Slot dst takes a copy of the lattice of slot src as it was uploaded
    // (src must not have run) with its own sphere: admitting a trial of a
    // lattice already on the device is a device-to-device copy.
    void clone(uint32_t dst, uint32_t src, const blx_sphere &sp) {
        HIP_OK(hipSetDevice(device_));
        if (dst == src) throw Invalid("a slot cannot be cloned into itself");
        Slot<R> &d = slot(dst);
        const Slot<R> &s = slot(src);
        if (!s.uploaded || !s.pristine) throw Logic("clone source must be an uploaded lattice that has not run");
        HIP_OK(hipStreamSynchronize(stream_));
        d.uploaded = false;
        d.cloneFrom(s, stream_, ctls_.as<DevCtl>() + dst, spheres_.as<Sphere<R>>() + dst);
        DevCtl c = readCtl(src);
        writeCtl(dst, c);
        d.ctl_host = c;
        const Sphere<R> s0{{R(sp.center[0]), R(sp.center[1]), R(sp.center[2])},
                           {R(sp.velocity[0]), R(sp.velocity[1]), R(sp.velocity[2])},
                           {R(sp.angular_velocity[0]), R(sp.angular_velocity[1]), R(sp.angular_velocity[2])},
                           R(sp.radius), R(sp.mass), R(sp.inertia)};
        HIP_OK(hipMemcpy(spheres_.as<Sphere<R>>() + dst, &s0, sizeof(s0), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(ls_.as<DevLattice<R>>() + dst, &d.L, sizeof(DevLattice<R>), hipMemcpyHostToDevice));
        maps_dirty_ = true;
    }

    // The kinematic blade of every slot (an extension of the support planes).
    void setBlade(const blx_blade &b) {
        HIP_OK(hipSetDevice(device_));
        HIP_OK(hipStreamSynchronize(stream_));
        Blade<R> o{};
        o.enabled = b.enabled ? 1U : 0U;
        if (o.enabled) {
            if (!(b.speed >= 0.0) || !std::isfinite(b.speed) || !(b.half_width >= 0.0) || !(b.node_radius >= 0.0))
                throw Invalid("blade speed, half width and node radius must be finite and non-negative");
            auto v = [](const double *x) { return V3<R>{R(x[0]), R(x[1]), R(x[2])}; };
            o.c0 = v(b.edge_point);
            o.f = v(b.flat_normal);
            o.m = v(b.direction);
            o.speed = R(b.speed);
            o.hw = R(b.half_width);
            o.r = R(b.node_radius);
        }
        S_.blade = o;
    }

    // The slot leaves the batch and keeps its arena for the next upload.
    void release(uint32_t k) {
        Slot<R> &sl = slot(k);
        HIP_OK(hipStreamSynchronize(stream_));
        sl.uploaded = false;
        sl.armed = false;
        maps_dirty_ = true;
        updateShared();
    }

    // ExternalLoads.hpp:28-50: validate the whole field, then replace the load.
    void setForces(uint32_t k, const double *f, uint64_t substeps) {
        Slot<R> &sl = slot(k);
        if (!sl.uploaded) throw Logic("external forces need an uploaded lattice");
        HIP_OK(hipSetDevice(device_));
        HIP_OK(hipStreamSynchronize(stream_));
        if (f == nullptr) {
            if (substeps != 0) throw Invalid("empty external load needs zero substeps");
            setRemaining(k, 0);
            sl.loads_set = false;
            maps_dirty_ = true;
            return;
        }
        if (substeps == 0) throw Invalid("external load needs one force per node and a finite nonzero span");
        const uint32_t N = sl.N;
        std::vector<R4<R>> fe(N);
        std::vector<double> fr(3 * size_t(N));
        bool any = false;
        for (uint32_t i = 0; i < N; ++i) {
            const double x = f[3 * size_t(i)], y = f[3 * size_t(i) + 1], z = f[3 * size_t(i) + 2];
            const R rx = R(x), ry = R(y), rz = R(z);
            if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z) || !std::isfinite(std::hypot(x, y, z)) ||
                !std::isfinite(rx) || !std::isfinite(ry) || !std::isfinite(rz))
                throw Invalid("external force is not finite in backend precision");
            if (x == 0 && y == 0 && z == 0) {
                fe[i] = {R(0), R(0), R(0), R(0)};
                continue;
            }
            if (!(sl.mass_host[i] > 0) || !(sl.inv_mass_host[i] > 0) || !std::isfinite(sl.mass_host[i]) ||
                !std::isfinite(sl.inv_mass_host[i]))
                throw Invalid("external force needs a finite positive movable node mass");
            fe[i] = {rx, ry, rz, R(1)};
            fr[3 * size_t(i)] = x;
            fr[3 * size_t(i) + 1] = y;
            fr[3 * size_t(i) + 2] = z;
            any = true;
        }
        HIP_OK(hipMemcpy(sl.L.fext, fe.data(), fe.size() * sizeof(R4<R>), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(const_cast<double *>(sl.L.freq), fr.data(), fr.size() * sizeof(double), hipMemcpyHostToDevice));
        sl.loads_set = any;
        maps_dirty_ = true;
        setRemaining(k, any ? substeps : 0);
    }

    // The run control of a slot's next advance: its exit rules and its step
    // budget (RunControl, FastLattice.hpp:79-128).
    void arm(uint32_t k, const blx_run_control &c) {
        Slot<R> &sl = slot(k);
        if (!sl.uploaded) throw Logic("run before upload");
        HIP_OK(hipSetDevice(device_));
        HIP_OK(hipStreamSynchronize(stream_));
        DevCtl ctl = readCtl(k);
        ctl.stop = c.max_steps == 0 ? 1U : 0U;
        ctl.exit_reason = 0;
        ctl.budget = c.max_steps;
        ctl.quiet = c.quiet_steps;
        ctl.min_steps = c.min_steps;
        ctl.no_failure = c.no_failure_steps;
        ctl.energy_flat = c.energy_flat_steps;
        ctl.calm = c.calm_steps;
        ctl.flat_fraction = c.energy_flat_fraction;
        ctl.calm_margin = c.calm_damage_margin;
        ctl.removable = c.removable_energy_j;
        writeCtl(k, ctl);
        sl.ctl_host = ctl;
        sl.armed = c.max_steps > 0;
        sl.pristine = false;
        maps_dirty_ = true;
    }

    // Advances every armed slot until it stops (its exit rule or its budget)
    // or `substeps` substeps have been enqueued; returns the number of slots
    // still running. Abandon is polled between device batches.
    // RunControl::abandon (FastLattice.hpp:122-127), read between device
    // batches: a byte flag, or a caller's poll function (an adapter holding a
    // std::atomic<bool> passes one that loads it).
    struct Abandon {
        const volatile uint8_t *flag = nullptr;
        int (*poll)(const void *) = nullptr;
        const void *context = nullptr;
        bool operator()() const { return (flag && *flag) || (poll && poll(context) != 0); }
    };

    uint32_t advance(uint64_t substeps, uint64_t batch_len, const Abandon &abandon, double *kernel) {
        HIP_OK(hipSetDevice(device_));
        dropGraph();  // settings, maps or slots may have changed since the last call
        uint64_t done = 0;
        uint32_t running = countRunning();
        double ksec = 0.0;
        while (running > 0 && done < substeps) {
            if (maps_dirty_) {
                // The captured substeps hold the old maps' offsets, counts and
                // grids as launch arguments: capture again after a rebuild.
                buildMaps();
                dropGraph();
            }
            const uint64_t n = std::min<uint64_t>(profilingBatch(batch_len), substeps - done);
            HIP_OK(hipEventRecord(ev0_, stream_));
            uint64_t q = 0;
            const uint64_t gl = tuning_.graph_substeps;
            if (gl != 0 && S_.nc.mode == 0U && !prof_ && n >= gl) {
                if (!graph_exec_ || graph_len_ != gl) captureSubsteps(gl);
                for (; q + gl <= n; q += gl) HIP_OK(hipGraphLaunch(graph_exec_, stream_));
            }
            const uint64_t q_sample = q;
            for (; q < n; ++q) {
                // The batch's first enqueued substep is timed phase by phase
                // (phase_seconds); the rest run without markers.
                sample_ = !prof_ && q == q_sample;
                enqueueSubstepA();
                if (S_.nc.mode != 0U) {
                    // The rebuild decision of each slot, read back once a substep.
                    HIP_OK(hipStreamSynchronize(stream_));
                    readAllCtls();
                    rebuildContacts();
                }
                enqueueSubstepB();
                sample_ = false;
            }
            // side_ ends each batch joined: its last k_rinv is done when the batch is.
            HIP_OK(hipStreamWaitEvent(stream_, ev_rinv_, 0));
            HIP_OK(hipEventRecord(ev1_, stream_));
            HIP_OK(hipStreamSynchronize(stream_));
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, ev0_, ev1_));
            ksec += double(ms) * 1e-3;
            profCollect(double(ms) * 1e-3);
            ++launches_;
            done += n;
            readAllCtls();
            uint32_t now = 0;
            for (auto &s : slots_) {
                Slot<R> &sl = *s;
                if (!sl.armed) continue;
                if (sl.ctl_host.error) throw Overflow("external force kick exceeds backend precision");
                sl.status.kernel_seconds += double(ms) * 1e-3;
                ++sl.status.launches;
                if (sl.ctl_host.stop) {
                    sl.armed = false;
                    maps_dirty_ = true;
                } else {
                    ++now;
                }
            }
            running = now;
            if (abandon()) break;
        }
        if (kernel) *kernel = ksec;
        return running;
    }

    // LatticeBackend::run on slot 0 of a single-lattice engine, with frames.
    void run(const blx_run_control &c, blx_run_status &out) {
        Slot<R> &sl = slot(0);
        if (!sl.uploaded) throw Logic("run before upload");
        HIP_OK(hipSetDevice(device_));
        const auto t0 = std::chrono::steady_clock::now();
        arm(0, c);
        // RunControl::steps_per_launch (FastLattice.hpp:120-121): substeps per
        // device batch, 0 for the whole run in one. The host enqueues a batch
        // without waiting on the device; a run that stops early (its exit
        // rule) turns the rest of its batch into no-ops, and abandon is polled
        // between batches.
        const uint64_t batch = c.steps_per_launch ? c.steps_per_launch : std::max<uint64_t>(1, c.max_steps);
        const Abandon abandon{c.abandon, c.abandon_poll, c.abandon_context};
        uint64_t done = 0;
        unsigned reason = 0;
        while (done < c.max_steps) {
            const uint64_t step = sl.ctl_host.step;
            uint64_t n = c.max_steps - done;
            if (c.capture_stride != 0) {
                if (step % c.capture_stride == 0 && sl.status.frames_captured < c.max_frames) {
                    captureFrame(0, step);
                    ++sl.status.frames_captured;
                }
                n = std::min(n, c.capture_stride - step % c.capture_stride);
            }
            advance(n, batch, abandon, nullptr);
            done += sl.ctl_host.step - step;
            if (sl.ctl_host.stop && sl.ctl_host.exit_reason != kBudgetSpent) {
                reason = sl.ctl_host.exit_reason;
                break;
            }
            if (done < c.max_steps && abandon()) {
                reason = 7;
                break;
            }
        }
        if (reason == 0 && done >= c.max_steps) reason = 3;
        if (sl.armed) {
            sl.armed = false;
            maps_dirty_ = true;
        }
        sl.status.exit_reason = reason;
        sl.status.wall_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        status(0, out);
        out.exit_reason = reason;
    }

    // Launch tuning. Every setting gives the same bits.
    const blx_tuning &tuning() const { return tuning_; }
    void setTuning(const blx_tuning &t) {
        if (t.colours_per_launch != 1 && t.colours_per_launch != 2)
            throw Invalid("colours_per_launch is 1 or 2");
        if (t.lds_threads != 256 && t.lds_threads != 512 && t.lds_threads != 1024)
            throw Invalid("lds_threads is 256, 512 or 1024");
        if (t.graph_substeps > kMaxGraphSubsteps) throw Invalid("graph_substeps is at most 256");
        if (t.streams > 2) throw Invalid("streams is 0, 1 or 2");
        tuning_ = t;
        maps_dirty_ = true;
        dropGraph();
    }

    // The autotuner: every candidate tuning advances the armed slots (after
    // `prepare`, which may arm them) from the same state, once by warm
    // substeps and once by warm + probe, on the host clock; the difference is
    // the probe's time without the set-up both pay (a graph capture, a map
    // build). Two rounds over the candidates, and the fastest is kept. The
    // state of every slot, on the device and on the host, is put back as it
    // was before, so tuning changes no result. Takes a device copy of every
    // uploaded slot.
    template <class Prepare> void autotune(uint64_t warm, uint64_t probe, Prepare prepare, blx_autotune_report &rep) {
        if (probe == 0) throw Invalid("the autotuner needs at least one probe substep");
        HIP_OK(hipSetDevice(device_));
        HIP_OK(hipStreamSynchronize(stream_));
        HIP_OK(hipStreamSynchronize(side_));
        HIP_OK(hipStreamSynchronize(early_));
        struct Saved {
            DevBuf arena;
            DevCtl ctl;
            Sphere<R> sphere;
            blx_run_status status;
            bool armed, pristine;
        };
        std::vector<std::unique_ptr<Saved>> saved(slots_.size());
        size_t need = 0;
        for (auto &sl : slots_)
            if (sl->uploaded) need += sl->used;
        size_t free_bytes = 0, total_bytes = 0;
        HIP_OK(hipMemGetInfo(&free_bytes, &total_bytes));
        if (need + (size_t(64) << 20) > free_bytes)
            throw Runtime("the autotuner's copy of the lattices (" + std::to_string(need >> 20) +
                          " MiB) does not fit in free device memory");
        for (uint32_t k = 0; k < slots_.size(); ++k) {
            Slot<R> &sl = *slots_[k];
            if (!sl.uploaded) continue;
            auto sv = std::make_unique<Saved>();
            sv->arena.alloc(sl.used);
            HIP_OK(hipMemcpyAsync(sv->arena.p, sl.arena.template at<char>(0), sl.used, hipMemcpyDeviceToDevice, stream_));
            sv->ctl = readCtl(k);
            sv->sphere = readSphere(k);
            sv->status = sl.status;
            sv->armed = sl.armed;
            sv->pristine = sl.pristine;
            saved[k] = std::move(sv);
        }
        HIP_OK(hipStreamSynchronize(stream_));
        auto restore = [&] {
            HIP_OK(hipStreamSynchronize(stream_));
            for (uint32_t k = 0; k < slots_.size(); ++k) {
                if (!saved[k]) continue;
                Slot<R> &sl = *slots_[k];
                const Saved &sv = *saved[k];
                HIP_OK(hipMemcpyAsync(sl.arena.template at<char>(0), sv.arena.p, sl.used, hipMemcpyDeviceToDevice,
                                      stream_));
                writeCtl(k, sv.ctl);
                HIP_OK(hipMemcpy(spheres_.as<Sphere<R>>() + k, &sv.sphere, sizeof(Sphere<R>), hipMemcpyHostToDevice));
                sl.ctl_host = sv.ctl;
                sl.status = sv.status;
                sl.armed = sv.armed;
                sl.pristine = sv.pristine;
            }
            HIP_OK(hipStreamSynchronize(stream_));
            maps_dirty_ = true;
            dropGraph();
        };
        // The candidates: colour stages per launch, the LDS sweep and its
        // workgroup size where a lattice fits, and substeps per graph launch.
        const blx_tuning before = tuning_;
        bool any_fits = false, all_fit = true, any_run = false;
        prepare();
        for (auto &sl : slots_) {
            if (!sl->armed) continue;
            any_run = true;
            blx_tuning t = tuning_;
            t.lds_sweep = 1U;
            std::swap(t, tuning_);
            const bool fits = ldsSweeps(*sl);
            std::swap(t, tuning_);
            any_fits |= fits;
            all_fit &= fits;
        }
        restore();
        if (!any_run) throw Logic("the autotuner needs an armed lattice");
        std::vector<blx_tuning> cand;
        const uint32_t graphs[] = {0U, 8U, 32U};
        for (uint32_t g : graphs) {
            if (g != 0 && (g > warm + probe || S_.nc.mode != 0U)) continue;
            for (uint32_t c = 1; c <= 2; ++c) cand.push_back(blx_tuning{c, 0U, 1024U, g, tuning_.streams});
            if (any_fits)
                for (uint32_t th : {256U, 512U, 1024U})
                    for (uint32_t c = 1; c <= (all_fit ? 1U : 2U); ++c)
                        cand.push_back(blx_tuning{c, 1U, th, g, tuning_.streams});
        }
        if (cand.size() > BLX_TUNE_MAX) throw Logic("too many tuning candidates");
        std::vector<double> best(cand.size(), std::numeric_limits<double>::infinity());
        const Abandon none{};
        auto timed = [&](uint64_t n) {
            prepare();
            const auto t0 = std::chrono::steady_clock::now();
            advance(n, n, none, nullptr);
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            restore();
            return sec;
        };
        try {
            for (int round = 0; round < 2; ++round)
                for (size_t i = 0; i < cand.size(); ++i) {
                    setTuning(cand[i]);
                    const double a = warm ? timed(warm) : 0.0;
                    const double b = timed(warm + probe);
                    best[i] = std::min(best[i], std::max(0.0, b - a) / double(probe));
                }
        } catch (...) {
            restore();
            tuning_ = before;
            throw;
        }
        size_t pick = 0;
        for (size_t i = 1; i < cand.size(); ++i)
            if (best[i] < best[pick]) pick = i;
        setTuning(cand[pick]);
        rep = blx_autotune_report{};
        rep.chosen = cand[pick];
        rep.candidates = uint32_t(cand.size());
        for (size_t i = 0; i < cand.size(); ++i) {
            rep.tried[i] = cand[i];
            rep.seconds_per_substep[i] = best[i];
        }
    }

    void status(uint32_t k, blx_run_status &o) const {
        const Slot<R> &sl = slot(k);
        const DevCtl &c = sl.ctl_host;
        o = sl.status;
        // RunStatus::shared_memory_* (FastLattice.hpp:229-231): whether this
        // lattice's sweep keeps its nodes in LDS, and how many bytes.
        o.shared_memory_positions = ldsSweeps(sl) ? 1U : 0U;
        o.shared_memory_bytes = o.shared_memory_positions ? uint64_t(sl.N) * sizeof(R4<R>) : 0U;
        o.total_steps = c.step;
        o.first_failure_step = c.first_failure_step;
        o.last_failure_step = c.last_failure_step;
        o.failure_rounds = c.failure_rounds;
        o.broken_bonds = c.broken_bonds;
        o.removed_energy_j = c.removed_energy_j;
        o.last_energy_gain_step = c.last_energy_gain_step;
        o.max_damage = c.max_damage;
        o.last_damage_gain_step = c.last_damage_gain_step;
        o.contact.impulse_contacts = c.impulse_contacts;
        for (int q = 0; q < 3; ++q) o.contact.impulse_to_material[q] = c.impulse[q];
        o.contact.dissipated_kinetic_energy_j = c.dissipated;
        o.contact.maximum_penetration_m = c.max_penetration;
        o.contact.maximum_position_correction_m = c.max_correction;
        o.contact.maximum_center_shift_m = c.max_shift;
        o.contact.candidate_overflow = c.candidate_overflow;
        o.contact.ball_support_events = c.ball_support_events;
        o.external_load.steps = c.ledger_steps;
        o.external_load.elapsed_s = c.ledger_elapsed;
        for (int q = 0; q < 3; ++q) {
            o.external_load.requested_impulse_n_s[q] = c.ledger_req[q];
            o.external_load.impulse_n_s[q] = c.ledger_imp[q];
            o.external_load.requested_angular_impulse_kg_m2_s[q] = c.ledger_areq[q];
            o.external_load.angular_impulse_kg_m2_s[q] = c.ledger_aimp[q];
        }
        o.external_load.work_j = c.ledger_work;
        std::memcpy(&o.max_tensile_stretch, &c.peak_bits[0], 4);
        std::memcpy(&o.max_compressive_strain, &c.peak_bits[1], 4);
        std::memcpy(&o.max_shear_strain, &c.peak_bits[2], 4);
        std::memcpy(&o.max_plastic_stretch, &c.pstretch_bits, 4);
        o.rank_deficient_nodes = c.rank_deficient;
        o.plastic_work_j = c.plastic_work;
        o.node_contact.contacts = c.nc_contacts;
        o.node_contact.pair_tests = c.nc_tests;
        o.node_contact.rebuilds = c.nc_rebuilds;
        o.node_contact.pairs_listed = c.nc_listed;
        o.node_contact.pair_overflow = c.nc_overflow;
        o.node_contact.dissipated_kinetic_energy_j = c.nc_dissipated;
        o.node_contact.maximum_overlap_m = c.nc_overlap;
        o.node_contact.maximum_position_correction_m = c.nc_correction;
        for (int q = 0; q < 3; ++q) o.node_contact.momentum_residual[q] = c.nc_momentum[q];
        // A batch slot reports how its last advance ended; a spent budget is
        // the reference's max-steps reason.
        if (c.stop && c.exit_reason) o.exit_reason = c.exit_reason == kBudgetSpent ? 3U : c.exit_reason;
    }

    // The exit reason status() reports, for slots 0 .. count - 1 (0 while a
    // slot runs or holds no lattice).
    void exitReasons(uint32_t *out, uint32_t count) const {
        if (count > slots_.size()) throw Invalid("more slots asked for than the batch holds");
        for (uint32_t k = 0; k < count; ++k) {
            const Slot<R> &sl = *slots_[k];
            const DevCtl &c = sl.ctl_host;
            uint32_t r = sl.status.exit_reason;
            if (c.stop && c.exit_reason) r = c.exit_reason == kBudgetSpent ? 3U : c.exit_reason;
            out[k] = r;
        }
    }

    void download(uint32_t k, blx_state_view &st, blx_sphere &sp) {
        Slot<R> &sl = slot(k);
        if (!sl.uploaded) throw Logic("download before upload");
        HIP_OK(hipSetDevice(device_));
        HIP_OK(hipStreamSynchronize(stream_));
        const uint32_t N = sl.N, B = sl.B;
        if (sl.tensile_only && B > 0 && readCtl(k).step > 0) {
            // The live bonds' compressive and shear samples are not kept on a
            // tensile-only lattice: rebuild what the last end sample stored.
            if (shared_.direct)
                hipLaunchKernelGGL((k_prev_sample<R, true>), dim3(gridFor(B)), dim3(kWG), 0, stream_, sl.L);
            else
                hipLaunchKernelGGL((k_prev_sample<R, false>), dim3(gridFor(B)), dim3(kWG), 0, stream_, sl.L);
            HIP_OK(hipGetLastError());
            HIP_OK(hipStreamSynchronize(stream_));
        }
        std::vector<R4<R>> u4(N), up4(N), v4(N);
        std::vector<R2<R>> pv(B), pyz(B);
        std::vector<R> pext(B), pstr(B);
        std::vector<uint8_t> mode(B);
        copyDown(u4, sl.L.u4);
        copyDown(up4, sl.L.up4);
        copyDown(v4, sl.L.v4);
        copyDown(pv, sl.L.pv);
        copyDown(pyz, sl.L.pyz);
        copyDown(pext, sl.L.pext);
        copyDown(pstr, sl.L.pstr);
        copyDown(mode, sl.L.mode);
        HIP_OK(hipMemcpy(st.alive, sl.L.alive, B, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < N; ++i) {
            const size_t q = 3 * size_t(i);
            st.u[q] = double(u4[i].x);
            st.u[q + 1] = double(u4[i].y);
            st.u[q + 2] = double(u4[i].z);
            st.u_prev[q] = double(up4[i].x);
            st.u_prev[q + 1] = double(up4[i].y);
            st.u_prev[q + 2] = double(up4[i].z);
            st.v[q] = double(v4[i].x);
            st.v[q + 1] = double(v4[i].y);
            st.v[q + 2] = double(v4[i].z);
        }
        for (uint32_t q = 0; q < B; ++q) st.plastic_extension[q] = double(pext[q]);
        // The end-sample arrays are node-major on the device.
        for (uint32_t o = 0; o < B; ++o) {
            const uint32_t q = sl.node_j_host[o];
            st.prev_tensile[q] = double(pv[o].x);
            st.prev_compressive[q] = double(pyz[o].x);
            st.prev_shear[q] = double(pyz[o].y);
            st.damage[q] = double(pv[o].y);
            st.plastic_strain[q] = double(pstr[o]);
            st.failure_mode[q] = mode[o];
        }
        sphereOut(readSphere(k), sp);
    }

    void sphereOf(uint32_t k, blx_sphere &sp) {
        if (!slot(k).uploaded) throw Logic("sphere of an empty slot");
        HIP_OK(hipSetDevice(device_));
        HIP_OK(hipStreamSynchronize(stream_));
        sphereOut(readSphere(k), sp);
    }

    uint64_t firstFailure(uint32_t k, uint32_t *out, uint64_t cap) const {
        const Slot<R> &sl = slot(k);
        const uint64_t n = sl.ctl_host.first_failure_count;
        if (out && cap)
            HIP_OK(hipMemcpy(out, sl.L.first_fail, std::min(n, cap) * sizeof(uint32_t), hipMemcpyDeviceToHost));
        return n;
    }

    uint32_t frameCount(uint32_t k) const { return uint32_t(slot(k).frames.size()); }
    void takeFrame(uint32_t k, uint32_t f, uint64_t *step, float *u, uint8_t *alive, float *damage, blx_sphere *sp) {
        const Slot<R> &sl = slot(k);
        if (f >= sl.frames.size()) throw Invalid("frame index out of range");
        const HostFrame<R> &fr = sl.frames[f];
        *step = fr.step;
        std::copy(fr.u.begin(), fr.u.end(), u);
        std::copy(fr.alive.begin(), fr.alive.end(), alive);
        std::copy(fr.damage.begin(), fr.damage.end(), damage);
        if (sp) sphereOut(fr.sphere, *sp);
    }
    void clearFrames(uint32_t k) {
        slot(k).frames.clear();
        slot(k).status.frames_captured = 0;
    }

    // The device buffers of one label computation: N-node parent, root, size
    // and rank arrays, N 64-bit sort keys twice, the component counter and the
    // radix sort's temporary storage.
    struct LabelScratch {
        uint32_t *parent, *root, *size, *rank, *count;
        uint64_t *keys, *sorted;
        void *temp;
        size_t temp_bytes;
    };
    // The radix sort's temporary bytes for n keys (a size query; allocates nothing).
    size_t labelTempBytes(uint32_t n) {
        size_t temp = 0;
        HIP_OK(rocprim::radix_sort_keys(nullptr, temp, static_cast<uint64_t *>(nullptr),
                                        static_cast<uint64_t *>(nullptr), n, 0, 64, stream_));
        return temp;
    }

    // Union-find, roots, canonical rank and labels of one slot, on the device,
    // in the buffers of s; the labels go to label_dev (device memory).
    uint32_t labelsInto(uint32_t k, const LabelScratch &s, uint32_t *label_dev) {
        Slot<R> &sl = slot(k);
        const uint32_t N = sl.N;
        const uint32_t g = gridFor(N);
        hipLaunchKernelGGL(k_cc_zero, dim3(1), dim3(64), 0, stream_, s.count);
        const uint32_t ranges = std::max(1U, (N + kCcRange - 1) / kCcRange);
        hipLaunchKernelGGL(k_cc_local<R>, dim3(ranges), dim3(kCcRange), 0, stream_, sl.L, s.parent, s.size);
        hipLaunchKernelGGL(k_cc_cross<R>, dim3(ranges), dim3(kCcRange), 0, stream_, sl.L, s.parent);
        hipLaunchKernelGGL(k_cc_flatten<true>, dim3(gridFew(N)), dim3(kWG), 0, stream_, s.parent, s.root, s.size, N);
        hipLaunchKernelGGL(k_cc_keys, dim3(g), dim3(kWG), 0, stream_, s.root, s.size, s.keys, s.count, N);
        HIP_OK(hipGetLastError());
        uint32_t c = 0;
        HIP_OK(hipMemcpyAsync(&c, s.count, 4, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        size_t bytes = s.temp_bytes;
        HIP_OK(rocprim::radix_sort_keys(s.temp, bytes, s.keys, s.sorted, c, 0, 64, stream_));
        hipLaunchKernelGGL(k_cc_rank, dim3(gridFor(c)), dim3(kWG), 0, stream_, s.sorted, s.rank, c);
        hipLaunchKernelGGL(k_cc_label, dim3(g), dim3(kWG), 0, stream_, s.root, s.rank, label_dev, N);
        HIP_OK(hipGetLastError());
        return c;
    }

    // Labels of one slot in the slot's own buffers (allocated on first use);
    // label (host memory, or NULL) receives them.
    uint32_t labels(uint32_t k, uint32_t *label) {
        Slot<R> &sl = slot(k);
        if (!sl.uploaded) throw Logic("fragments before upload");
        HIP_OK(hipSetDevice(device_));
        const uint32_t N = sl.N;
        if (N == 0) return 0;
        const size_t n4 = size_t(N) * 4;
        sl.cc_parent.alloc(n4);
        sl.cc_root.alloc(n4);
        sl.cc_size.alloc(n4);
        sl.cc_rank.alloc(n4);
        sl.cc_label.alloc(n4);
        sl.cc_keys.alloc(size_t(N) * 8);
        sl.cc_sorted.alloc(size_t(N) * 8);
        sl.cc_count.alloc(4);
        sl.cc_temp.alloc(labelTempBytes(N));
        const LabelScratch s{sl.cc_parent.template as<uint32_t>(), sl.cc_root.template as<uint32_t>(),
                             sl.cc_size.template as<uint32_t>(), sl.cc_rank.template as<uint32_t>(),
                             sl.cc_count.template as<uint32_t>(), sl.cc_keys.template as<uint64_t>(),
                             sl.cc_sorted.template as<uint64_t>(), sl.cc_temp.p, sl.cc_temp.bytes};
        const uint32_t c = labelsInto(k, s, sl.cc_label.template as<uint32_t>());
        if (label) HIP_OK(hipMemcpyAsync(label, sl.cc_label.p, n4, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        return c;
    }

    // Labels of one slot with every buffer carved from `scratch` (bytes long,
    // device memory): allocates nothing. label_dev is the caller's device buffer.
    uint32_t labelsNoAlloc(uint32_t k, void *scratch, size_t bytes, uint32_t max_nodes, uint32_t *label_dev) {
        Slot<R> &sl = slot(k);
        if (!sl.uploaded) throw Logic("fragments before upload");
        if (!label_dev) throw Invalid("the labels need a device buffer of node_count entries");
        const uint32_t N = sl.N;
        if (N > max_nodes)
            throw Invalid("the allocation-free labels hold at most " + std::to_string(max_nodes) + " nodes; this lattice has " +
                          std::to_string(N));
        HIP_OK(hipSetDevice(device_));
        if (N == 0) return 0;
        // Carve at the maximum size, so every call uses the same offsets.
        auto at = [&, off = size_t(0)](size_t n) mutable {
            void *p = static_cast<char *>(scratch) + off;
            off += (n + 255) & ~size_t(255);
            return p;
        };
        const size_t m4 = size_t(max_nodes) * 4, m8 = size_t(max_nodes) * 8;
        LabelScratch s{};
        s.parent = static_cast<uint32_t *>(at(m4));
        s.root = static_cast<uint32_t *>(at(m4));
        s.size = static_cast<uint32_t *>(at(m4));
        s.rank = static_cast<uint32_t *>(at(m4));
        s.keys = static_cast<uint64_t *>(at(m8));
        s.sorted = static_cast<uint64_t *>(at(m8));
        s.count = static_cast<uint32_t *>(at(4));
        const size_t used = 4 * ((m4 + 255) & ~size_t(255)) + 2 * ((m8 + 255) & ~size_t(255)) + 256;
        const size_t temp = labelTempBytes(N);
        if (used + temp > bytes) throw Overflow("the radix sort needs more temporary storage than the scratch array holds");
        s.temp = static_cast<char *>(scratch) + used;
        s.temp_bytes = bytes - used;
        const uint32_t c = labelsInto(k, s, label_dev);
        HIP_OK(hipStreamSynchronize(stream_));
        return c;
    }

    // calculateFragmentMassProperties (FragmentMassProperties.cpp:7-80) of
    // every fragment of slot k, on the device: canonical labels, a stable
    // sort of the nodes by label, two fixed-order segmented sums and the
    // per-fragment closing algebra. Returns the fragment count; out[f] for
    // f < cap.
    uint32_t massProperties(uint32_t k, double cell, const double *spin, blx_fragment_mass *out, uint32_t cap) {
        const uint32_t c = labels(k, nullptr);
        Slot<R> &sl = slot(k);
        const uint32_t N = sl.N;
        if (c == 0) return 0;
        if (!sl.x0m_host || sl.x0m_host->size() != size_t(N) * 4) throw Logic("fragment mass before upload");
        if (!sl.mp_ready) {
            sl.mp_x0m.alloc(size_t(N) * 32);
            HIP_OK(hipMemcpy(sl.mp_x0m.p, sl.x0m_host->data(), size_t(N) * 32, hipMemcpyHostToDevice));
            sl.mp_ready = true;
        }
        if (spin) {
            sl.mp_spin.alloc(size_t(N) * 24);
            HIP_OK(hipMemcpyAsync(sl.mp_spin.p, spin, size_t(N) * 24, hipMemcpyHostToDevice, stream_));
        }
        const uint32_t tiles = (N + kMpTile - 1) / kMpTile;
        const size_t n4 = size_t(N) * 4;
        sl.mp_key.alloc(n4);
        sl.mp_order.alloc(n4);
        sl.mp_iota.alloc(n4);
        sl.mp_bounds.alloc(size_t(c) * 8);
        sl.mp_sum.alloc(size_t(c) * 10 * 8);
        sl.mp_first.alloc(size_t(tiles) * 10 * 8);
        sl.mp_last.alloc(size_t(tiles) * 10 * 8);
        sl.mp_cross.alloc(size_t(tiles) * 4);
        sl.mp_frag.alloc(size_t(c) * sizeof(blx_fragment_mass));
        sl.mp_bad.alloc(4);
        const unsigned bits = std::max(1U, unsigned(std::bit_width(c - 1)));
        uint32_t *label = sl.cc_label.template as<uint32_t>(), *key = sl.mp_key.template as<uint32_t>(),
                 *iota = sl.mp_iota.template as<uint32_t>(), *order = sl.mp_order.template as<uint32_t>();
        size_t temp = 0;
        HIP_OK(rocprim::radix_sort_pairs(nullptr, temp, label, key, iota, order, N, 0, bits, stream_));
        sl.mp_temp.alloc(temp);
        temp = sl.mp_temp.bytes;
        hipLaunchKernelGGL(k_nc_iota, dim3(gridFor(N)), dim3(kWG), 0, stream_, iota, N);
        HIP_OK(rocprim::radix_sort_pairs(sl.mp_temp.p, temp, label, key, iota, order, N, 0, bits, stream_));
        const uint32_t bad0 = kNone;
        HIP_OK(hipMemcpyAsync(sl.mp_bad.p, &bad0, 4, hipMemcpyHostToDevice, stream_));
        MpArgs<R> a{};
        a.u4 = sl.L.u4;
        a.v4 = sl.L.v4;
        a.x0m = sl.mp_x0m.template as<double4>();
        a.spin = spin ? sl.mp_spin.template as<double>() : nullptr;
        a.key = key;
        a.order = order;
        a.bounds = sl.mp_bounds.template as<uint2>();
        a.sum = sl.mp_sum.template as<double>();
        a.first = sl.mp_first.template as<double>();
        a.last = sl.mp_last.template as<double>();
        a.cross = sl.mp_cross.template as<uint32_t>();
        a.frag = sl.mp_frag.template as<blx_fragment_mass>();
        a.bad = sl.mp_bad.template as<uint32_t>();
        for (int q = 0; q < 3; ++q) a.origin[q] = sl.L.origin[q];
        a.cell = cell;
        a.n = N;
        a.k = c;
        hipLaunchKernelGGL(k_mp_bounds, dim3(gridFor(N)), dim3(kWG), 0, stream_, key,
                           sl.mp_bounds.template as<uint2>(), N);
        hipLaunchKernelGGL((k_mp_tile<R, 1>), dim3(tiles), dim3(kMpTile), 0, stream_, a);
        hipLaunchKernelGGL((k_mp_combine<R, 1>), dim3(tiles), dim3(kMpTile), 0, stream_, a);
        hipLaunchKernelGGL((k_mp_final1<R>), dim3(gridFor(c)), dim3(kWG), 0, stream_, a);
        hipLaunchKernelGGL((k_mp_tile<R, 2>), dim3(tiles), dim3(kMpTile), 0, stream_, a);
        hipLaunchKernelGGL((k_mp_combine<R, 2>), dim3(tiles), dim3(kMpTile), 0, stream_, a);
        hipLaunchKernelGGL((k_mp_final2<R>), dim3(gridFor(c)), dim3(kWG), 0, stream_, a);
        HIP_OK(hipGetLastError());
        uint32_t bad = kNone;
        HIP_OK(hipMemcpyAsync(&bad, sl.mp_bad.p, 4, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        if (bad != kNone) throw Runtime(bad & 1U ? "fragment inertia is singular" : "fragment has no mass");
        if (out && cap)
            HIP_OK(hipMemcpy(out, sl.mp_frag.p, size_t(std::min(c, cap)) * sizeof(blx_fragment_mass),
                             hipMemcpyDeviceToHost));
        return c;
    }

    // Device bytes one slot holds (its arena).
    size_t slotBytes(uint32_t k) const { return slot(k).arena.bytes(); }

    // Per-launch HIP-event timing of every kernel kind (off by default).
    void setProfiling(bool on) { prof_ = on; }
    uint32_t kernelStats(uint32_t k, std::string &name, double &seconds, uint64_t &launches) const {
        if (k < KCount) {
            name = kernelName(k);
            seconds = kernel_seconds_[k];
            launches = kernel_launches_[k];
        }
        return KCount;
    }

private:
    int device_;
    std::vector<std::unique_ptr<Slot<R>>> slots_;
    Shared shared_;
    Step<R> S_{};
    // stream_ carries the substep; side_ the inverse rest covariances that
    // overlap the folds; early_ the end sample of the nodes and bonds striker
    // pass 2 cannot reach, which overlaps pass 2.
    hipStream_t stream_ = nullptr, side_ = nullptr, early_ = nullptr;
    // lds_ runs the LDS sweep of the lattices that fit (k_sweep_lds) beside the
    // staged sweep of the rest, between pass 1 and the sweep's end.
    hipStream_t lds_ = nullptr;
    hipEvent_t ev_predict_ = nullptr, ev_start_ = nullptr, ev_sweep_ = nullptr;
    hipEvent_t ev_bond_end_ = nullptr, ev_rinv_ = nullptr, ev_pass1_ = nullptr, ev_velocity_ = nullptr,
               ev_early_ = nullptr, ev_strain_ = nullptr, ev_lds_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    uint32_t cus_ = 256;
    uint64_t launches_ = 0;
    DevBuf ls_, ctls_, spheres_, upload_map_;
    std::vector<DevCtl> ctl_all_;

    // Launch maps of the running slots, all in one device buffer.
    struct MapRef {
        uint32_t offset = 0, count = 0;
    };
    bool maps_dirty_ = true;
    DevBuf maps_;
    MapRef node_tiles_map_, bond_tiles_map_, list_map_, small_map_, mark_map_, lattice_map_;
    std::vector<MapRef> stage_maps_;
    std::vector<MapRef> damp_maps_;      // colour stages of every running slot, for radial damping
    size_t interior_stages_ = 0;
    std::vector<MapRef> sweep_maps_;     // the paired sweep's launch groups
    std::vector<uint32_t> sweep_phase_;  // reference phase id of each group (5 interior, 6 boundary)
    std::vector<uint32_t> sweep_lds_;    // LDS bytes of each group's launch
    std::vector<uint32_t> sweep_k_;      // 3 where a group holds a triple (k_sweep_triple), else 2
    std::vector<int64_t> sweep_stage_;   // the stage a group of single entries runs by stage map, or -1
    // A striker substep splits across streams (striker pass 2 beside the end
    // sample of what it cannot reach) when a running lattice has at least
    // kSplitNodes nodes (tuning streams 0; 1 always, 2 never): a smaller
    // lattice's pass 2 is shorter than the cross-stream waits the split adds,
    // so its substep stays on one stream.
    static constexpr uint32_t kSplitNodes = 1U << 17;
    bool split_nodes_ = true;
    // Launch tuning (blx_tuning). Every setting gives the same bits.
    // A batch's default: colour stages one per launch and no LDS sweep, the
    // faster setting for many small lattices advancing together (a lattice
    // sweeping in LDS holds its launch for the whole sweep); a single-lattice
    // backend starts from kBackendTuning instead.
    blx_tuning tuning_{1U, 0U, 1024U, 0U, 0U};
    MapRef lds_map_;                     // one entry per lattice that sweeps in LDS (k_sweep_lds)
    uint32_t lds_bytes_ = 0;             // that launch's dynamic LDS: the largest such lattice's nodes
    size_t lds_limit_ = 0;               // LDS one workgroup may hold on this device
    bool any_loads_ = false;

    uint32_t gridFor(uint32_t n) const {
        const uint32_t need = (n + kWG - 1) / kWG;
        return std::max(1U, std::min(need, cus_ * 16U));
    }
    // Two workgroups per CU: for launches that end in one atomic per workgroup.
    uint32_t gridFew(uint32_t n) const {
        const uint32_t need = (n + kWG - 1) / kWG;
        return std::max(1U, std::min(need, cus_ * 2U));
    }

    uint32_t countRunning() const {
        uint32_t n = 0;
        for (auto &s : slots_) n += s->armed ? 1U : 0U;
        return n;
    }

    void updateShared() {
        bool any = false;
        for (auto &s : slots_) any |= s->uploaded;
        if (!any) shared_.set = false;
    }

    void buildMaps() {
        std::vector<WgMap> all;
        auto begin = [&](MapRef &r) { r.offset = uint32_t(all.size()); };
        auto end = [&](MapRef &r) { r.count = uint32_t(all.size()) - r.offset; };
        std::vector<uint32_t> run;
        uint64_t total_nodes = 0;
        any_loads_ = false;
        for (uint32_t k = 0; k < slots_.size(); ++k)
            if (slots_[k]->armed) {
                run.push_back(k);
                total_nodes += slots_[k]->N;
                any_loads_ |= slots_[k]->loads_set;
            }
        // List launches share a workgroup budget among the slots by size.
        auto slices = [&](uint32_t n, uint32_t cap) {
            const uint64_t need = std::max(1U, (n + kWG - 1) / kWG);
            const uint64_t fair = total_nodes ? (uint64_t(cap) * n + total_nodes - 1) / total_nodes : cap;
            return uint32_t(std::max<uint64_t>(1, std::min(need, fair)));
        };
        begin(node_tiles_map_);
        for (uint32_t k : run)
            for (uint32_t t = 0; t < slots_[k]->node_tiles; ++t) all.push_back(WgMap{k, t, 0, 0});
        end(node_tiles_map_);
        // End-sample spans: up to kEndSpan consecutive bond tiles per workgroup.
        begin(bond_tiles_map_);
        for (uint32_t k : run)
            for (uint32_t t = 0; t < slots_[k]->bond_tiles; t += kEndSpan)
                all.push_back(WgMap{k, t, std::min(kEndSpan, slots_[k]->bond_tiles - t), 0});
        end(bond_tiles_map_);
        begin(list_map_);
        for (uint32_t k : run) {
            const uint32_t s = slices(slots_[k]->N, cus_ * 16U);
            for (uint32_t q = 0; q < s; ++q) all.push_back(WgMap{k, q, s, 0});
        }
        end(list_map_);
        begin(small_map_);
        for (uint32_t k : run) {
            const uint32_t s = slices(slots_[k]->N, cus_ * 4U);
            for (uint32_t q = 0; q < s; ++q) all.push_back(WgMap{k, q, s, 0});
        }
        end(small_map_);
        begin(mark_map_);
        for (uint32_t k : run) {
            const Slot<R> &sl = *slots_[k];
            const uint64_t work = uint64_t(sl.sch.blocks) * kCap * (sl.D + 1);
            const uint64_t cap = std::max<uint64_t>(1, cus_ * 4U / std::max<size_t>(1, run.size()));
            const uint32_t s = uint32_t(std::max<uint64_t>(1, std::min((work + kWG - 1) / kWG, cap)));
            for (uint32_t q = 0; q < s; ++q) all.push_back(WgMap{k, q, s, 0});
        }
        end(mark_map_);
        begin(lattice_map_);
        for (uint32_t k : run) all.push_back(WgMap{k, 0, 0, 0});
        end(lattice_map_);
        // A one-iteration run sweeps every lattice whose nodes fit in one
        // workgroup's LDS in one launch (k_sweep_lds, one workgroup each); the
        // rest sweep by stages, or by stage pairs.
        begin(lds_map_);
        lds_bytes_ = 0;
        std::vector<uint32_t> staged;
        for (uint32_t k : run) {
            const size_t bytes = kLdsPal * sizeof(EdgePal<R>) + size_t(slots_[k]->N) * sizeof(R4<R>);
            if (ldsSweeps(*slots_[k])) {
                all.push_back(WgMap{k, 0, 0, 0});
                lds_bytes_ = std::max(lds_bytes_, uint32_t(bytes));
            } else {
                staged.push_back(k);
            }
        }
        end(lds_map_);
        // Stage s of the batch is stage s of every running slot: lattices are
        // independent, so only each slot's own stage order matters.
        size_t stages = 0;
        interior_stages_ = 0;
        for (uint32_t k : staged) {
            stages = std::max(stages, slots_[k]->stages.size());
            interior_stages_ = std::max(interior_stages_, slots_[k]->interior_stages);
        }
        stage_maps_.assign(stages, MapRef{});
        for (size_t s = 0; s < stages; ++s) {
            begin(stage_maps_[s]);
            for (uint32_t k : staged) {
                const Slot<R> &sl = *slots_[k];
                if (s < sl.stages.size())
                    for (const uint2 &r : sl.stages[s]) all.push_back(WgMap{k, r.x, r.y, 0});
            }
            end(stage_maps_[s]);
        }
        // Radial damping runs by colour stages on every running slot, the
        // ones that sweep in LDS included.
        size_t damp_stages = 0;
        if (S_.damping > R(0))
            for (uint32_t k : run) damp_stages = std::max(damp_stages, slots_[k]->stages.size());
        damp_maps_.assign(damp_stages, MapRef{});
        for (size_t s = 0; s < damp_stages; ++s) {
            begin(damp_maps_[s]);
            for (uint32_t k : run) {
                const Slot<R> &sl = *slots_[k];
                if (s < sl.stages.size())
                    for (const uint2 &r : sl.stages[s]) all.push_back(WgMap{k, r.x, r.y, 0});
            }
            end(damp_maps_[s]);
        }
        // The paired sweep's launch group g of the batch is group g of every
        // running slot (each slot's groups keep its own stage order and its
        // own even number of ping-pong pairs).
        size_t groups = 0;
        for (uint32_t k : staged) groups = std::max(groups, slots_[k]->sweep_groups.size());
        sweep_maps_.assign(groups, MapRef{});
        sweep_phase_.assign(groups, 5U);
        sweep_lds_.assign(groups, 0U);
        sweep_k_.assign(groups, 2U);
        for (size_t g = 0; g < groups; ++g) {
            begin(sweep_maps_[g]);
            for (uint32_t k : staged) {
                const Slot<R> &sl = *slots_[k];
                if (g >= sl.sweep_groups.size()) continue;
                if (sl.sweep_first_stage[g] >= sl.interior_stages) sweep_phase_[g] = 6U;
                sweep_lds_[g] = std::max(sweep_lds_[g], uint32_t(sl.sweep_lds_nodes[g] * sizeof(R4<R>)));
                if (((sl.sweep_groups[g][0].pad >> 16) & 0xfU) == kSweepTriple) sweep_k_[g] = 3U;
                // Two entries per workgroup: the second holds counts, not a lattice.
                const std::vector<WgMap> &es = sl.sweep_groups[g];
                for (size_t q = 0; q < es.size(); q += 2) {
                    WgMap e = es[q];
                    e.lat = k;
                    all.push_back(e);
                    all.push_back(es[q + 1]);
                }
            }
            end(sweep_maps_[g]);
        }
        // A group that is the same single stage s in every staged slot that
        // has a stage s runs as k_solve over stage s's map, the lighter kernel.
        sweep_stage_.assign(groups, -1);
        for (size_t g = 0; g < groups; ++g) {
            int64_t st = -1;
            bool ok = true;
            for (uint32_t k : staged) {
                const Slot<R> &sl = *slots_[k];
                if (g >= sl.sweep_groups.size()) continue;
                const bool single = ((sl.sweep_groups[g][0].pad >> 16) & 0xfU) == kSweepSingle;
                if (!single || (st >= 0 && st != int64_t(sl.sweep_first_stage[g]))) ok = false;
                st = sl.sweep_first_stage[g];
            }
            for (uint32_t k : staged) {
                const Slot<R> &sl = *slots_[k];
                if (!ok || st < 0 || size_t(st) >= sl.stages.size()) continue;
                if (g >= sl.sweep_groups.size() || sl.sweep_first_stage[g] != uint32_t(st) ||
                    ((sl.sweep_groups[g][0].pad >> 16) & 0xfU) != kSweepSingle)
                    ok = false;
            }
            if (ok && st >= 0 && size_t(st) < stage_maps_.size()) sweep_stage_[g] = st;
        }

        split_nodes_ = tuning_.streams == 1U;
        if (tuning_.streams == 0U)
            for (uint32_t k : run) split_nodes_ |= slots_[k]->N >= kSplitNodes;
        maps_.alloc(std::max<size_t>(1, all.size()) * sizeof(WgMap));
        if (!all.empty()) HIP_OK(hipMemcpy(maps_.p, all.data(), all.size() * sizeof(WgMap), hipMemcpyHostToDevice));
        maps_dirty_ = false;
    }

    // Whether slot sl sweeps in LDS (k_sweep_lds): on request, in a
    // one-iteration run, with its nodes within one workgroup's LDS.
    bool ldsSweeps(const Slot<R> &sl) const {
        return tuning_.lds_sweep != 0 && shared_.iterations == 1 && sl.N > 0 &&
               kLdsPal * sizeof(EdgePal<R>) + size_t(sl.N) * sizeof(R4<R>) <= lds_limit_;
    }

    Batch<R> batch(const MapRef &r) const {
        return Batch<R>{ls_.as<DevLattice<R>>(), maps_.as<WgMap>() + r.offset, r.count};
    }

    DevCtl readCtl(uint32_t k) {
        DevCtl c;
        HIP_OK(hipMemcpy(&c, ctls_.as<DevCtl>() + k, sizeof(DevCtl), hipMemcpyDeviceToHost));
        return c;
    }
    void writeCtl(uint32_t k, const DevCtl &c) {
        HIP_OK(hipMemcpy(ctls_.as<DevCtl>() + k, &c, sizeof(DevCtl), hipMemcpyHostToDevice));
    }
    void readAllCtls() {
        ctl_all_.resize(slots_.size());
        HIP_OK(hipMemcpy(ctl_all_.data(), ctls_.p, slots_.size() * sizeof(DevCtl), hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < slots_.size(); ++k)
            if (slots_[k]->uploaded) slots_[k]->ctl_host = ctl_all_[k];
    }
    void setRemaining(uint32_t k, uint64_t n) {
        DevCtl c = readCtl(k);
        c.load_remaining = n;
        writeCtl(k, c);
        slot(k).ctl_host.load_remaining = n;
    }
    Sphere<R> readSphere(uint32_t k) {
        Sphere<R> s;
        HIP_OK(hipMemcpy(&s, spheres_.as<Sphere<R>>() + k, sizeof(s), hipMemcpyDeviceToHost));
        return s;
    }
    template <class T> void copyDown(std::vector<T> &dst, const T *src) {
        if (!dst.empty()) HIP_OK(hipMemcpy(dst.data(), src, dst.size() * sizeof(T), hipMemcpyDeviceToHost));
    }

    void captureFrame(uint32_t k, uint64_t step) {
        Slot<R> &sl = slot(k);
        HostFrame<R> f;
        f.step = step;
        std::vector<R4<R>> u4(sl.N);
        std::vector<R2<R>> pv(sl.B);
        copyDown(u4, sl.L.u4);
        copyDown(pv, sl.L.pv);
        f.u.resize(3 * size_t(sl.N));
        for (uint32_t i = 0; i < sl.N; ++i) {
            f.u[3 * size_t(i)] = float(u4[i].x);
            f.u[3 * size_t(i) + 1] = float(u4[i].y);
            f.u[3 * size_t(i) + 2] = float(u4[i].z);
        }
        f.alive.resize(sl.B);
        HIP_OK(hipMemcpy(f.alive.data(), sl.L.alive, sl.B, hipMemcpyDeviceToHost));
        f.damage.resize(sl.B);
        for (uint32_t o = 0; o < sl.B; ++o) f.damage[sl.node_j_host[o]] = float(pv[o].y);
        f.sphere = readSphere(k);
        sl.frames.push_back(std::move(f));
    }

    // Kernel kinds for the optional per-launch event timing, each with the
    // reference phase id it belongs to (FastLattice.cpp:22-26).
    enum Kernel : uint32_t { KStrainStart, KBondStart, KPredict, KPass1, KSolve, KProject, KVelocity, KDamp,
                             KSupportVelocity, KPass2, KStrain, KBondEnd, KFinish, KRinv, KMark, KStrainLate,
                             KBondEndLate, KContact, KSweepPair, KSweepTriple, KCount };
    static const char *kernelName(uint32_t k) {
        static const char *names[KCount] = {"k_strain_start", "k_bond_start", "k_predict", "k_pass1", "k_solve",
                                            "k_project", "k_velocity", "k_damp", "k_support_velocity", "k_pass2",
                                            "k_strain", "k_bond_end", "k_finish", "k_rinv", "k_mark_affected",
                                            "k_strain_late", "k_bond_end_late", "k_node_contact", "k_sweep_pair",
                                            "k_sweep_triple"};
        return k < KCount ? names[k] : "unknown";
    }
    struct ProfRecord { uint32_t kernel, phase; hipEvent_t a, b; };
    bool prof_ = false;
    bool sample_ = false;  // the substep being enqueued is the batch's phase-timed one
    double sample_share_[BLX_PHASE_COUNT] = {};  // phase shares of the last timed substep
    bool have_share_ = false;
    std::vector<hipEvent_t> pool_;
    std::vector<ProfRecord> records_;
    double kernel_seconds_[KCount] = {};
    uint64_t kernel_launches_[KCount] = {};

    void profBegin(uint32_t kernel, uint32_t phase, hipStream_t on = nullptr) {
        if (!prof_ && !sample_) return;
        if (!on) on = stream_;
        const size_t need = 2 * (records_.size() + 1);
        while (pool_.size() < need) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            pool_.push_back(e);
        }
        ProfRecord r{kernel, phase, pool_[need - 2], pool_[need - 1]};
        HIP_OK(hipEventRecord(r.a, on));
        records_.push_back(r);
    }
    // A record whose two events a launch carries itself (hipExtLaunchKernelGGL
    // start and stop events), so that it times the kernel alone, as a kernel
    // trace does; nullptr when nothing is timed.
    ProfRecord *profLaunch(uint32_t kernel, uint32_t phase) {
        if ((!prof_ && !sample_) || capturing_) return nullptr;
        const size_t need = 2 * (records_.size() + 1);
        while (pool_.size() < need) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            pool_.push_back(e);
        }
        records_.push_back(ProfRecord{kernel, phase, pool_[need - 2], pool_[need - 1]});
        return &records_.back();
    }
    void profEnd(hipStream_t on = nullptr) {
        if (!prof_ && !sample_) return;
        HIP_OK(hipEventRecord(records_.back().b, on ? on : stream_));
    }
    void profAddLaunches(uint32_t kernel, uint32_t extra) {
        if (prof_) kernel_launches_[kernel] += extra;
    }
    // RunStatus::phase_seconds (FastLattice.hpp:225-228). With profiling on,
    // every launch is timed and each phase gets its kernels' event time. With
    // it off, each batch times one substep phase by phase and splits the
    // batch's device time (batch_seconds) in those proportions; kernels of
    // overlapping streams share the wall time as they share the sampled sum.
    void profCollect(double batch_seconds) {
        if (prof_) {
            for (const ProfRecord &r : records_) {
                float ms = 0.f;
                HIP_OK(hipEventElapsedTime(&ms, r.a, r.b));
                kernel_seconds_[r.kernel] += double(ms) * 1e-3;
                kernel_launches_[r.kernel] += 1;
                for (auto &s : slots_)
                    if (s->armed) s->status.phase_seconds[r.phase] += double(ms) * 1e-3;
            }
            records_.clear();
            return;
        }
        if (!records_.empty()) {
            double ph[BLX_PHASE_COUNT] = {};
            double total = 0.0;
            for (const ProfRecord &r : records_) {
                float ms = 0.f;
                HIP_OK(hipEventElapsedTime(&ms, r.a, r.b));
                if (r.phase < BLX_PHASE_COUNT) ph[r.phase] += double(ms);
                total += double(ms);
            }
            records_.clear();
            if (total > 0.0) {
                for (uint32_t p = 0; p < BLX_PHASE_COUNT; ++p) sample_share_[p] = ph[p] / total;
                have_share_ = true;
            }
        }
        if (!have_share_) return;
        for (auto &s : slots_)
            if (s->armed)
                for (uint32_t p = 0; p < BLX_PHASE_COUNT; ++p) s->status.phase_seconds[p] += sample_share_[p] * batch_seconds;
    }
    uint64_t profilingBatch(uint64_t batch) const { return prof_ ? std::min<uint64_t>(batch, 4) : batch; }

    // Substep graphs (blx_tuning::graph_substeps). Every launch argument of a
    // substep is constant (the run state lives on the device), so a block of
    // substeps is captured once per advance call (and after every map rebuild)
    // and replayed; the side streams fork from stream_ and join back into it
    // through the substep's own events, which become graph edges. Nothing in a
    // substep records a memset. Replays pay where the host, not the device,
    // sets the pace: small lattices with short kernels.
    static constexpr uint32_t kMaxGraphSubsteps = 256;
    bool capturing_ = false;
    uint64_t graph_len_ = 0;  // substeps in graph_exec_
    hipGraphExec_t graph_exec_ = nullptr;

    void dropGraph() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
    }
    void captureSubsteps(uint64_t len) {
        dropGraph();
        HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
        capturing_ = true;
        hipError_t err = hipSuccess;
        try {
            for (uint64_t q = 0; q < len; ++q) {
                enqueueSubstepA();
                enqueueSubstepB();
            }
            err = hipStreamWaitEvent(stream_, ev_rinv_, 0);
        } catch (...) {
            capturing_ = false;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        capturing_ = false;
        hipGraph_t g = nullptr;
        const hipError_t end = hipStreamEndCapture(stream_, &g);
        // Every stream the capture forked into must have been joined back into
        // stream_ by an event (hipStreamEndCapture reports an unjoined fork,
        // and a stream left capturing would capture the next eager launches).
        bool left = false;
        for (hipStream_t s : {side_, early_, lds_}) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) left = true;
        }
        hipError_t inst = hipSuccess;
        if (err == hipSuccess && end == hipSuccess && !left)
            inst = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (inst != hipSuccess || left) graph_exec_ = nullptr;
        graph_len_ = graph_exec_ ? len : 0;
        HIP_OK(err);
        HIP_OK(end);
        if (left) throw Runtime("a side stream was left unjoined in the substep capture");
        HIP_OK(inst);
    }

// A launch whose end another stream waits on (ev; nullptr for none). The
// event rides on the launch as its stop event: a separate record puts a marker
// between two kernels of the stream, which costs about 5 us. While a substep
// graph is being captured the record is a graph edge instead.
#define BLX_SIGNAL_LDS(kernel, grid, block, lds, stream, ev, ...)                                  \
    do {                                                                                           \
        hipEvent_t ev_ = (ev);                                                                     \
        if (capturing_) {                                                                          \
            hipLaunchKernelGGL(kernel, grid, block, lds, stream, __VA_ARGS__);                     \
            HIP_OK(hipGetLastError());                                                             \
            if (ev_) HIP_OK(hipEventRecord(ev_, stream));                                          \
        } else {                                                                                   \
            hipExtLaunchKernelGGL(kernel, grid, block, lds, stream, nullptr, ev_, 0, __VA_ARGS__); \
        }                                                                                          \
    } while (0)
#define BLX_SIGNAL(kernel, grid, block, stream, ev, ...) BLX_SIGNAL_LDS(kernel, grid, block, 0, stream, ev, __VA_ARGS__)

#define BLX_LAUNCH(kernel, phase, ...) \
    do {                                \
        profBegin(kernel, phase);       \
        __VA_ARGS__;                    \
        profEnd();                      \
    } while (0)

    static dim3 tiles(const MapRef &r) { return dim3(std::max(1U, xcdGrid(r.count))); }
    static dim3 entries(const MapRef &r) { return dim3(std::max(1U, r.count)); }

    // A substep in two parts. Part A runs through striker pass 2 (and, with
    // node contact on, the staleness test of the pair list); part B runs the
    // contact narrow phase, the end sample and the folds. With contact on the
    // host reads the rebuild flags between them and rebuilds the pair lists
    // that need it (CpuLatticeBackend.cpp:424-437); otherwise the parts are
    // enqueued back to back.
    template <bool DIRECT> void enqueueA() {
        const Step<R> S = S_;
        const bool sphere = S.sphere_enabled != 0;
        const bool plastic = shared_.plastic;
        const bool contact = S.nc.mode != 0U;
        // With a striker, the end sample of what pass 2 cannot reach overlaps
        // pass 2, unless node contact moves nodes between pass 2 and the sample,
        // or no running lattice is large enough for the overlap to repay the
        // cross-stream waits it adds (split_nodes_).
        const bool split = sphere && !plastic && !contact && split_nodes_;
        // Without damping between the velocity update and the end sample,
        // striker pass 2 need not wait for the velocity update of the rest.
        const bool own = split && !(S.damping > R(0));
        const Batch<R> nodes = batch(node_tiles_map_), bonds = batch(bond_tiles_map_), list = batch(list_map_),
                       small = batch(small_map_), lats = batch(lattice_map_);
        const dim3 ng = tiles(node_tiles_map_), bg = tiles(bond_tiles_map_), lg = entries(list_map_),
                   sg = entries(small_map_), fg = entries(lattice_map_);
        // An event another stream waits on rides on the kernel it follows
        // (hipExtLaunchKernelGGL's stop event): a separate hipEventRecord puts a
        // marker between two kernels of the stream, which cost about 6 us each.
        profBegin(KPredict, 3);
        if (sphere && any_loads_) BLX_SIGNAL((k_predict<R, true, true, DIRECT>), ng, dim3(kWG), stream_, ev_predict_, nodes, S);
        else if (sphere) BLX_SIGNAL((k_predict<R, true, false, DIRECT>), ng, dim3(kWG), stream_, ev_predict_, nodes, S);
        else if (any_loads_) BLX_SIGNAL((k_predict<R, false, true, DIRECT>), ng, dim3(kWG), stream_, ev_predict_, nodes, S);
        else BLX_SIGNAL((k_predict<R, false, false, DIRECT>), ng, dim3(kWG), stream_, ev_predict_, nodes, S);
        profEnd();
        // The start sample of the nodes and bonds the last failure round made
        // dirty reads the positions at substep start from up4, so it runs on
        // side_ after k_predict, behind the previous substep's k_rinv on the
        // same stream, and overlaps the striker pass and the sweep; nothing
        // before the end sample reads what it writes.
        HIP_OK(hipStreamWaitEvent(side_, ev_predict_, 0));
        profBegin(KStrainStart, 1, side_);
        hipLaunchKernelGGL((k_strain<R, DIRECT, kStrainDirty>), lg, dim3(kWG), 0, side_, list);
        profEnd(side_);
        profBegin(KBondStart, 2, side_);
        BLX_SIGNAL((k_bond_start<R, DIRECT>), lg, dim3(kWG), side_, ev_start_, list);
        profEnd(side_);
        const uint32_t iterations = shared_.iterations;
        const bool pairs = iterations == 1 && tuning_.colours_per_launch == 2;
        // The lattices that sweep in LDS do so on lds_, beside the staged sweep
        // of the rest, once pass 1 (or, without a striker, k_predict) is done.
        const bool fork = lds_map_.count > 0 && (pairs ? !sweep_maps_.empty() : !stage_maps_.empty());
        if (sphere)
            BLX_LAUNCH(KPass1, 4, BLX_SIGNAL((k_pass<R, 1>), fg, dim3(kFinishThreads), stream_,
                                             (split && !own) || fork ? ev_pass1_ : nullptr, lats, S));
        // Without damping the velocity update on the early stream marks the
        // nodes pass 2 can reach; otherwise k_mark_affected does it there
        // during the sweep, after pass 1.
        if (split && !own) {
            HIP_OK(hipStreamWaitEvent(early_, ev_pass1_, 0));
            profBegin(KMark, 4, early_);
            hipLaunchKernelGGL((k_mark_affected<R>), entries(mark_map_), dim3(kWG), 0, early_, batch(mark_map_));
            profEnd(early_);
        }
        if (lds_map_.count > 0) {
            // Every lattice that fits sweeps in LDS, one workgroup each: on lds_
            // when other lattices sweep by stages (joined after their sweep),
            // else on stream_, where it is the whole sweep.
            const bool last = own && !fork;
            const Batch<R> lb = batch(lds_map_);
            const dim3 g(lds_map_.count), t(tuning_.lds_threads);
            hipStream_t on = fork ? lds_ : stream_;
            if (fork) HIP_OK(hipStreamWaitEvent(lds_, sphere ? ev_pass1_ : ev_predict_, 0));
            profBegin(KSolve, 5, on);
            if (fork && plastic)
                BLX_SIGNAL_LDS((k_sweep_lds<R, DIRECT, true>), g, t, lds_bytes_, on, ev_lds_, lb, S);
            else if (fork)
                BLX_SIGNAL_LDS((k_sweep_lds<R, DIRECT, false>), g, t, lds_bytes_, on, ev_lds_, lb, S);
            else if (last)
                BLX_SIGNAL_LDS((k_sweep_lds<R, DIRECT, false>), g, t, lds_bytes_, on, ev_sweep_, lb, S);
            else if (plastic)
                hipLaunchKernelGGL((k_sweep_lds<R, DIRECT, true>), g, t, lds_bytes_, on, lb, S);
            else
                hipLaunchKernelGGL((k_sweep_lds<R, DIRECT, false>), g, t, lds_bytes_, on, lb, S);
            profEnd(on);
        }
        if (pairs) {
            // The paired sweep: two stages per launch on LDS tiles (k_sweep_pair).
            // When timed, each launch carries its own start and stop events.
            for (size_t g = 0; g < sweep_maps_.size(); ++g) {
                const MapRef &m = sweep_maps_[g];
                const uint32_t lds = sweep_lds_[g];
                const bool last = own && g + 1 == sweep_maps_.size();
                const bool single = sweep_stage_[g] >= 0;
                const bool triple = sweep_k_[g] == 3U;
                ProfRecord *pr = profLaunch(single ? KSolve : (triple ? KSweepTriple : KSweepPair), sweep_phase_[g]);
                hipEvent_t start = pr ? pr->a : nullptr, stop = last ? ev_sweep_ : (pr ? pr->b : nullptr);
                if (capturing_) {
                    // A captured launch takes no events; the signal is a graph edge.
                    start = nullptr;
                    stop = nullptr;
                }
                const dim3 grid(std::max(1U, xcdGrid(m.count / 2)));
                if (single) {
                    const MapRef &sm = stage_maps_[size_t(sweep_stage_[g])];
                    if (plastic)
                        hipExtLaunchKernelGGL((k_solve<R, DIRECT, true>), tiles(sm), dim3(kWG), 0, stream_, start, stop, 0,
                                              batch(sm), S, kSolveFirst);
                    else
                        hipExtLaunchKernelGGL((k_solve<R, DIRECT, false>), tiles(sm), dim3(kWG), 0, stream_, start, stop,
                                              0, batch(sm), S, kSolveFirst);
                } else if (triple && plastic) {
                    hipExtLaunchKernelGGL((k_sweep_triple<R, DIRECT, true>), grid, dim3(kPairThreads), lds, stream_,
                                          start, stop, 0, batch(m), S);
                } else if (triple) {
                    hipExtLaunchKernelGGL((k_sweep_triple<R, DIRECT, false>), grid, dim3(kPairThreads), lds, stream_,
                                          start, stop, 0, batch(m), S);
                } else if (plastic) {
                    hipExtLaunchKernelGGL((k_sweep_pair<R, DIRECT, true>), grid, dim3(kPairThreads), lds, stream_, start,
                                          stop, 0, batch(m), S);
                } else {
                    hipExtLaunchKernelGGL((k_sweep_pair<R, DIRECT, false>), grid, dim3(kPairThreads), lds, stream_, start,
                                          stop, 0, batch(m), S);
                }
                HIP_OK(hipGetLastError());
                if (capturing_ && last) HIP_OK(hipEventRecord(ev_sweep_, stream_));
                if (pr && last) HIP_OK(hipEventRecord(pr->b, stream_));
            }
        }
        for (uint32_t it = 0; it < (pairs ? 0U : iterations); ++it) {
            // With more than one iteration every solve stores its multiplier, as
            // the reference does (LatticePhysics.hpp:1163): a bond that skips
            // iteration 0 (a zero-length bond) reads the previous substep's
            // last value in iteration 1.
            const uint32_t flags = (it == 0 ? kSolveFirst : 0U) | (iterations > 1 ? kSolveKeepLambda : 0U);
            // One timing bracket over the back-to-back stage launches of the
            // interior phase and one over the boundary phases.
            for (int phase = 0; phase < 2; ++phase) {
                const size_t lo = phase == 0 ? 0 : interior_stages_;
                const size_t hi = phase == 0 ? std::min(interior_stages_, stage_maps_.size()) : stage_maps_.size();
                if (lo >= hi) continue;
                profBegin(KSolve, phase == 0 ? 5 : 6);
                for (size_t k = lo; k < hi; ++k) {
                    const MapRef &m = stage_maps_[k];
                    if (own && it + 1 == iterations && k + 1 == stage_maps_.size())
                        BLX_SIGNAL((k_solve<R, DIRECT, false>), tiles(m), dim3(kWG), stream_, ev_sweep_, batch(m), S, flags);
                    else if (plastic)
                        hipLaunchKernelGGL((k_solve<R, DIRECT, true>), tiles(m), dim3(kWG), 0, stream_, batch(m), S, flags);
                    else
                        hipLaunchKernelGGL((k_solve<R, DIRECT, false>), tiles(m), dim3(kWG), 0, stream_, batch(m), S, flags);
                }
                profEnd();
                profAddLaunches(KSolve, uint32_t(hi - lo) - 1U);
            }
            if ((S.plane_count > 0 || S.blade.enabled) && it + 1 < iterations)
                BLX_LAUNCH(KProject, 7, hipLaunchKernelGGL((k_project<R>), ng, dim3(kWG), 0, stream_, nodes, S));
        }
        // The LDS sweep on lds_ joins here: what follows on stream_ reads every
        // lattice's swept nodes, and so does striker pass 2 on the early stream.
        if (fork) HIP_OK(hipStreamWaitEvent(stream_, ev_lds_, 0));
        if (own) {
            if (fork) HIP_OK(hipStreamWaitEvent(early_, ev_lds_, 0));
            // Striker pass 2 forks onto the early stream after the sweep and
            // takes its listed candidates' velocity update itself; the strain
            // of the nodes it can reach follows it there. The velocity update
            // of the rest, the end sample of what pass 2 cannot reach and the
            // affected bonds' sample stay on stream_, the longer branch, which
            // joins the early stream only before that last sample: a join the
            // early stream has usually reached by then, so stream_ does not
            // wait on a cross-stream signal at either end of the branch. The
            // host enqueues k_velocity (which signals ev_velocity_) before the
            // early stream's wait on it.
            HIP_OK(hipStreamWaitEvent(early_, ev_sweep_, 0));
            profBegin(KPass2, 10, early_);
            hipLaunchKernelGGL((k_pass<R, 2, true>), fg, dim3(kPass2Threads), 0, early_, lats, S);
            profEnd(early_);
            HIP_OK(hipStreamWaitEvent(stream_, ev_start_, 0));
            profBegin(KVelocity, 8);
            const Batch<R> mark = batch(mark_map_);
            if (S.plane_count > 0 || S.blade.enabled)
                BLX_SIGNAL((k_velocity<R, true, true>), ng, dim3(kWG), stream_, ev_velocity_, nodes, S, mark);
            else BLX_SIGNAL((k_velocity<R, false, true>), ng, dim3(kWG), stream_, ev_velocity_, nodes, S, mark);
            profEnd();
            HIP_OK(hipStreamWaitEvent(early_, ev_velocity_, 0));
            profBegin(KStrainLate, 11, early_);
            hipLaunchKernelGGL((k_strain<R, DIRECT, kStrainAffected>), sg, dim3(kWG), 0, early_, small);
            profEnd(early_);
            BLX_LAUNCH(KStrain, 11, BLX_SIGNAL((k_strain<R, DIRECT, kStrainUnaffected>), ng, dim3(kWG), stream_, ev_strain_, nodes));
            // The affected bonds' sample reads the strains of both ends, so it
            // follows both strain passes; it shares no bond with the unaffected
            // sample (whose shared words it updates only with atomics), so the
            // two run side by side and stream_ joins the early stream once,
            // before the finish.
            HIP_OK(hipStreamWaitEvent(early_, ev_strain_, 0));
            profBegin(KBondEndLate, 12, early_);
            BLX_SIGNAL((k_bond_end<R, DIRECT, false, kBondAffected>), sg, dim3(kWG), early_, ev_early_, small, S);
            profEnd(early_);
            BLX_LAUNCH(KBondEnd, 12,
                       BLX_SIGNAL((k_bond_end<R, DIRECT, false, kBondUnaffected>), bg, dim3(kWG), stream_, ev_bond_end_, bonds, S));
            HIP_OK(hipStreamWaitEvent(stream_, ev_early_, 0));
            HIP_OK(hipGetLastError());
            return;
        }
        HIP_OK(hipStreamWaitEvent(stream_, ev_start_, 0));
        profBegin(KVelocity, 8);
        // With no damping after it, the velocity update carries the event the early stream waits on.
        hipEvent_t after_velocity = split && !(S.damping > R(0)) ? ev_velocity_ : nullptr;
        if (S.plane_count > 0 || S.blade.enabled)
            BLX_SIGNAL((k_velocity<R, true>), ng, dim3(kWG), stream_, after_velocity, nodes, S, Batch<R>{});
        else BLX_SIGNAL((k_velocity<R, false>), ng, dim3(kWG), stream_, after_velocity, nodes, S, Batch<R>{});
        profEnd();
        if (S.damping > R(0)) {
            for (const MapRef &m : damp_maps_)
                BLX_LAUNCH(KDamp, 9, hipLaunchKernelGGL((k_damp<R, DIRECT>), tiles(m), dim3(kWG), 0, stream_, batch(m), S));
            BLX_LAUNCH(KSupportVelocity, 9, hipLaunchKernelGGL((k_support_velocity<R>), ng, dim3(kWG), 0, stream_, nodes, S));
        }
        if (split) {
            if (!after_velocity) HIP_OK(hipEventRecord(ev_velocity_, stream_));
            HIP_OK(hipStreamWaitEvent(early_, ev_velocity_, 0));
            profBegin(KStrain, 11, early_);
            hipLaunchKernelGGL((k_strain<R, DIRECT, kStrainUnaffected>), ng, dim3(kWG), 0, early_, nodes);
            profEnd(early_);
            profBegin(KBondEnd, 12, early_);
            BLX_SIGNAL((k_bond_end<R, DIRECT, false, kBondUnaffected>), bg, dim3(kWG), early_, ev_early_, bonds, S);
            profEnd(early_);
            BLX_LAUNCH(KPass2, 10, hipLaunchKernelGGL((k_pass<R, 2>), fg, dim3(kPass2Threads), 0, stream_, lats, S));
            BLX_LAUNCH(KStrainLate, 11, hipLaunchKernelGGL((k_strain<R, DIRECT, kStrainAffected>), sg, dim3(kWG), 0, stream_, small));
            HIP_OK(hipStreamWaitEvent(stream_, ev_early_, 0));
            BLX_LAUNCH(KBondEndLate, 12,
                       BLX_SIGNAL((k_bond_end<R, DIRECT, false, kBondAffected>), sg, dim3(kWG), stream_, ev_bond_end_, small, S));
        } else {
            if (sphere) BLX_LAUNCH(KPass2, 10, hipLaunchKernelGGL((k_pass<R, 2>), fg, dim3(kPass2Threads), 0, stream_, lats, S));
            if (contact) BLX_LAUNCH(KContact, 14, hipLaunchKernelGGL((k_nc_stale<R>), ng, dim3(kWG), 0, stream_, nodes, S));
        }
        HIP_OK(hipGetLastError());
    }

    template <bool DIRECT> void enqueueB() {
        const Step<R> S = S_;
        const bool sphere = S.sphere_enabled != 0;
        const bool plastic = shared_.plastic;
        const bool contact = S.nc.mode != 0U;
        const bool split = sphere && !plastic && !contact && split_nodes_;
        const Batch<R> nodes = batch(node_tiles_map_), bonds = batch(bond_tiles_map_), list = batch(list_map_),
                       lats = batch(lattice_map_);
        const dim3 ng = tiles(node_tiles_map_), bg = tiles(bond_tiles_map_), lg = entries(list_map_),
                   fg = entries(lattice_map_);
        if (!split) {
            if (contact)
                BLX_LAUNCH(KContact, 15, hipLaunchKernelGGL((k_nc_narrow<R>), fg, dim3(kNarrowThreads), 0, stream_, lats, S));
            BLX_LAUNCH(KStrain, 11, hipLaunchKernelGGL((k_strain<R, DIRECT, kStrainAll>), ng, dim3(kWG), 0, stream_, nodes));
            profBegin(KBondEnd, 12);
            if (plastic) hipLaunchKernelGGL((k_bond_end<R, DIRECT, true, kBondAll>), bg, dim3(kWG), 0, stream_, bonds, S);
            else hipLaunchKernelGGL((k_bond_end<R, DIRECT, false, kBondAll>), bg, dim3(kWG), 0, stream_, bonds, S);
            profEnd();
        }
        // The dirty nodes' inverse rest covariances overlap the ordered folds
        // and the next substep's k_predict; the next start sample follows them
        // on side_.
        if (!split) HIP_OK(hipEventRecord(ev_bond_end_, stream_));
        HIP_OK(hipStreamWaitEvent(side_, ev_bond_end_, 0));
        // Striker pass 2's branch ends with the affected bonds' sample on the
        // early stream, beside the unaffected one.
        if (split && !(S.damping > R(0))) HIP_OK(hipStreamWaitEvent(side_, ev_early_, 0));
        profBegin(KRinv, 1, side_);
        BLX_SIGNAL((k_rinv<R>), lg, dim3(kWG), side_, ev_rinv_, list);
        profEnd(side_);
        profBegin(KFinish, 13);
        if (any_loads_ && plastic) hipLaunchKernelGGL((k_finish<R, true, true>), fg, dim3(kFinishThreads), 0, stream_, lats, S);
        else if (any_loads_) hipLaunchKernelGGL((k_finish<R, true, false>), fg, dim3(kFinishThreads), 0, stream_, lats, S);
        else if (plastic) hipLaunchKernelGGL((k_finish<R, false, true>), fg, dim3(kFinishThreads), 0, stream_, lats, S);
        else hipLaunchKernelGGL((k_finish<R, false, false>), fg, dim3(kFinishThreads), 0, stream_, lats, S);
        profEnd();
        HIP_OK(hipGetLastError());
    }

    void enqueueSubstepA() {
        if (shared_.direct) enqueueA<true>();
        else enqueueA<false>();
    }
    void enqueueSubstepB() {
        if (shared_.direct) enqueueB<true>();
        else enqueueB<false>();
    }

    // The contact pair list of every running slot whose flag asks for it
    // (LatticePhysics.hpp:1430-1561): cells, a stable sort by cell, the cell
    // table, the per-node gather and the pair list in order.
    void rebuildContacts() {
        for (uint32_t k = 0; k < slots_.size(); ++k) {
            Slot<R> &sl = *slots_[k];
            if (!sl.armed || !sl.contact || sl.ctl_host.stop || !sl.ctl_host.nc_rebuild) continue;
            if (sl.ctl_host.nc_error) throw Invalid("a node's contact cell is out of range (positions beyond 2^20 cells)");
            const DevLattice<R> &L = sl.L;
            const uint32_t g = gridFor(sl.N);
            profBegin(KContact, 14);
            hipLaunchKernelGGL((k_nc_cells<R>), dim3(g), dim3(kWG), 0, stream_, L, S_);
            size_t bytes = sl.nc_temp.bytes;
            // Stable sort by bucket: ascending node index within a bucket. The
            // buckets take log2(table) bits, plus one for the massless sentinel.
            const int bucket_bits = std::countr_zero(sl.nc_table_mask + 1U) + 1;
            HIP_OK(rocprim::radix_sort_pairs(sl.nc_temp.p, bytes, L.nc_bucket, L.nc_bucket_sorted, L.nc_node,
                                             L.nc_node_sorted, sl.N, 0, bucket_bits, stream_));
            hipLaunchKernelGGL(k_nc_table_clear, dim3(gridFor(sl.nc_table_mask + 1U)), dim3(kWG), 0, stream_,
                               L.nc_table_val, sl.nc_table_mask + 1U);
            hipLaunchKernelGGL((k_nc_table_insert<R>), dim3(g), dim3(kWG), 0, stream_, L);
            hipLaunchKernelGGL((k_nc_gather<R>), dim3(g), dim3(kWG), 0, stream_, L, S_);
            bytes = sl.nc_temp.bytes;
            HIP_OK(rocprim::exclusive_scan(sl.nc_temp.p, bytes, L.pair_fill, L.pair_offset, 0U, sl.N,
                                           rocprim::plus<uint32_t>(), stream_));
            hipLaunchKernelGGL((k_nc_flatten<R>), dim3(g), dim3(kWG), 0, stream_, L);
            // The narrow phase's schedule: pair ranks at their partners, and each
            // node's count of pairs naming it as partner.
            HIP_OK(hipStreamSynchronize(stream_));
            const uint32_t P = readCtl(k).nc_pairs;
            hipLaunchKernelGGL(k_nc_clear, dim3(g), dim3(kWG), 0, stream_, L.nc_partners, sl.N);
            if (P > 0) {
                hipLaunchKernelGGL(k_nc_iota, dim3(gridFor(P)), dim3(kWG), 0, stream_, L.nc_sort_in, P);
                bytes = sl.nc_temp.bytes;
                HIP_OK(rocprim::radix_sort_pairs(sl.nc_temp.p, bytes, L.flat_b, L.nc_sort_key, L.nc_sort_in,
                                                 L.nc_sort_out, P, 0, 32, stream_));
                hipLaunchKernelGGL((k_nc_ranks<R>), dim3(gridFor(P)), dim3(kWG), 0, stream_, L, P);
            }
            profEnd();
            HIP_OK(hipGetLastError());
        }
    }
};

// Opaque handles of the C ABI.
struct EngineBase {
    virtual ~EngineBase() = default;
};
template <class R> struct EngineHolder final : EngineBase {
    Engine<R> e;
    EngineHolder(int device, uint32_t capacity) : e(device, capacity) {}
};
struct Handle {
    int precision = BLX_PRECISION_FLOAT;
    std::unique_ptr<EngineBase> impl;
};

template <class F> auto dispatch(Handle *h, F &&f) {
    if (!h) throw Invalid("null handle");
    if (h->precision == BLX_PRECISION_FLOAT) return f(static_cast<EngineHolder<float> *>(h->impl.get())->e);
    return f(static_cast<EngineHolder<double> *>(h->impl.get())->e);
}
template <class F> auto dispatch(const Handle *h, F &&f) { return dispatch(const_cast<Handle *>(h), std::forward<F>(f)); }

Handle *makeHandle(int device, int precision, uint32_t capacity) {
    if (deviceCount() <= device || device < 0)
        throw Runtime("no HIP device " + std::to_string(device) + " for the lattice backend");
    // Set-up of this device's label scratch address, once per device.
    (void)scratchOn(device);
    auto h = std::make_unique<Handle>();
    h->precision = precision;
    if (precision == BLX_PRECISION_FLOAT) h->impl = std::make_unique<EngineHolder<float>>(device, capacity);
    else if (precision == BLX_PRECISION_DOUBLE) h->impl = std::make_unique<EngineHolder<double>>(device, capacity);
    else throw Invalid("unknown precision");
    return h.release();
}

void copyName(const std::string &s, char *out, size_t cap) {
    if (out && cap) {
        std::strncpy(out, s.c_str(), cap - 1);
        out[cap - 1] = 0;
    }
}

} // namespace
} // namespace blx

using namespace blx;

#define H(b) reinterpret_cast<blx::Handle *>(b)
#define CH(b) reinterpret_cast<const blx::Handle *>(b)

extern "C" {

const char *blx_last_error(void) { return blx::lastError().c_str(); }

int blx_available(void) {
    const int n = deviceCount();
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t p{};
        if (hipGetDeviceProperties(&p, d) == hipSuccess && std::strncmp(p.gcnArchName, "gfx950", 6) == 0) return 1;
    }
    return 0;
}

int blx_description(char *out, size_t cap) {
    return guard([&] {
        std::string s;
        const int n = deviceCount();
        if (n == 0) s = "no HIP device visible";
        for (int d = 0; d < n; ++d) {
            hipDeviceProp_t p{};
            HIP_OK(hipGetDeviceProperties(&p, d));
            char line[512];
            std::snprintf(line, sizeof line, "%sdevice %d: %s (%s), %d CUs, %.1f GiB, %zu KiB LDS per workgroup",
                          d ? "; " : "", d, p.name, p.gcnArchName, p.multiProcessorCount,
                          double(p.totalGlobalMem) / (1024.0 * 1024.0 * 1024.0), size_t(p.sharedMemPerBlock) >> 10);
            s += line;
        }
        copyName(s, out, cap);
    });
}

uint32_t blx_multiprocessor_count(void) {
    if (deviceCount() == 0) return 0;
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 0;
    return uint32_t(p.multiProcessorCount);
}

// ---- single-lattice backend: an engine with one slot ----------------------

// A single lattice's default tuning: stage pairs for a lattice too large for
// LDS, the whole sweep in one workgroup for one that fits.
constexpr blx_tuning kBackendTuning{2U, 1U, 1024U, 0U, 0U};

int blx_backend_create(int device, int precision, const blx_schedule_view *sv, blx_backend **out) {
    return guard([&] {
        if (!sv || !out) throw Invalid("null argument");
        std::unique_ptr<Handle> h(makeHandle(device, precision, 1));
        dispatch(h.get(), [&](auto &e) {
            e.setSchedule(0, scheduleOf(*sv));
            e.setTuning(kBackendTuning);
        });
        *out = reinterpret_cast<blx_backend *>(h.release());
    });
}
int blx_backend_destroy(blx_backend *b) { return guard([&] { delete H(b); }); }
int blx_backend_name(const blx_backend *b, char *out, size_t cap) {
    return guard([&] { copyName(dispatch(CH(b), [&](auto &e) { return e.name(); }), out, cap); });
}
int blx_backend_upload(blx_backend *b, const blx_state_view *st, const blx_settings *s, const blx_sphere *sp) {
    return guard([&] {
        if (!b || !st || !s || !sp) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) { e.upload(0, *st, *s, *sp); });
    });
}
int blx_backend_set_blade(blx_backend *b, const blx_blade *blade) {
    return guard([&] {
        if (!blade) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) { e.setBlade(*blade); });
    });
}
int blx_backend_set_external_forces(blx_backend *b, const double *f, uint64_t n) {
    return guard([&] { dispatch(H(b), [&](auto &e) { e.setForces(0, f, n); }); });
}
int blx_backend_run(blx_backend *b, const blx_run_control *c, blx_run_status *st) {
    return guard([&] {
        if (!b || !c || !st) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) { e.run(*c, *st); });
    });
}
int blx_backend_status(const blx_backend *b, blx_run_status *st) {
    return guard([&] { dispatch(CH(b), [&](auto &e) { e.status(0, *st); }); });
}
int blx_backend_download(blx_backend *b, blx_state_view *st, blx_sphere *sp) {
    return guard([&] { dispatch(H(b), [&](auto &e) { e.download(0, *st, *sp); }); });
}
int blx_backend_first_failure_bonds(const blx_backend *b, uint32_t *out, uint64_t cap, uint64_t *count) {
    return guard([&] { *count = dispatch(CH(b), [&](auto &e) { return e.firstFailure(0, out, cap); }); });
}
int blx_backend_frame_count(const blx_backend *b, uint32_t *n) {
    return guard([&] { *n = dispatch(CH(b), [&](auto &e) { return e.frameCount(0); }); });
}
int blx_backend_take_frame(blx_backend *b, uint32_t k, uint64_t *step, float *u, uint8_t *alive, float *damage,
                           blx_sphere *sp) {
    return guard([&] { dispatch(H(b), [&](auto &e) { e.takeFrame(0, k, step, u, alive, damage, sp); }); });
}
int blx_backend_clear_frames(blx_backend *b) {
    return guard([&] { dispatch(H(b), [&](auto &e) { e.clearFrames(0); }); });
}
int blx_backend_fragment_labels(blx_backend *b, uint32_t *label, uint32_t *count) {
    return guard([&] { *count = dispatch(H(b), [&](auto &e) { return e.labels(0, label); }); });
}

int blx_backend_fragment_labels_noalloc(blx_backend *b, uint32_t *label_device, uint32_t *count) {
    return guard([&] {
        *count = dispatch(H(b), [&](auto &e) {
            return e.labelsNoAlloc(0, scratchOn(e.device()), kNoAllocBytes, kNoAllocMaxNodes, label_device);
        });
    });
}

int blx_backend_fragment_mass_properties(blx_backend *b, double cell_size_m, const double *spin,
                                         blx_fragment_mass *out, uint32_t capacity, uint32_t *count) {
    return guard([&] {
        *count = dispatch(H(b), [&](auto &e) { return e.massProperties(0, cell_size_m, spin, out, capacity); });
    });
}
int blx_backend_set_profiling(blx_backend *b, int on) {
    return guard([&] { dispatch(H(b), [&](auto &e) { e.setProfiling(on != 0); }); });
}
int blx_backend_get_tuning(const blx_backend *b, blx_tuning *t) {
    return guard([&] {
        if (!t) throw Invalid("null argument");
        *t = dispatch(CH(b), [&](auto &e) { return e.tuning(); });
    });
}
int blx_backend_set_tuning(blx_backend *b, const blx_tuning *t) {
    return guard([&] {
        if (!t) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) { e.setTuning(*t); });
    });
}
int blx_backend_autotune(blx_backend *b, uint64_t warm, uint64_t probe, blx_autotune_report *rep) {
    return guard([&] {
        if (!rep) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) {
            blx_run_control c{};
            c.max_steps = warm + probe;
            e.autotune(warm, probe, [&] { e.arm(0, c); }, *rep);
        });
    });
}
int blx_batch_get_tuning(const blx_batch *b, blx_tuning *t) {
    return guard([&] {
        if (!t) throw Invalid("null argument");
        *t = dispatch(CH(b), [&](auto &e) { return e.tuning(); });
    });
}
int blx_batch_set_tuning(blx_batch *b, const blx_tuning *t) {
    return guard([&] {
        if (!t) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) { e.setTuning(*t); });
    });
}
int blx_batch_autotune(blx_batch *b, uint64_t warm, uint64_t probe, blx_autotune_report *rep) {
    return guard([&] {
        if (!rep) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) { e.autotune(warm, probe, [] {}, *rep); });
    });
}
int blx_backend_kernel_stats(const blx_backend *b, uint32_t index, char *name, size_t cap, double *seconds,
                             uint64_t *launches, uint32_t *count) {
    return guard([&] {
        std::string n;
        double s = 0.0;
        uint64_t l = 0;
        *count = dispatch(CH(b), [&](auto &e) { return e.kernelStats(index, n, s, l); });
        if (index < *count) {
            copyName(n, name, cap);
            if (seconds) *seconds = s;
            if (launches) *launches = l;
        }
    });
}

// ---- batch of lattices ------------------------------------------------------

int blx_batch_create(int device, int precision, uint32_t capacity, blx_batch **out) {
    return guard([&] {
        if (!out) throw Invalid("null argument");
        *out = reinterpret_cast<blx_batch *>(makeHandle(device, precision, capacity));
    });
}
int blx_batch_destroy(blx_batch *b) { return guard([&] { delete H(b); }); }
int blx_batch_capacity(const blx_batch *b, uint32_t *capacity) {
    return guard([&] { *capacity = dispatch(CH(b), [&](auto &e) { return e.capacity(); }); });
}
int blx_batch_upload(blx_batch *b, uint32_t slot, const blx_schedule_view *sv, const blx_state_view *st,
                     const blx_settings *s, const blx_sphere *sp) {
    return guard([&] {
        if (!b || !sv || !st || !s || !sp) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) {
            e.setSchedule(slot, scheduleOf(*sv));
            e.upload(slot, *st, *s, *sp);
        });
    });
}
int blx_batch_clone(blx_batch *b, uint32_t dst, uint32_t src, const blx_sphere *sp) {
    return guard([&] {
        if (!sp) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) { e.clone(dst, src, *sp); });
    });
}
int blx_batch_set_external_forces(blx_batch *b, uint32_t slot, const double *f, uint64_t n) {
    return guard([&] { dispatch(H(b), [&](auto &e) { e.setForces(slot, f, n); }); });
}
int blx_batch_arm(blx_batch *b, uint32_t slot, const blx_run_control *c) {
    return guard([&] {
        if (!c) throw Invalid("null argument");
        if (c->capture_stride != 0) throw Invalid("frames are captured by the single-lattice backend only");
        dispatch(H(b), [&](auto &e) { e.arm(slot, *c); });
    });
}
int blx_batch_advance(blx_batch *b, uint64_t substeps, uint64_t steps_per_launch, const volatile uint8_t *abandon,
                      uint32_t *running) {
    return guard([&] {
        const uint32_t r = dispatch(H(b), [&](auto &e) {
            return e.advance(substeps, steps_per_launch ? steps_per_launch : 240,
                             typename std::decay_t<decltype(e)>::Abandon{abandon, nullptr, nullptr}, nullptr);
        });
        if (running) *running = r;
    });
}
int blx_batch_status(const blx_batch *b, uint32_t slot, blx_run_status *st) {
    return guard([&] { dispatch(CH(b), [&](auto &e) { e.status(slot, *st); }); });
}
int blx_batch_exit_reasons(const blx_batch *b, uint32_t *reasons, uint32_t count) {
    return guard([&] {
        if (!reasons && count) throw Invalid("null argument");
        dispatch(CH(b), [&](auto &e) { e.exitReasons(reasons, count); });
    });
}
int blx_batch_download(blx_batch *b, uint32_t slot, blx_state_view *st, blx_sphere *sp) {
    return guard([&] { dispatch(H(b), [&](auto &e) { e.download(slot, *st, *sp); }); });
}
int blx_batch_sphere(blx_batch *b, uint32_t slot, blx_sphere *sp) {
    return guard([&] {
        if (!sp) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) { e.sphereOf(slot, *sp); });
    });
}
int blx_batch_first_failure_bonds(const blx_batch *b, uint32_t slot, uint32_t *out, uint64_t cap, uint64_t *count) {
    return guard([&] { *count = dispatch(CH(b), [&](auto &e) { return e.firstFailure(slot, out, cap); }); });
}
int blx_batch_fragment_labels(blx_batch *b, uint32_t slot, uint32_t *label, uint32_t *count) {
    return guard([&] { *count = dispatch(H(b), [&](auto &e) { return e.labels(slot, label); }); });
}

int blx_batch_fragment_mass_properties(blx_batch *b, uint32_t slot, double cell_size_m, const double *spin,
                                       blx_fragment_mass *out, uint32_t capacity, uint32_t *count) {
    return guard([&] {
        *count = dispatch(H(b), [&](auto &e) { return e.massProperties(slot, cell_size_m, spin, out, capacity); });
    });
}
int blx_batch_set_blade(blx_batch *b, const blx_blade *blade) {
    return guard([&] {
        if (!blade) throw Invalid("null argument");
        dispatch(H(b), [&](auto &e) { e.setBlade(*blade); });
    });
}
int blx_batch_release(blx_batch *b, uint32_t slot) {
    return guard([&] { dispatch(H(b), [&](auto &e) { e.release(slot); }); });
}
int blx_batch_slot_bytes(const blx_batch *b, uint32_t slot, uint64_t *bytes) {
    return guard([&] { *bytes = dispatch(CH(b), [&](auto &e) { return uint64_t(e.slotBytes(slot)); }); });
}

// ---- measurement ------------------------------------------------------------

int blx_hbm_copy_probe(int device, uint64_t bytes, uint32_t repeats, double *seconds_per_copy) {
    return guard([&] {
        if (deviceCount() <= device || device < 0) throw Runtime("no HIP device for the copy probe");
        if (bytes < 16 || repeats == 0) throw Invalid("copy probe needs at least 16 bytes and one repeat");
        HIP_OK(hipSetDevice(device));
        const size_t n = size_t(bytes / 16);
        DevBuf src, dst;
        src.alloc(n * 16);
        dst.alloc(n * 16);
        HIP_OK(hipMemset(src.p, 1, n * 16));
        const size_t grid = (n + kWG - 1) / kWG;
        if (grid > 0x7fffffffULL) throw Invalid("copy probe larger than one launch");
        hipEvent_t a, b;
        HIP_OK(hipEventCreate(&a));
        HIP_OK(hipEventCreate(&b));
        hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(kWG), 0, nullptr, src.as<uint4>(), dst.as<uint4>(), n);
        HIP_OK(hipEventRecord(a, nullptr));
        for (uint32_t r = 0; r < repeats; ++r)
            hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(kWG), 0, nullptr, src.as<uint4>(), dst.as<uint4>(), n);
        HIP_OK(hipEventRecord(b, nullptr));
        HIP_OK(hipEventSynchronize(b));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *seconds_per_copy = double(ms) * 1e-3 / repeats;
    });
}

} // extern "C"
