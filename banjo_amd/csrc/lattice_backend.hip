// The MI355X lattice backend: device memory, the substep launch sequence and
// the C ABI of include/banjo_lattice.h. It is the HIP counterpart of the
// reference's LatticeBackend (src/fastlattice/FastLattice.hpp:255-369), whose
// CUDA implementation the reference no longer carries
// (src/fastlattice/CudaLatticeUnavailable.cpp:1-27).

#include "banjo_lattice.h"
#include "lattice_errors.hpp"
#include "lattice_kernels.hpp"

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
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

// One device allocation, freed with the backend.
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

template <class T> void toDevice(DevBuf &b, const std::vector<T> &v) {
    b.alloc(v.size() * sizeof(T));
    if (!v.empty()) HIP_OK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

template <class T> void zero(DevBuf &b, size_t count) {
    b.alloc(count * sizeof(T));
    HIP_OK(hipMemset(b.p, 0, count * sizeof(T)));
}

class BackendBase {
public:
    virtual ~BackendBase() = default;
    virtual std::string name() const = 0;
    virtual void upload(const blx_state_view &, const blx_settings &, const blx_sphere &) = 0;
    virtual void setForces(const double *, uint64_t) = 0;
    virtual void run(const blx_run_control &, blx_run_status &) = 0;
    virtual void status(blx_run_status &) const = 0;
    virtual void download(blx_state_view &, blx_sphere &) = 0;
    virtual uint64_t firstFailure(uint32_t *, uint64_t) const = 0;
    virtual uint32_t frameCount() const = 0;
    virtual void takeFrame(uint32_t, uint64_t *, float *, uint8_t *, float *, blx_sphere *) = 0;
    virtual void clearFrames() = 0;
    virtual uint32_t labels(uint32_t *) = 0;
    virtual void setProfiling(bool) = 0;
    virtual uint32_t kernelStats(uint32_t, std::string &, double &, uint64_t &) const = 0;
};

struct Schedule {
    uint32_t N = 0, B = 0, blocks = 0, colors = 0;
    std::vector<uint32_t> nbb, range_begin, range_end;
};

int deviceCount() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

template <class R> struct HostFrame {
    uint64_t step;
    std::vector<float> u;
    std::vector<uint8_t> alive;
    std::vector<float> damage;
    Sphere<R> sphere;
};

template <class R> class Backend final : public BackendBase {
public:
    Backend(int device, Schedule s) : device_(device), sch_(std::move(s)) {
        HIP_OK(hipSetDevice(device_));
        HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&early_, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&ev_bond_end_, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&ev_rinv_, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&ev_pass1_, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&ev_velocity_, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&ev_early_, hipEventDisableTiming));
        HIP_OK(hipEventCreate(&ev0_));
        HIP_OK(hipEventCreate(&ev1_));
        hipDeviceProp_t prop{};
        HIP_OK(hipGetDeviceProperties(&prop, device_));
        cus_ = uint32_t(prop.multiProcessorCount);
        buildStages();
    }
    ~Backend() override {
        (void)hipSetDevice(device_);
        if (stream_) (void)hipStreamSynchronize(stream_);
        if (ev0_) (void)hipEventDestroy(ev0_);
        if (ev1_) (void)hipEventDestroy(ev1_);
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
        if (side_) (void)hipStreamSynchronize(side_);
        if (ev_bond_end_) (void)hipEventDestroy(ev_bond_end_);
        if (ev_rinv_) (void)hipEventDestroy(ev_rinv_);
        if (early_) (void)hipStreamSynchronize(early_);
        for (hipEvent_t e : {ev_pass1_, ev_velocity_, ev_early_})
            if (e) (void)hipEventDestroy(e);
        if (early_) (void)hipStreamDestroy(early_);
        if (side_) (void)hipStreamDestroy(side_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    std::string name() const override { return std::string("hip-gfx950-") + (sizeof(R) == 4 ? "float" : "double"); }

    void upload(const blx_state_view &st, const blx_settings &s, const blx_sphere &sp) override {
        HIP_OK(hipSetDevice(device_));
        validate(st, s);
        const uint32_t N = st.node_count, B = st.bond_count, D = st.max_degree;
        N_ = N;
        B_ = B;
        D_ = D;
        direct_ = s.direct_arithmetic != 0;
        plastic_ = s.plastic_yield_stretch > 0.0;
        S_ = convertStep(s);
        iterations_ = s.constraint_iterations;
        const size_t P = size_t(D) * N;
        std::vector<R4<R>> x04(N), u4(N), up4(N), v4(N);
        for (uint32_t i = 0; i < N; ++i) {
            x04[i] = {R(st.x0[3 * size_t(i)]), R(st.x0[3 * size_t(i) + 1]), R(st.x0[3 * size_t(i) + 2]), R(st.mass[i])};
            u4[i] = {R(st.u[3 * size_t(i)]), R(st.u[3 * size_t(i) + 1]), R(st.u[3 * size_t(i) + 2]), R(st.inv_mass[i])};
            up4[i] = {R(st.u_prev[3 * size_t(i)]), R(st.u_prev[3 * size_t(i) + 1]), R(st.u_prev[3 * size_t(i) + 2]), R(0)};
            v4[i] = {R(st.v[3 * size_t(i)]), R(st.v[3 * size_t(i) + 1]), R(st.v[3 * size_t(i) + 2]), R(0)};
        }
        mass_host_.resize(N);
        inv_mass_host_.resize(N);
        for (uint32_t i = 0; i < N; ++i) {
            mass_host_[i] = R(st.mass[i]);
            inv_mass_host_[i] = R(st.inv_mass[i]);
        }
        toDevice(x04_, x04);
        toDevice(u4_, u4);
        toDevice(up4_, up4);
        toDevice(v4_, v4);
        // Bonds in node-major order for the end sample: a counting sort by
        // the first end, stable in schedule index.
        std::vector<uint32_t> node_of(B);  // node-major position -> schedule index
        std::vector<uint32_t> pos_of(B);   // schedule index -> node-major position
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
        node_j_host_ = node_of;
        // Neighbour slots as slot palette indices.
        {
            Dedup<4 * sizeof(R) / 4 + 1> table;
            std::vector<uint32_t> key(P), bond(P);
            for (size_t k = 0; k < P; ++k) {
                if (st.nbr_bond[k] == BLX_NO_BOND) {
                    key[k] = kNone;
                    bond[k] = kNone;
                    continue;
                }
                if (st.nbr_bond[k] >= B || st.nbr_other[k] >= N) throw Invalid("neighbour slot out of range");
                const uint32_t i = uint32_t(k % N);
                const R rw[4] = {R(st.nbr_rest[3 * k]), R(st.nbr_rest[3 * k + 1]), R(st.nbr_rest[3 * k + 2]),
                                 R(st.nbr_weight[k])};
                const int32_t delta = int32_t(int64_t(st.nbr_other[k]) - int64_t(i));
                typename Dedup<4 * sizeof(R) / 4 + 1>::Key kk{};
                std::memcpy(kk.data(), rw, sizeof(rw));
                std::memcpy(kk.data() + 4 * sizeof(R) / 4, &delta, 4);
                const uint32_t id = table.index(kk);
                if (id >= kDeadSlot) throw Overflow("more than 2^31 distinct neighbour slots");
                key[k] = id | (st.nbr_alive[k] ? 0U : kDeadSlot);
                bond[k] = pos_of[st.nbr_bond[k]];
            }
            const auto &keys = table.keys();
            std::vector<R4<R>> pal(keys.size());
            std::vector<int32_t> dl(keys.size());
            for (size_t q = 0; q < keys.size(); ++q) {
                std::memcpy(&pal[q], keys[q].data(), 4 * sizeof(R));
                std::memcpy(&dl[q], keys[q].data() + 4 * sizeof(R) / 4, 4);
            }
            slot_palette_ = uint32_t(keys.size());
            toDevice(nbr_key_, key);
            toDevice(nbr_bond_, bond);
            toDevice(spal_, pal);
            toDevice(sdelta_, dl);
        }
        {
            // Edge palette: the distinct (rest edge, rest length, |e|^2 - L^2) in
            // backend precision; the direction terms are filled on the device.
            Dedup<5 * sizeof(R) / 4> table;
            std::vector<uint32_t> pal_of(B);
            for (uint32_t k = 0; k < B; ++k) {
                const R v[5] = {R(st.rest_edge[3 * size_t(k)]), R(st.rest_edge[3 * size_t(k) + 1]),
                                R(st.rest_edge[3 * size_t(k) + 2]), R(st.rest_length[k]), R(st.rest_length_sq_minus[k])};
                typename Dedup<5 * sizeof(R) / 4>::Key kk{};
                std::memcpy(kk.data(), v, sizeof(v));
                pal_of[k] = table.index(kk);
                if (pal_of[k] >= kNoDir) throw Overflow("more than 2^31 distinct bond edges");
            }
            const auto &keys = table.keys();
            const uint32_t np = uint32_t(keys.size());
            edge_palette_ = np;
            std::vector<EdgePal<R>> pal(np);
            for (uint32_t q = 0; q < np; ++q) {
                R v[5];
                std::memcpy(v, keys[q].data(), sizeof(v));
                pal[q].e = {v[0], v[1], v[2], v[3]};
                pal[q].d = {v[4], R(0), R(0), R(0)};
            }
            toDevice(epal_, pal);
            DevBuf nodir;
            nodir.alloc(size_t(np) * 4);
            hipLaunchKernelGGL((k_edge_palette<R>), dim3(gridFor(np)), dim3(kWG), 0, stream_, epal_.as<EdgePal<R>>(),
                               nodir.as<uint32_t>(), np);
            HIP_OK(hipGetLastError());
            HIP_OK(hipStreamSynchronize(stream_));
            std::vector<uint32_t> no_dir(np);
            HIP_OK(hipMemcpy(no_dir.data(), nodir.p, size_t(np) * 4, hipMemcpyDeviceToHost));
            for (uint32_t k = 0; k < B; ++k) pal_of[k] |= no_dir[pal_of[k]] ? kNoDir : 0U;

            std::vector<BondRec<R>> rec(B);
            std::vector<BondNode> bnd(B);
            std::vector<R4<R>> pvd(B);
            std::vector<R> pext(B), pstr(B);
            std::vector<uint2> slots(B);
            std::vector<uint8_t> alive_n(B), mode(B);
            bool tensile = true;
            for (uint32_t k = 0; k < B && tensile; ++k) {
                const double *t = st.threshold + 6 * size_t(k);
                tensile = !std::isfinite(R(t[2])) && !std::isfinite(R(t[4]));
            }
            tensile_only_ = tensile;
            std::vector<Thresholds<R>> thr(tensile ? 0 : B);
            std::vector<Tensile<R>> thr2(tensile ? B : 0);
            for (uint32_t k = 0; k < B; ++k) {
                BondRec<R> r{};
                r.a = st.bond_a[k];
                r.b = st.bond_b[k];
                r.p = pal_of[k];
                r.c = R(st.compliance[k]);
                rec[k] = r;
                pext[k] = R(st.plastic_extension[k]);
            }
            for (uint32_t o = 0; o < B; ++o) {
                const uint32_t k = node_of[o];
                bnd[o] = BondNode{st.bond_a[k], st.bond_b[k], k, pal_of[k]};
                const double *t = st.threshold + 6 * size_t(k);
                if (tensile) thr2[o] = {R(t[0]), R(t[1])};
                else thr[o] = {R(t[0]), R(t[1]), R(t[2]), R(t[3]), R(t[4]), R(t[5])};
                pvd[o] = {R(st.prev_tensile[k]), R(st.prev_compressive[k]), R(st.prev_shear[k]), R(st.damage[k])};
                pstr[o] = R(st.plastic_strain[k]);
                slots[o] = make_uint2(st.bond_slot_a[k], st.bond_slot_b[k]);
                alive_n[o] = st.alive[k];
                mode[o] = st.failure_mode[k];
            }
            toDevice(rec_, rec);
            toDevice(bnd_, bnd);
            toDevice(thr_, thr);
            toDevice(thr2_, thr2);
            toDevice(pvd_, pvd);
            toDevice(pext_, pext);
            toDevice(pstr_, pstr);
            toDevice(slots_, slots);
            toDevice(alive_, std::vector<uint8_t>(st.alive, st.alive + B));
            toDevice(alive_n_, alive_n);
            toDevice(mode_, mode);
        }
        zero<R>(lambda_, B);
        zero<uint32_t>(first_fail_, B);
        zero<R>(rinv_, 9 * size_t(N));
        zero<R>(unm_, 3 * size_t(N));
        zero<R4<R>>(sa_, N);
        zero<R4<R>>(sb_, N);
        zero<uint8_t>(valid_, N);
        zero<R>(approach_, 2 * size_t(N));
        zero<uint8_t>(engaged_, 2 * size_t(N));
        zero<uint8_t>(cand_, N);
        node_tiles_ = (N + kWG - 1) / kWG;
        bond_tiles_ = (B + kWG - 1) / kWG;
        const uint32_t event_words = (B + 63) / 64;
        group_words_ = (event_words + 63) / 64;
        zero<uint32_t>(tile_cand_, node_tiles_);
        zero<uint64_t>(ev_bits_, event_words);
        zero<uint64_t>(brk_bits_, event_words);
        zero<uint64_t>(grp_bits_, group_words_);
        zero<float>(ev_d_, B);
        zero<R>(ev_e_, B);
        zero<uint32_t>(flag_list_, event_words);
        zero<double>(tile_plastic_, bond_tiles_);
        zero<double>(load_part_, 13 * size_t(node_tiles_));
        zero<R4<R>>(fext_, N);
        zero<double>(freq_, 3 * size_t(N));
        zero<uint32_t>(cand_all_, std::max<size_t>(N, size_t(sch_.blocks) * kCap));
        zero<uint32_t>(rank_stamp_, N);
        zero<uint32_t>(affected_, N);
        zero<uint32_t>(affected_list_, N);
        zero<uint32_t>(cand_tile_, size_t(node_tiles_) * kWG);
        {
            std::vector<uint32_t> all(N);
            for (uint32_t i = 0; i < N; ++i) all[i] = i;
            toDevice(dirty_list_, all);  // upload marks every node dirty
        }
        zero<uint32_t>(dirty_stamp_, N);
        zero<uint32_t>(cand_lo_, sch_.blocks);
        zero<uint32_t>(cand_cnt_, sch_.blocks);
        toDevice(nbb_, sch_.nbb);
        // Control block and sphere.
        DevCtl c{};
        c.first_failure_step = c.last_failure_step = c.last_energy_gain_step = ~0ULL;
        c.dirty_start = 1;
        c.dirty_count = N;
        c.rank_slot = 1;            // booked by the first substep's k_finish
        c.rinv_stamp = 0xffffffffU;
        c.flat_fraction = 1.0e-3;
        c.calm_margin = 0.5;
        ctl_host_ = c;
        toDevice(ctl_, std::vector<DevCtl>{c});
        Sphere<R> sph{{R(sp.center[0]), R(sp.center[1]), R(sp.center[2])},
                      {R(sp.velocity[0]), R(sp.velocity[1]), R(sp.velocity[2])},
                      {R(sp.angular_velocity[0]), R(sp.angular_velocity[1]), R(sp.angular_velocity[2])},
                      R(sp.radius), R(sp.mass), R(sp.inertia)};
        toDevice(sphere_, std::vector<Sphere<R>>{sph});
        for (int k = 0; k < 3; ++k) origin_[k] = st.origin[k];
        loads_set_ = false;
        frames_.clear();
        run_status_ = blx_run_status{};
        run_status_.first_failure_step = run_status_.last_failure_step = run_status_.last_energy_gain_step = ~0ULL;
        bind();
        // Every node starts dirty: its inverse rest covariance before the first substep.
        hipLaunchKernelGGL((k_rinv<R>), dim3(gridFor(N)), dim3(kWG), 0, stream_, L_);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(stream_));
        uploaded_ = true;
        HIP_OK(hipDeviceSynchronize());
    }

    // ExternalLoads.hpp:28-50: validate the whole field, then replace the load.
    void setForces(const double *f, uint64_t substeps) override {
        if (!uploaded_) throw Logic("external forces need an uploaded lattice");
        HIP_OK(hipSetDevice(device_));
        if (f == nullptr) {
            if (substeps != 0) throw Invalid("empty external load needs zero substeps");
            setRemaining(0);
            return;
        }
        if (substeps == 0) throw Invalid("external load needs one force per node and a finite nonzero span");
        std::vector<R4<R>> fe(N_);
        std::vector<double> fr(3 * size_t(N_));
        bool any = false;
        for (uint32_t i = 0; i < N_; ++i) {
            const double x = f[3 * size_t(i)], y = f[3 * size_t(i) + 1], z = f[3 * size_t(i) + 2];
            const R rx = R(x), ry = R(y), rz = R(z);
            if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z) || !std::isfinite(std::hypot(x, y, z)) ||
                !std::isfinite(rx) || !std::isfinite(ry) || !std::isfinite(rz))
                throw Invalid("external force is not finite in backend precision");
            if (x == 0 && y == 0 && z == 0) {
                fe[i] = {R(0), R(0), R(0), R(0)};
                continue;
            }
            if (!(mass_host_[i] > 0) || !(inv_mass_host_[i] > 0) || !std::isfinite(mass_host_[i]) ||
                !std::isfinite(inv_mass_host_[i]))
                throw Invalid("external force needs a finite positive movable node mass");
            fe[i] = {rx, ry, rz, R(1)};
            fr[3 * size_t(i)] = x;
            fr[3 * size_t(i) + 1] = y;
            fr[3 * size_t(i) + 2] = z;
            any = true;
        }
        HIP_OK(hipMemcpy(fext_.p, fe.data(), fe.size() * sizeof(R4<R>), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(freq_.p, fr.data(), fr.size() * sizeof(double), hipMemcpyHostToDevice));
        loads_set_ = any;
        setRemaining(any ? substeps : 0);
    }

    void run(const blx_run_control &c, blx_run_status &out) override {
        if (!uploaded_) throw Logic("run before upload");
        HIP_OK(hipSetDevice(device_));
        const auto t0 = std::chrono::steady_clock::now();
        DevCtl ctl = readCtl();
        ctl.stop = 0;
        ctl.exit_reason = 0;
        ctl.quiet = c.quiet_steps;
        ctl.min_steps = c.min_steps;
        ctl.no_failure = c.no_failure_steps;
        ctl.energy_flat = c.energy_flat_steps;
        ctl.calm = c.calm_steps;
        ctl.flat_fraction = c.energy_flat_fraction;
        ctl.calm_margin = c.calm_damage_margin;
        ctl.removable = c.removable_energy_j;
        writeCtl(ctl);
        const uint64_t batch = profilingBatch(c.steps_per_launch ? c.steps_per_launch : kDefaultBatch);
        uint64_t done = 0;
        unsigned reason = 0;
        while (done < c.max_steps) {
            const uint64_t step = ctl.step;
            uint64_t n = std::min<uint64_t>(batch, c.max_steps - done);
            if (c.capture_stride != 0) {
                if (step % c.capture_stride == 0 && run_status_.frames_captured < c.max_frames) {
                    captureFrame(step);
                    ++run_status_.frames_captured;
                }
                const uint64_t to_next = c.capture_stride - step % c.capture_stride;
                n = std::min(n, to_next);
            }
            HIP_OK(hipEventRecord(ev0_, stream_));
            for (uint64_t k = 0; k < n; ++k) enqueueSubstep();
            HIP_OK(hipEventRecord(ev1_, stream_));
            HIP_OK(hipStreamSynchronize(stream_));
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, ev0_, ev1_));
            run_status_.kernel_seconds += double(ms) * 1e-3;
            profCollect();
            ++run_status_.launches;
            ctl = readCtl();
            done += ctl.step - step;
            if (ctl.error) throw Overflow("external force kick exceeds backend precision");
            if (ctl.stop) {
                reason = ctl.exit_reason;
                break;
            }
            if (c.abandon && *c.abandon) {
                reason = 7;
                break;
            }
        }
        if (reason == 0 && done >= c.max_steps) reason = 3;
        run_status_.exit_reason = reason;
        run_status_.wall_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        ctl_host_ = ctl;
        status(out);
    }

    void status(blx_run_status &o) const override {
        const DevCtl &c = ctl_host_;
        o = run_status_;
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
        for (int k = 0; k < 3; ++k) o.contact.impulse_to_material[k] = c.impulse[k];
        o.contact.dissipated_kinetic_energy_j = c.dissipated;
        o.contact.maximum_penetration_m = c.max_penetration;
        o.contact.maximum_position_correction_m = c.max_correction;
        o.contact.maximum_center_shift_m = c.max_shift;
        o.contact.candidate_overflow = c.candidate_overflow;
        o.contact.ball_support_events = c.ball_support_events;
        o.external_load.steps = c.ledger_steps;
        o.external_load.elapsed_s = c.ledger_elapsed;
        for (int k = 0; k < 3; ++k) {
            o.external_load.requested_impulse_n_s[k] = c.ledger_req[k];
            o.external_load.impulse_n_s[k] = c.ledger_imp[k];
            o.external_load.requested_angular_impulse_kg_m2_s[k] = c.ledger_areq[k];
            o.external_load.angular_impulse_kg_m2_s[k] = c.ledger_aimp[k];
        }
        o.external_load.work_j = c.ledger_work;
        std::memcpy(&o.max_tensile_stretch, &c.peak_bits[0], 4);
        std::memcpy(&o.max_compressive_strain, &c.peak_bits[1], 4);
        std::memcpy(&o.max_shear_strain, &c.peak_bits[2], 4);
        std::memcpy(&o.max_plastic_stretch, &c.pstretch_bits, 4);
        o.rank_deficient_nodes = c.rank_deficient;
        o.plastic_work_j = c.plastic_work;
    }

    void download(blx_state_view &st, blx_sphere &sp) override {
        if (!uploaded_) throw Logic("download before upload");
        HIP_OK(hipSetDevice(device_));
        HIP_OK(hipStreamSynchronize(stream_));
        std::vector<R4<R>> u4(N_), up4(N_), v4(N_), pvd(B_);
        std::vector<R> pext(B_), pstr(B_);
        std::vector<uint8_t> mode(B_);
        copyDown(u4, u4_);
        copyDown(up4, up4_);
        copyDown(v4, v4_);
        copyDown(pvd, pvd_);
        copyDown(pext, pext_);
        copyDown(pstr, pstr_);
        copyDown(mode, mode_);
        HIP_OK(hipMemcpy(st.alive, alive_.p, B_, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < N_; ++i) {
            st.u[3 * size_t(i)] = double(u4[i].x);
            st.u[3 * size_t(i) + 1] = double(u4[i].y);
            st.u[3 * size_t(i) + 2] = double(u4[i].z);
            st.u_prev[3 * size_t(i)] = double(up4[i].x);
            st.u_prev[3 * size_t(i) + 1] = double(up4[i].y);
            st.u_prev[3 * size_t(i) + 2] = double(up4[i].z);
            st.v[3 * size_t(i)] = double(v4[i].x);
            st.v[3 * size_t(i) + 1] = double(v4[i].y);
            st.v[3 * size_t(i) + 2] = double(v4[i].z);
        }
        for (uint32_t k = 0; k < B_; ++k) st.plastic_extension[k] = double(pext[k]);
        // The end-sample arrays are node-major on the device.
        for (uint32_t o = 0; o < B_; ++o) {
            const uint32_t k = node_j_host_[o];
            st.prev_tensile[k] = double(pvd[o].x);
            st.prev_compressive[k] = double(pvd[o].y);
            st.prev_shear[k] = double(pvd[o].z);
            st.damage[k] = double(pvd[o].w);
            st.plastic_strain[k] = double(pstr[o]);
            st.failure_mode[k] = mode[o];
        }
        sphereOut(readSphere(), sp);
    }

    uint64_t firstFailure(uint32_t *out, uint64_t cap) const override {
        const uint64_t n = ctl_host_.first_failure_count;
        if (out && cap) HIP_OK(hipMemcpy(out, first_fail_.p, std::min(n, cap) * sizeof(uint32_t), hipMemcpyDeviceToHost));
        return n;
    }

    uint32_t frameCount() const override { return uint32_t(frames_.size()); }
    void takeFrame(uint32_t k, uint64_t *step, float *u, uint8_t *alive, float *damage, blx_sphere *sp) override {
        if (k >= frames_.size()) throw Invalid("frame index out of range");
        const HostFrame<R> &f = frames_[k];
        *step = f.step;
        std::copy(f.u.begin(), f.u.end(), u);
        std::copy(f.alive.begin(), f.alive.end(), alive);
        std::copy(f.damage.begin(), f.damage.end(), damage);
        if (sp) sphereOut(f.sphere, *sp);
    }
    void clearFrames() override {
        frames_.clear();
        run_status_.frames_captured = 0;
    }

    uint32_t labels(uint32_t *label) override {
        if (!uploaded_) throw Logic("fragments before upload");
        HIP_OK(hipSetDevice(device_));
        const uint32_t c = labelsOnDevice();
        HIP_OK(hipMemcpyAsync(label, cc_label_.p, size_t(N_) * 4, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        return c;
    }

    // Union-find, roots, canonical rank and labels, all on the device; the
    // labels are left in cc_label_. Returns the component count.
    uint32_t labelsOnDevice() {
        const size_t n4 = size_t(N_) * 4;
        cc_parent_.alloc(n4);
        cc_root_.alloc(n4);
        cc_size_.alloc(n4);
        cc_rank_.alloc(n4);
        cc_label_.alloc(n4);
        cc_keys_.alloc(size_t(N_) * 8);
        cc_sorted_.alloc(size_t(N_) * 8);
        cc_count_.alloc(4);
        size_t temp = 0;
        HIP_OK(rocprim::radix_sort_keys(nullptr, temp, cc_keys_.as<uint64_t>(), cc_sorted_.as<uint64_t>(), N_, 0, 64,
                                        stream_));
        cc_temp_.alloc(temp);
        const uint32_t g = gridFor(N_);
        auto *parent = cc_parent_.as<uint32_t>(), *root = cc_root_.as<uint32_t>(), *size = cc_size_.as<uint32_t>();
        HIP_OK(hipMemsetAsync(cc_count_.p, 0, 4, stream_));
        hipLaunchKernelGGL(k_cc_init, dim3(g), dim3(kWG), 0, stream_, parent, size, N_);
        hipLaunchKernelGGL(k_cc_hook<R>, dim3(gridFor(B_)), dim3(kWG), 0, stream_, L_, parent);
        hipLaunchKernelGGL(k_cc_roots, dim3(g), dim3(kWG), 0, stream_, parent, root, size, N_);
        hipLaunchKernelGGL(k_cc_keys, dim3(g), dim3(kWG), 0, stream_, root, size, cc_keys_.as<uint64_t>(),
                           cc_count_.as<uint32_t>(), N_);
        HIP_OK(hipGetLastError());
        uint32_t c = 0;
        HIP_OK(hipMemcpyAsync(&c, cc_count_.p, 4, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        size_t bytes = cc_temp_.bytes;
        HIP_OK(rocprim::radix_sort_keys(cc_temp_.p, bytes, cc_keys_.as<uint64_t>(), cc_sorted_.as<uint64_t>(), c, 0, 64,
                                        stream_));
        hipLaunchKernelGGL(k_cc_rank, dim3(gridFor(c)), dim3(kWG), 0, stream_, cc_sorted_.as<uint64_t>(),
                           cc_rank_.as<uint32_t>(), c);
        hipLaunchKernelGGL(k_cc_label, dim3(g), dim3(kWG), 0, stream_, root, cc_rank_.as<uint32_t>(),
                           cc_label_.as<uint32_t>(), N_);
        HIP_OK(hipGetLastError());
        return c;
    }

private:
    static constexpr uint64_t kDefaultBatch = 240;

    int device_;
    Schedule sch_;
    // stream_ carries the substep; side_ the inverse rest covariances that
    // overlap the folds; early_ the end sample of the nodes and bonds striker
    // pass 2 cannot reach, which overlaps pass 2.
    hipStream_t stream_ = nullptr, side_ = nullptr, early_ = nullptr;
    hipEvent_t ev_bond_end_ = nullptr, ev_rinv_ = nullptr, ev_pass1_ = nullptr, ev_velocity_ = nullptr,
               ev_early_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    uint32_t cus_ = 256;
    uint32_t N_ = 0, B_ = 0, D_ = 0, iterations_ = 1;
    uint32_t node_tiles_ = 0, bond_tiles_ = 0, group_words_ = 0;
    uint32_t slot_palette_ = 0, edge_palette_ = 0;
    bool direct_ = false, plastic_ = false, uploaded_ = false, loads_set_ = false, tensile_only_ = false;
    std::vector<uint32_t> node_j_host_;  // node-major bond position -> schedule index
    Step<R> S_{};
    double origin_[3]{};
    std::vector<R> mass_host_, inv_mass_host_;
    DevLattice<R> L_{};
    DevCtl ctl_host_{};
    blx_run_status run_status_{};
    std::vector<HostFrame<R>> frames_;
    // Stage tables: per colour stage, a run of workgroups over [begin, end) bond ranges.
    std::vector<uint2> wg_host_;
    std::vector<std::pair<uint32_t, uint32_t>> stages_; // (first workgroup, workgroup count)
    size_t interior_stages_ = 0;
    DevBuf wg_;
    DevBuf x04_, u4_, up4_, v4_, rinv_, unm_, sa_, sb_, valid_, approach_, engaged_, cand_;
    DevBuf affected_, affected_list_;
    DevBuf rank_stamp_, dirty_list_, dirty_stamp_, cand_tile_, nbr_bond_, ev_bits_, brk_bits_, grp_bits_, ev_d_, ev_e_;
    DevBuf flag_list_, nbr_key_, spal_, sdelta_, epal_;
    DevBuf rec_, alive_, lambda_, pext_, bnd_, alive_n_, pstr_, thr_, thr2_, pvd_, mode_, slots_;
    DevBuf tile_cand_, tile_plastic_, load_part_, fext_, freq_, cand_all_, cand_lo_, cand_cnt_, nbb_;
    DevBuf first_fail_, ctl_, sphere_;
    DevBuf cc_parent_, cc_root_, cc_size_, cc_rank_, cc_label_, cc_keys_, cc_sorted_, cc_count_, cc_temp_;

    template <class T> void copyDown(std::vector<T> &dst, const DevBuf &src) {
        if (!dst.empty()) HIP_OK(hipMemcpy(dst.data(), src.p, dst.size() * sizeof(T), hipMemcpyDeviceToHost));
    }

    uint32_t gridFor(uint32_t n) const {
        const uint32_t need = (n + kWG - 1) / kWG;
        return std::max(1U, std::min(need, cus_ * 16U));
    }

    void validate(const blx_state_view &st, const blx_settings &s) const {
        if (s.bond_integrator == BLX_BOND_VELOCITY_VERLET)
            throw Invalid("the Verlet bond integrator is the serial double CPU reference only");
        if (s.bond_integrator != BLX_BOND_XPBD) throw Invalid("unknown bond integrator");
        if (s.node_contact.mode != BLX_NODE_CONTACT_OFF) throw Invalid("node-node contact is not available on the HIP backend");
        if (s.plate_half_thickness != 0.0) throw Invalid("plate bending is not available on the HIP backend");
        if (s.audit_energy) throw Invalid("the kinetic energy audit is not available on the HIP backend");
        if (s.plane_count > BLX_MAX_SUPPORT_PLANES) throw Invalid("at most two support planes");
        for (uint32_t p = 0; p < s.plane_count; ++p)
            if (s.planes[p].footprint_count > BLX_MAX_FOOTPRINTS) throw Invalid("at most four footprints per plane");
        if (s.constraint_iterations == 0) throw Invalid("at least one constraint iteration");
        if (st.node_count != sch_.N || st.bond_count != sch_.B) throw Invalid("state does not match the schedule");
        if (size_t(st.max_degree) * st.node_count > 0xffffffffULL) throw Invalid("neighbour slots exceed 32-bit indices");
        if (st.bond_count >= 0x80000000U) throw Invalid("at most 2^31 - 1 bonds");
        if (!(s.dt > 0.0) || !std::isfinite(s.dt)) throw Invalid("the substep must be finite and positive");
        if (s.sphere_enabled && (st.node_count + kWG - 1) / kWG > kMaxNodeTiles)
            throw Invalid("the striker pass holds at most 4,194,304 nodes");
    }

    static Step<R> convertStep(const blx_settings &s) {
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

    // Colour stages in sweep order (CpuLatticeBackend.cpp:395-404): interior
    // bonds of every block colour by colour, then the boundary bonds of even
    // blocks, then of odd blocks. Interior bonds of different blocks share no
    // node, nor do boundary bonds of blocks of one parity, so the ranges of
    // one colour across those blocks form one node-disjoint stage.
    void buildStages() {
        const uint32_t cc = std::max(sch_.colors, 1U);
        auto add = [&](uint32_t boundary, uint32_t parity_first, uint32_t step) {
            for (uint32_t c = 0; c < sch_.colors; ++c) {
                const uint32_t first = uint32_t(wg_host_.size());
                for (uint32_t b = parity_first; b < sch_.blocks; b += step) {
                    const size_t key = (size_t(b) * 2 + boundary) * cc + c;
                    const uint32_t lo = sch_.range_begin[key], hi = sch_.range_end[key];
                    for (uint32_t j = lo; j < hi; j += kWG) wg_host_.push_back(make_uint2(j, std::min(hi, j + kWG)));
                }
                const uint32_t count = uint32_t(wg_host_.size()) - first;
                if (count) stages_.push_back({first, count});
            }
        };
        add(0, 0, 1);
        interior_stages_ = stages_.size();
        add(1, 0, 2);
        add(1, 1, 2);
        toDevice(wg_, wg_host_);
    }

    void bind() {
        DevLattice<R> &L = L_;
        L.N = N_;
        L.B = B_;
        L.D = D_;
        L.blocks = sch_.blocks;
        L.node_tiles = node_tiles_;
        L.bond_tiles = bond_tiles_;
        L.group_words = group_words_;
        L.tensile_only = tensile_only_ ? 1U : 0U;
        L.u4 = u4_.as<R4<R>>();
        L.up4 = up4_.as<R4<R>>();
        L.v4 = v4_.as<R4<R>>();
        L.x04 = x04_.as<R4<R>>();
        L.rinv = rinv_.as<R>();
        L.unm = unm_.as<R>();
        L.sa = sa_.as<R4<R>>();
        L.sb = sb_.as<R4<R>>();
        L.valid = valid_.as<uint8_t>();
        L.approach = approach_.as<R>();
        L.engaged = engaged_.as<uint8_t>();
        L.cand = cand_.as<uint8_t>();
        L.nbr_key = nbr_key_.as<uint32_t>();
        L.spal = spal_.as<R4<R>>();
        L.sdelta = sdelta_.as<int>();
        L.nbr_bond = nbr_bond_.as<uint32_t>();
        L.epal = epal_.as<EdgePal<R>>();
        L.rec = rec_.as<BondRec<R>>();
        L.alive = alive_.as<uint8_t>();
        L.lambda = lambda_.as<R>();
        L.pext = pext_.as<R>();
        L.bnd = bnd_.as<BondNode>();
        L.alive_n = alive_n_.as<uint8_t>();
        L.thr = thr_.as<Thresholds<R>>();
        L.thr2 = thr2_.as<Tensile<R>>();
        L.pvd = pvd_.as<R4<R>>();
        L.mode = mode_.as<uint8_t>();
        L.pstr = pstr_.as<R>();
        L.slots = slots_.as<uint2>();
        L.ev_bits = ev_bits_.as<uint64_t>();
        L.brk_bits = brk_bits_.as<uint64_t>();
        L.grp_bits = grp_bits_.as<uint64_t>();
        L.ev_d = ev_d_.as<float>();
        L.ev_e = ev_e_.as<R>();
        L.flag_list = flag_list_.as<uint32_t>();
        L.tile_cand = tile_cand_.as<uint32_t>();
        L.tile_plastic = tile_plastic_.as<double>();
        L.load_part = load_part_.as<double>();
        L.fext = fext_.as<R4<R>>();
        L.freq = freq_.as<double>();
        L.cand_all = cand_all_.as<uint32_t>();
        L.cand_tile = cand_tile_.as<uint32_t>();
        L.dirty_list = dirty_list_.as<uint32_t>();
        L.dirty_stamp = dirty_stamp_.as<uint32_t>();
        L.rank_stamp = rank_stamp_.as<uint32_t>();
        L.affected = affected_.as<uint32_t>();
        L.affected_list = affected_list_.as<uint32_t>();
        L.cand_lo = cand_lo_.as<uint32_t>();
        L.cand_cnt = cand_cnt_.as<uint32_t>();
        L.nbb = nbb_.as<uint32_t>();
        L.first_fail = first_fail_.as<uint32_t>();
        L.ctl = ctl_.as<DevCtl>();
        L.sphere = sphere_.as<Sphere<R>>();
        for (int k = 0; k < 3; ++k) L.origin[k] = origin_[k];
    }

    DevCtl readCtl() {
        DevCtl c;
        HIP_OK(hipMemcpy(&c, ctl_.p, sizeof(DevCtl), hipMemcpyDeviceToHost));
        return c;
    }
    void writeCtl(const DevCtl &c) { HIP_OK(hipMemcpy(ctl_.p, &c, sizeof(DevCtl), hipMemcpyHostToDevice)); }
    void setRemaining(uint64_t n) {
        DevCtl c = readCtl();
        c.load_remaining = n;
        writeCtl(c);
        ctl_host_.load_remaining = n;
    }
    Sphere<R> readSphere() {
        Sphere<R> s;
        HIP_OK(hipMemcpy(&s, sphere_.p, sizeof(s), hipMemcpyDeviceToHost));
        return s;
    }
    static void sphereOut(const Sphere<R> &s, blx_sphere &o) {
        o.center[0] = double(s.center.x); o.center[1] = double(s.center.y); o.center[2] = double(s.center.z);
        o.velocity[0] = double(s.velocity.x); o.velocity[1] = double(s.velocity.y); o.velocity[2] = double(s.velocity.z);
        o.angular_velocity[0] = double(s.omega.x); o.angular_velocity[1] = double(s.omega.y); o.angular_velocity[2] = double(s.omega.z);
        o.radius = double(s.radius);
        o.mass = double(s.mass);
        o.inertia = double(s.inertia);
    }

    void captureFrame(uint64_t step) {
        HostFrame<R> f;
        f.step = step;
        std::vector<R4<R>> u4(N_), pvd(B_);
        copyDown(u4, u4_);
        copyDown(pvd, pvd_);
        f.u.resize(3 * size_t(N_));
        for (uint32_t i = 0; i < N_; ++i) {
            f.u[3 * size_t(i)] = float(u4[i].x);
            f.u[3 * size_t(i) + 1] = float(u4[i].y);
            f.u[3 * size_t(i) + 2] = float(u4[i].z);
        }
        f.alive.resize(B_);
        HIP_OK(hipMemcpy(f.alive.data(), alive_.p, B_, hipMemcpyDeviceToHost));
        f.damage.resize(B_);
        for (uint32_t o = 0; o < B_; ++o) f.damage[node_j_host_[o]] = float(pvd[o].w);
        f.sphere = readSphere();
        frames_.push_back(std::move(f));
    }

    // Kernel kinds for the optional per-launch event timing, each with the
    // reference phase id it belongs to (FastLattice.cpp:22-26).
    enum Kernel : uint32_t { KStrainStart, KBondStart, KPredict, KPass1, KSolve, KProject, KVelocity, KDamp,
                             KSupportVelocity, KPass2, KStrain, KBondEnd, KFinish, KRinv, KMark, KStrainLate,
                             KBondEndLate, KCount };
    static const char *kernelName(uint32_t k) {
        static const char *names[KCount] = {"k_strain_start", "k_bond_start", "k_predict", "k_pass1", "k_solve",
                                            "k_project", "k_velocity", "k_damp", "k_support_velocity", "k_pass2",
                                            "k_strain", "k_bond_end", "k_finish", "k_rinv", "k_mark_affected",
                                            "k_strain_late", "k_bond_end_late"};
        return k < KCount ? names[k] : "unknown";
    }
    struct ProfRecord { uint32_t kernel, phase; hipEvent_t a, b; };
    bool prof_ = false;
    std::vector<hipEvent_t> pool_;
    std::vector<ProfRecord> records_;
    double kernel_seconds_[KCount] = {};
    uint64_t kernel_launches_[KCount] = {};

    void profBegin(uint32_t kernel, uint32_t phase, hipStream_t on = nullptr) {
        if (!prof_) return;
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
    void profEnd(hipStream_t on = nullptr) {
        if (!prof_) return;
        HIP_OK(hipEventRecord(records_.back().b, on ? on : stream_));
    }
    void profAddLaunches(uint32_t kernel, uint32_t extra) {
        if (prof_) kernel_launches_[kernel] += extra;
    }
    void profCollect() {
        for (const ProfRecord &r : records_) {
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, r.a, r.b));
            kernel_seconds_[r.kernel] += double(ms) * 1e-3;
            kernel_launches_[r.kernel] += 1;
            run_status_.phase_seconds[r.phase] += double(ms) * 1e-3;
        }
        records_.clear();
    }

public:
    void setProfiling(bool on) override { prof_ = on; }
    uint32_t kernelStats(uint32_t k, std::string &name, double &seconds, uint64_t &launches) const override {
        if (k < KCount) {
            name = kernelName(k);
            seconds = kernel_seconds_[k];
            launches = kernel_launches_[k];
        }
        return KCount;
    }
    uint64_t profilingBatch(uint64_t batch) const { return prof_ ? std::min<uint64_t>(batch, 4) : batch; }

private:
#define BLX_LAUNCH(kernel, phase, ...) \
    do {                                \
        profBegin(kernel, phase);       \
        __VA_ARGS__;                    \
        profEnd();                      \
    } while (0)

    template <bool DIRECT> void enqueueT() {
        const DevLattice<R> L = L_;
        const Step<R> S = S_;
        const uint32_t gn = gridFor(N_);
        const bool sphere = S.sphere_enabled != 0;
        // With a striker, the end sample of what pass 2 cannot reach overlaps pass 2.
        const bool split = sphere && !plastic_;
        const uint32_t gnt = xcdGrid(node_tiles_), gbt = xcdGrid(bond_tiles_);
        BLX_LAUNCH(KStrainStart, 1, hipLaunchKernelGGL((k_strain<R, DIRECT, kStrainDirty>), dim3(gn), dim3(kWG), 0, stream_, L));
        BLX_LAUNCH(KBondStart, 2, hipLaunchKernelGGL((k_bond_start<R, DIRECT>), dim3(gn), dim3(kWG), 0, stream_, L));
        const dim3 pg(gnt);
        profBegin(KPredict, 3);
        if (sphere && loads_set_) hipLaunchKernelGGL((k_predict<R, true, true, DIRECT>), pg, dim3(kWG), 0, stream_, L, S);
        else if (sphere) hipLaunchKernelGGL((k_predict<R, true, false, DIRECT>), pg, dim3(kWG), 0, stream_, L, S);
        else if (loads_set_) hipLaunchKernelGGL((k_predict<R, false, true, DIRECT>), pg, dim3(kWG), 0, stream_, L, S);
        else hipLaunchKernelGGL((k_predict<R, false, false, DIRECT>), pg, dim3(kWG), 0, stream_, L, S);
        profEnd();
        if (sphere) BLX_LAUNCH(KPass1, 4, hipLaunchKernelGGL((k_pass<R, 1>), dim3(1), dim3(kFinishThreads), 0, stream_, L, S));
        if (split) {
            HIP_OK(hipEventRecord(ev_pass1_, stream_));
            HIP_OK(hipStreamWaitEvent(early_, ev_pass1_, 0));
            const uint32_t marks = std::min<uint32_t>(cus_ * 4U, (sch_.blocks * kCap * (D_ + 1) + kWG - 1) / kWG);
            profBegin(KMark, 4, early_);
            hipLaunchKernelGGL((k_mark_affected<R>), dim3(std::max(1U, marks)), dim3(kWG), 0, early_, L);
            profEnd(early_);
        }
        const uint2 *wg = wg_.as<uint2>();
        for (uint32_t it = 0; it < iterations_; ++it) {
            const uint32_t flags = (it == 0 ? kSolveFirst : 0U) | (it + 1 < iterations_ ? kSolveKeepLambda : 0U);
            // One timing bracket over the back-to-back stage launches of the
            // interior phase and one over the boundary phases.
            for (int phase = 0; phase < 2; ++phase) {
                const size_t lo = phase == 0 ? 0 : interior_stages_, hi = phase == 0 ? interior_stages_ : stages_.size();
                if (lo == hi) continue;
                profBegin(KSolve, phase == 0 ? 5 : 6);
                for (size_t k = lo; k < hi; ++k) {
                    const auto &st = stages_[k];
                    if (plastic_)
                        hipLaunchKernelGGL((k_solve<R, DIRECT, true>), dim3(xcdGrid(st.second)), dim3(kWG), 0, stream_, L,
                                           S, wg, st.first, st.second, flags);
                    else
                        hipLaunchKernelGGL((k_solve<R, DIRECT, false>), dim3(xcdGrid(st.second)), dim3(kWG), 0, stream_, L,
                                           S, wg, st.first, st.second, flags);
                }
                profEnd();
                profAddLaunches(KSolve, uint32_t(hi - lo) - 1U);
            }
            if (S.plane_count > 0 && it + 1 < iterations_)
                BLX_LAUNCH(KProject, 7, hipLaunchKernelGGL((k_project<R>), dim3(gn), dim3(kWG), 0, stream_, L, S));
        }
        profBegin(KVelocity, 8);
        if (S.plane_count > 0) hipLaunchKernelGGL((k_velocity<R, true>), dim3(gnt), dim3(kWG), 0, stream_, L, S);
        else hipLaunchKernelGGL((k_velocity<R, false>), dim3(gnt), dim3(kWG), 0, stream_, L, S);
        profEnd();
        if (S.damping > R(0)) {
            for (const auto &st : stages_)
                BLX_LAUNCH(KDamp, 9, hipLaunchKernelGGL((k_damp<R, DIRECT>), dim3(st.second), dim3(kWG), 0, stream_, L, S, wg, st.first));
            BLX_LAUNCH(KSupportVelocity, 9, hipLaunchKernelGGL((k_support_velocity<R>), dim3(gn), dim3(kWG), 0, stream_, L, S));
        }
        if (split) {
            HIP_OK(hipEventRecord(ev_velocity_, stream_));
            HIP_OK(hipStreamWaitEvent(early_, ev_velocity_, 0));
            profBegin(KStrain, 11, early_);
            hipLaunchKernelGGL((k_strain<R, DIRECT, kStrainUnaffected>), dim3(gnt), dim3(kWG), 0, early_, L);
            profEnd(early_);
            profBegin(KBondEnd, 12, early_);
            hipLaunchKernelGGL((k_bond_end<R, DIRECT, false, kBondUnaffected>), dim3(gbt), dim3(kWG), 0, early_, L, S);
            profEnd(early_);
            HIP_OK(hipEventRecord(ev_early_, early_));
            BLX_LAUNCH(KPass2, 10, hipLaunchKernelGGL((k_pass<R, 2>), dim3(1), dim3(kFinishThreads), 0, stream_, L, S));
            const uint32_t late = std::min<uint32_t>(cus_ * 4U, gridFor(N_));
            BLX_LAUNCH(KStrainLate, 11, hipLaunchKernelGGL((k_strain<R, DIRECT, kStrainAffected>), dim3(late), dim3(kWG), 0, stream_, L));
            HIP_OK(hipStreamWaitEvent(stream_, ev_early_, 0));
            BLX_LAUNCH(KBondEndLate, 12, hipLaunchKernelGGL((k_bond_end<R, DIRECT, false, kBondAffected>), dim3(late), dim3(kWG), 0, stream_, L, S));
        } else {
            if (sphere) BLX_LAUNCH(KPass2, 10, hipLaunchKernelGGL((k_pass<R, 2>), dim3(1), dim3(kFinishThreads), 0, stream_, L, S));
            BLX_LAUNCH(KStrain, 11, hipLaunchKernelGGL((k_strain<R, DIRECT, kStrainAll>), dim3(gnt), dim3(kWG), 0, stream_, L));
            profBegin(KBondEnd, 12);
            if (plastic_) hipLaunchKernelGGL((k_bond_end<R, DIRECT, true, kBondAll>), dim3(gbt), dim3(kWG), 0, stream_, L, S);
            else hipLaunchKernelGGL((k_bond_end<R, DIRECT, false, kBondAll>), dim3(gbt), dim3(kWG), 0, stream_, L, S);
            profEnd();
        }
        // The dirty nodes' inverse rest covariances overlap the ordered folds.
        HIP_OK(hipEventRecord(ev_bond_end_, stream_));
        HIP_OK(hipStreamWaitEvent(side_, ev_bond_end_, 0));
        profBegin(KRinv, 1, side_);
        hipLaunchKernelGGL((k_rinv<R>), dim3(gridFor(N_)), dim3(kWG), 0, side_, L);
        profEnd(side_);
        HIP_OK(hipEventRecord(ev_rinv_, side_));
        profBegin(KFinish, 13);
        if (loads_set_ && plastic_) hipLaunchKernelGGL((k_finish<R, true, true>), dim3(1), dim3(kFinishThreads), 0, stream_, L, S);
        else if (loads_set_) hipLaunchKernelGGL((k_finish<R, true, false>), dim3(1), dim3(kFinishThreads), 0, stream_, L, S);
        else if (plastic_) hipLaunchKernelGGL((k_finish<R, false, true>), dim3(1), dim3(kFinishThreads), 0, stream_, L, S);
        else hipLaunchKernelGGL((k_finish<R, false, false>), dim3(1), dim3(kFinishThreads), 0, stream_, L, S);
        profEnd();
        HIP_OK(hipStreamWaitEvent(stream_, ev_rinv_, 0));
        HIP_OK(hipGetLastError());
    }

    void enqueueSubstep() {
        if (direct_) enqueueT<true>();
        else enqueueT<false>();
    }
};

struct Handle {
    std::unique_ptr<BackendBase> impl;
};

} // namespace
} // namespace blx

using namespace blx;

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
            std::snprintf(line, sizeof line, "%sdevice %d: %s (%s), %d CUs, %.1f GiB", d ? "; " : "", d, p.name,
                          p.gcnArchName, p.multiProcessorCount, double(p.totalGlobalMem) / (1024.0 * 1024.0 * 1024.0));
            s += line;
        }
        if (out && cap) {
            std::strncpy(out, s.c_str(), cap - 1);
            out[cap - 1] = 0;
        }
    });
}

uint32_t blx_multiprocessor_count(void) {
    if (deviceCount() == 0) return 0;
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 0;
    return uint32_t(p.multiProcessorCount);
}

int blx_backend_create(int device, int precision, const blx_schedule_view *sv, blx_backend **out) {
    return guard([&] {
        if (!sv || !out) throw Invalid("null argument");
        if (deviceCount() <= device || device < 0)
            throw Runtime("no HIP device " + std::to_string(device) + " for the lattice backend");
        blx::Schedule s;
        s.N = sv->node_count;
        s.B = sv->bond_count;
        s.blocks = sv->block_count;
        s.colors = sv->color_count;
        s.nbb.assign(sv->node_block_begin, sv->node_block_begin + sv->block_count + 1);
        const size_t ranges = size_t(sv->block_count) * 2 * std::max(sv->color_count, 1U);
        s.range_begin.assign(sv->range_begin, sv->range_begin + ranges);
        s.range_end.assign(sv->range_end, sv->range_end + ranges);
        auto h = std::make_unique<Handle>();
        if (precision == BLX_PRECISION_FLOAT) h->impl = std::make_unique<Backend<float>>(device, std::move(s));
        else if (precision == BLX_PRECISION_DOUBLE) h->impl = std::make_unique<Backend<double>>(device, std::move(s));
        else throw Invalid("unknown precision");
        *out = reinterpret_cast<blx_backend *>(h.release());
    });
}

#define IMPL(b) (reinterpret_cast<Handle *>(b)->impl)

int blx_backend_destroy(blx_backend *b) { return guard([&] { delete reinterpret_cast<Handle *>(b); }); }
int blx_backend_name(const blx_backend *b, char *out, size_t cap) {
    return guard([&] {
        const std::string s = reinterpret_cast<const Handle *>(b)->impl->name();
        if (out && cap) {
            std::strncpy(out, s.c_str(), cap - 1);
            out[cap - 1] = 0;
        }
    });
}
int blx_backend_upload(blx_backend *b, const blx_state_view *st, const blx_settings *s, const blx_sphere *sp) {
    return guard([&] {
        if (!b || !st || !s || !sp) throw Invalid("null argument");
        IMPL(b)->upload(*st, *s, *sp);
    });
}
int blx_backend_set_external_forces(blx_backend *b, const double *f, uint64_t n) {
    return guard([&] { IMPL(b)->setForces(f, n); });
}
int blx_backend_run(blx_backend *b, const blx_run_control *c, blx_run_status *st) {
    return guard([&] {
        if (!b || !c || !st) throw Invalid("null argument");
        IMPL(b)->run(*c, *st);
    });
}
int blx_backend_status(const blx_backend *b, blx_run_status *st) {
    return guard([&] { reinterpret_cast<const Handle *>(b)->impl->status(*st); });
}
int blx_backend_download(blx_backend *b, blx_state_view *st, blx_sphere *sp) {
    return guard([&] { IMPL(b)->download(*st, *sp); });
}
int blx_backend_first_failure_bonds(const blx_backend *b, uint32_t *out, uint64_t cap, uint64_t *count) {
    return guard([&] { *count = reinterpret_cast<const Handle *>(b)->impl->firstFailure(out, cap); });
}
int blx_backend_frame_count(const blx_backend *b, uint32_t *n) {
    return guard([&] { *n = reinterpret_cast<const Handle *>(b)->impl->frameCount(); });
}
int blx_backend_take_frame(blx_backend *b, uint32_t k, uint64_t *step, float *u, uint8_t *alive, float *damage,
                           blx_sphere *sp) {
    return guard([&] { IMPL(b)->takeFrame(k, step, u, alive, damage, sp); });
}
int blx_backend_clear_frames(blx_backend *b) { return guard([&] { IMPL(b)->clearFrames(); }); }
int blx_backend_fragment_labels(blx_backend *b, uint32_t *label, uint32_t *count) {
    return guard([&] { *count = IMPL(b)->labels(label); });
}

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

int blx_backend_set_profiling(blx_backend *b, int on) { return guard([&] { IMPL(b)->setProfiling(on != 0); }); }
int blx_backend_kernel_stats(const blx_backend *b, uint32_t index, char *name, size_t cap, double *seconds,
                             uint64_t *launches, uint32_t *count) {
    return guard([&] {
        std::string n;
        double s = 0.0;
        uint64_t l = 0;
        *count = reinterpret_cast<const Handle *>(b)->impl->kernelStats(index, n, s, l);
        if (index < *count) {
            if (name && cap) {
                std::strncpy(name, n.c_str(), cap - 1);
                name[cap - 1] = 0;
            }
            if (seconds) *seconds = s;
            if (launches) *launches = l;
        }
    });
}

} // extern "C"
