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

#include "lattice_device.hpp"

namespace blx {

constexpr int kWG = 256;
constexpr u32 kCap = 1024;        // striker candidates per block (LatticePhysics.hpp:105)
constexpr u32 kStage = 512;       // candidates staged in LDS at a time
constexpr int kFinishThreads = 1024;

// Device-resident run state. Everything that changes from one substep to the
// next lives here, so kernel arguments are constant and a batch of substeps
// can be enqueued (or captured) without the host in the loop.
struct DevCtl {
    u64 step;            // RunStatus::total_steps
    u32 stop;            // nonzero: the remaining substeps of the batch are no-ops
    u32 error;           // bit 0: external kick overflowed backend precision
    u32 dirty_start;     // recompute the start sample before the next substep
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
    double plastic_work;
    u64 first_failure_count;
    // RunControl of the current run() call.
    u64 quiet, min_steps, no_failure, energy_flat, calm;
    double flat_fraction, calm_margin, removable;
};

template <class R> struct DevLattice {
    u32 N, B, D, blocks, node_tiles, bond_tiles, bitmap_words;
    R4<R> *u4;        // displacement, w = inv_mass
    R4<R> *up4;       // displacement at substep start
    R4<R> *v4;        // velocity
    const R4<R> *x04; // reference position, w = mass
    R *rinv;          // 9 planes of N: cached inverse rest covariance
    R *unm;           // 3 planes of N: unmeasured normal
    R4<R> *sa, *sb;   // strain (xx, yy, zz, xy), (xz, yz, valid, 0)
    u8 *valid, *dirty;
    R *approach;      // 2 planes of N
    u8 *engaged;      // 2 planes of N
    u8 *cand;         // N
    const u32 *nbr_other; // D planes of N (k-major), kNone where no bond
    const R *nbr_rx, *nbr_ry, *nbr_rz, *nbr_w;
    u8 *nbr_alive;
    const BondRec<R> *rec;
    u8 *alive;
    R *lambda, *pext, *pstr;
    const Thresholds<R> *thr;
    R4<R> *pvd;       // prev tensile, compressive, shear; damage
    u8 *mode;
    const uint2 *slots;
    u32 *stamp;       // substep (+1, low 32 bits) a bond broke in
    R *removed;       // stored energy of a bond at its break
    u32 *tile_cand;   // candidates per node tile
    u64 *tile_bits;   // bond tiles with events this substep
    double *tile_plastic; // plastic work per bond tile
    double *load_part;    // ledger partial sums per node tile (13 each)
    R4<R> *fext;      // applied external force in backend precision, w = 1 where loaded
    const double *freq; // requested external force (3N double)
    u32 *cand_all;    // ordered candidates, all blocks
    u32 *cand_lo, *cand_cnt;
    const u32 *nbb;   // node block begin
    u32 *first_fail;  // schedule indices broken in the first failure round
    DevCtl *ctl;
    Sphere<R> *sphere;
    double origin[3];
};

// ---------------------------------------------------------------------------
// Nodes.

template <class R, bool DIRECT> __device__ __forceinline__ V3<R> nodePos(const DevLattice<R> &L, u32 i, const R4<R> &u) {
    if (DIRECT) return xyz(L.x04[i]) + xyz(u);
    return xyz(u);
}

// LatticePhysics.hpp:792-907 (plate bending is refused at upload).
// Dead and absent neighbour slots contribute an exact zero to every sum in the
// reference, so they are skipped; the result is the same bits.
template <class R, bool DIRECT> __device__ void strainNode(const DevLattice<R> &L, u32 i) {
    const u32 N = L.N, D = L.D;
    if (L.dirty[i]) {
        R cov[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        u32 live = 0;
        for (u32 k = 0; k < D; ++k) {
            const size_t slot = size_t(k) * N + i;
            if (L.nbr_other[slot] == kNone) break;
            if (!L.nbr_alive[slot]) continue;
            const R w = L.nbr_w[slot];
            if (w <= R(0)) continue;
            const R rr[3] = {L.nbr_rx[slot], L.nbr_ry[slot], L.nbr_rz[slot]};
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) cov[r * 3 + q] += w * rr[r] * rr[q];
            ++live;
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
            if (rank < 3) atomicAdd(&L.ctl->rank_deficient, 1U);
        }
        L.valid[i] = valid;
        L.dirty[i] = 0;
        if (!valid) {
            R4<R> sb = L.sb[i];
            sb.z = R(0);
            L.sb[i] = sb;
        }
    }
    if (!L.valid[i]) return;
    R acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const V3<R> xi = nodePos<R, DIRECT>(L, i, L.u4[i]);
    for (u32 k = 0; k < D; ++k) {
        const size_t slot = size_t(k) * N + i;
        const u32 other = L.nbr_other[slot];
        if (other == kNone) break;
        if (!L.nbr_alive[slot]) continue;
        const R w = L.nbr_w[slot];
        const V3<R> rest{L.nbr_rx[slot], L.nbr_ry[slot], L.nbr_rz[slot]};
        const V3<R> xo = nodePos<R, DIRECT>(L, other, L.u4[other]);
        const V3<R> cur = DIRECT ? (xo - xi) - rest : xo - xi;
        const R cc[3] = {cur.x, cur.y, cur.z}, rr[3] = {rest.x, rest.y, rest.z};
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) acc[r * 3 + q] += w * cc[r] * rr[q];
    }
    R inv[9], un[3], e[6];
    for (int k = 0; k < 9; ++k) inv[k] = L.rinv[size_t(k) * N + i];
    for (int k = 0; k < 3; ++k) un[k] = L.unm[size_t(k) * N + i];
    greenStrain(acc, inv, un, e);
    L.sa[i] = R4<R>{e[0], e[1], e[2], e[3]};
    L.sb[i] = R4<R>{e[4], e[5], R(1), R(0)};
}

template <class R, bool DIRECT, bool GATED>
__global__ __launch_bounds__(kWG) void k_strain(DevLattice<R> L) {
    if (L.ctl->stop) return;
    if (GATED && !L.ctl->dirty_start) return;
    for (u32 i = blockIdx.x * kWG + threadIdx.x; i < L.N; i += gridDim.x * kWG) strainNode<R, DIRECT>(L, i);
}

// ---------------------------------------------------------------------------
// Bonds: the strain sample (LatticePhysics.hpp:924-959).

template <class R, bool DIRECT>
__device__ __forceinline__ V3<R> bondVec(const DevLattice<R> &L, const BondRec<R> &r, const R4<R> &ua, const R4<R> &ub) {
    if (DIRECT) return (xyz(L.x04[r.b]) + xyz(ub)) - (xyz(L.x04[r.a]) + xyz(ua));
    return V3<R>{r.ex, r.ey, r.ez} + (xyz(ub) - xyz(ua));
}

template <class R, bool DIRECT>
__device__ void sampleBond(const DevLattice<R> &L, const BondRec<R> &r, const R4<R> &ua, const R4<R> &ub, R &t, R &c,
                           R &s) {
    const V3<R> delta = bondVec<R, DIRECT>(L, r, ua, ub);
    const R len = length(delta);
    const R rest = r.L;
    const V3<R> e{r.ex, r.ey, r.ez};
    R stretch;
    if (DIRECT) {
        stretch = len / rest - R(1);
    } else {
        const V3<R> du = xyz(ub) - xyz(ua);
        const R l2 = r.rsm + R(2) * dot(e, du) + dot(du, du);
        stretch = l2 / ((len + rest) * rest);
    }
    t = stretch;
    c = -stretch;
    s = R(0);
    const R rl = length(e);
    if (rl > R(1.0e-9)) {
        const V3<R> dir = e * (R(1) / rl);
        const u32 ends[2] = {r.a, r.b};
        for (int q = 0; q < 2; ++q) {
            const R4<R> sb = L.sb[ends[q]];
            if (sb.z == R(0)) continue;
            const R4<R> sa = L.sa[ends[q]];
            R nn, sh;
            resolve(sa, sb, dir, nn, sh);
            t = maxR(t, nn);
            c = maxR(c, -nn);
            s = maxR(s, sh);
        }
    }
}

template <class R, bool DIRECT> __global__ __launch_bounds__(kWG) void k_bond_start(DevLattice<R> L) {
    if (L.ctl->stop || !L.ctl->dirty_start) return;
    for (u32 j = blockIdx.x * kWG + threadIdx.x; j < L.B; j += gridDim.x * kWG) {
        if (!L.alive[j]) continue;
        const BondRec<R> r = L.rec[j];
        R t, c, s;
        sampleBond<R, DIRECT>(L, r, L.u4[r.a], L.u4[r.b], t, c, s);
        R4<R> p = L.pvd[j];
        p.x = maxR(R(0), t);
        p.y = maxR(R(0), c);
        p.z = maxR(R(0), s);
        L.pvd[j] = p;
    }
}

// ---------------------------------------------------------------------------
// Kick, predict, classify.

template <class R> struct Ledger13 { double v[13]; };

template <class R> __device__ void blockSum13(double (*sh)[kWG], double v[13]) {
    for (int k = 0; k < 13; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = kWG / 2; w > 0; w >>= 1) {
        if (threadIdx.x < unsigned(w))
            for (int k = 0; k < 13; ++k) sh[k][threadIdx.x] = sh[k][threadIdx.x] + sh[k][threadIdx.x + w];
        __syncthreads();
    }
}

template <class R, bool SPHERE, bool LOADS, bool DIRECT>
__global__ __launch_bounds__(kWG) void k_predict(DevLattice<R> L, Step<R> S) {
    DevCtl *ctl = L.ctl;
    if (ctl->stop) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->dirty_start = 0;
    const bool loads = LOADS && ctl->load_remaining > 0;
    const u32 i = blockIdx.x * kWG + threadIdx.x;
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
        const int n = __syncthreads_count(cand);
        if (threadIdx.x == 0) L.tile_cand[blockIdx.x] = u32(n);
    }
    if (LOADS && loads) {
        __shared__ double sh[13][kWG];
        blockSum13<R>(sh, led);
        if (threadIdx.x == 0)
            for (int k = 0; k < 13; ++k) L.load_part[size_t(blockIdx.x) * 13 + k] = sh[k][0];
    }
}

// ---------------------------------------------------------------------------
// Striker contact (LatticePhysics.hpp:1324-1387), one node, serial caller.

template <class R> struct ContactAcc {
    u64 contacts;
    double imp[3], dissipated, max_pen, max_corr;
};

template <class R>
__device__ void contactNode(const Step<R> &S, Sphere<R> &sp, V3<R> x0, V3<R> &u, V3<R> &v, R mass, bool correct,
                            ContactAcc<R> &acc) {
    if (mass <= R(0)) return;
    const R dt = S.dt;
    const R ism = R(1) / sp.mass;
    const R iin = R(1) / sp.inertia;
    const R cr = sp.radius + S.c_node_radius;
    const V3<R> r = (x0 + u) - sp.center;
    const R dist = length(r);
    const V3<R> n = normalized(r);
    const R gap = dist - cr;
    V3<R> nv = v;
    const V3<R> rel = nv - sp.velocity - cross(sp.omega, r);
    const R vn = dot(rel, n);
    if (gap > maxR(S.c_margin, -vn * dt)) return;
    if (-gap > R(acc.max_pen)) acc.max_pen = double(-gap);
    const R inm = R(1) / mass;
    const R rest = -vn > S.c_rest_thr ? S.c_rest : R(0);
    const R dvn = gap > S.c_margin ? -gap / dt : -rest * minR(vn, R(0));
    const R ni = maxR(R(0), (dvn - vn) / (inm + ism + iin * length2(cross(r, n))));
    if (ni > R(0)) {
        const R before = R(0.5) * mass * length2(nv);
        V3<R> imp = ni * n;
        nv = nv + inm * imp;
        sp.velocity = sp.velocity - ism * imp;
        sp.omega = sp.omega + iin * (-cross(r, imp));
        acc.imp[0] += double(imp.x);
        acc.imp[1] += double(imp.y);
        acc.imp[2] += double(imp.z);
        const V3<R> anr = nv - sp.velocity - cross(sp.omega, r);
        const V3<R> tan = anr - dot(anr, n) * n;
        const R ts = length(tan);
        if (ts > R(1.0e-12)) {
            const V3<R> dir = tan / ts;
            const R stick = ts / (inm + ism + iin * length2(cross(r, dir)));
            const R fi = stick <= S.c_sf * ni ? stick : minR(stick, S.c_df * ni);
            imp = -fi * dir;
            nv = nv + inm * imp;
            sp.velocity = sp.velocity - ism * imp;
            sp.omega = sp.omega + iin * (-cross(r, imp));
            acc.imp[0] += double(imp.x);
            acc.imp[1] += double(imp.y);
            acc.imp[2] += double(imp.z);
        }
        v = nv;
        acc.dissipated += double(before - R(0.5) * mass * length2(nv));
        ++acc.contacts;
    }
    if (correct && gap < -S.c_margin) {
        const R corr = minR(-gap - S.c_margin, R(0.2) * maxR(S.c_node_radius, sp.radius * R(0.01)));
        const R sum = inm + ism;
        const V3<R> ns = (corr * inm / sum) * n;
        const V3<R> ss = -((corr * ism / sum) * n);
        u = u + ns;
        sp.center = sp.center + ss;
        if (corr > R(acc.max_corr)) acc.max_corr = double(corr);
    }
}

// The serial striker pass of one substep (pass 1 or 2) on one workgroup:
// pass 1 first builds the ordered per-block candidate lists from the tile
// counts k_predict left (CpuLatticeBackend.cpp:375-386, cap 1024 per block),
// then walks them; every node's data is staged into LDS by the whole
// workgroup and one thread runs the order-dependent chain.
template <class R, int PASS>
__global__ __launch_bounds__(kFinishThreads) void k_pass(DevLattice<R> L, Step<R> S) {
    DevCtl *ctl = L.ctl;
    if (ctl->stop) return;
    __shared__ R4<R> s_x0[kStage], s_u[kStage], s_v[kStage];
    __shared__ u32 s_scan[kFinishThreads];
    __shared__ Sphere<R> s_sp;
    __shared__ u32 s_total;
    const u32 tid = threadIdx.x;
    if (tid == 0) {
        Sphere<R> sp = *L.sphere;
        if (PASS == 1) sp.velocity = sp.velocity + S.dt * S.gravity;
        s_sp = sp;
    }
    if (PASS == 1) {
        // Ordered compaction of the candidate flags, tile by tile.
        u32 running = 0;
        for (u32 base = 0; base < L.node_tiles; base += kFinishThreads) {
            const u32 t = base + tid;
            const u32 cnt = t < L.node_tiles ? L.tile_cand[t] : 0U;
            s_scan[tid] = cnt;
            __syncthreads();
            for (u32 w = 1; w < kFinishThreads; w <<= 1) {
                const u32 add = tid >= w ? s_scan[tid - w] : 0U;
                __syncthreads();
                s_scan[tid] += add;
                __syncthreads();
            }
            const u32 total = s_scan[kFinishThreads - 1];
            if (cnt > 0) {
                u32 pos = running + s_scan[tid] - cnt;
                const u32 end = min(L.N, (t + 1) * kWG);
                for (u32 i = t * kWG; i < end; ++i)
                    if (L.cand[i]) L.cand_all[pos++] = i;
            }
            running += total;
            __syncthreads();
        }
        if (tid == 0) s_total = running;
        __syncthreads();
        __threadfence_block();
        for (u32 b = tid; b < L.blocks; b += kFinishThreads) {
            const u32 K = s_total;
            auto lower = [&](u32 key) {
                u32 lo = 0, hi = K;
                while (lo < hi) {
                    const u32 mid = (lo + hi) >> 1;
                    if (L.cand_all[mid] < key) lo = mid + 1;
                    else hi = mid;
                }
                return lo;
            };
            const u32 lo = lower(L.nbb[b]), hi = lower(L.nbb[b + 1]);
            const u32 count = hi - lo;
            L.cand_lo[b] = lo;
            L.cand_cnt[b] = count < kCap ? count : kCap;
            if (count > kCap) atomicAdd((unsigned long long *)&ctl->candidate_overflow, (unsigned long long)(count - kCap));
        }
        __threadfence_block();
    }
    __syncthreads();
    ContactAcc<R> acc;
    Sphere<R> sp;
    V3<R> center_before;
    R energy_before;
    if (tid == 0) {
        acc.contacts = ctl->impulse_contacts;
        for (int k = 0; k < 3; ++k) acc.imp[k] = ctl->impulse[k];
        acc.dissipated = ctl->dissipated;
        acc.max_pen = ctl->max_penetration;
        acc.max_corr = ctl->max_correction;
        sp = s_sp;
        center_before = sp.center;
        energy_before = R(0.5) * sp.mass * length2(sp.velocity) + R(0.5) * sp.inertia * length2(sp.omega);
    }
    for (u32 b = 0; b < L.blocks; ++b) {
        const u32 lo = L.cand_lo[b], cnt = L.cand_cnt[b];
        for (u32 c0 = 0; c0 < cnt; c0 += kStage) {
            const u32 n = min(kStage, cnt - c0);
            for (u32 k = tid; k < n; k += kFinishThreads) {
                const u32 i = L.cand_all[lo + c0 + k];
                s_x0[k] = L.x04[i];
                s_u[k] = L.u4[i];
                s_v[k] = L.v4[i];
            }
            __syncthreads();
            if (tid == 0) {
                for (u32 k = 0; k < n; ++k) {
                    V3<R> u = xyz(s_u[k]), v = xyz(s_v[k]);
                    contactNode<R>(S, sp, xyz(s_x0[k]), u, v, s_x0[k].w, PASS == 2, acc);
                    s_u[k] = with(u, s_u[k].w);
                    s_v[k] = with(v, R(0));
                }
            }
            __syncthreads();
            for (u32 k = tid; k < n; k += kFinishThreads) {
                const u32 i = L.cand_all[lo + c0 + k];
                V3<R> v = xyz(s_v[k]);
                R4<R> u = s_u[k];
                if (PASS == 1) {
                    // nodeRecordApproach and the prediction (LatticePhysics.hpp:1408-1411).
                    for (u32 p = 0; p < S.plane_count; ++p) L.approach[size_t(p) * L.N + i] = dot(v, S.planes[p].normal);
                    u = with(xyz(u) + S.dt * v, u.w);
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

template <class R, bool DIRECT, bool PLASTIC>
__global__ __launch_bounds__(kWG) void k_solve(DevLattice<R> L, Step<R> S, const uint2 *wg_range, u32 wg_base, int first) {
    if (L.ctl->stop) return;
    const uint2 rg = wg_range[wg_base + blockIdx.x];
    const u32 j = rg.x + threadIdx.x;
    if (j >= rg.y) return;
    if (!L.alive[j]) return;
    const BondRec<R> r = L.rec[j];
    const R4<R> ua = L.u4[r.a], ub = L.u4[r.b];
    const V3<R> du = xyz(ub) - xyz(ua);
    const V3<R> e{r.ex, r.ey, r.ez};
    const V3<R> delta = DIRECT ? (xyz(L.x04[r.b]) + xyz(ub)) - (xyz(L.x04[r.a]) + xyz(ua)) : e + du;
    const R len = length(delta);
    if (len <= R(1.0e-12)) return;
    const R orig = r.L;
    R C;
    if (PLASTIC) {
        const R plastic = L.pext[j];
        const R rest = orig + plastic;
        C = DIRECT ? len - rest : (r.rsm + R(2) * dot(e, du) + dot(du, du) - plastic * (rest + orig)) / (len + rest);
    } else {
        // plastic == 0: rest == orig and the subtracted term is an exact +0.
        C = DIRECT ? len - orig : (r.rsm + R(2) * dot(e, du) + dot(du, du)) / (len + orig);
    }
    const V3<R> dir = delta / len;
    const R wa = ua.w, wb = ub.w;
    const R alpha = r.c / (S.dt * S.dt);
    const R lam = first ? R(0) : L.lambda[j];
    const R dl = (-C - alpha * lam) / (wa + wb + alpha);
    L.lambda[j] = lam + dl;
    L.u4[r.a] = with(xyz(ua) - wa * dl * dir, ua.w);
    L.u4[r.b] = with(xyz(ub) + wb * dl * dir, ub.w);
}

// Radial pair damping, one colour stage (LatticePhysics.hpp:1171-1184).
template <class R, bool DIRECT>
__global__ __launch_bounds__(kWG) void k_damp(DevLattice<R> L, Step<R> S, const uint2 *wg_range, u32 wg_base) {
    if (L.ctl->stop) return;
    const uint2 rg = wg_range[wg_base + blockIdx.x];
    const u32 j = rg.x + threadIdx.x;
    if (j >= rg.y) return;
    if (!L.alive[j]) return;
    const BondRec<R> r = L.rec[j];
    const R ma = L.x04[r.a].w, mb = L.x04[r.b].w;
    if (ma <= R(0) || mb <= R(0)) return;
    const V3<R> n = normalized(bondVec<R, DIRECT>(L, r, L.u4[r.a], L.u4[r.b]));
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
}

template <class R> __global__ __launch_bounds__(kWG) void k_project(DevLattice<R> L, Step<R> S) {
    if (L.ctl->stop) return;
    for (u32 i = blockIdx.x * kWG + threadIdx.x; i < L.N; i += gridDim.x * kWG) {
        R4<R> u = L.u4[i];
        bool moved = false;
        projectNode(L, S, i, u, moved);
        if (moved) L.u4[i] = u;
    }
}

// Last iteration's support projection fused with nodeVelocityUpdate
// (LatticePhysics.hpp:1726-1730): both only touch node i.
template <class R, bool PROJECT>
__global__ __launch_bounds__(kWG) void k_velocity(DevLattice<R> L, Step<R> S) {
    if (L.ctl->stop) return;
    for (u32 i = blockIdx.x * kWG + threadIdx.x; i < L.N; i += gridDim.x * kWG) {
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
template <class R> __global__ __launch_bounds__(kWG) void k_support_velocity(DevLattice<R> L, Step<R> S) {
    if (L.ctl->stop) return;
    for (u32 i = blockIdx.x * kWG + threadIdx.x; i < L.N; i += gridDim.x * kWG) {
        if (L.cand[i]) continue;
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
// One bond tile per workgroup iteration; a tile with a break, or with a damage
// that could raise the running maximum, sets its bit for the ordered folds.

template <class R, bool DIRECT, bool PLASTIC>
__global__ __launch_bounds__(kWG) void k_bond_end(DevLattice<R> L, Step<R> S) {
    DevCtl *ctl = L.ctl;
    if (ctl->stop) return;
    __shared__ float s_max[4][kWG / 64];
    __shared__ double s_pw[kWG];
    const double m0 = ctl->max_damage;
    const u32 stamp = u32(ctl->step) + 1U;
    const u32 lane = threadIdx.x & 63U, wave = threadIdx.x >> 6;
    float mt = 0.f, mc = 0.f, ms = 0.f, mp = 0.f;
    for (u32 tile = blockIdx.x; tile < L.bond_tiles; tile += gridDim.x) {
        const u32 j = tile * kWG + threadIdx.x;
        bool event = false;
        double pw = 0.0;
        if (j < L.B && L.alive[j]) {
            const BondRec<R> r = L.rec[j];
            const R4<R> ua = L.u4[r.a], ub = L.u4[r.b];
            R plastic = R(0);
            if (PLASTIC) {
                // bondPlasticReturn (LatticePhysics.hpp:1018-1039).
                plastic = L.pext[j];
                R ps = L.pstr[j];
                const R elastic = length(bondVec<R, DIRECT>(L, r, ua, ub)) - r.L - plastic;
                const R mag = absR(elastic);
                const R h = S.hardening;
                const R ye = S.yield * r.L + h * ps;
                if (mag > ye) {
                    const R inc = (mag - ye) / (R(1) + h);
                    plastic = plastic + (elastic < R(0) ? -inc : inc);
                    ps = ps + inc;
                    L.pext[j] = plastic;
                    L.pstr[j] = ps;
                    if (r.c > R(0)) {
                        const R yend = ye + h * inc;
                        pw = double(R(0.5) * (ye + yend) * inc / r.c);
                    }
                }
                mp = fmaxf(mp, float(absR(plastic) / r.L));
            }
            R t, c, s;
            sampleBond<R, DIRECT>(L, r, ua, ub, t, c, s);
            R4<R> pv = L.pvd[j];
            const R pt = maxR(R(0), maxR(pv.x, t)), pc = maxR(R(0), maxR(pv.y, c)), pss = maxR(R(0), maxR(pv.z, s));
            mt = fmaxf(mt, float(pt));
            mc = fmaxf(mc, float(pc));
            ms = fmaxf(ms, float(pss));
            pv.x = maxR(R(0), t);
            pv.y = maxR(R(0), c);
            pv.z = maxR(R(0), s);
            const Thresholds<R> th = L.thr[j];
            const R td = progress(pt, th.ds, th.de), cd = progress(pc, th.cs, th.ce), sd = progress(pss, th.ss, th.se);
            R dmg = td;
            u8 md = 1;
            if (cd > dmg) { dmg = cd; md = 2; }
            if (sd > dmg) { dmg = sd; md = 3; }
            if (dmg > pv.w) {
                pv.w = dmg;
                L.mode[j] = md;
            }
            L.pvd[j] = pv;
            event = double(float(pv.w)) > m0 + 1.0e-6;
            if (pv.w >= R(1)) {
                const R ext = length(bondVec<R, DIRECT>(L, r, ua, ub)) - r.L - plastic;
                L.removed[j] = r.c <= R(0) ? R(0) : R(0.5) * ext * ext / r.c;
                L.alive[j] = 0;
                const uint2 sl = L.slots[j];
                L.nbr_alive[sl.x] = 0;
                L.nbr_alive[sl.y] = 0;
                L.dirty[r.a] = 1;
                L.dirty[r.b] = 1;
                L.stamp[j] = stamp;
                event = true;
            }
        }
        if (__syncthreads_or(event) && threadIdx.x == 0)
            atomicOr((unsigned long long *)&L.tile_bits[tile >> 6], 1ULL << (tile & 63U));
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
    // Running maxima of non-negative floats: integer atomicMax on the bits.
    for (int o = 32; o > 0; o >>= 1) {
        mt = fmaxf(mt, __shfl_xor(mt, o));
        mc = fmaxf(mc, __shfl_xor(mc, o));
        ms = fmaxf(ms, __shfl_xor(ms, o));
        mp = fmaxf(mp, __shfl_xor(mp, o));
    }
    if (lane == 0) {
        s_max[0][wave] = mt;
        s_max[1][wave] = mc;
        s_max[2][wave] = ms;
        s_max[3][wave] = mp;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        float m = 0.f;
        for (int w = 0; w < kWG / 64; ++w) m = fmaxf(m, s_max[threadIdx.x][w]);
        if (m > 0.f) {
            u32 *dst = threadIdx.x < 3 ? &ctl->peak_bits[threadIdx.x] : &ctl->pstretch_bits;
            atomicMax(dst, __float_as_uint(m));
        }
    }
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

template <class R, bool LOADS, bool PLASTIC>
__global__ __launch_bounds__(kFinishThreads) void k_finish(DevLattice<R> L, Step<R> S) {
    DevCtl *ctl = L.ctl;
    if (ctl->stop) return;
    __shared__ u64 s_words[kFinishThreads];
    __shared__ double s_red[kFinishThreads];
    const u32 tid = threadIdx.x, lane = tid & 63U;
    const u64 step = ctl->step;
    const u32 stamp = u32(step) + 1U;
    // Wave 0 carries the folds; every lane holds the same running values.
    double m = ctl->max_damage, removed = ctl->removed_energy_j;
    bool damage_gain = false, energy_gain = false;
    u32 broken_now = 0;
    const bool first_round = ctl->broken_bonds == 0;
    u64 ff = ctl->first_failure_count;
    const double flat = ctl->flat_fraction;
    for (u32 base = 0; base < L.bitmap_words; base += kFinishThreads) {
        const u32 w = base + tid;
        u64 bits = 0;
        if (w < L.bitmap_words) {
            bits = L.tile_bits[w];
            if (bits) L.tile_bits[w] = 0;
        }
        s_words[tid] = bits;
        __syncthreads();
        if (tid < 64) {
            const u32 nw = min(u32(kFinishThreads), L.bitmap_words - base);
            for (u32 k = 0; k < nw; ++k) {
                u64 word = s_words[k];
                while (word) {
                    const u32 tile = (base + k) * 64U + u32(__ffsll((long long)word) - 1);
                    word &= word - 1;
                    for (u32 sub = 0; sub < kWG; sub += 64) {
                        const u32 j = tile * kWG + sub + lane;
                        bool eval = false, broke = false;
                        float dmg = 0.f;
                        double e = 0.0;
                        if (j < L.B) {
                            broke = L.stamp[j] == stamp && !L.alive[j];
                            eval = L.alive[j] || broke;
                            dmg = float(L.pvd[j].w);
                            if (broke) e = double(L.removed[j]);
                        }
                        // max_damage rises only by more than 1e-6 (CpuLatticeBackend.cpp:462-465).
                        u64 cand = __ballot(eval && double(dmg) > m + 1.0e-6);
                        while (cand) {
                            const int l = __ffsll((long long)cand) - 1;
                            m = double(__shfl(dmg, l));
                            damage_gain = true;
                            cand = __ballot(int(lane) > l && eval && double(dmg) > m + 1.0e-6);
                        }
                        // Removed energy summed in bond order (CpuLatticeBackend.cpp:466-477).
                        u64 bmask = __ballot(broke);
                        while (bmask) {
                            const int l = __ffsll((long long)bmask) - 1;
                            bmask &= bmask - 1;
                            const double el = __shfl(e, l);
                            removed += el;
                            if (el > flat * (removed > 1.0e-12 ? removed : 1.0e-12)) energy_gain = true;
                            if (first_round && lane == 0) L.first_fail[ff] = tile * kWG + sub + u32(l);
                            ++ff;
                            ++broken_now;
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    double ledger[13];
    const bool loads = LOADS && ctl->load_remaining > 0;
    if (LOADS && loads)
        for (int k = 0; k < 13; ++k) ledger[k] = fixedOrderSum<R>(L.load_part, L.node_tiles, 13, k, s_red);
    double plastic = 0.0;
    if (PLASTIC) plastic = fixedOrderSum<R>(L.tile_plastic, L.bond_tiles, 1, 0, s_red);
    if (tid != 0) return;
    if (damage_gain) {
        ctl->max_damage = m;
        ctl->last_damage_gain_step = step;
    }
    const bool any = broken_now > 0;
    if (any) {
        ctl->removed_energy_j = removed;
        if (energy_gain) ctl->last_energy_gain_step = step;
        ctl->broken_bonds += broken_now;
        if (first_round) ctl->first_failure_count = ff;
    }
    if (PLASTIC) ctl->plastic_work += plastic;
    if (LOADS && loads) {
        ctl->ledger_steps += 1;
        ctl->ledger_elapsed += double(S.dt);
        for (int k = 0; k < 3; ++k) {
            ctl->ledger_req[k] += ledger[k];
            ctl->ledger_imp[k] += ledger[3 + k];
            ctl->ledger_areq[k] += ledger[6 + k];
            ctl->ledger_aimp[k] += ledger[9 + k];
        }
        ctl->ledger_work += ledger[12];
        ctl->load_remaining -= 1;
    }
    // Sphere drift and sphere against the supports (CpuLatticeBackend.cpp:481-482).
    Sphere<R> sp = *L.sphere;
    if (!S.sphere_enabled) sp.velocity = sp.velocity + S.dt * S.gravity;
    sp.center = sp.center + S.dt * sp.velocity;
    u64 ev = ctl->ball_support_events;
    sphereSupport(S, sp, ev);
    ctl->ball_support_events = ev;
    *L.sphere = sp;
    if (any) ctl->dirty_start = 1;
    // Run loop bookkeeping (CpuLatticeBackend.cpp:279-298).
    const u64 total = step + 1;
    ctl->step = total;
    if (any) {
        if (ctl->first_failure_step == ~0ULL) ctl->first_failure_step = step;
        ctl->last_failure_step = step;
        ctl->failure_rounds += 1;
    }
    const double available = ctl->removable + ctl->ledger_work;
    const unsigned reason = ctl->removable > 0 && available <= 0
        ? 6U
        : exitReason(total, ctl->broken_bonds, ctl->last_failure_step, ctl->quiet, ctl->min_steps, ctl->no_failure,
                     ctl->energy_flat, ctl->last_energy_gain_step, ctl->calm, ctl->last_damage_gain_step, ctl->max_damage,
                     ctl->calm_margin, ctl->removed_energy_j, ctl->removable > 0 ? available : 0.0);
    ctl->exit_reason = reason;
    if (reason != 0 || ctl->error) ctl->stop = 1;
}

// ---------------------------------------------------------------------------
// Fragments: integer union-find over the live bonds, then the canonical
// relabel of findConnectedComponents (ConnectedComponents.cpp:77-85). Roots are
// linked larger index under smaller with a compare-and-swap, so every root
// ends as its component's smallest node: the tie-break of the canonical order.
// Parent reads are agent-scope atomic loads, never a stale per-CU cache line.

__device__ __forceinline__ u32 loadParent(u32 *parent, u32 x) {
    return __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ u32 findRoot(u32 *parent, u32 x) {
    u32 p = loadParent(parent, x);
    while (p != x) {
        const u32 g = loadParent(parent, p);
        // Path halving: g is an ancestor of x, so the store only shortens the path.
        if (g != p) __hip_atomic_store(&parent[x], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
        p = g;
    }
    return x;
}

__global__ void k_cc_init(u32 *parent, u32 *size, u32 n) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        parent[i] = i;
        size[i] = 0;
    }
}

template <class R> __global__ void k_cc_hook(DevLattice<R> L, u32 *parent) {
    for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < L.B; j += gridDim.x * blockDim.x) {
        if (!L.alive[j]) continue;
        const BondRec<R> r = L.rec[j];
        u32 a = r.a, b = r.b;
        while (true) {
            a = findRoot(parent, a);
            b = findRoot(parent, b);
            if (a == b) break;
            const u32 hi = a > b ? a : b, lo = a > b ? b : a;
            if (atomicCAS(&parent[hi], hi, lo) == hi) break;
            a = hi;
            b = lo;
        }
    }
}

// Roots go to their own array: a root stored into parent[] could be overwritten
// by another thread's path halving of the same node.
__global__ void k_cc_roots(u32 *parent, u32 *root, u32 *size, u32 n) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const u32 r = findRoot(parent, i);
        root[i] = r;
        atomicAdd(&size[r], 1U);
    }
}

// One sort key per component: (n - size) << 32 | root. Ascending keys are the
// canonical order, size descending then smallest node; the append order of
// the keys does not matter because the sort is on the whole key.
__global__ void k_cc_keys(const u32 *root, const u32 *size, u64 *keys, u32 *count, u32 n) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (root[i] == i) keys[atomicAdd(count, 1U)] = (u64(n - size[i]) << 32) | u64(i);
}

__global__ void k_cc_rank(const u64 *sorted, u32 *rank, u32 c) {
    for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < c; k += gridDim.x * blockDim.x)
        rank[u32(sorted[k] & 0xffffffffULL)] = k;
}

__global__ void k_cc_label(const u32 *root, const u32 *rank, u32 *label, u32 n) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) label[i] = rank[root[i]];
}

// A plain streaming copy, 16 bytes per lane per iteration: the known-size
// workload rocprofv3's FETCH_SIZE and WRITE_SIZE counters are scaled against.
__global__ __launch_bounds__(kWG) void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    for (size_t i = size_t(blockIdx.x) * kWG + threadIdx.x; i < n; i += size_t(gridDim.x) * kWG) dst[i] = src[i];
}

} // namespace blx
