// Register-resident tiled passes (the fast path for tiled plans).
//
// A tile of 2^12 amplitudes lives in the registers of 256 threads, 16
// amplitudes each. Which 4 of the 12 tile-local index bits select the
// register ("register bits") and which 8 select the thread is the tile's
// layout. A gate whose non-diagonal support lies in the register bits runs
// entirely inside each thread: no LDS traffic, no barrier. The pass is split
// into phases; between phases the tile is transposed through LDS into the
// next phase's layout (one LDS write + read of the tile, two barriers).
//
// Global load and store use the first and last phases' layouts, padded with
// high tile bits so that lanes walk the contiguous low-qubit runs.
#include "common.hpp"
#include "engine.hpp"
#include "tile.hpp"

#include <cstring>

namespace lqcs_dev {

// Persistent register tiles (default) and their workgroups per CU.
bool g_rtile_persistent = true;
uint32_t g_rtile_blocks_per_cu = 2;

namespace {

constexpr int kRL = 12;   // tile bits
constexpr int kRT = 256;  // threads
constexpr int kRR = 16;   // registers per thread

__device__ __forceinline__ uint64_t pdep64(uint64_t v, uint64_t mask) {
    uint64_t r = 0;
    for (uint64_t bit = 1; mask; bit <<= 1) {
        const uint64_t low = mask & (~mask + 1);
        if (v & bit) r |= low;
        mask ^= low;
    }
    return r;
}

__device__ __forceinline__ uint32_t pdep32(uint32_t v, uint32_t mask) {
    uint32_t r = 0;
    for (uint32_t bit = 1; mask; bit <<= 1) {
        const uint32_t low = mask & (~mask + 1);
        if (v & bit) r |= low;
        mask ^= low;
    }
    return r;
}

// LDS slot of tile index j: low 3 bits XOR-folded with higher bits so that
// lanes differing only in high bits still spread over the LDS banks. The map
// is linear over XOR, so slot(jt | jr) = slot(jt) ^ slot(jr) for disjoint jt,
// jr; the host precomputes slot(jr) per register.
__host__ __device__ __forceinline__ uint32_t swz(uint32_t j) {
    return j ^ (((j >> 3) ^ (j >> 6) ^ (j >> 9)) & 7u);
}

using Regs = amp_t[kRR];

// Select on values already loaded (a select between an array element and a
// computed value would otherwise become a select of addresses and push the
// register arrays to scratch).
__device__ __forceinline__ amp_t pick(bool c, amp_t p, amp_t q) {
    return make_double2(c ? p.x : q.x, c ? p.y : q.y);
}

__device__ __forceinline__ void copy_regs(const Regs& a, Regs& o) {
#pragma unroll
    for (int r = 0; r < kRR; ++r) o[r] = a[r];
}

// Every op reads register set `a` and writes all of set `o` (ping-pong), so
// results land in their final registers without copies.
template <int B>
__device__ __forceinline__ void d1_all(const Regs& a, Regs& o, const RegOp& op) {
    const amp_t m0 = op.m[0], m1 = op.m[1], m2 = op.m[2], m3 = op.m[3];
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        if (r & (1 << B)) continue;
        const amp_t a0 = a[r], a1 = a[r | (1 << B)];
        o[r] = cmac2(m0, a0, m1, a1);
        o[r | (1 << B)] = cmac2(m2, a0, m3, a1);
    }
}

// Controlled variant: pairs whose register bits miss a control, or lanes
// whose thread bits miss one, keep their values (selects, no branches).
template <int B>
__device__ __forceinline__ void d1_masked(const Regs& a, Regs& o, const RegOp& op, bool tp) {
    const amp_t m0 = op.m[0], m1 = op.m[1], m2 = op.m[2], m3 = op.m[3];
    const uint32_t rc = op.rctl;
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        if (r & (1 << B)) continue;
        const bool on = tp && ((uint32_t(r) & rc) == rc);
        const amp_t a0 = a[r], a1 = a[r | (1 << B)];
        const amp_t n0 = cmac2(m0, a0, m1, a1), n1 = cmac2(m2, a0, m3, a1);
        o[r] = pick(on, n0, a0);
        o[r | (1 << B)] = pick(on, n1, a1);
    }
}

template <int B>
__device__ __forceinline__ void x_all(const Regs& a, Regs& o) {
#pragma unroll
    for (int r = 0; r < kRR; ++r) o[r] = a[r ^ (1 << B)];
}

// X on register bit B controlled by register bit C (compile-time pairs).
template <int B, int C>
__device__ __forceinline__ void x_rc(const Regs& a, Regs& o) {
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        if (r & (1 << C)) o[r] = a[r ^ (1 << B)];
        else o[r] = a[r];
    }
}

template <int B>
__device__ __forceinline__ void x_masked(const Regs& a, Regs& o, const RegOp& op, bool tp) {
    const uint32_t rc = op.rctl;
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        const bool on = tp && ((uint32_t(r) & rc) == rc);
        const amp_t p = a[r ^ (1 << B)], q = a[r];
        o[r] = pick(on, p, q);
    }
}

template <int B0, int B1>
__device__ __forceinline__ void d2(const Regs& a, Regs& o, const RegOp& op, bool tp) {
    const uint32_t rc = op.rctl;
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        if (r & ((1 << B0) | (1 << B1))) continue;
        const int idx[4] = {r, r | (1 << B0), r | (1 << B1), r | (1 << B0) | (1 << B1)};
        const bool on = tp && ((uint32_t(r) & rc) == rc);
        amp_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = a[idx[k]];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const amp_t acc = cmac2(op.m[4 * k], x[0], op.m[4 * k + 1], x[1]);
            const amp_t n = cadd(acc, cmac2(op.m[4 * k + 2], x[2], op.m[4 * k + 3], x[3]));
            o[idx[k]] = pick(on, n, x[k]);
        }
    }
}

__device__ __forceinline__ void r_diag(const Regs& a, Regs& o, const RegOp& op, uint32_t jt, const RPhase& ph,
                                       uint64_t base, bool tp) {
    const uint32_t g = ((base & op.dg0) ? 1u : 0u) | ((base & op.dg1) ? 2u : 0u);
    const amp_t d0 = op.m[0], d1 = op.m[1], d2v = op.m[2], d3 = op.m[3];
    const uint32_t rc = op.rctl;
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        const bool on = tp && ((uint32_t(r) & rc) == rc);
        const uint32_t j = jt | ph.jr[r];
        const uint32_t sel = g | ((j & op.dl0) ? 1u : 0u) | ((j & op.dl1) ? 2u : 0u);
        const amp_t d = sel == 0 ? d0 : sel == 1 ? d1 : sel == 2 ? d2v : d3;
        const amp_t q = a[r];
        o[r] = pick(on, cmul(q, d), q);
    }
}

template <int XR>
__device__ __forceinline__ void r_pauli(const Regs& a, Regs& o, const RegOp& op, uint32_t jt, const RPhase& ph,
                                        uint64_t base) {
    constexpr int low = XR & -XR;
    const amp_t c = op.m[0], yp = op.m[1];
    const uint32_t gp = __popcll(base & op.zg) & 1u;
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        if (r & low) continue;
        const int r2 = r ^ XR;
        const uint32_t j1 = jt | ph.jr[r], j2 = jt | ph.jr[r2];
        const double s1 = ((__popc(j1 & op.zl) ^ gp) & 1u) ? -1.0 : 1.0;
        const double s2 = ((__popc(j2 & op.zl) ^ gp) & 1u) ? -1.0 : 1.0;
        const amp_t a1 = a[r], a2 = a[r2];
        o[r] = cmac2(c, a1, cscale(yp, s2), a2);
        o[r2] = cmac2(cscale(yp, s1), a1, c, a2);
    }
}

__device__ __forceinline__ void r_pauli_diag(const Regs& a, Regs& o, const RegOp& op, uint32_t jt,
                                             const RPhase& ph, uint64_t base) {
    const amp_t c = op.m[0], yp = op.m[1];
    const amp_t dp = cadd(c, yp), dm = cadd(c, cscale(yp, -1.0));
    const uint32_t gp = __popcll(base & op.zg) & 1u;
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        const uint32_t j = jt | ph.jr[r];
        o[r] = cmul(a[r], ((__popc(j & op.zl) ^ gp) & 1u) ? dm : dp);
    }
}

__device__ __forceinline__ void apply_reg_op(const Regs& a, Regs& o, const RegOp& op, uint32_t jt,
                                             const RPhase& ph, uint64_t base) {
    if ((base & op.gctrl) != op.gctrl) {  // tile-uniform control not met
        copy_regs(a, o);
        return;
    }
    const bool free = (op.rctl | op.tctl) == 0;
    const bool tp = (jt & op.tctl) == op.tctl;
    switch (op.kind) {
        case RK_DENSE1:
            if (free) {
                switch (op.b0) {
                    case 0: d1_all<0>(a, o, op); return;
                    case 1: d1_all<1>(a, o, op); return;
                    case 2: d1_all<2>(a, o, op); return;
                    default: d1_all<3>(a, o, op); return;
                }
            }
            switch (op.b0) {
                case 0: d1_masked<0>(a, o, op, tp); return;
                case 1: d1_masked<1>(a, o, op, tp); return;
                case 2: d1_masked<2>(a, o, op, tp); return;
                default: d1_masked<3>(a, o, op, tp); return;
            }
        case RK_X:
            if (free) {
                switch (op.b0) {
                    case 0: x_all<0>(a, o); return;
                    case 1: x_all<1>(a, o); return;
                    case 2: x_all<2>(a, o); return;
                    default: x_all<3>(a, o); return;
                }
            }
            if (op.tctl == 0 && __builtin_popcount(op.rctl) == 1) {
                switch (op.b0 * 4 + __builtin_ctz(op.rctl)) {
                    case 1: x_rc<0, 1>(a, o); return;
                    case 2: x_rc<0, 2>(a, o); return;
                    case 3: x_rc<0, 3>(a, o); return;
                    case 4: x_rc<1, 0>(a, o); return;
                    case 6: x_rc<1, 2>(a, o); return;
                    case 7: x_rc<1, 3>(a, o); return;
                    case 8: x_rc<2, 0>(a, o); return;
                    case 9: x_rc<2, 1>(a, o); return;
                    case 11: x_rc<2, 3>(a, o); return;
                    case 12: x_rc<3, 0>(a, o); return;
                    case 13: x_rc<3, 1>(a, o); return;
                    default: x_rc<3, 2>(a, o); return;
                }
            }
            switch (op.b0) {
                case 0: x_masked<0>(a, o, op, tp); return;
                case 1: x_masked<1>(a, o, op, tp); return;
                case 2: x_masked<2>(a, o, op, tp); return;
                default: x_masked<3>(a, o, op, tp); return;
            }
        case RK_DENSE2:
            switch (op.b0 * 4 + op.b1) {
                case 1: d2<0, 1>(a, o, op, tp); return;
                case 2: d2<0, 2>(a, o, op, tp); return;
                case 3: d2<0, 3>(a, o, op, tp); return;
                case 4: d2<1, 0>(a, o, op, tp); return;
                case 6: d2<1, 2>(a, o, op, tp); return;
                case 7: d2<1, 3>(a, o, op, tp); return;
                case 8: d2<2, 0>(a, o, op, tp); return;
                case 9: d2<2, 1>(a, o, op, tp); return;
                case 11: d2<2, 3>(a, o, op, tp); return;
                case 12: d2<3, 0>(a, o, op, tp); return;
                case 13: d2<3, 1>(a, o, op, tp); return;
                default: d2<3, 2>(a, o, op, tp); return;
            }
        case RK_DIAG:
            r_diag(a, o, op, jt, ph, base, tp);
            return;
        case RK_PAULI:
            switch (op.xr) {
                case 0: r_pauli_diag(a, o, op, jt, ph, base); return;
                case 1: r_pauli<1>(a, o, op, jt, ph, base); return;
                case 2: r_pauli<2>(a, o, op, jt, ph, base); return;
                case 3: r_pauli<3>(a, o, op, jt, ph, base); return;
                case 4: r_pauli<4>(a, o, op, jt, ph, base); return;
                case 5: r_pauli<5>(a, o, op, jt, ph, base); return;
                case 6: r_pauli<6>(a, o, op, jt, ph, base); return;
                case 7: r_pauli<7>(a, o, op, jt, ph, base); return;
                case 8: r_pauli<8>(a, o, op, jt, ph, base); return;
                case 9: r_pauli<9>(a, o, op, jt, ph, base); return;
                case 10: r_pauli<10>(a, o, op, jt, ph, base); return;
                case 11: r_pauli<11>(a, o, op, jt, ph, base); return;
                case 12: r_pauli<12>(a, o, op, jt, ph, base); return;
                case 13: r_pauli<13>(a, o, op, jt, ph, base); return;
                case 14: r_pauli<14>(a, o, op, jt, ph, base); return;
                default: r_pauli<15>(a, o, op, jt, ph, base); return;
            }
        default:
            copy_regs(a, o);
            return;
    }
}

__global__ void __launch_bounds__(kRT) k_rtile(amp_t* __restrict__ s, uint32_t nq, uint64_t smask,
                                              const RPhase* __restrict__ phases, uint32_t n_phases,
                                              const RegOp* __restrict__ ops) {
    __shared__ amp_t lds[1 << kRL];
    amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint64_t full = (uint64_t{1} << nq) - 1;
    const uint64_t base = pdep64(blockIdx.x, full & ~smask);
    Regs v, w;
    {
        const RPhase& p0 = phases[0];
        const uint64_t gt = base | pdep64(threadIdx.x, p0.tmask_global);
#pragma unroll
        for (int r = 0; r < kRR; ++r) v[r] = b[gt | p0.gofs[r]];
    }
    uint32_t sjt_prev = 0;
    for (uint32_t p = 0; p < n_phases; ++p) {
        const RPhase& ph = phases[p];
        const uint32_t jt = pdep32(threadIdx.x, ph.tmask);
        const uint32_t sjt = swz(jt);
        if (p > 0) {
            const RPhase& pp = phases[p - 1];
            if (p > 1) __syncthreads();  // every thread has read the previous layout
#pragma unroll
            for (int r = 0; r < kRR; ++r) lds[sjt_prev ^ pp.sjr[r]] = v[r];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kRR; ++r) v[r] = lds[sjt ^ ph.sjr[r]];
        }
        sjt_prev = sjt;
        uint32_t o = ph.first_op;
        const uint32_t end = ph.first_op + ph.n_ops;
        for (; o + 1 < end; o += 2) {
            apply_reg_op(v, w, ops[o], jt, ph, base);
            apply_reg_op(w, v, ops[o + 1], jt, ph, base);
        }
        if (o < end) {
            apply_reg_op(v, w, ops[o], jt, ph, base);
            copy_regs(w, v);
        }
    }
    {
        const RPhase& pl = phases[n_phases - 1];
        const uint64_t gt = base | pdep64(threadIdx.x, pl.tmask_global);
#pragma unroll
        for (int r = 0; r < kRR; ++r) b[gt | pl.gofs[r]] = v[r];
    }
}

// Persistent form: each workgroup walks tiles blockIdx.x, +gridDim.x, ...
// and issues the global loads of its next tile before computing the current
// one, so HBM reads overlap the in-register gate work.
__global__ void __launch_bounds__(kRT) k_rtile_persist(amp_t* __restrict__ s, uint32_t nq, uint64_t smask,
                                                      const RPhase* __restrict__ phases, uint32_t n_phases,
                                                      const RegOp* __restrict__ ops, uint32_t tiles) {
    __shared__ amp_t lds[1 << kRL];
    amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint64_t full = (uint64_t{1} << nq) - 1;
    const uint64_t outside = full & ~smask;
    const RPhase& p0 = phases[0];
    const RPhase& pl = phases[n_phases - 1];
    const uint64_t gt0 = pdep64(threadIdx.x, p0.tmask_global);
    const uint64_t gtl = pdep64(threadIdx.x, pl.tmask_global);
    Regs v, w, nx;
    uint32_t tile = blockIdx.x;
    if (tile >= tiles) return;
    {
        const uint64_t gt = pdep64(tile, outside) | gt0;
#pragma unroll
        for (int r = 0; r < kRR; ++r) v[r] = b[gt | p0.gofs[r]];
    }
    for (; tile < tiles; tile += gridDim.x) {
        const uint64_t base = pdep64(tile, outside);
        const uint32_t next = tile + gridDim.x;
        if (next < tiles) {
            const uint64_t gt = pdep64(next, outside) | gt0;
#pragma unroll
            for (int r = 0; r < kRR; ++r) nx[r] = b[gt | p0.gofs[r]];
        }
        uint32_t sjt_prev = 0;
        for (uint32_t p = 0; p < n_phases; ++p) {
            const RPhase& ph = phases[p];
            const uint32_t jt = pdep32(threadIdx.x, ph.tmask);
            const uint32_t sjt = swz(jt);
            if (p > 0) {
                const RPhase& pp = phases[p - 1];
                __syncthreads();  // every thread is done with the LDS tile
#pragma unroll
                for (int r = 0; r < kRR; ++r) lds[sjt_prev ^ pp.sjr[r]] = v[r];
                __syncthreads();
#pragma unroll
                for (int r = 0; r < kRR; ++r) v[r] = lds[sjt ^ ph.sjr[r]];
            }
            sjt_prev = sjt;
            uint32_t o = ph.first_op;
            const uint32_t end = ph.first_op + ph.n_ops;
            for (; o + 1 < end; o += 2) {
                apply_reg_op(v, w, ops[o], jt, ph, base);
                apply_reg_op(w, v, ops[o + 1], jt, ph, base);
            }
            if (o < end) {
                apply_reg_op(v, w, ops[o], jt, ph, base);
                copy_regs(w, v);
            }
        }
        {
            const uint64_t gt = base | gtl;
#pragma unroll
            for (int r = 0; r < kRR; ++r) b[gt | pl.gofs[r]] = v[r];
        }
        if (next < tiles) copy_regs(nx, v);
    }
}

uint64_t pdep_host(uint64_t v, uint64_t mask) {
    uint64_t r = 0;
    for (uint64_t bit = 1; mask; bit <<= 1) {
        const uint64_t low = mask & (~mask + 1);
        if (v & bit) r |= low;
        mask ^= low;
    }
    return r;
}

int popc(uint64_t m) { return __builtin_popcountll(m); }

bool is_x(const TileOp& t) {
    return t.m[0].x == 0 && t.m[0].y == 0 && t.m[3].x == 0 && t.m[3].y == 0 && t.m[1].x == 1 &&
           t.m[1].y == 0 && t.m[2].x == 1 && t.m[2].y == 0;
}

// Tile-local bits an op reads or writes, and the bits it needs as register bits.
void local_masks(const TileOp& t, uint32_t& touched, uint32_t& dense) {
    switch (t.kind) {
        case TK_DENSE1: dense = 1u << t.t0; touched = dense | t.lctrl; break;
        case TK_DENSE2: dense = (1u << t.t0) | (1u << t.t1); touched = dense | t.lctrl; break;
        case TK_DIAG: dense = 0; touched = t.dl0 | t.dl1 | t.lctrl; break;
        default: dense = t.xl; touched = t.xl | t.zl; break;
    }
}

}  // namespace

bool build_rtile(const std::vector<TileOp>& tops, uint32_t L, uint64_t smask, RTileProgram& out) {
    if (L != static_cast<uint32_t>(kRL)) return false;
    out.phases.clear();
    out.ops.clear();
    const uint32_t all = (1u << kRL) - 1;
    std::vector<uint32_t> touched(tops.size()), dense(tops.size());
    for (size_t i = 0; i < tops.size(); ++i) {
        local_masks(tops[i], touched[i], dense[i]);
        if (popc(dense[i]) > 4) return false;
    }
    std::vector<uint32_t> pending(tops.size());
    for (uint32_t i = 0; i < tops.size(); ++i) pending[i] = i;
    std::vector<std::pair<uint32_t, std::vector<uint32_t>>> groups;  // (register bits, ops)
    while (!pending.empty()) {
        uint32_t C = 0, blocked = 0;
        std::vector<uint32_t> taken, rest;
        for (uint32_t i : pending) {
            if ((touched[i] & blocked) || popc(C | dense[i]) > 4) {
                blocked |= touched[i];
                rest.push_back(i);
                continue;
            }
            C |= dense[i];
            taken.push_back(i);
        }
        groups.emplace_back(C, std::move(taken));
        pending.swap(rest);
    }
    if (groups.empty()) groups.emplace_back(0u, std::vector<uint32_t>{});
    for (auto& g : groups) {  // pad with the highest free tile bits
        for (int p = kRL - 1; p >= 0 && popc(g.first) < 4; --p) g.first |= (1u << p) & ~g.first & all;
    }
    for (auto& g : groups) {
        RPhase ph;
        std::memset(&ph, 0, sizeof ph);
        const uint32_t rmask = g.first;
        ph.tmask = all & ~rmask;
        ph.tmask_global = pdep_host(ph.tmask, smask);
        int rb[4], k = 0;
        for (int p = 0; p < kRL; ++p)
            if (rmask & (1u << p)) rb[k++] = p;
        for (int r = 0; r < kRR; ++r) {
            uint32_t j = 0;
            for (int i = 0; i < 4; ++i)
                if (r & (1 << i)) j |= 1u << rb[i];
            ph.jr[r] = j;
            ph.sjr[r] = swz(j);
            ph.gofs[r] = pdep_host(j, smask);
        }
        auto ridx = [&](uint32_t pos) {
            for (int i = 0; i < 4; ++i)
                if (rb[i] == static_cast<int>(pos)) return static_cast<uint32_t>(i);
            return 0u;
        };
        auto rsplit = [&](uint32_t lmask, uint32_t& rctl, uint32_t& tctl) {
            rctl = 0;
            for (int i = 0; i < 4; ++i)
                if (lmask & (1u << rb[i])) rctl |= 1u << i;
            tctl = lmask & ph.tmask;
        };
        ph.first_op = static_cast<uint32_t>(out.ops.size());
        for (uint32_t i : g.second) {
            const TileOp& t = tops[i];
            RegOp o;
            std::memset(&o, 0, sizeof o);
            o.gctrl = t.gctrl;
            rsplit(t.lctrl, o.rctl, o.tctl);
            switch (t.kind) {
                case TK_DENSE1:
                    o.kind = is_x(t) ? RK_X : RK_DENSE1;
                    o.b0 = ridx(t.t0);
                    for (int e = 0; e < 4; ++e) o.m[e] = t.m[e];
                    break;
                case TK_DENSE2:
                    o.kind = RK_DENSE2;
                    o.b0 = ridx(t.t0);
                    o.b1 = ridx(t.t1);
                    for (int e = 0; e < 16; ++e) o.m[e] = t.m[e];
                    break;
                case TK_DIAG:
                    o.kind = RK_DIAG;
                    o.dl0 = t.dl0;
                    o.dl1 = t.dl1;
                    o.dg0 = t.dg0;
                    o.dg1 = t.dg1;
                    for (int e = 0; e < 4; ++e) o.m[e] = t.m[e];
                    break;
                default: {
                    o.kind = RK_PAULI;
                    uint32_t xr = 0;
                    for (int bit = 0; bit < kRL; ++bit)
                        if (t.xl & (1u << bit)) xr |= 1u << ridx(bit);
                    o.xr = xr;
                    o.zl = t.zl;
                    o.zg = t.zg;
                    o.m[0] = t.m[0];
                    o.m[1] = t.m[1];
                    break;
                }
            }
            out.ops.push_back(o);
        }
        ph.n_ops = static_cast<uint32_t>(out.ops.size()) - ph.first_op;
        out.phases.push_back(ph);
    }
    return true;
}

void launch_rtile(amp_t* s, uint32_t nq, uint64_t smask, const RPhase* d_phases, uint32_t n_phases,
                  const RegOp* d_ops, uint32_t batch) {
    const uint64_t tiles = uint64_t{1} << (nq - kRL);
    const uint64_t resident = uint64_t(num_cus()) * g_rtile_blocks_per_cu;
    if (g_rtile_persistent && tiles > resident) {
        k_rtile_persist<<<dim3(static_cast<unsigned>(resident), batch), kRT, 0, stream()>>>(
            s, nq, smask, d_phases, n_phases, d_ops, static_cast<uint32_t>(tiles));
        return;
    }
    k_rtile<<<dim3(static_cast<unsigned>(tiles), batch), kRT, 0, stream()>>>(s, nq, smask, d_phases, n_phases,
                                                                            d_ops);
}

}  // namespace lqcs_dev
