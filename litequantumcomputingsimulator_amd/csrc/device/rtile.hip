// Register-resident tiled passes (the fast path for tiled plans).
//
// A tile of 2^12 amplitudes lives in the registers of 256 threads, 16
// amplitudes each. Which 4 of the 12 tile-local index bits select the
// register ("register bits") and which 8 select the thread is the tile's
// layout. A gate whose non-diagonal support lies in the register bits runs
// entirely inside each thread: no LDS traffic, no barrier. The pass is split
// into phases; between phases the tile is transposed through LDS into the
// next phase's layout, half a tile at a time (32 KiB of LDS), so that three
// workgroups fit on a CU.
//
// Global load and store use the first and last phases' layouts. Their
// register bits avoid the tile's contiguous low qubits, so that the lanes of
// a wave walk contiguous runs of amplitudes: a layout whose lanes step over
// high qubits touches a separate cache line (and page) per lane and runs the
// pass several times slower.
//
// Op records (a 32-byte header plus matrix entries) are read from global
// memory with scalar loads: every lane uses the same address, so they land
// in scalar registers, where the FMAs take the matrix entries as operands.
#include "common.hpp"
#include "engine.hpp"
#include "tile.hpp"

#include <cstring>

namespace lqcs_dev {

// Streaming register tiles (next-tile prefetch) and their workgroups per CU.
bool g_rtile_stream = false;
// Landing-buffer register tiles (LDS DMA prefetch of the next tile).
bool g_rtile_land = false;

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

using Regs = amp_t[kRR];

// An op's uniform fields, decoded from its header.
struct Hdr {
    uint32_t kind, b0, b1, xr, rctl, tctl, la, lb, gctrl, ga, gb, member;
    bool restricted;  // a tile-number control or a batch-register restriction to check
    bool free;        // no controls among the register or thread bits
};

__device__ __forceinline__ uint32_t rfl(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

__device__ __forceinline__ double rfl(double v) {
    const uint64_t u = static_cast<uint64_t>(__double_as_longlong(v));
    const uint64_t lo = rfl(static_cast<uint32_t>(u)), hi = rfl(static_cast<uint32_t>(u >> 32));
    return __longlong_as_double(static_cast<long long>(lo | (hi << 32)));
}

// A matrix entry of the op words (same address in every lane): a scalar
// load, kept in scalar registers (readfirstlane folds away on them).
__device__ __forceinline__ amp_t uni(const amp_t* p) {
    const amp_t v = *p;
    return make_double2(rfl(v.x), rfl(v.y));
}

// Select on values already loaded (a select between an array element and a
// computed value would otherwise become a select of addresses and push the
// register arrays to scratch).
__device__ __forceinline__ amp_t pick(bool c, amp_t p, amp_t q) {
    return make_double2(c ? p.x : q.x, c ? p.y : q.y);
}

// Register sets in fixed physical registers: set A in v[2:65], set B in
// v[66:129]. Pinning an op's output set at the end of the op makes every case
// of the op dispatch produce its results in the same registers, so the joins
// of the dispatch need no copies (register allocation otherwise tends to
// leave results where one case put them and copy whole sets at the joins).
__device__ __forceinline__ void pin_a(Regs& x) {
    asm volatile("" : "+{v[2:3]}"(x[0].x), "+{v[4:5]}"(x[0].y), "+{v[6:7]}"(x[1].x), "+{v[8:9]}"(x[1].y), "+{v[10:11]}"(x[2].x), "+{v[12:13]}"(x[2].y), "+{v[14:15]}"(x[3].x), "+{v[16:17]}"(x[3].y));
    asm volatile("" : "+{v[18:19]}"(x[4].x), "+{v[20:21]}"(x[4].y), "+{v[22:23]}"(x[5].x), "+{v[24:25]}"(x[5].y), "+{v[26:27]}"(x[6].x), "+{v[28:29]}"(x[6].y), "+{v[30:31]}"(x[7].x), "+{v[32:33]}"(x[7].y));
    asm volatile("" : "+{v[34:35]}"(x[8].x), "+{v[36:37]}"(x[8].y), "+{v[38:39]}"(x[9].x), "+{v[40:41]}"(x[9].y), "+{v[42:43]}"(x[10].x), "+{v[44:45]}"(x[10].y), "+{v[46:47]}"(x[11].x), "+{v[48:49]}"(x[11].y));
    asm volatile("" : "+{v[50:51]}"(x[12].x), "+{v[52:53]}"(x[12].y), "+{v[54:55]}"(x[13].x), "+{v[56:57]}"(x[13].y), "+{v[58:59]}"(x[14].x), "+{v[60:61]}"(x[14].y), "+{v[62:63]}"(x[15].x), "+{v[64:65]}"(x[15].y));
}

__device__ __forceinline__ void pin_b(Regs& x) {
    asm volatile("" : "+{v[66:67]}"(x[0].x), "+{v[68:69]}"(x[0].y), "+{v[70:71]}"(x[1].x), "+{v[72:73]}"(x[1].y), "+{v[74:75]}"(x[2].x), "+{v[76:77]}"(x[2].y), "+{v[78:79]}"(x[3].x), "+{v[80:81]}"(x[3].y));
    asm volatile("" : "+{v[82:83]}"(x[4].x), "+{v[84:85]}"(x[4].y), "+{v[86:87]}"(x[5].x), "+{v[88:89]}"(x[5].y), "+{v[90:91]}"(x[6].x), "+{v[92:93]}"(x[6].y), "+{v[94:95]}"(x[7].x), "+{v[96:97]}"(x[7].y));
    asm volatile("" : "+{v[98:99]}"(x[8].x), "+{v[100:101]}"(x[8].y), "+{v[102:103]}"(x[9].x), "+{v[104:105]}"(x[9].y), "+{v[106:107]}"(x[10].x), "+{v[108:109]}"(x[10].y), "+{v[110:111]}"(x[11].x), "+{v[112:113]}"(x[11].y));
    asm volatile("" : "+{v[114:115]}"(x[12].x), "+{v[116:117]}"(x[12].y), "+{v[118:119]}"(x[13].x), "+{v[120:121]}"(x[13].y), "+{v[122:123]}"(x[14].x), "+{v[124:125]}"(x[14].y), "+{v[126:127]}"(x[15].x), "+{v[128:129]}"(x[15].y));
}

template <int OUT>
__device__ __forceinline__ void pin_out(Regs& x) {
    if constexpr (OUT == 0) pin_a(x);
    else pin_b(x);
}

__device__ __forceinline__ void copy_regs(const Regs& a, Regs& o) {
#pragma unroll
    for (int r = 0; r < kRR; ++r) o[r] = a[r];
}

// Every op reads register set `a` and writes all of set `o` (ping-pong), so
// results land in their final registers without copies. `m` points at the
// op's matrix entries in LDS.
template <int B>
__device__ __forceinline__ void d1_vals(const Regs& a, Regs& o, amp_t m0, amp_t m1, amp_t m2, amp_t m3) {
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        if (r & (1 << B)) continue;
        const amp_t a0 = a[r], a1 = a[r | (1 << B)];
        o[r] = cmac2(m0, a0, m1, a1);
        o[r | (1 << B)] = cmac2(m2, a0, m3, a1);
    }
}

template <int B>
__device__ __forceinline__ void d1_all(const Regs& a, Regs& o, const amp_t* m) {
    d1_vals<B>(a, o, uni(m), uni(m + 1), uni(m + 2), uni(m + 3));
}

// 2x2 op on register bit B with the matrix chosen by register bit S: the
// first four entries where bit S is 0, the next four where it is 1.
template <int B, int S>
__device__ __forceinline__ void d1_sel(const Regs& a, Regs& o, const amp_t* m) {
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        if (r & (1 << B)) continue;
        const amp_t* mr = (r & (1 << S)) ? m + 4 : m;
        const amp_t a0 = a[r], a1 = a[r | (1 << B)];
        o[r] = cmac2(uni(mr), a0, uni(mr + 1), a1);
        o[r | (1 << B)] = cmac2(uni(mr + 2), a0, uni(mr + 3), a1);
    }
}

// The same with the matrix chosen per lane (select bit among the thread's
// tile bits): the entries are selected once into vector registers.
template <int B>
__device__ __forceinline__ void d1_lanesel(const Regs& a, Regs& o, const amp_t* m, bool s) {
    d1_vals<B>(a, o, pick(s, uni(m + 4), uni(m)), pick(s, uni(m + 5), uni(m + 1)), pick(s, uni(m + 6), uni(m + 2)),
               pick(s, uni(m + 7), uni(m + 3)));
}

// Controlled variant: pairs whose register bits miss a control, or lanes
// whose thread bits miss one, keep their values (selects, no branches).
template <int B>
__device__ __forceinline__ void d1_masked(const Regs& a, Regs& o, const amp_t* m, uint32_t rc, bool tp) {
    const amp_t m0 = uni(m), m1 = uni(m + 1), m2 = uni(m + 2), m3 = uni(m + 3);
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
__device__ __forceinline__ void x_masked(const Regs& a, Regs& o, uint32_t rc, bool tp) {
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        const bool on = tp && ((uint32_t(r) & rc) == rc);
        const amp_t p = a[r ^ (1 << B)], q = a[r];
        o[r] = pick(on, p, q);
    }
}

// Row k of a 4x4 op times x: one chain of multiply-adds per component.
__device__ __forceinline__ amp_t row4(const amp_t* m, const amp_t (&x)[4]) {
    const amp_t m0 = uni(m), m1 = uni(m + 1), m2 = uni(m + 2), m3 = uni(m + 3);
    return make_double2(m0.x * x[0].x - m0.y * x[0].y + m1.x * x[1].x - m1.y * x[1].y + m2.x * x[2].x -
                            m2.y * x[2].y + m3.x * x[3].x - m3.y * x[3].y,
                        m0.x * x[0].y + m0.y * x[0].x + m1.x * x[1].y + m1.y * x[1].x + m2.x * x[2].y +
                            m2.y * x[2].x + m3.x * x[3].y + m3.y * x[3].x);
}

// Row k of a real 4x4 op times x (entries' imaginary parts zero): half the
// multiply-adds of a complex row.
__device__ __forceinline__ amp_t row4r(const amp_t* m, const amp_t (&x)[4]) {
    const double m0 = uni(m).x, m1 = uni(m + 1).x, m2 = uni(m + 2).x, m3 = uni(m + 3).x;
    return make_double2(m0 * x[0].x + m1 * x[1].x + m2 * x[2].x + m3 * x[3].x,
                        m0 * x[0].y + m1 * x[1].y + m2 * x[2].y + m3 * x[3].y);
}

// 4x4 op on register bits B0 (local bit 0) and B1. `masked` keeps the
// amplitudes of groups whose register bits miss a control, or of lanes whose
// thread bits miss one; without controls there is nothing to select.
template <int B0, int B1, bool MASKED, bool REAL = false>
__device__ __forceinline__ void d2(const Regs& a, Regs& o, const amp_t* m, uint32_t rc, bool tp) {
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        if (r & ((1 << B0) | (1 << B1))) continue;
        const int idx[4] = {r, r | (1 << B0), r | (1 << B1), r | (1 << B0) | (1 << B1)};
        amp_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = a[idx[k]];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const amp_t n = REAL ? row4r(m + 4 * k, x) : row4(m + 4 * k, x);
            if constexpr (MASKED) {
                const bool on = tp && ((uint32_t(r) & rc) == rc);
                o[idx[k]] = pick(on, n, x[k]);
            } else {
                o[idx[k]] = n;
            }
        }
    }
}

__device__ __forceinline__ void r_diag(const Regs& a, Regs& o, const Hdr& h, const amp_t* m, uint32_t jt,
                                       const RPhase& ph, uint32_t tile, bool tp) {
    const uint32_t g = ((tile & h.ga) ? 1u : 0u) | ((tile & h.gb) ? 2u : 0u);
    const amp_t d0 = uni(m), d1 = uni(m + 1), d2v = uni(m + 2), d3 = uni(m + 3);
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        const bool on = tp && ((uint32_t(r) & h.rctl) == h.rctl);
        const uint32_t j = jt | ph.jr[r];
        const uint32_t sel = g | ((j & h.la) ? 1u : 0u) | ((j & h.lb) ? 2u : 0u);
        const amp_t d = sel == 0 ? d0 : sel == 1 ? d1 : sel == 2 ? d2v : d3;
        const amp_t q = a[r];
        o[r] = pick(on, cmul(q, d), q);
    }
}

// diag(d0, d1) on register bit B. ONE0: d0 is exactly 1, so the registers
// whose bit B is 0 pass through. MASKED: controls select per lane and
// register which amplitudes change.
template <int B, bool ONE0, bool MASKED>
__device__ __forceinline__ void dg1_reg(const Regs& a, Regs& o, const amp_t* m, uint32_t rc, bool tp) {
    const amp_t d0 = uni(m), d1 = uni(m + 1);
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        const bool hi = ((r >> B) & 1) != 0;
        if (ONE0 && !hi) {
            o[r] = a[r];
            continue;
        }
        const amp_t n = cmul(a[r], hi ? d1 : d0);
        if constexpr (MASKED) {
            const bool on = tp && ((uint32_t(r) & rc) == rc);
            o[r] = pick(on, n, a[r]);
        } else {
            o[r] = n;
        }
    }
}

// diag(d0, d1) on a qubit that is the same for all of a lane's registers (a
// thread bit, or a bit of the tile number): one factor per lane.
template <bool MASKED>
__device__ __forceinline__ void dg1_uniform(const Regs& a, Regs& o, const amp_t* m, bool s, uint32_t rc, bool tp) {
    const amp_t d = pick(s, uni(m + 1), uni(m));
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        const amp_t n = cmul(a[r], d);
        if constexpr (MASKED) {
            const bool on = tp && ((uint32_t(r) & rc) == rc);
            o[r] = pick(on, n, a[r]);
        } else {
            o[r] = n;
        }
    }
}

template <int XR>
__device__ __forceinline__ void r_pauli(const Regs& a, Regs& o, const Hdr& h, const amp_t* m, uint32_t jt,
                                        const RPhase& ph, uint32_t tile) {
    constexpr int low = XR & -XR;
    const amp_t c = uni(m), yp = uni(m + 1);
    const uint32_t gp = __popc(tile & h.ga) & 1u;
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        if (r & low) continue;
        const int r2 = r ^ XR;
        const uint32_t j1 = jt | ph.jr[r], j2 = jt | ph.jr[r2];
        const double s1 = ((__popc(j1 & h.la) ^ gp) & 1u) ? -1.0 : 1.0;
        const double s2 = ((__popc(j2 & h.la) ^ gp) & 1u) ? -1.0 : 1.0;
        const amp_t a1 = a[r], a2 = a[r2];
        o[r] = cmac2(c, a1, cscale(yp, s2), a2);
        o[r2] = cmac2(cscale(yp, s1), a1, c, a2);
    }
}

__device__ __forceinline__ void r_pauli_diag(const Regs& a, Regs& o, const Hdr& h, const amp_t* m, uint32_t jt,
                                             const RPhase& ph, uint32_t tile) {
    const amp_t c = uni(m), yp = uni(m + 1);
    const amp_t dp = cadd(c, yp), dm = cadd(c, cscale(yp, -1.0));
    const uint32_t gp = __popc(tile & h.ga) & 1u;
#pragma unroll
    for (int r = 0; r < kRR; ++r) {
        const uint32_t j = jt | ph.jr[r];
        o[r] = cmul(a[r], ((__popc(j & h.la) ^ gp) & 1u) ? dm : dp);
    }
}

template <int OUT>
__device__ __forceinline__ void apply_reg_op(const Regs& a, Regs& o, const Hdr& h, const amp_t* m, uint32_t jt,
                                             const RPhase& ph, uint32_t tile) {
    if (h.restricted &&
        ((tile & h.gctrl) != h.gctrl ||  // tile-uniform control not met
         (h.member && ((h.member & 0x7fffffffu) - 1 == blockIdx.y) != (h.member >> 31 != 0)))) {  // batch register
        copy_regs(a, o); pin_out<OUT>(o);
        return;
    }
    const bool free = h.free;
    const bool tp = (jt & h.tctl) == h.tctl;
    // A select op whose select bit is outside the tile is a plain 2x2 op with
    // one of its two matrices, the same for the whole tile.
    uint32_t kind = h.kind;
    const amp_t* mp = m;
    if (kind == RK_SEL1 && h.xr == 0 && h.la == 0) {
        kind = RK_DENSE1;
        if (tile & h.ga) mp = m + 4;
    }
    // The uncontrolled 2x2 op first, ahead of the other kinds: it is most ops
    // of most circuits, and a shorter dispatch for it is measurable.
    if (kind == RK_DENSE1 && free) {
        switch (h.b0) {
            case 0: d1_all<0>(a, o, mp); pin_out<OUT>(o); return;
            case 1: d1_all<1>(a, o, mp); pin_out<OUT>(o); return;
            case 2: d1_all<2>(a, o, mp); pin_out<OUT>(o); return;
            default: d1_all<3>(a, o, mp); pin_out<OUT>(o); return;
        }
    }
    // Then select ops (a 2x2 op merged with a singly-controlled X or diagonal).
    if (kind == RK_SEL1) {
        if (h.xr) {
            switch (h.b0 * 4 + h.xr - 1) {
                case 1: d1_sel<0, 1>(a, o, m); pin_out<OUT>(o); return;
                case 2: d1_sel<0, 2>(a, o, m); pin_out<OUT>(o); return;
                case 3: d1_sel<0, 3>(a, o, m); pin_out<OUT>(o); return;
                case 4: d1_sel<1, 0>(a, o, m); pin_out<OUT>(o); return;
                case 6: d1_sel<1, 2>(a, o, m); pin_out<OUT>(o); return;
                case 7: d1_sel<1, 3>(a, o, m); pin_out<OUT>(o); return;
                case 8: d1_sel<2, 0>(a, o, m); pin_out<OUT>(o); return;
                case 9: d1_sel<2, 1>(a, o, m); pin_out<OUT>(o); return;
                case 11: d1_sel<2, 3>(a, o, m); pin_out<OUT>(o); return;
                case 12: d1_sel<3, 0>(a, o, m); pin_out<OUT>(o); return;
                case 13: d1_sel<3, 1>(a, o, m); pin_out<OUT>(o); return;
                default: d1_sel<3, 2>(a, o, m); pin_out<OUT>(o); return;
            }
        }
        {
            const bool s = (jt & h.la) != 0;
            switch (h.b0) {
                case 0: d1_lanesel<0>(a, o, m, s); pin_out<OUT>(o); return;
                case 1: d1_lanesel<1>(a, o, m, s); pin_out<OUT>(o); return;
                case 2: d1_lanesel<2>(a, o, m, s); pin_out<OUT>(o); return;
                default: d1_lanesel<3>(a, o, m, s); pin_out<OUT>(o); return;
            }
        }
    }
    switch (kind) {
        case RK_DENSE1:
            switch (h.b0) {
                case 0: d1_masked<0>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 1: d1_masked<1>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 2: d1_masked<2>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                default: d1_masked<3>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
            }
        case RK_X:
            if (free) {
                switch (h.b0) {
                    case 0: x_all<0>(a, o); pin_out<OUT>(o); return;
                    case 1: x_all<1>(a, o); pin_out<OUT>(o); return;
                    case 2: x_all<2>(a, o); pin_out<OUT>(o); return;
                    default: x_all<3>(a, o); pin_out<OUT>(o); return;
                }
            }
            if (h.tctl == 0 && __builtin_popcount(h.rctl) == 1) {
                switch (h.b0 * 4 + __builtin_ctz(h.rctl)) {
                    case 1: x_rc<0, 1>(a, o); pin_out<OUT>(o); return;
                    case 2: x_rc<0, 2>(a, o); pin_out<OUT>(o); return;
                    case 3: x_rc<0, 3>(a, o); pin_out<OUT>(o); return;
                    case 4: x_rc<1, 0>(a, o); pin_out<OUT>(o); return;
                    case 6: x_rc<1, 2>(a, o); pin_out<OUT>(o); return;
                    case 7: x_rc<1, 3>(a, o); pin_out<OUT>(o); return;
                    case 8: x_rc<2, 0>(a, o); pin_out<OUT>(o); return;
                    case 9: x_rc<2, 1>(a, o); pin_out<OUT>(o); return;
                    case 11: x_rc<2, 3>(a, o); pin_out<OUT>(o); return;
                    case 12: x_rc<3, 0>(a, o); pin_out<OUT>(o); return;
                    case 13: x_rc<3, 1>(a, o); pin_out<OUT>(o); return;
                    default: x_rc<3, 2>(a, o); pin_out<OUT>(o); return;
                }
            }
            switch (h.b0) {
                case 0: x_masked<0>(a, o, h.rctl, tp); pin_out<OUT>(o); return;
                case 1: x_masked<1>(a, o, h.rctl, tp); pin_out<OUT>(o); return;
                case 2: x_masked<2>(a, o, h.rctl, tp); pin_out<OUT>(o); return;
                default: x_masked<3>(a, o, h.rctl, tp); pin_out<OUT>(o); return;
            }
        case RK_DENSE2R:  // real entries, no controls
            switch (h.b0 * 4 + h.b1) {
                case 1: d2<0, 1, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 2: d2<0, 2, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 3: d2<0, 3, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 4: d2<1, 0, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 6: d2<1, 2, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 7: d2<1, 3, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 8: d2<2, 0, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 9: d2<2, 1, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 11: d2<2, 3, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 12: d2<3, 0, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 13: d2<3, 1, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                default: d2<3, 2, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
            }
        case RK_DENSE2:
            if (free) {
                switch (h.b0 * 4 + h.b1) {
                    case 1: d2<0, 1, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 2: d2<0, 2, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 3: d2<0, 3, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 4: d2<1, 0, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 6: d2<1, 2, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 7: d2<1, 3, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 8: d2<2, 0, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 9: d2<2, 1, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 11: d2<2, 3, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 12: d2<3, 0, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 13: d2<3, 1, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    default: d2<3, 2, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                }
            }
            switch (h.b0 * 4 + h.b1) {
                case 1: d2<0, 1, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 2: d2<0, 2, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 3: d2<0, 3, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 4: d2<1, 0, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 6: d2<1, 2, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 7: d2<1, 3, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 8: d2<2, 0, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 9: d2<2, 1, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 11: d2<2, 3, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 12: d2<3, 0, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                case 13: d2<3, 1, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                default: d2<3, 2, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
            }
        case RK_DIAG:
            r_diag(a, o, h, m, jt, ph, tile, tp); pin_out<OUT>(o);
            return;
        case RK_DIAG1:
            if (h.xr) {
                const bool one0 = h.b1 != 0;
                switch ((h.xr - 1) | (one0 ? 4u : 0u) | (free ? 0u : 8u)) {
                    case 0: dg1_reg<0, false, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 1: dg1_reg<1, false, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 2: dg1_reg<2, false, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 3: dg1_reg<3, false, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 4: dg1_reg<0, true, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 5: dg1_reg<1, true, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 6: dg1_reg<2, true, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 7: dg1_reg<3, true, false>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 8: dg1_reg<0, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 9: dg1_reg<1, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 10: dg1_reg<2, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 11: dg1_reg<3, false, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 12: dg1_reg<0, true, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 13: dg1_reg<1, true, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    case 14: dg1_reg<2, true, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                    default: dg1_reg<3, true, true>(a, o, m, h.rctl, tp); pin_out<OUT>(o); return;
                }
            }
            {
                const bool s = h.la ? (jt & h.la) != 0 : (tile & h.ga) != 0;
                if (free) {
                    dg1_uniform<false>(a, o, m, s, h.rctl, tp);
                } else {
                    dg1_uniform<true>(a, o, m, s, h.rctl, tp);
                }
                pin_out<OUT>(o);
                return;
            }
        case RK_PAULI:
            switch (h.xr) {
                case 0: r_pauli_diag(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 1: r_pauli<1>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 2: r_pauli<2>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 3: r_pauli<3>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 4: r_pauli<4>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 5: r_pauli<5>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 6: r_pauli<6>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 7: r_pauli<7>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 8: r_pauli<8>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 9: r_pauli<9>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 10: r_pauli<10>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 11: r_pauli<11>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 12: r_pauli<12>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 13: r_pauli<13>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                case 14: r_pauli<14>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
                default: r_pauli<15>(a, o, h, m, jt, ph, tile); pin_out<OUT>(o); return;
            }
        default:
            copy_regs(a, o); pin_out<OUT>(o);
            return;
    }
}

// Applies the op at word `off` of the pass's words; returns the next op's word.
template <int OUT>
__device__ __forceinline__ uint32_t apply_op(const Regs& a, Regs& o, const amp_t* words, uint32_t off,
                                                 uint32_t jt, const RPhase& ph, uint32_t tile) {
    const uint4* hp = reinterpret_cast<const uint4*>(words + off);
    const uint4 h0 = hp[0], h1 = hp[1];
    const uint32_t code = rfl(h0.x);
    Hdr h;
    h.kind = code & 15u;
    h.b0 = (code >> 4) & 3u;
    h.b1 = (code >> 6) & 3u;
    h.xr = (code >> 8) & 15u;
    h.rctl = (code >> 12) & 15u;
    h.restricted = (code & kCodeRestricted) != 0;
    h.free = (code & kCodeFree) != 0;
    h.tctl = rfl(h0.y);
    h.la = rfl(h0.z);
    h.lb = rfl(h0.w);
    h.gctrl = rfl(h1.x);
    h.ga = rfl(h1.y);
    h.gb = rfl(h1.z);
    h.member = rfl(h1.w);
    apply_reg_op<OUT>(a, o, h, words + off + 2, jt, ph, tile);
    return off + 2 + ((code >> 16) & 0xFFFu);
}

// s_waitcnt vmcnt(0) (gfx9 encoding: expcnt and lgkmcnt fields at their
// maxima, i.e. not waited for).
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0xF70); }

// Runs every phase of the pass on the tile held in v (phase 0's layout on
// entry, the last phase's on exit). Transposes go through LDS half a tile at
// a time (32 KiB for four-wave tiles, 8 KiB for wave tiles).
__device__ __forceinline__ void run_phases(Regs& v, Regs& w, amp_t* lds, const RPhase* __restrict__ phases,
                                           uint32_t n_phases, const amp_t* __restrict__ words, uint32_t tile,
                                           bool& lds_used) {
    uint32_t jt_prev = 0;
    for (uint32_t p = 0; p < n_phases; ++p) {
        const RPhase& ph = phases[p];
        const uint32_t jt = rtile_thread_bits(threadIdx.x, ph);
        if (p > 0 && ph.transpose) {
            // Half a tile per round: the amplitudes whose bit h is `half`.
            // Bit h is a thread bit of both layouts, so a thread writes all
            // or none of its registers in a round, and reads likewise; reads
            // land in w so that nothing still to be written is overwritten.
            // Addresses are byte offsets: the thread's part XOR the register's.
            const uint32_t h = ph.h;
            const uint32_t hp = (jt_prev >> h) & 1u, hc = (jt >> h) & 1u;
            const uint32_t sp = rtile_slot(jt_prev, h) << 4, sc = rtile_slot(jt, h) << 4;
            char* lb = reinterpret_cast<char*>(lds);
            for (uint32_t half = 0; half < 2; ++half) {
                if (lds_used) __syncthreads();  // every thread has read the previous round
                if (hp == half) {
#pragma unroll
                    for (int r = 0; r < kRR; ++r) *reinterpret_cast<amp_t*>(lb + (sp ^ ph.cs_prev[r])) = v[r];
                }
                __syncthreads();
                if (hc == half) {
#pragma unroll
                    for (int r = 0; r < kRR; ++r) w[r] = *reinterpret_cast<const amp_t*>(lb + (sc ^ ph.cs_cur[r]));
                }
                lds_used = true;
            }
        }
        jt_prev = jt;
        uint32_t off = ph.op_off;
        uint32_t i = 0;
        if (p > 0 && ph.transpose) {
            // The transposed tile is in w: the phase's first op reads it from
            // there (no copy back), or it is copied when the phase has none.
            if (ph.n_ops > 0) {
                off = apply_op<0>(w, v, words, off, jt, ph, tile);
                i = 1;
            } else {
                copy_regs(w, v);
                pin_a(v);
            }
        }
        for (; i + 1 < ph.n_ops; i += 2) {
            off = apply_op<1>(v, w, words, off, jt, ph, tile);
            off = apply_op<0>(w, v, words, off, jt, ph, tile);
        }
        if (i < ph.n_ops) {
            apply_op<1>(v, w, words, off, jt, ph, tile);
            copy_regs(w, v);
            pin_a(v);
        }
    }
}

// Global loads and stores of one amplitude, optionally non-temporal (a cache
// hint: streamed data is not kept for reuse).
typedef double dbl2 __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ amp_t ld_amp(const amp_t* p) {
    if constexpr (NT) {
        const dbl2 v = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(p));
        return make_double2(v.x, v.y);
    } else {
        return *p;
    }
}

template <bool NT>
__device__ __forceinline__ void st_amp(amp_t* p, amp_t x) {
    if constexpr (NT) {
        dbl2 v;
        v.x = x.x;
        v.y = x.y;
        __builtin_nontemporal_store(v, reinterpret_cast<dbl2*>(p));
    } else {
        *p = x;
    }
}

// Tile number of workgroup b among `count`. Workgroups go to the 8 XCDs
// round-robin (b % 8), each XCD with its own L2: giving every XCD a
// contiguous range of tile numbers, walked in order, keeps neighbouring
// tiles (adjacent in memory through the lowest outside qubits) on one L2
// at about the same time. A 30-qubit pass ran 25-35% faster this way in a
// probe with scattered tile qubits.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t count) {
    if (count < 8) return b;
    return (b & 7u) * (count >> 3) + (b >> 3);
}

// One workgroup per tile.
template <bool NT>
__global__ void __launch_bounds__(kRT) k_rtile(amp_t* __restrict__ s, uint32_t nq, uint64_t smask,
                                              const RPhase* __restrict__ phases, uint32_t n_phases,
                                              const amp_t* __restrict__ words) {
    __shared__ amp_t lds[1 << (kRL - 1)];
    amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint64_t base = rtile_tile_base(tile, phases[0], ((uint64_t{1} << nq) - 1) & ~smask);
    Regs v, w;
    {
        // base, the thread's bits and the register's bits are disjoint: the
        // register part is added to a per-thread pointer.
        const RPhase& p0 = phases[0];
        const amp_t* bt = b + (base | rtile_thread_global(threadIdx.x, p0));
#pragma unroll
        for (int r = 0; r < kRR; ++r) v[r] = ld_amp<NT>(bt + p0.gofs[r]);
    }
    bool lds_used = false;
    run_phases(v, w, lds, phases, n_phases, words, tile, lds_used);
    {
        const RPhase& pl = phases[n_phases - 1];
        amp_t* bt = b + (base | rtile_thread_global(threadIdx.x, pl));
#pragma unroll
        for (int r = 0; r < kRR; ++r) st_amp<NT>(bt + pl.gofs[r], v[r]);
    }
}

// Wave tiles: one 64-thread workgroup per tile of 2^10 amplitudes (16 per
// lane), tile numbers mapped XCD by XCD. Every layout change stays within the
// wave, so no barrier ties waves together (a one-wave barrier only orders the
// wave's own LDS accesses). Transposes use 8 KiB of LDS in two halves, so the
// CU's LDS holds twelve waves, as many as the registers allow.
template <bool NT>
__global__ void __launch_bounds__(64) k_wtile(amp_t* __restrict__ s, uint32_t nq, uint64_t smask,
                                             const RPhase* __restrict__ phases, uint32_t n_phases,
                                             const amp_t* __restrict__ words) {
    __shared__ amp_t lds[1 << 9];
    amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint64_t base = rtile_tile_base(tile, phases[0], ((uint64_t{1} << nq) - 1) & ~smask);
    Regs v, w;
    {
        const RPhase& p0 = phases[0];
        const amp_t* bt = b + (base | rtile_thread_global(threadIdx.x, p0));
#pragma unroll
        for (int r = 0; r < kRR; ++r) v[r] = ld_amp<NT>(bt + p0.gofs[r]);
    }
    bool lds_used = false;
    run_phases(v, w, lds, phases, n_phases, words, tile, lds_used);
    {
        const RPhase& pl = phases[n_phases - 1];
        amp_t* bt = b + (base | rtile_thread_global(threadIdx.x, pl));
#pragma unroll
        for (int r = 0; r < kRR; ++r) st_amp<NT>(bt + pl.gofs[r], v[r]);
    }
}

// Streaming form: each workgroup walks tiles blockIdx.x, + gridDim.x, ...
// and issues the global loads of its next tile before running the phases
// of the current one, so HBM reads overlap the in-register gate work.
__global__ void __launch_bounds__(kRT) k_rtile_stream(amp_t* __restrict__ s, uint32_t nq, uint64_t smask,
                                                     const RPhase* __restrict__ phases, uint32_t n_phases,
                                                     const amp_t* __restrict__ words, uint32_t tiles) {
    __shared__ amp_t lds[1 << (kRL - 1)];
    amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint64_t outside = ((uint64_t{1} << nq) - 1) & ~smask;
    const RPhase& p0 = phases[0];
    const RPhase& pl = phases[n_phases - 1];
    const uint64_t gt0 = rtile_thread_global(threadIdx.x, p0);
    const uint64_t gtl = rtile_thread_global(threadIdx.x, pl);
    uint32_t tile = blockIdx.x;
    Regs v, w, nx;
    {
        const uint64_t gt = rtile_tile_base(tile, p0, outside) | gt0;
#pragma unroll
        for (int r = 0; r < kRR; ++r) v[r] = b[gt | p0.gofs[r]];
    }
    // Explicit waits for the vector-memory counter (vmcnt 0) keep the
    // compiler's own waits out of the phases: without them it cannot tell
    // the next tile's loads from the current tile's and waits for both
    // before the first gate.
    wait_vmem();
    bool lds_used = false;
    for (;;) {
        const uint64_t base = rtile_tile_base(tile, p0, outside);
        const uint32_t next = tile + gridDim.x;
        if (next < tiles) {
            const uint64_t gt = rtile_tile_base(next, p0, outside) | gt0;
#pragma unroll
            for (int r = 0; r < kRR; ++r) nx[r] = b[gt | p0.gofs[r]];
        }
        run_phases(v, w, lds, phases, n_phases, words, tile, lds_used);
        wait_vmem();  // the next tile has landed
        {
            const uint64_t gt = base | gtl;
#pragma unroll
            for (int r = 0; r < kRR; ++r) b[gt | pl.gofs[r]] = v[r];
        }
        if (next >= tiles) break;
        copy_regs(nx, v);
        tile = next;
    }
}

// One 16-byte-per-lane DMA from global memory into LDS: lane i's 16 bytes
// land at LDS byte address lds_addr + 16 i (lds_addr wave-uniform). Written
// in asm so that the compiler's own wait counting does not treat every
// later barrier as needing the transfer complete; the kernel below waits for
// it explicitly. No vector register is written, so nothing is in flight in
// a register.
__device__ __forceinline__ void dma16(const void* g, uint32_t lds_addr) {
    asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "s"(lds_addr) : "memory", "m0");
}

// s_waitcnt vmcnt(16): all but the 16 most recent vector-memory operations
// done (vmcnt low bits [3:0] = 0, high bits [15:14] = 1; expcnt and lgkmcnt
// at their maxima).
__device__ __forceinline__ void wait_vmem_but_16() { __builtin_amdgcn_s_waitcnt(0x4F70); }

// Landing form: one workgroup per CU walks tiles blockIdx.x, + gridDim.x, ...
// The next tile streams from HBM straight into an LDS landing buffer (64 KiB,
// tile index order) by DMA while the current tile's phases run in
// registers; the tile is then read from LDS in phase 0's layout, so loads
// need no registers and no particular layout.
__global__ void __launch_bounds__(kRT) k_rtile_land(amp_t* __restrict__ s, uint32_t nq, uint64_t smask,
                                                   const RPhase* __restrict__ phases, uint32_t n_phases,
                                                   const amp_t* __restrict__ words, uint32_t tiles) {
    __shared__ amp_t land[1 << kRL];
    __shared__ amp_t lds[1 << (kRL - 1)];
    amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint64_t outside = ((uint64_t{1} << nq) - 1) & ~smask;
    // Tile index j = 256 k + thread: the thread's part over the 8 lowest
    // tile qubits, k's part over the 4 highest.
    uint64_t lo_mask = 0, hi_mask = smask;
    for (int i = 0; i < 8; ++i) {
        const uint64_t bit = hi_mask & (~hi_mask + 1);
        lo_mask |= bit;
        hi_mask ^= bit;
    }
    const uint64_t lane_off = pdep64(threadIdx.x, lo_mask);
    const uint32_t wave = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6)));
    const uint32_t land_base = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(land)) + wave * 64 * 16;
    auto issue = [&](uint32_t t) {
        const amp_t* g = b + (rtile_tile_base(t, phases[0], outside) | lane_off);
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) dma16(g + pdep64(k, hi_mask), land_base + k * 256 * 16);
    };
    const RPhase& p0 = phases[0];
    const RPhase& pl = phases[n_phases - 1];
    const uint32_t jt0 = rtile_thread_bits(threadIdx.x, p0);
    const uint64_t gtl = rtile_thread_global(threadIdx.x, pl);
    uint32_t tile = blockIdx.x;
    issue(tile);
    wait_vmem();
    __syncthreads();
    bool lds_used = false;
    Regs v, w;
    for (;;) {
#pragma unroll
        for (int r = 0; r < kRR; ++r) v[r] = land[jt0 | p0.jr[r]];
        __syncthreads();  // every thread has read the landing buffer
        const uint32_t next = tile + gridDim.x;
        if (next < tiles) issue(next);
        run_phases(v, w, lds, phases, n_phases, words, tile, lds_used);
        {
            const uint64_t gt = rtile_tile_base(tile, p0, outside) | gtl;
#pragma unroll
            for (int r = 0; r < kRR; ++r) b[gt | pl.gofs[r]] = v[r];
        }
        if (next >= tiles) break;
        wait_vmem_but_16();  // the next tile's DMA (older than the 16 stores) has landed
        __syncthreads();
        tile = next;
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

uint32_t pext_host(uint64_t v, uint64_t mask) {
    uint32_t r = 0, bit = 1;
    for (; mask; bit <<= 1) {
        const uint64_t low = mask & (~mask + 1);
        if (v & low) r |= bit;
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

struct Group {
    uint32_t reg = 0;  // register bits
    std::vector<uint32_t> ops;
};

// First-fit grouping in program order: an op joins while its register bits
// fit (at most 4, all within `allowed`) and no skipped earlier op shares a
// bit with it. Returns the group and leaves the skipped ops in `pending`.
Group take_forward(std::vector<uint32_t>& pending, const std::vector<uint32_t>& touched,
                   const std::vector<uint32_t>& dense, uint32_t allowed) {
    Group g;
    uint32_t blocked = 0;
    std::vector<uint32_t> rest;
    for (uint32_t i : pending) {
        if ((touched[i] & blocked) || (dense[i] & ~allowed) || popc(g.reg | dense[i]) > 4) {
            blocked |= touched[i];
            rest.push_back(i);
            continue;
        }
        g.reg |= dense[i];
        g.ops.push_back(i);
    }
    pending.swap(rest);
    return g;
}

// The same from the end: ops that can run after everything left in `pending`.
Group take_backward(std::vector<uint32_t>& pending, const std::vector<uint32_t>& touched,
                    const std::vector<uint32_t>& dense, uint32_t allowed) {
    Group g;
    uint32_t blocked = 0;
    std::vector<uint32_t> rest;
    for (auto it = pending.rbegin(); it != pending.rend(); ++it) {
        const uint32_t i = *it;
        if ((touched[i] & blocked) || (dense[i] & ~allowed) || popc(g.reg | dense[i]) > 4) {
            blocked |= touched[i];
            rest.push_back(i);
            continue;
        }
        g.reg |= dense[i];
        g.ops.push_back(i);
    }
    std::reverse(g.ops.begin(), g.ops.end());
    std::reverse(rest.begin(), rest.end());
    pending.swap(rest);
    return g;
}

uint32_t pad_high(uint32_t reg, uint32_t L) {
    for (int p = static_cast<int>(L) - 1; p >= 0 && popc(reg) < 4; --p) reg |= 1u << p;
    return reg;
}

uint32_t matrix_words(uint32_t kind) {
    switch (kind) {
        case RK_DENSE1: case RK_DIAG: return 4;
        case RK_DENSE2: return 16;
        case RK_PAULI: return 2;
        case RK_SEL1: return 8;
        case RK_DIAG1: return 2;
        case RK_DENSE2R: return 16;
        default: return 0;
    }
}

// Merges each singly-controlled 2x2 op C-V (V = X, as in CX, or a diagonal,
// as in CZ and CP) with the uncontrolled dense 2x2 op U on its target that
// comes next on that qubit, or else with the one just before it: U C-V (or
// C-V U) is one 2x2 op on the target whose matrix the control bit selects, U
// where it is 0 and U V (or V U) where it is 1. The merged op stands where
// C-V stood; U moves there past ops that do not touch its qubit. A
// controlled phase (V = diag(1, p)) is symmetric in its two qubits, so it may
// also merge with a U on its control, selected by its target. The merged op
// costs what U alone costs: the X's register moves, or the diagonal's
// multiplications, are gone. `sel` marks the merged ops (their entries 4..7
// hold the second matrix). Returns how many ops were absorbed.
uint32_t merge_selects(std::vector<TileOp>& ops, std::vector<uint8_t>& sel) {
    const size_t n = ops.size();
    std::vector<uint8_t> dead(n, 0);
    sel.assign(n, 0);
    std::vector<uint32_t> touched(n), dense(n);
    for (size_t i = 0; i < n; ++i) local_masks(ops[i], touched[i], dense[i]);
    auto plain_u = [&](size_t i, uint32_t t) {
        const TileOp& u = ops[i];
        return !sel[i] && u.kind == TK_DENSE1 && u.t0 == t && u.lctrl == 0 && u.gctrl == 0 && u.member < 0;
    };
    // The plain 2x2 op on tile bit t next to op k: the next op on t if it is
    // one, else the previous op on t if that is one; n when neither is.
    auto partner = [&](size_t k, uint32_t t, bool& fwd) -> size_t {
        const uint32_t tb = 1u << t;
        for (size_t j = k + 1; j < n; ++j) {
            if (dead[j] || !(touched[j] & tb)) continue;
            if (plain_u(j, t)) {
                fwd = true;
                return j;
            }
            break;
        }
        for (size_t j = k; j-- > 0;) {
            if (dead[j] || !(touched[j] & tb)) continue;
            if (plain_u(j, t)) {
                fwd = false;
                return j;
            }
            break;
        }
        return n;
    };
    auto mul2 = [](const amp_t* a, const amp_t* b, amp_t* c) {
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) c[2 * i + j] = cadd(cmul(a[2 * i], b[j]), cmul(a[2 * i + 1], b[2 + j]));
    };
    const amp_t zero = make_double2(0.0, 0.0), one = make_double2(1.0, 0.0);
    uint32_t merged = 0;
    for (size_t k = 0; k < n; ++k) {
        const TileOp& x = ops[k];
        if (dead[k] || sel[k] || x.member >= 0) continue;
        if (popc(x.lctrl) + popc(x.gctrl) != 1) continue;
        amp_t V[4];
        uint32_t t = 0;
        bool symmetric = false;
        if (x.kind == TK_DENSE1 && is_x(x)) {
            t = x.t0;
            V[0] = zero, V[1] = one, V[2] = one, V[3] = zero;
        } else if (x.kind == TK_DIAG && popc(x.dl0) == 1 && x.dl1 == 0 && x.dg0 == 0 && x.dg1 == 0) {
            t = static_cast<uint32_t>(__builtin_ctz(x.dl0));
            V[0] = x.m[0], V[1] = zero, V[2] = zero, V[3] = x.m[1];
            symmetric = x.m[0].x == 1.0 && x.m[0].y == 0.0 && x.lctrl != 0;
        } else {
            continue;
        }
        uint32_t lsel = x.lctrl;
        uint64_t gsel = x.gctrl;
        bool fwd = false;
        size_t u = partner(k, t, fwd);
        if (u == n && symmetric) {  // diag(1, p) on the control, selected by the target
            const uint32_t t2 = static_cast<uint32_t>(__builtin_ctz(x.lctrl));
            u = partner(k, t2, fwd);
            if (u != n) {
                lsel = 1u << t;
                gsel = 0;
                t = t2;
            }
        }
        if (u == n) continue;
        TileOp m = x;
        m.kind = TK_DENSE1;
        m.t0 = t;
        m.lctrl = lsel;
        m.gctrl = gsel;
        m.dl0 = m.dl1 = 0;
        m.dg0 = m.dg1 = 0;
        for (int e = 0; e < 4; ++e) m.m[e] = ops[u].m[e];
        if (fwd)  // U after C-V: U V where the bit is 1
            mul2(ops[u].m, V, m.m + 4);
        else      // U before C-V: V U
            mul2(V, ops[u].m, m.m + 4);
        ops[k] = m;
        local_masks(ops[k], touched[k], dense[k]);
        sel[k] = 1;
        dead[u] = 1;
        ++merged;
    }
    if (merged) {
        size_t w = 0;
        for (size_t i = 0; i < n; ++i) {
            if (dead[i]) continue;
            ops[w] = ops[i];
            sel[w] = sel[i];
            ++w;
        }
        ops.resize(w);
        sel.resize(w);
    }
    return merged;
}

}  // namespace

bool build_rtile(const std::vector<TileOp>& tops_in, uint32_t L, uint64_t smask, uint32_t nq, RTileProgram& out) {
    if (L != static_cast<uint32_t>(kRL) && L != 10u && L != 13u) return false;
    // 13-qubit programs run only as specialized passes: the fields below that
    // only the interpreting kernels read (thread positions, tile-number
    // insertion masks, transpose slots) stay zero for them.
    const bool interp = L <= static_cast<uint32_t>(kRL);
    out.phases.clear();
    out.words.clear();
    std::vector<TileOp> tops = tops_in;
    std::vector<uint8_t> sel;
    out.merged = merge_selects(tops, sel);
    const uint32_t all = (1u << L) - 1;
    const uint64_t outside = ((nq >= 64 ? ~uint64_t{0} : (uint64_t{1} << nq) - 1)) & ~smask;
    std::vector<uint32_t> touched(tops.size()), dense(tops.size());
    for (size_t i = 0; i < tops.size(); ++i) {
        local_masks(tops[i], touched[i], dense[i]);
        if (popc(dense[i]) > 4) return false;
    }
    // Register bits of the first and last phases avoid the tile's contiguous
    // low qubits (tile bit p is qubit p for p below the run length), which
    // then index the lanes and keep global accesses in contiguous runs.
    const uint32_t run = static_cast<uint32_t>(__builtin_ctzll(~smask));
    const uint32_t avoid = (1u << std::min<uint32_t>(run, 6)) - 1;
    const uint32_t edge = all & ~avoid;

    std::vector<uint32_t> pending(tops.size());
    for (uint32_t i = 0; i < tops.size(); ++i) pending[i] = i;
    std::vector<Group> groups;
    groups.push_back(take_forward(pending, touched, dense, edge));
    Group last;
    if (!pending.empty()) last = take_backward(pending, touched, dense, edge);
    while (!pending.empty()) groups.push_back(take_forward(pending, touched, dense, all));
    if (!last.ops.empty()) groups.push_back(std::move(last));
    for (auto& g : groups) g.reg = pad_high(g.reg, L);
    if (groups.size() > 1 && groups.front().ops.empty() && (groups[1].reg & avoid) == 0)
        groups.erase(groups.begin());  // the second layout loads as well
    if (groups.back().reg & avoid) groups.push_back(Group{pad_high(0, L), {}});  // store layout
    // Neighbours with one layout run as one phase.
    std::vector<Group> merged;
    for (auto& g : groups) {
        if (!merged.empty() && merged.back().reg == g.reg) {
            merged.back().ops.insert(merged.back().ops.end(), g.ops.begin(), g.ops.end());
        } else {
            merged.push_back(std::move(g));
        }
    }

    // Encode: per phase its layout, per op a header and its matrix words.
    auto put_hdr = [&](const RegOpHdr& h) {
        amp_t w[2];
        std::memcpy(w, &h, sizeof h);
        out.words.push_back(w[0]);
        out.words.push_back(w[1]);
    };
    uint32_t prev_reg = ~0u;
    for (const Group& g : merged) {
        const uint32_t rmask = g.reg;
        RPhase ph;
        std::memset(&ph, 0, sizeof ph);
        ph.tmask = all & ~rmask;
        ph.tmask_global = pdep_host(ph.tmask, smask);
        ph.transpose = (prev_reg != ~0u && prev_reg != rmask) ? 1u : 0u;
        prev_reg = rmask;
        int rb[4], k = 0;
        for (int p = 0; p < static_cast<int>(L); ++p)
            if (rmask & (1u << p)) rb[k++] = p;
        for (int i = 0; i < 4; ++i) ph.hole_low[i] = (1u << rb[i]) - 1u;
        if (interp) {
            int t = 0;
            for (int q = 0; q < 64; ++q)
                if (ph.tmask_global & (uint64_t{1} << q)) ph.tpos[t++] = static_cast<uint8_t>(q);
            t = 0;
            for (int q = 0; q < 64; ++q)
                if (smask & (uint64_t{1} << q)) ph.tile_low[t++] = (uint64_t{1} << q) - 1;
            for (; t < 12; ++t) ph.tile_low[t] = ~uint64_t{0};  // no insertion
        }
        for (int r = 0; r < kRR; ++r) {
            uint32_t j = 0;
            for (int i = 0; i < 4; ++i)
                if (r & (1 << i)) j |= 1u << rb[i];
            ph.jr[r] = j;
            ph.gofs[r] = pdep_host(j, smask);
        }
        if (ph.transpose && interp) {
            // Split bit: a thread bit of both layouts, preferring one that is
            // the top thread-id bit in both (then whole waves take turns).
            const RPhase& pp = out.phases.back();
            const uint32_t both = pp.tmask & ph.tmask;
            auto rank = [](uint32_t mask, uint32_t bit) { return popc(mask & ((1u << bit) - 1u)); };
            int best = -1, best_score = -1;
            for (uint32_t b = 0; b < static_cast<uint32_t>(kRL); ++b) {
                if (!(both & (1u << b))) continue;
                const int score = rank(pp.tmask, b) + rank(ph.tmask, b);
                if (score > best_score) {
                    best_score = score;
                    best = static_cast<int>(b);
                }
            }
            ph.h = static_cast<uint32_t>(best);
            for (int r = 0; r < kRR; ++r) {
                ph.cs_prev[r] = rtile_slot(pp.jr[r], ph.h) << 4;
                ph.cs_cur[r] = rtile_slot(ph.jr[r], ph.h) << 4;
            }
        }
        auto ridx = [&](uint32_t pos) {
            for (int i = 0; i < 4; ++i)
                if (rb[i] == static_cast<int>(pos)) return static_cast<uint32_t>(i);
            return 0u;
        };
        ph.op_off = static_cast<uint32_t>(out.words.size());
        for (uint32_t i : g.ops) {
            const TileOp& t = tops[i];
            RegOpHdr h;
            std::memset(&h, 0, sizeof h);
            uint32_t kind = 0, b0 = 0, b1 = 0, xr = 0, rctl = 0;
            for (int q = 0; q < 4; ++q)
                if (t.lctrl & (1u << rb[q])) rctl |= 1u << q;
            h.tctl = t.lctrl & ph.tmask;
            h.gctrl = pext_host(t.gctrl, outside);
            if (t.member >= 0) h.member = (static_cast<uint32_t>(t.member) + 1) | (t.only ? 1u << 31 : 0u);
            switch (t.kind) {
                case TK_DENSE1:
                    b0 = ridx(t.t0);
                    if (sel[i]) {  // the control bit selects the matrix; no control left
                        kind = RK_SEL1;
                        if (t.lctrl & rmask)
                            xr = 1 + ridx(static_cast<uint32_t>(__builtin_ctz(t.lctrl)));
                        else if (t.lctrl)
                            h.la = t.lctrl;
                        else
                            h.ga = pext_host(t.gctrl, outside);
                        rctl = 0;
                        h.tctl = 0;
                        h.gctrl = 0;
                        break;
                    }
                    kind = is_x(t) ? RK_X : RK_DENSE1;
                    break;
                case TK_DENSE2: {
                    bool real = rctl == 0 && h.tctl == 0;
                    for (int e = 0; e < 16 && real; ++e) real = t.m[e].y == 0.0;
                    kind = real ? RK_DENSE2R : RK_DENSE2;
                    b0 = ridx(t.t0);
                    b1 = ridx(t.t1);
                    break;
                }
                case TK_DIAG:
                    if (t.dl1 == 0 && t.dg1 == 0 && popc(t.dl0) + popc(t.dg0) == 1) {  // one qubit
                        kind = RK_DIAG1;
                        if (t.dl0 & rmask)
                            xr = 1 + ridx(static_cast<uint32_t>(__builtin_ctz(t.dl0)));
                        else if (t.dl0)
                            h.la = t.dl0;
                        else
                            h.ga = pext_host(t.dg0, outside);
                        b1 = (t.m[0].x == 1.0 && t.m[0].y == 0.0) ? 1u : 0u;
                        break;
                    }
                    kind = RK_DIAG;
                    h.la = t.dl0;
                    h.lb = t.dl1;
                    h.ga = pext_host(t.dg0, outside);
                    h.gb = pext_host(t.dg1, outside);
                    break;
                default:
                    kind = RK_PAULI;
                    for (int bit = 0; bit < static_cast<int>(L); ++bit)
                        if (t.xl & (1u << bit)) xr |= 1u << ridx(bit);
                    h.la = t.zl;
                    h.ga = pext_host(t.zg, outside);
                    break;
            }
            const uint32_t nm = matrix_words(kind);
            h.code = kind | (b0 << 4) | (b1 << 6) | (xr << 8) | (rctl << 12) | (nm << 16);
            if (h.gctrl || h.member) h.code |= kCodeRestricted;
            if (rctl == 0 && h.tctl == 0) h.code |= kCodeFree;
            put_hdr(h);
            for (uint32_t e = 0; e < nm; ++e) out.words.push_back(t.m[e]);
            ++ph.n_ops;
        }
        out.phases.push_back(ph);
    }
    return true;
}

void launch_rtile(amp_t* s, uint32_t nq, uint32_t L, uint64_t smask, const RPhase* d_phases, uint32_t n_phases,
                  const amp_t* d_words, uint32_t batch) {
    if (L == 10u) {
        const uint64_t wt = uint64_t{1} << (nq - 10);
        const bool nt = (uint64_t(batch) << nq) * sizeof(amp_t) >= (uint64_t{1} << 30);
        const dim3 grid(static_cast<unsigned>(wt), batch);
        if (nt) k_wtile<true><<<grid, 64, 0, stream()>>>(s, nq, smask, d_phases, n_phases, d_words);
        else k_wtile<false><<<grid, 64, 0, stream()>>>(s, nq, smask, d_phases, n_phases, d_words);
        return;
    }
    const uint64_t tiles = uint64_t{1} << (nq - kRL);
    if (g_rtile_land && tiles > uint64_t(num_cus())) {
        k_rtile_land<<<dim3(static_cast<unsigned>(num_cus()), batch), kRT, 0, stream()>>>(
            s, nq, smask, d_phases, n_phases, d_words, static_cast<uint32_t>(tiles));
        return;
    }
    const uint64_t resident = uint64_t(num_cus()) * g_rtile_blocks_per_cu;
    if (g_rtile_stream && tiles > resident) {
        k_rtile_stream<<<dim3(static_cast<unsigned>(resident), batch), kRT, 0, stream()>>>(
            s, nq, smask, d_phases, n_phases, d_words, static_cast<uint32_t>(tiles));
        return;
    }
    // Non-temporal loads and stores once the registers passed over reach
    // 1 GiB: they cannot stay in the caches anyway, and a copy-like pass at
    // 30 qubits runs 4% faster. Smaller states keep the cached path, which
    // later passes and reductions can hit.
    const bool nt = (uint64_t(batch) << nq) * sizeof(amp_t) >= (uint64_t{1} << 30);
    if (nt)
        k_rtile<true><<<dim3(static_cast<unsigned>(tiles), batch), kRT, 0, stream()>>>(s, nq, smask, d_phases,
                                                                                      n_phases, d_words);
    else
        k_rtile<false><<<dim3(static_cast<unsigned>(tiles), batch), kRT, 0, stream()>>>(s, nq, smask, d_phases,
                                                                                       n_phases, d_words);
}

}  // namespace lqcs_dev
