// Device side of the specialized register-tile passes.
//
// A specialized pass is the register-tile pass of one program (rtile.hip)
// with its structure fixed at compile time: the tile qubits, every phase's
// register bits and thread-bit order, and every op's kind and register bits.
// The generator (spec_pass.hip) writes a kernel body that calls the
// templates below in program order, and compiles it at run time (hipRTC).
// Matrix entries stay in device memory, read with scalar loads from the
// pass's op words, so new angles need no new kernel.
//
// Fixed structure buys two things over the interpreting kernel:
// * no op dispatch: no scalar decode, no branch chain per op;
// * one register set updated in place, since no dispatch join forces a
//   second set: about 110 VGPRs instead of 163, so four waves per SIMD hide
//   the HBM latency behind other waves' FP64 work instead of three;
// * compile-time matrix entry flags (entries with a zero imaginary part, or
//   zero entries), whose products are left out.
//
// This header is self-contained: it is compiled by hipRTC (which provides
// the HIP device declarations) and, for the build-time check, by hipcc.

typedef unsigned int sp_u32;
typedef unsigned long long sp_u64;
typedef double2 sp_amp;

#define SP_INLINE __device__ __forceinline__

namespace sp {

typedef sp_amp Regs[16];

// m0*a0 + m1*a1 in the engine's operation order (bit-identical results).
SP_INLINE sp_amp cmac2(sp_amp m0, sp_amp a0, sp_amp m1, sp_amp a1) {
    return make_double2(m0.x * a0.x - m0.y * a0.y + m1.x * a1.x - m1.y * a1.y,
                        m0.x * a0.y + m0.y * a0.x + m1.x * a1.y + m1.y * a1.x);
}
SP_INLINE sp_amp cmul(sp_amp a, sp_amp b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
SP_INLINE sp_amp pick(bool c, sp_amp p, sp_amp q) { return make_double2(c ? p.x : q.x, c ? p.y : q.y); }

// The thread id as a value the compiler cannot see through: each phase
// derives its layout's thread bits and LDS offsets from a fresh copy, so
// they are computed where used instead of being kept live from the start of
// the kernel (every long-lived value costs a VGPR in every wave).
SP_INLINE sp_u32 fresh(sp_u32 x) {
    asm volatile("" : "+v"(x));
    return x;
}

// A wave-uniform address kept in scalar registers: loads and stores then take
// it as their base with one 32-bit per-lane offset.
typedef __attribute__((address_space(1))) char sp_gchar;
SP_INLINE sp_gchar* sbase(char* p) {
    sp_gchar* g = (sp_gchar*)p;
    asm volatile("" : "+s"(g));
    return g;
}

// One matrix entry of the op words: the same address in every lane, so a
// scalar load.
SP_INLINE sp_amp uni(const sp_amp* __restrict__ p) {
    const sp_amp v = *p;
    const sp_u64 ux = static_cast<sp_u64>(__double_as_longlong(v.x));
    const sp_u64 uy = static_cast<sp_u64>(__double_as_longlong(v.y));
    const sp_u64 rx = static_cast<sp_u64>(static_cast<sp_u32>(__builtin_amdgcn_readfirstlane(static_cast<int>(ux)))) |
                      (static_cast<sp_u64>(static_cast<sp_u32>(__builtin_amdgcn_readfirstlane(static_cast<int>(ux >> 32)))) << 32);
    const sp_u64 ry = static_cast<sp_u64>(static_cast<sp_u32>(__builtin_amdgcn_readfirstlane(static_cast<int>(uy)))) |
                      (static_cast<sp_u64>(static_cast<sp_u32>(__builtin_amdgcn_readfirstlane(static_cast<int>(uy >> 32)))) << 32);
    return make_double2(__longlong_as_double(static_cast<long long>(rx)), __longlong_as_double(static_cast<long long>(ry)));
}

// ---- entry flags ----------------------------------------------------------------
//
// Two bits per matrix entry, fixed when the pass is compiled: bit 0, the
// imaginary part is exactly zero (a U3's top-left entry, every entry of H or
// RY); bit 1, the entry is exactly zero. The products of such parts are left
// out; what remains is summed in the same order as cmac2.

template <int F>
SP_INLINE sp_amp cm1(sp_amp m, sp_amp a) {
    if constexpr ((F & 1) != 0) return make_double2(m.x * a.x, m.x * a.y);
    else return make_double2(m.x * a.x - m.y * a.y, m.x * a.y + m.y * a.x);
}

template <int F>
SP_INLINE sp_amp cma(sp_amp acc, sp_amp m, sp_amp a) {
    if constexpr ((F & 1) != 0) return make_double2(acc.x + m.x * a.x, acc.y + m.x * a.y);
    else return make_double2(acc.x + m.x * a.x - m.y * a.y, acc.y + m.x * a.y + m.y * a.x);
}

// m0*a0 + m1*a1 under the entries' flags F0, F1.
template <int F0, int F1>
SP_INLINE sp_amp cmac2f(sp_amp m0, sp_amp a0, sp_amp m1, sp_amp a1) {
    if constexpr ((F0 & 2) != 0 && (F1 & 2) != 0) return make_double2(0.0, 0.0);
    else if constexpr ((F0 & 2) != 0) return cm1<F1>(m1, a1);
    else if constexpr ((F1 & 2) != 0) return cm1<F0>(m0, a0);
    else return cma<F1>(cm1<F0>(m0, a0), m1, a1);
}

// ---- 2x2 ops on register bit B (F: flags of entries 0..3, two bits each) --------

template <int B, unsigned F>
SP_INLINE void d1v(Regs& v, sp_amp m0, sp_amp m1, sp_amp m2, sp_amp m3) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & (1 << B)) continue;
        const sp_amp a0 = v[r], a1 = v[r | (1 << B)];
        v[r] = cmac2f<F & 3, (F >> 2) & 3>(m0, a0, m1, a1);
        v[r | (1 << B)] = cmac2f<(F >> 4) & 3, (F >> 6) & 3>(m2, a0, m3, a1);
    }
}

template <int B, unsigned F>
SP_INLINE void d1(Regs& v, const sp_amp* __restrict__ m) {
    d1v<B, F>(v, uni(m), uni(m + 1), uni(m + 2), uni(m + 3));
}

// The matrix chosen by register bit S: entries 0..3 where it is 0, 4..7
// where it is 1 (F: flags of all eight entries).
template <int B, int S, unsigned F>
SP_INLINE void sel_r(Regs& v, const sp_amp* __restrict__ m) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & (1 << B)) continue;
        const sp_amp a0 = v[r], a1 = v[r | (1 << B)];
        if (r & (1 << S)) {
            v[r] = cmac2f<(F >> 8) & 3, (F >> 10) & 3>(uni(m + 4), a0, uni(m + 5), a1);
            v[r | (1 << B)] = cmac2f<(F >> 12) & 3, (F >> 14) & 3>(uni(m + 6), a0, uni(m + 7), a1);
        } else {
            v[r] = cmac2f<F & 3, (F >> 2) & 3>(uni(m), a0, uni(m + 1), a1);
            v[r | (1 << B)] = cmac2f<(F >> 4) & 3, (F >> 6) & 3>(uni(m + 2), a0, uni(m + 3), a1);
        }
    }
}

// The second matrix is the first with its columns swapped (U X) or its rows
// swapped (X U), as when a CX is merged into the 2x2 op on its target: the lanes whose select bit is set swap each pair's inputs
// (COLS) or outputs before or after the first matrix, so the matrix stays in
// scalar registers instead of a per-lane copy.
template <int B, unsigned F, bool COLS>
SP_INLINE void sel_swap(Regs& v, const sp_amp* __restrict__ m, bool s) {
    const sp_amp m0 = uni(m), m1 = uni(m + 1), m2 = uni(m + 2), m3 = uni(m + 3);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & (1 << B)) continue;
        sp_amp a0 = v[r], a1 = v[r | (1 << B)];
        if (COLS) {
            const sp_amp t0 = pick(s, a1, a0), t1 = pick(s, a0, a1);
            a0 = t0;
            a1 = t1;
        }
        const sp_amp n0 = cmac2f<F & 3, (F >> 2) & 3>(m0, a0, m1, a1);
        const sp_amp n1 = cmac2f<(F >> 4) & 3, (F >> 6) & 3>(m2, a0, m3, a1);
        if (COLS) {
            v[r] = n0;
            v[r | (1 << B)] = n1;
        } else {
            v[r] = pick(s, n1, n0);
            v[r | (1 << B)] = pick(s, n0, n1);
        }
    }
}

// The matrix chosen per lane (select bit among the thread's tile bits); F:
// the flags both matrices share.
template <int B, unsigned F>
SP_INLINE void sel_l(Regs& v, const sp_amp* __restrict__ m, bool s) {
    d1v<B, F>(v, pick(s, uni(m + 4), uni(m)), pick(s, uni(m + 5), uni(m + 1)), pick(s, uni(m + 6), uni(m + 2)),
              pick(s, uni(m + 7), uni(m + 3)));
}

// Controlled 2x2: pairs whose register bits miss a control (RC), or lanes
// whose thread bits miss one (tp false), keep their values.
template <int B, int RC, unsigned F>
SP_INLINE void d1_c(Regs& v, const sp_amp* __restrict__ m, bool tp) {
    const sp_amp m0 = uni(m), m1 = uni(m + 1), m2 = uni(m + 2), m3 = uni(m + 3);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & (1 << B)) continue;
        if ((r & RC) != RC) continue;
        const sp_amp a0 = v[r], a1 = v[r | (1 << B)];
        const sp_amp n0 = cmac2f<F & 3, (F >> 2) & 3>(m0, a0, m1, a1);
        const sp_amp n1 = cmac2f<(F >> 4) & 3, (F >> 6) & 3>(m2, a0, m3, a1);
        v[r] = pick(tp, n0, a0);
        v[r | (1 << B)] = pick(tp, n1, a1);
    }
}

// X on register bit B for the registers that meet the register controls RC,
// in lanes where tp holds (TP_ALL: every lane).
template <int B, int RC, bool TP_ALL>
SP_INLINE void x_c(Regs& v, bool tp) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & (1 << B)) continue;
        if ((r & RC) != RC) continue;
        const sp_amp a0 = v[r], a1 = v[r | (1 << B)];
        if (TP_ALL) {
            v[r] = a1;
            v[r | (1 << B)] = a0;
        } else {
            v[r] = pick(tp, a1, a0);
            v[r | (1 << B)] = pick(tp, a0, a1);
        }
    }
}

// ---- 4x4 ops on register bits B0 (matrix bit 0) and B1 -------------------------

// Row of a 4x4 op times x under the flags F of its four entries.
template <unsigned F>
SP_INLINE sp_amp row4(const sp_amp* __restrict__ m, const sp_amp (&x)[4]) {
    const sp_amp m0 = uni(m), m1 = uni(m + 1), m2 = uni(m + 2), m3 = uni(m + 3);
    if constexpr ((F & 0xFFu) == 0xAAu) {
        return make_double2(0.0, 0.0);
    } else {
        // The first nonzero entry starts the sum; the others follow in order.
        constexpr int f0 = F & 3, f1 = (F >> 2) & 3, f2 = (F >> 4) & 3, f3 = (F >> 6) & 3;
        sp_amp acc;
        if constexpr ((f0 & 2) == 0) acc = cm1<f0>(m0, x[0]);
        if constexpr ((f1 & 2) == 0) {
            if constexpr ((f0 & 2) == 0) acc = cma<f1>(acc, m1, x[1]);
            else acc = cm1<f1>(m1, x[1]);
        }
        if constexpr ((f2 & 2) == 0) {
            if constexpr ((f0 & 2) == 0 || (f1 & 2) == 0) acc = cma<f2>(acc, m2, x[2]);
            else acc = cm1<f2>(m2, x[2]);
        }
        if constexpr ((f3 & 2) == 0) {
            if constexpr ((f0 & 2) == 0 || (f1 & 2) == 0 || (f2 & 2) == 0) acc = cma<f3>(acc, m3, x[3]);
            else acc = cm1<f3>(m3, x[3]);
        }
        return acc;
    }
}

// 4x4 op on register bits B0 (matrix bit 0) and B1; F: flags of the 16
// entries, row-major.
template <int B0, int B1, int RC, unsigned F>
SP_INLINE void d2(Regs& v, const sp_amp* __restrict__ m, bool tp) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & ((1 << B0) | (1 << B1))) continue;
        if ((r & RC) != RC) continue;
        const int idx[4] = {r, r | (1 << B0), r | (1 << B1), r | (1 << B0) | (1 << B1)};
        sp_amp x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = v[idx[k]];
        const sp_amp n0 = row4<F & 0xFFu>(m, x), n1 = row4<(F >> 8) & 0xFFu>(m + 4, x);
        const sp_amp n2 = row4<(F >> 16) & 0xFFu>(m + 8, x), n3 = row4<(F >> 24) & 0xFFu>(m + 12, x);
        v[idx[0]] = pick(tp, n0, x[0]);
        v[idx[1]] = pick(tp, n1, x[1]);
        v[idx[2]] = pick(tp, n2, x[2]);
        v[idx[3]] = pick(tp, n3, x[3]);
    }
}

// ---- diagonals -------------------------------------------------------------------

// diag(d0, d1) on register bit B (ONE0: d0 is exactly 1).
template <int B, bool ONE0, int RC>
SP_INLINE void dg1_r(Regs& v, const sp_amp* __restrict__ m, bool tp) {
    const sp_amp d0 = uni(m), d1v_ = uni(m + 1);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if ((r & RC) != RC) continue;
        const bool hi = ((r >> B) & 1) != 0;
        if (ONE0 && !hi) continue;
        v[r] = pick(tp, cmul(v[r], hi ? d1v_ : d0), v[r]);
    }
}

// diag(d0, d1) on a qubit that is the same for all of a lane's registers.
template <int RC>
SP_INLINE void dg1_u(Regs& v, const sp_amp* __restrict__ m, bool s, bool tp) {
    const sp_amp d = pick(s, uni(m + 1), uni(m));
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if ((r & RC) != RC) continue;
        v[r] = pick(tp, cmul(v[r], d), v[r]);
    }
}

// General diagonal over two qubits: factor index g (the tile-number part,
// uniform) | bit LA of the register's tile index | bit LB << 1. JR[r] holds
// register r's tile-index bits, JT the thread's.
template <int RC>
SP_INLINE void dg_gen(Regs& v, const sp_amp* __restrict__ m, sp_u32 jt, const sp_u32 (&jr)[16], sp_u32 la, sp_u32 lb,
                      sp_u32 g, bool tp) {
    const sp_amp d0 = uni(m), d1 = uni(m + 1), d2 = uni(m + 2), d3 = uni(m + 3);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if ((r & RC) != RC) continue;
        const sp_u32 j = jt | jr[r];
        const sp_u32 sel = g | ((j & la) ? 1u : 0u) | ((j & lb) ? 2u : 0u);
        const sp_amp d = sel == 0 ? d0 : sel == 1 ? d1 : sel == 2 ? d2 : d3;
        v[r] = pick(tp, cmul(v[r], d), v[r]);
    }
}

// ---- Pauli rotations exp(-i theta/2 P): c = m[0], yphase = m[1] -------------------

// X part on register bits XR; Z sign from the tile-index bits ZL of each
// amplitude and the parity gp of the tile-number Z bits.
template <int XR>
SP_INLINE void pauli_x(Regs& v, const sp_amp* __restrict__ m, sp_u32 jt, const sp_u32 (&jr)[16], sp_u32 zl, sp_u32 gp) {
    constexpr int low = XR & -XR;
    const sp_amp c = uni(m), yp = uni(m + 1);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & low) continue;
        const int r2 = r ^ XR;
        const sp_u32 j1 = jt | jr[r], j2 = jt | jr[r2];
        const double s1 = ((__popc(j1 & zl) ^ gp) & 1u) ? -1.0 : 1.0;
        const double s2 = ((__popc(j2 & zl) ^ gp) & 1u) ? -1.0 : 1.0;
        const sp_amp a1 = v[r], a2 = v[r2];
        v[r] = cmac2(c, a1, make_double2(yp.x * s2, yp.y * s2), a2);
        v[r2] = cmac2(make_double2(yp.x * s1, yp.y * s1), a1, c, a2);
    }
}

SP_INLINE void pauli_z(Regs& v, const sp_amp* __restrict__ m, sp_u32 jt, const sp_u32 (&jr)[16], sp_u32 zl, sp_u32 gp) {
    const sp_amp c = uni(m), yp = uni(m + 1);
    const sp_amp dp = make_double2(c.x + yp.x, c.y + yp.y), dm = make_double2(c.x + yp.x * -1.0, c.y + yp.y * -1.0);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const sp_u32 j = jt | jr[r];
        v[r] = cmul(v[r], ((__popc(j & zl) ^ gp) & 1u) ? dm : dp);
    }
}

// ---- memory ------------------------------------------------------------------------

typedef double sp_dbl2 __attribute__((ext_vector_type(2)));

template <bool NT>
SP_INLINE sp_amp ld(const sp_amp* p) {
    if (NT) {
        const sp_dbl2 x = __builtin_nontemporal_load(reinterpret_cast<const sp_dbl2*>(p));
        return make_double2(x.x, x.y);
    }
    return *p;
}

template <bool NT>
SP_INLINE void st(sp_amp* p, sp_amp v) {
    if (NT) {
        sp_dbl2 x;
        x.x = v.x;
        x.y = v.y;
        __builtin_nontemporal_store(x, reinterpret_cast<sp_dbl2*>(p));
    } else {
        *p = v;
    }
}

// Tile number of workgroup b among `count`: every XCD (workgroups go to the
// 8 XCDs round-robin) walks a contiguous range of tile numbers.
SP_INLINE sp_u32 xcd_tile(sp_u32 b, sp_u32 count) {
    if (count < 8) return b;
    return (b & 7u) * (count >> 3) + (b >> 3);
}

}  // namespace sp
