// CPU check of the register-tile program builder (build_rtile).
//
// Random tile-op sequences are turned into register programs; the program
// is then executed on the host exactly as k_rtile reads it (op words,
// per-phase layouts, the LDS slots of both half-tile transpose rounds,
// header decoding) and compared with the tile ops applied directly in
// tile-index space. Also checks the global load/store offsets of the first
// and last layouts.
// Exit status 0 and a line "ok <cases>" when every case agrees.
#include "tile.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

using lqcs_dev::amp_t;
using lqcs_dev::RegOpHdr;
using lqcs_dev::RPhase;
using lqcs_dev::RTileProgram;
using lqcs_dev::TileOp;
using cd = std::complex<double>;

namespace {

// Tile width of the current case: 12 (four-wave tiles) or 10 (wave tiles);
// both transpose in two half-tile rounds.
int kL = 12;
uint32_t kN = 1u << 12;
uint32_t kT = 256;  // threads per tile

uint64_t pdep(uint64_t v, uint64_t mask) {
    uint64_t r = 0;
    for (uint64_t bit = 1; mask; bit <<= 1) {
        const uint64_t low = mask & (~mask + 1);
        if (v & bit) r |= low;
        mask ^= low;
    }
    return r;
}

cd C(amp_t a) { return {a.x, a.y}; }
amp_t A(cd z) { return make_double2(z.real(), z.imag()); }
int parity(uint64_t v) { return __builtin_popcountll(v) & 1; }

// Tile ops applied directly: amplitude j of the tile sits at global index
// base | pdep(j, smask).
void apply_direct(std::vector<cd>& a, const TileOp& t, uint64_t base) {
    std::vector<cd> o = a;
    const bool gon = (base & t.gctrl) == t.gctrl;
    if (!gon && t.kind != lqcs_dev::TK_PAULI) return;
    for (uint32_t j = 0; j < kN; ++j) {
        switch (t.kind) {
            case lqcs_dev::TK_DENSE1: {
                if ((j & t.lctrl) != t.lctrl) break;
                const uint32_t b = 1u << t.t0;
                const uint32_t j0 = j & ~b, j1 = j | b;
                const int row = (j & b) ? 1 : 0;
                o[j] = C(t.m[2 * row]) * a[j0] + C(t.m[2 * row + 1]) * a[j1];
                break;
            }
            case lqcs_dev::TK_DENSE2: {
                if ((j & t.lctrl) != t.lctrl) break;
                const uint32_t b0 = 1u << t.t0, b1 = 1u << t.t1, z = j & ~(b0 | b1);
                const int row = ((j & b0) ? 1 : 0) | ((j & b1) ? 2 : 0);
                const uint32_t idx[4] = {z, z | b0, z | b1, z | b0 | b1};
                cd acc = 0;
                for (int c = 0; c < 4; ++c) acc += C(t.m[4 * row + c]) * a[idx[c]];
                o[j] = acc;
                break;
            }
            case lqcs_dev::TK_DIAG: {
                if ((j & t.lctrl) != t.lctrl) break;
                const int sel = (((j & t.dl0) || (base & t.dg0)) ? 1 : 0) | (((j & t.dl1) || (base & t.dg1)) ? 2 : 0);
                o[j] = a[j] * C(t.m[sel]);
                break;
            }
            default: {
                auto sgn = [&](uint32_t k) { return (parity(k & t.zl) ^ parity(base & t.zg)) ? -1.0 : 1.0; };
                o[j] = C(t.m[0]) * a[j] + C(t.m[1]) * sgn(j ^ t.xl) * a[j ^ t.xl];
                break;
            }
        }
    }
    a.swap(o);
}

int fail(const char* what, int c) {
    std::printf("FAIL case %d: %s\n", c, what);
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    const int cases = argc > 1 ? std::atoi(argv[1]) : 400;
    std::mt19937_64 rng(20240611);
    std::uniform_real_distribution<double> U(-1, 1);
    int total_phases = 0, total_transposes = 0;
    size_t merged_total = 0;
    for (int c = 0; c < cases; ++c) {
        kL = (c % 2) ? 10 : 12;
        kN = 1u << kL;
        kT = kN >> 4;
        const uint32_t nq = kL + rng() % 9;
        // Tile set: `low` contiguous low qubits plus random others.
        const uint32_t low = rng() % 7;
        uint64_t smask = (uint64_t{1} << low) - 1;
        while (__builtin_popcountll(smask) < kL) smask |= uint64_t{1} << (rng() % nq);
        const uint64_t outside = ((uint64_t{1} << nq) - 1) & ~smask;
        const uint32_t tile = static_cast<uint32_t>(rng() & ((uint64_t{1} << (nq - kL)) - 1));
        const uint64_t base = pdep(tile, outside);
        auto outside_bit = [&]() -> uint64_t {
            if (!outside) return 0;
            uint64_t b;
            do b = uint64_t{1} << (rng() % nq); while (!(b & outside));
            return b;
        };
        // Ops: long sequences now and then (long phases, many op words).
        const int n_ops = (c % 10 == 9) ? 300 + static_cast<int>(rng() % 200) : 1 + static_cast<int>(rng() % 60);
        // Every tenth case keeps its dense ops on four tile bits, so one
        // layout collects more op words than one staging segment holds.
        const bool narrow = c % 10 == 4;
        uint32_t four[4];
        for (int e = 0; e < 4; ++e) {
            bool again;
            do {
                four[e] = rng() % kL;
                again = false;
                for (int f = 0; f < e; ++f) again |= four[f] == four[e];
            } while (again);
        }
        auto pick_bit = [&]() { return narrow ? four[rng() % 4] : static_cast<uint32_t>(rng() % kL); };
        std::vector<TileOp> ops;
        // Every tenth case is layers of dense 2x2 ops and singly-controlled X
        // (controls on register, thread or outside bits), which the builder
        // merges into select ops.
        const bool layered = c % 10 == 6;
        for (int k = 0; layered && k < 160; ++k) {
            TileOp t;
            std::memset(&t, 0, sizeof t);
            t.member = -1;
            t.kind = lqcs_dev::TK_DENSE1;
            t.t0 = static_cast<uint32_t>(rng() % kL);
            const int two = static_cast<int>(rng() % 6);
            if (two == 1 || two == 2) {  // controlled diagonal (CZ, CP or a general one)
                t.kind = lqcs_dev::TK_DIAG;
                t.dl0 = 1u << t.t0;
                t.m[0] = two == 1 ? make_double2(1, 0) : make_double2(U(rng), U(rng));
                t.m[1] = make_double2(U(rng), U(rng));
                if (rng() % 4 == 0 && outside) {
                    t.gctrl = outside_bit();
                } else {
                    uint32_t b;
                    do b = static_cast<uint32_t>(rng() % kL); while (b == t.t0);
                    t.lctrl = 1u << b;
                }
                t.t0 = 0;
            } else if (two == 0) {
                t.m[1] = make_double2(1, 0);
                t.m[2] = make_double2(1, 0);
                if (rng() % 4 == 0 && outside) {
                    t.gctrl = outside_bit();
                } else {
                    uint32_t b;
                    do b = static_cast<uint32_t>(rng() % kL); while (b == t.t0);
                    t.lctrl = 1u << b;
                }
            } else {
                for (int e = 0; e < 4; ++e) t.m[e] = make_double2(U(rng), U(rng));
            }
            ops.push_back(t);
        }
        for (int k = 0; !layered && k < (narrow ? 400 : n_ops); ++k) {
            TileOp t;
            std::memset(&t, 0, sizeof t);
            t.member = -1;  // every batch register
            const int kind = static_cast<int>(rng() % 10);
            auto rand_m = [&](int n) {
                for (int e = 0; e < n; ++e) t.m[e] = make_double2(U(rng), U(rng));
            };
            auto ctrl = [&](uint32_t used) {
                if (rng() % 3 == 0) {
                    const uint32_t b = 1u << (rng() % kL);
                    if (!(b & used)) t.lctrl |= b;
                }
                if (rng() % 4 == 0) t.gctrl |= outside_bit();
                if (rng() % 8 == 0) {
                    const uint32_t b = 1u << (rng() % kL);
                    if (!((b | t.lctrl) & used)) t.lctrl |= b;
                }
            };
            if (kind <= 3) {
                t.kind = lqcs_dev::TK_DENSE1;
                t.t0 = pick_bit();
                if (kind == 3) {  // X
                    t.m[1] = make_double2(1, 0);
                    t.m[2] = make_double2(1, 0);
                } else {
                    rand_m(4);
                }
                ctrl(1u << t.t0);
            } else if (kind <= 5) {
                t.kind = lqcs_dev::TK_DENSE2;
                t.t0 = pick_bit();
                do t.t1 = pick_bit(); while (t.t1 == t.t0);
                rand_m(16);
                if (rng() % 2)  // real entries (a real channel's superoperator)
                    for (int e = 0; e < 16; ++e) t.m[e].y = 0.0;
                ctrl((1u << t.t0) | (1u << t.t1));
            } else if (kind <= 7) {
                t.kind = lqcs_dev::TK_DIAG;
                rand_m(4);
                if (rng() % 3 == 0) t.m[0] = make_double2(1, 0);  // diag(1, p): CZ, CP, P, T
                if (rng() % 3) t.dl0 = 1u << (rng() % kL); else t.dg0 = outside_bit();
                if (rng() % 2) {
                    if (rng() % 3) {
                        do t.dl1 = 1u << (rng() % kL); while (t.dl1 == t.dl0);
                    } else {
                        if (__builtin_popcountll(outside) >= 2) {
                            do t.dg1 = outside_bit(); while (t.dg1 == t.dg0);
                        }
                    }
                }
                ctrl(t.dl0 | t.dl1);
            } else {
                t.kind = lqcs_dev::TK_PAULI;
                const int nx = static_cast<int>(rng() % 4);
                for (int e = 0; e < nx; ++e) t.xl |= 1u << pick_bit();
                for (int e = 0; e < 3; ++e)
                    if (rng() % 2) t.zl |= 1u << (rng() % kL);
                t.zl &= ~t.xl;
                if (rng() % 2) t.zg = outside_bit();
                rand_m(2);
            }
            ops.push_back(t);
        }
        std::vector<cd> init(kN);
        for (uint32_t j = 0; j < kN; ++j) init[j] = cd(U(rng), U(rng));
        const char* why = nullptr;
        int n_ph = 0, n_tr = 0;
        auto check = [&](const std::vector<TileOp>& ops) -> double {
            RTileProgram prog;
            if (!lqcs_dev::build_rtile(ops, kL, smask, nq, prog)) { why = "build_rtile refused"; return -1; }
            if (prog.phases.empty()) { why = "no phases"; return -1; }
            n_ph = static_cast<int>(prog.phases.size());
            n_tr = 0;
            std::vector<cd> want = init, a = init;
            for (const TileOp& t : ops) apply_direct(want, t, base);
            for (const RPhase* ph : {&prog.phases.front(), &prog.phases.back()}) {
                for (uint32_t t = 0; t < kT; ++t)
                    for (int r = 0; r < 16; ++r) {
                        const uint32_t j = static_cast<uint32_t>(pdep(t, ph->tmask)) | ph->jr[r];
                        if ((pdep(t, ph->tmask_global) | ph->gofs[r]) != pdep(j, smask)) { why = "gofs"; return -1; }
                    }
                const uint32_t run = static_cast<uint32_t>(__builtin_ctzll(~smask));
                const uint32_t avoid = (1u << std::min<uint32_t>(run, 6)) - 1;
                if ((~ph->tmask & (kN - 1)) & avoid) { why = "edge layout keeps low qubits in registers"; return -1; }
            }
            size_t seen_ops = 0;
            for (uint64_t t : {uint64_t{0}, uint64_t{1}, uint64_t{tile}, (outside >> 12) ^ tile, uint64_t{0x2aaaa}})
                if (lqcs_dev::rtile_tile_base(t & ((uint64_t{1} << (nq - kL)) - 1), prog.phases[0], outside) !=
                    pdep(t & ((uint64_t{1} << (nq - kL)) - 1), outside)) { why = "tile base"; return -1; }
            for (size_t p = 0; p < prog.phases.size(); ++p) {
                const RPhase& ph = prog.phases[p];
                for (uint32_t t = 0; t < kT; ++t) {
                    if (lqcs_dev::rtile_thread_bits(t, ph) != pdep(t, ph.tmask)) { why = "thread bits"; return -1; }
                    if (lqcs_dev::rtile_thread_global(t, ph) != pdep(t, ph.tmask_global)) { why = "thread global"; return -1; }
                }
                if (__builtin_popcount(ph.tmask) != kL - 4 || (ph.tmask & ~(kN - 1))) { why = "tmask"; return -1; }
                if (p > 0 && !ph.transpose && ph.tmask != prog.phases[p - 1].tmask) { why = "layout change without transpose"; return -1; }
                if (p > 0 && ph.transpose) {
                    ++n_tr;
                    // The two half-tile rounds through LDS, slot by slot: every
                    // slot written once per round, and every read finds the
                    // amplitude of the tile index it stands for.
                    const RPhase& pp = prog.phases[p - 1];
                    const uint32_t h = ph.h;
                    if (h >= 12 || !(pp.tmask & ph.tmask & (1u << h))) { why = "split bit is not a thread bit of both layouts"; return -1; }
                    for (uint32_t half = 0; half < 2; ++half) {
                        std::vector<int64_t> owner(kN / 2, -1);
                        for (uint32_t t = 0; t < kT; ++t) {
                            const uint32_t jp = static_cast<uint32_t>(pdep(t, pp.tmask));
                            if (((jp >> h) & 1u) != half) continue;
                            for (int r = 0; r < 16; ++r) {
                                const uint32_t slot = lqcs_dev::rtile_slot(jp, h) ^ (ph.cs_prev[r] >> 4);
                                if (ph.cs_prev[r] & 15u) { why = "LDS register offset not in bytes"; return -1; }
                                if (slot >= kN / 2 || owner[slot] != -1) { why = "LDS slot collision"; return -1; }
                                owner[slot] = jp | pp.jr[r];
                            }
                        }
                        for (uint32_t t = 0; t < kT; ++t) {
                            const uint32_t jc = static_cast<uint32_t>(pdep(t, ph.tmask));
                            if (((jc >> h) & 1u) != half) continue;
                            for (int r = 0; r < 16; ++r) {
                                const uint32_t slot = lqcs_dev::rtile_slot(jc, h) ^ (ph.cs_cur[r] >> 4);
                                if (slot >= kN / 2 || owner[slot] != int64_t(jc | ph.jr[r])) { why = "LDS read of the wrong amplitude"; return -1; }
                            }
                        }
                    }
                }
                uint32_t off = ph.op_off;
                const size_t total = prog.words.size();
                for (uint32_t k = 0; k < ph.n_ops; ++k) {
                    if (off + 2 > total) { why = "header outside the op words"; return -1; }
                    RegOpHdr h;
                    std::memcpy(&h, &prog.words[off], sizeof h);
                    const uint32_t kind = h.code & 15u, b0 = (h.code >> 4) & 3u, b1 = (h.code >> 6) & 3u;
                    const uint32_t xr = (h.code >> 8) & 15u, rctl = (h.code >> 12) & 15u, nm = (h.code >> 16) & 0xFFFu;
                    if (off + 2 + nm > total) { why = "matrix outside the op words"; return -1; }
                    if (((h.code & lqcs_dev::kCodeFree) != 0) != (rctl == 0 && h.tctl == 0)) {
                        why = "free flag disagrees with the controls";
                        return -1;
                    }
                    if (((h.code & lqcs_dev::kCodeRestricted) != 0) != (h.gctrl != 0 || h.member != 0)) {
                        why = "restricted flag disagrees with the header";
                        return -1;
                    }
                    const amp_t* m = &prog.words[off + 2];
                    std::vector<cd> o = a;
                    if ((tile & h.gctrl) == h.gctrl) {
                        for (uint32_t t = 0; t < kT; ++t) {
                            const uint32_t jt = static_cast<uint32_t>(pdep(t, ph.tmask));
                            const bool tp = (jt & h.tctl) == h.tctl;
                            auto J = [&](int r) { return jt | ph.jr[r]; };
                            for (int r = 0; r < 16; ++r) {
                                const bool on = tp && ((uint32_t(r) & rctl) == rctl);
                                switch (kind) {
                                    case lqcs_dev::RK_DENSE1: {
                                        const int row = (r >> b0) & 1, r0 = r & ~(1 << b0), r1 = r | (1 << b0);
                                        const bool on0 = tp && ((uint32_t(r0) & rctl) == rctl);
                                        if (on0) o[J(r)] = C(m[2 * row]) * a[J(r0)] + C(m[2 * row + 1]) * a[J(r1)];
                                        break;
                                    }
                                    case lqcs_dev::RK_X:
                                        if (on) o[J(r)] = a[J(r ^ (1 << b0))];
                                        break;
                                    case lqcs_dev::RK_SEL1: {
                                        const bool s = xr ? ((r >> (xr - 1)) & 1) != 0
                                                          : h.la ? (J(r) & h.la) != 0 : (tile & h.ga) != 0;
                                        const amp_t* ms = s ? m + 4 : m;
                                        const int row = (r >> b0) & 1, r0 = r & ~(1 << b0), r1 = r | (1 << b0);
                                        o[J(r)] = C(ms[2 * row]) * a[J(r0)] + C(ms[2 * row + 1]) * a[J(r1)];
                                        break;
                                    }
                                    case lqcs_dev::RK_DENSE2R:
                                        if (rctl || h.tctl) {
                                            why = "DENSE2R with controls";
                                            return -1;
                                        }
                                        for (int e = 0; e < 16; ++e)
                                            if (m[e].y != 0.0) {
                                                why = "DENSE2R with a complex entry";
                                                return -1;
                                            }
                                        [[fallthrough]];
                                    case lqcs_dev::RK_DENSE2: {
                                        const int z = r & ~((1 << b0) | (1 << b1));
                                        const bool onz = tp && ((uint32_t(z) & rctl) == rctl);
                                        const int row = ((r >> b0) & 1) | (((r >> b1) & 1) << 1);
                                        const int idx[4] = {z, z | (1 << b0), z | (1 << b1), z | (1 << b0) | (1 << b1)};
                                        if (onz) {
                                            cd acc = 0;
                                            for (int e = 0; e < 4; ++e) acc += C(m[4 * row + e]) * a[J(idx[e])];
                                            o[J(r)] = acc;
                                        }
                                        break;
                                    }
                                    case lqcs_dev::RK_DIAG: {
                                        const uint32_t g = ((tile & h.ga) ? 1u : 0u) | ((tile & h.gb) ? 2u : 0u);
                                        const uint32_t sel = g | ((J(r) & h.la) ? 1u : 0u) | ((J(r) & h.lb) ? 2u : 0u);
                                        if (on) o[J(r)] = a[J(r)] * C(m[sel]);
                                        break;
                                    }
                                    case lqcs_dev::RK_DIAG1: {
                                        const bool hi = xr ? ((r >> (xr - 1)) & 1) != 0
                                                           : h.la ? (J(r) & h.la) != 0 : (tile & h.ga) != 0;
                                        if (b1 && !(m[0].x == 1.0 && m[0].y == 0.0)) {
                                            why = "DIAG1 marked d0 = 1 with another d0";
                                            return -1;
                                        }
                                        if (on) o[J(r)] = a[J(r)] * C(m[hi ? 1 : 0]);
                                        break;
                                    }
                                    case lqcs_dev::RK_PAULI: {
                                        const int gp = __builtin_popcount(tile & h.ga) & 1;
                                        const int r2 = r ^ static_cast<int>(xr);
                                        const double s2 = ((__builtin_popcount(J(r2) & h.la) ^ gp) & 1) ? -1.0 : 1.0;
                                        o[J(r)] = C(m[0]) * a[J(r)] + C(m[1]) * s2 * a[J(r2)];
                                        break;
                                    }
                                    default:
                                        why = "unknown kind";
                                        return -1;
                                }
                            }
                        }
                    }
                    a.swap(o);
                    off += 2 + nm;
                    ++seen_ops;
                }
            }
            if (seen_ops + prog.merged != ops.size()) { why = "op count"; return -1; }
            merged_total += prog.merged;
            // Relative to the largest amplitude: the random matrices are not
            // unitary, so magnitudes (and rounding) grow along the sequence.
            double err = 0, scale = 0;
            for (uint32_t j = 0; j < kN; ++j) {
                err = std::max(err, std::abs(a[j] - want[j]));
                scale = std::max(scale, std::abs(want[j]));
            }
            return err / std::max(scale, 1e-300);
        };
        const double err = check(ops);
        total_phases += n_ph;
        total_transposes += n_tr;
        if (!(err >= 0 && err <= 1e-11)) {
            std::printf("FAIL case %d: %s error %g (nq=%u smask=%llx ops=%zu phases=%d transposes=%d)\n", c,
                        why ? why : "values", err, nq, static_cast<unsigned long long>(smask), ops.size(), n_ph, n_tr);
            size_t lo = 1, hi = ops.size();  // smallest failing prefix
            while (lo < hi) {
                const size_t mid = (lo + hi) / 2;
                const double e = check(std::vector<TileOp>(ops.begin(), ops.begin() + mid));
                if (e >= 0 && e <= 1e-11) lo = mid + 1; else hi = mid;
            }
            const double e = check(std::vector<TileOp>(ops.begin(), ops.begin() + lo));
            std::printf("  smallest failing prefix %zu ops: %s error %g phases=%d transposes=%d\n", lo,
                        why ? why : "values", e, n_ph, n_tr);
            return 1;
        }
    }
    std::printf("ok %d phases=%d transposes=%d merged=%zu\n", cases, total_phases, total_transposes, merged_total);
    return 0;
}
