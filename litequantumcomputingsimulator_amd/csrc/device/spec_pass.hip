// Specialized register-tile passes: kernel text generated per pass structure
// and compiled at run time with hipRTC.
//
// The interpreting kernel (rtile.hip) decodes every op of a pass at run time
// and keeps two register sets, because every case of its op dispatch must
// leave the tile in the same registers. A specialized pass knows each op's
// kind and register bits when it is compiled: it calls the op templates of
// spec_pass.hpp in program order on one register set updated in place. That
// removes the dispatch (scalar decode and a branch chain per op) and about 40
// VGPRs, so four waves per SIMD fit instead of three.
//
// Tiles of 12 qubits (256 threads) or 13 (512 threads; the interpreting
// kernels stop at 12): with 13, a plan can keep five contiguous low qubits in
// every tile (512-byte runs in HBM) and still have eight free tile qubits.
//
// In-place transposes. The tile moves through LDS (32 or 64 KiB) in two halves,
// split by a tile bit h that is a thread bit of both layouts. A thread writes
// its registers in the round of its old half and reads in the round of its
// new half; in place, no thread may read before it has written, so h must sit
// at the same thread-id bit in both layouts. Every layout orders its thread
// bits so that the split bits of the transposes into and out of it sit at
// the top two thread-id bits (wave bits), alternating so the two never collide;
// the other thread bits fill the remaining positions in ascending order,
// which keeps the tile's contiguous low qubits on the lanes of the load and
// store layouts. When two layouts share no thread bit but such low qubits,
// the split is a low qubit at its own lane position. Each transpose has its
// own linear LDS slot map, conflict-free for the lanes of both layouts.
//
// The generated kernel reads its matrix entries from the pass's op words (the
// same words the interpreting kernel reads), so passes that differ only in
// angles share one kernel. Kernels are cached by source text for the life of
// the process; distinct sources of one plan compile in parallel, in helper
// processes (lqcs_rtc_worker) where it is installed beside the library.
#include "common.hpp"
#include "engine.hpp"
#include "spec_pass.hpp"
#include "tile.hpp"

#include "spec_compile.hpp"

#include <dlfcn.h>
#include <fcntl.h>
#include <poll.h>
#include <pthread.h>
#include <spawn.h>
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <csignal>
#include <array>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <future>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

extern char** environ;

namespace lqcs_dev {

// Specialized passes for the calls that ask for them (tile kernel option).
bool g_rtile_spec = false;

namespace {

// The device templates, embedded at build time (spec_pass_hpp.inc holds the
// text of spec_pass.hpp as a raw string literal).
const char* const kSpecHeader =
#include "spec_pass_hpp.inc"
    ;

// Tiles of L = 12 or 13 qubits: 16 registers per thread, T = L - 4 thread
// bits (256 or 512 threads).
constexpr uint32_t kLMax = 13;
constexpr uint32_t kTMax = 9;

std::string hex(uint64_t v, bool wide = true) {
    char buf[32];
    std::snprintf(buf, sizeof buf, wide ? "0x%llxull" : "0x%llxu", static_cast<unsigned long long>(v));
    return buf;
}

std::string dec(uint64_t v) { return std::to_string(v); }

// Expression for sum_i ((x >> i) & 1) << dst[i] over the entries of dst
// (dst[i] < 0: bit i unused), grouping runs of consecutive bits.
std::string scatter_expr(const std::string& x, const std::vector<int>& dst, bool wide) {
    std::string out;
    const std::string cast = wide ? "(sp_u64)" : "";
    for (size_t i = 0; i < dst.size();) {
        if (dst[i] < 0) {
            ++i;
            continue;
        }
        size_t j = i + 1;
        while (j < dst.size() && dst[j] == dst[j - 1] + 1) ++j;
        const uint64_t mask = ((uint64_t{1} << (j - i)) - 1) << i;
        const int shift = dst[i] - static_cast<int>(i);
        std::string term = "(" + cast + "(" + x + " & " + hex(mask, wide) + ")";
        if (shift > 0) term += " << " + dec(static_cast<uint64_t>(shift));
        if (shift < 0) term += " >> " + dec(static_cast<uint64_t>(-shift));
        term += ")";
        out += (out.empty() ? "" : " | ") + term;
        i = j;
    }
    return out.empty() ? std::string(wide ? "(sp_u64)0" : "0u") : out;
}

// Register bits (ascending) of a phase, read off its register tile indices.
void reg_bits(const RPhase& ph, int (&rb)[4]) {
    for (int i = 0; i < 4; ++i) rb[i] = __builtin_ctz(ph.jr[1 << i]);
}

struct Layout {
    int rb[4];
    uint32_t tmask;
    uint32_t nt;            // thread bits (L - 4)
    int tpos[kTMax];        // tile bit at each thread-id bit
    uint32_t jr[16];        // tile index bits of register r
};

// The pass's layouts with thread-bit orders for in-place transposes, and the
// split bit of each transpose (split[k]: into layout k; -1 none) with its
// thread-id position (spos[k], the same in both layouts). `run` is the number
// of the tile's contiguous low qubits (tile bits 0..run-1): in the load and
// store layouts they stay on lanes 0..run-1, in order, so that each group of
// 2^run lanes reads or writes one contiguous run of amplitudes.
//
// A split bit goes to thread-id bit 7 or 6 (a wave bit: whole waves take
// turns), alternating so that the splits into and out of a layout never share
// a position. When the only thread bits the two layouts share are low bits of
// an edge layout, the split is such a bit at its own lane position; the two
// rounds then run on half the lanes of every wave.
bool plan_layouts(const RTileProgram& rp, uint32_t L, uint32_t run, std::vector<Layout>& lay, std::vector<int>& split,
                  std::vector<int>& spos) {
    const size_t P = rp.phases.size();
    const uint32_t kT = L - 4;
    lay.assign(P, Layout{});
    split.assign(P, -1);
    spos.assign(P, -1);
    const uint32_t all = (1u << L) - 1;
    run = std::min<uint32_t>(run, 6);
    for (size_t k = 0; k < P; ++k) {
        reg_bits(rp.phases[k], lay[k].rb);
        uint32_t rm = 0;
        for (int i = 0; i < 4; ++i) rm |= 1u << lay[k].rb[i];
        lay[k].tmask = all & ~rm;
        lay[k].nt = kT;
        for (int r = 0; r < 16; ++r) lay[k].jr[r] = rp.phases[k].jr[r];
    }
    // pins[k][p]: tile bit fixed at thread-id position p of layout k (-1 free)
    std::vector<std::array<int, kTMax>> pins(P);
    for (auto& a : pins) a.fill(-1);
    auto edge = [&](size_t k) { return k == 0 || k == P - 1; };
    // A layout without a transpose of its own repeats the previous one: it
    // shares that layout's pins (handled by copying orders below).
    int wave_pos = static_cast<int>(kT) - 1;  // the top two thread-id bits, alternating
    const int wave_sum = 2 * static_cast<int>(kT) - 3;
    for (size_t k = 1; k < P; ++k) {
        if (!rp.phases[k].transpose) continue;
        size_t prev = k - 1;
        while (prev > 0 && !rp.phases[prev].transpose) --prev;  // the layout in force before k
        const uint32_t common = lay[prev].tmask & lay[k].tmask;
        auto pinned_bit = [&](size_t l, int bit) {
            for (int p = 0; p < static_cast<int>(kT); ++p)
                if (pins[l][p] == bit) return true;
            return false;
        };
        auto low_in_edge = [&](int bit) {
            return bit < static_cast<int>(run) && (edge(prev) || edge(k));
        };
        // Candidates: not pinned in the old layout already; high bits first
        // (highest first), then low bits.
        std::vector<int> cand;
        for (int b = static_cast<int>(L) - 1; b >= 0; --b)
            if ((common >> b & 1u) && !pinned_bit(prev, b) && !low_in_edge(b)) cand.push_back(b);
        for (int b = static_cast<int>(L) - 1; b >= 0; --b)
            if ((common >> b & 1u) && !pinned_bit(prev, b) && low_in_edge(b)) cand.push_back(b);
        bool placed = false;
        for (int h : cand) {
            int positions[2] = {wave_pos, wave_sum - wave_pos};
            int np = 2;
            if (low_in_edge(h)) {
                positions[0] = h;
                np = 1;
            }
            for (int i = 0; i < np && !placed; ++i) {
                const int p = positions[i];
                if (pins[prev][p] >= 0 || pins[k][p] >= 0) continue;
                // keep the edge layouts' low lanes for their own bits
                if (p < static_cast<int>(run) && p != h && (edge(prev) || edge(k))) continue;
                pins[prev][p] = h;
                pins[k][p] = h;
                split[k] = h;
                spos[k] = p;
                placed = true;
                if (p == wave_pos) wave_pos = wave_sum - wave_pos;
            }
            if (placed) break;
        }
        if (!placed) return false;
    }
    // Thread-bit orders: pinned bits, the rest ascending in the free positions.
    for (size_t k = 0; k < P; ++k) {
        if (k > 0 && !rp.phases[k].transpose) {
            std::memcpy(lay[k].tpos, lay[k - 1].tpos, sizeof lay[k].tpos);
            continue;
        }
        uint32_t used = 0;
        for (uint32_t p = 0; p < kT; ++p)
            if (pins[k][p] >= 0) used |= 1u << pins[k][p];
        uint32_t rest = lay[k].tmask & ~used;
        for (uint32_t p = 0; p < kT; ++p) {
            if (pins[k][p] >= 0) {
                lay[k].tpos[p] = pins[k][p];
                continue;
            }
            if (!rest) return false;
            lay[k].tpos[p] = __builtin_ctz(rest);
            rest &= rest - 1;
        }
    }
    return true;
}

std::string jt_expr(const Layout& l, const char* tid = "t") {
    std::vector<int> dst(l.nt);
    for (uint32_t p = 0; p < l.nt; ++p) dst[p] = l.tpos[p];
    return scatter_expr(tid, dst, false);
}

// Global index bits of the thread in layout l (tile bit b is qubit gq[b]).
std::string tglobal_expr(const Layout& l, const int (&gq)[kLMax]) {
    std::vector<int> dst(l.nt);
    for (uint32_t p = 0; p < l.nt; ++p) dst[p] = gq[l.tpos[p]];
    return scatter_expr("tid", dst, true);
}

uint64_t reg_global(const Layout& l, int r, const int (&gq)[kLMax]) {
    uint64_t g = 0;
    for (int i = 0; i < 4; ++i)
        if (r & (1 << i)) g |= uint64_t{1} << gq[l.rb[i]];
    return g;
}

// Global loads (or stores) of the tile in layout l. When the thread's part
// of the byte offset fits 32 bits, each register's address is a scalar base
// (the tile's base plus the register's constant offset) and one 32-bit
// per-lane offset, the form the load and store instructions take directly
// (no 64-bit address arithmetic per lane); otherwise 64-bit per-lane
// addresses.
std::string edge_access(const Layout& l, const int (&gq)[kLMax], bool load, const std::string& NT) {
    uint64_t tmax = 0;
    for (uint32_t p = 0; p < l.nt; ++p) tmax |= uint64_t{1} << gq[l.tpos[p]];
    std::string s;
    if ((tmax * sizeof(amp_t)) >> 32 == 0) {
        std::vector<int> dst(l.nt);
        for (uint32_t p = 0; p < l.nt; ++p) dst[p] = gq[l.tpos[p]] + 4;  // bytes
        s += "        const sp_u32 o = " + scatter_expr("tid", dst, false) + ";\n";
        s += "        char* const pb = reinterpret_cast<char*>(b);\n";
        for (int r = 0; r < 16; ++r) {
            const std::string a = "(__attribute__((address_space(1))) sp_amp*)(sp::sbase(pb + " +
                                  hex(reg_global(l, r, gq) * sizeof(amp_t)) + ") + o)";
            if (load) s += "        v[" + dec(r) + "] = sp::ld<" + NT + ">(" + a + ");\n";
            else s += "        sp::st<" + NT + ">(" + a + ", v[" + dec(r) + "]);\n";
        }
        return s;
    }
    s += "        sp_amp* const p = b + (" + tglobal_expr(l, gq) + ");\n";
    for (int r = 0; r < 16; ++r) {
        const std::string a = "p + " + hex(reg_global(l, r, gq));
        if (load) s += "        v[" + dec(r) + "] = sp::ld<" + NT + ">(" + a + ");\n";
        else s += "        sp::st<" + NT + ">(" + a + ", v[" + dec(r) + "]);\n";
    }
    return s;
}

// LDS slots of one half-tile transpose split by tile bit h. The index of
// tile index j is c(j), j without bit h (L - 1 bits); its slot is M c(j) for an
// invertible matrix M over GF(2) chosen per transpose so that the 16 lanes
// of a writing or reading group (thread-id bits 0..3 of the old and of the
// new layout) fall on 16 distinct slots modulo 16: 16-byte accesses then hit
// distinct banks. Rows 0..3 of M: row b picks the old layout's lane bit b,
// plus one new-layout lane bit that is not an old lane bit, placed in a row
// that no shared lane bit occupies, so that both 4x4 blocks are permutations;
// rows 4 and up: the bits that are not old lane bits, in order. M is linear, so
// slot(jt | jr) = slot(jt) ^ slot(jr): the thread's part is computed once
// per layout, the register's part is a constant.
struct SlotMap {
    uint32_t n = 11;              // index bits (L - 1)
    uint32_t col[kLMax - 1] = {};  // M's column of each index bit
    uint32_t cidx(int tile_bit, int h) const { return static_cast<uint32_t>(tile_bit < h ? tile_bit : tile_bit - 1); }
    uint32_t bytes(uint32_t j, int h) const {
        uint32_t c = (j & ((1u << h) - 1u)) | ((j >> (h + 1)) << h), s = 0;
        for (uint32_t x = 0; x < n; ++x)
            if (c & (1u << x)) s ^= col[x];
        return s << 4;
    }
};

SlotMap slot_map(const Layout& prev, const Layout& cur, int h, uint32_t L) {
    SlotMap sm;
    sm.n = L - 1;
    const int nb = static_cast<int>(sm.n);
    uint32_t row[kLMax - 1];
    // The first four lane bits of each layout other than the split bit (a
    // split among the lanes halves the active lanes of each group).
    int W[4], R[4];
    for (int b = 0, p = 0; b < 4; ++p)
        if (prev.tpos[p] != h) W[b++] = static_cast<int>(sm.cidx(prev.tpos[p], h));
    for (int b = 0, p = 0; b < 4; ++p)
        if (cur.tpos[p] != h) R[b++] = static_cast<int>(sm.cidx(cur.tpos[p], h));
    bool taken[4] = {false, false, false, false};
    for (int b = 0; b < 4; ++b) row[b] = 1u << W[b];
    for (int j = 0; j < 4; ++j)
        for (int b = 0; b < 4; ++b)
            if (R[j] == W[b]) taken[b] = true;
    for (int j = 0; j < 4; ++j) {
        bool shared = false;
        for (int b = 0; b < 4; ++b) shared |= R[j] == W[b];
        if (shared) continue;
        for (int b = 0; b < 4; ++b)
            if (!taken[b]) {
                taken[b] = true;
                row[b] |= 1u << R[j];
                break;
            }
    }
    int n = 4;
    for (int x = 0; x < nb; ++x) {
        bool lane = false;
        for (int b = 0; b < 4; ++b) lane |= W[b] == x;
        if (!lane) row[n++] = 1u << x;
    }
    for (int x = 0; x < nb; ++x) {
        sm.col[x] = 0;
        for (int b = 0; b < nb; ++b)
            if (row[b] & (1u << x)) sm.col[x] |= 1u << b;
    }
    return sm;
}

// The thread's part of the slot byte offset in layout l: the XOR of the
// columns of its set thread-id bits (the split bit h excluded).
std::string slot_thread_expr(const Layout& l, const SlotMap& sm, int h) {
    std::string e;
    for (uint32_t p = 0; p < l.nt; ++p) {
        if (l.tpos[p] == h) continue;
        const uint32_t c = sm.col[sm.cidx(l.tpos[p], h)] << 4;
        e += (e.empty() ? "" : " ^ ") + std::string("(((t >> ") + dec(p) + ") & 1u) * " + dec(c) + "u)";
    }
    return e.empty() ? "0u" : e;
}

std::string jr_array(const Layout& l, const std::string& name) {
    std::string s = "    const sp_u32 " + name + "[16] = {";
    for (int r = 0; r < 16; ++r) s += (r ? ", " : "") + dec(l.jr[r]) + "u";
    return s + "};\n";
}

// Flags of n matrix entries, two bits each: bit 0, the imaginary part is
// exactly zero; bit 1, the entry is exactly zero.
uint32_t entry_flags(const amp_t* m, int n) {
    uint32_t f = 0;
    for (int i = 0; i < n; ++i) {
        const uint32_t e = (m[i].y == 0.0 ? 1u : 0u) | (m[i].x == 0.0 && m[i].y == 0.0 ? 2u : 0u);
        f |= e << (2 * i);
    }
    return f;
}

// One op of the pass as a call of the templates (m: its matrix entries);
// false if not handled.
bool emit_op(std::string& out, const RegOpHdr& h, uint32_t off, const amp_t* m, const std::string& jr_name) {
    const uint32_t kind = h.code & 15u, b0 = (h.code >> 4) & 3u, b1 = (h.code >> 6) & 3u, xr = (h.code >> 8) & 15u,
                   rctl = (h.code >> 12) & 15u;
    const std::string M = "w + " + dec(off);
    const std::string tp = h.tctl ? "((jt & " + dec(h.tctl) + "u) == " + dec(h.tctl) + "u)" : "true";
    std::string call;
    switch (kind) {
        case RK_DENSE1: {
            const std::string F = dec(entry_flags(m, 4)) + "u";
            if (rctl == 0 && h.tctl == 0) call = "sp::d1<" + dec(b0) + ", " + F + ">(v, " + M + ");";
            else call = "sp::d1_c<" + dec(b0) + ", " + dec(rctl) + ", " + F + ">(v, " + M + ", " + tp + ");";
            break;
        }
        case RK_SEL1: {
            if (rctl || h.tctl) return false;
            const uint32_t f0 = entry_flags(m, 4), f1 = entry_flags(m + 4, 4);
            const std::string both = dec(f0 & f1) + "u";
            auto same = [](amp_t a, amp_t b) { return a.x == b.x && a.y == b.y; };
            const bool cols = same(m[4], m[1]) && same(m[5], m[0]) && same(m[6], m[3]) && same(m[7], m[2]);
            const bool rows = same(m[4], m[2]) && same(m[5], m[3]) && same(m[6], m[0]) && same(m[7], m[1]);
            if (xr) {
                call = "sp::sel_r<" + dec(b0) + ", " + dec(xr - 1) + ", " + dec(f0 | (f1 << 8)) + "u>(v, " + M + ");";
            } else if (h.la) {
                const std::string sel = "(jt & " + dec(h.la) + "u) != 0";
                if (cols || rows)
                    call = "sp::sel_swap<" + dec(b0) + ", " + dec(f0) + "u, " + (cols ? "true" : "false") + ">(v, " + M +
                           ", " + sel + ");";
                else
                    call = "sp::sel_l<" + dec(b0) + ", " + both + ">(v, " + M + ", " + sel + ");";
            } else {
                // the select bit is a tile-number bit: one matrix per tile
                call = "sp::d1<" + dec(b0) + ", " + both + ">(v, " + M + " + ((tile & " + dec(h.ga) + "u) ? 4 : 0));";
            }
            break;
        }
        case RK_X:
            call = "sp::x_c<" + dec(b0) + ", " + dec(rctl) + ", " + (h.tctl ? "false" : "true") + ">(v, " + tp + ");";
            break;
        case RK_DENSE2:
        case RK_DENSE2R:
            call = "sp::d2<" + dec(b0) + ", " + dec(b1) + ", " + dec(rctl) + ", " + dec(entry_flags(m, 16)) + "u>(v, " + M +
                   ", " + tp + ");";
            break;
        case RK_DIAG1:
            if (xr) {
                call = "sp::dg1_r<" + dec(xr - 1) + ", " + (b1 ? "true" : "false") + ", " + dec(rctl) + ">(v, " + M + ", " +
                       tp + ");";
            } else {
                const std::string s = h.la ? "((jt & " + dec(h.la) + "u) != 0)" : "((tile & " + dec(h.ga) + "u) != 0)";
                call = "sp::dg1_u<" + dec(rctl) + ">(v, " + M + ", " + s + ", " + tp + ");";
            }
            break;
        case RK_DIAG: {
            const std::string g = "(((tile & " + dec(h.ga) + "u) ? 1u : 0u) | ((tile & " + dec(h.gb) + "u) ? 2u : 0u))";
            call = "sp::dg_gen<" + dec(rctl) + ">(v, " + M + ", jt, " + jr_name + ", " + dec(h.la) + "u, " + dec(h.lb) +
                   "u, " + g + ", " + tp + ");";
            break;
        }
        case RK_PAULI: {
            const std::string gp = "(__popc(tile & " + dec(h.ga) + "u) & 1u)";
            if (xr == 0) call = "sp::pauli_z(v, " + M + ", jt, " + jr_name + ", " + dec(h.la) + "u, " + gp + ");";
            else call = "sp::pauli_x<" + dec(xr) + ">(v, " + M + ", jt, " + jr_name + ", " + dec(h.la) + "u, " + gp + ");";
            break;
        }
        default:
            return false;
    }
    std::string cond;
    if (h.gctrl) cond = "(tile & " + dec(h.gctrl) + "u) == " + dec(h.gctrl) + "u";
    if (h.member) {
        const uint32_t mid = (h.member & 0x7fffffffu) - 1;
        const std::string m = (h.member >> 31) ? "blockIdx.y == " + dec(mid) + "u" : "blockIdx.y != " + dec(mid) + "u";
        cond = cond.empty() ? m : cond + " && " + m;
    }
    if (cond.empty()) out += "    " + call + "\n";
    else out += "    if (" + cond + ") { " + call + " }\n";
    return true;
}

// Waves per SIMD the compiler must fit: four (at most 128 VGPRs; most passes
// need 95-120, and a pass that needs more would otherwise drop to three).
// Five (96 VGPRs) spills a few registers and measured slower.
std::string waves_attr() { return " __attribute__((amdgpu_waves_per_eu(4, 8)))"; }

}  // namespace

std::string spec_source(const RTileProgram& rp, uint32_t nq, uint32_t L, uint64_t smask, bool nt) {
    if (rp.phases.empty() || (L != 12 && L != 13) || nq < L) return {};
    std::vector<Layout> lay;
    std::vector<int> split, spos;
    if (!plan_layouts(rp, L, static_cast<uint32_t>(__builtin_ctzll(~smask)), lay, split, spos)) return {};
    int gq[kLMax];
    {
        int t = 0;
        for (int q = 0; q < 64 && t < static_cast<int>(L); ++q)
            if (smask & (uint64_t{1} << q)) gq[t++] = q;
        if (t != static_cast<int>(L)) return {};
    }
    const uint32_t threads = 1u << (L - 4);

    // Tile number -> base index: its bits on the qubits outside the tile set.
    std::vector<int> outside_dst;
    for (uint32_t q = 0; q < nq; ++q)
        if (!(smask & (uint64_t{1} << q))) outside_dst.push_back(static_cast<int>(q));

    const size_t P = rp.phases.size();
    const std::string NT = nt ? "true" : "false";
    std::string s;
    s += "extern \"C\" __global__ void __launch_bounds__(" + dec(threads) + ")" + waves_attr() +
         " lqcs_pass(sp_amp* __restrict__ s, const sp_amp* __restrict__ w) {\n";
    s += "    __shared__ sp_amp lds[" + dec(uint64_t{1} << (L - 1)) + "];\n";
    s += "    char* const lb = reinterpret_cast<char*>(lds);\n";
    s += "    const sp_u32 tid = threadIdx.x;\n";
    s += "    const sp_u32 tile = sp::xcd_tile(blockIdx.x, gridDim.x);\n";
    s += "    sp_amp* const b = s + ((sp_u64)blockIdx.y << " + dec(nq) + ") + (" +
         scatter_expr("(sp_u64)tile", outside_dst, true) + ");\n";
    s += "    sp::Regs v;\n";
    s += "    {\n" + edge_access(lay[0], gq, true, NT) + "    }\n";
    bool lds_used = false;
    for (size_t k = 0; k < P; ++k) {
        const RPhase& ph = rp.phases[k];
        {
            std::string d = "    {  // phase " + dec(k) + ": register bits";
            for (int i = 0; i < 4; ++i) d += " " + dec(lay[k].rb[i]);
            d += "; thread bits";
            for (uint32_t p = 0; p < lay[k].nt; ++p) d += " " + dec(lay[k].tpos[p]);
            s += d + "\n";
        }
        s += "        const sp_u32 t = sp::fresh(tid);\n";
        if (k > 0 && ph.transpose) {
            const int h = split[k];
            const uint32_t pbit = 1u << spos[k];
            const SlotMap sm = slot_map(lay[k - 1], lay[k], h, L);
            s += "        const sp_u32 sp0 = " + slot_thread_expr(lay[k - 1], sm, h) + ";\n";
            s += "        const sp_u32 sc0 = " + slot_thread_expr(lay[k], sm, h) + ";\n";
            for (int half = 0; half < 2; ++half) {
                const std::string cond = half == 0 ? "!(t & " + dec(pbit) + "u)" : "(t & " + dec(pbit) + "u)";
                if (lds_used) s += "        __syncthreads();\n";
                s += "        if (" + cond + ") {\n";
                for (int r = 0; r < 16; ++r)
                    s += "            *reinterpret_cast<sp_amp*>(lb + (sp0 ^ " + dec(sm.bytes(lay[k - 1].jr[r], h)) +
                         "u)) = v[" + dec(r) + "];\n";
                s += "        }\n        __syncthreads();\n        if (" + cond + ") {\n";
                for (int r = 0; r < 16; ++r)
                    s += "            v[" + dec(r) + "] = *reinterpret_cast<const sp_amp*>(lb + (sc0 ^ " +
                         dec(sm.bytes(lay[k].jr[r], h)) + "u));\n";
                s += "        }\n";
                lds_used = true;
            }
        }
        s += "        const sp_u32 jt = " + jt_expr(lay[k]) + ";\n";
        s += "        (void)jt;\n";
        std::string ops;
        const std::string jr_name = "jr" + dec(k);
        bool need_jr = false;
        uint32_t off = ph.op_off;
        for (uint32_t i = 0; i < ph.n_ops; ++i) {
            RegOpHdr h;
            std::memcpy(&h, &rp.words[off], sizeof h);
            const uint32_t kind = h.code & 15u;
            if (kind == RK_DIAG || kind == RK_PAULI) need_jr = true;
            if (!emit_op(ops, h, off + 2, &rp.words[off + 2], jr_name)) return {};
            off += 2 + ((h.code >> 16) & 0xFFFu);
        }
        if (need_jr) s += jr_array(lay[k], jr_name);
        s += ops;
        s += "    }\n";
    }
    s += "    {\n" + edge_access(lay[P - 1], gq, false, NT) + "    }\n}\n";
    return s;
}

namespace {

struct SpecCache {
    std::unordered_map<std::string, hipFunction_t> fn;
    std::vector<hipModule_t> modules;
};
SpecCache& cache() {
    static SpecCache c;
    return c;
}

// hipRTC compile of one pass in this process; the code object, or an
// exception with the log.
std::vector<char> compile(const std::string& body) {
    std::vector<char> code;
    std::string log;
    if (!lqcs_rtc::compile_pass(kSpecHeader, body, code, log))
        throw Error(LQCS_ERR_DEVICE, "hipRTC compile of a specialized pass failed: " + log.substr(0, 2000));
    return code;
}

// hipRTC serialises the compiles of one process, so the passes of a plan
// compile in parallel in helper processes: lqcs_rtc_worker, next to this
// library, reads one kernel text on stdin and writes its code object on stdout.
// Empty when the worker is missing.
const std::string& worker_path() {
    static const std::string path = [] {
        Dl_info info{};
        if (!dladdr(reinterpret_cast<void*>(&spec_cache_size), &info) || !info.dli_fname) return std::string();
        const std::string lib = info.dli_fname;
        const size_t slash = lib.rfind('/');
        const std::string w = (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/lqcs_rtc_worker";
        return access(w.c_str(), X_OK) == 0 ? w : std::string();
    }();
    return path;
}

// One pass compiled by a worker process; false (code empty) when the worker
// cannot be started or fails, so that the caller compiles in process.
bool compile_in_worker(const std::string& body, std::vector<char>& code) {
    const std::string& path = worker_path();
    if (path.empty()) return false;
    int in[2], out[2];
    if (pipe2(in, O_CLOEXEC) != 0) return false;
    if (pipe2(out, O_CLOEXEC) != 0) {
        close(in[0]);
        close(in[1]);
        return false;
    }
    posix_spawn_file_actions_t fa;
    posix_spawn_file_actions_init(&fa);
    posix_spawn_file_actions_adddup2(&fa, in[0], 0);
    posix_spawn_file_actions_adddup2(&fa, out[1], 1);
    // The child keeps stdin, stdout and stderr only (no device or other
    // descriptors of this process).
    posix_spawn_file_actions_addclosefrom_np(&fa, 3);
    char* argv[] = {const_cast<char*>(path.c_str()), nullptr};
    pid_t pid = -1;
    const int rc = posix_spawn(&pid, path.c_str(), &fa, nullptr, argv, environ);
    posix_spawn_file_actions_destroy(&fa);
    close(in[0]);
    close(out[1]);
    if (rc != 0) {
        close(in[1]);
        close(out[0]);
        return false;
    }
    // A worker that exits early must not raise SIGPIPE in this process: block
    // it in this thread and take any pending one before unblocking.
    sigset_t pipe_set, old;
    sigemptyset(&pipe_set);
    sigaddset(&pipe_set, SIGPIPE);
    pthread_sigmask(SIG_BLOCK, &pipe_set, &old);
    // A worker that takes or sends nothing for 10 minutes is stopped (the pass
    // then compiles in process); a compile takes about a second.
    constexpr int kQuietMs = 600 * 1000;
    bool ok = true;
    for (size_t done = 0; ok && done < body.size();) {
        pollfd pf{in[1], POLLOUT, 0};
        const int ready = poll(&pf, 1, kQuietMs);
        if (ready < 0 && errno == EINTR) continue;
        if (ready <= 0) {
            ok = false;
            break;
        }
        const ssize_t n = write(in[1], body.data() + done, body.size() - done);
        if (n > 0) done += static_cast<size_t>(n);
        else if (n < 0 && errno == EINTR) continue;
        else ok = false;
    }
    close(in[1]);
    char buf[65536];
    for (;;) {
        if (!ok) {
            kill(pid, SIGKILL);
            break;
        }
        pollfd pf{out[0], POLLIN, 0};
        const int ready = poll(&pf, 1, kQuietMs);
        if (ready < 0 && errno == EINTR) continue;
        if (ready <= 0) {
            ok = false;
            continue;
        }
        const ssize_t n = read(out[0], buf, sizeof buf);
        if (n > 0) code.insert(code.end(), buf, buf + n);
        else if (n < 0 && errno == EINTR) continue;
        else break;
    }
    close(out[0]);
    int status = 0;
    while (waitpid(pid, &status, 0) < 0 && errno == EINTR) {
    }
    const timespec zero{0, 0};
    while (sigtimedwait(&pipe_set, nullptr, &zero) > 0) {
    }
    pthread_sigmask(SIG_SETMASK, &old, nullptr);
    ok = ok && WIFEXITED(status) && WEXITSTATUS(status) == 0 && code.size() > 4 && std::memcmp(code.data(), "\x7f" "ELF", 4) == 0;
    if (!ok) code.clear();
    return ok;
}

}  // namespace

std::vector<hipFunction_t> spec_kernels(const std::vector<std::string>& sources) {
    SpecCache& c = cache();
    std::vector<std::string> todo;
    for (const std::string& s : sources)
        if (!s.empty() && !c.fn.count(s) && std::find(todo.begin(), todo.end(), s) == todo.end()) todo.push_back(s);
    if (!todo.empty()) {
        const size_t workers = std::max<size_t>(1, std::min<size_t>(todo.size(), std::min(16u, std::thread::hardware_concurrency())));
        std::vector<std::vector<char>> code(todo.size());
        std::vector<std::string> err(todo.size());
        std::vector<std::thread> pool;
        std::atomic<size_t> next{0};
        for (size_t t = 0; t < workers; ++t)
            pool.emplace_back([&] {
                for (size_t i; (i = next.fetch_add(1)) < todo.size();) {
                    try {
                        if (todo.size() < 2 || !compile_in_worker(todo[i], code[i])) code[i] = compile(todo[i]);
                    } catch (const std::exception& e) {
                        err[i] = e.what();
                    }
                }
            });
        for (auto& th : pool) th.join();
        for (size_t i = 0; i < todo.size(); ++i) {
            if (!err[i].empty()) throw Error(LQCS_ERR_DEVICE, err[i]);
            hipModule_t mod = nullptr;
            hipFunction_t fn = nullptr;
            LQCS_HIP(hipModuleLoadData(&mod, code[i].data()));
            c.modules.push_back(mod);
            LQCS_HIP(hipModuleGetFunction(&fn, mod, "lqcs_pass"));
            c.fn.emplace(todo[i], fn);
        }
    }
    std::vector<hipFunction_t> out(sources.size(), nullptr);
    for (size_t i = 0; i < sources.size(); ++i)
        if (!sources[i].empty()) out[i] = c.fn.at(sources[i]);
    return out;
}

size_t spec_cache_size() { return cache().fn.size(); }

void launch_spec(hipFunction_t f, amp_t* s, uint32_t nq, uint32_t L, const amp_t* d_words, uint32_t batch) {
    const uint64_t tiles = uint64_t{1} << (nq - L);
    void* args[] = {&s, &d_words};
    LQCS_HIP(hipModuleLaunchKernel(f, static_cast<unsigned>(tiles), batch, 1, 1u << (L - 4), 1, 1, 0, stream(), args,
                                   nullptr));
}

}  // namespace lqcs_dev
