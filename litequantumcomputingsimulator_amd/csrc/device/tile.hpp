// Tile-pass operation encoding shared by the planner and the tile kernel.
#pragma once

#include "common.hpp"

#include <vector>

namespace lqcs_dev {

enum TileKind : uint32_t {
    TK_DENSE1 = 0,  // 2x2 on local bit t0
    TK_DENSE2 = 1,  // 4x4 on local bits (t0 = matrix bit 0, t1 = matrix bit 1)
    TK_DIAG = 2,    // diagonal over up to two qubits, each local (dl*) or outside (dg*)
    TK_PAULI = 3    // exp(-i theta/2 P): local X bits xl, Z parity from zl and zg
};

// One gate inside a tile pass. Controls: lctrl over the tile-local index,
// gctrl over the global index (qubits outside the tile).
struct TileOp {
    uint32_t kind;
    uint32_t t0, t1;
    // Batch-member restriction (trajectories): member >= 0 with only = 1:
    // that register alone; with only = 0: every register but that one.
    int32_t member;
    uint32_t only;
    uint32_t lctrl;
    uint64_t gctrl;
    uint32_t dl0, dl1;
    uint64_t dg0, dg1;
    uint32_t xl, zl;
    uint64_t zg;
    amp_t m[16];
};

// Register-resident tile (rtile.hip): per-phase layout and in-register ops.
enum RegKind : uint32_t { RK_DENSE1 = 0, RK_DENSE2 = 1, RK_X = 2, RK_DIAG = 3, RK_PAULI = 4, RK_SEL1 = 5, RK_DIAG1 = 6,
                          RK_DENSE2R = 7 };

// Header of one in-register op in a pass's op words: two 16-byte words, then
// the op's matrix entries (4 for DENSE1 and DIAG, 16 for DENSE2, 2 for
// PAULI, none for X, 8 for SEL1). SEL1 is a 2x2 op on register bit b0 whose
// matrix is chosen by one select bit: the first four entries where it is 0,
// the next four where it is 1. The select bit is register bit xr - 1 when xr
// is nonzero, else the tile-local thread bit la when nonzero, else the
// tile-number bit ga. DIAG1 is diag(d0, d1) on one qubit (2 entries): the
// qubit is register bit xr - 1 when xr is nonzero, else thread bit la when
// nonzero, else tile-number bit ga; b1 = 1 marks d0 exactly 1. DENSE2R is
// a DENSE2 without controls whose entries are all real (16 entries, imaginary
// parts zero), such as the superoperator of a real Kraus channel. Masks named "tile number" are over the qubits outside
// the tile, compressed to the tile's index. The kernel reads them with
// scalar loads (the same address in every lane).
// Flags of an op's code word, precomputed by the builder so that the kernel's
// dispatch tests one bit: the op has a tile-number control or a batch-register
// restriction to check; the op has no controls among the register or thread bits.
constexpr uint32_t kCodeRestricted = 1u << 28;
constexpr uint32_t kCodeFree = 1u << 29;

struct RegOpHdr {
    uint32_t code;    // kind | b0 << 4 | b1 << 6 | xr << 8 | rctl << 12 | matrix words << 16 | flags
    uint32_t tctl;    // controls among the thread's tile bits (tile-local mask)
    uint32_t la, lb;  // DIAG: tile-local diagonal bits; PAULI: la = tile-local Z bits
    uint32_t gctrl;   // controls on the tile number
    uint32_t ga, gb;  // DIAG: tile-number diagonal bits; PAULI: ga = tile-number Z bits
    uint32_t member;  // 0: every batch register; else (index + 1) | 1 << 31 for that one alone,
                      // (index + 1) for every one but it
};
static_assert(sizeof(RegOpHdr) == 2 * sizeof(amp_t), "op header is two 16-byte words");

struct RPhase {
    uint32_t n_ops;
    uint32_t op_off;         // first op word within the pass's words
    uint32_t tmask;          // tile bits indexed by the thread id (ascending)
    uint32_t transpose;      // 1: the tile is transposed into this layout first
    uint64_t tmask_global;   // their global qubit positions
    uint32_t jr[16];         // tile index of register r (thread bits zero)
    uint64_t gofs[16];       // global offset of register r
    // The transpose into this layout moves the tile through LDS in two
    // halves, split by tile bit h (a thread bit of both layouts): the LDS
    // slot of tile index j is slot(j without bit h), XOR-linear, so slot(jt |
    // jr) = slot(jt) ^ slot(jr); these are the register parts, as LDS byte
    // offsets (slot * 16).
    uint32_t h;
    uint32_t pad2;
    uint32_t cs_prev[16];    // previous layout's register r
    uint32_t cs_cur[16];     // this layout's register r
    // Index arithmetic without bit loops: (1 << p) - 1 for the 4 register
    // bits p, ascending (the thread's tile bits are the thread id with zeros
    // inserted there); the global qubit of each of the 8 thread bits; and
    // (1 << q) - 1 for the 12 tile qubits q, ascending (the tile number with
    // zeros inserted there is the tile's base index).
    uint32_t hole_low[4];
    uint8_t tpos[8];
    uint64_t tile_low[12];
};

// Tile bits of thread `tid` in layout ph (pdep of tid over ph.tmask): each
// insertion keeps the bits below the register bit and shifts the rest up;
// the bit left at the register position is cleared by the final mask.
__host__ __device__ __forceinline__ uint32_t rtile_thread_bits(uint32_t tid, const RPhase& ph) {
    uint32_t x = tid;
#pragma unroll
    for (int i = 0; i < 4; ++i) x = (x & ph.hole_low[i]) | ((x << 1) & ~ph.hole_low[i]);
    return x & ph.tmask;
}

// Global index bits of thread `tid` in layout ph (pdep of tid over
// ph.tmask_global).
__host__ __device__ __forceinline__ uint64_t rtile_thread_global(uint32_t tid, const RPhase& ph) {
    uint64_t g = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) g |= uint64_t((tid >> k) & 1u) << ph.tpos[k];
    return g;
}

// Base index of tile number `tile` (pdep over the qubits outside the tile
// set; `outside` = every qubit of the register not in the tile set).
__host__ __device__ __forceinline__ uint64_t rtile_tile_base(uint64_t tile, const RPhase& ph, uint64_t outside) {
    uint64_t x = tile;
#pragma unroll
    for (int i = 0; i < 12; ++i) x = (x & ph.tile_low[i]) | ((x << 1) & ~ph.tile_low[i]);
    return x & outside;
}

// LDS slot of tile index j in a half-tile transpose split by bit h: j with
// bit h removed (11 bits), low 3 bits XOR-folded with higher ones so that
// lanes differing only in high bits spread over the banks.
__host__ __device__ __forceinline__ uint32_t rtile_slot(uint32_t j, uint32_t h) {
    const uint32_t c = (j & ((1u << h) - 1u)) | ((j >> (h + 1)) << h);
    return c ^ (((c >> 3) ^ (c >> 6) ^ (c >> 9)) & 7u);
}

struct RTileProgram {
    std::vector<RPhase> phases;
    std::vector<amp_t> words;  // the pass's op headers and matrices, 16 bytes each
    uint32_t merged = 0;       // tile ops absorbed into SEL1 ops
};

// Splits a tiled step into register phases; false when some op needs more
// than 4 register bits or the tile width is not the register tile's.
// Register programs: L = 12 (four waves per tile) or L = 10 (one wave per
// tile, so transposes involve no other wave); both transpose through LDS in
// two halves. L = 13 (eight waves per tile) builds programs that only the
// specialized passes run. False for other widths or ops that do not fit.
bool build_rtile(const std::vector<TileOp>& tops, uint32_t L, uint64_t smask, uint32_t nq, RTileProgram& out);
void launch_rtile(amp_t* s, uint32_t nq, uint32_t L, uint64_t smask, const RPhase* d_phases, uint32_t n_phases,
                  const amp_t* d_words, uint32_t batch = 1);

void launch_tile_pass(amp_t* s, uint32_t nq, uint32_t L, uint64_t smask, const TileOp* d_ops,
                      uint32_t n_ops, uint32_t batch = 1);

}  // namespace lqcs_dev
