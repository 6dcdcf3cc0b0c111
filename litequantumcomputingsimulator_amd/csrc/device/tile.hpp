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
    uint32_t lctrl;
    uint64_t gctrl;
    uint32_t dl0, dl1;
    uint64_t dg0, dg1;
    uint32_t xl, zl;
    uint64_t zg;
    amp_t m[16];
};

// Register-resident tile (rtile.hip): per-phase layout and in-register ops.
enum RegKind : uint32_t { RK_DENSE1 = 0, RK_DENSE2 = 1, RK_X = 2, RK_DIAG = 3, RK_PAULI = 4 };

struct RegOp {
    uint32_t kind;
    uint32_t b0, b1;     // register-index bits of the targets
    uint32_t rctl;       // controls among the register-index bits
    uint32_t tctl;       // controls among the thread's tile bits (tile-local mask)
    uint32_t dl0, dl1;   // diagonal qubits inside the tile (tile-local masks)
    uint32_t xr;         // Pauli X bits, register-index mask
    uint32_t zl;         // Pauli Z bits, tile-local mask
    uint32_t pad;
    uint64_t gctrl, dg0, dg1, zg;  // the same, for qubits outside the tile
    amp_t m[16];
};

struct RPhase {
    uint32_t first_op, n_ops;
    uint32_t tmask;          // tile bits indexed by the thread id (ascending)
    uint32_t pad;
    uint64_t tmask_global;   // their global qubit positions
    uint32_t jr[16];         // tile index of register r (thread bits zero)
    uint32_t sjr[16];        // its LDS slot contribution (XOR-linear swizzle)
    uint64_t gofs[16];       // global offset of register r
};

struct RTileProgram {
    std::vector<RPhase> phases;
    std::vector<RegOp> ops;
};

// Splits a tiled step into register phases; false when some op needs more
// than 4 register bits or the tile width is not the register tile's.
bool build_rtile(const std::vector<TileOp>& tops, uint32_t L, uint64_t smask, RTileProgram& out);
void launch_rtile(amp_t* s, uint32_t nq, uint64_t smask, const RPhase* d_phases, uint32_t n_phases,
                  const RegOp* d_ops, uint32_t batch = 1);

void launch_tile_pass(amp_t* s, uint32_t nq, uint32_t L, uint64_t smask, const TileOp* d_ops,
                      uint32_t n_ops, uint32_t batch = 1);

}  // namespace lqcs_dev
