// Tile-pass operation encoding shared by the planner and the tile kernel.
#pragma once

#include "common.hpp"

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

void launch_tile_pass(amp_t* s, uint32_t nq, uint32_t L, uint64_t smask, const TileOp* d_ops,
                      uint32_t n_ops, uint32_t batch = 1);

}  // namespace lqcs_dev
