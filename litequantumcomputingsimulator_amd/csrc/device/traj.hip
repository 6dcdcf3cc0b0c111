// Batched noisy trajectories (the per-shot path of run_per_shot,
// src/backend/statevector_simulator.cpp:122-190).
//
// B shots evolve side by side as B registers in one allocation; every gate is
// one launch over the whole batch (blockIdx.y = shot). A Kraus channel on a
// qubit costs one read pass for the qubit's 2x2 statistics (S00, S11, S01),
// from which every branch probability ||K psi||^2 follows, and one pass that
// picks the branch with the shot's own uniform draw and applies K/sqrt(p).
// Draw order is the reference's: per unitary instruction, channel-major then
// qubit order; one draw per Measure/Reset; shot s owns draws [s*D, (s+1)*D).
#include "common.hpp"
#include "engine.hpp"
#include "reduce.hpp"

#include <cstring>
#include <vector>

namespace lqcs_dev {
namespace {

constexpr int kMaxKraus = 8;
struct KrausSet {
    amp_t k[kMaxKraus][4];
    uint32_t n;
};

__global__ void __launch_bounds__(kBlock) k_kraus_apply(amp_t* __restrict__ s, uint32_t nq,
                                                        uint32_t q, const double* __restrict__ stats,
                                                        KrausSet ks,
                                                        const double* __restrict__ draws,
                                                        uint64_t draw_stride, uint64_t draw_offset) {
    const uint32_t r = blockIdx.y;
    const double* st = stats + 4 * r;
    const double u = draws[uint64_t(r) * draw_stride + draw_offset];
    // Branch choice as sample_kraus_branch: skip p <= 0, first with u < cum,
    // falling back to the last non-zero branch.
    int chosen = -1;
    double chosen_p = 0.0, cum = 0.0;
    for (uint32_t i = 0; i < ks.n; ++i) {
        const amp_t* K = ks.k[i];
        const double c00 = cnorm(K[0]) + cnorm(K[2]);
        const double c11 = cnorm(K[1]) + cnorm(K[3]);
        const double xr = K[0].x * K[1].x + K[0].y * K[1].y + K[2].x * K[3].x + K[2].y * K[3].y;
        const double xi = K[0].x * K[1].y - K[0].y * K[1].x + K[2].x * K[3].y - K[2].y * K[3].x;
        const double p = c00 * st[0] + c11 * st[1] + 2.0 * (xr * st[2] - xi * st[3]);
        if (p <= 0.0) continue;
        chosen = static_cast<int>(i);
        chosen_p = p;
        cum += p;
        if (u < cum) break;
    }
    if (chosen < 0) return;
    const double inv = 1.0 / sqrt(chosen_p);
    const amp_t m0 = cscale(ks.k[chosen][0], inv), m1 = cscale(ks.k[chosen][1], inv);
    const amp_t m2 = cscale(ks.k[chosen][2], inv), m3 = cscale(ks.k[chosen][3], inv);
    amp_t* b = s + (uint64_t(r) << nq);
    const uint64_t half = uint64_t{1} << (nq - 1);
    for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < half;
         g += uint64_t(gridDim.x) * kBlock) {
        const uint64_t i0 = insert_zero_bit(g, q), i1 = i0 | (uint64_t{1} << q);
        const amp_t a0 = b[i0], a1 = b[i1];
        b[i0] = cmac2(m0, a0, m1, a1);
        b[i1] = cmac2(m2, a0, m3, a1);
    }
}

unsigned pass_blocks(uint32_t nq) {
    const uint64_t want = ((uint64_t{1} << (nq - 1)) + kBlock * 4 - 1) / (kBlock * 4);
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(want, 1u << 20)));
}

}  // namespace
}  // namespace lqcs_dev

using namespace lqcs_dev;

extern "C" lqcs_status lqcs_trajectories_run(uint32_t nq, uint32_t n_clbits, const lqcs_program* prog,
                                             const lqcs_traj_step* steps, uint64_t n_steps,
                                             const uint32_t* noise_qubits, const lqcs_channel* channels,
                                             uint32_t n_channels, const double* draws,
                                             uint64_t D, uint64_t n_shots, uint32_t max_batch,
                                             uint8_t* clbits_out) {
    return run_guarded([&] {
        if (nq == 0 || nq > LQCS_DEVICE_MAX_QUBITS) throw Error(LQCS_ERR_INVALID_ARGUMENT, "bad register width");
        if (!prog || (n_steps && !steps)) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null program or steps");
        if (n_shots == 0) return;
        if (!draws || !clbits_out) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null draws or output");
        if (n_channels && !channels) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null channels");
        ensure_ready();
        arena_reset();
        std::vector<DecodedOp> ops;
        for (uint64_t k = 0; k < prog->n_ops; ++k) ops.push_back(decode(*prog, k, nq));
        std::vector<KrausSet> sets(n_channels);
        for (uint32_t c = 0; c < n_channels; ++c) {
            if (channels[c].n_ops == 0 || channels[c].n_ops > kMaxKraus)
                throw Error(LQCS_ERR_INVALID_ARGUMENT, "channel needs 1..8 Kraus operators");
            std::memset(&sets[c], 0, sizeof(KrausSet));
            sets[c].n = channels[c].n_ops;
            for (uint32_t i = 0; i < channels[c].n_ops; ++i)
                for (int e = 0; e < 4; ++e)
                    sets[c].k[i][e] = make_double2(channels[c].ops[8 * i + 2 * e], channels[c].ops[8 * i + 2 * e + 1]);
        }
        // Draw offsets per step and a consistency check against D.
        std::vector<uint64_t> off(n_steps);
        uint64_t d = 0;
        for (uint64_t i = 0; i < n_steps; ++i) {
            off[i] = d;
            const lqcs_traj_step& st = steps[i];
            if (st.kind == LQCS_TOP_UNITARY) {
                if (st.op_index >= ops.size()) throw Error(LQCS_ERR_INVALID_ARGUMENT, "step op index out of range");
                d += uint64_t(n_channels) * st.n_noise_qubits;
                for (uint32_t j = 0; j < st.n_noise_qubits; ++j)
                    if (noise_qubits[st.noise_qubit_offset + j] >= nq)
                        throw Error(LQCS_ERR_OUT_OF_RANGE, "noise qubit out of range");
            } else if (st.kind == LQCS_TOP_MEASURE || st.kind == LQCS_TOP_RESET) {
                if (st.qubit >= nq) throw Error(LQCS_ERR_OUT_OF_RANGE, "measured qubit out of range");
                if (st.kind == LQCS_TOP_MEASURE && st.clbit >= n_clbits)
                    throw Error(LQCS_ERR_OUT_OF_RANGE, "clbit out of range");
                d += 1;
            } else {
                throw Error(LQCS_ERR_INVALID_ARGUMENT, "unknown trajectory step kind");
            }
        }
        if (d != D) throw Error(LQCS_ERR_INVALID_ARGUMENT, "draws_per_shot does not match the steps");

        // Batch size: at most max_batch, at most ~1/2 of free HBM.
        size_t free_b = 0, total_b = 0;
        LQCS_HIP(hipMemGetInfo(&free_b, &total_b));
        const uint64_t reg_bytes = (uint64_t{1} << nq) * sizeof(amp_t);
        uint64_t B = std::max<uint64_t>(1, (free_b / 2) / reg_bytes);
        B = std::min<uint64_t>(B, max_batch ? max_batch : 65535);
        B = std::min<uint64_t>(B, n_shots);
        B = std::min<uint64_t>(B, 65535);
        amp_t* states = nullptr;
        LQCS_HIP(hipMalloc(&states, B * reg_bytes));
        struct Free {
            amp_t* p;
            ~Free() { if (p) (void)hipFree(p); }
        } guard{states};
        const uint32_t creg_w = std::max<uint32_t>(1, n_clbits);
        uint8_t* d_creg = static_cast<uint8_t*>(arena(B * creg_w));
        double* d_draws = static_cast<double*>(arena(B * std::max<uint64_t>(D, 1) * sizeof(double)));
        double* d_stats = static_cast<double*>(arena(B * 4 * sizeof(double)));
        std::vector<uint8_t> host_creg(B * creg_w);

        for (uint64_t s0 = 0; s0 < n_shots; s0 += B) {
            const uint32_t nb = static_cast<uint32_t>(std::min<uint64_t>(B, n_shots - s0));
            launch_fill_basis(states, nq, 0, nb);
            LQCS_HIP(hipMemsetAsync(d_creg, 0, uint64_t(nb) * creg_w, stream()));
            if (D)
                LQCS_HIP(hipMemcpyAsync(d_draws, draws + s0 * D, uint64_t(nb) * D * sizeof(double),
                                        hipMemcpyHostToDevice, stream()));
            for (uint64_t i = 0; i < n_steps; ++i) {
                const lqcs_traj_step& st = steps[i];
                if (st.kind == LQCS_TOP_UNITARY) {
                    run_single_op(states, nq, ops[st.op_index], nb);
                    uint64_t dd = off[i];
                    for (uint32_t c = 0; c < n_channels; ++c)
                        for (uint32_t j = 0; j < st.n_noise_qubits; ++j, ++dd) {
                            const uint32_t q = noise_qubits[st.noise_qubit_offset + j];
                            reduce_qstats(states, nq, q, nb, d_stats);
                            k_kraus_apply<<<dim3(pass_blocks(nq), nb), kBlock, 0, stream()>>>(
                                states, nq, q, d_stats, sets[c], d_draws, D, dd);
                        }
                } else {
                    reduce_p1(states, nq, st.qubit, nb, d_stats);
                    const int reset = st.kind == LQCS_TOP_RESET;
                    launch_collapse(states, nq, st.qubit, d_stats, d_draws, D, off[i], reset, d_creg, creg_w,
                                    reset ? -1 : static_cast<int32_t>(st.clbit), nb);
                }
            }
            LQCS_HIP(hipMemcpyAsync(host_creg.data(), d_creg, uint64_t(nb) * creg_w, hipMemcpyDeviceToHost,
                                    stream()));
            sync();
            for (uint32_t b = 0; b < nb; ++b)
                std::memcpy(clbits_out + (s0 + b) * n_clbits, host_creg.data() + uint64_t(b) * creg_w, n_clbits);
        }
    });
}
