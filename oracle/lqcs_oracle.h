/*
 * lqcs_oracle.h -- CPU restatement of LiteQuantumComputingSimulator's
 * statevector path, used ONLY as a correctness checker and CPU baseline.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library. The product (the HIP engine) never
 * links, loads or calls it.
 *
 * Every function restates one piece of the reference (file:line cited next to
 * it) with the same per-element arithmetic, loop order and RNG consumption, so
 * element-wise results are bit-identical to the reference built without
 * -march flags and reductions match the reference run with one OpenMP thread.
 *
 * Pinning (what fixes each function independently of this file):
 *   - mt19937_64 / uniform draw: the C++ standard's 10000th-output value and a
 *     comparison against libstdc++'s std::uniform_real_distribution
 *     (tests/test_oracle.py);
 *   - gate matrices, kernels, dispatch, fusion, sampling, trajectories,
 *     expectation, parameter-shift, density matrix: golden vectors written by
 *     the reference itself compiled from its sources (oracle/build_ref.sh,
 *     tests/golden/make_golden.py), closed forms (QFT|x> = DFT, Bell/GHZ,
 *     RY expectations, depolarised-Bell fidelity), special-case identities
 *     (rp("XX") == RXX ...), and a brute-force dense embedding at n <= 6.
 */
#ifndef LQCS_ORACLE_H
#define LQCS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc_c;

/* Gate type numbering follows include/lqcs/circuit/gate.hpp:14-28. */
enum {
    ORC_I, ORC_X, ORC_Y, ORC_Z, ORC_H, ORC_S, ORC_SDG, ORC_T, ORC_TDG, ORC_SX,
    ORC_RX, ORC_RY, ORC_RZ, ORC_P, ORC_U,
    ORC_SWAP, ORC_ISWAP, ORC_RXX, ORC_RYY, ORC_RZZ,
    ORC_PAULI_ROTATION, ORC_UNITARY, ORC_PERMUTATION,
    ORC_MEASURE, ORC_RESET, ORC_BARRIER
};

/* One instruction: gate + qubits (controls first) + clbit for Measure
 * (include/lqcs/circuit/instruction.hpp:9-13, gate.hpp:30-43). */
typedef struct {
    int32_t type;
    uint32_t n_controls;
    uint32_t n_qubits;
    uint32_t n_params;
    const uint32_t* qubits;
    const double* params;
    uint64_t mat_len;          /* Unitary: dim*dim entries */
    const orc_c* mat;
    uint64_t perm_len;         /* Permutation: 2^k entries */
    const uint64_t* perm;
    uint32_t pauli_len;        /* PauliRotation */
    const char* pauli;
    uint32_t clbit;            /* Measure */
} orc_inst;

typedef struct {
    uint32_t n_qubits;
    uint32_t n_clbits;
    uint64_t n_inst;
    const orc_inst* inst;
} orc_circuit;

typedef struct {
    uint32_t n_ops;
    const orc_c* ops;          /* n_ops * 4 row-major 2x2 */
} orc_channel;

typedef struct {
    uint32_t n_channels;
    const orc_channel* channels;
} orc_noise;

typedef struct {
    double coeff;
    const char* pauli;         /* full width, leftmost char = highest qubit */
} orc_term;

/* ---- RNG: mt19937_64 + libstdc++ uniform_real_distribution<double>(0,1) */
typedef struct { uint64_t mt[312]; uint32_t idx; } orc_rng;
void orc_rng_seed(orc_rng* r, uint64_t seed);
uint64_t orc_rng_next(orc_rng* r);
double orc_rng_uniform(orc_rng* r);
/* Fill out[0..n) with consecutive uniform draws of a fresh stream. */
void orc_uniform_stream(uint64_t seed, uint64_t n, double* out);

/* ---- gate model */
uint32_t orc_base_qubits(const orc_inst* g);
int orc_is_diagonal(const orc_inst* g);
/* Writes the 2^k x 2^k base matrix; returns dim or 0 for non-unitary. */
uint64_t orc_base_matrix(const orc_inst* g, orc_c* out);

/* ---- kernels (in place on state of n_amps amplitudes) */
void orc_apply_single_qubit(orc_c* s, uint64_t n, uint32_t t, const orc_c m[4]);
void orc_apply_diagonal_single_qubit(orc_c* s, uint64_t n, uint32_t t, orc_c d0, orc_c d1);
void orc_apply_controlled_single_qubit(orc_c* s, uint64_t n, uint32_t c, uint32_t t, const orc_c m[4]);
void orc_apply_swap(orc_c* s, uint64_t n, uint32_t a, uint32_t b);
void orc_apply_two_qubit(orc_c* s, uint64_t n, uint32_t q0, uint32_t q1, const orc_c m[16]);
void orc_apply_controlled_unitary(orc_c* s, uint64_t n, const uint32_t* ctl, uint32_t nc,
                                  const uint32_t* tgt, uint32_t nt, const orc_c* m);
void orc_apply_controlled_permutation(orc_c* s, uint64_t n, const uint32_t* ctl, uint32_t nc,
                                      const uint32_t* tgt, uint32_t nt, const uint64_t* perm);
void orc_apply_pauli_rotation(orc_c* s, uint64_t n, uint64_t x, uint64_t zy, uint32_t n_y,
                              double theta);
int orc_collapse_qubit(orc_c* s, uint64_t n, uint32_t q, double u);
void orc_reset_qubit(orc_c* s, uint64_t n, uint32_t q, double u);
double orc_kraus_probability(const orc_c* s, uint64_t n, uint32_t t, const orc_c k[4]);

/* ---- simulator */
/* Applies one unitary instruction with the reference dispatcher. */
int orc_apply_instruction(orc_c* s, uint64_t n, const orc_inst* g);
/* Greedy single-qubit fusion. Returns an opaque fused circuit. */
typedef struct orc_fused orc_fused;
orc_fused* orc_fuse(const orc_circuit* c);
const orc_circuit* orc_fused_circuit(const orc_fused* f);
void orc_fused_free(orc_fused* f);
/* run_statevector: out gets 2^n amplitudes. Returns 0 or -1 (Reset). */
int orc_run_statevector(const orc_circuit* c, int fuse, orc_c* out);
/* Terminal-measurement sampling on an evolved state: idx_out[shots]
 * receives the sampled basis indices, cum_out (optional, 2^n) the
 * cumulative distribution. Returns 0, or -1 when the circuit needs the
 * per-shot path, -2 when it has no Measure. */
int orc_run_sampling(const orc_circuit* c, int fuse, uint64_t seed, uint64_t shots,
                     uint64_t* idx_out, double* cum_out);
/* Per-shot path (mid-circuit measure / reset / noise): clbits_out is
 * shots * n_clbits bytes, 0/1 per classical bit. */
int orc_run_per_shot(const orc_circuit* c, const orc_noise* noise, uint64_t seed,
                     uint64_t shots, uint8_t* clbits_out);
/* StatevectorSimulator::run: 0 = sampling path, 1 = per-shot path,
 * -2 = no Measure, -3 = shots == 0. clbits_out: shots * n_clbits bytes. */
int orc_run(const orc_circuit* c, const orc_noise* noise, int fuse, uint64_t seed,
            uint64_t shots, uint8_t* clbits_out);
/* Number of uniform draws one shot of run_per_shot consumes. */
uint64_t orc_draws_per_shot(const orc_circuit* c, const orc_noise* noise);

/* ---- observables */
double orc_norm(const orc_c* s, uint64_t n);
double orc_expectation_value(const orc_c* s, uint32_t n_qubits, const char* pauli);
double orc_expectation(const orc_c* s, uint32_t n_qubits, const orc_term* h, uint64_t n_terms);
/* Parameter-shift gradient of <H> for a parametric template: param_inst[p]
 * is the template instruction whose params[0] parameter p sets. */
int orc_parameter_shift_gradient(const orc_circuit* tmpl, const uint64_t* param_inst,
                                 uint64_t n_params, const double* theta,
                                 const orc_term* h, uint64_t n_terms, double* grad);
int orc_bound_energy(const orc_circuit* tmpl, const uint64_t* param_inst, uint64_t n_params,
                     const double* theta, const orc_term* h, uint64_t n_terms, double* energy);

/* ---- density matrix (4^n vectorised rho, ket low bits / bra high bits) */
int orc_run_density_matrix(const orc_circuit* c, const orc_noise* noise, orc_c* rho_out);

#ifdef __cplusplus
}
#endif
#endif
