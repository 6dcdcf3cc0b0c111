// Launchers of the read-mostly passes (reduce.hip). Results land in device
// memory; `batch` registers of 2^nq amplitudes lie back to back.
#pragma once

#include "common.hpp"

namespace lqcs_dev {

void reduce_norm_sq(const amp_t* s, uint32_t nq, uint32_t batch, double* d_out);
// max_i |a_i - b_i| of two registers of width nq (NaN if any difference is NaN).
void reduce_max_abs_diff(const amp_t* a, const amp_t* b, uint32_t nq, double* d_out);
// d_out[2*t .. 2*t+1] = (re, im) of sum_i sign_t(i) a_i conj(a_{i ^ x_t}).
void reduce_pauli_terms(const amp_t* s, uint32_t nq, const uint64_t* d_x, const uint64_t* d_zy,
                        uint32_t n_terms, double* d_out);
// The same with the partial sums in a caller-given buffer of
// pauli_term_partials(nq) * n_terms doubles (no arena use).
uint64_t pauli_term_partials(uint32_t nq);
void reduce_pauli_terms_with(const amp_t* s, uint32_t nq, const uint64_t* d_x, const uint64_t* d_zy,
                             uint32_t n_terms, double* part, double* d_out);
void reduce_p1(const amp_t* s, uint32_t nq, uint32_t q, uint32_t batch, double* d_out);
// d_out[4*r ..] = S00, S11, Re S01, Im S01 of qubit q in register r.
void reduce_qstats(const amp_t* s, uint32_t nq, uint32_t q, uint32_t batch, double* d_out);
void launch_collapse(amp_t* s, uint32_t nq, uint32_t q, const double* d_p1, const double* d_draws,
                     uint64_t draw_stride, uint64_t draw_offset, int reset, uint8_t* d_creg,
                     uint32_t creg_stride, int32_t clbit, uint32_t batch);
void launch_probabilities(const amp_t* s, uint32_t nq, double* d_out);
void launch_sample(const amp_t* s, uint32_t nq, const double* d_draws, uint64_t n_draws,
                   uint64_t* d_out);
// Density matrices as 2n-qubit registers (rho[r][c] at r | c << n):
// d_out[r] = Re rho[r][r]; and <psi|rho|psi> for an n-qubit psi.
void launch_dm_diagonal(const amp_t* rho, uint32_t n, double* d_out);
void reduce_dm_fidelity(const amp_t* rho, const amp_t* psi, uint32_t n, double* d_out);

}  // namespace lqcs_dev
