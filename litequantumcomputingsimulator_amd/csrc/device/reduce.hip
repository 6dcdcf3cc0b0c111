// Read-mostly passes: norms, Pauli expectations, Born sampling, collapse,
// and the per-qubit 2x2 statistics that Kraus branch probabilities need.
// Reductions are two-level (per-block partials, then one block) with a fixed
// grid, so results are run-to-run deterministic. blockIdx.y = batch member.
#include "common.hpp"
#include "reduce.hpp"

namespace lqcs_dev {
namespace {

constexpr int kRedBlocks = 1024;  // partial-sum blocks per register

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of up to 4 values per thread; result valid in thread 0.
template <int V>
__device__ __forceinline__ void block_sum(double (&v)[V]) {
    __shared__ double red[V][kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = wave_sum(v[i]);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < V; ++i) red[i][wid] = v[i];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < V; ++i) {
            double acc = 0.0;
            for (int w = 0; w < kBlock / 64; ++w) acc += red[i][w];
            v[i] = acc;
        }
    }
    __syncthreads();
}

// ---- sum of |a|^2 ---------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_norm_partial(const amp_t* __restrict__ s, uint32_t nq,
                                                         double* __restrict__ part) {
    const amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint64_t n = uint64_t{1} << nq;
    double acc[1] = {0.0};
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * kBlock)
        acc[0] += cnorm(b[i]);
    block_sum(acc);
    if (threadIdx.x == 0) part[uint64_t(blockIdx.y) * gridDim.x + blockIdx.x] = acc[0];
}

// ---- max |a_i - b_i| over two registers (NaN propagates) ----------------------------
__device__ __forceinline__ double nan_max(double m, double d) { return (d > m || d != d) ? d : m; }

__device__ __forceinline__ double block_max(double m) {
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) sh[threadIdx.x] = nan_max(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    return sh[0];
}

__global__ void __launch_bounds__(kBlock) k_maxdiff_partial(const amp_t* __restrict__ a, const amp_t* __restrict__ b,
                                                            uint64_t n, double* __restrict__ part) {
    double m = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
        const amp_t x = a[i], y = b[i];
        const double dx = x.x - y.x, dy = x.y - y.y;
        m = nan_max(m, sqrt(dx * dx + dy * dy));
    }
    m = block_max(m);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void __launch_bounds__(kBlock) k_max_finish(const double* __restrict__ part, uint32_t count,
                                                       double* __restrict__ out) {
    double m = 0.0;
    for (uint32_t j = threadIdx.x; j < count; j += kBlock) m = nan_max(m, part[j]);
    m = block_max(m);
    if (threadIdx.x == 0) out[0] = m;
}

// Sums `count` partials of each of gridDim.x rows (V values interleaved).
template <int V>
__global__ void __launch_bounds__(kBlock) k_finish(const double* __restrict__ part,
                                                   uint32_t count, double* __restrict__ out) {
    const double* row = part + uint64_t(blockIdx.x) * count * V;
    double acc[V];
#pragma unroll
    for (int i = 0; i < V; ++i) acc[i] = 0.0;
    for (uint32_t j = threadIdx.x; j < count; j += kBlock)
#pragma unroll
        for (int i = 0; i < V; ++i) acc[i] += row[uint64_t(j) * V + i];
    block_sum(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < V; ++i) out[uint64_t(blockIdx.x) * V + i] = acc[i];
}

// ---- Pauli expectation: sum_i sign(i) a_i conj(a_{i^x}); term = blockIdx.y ----
__global__ void __launch_bounds__(kBlock) k_pauli_expect_partial(
    const amp_t* __restrict__ s, uint32_t nq, const uint64_t* __restrict__ xs,
    const uint64_t* __restrict__ zys, double* __restrict__ part) {
    const uint64_t x = xs[blockIdx.y], zy = zys[blockIdx.y];
    double acc[2] = {0.0, 0.0};
    if (x == 0) {
        const uint64_t n = uint64_t{1} << nq;
        for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
             i += uint64_t(gridDim.x) * kBlock) {
            const double v = cnorm(s[i]);
            acc[0] += (__popcll(i & zy) & 1) ? -v : v;
        }
    } else {
        const uint32_t lsb = __builtin_ctzll(x);
        const uint64_t pairs = uint64_t{1} << (nq - 1);
        for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < pairs;
             g += uint64_t(gridDim.x) * kBlock) {
            const uint64_t i = insert_zero_bit(g, lsb), j = i ^ x;
            const amp_t ai = s[i], aj = s[j];
            const double si = (__popcll(i & zy) & 1) ? -1.0 : 1.0;
            const double sj = (__popcll(j & zy) & 1) ? -1.0 : 1.0;
            // si * ai * conj(aj) + sj * aj * conj(ai)
            const double re = ai.x * aj.x + ai.y * aj.y;
            const double im = ai.y * aj.x - ai.x * aj.y;
            acc[0] += (si + sj) * re;
            acc[1] += (si - sj) * im;
        }
    }
    block_sum(acc);
    if (threadIdx.x == 0) {
        double* o = part + (uint64_t(blockIdx.y) * gridDim.x + blockIdx.x) * 2;
        o[0] = acc[0];
        o[1] = acc[1];
    }
}

// ---- Born sampling ------------------------------------------------------------
constexpr int kChunkLog = 12;  // amplitudes per sampling chunk (when n >= 12)

__global__ void __launch_bounds__(kBlock) k_chunk_sums(const amp_t* __restrict__ s,
                                                       uint32_t chunk_log,
                                                       double* __restrict__ sums) {
    const uint64_t base = uint64_t(blockIdx.x) << chunk_log;
    const uint32_t len = 1u << chunk_log;
    double acc[1] = {0.0};
    for (uint32_t i = threadIdx.x; i < len; i += kBlock) acc[0] += cnorm(s[base + i]);
    block_sum(acc);
    if (threadIdx.x == 0) sums[blockIdx.x] = acc[0];
}

// Inclusive prefix of chunk sums by one block: serial runs then a block scan.
__global__ void __launch_bounds__(1024) k_scan_chunks(const double* __restrict__ sums,
                                                      uint64_t count,
                                                      double* __restrict__ incl) {
    __shared__ double carry[1024];
    const uint64_t per = (count + 1023) / 1024;
    const uint64_t b0 = uint64_t(threadIdx.x) * per;
    double run = 0.0;
    for (uint64_t i = b0; i < b0 + per && i < count; ++i) run += sums[i];
    carry[threadIdx.x] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int t = 0; t < 1024; ++t) {
            const double v = carry[t];
            carry[t] = acc;
            acc += v;
        }
    }
    __syncthreads();
    run = carry[threadIdx.x];
    for (uint64_t i = b0; i < b0 + per && i < count; ++i) {
        run += sums[i];
        incl[i] = run;
    }
}

// One wave per draw: r = u * total; index = first i whose running sum exceeds r.
__global__ void __launch_bounds__(64) k_sample(const amp_t* __restrict__ s, uint32_t nq,
                                               uint32_t chunk_log,
                                               const double* __restrict__ incl,
                                               uint64_t n_chunks,
                                               const double* __restrict__ draws,
                                               uint64_t n_draws, uint64_t* __restrict__ out) {
    const uint64_t d = blockIdx.x;
    if (d >= n_draws) return;
    const double total = incl[n_chunks - 1];
    const double r = draws[d] * total;
    uint64_t lo = 0, hi = n_chunks;  // first chunk with incl > r
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (r < incl[mid]) hi = mid; else lo = mid + 1;
    }
    if (lo == n_chunks) {  // r beyond the last cumulative value: upper_bound == end
        if (threadIdx.x == 0) out[d] = uint64_t{1} << nq;
        return;
    }
    const uint64_t chunk = lo;
    const double before = chunk ? incl[chunk - 1] : 0.0;
    const uint32_t len = 1u << chunk_log;
    const uint32_t per = len >= 64 ? len / 64 : 1;
    const uint32_t lanes = len >= 64 ? 64 : len;
    const uint64_t first = (chunk << chunk_log) + uint64_t(threadIdx.x) * per;
    double mine = 0.0;
    if (threadIdx.x < lanes)
        for (uint32_t i = 0; i < per; ++i) mine += cnorm(s[first + i]);
    // inclusive wave scan of lane sums
    double scan = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double up = __shfl_up(scan, off, 64);
        if (static_cast<int>(threadIdx.x) >= off) scan += up;
    }
    const bool over = threadIdx.x < lanes && before + scan > r;
    const uint64_t ballot = __ballot(over);
    if (ballot == 0) {  // rounding put r past this chunk's own sum: last index of chunk
        if (threadIdx.x == 0) out[d] = ((chunk + 1) << chunk_log) - 1;
        return;
    }
    const uint32_t lane = __builtin_ctzll(ballot);
    if (threadIdx.x == lane) {
        double run = before + scan - mine;
        uint64_t idx = first + per - 1;
        for (uint32_t i = 0; i < per; ++i) {
            run += cnorm(s[first + i]);
            if (run > r) { idx = first + i; break; }
        }
        out[d] = idx;
    }
}

// ---- collapse / reset -----------------------------------------------------------
// Per register: p1 = sum over amplitudes with bit q set.
__global__ void __launch_bounds__(kBlock) k_p1_partial(const amp_t* __restrict__ s, uint32_t nq,
                                                       uint32_t q, double* __restrict__ part) {
    const amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint64_t half = uint64_t{1} << (nq - 1);
    double acc[1] = {0.0};
    for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < half;
         g += uint64_t(gridDim.x) * kBlock)
        acc[0] += cnorm(b[insert_zero_bit(g, q) | (uint64_t{1} << q)]);
    block_sum(acc);
    if (threadIdx.x == 0) part[uint64_t(blockIdx.y) * gridDim.x + blockIdx.x] = acc[0];
}

// outcome = u < p1; keep that branch scaled by 1/sqrt(p_keep), zero the other;
// with `reset`, a |1> outcome is moved back to |0>.
__global__ void __launch_bounds__(kBlock) k_collapse(amp_t* __restrict__ s, uint32_t nq, uint32_t q,
                                                     const double* __restrict__ p1,
                                                     const double* __restrict__ draws,
                                                     uint64_t draw_stride, uint64_t draw_offset,
                                                     int reset, uint8_t* __restrict__ creg,
                                                     uint32_t creg_stride, int32_t clbit) {
    const uint32_t r = blockIdx.y;
    amp_t* b = s + (uint64_t(r) << nq);
    const double p = p1[r];
    const double u = draws[uint64_t(r) * draw_stride + draw_offset];
    const bool one = u < p;
    const double scale = 1.0 / sqrt(one ? p : 1.0 - p);
    const uint64_t half = uint64_t{1} << (nq - 1);
    const uint64_t g0 = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (g0 == 0 && clbit >= 0 && creg) creg[uint64_t(r) * creg_stride + clbit] = one ? 1 : 0;
    const amp_t z = make_double2(0.0, 0.0);
    for (uint64_t g = g0; g < half; g += uint64_t(gridDim.x) * kBlock) {
        const uint64_t i0 = insert_zero_bit(g, q), i1 = i0 | (uint64_t{1} << q);
        const amp_t a0 = b[i0], a1 = b[i1];
        if (one) {
            if (reset) { b[i0] = cscale(a1, scale); b[i1] = z; }
            else { b[i0] = z; b[i1] = cscale(a1, scale); }
        } else {
            b[i0] = cscale(a0, scale);
            b[i1] = z;
        }
    }
}

// ---- 2x2 single-qubit statistics for Kraus probabilities --------------------------
// S00 = sum|a0|^2, S11 = sum|a1|^2, S01 = sum conj(a0) a1 over pairs of qubit q.
__global__ void __launch_bounds__(kBlock) k_qstats_partial(const amp_t* __restrict__ s,
                                                           uint32_t nq, uint32_t q,
                                                           double* __restrict__ part) {
    const amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint64_t half = uint64_t{1} << (nq - 1);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < half;
         g += uint64_t(gridDim.x) * kBlock) {
        const uint64_t i0 = insert_zero_bit(g, q);
        const amp_t a0 = b[i0], a1 = b[i0 | (uint64_t{1} << q)];
        acc[0] += cnorm(a0);
        acc[1] += cnorm(a1);
        acc[2] += a0.x * a1.x + a0.y * a1.y;
        acc[3] += a0.x * a1.y - a0.y * a1.x;
    }
    block_sum(acc);
    if (threadIdx.x == 0) {
        double* o = part + (uint64_t(blockIdx.y) * gridDim.x + blockIdx.x) * 4;
        o[0] = acc[0];
        o[1] = acc[1];
        o[2] = acc[2];
        o[3] = acc[3];
    }
}

__global__ void __launch_bounds__(kBlock) k_probabilities(const amp_t* __restrict__ s, uint64_t n,
                                                          double* __restrict__ out) {
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * kBlock)
        out[i] = cnorm(s[i]);
}

unsigned red_blocks(uint32_t nq, uint32_t shift = 0) {
    const uint64_t items = (uint64_t{1} << nq) >> shift;
    const uint64_t want = (items + kBlock * 8 - 1) / (kBlock * 8);
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(kRedBlocks, want)));
}

// ---- density matrices held as 2n-qubit registers: rho[r][c] at r | (c << n) ----------
// (reference include/lqcs/backend/density_matrix.hpp: ket bits low, bra bits high)
__global__ void __launch_bounds__(kBlock) k_dm_diagonal(const amp_t* __restrict__ s, uint32_t n,
                                                        double* __restrict__ out) {
    const uint64_t d = uint64_t{1} << n;
    for (uint64_t r = uint64_t(blockIdx.x) * kBlock + threadIdx.x; r < d; r += uint64_t(gridDim.x) * kBlock)
        out[r] = s[(r << n) | r].x;
}

// Partial sums of Re(conj(psi_r) rho[r][c] psi_c) over all (r, c).
__global__ void __launch_bounds__(kBlock) k_dm_fidelity_partial(const amp_t* __restrict__ rho,
                                                                const amp_t* __restrict__ psi, uint32_t n,
                                                                double* __restrict__ part) {
    const uint64_t total = uint64_t{1} << (2 * n), mask = (uint64_t{1} << n) - 1;
    double acc[1] = {0.0};
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < total; i += uint64_t(gridDim.x) * kBlock) {
        const amp_t t = cmul(cmul(cconj(psi[i & mask]), rho[i]), psi[i >> n]);
        acc[0] += t.x;
    }
    block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

}  // namespace

void launch_dm_diagonal(const amp_t* rho, uint32_t n, double* d_out) {
    const unsigned g = std::min<unsigned>(grid_for(uint64_t{1} << n), 4096);
    k_dm_diagonal<<<g, kBlock, 0, stream()>>>(rho, n, d_out);
}

void reduce_dm_fidelity(const amp_t* rho, const amp_t* psi, uint32_t n, double* d_out) {
    const unsigned g = red_blocks(2 * n);
    double* part = static_cast<double*>(arena(sizeof(double) * g));
    k_dm_fidelity_partial<<<g, kBlock, 0, stream()>>>(rho, psi, n, part);
    k_finish<1><<<1, kBlock, 0, stream()>>>(part, g, d_out);
}

void reduce_norm_sq(const amp_t* s, uint32_t nq, uint32_t batch, double* d_out) {
    const unsigned g = red_blocks(nq);
    double* part = static_cast<double*>(arena(sizeof(double) * g * batch));
    k_norm_partial<<<dim3(g, batch), kBlock, 0, stream()>>>(s, nq, part);
    k_finish<1><<<batch, kBlock, 0, stream()>>>(part, g, d_out);
}

void reduce_max_abs_diff(const amp_t* a, const amp_t* b, uint32_t nq, double* d_out) {
    const unsigned g = red_blocks(nq);
    double* part = static_cast<double*>(arena(sizeof(double) * g));
    k_maxdiff_partial<<<g, kBlock, 0, stream()>>>(a, b, uint64_t{1} << nq, part);
    k_max_finish<<<1, kBlock, 0, stream()>>>(part, g, d_out);
}

void reduce_pauli_terms(const amp_t* s, uint32_t nq, const uint64_t* d_x, const uint64_t* d_zy,
                        uint32_t n_terms, double* d_out) {
    double* part = static_cast<double*>(arena(sizeof(double) * pauli_term_partials(nq) * n_terms));
    reduce_pauli_terms_with(s, nq, d_x, d_zy, n_terms, part, d_out);
}

uint64_t pauli_term_partials(uint32_t nq) { return 2 * uint64_t{red_blocks(nq, 1)}; }

void reduce_pauli_terms_with(const amp_t* s, uint32_t nq, const uint64_t* d_x, const uint64_t* d_zy,
                             uint32_t n_terms, double* part, double* d_out) {
    const unsigned g = red_blocks(nq, 1);
    k_pauli_expect_partial<<<dim3(g, n_terms), kBlock, 0, stream()>>>(s, nq, d_x, d_zy, part);
    k_finish<2><<<n_terms, kBlock, 0, stream()>>>(part, g, d_out);
}

void reduce_p1(const amp_t* s, uint32_t nq, uint32_t q, uint32_t batch, double* d_out) {
    const unsigned g = red_blocks(nq, 1);
    double* part = static_cast<double*>(arena(sizeof(double) * g * batch));
    k_p1_partial<<<dim3(g, batch), kBlock, 0, stream()>>>(s, nq, q, part);
    k_finish<1><<<batch, kBlock, 0, stream()>>>(part, g, d_out);
}

void reduce_qstats(const amp_t* s, uint32_t nq, uint32_t q, uint32_t batch, double* d_out) {
    const unsigned g = red_blocks(nq, 1);
    double* part = static_cast<double*>(arena(sizeof(double) * 4 * g * batch));
    k_qstats_partial<<<dim3(g, batch), kBlock, 0, stream()>>>(s, nq, q, part);
    k_finish<4><<<batch, kBlock, 0, stream()>>>(part, g, d_out);
}

void launch_collapse(amp_t* s, uint32_t nq, uint32_t q, const double* d_p1, const double* d_draws,
                     uint64_t draw_stride, uint64_t draw_offset, int reset, uint8_t* d_creg,
                     uint32_t creg_stride, int32_t clbit, uint32_t batch) {
    k_collapse<<<dim3(std::min<unsigned>(grid_for(uint64_t{1} << (nq - 1)), 1u << 22), batch), kBlock, 0, stream()>>>(
        s, nq, q, d_p1, d_draws, draw_stride, draw_offset, reset, d_creg, creg_stride, clbit);
}

void launch_probabilities(const amp_t* s, uint32_t nq, double* d_out) {
    const uint64_t n = uint64_t{1} << nq;
    k_probabilities<<<std::min<unsigned>(grid_for(n), 65536), kBlock, 0, stream()>>>(s, n, d_out);
}

void launch_sample(const amp_t* s, uint32_t nq, const double* d_draws, uint64_t n_draws,
                   uint64_t* d_out) {
    const uint32_t chunk_log = nq < static_cast<uint32_t>(kChunkLog) ? nq : kChunkLog;
    const uint64_t n_chunks = uint64_t{1} << (nq - chunk_log);
    double* sums = static_cast<double*>(arena(sizeof(double) * 2 * n_chunks));
    double* incl = sums + n_chunks;
    k_chunk_sums<<<static_cast<unsigned>(n_chunks), kBlock, 0, stream()>>>(s, chunk_log, sums);
    k_scan_chunks<<<1, 1024, 0, stream()>>>(sums, n_chunks, incl);
    if (n_draws)
        k_sample<<<static_cast<unsigned>(n_draws), 64, 0, stream()>>>(s, nq, chunk_log, incl,
                                                                     n_chunks, d_draws, n_draws,
                                                                     d_out);
}

}  // namespace lqcs_dev
