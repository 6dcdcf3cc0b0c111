// Allocation-free Hamiltonian expectation: the entry point
// lqcs_state_expectation_terms_noalloc computes what lqcs_state_expectation_terms
// computes (sum_t coeff_t Re(i^ny_t <psi|P_t|psi>)) with every piece of device
// memory beyond the register taken from one file-scope array, so it never
// calls hipMalloc or hipFree. It serves callers that cannot allocate (code
// under HIP graph capture, tight serving loops).
//
// The array holds, in order: the staged X and Z/Y masks, coefficients and Y
// counts of up to kMaxTerms terms, the per-block partial sums of every term,
// the per-term sums and the final value. Each GPU holds its own copy, at its
// own address, looked up once per device with hipGetSymbolAddress. Nothing in
// it accumulates across calls (every slot a call reads was written earlier in
// the same call), so nothing is cleared. Every call shares the array: calls
// hold the engine lock and return only after their work has completed, so
// they never overlap on one GPU; work that did overlap would race on it.
#include "common.hpp"
#include "engine.hpp"
#include "reduce.hpp"

#include <mutex>

namespace lqcs_dev {
namespace {

constexpr uint64_t kMaxTerms = LQCS_NOALLOC_MAX_TERMS;
constexpr uint64_t kMaxPartials = 2 * 1024;  // pauli_term_partials(nq) is at most this
constexpr uint64_t kOffX = 0;
constexpr uint64_t kOffZ = kOffX + kMaxTerms;
constexpr uint64_t kOffC = kOffZ + kMaxTerms;
constexpr uint64_t kOffY = kOffC + kMaxTerms;
constexpr uint64_t kOffPart = kOffY + kMaxTerms;
constexpr uint64_t kOffSum = kOffPart + kMaxPartials * kMaxTerms;
constexpr uint64_t kOffOut = kOffSum + 2 * kMaxTerms;
constexpr uint64_t kScratchWords = kOffOut + 32;  // 8-byte words, about 8 MiB at 512 terms

static __device__ __align__(256) double scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// The array's device address on the current device (looked up once per device).
double* scratch_ptr() {
    int dev = 0;
    LQCS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(LQCS_ERR_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        LQCS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_ptr[dev] = static_cast<double*>(p);
    }
    return g_ptr[dev];
}

// out = sum_t coeff_t * Re(i^ny_t * (re_t + i im_t)), in term order, with
// separate rounding of each product and sum (as the host combine does).
__global__ void k_combine_terms(const double* __restrict__ sums, const double* __restrict__ coeffs,
                                const uint32_t* __restrict__ n_ys, uint32_t n_terms, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double e = 0.0;
    for (uint32_t t = 0; t < n_terms; ++t) {
        const double re = sums[2 * t], im = sums[2 * t + 1];
        double v;
        switch (n_ys[t] % 4) {
            case 0: v = re; break;
            case 1: v = -im; break;
            case 2: v = -re; break;
            default: v = im; break;
        }
        e = __dadd_rn(e, __dmul_rn(coeffs[t], v));
    }
    *out = e;
}

}  // namespace
}  // namespace lqcs_dev

using namespace lqcs_dev;

extern "C" lqcs_status lqcs_state_expectation_terms_noalloc(const lqcs_dev_state* s, const uint64_t* x_masks,
                                                           const uint64_t* zy_masks, const uint32_t* n_ys,
                                                           const double* coeffs, uint64_t n_terms, double* out) {
    return run_guarded([&] {
        if (!s) throw Error(LQCS_ERR_INVALID_ARGUMENT, "state is null");
        if (!out) throw Error(LQCS_ERR_INVALID_ARGUMENT, "out is null");
        *out = 0.0;
        if (n_terms == 0) return;
        if (n_terms > kMaxTerms)
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "expectation_terms_noalloc: at most " +
                                                       std::to_string(kMaxTerms) + " terms");
        if (!x_masks || !zy_masks || !n_ys || !coeffs) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null term array");
        const uint64_t full = (uint64_t{1} << s->nq) - 1;
        for (uint64_t t = 0; t < n_terms; ++t)
            if ((x_masks[t] | zy_masks[t]) & ~full) throw Error(LQCS_ERR_OUT_OF_RANGE, "pauli mask out of range");
        if (pauli_term_partials(s->nq) > kMaxPartials) throw Error(LQCS_ERR_LOGIC, "partial-sum count beyond scratch");
        ensure_ready();
        double* w = scratch_ptr();
        auto* dx = reinterpret_cast<uint64_t*>(w + kOffX);
        auto* dz = reinterpret_cast<uint64_t*>(w + kOffZ);
        auto* dy = reinterpret_cast<uint32_t*>(w + kOffY);
        double* dc = w + kOffC;
        const hipStream_t st = stream();
        LQCS_HIP(hipMemcpyAsync(dx, x_masks, n_terms * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        LQCS_HIP(hipMemcpyAsync(dz, zy_masks, n_terms * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        LQCS_HIP(hipMemcpyAsync(dc, coeffs, n_terms * sizeof(double), hipMemcpyHostToDevice, st));
        LQCS_HIP(hipMemcpyAsync(dy, n_ys, n_terms * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        reduce_pauli_terms_with(s->d, s->nq, dx, dz, static_cast<uint32_t>(n_terms), w + kOffPart, w + kOffSum);
        k_combine_terms<<<1, 64, 0, st>>>(w + kOffSum, dc, dy, static_cast<uint32_t>(n_terms), w + kOffOut);
        LQCS_HIP(hipGetLastError());
        LQCS_HIP(hipMemcpyAsync(out, w + kOffOut, sizeof(double), hipMemcpyDeviceToHost, st));
        LQCS_HIP(hipStreamSynchronize(st));
    });
}
