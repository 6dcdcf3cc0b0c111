// Cost-model network: a multilayer perceptron F -> 64 -> 64 -> 1 with ReLU,
// in FP64, trained and evaluated on the device by the kernels below.
//
// Parameters are one packed array:
//   W1 [F][64] (input-major), b1 [64], W2 [64][64], b2 [64], w3 [64], b3 [1].
// Loss: mean squared error over the rows, L = (1/B) sum_r (y_r - t_r)^2.
// Every reduction runs in a fixed order (one thread per output, rows in
// index order), so gradients are reproducible run to run.
#include "common.hpp"
#include "engine.hpp"

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace lqcs_dev {
namespace {

constexpr int kH = 64;  // hidden width

struct Layout {
    uint32_t F;
    size_t w1, b1, w2, b2, w3, b3, n;
};

Layout layout(uint32_t F) {
    Layout l;
    l.F = F;
    l.w1 = 0;
    l.b1 = l.w1 + size_t(F) * kH;
    l.w2 = l.b1 + kH;
    l.b2 = l.w2 + size_t(kH) * kH;
    l.w3 = l.b2 + kH;
    l.b3 = l.w3 + kH;
    l.n = l.b3 + 1;
    return l;
}

// Z[r][o] = b[o] + sum_i X[r][i] W[i][o]; A = max(Z, 0). Block = one row.
__global__ void __launch_bounds__(kH) k_dense_relu(const double* __restrict__ X, uint32_t in,
                                                   const double* __restrict__ W, const double* __restrict__ b,
                                                   double* __restrict__ Z, double* __restrict__ A) {
    const uint64_t r = blockIdx.x;
    const uint32_t o = threadIdx.x;
    double z = b[o];
    for (uint32_t i = 0; i < in; ++i) z += X[r * in + i] * W[size_t(i) * kH + o];
    Z[r * kH + o] = z;
    A[r * kH + o] = z > 0.0 ? z : 0.0;
}

// y[r] = b3 + sum_o A[r][o] w3[o]; one thread per row.
__global__ void __launch_bounds__(kBlock) k_out(const double* __restrict__ A, const double* __restrict__ w3,
                                                const double* __restrict__ b3, uint64_t rows,
                                                double* __restrict__ y) {
    const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double acc = b3[0];
    for (int o = 0; o < kH; ++o) acc += A[r * kH + o] * w3[o];
    y[r] = acc;
}

// dy[r] = 2 (y[r] - t[r]) / B.
__global__ void __launch_bounds__(kBlock) k_mse_grad(const double* __restrict__ y, const double* __restrict__ t,
                                                     uint64_t rows, double* __restrict__ dy) {
    const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    dy[r] = 2.0 * (y[r] - t[r]) / double(rows);
}

// loss = (1/B) sum (y - t)^2, one block, rows strided by thread, then a
// fixed-order tree.
__global__ void __launch_bounds__(kBlock) k_mse(const double* __restrict__ y, const double* __restrict__ t,
                                                uint64_t rows, double* __restrict__ loss) {
    __shared__ double part[kBlock];
    double acc = 0.0;
    for (uint64_t r = threadIdx.x; r < rows; r += kBlock) {
        const double d = y[r] - t[r];
        acc += d * d;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = part[0] / double(rows);
}

// Output layer backward: dw3[o] = sum_r A2[r][o] dy[r] (o < 64), db3 = sum_r dy[r]
// (o = 64); and dZ2[r][o] = dy[r] w3[o] [Z2 > 0].
__global__ void __launch_bounds__(kH + 1) k_out_grad(const double* __restrict__ A2, const double* __restrict__ dy,
                                                     uint64_t rows, double* __restrict__ dw3,
                                                     double* __restrict__ db3) {
    const uint32_t o = threadIdx.x;
    double acc = 0.0;
    if (o < kH) {
        for (uint64_t r = 0; r < rows; ++r) acc += A2[r * kH + o] * dy[r];
        dw3[o] = acc;
    } else {
        for (uint64_t r = 0; r < rows; ++r) acc += dy[r];
        db3[0] = acc;
    }
}

__global__ void __launch_bounds__(kH) k_out_back(const double* __restrict__ Z2, const double* __restrict__ dy,
                                                 const double* __restrict__ w3, double* __restrict__ dZ2) {
    const uint64_t r = blockIdx.x;
    const uint32_t o = threadIdx.x;
    dZ2[r * kH + o] = Z2[r * kH + o] > 0.0 ? dy[r] * w3[o] : 0.0;
}

// Weight gradient of a dense layer: dW[i][o] = sum_r X[r][i] dZ[r][o] and
// db[o] = sum_r dZ[r][o]. Block = one input i (blockIdx.x == in: the bias).
__global__ void __launch_bounds__(kH) k_dense_wgrad(const double* __restrict__ X, uint32_t in,
                                                    const double* __restrict__ dZ, uint64_t rows,
                                                    double* __restrict__ dW, double* __restrict__ db) {
    const uint32_t i = blockIdx.x;
    const uint32_t o = threadIdx.x;
    double acc = 0.0;
    if (i < in) {
        for (uint64_t r = 0; r < rows; ++r) acc += X[r * in + i] * dZ[r * kH + o];
        dW[size_t(i) * kH + o] = acc;
    } else {
        for (uint64_t r = 0; r < rows; ++r) acc += dZ[r * kH + o];
        db[o] = acc;
    }
}

// Input gradient through a hidden layer: dZp[r][i] = (sum_o dZ[r][o] W[i][o]) [Zp[r][i] > 0].
__global__ void __launch_bounds__(kH) k_dense_back(const double* __restrict__ dZ, const double* __restrict__ W,
                                                   const double* __restrict__ Zp, double* __restrict__ dZp) {
    const uint64_t r = blockIdx.x;
    const uint32_t i = threadIdx.x;
    double acc = 0.0;
    for (int o = 0; o < kH; ++o) acc += dZ[r * kH + o] * W[size_t(i) * kH + o];
    dZp[r * kH + i] = Zp[r * kH + i] > 0.0 ? acc : 0.0;
}

// Adam (Kingma & Ba, 2015): m <- b1 m + (1-b1) g; v <- b2 v + (1-b2) g^2;
// p <- p - lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
__global__ void __launch_bounds__(kBlock) k_adam(double* __restrict__ p, const double* __restrict__ g,
                                                 double* __restrict__ m, double* __restrict__ v, uint64_t n,
                                                 double lr, double c1, double c2) {
    const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    constexpr double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    const double gk = g[k];
    const double mk = b1 * m[k] + (1.0 - b1) * gk;
    const double vk = b2 * v[k] + (1.0 - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrt(vk / c2) + eps);
}

// Rows idx[k] of X (in columns) and t into contiguous minibatch buffers.
__global__ void __launch_bounds__(kBlock) k_gather_rows(const double* __restrict__ X, const double* __restrict__ t,
                                                        uint32_t in, const uint64_t* __restrict__ idx,
                                                        uint64_t rows, double* __restrict__ Xb,
                                                        double* __restrict__ tb) {
    const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= rows * in) return;
    const uint64_t r = k / in, c = k % in;
    Xb[k] = X[idx[r] * in + c];
    if (c == 0) tb[r] = t[idx[r]];
}

// Device buffers of one forward/backward evaluation over `rows` rows.
struct Work {
    uint64_t rows = 0;
    double *z1 = nullptr, *a1 = nullptr, *z2 = nullptr, *a2 = nullptr, *y = nullptr, *dy = nullptr;
    double *dz2 = nullptr, *dz1 = nullptr, *loss = nullptr;
    explicit Work(uint64_t r) : rows(r) {
        z1 = static_cast<double*>(device_alloc(r * kH * sizeof(double)));
        a1 = static_cast<double*>(device_alloc(r * kH * sizeof(double)));
        z2 = static_cast<double*>(device_alloc(r * kH * sizeof(double)));
        a2 = static_cast<double*>(device_alloc(r * kH * sizeof(double)));
        dz2 = static_cast<double*>(device_alloc(r * kH * sizeof(double)));
        dz1 = static_cast<double*>(device_alloc(r * kH * sizeof(double)));
        y = static_cast<double*>(device_alloc(r * sizeof(double)));
        dy = static_cast<double*>(device_alloc(r * sizeof(double)));
        loss = static_cast<double*>(device_alloc(sizeof(double)));
    }
    ~Work() {
        for (double* p : {z1, a1, z2, a2, dz2, dz1, y, dy, loss}) device_free(p);
    }
    Work(const Work&) = delete;
    Work& operator=(const Work&) = delete;
};

void forward(const Layout& L, const double* P, const double* X, Work& w) {
    const unsigned rows = static_cast<unsigned>(w.rows);
    k_dense_relu<<<rows, kH, 0, stream()>>>(X, L.F, P + L.w1, P + L.b1, w.z1, w.a1);
    k_dense_relu<<<rows, kH, 0, stream()>>>(w.a1, kH, P + L.w2, P + L.b2, w.z2, w.a2);
    k_out<<<grid_for(w.rows), kBlock, 0, stream()>>>(w.a2, P + L.w3, P + L.b3, w.rows, w.y);
    LQCS_HIP(hipGetLastError());
}

// Loss and gradient (packed like the parameters) of the rows in X, t.
void backward(const Layout& L, const double* P, const double* X, const double* t, Work& w, double* G) {
    const unsigned rows = static_cast<unsigned>(w.rows);
    k_mse<<<1, kBlock, 0, stream()>>>(w.y, t, w.rows, w.loss);
    k_mse_grad<<<grid_for(w.rows), kBlock, 0, stream()>>>(w.y, t, w.rows, w.dy);
    k_out_grad<<<1, kH + 1, 0, stream()>>>(w.a2, w.dy, w.rows, G + L.w3, G + L.b3);
    k_out_back<<<rows, kH, 0, stream()>>>(w.z2, w.dy, P + L.w3, w.dz2);
    k_dense_wgrad<<<kH + 1, kH, 0, stream()>>>(w.a1, kH, w.dz2, w.rows, G + L.w2, G + L.b2);
    k_dense_back<<<rows, kH, 0, stream()>>>(w.dz2, P + L.w2, w.z1, w.dz1);
    k_dense_wgrad<<<L.F + 1, kH, 0, stream()>>>(X, L.F, w.dz1, w.rows, G + L.w1, G + L.b1);
    LQCS_HIP(hipGetLastError());
}

void check_shape(uint32_t F, uint64_t rows) {
    if (F == 0 || F > 256) throw Error(LQCS_ERR_INVALID_ARGUMENT, "n_features must be 1..256");
    if (rows == 0 || rows > (uint64_t{1} << 24)) throw Error(LQCS_ERR_INVALID_ARGUMENT, "n_rows must be 1..2^24");
}

template <typename T>
T* to_device(const T* h, size_t n) {
    T* d = static_cast<T*>(device_alloc(n * sizeof(T)));
    LQCS_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, stream()));
    return d;
}

// ---- installed cost model (read by the planner) ----
struct CostModel {
    bool on = false;
    uint32_t F = 0;
    std::vector<double> params, x_mean, x_scale;
    double t_mean = 0.0, t_scale = 1.0;
};
// Two models: one for passes on the interpreting register tile, one for
// specialized passes (whose costs differ); the planner consults the one of
// the kernel a call selects.
CostModel& cost_model_slot(bool specialized) {
    static CostModel m[2];
    return m[specialized ? 1 : 0];
}
CostModel& cost_model() { return cost_model_slot(g_rtile_spec); }

}  // namespace

// Counts installs and removals, so that decisions taken with one model are
// not reused under another.
uint64_t& cost_model_counter() {
    static uint64_t n = 0;
    return n;
}

bool cost_model_installed() { return cost_model().on; }
uint64_t cost_model_version() { return cost_model_counter(); }
uint32_t cost_model_features() { return cost_model().F; }

// Predicted milliseconds of each feature row (raw features, normalised here),
// evaluated on the device by the same forward kernels that train the model.
std::vector<double> cost_model_predict_ms(const std::vector<double>& raw, uint64_t rows) {
    const CostModel& cm = cost_model();
    const Layout L = layout(cm.F);
    std::vector<double> x(raw.size());
    for (uint64_t r = 0; r < rows; ++r)
        for (uint32_t f = 0; f < cm.F; ++f)
            x[r * cm.F + f] = (raw[r * cm.F + f] - cm.x_mean[f]) / cm.x_scale[f];
    double* dP = to_device(cm.params.data(), L.n);
    double* dX = to_device(x.data(), x.size());
    std::vector<double> y(rows);
    {
        Work w(rows);
        forward(L, dP, dX, w);
        LQCS_HIP(hipMemcpyAsync(y.data(), w.y, rows * sizeof(double), hipMemcpyDeviceToHost, stream()));
        sync();
    }
    device_free(dP);
    device_free(dX);
    for (double& v : y) v = std::exp(v * cm.t_scale + cm.t_mean);
    return y;
}

}  // namespace lqcs_dev

using namespace lqcs_dev;

lqcs_status lqcs_mlp_num_params(uint32_t n_features, uint64_t* n_params) {
    return run_guarded([&] {
        if (!n_params) throw Error(LQCS_ERR_INVALID_ARGUMENT, "n_params is null");
        check_shape(n_features, 1);
        *n_params = layout(n_features).n;
    });
}

lqcs_status lqcs_mlp_forward(uint32_t n_features, const double* params, const double* x, uint64_t n_rows,
                             double* y) {
    return run_guarded([&] {
        if (!params || !x || !y) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
        check_shape(n_features, n_rows);
        ensure_ready();
        const Layout L = layout(n_features);
        double* dP = to_device(params, L.n);
        double* dX = to_device(x, n_rows * n_features);
        {
            Work w(n_rows);
            forward(L, dP, dX, w);
            LQCS_HIP(hipMemcpyAsync(y, w.y, n_rows * sizeof(double), hipMemcpyDeviceToHost, stream()));
            sync();
        }
        device_free(dP);
        device_free(dX);
    });
}

lqcs_status lqcs_mlp_loss_grad(uint32_t n_features, const double* params, const double* x, const double* t,
                               uint64_t n_rows, double* loss, double* grad) {
    return run_guarded([&] {
        if (!params || !x || !t || !loss || !grad) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
        check_shape(n_features, n_rows);
        ensure_ready();
        const Layout L = layout(n_features);
        double* dP = to_device(params, L.n);
        double* dX = to_device(x, n_rows * n_features);
        double* dT = to_device(t, n_rows);
        double* dG = static_cast<double*>(device_alloc(L.n * sizeof(double)));
        {
            Work w(n_rows);
            forward(L, dP, dX, w);
            backward(L, dP, dX, dT, w, dG);
            LQCS_HIP(hipMemcpyAsync(grad, dG, L.n * sizeof(double), hipMemcpyDeviceToHost, stream()));
            LQCS_HIP(hipMemcpyAsync(loss, w.loss, sizeof(double), hipMemcpyDeviceToHost, stream()));
            sync();
        }
        for (double* p : {dP, dX, dT, dG}) device_free(p);
    });
}

lqcs_status lqcs_mlp_train(uint32_t n_features, double* params, const double* x, const double* t, uint64_t n_rows,
                           uint32_t epochs, uint32_t batch_rows, double lr, uint64_t seed, double* final_loss) {
    return run_guarded([&] {
        if (!params || !x || !t) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
        check_shape(n_features, n_rows);
        if (batch_rows == 0 || batch_rows > n_rows) batch_rows = static_cast<uint32_t>(std::min<uint64_t>(n_rows, 256));
        if (!(lr > 0.0)) throw Error(LQCS_ERR_INVALID_ARGUMENT, "learning rate must be positive");
        ensure_ready();
        const Layout L = layout(n_features);
        double* dP = to_device(params, L.n);
        double* dX = to_device(x, n_rows * n_features);
        double* dT = to_device(t, n_rows);
        double* dG = static_cast<double*>(device_alloc(L.n * sizeof(double)));
        double* dM = static_cast<double*>(device_alloc(L.n * sizeof(double)));
        double* dV = static_cast<double*>(device_alloc(L.n * sizeof(double)));
        double* dXb = static_cast<double*>(device_alloc(size_t(batch_rows) * n_features * sizeof(double)));
        double* dTb = static_cast<double*>(device_alloc(size_t(batch_rows) * sizeof(double)));
        uint64_t* dIdx = static_cast<uint64_t*>(device_alloc(size_t(batch_rows) * sizeof(uint64_t)));
        LQCS_HIP(hipMemsetAsync(dM, 0, L.n * sizeof(double), stream()));
        LQCS_HIP(hipMemsetAsync(dV, 0, L.n * sizeof(double), stream()));
        std::mt19937_64 rng(seed);
        std::vector<uint64_t> perm(n_rows);
        for (uint64_t i = 0; i < n_rows; ++i) perm[i] = i;
        uint64_t step = 0;
        {
            Work w(batch_rows);
            for (uint32_t e = 0; e < epochs; ++e) {
                for (uint64_t i = n_rows - 1; i > 0; --i) std::swap(perm[i], perm[rng() % (i + 1)]);
                for (uint64_t s0 = 0; s0 + batch_rows <= n_rows; s0 += batch_rows) {
                    LQCS_HIP(hipMemcpyAsync(dIdx, perm.data() + s0, batch_rows * sizeof(uint64_t),
                                            hipMemcpyHostToDevice, stream()));
                    k_gather_rows<<<grid_for(uint64_t(batch_rows) * n_features), kBlock, 0, stream()>>>(
                        dX, dT, n_features, dIdx, batch_rows, dXb, dTb);
                    forward(L, dP, dXb, w);
                    backward(L, dP, dXb, dTb, w, dG);
                    ++step;
                    const double c1 = 1.0 - std::pow(0.9, double(step));
                    const double c2 = 1.0 - std::pow(0.999, double(step));
                    k_adam<<<grid_for(L.n), kBlock, 0, stream()>>>(dP, dG, dM, dV, L.n, lr, c1, c2);
                    LQCS_HIP(hipGetLastError());
                    // the host permutation buffer is reused next step: keep the copy ordered
                    sync();
                }
            }
        }
        if (final_loss) {  // loss over every row with the trained parameters
            Work w(n_rows);
            forward(L, dP, dX, w);
            k_mse<<<1, kBlock, 0, stream()>>>(w.y, dT, n_rows, w.loss);
            LQCS_HIP(hipMemcpyAsync(final_loss, w.loss, sizeof(double), hipMemcpyDeviceToHost, stream()));
        }
        LQCS_HIP(hipMemcpyAsync(params, dP, L.n * sizeof(double), hipMemcpyDeviceToHost, stream()));
        sync();
        for (void* p : {static_cast<void*>(dP), static_cast<void*>(dX), static_cast<void*>(dT),
                        static_cast<void*>(dG), static_cast<void*>(dM), static_cast<void*>(dV),
                        static_cast<void*>(dXb), static_cast<void*>(dTb), static_cast<void*>(dIdx)})
            device_free(p);
    });
}

namespace {
lqcs_status install_model(bool specialized, uint32_t n_features, const double* params, const double* x_mean,
                          const double* x_scale, double t_mean, double t_scale) {
    return run_guarded([&] {
        CostModel& cm = cost_model_slot(specialized);
        ++cost_model_counter();
        if (!params) {
            cm = CostModel{};
            return;
        }
        if (n_features != LQCS_COST_FEATURES)
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "cost model must take LQCS_COST_FEATURES features");
        if (!x_mean || !x_scale) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null normalisation");
        const Layout L = layout(n_features);
        for (uint32_t f = 0; f < n_features; ++f)
            if (!(x_scale[f] > 0.0)) throw Error(LQCS_ERR_INVALID_ARGUMENT, "feature scales must be positive");
        if (!(t_scale > 0.0)) throw Error(LQCS_ERR_INVALID_ARGUMENT, "target scale must be positive");
        cm.on = true;
        cm.F = n_features;
        cm.params.assign(params, params + L.n);
        cm.x_mean.assign(x_mean, x_mean + n_features);
        cm.x_scale.assign(x_scale, x_scale + n_features);
        cm.t_mean = t_mean;
        cm.t_scale = t_scale;
    });
}
}  // namespace

lqcs_status lqcs_costmodel_install(uint32_t n_features, const double* params, const double* x_mean,
                                   const double* x_scale, double t_mean, double t_scale) {
    return install_model(false, n_features, params, x_mean, x_scale, t_mean, t_scale);
}

lqcs_status lqcs_costmodel_install_kernel(int32_t tile_kernel, uint32_t n_features, const double* params,
                                          const double* x_mean, const double* x_scale, double t_mean,
                                          double t_scale) {
    if (tile_kernel != LQCS_TILE_REGISTER && tile_kernel != LQCS_TILE_SPECIALIZED)
        return run_guarded([] {
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "cost models exist for LQCS_TILE_REGISTER and LQCS_TILE_SPECIALIZED");
        });
    return install_model(tile_kernel == LQCS_TILE_SPECIALIZED, n_features, params, x_mean, x_scale, t_mean, t_scale);
}
