// Learned cost model of the autotuner (symcost.h): FP32 MLP trained and evaluated by HIP
// kernels written here and compiled by hipRTC.
#include "symcost.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

// One thread per output value in every kernel; loops run over the (small) reduced dimension
// in a fixed order, so training is deterministic for a given seed.
const char* kSource = R"(
extern "C" __global__ void mlp_layer(const float* __restrict__ in, int inDim, const float* __restrict__ W,
                                     const float* __restrict__ b, float* __restrict__ out, int outDim, int n, int relu) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * outDim) return;
    const int s = idx / outDim, j = idx - s * outDim;
    const float* __restrict__ x = in + (size_t)s * inDim;
    const float* __restrict__ w = W + (size_t)j * inDim;
    float acc = b[j];
    for (int k = 0; k < inDim; ++k) acc = fmaf(w[k], x[k], acc);
    out[idx] = relu ? fmaxf(acc, 0.0f) : acc;
}

extern "C" __global__ void mlp_loss_grad(const float* __restrict__ Y, const float* __restrict__ t, float* __restrict__ dY,
                                         float* __restrict__ sq, int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float e = Y[s] - t[s];
    dY[s] = 2.0f * e / (float)n;
    sq[s] = e * e;
}

extern "C" __global__ void mlp_grad_w(const float* __restrict__ dOut, int outDim, const float* __restrict__ In, int inDim,
                                      int n, float* __restrict__ dW, float* __restrict__ db) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= outDim * (inDim + 1)) return;
    const int o = idx / (inDim + 1), i = idx - o * (inDim + 1);
    float acc = 0.0f;
    if (i < inDim) {
        for (int s = 0; s < n; ++s) acc = fmaf(dOut[(size_t)s * outDim + o], In[(size_t)s * inDim + i], acc);
        dW[(size_t)o * inDim + i] = acc;
    } else {
        for (int s = 0; s < n; ++s) acc += dOut[(size_t)s * outDim + o];
        db[o] = acc;
    }
}

extern "C" __global__ void mlp_grad_in(const float* __restrict__ dOut, int outDim, const float* __restrict__ W, int inDim,
                                       const float* __restrict__ act, float* __restrict__ dIn, int n) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * inDim) return;
    const int s = idx / inDim, i = idx - s * inDim;
    float acc = 0.0f;
    for (int o = 0; o < outDim; ++o) acc = fmaf(dOut[(size_t)s * outDim + o], W[(size_t)o * inDim + i], acc);
    dIn[idx] = act[idx] > 0.0f ? acc : 0.0f;
}

extern "C" __global__ void mlp_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                                    float* __restrict__ V, int count, float lr, float b1, float b2, float eps,
                                    float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float g = G[i];
    const float m = b1 * M[i] + (1.0f - b1) * g;
    const float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

extern "C" __global__ void mlp_gather(const float* __restrict__ X, const float* __restrict__ y, const int* __restrict__ idx,
                                      int n, int F, float* __restrict__ Xb, float* __restrict__ yb) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n * F) return;
    const int s = k / F, f = k - s * F;
    const int r = idx[s];
    Xb[k] = X[(size_t)r * F + f];
    if (f == 0) yb[s] = y[r];
}
)";

#define TRY(call, what)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Module {
    hipModule_t mod = nullptr;
    hipFunction_t layer = nullptr, lossGrad = nullptr, gradW = nullptr, gradIn = nullptr, adam = nullptr, gather = nullptr;

    void load() {
        int dev = 0;
        TRY(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop;
        TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        std::string arch = prop.gcnArchName;
        arch = arch.substr(0, arch.find(':'));
        hiprtcProgram prog;
        if (hiprtcCreateProgram(&prog, kSource, "symcost.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
            throw std::runtime_error("hiprtcCreateProgram failed");
        const std::string a = "--offload-arch=" + arch;
        const char* opts[] = {a.c_str(), "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
        if (hiprtcCompileProgram(prog, 6, opts) != HIPRTC_SUCCESS) {
            size_t n = 0;
            hiprtcGetProgramLogSize(prog, &n);
            std::string log(n, '\0');
            if (n) hiprtcGetProgramLog(prog, &log[0]);
            hiprtcDestroyProgram(&prog);
            throw std::runtime_error("cost model kernels failed to compile: " + log);
        }
        size_t sz = 0;
        hiprtcGetCodeSize(prog, &sz);
        std::vector<char> code(sz);
        hiprtcGetCode(prog, code.data());
        hiprtcDestroyProgram(&prog);
        TRY(hipModuleLoadData(&mod, code.data()), "hipModuleLoadData");
        TRY(hipModuleGetFunction(&layer, mod, "mlp_layer"), "mlp_layer");
        TRY(hipModuleGetFunction(&lossGrad, mod, "mlp_loss_grad"), "mlp_loss_grad");
        TRY(hipModuleGetFunction(&gradW, mod, "mlp_grad_w"), "mlp_grad_w");
        TRY(hipModuleGetFunction(&gradIn, mod, "mlp_grad_in"), "mlp_grad_in");
        TRY(hipModuleGetFunction(&adam, mod, "mlp_adam"), "mlp_adam");
        TRY(hipModuleGetFunction(&gather, mod, "mlp_gather"), "mlp_gather");
    }
    ~Module() {
        if (mod) (void)hipModuleUnload(mod);
    }
};

void launch(hipFunction_t f, size_t threads, std::vector<void*> args) {
    const unsigned block = 256;
    const unsigned grid = (unsigned)((threads + block - 1) / block);
    if (!grid) return;
    TRY(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, nullptr, args.data(), nullptr), "hipModuleLaunchKernel");
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void resize(size_t m) {
        if (m <= n) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        TRY(hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(m, 1)), "hipMalloc");
        n = m;
    }
    void upload(const T* h, size_t m) {
        resize(m);
        if (m) TRY(hipMemcpy(p, h, sizeof(T) * m, hipMemcpyHostToDevice), "hipMemcpy");
    }
    void download(T* h, size_t m) const {
        if (m) TRY(hipMemcpy(h, p, sizeof(T) * m, hipMemcpyDeviceToHost), "hipMemcpy");
    }
    void zero(size_t m) {
        resize(m);
        if (m) TRY(hipMemset(p, 0, sizeof(T) * m), "hipMemset");
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

struct symcost_model {
    int F = 0, Hd = 0;
    Module mod;
    size_t numParams = 0;
    DevBuf<float> P, G, M, V;
    std::vector<float> mean, dev;
    float yMean = 0.0f, yDev = 1.0f;
    bool scaled = false;
    int adamStep = 0;
    // batch buffers
    DevBuf<float> X, Y, H1, H2, Out, dOut, dH2, dH1, sq, allX, allY, bX, bY;
    DevBuf<int> idx;

    size_t oW1() const { return 0; }
    size_t ob1() const { return (size_t)Hd * F; }
    size_t oW2() const { return ob1() + Hd; }
    size_t ob2() const { return oW2() + (size_t)Hd * Hd; }
    size_t ow3() const { return ob2() + Hd; }
    size_t ob3() const { return ow3() + Hd; }

    std::vector<float> standardize(const float* Xh, size_t n) const {
        std::vector<float> out((size_t)n * F);
        for (size_t s = 0; s < n; ++s)
            for (int f = 0; f < F; ++f) out[s * F + f] = (Xh[s * F + f] - mean[f]) / dev[f];
        return out;
    }

    // forward on device rows X (already standardised), n samples; result in Out
    void forward(const float* Xd, size_t n) {
        H1.resize(n * Hd);
        H2.resize(n * Hd);
        Out.resize(n);
        int in = F, h = Hd, one = 1, nn = (int)n, relu = 1, lin = 0;
        const float* xp = Xd;
        float *w1 = P.p + oW1(), *b1 = P.p + ob1(), *w2 = P.p + oW2(), *b2 = P.p + ob2(), *w3 = P.p + ow3(),
              *b3 = P.p + ob3();
        float *h1 = H1.p, *h2 = H2.p, *o = Out.p;
        launch(mod.layer, n * Hd, {&xp, &in, &w1, &b1, &h1, &h, &nn, &relu});
        launch(mod.layer, n * Hd, {&h1, &h, &w2, &b2, &h2, &h, &nn, &relu});
        launch(mod.layer, n, {&h2, &h, &w3, &b3, &o, &one, &nn, &lin});
    }

    // gradients of the mean squared error for n samples on device (X standardised, y
    // standardised), then one Adam step
    float step(const float* Xd, const float* yd, size_t n, float lr) {
        forward(Xd, n);
        dOut.resize(n);
        sq.resize(n);
        dH2.resize(n * Hd);
        dH1.resize(n * Hd);
        int nn = (int)n, h = Hd, in = F, one = 1;
        float *o = Out.p, *dy = dOut.p, *sqp = sq.p;
        const float* yp = yd;
        launch(mod.lossGrad, n, {&o, &yp, &dy, &sqp, &nn});
        float *gW1 = G.p + oW1(), *gb1 = G.p + ob1(), *gW2 = G.p + oW2(), *gb2 = G.p + ob2(), *gw3 = G.p + ow3(),
              *gb3 = G.p + ob3();
        float *w2 = P.p + oW2(), *w3 = P.p + ow3();
        float *h1 = H1.p, *h2 = H2.p, *dh2 = dH2.p, *dh1 = dH1.p;
        const float* xp = Xd;
        launch(mod.gradW, 1 * (Hd + 1), {&dy, &one, &h2, &h, &nn, &gw3, &gb3});
        launch(mod.gradIn, n * Hd, {&dy, &one, &w3, &h, &h2, &dh2, &nn});
        launch(mod.gradW, (size_t)Hd * (Hd + 1), {&dh2, &h, &h1, &h, &nn, &gW2, &gb2});
        launch(mod.gradIn, n * Hd, {&dh2, &h, &w2, &h, &h1, &dh1, &nn});
        launch(mod.gradW, (size_t)Hd * (F + 1), {&dh1, &h, &xp, &in, &nn, &gW1, &gb1});
        ++adamStep;
        float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        float c1 = 1.0f / (1.0f - std::pow(b1, (float)adamStep)), c2 = 1.0f / (1.0f - std::pow(b2, (float)adamStep));
        int cnt = (int)numParams;
        float *pp = P.p, *gp = G.p, *mp = M.p, *vp = V.p;
        launch(mod.adam, numParams, {&pp, &gp, &mp, &vp, &cnt, &lr, &b1, &b2, &eps, &c1, &c2});
        return 0.0f;
    }
};

namespace {

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    } catch (...) {
        g_err = "unknown C++ exception";
        return 1;
    }
}

void initModel(symcost_model* m, int F, int H) {
    if (F <= 0 || H <= 0) throw std::invalid_argument("symcost: feature and hidden sizes must be positive");
    m->F = F;
    m->Hd = H;
    m->numParams = (size_t)H * F + H + (size_t)H * H + H + H + 1;
    m->mod.load();
    m->P.zero(m->numParams);
    m->G.zero(m->numParams);
    m->M.zero(m->numParams);
    m->V.zero(m->numParams);
    m->mean.assign(F, 0.0f);
    m->dev.assign(F, 1.0f);
}

}  // namespace

extern "C" {

const char* symcost_last_error(void) { return g_err.c_str(); }

int symcost_create(symcost_t* out, int F, int H, uint64_t seed) {
    if (!out) return 1;
    *out = nullptr;
    return guarded([&] {
        std::unique_ptr<symcost_model> m(new symcost_model());
        initModel(m.get(), F, H);
        std::vector<float> p(m->numParams, 0.0f);
        std::mt19937_64 rng(seed);
        auto fill = [&](size_t off, size_t count, int fanIn) {
            std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)fanIn));
            for (size_t i = 0; i < count; ++i) p[off + i] = nd(rng);
        };
        fill(m->oW1(), (size_t)H * F, F);
        fill(m->oW2(), (size_t)H * H, H);
        fill(m->ow3(), (size_t)H, H);
        m->P.upload(p.data(), p.size());
        *out = m.release();
    });
}

int symcost_num_params(symcost_t m, size_t* n) {
    if (!m || !n) return 1;
    *n = m->numParams;
    return 0;
}

int symcost_get_params(symcost_t m, float* params, size_t n) {
    if (!m) return 1;
    return guarded([&] {
        if (n != m->numParams) throw std::invalid_argument("symcost_get_params: wrong size");
        m->P.download(params, n);
    });
}

int symcost_get_gradient(symcost_t m, float* grad, size_t n) {
    if (!m) return 1;
    return guarded([&] {
        if (n != m->numParams) throw std::invalid_argument("symcost_get_gradient: wrong size");
        TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
        m->G.download(grad, n);
    });
}

int symcost_set_params(symcost_t m, const float* params, size_t n) {
    if (!m) return 1;
    return guarded([&] {
        if (n != m->numParams) throw std::invalid_argument("symcost_set_params: wrong size");
        m->P.upload(params, n);
        m->M.zero(n);
        m->V.zero(n);
        m->adamStep = 0;
    });
}

int symcost_get_scaling(symcost_t m, float* mean, float* dev, float* y) {
    if (!m) return 1;
    std::copy(m->mean.begin(), m->mean.end(), mean);
    std::copy(m->dev.begin(), m->dev.end(), dev);
    y[0] = m->yMean;
    y[1] = m->yDev;
    return 0;
}

int symcost_set_scaling(symcost_t m, const float* mean, const float* dev, const float* y) {
    if (!m) return 1;
    return guarded([&] {
        for (int f = 0; f < m->F; ++f)
            if (!(dev[f] > 0.0f)) throw std::invalid_argument("symcost_set_scaling: deviations must be positive");
        if (!(y[1] > 0.0f)) throw std::invalid_argument("symcost_set_scaling: target deviation must be positive");
        m->mean.assign(mean, mean + m->F);
        m->dev.assign(dev, dev + m->F);
        m->yMean = y[0];
        m->yDev = y[1];
        m->scaled = true;
    });
}

int symcost_train(symcost_t m, const float* X, const float* y, size_t n, int epochs, int batch, float lr, uint64_t seed,
                  float* loss_out) {
    if (!m) return 1;
    return guarded([&] {
        if (n == 0 || !X || !y) throw std::invalid_argument("symcost_train: no samples");
        if (batch <= 0 || epochs <= 0) throw std::invalid_argument("symcost_train: batch and epochs must be positive");
        const int F = m->F;
        if (!m->scaled) {
            for (int f = 0; f < F; ++f) {
                double s = 0, s2 = 0;
                for (size_t i = 0; i < n; ++i) {
                    s += X[i * F + f];
                    s2 += (double)X[i * F + f] * X[i * F + f];
                }
                const double mu = s / n, var = std::max(0.0, s2 / n - mu * mu);
                m->mean[f] = (float)mu;
                m->dev[f] = var > 1e-12 ? (float)std::sqrt(var) : 1.0f;
            }
            double s = 0, s2 = 0;
            for (size_t i = 0; i < n; ++i) {
                s += y[i];
                s2 += (double)y[i] * y[i];
            }
            const double mu = s / n, var = std::max(0.0, s2 / n - mu * mu);
            m->yMean = (float)mu;
            m->yDev = var > 1e-12 ? (float)std::sqrt(var) : 1.0f;
            m->scaled = true;
        }
        const std::vector<float> Xs = m->standardize(X, n);
        std::vector<float> ys(n);
        for (size_t i = 0; i < n; ++i) ys[i] = (y[i] - m->yMean) / m->yDev;
        m->allX.upload(Xs.data(), Xs.size());
        m->allY.upload(ys.data(), ys.size());
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::mt19937_64 rng(seed);
        const size_t B = std::min<size_t>((size_t)batch, n);
        m->bX.resize(B * F);
        m->bY.resize(B);
        m->idx.resize(n);
        for (int ep = 0; ep < epochs; ++ep) {
            std::shuffle(order.begin(), order.end(), rng);
            m->idx.upload(order.data(), n);
            for (size_t b0 = 0; b0 < n; b0 += B) {
                const size_t nb = std::min(B, n - b0);
                const float *ax = m->allX.p, *ay = m->allY.p;
                const int* ip = m->idx.p + b0;
                int nn = (int)nb, ff = F;
                float *bx = m->bX.p, *by = m->bY.p;
                launch(m->mod.gather, nb * F, {&ax, &ay, &ip, &nn, &ff, &bx, &by});
                m->step(m->bX.p, m->bY.p, nb, lr);
            }
        }
        if (loss_out) {
            // loss over the whole training set
            m->forward(m->allX.p, n);
            std::vector<float> out(n);
            m->Out.download(out.data(), n);
            double l = 0;
            for (size_t i = 0; i < n; ++i) l += (double)(out[i] - ys[i]) * (out[i] - ys[i]);
            *loss_out = (float)(l / n);
        }
        TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

int symcost_step(symcost_t m, const float* X, const float* y, size_t n, float lr) {
    if (!m) return 1;
    return guarded([&] {
        if (!m->scaled) throw std::invalid_argument("symcost_step: set or fit the scaling first");
        const std::vector<float> Xs = m->standardize(X, n);
        std::vector<float> ys(n);
        for (size_t i = 0; i < n; ++i) ys[i] = (y[i] - m->yMean) / m->yDev;
        m->bX.upload(Xs.data(), Xs.size());
        m->bY.upload(ys.data(), n);
        m->step(m->bX.p, m->bY.p, n, lr);
        TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

int symcost_predict(symcost_t m, const float* X, size_t n, float* y_out) {
    if (!m) return 1;
    return guarded([&] {
        if (!n) return;
        const std::vector<float> Xs = m->standardize(X, n);
        m->X.upload(Xs.data(), Xs.size());
        m->forward(m->X.p, n);
        m->Out.download(y_out, n);
        for (size_t i = 0; i < n; ++i) y_out[i] = y_out[i] * m->yDev + m->yMean;
    });
}

int symcost_save(symcost_t m, const char* path) {
    if (!m || !path) return 1;
    return guarded([&] {
        std::vector<float> p(m->numParams);
        m->P.download(p.data(), p.size());
        std::ofstream f(path);
        if (!f) throw std::runtime_error(std::string("cannot write ") + path);
        f.precision(9);
        f << "symcost-mlp " << m->F << " " << m->Hd << "\n";
        for (float v : m->mean) f << v << " ";
        f << "\n";
        for (float v : m->dev) f << v << " ";
        f << "\n" << m->yMean << " " << m->yDev << "\n";
        for (float v : p) f << v << "\n";
        if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    });
}

int symcost_load(symcost_t* out, const char* path) {
    if (!out || !path) return 1;
    *out = nullptr;
    return guarded([&] {
        std::ifstream f(path);
        std::string tag;
        int F = 0, H = 0;
        if (!(f >> tag >> F >> H) || tag != "symcost-mlp") throw std::runtime_error(std::string("not a cost model: ") + path);
        std::unique_ptr<symcost_model> m(new symcost_model());
        initModel(m.get(), F, H);
        for (auto& v : m->mean) f >> v;
        for (auto& v : m->dev) f >> v;
        f >> m->yMean >> m->yDev;
        std::vector<float> p(m->numParams);
        for (auto& v : p) f >> v;
        if (!f) throw std::runtime_error(std::string("truncated cost model: ") + path);
        m->P.upload(p.data(), p.size());
        m->scaled = true;
        *out = m.release();
    });
}

void symcost_destroy(symcost_t m) { delete m; }

}  // extern "C"
