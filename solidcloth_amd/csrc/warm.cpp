// Host side of the learned warm-start: the network object (FP32 masters and
// Adam state on the device, BF16 MFMA operands), its host-array entry points,
// and the cloth model's per-vertex buffers, residual and state snapshots.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "hiputil.h"
#include "model.h"

namespace sc {

// splitmix64 (Steele, Lea & Flood 2014) as the seeded generator of the
// initial weights
static uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

std::vector<float> WarmStart::initial_params(unsigned seed)
{
    // Glorot-uniform weights, zero biases; the last layer scaled by 0.1 so
    // that an untrained network predicts corrections close to zero
    std::vector<float> p(kMlpParams, 0.f);
    uint64_t s = 0x5eed0000ull + seed;
    auto uni = [&](float a) {
        const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
        return (float)((2.0 * u - 1.0) * a);
    };
    const float a1 = std::sqrt(6.f / (kMlpIn + kMlpHidden));
    const float a2 = std::sqrt(6.f / (kMlpHidden + kMlpHidden));
    const float a3 = 0.1f * std::sqrt(6.f / (kMlpHidden + kMlpOut));
    for (int i = 0; i < kMlpIn * kMlpHidden; i++) p[kOffW1 + i] = uni(a1);
    for (int i = 0; i < kMlpHidden * kMlpHidden; i++) p[kOffW2 + i] = uni(a2);
    for (int i = 0; i < kMlpHidden * kMlpOut; i++) p[kOffW3 + i] = uni(a3);
    return p;
}

WarmStart::WarmStart(unsigned seed)
{
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    dev_alloc(d_master, kMlpParams);
    dev_alloc(d_m, kMlpParams);
    dev_alloc(d_v, kMlpParams);
    dev_alloc(d_grad, kMlpStride);
    dev_alloc(d_partial, (size_t)kMlpMaxBlocks * kMlpStride);
    dev_alloc(d_packed, (size_t)kMlpOperands * 256);
    dev_alloc(d_step, 1);
    dev_alloc(d_lr, 1);
    set_lr(mLr);
    set_params(initial_params(seed));
}

void WarmStart::set_lr(float lr)
{
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(d_lr, &lr, sizeof(float), hipMemcpyHostToDevice));
    mLr = lr;
}

WarmStart::~WarmStart()
{
    dev_free(d_master); dev_free(d_m); dev_free(d_v); dev_free(d_grad); dev_free(d_partial);
    dev_free(d_packed); dev_free(d_step); dev_free(d_lr);
    dev_free(d_X); dev_free(d_Xf); dev_free(d_T); dev_free(d_Y);
    if (stream) (void)hipStreamDestroy(stream);
}

std::vector<float> WarmStart::params() const
{
    std::vector<float> p(kMlpParams);
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(p.data(), d_master, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return p;
}

// New parameters restart the optimiser (moments and step count zero).
void WarmStart::set_params(const std::vector<float>& p)
{
    if ((int)p.size() != kMlpParams) throw std::runtime_error("the warm-start network has 5187 parameters");
    for (float x : p)
        if (!std::isfinite(x)) throw std::runtime_error("warm-start parameters must be finite");
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(d_master, p.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(d_m, 0, kMlpParams * sizeof(float)));
    HIP_OK(hipMemset(d_v, 0, kMlpParams * sizeof(float)));
    HIP_OK(hipMemset(d_grad, 0, kMlpStride * sizeof(float)));
    HIP_OK(hipMemset(d_step, 0, sizeof(int)));
    launch_mlp_pack(stream, d_master, d_packed, nullptr);
    HIP_OK(hipStreamSynchronize(stream));
}

int WarmStart::steps() const
{
    int s = 0;
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(&s, d_step, sizeof(int), hipMemcpyDeviceToHost));
    return s;
}

float WarmStart::last_loss() const
{
    float l = 0.f;
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(&l, d_grad + kMlpParams, sizeof(float), hipMemcpyDeviceToHost));
    return l;
}

void WarmStart::enqueue_gradient(hipStream_t s, const uint16_t* X, const float4* T, int n, float invCount)
{
    const int blocks = mlp_backward_blocks(n);
    launch_mlp_backward(s, weights(), X, T, n, invCount, d_partial, blocks);
    launch_mlp_reduce(s, d_partial, blocks, d_grad);
}

void WarmStart::enqueue_train(hipStream_t s, const uint16_t* X, const float4* T, int n, float invCount)
{
    enqueue_gradient(s, X, T, n, invCount);
    launch_mlp_adam(s, d_master, d_m, d_v, d_grad, d_step, d_lr);
    launch_mlp_pack(s, d_master, d_packed, d_step);
}

void WarmStart::stage(const float* X, const float* T, const float* w, int n, float& invCount)
{
    if (n <= 0) throw std::runtime_error("no samples");
    if (n > cap) {
        dev_alloc(d_X, (size_t)n * 16);
        dev_alloc(d_Xf, (size_t)n * kMlpIn);
        dev_alloc(d_T, (size_t)n);
        dev_alloc(d_Y, (size_t)n * kMlpOut);
        cap = n;
    }
    HIP_OK(hipMemcpy(d_Xf, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    launch_pack_features(stream, d_Xf, n, d_X);
    invCount = 0.f;
    if (T) {
        std::vector<float4> t(n);
        double count = 0.0;
        for (int i = 0; i < n; i++) {
            const float wi = w ? w[i] : 1.f;
            if (!(wi >= 0.f)) throw std::runtime_error("sample weights must be non-negative");
            t[i] = make_float4(T[3 * i], T[3 * i + 1], T[3 * i + 2], wi);
            count += wi;
        }
        if (count <= 0.0) throw std::runtime_error("all sample weights are zero");
        invCount = (float)(1.0 / count);
        HIP_OK(hipMemcpy(d_T, t.data(), (size_t)n * sizeof(float4), hipMemcpyHostToDevice));
    }
}

void WarmStart::forward_host(const float* X, int n, float* Y)
{
    float inv;
    stage(X, nullptr, nullptr, n, inv);
    launch_mlp_forward(stream, weights(), d_X, n, d_Y);
    HIP_OK(hipStreamSynchronize(stream));
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(Y, d_Y, (size_t)n * kMlpOut * sizeof(float), hipMemcpyDeviceToHost));
}

float WarmStart::gradient_host(const float* X, const float* T, const float* w, int n, float* grad)
{
    float inv;
    stage(X, T, w, n, inv);
    enqueue_gradient(stream, d_X, d_T, n, inv);
    HIP_OK(hipStreamSynchronize(stream));
    HIP_OK(hipGetLastError());
    std::vector<float> g(kMlpStride);
    HIP_OK(hipMemcpy(g.data(), d_grad, kMlpStride * sizeof(float), hipMemcpyDeviceToHost));
    if (grad) memcpy(grad, g.data(), kMlpParams * sizeof(float));
    return g[kMlpParams] * inv;
}

float WarmStart::train_host(const float* X, const float* T, const float* w, int n, int count)
{
    float inv;
    stage(X, T, w, n, inv);
    for (int k = 0; k < count; k++) enqueue_train(stream, d_X, d_T, n, inv);
    HIP_OK(hipStreamSynchronize(stream));
    HIP_OK(hipGetLastError());
    return last_loss() * inv;
}

// ------------------------------------------------------------ the model ----

WarmBufs ClothModel::warm_bufs() const
{
    const GpuSolver::Warm& w = gpu.warm;
    return WarmBufs{w.xhat, w.corr, w.X, w.T, w.triStrain, w.adjOff, w.adj, w.vtOff, w.vt};
}

void ClothModel::set_warm(WarmStart* ws, int mode, int iterations)
{
    if (mode < 0 || mode > 3)
        throw std::runtime_error("warm-start mode must be 0 (off), 1 (train), 2 (solve) or 3 (track)");
    if (iterations < 1) throw std::runtime_error("warm-started iterations must be >= 1");
    if (mode != 0 && !ws) throw std::runtime_error("a warm-start mode needs a network");
    // the last correction is cloth state: kept across modes, cleared when
    // tracking starts afresh
    if (mWarmMode == 0) mWarmResetCorr = true;
    mWarm = mode ? ws : nullptr;
    mWarmMode = mode;
    mWarmIters = iterations;
    mWarmResidual = mWarmPlainResidual = mWarmWarmResidual = -1.0;
    mWarmFallbacks = mWarmSteps = 0;
}

void ClothModel::ensure_warm()
{
    GpuSolver& g = gpu;
    GpuSolver::Warm& w = g.warm;
    const int nv = g.nv, nt = g.nt;
    if (!w.built) {
        const bool sized = w.nv == nv;
        if (!sized) {
            dev_alloc(w.xhat, nv);
            dev_alloc(w.corr, nv);
            dev_alloc(w.T, nv);
            dev_alloc(w.X, (size_t)nv * 16);
            dev_alloc(w.snapPos, nv);
            dev_alloc(w.snapPrev, nv);
            dev_alloc(w.snapVel, nv);
            dev_alloc(w.snapCorr, nv);
            w.nv = nv;
            mWarmResetCorr = true;
        }
        dev_alloc(w.triStrain, nt);
        // one-ring neighbours over the cloth's edges, ascending
        std::vector<std::vector<int>> nb(nv);
        double lsum = 0.0;
        for (const MeshEdge& e : edges) {
            nb[e.i1].push_back((int)e.i2);
            nb[e.i2].push_back((int)e.i1);
            const double dx = (double)uv[e.i1].x - uv[e.i2].x, dy = (double)uv[e.i1].y - uv[e.i2].y;
            lsum += std::sqrt(dx * dx + dy * dy);
        }
        std::vector<int> off(nv + 1, 0), adj;
        for (int v = 0; v < nv; v++) {
            std::sort(nb[v].begin(), nb[v].end());
            nb[v].erase(std::unique(nb[v].begin(), nb[v].end()), nb[v].end());
            off[v + 1] = off[v] + (int)nb[v].size();
            adj.insert(adj.end(), nb[v].begin(), nb[v].end());
        }
        // incident triangles by stream position (colour order), ascending
        std::vector<std::vector<int>> vt(nv);
        for (int k = 0; k < nt; k++) {
            const Triangle& t = tris[triCol.order[k]];
            vt[t.i1].push_back(k);
            vt[t.i2].push_back(k);
            vt[t.i3].push_back(k);
        }
        std::vector<int> voff(nv + 1, 0), vlist;
        for (int v = 0; v < nv; v++) {
            voff[v + 1] = voff[v] + (int)vt[v].size();
            vlist.insert(vlist.end(), vt[v].begin(), vt[v].end());
        }
        auto upload = [](int*& d, const std::vector<int>& h) {
            dev_alloc(d, std::max<size_t>(h.size(), 1));
            if (!h.empty()) HIP_OK(hipMemcpy(d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
        };
        upload(w.adjOff, off);
        upload(w.adj, adj);
        upload(w.vtOff, voff);
        upload(w.vt, vlist);
        w.ell = edges.empty() ? 1.f : (float)(lsum / (double)edges.size());
        if (!(w.ell > 0.f)) w.ell = 1.f;
        w.resBlocks = residual_blocks(nt + g.nb);
        dev_alloc(w.resPartial, 3 * w.resBlocks);
        if (!w.resOut) dev_alloc(w.resOut, 3);
        w.built = true;
        g.release_graphs();
    }
    if (mWarmResetCorr) {
        HIP_OK(hipMemset(w.corr, 0, (size_t)nv * sizeof(float4)));
        HIP_OK(hipMemset(w.T, 0, (size_t)nv * sizeof(float4)));
        HIP_OK(hipMemset(w.X, 0, (size_t)nv * 16 * sizeof(uint16_t)));
        mWarmResetCorr = false;
    }
    int free = 0;
    for (float m : invMass) free += m != 0.f;
    w.nFree = free;
}

// The solve's residual: RMS over the stiff stretch constraints (warp and
// weft of every triangle) and the bend errors. The shear constraint is
// compliant (ClothPBD.cpp:96-101): its error after a substep is set by the
// compliance rather than by the iteration count, and it is reported apart.
void ClothModel::residual_terms(double out[3])
{
    sync_to_device();
    GpuSolver& g = gpu;
    out[0] = out[1] = out[2] = 0.0;
    if (g.nv == 0) return;
    ensure_warm();
    TriStreams tr{g.d_tidx, g.d_tA, g.d_tB, g.d_tC};
    launch_residual(g.stream, g.d_pos, tr, g.nt, g.d_bidx, g.nb, (double)(0.0005f * mBendStiff), g.warm.resPartial,
                    g.warm.resBlocks, g.warm.resOut);
    double sum[3];
    HIP_OK(hipStreamSynchronize(g.stream));
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(sum, g.warm.resOut, sizeof(sum), hipMemcpyDeviceToHost));
    out[0] = g.nt ? std::sqrt(sum[0] / (2.0 * g.nt)) : 0.0;
    out[1] = g.nt ? std::sqrt(sum[1] / g.nt) : 0.0;
    out[2] = g.nb ? std::sqrt(sum[2] / g.nb) : 0.0;
    const int terms = 2 * g.nt + g.nb;
    residualSum = terms ? std::sqrt((sum[0] + sum[2]) / terms) : 0.0;
}

double ClothModel::residual()
{
    double t[3];
    residual_terms(t);
    return residualSum;
}

void ClothModel::save_state()
{
    sync_to_device();
    ensure_warm();
    GpuSolver& g = gpu;
    launch_copy(g.stream, g.d_pos, g.warm.snapPos, g.nv);
    launch_copy(g.stream, g.d_prev, g.warm.snapPrev, g.nv);
    launch_copy(g.stream, g.d_vel, g.warm.snapVel, g.nv);
    launch_copy(g.stream, g.warm.corr, g.warm.snapCorr, g.nv);
    HIP_OK(hipStreamSynchronize(g.stream));
    g.warm.saved = true;
}

// The plain judge's result (positions, previous positions, velocities, the
// warm start's correction), kept while the warm-started step runs.
void ClothModel::keep_plain_result()
{
    GpuSolver& g = gpu;
    GpuSolver::Warm& w = g.warm;
    if (w.altNv != g.nv) {
        dev_alloc(w.altPos, std::max(g.nv, 1)); dev_alloc(w.altPrev, std::max(g.nv, 1));
        dev_alloc(w.altVel, std::max(g.nv, 1)); dev_alloc(w.altCorr, std::max(g.nv, 1));
        w.altNv = g.nv;
    }
    launch_copy(g.stream, g.d_pos, w.altPos, g.nv);
    launch_copy(g.stream, g.d_prev, w.altPrev, g.nv);
    launch_copy(g.stream, g.d_vel, w.altVel, g.nv);
    launch_copy(g.stream, w.corr, w.altCorr, g.nv);
}

void ClothModel::restore_plain_result()
{
    GpuSolver& g = gpu;
    GpuSolver::Warm& w = g.warm;
    launch_copy(g.stream, w.altPos, g.d_pos, g.nv);
    launch_copy(g.stream, w.altPrev, g.d_prev, g.nv);
    launch_copy(g.stream, w.altVel, g.d_vel, g.nv);
    launch_copy(g.stream, w.altCorr, w.corr, g.nv);
    HIP_OK(hipStreamSynchronize(g.stream));
}

void ClothModel::restore_state()
{
    GpuSolver& g = gpu;
    if (!g.warm.saved || g.warm.nv != (int)pos.size()) throw std::runtime_error("no saved state for this cloth");
    launch_copy(g.stream, g.warm.snapPos, g.d_pos, g.nv);
    launch_copy(g.stream, g.warm.snapPrev, g.d_prev, g.nv);
    launch_copy(g.stream, g.warm.snapVel, g.d_vel, g.nv);
    launch_copy(g.stream, g.warm.snapCorr, g.warm.corr, g.nv);
    HIP_OK(hipStreamSynchronize(g.stream));
    after_device_step();
}

int ClothModel::warm_data(int which, float* out, int count)
{
    sync_to_device();
    GpuSolver& g = gpu;
    if (g.nv == 0) return 0;
    ensure_warm();
    const int n = g.nv;
    if (which == 4) {
        if (count >= 1) out[0] = g.warm.ell;
        if (count >= 2) out[1] = (float)g.warm.nFree;
        return 2;
    }
    const int width = which == 0 ? kMlpIn : which == 3 ? 4 : 3;
    if (which < 0 || which > 3) throw std::runtime_error("warm-start data: 0 features, 1 xhat, 2 corr, 3 targets, 4 scale");
    std::vector<float> buf((size_t)n * width);
    HIP_OK(hipStreamSynchronize(g.stream));
    if (which == 0) {
        float* d = nullptr;
        dev_alloc(d, (size_t)n * kMlpIn);
        launch_unpack_features(g.stream, g.warm.X, n, d);
        HIP_OK(hipStreamSynchronize(g.stream));
        HIP_OK(hipMemcpy(buf.data(), d, buf.size() * sizeof(float), hipMemcpyDeviceToHost));
        dev_free(d);
    } else {
        std::vector<float4> b(n);
        const float4* src = which == 1 ? g.warm.xhat : which == 2 ? g.warm.corr : g.warm.T;
        HIP_OK(hipMemcpy(b.data(), src, n * sizeof(float4), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) {
            const float v4[4] = {b[i].x, b[i].y, b[i].z, b[i].w};
            for (int k = 0; k < width; k++) buf[(size_t)i * width + k] = v4[k];
        }
    }
    const int m = std::min(count, n);
    memcpy(out, buf.data(), (size_t)m * width * sizeof(float));
    return n;
}

}  // namespace sc
