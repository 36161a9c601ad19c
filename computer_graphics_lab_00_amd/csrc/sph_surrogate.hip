// sph_surrogate.hip — the learned surrogate behind the C ABI: the network (training and
// inference on hand-written BF16 MFMA kernels), a device-resident training set recorded
// from the plain SPH path, and the rollout that replaces four substeps by one network
// evaluation and falls back to plain SPH when the density check fails.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "sph_engine.h"
#include "sph_internal.h"
#include "sph_mlp.h"
#include "sph_surrogate.h"

using sphhost::fail;
using sphhost::hip_fail;

namespace {

// Parameter layout (floats): per layer W[out][in] then b[out].
struct LayerDesc {
    uint32_t in, out, w, b;
};
constexpr LayerDesc kLayers[4] = {
    {64, 128, 0, 8192}, {128, 128, 8320, 24704}, {128, 128, 24832, 41216}, {128, 32, 41344, 45440}};
constexpr uint32_t kParams = 45472;
constexpr uint32_t kSplitRows = 256;   // batch rows per split-K partial of a weight gradient

uint32_t pad32(uint32_t v) { return (v + 31u) & ~31u; }

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

uint32_t shape_hash() {
    uint32_t h = 2166136261u;
    for (const LayerDesc& l : kLayers) {
        for (uint32_t v : {l.in, l.out}) {
            h ^= v;
            h *= 16777619u;
        }
    }
    return h;
}

}  // namespace

struct SphNet {
    int device = 0;
    hipStream_t stream = nullptr;
    float *w = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr, *frozen = nullptr;
    uint16_t* wb = nullptr;
    uint32_t adamStep = 0;
    // training batch
    uint32_t cap = 0;
    uint16_t *x = nullptr, *h[3] = {nullptr, nullptr, nullptr}, *dy = nullptr, *dh[2] = {nullptr, nullptr};
    float *y = nullptr, *t = nullptr, *partial = nullptr, *rowErr = nullptr;
    uint32_t* order = nullptr;
    // inference
    uint32_t icap = 0;
    uint16_t *ix = nullptr, *ia = nullptr, *ib = nullptr;
    float* iy = nullptr;
    float4* sums = nullptr;   // neighbourhood sums, kSumRows x icap
    float4 *keepPos = nullptr, *keepVel = nullptr;   // state before the last macro-step (sorted order)
    float2* devPartial = nullptr;
    double* hostCheck = nullptr;       // pinned, mapped: the check's two means
    double* hostCheckDev = nullptr;    // its device address
    uint32_t devPartialCap = 0;
    // 0..3: begin/end of the network pass of prediction slots 0 and 1; 4: density check copied
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    uint32_t predictedSlot = 0;   // event pair of the prediction held in iy
    // One network pass in eight is timed (an event pair on the stream costs a few
    // microseconds of idle GPU); timed[slot]: the pass in that slot was.
    uint32_t predictions = 0;
    bool timed[2] = {false, false};
    // Engine epochs (unique in the process): the solver's sorted buffers and `sums` hold
    // the prepared state of epoch `preparedEpoch`, and `iy` the network's prediction for
    // the prepared state of epoch `predictedEpoch` (0: none). The launch tables of a prepared
    // state carry its substep length (dt and the speed clamp of the integrate): `preparedDt`.
    uint64_t preparedEpoch = 0, predictedEpoch = 0;
    float preparedDt = 0.0f;
    // Density statistics of the state of epoch `checkEpoch` (mean excess, mean |deviation|):
    // the previous check's, which describe the next macro-step's starting state.
    uint64_t checkEpoch = 0;
    double checkDev[2] = {0.0, 0.0};
    uint32_t checkSides = 3;   // SPH_CHECK_COMPRESSION | SPH_CHECK_RAREFACTION
    uint32_t featureKind = sphsur::kFeaturesMoments;   // what the input rows hold
    // The inference buffers (sums, ix, iy, keep*, the check's) are used on the stream of the
    // solver a call works with; a call on another solver's stream first waits for `done`,
    // recorded at the end of the previous call on `lastStream`.
    hipEvent_t done = nullptr;
    hipStream_t lastStream = nullptr;
    float smoothing = 0.0f;   // XSPH velocity smoothing at the start of a macro-step (0: off)
    uint32_t maxBackoff = 0;  // after a failed check, macro-steps run plain before the next try
    uint32_t backoff = 0, skip = 0;
    double mlpMs = 0.0;
    uint64_t mlpRows = 0;

    ~SphNet() {
        hipSetDevice(device);
        // rollouts queue work that uses these buffers on the solver's stream
        hipDeviceSynchronize();
        for (void* p : {(void*)w, (void*)m1, (void*)m2, (void*)grad, (void*)frozen, (void*)wb, (void*)x, (void*)h[0],
                        (void*)h[1], (void*)h[2], (void*)dy, (void*)dh[0], (void*)dh[1], (void*)y, (void*)t,
                        (void*)partial, (void*)rowErr, (void*)order, (void*)ix, (void*)ia, (void*)ib, (void*)iy,
                        (void*)sums, (void*)keepPos, (void*)keepVel, (void*)devPartial})
            hipFree(p);
        if (hostCheck) hipHostFree(hostCheck);
        for (hipEvent_t e : ev)
            if (e) hipEventDestroy(e);
        if (done) hipEventDestroy(done);
        if (stream) hipStreamDestroy(stream);
    }
};

struct SphDataset {
    int device = 0;
    uint32_t featureKind = sphsur::kFeaturesMoments;
    uint32_t cap = 0, size = 0;
    uint16_t* x = nullptr;
    float* t = nullptr;
    float* t0 = nullptr;   // SPH acceleration of the first substep alone (units of 10 g)
    // recording scratch, sized for the solver it records from
    uint32_t scratchN = 0;
    uint16_t* xAll = nullptr;
    float4* sums = nullptr;
    float4* acc = nullptr;
    uint32_t *slots = nullptr, *sampleIds = nullptr;

    ~SphDataset() {
        hipSetDevice(device);
        for (void* p : {(void*)x, (void*)t, (void*)t0, (void*)xAll, (void*)sums, (void*)acc, (void*)slots,
                        (void*)sampleIds})
            hipFree(p);
    }
};

namespace {

int net_ensure_batch(SphNet* n, uint32_t rowsPad) {
    if (rowsPad <= n->cap) return SPH_OK;
    for (void* p : {(void*)n->x, (void*)n->h[0], (void*)n->h[1], (void*)n->h[2], (void*)n->dy, (void*)n->dh[0],
                    (void*)n->dh[1], (void*)n->y, (void*)n->t, (void*)n->partial, (void*)n->rowErr, (void*)n->order})
        hipFree(p);
    const size_t R = rowsPad;
    const size_t splits = (R + kSplitRows - 1) / kSplitRows;
    SPH_HIP(hipMalloc(&n->x, R * 64 * 2));
    for (int i = 0; i < 3; ++i) SPH_HIP(hipMalloc(&n->h[i], R * 128 * 2));
    SPH_HIP(hipMalloc(&n->dy, R * 32 * 2));
    for (int i = 0; i < 2; ++i) SPH_HIP(hipMalloc(&n->dh[i], R * 128 * 2));
    SPH_HIP(hipMalloc(&n->y, R * 32 * 4));
    SPH_HIP(hipMalloc(&n->t, R * 3 * 4));
    SPH_HIP(hipMalloc(&n->partial, splits * 128 * 128 * 4));
    SPH_HIP(hipMalloc(&n->rowErr, R * 4));
    SPH_HIP(hipMalloc(&n->order, R * 4));
    n->cap = rowsPad;
    return SPH_OK;
}

int net_ensure_inference(SphNet* n, uint32_t rowsPad) {
    if (rowsPad <= n->icap) return SPH_OK;
    for (void* p : {(void*)n->ix, (void*)n->ia, (void*)n->ib, (void*)n->iy, (void*)n->sums, (void*)n->keepPos,
                    (void*)n->keepVel})
        hipFree(p);
    n->preparedEpoch = 0;
    n->predictedEpoch = 0;
    const size_t R = rowsPad;
    SPH_HIP(hipMalloc(&n->ix, R * 64 * 2));
    SPH_HIP(hipMemset(n->ix, 0, R * 64 * 2));   // padded rows stay zero
    SPH_HIP(hipMalloc(&n->ia, R * 128 * 2));
    SPH_HIP(hipMalloc(&n->ib, R * 128 * 2));
    SPH_HIP(hipMalloc(&n->iy, R * 32 * 4));
    SPH_HIP(hipMalloc(&n->sums, R * sphdev::kSumRows * sizeof(float4)));
    SPH_HIP(hipMalloc(&n->keepPos, R * sizeof(float4)));
    SPH_HIP(hipMalloc(&n->keepVel, R * sizeof(float4)));
    SPH_HIP(hipMemset(n->keepPos, 0, R * sizeof(float4)));
    SPH_HIP(hipMemset(n->keepVel, 0, R * sizeof(float4)));
    n->icap = rowsPad;
    return SPH_OK;
}

sphmlp::GemmArgs gemm_forward(const uint16_t* x, const SphNet* n, const LayerDesc& l, uint32_t rowsPad) {
    sphmlp::GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.A = x;
    g.B = n->wb + l.w;       // B(k, n) = W[n][k]
    g.M = rowsPad;
    g.N = l.out;
    g.K = l.in;
    g.sAm = l.in;
    g.sAk = 1;
    g.sBk = 1;
    g.sBn = l.in;
    g.splits = 1;
    g.kPerSplit = l.in;
    g.bias = n->w + l.b;
    g.ldOut = l.out;
    return g;
}

// X (rowsPad x 64, BF16) -> Y (rowsPad x 32, FP32); hidden activations into h1, h2, h3.
// One launch of the fused network kernel (weights staged in LDS, all four layers per row group).
int net_forward(SphNet* n, const uint16_t* x, uint32_t rowsPad, uint16_t* h1, uint16_t* h2, uint16_t* h3, float* y,
                hipStream_t s) {
    SPH_HIP(sphmlp::launch_mlp_fused_train(x, rowsPad, n->wb, n->w, h1, h2, h3, y, s));
    return SPH_OK;
}

// The same forward pass by four tile GEMM launches (one per layer).
int net_forward_gemm(SphNet* n, const uint16_t* x, uint32_t rowsPad, uint16_t* h1, uint16_t* h2, uint16_t* h3,
                     float* y, hipStream_t s) {
    sphmlp::GemmArgs g = gemm_forward(x, n, kLayers[0], rowsPad);
    g.relu = 1;
    g.outBf16 = h1;
    SPH_HIP(sphmlp::launch_gemm(g, s));
    g = gemm_forward(h1, n, kLayers[1], rowsPad);
    g.relu = 1;
    g.outBf16 = h2;
    SPH_HIP(sphmlp::launch_gemm(g, s));
    g = gemm_forward(h2, n, kLayers[2], rowsPad);
    g.relu = 1;
    g.outBf16 = h3;
    SPH_HIP(sphmlp::launch_gemm(g, s));
    g = gemm_forward(h3, n, kLayers[3], rowsPad);
    g.outF32 = y;
    SPH_HIP(sphmlp::launch_gemm(g, s));
    return SPH_OK;
}

// dW[out][in] = sum over rows of dZ[row][out] * A[row][in], in fixed-order partial sums.
int weight_grad(SphNet* n, const uint16_t* dz, uint32_t outs, const uint16_t* a, uint32_t ins, uint32_t rowsPad,
                float* out, hipStream_t s) {
    sphmlp::GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.A = dz;
    g.sAm = 1;
    g.sAk = outs;
    g.B = a;
    g.sBk = ins;
    g.sBn = 1;
    g.M = outs;
    g.N = ins;
    g.K = rowsPad;
    g.kPerSplit = kSplitRows < rowsPad ? kSplitRows : rowsPad;
    g.splits = (rowsPad + g.kPerSplit - 1) / g.kPerSplit;
    g.partial = n->partial;
    SPH_HIP(sphmlp::launch_gemm(g, s));
    SPH_HIP(sphmlp::launch_reduce_partials(n->partial, g.splits, outs * ins, out, s));
    return SPH_OK;
}

// dX[row][in] = (sum over out of dZ[row][out] W[out][in]) masked where the forward
// activation of this layer's input was zero.
int input_grad(SphNet* n, const uint16_t* dz, const LayerDesc& l, const uint16_t* mask, uint32_t rowsPad,
               uint16_t* dx, hipStream_t s) {
    sphmlp::GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.A = dz;
    g.sAm = l.out;
    g.sAk = 1;
    g.B = n->wb + l.w;        // B(k = out, n = in) = W[out][in]
    g.sBk = l.in;
    g.sBn = 1;
    g.M = rowsPad;
    g.N = l.in;
    g.K = l.out;
    g.splits = 1;
    g.kPerSplit = l.out;
    g.reluMask = mask;
    g.outBf16 = dx;
    g.ldOut = l.in;
    SPH_HIP(sphmlp::launch_gemm(g, s));
    return SPH_OK;
}

// Forward + backward on the batch buffers (rows valid of rowsPad); gradient into n->grad.
int net_backprop(SphNet* n, uint32_t rows, uint32_t rowsPad, hipStream_t s) {
    if (int st = net_forward(n, n->x, rowsPad, n->h[0], n->h[1], n->h[2], n->y, s)) return st;
    const float scale = 2.0f / (float)(rows * sphsur::kOutputs);
    SPH_HIP(sphmlp::launch_mse_grad(n->y, n->t, rows, rowsPad, 32, sphsur::kOutputs, scale, n->dy, n->rowErr, s));
    // layer 3 (128 -> 32)
    if (int st = weight_grad(n, n->dy, 32, n->h[2], 128, rowsPad, n->grad + kLayers[3].w, s)) return st;
    SPH_HIP(sphmlp::launch_column_sum_bf16(n->dy, rowsPad, 32, n->partial, n->grad + kLayers[3].b, s));
    if (int st = input_grad(n, n->dy, kLayers[3], n->h[2], rowsPad, n->dh[0], s)) return st;
    // layer 2
    if (int st = weight_grad(n, n->dh[0], 128, n->h[1], 128, rowsPad, n->grad + kLayers[2].w, s)) return st;
    SPH_HIP(sphmlp::launch_column_sum_bf16(n->dh[0], rowsPad, 128, n->partial, n->grad + kLayers[2].b, s));
    if (int st = input_grad(n, n->dh[0], kLayers[2], n->h[1], rowsPad, n->dh[1], s)) return st;
    // layer 1
    if (int st = weight_grad(n, n->dh[1], 128, n->h[0], 128, rowsPad, n->grad + kLayers[1].w, s)) return st;
    SPH_HIP(sphmlp::launch_column_sum_bf16(n->dh[1], rowsPad, 128, n->partial, n->grad + kLayers[1].b, s));
    if (int st = input_grad(n, n->dh[1], kLayers[1], n->h[0], rowsPad, n->dh[0], s)) return st;
    // layer 0 (64 -> 128)
    if (int st = weight_grad(n, n->dh[0], 128, n->x, 64, rowsPad, n->grad + kLayers[0].w, s)) return st;
    SPH_HIP(sphmlp::launch_column_sum_bf16(n->dh[0], rowsPad, 128, n->partial, n->grad + kLayers[0].b, s));
    return SPH_OK;
}

int net_loss(SphNet* n, uint32_t rows, float* loss) {
    std::vector<float> e(rows);
    SPH_HIP(hipMemcpyAsync(e.data(), n->rowErr, (size_t)rows * 4, hipMemcpyDeviceToHost, n->stream));
    SPH_HIP(hipStreamSynchronize(n->stream));
    double s = 0.0;
    for (float v : e) s += v;
    *loss = (float)(s / ((double)rows * sphsur::kOutputs));
    return SPH_OK;
}

int net_adam(SphNet* n, float lr) {
    n->adamStep += 1;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 - pow(b1, (double)n->adamStep));
    const float c2 = (float)(1.0 - pow(b2, (double)n->adamStep));
    SPH_HIP(sphmlp::launch_adam(n->w, n->m1, n->m2, n->grad, kParams, lr, 0.9f, 0.999f, 1e-8f, c1, c2, n->wb, n->frozen,
                                n->stream));
    n->predictedEpoch = 0;   // the weights changed
    return SPH_OK;
}

std::vector<float> frozen_mask() {
    std::vector<float> f(kParams, 0.0f);
    const LayerDesc& l = kLayers[3];
    for (uint32_t o = sphsur::kOutputs; o < l.out; ++o) {
        for (uint32_t i = 0; i < l.in; ++i) f[l.w + o * l.in + i] = 1.0f;
        f[l.b + o] = 1.0f;
    }
    return f;
}

int upload_params(SphNet* n, const float* host) {
    std::vector<float> p(host, host + kParams);
    const std::vector<float> fz = frozen_mask();
    for (uint32_t i = 0; i < kParams; ++i)
        if (fz[i] != 0.0f) p[i] = 0.0f;
    SPH_HIP(hipMemcpy(n->w, p.data(), kParams * 4, hipMemcpyHostToDevice));
    SPH_HIP(hipMemset(n->m1, 0, kParams * 4));
    SPH_HIP(hipMemset(n->m2, 0, kParams * 4));
    SPH_HIP(sphmlp::launch_to_bf16(n->w, kParams, n->wb, nullptr));
    SPH_HIP(hipDeviceSynchronize());
    n->adamStep = 0;
    n->predictedEpoch = 0;
    return SPH_OK;
}

}  // namespace

extern "C" uint32_t sph_net_param_count(void) { return kParams; }

extern "C" int sph_net_create(int32_t device, uint64_t seed, SphNet** out) {
    sphhost::clear_error();
    if (!out) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_create: null out");
    *out = nullptr;
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices <= 0) return fail(SPH_ERR_NO_DEVICE, "no HIP device is visible");
    if (device < 0 || device >= devices) return fail(SPH_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    SphNet* n = new SphNet();
    n->device = device;
    auto bail = [&](int st) {
        delete n;
        return st;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(fail(SPH_ERR_DEVICE, "hipSetDevice"));
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(SPH_ERR_DEVICE, "stream"));
    for (float** p : {&n->w, &n->m1, &n->m2, &n->grad, &n->frozen})
        if (hipMalloc(p, kParams * 4) != hipSuccess) return bail(fail(SPH_ERR_OUT_OF_MEMORY, "net params"));
    if (hipMalloc(&n->wb, kParams * 2) != hipSuccess) return bail(fail(SPH_ERR_OUT_OF_MEMORY, "net params"));
    for (hipEvent_t& e : n->ev) hipEventCreate(&e);
    if (hipEventCreateWithFlags(&n->done, hipEventDisableTiming) != hipSuccess)
        return bail(fail(SPH_ERR_DEVICE, "event"));
    const std::vector<float> fz = frozen_mask();
    if (hipMemcpy(n->frozen, fz.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SPH_ERR_DEVICE, "upload"));
    // He-uniform weights (bound sqrt(6 / fan_in)), zero biases, from a seeded splitmix64.
    std::vector<float> p(kParams, 0.0f);
    uint64_t s = seed;
    for (const LayerDesc& l : kLayers) {
        const double bound = sqrt(6.0 / (double)l.in);
        for (uint32_t i = 0; i < l.out * l.in; ++i) p[l.w + i] = (float)((2.0 * uniform01(s) - 1.0) * bound);
    }
    if (int st = upload_params(n, p.data())) return bail(st);
    *out = n;
    return SPH_OK;
}

extern "C" int sph_net_destroy(SphNet* n) {
    delete n;
    return SPH_OK;
}

extern "C" int sph_net_get_params(SphNet* n, float* out) {
    if (!n || !out) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_get_params: null argument");
    hipSetDevice(n->device);
    SPH_HIP(hipStreamSynchronize(n->stream));
    SPH_HIP(hipMemcpy(out, n->w, kParams * 4, hipMemcpyDeviceToHost));
    return SPH_OK;
}

extern "C" int sph_net_set_params(SphNet* n, const float* in) {
    if (!n || !in) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_set_params: null argument");
    hipSetDevice(n->device);
    SPH_HIP(hipStreamSynchronize(n->stream));
    return upload_params(n, in);
}

extern "C" int sph_net_save(SphNet* n, const char* path) {
    if (!n || !path) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_save: null argument");
    std::vector<float> p(kParams);
    if (int st = sph_net_get_params(n, p.data())) return st;
    for (float v : p)
        if (!std::isfinite(v)) return fail(SPH_ERR_STATE, "refusing to save non-finite weights");
    FILE* f = fopen(path, "wb");
    if (!f) return fail(SPH_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
    // version 2: the feature kind follows the header (version 1 files are moment networks)
    const uint32_t head[5] = {0x4E485053u /* "SPHN" */, 2u, shape_hash(), kParams, n->featureKind};
    bool ok = fwrite(head, 4, 5, f) == 5 && fwrite(p.data(), 4, kParams, f) == kParams;
    ok = (fclose(f) == 0) && ok;
    return ok ? SPH_OK : fail(SPH_ERR_STATE, "write failed");
}

extern "C" int sph_net_load(SphNet* n, const char* path) {
    if (!n || !path) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_load: null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(SPH_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
    uint32_t head[4] = {0, 0, 0, 0}, kind = sphsur::kFeaturesMoments;
    std::vector<float> p(kParams);
    bool ok = fread(head, 4, 4, f) == 4;
    ok = ok && head[0] == 0x4E485053u && (head[1] == 1u || head[1] == 2u) && head[2] == shape_hash() &&
         head[3] == kParams;
    if (ok && head[1] == 2u) ok = fread(&kind, 4, 1, f) == 1 && kind <= sphsur::kFeaturesKnn8;
    ok = ok && fread(p.data(), 4, kParams, f) == kParams;
    fclose(f);
    if (!ok) return fail(SPH_ERR_INVALID_ARGUMENT, "not a weight file of this network shape");
    if (int st = sph_net_set_params(n, p.data())) return st;
    n->featureKind = kind;
    n->predictedEpoch = 0;
    return SPH_OK;
}

// Forward pass on host rows (x: rows x 64 BF16 bits) -> y: rows x 3.
extern "C" int sph_net_forward(SphNet* n, const uint16_t* x, uint32_t rows, float* y) {
    sphhost::clear_error();
    if (!n || (rows && (!x || !y))) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_forward: null argument");
    if (rows == 0) return SPH_OK;
    hipSetDevice(n->device);
    const uint32_t R = pad32(rows);
    if (int st = net_ensure_batch(n, R)) return st;
    SPH_HIP(hipMemcpyAsync(n->x, x, (size_t)rows * 64 * 2, hipMemcpyHostToDevice, n->stream));
    // the inference kernel of the rollout; training uses the layer-by-layer path
    SPH_HIP(sphmlp::launch_mlp_fused(n->x, rows, n->wb, n->w, reinterpret_cast<float4*>(n->y), 0, n->stream));
    std::vector<float> yy((size_t)rows * 4);
    SPH_HIP(hipMemcpyAsync(yy.data(), n->y, (size_t)rows * 16, hipMemcpyDeviceToHost, n->stream));
    SPH_HIP(hipStreamSynchronize(n->stream));
    for (uint32_t r = 0; r < rows; ++r)
        for (uint32_t c = 0; c < 3; ++c) y[(size_t)r * 3 + c] = yy[(size_t)r * 4 + c];
    return SPH_OK;
}

static int upload_batch(SphNet* n, const uint16_t* x, const float* t, uint32_t rows, uint32_t R) {
    if (int st = net_ensure_batch(n, R)) return st;
    SPH_HIP(hipMemsetAsync(n->x, 0, (size_t)R * 64 * 2, n->stream));
    SPH_HIP(hipMemcpyAsync(n->x, x, (size_t)rows * 64 * 2, hipMemcpyHostToDevice, n->stream));
    SPH_HIP(hipMemcpyAsync(n->t, t, (size_t)rows * 3 * 4, hipMemcpyHostToDevice, n->stream));
    return SPH_OK;
}

// Gradient of the mean squared error over host rows, without an update.
extern "C" int sph_net_gradients(SphNet* n, const uint16_t* x, const float* t, uint32_t rows, float* grad,
                                 float* loss) {
    sphhost::clear_error();
    if (!n || !x || !t || !grad || rows == 0) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_gradients: bad argument");
    hipSetDevice(n->device);
    const uint32_t R = pad32(rows);
    if (int st = upload_batch(n, x, t, rows, R)) return st;
    if (int st = net_backprop(n, rows, R, n->stream)) return st;
    SPH_HIP(hipMemcpyAsync(grad, n->grad, kParams * 4, hipMemcpyDeviceToHost, n->stream));
    SPH_HIP(hipStreamSynchronize(n->stream));
    if (loss) return net_loss(n, rows, loss);
    return SPH_OK;
}

// The training forward pass on host rows, by the fused kernel (path 0) or the four tile GEMMs
// (path 1): h gets the three hidden activations (3 x rows x 128 BF16 bits), y the 32 outputs
// (rows x 32). A check that the two paths agree.
extern "C" int sph_debug_forward(SphNet* n, const uint16_t* x, uint32_t rows, uint32_t path, uint16_t* h, float* y) {
    sphhost::clear_error();
    if (!n || !x || !h || !y || rows == 0 || path > 1) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_debug_forward");
    hipSetDevice(n->device);
    const uint32_t R = pad32(rows);
    std::vector<float> t((size_t)rows * 3, 0.0f);
    if (int st = upload_batch(n, x, t.data(), rows, R)) return st;
    const int st = path == 0 ? net_forward(n, n->x, R, n->h[0], n->h[1], n->h[2], n->y, n->stream)
                             : net_forward_gemm(n, n->x, R, n->h[0], n->h[1], n->h[2], n->y, n->stream);
    if (st) return st;
    for (int l = 0; l < 3; ++l)
        SPH_HIP(hipMemcpyAsync(h + (size_t)l * rows * 128, n->h[l], (size_t)rows * 128 * 2, hipMemcpyDeviceToHost,
                               n->stream));
    SPH_HIP(hipMemcpyAsync(y, n->y, (size_t)rows * 32 * 4, hipMemcpyDeviceToHost, n->stream));
    SPH_HIP(hipStreamSynchronize(n->stream));
    return SPH_OK;
}

// One Adam step (lr, beta 0.9 / 0.999, eps 1e-8) on host rows.
extern "C" int sph_net_train_batch(SphNet* n, const uint16_t* x, const float* t, uint32_t rows, float lr, float* loss) {
    sphhost::clear_error();
    if (!n || !x || !t || rows == 0) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_train_batch: bad argument");
    hipSetDevice(n->device);
    const uint32_t R = pad32(rows);
    if (int st = upload_batch(n, x, t, rows, R)) return st;
    if (int st = net_backprop(n, rows, R, n->stream)) return st;
    if (int st = net_adam(n, lr)) return st;
    if (loss) return net_loss(n, rows, loss);
    SPH_HIP(hipStreamSynchronize(n->stream));
    return SPH_OK;
}

// ---- dataset ----

extern "C" int sph_dataset_create(int32_t device, uint32_t capacity, SphDataset** out) {
    sphhost::clear_error();
    if (!out || capacity == 0) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_dataset_create: bad argument");
    *out = nullptr;
    SphDataset* d = new SphDataset();
    d->device = device;
    if (hipSetDevice(device) != hipSuccess || hipMalloc(&d->x, (size_t)capacity * 64 * 2) != hipSuccess ||
        hipMalloc(&d->t, (size_t)capacity * 3 * 4) != hipSuccess || hipMalloc(&d->t0, (size_t)capacity * 3 * 4) != hipSuccess) {
        delete d;
        return fail(SPH_ERR_OUT_OF_MEMORY, "dataset");
    }
    d->cap = capacity;
    *out = d;
    return SPH_OK;
}

extern "C" int sph_dataset_destroy(SphDataset* d) {
    delete d;
    return SPH_OK;
}

extern "C" uint32_t sph_dataset_size(const SphDataset* d) { return d ? d->size : 0u; }

// Records one macro-step of the plain path: features of `samples` random particles at
// its start, and as target each one's mean SPH acceleration (force / density, gravity
// excluded) over the 4 substeps of length dtSub, in units of 10 g; also the first
// substep's acceleration alone, the "hold the last acceleration" prediction. Advances
// the solver by the 4 substeps.
extern "C" int sph_dataset_record(SphDataset* d, SphSolver* solver, const SphParams* p, float dtSub,
                                  uint32_t samples, uint64_t seed) {
    sphhost::clear_error();
    if (!d || !solver || !p) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_dataset_record: null argument");
    sphhost::Engine& e = solver->e;
    const uint32_t n = e.sc[0].n;
    if (samples == 0 || samples > n) return fail(SPH_ERR_INVALID_ARGUMENT, "samples must be in 1..particles");
    if (d->size + samples > d->cap) return fail(SPH_ERR_CAPACITY, "dataset full");
    if (int st = e.set_params(0, *p)) return st;
    if (int st = e.ensure_force_buffer()) return st;
    const uint32_t N = e.total_slots();
    if (d->scratchN < N) {
        for (void* q : {(void*)d->xAll, (void*)d->sums, (void*)d->acc, (void*)d->slots, (void*)d->sampleIds})
            hipFree(q);
        SPH_HIP(hipMalloc(&d->xAll, (size_t)N * 64 * 2));
        SPH_HIP(hipMalloc(&d->sums, (size_t)N * sphdev::kSumRows * sizeof(float4)));
        SPH_HIP(hipMalloc(&d->acc, (size_t)N * 16));
        SPH_HIP(hipMalloc(&d->slots, (size_t)N * 4));
        SPH_HIP(hipMalloc(&d->sampleIds, (size_t)N * 4));
        d->scratchN = N;
    }
    hipStream_t s = e.stream;
    const sphdev::NeighbourSums ns{d->sums};
    if (int st = e.prepare(dtSub, &ns)) return st;
    const sphdev::Launch L = e.plan_launch();
    if (d->featureKind == sphsur::kFeaturesKnn8)
        SPH_HIP(sphsur::launch_knn_features(L, e.sorted_pos(), e.sorted_vel(), e.bucket_table(),
                                            e.sorted_ids(), sphsur::kVelocityScale, d->xAll, s));
    else
        SPH_HIP(sphsur::launch_features(L, e.sorted_pos(), d->sums, sphsur::kVelocityScale, d->xAll, s));
    // `samples` distinct slots: a seeded partial Fisher-Yates shuffle.
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    uint64_t rs = seed;
    for (uint32_t i = 0; i < samples; ++i) {
        const uint32_t j = i + (uint32_t)(splitmix64(rs) % (uint64_t)(n - i));
        std::swap(perm[i], perm[j]);
    }
    SPH_HIP(hipMemcpyAsync(d->slots, perm.data(), (size_t)samples * 4, hipMemcpyHostToDevice, s));
    SPH_HIP(sphsur::launch_gather_rows(d->xAll, e.sorted_ids(), d->slots, samples, d->x + (size_t)d->size * 64,
                                       d->sampleIds, s));
    SPH_HIP(sphsur::launch_zero_f4(d->acc, n, s));
    for (uint32_t k = 0; k < sphsur::kMacro; ++k) {
        if (int st = e.step(dtSub, true)) return st;
        SPH_HIP(sphsur::launch_accumulate_accel(e.force_buffer(), e.state_pos(), e.state_ids(), n, d->acc, s));
        if (k == 0)
            SPH_HIP(sphsur::launch_targets(d->acc, d->sampleIds, samples, 1.0f / sphsur::kAccelScale,
                                           d->t0 + (size_t)d->size * 3, s));
    }
    SPH_HIP(sphsur::launch_targets(d->acc, d->sampleIds, samples, 1.0f / ((float)sphsur::kMacro * sphsur::kAccelScale),
                                   d->t + (size_t)d->size * 3, s));
    SPH_HIP(hipStreamSynchronize(s));
    d->size += samples;
    return SPH_OK;
}

extern "C" int sph_dataset_read(SphDataset* d, uint32_t begin, uint32_t count, uint16_t* x, float* t, float* t0) {
    if (!d || begin + count > d->size) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_dataset_read: range");
    hipSetDevice(d->device);
    if (t0) SPH_HIP(hipMemcpy(t0, d->t0 + (size_t)begin * 3, (size_t)count * 3 * 4, hipMemcpyDeviceToHost));
    if (x) SPH_HIP(hipMemcpy(x, d->x + (size_t)begin * 64, (size_t)count * 64 * 2, hipMemcpyDeviceToHost));
    if (t) SPH_HIP(hipMemcpy(t, d->t + (size_t)begin * 3, (size_t)count * 3 * 4, hipMemcpyDeviceToHost));
    return SPH_OK;
}

// `steps` Adam steps on batches of `batchRows` rows drawn (seeded, with replacement)
// from dataset rows [begin, end); losses[i] gets each step's loss when not NULL.
extern "C" int sph_net_train(SphNet* n, SphDataset* d, uint32_t begin, uint32_t end, uint32_t batchRows,
                             uint32_t steps, float lr, uint64_t seed, float* losses) {
    sphhost::clear_error();
    if (!n || !d || end > d->size || begin >= end || batchRows == 0)
        return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_train: bad argument");
    hipSetDevice(n->device);
    const uint32_t R = pad32(batchRows);
    if (int st = net_ensure_batch(n, R)) return st;
    std::vector<uint32_t> order(batchRows);
    uint64_t rs = seed;
    for (uint32_t i = 0; i < steps; ++i) {
        for (uint32_t r = 0; r < batchRows; ++r) order[r] = begin + (uint32_t)(splitmix64(rs) % (uint64_t)(end - begin));
        SPH_HIP(hipMemcpyAsync(n->order, order.data(), (size_t)batchRows * 4, hipMemcpyHostToDevice, n->stream));
        SPH_HIP(sphsur::launch_batch_gather(d->x, d->t, n->order, batchRows, R, n->x, n->t, n->stream));
        if (int st = net_backprop(n, batchRows, R, n->stream)) return st;
        if (int st = net_adam(n, lr)) return st;
        if (losses) {
            if (int st = net_loss(n, batchRows, &losses[i])) return st;
        } else {
            SPH_HIP(hipStreamSynchronize(n->stream));   // keeps `order` alive for the copy
        }
    }
    return SPH_OK;
}

// Mean squared errors on dataset rows [begin, end): out[0] of the network, out[1] of the
// no-network prediction of zero SPH acceleration (gravity only), out[2] of holding the
// first substep's SPH acceleration (which costs a plain force pass).
extern "C" int sph_net_evaluate(SphNet* n, SphDataset* d, uint32_t begin, uint32_t end, double out[3]) {
    sphhost::clear_error();
    if (!n || !d || end > d->size || begin >= end || !out) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_evaluate");
    hipSetDevice(n->device);
    const uint32_t chunk = 65536;
    if (int st = net_ensure_batch(n, chunk)) return st;
    std::vector<uint32_t> order(chunk);
    std::vector<float> err(chunk), tt((size_t)chunk * 3), t0((size_t)chunk * 3);
    double se = 0.0, base = 0.0, hold = 0.0;
    for (uint32_t b = begin; b < end; b += chunk) {
        const uint32_t rows = std::min(chunk, end - b), R = pad32(rows);
        for (uint32_t r = 0; r < rows; ++r) order[r] = b + r;
        SPH_HIP(hipMemcpyAsync(n->order, order.data(), (size_t)rows * 4, hipMemcpyHostToDevice, n->stream));
        SPH_HIP(sphsur::launch_batch_gather(d->x, d->t, n->order, rows, R, n->x, n->t, n->stream));
        if (int st = net_forward(n, n->x, R, n->h[0], n->h[1], n->h[2], n->y, n->stream)) return st;
        SPH_HIP(sphmlp::launch_mse_grad(n->y, n->t, rows, R, 32, 3, 0.0f, n->dy, n->rowErr, n->stream));
        SPH_HIP(hipMemcpyAsync(err.data(), n->rowErr, (size_t)rows * 4, hipMemcpyDeviceToHost, n->stream));
        SPH_HIP(hipMemcpyAsync(tt.data(), n->t, (size_t)rows * 12, hipMemcpyDeviceToHost, n->stream));
        SPH_HIP(hipMemcpyAsync(t0.data(), d->t0 + (size_t)b * 3, (size_t)rows * 12, hipMemcpyDeviceToHost, n->stream));
        SPH_HIP(hipStreamSynchronize(n->stream));
        for (uint32_t r = 0; r < rows; ++r) {
            se += err[r];
            for (int c = 0; c < 3; ++c) {
                const double v = tt[(size_t)r * 3 + c], w = t0[(size_t)r * 3 + c];
                base += v * v;
                hold += (w - v) * (w - v);
            }
        }
    }
    const double cnt = (double)(end - begin) * 3.0;
    out[0] = se / cnt;
    out[1] = base / cnt;
    out[2] = hold / cnt;
    return SPH_OK;
}

// ---- features and rollout ----

namespace {

// Enqueues the density statistics of the prepared state (reduced on the device into
// pinned host memory), then event 4.
// The check's buffers: mapped host memory for its two means and one workgroup sum pair per
// workgroup of the solver's launches.
int ensure_check(SphNet* n, sphhost::Engine& e) {
    // A single-scene solver's launches have one workgroup per kBlock particles; the plan may
    // not exist yet (a fresh solver), so the count comes from the particles, not the plan.
    const uint32_t blocks = (e.sc[0].n + sphdev::kBlock - 1) / sphdev::kBlock;
    if (!n->hostCheck) {
        SPH_HIP(hipHostMalloc(&n->hostCheck, 2 * sizeof(double), hipHostMallocMapped));
        SPH_HIP(hipHostGetDevicePointer((void**)&n->hostCheckDev, n->hostCheck, 0));
        SPH_HIP(hipDeviceSynchronize());
    }
    if (n->devPartialCap < blocks) {
        hipFree(n->devPartial);
        n->devPartial = nullptr;
        n->devPartialCap = 0;
        SPH_HIP(hipMalloc(&n->devPartial, (size_t)blocks * sizeof(float2)));
        n->devPartialCap = blocks;
    }
    return SPH_OK;
}

// The density check of the prepared state; `summed`: the prepare's density pass already wrote
// the workgroup sums into devPartial.
int launch_check(SphNet* n, sphhost::Engine& e, bool summed) {
    const sphdev::Launch L = e.plan_launch();
    if (int rc = ensure_check(n, e)) return rc;
    if (L.blocks > n->devPartialCap) return fail(SPH_ERR_STATE, "density check: more workgroups than partial sums");
    n->hostCheck[0] = 0.0;
    n->hostCheck[1] = 0.0;
    const uint32_t count = e.sc[0].n;
    const double inv = count ? 1.0 / (double)count : 0.0;
    if (summed)
        SPH_HIP(sphsur::launch_deviation_total(L, n->devPartial, n->hostCheckDev, inv, e.stream));
    else
        SPH_HIP(sphsur::launch_deviation(L, e.sorted_pos(), n->devPartial, n->hostCheckDev, inv, e.stream));
    SPH_HIP(hipEventRecord(n->ev[4], e.stream));
    return SPH_OK;
}

// Waits for the check enqueued last: out[0] = mean max(rho / rho0 - 1, 0), out[1] = mean
// |rho / rho0 - 1| (FP32 block sums added in FP64 in block order).
int finish_check(SphNet* n, sphhost::Engine& e, double out[2]) {
    (void)e;
    SPH_HIP(hipEventSynchronize(n->ev[4]));
    out[0] = n->hostCheck[0];
    out[1] = n->hostCheck[1];
    return SPH_OK;
}

int deviation_of(SphNet* n, sphhost::Engine& e, double out[2]) {
    if (int rc = launch_check(n, e, false)) return rc;
    return finish_check(n, e, out);
}

// The network pass on the prepared state into iy, its feature rows built in the kernel
// from the neighbourhood sums, bracketed by the events of prediction slot `slot`.
int predict(SphNet* n, sphhost::Engine& e, uint32_t slot) {
    const sphdev::Launch L = e.plan_launch();
    const SphParams& p = e.sc[0].params;
    hipStream_t s = e.stream;
    const bool timed = (n->predictions++ & 7u) == 7u;   // not the first pass of a run: it warms up
    if (timed) SPH_HIP(hipEventRecord(n->ev[2 * slot], s));
    if (n->featureKind == sphsur::kFeaturesKnn8) {
        // feature rows of the 8 nearest neighbours, then the network on the rows
        SPH_HIP(sphsur::launch_knn_features(L, e.sorted_pos(), e.sorted_vel(), e.bucket_table(),
                                            e.sorted_ids(), sphsur::kVelocityScale, n->ix, s));
        SPH_HIP(sphmlp::launch_mlp_fused(n->ix, e.sc[0].n, n->wb, n->w, reinterpret_cast<float4*>(n->iy), 0, s));
    } else {
        SPH_HIP(sphmlp::launch_mlp_fused_sums(e.sorted_pos(), n->sums, L.total, e.sc[0].n, p.restDensity,
                                              p.smoothingRadius, 1.0f / sphsur::kVelocityScale, n->wb, n->w,
                                              reinterpret_cast<float4*>(n->iy), 0, s));
    }
    if (timed) SPH_HIP(hipEventRecord(n->ev[2 * slot + 1], s));
    n->timed[slot] = timed;
    n->predictedEpoch = e.epoch();
    n->predictedSlot = slot;
    return SPH_OK;
}

// Orders this call's use of the inference buffers after the previous call's, when that one
// ran on another stream.
int net_enter(SphNet* n, hipStream_t s) {
    if (n->lastStream && n->lastStream != s) SPH_HIP(hipStreamWaitEvent(s, n->done, 0));
    return SPH_OK;
}

int net_leave(SphNet* n, hipStream_t s) {
    SPH_HIP(hipEventRecord(n->done, s));
    n->lastStream = s;
    return SPH_OK;
}

}  // namespace

// Features of every particle of the current state (index order), rows x 64 BF16 bits.
extern "C" int sph_features(SphSolver* solver, const SphParams* p, float dtSub, uint16_t* x) {
    sphhost::clear_error();
    if (!solver || !p || !x) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_features: null argument");
    sphhost::Engine& e = solver->e;
    if (int st = e.set_params(0, *p)) return st;
    const uint32_t n = e.sc[0].n, N = e.total_slots();
    uint16_t *all = nullptr, *ordered = nullptr;
    float4* sums = nullptr;
    SPH_HIP(hipMalloc(&all, (size_t)N * 64 * 2));
    SPH_HIP(hipMalloc(&ordered, (size_t)N * 64 * 2));
    SPH_HIP(hipMalloc(&sums, (size_t)N * sphdev::kSumRows * sizeof(float4)));
    const sphdev::NeighbourSums ns{sums};
    int st = e.prepare(dtSub, &ns);
    if (!st) {
        const sphdev::Launch L = e.plan_launch();
        hipError_t he = sphsur::launch_features(L, e.sorted_pos(), sums, sphsur::kVelocityScale, all, e.stream);
        if (he == hipSuccess) he = sphsur::launch_unpermute_rows(all, e.sorted_ids(), n, ordered, e.stream);
        if (he == hipSuccess) he = hipMemcpyAsync(x, ordered, (size_t)n * 64 * 2, hipMemcpyDeviceToHost, e.stream);
        if (he == hipSuccess) he = hipStreamSynchronize(e.stream);
        if (he != hipSuccess) st = hip_fail(he, "sph_features");
    }
    hipFree(all);
    hipFree(ordered);
    hipFree(sums);
    return st;
}

// The 8-nearest-neighbour rows (launch_knn_features) of every particle of the current state,
// in particle order, rows x 64 BF16 bits.
extern "C" int sph_knn_features(SphSolver* solver, const SphParams* p, float dtSub, uint16_t* x) {
    sphhost::clear_error();
    if (!solver || !p || !x) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_knn_features: null argument");
    sphhost::Engine& e = solver->e;
    if (int st = e.set_params(0, *p)) return st;
    const uint32_t n = e.sc[0].n, N = e.total_slots();
    uint16_t *all = nullptr, *ordered = nullptr;
    SPH_HIP(hipMalloc(&all, (size_t)N * 64 * 2));
    SPH_HIP(hipMalloc(&ordered, (size_t)N * 64 * 2));
    int st = e.prepare(dtSub, nullptr);
    if (!st) {
        const sphdev::Launch L = e.plan_launch();
        hipError_t he = sphsur::launch_knn_features(L, e.sorted_pos(), e.sorted_vel(), e.bucket_table(),
                                                    e.sorted_ids(), sphsur::kVelocityScale, all,
                                                    e.stream);
        if (he == hipSuccess) he = sphsur::launch_unpermute_rows(all, e.sorted_ids(), n, ordered, e.stream);
        if (he == hipSuccess) he = hipMemcpyAsync(x, ordered, (size_t)n * 64 * 2, hipMemcpyDeviceToHost, e.stream);
        if (he == hipSuccess) he = hipStreamSynchronize(e.stream);
        if (he != hipSuccess) st = hip_fail(he, "sph_knn_features");
    }
    hipFree(all);
    hipFree(ordered);
    return st;
}

// Which rows a network reads: SPH_FEATURES_MOMENTS (default) or SPH_FEATURES_KNN8. Weight
// files keep it.
extern "C" int sph_net_set_features(SphNet* n, uint32_t kind) {
    if (!n || kind > sphsur::kFeaturesKnn8) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_set_features: bad kind");
    n->featureKind = kind;
    n->predictedEpoch = 0;
    return SPH_OK;
}

extern "C" uint32_t sph_net_features(const SphNet* n) { return n ? n->featureKind : 0u; }

// The rows sph_dataset_record writes from now on (SPH_FEATURES_*).
extern "C" int sph_dataset_set_features(SphDataset* d, uint32_t kind) {
    if (!d || kind > sphsur::kFeaturesKnn8) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_dataset_set_features: bad kind");
    d->featureKind = kind;
    return SPH_OK;
}

// FluidSolver::step with the surrogate: each run of 4 substeps becomes one network
// evaluation (neighbourhood-sum features -> MLP -> 4 substeps with the predicted SPH
// acceleration), followed by the density pass of the next substep as the check; if the
// mean density excess max(rho / rho0 - 1, 0) exceeds `threshold`, the macro-step is redone
// with the plain path. Substeps left over (substeps % 4) and scenes with colliders run plain.
//
// The check's density pass is also the next macro-step's: the prepared state, and the
// network's prediction on it, are kept with the solver's epoch and re-used by the next call
// when nothing changed the solver in between. While the host waits for the check, the
// prediction for the next macro-step is already queued behind it (discarded if the check
// fails), so the GPU does not idle on the host round trip; that includes one prediction
// queued at the end of a call for the call after it.
extern "C" int sph_rollout(SphSolver* solver, SphNet* net, const SphParams* p, float deltaSeconds, uint32_t steps,
                           float threshold, SphRolloutStats* stats) {
    sphhost::clear_error();
    if (!solver || !net || !p) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_rollout: null argument");
    sphhost::Engine& e = solver->e;
    if (e.sc.size() != 1) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_rollout: single-scene solvers only");
    if (net->device != e.device) return fail(SPH_ERR_INVALID_ARGUMENT, "network and solver on different devices");
    SphRolloutStats local;
    memset(&local, 0, sizeof(local));
    SphRolloutStats& st = stats ? *stats : local;
    const uint32_t n = e.sc[0].n;
    if (n == 0 || deltaSeconds <= 0.0f || steps == 0) return SPH_OK;
    if (int rc = e.set_params(0, *p)) return rc;
    if (int rc = e.set_device()) return rc;
    const uint32_t substeps = sph_substeps(p, deltaSeconds);
    const float frameStep = SPH_MAX_FRAME_STEP < deltaSeconds ? SPH_MAX_FRAME_STEP : deltaSeconds;
    const float dtSub = frameStep / (float)substeps;
    const uint32_t macros = p->colliderCount ? 0u : substeps / sphsur::kMacro;
    const uint32_t rest = substeps - macros * sphsur::kMacro;
    const uint32_t R = pad32(e.total_slots());
    if (int rc = net_ensure_inference(net, R)) return rc;
    hipStream_t s = e.stream;
    if (int rc = net_enter(net, s)) return rc;
    if (int rc = ensure_check(net, e)) return rc;
    // every prepare here also writes the density check's workgroup sums
    const sphdev::NeighbourSums ns{net->sums, net->devPartial};
    const size_t stateBytes = (size_t)e.total_slots() * sizeof(float4);
    for (uint32_t step = 0; step < steps; ++step) {
        for (uint32_t m = 0; m < macros; ++m) {
            if (net->skip > 0) {   // backing off after a failed check: a plain macro-step
                net->skip -= 1;
                for (uint32_t k = 0; k < sphsur::kMacro; ++k)
                    if (int rc = e.step(dtSub, true)) return rc;
                st.plainSubsteps += sphsur::kMacro;
                st.backedOff += 1;
                continue;
            }
            // A prepared state of another substep length (the previous call's frame time
            // differed) has the wrong dt and speed clamp in its launch tables: prepare again.
            if (net->preparedEpoch != e.epoch() || net->preparedDt != dtSub) {
                if (int rc = e.prepare(dtSub, &ns)) return rc;
                net->preparedEpoch = e.epoch();
                net->preparedDt = dtSub;
            }
            // The starting state's density statistics, for the rarefaction side of the check:
            // the previous check's when it described this state, else reduced now from the
            // prepare's workgroup sums (after a fallback or on a fresh state).
            if ((net->checkSides & SPH_CHECK_RAREFACTION) && net->checkEpoch != e.epoch()) {
                if (int rc = launch_check(net, e, true)) return rc;
                if (int rc = finish_check(net, e, net->checkDev)) return rc;
                net->checkEpoch = e.epoch();
            }
            const double deficitBefore = net->checkDev[1] - net->checkDev[0];
            if (net->predictedEpoch != e.epoch())
                if (int rc = predict(net, e, 1u - net->predictedSlot)) return rc;
            const uint32_t used = net->predictedSlot;
            const sphdev::Launch L = e.plan_launch();
            // four substeps from the prepared (sorted) state; the state before them is kept
            SPH_HIP(sphsur::launch_integrate_macro(L, e.sorted_pos(), e.sorted_vel(), net->iy, 4, sphsur::kAccelScale,
                                                   sphsur::kMacro, net->sums, net->smoothing, e.state_pos(),
                                                   e.state_vel(), net->keepPos, net->keepVel, s));
            e.commit_sorted();
            if (int rc = e.prepare(dtSub, &ns)) return rc;
            net->preparedEpoch = e.epoch();
            net->preparedDt = dtSub;
            if (int rc = launch_check(net, e, true)) return rc;
            // the next macro-step's prediction, queued before the host waits for the check
            if (m + 1 < macros || rest == 0)
                if (int rc = predict(net, e, 1u - used)) return rc;
            double dev[2];
            if (int rc = finish_check(net, e, dev)) return rc;
            float ms = 0.0f;
            if (net->timed[used] && hipEventElapsedTime(&ms, net->ev[2 * used], net->ev[2 * used + 1]) == hipSuccess) {
                net->mlpMs += ms;
                net->mlpRows += n;
            }
            st.macroSteps += 1;
            st.lastDeviation = (float)dev[0];
            if ((float)dev[0] > st.maxDeviation) st.maxDeviation = (float)dev[0];
            // two-sided: compression (mean excess above the threshold) or rarefaction (the mean
            // deficit grew by more than the threshold over the macro-step)
            const double growth = (dev[1] - dev[0]) - deficitBefore;
            if ((float)growth > st.maxRarefactionGrowth) st.maxRarefactionGrowth = (float)growth;
            const bool compressed = (net->checkSides & SPH_CHECK_COMPRESSION) && dev[0] > (double)threshold;
            const bool rarefied = (net->checkSides & SPH_CHECK_RAREFACTION) && growth > (double)threshold;
            if (rarefied && !compressed) st.rarefactionFallbacks += 1;
            if (!compressed && !rarefied) {
                net->checkEpoch = e.epoch();   // these statistics describe the next starting state
                net->checkDev[0] = dev[0];
                net->checkDev[1] = dev[1];
            }
            if (compressed || rarefied) {
                // The check failed: back to the kept state, and the plain path redoes the
                // macro-step. The kept state is in the prepared (sorted) order, which is the
                // order of the committed ids.
                SPH_HIP(hipMemcpyAsync(e.state_pos(), net->keepPos, stateBytes, hipMemcpyDeviceToDevice, s));
                SPH_HIP(hipMemcpyAsync(e.state_vel(), net->keepVel, stateBytes, hipMemcpyDeviceToDevice, s));
                e.touch();
                for (uint32_t k = 0; k < sphsur::kMacro; ++k)
                    if (int rc = e.step(dtSub, true)) return rc;
                st.fallbacks += 1;
                st.plainSubsteps += sphsur::kMacro;
                if (net->maxBackoff > 0) {
                    net->backoff = net->backoff ? std::min(2 * net->backoff, net->maxBackoff) : 1u;
                    net->skip = net->backoff;
                }
            } else {
                net->backoff = 0;
            }
        }
        for (uint32_t k = 0; k < rest; ++k) {
            if (int rc = e.step(dtSub, true)) return rc;
            st.plainSubsteps += 1;
        }
    }
    return net_leave(net, s);
}

// After a failed check the next 1, 2, 4, ... (at most maxBackoff) macro-steps run plain
// before the network is tried again; a passed check resets the sequence. 0 disables it.
extern "C" int sph_net_set_backoff(SphNet* n, uint32_t maxBackoff) {
    if (!n) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_set_backoff: null net");
    n->maxBackoff = maxBackoff;
    n->backoff = 0;
    n->skip = 0;
    return SPH_OK;
}

// Which sides of the density check sph_rollout applies (SPH_CHECK_*; both by default).
extern "C" int sph_net_set_check(SphNet* n, uint32_t sides) {
    if (!n || sides == 0 || (sides & ~(uint32_t)(SPH_CHECK_COMPRESSION | SPH_CHECK_RAREFACTION)))
        return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_set_check: sides must be a non-empty set of SPH_CHECK_*");
    n->checkSides = sides;
    return SPH_OK;
}

// XSPH velocity smoothing toward the neighbourhood's s-weighted mean velocity at the start of
// each macro-step, 0 (the default) to 1.
extern "C" int sph_net_set_smoothing(SphNet* n, float smoothing) {
    if (!n || !(smoothing >= 0.0f && smoothing <= 1.0f))
        return fail(SPH_ERR_INVALID_ARGUMENT, "sph_net_set_smoothing: smoothing must be in [0, 1]");
    n->smoothing = smoothing;
    return SPH_OK;
}

// Time the timed network passes (one in eight) took in sph_rollout since the last call (ms)
// and the rows they processed; for the MFMA-utilisation figure.
extern "C" int sph_net_inference_time(SphNet* n, double* ms, uint64_t* rows) {
    if (!n) return fail(SPH_ERR_INVALID_ARGUMENT, "null net");
    if (ms) *ms = n->mlpMs;
    if (rows) *rows = n->mlpRows;
    n->mlpMs = 0.0;
    n->mlpRows = 0;
    return SPH_OK;
}

// Density statistics of the current state (after a density pass of length dtSub):
// out[0] = mean max(rho / rho0 - 1, 0), out[1] = mean |rho / rho0 - 1|.
extern "C" int sph_density_deviation(SphSolver* solver, SphNet* net, const SphParams* p, float dtSub, double out[2]) {
    sphhost::clear_error();
    if (!solver || !net || !p || !out) return fail(SPH_ERR_INVALID_ARGUMENT, "sph_density_deviation: null argument");
    sphhost::Engine& e = solver->e;
    if (int rc = e.set_params(0, *p)) return rc;
    if (int rc = e.set_device()) return rc;
    if (int rc = net_enter(net, e.stream)) return rc;
    if (int rc = e.prepare(dtSub, nullptr)) return rc;
    if (int rc = deviation_of(net, e, out)) return rc;
    return net_leave(net, e.stream);
}

// C[M x N] = A[M x K] * B[K x N] for host row-major BF16 operands through the network's
// GEMM kernel (split-K with `splits` > 1); a check of the MFMA tile layout.
extern "C" int sph_debug_gemm(const uint16_t* a, const uint16_t* b, uint32_t m, uint32_t n, uint32_t k,
                              uint32_t splits, float* c) {
    sphhost::clear_error();
    if (!a || !b || !c || m == 0 || n == 0 || k == 0 || m % 32 || n % 32 || splits == 0 || k % (16 * splits))
        return fail(SPH_ERR_INVALID_ARGUMENT, "sph_debug_gemm: M, N multiples of 32, K of 16 * splits");
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices <= 0) return fail(SPH_ERR_NO_DEVICE, "no HIP device");
    SPH_HIP(hipSetDevice(0));
    uint16_t *da = nullptr, *db = nullptr;
    float *dc = nullptr, *dp = nullptr;
    int st = SPH_OK;
    auto run = [&]() -> int {
        SPH_HIP(hipMalloc(&da, (size_t)m * k * 2));
        SPH_HIP(hipMalloc(&db, (size_t)k * n * 2));
        SPH_HIP(hipMalloc(&dc, (size_t)m * n * 4));
        SPH_HIP(hipMalloc(&dp, (size_t)splits * m * n * 4));
        SPH_HIP(hipMemcpy(da, a, (size_t)m * k * 2, hipMemcpyHostToDevice));
        SPH_HIP(hipMemcpy(db, b, (size_t)k * n * 2, hipMemcpyHostToDevice));
        sphmlp::GemmArgs g;
        memset(&g, 0, sizeof(g));
        g.A = da;
        g.B = db;
        g.M = m;
        g.N = n;
        g.K = k;
        g.sAm = k;
        g.sAk = 1;
        g.sBk = n;
        g.sBn = 1;
        g.splits = splits;
        g.kPerSplit = k / splits;
        g.ldOut = n;
        if (splits > 1) {
            g.partial = dp;
            SPH_HIP(sphmlp::launch_gemm(g, nullptr));
            SPH_HIP(sphmlp::launch_reduce_partials(dp, splits, m * n, dc, nullptr));
        } else {
            g.outF32 = dc;
            SPH_HIP(sphmlp::launch_gemm(g, nullptr));
        }
        SPH_HIP(hipMemcpy(c, dc, (size_t)m * n * 4, hipMemcpyDeviceToHost));
        return SPH_OK;
    };
    st = run();
    hipFree(da);
    hipFree(db);
    hipFree(dc);
    hipFree(dp);
    return st;
}
