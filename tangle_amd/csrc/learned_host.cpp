// Host side of the learned warm start: parameters, the trajectory dataset
// kept in device memory, minibatch training with Adam, and proposals applied
// to a world. See learned.hip for the network and its features.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/tangle_hip.h"
#include "errors.h"
#include "learned_device.h"
#include "world_internal.h"

using tgh::MlpDevice;

#define HIP_CHECK(expr)                                                                             \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess)                                                                       \
            return tgh::fail(TGH_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));   \
    } while (0)

struct tgh_mlp {
    int device = 0;  // the device current at tgh_mlp_create
    hipStream_t stream = nullptr;
    MlpDevice m{};
    float *grad = nullptr;     // kMlpPartialStride
    float *partial = nullptr;  // kMaxSlots * kMlpPartialStride
    uint64_t step = 0;
    // dataset
    uint16_t *X = nullptr;
    float4 *T = nullptr;
    uint64_t rows = 0, capacity = 0;
    // pending recording
    uint16_t *pending_x = nullptr;
    float4 *pending_mid = nullptr;
    float *pending_scale = nullptr;
    uint32_t pending_rows = 0, pending_capacity = 0;
    // scratch for standalone forward/gradient calls
    std::vector<void *> owned;

    ~tgh_mlp() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : owned) (void)hipFree(p);
        if (X) (void)hipFree(X);
        if (T) (void)hipFree(T);
        if (pending_x) (void)hipFree(pending_x);
        if (pending_mid) (void)hipFree(pending_mid);
        if (pending_scale) (void)hipFree(pending_scale);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

constexpr uint32_t kMaxSlots = 512;

template <typename T>
hipError_t owned_alloc(tgh_mlp *m, T **p, size_t count) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(1, count) * sizeof(T));
    if (e == hipSuccess) m->owned.push_back(*p);
    return e;
}

int init_parameters(std::vector<float> &w, uint64_t seed) {
    std::mt19937_64 rng(seed);
    w.assign(tgh::kMlpParams, 0.0f);
    auto fill = [&](uint32_t off, uint32_t out, uint32_t in, float gain) {
        std::uniform_real_distribution<float> u(-1.0f, 1.0f);
        const float bound = gain * std::sqrt(6.0f / (float)in);  // He-uniform for ReLU layers
        for (uint32_t i = 0; i < out * in; ++i) w[off + i] = bound * u(rng);
    };
    fill(tgh::kOffW1, 64, 64, 1.0f);
    fill(tgh::kOffW2, 64, 64, 1.0f);
    fill(tgh::kOffW3, 16, 64, 0.05f);  // small first proposals
    for (uint32_t o = 3; o < 16; ++o)
        for (uint32_t i = 0; i < 64; ++i) w[tgh::kOffW3 + o * 64 + i] = 0.0f;  // unused outputs
    return TGH_OK;
}

int upload_parameters(tgh_mlp *m, const float *params) {
    HIP_CHECK(hipMemcpyAsync(m->m.master, params, tgh::kMlpParams * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIP_CHECK(tgh::launch_mlp_publish(m->m, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    return TGH_OK;
}

int ensure_dataset(tgh_mlp *m, uint64_t rows) {
    if (rows <= m->capacity) return TGH_OK;
    uint64_t cap = std::max<uint64_t>(rows, 2 * m->capacity);
    uint16_t *X = nullptr;
    float4 *T = nullptr;
    HIP_CHECK(hipMalloc((void **)&X, cap * tgh::kFeatureWidth * sizeof(uint16_t)));
    HIP_CHECK(hipMalloc((void **)&T, cap * sizeof(float4)));
    if (m->rows) {
        HIP_CHECK(hipMemcpyAsync(X, m->X, m->rows * tgh::kFeatureWidth * sizeof(uint16_t), hipMemcpyDeviceToDevice,
                                 m->stream));
        HIP_CHECK(hipMemcpyAsync(T, m->T, m->rows * sizeof(float4), hipMemcpyDeviceToDevice, m->stream));
        HIP_CHECK(hipStreamSynchronize(m->stream));
    }
    if (m->X) (void)hipFree(m->X);
    if (m->T) (void)hipFree(m->T);
    m->X = X;
    m->T = T;
    m->capacity = cap;
    return TGH_OK;
}

int ensure_pending(tgh_mlp *m, uint32_t rows) {
    if (rows <= m->pending_capacity) return TGH_OK;
    if (m->pending_x) (void)hipFree(m->pending_x);
    if (m->pending_mid) (void)hipFree(m->pending_mid);
    if (m->pending_scale) (void)hipFree(m->pending_scale);
    m->pending_x = nullptr;
    m->pending_mid = nullptr;
    m->pending_scale = nullptr;
    m->pending_capacity = 0;
    HIP_CHECK(hipMalloc((void **)&m->pending_x, (size_t)rows * tgh::kFeatureWidth * sizeof(uint16_t)));
    HIP_CHECK(hipMalloc((void **)&m->pending_mid, (size_t)rows * sizeof(float4)));
    HIP_CHECK(hipMalloc((void **)&m->pending_scale, (size_t)rows * sizeof(float)));
    m->pending_capacity = rows;
    return TGH_OK;
}

// Per-world scratch for proposals: features, predictions, position backup.
struct ProposalScratch {
    uint16_t *x = nullptr;
    float4 *y = nullptr;
    float4 *backup = nullptr;
};

}  // namespace

extern "C" {

size_t tgh_mlp_parameter_count(void) { return tgh::kMlpParams; }

int tgh_mlp_create(uint64_t seed, tgh_mlp **out) {
    if (!out) return tgh::fail(TGH_ERR_INVALID, "output pointer is null");
    *out = nullptr;
    tgh_mlp *m = new tgh_mlp();
    auto bail = [&](int code) {
        delete m;
        return code;
    };
    hipError_t e = hipGetDevice(&m->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e)));
    float *master = nullptr, *am = nullptr, *av = nullptr;
    uint16_t *w1, *w2, *w3, *w2n, *w3n;
    e = owned_alloc(m, &master, tgh::kMlpParams);
    if (e == hipSuccess) e = owned_alloc(m, &am, tgh::kMlpParams);
    if (e == hipSuccess) e = owned_alloc(m, &av, tgh::kMlpParams);
    if (e == hipSuccess) e = owned_alloc(m, &w1, 64 * 64);
    if (e == hipSuccess) e = owned_alloc(m, &w2, 64 * 64);
    if (e == hipSuccess) e = owned_alloc(m, &w3, 16 * 64);
    if (e == hipSuccess) e = owned_alloc(m, &w2n, 64 * 64);
    if (e == hipSuccess) e = owned_alloc(m, &w3n, 64 * 16);
    if (e == hipSuccess) e = owned_alloc(m, &m->grad, tgh::kMlpPartialStride);
    if (e == hipSuccess) e = owned_alloc(m, &m->partial, (size_t)kMaxSlots * tgh::kMlpPartialStride);
    if (e != hipSuccess) return bail(tgh::fail(TGH_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e)));
    m->m.master = master;
    m->m.adam_m = am;
    m->m.adam_v = av;
    m->m.b1 = master + tgh::kOffB1;
    m->m.b2 = master + tgh::kOffB2;
    m->m.b3 = master + tgh::kOffB3;
    m->m.w1 = w1;
    m->m.w2 = w2;
    m->m.w3 = w3;
    m->m.w2n = w2n;
    m->m.w3n = w3n;
    e = hipMemsetAsync(am, 0, tgh::kMlpParams * sizeof(float), m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(av, 0, tgh::kMlpParams * sizeof(float), m->stream);
    if (e != hipSuccess) return bail(tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e)));
    std::vector<float> w;
    init_parameters(w, seed);
    int rc = upload_parameters(m, w.data());
    if (rc) return bail(rc);
    *out = m;
    return TGH_OK;
}

void tgh_mlp_destroy(tgh_mlp *mlp) { delete mlp; }

int tgh_mlp_get_parameters(const tgh_mlp *mlp, float *out, size_t count) {
    if (!mlp || !out) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (count < tgh::kMlpParams) return tgh::fail(TGH_ERR_SIZE, "parameter buffer too small");
    HIP_CHECK(hipStreamSynchronize(mlp->stream));
    HIP_CHECK(hipMemcpy(out, mlp->m.master, tgh::kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return TGH_OK;
}

int tgh_mlp_set_parameters(tgh_mlp *mlp, const float *in, size_t count) {
    if (!mlp || !in) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (count != tgh::kMlpParams) return tgh::fail(TGH_ERR_SIZE, "expected tgh_mlp_parameter_count() floats");
    mlp->step = 0;
    HIP_CHECK(hipMemsetAsync(mlp->m.adam_m, 0, tgh::kMlpParams * sizeof(float), mlp->stream));
    HIP_CHECK(hipMemsetAsync(mlp->m.adam_v, 0, tgh::kMlpParams * sizeof(float), mlp->stream));
    return upload_parameters(mlp, in);
}

int tgh_mlp_forward(tgh_mlp *mlp, const uint16_t *features, size_t rows, float *out) {
    if (!mlp || !features || !out) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (rows == 0) return TGH_OK;
    uint16_t *X = nullptr;
    float4 *Y = nullptr;
    HIP_CHECK(hipMalloc((void **)&X, rows * tgh::kFeatureWidth * sizeof(uint16_t)));
    hipError_t e = hipMalloc((void **)&Y, rows * sizeof(float4));
    if (e != hipSuccess) {
        (void)hipFree(X);
        return tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e));
    }
    std::vector<float4> host(rows);
    e = hipMemcpyAsync(X, features, rows * tgh::kFeatureWidth * sizeof(uint16_t), hipMemcpyHostToDevice, mlp->stream);
    if (e == hipSuccess) e = tgh::launch_mlp_forward(mlp->m, X, Y, (uint32_t)rows, mlp->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host.data(), Y, rows * sizeof(float4), hipMemcpyDeviceToHost, mlp->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(mlp->stream);
    (void)hipFree(X);
    (void)hipFree(Y);
    if (e != hipSuccess) return tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e));
    for (size_t r = 0; r < rows; ++r) {
        out[3 * r] = host[r].x;
        out[3 * r + 1] = host[r].y;
        out[3 * r + 2] = host[r].z;
    }
    return TGH_OK;
}

// ---- forward pass without device allocation -------------------------------
// tgh_mlp_forward_static stages the features, the network's float4 outputs
// and the packed three-float predictions in `scratch`, a static device array
// of this translation unit: every GPU the library runs on holds its own copy
// until the process ends. Host code never passes the array's name to a
// kernel or a copy: it uses the address hipGetSymbolAddress returns for the
// device, taken once per device with that device current.
namespace {
constexpr size_t kStaticRows = TGH_MLP_FORWARD_STATIC_MAX_ROWS;
constexpr size_t kStaticX = kStaticRows * tgh::kFeatureWidth * sizeof(uint16_t);  // 8 MiB
constexpr size_t kStaticY = kStaticRows * sizeof(float4);                          // 1 MiB
constexpr size_t kStaticP = kStaticRows * 3 * sizeof(float);                       // 0.75 MiB
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kStaticX + kStaticY + kStaticP];

namespace {

__global__ void k_pack_predictions(const float4 *__restrict__ Y, float *__restrict__ out, uint32_t rows) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) {
        const float4 y = Y[r];
        out[3 * (size_t)r] = y.x;
        out[3 * (size_t)r + 1] = y.y;
        out[3 * (size_t)r + 2] = y.z;
    }
}

// The device address of `scratch` on `device` (which must be current).
int scratch_address(int device, unsigned char **out) {
    if (device < 0 || device >= kMaxDevices) return tgh::fail(TGH_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return tgh::fail(TGH_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = (unsigned char *)ptr;
    }
    *out = g_scratch[device];
    return TGH_OK;
}

}  // namespace

int tgh_mlp_forward_static(tgh_mlp *mlp, const uint16_t *features, size_t rows, float *out) {
    if (!mlp || !features || !out) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (rows == 0) return TGH_OK;
    if (rows > kStaticRows)
        return tgh::fail(TGH_ERR_SIZE, "tgh_mlp_forward_static takes at most " + std::to_string(kStaticRows) +
                                           " rows (" + std::to_string(rows) + " given)");
    int current = 0;
    HIP_CHECK(hipGetDevice(&current));
    if (current != mlp->device) HIP_CHECK(hipSetDevice(mlp->device));
    unsigned char *base = nullptr;
    int rc = scratch_address(mlp->device, &base);
    hipError_t e = hipSuccess;
    if (rc == TGH_OK) {
        uint16_t *X = (uint16_t *)base;
        float4 *Y = (float4 *)(base + kStaticX);
        float *P = (float *)(base + kStaticX + kStaticY);
        e = hipMemcpyAsync(X, features, rows * tgh::kFeatureWidth * sizeof(uint16_t), hipMemcpyHostToDevice,
                           mlp->stream);
        if (e == hipSuccess) e = tgh::launch_mlp_forward(mlp->m, X, Y, (uint32_t)rows, mlp->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_pack_predictions, dim3((uint32_t)((rows + 255) / 256)), dim3(256), 0, mlp->stream, Y,
                               P, (uint32_t)rows);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, P, rows * 3 * sizeof(float), hipMemcpyDeviceToHost, mlp->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(mlp->stream);
    }
    if (current != mlp->device) (void)hipSetDevice(current);
    if (rc) return rc;
    if (e != hipSuccess) return tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e));
    return TGH_OK;
}

int tgh_mlp_gradients(tgh_mlp *mlp, const uint16_t *features, const float *targets, size_t rows, float *grad,
                      float *loss) {
    if (!mlp || !features || !targets || !grad || !loss) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (rows == 0) return tgh::fail(TGH_ERR_INVALID, "rows must be positive");
    uint16_t *X = nullptr;
    float4 *T = nullptr;
    HIP_CHECK(hipMalloc((void **)&X, rows * tgh::kFeatureWidth * sizeof(uint16_t)));
    hipError_t e = hipMalloc((void **)&T, rows * sizeof(float4));
    if (e != hipSuccess) {
        (void)hipFree(X);
        return tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e));
    }
    std::vector<float4> t4(rows);
    for (size_t r = 0; r < rows; ++r) t4[r] = make_float4(targets[3 * r], targets[3 * r + 1], targets[3 * r + 2], 0.f);
    std::vector<float> g(tgh::kMlpPartialStride);
    const uint32_t tiles = (uint32_t)((rows + 15) / 16);
    const uint32_t slots = std::min<uint32_t>(tiles, kMaxSlots);
    e = hipMemcpyAsync(X, features, rows * tgh::kFeatureWidth * sizeof(uint16_t), hipMemcpyHostToDevice, mlp->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(T, t4.data(), rows * sizeof(float4), hipMemcpyHostToDevice, mlp->stream);
    if (e == hipSuccess)
        e = tgh::launch_mlp_gradients(mlp->m, X, T, nullptr, (uint32_t)rows, slots, mlp->partial, mlp->grad, mlp->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g.data(), mlp->grad, g.size() * sizeof(float), hipMemcpyDeviceToHost, mlp->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(mlp->stream);
    (void)hipFree(X);
    (void)hipFree(T);
    if (e != hipSuccess) return tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e));
    std::memcpy(grad, g.data(), tgh::kMlpParams * sizeof(float));
    *loss = g[tgh::kOffLoss];
    return TGH_OK;
}

int tgh_world_features(tgh_world *world, uint16_t *out, size_t count) {
    if (!world || !out) return tgh::fail(TGH_ERR_INVALID, "null argument");
    const size_t S = world->d.S;
    if (count < S * tgh::kFeatureWidth) return tgh::fail(TGH_ERR_SIZE, "feature buffer holds fewer than 64 per segment");
    uint16_t *X = nullptr;
    HIP_CHECK(hipMalloc((void **)&X, S * tgh::kFeatureWidth * sizeof(uint16_t)));
    hipError_t e = tgh::launch_refresh_lists(world->d, world->stream);
    if (e == hipSuccess)
        e = tgh::launch_features(world->d, world->history_valid ? world->history : nullptr, X, nullptr,
                                 world->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, X, S * tgh::kFeatureWidth * sizeof(uint16_t), hipMemcpyDeviceToHost,
                                            world->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(world->stream);
    (void)hipFree(X);
    if (e != hipSuccess) return tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e));
    return TGH_OK;
}

int tgh_world_mark_history(tgh_world *world) {
    if (!world) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (!world->history) HIP_CHECK(world->alloc(&world->history, world->d.S));
    HIP_CHECK(tgh::launch_midpoints(world->d, nullptr, world->history, world->stream));
    HIP_CHECK(hipStreamSynchronize(world->stream));
    world->history_valid = true;
    return TGH_OK;
}

int tgh_mlp_record_begin(tgh_mlp *mlp, tgh_world *world) {
    if (!mlp || !world) return tgh::fail(TGH_ERR_INVALID, "null argument");
    int rc = ensure_pending(mlp, world->d.S);
    if (rc) return rc;
    HIP_CHECK(hipStreamSynchronize(mlp->stream));
    HIP_CHECK(tgh::launch_refresh_lists(world->d, world->stream));
    HIP_CHECK(tgh::launch_features(world->d, world->history_valid ? world->history : nullptr, mlp->pending_x,
                                   mlp->pending_scale, world->stream));
    HIP_CHECK(tgh::launch_midpoints(world->d, mlp->pending_scale, mlp->pending_mid, world->stream));
    HIP_CHECK(hipStreamSynchronize(world->stream));
    mlp->pending_rows = world->d.S;
    return TGH_OK;
}

int tgh_mlp_record_end(tgh_mlp *mlp, tgh_world *world) {
    if (!mlp || !world) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (mlp->pending_rows == 0 || mlp->pending_rows != world->d.S)
        return tgh::fail(TGH_ERR_INVALID, "record_end without a matching record_begin on this world");
    int rc = ensure_dataset(mlp, mlp->rows + mlp->pending_rows);
    if (rc) return rc;
    HIP_CHECK(hipMemcpyAsync(mlp->X + mlp->rows * tgh::kFeatureWidth, mlp->pending_x,
                             (size_t)mlp->pending_rows * tgh::kFeatureWidth * sizeof(uint16_t),
                             hipMemcpyDeviceToDevice, world->stream));
    HIP_CHECK(tgh::launch_targets(world->d, mlp->pending_mid, mlp->T + mlp->rows, world->stream));
    HIP_CHECK(hipStreamSynchronize(world->stream));
    mlp->rows += mlp->pending_rows;
    mlp->pending_rows = 0;
    return TGH_OK;
}

uint64_t tgh_mlp_dataset_rows(const tgh_mlp *mlp) { return mlp ? mlp->rows : 0; }

int tgh_mlp_dataset_clear(tgh_mlp *mlp) {
    if (!mlp) return tgh::fail(TGH_ERR_INVALID, "null argument");
    mlp->rows = 0;
    return TGH_OK;
}

int tgh_mlp_train(tgh_mlp *mlp, uint32_t epochs, uint32_t batch_rows, float learning_rate, uint64_t seed,
                  float *epoch_losses) {
    if (!mlp) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (mlp->rows == 0) return tgh::fail(TGH_ERR_INVALID, "the dataset is empty");
    if (batch_rows == 0 || epochs == 0) return tgh::fail(TGH_ERR_INVALID, "epochs and batch_rows must be positive");
    const uint64_t rows = mlp->rows;
    uint32_t *idx = nullptr;
    HIP_CHECK(hipMalloc((void **)&idx, rows * sizeof(uint32_t)));
    std::vector<uint32_t> perm(rows);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937_64 rng(seed);
    const uint64_t batches = (rows + batch_rows - 1) / batch_rows;
    float *losses = nullptr;
    hipError_t e = hipHostMalloc((void **)&losses, batches * sizeof(float), hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipFree(idx);
        return tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e));
    }
    for (uint32_t epoch = 0; epoch < epochs && e == hipSuccess; ++epoch) {
        std::shuffle(perm.begin(), perm.end(), rng);
        e = hipMemcpyAsync(idx, perm.data(), rows * sizeof(uint32_t), hipMemcpyHostToDevice, mlp->stream);
        for (uint64_t b = 0; b < batches && e == hipSuccess; ++b) {
            const uint64_t begin = b * batch_rows;
            const uint32_t n = (uint32_t)std::min<uint64_t>(batch_rows, rows - begin);
            const uint32_t slots = std::min<uint32_t>((n + 15) / 16, kMaxSlots);
            e = tgh::launch_mlp_gradients(mlp->m, mlp->X, mlp->T, idx + begin, n, slots, mlp->partial, mlp->grad,
                                          mlp->stream);
            if (e == hipSuccess) e = tgh::launch_mlp_adam(mlp->m, mlp->grad, learning_rate, ++mlp->step, mlp->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(&losses[b], mlp->grad + tgh::kOffLoss, sizeof(float), hipMemcpyDeviceToHost,
                                   mlp->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(mlp->stream);
        if (e == hipSuccess && epoch_losses) {
            double s = 0.0;
            for (uint64_t b = 0; b < batches; ++b) s += losses[b];
            epoch_losses[epoch] = (float)(s / (double)batches);
        }
    }
    (void)hipHostFree(losses);
    (void)hipFree(idx);
    if (e != hipSuccess) return tgh::fail(TGH_ERR_DEVICE, std::string("training: ") + hipGetErrorString(e));
    return TGH_OK;
}

int tgh_world_propose(tgh_world *world, tgh_mlp *mlp, float scale) {
    if (!world || !mlp) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (!(std::isfinite(scale))) return tgh::fail(TGH_ERR_INVALID, "scale must be finite");
    const size_t S = world->d.S, V = world->d.V;
    if (!world->proposal_x) {
        HIP_CHECK(world->alloc(&world->proposal_x, S * tgh::kFeatureWidth));
        HIP_CHECK(world->alloc(&world->proposal_y, S));
        HIP_CHECK(world->alloc(&world->proposal_backup, V));
        HIP_CHECK(world->alloc(&world->proposal_scale, S));
    }
    HIP_CHECK(hipStreamSynchronize(mlp->stream));
    HIP_CHECK(tgh::launch_refresh_lists(world->d, world->stream));
    HIP_CHECK(tgh::launch_features(world->d, world->history_valid ? world->history : nullptr, world->proposal_x,
                                   world->proposal_scale, world->stream));
    HIP_CHECK(tgh::launch_mlp_forward(mlp->m, world->proposal_x, world->proposal_y, (uint32_t)S, world->stream));
    HIP_CHECK(tgh::launch_apply_proposal(world->d, world->proposal_y, world->proposal_scale, world->proposal_backup,
                                         scale, world->stream));
    // the proposal moved vertices outside the displacement-tracked loop
    HIP_CHECK(tgh::launch_request_rebuild(world->d, world->stream));
    HIP_CHECK(hipStreamSynchronize(world->stream));
    world->proposal_pending = true;
    return TGH_OK;
}

int tgh_world_revert_proposal(tgh_world *world) {
    if (!world) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (!world->proposal_pending) return tgh::fail(TGH_ERR_INVALID, "no proposal to revert");
    HIP_CHECK(tgh::launch_restore(world->d, world->proposal_backup, world->stream));
    HIP_CHECK(tgh::launch_request_rebuild(world->d, world->stream));
    HIP_CHECK(hipStreamSynchronize(world->stream));
    world->proposal_pending = false;
    return TGH_OK;
}

int tgh_world_revert_proposal_specimens(tgh_world *world, const uint32_t *revert, size_t count) {
    if (!world || !revert) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (!world->proposal_pending) return tgh::fail(TGH_ERR_INVALID, "no proposal to revert");
    if (count != world->d.nspec) return tgh::fail(TGH_ERR_SIZE, "one flag per specimen");
    uint32_t *flags = nullptr;
    HIP_CHECK(hipMalloc((void **)&flags, count * sizeof(uint32_t)));
    hipError_t e = hipMemcpyAsync(flags, revert, count * sizeof(uint32_t), hipMemcpyHostToDevice, world->stream);
    if (e == hipSuccess) e = tgh::launch_restore_specimens(world->d, world->proposal_backup, flags, world->stream);
    if (e == hipSuccess) e = tgh::launch_request_rebuild_specimens(world->d, flags, world->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(world->stream);
    (void)hipFree(flags);
    if (e != hipSuccess) return tgh::fail(TGH_ERR_DEVICE, std::string("revert: ") + hipGetErrorString(e));
    // the backup still holds every specimen's pre-proposal positions, so a
    // later call may revert more specimens until the next proposal
    return TGH_OK;
}

int tgh_mlp_time_kernels(tgh_mlp *mlp, uint32_t rows, uint32_t repeats, float *forward_ms, float *gradient_ms) {
    if (!mlp || !forward_ms || !gradient_ms) return tgh::fail(TGH_ERR_INVALID, "null argument");
    if (rows == 0 || repeats == 0) return tgh::fail(TGH_ERR_INVALID, "rows and repeats must be positive");
    uint16_t *X = nullptr;
    float4 *Y = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    hipError_t e = hipMalloc((void **)&X, (size_t)rows * tgh::kFeatureWidth * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc((void **)&Y, (size_t)rows * sizeof(float4));
    // inputs: every feature 0.25 (bf16 0x3e80); the targets are the zeroed outputs
    if (e == hipSuccess) e = hipMemsetD16Async((hipDeviceptr_t)X, 0x3e80, (size_t)rows * tgh::kFeatureWidth, mlp->stream);
    if (e == hipSuccess) e = hipMemsetAsync(Y, 0, (size_t)rows * sizeof(float4), mlp->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventCreate(&e2);
    const uint32_t slots = std::min<uint32_t>((rows + 15) / 16, kMaxSlots);
    if (e == hipSuccess) e = tgh::launch_mlp_forward(mlp->m, X, Y, rows, mlp->stream);  // warm
    if (e == hipSuccess) e = hipEventRecord(e0, mlp->stream);
    for (uint32_t i = 0; i < repeats && e == hipSuccess; ++i) e = tgh::launch_mlp_forward(mlp->m, X, Y, rows, mlp->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, mlp->stream);
    for (uint32_t i = 0; i < repeats && e == hipSuccess; ++i)
        e = tgh::launch_mlp_gradients(mlp->m, X, Y, nullptr, rows, slots, mlp->partial, mlp->grad, mlp->stream);
    if (e == hipSuccess) e = hipEventRecord(e2, mlp->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e2);
    float f = 0.f, g = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&g, e1, e2);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e2) (void)hipEventDestroy(e2);
    if (X) (void)hipFree(X);
    if (Y) (void)hipFree(Y);
    if (e != hipSuccess) return tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e));
    *forward_ms = f / (float)repeats;
    *gradient_ms = g / (float)repeats;
    return TGH_OK;
}

int tgh_mfma_probe(const uint16_t *a, const uint16_t *b, float *d) {
    if (!a || !b || !d) return tgh::fail(TGH_ERR_INVALID, "null argument");
    uint16_t *A = nullptr, *B = nullptr;
    float *D = nullptr;
    HIP_CHECK(hipMalloc((void **)&A, 256 * sizeof(uint16_t)));
    HIP_CHECK(hipMalloc((void **)&B, 256 * sizeof(uint16_t)));
    HIP_CHECK(hipMalloc((void **)&D, 256 * sizeof(float)));
    hipError_t e = hipMemcpy(A, a, 512, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B, b, 512, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = tgh::launch_mfma_probe(A, B, D, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(d, D, 1024, hipMemcpyDeviceToHost);
    (void)hipFree(A);
    (void)hipFree(B);
    (void)hipFree(D);
    if (e != hipSuccess) return tgh::fail(TGH_ERR_DEVICE, hipGetErrorString(e));
    return TGH_OK;
}

}  // extern "C"
