// mlp_abi.cpp -- C ABI of the proposal network's training and inference (include/sgfim.h).
//
// A trainer owns FP32 master weights, Adam moments and BF16 copies on one device,
// plus the resident training set (features [n][24], labels [n]). Every step is
// one launch of the fused forward/backward/Adam kernel over a minibatch given as
// a slice of a permutation of the training rows.
#include <hip/hip_runtime.h>

#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/sgfim.h"
#include "sgfim_impl.h"

struct sgfim_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    SgTrainState T{};
    std::vector<void*> bufs;
    float* d_X = nullptr;
    float* d_Y = nullptr;
    int32_t* d_idx = nullptr;
    int64_t n = 0;
    int32_t step = 0;
    std::string err;
};

namespace {

int mfail(sgfim_mlp* m, int code, const std::string& msg)
{
    if (m) m->err = msg;
    return code;
}

#define MCHK(m, call)                                                                                \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return mfail((m), SGFIM_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int malloc_t(sgfim_mlp* m, T** p, size_t count)
{
    void* q = nullptr;
    if (hipMalloc(&q, (count ? count : 1) * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        return mfail(m, SGFIM_E_NOMEM, "hipMalloc failed");
    }
    m->bufs.push_back(q);
    *p = (T*)q;
    return SGFIM_OK;
}

void free_data(sgfim_mlp* m)
{
    for (void* p : {(void*)m->d_X, (void*)m->d_Y, (void*)m->d_idx}) {
        if (!p) continue;
        for (size_t i = 0; i < m->bufs.size(); i++)
            if (m->bufs[i] == p) {
                (void)hipFree(p);
                m->bufs.erase(m->bufs.begin() + i);
                break;
            }
    }
    m->d_X = nullptr;
    m->d_Y = nullptr;
    m->d_idx = nullptr;
    m->n = 0;
}

}  // namespace

extern "C" {

int sgfim_mlp_create(int32_t device, const float* W1, const float* b1, const float* w2, float b2, sgfim_mlp** out)
{
    if (!out || !W1 || !b1 || !w2) return SGFIM_E_INVAL;
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) return SGFIM_E_DEVICE;
    sgfim_mlp* m = new sgfim_mlp();
    m->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return SGFIM_E_DEVICE;
    }
    SgTrainState& T = m->T;
    int rc;
    if ((rc = malloc_t(m, &T.W1, SG_FEAT * SG_HID)) || (rc = malloc_t(m, &T.b1, SG_HID)) || (rc = malloc_t(m, &T.w2, SG_HID)) ||
        (rc = malloc_t(m, &T.b2, 1)) || (rc = malloc_t(m, &T.m, SG_NPARAM)) || (rc = malloc_t(m, &T.v, SG_NPARAM)) ||
        (rc = malloc_t(m, &T.w1t, SG_HID * 32)) || (rc = malloc_t(m, &T.w2b, SG_HID)) ||
        (rc = malloc_t(m, &T.grad, SG_NPARAM)) || (rc = malloc_t(m, &T.loss, 1))) {
        sgfim_mlp_destroy(m);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(T.W1, W1, SG_FEAT * SG_HID * 4, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(T.b1, b1, SG_HID * 4, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(T.w2, w2, SG_HID * 4, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(T.b2, &b2, 4, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(T.m, 0, SG_NPARAM * 4, m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(T.v, 0, SG_NPARAM * 4, m->stream);
    if (e == hipSuccess) e = sg_mlp_pack(T, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
        sgfim_mlp_destroy(m);
        return SGFIM_E_DEVICE;
    }
    *out = m;
    return SGFIM_OK;
}

void sgfim_mlp_destroy(sgfim_mlp* m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* p : m->bufs) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char* sgfim_mlp_last_error(const sgfim_mlp* m) { return m ? m->err.c_str() : "null trainer"; }

int sgfim_mlp_set_data(sgfim_mlp* m, const float* X, const float* Y, int64_t n)
{
    if (!m || n < 0 || (n > 0 && (!X || !Y))) return SGFIM_E_INVAL;
    if (n >= ((int64_t)1 << 31)) return mfail(m, SGFIM_E_INVAL, "too many rows");
    (void)hipSetDevice(m->device);
    MCHK(m, hipStreamSynchronize(m->stream));
    free_data(m);
    int rc;
    if ((rc = malloc_t(m, &m->d_X, (size_t)n * SG_FEAT)) || (rc = malloc_t(m, &m->d_Y, (size_t)n)) ||
        (rc = malloc_t(m, &m->d_idx, (size_t)n)))
        return rc;
    m->n = n;
    if (n > 0) {
        MCHK(m, hipMemcpyAsync(m->d_X, X, (size_t)n * SG_FEAT * 4, hipMemcpyHostToDevice, m->stream));
        MCHK(m, hipMemcpyAsync(m->d_Y, Y, (size_t)n * 4, hipMemcpyHostToDevice, m->stream));
    }
    MCHK(m, hipStreamSynchronize(m->stream));
    return SGFIM_OK;
}

// One step on the rows X (host, n x 24) and labels Y: gradient (and loss) out; Adam update if lr > 0.
int sgfim_mlp_step(sgfim_mlp* m, const float* X, const float* Y, int32_t n, float lr, float* loss, float* grad)
{
    if (!m || !X || !Y || n <= 0) return SGFIM_E_INVAL;
    (void)hipSetDevice(m->device);
    float *dX = nullptr, *dY = nullptr;
    if (hipMalloc(&dX, (size_t)n * SG_FEAT * 4) != hipSuccess) return mfail(m, SGFIM_E_NOMEM, "hipMalloc");
    if (hipMalloc(&dY, (size_t)n * 4) != hipSuccess) {
        (void)hipFree(dX);
        return mfail(m, SGFIM_E_NOMEM, "hipMalloc");
    }
    m->step++;
    hipError_t e = hipMemcpyAsync(dX, X, (size_t)n * SG_FEAT * 4, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dY, Y, (size_t)n * 4, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = sg_mlp_train_step(m->T, dX, dY, nullptr, n, lr, 0.9f, 0.999f, 1e-8f, m->step, m->stream);
    if (e == hipSuccess && lr > 0.0f) e = sg_mlp_pack(m->T, m->stream);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, m->T.loss, 4, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess && grad) e = hipMemcpyAsync(grad, m->T.grad, SG_NPARAM * 4, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dX);
    (void)hipFree(dY);
    if (e != hipSuccess) return mfail(m, SGFIM_E_DEVICE, hipGetErrorString(e));
    return SGFIM_OK;
}

// Minibatch Adam over the resident training set: `epochs` passes, each a fresh
// permutation (std::mt19937_64 seeded with seed + epoch); losses[epoch] = mean
// minibatch loss of the epoch (may be null).
int sgfim_mlp_train(sgfim_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* losses)
{
    if (!m || epochs < 0 || batch <= 0) return SGFIM_E_INVAL;
    if (m->n <= 0) return mfail(m, SGFIM_E_STATE, "no training data");
    (void)hipSetDevice(m->device);
    std::vector<int32_t> perm((size_t)m->n);
    const int64_t nb = (m->n + batch - 1) / batch;
    std::vector<float> bl((size_t)nb);
    for (int32_t ep = 0; ep < epochs; ep++) {
        std::iota(perm.begin(), perm.end(), 0);
        std::mt19937_64 rng(seed + (uint64_t)ep);
        for (int64_t i = m->n - 1; i > 0; i--) {
            const int64_t j = (int64_t)(rng() % (uint64_t)(i + 1));
            std::swap(perm[(size_t)i], perm[(size_t)j]);
        }
        MCHK(m, hipMemcpyAsync(m->d_idx, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, m->stream));
        for (int64_t b = 0; b < nb; b++) {
            const int32_t cnt = (int32_t)std::min<int64_t>(batch, m->n - b * batch);
            m->step++;
            MCHK(m, sg_mlp_train_step(m->T, m->d_X, m->d_Y, m->d_idx + b * batch, cnt, lr, 0.9f, 0.999f, 1e-8f,
                                      m->step, m->stream));
            MCHK(m, sg_mlp_pack(m->T, m->stream));
            if (losses) MCHK(m, hipMemcpyAsync(&bl[(size_t)b], m->T.loss, 4, hipMemcpyDeviceToHost, m->stream));
        }
        MCHK(m, hipStreamSynchronize(m->stream));
        if (losses) {
            double s = 0.0;
            for (int64_t b = 0; b < nb; b++) s += bl[(size_t)b];
            losses[ep] = (float)(s / (double)nb);
        }
    }
    return SGFIM_OK;
}

int sgfim_mlp_get(sgfim_mlp* m, float* W1, float* b1, float* w2, float* b2)
{
    if (!m) return SGFIM_E_INVAL;
    (void)hipSetDevice(m->device);
    if (W1) MCHK(m, hipMemcpyAsync(W1, m->T.W1, SG_FEAT * SG_HID * 4, hipMemcpyDeviceToHost, m->stream));
    if (b1) MCHK(m, hipMemcpyAsync(b1, m->T.b1, SG_HID * 4, hipMemcpyDeviceToHost, m->stream));
    if (w2) MCHK(m, hipMemcpyAsync(w2, m->T.w2, SG_HID * 4, hipMemcpyDeviceToHost, m->stream));
    if (b2) MCHK(m, hipMemcpyAsync(b2, m->T.b2, 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return SGFIM_OK;
}

// Logits (or sigmoid scores when probs != 0) of n feature rows, with the current weights.
int sgfim_mlp_forward(sgfim_mlp* m, const float* X, int32_t n, int32_t probs, float* out)
{
    if (!m || !X || !out || n < 0) return SGFIM_E_INVAL;
    if (n == 0) return SGFIM_OK;
    (void)hipSetDevice(m->device);
    float *dX = nullptr, *dO = nullptr;
    if (hipMalloc(&dX, (size_t)n * SG_FEAT * 4) != hipSuccess) return mfail(m, SGFIM_E_NOMEM, "hipMalloc");
    if (hipMalloc(&dO, (size_t)n * 4) != hipSuccess) {
        (void)hipFree(dX);
        return mfail(m, SGFIM_E_NOMEM, "hipMalloc");
    }
    hipError_t e = hipMemcpyAsync(dX, X, (size_t)n * SG_FEAT * 4, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = sg_mlp_forward(m->T.w1t, m->T.b1, m->T.w2b, m->T.b2, dX, n, dO, probs ? 0 : 1, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dO, (size_t)n * 4, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dX);
    (void)hipFree(dO);
    if (e != hipSuccess) return mfail(m, SGFIM_E_DEVICE, hipGetErrorString(e));
    return SGFIM_OK;
}

}  // extern "C"
