// Transactionisation front-end on the GPU: sensor matrix -> window medians -> global
// quantile bins -> one item per (window, sensor) -> the CSR the bitmap transpose takes.
//
// Restates the reference's discretisation and windowing (code/pipeline/skab_spike.py):
//   fit_bins (:37-50)          per sensor, edges = unique(quantile(column, linspace(0,1,B+1)))
//                              over the whole column (NumPy's default linear interpolation),
//                              a constant column gets [v - 0.5, v + 0.5];
//   bin_index (:53-56)         searchsorted(edges, v, side="right") - 1, clamped to
//                              [0, len(edges) - 2];
//   window_transactions (:60-80) windows [s, s + window) for s = 0, stride, ... while
//                              s + window <= n; one item per sensor: the bin of the
//                              window's median (the mean of the two middle values for an
//                              even window).
// The item of sensor s in bin b is s * B + b + 1. The column sort runs on the device
// (hipCUB segmented radix sort); the host turns the gathered order statistics into the
// edges with NumPy's interpolation arithmetic, so the bins are the reference's bit for bit.
#include "sgfim_impl.h"

#include <hipcub/hipcub.hpp>

#define SG_MAX_WINDOW 64

// Row-major x[n][S] -> column-major xc[S][n].
__global__ void k_sensor_columns(const double* __restrict__ x, int64_t n, int32_t S, double* __restrict__ xc)
{
    const int64_t total = n * S;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = i / S;
        const int32_t s = (int32_t)(i - t * S);
        xc[(int64_t)s * n + t] = x[i];
    }
}

__global__ void k_sensor_gather(const double* __restrict__ xs, const int64_t* __restrict__ pos, int64_t m,
                                double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = xs[pos[i]];
}

// One thread per (window t, sensor s): median of the window, its bin, the item id.
// edges[s][0..ne[s]) ascending; present[s*B + b] = 1 for every bin that occurs.
__global__ __launch_bounds__(256) void k_sensor_bins(const double* __restrict__ x, int32_t S, int32_t B,
                                                     int32_t window, int32_t stride, int64_t n_win,
                                                     const double* __restrict__ edges,
                                                     const int32_t* __restrict__ ne,
                                                     int32_t* __restrict__ items, uint32_t* __restrict__ present)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_win * S) return;
    const int64_t t = i / S;
    const int32_t s = (int32_t)(i - t * S);
    double v[SG_MAX_WINDOW];
    const double* col = x + (int64_t)t * stride * S + s;
    for (int k = 0; k < window; k++) {   // insertion sort of the window
        const double y = col[(int64_t)k * S];
        int j = k;
        while (j > 0 && v[j - 1] > y) {
            v[j] = v[j - 1];
            j--;
        }
        v[j] = y;
    }
    const double med = (window & 1) ? v[window / 2] : (v[window / 2 - 1] + v[window / 2]) / 2.0;
    const double* e = edges + (int64_t)s * (B + 1);
    const int32_t n_e = ne[s];
    int32_t c = 0;   // edges <= med
    while (c < n_e && e[c] <= med) c++;
    int32_t b = c - 1;
    b = b < 0 ? 0 : (b > n_e - 2 ? n_e - 2 : b);
    const int32_t id = s * B + b + 1;
    items[i] = id;
    present[id - 1] = 1u;
}

__global__ void k_sensor_offsets(int64_t n_win, int32_t S, int64_t* __restrict__ off)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= n_win) off[t] = t * S;
}

hipError_t sg_sensor_columns(const double* x, int64_t n, int32_t S, double* xc, hipStream_t s)
{
    hipLaunchKernelGGL(k_sensor_columns, dim3(4096), dim3(256), 0, s, x, n, S, xc);
    return hipGetLastError();
}

// Sort each column (S segments of n doubles); temp/temp_bytes as hipCUB's two-phase call.
hipError_t sg_sensor_sort(void* temp, size_t* temp_bytes, const double* xc, double* xs, int64_t n, int32_t S,
                          const int64_t* seg_begin, const int64_t* seg_end, hipStream_t s)
{
    return hipcub::DeviceSegmentedRadixSort::SortKeys(temp, *temp_bytes, xc, xs, n * S, S, seg_begin, seg_end, 0,
                                                      (int)(sizeof(double) * 8), s);
}

hipError_t sg_sensor_gather(const double* xs, const int64_t* pos, int64_t m, double* out, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sensor_gather, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, xs, pos, m, out);
    return hipGetLastError();
}

hipError_t sg_sensor_bins(const double* x, int32_t S, int32_t B, int32_t window, int32_t stride, int64_t n_win,
                          const double* edges, const int32_t* ne, int32_t* items, uint32_t* present, int64_t* off,
                          hipStream_t s)
{
    const int64_t m = n_win * S;
    if (m > 0)
        hipLaunchKernelGGL(k_sensor_bins, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, x, S, B, window, stride,
                           n_win, edges, ne, items, present);
    hipLaunchKernelGGL(k_sensor_offsets, dim3((unsigned)((n_win + 1 + 255) / 256)), dim3(256), 0, s, n_win, S, off);
    return hipGetLastError();
}
