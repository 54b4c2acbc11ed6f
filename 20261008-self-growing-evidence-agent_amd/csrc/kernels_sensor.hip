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
// The online variant (k_online_p2, k_online_bins) bins each window median against
// streaming P^2 estimates instead (code/pipeline/online_discretize.py:22-131).
// The item of sensor s in bin b is s * B + b + 1. The quantiles need two order statistics
// each (positions floor((n - 1) q) and the one after); a radix select on the device finds
// them without sorting the columns, and the host turns them into the edges with NumPy's
// interpolation arithmetic, so the bins are the reference's bit for bit.
#include "sgfim_impl.h"

#include <algorithm>

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

// Median of the window col[0], col[S], ..., col[(window - 1) S] (insertion sort; the mean
// of the two middle values for an even window, as NumPy and pandas compute it).
__device__ inline double sg_window_median(const double* col, int32_t S, int32_t window)
{
    double v[SG_MAX_WINDOW];
    for (int k = 0; k < window; k++) {
        const double y = col[(int64_t)k * S];
        int j = k;
        while (j > 0 && v[j - 1] > y) {
            v[j] = v[j - 1];
            j--;
        }
        v[j] = y;
    }
    return (window & 1) ? v[window / 2] : (v[window / 2 - 1] + v[window / 2]) / 2.0;
}

// searchsorted(e[0..n_e), v, side="right") - 1, clamped to [0, n_e - 2].
__device__ inline int32_t sg_bin_of(const double* e, int32_t n_e, double v)
{
    int32_t c = 0;
    while (c < n_e && e[c] <= v) c++;
    const int32_t b = c - 1;
    return b < 0 ? 0 : (b > n_e - 2 ? n_e - 2 : b);
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
    const double med = sg_window_median(x + (int64_t)t * stride * S + s, S, window);
    const int32_t id = s * B + sg_bin_of(edges + (int64_t)s * (B + 1), ne[s], med) + 1;
    items[i] = id;
    present[id - 1] = 1u;
}

// ---- Online discretisation: P^2 quantile estimators (code/pipeline/online_discretize.py) ----
//
// One estimator for quantile q keeps five marker heights h and positions n. The first
// five samples are kept sorted in h (stable insertion, as Python's sorted); at the fifth,
// n = (0, 1 + 2q, 1 + 4q, 3 + 2q, 4). Each later sample x finds its cell k (a new minimum
// or maximum replaces h[0] or h[4]), every position gains 1 and positions 1..k gain one
// more, then markers i = 1, 2, 3 in turn move by qp = (d + 1)(h[i+1] - h[i-1]) / span,
// d = n[i] - n[i-1], span = n[i+1] - n[i-1], when d >= 1, span > 1 and h[i] + qp lies
// strictly between h[i-1] and h[i+1] (P2Quantile.add, :37-64). The estimate is h[2], or
// the median of the fewer than five samples seen (0 before any) (value, :66-69).
// Each step is rounded as the reference rounds it: no multiply-add contraction.
struct SgP2 {
    double h[5];
    double n[5];
    int c;   // samples seen, up to 5
};

__device__ inline void sg_p2_add(SgP2& p, double q, double x)
{
#pragma clang fp contract(off)
    if (p.c < 5) {
        double y = x;
#pragma unroll
        for (int j = 0; j < 5; j++) {
            if (j < p.c) {
                if (p.h[j] > y) {
                    const double t = p.h[j];
                    p.h[j] = y;
                    y = t;
                }
            } else if (j == p.c) {
                p.h[j] = y;
            }
        }
        if (++p.c == 5) {
            p.n[0] = 0.0;
            p.n[1] = 1.0 + 2.0 * q;
            p.n[2] = 1.0 + 4.0 * q;
            p.n[3] = 3.0 + 2.0 * q;
            p.n[4] = 4.0;
        }
        return;
    }
    int k;
    if (x < p.h[0]) {
        p.h[0] = x;
        k = 0;
    } else if (x < p.h[1]) {
        k = 0;
    } else if (x < p.h[2]) {
        k = 1;
    } else if (x < p.h[3]) {
        k = 2;
    } else if (x <= p.h[4]) {
        k = 3;
    } else {
        p.h[4] = x;
        k = 3;
    }
#pragma unroll
    for (int i = 0; i < 5; i++) p.n[i] = p.n[i] + 1.0;
#pragma unroll
    for (int i = 1; i < 4; i++)
        if (i <= k) p.n[i] = p.n[i] + 1.0;
#pragma unroll
    for (int i = 1; i <= 3; i++) {
        const double d = p.n[i] - p.n[i - 1];
        const double span = p.n[i + 1] - p.n[i - 1];
        if (d >= 1.0 && span > 1.0) {
            const double qp = (d + 1.0) * (p.h[i + 1] - p.h[i - 1]) / span;
            const double m = p.h[i] + qp;
            if (p.h[i - 1] < m && m < p.h[i + 1]) p.h[i] = m;
        }
    }
}

__device__ inline double sg_p2_value(const SgP2& p)
{
#pragma clang fp contract(off)
    switch (p.c) {
    case 0: return 0.0;
    case 1: return p.h[0];
    case 2: return (p.h[0] + p.h[1]) / 2.0;
    case 3: return p.h[1];
    case 4: return (p.h[1] + p.h[2]) / 2.0;
    default: return p.h[2];
    }
}

// One lane per (sensor s, estimator e), e = 0..B-2 at quantile (e + 1) / B: the lane runs
// the sensor's stream window by window (every sample of [w stride, w stride + window) is
// added, overlapping windows add a sample again, as run_file does, :118-123) and writes
// its estimate after window w to E[w][s][1 + e]. Consecutive lanes take consecutive
// sensors, so the loads of one row coalesce.
__global__ __launch_bounds__(64) void k_online_p2(const double* __restrict__ x, int32_t S, int32_t B,
                                                  int32_t window, int32_t stride, int64_t n_win,
                                                  double* __restrict__ E)
{
#pragma clang fp contract(off)
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= S * (B - 1)) return;
    const int32_t s = i % S, e = i / S;
    const double q = (double)(e + 1) / (double)B;
    SgP2 p;
#pragma unroll
    for (int j = 0; j < 5; j++) p.h[j] = p.n[j] = 0.0;
    p.c = 0;
    for (int64_t w = 0; w < n_win; w++) {
        const double* col = x + w * stride * S + s;
        for (int32_t t = 0; t < window; t++) sg_p2_add(p, q, col[(int64_t)t * S]);
        E[(w * S + s) * (B + 1) + 1 + e] = sg_p2_value(p);
    }
}

// One thread per (window w, sensor s): the edges of OnlineBinner.edges (:86-103) from the
// B - 1 estimates in E[w][s][1..B-1] -- sorted, span = max(q_last - q_0, 1e-9), minimum
// gap 0.05 span (B = 3: both inner edges re-centred on their midpoint when closer than
// the gap; otherwise each edge pushed up to the previous one plus the gap), outer edges
// q_0 - span and q_last + span -- written back in place, then the window median's bin
// under them (bin_index, :106-108) and the item s * B + b + 1.
__global__ __launch_bounds__(256) void k_online_bins(const double* __restrict__ x, int32_t S, int32_t B,
                                                     int32_t window, int32_t stride, int64_t n_win,
                                                     double* __restrict__ E, int32_t* __restrict__ items,
                                                     uint32_t* __restrict__ present)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_win * S) return;
    const int64_t t = i / S;
    const int32_t s = (int32_t)(i - t * S);
    double* ed = E + i * (B + 1);
    for (int32_t k = 2; k < B; k++) {   // stable insertion sort of ed[1..B-1]
        const double y = ed[k];
        int32_t j = k;
        while (j > 1 && ed[j - 1] > y) {
            ed[j] = ed[j - 1];
            j--;
        }
        ed[j] = y;
    }
    double span = ed[B - 1] - ed[1];
    if (span < 1e-9) span = 1e-9;
    const double gap = 0.05 * span;
    if (B == 3) {
        if (ed[2] - ed[1] < gap) {
            const double mid = (ed[1] + ed[2]) / 2.0;
            ed[1] = mid - gap / 2.0;
            ed[2] = mid + gap / 2.0;
        }
    } else {
        for (int32_t k = 2; k < B; k++)
            if (ed[k] < ed[k - 1] + gap) ed[k] = ed[k - 1] + gap;
    }
    ed[0] = ed[1] - span;
    ed[B] = ed[B - 1] + span;
    const double med = sg_window_median(x + t * stride * S + s, S, window);
    const int32_t id = s * B + sg_bin_of(ed, B + 1, med) + 1;
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

// ------------------------------------------------------------------ order statistics
// Radix select of R ranks (ascending, the same for every column) in each of S columns of n
// doubles, 8 key bits per pass. Keys order the doubles as unsigned integers (sign bit set
// for non-negative values, all bits inverted for negative ones), the order a radix sort of
// the column gives. Before pass p every target holds the key's top 8p bits it has decided
// (its prefix) and its rank among the column's values with that prefix. Targets sorted by
// rank have ascending prefixes, so equal prefixes form runs, the column's groups; groups
// are disjoint key ranges, so a value falls in at most one of them. Pass p counts the next
// 8 bits of the values in each group; each target then takes the digit its rank falls in.
// After 8 passes the prefix is the whole key of the target's order statistic.
#define SG_SEL_LDS_GROUPS 64

__device__ __forceinline__ uint64_t sg_dkey(double v)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double sg_dval(uint64_t k)
{
    const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}

// The group of prefix hp among the column's ng ascending group prefixes P, or -1.
__device__ __forceinline__ int sg_sel_group(const uint64_t* P, int ng, uint64_t hp)
{
    int lo = 0, hi = ng;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (P[mid] < hp) lo = mid + 1;
        else hi = mid;
    }
    return (lo < ng && P[lo] == hp) ? lo : -1;
}

// Histograms of pass `pass`: hist[s][g][256] += values of column s in group g by digit.
// Grid (blocks per column, S); 256 threads.
__global__ __launch_bounds__(256) void k_sel_hist(const double* __restrict__ xc, int64_t n, int32_t pass, int32_t R,
                                                  const int32_t* __restrict__ ngrp, const uint64_t* __restrict__ gpref,
                                                  uint32_t* __restrict__ hist)
{
    __shared__ uint32_t lh[SG_SEL_LDS_GROUPS * 256];
    __shared__ uint64_t lp[SG_SEL_LDS_GROUPS];
    const int32_t s = blockIdx.y;
    const int ng = ngrp[s];
    const uint64_t* gp = gpref + (int64_t)s * R;
    uint32_t* gh = hist + (int64_t)s * R * 256;
    const bool local = ng <= SG_SEL_LDS_GROUPS;   // block-uniform
    if (local) {
        for (int i = threadIdx.x; i < ng * 256; i += 256) lh[i] = 0u;
        for (int i = threadIdx.x; i < ng; i += 256) lp[i] = gp[i];
    }
    __syncthreads();
    const int ds = 56 - 8 * pass;          // shift of this pass's digit
    const double* col = xc + (int64_t)s * n;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const uint64_t key = sg_dkey(col[t]);
        int g = 0;
        if (pass > 0) {
            const uint64_t hp = key >> (ds + 8);
            g = local ? sg_sel_group(lp, ng, hp) : sg_sel_group(gp, ng, hp);
            if (g < 0) continue;
        }
        const uint32_t d = (uint32_t)(key >> ds) & 255u;
        if (local) atomicAdd(&lh[g * 256 + d], 1u);
        else atomicAdd(&gh[(int64_t)g * 256 + d], 1u);
    }
    if (local) {
        __syncthreads();
        for (int i = threadIdx.x; i < ng * 256; i += 256)
            if (lh[i]) atomicAdd(&gh[i], lh[i]);
    }
}

// After pass `pass`: each target of column s (one workgroup per column) takes the digit its
// rank falls in, the histograms read are cleared, and the groups are rebuilt as runs of
// equal prefixes.
__global__ __launch_bounds__(256) void k_sel_refine(int32_t R, int64_t* __restrict__ rank, uint64_t* __restrict__ pref,
                                                    int32_t* __restrict__ grp, int32_t* __restrict__ ngrp,
                                                    uint64_t* __restrict__ gpref, uint32_t* __restrict__ hist)
{
    __shared__ int32_t sc[256];
    __shared__ int32_t s_base;
    const int32_t s = blockIdx.x;
    int64_t* rk = rank + (int64_t)s * R;
    uint64_t* pf = pref + (int64_t)s * R;
    int32_t* gr = grp + (int64_t)s * R;
    uint64_t* gp = gpref + (int64_t)s * R;
    uint32_t* gh = hist + (int64_t)s * R * 256;
    const int ng = ngrp[s];
    for (int r = threadIdx.x; r < R; r += 256) {
        const uint32_t* h = gh + (int64_t)gr[r] * 256;
        int64_t k = rk[r];
        uint32_t d = 0;
        while (d < 255u && k >= (int64_t)h[d]) k -= h[d++];
        rk[r] = k;
        pf[r] = (pf[r] << 8) | d;
    }
    if (threadIdx.x == 0) s_base = 0;
    __syncthreads();
    for (int64_t i = threadIdx.x; i < (int64_t)ng * 256; i += 256) gh[i] = 0u;
    for (int r0 = 0; r0 < R; r0 += 256) {
        const int r = r0 + threadIdx.x;
        const int f = (r < R && (r == 0 || pf[r] != pf[r - 1])) ? 1 : 0;
        sc[threadIdx.x] = f;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {   // inclusive scan of the run starts
            const int v = threadIdx.x >= o ? sc[threadIdx.x - o] : 0;
            __syncthreads();
            sc[threadIdx.x] += v;
            __syncthreads();
        }
        if (r < R) {
            const int g = s_base + sc[threadIdx.x] - 1;
            gr[r] = g;
            if (f) gp[g] = pf[r];
        }
        __syncthreads();
        if (threadIdx.x == 0) s_base += sc[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) ngrp[s] = s_base;
}

// out[s * J + j] = the value of target umap[j] of column s.
__global__ void k_sel_out(const uint64_t* __restrict__ pref, int32_t R, const int32_t* __restrict__ umap, int32_t J,
                          int32_t S, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)J * S) return;
    const int64_t s = i / J;
    out[i] = sg_dval(pref[s * R + umap[i - s * J]]);
}

hipError_t sg_sensor_select(const double* xc, int64_t n, int32_t S, int32_t R, int64_t* rank, uint64_t* pref,
                            int32_t* grp, int32_t* ngrp, uint64_t* gpref, uint32_t* hist, const int32_t* umap,
                            int32_t J, double* out, hipStream_t s)
{
    const int64_t per_col = std::max<int64_t>(1, std::min<int64_t>((n + 4095) / 4096, std::max<int64_t>(1, 4096 / S)));
    for (int32_t pass = 0; pass < 8; pass++) {
        hipLaunchKernelGGL(k_sel_hist, dim3((unsigned)per_col, (unsigned)S), dim3(256), 0, s, xc, n, pass, R, ngrp, gpref,
                           hist);
        hipLaunchKernelGGL(k_sel_refine, dim3((unsigned)S), dim3(256), 0, s, R, rank, pref, grp, ngrp, gpref, hist);
    }
    const int64_t m = (int64_t)J * S;
    hipLaunchKernelGGL(k_sel_out, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, pref, R, umap, J, S, out);
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

hipError_t sg_online_bins(const double* x, int32_t S, int32_t B, int32_t window, int32_t stride, int64_t n_win,
                          double* E, int32_t* items, uint32_t* present, int64_t* off, hipStream_t s)
{
    const int32_t lanes = S * (B - 1);
    if (n_win > 0) {
        hipLaunchKernelGGL(k_online_p2, dim3((unsigned)((lanes + 63) / 64)), dim3(64), 0, s, x, S, B, window, stride,
                           n_win, E);
        const int64_t m = n_win * S;
        hipLaunchKernelGGL(k_online_bins, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, x, S, B, window, stride,
                           n_win, E, items, present);
    }
    hipLaunchKernelGGL(k_sensor_offsets, dim3((unsigned)((n_win + 1 + 255) / 256)), dim3(256), 0, s, n_win, S, off);
    return hipGetLastError();
}
