// kron_adjoint.hpp -- adjoint of the batched Kronecker mat-vec (no counterpart in the reference).
//
// Forward, per entry k with factors A_0..A_{d-1} (A_0 owns the slowest index) and output o(k):
//     y_{o(k)} += K_k x_k,   K_k = A_0 (x) ... (x) A_{d-1}.
// For a cotangent g on the outputs (L = sum_o <g_o, y_o>), entry k contributes
//     d x_k  = K_k^T g_{o(k)} = (A_0^T (x) ... (x) A_{d-1}^T) g_{o(k)},
//     d A_t  = H_t,(t) . X_t,(t)^T      (n x n; (t) = mode-t unfolding, contraction over the other modes)
// with X_t = (I (x) .. (x) I (x) A_{t+1} (x) .. (x) A_{d-1}) x_k   (the forward sweep's intermediate once the
//      factors after t have been applied) and
//      H_t = (A_0^T (x) .. (x) A_{t-1}^T (x) I (x) .. (x) I) g     (the transposed sweep run from the front).
// H_0 = g, H_{t+1} = A_t^T applied to mode t of H_t, and d x_k = H_d.
//
// One workgroup per entry. Mode contractions use the fiber layout of the forward fiber sweep (each thread
// owns whole fibers, in place). The d A_t contraction is a small GEMM (n x n^(d-1)) . (n^(d-1) x n) and
// runs on the matrix cores: V_MFMA_F64_16X16X4_F64 / V_MFMA_F32_16X16X4_F32, K = 4 fibers per
// instruction, rows and columns beyond n zero-padded, one 16 x 16 accumulator per wavefront, reduced across
// wavefronts in LDS. Results are accumulated with hardware atomic adds (entries may share a factor, an
// input or an output: the gradients of shared storage are summed).
//
// Intermediates X_t: RECOMPUTE rebuilds X_t from x for every t (d(d-1)/2 extra mode steps, two vectors of
// scratch); STORE computes the chain X_{d-2}, .., X_0 once and keeps it (d-1 vectors of scratch).
// Scratch lives in LDS when it fits, otherwise in the caller's per-entry workspace.
#pragma once

#include <hip/hip_runtime.h>

#include "kron_reg.hpp"

namespace kronadj
{

constexpr int kThreads = 256;
constexpr int kMaxN = 8;  // compile-time fiber lengths 1..kMaxN

// One K = 4 step of a 16 x 16 product on the matrix cores. Operand layout (both types): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; the accumulator layout differs between FP64 and FP32 (row()).
template<typename T>
struct Mfma;
template<>
struct Mfma<double>
{
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t step(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // accumulator slot v of `lane` holds D[row][lane % 16] (measured on gfx950)
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template<>
struct Mfma<float>
{
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t step(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// first element of fiber f of the mode with stride st (fibers of that mode: n^(d-1))
template<int N>
__device__ __forceinline__ int fiber_base(int f, int st)
{
    int const hi = f / st;
    return hi * st * N + (f - hi * st);
}

// dst = (mode t of) A applied to src, or A^T when Trans; fiber-owned, so dst may equal src.
template<typename T, int N, bool Trans>
__device__ __forceinline__ void mode_step(T *dst, T const *src, T const *A_generic, int stride, int st, int nfib)
{
    kronreg::constant_ptr<T> A = kronreg::as_constant(A_generic);
    T a[N][N];
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int k = 0; k < N; ++k) a[r][k] = Trans ? A[k + r * stride] : A[r + k * stride];
    for (int f = threadIdx.x; f < nfib; f += blockDim.x)
    {
        int const base = fiber_base<N>(f, st);
        T v[N];
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = src[base + k * st];
#pragma unroll
        for (int r = 0; r < N; ++r)
        {
            T acc = T(0);
#pragma unroll
            for (int k = 0; k < N; ++k) acc = fma(v[k], a[r][k], acc);
            dst[base + r * st] = acc;
        }
    }
}

// gA[r + c*stride] += sum over fibers f of H[base_f + r*st] * X[base_f + c*st]  (r, c < N).
// red: LDS scratch of (blockDim.x / 64) * 256 elements (the 16 x 16 accumulator of every wavefront).
template<typename T, int N>
__device__ __forceinline__ void grad_vals(T const *H, T const *X, int st, int nfib, T *red, T (&out)[4])
{
    if constexpr (std::is_same<T, double>::value && (N <= 4 || N == 8))
    {
        // FP64, n <= 4 or n = 8: V_MFMA_F64_4X4X4_4B_F64, four 4 x 4 blocks per instruction; lane
        // 16 kb + 4 jb + i supplies A_jb[i][kb] and B_jb[kb][i], D_jb[i][s] lands in lane 16 i + 4 jb + s.
        // n <= 4: block jb sums its own fibers (16 fibers per instruction), the blocks are added by two lane
        // shuffles. n = 8: block jb = (rb, cb) is quarter (rb, cb) of dA, 4 fibers per instruction.
        int const lane = threadIdx.x & 63;
        int const wave = threadIdx.x >> 6;
        int const nwaves = blockDim.x >> 6;
        int const i = lane & 3, jb = (lane >> 2) & 3, kb = lane >> 4;
        double acc = 0.0;
        if constexpr (N <= 4)
        {
            int const ngroups = (nfib + 15) >> 4;
            for (int grp = wave; grp < ngroups; grp += nwaves)
            {
                int const f = grp * 16 + jb * 4 + kb;
                double a = 0.0, b = 0.0;
                if (i < N && f < nfib)
                {
                    int const base = fiber_base<N>(f, st);
                    a = H[base + i * st];
                    b = X[base + i * st];
                }
                acc = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc, 0, 0, 0);
            }
            acc += __shfl_xor(acc, 4);
            acc += __shfl_xor(acc, 8);
        }
        else
        {
            int const rb = jb >> 1, cb = jb & 1;
            int const ngroups = (nfib + 3) >> 2;
            for (int grp = wave; grp < ngroups; grp += nwaves)
            {
                int const f = grp * 4 + kb;
                double a = 0.0, b = 0.0;
                if (f < nfib)
                {
                    int const base = fiber_base<N>(f, st);
                    a = H[base + (4 * rb + i) * st];
                    b = X[base + (4 * cb + i) * st];
                }
                acc = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc, 0, 0, 0);
            }
        }
        if (nwaves > 1)
        {
            red[wave * 64 + lane] = acc;
            __syncthreads();
            if (wave == 0)
                for (int w = 1; w < nwaves; ++w) acc += red[w * 64 + lane];
        }
        out[0] = acc;
        out[1] = out[2] = out[3] = T(0);
        return;
    }
    using M = Mfma<T>;
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const nwaves = blockDim.x >> 6;
    int const i = lane & 15;  // row of the A operand / column of the B operand
    int const kk = lane >> 4; // fiber within the group of 4
    typename M::acc_t acc = {T(0), T(0), T(0), T(0)};
    int const ngroups = (nfib + 3) >> 2;
    for (int grp = wave; grp < ngroups; grp += nwaves)
    {
        int const f = grp * 4 + kk;
        T a = T(0), b = T(0);
        if (i < N && f < nfib)
        {
            int const base = fiber_base<N>(f, st);
            a = H[base + i * st];
            b = X[base + i * st];
        }
        acc = M::step(a, b, acc);
    }
    // lane holds D[M::row(lane, v)][lane % 16], v < 4
    if (nwaves > 1)
    {
#pragma unroll
        for (int v = 0; v < 4; ++v) red[wave * 256 + v * 64 + lane] = acc[v];
        __syncthreads();
        if (wave == 0)
        {
            for (int w = 1; w < nwaves; ++w)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[v] += red[w * 256 + v * 64 + lane];
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v] = acc[v];
}

// Adds the factor gradient held by wavefront 0 (grad_vals layout) to gA.
template<typename T, int N>
__device__ __forceinline__ void grad_flush(T *gA, int stride, T const (&val)[4])
{
    int const lane = threadIdx.x & 63;
    kronreg::global_ptr<T> g = kronreg::as_global(gA);
    if constexpr (std::is_same<T, double>::value && (N <= 4 || N == 8))
    {
        int const jb = (lane >> 2) & 3;
        int const r = lane >> 4, c = lane & 3;  // lane 16 r + 4 jb + c holds block jb's [r][c]
        int const row = N <= 4 ? r : 4 * (jb >> 1) + r;
        int const col = N <= 4 ? c : 4 * (jb & 1) + c;
        if (row < N && col < N && (N > 4 || jb == 0)) atomicAdd((T *)(g + row + col * stride), val[0]);
    }
    else
    {
        int const c = lane & 15;
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            int const r = Mfma<T>::row(lane, v);
            if (r < N && c < N) atomicAdd((T *)(g + r + c * stride), val[v]);
        }
    }
}

// InLds: the working vectors live in LDS (dynamic shared memory), otherwise in ws[b].
// Store: keep the chain X_0..X_{d-2} (d-1 vectors) instead of recomputing each X_t.
// Scratch layout (elements): H [size], then Store ? X_0..X_{d-2} [(d-1) size] : X [size].
// Factor gradients of up to kRunD factors are summed across the consecutive entries of a workgroup's run.
constexpr int kRunD = 8;

template<typename T, int N, bool InLds, bool Store>
__device__ __forceinline__ void kron_adjoint_entry(int b, int d, int size, T const *const *__restrict__ mats,
                                                   int stride, T const *const *__restrict__ in,
                                                   T const *const *__restrict__ gout, T *const *__restrict__ gin,
                                                   T *const *__restrict__ gmats, T *const *__restrict__ ws,
                                                   unsigned char *smem_raw, T (&dacc)[kRunD][4], T *(&dptr)[kRunD],
                                                   bool wave0)
{
    T const *const *A = mats + static_cast<size_t>(b) * d;
    T *const lds = reinterpret_cast<T *>(smem_raw);
    T *const H = InLds ? lds : ws[b];
    T *const chain = H + size;
    T *const red = InLds ? lds + (Store ? static_cast<size_t>(d) : 2) * size : lds;
    T const *const x = in[b];
    T const *const g = gout[b];
    bool const want_mats = gmats != nullptr;
    bool const want_input = gin != nullptr;
    int const nfib = size / N;

    for (int e = threadIdx.x; e < size; e += blockDim.x) H[e] = g[e];
    if (Store && want_mats)
    {
        // X_{d-1} = x; X_{t-1} = A_t on mode t of X_t
        T const *src = x;
        int st = 1;
        for (int t = d - 1; t >= 1; --t)
        {
            T *dst = chain + static_cast<size_t>(t - 1) * size;
            mode_step<T, N, false>(dst, src, A[t], stride, st, nfib);
            __syncthreads();
            src = dst;
            st *= N;
        }
    }
    __syncthreads();

    int st = size / N;  // stride of mode 0
    for (int t = 0; t < d; ++t)
    {
        if (want_mats)
        {
            T const *Xt;
            if (t == d - 1)
                Xt = x;
            else if (Store)
                Xt = chain + static_cast<size_t>(t) * size;
            else
            {
                T *X = chain;
                int s2 = 1;
                // modes d-1 .. t+1 applied to x; the first step reads x directly
                for (int u = d - 1; u > t; --u)
                {
                    mode_step<T, N, false>(X, u == d - 1 ? x : X, A[u], stride, s2, nfib);
                    __syncthreads();
                    s2 *= N;
                }
                Xt = X;
            }
            T vals[4];
            grad_vals<T, N>(H, Xt, st, nfib, red, vals);
            T *const gA = gmats[static_cast<size_t>(b) * d + t];
            if (t < kRunD)
            {
#pragma unroll
                for (int tt = 0; tt < kRunD; ++tt)
                    if (tt == t)
                    {
                        if (dptr[tt] != gA)
                        {
                            if (wave0 && dptr[tt] != nullptr) grad_flush<T, N>(dptr[tt], stride, dacc[tt]);
                            dptr[tt] = gA;
#pragma unroll
                            for (int v = 0; v < 4; ++v) dacc[tt][v] = T(0);
                        }
#pragma unroll
                        for (int v = 0; v < 4; ++v) dacc[tt][v] += vals[v];
                    }
            }
            else if (wave0)
                grad_flush<T, N>(gA, stride, vals);
            __syncthreads();
        }
        if (t < d - 1 || want_input)
        {
            mode_step<T, N, true>(H, H, A[t], stride, st, nfib);
            __syncthreads();
        }
        st /= N;
    }
    if (want_input)
    {
        kronreg::global_ptr<T> gi = kronreg::as_global(gin[b]);
        for (int e = threadIdx.x; e < size; e += blockDim.x) atomicAdd((T *)(gi + e), H[e]);
    }
}

// A workgroup walks a run of entries [e0, e1) (logical run index as in the forward kernels). Factor
// gradients of consecutive entries that write the same storage (shared factors) are summed in registers
// of wavefront 0 -- for factors t < kRunD -- and added with one atomic per element when the storage
// changes or the run ends.

template<typename T, int N, bool InLds, bool Store>
__global__ void __launch_bounds__(kThreads)
    kron_adjoint_kernel(int nb, int epw, int d, int size, T const *const *__restrict__ mats, int stride,
                        T const *const *__restrict__ in, T const *const *__restrict__ gout, T *const *__restrict__ gin,
                        T *const *__restrict__ gmats, T *const *__restrict__ ws)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int const e0 = kronreg::xcd_run_index() * epw;
    int const e1 = (e0 + epw < nb) ? e0 + epw : nb;
    bool const wave0 = threadIdx.x < 64;
    T dacc[kRunD][4];
    T *dptr[kRunD];
#pragma unroll
    for (int tt = 0; tt < kRunD; ++tt)
    {
        dptr[tt] = nullptr;
#pragma unroll
        for (int v = 0; v < 4; ++v) dacc[tt][v] = T(0);
    }
    for (int b = e0; b < e1; ++b)
    {
        kron_adjoint_entry<T, N, InLds, Store>(b, d, size, mats, stride, in, gout, gin, gmats, ws, smem_raw, dacc,
                                               dptr, wave0);
        __syncthreads();  // the next entry overwrites the scratch
    }
    if (wave0)
    {
#pragma unroll
        for (int tt = 0; tt < kRunD; ++tt)
            if (dptr[tt] != nullptr) grad_flush<T, N>(dptr[tt], stride, dacc[tt]);
    }
}


} // namespace kronadj
