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
__device__ __forceinline__ void factor_gradient(T *gA, int stride, T const *H, T const *X, int st, int nfib, T *red)
{
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
    if (wave == 0)
    {
        kronreg::global_ptr<T> g = kronreg::as_global(gA);
        int const c = lane & 15;
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            int const r = M::row(lane, v);
            if (r < N && c < N) atomicAdd((T *)(g + r + c * stride), acc[v]);
        }
    }
}

// InLds: the working vectors live in LDS (dynamic shared memory), otherwise in ws[b].
// Store: keep the chain X_0..X_{d-2} (d-1 vectors) instead of recomputing each X_t.
// Scratch layout (elements): H [size], then Store ? X_0..X_{d-2} [(d-1) size] : X [size].
template<typename T, int N, bool InLds, bool Store>
__global__ void __launch_bounds__(kThreads)
    kron_adjoint_kernel(int d, int size, T const *const *__restrict__ mats, int stride, T const *const *__restrict__ in,
                        T const *const *__restrict__ gout, T *const *__restrict__ gin, T *const *__restrict__ gmats,
                        T *const *__restrict__ ws)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int const b = blockIdx.x;
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
            factor_gradient<T, N>(gmats[static_cast<size_t>(b) * d + t], stride, H, Xt, st, nfib, red);
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

} // namespace kronadj
