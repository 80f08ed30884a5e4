// kronmult_hip.hip -- batched Kronecker mat-vec for MI355X (gfx950): kernels and the C ABI.
//
//   output[k] += kron(A_{k,0}, ..., A_{k,d-1}) · input[k]
//
// The arithmetic is the reference's backward sweep (kronmult_gpu/kronmult.cu:95-130, the "variant of the
// backward version of Algorithm 993", README.md:13-16): factors are consumed from the last (fastest index)
// to the first, each step is an n-term dot product per output, summed k ascending from zero with one fused
// multiply-add per term. Here a step is an in-place contraction of one mode: each thread owns whole fibers
// (the n values that differ only in that mode), so no rotation buffer is needed and the values produced
// are bit-identical to the reference sweep evaluated with fused multiply-adds.
//
// Kernel families:
//   * register-resident sweep (kron_reg.hpp) for the (d, n) plans listed in reg_plan(): the tensor lives
//     in registers, n^RHO values per thread, with LDS only between phases; a workgroup walks a run of
//     entries and accumulates same-output contributions in registers before one atomic flush;
//   * fiber sweep (any d, n <= 64): the n^d vector lives in LDS (<= 64 KiB) or, beyond that, in the
//     caller's per-entry workspace; one workgroup per entry; shared outputs are accumulated with
//     hardware FP64/FP32 atomics (global_atomic_add_f64/f32).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "kron_adjoint.hpp"
#include "kron_group.hpp"
#include "kron_mfma.hpp"
#include "kron_ordered.hpp"
#include "kron_pass.hpp"
#include "kron_split.hpp"
#include "kron_split_reg.hpp"
#include "kron_reg.hpp"
#include "kron_m8.hpp"
#include "kron_wave.hpp"
#include "kron_reg_adjoint.hpp"
#include "kron_reg_adjoint2.hpp"
#include "kronmult_hip.h"

namespace
{

constexpr int kMaxFiber = 64;           // largest n of the per-thread fiber kernels (wide-fiber kernel beyond)
constexpr int kLdsBudget = 64 * 1024;   // bytes of LDS one workgroup may take for the n^d vector
constexpr int kFiberThreads = 256;      // workgroup size of the fiber sweep

// ------------------------------------------------------------------------------------------------------
// Fiber sweep.
//
// Mode t (0 = slowest) of the canonical n^d layout has stride st = n^(d-1-t). Contracting it maps the
// fiber {base + k*st, k < n} to {base + r*st, r < n} with y_r = sum_k A_t[r, k] * x_k. There are n^(d-1)
// fibers per step; fiber f has base = (f / st) * st * n + f % st.
//
// N > 0 (n <= 10): compile-time fiber length, coefficients held in registers (uniform across the workgroup, so the
// compiler loads them with scalar loads). N == 0: runtime n <= kMaxFiber, coefficients staged in LDS.
// InLds: the vector lives in LDS; otherwise in ws[b] (the first step reads the input directly).
// ORDERED: each entry is a piece (kron_ordered.hpp): added to its output when it is the output's only
// entry, else left in ws[b] for the reduction.
// ------------------------------------------------------------------------------------------------------
template<typename T, int N, bool InLds, bool ORDERED>
__global__ void __launch_bounds__(kFiberThreads)
    kron_fiber_kernel(int d, int n_rt, int size, int nb, T const *const *__restrict__ mats, int stride,
                      T const *const *__restrict__ in, T *const *__restrict__ out, T *const *__restrict__ ws)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int const n = (N > 0) ? N : n_rt;
    int const b = blockIdx.x;
    T const *const *A_list = mats + static_cast<size_t>(b) * d;
    T const *src = in[b];
    T *x = InLds ? reinterpret_cast<T *>(smem_raw) : ws[b];
    T *coef = InLds ? reinterpret_cast<T *>(smem_raw) + size : reinterpret_cast<T *>(smem_raw);

    if (InLds)
    {
        for (int i = threadIdx.x; i < size; i += blockDim.x) x[i] = src[i];
    }

    int const nfib = size / n;
    int st = 1;
    for (int t = d - 1; t >= 0; --t)
    {
        T const *A = A_list[t];
        T const *rd = (!InLds && t == d - 1) ? src : x;
        if constexpr (N > 0)
        {
            T a[N][N];
#pragma unroll
            for (int r = 0; r < N; ++r)
#pragma unroll
                for (int k = 0; k < N; ++k) a[r][k] = A[r + static_cast<size_t>(k) * stride];
            __syncthreads();  // previous step's writes visible
            for (int f = threadIdx.x; f < nfib; f += blockDim.x)
            {
                int const hi = f / st;
                int const base = hi * st * N + (f - hi * st);
                T v[N];
#pragma unroll
                for (int k = 0; k < N; ++k) v[k] = rd[base + k * st];
#pragma unroll
                for (int r = 0; r < N; ++r)
                {
                    T acc = T(0);
#pragma unroll
                    for (int k = 0; k < N; ++k) acc = fma(v[k], a[r][k], acc);
                    x[base + r * st] = acc;
                }
            }
        }
        else
        {
            __syncthreads();  // previous step done with coef, its writes visible
            for (int i = threadIdx.x; i < n * n; i += blockDim.x)
            {
                int const r = i / n;
                int const k = i - r * n;
                coef[i] = A[r + static_cast<size_t>(k) * stride];
            }
            __syncthreads();
            for (int f = threadIdx.x; f < nfib; f += blockDim.x)
            {
                int const hi = f / st;
                int const base = hi * st * n + (f - hi * st);
                T v[kMaxFiber];
                for (int k = 0; k < n; ++k) v[k] = rd[base + k * st];
                for (int r = 0; r < n; ++r)
                {
                    T acc = T(0);
                    for (int k = 0; k < n; ++k) acc = fma(v[k], coef[r * n + k], acc);
                    x[base + r * st] = acc;
                }
            }
        }
        st *= n;
    }
    __syncthreads();
    T *y = out[b];
    if constexpr (ORDERED)
    {
        if (kronordered::whole_segment(out, nb, b, b + 1, y))
        {
            for (int i = threadIdx.x; i < size; i += blockDim.x) y[i] = y[i] + x[i];
        }
        else if (InLds)
        {
            T *w = ws[b];
            for (int i = threadIdx.x; i < size; i += blockDim.x) w[i] = x[i];
        }
    }
    else
    {
        for (int i = threadIdx.x; i < size; i += blockDim.x) atomicAdd(y + i, x[i]);
    }
}

// ------------------------------------------------------------------------------------------------------
// Wide-fiber sweep (n > kMaxFiber, any d): one workgroup per entry, no per-thread fiber copy. Each mode is
// contracted in chunks of G fibers: the chunk's n*G inputs are staged in LDS (thread idx -> fiber idx % G,
// element idx / G, so consecutive threads read consecutive fibers: contiguous in memory when G <= stride),
// then every (fiber, row) output is one fused chain over k ascending from zero -- the reference's step
// (kronmult_gpu/kronmult.cu:66-70) -- written in place into the entry's vector, which lives in LDS when it
// fits next to the chunk, else in the caller's workspace (the first mode reads the input). Coefficients
// come from the factor itself (global memory, read through the caches: n*n values shared by the workgroup).
// ------------------------------------------------------------------------------------------------------
constexpr int kWideThreads = 256;
constexpr int kWideChunkBytes = 32 * 1024;  // staged fibers per chunk (n <= 4096 in FP64, 8192 in FP32)

template<typename T, bool InLds, bool ORDERED>
__global__ void __launch_bounds__(kWideThreads)
    kron_wide_kernel(int d, int n, int size, int nb, int chunk, T const *const *__restrict__ mats, int stride,
                     T const *const *__restrict__ in, T *const *__restrict__ out, T *const *__restrict__ ws)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int const b = blockIdx.x;
    T *const buf = reinterpret_cast<T *>(smem_raw);             // chunk * n staged fiber values
    T *const x = InLds ? buf + static_cast<size_t>(chunk) * n : ws[b];
    T const *const src = in[b];
    T const *const *A_list = mats + static_cast<size_t>(b) * d;
    int const nfib = size / n;
    int st = 1;
    for (int t = d - 1; t >= 0; --t)
    {
        T const *const A = A_list[t];
        T const *const rd = (t == d - 1) ? src : x;
        for (int f0 = 0; f0 < nfib; f0 += chunk)
        {
            int const g = (nfib - f0 < chunk) ? nfib - f0 : chunk;
            __syncthreads();  // previous chunk (or mode, or the input copy) done with buf / x
            for (int i = threadIdx.x; i < g * n; i += blockDim.x)
            {
                int const fl = i % g, k = i / g;
                int const f = f0 + fl;
                int const hi = f / st;
                buf[fl * n + k] = rd[hi * st * n + (f - hi * st) + k * st];
            }
            __syncthreads();
            for (int i = threadIdx.x; i < g * n; i += blockDim.x)
            {
                int const fl = i % g, r = i / g;
                int const f = f0 + fl;
                int const hi = f / st;
                T const *const xf = buf + fl * n;
                T acc = T(0);
                for (int k = 0; k < n; ++k) acc = fma(xf[k], A[r + static_cast<size_t>(k) * stride], acc);
                x[hi * st * n + (f - hi * st) + r * st] = acc;
            }
        }
        st *= n;
    }
    __syncthreads();
    T *const y = out[b];
    if constexpr (ORDERED)
    {
        if (kronordered::whole_segment(out, nb, b, b + 1, y))
        {
            for (int i = threadIdx.x; i < size; i += blockDim.x) y[i] = y[i] + x[i];
        }
        else if (InLds)
        {
            T *const w = ws[b];
            for (int i = threadIdx.x; i < size; i += blockDim.x) w[i] = x[i];
        }
    }
    else
    {
        for (int i = threadIdx.x; i < size; i += blockDim.x) atomicAdd(y + i, x[i]);
    }
}

// ------------------------------------------------------------------------------------------------------
// Dispatch.
// ------------------------------------------------------------------------------------------------------
struct Shape
{
    int d, n, stride, nb;
    long size;
    bool ordered = false;  // ordered accumulation (kron_ordered.hpp) instead of atomics
};

hipError_t validate(int d, int n, int stride, int nb, Shape &s)
{
    if (d < 1 || n < 1 || nb < 0 || stride < n) return hipErrorInvalidValue;
    long size = 1;
    for (int i = 0; i < d; ++i)
    {
        size *= n;
        if (size > INT_MAX) return hipErrorInvalidValue;
    }
    s = Shape{d, n, stride, nb, size, false};
    return hipSuccess;
}

// Register-resident plans (d, n) -> RHO. Threads per workgroup P = n^(d-RHO) must be a multiple of 64
// and at most 1024; the n^d elements of an entry stay in LDS between phases.
constexpr int reg_rho(int d, int n)
{
    return (d == 6 && n == 4) ? 2 : (d == 5 && n == 4) ? 2 : (d == 4 && n == 4) ? 1
         : (d == 4 && n == 8) ? 1 : (d == 3 && n == 8) ? 1 : 0;
}

// n = 8 FP64 plans run on the matrix-core sweep (kron_m8.hpp): the vector sweep feeds each FMA one factor
// coefficient through an LDS broadcast and is LDS-bound there.
template<typename T>
constexpr bool m8_plan(int d, int n)
{
    return sizeof(T) == 8 && n == 8 && (d == 3 || d == 4);
}

// Number of compute units of the current device (cached per device).
int device_cu_count()
{
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v == 0)
    {
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cache[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}

// Resident workgroups per CU of one kernel (cached per device, kernel address and block size when the
// launch takes no dynamic LDS: then it depends only on the code object; otherwise queried per call, the LDS
// varying with the shape). Kernels of one plan family share a function type, so the key is the address.
template<typename Kernel>
int resident_per_cu(Kernel kernel, int block, size_t dyn_lds)
{
    static std::mutex lock;
    static std::map<std::tuple<int, void const *, int>, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = -1;
    auto const key = std::make_tuple(dev, reinterpret_cast<void const *>(kernel), block);
    if (dyn_lds == 0)
    {
        std::lock_guard<std::mutex> guard(lock);
        auto const it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    int v = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kernel, block, dyn_lds) != hipSuccess || v <= 0) v = 1;
    if (dyn_lds == 0)
    {
        std::lock_guard<std::mutex> guard(lock);
        cache[key] = v;
    }
    return v;
}

struct Launch
{
    int family = 0;
    long grid = 0;
    int block = 0;
    int epw = 1;
    size_t lds = 0;
    int resident = 0;
    bool workspace = false;
};

// ---- register family --------------------------------------------------------------------------------
// grid = CUs x workgroups per CU (a multiple of 8, so every XCD gets the same number of runs); each
// workgroup walks ceil(nb / grid) consecutive entries. The default number of workgroups per CU was
// measured on MI355X (launch-shape sweeps of the d=6 n=4, d=3 n=8 and d=4 n=8 batches): 4 workgroups of 4
// or 8 wavefronts per CU (the runs are then short enough to balance), 8 one-wavefront workgroups.
constexpr int default_workgroups_per_cu(int threads)
{
    return threads >= 512 ? 4 : threads >= 256 ? 4 : threads >= 128 ? 4 : 8;
}
template<typename T, int D, int N>
hipError_t reg_launch(Shape const &s, kronmult_tuning_t const &tu, Launch &L, T const *const *mats,
                      T const *const *in, T *const *out, T *const *ws, hipStream_t stream, bool run)
{
    using PL = kronplan::Plan<D, N, reg_rho(D, N)>;
    static_assert(PL::P % 64 == 0 && PL::P <= 1024, "plan must fill whole wavefronts");
    auto kernel = s.ordered ? kronreg::kron_reg_kernel<T, PL, true> : kronreg::kron_reg_kernel<T, PL, false>;
    L.lds = sizeof(T) * (PL::SIZE + (kronreg::fwd_coef_in_lds<T, PL>() ? kronreg::Coef<T, PL>::kCount : 0));
    if constexpr (m8_plan<T>(D, N))
    {
        // n = 8 FP64: the matrix-core sweep (kron_m8.hpp), same threads per entry and run structure
        using MP = kronm8::M8Plan<D>;
        static_assert(MP::P == PL::P, "one thread per fiber of eight values in both plans");
        kernel = s.ordered ? kronm8::kron_m8_kernel<D, true> : kronm8::kron_m8_kernel<D, false>;
        L.lds = sizeof(double) * kronm8::lds_doubles<D>();
    }
    int const block = PL::P;
    L.family = KRONMULT_FAMILY_REGISTER;
    L.block = block;
    L.resident = resident_per_cu(kernel, block, 0);
    // by default no more workgroups than fit at once: a second round of workgroups starts only as the first
    // drains (measured on the d=4 n=8 FP32 batch: 3 resident, 118 us at 3 per CU against 149 us at 4)
    int const per_cu = tu.workgroups_per_cu > 0 ? tu.workgroups_per_cu
                                                : std::min(default_workgroups_per_cu(block), L.resident);
    long const slots = static_cast<long>(device_cu_count()) * per_cu;
    int epw = tu.entries_per_workgroup > 0 ? tu.entries_per_workgroup : static_cast<int>((s.nb + slots - 1) / slots);
    if (epw < 1) epw = 1;
    L.epw = epw;
    long const runs = (s.nb + epw - 1) / epw;
    L.grid = (runs + 7) / 8 * 8;  // surplus runs exit at once
    if (!run) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(L.grid)), dim3(block), 0, stream, s.nb, epw, mats,
                       s.stride, in, out, ws);
    return hipGetLastError();
}

// ---- matrix-core family (d = 6, n = 4, FP64; opt-in) -----------------------------------------------------
// Same run structure as the register family. Not the default: on MI355X the FP64 matrix rate equals the
// vector rate, and this kernel measured 67 us against the register sweep's 53 us (48 us with LDS-DMA
// staging) on the d=6 n=4 batch.
hipError_t mfma_launch(Shape const &s, kronmult_tuning_t const &tu, Launch &L, double const *const *mats,
                       double const *const *in, double *const *out, hipStream_t stream, bool run)
{
    auto kernel = kronmfma::kron_mfma_d6n4_kernel;
    L.family = KRONMULT_FAMILY_MATRIX;
    L.block = 256;
    L.lds = sizeof(double) * 4096;
    L.resident = resident_per_cu(kernel, 256, 0);
    int const per_cu = tu.workgroups_per_cu > 0 ? tu.workgroups_per_cu : 3;  // measured best: 3 per CU
    long const slots = static_cast<long>(device_cu_count()) * per_cu;
    int epw = tu.entries_per_workgroup > 0 ? tu.entries_per_workgroup : static_cast<int>((s.nb + slots - 1) / slots);
    if (epw < 1) epw = 1;
    L.epw = epw;
    long const runs = (s.nb + epw - 1) / epw;
    L.grid = (runs + 7) / 8 * 8;
    if (!run) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(L.grid)), dim3(256), 0, stream, s.nb, epw, mats, s.stride,
                       in, out);
    return hipGetLastError();
}

template<typename T>
hipError_t reg_dispatch(Shape const &s, kronmult_tuning_t const &tu, Launch &L, T const *const *mats,
                        T const *const *in, T *const *out, T *const *ws, hipStream_t stream, bool run)
{
    if (s.d == 6 && s.n == 4) return reg_launch<T, 6, 4>(s, tu, L, mats, in, out, ws, stream, run);
    if (s.d == 5 && s.n == 4) return reg_launch<T, 5, 4>(s, tu, L, mats, in, out, ws, stream, run);
    if (s.d == 4 && s.n == 4) return reg_launch<T, 4, 4>(s, tu, L, mats, in, out, ws, stream, run);
    if (s.d == 4 && s.n == 8) return reg_launch<T, 4, 8>(s, tu, L, mats, in, out, ws, stream, run);
    if (s.d == 3 && s.n == 8) return reg_launch<T, 3, 8>(s, tu, L, mats, in, out, ws, stream, run);
    return hipErrorInvalidValue;
}

// ---- fiber family -------------------------------------------------------------------------------------
template<typename T>
bool fiber_in_lds(Shape const &s)
{
    long const coef_bytes = (s.n <= 8) ? 0 : static_cast<long>(s.n) * s.n * sizeof(T);
    return s.size * static_cast<long>(sizeof(T)) + coef_bytes <= kLdsBudget;
}

template<typename T, int N, bool InLds>
hipError_t fiber_launch(Shape const &s, Launch &L, T const *const *mats, T const *const *in, T *const *out,
                        T *const *ws, hipStream_t stream, bool run)
{
    long const nfib = s.size / s.n;
    int threads = static_cast<int>(((nfib + 63) / 64) * 64);
    if (threads > kFiberThreads) threads = kFiberThreads;
    size_t lds = 0;
    if (InLds) lds += static_cast<size_t>(s.size) * sizeof(T);
    if (N == 0) lds += static_cast<size_t>(s.n) * s.n * sizeof(T);
    auto kernel = s.ordered ? kron_fiber_kernel<T, N, InLds, true> : kron_fiber_kernel<T, N, InLds, false>;
    L.family = KRONMULT_FAMILY_FIBER;
    L.grid = s.nb;
    L.block = threads;
    L.epw = 1;
    L.lds = lds;
    L.workspace = !InLds;
    L.resident = resident_per_cu(kernel, threads, lds);
    if (!run) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3(s.nb), dim3(threads), lds, stream, s.d, s.n, static_cast<int>(s.size), s.nb,
                       mats, s.stride, in, out, ws);
    return hipGetLastError();
}

template<typename T, bool InLds>
hipError_t fiber_dispatch_n(Shape const &s, Launch &L, T const *const *mats, T const *const *in, T *const *out,
                            T *const *ws, hipStream_t stream, bool run)
{
    switch (s.n)
    {
    case 1: return fiber_launch<T, 1, InLds>(s, L, mats, in, out, ws, stream, run);
    case 2: return fiber_launch<T, 2, InLds>(s, L, mats, in, out, ws, stream, run);
    case 3: return fiber_launch<T, 3, InLds>(s, L, mats, in, out, ws, stream, run);
    case 4: return fiber_launch<T, 4, InLds>(s, L, mats, in, out, ws, stream, run);
    case 5: return fiber_launch<T, 5, InLds>(s, L, mats, in, out, ws, stream, run);
    case 6: return fiber_launch<T, 6, InLds>(s, L, mats, in, out, ws, stream, run);
    case 7: return fiber_launch<T, 7, InLds>(s, L, mats, in, out, ws, stream, run);
    case 8: return fiber_launch<T, 8, InLds>(s, L, mats, in, out, ws, stream, run);
    case 9: return fiber_launch<T, 9, InLds>(s, L, mats, in, out, ws, stream, run);
    case 10: return fiber_launch<T, 10, InLds>(s, L, mats, in, out, ws, stream, run);
    default: return fiber_launch<T, 0, InLds>(s, L, mats, in, out, ws, stream, run);
    }
}

// ---- multi-pass family (vectors beyond the on-chip size, n <= 8) ---------------------------------------
// n <= 8: for n = 9, 10 the pass kernels' coefficients overflow the scalar registers, and the fiber sweep
// with a compile-time fiber length is faster (measured, 4608 entries: d=5 n=9 5.9 ms against 42 ms)
constexpr int kPassMaxN = 8;

// Largest power of n not above `limit` (at least 1).
long pow_floor(int n, long limit)
{
    long p = 1;
    while (p * n <= limit) p *= n;
    return p;
}

// TI: consecutive inner indices per tile, a power of n up to 16 (1 when the group holds the fastest mode)
template<int N>
constexpr int pass_ti()
{
    int p = 1;
    while (p * N <= 16) p *= N;
    return p;
}

template<typename T, int N, int G, int TI>
hipError_t pass_launch(kronpass::PassArgs const &P, long grid, size_t lds, T const *const *mats, int stride,
                       T const *const *in, T *const *out, T *const *ws, hipStream_t stream)
{
    hipLaunchKernelGGL((kronpass::kron_pass_kernel<T, N, G, TI>), dim3(static_cast<unsigned>(grid)),
                       dim3(kronpass::kThreads), lds, stream, P, mats, stride, in, ws, out);
    return hipGetLastError();
}

template<typename T, int N, int G>
hipError_t pass_launch_ti(kronpass::PassArgs const &P, int ti, long grid, size_t lds, T const *const *mats,
                          int stride, T const *const *in, T *const *out, T *const *ws, hipStream_t stream)
{
    if (ti == 1) return pass_launch<T, N, G, 1>(P, grid, lds, mats, stride, in, out, ws, stream);
    return pass_launch<T, N, G, pass_ti<N>()>(P, grid, lds, mats, stride, in, out, ws, stream);
}

constexpr int kPassMaxGroup = 4;  // modes per pass (compile-time group sizes 1..4)
constexpr int kPassLastRun = 8;   // entries per workgroup in the last pass (register sums of equal outputs)

template<typename T, int N>
hipError_t pass_run(Shape const &s, Launch &L, T const *const *mats, T const *const *in, T *const *out,
                    T *const *ws, hipStream_t stream, bool run)
{
    long const E = kronpass::kTileBytes / static_cast<long>(sizeof(T));  // tile elements
    constexpr int TIW = pass_ti<N>();
    int gmax = 1;  // the tile must hold one sub-tensor times TIW inner indices
    for (long p = N; p * N * TIW <= E && gmax < kPassMaxGroup; p *= N) ++gmax;
    int const passes = (s.d + gmax - 1) / gmax;
    L.family = KRONMULT_FAMILY_FIBER;
    L.block = kronpass::kThreads;
    L.epw = 1;
    L.workspace = true;
    L.resident = 0;
    int b = s.d;  // groups from the fastest: [b - g, b)
    for (int q = 0; q < passes; ++q)
    {
        int const g = s.d / passes + (q < s.d % passes ? 1 : 0);
        int const a = b - g;
        long ng = 1;
        for (int i = 0; i < g; ++i) ng *= N;
        long inner = 1;
        for (int i = b; i < s.d; ++i) inner *= N;
        long outer = 1;
        for (int i = 0; i < a; ++i) outer *= N;
        int const TI = inner == 1 ? 1 : TIW;
        long TO = pow_floor(N, E / (ng * TI));
        if (TO > outer) TO = outer;
        kronpass::PassArgs P{};
        P.d = s.d;
        P.a = a;
        P.inner = static_cast<int>(inner);
        P.TO = static_cast<int>(TO);
        P.tiles_i = static_cast<int>(inner / TI);
        P.tiles_per_entry = static_cast<int>((outer / TO) * (inner / TI));
        P.first = q == 0;
        P.last = q == passes - 1 && !s.ordered;  // ordered: the last pass also writes the workspace
        P.nb = s.nb;
        P.epw = P.last ? kPassLastRun : 1;
        long const grid = static_cast<long>((s.nb + P.epw - 1) / P.epw) * P.tiles_per_entry;
        size_t const lds = static_cast<size_t>(TO * ng * TI) * sizeof(T);
        if (q == 0)
        {
            L.grid = grid;
            L.lds = lds;
        }
        if (grid > INT_MAX) return hipErrorInvalidValue;
        if (run && grid > 0)
        {
            hipError_t err;
            switch (g)
            {
            case 1: err = pass_launch_ti<T, N, 1>(P, TI, grid, lds, mats, s.stride, in, out, ws, stream); break;
            case 2: err = pass_launch_ti<T, N, 2>(P, TI, grid, lds, mats, s.stride, in, out, ws, stream); break;
            case 3: err = pass_launch_ti<T, N, 3>(P, TI, grid, lds, mats, s.stride, in, out, ws, stream); break;
            default: err = pass_launch_ti<T, N, 4>(P, TI, grid, lds, mats, s.stride, in, out, ws, stream); break;
            }
            if (err != hipSuccess) return err;
        }
        b = a;
    }
    return hipSuccess;
}

template<typename T>
hipError_t pass_dispatch(Shape const &s, Launch &L, T const *const *mats, T const *const *in, T *const *out,
                         T *const *ws, hipStream_t stream, bool run)
{
    switch (s.n)
    {
    case 2: return pass_run<T, 2>(s, L, mats, in, out, ws, stream, run);
    case 3: return pass_run<T, 3>(s, L, mats, in, out, ws, stream, run);
    case 4: return pass_run<T, 4>(s, L, mats, in, out, ws, stream, run);
    case 5: return pass_run<T, 5>(s, L, mats, in, out, ws, stream, run);
    case 6: return pass_run<T, 6>(s, L, mats, in, out, ws, stream, run);
    case 7: return pass_run<T, 7>(s, L, mats, in, out, ws, stream, run);
    case 8: return pass_run<T, 8>(s, L, mats, in, out, ws, stream, run);
    default: return hipErrorInvalidValue;
    }
}

template<typename T>
bool use_passes(Shape const &s)
{
    return !fiber_in_lds<T>(s) && s.n >= 2 && s.n <= kPassMaxN;
}

// ---- wide-fiber family (n > kMaxFiber) -----------------------------------------------------------------
template<typename T>
bool wide_ok(Shape const &s)
{
    return static_cast<long>(s.n) * static_cast<long>(sizeof(T)) <= kWideChunkBytes;
}

// chunk of fibers staged per step, and whether the vector fits in LDS beside it
template<typename T>
void wide_shape(Shape const &s, int &chunk, bool &in_lds)
{
    long const nfib = s.size / s.n;
    long const fiber_bytes = static_cast<long>(s.n) * static_cast<long>(sizeof(T));
    long g = kWideChunkBytes / fiber_bytes;
    if (g > nfib) g = nfib;
    chunk = static_cast<int>(g < 1 ? 1 : g);
    in_lds = s.size * static_cast<long>(sizeof(T)) + chunk * fiber_bytes <= kLdsBudget;
}

template<typename T, bool InLds>
hipError_t wide_launch(Shape const &s, Launch &L, int chunk, T const *const *mats, T const *const *in,
                       T *const *out, T *const *ws, hipStream_t stream, bool run)
{
    size_t const lds = static_cast<size_t>(chunk) * s.n * sizeof(T) + (InLds ? static_cast<size_t>(s.size) * sizeof(T) : 0);
    auto kernel = s.ordered ? kron_wide_kernel<T, InLds, true> : kron_wide_kernel<T, InLds, false>;
    L.family = KRONMULT_FAMILY_FIBER;
    L.grid = s.nb;
    L.block = kWideThreads;
    L.epw = 1;
    L.lds = lds;
    L.workspace = !InLds;
    L.resident = resident_per_cu(kernel, kWideThreads, lds);
    if (!run) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3(s.nb), dim3(kWideThreads), lds, stream, s.d, s.n, static_cast<int>(s.size), s.nb,
                       chunk, mats, s.stride, in, out, ws);
    return hipGetLastError();
}

template<typename T>
hipError_t wide_dispatch(Shape const &s, Launch &L, T const *const *mats, T const *const *in, T *const *out,
                         T *const *ws, hipStream_t stream, bool run)
{
    if (!wide_ok<T>(s)) return hipErrorInvalidValue;
    int chunk = 1;
    bool in_lds = false;
    wide_shape<T>(s, chunk, in_lds);
    if (in_lds) return wide_launch<T, true>(s, L, chunk, mats, in, out, ws, stream, run);
    if (run && ws == nullptr) return hipErrorInvalidValue;
    return wide_launch<T, false>(s, L, chunk, mats, in, out, ws, stream, run);
}

template<typename T>
hipError_t fiber_dispatch(Shape const &s, Launch &L, T const *const *mats, T const *const *in, T *const *out,
                          T *const *ws, hipStream_t stream, bool run)
{
    if (s.n > kMaxFiber) return wide_dispatch<T>(s, L, mats, in, out, ws, stream, run);
    if (fiber_in_lds<T>(s)) return fiber_dispatch_n<T, true>(s, L, mats, in, out, ws, stream, run);
    if (run && ws == nullptr) return hipErrorInvalidValue;
    if (use_passes<T>(s)) return pass_dispatch<T>(s, L, mats, in, out, ws, stream, run);
    return fiber_dispatch_n<T, false>(s, L, mats, in, out, ws, stream, run);
}

// ---- persistent (packed) launch helpers -------------------------------------------------------------------------------------
constexpr size_t kGroupLdsBudget = 64 * 1024;

// LDS bytes one entry of a persistent pass takes (padded vector and factors, kron_group.hpp)
template<typename T>
long group_entry_bytes(Shape const &s)
{
    if (s.n > 8) return LONG_MAX;
    return (static_cast<long>(krongroup::padded_size(s.n, static_cast<int>(s.size))) +
            krongroup::coef_pitch(s.d, s.n) + krongroup::ptr_elems<T>(s.d)) * static_cast<long>(sizeof(T));
}

template<typename T, typename GroupT>
bool group_persistent_ok(GroupT const &g, Shape const &s)
{
    return s.n <= 8 && group_entry_bytes<T>(s) <= static_cast<long>(kGroupLdsBudget);
}

// One group's descriptor; returns the LDS bytes its passes take.
template<typename T>
size_t fill_group(Shape const &s, T const *const *mats, T const *const *in, T *const *out, int entries_per_tile,
                  krongroup::GroupDesc<T> &D, bool wave = false)
{
    D.wave = wave && kronwave::wave_shape(s.d, s.n) ? 1 : 0;
    D.d = s.d;
    D.n = s.n;
    D.stride = s.stride;
    D.nb = s.nb;
    D.size = static_cast<int>(s.size);
    D.nfib = static_cast<int>(s.size / s.n);
    long const per_entry = group_entry_bytes<T>(s);
    int E = krongroup::kThreads / D.nfib;
    if (E < 1) E = 1;
    while (E > 1 && E * per_entry > static_cast<long>(kGroupLdsBudget)) --E;
    // a tile shorter than one pass shrinks the pass (and the LDS the launch needs, so more workgroups fit)
    if (entries_per_tile > 0 && entries_per_tile < E) E = entries_per_tile;
    D.E = E;
    D.epw = entries_per_tile > 0 ? entries_per_tile : E;
    D.mats = mats;
    D.in = in;
    D.out = out;
    // a wave group's tile is cut among the workgroup's wavefronts, each with its own LDS rows
    if (D.wave) return krongroup::kWaves * krongroup::wave_lds_elems() * sizeof(T);
    return static_cast<size_t>(E) * per_entry + sizeof(void *);  // + pointer alignment
}

// The packed fiber sweep for one batch of small vectors: the persistent kernel over a one-group table, so
// that a workgroup contracts several entries at once (one per n^(d-1) threads) instead of one.
constexpr int kPackedMaxFibers = 64;        // at least 4 entries per 256-thread pass
constexpr long kPackedLds = 40 * 1024;      // pass size for 4 resident workgroups per CU
constexpr int kPackedWorkgroupsPerCu = 8;
// Default tile: one pass, or, where that is at most half a pass, the tile that gives every CU
// kPackedTilesPerCu tiles. The kernel waits on its staging loads, so the workgroups that share a CU hide
// each other's latency; measured on MI355X (FP64): 20000 entries of d=2 n=8 26.8 -> 15.1 us, of d=2 n=6
// 11.3 -> 9.5 us, of d=2 n=4 8.8 -> 7.2 us, 8192 of d=3 n=4 8.1 -> 7.4 us, 6000 of d=3 n=6 12.6 -> 11.1 us;
// a tile above half a pass was slower than the full pass (20000 entries of d=5 n=2: 9.9 against 8.8 us).
constexpr long kPackedTilesPerCu = 8;
// The packed launch is the default from 20 entries per CU on: measured on MI355X, one workgroup per entry
// is faster for 4096 entries of d=3 n=4 (6.2 vs 7.9 us), the packed launch for 6000 of d=3 n=6 (12.4 vs
// 15.1 us) and 20000 of d=2 n=8 (26.5 vs 51 us).
constexpr int kPackedEntriesPerCu = 20;

template<typename T>
bool packed_ok(Shape const &s)
{
    return s.n <= 8 && s.size / s.n <= kPackedMaxFibers && group_entry_bytes<T>(s) <= kPackedLds;
}

template<typename T>
hipError_t packed_launch(Shape const &s, kronmult_tuning_t const &tu, Launch &L, T const *const *mats,
                         T const *const *in, T *const *out, hipStream_t stream, bool run)
{
    krongroup::GroupTable<T> table{};
    long const per_entry = group_entry_bytes<T>(s);
    int ept = tu.entries_per_workgroup;
    if (ept <= 0)
    {
        ept = krongroup::kThreads / static_cast<int>(s.size / s.n);
        while (ept > 1 && ept * per_entry > kPackedLds) --ept;
        long const tiles = kPackedTilesPerCu * device_cu_count();
        long const short_tile = (s.nb + tiles - 1) / tiles;
        if (2 * short_tile <= ept) ept = static_cast<int>(short_tile < 1 ? 1 : short_tile);
    }
    size_t const lds = fill_group<T>(s, mats, in, out, ept, table.g[0]);
    table.count = 1;
    table.tile_start[0] = 0;
    table.tiles = (s.nb + table.g[0].epw - 1) / table.g[0].epw;
    table.tile_start[1] = table.tiles;
    int const per_cu = tu.workgroups_per_cu > 0 ? tu.workgroups_per_cu : kPackedWorkgroupsPerCu;
    long grid = static_cast<long>(device_cu_count()) * per_cu;
    if (grid > table.tiles) grid = table.tiles;
    L.family = KRONMULT_FAMILY_PACKED;
    L.grid = grid;
    L.block = krongroup::kThreads;
    L.epw = table.g[0].epw;
    L.lds = lds;
    L.resident = resident_per_cu(krongroup::kron_group_kernel<T, false>, krongroup::kThreads, lds);
    if (!run || grid == 0) return hipSuccess;
    hipLaunchKernelGGL((krongroup::kron_group_kernel<T, false>), dim3(static_cast<unsigned>(grid)), dim3(krongroup::kThreads),
                       lds, stream, table);
    return hipGetLastError();
}

// ---- one wavefront per entry (n^d <= 64, 2 <= n <= 8; kron_wave.hpp) --------------------------------------
// Default run length: enough runs for 8 workgroups per CU, at most 4 entries. Measured on MI355X (graph
// replay, FP64): 20000 entries of d=2 n=8 7.9 us (packed launch 15.0), d=6 n=2 7.5 (14.0), d=2 n=4 5.8 (7.2);
// 8192 of d=3 n=4 4.2 (packed 7.3), 4096 3.5 (fiber 6.3); 2000 of d=2 n=8 3.3 (fiber 7.1), 800 of d=6 n=2
// 2.7 (fiber 3.5). Longer runs gave fewer workgroups than the latency needs, shorter ones more atomics.
constexpr long kWaveRunsPerCu = 8;
constexpr int kWaveMaxRun = 4;
// 2 to 4 elements per lane: 20000 entries of d=7 n=2 13.6 us (packed launch 27.1), d=8 n=2 27.5 (fiber 44.0),
// d=5 n=3 24.7 (fiber 47.3), d=4 n=3 10.7 (packed 13.5), d=3 n=5 14.2 (packed 17.4), d=3 n=6 29.3 (packed
// 29.7) at runs of 2; runs of 4 were 1-7% slower, runs of 1 10-50% slower.
constexpr int kWaveMaxRunWide = 2;

template<typename T, int D, int N>
hipError_t wave_launch_dn(Shape const &s, kronmult_tuning_t const &tu, Launch &L, T const *const *mats,
                          T const *const *in, T *const *out, hipStream_t stream, bool run)
{
    auto kernel = kronwave::kron_wave_kernel<T, D, N>;
    L.family = KRONMULT_FAMILY_WAVE;
    L.block = kronwave::kLanes;
    L.lds = sizeof(T) * kronwave::wave_lds<D, N>();
    L.resident = resident_per_cu(kernel, kronwave::kLanes, 0);
    long const cus = device_cu_count();
    long epw = tu.entries_per_workgroup;
    if (epw <= 0 && tu.workgroups_per_cu > 0)
        epw = (s.nb + cus * tu.workgroups_per_cu - 1) / (cus * tu.workgroups_per_cu);
    if (epw <= 0)
    {
        epw = (s.nb + cus * kWaveRunsPerCu - 1) / (cus * kWaveRunsPerCu);
        int const max_run = kronwave::wave_size<D, N>() <= kronwave::kLanes ? kWaveMaxRun : kWaveMaxRunWide;
        if (epw > max_run) epw = max_run;
    }
    if (epw < 1) epw = 1;
    L.epw = static_cast<int>(epw);
    long const runs = (s.nb + epw - 1) / epw;
    L.grid = (runs + 7) / 8 * 8;  // whole XCD stripes; surplus runs exit at once
    if (L.grid > INT_MAX) return hipErrorInvalidValue;
    if (!run || s.nb == 0) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(L.grid)), dim3(kronwave::kLanes), 0, stream, s.nb, L.epw,
                       mats, s.stride, in, out);
    return hipGetLastError();
}

template<typename T>
hipError_t wave_launch(Shape const &s, kronmult_tuning_t const &tu, Launch &L, T const *const *mats,
                       T const *const *in, T *const *out, hipStream_t stream, bool run)
{
#define KRON_WAVE_LAUNCH(D, N) \
    if (s.d == D && s.n == N) return wave_launch_dn<T, D, N>(s, tu, L, mats, in, out, stream, run);
    KRON_WAVE_SHAPES(KRON_WAVE_LAUNCH)
    KRON_WAVE_SHAPES_WIDE(KRON_WAVE_LAUNCH)
#undef KRON_WAVE_LAUNCH
    return hipErrorInvalidValue;
}

// ---- entry --------------------------------------------------------------------------------------------
template<typename T>
int choose_family(Shape const &s, kronmult_tuning_t const &tu)
{
    bool const reg_ok = reg_rho(s.d, s.n) > 0;
    bool const mfma_ok = sizeof(T) == 8 && s.d == 6 && s.n == 4;
    if (tu.family == KRONMULT_FAMILY_REGISTER) return reg_ok ? KRONMULT_FAMILY_REGISTER : -1;
    bool const fiber_ok = s.n <= kMaxFiber || wide_ok<T>(s);
    if (tu.family == KRONMULT_FAMILY_FIBER) return fiber_ok ? KRONMULT_FAMILY_FIBER : -1;
    if (tu.family == KRONMULT_FAMILY_MATRIX) return mfma_ok ? KRONMULT_FAMILY_MATRIX : -1;
    if (tu.family == KRONMULT_FAMILY_PACKED) return packed_ok<T>(s) ? KRONMULT_FAMILY_PACKED : -1;
    if (tu.family == KRONMULT_FAMILY_WAVE) return kronwave::wave_family_shape(s.d, s.n) ? KRONMULT_FAMILY_WAVE : -1;
    if (tu.family != KRONMULT_FAMILY_AUTO) return -1;
    if (reg_ok) return KRONMULT_FAMILY_REGISTER;
    // d=3 n=6 FP64 (4 elements per lane, 152 VGPRs) measured 3-4% behind the packed launch
    bool const wave_default = kronwave::wave_family_shape(s.d, s.n) && !(sizeof(T) == 8 && s.d == 3 && s.n == 6);
    if (wave_default) return KRONMULT_FAMILY_WAVE;
    if (packed_ok<T>(s) && s.nb >= kPackedEntriesPerCu * device_cu_count()) return KRONMULT_FAMILY_PACKED;
    return fiber_ok ? KRONMULT_FAMILY_FIBER : -1;
}

// The reduction of an ordered call (kron_ordered.hpp): ranges of L.epw entries (the pieces' length), every
// segment reduced when the compute kernels left all contributions in the workspace (multi-pass).
template<typename T>
hipError_t reduce_pieces(Shape const &s, Launch const &L, T *const *out, T *const *ws, hipStream_t stream)
{
    using namespace kronordered;
    bool const all = L.workspace && use_passes<T>(s);
    int const epw = all ? 1 : L.epw;
    long const ranges = (s.nb + epw - 1) / epw;
    long const chunks = (s.size + kReduceThreads * kReducePer - 1) / (kReduceThreads * kReducePer);
    long const gy = chunks < 65535 ? chunks : 65535;
    // about 1024 workgroups: ranges per workgroup from 1 to 256 (one per thread)
    long rpw = (ranges * gy + 1023) / 1024;
    rpw = rpw < 1 ? 1 : rpw > kReduceThreads ? kReduceThreads : rpw;
    long const gx = (ranges + rpw - 1) / rpw;
    hipLaunchKernelGGL(kron_reduce_pieces_kernel<T>, dim3(static_cast<unsigned>(gx), static_cast<unsigned>(gy)),
                       dim3(kReduceThreads), 0, stream, s.nb, epw, static_cast<int>(s.size), all ? 1 : 0,
                       static_cast<int>(rpw), out, reinterpret_cast<T const *const *>(ws));
    return hipGetLastError();
}

template<typename T>
hipError_t kronmult_launch(int d, int n, T const *const *mats, int stride, T *const *in, T *const *out,
                           T *const *ws, int nb, kronmult_tuning_t const *tuning, hipStream_t stream, Launch *info,
                           bool ordered = false)
{
    Shape s{};
    hipError_t err = validate(d, n, stride, nb, s);
    if (err != hipSuccess) return err;
    s.ordered = ordered;
    kronmult_tuning_t const tu = tuning ? *tuning : kronmult_tuning_t{0, 0, 0};
    if (tu.workgroups_per_cu < 0 || tu.entries_per_workgroup < 0) return hipErrorInvalidValue;
    int family = choose_family<T>(s, tu);
    // the ordered epilogue exists for the register and one-entry-per-workgroup kernels
    if (ordered && (family == KRONMULT_FAMILY_PACKED || family == KRONMULT_FAMILY_WAVE) &&
        tu.family == KRONMULT_FAMILY_AUTO)
        family = KRONMULT_FAMILY_FIBER;
    if (family < 0 || (ordered && (family == KRONMULT_FAMILY_MATRIX || family == KRONMULT_FAMILY_PACKED ||
                                   family == KRONMULT_FAMILY_WAVE)))
        return hipErrorInvalidValue;
    bool const run = info == nullptr;
    if (run)
    {
        if (nb == 0) return hipSuccess;
        if (mats == nullptr || in == nullptr || out == nullptr) return hipErrorInvalidValue;
        if (ordered && ws == nullptr) return hipErrorInvalidValue;
    }
    Launch local;
    Launch &L = info ? *info : local;
    T const *const *cin = reinterpret_cast<T const *const *>(in);
    if (family == KRONMULT_FAMILY_REGISTER)
        err = reg_dispatch<T>(s, tu, L, mats, cin, out, ws, stream, run);
    else if (family == KRONMULT_FAMILY_PACKED)
        err = packed_launch<T>(s, tu, L, mats, cin, out, stream, run);
    else if (family == KRONMULT_FAMILY_WAVE)
        err = wave_launch<T>(s, tu, L, mats, cin, out, stream, run);
    else if constexpr (sizeof(T) == 8)
        err = family == KRONMULT_FAMILY_MATRIX ? mfma_launch(s, tu, L, mats, cin, out, stream, run)
                                               : fiber_dispatch<T>(s, L, mats, cin, out, ws, stream, run);
    else
        err = fiber_dispatch<T>(s, L, mats, cin, out, ws, stream, run);
    if (err != hipSuccess || !ordered) return err;
    return run ? reduce_pieces<T>(s, L, out, ws, stream) : hipSuccess;
}

// ---- several Kronecker terms per entry -----------------------------------------------------------------
template<typename T, int D, int N>
hipError_t reg_terms_launch(Shape const &s, int nterm, T const *const *mats, T const *const *in, T *const *out,
                            hipStream_t stream)
{
    using PL = kronplan::Plan<D, N, reg_rho(D, N)>;
    auto kernel = kronreg::kron_reg_terms_kernel<T, PL>;
    long const slots = static_cast<long>(device_cu_count()) * default_workgroups_per_cu(PL::P);
    int epw = static_cast<int>((s.nb + slots - 1) / slots);
    if (epw < 1) epw = 1;
    long const runs = (s.nb + epw - 1) / epw;
    long const grid = (runs + 7) / 8 * 8;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(grid)), dim3(PL::P), 0, stream, s.nb, nterm, epw, mats,
                       s.stride, in, out);
    return hipGetLastError();
}

// The fused kernel where a register-resident plan exists, else one launch of the one-term path per term
// (term t's factor pointers start at mats + t*nb*d).
template<typename T>
hipError_t terms_launch(int d, int n, int nterm, T const *const *mats, int stride, T *const *in, T *const *out,
                        T *const *ws, int nb, hipStream_t stream)
{
    Shape s{};
    hipError_t err = validate(d, n, stride, nb, s);
    if (err != hipSuccess) return err;
    if (nterm < 1) return hipErrorInvalidValue;
    if (nb == 0) return hipSuccess;
    if (mats == nullptr || in == nullptr || out == nullptr) return hipErrorInvalidValue;
    T const *const *cin = reinterpret_cast<T const *const *>(in);
    if (s.d == 6 && s.n == 4) return reg_terms_launch<T, 6, 4>(s, nterm, mats, cin, out, stream);
    if (s.d == 5 && s.n == 4) return reg_terms_launch<T, 5, 4>(s, nterm, mats, cin, out, stream);
    if (s.d == 4 && s.n == 4) return reg_terms_launch<T, 4, 4>(s, nterm, mats, cin, out, stream);
    if (s.d == 4 && s.n == 8) return reg_terms_launch<T, 4, 8>(s, nterm, mats, cin, out, stream);
    if (s.d == 3 && s.n == 8) return reg_terms_launch<T, 3, 8>(s, nterm, mats, cin, out, stream);
    for (int t = 0; t < nterm; ++t)
    {
        err = kronmult_launch<T>(d, n, mats + static_cast<size_t>(t) * nb * d, stride, in, out, ws, nb, nullptr,
                                 stream, nullptr);
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

template<typename T>
int kronmult_sync(int d, int n, T const *const *mats, int stride, T **in, T **out, T **ws, int nb)
{
    Shape s{};
    hipError_t err = validate(d, n, stride, nb, s);
    if (err != hipSuccess || nb == 0) return static_cast<int>(err);
    err = kronmult_launch<T>(d, n, mats, stride, in, out, ws, nb, nullptr, nullptr, nullptr);
    if (err != hipSuccess) return static_cast<int>(err);
    return static_cast<int>(hipDeviceSynchronize());
}

template<typename T>
char const *kernel_name(int d, int n, int nb)
{
    Shape s{};
    if (validate(d, n, n, nb < 0 ? 0 : nb, s) != hipSuccess) return "invalid";
    bool const f64 = sizeof(T) == 8;
    switch (choose_family<T>(s, kronmult_tuning_t{0, 0, 0}))
    {
    case KRONMULT_FAMILY_MATRIX: return "mfma_sweep_f64";
    case KRONMULT_FAMILY_PACKED: return f64 ? "packed_f64" : "packed_f32";
    case KRONMULT_FAMILY_WAVE: return f64 ? "wave_f64" : "wave_f32";
    case KRONMULT_FAMILY_REGISTER:
        if (m8_plan<T>(s.d, s.n)) return "reg_sweep_mfma_f64";
        return f64 ? "reg_sweep_f64" : "reg_sweep_f32";
    case KRONMULT_FAMILY_FIBER:
        if (s.n > kMaxFiber)
        {
            int chunk = 1;
            bool in_lds = false;
            wide_shape<T>(s, chunk, in_lds);
            if (in_lds) return f64 ? "wide_lds_f64" : "wide_lds_f32";
            return f64 ? "wide_global_f64" : "wide_global_f32";
        }
        if (fiber_in_lds<T>(s)) return f64 ? "fiber_lds_f64" : "fiber_lds_f32";
        if (use_passes<T>(s)) return f64 ? "multipass_f64" : "multipass_f32";
        return f64 ? "fiber_global_f64" : "fiber_global_f32";
    default: return "unsupported";
    }
}

// ---- adjoint ------------------------------------------------------------------------------------------
// Largest dynamic LDS one workgroup may take on the current device (gfx950: 160 KiB).
size_t device_max_lds()
{
    static std::atomic<long> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 64 * 1024;
    long v = cache[dev].load(std::memory_order_relaxed);
    if (v == 0)
    {
        int a = 0;
        if (hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || a <= 0)
            a = 64 * 1024;
        v = a;
        cache[dev].store(v, std::memory_order_relaxed);
    }
    return static_cast<size_t>(v);
}

constexpr size_t kAdjointAutoStoreLds = 64 * 1024;

struct AdjointPlan
{
    bool in_lds = false, store = false;
    int threads = 64;
    size_t lds = 0;
    long ws_elems = 0;  // per entry, 0 when the workspace is not read
};

// intermediates: KRONMULT_ADJOINT_AUTO / _RECOMPUTE / _STORE
hipError_t adjoint_plan(Shape const &s, int elem, int intermediates, size_t max_lds, AdjointPlan &p)
{
    if (s.n > kronadj::kMaxN) return hipErrorInvalidValue;
    if (intermediates < KRONMULT_ADJOINT_AUTO || intermediates > KRONMULT_ADJOINT_STORE) return hipErrorInvalidValue;
    long const nfib = s.size / s.n;
    p.threads = static_cast<int>(((nfib + 63) / 64) * 64);
    if (p.threads > kronadj::kThreads) p.threads = kronadj::kThreads;
    size_t const red = static_cast<size_t>(p.threads / 64) * 256 * elem;  // per-wavefront 16 x 16 tiles
    auto fits = [&](long vectors) { return static_cast<size_t>(vectors * s.size * elem) + red <= max_lds; };
    long const store_vectors = s.d;  // H and X_0..X_{d-2}
    bool store = intermediates == KRONMULT_ADJOINT_STORE;
    // AUTO keeps the chain only when it leaves room for several workgroups per CU: measured on the d=6 n=4
    // batch in FP32, the chain in 96 KiB of LDS (one workgroup per CU) took 2.2 ms against 1.0 ms recomputing
    if (intermediates == KRONMULT_ADJOINT_AUTO)
        store = fits(store_vectors) &&
                static_cast<size_t>(store_vectors * s.size * elem) + red <= kAdjointAutoStoreLds;
    long const vectors = store ? store_vectors : 2;
    p.store = store;
    p.in_lds = fits(vectors);
    p.lds = (p.in_lds ? static_cast<size_t>(vectors * s.size * elem) : 0) + red;
    p.ws_elems = p.in_lds ? 0 : vectors * s.size;
    return hipSuccess;
}

// Register-resident adjoint (kron_reg_adjoint.hpp): one-mode-per-phase plans (for d = 5, 6 with n = 4 too:
// 256 / 1024 threads of four values; measured on the d=6 n=4 batch: FP32 0.50 ms against 1.03 ms for the
// generic kernel, FP64 0.80 ms either way). STORE keeps d-1 extra vectors in LDS and is offered where that
// fits comfortably; RECOMPUTE everywhere.
template<typename T>
bool reg_adjoint_store_ok(Shape const &s)
{
    return (s.d == 3 && s.n == 8) || (s.d == 4 && s.n == 4);
}

template<typename T>
bool reg_adjoint_ok(Shape const &s, int /*intermediates*/)
{
    return (s.d == 3 && s.n == 8) || (s.d == 4 && s.n == 4) || (s.d == 4 && s.n == 8) || (s.d == 5 && s.n == 4) ||
           (s.d == 6 && s.n == 4);
}

// How the register kernel obtains X_t: rebuilt from x for every t, the chain kept fiber-major in LDS, or the
// chain kept in registers (each X_t in the layout of the phase that holds mode t: (d-1) R more values).
enum class RegAdjMode
{
    Recompute,
    StoreLds,
    KeepRegisters
};

template<typename T, int D, int N>
hipError_t reg_adjoint_launch(Shape const &s, RegAdjMode mode, T const *const *mats, T const *const *in,
                              T const *const *gout, T *const *gin, T *const *gmats, hipStream_t stream)
{
    using PL = kronplan::Plan<D, N, 1>;
    // short runs (about one entry per workgroup at the d=3 n=8 batch size): the adjoint's per-entry chain is
    // long, so balance beats prefetch depth (measured: 46 us vs 59 us with the forward's run length). The
    // d=6 n=4 plan (1024 threads, one workgroup per CU) takes one run per CU instead: its run sums of the
    // input and factor gradients then cover 32 entries of the 8192-entry batch (FP64 569 -> 521 us, input-
    // major 562 -> 487 us; FP32 347 -> 304 us; two or four runs per CU: 524 / 533 us)
    int const per_cu = PL::P >= 1024 ? 1 : 4 * default_workgroups_per_cu(PL::P);
    long const slots = static_cast<long>(device_cu_count()) * per_cu;
    int epw = static_cast<int>((s.nb + slots - 1) / slots);
    if (epw < 1) epw = 1;
    long const runs = (s.nb + epw - 1) / epw;
    long const grid = (runs + 7) / 8 * 8;
    dim3 const g(static_cast<unsigned>(grid)), blk(PL::P);
    switch (mode)
    {
    case RegAdjMode::StoreLds:
        if constexpr ((D == 3 && N == 8) || (D == 4 && N == 4))
        {
            hipLaunchKernelGGL((kronregadj::kron_reg_adjoint_kernel<T, PL, true, false>), g, blk, 0, stream, s.nb,
                               epw, mats, s.stride, in, gout, gin, gmats);
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    case RegAdjMode::KeepRegisters:
        hipLaunchKernelGGL((kronregadj::kron_reg_adjoint_kernel<T, PL, false, true>), g, blk, 0, stream, s.nb, epw,
                           mats, s.stride, in, gout, gin, gmats);
        return hipGetLastError();
    default:
        hipLaunchKernelGGL((kronregadj::kron_reg_adjoint_kernel<T, PL, false, false>), g, blk, 0, stream, s.nb,
                           epw, mats, s.stride, in, gout, gin, gmats);
        return hipGetLastError();
    }
}

// d = 6, n = 4: the two-modes-per-phase adjoint (kron_reg_adjoint2.hpp), one run per resident workgroup.
// It serves AUTO (measured on the d=6 batch: FP64 258-266 us, FP32 142-145 us, against 522 / 304 us for
// the one-mode-per-phase kernel, which RECOMPUTE and STORE keep).
template<typename T>
hipError_t adjoint2_launch(Shape const &s, T const *const *mats, T const *const *in, T const *const *gout,
                           T *const *gin, T *const *gmats, hipStream_t stream)
{
    auto kernel = kronregadj2::kron_reg_adjoint2_kernel<T>;
    int const per_cu = resident_per_cu(kernel, kronregadj2::kP, 0);
    long const slots = static_cast<long>(device_cu_count()) * per_cu;
    int epw = static_cast<int>((s.nb + slots - 1) / slots);
    if (epw < 1) epw = 1;
    long const runs = (s.nb + epw - 1) / epw;
    long const grid = (runs + 7) / 8 * 8;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(grid)), dim3(kronregadj2::kP), 0, stream, s.nb, epw, mats,
                       s.stride, in, gout, gin, gmats);
    return hipGetLastError();
}

template<typename T>
hipError_t reg_adjoint_dispatch(Shape const &s, int intermediates, T const *const *mats, T const *const *in,
                                T const *const *gout, T *const *gin, T *const *gmats, hipStream_t stream)
{
    if (s.d == 6 && s.n == 4 && intermediates == KRONMULT_ADJOINT_AUTO)
        return adjoint2_launch<T>(s, mats, in, gout, gin, gmats, stream);
    // STORE: the chain in LDS where it fits comfortably (d=3 n=8, d=4 n=4), else in registers.
    // AUTO, measured per shape: d=4 n=8 and d=5/6 n=4 keep the chain in registers (d=6 n=4 FP64 0.58 ms
    // against 0.78 ms recomputing, FP32 0.38 against 0.49; d=4 n=8 FP64 1.01 against 1.15 ms); d=3 n=8
    // recomputes (FP64 three waves per SIMD: 41.7 us against 46.4 us keeping the chain; FP32 24.6 us
    // against 24.9 us with the chain in LDS); d=4 n=4 recomputes in FP64, keeps the chain in LDS in FP32
    bool const lds_ok = reg_adjoint_store_ok<T>(s);
    bool const keep_preferred = (s.d == 4 && s.n == 8) || (s.d == 5 && s.n == 4) || (s.d == 6 && s.n == 4);
    RegAdjMode mode;
    if (intermediates == KRONMULT_ADJOINT_RECOMPUTE)
        mode = RegAdjMode::Recompute;
    else if (intermediates == KRONMULT_ADJOINT_STORE)
        mode = lds_ok ? RegAdjMode::StoreLds : RegAdjMode::KeepRegisters;
    else if (keep_preferred)
        mode = RegAdjMode::KeepRegisters;
    else
        mode = (sizeof(T) == 4 && lds_ok && !(s.d == 3 && s.n == 8)) ? RegAdjMode::StoreLds : RegAdjMode::Recompute;
    if (s.d == 3 && s.n == 8) return reg_adjoint_launch<T, 3, 8>(s, mode, mats, in, gout, gin, gmats, stream);
    if (s.d == 5 && s.n == 4) return reg_adjoint_launch<T, 5, 4>(s, mode, mats, in, gout, gin, gmats, stream);
    if (s.d == 6 && s.n == 4) return reg_adjoint_launch<T, 6, 4>(s, mode, mats, in, gout, gin, gmats, stream);
    if (s.d == 4 && s.n == 4) return reg_adjoint_launch<T, 4, 4>(s, mode, mats, in, gout, gin, gmats, stream);
    if (s.d == 4 && s.n == 8) return reg_adjoint_launch<T, 4, 8>(s, mode, mats, in, gout, gin, gmats, stream);
    return hipErrorInvalidValue;
}

template<typename T, int N, bool InLds, bool Store>
hipError_t adjoint_launch_n(Shape const &s, AdjointPlan const &p, T const *const *mats, T const *const *in,
                            T const *const *gout, T *const *gin, T *const *gmats, T *const *ws, hipStream_t stream)
{
    // runs of entries, one per resident workgroup slot: consecutive entries that share factors sum their
    // factor gradients in registers (the d=6 n=4 batch shares one factor block among all its entries)
    auto kernel = kronadj::kron_adjoint_kernel<T, N, InLds, Store>;
    long const slots = static_cast<long>(device_cu_count()) * resident_per_cu(kernel, p.threads, p.lds);
    int epw = static_cast<int>((s.nb + slots - 1) / slots);
    if (epw < 1) epw = 1;
    long const runs = (s.nb + epw - 1) / epw;
    long const grid = (runs + 7) / 8 * 8;  // surplus workgroups exit at once
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(grid)), dim3(p.threads), p.lds, stream, s.nb, epw, s.d,
                       static_cast<int>(s.size), mats, s.stride, in, gout, gin, gmats, ws);
    return hipGetLastError();
}

template<typename T, int N>
hipError_t adjoint_launch_mode(Shape const &s, AdjointPlan const &p, T const *const *mats, T const *const *in,
                               T const *const *gout, T *const *gin, T *const *gmats, T *const *ws, hipStream_t stream)
{
    if (p.in_lds)
        return p.store ? adjoint_launch_n<T, N, true, true>(s, p, mats, in, gout, gin, gmats, ws, stream)
                       : adjoint_launch_n<T, N, true, false>(s, p, mats, in, gout, gin, gmats, ws, stream);
    return p.store ? adjoint_launch_n<T, N, false, true>(s, p, mats, in, gout, gin, gmats, ws, stream)
                   : adjoint_launch_n<T, N, false, false>(s, p, mats, in, gout, gin, gmats, ws, stream);
}

template<typename T>
hipError_t kronmult_adjoint_launch(int d, int n, T const *const *mats, int stride, T const *const *in,
                                   T const *const *gout, T *const *gin, T *const *gmats, T *const *ws, int nb,
                                   int intermediates, hipStream_t stream)
{
    Shape s{};
    hipError_t err = validate(d, n, stride, nb, s);
    if (err != hipSuccess) return err;
    AdjointPlan p;
    err = adjoint_plan(s, sizeof(T), intermediates, device_max_lds(), p);
    if (err != hipSuccess) return err;
    if (nb == 0 || (gin == nullptr && gmats == nullptr)) return hipSuccess;
    if (mats == nullptr || gout == nullptr || (gmats != nullptr && in == nullptr)) return hipErrorInvalidValue;
    if (reg_adjoint_ok<T>(s, intermediates))
        return reg_adjoint_dispatch<T>(s, intermediates, mats, in, gout, gin, gmats, stream);
    if (p.ws_elems > 0 && ws == nullptr) return hipErrorInvalidValue;
    switch (n)
    {
    case 1: return adjoint_launch_mode<T, 1>(s, p, mats, in, gout, gin, gmats, ws, stream);
    case 2: return adjoint_launch_mode<T, 2>(s, p, mats, in, gout, gin, gmats, ws, stream);
    case 3: return adjoint_launch_mode<T, 3>(s, p, mats, in, gout, gin, gmats, ws, stream);
    case 4: return adjoint_launch_mode<T, 4>(s, p, mats, in, gout, gin, gmats, ws, stream);
    case 5: return adjoint_launch_mode<T, 5>(s, p, mats, in, gout, gin, gmats, ws, stream);
    case 6: return adjoint_launch_mode<T, 6>(s, p, mats, in, gout, gin, gmats, ws, stream);
    case 7: return adjoint_launch_mode<T, 7>(s, p, mats, in, gout, gin, gmats, ws, stream);
    case 8: return adjoint_launch_mode<T, 8>(s, p, mats, in, gout, gin, gmats, ws, stream);
    default: return hipErrorInvalidValue;
    }
}

template<typename T>
int kronmult_adjoint_sync(int d, int n, T const *const *mats, int stride, T const *const *in, T const *const *gout,
                          T **gin, T **gmats, T **ws, int nb, int intermediates)
{
    hipError_t err = kronmult_adjoint_launch<T>(d, n, mats, stride, in, gout, gin, gmats, ws, nb, intermediates, nullptr);
    if (err != hipSuccess || nb == 0) return static_cast<int>(err);
    return static_cast<int>(hipDeviceSynchronize());
}

template<typename T, typename GroupT>
hipError_t kronmult_grouped_launch(GroupT const *groups, int count, kronmult_grouped_options_t const *options,
                                   hipStream_t stream)
{
    if (count < 0 || (count > 0 && groups == nullptr)) return hipErrorInvalidValue;
    kronmult_grouped_options_t const opt = options ? *options : kronmult_grouped_options_t{0, 0};
    if (opt.strategy < KRONMULT_GROUPED_AUTO || opt.strategy > KRONMULT_GROUPED_SEPARATE || opt.workgroups_per_cu < 0)
        return hipErrorInvalidValue;
    // validate everything before launching anything
    Shape shapes[krongroup::kMaxGroups];
    bool persistent[krongroup::kMaxGroups];
    int n_persistent = 0;
    for (int i = 0; i < count; ++i)
    {
        GroupT const &g = groups[i];
        Shape s{};
        hipError_t err = validate(g.matrix_count, g.matrix_size, g.matrix_stride, g.nb_batch, s);
        if (err != hipSuccess) return err;
        if (g.entries_per_tile < 0) return hipErrorInvalidValue;
        if (g.nb_batch > 0 && (g.matrix_list_batched == nullptr || g.input_batched == nullptr ||
                               g.output_batched == nullptr))
            return hipErrorInvalidValue;
        // groups past the first kMaxGroups always get launches of their own
        bool const pers = opt.strategy != KRONMULT_GROUPED_SEPARATE && i < krongroup::kMaxGroups &&
                          group_persistent_ok<T>(g, s) && g.nb_batch > 0;
        if (i < krongroup::kMaxGroups)
        {
            shapes[i] = s;
            persistent[i] = pers;
        }
        if (!pers)
        {
            Launch L;
            kronmult_tuning_t const tu = g.tuning;
            err = kronmult_launch<T>(g.matrix_count, g.matrix_size, g.matrix_list_batched, g.matrix_stride,
                                     g.input_batched, g.output_batched, g.workspace_batched, 0, &tu, nullptr, &L);
            if (err != hipSuccess) return err;
            if (L.workspace && g.nb_batch > 0 && g.workspace_batched == nullptr) return hipErrorInvalidValue;
        }
        n_persistent += pers ? 1 : 0;
    }
    // separate launches first (same stream, so the order does not matter for the sums)
    for (int i = 0; i < count; ++i)
    {
        GroupT const &g = groups[i];
        if (i < krongroup::kMaxGroups && persistent[i]) continue;
        if (g.nb_batch == 0) continue;
        hipError_t err = kronmult_launch<T>(g.matrix_count, g.matrix_size, g.matrix_list_batched, g.matrix_stride,
                                            g.input_batched, g.output_batched, g.workspace_batched, g.nb_batch,
                                            &g.tuning, stream, nullptr);
        if (err != hipSuccess) return err;
    }
    if (n_persistent == 0) return hipSuccess;

    krongroup::GroupTable<T> table{};
    size_t lds = 0;
    int tiles = 0;
    bool any_wave = false;
    for (int i = 0; i < count && i < krongroup::kMaxGroups; ++i)
    {
        if (!persistent[i]) continue;
        GroupT const &g = groups[i];
        Shape const &s = shapes[i];
        krongroup::GroupDesc<T> &D = table.g[table.count];
        size_t const need = fill_group<T>(s, g.matrix_list_batched, const_cast<T const *const *>(g.input_batched),
                                          g.output_batched, g.entries_per_tile, D, true);
        table.tile_start[table.count] = tiles;
        tiles += (s.nb + D.epw - 1) / D.epw;
        if (need > lds) lds = need;
        any_wave = any_wave || D.wave != 0;
        ++table.count;
    }
    table.tile_start[table.count] = tiles;
    table.tiles = tiles;
    int const per_cu = opt.workgroups_per_cu > 0 ? opt.workgroups_per_cu : 2;
    long grid = static_cast<long>(device_cu_count()) * per_cu;
    if (grid > tiles) grid = tiles;
    auto kernel = any_wave ? krongroup::kron_group_kernel<T, true> : krongroup::kron_group_kernel<T, false>;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(grid)), dim3(krongroup::kThreads), lds, stream, table);
    return hipGetLastError();
}

// ---- FP64 through FP32 matrix cores ---------------------------------------------------------------------
template<int MODE, int N>
hipError_t split_launch(Shape const &s, double const *const *mats, double const *const *in, double *const *out,
                        hipStream_t stream)
{
    long const ntiles = (s.size / s.n + 15) / 16;
    int const waves = ntiles >= 4 ? 4 : static_cast<int>(ntiles);
    size_t const lds = static_cast<size_t>(s.size) * sizeof(double) +
                       static_cast<size_t>(kronsplit::SplitTraits<MODE>::kSlices) * 16 * N * sizeof(float);
    hipLaunchKernelGGL((kronsplit::kron_split_kernel<MODE, N>), dim3(s.nb), dim3(64 * waves), lds, stream, s.d,
                       static_cast<int>(s.size), mats, s.stride, in, out);
    return hipGetLastError();
}

template<int MODE>
hipError_t split_dispatch_n(Shape const &s, double const *const *mats, double const *const *in,
                            double *const *out, hipStream_t stream)
{
    switch (s.n)
    {
    case 1: return split_launch<MODE, 1>(s, mats, in, out, stream);
    case 2: return split_launch<MODE, 2>(s, mats, in, out, stream);
    case 3: return split_launch<MODE, 3>(s, mats, in, out, stream);
    case 4: return split_launch<MODE, 4>(s, mats, in, out, stream);
    case 5: return split_launch<MODE, 5>(s, mats, in, out, stream);
    case 6: return split_launch<MODE, 6>(s, mats, in, out, stream);
    case 7: return split_launch<MODE, 7>(s, mats, in, out, stream);
    case 8: return split_launch<MODE, 8>(s, mats, in, out, stream);
    default: return hipErrorInvalidValue;
    }
}

// FIXED3 for the n = 4 register plans: the register-resident split sweep (kron_split_reg.hpp), runs of
// entries per workgroup, one workgroup per resident slot.
template<int D>
hipError_t split_reg_launch(Shape const &s, double const *const *mats, double const *const *in,
                            double *const *out, hipStream_t stream)
{
    using PL = kronplan::Plan<D, 4, reg_rho(D, 4)>;
    auto kernel = kronsplitreg::kron_split_reg_kernel<PL>;
    int const per_cu = resident_per_cu(kernel, PL::P, 0);
    long const slots = static_cast<long>(device_cu_count()) * per_cu;
    int epw = static_cast<int>((s.nb + slots - 1) / slots);
    if (epw < 1) epw = 1;
    long const runs = (s.nb + epw - 1) / epw;
    long const grid = (runs + 7) / 8 * 8;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(grid)), dim3(PL::P), 0, stream, s.nb, epw, mats,
                       s.stride, in, out);
    return hipGetLastError();
}

hipError_t kronmult_split_launch(int d, int n, double const *const *mats, int stride, double *const *in,
                                 double *const *out, int nb, int mode, hipStream_t stream)
{
    Shape s{};
    hipError_t err = validate(d, n, stride, nb, s);
    if (err != hipSuccess) return err;
    if (n > kronsplit::kMaxN || mode < KRONMULT_SPLIT_FP32 || mode > KRONMULT_SPLIT_FIXED3) return hipErrorInvalidValue;
    if (s.size * static_cast<long>(sizeof(double)) + 3 * 16 * 8 * 4 > 64 * 1024) return hipErrorInvalidValue;
    if (nb == 0) return hipSuccess;
    if (mats == nullptr || in == nullptr || out == nullptr) return hipErrorInvalidValue;
    double const *const *cin = const_cast<double const *const *>(in);
    if (mode == KRONMULT_SPLIT_FIXED3 && n == 4)
    {
        if (d == 6) return split_reg_launch<6>(s, mats, cin, out, stream);
        if (d == 5) return split_reg_launch<5>(s, mats, cin, out, stream);
        if (d == 4) return split_reg_launch<4>(s, mats, cin, out, stream);
    }
    switch (mode)
    {
    case KRONMULT_SPLIT_FP32: return split_dispatch_n<kronsplit::FP32>(s, mats, cin, out, stream);
    case KRONMULT_SPLIT_PAIR32: return split_dispatch_n<kronsplit::PAIR32>(s, mats, cin, out, stream);
    case KRONMULT_SPLIT_BF16: return split_dispatch_n<kronsplit::BF16>(s, mats, cin, out, stream);
    case KRONMULT_SPLIT_PAIRBF16: return split_dispatch_n<kronsplit::PAIRBF16>(s, mats, cin, out, stream);
    default: return split_dispatch_n<kronsplit::FIXED3>(s, mats, cin, out, stream);
    }
}

// ---- batch builder: element-pair operator batches ----------------------------------------------------
template<typename T>
__global__ void __launch_bounds__(256)
    build_element_batch_kernel(int d, int n, int base, T const *const *__restrict__ coeff, int ld,
                               int const *__restrict__ row_ids, int const *__restrict__ col_ids, int ncols,
                               long nb, T *const *__restrict__ inputs, T *const *__restrict__ outputs,
                               T const **__restrict__ mats, T **__restrict__ in, T **__restrict__ out)
{
    long const k = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    int const a = static_cast<int>(k / ncols);
    int const b = static_cast<int>(k - static_cast<long>(a) * ncols);
    int ra = row_ids[a], cb = col_ids[b];
    for (int t = 0; t < d; ++t)
    {
        int const dr = ra % base, dc = cb % base;
        ra /= base;
        cb /= base;
        mats[k * d + t] = coeff[t] + static_cast<long>(n) * dr + static_cast<long>(n) * dc * ld;
    }
    in[k] = inputs[b];
    out[k] = outputs[a];
}

template<typename T>
hipError_t build_element_batch(int d, int n, int base, T const *const *coeff, int ld, int const *row_ids,
                               int nrows, int const *col_ids, int ncols, T *const *inputs, T *const *outputs,
                               T const **mats, T **in, T **out, hipStream_t stream)
{
    if (d < 1 || n < 1 || base < 1 || nrows < 0 || ncols < 0) return hipErrorInvalidValue;
    if (static_cast<long>(ld) < static_cast<long>(n) * base) return hipErrorInvalidValue;
    long const nb = static_cast<long>(nrows) * ncols;
    if (nb > INT_MAX) return hipErrorInvalidValue;
    if (nb == 0) return hipSuccess;
    if (!coeff || !row_ids || !col_ids || !inputs || !outputs || !mats || !in || !out) return hipErrorInvalidValue;
    hipLaunchKernelGGL(build_element_batch_kernel<T>, dim3(static_cast<unsigned>((nb + 255) / 256)), dim3(256), 0,
                       stream, d, n, base, coeff, ld, row_ids, col_ids, ncols, nb, inputs, outputs, mats, in, out);
    return hipGetLastError();
}

// Precondition of the ordered entry points, checked on the host from copies of the pointer arrays:
// entries with one output are consecutive, every output and workspace pointer is non-null, and the output
// vectors of distinct segments and all workspace vectors are pairwise disjoint.
template<typename T>
hipError_t check_ordered_batch(T *const *out, T *const *ws, int nb, long size)
{
    if (nb == 0) return hipSuccess;
    if (out == nullptr || ws == nullptr) return hipErrorInvalidValue;
    std::vector<T *> o(static_cast<size_t>(nb)), w(static_cast<size_t>(nb));
    hipError_t err = hipMemcpy(o.data(), out, sizeof(T *) * nb, hipMemcpyDefault);
    if (err == hipSuccess) err = hipMemcpy(w.data(), ws, sizeof(T *) * nb, hipMemcpyDefault);
    if (err != hipSuccess) return err;
    std::vector<uintptr_t> starts;
    starts.reserve(static_cast<size_t>(nb) * 2);
    for (int k = 0; k < nb; ++k)
    {
        if (o[k] == nullptr || w[k] == nullptr) return hipErrorInvalidValue;
        if (k == 0 || o[k] != o[k - 1]) starts.push_back(reinterpret_cast<uintptr_t>(o[k]));
        starts.push_back(reinterpret_cast<uintptr_t>(w[k]));
    }
    std::sort(starts.begin(), starts.end());
    uintptr_t const bytes = static_cast<uintptr_t>(size) * sizeof(T);
    for (size_t i = 0; i + 1 < starts.size(); ++i)
        if (starts[i] + bytes > starts[i + 1]) return hipErrorInvalidValue;
    return hipSuccess;
}

template<typename T>
int kronmult_ordered_sync(int d, int n, T const *const *mats, int stride, T **in, T **out, T **ws, int nb)
{
    Shape s{};
    hipError_t err = validate(d, n, stride, nb, s);
    if (err != hipSuccess || nb == 0) return static_cast<int>(err);
    err = check_ordered_batch<T>(out, ws, nb, s.size);
    if (err != hipSuccess) return static_cast<int>(err);
    err = kronmult_launch<T>(d, n, mats, stride, in, out, ws, nb, nullptr, nullptr, nullptr, true);
    if (err != hipSuccess) return static_cast<int>(err);
    return static_cast<int>(hipDeviceSynchronize());
}

// ---- ordered accumulation without a workspace or an allocation ---------------------------------------------
// The partial sums of an ordered call, and the per-entry workspace pointer array the kernels read, live in one
// library-owned device array per GPU. Layout of a call: nb pointers (rounded up to 256 bytes), then slots of
// n^d values. A range of epw entries owns two slots: r for the piece that starts at the range's first entry,
// ranges + r for the piece that starts inside the range and continues past its end (the only other piece of
// a range that is not a whole segment). One-entry ranges (epw == 1) take one slot each, slot = entry.
static __device__ __align__(256) unsigned char scratch[KRONMULT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on the current device, looked up once per device with that device current:
// each GPU holds its own copy of the array. Host code passes this address, never the array's name.
hipError_t scratch_address(void **ptr)
{
    static std::mutex lock;
    static void *addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(lock);
    if (addr[dev] == nullptr)
    {
        void *p = nullptr;
        err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return err;
        if (p == nullptr) return hipErrorInvalidSymbol;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    return hipSuccess;
}

template<typename T>
__global__ void __launch_bounds__(256)
    scratch_workspace_kernel(int nb, int epw, long ranges, long size, T **wsp, T *slots)
{
    long const k = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    long const r = k / epw;
    wsp[k] = slots + ((k % epw == 0) ? r : ranges + r) * size;
}

template<typename T>
hipError_t kronmult_ordered_noalloc(int d, int n, T const *const *mats, int stride, T **in, T **out, int nb,
                                    kronmult_tuning_t const *tuning, hipStream_t stream)
{
    Shape s{};
    hipError_t err = validate(d, n, stride, nb, s);
    if (err != hipSuccess) return err;
    Launch L;  // launch shape without launching: the range length of the summation order
    err = kronmult_launch<T>(d, n, mats, stride, in, out, nullptr, nb, tuning, nullptr, &L, true);
    if (err != hipSuccess) return err;
    if (nb == 0) return hipSuccess;
    if (mats == nullptr || in == nullptr || out == nullptr) return hipErrorInvalidValue;
    int const epw = (L.workspace && use_passes<T>(s)) ? 1 : L.epw;  // multi-pass: every entry is a piece
    long const ranges = (static_cast<long>(nb) + epw - 1) / epw;
    long const slots = epw == 1 ? nb : 2 * ranges;
    size_t const ptr_bytes = (static_cast<size_t>(nb) * sizeof(T *) + 255) / 256 * 256;
    size_t const slot_bytes = static_cast<size_t>(s.size) * sizeof(T);
    if (ptr_bytes > KRONMULT_NOALLOC_SCRATCH_BYTES ||
        static_cast<size_t>(slots) > (KRONMULT_NOALLOC_SCRATCH_BYTES - ptr_bytes) / slot_bytes)
        return hipErrorOutOfMemory;
    void *base = nullptr;
    err = scratch_address(&base);
    if (err != hipSuccess) return err;
    T **wsp = static_cast<T **>(base);
    T *slot0 = reinterpret_cast<T *>(static_cast<unsigned char *>(base) + ptr_bytes);
    hipLaunchKernelGGL(scratch_workspace_kernel<T>, dim3(static_cast<unsigned>((nb + 255) / 256)), dim3(256), 0,
                       stream, nb, epw, ranges, static_cast<long>(s.size), wsp, slot0);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    return kronmult_launch<T>(d, n, mats, stride, in, out, wsp, nb, tuning, stream, nullptr, true);
}

} // namespace

extern "C"
{

int kronmult_batched_f64(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                         int matrix_stride, double **input_batched, double **output_batched,
                         double **workspace_batched, int nb_batch)
{
    return kronmult_sync<double>(matrix_count, matrix_size, matrix_list_batched, matrix_stride, input_batched,
                                 output_batched, workspace_batched, nb_batch);
}

int kronmult_batched_f32(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                         int matrix_stride, float **input_batched, float **output_batched,
                         float **workspace_batched, int nb_batch)
{
    return kronmult_sync<float>(matrix_count, matrix_size, matrix_list_batched, matrix_stride, input_batched,
                                output_batched, workspace_batched, nb_batch);
}

int kronmult_batched_f64_async(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                               int matrix_stride, double **input_batched, double **output_batched,
                               double **workspace_batched, int nb_batch, void *stream)
{
    return static_cast<int>(kronmult_launch<double>(matrix_count, matrix_size, matrix_list_batched, matrix_stride,
                                                    input_batched, output_batched, workspace_batched, nb_batch,
                                                    nullptr, static_cast<hipStream_t>(stream), nullptr));
}

int kronmult_batched_f32_async(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                               int matrix_stride, float **input_batched, float **output_batched,
                               float **workspace_batched, int nb_batch, void *stream)
{
    return static_cast<int>(kronmult_launch<float>(matrix_count, matrix_size, matrix_list_batched, matrix_stride,
                                                   input_batched, output_batched, workspace_batched, nb_batch,
                                                   nullptr, static_cast<hipStream_t>(stream), nullptr));
}

int kronmult_batched_f64_tuned(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                               int matrix_stride, double **input_batched, double **output_batched,
                               double **workspace_batched, int nb_batch, kronmult_tuning_t const *tuning,
                               void *stream)
{
    return static_cast<int>(kronmult_launch<double>(matrix_count, matrix_size, matrix_list_batched, matrix_stride,
                                                    input_batched, output_batched, workspace_batched, nb_batch,
                                                    tuning, static_cast<hipStream_t>(stream), nullptr));
}

int kronmult_batched_f32_tuned(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                               int matrix_stride, float **input_batched, float **output_batched,
                               float **workspace_batched, int nb_batch, kronmult_tuning_t const *tuning,
                               void *stream)
{
    return static_cast<int>(kronmult_launch<float>(matrix_count, matrix_size, matrix_list_batched, matrix_stride,
                                                   input_batched, output_batched, workspace_batched, nb_batch,
                                                   tuning, static_cast<hipStream_t>(stream), nullptr));
}

int kronmult_batched_ordered_f64(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                                 int matrix_stride, double **input_batched, double **output_batched,
                                 double **workspace_batched, int nb_batch)
{
    return kronmult_ordered_sync<double>(matrix_count, matrix_size, matrix_list_batched, matrix_stride,
                                         input_batched, output_batched, workspace_batched, nb_batch);
}

int kronmult_batched_ordered_f32(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                                 int matrix_stride, float **input_batched, float **output_batched,
                                 float **workspace_batched, int nb_batch)
{
    return kronmult_ordered_sync<float>(matrix_count, matrix_size, matrix_list_batched, matrix_stride,
                                        input_batched, output_batched, workspace_batched, nb_batch);
}

int kronmult_batched_ordered_f64_async(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                                       int matrix_stride, double **input_batched, double **output_batched,
                                       double **workspace_batched, int nb_batch, kronmult_tuning_t const *tuning,
                                       void *stream)
{
    return static_cast<int>(kronmult_launch<double>(matrix_count, matrix_size, matrix_list_batched, matrix_stride,
                                                    input_batched, output_batched, workspace_batched, nb_batch,
                                                    tuning, static_cast<hipStream_t>(stream), nullptr, true));
}

int kronmult_batched_ordered_f32_async(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                                       int matrix_stride, float **input_batched, float **output_batched,
                                       float **workspace_batched, int nb_batch, kronmult_tuning_t const *tuning,
                                       void *stream)
{
    return static_cast<int>(kronmult_launch<float>(matrix_count, matrix_size, matrix_list_batched, matrix_stride,
                                                   input_batched, output_batched, workspace_batched, nb_batch,
                                                   tuning, static_cast<hipStream_t>(stream), nullptr, true));
}

int kronmult_batched_ordered_f64_noalloc(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                                         int matrix_stride, double **input_batched, double **output_batched,
                                         int nb_batch, kronmult_tuning_t const *tuning, void *stream)
{
    return static_cast<int>(kronmult_ordered_noalloc<double>(matrix_count, matrix_size, matrix_list_batched,
                                                             matrix_stride, input_batched, output_batched, nb_batch,
                                                             tuning, static_cast<hipStream_t>(stream)));
}

int kronmult_batched_terms_f64(int matrix_count, int matrix_size, int term_count,
                               double const *const *matrix_list_batched, int matrix_stride, double **input_batched,
                               double **output_batched, double **workspace_batched, int nb_batch)
{
    hipError_t err = terms_launch<double>(matrix_count, matrix_size, term_count, matrix_list_batched, matrix_stride,
                                          input_batched, output_batched, workspace_batched, nb_batch, nullptr);
    return static_cast<int>(err != hipSuccess ? err : hipDeviceSynchronize());
}

int kronmult_batched_terms_f32(int matrix_count, int matrix_size, int term_count,
                               float const *const *matrix_list_batched, int matrix_stride, float **input_batched,
                               float **output_batched, float **workspace_batched, int nb_batch)
{
    hipError_t err = terms_launch<float>(matrix_count, matrix_size, term_count, matrix_list_batched, matrix_stride,
                                         input_batched, output_batched, workspace_batched, nb_batch, nullptr);
    return static_cast<int>(err != hipSuccess ? err : hipDeviceSynchronize());
}

int kronmult_batched_terms_f64_async(int matrix_count, int matrix_size, int term_count,
                                     double const *const *matrix_list_batched, int matrix_stride,
                                     double **input_batched, double **output_batched, double **workspace_batched,
                                     int nb_batch, void *stream)
{
    return static_cast<int>(terms_launch<double>(matrix_count, matrix_size, term_count, matrix_list_batched,
                                                 matrix_stride, input_batched, output_batched, workspace_batched,
                                                 nb_batch, static_cast<hipStream_t>(stream)));
}

int kronmult_batched_terms_f32_async(int matrix_count, int matrix_size, int term_count,
                                     float const *const *matrix_list_batched, int matrix_stride,
                                     float **input_batched, float **output_batched, float **workspace_batched,
                                     int nb_batch, void *stream)
{
    return static_cast<int>(terms_launch<float>(matrix_count, matrix_size, term_count, matrix_list_batched,
                                                matrix_stride, input_batched, output_batched, workspace_batched,
                                                nb_batch, static_cast<hipStream_t>(stream)));
}

int kronmult_hip_ordered_launch_info(int matrix_count, int matrix_size, int element_size, int nb_batch,
                                     kronmult_tuning_t const *tuning, kronmult_launch_info_t *info)
{
    if (info == nullptr) return static_cast<int>(hipErrorInvalidValue);
    Launch L;
    hipError_t err;
    if (element_size == 8)
        err = kronmult_launch<double>(matrix_count, matrix_size, nullptr, matrix_size, nullptr, nullptr, nullptr,
                                      nb_batch, tuning, nullptr, &L, true);
    else if (element_size == 4)
        err = kronmult_launch<float>(matrix_count, matrix_size, nullptr, matrix_size, nullptr, nullptr, nullptr,
                                     nb_batch, tuning, nullptr, &L, true);
    else
        err = hipErrorInvalidValue;
    if (err != hipSuccess) return static_cast<int>(err);
    *info = kronmult_launch_info_t{L.family, static_cast<int>(L.grid), L.block, L.epw, static_cast<int>(L.lds),
                                   L.resident, 1};
    return static_cast<int>(hipSuccess);
}

int kronmult_hip_launch_info(int matrix_count, int matrix_size, int element_size, int nb_batch,
                             kronmult_tuning_t const *tuning, kronmult_launch_info_t *info)
{
    if (info == nullptr) return static_cast<int>(hipErrorInvalidValue);
    Launch L;
    hipError_t err;
    if (element_size == 8)
        err = kronmult_launch<double>(matrix_count, matrix_size, nullptr, matrix_size, nullptr, nullptr, nullptr,
                                      nb_batch, tuning, nullptr, &L);
    else if (element_size == 4)
        err = kronmult_launch<float>(matrix_count, matrix_size, nullptr, matrix_size, nullptr, nullptr, nullptr,
                                     nb_batch, tuning, nullptr, &L);
    else
        err = hipErrorInvalidValue;
    if (err != hipSuccess) return static_cast<int>(err);
    *info = kronmult_launch_info_t{L.family, static_cast<int>(L.grid), L.block, L.epw, static_cast<int>(L.lds),
                                   L.resident, L.workspace ? 1 : 0};
    return 0;
}

int kronmult_batched_adjoint_f64(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                                 int matrix_stride, double const *const *input_batched,
                                 double const *const *grad_output_batched, double **grad_input_batched,
                                 double **grad_matrix_batched, double **workspace_batched, int nb_batch,
                                 int intermediates)
{
    return kronmult_adjoint_sync<double>(matrix_count, matrix_size, matrix_list_batched, matrix_stride, input_batched,
                                         grad_output_batched, grad_input_batched, grad_matrix_batched,
                                         workspace_batched, nb_batch, intermediates);
}

int kronmult_batched_adjoint_f32(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                                 int matrix_stride, float const *const *input_batched,
                                 float const *const *grad_output_batched, float **grad_input_batched,
                                 float **grad_matrix_batched, float **workspace_batched, int nb_batch,
                                 int intermediates)
{
    return kronmult_adjoint_sync<float>(matrix_count, matrix_size, matrix_list_batched, matrix_stride, input_batched,
                                        grad_output_batched, grad_input_batched, grad_matrix_batched,
                                        workspace_batched, nb_batch, intermediates);
}

int kronmult_batched_adjoint_f64_async(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                                       int matrix_stride, double const *const *input_batched,
                                       double const *const *grad_output_batched, double **grad_input_batched,
                                       double **grad_matrix_batched, double **workspace_batched, int nb_batch,
                                       int intermediates, void *stream)
{
    return static_cast<int>(kronmult_adjoint_launch<double>(
        matrix_count, matrix_size, matrix_list_batched, matrix_stride, input_batched, grad_output_batched,
        grad_input_batched, grad_matrix_batched, workspace_batched, nb_batch, intermediates,
        static_cast<hipStream_t>(stream)));
}

int kronmult_batched_adjoint_f32_async(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                                       int matrix_stride, float const *const *input_batched,
                                       float const *const *grad_output_batched, float **grad_input_batched,
                                       float **grad_matrix_batched, float **workspace_batched, int nb_batch,
                                       int intermediates, void *stream)
{
    return static_cast<int>(kronmult_adjoint_launch<float>(
        matrix_count, matrix_size, matrix_list_batched, matrix_stride, input_batched, grad_output_batched,
        grad_input_batched, grad_matrix_batched, workspace_batched, nb_batch, intermediates,
        static_cast<hipStream_t>(stream)));
}

long long kronmult_batched_adjoint_workspace(int matrix_count, int matrix_size, int element_size, int intermediates)
{
    Shape s{};
    if (element_size != 4 && element_size != 8) return -1;
    if (validate(matrix_count, matrix_size, matrix_size, 0, s) != hipSuccess) return -1;
    AdjointPlan p;
    if (adjoint_plan(s, element_size, intermediates, device_max_lds(), p) != hipSuccess) return -1;
    bool const reg = element_size == 8 ? reg_adjoint_ok<double>(s, intermediates) : reg_adjoint_ok<float>(s, intermediates);
    return reg ? 0 : p.ws_elems;  // the register-resident kernel keeps its scratch on chip
}

int kronmult_batched_grouped_f64(kronmult_group_f64_t const *groups, int group_count,
                                 kronmult_grouped_options_t const *options, void *stream)
{
    return static_cast<int>(
        kronmult_grouped_launch<double>(groups, group_count, options, static_cast<hipStream_t>(stream)));
}

int kronmult_batched_grouped_f32(kronmult_group_f32_t const *groups, int group_count,
                                 kronmult_grouped_options_t const *options, void *stream)
{
    return static_cast<int>(
        kronmult_grouped_launch<float>(groups, group_count, options, static_cast<hipStream_t>(stream)));
}

int kronmult_batched_split_f64(int matrix_count, int matrix_size, double const *const *matrix_list_batched,
                               int matrix_stride, double **input_batched, double **output_batched, int nb_batch,
                               int split_mode, void *stream)
{
    return static_cast<int>(kronmult_split_launch(matrix_count, matrix_size, matrix_list_batched, matrix_stride,
                                                  input_batched, output_batched, nb_batch, split_mode,
                                                  static_cast<hipStream_t>(stream)));
}

int kronmult_build_element_batch_f64(int matrix_count, int matrix_size, int base, double const *const *coeff,
                                     int coeff_ld, int const *row_ids, int nrows, int const *col_ids, int ncols,
                                     double *const *inputs, double *const *outputs,
                                     double const **matrix_list_batched, double **input_batched,
                                     double **output_batched, void *stream)
{
    return static_cast<int>(build_element_batch<double>(matrix_count, matrix_size, base, coeff, coeff_ld, row_ids,
                                                        nrows, col_ids, ncols, inputs, outputs, matrix_list_batched,
                                                        input_batched, output_batched,
                                                        static_cast<hipStream_t>(stream)));
}

int kronmult_build_element_batch_f32(int matrix_count, int matrix_size, int base, float const *const *coeff,
                                     int coeff_ld, int const *row_ids, int nrows, int const *col_ids, int ncols,
                                     float *const *inputs, float *const *outputs,
                                     float const **matrix_list_batched, float **input_batched,
                                     float **output_batched, void *stream)
{
    return static_cast<int>(build_element_batch<float>(matrix_count, matrix_size, base, coeff, coeff_ld, row_ids,
                                                       nrows, col_ids, ncols, inputs, outputs, matrix_list_batched,
                                                       input_batched, output_batched,
                                                       static_cast<hipStream_t>(stream)));
}

char const *kronmult_hip_version(void)
{
    return "kronmult_hip 0.4 (gfx950)";
}

char const *kronmult_hip_kernel_name(int matrix_count, int matrix_size, int element_size, int nb_batch)
{
    if (element_size == 8) return kernel_name<double>(matrix_count, matrix_size, nb_batch);
    if (element_size == 4) return kernel_name<float>(matrix_count, matrix_size, nb_batch);
    return "invalid";
}

} // extern "C"
