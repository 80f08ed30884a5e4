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

#include <atomic>
#include <climits>
#include <cstdint>
#include <cstdio>

#include "kron_reg.hpp"
#include "kronmult_hip.h"

namespace
{

constexpr int kMaxFiber = 64;           // largest n handled (runtime-n kernel)
constexpr int kLdsBudget = 64 * 1024;   // bytes of LDS one workgroup may take for the n^d vector
constexpr int kFiberThreads = 256;      // workgroup size of the fiber sweep

// ------------------------------------------------------------------------------------------------------
// Fiber sweep.
//
// Mode t (0 = slowest) of the canonical n^d layout has stride st = n^(d-1-t). Contracting it maps the
// fiber {base + k*st, k < n} to {base + r*st, r < n} with y_r = sum_k A_t[r, k] * x_k. There are n^(d-1)
// fibers per step; fiber f has base = (f / st) * st * n + f % st.
//
// N > 0: compile-time fiber length, coefficients held in registers (uniform across the workgroup, so the
// compiler loads them with scalar loads). N == 0: runtime n <= kMaxFiber, coefficients staged in LDS.
// InLds: the vector lives in LDS; otherwise in ws[b] (the first step reads the input directly).
// ------------------------------------------------------------------------------------------------------
template<typename T, int N, bool InLds>
__global__ void __launch_bounds__(kFiberThreads)
    kron_fiber_kernel(int d, int n_rt, int size, T const *const *__restrict__ mats, int stride,
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
    for (int i = threadIdx.x; i < size; i += blockDim.x) atomicAdd(y + i, x[i]);
}

// ------------------------------------------------------------------------------------------------------
// Dispatch.
// ------------------------------------------------------------------------------------------------------
struct Shape
{
    int d, n, stride, nb;
    long size;
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
    s = Shape{d, n, stride, nb, size};
    return hipSuccess;
}

enum class Path
{
    Reg,
    FiberLds,
    FiberGlobal,
    Unsupported
};

// Register-resident plans (d, n) -> RHO. Threads per workgroup P = n^(d-RHO) must be a multiple of 64
// and at most 1024; n^d elements per entry stay in LDS between phases.
constexpr int reg_rho(int d, int n)
{
    return (d == 6 && n == 4) ? 2 : (d == 5 && n == 4) ? 2 : (d == 4 && n == 4) ? 1
         : (d == 4 && n == 8) ? 1 : (d == 3 && n == 8) ? 1 : 0;
}

template<typename T>
Path choose_path(Shape const &s)
{
    if (reg_rho(s.d, s.n) > 0) return Path::Reg;
    if (s.n > kMaxFiber) return Path::Unsupported;
    long const coef_bytes = (s.n <= 8) ? 0 : static_cast<long>(s.n) * s.n * sizeof(T);
    if (s.size * static_cast<long>(sizeof(T)) + coef_bytes <= kLdsBudget) return Path::FiberLds;
    return Path::FiberGlobal;
}

template<typename T, int N, bool InLds>
hipError_t launch_fiber(Shape const &s, T const *const *mats, T const *const *in, T *const *out, T *const *ws,
                        hipStream_t stream)
{
    long const nfib = s.size / s.n;
    int threads = static_cast<int>(((nfib + 63) / 64) * 64);
    if (threads > kFiberThreads) threads = kFiberThreads;
    size_t lds = 0;
    if (InLds) lds += static_cast<size_t>(s.size) * sizeof(T);
    if (N == 0) lds += static_cast<size_t>(s.n) * s.n * sizeof(T);
    hipLaunchKernelGGL((kron_fiber_kernel<T, N, InLds>), dim3(s.nb), dim3(threads), lds, stream, s.d, s.n,
                       static_cast<int>(s.size), mats, s.stride, in, out, ws);
    return hipGetLastError();
}

template<typename T, bool InLds>
hipError_t dispatch_fiber_n(Shape const &s, T const *const *mats, T const *const *in, T *const *out,
                            T *const *ws, hipStream_t stream)
{
    switch (s.n)
    {
    case 1: return launch_fiber<T, 1, InLds>(s, mats, in, out, ws, stream);
    case 2: return launch_fiber<T, 2, InLds>(s, mats, in, out, ws, stream);
    case 3: return launch_fiber<T, 3, InLds>(s, mats, in, out, ws, stream);
    case 4: return launch_fiber<T, 4, InLds>(s, mats, in, out, ws, stream);
    case 5: return launch_fiber<T, 5, InLds>(s, mats, in, out, ws, stream);
    case 6: return launch_fiber<T, 6, InLds>(s, mats, in, out, ws, stream);
    case 7: return launch_fiber<T, 7, InLds>(s, mats, in, out, ws, stream);
    case 8: return launch_fiber<T, 8, InLds>(s, mats, in, out, ws, stream);
    default: return launch_fiber<T, 0, InLds>(s, mats, in, out, ws, stream);
    }
}

// Number of compute units of the current device (cached per device; the value never changes).
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

// Resident workgroups per CU of a kernel (cached per kernel; the occupancy does not depend on the device
// beyond the architecture this library is built for).
template<typename Kernel>
int resident_per_cu(Kernel kernel, int block)
{
    static std::atomic<int> cache{0};
    int v = cache.load(std::memory_order_relaxed);
    if (v == 0)
    {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kernel, block, 0) != hipSuccess || v <= 0) v = 1;
        cache.store(v, std::memory_order_relaxed);
    }
    return v;
}

// The register-resident sweep is launched as exactly one resident round: grid = CUs x resident
// workgroups per CU (a multiple of 8, so every XCD gets the same number of runs), and each workgroup walks
// a run of ceil(nb / grid) consecutive entries.
template<typename T, int D, int N>
hipError_t launch_reg(Shape const &s, T const *const *mats, T const *const *in, T *const *out, hipStream_t stream)
{
    using PL = kronplan::Plan<D, N, reg_rho(D, N)>;
    static_assert(PL::P % 64 == 0 && PL::P <= 1024, "plan must fill whole wavefronts");
    auto kernel = kronreg::kron_reg_kernel<T, PL>;
    long const slots = static_cast<long>(device_cu_count()) * resident_per_cu(kernel, PL::P);
    int epw = static_cast<int>((s.nb + slots - 1) / slots);
    if (epw < 1) epw = 1;
    long runs = (s.nb + epw - 1) / epw;
    runs = (runs + 7) / 8 * 8;  // whole rounds over the 8 XCDs; surplus runs exit at once
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(runs)), dim3(PL::P), 0, stream, s.nb, epw, mats, s.stride,
                       in, out);
    return hipGetLastError();
}

template<typename T>
hipError_t dispatch_reg(Shape const &s, T const *const *mats, T const *const *in, T *const *out, hipStream_t stream)
{
    if (s.d == 6 && s.n == 4) return launch_reg<T, 6, 4>(s, mats, in, out, stream);
    if (s.d == 5 && s.n == 4) return launch_reg<T, 5, 4>(s, mats, in, out, stream);
    if (s.d == 4 && s.n == 4) return launch_reg<T, 4, 4>(s, mats, in, out, stream);
    if (s.d == 4 && s.n == 8) return launch_reg<T, 4, 8>(s, mats, in, out, stream);
    if (s.d == 3 && s.n == 8) return launch_reg<T, 3, 8>(s, mats, in, out, stream);
    return hipErrorInvalidValue;
}

template<typename T>
hipError_t kronmult_launch(int d, int n, T const *const *mats, int stride, T *const *in, T *const *out,
                           T *const *ws, int nb, hipStream_t stream)
{
    Shape s{};
    hipError_t err = validate(d, n, stride, nb, s);
    if (err != hipSuccess) return err;
    if (nb == 0) return hipSuccess;
    if (mats == nullptr || in == nullptr || out == nullptr) return hipErrorInvalidValue;
    T const *const *cin = reinterpret_cast<T const *const *>(in);
    switch (choose_path<T>(s))
    {
    case Path::Reg: return dispatch_reg<T>(s, mats, cin, out, stream);
    case Path::FiberLds: return dispatch_fiber_n<T, true>(s, mats, cin, out, ws, stream);
    case Path::FiberGlobal:
        if (ws == nullptr) return hipErrorInvalidValue;
        return dispatch_fiber_n<T, false>(s, mats, cin, out, ws, stream);
    default: return hipErrorInvalidValue;
    }
}

template<typename T>
int kronmult_sync(int d, int n, T const *const *mats, int stride, T **in, T **out, T **ws, int nb)
{
    Shape s{};
    hipError_t err = validate(d, n, stride, nb, s);
    if (err != hipSuccess || nb == 0) return static_cast<int>(err);
    err = kronmult_launch<T>(d, n, mats, stride, in, out, ws, nb, nullptr);
    if (err != hipSuccess) return static_cast<int>(err);
    return static_cast<int>(hipDeviceSynchronize());
}

template<typename T>
char const *kernel_name(int d, int n, int nb)
{
    Shape s{};
    if (validate(d, n, n, nb < 0 ? 0 : nb, s) != hipSuccess) return "invalid";
    bool const f64 = sizeof(T) == 8;
    switch (choose_path<T>(s))
    {
    case Path::Reg: return f64 ? "reg_sweep_f64" : "reg_sweep_f32";
    case Path::FiberLds: return f64 ? "fiber_lds_f64" : "fiber_lds_f32";
    case Path::FiberGlobal: return f64 ? "fiber_global_f64" : "fiber_global_f32";
    default: return "unsupported";
    }
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
    return static_cast<int>(kronmult_launch<double>(matrix_count, matrix_size, matrix_list_batched,
                                                    matrix_stride, input_batched, output_batched,
                                                    workspace_batched, nb_batch,
                                                    static_cast<hipStream_t>(stream)));
}

int kronmult_batched_f32_async(int matrix_count, int matrix_size, float const *const *matrix_list_batched,
                               int matrix_stride, float **input_batched, float **output_batched,
                               float **workspace_batched, int nb_batch, void *stream)
{
    return static_cast<int>(kronmult_launch<float>(matrix_count, matrix_size, matrix_list_batched,
                                                   matrix_stride, input_batched, output_batched,
                                                   workspace_batched, nb_batch,
                                                   static_cast<hipStream_t>(stream)));
}

char const *kronmult_hip_version(void)
{
    return "kronmult_hip 0.1 (gfx950)";
}

char const *kronmult_hip_kernel_name(int matrix_count, int matrix_size, int element_size, int nb_batch)
{
    if (element_size == 8) return kernel_name<double>(matrix_count, matrix_size, nb_batch);
    if (element_size == 4) return kernel_name<float>(matrix_count, matrix_size, nb_batch);
    return "invalid";
}

} // extern "C"
