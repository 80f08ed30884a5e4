// kron_wave.hpp -- one wavefront per entry for small tensors (n^d <= 64 with 2 <= n <= 8, and a few shapes
// of up to 256 elements; no counterpart in the reference, whose kernel gives every entry a workgroup of its
// own, kronmult_gpu/kronmult.cu:95-130).
//
// A workgroup is one wavefront and walks a run of consecutive entries. Lane l holds elements l + 64 j of
// the entry (canonical order, mode d-1 fastest); lanes past n^d idle. A mode step writes the lanes'
// values to the wavefront's LDS row, and lane l reads the n values of each of its fibers and row r of the
// factor (staged beside them, column-major) and forms the fused chain over k ascending from zero. The
// wavefront's LDS instructions complete in program order, so there are no workgroup barriers: many
// independent wavefronts per CU hide each other's memory latency. The next entry's element and factor
// values are loaded one entry ahead. Consecutive entries with the same output pointer are summed per lane
// and flushed with one atomic add per element and run. Modes d-1 .. 0, fused multiply-adds from zero:
// bit-identical per entry to the reference's sweep (kronmult_gpu/kronmult.cu:66-70, :104).
#pragma once

#include <hip/hip_runtime.h>

#include "kron_reg.hpp"

namespace kronwave
{

constexpr int kLanes = 64;
constexpr int kMaxD = 8;          // n >= 2 and n^d <= kMaxElems
constexpr int kMaxElems = 256;    // elements per entry: at most 4 per lane
constexpr int kCoefPerLane = 2;   // d n^2 <= 128 for every one-element-per-lane shape
constexpr int kMaxCoef = kCoefPerLane * kLanes;

// orders this wavefront's LDS accesses around it (a wavefront's LDS instructions complete in order; the
// fences keep the compiler from moving accesses across)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__host__ __device__ constexpr int wave_elems(int d, int n)
{
    int size = 1;
    for (int t = 0; t < d; ++t) size *= n;
    return size;
}

// the shapes the one-element-per-lane path takes (the grouped kernel's wave groups)
__host__ __device__ constexpr bool wave_shape(int d, int n)
{
    return n >= 2 && n <= 8 && d >= 1 && d <= kMaxD && wave_elems(d, n) <= kLanes && d * n * n <= kMaxCoef;
}

template<int D, int N>
constexpr int wave_size()
{
    return wave_elems(D, N);
}

// LDS elements of one wavefront's rows for (D, N): the tensor and the entry's factors, whole lanes each
template<int D, int N>
constexpr int wave_lds()
{
    return (wave_size<D, N>() + kLanes - 1) / kLanes * kLanes + (D * N * N + kLanes - 1) / kLanes * kLanes;
}

// Entries [e0, e1) by this wavefront (lane = its lane index, xs / cs = its LDS rows: R = ceil(n^d / 64)
// lanes' worth of tensor and ceil(d n^2 / 64) of factors). Lane l holds elements l + 64 j, j < R. Used by
// the single-batch kernel below and by the grouped kernel (kron_group.hpp, one element per lane).
template<typename T, int D, int N, bool OPAQUE_LANE = false>
__device__ __forceinline__ void wave_run(int e0, int e1, T const *const *__restrict__ mats, int stride,
                                         T const *const *__restrict__ in, T *const *__restrict__ out, T *xs,
                                         T *cs, int lane)
{
    constexpr int NN = N * N;
    constexpr int NCOEF = D * NN;
    constexpr int SIZE = wave_size<D, N>();
    constexpr int R = (SIZE + kLanes - 1) / kLanes;
    constexpr int CPL = (NCOEF + kLanes - 1) / kLanes;
    static_assert(N >= 2 && N <= 10 && SIZE <= kMaxElems, "2 <= n <= 10, n^d <= 256");
    // OPAQUE_LANE (the grouped kernel): inside its tile loop the compiler would otherwise hoist every
    // shape's per-lane offsets out of the loop and keep them all live (215 VGPRs)
    if constexpr (OPAQUE_LANE) asm volatile("" : "+v"(lane));

    T x[R];
    T f[CPL];
    auto load = [&](int e) {
#pragma unroll
        for (int j = 0; j < R; ++j)
        {
            int const idx = lane + j * kLanes;
            if (idx < SIZE) x[j] = kronreg::as_global(in[e])[idx];
        }
#pragma unroll
        for (int c = 0; c < CPL; ++c)
        {
            int const i = lane + c * kLanes;  // factor value i: A_m(r, k) at m n^2 + r + k n
            if (i < NCOEF)
            {
                int const m = i / NN;
                int const rk = i - m * NN;
                int const k = rk / N;
                int const r = rk - k * N;
                f[c] = kronreg::as_global(mats[static_cast<size_t>(e) * D + m])[r + k * stride];
            }
        }
    };
#pragma unroll
    for (int j = 0; j < R; ++j) x[j] = T(0);
    load(e0);

    T run[R];
#pragma unroll
    for (int j = 0; j < R; ++j) run[j] = T(0);
    T *cur = out[e0];
    for (int e = e0; e < e1; ++e)
    {
#pragma unroll
        for (int j = 0; j < R; ++j) xs[lane + j * kLanes] = x[j];
#pragma unroll
        for (int c = 0; c < CPL; ++c)
        {
            int const i = lane + c * kLanes;
            if (i < NCOEF) cs[i] = f[c];
        }
        if (e + 1 < e1) load(e + 1);
        wave_sync();
        T y[R];
        int st = 1;  // stride of mode t: N^(d-1-t)
#pragma unroll
        for (int t = D - 1; t >= 0; --t)
        {
#pragma unroll
            for (int j = 0; j < R; ++j)
            {
                int const idx = lane + j * kLanes;
                int const l = idx < SIZE ? idx : 0;
                int const digit = (l / st) % N;
                T const *xf = xs + (l - digit * st);
                T const *a = cs + t * NN + digit;
                T acc = T(0);
#pragma unroll
                for (int k = 0; k < N; ++k) acc = fma(xf[k * st], a[k * N], acc);
                y[j] = acc;
            }
            wave_sync();
            if (t > 0)
            {
#pragma unroll
                for (int j = 0; j < R; ++j) xs[lane + j * kLanes] = y[j];  // after every lane's reads
            }
            wave_sync();
            st *= N;
        }
        T *dst = out[e];
        if (dst != cur)
        {
#pragma unroll
            for (int j = 0; j < R; ++j)
            {
                int const idx = lane + j * kLanes;
                if (idx < SIZE) atomicAdd((T *)(kronreg::as_global(cur) + idx), run[j]);
                run[j] = T(0);
            }
            cur = dst;
        }
#pragma unroll
        for (int j = 0; j < R; ++j) run[j] += y[j];
    }
#pragma unroll
    for (int j = 0; j < R; ++j)
    {
        int const idx = lane + j * kLanes;
        if (idx < SIZE) atomicAdd((T *)(kronreg::as_global(cur) + idx), run[j]);
    }
}

template<typename T, int D, int N>
__global__ void __launch_bounds__(kLanes)
    kron_wave_kernel(int nb, int epw, T const *const *__restrict__ mats, int stride, T const *const *__restrict__ in,
                     T *const *__restrict__ out)
{
    __shared__ T rows[wave_lds<D, N>()];
    int const e0 = kronreg::xcd_run_index() * epw;
    int const e1 = (e0 + epw < nb) ? e0 + epw : nb;
    if (e0 >= e1) return;
    constexpr int XS = (wave_size<D, N>() + kLanes - 1) / kLanes * kLanes;
    wave_run<T, D, N>(e0, e1, mats, stride, in, out, rows, rows + XS, static_cast<int>(threadIdx.x));
}

// The one-element-per-lane shapes (grouped kernel and single batch), as X(D, N) for a macro X.
#define KRON_WAVE_SHAPES(X)                                                                                      \
    X(1, 2) X(2, 2) X(3, 2) X(4, 2) X(5, 2) X(6, 2) X(1, 3) X(2, 3) X(3, 3) X(1, 4) X(2, 4) X(3, 4) X(1, 5)       \
        X(2, 5) X(1, 6) X(2, 6) X(1, 7) X(2, 7) X(1, 8) X(2, 8)
// Shapes with 2 to 4 elements per lane and no register-resident plan (single batch only).
#define KRON_WAVE_SHAPES_WIDE(X) X(7, 2) X(8, 2) X(4, 3) X(5, 3) X(3, 5) X(3, 6) X(2, 9) X(2, 10)

// whether (d, n) is one of the shapes above
__host__ __device__ constexpr bool wave_family_shape(int d, int n)
{
    if (wave_shape(d, n)) return true;
    return (d == 7 && n == 2) || (d == 8 && n == 2) || (d == 4 && n == 3) || (d == 5 && n == 3) ||
           (d == 3 && n == 5) || (d == 3 && n == 6) || (d == 2 && n == 9) || (d == 2 && n == 10);
}

} // namespace kronwave
