// kron_pass.hpp -- the sweep for vectors too large for one workgroup's LDS, in a few passes over HBM.
//
// The d modes are cut into groups of consecutive modes, fastest group first (e.g. d = 6, n = 8: modes 5..3,
// then 2..0). A pass applies every factor of one group, in the reference order (fastest mode first), to the
// whole batch: the n^d vector of an entry is a set of independent sub-tensors (the group's modes varying,
// the others fixed), and a workgroup loads a tile of them into LDS -- TO consecutive values of the slower
// modes times TI consecutive values of the faster modes, so every global access runs over TI (or whole
// sub-tensors of) contiguous elements -- contracts the group's modes in LDS (fiber per thread, n-term fused
// dot, k ascending from zero, factors from the last to the first) and writes the tile back. The first pass
// reads the input, the last one adds into the output (hardware atomics: outputs may alias), the passes in
// between work in place in the caller's per-entry workspace. Each entry's arithmetic is the reference
// sweep's, mode by mode, so its contribution is bit-identical to the single-workgroup kernels'.
#pragma once

#include <hip/hip_runtime.h>

#include "kron_reg.hpp"

namespace kronpass
{

constexpr int kThreads = 256;
constexpr int kTileBytes = 32 * 1024;

struct PassArgs
{
    int d;      // modes of the vector
    int a, g;   // this pass: modes a .. a+g-1
    int ng;     // n^g: elements of one sub-tensor
    int inner;  // n^(d-a-g): stride of the group's fastest mode
    int TO, TI; // tile: TO consecutive outer indices x TI consecutive inner indices
    int tiles_i;          // inner / TI
    int tiles_per_entry;  // (outer / TO) * tiles_i
    int first, last;      // read the input / add into the output
};

template<typename T, int N>
__global__ void __launch_bounds__(kThreads)
    kron_pass_kernel(PassArgs P, T const *const *__restrict__ mats, int stride, T const *const *__restrict__ in,
                     T *const *__restrict__ ws, T *const *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *const tile = reinterpret_cast<T *>(smem_raw);
    long const blk = blockIdx.x;
    int const e = static_cast<int>(blk / P.tiles_per_entry);
    int const w = static_cast<int>(blk - static_cast<long>(e) * P.tiles_per_entry);
    int const o0 = (w / P.tiles_i) * P.TO;
    int const i0 = (w % P.tiles_i) * P.TI;
    int const nelem = P.TO * P.ng * P.TI;
    long const sub = static_cast<long>(P.ng) * P.inner;  // elements between consecutive outer indices
    long const base = static_cast<long>(o0) * sub + i0;

    // tile layout [to][k][ti]; global element (o0 + to) * sub + k * inner + i0 + ti
    {
        kronreg::global_ptr<T const> src = kronreg::as_global(P.first ? in[e] : static_cast<T const *>(ws[e]));
        for (int idx = threadIdx.x; idx < nelem; idx += blockDim.x)
        {
            int const ti = idx % P.TI;
            int const rest = idx / P.TI;
            int const k = rest % P.ng;
            int const to = rest / P.ng;
            tile[idx] = src[base + static_cast<long>(to) * sub + static_cast<long>(k) * P.inner + ti];
        }
    }
    __syncthreads();

    int const kf_count = P.ng / N;                 // fibers per sub-tensor and inner index
    int const nfib = P.TO * kf_count * P.TI;
    int sk = 1;                                    // stride of the local mode within the sub-tensor
    for (int m = P.g - 1; m >= 0; --m)
    {
        kronreg::constant_ptr<T> A = kronreg::as_constant(mats[static_cast<size_t>(e) * P.d + P.a + m]);
        T c[N][N];
#pragma unroll
        for (int r = 0; r < N; ++r)
#pragma unroll
            for (int k = 0; k < N; ++k) c[r][k] = A[r + k * stride];
        for (int f = threadIdx.x; f < nfib; f += blockDim.x)
        {
            int const ti = f % P.TI;
            int const rest = f / P.TI;
            int const kf = rest % kf_count;
            int const to = rest / kf_count;
            int const kbase = (kf / sk) * (N * sk) + kf % sk;
            int const first = (to * P.ng + kbase) * P.TI + ti;
            int const step = sk * P.TI;
            T v[N];
#pragma unroll
            for (int k = 0; k < N; ++k) v[k] = tile[first + k * step];
#pragma unroll
            for (int r = 0; r < N; ++r)
            {
                T acc = T(0);
#pragma unroll
                for (int k = 0; k < N; ++k) acc = fma(v[k], c[r][k], acc);
                tile[first + r * step] = acc;
            }
        }
        __syncthreads();
        sk *= N;
    }

    kronreg::global_ptr<T> dst = kronreg::as_global(P.last ? out[e] : ws[e]);
    for (int idx = threadIdx.x; idx < nelem; idx += blockDim.x)
    {
        int const ti = idx % P.TI;
        int const rest = idx / P.TI;
        int const k = rest % P.ng;
        int const to = rest / P.ng;
        long const off = base + static_cast<long>(to) * sub + static_cast<long>(k) * P.inner + ti;
        if (P.last)
            atomicAdd((T *)(dst + off), tile[idx]);
        else
            dst[off] = tile[idx];
    }
}

} // namespace kronpass
