// kron_pass.hpp -- the sweep for vectors too large for one workgroup's LDS, in a few passes over HBM.
//
// The d modes are cut into groups of consecutive modes, fastest group first (e.g. d = 6, n = 8: modes 5..3,
// then 2..0). A pass applies every factor of one group, in the reference order (fastest mode first), to the
// whole batch: the n^d vector of an entry is a set of independent sub-tensors (the group's modes varying,
// the others fixed), and a workgroup loads a tile of them into LDS -- TO consecutive values of the slower
// modes times TI consecutive values of the faster modes, so every global access runs over TI (or whole
// sub-tensors of) contiguous elements -- contracts the group's modes in LDS (fiber per thread, n-term fused
// dot, k ascending from zero, factors from the last to the first) and writes the tile back. The first pass
// reads the input, the last one adds into the output (hardware atomics: outputs may alias; a workgroup walks
// a run of entries at one tile position and sums consecutive entries with the same output pointer in
// registers first), the passes in between work in place in the caller's per-entry workspace. Each entry's arithmetic is the reference
// sweep's, mode by mode, so its contribution is bit-identical to the single-workgroup kernels'.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kron_reg.hpp"

namespace kronpass
{

constexpr int kThreads = 256;
constexpr int kTileBytes = 32 * 1024;

struct PassArgs
{
    int d;      // modes of the vector
    int a;      // this pass: modes a .. a+G-1 (G a template parameter)
    int inner;  // n^(d-a-G): stride of the group's fastest mode
    int TO;     // tile: TO consecutive outer indices (x TI consecutive inner indices, a template parameter)
    int tiles_i;          // inner / TI
    int tiles_per_entry;  // (outer / TO) * tiles_i
    int first, last;      // read the input / add into the output
    int nb;               // entries
    int epw;              // consecutive entries per workgroup (same tile position; the last pass sums the
                          // runs of equal output pointers in registers before its atomic adds)
};

constexpr int ipow(int b, int e)
{
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

// contract local mode M (stride SK = N^(G-1-M) inside a sub-tensor) of every sub-tensor of the tile
template<typename T, int N, int G, int TI, int M>
__device__ __forceinline__ void contract_modes(T *tile, int TO, T const *const *A_entry, int a, int stride)
{
    if constexpr (M >= 0)
    {
        constexpr int NG = ipow(N, G), KF = NG / N, SK = ipow(N, G - 1 - M), STEP = SK * TI;
        kronreg::constant_ptr<T> A = kronreg::as_constant(A_entry[a + M]);
        T c[N][N];
#pragma unroll
        for (int r = 0; r < N; ++r)
#pragma unroll
            for (int k = 0; k < N; ++k) c[r][k] = A[r + k * stride];
        int const nfib = TO * KF * TI;
        for (int f = threadIdx.x; f < nfib; f += blockDim.x)
        {
            int const ti = f % TI;
            int const rest = f / TI;
            int const kf = rest % KF;
            int const to = rest / KF;
            int const first = (to * NG + (kf / SK) * (N * SK) + kf % SK) * TI + ti;
            T v[N];
#pragma unroll
            for (int k = 0; k < N; ++k) v[k] = tile[first + k * STEP];
#pragma unroll
            for (int r = 0; r < N; ++r)
            {
                T acc = T(0);
#pragma unroll
                for (int k = 0; k < N; ++k) acc = fma(v[k], c[r][k], acc);
                tile[first + r * STEP] = acc;
            }
        }
        __syncthreads();
        contract_modes<T, N, G, TI, M - 1>(tile, TO, A_entry, a, stride);
    }
}

template<typename T, int N, int G, int TI>
__global__ void __launch_bounds__(kThreads)
    kron_pass_kernel(PassArgs P, T const *const *__restrict__ mats, int stride, T const *const *__restrict__ in,
                     T *const *__restrict__ ws, T *const *__restrict__ out)
{
    constexpr int NG = ipow(N, G);
    constexpr int kPer = kTileBytes / static_cast<int>(sizeof(T)) / kThreads;  // tile elements per thread
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *const tile = reinterpret_cast<T *>(smem_raw);
    long const blk = blockIdx.x;
    int const run = static_cast<int>(blk / P.tiles_per_entry);
    int const w = static_cast<int>(blk - static_cast<long>(run) * P.tiles_per_entry);
    int const e_begin = run * P.epw;
    int const e_end = e_begin + P.epw < P.nb ? e_begin + P.epw : P.nb;
    int const o0 = (w / P.tiles_i) * P.TO;
    int const i0 = (w % P.tiles_i) * TI;
    int const nelem = P.TO * NG * TI;
    long const sub = static_cast<long>(NG) * P.inner;  // elements between consecutive outer indices
    long const base = static_cast<long>(o0) * sub + i0;

    // tile layout [to][k][ti]; global element base + to * sub + k * inner + ti. Runs of TI elements (the whole
    // tile when TI = 1, since then inner = 1) are contiguous in memory: 16-byte accesses when aligned.
    constexpr int GR = kronreg::Granule<T>::G;
    constexpr bool kVec = TI == 1 || TI % GR == 0;
    using V = typename kronreg::Granule<T>::type;
    auto offset = [&](int idx) {
        int const ti = idx % TI;
        int const rest = idx / TI;
        int const k = rest % NG;
        int const to = rest / NG;
        return base + static_cast<long>(to) * sub + static_cast<long>(k) * P.inner + ti;
    };

    T acc[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) acc[j] = T(0);

    for (int e = e_begin; e < e_end; ++e)
    {
        T const *src_generic = P.first ? in[e] : static_cast<T const *>(ws[e]);
        T *dst_generic = P.last ? out[e] : ws[e];
        bool const vec = kVec && nelem % GR == 0 && base % GR == 0 &&
                         ((reinterpret_cast<uintptr_t>(src_generic) | reinterpret_cast<uintptr_t>(dst_generic)) & 15) == 0;
        {
            kronreg::global_ptr<T const> src = kronreg::as_global(src_generic);
            if (vec)
            {
                for (int gi = threadIdx.x; gi < nelem / GR; gi += blockDim.x)
                    *reinterpret_cast<V *>(tile + gi * GR) = *(kronreg::global_ptr<V const>)(src + offset(gi * GR));
            }
            else
            {
                for (int idx = threadIdx.x; idx < nelem; idx += blockDim.x) tile[idx] = src[offset(idx)];
            }
        }
        __syncthreads();
        contract_modes<T, N, G, TI, G - 1>(tile, P.TO, mats + static_cast<size_t>(e) * P.d, P.a, stride);

        kronreg::global_ptr<T> dst = kronreg::as_global(dst_generic);
        if (P.last)
        {
#pragma unroll
            for (int j = 0; j < kPer; ++j)
            {
                int const idx = threadIdx.x + j * kThreads;
                if (idx < nelem) acc[j] += tile[idx];
            }
            if (e + 1 == e_end || out[e + 1] != dst_generic)
            {
#pragma unroll
                for (int j = 0; j < kPer; ++j)
                {
                    int const idx = threadIdx.x + j * kThreads;
                    if (idx < nelem) atomicAdd((T *)(dst + offset(idx)), acc[j]);
                    acc[j] = T(0);
                }
            }
        }
        else if (vec)
        {
            for (int gi = threadIdx.x; gi < nelem / GR; gi += blockDim.x)
                *(kronreg::global_ptr<V>)(dst + offset(gi * GR)) = *reinterpret_cast<V const *>(tile + gi * GR);
        }
        else
        {
            for (int idx = threadIdx.x; idx < nelem; idx += blockDim.x) dst[offset(idx)] = tile[idx];
        }
        __syncthreads();  // the tile is free for the next entry
    }
}

} // namespace kronpass
