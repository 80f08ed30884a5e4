// kron_group.hpp -- one persistent launch over several batches of different (d, n) (no counterpart in the
// reference, whose entry point takes one (d, n, stride) per call, kronmult_gpu/kronmult.cu:173-197).
//
// The launch covers a table of groups. Each group is cut into tiles of `epw` consecutive entries; the tiles
// of all groups are numbered one after the other and workgroup b takes tiles b, b + grid, b + 2 grid, ...
// (static round-robin: no counters to reset, so the launch can be captured in a graph and replayed).
// A tile is processed in passes of E entries at once: thread t works on entry slot t / nfib and fiber
// t % nfib (nfib = n^(d-1) fibers per mode), so small tensors still fill the workgroup. Per pass the E
// input vectors and their d factors are staged in LDS, the d modes are contracted in place (the fiber
// sweep: n-term fused dot per output, k ascending, factors from the last to the first -- bit-identical per
// entry to the reference's sweep with fused multiply-adds), and the results are added to the outputs:
// consecutive entries of the pass that target the same output are summed in LDS order first, then one
// hardware atomic add per element and run.
#pragma once

#include <hip/hip_runtime.h>

#include "kron_reg.hpp"

namespace krongroup
{

constexpr int kMaxGroups = 16;
constexpr int kThreads = 256;

template<typename T>
struct GroupDesc
{
    int d, n, stride, nb;
    int size;       // n^d
    int nfib;       // n^(d-1)
    int E;          // entries per pass
    int epw;        // entries per tile
    T const *const *mats;
    T const *const *in;
    T *const *out;
};

template<typename T>
struct GroupTable
{
    int count;
    int tiles;                         // total over the groups
    int tile_start[kMaxGroups + 1];    // first tile of each group (prefix sums)
    GroupDesc<T> g[kMaxGroups];
};

// process entries [e0, e0 + cnt) of group G (cnt <= G.E) with N = G.n fixed at compile time
template<typename T, int N>
__device__ __forceinline__ void pass(GroupDesc<T> const &G, int e0, int cnt, T *lds)
{
    int const d = G.d, size = G.size, nfib = G.nfib;
    int const nn = N * N;
    T *const x = lds;                               // cnt vectors of `size`
    T *const coef = lds + static_cast<size_t>(G.E) * size;  // cnt x d factors, row-major a[r][k]
    int const tid = threadIdx.x;

    for (int i = tid; i < cnt * size; i += blockDim.x)
    {
        int const q = i / size;
        int const e = i - q * size;
        x[i] = kronreg::as_global(G.in[e0 + q])[e];
    }
    int const ncoef = cnt * d * nn;
    for (int i = tid; i < ncoef; i += blockDim.x)
    {
        int const qm = i / nn;               // q * d + m
        int const rk = i - qm * nn;
        int const r = rk / N, k = rk - (rk / N) * N;
        coef[i] = kronreg::as_global(G.mats[static_cast<size_t>(e0) * d + qm])[r + k * G.stride];
    }
    __syncthreads();

    int st = 1;
    for (int t = d - 1; t >= 0; --t)
    {
        for (int w = tid; w < cnt * nfib; w += blockDim.x)
        {
            int const q = w / nfib;
            int const f = w - q * nfib;
            int const hi = f / st;
            int const base = q * size + hi * st * N + (f - hi * st);
            T const *a = coef + (q * d + t) * nn;
            T v[N];
#pragma unroll
            for (int k = 0; k < N; ++k) v[k] = x[base + k * st];
#pragma unroll
            for (int r = 0; r < N; ++r)
            {
                T acc = T(0);
#pragma unroll
                for (int k = 0; k < N; ++k) acc = fma(v[k], a[r * N + k], acc);
                x[base + r * st] = acc;
            }
        }
        __syncthreads();
        st *= N;
    }

    // accumulate: the first entry of every run of equal output pointers adds the run's sum
    for (int i = tid; i < cnt * size; i += blockDim.x)
    {
        int const q = i / size;
        int const e = i - q * size;
        T *const dst = G.out[e0 + q];
        if (q > 0 && G.out[e0 + q - 1] == dst) continue;
        T s = x[i];
        for (int q2 = q + 1; q2 < cnt && G.out[e0 + q2] == dst; ++q2) s += x[q2 * size + e];
        atomicAdd((T *)(kronreg::as_global(dst) + e), s);
    }
    __syncthreads();  // LDS free for the next pass
}

template<typename T>
__global__ void __launch_bounds__(kThreads) kron_group_kernel(GroupTable<T> const table)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *const lds = reinterpret_cast<T *>(smem_raw);
    for (int tile = blockIdx.x; tile < table.tiles; tile += gridDim.x)
    {
        int gi = 0;
        while (gi + 1 < table.count && tile >= table.tile_start[gi + 1]) ++gi;
        GroupDesc<T> const &G = table.g[gi];
        int const e_begin = (tile - table.tile_start[gi]) * G.epw;
        int const e_end = e_begin + G.epw < G.nb ? e_begin + G.epw : G.nb;
        for (int e0 = e_begin; e0 < e_end; e0 += G.E)
        {
            int const cnt = e_end - e0 < G.E ? e_end - e0 : G.E;
            switch (G.n)
            {
            case 1: pass<T, 1>(G, e0, cnt, lds); break;
            case 2: pass<T, 2>(G, e0, cnt, lds); break;
            case 3: pass<T, 3>(G, e0, cnt, lds); break;
            case 4: pass<T, 4>(G, e0, cnt, lds); break;
            case 5: pass<T, 5>(G, e0, cnt, lds); break;
            case 6: pass<T, 6>(G, e0, cnt, lds); break;
            case 7: pass<T, 7>(G, e0, cnt, lds); break;
            case 8: pass<T, 8>(G, e0, cnt, lds); break;
            default: break;  // rejected on the host
            }
        }
    }
}

} // namespace krongroup
