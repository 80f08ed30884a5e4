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
#include "kron_wave.hpp"

namespace krongroup
{

constexpr int kMaxGroups = 16;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kronwave::kLanes;

// LDS elements of one wavefront's rows on the wave path
__host__ __device__ constexpr int wave_lds_elems()
{
    return kronwave::kLanes + kronwave::kMaxCoef;  // one element per lane, d n^2 <= 128
}

template<typename T>
struct GroupDesc
{
    int wave;  // 1: the group's tiles run on the wave path (n^d <= 64, kron_wave.hpp)
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

// LDS layout of a pass. For n = 4 and 8 every row of n consecutive vector elements is followed by one
// element of padding, and every entry's d factors by 16 bytes: the fastest mode's fibers (n elements apart
// across lanes) and the factors of different entries (read as broadcasts, one address per entry) then fall
// into different 16-byte banks instead of 4 of them.
template<int N>
constexpr bool pad_rows()
{
    return N == 4 || N == 8;
}
__host__ __device__ constexpr int padded_size(int n, int size)
{
    return (n == 4 || n == 8) ? size + size / n : size;
}
__host__ __device__ constexpr int coef_pitch(int d, int n)
{
    return d * n * n + 2;
}
// LDS elements of type T holding one entry's d + 1 pointers
template<typename T>
__host__ __device__ constexpr int ptr_elems(int d)
{
    return static_cast<int>(((d + 1) * sizeof(void *) + sizeof(T) - 1) / sizeof(T));
}

// process entries [e0, e0 + cnt) of group G (cnt <= G.E) with N = G.n fixed at compile time
template<typename T, int N>
__device__ __forceinline__ void pass(GroupDesc<T> const &G, int e0, int cnt, T *lds, T **out_slot)
{
    int const d = G.d, size = G.size, nfib = G.nfib;
    int const nn = N * N;
    int const cp = coef_pitch(d, N);
    auto px = [](int i) { return pad_rows<N>() ? i + i / N : i; };  // element i of the entry-major vectors
    T *const x = lds;                                                           // cnt padded vectors
    T *const coef = lds + static_cast<size_t>(G.E) * padded_size(N, size);      // entry q at q * cp: a[m][r][k]
    // the pass's input and factor pointers (E and E*d slots after the factors), read once, in parallel
    constexpr int kPtrAlign = static_cast<int>(sizeof(void *) / sizeof(T));
    size_t const ptr_off = (static_cast<size_t>(G.E) * (padded_size(N, size) + cp) + kPtrAlign - 1) / kPtrAlign * kPtrAlign;
    T const **const ptr = reinterpret_cast<T const **>(lds + ptr_off);
    int const tid = threadIdx.x;

    int const nptr = cnt * (d + 1);
    for (int j = tid; j < nptr; j += blockDim.x)
        ptr[j] = j < cnt ? G.in[e0 + j] : G.mats[static_cast<size_t>(e0) * d + (j - cnt)];
    if (tid < cnt) out_slot[tid] = G.out[e0 + tid];  // the pass's output pointers
    __syncthreads();
    for (int i = tid; i < cnt * size; i += blockDim.x)
    {
        int const q = i / size;
        x[px(i)] = kronreg::as_global(ptr[q])[i - q * size];
    }
    int const ncoef = cnt * d * nn;
    for (int i = tid; i < ncoef; i += blockDim.x)
    {
        int const qm = i / nn;  // q * d + m
        int const rk = i - qm * nn;
        int const r = rk / N, k = rk - (rk / N) * N;
        int const q = qm / d;
        coef[q * cp + (qm - q * d) * nn + rk] = kronreg::as_global(ptr[cnt + qm])[r + k * G.stride];
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
            T const *a = coef + q * cp + t * nn;
            T v[N];
#pragma unroll
            for (int k = 0; k < N; ++k) v[k] = x[px(base + k * st)];
#pragma unroll
            for (int r = 0; r < N; ++r)
            {
                T acc = T(0);
#pragma unroll
                for (int k = 0; k < N; ++k) acc = fma(v[k], a[r * N + k], acc);
                x[px(base + r * st)] = acc;
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
        T *const dst = out_slot[q];
        if (q > 0 && out_slot[q - 1] == dst) continue;
        T s = x[px(i)];
        for (int q2 = q + 1; q2 < cnt && out_slot[q2] == dst; ++q2) s += x[px(q2 * size + e)];
        atomicAdd((T *)(kronreg::as_global(dst) + e), s);
    }
    __syncthreads();  // LDS free for the next pass
}

// WAVE: the table may hold wave groups (the grouped launch). Their inlined paths take the kernel from 48 to
// 82 VGPRs; the cap keeps 6 waves per SIMD without spills. The packed single-batch launch has no wave
// groups and keeps 8 waves per SIMD.
template<typename T, bool WAVE>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WAVE ? 6 : 1)))
    kron_group_kernel(GroupTable<T> const table)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ T *out_slot[kThreads];  // a pass holds at most kThreads entries (one fiber each at least)
    T *const lds = reinterpret_cast<T *>(smem_raw);
    for (int tile = blockIdx.x; tile < table.tiles; tile += gridDim.x)
    {
        int gi = 0;
        while (gi + 1 < table.count && tile >= table.tile_start[gi + 1]) ++gi;
        GroupDesc<T> const &G = table.g[gi];
        int const e_begin = (tile - table.tile_start[gi]) * G.epw;
        int const e_end = e_begin + G.epw < G.nb ? e_begin + G.epw : G.nb;
        if (WAVE && G.wave)
        {
            // n^d <= 64: each wavefront sweeps a contiguous quarter of the tile on its own (kron_wave.hpp)
            int const w = static_cast<int>(threadIdx.x) / kronwave::kLanes;
            int const lane = static_cast<int>(threadIdx.x) - w * kronwave::kLanes;
            int const chunk = (e_end - e_begin + kWaves - 1) / kWaves;
            int const a = e_begin + w * chunk;
            int const b = a + chunk < e_end ? a + chunk : e_end;
            T *const xs = lds + w * wave_lds_elems();
            T *const cs = xs + kronwave::kLanes;
            if (a < b)
            {
                switch (G.d * 16 + G.n)
                {
#define KRON_GROUP_WAVE(D, N)                                                                          \
    case D * 16 + N: kronwave::wave_run<T, D, N, true>(a, b, G.mats, G.stride, G.in, G.out, xs, cs, lane); break;
                    KRON_WAVE_SHAPES(KRON_GROUP_WAVE)
#undef KRON_GROUP_WAVE
                default: break;
                }
            }
            __syncthreads();  // LDS free for the next tile
            continue;
        }
        for (int e0 = e_begin; e0 < e_end; e0 += G.E)
        {
            int const cnt = e_end - e0 < G.E ? e_end - e0 : G.E;
            switch (G.n)
            {
            case 1: pass<T, 1>(G, e0, cnt, lds, out_slot); break;
            case 2: pass<T, 2>(G, e0, cnt, lds, out_slot); break;
            case 3: pass<T, 3>(G, e0, cnt, lds, out_slot); break;
            case 4: pass<T, 4>(G, e0, cnt, lds, out_slot); break;
            case 5: pass<T, 5>(G, e0, cnt, lds, out_slot); break;
            case 6: pass<T, 6>(G, e0, cnt, lds, out_slot); break;
            case 7: pass<T, 7>(G, e0, cnt, lds, out_slot); break;
            case 8: pass<T, 8>(G, e0, cnt, lds, out_slot); break;
            default: break;  // rejected on the host
            }
        }
    }
}

} // namespace krongroup
