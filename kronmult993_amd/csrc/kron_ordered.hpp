// kron_ordered.hpp -- run-to-run reproducible accumulation into shared outputs (no counterpart in the
// reference, whose epilogue is one atomicAdd per element and entry, kronmult_gpu/kronmult.cu:124-129, so
// the order of the additions into an aliased output changes from run to run).
//
// Precondition (checked on the host by the synchronous entry points): entries that share an output pointer
// are consecutive in the batch -- each output is one segment [s0, s1) of entries -- and distinct outputs do
// not overlap. The batch is cut into ranges of `epw` entries (the launch's entries per workgroup; 1 for the
// one-entry-per-workgroup kernels); a piece is a segment's part inside one range.
//   * The compute kernel sums each piece in entry order, P = ((0 + y_a) + y_a+1) + ..., and, when the piece
//     is its whole segment, adds it to the output (out = out + P, a plain read-modify-write: no other
//     workgroup touches that output); otherwise it stores P in workspace[first entry of the piece].
//   * kron_reduce_pieces_kernel then finishes every segment with more than one piece, in piece order:
//     S = P_0, S = S + P_1, ..., out = out + S. The workgroup of the range holding the segment's first
//     entry does it, so each output is written by exactly one workgroup.
// The result is therefore a fixed function of the inputs, nb and epw: bitwise the same on every run.
#pragma once

#include <hip/hip_runtime.h>

#include "kron_reg.hpp"

namespace kronordered
{

// True when piece [ps, pe) (entries with output `dst`) is its whole segment.
template<typename T>
__device__ __forceinline__ bool whole_segment(T *const *out, int nb, int ps, int pe, T const *dst)
{
    return (ps == 0 || out[ps - 1] != dst) && (pe == nb || out[pe] != dst);
}

// Segments with more than one piece (all = 0), or every segment (all = 1, epw must be 1: the compute
// kernels left every entry's contribution in its workspace).
// Grid: x = groups of `rpw` consecutive ranges, y = element chunks. A workgroup first checks its ranges in
// parallel (one per thread) for the segments they own -- a segment is owned by the range holding its first
// entry, and only if it continues past that range (or always, when all = 1) -- then, for each owned segment,
// looks up the workspaces of its further pieces (entries hi, hi + epw, ... while they share the output; 256
// positions per step, in parallel) and sums its chunks of elements. The chain of dependent loads therefore
// does not grow with the number of pieces, and few workgroups have nothing to do.
constexpr int kReduceThreads = 256;
constexpr int kReducePer = 1;  // elements per thread and chunk

template<typename T>
__global__ void __launch_bounds__(kReduceThreads)
    kron_reduce_pieces_kernel(int nb, int epw, int size, int all, int rpw, T *const *__restrict__ out,
                              T const *const *__restrict__ ws)
{
    __shared__ T const *piece[kReduceThreads];
    __shared__ int seg_first[kReduceThreads], seg_next[kReduceThreads];  // first entry; first entry after range
    __shared__ int nseg;
    int const t = threadIdx.x;
    long const nranges = (static_cast<long>(nb) + epw - 1) / epw;
    if (t == 0) nseg = 0;
    __syncthreads();
    for (int q = t; q < rpw; q += kReduceThreads)
    {
        long const r = static_cast<long>(blockIdx.x) * rpw + q;
        if (r >= nranges) break;
        int const lo = static_cast<int>(r * epw);
        int const hi = (nb - lo > epw) ? lo + epw : nb;
        T *const dst = out[hi - 1];
        int s0 = -1;
        if (all)
        {
            if (lo == 0 || out[lo - 1] != dst) s0 = lo;
        }
        else if (hi < nb && out[hi] == dst)
        {
            for (int k = lo; k < hi; ++k)  // the last segment start in the range, if any
                if (k == 0 || out[k - 1] != out[k]) s0 = k;
            if (s0 >= 0 && out[s0] != dst) s0 = -1;  // (cannot happen for a grouped batch)
        }
        if (s0 >= 0)
        {
            int const i = atomicAdd(&nseg, 1);
            seg_first[i] = s0;
            seg_next[i] = hi;
        }
    }
    __syncthreads();
    int const count = nseg;
    constexpr int kChunk = kReduceThreads * kReducePer;
    int const nchunks = (size + kChunk - 1) / kChunk;
    for (int g = 0; g < count; ++g)
    {
        int const s0 = seg_first[g], hi = seg_next[g];
        T *const dst = out[s0];
        kronreg::global_ptr<T> y = kronreg::as_global(dst);
        kronreg::global_ptr<T const> p0 = kronreg::as_global(ws[s0]);
        for (int c = blockIdx.y; c < nchunks; c += gridDim.y)
        {
            int const base = c * kChunk + t;
            T s[kReducePer];
#pragma unroll
            for (int j = 0; j < kReducePer; ++j)
            {
                int const i = base + j * kReduceThreads;
                s[j] = i < size ? p0[i] : T(0);
            }
            for (long m0 = hi;; m0 += static_cast<long>(kReduceThreads) * epw)
            {
                long const m = m0 + static_cast<long>(t) * epw;
                bool const valid = m < nb && out[m] == dst;
                __syncthreads();  // piece[] free
                piece[t] = valid ? ws[m] : nullptr;
                int const cnt = __syncthreads_count(valid);  // valid positions form a prefix
#pragma unroll 4
                for (int q = 0; q < cnt; ++q)
                {
                    kronreg::global_ptr<T const> p = kronreg::as_global(piece[q]);
#pragma unroll
                    for (int j = 0; j < kReducePer; ++j)
                    {
                        int const i = base + j * kReduceThreads;
                        if (i < size) s[j] = s[j] + p[i];
                    }
                }
                if (cnt < kReduceThreads) break;
            }
#pragma unroll
            for (int j = 0; j < kReducePer; ++j)
            {
                int const i = base + j * kReduceThreads;
                if (i < size) y[i] = y[i] + s[j];
            }
        }
    }
}

} // namespace kronordered
