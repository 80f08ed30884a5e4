// kron_reg_adjoint2.hpp -- adjoint on the two-modes-per-phase layout of the d = 6, n = 4 forward sweep
// (kron_plan.hpp Plan<6, 4, 2>: 256 threads, 16 values per thread; phase 0 holds modes 4, 5 in registers,
// phase 1 modes 2, 3, phase 2 modes 0, 1).
//
// Same mathematics as kron_reg_adjoint.hpp: d x = H_6 with H_0 = g, H_{t+1} = A_t^T on mode t of H_t, and
// dA_t = sum over the fibers of mode t of h_f x_f^T with X_t = (A_{t+1} .. A_5 on their modes) x. Per entry:
//   * X chain: x is read from the input in the phase-0 layout (each thread 16 consecutive elements),
//     modes 5, 4 contracted, exchanged through LDS into phase 1, modes 3, 2, into phase 2: X_1 is kept,
//     mode 1 contracted: X_0;
//   * H chain, phase 2: g read in the phase-2 layout (elements j * 256 + t, coalesced); dA_0 from (H_0,
//     X_0), A_0^T, dA_1 from (H_1, X_1), A_1^T, exchange into phase 1;
//   * X_3 kept from the first pass (below), X_2 = A_3 X_3, dA_2, A_2^T, dA_3, A_3^T, exchange into phase 0;
//     X_5 = x (read again) and X_4 (kept in FP32, A_5 x in FP64), dA_4, A_4^T, dA_5, A_5^T: d x in the
//     phase-0 layout.
// A factor gradient's per-thread partial (16 values: its 4 fibers of the mode) is reduce-scattered across
// the wavefront in four lane-shuffle steps to one value per lane and summed over consecutive entries that
// share the gradient pointer; when the pointer changes, two more steps, a sum over the four wavefronts in
// LDS and 16 atomics. d x is summed over entries that share the input gradient and leaves through one LDS
// exchange into the coalesced phase-2 layout.
#pragma once

#include <hip/hip_runtime.h>

#include "kron_reg.hpp"

namespace kronregadj2
{

using kronreg::as_constant;
using kronreg::as_global;
using kronreg::constant_ptr;
using kronreg::global_ptr;
using kronreg::lds_barrier;

using PL = kronplan::Plan<6, 4, 2>;
constexpr int kR = PL::R;  // 16
constexpr int kP = PL::P;  // 256
constexpr int kWaves = kP / 64;

// phase holding mode m in registers, and the stride of m in the register index
constexpr int phase_of(int m) { return m >= 4 ? 0 : m >= 2 ? 1 : 2; }
constexpr int stride_of(int m) { return PL::reg_stride(m - PL::lo(phase_of(m))); }

// contract register mode M with A_M (TRANS: A_M^T); the n-term fused chain, k ascending
template<typename T, int M, bool TRANS>
__device__ __forceinline__ void contract(T (&v)[kR], T const *A_generic, int stride)
{
    constexpr int S = stride_of(M);
    constant_ptr<T> A = as_constant(A_generic);
    T a[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int k = 0; k < 4; ++k) a[r][k] = TRANS ? A[k + r * stride] : A[r + k * stride];
#pragma unroll
    for (int base = 0; base < kR; ++base)
    {
        if ((base / S) % 4 != 0) continue;
        T x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = v[base + k * S];
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
            T acc = T(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) acc = fma(x[k], a[r][k], acc);
            v[base + r * S] = acc;
        }
    }
}

// This thread's partial of dA_M, p[4 r + c] = sum over its 4 fibers f of h[f][r] x[f][c], reduce-scattered
// across the wavefront: after step s (lane bit 5 - s) a lane keeps the half of its values selected by that
// bit, summed with the partner lane's copy of the same half. Lane l ends with value index
// 8 b5 + 4 b4 + 2 b3 + b2 (bits of l), summed over the 16 lanes that share its bits 0..1. The first step is
// fused with the partial (values k and k + 8 formed together and folded at once, so the 16 are never live
// together).
template<typename T, int M>
__device__ __forceinline__ void partial_half(T const (&h)[kR], T const (&x)[kR], T (&q8)[8], int lane)
{
    constexpr int S = stride_of(M);
    constexpr int O = S == 4 ? 1 : 4;
    bool const b5 = (lane >> 5) & 1;
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        T lo = T(0), hi = T(0);
        int const r0 = k >> 2, c0 = k & 3, r1 = (k + 8) >> 2, c1 = (k + 8) & 3;
#pragma unroll
        for (int f = 0; f < 4; ++f)
        {
            lo = fma(h[r0 * S + f * O], x[c0 * S + f * O], lo);
            hi = fma(h[r1 * S + f * O], x[c1 * S + f * O], hi);
        }
        T const mine = b5 ? hi : lo, other = b5 ? lo : hi;
        q8[k] = mine + __shfl_xor(other, 32);
    }
}

template<typename T>
__device__ __forceinline__ T reduce_scatter8(T (&q8)[8], int lane)
{
    T q4[4], q2[2];
    bool const b4 = (lane >> 4) & 1, b3 = (lane >> 3) & 1, b2 = (lane >> 2) & 1;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        T const mine = b4 ? q8[k + 4] : q8[k], other = b4 ? q8[k] : q8[k + 4];
        q4[k] = mine + __shfl_xor(other, 16);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)
    {
        T const mine = b3 ? q4[k + 2] : q4[k], other = b3 ? q4[k] : q4[k + 2];
        q2[k] = mine + __shfl_xor(other, 8);
    }
    T const mine = b2 ? q2[1] : q2[0], other = b2 ? q2[0] : q2[1];
    return mine + __shfl_xor(other, 4);
}

// add a run's dA partial (one value per lane, in the order above) to the gradient: lanes 0..3 of each
// quad hold partials of one index; the four wavefronts meet in red[]
template<typename T>
__device__ __forceinline__ void flush_gradient(T value, T *g_generic, int stride, T *red, int t)
{
    int const lane = t & 63, wave = t >> 6;
    value += __shfl_xor(value, 1);
    value += __shfl_xor(value, 2);
    int const idx = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
    if ((lane & 3) == 0) red[wave * 16 + idx] = value;
    lds_barrier();
    if (t < 16)
    {
        T s = red[t];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += red[w * 16 + t];
        global_ptr<T> g = as_global(g_generic);
        atomicAdd((T *)(g + (t >> 2) + (t & 3) * stride), s);  // index 4 r + c: dA[r][c]
    }
    lds_barrier();  // red[] free for the next flush
}

// exchange the vector from the layout of phase A to that of phase B through LDS (the forward's exchange
// swizzle on both sides)
template<typename T, int A, int B>
__device__ __forceinline__ void exchange(T *lds, T (&v)[kR], int t)
{
    using SW = kronplan::PlanSwizzle<PL, T, 1>;
    lds_barrier();
#pragma unroll
    for (int j = 0; j < kR; ++j) lds[SW::phys(PL::elem(A, t, j))] = v[j];
    lds_barrier();
#pragma unroll
    for (int j = 0; j < kR; ++j) v[j] = lds[SW::phys(PL::elem(B, t, j))];
}

// x in the phase-0 layout: thread t holds elements 16 t .. 16 t + 15
template<typename T>
__device__ __forceinline__ void load_x(T const *src_generic, T (&v)[kR], int t)
{
    static_assert(PL::L(0) == 1, "phase 0 holds 16 consecutive elements per thread");
    global_ptr<T const> src = as_global(src_generic);
#pragma unroll
    for (int j = 0; j < kR; ++j) v[j] = src[t * kR + j];
}

// X_3 (phase-1 layout) and, in FP32, X_4 (phase-0 layout) are kept from the first pass of the X chain in
// thread-private LDS slots (element j * P + t of the buffer: each thread reads back only what it wrote, so
// no barrier) instead of being rebuilt from x; FP64 keeps X_3 only, as 64 KiB of LDS per workgroup already
// matches its two resident workgroups per CU (223 VGPRs).
template<typename T>
constexpr bool keep_x4()
{
    return sizeof(T) == 4;
}

template<typename T>
__device__ __forceinline__ void keep(T *buf, T const (&v)[kR], int t)
{
#pragma unroll
    for (int j = 0; j < kR; ++j) buf[j * kP + t] = v[j];
}
template<typename T>
__device__ __forceinline__ void restore(T const *buf, T (&v)[kR], int t)
{
#pragma unroll
    for (int j = 0; j < kR; ++j) v[j] = buf[j * kP + t];
}

template<typename T>
__global__ void __launch_bounds__(kP)
    kron_reg_adjoint2_kernel(int nb, int epw, T const *const *__restrict__ mats, int stride,
                             T const *const *__restrict__ in, T const *const *__restrict__ gout,
                             T *const *__restrict__ gin, T *const *__restrict__ gmats)
{
    static_assert(PL::NPHASE == 3 && PL::elem(2, 1, 0) == 1 && PL::elem(2, 0, 1) == kP,
                  "phase 2 is element j * P + t");
    constexpr bool kKeep = true;
    constexpr bool kKeep4 = keep_x4<T>();
    __shared__ __attribute__((aligned(16))) T lds[PL::SIZE * (1 + (kKeep ? 1 : 0) + (kKeep4 ? 1 : 0))];
    __shared__ T red[kWaves * 16];
    T *const x3buf = lds + PL::SIZE;
    T *const x4buf = lds + 2 * PL::SIZE;
    int const t = threadIdx.x;
    int const lane = t & 63;
    int const e0 = kronreg::xcd_run_index() * epw;
    int const e1 = (e0 + epw < nb) ? e0 + epw : nb;
    if (e0 >= e1) return;
    bool const want_mats = gmats != nullptr;
    bool const want_input = gin != nullptr;

    T gacc[6], dxacc[kR];
#pragma unroll
    for (int m = 0; m < 6; ++m) gacc[m] = T(0);
#pragma unroll
    for (int j = 0; j < kR; ++j) dxacc[j] = T(0);

    // FP64: the next entry's cotangent is loaded while this entry runs (248 registers, still two waves per
    // SIMD; d=6 batch 273 -> 263 us). FP32 gains nothing from it (its three waves per SIMD hide the load),
    // and prefetching the input as well takes it to 180 registers and two waves (146 -> 161 us).
    constexpr bool kPreG = sizeof(T) == 8;
    T gpre[kPreG ? kR : 1];
    auto load_g = [&](int e, T (&dst)[kR]) {
        global_ptr<T const> g = as_global(gout[e]);
#pragma unroll
        for (int j = 0; j < kR; ++j) dst[j] = g[j * kP + t];
    };
    if constexpr (kPreG) load_g(e0, gpre);

    for (int e = e0; e < e1; ++e)
    {
        bool const last = e + 1 >= e1;
        T const *const *A = mats + static_cast<size_t>(e) * 6;
        T h[kR], v[kR], xs[kR];
        // one dA step: partial of mode M from (h, x), run sum, flush when the next entry's pointer differs
        auto grad = [&](auto mc, T const (&x)[kR]) {
            constexpr int M = decltype(mc)::value;
            T q8[8];
            partial_half<T, M>(h, x, q8, lane);
            gacc[M] += reduce_scatter8(q8, lane);
            T *const gA = gmats[static_cast<size_t>(e) * 6 + M];
            if (last || gmats[static_cast<size_t>(e + 1) * 6 + M] != gA)
            {
                flush_gradient(gacc[M], gA, stride, red, t);
                gacc[M] = T(0);
            }
        };
        // cotangent in the phase-2 layout: element j * 256 + t
        if constexpr (kPreG)
        {
#pragma unroll
            for (int j = 0; j < kR; ++j) h[j] = gpre[j];
            if (!last) load_g(e + 1, gpre);
        }
        else
            load_g(e, h);
        if (want_mats)
        {
            load_x(in[e], v, t);
            contract<T, 5, false>(v, A[5], stride);
            if constexpr (kKeep4) keep(x4buf, v, t);  // X_4
            contract<T, 4, false>(v, A[4], stride);
            exchange<T, 0, 1>(lds, v, t);
            if constexpr (kKeep) keep(x3buf, v, t);   // X_3
            contract<T, 3, false>(v, A[3], stride);
            contract<T, 2, false>(v, A[2], stride);
            exchange<T, 1, 2>(lds, v, t);
#pragma unroll
            for (int j = 0; j < kR; ++j) xs[j] = v[j];  // X_1
            contract<T, 1, false>(v, A[1], stride);      // X_0
            grad(std::integral_constant<int, 0>(), v);
        }
        contract<T, 0, true>(h, A[0], stride);  // H_1
        if (want_mats) grad(std::integral_constant<int, 1>(), xs);
        contract<T, 1, true>(h, A[1], stride);  // H_2
        exchange<T, 2, 1>(lds, h, t);
        if (want_mats)
        {
            if constexpr (kKeep)
                restore(x3buf, v, t);
            else
            {
                load_x(in[e], v, t);
                contract<T, 5, false>(v, A[5], stride);
                contract<T, 4, false>(v, A[4], stride);
                exchange<T, 0, 1>(lds, v, t);
            }
#pragma unroll
            for (int j = 0; j < kR; ++j) xs[j] = v[j];  // X_3
            contract<T, 3, false>(v, A[3], stride);      // X_2
            grad(std::integral_constant<int, 2>(), v);
        }
        contract<T, 2, true>(h, A[2], stride);  // H_3
        if (want_mats) grad(std::integral_constant<int, 3>(), xs);
        contract<T, 3, true>(h, A[3], stride);  // H_4
        exchange<T, 1, 0>(lds, h, t);
        if (want_mats)
        {
            load_x(in[e], xs, t);  // X_5
            if constexpr (kKeep4)
                restore(x4buf, v, t);
            else
            {
#pragma unroll
                for (int j = 0; j < kR; ++j) v[j] = xs[j];
                contract<T, 5, false>(v, A[5], stride);  // X_4
            }
            grad(std::integral_constant<int, 4>(), v);
        }
        contract<T, 4, true>(h, A[4], stride);  // H_5
        if (want_mats) grad(std::integral_constant<int, 5>(), xs);
        if (want_input)
        {
            contract<T, 5, true>(h, A[5], stride);  // d x, phase-0 layout
#pragma unroll
            for (int j = 0; j < kR; ++j) dxacc[j] += h[j];
            if (last || gin[e + 1] != gin[e])
            {
                exchange<T, 0, 2>(lds, dxacc, t);
                global_ptr<T> gi = as_global(gin[e]);
#pragma unroll
                for (int j = 0; j < kR; ++j)
                {
                    atomicAdd((T *)(gi + j * kP + t), dxacc[j]);
                    dxacc[j] = T(0);
                }
            }
        }
    }
}

} // namespace kronregadj2
