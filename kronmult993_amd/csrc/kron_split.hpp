// kron_split.hpp -- FP64 batched Kronecker mat-vec with the contractions on FP32 matrix cores (no counterpart
// in the reference, which has plain FP64 and FP32 only, kronmult_gpu/kronmult.cu:202-224).
//
// Inputs, outputs and the vector between modes are FP64; each mode contraction y_r = sum_k A[r][k] x_k is
// computed on V_MFMA_F32_16X16X4_F32 from FP32 (or BF16-valued) slices of the operands, and the partial
// products are recombined in FP64. Variants (SplitMode):
//   FP32       x and A rounded to FP32, one product, FP32 accumulation;
//   PAIR32     hi = fp32(v), lo = fp32(v - hi): hi*hi + hi*lo + lo*hi in one FP32 accumulator;
//   BF16       x and A rounded to BF16 (values exact in FP32), one product, FP32 accumulation;
//   PAIRBF16   BF16 hi + BF16 lo, three products, one FP32 accumulator;
//   FIXED3     error-free splitting (Ozaki): per fiber of x and per row of A a power-of-two scale 2^e with
//              max |v| < 2^e; hi and mid are b-bit fixed-point slices on the grids 2^(e-b), 2^(e-2b), lo the
//              FP32 remainder. Products of fixed-point slices are integers on a common grid, and b is
//              chosen so that n such products sum below 2^24 grid units (b = (24 - ceil(log2 n)) / 2):
//              the matrix-core sums of hi*hi, hi*mid + mid*hi and mid*mid are exact. The terms with lo
//              (hi*lo, lo*hi, mid*lo, lo*mid) go to a fourth FP32 accumulator whose rounding is 2^-24 of a
//              term already 2^-2b smaller. lo*lo (2^-4b) is dropped. The four accumulators are added in FP64.
//              Error per mode ~ 2^-(2b+24) of sum |A||x|: ~1e-13 relative for n <= 8.
// The plain FP32/BF16 variants exist to report what the splitting buys; FIXED3 is the accurate one.
//
// This generic kernel serves every shape except FIXED3 with n = 4 on the register plans (d = 4, 5, 6), which
// runs on kron_split_reg.hpp.
// One workgroup (4 wavefronts) per entry; the entry's vector lives in LDS as FP64 (n^d * 8 <= 64 KiB). A
// mode step takes tiles of 16 fibers: D (16 fibers x 16 outputs) = X^T (16 x 4) . A^T (4 x 16) per K-step of
// 4, rows r >= n and k >= n zero-padded. Output accumulation into the caller's outputs: FP64 atomics.
#pragma once

#include <hip/hip_runtime.h>

#include "kron_reg.hpp"

namespace kronsplit
{

enum SplitMode
{
    FP32 = 0,
    PAIR32 = 1,
    BF16 = 2,
    PAIRBF16 = 3,
    FIXED3 = 4
};

constexpr int kThreads = 256;
constexpr int kMaxN = 8;

typedef float f32x4 __attribute__((ext_vector_type(4)));

template<int MODE>
struct SplitTraits
{
    static constexpr int kSlices = (MODE == FP32 || MODE == BF16) ? 1 : (MODE == FIXED3 ? 3 : 2);
    static constexpr int kAcc = MODE == FIXED3 ? 4 : 1;
};

// round to the nearest BF16 value (ties to even), returned as FP32
__device__ __forceinline__ float round_bf16(float v)
{
    unsigned u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

constexpr int fixed_bits(int n)
{
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    return (24 - lg) / 2;
}

// slices of v; `e`: scale exponent of v's group (max |v| < 2^e), used by FIXED3 only
template<int MODE, int N>
__device__ __forceinline__ void split(double v, int e, float (&s)[SplitTraits<MODE>::kSlices])
{
    if constexpr (MODE == FP32)
        s[0] = static_cast<float>(v);
    else if constexpr (MODE == BF16)
        s[0] = round_bf16(static_cast<float>(v));
    else if constexpr (MODE == PAIR32)
    {
        s[0] = static_cast<float>(v);
        s[1] = static_cast<float>(v - static_cast<double>(s[0]));
    }
    else if constexpr (MODE == PAIRBF16)
    {
        s[0] = round_bf16(static_cast<float>(v));
        s[1] = round_bf16(static_cast<float>(v - static_cast<double>(s[0])));
    }
    else
    {
        constexpr int b = fixed_bits(N);
        double const hi = ldexp(rint(ldexp(v, b - e)), e - b);
        double const r1 = v - hi;
        double const mid = ldexp(rint(ldexp(r1, 2 * b - e)), e - 2 * b);
        s[0] = static_cast<float>(hi);
        s[1] = static_cast<float>(mid);
        s[2] = static_cast<float>(r1 - mid);
    }
}

// scale exponent: smallest e with max_abs < 2^e (0 for a zero group)
__device__ __forceinline__ int scale_exp(double max_abs)
{
    return max_abs > 0.0 ? ilogb(max_abs) + 1 : 0;
}

// accumulator of the product (A slice sa) x (x slice sx); -1 = dropped
template<int MODE>
constexpr int term_acc(int sa, int sx)
{
    if (MODE != FIXED3) return (sa + sx <= 1) ? 0 : -1;
    if (sa == 0 && sx == 0) return 0;
    if (sa + sx == 1) return 1;
    if (sa == 1 && sx == 1) return 2;
    if (sa == 2 && sx == 2) return -1;
    return 3;
}

// One mode step of the entry's vector x (LDS, FP64) with factor A (global, column-major, stride).
// aslice: LDS scratch of kSlices * 16 * N floats.
template<int MODE, int N>
__device__ __forceinline__ void mode_step(double *x, double const *A_generic, int stride, int st, int nfib,
                                          float *aslice)
{
    using TR = SplitTraits<MODE>;
    constexpr int S = TR::kSlices;
    constexpr int NK = (N + 3) / 4;  // K-steps of 4
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const nwaves = blockDim.x >> 6;

    // factor slices: row r scaled by its own max over k
    kronreg::global_ptr<double const> A = kronreg::as_global(A_generic);
    if (threadIdx.x < N)
    {
        int const r = threadIdx.x;
        double row[N];
        double m = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k)
        {
            row[k] = A[r + k * stride];
            m = fmax(m, fabs(row[k]));
        }
        int const e = scale_exp(m);
#pragma unroll
        for (int k = 0; k < N; ++k)
        {
            float s[S];
            split<MODE, N>(row[k], e, s);
#pragma unroll
            for (int q = 0; q < S; ++q) aslice[(q * 16 + r) * N + k] = s[q];
        }
    }
    __syncthreads();

    int const i = lane & 15;   // fiber within the tile (A operand row) / output r (B operand column)
    int const kk = lane >> 4;  // k within the K-step
    int const ntiles = (nfib + 15) >> 4;
    for (int tile = wave; tile < ntiles; tile += nwaves)
    {
        // operands of this lane for every K-step: x of fiber tile*16+i at k = ks*4+kk, A[r=i][k]
        int const f = tile * 16 + i;
        bool const fiber_ok = f < nfib;
        int base = 0;
        if (fiber_ok)
        {
            int const hi = f / st;
            base = hi * st * N + (f - hi * st);
        }
        double xv[NK];
        double m = 0.0;
#pragma unroll
        for (int ks = 0; ks < NK; ++ks)
        {
            int const k = ks * 4 + kk;
            xv[ks] = (fiber_ok && k < N) ? x[base + k * st] : 0.0;
            m = fmax(m, fabs(xv[ks]));
        }
        int e = 0;
        if constexpr (MODE == FIXED3)
        {
            m = fmax(m, __shfl_xor(m, 16));
            m = fmax(m, __shfl_xor(m, 32));
            e = scale_exp(m);
        }
        f32x4 acc[TR::kAcc];
#pragma unroll
        for (int a = 0; a < TR::kAcc; ++a) acc[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < NK; ++ks)
        {
            int const k = ks * 4 + kk;
            float xs[S];
            split<MODE, N>(xv[ks], e, xs);
            float as[S];
#pragma unroll
            for (int q = 0; q < S; ++q) as[q] = (i < N && k < N) ? aslice[(q * 16 + i) * N + k] : 0.f;
#pragma unroll
            for (int sa = 0; sa < S; ++sa)
#pragma unroll
                for (int sx = 0; sx < S; ++sx)
                {
                    int const a = term_acc<MODE>(sa, sx);
                    if (a >= 0)
                        acc[a] = __builtin_amdgcn_mfma_f32_16x16x4f32(xs[sx], as[sa], acc[a], 0, 0, 0);
                }
        }
        // lane holds D[4*(lane/16) + v][lane % 16]: fiber tile*16 + 4*(lane/16) + v, output r = lane % 16.
        // Every lane of the wave has finished reading these fibers (the MFMAs consumed them), so the
        // results can be written in place.
        int const r = lane & 15;
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            int const fo = tile * 16 + 4 * (lane >> 4) + v;
            if (fo < nfib && r < N)
            {
                int const hi2 = fo / st;
                int const b2 = hi2 * st * N + (fo - hi2 * st);
                double y = 0.0;
#pragma unroll
                for (int a = TR::kAcc - 1; a >= 0; --a) y += static_cast<double>(acc[a][v]);
                x[b2 + r * st] = y;
            }
        }
    }
    __syncthreads();
}

template<int MODE, int N>
__global__ void __launch_bounds__(kThreads)
    kron_split_kernel(int d, int size, double const *const *__restrict__ mats, int stride,
                      double const *const *__restrict__ in, double *const *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *const x = reinterpret_cast<double *>(smem_raw);
    float *const aslice = reinterpret_cast<float *>(x + size);
    int const b = blockIdx.x;
    double const *const *A = mats + static_cast<size_t>(b) * d;
    kronreg::global_ptr<double const> src = kronreg::as_global(in[b]);
    for (int e = threadIdx.x; e < size; e += blockDim.x) x[e] = src[e];
    __syncthreads();
    int const nfib = size / N;
    int st = 1;
    for (int t = d - 1; t >= 0; --t)
    {
        mode_step<MODE, N>(x, A[t], stride, st, nfib, aslice);
        st *= N;
    }
    kronreg::global_ptr<double> y = kronreg::as_global(out[b]);
    for (int e = threadIdx.x; e < size; e += blockDim.x) atomicAdd((double *)(y + e), x[e]);
}

} // namespace kronsplit
