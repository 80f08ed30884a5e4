// kron_mfma.hpp -- the d = 6, n = 4 FP64 sweep with every mode contraction on the matrix cores.
//
// A workgroup of 256 threads (4 wavefronts) holds one entry's 4096-element tensor, 16 values per thread, and
// walks a run of entries like the register-resident sweep (kron_reg.hpp): input prefetched one entry ahead
// with 16-byte loads and staged through LDS, same-output entries summed in registers, one atomic flush.
//
// The 12 index bits of an element (mode t = bits 2(5-t) .. 2(5-t)+1, mode 0 slowest) are spread over the
// wavefront (2 bits), the lane -- kb = lane bits 4-5, jb = bits 2-3, s = bits 0-1 -- and the register,
// j = 4*j1 + j0. V_MFMA_F64_4X4X4_4B_F64 with A = the factor (lane 16*kb + 4*jb + i holds A[i][kb]) and
// B = a register (lane 16*kb + 4*jb + s holds element kb of fiber (jb, s)) returns in lane 16*i + 4*jb + s
// output i of that fiber: one instruction contracts the mode held in kb for 16 fibers, in place. Measured
// on gfx950, its result equals the fused chain acc = fma(x_k, A[i][k], acc), k = 0..3 from zero, bit for
// bit -- the reference's step (kronmult_gpu/kronmult.cu:66-70 with fused multiply-adds) -- so each entry's
// contribution is bit-identical to the other kernels'.
//
// Modes reach kb in the reference order 5, 4, 3, 2, 1, 0:
//   layout A  (read from the staged input):  kb=5 j0=4 j1=3 s=2 jb=1 w=0
//     contract kb; swap kb<->j0; contract kb (mode 4); swap kb<->j1; contract kb (mode 3)
//   layout A2 (after those swaps):          kb=3 j0=5 j1=4 s=2 jb=1 w=0   -> LDS (16-byte pair writes)
//   layout B  (read back):                   kb=2 j0=1 j1=0 s=5 jb=4 w=3
//     contract kb; swap kb<->j0; contract (mode 1); swap kb<->j1; contract (mode 0)
//   layout B2 (result):                      kb=0 j0=2 j1=1 s=5 jb=4 w=3
// A swap of lane bits 4-5 with two register-index bits is a 4 x 4 transpose over 4 lanes x 4 registers,
// done with V_PERMLANE32_SWAP (lane bit 5 <-> a register bit) and V_PERMLANE16_SWAP (lane bit 4) on
// register pairs -- no LDS. One LDS round trip per entry instead of two; both LDS stages use XOR swizzles
// found by search under the measured bank model (conflict-free for all four access patterns).
//
// Measured on MI355X (8192-entry d=6 n=4 batch, FP64): 67 us per launch at 3 workgroups per CU against
// 53 us for the register-resident sweep with register staging (48 us with LDS-DMA staging). The FP64 matrix rate of gfx950 equals its vector rate (one shared
// FP64 datapath), so moving the FMAs to V_MFMA gains no throughput, and the permlane transposes and the
// 3-waves-per-SIMD register budget (a few spilled values) cost more than the LDS exchange they replace.
// The kernel is kept as an opt-in launch family (and an autotuner candidate) for parts whose FP64 matrix
// rate exceeds the vector rate.
#pragma once

#include <hip/hip_runtime.h>

#include "kron_plan.hpp"
#include "kron_reg.hpp"

namespace kronmfma
{

using PL = kronplan::Plan<6, 4, 2>;  // 256 threads, 16 values each (its input prefetch order is reused)

// swizzle masks (granule-index bits XORed into the bank bits), 16 bits per mask
constexpr unsigned long long kMaskIn = 908ull | (1793ull << 16) | (2030ull << 32) | (954ull << 48);
constexpr unsigned long long kMaskEx = 1503ull | (1534ull << 16) | (354ull << 32) | (1797ull << 48);
using SwIn = kronplan::Swizzle<kMaskIn, 2>;
using SwEx = kronplan::Swizzle<kMaskEx, 2>;

// element index of (wavefront w, lane fields kb/jb/s, register j) in each layout
__device__ __forceinline__ int elem_a(int w, int kb, int jb, int s, int j)
{
    return kb + 4 * (j & 3) + 16 * (j >> 2) + 64 * s + 256 * jb + 1024 * w;
}
__device__ __forceinline__ int elem_a2(int w, int kb, int jb, int s, int j)
{
    return 16 * kb + (j & 3) + 4 * (j >> 2) + 64 * s + 256 * jb + 1024 * w;
}
__device__ __forceinline__ int elem_b(int w, int kb, int jb, int s, int j)
{
    return 64 * kb + 256 * (j & 3) + 1024 * (j >> 2) + s + 4 * jb + 16 * w;
}
__device__ __forceinline__ int elem_b2(int w, int kb, int jb, int s, int j)
{
    return 1024 * kb + 64 * (j & 3) + 256 * (j >> 2) + s + 4 * jb + 16 * w;
}

// (a, b) -> values with lane bit 5 (SHIFT 32) or 4 (SHIFT 16) and the a/b index exchanged
template<int SHIFT>
__device__ __forceinline__ void swap_pair(double &a, double &b)
{
    unsigned const alo = static_cast<unsigned>(__double_as_longlong(a));
    unsigned const ahi = static_cast<unsigned>(static_cast<unsigned long long>(__double_as_longlong(a)) >> 32);
    unsigned const blo = static_cast<unsigned>(__double_as_longlong(b));
    unsigned const bhi = static_cast<unsigned>(static_cast<unsigned long long>(__double_as_longlong(b)) >> 32);
    if constexpr (SHIFT == 32)
    {
        auto const lo = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
        auto const hi = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
        a = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi[0]) << 32) | lo[0]));
        b = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi[1]) << 32) | lo[1]));
    }
    else
    {
        auto const lo = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
        auto const hi = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
        a = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi[0]) << 32) | lo[0]));
        b = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi[1]) << 32) | lo[1]));
    }
}

// exchange kb (lane bits 4-5) with the register bits selected by STEP (1: j0, 4: j1)
template<int STEP>
__device__ __forceinline__ void swap_kb(double (&v)[16])
{
#pragma unroll
    for (int g = 0; g < 4; ++g)
    {
        int const base = STEP == 1 ? 4 * g : g;
        double &r0 = v[base], &r1 = v[base + STEP], &r2 = v[base + 2 * STEP], &r3 = v[base + 3 * STEP];
        swap_pair<32>(r0, r2);
        swap_pair<32>(r1, r3);
        swap_pair<16>(r0, r1);
        swap_pair<16>(r2, r3);
    }
}

__device__ __forceinline__ void contract_kb(double (&v)[16], double a)
{
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, v[j], 0.0, 0, 0, 0);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
    kron_mfma_d6n4_kernel(int nb, int epw, double const *const *__restrict__ mats, int stride,
                          double const *const *__restrict__ in, double *const *__restrict__ out)
{
    __shared__ __attribute__((aligned(16))) double lds[4096];
    int const t = threadIdx.x;
    int const w = t >> 6, lane = t & 63;
    int const kb = lane >> 4, jb = (lane >> 2) & 3, s = lane & 3;
    int const e0 = kronreg::xcd_run_index() * epw;
    int const e1 = (e0 + epw < nb) ? e0 + epw : nb;
    if (e0 >= e1) return;

    // the factor operand of this lane for every mode: A_m[i = lane & 3][k = lane >> 4]
    int const a_off = (lane & 3) + (lane >> 4) * stride;
    auto load_factors = [&](int e, double (&a)[6]) {
#pragma unroll
        for (int m = 0; m < 6; ++m) a[m] = kronreg::as_global(mats[static_cast<size_t>(e) * 6 + m])[a_off];
    };

    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0;
    double pre[16], apre[6], a[6], v[16];
    kronreg::load_input<double, PL>(in[e0], pre, t);
    load_factors(e0, apre);
    double *cur = out[e0];

    for (int e = e0; e < e1; ++e)
    {
        kronreg::lds_barrier();  // LDS free: everyone has read the previous entry's layout B
#pragma unroll
        for (int c = 0; c < 8; ++c)  // granule c*256 + t (the prefetch order)
        {
            typedef double d2 __attribute__((ext_vector_type(2)));
            d2 const g = {pre[2 * c], pre[2 * c + 1]};
            *reinterpret_cast<d2 *>(lds + SwIn::phys(2 * (c * 256 + t))) = g;
        }
#pragma unroll
        for (int m = 0; m < 6; ++m) a[m] = apre[m];
        kronreg::lds_barrier();
        if (e + 1 < e1)
        {
            kronreg::load_input<double, PL>(in[e + 1], pre, t);
            load_factors(e + 1, apre);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = lds[SwIn::phys(elem_a(w, kb, jb, s, j))];

        contract_kb(v, a[5]);
        swap_kb<1>(v);
        contract_kb(v, a[4]);
        swap_kb<4>(v);
        contract_kb(v, a[3]);

        kronreg::lds_barrier();  // everyone has read layout A
#pragma unroll
        for (int j1 = 0; j1 < 4; ++j1)
#pragma unroll
            for (int q = 0; q < 2; ++q)
            {
                typedef double d2 __attribute__((ext_vector_type(2)));
                int const j = 4 * j1 + 2 * q;  // registers j, j+1: consecutive elements (j0 holds mode 5)
                d2 const g = {v[j], v[j + 1]};
                *reinterpret_cast<d2 *>(lds + SwEx::phys(elem_a2(w, kb, jb, s, j))) = g;
            }
        kronreg::lds_barrier();
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = lds[SwEx::phys(elem_b(w, kb, jb, s, j))];

        contract_kb(v, a[2]);
        swap_kb<1>(v);
        contract_kb(v, a[1]);
        swap_kb<4>(v);
        contract_kb(v, a[0]);

        double *dst = out[e];
        if (dst != cur)
        {
            kronreg::global_ptr<double> g = kronreg::as_global(cur);
#pragma unroll
            for (int j = 0; j < 16; ++j)
            {
                atomicAdd((double *)(g + elem_b2(w, kb, jb, s, j)), acc[j]);
                acc[j] = 0.0;
            }
            cur = dst;
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] += v[j];
    }
    kronreg::global_ptr<double> g = kronreg::as_global(cur);
#pragma unroll
    for (int j = 0; j < 16; ++j) atomicAdd((double *)(g + elem_b2(w, kb, jb, s, j)), acc[j]);
}

} // namespace kronmfma
