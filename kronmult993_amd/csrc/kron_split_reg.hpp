// kron_split_reg.hpp -- FP64 batched Kronecker mat-vec whose products run on the matrix cores through
// error-free fixed-point splitting, in the register-resident layout of kron_reg.hpp (plans with n = 4). No
// counterpart in the reference, which has plain FP64 and FP32 paths only (kronmult_gpu/kronmult.cu:202-224).
//
// Between modes a value is carried as an unevaluated pair of FP32 numbers (hi + lo, about 48 bits). For one
// mode step y_i = sum_k A[i][k] x_k of a fiber (the four values of the contracted mode, held by one thread):
//   * the factor row i is scaled by 2^(11 - e_i) (max_k |A[i][k]| < 2^e_i) and split into a' = ha + ma + ra:
//     ha an integer (|ha| <= 2^11), ma a multiple of 2^-11 (|ma| <= 1/2), ra the FP32 remainder (<= 2^-12);
//   * the fiber is scaled by 2^(10 - e) (max_k |x_k| < 2^e) and split the same way into x' = hx + mx + rx
//     (|hx| <= 2^10);
//   * ha hx, ha mx + ma hx and ma mx are exact on V_MFMA_F32_4X4X4_16B_F16 (integers and multiples of 2^-11
//     and 2^-22 whose partial sums stay below 2^24 grid units, so the FP32 accumulation never rounds); the
//     remainder terms (ha + ma) rx + ra x', 2^-11 of the result and below, accumulate in FP32 on
//     V_MFMA_F32_4X4X1_16B_F32 (their rounding is ~2^-35 of them);
//   * the integer part of the second sum joins the first exactly, the rest forms the low word of the new
//     pair, which is scaled back by 2^(e_i + e - 21).
// Per mode the error is about 2^-44 of sum_k |A[i][k]| |x_k|. Lane layout of both instructions (measured on
// gfx950): lane 4b + i supplies row i of the factor, lane 4b + j supplies fiber j of block b and receives its
// output i in result register i. Every lane keeps its own fibers, so the matrix cores take the place of the
// FMA chains of the FP64 sweep without moving data.
//
// The factors of an entry are split once, cooperatively (one element per thread), into a double-buffered
// LDS table that every lane reads for its row; the next entry's factor elements are loaded while the
// current entry runs. Input staging, LDS exchanges (an FP32 pair occupies the slot of a double, so the FP64
// swizzles apply), the run accumulation in FP64 and the atomic flush are those of the FP64 sweep.
#pragma once

#include <hip/hip_runtime.h>

#include "kron_plan.hpp"
#include "kron_reg.hpp"

namespace kronsplitreg
{

using kronreg::Lds;
using kronreg::lds_barrier;

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// one factor row prepared for the matrix cores (48 bytes)
struct RowData
{
    h4 hi;   // ha[k], exact integers
    h4 mid;  // ma[k], exact multiples of 2^-11
    f4 a32;  // ha[k] + ma[k], exact in FP32
    f4 rem;  // ra[k]
};
constexpr int kRowBytes = 48;
constexpr int kFactorBytes = 4 * kRowBytes + 16;  // four rows, then the four row exponents

struct Pair
{
    float hi, lo;
};

__device__ __forceinline__ double to_slot(float hi, float lo)
{
    return __builtin_bit_cast(double, Pair{hi, lo});
}
__device__ __forceinline__ Pair from_slot(double s)
{
    return __builtin_bit_cast(Pair, s);
}

// FP64 value -> FP32 pair (hi = nearest FP32, lo = nearest FP32 to the rest)
__device__ __forceinline__ double pair_of(double x)
{
    float const hi = static_cast<float>(x);
    return to_slot(hi, static_cast<float>(x - static_cast<double>(hi)));
}

__device__ __forceinline__ double value_of(double slot)
{
    Pair const p = from_slot(slot);
    return static_cast<double>(p.hi) + static_cast<double>(p.lo);
}

// exponent e with |m| < 2^e (0 for m = 0)
__device__ __forceinline__ int exp_above(float m)
{
    return __builtin_amdgcn_frexp_expf(m);
}

// The factor elements one thread prepares: element idx = t + q*P of the entry's d*16 (factor m = idx/16,
// row r = idx/4 % 4, column k = idx % 4; the four elements of a row sit in four consecutive lanes).
template<typename PL>
struct FactorPrep
{
    static constexpr int kCount = PL::kD * 16;
    static constexpr int kPer = (kCount + PL::P - 1) / PL::P;
    double val[kPer];

    __device__ __forceinline__ void load(double const *const *mats, int stride, int t)
    {
#pragma unroll
        for (int q = 0; q < kPer; ++q)
        {
            int const idx = t + q * PL::P;
            if (kCount % PL::P == 0 || idx < kCount)
            {
                double const *A = mats[idx >> 4];
                val[q] = kronreg::as_global(A)[((idx >> 2) & 3) + (idx & 3) * stride];
            }
        }
    }

    // split and store into the table at `tab` (kFactorBytes per factor)
    __device__ __forceinline__ void store(unsigned char *tab, int t) const
    {
#pragma unroll
        for (int q = 0; q < kPer; ++q)
        {
            int const idx = t + q * PL::P;
            // every lane of a quad takes part in the row maximum (quads never straddle kCount: 16 | kCount)
            double const a = (kCount % PL::P == 0 || idx < kCount) ? val[q] : 0.0;
            double m = fabs(a);
            m = fmax(m, __shfl_xor(m, 1));
            m = fmax(m, __shfl_xor(m, 2));
            if (kCount % PL::P != 0 && idx >= kCount) continue;
            int e = 0;
            (void)frexp(m, &e);  // m < 2^e (e = 0 for m = 0)
            double const s = ldexp(a, 11 - e);
            double const ha = rint(s);
            double const u = s - ha;
            double const ma = rint(u * 2048.0) * (1.0 / 2048.0);
            double const ra = u - ma;
            int const f = idx >> 4, r = (idx >> 2) & 3, k = idx & 3;
            unsigned char *row = tab + f * kFactorBytes + r * kRowBytes;
            reinterpret_cast<_Float16 *>(row)[k] = static_cast<_Float16>(ha);
            reinterpret_cast<_Float16 *>(row + 8)[k] = static_cast<_Float16>(ma);
            reinterpret_cast<float *>(row + 16)[k] = static_cast<float>(ha + ma);
            reinterpret_cast<float *>(row + 32)[k] = static_cast<float>(ra);
            if (k == 0) reinterpret_cast<int *>(tab + f * kFactorBytes + 4 * kRowBytes)[r] = e;
        }
    }
};

// the prepared factor of one mode, as this lane reads it (row lane & 3), and the four row exponents
struct Factor
{
    RowData row;
    int e[4];
};

__device__ __forceinline__ void read_factor(unsigned char const *tab, int m, int lane, Factor &F)
{
    unsigned char const *base = tab + m * kFactorBytes;
    F.row = *reinterpret_cast<RowData const *>(base + (lane & 3) * kRowBytes);
    int4 const e = *reinterpret_cast<int4 const *>(base + 4 * kRowBytes);
    F.e[0] = __builtin_amdgcn_readfirstlane(e.x);
    F.e[1] = __builtin_amdgcn_readfirstlane(e.y);
    F.e[2] = __builtin_amdgcn_readfirstlane(e.z);
    F.e[3] = __builtin_amdgcn_readfirstlane(e.w);
}

// One mode step on the fibers of register stride S (R / 4 fibers per thread).
//   x' = x * 2^(10 - e) = hx + mx + rx: |hx| <= 2^10 an integer, mx a multiple of 2^-11 (|mx| <= 1/2), rx the
//   FP32 remainder; the factor row a' = a * 2^(11 - e_i) = ha + ma + ra as prepared in the table.
//   acc0 = sum ha hx      (integers, |acc0| <= 2^23)
//   acc1 = sum ha mx + ma hx  (multiples of 2^-11, |acc1| <= 6144: below 2^24 grid units)
//   acc2 = sum ma mx      (multiples of 2^-22, |acc2| <= 1)
//   c    = sum (ha + ma) rx + ra x'   (FP32)
// The integer part of acc1 joins acc0 exactly (|sum| < 2^24), the rest forms the low word: the result
// pair is (acc0 + rint(acc1), (acc1 - rint(acc1)) + acc2 + c) scaled by 2^(e_i + e - 21). The pair is not
// renormalised: its low word is at most ~6 units of the high word's grid, which matters only for values
// that have cancelled to a few units, and there only at ~2^-24 of those few units.
template<int R, int S>
__device__ __forceinline__ void contract_mode(double (&v)[R], Factor const &F)
{
    constexpr float kMagic = 6144.0f;  // 1.5 * 2^12: (u + kMagic) - kMagic rounds |u| <= 1/2 to 2^-11
#pragma unroll
    for (int base = 0; base < R; ++base)
    {
        if ((base / S) % 4 != 0) continue;
        float xh[4], xl[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            Pair const p = from_slot(v[base + k * S]);
            xh[k] = p.hi;
            xl[k] = p.lo;
        }
        float const mx = fmaxf(fmaxf(fabsf(xh[0]), fabsf(xh[1])), fmaxf(fabsf(xh[2]), fabsf(xh[3])));
        // |x| < 2^ex; fibers below 2^-100 share that scale, so 2^(10 - ex) stays a finite FP32 power of two
        int const ex = max(exp_above(mx), -100);
        float const sc = ldexpf(1.0f, 10 - ex);
        float xs[4], rx[4];
        h4 hx, mxs;
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            xs[k] = xh[k] * sc;
            float const h = rintf(xs[k]);
            float const u = xs[k] - h;
            float const m = (u + kMagic) - kMagic;
            rx[k] = fmaf(xl[k], sc, u - m);
            hx[k] = static_cast<_Float16>(h);
            mxs[k] = static_cast<_Float16>(m);
        }
        f4 const zero = {0.f, 0.f, 0.f, 0.f};
        f4 const acc0 = __builtin_amdgcn_mfma_f32_4x4x4f16(F.row.hi, hx, zero, 0, 0, 0);
        f4 acc1 = __builtin_amdgcn_mfma_f32_4x4x4f16(F.row.hi, mxs, zero, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_4x4x4f16(F.row.mid, hx, acc1, 0, 0, 0);
        f4 const acc2 = __builtin_amdgcn_mfma_f32_4x4x4f16(F.row.mid, mxs, zero, 0, 0, 0);
        f4 c = zero;
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            c = __builtin_amdgcn_mfma_f32_4x4x1f32(F.row.a32[k], rx[k], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_4x4x1f32(F.row.rem[k], xs[k], c, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            float const ai = rintf(acc1[i]);
            float const hi = acc0[i] + ai;
            float const lo = (acc1[i] - ai) + (acc2[i] + c[i]);
            int const E = F.e[i] + ex - 21;
            v[base + i * S] = to_slot(ldexpf(hi, E), ldexpf(lo, E));
        }
        // one fiber at a time. Measured on the d=6 n=4 batch: 171 us per launch; two fibers interleaved, or
        // independent accumulators for the chained products, 184-205 us (more VALU work or fewer waves per
        // SIMD). On gfx950 these matrix-core instructions do not overlap VALU work of the same SIMD, so the
        // step costs the sum of both.
        __builtin_amdgcn_sched_barrier(0);
    }
}

// contract modes M, M-1, ..., LO of phase PH (fastest first), each with its factor read from the table just
// before use (the table is double-buffered per entry, so the reads need no barrier of their own)
template<typename PL, int PH, int M>
__device__ __forceinline__ void contract_phase(double (&v)[PL::R], unsigned char const *tab, int lane)
{
    if constexpr (M >= PL::lo(PH))
    {
        Factor F;
        read_factor(tab, M, lane, F);
        contract_mode<PL::R, PL::reg_stride(M - PL::lo(PH))>(v, F);
        contract_phase<PL, PH, M - 1>(v, tab, lane);
    }
}

// Phases from PH on. The factor table of this entry is `tab`; the next entry's factor elements (prep) are
// loaded at phase 1 and split into `next_tab` before the last phase. after_last_read runs once every
// thread holds the last phase's values.
template<typename PL, int PH, typename G>
__device__ __forceinline__ void run_phases(double *lds, double (&v)[PL::R], unsigned char const *tab,
                                           unsigned char *next_tab, FactorPrep<PL> &prep, bool has_next,
                                           double const *const *next_mats, int stride, int t, G const &after_last_read)
{
    if constexpr (PH + 1 == PL::NPHASE) after_last_read();
    if constexpr (PH == 1)
    {
        if (has_next) prep.load(next_mats, stride, t);
    }
    contract_phase<PL, PH, PL::c_hi(PH) - 1>(v, tab, t & 63);
    if constexpr (PH + 1 < PL::NPHASE)
    {
        if constexpr (PH + 2 == PL::NPHASE)
        {
            if (has_next) prep.store(next_tab, t);
        }
        lds_barrier();  // every thread has read phase PH
        Lds<double, PL>::template write_xor<PH>(lds, v, t);
        lds_barrier();
        Lds<double, PL>::template read_xor<PH + 1>(lds, v, t);
        run_phases<PL, PH + 1>(lds, v, tab, next_tab, prep, has_next, next_mats, stride, t, after_last_read);
    }
}

template<typename PL>
constexpr int table_bytes()
{
    return 2 * PL::kD * kFactorBytes;
}

// Input staging by LDS-DMA where the FP64 sweep uses it (d = 5, 6 n = 4), else through registers.
template<typename PL>
constexpr bool dma_staging()
{
    return kronreg::dma_staging<double, PL>();
}

// Three waves per SIMD: the d=6 n=4 kernel then takes 156 VGPRs without spills (four waves would spill 21).
template<typename PL>
__global__ void __launch_bounds__(PL::P) __attribute__((amdgpu_waves_per_eu(3)))
    kron_split_reg_kernel(int nb, int epw, double const *const *__restrict__ mats, int stride,
                          double const *const *__restrict__ in, double *const *__restrict__ out)
{
    static_assert(PL::kN == 4 && PL::NPHASE >= 3, "plans with n = 4 and at least three phases");
    constexpr int R = PL::R;
    __shared__ __attribute__((aligned(16))) double lds[PL::SIZE + table_bytes<PL>() / 8];
    unsigned char *const tabs = reinterpret_cast<unsigned char *>(lds + PL::SIZE);
    int const t = threadIdx.x;
    int const e0 = kronreg::xcd_run_index() * epw;
    int const e1 = (e0 + epw < nb) ? e0 + epw : nb;
    if (e0 >= e1) return;
    auto tab_of = [&](int e) { return tabs + (e & 1) * (PL::kD * kFactorBytes); };

    FactorPrep<PL> prep;
    prep.load(mats + static_cast<size_t>(e0) * PL::kD, stride, t);
    prep.store(tab_of(e0), t);

    double acc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) acc[j] = 0.0;
    double v[R];
    double *cur = out[e0];

    bool staged = false;
    auto issue = [&](int e) -> bool {
        if constexpr (dma_staging<PL>())
        {
            double const *src = in[e];
            if ((reinterpret_cast<uintptr_t>(src) & 15) != 0) return false;
            Lds<double, PL>::dma_input(lds, src, t);
            return true;
        }
        else
        {
            return false;
        }
    };
    staged = issue(e0);
    for (int e = e0; e < e1; ++e)
    {
        if (!staged)
        {
            if constexpr (!dma_staging<PL>()) lds_barrier();  // the previous entry's last phase has been read
            double tmp[R];
            kronreg::load_input<double, PL>(in[e], tmp, t);
            Lds<double, PL>::store_input(lds, tmp, t);
        }
        __builtin_amdgcn_s_waitcnt(kronreg::kWaitVmcnt0);
        lds_barrier();  // input staged; this entry's factor table written
        if constexpr (dma_staging<PL>())
            Lds<double, PL>::read0_dma(lds, v, t);
        else
            Lds<double, PL>::template read<0>(lds, v, t);
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = pair_of(v[j]);
        bool const has_next = e + 1 < e1;
        run_phases<PL, 0>(lds, v, tab_of(e), tab_of(e + 1), prep, has_next,
                          mats + static_cast<size_t>(e + 1) * PL::kD, stride, t, [&]() {
                              lds_barrier();  // every thread has read the last phase: LDS free
                              staged = has_next ? issue(e + 1) : true;
                          });
        double *dst = out[e];
        if (dst != cur)
        {
            kronreg::flush<double, PL>(cur, acc, t);
            cur = dst;
        }
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] += value_of(v[j]);
    }
    kronreg::flush<double, PL>(cur, acc, t);
}

}  // namespace kronsplitreg
