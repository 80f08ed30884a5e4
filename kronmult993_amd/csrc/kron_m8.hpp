// kron_m8.hpp -- matrix-core sweep for n = 8, FP64, d = 3 and 4 (layout in kron_m8_plan.hpp).
//
// The register-resident sweep feeds every FMA of an n = 8 FP64 plan one factor coefficient through an LDS
// broadcast (64 per mode and thread: 32 ds_read_b128 per wavefront and mode), which makes those plans
// LDS-bound. Here a mode contraction is 16 V_MFMA_F64_4X4X4_4B_F64 per wavefront whose factor operand is
// four values per lane, read once per mode; the data stays in the MFMA operand layout, two modes per LDS
// round trip. Same run structure as kron_reg_kernel: a workgroup walks a run of consecutive entries, the
// input and the factors of the next entry are prefetched into registers, entries with the same output are
// summed in registers and flushed with one atomic per element (or per ordered piece). Every output value
// is the reference's fused chain over k ascending from zero (kronmult_gpu/kronmult.cu:66-70), modes from
// d-1 down to 0 (kronmult.cu:104): bit-identical to the register-resident sweep and the oracle.
#pragma once

#include <hip/hip_runtime.h>

#include "kron_m8_plan.hpp"
#include "kron_reg.hpp"

namespace kronm8
{

using kronreg::as_global;
using kronreg::global_ptr;
using kronreg::lds_barrier;

// byte offset in LDS of thread part / register part of layout ST under swizzle SW (the swizzle is linear
// over the two disjoint fields, so an access address is base(t) ^ reg(j))
template<int D, int ST, typename SW>
struct Field
{
    using PL = M8Plan<D>;
    static constexpr unsigned reg(int j) { return static_cast<unsigned>(SW::phys(PL::reg_part(ST, j)) * 8); }
    static __device__ __forceinline__ unsigned base(int t)
    {
        return static_cast<unsigned>(SW::phys(PL::thread_part(ST, t)) * 8);
    }
};

// The base is made opaque: the compiler otherwise proves some base ^ reg(j) = base + reg(j) and pairs
// those reads into ds_read2_b64, whose two addresses per lane then share a bank (measured with
// SQ_LDS_BANK_CONFLICT: 32 cycles per P0 read of a wavefront; none as eight ds_read_b64).
template<int D, int ST, typename SW>
__device__ __forceinline__ void lds_read(char const *lds, unsigned base, double (&v)[8])
{
    asm volatile("" : "+v"(base));
#pragma unroll
    for (int j = 0; j < 8; ++j)
        v[j] = *reinterpret_cast<double const *>(lds + (base ^ Field<D, ST, SW>::reg(j)));
}

template<int D, int ST, typename SW>
__device__ __forceinline__ void lds_write(char *lds, unsigned base, double const (&v)[8])
{
#pragma unroll
    for (int j = 0; j < 8; ++j) *reinterpret_cast<double *>(lds + (base ^ Field<D, ST, SW>::reg(j))) = v[j];
}

// Staged factors: the four values lane (j = lane & 3, kb = (lane >> 4) & 3) takes in one mode are
// a[j + 4 rg][kb + 4 kg], (rg, kg) = (0,0), (0,1), (1,0), (1,1), stored together at slot kb*4 + j of the
// mode's 64 values: two 16-byte reads per lane, the 16 slots of a 16-lane group in distinct banks (the
// row-major copy read as ds_read2_b64 pairs cost 16 conflict cycles per mode and wavefront).
constexpr int coef_slot(int r, int k)
{
    return (((k & 3) * 4 + (r & 3)) * 4) + (r >> 2) * 2 + (k >> 2);
}

template<typename PL>
__device__ __forceinline__ void store_coef(double *coef, double const (&pre)[kronreg::Coef<double, PL>::kPerThread],
                                          int t)
{
    using CF = kronreg::Coef<double, PL>;
#pragma unroll
    for (int i = 0; i < CF::kPerThread; ++i)
    {
        int const idx = t + i * PL::P;  // factor m, row r, column k at m*64 + r*8 + k (Coef::load)
        if (CF::kCount % PL::P == 0 || idx < CF::kCount)
            coef[(idx >> 6) * 64 + coef_slot((idx >> 3) & 7, idx & 7)] = pre[i];
    }
}

// One mode: its low index bits in lanes L4L5, its high bit in register bit KG; cm = the mode's staged
// factor, cofs = this lane's slot offset ((lane >> 4) & 3) * 16 + (lane & 3) * 4.
// Lane operand B[kb][j] = a[j + 4 rg][kb + 4 kg]; the output row's high bit goes to register bit KG.
template<int KG>
__device__ __forceinline__ void step(double (&v)[8], double const *cm, int cofs)
{
    double2 const lo = *reinterpret_cast<double2 const *>(cm + cofs);
    double2 const hi = *reinterpret_cast<double2 const *>(cm + cofs + 2);
    double const c00 = lo.x, c01 = lo.y, c10 = hi.x, c11 = hi.y;
    constexpr int O0 = KG == 0 ? 1 : 0, O1 = KG == 2 ? 1 : 2;
#pragma unroll
    for (int f = 0; f < 4; ++f)
    {
        int const j0 = ((f & 1) << O0) | (((f >> 1) & 1) << O1);
        int const j1 = j0 | (1 << KG);
        double d0 = __builtin_amdgcn_mfma_f64_4x4x4f64(v[j0], c00, 0.0, 0, 0, 0);
        double d1 = __builtin_amdgcn_mfma_f64_4x4x4f64(v[j0], c10, 0.0, 0, 0, 0);
        d0 = __builtin_amdgcn_mfma_f64_4x4x4f64(v[j1], c01, d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f64_4x4x4f64(v[j1], c11, d1, 0, 0, 0);
        v[j0] = d0;
        v[j1] = d1;
    }
}

// final-layout offset (elements) of register j
template<int D>
constexpr int out_reg(int j)
{
    return M8Plan<D>::reg_part(kF, j);
}

template<int D>
__device__ __forceinline__ void flush(double *dst_generic, double (&acc)[8], int tf)
{
    global_ptr<double> dst = as_global(dst_generic) + tf;
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        atomicAdd((double *)(dst + out_reg<D>(j)), acc[j]);
        acc[j] = 0.0;
    }
}

// ordered accumulation: piece [ps, pe) into its output when it is the whole segment, else into ws[ps]
template<int D>
__device__ __forceinline__ void flush_piece(double *dst_generic, double (&acc)[8], int tf, double *const *out,
                                            double *const *ws, int nb, int ps, int pe)
{
    bool const whole = (ps == 0 || out[ps - 1] != dst_generic) && (pe == nb || out[pe] != dst_generic);
    global_ptr<double> dst = as_global(whole ? dst_generic : ws[ps]) + tf;
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        dst[out_reg<D>(j)] = whole ? dst[out_reg<D>(j)] + acc[j] : acc[j];
        acc[j] = 0.0;
    }
}

// d = 4 (eight wavefronts) keeps the staged input and the exchange in two LDS buffers, and the factors in
// two alternating ones: the next entry's stores then need no barrier of their own (the last readers of
// each buffer are ordered before them by the barriers in between), two barriers per entry instead of four
// (d=4 n=8 batch: 142.4 against 145.5 us per launch). A single buffer under an 80-register cap, three
// workgroups per CU, measured slower (152-158 against 140-148 us).
template<int D>
constexpr bool two_buffers()
{
    return D == 4;
}

template<int D>
constexpr int lds_doubles()
{
    return two_buffers<D>() ? 2 * (M8Plan<D>::SIZE + kronreg::Coef<double, M8Plan<D>>::kCount)
                            : M8Plan<D>::SIZE + kronreg::Coef<double, M8Plan<D>>::kCount;
}

template<int D, bool ORDERED>
__global__ void __launch_bounds__(M8Plan<D>::P)
    kron_m8_kernel(int nb, int epw, double const *const *__restrict__ mats, int stride,
                   double const *const *__restrict__ in, double *const *__restrict__ out,
                   double *const *__restrict__ ws)
{
    using PL = M8Plan<D>;
    using CF = kronreg::Coef<double, PL>;
    constexpr bool k2 = two_buffers<D>();
    __shared__ __attribute__((aligned(16))) double lds[lds_doubles<D>()];
    char *const bytes = reinterpret_cast<char *>(lds);
    char *const xbytes = k2 ? bytes + PL::SIZE * 8 : bytes;  // exchange buffer
    double *const coef0 = lds + (k2 ? 2 : 1) * PL::SIZE;
    int const t = threadIdx.x;
    int const lane = t & 63;
    int const e0 = kronreg::xcd_run_index() * epw;
    int const e1 = (e0 + epw < nb) ? e0 + epw : nb;
    if (e0 >= e1) return;

    unsigned const in_base = static_cast<unsigned>(SwizzleIn::phys(2 * t) * 8);  // granule t of each block
    unsigned const p0_rd = Field<D, kP0, SwizzleIn>::base(t);
    unsigned const p0_wr = Field<D, kP0, SwizzleEx>::base(t);
    unsigned const p1_rd = Field<D, kP1, SwizzleEx>::base(t);
    int const cofs = ((lane >> 4) & 3) * 16 + (lane & 3) * 4;
    int const tf = PL::thread_part(kF, t);

    double acc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = 0.0;
    double *cur = out[e0];
    int ps = e0;

    // entry e from its prefetched registers, which then take entry e + 1 (a second set of prefetch
    // registers, two entries in flight, measured no faster: d=4 153 against 149 us, d=3 14.7 against 14.6)
    auto entry = [&](int e, double (&pre)[8], double (&pre_coef)[CF::kPerThread]) {
        double *const dst = out[e];
        if (dst != cur)
        {
            if constexpr (ORDERED)
                flush_piece<D>(cur, acc, tf, out, ws, nb, ps, e);
            else
                flush<D>(cur, acc, tf);
            cur = dst;
            ps = e;
        }
        double *const coef = coef0 + (k2 ? (e & 1) * CF::kCount : 0);
        if constexpr (!k2) lds_barrier();  // LDS free: every thread has read the previous entry's P1 layout
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            // granule c*P + t: the swizzle is linear and the block bits lie above the thread bits
            unsigned const a = in_base ^ static_cast<unsigned>(SwizzleIn::phys(2 * c * PL::P) * 8);
            double2 g;
            g.x = pre[2 * c];
            g.y = pre[2 * c + 1];
            *reinterpret_cast<double2 *>(bytes + a) = g;
        }
        store_coef<PL>(coef, pre_coef, t);
        lds_barrier();
        int const en = e + 1;
        if (en < e1)
        {
            kronreg::load_input<double, PL>(in[en], pre, t);
            CF::load(mats + static_cast<size_t>(en) * D, stride, pre_coef, t);
        }
        double v[8];
        lds_read<D, kP0, SwizzleIn>(bytes, p0_rd, v);
        step<0>(v, coef + (D - 1) * 64, cofs);
        step<1>(v, coef + (D - 2) * 64, cofs);
        if constexpr (!k2) lds_barrier();  // every thread has read the P0 layout
        lds_write<D, kP0, SwizzleEx>(xbytes, p0_wr, v);
        lds_barrier();
        lds_read<D, kP1, SwizzleEx>(xbytes, p1_rd, v);
        if constexpr (D == 3)
            step<0>(v, coef, cofs);
        else
        {
            step<0>(v, coef + 64, cofs);
            step<1>(v, coef, cofs);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] += v[j];
    };

    double pre[8];
    double pre_coef[CF::kPerThread];
    kronreg::load_input<double, PL>(in[e0], pre, t);
    CF::load(mats + static_cast<size_t>(e0) * D, stride, pre_coef, t);
    for (int e = e0; e < e1; ++e) entry(e, pre, pre_coef);
    if constexpr (ORDERED)
        flush_piece<D>(cur, acc, tf, out, ws, nb, ps, e1);
    else
        flush<D>(cur, acc, tf);
}

} // namespace kronm8
