// kron_plan.hpp -- data layout of the register-resident sweep, shared by the HIP kernels and the CPU
// emulator (tests/cpp/plan_check.cpp). Plain constexpr C++: in HIP these functions are usable on host and
// device alike.
//
// The n^d vector of one batch entry is a tensor with d modes of size n; mode 0 is the slowest index
// (canonical element e = sum_t k_t n^(d-1-t)), as in the reference (tests/utils/kronmult_naive.h:99-101).
// A workgroup of P threads holds the whole tensor, R = n^RHO values per thread. The sweep runs in phases;
// in phase p a thread's registers hold RHO consecutive modes [lo(p), lo(p)+RHO) and the remaining modes
// index the thread. Phase p contracts the modes of its block that are due, from the fastest down, which
// keeps the reference order (factor d-1 first, kronmult_gpu/kronmult.cu:104). Between phases the tensor
// goes through LDS in canonical order (with an XOR swizzle against bank conflicts).
#pragma once

namespace kronplan
{

constexpr int ipow(int b, int e)
{
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

constexpr int ilog2(int v)
{
    int r = 0;
    while ((1 << r) < v) ++r;
    return r;
}

template<int D, int N, int RHO>
struct Plan
{
    static_assert(RHO >= 1 && RHO <= D, "register block must hold 1..d modes");
    static constexpr int kD = D;
    static constexpr int kN = N;
    static constexpr int kRho = RHO;
    static constexpr int SIZE = ipow(N, D);          // elements per entry
    static constexpr int R = ipow(N, RHO);           // values per thread
    static constexpr int P = SIZE / R;               // threads per entry
    static constexpr int NPHASE = (D + RHO - 1) / RHO;

    // first mode of the register block of phase p
    static constexpr int lo(int p) { return D - (p + 1) * RHO < 0 ? 0 : D - (p + 1) * RHO; }
    // modes contracted in phase p: c_hi(p)-1 down to lo(p)
    static constexpr int c_hi(int p) { return D - p * RHO; }
    // number of elements spanned by the modes faster than the register block
    static constexpr int L(int p) { return ipow(N, D - lo(p) - RHO); }
    // canonical element held by thread t in register j during phase p
    static constexpr int elem(int p, int t, int j)
    {
        return (t / L(p)) * (R * L(p)) + j * L(p) + (t % L(p));
    }
    // stride, in the register index, of register-block mode m (0 = slowest of the block)
    static constexpr int reg_stride(int m) { return ipow(N, RHO - 1 - m); }
};

// LDS placement of canonical element e. The tensor is stored in 16-byte granules (G = 16 / sizeof(T)
// elements); granule index g goes to g ^ x(g), where bit i (i < 4) of x is bit src_i of g. gfx950's LDS
// serves 256 bytes per cycle from 64 four-byte banks, i.e. 16 granules, so the four low granule bits
// select the bank group. The sources per plan come from an exhaustive search over single-bit XOR maps
// under that bank model (every plan below is conflict-free for all of its LDS accesses; the CPU check
// tests/cpp/plan_check.cpp verifies it). Code layout: 5 bits per target bit, value = source bit + 1,
// 0 = leave the bit alone.
constexpr unsigned swizzle_code(int d, int n, int rho, int elem_bytes)
{
    return (d == 6 && n == 4 && rho == 2) ? 0x41cc5u
         : (d == 5 && n == 4 && rho == 2) ? 0x41cc5u
         : (d == 4 && n == 4 && rho == 1) ? (elem_bytes == 8 ? 0x318a5u : 0x314c5u)
         : (d == 4 && n == 8 && rho == 1) ? (elem_bytes == 8 ? 0x398c5u : 0x398a5u)
         : (d == 3 && n == 8 && rho == 1) ? (elem_bytes == 8 ? 0x398c5u : 0x398a5u)
         : (d == 4 && n == 8 && rho == 2) ? (elem_bytes == 8 ? 0x4a0e6u : 0x41cc5u)
         : (d == 6 && n == 4 && rho == 3) ? (elem_bytes == 8 ? 0x4a0e6u : 0x41cc5u)
         : (d == 3 && n == 8 && rho == 2) ? (elem_bytes == 8 ? 0x020e6u : 0x01cc5u)
         : 0u;
}

template<unsigned CODE, int G>
struct Swizzle
{
    static constexpr int phys(int e)
    {
        int const g = e / G;
        int x = 0;
        for (int i = 0; i < 4; ++i)
        {
            int const src = static_cast<int>((CODE >> (5 * i)) & 31u);
            if (src) x |= ((g >> (src - 1)) & 1) << i;
        }
        return (g ^ x) * G + e % G;
    }
};

template<typename PL, typename T>
using PlanSwizzle = Swizzle<swizzle_code(PL::kD, PL::kN, PL::kRho, static_cast<int>(sizeof(T))),
                            static_cast<int>(16 / sizeof(T))>;

} // namespace kronplan
