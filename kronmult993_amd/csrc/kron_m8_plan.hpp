// kron_m8_plan.hpp -- data layout of the matrix-core sweep for n = 8, FP64, d = 3 and 4 (kron_m8.hpp),
// shared by the HIP kernel and the CPU checker (tests/cpp/plan_check.cpp). Plain constexpr C++.
//
// One workgroup of P = 8^(d-1) threads (one wavefront for d = 3, eight for d = 4) holds one entry's 8^d
// values, 8 per thread. Element e of the canonical order (mode 0 slowest, as in kron_plan.hpp) has 3d bits;
// mode t owns bits [3(d-1-t), 3(d-1-t)+3), so mode d-1 is e0..e2. A thread is (wave w, lane l), its values
// are registers j = 0..7; the bits of l are L0..L5, of j J0..J2, of w W0..W2.
//
// V_MFMA_F64_4X4X4_4B_F64 (measured with in-bounds probes): lane 16 kb + 4 jb + i supplies A_jb[i][kb],
// lane 16 kb + 4 jb + j supplies B_jb[kb][j]; D_jb[i][j] = C_jb[i][j] + sum_kb A_jb[i][kb] B_jb[kb][j]
// lands in lane 16 i + 4 jb + j, computed as the fused chain kb = 0, 1, 2, 3 from C (bit for bit). With
// the data as A (k = mode index) and the factor as B (B[kb][j] = a[r = j][k = kb]), one contraction step
// reads the mode's low index bits from L4L5 and its high bit from one register bit (two MFMAs chained,
// k = 0..3 then 4..7), and leaves the output index r in L0L1 (high bit in the same register bit), while
// the two bits that were in L0L1 move to L4L5. Two steps restore the layout:
//
//   P0 (modes d-1, d-2):  e0 L4  e1 L5  e2 J0 | e3 L0  e4 L1  e5 J1 | e6 L2  e7 L3  e8 J2 | e9..e11 W0..W2
//   step on mode d-1 (high bit J0), step on mode d-2 (high bit J1): P0 again.
//   LDS exchange into P1:
//   d = 3  P1:  e6 L4  e7 L5  e8 J0 | e0 L0  e1 L1  e2 L2  e3 L3  e4 J1  e5 J2
//          F (after the step on mode 0, high bit J0):  e6 L0  e7 L1  e8 J0 | e0 L4  e1 L5  e2 L2  e3 L3 ...
//   d = 4  P1:  e6 L4  e7 L5  e8 J0 | e9 L0  e10 L1  e11 J1 | e0 L2  e1 L3  e2 J2  e3..e5 W0..W2
//          step on mode 1 (J0), step on mode 0 (J1):
//          F:   e9 L0  e10 L1  e11 J1 | e6 L4  e7 L5  e8 J0 | e0 L2  e1 L3  e2 J2  e3..e5 W0..W2
//
// Every output value is the reference's fused chain over k ascending from zero (kronmult_gpu/kronmult.cu:
// 66-70), modes contracted from d-1 down to 0 (kronmult.cu:104), so the sweep is bit-identical to the
// register-resident one.
//
// LDS: the tensor is stored in canonical order, 16-byte granules g = e / 2, granule g at g ^ x(g) with
// x_i = parity(g & mask_i) (the Swizzle of kron_plan.hpp). Stage 0 (input store, b128 per thread, then
// P0 reads) and stage 1 (P0 writes, P1 reads) have their own masks; checked against the bank model of
// gfx950 in tests/cpp/plan_check.cpp (the P0 writes hit each bank twice per 16-lane group: e0 sits in a
// lane bit, so a 16-byte row is never written whole by one group -- the minimum for that layout).
#pragma once

#include "kron_plan.hpp"

namespace kronm8
{

// positions: 0..5 lane bits L0..L5, 6..8 register bits J0..J2, 9..11 wave bits W0..W2
constexpr int kL0 = 0, kL1 = 1, kL2 = 2, kL3 = 3, kL4 = 4, kL5 = 5, kJ0 = 6, kJ1 = 7, kJ2 = 8, kW0 = 9, kW1 = 10,
              kW2 = 11;

enum Stage
{
    kP0 = 0,
    kP1 = 1,
    kF = 2
};

template<int D>
struct M8Plan
{
    static_assert(D == 3 || D == 4, "matrix-core n = 8 plans exist for d = 3 and 4");
    static constexpr int kD = D;
    static constexpr int kN = 8;
    static constexpr int SIZE = 1 << (3 * D);  // elements per entry
    static constexpr int R = 8;                // values per thread
    static constexpr int P = SIZE / R;         // threads per entry
    static constexpr int WAVES = P / 64;
    static constexpr int BITS = 3 * D;

    // source position of element bit b in layout s
    static constexpr int pos(int s, int b)
    {
        constexpr int p0[12] = {kL4, kL5, kJ0, kL0, kL1, kJ1, kL2, kL3, kJ2, kW0, kW1, kW2};
        constexpr int p1_3[9] = {kL0, kL1, kL2, kL3, kJ1, kJ2, kL4, kL5, kJ0};
        constexpr int f_3[9] = {kL4, kL5, kL2, kL3, kJ1, kJ2, kL0, kL1, kJ0};
        constexpr int p1_4[12] = {kL2, kL3, kJ2, kW0, kW1, kW2, kL4, kL5, kJ0, kL0, kL1, kJ1};
        constexpr int f_4[12] = {kL2, kL3, kJ2, kW0, kW1, kW2, kL4, kL5, kJ0, kL0, kL1, kJ1};
        return s == kP0 ? p0[b] : (D == 3 ? (s == kP1 ? p1_3[b] : f_3[b]) : (s == kP1 ? p1_4[b] : f_4[b]));
    }

    // canonical element of thread part (wave w, lane l) and register part j, separately: the layout maps
    // disjoint bit fields, so elem(s, t, j) = thread_part(s, t) | reg_part(s, j)
    static constexpr int thread_part(int s, int t)
    {
        int e = 0;
        for (int b = 0; b < BITS; ++b)
        {
            int const p = pos(s, b);
            int const bit = p < 6 ? (t >> p) & 1 : (p >= 9 ? (t >> (6 + p - 9)) & 1 : 0);
            e |= bit << b;
        }
        return e;
    }
    static constexpr int reg_part(int s, int j)
    {
        int e = 0;
        for (int b = 0; b < BITS; ++b)
        {
            int const p = pos(s, b);
            if (p >= 6 && p <= 8) e |= ((j >> (p - 6)) & 1) << b;
        }
        return e;
    }
    static constexpr int elem(int s, int t, int j) { return thread_part(s, t) | reg_part(s, j); }
};

// XOR masks of the two LDS round trips (granule bits; x_i = parity(g & mask_i), mask_i in bits [16 i, 16 i + 16))
constexpr unsigned long long kM8MasksStage0 = (1ull << 5) | ((1ull << 6) << 16);
constexpr unsigned long long kM8MasksStage1 =
    (1ull << 5) | (((1ull << 8) | (1ull << 6)) << 16) | ((1ull << 9) << 32) | ((1ull << 5) << 48);

using SwizzleIn = kronplan::Swizzle<kM8MasksStage0, 2>;
using SwizzleEx = kronplan::Swizzle<kM8MasksStage1, 2>;

} // namespace kronm8
