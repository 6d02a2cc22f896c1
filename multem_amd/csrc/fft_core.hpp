// Register-resident Stockham FFT building blocks for CDNA4.
//
// A length-N sequence is held by T = N/E threads, E complex values each:
// thread t holds x[t + T*e], e = 0..E-1.  Each stage of radix R performs E/R
// butterflies per thread on registers v[q + r*(E/R)], then scatters its
// results through LDS so that, after the exchange, thread t again holds
// x'[t + T*e].  After the last stage the layout is the natural-order
// transform in the same t + T*e pattern, so a forward transform can be
// followed by a point-wise multiply and an inverse transform with no extra
// data movement, and loads/stores from HBM stay coalesced.
//
// Transforms are unnormalised in both directions (forward: exp(-2 pi i nk/N)).
#pragma once

#include <hip/hip_runtime.h>

#include "complex_types.hpp"

namespace mtg {

__device__ __forceinline__ float2 cmake(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ double2 cmake(double x, double y) { return make_double2(x, y); }

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// Complex product with a fixed rounding pattern (the a.y terms rounded, the
// a.x terms fused), so that every kernel instance rounds alike whatever the
// compiler's contraction choices.
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
#pragma clang fp contract(off)
    const float yy = a.y * b.y, yx = a.y * b.x;
    return make_float2(__builtin_fmaf(a.x, b.x, -yy), __builtin_fmaf(a.x, b.y, yx));
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
#pragma clang fp contract(off)
    const double yy = a.y * b.y, yx = a.y * b.x;
    return make_double2(__builtin_fma(a.x, b.x, -yy), __builtin_fma(a.x, b.y, yx));
}
// |a|^2 with a fixed rounding pattern
__device__ __forceinline__ float cnorm2(float2 a)
{
#pragma clang fp contract(off)
    const float yy = a.y * a.y;
    return __builtin_fmaf(a.x, a.x, yy);
}
__device__ __forceinline__ double cnorm2(double2 a)
{
#pragma clang fp contract(off)
    const double yy = a.y * a.y;
    return __builtin_fma(a.x, a.x, yy);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV, class C>
__device__ __forceinline__ C mul_mi(C a)
{
    return INV ? cmake(-a.y, a.x) : cmake(a.y, -a.x);
}

// radix constants written in the precision they are used in
template <class C> struct DftConst;
template <> struct DftConst<float2> {
    static constexpr float h = 0.70710678118654752440f, c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
};
template <> struct DftConst<double2> {
    static constexpr double h = 0.70710678118654752440, c1 = 0.92387953251128675613, s1 = 0.38268343236508977173;
};
template <class C>
__device__ __forceinline__ real_t<C> sqrt_half()
{
    return DftConst<C>::h;
}

template <bool INV, class C>
__device__ __forceinline__ void dft2(C &a0, C &a1)
{
    C t = csub(a0, a1);
    a0 = cadd(a0, a1);
    a1 = t;
}

template <bool INV, class C>
__device__ __forceinline__ void dft4(C &a0, C &a1, C &a2, C &a3)
{
    C s02 = cadd(a0, a2), d02 = csub(a0, a2);
    C s13 = cadd(a1, a3), d13 = mul_mi<INV>(csub(a1, a3));
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = cadd(d02, d13);
    a3 = csub(d02, d13);
}

// fused a + h*b with one rounding, written out so that every kernel instance
// rounds alike whatever the compiler's contraction choices
__device__ __forceinline__ float fmah(float b, float h, float a) { return __builtin_fmaf(b, h, a); }
__device__ __forceinline__ double fmah(double b, double h, double a) { return __builtin_fma(b, h, a); }

// e +/- W8^k d for k = 1 (K3 = false) or 3 (K3 = true), forward (W8 =
// exp(-i pi/4)) or inverse (conjugate): W8^k d = h (sx, sy) with the sums
// below, so each output component is one fused multiply-add.
template <bool INV, bool K3, class C>
__device__ __forceinline__ void w8_butterfly(C e, C d, C &plus, C &minus)
{
#pragma clang fp contract(off)
    using R = real_t<C>;
    const R h = sqrt_half<C>();
    R sx, sy;
    if constexpr (!K3) {
        // forward (1 - i)/sqrt2, inverse (1 + i)/sqrt2
        sx = INV ? d.x - d.y : d.x + d.y;
        sy = INV ? d.x + d.y : d.y - d.x;
    } else {
        // forward (-1 - i)/sqrt2, inverse (-1 + i)/sqrt2
        sx = INV ? -d.x - d.y : d.y - d.x;
        sy = INV ? d.x - d.y : -d.x - d.y;
    }
    plus = cmake(fmah(sx, h, e.x), fmah(sy, h, e.y));
    minus = cmake(fmah(-sx, h, e.x), fmah(-sy, h, e.y));
}

// 8-point DFT of v[0], v[S], ..., v[7S], result in natural order
template <bool INV, int S, int E, class C>
__device__ __forceinline__ void dft8(C (&v)[E], int o)
{
    C e0 = v[o], e1 = v[o + 2 * S], e2 = v[o + 4 * S], e3 = v[o + 6 * S];
    C d0 = v[o + S], d1 = v[o + 3 * S], d2 = v[o + 5 * S], d3 = v[o + 7 * S];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(d0, d1, d2, d3);
    d2 = mul_mi<INV>(d2);
    v[o] = cadd(e0, d0);         v[o + 4 * S] = csub(e0, d0);
    w8_butterfly<INV, false>(e1, d1, v[o + S], v[o + 5 * S]);
    v[o + 2 * S] = cadd(e2, d2); v[o + 6 * S] = csub(e2, d2);
    w8_butterfly<INV, true>(e3, d3, v[o + 3 * S], v[o + 7 * S]);
}

// 16-point DFT as 4 x 4 with internal twiddles W16^(q*kp)
template <bool INV, int S, int E, class C>
__device__ __forceinline__ void dft16(C (&v)[E], int o)
{
    using R = real_t<C>;
    C y[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        y[q][0] = v[o + q * S];
        y[q][1] = v[o + (q + 4) * S];
        y[q][2] = v[o + (q + 8) * S];
        y[q][3] = v[o + (q + 12) * S];
        dft4<INV>(y[q][0], y[q][1], y[q][2], y[q][3]);
    }
    // twiddles W16^(q*kp), q, kp in 1..3
    const R c1 = DftConst<C>::c1, s1 = DftConst<C>::s1, h = DftConst<C>::h;
    const R sg = INV ? R(1) : R(-1);
    const C w1 = cmake(c1, sg * s1);     // W16^1
    const C w2 = cmake(h, sg * h);       // W16^2
    const C w3 = cmake(s1, sg * c1);     // W16^3
    const C w6 = cmake(-h, sg * h);      // W16^6
    const C w9 = cmake(-c1, -sg * s1);   // W16^9
    y[1][1] = cmul(y[1][1], w1); y[1][2] = cmul(y[1][2], w2); y[1][3] = cmul(y[1][3], w3);
    y[2][1] = cmul(y[2][1], w2); y[2][2] = mul_mi<INV>(y[2][2]); y[2][3] = cmul(y[2][3], w6);
    y[3][1] = cmul(y[3][1], w3); y[3][2] = cmul(y[3][2], w6); y[3][3] = cmul(y[3][3], w9);
#pragma unroll
    for (int kp = 0; kp < 4; ++kp) {
        C a0 = y[0][kp], a1 = y[1][kp], a2 = y[2][kp], a3 = y[3][kp];
        dft4<INV>(a0, a1, a2, a3);
        v[o + kp * S] = a0;          // X[4*kq + kp], kq = 0..3
        v[o + (kp + 4) * S] = a1;
        v[o + (kp + 8) * S] = a2;
        v[o + (kp + 12) * S] = a3;
    }
}

// odd radices: constants cos / sin (2 pi k / R) in the precision they are used in
template <class C> struct OddConst;
template <> struct OddConst<float2> {
    static constexpr float s3 = 0.86602540378443864676f;
    static constexpr float c5a = 0.30901699437494742410f, c5b = -0.80901699437494742410f;
    static constexpr float s5a = 0.95105651629515357212f, s5b = 0.58778525229247312917f;
};
template <> struct OddConst<double2> {
    static constexpr double s3 = 0.86602540378443864676;
    static constexpr double c5a = 0.30901699437494742410, c5b = -0.80901699437494742410;
    static constexpr double s5a = 0.95105651629515357212, s5b = 0.58778525229247312917;
};

// 3-point DFT: X0 = a0 + s, X1,2 = (a0 - s/2) -/+ i sin(2pi/3) d (forward;
// signs of the i terms swapped for the inverse), s = a1 + a2, d = a1 - a2.
// Every output component is fixed to one rounding pattern (explicit FMAs).
template <bool INV, int S, int E, class C>
__device__ __forceinline__ void dft3(C (&v)[E], int o)
{
#pragma clang fp contract(off)
    using R = real_t<C>;
    const R h = OddConst<C>::s3;
    const C a0 = v[o], a1 = v[o + S], a2 = v[o + 2 * S];
    const C s = cadd(a1, a2), d = csub(a1, a2);
    const C m = cmake(fmah(s.x, R(-0.5), a0.x), fmah(s.y, R(-0.5), a0.y));
    // -i d (forward) or +i d (inverse), scaled by sin(2pi/3) inside the FMA
    const C e = mul_mi<INV>(d);
    v[o] = cadd(a0, s);
    v[o + S] = cmake(fmah(e.x, h, m.x), fmah(e.y, h, m.y));
    v[o + 2 * S] = cmake(fmah(-e.x, h, m.x), fmah(-e.y, h, m.y));
}

// 5-point DFT with s14 = a1 + a4, d14 = a1 - a4, s23 = a2 + a3, d23 = a2 - a3:
//   X0 = a0 + s14 + s23
//   X1,4 = m1 -/+ i n1,  m1 = a0 + c1 s14 + c2 s23,  n1 = s1 d14 + s2 d23
//   X2,3 = m2 -/+ i n2,  m2 = a0 + c2 s14 + c1 s23,  n2 = s2 d14 - s1 d23
// (c_k = cos(2 pi k / 5), s_k = sin(2 pi k / 5); forward signs, swapped for
// the inverse).  Explicit FMAs fix the rounding.
template <bool INV, int S, int E, class C>
__device__ __forceinline__ void dft5(C (&v)[E], int o)
{
#pragma clang fp contract(off)
    using R = real_t<C>;
    const R c1 = OddConst<C>::c5a, c2 = OddConst<C>::c5b, s1 = OddConst<C>::s5a, s2 = OddConst<C>::s5b;
    const C a0 = v[o], a1 = v[o + S], a2 = v[o + 2 * S], a3 = v[o + 3 * S], a4 = v[o + 4 * S];
    const C s14 = cadd(a1, a4), d14 = csub(a1, a4), s23 = cadd(a2, a3), d23 = csub(a2, a3);
    const C m1 = cmake(fmah(s23.x, c2, fmah(s14.x, c1, a0.x)), fmah(s23.y, c2, fmah(s14.y, c1, a0.y)));
    const C m2 = cmake(fmah(s23.x, c1, fmah(s14.x, c2, a0.x)), fmah(s23.y, c1, fmah(s14.y, c2, a0.y)));
    const C n1 = cmake(fmah(d23.x, s2, d14.x * s1), fmah(d23.y, s2, d14.y * s1));
    const C n2 = cmake(fmah(d23.x, -s1, d14.x * s2), fmah(d23.y, -s1, d14.y * s2));
    const C e1 = mul_mi<INV>(n1), e2 = mul_mi<INV>(n2);   // -i n (forward), +i n (inverse)
    v[o] = cadd(cadd(a0, s14), s23);
    v[o + S] = cadd(m1, e1);
    v[o + 4 * S] = csub(m1, e1);
    v[o + 2 * S] = cadd(m2, e2);
    v[o + 3 * S] = csub(m2, e2);
}

template <int R, bool INV, int S, int E, class C>
__device__ __forceinline__ void dftR(C (&v)[E], int o)
{
    static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 16, "radix");
    if constexpr (R == 2) dft2<INV>(v[o], v[o + S]);
    else if constexpr (R == 3) dft3<INV, S, E>(v, o);
    else if constexpr (R == 4) dft4<INV>(v[o], v[o + S], v[o + 2 * S], v[o + 3 * S]);
    else if constexpr (R == 5) dft5<INV, S, E>(v, o);
    else if constexpr (R == 8) dft8<INV, S, E>(v, o);
    else if constexpr (R == 16) dft16<INV, S, E>(v, o);
}

// Radix schedule of a length-N transform: E values per thread, NST stages
// of radices R[0..NST-1] (each dividing E, product N).
template <int N_, int E_, int R0, int R1, int R2 = 0, int R3 = 0>
struct Radix {
    static constexpr int N = N_, E = E_, T = N_ / E_;
    static constexpr int NST = (R3 ? 4 : R2 ? 3 : 2);
    static constexpr int R[4] = {R0, R1, R2, R3};
    static_assert((R0 * R1 * (R2 ? R2 : 1) * (R3 ? R3 : 1)) == N_, "radices must multiply to N");
    static_assert(E_ % R0 == 0 && E_ % R1 == 0 && (R2 == 0 || E_ % R2 == 0) && (R3 == 0 || E_ % R3 == 0),
                  "each radix divides E");
};

// Schedules per length: 0 is the default, 1 the alternative the tuner may pick.
// Lengths 2^k, 3 x 2^k and 5 x 2^k in [64, 2048]; the odd radix runs last.
template <int N, int SC> struct Sched;
template <> struct Sched<64, 0>   : Radix<64, 8, 8, 8> {};
template <> struct Sched<64, 1>   : Radix<64, 16, 16, 4> {};
template <> struct Sched<128, 0>  : Radix<128, 8, 8, 8, 2> {};
template <> struct Sched<128, 1>  : Radix<128, 16, 16, 8> {};
template <> struct Sched<256, 0>  : Radix<256, 8, 8, 8, 4> {};
template <> struct Sched<256, 1>  : Radix<256, 16, 16, 16> {};
template <> struct Sched<512, 0>  : Radix<512, 8, 8, 8, 8> {};
template <> struct Sched<512, 1>  : Radix<512, 16, 16, 8, 4> {};
template <> struct Sched<1024, 0> : Radix<1024, 16, 16, 8, 8> {};
template <> struct Sched<1024, 1> : Radix<1024, 8, 8, 8, 8, 2> {};
template <> struct Sched<2048, 0> : Radix<2048, 16, 16, 16, 8> {};
template <> struct Sched<2048, 1> : Radix<2048, 8, 8, 8, 8, 4> {};
template <> struct Sched<96, 0>   : Radix<96, 12, 4, 4, 2, 3> {};
template <> struct Sched<96, 1>   : Radix<96, 24, 8, 4, 3> {};
template <> struct Sched<192, 0>  : Radix<192, 12, 4, 4, 4, 3> {};
template <> struct Sched<192, 1>  : Radix<192, 24, 8, 8, 3> {};
template <> struct Sched<384, 0>  : Radix<384, 24, 8, 4, 4, 3> {};
template <> struct Sched<384, 1>  : Radix<384, 48, 16, 8, 3> {};
template <> struct Sched<768, 0>  : Radix<768, 24, 8, 8, 4, 3> {};
template <> struct Sched<768, 1>  : Radix<768, 48, 16, 16, 3> {};
template <> struct Sched<1536, 0> : Radix<1536, 24, 8, 8, 8, 3> {};
template <> struct Sched<1536, 1> : Radix<1536, 48, 16, 16, 2, 3> {};
template <> struct Sched<80, 0>   : Radix<80, 20, 4, 4, 5> {};
template <> struct Sched<80, 1>   : Radix<80, 40, 8, 2, 5> {};
template <> struct Sched<160, 0>  : Radix<160, 20, 4, 4, 2, 5> {};
template <> struct Sched<160, 1>  : Radix<160, 40, 8, 4, 5> {};
template <> struct Sched<320, 0>  : Radix<320, 20, 4, 4, 4, 5> {};
template <> struct Sched<320, 1>  : Radix<320, 40, 8, 8, 5> {};
template <> struct Sched<640, 0>  : Radix<640, 40, 8, 4, 4, 5> {};
template <> struct Sched<640, 1>  : Radix<640, 40, 8, 8, 2, 5> {};
template <> struct Sched<1280, 0> : Radix<1280, 40, 8, 8, 4, 5> {};
template <> struct Sched<1280, 1> : Radix<1280, 40, 4, 8, 8, 5> {};

template <class S, int ST>
struct StageNS {
    static constexpr int value = StageNS<S, ST - 1>::value * S::R[ST - 1];
};
template <class S>
struct StageNS<S, 0> { static constexpr int value = 1; };

// One Stockham stage ST of schedule S.  tw.get<ST>(q, r) is this thread's
// twiddle W_{NS R}^{r k} of butterfly q (forward sign; conjugated here for
// the inverse).
template <class S, int ST, bool INV, class TW, class C>
__device__ __forceinline__ void fft_stage(C (&v)[S::E], const TW &tw)
{
    constexpr int R = S::R[ST], Q = S::E / R;
    constexpr int NS = StageNS<S, ST>::value;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C w = tw.template get<ST>(q, r);
                if (INV) w.y = -w.y;
                v[q + r * Q] = cmul(v[q + r * Q], w);
            }
        }
        dftR<R, INV, Q, S::E>(v, q);
    }
}

// LDS index written by element (q, r) of stage ST: output j/NS*NS*R + j%NS + r*NS
// of butterfly j = t + T*q (Stockham autosort).
template <class S, int ST>
__device__ __forceinline__ int stage_out_index(int t, int q, int r)
{
    constexpr int R = S::R[ST], NS = StageNS<S, ST>::value;
    const unsigned j = t + S::T * q;
    return (j / NS) * NS * R + (j % NS) + r * NS;
}

// Per-thread twiddles of all stages with NS > 1, in stage order.
template <class S, int ST>
struct TwOffset {
    static constexpr int prev = (StageNS<S, ST - 1>::value > 1) ? (S::E / S::R[ST - 1]) * (S::R[ST - 1] - 1) : 0;
    static constexpr int value = TwOffset<S, ST - 1>::value + prev;
};
template <class S>
struct TwOffset<S, 0> { static constexpr int value = 0; };

// Twiddles held in registers: loaded once per thread, reused by the forward
// and the inverse transform.
template <class S, class C = float2>
struct TwReg {
    static constexpr int COUNT = TwOffset<S, S::NST>::value;
    C w[COUNT > 0 ? COUNT : 1];

    // tw: W_N^m = exp(-2 pi i m / N), m in [0, N); t: thread index in the sequence
    template <int ST = 0>
    __device__ __forceinline__ void load(int t, const C *__restrict__ tw)
    {
        constexpr int R = S::R[ST], NS = StageNS<S, ST>::value;
        if constexpr (NS > 1) {
            constexpr int Q = S::E / R, off = TwOffset<S, ST>::value;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int k = (unsigned)(t + S::T * q) % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) w[off + q * (R - 1) + r - 1] = tw[r * k * (S::N / (NS * R))];
            }
        }
        if constexpr (ST + 1 < S::NST) load<ST + 1>(t, tw);
    }
    template <int ST>
    __device__ __forceinline__ C get(int q, int r) const
    {
        return w[TwOffset<S, ST>::value + q * (S::R[ST] - 1) + r - 1];
    }
};

// Twiddles read from the cache-resident table at each use: fewer registers.
template <class S, class C = float2>
struct TwTable {
    const C *__restrict__ tab;
    int t;
    __device__ __forceinline__ void load(int t_, const C *__restrict__ tw)
    {
        t = t_;
        tab = tw;
    }
    template <int ST>
    __device__ __forceinline__ C get(int q, int r) const
    {
        constexpr int R = S::R[ST], NS = StageNS<S, ST>::value;
        return tab[r * ((unsigned)(t + S::T * q) % NS) * (S::N / (NS * R))];
    }
};

// Full transform.  lds_idx(i) maps a sequence index to this sequence's LDS
// slot; sync() is the barrier that orders LDS writes and reads of all threads
// sharing the buffer.  The data starts and ends in the t + T*e layout.
template <class S, bool INV, int ST = 0, class TW, class LdsIdx, class Sync, class C>
__device__ __forceinline__ void fft_run(C (&v)[S::E], int t, const TW &tw, C *lds, LdsIdx lds_idx, Sync sync)
{
    fft_stage<S, ST, INV>(v, tw);
    if constexpr (ST + 1 < S::NST) {
        constexpr int R = S::R[ST], Q = S::E / R;
        sync();   // previous readers of the buffer are done
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int r = 0; r < R; ++r)
                lds[lds_idx(stage_out_index<S, ST>(t, q, r))] = v[q + r * Q];
        sync();
#pragma unroll
        for (int e = 0; e < S::E; ++e) v[e] = lds[lds_idx(t + S::T * e)];
        fft_run<S, INV, ST + 1>(v, t, tw, lds, lds_idx, sync);
    }
}

} // namespace mtg
