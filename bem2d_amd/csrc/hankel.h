// H0^(1)(x), H1^(1)(x) of real x > 0 in FP64, for the device and the host.
//
// Replaces the reference's BesselH0/BesselH1(double) = GSL J + iY (lib/bem2d_mathroutines.cpp:20-42).
// Three regimes, coefficients in hankel_tables.h (written by gen_hankel_tables.py from mpmath):
//   x < 2    ascending series in t = x^2/4 (A&S 9.1.10 and 9.1.11, n = 0, 1);
//   x >= 2   H_nu = sqrt(2/(pi x)) (P_nu + i Q_nu) exp(i(x - (2nu+1) pi/4)) (A&S 9.2.5-9.2.6),
//            P_nu and x Q_nu as polynomials of z(y), y = (a/x)^2, on the pieces [2, 2.6),
//            [2.6, 3.5), [3.5, 5), [5, 8) (degree 12) and [8, inf) (degree 10), and for
//            x >= 40 (the asymptotic regime where most far-field pairs fall) degree-4
//            polynomials in w = 1/x^2.
// On the device every coefficient lives in one __constant__ table and every index is a
// compile-time constant, so the coefficients reach v_fma_f64 as SGPR operands through scalar loads
// (instruction literals would each cost two v_mov_b32 per use).  The last piece is evaluated in
// w = 1/x^2 directly (hankel_pq_far).  The phase x - pi/4 is reduced by whole turns of pi
// (sincos_half_turn), the sign (-1)^n of the reduction folded into the amplitude.  Accuracy: about
// 3e-16 relative to |H| (pinned against mpmath by the tests).
#pragma once

#include <hip/hip_runtime.h>

#include "hankel_tables.h"

namespace bem2d_dev {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoOverPi = 0.63661977236758134308;
constexpr double kInvPi = 0.31830988618379067154;
constexpr double kHankelFarX = BEM2D_HANKEL_FAR_X;                      // start of the far table
constexpr double kHankelMidX = BEM2D_PIECE_X[BEM2D_HANKEL_NPIECE - 1]; // start of the last piece

struct HankelTables {
    double ser[4][BEM2D_HANKEL_NSER];                          // J0, J1, Y0, Y1 series in t
    double zmap[BEM2D_HANKEL_NPIECE][3];                       // a^2, ym, 1/yh per piece
    double fit[BEM2D_HANKEL_NPIECE][4][BEM2D_HANKEL_DEGMAX + 1]; // P0, xQ0, P1, xQ1 per piece
    double far[4][BEM2D_HANKEL_FAR_DEG + 1];                   // last piece, in w = 1/x^2
    // 2/pi, pi/2 = P1 + P2 + P3, then S (degree 5) and C (degree 6) of
    // sin r = r + r^3 S(r^2), cos r = 1 + r^2 C(r^2), highest power first
    double trig[4 + BEM2D_TRIG_NS + BEM2D_TRIG_NC];
    // 1/pi, pi = P1 + P2 + P3, then S and C (degree 7 each) of the same on |r| <= pi/2
    double trig2[4 + BEM2D_TRIG2_NS + BEM2D_TRIG2_NC];
};

static constexpr HankelTables kHankelHost = {BEM2D_HANKEL_SER_INIT, BEM2D_HANKEL_ZMAP_INIT,
                                             BEM2D_HANKEL_FIT_INIT, BEM2D_HANKEL_FAR_INIT,
                                             BEM2D_TRIG_INIT, BEM2D_TRIG2_INIT};
static __constant__ HankelTables kHankelDev = {BEM2D_HANKEL_SER_INIT, BEM2D_HANKEL_ZMAP_INIT,
                                               BEM2D_HANKEL_FIT_INIT, BEM2D_HANKEL_FAR_INIT,
                                               BEM2D_TRIG_INIT, BEM2D_TRIG2_INIT};

#ifdef __HIP_DEVICE_COMPILE__
// Device view of the table: a constant-address-space (4) pointer, so that uniform reads become
// scalar loads.  The empty asm makes the pointer opaque at every use, which keeps the loads next to
// their use inside the quadrature loops -- hoisted out, the ~60 coefficients would exceed the SGPR
// file and be spilled to VGPR lanes.
typedef __attribute__((address_space(4))) const HankelTables DevTables;
__device__ __forceinline__ DevTables& hankel_tables()
{
    DevTables* p = (DevTables*)&kHankelDev;
    asm volatile("" : "+s"(p));
    return *p;
}

// a * b + c with c read from an SGPR pair: the VOP3 form takes the coefficient as an operand,
// where the compiler's two-address v_fmac_f64 would first copy it into VGPRs (two v_mov_b32).
__device__ __forceinline__ double fma_vvs(double a, double b, double c)
{
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

// a + c with c read from an SGPR pair, kept apart from the product that feeds it (a contracted
// fma would need two scalar operands, one more than a VALU instruction reads).
__device__ __forceinline__ double add_vs(double a, double c)
{
    double r;
    asm("v_add_f64 %0, %1, %2" : "=v"(r) : "v"(a), "s"(c));
    return r;
}
#else
__host__ __forceinline__ const HankelTables& hankel_tables() { return kHankelHost; }
__host__ __forceinline__ double fma_vvs(double a, double b, double c) { return fma(a, b, c); }
__host__ __forceinline__ double add_vs(double a, double c) { return a + c; }
#endif

// Horner's rule from the top, its first step as a multiply and an add so that no coefficient has
// to be copied into vector registers.
template <int N, typename Ptr>
__host__ __device__ __forceinline__ double horner(Ptr c, double z)
{
    double r = add_vs(z * c[N - 1], c[N - 2]);
#pragma unroll
    for (int i = N - 3; i >= 0; --i) r = fma_vvs(r, z, c[i]);
    return r;
}

// Leading operands of the phase and P/Q polynomials.  An instruction reads at most one SGPR
// operand, so a first Horner step c_top z + c_next with both coefficients from the table costs a
// multiply and an add (horner above); the quadrature loops hold these few values in VGPRs
// (hankel_tops_vgpr) and pay one FMA per polynomial instead, and the reduction's -1/4 no longer
// has to be copied into VGPRs at every use.
struct HankelTops {
    double mquarter; // -1/4 of the half-turn reduction
    double trig[2];  // leading coefficients of the half-turn S and C
    double pq[4];    // leading coefficients of P0, xQ0, P1, xQ1 of one table
};

// TABLE: a piece 0 .. BEM2D_HANKEL_NPIECE - 1, or BEM2D_HANKEL_NPIECE for the far table
template <int TABLE>
__host__ __device__ __forceinline__ HankelTops hankel_tops()
{
    const auto& T = hankel_tables();
    HankelTops h;
    h.mquarter = -0.25;
    h.trig[0] = T.trig2[4];
    h.trig[1] = T.trig2[4 + BEM2D_TRIG2_NS];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if constexpr (TABLE == BEM2D_HANKEL_NPIECE)
            h.pq[i] = T.far[i][BEM2D_HANKEL_FAR_DEG];
        else
            h.pq[i] = T.fit[TABLE][i][BEM2D_PIECE_DEG[TABLE]];
    }
    return h;
}

// The same values in VGPRs: the empty asm hides that they are uniform, so they stay in vector
// registers across the loop that uses them.
template <int TABLE>
__host__ __device__ __forceinline__ HankelTops hankel_tops_vgpr()
{
    HankelTops h = hankel_tops<TABLE>();
#ifdef __HIP_DEVICE_COMPILE__
    asm volatile("" : "+v"(h.mquarter), "+v"(h.trig[0]), "+v"(h.trig[1]));
    asm volatile("" : "+v"(h.pq[0]), "+v"(h.pq[1]), "+v"(h.pq[2]), "+v"(h.pq[3]));
#endif
    return h;
}

// Horner's rule with the leading coefficient given apart (top z + c[N - 2] as one FMA).
template <int N, typename Ptr>
__host__ __device__ __forceinline__ double horner_top(Ptr c, double top, double z)
{
    double r = fma_vvs(z, top, c[N - 2]);
#pragma unroll
    for (int i = N - 3; i >= 0; --i) r = fma_vvs(r, z, c[i]);
    return r;
}

// 1/a and 1/sqrt(a) for finite a > 0: the hardware estimates (V_RCP_F64 / V_RSQ_F64, within 2^29
// ulp = 2^-23 relative) refined by one third-order step to within ~1 ulp.  With e = 1 - a y
// (resp. 1 - a y^2), |e| <= 2^-22: 1/a = y (1 + e + e^2 + ...) and 1/sqrt(a) = y (1 + e/2 +
// 3e^2/8 + ...); the first omitted terms, e^3 and 5/16 e^3, are below 2^-65.  On the host, the C
// library.
__host__ __device__ __forceinline__ double rcp_refined(double a)
{
#ifdef __HIP_DEVICE_COMPILE__
    const double y = __builtin_amdgcn_rcp(a);
    const double e = fma(-a, y, 1.0);
    return fma(y, fma(e, e, e), y);
#else
    return 1.0 / a;
#endif
}

__host__ __device__ __forceinline__ double rsqrt_refined(double a)
{
#ifdef __HIP_DEVICE_COMPILE__
    const double y = __builtin_amdgcn_rsq(a);
    const double e = fma(-(a * y), y, 1.0); // 1 - a y^2
    return fma(y, e * fma(e, 0.375, 0.5), y);
#else
    return 1.0 / sqrt(a);
#endif
}

// sin r and cos r on |r| <= pi/4 from the near-minimax polynomials of the table (interpolated at
// Chebyshev nodes in r^2: 1.4e-17 relative on (sin r - r)/r^3, 1.2e-19 absolute on (cos r - 1)/r^2),
// then the quadrant q of the reduction applied.
__host__ __device__ __forceinline__ void sincos_kernel(double r, int q, double& s, double& c)
{
    const auto& T = hankel_tables().trig;
    constexpr int S0 = 4, C0 = 4 + BEM2D_TRIG_NS;
    const double r2 = r * r;
    double ps = add_vs(r2 * T[S0], T[S0 + 1]);
#pragma unroll
    for (int i = S0 + 2; i < C0; ++i) ps = fma_vvs(ps, r2, T[i]);
    const double sr = fma(ps * r2, r, r);
    double pc = add_vs(r2 * T[C0], T[C0 + 1]);
#pragma unroll
    for (int i = C0 + 2; i < C0 + BEM2D_TRIG_NC; ++i) pc = fma_vvs(pc, r2, T[i]);
    const double cr = fma(pc, r2, 1.0);
    const double a = (q & 1) ? cr : sr; // sin before the sign
    const double b = (q & 1) ? sr : cr; // cos before the sign
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}

// sin and cos of x for |x| < 2^30: Cody-Waite reduction by pi/2 split into three doubles
// (pi/2 = P1 + P2 + P3 to 160 bits), then sincos_kernel.  Both results within about 1 ulp.
__host__ __device__ __forceinline__ void sincos_reduced(double x, double& s, double& c)
{
    const auto& T = hankel_tables().trig;
    const double n = rint(x * T[0]);
    double r = fma(-n, T[1], x);
    r = fma(-n, T[2], r);
    r = fma(-n, T[3], r);
    sincos_kernel(r, ((int)n) & 3, s, c);
}

// Series regime, x < 2.
__host__ __device__ __forceinline__ void hankel01_series(double x, double& h0r, double& h0i,
                                                         double& h1r, double& h1i)
{
    const auto& T = hankel_tables();
    const double t = 0.25 * x * x;
    const double hx = 0.5 * x;
    const double lg = log(hx);
    const double j0 = horner<BEM2D_HANKEL_NSER>(T.ser[0], t);
    const double j1 = hx * horner<BEM2D_HANKEL_NSER>(T.ser[1], t);
    const double y0 = kTwoOverPi * lg * j0 + horner<BEM2D_HANKEL_NSER>(T.ser[2], t);
    const double y1 = kTwoOverPi * lg * j1 - kTwoOverPi * rcp_refined(x) +
                      hx * horner<BEM2D_HANKEL_NSER>(T.ser[3], t);
    h0r = j0; h0i = y0; h1r = j1; h1i = y1;
}

// P_nu, Q_nu of piece P at 1/x; h = hankel_tops<P>() (from the table or held in VGPRs).
template <int P>
__host__ __device__ __forceinline__ void hankel_pq(double invx, double& P0, double& Q0, double& P1,
                                                   double& Q1, const HankelTops& h)
{
    const auto& T = hankel_tables();
    constexpr int N = BEM2D_PIECE_DEG[P] + 1;
    const double y = T.zmap[P][0] * invx * invx;
    const double z = (y - T.zmap[P][1]) * T.zmap[P][2];
    P0 = horner_top<N>(T.fit[P][0], h.pq[0], z);
    Q0 = horner_top<N>(T.fit[P][1], h.pq[1], z) * invx;
    P1 = horner_top<N>(T.fit[P][2], h.pq[2], z);
    Q1 = horner_top<N>(T.fit[P][3], h.pq[3], z) * invx;
}

// sin r and cos r of x - pi/4 = r + n pi, |r| <= pi/2, for 0 <= x < 2^30: Cody-Waite reduction by
// pi = P1 + P2 + P3 (m = n + 1/4 is exact; with |m| < 2^29 the P3 term, |m P3| < 2e-24, lies far
// below the last bit of r and is left out), then the degree-7 polynomials of the half-turn table.  Returns n; sin(x - pi/4) = (-1)^n s and cos(x - pi/4) = (-1)^n c.  Against
// the quarter-turn reduction this trades the quadrant selects for three more polynomial terms.
// h supplies -1/4 and the leading coefficients (hankel_tops).
__host__ __device__ __forceinline__ int sincos_half_turn(double x, double& s, double& c,
                                                         const HankelTops& h)
{
    const auto& T = hankel_tables().trig2;
    constexpr int S0 = 4, C0 = 4 + BEM2D_TRIG2_NS;
    const double n = rint(fma(x, T[0], h.mquarter));
    const double m = n + 0.25;
    const double r = fma(-m, T[2], fma(-m, T[1], x));
    const double r2 = r * r;
    double ps = fma_vvs(r2, h.trig[0], T[S0 + 1]);
#pragma unroll
    for (int i = S0 + 2; i < C0; ++i) ps = fma_vvs(ps, r2, T[i]);
    s = fma(ps * r2, r, r);
    double pc = fma_vvs(r2, h.trig[1], T[C0 + 1]);
#pragma unroll
    for (int i = C0 + 2; i < C0 + BEM2D_TRIG2_NC; ++i) pc = fma_vvs(pc, r2, T[i]);
    c = fma(pc, r2, 1.0);
    return (int)n;
}

// a (-1)^n: the parity of n moved into the sign bit
__host__ __device__ __forceinline__ double with_parity_sign(double a, int n)
{
    const unsigned long long bits = __builtin_bit_cast(unsigned long long, a);
    return __builtin_bit_cast(double, bits ^ ((unsigned long long)(n & 1) << 63));
}

// P_nu, Q_nu for x >= BEM2D_PIECE_X[last] as polynomials in w = 1/x^2 (the last piece's map to z
// folded into the coefficients).
// h = hankel_tops<BEM2D_HANKEL_NPIECE>().
__host__ __device__ __forceinline__ void hankel_pq_far(double invx, double& P0, double& Q0,
                                                       double& P1, double& Q1, const HankelTops& h)
{
    const auto& T = hankel_tables();
    constexpr int N = BEM2D_HANKEL_FAR_DEG + 1;
    const double w = invx * invx;
    P0 = horner_top<N>(T.far[0], h.pq[0], w);
    Q0 = horner_top<N>(T.far[1], h.pq[1], w) * invx;
    P1 = horner_top<N>(T.far[2], h.pq[2], w);
    Q1 = horner_top<N>(T.far[3], h.pq[3], w) * invx;
}

// Modulus/phase combination for x >= 2:
//   H0 = sqrt(2/(pi x)) (P0 + i Q0) (C + i S),  H1 = sqrt(2/(pi x)) (P1 + i Q1) (S - i C),
// with C + i S = exp(i (x - pi/4)) = (-1)^n (c + i s) and the sign carried by the amplitude.
__host__ __device__ __forceinline__ void hankel01_combine(double x, double invx, double P0,
                                                          double Q0, double P1, double Q1,
                                                          double& h0r, double& h0i, double& h1r,
                                                          double& h1i)
{
    double S, C;
    const int n = sincos_half_turn(x, S, C, hankel_tops<BEM2D_HANKEL_NPIECE>());
#ifdef __HIP_DEVICE_COMPILE__
    const double A = with_parity_sign(rsqrt_refined(0.5 * kPi * x), n); // +-sqrt(2/(pi x))
#else
    const double A = with_parity_sign(sqrt(2.0 * kInvPi * invx), n);
#endif
    h0r = A * (P0 * C - Q0 * S);
    h0i = A * (P0 * S + Q0 * C);
    h1r = A * (P1 * S + Q1 * C);
    h1i = A * (Q1 * S - P1 * C);
}

// H0 = (h0r, h0i), H1 = (h1r, h1i) at x > 0, given invx = 1/x.
__host__ __device__ __forceinline__ void hankel01_inv(double x, double invx, double& h0r,
                                                      double& h0i, double& h1r, double& h1i)
{
    if (x < 2.0) {
        hankel01_series(x, h0r, h0i, h1r, h1i);
        return;
    }
    double P0, Q0, P1, Q1;
    if (x >= kHankelFarX)
        hankel_pq_far(invx, P0, Q0, P1, Q1, hankel_tops<BEM2D_HANKEL_NPIECE>());
    else if (x >= kHankelMidX)
        hankel_pq<4>(invx, P0, Q0, P1, Q1, hankel_tops<4>());
    else if (x >= 5.0)
        hankel_pq<3>(invx, P0, Q0, P1, Q1, hankel_tops<3>());
    else if (x >= 3.5)
        hankel_pq<2>(invx, P0, Q0, P1, Q1, hankel_tops<2>());
    else if (x >= 2.6)
        hankel_pq<1>(invx, P0, Q0, P1, Q1, hankel_tops<1>());
    else
        hankel_pq<0>(invx, P0, Q0, P1, Q1, hankel_tops<0>());
    hankel01_combine(x, invx, P0, Q0, P1, Q1, h0r, h0i, h1r, h1i);
}

// The piecewise tables staged in LDS for hankel01_gather: per piece, P0, xQ0, P1, xQ1 with
// BEM2D_HANKEL_DEGMAX + 1 coefficients each (the last piece zero-padded), then its z map.
constexpr int kPieceCoefs = BEM2D_HANKEL_DEGMAX + 1;
constexpr int kPieceStride = 4 * kPieceCoefs + 3;
constexpr int kPieceTableDoubles = BEM2D_HANKEL_NPIECE * kPieceStride;

// Copies the piece tables into lds (kPieceTableDoubles doubles) with threads lane = 0 .. nthr - 1.
__device__ __forceinline__ void stage_piece_tables(double* lds, int lane, int nthr)
{
#ifdef __HIP_DEVICE_COMPILE__
    const auto& T = hankel_tables();
    for (int e = lane; e < kPieceTableDoubles; e += nthr) {
        const int p = e / kPieceStride, o = e - p * kPieceStride;
        lds[e] = o < 4 * kPieceCoefs ? T.fit[p][o / kPieceCoefs][o % kPieceCoefs]
                                     : T.zmap[p][o - 4 * kPieceCoefs];
    }
#endif
}

// hankel01_inv for lanes whose x spread over several pieces: every lane with x >= 2 evaluates
// one set of polynomials with its own piece's coefficients, gathered from the LDS copy (the last
// piece serves x >= 8, the far table's range included).  Branching per piece instead costs a
// wavefront every piece that any of its lanes falls in.
__device__ __forceinline__ void hankel01_gather(double x, double invx, const double* tab,
                                                double& h0r, double& h0i, double& h1r, double& h1i)
{
    if (x < 2.0) {
        hankel01_series(x, h0r, h0i, h1r, h1i);
        return;
    }
    int p = 0;
#pragma unroll
    for (int q = 1; q < BEM2D_HANKEL_NPIECE; ++q) p += x >= BEM2D_PIECE_X[q] ? 1 : 0;
    const double* c = tab + p * kPieceStride;
    const double y = c[4 * kPieceCoefs] * invx * invx;
    const double z = (y - c[4 * kPieceCoefs + 1]) * c[4 * kPieceCoefs + 2];
    double v[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const double* cf = c + f * kPieceCoefs;
        double r = cf[kPieceCoefs - 1];
#pragma unroll
        for (int i = kPieceCoefs - 2; i >= 0; --i) r = fma(r, z, cf[i]);
        v[f] = r;
    }
    hankel01_combine(x, invx, v[0], v[1] * invx, v[2], v[3] * invx, h0r, h0i, h1r, h1i);
}

__host__ __device__ __forceinline__ void hankel01(double x, double& h0r, double& h0i, double& h1r,
                                                  double& h1i)
{
    hankel01_inv(x, rcp_refined(x), h0r, h0i, h1r, h1i);
}

} // namespace bem2d_dev
