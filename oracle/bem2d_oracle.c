/*
 * bem2d_oracle.c -- CPU restatement of the bem2d hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker for the HIP path: only tests/, the smoke entry and the
 * cpu_baseline leg of bench.py may load it.  The product library never links it.
 *
 * It restates, in plain C99 with long-double special functions, what the reference does for:
 *   - Gauss-Legendre rule on [0,1] by Golub-Welsch      lib/bem2d_quadrature.cpp:10-56
 *   - Rescale / MapPoints2d                            lib/bem2d_quadrature.cpp:58-76,
 *                                                      lib/bem2d_quadpoints.cpp:9-50
 *   - Gauss2D, AdaptedGauss1, AdaptedGauss2, AdaptedGauss3
 *                                                      lib/bem2d_quadpoints.cpp:31-183
 *   - curve / element Map, Deriv, Normal               lib/bem2d_curvetypes.h:13-99,
 *                                                      lib/bem2d_element.h:29-34,72-116
 *   - H0^(1), H1^(1) of real argument (GSL J/Y)        lib/bem2d_mathroutines.cpp:20-42
 *   - H0^(1), H1^(1) of complex argument (AMOS zbesh)  lib/bem2d_mathroutines.cpp:44-75
 *   - SingleLayer, DoubleLayer, ConjDoubleLayer, CombinedSingleConjDouble, IdKernel
 *                                                      lib/bem2d_kernel.cpp:15-112, bem2d_kernel.h:91-98
 *   - Integrate / IntegrateIdent / DiscreteKernel      lib/bem2d_quadrature.h:37-109,231-294
 *   - EvalIdent (mass matrix)                          lib/bem2d_quadrature.cpp:79-127
 *   - Integrate<F> 1-D, DotProdBasFuns                 lib/bem2d_quadrature.h:111-156,
 *     with PlaneWave / NormalPlaneWave / CombinedPlaneWave   lib/bem2d_fun.h:14-93
 *   - SolveSystem (zgetrf + zgetrs semantics)          lib/bem2d_mathroutines.cpp:323-351
 *   - EvalKernel<SingleLayer> (EvalSol's potential)    lib/bem2d_quadrature.h:184-228,
 *                                                      lib/bem2d_soundsoftscattering.h:149-168
 *   - far field of u = u_inc - S phi (not in the reference; derived from EvalSol's representation)
 *
 * Pinning (what fixes each function independently of this file):
 *   oracle_gauss          numpy leggauss; exactness on monomials; ascending order; sum w = 1
 *   oracle_rule           node counts, sum of weights, exact integrals of x^a y^b and of log-singular
 *                         integrands over the covered region (closed forms / mpmath)
 *   oracle_hankel         mpmath hankel1 at 50 digits (tests/golden/hankel_mpmath.npz); Wronskian
 *   oracle_hankel_complex mpmath hankel1 at 30 digits (tests/golden/hankel_complex_mpmath.npz);
 *                         scipy.special.hankel1 (AMOS zbesh itself); Wronskian
 *   oracle_assemble       entries of tiny meshes evaluated entirely in mpmath with the same rule nodes
 *                         (tests/golden/entries_mpmath.npz); analytic circle series end to end
 *   oracle_rhs            closed form of the 1-D plane-wave integral over flat elements
 *   oracle_lu_solve       numpy.linalg.solve; backward error
 *   oracle_eval_sl,
 *   oracle_farfield       analytic circle series (sound-soft unit circle) after a solve
 *
 * Element description shared with the HIP C-ABI: type[n], param[8*n], index/next/prev[n] (int64, -1 =
 * none).  Types: 0 flat ConstElement (p1x,p1y,dx,dy), and AnalyticCurveElement on
 * 1 Circle, 2 Trefoil, 3 Kite, 4 Ellipse(rho), 5 EllipseArc(a,b,theta1,theta2), 6 InvEllipse(alpha),
 * with param = (tstart, tend, c0, c1, c2, c3, 0, 0).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#include <quadmath.h>

#define O_PI 3.14159265358979323846 /* lib/bem2d_defs.h: PI */
#define O_PI_L 3.141592653589793238462643383279502884L
#define O_EULER_L 0.577215664901532860606512090082402431L

enum { E_FLAT = 0, E_CIRCLE = 1, E_TREFOIL = 2, E_KITE = 3, E_ELLIPSE = 4, E_ELLIPSE_ARC = 5,
       E_INV_ELLIPSE = 6 };
enum { K_NONE = 0, K_SL = 1, K_DL = 2, K_CDL = 3, K_COMBINED = 4, K_ID = 5 };
enum { F_PLANE = 0, F_NORMAL_PLANE = 1, F_COMBINED_PLANE = 2 };
enum { R_GAUSS2D = 0, R_AG1 = 1, R_AG2 = 2, R_AG3 = 3 };

typedef struct { double x, y; } pt;

/* ------------------------------------------------------------------------------------------ */
/* Gauss-Legendre on [0,1] by Golub-Welsch (lib/bem2d_quadrature.cpp:10-56): eigenvalues of the  */
/* symmetric tridiagonal Jacobi matrix with zero diagonal and beta_i = 0.5/sqrt(1-1/(4 i^2)),   */
/* x = (1+lambda)/2, w = (first eigenvector component)^2.  The eigen-solver is the implicit      */
/* QL iteration with Wilkinson shifts (textbook tql2), in long double.                           */
/* ------------------------------------------------------------------------------------------ */
static int tql2_ld(int n, long double* d, long double* e, long double* z /* n*n row-major z[r*n+c] */)
{
    for (int i = 1; i < n; i++) e[i - 1] = e[i];
    e[n - 1] = 0.0L;
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; m++) {
                long double dd = fabsl(d[m]) + fabsl(d[m + 1]);
                if (fabsl(e[m]) <= 1e-19L * dd) break;
            }
            if (m != l) {
                if (iter++ == 200) return 1;
                long double g = (d[l + 1] - d[l]) / (2.0L * e[l]);
                long double r = hypotl(g, 1.0L);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabsl(r) : -fabsl(r)));
                long double s = 1.0L, c = 1.0L, p = 0.0L;
                int i;
                for (i = m - 1; i >= l; i--) {
                    long double f = s * e[i], b = c * e[i];
                    r = hypotl(f, g);
                    e[i + 1] = r;
                    if (r == 0.0L) { d[i + 1] -= p; e[m] = 0.0L; break; }
                    s = f / r; c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0L * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int k = 0; k < n; k++) {
                        f = z[k * n + i + 1];
                        z[k * n + i + 1] = s * z[k * n + i] + c * f;
                        z[k * n + i] = c * z[k * n + i] - s * f;
                    }
                }
                if (r == 0.0L && i >= l) continue;
                d[l] -= p; e[l] = g; e[m] = 0.0L;
            }
        } while (m != l);
    }
    return 0;
}

int oracle_gauss(int n, double* x, double* w)
{
    if (n < 1) return 1; /* ParameterException, lib/bem2d_quadpoints.cpp:11 */
    if (n == 1) { x[0] = 0.5; w[0] = 1.0; return 0; } /* lib/bem2d_quadrature.cpp:25-29 */
    long double* d = calloc((size_t)n, sizeof(long double));
    long double* e = calloc((size_t)n, sizeof(long double));
    long double* z = calloc((size_t)n * n, sizeof(long double));
    for (int i = 0; i < n; i++) z[i * n + i] = 1.0L;
    for (int i = 0; i < n - 1; i++) {
        long double ip = (long double)(i + 1);
        e[i + 1] = 0.5L / sqrtl(1.0L - 1.0L / (4.0L * ip * ip));
    }
    int rc = tql2_ld(n, d, e, z);
    /* sort ascending (dstev returns eigenvalues in ascending order) */
    int* ord = malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) ord[i] = i;
    for (int i = 1; i < n; i++) {
        int v = ord[i], j = i - 1;
        while (j >= 0 && d[ord[j]] > d[v]) { ord[j + 1] = ord[j]; j--; }
        ord[j + 1] = v;
    }
    for (int i = 0; i < n; i++) {
        int c = ord[i];
        x[i] = (double)((1.0L + d[c]) / 2.0L);
        long double v0 = z[0 * n + c];
        w[i] = (double)(v0 * v0);
    }
    free(ord); free(d); free(e); free(z);
    return rc;
}

/* Rescale of a 2-D tensor rule onto [a,b]x[c,d] (lib/bem2d_quadrature.cpp:68-76). */
static void push_box(const double* gx, const double* gy, const double* gw, int m, double a, double b,
                     double c, double d, double* x, double* y, double* w, int* cnt)
{
    for (int i = 0; i < m; i++) {
        x[*cnt] = a + (b - a) * gx[i];
        y[*cnt] = c + (d - c) * gy[i];
        w[*cnt] = (b - a) * (d - c) * gw[i];
        (*cnt)++;
    }
}

/* Number of nodes of each rule (lib/bem2d_quadpoints.cpp:166,210,240,282). */
int oracle_rule_size(int which, int N, int L)
{
    switch (which) {
    case R_GAUSS2D: return N * N;
    case R_AG1: return N * N * L * L;
    case R_AG2: case R_AG3: return 3 * N * N * L;
    }
    return -1;
}

/* Gauss2D / AdaptedGauss1/2/3 node lists, in the reference's node order
 * (lib/bem2d_quadpoints.cpp:166-183, 210-235, 240-277, 282-318). */
int oracle_rule(int which, int N, int L, double sigma, double* x, double* y, double* w)
{
    if (N < 1) return -1;
    double* xt = malloc(sizeof(double) * N);
    double* wt = malloc(sizeof(double) * N);
    oracle_gauss(N, xt, wt);
    int m = N * N;
    double* gx = malloc(sizeof(double) * m);
    double* gy = malloc(sizeof(double) * m);
    double* gw = malloc(sizeof(double) * m);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            gx[i * N + j] = xt[i];
            gy[i * N + j] = xt[j];
            gw[i * N + j] = wt[i] * wt[j];
        }
    int cnt = 0;
    if (which == R_GAUSS2D) {
        memcpy(x, gx, sizeof(double) * m); memcpy(y, gy, sizeof(double) * m);
        memcpy(w, gw, sizeof(double) * m);
        cnt = m;
    } else if (which == R_AG1) {
        double* xl = malloc(sizeof(double) * (L + 1));
        for (int i = 0; i < L; i++) xl[i] = pow(sigma, (double)(L - i));
        xl[L] = 1.0;
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++)
                push_box(gx, gy, gw, m, xl[i], xl[i + 1], xl[j], xl[j + 1], x, y, w, &cnt);
        free(xl);
    } else if (which == R_AG2) {
        double a = sigma, b = 1;
        for (int l = L; l > 0; l--) {
            push_box(gx, gy, gw, m, 0, a, 1 - b, 1 - a, x, y, w, &cnt);
            push_box(gx, gy, gw, m, a, b, 1 - b, 1 - a, x, y, w, &cnt);
            push_box(gx, gy, gw, m, a, b, 1 - a, 1, x, y, w, &cnt);
            a *= sigma; b *= sigma;
        }
    } else if (which == R_AG3) {
        double a = 1, b = sigma;
        for (int l = L; l > 0; l--) {
            push_box(gx, gy, gw, m, 1 - a, 1 - b, 0, b, x, y, w, &cnt);
            push_box(gx, gy, gw, m, 1 - a, 1 - b, b, a, x, y, w, &cnt);
            push_box(gx, gy, gw, m, 1 - b, 1, b, a, x, y, w, &cnt);
            a *= sigma; b *= sigma;
        }
    } else {
        cnt = -1;
    }
    free(gx); free(gy); free(gw); free(xt); free(wt);
    return cnt;
}

/* ------------------------------------------------------------------------------------------ */
/* H0^(1), H1^(1) of real x > 0, in long double (stand-in for GSL J0/Y0/J1/Y1,                  */
/* lib/bem2d_mathroutines.cpp:20-42).  Textbook algorithms:                                      */
/*   x <= 2      ascending series, A&S 9.1.10 (J) and 9.1.11 with n=0,1 (Y)                       */
/*   2 < x < 25  Miller backward recurrence normalised by J0 + 2 sum J_2k = 1 (A&S 9.1.46), with  */
/*               Neumann series for Y0 (A&S 9.1.88) and its derivative for Y1 = -Y0'             */
/*   x >= 25     Hankel asymptotic expansion, A&S 9.2.5-9.2.10                                    */
/* ------------------------------------------------------------------------------------------ */
static void bessel01_ld(long double x, long double* J0, long double* Y0, long double* J1,
                        long double* Y1)
{
    const long double pi = O_PI_L, g = O_EULER_L;
    if (x <= 2.0L) {
        long double t = x * x / 4.0L; /* (x/2)^2 */
        long double lg = logl(x / 2.0L);
        /* J0 = sum (-t)^k/(k!)^2; J1 = (x/2) sum (-t)^k/(k!(k+1)!) */
        long double a0 = 1.0L, a1 = 1.0L; /* term magnitudes without sign */
        long double sj0 = 0, sj1 = 0, sy0 = 0, sy1 = 0;
        long double Hk = 0.0L, Hk1 = 1.0L; /* harmonic numbers H_k, H_{k+1} */
        for (int k = 0; k < 60; k++) {
            long double sg = (k & 1) ? -1.0L : 1.0L;
            sj0 += sg * a0;
            sj1 += sg * a1;
            sy0 += sg * a0 * Hk;              /* sum (-1)^k H_k t^k/(k!)^2 */
            sy1 += sg * a1 * (Hk + Hk1);      /* sum (-1)^k (H_k + H_{k+1}) t^k/(k!(k+1)!) */
            a0 *= t / (long double)((k + 1) * (k + 1));
            a1 *= t / (long double)((k + 1) * (k + 2));
            Hk += 1.0L / (long double)(k + 1);
            Hk1 += 1.0L / (long double)(k + 2);
            if (a0 < 1e-40L && a1 < 1e-40L) break;
        }
        *J0 = sj0;
        *J1 = (x / 2.0L) * sj1;
        /* Y0 = (2/pi)(ln(x/2)+gamma) J0 - (2/pi) sum (-1)^k H_k t^k/(k!)^2           (A&S 9.1.13) */
        *Y0 = (2.0L / pi) * (lg + g) * sj0 - (2.0L / pi) * sy0;
        /* Y1 = -2/(pi x) + (2/pi) ln(x/2) J1 - (1/pi)(x/2) sum (psi(k+1)+psi(k+2)) (-t)^k/(k!(k+1)!) */
        /* with psi(k+1) = -gamma + H_k                                                  (A&S 9.1.11) */
        long double s_psi = sy1 - 2.0L * g * sj1;
        *Y1 = -2.0L / (pi * x) + (2.0L / pi) * lg * (*J1) - (1.0L / pi) * (x / 2.0L) * s_psi;
        return;
    }
    if (x < 25.0L) {
        int M = 2 * (int)((x + 50.0L) / 2.0L);
        long double jp1 = 0.0L, jn = 1e-30L; /* j_{M+1}, j_M */
        long double snorm = 0.0L, sy0 = 0.0L, sy1 = 0.0L;
        /* running values: we walk n = M .. 1 computing j_{n-1} */
        long double jv[128]; /* M <= 74 for x < 25 */
        jv[M + 1] = 0.0L;
        jv[M] = jn;
        for (int n = M; n >= 1; n--) {
            long double jm1 = (2.0L * (long double)n / x) * jn - jp1;
            jv[n - 1] = jm1;
            jp1 = jn; jn = jm1;
        }
        snorm = jv[0];
        for (int k = 1; 2 * k <= M; k++) {
            snorm += 2.0L * jv[2 * k];
            long double sg = (k & 1) ? -1.0L : 1.0L;
            sy0 += sg * jv[2 * k] / (long double)k;
            long double jnext = (2 * k + 1 <= M) ? jv[2 * k + 1] : 0.0L;
            sy1 += sg * (jv[2 * k - 1] - jnext) / (long double)k;
        }
        long double j0 = jv[0] / snorm, j1 = jv[1] / snorm;
        long double lg = logl(x / 2.0L) + g;
        *J0 = j0; *J1 = j1;
        *Y0 = (2.0L / pi) * lg * j0 - (4.0L / pi) * sy0 / snorm;
        *Y1 = (2.0L / pi) * (lg * j1 - j0 / x) + (2.0L / pi) * sy1 / snorm;
        return;
    }
    /* Hankel asymptotic expansion: P, Q with a_k(nu) = prod_{m=1..k} (mu-(2m-1)^2) / (k! 8^k) */
    long double P[2], Q[2];
    for (int nu = 0; nu < 2; nu++) {
        long double mu = 4.0L * nu * nu;
        long double ak = 1.0L, p = 0.0L, q = 0.0L, last = 1e300L;
        long double xp = 1.0L;
        for (int k = 0; k < 200; k++) {
            long double term = ak / xp;
            if (fabsl(term) > last) break; /* asymptotic: stop at the smallest term */
            last = fabsl(term);
            int r = k % 4; /* sign pattern: P = a0 - a2/x^2 + a4/x^4..., Q = a1/x - a3/x^3 ... */
            if (r == 0) p += term;
            else if (r == 1) q += term;
            else if (r == 2) p -= term;
            else q -= term;
            if (last < 1e-24L) break;
            long double m2 = (long double)(2 * k + 1);
            ak *= (mu - m2 * m2) / (8.0L * (long double)(k + 1));
            xp *= x;
        }
        P[nu] = p; Q[nu] = q;
    }
    long double s = sinl(x), c = cosl(x);
    long double r2 = sqrtl(0.5L);
    long double c0 = (c + s) * r2, s0 = (s - c) * r2; /* cos, sin of x - pi/4 */
    long double c1 = s0, s1 = -c0;                    /* cos, sin of x - 3pi/4 */
    long double amp = sqrtl(2.0L / (pi * x));
    *J0 = amp * (P[0] * c0 - Q[0] * s0);
    *Y0 = amp * (P[0] * s0 + Q[0] * c0);
    *J1 = amp * (P[1] * c1 - Q[1] * s1);
    *Y1 = amp * (P[1] * s1 + Q[1] * c1);
}

void oracle_hankel(double x, double* out4)
{
    long double J0, Y0, J1, Y1;
    bessel01_ld((long double)x, &J0, &Y0, &J1, &Y1);
    out4[0] = (double)J0; out4[1] = (double)Y0; out4[2] = (double)J1; out4[3] = (double)Y1;
}

void oracle_hankel_many(const double* x, int64_t n, double* out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) oracle_hankel(x[i], out + 4 * i);
}

/* ------------------------------------------------------------------------------------------ */
/* H0^(1), H1^(1) of complex z = zr + i zi (stand-in for AMOS zbesh with kode = 1, n = 1,        */
/* lib/bem2d_mathroutines.cpp:44-75).  Textbook algorithms, chosen apart from the HIP path's:   */
/*   |z| <= 20  ascending series A&S 9.1.10 (J) and 9.1.11 with n = 0, 1 (Y) in __float128, so   */
/*              that the cancellation of terms up to e^|z| still leaves ~1e-17 relative;         */
/*   |z| > 20   Hankel asymptotic expansion A&S 9.2.5-9.2.10 in long double, summed until the    */
/*              terms fall below 1e-21 of the sum or stop decreasing (smallest term ~e^-2|z|).    */
/* ------------------------------------------------------------------------------------------ */
void oracle_hankel_complex(double zr, double zi, double* out4)
{
    if (zr * zr + zi * zi <= 400.0) {
        const __float128 PIQ = M_PIq;
        const __float128 EULQ = 0.577215664901532860606512090082402431Q;
        __complex128 z, t, term, J0, S0, J1s, S1;
        __real__ z = zr;
        __imag__ z = zi;
        t = -z * z / 4;
        term = 1;
        J0 = 1; S0 = 0; J1s = 1; S1 = 1;
        __float128 H = 0, big = 1;
        for (int m = 1; m < 400; m++) {
            term = term * t / ((__float128)m * m);
            H += 1.0Q / m;
            J0 += term;
            S0 += H * term;
            J1s += term / (m + 1);
            S1 += (2 * H + 1.0Q / (m + 1)) * term / (m + 1);
            __float128 a = cabsq(term);
            if (a > big) big = a;
            if (m > 4 && a < 1e-40Q * big) break;
        }
        __complex128 J1 = z / 2 * J1s;
        __complex128 L = clogq(z / 2) + EULQ;
        __complex128 Y0 = 2 / PIQ * (L * J0 - S0);
        __complex128 Y1 = 2 / PIQ * L * J1 - 2 / (PIQ * z) - z / (2 * PIQ) * S1;
        __complex128 H0 = J0 + 1.0Qi * Y0, H1 = J1 + 1.0Qi * Y1;
        out4[0] = (double)crealq(H0); out4[1] = (double)cimagq(H0);
        out4[2] = (double)crealq(H1); out4[3] = (double)cimagq(H1);
        return;
    }
    long double complex z = (long double)zr + I * (long double)zi;
    long double complex s[2];
    for (int nu = 0; nu < 2; nu++) {
        long double mu = 4.0L * nu * nu;
        long double complex term = 1, sum = 1;
        long double last = 1;
        for (int k = 1; k < 80; k++) {
            long double o = (2.0L * k - 1) * (2.0L * k - 1);
            term = term * I * (mu - o) / (8.0L * k * z);
            long double a = cabsl(term);
            if (a > last) break; /* past the smallest term: the series diverges from here */
            last = a;
            sum += term;
            if (a < 1e-21L * cabsl(sum)) break;
        }
        s[nu] = sum;
    }
    /* sqrt(2/(pi z)) e^{iz} e^{-i pi/4} (and e^{-i pi/2} more for nu = 1) */
    long double complex f = csqrtl(2.0L / (O_PI_L * z)) * cexpl(I * z) *
                            (1.0L - I) / sqrtl(2.0L);
    long double complex H0 = f * s[0], H1 = -I * f * s[1];
    out4[0] = (double)creall(H0); out4[1] = (double)cimagl(H0);
    out4[2] = (double)creall(H1); out4[3] = (double)cimagl(H1);
}

void oracle_hankel_complex_many(const double* z, int64_t n, double* out)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++) oracle_hankel_complex(z[2 * i], z[2 * i + 1], out + 4 * i);
}

/* ------------------------------------------------------------------------------------------ */
/* Geometry (lib/bem2d_curvetypes.h:13-99, lib/bem2d_element.h:29-34, 72-116).                    */
/* The association of the reference's products (e.g. (2*PI)*t) is kept.                          */
/* The curves' sin and cos are the C library's, as in the reference; built with ORACLE_CR_TRIG    */
/* they are correctly rounded instead (evaluated in __float128 and rounded once): the distance     */
/* between the two builds measures how far the reference's own matrix moves with the libm it is   */
/* linked against (near-field double-layer entries amplify a one-ulp change of a node position).  */
/* ------------------------------------------------------------------------------------------ */
#ifdef ORACLE_CR_TRIG
#define CSIN(x) ((double)sinq((__float128)(x)))
#define CCOS(x) ((double)cosq((__float128)(x)))
#else
#define CSIN(x) sin(x)
#define CCOS(x) cos(x)
#endif
static void curve_eval(int type, const double* c, double t, pt* M, pt* D)
{
    const double PI = O_PI;
    switch (type) {
    case E_CIRCLE:
        M->x = CCOS(2 * PI * t); M->y = CSIN(2 * PI * t);
        D->x = 2 * PI * (-1.0 * CSIN(2 * PI * t)); D->y = 2 * PI * CCOS(2 * PI * t);
        break;
    case E_TREFOIL: {
        double r = 1 + 0.3 * CCOS(3 * 2 * PI * t);
        M->x = r * CCOS(2 * PI * t); M->y = r * CSIN(2 * PI * t);
        double rp = -0.3 * 3 * 2 * PI * CSIN(3 * 2 * PI * t);
        D->x = rp * CCOS(2 * PI * t) - r * 2 * PI * CSIN(2 * PI * t);
        D->y = rp * CSIN(2 * PI * t) + r * 2 * PI * CCOS(2 * PI * t);
        break;
    }
    case E_KITE:
        M->x = CCOS(2 * PI * t) + 0.65 * CCOS(4 * PI * t) - 0.65; M->y = 1.5 * CSIN(2 * PI * t);
        D->x = -2 * PI * CSIN(2 * PI * t) - 0.65 * 4 * PI * CSIN(4 * PI * t);
        D->y = 1.5 * 2 * PI * CCOS(2 * PI * t);
        break;
    case E_ELLIPSE: {
        double rho = c[0];
        M->x = rho * CCOS(2 * PI * t); M->y = 1 / rho * CSIN(2 * PI * t);
        D->x = -rho * 2 * PI * CSIN(2 * PI * t); D->y = 2 * PI / rho * CCOS(2 * PI * t);
        break;
    }
    case E_ELLIPSE_ARC: {
        double a = c[0], b = c[1], t1 = c[2], t2 = c[3];
        M->x = a * CCOS(t1 + t * (t2 - t1)); M->y = b * CSIN(t1 + t * (t2 - t1));
        D->x = -a * (t2 - t1) * CSIN(t1 + t * (t2 - t1));
        D->y = b * (t2 - t1) * CCOS(t1 + t * (t2 - t1));
        break;
    }
    case E_INV_ELLIPSE: {
        double alpha = c[0];
        double complex s = 2.0 * PI * t * I;
        double complex e2 = alpha * cexp(2.0 * s);
        double complex res = cexp(s) / (1.0 + e2);
        M->x = creal(res); M->y = cimag(res);
        double complex dr = 2.0 * PI * I * cexp(s) / (1.0 + e2) / (1.0 + e2) * (1.0 - e2);
        D->x = creal(dr); D->y = cimag(dr);
        break;
    }
    default:
        M->x = M->y = D->x = D->y = NAN;
    }
}

typedef struct {
    const int32_t* type;
    const double* param;
    const int64_t* index;
    const int64_t* next;
    const int64_t* prev;
    int64_t n;
} geom_t;

/* Element Map/Deriv at local parameter t in [0,1]. */
static void elem_eval(const geom_t* g, int64_t e, double t, pt* M, pt* D)
{
    const double* p = g->param + 8 * e;
    int ty = g->type[e];
    if (ty == E_FLAT) {
        /* ConstElement: Map = p1 + d*t, Deriv = d (lib/bem2d_element.h:77-83) */
        M->x = p[0] + t * p[2]; M->y = p[1] + t * p[3];
        D->x = p[2]; D->y = p[3];
        return;
    }
    /* AnalyticCurveElement (lib/bem2d_element.h:101-106) */
    double ts = p[0], te = p[1];
    pt cm, cd;
    curve_eval(ty, p + 2, ts + t * (te - ts), &cm, &cd);
    *M = cm;
    D->x = (te - ts) * cd.x; D->y = (te - ts) * cd.y;
}

static inline double plen(pt p) { return sqrt(p.x * p.x + p.y * p.y); }
static inline pt pnormal(pt D) /* Normal = normalize(D.y, -D.x) (lib/bem2d_element.h:29-34) */
{
    double l = plen((pt){D.y, -D.x});
    pt r = {D.y / l, -D.x / l};
    return r;
}

/* Hankel pair in double complex */
static void hankel_pair(double x, double complex* h0, double complex* h1)
{
    double v[4];
    oracle_hankel(x, v);
    *h0 = v[0] + I * v[1];
    *h1 = v[2] + I * v[3];
}

/* Hankel pair of complex argument in double complex */
static void hankel_pair_c(double complex z, double complex* h0, double complex* h1)
{
    double v[4];
    oracle_hankel_complex(creal(z), cimag(z), v);
    *h0 = v[0] + I * v[1];
    *h1 = v[2] + I * v[3];
}

/* The k.im != 0 branches of the kernel functors (lib/bem2d_kernel.cpp:18-27, 44-53, 71-80, 97-111). */
static double complex kernel_eval_c(int kind, double complex kc, double eta, pt x, pt y, pt n1,
                                    pt n2)
{
    double complex h0, h1;
    switch (kind) {
    case K_SL: {
        double d = plen((pt){x.x - y.x, x.y - y.y});
        hankel_pair_c(kc * d, &h0, &h1);
        return I / 4.0 * h0;
    }
    case K_DL: {
        pt w = {y.x - x.x, y.y - x.y};
        double d = plen(w);
        hankel_pair_c(kc * d, &h0, &h1);
        return -I * kc / 4.0 * (h1 / d * (n2.x * w.x + n2.y * w.y));
    }
    case K_CDL: {
        pt w = {x.x - y.x, x.y - y.y};
        double d = plen(w);
        hankel_pair_c(kc * d, &h0, &h1);
        return -I * kc / 4.0 * (h1 / d * (n1.x * w.x + n1.y * w.y));
    }
    case K_COMBINED: {
        pt w = {x.x - y.x, x.y - y.y};
        double d = plen(w);
        hankel_pair_c(kc * d, &h0, &h1);
        double complex slayer = I / 4.0 * h0;
        double complex cdlayer = -I * kc / 4.0 * (h1 / d * (n1.x * w.x + n1.y * w.y));
        return cdlayer - I * eta * slayer;
    }
    case K_ID:
        return 1.0;
    }
    return 0.0;
}

/* Kernel functors (lib/bem2d_kernel.cpp:15-112, bem2d_kernel.h:91-98); ki = Im k (0: real k). */
static double complex kernel_eval(int kind, double k, double ki, double eta, pt x, pt y, pt n1,
                                  pt n2)
{
    if (ki != 0.0) return kernel_eval_c(kind, k + I * ki, eta, x, y, n1, n2);
    double complex h0, h1;
    switch (kind) {
    case K_SL: {
        double d = plen((pt){x.x - y.x, x.y - y.y});
        hankel_pair(k * d, &h0, &h1);
        return I / 4.0 * h0;
    }
    case K_DL: {
        pt w = {y.x - x.x, y.y - x.y};
        double d = plen(w);
        hankel_pair(k * d, &h0, &h1);
        return -I * k / 4.0 * (h1 / d * (n2.x * w.x + n2.y * w.y));
    }
    case K_CDL: {
        pt w = {x.x - y.x, x.y - y.y};
        double d = plen(w);
        hankel_pair(k * d, &h0, &h1);
        return -I * k / 4.0 * (h1 / d * (n1.x * w.x + n1.y * w.y));
    }
    case K_COMBINED: {
        pt w = {x.x - y.x, x.y - y.y};
        double d = plen(w);
        hankel_pair(k * d, &h0, &h1);
        double complex slayer = I / 4.0 * h0;
        double complex cdlayer = -I * k / 4.0 * (h1 / d * (n1.x * w.x + n1.y * w.y));
        return cdlayer - I * eta * slayer;
    }
    case K_ID:
        return 1.0;
    }
    return 0.0;
}

/* Integrate<kernel,T> over a regular or adjacent pair (lib/bem2d_quadrature.h:37-65). P0 basis. */
static double complex integrate_pair(const geom_t* g, int64_t i, int64_t j, int kind, double k,
                                     double ki, double eta, const double* rx, const double* ry,
                                     const double* rw, int m)
{
    double complex res = 0;
    for (int q = 0; q < m; q++) {
        pt Mx, Dx, My, Dy;
        elem_eval(g, i, rx[q], &Mx, &Dx);
        elem_eval(g, j, ry[q], &My, &Dy);
        double complex gv = kernel_eval(kind, k, ki, eta, Mx, My, pnormal(Dx), pnormal(Dy));
        double s1 = plen(Dx), s2 = plen(Dy);
        res += gv * s1 * s2 * rw[q];
    }
    return res;
}

/* IntegrateIdent: Duffy split of the coincident pair (lib/bem2d_quadrature.h:67-109). */
static double complex integrate_ident(const geom_t* g, int64_t i, int64_t j, int kind, double k,
                                      double ki, double eta, const double* rx, const double* ry,
                                      const double* rw, int m)
{
    double complex res = 0;
    for (int q = 0; q < m; q++) {
        double xq = rx[q], yq = ry[q];
        pt Mx, Dx, My, Dy;
        double s2p = (1 - yq) * xq;
        elem_eval(g, i, xq, &Mx, &Dx);
        elem_eval(g, j, s2p, &My, &Dy);
        double complex gv = kernel_eval(kind, k, ki, eta, Mx, My, pnormal(Dx), pnormal(Dy));
        res += xq * gv * plen(Dx) * plen(Dy) * rw[q];

        double s1b = 1 - xq, s2b = 1 - xq + yq * xq;
        elem_eval(g, i, s1b, &Mx, &Dx);
        elem_eval(g, j, s2b, &My, &Dy);
        gv = kernel_eval(kind, k, ki, eta, Mx, My, pnormal(Dx), pnormal(Dy));
        res += xq * gv * plen(Dx) * plen(Dy) * rw[q];
    }
    return res;
}

typedef struct {
    int N, L;
    double sigma;
    int m0, m1, m2, m3;
    double *x0, *y0, *w0, *x1, *y1, *w1, *x2, *y2, *w2, *x3, *y3, *w3;
    double *gx, *gw; /* Gauss1D */
} rules_t;

static int rules_make(rules_t* r, int N, int L, double sigma)
{
    if (N < 1 || L < 0 || !(sigma > 0 && sigma < 1)) return 1;
    r->N = N; r->L = L; r->sigma = sigma;
    r->m0 = oracle_rule_size(R_GAUSS2D, N, L);
    r->m1 = oracle_rule_size(R_AG1, N, L);
    r->m2 = oracle_rule_size(R_AG2, N, L);
    r->m3 = oracle_rule_size(R_AG3, N, L);
#define ALLOC3(X, Y, W, M) X = malloc(8 * (M)); Y = malloc(8 * (M)); W = malloc(8 * (M));
    ALLOC3(r->x0, r->y0, r->w0, r->m0)
    ALLOC3(r->x1, r->y1, r->w1, r->m1)
    ALLOC3(r->x2, r->y2, r->w2, r->m2)
    ALLOC3(r->x3, r->y3, r->w3, r->m3)
#undef ALLOC3
    oracle_rule(R_GAUSS2D, N, L, sigma, r->x0, r->y0, r->w0);
    oracle_rule(R_AG1, N, L, sigma, r->x1, r->y1, r->w1);
    oracle_rule(R_AG2, N, L, sigma, r->x2, r->y2, r->w2);
    oracle_rule(R_AG3, N, L, sigma, r->x3, r->y3, r->w3);
    r->gx = malloc(8 * N); r->gw = malloc(8 * N);
    oracle_gauss(N, r->gx, r->gw);
    return 0;
}

static void rules_free(rules_t* r)
{
    free(r->x0); free(r->y0); free(r->w0); free(r->x1); free(r->y1); free(r->w1);
    free(r->x2); free(r->y2); free(r->w2); free(r->x3); free(r->y3); free(r->w3);
    free(r->gx); free(r->gw);
}

/* One entry of DiscreteKernel with the rule selection of lib/bem2d_quadrature.h:271-287,
 * plus the mass-matrix entry of EvalIdent (lib/bem2d_quadrature.cpp:108-123). */
static double complex entry(const geom_t* g, const rules_t* r, int64_t i, int64_t j, int kind,
                            double k, double ki, double eta, double mass_scale,
                            double kernel_scale)
{
    double complex kv = 0, mv = 0;
    int ident = g->index[i] == g->index[j];
    if (kind != K_NONE && kernel_scale != 0.0) {
        if (ident)
            kv = integrate_ident(g, i, j, kind, k, ki, eta, r->x1, r->y1, r->w1, r->m1);
        else if (g->next[i] == g->index[j])
            kv = integrate_pair(g, i, j, kind, k, ki, eta, r->x3, r->y3, r->w3, r->m3);
        else if (g->prev[i] == g->index[j])
            kv = integrate_pair(g, i, j, kind, k, ki, eta, r->x2, r->y2, r->w2, r->m2);
        else
            kv = integrate_pair(g, i, j, kind, k, ki, eta, r->x0, r->y0, r->w0, r->m0);
    }
    if (ident && mass_scale != 0.0) {
        for (int t = 0; t < r->N; t++) {
            pt M, D;
            elem_eval(g, j, r->gx[t], &M, &D);
            mv += plen(D) * r->gw[t];
        }
    }
    /* A = mass_scale*M + kernel_scale*K: the facade's (*pA_)=(*pId_)+2.0*(*pA_)
     * (lib/bem2d_soundsoftscattering.h:102) is mass_scale=1, kernel_scale=2. */
    return kernel_scale * kv + mass_scale * mv;
}

/* DiscreteKernel (+ EvalIdent, + the facade's M + 2K) over the columns cols[0..ncols) (all columns
 * when cols == NULL).  out is column-major with leading dimension ld; column c of out holds matrix
 * column cols[c].  OpenMP over columns as lib/bem2d_quadrature.h:253. */
int oracle_assemble_ck(int64_t n, const int32_t* type, const double* param, const int64_t* index,
                       const int64_t* next, const int64_t* prev, int N, int L, double sigma,
                       int kind, double k, double ki, double eta, double mass_scale,
                       double kernel_scale, const int64_t* cols, int64_t ncols, double* out,
                       int64_t ld)
{
    geom_t g = {type, param, index, next, prev, n};
    rules_t r;
    if (rules_make(&r, N, L, sigma)) return 1;
    if (!cols) ncols = n;
    double complex* o = (double complex*)out;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < ncols; c++) {
        int64_t j = cols ? cols[c] : c;
        for (int64_t i = 0; i < n; i++)
            o[i + ld * c] = entry(&g, &r, i, j, kind, k, ki, eta, mass_scale, kernel_scale);
    }
    rules_free(&r);
    return 0;
}

int oracle_assemble(int64_t n, const int32_t* type, const double* param, const int64_t* index,
                    const int64_t* next, const int64_t* prev, int N, int L, double sigma,
                    int kind, double k, double eta, double mass_scale, double kernel_scale,
                    const int64_t* cols, int64_t ncols, double* out, int64_t ld)
{
    return oracle_assemble_ck(n, type, param, index, next, prev, N, L, sigma, kind, k, 0.0, eta,
                              mass_scale, kernel_scale, cols, ncols, out, ld);
}

/* Field functors (lib/bem2d_fun.h:21-24, 45-49, 71-75), k = complex(k.re, k.im). */
static double complex field_eval(int kind, double k, double ki, double eta, pt dir, pt p, pt n)
{
    if (ki != 0.0) {
        double complex kc = k + I * ki;
        double complex f = cexp(kc * I * (dir.x * p.x + dir.y * p.y));
        switch (kind) {
        case F_PLANE: return f;
        case F_NORMAL_PLANE: return I * kc * f * (n.x * dir.x + n.y * dir.y);
        case F_COMBINED_PLANE: return I * kc * f * (n.x * dir.x + n.y * dir.y) - I * eta * f;
        }
        return 0;
    }
    double complex f = cexp(k * I * (dir.x * p.x + dir.y * p.y));
    switch (kind) {
    case F_PLANE: return f;
    case F_NORMAL_PLANE: return I * k * f * (n.x * dir.x + n.y * dir.y);
    case F_COMBINED_PLANE: return I * k * f * (n.x * dir.x + n.y * dir.y) - I * eta * f;
    }
    return 0;
}

/* DotProdBasFuns (lib/bem2d_quadrature.h:111-156) for nrhs directions; out[i + ld*r].
 * scale multiplies the result (the facade's 2.0*rhs, lib/bem2d_soundsoftscattering.h:109). */
int oracle_rhs_ck(int64_t n, const int32_t* type, const double* param, int N, int kind, double k,
                  double ki, double eta, const double* dirs, int nrhs, double scale, double* out,
                  int64_t ld)
{
    geom_t g = {type, param, NULL, NULL, NULL, n};
    double* gx = malloc(8 * N);
    double* gw = malloc(8 * N);
    oracle_gauss(N, gx, gw);
    double complex* o = (double complex*)out;
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++) {
        for (int r = 0; r < nrhs; r++) {
            pt dir = {dirs[2 * r], dirs[2 * r + 1]};
            double complex res = 0;
            for (int t = 0; t < N; t++) {
                pt M, D;
                elem_eval(&g, i, gx[t], &M, &D);
                double s = plen(D);
                res += field_eval(kind, k, ki, eta, dir, M, pnormal(D)) * s * gw[t];
            }
            o[i + ld * r] = scale * res;
        }
    }
    free(gx); free(gw);
    return 0;
}

int oracle_rhs(int64_t n, const int32_t* type, const double* param, int N, int kind, double k,
               double eta, const double* dirs, int nrhs, double scale, double* out, int64_t ld)
{
    return oracle_rhs_ck(n, type, param, N, kind, k, 0.0, eta, dirs, nrhs, scale, out, ld);
}

/* SolveSystem: LU with partial pivoting (pivot = first max of |re|+|im|, as izamax inside zgetrf),
 * then forward/back substitution for nrhs columns (lib/bem2d_mathroutines.cpp:323-351).
 * A (n x n, column-major, lda=n) is not modified; B (n x nrhs, ldb=n) is overwritten with X.
 * Returns 0, or i+1 if U(i,i) is exactly zero (zgetrf info > 0 -> LapackError). */
int oracle_lu_solve(int64_t n, int64_t nrhs, const double* Ain, double* Bio)
{
    double complex* A = malloc(sizeof(double complex) * n * n);
    memcpy(A, Ain, sizeof(double complex) * n * n);
    double complex* B = (double complex*)Bio;
    int64_t* piv = malloc(sizeof(int64_t) * n);
    int info = 0;
    for (int64_t kk = 0; kk < n; kk++) {
        int64_t p = kk;
        double best = -1;
        for (int64_t i = kk; i < n; i++) {
            double v = fabs(creal(A[i + n * kk])) + fabs(cimag(A[i + n * kk]));
            if (v > best) { best = v; p = i; }
        }
        piv[kk] = p;
        if (A[p + n * kk] == 0) { if (!info) info = (int)kk + 1; continue; }
        if (p != kk)
            for (int64_t c = 0; c < n; c++) {
                double complex t = A[kk + n * c]; A[kk + n * c] = A[p + n * c]; A[p + n * c] = t;
            }
        double complex inv = 1.0 / A[kk + n * kk];
        for (int64_t i = kk + 1; i < n; i++) A[i + n * kk] *= inv;
#pragma omp parallel for if (n - kk > 256)
        for (int64_t c = kk + 1; c < n; c++) {
            double complex f = A[kk + n * c];
            if (f != 0)
                for (int64_t i = kk + 1; i < n; i++) A[i + n * c] -= A[i + n * kk] * f;
        }
    }
    if (!info) {
        for (int64_t r = 0; r < nrhs; r++) {
            double complex* b = B + n * r;
            for (int64_t kk = 0; kk < n; kk++)
                if (piv[kk] != kk) { double complex t = b[kk]; b[kk] = b[piv[kk]]; b[piv[kk]] = t; }
            for (int64_t c = 0; c < n; c++) /* L y = b, unit lower */
                for (int64_t i = c + 1; i < n; i++) b[i] -= A[i + n * c] * b[c];
            for (int64_t c = n - 1; c >= 0; c--) { /* U x = y */
                b[c] /= A[c + n * c];
                for (int64_t i = 0; i < c; i++) b[i] -= A[i + n * c] * b[c];
            }
        }
    }
    free(A); free(piv);
    return info;
}

/* EvalKernel<SingleLayer> with Gauss1D(N): vals[p] = sum_j coef[j] sum_t (i/4)H0(k|P-y_t|) s_t w_t
 * (lib/bem2d_quadrature.h:184-228).  EvalSol returns u_inc - vals (soundsoftscattering.h:149-168). */
int oracle_eval_sl_ck(int64_t n, const int32_t* type, const double* param, int N, double k,
                      double ki, const double* coef, const double* pts, int64_t npts, double* out)
{
    geom_t g = {type, param, NULL, NULL, NULL, n};
    double* gx = malloc(8 * N);
    double* gw = malloc(8 * N);
    oracle_gauss(N, gx, gw);
    const double complex* c = (const double complex*)coef;
    double complex* o = (double complex*)out;
#pragma omp parallel for
    for (int64_t p = 0; p < npts; p++) {
        pt P = {pts[2 * p], pts[2 * p + 1]};
        double complex v = 0;
        for (int64_t j = 0; j < n; j++)
            for (int t = 0; t < N; t++) {
                pt M, D;
                elem_eval(&g, j, gx[t], &M, &D);
                double s = plen(D);
                double complex gv = kernel_eval(K_SL, k, ki, 0, P, M, pnormal(D), pnormal(D));
                v += c[j] * gv * s * gw[t];
            }
        o[p] = v;
    }
    free(gx); free(gw);
    return 0;
}

int oracle_eval_sl(int64_t n, const int32_t* type, const double* param, int N, double k,
                   const double* coef, const double* pts, int64_t npts, double* out)
{
    return oracle_eval_sl_ck(n, type, param, N, k, 0.0, coef, pts, npts, out);
}

/* Far field of u = u_inc - S phi:  u_inf(theta) = -(e^{i pi/4}/sqrt(8 pi k)) sum_j c_j
 * sum_t e^{-ik xhat.y_t} s_t w_t,  xhat = (cos theta, sin theta).  out[a + nang*r]. */
int oracle_farfield(int64_t n, const int32_t* type, const double* param, int N, double k,
                    const double* coef, int nrhs, const double* theta, int64_t nang, double* out)
{
    geom_t g = {type, param, NULL, NULL, NULL, n};
    double* gx = malloc(8 * N);
    double* gw = malloc(8 * N);
    oracle_gauss(N, gx, gw);
    const double complex* c = (const double complex*)coef;
    double complex* o = (double complex*)out;
    double complex pref = -cexp(I * O_PI / 4.0) / sqrt(8.0 * O_PI * k);
#pragma omp parallel for collapse(2)
    for (int r = 0; r < nrhs; r++)
        for (int64_t a = 0; a < nang; a++) {
            double cx = cos(theta[a]), cy = sin(theta[a]);
            double complex v = 0;
            for (int64_t j = 0; j < n; j++) {
                double complex sj = 0;
                for (int t = 0; t < N; t++) {
                    pt M, D;
                    elem_eval(&g, j, gx[t], &M, &D);
                    sj += cexp(-I * k * (cx * M.x + cy * M.y)) * plen(D) * gw[t];
                }
                v += c[j + n * r] * sj;
            }
            o[a + nang * r] = pref * v;
        }
    free(gx); free(gw);
    return 0;
}

/* Element geometry sampled at local parameters (used by tests to build closed-form checks). */
void oracle_elem_eval(int64_t n, const int32_t* type, const double* param, int64_t e, double t,
                      double* out6 /* Mx My Dx Dy nx ny */)
{
    geom_t g = {type, param, NULL, NULL, NULL, n};
    pt M, D;
    elem_eval(&g, e, t, &M, &D);
    pt nn = pnormal(D);
    out6[0] = M.x; out6[1] = M.y; out6[2] = D.x; out6[3] = D.y; out6[4] = nn.x; out6[5] = nn.y;
}

/* ------------------------------------------------------------------------------------------ */
/* Several basis functions per element (lib/bem2d_basistypes.h:16-99, lib/bem2d_geometry.h:48-57) */
/* Basis kinds: 0 ConstBasis (1), 1 PolBasis(deg) (t^deg), 2 LinBasis(sign) (t for sign 1, 1-t    */
/* for sign -1; the sign travels in deg), 3 WaveBasis(dir, k) (exp(i k dir t)), 4                 */
/* WavePolBasis(deg, dir, k) (exp(i k dir t) t^deg), at the element-local parameter t.  Every      */
/* element carries the same nb bases, so FlatMap's order is row = element * nb + basis (the       */
/* order AddBasis gives when each basis is added to all elements, lib/bem2d_geometry.h:33-39).     */
/* ------------------------------------------------------------------------------------------ */
enum { B_CONST = 0, B_POL = 1, B_LIN = 2, B_WAVE = 3, B_WAVEPOL = 4 };
typedef struct {
    int kind, deg;
    double dir, kr, ki;
} basis_t;

static double complex basis_eval(const basis_t* b, double t)
{
    double r = 1;
    switch (b->kind) {
    case B_CONST:
        return 1.0;
    case B_POL:
        for (int i = 0; i < b->deg; i++) r = r * t;
        return r;
    case B_LIN:
        return b->deg == -1 ? 1 - t : t;
    case B_WAVE:
        return cexp(I * (b->kr + I * b->ki) * b->dir * t);
    case B_WAVEPOL:
        for (int i = 0; i < b->deg; i++) r = r * t;
        return cexp(I * (b->kr + I * b->ki) * b->dir * t) * r;
    }
    return NAN;
}

static void bases_make(basis_t* out, int nb, const int32_t* kind, const int32_t* deg,
                       const double* dir, const double* kk)
{
    for (int b = 0; b < nb; b++) {
        out[b].kind = kind[b];
        out[b].deg = deg[b];
        out[b].dir = dir[b];
        out[b].kr = kk[2 * b];
        out[b].ki = kk[2 * b + 1];
    }
}

/* Integrate<kernel,T> with basis functions (lib/bem2d_quadrature.h:37-65). */
static double complex integrate_pair_b(const geom_t* g, int64_t i, int64_t j, const basis_t* fa,
                                       const basis_t* fb, int kind, double k, double ki,
                                       double eta, const double* rx, const double* ry,
                                       const double* rw, int m)
{
    double complex res = 0;
    for (int q = 0; q < m; q++) {
        pt Mx, Dx, My, Dy;
        elem_eval(g, i, rx[q], &Mx, &Dx);
        elem_eval(g, j, ry[q], &My, &Dy);
        double complex gv = kernel_eval(kind, k, ki, eta, Mx, My, pnormal(Dx), pnormal(Dy));
        double complex f1 = conj(basis_eval(fa, rx[q]));
        double complex f2 = basis_eval(fb, ry[q]);
        double s1 = plen(Dx), s2 = plen(Dy);
        res += gv * f1 * f2 * s1 * s2 * rw[q];
    }
    return res;
}

/* IntegrateIdent with basis functions (lib/bem2d_quadrature.h:67-109). */
static double complex integrate_ident_b(const geom_t* g, int64_t i, int64_t j, const basis_t* fa,
                                        const basis_t* fb, int kind, double k, double ki,
                                        double eta, const double* rx, const double* ry,
                                        const double* rw, int m)
{
    double complex res = 0;
    for (int q = 0; q < m; q++) {
        double xq = rx[q], yq = ry[q];
        pt Mx, Dx, My, Dy;
        double s2p = (1 - yq) * xq;
        elem_eval(g, i, xq, &Mx, &Dx);
        elem_eval(g, j, s2p, &My, &Dy);
        double complex gv = kernel_eval(kind, k, ki, eta, Mx, My, pnormal(Dx), pnormal(Dy));
        double complex f1 = conj(basis_eval(fa, xq)), f2 = basis_eval(fb, s2p);
        res += xq * gv * f1 * f2 * plen(Dx) * plen(Dy) * rw[q];

        double s1b = 1 - xq, s2b = 1 - xq + yq * xq;
        elem_eval(g, i, s1b, &Mx, &Dx);
        elem_eval(g, j, s2b, &My, &Dy);
        gv = kernel_eval(kind, k, ki, eta, Mx, My, pnormal(Dx), pnormal(Dy));
        f1 = conj(basis_eval(fa, s1b));
        f2 = basis_eval(fb, s2b);
        res += xq * gv * f1 * f2 * plen(Dx) * plen(Dy) * rw[q];
    }
    return res;
}

/* Entry (row, col) of mass_scale*M + kernel_scale*K with nb bases per element: the rule selection
 * of lib/bem2d_quadrature.h:271-287 by the two elements, EvalIdent's entry
 * sum_t |Gamma_j'(x_t)| conj(f_a(x_t)) f_b(x_t) w_t for elements with the same index
 * (lib/bem2d_quadrature.cpp:108-123). */
static double complex entry_b(const geom_t* g, const rules_t* r, const basis_t* bs, int nb,
                              int64_t row, int64_t col, int kind, double k, double ki, double eta,
                              double mass_scale, double kernel_scale)
{
    const int64_t i = row / nb, j = col / nb;
    const basis_t* fa = bs + row % nb;
    const basis_t* fb = bs + col % nb;
    double complex kv = 0, mv = 0;
    int ident = g->index[i] == g->index[j];
    if (kind != K_NONE && kernel_scale != 0.0) {
        if (ident)
            kv = integrate_ident_b(g, i, j, fa, fb, kind, k, ki, eta, r->x1, r->y1, r->w1, r->m1);
        else if (g->next[i] == g->index[j])
            kv = integrate_pair_b(g, i, j, fa, fb, kind, k, ki, eta, r->x3, r->y3, r->w3, r->m3);
        else if (g->prev[i] == g->index[j])
            kv = integrate_pair_b(g, i, j, fa, fb, kind, k, ki, eta, r->x2, r->y2, r->w2, r->m2);
        else
            kv = integrate_pair_b(g, i, j, fa, fb, kind, k, ki, eta, r->x0, r->y0, r->w0, r->m0);
    }
    if (ident && mass_scale != 0.0) {
        for (int t = 0; t < r->N; t++) {
            pt M, D;
            elem_eval(g, j, r->gx[t], &M, &D);
            double s = plen(D);
            mv += s * conj(basis_eval(fa, r->gx[t])) * basis_eval(fb, r->gx[t]) * r->gw[t];
        }
    }
    return kernel_scale * kv + mass_scale * mv;
}

/* The (n nb) x (n nb) matrix (columns cols[0..ncols) of it, all when cols == NULL), column c of
 * out = matrix column cols[c], leading dimension ld; real k. */
int oracle_assemble_basis(int64_t n, const int32_t* type, const double* param,
                          const int64_t* index, const int64_t* next, const int64_t* prev, int nb,
                          const int32_t* bkind, const int32_t* bdeg, const double* bdir,
                          const double* bk, int N, int L, double sigma, int kind, double k,
                          double ki, double eta, double mass_scale, double kernel_scale,
                          const int64_t* cols, int64_t ncols, double* out, int64_t ld)
{
    if (nb < 1 || nb > 64) return 2;
    geom_t g = {type, param, index, next, prev, n};
    rules_t r;
    if (rules_make(&r, N, L, sigma)) return 1;
    basis_t bs[64];
    bases_make(bs, nb, bkind, bdeg, bdir, bk);
    const int64_t nn = n * nb;
    if (!cols) ncols = nn;
    double complex* o = (double complex*)out;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < ncols; c++) {
        int64_t col = cols ? cols[c] : c;
        for (int64_t row = 0; row < nn; row++)
            o[row + ld * c] = entry_b(&g, &r, bs, nb, row, col, kind, k, ki, eta, mass_scale,
                                      kernel_scale);
    }
    rules_free(&r);
    return 0;
}

/* DotProdBasFuns with basis functions (lib/bem2d_quadrature.h:111-128): out[(i nb + a) + ld r]. */
int oracle_rhs_basis(int64_t n, const int32_t* type, const double* param, int nb,
                     const int32_t* bkind, const int32_t* bdeg, const double* bdir,
                     const double* bk, int N, int kind, double k, double ki, double eta,
                     const double* dirs, int nrhs, double scale, double* out, int64_t ld)
{
    if (nb < 1 || nb > 64) return 2;
    geom_t g = {type, param, NULL, NULL, NULL, n};
    basis_t bs[64];
    bases_make(bs, nb, bkind, bdeg, bdir, bk);
    double* gx = malloc(8 * N);
    double* gw = malloc(8 * N);
    oracle_gauss(N, gx, gw);
    double complex* o = (double complex*)out;
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < nb; a++)
            for (int r = 0; r < nrhs; r++) {
                pt dir = {dirs[2 * r], dirs[2 * r + 1]};
                double complex res = 0;
                for (int t = 0; t < N; t++) {
                    pt M, D;
                    elem_eval(&g, i, gx[t], &M, &D);
                    double s = plen(D);
                    double complex tf = conj(basis_eval(bs + a, gx[t]));
                    res += field_eval(kind, k, ki, eta, dir, M, pnormal(D)) * s * tf * gw[t];
                }
                o[(i * nb + a) + ld * r] = scale * res;
            }
    free(gx); free(gw);
    return 0;
}

/* EvalKernel<SingleLayer> with basis functions (lib/bem2d_quadrature.h:184-228). */
int oracle_eval_sl_basis(int64_t n, const int32_t* type, const double* param, int nb,
                         const int32_t* bkind, const int32_t* bdeg, const double* bdir,
                         const double* bk, int N, double k, double ki, const double* coef,
                         const double* pts, int64_t npts, double* out)
{
    if (nb < 1 || nb > 64) return 2;
    geom_t g = {type, param, NULL, NULL, NULL, n};
    basis_t bs[64];
    bases_make(bs, nb, bkind, bdeg, bdir, bk);
    double* gx = malloc(8 * N);
    double* gw = malloc(8 * N);
    oracle_gauss(N, gx, gw);
    const double complex* c = (const double complex*)coef;
    double complex* o = (double complex*)out;
#pragma omp parallel for
    for (int64_t p = 0; p < npts; p++) {
        pt P = {pts[2 * p], pts[2 * p + 1]};
        double complex v = 0;
        for (int64_t j = 0; j < n; j++)
            for (int b = 0; b < nb; b++)
                for (int t = 0; t < N; t++) {
                    pt M, D;
                    elem_eval(&g, j, gx[t], &M, &D);
                    double s = plen(D);
                    double complex gv = kernel_eval(K_SL, k, ki, 0, P, M, pnormal(D), pnormal(D));
                    v += c[j * nb + b] * gv * s * basis_eval(bs + b, gx[t]) * gw[t];
                }
        o[p] = v;
    }
    free(gx); free(gw);
    return 0;
}

/* Far field of u = u_inc - S phi with basis functions (as oracle_farfield). */
int oracle_farfield_basis(int64_t n, const int32_t* type, const double* param, int nb,
                          const int32_t* bkind, const int32_t* bdeg, const double* bdir,
                          const double* bk, int N, double k, const double* coef, int nrhs,
                          const double* theta, int64_t nang, double* out)
{
    if (nb < 1 || nb > 64) return 2;
    geom_t g = {type, param, NULL, NULL, NULL, n};
    basis_t bs[64];
    bases_make(bs, nb, bkind, bdeg, bdir, bk);
    double* gx = malloc(8 * N);
    double* gw = malloc(8 * N);
    oracle_gauss(N, gx, gw);
    const double complex* c = (const double complex*)coef;
    double complex* o = (double complex*)out;
    double complex pref = -cexp(I * O_PI / 4.0) / sqrt(8.0 * O_PI * k);
    const int64_t nn = n * nb;
#pragma omp parallel for collapse(2)
    for (int r = 0; r < nrhs; r++)
        for (int64_t a = 0; a < nang; a++) {
            double cx = cos(theta[a]), cy = sin(theta[a]);
            double complex v = 0;
            for (int64_t j = 0; j < n; j++)
                for (int b = 0; b < nb; b++) {
                    double complex sj = 0;
                    for (int t = 0; t < N; t++) {
                        pt M, D;
                        elem_eval(&g, j, gx[t], &M, &D);
                        sj += cexp(-I * k * (cx * M.x + cy * M.y)) * plen(D) *
                              basis_eval(bs + b, gx[t]) * gw[t];
                    }
                    v += c[(j * nb + b) + nn * r] * sj;
                }
            o[a + nang * r] = pref * v;
        }
    free(gx); free(gw);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Arc-length parameterisation of a curve (AnalyticCurve, lib/bem2d_shape.cpp:247-368, and      */
/* Curve::Length, lib/bem2d_curve.cpp:11-25), restated without GSL: the arc length            */
/* s(t) = int_0^t |Gamma'(u)| du by adaptive Gauss-Legendre (20 points per panel, panels halved */
/* until two levels agree to 1e-15 relative), and t_i with s(t_i) = i L / n by Newton's method  */
/* on s(t) - s_i (s' = |Gamma'|), bracketed by bisection.  The reference integrates            */
/* dt/ds = 1/|Gamma'| with GSL's adaptive RK2 to an absolute tolerance of 1e-10 and takes L by  */
/* GSL qag to 1e-4 relative; this restatement computes the exact parameterisation instead.     */
/* ------------------------------------------------------------------------------------------ */
static double curve_speed(int type, const double* c, double t)
{
    pt M, D;
    curve_eval(type, c, t, &M, &D);
    return plen(D);
}

static double gl20(int type, const double* c, double a, double b)
{
    static double x[20], w[20];
    static int ready = 0;
    if (!ready) { /* benign race: every thread writes the same values */
        double gx[20], gw[20];
        oracle_gauss(20, gx, gw);
        for (int i = 0; i < 20; i++) { x[i] = gx[i]; w[i] = gw[i]; }
        ready = 1;
    }
    double s = 0;
    for (int i = 0; i < 20; i++) s += w[i] * curve_speed(type, c, a + (b - a) * x[i]);
    return s * (b - a);
}

static double arc_adapt(int type, const double* c, double a, double b, double whole, int depth)
{
    const double m = 0.5 * (a + b);
    const double l = gl20(type, c, a, m), r = gl20(type, c, m, b);
    if (depth >= 30 || fabs(l + r - whole) <= 1e-15 * fabs(l + r)) return l + r;
    return arc_adapt(type, c, a, m, l, depth + 1) + arc_adapt(type, c, m, b, r, depth + 1);
}

static double arc_length(int type, const double* c, double a, double b)
{
    if (b <= a) return 0.0;
    /* 16 panels first, so that a periodic integrand is resolved before the adaptive test */
    double s = 0;
    for (int p = 0; p < 16; p++) {
        const double pa = a + (b - a) * p / 16.0, pb = a + (b - a) * (p + 1) / 16.0;
        s += arc_adapt(type, c, pa, pb, gl20(type, c, pa, pb), 0);
    }
    return s;
}

/* Curve length (Curve::Length) of curve type `type` (1..6) with arguments c[4]. */
double oracle_curve_length(int type, const double* c) { return arc_length(type, c, 0.0, 1.0); }

/* The parameters t[0..n] with arc length s(t_i) = i L / n, t_0 = 0, t_n = 1 (AnalyticCurve(n)). */
int oracle_arclength_partition(int type, const double* c, int n, double* t)
{
    if (n < 1) return 1;
    const double L = arc_length(type, c, 0.0, 1.0);
    t[0] = 0.0;
    double tprev = 0.0, sprev = 0.0;
    for (int i = 1; i < n; i++) {
        const double target = i * L / n;
        double lo = tprev, hi = 1.0, x = tprev + (target - sprev) / L;
        if (x <= lo || x >= hi) x = 0.5 * (lo + hi);
        for (int it = 0; it < 100; it++) {
            const double f = sprev + arc_length(type, c, tprev, x) - target;
            if (f > 0) hi = x; else lo = x;
            const double step = f / curve_speed(type, c, x);
            double xn = x - step;
            if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
            if (fabs(xn - x) <= 1e-16 * fabs(x) + 1e-300) { x = xn; break; }
            x = xn;
        }
        sprev += arc_length(type, c, tprev, x);
        tprev = x;
        t[i] = x;
    }
    t[n] = 1.0;
    return 0;
}
