/*
 * parode_oracle.c -- CPU restatement of ParODE's fixed-step LTI solver path.
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  It is the checker: only the
 * tests, the smoke check and the CPU-baseline leg of bench.py load it (through
 * ctypes).  The product library (libParODEShared.so) never links, loads or
 * calls anything in this directory.
 *
 * Two arithmetic modes are restated:
 *
 *  mirror (FP32) -- follows the reference's own rounding:
 *    * M and Bbar from float matrix products with sequential-k float
 *      accumulation and double scalar combinations rounded once to float
 *      (Sources/RungeKutta4thOrder.cpp:36-70, Sources/AdamsMoulton.cpp:238-316);
 *    * the float time grid  t_s = float(t_b) + float(dt) * s   (Kernels/u.cl:8,
 *      Sources/RungeKutta4thOrder.cpp:255-257, Sources/AdamsMoulton.cpp:426-428);
 *    * FP32 leaves v_e = Bbar * [u...]  (Kernels/BTimesURK4thOrder.cl:15-38,
 *      Kernels/BTimesUAdamsMoulton.cl:17-39);
 *    * the exclusive-scan semantics of Sources/GPUFixedStepIterative.cpp:191-451
 *      (x_0 = seed, x_{e+1} = M x_e + v_e), evaluated as the equivalent
 *      sequential recurrence, and separately as the reference's per-element
 *      Blelloch up/down-sweep for small N (or_blelloch_mirror_f32);
 *    * the ABM start-up: three classic RK4 steps on a float state with the
 *      stage combinations formed in double (Sources/AdamsMoulton.cpp:29-104;
 *      odeint runge_kutta4<vector<float>> defaults to value_type double);
 *    * outputs, time vector and layouts (Sources/GPUFixedStepIterative.cpp:
 *      141-173, 401-451; Kernels/SystemOutput.cl:28-58; Sources/GPUODESolver.cpp:36-42).
 *
 *  FP64 -- the same formulas in double, with t_s = t0 + s*h/2 in double; the
 *    precision reference against which maximum relative error is stated.
 *
 * The reference's CPU path (the baseline timed beside the GPU) is restated in
 * or_cpu_rk4_odeint / or_cpu_abm_pece (TestsSources/TestParODEStaticLibCInterface.cpp
 * :158-202 and :336-461).
 *
 * Parity pins (what checks each function; the tests live in tests/test_oracle.py):
 *   or_rk4_build_*       scalar closed form (diagonal A gives the 4th-order
 *                        Taylor polynomial of h*lambda), RK4 stage expansion
 *                        checked against a brute-force classic RK4 step, exact
 *                        power-of-two inputs reproduce the float formula bitwise
 *   or_abm_build_*       PECE expansion: one companion step equals one explicit
 *                        ABM4 predictor-corrector step (coefficients 55,-59,37,-9
 *                        and 9,19,-5,1 over 24, TestAdamsMoulton.cpp:163-164)
 *   or_recurrence_*      u == 0 gives M^k x0 (matrix powers), constant u gives
 *                        the geometric-series closed form
 *   or_blelloch_mirror   equals the sequential recurrence for N = 1..64
 *   or_abm_startup_*     equals a brute-force classic RK4 on the same samples
 *   or_simulate_*        reference scenarios reproduced with glibc rand(),
 *                        passing the reference's own sqrt(sum err^2) < 0.1 rule
 *                        against or_cpu_rk4_odeint / or_cpu_abm_pece
 *   or_simulate_mirror_exact  agrees with or_simulate_mirror (FP32 scan) and
 *                        or_simulate_f64 within their rounding at short
 *                        horizons, and equals the FP64 recurrence driven by the
 *                        mirror's float M, Bbar and samples
 *   or_simulate_*_nsb    batches of NSB elements: with sampled inputs (no
 *                        restart) they equal one batch to rounding, and the
 *                        padded outputs of the power-of-two last batch are
 *                        dropped; with UFunc inputs every batch's samples
 *                        equal the float grid restarted at float(t_b) (checked
 *                        element by element against a direct evaluation)
 *   or_gauss_noise       moments of the generated samples
 *   all modes            golden fixtures (the .npz files of tests/golden) guard against
 *                        regressions once the checks above pass
 * What no check can pin (parity unpinned): the FP32 summation order and FMA
 * contraction inside the reference's OpenCL kernels, the OpenCL runtime's sin/cos
 * ulps, and Boost.uBLAS / Boost.Odeint internals beyond their documented
 * evaluation order (Boost 1.55 is not vendored).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* dense helpers, row-major                                                  */
/* ------------------------------------------------------------------------- */

/* Z = X*Y with float products, sequential-k float accumulation and no fused
 * multiply-add: uBLAS prod() on matrix<float> (RungeKutta4thOrder.cpp:42-46). */
static void mm_f32(int r, int k, int c, const float *X, const float *Y, float *Z)
{
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < c; ++j) {
            float t = 0.0f;
            for (int q = 0; q < k; ++q)
                t += X[(long)i * k + q] * Y[(long)q * c + j];
            Z[(long)i * c + j] = t;
        }
}

static void mm_f64(int r, int k, int c, const double *X, const double *Y, double *Z)
{
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < c; ++j) {
            double t = 0.0;
            for (int q = 0; q < k; ++q)
                t += X[(long)i * k + q] * Y[(long)q * c + j];
            Z[(long)i * c + j] = t;
        }
}

/* ------------------------------------------------------------------------- */
/* RK4 step map  (RungeKutta4thOrder.cpp:36-70)                               */
/* ------------------------------------------------------------------------- */

/* M = I + hA + h^2/2 A^2 + h^3/6 A^3 + h^4/24 A*A^3 ;
 * B1 = (hB + h^2 AB + h^3/2 A^2B + h^4/4 A^3B)/6, B2 = (4hB + 2h^2 AB + h^3/2 A^2B)/6,
 * B3 = hB/6 ; Bbar = [B1|B2|B3] (n x 3m, row-major).
 * Products in float, scalar combinations in double left to right, one rounding. */
OR_EXPORT void or_rk4_build_mirror(int n, int m, const float *A, const float *B,
                                   double h, float *M, float *Bbar)
{
    size_t nn = (size_t)n * n, nm = (size_t)n * (m > 0 ? m : 1);
    float *A2 = malloc(nn * 4), *A3 = malloc(nn * 4), *A4 = malloc(nn * 4);
    float *AB = malloc(nm * 4), *A2B = malloc(nm * 4), *A3B = malloc(nm * 4);
    mm_f32(n, n, n, A, A, A2);          /* A2 = prod(A, A)   :42 */
    mm_f32(n, n, n, A2, A, A3);         /* A3 = prod(A2, A)  :43 */
    mm_f32(n, n, n, A, A3, A4);         /* prod(A, A3)       :57 */
    const double c2 = 0.5 * h * h, c3 = (1.0 / 6.0) * h * h * h,
                 c4 = (1.0 / 24.0) * h * h * h * h;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            size_t q = (size_t)i * n + j;
            double id = (i == j) ? 1.0 : 0.0;
            double v = id + h * (double)A[q];
            v = v + c2 * (double)A2[q];
            v = v + c3 * (double)A3[q];
            v = v + c4 * (double)A4[q];
            M[q] = (float)v;
        }
    if (m > 0) {
        mm_f32(n, n, m, A, B, AB);      /* AB  = prod(A, B)  :44 */
        mm_f32(n, n, m, A2, B, A2B);    /* A2B = prod(A2, B) :45 */
        mm_f32(n, n, m, A3, B, A3B);    /* A3B = prod(A3, B) :46 */
        const double s1 = h * h, s2 = 0.5 * h * h * h, s3 = 0.25 * h * h * h * h;
        const double t0 = 4.0 * h, t1 = 2.0 * h * h, t2 = 0.5 * h * h * h;
        const double u0 = (1.0 / 6.0) * h;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < m; ++j) {
                size_t q = (size_t)i * m + j;
                double b = B[q], ab = AB[q], a2b = A2B[q], a3b = A3B[q];
                double b1 = (1.0 / 6.0) * (((h * b + s1 * ab) + s2 * a2b) + s3 * a3b);
                double b2 = (1.0 / 6.0) * ((t0 * b + t1 * ab) + t2 * a2b);
                double b3 = u0 * b;
                Bbar[(size_t)i * 3 * m + j] = (float)b1;
                Bbar[(size_t)i * 3 * m + m + j] = (float)b2;
                Bbar[(size_t)i * 3 * m + 2 * m + j] = (float)b3;
            }
    }
    free(A2); free(A3); free(A4); free(AB); free(A2B); free(A3B);
}

OR_EXPORT void or_rk4_build_f64(int n, int m, const double *A, const double *B,
                                double h, double *M, double *Bbar)
{
    size_t nn = (size_t)n * n, nm = (size_t)n * (m > 0 ? m : 1);
    double *A2 = malloc(nn * 8), *A3 = malloc(nn * 8), *A4 = malloc(nn * 8);
    double *AB = malloc(nm * 8), *A2B = malloc(nm * 8), *A3B = malloc(nm * 8);
    mm_f64(n, n, n, A, A, A2);
    mm_f64(n, n, n, A2, A, A3);
    mm_f64(n, n, n, A, A3, A4);
    const double c2 = 0.5 * h * h, c3 = (1.0 / 6.0) * h * h * h,
                 c4 = (1.0 / 24.0) * h * h * h * h;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            size_t q = (size_t)i * n + j;
            double v = ((i == j) ? 1.0 : 0.0) + h * A[q];
            v = v + c2 * A2[q];
            v = v + c3 * A3[q];
            v = v + c4 * A4[q];
            M[q] = v;
        }
    if (m > 0) {
        mm_f64(n, n, m, A, B, AB);
        mm_f64(n, n, m, A2, B, A2B);
        mm_f64(n, n, m, A3, B, A3B);
        const double s1 = h * h, s2 = 0.5 * h * h * h, s3 = 0.25 * h * h * h * h;
        const double t0 = 4.0 * h, t1 = 2.0 * h * h, t2 = 0.5 * h * h * h;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < m; ++j) {
                size_t q = (size_t)i * m + j;
                Bbar[(size_t)i * 3 * m + j] =
                    (1.0 / 6.0) * (((h * B[q] + s1 * AB[q]) + s2 * A2B[q]) + s3 * A3B[q]);
                Bbar[(size_t)i * 3 * m + m + j] = (1.0 / 6.0) * ((t0 * B[q] + t1 * AB[q]) + t2 * A2B[q]);
                Bbar[(size_t)i * 3 * m + 2 * m + j] = ((1.0 / 6.0) * h) * B[q];
            }
    }
    free(A2); free(A3); free(A4); free(AB); free(A2B); free(A3B);
}

/* ------------------------------------------------------------------------- */
/* ABM4 companion map  (AdamsMoulton.cpp:238-316)                              */
/* ------------------------------------------------------------------------- */

/* A_0 = I + 28h/24 A + 495h^2/576 A^2, A_1 = -5h/24 A - 531h^2/576 A^2,
 * A_2 = h/24 A + 333h^2/576 A^2, A_3 = -81h^2/576 A^2 ;
 * B_-1 = 9h/24 B, B_0 = 19h/24 B + 495h^2/576 AB, B_1 = -5h/24 B - 531h^2/576 AB,
 * B_2 = h/24 B + 333h^2/576 AB, B_3 = -81h^2/576 AB ;
 * M (4n x 4n) = [0 I 0 0; 0 0 I 0; 0 0 0 I; A_3 A_2 A_1 A_0],
 * Bbar (n x 5m) = [B_3 | B_2 | B_1 | B_0 | B_-1]. */
typedef struct {
    double a0A, a0A2, a1A, a1A2, a2A, a2A2, a3A2;
    double bm1, b0B, b0AB, b1B, b1AB, b2B, b2AB, b3AB;
} abm_coef;

static abm_coef abm_coefs(double h)
{
    abm_coef c;
    c.a0A = 28.0 * h / 24.0;        c.a0A2 = 495.0 * h * h / 576.0;   /* :258-260 */
    c.a1A = -5.0 * h / 24.0;        c.a1A2 = -531.0 * h * h / 576.0;  /* :261 */
    c.a2A = 1.0 * h / 24.0;         c.a2A2 = 333.0 * h * h / 576.0;   /* :262 */
    c.a3A2 = -81.0 * h * h / 576.0;                                    /* :263 */
    c.bm1 = 9.0 * h / 24.0;                                            /* :265 */
    c.b0B = 19.0 * h / 24.0;        c.b0AB = 495.0 * h * h / 576.0;   /* :266 */
    c.b1B = -5.0 * h / 24.0;        c.b1AB = -531.0 * h * h / 576.0;  /* :267 */
    c.b2B = h / 24.0;               c.b2AB = 333.0 * h * h / 576.0;   /* :268 */
    c.b3AB = -81.0 * h * h / 576.0;                                    /* :269 */
    return c;
}

OR_EXPORT void or_abm_build_mirror(int n, int m, const float *A, const float *B,
                                   double h, float *M, float *Bbar)
{
    const int na = 4 * n;
    size_t nn = (size_t)n * n, nm = (size_t)n * (m > 0 ? m : 1);
    float *A2 = malloc(nn * 4), *AB = malloc(nm * 4);
    mm_f32(n, n, n, A, A, A2);           /* :254 */
    abm_coef c = abm_coefs(h);
    memset(M, 0, (size_t)na * na * sizeof(float));
    for (int r = 0; r < n; ++r) {
        for (int q = 0; q < n; ++q) {
            size_t e = (size_t)r * n + q;
            double id = (r == q) ? 1.0 : 0.0, a = A[e], a2 = A2[e];
            float A0 = (float)((id + c.a0A * a) + c.a0A2 * a2);
            float A1 = (float)(c.a1A * a + c.a1A2 * a2);
            float A2v = (float)(c.a2A * a + c.a2A2 * a2);
            float A3 = (float)(c.a3A2 * a2);
            size_t row = (size_t)(3 * n + r) * na;
            M[row + q] = A3;
            M[row + n + q] = A2v;
            M[row + 2 * n + q] = A1;
            M[row + 3 * n + q] = A0;
        }
        M[(size_t)r * na + n + r] = 1.0f;
        M[(size_t)(n + r) * na + 2 * n + r] = 1.0f;
        M[(size_t)(2 * n + r) * na + 3 * n + r] = 1.0f;
    }
    if (m > 0) {
        mm_f32(n, n, m, A, B, AB);       /* :255 */
        for (int r = 0; r < n; ++r)
            for (int q = 0; q < m; ++q) {
                size_t e = (size_t)r * m + q;
                double b = B[e], ab = AB[e];
                float *row = Bbar + (size_t)r * 5 * m;
                row[q] = (float)(c.b3AB * ab);
                row[m + q] = (float)(c.b2B * b + c.b2AB * ab);
                row[2 * m + q] = (float)(c.b1B * b + c.b1AB * ab);
                row[3 * m + q] = (float)(c.b0B * b + c.b0AB * ab);
                row[4 * m + q] = (float)(c.bm1 * b);
            }
    }
    free(A2); free(AB);
}

OR_EXPORT void or_abm_build_f64(int n, int m, const double *A, const double *B,
                                double h, double *M, double *Bbar)
{
    const int na = 4 * n;
    size_t nn = (size_t)n * n, nm = (size_t)n * (m > 0 ? m : 1);
    double *A2 = malloc(nn * 8), *AB = malloc(nm * 8);
    mm_f64(n, n, n, A, A, A2);
    abm_coef c = abm_coefs(h);
    memset(M, 0, (size_t)na * na * sizeof(double));
    for (int r = 0; r < n; ++r) {
        for (int q = 0; q < n; ++q) {
            size_t e = (size_t)r * n + q;
            double id = (r == q) ? 1.0 : 0.0, a = A[e], a2 = A2[e];
            size_t row = (size_t)(3 * n + r) * na;
            M[row + q] = c.a3A2 * a2;
            M[row + n + q] = c.a2A * a + c.a2A2 * a2;
            M[row + 2 * n + q] = c.a1A * a + c.a1A2 * a2;
            M[row + 3 * n + q] = (id + c.a0A * a) + c.a0A2 * a2;
        }
        M[(size_t)r * na + n + r] = 1.0;
        M[(size_t)(n + r) * na + 2 * n + r] = 1.0;
        M[(size_t)(2 * n + r) * na + 3 * n + r] = 1.0;
    }
    if (m > 0) {
        mm_f64(n, n, m, A, B, AB);
        for (int r = 0; r < n; ++r)
            for (int q = 0; q < m; ++q) {
                size_t e = (size_t)r * m + q;
                double b = B[e], ab = AB[e];
                double *row = Bbar + (size_t)r * 5 * m;
                row[q] = c.b3AB * ab;
                row[m + q] = c.b2B * b + c.b2AB * ab;
                row[2 * m + q] = c.b1B * b + c.b1AB * ab;
                row[3 * m + q] = c.b0B * b + c.b0AB * ab;
                row[4 * m + q] = c.bm1 * b;
            }
    }
    free(A2); free(AB);
}

/* ------------------------------------------------------------------------- */
/* inputs                                                                    */
/* ------------------------------------------------------------------------- */

/* float time grid of Kernels/u.cl:8: tStamp = tStart + tStep * nStep with
 * fType tStart = t0, fType tStep = dt (both rounded from double on the host,
 * RungeKutta4thOrder.cpp:255-256) and the multiply-add contracted. */
OR_EXPORT void or_time_grid_f32(double t0, double dt, long s0, long count, float *t)
{
    const float tf = (float)t0, df = (float)dt;
    for (long i = 0; i < count; ++i)
        t[i] = fmaf(df, (float)(s0 + i), tf);
}

/* The synthetic input family used by the configurations, in the float form
 * that is registered as a UFunc source string:
 *   g = fma(w, t, phi);  u = fma(d, t, fma(b, cos(g), fma(a, sin(g), c)))
 * coef is m x 6 = (a, b, c, d, w, phi) per input.  U is S x m. */
OR_EXPORT void or_ufunc_eval_f32(int m, const float *coef, long S, const float *t, float *U)
{
    for (long s = 0; s < S; ++s)
        for (int j = 0; j < m; ++j) {
            const float *k = coef + 6 * j;
            float g = fmaf(k[4], t[s], k[5]);
            U[s * m + j] = fmaf(k[3], t[s], fmaf(k[1], cosf(g), fmaf(k[0], sinf(g), k[2])));
        }
}

static inline uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Deterministic N(0,1) sample for (seed, sample index s, input j):
 * Box-Muller on two splitmix64 draws of a counter keyed by (seed, s, j). */
OR_EXPORT double or_gauss_noise(uint64_t seed, long s, int j)
{
    uint64_t st = seed * 0xD1B54A32D192ED03ull + (uint64_t)s * 0x9E3779B97F4A7C15ull
                  + (uint64_t)j * 0xA24BAED4963EE407ull;
    double u1 = ((splitmix64(&st) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    double u2 = (splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* FP64 form of the family on t_s = t0 + (s0 + s)*dt, plus sigma * N(0,1) noise
 * keyed by the absolute sample index s0 + s. */
OR_EXPORT void or_ufunc_eval_f64(int m, const double *coef, double t0, double dt,
                                 long s0, long S, double sigma, uint64_t seed, double *U)
{
    for (long s = 0; s < S; ++s) {
        double t = t0 + (double)(s0 + s) * dt;
        for (int j = 0; j < m; ++j) {
            const double *k = coef + 6 * j;
            double g = fma(k[4], t, k[5]);
            double u = fma(k[3], t, fma(k[1], cos(g), fma(k[0], sin(g), k[2])));
            if (sigma != 0.0) u += sigma * or_gauss_noise(seed, s0 + s, j);
            U[s * m + j] = u;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* leaves  v_e = Bbar * [stage inputs]                                        */
/* ------------------------------------------------------------------------- */

/* RK4 (BTimesURK4thOrder.cl:15-38): v_e[r] = sum_{c<3m} Bbar[r][c] U[2e*m + c],
 * i.e. B1 u(t_e) + B2 u(t_e + h/2) + B3 u(t_e + h).  U holds 2N+1 samples. */
OR_EXPORT void or_rk4_leaves_f32(int n, int m, const float *Bbar, long N, const float *U, float *V)
{
    for (long e = 0; e < N; ++e)
        for (int r = 0; r < n; ++r) {
            float t = 0.0f;
            for (int c = 0; c < 3 * m; ++c)
                t += Bbar[(size_t)r * 3 * m + c] * U[2 * e * m + c];
            V[e * n + r] = t;
        }
}

OR_EXPORT void or_rk4_leaves_f64(int n, int m, const double *Bbar, long N, const double *U, double *V)
{
    for (long e = 0; e < N; ++e)
        for (int r = 0; r < n; ++r) {
            double t = 0.0;
            for (int c = 0; c < 3 * m; ++c)
                t += Bbar[(size_t)r * 3 * m + c] * U[2 * e * m + c];
            V[e * n + r] = t;
        }
}

/* ABM (BTimesUAdamsMoulton.cl:17-39 with the zero fill of AdamsMoulton.cpp:419):
 * v_e = [0; 0; 0; Bbar * [u_e .. u_{e+4}]] (4n), U on the h grid from t_0 - 3h. */
OR_EXPORT void or_abm_leaves_f32(int n, int m, const float *Bbar, long N, const float *U, float *V)
{
    const int na = 4 * n;
    for (long e = 0; e < N; ++e) {
        memset(V + e * na, 0, (size_t)3 * n * sizeof(float));
        for (int r = 0; r < n; ++r) {
            float t = 0.0f;
            for (int c = 0; c < 5 * m; ++c)
                t += Bbar[(size_t)r * 5 * m + c] * U[e * m + c];
            V[e * na + 3 * n + r] = t;
        }
    }
}

OR_EXPORT void or_abm_leaves_f64(int n, int m, const double *Bbar, long N, const double *U, double *V)
{
    const int na = 4 * n;
    for (long e = 0; e < N; ++e) {
        memset(V + e * na, 0, (size_t)3 * n * sizeof(double));
        for (int r = 0; r < n; ++r) {
            double t = 0.0;
            for (int c = 0; c < 5 * m; ++c)
                t += Bbar[(size_t)r * 5 * m + c] * U[e * m + c];
            V[e * na + 3 * n + r] = t;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* recurrences: the value of the reference's exclusive scan                  */
/* ------------------------------------------------------------------------- */

/* X[0] = x0 ; X[e+1] = (M X[e]) + V[e]  for e = 0..N-2  (N states out).
 * This is what GPUFixedStepIterative.cpp:191-398 computes: the seed (I, x0)
 * enters at the root (:309-332) and element e receives the composition of
 * leaves 0..e-1.  Row sums are sequential in the column index. */
OR_EXPORT void or_recurrence_f32(int na, const float *M, long N, const float *V,
                                 const float *x0, float *X)
{
    float *MT = malloc((size_t)na * na * sizeof(float));
    for (int r = 0; r < na; ++r)
        for (int c = 0; c < na; ++c) MT[(size_t)c * na + r] = M[(size_t)r * na + c];
    memcpy(X, x0, (size_t)na * sizeof(float));
    float *acc = malloc((size_t)na * sizeof(float));
    for (long e = 0; e + 1 < N; ++e) {
        const float *x = X + e * na;
        for (int r = 0; r < na; ++r) acc[r] = 0.0f;
        for (int c = 0; c < na; ++c) {
            const float xc = x[c], *col = MT + (size_t)c * na;
            for (int r = 0; r < na; ++r) acc[r] += col[r] * xc;
        }
        float *xn = X + (e + 1) * na;
        for (int r = 0; r < na; ++r) xn[r] = acc[r] + V[e * na + r];
    }
    free(acc); free(MT);
}

OR_EXPORT void or_recurrence_f64(int na, const double *M, long N, const double *V,
                                 const double *x0, double *X)
{
    double *MT = malloc((size_t)na * na * sizeof(double));
    for (int r = 0; r < na; ++r)
        for (int c = 0; c < na; ++c) MT[(size_t)c * na + r] = M[(size_t)r * na + c];
    memcpy(X, x0, (size_t)na * sizeof(double));
    double *acc = malloc((size_t)na * sizeof(double));
    for (long e = 0; e + 1 < N; ++e) {
        const double *x = X + e * na;
        for (int r = 0; r < na; ++r) acc[r] = 0.0;
        for (int c = 0; c < na; ++c) {
            const double xc = x[c], *col = MT + (size_t)c * na;
            for (int r = 0; r < na; ++r) acc[r] += col[r] * xc;
        }
        double *xn = X + (e + 1) * na;
        for (int r = 0; r < na; ++r) xn[r] = acc[r] + V[e * na + r];
    }
    free(acc); free(MT);
}

/* ABM recurrence on the companion form without storing the 4n x 4n matrix:
 * state history x_0..x_{N+2} (n each).  The first three block rows of M are
 * exact shifts and the last block row sums A_3 x_e, A_2 x_{e+1}, A_1 x_{e+2},
 * A_0 x_{e+3} in that column order, which is bitwise the dense row sum.
 * Mblk is the last block row of M (n x 4n).  Xh receives N+3 states. */
static void abm_hist_f32(int n, const float *Mblk, long N, const float *Vlast,
                         const float *X0, float *Xh)
{
    const int na = 4 * n;
    float *MT = malloc((size_t)na * n * sizeof(float));
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < na; ++c) MT[(size_t)c * n + r] = Mblk[(size_t)r * na + c];
    memcpy(Xh, X0, (size_t)na * sizeof(float));
    float *acc = malloc((size_t)n * sizeof(float));
    for (long e = 0; e + 1 < N; ++e) {
        const float *x = Xh + e * n;          /* x_e .. x_{e+3} contiguous */
        for (int r = 0; r < n; ++r) acc[r] = 0.0f;
        for (int c = 0; c < na; ++c) {
            const float xc = x[c], *col = MT + (size_t)c * n;
            for (int r = 0; r < n; ++r) acc[r] += col[r] * xc;
        }
        float *xn = Xh + (e + 4) * n;
        for (int r = 0; r < n; ++r) xn[r] = acc[r] + Vlast[e * n + r];
    }
    free(acc); free(MT);
}

static void abm_hist_f64(int n, const double *Mblk, long N, const double *Vlast,
                         const double *X0, double *Xh)
{
    const int na = 4 * n;
    double *MT = malloc((size_t)na * n * sizeof(double));
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < na; ++c) MT[(size_t)c * n + r] = Mblk[(size_t)r * na + c];
    memcpy(Xh, X0, (size_t)na * sizeof(double));
    double *acc = malloc((size_t)n * sizeof(double));
    for (long e = 0; e + 1 < N; ++e) {
        const double *x = Xh + e * n;
        for (int r = 0; r < n; ++r) acc[r] = 0.0;
        for (int c = 0; c < na; ++c) {
            const double xc = x[c], *col = MT + (size_t)c * n;
            for (int r = 0; r < n; ++r) acc[r] += col[r] * xc;
        }
        double *xn = Xh + (e + 4) * n;
        for (int r = 0; r < n; ++r) xn[r] = acc[r] + Vlast[e * n + r];
    }
    free(acc); free(MT);
}

/* ------------------------------------------------------------------------- */
/* the reference's per-element Blelloch scan, FP32, for small N              */
/* ------------------------------------------------------------------------- */

/* MVV row sum in the order of MVVKernelScanUpsweep.cl:29-48: lane l of a
 * work-group row of width cw accumulates columns l, l+cw, ... ; the cw
 * partial sums are then tree-reduced (stride cw/2, cw/4, ...). */
static float mvv_row(const float *Mrow, const float *v, int na, int cw, float *lanes)
{
    for (int l = 0; l < cw; ++l) {
        float t = 0.0f;
        for (int c0 = 0; c0 < na; c0 += cw)
            if (c0 + l < na) t += Mrow[c0 + l] * v[c0 + l];
        lanes[l] = t;
    }
    for (int s = cw >> 1; s > 0; s >>= 1)
        for (int l = 0; l < s; ++l) lanes[l] += lanes[l + s];
    return lanes[0];
}

/* Leaves (M, V[e]) for e < Np2 (Np2 a power of two); on return V[e] holds the
 * state x_e and tot the final carried accumulator vector.  Follows
 * GPUFixedStepIterative.cpp:236-398 with MM up (M'[R]=M[R]M[L],
 * MMKernelScanUpsweep.cl:24-53), MVV up (v'[R]=M[R]v[L]+v[R]), left-leaf copy,
 * root carry with seed (I, x0) (:294-337), MM down (M'[R]=M[L]M[R]),
 * MVV down (v'[R]=M[L]v[R]+v[L]) and root copy (M'[L]=M[R], v'[L]=v[R]).
 * cw is the MVV work-group width (_nLocalSizeMVVColumn). */
OR_EXPORT void or_blelloch_mirror_f32(int na, long Np2, const float *M, float *V,
                                      const float *x0, int cw, float *tot)
{
    size_t msz = (size_t)na * na;
    float *Ms = malloc(msz * Np2 * sizeof(float)), *Md = malloc(msz * Np2 * sizeof(float));
    float *Vs = malloc((size_t)na * Np2 * sizeof(float)), *Vd = malloc((size_t)na * Np2 * sizeof(float));
    float *lanes = malloc((size_t)(cw > 0 ? cw : 1) * sizeof(float));
    for (long e = 0; e < Np2; ++e) memcpy(Ms + e * msz, M, msz * sizeof(float));   /* DistributeA */
    memcpy(Vs, V, (size_t)na * Np2 * sizeof(float));
    memcpy(Md, Ms, msz * Np2 * sizeof(float));
    memcpy(Vd, Vs, (size_t)na * Np2 * sizeof(float));
    for (long d = 1; d < Np2; d <<= 1) {
        for (long k = 0; k < Np2 / (2 * d); ++k) {
            long R = (k + 1) * 2 * d - 1, L = R - d;
            const float *MR = Ms + R * msz, *ML = Ms + L * msz;
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < na; ++j) {
                    float t = 0.0f;
                    for (int q = 0; q < na; ++q) t += MR[(size_t)i * na + q] * ML[(size_t)q * na + j];
                    Md[R * msz + (size_t)i * na + j] = t;
                }
            for (int i = 0; i < na; ++i)
                Vd[R * na + i] = mvv_row(MR + (size_t)i * na, Vs + L * na, na, cw, lanes) + Vs[R * na + i];
            memcpy(Md + L * msz, ML, msz * sizeof(float));
            memcpy(Vd + L * na, Vs + L * na, (size_t)na * sizeof(float));
        }
        float *t;
        t = Ms; Ms = Md; Md = t;
        t = Vs; Vs = Vd; Vd = t;
    }
    /* root carry with accumulator (I, x0) */
    {
        float *accM = malloc(msz * sizeof(float)), *accV = malloc((size_t)na * sizeof(float));
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < na; ++j) accM[(size_t)i * na + j] = (i == j) ? 1.0f : 0.0f;
        memcpy(accV, x0, (size_t)na * sizeof(float));
        const float *MT = Ms + (Np2 - 1) * msz, *VT = Vs + (Np2 - 1) * na;
        for (int i = 0; i < na; ++i) {
            float t = 0.0f;
            for (int c = 0; c < na; ++c) t += MT[(size_t)i * na + c] * accV[c];
            tot[i] = t + VT[i];
        }
        memcpy(Ms + (Np2 - 1) * msz, accM, msz * sizeof(float));
        memcpy(Vs + (Np2 - 1) * na, accV, (size_t)na * sizeof(float));
        free(accM); free(accV);
    }
    memcpy(Md, Ms, msz * Np2 * sizeof(float));
    memcpy(Vd, Vs, (size_t)na * Np2 * sizeof(float));
    for (long d = Np2 / 2; d > 0; d >>= 1) {
        for (long k = 0; k < Np2 / (2 * d); ++k) {
            long R = (k + 1) * 2 * d - 1, L = R - d;
            const float *MR = Ms + R * msz, *ML = Ms + L * msz;
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < na; ++j) {
                    float t = 0.0f;
                    for (int q = 0; q < na; ++q) t += ML[(size_t)i * na + q] * MR[(size_t)q * na + j];
                    Md[R * msz + (size_t)i * na + j] = t;
                }
            for (int i = 0; i < na; ++i)
                Vd[R * na + i] = mvv_row(ML + (size_t)i * na, Vs + R * na, na, cw, lanes) + Vs[L * na + i];
            memcpy(Md + L * msz, MR, msz * sizeof(float));
            memcpy(Vd + L * na, Vs + R * na, (size_t)na * sizeof(float));
        }
        float *t;
        t = Ms; Ms = Md; Md = t;
        t = Vs; Vs = Vd; Vd = t;
    }
    memcpy(V, Vs, (size_t)na * Np2 * sizeof(float));
    free(Ms); free(Md); free(Vs); free(Vd); free(lanes);
}

/* ------------------------------------------------------------------------- */
/* ABM start-up: three classic RK4 steps  (AdamsMoulton.cpp:29-104)           */
/* ------------------------------------------------------------------------- */

/* dxdt = A x + B u_idx with float accumulation (AdamsMoulton.cpp:15-27). */
static void abm_sys_f32(int n, int m, const float *A, const float *B, const float *x,
                        const float *u, float *dx)
{
    for (int r = 0; r < n; ++r) {
        float f = 0.0f;
        for (int c = 0; c < n; ++c) f += A[(size_t)r * n + c] * x[c];
        for (int c = 0; c < m; ++c) f += B[(size_t)r * m + c] * u[c];
        dx[r] = f;
    }
}

/* One classic RK4 step with the stage combinations formed in double and
 * stored to the float state (odeint runge_kutta4 with value_type double):
 *   xt = x + dh*k1; xt = x + dh*k2; xt = x + dt*k3;
 *   x  = x + dt/6 k1 + dt/3 k2 + dt/3 k3 + dt/6 k4.
 * u0, uh, u1 are the inputs at t, t + dt/2 and t + dt. */
static void rk4_step_f32(int n, int m, const float *A, const float *B, float *x,
                         const float *u0, const float *uh, const float *u1, double dt,
                         float *k1, float *k2, float *k3, float *k4, float *xt)
{
    const double dh = dt / 2.0, dt6 = dt / 6.0, dt3 = dt / 3.0;
    abm_sys_f32(n, m, A, B, x, u0, k1);
    for (int r = 0; r < n; ++r) xt[r] = (float)(1.0 * (double)x[r] + dh * (double)k1[r]);
    abm_sys_f32(n, m, A, B, xt, uh, k2);
    for (int r = 0; r < n; ++r) xt[r] = (float)(1.0 * (double)x[r] + dh * (double)k2[r]);
    abm_sys_f32(n, m, A, B, xt, uh, k3);
    for (int r = 0; r < n; ++r) xt[r] = (float)(1.0 * (double)x[r] + dt * (double)k3[r]);
    abm_sys_f32(n, m, A, B, xt, u1, k4);
    for (int r = 0; r < n; ++r)
        x[r] = (float)((((1.0 * (double)x[r] + dt6 * (double)k1[r]) + dt3 * (double)k2[r])
                        + dt3 * (double)k3[r]) + dt6 * (double)k4[r]);
}

/* Uh: 9 float samples on the h/2 grid from tStart (AdamsMoulton.cpp:53-88),
 * indexed by floor(t/(h/2) + 0.5) = 2i, 2i+1, 2i+1, 2i+2 for step i.
 * X0 (4n) = [x0; x1; x2; x3]. */
OR_EXPORT void or_abm_startup_f32(int n, int m, const float *A, const float *B,
                                  const float *x0, const float *Uh, double h, float *X0)
{
    float *k = malloc((size_t)5 * n * sizeof(float)), *x = malloc((size_t)n * sizeof(float));
    float zero = 0.0f;
    memcpy(x, x0, (size_t)n * sizeof(float));
    memcpy(X0, x0, (size_t)n * sizeof(float));
    for (int i = 0; i < 3; ++i) {
        const float *u0 = m ? Uh + (size_t)(2 * i) * m : &zero;
        const float *uh = m ? Uh + (size_t)(2 * i + 1) * m : &zero;
        const float *u1 = m ? Uh + (size_t)(2 * i + 2) * m : &zero;
        rk4_step_f32(n, m, A, B, x, u0, uh, u1, h, k, k + n, k + 2 * n, k + 3 * n, k + 4 * n);
        memcpy(X0 + (size_t)(i + 1) * n, x, (size_t)n * sizeof(float));
    }
    free(k); free(x);
}

static void abm_sys_f64(int n, int m, const double *A, const double *B, const double *x,
                        const double *u, double *dx)
{
    for (int r = 0; r < n; ++r) {
        double f = 0.0;
        for (int c = 0; c < n; ++c) f += A[(size_t)r * n + c] * x[c];
        for (int c = 0; c < m; ++c) f += B[(size_t)r * m + c] * u[c];
        dx[r] = f;
    }
}

OR_EXPORT void or_abm_startup_f64(int n, int m, const double *A, const double *B,
                                  const double *x0, const double *Uh, double h, double *X0)
{
    double *k1 = malloc((size_t)5 * n * sizeof(double)), *x = malloc((size_t)n * sizeof(double));
    double *k2 = k1 + n, *k3 = k1 + 2 * n, *k4 = k1 + 3 * n, *xt = k1 + 4 * n, zero = 0.0;
    memcpy(x, x0, (size_t)n * sizeof(double));
    memcpy(X0, x0, (size_t)n * sizeof(double));
    for (int i = 0; i < 3; ++i) {
        const double *u0 = m ? Uh + (size_t)(2 * i) * m : &zero;
        const double *uh = m ? Uh + (size_t)(2 * i + 1) * m : &zero;
        const double *u1 = m ? Uh + (size_t)(2 * i + 2) * m : &zero;
        abm_sys_f64(n, m, A, B, x, u0, k1);
        for (int r = 0; r < n; ++r) xt[r] = x[r] + (h / 2.0) * k1[r];
        abm_sys_f64(n, m, A, B, xt, uh, k2);
        for (int r = 0; r < n; ++r) xt[r] = x[r] + (h / 2.0) * k2[r];
        abm_sys_f64(n, m, A, B, xt, uh, k3);
        for (int r = 0; r < n; ++r) xt[r] = x[r] + h * k3[r];
        abm_sys_f64(n, m, A, B, xt, u1, k4);
        for (int r = 0; r < n; ++r)
            x[r] = x[r] + (h / 6.0) * (k1[r] + 2.0 * k2[r] + 2.0 * k3[r] + k4[r]);
        memcpy(X0 + (size_t)(i + 1) * n, x, (size_t)n * sizeof(double));
    }
    free(k1); free(x);
}

/* ------------------------------------------------------------------------- */
/* whole-solver restatements                                                 */
/* ------------------------------------------------------------------------- */

/* Step counts (RungeKutta4thOrder.cpp:20; AdamsMoulton.cpp:356-358).
 * Returns the number of output points, or -1 when the ABM horizon is shorter
 * than its three start-up steps (the reference's unsigned cast is undefined there). */
OR_EXPORT long or_step_count(int solver, double tStart, double tEnd, double h)
{
    if (solver == 0) {
        double q = (tEnd - tStart) / h;
        if (q < 0.0) return -1;
        return (long)(unsigned long)q + 1;
    }
    double ts = tStart + 3.0 * h;
    double q = (tEnd - ts) / h;
    if (q < 0.0) return -1;
    return (long)(unsigned long)q + 1 + 3;
}

/* tVect (GPUFixedStepIterative.cpp:141-146): lead entries tStart' - (lead-i)h,
 * then tVect[i] = (float)((double)tVect[i-1] + h). */
OR_EXPORT void or_time_vector(int solver, double tStart, double h, long total, float *tv)
{
    long lead = solver == 0 ? 0 : 3;
    double ts = solver == 0 ? tStart : tStart + 3.0 * h;
    for (long i = 0; i < lead && i < total; ++i)
        tv[i] = (float)(ts - (double)(unsigned long)(lead - i) * h);
    if (total <= lead) return;
    tv[lead] = (float)ts;
    for (long i = 1; i < total - lead; ++i)
        tv[lead + i] = (float)((double)tv[lead + i - 1] + h);
}

/* y = C x + D u with float accumulation (Kernels/SystemOutput.cl:34-58, summed
 * sequentially); with C and D both absent the state itself is returned. */
static void outputs_f32(int n, int m, int p, const float *C, const float *D, int zC, int zD,
                        const float *x, const float *u, float *y)
{
    if (zC && zD) { memcpy(y, x, (size_t)n * sizeof(float)); return; }
    for (int j = 0; j < p; ++j) {
        float t = 0.0f;
        if (!zC) for (int c = 0; c < n; ++c) t += C[(size_t)j * n + c] * x[c];
        if (!zD) for (int c = 0; c < m; ++c) t += D[(size_t)j * m + c] * u[c];
        y[j] = t;
    }
}

static void outputs_f64(int n, int m, int p, const double *C, const double *D, int zC, int zD,
                        const double *x, const double *u, double *y)
{
    if (zC && zD) { memcpy(y, x, (size_t)n * sizeof(double)); return; }
    for (int j = 0; j < p; ++j) {
        double t = 0.0;
        if (!zC) for (int c = 0; c < n; ++c) t += C[(size_t)j * n + c] * x[c];
        if (!zD) for (int c = 0; c < m; ++c) t += D[(size_t)j * m + c] * u[c];
        y[j] = t;
    }
}

/* ------------------------------------------------------------------------- */
/* batches of the reference's driver                                         */
/* ------------------------------------------------------------------------- */

/* The reference runs the scan in batches of NSB elements (a power of two; the
 * last batch is padded up to the next power of two and its padded outputs are
 * dropped, GPUFixedStepIterative.cpp:191-221, :443).  Batch b starts at
 * t_b = tStart + (b NSB) h in double (:224) and samples its inputs on a float
 * grid that restarts there:
 *   RK4: u_b[s] = UFunc(float(t_b) + float(h/2) s), s = 0 .. 2 cnt
 *        (RungeKutta4thOrder.cpp:255-257, Kernels/u.cl:8);
 *   ABM: u_b[s] = UFunc(float(t_b - 3h) + float(h) s), s = 0 .. cnt + 3
 *        (AdamsMoulton.cpp:426-428; tStart already shifted by 3h, :356).
 * Element e = b NSB + i takes its leaf and output inputs from u_b at local
 * index i.  The carried accumulator (:294-337) continues the recurrence
 * across batches, so the state sequence is that of one scan; only the input
 * samples depend on the batching.  Sampled inputs (Uhalf on the h/2 grid from
 * tStart) are indexed globally and do not restart.
 * nsb <= 0 or nsb >= N: one batch (what the reference does whenever the
 * horizon fits in device memory, RungeKutta4thOrder.cpp:126-142). */
static long batch_len(long N, long nsb) { return (nsb <= 0 || nsb >= N) ? N : nsb; }

/* samples of the batch starting at element k0 with cnt real elements */
static void batch_samples(int solver, int m, const float *ucoef, const float *Uhalf,
                          double tStart, double h, long k0, long cnt, float *U)
{
    if (!m) return;
    if (solver == 0) {
        const long S = 2 * cnt + 1;
        if (ucoef) {
            float *t = malloc((size_t)S * 4);
            or_time_grid_f32(tStart + (double)k0 * h, h / 2.0, 0, S, t);
            or_ufunc_eval_f32(m, ucoef, S, t, U);
            free(t);
        } else memcpy(U, Uhalf + 2 * k0 * m, (size_t)S * m * 4);
        return;
    }
    const long S = cnt + 4;   /* tStart: the shifted t_0 */
    if (ucoef) {
        float *t = malloc((size_t)S * 4);
        or_time_grid_f32(tStart + (double)k0 * h - 3.0 * h, h, 0, S, t);
        or_ufunc_eval_f32(m, ucoef, S, t, U);
        free(t);
    } else {
        /* h/2-grid samples from the original tStart: u(t_{e}) of the shifted
         * grid's row s is row 2 (k0 + s) */
        for (long s = 0; s < S; ++s)
            memcpy(U + s * m, Uhalf + 2 * (k0 + s) * m, (size_t)m * 4);
    }
}

/* The reference's SimulateODE in FP32-mirror arithmetic with batches of nsb
 * elements (see batch_len).  The input is either the synthetic UFunc family
 * (ucoef, m x 6, evaluated on the reference's float time grids) or, when
 * ucoef is NULL, float samples Uhalf on the h/2 grid from tStart
 * (Uhalf[s] = u(tStart + s h/2), 2*total+1 rows).
 * m = 0 stands for ZeroB (no inputs).  p = rows of C/D (ignored when both are
 * absent; the output is then the state, n wide).  xVect is time-major,
 * xVect[k*nOut + j] (GPUODESolver.cpp:36-42).  Returns the number of points. */
OR_EXPORT long or_simulate_mirror_nsb(int solver, int n, int m, int p,
                                      const float *A, const float *B, const float *C, const float *D,
                                      int zC, int zD, const float *ucoef, const float *Uhalf,
                                      double tStart, double tEnd, double h, const float *x0,
                                      long nsb, float *tVect, float *xVect)
{
    long total = or_step_count(solver, tStart, tEnd, h);
    if (total < 1) return -1;
    int nOut = (zC && zD) ? n : p;
    or_time_vector(solver, tStart, h, total, tVect);
    if (solver == 0) {
        const long N = total, NSB = batch_len(N, nsb);
        float *M = malloc((size_t)n * n * 4), *Bb = malloc((size_t)n * 3 * (m ? m : 1) * 4);
        or_rk4_build_mirror(n, m, A, B, h, M, Bb);
        float *U = malloc((size_t)(2 * NSB + 1) * (m ? m : 1) * 4);
        float *V = calloc((size_t)N * n, 4), *X = malloc((size_t)N * n * 4);
        for (long k0 = 0; k0 < N; k0 += NSB) {
            const long cnt = N - k0 < NSB ? N - k0 : NSB;
            batch_samples(0, m, ucoef, Uhalf, tStart, h, k0, cnt, U);
            if (m) or_rk4_leaves_f32(n, m, Bb, cnt, U, V + k0 * n);
        }
        or_recurrence_f32(n, M, N, V, x0, X);
        for (long k0 = 0; k0 < N; k0 += NSB) {
            const long cnt = N - k0 < NSB ? N - k0 : NSB;
            batch_samples(0, m, ucoef, Uhalf, tStart, h, k0, cnt, U);
            for (long i = 0; i < cnt; ++i)
                outputs_f32(n, m, p, C, D, zC, zD, X + (k0 + i) * n, m ? U + 2 * i * m : NULL,
                            xVect + (k0 + i) * nOut);
        }
        free(M); free(Bb); free(U); free(V); free(X);
        return total;
    }
    /* ABM */
    const long N = total - 3, NSB = batch_len(N, nsb);
    const int na = 4 * n;
    float *M = malloc((size_t)na * na * 4), *Bb = malloc((size_t)n * 5 * (m ? m : 1) * 4);
    or_abm_build_mirror(n, m, A, B, h, M, Bb);
    float Uh0[9 * 64];
    float *Uh = m <= 64 ? Uh0 : malloc((size_t)9 * m * 4);
    const double ts = tStart + 3.0 * h;
    if (m) {
        if (ucoef) {
            float t9[9];
            or_time_grid_f32(tStart, h / 2.0, 0, 9, t9);          /* AdamsMoulton.cpp:64-65 */
            or_ufunc_eval_f32(m, ucoef, 9, t9, Uh);
        } else memcpy(Uh, Uhalf, (size_t)9 * m * 4);
    }
    float *U = malloc((size_t)(NSB + 4) * (m ? m : 1) * 4);
    float *X0 = malloc((size_t)na * 4);
    or_abm_startup_f32(n, m, A, B, x0, Uh, h, X0);
    float *V = calloc((size_t)NSB * na, 4);
    float *Vl = calloc((size_t)N * n, 4), *Xh = malloc((size_t)(N + 3) * n * 4);
    for (long k0 = 0; k0 < N; k0 += NSB) {
        const long cnt = N - k0 < NSB ? N - k0 : NSB;
        batch_samples(1, m, ucoef, Uhalf, ts, h, k0, cnt, U);
        if (m) or_abm_leaves_f32(n, m, Bb, cnt, U, V);
        for (long e = 0; m && e < cnt; ++e)
            memcpy(Vl + (k0 + e) * n, V + e * na + 3 * n, (size_t)n * 4);
    }
    abm_hist_f32(n, M + (size_t)3 * n * na, N, Vl, X0, Xh);
    /* first three outputs from the start-up states and u(t0), u(t0+h), u(t0+2h)
     * (GPUFixedStepIterative.cpp:150-173; AdamsMoulton.cpp:106-116) */
    for (int i = 0; i < 3 && i < total; ++i)
        outputs_f32(n, m, p, C, D, zC, zD, X0 + (size_t)i * n, m ? Uh + (size_t)2 * i * m : NULL,
                    xVect + (size_t)i * nOut);
    /* scan outputs: state block 3 of element e (= x_{e+3}) and input sample e+3
     * of its batch */
    for (long k0 = 0; k0 < N; k0 += NSB) {
        const long cnt = N - k0 < NSB ? N - k0 : NSB;
        batch_samples(1, m, ucoef, Uhalf, ts, h, k0, cnt, U);
        for (long i = 0; i < cnt; ++i)
            outputs_f32(n, m, p, C, D, zC, zD, Xh + (k0 + i + 3) * n, m ? U + (i + 3) * m : NULL,
                        xVect + (3 + k0 + i) * nOut);
    }
    if (Uh != Uh0) free(Uh);
    free(M); free(Bb); free(U); free(X0); free(V); free(Vl); free(Xh);
    return total;
}

OR_EXPORT long or_simulate_mirror(int solver, int n, int m, int p,
                                  const float *A, const float *B, const float *C, const float *D,
                                  int zC, int zD, const float *ucoef, const float *Uhalf,
                                  double tStart, double tEnd, double h, const float *x0,
                                  float *tVect, float *xVect)
{
    return or_simulate_mirror_nsb(solver, n, m, p, A, B, C, D, zC, zD, ucoef, Uhalf, tStart, tEnd,
                                  h, x0, 0, tVect, xVect);
}

/* The reference's algorithm in exact arithmetic on the reference's float
 * inputs: M, Bbar, the input samples (float grids + float UFunc, restarted
 * per batch of nsb elements as in or_simulate_mirror_nsb) and the ABM
 * start-up states are formed exactly as or_simulate_mirror_nsb forms them
 * (FP32), then promoted to double; leaves, the recurrence and the outputs run
 * in FP64.  Any correct evaluation order of the reference's scan on those
 * inputs (its Blelloch tree, a sequential loop, the collapsed power-table
 * scan) agrees with this to O(log N) float roundings, which is what the FP32
 * parity bar measures.  yOut is time-major FP64. */
OR_EXPORT long or_simulate_mirror_exact_nsb(int solver, int n, int m, int p,
                                            const float *A, const float *B, const float *C,
                                            const float *D, int zC, int zD, const float *ucoef,
                                            const float *Uhalf, double tStart, double tEnd,
                                            double h, const float *x0, long nsb, double *yOut)
{
    long total = or_step_count(solver, tStart, tEnd, h);
    if (total < 1) return -1;
    int nOut = (zC && zD) ? n : p;
    const int mm = m ? m : 1;
    double *Cd = malloc((size_t)(p > 0 ? p : 1) * n * 8), *Dd = malloc((size_t)(p > 0 ? p : 1) * mm * 8);
    for (long i = 0; !zC && i < (long)p * n; ++i) Cd[i] = C[i];
    for (long i = 0; !zD && i < (long)p * m; ++i) Dd[i] = D[i];
    if (solver == 0) {
        const long N = total, NSB = batch_len(N, nsb), S = 2 * NSB + 1;
        float *M = malloc((size_t)n * n * 4), *Bb = malloc((size_t)n * 3 * mm * 4);
        or_rk4_build_mirror(n, m, A, B, h, M, Bb);
        double *Md = malloc((size_t)n * n * 8), *Bd = malloc((size_t)n * 3 * mm * 8);
        double *xd = malloc((size_t)n * 8);
        for (long i = 0; i < (long)n * n; ++i) Md[i] = M[i];
        for (long i = 0; i < (long)n * 3 * m; ++i) Bd[i] = Bb[i];
        for (int i = 0; i < n; ++i) xd[i] = x0[i];
        float *U = malloc((size_t)S * mm * 4);
        double *Ud = malloc((size_t)S * mm * 8);
        double *V = calloc((size_t)N * n, 8), *X = malloc((size_t)N * n * 8);
        for (long k0 = 0; m && k0 < N; k0 += NSB) {
            const long cnt = N - k0 < NSB ? N - k0 : NSB;
            batch_samples(0, m, ucoef, Uhalf, tStart, h, k0, cnt, U);
            for (long i = 0; i < (2 * cnt + 1) * m; ++i) Ud[i] = U[i];
            or_rk4_leaves_f64(n, m, Bd, cnt, Ud, V + k0 * n);
        }
        or_recurrence_f64(n, Md, N, V, xd, X);
        for (long k0 = 0; k0 < N; k0 += NSB) {
            const long cnt = N - k0 < NSB ? N - k0 : NSB;
            batch_samples(0, m, ucoef, Uhalf, tStart, h, k0, cnt, U);
            for (long i = 0; m && i < (2 * cnt + 1) * m; ++i) Ud[i] = U[i];
            for (long i = 0; i < cnt; ++i)
                outputs_f64(n, m, p, Cd, Dd, zC, zD, X + (k0 + i) * n, m ? Ud + 2 * i * m : NULL,
                            yOut + (k0 + i) * nOut);
        }
        free(M); free(Bb); free(U); free(Md); free(Bd); free(Ud); free(xd); free(V); free(X);
        free(Cd); free(Dd);
        return total;
    }
    const long N = total - 3, NSB = batch_len(N, nsb), S = NSB + 4;
    const int na = 4 * n;
    float *M = malloc((size_t)na * na * 4), *Bb = malloc((size_t)n * 5 * mm * 4);
    or_abm_build_mirror(n, m, A, B, h, M, Bb);
    float *Uh = malloc((size_t)9 * mm * 4), *U = malloc((size_t)S * mm * 4);
    const double ts = tStart + 3.0 * h;
    if (m) {
        if (ucoef) {
            float t9[9];
            or_time_grid_f32(tStart, h / 2.0, 0, 9, t9);
            or_ufunc_eval_f32(m, ucoef, 9, t9, Uh);
        } else memcpy(Uh, Uhalf, (size_t)9 * m * 4);
    }
    float *X0 = malloc((size_t)na * 4);
    or_abm_startup_f32(n, m, A, B, x0, Uh, h, X0);
    double *Mb = malloc((size_t)n * na * 8), *Bd = malloc((size_t)n * 5 * mm * 8);
    double *Ud = malloc((size_t)S * mm * 8), *Uhd = malloc((size_t)9 * mm * 8);
    double *X0d = malloc((size_t)na * 8);
    for (long i = 0; i < (long)n * na; ++i) Mb[i] = M[(size_t)3 * n * na + i];
    for (long i = 0; i < (long)n * 5 * m; ++i) Bd[i] = Bb[i];
    for (long i = 0; i < 9L * m; ++i) Uhd[i] = Uh[i];
    for (int i = 0; i < na; ++i) X0d[i] = X0[i];
    double *V = calloc((size_t)NSB * na, 8);
    double *Vl = calloc((size_t)N * n, 8), *Xh = malloc((size_t)(N + 3) * n * 8);
    for (long k0 = 0; m && k0 < N; k0 += NSB) {
        const long cnt = N - k0 < NSB ? N - k0 : NSB;
        batch_samples(1, m, ucoef, Uhalf, ts, h, k0, cnt, U);
        for (long i = 0; i < (cnt + 4) * m; ++i) Ud[i] = U[i];
        or_abm_leaves_f64(n, m, Bd, cnt, Ud, V);
        for (long e = 0; e < cnt; ++e) memcpy(Vl + (k0 + e) * n, V + e * na + 3 * n, (size_t)n * 8);
    }
    abm_hist_f64(n, Mb, N, Vl, X0d, Xh);
    for (int i = 0; i < 3; ++i)
        outputs_f64(n, m, p, Cd, Dd, zC, zD, X0d + (size_t)i * n, m ? Uhd + (size_t)2 * i * m : NULL,
                    yOut + (size_t)i * nOut);
    for (long k0 = 0; k0 < N; k0 += NSB) {
        const long cnt = N - k0 < NSB ? N - k0 : NSB;
        batch_samples(1, m, ucoef, Uhalf, ts, h, k0, cnt, U);
        for (long i = 0; m && i < (cnt + 4) * m; ++i) Ud[i] = U[i];
        for (long i = 0; i < cnt; ++i)
            outputs_f64(n, m, p, Cd, Dd, zC, zD, Xh + (k0 + i + 3) * n, m ? Ud + (i + 3) * m : NULL,
                        yOut + (3 + k0 + i) * nOut);
    }
    free(M); free(Bb); free(Uh); free(U); free(X0); free(Mb); free(Bd); free(Ud); free(Uhd);
    free(X0d); free(V); free(Vl); free(Xh); free(Cd); free(Dd);
    return total;
}

OR_EXPORT long or_simulate_mirror_exact(int solver, int n, int m, int p,
                                        const float *A, const float *B, const float *C, const float *D,
                                        int zC, int zD, const float *ucoef, const float *Uhalf,
                                        double tStart, double tEnd, double h, const float *x0,
                                        double *yOut)
{
    return or_simulate_mirror_exact_nsb(solver, n, m, p, A, B, C, D, zC, zD, ucoef, Uhalf, tStart,
                                        tEnd, h, x0, 0, yOut);
}
/* FP64 precision reference.  Uhalf: FP64 samples on the h/2 grid from tStart,
 * 2*total+1 rows; time t_s = tStart + s h/2 is the caller's.  Writes y (time-
 * major, FP64) and, when Xout is not NULL, the states (n per point). */
OR_EXPORT long or_simulate_f64(int solver, int n, int m, int p,
                               const double *A, const double *B, const double *C, const double *D,
                               int zC, int zD, const double *Uhalf,
                               long total, double h, const double *x0, double *yOut, double *Xout)
{
    if (total < 1 || (solver == 1 && total < 4)) return -1;
    int nOut = (zC && zD) ? n : p;
    if (solver == 0) {
        long N = total;
        double *M = malloc((size_t)n * n * 8), *Bb = malloc((size_t)n * 3 * (m ? m : 1) * 8);
        or_rk4_build_f64(n, m, A, B, h, M, Bb);
        double *V = calloc((size_t)N * n, 8), *X = malloc((size_t)N * n * 8);
        if (m) or_rk4_leaves_f64(n, m, Bb, N, Uhalf, V);
        or_recurrence_f64(n, M, N, V, x0, X);
        for (long e = 0; e < N; ++e)
            outputs_f64(n, m, p, C, D, zC, zD, X + e * n, m ? Uhalf + 2 * e * m : NULL, yOut + e * nOut);
        if (Xout) memcpy(Xout, X, (size_t)N * n * 8);
        free(M); free(Bb); free(V); free(X);
        return total;
    }
    long N = total - 3;
    const int na = 4 * n;
    double *M = malloc((size_t)na * na * 8), *Bb = malloc((size_t)n * 5 * (m ? m : 1) * 8);
    or_abm_build_f64(n, m, A, B, h, M, Bb);
    long S = N + 4;
    double *U = malloc((size_t)S * (m ? m : 1) * 8);
    for (long s = 0; m && s < S; ++s) memcpy(U + s * m, Uhalf + 2 * s * m, (size_t)m * 8);
    double *X0 = malloc((size_t)na * 8);
    or_abm_startup_f64(n, m, A, B, x0, Uhalf, h, X0);
    double *V = calloc((size_t)N * na, 8);
    if (m) or_abm_leaves_f64(n, m, Bb, N, U, V);
    double *Vl = malloc((size_t)N * n * 8), *Xh = malloc((size_t)(N + 3) * n * 8);
    for (long e = 0; e < N; ++e) memcpy(Vl + e * n, V + e * na + 3 * n, (size_t)n * 8);
    abm_hist_f64(n, M + (size_t)3 * n * na, N, Vl, X0, Xh);
    for (int i = 0; i < 3; ++i)
        outputs_f64(n, m, p, C, D, zC, zD, X0 + (size_t)i * n, m ? Uhalf + (size_t)2 * i * m : NULL,
                    yOut + (size_t)i * nOut);
    for (long e = 0; e < N; ++e)
        outputs_f64(n, m, p, C, D, zC, zD, Xh + (e + 3) * n, m ? U + (e + 3) * m : NULL,
                    yOut + (3 + e) * nOut);
    if (Xout) memcpy(Xout, Xh, (size_t)(N + 3) * n * 8);
    free(M); free(Bb); free(U); free(X0); free(V); free(Vl); free(Xh);
    return total;
}

/* ------------------------------------------------------------------------- */
/* the reference's CPU path (the baseline)                                   */
/* ------------------------------------------------------------------------- */

/* TestParODEStaticLibCInterface.cpp:158-202: odeint runge_kutta4 on
 * f(t,x) = A x + B u(t) with a float state, then y = C x + D u(t) in float.
 * u comes from float samples on the h/2 grid (Uhalf).  y is time-major; the
 * first point is the initial state's output.  Returns the points produced. */
OR_EXPORT long or_cpu_rk4_odeint(int n, int m, int p, const float *A, const float *B,
                                 const float *C, const float *D, int zC, int zD,
                                 const float *Uhalf, long total, double h,
                                 const float *x0, float *y)
{
    int nOut = (zC && zD) ? n : p;
    float *k = malloc((size_t)5 * n * sizeof(float)), *x = malloc((size_t)n * sizeof(float));
    float zero = 0.0f;
    memcpy(x, x0, (size_t)n * sizeof(float));
    outputs_f32(n, m, p, C, D, zC, zD, x, m ? Uhalf : &zero, y);
    for (long i = 1; i < total; ++i) {
        const float *u0 = m ? Uhalf + (size_t)(2 * (i - 1)) * m : &zero;
        const float *uh = m ? Uhalf + (size_t)(2 * (i - 1) + 1) * m : &zero;
        const float *u1 = m ? Uhalf + (size_t)(2 * i) * m : &zero;
        rk4_step_f32(n, m, A, B, x, u0, uh, u1, h, k, k + n, k + 2 * n, k + 3 * n, k + 4 * n);
        outputs_f32(n, m, p, C, D, zC, zD, x, u1, y + (size_t)i * nOut);
    }
    free(k); free(x);
    return total;
}

/* TestParODEStaticLibCInterface.cpp:336-461: three RK4 steps, then the
 * ABM4 predictor-corrector loop
 *   p  = xc + h (55 f_k - 59 f_{k-1} + 37 f_{k-2} - 9 f_{k-3})/24
 *   xc = xc + h (9 f(p) + 19 f_k - 5 f_{k-1} + f_{k-2})/24
 * with float vectors and double coefficients.  y (time-major) receives the
 * outputs of points 0..total-1 (points 0..3 from the RK4 start). */
OR_EXPORT long or_cpu_abm_pece(int n, int m, int p, const float *A, const float *B,
                               const float *C, const float *D, int zC, int zD,
                               const float *Uhalf, long total, double h,
                               const float *x0, float *y)
{
    int nOut = (zC && zD) ? n : p;
    float zero = 0.0f;
    float *buf = malloc((size_t)12 * n * sizeof(float));
    float *x = buf, *k = buf + n, *db0 = buf + 6 * n, *db1 = buf + 7 * n, *db2 = buf + 8 * n,
          *db3 = buf + 9 * n, *pr = buf + 10 * n, *dp = buf + 11 * n;
    memcpy(x, x0, (size_t)n * sizeof(float));
    const double ab0 = 55.0 / 24.0, ab1 = -59.0 / 24.0, ab2 = 37.0 / 24.0, ab3 = -9.0 / 24.0;
    const double am0 = 9.0 / 24.0, am1 = 19.0 / 24.0, am2 = -5.0 / 24.0, am3 = 1.0 / 24.0;
#define UH(s) (m ? Uhalf + (size_t)(s) * m : &zero)
    outputs_f32(n, m, p, C, D, zC, zD, x, UH(0), y);
    /* start-up: derivatives at t0, t1, t2 and RK4 steps to x3 */
    abm_sys_f32(n, m, A, B, x, UH(0), db2);
    for (int i = 0; i < 3; ++i) {
        if (i == 1) abm_sys_f32(n, m, A, B, x, UH(2), db1);
        if (i == 2) abm_sys_f32(n, m, A, B, x, UH(4), db0);
        rk4_step_f32(n, m, A, B, x, UH(2 * i), UH(2 * i + 1), UH(2 * i + 2), h,
                     k, k + n, k + 2 * n, k + 3 * n, k + 4 * n);
        if (i + 1 < total)
            outputs_f32(n, m, p, C, D, zC, zD, x, UH(2 * i + 2), y + (size_t)(i + 1) * nOut);
    }
    for (long i = 4; i < total; ++i) {
        float *t = db3; db3 = db2; db2 = db1; db1 = db0; db0 = t;
        abm_sys_f32(n, m, A, B, x, UH(2 * (i - 1)), db0);           /* f_k at t_{i-1} */
        for (int r = 0; r < n; ++r)
            pr[r] = (float)((double)x[r] + h * ((((ab0 * (double)db0[r] + ab1 * (double)db1[r])
                                                 + ab2 * (double)db2[r]) + ab3 * (double)db3[r])));
        abm_sys_f32(n, m, A, B, pr, UH(2 * i), dp);                  /* f(p) at t_i */
        for (int r = 0; r < n; ++r)
            x[r] = (float)((double)x[r] + h * ((((am0 * (double)dp[r] + am1 * (double)db0[r])
                                                + am2 * (double)db1[r]) + am3 * (double)db2[r])));
        outputs_f32(n, m, p, C, D, zC, zD, x, UH(2 * i), y + (size_t)i * nOut);
    }
#undef UH
    free(buf);
    return total;
}
