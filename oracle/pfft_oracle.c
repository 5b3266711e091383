/*
 * pfft_oracle.c -- CPU restatement of rocket-fft's transform engine.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker.  The product (rocket-fft_amd/csrc) never links or calls it.
 *
 * What is restated, with the reference location it follows
 * (H = rocket_fft/_pocketfft_hdronly.h):
 *   - integer plan facts, bit-exact: largest_prime_factor (H:548-563),
 *     cost_guess (H:565-586), good_size_cmplx (H:588-623), good_size_real
 *     (H:625-650), cfftp factor order (H:1750-1778), rfftp factor order
 *     (H:2629-2652), FFTPACK-vs-Bluestein choice (H:2843-2861 complex,
 *     H:2885-2903 real) and the Bluestein length n2 (H:2765-2767);
 *   - the 1-D complex DFT: mixed-radix decimation in time over the reference's
 *     factor list (H:1690-1740), or Bluestein with the exact integer chirp
 *     index m^2 mod 2n and 1/n2-normalised kernel (H:2723-2793) where the
 *     reference picks Bluestein;
 *   - real transforms in FFTPACK halfcomplex order [r0, r1, i1, r2, i2, ...]
 *     (H:1836-2717, H:2801-2827);
 *   - DCT/DST I-IV from their defining sums evaluated through 2n-point DFTs,
 *     with the reference's non-textbook ortho rules (H:2918-3163).
 * All arithmetic is long double (64-bit mantissa on x86-64) so the oracle is
 * more accurate than either the reference or the GPU path.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef long double R;
typedef long double complex C;

static const R PI_L = 3.141592653589793238462643383279502884197L;

/* ---------------------------------------------------------------- integers */

uint64_t or_largest_prime_factor(uint64_t n) /* H:548-563 */
{
    uint64_t res = 1;
    while ((n & 1) == 0) { res = 2; n >>= 1; }
    for (uint64_t x = 3; x * x <= n; x += 2)
        while (n % x == 0) { res = x; n /= x; }
    if (n > 1) res = n;
    return res;
}

double or_cost_guess(uint64_t n) /* H:565-586: cost = n * sum(p<=5 ? p : 1.1p) */
{
    const double lfp = 1.1;
    uint64_t ni = n;
    double result = 0.;
    while ((n & 1) == 0) { result += 2; n >>= 1; }
    for (uint64_t x = 3; x * x <= n; x += 2)
        while (n % x == 0) {
            result += (x <= 5) ? (double)x : lfp * (double)x;
            n /= x;
        }
    if (n > 1) result += (n <= 5) ? (double)n : lfp * (double)n;
    return result * (double)ni;
}

/* Smallest 2^a 3^b 5^c 7^d 11^e >= n, n itself when n <= 12 (H:588-623).
 * Restated as an enumeration over the 7/11/5 powers with the 2/3 part
 * completed greedily; the result is the same minimum. */
uint64_t or_good_size_cmplx(uint64_t n)
{
    if (n <= 12) return n;
    uint64_t best = 2 * n;
    for (uint64_t f11 = 1; f11 < best; f11 *= 11)
        for (uint64_t f7 = f11; f7 < best; f7 *= 7)
            for (uint64_t f5 = f7; f5 < best; f5 *= 5) {
                uint64_t x = f5;
                while (x < n) x *= 2;
                for (;;) {
                    if (x < n) x *= 3;
                    else if (x > n) {
                        if (x < best) best = x;
                        if (x & 1) break;
                        x >>= 1;
                    } else return n;
                }
            }
    return best;
}

uint64_t or_good_size_real(uint64_t n) /* H:625-650: 2^a 3^b 5^c, n<=6 -> n */
{
    if (n <= 6) return n;
    uint64_t best = 2 * n;
    for (uint64_t f5 = 1; f5 < best; f5 *= 5) {
        uint64_t x = f5;
        while (x < n) x *= 2;
        for (;;) {
            if (x < n) x *= 3;
            else if (x > n) {
                if (x < best) best = x;
                if (x & 1) break;
                x >>= 1;
            } else return n;
        }
    }
    return best;
}

/* H:1750-1778: 8s, then 4s, then one 2 moved to the front, then odd
 * divisors ascending, then the remaining prime. */
int or_cfftp_factors(uint64_t n, uint64_t *out, int cap)
{
    uint64_t f[64];
    int k = 0;
    uint64_t len = n;
    if (n <= 1) return 0;
    while ((len & 7) == 0) { f[k++] = 8; len >>= 3; }
    while ((len & 3) == 0) { f[k++] = 4; len >>= 2; }
    if ((len & 1) == 0) {
        len >>= 1;
        f[k++] = 2;
        uint64_t t = f[0]; f[0] = f[k - 1]; f[k - 1] = t;
    }
    for (uint64_t d = 3; d * d <= len; d += 2)
        while (len % d == 0) { f[k++] = d; len /= d; }
    if (len > 1) f[k++] = len;
    for (int i = 0; i < k && i < cap; ++i) out[i] = f[i];
    return k;
}

/* H:2629-2652: 4s, one 2 moved to the front, odd divisors ascending. */
int or_rfftp_factors(uint64_t n, uint64_t *out, int cap)
{
    uint64_t f[64];
    int k = 0;
    uint64_t len = n;
    if (n <= 1) return 0;
    while ((len % 4) == 0) { f[k++] = 4; len >>= 2; }
    if ((len % 2) == 0) {
        len >>= 1;
        f[k++] = 2;
        uint64_t t = f[0]; f[0] = f[k - 1]; f[k - 1] = t;
    }
    for (uint64_t d = 3; d * d <= len; d += 2)
        while (len % d == 0) { f[k++] = d; len /= d; }
    if (len > 1) f[k++] = len;
    for (int i = 0; i < k && i < cap; ++i) out[i] = f[i];
    return k;
}

/* H:2843-2861 (complex) and H:2885-2903 (real, FFTPACK cost halved).
 * Returns 1 when Bluestein is chosen and writes n2 = good_size_cmplx(2n-1). */
static int plan_choice(uint64_t n, uint64_t *n2, double half)
{
    uint64_t lpf = (n < 50) ? 0 : or_largest_prime_factor(n);
    *n2 = 0;
    if (lpf * lpf <= n) return 0;
    double comp1 = half * or_cost_guess(n);
    double comp2 = 2 * or_cost_guess(or_good_size_cmplx(2 * n - 1));
    comp2 *= 1.5;
    if (comp2 < comp1) {
        *n2 = or_good_size_cmplx(2 * n - 1);
        return 1;
    }
    return 0;
}

int or_plan_c(uint64_t n, uint64_t *n2) { return plan_choice(n, n2, 1.0); }
int or_plan_r(uint64_t n, uint64_t *n2) { return plan_choice(n, n2, 0.5); }

/* ------------------------------------------------------------- twiddles */

/* e^{+2 pi i k / n}, octant-reduced so the argument of sinl/cosl is at most
 * pi/4 (the reduction of H:472-510, here in long double). */
static C expi_2pi(uint64_t k, uint64_t n)
{
    k %= n;
    /* x = 8k in units of n/... : work with 8k vs n multiples exactly. */
    unsigned __int128 x = (unsigned __int128)k * 8u;
    unsigned __int128 N = n;
    R c, s;
    int oct = (int)(x / N);            /* 0..7 */
    unsigned __int128 rem = x - (unsigned __int128)oct * N; /* in [0, n) */
    /* angle within octant = rem / n * pi/4; odd octants count from the top. */
    R a;
    if (oct & 1) a = (R)(N - rem) / (R)n * (PI_L / 4);
    else a = (R)rem / (R)n * (PI_L / 4);
    R ca = cosl(a), sa = sinl(a);
    switch (oct) {
    case 0: c = ca; s = sa; break;
    case 1: c = sa; s = ca; break;
    case 2: c = -sa; s = ca; break;
    case 3: c = -ca; s = sa; break;
    case 4: c = -ca; s = -sa; break;
    case 5: c = -sa; s = -ca; break;
    case 6: c = sa; s = -ca; break;
    default: c = ca; s = -sa; break;
    }
    return c + I * s;
}

void or_twiddle(uint64_t n, uint64_t count, double *out)
{
    for (uint64_t k = 0; k < count; ++k) {
        C w = expi_2pi(k, n);
        out[2 * k] = (double)creall(w);
        out[2 * k + 1] = (double)cimagl(w);
    }
}

/* ------------------------------------------------------- complex 1-D DFT */

/* A per-length plan: factor list, table W[e] = e^{+2 pi i e/n}, and for
 * Bluestein lengths the inner plan, chirp and transformed kernel. */
typedef struct oplan {
    uint64_t n;
    int nf;
    uint64_t fac[64];
    C *tw;
    int blue;
    uint64_t n2;
    struct oplan *sub;
    C *bk, *bkf, *scr, *ba, *bb;
} oplan;

static void plan_exec(const oplan *p, const C *x, C *y, int forward);

static oplan *plan_new(uint64_t n)
{
    oplan *p = (oplan *)calloc(1, sizeof(oplan));
    p->n = n;
    if (n > 1 && or_plan_c(n, &p->n2)) {
        /* H:2765-2793 */
        p->blue = 1;
        p->sub = plan_new(p->n2);
        p->bk = (C *)malloc(sizeof(C) * n);
        p->bkf = (C *)malloc(sizeof(C) * p->n2);
        p->ba = (C *)malloc(sizeof(C) * p->n2);
        p->bb = (C *)malloc(sizeof(C) * p->n2);
        uint64_t coeff = 0;
        p->bk[0] = 1;
        for (uint64_t m = 1; m < n; ++m) {
            coeff += 2 * m - 1;
            if (coeff >= 2 * n) coeff -= 2 * n;
            p->bk[m] = expi_2pi(coeff, 2 * n);
        }
        R xn2 = 1.0L / (R)p->n2;
        for (uint64_t m = 0; m < p->n2; ++m) p->ba[m] = 0;
        p->ba[0] = p->bk[0] * xn2;
        for (uint64_t m = 1; m < n; ++m) p->ba[m] = p->ba[p->n2 - m] = p->bk[m] * xn2;
        plan_exec(p->sub, p->ba, p->bkf, 1);
        return p;
    }
    p->nf = n > 1 ? or_cfftp_factors(n, p->fac, 64) : 0;
    p->tw = (C *)malloc(sizeof(C) * (n ? n : 1));
    for (uint64_t e = 0; e < n; ++e) p->tw[e] = expi_2pi(e, n);
    p->scr = (C *)malloc(sizeof(C) * (2 * n + 8));
    return p;
}

static void plan_free(oplan *p)
{
    if (!p) return;
    plan_free(p->sub);
    free(p->tw); free(p->bk); free(p->bkf); free(p->scr); free(p->ba); free(p->bb);
    free(p);
}

/* Mixed-radix decimation in time, radices taken from the front of the
 * reference's factor list (H:1750-1778, executed in that order as in
 * H:1701-1727): out[k + m q] = sum_r W_n^{r(k+mq)} Y_r[k], Y_r the m-point
 * DFT of x[r::p].  W_n^e is read from the top-level table with stride N/n. */
static void dft_mixed(const oplan *P, const C *x, size_t xs, C *y, uint64_t n,
                      const uint64_t *fac, int sign, C *scratch)
{
    if (n == 1) { y[0] = x[0]; return; }
    uint64_t p = fac[0], m = n / p, st = P->n / n;
    for (uint64_t r = 0; r < p; ++r)
        dft_mixed(P, x + r * xs, xs * p, y + r * m, m, fac + 1, sign, scratch + n);
    C *t = scratch;
    for (uint64_t k = 0; k < m; ++k)
        for (uint64_t q = 0; q < p; ++q) {
            uint64_t kk = k + m * q;
            C acc = y[k];
            for (uint64_t r = 1; r < p; ++r) {
                C w = P->tw[((r * kk) % n) * st];
                acc += (sign < 0 ? conjl(w) : w) * y[r * m + k];
            }
            t[kk] = acc;
        }
    memcpy(y, t, n * sizeof(C));
}

/* Bluestein execution (H:2732-2762): forward multiplies by conj(bk), by bkf
 * in the frequency domain, and by conj(bk) again; backward conjugates all. */
static void dft_bluestein(const oplan *p, const C *x, C *y, int forward)
{
    uint64_t n = p->n, n2 = p->n2;
    C *a = p->ba, *af = p->bb;
    for (uint64_t m = 0; m < n2; ++m) a[m] = 0;
    for (uint64_t m = 0; m < n; ++m) a[m] = x[m] * (forward ? conjl(p->bk[m]) : p->bk[m]);
    plan_exec(p->sub, a, af, 1);
    for (uint64_t m = 0; m < n2; ++m) af[m] *= forward ? p->bkf[m] : conjl(p->bkf[m]);
    plan_exec(p->sub, af, a, 0);
    for (uint64_t m = 0; m < n; ++m) y[m] = a[m] * (forward ? conjl(p->bk[m]) : p->bk[m]);
}

static void plan_exec(const oplan *p, const C *x, C *y, int forward)
{
    if (p->n == 1) { y[0] = x[0]; return; }
    if (p->blue) { dft_bluestein(p, x, y, forward); return; }
    dft_mixed(p, x, 1, y, p->n, p->fac, forward ? -1 : +1, p->scr);
}

/* batch of contiguous complex lines (interleaved doubles), in may equal out */
void or_c2c(uint64_t batch, uint64_t n, const double *in, double *out, int forward,
            double fct)
{
    oplan *p = plan_new(n);
    C *x = (C *)malloc(sizeof(C) * n), *y = (C *)malloc(sizeof(C) * n);
    for (uint64_t b = 0; b < batch; ++b) {
        const double *pi = in + 2 * n * b;
        double *po = out + 2 * n * b;
        for (uint64_t j = 0; j < n; ++j) x[j] = (R)pi[2 * j] + I * (R)pi[2 * j + 1];
        plan_exec(p, x, y, forward);
        for (uint64_t j = 0; j < n; ++j) {
            po[2 * j] = (double)(creall(y[j]) * (R)fct);
            po[2 * j + 1] = (double)(cimagl(y[j]) * (R)fct);
        }
    }
    free(x); free(y);
    plan_free(p);
}

/* ------------------------------------------------------------ real 1-D */

/* r2hc: forward real DFT packed as [r0, r1, i1, ..., r_{n/2} (n even)]
 * (rfftp forward, H:2584-2605; Bluestein exec_r, H:2801-2812). */
static void r2hc_line(const oplan *P, const R *x, R *y, uint64_t n, C *tx, C *ty)
{
    for (uint64_t j = 0; j < n; ++j) tx[j] = x[j];
    plan_exec(P, tx, ty, 1);
    y[0] = creall(ty[0]);
    uint64_t i = 1, k = 1;
    for (; i + 1 < n; i += 2, ++k) { y[i] = creall(ty[k]); y[i + 1] = cimagl(ty[k]); }
    if (i < n) y[i] = creall(ty[k]);
}

/* hc2r: backward real DFT of halfcomplex data (rfftp backward,
 * H:2606-2623; Bluestein exec_r, H:2813-2826). */
static void hc2r_line(const oplan *P, const R *x, R *y, uint64_t n, C *tx, C *ty)
{
    tx[0] = x[0];
    uint64_t i = 1, k = 1;
    for (; i + 1 < n; i += 2, ++k) {
        tx[k] = x[i] + I * x[i + 1];
        tx[n - k] = x[i] - I * x[i + 1];
    }
    if (i < n) tx[k] = x[i];
    plan_exec(P, tx, ty, 0);
    for (uint64_t j = 0; j < n; ++j) y[j] = creall(ty[j]);
}

void or_r2hc(uint64_t batch, uint64_t n, const double *in, double *out, double fct)
{
    oplan *P = plan_new(n);
    R *x = (R *)malloc(sizeof(R) * n), *y = (R *)malloc(sizeof(R) * n);
    C *tx = (C *)malloc(sizeof(C) * n), *ty = (C *)malloc(sizeof(C) * n);
    for (uint64_t b = 0; b < batch; ++b) {
        for (uint64_t j = 0; j < n; ++j) x[j] = in[n * b + j];
        r2hc_line(P, x, y, n, tx, ty);
        for (uint64_t j = 0; j < n; ++j) out[n * b + j] = (double)(y[j] * (R)fct);
    }
    free(x); free(y); free(tx); free(ty);
    plan_free(P);
}

void or_hc2r(uint64_t batch, uint64_t n, const double *in, double *out, double fct)
{
    oplan *P = plan_new(n);
    R *x = (R *)malloc(sizeof(R) * n), *y = (R *)malloc(sizeof(R) * n);
    C *tx = (C *)malloc(sizeof(C) * n), *ty = (C *)malloc(sizeof(C) * n);
    for (uint64_t b = 0; b < batch; ++b) {
        for (uint64_t j = 0; j < n; ++j) x[j] = in[n * b + j];
        hc2r_line(P, x, y, n, tx, ty);
        for (uint64_t j = 0; j < n; ++j) out[n * b + j] = (double)(y[j] * (R)fct);
    }
    free(x); free(y); free(tx); free(ty);
    plan_free(P);
}

/* ------------------------------------------------------------- DCT / DST */

/* Unnormalised transforms (the reference's fct = 1 output):
 *   DCT-I   y_k = x_0 + (-1)^k x_{n-1} + 2 sum_{j=1}^{n-2} x_j cos(pi j k/(n-1))
 *   DST-I   y_k = 2 sum_j x_j sin(pi (j+1)(k+1)/(n+1))
 *   DCT-II  y_k = 2 sum_j x_j cos(pi k (2j+1)/(2n))
 *   DCT-III y_k = x_0 + 2 sum_{j>=1} x_j cos(pi j (2k+1)/(2n))
 *   DST-II  y_k = 2 sum_j x_j sin(pi (k+1)(2j+1)/(2n))
 *   DST-III y_k = (-1)^k x_{n-1} + 2 sum_{j<n-1} x_j sin(pi (j+1)(2k+1)/(2n))
 *   DCT-IV  y_k = 2 sum_j x_j cos(pi (2k+1)(2j+1)/(4n))
 *   DST-IV  y_k = 2 sum_j x_j sin(pi (2k+1)(2j+1)/(4n))
 * ortho follows the reference, not the textbook (H:2931-2944 DCT-I scales
 * x_0, x_{n-1} by sqrt2 and y_0, y_{n-1} by 1/sqrt2; H:3030-3034 types II
 * scale y_0 by 1/sqrt2 -- for DST-II that is y_0, not y_{n-1}; H:3038-3039
 * types III scale x_0 by sqrt2 -- for DST-III x_0, not x_{n-1}; DST-I and
 * types IV ignore ortho, H:2968, H:3089). */
static void dcst_line(const oplan *P, const R *xin, R *y, uint64_t n, int type, int cosine,
                      int ortho, C *a, C *b)
{
    const R s2 = sqrtl(2.0L);
    R *x = (R *)malloc(sizeof(R) * (n + 1));
    for (uint64_t j = 0; j < n; ++j) x[j] = xin[j];
    if (type == 1) {
        if (cosine) {
            if (n == 1) { y[0] = x[0]; free(x); return; }  /* not reached: n>=2 */
            uint64_t N = 2 * (n - 1);
            if (ortho) { x[0] *= s2; x[n - 1] *= s2; }
            for (uint64_t j = 0; j < N; ++j) a[j] = 0;
            a[0] = x[0];
            for (uint64_t j = 1; j < n; ++j) a[j] = a[N - j] = x[j];
            plan_exec(P, a, b, 1);
            for (uint64_t k = 0; k < n; ++k) y[k] = creall(b[k]);
            if (ortho) { y[0] /= s2; y[n - 1] /= s2; }
        } else {
            uint64_t N = 2 * (n + 1);
            for (uint64_t j = 0; j < N; ++j) a[j] = 0;
            for (uint64_t j = 0; j < n; ++j) { a[j + 1] = x[j]; a[N - 1 - j] = -x[j]; }
            plan_exec(P, a, b, 1);
            for (uint64_t k = 0; k < n; ++k) y[k] = -cimagl(b[k + 1]);
        }
        free(x);
        return;
    }
    uint64_t M = 2 * n;
    for (uint64_t j = 0; j < M; ++j) a[j] = 0;
    if (type == 2) {
        /* sum_j x_j e^{-i pi k(2j+1)/(2n)} = e^{-i pi k/(2n)} DFT_2n(x)[k] */
        for (uint64_t j = 0; j < n; ++j) a[j] = x[j];
        plan_exec(P, a, b, 1);
        for (uint64_t k = 0; k < n; ++k) {
            uint64_t kk = cosine ? k : k + 1;
            C z = 2 * b[kk] * conjl(expi_2pi(kk, 4 * n));
            y[k] = cosine ? creall(z) : -cimagl(z);
        }
        if (ortho) y[0] /= s2;
    } else if (type == 3) {
        if (ortho) x[0] *= s2;
        if (cosine) {
            /* y_k = Re sum_j w_j x_j e^{i pi j(2k+1)/(2n)}, w_0 = 1, else 2 */
            for (uint64_t j = 0; j < n; ++j)
                a[j] = (j == 0 ? 1 : 2) * x[j] * expi_2pi(j, 4 * n);
            plan_exec(P, a, b, 0);
            for (uint64_t k = 0; k < n; ++k) y[k] = creall(b[k]);
        } else {
            /* y_k = Im sum_{j'=1}^{n} w x_{j'-1} e^{i pi j'(2k+1)/(2n)},
             * w = 2 except j' = n (weight 1) */
            for (uint64_t jp = 1; jp <= n; ++jp)
                a[jp % M] = (jp == n ? 1 : 2) * x[jp - 1] * expi_2pi(jp, 4 * n);
            plan_exec(P, a, b, 0);
            for (uint64_t k = 0; k < n; ++k) y[k] = cimagl(b[k]);
        }
    } else { /* type 4 */
        /* sum_j x_j e^{-i pi (2k+1)(2j+1)/(4n)}
         *   = e^{-i pi (2k+1)/(4n)} sum_j (x_j e^{-i pi j/(2n)}) e^{-2 pi i jk/(2n)} */
        for (uint64_t j = 0; j < n; ++j) a[j] = x[j] * conjl(expi_2pi(j, 4 * n));
        plan_exec(P, a, b, 1);
        for (uint64_t k = 0; k < n; ++k) {
            C z = 2 * b[k] * conjl(expi_2pi(2 * k + 1, 8 * n));
            y[k] = cosine ? creall(z) : -cimagl(z);
        }
    }
    free(x);
}

/* type 1..4; returns -1 for an invalid type (the reference throws,
 * H:3896-3897, H:3919-3920). */
int or_dcst(uint64_t batch, uint64_t n, const double *in, double *out, int type,
            int cosine, int ortho, double fct)
{
    if (type < 1 || type > 4) return -1;
    if (type == 1 && cosine && n < 2) return -1; /* zero-length rfft plan */
    uint64_t M = 2 * (n + 1) + 8;
    uint64_t L = type == 1 ? (cosine ? 2 * (n - 1) : 2 * (n + 1)) : 2 * n;
    oplan *P = plan_new(L);
    R *x = (R *)malloc(sizeof(R) * n), *y = (R *)malloc(sizeof(R) * n);
    C *a = (C *)malloc(sizeof(C) * M), *b = (C *)malloc(sizeof(C) * M);
    for (uint64_t bb = 0; bb < batch; ++bb) {
        for (uint64_t j = 0; j < n; ++j) x[j] = in[n * bb + j];
        dcst_line(P, x, y, n, type, cosine, ortho, a, b);
        for (uint64_t j = 0; j < n; ++j) out[n * bb + j] = (double)(y[j] * (R)fct);
    }
    free(x); free(y); free(a); free(b);
    plan_free(P);
    return 0;
}

/* Direct O(n^2) DFT in long double, an independent check on the above. */
void or_brute_dft(uint64_t batch, uint64_t n, const double *in, double *out, int forward)
{
    for (uint64_t b = 0; b < batch; ++b) {
        const double *pi = in + 2 * n * b;
        double *po = out + 2 * n * b;
        for (uint64_t k = 0; k < n; ++k) {
            C acc = 0;
            for (uint64_t j = 0; j < n; ++j) {
                C w = expi_2pi((j * k) % n, n);
                if (forward) w = conjl(w);
                acc += w * ((R)pi[2 * j] + I * (R)pi[2 * j + 1]);
            }
            po[2 * k] = (double)creall(acc);
            po[2 * k + 1] = (double)cimagl(acc);
        }
    }
}
