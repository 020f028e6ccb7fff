/*
 * tcfft_oracle.c -- FP64 CPU restatement of tcFFT's FFT path (checker only).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The HIP product never links or calls it.
 * See tcfft_oracle.h for what pins each function.
 */
#include "tcfft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* tcfft_doit_half.cpp:6-29.  The reference builds the permutation by peeling
 * radices from the last to the first: at each level a block of length `len`
 * is split into `r` interleaved sub-sequences (element l*r+k of the block goes
 * to slot k*(len/r)+l), then each sub-block is processed with the next radix.
 * The same construction, written out level by level. */
void oracle_gen_rev(int n, int *rev, const int *radices, int n_radices)
{
    int *cur = (int *)malloc(sizeof(int) * (size_t)n);
    int *nxt = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i)
        cur[i] = i;
    int len = n;
    for (int lvl = n_radices - 1; lvl >= 0; --lvl) {
        const int r = radices[lvl];
        const int sub = len / r;
        for (int base = 0; base < n; base += len)
            for (int k = 0; k < r; ++k)
                for (int l = 0; l < sub; ++l)
                    nxt[base + k * sub + l] = cur[base + l * r + k];
        int *t = cur;
        cur = nxt;
        nxt = t;
        len = sub;
    }
    memcpy(rev, cur, sizeof(int) * (size_t)n);
    free(cur);
    free(nxt);
}

/* doit.h:7-19: srand(seed), then re/im pairs of 0.0001f*rand()/RAND_MAX,
 * evaluated in float as the reference's expression is. */
void oracle_generate_data(double *data, int n, int n_batch, int seed)
{
    srand((unsigned)seed);
    for (int b = 0; b < n_batch; ++b)
        for (int j = 0; j < n; ++j) {
            float re = 0.0001f * (float)rand() / (float)RAND_MAX;
            float im = 0.0001f * (float)rand() / (float)RAND_MAX;
            data[2 * ((size_t)b * n + j) + 0] = re;
            data[2 * ((size_t)b * n + j) + 1] = im;
        }
}

/* accuracy.cpp:25-41.  std::min(1.0, v) returns 1.0 unless v < 1.0, so a NaN
 * (0/0) or an infinity (x/0) counts as 1.0; reproduced with the same test. */
double oracle_get_error(const double *tested, const double *standard, int n, int n_batch)
{
    double acc = 0.0;
    size_t count = (size_t)n * (size_t)n_batch * 2;
    for (size_t i = 0; i < count; ++i) {
        double v = fabs((tested[i] - standard[i]) / standard[i]);
        acc += (v < 1.0) ? v : 1.0;
    }
    return acc / n / n_batch;
}

/* ------------------------------------------------------------------------ */
/* tcfft_half.cu:1235-1330: the ten sizes tcfftCreate knows.  radices default
 * to 16 (tcfft_half.h:10) and mergings to 0 (tcfft_half.h:12); each case
 * overrides a few entries.  merging code m means a pass of radix 256<<m. */
struct ref_entry {
    int n, n_radices, n_mergings;
    int small_radix_at[3]; /* positions that hold a 2 or a 4 (or -1) */
    int small_radix[3];
    int mergings[3];
};

static const struct ref_entry REF_TABLE_1D[] = {
    {256, 2, 1, {-1, -1, -1}, {0, 0, 0}, {0, 0, 0}},
    {512, 3, 1, {2, -1, -1}, {2, 0, 0}, {1, 0, 0}},
    {1024, 3, 1, {2, -1, -1}, {4, 0, 0}, {2, 0, 0}},
    {131072, 5, 2, {2, -1, -1}, {2, 0, 0}, {1, 0, 0}},
    {262144, 6, 2, {2, 5, -1}, {2, 2, 0}, {1, 1, 0}},
    {524288, 6, 2, {2, 5, -1}, {4, 2, 0}, {2, 1, 0}},
    {16777216, 6, 3, {-1, -1, -1}, {0, 0, 0}, {0, 0, 0}},
    {33554432, 7, 3, {2, -1, -1}, {2, 0, 0}, {1, 0, 0}},
    {67108864, 8, 3, {2, 5, -1}, {2, 2, 0}, {1, 1, 0}},
    {134217728, 9, 3, {2, 5, 8}, {2, 2, 2}, {1, 1, 1}},
};

int oracle_ref_plan_1d(int n, int *radices9, int *n_radices, int *mergings3, int *n_mergings)
{
    for (int i = 0; i < 9; ++i)
        radices9[i] = 16;
    for (int i = 0; i < 3; ++i)
        mergings3[i] = 0;
    *n_radices = 0;
    *n_mergings = 0;
    for (size_t e = 0; e < sizeof(REF_TABLE_1D) / sizeof(REF_TABLE_1D[0]); ++e) {
        const struct ref_entry *t = &REF_TABLE_1D[e];
        if (t->n != n)
            continue;
        *n_radices = t->n_radices;
        *n_mergings = t->n_mergings;
        for (int k = 0; k < 3; ++k) {
            if (t->small_radix_at[k] >= 0)
                radices9[t->small_radix_at[k]] = t->small_radix[k];
            mergings3[k] = t->mergings[k];
        }
        return 0;
    }
    return -1;
}

/* tcfft_half_2d.cu:1151-1206 (radix defaults {16,16,2}, tcfft_half_2d.h:10-11).
 * Row pass uses mergings[0] (from ny), column pass mergings[1] (from nx);
 * there is no column kernel for nx = 1024 (tcfft_half_2d.cu:1164-1165). */
static int ref_axis_2d(int len, int *radices3, int *n_rad, int *merging)
{
    radices3[0] = 16;
    radices3[1] = 16;
    radices3[2] = 2;
    *merging = 0;
    switch (len) {
    case 256: *n_rad = 2; return 0;
    case 512: *n_rad = 3; *merging = 1; return 0;
    case 1024: *n_rad = 3; radices3[2] = 4; *merging = 2; return 0;
    default: *n_rad = 0; return -1;
    }
}

int oracle_ref_plan_2d(int nx, int ny, int *radices_x3, int *n_radices_x,
                       int *radices_y3, int *n_radices_y, int *mergings2)
{
    int ok_x = ref_axis_2d(nx, radices_x3, n_radices_x, &mergings2[1]);
    int ok_y = ref_axis_2d(ny, radices_y3, n_radices_y, &mergings2[0]);
    if (ok_x || ok_y)
        return -1;
    if (mergings2[1] == 2)
        return -1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* The DIT arithmetic.  Pass kernels in the reference fuse several stages
 * (layer_256_1 does two radix-16 stages with twiddles W_{16*step}^{col*j}
 * then W_{256*step}^{(col+k1*step)*j}, tcfft_half.cu:299-321,346-381); as a
 * sequence of single-radix stages that is: for stage i with radix r and
 * step s = r_0*...*r_{i-1}, every column c < s of every block of s*r
 * elements gets v_j = x[j*s+c] * W_{s*r}^{c*j} and then a length-r DFT
 * written back to rows k*s+c (the DFT-16-as-matrix of complex_mul,
 * tcfft_half.cu:5-20, and the radix-2/4 epilogues, :690-698, :1078-1094). */

/* exp(sign * 2*pi*i * m / M) with m reduced exactly. */
static void unit_root(long long m, long long M, int sign, double *re, double *im)
{
    long long q = m % M;
    if (q < 0)
        q += M;
    double a = 2.0 * M_PI * (double)q / (double)M;
    *re = cos(a);
    *im = sign * sin(a);
}

static void dit_stage(double *x, int n, int r, int s, int sign, double *scratch)
{
    const int sr = s * r;
    /* DFT-r matrix */
    double *fr = scratch;            /* r*r */
    double *fi = fr + r * r;         /* r*r */
    double *vr = fi + r * r;         /* r */
    double *vi = vr + r;             /* r */
    for (int j = 0; j < r; ++j)
        for (int k = 0; k < r; ++k)
            unit_root((long long)j * k, r, sign, &fr[j * r + k], &fi[j * r + k]);
    /* twiddle table W_{sr}^m for m < sr when it is small enough */
    double *tw = NULL;
    if (sr <= (1 << 22) && s > 1) {
        tw = (double *)malloc(sizeof(double) * 2 * (size_t)sr);
        for (int m = 0; m < sr; ++m)
            unit_root(m, sr, sign, &tw[2 * m], &tw[2 * m + 1]);
    }
    for (int base = 0; base < n; base += sr) {
        for (int c = 0; c < s; ++c) {
            for (int j = 0; j < r; ++j) {
                const double *e = x + 2 * (size_t)(base + j * s + c);
                double wr = 1.0, wi = 0.0;
                if (s > 1) {
                    long long m = ((long long)c * j) % sr;
                    if (tw) {
                        wr = tw[2 * m];
                        wi = tw[2 * m + 1];
                    } else {
                        unit_root(m, sr, sign, &wr, &wi);
                    }
                }
                vr[j] = e[0] * wr - e[1] * wi;
                vi[j] = e[0] * wi + e[1] * wr;
            }
            for (int k = 0; k < r; ++k) {
                double ar = 0.0, ai = 0.0;
                for (int j = 0; j < r; ++j) {
                    ar += vr[j] * fr[j * r + k] - vi[j] * fi[j * r + k];
                    ai += vr[j] * fi[j * r + k] + vi[j] * fr[j * r + k];
                }
                double *o = x + 2 * (size_t)(base + k * s + c);
                o[0] = ar;
                o[1] = ai;
            }
        }
    }
    free(tw);
}

void oracle_dit(double *data, int n, int n_batch, const int *radices, int n_radices, int sign)
{
    int rmax = 1;
    for (int i = 0; i < n_radices; ++i)
        if (radices[i] > rmax)
            rmax = radices[i];
    double *scratch = (double *)malloc(sizeof(double) * (size_t)(2 * rmax * rmax + 2 * rmax));
    for (int b = 0; b < n_batch; ++b) {
        double *x = data + 2 * (size_t)n * b;
        int s = 1;
        for (int i = 0; i < n_radices; ++i) {
            dit_stage(x, n, radices[i], s, sign, scratch);
            s *= radices[i];
        }
    }
    free(scratch);
}

void oracle_dit_2d(double *data, int nx, int ny, int n_batch,
                   const int *radices_x, int n_radices_x,
                   const int *radices_y, int n_radices_y, int sign)
{
    double *col = (double *)malloc(sizeof(double) * 2 * (size_t)nx);
    for (int b = 0; b < n_batch; ++b) {
        double *img = data + 2 * (size_t)nx * ny * b;
        /* row pass: nx rows of length ny, contiguous */
        oracle_dit(img, ny, nx, radices_y, n_radices_y, sign);
        /* column pass: ny columns of length nx, stride ny */
        for (int y = 0; y < ny; ++y) {
            for (int x = 0; x < nx; ++x) {
                col[2 * x] = img[2 * ((size_t)x * ny + y)];
                col[2 * x + 1] = img[2 * ((size_t)x * ny + y) + 1];
            }
            oracle_dit(col, nx, 1, radices_x, n_radices_x, sign);
            for (int x = 0; x < nx; ++x) {
                img[2 * ((size_t)x * ny + y)] = col[2 * x];
                img[2 * ((size_t)x * ny + y) + 1] = col[2 * x + 1];
            }
        }
    }
    free(col);
}

/* ------------------------------------------------------------------------ */
int oracle_default_radices(int n, int *radices, int capacity)
{
    int count = 0;
    int rem = n;
    while (rem > 1) {
        int r = rem >= 16 ? 16 : rem;
        if (rem % r != 0 || count >= capacity)
            return -1;
        radices[count++] = r;
        rem /= r;
    }
    return count;
}

static void permute_into(const double *in, double *out, const int *rev, int n, int n_batch)
{
    for (int b = 0; b < n_batch; ++b)
        for (int i = 0; i < n; ++i) {
            out[2 * ((size_t)b * n + i)] = in[2 * ((size_t)b * n + rev[i])];
            out[2 * ((size_t)b * n + i) + 1] = in[2 * ((size_t)b * n + rev[i]) + 1];
        }
}

void oracle_fft(const double *in, double *out, int n, int n_batch, int sign)
{
    int radices[32];
    int nr = oracle_default_radices(n, radices, 32);
    if (nr < 0)
        return;
    int *rev = (int *)malloc(sizeof(int) * (size_t)n);
    oracle_gen_rev(n, rev, radices, nr);
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)n * n_batch);
    permute_into(in, tmp, rev, n, n_batch);
    oracle_dit(tmp, n, n_batch, radices, nr, sign);
    memcpy(out, tmp, sizeof(double) * 2 * (size_t)n * n_batch);
    free(tmp);
    free(rev);
}

void oracle_rfft(const double *in, double *out, int n, int n_batch)
{
    double *full = (double *)malloc(sizeof(double) * 2 * (size_t)n * n_batch);
    for (size_t i = 0; i < (size_t)n * n_batch; ++i) {
        full[2 * i] = in[i];
        full[2 * i + 1] = 0.0;
    }
    oracle_fft(full, full, n, n_batch, -1);
    const int bins = n / 2 + 1;
    for (int b = 0; b < n_batch; ++b)
        memcpy(out + 2 * (size_t)b * bins, full + 2 * (size_t)b * n, sizeof(double) * 2 * (size_t)bins);
    free(full);
}

void oracle_irfft(const double *in, double *out, int n, int n_batch)
{
    const int h = n / 2, bins = h + 1;
    double *full = (double *)malloc(sizeof(double) * 2 * (size_t)n * n_batch);
    for (int b = 0; b < n_batch; ++b) {
        const double *x = in + 2 * (size_t)b * bins;
        double *f = full + 2 * (size_t)b * n;
        for (int k = 0; k <= h; ++k) {
            f[2 * k] = x[2 * k];
            f[2 * k + 1] = (k == 0 || k == h) ? 0.0 : x[2 * k + 1];
        }
        for (int k = 1; k < h; ++k) {
            f[2 * (n - k)] = x[2 * k];
            f[2 * (n - k) + 1] = -x[2 * k + 1];
        }
    }
    oracle_fft(full, full, n, n_batch, +1);
    for (size_t i = 0; i < (size_t)n * n_batch; ++i)
        out[i] = full[2 * i];
    free(full);
}

void oracle_fft_2d(const double *in, double *out, int nx, int ny, int n_batch, int sign)
{
    int rx[32], ry[32];
    int nrx = oracle_default_radices(nx, rx, 32);
    int nry = oracle_default_radices(ny, ry, 32);
    if (nrx < 0 || nry < 0)
        return;
    int *revx = (int *)malloc(sizeof(int) * (size_t)nx);
    int *revy = (int *)malloc(sizeof(int) * (size_t)ny);
    oracle_gen_rev(nx, revx, rx, nrx);
    oracle_gen_rev(ny, revy, ry, nry);
    size_t img = (size_t)nx * ny;
    double *tmp = (double *)malloc(sizeof(double) * 2 * img * n_batch);
    for (int b = 0; b < n_batch; ++b)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                size_t dst = b * img + (size_t)x * ny + y;
                size_t src = b * img + (size_t)revx[x] * ny + revy[y];
                tmp[2 * dst] = in[2 * src];
                tmp[2 * dst + 1] = in[2 * src + 1];
            }
    oracle_dit_2d(tmp, nx, ny, n_batch, rx, nrx, ry, nry, sign);
    memcpy(out, tmp, sizeof(double) * 2 * img * n_batch);
    free(tmp);
    free(revx);
    free(revy);
}

/* ------------------------------------------------------------------------ */
void oracle_dft_ld(const double *in, double *out, int n, int sign)
{
    long double *acc = (long double *)malloc(sizeof(long double) * 2 * (size_t)n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < n; ++k) {
        long double ar = 0.0L, ai = 0.0L;
        for (int j = 0; j < n; ++j) {
            long long m = ((long long)j * k) % n;
            long double a = two_pi * (long double)m / (long double)n;
            long double c = cosl(a), s = sign * sinl(a);
            ar += (long double)in[2 * j] * c - (long double)in[2 * j + 1] * s;
            ai += (long double)in[2 * j] * s + (long double)in[2 * j + 1] * c;
        }
        acc[2 * k] = ar;
        acc[2 * k + 1] = ai;
    }
    for (int i = 0; i < 2 * n; ++i)
        out[i] = (double)acc[i];
    free(acc);
}

double oracle_rel_l2(const double *a, const double *ref, size_t count)
{
    long double num = 0.0L, den = 0.0L;
    for (size_t i = 0; i < count; ++i) {
        long double d = (long double)a[i] - (long double)ref[i];
        num += d * d;
        den += (long double)ref[i] * ref[i];
    }
    if (den == 0.0L)
        return num == 0.0L ? 0.0 : INFINITY;
    return (double)sqrtl(num / den);
}
