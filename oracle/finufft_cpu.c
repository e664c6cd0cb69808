/*
 * CPU restatement of the FINUFFT method (spread / interpolate / deconvolve /
 * amplify with the exponential-of-semicircle kernel), FP64, OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a stage-level checker and by
 * bench.py's cpu_baseline leg as the "FINUFFT-method CPU restatement".  It is
 * never linked into or called by the product library.
 *
 * The reference delegates this arithmetic to the FINUFFT library
 * (vendor/finufft, an empty git submodule in the reference tree,
 * .gitmodules:1-3; version evidence 2.4.1 in README.md:501; call sites
 * lib/jax_finufft_cpu.h:27-98 and lib/cufinufft_wrapper.cc:48-99).  The method
 * is restated here from its publications:
 *   - Barnett, Magland, af Klinteberg, "A parallel non-uniform fast Fourier
 *     transform library based on an exponential of semicircle kernel",
 *     SIAM J. Sci. Comput. 41(5), 2019: kernel phi(z)=exp(beta(sqrt(1-z^2)-1)),
 *     width w ~ log10(1/eps)+1, beta ~ 2.30 w at upsampling 2, fine grid
 *     nf = next 2,3,5-smooth even size >= max(sigma N, 2w), deconvolution by
 *     the kernel's Fourier transform evaluated by Gauss-Legendre quadrature.
 *   - Barnett, "Aliasing error of the exp(beta sqrt(1-z^2)) kernel in the
 *     nonuniform fast Fourier transform", ACHA 2021: beta for sigma != 2.
 * The kernel is evaluated directly (exp/sqrt per tap), independent of the
 * product's plan-time Horner fit.
 *
 * Conventions (shared with the product so stage outputs can be compared):
 *   grid spacing h = 2 pi / nf; a point x is folded to u = (x/h) mod nf in
 *   [0, nf); its ns taps are l0 .. l0+ns-1 with l0 = ceil(u - ns/2); tap l has
 *   weight psi(u - l) = phi(2 (u - l) / ns) and lands at l mod nf.  Fine grid
 *   arrays are interleaved complex, x fastest: fw[l1 + nf1*(l2 + nf2*l3)].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FCPU_MAX_NS 16

/* ---------------------------------------------------------------- params */

/* Kernel width and shape parameter for tolerance eps at upsampling sigma.
 * Returns 0, or 1 when eps was below what the width cap can reach (the
 * "eps too small" warning, which is not an error). */
int fcpu_setup(double eps, double sigma, int *ns_out, double *beta_out) {
    int ret = 0;
    int ns;
    if (sigma == 2.0)
        ns = (int)ceil(-log10(eps / 10.0));
    else
        ns = (int)ceil(-log(eps) / (M_PI * sqrt(1.0 - 1.0 / sigma)));
    if (ns < 2) ns = 2;
    if (ns > FCPU_MAX_NS) {
        ns = FCPU_MAX_NS;
        ret = 1;
    }
    double bovns = 2.30;
    if (ns == 2) bovns = 2.20;
    if (ns == 3) bovns = 2.26;
    if (ns == 4) bovns = 2.38;
    if (sigma != 2.0) bovns = 0.97 * M_PI * (1.0 - 1.0 / (2.0 * sigma));
    *ns_out = ns;
    *beta_out = bovns * ns;
    return ret;
}

/* Smallest even n' >= n whose only prime factors are 2, 3, 5. */
int64_t fcpu_next235even(int64_t n) {
    if (n <= 2) return 2;
    if (n % 2 == 1) n += 1;
    for (int64_t m = n;; m += 2) {
        int64_t r = m;
        while (r % 2 == 0) r /= 2;
        while (r % 3 == 0) r /= 3;
        while (r % 5 == 0) r /= 5;
        if (r == 1) return m;
    }
}

int64_t fcpu_fine_size(int64_t n, double sigma, int ns) {
    int64_t nf = (int64_t)(sigma * (double)n);
    if (nf < 2 * ns) nf = 2 * ns;
    return fcpu_next235even(nf);
}

static inline double es_kernel(double z, double beta) {
    double a = 1.0 - z * z;
    if (a < 0.0) return 0.0;
    return exp(beta * (sqrt(a) - 1.0));
}

/* Gauss-Legendre nodes/weights on [-1, 1] by Newton iteration on P_n. */
void fcpu_gauss_legendre(int n, double *x, double *w) {
    for (int i = 0; i < n; i++) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0; }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        /* recompute derivative at the converged node */
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; k++) {
            double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

/* psi_hat(k h) = ns * int_0^1 phi(z) cos(pi k ns z / nf) dz for the n
 * frequencies k = kmin .. kmin+n-1 (out[q] for k = kmin + q).
 * The substitution z = sin(theta) turns the sqrt(1-z^2) endpoint into the
 * analytic integrand exp(beta(cos t - 1)) cos(a sin t) cos t on [0, pi/2], so
 * Gauss-Legendre converges exponentially. */
void fcpu_phihat(int ns, double beta, int64_t nf, int64_t kmin, int64_t n, double *out) {
    int q = 2 * ns + 40;
    double *gx = (double *)malloc(sizeof(double) * 2 * q);
    double *gw = gx + q;
    fcpu_gauss_legendre(q, gx, gw);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double a = M_PI * (double)(kmin + i) * ns / (double)nf;
        double acc = 0.0;
        for (int p = 0; p < q; p++) {
            double t = 0.25 * M_PI * (gx[p] + 1.0); /* [-1,1] -> [0, pi/2] */
            acc += 0.25 * M_PI * gw[p] * exp(beta * (cos(t) - 1.0)) * cos(a * sin(t)) * cos(t);
        }
        out[i] = ns * acc;
    }
    free(gx);
}

/* ---------------------------------------------------------------- points */

static inline double fold_u(double x, int64_t nf) {
    double u = x * ((double)nf / (2.0 * M_PI));
    u -= (double)nf * floor(u / (double)nf);
    if (u >= (double)nf) u -= (double)nf;
    if (u < 0.0) u += (double)nf;
    return u;
}

typedef struct {
    int64_t l0[3];
    double w[3][FCPU_MAX_NS];
} tapset_t;

static inline void point_taps(int dim, const double *xs[3], int64_t j, const int64_t nf[3], int ns,
                              double beta, tapset_t *t) {
    for (int d = 0; d < dim; d++) {
        double u = fold_u(xs[d][j], nf[d]);
        int64_t l0 = (int64_t)ceil(u - 0.5 * ns);
        t->l0[d] = l0;
        for (int i = 0; i < ns; i++) {
            double r = u - (double)(l0 + i); /* in (-ns/2, ns/2] */
            t->w[d][i] = es_kernel(2.0 * r / ns, beta);
        }
    }
    for (int d = dim; d < 3; d++) {
        t->l0[d] = 0;
        t->w[d][0] = 1.0;
    }
}

static inline int64_t wrapi(int64_t l, int64_t n) {
    l %= n;
    return l < 0 ? l + n : l;
}

/* ---------------------------------------------------------------- spread */

/* Type-1 spreading: fw = sum_j c_j psi(u_j - l), periodic.  Points are
 * bucketed by the block of their first tap along the slowest dimension; each
 * thread owns whole blocks and accumulates into a private slab, so the global
 * grid is only written under an atomic for slab overlaps. */
void fcpu_spread(int dim, int64_t M, const double *x, const double *y, const double *z,
                 const double *c, int ns, double beta, int64_t nf1, int64_t nf2, int64_t nf3,
                 double *fw) {
    const double *xs[3] = {x, y, z};
    int64_t nf[3] = {nf1, dim > 1 ? nf2 : 1, dim > 2 ? nf3 : 1};
    int64_t F = nf[0] * nf[1] * nf[2];
    memset(fw, 0, sizeof(double) * 2 * F);
    int sd = dim - 1; /* slowest dimension */
    int64_t bw = 16;  /* block width along the slowest dimension */
    if (dim == 1) bw = 256;
    int64_t nb = (nf[sd] + bw - 1) / bw;
    int64_t *cnt = (int64_t *)calloc(nb + 1, sizeof(int64_t));
    int64_t *bin = (int64_t *)malloc(sizeof(int64_t) * M);
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * M);
    for (int64_t j = 0; j < M; j++) {
        double u = fold_u(xs[sd][j], nf[sd]);
        int64_t l0 = wrapi((int64_t)ceil(u - 0.5 * ns), nf[sd]);
        bin[j] = l0 / bw;
        cnt[bin[j] + 1]++;
    }
    for (int64_t b = 0; b < nb; b++) cnt[b + 1] += cnt[b];
    {
        int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * nb);
        memcpy(pos, cnt, sizeof(int64_t) * nb);
        for (int64_t j = 0; j < M; j++) perm[pos[bin[j]]++] = j;
        free(pos);
    }
    int64_t slab_len = bw + ns; /* cells along sd touched by one block */
    int64_t inner = F / nf[sd]; /* cells per slowest-dim index */
#pragma omp parallel
    {
        double *slab = (double *)malloc(sizeof(double) * 2 * slab_len * inner);
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < nb; b++) {
            if (cnt[b + 1] == cnt[b]) continue;
            memset(slab, 0, sizeof(double) * 2 * slab_len * inner);
            int64_t s0 = b * bw; /* slab origin along sd */
            tapset_t t;
            for (int64_t p = cnt[b]; p < cnt[b + 1]; p++) {
                int64_t j = perm[p];
                point_taps(dim, xs, j, nf, ns, beta, &t);
                double cr = c[2 * j], ci = c[2 * j + 1];
                int n1 = ns, n2 = dim > 1 ? ns : 1, n3 = dim > 2 ? ns : 1;
                for (int i3 = 0; i3 < n3; i3++)
                    for (int i2 = 0; i2 < n2; i2++) {
                        double w23 = t.w[2][dim > 2 ? i3 : 0] * t.w[1][dim > 1 ? i2 : 0];
                        int64_t base;
                        if (dim == 1) base = 0;
                        else if (dim == 2) base = (wrapi(t.l0[1], nf[1]) - s0 + i2) * nf[0];
                        else
                            base = ((wrapi(t.l0[2], nf[2]) - s0 + i3) * nf[1] +
                                    wrapi(t.l0[1] + i2, nf[1])) * nf[0];
                        for (int i1 = 0; i1 < n1; i1++) {
                            double wv = w23 * t.w[0][i1];
                            int64_t l1;
                            if (dim == 1) l1 = wrapi(t.l0[0], nf[0]) - s0 + i1;
                            else l1 = wrapi(t.l0[0] + i1, nf[0]);
                            int64_t idx = base + l1;
                            slab[2 * idx] += wv * cr;
                            slab[2 * idx + 1] += wv * ci;
                        }
                    }
            }
            /* add the slab into the periodic grid */
            for (int64_t s = 0; s < slab_len; s++) {
                int64_t g = wrapi(s0 + s, nf[sd]);
                double *src = slab + 2 * s * inner;
                double *dst = fw + 2 * g * inner;
                for (int64_t q = 0; q < 2 * inner; q++) {
                    if (src[q] != 0.0) {
#pragma omp atomic
                        dst[q] += src[q];
                    }
                }
            }
        }
        free(slab);
    }
    free(cnt);
    free(bin);
    free(perm);
}

/* ---------------------------------------------------------------- interp */

/* Type-2 interpolation: c_j = sum_l fw[l] psi(u_j - l). */
void fcpu_interp(int dim, int64_t M, const double *x, const double *y, const double *z,
                 const double *fw, int ns, double beta, int64_t nf1, int64_t nf2, int64_t nf3,
                 double *c) {
    const double *xs[3] = {x, y, z};
    int64_t nf[3] = {nf1, dim > 1 ? nf2 : 1, dim > 2 ? nf3 : 1};
#pragma omp parallel for schedule(static, 1024)
    for (int64_t j = 0; j < M; j++) {
        tapset_t t;
        point_taps(dim, xs, j, nf, ns, beta, &t);
        int n1 = ns, n2 = dim > 1 ? ns : 1, n3 = dim > 2 ? ns : 1;
        double ar = 0, ai = 0;
        for (int i3 = 0; i3 < n3; i3++) {
            int64_t l3 = dim > 2 ? wrapi(t.l0[2] + i3, nf[2]) : 0;
            for (int i2 = 0; i2 < n2; i2++) {
                int64_t l2 = dim > 1 ? wrapi(t.l0[1] + i2, nf[1]) : 0;
                double w23 = t.w[2][dim > 2 ? i3 : 0] * t.w[1][dim > 1 ? i2 : 0];
                const double *row = fw + 2 * ((l3 * nf[1] + l2) * nf[0]);
                for (int i1 = 0; i1 < n1; i1++) {
                    int64_t l1 = wrapi(t.l0[0] + i1, nf[0]);
                    double wv = w23 * t.w[0][i1];
                    ar += wv * row[2 * l1];
                    ai += wv * row[2 * l1 + 1];
                }
            }
        }
        c[2 * j] = ar;
        c[2 * j + 1] = ai;
    }
}

/* ------------------------------------------------------- deconv / amplify */

static inline int64_t freq_of(int64_t i, int64_t n, int modeord) {
    if (modeord == 0) return i - n / 2;
    return (i < (n + 1) / 2) ? i : i - n;
}

/* Type 1: out[k] = B[k mod nf] / prod_d psi_hat_d(k_d).  ph_d holds
 * psi_hat for k = -(n_d/2) .. (n_d-1)/2 (CMCL order). */
void fcpu_deconv(int dim, int64_t n1, int64_t n2, int64_t n3, int64_t nf1, int64_t nf2,
                 int64_t nf3, const double *ph1, const double *ph2, const double *ph3,
                 int modeord, const double *B, double *out) {
    if (dim < 2) { n2 = 1; nf2 = 1; }
    if (dim < 3) { n3 = 1; nf3 = 1; }
    int64_t N = n1 * n2 * n3;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < N; k++) {
        int64_t i1 = k % n1, i2 = (k / n1) % n2, i3 = k / (n1 * n2);
        int64_t k1 = freq_of(i1, n1, modeord), k2 = freq_of(i2, n2, modeord),
                k3 = freq_of(i3, n3, modeord);
        double p = ph1[k1 + n1 / 2];
        if (dim > 1) p *= ph2[k2 + n2 / 2];
        if (dim > 2) p *= ph3[k3 + n3 / 2];
        int64_t l = wrapi(k1, nf1) + nf1 * (wrapi(k2, nf2) + nf2 * wrapi(k3, nf3));
        out[2 * k] = B[2 * l] / p;
        out[2 * k + 1] = B[2 * l + 1] / p;
    }
}

/* Type 2: fine-grid spectrum bhat[k mod nf] = f[k] / prod psi_hat, zero elsewhere. */
void fcpu_amplify(int dim, int64_t n1, int64_t n2, int64_t n3, int64_t nf1, int64_t nf2,
                  int64_t nf3, const double *ph1, const double *ph2, const double *ph3,
                  int modeord, const double *f, double *bhat) {
    if (dim < 2) { n2 = 1; nf2 = 1; }
    if (dim < 3) { n3 = 1; nf3 = 1; }
    int64_t N = n1 * n2 * n3;
    memset(bhat, 0, sizeof(double) * 2 * nf1 * nf2 * nf3);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < N; k++) {
        int64_t i1 = k % n1, i2 = (k / n1) % n2, i3 = k / (n1 * n2);
        int64_t k1 = freq_of(i1, n1, modeord), k2 = freq_of(i2, n2, modeord),
                k3 = freq_of(i3, n3, modeord);
        double p = ph1[k1 + n1 / 2];
        if (dim > 1) p *= ph2[k2 + n2 / 2];
        if (dim > 2) p *= ph3[k3 + n3 / 2];
        int64_t l = wrapi(k1, nf1) + nf1 * (wrapi(k2, nf2) + nf2 * wrapi(k3, nf3));
        bhat[2 * l] = f[2 * k] / p;
        bhat[2 * l + 1] = f[2 * k + 1] / p;
    }
}
