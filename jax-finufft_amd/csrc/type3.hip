// Type-3 driver: sizes the spread grid from the extents of sources and
// targets, spreads pre-phased strengths at the rescaled sources, evaluates the
// grid at the rescaled targets with an inner type-2 plan, and post-phases.
// Replaces the type-3 path of cuFINUFFT's setpts/execute
// (lib/kernels.cc.cu:65-75).
//
// Grid rule (published FINUFFT type-3 scheme), per dimension with source
// half-width X and target half-width S:
//   X, S made safe so that X*S >= 1 (X = 0 and/or S = 0 handled);
//   nf  = next235even(max(2 sigma X S / pi + ns + 1, 2 ns));
//   h   = 2 pi / nf;   gamma = nf / (2 sigma S).
// Setpts has to read the extents back to size the grid, so type-3 setpts
// synchronises the stream once (types 1 and 2 never do).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "jfa_internal.hpp"
#include "plan_state.hpp"

using namespace jfa;

namespace {

void free_ptr(void *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void type3_grid(double X, double S, double sigma, int ns, int64_t *nf, double *h,
                double *gam) {
    double Xs = X, Ss = S;
    if (X == 0.0) {
        if (S == 0.0) {
            Xs = 1.0;
            Ss = 1.0;
        } else {
            Xs = std::max(Xs, 1.0 / S);
        }
    } else {
        Ss = std::max(Ss, 1.0 / X);
    }
    double nfd = 2.0 * sigma * Ss * Xs / M_PI + (ns + 1);
    if (!std::isfinite(nfd)) nfd = 0.0;
    int64_t n = (int64_t)nfd;
    if (n < 2 * ns) n = 2 * ns;
    n = next235even(n);
    *nf = n;
    *h = 2.0 * M_PI / (double)n;
    *gam = (double)n / (2.0 * sigma * Ss);
}

}  // namespace

int jfa::type3_init(jfa_plan p) {
    p->geom.dim = p->dim;
    p->geom.ns = p->kp.ns;
    // Horner table (the spread kernel) and the quadrature table for psi_hat
    int rc = upload_kernel_tables(p, nullptr);
    if (rc) return rc;
    const int ns = p->kp.ns;
    const int nq = 2 * ns + 40;
    std::vector<double> gx(nq), gw(nq), quad(2 * nq);
    gauss_legendre(nq, gx.data(), gw.data());
    for (int q = 0; q < nq; ++q) {
        double th = 0.25 * M_PI * (gx[q] + 1.0);
        quad[q] = 0.25 * M_PI * gw[q] * std::cos(th) * std::exp(p->kp.beta * (std::cos(th) - 1.0));
        quad[nq + q] = std::sin(th);
    }
    JFA_HIP_CHECK(hipMalloc(&p->d_quad, sizeof(double) * 2 * nq));
    JFA_HIP_CHECK(hipMemcpy(p->d_quad, quad.data(), sizeof(double) * 2 * nq, hipMemcpyHostToDevice));
    JFA_HIP_CHECK(hipMalloc(&p->d_minmax, sizeof(double) * (2 + 2 * 1024)));
    return JFA_OK;
}

void jfa::type3_free(jfa_plan p) {
    free_ptr(p->d_cp);
    free_ptr(p->d_t3_deconv);
    free_ptr(reinterpret_cast<void *&>(p->d_minmax));
    free_ptr(p->d_quad);
}

template <typename T>
static int extents(jfa_plan p, int64_t n, const T *a, double *mn, double *mx) {
    if (n == 0) {
        *mn = *mx = 0.0;
        return JFA_OK;
    }
    int rc = launch_minmax<T>(n, a, p->d_minmax, p->stream);
    if (rc) return rc;
    double h[2];
    JFA_HIP_CHECK(hipMemcpyAsync(h, p->d_minmax, sizeof(h), hipMemcpyDeviceToHost, p->stream));
    JFA_HIP_CHECK(hipStreamSynchronize(p->stream));
    *mn = h[0];
    *mx = h[1];
    return JFA_OK;
}

template <typename T>
static int type3_setpts_t(jfa_plan p, int64_t M, const T *const xs[3], int64_t N,
                          const T *const ss[3]) {
    const int dim = p->dim;
    const int ns = p->kp.ns;
    const double sigma = p->opts.upsampfac;
    GridGeom &g = p->geom;
    double hg[3] = {1, 1, 1};
    for (int d = 0; d < 3; ++d) {
        g.nf[d] = 1;
        g.scale[d] = 1.0;
        g.shift[d] = 0.0;
        p->t3_C[d] = p->t3_D[d] = 0.0;
    }
    for (int d = 0; d < dim; ++d) {
        double xmn, xmx, smn, smx;
        int rc = extents<T>(p, M, xs[d], &xmn, &xmx);
        if (!rc) rc = extents<T>(p, N, ss[d], &smn, &smx);
        if (rc) return rc;
        const double C = 0.5 * (xmx + xmn), X = 0.5 * (xmx - xmn);
        const double D = 0.5 * (smx + smn), S = 0.5 * (smx - smn);
        int64_t nf;
        double h, gam;
        type3_grid(X, S, sigma, ns, &nf, &h, &gam);
        if (nf > (int64_t)1 << 30) {
            set_error("type-3 grid of %lld cells along dim %d is too large (space-bandwidth "
                      "product X*S = %g)", (long long)nf, d, X * S);
            return JFA_ERR_TOO_LARGE;
        }
        g.nf[d] = nf;
        g.scale[d] = (double)nf / (2.0 * M_PI) / gam;  // u = (x - C)/gam / h
        g.shift[d] = C;
        p->t3_C[d] = C;
        p->t3_D[d] = D;
        p->t3_gam[d] = gam;
        p->t3_h[d] = h;
        hg[d] = h * gam;
    }
    p->F = g.nf[0] * g.nf[1] * g.nf[2];
    p->Ntot = N;
    {
        // the inner type-2 plan upsamples the spread grid once more; check both
        // before allocating anything
        int64_t inner_cells = 1;
        for (int d = 0; d < dim; ++d) inner_cells *= fine_size(g.nf[d], sigma, ns);
        if (p->F > kMaxFineCells || inner_cells > kMaxFineCells) {
            set_error("type-3 grids of %lld / %lld cells exceed the supported 2^32 "
                      "(space-bandwidth product too large for one transform)",
                      (long long)p->F, (long long)inner_cells);
            return JFA_ERR_TOO_LARGE;
        }
    }
    // bins and method for the new grid
    {
        const int b2 = 32, b1 = 1024;
        int bs[3] = {1, 1, 1};
        if (dim == 1) bs[0] = b1;
        if (dim == 2) bs[0] = bs[1] = b2;
        if (dim == 3) bs[0] = bs[1] = bs[2] = 8;
        g.nbins = 1;
        for (int d = 0; d < 3; ++d) {
            int b = d < dim ? bs[d] : 1;
            if (d < dim) b = (int)std::min<int64_t>(b, std::max<int64_t>(1, g.nf[d] / 2));
            g.bs[d] = b;
            g.nb[d] = (int)((g.nf[d] + b - 1) / b);
            g.nbins *= g.nb[d];
        }
        size_t cells = 1;
        for (int d = 0; d < dim; ++d) cells *= (size_t)(g.bs[d] + ns - 1);
        const bool fits = cells * 2 * sizeof(T) <= 64 * 1024;
        g.ymajor = 0;
        p->method = p->opts.gpu_method == JFA_METHOD_GM_SORT || !fits ? JFA_METHOD_GM_SORT
                                                                      : JFA_METHOD_SM;
        p->method_by_density = p->opts.gpu_method == JFA_METHOD_AUTO && fits;
        const int om = p->opts.gpu_method;
        if ((om == JFA_METHOD_AUTO || om == JFA_METHOD_STRIP) && strip_applicable(p)) {
            int rc = setup_strip(p);
            if (rc) return rc;
            p->method = JFA_METHOD_STRIP;
            p->method_by_density = false;  // the density choice is between the other two
        }
    }
    // fine grid and prephase buffers
    if (p->F * p->batch > p->t3_nf_cap) {
        free_ptr(p->d_fw);
        const size_t bytes = (size_t)p->F * p->batch * 2 * sizeof(T);
        if (hipMalloc(&p->d_fw, bytes) != hipSuccess) {
            (void)hipGetLastError();
            set_error("could not allocate the %zu-byte type-3 grid", bytes);
            return JFA_ERR_ALLOC;
        }
        p->t3_nf_cap = p->F * p->batch;
    }
    if (M * p->batch > p->cap_cp) {
        free_ptr(p->d_cp);
        JFA_HIP_CHECK(hipMalloc(&p->d_cp, (size_t)std::max<int64_t>(M * p->batch, 1) * 2 * sizeof(T)));
        p->cap_cp = M * p->batch;
    }
    if (N > p->cap_t3_deconv) {
        free_ptr(p->d_t3_deconv);
        JFA_HIP_CHECK(hipMalloc(&p->d_t3_deconv, (size_t)std::max<int64_t>(N, 1) * sizeof(T)));
        p->cap_t3_deconv = N;
    }
    // sort the sources on the spread grid
    int rc = setpts_12(p, M, xs[0], xs[1], xs[2]);
    if (rc) return rc;
    // inner type-2 plan from the spread grid (FFT-ordered modes) to the targets
    const int64_t nm[3] = {g.nf[0], g.nf[1], g.nf[2]};
    if (p->inner) {
        bool same = true;
        for (int d = 0; d < dim; ++d) same = same && p->inner->nmodes[d] == nm[d];
        if (!same) {
            jfa_destroy(p->inner);
            p->inner = nullptr;
        }
    }
    if (!p->inner) {
        jfa_opts io = p->opts;
        io.modeord = 1;
        io.gpu_spreadinterponly = 0;
        io.kernel_width = ns;
        io.gpu_stream = p->stream;
        io.gpu_device_id = p->device;
        io.gpu_binsizex = io.gpu_binsizey = io.gpu_binsizez = 0;
        rc = jfa_makeplan(2, dim, nm, p->iflag, p->batch, p->eps, p->is_double, &io, &p->inner);
        if (rc > 1) return rc;
    }
    p->inner->stream = p->stream;
    p->inner->used = true;
    for (int d = 0; d < dim; ++d) {
        p->inner->geom.shift[d] = p->t3_D[d];
        p->inner->geom.scale[d] = (double)p->inner->geom.nf[d] / (2.0 * M_PI) * hg[d];
    }
    rc = setpts_12(p->inner, N, ss[0], ss[1], ss[2]);
    if (rc) return rc;
    rc = launch_t3_deconv_factors<T>(dim, N, ss[0], ss[1], ss[2], p->t3_D, hg, ns,
                                     2 * ns + 40, (const double *)p->d_quad,
                                     (T *)p->d_t3_deconv, p->stream);
    if (rc) return rc;
    for (int d = 0; d < 3; ++d) {
        p->t3_x[d] = xs[d];
        p->t3_s[d] = ss[d];
    }
    return JFA_OK;
}

int jfa::type3_setpts(jfa_plan p, int64_t M, const void *x, const void *y, const void *z,
                      int64_t N, const void *s, const void *t, const void *u) {
    if (M < 0 || M > INT32_MAX || N < 0 || N > INT32_MAX) {
        set_error("type-3 point counts must be in [0, 2^31)");
        return JFA_ERR_TOO_LARGE;
    }
    const void *xs[3] = {x, y, z}, *ss[3] = {s, t, u};
    for (int d = 0; d < p->dim; ++d) {
        if ((M > 0 && !xs[d]) || (N > 0 && !ss[d])) {
            set_error("type-3 coordinates for dimension %d are NULL", d);
            return JFA_ERR_BAD_ARG;
        }
    }
    for (int d = p->dim; d < 3; ++d) xs[d] = ss[d] = nullptr;
    if (p->is_double) {
        const double *xd[3] = {(const double *)xs[0], (const double *)xs[1], (const double *)xs[2]};
        const double *sd[3] = {(const double *)ss[0], (const double *)ss[1], (const double *)ss[2]};
        return type3_setpts_t<double>(p, M, xd, N, sd);
    }
    const float *xf[3] = {(const float *)xs[0], (const float *)xs[1], (const float *)xs[2]};
    const float *sf[3] = {(const float *)ss[0], (const float *)ss[1], (const float *)ss[2]};
    return type3_setpts_t<float>(p, M, xf, N, sf);
}

template <typename T>
static int type3_execute_t(jfa_plan p, const T *c, T *f) {
    const int dim = p->dim;
    const int64_t M = p->M, N = p->Ntot;
    if (N == 0) return JFA_OK;
    if (M == 0) return launch_fill_zero(f, (int64_t)sizeof(T) * 2 * N * p->ntransf, p->stream);
    const T *x[3] = {(const T *)p->t3_x[0], (const T *)p->t3_x[1], (const T *)p->t3_x[2]};
    const T *s[3] = {(const T *)p->t3_s[0], (const T *)p->t3_s[1], (const T *)p->t3_s[2]};
    const SortedPts<T> pts = sorted_view<T>(p);
    const BinIndex bi = bin_view(p);
    const int kem = p->opts.gpu_kerevalmeth ? 1 : 0;
    bool phase = false;
    for (int d = 0; d < dim; ++d) phase = phase || p->t3_D[d] != 0.0;
    T *fw = (T *)p->d_fw;
    int rc;
    for (int t0 = 0; t0 < p->ntransf; t0 += p->batch) {
        const int nb = std::min(p->batch, p->ntransf - t0);
        const T *cb = c + 2 * (int64_t)t0 * M;
        T *fb = f + 2 * (int64_t)t0 * N;
        const T *src = cb;
        if (phase) {
            if ((rc = launch_t3_prephase<T>(dim, M, nb, x[0], x[1], x[2], p->t3_C, p->t3_D,
                                            p->iflag, cb, (T *)p->d_cp, p->stream)))
                return rc;
            src = (const T *)p->d_cp;
        }
        // the strip spreader writes every cell itself
        if (p->method != JFA_METHOD_STRIP)
            if ((rc = launch_fill_zero(fw, (int64_t)sizeof(T) * 2 * p->F * nb, p->stream)))
                return rc;
        if ((rc = spread_any<T>(p, src, M, fw, p->F, nb))) return rc;
        const int saved = p->inner->ntransf;
        p->inner->ntransf = nb;
        rc = execute_12<T>(p->inner, fb, fw);
        p->inner->ntransf = saved;
        if (rc) return rc;
        if ((rc = launch_t3_postphase<T>(dim, N, nb, s[0], s[1], s[2], p->t3_C, p->iflag,
                                         (const T *)p->d_t3_deconv, fb, p->stream)))
            return rc;
    }
    return JFA_OK;
}

int jfa::type3_execute(jfa_plan p, void *c, void *fk) {
    if ((p->M > 0 && !c) || (p->Ntot > 0 && !fk)) {
        set_error("execute: NULL data pointer");
        return JFA_ERR_BAD_ARG;
    }
    if (!p->inner && p->M > 0 && p->Ntot > 0) {
        set_error("type-3 execute before setpts");
        return JFA_ERR_BAD_ARG;
    }
    if (p->is_double) return type3_execute_t<double>(p, (const double *)c, (double *)fk);
    return type3_execute_t<float>(p, (const float *)c, (float *)fk);
}
