// setpts: fold points onto the fine grid, bin them, and counting-sort them
// into bin order.  Replaces the point sort cuFINUFFT runs inside setpts
// (called at lib/kernels.cc.cu:63-68; device code absent from the reference).
//
// Bin of a point = bin of its first tap l0 (wrapped), so every tap of a point
// in bin b lies in the tile [origin(b), origin(b) + bs + ns - 1).  The integer
// tap position is produced by fold_point() from the coordinate in double, and
// the same function is called by the binning and by the scatter pass, so the
// bin a point is counted in and the tap window the spreader uses agree exactly.
#include "jfa_device.hpp"

namespace jfa {

namespace {

constexpr int kBinThreads = 256;
constexpr int kBinPPT = 16;  // points per thread in the LDS-histogram pass

template <typename T, int DIM>
__device__ __forceinline__ int bin_of_point(const GridGeom &g, const T *x, const T *y,
                                            const T *z, int64_t j) {
    Folded f0 = fold_point((double)x[j], g.nf[0], g.scale[0], g.shift[0], g.ns);
    int b = f0.l0w / g.bs[0];
    if constexpr (DIM > 1) {
        Folded f1 = fold_point((double)y[j], g.nf[1], g.scale[1], g.shift[1], g.ns);
        if (DIM == 2 && g.ymajor)
            b = b * g.nb[1] + f1.l0w / g.bs[1];
        else
            b += (f1.l0w / g.bs[1]) * g.nb[0];
    }
    if constexpr (DIM > 2) {
        Folded f2 = fold_point((double)z[j], g.nf[2], g.scale[2], g.shift[2], g.ns);
        b += (f2.l0w / g.bs[2]) * (g.nb[0] * g.nb[1]);
    }
    return b;
}

// Direct version: one global atomic per point (used when the bin histogram
// does not fit in LDS).
template <typename T, int DIM>
__global__ void __launch_bounds__(kBinThreads)
    fold_bin_global_kernel(GridGeom g, int64_t M, const T *__restrict__ x,
                           const T *__restrict__ y, const T *__restrict__ z, int *bin_of,
                           int *rank_of, int *bin_count) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    int b = bin_of_point<T, DIM>(g, x, y, z, j);
    bin_of[j] = b;
    rank_of[j] = atomicAdd(&bin_count[b], 1);
}

// LDS-histogram version: ranks within the block come from LDS atomics, then
// one global atomic per (block, non-empty bin) reserves the block's slice of
// each bin.  Clustered inputs hit a handful of bins; this keeps their global
// atomic traffic to one per block instead of one per point.
template <typename T, int DIM>
__global__ void __launch_bounds__(kBinThreads)
    fold_bin_lds_kernel(GridGeom g, int64_t M, const T *__restrict__ x, const T *__restrict__ y,
                        const T *__restrict__ z, int *bin_of, int *rank_of, int *bin_count) {
    extern __shared__ int hist[];
    const int nbins = (int)g.nbins;
    for (int b = threadIdx.x; b < nbins; b += kBinThreads) hist[b] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * (kBinThreads * kBinPPT);
    int bins[kBinPPT], ranks[kBinPPT];
#pragma unroll
    for (int k = 0; k < kBinPPT; ++k) {
        int64_t j = base + (int64_t)k * kBinThreads + threadIdx.x;
        bins[k] = -1;
        if (j < M) {
            bins[k] = bin_of_point<T, DIM>(g, x, y, z, j);
            ranks[k] = atomicAdd(&hist[bins[k]], 1);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += kBinThreads) {
        int h = hist[b];
        hist[b] = h ? atomicAdd(&bin_count[b], h) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kBinPPT; ++k) {
        int64_t j = base + (int64_t)k * kBinThreads + threadIdx.x;
        if (bins[k] >= 0) {
            bin_of[j] = bins[k];
            rank_of[j] = ranks[k] + hist[bins[k]];
        }
    }
}

// One workgroup: exclusive prefix sums of the bin counts and of the number of
// subproblems per bin (ceil(count / maxsub)); writes the totals at [nbins] and
// the subproblem total to *nsub.
__global__ void __launch_bounds__(1024)
    bin_scan_kernel(int64_t nbins, const int *__restrict__ count, int *bin_start,
                    int *sub_start, int *nsub, int maxsub) {
    __shared__ int sa[1024];
    __shared__ int sb[1024];
    const int tid = threadIdx.x;
    const int64_t per = (nbins + 1023) / 1024;
    const int64_t lo = tid * per;
    const int64_t hi = lo + per < nbins ? lo + per : nbins;
    int a = 0, s = 0;
    for (int64_t i = lo; i < hi; ++i) {
        int cnt = count[i];
        a += cnt;
        s += (cnt + maxsub - 1) / maxsub;
    }
    sa[tid] = a;
    sb[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int va = tid >= off ? sa[tid - off] : 0;
        int vb = tid >= off ? sb[tid - off] : 0;
        __syncthreads();
        sa[tid] += va;
        sb[tid] += vb;
        __syncthreads();
    }
    int ea = sa[tid] - a, eb = sb[tid] - s;
    for (int64_t i = lo; i < hi; ++i) {
        int cnt = count[i];
        bin_start[i] = ea;
        sub_start[i] = eb;
        ea += cnt;
        eb += (cnt + maxsub - 1) / maxsub;
    }
    if (tid == 1023) {
        bin_start[nbins] = sa[1023];
        sub_start[nbins] = sb[1023];
        *nsub = sb[1023];
    }
}

// Writes each point's original index, wrapped first tap and Horner variable at
// its sorted position (or in input order when `sorted` is 0).
template <typename T, int DIM>
__global__ void __launch_bounds__(256)
    scatter_kernel(GridGeom g, int64_t M, const T *__restrict__ x, const T *__restrict__ y,
                   const T *__restrict__ z, const int *__restrict__ bin_of,
                   const int *__restrict__ rank_of, const int *__restrict__ bin_start,
                   int sorted, int *idx, int *l0, T *tv) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    int64_t pos = sorted ? (int64_t)bin_start[bin_of[j]] + rank_of[j] : j;
    idx[pos] = (int)j;
    const T *xs[3] = {x, y, z};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        Folded f = fold_point((double)xs[d][j], g.nf[d], g.scale[d], g.shift[d], g.ns);
        l0[d * M + pos] = f.l0w;
        tv[d * M + pos] = (T)f.t;
    }
}

__global__ void fill_zero_kernel(uint4 *p, int64_t n16, unsigned char *tail, int ntail) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

}  // namespace

int launch_fill_zero(void *p, int64_t bytes, hipStream_t stream) {
    if (bytes <= 0) return JFA_OK;
    if (((uintptr_t)p & 15) != 0) {
        set_error("fill_zero: pointer not 16-byte aligned");
        return JFA_ERR_BAD_ARG;
    }
    int64_t n16 = bytes / 16;
    int ntail = (int)(bytes - n16 * 16);
    int64_t blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    fill_zero_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(
        (uint4 *)p, n16, (unsigned char *)p + n16 * 16, ntail);
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

template <typename T>
int launch_fold_bin(const GridGeom &g, int64_t M, const T *x, const T *y, const T *z,
                    int *bin_of, int *rank_of, int *bin_count, hipStream_t stream) {
    if (M == 0) return JFA_OK;
    const size_t lds = sizeof(int) * (size_t)g.nbins;
    if (g.nbins <= 8192) {
        int64_t per_block = (int64_t)kBinThreads * kBinPPT;
        unsigned blocks = (unsigned)((M + per_block - 1) / per_block);
        JFA_DISPATCH_DIM(g.dim, fold_bin_lds_kernel<T, DIM_><<<blocks, kBinThreads, lds, stream>>>(
                                    g, M, x, y, z, bin_of, rank_of, bin_count));
    } else {
        unsigned blocks = (unsigned)((M + kBinThreads - 1) / kBinThreads);
        JFA_DISPATCH_DIM(g.dim, fold_bin_global_kernel<T, DIM_><<<blocks, kBinThreads, 0, stream>>>(
                                    g, M, x, y, z, bin_of, rank_of, bin_count));
    }
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

int launch_bin_scan(const GridGeom &g, const int *bin_count, int *bin_start, int *sub_start,
                    int *nsub, int maxsub, hipStream_t stream) {
    bin_scan_kernel<<<1, 1024, 0, stream>>>(g.nbins, bin_count, bin_start, sub_start, nsub,
                                            maxsub);
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

template <typename T>
int launch_scatter(const GridGeom &g, int64_t M, const T *x, const T *y, const T *z,
                   const int *bin_of, const int *rank_of, const int *bin_start, int sorted,
                   int *idx, int *l0, T *tv, hipStream_t stream) {
    if (M == 0) return JFA_OK;
    unsigned blocks = (unsigned)((M + 255) / 256);
    JFA_DISPATCH_DIM(g.dim, scatter_kernel<T, DIM_><<<blocks, 256, 0, stream>>>(
                                g, M, x, y, z, bin_of, rank_of, bin_start, sorted, idx, l0, tv));
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

template int launch_fold_bin<float>(const GridGeom &, int64_t, const float *, const float *,
                                    const float *, int *, int *, int *, hipStream_t);
template int launch_fold_bin<double>(const GridGeom &, int64_t, const double *, const double *,
                                     const double *, int *, int *, int *, hipStream_t);
template int launch_scatter<float>(const GridGeom &, int64_t, const float *, const float *,
                                   const float *, const int *, const int *, const int *, int,
                                   int *, int *, float *, hipStream_t);
template int launch_scatter<double>(const GridGeom &, int64_t, const double *, const double *,
                                    const double *, const int *, const int *, const int *, int,
                                    int *, int *, double *, hipStream_t);

}  // namespace jfa
