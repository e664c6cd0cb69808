// Deconvolution (type 1) and amplification (type 2) between the fine grid and
// the mode array, with the modeord reordering fused in.  Replaces the
// deconvolve/amplify stage cuFINUFFT runs inside execute (called at
// lib/kernels.cc.cu:75; source absent from the reference).
//
// Mode index i along a dimension of size n maps to frequency k:
//   modeord 0: k = i - n/2 (floor)       modeord 1: k = i (i < (n+1)/2), i - n
// (src/jax_finufft/ops.py:114-123).  Frequency k lives at fine-grid cell
// k mod nf.  dk_d[k + n/2] holds 1 / psi_hat_d(k) for the plan.
#include "jfa_device.hpp"

namespace jfa {
namespace {

__device__ __forceinline__ int64_t freq_of(int64_t i, int64_t n, int modeord) {
    if (modeord == 0) return i - n / 2;
    return i < (n + 1) / 2 ? i : i - n;
}

// out[t][i] = fw[t][k mod nf] * prod_d dk_d[k_d + n_d/2]
template <typename T, int DIM>
__global__ void __launch_bounds__(256)
    deconv_kernel(int64_t n0, int64_t n1, int64_t n2, int64_t nf0, int64_t nf1, int64_t nf2,
                  const T *__restrict__ dk0, const T *__restrict__ dk1,
                  const T *__restrict__ dk2, int modeord, const T *__restrict__ fw,
                  int64_t fw_stride, T *__restrict__ out, int64_t out_stride) {
    const int64_t ntot = n0 * n1 * n2;
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= ntot) return;
    fw += 2 * fw_stride * blockIdx.y;
    out += 2 * out_stride * blockIdx.y;
    const int64_t i0 = q % n0;
    const int64_t k0 = freq_of(i0, n0, modeord);
    T s = dk0[k0 + n0 / 2];
    int64_t l = k0 < 0 ? k0 + nf0 : k0;
    if constexpr (DIM > 1) {
        const int64_t i1 = (q / n0) % n1;
        const int64_t k1 = freq_of(i1, n1, modeord);
        s *= dk1[k1 + n1 / 2];
        l += nf0 * (k1 < 0 ? k1 + nf1 : k1);
    }
    if constexpr (DIM > 2) {
        const int64_t i2 = q / (n0 * n1);
        const int64_t k2 = freq_of(i2, n2, modeord);
        s *= dk2[k2 + n2 / 2];
        l += nf0 * nf1 * (k2 < 0 ? k2 + nf2 : k2);
    }
    out[2 * q] = fw[2 * l] * s;
    out[2 * q + 1] = fw[2 * l + 1] * s;
}

// fw[t][l] = f[t][i(k)] * prod_d dk_d[k_d + n_d/2] where cell l holds a mode,
// 0 elsewhere.  Every fine-grid cell is written, so no separate zero pass.
template <typename T, int DIM>
__global__ void __launch_bounds__(256)
    amplify_kernel(int64_t n0, int64_t n1, int64_t n2, int64_t nf0, int64_t nf1, int64_t nf2,
                   const T *__restrict__ dk0, const T *__restrict__ dk1,
                   const T *__restrict__ dk2, int modeord, const T *__restrict__ f,
                   int64_t f_stride, T *__restrict__ fw, int64_t fw_stride) {
    const int64_t F = nf0 * nf1 * nf2;
    const int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (l >= F) return;
    f += 2 * f_stride * blockIdx.y;
    fw += 2 * fw_stride * blockIdx.y;
    bool ok = true;
    T s = T(1);
    int64_t idx = 0;
    {
        const int64_t l0 = l % nf0;
        const int64_t k = l0 < (nf0 + 1) / 2 ? l0 : l0 - nf0;
        ok = ok && (k >= -(n0 / 2)) && (k <= (n0 - 1) / 2);
        if (ok) {
            s *= dk0[k + n0 / 2];
            idx += modeord == 0 ? k + n0 / 2 : (k < 0 ? k + n0 : k);
        }
    }
    if constexpr (DIM > 1) {
        const int64_t l1 = (l / nf0) % nf1;
        const int64_t k = l1 < (nf1 + 1) / 2 ? l1 : l1 - nf1;
        ok = ok && (k >= -(n1 / 2)) && (k <= (n1 - 1) / 2);
        if (ok) {
            s *= dk1[k + n1 / 2];
            idx += n0 * (modeord == 0 ? k + n1 / 2 : (k < 0 ? k + n1 : k));
        }
    }
    if constexpr (DIM > 2) {
        const int64_t l2 = l / (nf0 * nf1);
        const int64_t k = l2 < (nf2 + 1) / 2 ? l2 : l2 - nf2;
        ok = ok && (k >= -(n2 / 2)) && (k <= (n2 - 1) / 2);
        if (ok) {
            s *= dk2[k + n2 / 2];
            idx += n0 * n1 * (modeord == 0 ? k + n2 / 2 : (k < 0 ? k + n2 : k));
        }
    }
    T vr = T(0), vi = T(0);
    if (ok) {
        vr = f[2 * idx] * s;
        vi = f[2 * idx + 1] * s;
    }
    fw[2 * l] = vr;
    fw[2 * l + 1] = vi;
}

// Ragged batch: block b deconvolves a 32 x 32 tile of modes of the virtual
// problem v with mode_blk[v] <= b < mode_blk[v + 1], reading its pruned-FFT
// buffer G[j0][y] (j0 the compact column of k0, y = k1 mod nf1) along y and
// writing the modes along x through LDS; the output of problem i is f_i
// (pointer table entry 3*nprob + i), transform-major.
template <typename T>
__global__ void __launch_bounds__(256)
    ragged_deconv_kernel(const RaggedVirt *__restrict__ virt, const int64_t *__restrict__ mode_blk,
                         int nv, const T *__restrict__ dk, int modeord,
                         const T *__restrict__ g_all, const void *const *__restrict__ ptrs,
                         int nprob) {
    __shared__ T tr[32][33], ti[32][33];
    int lo = 0, hi = nv;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (mode_blk[mid] <= (int64_t)blockIdx.x)
            lo = mid;
        else
            hi = mid;
    }
    const RaggedVirt d = virt[lo];
    const int64_t n0 = d.n0, n1 = d.n1, nf1 = d.sg.nf1;
    const int64_t tiles0 = (n0 + 31) / 32;
    const int64_t tb = (int64_t)blockIdx.x - mode_blk[lo];
    const int64_t xb = tb % tiles0, yb = tb / tiles0;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const T *G = g_all + 2 * d.g_off;
    // read along y
    const int64_t i1 = yb * 32 + tx;
    int64_t y = -1;
    if (i1 < n1) {
        const int64_t k1 = freq_of(i1, n1, modeord);
        y = k1 < 0 ? k1 + nf1 : k1;
    }
    for (int k = ty; k < 32; k += 8) {
        const int64_t i0 = xb * 32 + k;
        T vr = T(0), vi = T(0);
        if (y >= 0 && i0 < n0) {
            const int64_t k0 = freq_of(i0, n0, modeord);
            const int64_t l = (k0 < 0 ? k0 + n0 : k0) * nf1 + y;
            vr = G[2 * l];
            vi = G[2 * l + 1];
        }
        tr[k][tx] = vr;
        ti[k][tx] = vi;
    }
    __syncthreads();
    // write along x
    const int64_t i0 = xb * 32 + tx;
    if (i0 >= n0) return;
    const T *dk0 = dk + d.dk_off;
    const T *dk1 = dk0 + n0;
    const T s0 = dk0[freq_of(i0, n0, modeord) + n0 / 2];
    T *out = (T *)ptrs[3 * nprob + d.prob] + 2 * (int64_t)d.t * n0 * n1;
    for (int k = ty; k < 32; k += 8) {
        const int64_t ii = yb * 32 + k;
        if (ii >= n1) break;
        const T s = s0 * dk1[freq_of(ii, n1, modeord) + n1 / 2];
        const int64_t q = i0 + n0 * ii;
        out[2 * q] = tr[tx][k] * s;
        out[2 * q + 1] = ti[tx][k] * s;
    }
}

}  // namespace

template <typename T>
int launch_ragged_deconv(const RaggedLaunch &rl, const T *dk, int modeord, const T *g_all,
                         hipStream_t stream) {
    if (rl.mode_blocks == 0) return JFA_OK;
    ragged_deconv_kernel<T><<<(unsigned)rl.mode_blocks, 256, 0, stream>>>(
        rl.virt, rl.mode_blk, rl.nv, dk, modeord, g_all, rl.ptrs, rl.nprob);
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

template <typename T>
int launch_deconv(int dim, const int64_t *n, const int64_t *nf, const T *dk0, const T *dk1,
                  const T *dk2, int modeord, const T *fw, int64_t fw_stride, T *out,
                  int64_t out_stride, int ntr, hipStream_t stream) {
    const int64_t ntot = n[0] * n[1] * n[2];
    if (ntot == 0 || ntr == 0) return JFA_OK;
    dim3 grid((unsigned)((ntot + 255) / 256), ntr);
    JFA_DISPATCH_DIM(dim, deconv_kernel<T, DIM_><<<grid, 256, 0, stream>>>(
                              n[0], n[1], n[2], nf[0], nf[1], nf[2], dk0, dk1, dk2, modeord, fw,
                              fw_stride, out, out_stride));
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

template <typename T>
int launch_amplify(int dim, const int64_t *n, const int64_t *nf, const T *dk0, const T *dk1,
                   const T *dk2, int modeord, const T *f, int64_t f_stride, T *fw,
                   int64_t fw_stride, int ntr, hipStream_t stream) {
    const int64_t F = nf[0] * nf[1] * nf[2];
    if (F == 0 || ntr == 0) return JFA_OK;
    dim3 grid((unsigned)((F + 255) / 256), ntr);
    JFA_DISPATCH_DIM(dim, amplify_kernel<T, DIM_><<<grid, 256, 0, stream>>>(
                              n[0], n[1], n[2], nf[0], nf[1], nf[2], dk0, dk1, dk2, modeord, f,
                              f_stride, fw, fw_stride));
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

#define JFA_INST(T)                                                                           \
    template int launch_ragged_deconv<T>(const RaggedLaunch &, const T *, int, const T *,      \
                                         hipStream_t);                                         \
    template int launch_deconv<T>(int, const int64_t *, const int64_t *, const T *, const T *, \
                                  const T *, int, const T *, int64_t, T *, int64_t, int,       \
                                  hipStream_t);                                                \
    template int launch_amplify<T>(int, const int64_t *, const int64_t *, const T *,          \
                                   const T *, const T *, int, const T *, int64_t, T *,         \
                                   int64_t, int, hipStream_t);
JFA_INST(float)
JFA_INST(double)
#undef JFA_INST

}  // namespace jfa
