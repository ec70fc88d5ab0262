// Row passes with the block's rows staged through LDS.
//
// A workgroup's S row pairs are 2S consecutive rows, one contiguous span of
// memory both on input and on output, so it is moved between HBM and LDS as
// one linear copy of 16-byte vectors (every lane of a wavefront touching
// consecutive 16 bytes).  The four-step transform itself is the register
// one of fft2.hpp, reading and writing the staged rows in LDS.  Same
// arguments and results as k_rows_r2c_v2 / k_rows_c2r_v2.
#pragma once

#include "fft2.hpp"

namespace lg {

template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };

// n elements of T, 16-byte aligned at both ends, src -> dst by the block
template <typename T>
__device__ __forceinline__ void block_copy16(const T* __restrict__ src, T* __restrict__ dst,
                                             long n) {
    using V = typename Vec16<T>::type;
    constexpr int per = (int)(16 / sizeof(T));
    const long nv = n / per;
    const V* s = reinterpret_cast<const V*>(src);
    V* d = reinterpret_cast<V*>(dst);
    for (long q = threadIdx.x; q < nv; q += blockDim.x) d[q] = s[q];
}

// elements of the row stage: 2S rows of max(ld_real, 2 * ncols) (input and
// output rows share it, each at its own contiguous row width)
__host__ __device__ inline long stage_elems(int S, int ld_real, int ncols) {
    const int w = ld_real > 2 * ncols ? ld_real : 2 * ncols;
    return 2L * S * w;
}

// r2c: rows of src (src_ld reals, N used) -> half spectra (ncols complex).
// Launcher: S * N1 <= blockDim.x; LDS = stage + S * YS complex.
template <typename T, int N1, int N2>
__global__ void __launch_bounds__(256) k_rows_r2c_v3(RowR2CArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    extern __shared__ __align__(16) unsigned char smem_r3[];
    const auto* tw = plan_tw<T>(a.plan);
    const int job = blockIdx.z;
    const int mem = blockIdx.y / a.nlev;
    const int plane = blockIdx.y - mem * a.nlev;
    const int npairs = a.nrows / 2;
    const int pair0 = blockIdx.x * a.S;
    const int S = min(a.S, npairs - pair0);
    const int nc = a.ncols;
    const int sld = a.src_ld;
    T* stage = reinterpret_cast<T*>(smem_r3);
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_r3 + ((stage_elems(a.S, sld, nc) * sizeof(T) + 15) & ~15L));
    const T* src = a.src[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane +
                   (long)(2 * pair0) * sld;
    cplx<T>* dst = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane +
                   (long)(2 * pair0) * nc;
    // rows 2 pair0 .. 2 (pair0 + S) - 1 are contiguous: stage them as they are
    block_copy16<T>(src, stage, 2L * S * sld);
    __syncthreads();
    const T sc = a.mscale ? (T)a.mscale[(long)mem * a.mscale_stride] : a.scale[job];
    for (int t = threadIdx.x; t < S * N2; t += blockDim.x) {
        const int s = t / N2, n2 = t - s * N2;
        const T* ra = stage + (long)(2 * s) * sld;
        const T* rb = ra + sld;
        cplx<T> x[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) x[n1] = cplx<T>{sc * ra[n2 + N2 * n1], sc * rb[n2 + N2 * n1]};
        dft_reg<T, N1, -1>(x);
        cplx<T>* Y = lds + s * SP::YS;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1)
            Y[k1 * SP::YR + n2] = (k1 == 0 || n2 == 0) ? x[k1] : cmul(x[k1], wbig<T, -1>(tw, n2 * k1));
    }
    __syncthreads();
    // step 2 in one round (S * N1 <= blockDim.x), X over Y in natural order
    cplx<T> y[N2];
    const int t2 = threadIdx.x;
    const bool act = t2 < S * N1;
    const int s2 = act ? t2 / N1 : 0;
    const int k1 = act ? t2 - s2 * N1 : 0;
    if (act) {
        const cplx<T>* Y = lds + s2 * SP::YS + k1 * SP::YR;
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) y[n2] = Y[n2];
        dft_reg<T, N2, -1>(y);
    }
    __syncthreads();
    if (act) {
        cplx<T>* X = lds + s2 * SP::YS;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) X[k1 + N1 * k2] = y[k2];
    }
    __syncthreads();
    // split the paired transform into the two half spectra, staged as
    // 2S rows of nc complex (contiguous, as in dst)
    cplx<T>* ostage = reinterpret_cast<cplx<T>*>(stage);
    for (int idx = threadIdx.x; idx < S * nc; idx += blockDim.x) {
        const int s = idx / nc;
        const int k = idx - s * nc;
        const cplx<T>* X = lds + s * SP::YS;
        const cplx<T> z = X[k];
        const cplx<T> zc = X[(N - k) % N];
        ostage[(long)(2 * s) * nc + k] = cplx<T>{(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y - zc.y)};
        const cplx<T> Bm{(T)0.5 * (z.x - zc.x), (T)0.5 * (z.y + zc.y)};
        ostage[(long)(2 * s + 1) * nc + k] = cplx<T>{Bm.y, -Bm.x};
    }
    __syncthreads();
    block_copy16<T>(reinterpret_cast<const T*>(ostage), reinterpret_cast<T*>(dst), 2L * S * 2 * nc);
}

// c2r: half spectra (ncols complex) -> rows of dst (dst_ld reals, N written,
// padding zero).  LDS = stage + S * YS complex.
template <typename T, int N1, int N2>
__global__ void __launch_bounds__(256) k_rows_c2r_v3(RowC2RArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    constexpr int half = N / 2;
    extern __shared__ __align__(16) unsigned char smem_c3[];
    const auto* tw = plan_tw<T>(a.plan);
    const int job = blockIdx.z;
    const int mem = blockIdx.y / a.nlev;
    const int plane = blockIdx.y - mem * a.nlev;
    const int npairs = a.nrows / 2;
    const int pair0 = blockIdx.x * a.S;
    const int S = min(a.S, npairs - pair0);
    const int nc = a.ncols;
    const int nce = a.nce[job] ? a.nce[job] : nc;
    const int ld = a.dst_ld;
    T* stage = reinterpret_cast<T*>(smem_c3);
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_c3 + ((stage_elems(a.S, ld, nc) * sizeof(T) + 15) & ~15L));
    const cplx<T>* src = a.src[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane +
                         (long)(2 * pair0) * nc;
    T* dst = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane +
             (long)(2 * pair0) * ld;
    // the 2S input rows (nc complex each) are contiguous: stage them as they are
    block_copy16<T>(reinterpret_cast<const T*>(src), stage, 2L * S * 2 * nc);
    __syncthreads();
    for (int t = threadIdx.x; t < S * N2; t += blockDim.x) {
        const int s = t / N2, n2 = t - s * N2;
        const cplx<T>* ra = reinterpret_cast<const cplx<T>*>(stage) + (long)(2 * s) * nc;
        const cplx<T>* rb = ra + nc;
        cplx<T> x[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
            const int k = n2 + N2 * n1;
            cplx<T> z{(T)0, (T)0};
            if (k == 0) {
                z = cplx<T>{ra[0].x, rb[0].x};
            } else if (k < half) {
                if (k < nce) {
                    const cplx<T> A = ra[k], B = rb[k];
                    z = cplx<T>{A.x - B.y, A.y + B.x};
                }
            } else if (k > half) {
                const int kk = N - k;
                if (kk < nce) {
                    const cplx<T> A = ra[kk], B = rb[kk];
                    z = cplx<T>{A.x + B.y, B.x - A.y};
                }
            }
            x[n1] = z;
        }
        dft_reg<T, N1, +1>(x);
        cplx<T>* Y = lds + s * SP::YS;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1)
            Y[k1 * SP::YR + n2] = (k1 == 0 || n2 == 0) ? x[k1] : cmul(x[k1], wbig<T, +1>(tw, n2 * k1));
    }
    __syncthreads();
    // the input stage is consumed: the output rows (width ld) go over it
    for (int t = threadIdx.x; t < S * N1; t += blockDim.x) {
        const int s = t / N1, k1 = t - s * N1;
        const cplx<T>* Y = lds + s * SP::YS + k1 * SP::YR;
        cplx<T> y[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) y[n2] = Y[n2];
        dft_reg<T, N2, +1>(y);
        T* oa = stage + (long)(2 * s) * ld;
        T* ob = oa + ld;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            oa[k1 + N1 * k2] = y[k2].x;
            ob[k1 + N1 * k2] = y[k2].y;
        }
    }
    const int npad = ld - N;
    for (int idx = threadIdx.x; idx < 2 * S * npad; idx += blockDim.x) {
        const int r = idx / npad, q = idx - r * npad;
        stage[(long)r * ld + N + q] = (T)0;
    }
    __syncthreads();
    block_copy16<T>(stage, dst, 2L * S * ld);
}

}  // namespace lg
