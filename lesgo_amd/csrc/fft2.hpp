// Register-resident plane FFT passes for the common transform lengths.
//
// A length N = N1 * N2 transform is done as a four-step FFT: every task
// first takes the N1 points x[n2 + N2 n1] (n1 < N1) into registers and does
// a length-N1 DFT there, multiplies by W_N^(n2 k1), and writes them to LDS;
// after one barrier every task takes the N2 points of one k1 and does the
// length-N2 DFT in registers, giving X[k1 + N1 k2].  Both small DFTs are
// unrolled at compile time (recursive radix 4/3/2 with constant twiddles),
// so one LDS transpose replaces the log(N) LDS stages of the generic pass.
// LDS rows are padded so that consecutive lanes fall in different banks.
// Same conventions and arguments as the generic passes in fft.hpp.
#pragma once

#include "fft.hpp"
#include "twiddle144.hpp"

namespace lg {

// W_n^j for the small DFTs (n divides 144); DIR = -1 forward, +1 inverse.
template <typename T, int n, int j, int DIR>
__device__ __forceinline__ cplx<T> wsmall() {
    constexpr int m = ((j % n) * (144 / n)) % 144;
    return cplx<T>{(T)kCos144[m], (T)(DIR * kSin144[m])};
}

template <typename T, int DIR>
__device__ __forceinline__ void bf2(cplx<T>& a, cplx<T>& b) {
    const cplx<T> t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <typename T, int DIR>
__device__ __forceinline__ void bf3(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2) {
    const T c = (T)-0.5;
    const T sn = (T)(DIR * 0.86602540378443864676);
    const cplx<T> s = cadd(a1, a2), d = csub(a1, a2);
    const cplx<T> m{a0.x + c * s.x, a0.y + c * s.y};
    const cplx<T> id{-sn * d.y, sn * d.x};
    a0 = cadd(a0, s);
    a1 = cadd(m, id);
    a2 = csub(m, id);
}

template <typename T, int DIR>
__device__ __forceinline__ void bf4(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3) {
    const cplx<T> s02 = cadd(a0, a2), d02 = csub(a0, a2);
    const cplx<T> s13 = cadd(a1, a3), d13 = csub(a1, a3);
    const cplx<T> jd13 = (DIR < 0) ? cplx<T>{d13.y, -d13.x} : cplx<T>{-d13.y, d13.x};
    a0 = cadd(s02, s13);
    a1 = cadd(d02, jd13);
    a2 = csub(s02, s13);
    a3 = csub(d02, jd13);
}

template <typename T, int N, int R, int M, int DIR, int K1 = 0>
__device__ __forceinline__ void combine(cplx<T> (&x)[N], cplx<T> (&sub)[R][M]);

// In-place DFT of N register values (natural order in and out).
template <typename T, int N, int DIR>
__device__ __forceinline__ void dft_reg(cplx<T> (&x)[N]) {
    if constexpr (N == 1) {
        return;
    } else if constexpr (N == 2) {
        bf2<T, DIR>(x[0], x[1]);
    } else if constexpr (N == 3) {
        bf3<T, DIR>(x[0], x[1], x[2]);
    } else if constexpr (N == 4) {
        bf4<T, DIR>(x[0], x[1], x[2], x[3]);
    } else {
        constexpr int R = (N % 4 == 0) ? 4 : ((N % 3 == 0) ? 3 : 2);
        constexpr int M = N / R;
        cplx<T> sub[R][M];
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
            for (int j = 0; j < M; ++j) sub[q][j] = x[q + R * j];
#pragma unroll
        for (int q = 0; q < R; ++q) dft_reg<T, M, DIR>(sub[q]);
        // X[k1 + M k2] = sum_q W_N^(q k1) W_R^(q k2) sub[q][k1]
        combine<T, N, R, M, DIR>(x, sub);
    }
}

template <typename T, int N, int R, int M, int DIR, int K1>
__device__ __forceinline__ void combine(cplx<T> (&x)[N], cplx<T> (&sub)[R][M]) {
    if constexpr (K1 < M) {
        cplx<T> t[R];
        t[0] = sub[0][K1];
        if constexpr (R > 1) t[1] = (K1 == 0) ? sub[1][K1] : cmul(sub[1][K1], wsmall<T, N, 1 * K1, DIR>());
        if constexpr (R > 2) t[2] = (K1 == 0) ? sub[2][K1] : cmul(sub[2][K1], wsmall<T, N, 2 * K1, DIR>());
        if constexpr (R > 3) t[3] = (K1 == 0) ? sub[3][K1] : cmul(sub[3][K1], wsmall<T, N, 3 * K1, DIR>());
        dft_reg<T, R, DIR>(t);
#pragma unroll
        for (int k2 = 0; k2 < R; ++k2) x[K1 + M * k2] = t[k2];
        combine<T, N, R, M, DIR, K1 + 1>(x, sub);
    }
}

// outer twiddle W_N^(e) from the plan table (e < N)
template <typename T, int DIR>
__device__ __forceinline__ cplx<T> wbig(const typename TwT<T>::type* tw, int e) {
    const auto w = tw[e];
    return cplx<T>{(T)w.x, (T)(-DIR) * (T)w.y};
}

template <int N1, int N2>
struct Split {
    static constexpr int n1 = N1, n2 = N2, N = N1 * N2;
    static constexpr int YR = N2 + 1;          // LDS row stride (complex)
    static constexpr int YS = N1 * YR + 1;     // LDS stride per sequence / column
};

// --------------------------------------------------------------------------
// Row pass r2c (paired rows), register four-step.
// --------------------------------------------------------------------------
template <typename T, int N1, int N2>
__global__ void __launch_bounds__(256) k_rows_r2c_v2(RowR2CArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    extern __shared__ __align__(16) unsigned char smem_r2[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_r2);
    const auto* tw = plan_tw<T>(a.plan);
    const int job = blockIdx.z;
    const int mem = blockIdx.y / a.nlev;
    const int plane = blockIdx.y - mem * a.nlev;
    const int npairs = a.nrows / 2;
    const int pair0 = blockIdx.x * a.S;
    const int S = min(a.S, npairs - pair0);
    const T* src = a.src[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane;
    cplx<T>* dst = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane;
    const T sc = a.mscale ? (T)a.mscale[(long)mem * a.mscale_stride] : a.scale[job];
    for (int t = threadIdx.x; t < S * N2; t += blockDim.x) {
        const int s = t / N2, n2 = t - s * N2;
        const int j = 2 * (pair0 + s);
        const T* ra = src + (long)j * a.src_ld;
        const T* rb = ra + a.src_ld;
        cplx<T> x[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) x[n1] = cplx<T>{sc * gld(ra + n2 + N2 * n1), sc * gld(rb + n2 + N2 * n1)};
        dft_reg<T, N1, -1>(x);
        cplx<T>* Y = lds + s * SP::YS;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1)
            Y[k1 * SP::YR + n2] = (k1 == 0 || n2 == 0) ? x[k1] : cmul(x[k1], wbig<T, -1>(tw, n2 * k1));
    }
    __syncthreads();
    // step 2 in one round (the launcher keeps S * N1 <= blockDim.x): all LDS
    // reads finish before the barrier, then X overwrites Y in natural order
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
    const int nc = a.ncols;
    const int nw = a.nwr ? a.nwr : nc;
    for (int idx = threadIdx.x; idx < S * nw; idx += blockDim.x) {
        const int s = idx / nw;
        const int k = idx - s * nw;
        const int j = 2 * (pair0 + s);
        const cplx<T>* X = lds + s * SP::YS;
        const cplx<T> z = X[k];
        const cplx<T> zc = X[(N - k) % N];
        const cplx<T> A{(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y - zc.y)};
        const cplx<T> Bm{(T)0.5 * (z.x - zc.x), (T)0.5 * (z.y + zc.y)};
        gst(dst + (long)j * nc + k, A);
        gst(dst + (long)(j + 1) * nc + k, cplx<T>{Bm.y, -Bm.x});
    }
}

// two consecutive reals of an output row (rows hold an even number of
// values and start at even offsets, so the pair is aligned to its size)
template <typename T>
__device__ __forceinline__ void st_pair2(T* p, T a, T b);
template <>
__device__ __forceinline__ void st_pair2<float>(float* p, float a, float b) {
    *reinterpret_cast<float2*>(p) = make_float2(a, b);
}
template <>
__device__ __forceinline__ void st_pair2<double>(double* p, double a, double b) {
    *reinterpret_cast<double2*>(p) = make_double2(a, b);
}

// --------------------------------------------------------------------------
// Row pass c2r (paired rows), register four-step.
// --------------------------------------------------------------------------
template <typename T, int N1, int N2>
__global__ void __launch_bounds__(256) k_rows_c2r_v2(RowC2RArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    constexpr int half = N / 2;
    extern __shared__ __align__(16) unsigned char smem_c2[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_c2);
    const auto* tw = plan_tw<T>(a.plan);
    const int job = blockIdx.z;
    const int mem = blockIdx.y / a.nlev;
    const int plane = blockIdx.y - mem * a.nlev;
    const int npairs = a.nrows / 2;
    const int pair0 = blockIdx.x * a.S;
    const int S = min(a.S, npairs - pair0);
    const cplx<T>* src = a.src[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane;
    T* dst = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane;
    const int nc = a.ncols;
    const int nce = a.nce[job] ? a.nce[job] : nc;
    T* const dst2 = a.dst2[job] ? a.dst2[job] + mem * a.dst_mstride +
                                      (long)(a.lev0_dst + plane) * a.dst_plane
                                : nullptr;
    // pass 1 (x derivative): the half spectra times i kx of their column
    auto ikx2 = [&](cplx<T>& A, cplx<T>& B, int col) {
        const T kx = a.kxs * (T)col;
        A = cplx<T>{-A.y * kx, A.x * kx};
        B = cplx<T>{-B.y * kx, B.x * kx};
    };
    for (int pass = 0; pass < (dst2 ? 2 : 1); ++pass) {
        T* const out = pass ? dst2 : dst;
        for (int t = threadIdx.x; t < S * N2; t += blockDim.x) {
            const int s = t / N2, n2 = t - s * N2;
            const int j = 2 * (pair0 + s);
            const cplx<T>* ra = src + (long)j * nc;
            const cplx<T>* rb = ra + nc;
            cplx<T> x[N1];
    #pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) {
                const int k = n2 + N2 * n1;
                cplx<T> z{(T)0, (T)0};
                if (k == 0) {
                    z = pass ? cplx<T>{(T)0, (T)0} : cplx<T>{gld(ra).x, gld(rb).x};
                } else if (k < half) {
                    if (k < nce) {
                        cplx<T> A = gld(ra + k), B = gld(rb + k);
                        if (pass) ikx2(A, B, k);
                        z = cplx<T>{A.x - B.y, A.y + B.x};
                    }
                } else if (k > half) {
                    const int kk = N - k;
                    if (kk < nce) {
                        cplx<T> A = gld(ra + kk), B = gld(rb + kk);
                        if (pass) ikx2(A, B, kk);
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
        const int ld = a.dst_ld;
        // N2 <= 16 with step 2 in one round: the results go back to LDS in
        // natural order (every read of the slot done before the barrier),
        // then rows are stored two values per lane, consecutive lanes on
        // consecutive pairs of a row.  (For N2 = 24 the N2 values held across
        // the barrier cost more occupancy than the wider stores gain:
        // 346 -> 376 us for the 384-point float pass.)
        bool staged = false;
        if constexpr (N2 <= 16) staged = S * N1 <= (int)blockDim.x;
        if (staged) {
            const int t = threadIdx.x;
            const bool act = t < S * N1;
            const int s = act ? t / N1 : 0, k1 = act ? t - s * N1 : 0;
            cplx<T> y[N2];
            if (act) {
                const cplx<T>* Y = lds + s * SP::YS + k1 * SP::YR;
    #pragma unroll
                for (int n2 = 0; n2 < N2; ++n2) y[n2] = Y[n2];
                dft_reg<T, N2, +1>(y);
            }
            __syncthreads();
            if (act) {
                cplx<T>* X = lds + s * SP::YS;
    #pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) X[k1 + N1 * k2] = y[k2];
            }
            __syncthreads();
            constexpr int NP = N / 2;
            for (int idx = threadIdx.x; idx < S * NP; idx += blockDim.x) {
                const int s2 = idx / NP, q = idx - s2 * NP;
                const int j = 2 * (pair0 + s2);
                const cplx<T>* X = lds + s2 * SP::YS + 2 * q;
                const cplx<T> p0 = X[0], p1 = X[1];
                T* oa = out + (long)j * ld + 2 * q;
                st_pair2(oa, p0.x, p1.x);
                st_pair2(oa + ld, p0.y, p1.y);
            }
        } else {
            for (int t = threadIdx.x; t < S * N1; t += blockDim.x) {
                const int s = t / N1, k1 = t - s * N1;
                const int j = 2 * (pair0 + s);
                const cplx<T>* Y = lds + s * SP::YS + k1 * SP::YR;
                cplx<T> y[N2];
    #pragma unroll
                for (int n2 = 0; n2 < N2; ++n2) y[n2] = Y[n2];
                dft_reg<T, N2, +1>(y);
                T* oa = out + (long)j * ld;
                T* ob = oa + ld;
    #pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) {
                    gst(oa + k1 + N1 * k2, y[k2].x);
                    gst(ob + k1 + N1 * k2, y[k2].y);
                }
            }
        }
        // padding columns N..ld-1
        const int npad = ld - N;
        for (int idx = threadIdx.x; idx < S * npad; idx += blockDim.x) {
            const int s = idx / npad, q = idx - s * npad;
            const int j = 2 * (pair0 + s);
            out[(long)j * ld + N + q] = (T)0;
            out[(long)(j + 1) * ld + N + q] = (T)0;
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------------------
// Column pass, register four-step; lanes run over columns for coalescing.
// --------------------------------------------------------------------------
template <typename T, int N1, int N2, int DIR>
__global__ void __launch_bounds__(256) k_cols_v2(ColArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    extern __shared__ __align__(16) unsigned char smem_cv[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_cv);
    const auto* tw = plan_tw<T>(a.plan);
    const int job = blockIdx.x;
    const int TC = a.TC;
    const int c0 = blockIdx.y * TC;
    const int nzc = a.nzc[job] ? a.nzc[job] : a.ncols;
    if (c0 >= nzc) return;
    const int tc = min(TC, nzc - c0);
    const int ykeep = a.ykeep[job];
    const int mem = blockIdx.z / a.nlev;
    const int plane = blockIdx.z - mem * a.nlev;
    const cplx<T>* src = a.src[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane;
    cplx<T>* dst = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane;
    const int op = a.op[job];
    const int ny = a.ny;
    const int nyh = ny / 2;
    const int CS = SP::YS;                     // per-column LDS block (odd-padded)
    for (int t = threadIdx.x; t < TC * N2; t += blockDim.x) {
        const int c = t % TC, n2 = t / TC;
        if (c >= tc) continue;
        cplx<T> x[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
            const int jb = n2 + N2 * n1;
            cplx<T> v{(T)0, (T)0};
            if (a.map == MAP_PAD) {
                int j = -1;
                if (jb < nyh) j = jb;
                else if (jb > N - nyh) j = jb - (N - ny);
                if (j >= 0) v = apply_op<T>(op, gld(src + (long)j * a.ncols + c0 + c), c0 + c, j, a);
            } else {
                v = (DIR > 0 && !in_yband(jb, a.ny, ykeep)) ? cplx<T>{(T)0, (T)0}
                                                              : gld(src + (long)jb * a.ncols + c0 + c);
                if (DIR > 0 && a.map == MAP_ID) {
                    if (a.nyq0[job] && jb == nyh) v = cplx<T>{(T)0, (T)0};
                    if (op == OP_DIVXY) {
                        const cplx<T>* s2 = a.src2[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane;
                        v = divxy<T>(v, gld(s2 + (long)jb * a.ncols + c0 + c), c0 + c, jb, a);
                    } else {
                        v = apply_op<T>(op, v, c0 + c, jb, a);
                    }
                }
            }
            x[n1] = v;
        }
        dft_reg<T, N1, DIR>(x);
        cplx<T>* Y = lds + c * CS;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1)
            Y[k1 * SP::YR + n2] = (k1 == 0 || n2 == 0) ? x[k1] : cmul(x[k1], wbig<T, DIR>(tw, n2 * k1));
    }
    __syncthreads();
    for (int t = threadIdx.x; t < TC * N1; t += blockDim.x) {
        const int c = t % TC, k1 = t / TC;
        if (c >= tc) continue;
        const cplx<T>* Y = lds + c * CS + k1 * SP::YR;
        cplx<T> y[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) y[n2] = Y[n2];
        dft_reg<T, N2, DIR>(y);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            const int jb = k1 + N1 * k2;
            cplx<T> v = y[k2];
            if (a.map == MAP_UNPAD) {
                int j;
                if (jb < nyh) j = jb;
                else if (jb > N - nyh) j = jb - (N - ny);
                else continue;
                if (DIR < 0) v = apply_op<T>(op, v, c0 + c, j, a);
                gst(dst + (long)j * a.ncols + c0 + c, v);
            } else {
                if (DIR < 0 && a.map == MAP_ID) {
                    if (a.zero_nyq && jb == nyh) v = cplx<T>{(T)0, (T)0};
                    v = apply_op<T>(op, v, c0 + c, jb, a);
                }
                if (DIR > 0 || in_yband(jb, ny, ykeep)) gst(dst + (long)jb * a.ncols + c0 + c, v);
            }
        }
    }
    if (a.map == MAP_UNPAD) {
        // the ky = ny/2 row of the small output is zero (fft.f90:84-86)
        for (int c = threadIdx.x; c < tc; c += blockDim.x)
            dst[(long)nyh * a.ncols + c0 + c] = cplx<T>{(T)0, (T)0};
    }
}

// Splits used for the transform lengths on the step: N -> (N1, N2).
// Returns false when N has no specialised kernel (generic passes are used).
inline bool split_of(int N, int& n1, int& n2) {
    switch (N) {
        case 16: n1 = 4; n2 = 4; return true;
        case 24: n1 = 4; n2 = 6; return true;
        case 32: n1 = 4; n2 = 8; return true;
        case 48: n1 = 6; n2 = 8; return true;
        case 64: n1 = 8; n2 = 8; return true;
        case 72: n1 = 8; n2 = 9; return true;
        case 96: n1 = 8; n2 = 12; return true;
        case 128: n1 = 8; n2 = 16; return true;
        case 144: n1 = 12; n2 = 12; return true;
        case 192: n1 = 12; n2 = 16; return true;
        case 256: n1 = 16; n2 = 16; return true;
        case 384: n1 = 16; n2 = 24; return true;
        default: return false;
    }
}

}  // namespace lg
