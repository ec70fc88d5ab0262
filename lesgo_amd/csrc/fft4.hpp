// Inverse column pass specialised by index map and spectral multiplier.
//
// The same four-step transform as k_cols_v2<T, N1, N2, +1> (and the same
// arithmetic, so the results are the same bits), arranged so that little
// besides the loads, the butterflies and the stores is executed per element:
//
// * the index map (MAP_ID: N = ny rows; MAP_PAD: N = 3 ny / 2 with the
//   small grid's rows at the ends, padd of fft.f90:43-66) is a template
//   parameter, and each unrolled element's row class (below / above the
//   Nyquist row, or inside the padded gap) is known at compile time except
//   for the one element per task that straddles it;
// * the multiplier (none, i kx, i ky, G_test, G_test_test, G_test's band,
//   i kx A + i ky B) is uniform over a block and chosen once per task, not
//   per element;
// * addresses are 32-bit byte offsets from a block-uniform base;
// * the plan's twiddles are read from a copy in LDS.
#pragma once

#include "fft2.hpp"

namespace lg {

template <typename T>
__device__ __forceinline__ const cplx<T>& at_bytes(const cplx<T>* base, unsigned off) {
    return *reinterpret_cast<const cplx<T>*>(reinterpret_cast<const char*>(base) + off);
}

template <typename T>
__device__ __forceinline__ void st_bytes(cplx<T>* base, unsigned off, cplx<T> v) {
    *reinterpret_cast<cplx<T>*>(reinterpret_cast<char*>(base) + off) = v;
}

// ky index of small-grid row j (j != ny/2)
__device__ __forceinline__ int ky_index(int j, int ny) { return j < ny / 2 ? j : j - ny; }

// Multiplier kinds as dispatched per block (OP_DIVXY has its own loader).
enum : int { K_NONE = 0, K_IKX = 1, K_IKY = 2, K_G = 3, K_GB = 4 };

// Element n1 of a task (rows jb = n2 + N2 n1, n2 < N2) and where the row
// falls: 0 wholly below the split row, 2 wholly above, 1 straddling it.
template <int N2>
__host__ __device__ constexpr int row_class(int n1, int split) {
    return (N2 * n1 + N2 - 1 < split) ? 0 : ((N2 * n1 > split) ? 2 : 1);
}

// x[n1] = op(src(row of jb, column c)) for one task of the inverse pass.
// MAP_ID: small-grid row jb, zero outside the band |m| < ykeep (ykeep > 0)
// and on the Nyquist row when nyq0; MAP_PAD: small-grid row jb for jb <
// ny/2, jb - ny/2 for jb > ny, zero between.
template <typename T, int N1, int N2, int MAP, int K>
__device__ __forceinline__ void cols_inv_load(cplx<T> (&x)[N1], const cplx<T>* sb,
                                              const ColArgs<T>& a, const T* gt, int c, int n2,
                                              int ykeep, bool nyq0) {
    constexpr int N = N1 * N2;
    const unsigned nc = (unsigned)a.ncols;
    const T kxv = a.kxs * (T)c;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int jb = n2 + N2 * n1;
        cplx<T> v{(T)0, (T)0};
        int j = jb;          // small-grid row
        bool load = true;
        bool nyq = false;    // the small grid's Nyquist row (MAP_ID only)
        if constexpr (MAP == MAP_ID) {
            constexpr int NH = N / 2;
            const int cl = row_class<N2>(n1, NH);   // constant once unrolled
            if (cl == 1) nyq = jb == NH;
            if (ykeep) {
                const int m = (cl == 0) ? jb : ((cl == 2) ? N - jb : (jb < NH ? jb : N - jb));
                load = !nyq && m < ykeep;
            }
            if (nyq && nyq0) load = false;
        } else {
            // MAP_PAD: N = 3 ny / 2; rows [0, N/3) and (2N/3, N) carry data
            constexpr int LO = N / 3, HI = 2 * N / 3;
            // (constant once unrolled)
            const bool low = N2 * n1 + N2 - 1 < LO;
            const bool high = N2 * n1 > HI;
            const bool gap = N2 * n1 >= LO && N2 * n1 + N2 - 1 <= HI;
            if (gap) {
                load = false;
            } else if (high) {
                j = jb - LO;
            } else if (!low) {
                if (jb < LO) j = jb;
                else if (jb > HI) j = jb - LO;
                else load = false;
            }
        }
        if (load) {
            const unsigned off = ((unsigned)j * nc + (unsigned)c) * (unsigned)sizeof(cplx<T>);
            v = at_bytes(sb, off);
            if constexpr (K == K_IKX) {
                const T k = nyq ? (T)0 : kxv;
                v = cplx<T>{-v.y * k, v.x * k};
            } else if constexpr (K == K_IKY) {
                const T k = nyq ? (T)0 : a.kys * (T)ky_index(j, a.ny);
                v = cplx<T>{-v.y * k, v.x * k};
            } else if constexpr (K == K_G) {
                const T g = gt[(unsigned)j * nc + (unsigned)c];
                v = cplx<T>{g * v.x, g * v.y};
            } else if constexpr (K == K_GB) {
                const bool in = gt[(unsigned)j * nc + (unsigned)c] != (T)0 && c != a.ncols - 1;
                if (!in) v = cplx<T>{(T)0, (T)0};
            }
        }
        x[n1] = v;
    }
}

// OP_DIVXY (MAP_ID): i kx A + i ky B, the Nyquist row zero.
template <typename T, int N1, int N2>
__device__ __forceinline__ void cols_inv_load_divxy(cplx<T> (&x)[N1], const cplx<T>* sb,
                                                    const cplx<T>* sb2, const ColArgs<T>& a,
                                                    int c, int n2, int ykeep, bool nyq0) {
    constexpr int N = N1 * N2;
    constexpr int NH = N / 2;
    const unsigned nc = (unsigned)a.ncols;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int jb = n2 + N2 * n1;
        const int cl = row_class<N2>(n1, NH);
        const bool nyq = (cl == 1) && jb == NH;
        bool load = true;
        if (ykeep) {
            const int m = (cl == 0) ? jb : ((cl == 2) ? N - jb : (jb < NH ? jb : N - jb));
            load = !nyq && m < ykeep;
        }
        if (nyq && nyq0) load = false;
        cplx<T> v{(T)0, (T)0};
        if (load) {
            const unsigned off = ((unsigned)jb * nc + (unsigned)c) * (unsigned)sizeof(cplx<T>);
            v = divxy<T>(at_bytes(sb, off), at_bytes(sb2, off), c, jb, a);
        }
        x[n1] = v;
    }
}

template <typename T, int N1, int N2, int MAP>
__global__ void __launch_bounds__(256) k_cols_inv_v4(ColArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    extern __shared__ __align__(16) unsigned char smem_c4[];
    cplx<T>* ltw = reinterpret_cast<cplx<T>*>(smem_c4);
    cplx<T>* lds = ltw + N;
    const int job = blockIdx.x;
    const int TC = a.TC;
    const int c0 = blockIdx.y * TC;
    const int nzc = a.nzc[job] ? a.nzc[job] : a.ncols;
    if (c0 >= nzc) return;
    const int tc = min(TC, nzc - c0);
    const int mem = blockIdx.z / a.nlev;
    const int plane = blockIdx.z - mem * a.nlev;
    const long so = mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane;
    const cplx<T>* sb = a.src[job] + so;
    cplx<T>* db = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane;
    const int op = a.op[job];
    const int ykeep = MAP == MAP_ID ? a.ykeep[job] : 0;
    const bool nyq0 = MAP == MAP_ID && a.nyq0[job] != 0;
    {
        const auto* tw = plan_tw<T>(a.plan);
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            const auto w = tw[i];
            ltw[i] = cplx<T>{(T)w.x, (T)w.y};
        }
    }
    __syncthreads();
    const int CS = SP::YS;
    for (int t = threadIdx.x; t < TC * N2; t += blockDim.x) {
        const int cl = t % TC, n2 = t / TC;
        if (cl >= tc) continue;
        const int c = c0 + cl;
        cplx<T> x[N1];
        switch (op) {
            case OP_IKX: cols_inv_load<T, N1, N2, MAP, K_IKX>(x, sb, a, nullptr, c, n2, ykeep, nyq0); break;
            case OP_IKY: cols_inv_load<T, N1, N2, MAP, K_IKY>(x, sb, a, nullptr, c, n2, ykeep, nyq0); break;
            case OP_G1: cols_inv_load<T, N1, N2, MAP, K_G>(x, sb, a, a.G1, c, n2, ykeep, nyq0); break;
            case OP_G2: cols_inv_load<T, N1, N2, MAP, K_G>(x, sb, a, a.G2, c, n2, ykeep, nyq0); break;
            case OP_G1B: cols_inv_load<T, N1, N2, MAP, K_GB>(x, sb, a, a.G1, c, n2, ykeep, nyq0); break;
            case OP_DIVXY:
                if constexpr (MAP == MAP_ID) {
                    cols_inv_load_divxy<T, N1, N2>(x, sb, a.src2[job] + so, a, c, n2, ykeep, nyq0);
                    break;
                }
                [[fallthrough]];
            default: cols_inv_load<T, N1, N2, MAP, K_NONE>(x, sb, a, nullptr, c, n2, ykeep, nyq0); break;
        }
        dft_reg<T, N1, +1>(x);
        cplx<T>* Y = lds + cl * CS;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {
            cplx<T> v = x[k1];
            if (k1 != 0 && n2 != 0) {
                const cplx<T> w = ltw[n2 * k1];
                v = cmul(v, cplx<T>{w.x, (T)(-1) * w.y});
            }
            Y[k1 * SP::YR + n2] = v;
        }
    }
    __syncthreads();
    const unsigned nc = (unsigned)a.ncols;
    for (int t = threadIdx.x; t < TC * N1; t += blockDim.x) {
        const int cl = t % TC, k1 = t / TC;
        if (cl >= tc) continue;
        const cplx<T>* Y = lds + cl * CS + k1 * SP::YR;
        cplx<T> y[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) y[n2] = Y[n2];
        dft_reg<T, N2, +1>(y);
        const unsigned cb = (unsigned)(c0 + cl) * (unsigned)sizeof(cplx<T>);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            const unsigned jb = (unsigned)(k1 + N1 * k2);
            st_bytes(db, jb * nc * (unsigned)sizeof(cplx<T>) + cb, y[k2]);
        }
    }
}

}  // namespace lg
