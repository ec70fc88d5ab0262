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
// * the plan's twiddles are read from a copy in LDS (TWL) or from the plan's
//   table in global memory, which leaves the LDS to more resident blocks.
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

// Twiddle W_N^e of the plan: from the block's LDS copy (TWL) or the plan's
// table in global memory (cached; leaves the LDS to more resident blocks).
template <typename T, bool TWL>
__device__ __forceinline__ cplx<T> tw_at(const cplx<T>* ltw, const typename TwT<T>::type* gtw,
                                         int e) {
    if constexpr (TWL) {
        return ltw[e];
    } else {
        const auto w = gtw[e];
        return cplx<T>{(T)w.x, (T)w.y};
    }
}

// Copies the plan's N twiddles into LDS (TWL) and waits for the block.
template <typename T, bool TWL>
__device__ __forceinline__ void tw_stage(cplx<T>* ltw, const typename TwT<T>::type* gtw, int N) {
    if constexpr (TWL) {
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            const auto w = gtw[i];
            ltw[i] = cplx<T>{(T)w.x, (T)w.y};
        }
        __syncthreads();
    }
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
    // every load is issued before any is used (the band and the Nyquist
    // choice are runtime values: a use inside its branch would wait for each
    // load in turn), then the multiplier
    int jr[N1];
    bool ld[N1], nq[N1];
    cplx<T> raw[N1];
    T gv[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int jb = n2 + N2 * n1;
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
        jr[n1] = j;
        ld[n1] = load;
        nq[n1] = nyq;
        raw[n1] = cplx<T>{(T)0, (T)0};
        gv[n1] = (T)0;
        if (load) {
            const unsigned off = ((unsigned)j * nc + (unsigned)c) * (unsigned)sizeof(cplx<T>);
            raw[n1] = at_bytes(sb, off);
            if constexpr (K == K_G || K == K_GB) gv[n1] = gt[(unsigned)j * nc + (unsigned)c];
        }
    }
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        cplx<T> v{(T)0, (T)0};
        if (ld[n1]) {
            v = raw[n1];
            const int j = jr[n1];
            const bool nyq = nq[n1];
            if constexpr (K == K_IKX) {
                const T k = nyq ? (T)0 : kxv;
                v = cplx<T>{-v.y * k, v.x * k};
            } else if constexpr (K == K_IKY) {
                const T k = nyq ? (T)0 : a.kys * (T)ky_index(j, a.ny);
                v = cplx<T>{-v.y * k, v.x * k};
            } else if constexpr (K == K_G) {
                const T g = gv[n1];
                v = cplx<T>{g * v.x, g * v.y};
            } else if constexpr (K == K_GB) {
                const bool in = gv[n1] != (T)0 && c != a.ncols - 1;
                if (!in) v = cplx<T>{(T)0, (T)0};
            }
            (void)j;
            (void)nyq;
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
    bool ld[N1];
    cplx<T> ra[N1], rb[N1];
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
        ld[n1] = load;
        ra[n1] = rb[n1] = cplx<T>{(T)0, (T)0};
        if (load) {
            // (issued before any is used: see cols_inv_load)
            const unsigned off = ((unsigned)jb * nc + (unsigned)c) * (unsigned)sizeof(cplx<T>);
            ra[n1] = at_bytes(sb, off);
            rb[n1] = at_bytes(sb2, off);
        }
    }
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int jb = n2 + N2 * n1;
        x[n1] = ld[n1] ? divxy<T>(ra[n1], rb[n1], c, jb, a) : cplx<T>{(T)0, (T)0};
    }
}

template <typename T, int N1, int N2, int MAP, bool TWL>
__global__ void __launch_bounds__(256) k_cols_inv_v4(ColArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    extern __shared__ __align__(16) unsigned char smem_c4[];
    cplx<T>* ltw = reinterpret_cast<cplx<T>*>(smem_c4);
    cplx<T>* lds = TWL ? ltw + N : ltw;
    const auto* gtw = plan_tw<T>(a.plan);
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
    tw_stage<T, TWL>(ltw, gtw, N);
    const int CS = SP::YS;
    for (int t = threadIdx.x; t < TC * N2; t += blockDim.x) {
        const int cl = t % TC, n2 = t / TC;
        if (cl >= tc) continue;
        __builtin_assume(n2 >= 0 && n2 < N2);
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
                const cplx<T> w = tw_at<T, TWL>(ltw, gtw, n2 * k1);
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

// ---------------------------------------------------------------------------
// Inverse row pass (paired rows) specialised by band: the same four-step
// transform and arithmetic as k_rows_c2r_v2, with the Hermitian extension's
// cases (DC, below / above the Nyquist column, outside the stored band)
// resolved per unrolled element at compile time wherever the element's
// columns lie on one side, 32-bit offsets from a block-uniform base and the
// twiddles in LDS.  BAND: 0 every stored column (ncols = N/2), 3 the 3/2
// grid's band (ncols = N/3), 1 a per-job band nce < ncols.
// ---------------------------------------------------------------------------
enum : int { BAND_FULL = 0, BAND_JOB = 1, BAND_THIRD = 3 };

template <typename T, int N1, int N2, int BAND, bool TWL>
__device__ __forceinline__ void rows_c2r_v4_body(const RowC2RArgs<T>& a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    constexpr int half = N / 2;
    extern __shared__ __align__(16) unsigned char smem_r4[];
    cplx<T>* ltw = reinterpret_cast<cplx<T>*>(smem_r4);
    cplx<T>* lds = TWL ? ltw + N : ltw;
    const auto* gtw = plan_tw<T>(a.plan);
    const int job = blockIdx.z;
    const int mem = blockIdx.y / a.nlev;
    const int plane = blockIdx.y - mem * a.nlev;
    const int npairs = a.nrows / 2;
    const int pair0 = blockIdx.x * a.S;
    const int S = min(a.S, npairs - pair0);
    const unsigned nc = (unsigned)a.ncols;
    const int ld = a.dst_ld;
    // bases at the block's first row
    const cplx<T>* sb = a.src[job] + mem * a.src_mstride +
                        (long)(a.lev0_src + plane) * a.src_plane + (long)(2 * pair0) * nc;
    const long dofs = mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane +
                      (long)(2 * pair0) * ld;
    T* const ob1 = a.dst[job] + dofs;
    T* const ob2 = a.dst2[job] ? a.dst2[job] + dofs : nullptr;
    const int nce = BAND == BAND_JOB ? (a.nce[job] ? a.nce[job] : (int)nc)
                                     : (BAND == BAND_THIRD ? N / 3 : half);
    tw_stage<T, TWL>(ltw, gtw, N);
    for (int pass = 0; pass < (ob2 ? 2 : 1); ++pass) {
        T* const out = pass ? ob2 : ob1;
        for (int t = threadIdx.x; t < S * N2; t += blockDim.x) {
            const int s = t / N2, n2 = t - s * N2;
            __builtin_assume(n2 >= 0 && n2 < N2);
            const unsigned ra = 2u * (unsigned)s * nc;          // row a, in elements
            auto ld2 = [&](int col, cplx<T>& A, cplx<T>& B) {
                const unsigned o = (ra + (unsigned)col) * (unsigned)sizeof(cplx<T>);
                A = at_bytes(sb, o);
                B = at_bytes(sb, o + nc * (unsigned)sizeof(cplx<T>));
                if (pass) {
                    const T kx = a.kxs * (T)col;
                    A = cplx<T>{-A.y * kx, A.x * kx};
                    B = cplx<T>{-B.y * kx, B.x * kx};
                }
            };
            cplx<T> x[N1];
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) {
                const int k = n2 + N2 * n1;
                // (lo, hi: the element's column range over the tasks; constant
                // once unrolled)
                const int lo = N2 * n1, hi = N2 * n1 + N2 - 1;
                cplx<T> z{(T)0, (T)0};
                // column kk of the stored band, kk in [klo, khi] over the
                // tasks: the band test is decided at compile time unless
                // the range straddles the band's edge
                auto in_band = [&](int kk, int klo, int khi) {
                    if (BAND == BAND_FULL) return true;
                    if (BAND == BAND_THIRD && khi < N / 3) return true;
                    if (BAND == BAND_THIRD && klo >= N / 3) return false;
                    return kk < nce;
                };
                auto below = [&](int kk, int klo, int khi) {        // 0 < kk < half
                    if (in_band(kk, klo, khi)) {
                        cplx<T> A, B;
                        ld2(kk, A, B);
                        z = cplx<T>{A.x - B.y, A.y + B.x};
                    }
                };
                auto above = [&](int kk, int klo, int khi) {        // kk = N - k
                    if (in_band(kk, klo, khi)) {
                        cplx<T> A, B;
                        ld2(kk, A, B);
                        z = cplx<T>{A.x + B.y, B.x - A.y};
                    }
                };
                // FP32 over the 3/2 grid's third: the elements whose case
                // differs between lanes (k = 0 and the third's edges) load
                // unconditionally from a valid column and select the result,
                // so that no load and its use sit in a per-lane branch, which
                // would hold the loads after it until it returns (measured
                // faster there; the FP64 instance and the whole half spectrum
                // ran slower this way and keep the branches).  n1 = 0 lies
                // inside the band: N2 - 1 < N / 3.
                constexpr bool SEL = BAND == BAND_THIRD && sizeof(T) == 4 && N2 - 1 < N / 3;
                if (hi < half) {
                    if (lo == 0 && SEL) {
                        cplx<T> A, B;
                        ld2(k, A, B);
                        const cplx<T> zb{A.x - B.y, A.y + B.x};
                        const cplx<T> z0 = pass ? cplx<T>{(T)0, (T)0} : cplx<T>{A.x, B.x};
                        z = k == 0 ? z0 : zb;
                    } else if (lo == 0 && k == 0) {
                        if (!pass) {
                            const cplx<T> A = at_bytes(sb, ra * (unsigned)sizeof(cplx<T>));
                            const cplx<T> B = at_bytes(sb, (ra + nc) * (unsigned)sizeof(cplx<T>));
                            z = cplx<T>{A.x, B.x};
                        }
                    } else if (SEL && lo < N / 3 && hi >= N / 3) {
                        // the third's edge from below: column k, kept below nce
                        cplx<T> A, B;
                        ld2(k < nce ? k : nce - 1, A, B);
                        const cplx<T> zb{A.x - B.y, A.y + B.x};
                        z = k < nce ? zb : cplx<T>{(T)0, (T)0};
                    } else {
                        below(k, lo, hi);
                    }
                } else if (lo > half) {
                    if (SEL && N - hi < N / 3 && N - lo >= N / 3) {
                        // the third's edge from above: column N - k, kept below nce
                        const int kk = N - k;
                        cplx<T> A, B;
                        ld2(kk < nce ? kk : nce - 1, A, B);
                        const cplx<T> za{A.x + B.y, B.x - A.y};
                        z = kk < nce ? za : cplx<T>{(T)0, (T)0};
                    } else {
                        above(N - k, N - hi, N - lo);
                    }
                } else {
                    if (k < half) below(k, lo, half - 1);
                    else if (k > half) above(N - k, N - hi, half - 1);
                }
                x[n1] = z;
            }
            dft_reg<T, N1, +1>(x);
            cplx<T>* Y = lds + s * SP::YS;
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) {
                cplx<T> v = x[k1];
                if (k1 != 0 && n2 != 0) {
                    const cplx<T> w = tw_at<T, TWL>(ltw, gtw, n2 * k1);
                    v = cmul(v, cplx<T>{w.x, (T)(-1) * w.y});
                }
                Y[k1 * SP::YR + n2] = v;
            }
        }
        __syncthreads();
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
                const cplx<T>* X = lds + s2 * SP::YS + 2 * q;
                const cplx<T> p0 = X[0], p1 = X[1];
                const unsigned o = 2u * (unsigned)s2 * (unsigned)ld + 2u * (unsigned)q;
                st_pair2(out + o, p0.x, p1.x);
                st_pair2(out + o + ld, p0.y, p1.y);
            }
        } else {
            for (int t = threadIdx.x; t < S * N1; t += blockDim.x) {
                const int s = t / N1, k1 = t - s * N1;
                const cplx<T>* Y = lds + s * SP::YS + k1 * SP::YR;
                cplx<T> y[N2];
#pragma unroll
                for (int n2 = 0; n2 < N2; ++n2) y[n2] = Y[n2];
                dft_reg<T, N2, +1>(y);
                const unsigned o = 2u * (unsigned)s * (unsigned)ld + (unsigned)k1;
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) {
                    out[o + N1 * k2] = y[k2].x;
                    out[o + ld + N1 * k2] = y[k2].y;
                }
            }
        }
        // padding columns N..ld-1
        const int npad = ld - N;
        for (int idx = threadIdx.x; idx < S * npad; idx += blockDim.x) {
            const int s = idx / npad, q = idx - s * npad;
            const unsigned o = 2u * (unsigned)s * (unsigned)ld + (unsigned)(N + q);
            out[o] = (T)0;
            out[o + ld] = (T)0;
        }
        __syncthreads();
    }
}

template <typename T, int N1, int N2, int BAND, bool TWL>
__global__ void __launch_bounds__(256) k_rows_c2r_v4(RowC2RArgs<T> a) {
    rows_c2r_v4_body<T, N1, N2, BAND, TWL>(a);
}


// ---------------------------------------------------------------------------
// Forward column pass (no multiplier, the form every forward pass of the
// step uses) specialised by index map: MAP_ID (N = ny rows, the ky = ny/2
// row zeroed when zero_nyq, only rows |m| < ykeep written when ykeep > 0)
// and MAP_UNPAD (N = 3 ny / 2 rows in, the small grid's rows out, unpadd of
// fft.f90:69-99, its ky = ny/2 row zero).  The same arithmetic as
// k_cols_v2<T, N1, N2, -1>.
// ---------------------------------------------------------------------------
template <typename T, int N1, int N2, int MAP, bool TWL>
__global__ void __launch_bounds__(256) k_cols_fwd_v4(ColArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    extern __shared__ __align__(16) unsigned char smem_f4[];
    cplx<T>* ltw = reinterpret_cast<cplx<T>*>(smem_f4);
    cplx<T>* lds = TWL ? ltw + N : ltw;
    const auto* gtw = plan_tw<T>(a.plan);
    const int job = blockIdx.x;
    const int TC = a.TC;
    const int c0 = blockIdx.y * TC;
    const int nzc = a.nzc[job] ? a.nzc[job] : a.ncols;
    if (c0 >= nzc) return;
    const int tc = min(TC, nzc - c0);
    const int mem = blockIdx.z / a.nlev;
    const int plane = blockIdx.z - mem * a.nlev;
    const cplx<T>* sb = a.src[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane;
    cplx<T>* db = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane;
    const int ykeep = MAP == MAP_ID ? a.ykeep[job] : 0;
    const bool znyq = a.zero_nyq != 0;
    const unsigned nc = (unsigned)a.ncols;
    constexpr unsigned CB = (unsigned)sizeof(cplx<T>);
    tw_stage<T, TWL>(ltw, gtw, N);
    const int CS = SP::YS;
    for (int t = threadIdx.x; t < TC * N2; t += blockDim.x) {
        const int cl = t % TC, n2 = t / TC;
        if (cl >= tc) continue;
        const unsigned cb = (unsigned)(c0 + cl) * CB;
        cplx<T> x[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1)
            x[n1] = at_bytes(sb, (unsigned)(n2 + N2 * n1) * nc * CB + cb);
        dft_reg<T, N1, -1>(x);
        cplx<T>* Y = lds + cl * CS;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {
            cplx<T> v = x[k1];
            if (k1 != 0 && n2 != 0) v = cmul(v, tw_at<T, TWL>(ltw, gtw, n2 * k1));
            Y[k1 * SP::YR + n2] = v;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < TC * N1; t += blockDim.x) {
        const int cl = t % TC, k1 = t / TC;
        if (cl >= tc) continue;
        __builtin_assume(k1 >= 0 && k1 < N1);
        const unsigned cb = (unsigned)(c0 + cl) * CB;
        const cplx<T>* Y = lds + cl * CS + k1 * SP::YR;
        cplx<T> y[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) y[n2] = Y[n2];
        dft_reg<T, N2, -1>(y);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            const int jb = k1 + N1 * k2;
            // (row range of this element over the tasks: constant once unrolled)
            const int lo = N1 * k2, hi = N1 * k2 + N1 - 1;
            cplx<T> v = y[k2];
            if constexpr (MAP == MAP_ID) {
                constexpr int NH = N / 2;
                const bool can_nyq = lo <= NH && NH <= hi;
                if (can_nyq && znyq && jb == NH) v = cplx<T>{(T)0, (T)0};
                bool keep = true;
                if (ykeep) {
                    const int m = hi < NH ? jb : (lo > NH ? N - jb : (jb < NH ? jb : N - jb));
                    keep = !(can_nyq && jb == NH) && m < ykeep;
                }
                if (keep) st_bytes(db, (unsigned)jb * nc * CB + cb, v);
            } else {
                // MAP_UNPAD: rows [0, N/3) -> j = jb, (2N/3, N) -> jb - N/3
                constexpr int LO = N / 3, HI = 2 * N / 3;
                int j = -1;
                if (hi < LO) j = jb;
                else if (lo > HI) j = jb - LO;
                else if (!(lo >= LO && hi <= HI)) j = jb < LO ? jb : (jb > HI ? jb - LO : -1);
                if (j >= 0) st_bytes(db, (unsigned)j * nc * CB + cb, v);
            }
        }
    }
    if constexpr (MAP == MAP_UNPAD) {
        // the ky = ny/2 row of the small output is zero (fft.f90:84-86)
        constexpr int NYH = N / 3;
        for (int c = threadIdx.x; c < tc; c += blockDim.x)
            st_bytes(db, (unsigned)NYH * nc * CB + (unsigned)(c0 + c) * CB, cplx<T>{(T)0, (T)0});
    }
}

// ---------------------------------------------------------------------------
// Forward row pass (paired rows): the same four-step transform and the same
// arithmetic as k_rows_r2c_v2, with 32-bit offsets from block-uniform bases
// and the twiddles from LDS (TWL) or global memory.
// ---------------------------------------------------------------------------
template <typename T, int N1, int N2, bool TWL>
__global__ void __launch_bounds__(256) k_rows_r2c_v4(RowR2CArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    extern __shared__ __align__(16) unsigned char smem_rf4[];
    cplx<T>* ltw = reinterpret_cast<cplx<T>*>(smem_rf4);
    cplx<T>* lds = TWL ? ltw + N : ltw;
    const auto* gtw = plan_tw<T>(a.plan);
    const int job = blockIdx.z;
    const int mem = blockIdx.y / a.nlev;
    const int plane = blockIdx.y - mem * a.nlev;
    const int npairs = a.nrows / 2;
    const int pair0 = blockIdx.x * a.S;
    const int S = min(a.S, npairs - pair0);
    const unsigned sld = (unsigned)a.src_ld;
    const unsigned nc = (unsigned)a.ncols;
    const T* sb = a.src[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane +
                  (long)(2 * pair0) * sld;
    cplx<T>* db = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane +
                  (long)(2 * pair0) * nc;
    const T sc = a.mscale ? (T)a.mscale[(long)mem * a.mscale_stride] : a.scale[job];
    tw_stage<T, TWL>(ltw, gtw, N);
    for (int t = threadIdx.x; t < S * N2; t += blockDim.x) {
        const int s = t / N2, n2 = t - s * N2;
        const unsigned ra = 2u * (unsigned)s * sld + (unsigned)n2;
        cplx<T> x[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1)
            x[n1] = cplx<T>{sc * sb[ra + N2 * n1], sc * sb[ra + sld + N2 * n1]};
        dft_reg<T, N1, -1>(x);
        cplx<T>* Y = lds + s * SP::YS;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {
            cplx<T> v = x[k1];
            if (k1 != 0 && n2 != 0) v = cmul(v, tw_at<T, TWL>(ltw, gtw, n2 * k1));
            Y[k1 * SP::YR + n2] = v;
        }
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
    const int nw = a.nwr ? a.nwr : (int)nc;
    for (int idx = threadIdx.x; idx < S * nw; idx += blockDim.x) {
        const int s = idx / nw;
        const int k = idx - s * nw;
        const cplx<T>* X = lds + s * SP::YS;
        const cplx<T> z = X[k];
        const cplx<T> zc = X[(N - k) % N];
        const cplx<T> A{(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y - zc.y)};
        const cplx<T> Bm{(T)0.5 * (z.x - zc.x), (T)0.5 * (z.y + zc.y)};
        const unsigned o = (2u * (unsigned)s * nc + (unsigned)k) * (unsigned)sizeof(cplx<T>);
        st_bytes(db, o, A);
        st_bytes(db, o + nc * (unsigned)sizeof(cplx<T>), cplx<T>{Bm.y, -Bm.x});
    }
}

}  // namespace lg
