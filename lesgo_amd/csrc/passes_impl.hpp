// Definitions of the transform-pass launchers (included by one translation
// unit per precision, which instantiates them).
#pragma once

#include <algorithm>
#include <vector>

#include "fft3.hpp"
#include "fft4.hpp"
#include "lag_rows.hpp"
#include "passes.hpp"
#include "prod_rows.hpp"

namespace lg {

static constexpr int PBLK = 256;

// Block size of a four-step pass with S sequences (rows pairs or columns)
// per block: its larger phase's task count S max(N1, N2) in whole waves, at
// most PBLK, so that neither phase leaves most of the block's waves idle
// (S = 8 at N = 16 x 24: 192 and 128 tasks, 192 threads).
inline unsigned four_step_block(int S, int n1, int n2) {
    const int t = ((S * std::max(n1, n2) + 63) / 64) * 64;
    return (unsigned)std::min(std::max(t, 64), PBLK);
}

// ---------------------------------------------------------------------------
// Four-step kernel instances: every (N1, N2) split a transform length may
// use.  split_of() gives the default; the tuner may pick another.
// ---------------------------------------------------------------------------
#define LG_SPLITS(X)                                                                      \
    X(4, 4) X(4, 6) X(6, 4) X(4, 8) X(8, 4) X(6, 8) X(8, 6) X(4, 12) X(8, 8) X(4, 16)    \
    X(16, 4) X(8, 9) X(9, 8) X(8, 12) X(12, 8) X(6, 16) X(8, 16) X(16, 8) X(12, 12)      \
    X(9, 16) X(16, 9) X(12, 16) X(16, 12) X(8, 24) X(16, 16) X(16, 24) X(24, 16)

static std::vector<std::pair<int, int>> splits_for(int N) {
    std::vector<std::pair<int, int>> v;
#define LG_CASE(A, B) \
    if (A * B == N) v.push_back({A, B});
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return v;
}

template <typename T>
static bool v2_r2c(int n1, int n2, dim3 g, size_t lds, hipStream_t s, const RowR2CArgs<T>& a) {
#define LG_CASE(A, B)                                                                    \
    if (n1 == A && n2 == B) {                                                            \
        LG_LAUNCH((k_rows_r2c_v2<T, A, B>), g, dim3(four_step_block(a.S, A, B)), lds, s, a);           \
        return true;                                                                     \
    }
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return false;
}

template <typename T>
static bool v2_c2r(int n1, int n2, dim3 g, size_t lds, hipStream_t s, const RowC2RArgs<T>& a) {
#define LG_CASE(A, B)                                                                    \
    if (n1 == A && n2 == B) {                                                            \
        LG_LAUNCH((k_rows_c2r_v2<T, A, B>), g, dim3(four_step_block(a.S, A, B)), lds, s, a);           \
        return true;                                                                     \
    }
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return false;
}

// The forward row pass with 32-bit offsets (fft4.hpp); twl: twiddles in LDS.
template <typename T>
static bool v4_r2c(int n1, int n2, bool twl, dim3 g, size_t lds, hipStream_t s,
                   const RowR2CArgs<T>& a) {
    const size_t l4 = lds + (twl ? (size_t)n1 * n2 * sizeof(cplx<T>) : 0);
    if (l4 > LDS_MAX) return false;
    const dim3 blk(four_step_block(a.S, n1, n2));
#define LG_CASE(A, B)                                                                    \
    if (n1 == A && n2 == B) {                                                            \
        if (twl) LG_LAUNCH((k_rows_r2c_v4<T, A, B, true>), g, blk, l4, s, a);            \
        else LG_LAUNCH((k_rows_r2c_v4<T, A, B, false>), g, blk, l4, s, a);               \
        return true;                                                                     \
    }
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return false;
}

template <typename T>
static bool v3_r2c(int n1, int n2, dim3 g, size_t lds, hipStream_t s, const RowR2CArgs<T>& a) {
#define LG_CASE(A, B)                                                                    \
    if (n1 == A && n2 == B) {                                                            \
        LG_LAUNCH((k_rows_r2c_v3<T, A, B>), g, dim3(PBLK), lds, s, a);           \
        return true;                                                                     \
    }
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return false;
}

template <typename T>
static bool v3_c2r(int n1, int n2, dim3 g, size_t lds, hipStream_t s, const RowC2RArgs<T>& a) {
#define LG_CASE(A, B)                                                                    \
    if (n1 == A && n2 == B) {                                                            \
        LG_LAUNCH((k_rows_c2r_v3<T, A, B>), g, dim3(PBLK), lds, s, a);           \
        return true;                                                                     \
    }
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return false;
}

// The inverse row pass specialised by band (fft4.hpp); band: BAND_FULL,
// BAND_THIRD or BAND_JOB; twl: twiddles staged in LDS.
template <typename T, int A, int B, bool TWL>
static void launch_c2r_v4(int band, dim3 g, size_t lds, hipStream_t s, const RowC2RArgs<T>& a) {
    const dim3 blk(four_step_block(a.S, A, B));
    if constexpr ((A * B) % 3 == 0) {
        if (band == BAND_THIRD) {
            LG_LAUNCH((k_rows_c2r_v4<T, A, B, BAND_THIRD, TWL>), g, blk, lds, s, a);
            return;
        }
    }
    if (band == BAND_FULL) LG_LAUNCH((k_rows_c2r_v4<T, A, B, BAND_FULL, TWL>), g, blk, lds, s, a);
    else LG_LAUNCH((k_rows_c2r_v4<T, A, B, BAND_JOB, TWL>), g, blk, lds, s, a);
}

template <typename T>
static bool v4_c2r(int n1, int n2, bool twl, dim3 g, size_t lds, hipStream_t s,
                   const RowC2RArgs<T>& a) {
    const int N = n1 * n2;
    const size_t l4 = lds + (twl ? (size_t)N * sizeof(cplx<T>) : 0);
    if (l4 > LDS_MAX) return false;
    bool per_job = false;
    for (int i = 0; i < (int)g.z; ++i) per_job |= a.nce[i] != 0;
    const int band = per_job ? BAND_JOB
                             : (2 * a.ncols == N ? BAND_FULL
                                                 : ((3 * a.ncols == N && N % 3 == 0) ? BAND_THIRD
                                                                                      : BAND_JOB));
#define LG_CASE(A, B)                                                                    \
    if (n1 == A && n2 == B) {                                                            \
        if (twl) launch_c2r_v4<T, A, B, true>(band, g, l4, s, a);                        \
        else launch_c2r_v4<T, A, B, false>(band, g, l4, s, a);                           \
        return true;                                                                     \
    }
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return false;
}

template <typename T, int DIR>
static bool v2_cols(int n1, int n2, dim3 g, size_t lds, hipStream_t s, const ColArgs<T>& a) {
#define LG_CASE(A, B)                                                                    \
    if (n1 == A && n2 == B) {                                                            \
        LG_LAUNCH((k_cols_v2<T, A, B, DIR>), g, dim3(four_step_block(a.TC, A, B)), lds, s, a);          \
        return true;                                                                     \
    }
    LG_SPLITS(LG_CASE)
#undef LG_CASE
    return false;
}

// The inverse column pass specialised by index map (fft4.hpp); false when
// the split or the map has no instance (the caller launches k_cols_v2).
template <typename T, int A, int B, int MAP>
static void launch_cols_v4(bool twl, dim3 g, size_t lds, hipStream_t s, const ColArgs<T>& a) {
    const dim3 blk(four_step_block(a.TC, A, B));
    if (twl) LG_LAUNCH((k_cols_inv_v4<T, A, B, MAP, true>), g, blk, lds, s, a);
    else LG_LAUNCH((k_cols_inv_v4<T, A, B, MAP, false>), g, blk, lds, s, a);
}

template <typename T>
static bool v4_cols_inv(int n1, int n2, bool twl, dim3 g, size_t lds, hipStream_t s,
                        const ColArgs<T>& a) {
    const int N = n1 * n2;
    const size_t l4 = lds + (twl ? (size_t)N * sizeof(cplx<T>) : 0);
    if (l4 > LDS_MAX) return false;
    if (a.map == MAP_ID && a.ny == N) {
#define LG_CASE(A, B)                                                                    \
        if (n1 == A && n2 == B) {                                                        \
            launch_cols_v4<T, A, B, MAP_ID>(twl, g, l4, s, a);                           \
            return true;                                                                 \
        }
        LG_SPLITS(LG_CASE)
#undef LG_CASE
    } else if (a.map == MAP_PAD && 2 * N == 3 * a.ny) {
#define LG_CASE(A, B)                                                                    \
        if (n1 == A && n2 == B) {                                                        \
            launch_cols_v4<T, A, B, MAP_PAD>(twl, g, l4, s, a);                          \
            return true;                                                                 \
        }
        LG_SPLITS(LG_CASE)
#undef LG_CASE
    }
    return false;
}

// The forward column pass specialised by index map (fft4.hpp), for jobs
// without a multiplier; false when the split, the map or an op has none.
template <typename T, int A, int B, int MAP>
static void launch_cols_fwd_v4(bool twl, dim3 g, size_t lds, hipStream_t s, const ColArgs<T>& a) {
    const dim3 blk(four_step_block(a.TC, A, B));
    if (twl) LG_LAUNCH((k_cols_fwd_v4<T, A, B, MAP, true>), g, blk, lds, s, a);
    else LG_LAUNCH((k_cols_fwd_v4<T, A, B, MAP, false>), g, blk, lds, s, a);
}

template <typename T>
static bool v4_cols_fwd(int n1, int n2, bool twl, dim3 g, size_t lds, hipStream_t s,
                        const ColArgs<T>& a) {
    const int N = n1 * n2;
    for (int i = 0; i < (int)g.x; ++i)
        if (a.op[i] != OP_NONE) return false;
    const size_t l4 = lds + (twl ? (size_t)N * sizeof(cplx<T>) : 0);
    if (l4 > LDS_MAX) return false;
    if (a.map == MAP_ID && a.ny == N) {
#define LG_CASE(A, B)                                                                    \
        if (n1 == A && n2 == B) {                                                        \
            launch_cols_fwd_v4<T, A, B, MAP_ID>(twl, g, l4, s, a);                       \
            return true;                                                                 \
        }
        LG_SPLITS(LG_CASE)
#undef LG_CASE
    } else if (a.map == MAP_UNPAD && 2 * N == 3 * a.ny) {
#define LG_CASE(A, B)                                                                    \
        if (n1 == A && n2 == B) {                                                        \
            launch_cols_fwd_v4<T, A, B, MAP_UNPAD>(twl, g, l4, s, a);                    \
            return true;                                                                 \
        }
        LG_SPLITS(LG_CASE)
#undef LG_CASE
    }
    return false;
}

// Is the choice launchable: LDS within bounds, and for the four-step r2c
// pass every row pair's first-step tasks in one round of the block
// (S * N1 <= block size).
template <typename T>
bool choice_ok(int kind, const PassChoice& c, int N, int other) {
    if (kind == PK_TRIDAG) return c.S == 64 || c.S == 128 || c.S == 256;
    if (c.S < 1) return false;
    if (c.n1) {
        if (c.n1 * c.n2 != N) return false;
        bool known = false;
        for (auto& p : splits_for(N)) known |= (p.first == c.n1 && p.second == c.n2);
        if (!known) return false;
    }
    if (c.v != 0 && !(c.v == 1 && c.n1 && (kind == PK_R2C || kind == PK_C2R)) &&
        !((c.v == 2 || c.v == 3) && c.n1 &&
          (kind == PK_COLS_INV || kind == PK_COLS_PAD || kind == PK_C2R || kind == PK_R2C ||
           kind == PK_COLS_FWD || kind == PK_COLS_UNPAD)))
        return false;
    // real rows of a length-N row pass have stride N + 2; ncols <= N / 2
    const size_t lds = row_lds_bytes<T>(c, N, N + 2, N / 2);
    if (lds > LDS_MAX) return false;
    if (kind == PK_R2C || kind == PK_C2R) {
        if (c.S > std::max(1, other / 2)) return false;
        if (kind == PK_R2C && c.n1 && c.S * c.n1 > PBLK) return false;
    } else {
        if (c.S > other) return false;
    }
    return true;
}

template <typename T>
PassChoice default_choice(int kind, int N, int other) {
    PassChoice c;
    if (kind == PK_TRIDAG) {
        c.S = 128;
        return c;
    }
    int n1, n2;
    if (split_of(N, n1, n2)) {
        c.n1 = n1;
        c.n2 = n2;
    }
    const int YS = seq_stride(c, N);
    if (kind == PK_R2C || kind == PK_C2R) {
        int S;
        if (c.n1) S = std::min(kind == PK_R2C ? std::min(PBLK / c.n1, 16) : 16, other / 2);
        else S = std::min(std::max(1, 2048 / N), other / 2);
        S = std::max(S, 1);
        while (S > 1 && (size_t)S * YS * sizeof(cplx<T>) > LDS_DEFAULT) --S;
        c.S = S;
    } else {
        int TC = 16;
        while (TC > 1 && (size_t)TC * YS * sizeof(cplx<T>) > LDS_DEFAULT) TC /= 2;
        c.S = std::min(TC, other);
    }
    return c;
}

template <typename T>
PassChoice choice_for(int kind, int N, int other) {
    PassChoice c;
    if (tuned_choice(kind, (int)sizeof(T), N, other, c) && choice_ok<T>(kind, c, N, other)) return c;
    return default_choice<T>(kind, N, other);
}

// ---------------------------------------------------------------------------
// Pass launchers.  grid.y (rows) / grid.z (columns) = nlev * M: member m
// covers planes lev0..lev0+nlev-1 at member offset m * mstride.
// ---------------------------------------------------------------------------
template <typename T>
void launch_rows_r2c(KernelTimer& tm, const FftPlan& pl, int njobs, const T* const* src,
                     const T* scales, long src_plane, int src_ld, int src_lev0, long src_ms,
                     cplx<T>* const* dst, long dst_plane, int dst_lev0, long dst_ms, int nrows,
                     int ncols, int nlev, int M, hipStream_t s, const PassChoice* force,
                     const double* mscale, int mscale_stride, int nwr) {
    const PassChoice ch = force ? *force : choice_for<T>(PK_R2C, pl.N, nrows);
    for (int j0 = 0; j0 < njobs; j0 += MAX_JOBS) {
        const int nj = std::min(MAX_JOBS, njobs - j0);
        RowR2CArgs<T> a{};
        a.plan = pl;
        for (int i = 0; i < nj; ++i) {
            a.src[i] = src[j0 + i];
            a.dst[i] = dst[j0 + i];
            a.scale[i] = scales ? scales[j0 + i] : (T)1;
        }
        a.src_plane = src_plane;
        a.dst_plane = dst_plane;
        a.src_ld = src_ld;
        a.nrows = nrows;
        a.ncols = ncols;
        a.lev0_src = src_lev0;
        a.lev0_dst = dst_lev0;
        a.nlev = nlev;
        a.src_mstride = src_ms;
        a.dst_mstride = dst_ms;
        a.mscale = mscale;
        a.mscale_stride = mscale_stride;
        a.nwr = (nwr > 0 && nwr < ncols) ? nwr : 0;
        a.S = ch.S;
        a.stride = seq_stride(ch, pl.N);
        const size_t lds = row_lds_bytes<T>(ch, pl.N, src_ld, ncols);
        dim3 grid((nrows / 2 + a.S - 1) / a.S, nlev * M, nj);
        const double bytes = (double)nj * nlev * M * nrows *
                             ((double)pl.N * sizeof(T) + (double)(a.nwr ? a.nwr : ncols) * sizeof(cplx<T>));
        tm.run(s, 3 * ncols == pl.N ? KK_ROWS_R2C_PAD : KK_ROWS_R2C, bytes, [&] {
            if (ch.n1 && ch.v == 1 && v3_r2c<T>(ch.n1, ch.n2, grid, lds, s, a)) return;
            if (ch.n1 && (ch.v == 0 || ch.v == 3) &&
                v4_r2c<T>(ch.n1, ch.n2, ch.v == 0, grid, lds, s, a))
                return;
            if (!(ch.n1 && v2_r2c<T>(ch.n1, ch.n2, grid, lds, s, a)))
                LG_LAUNCH(k_rows_r2c<T>, grid, dim3(PBLK), lds, s, a);
        });
    }
}

template <typename T>
void launch_rows_c2r(KernelTimer& tm, const FftPlan& pl, int njobs, const cplx<T>* const* src,
                     long src_plane, int src_lev0, long src_ms, T* const* dst, long dst_plane,
                     int dst_ld, int dst_lev0, long dst_ms, int nrows, int ncols, int nlev, int M,
                     hipStream_t s, const PassChoice* force, const int* nce, T* const* dst2,
                     T kxs) {
    PassChoice ch = force ? *force : choice_for<T>(PK_C2R, pl.N, nrows);
    bool dual = false;
    for (int i = 0; dst2 && i < njobs; ++i) dual |= dst2[i] != nullptr;
    // the LDS-staged kernel writes its output over its input: a job with a
    // second output takes the four-step kernel of the same split
    if (dual && ch.v == 1) ch.v = 0;
    for (int j0 = 0; j0 < njobs; j0 += MAX_JOBS) {
        const int nj = std::min(MAX_JOBS, njobs - j0);
        RowC2RArgs<T> a{};
        a.plan = pl;
        double cin = 0.0;      // complex columns read, summed over the jobs
        int nout = 0;          // real rows written per row of a job, summed
        for (int i = 0; i < nj; ++i) {
            a.src[i] = src[j0 + i];
            a.dst[i] = dst[j0 + i];
            a.dst2[i] = dst2 ? dst2[j0 + i] : nullptr;
            const int e = nce ? nce[j0 + i] : 0;
            a.nce[i] = (e > 0 && e < ncols) ? e : 0;
            cin += a.nce[i] ? a.nce[i] : ncols;
            nout += a.dst2[i] ? 2 : 1;
        }
        a.kxs = kxs;
        a.src_plane = src_plane;
        a.dst_plane = dst_plane;
        a.dst_ld = dst_ld;
        a.nrows = nrows;
        a.ncols = ncols;
        a.lev0_src = src_lev0;
        a.lev0_dst = dst_lev0;
        a.nlev = nlev;
        a.src_mstride = src_ms;
        a.dst_mstride = dst_ms;
        a.S = ch.S;
        a.stride = seq_stride(ch, pl.N);
        const size_t lds = row_lds_bytes<T>(ch, pl.N, dst_ld, ncols);
        dim3 grid((nrows / 2 + a.S - 1) / a.S, nlev * M, nj);
        const double bytes = (double)nlev * M * nrows *
                             (cin * sizeof(cplx<T>) + (double)nout * dst_ld * sizeof(T));
        tm.run(s, 3 * ncols == pl.N ? KK_ROWS_C2R_PAD : KK_ROWS_C2R, bytes, [&] {
            if (ch.n1 && ch.v == 1 && v3_c2r<T>(ch.n1, ch.n2, grid, lds, s, a)) return;
            if (ch.n1 && (ch.v == 0 || ch.v == 3) &&
                v4_c2r<T>(ch.n1, ch.n2, ch.v == 0, grid, lds, s, a))
                return;
            if (!(ch.n1 && v2_c2r<T>(ch.n1, ch.n2, grid, lds, s, a)))
                LG_LAUNCH(k_rows_c2r<T>, grid, dim3(PBLK), lds, s, a);
        });
    }
}

// Column pass of nj jobs (c.src/dst/op and c.plan filled): picks the tuned
// kernel, sets TC / stride / nlev and launches over nlev * M planes.
template <typename T, int DIR>
void launch_cols(KernelTimer& tm, ColArgs<T>& c, int nj, int nlev, int M, int ny_in,
                 hipStream_t s, const PassChoice* force) {
    const int N = c.plan.N;
    const PassChoice ch = force ? *force : choice_for<T>(cols_kind(DIR, c.map), N, c.ncols);
    c.TC = ch.S;
    c.stride = seq_stride(ch, N);
    c.nlev = nlev;
    const size_t lds = (size_t)c.TC * c.stride * sizeof(cplx<T>);
    // band-limited jobs: columns < nzc, and (MAP_ID) rows |ky| < ykeep read by
    // an inverse pass or written by a forward one
    int cmax = 0;
    double cols_rw = 0.0;      // sum over jobs of columns x rows moved
    for (int i = 0; i < nj; ++i) {
        const int nzc = (c.nzc[i] > 0 && c.nzc[i] < c.ncols) ? c.nzc[i] : c.ncols;
        c.nzc[i] = nzc < c.ncols ? nzc : 0;
        cmax = std::max(cmax, nzc);
        const int band = c.ykeep[i] > 0 ? std::min(2 * c.ykeep[i] - 1, ny_in) : 0;
        const double rin = (DIR > 0 && band) ? band : ny_in;
        // MAP_UNPAD writes the small grid's ny rows, not the N it transforms
        const double rout = c.map == MAP_UNPAD ? c.ny : ((DIR < 0 && band) ? band : N);
        cols_rw += (double)nzc * (rin + rout);
    }
    dim3 grid(nj, (cmax + c.TC - 1) / c.TC, nlev * M);
    const double bytes = (double)nlev * M * sizeof(cplx<T>) * cols_rw;
    const int kk = c.map == MAP_PAD ? KK_COLS_PAD
                   : (c.map == MAP_UNPAD ? KK_COLS_UNPAD : (DIR < 0 ? KK_COLS_FWD : KK_COLS_INV));
    tm.run(s, kk, bytes, [&] {
        if (DIR > 0 && ch.n1 && (ch.v == 0 || ch.v == 3) &&
            v4_cols_inv<T>(ch.n1, ch.n2, ch.v == 0, grid, lds, s, c))
            return;
        if (DIR < 0 && ch.n1 && (ch.v == 0 || ch.v == 3) &&
            v4_cols_fwd<T>(ch.n1, ch.n2, ch.v == 0, grid, lds, s, c))
            return;
        if (!(ch.n1 && v2_cols<T, DIR>(ch.n1, ch.n2, grid, lds, s, c)))
            LG_LAUNCH((k_cols<T, DIR>), grid, dim3(PBLK), lds, s, c);
    });
}

// The Lagrangian model's row pass with contractions (lag_rows.hpp), with
// the default split of nx.  Returns false, launching nothing, when nx has
// no split that fits the kernel's block.
#define LG_LAGROW_SPLITS(X) \
    X(4, 4) X(4, 6) X(4, 8) X(6, 8) X(8, 8) X(8, 9) X(8, 12) X(8, 16) X(12, 12) X(12, 16) X(16, 16)

template <typename T>
bool lag_rows_ok(int N, int nrows) {
    int n1, n2;
    if (!split_of(N, n1, n2) || LAGROW_Q * n1 > LAGROW_BLK || (nrows & 1)) return false;
    bool known = false;
#define LG_CASE(A, B) known |= (n1 == A && n2 == B);
    LG_LAGROW_SPLITS(LG_CASE)
#undef LG_CASE
    return known && (size_t)LAGROW_Q * (n1 * (n2 + 1) + 1) * sizeof(cplx<T>) <= LDS_MAX;
}

template <typename T>
bool launch_lag_rows(KernelTimer& tm, const LagRowArgs<T>& a, int M, hipStream_t s) {
    const int N = a.plan.N;
    if (!lag_rows_ok<T>(N, a.nrows)) return false;
    int n1, n2;
    split_of(N, n1, n2);
    const size_t lds = (size_t)LAGROW_Q * (n1 * (n2 + 1) + 1) * sizeof(cplx<T>);
    const dim3 grid(a.nrows / 2, a.nlev * M);
    // algorithmic bytes: the 21 band-limited half spectra read, two FP64
    // fields written
    const double bytes = (double)M * a.nlev * a.nrows *
                         (LAGROW_Q * (double)a.nce * sizeof(cplx<T>) + 2.0 * N * sizeof(double));
    tm.run(s, KK_LAG_ROWS, bytes, [&] {
#define LG_CASE(A, B)                                                                \
    if (n1 == A && n2 == B) LG_LAUNCH((k_lag_rows<T, A, B>), grid, dim3(LAGROW_BLK), lds, s, a);
        LG_LAGROW_SPLITS(LG_CASE)
#undef LG_CASE
    });
    return true;
}

// The Lagrangian model's forward row pass forming its quantities
// (lag_rows.hpp); launchable where lag_rows_ok is.
template <typename T>
bool launch_lag_prep_rows(KernelTimer& tm, const LagPrepArgs<T>& a, int M, hipStream_t s) {
    const int N = a.plan.N;
    int n1, n2;
    if (!lag_rows_ok<T>(N, 2) || !split_of(N, n1, n2) || LAGROW_Q * n2 > LAGROW_BLK ||
        a.nwr < 1 || a.nwr > a.ncols)
        return false;
    const size_t lds = (size_t)LAGROW_Q * (n1 * (n2 + 1) + 1) * sizeof(cplx<T>);
    const dim3 grid(a.fplane / a.ld / 2, a.nlev * M);
    // algorithmic bytes: u, v, w and the six S_ij read once, the 21 half
    // spectra written over the band
    const double rows = (double)M * a.nlev * (a.fplane / a.ld);
    const double bytes = rows * (9.0 * N * sizeof(T) + LAGROW_Q * (double)a.nwr * sizeof(cplx<T>));
    tm.run(s, KK_LAG_PREP, bytes, [&] {
#define LG_CASE(A, B)                                                                \
    if (n1 == A && n2 == B) LG_LAUNCH((k_lag_prep_rows<T, A, B>), grid, dim3(LAGROW_BLK), lds, s, a);
        LG_LAGROW_SPLITS(LG_CASE)
#undef LG_CASE
    });
    return true;
}

// The split of the fused products pass (prod_rows.hpp) for length N: the
// forward row pass's recorded or heuristic four-step split, else the first
// instantiated one; one row pair per block (measured on the 384-point 3/2
// grid of 256 x 256 x 192: 234, 236 and 245 us per pass in mixed precision
// for 1, 2 and 3 pairs, 406, 415 and 435 us in FP64).  False: N has no
// four-step split.
template <typename T>
static bool prod_rows_choice(int N, int nrows, int& n1, int& n2, int& P) {
    const PassChoice ch = choice_for<T>(PK_R2C, N, nrows);
    n1 = ch.n1;
    n2 = ch.n2;
    if (!n1) {
        const auto sp = splits_for(N);
        if (sp.empty()) return false;
        n1 = sp[0].first;
        n2 = sp[0].second;
    }
    P = 1;
    return 3 * P * std::max(n1, n2) <= (int)four_step_block(3 * P, n1, n2);
}

// Depends on N only (an instantiated split exists), not on the launch table.
template <typename T>
bool prod_rows_ok(int N, int nrows) {
    int n1, n2, P;
    return !splits_for(N).empty() && prod_rows_choice<T>(N, nrows, n1, n2, P);
}

template <typename T>
bool launch_rows_r2c_prod(KernelTimer& tm, const FftPlan& pl, const ProdSrc<T>& p,
                          long src_plane, int src_ld, int lev0, long src_ms,
                          cplx<T>* const* dst, long dst_plane, long dst_ms, int nrows, int ncols,
                          int nlev, int M, hipStream_t s) {
    int n1, n2, P;
    if (!prod_rows_choice<T>(pl.N, nrows, n1, n2, P)) return false;
    RowR2CArgs<T> a{};
    a.plan = pl;
    for (int i = 0; i < 3; ++i) a.dst[i] = dst[i];
    a.src_plane = src_plane;
    a.dst_plane = dst_plane;
    a.src_ld = src_ld;
    a.nrows = nrows;
    a.ncols = ncols;
    a.lev0_src = lev0;
    a.lev0_dst = lev0;
    a.nlev = nlev;
    a.src_mstride = src_ms;
    a.dst_mstride = dst_ms;
    a.S = P;
    a.stride = n1 * (n2 + 1) + 1;
    const size_t lds = (size_t)3 * P * a.stride * sizeof(cplx<T>);
    const bool twl = lds + (size_t)pl.N * sizeof(cplx<T>) <= LDS_DEFAULT;
    const size_t l4 = lds + (twl ? (size_t)pl.N * sizeof(cplx<T>) : 0);
    const dim3 grid((nrows / 2 + P - 1) / P, nlev * M);
    const dim3 blk(four_step_block(3 * P, n1, n2));
    // algorithmic bytes: the six fields read once on the levels that have a
    // product (1..nz-1 of lev0..lev0+nlev-1), the three half spectra written
    const int kr = std::max(0, std::min(lev0 + nlev - 1, p.nz - 1) - std::max(lev0, 1) + 1);
    const double bytes = (double)M * nrows *
                         (6.0 * kr * pl.N * sizeof(T) + 3.0 * nlev * ncols * sizeof(cplx<T>));
    bool ok = false;
    tm.run(s, KK_ROWS_R2C_PROD, bytes, [&] {
#define LG_CASE(A, B)                                                                      \
    if (n1 == A && n2 == B) {                                                              \
        if (twl) LG_LAUNCH((k_rows_r2c_prod<T, A, B, true>), grid, blk, l4, s, a, p);      \
        else LG_LAUNCH((k_rows_r2c_prod<T, A, B, false>), grid, blk, l4, s, a, p);         \
        ok = true;                                                                         \
    }
        LG_SPLITS(LG_CASE)
#undef LG_CASE
    });
    return ok;
}

}  // namespace lg
