// Transform-pass launchers with tunable launch choices.  A pass moves
// njobs arrays of nlev planes for each of M members: member m's planes start
// at m * mstride.
#pragma once

#include <utility>
#include <vector>

#include "fft.hpp"
#include "timer.hpp"

namespace lg {

// LDS a workgroup may take: the tuner's bound (gfx950 gives a workgroup the
// CU's 160 KB) and the built-in heuristics' budget.
constexpr size_t LDS_MAX = 160 * 1024;
constexpr size_t LDS_DEFAULT = 64 * 1024;

// Launch choice of one transform pass: the four-step split (n1 = 0: the
// generic LDS kernel) and the rows (pairs) or columns per workgroup.
struct PassChoice {
    int n1 = 0, n2 = 0, S = 0;
    int v = 0;          // rows: 1 = the block's rows staged through LDS (fft3.hpp);
                        // inverse rows: 0 = specialised by band (fft4.hpp);
                        // columns: 0 = specialised by map (fft4.hpp);
                        // forward rows: 0 = 32-bit offsets (fft4.hpp);
                        // rows and columns: 2 = the unspecialised four-step
                        // kernel (fft2.hpp), 3 = the fft4.hpp kernel with
                        // twiddles from global memory (0: an LDS copy)
};

// Pass kinds of the tuning table (column passes by direction and map).
enum PassKind : int {
    PK_R2C = 0,
    PK_C2R = 1,
    PK_COLS_FWD = 2,       // forward, in place (MAP_ID)
    PK_COLS_INV = 3,       // inverse (MAP_ID)
    PK_TRIDAG = 4,
    PK_COLS_UNPAD = 5,     // forward, big grid -> small (MAP_UNPAD)
    PK_COLS_PAD = 6,       // inverse, small -> big grid (MAP_PAD)
    PK_NKINDS = 7
};

inline int cols_kind(int dir, int map) {
    if (map == MAP_PAD) return PK_COLS_PAD;
    if (map == MAP_UNPAD) return PK_COLS_UNPAD;
    return dir < 0 ? PK_COLS_FWD : PK_COLS_INV;
}

// LDS complex elements per sequence (row or column) of a pass choice.
inline int seq_stride(const PassChoice& c, int N) { return c.n1 ? c.n1 * (c.n2 + 1) + 1 : N + 1; }

// LDS bytes of a row pass choice (ld_real: the real rows' stride).
template <typename T>
inline size_t row_lds_bytes(const PassChoice& c, int N, int ld_real, int ncols) {
    size_t b = (size_t)c.S * seq_stride(c, N) * sizeof(cplx<T>);
    if (c.v == 1) {
        const int w = ld_real > 2 * ncols ? ld_real : 2 * ncols;
        b += ((2ull * c.S * w * sizeof(T) + 15) & ~15ull);
    }
    return b;
}

// Process-wide table of tuned launch choices keyed by (kind, element size,
// transform length, rows or columns).  Choices depend on the grid only,
// never on the member count, so a member's arithmetic is the same alone
// and in any batch.
bool tuned_choice(int kind, int tsize, int N, int other, PassChoice& out);
void set_tuned_choice(int kind, int tsize, int N, int other, const PassChoice& c);

void clear_tuned_choices();
// Entries of the table in key order: key = (kind, element size, N, other).
int tuned_count();
bool tuned_entry(int i, int* key, PassChoice& c);

// Launchable (LDS, split instantiated, r2c one-round rule) / heuristic choice.
template <typename T>
bool choice_ok(int kind, const PassChoice& c, int N, int other);
template <typename T>
PassChoice default_choice(int kind, int N, int other);
// Tuned choice if one is recorded and launchable, the heuristic otherwise.
template <typename T>
PassChoice choice_for(int kind, int N, int other);
// Four-step splits instantiated for length N.
std::vector<std::pair<int, int>> pass_splits(int N);

// Row pass, real to complex (paired rows) / complex to real.  force, when
// given, overrides the table (the tuner's trial launches).  Band-limited
// spectra: r2c writes only columns < nwr (0: all), c2r reads only columns
// < nce[job] (null or 0: all) and takes the rest as zero.  c2r jobs with
// dst2[job] set also write the x derivative (i kx times the input, kxs =
// 2 pi / Lx) there.
template <typename T>
void launch_rows_r2c(KernelTimer& tm, const FftPlan& pl, int njobs, const T* const* src,
                     const T* scales, long src_plane, int src_ld, int src_lev0, long src_ms,
                     cplx<T>* const* dst, long dst_plane, int dst_lev0, long dst_ms, int nrows,
                     int ncols, int nlev, int M, hipStream_t s, const PassChoice* force,
                     const double* mscale = nullptr, int mscale_stride = 0, int nwr = 0);
template <typename T>
void launch_rows_c2r(KernelTimer& tm, const FftPlan& pl, int njobs, const cplx<T>* const* src,
                     long src_plane, int src_lev0, long src_ms, T* const* dst, long dst_plane,
                     int dst_ld, int dst_lev0, long dst_ms, int nrows, int ncols, int nlev, int M,
                     hipStream_t s, const PassChoice* force, const int* nce = nullptr,
                     T* const* dst2 = nullptr, T kxs = T(0));
// Column pass of nj jobs over nlev * M planes (c.src/dst/op/plan filled;
// sets c.TC, c.stride, c.nlev).
template <typename T, int DIR>
void launch_cols(KernelTimer& tm, ColArgs<T>& c, int nj, int nlev, int M, int ny_in,
                 hipStream_t s, const PassChoice* force);
// The Lagrangian model's inverse row pass with its tensor contractions
// (levels a.nlev of M members).  False: not launchable for this nx.
template <typename T>
bool launch_lag_rows(KernelTimer& tm, const LagRowArgs<T>& a, int M, hipStream_t s);
template <typename T>
bool lag_rows_ok(int N, int nrows);
// The 3/2 grid's forward row pass with the convective products formed in
// its loads (prod_rows.hpp): the x, y, z products of levels lev0 ..
// lev0 + nlev - 1 into dst[0..2].  False: no four-step split for this length.
template <typename T>
bool launch_rows_r2c_prod(KernelTimer& tm, const FftPlan& pl, const ProdSrc<T>& p,
                          long src_plane, int src_ld, int lev0, long src_ms,
                          cplx<T>* const* dst, long dst_plane, long dst_ms, int nrows, int ncols,
                          int nlev, int M, hipStream_t s);
template <typename T>
bool prod_rows_ok(int N, int nrows);
// The Lagrangian model's forward row pass forming its 21 quantities.
template <typename T>
bool launch_lag_prep_rows(KernelTimer& tm, const LagPrepArgs<T>& a, int M, hipStream_t s);

}  // namespace lg
