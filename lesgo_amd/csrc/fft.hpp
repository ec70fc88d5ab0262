// Plane FFT passes for the x-y pseudo-spectral layer (replaces the FFTW
// plans of fft.f90:102-127 and the in-place dfftw_execute_dft_r2c/c2r calls).
//
// A 2-D real transform of an (nx, ny) plane is done in two launches:
//   * a row pass (x, contiguous): two real rows are packed into one complex
//     sequence, transformed, and split (r2c), or the reverse (c2r);
//   * a column pass (y, strided): tiles of TC complex columns are staged in
//     LDS and transformed along y, with fused spectral multipliers, Nyquist
//     zeroing and the 3/2-rule pad/unpad index maps.
// Each sequence is transformed in place in LDS by an iterative mixed-radix
// decimation-in-time FFT: the load scatters natural index n to its
// digit-reversed LDS slot, then every stage combines r sub-transforms of
// length L in place.  Conventions follow FFTW: forward sign -1, unnormalised.
//
// Spectral planes are stored compactly as [ky][kx] complex with kx in
// [0, nx/2): the kx = nx/2 column is zero in every use on the time step
// (derivatives.f90:50-52, press_stag_array.f90:130-146, fft.f90:43-99,
// test_filtermodule.f90:74-75) and is never stored.
#pragma once

#include <utility>

#include "common.hpp"

namespace lg {

constexpr int MAX_STAGES = 12;
constexpr int MAX_JOBS = 12;

struct FftPlan {
    int N = 0;
    int nstage = 0;
    int radix[MAX_STAGES] = {};
    const int* perm = nullptr;          // natural index -> LDS slot
    const double2* tw64 = nullptr;      // exp(-2 pi i j / N), j < N
    const float2* tw32 = nullptr;
};

// Global loads and stores of the transform passes (non-temporal when
// LG_NT is set at build time).
#ifndef LG_NT
#define LG_NT 0
#endif
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
#if LG_NT
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
template <typename T>
__device__ __forceinline__ cplx<T> gld(const cplx<T>* p) {
#if LG_NT
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return cplx<T>{v.x, v.y};
#else
    return *p;
#endif
}
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
#if LG_NT
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}
template <typename T>
__device__ __forceinline__ void gst(cplx<T>* p, cplx<T> c) {
#if LG_NT
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 v;
    v.x = c.x;
    v.y = c.y;
    __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p));
#else
    *p = c;
#endif
}

template <typename T> struct TwT;
template <> struct TwT<double> { using type = double2; };
template <> struct TwT<float> { using type = float2; };

template <typename T>
__device__ __forceinline__ const typename TwT<T>::type* plan_tw(const FftPlan& p);
template <>
__device__ __forceinline__ const double2* plan_tw<double>(const FftPlan& p) { return p.tw64; }
template <>
__device__ __forceinline__ const float2* plan_tw<float>(const FftPlan& p) { return p.tw32; }

// --------------------------------------------------------------------------
// In-place mixed-radix DIT stages on S sequences stored at lds + s*stride.
// DIR = -1 forward, +1 inverse.  Caller has loaded data digit-reversed and
// synchronised.  Ends with a barrier.
// --------------------------------------------------------------------------
template <typename T, int DIR>
__device__ void fft_stages(cplx<T>* lds, int S, int stride, const FftPlan& pl) {
    const auto* tw = plan_tw<T>(pl);
    const int N = pl.N;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    int L = 1;
    for (int st = 0; st < pl.nstage; ++st) {
        const int r = pl.radix[st];
        const int span = L * r;
        const int B = N / r;
        const int total = S * B;
        const int tws = N / span;
        for (int t = tid; t < total; t += nthr) {
            const int seq = t / B;
            const int rem = t - seq * B;
            const int b = rem / L;
            const int p = rem - b * L;
            cplx<T>* base = lds + seq * stride + b * span + p;
            if (r == 4) {
                cplx<T> a0 = base[0], a1 = base[L], a2 = base[2 * L], a3 = base[3 * L];
                if (p != 0) {
                    auto w1 = tw[p * tws];
                    auto w2 = tw[2 * p * tws];
                    auto w3 = tw[3 * p * tws];
                    cplx<T> c1{(T)w1.x, (T)(-DIR) * (T)w1.y};
                    cplx<T> c2{(T)w2.x, (T)(-DIR) * (T)w2.y};
                    cplx<T> c3{(T)w3.x, (T)(-DIR) * (T)w3.y};
                    a1 = cmul(a1, c1);
                    a2 = cmul(a2, c2);
                    a3 = cmul(a3, c3);
                }
                cplx<T> s02 = cadd(a0, a2), d02 = csub(a0, a2);
                cplx<T> s13 = cadd(a1, a3), d13 = csub(a1, a3);
                // multiply d13 by DIR * i  (forward: -i)
                cplx<T> jd13 = (DIR < 0) ? cplx<T>{d13.y, -d13.x} : cplx<T>{-d13.y, d13.x};
                base[0] = cadd(s02, s13);
                base[L] = cadd(d02, jd13);
                base[2 * L] = csub(s02, s13);
                base[3 * L] = csub(d02, jd13);
            } else if (r == 2) {
                cplx<T> a0 = base[0], a1 = base[L];
                if (p != 0) {
                    auto w1 = tw[p * tws];
                    a1 = cmul(a1, cplx<T>{(T)w1.x, (T)(-DIR) * (T)w1.y});
                }
                base[0] = cadd(a0, a1);
                base[L] = csub(a0, a1);
            } else {  // r == 3
                cplx<T> a0 = base[0], a1 = base[L], a2 = base[2 * L];
                if (p != 0) {
                    auto w1 = tw[p * tws];
                    auto w2 = tw[2 * p * tws];
                    a1 = cmul(a1, cplx<T>{(T)w1.x, (T)(-DIR) * (T)w1.y});
                    a2 = cmul(a2, cplx<T>{(T)w2.x, (T)(-DIR) * (T)w2.y});
                }
                const T c = (T)-0.5;
                const T sn = (T)(DIR * 0.86602540378443864676);  // sin(2 pi/3) * DIR
                cplx<T> s = cadd(a1, a2), d = csub(a1, a2);
                cplx<T> m{a0.x + c * s.x, a0.y + c * s.y};
                // i * sn * d
                cplx<T> id{-sn * d.y, sn * d.x};
                base[0] = cadd(a0, s);
                base[L] = cadd(m, id);
                base[2 * L] = csub(m, id);
            }
        }
        __syncthreads();
        L = span;
    }
}

// --------------------------------------------------------------------------
// Row pass, r2c.  Rows j and j+1 of a plane are packed z = a + i b, so one
// complex FFT of length N = nx gives both half spectra:
//   A[k] = (Z[k] + conj Z[N-k]) / 2,  B[k] = (Z[k] - conj Z[N-k]) / (2i).
// Output: complex rows of width `ncols` (kx < ncols <= N/2).
// --------------------------------------------------------------------------
// The Lagrangian model's inverse row pass with its contractions
// (lag_rows.hpp)
template <typename T>
struct LagRowArgs {
    FftPlan plan;              // length nx
    const cplx<T>* src;        // column-pass output of quantity q at src + q * qstride:
                               // [member][level][ky][kx < ncols]
    long qstride;              // complex elements between quantities
    long plane;                // complex elements per level (ny * ncols)
    long mstride;              // complex elements per member
    int nrows, ncols, nce;     // ny, nx / 2, columns inside the filter's band
    int nlev;                  // levels; grid.y = nlev * members
    double* out0;              // L_ij M_ij (or Q_ij N_ij): [member][level][j][ld]
    double* out1;              // M_ij M_ij (or N_ij N_ij)
    long oplane, omstride;     // elements per level / member of out0, out1
    int ld;
    double cn;                 // 2 delta^2
    double tf2;                // (filter width / delta)^2
};

// The Lagrangian model's forward row pass forming its quantities
// (lag_rows.hpp)
template <typename T>
struct LagPrepArgs {
    FftPlan plan;              // length nx
    const T *u, *v, *w;        // physical fields [member][0..nz][ny][ld]
    const T *S11, *S12, *S13, *S22, *S23, *S33;
    long fplane, fmstride;     // elements per level / member of those fields
    int ld, nz, k0, nlev;      // levels k0 .. k0 + nlev - 1; grid.y = nlev * members
    cplx<T>* out;              // quantity q at out + q * qstride: [member][level][j][kx < ncols]
    long qstride, plane, mstride;   // complex elements
    int ncols, nwr;            // nx / 2, columns written (the filters' band)
};

template <typename T>
struct RowR2CArgs {
    FftPlan plan;
    const T* src[MAX_JOBS];
    cplx<T>* dst[MAX_JOBS];
    T scale[MAX_JOBS];
    long src_plane;        // elements between planes of src
    long dst_plane;        // complex elements between planes of dst
    int src_ld;            // row stride of src (ld or ld_big)
    int nrows;             // rows per plane (ny or ny2)
    int ncols;             // complex columns stored per dst row
    int lev0_src;          // first source plane index
    int lev0_dst;          // first destination plane index
    int S;                 // row pairs per workgroup
    int stride;            // LDS stride per sequence (complex elements)
    int nlev;              // planes per member (grid.y = nlev * members)
    long src_mstride;      // elements between members of src
    long dst_mstride;      // complex elements between members of dst
    // per-member input scale (replaces scale[] when set): mscale[m * stride]
    const double* mscale;
    int mscale_stride;
    int nwr;               // columns of dst written (0: ncols; a band-limited consumer
                           // reads no more; the four-step and generic kernels honour it)
};

// The fields the convective products are formed from (prod_rows.hpp): u, v,
// w and the vorticity on the 3/2 grid, [member][0..nz][ny2][ld_big].
template <typename T>
struct ProdSrc {
    const T *ub, *vb, *wb, *o1, *o2, *o3;
    double c;              // 1 / (nx2 ny2)
    int nz;
};

template <typename T>
__global__ void __launch_bounds__(256) k_rows_r2c(RowR2CArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_r2c[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_r2c);
    const int N = a.plan.N;
    const int job = blockIdx.z;
    const int mem = blockIdx.y / a.nlev;
    const int plane = blockIdx.y - mem * a.nlev;
    const int npairs = a.nrows / 2;
    const int pair0 = blockIdx.x * a.S;
    const int S = min(a.S, npairs - pair0);
    const T* src = a.src[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane;
    cplx<T>* dst = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane;
    const T sc = a.mscale ? (T)a.mscale[(long)mem * a.mscale_stride] : a.scale[job];
    const int* perm = a.plan.perm;
    for (int idx = threadIdx.x; idx < S * N; idx += blockDim.x) {
        const int s = idx / N;
        const int n = idx - s * N;
        const int j = 2 * (pair0 + s);
        const T av = sc * src[(long)j * a.src_ld + n];
        const T bv = sc * src[(long)(j + 1) * a.src_ld + n];
        lds[s * a.stride + perm[n]] = cplx<T>{av, bv};
    }
    __syncthreads();
    fft_stages<T, -1>(lds, S, a.stride, a.plan);
    const int nc = a.ncols;
    const int nw = a.nwr ? a.nwr : nc;
    for (int idx = threadIdx.x; idx < S * nw; idx += blockDim.x) {
        const int s = idx / nw;
        const int k = idx - s * nw;
        const int j = 2 * (pair0 + s);
        const cplx<T> z = lds[s * a.stride + k];
        const cplx<T> zc = lds[s * a.stride + ((N - k) % N)];
        // A = (z + conj zc)/2 ; B = (z - conj zc)/(2i)
        const cplx<T> A{(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y - zc.y)};
        const cplx<T> Bm{(T)0.5 * (z.x - zc.x), (T)0.5 * (z.y + zc.y)};  // (z - conj zc)/2
        const cplx<T> B{Bm.y, -Bm.x};                                   // / i
        dst[(long)j * nc + k] = A;
        dst[(long)(j + 1) * nc + k] = B;
    }
}

// --------------------------------------------------------------------------
// Row pass, c2r.  Input half spectra of width ncols (kx >= ncols are zero,
// the Nyquist entry is zero, the imaginary part of kx = 0 is ignored as in
// FFTW's c2r).  Z = A + i B over the full Hermitian extension; inverse FFT;
// Re -> row j, Im -> row j+1.  Padding columns N..ld-1 are written as 0.
// --------------------------------------------------------------------------
template <typename T>
struct RowC2RArgs {
    FftPlan plan;
    const cplx<T>* src[MAX_JOBS];
    T* dst[MAX_JOBS];
    long src_plane;
    long dst_plane;
    int dst_ld;
    int nrows;
    int ncols;
    int lev0_src;
    int lev0_dst;
    int S;
    int stride;
    int nlev;
    long src_mstride;      // complex elements between members of src
    long dst_mstride;      // elements between members of dst
    int nce[MAX_JOBS];     // columns of src read per job (0: ncols); the rest are
                           // taken as zero (band-limited spectra)
    T* dst2[MAX_JOBS];     // when set, the job also writes the x derivative: the
                           // transform of i kx times its input (a second pass)
    T kxs;                 // 2 pi / Lx
};

template <typename T>
__global__ void __launch_bounds__(256) k_rows_c2r(RowC2RArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_c2r[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_c2r);
    const int N = a.plan.N;
    const int job = blockIdx.z;
    const int mem = blockIdx.y / a.nlev;
    const int plane = blockIdx.y - mem * a.nlev;
    const int npairs = a.nrows / 2;
    const int pair0 = blockIdx.x * a.S;
    const int S = min(a.S, npairs - pair0);
    const cplx<T>* src = a.src[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane;
    T* dst = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane;
    const int* perm = a.plan.perm;
    const int nc = a.ncols;
    const int nce = a.nce[job] ? a.nce[job] : nc;
    const int half = N / 2;
    T* const dst2 = a.dst2[job] ? a.dst2[job] + mem * a.dst_mstride +
                                      (long)(a.lev0_dst + plane) * a.dst_plane
                                : nullptr;
    for (int pass = 0; pass < (dst2 ? 2 : 1); ++pass) {
        T* const out = pass ? dst2 : dst;
        for (int idx = threadIdx.x; idx < S * N; idx += blockDim.x) {
            const int s = idx / N;
            const int k = idx - s * N;
            const int j = 2 * (pair0 + s);
            const cplx<T>* ra = src + (long)j * nc;
            const cplx<T>* rb = src + (long)(j + 1) * nc;
            // pass 1: the half spectra times i kx of their column
            auto col = [&](int q) {
                cplx<T> A = ra[q], B = rb[q];
                if (pass) {
                    const T kx = a.kxs * (T)q;
                    A = cplx<T>{-A.y * kx, A.x * kx};
                    B = cplx<T>{-B.y * kx, B.x * kx};
                }
                return std::make_pair(A, B);
            };
            cplx<T> z;
            if (k == 0) {
                const auto AB = col(0);
                z = cplx<T>{AB.first.x, AB.second.x};
            } else if (k < half) {
                if (k < nce) {
                    const auto AB = col(k);
                    const cplx<T> A = AB.first, B = AB.second;
                    z = cplx<T>{A.x - B.y, A.y + B.x};
                } else {
                    z = cplx<T>{(T)0, (T)0};
                }
            } else if (k == half) {
                z = cplx<T>{(T)0, (T)0};
            } else {
                const int kk = N - k;
                if (kk < nce) {
                    const auto AB = col(kk);
                    const cplx<T> A = AB.first, B = AB.second;
                    // conj(A) + i conj(B)
                    z = cplx<T>{A.x + B.y, B.x - A.y};
                } else {
                    z = cplx<T>{(T)0, (T)0};
                }
            }
            lds[s * a.stride + perm[k]] = z;
        }
        __syncthreads();
        fft_stages<T, +1>(lds, S, a.stride, a.plan);
        const int ld = a.dst_ld;
        for (int idx = threadIdx.x; idx < S * ld; idx += blockDim.x) {
            const int s = idx / ld;
            const int n = idx - s * ld;
            const int j = 2 * (pair0 + s);
            T va = (T)0, vb = (T)0;
            if (n < N) {
                const cplx<T> z = lds[s * a.stride + n];
                va = z.x;
                vb = z.y;
            }
            out[(long)j * ld + n] = va;
            out[(long)(j + 1) * ld + n] = vb;
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------------------
// Column pass.  Ops are applied at the small-grid spectral index (kx, ky)
// either before an inverse transform or after a forward one.
// --------------------------------------------------------------------------
enum ColOp : int {
    OP_NONE = 0,
    OP_IKX = 1,        // .MULI. kx   (emul_complex.f90:154-206)
    OP_IKY = 2,        // .MULI. ky
    OP_G1 = 3,         // .MULR. G_test       (test_filtermodule.f90:141)
    OP_G2 = 4,         // .MULR. G_test_test  (test_filtermodule.f90:164)
    OP_DIVXY = 5,      // (.MULI. kx) src + (.MULI. ky) src2: d/dx + d/dy of two fields
    OP_G1B = 6,        // G_test's band with unit weight, kx Nyquist column zero: the
                       // test-filtered field from spectra already normalised by 1/(nx ny)
};

enum ColMap : int {
    MAP_ID = 0,        // N == rows stored
    MAP_PAD = 1,       // input small (ny rows), transform on ny2 (padd, fft.f90:43-66)
    MAP_UNPAD = 2,     // transform on ny2, output small (unpadd, fft.f90:69-99)
};

template <typename T>
struct ColArgs {
    FftPlan plan;
    const cplx<T>* src[MAX_JOBS];
    const cplx<T>* src2[MAX_JOBS];     // second input of OP_DIVXY
    cplx<T>* dst[MAX_JOBS];
    int op[MAX_JOBS];
    long src_plane;
    long dst_plane;
    int lev0_src;
    int lev0_dst;
    int ncols;         // stored complex columns per row (nx/2)
    int ny;            // small-grid rows
    int map;           // ColMap
    int zero_nyq;      // forward: zero the ky = ny/2 row of the (small) output
    int TC;            // columns per tile
    int stride;        // LDS stride per column sequence
    int nlev;          // planes per member (grid.z = nlev * members)
    long src_mstride;  // complex elements between members (src, src2)
    long dst_mstride;
    T kxs;             // 2 pi / Lx
    T kys;             // 2 pi / Ly
    const T* G1;       // [ny][ncols] filter tables
    const T* G2;
    // band limits of band-limited (test-filtered) spectra, per job, 0: none.
    // nzc: only columns c < nzc are transformed (tiles past it do nothing).
    // ykeep: only rows with |ky index| < ykeep (not ny/2) carry data: an
    // inverse MAP_ID pass reads no others (they are zero), a forward MAP_ID
    // pass writes no others.
    int nzc[MAX_JOBS];
    int ykeep[MAX_JOBS];
    // inverse MAP_ID jobs: the input's ky = ny/2 row is taken as zero (as
    // filt_da's is) when set
    int nyq0[MAX_JOBS];
};

// Row jb of a small-grid spectrum inside the band |m| < ykeep (0: all rows).
__device__ __forceinline__ bool in_yband(int jb, int ny, int ykeep) {
    if (!ykeep) return true;
    const int nyh = ny / 2;
    if (jb == nyh) return false;
    const int m = jb < nyh ? jb : ny - jb;
    return m < ykeep;
}

// i kx v + i ky v2 at small-grid index (c, j) (Nyquist row: both 0)
template <typename T>
__device__ __forceinline__ cplx<T> divxy(cplx<T> v, cplx<T> v2, int c, int j, const ColArgs<T>& a) {
    if (j == a.ny / 2) return cplx<T>{(T)0, (T)0};
    const T kx = a.kxs * (T)c;
    const int m = (j < a.ny / 2) ? j : j - a.ny;
    const T ky = a.kys * (T)m;
    return cplx<T>{-v.y * kx + -v2.y * ky, v.x * kx + v2.x * ky};
}

template <typename T>
__device__ __forceinline__ cplx<T> apply_op(int op, cplx<T> v, int c, int j, const ColArgs<T>& a) {
    if (op == OP_NONE) return v;
    if (op == OP_IKX || op == OP_IKY) {
        T k;
        if (j == a.ny / 2) {
            k = (T)0;
        } else if (op == OP_IKX) {
            k = a.kxs * (T)c;
        } else {
            const int m = (j < a.ny / 2) ? j : j - a.ny;
            k = a.kys * (T)m;
        }
        return cplx<T>{-v.y * k, v.x * k};
    }
    if (op == OP_G1B) {
        const bool in = a.G1[(long)j * a.ncols + c] != (T)0 && c != a.ncols - 1;
        return in ? v : cplx<T>{(T)0, (T)0};
    }
    const T g = (op == OP_G1 ? a.G1 : a.G2)[(long)j * a.ncols + c];
    return cplx<T>{g * v.x, g * v.y};
}

template <typename T, int DIR>
__global__ void __launch_bounds__(256) k_cols(ColArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_col[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_col);
    const int N = a.plan.N;
    const int job = blockIdx.x;
    const int c0 = blockIdx.y * a.TC;
    const int nzc = a.nzc[job] ? a.nzc[job] : a.ncols;
    if (c0 >= nzc) return;
    const int tc = min(a.TC, nzc - c0);
    const int ykeep = a.ykeep[job];
    const int mem = blockIdx.z / a.nlev;
    const int plane = blockIdx.z - mem * a.nlev;
    const cplx<T>* src = a.src[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane;
    cplx<T>* dst = a.dst[job] + mem * a.dst_mstride + (long)(a.lev0_dst + plane) * a.dst_plane;
    const int op = a.op[job];
    const int* perm = a.plan.perm;
    const int ny = a.ny;
    const int nyh = ny / 2;
    // load (transform rows jb = 0..N-1)
    for (int idx = threadIdx.x; idx < N * tc; idx += blockDim.x) {
        const int jb = idx / tc;
        const int c = idx - jb * tc;
        cplx<T> v{(T)0, (T)0};
        if (a.map == MAP_PAD) {
            int j = -1;
            if (jb < nyh) j = jb;
            else if (jb > N - nyh) j = jb - (N - ny);
            if (j >= 0) v = apply_op<T>(op, src[(long)j * a.ncols + c0 + c], c0 + c, j, a);
        } else {
            v = (DIR > 0 && !in_yband(jb, ny, ykeep)) ? cplx<T>{(T)0, (T)0}
                                                      : src[(long)jb * a.ncols + c0 + c];
            if (DIR > 0 && a.map == MAP_ID) {
                if (a.nyq0[job] && jb == nyh) v = cplx<T>{(T)0, (T)0};
                if (op == OP_DIVXY) {
                    const cplx<T>* s2 = a.src2[job] + mem * a.src_mstride + (long)(a.lev0_src + plane) * a.src_plane;
                    v = divxy<T>(v, s2[(long)jb * a.ncols + c0 + c], c0 + c, jb, a);
                } else {
                    v = apply_op<T>(op, v, c0 + c, jb, a);
                }
            }
        }
        lds[c * a.stride + perm[jb]] = v;
    }
    __syncthreads();
    fft_stages<T, DIR>(lds, tc, a.stride, a.plan);
    if (a.map == MAP_UNPAD) {
        for (int idx = threadIdx.x; idx < ny * tc; idx += blockDim.x) {
            const int j = idx / tc;
            const int c = idx - j * tc;
            cplx<T> v{(T)0, (T)0};
            if (j < nyh) v = lds[c * a.stride + j];
            else if (j > nyh) v = lds[c * a.stride + j + (N - ny)];
            if (DIR < 0) v = apply_op<T>(op, v, c0 + c, j, a);
            dst[(long)j * a.ncols + c0 + c] = v;
        }
    } else {
        for (int idx = threadIdx.x; idx < N * tc; idx += blockDim.x) {
            const int j = idx / tc;
            const int c = idx - j * tc;
            cplx<T> v = lds[c * a.stride + j];
            if (DIR < 0 && a.map == MAP_ID) {
                if (a.zero_nyq && j == nyh) v = cplx<T>{(T)0, (T)0};
                v = apply_op<T>(op, v, c0 + c, j, a);
            }
            if (DIR > 0 || in_yband(j, ny, ykeep)) dst[(long)j * a.ncols + c0 + c] = v;
        }
    }
}

}  // namespace lg
