// Forward row pass on the 3/2 grid that forms the convective products in
// its loads (convec.f90:172-305): the three components of u x omega are made
// from the six 3/2-grid fields (u, v, w and the vorticity, levels 0..nz) as
// the pass loads them and are transformed along x at once, so the products
// never reach HBM.
//
// A block takes P row pairs of one level k: each thread takes point pairs
// (i, i + 1) of one row (8-byte loads, consecutive threads on consecutive
// points), loads u, v, o3 at k, u, v at k - 1 and w, o1, o2 at k and k + 1
// (the blocks of the levels either side, launched nrows / 2P blocks before
// and after on the same XCD, read the same planes, so those loads mostly hit
// the caches), and writes the three products of its points into LDS in
// natural order: sequence q * P + s holds product q (0: x, 1: y, 2: z) of
// pair s, row a in .x and row b in .y.  The 3P sequences then take
// k_rows_r2c_v4's four-step arithmetic.  The products are
// k_products' expressions (same operands, same order, FP64 with no
// contraction, rounded once to T), so the transformed values are those of
// the unfused passes.
#pragma once

#include "fft4.hpp"
#include "pointwise.hpp"

namespace lg {

// One point's values at level k: u, v, o3 at k; w, o1, o2 at k + 1; u, v at
// k - 1 and w, o1, o2 at k.
template <typename T>
struct ProdPt {
    T u, v, o3, w1, o11, o21;
    T um, vm, w0, o10, o20;
};

// The three products at level k of one point (k_products' expressions): x, y
// on 1..nz-1, z on 2..nz-1, zero elsewhere (where k_products stores none or
// zero).
template <typename T>
__device__ __forceinline__ void conv_products(const ProdPt<T>& a, int k, int nz, double c,
                                              T& px, T& py, T& pz) {
#pragma clang fp contract(off)
    const double h = 0.5;
    double x = 0.0, y = 0.0, z = 0.0;
    const double ub0 = a.u, vb0 = a.v;
    if (k <= nz - 1) {
        const double o30 = a.o3;
        const double wb1 = a.w1, o11 = a.o11, o21 = a.o21;
        const double wb0 = a.w0, o10 = a.o10, o20 = a.o20;
        if (k == 1) {
            x = c * (vb0 * (-o30) + h * wb1 * o21);
            y = c * (ub0 * o30 + h * wb1 * (-o11));
        } else if (k == nz - 1) {
            x = c * (vb0 * (-o30) + h * wb0 * o20);
            y = c * (ub0 * o30 + h * wb0 * (-o10));
        } else {
            x = c * (vb0 * (-o30) + h * (wb1 * o21 + wb0 * o20));
            y = c * (ub0 * o30 + h * (wb1 * (-o11) + wb0 * (-o10)));
        }
    }
    if (k >= 2 && k <= nz - 1) {
        const double ubm = a.um, vbm = a.vm, o10 = a.o10, o20 = a.o20;
        z = (c * h) * ((ub0 + ubm) * (-o20) + (vb0 + vbm) * o10);
    }
    px = (T)x;
    py = (T)y;
    pz = (T)z;
}

// a.S: row pairs per block (P); a.src_* describe the fields' layout
// (a.src[] unused), a.dst[0..2] the x, y, z spectra; grid = (row-pair
// blocks, nlev * members).  Launched with
// 3P * max(N1, N2) <= blockDim.x.
template <typename T, int N1, int N2, bool TWL>
__global__ void __launch_bounds__(256) k_rows_r2c_prod(RowR2CArgs<T> a, ProdSrc<T> p) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    constexpr int NH = N / 2;
    extern __shared__ __align__(16) unsigned char smem_rp[];
    cplx<T>* ltw = reinterpret_cast<cplx<T>*>(smem_rp);
    cplx<T>* lds = TWL ? ltw + N : ltw;
    const auto* gtw = plan_tw<T>(a.plan);
    const int mem = blockIdx.y / a.nlev;
    const int pl = blockIdx.y - mem * a.nlev;
    const int k = a.lev0_src + pl;
    const int npairs = a.nrows / 2;
    const int pair0 = blockIdx.x * a.S;
    const int P = min(a.S, npairs - pair0);
    const int NS = 3 * P;
    const int ntask = 2 * P * NH;           // (row, point pair) tasks of a level
    const unsigned sld = (unsigned)a.src_ld;
    const unsigned nc = (unsigned)a.ncols;
    const long PL = a.src_plane;
    // fields at level k, the block's first row
    const long lk = mem * a.src_mstride + (long)k * PL + (long)(2 * pair0) * sld;
    const int nz = p.nz;
    tw_stage<T, TWL>(ltw, gtw, N);
    {
        const int tid = threadIdx.x;
        // products of the level into LDS, natural order; a thread issues the
        // loads of its next task before it forms the products of the current
        // one (the blocks are bound by LDS, so the registers this holds cost
        // no waves)
        auto load = [&](int task, ProdPt<T>& q0, ProdPt<T>& q1) {
            const int r = task / NH, ip = task - r * NH;
            const long o = lk + (long)r * sld + 2 * ip;
            ld_pair(p.ub + o, q0.u, q1.u);
            ld_pair(p.vb + o, q0.v, q1.v);
            ld_pair(p.o3 + o, q0.o3, q1.o3);
            ld_pair(p.ub + o - PL, q0.um, q1.um);
            ld_pair(p.vb + o - PL, q0.vm, q1.vm);
            ld_pair(p.wb + o, q0.w0, q1.w0);
            ld_pair(p.o1 + o, q0.o10, q1.o10);
            ld_pair(p.o2 + o, q0.o20, q1.o20);
            q0.w1 = q1.w1 = q0.o11 = q1.o11 = q0.o21 = q1.o21 = (T)0;
            if (k + 1 <= nz) {
                ld_pair(p.wb + o + PL, q0.w1, q1.w1);
                ld_pair(p.o1 + o + PL, q0.o11, q1.o11);
                ld_pair(p.o2 + o + PL, q0.o21, q1.o21);
            }
        };
        ProdPt<T> n0{}, n1{};
        if (tid < ntask) load(tid, n0, n1);
#pragma unroll 1
        for (int task = tid; task < ntask; task += blockDim.x) {
            // task: (row r of the block's 2P, point pair ip)
            const int r = task / NH, ip = task - r * NH;
            const int s = r >> 1, e = r & 1;
            const ProdPt<T> q0 = n0, q1 = n1;
            if (task + (int)blockDim.x < ntask) load(task + blockDim.x, n0, n1);
            T px[2], py[2], pz[2];
            conv_products(q0, k, nz, p.c, px[0], py[0], pz[0]);
            conv_products(q1, k, nz, p.c, px[1], py[1], pz[1]);
            // row a of the pair in .x, row b in .y
            T* X0 = reinterpret_cast<T*>(lds + s * SP::YS + 2 * ip) + e;
            T* X1 = X0 + 2 * P * SP::YS;
            T* X2 = X1 + 2 * P * SP::YS;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                X0[2 * j] = px[j];
                X1[2 * j] = py[j];
                X2[2 * j] = pz[j];
            }
        }
        __syncthreads();
        // step 1 (one round): each task reads its N1 inputs first; the slot
        // takes the four-step layout after a barrier
        cplx<T> x[N1];
        const int t1 = tid;
        const bool act1 = t1 < NS * N2;
        const int v1 = act1 ? t1 / N2 : 0, n2 = act1 ? t1 - v1 * N2 : 0;
        if (act1) {
            const cplx<T>* Xs = lds + v1 * SP::YS;
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) x[n1] = Xs[n2 + N2 * n1];
            dft_reg<T, N1, -1>(x);
        }
        __syncthreads();
        if (act1) {
            cplx<T>* Y = lds + v1 * SP::YS;
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) {
                cplx<T> w = x[k1];
                if (k1 != 0 && n2 != 0) w = cmul(w, tw_at<T, TWL>(ltw, gtw, n2 * k1));
                Y[k1 * SP::YR + n2] = w;
            }
        }
        __syncthreads();
        cplx<T> y[N2];
        const int t2 = tid;
        const bool act = t2 < NS * N1;
        const int v2 = act ? t2 / N1 : 0;
        const int k1 = act ? t2 - v2 * N1 : 0;
        if (act) {
            const cplx<T>* Y = lds + v2 * SP::YS + k1 * SP::YR;
#pragma unroll
            for (int m2 = 0; m2 < N2; ++m2) y[m2] = Y[m2];
            dft_reg<T, N2, -1>(y);
        }
        __syncthreads();
        if (act) {
            cplx<T>* X = lds + v2 * SP::YS;
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) X[k1 + N1 * k2] = y[k2];
        }
        __syncthreads();
        const int nw = a.nwr ? a.nwr : (int)nc;
        const long dofs = mem * a.dst_mstride + (long)(a.lev0_dst + pl) * a.dst_plane +
                          (long)(2 * pair0) * nc;
        for (int idx = tid; idx < NS * nw; idx += blockDim.x) {
            const int v = idx / nw;
            const int kx = idx - v * nw;
            const int q = v / P, s = v - q * P;
            const cplx<T>* X = lds + v * SP::YS;
            const cplx<T> z = X[kx];
            const cplx<T> zc = X[(N - kx) % N];
            const cplx<T> A{(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y - zc.y)};
            const cplx<T> Bm{(T)0.5 * (z.x - zc.x), (T)0.5 * (z.y + zc.y)};
            cplx<T>* db = a.dst[q] + dofs;
            const unsigned o = (2u * (unsigned)s * nc + (unsigned)kx) * (unsigned)sizeof(cplx<T>);
            st_bytes(db, o, A);
            st_bytes(db, o + nc * (unsigned)sizeof(cplx<T>), cplx<T>{Bm.y, -Bm.x});
        }
    }
}

}  // namespace lg
