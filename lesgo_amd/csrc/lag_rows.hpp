// The Lagrangian model's row passes: the inverse pass with its tensor
// contractions, and the forward pass that forms the quantities to filter.
//
// Inverse: the Lagrangian model's inverse row pass with its tensor contractions
// (lagrange_Sdep.f90:137-167, 203-259): for one row pair (j, j+1) of one
// level, the 21 quantities filtered at one width (the column pass's output
// of ubar, vbar, wbar, the six products, S_ij and |S| S_ij) are transformed
// back along x in LDS and contracted at once into L_ij M_ij and M_ij M_ij
// (2 delta) or Q_ij N_ij and N_ij N_ij (4 delta), stored in FP64.  The 21
// filtered fields never reach HBM.
//
// The transforms are the paired c2r of k_rows_c2r_v2 (same four-step
// arithmetic); the contractions are k_lag_update's.  LDS: 21 slots of
// Split<N1, N2>::YS complex elements; a slot holds a row pair's
// intermediate, then its physical values (row j in .x, row j + 1 in .y).
#pragma once

#include "fft2.hpp"
#include "pointwise.hpp"

namespace lg {

constexpr int LAGROW_BLK = 384;
constexpr int LAGROW_Q = 21;

template <typename T, int N1, int N2>
__global__ void __launch_bounds__(LAGROW_BLK) k_lag_rows(LagRowArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    constexpr int half = N / 2;
    constexpr int YS = SP::YS;
    static_assert(LAGROW_Q * N1 <= LAGROW_BLK, "step 2 in one round");
    extern __shared__ __align__(16) unsigned char smem_lr[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_lr);
    const auto* tw = plan_tw<T>(a.plan);
    const int mem = blockIdx.y / a.nlev;
    const int c = blockIdx.y - mem * a.nlev;
    const int j = 2 * blockIdx.x;
    const int nc = a.ncols, nce = a.nce;
    const int tid = threadIdx.x;
    const cplx<T>* base = a.src + (long)mem * a.mstride + (long)c * a.plane + (long)j * nc;
    // step 1: N1-point DFTs of the Hermitian-extended packed rows
    for (int t = tid; t < LAGROW_Q * N2; t += blockDim.x) {
        const int q = t / N2, n2 = t - q * N2;
        const cplx<T>* ra = base + (long)q * a.qstride;
        const cplx<T>* rb = ra + nc;
        // every in-band load is issued before any is used (the band is a
        // runtime value: a use inside its branch would wait for each load in
        // turn), then the Hermitian extension
        cplx<T> A[N1], B[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
            const int k = n2 + N2 * n1;
            const int kk = k < half ? k : N - k;
            A[n1] = B[n1] = cplx<T>{(T)0, (T)0};
            if (k != half && kk < nce) {
                A[n1] = gld(ra + kk);
                B[n1] = gld(rb + kk);
            }
        }
        cplx<T> x[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
            const int k = n2 + N2 * n1;
            const int kk = k < half ? k : N - k;
            cplx<T> z{(T)0, (T)0};
            if (k == 0) {
                z = cplx<T>{A[n1].x, B[n1].x};
            } else if (k < half) {
                if (kk < nce) z = cplx<T>{A[n1].x - B[n1].y, A[n1].y + B[n1].x};
            } else if (k > half) {
                if (kk < nce) z = cplx<T>{A[n1].x + B[n1].y, B[n1].x - A[n1].y};
            }
            x[n1] = z;
        }
        dft_reg<T, N1, +1>(x);
        cplx<T>* Y = lds + q * YS;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1)
            Y[k1 * SP::YR + n2] = (k1 == 0 || n2 == 0) ? x[k1] : cmul(x[k1], wbig<T, +1>(tw, n2 * k1));
    }
    __syncthreads();
    // step 2 (one round): N2-point DFTs; every LDS read is done before the
    // barrier, then the physical values overwrite the slot in natural order
    {
        const int q = tid / N1, k1 = tid - q * N1;
        const bool act = q < LAGROW_Q;
        cplx<T> y[N2];
        if (act) {
            const cplx<T>* Y = lds + q * YS + k1 * SP::YR;
#pragma unroll
            for (int n2 = 0; n2 < N2; ++n2) y[n2] = Y[n2];
            dft_reg<T, N2, +1>(y);
        }
        __syncthreads();
        if (act) {
            cplx<T>* X = lds + q * YS;
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) X[k1 + N1 * k2] = y[k2];
        }
        __syncthreads();
    }
    // the contractions at each point of the two rows (k_lag_update's
    // formulas, FP64 from the filtered values); a thread takes both rows of
    // a point, so each slot element is read from LDS once
    const double cn = a.cn, tf2 = a.tf2;
    for (int n = tid; n < N; n += blockDim.x) {
        cplx<T> val[LAGROW_Q];
#pragma unroll
        for (int q = 0; q < LAGROW_Q; ++q) val[q] = lds[q * YS + n];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            auto V = [&](int q) -> double { return (double)(r ? val[q].y : val[q].x); };
            const double ub = V(0), vb = V(1), wb = V(2);
            const double L11 = V(3) - ub * ub, L12 = V(4) - ub * vb, L13 = V(5) - ub * wb;
            const double L22 = V(6) - vb * vb, L23 = V(7) - vb * wb, L33 = V(8) - wb * wb;
            const double Sn = smag_norm(V(9), V(10), V(11), V(12), V(13), V(14));
            const double M11 = cn * (V(15) - tf2 * Sn * V(9));
            const double M12 = cn * (V(16) - tf2 * Sn * V(10));
            const double M13 = cn * (V(17) - tf2 * Sn * V(11));
            const double M22 = cn * (V(18) - tf2 * Sn * V(12));
            const double M23 = cn * (V(19) - tf2 * Sn * V(13));
            const double M33 = cn * (V(20) - tf2 * Sn * V(14));
            const double LM = L11 * M11 + L22 * M22 + L33 * M33 + 2.0 * (L12 * M12 + L13 * M13 + L23 * M23);
            const double MM = M11 * M11 + M22 * M22 + M33 * M33 + 2.0 * (M12 * M12 + M13 * M13 + M23 * M23);
            const long o = (long)mem * a.omstride + (long)c * a.oplane + (long)(j + r) * a.ld + n;
            a.out0[o] = LM;
            a.out1[o] = MM;
        }
    }
}

// Forward: the 21 quantities the test filters act on (lagrange_Sdep.f90:
// 93-151, 211-240; k_lag_prep's formulas) formed for one row pair of one
// level from u, v, w and S_ij, and transformed along x in LDS (the paired
// r2c of k_rows_r2c_v2), half spectra written for kx < nwr only (the
// filters' band).  Quantities: 0-2 ubar, vbar, wbar; 3-8 the products 11,
// 12, 13, 22, 23, 33; 9-14 S_ij; 15-20 |S| S_ij.
template <typename T, int N1, int N2>
__global__ void __launch_bounds__(LAGROW_BLK) k_lag_prep_rows(LagPrepArgs<T> a) {
    using SP = Split<N1, N2>;
    constexpr int N = SP::N;
    constexpr int YS = SP::YS;
    static_assert(LAGROW_Q * N1 <= LAGROW_BLK && LAGROW_Q * N2 <= LAGROW_BLK,
                  "each step in one round");
    extern __shared__ __align__(16) unsigned char smem_lp[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_lp);
    const auto* tw = plan_tw<T>(a.plan);
    const int mem = blockIdx.y / a.nlev;
    const int c = blockIdx.y - mem * a.nlev;
    const int k = a.k0 + c;
    const int j = 2 * blockIdx.x;
    const int tid = threadIdx.x;
    const long P = a.fplane;
    // the quantities at every point of the two rows, into slot q (row j in
    // .x, row j + 1 in .y), natural order.  A thread takes points t and
    // t + blockDim.x (2N <= 2 blockDim.x) and loads both before forming
    // either, so the second round's loads do not wait for the first's stores
    static_assert(2 * N <= 2 * LAGROW_BLK, "two rounds at most");
    T in[2][9];
    bool has[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int t = tid + h * (int)blockDim.x;
        has[h] = t < 2 * N;
        const int tt = has[h] ? t : tid;
        const int r = tt / N, i = tt - r * N;
        const long o = (long)mem * a.fmstride + (long)k * P + (long)(j + r) * a.ld + i;
        // u, v at k and k - 1 (level 1: k only), w at k + 1 (level 1) or k
        const long ou0 = o, ou1 = k == 1 ? o : o - P;
        const long ow = k == 1 ? o + P : o;
        in[h][0] = a.u[ou0];
        in[h][1] = a.u[ou1];
        in[h][2] = a.v[ou0];
        in[h][3] = a.v[ou1];
        in[h][4] = a.w[ow];
        in[h][5] = a.S11[o];
        in[h][6] = a.S12[o];
        in[h][7] = a.S13[o];
        in[h][8] = a.S22[o];
    }
    T s23[2], s33[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int t = has[h] ? tid + h * (int)blockDim.x : tid;
        const long o = (long)mem * a.fmstride + (long)k * P + (long)(j + t / N) * a.ld + (t % N);
        s23[h] = a.S23[o];
        s33[h] = a.S33[o];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (!has[h]) continue;
        const int t = tid + h * (int)blockDim.x;
        const int r = t / N, i = t - r * N;
        T ub, vb, wb;
        if (k == 1) {
            ub = in[h][0];
            vb = in[h][2];
            wb = (T)0.25 * in[h][4];
        } else if (k == a.nz) {
            ub = in[h][1];
            vb = in[h][3];
            wb = (T)0;
        } else {
            ub = (T)0.5 * (in[h][0] + in[h][1]);
            vb = (T)0.5 * (in[h][2] + in[h][3]);
            wb = in[h][4];
        }
        const T s11 = in[h][5], s12 = in[h][6], s13 = in[h][7], s22 = in[h][8],
                s23v = s23[h], s33v = s33[h];
        const T S = smag_norm(s11, s12, s13, s22, s23v, s33v);
        const T q[LAGROW_Q] = {ub,      vb,      wb,      ub * ub, ub * vb, ub * wb, vb * vb,
                               vb * wb, wb * wb, s11,     s12,     s13,     s22,     s23v,
                               s33v,    S * s11, S * s12, S * s13, S * s22, S * s23v, S * s33v};
#pragma unroll
        for (int qq = 0; qq < LAGROW_Q; ++qq) {
            T* e = reinterpret_cast<T*>(lds + qq * YS + i) + r;
            *e = q[qq];
        }
    }
    __syncthreads();
    // forward transforms: step 1 from natural order into the four-step
    // layout of a second copy would need twice the LDS, so each task reads
    // its N1 inputs into registers first and the slot is rewritten after a
    // barrier
    {
        cplx<T> x[N1];
        const int q = tid / N2, n2 = tid - (tid / N2) * N2;
        const bool act = q < LAGROW_Q;
        if (act) {
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) x[n1] = lds[q * YS + n2 + N2 * n1];
            dft_reg<T, N1, -1>(x);
        }
        __syncthreads();
        if (act) {
            cplx<T>* Y = lds + q * YS;
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1)
                Y[k1 * SP::YR + n2] = (k1 == 0 || n2 == 0) ? x[k1] : cmul(x[k1], wbig<T, -1>(tw, n2 * k1));
        }
        __syncthreads();
    }
    {
        const int q = tid / N1, k1 = tid - (tid / N1) * N1;
        const bool act = q < LAGROW_Q;
        cplx<T> y[N2];
        if (act) {
            const cplx<T>* Y = lds + q * YS + k1 * SP::YR;
#pragma unroll
            for (int n2 = 0; n2 < N2; ++n2) y[n2] = Y[n2];
            dft_reg<T, N2, -1>(y);
        }
        __syncthreads();
        if (act) {
            cplx<T>* X = lds + q * YS;
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) X[k1 + N1 * k2] = y[k2];
        }
        __syncthreads();
    }
    // split each packed transform into the two rows' half spectra, kx < nwr
    const int nc = a.ncols, nw = a.nwr;
    for (int t = tid; t < LAGROW_Q * nw; t += blockDim.x) {
        const int q = t / nw, kx = t - q * nw;
        const cplx<T>* X = lds + q * YS;
        const cplx<T> z = X[kx];
        const cplx<T> zc = X[(N - kx) % N];
        const cplx<T> A{(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y - zc.y)};
        const cplx<T> Bm{(T)0.5 * (z.x - zc.x), (T)0.5 * (z.y + zc.y)};
        cplx<T>* o = a.out + (long)q * a.qstride + (long)mem * a.mstride + (long)c * a.plane +
                     (long)j * nc;
        gst(o + kx, A);
        gst(o + nc + kx, cplx<T>{Bm.y, -Bm.x});
    }
}

}  // namespace lg
