// tridag_array as a standalone entry point (tridag_array.f90:165-246): the
// batched Thomas solve over (kx, ky) with caller-supplied real coefficients
// a, b, c (lh, ny, nz+1) and an interleaved complex right-hand side r and
// solution u (ld, ny, nz+1), all FP64 in the reference's column-major layout
// (C order [nz+1][ny][lh] and [nz+1][ny][ld]).  The time step itself uses
// the fused k_tridag of pressure.hpp, which assembles the same coefficients
// from the wavenumbers; this is the drop-in for callers that pass their own.
#include <mutex>
#include <string>

#include "common.hpp"
#include "../../include/lesgo_amd.h"

namespace lg {

// Status words at the head of the kernel's status buffer: [0] set to 1 by the
// first system whose pivot is zero, [1..3] that system's jx, jy (1-based,
// the reference's indices) and row j.
constexpr int TRI_STATUS_INTS = 64;

// One thread per system (jx < lh-1, jy < ny): adjacent threads take adjacent
// jx, so every level's a, b, c loads are contiguous doubles and the r, u
// accesses contiguous 16-byte complex values.  Row 1 is solved for every
// system (tridag_array.f90:190-200); rows 2..nz+1 skip jy = ny/2+1 and
// (jx, jy) = (1, 1) as the reference does (207-209), leaving those entries
// of u as they were.  gam holds the forward-sweep ratios (lh, ny, nz+1).
__global__ void __launch_bounds__(256) k_tridag_array(const double* __restrict__ a,
                                                      const double* __restrict__ b,
                                                      const double* __restrict__ c,
                                                      const double2* __restrict__ r,
                                                      double2* __restrict__ u, double* gam,
                                                      int lh, int ny, int n, int* status) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nsys = (lh - 1) * ny;
    if (gid >= nsys) return;
    const int jy = gid / (lh - 1);
    const int jx = gid - jy * (lh - 1);
    const long plane = (long)ny * lh;
    const long o = (long)jy * lh + jx;
    auto fail = [&](int row) {
        if (atomicCAS(status, 0, 1) == 0) {
            status[1] = jx + 1;
            status[2] = jy + 1;
            status[3] = row;
        }
    };
    const double b0 = b[o];
    if (b0 == 0.0) {
        fail(1);
        return;
    }
    double2 up = r[o];
    up = make_double2(up.x / b0, up.y / b0);
    u[o] = up;
    if (jy == ny / 2 || (jx == 0 && jy == 0)) return;
    double bet = b0;
    for (int j = 1; j < n; ++j) {
        const long oj = (long)j * plane + o;
        const double gm = c[oj - plane] / bet;
        gam[oj] = gm;
        const double aj = a[oj];
        bet = b[oj] - aj * gm;
        if (bet == 0.0) {
            fail(j + 1);
            return;
        }
        const double2 rj = r[oj];
        up = make_double2((rj.x - aj * up.x) / bet, (rj.y - aj * up.y) / bet);
        u[oj] = up;
    }
    for (int j = n - 2; j >= 0; --j) {
        const long oj = (long)j * plane + o;
        const double gm = gam[oj + plane];
        const double2 uj = u[oj];
        up = make_double2(uj.x - gm * up.x, uj.y - gm * up.y);
        u[oj] = up;
    }
}

__global__ void k_clear_ints(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Launches the solve on `stream` with gam and status supplied by the caller,
// then waits for the stream and turns a recorded zero pivot into
// LG_ERR_PIVOT with the reference's message (tridag_array.f90:190-193,
// 217-222).
int tridag_array_run(int lh, int ny, int n, const double* a, const double* b, const double* c,
                     const double* r, double* u, double* gam, int* status, hipStream_t stream) {
    hipLaunchKernelGGL(k_clear_ints, dim3(1), dim3(TRI_STATUS_INTS), 0, stream, status,
                       TRI_STATUS_INTS);
    const int nsys = (lh - 1) * ny;
    hipLaunchKernelGGL(k_tridag_array, dim3((nsys + 255) / 256), dim3(256), 0, stream, a, b, c,
                       reinterpret_cast<const double2*>(r), reinterpret_cast<double2*>(u), gam,
                       lh, ny, n, status);
    LG_HIP(hipGetLastError());
    int st[4] = {0, 0, 0, 0};
    LG_HIP(hipMemcpyAsync(st, status, sizeof(st), hipMemcpyDeviceToHost, stream));
    LG_HIP(hipStreamSynchronize(stream));
    if (st[0]) {
        set_error("tridag_array: zero pivot at jx, jy, j = " + std::to_string(st[1]) + ", " +
                  std::to_string(st[2]) + ", " + std::to_string(st[3]) +
                  (st[3] == 1 ? " (rewrite eqs)" : ""));
        return LG_ERR_PIVOT;
    }
    return LG_OK;
}

bool tridag_array_args(int lh, int ny, int nzp1, const void* a, const void* b, const void* c,
                       const void* r, const void* u) {
    if (lh < 2 || ny < 2 || ny % 2 != 0 || nzp1 < 2 || !a || !b || !c || !r || !u) {
        set_error("tridag_array: need lh >= 2, even ny >= 2, nz+1 >= 2 and non-null arrays");
        return false;
    }
    return true;
}

}  // namespace lg

// Device arrays in, device array out; the library allocates the forward-sweep
// scratch gam (lh * ny * (nz+1) doubles) and the status words for the call.
extern "C" int lg_tridag_array_device(int lh, int ny, int nzp1, const double* a,
                                      const double* b, const double* c, const double* r,
                                      double* u, void* stream) {
    if (!lg::tridag_array_args(lh, ny, nzp1, a, b, c, r, u)) return LG_ERR_ARG;
    double* gam = nullptr;
    int* status = nullptr;
    const size_t ng = (size_t)lh * ny * nzp1;
    if (hipMalloc(&gam, ng * sizeof(double)) != hipSuccess) {
        lg::set_error("tridag_array: hipMalloc of the Thomas scratch failed");
        return LG_ERR_ALLOC;
    }
    if (hipMalloc(&status, lg::TRI_STATUS_INTS * sizeof(int)) != hipSuccess) {
        (void)hipFree(gam);
        lg::set_error("tridag_array: hipMalloc of the status words failed");
        return LG_ERR_ALLOC;
    }
    int rc;
    try {
        rc = lg::tridag_array_run(lh, ny, nzp1, a, b, c, r, u, gam, status,
                                  reinterpret_cast<hipStream_t>(stream));
    } catch (const lg::HipError&) {
        rc = LG_ERR_HIP;
    }
    (void)hipFree(gam);
    (void)hipFree(status);
    return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free variant of lg_tridag_array_device.  Its scratch (the status
// words, then gam) is one file-scope device array, so a call makes no
// hipMalloc/hipFree.  Its limit: lh * ny * (nz+1) <= TRI_NOALLOC_MAX_POINTS
// (2^23 forward-sweep ratios, 64 MiB, held on every GPU the library runs on
// until the process ends; the 256 x 256 x 192 grid needs 129 * 256 * 194 =
// 6.4 M).  Every call on one GPU uses that GPU's copy of the array: calls
// that overlap on one GPU (other host threads, other streams) race.
namespace lg {
constexpr long TRI_NOALLOC_MAX_POINTS = 1L << 23;
constexpr long TRI_SCRATCH_HEAD = TRI_STATUS_INTS / 2;   // status words, in doubles
constexpr int TRI_MAX_DEVICES = 64;
}  // namespace lg

static __device__ __align__(256) double scratch[lg::TRI_SCRATCH_HEAD + lg::TRI_NOALLOC_MAX_POINTS];

namespace {
// Device address of `scratch` per device ordinal, looked up once per device
// with that device current (the array's host-side name is not a device
// address and is never passed to a kernel or a copy).
std::mutex g_scratch_mu;
double* g_scratch_dev[lg::TRI_MAX_DEVICES] = {};

double* scratch_on_current_device() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= lg::TRI_MAX_DEVICES) {
        lg::set_error("tridag_array (no-alloc): no usable current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        double* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            lg::set_error(std::string("tridag_array (no-alloc): hipGetSymbolAddress failed: ") +
                          hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}
}  // namespace

// The same solve as lg_tridag_array_device (the same kernels) with its
// scratch in the library's static device array; sizes above the limit are
// refused with LG_ERR_ARG.  The status words are cleared by a kernel at the
// start of every call; gam needs no clearing (every entry the backward sweep
// reads is written by the same call's forward sweep).
extern "C" int lg_tridag_array_device_noalloc(int lh, int ny, int nzp1, const double* a,
                                              const double* b, const double* c, const double* r,
                                              double* u, void* stream) {
    if (!lg::tridag_array_args(lh, ny, nzp1, a, b, c, r, u)) return LG_ERR_ARG;
    const long ng = (long)lh * ny * nzp1;
    if (ng > lg::TRI_NOALLOC_MAX_POINTS) {
        lg::set_error("tridag_array (no-alloc): lh * ny * (nz+1) = " + std::to_string(ng) +
                      " exceeds the static scratch of " +
                      std::to_string(lg::TRI_NOALLOC_MAX_POINTS) +
                      " points; use lg_tridag_array_device");
        return LG_ERR_ARG;
    }
    double* base = scratch_on_current_device();
    if (!base) return LG_ERR_HIP;
    int* status = reinterpret_cast<int*>(base);
    double* gam = base + lg::TRI_SCRATCH_HEAD;
    try {
        return lg::tridag_array_run(lh, ny, nzp1, a, b, c, r, u, gam, status,
                                    reinterpret_cast<hipStream_t>(stream));
    } catch (const lg::HipError&) {
        return LG_ERR_HIP;
    }
}

// Host arrays in and out, the reference's own call (tridag_array(a, b, c,
// r, u)): copies to the device, solves, copies u back.  Entries of u the
// reference does not write keep the caller's values.
extern "C" int lg_tridag_array(int lh, int ny, int nzp1, const double* a, const double* b,
                               const double* c, const double* r, double* u) {
    if (!lg::tridag_array_args(lh, ny, nzp1, a, b, c, r, u)) return LG_ERR_ARG;
    const size_t nc = (size_t)lh * ny * nzp1, nr = 2 * nc;
    double* d = nullptr;
    if (hipMalloc(&d, (3 * nc + 2 * nr) * sizeof(double)) != hipSuccess) {
        lg::set_error("tridag_array: hipMalloc of the device copies failed");
        return LG_ERR_ALLOC;
    }
    double *da = d, *db = da + nc, *dc = db + nc, *dr = dc + nc, *du = dr + nr;
    int rc = LG_OK;
    if (hipMemcpy(da, a, nc * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, b, nc * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dc, c, nc * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dr, r, nr * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(du, u, nr * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        lg::set_error("tridag_array: host to device copy failed");
        rc = LG_ERR_HIP;
    }
    if (rc == LG_OK) rc = lg_tridag_array_device(lh, ny, nzp1, da, db, dc, dr, du, nullptr);
    if (rc == LG_OK &&
        hipMemcpy(u, du, nr * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        lg::set_error("tridag_array: device to host copy failed");
        rc = LG_ERR_HIP;
    }
    (void)hipFree(d);
    return rc;
}
