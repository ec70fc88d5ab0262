// Launch-choice table and the autotuner.  The tuner times every launchable
// choice (four-step split or generic LDS kernel; rows / columns per
// workgroup; tridiagonal block size) of each pass a case's step uses, on
// the case's own scratch buffers at its real shape, and records the fastest
// in the process-wide table.  Choices are keyed by the grid only, so every
// case of that grid (alone or batched) then launches the same kernels.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "case.hpp"
#include "pressure.hpp"

namespace lg {

namespace {
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, PassChoice> g_tune;
}  // namespace

bool tuned_choice(int kind, int tsize, int N, int other, PassChoice& out) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(std::make_tuple(kind, tsize, N, other));
    if (it == g_tune.end()) return false;
    out = it->second;
    return true;
}

void set_tuned_choice(int kind, int tsize, int N, int other, const PassChoice& c) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[std::make_tuple(kind, tsize, N, other)] = c;
}

void clear_tuned_choices() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
}

int tuned_count() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return (int)g_tune.size();
}

bool tuned_entry(int i, int* key, PassChoice& c) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (i < 0 || i >= (int)g_tune.size()) return false;
    auto it = g_tune.begin();
    std::advance(it, i);
    key[0] = std::get<0>(it->first);
    key[1] = std::get<1>(it->first);
    key[2] = std::get<2>(it->first);
    key[3] = std::get<3>(it->first);
    c = it->second;
    return true;
}

namespace {

// Candidate choices of one pass: generic kernel and every instantiated
// split, each with power-of-two rows / columns per workgroup.
template <typename T>
std::vector<PassChoice> candidates(int kind, int N, int other) {
    std::vector<PassChoice> out;
    if (kind == PK_TRIDAG) {
        for (int b : {64, 128, 256}) out.push_back(PassChoice{0, 0, b});
        return out;
    }
    std::vector<std::pair<int, int>> sp = pass_splits(N);
    sp.insert(sp.begin(), {0, 0});
    const bool rows = kind == PK_R2C || kind == PK_C2R;
    const bool cols = kind == PK_COLS_INV || kind == PK_COLS_PAD || kind == PK_COLS_FWD ||
                          kind == PK_COLS_UNPAD;
    for (auto& s : sp)
        for (int S : {1, 2, 4, 8, 12, 16, 24, 32, 48, 64})
            for (int v : {0, 1, 2, 3}) {
                if (v == 1 && !(rows && s.first)) continue;
                if (v >= 2 && !((cols || rows) && s.first)) continue;
                PassChoice c{s.first, s.second, S, v};
                if (choice_ok<T>(kind, c, N, other)) out.push_back(c);
            }
    return out;
}

// Median of reps timed launches (after one untimed) of launch(); a choice
// the runtime refuses to launch (resources) times as +inf.
template <class F>
float time_launch(hipStream_t s, int reps, F&& launch) {
    launch();
    if (hipGetLastError() != hipSuccess) {
        (void)hipStreamSynchronize(s);
        return 1e30f;
    }
    hipEvent_t a, b;
    LG_HIP(hipEventCreate(&a));
    LG_HIP(hipEventCreate(&b));
    std::vector<float> t;
    for (int r = 0; r < reps; ++r) {
        LG_HIP(hipEventRecord(a, s));
        launch();
        LG_HIP(hipEventRecord(b, s));
        LG_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        LG_HIP(hipEventElapsedTime(&ms, a, b));
        t.push_back(ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

template <typename T, class F>
PassChoice pick(int kind, int N, int other, hipStream_t s, int reps, F&& trial) {
    PassChoice best = default_choice<T>(kind, N, other);
    float tbest = time_launch(s, reps, [&] { trial(best); });
    for (const PassChoice& c : candidates<T>(kind, N, other)) {
        const float t = time_launch(s, reps, [&] { trial(c); });
        if (t < tbest) {
            tbest = t;
            best = c;
        }
    }
    set_tuned_choice(kind, (int)sizeof(T), N, other, best);
    return best;
}

}  // namespace

// Tune every pass of c's step; returns the number of passes tuned.  Trials
// write only scratch arrays (tmp, spec, big, bigspec, usc and the
// spectra of the pressure solve), never the case's fields.
template <typename T>
int autotune(Case<T>& c, int reps) {
    const Dims& d = c.d;
    KernelTimer quiet;
    const hipStream_t s = c.stream;
    const int Ma = c.Ma;
    const int nl = d.nz - 1;
    LG_HIP(hipStreamSynchronize(s));
    int n = 0;
    {
        const T* src[3] = {c.tmp[0], c.tmp[1], c.tmp[2]};
        cplx<T>* dst[3] = {c.spec[0], c.spec[1], c.spec[2]};
        pick<T>(PK_R2C, d.nx, d.ny, s, reps, [&](const PassChoice& ch) {
            launch_rows_r2c<T>(quiet, c.px.plan, 3, src, nullptr, d.plane, d.ld, 1, d.field, dst,
                               d.splane, 1, c.msp, d.ny, d.NK, nl, Ma, s, &ch);
        });
        const cplx<T>* csrc[3] = {c.spec[0], c.spec[1], c.spec[2]};
        T* cdst[3] = {c.tmp[0], c.tmp[1], c.tmp[2]};
        pick<T>(PK_C2R, d.nx, d.ny, s, reps, [&](const PassChoice& ch) {
            launch_rows_c2r<T>(quiet, c.px.plan, 3, csrc, d.splane, 1, c.msp, cdst, d.plane, d.ld,
                               1, d.field, d.ny, d.NK, nl, Ma, s, &ch);
        });
        n += 2;
    }
    {
        const T* src[3] = {c.big[0], c.big[1], c.big[2]};
        cplx<T>* dst[3] = {c.bigspec[0], c.bigspec[1], c.bigspec[2]};
        pick<T>(PK_R2C, d.nx2, d.ny2, s, reps, [&](const PassChoice& ch) {
            launch_rows_r2c<T>(quiet, c.pbx.plan, 3, src, nullptr, d.bplane, d.ld_big, 1, c.mbp,
                               dst, d.bsplane, 1, c.mbsp, d.ny2, d.NK, nl, Ma, s, &ch);
        });
        const cplx<T>* csrc[3] = {c.bigspec[0], c.bigspec[1], c.bigspec[2]};
        T* cdst[3] = {c.big[0], c.big[1], c.big[2]};
        pick<T>(PK_C2R, d.nx2, d.ny2, s, reps, [&](const PassChoice& ch) {
            launch_rows_c2r<T>(quiet, c.pbx.plan, 3, csrc, d.bsplane, 1, c.mbsp, cdst, d.bplane,
                               d.ld_big, 1, c.mbp, d.ny2, d.NK, nl, Ma, s, &ch);
        });
        n += 2;
    }
    auto cols = [&](int kind, const FftPlan& pl, int map, cplx<T>* const* src, long sp,
                    long sms, cplx<T>* const* dst, long dp, long dms, int ny_in) {
        pick<T>(kind, pl.N, d.NK, s, reps, [&](const PassChoice& ch) {
            ColArgs<T> a{};
            a.plan = pl;
            for (int i = 0; i < 3; ++i) {
                a.src[i] = src[i];
                a.dst[i] = dst[i];
                a.op[i] = OP_NONE;
            }
            a.src_plane = sp;
            a.dst_plane = dp;
            a.lev0_src = a.lev0_dst = 1;
            a.src_mstride = sms;
            a.dst_mstride = dms;
            a.map = map;
            a.zero_nyq = (kind == PK_COLS_FWD || kind == PK_COLS_UNPAD) ? 1 : 0;
            a.ncols = d.NK;
            a.ny = d.ny;
            a.G1 = c.G1;
            a.G2 = c.G2;
            if (kind == PK_COLS_FWD || kind == PK_COLS_UNPAD)
                launch_cols<T, -1>(quiet, a, 3, nl, Ma, ny_in, s, &ch);
            else
                launch_cols<T, +1>(quiet, a, 3, nl, Ma, ny_in, s, &ch);
        });
        ++n;
    };
    cplx<T>* sm[3] = {c.spec[0], c.spec[1], c.spec[2]};
    cplx<T>* sm2[3] = {c.spec[3], c.spec[4], c.spec[5]};
    cplx<T>* bg[3] = {c.bigspec[0], c.bigspec[1], c.bigspec[2]};
    cols(PK_COLS_FWD, c.py.plan, MAP_ID, sm, d.splane, c.msp, sm, d.splane, c.msp, d.ny);
    cols(PK_COLS_INV, c.py.plan, MAP_ID, sm, d.splane, c.msp, sm2, d.splane, c.msp, d.ny);
    cols(PK_COLS_PAD, c.pby.plan, MAP_PAD, sm, d.splane, c.msp, bg, d.bsplane, c.mbsp, d.ny);
    cols(PK_COLS_UNPAD, c.pby.plan, MAP_UNPAD, bg, d.bsplane, c.mbsp, sm, d.splane, c.msp,
         d.ny2);
    {
        const double pi = 3.1415926535897932384626433;
        TriGeo tg{d.ny, d.NK, d.nz, d.splane, c.msp, Ma, 2.0 * pi / d.prm.lx, 2.0 * pi / d.prm.ly,
                  d.dz, 1.0 / (d.dz * d.dz), 1.0 / d.dz};
        const int nsys = Ma * d.ny * d.NK;
        // the coefficients are the step's own (the pivots depend on them
        // only); zero right-hand sides
        LG_HIP(hipMemsetAsync(c.spec[0], 0, c.M * c.msp * sizeof(cplx<T>), s));
        pick<T>(PK_TRIDAG, d.ny, d.NK, s, reps, [&](const PassChoice& ch) {
            const int tb = ch.S;
            hipLaunchKernelGGL(k_tridag<T>, dim3((nsys + tb - 1) / tb), dim3(tb), 0, s,
                               (const cplx<T>*)c.spec[0], (const cplx<T>*)c.spec[0],
                               (const cplx<T>*)c.spec[0], (const cplx<T>*)c.spec[0],
                               (const cplx<T>*)(c.spec[0] + d.splane), c.spec[4], c.gam, c.usc,
                               tg, c.dev_err);
        });
        ++n;
    }
    LG_HIP(hipStreamSynchronize(s));
    LG_HIP(hipGetLastError());
    c.drop_graphs();
    return n;
}

template int autotune<double>(Case<double>&, int);
template int autotune<float>(Case<float>&, int);

}  // namespace lg
