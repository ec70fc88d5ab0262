// Optional per-kernel-class timing (used by bench.py for the roofline
// figure).  A timed launch goes through hipExtLaunchKernelGGL with a
// start/stop event pair, which the dispatch itself records, so the interval
// is the kernel's execution, not the queue gaps around it.
#pragma once

#include <hip/hip_ext.h>

#include <vector>

#include "common.hpp"

namespace lg {

// events for the next kernel launch made through LG_LAUNCH (none: untimed)
struct LaunchEvents {
    hipEvent_t a = nullptr, b = nullptr;
};
inline thread_local LaunchEvents g_launch_events;

#define LG_LAUNCH(K, G, B, L, S, ...)                                                         \
    do {                                                                                      \
        if (::lg::g_launch_events.a)                                                          \
            hipExtLaunchKernelGGL(K, G, B, L, S, ::lg::g_launch_events.a,                     \
                                  ::lg::g_launch_events.b, 0, __VA_ARGS__);                   \
        else                                                                                  \
            hipLaunchKernelGGL(K, G, B, L, S, __VA_ARGS__);                                   \
    } while (0)

// Timing classes: one per transform kernel family and grid, so a class
// holds launches of one kernel on one shape (the 3/2-grid passes and the
// Lagrangian model's row passes have classes of their own).
enum KernelKind : int {
    KK_ROWS_R2C = 0,          // forward rows, nx points
    KK_ROWS_C2R = 1,          // inverse rows, nx points
    KK_COLS_FWD = 2,          // forward columns, ny points
    KK_COLS_INV = 3,          // inverse columns, ny points
    KK_TRIDAG = 4,            // pressure tridiagonal
    KK_ROWS_R2C_PAD = 5,      // forward rows on the 3/2 grid
    KK_ROWS_C2R_PAD = 6,      // inverse rows on the 3/2 grid
    KK_COLS_UNPAD = 7,        // forward columns 3/2 grid -> small
    KK_COLS_PAD = 8,          // inverse columns small -> 3/2 grid
    KK_LAG_ROWS = 9,          // the Lagrangian model's inverse rows + contractions
    KK_LAG_PREP = 10,         // the Lagrangian model's forward rows of its quantities
    KK_ROWS_R2C_PROD = 11,    // forward rows on the 3/2 grid forming the convective products
    KK_NKINDS = 12,
};

struct KernelTimer {
    struct Rec {
        int kind;
        hipEvent_t a, b;
        double bytes;
    };
    bool on = false;
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t used = 0;

    hipEvent_t get() {
        if (used == pool.size()) {
            hipEvent_t e;
            LG_HIP(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }

    template <class F>
    void run(hipStream_t s, int kind, double bytes, F&& launch) {
        if (!on) {
            launch();
            return;
        }
        hipEvent_t a = get(), b = get();
        g_launch_events = LaunchEvents{a, b};
        launch();
        g_launch_events = LaunchEvents{};
        recs.push_back({kind, a, b, bytes});
    }

    void reset() {
        recs.clear();
        used = 0;
    }

    // total milliseconds, launch count and algorithmic bytes of one kind
    void report(int kind, double& ms, double& n, double& bytes) {
        ms = n = bytes = 0.0;
        for (auto& r : recs) {
            if (r.kind != kind) continue;
            float t = 0.f;
            LG_HIP(hipEventSynchronize(r.b));
            LG_HIP(hipEventElapsedTime(&t, r.a, r.b));
            ms += t;
            n += 1.0;
            bytes += r.bytes;
        }
    }

    ~KernelTimer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace lg
