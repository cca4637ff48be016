// sp_handles.hpp -- the state behind the C ABI's opaque handles (contexts and
// propagators), shared by the translation units that implement the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <utility>
#include <vector>

#include "sp_internal.hpp"

// Handles are reference counted: a propagator holds its context, a Parareal driver holds
// its propagator, so destroying them in any order (as garbage collectors do) is safe.
struct sp_ctx {
    int refs = 1;
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    double* red = nullptr;       // device partials
    double* red_host = nullptr;  // pinned host copy
    int red_cap = 0;             // doubles
};

struct Buf {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= cap) return SP_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) return sp::fail(SP_ERR_NOMEM, "hipMalloc of scratch failed");
        cap = bytes;
        return SP_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct sp_prop {
    int refs = 1;
    sp_ctx* ctx = nullptr;
    int ni = 0, nj = 0, nk = 0;
    double dx = 0, nu0 = 0, nufreq = 0, cx = 0, cy = 0, cz = 0;
    int upwind = 1;
    int path = SP_PATH_FUSED;
    sp::FusedTile tile{0, 0, 0, -1, SP_KERNEL_AUTO};  // tx == 0: automatic
    Buf scratch[2];                // ping-pong for the fused steppers
    Buf split;                     // work field of SP_KERNEL_SPLIT steps (FP64, compact slices)
    Buf qI, k1, k2, k3, k4;         // staged path temporaries (f64, compact)
    Buf nu_tab;                     // per-step nu table
    std::vector<double> nu_host;
    std::map<std::pair<int, int>, sp::FusedTile> tuned;  // (slices, FP64 compute) -> autotuned tile
};
