// Fused one-kernel-per-step path: derived state, tile kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpdugks.h"
#include "device_state.h"

namespace mpd {

struct Consts;

struct FusedState {
    bool ready = false;
    int64_t bytes = 0;
    void *mem[16] = {nullptr};
    int nmem = 0;
    // ping-pong state planes (element type depends on precision)
    void *st[2] = {nullptr, nullptr};
    int cur = 0;
    int nplanes = 0;
};

bool fused_supported(const mpd_config &cfg);
int fused_alloc(FusedState &F, const DevState &S, const Consts &C, const mpd_config &cfg);
void fused_free(FusedState &F);
int64_t fused_bytes(const FusedState &F);
// first != 0: take the explicit macros of S as the state of this step
void fused_step(FusedState &F, const DevState &S, const Consts &C, const mpd_config &cfg, int first,
                hipStream_t st);
// write the full macro set (and tilde planes) of the current fused state into S
int fused_materialize(FusedState &F, const DevState &S, const Consts &C, const mpd_config &cfg,
                      hipStream_t st);

}  // namespace mpd
