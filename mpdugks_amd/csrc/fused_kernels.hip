// Fused step kernels (one launch per time step).
#include "fused.h"
#include "physics.h"

namespace mpd {

bool fused_supported(const mpd_config &cfg) { (void)cfg; return false; }
int fused_alloc(FusedState &F, const DevState &, const Consts &, const mpd_config &) { F.ready = false; return -1; }
void fused_free(FusedState &F) { for (int k = 0; k < F.nmem; ++k) (void)hipFree(F.mem[k]); F.nmem = 0; F.ready = false; }
int64_t fused_bytes(const FusedState &F) { return F.bytes; }
void fused_step(FusedState &, const DevState &, const Consts &, const mpd_config &, int, hipStream_t) {}
int fused_materialize(FusedState &, const DevState &, const Consts &, const mpd_config &, hipStream_t) { return 0; }

}  // namespace mpd
