/* A C99 translation unit against include/mpdugks.h, linked with
 * libmpdugks_hip.so: checks that the header is valid C and that the host-only
 * entry points work through it (no GPU is needed).  Built and run by
 * tests/test_abi_cpu.py. */
#include <stdio.h>
#include <stdlib.h>

#include "mpdugks.h"

int main(void)
{
    mpd_config cfg;
    mpd_handle *h = NULL;
    size_t n;
    double *th, *tf, *m;
    double mass = 0.0;
    size_t k;

    if (mpd_abi_version() != MPDUGKS_ABI_VERSION) return 1;
    if (mpd_default_config(MPD_MODEL_AC_HCZ, &cfg) != MPD_OK) return 2;
    if (cfg.nx != 128 || cfg.ny != 128 || cfg.model != MPD_MODEL_AC_HCZ) return 3;
    cfg.nx = 32;
    cfg.ny = 24;
    n = (size_t)cfg.nx * (size_t)cfg.ny;
    th = (double *)calloc(9 * n, sizeof(double));
    tf = (double *)calloc(9 * n, sizeof(double));
    m = (double *)calloc((size_t)MPD_M_COUNT * n, sizeof(double));
    if (!th || !tf || !m) return 4;
    if (mpd_init_ac_drop(&cfg, 16.0, 12.0, 8.0, -1, 1.0, 0.1, th, tf, m) != MPD_OK) return 5;
    for (k = 0; k < n; ++k) mass += m[MPD_M_PHI * n + k];
    /* an unsupported combination is refused with a message, before any device call */
    cfg.model = 99;
    if (mpd_create(&cfg, &h) != MPD_ERR_INVALID || h != NULL || mpd_last_error()[0] == '\0') return 6;
    printf("abi_smoke ok %.6f\n", mass);
    free(th);
    free(tf);
    free(m);
    return 0;
}
