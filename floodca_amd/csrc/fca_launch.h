/*
 * fca_launch.h -- C-visible interface between the C host (hip_backend.c) and
 * the HIP translation unit (fca_kernels.hip) of one precision.
 */
#ifndef FCA_LAUNCH_H
#define FCA_LAUNCH_H

#include "fca_device.h"

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Output columns of one wavefront tile with cpl (1 or 2) columns per
 * work-item: 64*cpl lanes less 3 halo lanes a side. */
#define FCA_TILE_COLS(cpl) (64 * (cpl) - 6)
/* Whether this precision's library has the two-column walk (float only: in
 * double it would spill registers to scratch). */
#ifdef FCA_DOUBLE
#define FCA_CPL2_BUILT 0
#else
#define FCA_CPL2_BUILT 1
#endif

/* Device-resident state of one grid. */

typedef struct FcaDev {
    int nr, nc, K, kernel;
    real *z, *n, *C, *Rday, *cr, *h[2], *hmax, *tmax;
    uint8_t *mask;
    /* FCA_HIP_SPLIT scratch: faces laid out as in fca_types.h:121-125 */
    real *qx, *qy, *qd1, *qd2, *qo, *s;
    void *sv;               /* split: device StepVals of the current step */
    FcaDevCtrl *ctrl;       /* [2], indexed by launch parity */
    void *cand;             /* [3] time-step candidate keys, indexed by launch % 3 */
    FcaAccSlot *slots;      /* fused: one per workgroup; split: [0] */
    int nslots;
    FcaHipStepRec *rec;
    int rec_cap;
    FcaPhys phys;
    FcaCtlParams ctl;
    /* a/d as two fmas after a multiply by RN(1/d) is exact for d = dx, sqrt2*dx,
     * A: verified on the host for every float mantissa (div_const_exact) */
    int fdiv[3];
    /* the fused walk may take its LAND form (fca_land_form_ok and the
     * division flags it relies on) */
    int land_form;
    /* fused geometry; max_blocks caps the persistent grid (0 = one wave per tile),
     * tiles_per_wave > 1 sizes it to walk that many tiles per wavefront,
     * cols_per_lane (1 or 2) sets the tile width, FCA_TILE_COLS(cols_per_lane) */
    int strip_rows, ngroups, nstrips, ntiles, nblocks, max_blocks, tiles_per_wave;
    int cols_per_lane;
    /* quiet-tile skipping: flags [2][ntiles], rain-free tiles [ntiles], on/off */
    uint8_t *tflags, *tnorain;
    int skip;
    double *red;            /* scratch of the volume / ensemble reductions */
    int nred;
    void *arena;            /* one allocation holding every buffer above */
} FcaDev;

/* Workspace of fca_hip_run_noalloc for grids of up to
 * FCA_HIP_NOALLOC_MAX_CELLS cells: the second depth buffer and cr (one real per
 * cell each), plus room for the controller records, candidate keys, balance
 * slots, tile flags and reduction partials of the worst tile count such a
 * grid can have (strips of at least 16 rows: at most 2^18 tiles, 15 B each). */
#define FCA_NOALLOC_SMALL_BYTES ((size_t)4 << 20)
#define FCA_NOALLOC_SCRATCH_BYTES \
    ((size_t)2 * FCA_HIP_NOALLOC_MAX_CELLS * sizeof(real) + FCA_NOALLOC_SMALL_BYTES)
/* Partial sums of the volume reduction in the no-alloc path. */
#define FCA_NOALLOC_NRED 4096
/* Device address and size of this device's workspace (hipGetSymbolAddress
 * once per device). Returns 0, or -1 if the address cannot be had. */
int fca_k_noalloc_scratch(void **ptr, size_t *bytes);
/* hmax = h and tmax = tlast on the LAND cells whose depth h[hbuf] exceeds
 * their stored hmax (the fused walk stores hmax and tmax only when a cell
 * stops rising); the controller record's tmax_last is tlast. */
int fca_k_hmax_finalize(const FcaDev *d, int hbuf, double tlast, hipStream_t st);
/* Zeroes bytes (rounded up to a multiple of 4) at p with a fill kernel. */
int fca_k_zero(void *p, size_t bytes, hipStream_t st);
/* out[0] = nonzero if a LAND roughness leaves [1e-12, 2^100]; out[1] = open
 * faces of LAND edge cells (zeroed first). */
int fca_k_grid_scan(const FcaDev *d, unsigned long long *out, hipStream_t st);

/* Device table of n grids for the batched fused step (NULL on failure); free
 * with hipFree. */
void *fca_k_batch_table(const FcaDev *const *devs, int n);
/* One batched launch: nwg workgroups listed in wgs (device array of int2
 * {grid index in the table, workgroup within that grid}); launch index k
 * selects every grid's controller parity and candidate slots; every grid of
 * the table has cpl columns per work-item. */
int fca_k_step_batch(const void *tab, const void *wgs, int nwg, int K, int cpl, long long k,
                     hipStream_t st);
/* Fills the fused geometry of d from d->strip_rows. */
void fca_k_geometry(FcaDev *d);
/* Time-step candidates of the state in h[hbuf] into cand slot `slot`
 * (resets the slot first). */
int fca_k_cand_init(const FcaDev *d, int hbuf, int slot, hipStream_t st);
/* One step launch with launch index k (ctrl parity, candidate rotation). */
int fca_k_step(const FcaDev *d, long long k, hipStream_t st);
/* Rain-free tiles of the current cr (for quiet-tile skipping). */
int fca_k_tile_norain(const FcaDev *d, hipStream_t st);
/* cr = C*Rday on every cell. */
int fca_k_cr(const FcaDev *d, hipStream_t st);
/* Fills d's z, mask, n, C, Rday and a dry state from a window of a resident
 * base grid (bz, bmask, bcls with bnc columns), as crop describes. */
int fca_k_crop(const FcaDev *d, const real *bz, const uint8_t *bmask, const uint8_t *bcls,
               int bnc, const FcaHipCrop *crop, hipStream_t st);
/* Sets candidate slot `slot` to +inf. */
int fca_k_cand_reset(const FcaDev *d, int slot, hipStream_t st);
/* Per-block partial LAND volumes sum(h*A) of h[hbuf] into d->red[0..nblk). Returns nblk. */
int fca_k_volume(const FcaDev *d, int hbuf, hipStream_t st);
/* Ensemble count: counts[j*N+a] += (LAND && hmax >= thr[j]) (uint32, saturating
 * uint16 copy kept alongside), plus per-block flooded counts and max LAND hmax into
 * d->red. Returns the number of blocks. */
int fca_k_ens_count(const FcaDev *d, uint32_t *counts, int nthr, const double *thr_host,
                    hipStream_t st);

#ifdef __cplusplus
}
#endif

#endif
