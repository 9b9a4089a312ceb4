// dme_device.h -- device-side job descriptors shared by the kernels
// (dme_kernels.hip) and the host orchestration (dme_engine.hip).
#ifndef DME_DEVICE_H
#define DME_DEVICE_H

#include <cstdint>

#include "../../include/dme.h"

namespace dme {

// Everything one picture's motion estimation needs on the device.  Built on
// the host, copied to device memory once per call, read by every kernel.
struct JobDesc {
    // geometry (host-validated)
    int W, H, CW, CH;
    int xnb, ynb, xsb, ysb, depth, nrefs;
    int prec;      // MVPrecision of the refinement (0, 1, 2)
    int step;      // sampling precision of mode decision (1 half, 2 quarter)
    int fp_sem;    // DME_FP_MMX or DME_FP_C
    int lw[DME_MAX_DEPTH + 1], lh[DME_MAX_DEPTH + 1];
    int lxb[DME_MAX_DEPTH + 1], lyb[DME_MAX_DEPTH + 1];
    int blen[3][2], bsep[3][2], boff[3][2];
    int cblen[2], cbsep[2], cboff[2];
    float lf[3];
    // planes
    const int16_t *pic[DME_MAX_DEPTH + 1];      // [0] = Y of the picture
    const int16_t *ref[2][DME_MAX_DEPTH + 1];   // reference pyramids
    const int16_t *up[2];                       // 2W x 2H up-converted refs
    const int16_t *cu, *cv;                     // chroma of the picture
    // pel search, per reference and level ([0] aliases the MEData arrays)
    dme_mv *lvec[2][DME_MAX_DEPTH + 1];
    dme_cost *lpc[2][DME_MAX_DEPTH + 1];
    int *bsad[2];                               // zero/guide-window best SAD
    dme_mv *bmv[2];                             // ... and its vector
    int *ptab[2];                               // pel SAD tables (see k_pel_bulk)
    int *qtab[2];                               // sub-pel SAD tables (k_subpel_bulk)
    int *m1tab;                                 // level-1 unit SAD tables (k_mode_tables)
    int *m0tab;                                 // level-0 unit SAD tables
    // MEData of the picture
    dme_mv *vec[2];
    dme_cost *pc[2];
    int32_t *mode;
    float *intra;
    dme_cost *bi;
    int16_t *dc[3];
    int32_t *split;
    float *sbc;
    float *lam;
    // mode decision temporaries (levels 1 and 0)
    float *lam1, *lam0;
    dme_mv *v1[2], *v0[2];
    dme_cost *c1[2], *c0[2];
    float *ucost;                               // level-2 unit costs
    int16_t *dct[3];                            // block DCs before smoothing
};

// One unit of a batched launch: a (job, reference) pair covering `count`
// consecutive work items starting at `offset` of the launch's index space.
struct Task {
    int job, ref, count, offset;
};

}  // namespace dme

#endif
