/*
 * cpu_ref.h -- the oracle: a CPU restatement of the Dirac reference encoder's
 * live motion-estimation path, used ONLY as test infrastructure (parity
 * checker in tests/, smoke() and the cpu_baseline leg of bench.py).  The
 * product library never links, loads or calls anything under oracle/.
 *
 * Every function in cpu_ref.cpp cites the reference file:line it restates.
 * Struct layouts come from the public header include/dme.h.
 */
#ifndef DME_ORACLE_CPU_REF_H
#define DME_ORACLE_CPU_REF_H

#include "../include/dme.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One picture handed to the oracle: luma plus (optional) 4:2:0 chroma. */
typedef struct oracle_picture {
    int32_t pic_num;
    int32_t width, height;
    const int16_t *y, *u, *v;
} oracle_picture;

/* Optional capture of intermediate pel-search state (per pyramid level and
 * reference) and of the pyramid pictures, for per-stage parity pins.  Any
 * pointer may be NULL.  Level l arrays have geometry level_xblocks[l] x
 * level_yblocks[l] (level 0 = the picture's own MEData grid).               */
typedef struct oracle_capture {
    dme_mv   *level_vectors[2][DME_MAX_DEPTH + 1];
    dme_cost *level_costs[2][DME_MAX_DEPTH + 1];
    int16_t  *cur_pyramid[DME_MAX_DEPTH + 1];   /* level l>=1, W_l*H_l      */
    int16_t  *ref_up[2];                         /* 2W x 2H up-converted ref */
} oracle_capture;

/* Stage selectors for oracle_me(stop_after). */
#define ORACLE_STOP_PEL     1
#define ORACLE_STOP_LAMBDA  2
#define ORACLE_STOP_SUBPEL  3
#define ORACLE_STOP_MODE    4

int oracle_geometry(const dme_params *p, int width, int height,
                    dme_geometry *out);

/* DownConverter::DoDownConvert on an in_w x in_h plane into out_w x out_h. */
int oracle_downconvert(const int16_t *in, int in_w, int in_h,
                       int16_t *out, int out_w, int out_h);

/* UpConverter::DoUpConverter(-128, 127) on a w x h plane into 2w x 2h. */
int oracle_upconvert(const int16_t *in, int w, int h, int16_t *out);

/* EncPicture::Combine (enc_picture.cpp:285-367), 4:2:0: the combined-
 * component picture of w x h luma and w/2 x h/2 chroma planes.              */
int oracle_combine(const int16_t *y, const int16_t *u, const int16_t *v, int w, int h, int16_t *out);

/* MEData::SetLambdaMap(num_refs, lambda) (host long double statistics). */
int oracle_lambda_map(int xnb, int ynb, int num_refs, const dme_mv *v1,
                      const dme_mv *v2, float lambda, float *out);

/* The live ME sequence of PictureCompressor for one inter picture:
 * PixelME -> SetLambdaMap -> SubPixelME -> ModeDecisionME, stopping after
 * stage stop_after.  If lambda_override is non-NULL the lambda map is taken
 * from it instead of being computed.  sad_terms (may be NULL) receives the
 * algorithmic SAD-term count of the stages run.                              */
int oracle_me(const dme_params *p, const oracle_picture *cur,
              const oracle_picture *refs, int nrefs, int stop_after,
              const float *lambda_override, dme_medata *out,
              oracle_capture *cap, double *sad_terms);

/* MotionCompensator::CompensatePicture (mot_comp.cpp:63-600) for one
 * 4:2:0 picture of width x height: planes[c] (row length widths[c], rows
 * heights[c], at least the component's size) are compensated in place in
 * `direction` (DME_MC_SUBTRACT / DME_MC_ADD) from the vectors, modes, DCs
 * and superblock splits of `me` and the reference pictures' planes, with
 * the block parameters and precision of p and default reference weights. */
int oracle_motion_compensate(const dme_params *p, int width, int height,
                             const oracle_picture *refs, int nrefs,
                             const dme_medata *me, int direction,
                             int16_t *const planes[3], const int widths[3],
                             const int heights[3]);
/* The fast form of the learned path (the build's own algorithm): from the
 * level-0 integer vectors v0[r] (blocks with no samples hold (0, 0)), the
 * lambda map, sub-pel refinement and mode decision with predictors that do
 * not depend on decisions of the same pass, then the reference Lagrangian
 * of the final vectors with the reference's predictors, and SetDC.
 * v2_out[r] (may be NULL) receive the pre-merge sub-pel vectors.          */
int oracle_fast_me(const dme_params *p, const oracle_picture *cur,
                   const oracle_picture *refs, int nrefs,
                   const dme_mv *const *v0, dme_medata *out,
                   dme_mv *const *v2_out, double *sad_terms);
const char *oracle_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
