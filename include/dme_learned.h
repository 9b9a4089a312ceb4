/* dme_learned.h -- learned-proposal motion estimation on the MI355X.
 *
 * No function here has a counterpart in the reference encoder: this is the
 * build's own addition beside the bit-exact path of dme.h.  A 4-layer 3x3
 * CNN with 32 channels (BF16 operands on MFMA, FP32 accumulation) runs on
 * the level-0 block grid.  Its inputs are the plain search's level-2
 * (quarter resolution) vectors and costs of both references and the
 * picture's texture; its outputs are four candidate vectors per block and
 * reference (offsets from 4 x the level-2 vector) and a predicted SAD.
 * Exact integer SADs over +-2 pel windows around the candidates choose the
 * level-0 pel vector; where that SAD exceeds 1.2 x the predicted one, a wider
 * exact search (+-4 around the level-2 guide, +-2 around zero) joins in.
 * Sub-pel refinement and mode decision follow in one of two forms
 * (dme_lp_set_form).
 *
 * Tensors passed to the layer entry points are device pointers, NHWC over
 * the block grid with 32 channels; activations and gradients fed to MFMA are
 * BF16 (uint16_t), everything else f32.  All work is queued on the context's
 * stream; dme_synchronize waits for it.
 */
#ifndef DME_LEARNED_H
#define DME_LEARNED_H

#include <stddef.h>
#include <stdint.h>

#include "dme.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DME_LP_CHANNELS 32
/* one layer: w[9][32][32] (tap = (dy+1)*3 + dx+1, input channel, output
 * channel) then bias[32]; the network is four layers in order */
#define DME_LP_LAYER_PARAMS (9 * 32 * 32 + 32)
#define DME_LP_PARAMS (4 * DME_LP_LAYER_PARAMS)
/* BF16 B-fragments of one layer's weights (per direction), in uint16_t */
#define DME_LP_FRAG_ELEMS (9 * 2 * 64 * 8)

/* Loads the network (host array of DME_LP_PARAMS floats).  Required before
 * the learned entry points below. */
int dme_lp_set_weights(dme_ctx *ctx, const float *params, size_t count);

/* Motion estimation of one picture with learned proposals for the pel
 * stage (pyramid depth >= 2).  Same outputs and layouts as
 * dme_motion_estimate. */
int dme_motion_estimate_learned(dme_ctx *ctx, int pic_num, const int *refs, int nrefs, dme_medata *out);

/* Batched form of dme_motion_estimate_learned (as dme_run_batch). */
int dme_run_batch_learned(dme_ctx *ctx, const dme_job *jobs, int njobs, dme_medata *outs);

/* The form of the learned entry points.
 *   DME_LP_FORM_FAST (the default): the network's level-0 vectors, then
 *     sub-pel refinement and mode decision with every block and superblock
 *     decided at once (the predictors come from the integer vectors and from
 *     the pre-merge sub-pel vectors), then the reference's Lagrangian of the
 *     final vectors with the reference's own predictors (SubpelRefine's
 *     GetPred, me_subpel.cpp:145-176; ModeDecider::DoME's predictor in
 *     DoModeDecn's raster order, me_mode_decn.cpp:308-321): the PredCosts,
 *     SBCosts and modes it writes are those the reference assigns to these
 *     vectors.  Half or quarter pel, MMX float semantics.
 *   DME_LP_FORM_EXACT_STAGES: the network's level-0 vectors, then the plain
 *     path's sub-pel and mode-decision stages unchanged. */
#define DME_LP_FORM_FAST          0
#define DME_LP_FORM_EXACT_STAGES  1
int dme_lp_set_form(dme_ctx *ctx, int form);

/* After a fast-form estimate of pic_num: its level-0 integer vectors (which
 * 0) or its sub-pel vectors before mode decision merged them (which 1), one
 * array of ynum_blocks * xnum_blocks per reference (NULL skips). */
int dme_lp_get_vectors(dme_ctx *ctx, int pic_num, int which, dme_mv *out_ref1, dme_mv *out_ref2);

/* After a fast-form estimate of pic_num: the network's outputs for its
 * blocks, ynum_blocks * xnum_blocks rows of 32 floats (per reference r and
 * hypothesis h the offset at 8r + 2h, 8r + 2h + 1 in units of 8 pel; the
 * predicted log(1 + SAD) of reference r at 16 + r).  Valid until the next
 * learned call of the context; DME_ERR_STATE after that. */
int dme_lp_get_outputs(dme_ctx *ctx, int pic_num, float *out);

/* Blocks refined and blocks that took the fallback search since the
 * context was created. */
int dme_lp_get_stats(dme_ctx *ctx, long long *blocks, long long *fallbacks);

/* Training pair of a picture whose plain pixel search (dme_pixel_search or
 * dme_motion_estimate) has run: network inputs d_features[xnb*ynb][32]
 * (BF16) and labels d_labels[xnb*ynb][8] (per reference: target offset of
 * the plain level-0 vector from 4 x its level-2 vector in pel (x, y),
 * log(1 + its SAD), valid). */
int dme_lp_training_pair(dme_ctx *ctx, int pic_num, uint16_t *d_features, float *d_labels);

/* ---- layer kernels (training and tests) ---------------------------------- */

/* BF16 B-fragments of one layer's weights d_w[9][32][32] for the forward
 * convolution and for its data gradient. */
int dme_lp_prepare(dme_ctx *ctx, const float *d_w, uint16_t *d_frag_fwd, uint16_t *d_frag_bwd);

/* out[p][co] = bias[co] + sum over taps and ci of in[p + tap][ci] * w, zero
 * padding, over n grids of h x w positions; ReLU if relu; d_out (BF16)
 * and/or d_out_f32 may be NULL; d_bias may be NULL (no bias).  With the
 * backward fragments this is the data gradient. */
int dme_lp_conv(dme_ctx *ctx, const uint16_t *d_in, int n, int h, int w, const uint16_t *d_frag, const float *d_bias,
                int relu, uint16_t *d_out, float *d_out_f32);

/* d_dw[9][32][32] += weight gradient, d_db[32] += bias gradient, from the
 * layer input d_in and the pre-activation gradient d_g (BF16). */
int dme_lp_wgrad(dme_ctx *ctx, const uint16_t *d_in, const uint16_t *d_g, int n, int h, int w, float *d_dw,
                 float *d_db);

/* d_out = d_act > 0 ? bf16(d_g) : 0 elementwise (d_act NULL: no mask). */
int dme_lp_relu_backward(dme_ctx *ctx, const uint16_t *d_act, const float *d_g, size_t count, uint16_t *d_out);

/* Winner-takes-all proposal loss of the last layer's f32 outputs against
 * d_labels (layout of dme_lp_training_pair); d_dout gets the gradient times
 * scale, *d_loss (device) accumulates the loss times scale. */
int dme_lp_wta_loss(dme_ctx *ctx, const float *d_out, const float *d_labels, int npos, float alpha, float beta,
                    float scale, float *d_dout, float *d_loss);

/* Adam step t (1-based) over count parameters. */
int dme_lp_adam(dme_ctx *ctx, float *d_param, const float *d_grad, float *d_m, float *d_v, int count, float lr,
                float beta1, float beta2, float eps, int t);

/* Zeroes bytes (a multiple of 4) of device memory on the context's stream. */
int dme_lp_zero(dme_ctx *ctx, void *d_ptr, size_t bytes);

/* Device memory for training tensors on the context's device, so that a
 * trainer needs no other GPU runtime in the process: allocate, free, and
 * copy (kind 1 host to device, 2 device to host, 3 device to device; the
 * copy is complete on return). */
int dme_lp_alloc(dme_ctx *ctx, size_t bytes, void **d_ptr);
int dme_lp_free(dme_ctx *ctx, void *d_ptr);
int dme_lp_copy(dme_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);

#ifdef __cplusplus
}
#endif

#endif /* DME_LEARNED_H */
