// dme_internal.h -- shared declarations of the MI355X motion-estimation
// engine (host and device).  Not part of the public ABI (include/dme.h).
#ifndef DME_INTERNAL_H
#define DME_INTERNAL_H

#include <cstdint>
#include <string>

#include "../../include/dme.h"

namespace dme {

// Records msg as the calling thread's last error and returns code.
int set_error(int code, const std::string &msg);

// Inside a catch handler: records the exception being handled as the last
// error and returns its status (DME_ERR_NOMEM for an allocation failure), so
// that no C++ exception leaves a C entry point.
int exception_status() noexcept;

// Geometry with everything the kernels need (host-computed).
struct Geometry {
    int W, H, CW, CH;
    int xsb, ysb, xnb, ynb, depth;
    int lw[DME_MAX_DEPTH + 1], lh[DME_MAX_DEPTH + 1];
    int lxb[DME_MAX_DEPTH + 1], lyb[DME_MAX_DEPTH + 1];
    int blen[3][2], bsep[3][2], boff[3][2];   // luma, levels 0..2, [x,y]
    int cblen[2], cbsep[2], cboff[2];         // chroma level 2
    float level_factor[3];
};

// Validates params for a width x height picture and fills g.
int make_geometry(const dme_params &p, int width, int height, Geometry &g);

// The launch limits (LDS of the table and wavefront kernels) of one job's
// stages: "" if they hold, else the limit that does not (dme_kernels.hip).
std::string launch_limits(const Geometry &g, bool full_search, bool learned, bool combined);

// Host lambda maps: level 2 from the pel-search vectors (MEData::
// SetLambdaMap, motion.cpp:128-171), levels 1 and 0 by max-pooling with the
// ModeDecider weights (motion.cpp:173-199, me_mode_decn.cpp:125-126).
void lambda_map_l2(int X, int Y, int nrefs, const dme_mv *v1, const dme_mv *v2, float lambda, float *out);
void lambda_map_level(const float *l2, int X2, int level, float wt, int Xl, int Yl, float *out);

}  // namespace dme

#endif
