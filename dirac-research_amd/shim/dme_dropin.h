// dme_dropin.h -- drop-in replacements of the reference encoder's motion
// estimation classes, for a maintainer who builds them into the reference
// tree (they include its headers and link libdme.so).
//
// Same names, constructors and method shapes as the reference's
// libdirac_motionest classes:
//   PixelMatcher(const EncoderParams&) + DoSearch(EncQueue&, int)   (pixel_match.h:71,81)
//   SubpelRefine(const EncoderParams&) + DoSubpel(EncQueue&, int)   (me_subpel.h:66,77)
//   ModeDecider(const EncoderParams&)  + DoModeDecn(EncQueue&, int) (me_mode_decn.h:79,93)
// Each call reads what the reference's class reads -- the picture and
// reference planes of the EncQueue, and the MEData state the caller owns
// (so the caller's own SetLambdaMap and pixel-precision doubling between
// the stages are honoured) -- runs the stage on the GPU through the C ABI
// (include/dme.h) and writes the MEData arrays in place (TwoDArray is
// contiguous, so the arrays are passed as plain pointers).  The reference's
// calls return void; failures here throw std::runtime_error with the
// library's message.  One GPU context per host thread, created on first use
// on device 0 (DME_DROPIN_DEVICE selects another).
#ifndef DME_DROPIN_H
#define DME_DROPIN_H

#include <libdirac_common/common.h>
#include <libdirac_encoder/enc_queue.h>

namespace dme_dropin {

class PixelMatcher {
public:
    explicit PixelMatcher(const dirac::EncoderParams &encp) : m_encparams(encp) {}
    void DoSearch(dirac::EncQueue &my_buffer, int pic_num);

private:
    const dirac::EncoderParams &m_encparams;
};

class SubpelRefine {
public:
    explicit SubpelRefine(const dirac::EncoderParams &encp) : m_encparams(encp) {}
    void DoSubpel(dirac::EncQueue &my_buffer, int pic_num);

private:
    const dirac::EncoderParams &m_encparams;
};

class ModeDecider {
public:
    explicit ModeDecider(const dirac::EncoderParams &encp) : m_encparams(encp) {}
    void DoModeDecn(dirac::EncQueue &my_buffer, int pic_num);

private:
    const dirac::EncoderParams &m_encparams;
};

}  // namespace dme_dropin

#endif
