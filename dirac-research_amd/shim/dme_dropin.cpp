// dme_dropin.cpp -- the drop-in classes of dme_dropin.h over the C ABI.
#include "dme_dropin.h"

#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include <libdirac_common/motion.h>
#include <libdirac_encoder/enc_picture.h>

#include "../../include/dme.h"

using namespace dirac;

namespace dme_dropin {
namespace {

// The C ABI's MEData records are the reference's own layouts (motion.h).
static_assert(sizeof(MVector) == sizeof(dme_mv), "MVector is {int x, y}");
static_assert(sizeof(MvCostData) == sizeof(dme_cost), "MvCostData is {float SAD, mvcost, total}");
static_assert(sizeof(PredMode) == sizeof(int32_t), "PredMode is an int-sized enum");
static_assert(sizeof(ValueType) == sizeof(int16_t), "ValueType is short");

void check(int rc, const char *what) {
    if (rc != DME_OK) throw std::runtime_error(std::string(what) + ": " + dme_last_error());
}

dme_ctx *context() {
    static thread_local struct Holder {
        dme_ctx *c = nullptr;
        ~Holder() {
            if (c) dme_destroy(c);
        }
    } h;
    if (!h.c) {
        const char *dev = std::getenv("DME_DROPIN_DEVICE");
        check(dme_create(dev ? std::atoi(dev) : 0, &h.c), "dme_create");
    }
    return h.c;
}

// The estimation parameters of picture pic_num: its MEData's block
// parameters (PicturePredParams, set by SetBlockSizes) and current vector
// precision, the encoder's search options and the build's float semantics.
void set_params(const EncoderParams &encp, const MEData &me) {
    const PicturePredParams &pp = me.GetPicPredParams();
    const OLBParams &b = pp.LumaBParams(2);
    dme_params p = {};
    p.xblen = b.Xblen();
    p.yblen = b.Yblen();
    p.xbsep = b.Xbsep();
    p.ybsep = b.Ybsep();
    p.mv_precision = (int32_t)pp.MVPrecision();
    p.depth = -1;  // PixelMatcher::DoSearch's own (pixel_match.cpp:94-95)
    p.lambda = encp.L2MELambda();
#ifdef HAVE_MMX
    p.fp_semantics = DME_FP_MMX;
#else
    p.fp_semantics = DME_FP_C;
#endif
    p.chroma_format = 0;
    p.full_search = encp.FullSearch() ? 1 : 0;
    if (p.full_search) {
        p.x_range_me = encp.XRangeME();
        p.y_range_me = encp.YRangeME();
    }
    p.combined_me = encp.CombinedME() ? 1 : 0;
    check(dme_set_params(context(), &p), "dme_set_params");
}

const int16_t *plane(const PicArray &a) { return reinterpret_cast<const int16_t *>(&a[0][0]); }

void upload(EncQueue &q, int n) {
    const EncPicture &P = q.GetPicture(n);
    const PicArray &y = P.DataForME(false);
    const PicArray &u = P.OrigData(U_COMP), &v = P.OrigData(V_COMP);
    check(dme_upload_picture(context(), n, y.LengthX(), y.LengthY(), plane(y), plane(u), plane(v)),
          "dme_upload_picture");
}

std::vector<int> refs_of(EncQueue &q, int pic_num) {
    return q.GetPicture(pic_num).GetPparams().Refs();
}

// Pointers into the reference's MEData arrays.
dme_medata view(MEData &me, int nrefs, bool all) {
    dme_medata m = {};
    for (int r = 0; r < nrefs; ++r) {
        m.vectors[r] = reinterpret_cast<dme_mv *>(&me.Vectors(r + 1)[0][0]);
        m.pred_costs[r] = reinterpret_cast<dme_cost *>(&me.PredCosts(r + 1)[0][0]);
    }
    if (all) {
        m.mode = reinterpret_cast<int32_t *>(&me.Mode()[0][0]);
        m.intra_costs = &me.IntraCosts()[0][0];
        m.bipred_costs = reinterpret_cast<dme_cost *>(&me.BiPredCosts()[0][0]);
        m.dc[0] = reinterpret_cast<int16_t *>(&me.DC(Y_COMP)[0][0]);
        m.dc[1] = reinterpret_cast<int16_t *>(&me.DC(U_COMP)[0][0]);
        m.dc[2] = reinterpret_cast<int16_t *>(&me.DC(V_COMP)[0][0]);
        m.sb_split = &me.SBSplit()[0][0];
        m.sb_costs = &me.SBCosts()[0][0];
    }
    return m;
}

// The MEData state a stage reads, handed back to the device.
void load(EncQueue &q, int pic_num, MEData &me, int stage) {
    const std::vector<int> refs = refs_of(q, pic_num);
    dme_medata in = view(me, (int)refs.size(), false);
    in.lambda_map = const_cast<float *>(&me.LambdaMap()[0][0]);
    check(dme_load_medata(context(), pic_num, refs.data(), (int)refs.size(), &in, stage), "dme_load_medata");
}

}  // namespace

void PixelMatcher::DoSearch(EncQueue &my_buffer, int pic_num) {
    MEData &me = my_buffer.GetPicture(pic_num).GetMEData();
    set_params(m_encparams, me);
    const std::vector<int> refs = refs_of(my_buffer, pic_num);
    upload(my_buffer, pic_num);
    for (int r : refs) upload(my_buffer, r);
    dme_medata out = view(me, (int)refs.size(), false);
    check(dme_pixel_search(context(), pic_num, refs.data(), (int)refs.size(), &out), "dme_pixel_search");
}

void SubpelRefine::DoSubpel(EncQueue &my_buffer, int pic_num) {
    MEData &me = my_buffer.GetPicture(pic_num).GetMEData();
    set_params(m_encparams, me);
    load(my_buffer, pic_num, me, 2);
    dme_medata out = view(me, (int)refs_of(my_buffer, pic_num).size(), false);
    check(dme_subpel(context(), pic_num, &out), "dme_subpel");
}

void ModeDecider::DoModeDecn(EncQueue &my_buffer, int pic_num) {
    MEData &me = my_buffer.GetPicture(pic_num).GetMEData();
    set_params(m_encparams, me);
    load(my_buffer, pic_num, me, 3);
    dme_medata out = view(me, (int)refs_of(my_buffer, pic_num).size(), true);
    check(dme_mode_decide(context(), pic_num, &out), "dme_mode_decide");
}

}  // namespace dme_dropin
