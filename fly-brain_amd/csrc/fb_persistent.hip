// fb_persistent.hip -- placeholder until the lookahead engine lands.
#include "fb_internal.h"

namespace fb {

fb_status persistent_supported(const DevBrain &d, bool serialDeliver, std::string *why) {
    (void)d; (void)serialDeliver;
    *why = "not built yet";
    return FB_ERR_UNSUPPORTED;
}

fb_status launch_persistent(const DevBrain &d, int32_t head, int32_t nsteps, hipStream_t st) {
    (void)d; (void)head; (void)nsteps; (void)st;
    set_error("persistent engine not built");
    return FB_ERR_UNSUPPORTED;
}

}  // namespace fb
