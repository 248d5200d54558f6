// Multilevel foreground estimation (placeholder until the sweep kernels land).
#include "common.h"
#include "internal.h"

extern "C" int hm_fb_ml(const void*, const void*, int, int, int, const int*, const int*, double,
                        int, int, int, double, int, void*, void*, void*) {
    hm::set_error("hm_fb_ml is not implemented yet");
    return HM_ERR_UNSUPPORTED;
}
