// nbx_mlp.hip - learned Barnes-Hut residual correction (placeholder: plain BH).
#include "nbx_internal.h"

namespace nbx {

NBodySimProStatus bh_learned(Engine* e, const NBodySimProGravity* g, NBodySimProError* error) {
    return barnes_hut(e, g, false, error);
}

}  // namespace nbx

extern "C" void nbx_mlp_release(void* mlp) { (void)mlp; }
