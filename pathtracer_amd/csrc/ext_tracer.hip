// ext_tracer.hip — PT_MODE_EXT (extended physically based mode).
#include "tracer.hpp"

namespace ptk {

int ext_scene_build(pt_ctx* ctx, pt_scene*) {
    ctx->err = "extended mode is not built yet";
    return PT_ERR_STATE;
}
void ext_scene_release(pt_ctx*, pt_scene*) {}
int ext_dispatch(pt_ctx* ctx, const pt_scene*, pt_accum*, const pt_push_constants*) {
    ctx->err = "extended mode is not built yet";
    return PT_ERR_STATE;
}
void ext_wave_release(pt_ctx*) {}

}  // namespace ptk
