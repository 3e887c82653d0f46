// batch.hip -- batched multi-view rendering.
#include "vrx_internal.h"

extern "C" {
size_t vrx_batch_scratch_bytes(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b)
{
  return 0;
}
int vrx_render_batch(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b, vrx_stream s)
{
  vrx::set_error("not implemented");
  return VRX_ERR_ARG;
}
}
