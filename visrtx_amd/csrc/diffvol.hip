// diffvol.hip -- differentiable emission-absorption volume integral.
#include "vrx_internal.h"

extern "C" {
size_t vrx_diffvol_ckpt_floats(const vrx_diffvol *v, uint32_t n) { return 0; }
int vrx_diffvol_forward(const vrx_diffvol *v, uint32_t n, const float *d_org, const float *d_dir,
    float *d_rgba_out, float *d_ckpt, vrx_stream s)
{
  vrx::set_error("not implemented");
  return VRX_ERR_ARG;
}
int vrx_diffvol_backward(const vrx_diffvol *v, uint32_t n, const float *d_org, const float *d_dir,
    const float *d_rgba, const float *d_grad_rgba, const float *d_ckpt, float *d_grad_density,
    float *d_grad_lut, vrx_stream s)
{
  vrx::set_error("not implemented");
  return VRX_ERR_ARG;
}
int vrx_make_rays(const vrx_camera *cams, uint32_t nviews, uint32_t w, uint32_t h, float *d_org,
    float *d_dir, vrx_stream s)
{
  vrx::set_error("not implemented");
  return VRX_ERR_ARG;
}
}
