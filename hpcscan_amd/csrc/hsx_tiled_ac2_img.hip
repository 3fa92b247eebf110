// The fused Ac2 step with the adjoint's imaging condition accumulated from
// its output (MODE_AC2_IMG), in a translation unit of its own.
#include "hsx_tiled.h"

namespace hsx {

thread_local ImgReq t_img = {{nullptr, nullptr, nullptr}, nullptr, 0.0, {0, 0, 0, 0, 0, 0}};

int tiled_ac2_step_img(const hsx_grid *g, void *prn, const void *prc, const void *coef, double *grad,
                       const void *q0, const void *q1, const void *q2, const int64_t box[6], double scale,
                       hipStream_t s)
{
    if (((uintptr_t)q0 | (uintptr_t)q1 | (uintptr_t)q2 | (uintptr_t)grad) % 16)
        return fail(HSX_ERR_UNSUPPORTED, "imaging buffers not 16-byte aligned");
    for (int a = 0; a < 3; a++)
        if (box[2 * a] < g->inner[2 * a] || box[2 * a + 1] > g->inner[2 * a + 1])
            return fail(HSX_ERR_INVALID, "imaging box outside the inner points");
    t_img.q[0] = q0;
    t_img.q[1] = q1;
    t_img.q[2] = q2;
    t_img.grad = grad;
    t_img.scale = scale;
    for (int k = 0; k < 6; k++) t_img.box[k] = box[k];
    const int rc = run_tiled<MODE_AC2_IMG>(g, prn, prn, prc, coef, s);
    t_img.grad = nullptr;
    return rc;
}

} // namespace hsx
