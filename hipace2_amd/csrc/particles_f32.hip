// FP32 plasma push of the mixed-precision path (k_push_plasma<O, float>), in a translation
// unit of its own so that it is built with FMA contraction (-ffp-contract=fast): the FP64
// path's unit keeps every multiply and add separately rounded, as the reference CPU build
// computes them, and is not affected. The kernels are those of particles.hip, whose host
// code is left out here.
#define HPQ_F32_TU 1
#include "particles.hip"

namespace hpq {

void push_plasma_f32(hpq_ctx* c, const hpq_plasma* p, const double* store, double qmc, const BC& bc, bool temp_slice) {
    const Layout& L = c->L;
    dispatch_order(c->p.depos_order, [&](auto O) {
        hipLaunchKernelGGL((k_push_plasma<decltype(O)::value, float>), grid1d(p->n, 256), dim3(256), 0, c->stream,
                           c->g, *p, store, L.Psi, L.Ez, L.Bx, L.By, L.Bz, qmc, bc, int(temp_slice));
    });
}

}  // namespace hpq
