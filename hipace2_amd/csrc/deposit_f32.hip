// FP32 plasma current and explicit depositions of the mixed-precision path
// (k_deposit_plasma<O, float>, k_explicit_deposit<O, T, float>), in a translation unit of their
// own so that they are built with FMA contraction (-ffp-contract=fast) while the FP64 path
// keeps every multiply and add separately rounded. The kernels are those of deposit.hip,
// whose host code is left out here.
#define HPQ_F32_TU 1
#include "deposit.hip"

namespace hpq {

void deposit_plasma_f32(hpq_ctx* c, const hpq_plasma* p, const int32_t* order, double* store, const int* cm,
                        double charge_invvol, double cmr) {
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid1d(p->n, DCHUNK), dim3(DT), 0, c->stream, c->g, *p, order, store, cm[0], cm[1],
                               cm[2], cm[3], cm[4], cm[5], charge_invvol, cmr, c->p.max_qsa, c->d_count);
        };
        if (c->p.mixed_sums == HPQ_SUM_KAHAN32) launch(k_deposit_plasma<o, float, HPQ_SUM_KAHAN32>);
        else launch(k_deposit_plasma<o, float, HPQ_SUM_FP64>);
    });
}

void explicit_deposit_f32(hpq_ctx* c, const hpq_plasma* p, const int32_t* order, double* store, double qimu0,
                          double qmr) {
    const Layout& L = c->L;
    const int T = c->p.derivative_type;
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid1d(p->n, DCHUNK_EXPLICIT), dim3(DT), 0, c->stream, c->g, *p, order, store,
                               L.Bz, L.Ez, L.ExmBy, L.EypBx, L.Sy, L.Sx, qimu0, qmr);
        };
        const bool k32 = c->p.mixed_sums == HPQ_SUM_KAHAN32;
        if (T == 0) k32 ? launch(k_explicit_deposit<o, 0, float, HPQ_SUM_KAHAN32>) : launch(k_explicit_deposit<o, 0, float>);
        else if (T == 1) k32 ? launch(k_explicit_deposit<o, 1, float, HPQ_SUM_KAHAN32>) : launch(k_explicit_deposit<o, 1, float>);
        else k32 ? launch(k_explicit_deposit<o, 2, float, HPQ_SUM_KAHAN32>) : launch(k_explicit_deposit<o, 2, float>);
    });
}

}  // namespace hpq
