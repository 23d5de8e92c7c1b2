// Host side of the slice loop, mirroring the reference's call surface:
//   Fields (InitializeSlices, ShiftSlices, AddRhoIons, SolvePoissonPsiExmByEypBxEzBz,
//           SolvePoissonBxBy, InitialBfieldGuess, MixAndShiftBfields, ComputeRelBFieldError)
//   FFTPoissonSolverDirichlet (SolvePoissonEquation on the staging area)
//   hpmg::MultiGrid::solve1
//   DepositCurrent, ExplicitDeposition, DepositCurrentSlice, AdvancePlasmaParticles
//   Hipace (SolveOneSlice, InitializeSxSyWithBeam, ExplicitMGSolveBxBy,
//           PredictorCorrectorLoopToSolveBxBy)
// and the C ABI of include/hpq.h on top of them. Everything runs on the context's
// stream; the only host synchronisations are the ones the reference also has
// (MG residual norm per V-cycle, PC-loop B-field error, QSA counter on request).
#include <cmath>
#include <cstring>
#include <functional>

#include "internal.hpp"

namespace hpq {

TraceScope::TraceScope(hpq_ctx* c, int kind) : c_(c), kind_(kind), slot_(0) {
    Trace& t = c_->trace[kind_];
    if (!t.on) return;
    if (t.used == t.start.size()) {
        hipEvent_t a, b;
        HPQ_HIP(hipEventCreate(&a));
        HPQ_HIP(hipEventCreate(&b));
        t.start.push_back(a);
        t.stop.push_back(b);
    }
    slot_ = t.used++;
    HPQ_HIP(hipEventRecord(t.start[slot_], c_->stream));
}

TraceScope::~TraceScope() {
    Trace& t = c_->trace[kind_];
    if (t.on) (void)hipEventRecord(t.stop[slot_], c_->stream);
}

void layout_init(hpq_ctx* c) {
    Layout L;
    int n = 0;
    const bool ions = c->p.neutralize != 0;
    if (c->p.explicit_solver) {
        L.next_jx_beam = n++; L.next_jy_beam = n++;
        L.chi = n++; L.Sy = n++; L.Sx = n++; L.ExmBy = n++; L.EypBx = n++; L.Ez = n++;
        L.Bx = n++; L.By = n++; L.Bz = n++; L.Psi = n++;
        L.jx_beam = n++; L.jy_beam = n++; L.jz_beam = n++; L.jx = n++; L.jy = n++; L.rhomjz = n++;
        if (c->p.deposit_rho) L.rho = n++;
        L.prev_jx_beam = n++; L.prev_jy_beam = n++;
        if (ions) L.ions_rhomjz = n++;
        if (c->p.salame) {
            L.sal_Ez_target = n++; L.sal_Ez_no_salame = n++; L.sal_Ez = n++; L.sal_jx = n++; L.sal_jy = n++;
            L.sal_jz_beam = n++; L.sal_Bx = n++; L.sal_By = n++; L.sal_Sy = n++; L.sal_Sx = n++;
            L.sal_Sy_back = n++; L.sal_Sx_back = n++;
        }
    } else {
        L.next_jx = n++; L.next_jy = n++;
        L.ExmBy = n++; L.EypBx = n++; L.Ez = n++; L.Bx = n++; L.By = n++; L.Bz = n++; L.Psi = n++;
        L.jx = n++; L.jy = n++; L.jz = n++; L.rhomjz = n++;
        if (c->p.deposit_rho) L.rho = n++;
        L.prev_Bx = n++; L.prev_By = n++; L.prev_jx = n++; L.prev_jy = n++;
        if (ions) L.ions_rhomjz = n++;
        L.pciter_Bx = n++; L.pciter_By = n++;
        L.pcprev_Bx = n++; L.pcprev_By = n++;
    }
    L.ncomp = n;
    c->L = L;
}

// the second stream and its events, created on first use (default priority)
void side_init(hpq_ctx* c) {
    if (c->side) return;
    HPQ_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    for (hipEvent_t* e : {&c->ev_start, &c->ev_beam, &c->ev_fields, &c->ev_end})
        HPQ_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
}

// ---------------------------------------------------------------------------------
// Fields (src/fields/Fields.cpp)
// ---------------------------------------------------------------------------------
class Fields {
public:
    Fields(hpq_ctx* c, double* store) : c_(c), s_(store), L_(c->L) {}

    // Fields::InitializeSlices, :536-586 (level 0): the planes zeroed before a slice
    int InitZeros(int* z) const {
        if (c_->p.explicit_solver) {
            const int v[] = {L_.chi, L_.Sy, L_.Sx, L_.ExmBy, L_.EypBx, L_.jz_beam, L_.rhomjz, L_.next_jx_beam,
                             L_.next_jy_beam, L_.rho};
            for (int k = 0; k < 10; ++k) z[k] = v[k];
            return 10;
        }
        const int v[] = {L_.ExmBy, L_.EypBx, L_.jx, L_.jy, L_.jz, L_.rhomjz, L_.rho};
        for (int k = 0; k < 7; ++k) z[k] = v[k];
        return 7;
    }
    void InitializeSlices() {
        int z[16];
        const int n = InitZeros(z);
        plane_fill(c_, s_, z, n, 0.0);
    }
    // Fields::ShiftSlices, :589-604: copies in the order they apply at every cell
    int ShiftCopies(int* d, int* s) const {
        if (c_->p.explicit_solver) {
            const int dv[] = {L_.prev_jx_beam, L_.prev_jy_beam, L_.jx_beam, L_.jy_beam, L_.jx, L_.jy};
            const int sv[] = {L_.jx_beam, L_.jy_beam, L_.next_jx_beam, L_.next_jy_beam, L_.next_jx_beam,
                              L_.next_jy_beam};
            for (int k = 0; k < 6; ++k) { d[k] = dv[k]; s[k] = sv[k]; }
            return 6;
        }
        const int dv[] = {L_.pcprev_Bx, L_.pcprev_By, L_.prev_Bx, L_.prev_By, L_.prev_jx, L_.prev_jy};
        const int sv[] = {L_.prev_Bx, L_.prev_By, L_.Bx, L_.By, L_.jx, L_.jy};
        for (int k = 0; k < 6; ++k) { d[k] = dv[k]; s[k] = sv[k]; }
        return 6;
    }
    void ShiftSlices() {
        int d[8], s[8];
        const int n = ShiftCopies(d, s);
        plane_copy(c_, s_, d, s, n);
    }
    // Fields::AddRhoIons, :607-615
    void AddRhoIons() {
        if (L_.ions_rhomjz < 0) return;
        const int d[2] = {L_.rhomjz, L_.rho};
        const int s[2] = {L_.ions_rhomjz, L_.ions_rhomjz};
        plane_add(c_, s_, d, s, 2);
    }
    // The field diagnostics of this slice (sum |F|, FillDiagnostics, InSituComputeDiags), and, when another
    // slice follows in the same pass, ShiftSlices plus that slice's InitializeSlices: one
    // pass over the store. The plasma push that comes next reads none of these planes.
    void EndSlice(int islice, bool chain) {
        captures(c_, s_, islice);
        if (chain) {
            int d[8], s[8], z[16];
            const int ns = ShiftCopies(d, s);
            const int nz = InitZeros(z);
            slice_end(c_, s_, islice, d, s, ns, z, nz);
        } else {
            slice_end(c_, s_, islice, nullptr, nullptr, 0, nullptr, 0);
        }
    }
    // Fields::SolvePoissonPsiExmByEypBxEzBz, :841-957
    void SolvePoissonPsiExmByEypBxEzBz() {
        build_rhs_psi_ez_bz(c_, s_);
        double* lhs[3] = {s_ + long(L_.Psi) * c_->g.plane, s_ + long(L_.Ez) * c_->g.plane,
                          s_ + long(L_.Bz) * c_->g.plane};
        poisson_solve_staged(c_, 3, lhs);
        exmby_eypbx(c_, s_);
    }
    // Fields::SolvePoissonBxBy, :1009-1078 (into PCIter)
    void SolvePoissonBxBy() {
        build_rhs_bxby(c_, s_);
        double* lhs[3] = {s_ + long(L_.pciter_Bx) * c_->g.plane, s_ + long(L_.pciter_By) * c_->g.plane, nullptr};
        poisson_solve_staged(c_, 2, lhs);
    }
    // Fields::InitialBfieldGuess, :1150-1171
    void InitialBfieldGuess(double err, double tol) {
        const double m = std::exp(-0.5 * std::pow(err / (2.5 * tol), 2));
        plane_lincomb(c_, s_, L_.Bx, 1. + m, L_.prev_Bx, -m, L_.pcprev_Bx);
        plane_lincomb(c_, s_, L_.By, 1. + m, L_.prev_By, -m, L_.pcprev_By);
    }
    // Fields::MixAndShiftBfields, :1174-1226
    void MixAndShiftBfields(double err, double err_prev, double mix) {
        double wi, wp;
        if (err != 0. || err_prev != 0.) {
            wi = err_prev / (err + err_prev);
            wp = err / (err + err_prev);
        } else {
            wi = 0.5;
            wp = 0.5;
        }
        plane_lincomb(c_, s_, L_.pcprev_Bx, wi, L_.pciter_Bx, wp, L_.pcprev_Bx);
        plane_lincomb(c_, s_, L_.pcprev_By, wi, L_.pciter_By, wp, L_.pcprev_By);
        plane_lincomb(c_, s_, L_.Bx, 1. - mix, L_.Bx, mix, L_.pcprev_Bx);
        plane_lincomb(c_, s_, L_.By, 1. - mix, L_.By, mix, L_.pcprev_By);
        const int d[] = {L_.pcprev_Bx, L_.pcprev_By};
        const int s[] = {L_.pciter_Bx, L_.pciter_By};
        plane_copy(c_, s_, d, s, 2);
    }
    // Fields::ComputeRelBFieldError, :1229-1286
    double ComputeRelBFieldError(int bx, int by, int bxi, int byi) { return rel_b_error(c_, s_, bx, by, bxi, byi); }

private:
    hpq_ctx* c_;
    double* s_;
    const Layout& L_;
};

namespace hpmg {
// hpmg::MultiGrid (src/mg_solver/HpMultiGrid.H:37-268), type I; the hierarchy lives in
// the context and is built once per grid.
class MultiGrid {
public:
    explicit MultiGrid(hpq_ctx* c) : c_(c) {}
    int solve1(double* sol, const double* rhs, const double* acf, long comp_stride, double tol_rel, double tol_abs,
               int nummaxiter, int* iters) {
        return mg_solve1(c_, sol, rhs, acf, comp_stride, tol_rel, tol_abs, nummaxiter, iters);
    }

private:
    hpq_ctx* c_;
};
}  // namespace hpmg

// ::DepositCurrent (src/particles/deposition/PlasmaDepositCurrent.cpp:22)
void DepositCurrent(hpq_ctx* c, const hpq_plasma* p, double* store, int jx, int jy, int jz, int rho, int chi,
                    int rhomjz, bool ions = false) {
    const int32_t comps[6] = {jx, jy, jz, rho, chi, rhomjz};
    deposit_current(c, p, store, comps, ions);
}

// ::DepositCurrentSlice (src/particles/deposition/BeamDepositCurrent.cpp:22)
void DepositCurrentSlice(hpq_ctx* c, const hpq_beam* b, int slice, double* store, int jx, int jy, int jz) {
    const int32_t comps[3] = {jx, jy, jz};
    deposit_beam(c, b, slice, store, comps);
}

// ---------------------------------------------------------------------------------
// Hipace (src/Hipace.cpp)
// ---------------------------------------------------------------------------------
class Hipace {
public:
    Hipace(hpq_ctx* c, const hpq_plasma* p, const hpq_beam* b, double* store)
        : c_(c), p_(p), b_(b), s_(store), fields_(c, store), L_(c->L) {}

    // Hipace::SolveOneSlice, :557-728 (level 0; no laser, ionisation, collisions).
    // initialized: the previous slice of the same pass has already run InitializeSlices for
    // this one; chain: another slice follows in this pass. two: the beam deposition and the
    // end-of-slice pass run on the context's second stream (explicit solver, chained slices):
    //   side: [jz_beam of this slice, jx/jy_beam of the next]          after the planes are zeroed
    //   main: [plasma deposit, Poisson solves] wait(beam) [Sx/Sy, explicit deposit, MG, SALAME]
    //   side: wait(fields) [diagnostics, ShiftSlices, next InitializeSlices]
    //   main: [plasma push] wait(end)
    // The push reads Psi, Ez, Bx, By, Bz, which the end-of-slice pass neither writes nor
    // zeroes, so the two overlap; the next slice's deposits wait for the zeroing.
    void SolveOneSlice(int islice, bool initialized = false, bool chain = false,
                       const std::function<void()>& before_push = {}, bool two = false) {
        const hpq_params& P = c_->p;
        two = two && P.explicit_solver;
        insitu_plasma(c_, p_, islice);
        // MultiPlasma::ReorderParticles (Hipace.cpp:594-595)
        if (plasma_reordered(c_) && islice % P.reorder_period == 0) reorder_plasma(c_, p_);
        if (!initialized) fields_.InitializeSlices();
        auto on_side = [&](auto&& f) {
            hipStream_t main = c_->stream;
            c_->stream = c_->side;
            try {
                f();
            } catch (...) {
                c_->stream = main;
                throw;
            }
            c_->stream = main;
        };
        auto beam_deposits = [&] {
            DepositCurrentSlice(c_, b_, islice, s_, -1, -1, L_.jz_beam);
            if (P.do_beam_jx_jy) {
                // SALAME beams deposit no jx, jy while SALAME runs (MultiBeam.cpp:50-56)
                const int32_t comps[3] = {L_.next_jx_beam, L_.next_jy_beam, -1};
                deposit_beam_select(c_, b_, islice - 1, s_, comps, c_->salame.mask, 0);
            }
        };
        if (two) {
            HPQ_HIP(hipEventRecord(c_->ev_start, c_->stream));
            on_side([&] {
                HPQ_HIP(hipStreamWaitEvent(c_->stream, c_->ev_start, 0));
                beam_deposits();
                HPQ_HIP(hipEventRecord(c_->ev_beam, c_->stream));
            });
        }
        if (P.explicit_solver) {
            DepositCurrent(c_, p_, s_, L_.jx, L_.jy, -1, L_.rho, L_.chi, L_.rhomjz);
            if (!two) DepositCurrentSlice(c_, b_, islice, s_, -1, -1, L_.jz_beam);
        } else {
            DepositCurrent(c_, p_, s_, L_.jx, L_.jy, L_.jz, L_.rho, -1, L_.rhomjz);
            DepositCurrentSlice(c_, b_, islice, s_, P.do_beam_jx_jy ? L_.jx : -1, P.do_beam_jx_jy ? L_.jy : -1,
                                L_.jz);
        }
        fields_.AddRhoIons();
        fields_.SolvePoissonPsiExmByEypBxEzBz();
        if (P.explicit_solver) {
            if (two) {
                HPQ_HIP(hipStreamWaitEvent(c_->stream, c_->ev_beam, 0));
            } else if (P.do_beam_jx_jy) {
                const int32_t comps[3] = {L_.next_jx_beam, L_.next_jy_beam, -1};
                deposit_beam_select(c_, b_, islice - 1, s_, comps, c_->salame.mask, 0);
            }
            InitializeSxSyWithBeam();
            explicit_deposition(c_, p_, s_);
            if (P.mg_guess == HPQ_MG_GUESS_LINEAR) mg_guess_linear(c_, s_ + long(L_.Bx) * c_->g.plane, islice);
            if (P.mg_guess == HPQ_MG_GUESS_DST) bxby_guess_dst(c_, s_, P.mg_guess_iters);
            ExplicitMGSolveBxBy();
            if (c_->salame.mask && salame_count(c_, b_, islice) > 0) SalameModule(islice);
        } else {
            PredictorCorrectorLoopToSolveBxBy(islice);
        }
        insitu_beam(c_, b_, islice);
        if (two) {
            HPQ_HIP(hipEventRecord(c_->ev_fields, c_->stream));
            on_side([&] {
                HPQ_HIP(hipStreamWaitEvent(c_->stream, c_->ev_fields, 0));
                fields_.EndSlice(islice, chain);
                HPQ_HIP(hipEventRecord(c_->ev_end, c_->stream));
            });
        } else {
            fields_.EndSlice(islice, chain);
        }
        if (before_push) before_push();
        advance_plasma(c_, p_, s_, false);
        c_->stats[3] += double(p_->n);
        if (two) HPQ_HIP(hipStreamWaitEvent(c_->stream, c_->ev_end, 0));
        if (!chain) fields_.ShiftSlices();
    }

    // Hipace::InitializeSxSyWithBeam, :745-790
    void InitializeSxSyWithBeam() { sxsy_from_beam(c_, s_); }

    // Hipace::ExplicitMGSolveBxBy, :794-933 (hpmg branch, level 0); chi always from This,
    // Bx/By and Sy/Sx from This or from the SALAME planes
    void ExplicitMGSolveBxBy(bool salame_slice = false) {
        hpmg::MultiGrid mg(c_);
        const long ps = c_->g.plane;
        const int bx = salame_slice ? L_.sal_Bx : L_.Bx, sy = salame_slice ? L_.sal_Sy : L_.Sy;
        int iters = 0;
        const int st = mg.solve1(s_ + long(bx) * ps, s_ + long(sy) * ps, s_ + long(L_.chi) * ps, ps,
                                 c_->p.mg_tol_rel, c_->p.mg_tol_abs, c_->p.mg_maxiter, &iters);
        c_->stats[1] += iters;
        if (st == HPQ_E_MG_DIVERGED) throw HpqError(st, "hpmg failing so lets stop here");
        if (st == HPQ_E_MG_NOT_CONVERGED) throw HpqError(st, "hpmg failed");
    }

    // Fields::SolvePoissonEz, src/fields/Fields.cpp:960-1006, on the SALAME planes
    void SolvePoissonEzSalame() {
        salame_rhs_ez(c_, s_);
        double* lhs[3] = {s_ + long(L_.sal_Ez) * c_->g.plane, nullptr, nullptr};
        poisson_solve_staged(c_, 1, lhs);
    }

    // SalameModule, src/salame/Salame.cpp:14-184 (level 0, no grid current)
    void SalameModule(int islice) {
        hpq_ctx* c = c_;
        const Layout& L = L_;
        const Geom& g = c->g;
        const hpq_salame& S = c->salame;
        auto fill0 = [&](std::initializer_list<int> comps) {
            std::vector<int> v(comps);
            plane_fill(c, s_, v.data(), int(v.size()), 0.0);
        };
        auto copy = [&](std::initializer_list<int> dst, std::initializer_list<int> src) {
            std::vector<int> d(dst), s(src);
            plane_copy(c, s_, d.data(), s.data(), int(d.size()));
        };
        // GetPosOffset(2, geom, domain), src/fields/Fields.H:72-77
        const double zoff = 0.5 * (g.zlo + g.zhi - g.dz * double(0 + g.nz - 1));
        // the Ez field from before SALAME started on this run of slices (:21-29)
        if (islice + 1 != c->sal_last_slice) {
            copy({L.sal_Ez_target}, {L.Ez});
            c->sal_overloaded = false;
            c->sal_zeta_initial = islice * g.dz + zoff;
        }
        c->sal_last_slice = islice;
        // plasma-only Sx, Sy, kept for step 4 (:31-38)
        fill0({L.Sy, L.Sx});
        explicit_deposition(c, p_, s_);
        copy({L.sal_Sy_back, L.sal_Sx_back}, {L.Sy, L.Sx});
        double W = 1.0;
        for (int iter = 0; iter < S.n_iter; ++iter) {
            // STEP 1: Ez of the next slice with the present SALAME weights (:44-75)
            advance_plasma(c, p_, s_, true);
            copy({L.sal_jx, L.sal_jy}, {L.next_jx_beam, L.next_jy_beam});
            DepositCurrent(c, p_, s_, L.sal_jx, L.sal_jy, -1, -1, -1, -1);
            fill0({L.sal_Ez, L.sal_jz_beam, L.sal_Sy, L.sal_Sx, L.sal_Bx, L.sal_By});
            SolvePoissonEzSalame();
            copy({L.sal_Ez_no_salame}, {L.sal_Ez});
            // STEP 2: Ez of the SALAME beam alone (:77-120)
            const int32_t cz[3] = {-1, -1, L.sal_jz_beam};
            deposit_beam_select(c, b_, islice, s_, cz, S.mask, 1);
            salame_sxsy(c, s_);
            ExplicitMGSolveBxBy(true);
            fill0({L.sal_Ez, L.sal_jx, L.sal_jy});
            if (S.do_advance) {
                salame_only_advance(c, p_, s_);
                DepositCurrent(c, p_, s_, L.sal_jx, L.sal_jy, -1, -1, -1, -1);
            } else {
                salame_jxjy_from_b(c, s_);
            }
            SolvePoissonEzSalame();
            // STEP 3: the weight factor (:122-159, SalameGetW :332-407)
            fill0({L.sal_jz_beam});
            deposit_beam_select(c, b_, islice, s_, cz, S.mask, 1);
            double sum[4];
            salame_sums(c, s_, sum);
            double sum_Ez_target = sum[0], sum_Ez_no_salame = sum[1], sum_Ez_only_salame = sum[2];
            const double sum_jz = sum[3];
            sum_Ez_target /= sum_jz;
            sum_Ez_no_salame /= sum_jz;
            sum_Ez_only_salame /= sum_jz;
            const double zeta = (islice - 1) * g.dz + zoff;
            if (S.target) sum_Ez_target = S.target(zeta, c->sal_zeta_initial, sum_Ez_target, S.user);
            W = (sum_Ez_target - sum_Ez_no_salame) / sum_Ez_only_salame + 1.0;
            if (W < 0 || c->sal_overloaded) {
                W = 0;
                iter = S.n_iter - 1;
                c->sal_overloaded = true;
            }
            if (!c->sal_overloaded && iter >= 1 && std::abs(W - 1.) < S.relative_tolerance) iter = S.n_iter - 1;
            salame_multiply(c, b_, islice, W);
            // STEP 4: Bx, By with the new weights (:161-177)
            fill0({L.jz_beam, L.Sy, L.Sx});
            DepositCurrentSlice(c, b_, islice, s_, -1, -1, L.jz_beam);
            InitializeSxSyWithBeam();
            const int d[2] = {L.Sy, L.Sx}, sb[2] = {L.sal_Sy_back, L.sal_Sx_back};
            plane_add(c, s_, d, sb, 2);
            ExplicitMGSolveBxBy();
        }
        if (S.weight_factor) S.weight_factor[islice] = W;
    }

    // Hipace::PredictorCorrectorLoopToSolveBxBy, :936-1031
    void PredictorCorrectorLoopToSolveBxBy(int islice) {
        const hpq_params& P = c_->p;
        double err_prev = 1.0;
        double err = fields_.ComputeRelBFieldError(L_.prev_Bx, L_.prev_By, L_.pcprev_Bx, L_.pcprev_By);
        fields_.InitialBfieldGuess(err, P.pc_tol);
        const int z[] = {L_.pciter_Bx, L_.pciter_By};
        plane_fill(c_, s_, z, 2, 0.0);
        const int d[] = {L_.pcprev_Bx, L_.pcprev_By};
        const int s[] = {L_.Bx, L_.By};
        plane_copy(c_, s_, d, s, 2);
        int it = 0;
        err = 1.0;
        while (err > P.pc_tol && it < P.pc_maxiter) {
            ++it;
            advance_plasma(c_, p_, s_, true);
            DepositCurrent(c_, p_, s_, L_.next_jx, L_.next_jy, -1, -1, -1, -1);
            if (P.do_beam_jx_jy) DepositCurrentSlice(c_, b_, islice - 1, s_, L_.next_jx, L_.next_jy, -1);
            fields_.SolvePoissonBxBy();
            err = fields_.ComputeRelBFieldError(L_.Bx, L_.By, L_.pciter_Bx, L_.pciter_By);
            if (it == 1) err_prev = err;
            fields_.MixAndShiftBfields(err, err_prev, P.pc_mix);
            const int zz[] = {L_.next_jx, L_.next_jy};
            plane_fill(c_, s_, zz, 2, 0.0);
            err_prev = err;
        }
        c_->stats[2] += it;
    }

private:
    hpq_ctx* c_;
    const hpq_plasma* p_;
    const hpq_beam* b_;
    double* s_;
    Fields fields_;
    const Layout& L_;
};

}  // namespace hpq

// ---------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------
using namespace hpq;

namespace {

template <class F>
int guarded(hpq_ctx* c, F&& f) {
    // the caller may have changed the particle arrays since the last call
    if (c) c->cl_plasma.valid = false;
    try {
        f();
        return HPQ_OK;
    } catch (const HpqError& e) {
        if (c) c->err = e.msg;
        return e.code;
    } catch (const HipError& e) {
        if (c) c->err = e.msg;
        return HPQ_E_HIP;
    } catch (const std::bad_alloc&) {
        if (c) c->err = "out of host memory";
        return HPQ_E_NOMEM;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        return HPQ_E_INVALID;
    }
}

void check_store(const hpq_ctx* c, const void* store) {
    if (!store) throw HpqError(HPQ_E_INVALID, "slice store pointer is null");
}
void check_plasma(const hpq_plasma* p) {
    if (!p) throw HpqError(HPQ_E_INVALID, "plasma descriptor is null");
    if (p->n > 0 && (!p->x || !p->y || !p->w || !p->ux || !p->uy || !p->psi || !p->x_prev || !p->y_prev ||
                     !p->ux_half_step || !p->uy_half_step || !p->psi_half_step || !p->valid))
        throw HpqError(HPQ_E_INVALID, "plasma descriptor has a null array");
}
void check_comp(const hpq_ctx* c, int k) {
    if (k < -1 || k >= c->L.ncomp) throw HpqError(HPQ_E_INVALID, "component index out of range");
}

}  // namespace

// ---------------------------------------------------------------------------------
// Adjoint slice loop (no counterpart in the reference): a checkpointed forward pass, then a
// backward march over the slices in reverse order. Each checkpoint segment of k slices is
// recomputed from its checkpoint, saving the slice store and the plasma right before each
// push; the adjoint slice step then runs the vector-Jacobian products of push, Bx/By solve,
// explicit deposition, beam source, Poisson solves and plasma deposition in reverse.
// ---------------------------------------------------------------------------------
namespace {

void plasma_buf_alloc(PlasmaBuf& b, long n) {
    b.n = n;
    for (int k = 0; k < PLASMA_REALS; ++k) HPQ_HIP(hipMalloc(&b.r[k], sizeof(double) * std::max(n, 1L)));
    HPQ_HIP(hipMalloc(&b.valid, sizeof(int32_t) * std::max(n, 1L)));
}

void plasma_buf_free(PlasmaBuf& b) {
    for (double*& r : b.r)
        if (r) (void)hipFree(r), r = nullptr;
    if (b.valid) (void)hipFree(b.valid), b.valid = nullptr;
}

double* const* plasma_reals(const hpq_plasma* p, double* (&a)[PLASMA_REALS]) {
    double* v[PLASMA_REALS] = {p->x, p->y, p->w, p->ux, p->uy, p->psi, p->x_prev, p->y_prev,
                               p->ux_half_step, p->uy_half_step, p->psi_half_step};
    for (int k = 0; k < PLASMA_REALS; ++k) a[k] = v[k];
    return a;
}

void plasma_save(hpq_ctx* c, const hpq_plasma* p, PlasmaBuf& b) {
    double* a[PLASMA_REALS];
    plasma_reals(p, a);
    for (int k = 0; k < PLASMA_REALS; ++k)
        HPQ_HIP(hipMemcpyAsync(b.r[k], a[k], sizeof(double) * p->n, hipMemcpyDeviceToDevice, c->stream));
    HPQ_HIP(hipMemcpyAsync(b.valid, p->valid, sizeof(int32_t) * p->n, hipMemcpyDeviceToDevice, c->stream));
}

void plasma_restore(hpq_ctx* c, const PlasmaBuf& b, const hpq_plasma* p) {
    c->cl_plasma.valid = false;
    double* a[PLASMA_REALS];
    plasma_reals(p, a);
    for (int k = 0; k < PLASMA_REALS; ++k)
        HPQ_HIP(hipMemcpyAsync(a[k], b.r[k], sizeof(double) * p->n, hipMemcpyDeviceToDevice, c->stream));
    HPQ_HIP(hipMemcpyAsync(p->valid, b.valid, sizeof(int32_t) * p->n, hipMemcpyDeviceToDevice, c->stream));
}

hpq_plasma plasma_view(const PlasmaBuf& b) {
    hpq_plasma v{};
    v.x = b.r[0]; v.y = b.r[1]; v.w = b.r[2]; v.ux = b.r[3]; v.uy = b.r[4]; v.psi = b.r[5];
    v.x_prev = b.r[6]; v.y_prev = b.r[7]; v.ux_half_step = b.r[8]; v.uy_half_step = b.r[9];
    v.psi_half_step = b.r[10];
    v.valid = b.valid;
    v.n = b.n;
    return v;
}

hpq_plasma_adj adj_view(double* const (&a)[8]) {
    return hpq_plasma_adj{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]};
}

size_t store_bytes(const hpq_ctx* c) { return sizeof(double) * size_t(c->L.ncomp) * c->g.plane; }

void adjoint_alloc(hpq_ctx* c, int k, long np) {
    AdjointState& A = c->adj;
    const int nck = (c->g.nz + k - 1) / k;
    if (A.k == k && A.np == np && A.nck == nck) return;
    adjoint_free(c);
    A.k = k;
    A.nck = nck;
    A.np = np;
    A.ck_store.assign(nck, nullptr);
    A.ck_plasma.assign(nck, PlasmaBuf{});
    A.sl_store.assign(k, nullptr);
    A.sl_plasma.assign(k, PlasmaBuf{});
    for (int i = 0; i < nck; ++i) {
        HPQ_HIP(hipMalloc(&A.ck_store[i], store_bytes(c)));
        plasma_buf_alloc(A.ck_plasma[i], np);
    }
    for (int i = 0; i < k; ++i) {
        HPQ_HIP(hipMalloc(&A.sl_store[i], store_bytes(c)));
        plasma_buf_alloc(A.sl_plasma[i], np);
    }
    HPQ_HIP(hipMalloc(&A.store_adj, store_bytes(c)));
    HPQ_HIP(hipMalloc(&A.mu, 2 * sizeof(double) * c->g.plane));
    for (auto& set : A.padj)
        for (double*& a : set) HPQ_HIP(hipMalloc(&a, sizeof(double) * std::max(np, 1L)));
}

}  // namespace

void hpq::adjoint_free(hpq_ctx* c) {
    AdjointState& A = c->adj;
    for (double* s : A.ck_store)
        if (s) (void)hipFree(s);
    for (double* s : A.sl_store)
        if (s) (void)hipFree(s);
    for (PlasmaBuf& b : A.ck_plasma) plasma_buf_free(b);
    for (PlasmaBuf& b : A.sl_plasma) plasma_buf_free(b);
    if (A.store_adj) (void)hipFree(A.store_adj);
    for (auto& set : A.padj)
        for (double* a : set)
            if (a) (void)hipFree(a);
    if (A.push_g) (void)hipFree(A.push_g);
    if (A.mu) (void)hipFree(A.mu);
    A = AdjointState{};
}

namespace {

void adjoint_check(const hpq_ctx* c) {
    const hpq_params& P = c->p;
    if (!P.explicit_solver) throw HpqError(HPQ_E_UNSUPPORTED, "adjoint: explicit Bx/By solver only");
    if (P.nx % 2 == 0 || P.ny % 2 == 0)
        throw HpqError(HPQ_E_UNSUPPORTED, "adjoint: odd nx, ny (node-centred multigrid, a symmetric operator)");
    if (P.do_beam_jx_jy) throw HpqError(HPQ_E_UNSUPPORTED, "adjoint: do_beam_jx_jy must be 0");
    if (c->salame.mask) throw HpqError(HPQ_E_UNSUPPORTED, "adjoint: SALAME must be off");
    if (plasma_reordered(c))
        throw HpqError(HPQ_E_UNSUPPORTED, "adjoint: plasma reordering (reorder_period) must be off");
    if (P.mg_guess != HPQ_MG_GUESS_PREVIOUS)
        throw HpqError(HPQ_E_UNSUPPORTED, "adjoint: mg_guess must be HPQ_MG_GUESS_PREVIOUS");
}

// the forward pass with checkpoints every k slices
void adjoint_forward(hpq_ctx* c, const hpq_plasma* p, const hpq_beam* b, double* store, int k) {
    // a forward pass that stops part way leaves no usable checkpoints
    c->adj.ready = false;
    adjoint_check(c);
    adjoint_alloc(c, k, p->n);
    AdjointState& A = c->adj;
    Hipace h(c, p, b, store);
    const int nz = c->g.nz;
    for (int isl = nz - 1; isl >= 0; --isl) {
        const int step = nz - 1 - isl;
        if (step % k == 0) {
            HPQ_HIP(hipMemcpyAsync(A.ck_store[step / k], store, store_bytes(c), hipMemcpyDeviceToDevice, c->stream));
            plasma_save(c, p, A.ck_plasma[step / k]);
        }
        h.SolveOneSlice(isl);
    }
    diag_finalize(c);
    A.ready = true;
}

// the backward march; seed: d loss / d Ez per slice ([slice][j][i], valid cells) or null;
// w_adj: beam weight adjoints, accumulated
void adjoint_backward(hpq_ctx* c, const hpq_plasma* p, const hpq_beam* b, double* store, const double* seed,
                      const double* jzb_seed, double* w_adj, double mg_tol) {
    adjoint_check(c);
    AdjointState& A = c->adj;
    if (!A.ready || A.np != p->n) throw HpqError(HPQ_E_INVALID, "adjoint: no matching forward pass");
    const Geom& g = c->g;
    const Layout& L = c->L;
    const int nz = g.nz, k = A.k;
    const long ps = g.plane;
    // the recomputation must not overwrite the caller's captures or diagnostics
    std::vector<Capture> caps;
    caps.swap(c->captures);
    double* checksums = c->checksums;
    c->checksums = nullptr;
    const hpq_insitu insitu = c->insitu;
    c->insitu = hpq_insitu{};
    try {
        for (double* a : A.padj[0]) HPQ_HIP(hipMemsetAsync(a, 0, sizeof(double) * p->n, c->stream));
        HPQ_HIP(hipMemsetAsync(A.mu, 0, 2 * sizeof(double) * ps, c->stream));
        int cur = 0;
        for (int seg = A.nck - 1; seg >= 0; --seg) {
            const int s_first = nz - 1 - seg * k;
            const int s_last = std::max(0, s_first - k + 1);
            HPQ_HIP(hipMemcpyAsync(store, A.ck_store[seg], store_bytes(c), hipMemcpyDeviceToDevice, c->stream));
            plasma_restore(c, A.ck_plasma[seg], p);
            {
                Hipace h(c, p, b, store);
                for (int isl = s_first; isl >= s_last; --isl) {
                    const int slot = s_first - isl;
                    h.SolveOneSlice(isl, false, false, [&] {
                        HPQ_HIP(hipMemcpyAsync(A.sl_store[slot], store, store_bytes(c), hipMemcpyDeviceToDevice,
                                               c->stream));
                        plasma_save(c, p, A.sl_plasma[slot]);
                    });
                }
            }
            for (int isl = s_last; isl <= s_first; ++isl) {
                const int slot = s_first - isl;
                const int nxt = 1 - cur;
                const hpq_plasma sp = plasma_view(A.sl_plasma[slot]);
                const double* st = A.sl_store[slot];
                double* adj = A.store_adj;
                HPQ_HIP(hipMemsetAsync(adj, 0, store_bytes(c), c->stream));
                for (double* a : A.padj[nxt]) HPQ_HIP(hipMemsetAsync(a, 0, sizeof(double) * p->n, c->stream));
                hpq_plasma_adj out = adj_view(A.padj[cur]), in = adj_view(A.padj[nxt]);
                // push: adjoints of its inputs and of the Psi, Ez, Bx, By, Bz planes
                advance_plasma_vjp(c, &sp, st, &out, &in, adj);
                if (seed) adj_seed(c, seed + long(isl) * g.nx * g.ny, adj);
                // (Lap - chi) B = S is self-adjoint for the node-centred operator:
                // S_adj = mu with (Lap - chi) mu = B_adj, chi_adj = mu . B. The previous
                // slice's mu starts the iteration, as the previous B does in the forward.
                if (L.Sx != L.Sy + 1 || L.By != L.Bx + 1) throw HpqError(HPQ_E_INVALID, "adjoint: layout");
                HPQ_HIP(hipMemcpyAsync(adj + long(L.Sy) * ps, A.mu, 2 * sizeof(double) * ps, hipMemcpyDeviceToDevice,
                                       c->stream));
                int iters = 0;
                const int st_mg = mg_solve1(c, adj + long(L.Sy) * ps, adj + long(L.Bx) * ps, st + long(L.chi) * ps,
                                            ps, mg_tol, 0.0, 200, &iters);
                if (st_mg != HPQ_OK) throw HpqError(st_mg, "adjoint: Bx/By adjoint solve did not converge");
                HPQ_HIP(hipMemcpyAsync(A.mu, adj + long(L.Sy) * ps, 2 * sizeof(double) * ps, hipMemcpyDeviceToDevice,
                                       c->stream));
                adj_chi(c, st, adj);
                explicit_deposition_vjp(c, &sp, st, adj, &in);
                adj_beam_weights(c, b, isl, adj, jzb_seed ? jzb_seed + long(isl) * g.nx * g.ny : nullptr, w_adj);
                adj_poisson(c, adj);
                const int32_t comps[6] = {L.jx, L.jy, -1, -1, L.chi, L.rhomjz};
                deposit_vjp(c, &sp, adj, comps, &in);
                cur = nxt;
            }
        }
    } catch (...) {
        c->captures.swap(caps);
        c->checksums = checksums;
        c->insitu = insitu;
        throw;
    }
    c->captures.swap(caps);
    c->checksums = checksums;
    c->insitu = insitu;
}

}  // namespace

extern "C" {

int hpq_create(const hpq_params* p, hpq_ctx** out) {
    if (!p || !out) return HPQ_E_INVALID;
    *out = nullptr;
    hpq_ctx* c = new hpq_ctx();
    const int rc = guarded(c, [&] {
        if (p->nx < 4 || p->ny < 4 || p->nz < 1) throw HpqError(HPQ_E_INVALID, "grid too small");
        if (!(p->xhi > p->xlo && p->yhi > p->ylo && p->zhi > p->zlo)) throw HpqError(HPQ_E_INVALID, "empty domain");
        if (p->depos_order < 0 || p->depos_order > 3) throw HpqError(HPQ_E_UNSUPPORTED, "depos_order_xy must be 0..3");
        if (p->precision != HPQ_PREC_FP64 && p->precision != HPQ_PREC_MIXED)
            throw HpqError(HPQ_E_INVALID, "precision must be HPQ_PREC_FP64 or HPQ_PREC_MIXED");
        if (p->poisson_solver != HPQ_POISSON_FFT && p->poisson_solver != HPQ_POISSON_BF16X3)
            throw HpqError(HPQ_E_INVALID, "poisson_solver must be HPQ_POISSON_FFT or HPQ_POISSON_BF16X3");
        if (p->derivative_type < 0 || p->derivative_type > 2)
            throw HpqError(HPQ_E_UNSUPPORTED, "depos_derivative_type must be 0..2");
        if (p->depos_order == 0 && p->derivative_type == 0)
            throw HpqError(HPQ_E_INVALID, "Analytic derivative with depos_order=0 would vanish");
        if (p->salame && !p->explicit_solver)
            throw HpqError(HPQ_E_UNSUPPORTED, "Cannot use SALAME algorithm with predictor-corrector solver");
        if (p->deposition != HPQ_DEPOSIT_ORDERED && p->deposition != HPQ_DEPOSIT_ATOMIC)
            throw HpqError(HPQ_E_INVALID, "deposition must be HPQ_DEPOSIT_ORDERED or HPQ_DEPOSIT_ATOMIC");
        if (p->reorder_period < 0) throw HpqError(HPQ_E_INVALID, "reorder_period must be >= 0");
        if (p->mixed_sums != HPQ_SUM_FP64 && p->mixed_sums != HPQ_SUM_KAHAN32)
            throw HpqError(HPQ_E_INVALID, "mixed_sums must be HPQ_SUM_FP64 or HPQ_SUM_KAHAN32");
        if (p->mg_guess != HPQ_MG_GUESS_PREVIOUS && p->mg_guess != HPQ_MG_GUESS_LINEAR &&
            p->mg_guess != HPQ_MG_GUESS_DST)
            throw HpqError(HPQ_E_INVALID, "mg_guess must be HPQ_MG_GUESS_PREVIOUS, _LINEAR or _DST");
        if (p->mg_guess != HPQ_MG_GUESS_PREVIOUS && !p->explicit_solver)
            throw HpqError(HPQ_E_UNSUPPORTED, "mg_guess applies to the explicit Bx/By solver");
        if (p->mg_guess == HPQ_MG_GUESS_DST) {
            auto pow2 = [](int v) { return v > 1 && (v & (v - 1)) == 0 && v <= 2048; };
            if (!pow2(p->nx + 1) || !pow2(p->ny + 1))
                throw HpqError(HPQ_E_UNSUPPORTED, "mg_guess DST needs nx + 1 and ny + 1 powers of two");
            if (p->mg_guess_iters < 1) throw HpqError(HPQ_E_INVALID, "mg_guess_iters must be >= 1");
        }
        {
            // the widest stencil the ordered deposition walks: O + 1 cells for the current and
            // beam deposition, the derivative stencil for the explicit one
            const int o = p->depos_order, t = p->derivative_type;
            const int wide = std::max(o + 1, t == 0 ? o + 1 : (t == 1 ? o + 2 : o + 3));
            const int T = p->tile_size > 0 ? p->tile_size : 32;
            if (p->tile_size < 0 || T < wide)
                throw HpqError(HPQ_E_INVALID, "tile_size must be at least the width of the deposition stencil");
        }
        c->p = *p;
        Geom& g = c->g;
        g.nx = p->nx; g.ny = p->ny; g.nz = p->nz; g.ng = 2;
        g.ldx = g.nx + 2 * g.ng; g.ldy = g.ny + 2 * g.ng; g.plane = long(g.ldx) * g.ldy;
        g.xlo = p->xlo; g.xhi = p->xhi; g.ylo = p->ylo; g.yhi = p->yhi; g.zlo = p->zlo; g.zhi = p->zhi;
        g.dx = (g.xhi - g.xlo) / double(g.nx);
        g.dy = (g.yhi - g.ylo) / double(g.ny);
        g.dz = (g.zhi - g.zlo) / double(g.nz);
        g.dxi = 1.0 / g.dx; g.dyi = 1.0 / g.dy; g.dzi = 1.0 / g.dz;
        g.xoff = 0.5 * (g.xlo + g.xhi - g.dx * double(g.nx - 1));
        g.yoff = 0.5 * (g.ylo + g.yhi - g.dy * double(g.ny - 1));
        layout_init(c);
        HPQ_HIP(hipGetDevice(&c->device));
        bins_noalloc_setup(c->device);
        HPQ_HIP(hipMalloc(&c->d_count, sizeof(int) * 4));
        HPQ_HIP(hipHostMalloc(&c->h_count, sizeof(int) * 4, hipHostMallocDefault));
        HPQ_HIP(hipMalloc(&c->d_red, sizeof(double) * 8));
        HPQ_HIP(hipHostMalloc(&c->h_red, sizeof(double) * 8, hipHostMallocDefault));
        HPQ_HIP(hipMemset(c->d_count, 0, sizeof(int) * 4));
        poisson_init(c);
        mg_init(c);
        HPQ_HIP(hipDeviceSynchronize());
    });
    if (rc != HPQ_OK) {
        // keep the context so the caller can read the message; it owns partial state
        *out = c;
        return rc;
    }
    *out = c;
    return HPQ_OK;
}

void hpq_destroy(hpq_ctx* c) {
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    for (Trace& t : c->trace) {
        for (hipEvent_t e : t.start) (void)hipEventDestroy(e);
        for (hipEvent_t e : t.stop) (void)hipEventDestroy(e);
    }
    poisson_free(c);
    mg_free(c);
    bins_free(c);
    cell_lists_free(c->cl_plasma);
    cell_lists_free(c->cl_beam);
    init_scratch_free(c);
    if (c->relb_rows) (void)hipFree(c->relb_rows);
    if (c->diag_partial) (void)hipFree(c->diag_partial);
    if (c->diag_comp_of) (void)hipFree(c->diag_comp_of);
    if (c->d_count) (void)hipFree(c->d_count);
    if (c->h_count) (void)hipHostFree(c->h_count);
    if (c->d_red) (void)hipFree(c->d_red);
    if (c->h_red) (void)hipHostFree(c->h_red);
    adjoint_free(c);
    beam_slip_free(c);
    insitu_free(c);
    salame_free(c);
    for (hipEvent_t e : {c->ev_start, c->ev_beam, c->ev_fields, c->ev_end})
        if (e) (void)hipEventDestroy(e);
    if (c->side) (void)hipStreamDestroy(c->side);
    delete c;
}

const char* hpq_last_error(const hpq_ctx* c) { return c ? c->err.c_str() : "null context"; }

int hpq_set_stream(hpq_ctx* c, void* s) {
    if (!c) return HPQ_E_INVALID;
    c->stream = static_cast<hipStream_t>(s);
    return HPQ_OK;
}

int hpq_layout(const hpq_ctx* c, int32_t* out) {
    if (!c || !out) return HPQ_E_INVALID;
    const Layout& L = c->L;
    const int v[HPQ_LAYOUT_N] = {L.ncomp, L.next_jx_beam, L.next_jy_beam, L.chi, L.Sy, L.Sx, L.ExmBy, L.EypBx,
                                 L.Ez, L.Bx, L.By, L.Bz, L.Psi, L.jx_beam, L.jy_beam, L.jz_beam, L.jx, L.jy,
                                 L.rhomjz, L.rho, L.prev_jx_beam, L.prev_jy_beam, L.ions_rhomjz, L.next_jx,
                                 L.next_jy, L.jz, L.prev_Bx, L.prev_By, L.prev_jx, L.prev_jy, L.pciter_Bx,
                                 L.pciter_By, L.sal_Ez_target, L.sal_Ez_no_salame, L.sal_Ez, L.sal_jx,
                                 L.sal_jy, L.sal_jz_beam, L.sal_Bx, L.sal_By, L.sal_Sy, L.sal_Sx,
                                 L.sal_Sy_back, L.sal_Sx_back};
    for (int k = 0; k < HPQ_LAYOUT_N; ++k) out[k] = v[k];
    return HPQ_OK;
}

int64_t hpq_store_doubles(const hpq_ctx* c) { return c ? int64_t(c->L.ncomp) * c->g.plane : -1; }

int hpq_deposit_current(hpq_ctx* c, const hpq_plasma* p, double* store, const int32_t* comps, int ions,
                        int64_t* n_qsa) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_plasma(p);
        if (!comps) throw HpqError(HPQ_E_INVALID, "comps is null");
        for (int k = 0; k < 6; ++k) check_comp(c, comps[k]);
        if (n_qsa) HPQ_HIP(hipMemsetAsync(c->d_count, 0, sizeof(int), c->stream));
        deposit_current(c, p, store, comps, ions != 0);
        if (n_qsa) {
            HPQ_HIP(hipMemcpyAsync(c->h_count, c->d_count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HPQ_HIP(hipStreamSynchronize(c->stream));
            *n_qsa = c->h_count[0];
        }
    });
}

int hpq_explicit_deposition(hpq_ctx* c, const hpq_plasma* p, double* store) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_plasma(p);
        if (!c->p.explicit_solver) throw HpqError(HPQ_E_UNSUPPORTED, "explicit deposition needs the explicit solver");
        explicit_deposition(c, p, store);
    });
}

int hpq_deposit_current_slice(hpq_ctx* c, const hpq_beam* b, int slice, double* store, const int32_t* comps) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        if (!comps) throw HpqError(HPQ_E_INVALID, "comps is null");
        for (int k = 0; k < 3; ++k) check_comp(c, comps[k]);
        deposit_beam(c, b, slice, store, comps);
    });
}

int hpq_advance_plasma_particles(hpq_ctx* c, const hpq_plasma* p, const double* store, int temp_slice) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_plasma(p);
        advance_plasma(c, p, store, temp_slice != 0);
    });
}

static void check_adj(const hpq_plasma_adj* a, bool half_step) {
    if (!a || !a->x || !a->y || !a->ux || !a->uy || !a->psi) throw HpqError(HPQ_E_INVALID, "null adjoint array");
    if (half_step && (!a->ux_half_step || !a->uy_half_step || !a->psi_half_step))
        throw HpqError(HPQ_E_INVALID, "null adjoint half-step array");
}

int hpq_deposit_current_vjp(hpq_ctx* c, const hpq_plasma* p, const double* store_adj, const int32_t* comps,
                            hpq_plasma_adj* adj) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store_adj);
        check_plasma(p);
        check_adj(adj, false);
        if (!comps) throw HpqError(HPQ_E_INVALID, "null comps");
        for (int k = 0; k < 6; ++k)
            if (comps[k] >= c->L.ncomp) throw HpqError(HPQ_E_INVALID, "component index out of range");
        deposit_vjp(c, p, store_adj, comps, adj);
    });
}

int hpq_explicit_deposition_vjp(hpq_ctx* c, const hpq_plasma* p, const double* store, double* store_adj,
                                hpq_plasma_adj* adj) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_store(c, store_adj);
        check_plasma(p);
        check_adj(adj, false);
        if (!c->p.explicit_solver) throw HpqError(HPQ_E_UNSUPPORTED, "ExplicitDeposition needs the explicit layout");
        explicit_deposition_vjp(c, p, store, store_adj, adj);
    });
}

int hpq_advance_plasma_particles_vjp(hpq_ctx* c, const hpq_plasma* p, const double* store,
                                     const hpq_plasma_adj* out_adj, hpq_plasma_adj* in_adj, double* store_adj) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_store(c, store_adj);
        check_plasma(p);
        check_adj(out_adj, true);
        check_adj(in_adj, true);
        advance_plasma_vjp(c, p, store, out_adj, in_adj, store_adj);
    });
}

static void check_beam_state(const hpq_ctx* c, const hpq_beam_state* b) {
    if (!b) throw HpqError(HPQ_E_INVALID, "beam state is null");
    if (b->n > 0 && (!b->x || !b->y || !b->z || !b->ux || !b->uy || !b->uz || !b->w || !b->valid || !b->nsubcycles))
        throw HpqError(HPQ_E_INVALID, "beam state has a null array");
    if (!b->off) throw HpqError(HPQ_E_INVALID, "beam state has no slice offsets");
    if (b->off[0] != 0 || b->off[c->g.nz] > b->n) throw HpqError(HPQ_E_INVALID, "beam slice offsets exceed n");
    if (!(b->mass != 0.0)) throw HpqError(HPQ_E_INVALID, "beam mass is zero");
}

int hpq_beam_begin_step(hpq_ctx* c, hpq_beam_state* b) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_beam_state(c, b);
        beam_begin_step(c, b);
    });
}

int hpq_advance_beam_slice(hpq_ctx* c, hpq_beam_state* b, int slice, const double* store, double dt_step,
                           int n_subcycles, int do_z_push) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_beam_state(c, b);
        check_store(c, store);
        if (slice < 0 || slice >= c->g.nz) throw HpqError(HPQ_E_INVALID, "slice index out of range");
        if (n_subcycles < 1) throw HpqError(HPQ_E_INVALID, "n_subcycles must be >= 1");
        advance_beam(c, b, slice, store, dt_step, n_subcycles, do_z_push != 0);
    });
}

int hpq_sort_beam_slices(hpq_ctx* c, hpq_beam_state* b, int64_t* off_out) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        if (!b || !off_out) throw HpqError(HPQ_E_INVALID, "null pointer");
        if (b->n > 0 && (!b->x || !b->y || !b->z || !b->ux || !b->uy || !b->uz || !b->w || !b->valid || !b->nsubcycles))
            throw HpqError(HPQ_E_INVALID, "beam state has a null array");
        sort_beam_slices(c, b, off_out);
    });
}

int hpq_run_step_with_beam(hpq_ctx* c, const hpq_plasma* p, hpq_beam_state* b, double* store, double dt_step,
                           int n_subcycles, int do_z_push) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_plasma(p);
        check_beam_state(c, b);
        if (n_subcycles < 1) throw HpqError(HPQ_E_INVALID, "n_subcycles must be >= 1");
        const hpq_beam view{b->x, b->y, b->z, b->ux, b->uy, b->uz, b->w, b->valid, b->off, b->n, b->charge, b->mass};
        beam_begin_step(c, b);
        Hipace h(c, p, &view, store);
        for (int isl = c->g.nz - 1; isl >= 0; --isl) {
            h.SolveOneSlice(isl);
            // AdvanceBeamParticlesSlice after the plasma push (src/Hipace.cpp:700-709)
            advance_beam(c, b, isl, store, dt_step, n_subcycles, do_z_push != 0);
        }
        diag_finalize(c);
    });
}

int hpq_adjoint_forward(hpq_ctx* c, const hpq_plasma* p, const hpq_beam* b, double* store, int checkpoint_every) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_plasma(p);
        if (checkpoint_every < 1) throw HpqError(HPQ_E_INVALID, "checkpoint_every must be >= 1");
        adjoint_forward(c, p, b, store, std::min(checkpoint_every, c->g.nz));
    });
}

int hpq_adjoint_backward(hpq_ctx* c, const hpq_plasma* p, const hpq_beam* b, double* store, const double* ez_seed,
                         const double* jz_beam_seed, double* beam_w_adj, double mg_tol) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_plasma(p);
        if (!(mg_tol > 0.0)) throw HpqError(HPQ_E_INVALID, "mg_tol must be positive");
        if (jz_beam_seed && (!b || !beam_w_adj))
            throw HpqError(HPQ_E_INVALID, "a jz_beam seed needs the beam and its weight adjoints");
        adjoint_backward(c, p, b, store, ez_seed, jz_beam_seed, beam_w_adj, mg_tol);
    });
}

int hpq_poisson_dirichlet(hpq_ctx* c, const double* staging, double* lhs) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        if (!staging || !lhs) throw HpqError(HPQ_E_INVALID, "null pointer");
        stage_rhs(c, staging);
        double* planes[3] = {lhs, lhs, lhs};
        poisson_solve_staged(c, 1, planes);
    });
}

int hpq_solve_poisson_psi_exmby_eypbx_ez_bz(hpq_ctx* c, double* store) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        Fields(c, store).SolvePoissonPsiExmByEypBxEzBz();
    });
}

int hpq_solve_poisson_bxby(hpq_ctx* c, double* store) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        if (c->p.explicit_solver) throw HpqError(HPQ_E_UNSUPPORTED, "SolvePoissonBxBy needs the predictor-corrector layout");
        Fields(c, store).SolvePoissonBxBy();
    });
}

int hpq_initialize_sxsy_with_beam(hpq_ctx* c, double* store) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        if (!c->p.explicit_solver) throw HpqError(HPQ_E_UNSUPPORTED, "Sx/Sy exist only with the explicit solver");
        sxsy_from_beam(c, store);
    });
}

int hpq_mg_solve1(hpq_ctx* c, double* sol, const double* rhs, const double* acf, int64_t comp_stride, double tol_rel,
                  double tol_abs, int maxiter, int* iters) {
    if (!c) return HPQ_E_INVALID;
    int st = HPQ_OK;
    const int rc = guarded(c, [&] {
        if (!sol || !rhs || !acf) throw HpqError(HPQ_E_INVALID, "null pointer");
        if (comp_stride < c->g.plane) throw HpqError(HPQ_E_INVALID, "comp_stride smaller than a plane");
        st = hpmg::MultiGrid(c).solve1(sol, rhs, acf, comp_stride, tol_rel, tol_abs, maxiter, iters);
        if (st == HPQ_E_MG_NOT_CONVERGED) c->err = "hpmg failed";
        if (st == HPQ_E_MG_DIVERGED) c->err = "hpmg failing so lets stop here";
    });
    return rc != HPQ_OK ? rc : st;
}

int64_t hpq_plasma_count(const hpq_ctx* c, const hpq_plasma_profile* prof) {
    if (!c || !prof) return -1;
    hpq_ctx* mc = const_cast<hpq_ctx*>(c);
    int64_t n = -1;
    const int rc = guarded(mc, [&] {
        if (prof->ppc[0] < 1 || prof->ppc[1] < 1) throw HpqError(HPQ_E_INVALID, "ppc must be >= 1");
        n = plasma_count(mc, prof);
    });
    return rc == HPQ_OK ? n : -1;
}

int hpq_init_plasma(hpq_ctx* c, const hpq_plasma_profile* prof, const hpq_plasma* p) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        if (!prof) throw HpqError(HPQ_E_INVALID, "profile is null");
        check_plasma(p);
        if (prof->ppc[0] < 1 || prof->ppc[1] < 1) throw HpqError(HPQ_E_INVALID, "ppc must be >= 1");
        init_plasma(c, prof, p);
        c->bins.stale = true;
    });
}

int hpq_bin_tiles(hpq_ctx* c, const hpq_plasma* p, int tile_size, int32_t* perm, int32_t* offsets) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_plasma(p);
        if (!perm && p->n > 0) throw HpqError(HPQ_E_INVALID, "perm is null");
        bin_tiles(c, p, tile_size, perm, offsets);
    });
}

int hpq_bin_tiles_noalloc(hpq_ctx* c, const hpq_plasma* p, int tile_size, int32_t* perm, int32_t* offsets) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_plasma(p);
        if (!perm && p->n > 0) throw HpqError(HPQ_E_INVALID, "perm is null");
        bin_tiles_noalloc(c, p, tile_size, perm, offsets);
    });
}

int hpq_begin_step(hpq_ctx* c, const hpq_plasma* p, double* store) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_plasma(p);
        HPQ_HIP(hipMemsetAsync(store, 0, sizeof(double) * c->L.ncomp * c->g.plane, c->stream));
        HPQ_HIP(hipMemsetAsync(c->d_count, 0, sizeof(int) * 4, c->stream));
        for (double& s : c->stats) s = 0;
        c->bins.stale = true;
        c->bins.reordered = false;   // a step starts from the plasma in injection order
        c->mg.hist_slice = -1;       // no Bx, By history across steps
        if (c->L.ions_rhomjz >= 0) DepositCurrent(c, p, store, -1, -1, -1, -1, -1, c->L.ions_rhomjz, true);
    });
}

int hpq_solve_one_slice(hpq_ctx* c, const hpq_plasma* p, const hpq_beam* b, double* store, int islice) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_plasma(p);
        if (islice < 0 || islice >= c->g.nz) throw HpqError(HPQ_E_INVALID, "slice index out of range");
        Hipace(c, p, b, store).SolveOneSlice(islice);
        diag_finalize(c);
    });
}

int hpq_run_slices(hpq_ctx* c, const hpq_plasma* p, const hpq_beam* b, double* store, int first, int last) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        check_store(c, store);
        check_plasma(p);
        if (first >= c->g.nz || last < 0 || first < last) throw HpqError(HPQ_E_INVALID, "bad slice range");
        Hipace h(c, p, b, store);
        const bool two = c->p.explicit_solver && first > last;
        if (two) side_init(c);
        for (int isl = first; isl >= last; --isl) h.SolveOneSlice(isl, isl != first, isl != last, {}, two);
        diag_finalize(c);
    });
}

int hpq_set_checksums(hpq_ctx* c, double* checksums) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        diag_finalize(c);
        c->checksums = checksums;
    });
}

int hpq_set_insitu(hpq_ctx* c, const hpq_insitu* d) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        if (!d) {
            c->insitu = hpq_insitu{};
            return;
        }
        if (d->field_rdata && !d->field_sum) throw HpqError(HPQ_E_INVALID, "insitu: field_sum missing");
        if (d->field_rdata && !c->p.explicit_solver)
            throw HpqError(HPQ_E_UNSUPPORTED, "Must use explicit solver for field insitu diagnostic");
        if (d->plasma_rdata && (!d->plasma_idata || !d->plasma_sum || !d->plasma_isum))
            throw HpqError(HPQ_E_INVALID, "insitu: plasma arrays missing");
        if (d->beam_rdata && (!d->beam_idata || !d->beam_sum || !d->beam_isum))
            throw HpqError(HPQ_E_INVALID, "insitu: beam arrays missing");
        if ((d->plasma_rdata && !(d->plasma_radius >= 0.0)) || (d->beam_rdata && !(d->beam_radius >= 0.0)))
            throw HpqError(HPQ_E_INVALID, "insitu: radius must be >= 0");
        if (d->field_rdata || d->plasma_rdata || d->beam_rdata) insitu_reserve(c);
        c->insitu = *d;
    });
}

int hpq_set_salame(hpq_ctx* c, const hpq_salame* s) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        c->sal_last_slice = -1;
        c->sal_overloaded = false;
        if (!s) {
            c->salame = hpq_salame{};
            return;
        }
        if (!c->p.salame || !c->p.explicit_solver)
            throw HpqError(HPQ_E_UNSUPPORTED, "SALAME needs hpq_params.salame = 1 and the explicit solver");
        if (!s->mask || !s->w || !s->valid) throw HpqError(HPQ_E_INVALID, "SALAME: mask, w and valid are required");
        if (s->n_iter < 0) throw HpqError(HPQ_E_INVALID, "SALAME: n_iter must be >= 0");
        salame_alloc(c);
        c->salame = *s;
    });
}

int hpq_add_capture(hpq_ctx* c, int comp, double* dst) {
    if (!c || !dst || comp < 0 || comp >= c->L.ncomp) return HPQ_E_INVALID;
    c->captures.push_back({comp, dst});
    return HPQ_OK;
}

int hpq_clear_captures(hpq_ctx* c) {
    if (!c) return HPQ_E_INVALID;
    c->captures.clear();
    return HPQ_OK;
}

int hpq_trace_enable(hpq_ctx* c, int kernel, int on) {
    if (!c || kernel < 0 || kernel >= HPQ_TRACE_N) return HPQ_E_INVALID;
    c->trace[kernel].on = on != 0;
    return HPQ_OK;
}

int hpq_trace_read(hpq_ctx* c, int kernel, double* total_ms, int64_t* launches) {
    if (!c || kernel < 0 || kernel >= HPQ_TRACE_N || !total_ms || !launches) return HPQ_E_INVALID;
    return guarded(c, [&] {
        HPQ_HIP(hipStreamSynchronize(c->stream));
        Trace& t = c->trace[kernel];
        double sum = 0.0;
        for (size_t k = 0; k < t.used; ++k) {
            float ms = 0.f;
            HPQ_HIP(hipEventElapsedTime(&ms, t.start[k], t.stop[k]));
            sum += ms;
        }
        *total_ms = sum;
        *launches = int64_t(t.used);
    });
}

int hpq_trace_reset(hpq_ctx* c) {
    if (!c) return HPQ_E_INVALID;
    return guarded(c, [&] {
        HPQ_HIP(hipStreamSynchronize(c->stream));
        for (Trace& t : c->trace) t.used = 0;
    });
}

int hpq_stats(const hpq_ctx* c, double* out4) {
    if (!c || !out4) return HPQ_E_INVALID;
    hpq_ctx* mc = const_cast<hpq_ctx*>(c);
    const int rc = guarded(mc, [&] {
        HPQ_HIP(hipMemcpyAsync(mc->h_count, mc->d_count, sizeof(int), hipMemcpyDeviceToHost, mc->stream));
        HPQ_HIP(hipStreamSynchronize(mc->stream));
    });
    out4[0] = rc == HPQ_OK ? double(mc->h_count[0]) : -1.0;
    out4[1] = c->stats[1];
    out4[2] = c->stats[2];
    out4[3] = c->stats[3];
    return rc;
}

}  // extern "C"
