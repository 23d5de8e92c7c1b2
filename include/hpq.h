/* hpq — MI355X-native per-slice quasi-static loop (gfx950 HIP), C ABI.
 *
 * This is the drop-in boundary for HiPACE++'s per-slice loop. Each entry point
 * replaces one reference function (cited below as path:line inside the reference
 * source tree). All particle and field buffers are caller-owned device pointers;
 * the context owns only scratch (DST work arrays, multigrid hierarchy, counters).
 * Calls are ordered on the context's HIP stream (hpq_set_stream); entry points that
 * must read back a scalar (MG convergence, QSA count) synchronise that stream.
 * Return value: 0 (HPQ_OK) or an HPQ_E* code; hpq_last_error() gives the message.
 * A context is not re-entrant: use one context per stream.
 *
 * Slice store layout (the reference's Fields::m_slices, src/fields/Fields.cpp:50-220):
 * ncomp planes of (nx+4) x (ny+4) doubles, x fastest, 2 guard cells on each side,
 * plane k at store + k*(nx+4)*(ny+4). Component indices: hpq_layout().
 */
#ifndef HPQ_H
#define HPQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hpq_ctx hpq_ctx;

enum {
    HPQ_OK = 0,
    HPQ_E_INVALID = 1,          /* bad argument or size */
    HPQ_E_HIP = 2,              /* HIP runtime error */
    HPQ_E_MG_NOT_CONVERGED = 3, /* hpmg: Failed to converge (HpMultiGrid.cpp:1408-1416) */
    HPQ_E_MG_DIVERGED = 4,      /* hpmg: failing (HpMultiGrid.cpp:1397-1405) */
    HPQ_E_UNSUPPORTED = 5,      /* option outside this path */
    HPQ_E_NOMEM = 6
};

enum { HPQ_BC_REFLECTING = 0, HPQ_BC_PERIODIC = 1, HPQ_BC_ABSORBING = 2 };

/* Arithmetic of the particle kernels: FP64 throughout (the drop-in path), or FP32 gather,
 * push increments and deposition (FP32 contributions summed in FP64, FP64 slice store and particle state,
 * FP64 multigrid). */
enum { HPQ_PREC_FP64 = 0, HPQ_PREC_MIXED = 1 };

/* Order of the FP64 deposition sums (plasma current, explicit Sx/Sy, beam current). ORDERED:
 * the order of the reference CPU build's tiled deposition (32-cell tiles in four colours,
 * particles by index inside a tile; src/particles/deposition/DepositionUtil.H:204-253), every
 * cell summed by one thread, no atomics: results repeat bit for bit, and equal a CPU build that
 * sums in that order. ATOMIC: LDS windows flushed with global atomics, sums in hardware order
 * (differences of a few ulp per cell between runs). The mixed-precision path always uses the
 * atomic kernels. */
enum { HPQ_DEPOSIT_ORDERED = 0, HPQ_DEPOSIT_ATOMIC = 1 };

/* Window sums of the mixed path's FP32 depositions: each FP32 contribution added to an FP64 LDS
 * word (SUM_FP64, exact to FP64 rounding), or to an FP32 word with a second FP32 word that
 * collects the rounding error of every addition (SUM_KAHAN32: the compensated FP32 tile sums;
 * the error of each atomic addition is recovered exactly from its returned old value). */
enum { HPQ_SUM_FP64 = 0, HPQ_SUM_KAHAN32 = 1 };

/* Initial guess of the explicit solver's multigrid for Bx, By. PREVIOUS: the previous
 * slice's solution, which the slice store still holds (the reference: Hipace.cpp:920 solves
 * in place). LINEAR: extrapolated from the two previous slices, 2 B(i+1) - B(i+2), as
 * Fields::InitialBfieldGuess (Fields.cpp:1150-1171) extrapolates for the predictor-corrector
 * loop. DST: the previous slice's solution corrected mg_guess_iters times by the exact
 * solve of the same system with chi replaced by its mean over the slice,
 * B += (Lap - <chi>)^-1 [S - (Lap - chi) B], each a shifted FFT DST solve (needs nx + 1,
 * ny + 1 powers of two). LINEAR and DST change only where the multigrid starts: it still
 * iterates to its tolerance, so results agree with PREVIOUS's within that tolerance. */
enum { HPQ_MG_GUESS_PREVIOUS = 0, HPQ_MG_GUESS_LINEAR = 1, HPQ_MG_GUESS_DST = 2 };

/* Transverse Dirichlet Poisson solver (fields.poisson_solver, src/fields/Fields.cpp:179-208):
 * DST-I by FFT in FP64, or DST-I as dense sine-matrix GEMMs on the matrix cores with each
 * FP32 operand split into two BF16 terms (hi + lo; three products hi*hi + hi*lo + lo*hi,
 * FP32 accumulation), about 1e-5 relative. */
enum { HPQ_POISSON_FFT = 0, HPQ_POISSON_BF16X3 = 1 };

/* Run parameters: the reference input keys this path honours
 * (src/Hipace.cpp:97-213, src/Hipace.H:208-252). Units are normalised (c=ep0=mu0=1). */
typedef struct hpq_params {
    int nx, ny, nz;                      /* amr.n_cell */
    double xlo, xhi, ylo, yhi, zlo, zhi; /* geometry.prob_lo / prob_hi */
    int depos_order;                     /* hipace.depos_order_xy: 0..3 */
    int derivative_type;                 /* hipace.depos_derivative_type: 0..2 */
    int explicit_solver;                 /* hipace.bxby_solver: 1 explicit, 0 predictor-corrector */
    int deposit_rho;                     /* allocate and deposit rho */
    int do_beam_jx_jy;                   /* hipace.do_beam_jx_jy_deposition */
    int neutralize;                      /* plasma.neutralize_background */
    double mg_tol_rel, mg_tol_abs;       /* hipace.MG_tolerance_rel / _abs */
    int mg_maxiter;                      /* 200 in Hipace::ExplicitMGSolveBxBy */
    int particle_bc;                     /* boundary.particle: HPQ_BC_* */
    double plo[2], phi[2];               /* boundary.particle_lo / particle_hi */
    double pc_tol;                       /* hipace.predcorr_B_error_tolerance */
    int pc_maxiter;                      /* hipace.predcorr_max_iterations */
    double pc_mix;                       /* hipace.predcorr_B_mixing_factor */
    double plasma_charge, plasma_mass;   /* species charge and mass */
    double max_qsa;                      /* plasma.max_qsa_weighting_factor */
    int bin_period;                      /* atomic deposition without reordering: walks the
                                            particles through a permutation sorted by 32x8-cell
                                            tiles, re-sorted every bin_period slices; 0: index
                                            order. Changes only the order of floating-point
                                            sums. */
    int precision;                       /* HPQ_PREC_FP64 or HPQ_PREC_MIXED */
    int poisson_solver;                  /* HPQ_POISSON_FFT or HPQ_POISSON_BF16X3 */
    int salame;                          /* a beam has <beam>.do_salame: allocate the SALAME
                                            planes (explicit solver only) */
    int deposition;                      /* HPQ_DEPOSIT_ORDERED or HPQ_DEPOSIT_ATOMIC */
    int tile_size;                       /* hipace.tile_size: tile width of the ordered deposition
                                            (src/Hipace.cpp:160); 0: 32, the reference's default;
                                            at least the widest deposition stencil */
    int reorder_period;                  /* plasmas.reorder_period (PlasmaParticleContainer.cpp:
                                            197-208, called at Hipace.cpp:594-595): every slice
                                            with islice % reorder_period == 0, the plasma arrays
                                            are reordered in place by deposition tile (32x8
                                            cells), so the depositions read them contiguously;
                                            0: never (the reference's default). Applied with
                                            atomic deposition only: the ordered deposition keeps
                                            the index order, whose sums it reproduces. */
    int mixed_sums;                      /* HPQ_SUM_FP64 or HPQ_SUM_KAHAN32 (mixed path) */
    int mg_guess;                        /* HPQ_MG_GUESS_PREVIOUS, _LINEAR or _DST */
    int mg_guess_iters;                  /* HPQ_MG_GUESS_DST: DST corrections per slice (>= 1) */
} hpq_params;

/* Plasma species, SoA device arrays in PlasmaIdx order
 * (src/particles/plasma/PlasmaParticleContainer.H:21-46). valid: 1 live, 0 removed. */
typedef struct hpq_plasma {
    double *x, *y, *w, *ux, *uy, *psi;
    double *x_prev, *y_prev, *ux_half_step, *uy_half_step, *psi_half_step;
    int32_t *valid;
    int64_t n;
} hpq_plasma;

/* Beam particles grouped by ζ-slice (BoxSort.cpp:14-78): the particles of slice s are
 * [off[s], off[s+1]). Device arrays; `off` is a host array of nz+1 offsets. */
typedef struct hpq_beam {
    const double *x, *y, *z, *ux, *uy, *uz, *w;
    const int32_t *valid;
    const int64_t *off;
    int64_t n;
    double charge, mass;
} hpq_beam;

/* Beam particles the beam push may move: hpq_beam's arrays, mutable, plus the sub-cycle
 * counter of every particle (BeamIdx::nsubcycles, src/particles/beam/BeamParticleContainer.H:23-41).
 * off stays a host array: particles of slice s are [off[s], off[s+1]). */
typedef struct hpq_beam_state {
    double *x, *y, *z, *ux, *uy, *uz, *w;
    int32_t *valid, *nsubcycles;
    const int64_t *off;
    int64_t n;
    double charge, mass;
    /* beams.external_E(x,y,z,t) / external_B(x,y,z,t) (BeamParticleContainer.cpp:72-87), added
     * to the gathered fields by the push as ApplyExternalField does
     * (src/particles/pusher/ExternalFields.H:28-56), for fields affine in position: component k
     * of (Ex, Ey, Ez, Bx, By, Bz) is ((e[4k] + e[4k+1]*x) + e[4k+2]*y) + e[4k+3]*z.
     * use_external_fields = 0: none. */
    double external_fields[24];
    int use_external_fields;
} hpq_beam_state;

/* Adjoints of plasma quantities for the vector-Jacobian products below: device arrays of
 * n doubles (n of the matching hpq_plasma), accumulated with +=. */
typedef struct hpq_plasma_adj {
    double *x, *y, *ux, *uy, *psi, *ux_half_step, *uy_half_step, *psi_half_step;
} hpq_plasma_adj;

/* Plasma injection profile: n(x,y) = density * (1 + ramp_a (x^2+y^2)) inside `radius`, or,
 * when density_table is not NULL, the density the caller evaluated at every lattice position
 * (<plasma>.density(x,y,z), a parser expression in the reference,
 * src/particles/plasma/PlasmaParticleContainer.cpp:90-97): device array of nx*ny*ppc[0]*ppc[1]
 * doubles, entry (j*nx + i)*ppc[0]*ppc[1] + i_part for the particle i_part = ix + iy*ppc[0] of
 * cell (i, j), at x = xlo + (i + (0.5 + ix)/ppc[0])*dx, y likewise
 * (ParticleUtil.H:65-104). Positions where the density is <= 0 inject nothing. */
typedef struct hpq_plasma_profile {
    int ppc[2];
    double density, ramp_a, radius;
    double u_mean[3];
    const double *density_table;
} hpq_plasma_profile;

/* Number of hpq_layout() entries. */
#define HPQ_LAYOUT_N 44

/* ---- context ---------------------------------------------------------------- */
int hpq_create(const hpq_params *p, hpq_ctx **out);
void hpq_destroy(hpq_ctx *ctx);
const char *hpq_last_error(const hpq_ctx *ctx);
int hpq_set_stream(hpq_ctx *ctx, void *hip_stream);
/* out[HPQ_LAYOUT_N]: ncomp, next_jx_beam, next_jy_beam, chi, Sy, Sx, ExmBy, EypBx, Ez, Bx,
 * By, Bz, Psi, jx_beam, jy_beam, jz_beam, jx, jy, rhomjz, rho, prev_jx_beam, prev_jy_beam,
 * ions_rhomjz, next_jx, next_jy, jz, prev_Bx, prev_By, prev_jx, prev_jy, pciter_Bx, pciter_By,
 * then the SALAME planes (WhichSlice::Salame, src/fields/Fields.cpp:111-115): Ez_target,
 * Ez_no_salame, Ez, jx, jy, jz_beam, Bx, By, Sy, Sx, Sy_back, Sx_back
 * (-1: not allocated). Replaces the Comps map, src/fields/Fields.H:27-56. */
int hpq_layout(const hpq_ctx *ctx, int32_t *out);
int64_t hpq_store_doubles(const hpq_ctx *ctx);

/* ---- per-slice operations (reference free functions and methods) -------------- */

/* ::DepositCurrent, src/particles/deposition/PlasmaDepositCurrent.cpp:22-259.
 * comps[6] = jx, jy, jz, rho, chi, rhomjz component indices (-1: skip). ions != 0 flips
 * the charge (RhomJzIons background). QSA violators are invalidated; their count is
 * written to *n_qsa (host) when n_qsa is not NULL (synchronises the stream). */
int hpq_deposit_current(hpq_ctx *ctx, const hpq_plasma *p, double *store, const int32_t *comps,
                        int ions, int64_t *n_qsa);

/* ::ExplicitDeposition, src/particles/deposition/ExplicitDeposition.cpp:20-265:
 * adds the plasma part of Sy, Sx from Bz, Ez, ExmBy, EypBx. */
int hpq_explicit_deposition(hpq_ctx *ctx, const hpq_plasma *p, double *store);

/* ::DepositCurrentSlice, src/particles/deposition/BeamDepositCurrent.cpp:22-196:
 * beam particles of slice `slice` deposit into comps[3] = jx, jy, jz (-1: skip). */
int hpq_deposit_current_slice(hpq_ctx *ctx, const hpq_beam *b, int slice, double *store,
                              const int32_t *comps);

/* AdvancePlasmaParticles, src/particles/pusher/PlasmaParticleAdvance.cpp:29-240
 * (leapfrog, gather FieldGather.H:45-96, EnforceBC GetAndSetPosition.H:29-99). */
int hpq_advance_plasma_particles(hpq_ctx *ctx, const hpq_plasma *p, const double *store,
                                 int temp_slice);

/* ---- adjoint: vector-Jacobian products of the slice operations -----------------------
 * No counterpart in the reference (its slice loop has no derivative); they make the
 * adjoint slice loop. Adjoint planes use the slice-store layout: store_adj[c] holds the
 * adjoint of component c. Evaluated in FP64 at the given state; branches (QSA removal,
 * particle boundaries, shape-factor pieces) are held fixed. */

/* Adjoint of hpq_deposit_current (ions = 0): adds to adj->x, y, ux, uy, psi the derivative
 * of sum_c <store_adj[comps[c]], deposited plane c>. */
int hpq_deposit_current_vjp(hpq_ctx *ctx, const hpq_plasma *p, const double *store_adj,
                            const int32_t *comps, hpq_plasma_adj *adj);

/* Adjoint of hpq_explicit_deposition: reads the Sy, Sx planes of store_adj; adds to its
 * Bz, Ez, ExmBy, EypBx planes and to adj->x, y, ux, uy, psi. */
int hpq_explicit_deposition_vjp(hpq_ctx *ctx, const hpq_plasma *p, const double *store,
                                double *store_adj, hpq_plasma_adj *adj);

/* Adjoint of hpq_advance_plasma_particles (temp_slice = 0) at the state p before the push.
 * out_adj: adjoints of the outputs x (also x_prev), y (also y_prev), ux/uy/psi_half_step
 * and ux, uy, psi. Adds the adjoints of the inputs x_prev, y_prev (to in_adj->x, y) and
 * of the half-step values (to in_adj->*_half_step), and those of the Psi, Ez, Bx, By, Bz
 * planes to store_adj. */
int hpq_advance_plasma_particles_vjp(hpq_ctx *ctx, const hpq_plasma *p, const double *store,
                                     const hpq_plasma_adj *out_adj, hpq_plasma_adj *in_adj,
                                     double *store_adj);

/* Adjoint slice loop, step 1: after hpq_begin_step, runs the slices nz-1 .. 0 as
 * hpq_run_slices does, keeping the slice store and the plasma every checkpoint_every slices
 * (device memory held by the context). Explicit solver, odd nx and ny, do_beam_jx_jy = 0. */
int hpq_adjoint_forward(hpq_ctx *ctx, const hpq_plasma *p, const hpq_beam *b, double *store,
                        int checkpoint_every);

/* Adjoint slice loop, step 2: for a loss L(Ez, jz_beam) with ez_seed = dL/dEz and
 * jz_beam_seed = dL/d(jz_beam) on every slice (each nz*ny*nx doubles, [slice][j][i] over the
 * valid box, the planes hpq_add_capture records; null: zero), adds dL/dw of every beam
 * macro-particle weight to beam_w_adj (b->n doubles). A jz_beam seed differentiates
 * quantities such as the beam's energy gain, sum Ez*jz_beam dx dy dz per slice (the in-situ
 * [Ez*jz_beam], src/fields/Fields.cpp:1320-1346). Marches the slices in reverse,
 * recomputing each checkpoint segment; the plasma and the store are scratch afterwards.
 * mg_tol: relative tolerance of the adjoint Bx/By solves. */
int hpq_adjoint_backward(hpq_ctx *ctx, const hpq_plasma *p, const hpq_beam *b, double *store,
                         const double *ez_seed, const double *jz_beam_seed, double *beam_w_adj,
                         double mg_tol);

/* BeamParticleContainer::initializeSlice's counter reset and an empty slipped-particle list,
 * before the beam pushes of a step (src/particles/beam/BeamParticleContainer.cpp:353-390). */
int hpq_beam_begin_step(hpq_ctx *ctx, hpq_beam_state *b);

/* AdvanceBeamParticlesSlice (src/particles/pusher/BeamParticleAdvance.cpp:20-336; no radiation
 * reaction, spin tracking or external fields) followed by shiftSlippedParticles
 * (src/particles/sorting/SliceSort.cpp:13-70): pushes the particles of `slice` and those that
 * slipped into it from the slice above, n_subcycles sub-steps of dt_step / n_subcycles, with
 * the fields of the slice store. A particle that falls below the slice stops; with sub-cycles
 * left it is listed for the next slice's push. Slices are pushed in decreasing order. */
int hpq_advance_beam_slice(hpq_ctx *ctx, hpq_beam_state *b, int slice, const double *store, double dt_step,
                           int n_subcycles, int do_z_push);

/* BoxSorter::sortParticlesByBox (src/particles/sorting/BoxSort.cpp:20-60): sorts the beam
 * arrays in place by zeta-slice (stable), particles outside the domain or invalid last (made
 * invalid), and writes the slice offsets to off_out (host, nz+1 entries). Between steps. */
int hpq_sort_beam_slices(hpq_ctx *ctx, hpq_beam_state *b, int64_t *off_out);

/* One time step with a moving beam: hpq_beam_begin_step, then for slices nz-1 .. 0
 * Hipace::SolveOneSlice followed by AdvanceBeamParticlesSlice (src/Hipace.cpp:557-728), the
 * beam of every slice deposited from its current positions. After hpq_begin_step. */
int hpq_run_step_with_beam(hpq_ctx *ctx, const hpq_plasma *p, hpq_beam_state *b, double *store,
                           double dt_step, int n_subcycles, int do_z_push);

/* FFTPoissonSolverDirichlet*::SolvePoissonEquation
 * (src/fields/fft_poisson_solver/FFTPoissonSolverDirichletDirect.cpp:101-139,
 * ...DirichletFast.cpp:291-328): solves Lap(lhs) = staging with zero Dirichlet values
 * one cell outside the box. staging: nx*ny contiguous; lhs: a guarded plane, valid cells
 * written. */
int hpq_poisson_dirichlet(hpq_ctx *ctx, const double *staging, double *lhs);

/* Fields::SolvePoissonPsiExmByEypBxEzBz, src/fields/Fields.cpp:841-957. */
int hpq_solve_poisson_psi_exmby_eypbx_ez_bz(hpq_ctx *ctx, double *store);

/* Fields::SolvePoissonBxBy, src/fields/Fields.cpp:1009-1078 (into PCIter Bx, By). */
int hpq_solve_poisson_bxby(hpq_ctx *ctx, double *store);

/* Hipace::InitializeSxSyWithBeam, src/Hipace.cpp:745-790. */
int hpq_initialize_sxsy_with_beam(hpq_ctx *ctx, double *store);

/* hpmg::MultiGrid::solve1, src/mg_solver/HpMultiGrid.cpp:1170-1190 + solve_doit
 * :1308-1427: solves -acf*sol + Lap(sol) = rhs (two components comp_stride apart,
 * guarded planes as in the slice store). sol holds the initial guess. *iters (host)
 * receives the number of V-cycles. Synchronises the stream once per V-cycle. */
int hpq_mg_solve1(hpq_ctx *ctx, double *sol, const double *rhs, const double *acf,
                  int64_t comp_stride, double tol_rel, double tol_abs, int maxiter, int *iters);

/* ---- slice loop (Hipace::Evolve / SolveOneSlice) ---------------------------------- */

/* Number of plasma particles the profile injects (PlasmaParticleContainerInit.cpp:143-176);
 * -1 on error. Synchronises the context's stream. */
int64_t hpq_plasma_count(const hpq_ctx *ctx, const hpq_plasma_profile *prof);
/* PlasmaParticleContainer::InitParticles (PlasmaParticleContainerInit.cpp:18-378): fills
 * p (capacity p->n must equal hpq_plasma_count) with the lattice in the CPU build's order:
 * cells x fastest, the particles of a cell in i_part order (the GPU build's injection order
 * after SortParticlesByCell, :372-377). */
int hpq_init_plasma(hpq_ctx *ctx, const hpq_plasma_profile *prof, const hpq_plasma *p);

/* Tile binning of the plasma: the DenseBins of the CPU tiled deposition,
 * src/particles/deposition/DepositionUtil.H:213-227 (key of get_start_cell,
 * PlasmaDepositCurrent.cpp:137-155). perm[p->n] (device) = particle indices sorted stably
 * by key = ((cx+2)/T)*ntile_y + (cy+2)/T, where (cx, cy) is the lowest cell of the particle's
 * order-depos_order stencil and ntile_d = ceil((n_cell_d + 4)/T); invalid particles take
 * key ntile_x*ntile_y. offsets (device, may be NULL) gets ntile_x*ntile_y + 2 entries: the
 * first slot of every bin, then p->n. At most 16383 tiles. */
int hpq_bin_tiles(hpq_ctx *ctx, const hpq_plasma *p, int tile_size, int32_t *perm, int32_t *offsets);

/* hpq_bin_tiles without allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). Same arguments and results as hpq_bin_tiles; its work arrays (tile keys,
 * per-chunk histograms, scan partials) live in one static device array of 32 MiB per device,
 * held from the library's load to the end of the process, instead of context memory. Limits:
 * p->n <= 4194304 particles and (tiles + 1) * ceil(p->n / 2048) <= 4194304; larger inputs
 * return HPQ_E_INVALID. Every call on one device shares that array: calls that overlap in
 * time on one device (several streams or host threads) race. Calls no hipMalloc/hipFree and
 * no hipMemset. */
int hpq_bin_tiles_noalloc(hpq_ctx *ctx, const hpq_plasma *p, int tile_size, int32_t *perm, int32_t *offsets);

/* Start of a time step (src/Hipace.cpp:403-471): zero the slice store and deposit the
 * neutralising ion background of p (MultiPlasma::DepositNeutralizingBackground). */
int hpq_begin_step(hpq_ctx *ctx, const hpq_plasma *p, double *store);

/* Hipace::SolveOneSlice, src/Hipace.cpp:557-728, for slice `islice`. */
int hpq_solve_one_slice(hpq_ctx *ctx, const hpq_plasma *p, const hpq_beam *b, double *store,
                        int islice);

/* Slices first..last (descending) of one step. */
int hpq_run_slices(hpq_ctx *ctx, const hpq_plasma *p, const hpq_beam *b, double *store,
                   int first, int last);

/* Diagnostics at the field-output point of each slice (Hipace::FillFieldDiagnostics):
 * checksums (device, ncomp doubles) accumulates sum |F| over valid cells per component
 * (the reference's checksum, tests/checksum/backend/openpmd_backend.py:40-45); NULL
 * disables. A capture copies the valid region of component `comp` of every slice into
 * dst (device, nz*ny*nx doubles, slice-major). */
int hpq_set_checksums(hpq_ctx *ctx, double *checksums);
int hpq_add_capture(hpq_ctx *ctx, int comp, double *dst);
int hpq_clear_captures(hpq_ctx *ctx);

/* In-situ reduced diagnostics, per slice, into caller-owned device arrays; a group whose
 * rdata is NULL is off. Quantity q of slice s goes to rdata[q*nz + s]; the sums over slices
 * are added (+=) into sum/isum, which the caller clears after writing a record.
 * - fields (Fields::InSituComputeDiags, src/fields/Fields.cpp:1289-1346), after the field
 *   solves: [Ex^2] [Ey^2] [Ez^2] [Bx^2] [By^2] [Bz^2] [ExmBy^2] [EypBx^2] [jz_beam]
 *   [Ez*jz_beam], each summed over valid cells times dx*dy*dz (explicit solver only);
 * - plasma (PlasmaParticleContainer::InSituComputeDiags,
 *   src/particles/plasma/PlasmaParticleContainer.cpp:442-527), at the start of the slice:
 *   sum(w) [x] [x^2] [y] [y^2] [ux] [ux^2] [uy] [uy^2] [uz] [uz^2] [ga] [ga^2]
 *   [(ga-1)*(1-vz)] with quasi-static weights w*ga/psi; idata = Np;
 * - beam (BeamParticleContainer::InSituComputeDiags,
 *   src/particles/beam/BeamParticleContainer.cpp:475-594), before the beam push, over the
 *   particles of the slice: sum(w) [x] [x^2] [y] [y^2] [z] [z^2] [ux] [ux^2] [uy] [uy^2]
 *   [uz] [uz^2] [x*ux] [y*uy] [z*uz] [x*uy] [y*ux] [ux/uz] [uy/uz] [ga] [ga^2]; idata = Np.
 * Per-slice particle moments are divided by the slice's sum(w) (not sum(w) itself, nor
 * the plasma's [(ga-1)*(1-vz)]); the sums are not. Only valid particles with
 * x^2 + y^2 <= radius^2 count (insitu_radius; infinity takes all). */
#define HPQ_INSITU_FIELD_NR 10
#define HPQ_INSITU_PLASMA_NR 14
#define HPQ_INSITU_BEAM_NR 22
typedef struct hpq_insitu {
    double *field_rdata;   /* [10*nz] */
    double *field_sum;     /* [10] */
    double *plasma_rdata;  /* [14*nz] */
    int32_t *plasma_idata; /* [nz] */
    double *plasma_sum;    /* [14] */
    int32_t *plasma_isum;  /* [1] */
    double plasma_radius;
    double *beam_rdata;    /* [22*nz] */
    int32_t *beam_idata;   /* [nz] */
    double *beam_sum;      /* [22] */
    int32_t *beam_isum;    /* [1] */
    double beam_radius;
} hpq_insitu;
/* NULL turns every group off. */
int hpq_set_insitu(hpq_ctx *ctx, const hpq_insitu *d);

/* SALAME, the reference's beam-loading algorithm (SalameModule, src/salame/Salame.cpp:14-184,
 * called from src/Hipace.cpp:673-678): on every slice that holds particles of a SALAME beam,
 * after the Bx/By solve, rescale those particles' weights by W so that the Ez field seen by
 * the beam on the next slice matches the target, iterating n_iter times (or until
 * |W - 1| < relative_tolerance); SALAME beams deposit no jx/jy meanwhile
 * (MultiBeam::DepositCurrentSlice, src/particles/beam/MultiBeam.cpp:41-58). The reference runs it
 * in the first time step only: set it before that step and clear it (NULL) after.
 * mask: device int32 per beam particle, 1 for particles of a SALAME beam (<beam>.do_salame).
 * w, valid: the beam's weight and valid arrays (device), updated in place (W = 0 invalidates).
 * target: hipace.salame_Ez_target(zeta, zeta_initial, Ez_initial); NULL is the default
 * "Ez_initial". weight_factor (host, nz doubles, may be NULL) receives the last W of every
 * slice SALAME ran on. Needs hpq_params.salame = 1 and the explicit solver. */
typedef double (*hpq_salame_target_fn)(double zeta, double zeta_initial, double ez_initial, void *user);
typedef struct hpq_salame {
    const int32_t *mask;
    double *w;
    int32_t *valid;
    int n_iter;                   /* hipace.salame_n_iter (5) */
    int do_advance;               /* hipace.salame_do_advance (1) */
    double relative_tolerance;    /* hipace.salame_relative_tolerance (1e-4) */
    hpq_salame_target_fn target;
    void *user;
    double *weight_factor;
} hpq_salame;
int hpq_set_salame(hpq_ctx *ctx, const hpq_salame *s);

/* Kernel timing (the reference's HIPACE_PROFILE regions, src/utils/HipaceProfilerWrapper.H):
 * when enabled for a kernel family, every launch of it is bracketed by HIP events on the
 * context's stream; hpq_trace_read synchronises the stream and returns the summed kernel
 * time (ms) and the number of launches since the last reset. */
enum {
    HPQ_TRACE_PUSH = 0,             /* AdvancePlasmaParticles kernel */
    HPQ_TRACE_DEPOSIT = 1,          /* DepositCurrent kernel */
    HPQ_TRACE_EXPLICIT_DEPOSIT = 2, /* ExplicitDeposition kernel */
    HPQ_TRACE_BEAM_DEPOSIT = 3,     /* DepositCurrentSlice kernel */
    HPQ_TRACE_MG_SMOOTH = 4,        /* four red-black sweeps (+ residual), tiled levels */
    HPQ_TRACE_MG_COARSE = 5,        /* coarse levels in one workgroup */
    HPQ_TRACE_DST = 6,              /* batched row DST-I */
    HPQ_TRACE_N = 7
};
int hpq_trace_enable(hpq_ctx *ctx, int kernel, int on);
int hpq_trace_read(hpq_ctx *ctx, int kernel, double *total_ms, int64_t *launches);
int hpq_trace_reset(hpq_ctx *ctx);

/* Counters of the last hpq_begin_step..hpq_run_slices sequence:
 * out[0] QSA removals, out[1] MG V-cycles, out[2] PC iterations, out[3] plasma pushes
 * (non-temporary, as src/particles/pusher/PlasmaParticleAdvance.cpp:70-73 counts). */
int hpq_stats(const hpq_ctx *ctx, double *out4);

#ifdef __cplusplus
}
#endif

#endif /* HPQ_H */
