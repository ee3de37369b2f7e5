/*
 * fresco_hip.h -- C ABI of the MI355X (gfx950) structure-factor energy/gradient
 * library libfresco_hip.so.
 *
 * Conventions
 *   - Every function returns an int status: 0 = OK, otherwise a FRESCO_E_* code;
 *     fresco_last_error() returns a thread-local message for the last failure.
 *   - Constructor arrays (radii, Sk, V, L, k vectors) are HOST pointers; the
 *     handle copies what it needs to device memory and owns it.
 *   - Evaluation arrays (x, grad) are DEVICE pointers to float64, owned by the
 *     caller.  `stream` is a hipStream_t (NULL = default stream).  The energy is
 *     returned through a HOST double*; the call synchronises the stream once to
 *     deliver it.
 *   - A handle is not re-entrant; distinct handles are independent.
 *
 * Every entry point cites the reference interface it replaces.
 */
#ifndef FRESCO_HIP_H
#define FRESCO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FRESCO_OK 0
#define FRESCO_E_INVALID 1   /* bad argument: maps to ValueError (std::invalid_argument) */
#define FRESCO_E_RUNTIME 2   /* runtime failure: maps to RuntimeError (std::runtime_error) */
#define FRESCO_E_HIP 3       /* HIP / rocFFT failure: maps to RuntimeError */

/* precision / algorithm of a structure-factor potential */
#define FRESCO_PREC_FP64_PLAIN 0   /* FP64 grids + FFTs, d separate type-2 transforms (reference structure) */
#define FRESCO_PREC_FP32_ADJOINT 1 /* FP32 grids + FFTs, one inverse transform + kernel-derivative interpolation */
#define FRESCO_PREC_FP64_ADJOINT 2 /* FP64 grids + FFTs, one inverse transform + kernel-derivative interpolation */

typedef struct fresco_pot fresco_pot; /* opaque potential handle */
typedef struct fresco_opt fresco_opt; /* opaque optimizer handle */

/* Host energy(+gradient) callback: x[ndof] (host) -> *energy and, when grad is
 * not NULL, grad[ndof] (host).  Returns 0 on success. */
typedef int (*fresco_eg_callback)(const double* x, double* grad, double* energy, void* user);

/* optimizer parameters for fresco_opt_set_param */
#define FRESCO_OPT_TOL 0            /* optimizer.hpp:173 set_tol */
#define FRESCO_OPT_MAXSTEP 1        /* optimizer.hpp:174 set_maxstep */
#define FRESCO_OPT_MAX_ITER 2       /* optimizer.hpp:175 set_max_iter */
#define FRESCO_OPT_IPRINT 3         /* optimizer.hpp:176 set_iprint */
#define FRESCO_OPT_VERBOSITY 4      /* optimizer.hpp:177 set_verbosity */
#define FRESCO_OPT_H0 5             /* lbfgs.hpp:70-76 set_H0 */
#define FRESCO_OPT_MAX_F_RISE 6     /* lbfgs.hpp:77 set_max_f_rise */
#define FRESCO_OPT_USE_RELATIVE_F 7 /* lbfgs.hpp:79-82 set_use_relative_f */
/* No counterpart in the reference: 1 (default) lets LBFGS replay each
 * iteration as one recorded HIP graph when the potential supports it; 0
 * launches the same kernels one by one.  Both give identical results. */
#define FRESCO_OPT_USE_GRAPH 8
/* No counterpart in the reference: LBFGS two-loop recursion form, set
 * before the first iteration.  0 (default): vector-free from 65536 unknowns
 * on when M <= 16, else slot by slot; 1: slot by slot (lbfgs.cpp:128-160
 * as written); 2: vector-free (Chen, Wang & Zhou 2014: the recursion on the
 * coefficients of {s, y, g} with their dot products; same step up to
 * rounding). */
#define FRESCO_OPT_LBFGS_FORM 9

/* Message of the last failed call on this thread ("" if none). */
const char* fresco_last_error(void);

/* Library version string. */
const char* fresco_version(void);

/*
 * NUwU: non-uniform points, uniform k grid.
 * Replaces fresco::NUwU::NUwU (source/fresco/nuwu.hpp:35-52) as constructed by
 * fresco/potentials/nuwu.pyx:10-23.
 *   radii[n]      point radii (weights c_i = (pi r^2)^d or ((4/3) pi r^3)^d, normalised to mean 1)
 *   d             dimension 1..3 (len(L))
 *   K             stored, unused by the arithmetic (as in the reference)
 *   Sk[m], V[m]   target S0(k) and weights on the (2K'+1)^d mode grid, flat order with x fastest
 *   L[d]          box
 *   eps           NUFFT tolerance
 *   precision     FRESCO_PREC_*
 *   sigma         fine-grid upsampling factor (0 = default 2.0)
 * m must equal (2K'+1)^d exactly (the reference silently truncates other sizes).
 */
int fresco_nuwu_create(const double* radii, int64_t n, int d, double K,
                       const double* Sk, const double* V, int64_t m,
                       const double* L, double eps, int precision, double sigma,
                       fresco_pot** out);

/*
 * UwU: grid intensities, uniform k grid (FFT only).
 * Replaces fresco::UwU::UwU (source/fresco/uwu.hpp:34-50) as constructed by
 * fresco/potentials/uwu.pyx:10-22.
 *   shape[d]      grid N, first entry slowest (row-major, the FFTW plan of uwu.hpp:48)
 *   Sk[m], V[m]   targets and weights in FFT order (the caller applies the
 *                 wrapper's ifftshift, uwu.pyx:14-15); m = prod(shape)
 *   precision     FRESCO_PREC_FP32_ADJOINT = FP32 transforms, otherwise FP64
 *   exact_grad    0: the reference's gradient Re(FFT_{-1}((2/sum x) V D rho/S0))
 *                 (uwu.hpp:152-194); 1: the exact derivative of the energy
 *                 (factor 4 and the d(sum x) term)
 * ndof = m.  x is the flattened grid.
 */
int fresco_uwu_create(const int64_t* shape, int d, const double* Sk, const double* V, int64_t m,
                      int precision, int exact_grad, fresco_pot** out);

/*
 * UwNU: grid values as NUFFT modes, non-uniform k vectors.
 * Replaces fresco::UwNU::UwNU (source/fresco/uwnu.hpp:34-83) as constructed by
 * fresco/potentials/uwnu.pyx:10-22.
 *   shape[d]      mode grid N, first entry fastest (FINUFFT n_modes, modes -N/2..)
 *   K[nkd]        k vectors interleaved (k0x, k0y, ...), scaled by 2 pi / L
 *   Sk[nk], V[nk] one per k vector; nkd >= nk * d
 *   eps, sigma    NUFFT tolerance and upsampling (0 = 2.0)
 *   precision     FRESCO_PREC_FP32_ADJOINT = FP32 grids, otherwise FP64
 * ndof = prod(shape).
 */
int fresco_uwnu_create(const int64_t* shape, int d, const double* K, int64_t nkd, const double* Sk,
                       const double* V, int64_t nk, const double* L, double eps, int precision,
                       double sigma, fresco_pot** out);

/*
 * NUwNU: non-uniform points, non-uniform k vectors.
 * Replaces fresco::NUwNU::NUwNU (source/fresco/nuwnu.hpp:36-79) as constructed
 * by fresco/potentials/nuwnu.pyx:10-23.
 *   radii[n]      point radii (weights as for NUwU)
 *   K[nkd]        k vectors interleaved, unscaled (as in the reference); Nk = nkd / d
 *   Sk[m], V[m]   m >= Nk
 *   eps, sigma    type-3 NUFFT tolerance and upsampling (0 = 2.0)
 *   precision     FRESCO_PREC_FP32_ADJOINT = FP32 grids / sums, otherwise FP64
 *   method        FRESCO_NUWNU_TYPE3 (the reference's FINUFFT type 3, nuwnu.hpp:77-78)
 *                 or FRESCO_NUWNU_DIRECT (exact sums with float64 phases)
 * ndof = n * d.
 */
#define FRESCO_NUWNU_TYPE3 0
#define FRESCO_NUWNU_DIRECT 1
int fresco_nuwnu_create(const double* radii, int64_t n, int d, const double* K, int64_t nkd,
                        const double* Sk, const double* V, int64_t m, const double* L, double eps,
                        int precision, int method, double sigma, fresco_pot** out);

/*
 * Batch of npat independent NUwU patterns (1D / 2D), evaluated together: one
 * sort, one spreading and one interpolation launch over every pattern's bins,
 * batched FFTs.  Each pattern follows NUwU::NUwU (source/fresco/nuwu.hpp:35-52):
 *   npts[p]          points of pattern p; radii[sum npts] concatenated
 *   nmodes[p]        (2K_p+1)^d modes of pattern p; Sk, V [sum nmodes] concatenated
 *   L[d]             box (shared)
 *   precision        FRESCO_PREC_FP32_ADJOINT or FRESCO_PREC_FP64_ADJOINT (the
 *                    one-transform adjoint gradient); FP64_PLAIN maps to FP64_ADJOINT
 * x / grad: the patterns' coordinate vectors concatenated (ndof = d sum npts);
 * the energy is the sum of the patterns' energies.  No reference counterpart
 * (the reference evaluates one pattern per object).
 */
int fresco_nuwu_batch_create(int npat, const int64_t* npts, const double* radii, int d,
                             const int64_t* nmodes, const double* Sk, const double* V, const double* L,
                             double eps, int precision, double sigma, fresco_pot** out);

/* Energy of every pattern of a batched potential at its last evaluation
 * (host double[npat]). */
int fresco_pot_pattern_energies(fresco_pot* pot, double* energies, int npat);

/* pair interactions and distance methods for fresco_pair_create
 * (method values = fresco.distances.Distance, distance_enum.py) */
#define FRESCO_PAIR_INVERSE_POWER 0
#define FRESCO_PAIR_WCA 1
#define FRESCO_DIST_CARTESIAN 1
#define FRESCO_DIST_PERIODIC 2

/*
 * Pair potential of n particles in d = 1..4 dimensions.
 * Replaces InversePowerCartesian / InversePowerPeriodic /
 * InversePowerPeriodicCellLists (source/fresco/inversepower_potential.hpp:81-121)
 * and WCA / WCAPeriodic / WCAPeriodicCellLists (source/fresco/wca.hpp:85-135)
 * as constructed by fresco/potentials/inversepower_potential.pyx:15-83 and
 * wca_potential.pyx:14-60.
 *   kind            FRESCO_PAIR_INVERSE_POWER: E = eps/a (1 - r/(r_i+r_j))^a for r < r_i+r_j,
 *                   p1 = a;  FRESCO_PAIR_WCA: p1 = sigma (radii unused by the energy)
 *   box[d]          periodic box (may be NULL for FRESCO_DIST_CARTESIAN)
 *   use_cell_lists  periodic only; cell sides are at least the interaction range
 *                   (also for WCA, whose cutoff 2^(1/6) sigma the reference's
 *                   rcut = 2 max(r) can miss); ncellx_scale is accepted and unused
 * ndof = n * d.  The gradient is dE/dx (simple_pairwise_potential.hpp:99-129).
 */
int fresco_pair_create(int kind, double p1, double eps, const double* radii, int64_t n, int d,
                       const double* box, int method, int use_cell_lists, double ncellx_scale,
                       fresco_pot** out);

/*
 * Energy of coordinates x[n*d] (device, interleaved x0,y0,z0,x1,...).
 * Replaces fresco::BasePotential::get_energy (base_potential.hpp:28-31), i.e.
 * NUwU::get_energy (nuwu.hpp:152-208) and its siblings.
 */
int fresco_pot_energy(fresco_pot* pot, const double* x, double* energy, void* stream);

/*
 * Energy and gradient; grad[n*d] (device) is overwritten.
 * Replaces fresco::BasePotential::get_energy_gradient (base_potential.hpp:38-43),
 * i.e. NUwU::get_energy_gradient (nuwu.hpp:210-317) and its siblings; called by the
 * optimizers at optimizer.hpp:208-215.
 */
int fresco_pot_energy_gradient(fresco_pot* pot, const double* x, double* grad,
                               double* energy, void* stream);

/*
 * fresco_pot_energy_gradient for stream-ordered callers: returns once the
 * energy is on the host; the rest of the gradient pass may still be running
 * on `stream`, so grad is read after work ordered on that stream (torch
 * tensors on the caller's current stream; the next evaluation's kernels
 * queue behind it instead of waiting for the host).  Potentials without
 * that form synchronise as fresco_pot_energy_gradient.  No reference
 * counterpart (the reference's get_energy_gradient is host-synchronous,
 * base_potential.hpp:38-43).
 */
int fresco_pot_energy_gradient_stream(fresco_pot* pot, const double* x, double* grad,
                                      double* energy, void* stream);

/*
 * The energy of fresco_pot_energy_gradient at x (device, float64), keeping the
 * state its gradient needs (for NUwU: the sorted points and the pullback
 * spectrum) so that fresco_pot_gradient_kept can finish the pass later: the
 * forward and backward halves of a differentiable loss.  *token (nonzero)
 * names this evaluation; *token = 0 (energy untouched) when the potential has
 * no such split.  No reference counterpart (the reference evaluates energy and
 * gradient together, nuwu.hpp:210-317).
 */
int fresco_pot_energy_keep(fresco_pot* pot, const double* x, double* energy, uint64_t* token, void* stream);

/*
 * The gradient (device, float64, ndof) of the kept evaluation named token;
 * *ok = 0 (grad untouched) when the potential no longer keeps it (another
 * evaluation or a target update since).  grad may be NULL: only the check.
 * The coordinates passed to fresco_pot_energy_keep must still hold their values.
 */
int fresco_pot_gradient_kept(fresco_pot* pot, uint64_t token, double* grad, int* ok, void* stream);

/*
 * Copy the last evaluation's rho_k (complex, interleaved re/im, reference flat mode
 * order, length = number of modes) into rho[2*m] (device, float64).  The S(k) the
 * reference would compute from it is |rho_k|^2 / N (nuwu.hpp:262-264).
 * No reference counterpart: exposes an intermediate of nuwu.hpp:258 for the
 * final-S(k) report and the autograd backward.
 * When the fine grid only covers the active band (NUwU adjoint precisions with
 * V = 0 outside |k_a| <= band), this re-reads the last evaluation's x, which
 * must still hold those coordinates; fresco_pot_backward does the same.
 */
int fresco_pot_get_rho(fresco_pot* pot, double* rho, void* stream);

/*
 * Derivatives of the last evaluation's energy with respect to the targets:
 * dS0[m], dV[m] (device, float64, the potential's mode order; m = nmodes):
 *   dphi/dV_k  = D_k^2
 *   dphi/dS0_k = -2 V_k D_k s_k / S0_k^2   (S0_k != 0),   -2 V_k D_k   (S0_k == 0)
 * with s_k = |rho_k|^2 / norm and D_k as in the loss (nuwu.hpp:260-276).
 * The backward of the autograd function; no reference counterpart (the
 * reference's S0 / V are constructor constants, nuwu.hpp:25-26).
 */
int fresco_pot_backward(fresco_pot* pot, double* dS0, double* dV, void* stream);

/*
 * Replace the targets S0 and weights V (device, float64, nmodes each, the
 * order the constructor takes them in after the wrapper's preprocessing).
 * No reference counterpart (constructor constants there).  For NUwU a mode
 * switched on outside the constructor's active set (V_k = V_-k = 0 before)
 * rebuilds the mode-pair tables on the host; otherwise the update stays on
 * the device.  Synchronises the stream once.  V may be NULL: the weights stay
 * and only S0 changes (no synchronisation).
 */
int fresco_pot_set_targets(fresco_pot* pot, const double* Sk, const double* V, void* stream);

/*
 * Spreader configuration of a NUFFT-based potential after its autotuning
 * (done on the first evaluation): mode 0 = register / owner tiles, 1 = halo
 * tiles with `parts` point-range slices per bin; -1 when not applicable.
 * No reference counterpart (FINUFFT chooses its spreading internally).
 */
int fresco_pot_spread_config(fresco_pot* pot, int* mode, int* parts);

/*
 * Where that spreader configuration came from: 0 fixed (one candidate), 1 the
 * autotuning cache or the tuning table shipped with the library (same plan
 * shape), 2 timed on this plan's first evaluation (then stored in the
 * cache), 3 pinned by FRESCO_SPREAD3.  No reference counterpart.
 */
int fresco_pot_spread_source(fresco_pot* pot, int* source);

/* Number of degrees of freedom the potential expects (n*d for points, prod(N) for grids). */
int fresco_pot_ndof(fresco_pot* pot, int64_t* ndof);

/* Number of modes / k vectors of the potential. */
int fresco_pot_nmodes(fresco_pot* pot, int64_t* nmodes);

/*
 * Introspection of the NUFFT set-up: fine grid sizes nf[3], kernel width w,
 * kernel beta and the tile sizes used by the spreader.
 */
int fresco_pot_plan_info(fresco_pot* pot, int64_t* nf, int* w, double* beta, int* tile);

/* Release a handle and all device memory / FFT plans it owns (the reference never frees its plans). */
int fresco_pot_destroy(fresco_pot* pot);

/*
 * Structure factor and pair correlation of a pattern: replaces the finufft
 * nufft{2,3}d1 call and the numpy inverse FFT of export_structure
 * (fresco/utils/utils.py:38-46).
 *   x[n*d]: device float64 positions; c[n]: device float64 weights (the
 *   caller's ones or normalised volumes, utils.py:26-35); L[d]: host.
 *   Sk[nk^d] (device): |sum_j c_j e^{i k.x_j}|^2 / n, x = 2 pi p / L,
 *   k_a = i_a - nk/2, first index slowest (x), as the Python finufft returns.
 *   gr[nk^d] (device, may be NULL): fftshift(real(ifftn(ifftshift(Sk) - 1)))
 *   * nk^d * prod(L) / n.
 *   eps: NUFFT tolerance (utils.py:16 tol); precision 0: float64, 1: float32.
 *   Returns after the results are complete.
 */
int fresco_structure_factor(const double* x, const double* c, int64_t n, int d, const double* L, int64_t nk,
                            double eps, int precision, double* Sk, double* gr, void* stream);

/*
 * Shell sums of radial_profile (fresco/utils/utils.py:138-152).
 *   data: device float64 grid, C order, shape[d], d <= 3; center[d]: host.
 *   sums[b], counts[b] (device, b = 0 .. shape[0]/2 + 1): sum and number of
 *   cells with floor(|i - center|) = b; cells beyond R = shape[0]/2 fold into
 *   b = R + 1.  The profile is sums / counts over the shells present.
 *   Returns after the results are complete.
 */
int fresco_radial_profile(const double* data, int d, const int64_t* shape, const double* center, double* sums,
                          int64_t* counts, void* stream);

/*
 * fresco_radial_profile without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Same arguments, kernels and summation order as fresco_radial_profile, so
 * the results are equal; its working buffers live in a fixed scratch array
 * of the library (20 MiB per GPU, held until the process ends).
 * Limits: at most FRESCO_RP_NOALLOC_MAX_CELLS cells and shape[0] <=
 * FRESCO_RP_NOALLOC_MAX_SHAPE0; larger grids return FRESCO_E_INVALID.
 * Every call on one GPU shares that scratch: calls that overlap on one GPU
 * (other host threads, other streams) race.  Returns after the results are
 * complete.
 */
#define FRESCO_RP_NOALLOC_MAX_CELLS 1048576
#define FRESCO_RP_NOALLOC_MAX_SHAPE0 4093
int fresco_radial_profile_noalloc(const double* data, int d, const int64_t* shape, const double* center,
                                  double* sums, int64_t* counts, void* stream);

/*
 * Sum of potentials over the same coordinates.
 * Replaces fresco::CombinedPotential (source/fresco/combine_potentials.hpp:16-59),
 * fresco/potentials/combine_potentials.pyx:10-30.  The children stay alive while
 * the combination exists (shared ownership); destroying a child handle is allowed.
 */
int fresco_combined_create(fresco_pot* const* pots, int count, fresco_pot** out);

/*
 * Potential implemented by a host callback (ndof coordinates).
 * Replaces the PythonPotential route (source/fresco/python_potential_wrapper.hpp:43-184,
 * fresco/potentials/_pythonpotential.pyx:58-65).  x is copied device->host before
 * and grad host->device after each call.
 */
int fresco_callback_create(int64_t ndof, fresco_eg_callback cb, void* user, fresco_pot** out);

/*
 * LBFGS with backtracking line search on the device.
 * Replaces fresco::LBFGS::LBFGS (source/lbfgs.cpp:9-31) as built by
 * fresco/optimize/_lbfgs_cpp.pyx:43-70.  x0[n] is a host array; other settings via
 * fresco_opt_set_param.
 */
int fresco_lbfgs_create(fresco_pot* pot, const double* x0, int64_t n, double tol, int M, fresco_opt** out);

/*
 * Modified FIRE on the device.
 * Replaces fresco::MODIFIED_FIRE::MODIFIED_FIRE (source/modified_fire.cpp:43-68) as
 * built by fresco/optimize/_modified_fire_cpp.pyx:18-41.
 */
int fresco_fire_create(fresco_pot* pot, const double* x0, int64_t n, double dtstart, double dtmax,
                       double maxstep, int64_t Nmin, double finc, double fdec, double fa, double astart,
                       double tol, int stepback, fresco_opt** out);

/* Setters of optimizer.hpp:173-177 and lbfgs.hpp:70-82 (FRESCO_OPT_*). */
int fresco_opt_set_param(fresco_opt* opt, int param, double value);

/* GradientOptimizer::run(niter) (optimizer.hpp:118-133); niter < 0 runs run()
 * (optimizer.hpp:139-142, up to max_iter). */
int fresco_opt_run(fresco_opt* opt, int niter);

/* one_iteration (lbfgs.cpp:36-64, modified_fire.hpp:160-230). */
int fresco_opt_one_iteration(fresco_opt* opt);

/* reset(x0) (lbfgs.cpp:247-257, modified_fire.hpp:84-112); x0[n] host. */
int fresco_opt_reset(fresco_opt* opt, const double* x0, int64_t n);

/* set_func_gradient (optimizer.hpp:148-165, modified_fire.cpp:94-118); grad[n] host. */
int fresco_opt_set_func_gradient(fresco_opt* opt, double f, const double* grad, int64_t n);

/* stop_criterion_satisfied (optimizer.hpp:195-201). */
int fresco_opt_stop_criterion(fresco_opt* opt, int* satisfied);

/* No counterpart in the reference: how many times LBFGS recorded its
 * iteration as a HIP graph and how many iterations replayed one (0, 0 for
 * FIRE and for potentials that cannot be recorded). */
int fresco_opt_graph_stats(fresco_opt* opt, int64_t* captures, int64_t* replays);

/* get_f / get_rms / get_niter / get_nfev / get_maxiter / get_H0 (optimizer.hpp:181-189,
 * lbfgs.hpp:86); any pointer may be NULL. */
int fresco_opt_get_state(fresco_opt* opt, double* f, double* rms, int* niter, int* nfev, int* maxiter,
                         double* H0);

/* get_x / get_g (optimizer.hpp:181-182) into host arrays of size n. */
int fresco_opt_get_x(fresco_opt* opt, double* x);
int fresco_opt_get_g(fresco_opt* opt, double* g);

int fresco_opt_destroy(fresco_opt* opt);

/*
 * Stage profiler (HIP events recorded on the evaluation's stream around each
 * stage: 0 bin_sort, 1 spread, 2 fft_fwd, 3 zero_fill, 4 loss, 5 fft_inv,
 * 6 interp, 7 spread_merge = the partial-tile sum kernel of two-kernel spreads).  fresco_prof_get returns the stage name, the summed duration in
 * milliseconds and the number of timed launches since the last reset.
 * No reference counterpart (the reference has no timers; its only progress
 * output is the optimizer's iprint line, lbfgs.cpp:56-62).
 */
int fresco_prof_enable(int on);
/* Time only the stages whose bit (1 << stage) is set (0 = off): fewer events
 * on the stream when one kernel's duration is wanted. */
int fresco_prof_enable_stages(unsigned mask);
/* Time one launch of every `period` of each enabled stage (1 = every launch,
 * the default): the stream carries fewer events when a kernel's average
 * duration is sampled during a throughput run. */
int fresco_prof_sample_period(int period);
int fresco_prof_reset(void);
int fresco_prof_get(int stage, const char** name, double* total_ms, int64_t* count);

/*
 * Failure injection for tests: the next `count` device allocations fail with
 * FRESCO_E_HIP.  No reference counterpart (the reference has no fault hooks).
 */
int fresco_set_fault(int count);

#ifdef __cplusplus
}
#endif

#endif /* FRESCO_HIP_H */
