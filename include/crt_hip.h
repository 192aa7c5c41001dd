/* C-ABI of the MI355X (gfx950) CRMod/CRTomo forward + sensitivity path.
 *
 * The entry points replace, stage for stage, the serial-region calls of the
 * reference programs (CRMod src/fem.f90, CRTomo src/inv.f90 and update.f90).
 * Arrays keep the layout of the reference's Fortran module arrays: column
 * major, with 1-based node / element / electrode numbers as VALUES inside
 * integer arrays.  Complex arrays are interleaved (re, im) doubles, i.e.
 * COMPLEX(KIND(0D0)) / C double _Complex, passed as double*.
 *
 * Every function returns the reference's error number semantics (errnr):
 * 0 = success, otherwise the code from the reference's error table (e.g. 19
 * band width too small, 26 bad element geometry, 27 zero diagonal, 28 zero
 * Cholesky pivot, 35 beta geometry, 73 duplicate electrodes, 82 zero voltage,
 * 101 mixed and Dirichlet boundaries together).  Codes >= 1000 are errors of
 * this library (1000 invalid argument, 1001 device/HIP failure, 1002 stage
 * called out of order, 1003 band width above the supported limit).
 * Calls are synchronous on return; a handle is not re-entrant, distinct handles
 * may be used from different threads.
 */
#ifndef CRT_HIP_H
#define CRT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct crt_handle crt_handle;

/* Mesh + electrodes as read by relem/relectr (src/relem.f90:62-151,
 * src/relectr.f90:46-68); the elemmod/electrmod module arrays. */
typedef struct crt_mesh {
  int32_t sanz;           /* number of nodes */
  int32_t typanz;         /* number of element types */
  int32_t mb;             /* band width */
  int32_t eanz;           /* number of electrodes */
  const int32_t *typ;     /* typanz: element type (3, 5, 8 area; 11, 12, 13 boundary) */
  const int32_t *nelanz;  /* typanz: elements of the type */
  const int32_t *selanz;  /* typanz: nodes per element of the type */
  const int32_t *nrel;    /* (elanz+relanz) x maxnodes, column major, 1-based nodes */
  const int32_t *rnr;     /* relanz: area element adjacent to each boundary element */
  const int32_t *snr;     /* sanz: coordinate pointer */
  const double *sx;       /* sanz */
  const double *sy;       /* sanz */
  const int32_t *enr;     /* eanz: electrode nodes */
  int32_t lsytop;         /* mirror sources about the mean no-flow height (bsytop) */
  int32_t lsink;          /* fictitious sink (crmod.cfg) */
  int32_t nsink;          /* sink node (1-based) */
} crt_mesh;

/* Create a handle for one mesh: copies the arrays, builds the device-side
 * row-assembly plan.  Replaces the mesh set-up of src/fem.f90:208-215.
 * Returns 19 when an element spans more than mb nodes (kompab's check). */
int crt_create(crt_handle **h, const crt_mesh *mesh);
int crt_destroy(crt_handle *h);

/* Wavenumbers and weights, src/rwaven.f90:22-171 (host, bit-identical to the
 * reference).  kwn/kwnwi must hold 30 values; the set is also kept in h. */
int crt_rwaven(crt_handle *h, int32_t nanz, const int32_t *strnr, const int32_t *vnr,
               int32_t *kwnanz, double *kwn, double *kwnwi);
/* Use a given set instead (swrtr = 0: 2-D, one wavenumber 0, src/fem.f90:240-251). */
int crt_set_wavenumbers(crt_handle *h, int32_t kwnanz, const double *kwn,
                        const double *kwnwi, int32_t swrtr);

/* Element and mixed-boundary pre-computation on the device, src/precal.f90:13-222
 * (elem1/3/5/8, beta with the reference's Bessel polynomials, bsytop). */
int crt_precal(crt_handle *h, int32_t *lbeta, int32_t *lrandb);
/* Prescribed boundary potentials (lrandb2: rrandb's rwdanz, rwdnr, rwd,
 * src/rrandb.f90), applied by randb2 (src/randb2.f90) after kompab/randb in
 * every assembled system: unit row and column at each node, b = -value there,
 * and the coupling terms moved to the right-hand side.  rwd: rwdanz complex
 * values (the DC twins take the real parts, as randbdc2 reads rwddc).
 * rwdanz = 0 switches them off.  Returns 30 for a node outside 1..sanz. */
int crt_set_boundary_values(crt_handle *h, int32_t rwdanz, const int32_t *rwdnr, const double *rwd);

/* Copy the precomputed arrays back in the reference layout (for checking):
 * elbg elanz x nimn x K, relbg relanz x nimn, kg relanz x eanz x K (column major,
 * nimn = maxnodes(maxnodes+1)/2).  Any pointer may be NULL. */
int crt_get_precal(crt_handle *h, double *elbg, double *relbg, double *kg);

/* Forward potentials for all (wavenumber, source electrode) systems: assembly
 * (kompab, randb), scaling (scalab), band Cholesky (chol) and substitution (vre),
 * src/fem.f90:328-397, then the back-transform (rtrafo, src/rtrafo.f90).
 * sigma: elanz complex conductivities.  kpot (sanz x eanz x K) and hpot
 * (sanz x eanz) are optional host outputs (NULL to keep them on the device only). */
int crt_forward(crt_handle *h, const double *sigma, double *kpot, double *hpot);
/* Same with sigma already in device memory (elanz complex). */
int crt_forward_dev(crt_handle *h, const double *sigma_dev);

/* The DC twins (CRTomo with ldc): kompadc (real part of sigma), randdc, scaldc
 * (unit diagonal; error 27 when a_ii <= 0), choldc (error 28 when a pivot is
 * <= 0), vredc or kompbdc + vredc for a shared factor, kpotdc = pot*fak, then
 * rtrafo's DC branch; replaces src/inv.f90:343-400 and src/rtrafo.f90:48-66.
 * sigma: elanz complex (as the reference's sigma array; kept for bsendc).
 * kpotdc (sanz x eanz x K) and hpotdc (sanz x eanz) are optional real host
 * outputs.  crt_voltages then gives bvolti's DC branch (src/bvolti.f90:56-66,
 * volt = dcmplx(V)). */
int crt_forward_dc(crt_handle *h, const double *sigma, double *kpotdc, double *hpotdc);

/* Voltages and apparent values, src/bvolt.f90 (mode 0: sigmaa = kfak*V) and
 * src/bvolti.f90 (mode 1: sigmaa = -ln V, error 82 on V = 0).  kfak may be NULL
 * (1).  The configuration is kept in h for the sensitivity/gradient stages. */
int crt_voltages(crt_handle *h, int32_t nanz, const int32_t *strnr, const int32_t *vnr,
                 int32_t mode, const double *kfak, double *volt, double *sigmaa);
/* Device-pointer variant of the outputs (volt_dev / sigmaa_dev: nanz complex). */
int crt_voltages_dev(crt_handle *h, int32_t nanz, const int32_t *strnr, const int32_t *vnr,
                     int32_t mode, double *volt_dev, double *sigmaa_dev);

/* Explicit Jacobian, src/bsens.f90 (mode 0: dV/dsigma) and src/bsensi.f90
 * (mode 1: d(-ln V)/d(ln sigma)); sens: nanz x elanz column major complex, or NULL
 * to keep it on the device for crt_cg_solve. */
int crt_sensitivities(crt_handle *h, int32_t mode, double *sens);

/* bsendc (src/bsendc.f90:23-187) after crt_forward_dc and crt_voltages(mode 1):
 * sensdc nanz x elanz column major, real. */
int crt_sensitivities_dc(crt_handle *h, double *sensdc);

/* Data-misfit gradient without forming J: grad_e = sum_i conj(J_ie) w_i r_i with
 * J as crt_sensitivities(mode) defines it (the data term of src/update.f90:235-289).
 * w: nanz weights (wmatd*wdfak), resid: nanz complex (d - f). */
int crt_adjoint_gradient(crt_handle *h, int32_t mode, const double *w, const double *resid,
                         double *grad);
int crt_adjoint_gradient_dev(crt_handle *h, int32_t mode, const double *w_dev,
                             const double *resid_dev, double *grad_dev);

/* The same data term from the explicit Jacobian of the last crt_sensitivities
 * call: grad_j = sum_i conj(J_ij) w_i r_i (src/update.f90:263-266, before the
 * cgfac scaling).  The explicit-J path the adjoint gradient is compared with. */
int crt_jhw_r(crt_handle *h, const double *w, const double *resid, double *grad);

/* The regularisation's smoothness operator, built on the host once per
 * inversion (bsmatm, src/bsmatm_mod.f90): ltri = 0 the rectangular stencil of
 * bsmatmreg (:190-362; elements numbered row by row, nx x nz; smatm manz x 3),
 * ltri = 1 or 2 the triangular operator of bsmatmtri (:364-461, with the
 * decoupled edges of crt_set_decouplings; smatm manz x (maxnodes+1)) with the edge neighbours of bnachbar
 * (src/bnachbar.f90) in nachbar (manz x (maxnodes+1), may be NULL).  Column
 * major, the reference's layout; the input crt_cg_solve takes. */
int crt_smoothness(crt_handle *h, int32_t ltri, double alfx, double alfz, int32_t nx, int32_t nz,
                   double *smatm, int32_t *nachbar);

/* Decoupled edges for the triangular operator (ltri 1/2) of the next
 * crt_smoothness calls: edecoup decanz x 2 column major (1-based element
 * pairs), edecstr decanz weight factors, as rall reads decouplings.dat
 * (src/rall.f90:648-665) and bsmatmtri applies them (src/bsmatm_mod.f90:434-445:
 * the weight of the edge between the pair, either order, times the factor).
 * decanz = 0 clears them.  Replaces the decoupling part of those two sites. */
int crt_set_decouplings(crt_handle *h, int32_t decanz, const int32_t *edecoup, const double *edecstr);

/* One model update's linear solve, src/update.f90:179-307 with cjggra
 * (src/cg_mod.f90:426-518): preconditioner cgfac, scaled RHS, then Fletcher-Reeves
 * CG on (J^H W J + lam R) with products bap/bp/bptri/bplma/bb, using the J of the
 * last crt_sensitivities call.  smatm: manz x (3 for ltri=0 | 1 for ltri=3,4 |
 * maxnodes+1 for triangular), nachbar: manz x maxnodes (ltri 1,2,5..14, else NULL),
 * rdiag: the diagonal of R the preconditioner adds, rm: R m (manz complex, NULL for
 * m = 0).  Outputs dpar (manz complex, scaled by cgfac as update does) and cgres
 * (ncgmax+2: cgres[0] = iterations, cgres[k] = relative residual). */
int crt_cg_solve(crt_handle *h, const double *wmatd_wdfak, const double *dat,
                 const double *sigmaa, double lam, int32_t ltri, int32_t nx,
                 const double *smatm, int32_t smatm_cols, const int32_t *nachbar,
                 int32_t maxnodes, const double *rdiag, const double *rm, int32_t ncgmax,
                 double eps, double *dpar, double *cgfac, double *cgres);

/* No reference counterpart.  Caps the device memory a forward keeps for its
 * factors in flight (band factors, scaled rows, sweeps; default 64 GiB, and
 * never more than half the free memory): a larger problem runs in several
 * factor batches.  For callers keeping many handles alive at once. */
int crt_set_device_budget(crt_handle *h, uint64_t bytes);

/* Launch choices for the handle's forwards, 0 = the built-in rule: threads
 * per factor workgroup (256 or 512; bands whose LDS window fits twice per CU
 * take 256 by default) and the back substitution's prefetch depth in rows (8
 * or 16).  The results are the same bits either way; an autotuner
 * (titatest_amd/autotune.py) times them.  No counterpart in the reference. */
int crt_set_tuning(crt_handle *h, int32_t factor_threads, int32_t backsolve_depth);

/* Kernel form of crt_sensitivities / crt_sensitivities_dc (bsens, bsensi,
 * bsendc): 0 = the built-in rule, 1 = the pair form (per element and
 * wavenumber, each distinct electrode pair's node differences, then one
 * bilinear term per measurement), 2 = the Gram form (per element, the E x E
 * matrix of every source pair's bilinear term summed over the wavenumbers on
 * the FP64 matrix cores, then one double difference per measurement; E <=
 * 128, else 1000).  The built-in rule takes the Gram form when E <= 128 and
 * its lower triangle of 16 x 16 tiles holds at most twice as many entries as
 * there are measurements.  Both forms agree to rounding.  No counterpart in
 * the reference. */
int crt_set_sens_form(crt_handle *h, int32_t form);

/* Per-wavenumber recomputation (checkpointing) of the potentials for
 * problems whose K x eanz x sanz potentials should not stay in device memory.
 * wavenumbers_per_pass 0 (the default) stores every wavenumber's potentials.
 * k > 0: a forward solves the systems in passes of k wavenumbers and keeps,
 * from each pass, only its contribution to the electrode rows of the
 * back-transform (the voltages are the stored mode's bit for bit); the
 * adjoint gradient (crt_adjoint_gradient[_dev]) re-solves each pass's
 * systems and adds its contribution, so at most k wavenumbers' potentials are
 * held at once.  In this mode crt_forward and crt_forward_dc take no kpot or
 * hpot output (1000), crt_sensitivities, crt_sensitivities_dc and
 * crt_kpot_device need the stored potentials (1002), and crt_forward_many
 * does not take the handle (1000).  The reference stores kpot(sanz, eanz,
 * kwnanz) (src/femmod.f90); no counterpart. */
int crt_set_recompute(crt_handle *h, int32_t wavenumbers_per_pass);

/* The forward of several models in one pass (the reference runs one model per
 * program run; this is the loop of src/fem.f90:328-397 over many handles).
 * hs: nmodels distinct handles on one device, after crt_precal; sigma_dev:
 * each model's conductivities on the device (elanz complex).  Each model's
 * assembly runs on its own stream; then one persistent factor launch per
 * factor-kernel form takes the (k, l) systems of all its models from a work
 * queue, costliest models first, and one back-substitution launch per form
 * takes all their right-hand sides.  Every system's factors stay resident
 * (the device budget does not apply).  Models without mixed boundaries (one
 * factor per wavenumber) run their own forward.  Afterwards each handle is as
 * after crt_forward_dev: crt_voltages_dev, the sensitivities and the adjoint
 * gradient follow per model.  Returns the first error (the handle's
 * crt_error_info names the system). */
int crt_forward_many(int32_t nmodels, crt_handle *const *hs, const double *const *sigma_dev);

/* No reference counterpart.  enable != 0: crt_cg_solve runs its iterations as
 * one HIP graph, captured on the first call and replayed while the problem's
 * sizes, buffers and parameters stay the same (for callers running many short
 * solves, where the per-kernel launches dominate).  The results are those of
 * the launch-by-launch path, bit for bit. */
int crt_set_graph(crt_handle *h, int32_t enable);

/* First failing system of the last call: error code, wavenumber and electrode
 * (1-based) where known. */
int crt_error_info(crt_handle *h, int32_t *code, int32_t *k, int32_t *l);

/* Per-kernel device time of the calls since the last reset (HIP events on the
 * handle's stream).  names: n_max slots of 48 chars; returns the count. */
int crt_set_timing(crt_handle *h, int32_t enable);
int crt_stage_times(crt_handle *h, int32_t n_max, char *names, double *ms, int32_t *launches);
/* Phase clock of the factor kernel: enable != 0 turns it on for the next
 * forwards; ticks8 (may be NULL) receives the accumulated wall-clock ticks
 * (100 MHz) since the last call, summed over workgroups: for the panel wave
 * (wave 0) the prologue, the wait for the look-ahead update, the panel and the
 * wait for the trailing update; for the trailing waves (wave 1) the prologue
 * with the look-ahead update, its barrier wait, the trailing update and the
 * wait for the panel. */
int crt_factor_phase_ticks(crt_handle *h, int32_t enable, uint64_t *ticks8);
/* Device pointer of the last forward's potentials kpot (sanz x eanz x K complex). */
int crt_kpot_device(crt_handle *h, const double **kpot_dev);

/* ---- Inversion-loop pieces around the path (one Gauss-Newton iteration can
 * stay on the device: forward -> voltages -> misfit -> update -> sigma).
 * Complex arrays are interleaved (re, im) doubles. ---- */

/* dmisft switches: the reference's lpol, lfpi, ldiff, lrobust */
#define CRT_MISFIT_LPOL 1
#define CRT_MISFIT_LFPI 2
#define CRT_MISFIT_LDIFF 4
#define CRT_MISFIT_LROBUST 8

/* Replaces dmisft (src/dmisft.f90:17-208, its chkpo2 polarity folding at
 * :213-300) as called from src/inv.f90:524 and :677, without the eps.ctr text
 * output.  dat, sigmaa (nanz complex): measured and modelled -ln V; wmatd
 * (in/out: the robust reweighting replaces it when l1rat > l1min), wmatdr,
 * wmatdp: the error weights of datmod; d0, fm0: difference-inversion data
 * (read only with CRT_MISFIT_LDIFF, may be NULL otherwise).  stats (6):
 * nrmsd, betrms, pharms, rmssum, l1rat, npol.  wdlok (nanz, may be NULL): 0
 * for a datum excluded by its polarity.  Returns 99 (the reference's errnr)
 * when every datum is excluded.  The sums run in a fixed tree order, not
 * sequentially: results agree with the reference to rounding. */
int crt_dmisft(crt_handle *h, int32_t nanz, const double *dat, const double *sigmaa, double *wmatd,
               const double *wmatdr, const double *wmatdp, const double *d0, const double *fm0,
               int32_t flags, double l1min, double *stats, int32_t *wdlok);
/* The same on device pointers, queued on the handle's stream without a host
 * synchronisation (capturable).  wdlok is required; stats[5] == nanz
 * signals errnr 99. */
int crt_dmisft_dev(crt_handle *h, int32_t nanz, const double *dat_dev, const double *sigmaa_dev,
                   double *wmatd_dev, const double *wmatdr_dev, const double *wmatdp_dev, const double *d0_dev,
                   const double *fm0_dev, int32_t flags, double l1min, double *stats_dev, int32_t *wdlok_dev);

/* No counterpart in the reference: crt_dmisft_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and stream; the
 * misfit kernel's per-datum scratch lives in a static device array of
 * 262,144 doubles (2 MiB per device, resolved by crt_create), so the call
 * allocates nothing.  nanz above 262,144 returns 1004.  Calls that overlap on
 * one device share that array and race: serialise them (one stream, or
 * events between streams).  h must come from crt_create on the current
 * device. */
int crt_dmisft_dev_noalloc(crt_handle *h, int32_t nanz, const double *dat_dev, const double *sigmaa_dev,
                           double *wmatd_dev, const double *wmatdr_dev, const double *wmatdp_dev,
                           const double *d0_dev, const double *fm0_dev, int32_t flags, double l1min,
                           double *stats_dev, int32_t *wdlok_dev);

/* Element-to-parameter map of modelmod (mnr, 1-based, elanz entries; NULL =
 * identity with manz = elanz), used by crt_bpar and crt_bsigma.  Replaces the
 * set-up of modelmod::mnr and manz (src/rall.f90:686-692, :719-728). */
int crt_set_model_map(crt_handle *h, int32_t manz, const int32_t *mnr);
/* Replaces bpar (src/bpar.f90, called at src/inv.f90:924): par(mnr(i)) =
 * log(sigma(i)) in element order (the last element of a parameter wins);
 * parameters without elements keep their value in par (in/out, manz
 * complex). */
int crt_bpar(crt_handle *h, const double *sigma, double *par);
int crt_bpar_dev(crt_handle *h, const double *sigma_dev, double *par_dev);
/* Replaces bsigma (src/bsigma.f90, called at src/inv.f90:933): sigma(i) =
 * exp(par(mnr(i))). */
int crt_bsigma(crt_handle *h, const double *par, double *sigma);
int crt_bsigma_dev(crt_handle *h, const double *par_dev, double *sigma_dev);
/* Replaces the model update of update (src/update.f90:332-351, called at
 * src/inv.f90:929): par += step dpar (manz complex, rounded as the reference's
 * complex product then sum), with lphi0 a negative phase set to zero; out
 * (2): bdpar = step sum |dpar|^2 and the count of zeroed phases. */
int crt_update_step(crt_handle *h, int32_t manz, double *par, const double *dpar, double step, int32_t lphi0,
                    double *out);
int crt_update_step_dev(crt_handle *h, int32_t manz, double *par_dev, const double *dpar_dev, double step,
                        int32_t lphi0, double *out_dev);

/* Replaces blam0 (src/blam0.f90:69-170, called at src/inv.f90:813) after
 * crt_sensitivities (or crt_sensitivities_dc): the starting regularisation
 * weight lammax = mean_j |(J^H W J 1)_j| * 2 / (alfx + alfz) * 5, formed as
 * J^H (w o (J 1)) in O(nanz manz) instead of the reference's O(nanz manz^2)
 * loop.  w = wmatd * wdfak; mode 0 complex J, 1 real parts (the lfpi branch,
 * and the ldc branch after crt_sensitivities_dc).  The preset-lambda branches
 * (:47-62) stay with the caller. */
int crt_blam0(crt_handle *h, const double *w, int32_t mode, double alfx, double alfz, double *lammax);

#ifdef __cplusplus
}
#endif

#endif
