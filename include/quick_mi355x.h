/*
 * quick_mi355x.h -- C-ABI of the MI355X (gfx950) Fock-build library.
 *
 * Two layers:
 *
 *  1. The reference's Fortran -> C boundary, symbol for symbol (QUICK
 *     src/cuda/gpu.h:17-50 and the definitions in src/cuda/gpu.cu cited below).
 *     Every argument is passed by reference, as gfortran/flang pass implicit-
 *     interface arguments; arrays are Fortran column-major with 1-based index
 *     contents.  These calls are synchronous, keep one process-global context,
 *     and never return an error code (as the reference); a failure is recorded
 *     and readable through qk_last_error(), and fatal ones abort with a message.
 *
 *  2. An explicit handle-based API (qk_*) for callers that keep the density and
 *     the operator resident in device memory (the test harness, bench.py and
 *     any host that binds through ctypes/iso_c_binding).  These return 0 on
 *     success and a negative code on failure (message in qk_last_error()).
 */
#ifndef QUICK_MI355X_H
#define QUICK_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- reference boundary (Fortran, by reference) ---------------- */

/* gpu.cu:29   create the process-global context */
void gpu_startup_(void);
/* gpu.cu:18   choose the device (-1: first visible device) */
void gpu_set_device_(int* gpu_dev_id);
/* gpu.cu:43   initialise the chosen device; aborts if none is present */
void gpu_init_(void);
/* gpu.cu:175  device information for the output header (name is written
 * without a terminator, namelen receives its length) */
void gpu_get_device_info_(int* gpu_dev_count, int* gpu_dev_id, int* gpu_dev_mem,
                          int* gpu_num_proc, double* gpu_core_freq, char* gpu_dev_name,
                          int* name_len, int* majorv, int* minorv);
/* gpu.cu:221  system sizes */
void gpu_setup_(int* natom, int* nbasis, int* nElec, int* imult, int* molchg, int* iAtomType);
/* gpu.cu:283  coordinates, 3 x natom, bohr */
void gpu_upload_xyz_(double* atom_xyz);
/* gpu.cu:334  atom types (atomic numbers) and charges, kept for the XC step */
void gpu_upload_atom_and_chg_(int* atom_type, double* atom_chg);
/* gpu.cu:1070 basis: QUICK's shell tables (quick_basis_module) */
void gpu_upload_basis_(int* nshell, int* nprim, int* jshell, int* jbasis, int* maxcontract,
                       int* ncontract, int* itype, double* aexp, double* dcoeff,
                       int* first_basis_function, int* last_basis_function,
                       int* first_shell_basis_function, int* last_shell_basis_function,
                       int* ncenter, int* kstart, int* katom, int* ktype, int* kprim, int* kshell,
                       int* Ksumtype, int* Qnumber, int* Qstart, int* Qfinal, int* Qsbasis,
                       int* Qfbasis, double* gccoeff, double* cons, double* gcexpo, int* KLMN);
/* gpu.cu:401  Schwarz bounds Ycutoff(nshell,nshell) and cutPrim(jbasis,jbasis);
 * builds the class-sorted, bound-sorted pair lists */
void gpu_upload_cutoff_matrix_(double* YCutoff, double* cutPrim);
/* gpu.cu:269  0: HF (K weight 1), 1: B3LYP (0.2), 2: pure DFT (0) */
void gpu_upload_method_(int* quick_method);
/* gpu.cu:1008 o: operator to accumulate into (H on entry), dense: density;
 * co and vec are accepted and unused, as in the reference */
void gpu_upload_calculated_(double* o, double* co, double* vec, double* dense);
/* gpu.cu:356  cutMatrix(nshell,nshell) = max |D| over each shell block, used by
 * the density screen of the next gpu_get2e_; the fourth argument is never read
 * (three-argument callers exist: hfoperator.f90:81, gradient.f90:283) */
void gpu_upload_cutoff_(double* cutMatrix, double* integralCutoff, double* primLimit,
                        double* DMCutoff);
/* gpu.cu:1806 forward: o <- o + G(D), symmetric, written in place */
void gpu_get2e_(double* o);
/* gpu.cu:1393 gradient seed (3*natom, the caller's partial gradient, copied
 * here) and gradient cutoff */
void gpu_upload_grad_(double* grad, double* gradCutoff);
/* gpu.cu:1481 backward: grad <- seed + dE_2e/dR (four-centre term) for the
 * uploaded density.  Screening as the reference CPU path (gradient.f90:289-316,
 * shell.f90:1941): quartet and primitive tests both against gradCutoff. */
void gpu_grad_(double* grad);
/* gpu.cu:1866 B3LYP XC for the uploaded density on the SG-1 grid (isg = 1; the
 * grid is built on the device once per geometry): o += V_xc, Eelxc += E_xc,
 * aelec and belec += N/2.  sigrad2: squared radii of significance of the basis
 * functions (get_sigrad).  The basis arrays it needs (ncontract, itype, aexp,
 * dcoeff, ncenter) and the atom types come from gpu_upload_basis_ and
 * gpu_upload_atom_and_chg_. */
void gpu_getxc_(int* isg, double* sigrad2, double* Eelxc, double* aelec, double* belec,
                double* o);
/* gpu.cu:1991 / gpu.cu:1559 NODIRECT (stored-integral) mode: not part of this
 * library; records an error */
void gpu_aoint_(double* leastIntegralCutoff, double* maxIntegralCutoff, int* intNum,
                char* intFileName);
void gpu_addint_(double* o, int* intindex, char* intFileName);
/* gpu.cu:1445 free per-geometry state; gpu.cu:206 free everything */
void gpu_cleanup_(void);
void gpu_shutdown_(void);

/* ---------------- explicit API ---------------- */

typedef struct qk_ctx qk_ctx;

enum qk_precision { QK_FP64 = 0, QK_MIXED = 1 };
enum qk_boys_order { QK_BOYS_ORDER_SHELL = 0, QK_BOYS_ORDER_QSHELL = 1 };

typedef struct qk_opts {
    double tau;            /* quartet cutoff (QUICK integralCutoff), e.g. 1e-11 */
    double prim_limit;     /* primitive cutoff (QUICK primLimit), e.g. 1e-20 */
    double pair_cut;       /* drop shell pairs with Ycutoff <= pair_cut (0: keep all,
                              as the reference CPU path; 1e-10: the reference GPU) */
    double hybrid;         /* exchange weight: 1 HF, 0.2 B3LYP, 0 pure DFT */
    int boys_order;        /* which order QUICK's FmT is asked for (branch selection
                              near T = 0): QK_BOYS_ORDER_SHELL = CPU path
                              (shell.f90:509), QK_BOYS_ORDER_QSHELL = GPU path */
    int precision;         /* QK_FP64 or QK_MIXED */
    double fp32_threshold; /* QK_MIXED: quartets with Y_ij*Y_kl below it run in FP32 */
} qk_opts;

/* fills the defaults: tau 1e-11, prim_limit 1e-20, pair_cut 0, hybrid 1,
 * shell Boys order, FP64, fp32_threshold 1e-5 */
void qk_default_opts(qk_opts* o);

const char* qk_last_error(void);
qk_ctx* qk_create(int device);
void qk_destroy(qk_ctx* c);
/* the process-global context the gpu_*_ entry points use (NULL before gpu_startup_) */
qk_ctx* qk_global_context(void);

/* Same arrays as gpu_upload_xyz_/gpu_upload_basis_ (column-major, 1-based contents). */
int qk_set_system(qk_ctx* c, int natom, const double* xyz, int nshell, int nbasis, int jbasis,
                  const int* kstart, const int* kprim, const int* katom, const int* ktype,
                  const int* Qstart, const int* Qfinal, const int* Qsbasis, const int* Qfbasis,
                  const int* ksumtype, const int* KLMN, const double* cons,
                  const double* gccoeff, const double* gcexpo);
/* Schwarz bounds from the caller (as gpu_upload_cutoff_matrix_) */
int qk_set_schwarz(qk_ctx* c, const double* Ycutoff, const double* cutprim);
/* Schwarz bounds computed on the device with QUICK's definitions
 * (schwarz.f90:12-548); outputs (host, may be NULL) receive them too */
int qk_compute_schwarz(qk_ctx* c, double* Ycutoff_out, double* cutprim_out);
/* Build the class/bound-sorted work lists for the given options (tau, pair_cut,
 * precision split).  Called implicitly by the builds when the options change. */
int qk_prepare(qk_ctx* c, const qk_opts* o);

/* Forward: dG <- G(dD), the two-electron operator (full symmetric n x n) for
 * the density dD; both device pointers, n x n doubles.  Runs on `stream`
 * (a hipStream_t, NULL = default) and returns after enqueueing; part of the
 * work runs on a second stream of the context, forked from `stream` and joined
 * back to it with events. */
int qk_fock_device(qk_ctx* c, const double* dD, double* dG, const qk_opts* o, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory during the call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): qk_fock_device with its per-build scratch
 * (one int64 accumulator, the pair-indexed Coulomb buffer, counters, scaled
 * density, shell-block density maxima) in a fixed device array of 64 MiB that
 * each GPU holds until the process ends,
 * and no hipMalloc, hipFree or allocating library call.  dG is bitwise the
 * operator qk_fock_device writes.  Requires the work lists for the same options
 * on the device (qk_prepare, or an earlier build): -6 otherwise; -7 when the
 * system needs more scratch than the array holds: about 20 n^2 + 8 nshell^2
 * bytes in FP64 and 24 n^2 + 8 nshell^2 in mixed precision when every shell pair
 * is significant (6-31G** water: 1770 and 1620 functions), less when pairs are
 * screened out.  Calls that overlap on one GPU (other streams, other
 * contexts) share the array and race.  The per-class statistics and the range
 * flag of this build are kept in the array, not reported. */
int qk_fock_device_noalloc(qk_ctx* c, const double* dD, double* dG, const qk_opts* o, void* stream);
/* Host convenience: F <- F + G(D) (host arrays), synchronous */
int qk_fock_host(qk_ctx* c, const double* D, double* F, const qk_opts* o);

/* Backward: dgrad (3*natom, device) <- dgrad + dE2/dR at fixed density dD,
 * E2 = 1/2 sum D G(D): the four-centre term of hfgrad (gradient.f90:261-320,
 * shell.f90:1863-2535, hrr.f90:862-1014).  opts->tau is the gradient cutoff
 * (quick_method%gradCutoff).  With opts->precision QK_MIXED, quartets whose
 * Schwarz bound Y_ij*Y_kl lies below opts->fp32_threshold are differentiated
 * in FP32 (accumulated in FP64 fixed point).  Runs on `stream`, with a second
 * stream of the context forked from it and joined back, as qk_fock_device. */
int qk_grad_device(qk_ctx* c, const double* dD, double* dgrad, const qk_opts* o, void* stream);
int qk_grad_host(qk_ctx* c, const double* D, double* grad, const qk_opts* o);

/* One-electron integrals and their gradient terms (get1e, quick_one_electron_
 * integral.f90:1633-1925; the one-electron part of hfgrad, gradient.f90:40-250).
 * qk_set_charges: the nuclear charges (natom, host), kept with the system.
 * qk_oneint_device: dS <- S, dH <- H_core = T + V (n x n device arrays; either
 *   may be NULL); asynchronous on `stream`.
 * qk_nuclear_repulsion: E <- sum_{A<B} Z_A Z_B / R_AB (host).
 * qk_oneint_grad_device: dgrad (3 natom, device) += nuclear repulsion gradient
 *   + sum D dH/dR - sum W dS/dR for the density dD and the energy-weighted
 *   density dW = 2 sum_occ e_i c_i c_i^T (n x n device arrays); asynchronous.
 * qk_oneint_host / qk_oneint_grad_host: the same on host arrays, synchronous. */
int qk_set_charges(qk_ctx* c, const double* charges);
int qk_oneint_device(qk_ctx* c, double* dS, double* dH, void* stream);
int qk_nuclear_repulsion(qk_ctx* c, double* E);
int qk_oneint_grad_device(qk_ctx* c, const double* dD, const double* dW, double* dgrad, void* stream);
int qk_oneint_host(qk_ctx* c, double* S, double* H);
int qk_oneint_grad_host(qk_ctx* c, const double* D, const double* W, double* grad);

/* The RHF SCF step beside the Fock build (electdiis, scf.f90:56-645): the core-
 * Hamiltonian guess, then per cycle F = H + G(D) from qk_fock_device with the
 * options in `fock`, E = 1/2 sum D (H + F) + E_nn, Pulay DIIS on F D S - S D F
 * (up to max_diis stored cycles), X F X diagonalised in FP64 (X = S^{-1/2}),
 * D = 2 C_occ C_occ^T, until the RMS change of D falls below dens_rms (QUICK's
 * DENSERMS).  D, F and the DIIS history stay on the device; needs the system,
 * its Schwarz bounds and the nuclear charges (qk_set_charges).  Outputs (host,
 * each may be NULL): the final density (n x n), its energy, the cycles run and
 * the last RMS change.  Returns -4 when the cycle limit is reached unconverged. */
typedef struct qk_scf_opts {
    int max_cycles;        /* default 100 */
    int max_diis;          /* stored cycles for DIIS, 1..10 (default 10) */
    double dens_rms;       /* convergence: RMS change of D (default 1e-7) */
    qk_opts fock;          /* options of the Fock builds */
} qk_scf_opts;
void qk_scf_default_opts(qk_scf_opts* o);
int qk_scf(qk_ctx* c, int nocc, const qk_scf_opts* o, double* D_out, double* E_out, int* cycles_out,
           double* rms_out);

/* The forward/backward pair under the names the test harness binds (the
 * torch.autograd.Function there wraps them); host arrays, synchronous.
 *   qk_fock_build: F <- F + G(D), as qk_fock_host (gpu_get2e_, gpu.cu:1806).
 *   qk_fock_grad:  grad <- grad + dE2/dR, E2 = 1/2 sum D G(D), as qk_grad_host
 *                  (gpu_grad_, gpu.cu:1481), when W is NULL; with W, the
 *                  energy-weighted density, the complete dE/dR of hfgrad
 *                  (gradient.f90:40-320): nuclear repulsion, the one-electron
 *                  terms and the four-centre term (needs qk_set_charges). */
int qk_fock_build(qk_ctx* c, const double* D, double* F, const qk_opts* o);
int qk_fock_grad(qk_ctx* c, const double* D, const double* W, double* grad, const qk_opts* o);

/* Range check of the fixed-point accumulators: waits for the device, returns 0
 * when the last Fock build's and the last gradient's sums stayed well inside
 * the int64 range, -5 (message in qk_last_error) when one came within a factor
 * of two of wrapping.  The host entry points (qk_fock_host, qk_grad_host,
 * gpu_get2e_, gpu_grad_) run this check themselves. */
int qk_check_range(qk_ctx* c);

/* Statistics of the last build: surviving contracted Qshell quartets, primitive
 * quartets evaluated (FP64 and FP32), algorithmic flops. */
typedef struct qk_stats {
    long long quartets;
    long long quartets_fp32;
    long long prim_quartets;
    double flops_fp64;
    double flops_fp32;
} qk_stats;
int qk_last_stats(qk_ctx* c, qk_stats* s);

/* Per-class launch bookkeeping for profiling: number of classes, and the
 * (la,lb,lc,ld, quartet-space size) of class k. */
int qk_num_classes(void);
int qk_class_info(qk_ctx* c, int k, int* lalblcld, long long* work);
/* Fraction of wavefront lanes the work tiles of class k occupy (0..1). */
double qk_class_lane_fill(qk_ctx* c, int k);
/* Record HIP events around every class kernel of the following builds (on the
 * stream the kernels run on). */
int qk_set_timing(qk_ctx* c, int on);
/* Figures of class k's FP64 kernel in the last build: surviving quartets,
 * primitive quartets, algorithmic flops; and the kernel's average time in ms
 * over all builds since qk_set_timing(c, 1) (0 unless timing was on). */
int qk_class_stats(qk_ctx* c, int k, long long* quartets, long long* prim_quartets, double* flops, float* ms);
/* Surviving quartets and primitive quartets of class k in the last gradient,
 * and its kernel's average time in ms (0 unless timing is on). */
int qk_grad_class_stats(qk_ctx* c, int k, long long* quartets, long long* prim_quartets, float* ms);
/* The same for the FP32 part of the last mixed-precision gradient. */
int qk_grad_class_stats_fp32(qk_ctx* c, int k, long long* quartets, long long* prim_quartets, float* ms);
/* The same figures for class k's FP32 kernel (QK_MIXED; zero otherwise). */
int qk_class_stats_fp32(qk_ctx* c, int k, long long* quartets, long long* prim_quartets, double* flops, float* ms);

/* Largest grid (blocks of waves_per_block wavefronts) the library launches for
 * a kernel with lane_bytes of scratch per lane: the dispatch's scratch backing,
 * lane_bytes x 64 x waves in flight, stays under 256 MiB.  No device needed. */
long long qk_scratch_grid_cap(long long lane_bytes, int waves_per_block);

/* The FP32 kernels' Boys table: F_m(k / per_unit) for k < rows, m < cols,
 * row-major.  No device needed. */
void qk_boys_table_shape(int* rows, int* cols, int* per_unit);
int qk_boys_table_f32(float* out, int rows, int cols);

/* ---------------- exchange-correlation (B3LYP) ---------------- */

/* The XC step of QUICK's B3LYP operator (dftoperator, dft.f90:894-1018; GPU
 * getxc_kernel, gpu_getxc.cu:92-267) for a closed-shell density: grid,
 * Stratmann-Scuseria-Frisch partition, basis values, density, functional and
 * potential on the device, the two contractions with the density and the
 * potential on FP64 MFMA, V_xc accumulated in int64 fixed point (bitwise
 * reproducible). */
typedef struct qk_xc qk_xc;

enum qk_grid_kind {
    QK_GRID_SG1 = 1,      /* SG-1: 50 Euler-Maclaurin radial points, pruned Lebedev
                             6/38/86/194 by region, SG-1 radii (gridformnew) */
    QK_GRID_PRODUCT = 2   /* nrad Euler-Maclaurin radial points x an nang-point
                             Lebedev rule (6, 38, 86, 194 or 302), SG-1 radii */
};

typedef struct qk_xc_opts {
    int grid;              /* qk_grid_kind */
    int nrad;              /* QK_GRID_PRODUCT: radial points per atom */
    int nang;              /* QK_GRID_PRODUCT: Lebedev points per shell */
    double weight_cut;     /* drop points of smaller weight (QUICK DMCutoff, 1e-10) */
    double basis_cut;      /* basis-value significance (QUICK basisCufoff, 1e-10) */
    double density_cut;    /* skip points with rho_alpha below it (QUICK DMCutoff) */
} qk_xc_opts;

/* SG-1, cutoffs 1e-10 */
void qk_xc_default_opts(qk_xc_opts* o);
qk_xc* qk_xc_create(int device);
void qk_xc_destroy(qk_xc* x);
/* Atoms (xyz 3 x natom in bohr, atomic numbers) and the Cartesian basis
 * functions as gpu_upload_basis_ passes them (gpu.cu:1070): ncontract(n),
 * itype(3, n), aexp(maxcontract, n), dcoeff(maxcontract, n) normalised per
 * function (normalize_basis), ncenter(n) 1-based.  sigrad2: the squared radii
 * of significance (get_sigrad, dft.f90:536-613), or NULL to compute them here
 * with basis_cut. */
int qk_xc_set_system(qk_xc* x, int natom, const double* xyz, const int* Z, int nbasis, int maxcontract,
                     const int* ncontract, const int* itype, const double* aexp, const double* dcoeff,
                     const int* ncenter, const double* sigrad2, double basis_cut);
/* the squared radii of significance in use (n) */
int qk_xc_sigrad2(qk_xc* x, double* out);
/* Build the grid on the device (points, weights, partition, batches). */
int qk_xc_build_grid(qk_xc* x, const qk_xc_opts* o);
/* Use a caller's grid instead: npts points as (x, y, z, w) rows, host array;
 * points outside every function's sphere of significance are dropped. */
int qk_xc_set_grid(qk_xc* x, long long npts, const double* pts);
/* points of the grid in use, and the points themselves ((x, y, z, w) rows in
 * the device's batch order) with their parent atoms (-1 for a caller's grid) */
long long qk_xc_grid_size(qk_xc* x);
int qk_xc_get_grid(qk_xc* x, double* pts, int* atom);
/* batches, mean functions per batch, FP64 MFMA flops of one evaluation */
int qk_xc_batch_info(qk_xc* x, int* nbatch, double* mean_functions, double* flops);
/* dV (n x n, device) <- V_xc of the density dD (n x n, device); E_xc and the
 * electron count (host scalars, may be NULL; when given the call waits for the
 * stream).  Runs on `stream`. */
int qk_xc_device(qk_xc* x, const double* dD, double* dV, double* Exc, double* nelec, void* stream);
/* host arrays, synchronous */
int qk_xc_host(qk_xc* x, const double* D, double* V, double* Exc, double* nelec);

#ifdef __cplusplus
}
#endif
#endif
