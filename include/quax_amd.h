/*
 * quax_amd.h -- C ABI of the MI355X-native two-electron integral library
 * (libquax_hip.so).  It replaces the pybind11 module ``libint_interface`` of
 * Quax for the ERI path (reference: quax/integrals/libint_interface.cc) and adds
 * the integral-direct consumers the RHF energy and gradient need.
 *
 * Conventions
 *  - Every entry point returns an int status: QX_OK (0) or one of the QX_E*
 *    codes; nothing aborts.  qx_last_error() gives the message.  The reference
 *    aborts on a bad deriv_vec length (libint_interface.cc:1098) and throws
 *    std::invalid_argument for an unknown operator (:520); both become
 *    QX_EINVAL here.
 *  - A context holds one molecule and one basis (the reference keeps them in
 *    module globals, libint_interface.cc:24-35).  Contexts are independent; one
 *    context must not be used from two threads at once.
 *  - Output and density buffers are DEVICE pointers owned by the caller, on the
 *    context's device; ``stream`` is a hipStream_t (NULL = default stream).
 *    Calls are asynchronous with respect to the host unless stated.
 *  - Tensors use the AO basis of the context: Cartesian when pure == 0 (the
 *    reference forces Cartesian shells, libint_interface.cc:49), spherical
 *    (l >= 2 shells as 2l+1 real solid harmonics, m = -l..l) when pure == 1.
 *  - Layouts follow the reference: ERI tensor row-major (p,q,r,s) in chemists'
 *    notation (libint_interface.cc:570-585); derivative tensor
 *    (n_unique, p, q, r, s) with the nuclear index outermost (:2653).
 */
#ifndef QUAX_AMD_H
#define QUAX_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qx_ctx qx_ctx;

enum {
  QX_OK = 0,
  QX_EINVAL = 1,        /* bad argument */
  QX_ENOMEM = 2,        /* device allocation failed */
  QX_EHIP = 3,          /* HIP runtime error */
  QX_EUNSUPPORTED = 4,  /* valid request outside what this build implements */
  QX_ENODEVICE = 5      /* no usable GPU */
};

enum { QX_PREC_FP64 = 0, QX_PREC_MIXED = 1 };

/* Replaces initialize(xyzfilename, basis1, basis2, basis3, basis4, ints_tol)
 * (libint_interface.cc:38-95).  Takes the geometry in bohr and the shells
 * directly instead of an xyz file and basis names; raw contraction
 * coefficients are normalised the way Libint's Shell::renorm does.
 * Computes the Schwarz table (libint_interface.cc:158-229) on the device and
 * builds the screened, class-sorted quartet lists.  ints_tol is the
 * reference's ints_tolerance (quartets with Q_ab*Q_cd < tol^2 are skipped). */
int qx_initialize(int natom, const double* xyz_bohr, const double* nuclear_charges, int nshell,
                  const int* shell_l, const int* shell_atom, const int* shell_nprim, const double* exps,
                  const double* coefs, int pure, double ints_tol, int device, qx_ctx** out);

/* Replaces finalize() (libint_interface.cc:97-99). */
int qx_finalize(qx_ctx* ctx);

/* Batched form (no reference counterpart; Quax runs one molecule per
 * initialize): as qx_initialize, plus a fragment id per atom.  Shell pairs
 * and quartets never span two fragments, so one context -- and one kernel
 * launch per class -- evaluates the independent molecules of a batch
 * side by side; J/K, gradients and tensors are block diagonal over
 * fragments.  One-electron integrals (qx_oei*) still couple all nuclei:
 * take them from a context per molecule.  The Schwarz cut tol^2/max Q uses
 * the largest Q over the whole batch. */
int qx_initialize_fragments(int natom, const double* xyz_bohr, const double* nuclear_charges, int nshell,
                            const int* shell_l, const int* shell_atom, const int* shell_nprim, const double* exps,
                            const double* coefs, int pure, double ints_tol, int device, const int* atom_fragment,
                            qx_ctx** out);

const char* qx_last_error(const qx_ctx* ctx);

/* Basis sizes: AO functions (spherical count when pure) and Cartesian. */
int qx_nbf(const qx_ctx* ctx, int* nbf, int* ncart);

/* schwarz_screening (libint_interface.cc:158-229), host outputs:
 * values[nshell*nshell] = max_f |(f1 f2|f1 f2)| (Cartesian, symmetric);
 * pairs[2*k] = (s1, s2), s2 <= s1, kept by Q >= tol^2/max Q, reference order.
 * pairs needs room for nshell*(nshell+1) ints. */
int qx_schwarz_screening(qx_ctx* ctx, double* values, int* pairs, int* npairs);

/* Number of canonical (8-fold unique) screened shell quartets. */
int qx_num_quartets(const qx_ctx* ctx, long long* nquartets);

/* compute_2e_int("eri", 0.) (libint_interface.cc:486-647): full nbf^4 tensor
 * written to out (device, nbf^4 doubles).  Screened quartets are exact zeros. */
int qx_compute_2e_int(qx_ctx* ctx, double* out, void* stream);

/* eri_deriv_core(deriv_order) (libint_interface.cc:2582-2779): out holds
 * how_many_derivs(natom, order) * nbf^4 doubles, slices in
 * generate_multi_index_lookup(3*natom, order) order.  deriv_order 1 and 2
 * (QX_EUNSUPPORTED above; LIBINT2_MAX_DERIV_ORDER of this build is 2).
 * Quartets on a single atom are skipped.  Order 2 runs the first-derivative
 * kernels in forward mode along each nuclear coordinate. */
int qx_eri_deriv_core(qx_ctx* ctx, int deriv_order, double* out, void* stream);

/* compute_2e_deriv("eri", 0., deriv_vec) (libint_interface.cc:1097-1335):
 * one nbf^4 slice.  len must equal 3*natom (QX_EINVAL otherwise); total order
 * 0, 1 or 2 (QX_EUNSUPPORTED above).  Order 0 is the integral tensor with
 * single-atom quartets left out, as the reference's deriv-0 engine gives. */
int qx_compute_2e_deriv(qx_ctx* ctx, const int* deriv_vec, int len, double* out, void* stream);

/* how_many_derivs (utils.h:131-140): C(3*natom + order - 1, order). */
long long qx_how_many_derivs(int natom, int deriv_order);

/* generate_multi_index_lookup (utils.h:142-172): combinations with
 * replacement of range(nparams), lexicographic, written row-major into
 * out[count*deriv_order] (host).  Returns the count, or -1 if cap is short. */
long long qx_generate_multi_index_lookup(int nparams, int deriv_order, int* out, long long cap);

/* Integral-direct J and K (hartree_fock.py:25-28, 79-80 contract the stored
 * tensor the same way): J_pq = sum_rs (pq|rs) D_rs, K_pq = sum_rs (pr|qs) D_rs.
 * D, J, K: nbf x nbf device buffers.  precision QX_PREC_MIXED evaluates
 * quartets whose Schwarz bound sqrt(Q_ab Q_cd) is below fp32_threshold
 * (finite, >= 0) with an FP32 recursion (Boys function and quartet set-up in
 * FP64), each integral contracted into FP64 accumulators.  The split lists
 * are rebuilt when the threshold changes (that call synchronises the
 * device). */
int qx_jk_build(qx_ctx* ctx, const double* D, double* J, double* K, int precision, double fp32_threshold,
                void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): qx_jk_build with precision QX_PREC_FP64 that allocates
 * nothing.  Its J'/K' work matrices live in a file-scope device array of
 * qx_jk_noalloc_capacity() doubles (24 MiB, held by every GPU the library
 * runs on until the process ends): 2 n_cart^2 doubles for Cartesian bases,
 * 5 n_cart^2 for spherical ones; larger contexts return QX_EINVAL with a
 * message.  All kernels run in turn on `stream`.  Calls that overlap on one
 * GPU share the array and race; the array is cleared on every call. */
int qx_jk_build_noalloc(qx_ctx* ctx, const double* D, double* J, double* K, void* stream);
long long qx_jk_noalloc_capacity(void);

/* Adjoint two-electron gradient (no reference counterpart: the reference
 * differentiates forward through eri_deriv_core, tei.py:25-34):
 * grad[3*natom] += sum_pqrs (2 D_pq D_rs - D_pr D_qs) d(pq|rs)/dR.
 * Integrals are recomputed, the derivative tensor is never stored. */
int qx_eri_grad(qx_ctx* ctx, const double* D, double* grad, int precision, double fp32_threshold, void* stream);

/* Hessian-vector products by forward-over-reverse (no reference
 * counterpart: the reference forms Hessians with jacfwd(jacfwd) over
 * eri_deriv_core(2), core.py:89-91).  v: 3*natom device vector (a geometry
 * direction); D, Pd, Wd: nbf x nbf device buffers.  The adjoint kernels run
 * in forward mode along v, recomputing integrals; nothing n^4 is stored.
 *   qx_eri_hess_dir: out[3N] += sum_pqrs (2 D_pq D_rs - D_pr D_qs)
 *                    d/dv d(pq|rs)/dR.
 *   qx_jk_deriv_dir: J_pq = sum_rs d/dv (pq|rs) D_rs, K_pq = sum_rs
 *                    d/dv (pr|qs) D_rs (overwritten), the skeleton derivative
 *                    Fock matrix of the coupled-perturbed equations.
 *   qx_oei_hess_dir: grad[3N] += sum_pq Pd_pq d/dv d(T+V)_pq/dR +
 *                    Wd_pq d/dv dS_pq/dR. */
int qx_eri_hess_dir(qx_ctx* ctx, const double* D, const double* v, double* out, void* stream);
int qx_jk_deriv_dir(qx_ctx* ctx, const double* D, const double* v, double* J, double* K, void* stream);
int qx_oei_hess_dir(qx_ctx* ctx, const double* Pd, const double* Wd, const double* v, double* grad, void* stream);

/* One-electron integrals S, T, V (nbf x nbf device buffers); Libint
 * overlap / kinetic / nuclear operators as compute_1e_int uses them
 * (libint_interface.cc:232-299). */
int qx_oei(qx_ctx* ctx, double* S, double* T, double* V, void* stream);

/* One-electron gradient contraction (oei_deriv_core, libint_interface.cc:
 * 2403-2579, contracted on the fly): grad[3*natom] +=
 * sum_pq Pd_pq d(T+V)_pq/dR + Wd_pq dS_pq/dR.  Pd, Wd: nbf x nbf. */
int qx_oei_grad(qx_ctx* ctx, const double* Pd, const double* Wd, double* grad, void* stream);

/* One-electron derivative tensors (oei_deriv_core(1)): dS, dT, dV each
 * 3*natom * nbf * nbf doubles on the device. */
int qx_oei_deriv(qx_ctx* ctx, double* dS, double* dT, double* dV, void* stream);

/* Materialise the 8-fold unique Cartesian integrals in the class-blocked
 * layout: for every quartet list, element (component k, quartet t) at
 * base + k * nquart + t.  Requires single-primitive shells; out must hold
 * qx_store_size() doubles. */
int qx_store_size(const qx_ctx* ctx, long long* ndoubles);
int qx_eri_store(qx_ctx* ctx, double* out, void* stream);

/* Map of the stored layout (no reference counterpart; lets a caller read
 * the class-blocked integrals of qx_eri_store), host outputs.
 * qx_store_lists: number of quartet lists and, per list, its class
 * (l4[4*k..4*k+3] = la, lb, lc, ld), element offset (base) and quartet
 * count; arrays need qx_num_classes() entries.
 * qx_store_quartet_pairs: for every quartet t of list `list`, in stored
 * order, the bra and ket shell pair (indices into qx_pair_shells).
 * qx_pair_shells: the screened shell pairs, oriented (l(sa) >= l(sb)); the
 * stored components of a quartet run over (sa, sb | sa', sb') row-major. */
int qx_store_lists(const qx_ctx* ctx, int* nlists, int* l4, long long* base, long long* nquart);
int qx_store_quartet_pairs(const qx_ctx* ctx, int list, int* p1, int* p2);
int qx_pair_shells(const qx_ctx* ctx, int* npairs, int* sa, int* sb);

/* AO -> MO transformation of a two-electron tensor (replaces
 * partial_tei_transformation and tei_transformation, quax/methods/
 * energy_utils.py:46-73, called from mp2.py:29 and ccsd.py:32):
 *   out[I,J,K,L] = sum_pqrs Ci[p,I] Cj[q,J] Ck[r,K] Cl[s,L] G[p,q,r,s].
 * n[4] = dimensions of G (row-major, device), m[4] = MO counts; Ci is
 * n[0] x m[0] row-major (device), likewise Cj, Ck, Cl; out holds
 * m[0]*m[1]*m[2]*m[3] doubles (device).  Four FP64 MFMA GEMMs in the
 * reference's order (l, k, j, i), each contracting the last index and
 * putting the new one first.  work: device buffer of
 * qx_tei_transform_work(n, m) doubles.  No context; errors in
 * qx_gemm_last_error(). */
long long qx_tei_transform_work(const int* n, const int* m);
int qx_partial_tei_transformation(const double* G, const int* n, const double* Ci, const double* Cj,
                                  const double* Ck, const double* Cl, const int* m, double* out, double* work,
                                  void* stream);
const char* qx_gemm_last_error(void);

/* Mixed precision (no reference counterpart).  With QX_PREC_MIXED a pass
 * evaluates, in the classes its FP32 policy admits and in which at least a
 * quarter of the primitive quartets have a Schwarz bound below
 * fp32_threshold, those quartets with FP32 recursion and the rest in FP64
 * (two launches over one quartet list, each masked to its part); every other
 * class runs in FP64.
 * qx_mixed_quartets: contracted quartets above / below the threshold over
 * every class.  qx_mixed_pass_quartets: FP64 / FP32 contracted quartets of
 * a pass (0 = J/K, 1 = adjoint gradient; -1 = as qx_mixed_quartets).
 * qx_set_fp32_classes: per-class policy of a pass, flags[qx_num_classes()]
 * (nonzero = FP32 allowed), or the built-in default (from timings of FP32
 * against FP64 kernels per class) when flags is NULL; qx_get_fp32_classes
 * reads it. */
int qx_mixed_quartets(qx_ctx* ctx, double fp32_threshold, long long* n_fp64, long long* n_fp32);
int qx_mixed_pass_quartets(qx_ctx* ctx, double fp32_threshold, int pass, long long* n_fp64, long long* n_fp32);
int qx_set_fp32_classes(qx_ctx* ctx, int pass, const int* flags);
int qx_get_fp32_classes(const qx_ctx* ctx, int pass, int* flags);

/* RHF SCF iterations captured into a HIP graph (replaces the SCF loop of
 * quax/methods/hartree_fock.py:83-99, a jax.lax.while_loop: J/K, Fock
 * matrix, energy, DIIS error A (F D S - S D F) A, density of the
 * orthogonalised Fock matrix), for one molecule or a batch of molecules
 * that are fragments of ctx.  Matrices are nmol x nm x nm device arrays,
 * molecule b in the leading nbf[b] x nbf[b] block of its slice (zero
 * elsewhere); offset[b] is its first AO in ctx.  H = T + V, S, A =
 * inv(chol(S))^T per molecule; nocc: doubly occupied orbitals; enuc: host
 * nuclear repulsion energies.  The density comes from trace-correcting
 * purification (at most purify_iters steps of X <- X^2 or 2X - X^2, the
 * rest skipped once tr(X - X^2) < 1e-8, then two McWeeny steps
 * X <- 3X^2 - 2X^3; FP64 MFMA GEMMs) instead of an eigensolver, so an
 * iteration has no host synchronisation.
 * diis: Pulay DIIS on the error above with diis_space (2..32) vectors;
 * spectral_shift: + diag(linspace(0, 1, n)) 1e-9 on the orthogonalised Fock
 * matrix (hartree_fock.py:38-40); convergence |dE| <= conv and RMS error
 * <= conv, or maxit iterations.  precision / fp32_threshold: the J/K pass.
 *   qx_scf_init: the first step from D0 (eager; allocates what capture needs).
 *   qx_scf_step: n iterations, replays of one captured graph when use_graph
 *     (captured on first use; a molecule that has converged is skipped by
 *     every dense kernel), else the same kernels launched directly.
 *   qx_scf_status: synchronises the stream; per molecule E, dE, RMS error,
 *     iterations, done flag and the idempotency defect tr(X - X^2) before
 *     the last McWeeny step (host arrays, any may be NULL).
 *   qx_scf_matrix: 0 = D, 1 = D_old (the density that built F), 2 = F (as
 *     used for D, DIIS-extrapolated with DIIS), 3 = F before extrapolation.
 * Graph replays and other passes of ctx must run on one stream (or the
 * caller synchronises between them).  Errors: qx_scf_last_error. */
typedef struct qx_scf qx_scf;
int qx_scf_create(qx_ctx* ctx, int nmol, int nm, const int* nbf, const int* offset, const int* nocc,
                  const double* enuc, const double* H, const double* S, const double* A, int diis, int diis_space,
                  int spectral_shift, double conv, int maxit, int purify_iters, int precision,
                  double fp32_threshold, qx_scf** out);
int qx_scf_init(qx_scf* scf, const double* D0, void* stream);
int qx_scf_capture(qx_scf* scf, void* stream);
int qx_scf_step(qx_scf* scf, int n, int use_graph, void* stream);
int qx_scf_status(qx_scf* scf, double* energy, double* de, double* drms, int* iterations, int* done,
                  double* idempotency, void* stream);
int qx_scf_matrix(qx_scf* scf, int which, double* out, void* stream);
void qx_scf_destroy(qx_scf* scf);
const char* qx_scf_last_error(const qx_scf* scf);

/* Quartet tiles (no reference counterpart): threads per workgroup -- the
 * primitive quartets one workgroup evaluates, and the quartets that share
 * its LDS K' rows -- of the streaming kernels of each class, per pass
 * (0 = J/K, 1 = adjoint gradient); block[qx_num_classes()], multiples of
 * 64 up to 256, or NULL for the built-in table (picked by the autotuner,
 * quax_amd/csrc/tune.py --tiles, from GPU timings). */
int qx_set_class_tiles(qx_ctx* ctx, int pass, const int* block);
int qx_get_class_tiles(const qx_ctx* ctx, int pass, int* block);
/* Per class (flags[qx_num_classes()], NULL = built-in table): whether the
 * adjoint-gradient kernels sum the terms of centres that are the same atom
 * on every lane of a wavefront across the wavefront before their LDS
 * accumulators (cheaper LDS atomics, ~200 cross-lane operations per wave);
 * the autotuner picks it per class. */
int qx_set_grad_wave_reduce(qx_ctx* ctx, const int* flags);
int qx_get_grad_wave_reduce(const qx_ctx* ctx, int* flags);
/* Per class (copies[qx_num_classes()], 1..64, NULL = built-in table): how
 * many copies of the 3N gradient accumulators the blocks of its
 * adjoint-gradient kernels spread their flushes over (block index mod
 * copies); the copies are summed once per pass. */
int qx_set_grad_copies(qx_ctx* ctx, const int* copies);
int qx_get_grad_copies(const qx_ctx* ctx, int* copies);
/* Per class (flags[qx_num_classes()], NULL = all on): whether its J/K
 * kernels stage the K' rows of a block's bra pair in LDS and flush them once
 * per block (where the kernel's rows fit and the context's row width allows
 * it), or add each K' term to global memory directly. */
int qx_set_jk_rows(qx_ctx* ctx, const int* flags);
int qx_get_jk_rows(const qx_ctx* ctx, int* flags);
/* Widest atom window (0..256 atoms; 0 = 256) of the adjoint-gradient
 * kernels' LDS accumulators.  A block accumulates its quartets' gradient
 * terms in LDS over the atoms from the lowest to the highest it touches when
 * that window fits, else adds them to global memory; in a batch context a
 * block's quartets lie within one or two molecules, so the window stays
 * narrow however many atoms the batch holds.  A narrower window only moves
 * blocks to the global path (used by the tests to exercise it). */
int qx_set_grad_window(qx_ctx* ctx, int atoms);
int qx_get_grad_window(const qx_ctx* ctx, int* atoms);

/* qx_jk_build for a batch context (qx_initialize_fragments) that leaves out
 * the quartets of every fragment f with frag_skip[f] != 0 (device int array,
 * one entry per fragment id): their blocks of J and K come out zero.  Blocks
 * of quartets that are all left out exit at once, so a batch whose
 * molecules converge at different iterations stops paying for the
 * converged ones (no reference counterpart). */
int qx_jk_build_masked(qx_ctx* ctx, const double* D, double* J, double* K, int precision, double fp32_threshold,
                       const int* frag_skip, void* stream);

/* Independent class kernels of a J/K or gradient pass are spread over
 * nstreams streams (1..8; default 4): the caller's stream plus side streams
 * of default priority that join back into it before the call returns. */
int qx_set_streams(qx_ctx* ctx, int nstreams);

/* Measurement support (no reference counterpart).
 * qx_class_quartets: canonical screened quartets per kernel class (eri[]) and
 * those not on a single atom (deri[], the derivative work); arrays of
 * qx_num_classes() entries.  qx_class_info: angular momenta (l[4]) and the
 * number of derivative kernel parts of class cls.
 * qx_profile_enable(1) resets and starts recording a HIP event pair around
 * every kernel launch, on the stream the kernel runs on (with
 * qx_set_streams(ctx, 1) every kernel is timed alone); qx_profile_read
 * synchronises and returns, per
 * slot, the summed kernel time in ms, the launch count and the primitive-pair
 * quartets processed.  Arrays hold qx_profile_slots() entries;
 * qx_profile_slot names a slot: kind 0 = array-form ERI kernel of class
 * cls, kind 1 = derivative part group `part` of class cls, kind 2 = J/K part
 * `part` of class cls; fp32 = FP32 variant. */
int qx_num_classes(void);
int qx_profile_slots(void);
int qx_profile_slot(int slot, int* kind, int* cls, int* part, int* fp32);
int qx_class_info(int cls, int* l4, int* nparts);
int qx_class_quartets(const qx_ctx* ctx, long long* eri, long long* deri);
/* Algorithmic FP operations (mul/add 1, fma 2) of one primitive-quartet ERI
 * body of class cls (J/K epilogue excluded) and of each derivative part
 * (adjoint-gradient epilogue included), and the number of derivative blocks
 * each part produces.  parts / ncomb hold qx_max_parts() entries. */
int qx_class_flops(int cls, double* eri, double* parts, int* ncomb);
/* Most derivative parts (and part groups) one class can have: the size of
 * the per-part arrays of qx_class_flops and qx_class_groups. */
int qx_max_parts(void);
/* The adjoint-gradient pass launches one kernel per group of derivative
 * parts of a class (blockIdx.y = part): number of groups, FP operations per
 * primitive quartet of each group (sum over its parts) and parts per group;
 * arrays of qx_max_parts() entries.  Derivative profile slots of the gradient
 * pass are indexed by group. */
int qx_class_groups(int cls, int* ngroups, double* flops, int* nparts);
/* The J/K pass launches one kernel per range of a components of a class:
 * number of parts and FP operations per primitive quartet of each (J/K
 * epilogue included); flops holds up to 8 entries. */
int qx_class_jk(int cls, int* nparts, double* flops);
int qx_profile_enable(qx_ctx* ctx, int on);
int qx_profile_read(qx_ctx* ctx, double* ms, long long* launches, long long* pquartets);

#ifdef __cplusplus
}
#endif
#endif
