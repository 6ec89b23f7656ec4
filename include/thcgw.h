/*
 * thcgw.h - C ABI of the MI355X THC imaginary-time GW/RPA library (libthcgw.so).
 *
 * The library computes the inner loop of CoQui's THC GW/RPA drivers on one GPU:
 * the polarizability chi0(tau,q), the screened interaction W (Dyson), the
 * self-energy Sigma(tau,k), the RPA correlation energy and the G0W0
 * quasiparticle energies.  Each entry point replaces one host routine of the
 * reference (paths relative to the CoQui repository root) and keeps its array
 * layout and conventions.
 *
 * Conventions
 *   - complex arrays are interleaved complex<double> (re, im), C order;
 *   - device pointers unless stated otherwise; the caller owns every buffer;
 *   - `stream` is a hipStream_t (0 = default stream); work is enqueued on it.
 *     Entry points that return a host value (Ec, energies) synchronise the stream;
 *     the others return once the work is enqueued unless stated otherwise;
 *   - every function returns THCGW_OK (0) or an error code; the message is
 *     available from thcgw_last_error(ctx).  No C++ exception crosses the ABI;
 *   - the entry points that factorise (thcgw_dyson_w, thcgw_rpa_ec, thcgw_gw_iteration,
 *     thcgw_batch_gw_iteration, thcgw_dyson_g, thcgw_nelec, thcgw_update_mu,
 *     thcgw_lri_dyson) first check their input arrays and return THCGW_ERR_NONFINITE,
 *     before any factorisation, when one holds a NaN or an infinity (one read pass);
 *   - ns = 1 (no spin polarisation), npol = 1: k-points form a Gamma-centred
 *     kgrid[0] x kgrid[1] x kgrid[2] mesh in C order with lattice vectors = identity,
 *     and q-points = k-points; symmetry only where flag 9, thcgw_set_qsym or
 *     thcgw_set_bz says so;
 *   - work enqueued on `stream` may fork to the context's own streams and joins back to
 *     `stream` before the call returns or its later work runs (flag 11).
 *
 * Layouts (names as in the reference):
 *   G_tskij     (nt, 1, nk, no, no)      Green's function on the fermionic IR tau mesh
 *   X_skPa      (1, nk, nmu, no)         THC collocation X^k[P,a]   (thc_reader_t::X)
 *   V_qPQ       (nk, nmu, nmu)           THC Coulomb, incl. the N_k factor (thc_reader_t::Z)
 *   Pi_tqPQ     (nt_half, nk, nmu, nmu)  polarizability on tau in [0, beta/2]
 *   Pi_wqPQ     (nw_half, nk, nmu, nmu)  polarizability / W - V on bosonic iW_n >= 0
 *   W_qtPQ      (nk, nt_half, nmu, nmu)  W - V on tau (MBState::dW_qtPQ)
 *   Sigma_tskij (nt, 1, nk, no, no)      self-energy on the fermionic tau mesh
 * with nt_half = ceil(nt/2), nw_half = ceil(nw_b/2).
 */
#ifndef THCGW_H
#define THCGW_H

#ifdef __cplusplus
extern "C" {
#endif

#define THCGW_OK 0
#define THCGW_ERR_ARG 1
#define THCGW_ERR_HIP 2
#define THCGW_ERR_SINGULAR 3
#define THCGW_ERR_UNSUPPORTED 4
#define THCGW_ERR_ALLOC 5
#define THCGW_ERR_NONFINITE 6 /* an input array holds a NaN or an infinity */

typedef struct thcgw_ctx thcgw_ctx;

typedef struct thcgw_dims {
  int kgrid[3]; /* k-mesh; nk = kgrid[0]*kgrid[1]*kgrid[2] */
  int no;       /* primary basis size (MF::nbnd) */
  int nmu;      /* THC auxiliary basis size (thc_reader_t::Np) */
  int ns;       /* spin channels; must be 1 */
  int nt;       /* IR tau points, fermionic = bosonic (IAFT::nt_f, nt_b) */
  int nw_b;     /* bosonic Matsubara points (IAFT::nw_b) */
  int nw_f;     /* fermionic Matsubara points (IAFT::nw_f) */
  double beta;  /* inverse temperature (Ha^-1) */
} thcgw_dims;

/* IR tables in the dimensionless form stored in the reference's ir/data/<lambda>.<eps>.h5 files.
 * They are rescaled with beta as in src/numerics/imag_axes_ft/ir/ir_driver.hpp:68-73.
 * Complex tables are interleaved (re, im) host arrays. */
typedef struct thcgw_ir_tables {
  const double* tau_mesh_f; /* nt, x in [-1, 1] (fermion/tau_mesh) */
  const long* wn_mesh_f;    /* nw_f Matsubara indices (fermion/wn_mesh) */
  const double* Ttw_ff;     /* nt x nw_f */
  const double* Twt_ff;     /* nw_f x nt */
  const double* Ttw_bb;     /* nt x nw_b */
  const double* Twt_bb;     /* nw_b x nt */
  const double* T_beta_t_ff; /* nt (fermion/T_betat) */
} thcgw_ir_tables;

/* Create a context for one system shape.  Replaces the construction of
 * imag_axes_ft::IAFT (src/numerics/imag_axes_ft/IAFT.hpp:34-42) and the shape checks
 * of scr_coulomb_t (src/methods/scr_coulomb/scr_coulomb_t.cpp:12-35): the tau mesh must
 * be symmetric about beta/2 and nt_f == nt_b.  Selects the current HIP device. */
int thcgw_create(const thcgw_dims* dims, const thcgw_ir_tables* tables, thcgw_ctx** out);
void thcgw_destroy(thcgw_ctx* ctx);
/* Message of the last failed call on ctx (or of the last failed thcgw_create if ctx is NULL). */
const char* thcgw_last_error(const thcgw_ctx* ctx);

/* a1: G_tskij from MO coefficients C (nk, no, no) [C_ia: column a = orbital] and
 * energies E (nk, no, double) at chemical potential mu, by the Fermi-function formula.
 * Replaces update_G, src/methods/SCF/qp_scf_common.cpp:82-118. */
int thcgw_update_G(thcgw_ctx* ctx, const double* C, const double* E, double mu, double* G_tskij, void* stream);

/* a2: O_PQ = X O_ab X^H for nb blocks of nk matrices: O_in (nb, nk, no, no) ->
 * O_out (nb, nk, nmu, nmu).  Replaces thc_solver_comm::primary_to_aux,
 * src/methods/HF/thc_solver_comm.hpp:98-132,363-429. */
int thcgw_primary_to_aux(thcgw_ctx* ctx, long nb, const double* X_skPa, const double* O_in, double* O_out,
                         void* stream);

/* No counterpart in the reference: thcgw_primary_to_aux for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results.  The intermediate X_k O_{b,k} is staged in a
 * static device buffer of 64 MiB per GPU (held until the process ends) instead of the
 * context's workspace, and the batch runs in chunks that fit it; no hipMalloc/hipFree.
 * THCGW_ERR_ARG if one N_mu x N_orb complex<double> block exceeds 64 MiB.  Every call on a
 * GPU shares that GPU's buffer: calls that overlap on one GPU (other streams, threads or
 * contexts) race, so serialise them. */
int thcgw_primary_to_aux_noalloc(thcgw_ctx* ctx, long nb, const double* X_skPa, const double* O_in, double* O_out,
                                 void* stream);

/* a10: O_ab = X^H O_PQ X: O_in (nb, nk, nmu, nmu) -> O_out (nb, nk, no, no).
 * Replaces thc_solver_comm::aux_to_primary, thc_solver_comm.hpp:240-304,516-569. */
int thcgw_aux_to_primary(thcgw_ctx* ctx, long nb, const double* X_skPa, const double* O_in, double* O_out,
                         void* stream);

/* a2-a4: RPA polarizability Pi_tqPQ from G_tskij and X.  Replaces
 * scr_coulomb_t::eval_Pi_rpa_Rspace, src/methods/scr_coulomb/rpa_pi.icc:22-180. */
int thcgw_pi(thcgw_ctx* ctx, const double* G_tskij, const double* X_skPa, double* Pi_tqPQ, void* stream);

/* k-space variants (no symmetry: isym = 0, no time reversal) of a4 and a9, the reference's
 * algorithms for symmetry-reduced grids: Pi^k_PQ(tau_i, q) = (sp/N_k) sum_k
 * conj(G_PQ(tau_i, k-q)) G_PQ(beta-tau_i, k) (eval_Pi_rpa_kspace, src/methods/GW/rpa_pi.icc:182-336;
 * equal to Pi^R_QP(-q) of thcgw_pi for Hermitian G) and Sigma_PQ(tau, k) = -(1/N_k) sum_q
 * G_PQ(tau, k-q) W_PQ(tau, q) in both particle-hole passes (eval_Sigma_all_kspace,
 * src/methods/GW/thc_gw.icc:342-443).  Same layouts as thcgw_pi / thcgw_sigma; N_k <= 64.
 * O(N_k^2) per (P,Q) column.  Enqueue only. */
int thcgw_pi_kspace(thcgw_ctx* ctx, const double* G_tskij, const double* X_skPa, double* Pi_tqPQ, void* stream);
int thcgw_sigma_kspace(thcgw_ctx* ctx, const double* G_tskij, const double* W_qtPQ, const double* X_skPa,
                       double* Sigma_tskij, int hermitize, void* stream);

/* f4: Brillouin-zone symmetry of the mean field, in the layout of the reference's
 * bz_symm (src/mean_field/symmetry/bz_symmetry.hpp:497-576).  The N_k k-points are listed
 * with the nk_ibz irreducible ones first and the nk_trev_pairs points that need time
 * reversal last; the q-points are listed the same way.  All tables are host int arrays:
 *   k_grid_index  (N_k)        C-order position (i n2 + j) n3 + l of each listed point on
 *                              the Gamma-centred k-grid of thcgw_dims;
 *   kp_to_ibz, kp_trev, kp_trev_pair (N_k)    as in bz_symm (kp_trev_pair < 0: no pair);
 *   qk_to_k2      (N_k x N_k)  [q][a] = b with k_b = k_a - q;
 *   qminus, qp_to_ibz, qp_trev (N_k);
 *   nq_per_s      (nsym), Qs (nsym x N_k, the first nq_per_s[s] of a row used),
 *   ks_to_k       (nsym x nk_ibz): the operations qsymms of bz_symm, identity first;
 *   D             (nsym x nk_ibz x N_orb x N_orb complex, host doubles): the orbital rotation
 *                 of symmetry_rotation(s, k) (MF.hpp), applied as Sigma(k) += D^H Sigma(k S) D
 *                 (ignored for s = 0).
 * thcgw_set_bz checks every index (N_k <= 64, nq_ibz <= 64, irreducible points first, time-
 * reversal points last with partners among the others, every q in exactly one Qs row,
 * k S never a time-reversal point); bz = NULL clears it. */
typedef struct thcgw_bz {
  int nk_ibz, nq_ibz, nk_trev_pairs, nsym;
  const int* k_grid_index;
  const int* kp_to_ibz;
  const int* kp_trev;
  const int* kp_trev_pair;
  const int* qk_to_k2;
  const int* qminus;
  const int* qp_to_ibz;
  const int* qp_trev;
  const int* nq_per_s;
  const int* Qs;
  const int* ks_to_k;
  const double* D;
} thcgw_bz;
int thcgw_set_bz(thcgw_ctx* ctx, const thcgw_bz* bz);

/* With a symmetry set (FP64, flags 7 and 9 and the q map off): G_tskij holds the nk_ibz
 * irreducible k-points (nt x nk_ibz x N_orb^2), X_skPa all N_k listed points.
 *   thcgw_pi_ibz: Pi_tqPQ (nt_half x nq_ibz x N_mu^2) on the irreducible q by the R-space
 *     algorithm: G_PQ on the points without time reversal, the conjugate of the partner's
 *     on the others, the k-grid transform, the Hadamard product and the transform back to
 *     the irreducible q only (rpa_pi.icc:22-180 with nqpts_ibz < nqpts).
 *   thcgw_sigma_ibz: Sigma_tskij (nt x nk_ibz x N_orb^2) from W_qtPQ (nq_ibz x nt_half x
 *     N_mu^2) by the k-space algorithm: per operation s, the Hadamard sums over the q of
 *     Qs[s] at k S with the time-reversal conjugates, back-projected with X_{kS}, rotated by
 *     D and summed, in both particle-hole passes (thc_gw.icc:152-178, 216-443).
 *   thcgw_gw_iteration_ibz: the iteration of thcgw_gw_iteration on the wedge: Pi, tau ->
 *     iW, Dyson for the irreducible q with V_qPQ (nq_ibz x N_mu^2), iW -> tau, Sigma; Ec
 *     weighs each irreducible q by its multiplicity (thc_rpa.icc:130-160).
 * Enqueue only, except the iteration, which synchronises for the Dyson flags and Ec. */
int thcgw_pi_ibz(thcgw_ctx* ctx, const double* G_tskij, const double* X_skPa, double* Pi_tqPQ, void* stream);
int thcgw_sigma_ibz(thcgw_ctx* ctx, const double* G_tskij, const double* W_qtPQ, const double* X_skPa,
                    double* Sigma_tskij, int hermitize, void* stream);
int thcgw_gw_iteration_ibz(thcgw_ctx* ctx, const double* G_tskij, const double* X_skPa, const double* V_qPQ,
                           double* Sigma_tskij, double* W_qtPQ, double* Ec, int hermitize, void* stream);

/* a5: Pi_tqPQ -> Pi_wqPQ with the particle-hole folded bosonic transform.  Replaces
 * scr_coulomb_t::tau_to_w (scr_coulomb_t.cpp:349-376) -> IAFT::tau_to_w_PHsym (IAFT.icc:49-73). */
int thcgw_tau_to_w(thcgw_ctx* ctx, const double* Pi_tqPQ, double* Pi_wqPQ, void* stream);

/* a6: in place Pi_wqPQ -> W_wqPQ = [(I - V Pi)^{-1} - I] V for every (iW_n, q).
 * Replaces scr_coulomb_t::dyson_W_in_place, scr_coulomb_t.cpp:121-206.
 * If trace_out / logdet_out are non-NULL they receive, per (n, q) in Pi_wqPQ order,
 * Tr(V Pi) (complex, host array of 2*nw_half*nk doubles) and ln|Re det(I - V Pi)|
 * (host array of nw_half*nk doubles), the two terms of the RPA energy
 * (thc_rpa.icc:141-163).  Synchronises the stream; THCGW_ERR_SINGULAR if a factor is singular. */
int thcgw_dyson_w(thcgw_ctx* ctx, double* PiW_wqPQ, const double* V_qPQ, double* trace_out, double* logdet_out,
                  void* stream);

/* a7 + a8: W_wqPQ -> W_qtPQ (q-major, the layout of MBState::dW_qtPQ).  Replaces
 * scr_coulomb_t::w_to_tau (scr_coulomb_t.cpp:421-448) -> IAFT::w_to_tau_PHsym
 * (IAFT.icc:104-128) and the transpose in update_w (scr_coulomb_t.cpp:73-86). */
int thcgw_w_to_tau(thcgw_ctx* ctx, const double* W_wqPQ, double* W_qtPQ, void* stream);

/* a9-a12: Sigma_tskij from G_tskij, W_qtPQ and X, R-space algorithm with both
 * particle-hole passes; hermitize != 0 applies (S + S^H)/2 afterwards.  Replaces
 * gw_t::eval_Sigma_all (src/methods/GW/thc_gw.cpp:152-178 -> thc_gw.icc:70-213),
 * gw_t::Sigma_div_correction for div_treatment = ignore_g0 (thc_gw.icc:460-461) and
 * hermitize (src/methods/SCF/scf_common.hpp:70-86). */
int thcgw_sigma(thcgw_ctx* ctx, const double* G_tskij, const double* W_qtPQ, const double* X_skPa,
                double* Sigma_tskij, int hermitize, void* stream);

/* a12: (A + A^H)/2 for nb square n x n matrices, in place. */
int thcgw_hermitize(thcgw_ctx* ctx, long nb, int n, double* A, void* stream);

/* a13: RPA correlation energy from Pi_wqPQ and V.  Replaces gw_t::thc_rpa_energy_all_impl,
 * src/methods/GW/thc_rpa.icc:72-181 (A = Pi V, Tr A + ln Re det(I - A), w_to_tau, tau_to_beta).
 * Synchronises the stream. */
int thcgw_rpa_ec(thcgw_ctx* ctx, const double* Pi_wqPQ, const double* V_qPQ, double* Ec, void* stream);

/* One GW iteration (the timed unit): chi0 -> tau_to_w -> Dyson -> w_to_tau -> Sigma, plus
 * the RPA energy from the same LU factors when Ec != NULL.  Mirrors
 * scr_coulomb_t::update_w + gw_t::evaluate as called by scf_loop
 * (src/methods/SCF/scf_driver.cpp:119-122) and gw_t::rpa_energy (thc_rpa.cpp:22-61).
 * W_qtPQ (caller-owned, nk*nt_half*nmu^2 complex) receives W - V; it also serves as
 * scratch for Pi_tqPQ.  Synchronises the stream. */
int thcgw_gw_iteration(thcgw_ctx* ctx, const double* G_tskij, const double* X_skPa, const double* V_qPQ,
                       double* Sigma_tskij, double* W_qtPQ, double* Ec, int hermitize, void* stream);

/* a15: G0W0 quasiparticle energies.  Sigma_tskij (nt,1,nk,no,no), MO coefficients C
 * (nk,no,no) and Vhf (nk,no,no, primary basis) are device arrays; E0 (nk,no) starting
 * energies and E_out (nk,no) are host arrays.  Sigma and Vhf are projected on the MO
 * diagonal, Sigma is taken to iw_n, continued by an nfit-point Thiele Pade and the QP
 * equation is solved by bisection with tolerance tol at broadening eta.  Replaces
 * solve_qp_eqn (src/methods/SCF/qp_scf_common.cpp:150-311) with pade_driver / pade_t
 * (src/numerics/ac/pade/pade_driver.hpp:24-65, pade_t.hpp:34-168) and qp_eqn_bisection
 * (qp_scf_common.cpp:679-728).  Synchronises the stream. */
int thcgw_solve_qp(thcgw_ctx* ctx, const double* Sigma_tskij, const double* C, const double* Vhf,
                   const double* E0, double mu, int nfit, double eta, double tol, double* E_out, void* stream);

/* a16: Galitskii-Migdal correlation energy from device G_tskij and Sigma_tskij.  Replaces
 * eval_corr_energy, src/methods/SCF/scf_common.cpp:87-136 (k weights 1/nk).
 * Synchronises the stream. */
int thcgw_gm_energy(thcgw_ctx* ctx, const double* G_tskij, const double* Sigma_tskij, double* e_corr,
                    void* stream);

/* THC Hartree-Fock potential F_skij (1,nk,no,no) of a density matrix Dm_skij (1,nk,no,no)
 * with collocation X_skPa and THC Coulomb V_qPQ (the same V as the GW calls): Hartree
 * J_P = sum_Q V(q=0)_PQ 2 Dm^P_{R=0,QQ} on the PQ diagonal and exchange
 * -(1/N_k) sum_q Dm^P(k-q) o V(q), projected back by X^H . X.  Either part can be
 * switched off.  Replaces hf_t::evaluate -> thc_hf_Xqindep
 * (src/methods/HF/thc_hf.cpp:20-60, src/methods/HF/thc_hf.icc:18-300) for ns = npol = 1,
 * no symmetry and ignore_g0 (HF_K_correction returns at once, hf_t.cpp:49-52).
 * Enqueues only. */
int thcgw_hf(thcgw_ctx* ctx, const double* Dm_skij, const double* X_skPa, const double* V_qPQ, int hartree,
             int exchange, double* F_skij, void* stream);

/* Dyson equation for G on the Matsubara axis: Sigma_tskij (nt,1,nk,no,no) -> Sigma(iw_n)
 * (fermionic), G(iw_n,k) = [(iw_n + mu) S_k - H0_k - F_k - Sigma(iw_n,k)]^-1,
 * G_tskij = w_to_tau(G(iw)), and, if Dm_skij is not NULL, Dm = -G(beta^-).  H0, F, S are
 * (1,nk,no,no); H0 = NULL means zero and S = NULL the identity.  Replaces
 * simple_dyson::solve_dyson (src/methods/SCF/simple_dyson.cpp:11-127) for ns = 1.
 * Returns THCGW_ERR_SINGULAR for a singular matrix.  Synchronises the stream. */
int thcgw_dyson_g(thcgw_ctx* ctx, const double* H0_skij, const double* F_skij, const double* S_skij,
                  const double* Sigma_tskij, double mu, double* G_tskij, double* Dm_skij, void* stream);

/* Electron number per cell at chemical potential mu for the same Dyson equation:
 * -2 tau_to_beta(w_to_tau(sum_k w_k Tr[G(iw_n,k) S_k])).  Replaces compute_Nelec
 * (src/methods/SCF/scf_common.cpp:10-42) with the spectral sum of compute_eigenspectra
 * (simple_dyson.cpp:130-172), sum_b 1/(iw_n + mu - lambda_b), taken as that resolvent
 * trace.  Synchronises the stream. */
int thcgw_nelec(thcgw_ctx* ctx, const double* H0_skij, const double* F_skij, const double* S_skij,
                const double* Sigma_tskij, double mu, double* nelec, void* stream);

/* Chemical potential for nelec_target: bracket from old_mu in steps of 0.2 Ha, then bisect
 * until |nelec - target| < mu_tol (at most 200 halvings).  Replaces update_mu
 * (src/methods/SCF/scf_common.cpp:147-202).  nelec_out may be NULL.  Synchronises. */
int thcgw_update_mu(thcgw_ctx* ctx, const double* H0_skij, const double* F_skij, const double* S_skij,
                    const double* Sigma_tskij, double nelec_target, double old_mu, double mu_tol, double* mu_out,
                    double* nelec_out, void* stream);

/* Batched form (no counterpart in the reference, which runs one system per job): a batch of
 * independent systems of uneven size, one context each, iterated together on one stream.
 * The contexts keep their own flags (precision, Dyson form) and share one scratch pool
 * (they must live on the same device and may not be used outside the batch meanwhile).
 * The first thcgw_batch_gw_iteration runs eagerly: it allocates the scratch, fixes the
 * chunk sizes and, unless switched off, times the three FP32 tile shapes on every
 * mixed-precision system and keeps the fastest.  Later calls capture the whole batch
 * iteration into one HIP graph (recaptured when any pointer, the stream or hermitize
 * changes) and replay it; the Cholesky/LU flags and Ec are read after the replay, and
 * the captured Hermitian form has no LU fallback (a failure is reported as
 * THCGW_ERR_SINGULAR).  G[i], X[i], V[i], Sigma[i], W_qt[i] are the per-system arrays of
 * thcgw_gw_iteration; Ec (nsys doubles) may be NULL.  Synchronises the stream.  Destroy
 * the batch before its contexts. */
typedef struct thcgw_batch thcgw_batch;
int thcgw_batch_create(int nsys, thcgw_ctx* const* ctxs, thcgw_batch** out);
void thcgw_batch_destroy(thcgw_batch* b);
const char* thcgw_batch_last_error(const thcgw_batch* b);
int thcgw_batch_set(thcgw_batch* b, int use_graph, int tune);
/* Spreads the systems over `lanes` (1..8) streams, largest Dyson work first onto the least
 * loaded lane; the systems of a lane run one after another and share that lane's scratch
 * pool, and the captured graph holds the lanes as parallel branches (the eager, uncaptured
 * path stays one system after another).  Each lane's automatic chunk sizes take 1/lanes of
 * the memory a single pool would.  Only before the first thcgw_batch_gw_iteration. */
int thcgw_batch_set_lanes(thcgw_batch* b, int lanes);
int thcgw_batch_gw_iteration(thcgw_batch* b, const double* const* G_tskij, const double* const* X_skPa,
                             const double* const* V_qPQ, double* const* Sigma_tskij, double* const* W_qtPQ,
                             double* Ec, int hermitize, void* stream);

/* a8 + a11: treatment of the long-wavelength divergence (Gygi).  div_treatment: 0 = ignore_g0,
 * 1 = gygi (head of the smallest nonzero |q|), 2 = gygi_extrplt (per bosonic frequency a
 * least-squares quadratic in |q| through the q with 0 < |q| <= 0.8, distinct |q|,
 * evaluated at q = 0; the first such q when there are fewer than 3), 3 = gygi_extrplt_2d
 * (the same over q_z = 0).
 * thcgw_eps_inv_head: eps^-1_head(tau, q) = (|q|^2/4pi) Vol chi_bar_q^T W(tau, q) conj(chi_bar_q)
 * from W_qtPQ (thcgw_w_to_tau's output), chi_bar_qP = thc_reader_t::basis_bar_head()
 * (device, (nk, nmu) complex), q-points in Cartesian units (host, (nk, 3)) and the cell
 * volume; eps_inv_tq (host (nt_half, nk) complex, may be NULL) receives every q and
 * eps_inv_t (host nt_half complex) the q -> 0 head.  Replaces div_utils::eps_inv_head_t /
 * eval_eps_inv_q (src/methods/GW/g0_div_utils.hpp:55-111,180-214).  Synchronises.
 * thcgw_sigma_div_correction: Sigma(tau, k) += -madelung Re eps^-1_head(tau') T^k G(tau, k) T^k^H
 * with T^k = X^k^H diag(conj chi_head) X^k, tau' = tau folded to [0, beta/2], chi_head =
 * basis_head()(0, :) (device nmu complex); a no-op for ignore_g0 or one k-point.  Replaces
 * gw_t::Sigma_div_correction (src/methods/GW/thc_gw.icc:446-526).  Synchronises. */
int thcgw_eps_inv_head(thcgw_ctx* ctx, const double* W_qtPQ, const double* chi_bar_qP, const double* qpts_cart,
                       double volume, int div_treatment, double* eps_inv_tq, double* eps_inv_t, void* stream);
int thcgw_sigma_div_correction(thcgw_ctx* ctx, double* Sigma_tskij, const double* G_tskij, const double* X_skPa,
                               const double* chi_head_P, double madelung, const double* eps_inv_t, int div_treatment,
                               void* stream);

/* Learned form of the Dyson solve (no counterpart in the reference; it replaces the
 * getrf/getri of scr_coulomb_t.cpp:180-196 inside a self-consistent loop).  With
 * V = L L^H and A = -L^H Pi(iW_n) L, W - V = L M^-1 L^H - V, M = I + A.  The Krylov solve
 * runs preconditioned conjugate gradients in FP64 on M X = [L^H | Q] column by column,
 * where Q is an orthonormal basis (n x rank) of the sketch A Omega (Omega a fixed
 * Gaussian) and the preconditioner is the rank-`rank` approximate inverse
 * P^-1 = I - Q C Q^H, also used as the warm start X0 = P^-1 B.  The core C comes from
 *   mode 0: none (C = 0, X0 = 0); mode 1: Woodbury on span(Q), C_W = T (I+T)^-1 with
 *   T = Q^H A Q; mode 2: the exact core C* = I - Q^H M^-1 Q of the previous call carried
 *   into the current basis, C_tr = O^H C*_prev O with O = Q_prev^H Q (mode 1 on the
 *   first call); mode 3 (default): C_W plus the correction predicted by a three-layer
 *   network (width `width`, BF16 weights, FP32 accumulation, SiLU) from (T, C_W, C_tr)
 *   -- trained by thcgw_lri_train on pairs recorded by direct solves.
 * A Krylov solve is accepted only when every column's true FP64 residual
 * ||B - M X|| / ||B|| is <= tol; otherwise its chunk is solved directly (Cholesky, as
 * flag 6 = 1).  krylov = 0 always solves directly.  record bit 0: store the pair
 * (features, C* - C_W) of every (n, q) for training; bit 1: keep the previous state
 * (otherwise every call leaves the exact core and the basis Q of each (n, q) as the
 * "previous" state of the next call, as one iteration of a self-consistent loop).  PiW and V as
 * in thcgw_dyson_w; synchronises.  stats may be NULL; core_err is
 * ||C_mode - C*||_F / ||C*||_F over the call's direct solves. */
typedef struct thcgw_lri thcgw_lri;
typedef struct thcgw_lri_stats {
  int solves;           /* (n, q) matrices processed */
  int krylov_accepted;  /* of which solved by the Krylov form */
  int fallbacks;        /* of which solved directly after the residual check failed */
  int iterations;       /* largest CG iteration count of a chunk */
  double iter_mean;     /* mean CG iterations per chunk */
  double resid_max;     /* largest accepted relative residual */
  double core_err;      /* see above */
} thcgw_lri_stats;
int thcgw_lri_create(thcgw_ctx* ctx, int rank, int width, unsigned long long seed, thcgw_lri** out);
void thcgw_lri_destroy(thcgw_lri* L);
const char* thcgw_lri_last_error(const thcgw_lri* L);
int thcgw_lri_set(thcgw_lri* L, int mode, int max_iter, double tol);
int thcgw_lri_dyson(thcgw_lri* L, double* PiW_wqPQ, const double* V_qPQ, int krylov, int record,
                    thcgw_lri_stats* stats, void* stream);
/* Adam on the mean squared error of the recorded pairs (targets scaled to unit RMS at the
 * first call), shuffled minibatches of `batch`; mean loss of the first and last epoch. */
int thcgw_lri_train(thcgw_lri* L, int epochs, int batch, float lr, unsigned long long seed, float* loss_first,
                    float* loss_last, void* stream);
long thcgw_lri_num_pairs(const thcgw_lri* L);
int thcgw_lri_clear_pairs(thcgw_lri* L);
int thcgw_lri_reset_state(thcgw_lri* L);

/* The network alone, for checks of its kernels: y = silu(silu(x W0^T + b0) W1^T + b1) W2^T + b2
 * (row-major batch x n_in input); parameters as one FP32 array W0, b0, W1, b1, W2, b2
 * (torch.nn.Linear layout) with a BF16 mirror used by the products.  backward writes the
 * gradient of sum(Y * dY) w.r.t. every parameter (device array of num_params floats) for
 * the X of the last forward; adam_step applies one Adam step (beta 0.9/0.999, eps 1e-8);
 * train_mse does forward, MSE gradient, backward and Adam, returning the loss before the
 * step. */
typedef struct thcgw_mlp thcgw_mlp;
int thcgw_mlp_create(int n_in, int width, int n_out, unsigned long long seed, thcgw_mlp** out);
void thcgw_mlp_destroy(thcgw_mlp* m);
const char* thcgw_mlp_last_error(const thcgw_mlp* m);
long thcgw_mlp_num_params(const thcgw_mlp* m);
int thcgw_mlp_get_params(thcgw_mlp* m, float* host);
int thcgw_mlp_set_params(thcgw_mlp* m, const float* host);
int thcgw_mlp_forward(thcgw_mlp* m, int batch, const float* X, float* Y, void* stream);
int thcgw_mlp_backward(thcgw_mlp* m, int batch, const float* X, const float* dY, float* grad, void* stream);
int thcgw_mlp_adam_step(thcgw_mlp* m, const float* grad, float lr, void* stream);
int thcgw_mlp_train_mse(thcgw_mlp* m, int batch, const float* X, const float* T, float lr, float* loss, void* stream);

/* q-point symmetry of the THC auxiliary basis (the IBZ reduction of scr_coulomb_t, which
 * solves Dyson for the irreducible q only, src/methods/scr_coulomb/scr_coulomb_t.cpp:121-206
 * over nqpts_ibz, with the basis rotations of thc_gw.icc:314-317 restricted to phased
 * permutations of the THC points).  For every q: src[q] is the irreducible q it is
 * generated from (src[src[q]] == src[q]); perm (nk x nmu, host) and phase (nk x nmu
 * complex, host, may be NULL for 1) give
 *   A(q)[P][Q] = phase[q][P] conj(phase[q][Q]) A^(src q)[perm[q][P]][perm[q][Q]]
 * for A = V, Pi and W, where A^ = conj A if conj[q] != 0 (time reversal combined; conj may
 * be NULL).  Irreducible q map to themselves with unit phases.  With a map set, the Dyson
 * step (thcgw_dyson_w, thcgw_rpa_ec, thcgw_gw_iteration) solves the irreducible q only and
 * fills the others; it refuses V or Pi(iW_0) that do not follow the map (1e-9 relative in FP64,
 * 1e-5 in the mixed forms, whose Pi carries FP32 rounding).
 * It takes precedence over flag 9.  src = NULL clears the map. */
int thcgw_set_qsym(thcgw_ctx* ctx, const int* src, const int* perm, const double* phase, const int* conj);

/* Diagnostics (no counterpart in the reference): bytes of device workspace held by ctx,
 * control flags and GEMM launch timing.
 *   flag 1: use the generic (LDS) k-grid transform kernel instead of the specialised ones;
 *   flag 2: tau pairs per projection batch (0 = automatic);
 *   flag 3: matrices per Dyson batch (0 = automatic: equal batches within 60% of free device
 *           memory, at most 64 GiB, counting the batch buffers ctx already holds);
 *   flag 4: record HIP events around every complex-FP64 GEMM launch (process-wide);
 *   flag 5: GEMM tile shape (process-wide): 0 automatic, 1 = 64x64, 2 = 128x64, 3 = 128x128;
 *   flag 6: Dyson algorithm: 0 = LU with partial pivoting (the reference's algorithm,
 *           default), 1 = Hermitian form V = L L^H, M = I - L^H Pi L = L_M L_M^H (two
 *           Cholesky factorisations, triangular products; valid for Hermitian positive
 *           definite V and Hermitian negative semidefinite Pi, with per-chunk fallback to LU);
 *           2 = resolvent form W + V = (V^-1 - Pi)^-1: V_q^-1 and ln det V_q once per q, then
 *           per matrix one Cholesky factorisation of A = V^-1 - Pi, the triangular inverse
 *           L_A^-H and the product L_A^-H L_A^-1, all FP64 (n^3/2 complex multiply-adds
 *           against 7n^3/3 for LU), with ln det(I - V Pi) = ln det V + ln det A.  Same
 *           validity and fallback as 1; its rounding error grows with the condition number
 *           of V (W to ~1e-9 relative for cond(V) ~ 1e7).  With flag 7 = 1 or 2 the Dyson
 *           step stays in this FP64 form and only the projections and k-grid transforms run
 *           in FP32;
 *   flag 11: 2 = the resolvent form's (flag 6 = 2) matrix chunks pipelined over two streams
 *           of the context, each with half the automatic chunk size: one chunk's
 *           factorisation runs while the host reads the previous chunk's flags and that
 *           chunk's Z Z^H runs; 1 (default) = one stream.  Results are bit-identical.  A
 *           batch's captured graph keeps one stream per system;
 *   flag 10: G(tau, k) is Hermitian (as for any G built from a Hermitian H(k) and Fermi
 *           occupations): the projections X G X^H for chi0 and Sigma form their lower tiles
 *           and mirror them, half the work of that product, and Sigma's back-projection
 *           X^H S X of the then Hermitian S is formed as Y + Y^H, Y = X^H L X with L the
 *           lower half of S (Sigma is returned Hermitian).  thcgw_pi, thcgw_sigma and
 *           thcgw_gw_iteration refuse G with max |G - G^H| > 1e-12 max |G| (THCGW_ERR_ARG);
 *   flag 8: FP32 GEMM tile shape (process-wide; products without FP64 chunk accumulation):
 *           0 = 64x64, 1 = 128x64, 2 = 128x128;
 *   flag 12: complex GEMMs on 64x64 tiles by three real MFMA products per complex product
 *           instead of four (process-wide mask; 0 = four everywhere, the default): bit 0 the
 *           FP64 products, bit 1 the FP32 products without FP64 folding, bit 2 the FP32
 *           products with FP64 folding (the mixed forms' back-projection of Sigma and the
 *           mixed Hermitian Dyson products), bit 3 the FP32 projections of G for Sigma
 *           only (bit 1 also covers those for chi0, whose rounding feeds Ec).  Three products (Gauss: Re = ArBr - AiBi,
 *           Im = (Ar + Ai)(Br + Bi) - ArBr - AiBi) take a quarter fewer MFMA cycles; the real
 *           part rounds as before, the imaginary part's error is bounded by
 *           eps (|Ar| + |Ai|)(|Br| + |Bi|) per term rather than eps |A||B| (normwise stable,
 *           Higham 1992).  Products on the 128-wide tiles (flags 5, 8) and split-BF16
 *           products keep four;
 *   flag 7: precision: 0 = FP64 throughout (default); 1 = mixed: the projections X G X^H and
 *           X^H Sigma X, the k-grid transforms with the Hadamard products, and the products
 *           of the Hermitian Dyson form run in FP32 (complex<float> intermediates), while
 *           the tau <-> iW transforms, the Cholesky factorisations and ln det stay FP64;
 *           2 = as 1 with every FP32 product done as split-BF16 MFMA products (three bf16
 *           parts per value, six part products, FP32 accumulation).
 *           Inputs and outputs keep their complex<double> layouts.
 * thcgw_profile_read fills, for each of the 17 GEMM classes, the summed kernel time ms[c]
 * of the launches recorded since the last read, their modelled flops[c] (8 M N K per
 * complex product, reduced for triangular products), their modelled algorithmic HBM
 * bytes[c] (each operand read once, C written once and read once when beta != 0, halved
 * for a triangular operand or output; bytes may be NULL) and launches[c], then drops the
 * records.  Classes 0..8 are the FP64 products c = opA*3 + opB (op 0 = N, 1 = T, 2 = C);
 * 9..12 the FP32 products NN, NC, CN, CC; 13..16 the split-BF16 products NN, NC, CN, CC.
 * It waits for the recorded launches to finish. */
long thcgw_workspace_bytes(const thcgw_ctx* ctx);
int thcgw_set_flag(thcgw_ctx* ctx, int flag, long value);
int thcgw_profile_read(thcgw_ctx* ctx, double* ms, double* flops, double* bytes, long* launches);

#ifdef __cplusplus
}
#endif

#endif /* THCGW_H */
