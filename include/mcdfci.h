/*
 * mcdfci.h -- C ABI of the MI355X coordinate-descent FCI (CDFCI) library.
 *
 * The reference (ninotreve/mcdfci) is a header-only C++ program with no
 * plugin or FFI surface; its hot path is the solver loop of
 * src/solver.h:672-893 over the interfaces
 *   Hamiltonian<N>        src/hamiltonian.h:25-54
 *   WaveFunction<N>       src/wavefunction.h:163-221
 *   CoordinatePick<N>     src/solver.h:94-101
 *   CoordinateUpdate<N>   src/solver.h:230-250
 * Each entry point below names the reference code it replaces.  Everything
 * runs on one HIP device; host arrays belong to the caller and are copied.
 *
 * Determinants cross this boundary in the reference's interleaved format:
 * spin-orbital 2p is alpha of spatial orbital p and 2p+1 is beta
 * (src/input.h:242-243); bit s%64 of word s/64 is set when spin-orbital s is
 * occupied (src/determinant.h:24-100).  Arrays of determinants hold
 * `nwords` uint64 words per determinant (mcdfci_info).
 *
 * Every call returns a status (MCDFCI_OK = 0); mcdfci_last_error(ctx)
 * gives the message.  Calls on one context are not re-entrant.
 */
#ifndef MCDFCI_H
#define MCDFCI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MCDFCI_OK = 0,
  MCDFCI_E_INPUT = 1,       /* invalid argument or file (std::invalid_argument in the reference) */
  MCDFCI_E_OOM = 2,         /* device allocation failed */
  MCDFCI_E_TABLE_FULL = 3,  /* std::overflow_error of src/wavefunction.h:594-599 */
  MCDFCI_E_HIP = 4,         /* HIP runtime error */
  MCDFCI_E_NUMERIC = 5,     /* subspace eigenvalue >= 0: sqrt(-lambda) undefined (src/solver.h:524,591) */
  MCDFCI_E_CANDIDATES = 6,  /* fewer than k distinct candidates (src/solver.h:218-223 reads past the end) */
  MCDFCI_E_UNSUPPORTED = 7  /* a configuration outside what the device path implements */
};

enum { MCDFCI_PREC_FP64 = 0, MCDFCI_PREC_MIXED = 1 };

/* coordinate_update (src/run_wrapper.h:145-150): the subspace eigen-update
 * (src/solver.h:319-610) or the line search on one coordinate
 * (src/solver.h:254-314; num_of_coordinate must be 1; |x|^2 and scale are
 * maintained, which the reference omits). */
enum { MCDFCI_UPDATE_EIG = 0, MCDFCI_UPDATE_LS = 1 };

typedef struct mcdfci_ctx mcdfci_ctx;

/* Options.  Meanings and defaults follow src/run_wrapper.h:64-87. */
typedef struct {
  int num_of_coordinate; /* k */
  int estimate;          /* M: per-column candidate list length */
  double z_threshold;    /* compression gate |dz*scale| > z_threshold (src/wavefunction.h:518) */
  double max_memory;     /* GB for the wavefunction table (src/run_wrapper.h:203-213) */
  double threshold;      /* hamiltonian.threshold: table entries need |v| > threshold */
  int precision;         /* MCDFCI_PREC_FP64 or MCDFCI_PREC_MIXED */
  int device;            /* HIP device ordinal */
  int coordinate_update; /* MCDFCI_UPDATE_EIG (default) or MCDFCI_UPDATE_LS */
} mcdfci_opts;

/* Spin-restricted Hamiltonian in spatial orbitals (what an RHF FCIDUMP holds). */
typedef struct {
  int norb;                      /* spatial orbitals */
  int nelec;
  int ms2;
  double core;
  const double* h;               /* norb*norb one-electron integrals */
  const double* eri;             /* norb^4 chemists' (pq|rs), fully 8-fold symmetric */
  const double* orbital_energy;  /* norb entries, or NULL (HF = lowest orbitals) */
} mcdfci_ham;

/* Spin-unrestricted Hamiltonian by spin block (what a UHF FCIDUMP holds,
 * src/input.h:125-127,203-237, with spin-orbital 2p the alpha and 2p+1 the
 * beta function of orbital p).  The (beta beta|alpha alpha) block is eri_ab
 * with the pairs exchanged. */
typedef struct {
  int norb;                      /* orbitals per spin (half the spin-orbitals) */
  int nelec;
  int ms2;
  double core;
  const double* h_alpha;         /* norb*norb */
  const double* h_beta;          /* norb*norb */
  const double* eri_aa;          /* norb^4 (p_a q_a|r_a s_a) */
  const double* eri_bb;          /* norb^4 (p_b q_b|r_b s_b) */
  const double* eri_ab;          /* norb^4 (p_a q_a|r_b s_b) */
  const double* orbital_energy;  /* 2*norb spin-orbital entries, or NULL */
} mcdfci_ham_uhf;

typedef struct {
  double energy;        /* xz / xx (src/wavefunction.h:195-198) */
  double dx_max;        /* max |dx| of the last iteration */
  int64_t x_size;       /* |x|_0 */
  int64_t z_size;       /* |z|_0 */
  int64_t cols_size;    /* |H[I]|_0: sum of the last iteration's column lengths */
  int64_t updates;      /* upds */
  double scale;         /* scale (hi part) */
  double xx;            /* ||x||^2 (hi part) */
  double xz;            /* x.z */
  int64_t iterations;   /* completed iterations after initialisation */
  /* eig statistics of the [LOG] line (src/solver.h:787-798) */
  double max_residual;
  double curr_max_abs_evec;
  int64_t sum_iterations;
  int64_t sum_inv_iter;
  int64_t curr_inv_iter;
  double sum_diff_inv_iter;
  double curr_diff_inv_iter;
  double max_diff_inv_iter;
  int64_t table_slots;  /* device table capacity (slots) */
  double max_size;      /* the reference's max_wavefunction_size */
  int64_t candidates;   /* length of the candidate list for the next pick (sub_xz.size()) */
  /* subspace solves whose Ritz vector failed the residual test after the
   * last restart and was used regardless (no counterpart: the reference
   * uses its LAPACK eigenvector unchecked, src/solver.h:384-445) */
  int64_t eig_unconverged;
} mcdfci_report;

const char* mcdfci_last_error(const mcdfci_ctx* ctx);

/* Replaces Fcidump + HamiltonianMolecule<N>(fci, threshold) construction
 * (src/input.h:88-152, src/hamiltonian.h:98-109) and the table sizing of
 * src/run_wrapper.h:181,203-213.  mcdfci_create_uhf takes unrestricted
 * integrals (the Fcidump::read_integral_uhf path, src/input.h:203-237);
 * mcdfci_create_from_fcidump reads either kind of file and refuses
 * integrals that change an electron's spin (MCDFCI_E_INPUT) and an
 * unrestricted file with an odd number of spin-orbitals. */
int mcdfci_create(const mcdfci_ham* ham, const mcdfci_opts* opts, mcdfci_ctx** out);
int mcdfci_create_uhf(const mcdfci_ham_uhf* ham, const mcdfci_opts* opts, mcdfci_ctx** out);
int mcdfci_create_from_fcidump(const char* path, const mcdfci_opts* opts, mcdfci_ctx** out);
int mcdfci_destroy(mcdfci_ctx* ctx);

/* info[0..7] = {nwords, spin-orbitals, nelec, ms2, column enumeration size,
 * k, 1 if the FCIDUMP carried orbital energies, 1 if the integrals are
 * unrestricted} */
int mcdfci_info(const mcdfci_ctx* ctx, int64_t* info);

/* HamiltonianMolecule::get_hartree_fock (src/hamiltonian.h:469-492). */
int mcdfci_hartree_fock(mcdfci_ctx* ctx, uint64_t* det);

/* Outer CDFCI::solve initialisation (src/solver.h:853-873): x = e_ref,
 * z = H[:,ref], candidates = top-M of that column.  ref_det NULL -> HF. */
int mcdfci_init(mcdfci_ctx* ctx, const uint64_t* ref_det);

/* Compression threshold used by the next mcdfci_init; the reference's
 * z_threshold_search restart doubles it (src/solver.h:882-888). */
int mcdfci_set_z_threshold(mcdfci_ctx* ctx, double z_threshold);

/* n_iter iterations of the inner loop (src/solver.h:701-759) on the device. */
int mcdfci_step(mcdfci_ctx* ctx, int64_t n_iter);

/* The report row (src/solver.h:761-798). Synchronises the device. */
int mcdfci_report_get(mcdfci_ctx* ctx, mcdfci_report* r);

/* Batching additions around mcdfci_step; no counterpart in the reference,
 * whose inner loop (src/solver.h:701-759) they execute.  Neither changes a
 * result bit.
 *
 * mcdfci_set_graph: iterations_per_graph > 0 makes mcdfci_step replay a HIP
 * graph holding that many captured iterations (captured on first use and
 * after mcdfci_init or mcdfci_autotune; the remainder of a step runs
 * uncaptured); 0 launches every kernel directly (the default).  Steps with
 * timing enabled are never captured.
 *
 * mcdfci_autotune: runs `rounds` rounds of real iterations (they advance the
 * solver like mcdfci_step), timing the scatter at 1, 2, 4, 8 and 16 column
 * elements per thread and the pick with one-workgroup and merge-tree sorts,
 * and keeps the fastest of each (median kernel time).  choice (may be NULL)
 * receives {elements per thread, 1 if merge-tree pick, iterations run}.
 *
 * mcdfci_set_launch: sets the same two launch forms directly (elements per
 * thread 1, 2, 4, 8 or 16; tree_pick 0 or 1, where 0 is refused when
 * num_of_coordinate * estimate exceeds one workgroup's sort of 4096).  The
 * forms never change results. */
int mcdfci_set_graph(mcdfci_ctx* ctx, int iterations_per_graph);
int mcdfci_autotune(mcdfci_ctx* ctx, int rounds, int64_t* choice);
int mcdfci_set_launch(mcdfci_ctx* ctx, int elements_per_thread, int tree_pick);

/* Checkpoint and resume (no counterpart in the reference, which restarts
 * from the Hartree-Fock determinant only, src/solver.h:853-873): save the
 * wavefunction table ({D: x_D, z_D} with z exact), the scalars, counters and
 * candidate lists to a file; load restores them into a context created for
 * the same Hamiltonian and num_of_coordinate / estimate / precision /
 * threshold options (MCDFCI_E_INPUT otherwise), in place of mcdfci_init.
 * Stepping a loaded context continues the saved run bit for bit.  The
 * loaded run keeps its own table size (max_memory). */
int mcdfci_checkpoint_save(mcdfci_ctx* ctx, const char* path);
int mcdfci_checkpoint_load(mcdfci_ctx* ctx, const char* path);

/* Measurement (no counterpart in the reference): the device time of n_ops
 * random-slot operations on this context's own table -- mode 0 a new-key
 * insert (key load, CAS, one 64-bit atomic add), 1 an update (key load, one
 * int64 atomic add: the scatter's exact z update for a contribution below
 * 2^-11), 2 a read of key, x and z, 3 an update by one FP64 atomic add (the
 * cost of a z kept as an order-dependent double), 4 an update by two int64
 * atomic adds (a large contribution) -- the access ceiling of the column
 * scatter and pass 2 on the same footprint.  Overwrites the table:
 * mcdfci_init must follow before mcdfci_step. */
int mcdfci_access_ceiling(mcdfci_ctx* ctx, int mode, int64_t n_ops, double* seconds);

/* Phase times in seconds (pick, submatrix, columns, eig, update_x, update_z),
 * the step_timer of src/solver.h:683,704-754; device phases are measured
 * with HIP events when timing is enabled (mcdfci_set_timing). */
int mcdfci_set_timing(mcdfci_ctx* ctx, int enable);
int mcdfci_timers(mcdfci_ctx* ctx, double* t6);
/* Per-kernel device times in seconds over the timed iterations:
 * {pick, submatrix, eig (+update_x), scatter, pass2, merge}. */
int mcdfci_kernel_times(mcdfci_ctx* ctx, double* t6, int64_t* iterations);

/* Test hooks. */
/* HamiltonianMolecule::get_column (src/hamiltonian.h:412-418): writes the
 * column of det (diagonal first, then the rest sorted by determinant). */
int mcdfci_column(mcdfci_ctx* ctx, const uint64_t* det, uint64_t* dets, double* vals, int64_t cap,
                  int64_t* n);
/* CDFCI::get_hamiltonian_submatrix (src/solver.h:637-670); H is k*k
 * row-major, entries with j >= i are filled (the triangle the update reads). */
int mcdfci_submatrix(mcdfci_ctx* ctx, int k, const uint64_t* dets, double* H);
/* CoordinateUpdateEig::operator() (src/solver.h:333-609) on given inputs.
 * s[6] = {xx.hi, xx.lo, scale.hi, scale.lo, xz.hi, xz.lo} in, xx/scale out. */
int mcdfci_eig_update(mcdfci_ctx* ctx, int k, const double* H, const double* x, const double* z,
                      double* s, double* dx);
/* mcdfci_eig_update for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, results
 * and errors as mcdfci_eig_update for k <= 256; k above 256 returns
 * MCDFCI_E_INPUT.  Calls no hipMalloc / hipFree: H_sub, x, z, dx and the
 * solver state live in a fixed 528 KiB device array that each GPU holds
 * until the process ends.  Calls that overlap on one GPU (any context)
 * race on that array. */
int mcdfci_eig_update_noalloc(mcdfci_ctx* ctx, int k, const double* H, const double* x, const double* z,
                              double* s, double* dx);
/* Contents of the wavefunction table, sorted by determinant
 * (WaveFunctionCuckoo::fetch_val over all entries, src/wavefunction.h:631-651). */
int mcdfci_table_dump(mcdfci_ctx* ctx, uint64_t* dets, double* x, double* z, int64_t cap, int64_t* n);
/* Order-independent digest of the whole table (the same entries as
 * mcdfci_table_dump, with z exact): out3 = {entries, sum of key hashes,
 * sum of (key, x, z) hashes}, sums modulo 2^64.  The entry hash is
 *   e = mix64(w0 + 0x9e3779b97f4a7c15); e = mix64(e ^ w1); key hash = e;
 *   e = mix64(e ^ bits(x), with 0 for x == 0); e = mix64(e ^ z_hi);
 *   entry hash = mix64(e ^ z_lo)
 * over the determinant words w0, w1 (w1 = 0 for one word) and the 128-bit
 * z in units of 2^-64; mix64 is the MurmurHash3 64-bit finaliser.  Compares
 * tables too large to dump (WaveFunctionCuckoo contents,
 * src/wavefunction.h:631-651). */
int mcdfci_table_digest(mcdfci_ctx* ctx, uint64_t* out3);
/* The candidate list passed to the next pick (sub_xz, src/wavefunction.h:567-589). */
int mcdfci_candidates(mcdfci_ctx* ctx, uint64_t* dets, double* x, double* z, int64_t cap, int64_t* n);

/* Adjoint of the energy.  The reference has no counterpart (it only
 * computes energies, src/solver.h:672-829); the Hamiltonian terms
 * differentiated are those of src/hamiltonian.h:257-410.  For the current
 * wavefunction x (normalised c = x/|x|) returns the gradient of the
 * variational energy E = c.Hc / c.c with respect to every unique FCIDUMP
 * parameter, at its canonical position:
 *   grad_h[p*norb + q], p >= q                 (d E / d h_pq, h_pq = h_qp)
 *   grad_eri[((p*norb + q)*norb + r)*norb + s]  (d E / d (pq|rs) for the
 *       8-fold orbit; p >= q, r >= s, pair(pq) >= pair(rs))
 * with all other entries 0, and energy = core + sum theta * dE/dtheta (E at
 * fixed c, exact by linearity).  At an eigenvector this is the
 * Hellmann-Feynman gradient; it equals the spin-summed 1-RDM and 2-RDM
 * folded onto the unique parameters.  grad_h has norb^2 entries and
 * grad_eri norb^4; energy may be NULL.  A context with unrestricted
 * integrals returns MCDFCI_E_UNSUPPORTED (its parameters are per spin). */
int mcdfci_energy_gradient(mcdfci_ctx* ctx, double* grad_h, double* grad_eri, double* energy);

/* The FP64 Rayleigh quotient x.Hx / x.x of the current wavefunction x, with
 * H built from the fp64 integrals whatever the context's precision (the
 * energy a mixed-precision run would report with exact H elements).  The
 * reference reports xz / xx only (src/wavefunction.h:195-198); this is the
 * measure of the mixed form's energy error.  ham NULL takes the context's
 * integrals; otherwise ham's h, (pq|rs) and core (same norb) are used for
 * this call, which evaluates the same x under other integrals (e.g. finite
 * differences of the energy at a fixed wavefunction).  Sums are exact fixed
 * point, so the value does not depend on thread order.  ham must be NULL
 * for a context with unrestricted integrals (MCDFCI_E_UNSUPPORTED). */
int mcdfci_energy_fp64(mcdfci_ctx* ctx, const mcdfci_ham* ham, double* energy);

#ifdef __cplusplus
}
#endif

#endif /* MCDFCI_H */
