/* blqh.h -- C ABI of the MI355X exact-diagonalisation engine for bilayer
 * quantum Hall systems on the torus (drop-in for the basis -> hamiltonian ->
 * lanczos_hamiltonian -> matrix path of qiongz/BLQH).
 *
 * Conventions
 *  - Every function returns BLQH_OK (0) or a negative error code;
 *    blqh_last_error() returns the calling thread's last message.
 *  - Handles are opaque and owned by the library; output buffers are owned by
 *    the caller.  Complex arrays are interleaved (re, im) doubles.
 *  - One process drives one GPU.  Each host thread queues its library work
 *    on a HIP stream of its own (non-blocking, default priority), so work
 *    issued from two threads can overlap on the GPU (a build on one, a
 *    solve on the other).  A call returns when its results are complete,
 *    except blqh_spmv_device, which leaves its work queued on the calling
 *    thread's stream.  A handle must not be used from two threads at once.
 *  - No exceptions cross this boundary.
 */
#ifndef BLQH_H
#define BLQH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLQH_OK 0
#define BLQH_EINVAL (-1)   /* invalid argument or unsupported sector */
#define BLQH_ENOMEM (-2)   /* device allocation failed */
#define BLQH_EDEVICE (-3)  /* HIP runtime error or no GPU */
#define BLQH_ERANGE (-4)   /* caller buffer too small / size out of range */
#define BLQH_EINTERNAL (-5) /* consistency check failed */

/* A (J,K) sector.  Reference: basis::init(nphi, nel, nel_up, J, K),
 * BLQH_without_spin/basis.cpp:141-149; spinful
 * basis::init(nphi, nel, nel_up, nel_su, J, K), BLQH_with_spin/basis.cpp:177-188;
 * monolayer basis::init(nphi, nel, J, K), Monolayer/basis.cpp:55-131 (its
 * translations rotate by nphi / gcd(nel, nphi) orbitals, so K < gcd(nel, nphi);
 * nel_up = nel_su = -1; the Hamiltonian takes only lx, ly and nLL). */
typedef struct blqh_sector {
    int32_t nspecies; /* 1: monolayer; 2: spinless bilayer; 4: spinful bilayer (upper-up, upper-down,
                         lower-up, lower-down) */
    int32_t nphi;     /* flux quanta (orbitals) per species */
    int32_t nel;      /* electrons */
    int32_t nel_up;   /* electrons in the upper layer; -1 = all layer densities */
    int32_t nel_su;   /* spin-up electrons (nspecies = 4); -1 = unconstrained */
    int32_t J;        /* total momentum mod nphi; -1 = no filter */
    int32_t K;        /* translation momentum; -1 = no translation reduction */
} blqh_sector;

/* Hamiltonian parameters.  Reference: lhamil::set_hamil(lx, ly, nphi, nLL, d,
 * Delta_SAS, Delta_V, theta_x, theta_y, nthread),
 * BLQH_without_spin/lanczos_hamiltonian.cpp:361-385; spinful adds Delta_Z
 * and theta_B, BLQH_with_spin/lanczos_hamiltonian.cpp:930-954. */
typedef struct blqh_params {
    double lx, ly;      /* torus lengths: ly = sqrt(2 pi nphi / gamma), lx = gamma ly */
    int32_t nLL;        /* Landau level index */
    double d;           /* layer distance d/l */
    double delta_sas;   /* tunnelling gap (applied when > 0) */
    double delta_v;     /* bias (spinless: applied when > 0) */
    double delta_z;     /* Zeeman energy (spinful) */
    double theta_b;     /* in-plane field phase (spinful) */
    double theta_x;     /* twist entering the interlayer table; as in the reference, which reads
                           that table at |s| (lanczos_hamiltonian.cpp:263), H is Hermitian
                           with theta_x != 0 only in K = 0 sectors */
    double theta_y;     /* twist on interlayer pair hops */
} blqh_params;

typedef struct blqh_basis blqh_basis;
typedef struct blqh_hamil blqh_hamil;

const char* blqh_last_error(void);

/* Select the GPU (device index).  Optional: every other entry point
 * initialises device 0 on first use. */
int blqh_init(int device);

/* ---- basis: basis::init + id[] / basis_C[]  (BLQH_without_spin/basis.cpp:61-120,
 *      basis.h:32-34; BLQH_with_spin/basis.cpp:74-141) ---- */
int blqh_basis_create(const blqh_sector* sector, blqh_basis** out);
int64_t blqh_basis_size(const blqh_basis* b);          /* nbasis; < 0 on error */
int blqh_basis_export(const blqh_basis* b, uint64_t* id, int16_t* period);
void blqh_basis_destroy(blqh_basis* b);

/* ---- Hamiltonian build into CSR: lhamil::set_hamil / peer_set_hamil
 *      (BLQH_without_spin/lanczos_hamiltonian.cpp:87-385; matrix.h:71-88) ---- */
int blqh_hamil_build(const blqh_basis* b, const blqh_params* p, blqh_hamil** out);
/* Parts of the same Hamiltonian for parameter sweeps (no counterpart in the
 * reference, whose sweeps rebuild H at every point, BLQH_without_spin/main.cpp:76-165).
 * terms = 3: as blqh_hamil_build; 1: without the interlayer tunnelling
 * (Delta_SAS taken as 0); 2: the tunnelling alone at Delta_SAS = 1 (the
 * Delta_SAS of p is ignored; the theta_B phases use p's d).
 * blqh_hamil_combine(base, tun, Delta_SAS): H(Delta_SAS) = base + Delta_SAS tun
 * for a base without tunnelling (terms 1, or terms 3 at Delta_SAS = 0) and a
 * tunnelling-only build of the same basis and parameters (theta_B, theta_x,
 * theta_y, lx, ly, nLL, and d when theta_B != 0: the tunnelling depends on d
 * only through the theta_B phases): the tunnelling moves one electron between the layers,
 * so its entries never meet the layer-conserving ones and the result is a
 * merge of sorted rows with the build's 1e-10 drop.  It equals
 * blqh_hamil_build with that Delta_SAS: same rows, columns and Coulomb and
 * diagonal values; tunnelling values to rounding (Delta_SAS * sum instead of
 * the sum of Delta_SAS * terms).  The result owns its arrays; base and tun may
 * be destroyed afterwards.  A terms-1 or terms-2 part gets its SpMV formats
 * (internal row order, panels) on its first SpMV, solve or batch, or from
 * blqh_hamil_prepare; a part that only feeds a merge never builds them. */
int blqh_hamil_build_terms(const blqh_basis* b, const blqh_params* p, int32_t terms, blqh_hamil** out);
int blqh_hamil_combine(const blqh_hamil* base, const blqh_hamil* tun, double delta_sas, blqh_hamil** out);
/* Build the SpMV formats of h now (no-op when present; no reference
 * counterpart): lets a caller put that cost on a build thread. */
int blqh_hamil_prepare(blqh_hamil* h);
int blqh_hamil_info(const blqh_hamil* h, int64_t* nrows, int64_t* nnz);
/* rowptr[nrows+1], col[nnz], val[2*nnz]; any pointer may be NULL. */
int blqh_hamil_export(const blqh_hamil* h, int64_t* rowptr, int32_t* col, double* val);
void blqh_hamil_destroy(blqh_hamil* h);

/* ---- y = H x: Mat::operator*  (BLQH_without_spin/matrix.cpp:244-267) ---- */
int blqh_spmv(const blqh_hamil* h, const double* x, double* y);            /* host arrays, 2*nrows doubles */
int blqh_spmv_device(const blqh_hamil* h, const void* x_dev, void* y_dev); /* device arrays, library stream */

/* SpMV variant used for this matrix (engine tuning; no counterpart in the
 * reference).  kind 0: row-major CSR with `lanes` lanes per row; kind 1:
 * panel-major (65536-column panels, x slice resident in L2) with `lanes`
 * lanes per row segment, `threads` per workgroup (512 or 1024) and
 * `wgs_per_cu` persistent workgroups per CU (1 or 2) and `rows` rows per lane
 * group in flight (2 or 4); kind 2: panel-major slice streaming (a wavefront
 * owns 64-row slices spread over the matrix, lanes = 64) with `threads` 512
 * or 1024, `wgs_per_cu` 1-4 and `rows` = slices per wavefront and round (1, 2,
 * 4 or 8); kind 3: the same slices over panels stored step-major (one lane per
 * row, coalesced, lanes = 64) with `threads` 256, 512 or 1024, `wgs_per_cu`
 * 1-4 and `rows` 1, 2, 4 or 8.  Switching between kind 3 and kinds 1-2
 * reorders the panel entries on the device (temporary copy of the entries).
 * kind 4: column tiles of 8192 staged in LDS, one workgroup of 1024 threads
 * per CU owning a row block (a separate copy of the matrix, built on first
 * use; `lanes`, `threads`, `wgs_per_cu` and `rows` are set by the layout and
 * reported by get; BLQH_EINVAL when a row has more than 255 entries in one
 * tile or the matrix has more than 4.2M rows).
 * kind 5: step-major panels with `lanes` = R rows per lane (2, 4 or 8): lane
 * l sums rows l, l+64, .., l+64(R-1) of a 64R-row block in one stream per
 * panel, which evens out the lanes' work per step; `threads` 512 or 1024,
 * `wgs_per_cu` 1-4, `rows` = blocks per wavefront and round with
 * R * rows <= 8.  Kind 3 orders each step greedily (lanes that read the same
 * x cache line share a step), kind 5 keeps each row's entries in order.
 * get: kind = -1 until
 * the autotuner has run (first SpMV or solve).  set: pins a variant (returns
 * BLQH_EINVAL for an unsupported one); kind = -1 re-arms the autotuner. */
int blqh_hamil_get_spmv(const blqh_hamil* h, int32_t* kind, int32_t* lanes, int32_t* threads, int32_t* wgs_per_cu,
                        int32_t* rows);
int blqh_hamil_set_spmv(blqh_hamil* h, int32_t kind, int32_t lanes, int32_t threads, int32_t wgs_per_cu,
                        int32_t rows);

/* ---- Lanczos: coeff_explicit_update + diag + eigenstates_reconstruction +
 *      ground_state_energy (BLQH_without_spin/lanczos_hamiltonian.cpp:453-656) ---- */
typedef struct blqh_lanczos_opts {
    int32_t lambda;       /* step cap (-m, lanczos_hamiltonian.h:15) */
    uint32_t seed;        /* std::mt19937 seed of the start vector */
    double tol;           /* stop when the lowest Ritz value moves by < tol; <= 0: run lambda steps */
    int32_t check_every;  /* iterations between convergence checks (>= 1) */
    int32_t store_basis;  /* 1: keep the Krylov vectors in HBM (two passes); 0: replay (three passes); -1: auto */
    const double* start;  /* optional start vector (2*nrows doubles), normalised here; NULL = seeded mt19937 */
    double res_tol;       /* also require the Ritz residual beta_{m} |s_m| < res_tol (bounds the state's
                             error; <= 0: energy criterion only) */
    int32_t skip_states;  /* 1: stop after passes 1 and 2 (alpha, beta, E0; psi untouched, E_rq = NaN),
                             for callers that reconstruct states with blqh_lanczos_reconstruct;
                             0: reconstruct the ground state.  Ignored by blqh_batch_lanczos. */
    double start_mix;     /* with a caller's start vector: start/|start| + start_mix * r, r the seeded
                             mt19937 unit vector, then normalised (0: the start as given).  A warm start
                             orthogonal to the ground state (the previous parameter point's ground state
                             across a level crossing between symmetry classes of the sector) otherwise
                             converges to an excited level; with start_mix = 0.1 the ground-state
                             component starts near 0.1/sqrt(N), far above the stopping tolerance. */
} blqh_lanczos_opts;

typedef struct blqh_lanczos_result {
    double E0;            /* lowest Ritz value (eigenvalues[0]) */
    double E_rq;          /* Re<psi|H psi> of the reconstructed state (ground_state_energy) */
    double dE;            /* last change of the lowest Ritz value */
    int32_t iterations;   /* Lanczos vectors generated (tridiagonal order) */
    int32_t converged;    /* 1 when |dE| < tol was reached */
    int32_t passes;       /* 2 (stored basis) or 3 (replay) */
    double seconds;       /* wall time of the solve */
    double residual;      /* Ritz residual estimate beta_{m} |s_m| at the stop */
} blqh_lanczos_result;

/* psi (2*nrows, may be NULL): reconstructed ground state, psi = sum_l c_l conj(v_l),
 * normalised, as the reference's psir_0.  alpha/beta (may be NULL): lambda and
 * lambda+1 doubles (beta[0] = 1 as in the reference). */
int blqh_lanczos(const blqh_hamil* h, const blqh_lanczos_opts* o, blqh_lanczos_result* r,
                 double* psi, double* alpha, double* beta);

/* blqh_lanczos without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops that forbid hipMalloc).  Same arguments, result
 * and output as blqh_lanczos with store_basis = 0 (three passes, pass 3 a
 * bit-identical replay); every device array of the solve lives in one static
 * array per GPU of about 12.3 MiB, held until the process ends.  Limits
 * (BLQH_ERANGE beyond them): at most 131072 rows and lambda <= 1024;
 * start_mix must be 0 (BLQH_EINVAL).  The SpMV is the tuned variant when the
 * Hamiltonian's layout is already in place for it (after a blqh_lanczos or
 * blqh_spmv on it), else the CSR kernel.  Calls in one process are serialised
 * (they share the array, whose contents each call re-initialises). */
int blqh_lanczos_noalloc(const blqh_hamil* h, const blqh_lanczos_opts* o, blqh_lanczos_result* r, double* psi,
                         double* alpha, double* beta);

/* Pass 3 on its own: eigenstates_reconstruction of both reference trees
 * (BLQH_without_spin/lanczos_hamiltonian.cpp:610-644; the spinful one,
 * BLQH_with_spin/lanczos_hamiltonian.cpp:1143-1191, builds psir_0 and psir_1
 * in one replay).  The Lanczos vectors v_0..v_{m-1} are regenerated from the
 * start vector (seeded mt19937 when start is NULL, else start, normalised; a
 * pass 1 run with start_mix > 0 replays from the mixed vector, so pass that),
 * and for each of the ns coefficient vectors coefs[s*m .. s*m+m-1] the state
 * psi_s = sum_l coefs[s*m+l] conj(v_l) is formed, normalised and written to
 * psi[s*2*nrows ..] (complex, 2*nrows doubles each; psi may be NULL), with
 * erq[s] = Re<psi_s|H psi_s> (ground_state_energy / first_excited_state_energy,
 * :1193-1215).  The replay recomputes alpha and beta with the engine's own
 * recurrence rather than reusing the stored values; *max_dev (may be NULL)
 * receives the largest relative difference between the recomputed and the
 * given alpha[0..m-2] / beta[1..m-1] (either array may be NULL), which is 0
 * for coefficients from blqh_lanczos with the same start and about 1e-8 for a
 * record read back from its 8-digit text (resume of pass 3 from a record,
 * read_from_file, :768-819).  1 <= ns <= 16. */
int blqh_lanczos_reconstruct(const blqh_hamil* h, int32_t m, uint32_t seed, const double* start,
                             const double* alpha, const double* beta, int32_t ns, const double* coefs, double* psi,
                             double* erq, double* max_dev);

/* ---- Batched solves: the sweeps over sectors and parameter points of
 *      BLQH_without_spin/main.cpp:76-165 and BLQH_with_spin/main.cpp:158-302,
 *      each of which runs the three passes above on one Hamiltonian at a time.
 * A batch groups resident Hamiltonians of any sizes (they stay owned by the
 * caller and must outlive the batch).  blqh_batch_lanczos advances every
 * problem through the same launches (steps captured in a hipGraph, replayed
 * every `check_every` steps), stops each problem by the rule of blqh_lanczos,
 * keeps all Krylov vectors in HBM ((lambda+1) * nrows * 16 bytes per problem;
 * BLQH_ENOMEM when they do not fit) and reconstructs each ground state.
 * seeds: per-problem mt19937 seeds (NULL: o->seed for all); starts: per-problem
 * start vectors (NULL array or NULL entry: seeded); psi: per-problem outputs
 * (2*nrows doubles; NULL array or entry: not copied); res[count]. ---- */
typedef struct blqh_batch blqh_batch;
/* lanes: 0 = autotune the batched SpMV on this batch, else as blqh_batch_set_lanes
 * (a sweep tunes its first batch and passes the choice to the next ones). */
int blqh_batch_create(blqh_hamil* const* hs, int32_t count, int32_t lanes, blqh_batch** out);
/* Replace the members by Hamiltonians of the same sizes in the same order (the
 * next point of a parameter sweep); device pools, the captured replay and the
 * lane choice are kept. */
int blqh_batch_rebind(blqh_batch* b, blqh_hamil* const* hs, int32_t count);
int blqh_batch_lanczos(blqh_batch* b, const blqh_lanczos_opts* o, const uint32_t* seeds,
                       const double* const* starts, blqh_lanczos_result* res, double* const* psi);
/* Continuation solve (no counterpart in the reference, which starts every
 * point of its sweeps from a seeded vector; for parameter sweeps whose
 * callers would otherwise copy each ground state out and back in as the next
 * start).  Every problem starts from conj(psi) of the previous finished solve
 * on this batch (kept on the device in the reference row order, across
 * blqh_batch_rebind / blqh_batch_rebind_scaled) plus o->start_mix times its
 * seeded unit vector, normalised: with excited_mix = 0 the start
 * blqh_batch_lanczos builds from starts[p] = conj(previous psi[p]).
 * excited_mix > 0 also adds that weight times conj of the previous solve's
 * excited mix (the normalised sum of its second and third lowest Ritz
 * vectors, orthogonal to its ground Ritz vector), against level crossings
 * between the points.  Needs the same o->lambda as that solve (BLQH_EINVAL
 * otherwise, and on a batch that has not finished a solve). */
int blqh_batch_lanczos_continue(blqh_batch* b, const blqh_lanczos_opts* o, const uint32_t* seeds,
                                double excited_mix, blqh_lanczos_result* res, double* const* psi);
/* Lanes per row of the batched SpMV (4, 8, 16 or 32 over row-major panels;
 * 64: one lane per row over step-major panels; autotuned at create). */
int blqh_batch_get_lanes(const blqh_batch* b, int32_t* lanes);
int blqh_batch_set_lanes(blqh_batch* b, int32_t lanes);
/* Shared-scale batch: the Delta_SAS points of one d, the inner loop of
 * BLQH_without_spin/main.cpp:133-165.  For nsector sectors and nscale
 * (1..4) values s_g, problem p = q * nscale + g is H0[q] + s_g T[q], with
 * base[q] a build without tunnelling (blqh_hamil_build_terms terms = 1) and
 * tun[q] the unit tunnelling matrix of the same sector (terms = 2), as for
 * blqh_hamil_combine; sectors of at most 32768 states.  The batch keeps its
 * own panel copy of H0 + T (base and tun may be destroyed afterwards) and
 * reads each sector's entries once per Lanczos step for all of its problems.
 * Results follow the problem order.  The entries that blqh_hamil_combine
 * drops (|s_g t| <= 1e-10) contribute their values here. */
int blqh_batch_create_scaled(const blqh_hamil* const* base, const blqh_hamil* const* tun, int32_t nsector,
                             const double* scales, int32_t nscale, blqh_batch** out);
/* The next d of a shared-scale batch: H0 and T of the same sectors, same order. */
int blqh_batch_rebind_scaled(blqh_batch* b, const blqh_hamil* const* base, const blqh_hamil* const* tun,
                             int32_t nsector);
void blqh_batch_destroy(blqh_batch* b);

/* ---- Learned warm start (no counterpart in the reference: it supplies the
 *      start vector that lanczos_hamiltonian.cpp:466-483 draws at random).
 * An FP32 MLP over the nbits occupation bits of a basis word (as +-1) plus
 * (d/3, 10 Delta_SAS): two hidden layers of 64 tanh units, outputs log|psi|
 * and the logit of sign(psi) > 0.  Trained on the GPU on ground states the
 * engine computed; its output is passed as blqh_lanczos_opts.start (or the
 * batch's starts).  Parameters, flat FP32: W1[nbits+2][64], b1[64],
 * W2[64][64], b2[64], W3[64][2], b3[2]. ---- */
typedef struct blqh_mlp blqh_mlp;
typedef struct blqh_mlp_train_opts {
    int32_t steps;    /* optimiser steps (Adam) */
    int32_t batch;    /* states per step, drawn with probability |psi|^2 from a uniformly chosen set */
    float lr;         /* learning rate of the first step */
    float lr_final;   /* learning rate of the last step (geometric decay); <= 0: constant */
    uint32_t seed;    /* sampling seed */
} blqh_mlp_train_opts;
typedef struct blqh_mlp_train_result {
    double loss_first;  /* minibatch loss before the first step */
    double loss_last;   /* mean minibatch loss of the last (up to) 20 steps */
    int64_t samples;    /* states in the training sets */
    int32_t steps;
    double seconds;
} blqh_mlp_train_result;
int blqh_mlp_create(int32_t nbits, uint32_t seed, blqh_mlp** out);  /* Glorot init from mt19937(seed) */
int blqh_mlp_nparams(const blqh_mlp* m);
int blqh_mlp_get_params(const blqh_mlp* m, float* out);
int blqh_mlp_set_params(blqh_mlp* m, const float* in);             /* also resets the optimiser state */
/* Mean loss over explicit states and its gradient (grad may be NULL):
 * L = mean_s (o0_s - ya_s)^2 + softplus(o1_s) - ys_s o1_s. */
int blqh_mlp_loss_grad(blqh_mlp* m, int64_t count, const uint64_t* words, const float* d, const float* dsas,
                       const float* ya, const float* ys, double* loss, float* grad);
/* Train on `nsets` ground states: set k = (bases[k], d = params[2k],
 * Delta_SAS = params[2k+1], psi[k] = 2*nrows doubles as returned by
 * blqh_lanczos); each state is gauge-fixed to real first. */
int blqh_mlp_train(blqh_mlp* m, int32_t nsets, const blqh_basis* const* bases, const double* params,
                   const double* const* psi, const blqh_mlp_train_opts* o, blqh_mlp_train_result* r);
/* Normalised real start vector (2*nrows doubles) for a basis at (d, Delta_SAS). */
int blqh_mlp_start_vector(const blqh_mlp* m, const blqh_basis* b, double d, double dsas, double* out);
void blqh_mlp_destroy(blqh_mlp* m);

/* Dense cross-check: hamil::diag + ground_state_energy
 * (BLQH_without_spin/hamiltonian.cpp:341-353,413-432).  Diagonalises the dense
 * matrix with rocSOLVER zheevd in the reference's convention (row-major H read
 * column-major, i.e. H^T); e0 = lowest eigenvalue, e_printed = Re<H v|v> for
 * the first eigenvector v (the value the reference prints), vec (2*nrows, may
 * be NULL) = v.  Limited to 16384 states. */
int blqh_dense_ground_state(const blqh_hamil* h, double* e0, double* e_printed, double* vec);

/* Start vector of the reference: std::mt19937(seed), entry i =
 * (draw_{2i+1}, draw_{2i}) / max - 0.5, normalised
 * (lanczos_hamiltonian.cpp:466-483, matrix.cpp:54-75). */
int blqh_start_vector(int64_t n, uint32_t seed, double* out);

/* Lowest eigenpair of the symmetric tridiagonal matrix (alpha[m], beta[m-1]);
 * the reference's diag() pass 2 (lanczos_hamiltonian.cpp:581-606). */
int blqh_tridiag_lowest(int32_t m, const double* alpha, const double* offdiag, double* e0, double* vec);

/* ---- Ground-state observables: lhamil::occupatation_number, pseudospin_Sz,
 *      pseudospin_Sx (BLQH_without_spin/lanczos_hamiltonian.cpp:658-735) and
 *      the spinful Sz, upper_Sz, down_Sz (BLQH_with_spin/lanczos_hamiltonian.cpp:1217-1500),
 *      of a state psi (2*nrows doubles, e.g. from blqh_lanczos) in basis b. ---- */
typedef struct blqh_observables {
    double pseudospin_sz;       /* (2 <N_upper> - Ne) / (2 Ne) */
    double pseudospin_sx;       /* |<psi|T|psi>| / Ne, T the interlayer tunnelling with 1/2 per pair */
    double sz;                  /* spinful: (2 <N_up> - Ne) / (2 Ne); spinless: 0 */
    double upper_sz, lower_sz;  /* spinful upper_Sz / down_Sz; spinless: 0 */
    double occupation[2][64];   /* [a][j]: orbital j of layer a (spinful: species 0 / 1) */
    double spinflip;            /* spinful spinflip_tunneling (BLQH_with_spin/lanczos_hamiltonian.cpp:1519-1592):
                                   |<psi|T_sf|psi>| / Ne, T_sf the spin-flip interlayer hops with 1/2
                                   per pair (upper-up <-> lower-down, upper-down <-> lower-up); spinless: 0 */
} blqh_observables;
int blqh_compute_observables(const blqh_basis* b, const double* psi, blqh_observables* out);

/* ---- Spectra and the Lanczos record ----
 * All eigenpairs of the Lanczos tridiagonal (alpha[m], off-diagonal beta_1..beta_{m-1}):
 * the reference's diag() (BLQH_without_spin/lanczos_hamiltonian.cpp:581-606).
 * eig[m] ascending; first[m] = first component of every eigenvector (psi_n0);
 * vec0[m] = eigenvector of eig[0] (psi_0, first component >= 0); any may be NULL. */
int blqh_tridiag_spectrum(int32_t m, const double* alpha, const double* offdiag, double* eig, double* first,
                          double* vec0);
/* Eigenvectors k0 .. k0+count-1 (ascending eigenvalues) of the same tridiagonal,
 * each with first component >= 0, into vecs[(k - k0)*m ..]: the Krylov
 * coefficients psi_0 / psi_1 of diag() (BLQH_with_spin/lanczos_hamiltonian.cpp:1106-1139). */
int blqh_tridiag_vectors(int32_t m, const double* alpha, const double* offdiag, int32_t k0, int32_t count,
                         double* vecs);
/* Low-lying distinct levels: eig[0], then every eig[i] with eig[i] - eig[i-1] > gap,
 * at most kmax (the ghost filter of BLQH_without_spin/main.cpp:100-110, gap 1e-5).
 * Returns the count (>= 0) or an error code. */
int blqh_spectrum_distinct(int32_t m, const double* eig, int32_t kmax, double gap, double* out);
/* The record of lhamil::save_to_file / read_from_file
 * (BLQH_without_spin/lanczos_hamiltonian.cpp:738-819): nHilbert, lambda (= m),
 * seed, eigenvalues[m], norm = beta[m+1] (beta[0] = 1), overlap = alpha[m],
 * psi_0[m], psi_n0[m], in the reference's text format.  load: arrays of
 * capacity cap (beta cap+1); BLQH_ERANGE when the record is longer. */
int blqh_record_save(const char* path, int64_t nhilbert, int32_t m, uint32_t seed, const double* eig,
                     const double* beta, const double* alpha, const double* psi0, const double* psin0);
int blqh_record_load(const char* path, int32_t cap, int64_t* nhilbert, int32_t* m, uint32_t* seed, double* eig,
                     double* beta, double* alpha, double* psi0, double* psin0);

/* Kernel timing (engine instrumentation; no counterpart in the reference).
 * When enabled, every fused Lanczos SpMV launch (the solver's dominant kernel)
 * is bracketed by HIP events on the library stream; blqh_profile_read returns
 * the launch count, the summed kernel time and the algorithmic bytes summed
 * over those launches (20 nnz + 56 n + 8 per launch: CSR values, columns,
 * row pointers, x read once, previous vector read, output written). */
/* Plain device-to-device copy kernel of `bytes` bytes (multiple of 16) on the
 * library stream: the reference point for scaling hardware byte counters. */
int blqh_device_copy(const void* src_dev, void* dst_dev, int64_t bytes);
/* Device allocation helpers for callers without a HIP runtime of their own. */
int blqh_device_alloc(void** ptr, int64_t bytes);
int blqh_device_free(void* ptr);
/* Wait for the work the calling thread has queued (its library stream). */
int blqh_synchronize(void);
/* Release the device blocks the library keeps for reuse.  Engine memory
 * management, no reference counterpart: blocks of 1 MiB or more freed by the
 * library (matrices, panels, Krylov blocks of destroyed handles) are kept,
 * up to a quarter of device memory, and reused for allocations of the same
 * size; they are also released by the library itself when an allocation
 * fails.  Callers sharing the device with other allocators call this to
 * return the memory.  *bytes (may be NULL) receives the amount released. */
int blqh_release_cached(int64_t* bytes);
int blqh_profile_enable(int on);
int blqh_profile_read(int64_t* launches, double* kernel_ms, double* bytes);
int blqh_profile_reset(void);

#ifdef __cplusplus
}
#endif
#endif /* BLQH_H */
