/*
 * aimdhf -- C ABI of the MI355X (gfx950) Hartree-Fock energy-and-forces engine.
 *
 * The reference (Nathan-W123/Ab-Initio-MD, package aimd) has no native ABI:
 * its hot path is numba-jitted Python. Each entry point below replaces the
 * reference function cited next to it; the Python host layer
 * (aimd_amd.qc.engine) binds these through ctypes and mirrors the
 * reference's Python API on top of them.
 *
 * Conventions
 *   - Return value: 0 on success, otherwise one of the AIMDHF_E_* codes; the
 *     message is available from aimdhf_last_error(). Nothing throws across
 *     the ABI.
 *   - Host pointers carry basis data and positions; "device" pointers are
 *     HBM buffers owned by the caller (PyTorch tensors), row-major doubles.
 *   - Every computation is enqueued on the caller's hipStream_t (passed as
 *     void*; NULL = the default stream) and is asynchronous unless noted.
 *   - One plan is used by one host thread at a time; plans are independent.
 *   - Units: bohr, hartree. AOs: unit-normalised Cartesian functions in
 *     PySCF order, shells of an atom sorted by l (the reference's layout,
 *     aimd/qc/basis.py:17-68).
 */
#ifndef AIMDHF_H
#define AIMDHF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AIMDHF_OK 0
#define AIMDHF_E_ARG 1          /* invalid argument (Python: ValueError)        */
#define AIMDHF_E_NOMEM 2        /* allocation failed (Python: MemoryError)      */
#define AIMDHF_E_HIP 3          /* HIP runtime error (Python: RuntimeError)     */
#define AIMDHF_E_UNSUPPORTED 4  /* e.g. l > 2 in this build (Python: ValueError) */
#define AIMDHF_E_STATE 5        /* call order (e.g. no geometry set yet)        */

typedef struct aimdhf_plan aimdhf_plan;

/* Copy the message of the last failed call on this thread into buf. */
int aimdhf_last_error(char* buf, size_t len);

/* Highest angular momentum this build's kernels are instantiated for. */
int aimdhf_lmax(void);

/*
 * Stage a basis (flat shell arrays of aimd/qc/basis.py:140-183 plus the
 * contraction-block index of every shell, aimd/qc/basis.py:185-228) and the
 * Boys table (aimd/qc/boys.py:57-75) on the current device.
 * Replaces: BasisSet / build_basis (aimd/qc/basis.py:140-312) and the
 * geometry-independent part of _Pairs (aimd/qc/integrals.py:602-621).
 * shell_prim has nshell+1 entries; charges are the nuclear charges Z_A.
 */
int aimdhf_plan_create(int natm, int nshell, const int64_t* shell_atom, const int64_t* shell_l,
                       const int64_t* shell_prim, const int64_t* shell_block, const double* prim_exp,
                       const double* prim_coef, const double* charges, aimdhf_plan** out);

int aimdhf_plan_destroy(aimdhf_plan* plan);

/*
 * A batched plan: nrep independent molecules (replicas) of any sizes in one plan,
 * their shells and atoms concatenated replica by replica (rep_atom / rep_shell:
 * nrep + 1 offsets; every shell on an atom of its own replica), the other arrays
 * as in aimdhf_plan_create. Shell pairs and quartets form only within a replica
 * and each replica's nuclei act only on its own electrons, so every replica is
 * the molecule it would be alone. The entry points then work on all replicas at
 * once: set_geometry takes the concatenated positions and evaluates every
 * screened quartet of every replica (one launch per quartet class) into one dense
 * two-electron tensor per replica; jk digests all tensors in one launch;
 * one_electron, grad_one(_terms) and grad_two take and fill matrices in the
 * padded layout: nset (or 1) matrices of ld x ld per replica, replica by replica
 * ((nrep, nset, ld, ld); ld = the largest replica's AO count, see
 * aimdhf_batch_layout), rows and columns past a replica's AO count zero; the
 * gradient stays (natm, 3) over all atoms. Replaces the per-molecule backend
 * calls that the reference's MD loop makes one replica at a time
 * (aimd/integrators.py:419-491 via aimd/backends/hf.py:293-363).
 */
int aimdhf_batch_create(int nrep, const int64_t* rep_atom, const int64_t* rep_shell, int natm, int nshell,
                        const int64_t* shell_atom, const int64_t* shell_l, const int64_t* shell_prim,
                        const int64_t* shell_block, const double* prim_exp, const double* prim_coef,
                        const double* charges, aimdhf_plan** out);

/* Replicas (0 for a single-molecule plan), padded matrix dimension ld (nao for a
 * single molecule) and bytes of the dense tensors of a batch (any pointer may be
 * NULL). No counterpart in the reference. */
int aimdhf_batch_layout(aimdhf_plan* plan, int* nrep, int* ld, int64_t* tensor_bytes);

/*
 * Move the basis to positions (host, (natm,3) bohr) and build everything that
 * depends on the geometry: primitive-pair records (mu |A-B|^2 <= 50 kept),
 * Schwarz bounds Q = max |(ab|ab)|^(1/2), and the class-binned list of unique
 * shell quartets whose contraction-block pairs satisfy Q_P Q_Q >= threshold.
 * In store mode (aimdhf_set_jk_mode) the screened quartets are also evaluated
 * here into a device-resident integral store, with the gather lists of the
 * J and K kernels. Synchronises the stream (the list sizes come back to the host).
 * Replaces: BasisSet.with_positions (aimd/qc/basis.py:242-246), _build_pairs
 * (aimd/qc/integrals.py:525-599), schwarz_bounds (aimd/qc/integrals.py:970-983)
 * and the screening loop of _eri_tensor_kernel (aimd/qc/integrals.py:900-903);
 * in store mode also eri_tensor (aimd/qc/integrals.py:986-1006), holding only
 * the screened unique blocks (8 bytes per integral) instead of nao^4.
 */
int aimdhf_set_geometry(aimdhf_plan* plan, const double* positions, double schwarz_threshold, void* stream);

/* Schwarz bound of every contraction-block pair, index i(i+1)/2+j (i >= j),
 * copied to host memory (npair_block doubles). Synchronous.
 * Replaces: schwarz_bounds (aimd/qc/integrals.py:970-983). */
int aimdhf_schwarz_bounds(aimdhf_plan* plan, double* q_block_host, int64_t cap);

/* Number of surviving unique shell quartets of the current geometry. */
int aimdhf_quartet_count(aimdhf_plan* plan, int64_t* count);

/* The surviving quartets as (P, Q) shell-pair indices with P >= Q, copied to
 * host memory (2*cap int64), grouped by class. Synchronous. */
int aimdhf_quartets(aimdhf_plan* plan, int64_t* out_host, int64_t cap, int64_t* count);

/*
 * Coulomb and exchange matrices of nset densities (device (nset,nao,nao)):
 *   J_mn = sum_ls (mn|ls) D_ls,   K_mn = sum_ls (ml|ns) D_ls
 * of the symmetric part of each D over the screened quartets. Every output
 * block is owned by one wavefront that sums its contributions in a fixed
 * order (no atomics), so J and K are bitwise reproducible run to run.
 * Direct mode recomputes the integrals (J-engine in Hermite space for J,
 * per-block exchange owners for K); store mode reads them from the store.
 * Replaces: jk_from_eri (aimd/qc/integrals.py:1009-1044) and, in direct mode,
 * eri_tensor (aimd/qc/integrals.py:986-1006).
 */
int aimdhf_jk(aimdhf_plan* plan, int nset, const double* dens, double* J, double* K, void* stream);

/*
 * J and K of density differences dD (device (nset,nao,nao)) for an
 * incremental Fock build, J(D) = J(D_prev) + J(dD). In store mode a work unit
 * whose largest Q_P Q_Q times max |dD| is below threshold is skipped (each of
 * its integral-density products is below threshold), and in mixed precision
 * every stored integral is digested in FP32, the large-bound ones from FP32
 * copies made at set_geometry; threshold 0 skips nothing. In direct mode the
 * same as aimdhf_jk. Replaces the full jk_from_eri rebuild of every SCF
 * iteration (aimd/qc/scf.py:863-867, aimd/qc/integrals.py:1009-1044).
 */
int aimdhf_jk_delta(aimdhf_plan* plan, int nset, const double* ddens, double* dJ, double* dK, double threshold,
                    void* stream);

/*
 * J/K mode for the next aimdhf_set_geometry: 0 auto (store when its estimated
 * size is within max_store_bytes, 0 = half the free device memory), 1 direct,
 * 2 store. No counterpart in the reference, which always stores the full
 * tensor and refuses nao > max_nao (aimd/qc/integrals.py:986-1006).
 */
int aimdhf_set_jk_mode(aimdhf_plan* plan, int mode, int64_t max_store_bytes);

/*
 * Arithmetic precision. mode 0: FP64 everywhere (default). mode 1 (mixed):
 * derivative integrals of quartets with Q_P Q_Q max|Gamma| < grad_switch are
 * evaluated in FP32 (accumulated per quartet in FP32, summed in exact fixed
 * point), and, in store mode, integrals of quartets with Q_P Q_Q < jk_switch
 * are held and digested in FP32 (taking effect at the next set_geometry);
 * aimdhf_jk_delta digests every stored integral in FP32 once max |dD| <=
 * delta_switch. Large-bound quartets, the SCF algebra and all sums stay FP64.
 * No counterpart in the reference (FP64 only).
 */
int aimdhf_set_precision(aimdhf_plan* plan, int mode, double jk_switch, double grad_switch, double delta_switch);

/*
 * One trace-correcting purification step on device matrices (n x n, row-major):
 * Xout = X2 if tr X > nocc, else 2 X - X2, where X2 = X X (computed by the
 * caller); with freeze > 0, Xout = X when |tr X - tr X2| < freeze (a converged
 * iterate is held: further steps would amplify its rounding errors). Needs no
 * plan; issues no memory clearing, so it may be captured in a graph. Building
 * block of the SCF density from the orthogonal-basis Fock matrix, which
 * replaces the eigenvector aufbau of aimd/qc/scf.py:913-921.
 */
int aimdhf_tc2_step(int n, const double* X, const double* X2, double* Xout, double nocc, double freeze,
                    void* stream);

/*
 * Start of the purification: X0 = (e_max - F) / (e_max - e_min) for the
 * symmetric device matrix F (n x n), e_min / e_max its Gershgorin bounds
 * (the start of aimd/qc/scf.py's aufbau replaced by purification, see
 * aimdhf_tc2_step). Two launches (row bounds, then the fill); the row bounds
 * go through a file-scope device array, so calls on one device must not
 * overlap. No plan.
 */
int aimdhf_tc2_init(int n, const double* F, double* X0, void* stream);

/*
 * Aufbau density projectors of nmat symmetric ld x ld matrices at once (ld <= 64;
 * device F (nmat, ld, ld), nocc (nmat) doubles): trace-correcting purification
 * from the Gershgorin start, each matrix in one workgroup's LDS for all steps;
 * `steps` held steps (an iterate with idempotency error below tol is held), then
 * two unheld steps; X^2 on the matrix cores (FP64 MFMA). P (nmat, ld, ld)
 * receives the symmetrised projectors, err (nmat) the idempotency error before
 * the last two steps. No plan; no memory
 * clearing, so a graph may capture it. The batched counterpart of the aufbau of
 * aimd/qc/scf.py:913-921 (there: eigenvectors of one Fock matrix at a time).
 */
int aimdhf_batch_purify(int nmat, int ld, const double* F, const double* nocc, double* P, int steps, double tol,
                        double* err, void* stream);

/*
 * Lanes (16, 32 or 64; default 64) that digest one output element of a
 * batched plan's J/K, for batches of small replicas where a full wavefront per
 * element leaves lanes idle. A tuning knob with no counterpart in the
 * reference; aimd_amd.batch.ReplicaBatch.autotune times the choices.
 */
int aimdhf_batch_set_jk_lanes(aimdhf_plan* plan, int lanes);

/*
 * Pulay DIIS coefficients of nrep independent problems (the solve of
 * aimd/qc/scf.py:540-560 for many SCFs at once): device G (nrep, m, m) and r
 * (nrep, m), c (nrep, m) = the solution of G c = r by Gaussian elimination with
 * partial pivoting, unknowns of singular pivots 0. m <= 12. No plan; capturable.
 */
int aimdhf_batch_diis_solve(int nrep, int m, const double* G, const double* r, double* c, void* stream);

/* err[0] (device) = |tr X - sum_ij X_ij^2|, the idempotency error of a symmetric
 * purification iterate X (n x n, device). Row sums, then their sum in a fixed
 * order (bitwise reproducible); the row sums go through the same device array
 * as aimdhf_tc2_init's bounds, so calls on one device must not overlap. No plan. */
int aimdhf_tc2_error(int n, const double* X, double* err, void* stream);

/*
 * Independent launches (the J and K kernels of a build, the store classes,
 * the derivative-ERI classes) run on the caller's stream and three extra
 * library streams of default priority, joined back into the caller's stream
 * (concurrent = 1, default), or all on the caller's stream (0: each kernel's
 * timing is then its own, e.g. for per-kernel profiling). Results are
 * identical either way. No counterpart in the reference.
 */
int aimdhf_set_streams(aimdhf_plan* plan, int concurrent);

/* Whether the current geometry has an integral store, its size in bytes, the
 * gather-list lengths of the K and J kernels and the bytes of their
 * gather-ordered integral data (FP64 and, in mixed precision, FP32), i.e. what
 * one J/K build streams (any pointer may be NULL). */
/*
 * Lane items per work unit of the store-mode J/K digestion (64 .. 65536,
 * default 1024; one wavefront digests one unit), from the next
 * aimdhf_set_geometry on. A tuning knob with no counterpart in the
 * reference; results do not depend on it beyond the order of the unit sums.
 */
int aimdhf_set_store_units(aimdhf_plan* plan, int items);

int aimdhf_jk_info(aimdhf_plan* plan, int* stored, int64_t* store_bytes, int64_t* k_entries, int64_t* j_entries,
                   int64_t* gather_bytes);

/*
 * The integral store of the current geometry broken down for traffic
 * accounting: per exchange output type t = 3 l_X + l_Y (9) the gather-ordered
 * bytes one build streams and the gather-list entries, the same per Coulomb
 * pair class (6: ss ps pp ds dp dd), and the bytes of the unique screened
 * integrals per quartet class (21). All zero in direct mode (any pointer may
 * be NULL). No counterpart in the reference (its integrals are not stored);
 * for profilers and benchmarks.
 */
int aimdhf_jk_type_info(aimdhf_plan* plan, int64_t* k_bytes, int64_t* k_entries, int64_t* j_bytes,
                        int64_t* j_entries, int64_t* unique_bytes);

/*
 * One-electron matrices of the current geometry into device buffers:
 * S, T, V (nao,nao) and the dipole integrals M (3,nao,nao) about the origin,
 * V over all nuclei as point charges Z_A.
 * Replaces: one_electron_matrices and dipole_integrals
 * (aimd/qc/integrals.py:1379-1418).
 */
int aimdhf_one_electron(aimdhf_plan* plan, double* S, double* T, double* V, double* M, void* stream);

/*
 * grad (device (natm,3)) = d/dR [sum D (T+V)] - d/dR [sum W S] at fixed D, W
 * (symmetric parts used), including the Hellmann-Feynman term.
 * Replaces: one_electron_gradient (aimd/qc/integrals.py:1483-1495).
 */
int aimdhf_grad_one(aimdhf_plan* plan, const double* D, const double* W, double* grad, void* stream);

/*
 * aimdhf_grad_one without any device allocation: the same kernels and the same
 * (bitwise identical) result, with the symmetrised D and -W and the fixed-point
 * gradient held in a fixed library array of 16 MiB + 64 KiB per device instead
 * of plan buffers. Needs 16 nao^2 + 48 natm bytes (plus alignment) of it, i.e.
 * nao <= 1024 for small natm; larger plans get AIMDHF_E_ARG (use
 * aimdhf_grad_one). Calls no hipMalloc / hipFree and clears with a fill kernel,
 * so code under HIP graph capture, tight serving loops and harnesses that
 * forbid allocation can call it. Every call on one device shares that device's
 * array and leaves its contents behind: calls that overlap on one device
 * (different streams or threads) race. No counterpart in the reference (its
 * one_electron_gradient, aimd/qc/integrals.py:1483-1495, is the function this
 * entry point and aimdhf_grad_one replace); it exists for callers that cannot
 * allocate.
 */
int aimdhf_grad_one_noalloc(aimdhf_plan* plan, const double* D, const double* W, double* grad, void* stream);

/*
 * The terms of aimdhf_grad_one separately, each device (natm,3):
 *   g_kin = d/dR sum D T,  g_nuc = d/dR sum D V (basis-function and operator
 *   centres),  g_ovl = -d/dR sum W S  (g_kin + g_nuc + g_ovl = aimdhf_grad_one).
 * Replaces: kinetic_gradient, nuclear_attraction_gradient and -overlap_gradient
 * (aimd/qc/integrals.py:1452-1480) as gradient_terms uses them
 * (aimd/qc/gradients.py:196-223).
 */
int aimdhf_grad_one_terms(aimdhf_plan* plan, const double* D, const double* W, double* g_kin, double* g_nuc,
                          double* g_ovl, void* stream);

/*
 * grad (device (natm,3)) = d/dR of
 *   E2 = 1/2 sum Dc_mn Dc_ls (mn|ls) - k/2 sum_x sum Dx_ml Dx_ns (mn|ls)
 * at fixed densities (RHF: Dc = P, Dx = [P], k = 0.5; UHF: Dc = Pa+Pb,
 * Dx = [Pa, Pb], k = 1), derivative integrals recomputed on the fly; quartets
 * with Q_P Q_Q max|Gamma| < threshold are skipped. Synchronises the stream.
 * Replaces: two_electron_gradient (aimd/qc/integrals.py:1498-1520).
 */
int aimdhf_grad_two(aimdhf_plan* plan, const double* Dc, int nx, const double* Dx, double kfac,
                    double threshold, double* grad, void* stream);

/*
 * Kernel timing with HIP events recorded on the caller's stream around every
 * launch while enabled (no counterpart in the reference, whose timings are
 * wall-clock dicts, aimd/backends/hf.py:347-354). Kinds: 0..5 direct J-engine
 * per bra pair class (ss ps pp ds dp dd), 6..14 direct exchange per output
 * type (l_X, l_Y), 21..41 derivative-ERI kernels of the 21 quartet classes,
 * 42 geometry set-up, 43 one-electron integrals, 44 one-electron gradient,
 * 45 J/K prologue and epilogue, 46 Hermite densities, 47 integral store
 * evaluation, 15 gather-ordered copies, 48 gather-list build, 49..54 store J
 * per pair class, 55..63 store K per output type, 64..84 derivative-ERI
 * screening pre-pass per class; 85..88 host wall-clock phases (energy list,
 * owner lists, store build incl. its device work, gradient list) and 89..92
 * sub-phases of the store build (list counts, band tables, list fill, unit
 * tables; milliseconds of the calling thread, not events). aimdhf_profile_read synchronises, returns the
 * summed milliseconds and launch counts per kind (nkind <= 96) and optionally
 * resets.
 */
int aimdhf_profile_enable(aimdhf_plan* plan, int on);
int aimdhf_profile_read(aimdhf_plan* plan, double* ms, int64_t* launches, int nkind, int reset);

/*
 * Average duration in milliseconds of `reps` back-to-back launches, on
 * `stream`, of one store-mode digestion kernel with the arguments of the last
 * J/K build at the current geometry: kind 55 + t (exchange, output type t) or
 * 49 + c (Coulomb, pair class c). The kernels write only their unit partials,
 * so J and K are left as they were. No counterpart in the reference; for
 * benchmarks (one event pair around many launches times the kernel without
 * the per-launch event and dispatch gaps).
 */
int aimdhf_profile_replay(aimdhf_plan* plan, int kind, int reps, void* stream, double* avg_ms);

/* Per-class sizes (21 each) of the energy quartet list and of the last
 * gradient quartet list (either pointer may be NULL). */
int aimdhf_class_counts(aimdhf_plan* plan, int64_t* energy_counts, int64_t* gradient_counts);

/* The quartets the last aimdhf_grad_two differentiated (those passing its
 * Q_P Q_Q max|Gamma| screening) as (P, Q) shell-pair indices, P >= Q, copied
 * to host memory (2*cap int64); *count receives their number. Synchronous.
 * No counterpart in the reference (used to price the derivative kernels). */
int aimdhf_grad_quartets(aimdhf_plan* plan, int64_t* out_host, int64_t cap, int64_t* count);

/* Kept primitive pairs of every shell pair (npair = nshell (nshell+1)/2),
 * copied to host memory. Synchronous. */
int aimdhf_pair_npp(aimdhf_plan* plan, int32_t* npp_host, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* AIMDHF_H */
