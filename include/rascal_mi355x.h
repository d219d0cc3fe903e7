/*
 * C ABI of the MI355X (gfx950) SOAP + sparse-GAP library librascal_mi355x.so.
 *
 * Drop-in boundary for librascal's SOAP prediction path.  One context owns
 * the calculator (hypers, radial tables), the current batch of structures
 * (neighbour list, expansion, power spectrum) and the sparse GAP model.
 *
 * Conventions
 *   - All functions return RGPU_OK (0) or an RGPU_ERR_* code; the message of
 *     the last failure on the calling thread is rgpu_last_error().
 *   - Array arguments are plain pointers; output pointers may be host or
 *     device memory (copies use hipMemcpyDefault).  Nothing is retained from
 *     caller memory after a call returns.
 *   - positions [n_atoms][3]; cells [n_struct][9] with cell[d*3+k] the
 *     component d of lattice vector k (lattice vectors are columns, as the
 *     reference AtomicStructure); pbc [n_struct][3].
 *   - Dense feature layout over the context's sorted species list S:
 *       expansion  [n_centres][S][N][(L+1)^2]   (n slowest, lm = l*l+l+m)
 *       invariants [n_centres][K][N*N][L+1]     (K = S(S+1)/2 keys a<=b)
 *     which is the reference's dense export (property_block_sparse.hh
 *     :911-932) before dropping keys absent from the whole batch.
 *   - Centres: every atom, or the atoms selected by a centre mask, in atom
 *     order.  Per-centre outputs (features, expansion, key masks, Atom K_NM
 *     rows) have one row per centre; per-atom outputs (gradients) one row per
 *     atom.  Gradients of masked batches are refused, as in the reference.
 *   - A sparsified power spectrum (hypers "coefficient_subselection") has
 *     one feature per selected coefficient, in the order given (FP64 only).
 *   - Calls on one context are not re-entrant (as the reference calculators).
 */
#ifndef RASCAL_MI355X_H_
#define RASCAL_MI355X_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RGPU_OK 0
#define RGPU_ERR_INVALID 1 /* bad argument or hypers (reference: std::logic_error) */
#define RGPU_ERR_RUNTIME 2 /* bad structure etc. (reference: std::runtime_error) */
#define RGPU_ERR_HIP 3     /* HIP runtime failure */
#define RGPU_ERR_STATE 4   /* call out of order (e.g. no structures set) */

#define RGPU_FP64 0  /* FP64 everywhere (parity path) */
#define RGPU_MIXED 1 /* FP32 expansion/invariants, BF16 hi+lo GEMM, FP64 reductions */

#define RGPU_INFO_MAX_RADIAL 0
#define RGPU_INFO_MAX_ANGULAR 1
#define RGPU_INFO_N_SPECIES 2
#define RGPU_INFO_FEATURE_SIZE 3 /* D = K N^2 (L+1) (PowerSpectrum; the number of selected
                                    coefficients with coefficient_subselection) or S N (L+1)^2 */
#define RGPU_INFO_N_CENTRES 4
#define RGPU_INFO_N_PAIRS 5      /* strict pairs including the self pairs */
#define RGPU_INFO_N_STRUCT 6
#define RGPU_INFO_SPLINE_GRID 7  /* 0 when the radial integral is not splined */
#define RGPU_INFO_N_SPARSE 8
#define RGPU_INFO_IS_INVARIANTS 9 /* 1 PowerSpectrum, 0 SphericalExpansion */
#define RGPU_INFO_PRECISION 10
#define RGPU_INFO_GRAPH_CAPTURES 11 /* graphs captured by rgpu_compute */
#define RGPU_INFO_GRAPH_LAUNCHES 12 /* graph replays by rgpu_compute */
#define RGPU_INFO_TILE_PAIRS_FWD 13 /* autotuned pairs per batch, expansion (0 = default) */
#define RGPU_INFO_TILE_KEYS_PS 14   /* autotuned key blocks per pass, power spectrum */
#define RGPU_INFO_TILE_PAIRS_ADJ 15 /* autotuned pairs per batch, adjoint */
#define RGPU_INFO_TILE_KEYS_ADJ 16  /* key blocks per pass of the FP32 adjoint (0 = auto) */
#define RGPU_INFO_NL_GRAPH_LAUNCHES 17 /* neighbour-list passes replayed from a graph */
#define RGPU_INFO_N_ATOMS 18       /* atoms of the batch (N_CENTRES counts the centres) */
#define RGPU_INFO_TILE_ADJ_KIND 19 /* mixed-precision adjoint: 0 vector ALUs, 1 matrix cores */
#define RGPU_INFO_HALF_LIST 20     /* 1: the expansion runs over the half neighbour list */
#define RGPU_INFO_FUSED 21         /* 1: expansion and invariants in one persistent launch */
#define RGPU_INFO_FWD_LDS 22       /* 1: the expansion accumulates c in LDS and writes it once */
#define RGPU_INFO_FWD_GROUP 23     /* 1: the FP32 expansion runs on pair groups where it applies */

/* stages timed by rgpu_enable_timing */
#define RGPU_STAGE_NEIGHBOURS 0
#define RGPU_STAGE_EXPANSION 1
#define RGPU_STAGE_POWER_SPECTRUM 2
#define RGPU_STAGE_GAP_KERNEL 3  /* K = (X T^T)^zeta GEMM + epilogue */
#define RGPU_STAGE_GAP_FORCE 4   /* T~ = W~ T GEMM */
#define RGPU_STAGE_BACKWARD 5    /* dL/dX -> dL/dc -> per-pair forces */
#define RGPU_N_STAGES 6

typedef struct rgpu_context rgpu_context;

const char* rgpu_last_error(void);

/* Calculator from the reference's JSON hypers: a SphericalInvariants
 * (soap_type "PowerSpectrum") or SphericalExpansion hypers dict.
 * Replaces CalculatorSphericalInvariants(const json&)
 * (representations/calculator_spherical_invariants.hh:171-176) and
 * CalculatorSphericalExpansion(const json&) (calculator_spherical_expansion.hh
 * :1779-1783), including the radial-spline set-up (:1107-1135). */
int rgpu_context_create(const char* hypers_json, rgpu_context** ctx);
void rgpu_context_destroy(rgpu_context* ctx);

/* HIP stream used by every later call (default: the null stream).  A
 * prediction with forces also runs the reverse pair lists of its force sum
 * on a second stream owned by the context (default priority), joined into
 * this stream by an event before the forces. */
int rgpu_set_stream(rgpu_context* ctx, void* hip_stream);
/* Replay the representation stage (cutoff, expansion, invariants) of
 * rgpu_compute as a captured HIP graph while the batch's buffers and sizes
 * are unchanged (positions may change; not used while stage timing is on),
 * and the device pass of the neighbour list built by rgpu_update_positions
 * likewise.  Creates a private stream when no stream was set.  No
 * counterpart in the reference (a CPU library). */
int rgpu_enable_graphs(rgpu_context* ctx, int enable);
/* Autotuner: times the candidate work partitions (pairs staged per batch in
 * the expansion and the adjoint, key blocks per pass in the power spectrum
 * and in the mixed-precision adjoint) on the current batch and precision and
 * keeps the fastest for later calls
 * (RGPU_INFO_TILE_*).  No counterpart in the reference. */
int rgpu_autotune(rgpu_context* ctx);
/* Sets the work partitions of the current precision directly (values <= 0:
 * built-in choice; keys_adj: key blocks of dL/dX staged in LDS per pass of
 * the mixed-precision adjoint; adj_kind: its staged vector-ALU kernel (0),
 * matrix-core kernel (1), register-tile kernel (2), or register-tile key
 * contraction followed by a pair kernel of its own (3, the default; 2 and 3
 * are built for max_radial 12 / max_angular 9 and 8 / 6 and up to 8
 * species, kind 0 runs elsewhere); fused: expansion and invariants in one persistent launch (1),
 * two launches (0), two launches with the expansion writing c once from
 * LDS (2), two launches with the FP32 expansion on pair groups (3),
 * built-in (-1: two launches)), e.g. to reuse an autotuned
 * set.  Results agree between partitions to rounding.  No counterpart in
 * the reference. */
int rgpu_set_tiles(rgpu_context* ctx, int pairs_fwd, int keys_ps, int pairs_adj, int keys_adj,
                   int adj_kind, int fused);
/* Expansion over the half neighbour list (enable 1) or the full list (0,
 * the default): each unordered pair's contribution is computed once and
 * given to the other centre with the parity (-1)^l; the results equal the
 * full list's.  The gradients (adjoint) always use the full list.  Refused,
 * with the reference's messages, for masked centres and for periodic cell
 * vectors shorter than twice the cutoff.  Replaces AdaptorHalfList in the
 * manager stack (adaptor_half_neighbour_list.hh:319-380) with the
 * expansion's half-list branch (calculator_spherical_expansion.hh
 * :2159-2185, 2282-2299).  rgpu_autotune picks it where it is faster. */
int rgpu_set_half_list(rgpu_context* ctx, int enable);
/* RGPU_FP64 (default) or RGPU_MIXED; invalidates computed results. */
int rgpu_set_precision(rgpu_context* ctx, int precision);
int rgpu_get_info(rgpu_context* ctx, int key, int64_t* value);

/* Batch of structures (host arrays).  Builds the neighbour list on the GPU.
 * species[n_species] sorted ascending must contain every atom type (with
 * expansion_by_species_method "user defined" it must contain the
 * global_species, and every centre's type must be one of them).
 * centre_mask[n_atoms] (may be NULL: every atom) selects the centres, as
 * the reference's center_atoms_mask (neighbourlist/structure_manager.py
 * :337-340; structure_manager_centers.cc:37-100).  Sparse points set
 * earlier are dropped when the species list changes.
 * Replaces AtomsList / ManagerCollection::add_structures with the stack
 * centers -> neighbourlist -> centercontribution -> strict
 * (structure_manager_collection.hh:177-236; adaptor_neighbour_list.hh
 * :904-1118; adaptor_center_contribution.hh:334-403; adaptor_strict.hh
 * :373-455). */
int rgpu_set_structures(rgpu_context* ctx, int n_struct, const int64_t* atom_offsets,
                        const double* positions, const int32_t* types, const double* cells,
                        const int32_t* pbc, int n_species, const int32_t* species,
                        const uint8_t* centre_mask);
/* New positions (and cells, may be NULL) for the same atoms and types, e.g.
 * one MD step; rebuilds the neighbour list. */
int rgpu_update_positions(rgpu_context* ctx, const double* positions, const double* cells);

/* Strict pair list of every atom, self pair first: centre_offsets[n_atoms+1],
 * pair_j[n_pairs] (original atom index within the batch), pair_vec[n_pairs][3]
 * = r_j(image) - r_i.  pair_vec may be NULL. */
int rgpu_get_neighbours(rgpu_context* ctx, int64_t* centre_offsets, int32_t* pair_j,
                        double* pair_vec);

/* Expansion (and power spectrum for invariants hypers) of the batch.
 * Replaces compute() (calculator_spherical_invariants.hh:506-531,
 * 693-931; calculator_spherical_expansion.hh:2139-2411). */
int rgpu_compute(rgpu_context* ctx);
/* expansion coefficients (dense layout) and per-centre key masks (bit s set
 * when species s is among the centre's expansion keys, by the hypers'
 * expansion_by_species_method: the centre's and its neighbours' species,
 * those of its structure, or the global species;
 * calculator_spherical_expansion.hh:2414-2570); either may be NULL */
int rgpu_get_expansion(rgpu_context* ctx, double* coeffs, uint64_t* key_mask);
/* power spectrum [n_centres][D], normalised when the hypers say so, and the
 * per-centre inverse norms (1 when not normalising); either may be NULL */
int rgpu_get_features(rgpu_context* ctx, double* features, double* inv_norms);

/* Gradient of the features with respect to the positions, the reference's
 * dense get_features_gradient matrix: out[3 n_pairs][D], rows 3 r + x for
 * the pairs r of rgpu_get_neighbours (self pair first per centre), row
 * 3 r + x = d X_i / d r_j,x for the pair (i, j); for the self pair the
 * derivative with respect to r_i (bind_py_structure_manager.cc:622-744;
 * calculator_spherical_invariants.hh:796-921).  FP64 precision, no centre
 * mask.  out may be host or device memory. */
int rgpu_get_features_gradient(rgpu_context* ctx, double* out);

/* Adjoint of the features: given dL/dX [n_centres][D] (X as returned by
 * rgpu_get_features), writes dL/dr [n_atoms][3] and the negative virial
 * [n_struct][6] (Voigt xx,yy,zz,yz,xz,xy, divided by the cell volume); either
 * output may be NULL.  Uses the same pair-slot conventions as the reference's
 * gradient properties (calculator_spherical_expansion.hh:2306-2401;
 * sparse_kernel_predict.hh:206-377). */
int rgpu_features_backward(rgpu_context* ctx, const double* dL_dX, double* dL_dr,
                           double* neg_stress);

/* Sparse points of a GAP model: values [n_sparse][D] in the context's dense
 * layout, species[n_sparse] centre types, weights[n_sparse] (may be NULL for
 * kernel-only use).  Replaces SparsePointsBlockSparse::push_back / JSON load
 * (models/sparse_points.hh:372-500) and SparseKernel{"name":"GAP","zeta":z}
 * (models/sparse_kernels.hh:517-742). */
int rgpu_set_sparse_points(rgpu_context* ctx, int zeta, int64_t n_sparse, const int32_t* species,
                           const double* values, const double* weights);
/* K_NM: target 0 = Structure [n_struct][n_sparse] (sum over centres),
 * 1 = Atom [n_centres][n_sparse]; columns in the order the sparse points were
 * given (SparseKernelImpl<GAP>::compute, sparse_kernels.hh:104-190). */
int rgpu_gap_knm(rgpu_context* ctx, int target, double* knm);
/* Gradient rows of K_NM (training with forces and stress): out
 * [3 n_atoms (+ 6 n_struct)][n_sparse], row 3 j + x = d/dr_j,x of
 * sum_i k(X_i, T_m) over the centres i of m's species; with neg_stress the
 * per-structure Voigt rows (xx, yy, zz, yz, xz, xy) of r_ji (x) dk/dr_j
 * divided by the cell volume follow.  FP64 precision only.  Replaces
 * SparseKernelImpl<GAP>::compute_derivative (models/sparse_kernels.hh
 * :258-499). */
int rgpu_gap_knm_gradient(rgpu_context* ctx, int neg_stress, double* out);
/* K_MM [n_sparse][n_sparse] in FP64: (T_a T_a^T)^zeta between sparse points
 * of the same species a, 0 across species; rows and columns in the order the
 * sparse points were given (SparseKernelImpl<GAP>::compute(sparse_points),
 * models/sparse_kernels.hh:136-150). */
int rgpu_gap_kmm(rgpu_context* ctx, double* kmm);
/* out [cols][cols] = A^T A for A [rows][cols], row-major FP64 host arrays,
 * on the FP64 matrix cores (no context needed).  Replaces the normal-matrix
 * product np.dot(KNM.T, KNM) of train_gap_model (models/krr.py:468). */
int rgpu_gram_f64(int64_t rows, int64_t cols, const double* A, double* out);
/* Prediction without K_NM rows: energy[n_struct] = K_NM w, gradient
 * [n_atoms][3] = dE/dr, neg_stress [n_struct][6] (/ volume).  Any output may
 * be NULL; gradient/neg_stress trigger the adjoint
 * (compute_sparse_kernel_gradients / _neg_stress,
 * sparse_kernel_predict.hh:206-377).  Every floating-point sum runs in a
 * fixed order (no floating-point atomics): repeated calls on the same
 * inputs return identical bits. */
int rgpu_gap_predict(rgpu_context* ctx, double* energy, double* gradient, double* neg_stress);

/* Allocation-free GAP energies: energy[n_struct] (device memory) = K_NM w,
 * as rgpu_gap_predict's energies.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree or other allocating routine: the per-centre energies and the
 * energy GEMM's per-column-tile partial sums live in a static device array
 * of 2^23 doubles (64 MiB per device), so a batch of n centres is accepted
 * when ceil(n/32)*32 + n*T <= 2^23, T being the column tiles of the widest
 * species block (ceil(M_s/128) mixed, ceil(M_s/64) FP64; larger batches
 * are rejected with an error), and every context buffer must already be
 * sized for the batch
 * (one rgpu_gap_predict on it first; otherwise an error, nothing allocated).
 * The call is stream-ordered (no synchronisation): the energies are ready
 * when the context's stream reaches them.  The scratch array is shared by
 * every call on a device: calls on one device that overlap in time race. */
int rgpu_gap_energy_noalloc(rgpu_context* ctx, double* energy);

/* Stage timers: HIP events recorded on the context stream around each
 * stage's kernels.  rgpu_get_timing synchronises and returns the summed
 * milliseconds and the number of launches since the last reset.  No
 * counterpart in the reference (its profilers time whole calls,
 * performance/profiles/profile_spherical_expansion.cc:110-171). */
int rgpu_enable_timing(rgpu_context* ctx, int enable);
int rgpu_get_timing(rgpu_context* ctx, int stage, double* total_ms, int64_t* launches);
int rgpu_reset_timing(rgpu_context* ctx);

/* Diagnostics (unit-level parity).  Radial functions at n distances r[n]:
 * R[n][N][L+1] and dR/dr (dR may be NULL), evaluated by the GPU path of the
 * current precision: the splined O^T R(r) for Spline hypers
 * (calculator_spherical_expansion.hh:1107-1156), the raw GTO integrals
 * before orthonormalisation for the direct Hyp1f1 basis (:518-584). */
int rgpu_eval_radial(rgpu_context* ctx, int64_t n, const double* r, double* R, double* dR);
/* Real spherical harmonics at n unit vectors dirs[n][3]: Y[n][(L+1)^2] and
 * the gradient on the unit sphere dY[n][3][(L+1)^2] (may be NULL)
 * (math/spherical_harmonics.cc:135-363). */
int rgpu_eval_ylm(rgpu_context* ctx, int64_t n, const double* dirs, double* Y, double* dY);

#ifdef __cplusplus
}
#endif
#endif /* RASCAL_MI355X_H_ */
