/*
 * FP64 CPU oracle for the SOAP + sparse-GAP path of librascal.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain serial C++,
 * the algorithms of the reference files cited in rascal_oracle.cc.  It is
 * the checker for the HIP implementation and the timed CPU baseline; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product (librascal_amd) never links or calls it.
 *
 * How each part is pinned (tests/test_oracle_golden.py, test_oracle_math.py):
 *   Hyp1f1 - 4800 mpmath values of the reference (15 DBL_FTOL), Kummer
 *     identity, recursion = direct; Y_lm / ALPs - reference ubjson (2 / 10
 *     DBL_FTOL), scipy closed form, finite differences; cutoffs, radial
 *     derivatives, representation gradients - finite differences; GTO
 *     radial - closed form with scaled Bessel functions; expansion - 32 GTO
 *     rows of spherical_expansion_reference (1e-8); power spectrum - cosine
 *     kernels of kernel_reference (1e-10) and the 30 sparse points of
 *     simple_gap_model.json; neighbour list - brute force and cell-skew
 *     invariance; K_NM, K_MM - closed forms; GAP forces / stress / K_NM
 *     derivative - finite differences and K_NM_der . w = gradient.
 *   Parity unpinned: none.
 *
 * Data conventions (shared with the HIP library):
 *   - positions: [n_atoms][3] row-major, Cartesian, Angstrom
 *   - cell: [3][3] with cell[d*3 + k] = component d of lattice vector k
 *     (lattice vectors are the COLUMNS, as in the reference AtomicStructure)
 *   - a batch of structures is given by atom_offsets[n_struct+1]
 *   - dense species layout: species_list[n_species] sorted ascending;
 *     expansion  c[centre][s][n][l*l + l + m]      (S * N * (L+1)^2)
 *     invariants p[centre][key(a<=b)][n1*N + n2][l] (K * N*N * (L+1))
 */
#ifndef RASCAL_ORACLE_H_
#define RASCAL_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_hypers {
  double cutoff;          /* interaction_cutoff (A) */
  double smooth_width;    /* cutoff_smooth_width (A) */
  double gaussian_sigma;  /* gaussian_sigma_constant (A) */
  int max_radial;
  int max_angular;
  int cutoff_type;        /* 0 ShiftedCosine, 1 RadialScaling */
  double rs_rate, rs_scale, rs_exponent;
  int optimization;       /* 0 none (direct Hyp1f1), 1 Spline */
  double spline_accuracy;
  int normalize;
  int compute_gradients;
} orc_hypers;

/* Last error message (thread-local), empty when no error. */
const char* orc_last_error(void);

/* ---- math unit functions (golden-vector checks) ---- */
double orc_hyp1f1(double a, double b, double z, int derivative, int mmax,
                  double tol);
double orc_hyp1f1_switch_point(double a, double b);
/* values/derivatives [N][L+1] of G(n,l) as in the GTO radial integral */
int orc_hyp1f1_sph(int max_radial, int max_angular, double r, double alpha,
                   const double* fac_b, int recursion, double tol, double* values,
                   double* derivatives);
/* harmonics[(L+1)^2]; derivatives[3][(L+1)^2] (may be NULL); alp[(L+1)*(L+2)]
 * (may be NULL) */
int orc_ylm(int max_angular, const double* direction, double* harmonics,
            double* derivatives, double* alp);
double orc_cutoff_fc(const orc_hypers* h, double r);
double orc_cutoff_dfc(const orc_hypers* h, double r);

/* ---- calculator ---- */
void* orc_calc_create(const orc_hypers* h);
void orc_calc_destroy(void* calc);
/* radial tables: fac_b[N], ortho_norm O[N][N] (row-major), center[N]
 * (finalised or not as the reference), spline grid size (0 if no spline) */
int orc_calc_radial_info(void* calc, double* fac_b, double* ortho_norm,
                         double* center, int* spline_grid_size,
                         double* spline_x1, double* spline_x2);
/* radial function at r: R[N][L+1] and dR/dr (either may be NULL) */
int orc_calc_radial(void* calc, double r, double* R, double* dR);
/* full spline tables: y[grid][N*(L+1)] and y2*h^2/6 [grid][N*(L+1)]; the
 * column index is l*N + n (Eigen column-major flattening of the N x (L+1)
 * radial matrix) */
int orc_calc_spline_tables(void* calc, double* y, double* d2);

/* ---- neighbour list (one structure) ---- */
/* returns the number of strict pairs including self pairs, or -1 on error */
int64_t orc_nl_count(int n_atoms, const double* pos, const double* cell,
                     const int* pbc, double cutoff);
/* centre_offsets[n_atoms+1], pair_j[n_pairs] (original atom index),
 * pair_vec[n_pairs][3] (r_j(image) - r_i), pair_dist[n_pairs] */
int orc_nl_fill(int n_atoms, const double* pos, const double* cell,
                const int* pbc, double cutoff, int64_t* centre_offsets,
                int* pair_j, double* pair_vec, double* pair_dist);

/* ---- representation (batch of structures) ---- */
typedef struct orc_batch {
  int n_struct;
  const int64_t* atom_offsets; /* [n_struct+1] */
  const double* positions;     /* [n_atoms][3] */
  const int* types;            /* [n_atoms] */
  const double* cells;         /* [n_struct][9] */
  const int* pbc;              /* [n_struct][3] */
  int n_species;
  const int* species;          /* sorted ascending */
} orc_batch;

/* total strict pairs (with self pairs) over the batch */
int64_t orc_batch_pairs(void* calc, const orc_batch* b);
/* expansion coefficients [n_atoms][S][N][(L+1)^2] and the per-centre key
 * mask (bit s set when species s is in the centre's key set) */
int orc_expansion(void* calc, const orc_batch* b, double* coeffs,
                  uint64_t* key_mask);
/* expansion gradients per pair slot [n_pairs][3][S][N][(L+1)^2] */
int orc_expansion_gradients(void* calc, const orc_batch* b, double* grads);
/* power spectrum [n_atoms][K][N*N][L+1] (normalised if hypers.normalize),
 * inv_norms[n_atoms] (1/|p| or 1 when not normalising); only the first
 * max_centres centres of the batch when max_centres >= 0 */
int orc_power_spectrum(void* calc, const orc_batch* b, double* ps,
                       double* inv_norms, int64_t max_centres);
/* power-spectrum gradients per pair slot [n_pairs][3][D] and the gradient
 * info rows [n_pairs][5] = (struct, i, j, type_i, type_j) with j the original
 * atom index within the structure */
int orc_power_spectrum_gradients(void* calc, const orc_batch* b, double* grads,
                                 int* info);

/* ---- sparse GAP ---- */
typedef struct orc_sparse {
  int n_sparse;            /* M */
  const int* sp_species;   /* [M] centre species of each point, grouped in
                              ascending species order */
  const double* values;    /* [M][D] dense features in the batch layout */
  int zeta;
} orc_sparse;

/* K_NM: target 0 = Structure ([n_struct][M]), 1 = Atom ([n_atoms][M]) */
int orc_gap_knm(void* calc, const orc_batch* b, const orc_sparse* sp,
                int target, double* knm);
/* K_MM [M][M] */
int orc_gap_kmm(const orc_sparse* sp, int D, double* kmm);
/* gradient rows [3*n_atoms (+6*n_struct)][M] */
int orc_gap_knm_derivative(void* calc, const orc_batch* b,
                           const orc_sparse* sp, int neg_stress, double* out);
/* prediction: energy[n_struct] = K_NM . w (no baseline), gradient
 * [n_atoms][3] = dE/dr, neg_stress [n_struct][6] (Voigt, / volume).  Any
 * output may be NULL.  The work covers max_centres centres (-1 = all) from
 * the centre with ordinal first_centre in the batch (timing subsets and
 * parallel workers). */
int orc_gap_predict(void* calc, const orc_batch* b, const orc_sparse* sp,
                    const double* weights, double* energy, double* gradient,
                    double* neg_stress, int64_t max_centres, int64_t first_centre);

#ifdef __cplusplus
}
#endif
#endif  // RASCAL_ORACLE_H_
