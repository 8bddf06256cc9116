/*
 * oracle.h -- FP64 CPU restatement of mtcode's per-timestep Langevin step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the mtcode_amd package,
 * libmdx.so, the drop-in host classes) may include, link or call this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only to check results.
 *
 * Every function restates one reference function; the citation is given next
 * to its declaration (paths relative to the reference's Code/ directory).
 * Arithmetic follows the reference expression by expression (operation order,
 * the long-double `pi` of Basic/basic.h:32, the literal 1.12246 WCA cut-off,
 * the 0.999999 acos clamp) so that FP64 results agree to rounding.
 *
 * Pinning (what checks each function independently of this file):
 *   orc_fasterfloor / orc_assign_box / orc_boxes_relationships /
 *   orc_pairs_reference          : golden fixtures from the compiled reference
 *                                  (oracle/_ref), brute force O(N^2) pair search
 *   orc_hs_force / orc_wca_energy: closed form, finite difference of energy
 *   orc_kf2_*                    : golden fixtures, finite differences of
 *                                  orc_kf2_energy in position and rotation,
 *                                  rotational covariance, Newton's third law
 *   orc_bond_* / orc_bend_*      : golden fixtures (harmonic, FENE, harmonic
 *                                  angle), finite differences; the cosine angle
 *                                  form and the active force have no reference
 *                                  counterpart and are pinned by closed form
 *                                  and finite differences only
 *   integrators                  : 1000-step trajectory fixtures from the
 *                                  compiled reference with injected noise,
 *                                  rotation-matrix orthonormality, affine-in-
 *                                  noise checks
 *   orc_philox4x32_10            : the Random123 known-answer vectors
 *   orc_noise_uniform            : moments / support (its own definition)
 *   orc_calcg1                   : golden fixture from the compiled reference
 *   orc_bind_forces              : golden fixtures from the compiled reference
 *                                  (bound sets, forces, torques after one
 *                                  call with the reference's rand() draws)
 *   orc_patch_list,
 *   orc_bound_pairs              : golden fixtures from the compiled reference
 *                                  (calculate_patch_list and
 *                                  calculate_bound_pairs on the reference's
 *                                  own pairs, rows bit-exact)
 *   orc_dgraph_labels            : golden fixture from the compiled reference
 *                                  (distance_graph); scipy's connected
 *                                  components in the tests
 */
#ifndef MTCODE_ORACLE_H
#define MTCODE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- geometry */

/* Box: orthorhombic generalisation of `cube` (MDBase/geometry.h:43-66).
 * For a cube l[0]=l[1]=l[2].  `periodic` follows the reference's single
 * flag (`periodic = pb[0]`, geometry.h:53) for distances and wrapping. */
typedef struct {
    double l[3];
    int periodic;
} orc_box;

/* Basic/basic.h:58 */
int orc_fasterfloor(double x);

/* MDBase/geometry.h:508-521 with dim=[1,n0,n0*n1] (MD.cpp:1071-1078).
 * n per axis; for a cube all three are the reference's single n. */
int orc_assign_box(const orc_box *b, const double *x3, const int n[3]);
void orc_assign_all(const orc_box *b, int N, const double *x, const int n[3], int *cell);

/* MDBase/geometry.h:523-609: 27-cell stencil, wrapped, sorted and deduped per
 * row.  Writes ncells*ncols ints, returns ncols (27 when every n>=3). */
int orc_boxes_relationships(const int n[3], int *out);

/* MDBase/geometry.h:145-164 (fabs(dx) > l/2 test; strict r2 < cut^2) */
int orc_distance_less_than(const orc_box *b, const double *xi, const double *xj, double cut);

/* MDBase/geometry.h:277-294 (SQR(dx) > (l/2)^2 test); un = xi - xj min-imaged */
void orc_distance_vector(const orc_box *b, const double *xi, const double *xj, double un[3], double *r2);

/* MD::calculatepairs (MD.cpp:1035-1098) + precalculatepairs (MD.cpp:575-713),
 * serial order.  cells per side = n (the reference derives it from the
 * stencil table).  Writes up to cap pairs (i,j) into ij, returns the count
 * (which may exceed cap; then only cap were written). */
long long orc_pairs_reference(const orc_box *b, int N, const double *x, const int n[3],
                              double cut, int *ij, long long cap);

/* --------------------------------------------------------------- potentials */

/* HSPotential::force (Potentials/potential.h:112-125), paramlimit 1.12246*sigma */
double orc_hs_force(double r, double eps, double sigma, double rcut);
/* WCAPotential::energy inner branch (potential.h:179-185) with attrepsilon=0 */
double orc_wca_energy(double r, double eps, double sigma, double rcut);

/* pair-list force loop MD::calculateforces (MD.cpp:1684-1764).
 * kind: 0 = HS/WCA (p0=eps, p1=sigma, p2=rcut<=0 -> 1.12246 sigma)
 *       1 = harmonic (HarmonicPotential, potential.h:265-297; per-pair k, r0
 *           via kk/r0 arrays when non-NULL, else p0,p1) -- calculateforcesharmonic
 *           (MD.cpp:2037-2111) when per-pair r0 are given
 *       2 = FENE (potential.h:345-376; p0=k, p1=R0)
 * F is accumulated (+=). */
void orc_pair_forces(const orc_box *b, int N, const double *x, const int *ij, long long np,
                     int kind, double p0, double p1, double p2,
                     const double *kk, const double *r0, double *F);

/* three-body loop MD::calculateforces_threebody (MD.cpp:1925-1983) with
 * form 0 = BendingPotential::force (Potentials/potential3.h:43-114),
 * form 1 = cosine bending U = kappa*(1 - cos theta) (no reference form).
 * k and theta0 per triplet (arrays) ; F accumulated. */
void orc_angle_forces(const orc_box *b, int N, const double *x, const int *ijk, long long nt,
                      int form, const double *k, const double *theta0, double *F);

/* Tangential active force (no 3-D reference): F_i += fa * unit(x_next - x_prev)
 * with prev/next = -1 meaning "self".  Accumulated. */
void orc_active_forces(const orc_box *b, int N, const double *x, const int *prev,
                       const int *next, double fa, double *F);

/* Patch set: GeneralPatch (Potentials/ComboPatch/GeneralPatch.cpp:7-121)
 * layout; TetrahedralPatch / BivalentPatch are single-type instances.
 * params: rows (eps, d, thetam) in GeneralPatch order: for ti<=tj, for a in
 * patches(ti), for b in patches(tj). */
#define ORC_MAXTYPES 8
typedef struct {
    int ntypes;
    const int *type;            /* per particle */
    int npatch[ORC_MAXTYPES];
    const double *dirs;         /* body-frame unit vectors, type-major, sum(npatch) x 3 */
    const double *params;       /* rows x 3 */
} orc_patchset;

/* KernFrenkelOnePatch2::force_and_torque (Potentials/potentialtheta.h:517-750)
 * un = unit(x_p1 - x_p2), rij = |x_p1 - x_p2|.  Q1,Q2 row-major 3x3 (lab->body).
 * out: fx fy fz tix tiy tiz tjx tjy tjz */
void orc_kf2_force_torque(const double un[3], double rij, const double *Q1, const double *Q2,
                          const double nb1[3], const double nb2[3], double att, double dis,
                          double thetam, double out[9]);
/* KernFrenkelOnePatch2::energy (potentialtheta.h:448-515) */
double orc_kf2_energy(const double un[3], double rij, const double *Q1, const double *Q2,
                      const double nb1[3], const double nb2[3], double att, double dis,
                      double thetam);

/* LangevinNVTR::calculate_forces_and_torques3D(pairs, ComboPatch, F, T)
 * (Rotational/LangevinR.cpp:911-1047).  F, T (lab frame) accumulated. */
void orc_patch_forces(const orc_box *b, int N, const double *x, const double *Q,
                      const int *ij, long long np, const orc_patchset *ps, double maxd2,
                      double *F, double *T);
/* sum of pair energies over a pair list (particle_energy, LangevinR.cpp:697-802,
 * without its extra r2>2.5 test: the maxd2 test of the force loop is used) */
double orc_patch_energy(const orc_box *b, int N, const double *x, const double *Q,
                        const int *ij, long long np, const orc_patchset *ps, double maxd2);

/* -------------------------------------------------------------- integrators */

typedef struct {
    double dt, m, gamma, gammar, kT;
    double I[3];                /* im[0], im[4], im[8] (LangevinR.cpp:44-47) */
} orc_params;

/* LangevinNVTR::create_forces_and_torques_sphere (LangevinR.cpp:475-592).
 * F lab (in/out), T lab in -> body out, RT N x 6. */
void orc_sphere(const orc_params *P, int N, const double *p, const double *L, const double *Q,
                double *F, double *T, const double *RT);
/* LangevinNVTR::advancemom_halfstep (LangevinR.cpp:327-347) */
void orc_halfkick(const orc_params *P, int N, double *p, double *L, const double *F,
                  const double *T);
/* LangevinNVT::advance_pos (Langevin.cpp:205-223) + cube::correct_position_and_momentum
 * (geometry.h:403-445) */
void orc_advance_pos(const orc_params *P, const orc_box *b, int N, double *x, double *p);
/* LangevinNVTR::genfullmat (LangevinR.cpp:392-461) */
void orc_genfullmat(const orc_params *P, const double *L3, double up[9]);
/* LangevinNVTR::rotate (LangevinR.cpp:595-695) */
void orc_rotate(const orc_params *P, int N, double *L, double *Q);
/* LangevinNVT::advance_mom (Langevin.cpp:254-269), coefficients Langevin.cpp:29-53 */
void orc_advance_mom(const orc_params *P, int N, double *p, const double *F, const double *R);

/* ------------------------------------------------------------- noise & RNG */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random numbers:
 * as easy as 1, 2, 3").  ctr[4], key[2] -> out[4]. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Shared injected-noise generator used by the parity trajectories: a pure
 * function of (seed, idx) with the reference's uniform transform
 * 3.464101615*u - 1.732050808 (DataStructures/matrix2.cpp:1807-1821), u from a
 * splitmix64 hash of seed and idx (53-bit mantissa). */
double orc_noise_uniform(uint64_t seed, uint64_t idx);
/* fills out[count] with orc_noise_uniform(seed, first + k) */
void orc_noise_fill(uint64_t seed, uint64_t first, long long count, double *out);

/* -------------------------------------------------------- step compositions */

/* Interaction description for the step compositions. */
typedef struct {
    /* WCA (HSPotential) on all listed pairs minus exclusions */
    int wca_on;
    double wca_eps, wca_sigma, wca_rcut;
    /* patches */
    const orc_patchset *patches;  /* NULL: none */
    double maxd2;
    /* bonds */
    const int *bonds; long long nbonds; int bond_kind;     /* 1 harmonic, 2 FENE */
    const double *bond_k, *bond_r0;
    /* angles */
    const int *angles; long long nangles; int angle_form;
    const double *angle_k, *angle_theta0;
    /* exclusions from WCA (pairs) */
    const int *excl; long long nexcl;
    /* active */
    double fa; const int *prev; const int *next;
    /* neighbour list: cut-off and rebuild period (Condensate.cpp:311,345) */
    double list_cut; int rebuild_every; int cells_per_side[3];
} orc_interactions;

/* Conservative forces F (lab), patch torques T (lab) on a frozen
 * configuration (F, T zeroed here).  Pair list built with list_cut. */
int orc_compute_forces(const orc_box *b, int N, const double *x, const double *Q,
                       const orc_interactions *it, double *F, double *T);

/* Condensate::run step order (Condensate.cpp:288-377) with injected noise:
 * noise holds (nsteps+1) x N x 6 values; row 0 primes F,T before the loop.
 * x,p,Q,L updated in place.  If snap_every>0, writes x after every
 * snap_every-th step into snaps (nsteps/snap_every x N x 3). */
int orc_run_rotors(const orc_params *P, const orc_box *b, int N, double *x, double *p,
                   double *Q, double *L, const orc_interactions *it, int nsteps,
                   const double *noise, int snap_every, double *snaps);

/* Bead step order of mainActivePolymer.cpp:381-426 (forces, advance_mom,
 * advance_pos) with injected noise nsteps x N x 3. */
int orc_run_beads(const orc_params *P, const orc_box *b, int N, double *x, double *p,
                  const orc_interactions *it, int nsteps, const double *noise,
                  int snap_every, double *snaps);

/* calcg1 (Analysis/PairCorrelation.cpp:9-155): counts[k] = number of pairs
 * i < j with floor(dis / dx) = k < floor(max / dx), dis = cube::distance
 * (geometry.h:107-123).  All pairs are visited (the reference visits the
 * pairs of stencil-adjacent cells of size >= max; the others have
 * dis >= max and fall outside the histogram).  Returns the bin count. */
int orc_calcg1(const orc_box *b, int N, const double *x, double max, double dx,
               long long *counts, int cap);

/* distance_graph (Analysis/AnalyzeGrowth.cpp:136-177): the pairs of
 * calculatepairs with n cells per side within cut (orc_pairs_reference),
 * kept when both indices lie outside [lo, hi] (lo > hi keeps all; the
 * excludea predicate, AnalyzeGrowth.cpp:123-134), and the label each particle
 * ends with in ConnectedComponentsParallel_Old
 * (Bindings/GraphAlgorithms.cpp:672-810).  That iteration starts from
 * ascending labels and only ever moves a root to a smaller label of an
 * adjacent vertex, so its fixed point is the smallest index of each connected
 * component; computed here by union-find with the smaller index as root.
 * Returns the number of pairs, or -1 when out of memory. */
long long orc_dgraph_labels(const orc_box *b, int N, const double *x, const int n[3], double cut,
                            int lo, int hi, int *labels);

/* MD::calculatestress (MD.cpp:2381-2424) over the pairs ij with the MD's
 * interaction an HSPotential(eps, sigma) (force2mdx, potential.h:126-139;
 * rcut <= 0: 1.12246 sigma): out[3k+j] = sum f(r^2) sqrt(r^2) d_k d_j with
 * d = x_i - x_j minimum-imaged (distance3v, geometry.h:250-276) */
void orc_stress(const orc_box *b, int N, const double *x, const int *ij, long long np,
                double eps, double sigma, double rcut, double out[9]);

/* Energy and virial of the conservative terms on a frozen configuration
 * (pairs built as orc_compute_forces builds them): U = WCA energy
 * (WCAPotential::energy, attrepsilon 0) on the listed pairs minus exclusions
 * + KernFrenkelOnePatch2::energy on the listed pairs within maxd2 + bond
 * energies (HarmonicPotential / FENEPotential::energy) + angle energies
 * (BendingPotential::energy, potential3.h:32-41, or k (1 - cos theta));
 * W[3k+l] = sum_i F_i,k x_i,l over the terms: pairs F_a,k (x_a - x_c)_l
 * minimum-imaged, triplets -F1,k ab_l + F3,k bc_l.  The active force is not
 * conservative and is left out.  Returns 0, or -1 when out of memory. */
int orc_energy_virial(const orc_box *b, int N, const double *x, const double *Q,
                      const orc_interactions *it, double *U, double W[9]);

/* measured_temperature (LangevinR.cpp:170-193): out[0..2] <p^2/2m>, out[3..5] <L^2/2I0> */
/* ------------------------------------------------ one-bond-per-patch binding */
/* LangevinNVTR::calculate_forces_and_torques3D_onlyone_nonlets
 * (Rotational/LangevinRforce.cpp:2874-4136) with BindingModelSingle
 * (Bindings/BindingModelSingle.cpp:1-119) on a frozen configuration: the
 * candidate patch pairs (both cones, a distance bound of 1, 0.7 or 0.5 x
 * the interaction distance by bound state, a score), the bonds whose partner
 * is no longer a candidate broken, the candidate graph's components (labelled
 * by their smallest patch), each component of 1-3 patches decided by the
 * model's doublet / triplet rule, larger ones cut by removing their
 * highest-scored edge until no part exceeds 3, and the Kern-Frenkel force
 * and torques of every bound pair added to F, T.  Patches are numbered per
 * particle in particle order, each particle's patches in type order
 * (GeneralPatch::which_patch, GeneralPatch.cpp:292-307, for particles sorted
 * by type).  Each patch's candidates are kept in ascending partner order (the
 * reference keeps its pair-list order; the outcome differs only where two
 * edges of a large cluster have exactly equal scores).
 * rng 0: the reference's draws -- srand(seed), one rand()/RAND_MAX per
 * component of 2 or more patches in ascending order of its smallest patch,
 * then a (large clusters) x (largest size / 2 + 1) matrix
 * (Generate_Random_Numbers_Needed, GraphAlgorithms.cpp:431-438;
 * generate_random_matrix, matrix2.cpp:1788-1805).  rng 1: Philox4x32-10,
 * counter {smallest patch, draw lo, draw hi, k} (k = 0 for a component of
 * 2-3 patches, 1 + i for the i-th decision inside a larger one), the 53-bit
 * uniform of orc_bind_uniform.
 * isbound, boundto: per patch, updated in place.  stats (5, or NULL):
 * candidate edges, components of 2+ patches, of 4+, the largest, bound
 * pairs.  Returns 0, or < 0 for a malformed three-patch component. */
typedef struct {
    double on_rate, off_rate;
    int rng;
    unsigned int seed;
    uint32_t key[2];
    uint64_t draw;
    /* table: 0 BindingModelSingle(on_rate, off_rate); 1 the species-table
     * form of BindingModelBinary / BindingModelTernary
     * (Bindings/BindingModelBinary.cpp:97-470, BindingModelTernary.cpp:
     * 492-870): species[patch] in 1..3; dtab[pair * 4 + 2 * before + k]
     * (k 0: rate to unbound, 1: to bound) by sorted species pair
     * 11,12,13,22,23,33; ttab[triple * 16 + 4 * from + to] (from: bound 12,
     * 23, 13, none; to: 12, 23, 13) and tr4[triple * 4 + from] (the rate to
     * none) by sorted species triple 111,112,113,122,123,133,222,223,233,333;
     * a triplet is decided in the frame of its sorted species
     * (sort_triplet_and_save_permutation, save_to_type, basic.h:166-468) */
    int table;
    const int *species;
    const double *dtab, *ttab, *tr4;
} orc_bind_model;
double orc_bind_uniform(const uint32_t key[2], uint64_t draw, int root, int k);
int orc_bind_forces(const orc_box *b, int N, const double *x, const double *Q,
                    const orc_patchset *ps, const orc_bind_model *m, int *isbound, int *boundto,
                    double *F, double *T, long long *stats);
void orc_measured_temperature(const orc_params *P, int N, const double *p, const double *L,
                              double out[6]);
/* LangevinNVTR::calculate_patch_list (Rotational/LangevinR.cpp:1702-1853) on
 * one thread: for each listed pair (sorted p1 < p2) closer than maxdis
 * (ComboPatch::max_check + 1), each patch pair of the pair's potentials whose
 * cone arguments satisfy (a_i - maxang) + (a_j - maxang) > -0.4 (maxang =
 * ComboPatch::max_ang: 0.6 lowered to the smallest cos theta_m,
 * combopatch.h:15, GeneralPatch.cpp:93-99) gives a row (p1, p2, potn, patch
 * of p1, patch of p2) (GeneralPatch::which_patch, GeneralPatch.cpp:297-320),
 * in pair order and ascending potn.  Particles must be numbered by type
 * (GeneralPatch's type ranges).  Returns the row count (rows beyond cap are
 * counted, not written), or -1 for a pair of descending types. */
long long orc_patch_list(const orc_box *b, int N, const double *x, const double *Q,
                         const int *ij, long long np, const orc_patchset *ps, double maxdis,
                         double maxang, int *out, long long cap);
/* LangevinNVTR::calculate_bound_pairs (Rotational/LangevinR.cpp:1854-1998):
 * as orc_patch_list with maxdis = max_check and the test a_i > cos theta_m
 * and a_j > cos theta_m of each row */
long long orc_bound_pairs(const orc_box *b, int N, const double *x, const double *Q,
                          const int *ij, long long np, const orc_patchset *ps, double maxdis,
                          int *out, long long cap);

#ifdef __cplusplus
}
#endif
#endif
