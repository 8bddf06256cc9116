/*
 * mdx.h -- C ABI of the MI355X (gfx950) Langevin step engine.
 *
 * One engine owns one HIP stream on one GPU and all device memory of a batch
 * of independent systems ("replicas"; a single system is a batch of one).
 * Host arrays use the reference's row-major FP64 layout
 * (DataStructures/matrix2.h:33-36): positions, momenta, forces, torques and
 * angular momenta N x 3, orientations N x 9 (lab -> body,
 * MDBase/Rotational/LangevinR.cpp:1-2).  Particles of replica r are stored
 * after those of replicas 0..r-1.  Host pointers are caller-owned and only
 * touched during the call.  Every function returns MDX_OK (0) or an error
 * code; mdx_last_error() gives the text (the reference's error() printed,
 * blocked on stdin and exited: Basic/basic.h:61-71).  Calls on one engine
 * must not overlap; different engines are independent.
 *
 * Reference paths cited below are relative to the reference's Code/ folder.
 */
#ifndef MDX_H
#define MDX_H

#ifdef __cplusplus
extern "C" {
#endif

#define MDX_OK 0
#define MDX_ERR_ARG 1       /* invalid argument or size */
#define MDX_ERR_HIP 2       /* HIP runtime failure */
#define MDX_ERR_STATE 3     /* call out of order (e.g. step before prime) */
#define MDX_ERR_FORCE 4     /* |f| of one pair term above the guard (MD.cpp:1711-1749) */
#define MDX_ERR_NAN 5       /* non-finite force (BasicGPU/basic.cu:14-27) */
#define MDX_ERR_CAPACITY 6  /* output buffer too small; the count is still returned */

#define MDX_PREC_FP64 0     /* everything in FP64 */
#define MDX_PREC_MIXED 1    /* FP32 pair math and p/Q/L; FP64 positions and decisions */

#define MDX_INTEG_ROTOR 0   /* LangevinNVTR step (Condensate.cpp:335-377) */
#define MDX_INTEG_BEAD 1    /* LangevinNVT step (mainActivePolymer.cpp:381-426) */

#define MDX_BOND_HARMONIC 1 /* HarmonicPotential / calculateforcesharmonic */
#define MDX_BOND_FENE 2     /* FENEPotential */

#define MDX_ANGLE_HARMONIC 0 /* BendingPotential: 0.5 k (theta - theta0)^2 */
#define MDX_ANGLE_COSINE 1   /* k (1 - cos theta) (no reference counterpart) */

typedef struct mdx_engine mdx_engine;

typedef struct mdx_desc {
    int nreplicas;            /* >= 1 */
    const int *n_per_replica; /* particle count of each replica */
    const double *box;        /* nreplicas x 3 box edges (a cube: equal edges; cube, geometry.h:43-66) */
    int periodic;             /* 1 periodic (cube pb), 0 reflecting walls (geometry.h:433-444) */
    const double *kT;         /* nreplicas temperatures (LangevinNVT::setkT) */
    double dt, m, gamma, gammar;  /* LangevinNVT::setdt/setm/setgamma, LangevinNVTR::setgammar */
    double inertia[3];        /* im[0], im[4], im[8] (LangevinR.cpp:44-47; 0.1 each) */
    int precision;            /* MDX_PREC_* */
    int integrator;           /* MDX_INTEG_* */
    unsigned long long seed;  /* Philox4x32-10 key for the engine's own noise */
    int device;               /* HIP device ordinal */
} mdx_desc;

/* Lifecycle.  Replaces constructing LangevinNVT(cube&) / LangevinNVTR(cube&)
 * (Langevin.cpp:34-52, LangevinR.cpp:41-65) plus the parameter setters
 * (LangevinR.h:109-137). */
int mdx_create(const mdx_desc *desc, mdx_engine **out);
void mdx_destroy(mdx_engine *e);
const char *mdx_last_error(void);

/* State.  Replaces MD::setdat (MD.cpp:64), LangevinNVT::setmom
 * (Langevin.cpp:170-174), LangevinNVTR::setorientation / setangularmomenta /
 * initialize (LangevinR.h:141-163, LangevinR.cpp:308-325); p, Q, L may be
 * NULL (zero momenta, identity orientation).  Upload drops the primed forces.
 * Download replaces getdat/getmom/getorientation/getangmom. */
int mdx_upload_state(mdx_engine *e, const double *x, const double *p, const double *Q,
                     const double *L);
int mdx_download_state(mdx_engine *e, double *x, double *p, double *Q, double *L);
/* total force (lab) and torque (body) carried between steps: the F and T of
 * Condensate::run after create_forces_and_torques_sphere (Condensate.cpp:370) */
int mdx_download_forces(mdx_engine *e, double *F, double *T);
/* Resume.  The reference continues a run from its newest CSV files with fresh
 * rand() noise (Condensate::run_singlebond_continue, Condensate.cpp:620-660;
 * Basic/Scripts/continue_run.sh:9-20).  Here the noise of force evaluation k
 * (k counts mdx_prime and every step since mdx_create) is a pure function of
 * the seed, k and the particle's original index, so a resumed run sets the
 * counter and continues the stream instead of repeating it.
 * mdx_get_noise_counter gives the k of the next force evaluation. */
int mdx_get_noise_counter(mdx_engine *e, unsigned long long *counter);
int mdx_set_noise_counter(mdx_engine *e, unsigned long long counter);
/* carried F (lab) and T (body) in the layout of mdx_download_forces, in place
 * of mdx_prime: after mdx_upload_state + mdx_set_noise_counter, a run
 * continued from downloaded state, forces and counter takes the same steps
 * as the uninterrupted run (FP64: bit for bit when the uninterrupted run
 * re-sorts its state at the step where the continuation starts, as the
 * continuation does at its first step; mixed precision: to FP32 rounding,
 * since its orientations are unit quaternions whose FP32 norm the N x 9
 * matrices do not carry) */
int mdx_upload_forces(mdx_engine *e, const double *F, const double *T);

/* Interactions. */
/* HSPotential(eps, sigma) as used by MD::calculateforces (potential.h:79-125);
 * rcut <= 0 selects the reference's literal 1.12246 sigma (potential.h:84). */
int mdx_set_pair_wca(mdx_engine *e, double eps, double sigma, double rcut);
/* GeneralPatch (Potentials/ComboPatch/GeneralPatch.cpp:7-121) with
 * KernFrenkelOnePatch2 rows (potentialtheta.h:379-797): ntypes <= 8,
 * type_of_particle N ints, npatch_per_type ntypes ints, dirs sum(npatch) x 3
 * body-frame unit vectors (type-major), params rows (eps, d, theta_m) in
 * GeneralPatch order (for ti <= tj, for patch a of ti, for patch b of tj),
 * maxd2 = LangevinNVTR::max_distance_check (LangevinR.h:70, default 2.4),
 * compared with r^2.  TetrahedralPatch / BivalentPatch are one-type cases. */
int mdx_set_patches(mdx_engine *e, int ntypes, const int *type_of_particle,
                    const int *npatch_per_type, const double *dirs, const double *params,
                    double maxd2);
/* per-replica patch rows (same layout as params above) */
int mdx_set_replica_patch_params(mdx_engine *e, int replica, const double *params);
/* bonds: MD::calculateforces(bonds, HarmonicPotential|FENEPotential)
 * (MD.cpp:1684-1764) and calculateforcesharmonic (MD.cpp:2037-2111): per
 * bond k and r0 (FENE: r0 is R_0). */
int mdx_set_bonds(mdx_engine *e, long long nbonds, const int *ij, const double *k,
                  const double *r0, int kind);
/* angles: MD::calculateforces_threebody(triplets, BendingPotential)
 * (MD.cpp:1925-1983, potential3.h:43-114), or the cosine form */
int mdx_set_angles(mdx_engine *e, long long nangles, const int *ijk, const double *k,
                   const double *theta0, int form);
/* pairs excluded from the WCA term (no reference counterpart: the reference
 * has no exclusion list) */
int mdx_set_exclusions(mdx_engine *e, long long npairs, const int *ij);
/* tangential active force fa * unit(x_next - x_prev), -1 = the particle itself
 * (no 3-D reference; cf. PositionForcesDueToAnglesGPUx, MicrotubuleGPU.cu:357-411) */
int mdx_set_active(mdx_engine *e, double fa, const int *prev, const int *next);
/* Neighbour list of mdx_step (the reference keeps a Verlet list of cut 3.5
 * rebuilt every 20 steps: Condensate.cpp:4, 345).  skin > 0: pairs within
 * reach + skin are listed at a build, and the list is rebuilt on the device
 * as soon as twice the summed per-step maximum drift reaches the skin, so the
 * pairs evaluated each step are exactly those within reach.  kmax list slots
 * per particle (at least 4; a particle with more walks its cells), state re-sorted into
 * cell order every resort_every steps.  skin 0: every step bins and walks
 * the 27 cells.  Defaults 0.4, 64, 200. */
int mdx_set_neighbor_list(mdx_engine *e, double skin, int kmax, int resort_every);
/* switch interaction terms off again (MDX_TERM_* bits); set them again with
 * the setters above */
#define MDX_TERM_WCA 1
#define MDX_TERM_PATCHES 2
#define MDX_TERM_BONDS 4
#define MDX_TERM_ANGLES 8
#define MDX_TERM_EXCLUSIONS 16
#define MDX_TERM_ACTIVE 32
#define MDX_TERM_ALL 63
int mdx_clear_interactions(mdx_engine *e, int terms);
/* Autotuner (no reference counterpart): after 4 x trial_steps warm-up steps
 * at the current settings (a system started from rest heats up for about
 * 1 / gamma), runs 2 x trial_steps steps with each
 * candidate neighbour-list skin (0.25, 0.4, 0.6, 0.9) and walk form
 * (compacted lanes, one lane per particle), timing the second half with HIP
 * events, and keeps the fastest.  The steps are real steps (the state
 * advances; rotors are primed again after each change).  report: 3 doubles
 * per candidate (skin, walk form 1/0, ms per step), cap candidates;
 * *nconfigs receives the number tried.  Rotors: after mdx_prime. */
int mdx_autotune(mdx_engine *e, int trial_steps, int cap, double *report, int *nconfigs);
/* |pair force| guard; 0 disables.  The reference prints "force too big" and
 * pauses above 1e4 (MD.cpp:1711-1749); here mdx_step returns MDX_ERR_FORCE. */
int mdx_set_force_guard(mdx_engine *e, double fmax);

/* Frozen-configuration entry points (parity). */
/* cube::assign_box with dim [1,n,n^2] (geometry.h:508-521, MD.cpp:1071-1078)
 * for replica r with n[3] cells per side; cell_of receives N_r ints */
int mdx_cell_assign(mdx_engine *e, int replica, const int n[3], int *cell_of);
/* MD::calculatepairs (MD.cpp:1035-1098, 575-713) as a set: every i<j of
 * replica r (local indices) with cube::distance_less_than(i, j, cut)
 * (geometry.h:145-164), sorted by (i, j).  n[3] cells per side, n >= 1. */
int mdx_neighbor_pairs(mdx_engine *e, int replica, const int n[3], double cut, int *ij,
                       long long cap, long long *npairs);
/* The Verlet list of the step (the reference's MD::calculatepairs(boxes, 3.5)
 * rebuilt every 20 steps, Condensate.cpp:311, 345) as a set: the list the
 * step's own kernels build at the current positions (pairs within reach +
 * skin; type pairs without patch rows within the WCA reach + skin), as
 * unordered pairs of original indices (i < j, ascending), at most cap;
 * radii (nullable, 2 doubles) receives the list radius and the radius of the
 * type pairs without patch rows (a pair is listed when r^2 is below the
 * square of its radius).  Checks that the full list holds each pair from
 * both sides.  The next mdx_step re-sorts and rebuilds its list. */
int mdx_list_pairs(mdx_engine *e, int *ij, long long cap, long long *npairs, double *radii);
/* conservative force F (lab) and patch torque T (lab) at the current state:
 * calculateforces(pairs, HSPotential) + bonds + angles + active +
 * calculate_forces_and_torques3D(pairs, ComboPatch, F, T) (LangevinR.cpp:911-1047) */
int mdx_compute_forces(mdx_engine *e, double *F, double *T);

/* Analysis of the current state (the scope's next row: observables). */
/* calcg1 (Analysis/PairCorrelation.cpp:9-155) for replica r: counts[k] =
 * number of particle pairs with floor(dis / dx) = k, k < floor(max / dx)
 * (*nbins; at most 4096), dis the minimum-image distance of cube::distance
 * (geometry.h:107-123), on floor(l / max) cells per side as the reference
 * builds them.  Bit-exact integer counts. */
int mdx_pair_histogram(mdx_engine *e, int replica, double max, double dx, long long *counts,
                       int cap, int *nbins);
/* patch bonds per particle (N ints, original order): partners with a patch
 * pair whose Kern-Frenkel term is non-zero (eps != 0, r <= 1.4 d, both cone
 * tests, potentialtheta.h:517-750, within max_distance_check), decided with
 * the force loop's own FP64 expressions.  No reference counterpart (the
 * reference's analysis counts bonds from its binding store). */
int mdx_patch_bonds(mdx_engine *e, int *bonds);
/* distance_graph (Analysis/AnalyzeGrowth.cpp:136-177) for replica r: the
 * pairs of MD::calculatepairs_parallel with n[3] cells per side within cut
 * (cube::distance_less_than), kept when both local indices lie outside
 * [ex_lo, ex_hi] (the excludea predicate, AnalyzeGrowth.cpp:123-134;
 * ex_lo > ex_hi keeps every pair, index_test), and labels[i] (N_r ints) = the
 * label ConnectedComponentsParallel_Old (Bindings/GraphAlgorithms.cpp:672-810)
 * converges to: the smallest local index of i's connected component.
 * *nclusters (nullable) = number of components.  Bit-exact. */
int mdx_clusters(mdx_engine *e, int replica, const int n[3], double cut, int ex_lo, int ex_hi,
                 int *labels, int *nclusters);

/* One-bond-per-patch binding (the reference's run_singlebond drivers,
 * Condensate.cpp:422-618).  Patches are numbered per particle in original
 * index order, each particle's patches in its type's order: the numbering of
 * GeneralPatch::which_patch (GeneralPatch.cpp:292-307) when particles are
 * sorted by type.  npatches (nullable) receives the number of patches.
 * BindingModelSingle(on_rate, off_rate) (Bindings/BindingModelSingle.cpp:1-8);
 * every patch starts unbound with boundto 0 (the BinaryBindStore of
 * Condensate.cpp:434-441).  Call after mdx_set_patches (which clears it). */
int mdx_set_binding(mdx_engine *e, double on_rate, double off_rate, int *npatches);
/* the species-table binding models in place of BindingModelSingle:
 * BindingModelBinary (Bindings/BindingModelBinary.cpp:97-470) and
 * BindingModelTernary (Bindings/BindingModelTernary.cpp:492-870).
 * species[nh]: 1..3 per patch (the model's rule: index < div, or its
 * sorting function); dtab[24]: per sorted species pair 11,12,13,22,23,33
 * the four rates doubrXY (before unbound: to unbound, to bound; before bound:
 * to unbound, to bound); ttab[160]: per sorted species triple
 * 111,112,113,122,123,133,222,223,233,333 the 16 rates triprXYZ (from bound
 * 12, 23, 13, none x to 12, 23, 13, none); tr4[40]: per triple and "from"
 * the rate to none the model reads (its own table's, except the binary
 * model's 122 and 222, which read 111's).  A triplet is decided in the
 * frame of its sorted species (sort_triplet_and_save_permutation,
 * save_to_type, basic.h:166-468).  After mdx_set_binding; mdx_set_binding
 * returns to BindingModelSingle. */
int mdx_set_binding_tables(mdx_engine *e, const int *species, const double *dtab,
                           const double *ttab, const double *tr4);
/* the bound state, per patch: isbound (0/1) and boundto (the partner patch;
 * kept, like the reference's, when the patch is unbound) --
 * BinaryBindStore (Bindings/BinaryBindStore.h:7-11) */
int mdx_set_bindings(mdx_engine *e, const int *isbound, const int *boundto);
int mdx_get_bindings(mdx_engine *e, int *isbound, int *boundto);
/* LangevinNVTR::calculate_forces_and_torques3D_onlyone_nonlets
 * (Rotational/LangevinRforce.cpp:2874-4136) at the current state: the
 * candidate patch pairs (both cones; a distance bound of 1, 0.7 or 0.5 x the
 * interaction distance by bound state; the reference's score), bonds to
 * non-candidates broken, each component of the candidate graph decided (1
 * patch: unbound; 2, 3: the model's doublet and triplet rules; more: the
 * highest-scored edge dropped until no part exceeds 3), the bound state
 * updated on the device, and F, T (N x 3 each, nullable, original order)
 * set to the Kern-Frenkel force and torques of the bound pairs alone (the
 * reference adds them to its F, T).  The random numbers are Philox4x32-10
 * keyed by the engine seed, at {smallest patch of the component, draw lo,
 * draw hi, k} (the reference draws rand() per component in the same order
 * of smallest patches).  MDX_ERR_CAPACITY when a patch has more than 8
 * candidates or a component more than 64 patches. */
int mdx_bind_forces(mdx_engine *e, unsigned long long draw, double *F, double *T);
/* LangevinNVTR::calculate_patch_list(pairs, ComboPatch&)
 * (Rotational/LangevinR.cpp:1702-1853): for each of the npairs particle
 * pairs (original indices, 2 per row) closer than max_check + 1, each patch
 * pair whose cone arguments satisfy (a_i - max_ang) + (a_j - max_ang) > -0.4
 * (max_check: 1 raised to the largest 1.4 d; max_ang: 0.6 lowered to the
 * smallest cos theta_m) gives a row of out (5 ints: p1 < p2, the potential
 * index in the patch-row order of mdx_set_patches, the patch of p1, the
 * patch of p2, patches numbered by particle as in mdx_set_bindings), in pair
 * order and ascending potential.  *count = the number of rows;
 * MDX_ERR_CAPACITY (and *count) when it exceeds cap.  A single rotor system
 * whose particles are numbered by type (GeneralPatch's type ranges). */
int mdx_patch_list(mdx_engine *e, const int *pairs, long long npairs, int *out, long long cap,
                   long long *count);
/* LangevinNVTR::calculate_bound_pairs(pairs, ComboPatch&)
 * (Rotational/LangevinR.cpp:1854-1998): as mdx_patch_list, with the pairs
 * closer than max_check and the patch pairs whose two cone arguments both
 * exceed their potential's cos(theta_m). */
int mdx_bound_pairs(mdx_engine *e, const int *pairs, long long npairs, int *out, long long cap,
                    long long *count);

/* Time stepping. */
/* Rotor integrator: computes the carried F, T at the current state
 * (Condensate.cpp:318-331).  noise0 (N x 6, nullable: Philox) */
int mdx_prime(mdx_engine *e, const double *noise0);
/* nsteps steps; rotor: the loop body of Condensate::run (Condensate.cpp:352-377),
 * bead: mainActivePolymer.cpp:390-425.  noise: nullable; else nsteps x N x 6
 * (rotor, row s is used by step s) or nsteps x N x 3 (bead) injected values in
 * place of generate_uniform_random_matrix (matrix2.cpp:1807-1821). */
int mdx_step(mdx_engine *e, int nsteps, const double *noise);
/* FP64 reductions of replica r (or -1: all): out[0..2] <p_k^2/2m>,
 * out[3..5] <L_k^2/2I_0> (LangevinR.cpp:170-193), out[6] potential energy,
 * out[7..15] virial sum_pairs r_ij (x) f_ij (calculatestress, MD.cpp:2381-2424) */
int mdx_observables(mdx_engine *e, int replica, double *out);

/* MD::calculatestress (MD.cpp:2381-2424) of replica r (or -1: all) with the
 * pair interaction of mdx_set_pair_wca as the MD's interaction
 * (HSPotential::force2mdx, potential.h:126-139): out[3k+j] = sum over pairs
 * within the cut-off of f(r^2) sqrt(r^2) d_k d_j, d the minimum-imaged
 * separation (the reference's own expression: r times the pair's virial
 * term; the virial itself is mdx_observables out[7..15]) */
int mdx_stress(mdx_engine *e, int replica, double *out);

/* mdx_stress without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same result as
 * mdx_stress (FP64 sums in another order) for any N: its kernels reduce into
 * a fixed static device scratch of 9 x 2049 doubles (147 KB) per GPU, held
 * until the process ends.  Needs the engine's cell grid built (after
 * mdx_prime, mdx_step, mdx_stress or mdx_observables), else MDX_ERR_STATE.
 * Calls on one GPU share that scratch: calls that overlap (from several
 * threads or engines on one GPU) race. */
int mdx_stress_noalloc(mdx_engine *e, int replica, double *out);

/* Single stages of the step, one reference call each, on the engine's state
 * (host arrays by original index, the reference's layouts).  For drivers that
 * compose the step themselves; afterwards mdx_step needs mdx_prime again. */
/* LangevinNVTR::advancemom_halfstep (LangevinR.cpp:327-347): p += dt/2 F,
 * L += dt/2 T (T ignored for the bead integrator) */
int mdx_stage_halfkick(mdx_engine *e, const double *F, const double *T);
/* LangevinNVT::advance_pos (Langevin.cpp:205-223) with
 * cube::correct_position_and_momentum (geometry.h:403-445) */
int mdx_stage_advance_pos(mdx_engine *e);
/* LangevinNVTR::rotate (LangevinR.cpp:595-695) */
int mdx_stage_rotate(mdx_engine *e);
/* LangevinNVTR::create_forces_and_torques_sphere (LangevinR.cpp:475-578):
 * F (lab) in/out; T lab torque in, body torque out; RT N x 6 */
int mdx_stage_sphere(mdx_engine *e, double *F, double *T, const double *RT);
/* LangevinNVT::advance_mom (Langevin.cpp:254-269): R N x 3 */
int mdx_stage_advance_mom(mdx_engine *e, const double *F, const double *R);

/* Diagnostics. */
/* per-kernel average device time (ms) over nsteps eagerly launched steps,
 * timed with HIP events on the engine's stream; names[k] are static strings.
 * Returns the kernel count in *nk (<= cap). */
int mdx_profile_step(mdx_engine *e, int nsteps, int cap, const char **names, double *ms,
                     int *nk);
/* engine facts (out holds 8): out[0] N, out[1] total cells, out[2] cell size
 * (replica 0, x), out[3] bytes of device memory held, out[4] steps taken,
 * out[5] neighbour-list builds so far, out[6] skin, out[7] list slots per
 * particle */
int mdx_info(mdx_engine *e, double *out);

#ifdef __cplusplus
}
#endif
#endif
