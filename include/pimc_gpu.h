/*
 * pimc_gpu.h -- C ABI of the MI355X action-evaluation library for PIMC++.
 *
 * This is the drop-in boundary for PIMC++'s action hot path: the short-range
 * pair action, the kinetic action and the Ewald long-range (rho_k) action that
 * the bisection stages evaluate once in OLDMODE and once in NEWMODE.  Every
 * entry point takes plain pointers and sizes; nothing here mentions a GPU
 * runtime type.  All calls return PIMC_OK or an error code; the message of the
 * last error on the calling thread is in pimc_last_error().  The caller owns
 * every host buffer.  A context is bound to one device and one stream and is
 * not reentrant: calls on one context must be serialised by the caller.
 *
 * Reference interfaces replaced (paths relative to the PIMC++ source tree):
 *   ActionBaseClass::SingleAction          PIMC++/src/Actions/ActionBase.h:43
 *   ShortRangeClass::SingleAction          PIMC++/src/Actions/ShortRangeClass.cc:131-176
 *   KineticClass::SingleAction             PIMC++/src/Actions/KineticClass.cc:33-90
 *   LongRangeClass::SingleAction           PIMC++/src/Actions/LongRangeClass.cc:273-371
 *   PairActionFitClass::U / DavidPAClass   Common/PairAction/PAFitBase.h:78,
 *                                          Common/PairAction/DavidPAClass.cc:92-102,468-511
 *   PathClass::DistDisp                    PIMC++/src/PathClass.h:696-746
 *   PathClass::UpdateRho_ks                PIMC++/src/PathClass.cc:649-723
 *   PathClass::AcceptCopy / RejectCopy     PIMC++/src/PathClass.cc:797-899
 *   PathClass::MoveJoin                    PIMC++/src/PathClass.cc:727-793
 *   BisectionBlockClass::MakeMove (NONE)   PIMC++/src/Moves/BisectionBlock.cc:380-451,
 *                                          PIMC++/src/Moves/StageClass.cc:49-95,
 *                                          PIMC++/src/Moves/BisectionStage.cc:128-289
 *
 * Conventions shared with the reference:
 *   - a replica's path holds M+1 time slices [0, M]; bead (slice, ptcl) is a
 *     3-vector of doubles; host buffers use the reference's layout
 *     [slice][ptcl][xyz] (blitz Array<dVec,2>, row-major);
 *   - particles of one species are contiguous and species are in order
 *     (SpeciesClass FirstPtcl/LastPtcl);
 *   - OLDMODE = 0, NEWMODE = 1 select the mirrored copy an action reads
 *     (MirroredClass ModeType, PIMC++/src/MirroredClass.h:23-39).
 */
#ifndef PIMC_GPU_H
#define PIMC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIMC_ABI_VERSION 2
#define PIMC_MAX_ACTIVE 4   /* particles moved together (PermuteType TABLE moves up to 4) */
#define PIMC_MAX_LEVELS 8   /* bisection levels */
#define PIMC_MAX_SPECIES 4

enum pimc_status {
  PIMC_OK = 0,
  PIMC_ERR_ARG = 1,         /* bad argument: the message names it */
  PIMC_ERR_STATE = 2,       /* object missing or in the wrong state */
  PIMC_ERR_DEVICE = 3,      /* the device runtime reported an error */
  PIMC_ERR_NOMEM = 4,       /* host or device allocation failed */
  PIMC_ERR_UNSUPPORTED = 5  /* valid request this build does not implement */
};

enum pimc_mode { PIMC_OLDMODE = 0, PIMC_NEWMODE = 1, PIMC_BOTHMODES = 2 };

/* Arithmetic of the spline evaluation and the rho_k phase sums.
 * PIMC_FP64: everything in FP64 (the correctness path).
 * PIMC_MIXED: pair geometry and images in FP64, spline and polynomial in FP32,
 * per-call sums in FP64 with Kahan compensation; rho_k totals, sum_k u_k|rho_k|^2
 * and the Metropolis difference stay FP64. */
enum pimc_precision { PIMC_FP64 = 0, PIMC_MIXED = 1 };

enum pimc_grid_type { PIMC_GRID_LINEAR = 0, PIMC_GRID_LOG = 1 };

typedef struct pimc_ctx pimc_ctx;

/* ------------------------------------------------------------------ */
/* Library and context                                                 */
/* ------------------------------------------------------------------ */

int pimc_abi_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* pimc_last_error(void);

/* Binds device `device`, creates the context's stream. */
int pimc_ctx_create(int32_t device, pimc_ctx** out);
int pimc_ctx_destroy(pimc_ctx* ctx);
/* Waits for all work queued on the context's stream. */
int pimc_ctx_synchronize(pimc_ctx* ctx);

/* ------------------------------------------------------------------ */
/* Pair-action fits (PairActionFitClass)                               */
/* ------------------------------------------------------------------ */

/* A DavidPAClass fit: one MultiCubicSpline per tau level over an r grid, with
 * column 0 the potential, column 1 u00 and columns 2.. the off-diagonal
 * u_kj(q) in the order k(k+1)/2+j+1 (DavidPAClass.cc:440-456).  Splines are
 * built as the reference's .dm reader builds them (DavidPAClass.cc:755-773):
 * natural second derivative at the start, first derivative 0 at the end.
 * tau_i = smallest_tau * 2^i; TauPos is the i with |tau_i - desired_tau| < 1e-4
 * (DavidPAClass.cc:859-868).  The LOG grid is r_i = start * spacing^i with
 * spacing = (end/start)^(1/(n-1)) (DavidPAClass.cc:714-719, Grid.h:537-547). */
typedef struct pimc_david_fit_desc {
  int32_t grid_type;       /* pimc_grid_type */
  int32_t num_points;      /* r-grid points, >= 3 */
  double grid_start;
  double grid_end;
  int32_t num_tau;         /* tau levels stored */
  int32_t num_cols;        /* NumUKJ + 1 */
  int32_t num_terms;       /* NumOffDiagonalTerms: needs num_cols >= (n+1)(n+2)/2 + 1 */
  int32_t reserved;
  double smallest_tau;
  double desired_tau;
  const double* y;         /* [num_tau][num_points][num_cols] */
  const double* dy;        /* the beta-derivative table dukj in the same layout, or NULL
                              (needed by pimc_estimator_energy only; DavidPAClass.cc:840-877) */
} pimc_david_fit_desc;

/* Creates a fit; replaces ReadPAFit + DavidPAClass::Read for table data. */
int pimc_fit_create_david(pimc_ctx* ctx, const pimc_david_fit_desc* desc, int32_t* fit_id);

/* Batched PairActionFitClass::U(q, z, s2, level) for n points (DavidPAClass::U).
 * level is the bisection level; the fit adds its TauPos. */
int pimc_fit_eval_u(pimc_ctx* ctx, int32_t fit_id, int32_t level, int64_t n,
                    const double* q, const double* z, const double* s2,
                    int32_t precision, double* out);

/* A PAcoulombBCFitClass fit (Common/PairAction/PAcoulombBCFit.cc:159-176,
 * 242-311): per level (beta_l = smallest_beta 2^l) a BicubicSpline of U over
 * (q, t = s/2q), z-independent; beyond the q grid U = beta_l Z1Z2 / q.  The
 * bicubic's node derivatives are built as BicubicSpline::Init does
 * (Common/Splines/BicubicSpline.cc:20-154).  The fit is long-ranged: when the
 * replica uses long range, ShortRange subtracts (Ulong(r) + Ulong(r'))/2
 * (ShortRangeClass.cc:165-166) with the replica's Ulong tables. */
typedef struct pimc_coulomb_bc_fit_desc {
  int32_t q_grid_type;     /* pimc_grid_type */
  int32_t q_points;
  double q_start;
  double q_end;
  int32_t t_grid_type;     /* pimc_grid_type */
  int32_t t_points;
  double t_start;
  double t_end;
  int32_t num_betas;       /* levels stored */
  int32_t reserved;
  double smallest_beta;    /* the level-0 tau */
  double z1z2;             /* Coulomb potential Z1 Z2 / r beyond the grid */
  const double* umat;      /* [num_betas][q_points][t_points] */
  const double* dumat;     /* dU/dbeta in the same layout (dUsplines), or NULL */
} pimc_coulomb_bc_fit_desc;

int pimc_fit_create_coulomb_bc(pimc_ctx* ctx, const pimc_coulomb_bc_fit_desc* desc, int32_t* fit_id);

/* ------------------------------------------------------------------ */
/* Replicas (PathClass + the actions' configuration)                   */
/* ------------------------------------------------------------------ */

typedef struct pimc_replica_desc {
  int32_t num_particles;
  int32_t num_time_slices;   /* M: the path holds M+1 slices (PathClass.cc:305,410) */
  int32_t num_species;
  int32_t num_images;        /* KineticClass NumImages */
  double box[3];
  int32_t periodic[3];       /* 1 = periodic in that dimension */
  double tau;
  const int32_t* species_count;  /* [num_species] */
  const double* lambda;          /* [num_species] hbar^2/2m */
  const int32_t* pair_fit;       /* [num_species*num_species], PairMatrix(s1,s2) */
  double k_cutoff;               /* > 0: build the k set and keep rho_k (PathClass.cc:534-586) */
  int32_t use_long_range;        /* ActionsClass UseLongRange */
  int32_t use_background;        /* LongRangeClass UseBackground */
  uint32_t rng_stream;           /* this replica's random stream (a clone's Local stream) */
  int32_t reserved;
} pimc_replica_desc;

int pimc_replica_create(pimc_ctx* ctx, const pimc_replica_desc* desc, int32_t* replica_id);
/* Number of k vectors of the replica's k set (0 if k_cutoff <= 0). */
int pimc_replica_num_kvecs(pimc_ctx* ctx, int32_t replica, int32_t* nk);
/* The k set: kvecs [nk][3] (Cartesian), kidx [nk][3] (integer index + MaxkIndex). Either may be NULL. */
int pimc_replica_kvecs(pimc_ctx* ctx, int32_t replica, double* kvecs, int32_t* kidx);

/* Long-range tables of species pair (sa, sb) on this replica's k set
 * (PairActionFitClass Ulong_k, Ulong_r0, Ushort_k0, Ulong_k0, PAFitBase.h:38-46,
 * and their beta derivatives dUlong_k, dUlong_r0, dUshort_k0, dUlong_k0).
 * ulong_k and dulong_k are [num_levels][nk]; the scalars are [num_levels].
 * The four derivative tables may be NULL together (needed by the energy estimator only). */
int pimc_replica_set_long_range(pimc_ctx* ctx, int32_t replica, int32_t sa, int32_t sb,
                                int32_t num_levels, const double* ulong_k,
                                const double* ulong_r0, const double* ushort_k0,
                                const double* ulong_k0, const double* dulong_k,
                                const double* dulong_r0, const double* dushort_k0,
                                const double* dulong_k0);

/* Ulong(level)(r) of species pair (sa, sb) (PairActionFitClass::Ulong, PAFitBase.h:38): a
 * QuinticSpline per level on LinearGrid(grid_start, grid_end, num_points), given as the
 * polynomial coefficients the reference evaluates (QuinticSpline.h:204-218):
 * coefs [num_levels][num_points][6] = {Y, B, C, D, E, F} of each knot,
 * S(x) = ((((F P + E) P + D) P + C) P + B) P + Y with P = x - X_I. */
int pimc_replica_set_ulong(pimc_ctx* ctx, int32_t replica, int32_t sa, int32_t sb, int32_t num_levels,
                           double grid_start, double grid_end, int32_t num_points, const double* coefs,
                           const double* dcoefs);
/* dcoefs: dUlong(level)(r) in the same layout, or NULL (needed by the energy estimator only). */

/* The long-range part of the pair potential of (sa, sb), for the potential-energy
 * estimator: PairActionFitClass's Vlong (a QuinticSpline on the grid of the pair's Ulong
 * tables, coefs [num_points][6] as above), Vlong_k [nk], Vlong_r0, Vshort_k0 and Vlong_k0
 * (PAFitBase.h:39-53), as LongRangeClass::OptimizedBreakup_V builds them
 * (PIMC++/src/Actions/LongRangeClass.cc:979-1136).  grid_start, grid_end and num_points
 * must match the replica's Ulong grid when one is set. */
int pimc_replica_set_vlong(pimc_ctx* ctx, int32_t replica, int32_t sa, int32_t sb, const double* vlong_k,
                           double vlong_r0, double vshort_k0, double vlong_k0, double grid_start,
                           double grid_end, int32_t num_points, const double* coefs);

/* Path data. mode: PIMC_OLDMODE, PIMC_NEWMODE or PIMC_BOTHMODES (set only).
 * beads is [(M+1)][N][3]. */
int pimc_path_set(pimc_ctx* ctx, int32_t replica, int32_t mode, const double* beads);
int pimc_path_get(pimc_ctx* ctx, int32_t replica, int32_t mode, double* beads);
/* Writes beads of one particle for slices [slice1, slice2] (Path.SetPos). xyz is [(slice2-slice1+1)][3]. */
int pimc_path_set_beads(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t ptcl,
                        int32_t slice1, int32_t slice2, const double* xyz);
/* Permutation vector of a mirrored copy, [N]. */
int pimc_permutation_get(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t* perm);

/* PathClass::AcceptCopy / RejectCopy for the listed particles over slices [slice1, slice2]. */
int pimc_accept_copy(pimc_ctx* ctx, int32_t replica, int32_t slice1, int32_t slice2,
                     const int32_t* ptcls, int32_t nptcls);
int pimc_reject_copy(pimc_ctx* ctx, int32_t replica, int32_t slice1, int32_t slice2,
                     const int32_t* ptcls, int32_t nptcls);
/* PathClass::MoveJoin(old_join, new_join) (PIMC++/src/PathClass.cc:727-793): for new_join >
 * old_join the slices (old_join, new_join] of copy OLD take, at particle p, the NEW bead of
 * particle Permutation(p); for new_join < old_join the slices (new_join, old_join] of OLD take,
 * at particle Permutation(p), the NEW bead of p; then NEW <- OLD over those slices.  The
 * replica's join becomes new_join (PathDataClass::MoveJoin, PIMC++/src/PathDataClass.h:106-111). */
int pimc_move_join(pimc_ctx* ctx, int32_t replica, int32_t old_join, int32_t new_join);
/* The replica's current join (PathDataClass::Join): 0 at creation, new_join after
 * pimc_move_join, and slice2 of the last move after pimc_bisection_run, which moves the join
 * at the start of every move as BisectionBlockClass::MakeMove does
 * (PIMC++/src/Moves/BisectionBlock.cc:423-424). */
int pimc_replica_join(pimc_ctx* ctx, int32_t replica, int32_t* join);

/* PermuteTableClass::ConstructHTable (PIMC++/src/Moves/PermuteTableClass.cc:122-150)
 * for the particles of `species`, read from copy `mode`:
 *   HTable(i,j) = exp((|PutInBox(r_i(s2) - r_i(s1))|^2 - |PutInBox(r_j(s2) - r_i(s1))|^2)
 *                     / (4 lambda tau (s2 - s1)))
 * stored as float, or 0 where the exponent is below log(epsilon).  htable is
 * [n][n] row-major (i = row), n = the species' particle count; i and j are
 * indices within the species. */
int pimc_permute_htable(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t species,
                        int32_t slice1, int32_t slice2, double epsilon, float* htable);
/* CycleClass::Apply (PIMC++/src/Moves/PermuteTableClass.cc:172-187) on the NEW copy
 * at one slice: particle first_ptcl + cycle[i] takes the bead and permutation entry
 * of first_ptcl + cycle[i+1], and the last takes those of first_ptcl + cycle[0].
 * length 1..4, entries distinct. */
int pimc_cycle_apply(pimc_ctx* ctx, int32_t replica, int32_t first_ptcl, int32_t slice,
                     int32_t length, const int32_t* cycle);
/* PermuteType TABLE for the bisection moves of `species` on this replica, with the cycle
 * weights gamma[0..3] (Gamma of 1- to 4-cycles) and the cut-off epsilon
 * (TablePermuteStageClass::Read, PIMC++/src/Moves/TablePermuteStage.cc:142-166).  Each
 * pimc_bisection_run move of that species then runs the stage list of
 * BisectionBlockClass with TABLE (PIMC++/src/Moves/BisectionBlock.cc:51-64,126): the
 * HTable of the window (PermuteTableClass::ConstructHTable), the boson cycle table
 * (ConstructBosonCycleTable, PermuteTableClass.cc:339-449, no excluded particle, no z
 * focus), a cycle drawn from it and applied at slice2 (AttemptPermutation), the bisection
 * stages moving every particle of the cycle, and the final permutation test with the
 * reverse table (TablePermuteStageClass::Attempt, TablePermuteStage.cc:200-260,
 * CalcReverseProb, UpdateHTable).  gamma == NULL switches the replica back to no
 * permutations.  TABLE moves run 256-thread blocks, each taking replicas in turn, whatever
 * the plan's block size and persistence; of the plan they use the link split.  Replicas with
 * the long-range action, cluster plans, num_levels above 5, a moved species whose n x n float
 * HTable and the staged spline level exceed a block's 64 KB of LDS (about 100 particles), and
 * sweep graphs refuse TABLE moves (PIMC_ERR_UNSUPPORTED); the autotuner only tries their
 * link split.  Every replica of one
 * call moves with TABLE or none does.  The GPU builds the cycle table in the reference's
 * order in LDS or, when larger, in device scratch of 65,536 entries per block; a larger table
 * is walked in that order without being stored, so its size is not bounded. */
int pimc_replica_set_permute(pimc_ctx* ctx, int32_t replica, int32_t species, const double* gamma,
                             double epsilon);

/* PathClass::UpdateRho_ks(): recompute rho_k of every slice and species from
 * the OLD copy, then NEW <- OLD (PathClass.cc:709-723). */
int pimc_rho_k_update_all(pimc_ctx* ctx, int32_t replica);
/* rho_k of a mirrored copy as [(M+1)][num_species][nk][re,im]. */
int pimc_rho_k_get(pimc_ctx* ctx, int32_t replica, int32_t mode, double* rho);

/* ------------------------------------------------------------------ */
/* Actions (ActionBaseClass::SingleAction)                              */
/* ------------------------------------------------------------------ */

/* One SingleAction(slice1, slice2, activeParticles, level) call evaluated on
 * copy `mode` of `replica`. */
typedef struct pimc_action_req {
  int32_t replica;
  int32_t mode;              /* PIMC_OLDMODE or PIMC_NEWMODE */
  int32_t slice1;
  int32_t slice2;
  int32_t level;
  int32_t nptcls;            /* 1..PIMC_MAX_ACTIVE */
  int32_t ptcls[PIMC_MAX_ACTIVE];
} pimc_action_req;

/* ShortRangeClass::SingleAction for each request; out[i] is request i's total. */
int pimc_action_short_range(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq,
                            int32_t precision, double* out);
/* pimc_action_short_range without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, hosts that forbid hipMalloc).
 * No counterpart in the reference.  Same requests, arithmetic and results as
 * pimc_action_short_range (the same kernels, bit-identical outputs); its request staging,
 * per-tile partial sums and outputs live in a fixed device array of
 * PIMC_NOALLOC_SCRATCH_BYTES per GPU, which the library holds until the process ends.
 * Limits: nreq <= PIMC_NOALLOC_MAX_REQS and nreq * ceil(N_max / 256) * 8 + nreq * 48 bytes
 * within the array (N_max: the largest particle count of the requested replicas); larger
 * calls return PIMC_ERR_ARG.  Fits and replicas created or changed since the last call of
 * another action entry point must be uploaded by such a call first (PIMC_ERR_STATE
 * otherwise).  Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (two contexts on different host threads) race. */
#define PIMC_NOALLOC_SCRATCH_BYTES (4u << 20)
#define PIMC_NOALLOC_MAX_REQS 1024
int pimc_action_short_range_noalloc(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq,
                                    int32_t precision, double* out);
/* KineticClass::SingleAction. */
int pimc_action_kinetic(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq, double* out);
/* LongRangeClass::SingleAction (PIMC++/src/Actions/LongRangeClass.cc:273-371): a NEWMODE
 * request first updates rho_k of copy NEW over [slice1, slice2] step 2^level from OLD
 * (PathClass::UpdateRho_ks, PIMC++/src/PathClass.cc:649-707).  Requests are processed in
 * order; requests on distinct replicas share a launch.  A NEWMODE request whose k set is too
 * large for its moved particles' phase rows (2 nptcls rows of 2(maxk0+maxk1+maxk2)+3
 * entries of 16 B beyond 24 KiB) returns PIMC_ERR_UNSUPPORTED. */
int pimc_action_long_range(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq,
                           int32_t precision, double* out);

/* PathClass::DistDisp(sliceA, sliceB, p1, p2, ...) for n index quadruples
 * idx[n][4] = {sliceA, sliceB, p1, p2} on copy `mode`.  out[n][8] =
 * {distA, distB, dispA[3], dispB[3]}; images[n][6] = {n[3], m[3]} as the
 * reference computes them (n = -nearbyint(dispA/L), m = nearbyint((dispA-dispB)/L)). */
int pimc_dist_disp(pimc_ctx* ctx, int32_t replica, int32_t mode, int64_t n,
                   const int32_t* idx, double* out, int32_t* images);

/* PathClass::PutInBox (PIMC++/src/PathClass.h:897-921): v_d += n L_d IsPeriodic_d with
 * n = -floor(v_d / L_d + 1/2), so a displacement of exactly half a box goes to -L/2 (ties go
 * up, unlike DistDisp's nearbyint); or, with fast = 1, PathClass::PutInBoxFast
 * (PIMC++/src/PathClass.cc:1389-1407): subtract L where v > L/2, then add L where v < -L/2.
 * v is [n][3], with the replica's box, replaced in place. */
int pimc_put_in_box(pimc_ctx* ctx, int32_t replica, int32_t fast, int64_t n, double* v);

/* ------------------------------------------------------------------ */
/* Estimators (ActionsClass::Energy, PairCorrelationClass)             */
/* ------------------------------------------------------------------ */

/* ActionsClass::Energy (PIMC++/src/Actions/ActionsClass.cc:813-871) of copy `mode`
 * over the whole path (slices [0, M], level 0):
 *   kinetic   KineticClass::d_dBeta        (PIMC++/src/Actions/KineticClass.cc:187-256)
 *   du_short  ShortRangeClass::d_dBeta     (PIMC++/src/Actions/ShortRangeClass.cc:256-295;
 *             DavidPAClass::dU via calcUsqz, DavidPAClass.cc:115-126, 291-463;
 *             PAcoulombBCFitClass::dU, PAcoulombBCFit.cc:178-192)
 *   du_long   LongRangeClass::d_dBeta      (PIMC++/src/Actions/LongRangeClass.cc:374-442), 0 without long range
 *   v_short   sum over slices (1/2 at the ends) of ShortRangePotClass::V
 *             (PIMC++/src/Actions/ShortRangePotClass.cc:27-47): V(r) of each pair (the fit's
 *             column 0 for DavidPA fits, Z1Z2/r for Coulomb fits), minus Vlong(r) for a
 *             long-ranged fit when long range is used; NaN when such a pair has no Vlong tables
 *   v_long    the same sum of LongRangePotClass::V (PIMC++/src/Actions/LongRangePotClass.cc:32-85);
 *             0 without long range, NaN when a long-ranged pair has no Vlong tables
 *   total     kinetic + du_short + du_long (EnergyClass::Accumulate's sum without node terms,
 *             PIMC++/src/Observables/ObservableEnergy.cc:31-60)
 * Needs the fits' beta-derivative tables (and the long-range ones when long range is used). */
typedef struct pimc_energy {
  double kinetic;
  double du_short;
  double du_long;
  double v_short;
  double v_long;
  double total;
} pimc_energy;

int pimc_estimator_energy(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, int32_t mode, pimc_energy* out);

/* ActionBaseClass::d_dBeta(slice1, slice2, level) (PIMC++/src/Actions/ActionBase.h:57) of one
 * action on copy `mode` of a replica, for each request:
 *   PIMC_ACTION_SHORT_RANGE  ShortRangeClass::d_dBeta (PIMC++/src/Actions/ShortRangeClass.cc:256-295):
 *                            every pair over links [slice1, slice2) at `level`, the fits' dU minus
 *                            dUlong of long-ranged pairs when long range is used
 *   PIMC_ACTION_KINETIC      KineticClass::d_dBeta (PIMC++/src/Actions/KineticClass.cc:187-256)
 *   PIMC_ACTION_LONG_RANGE   LongRangeClass::d_dBeta (PIMC++/src/Actions/LongRangeClass.cc:374-442):
 *                            slices slice1 .. slice2 step 2^level (ends weighted 1/2), with rho_k of
 *                            those slices recomputed from the positions as CalcRho_ks_Fast does (the
 *                            stored rho_k is left as it is)
 * The short- and long-range parts need the beta-derivative tables (pimc_estimator_energy). */
enum pimc_action_kind { PIMC_ACTION_SHORT_RANGE = 0, PIMC_ACTION_KINETIC = 1, PIMC_ACTION_LONG_RANGE = 2 };
typedef struct pimc_range_req {
  int32_t replica;
  int32_t mode;              /* PIMC_OLDMODE or PIMC_NEWMODE */
  int32_t slice1;
  int32_t slice2;
  int32_t level;
  int32_t reserved;
} pimc_range_req;
int pimc_action_d_dbeta(pimc_ctx* ctx, int32_t action, const pimc_range_req* reqs, int32_t nreq, double* out);

/* PairCorrelationClass::Accumulate (PIMC++/src/Observables/ObservableCorrelation.cc:156-221):
 * adds to hist[num_points-1] the count of minimum-image distances between
 * particles of species s1 and s2 (pairs counted once when s1 == s2) on slices
 * [0, M-1] of copy `mode` that fall below `end`, binned by the ReverseMap of
 * LinearGrid(start, end, num_points).  Distances below `start` are not counted. */
int pimc_estimator_pair_correlation(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t s1, int32_t s2,
                                    double start, double end, int32_t num_points, int64_t* hist);

/* ------------------------------------------------------------------ */
/* Bisection moves (BisectionBlockClass with PermuteType NONE)          */
/* ------------------------------------------------------------------ */

/* Stage actions: Kinetic + ShortRange at every level, + LongRange at the lowest
 * level when the replica uses long range (BisectionBlock.cc:66-104 with
 * HigherLevelActions = SamplingActions = ShortRange). */
typedef struct pimc_bisection_desc {
  int32_t num_levels;        /* NumLevels: the window is 2^num_levels slices */
  int32_t lowest_level;      /* LowestLevel */
  int32_t species;           /* species whose particles are moved */
  int32_t reserved;
  uint64_t seed;             /* random seed (the clone's Local stream key) */
} pimc_bisection_desc;

/* What one move did. Stage s is bisection level num_levels-1-s. */
typedef struct pimc_move_record {
  int32_t slice1;
  int32_t ptcl;              /* the moved particle (with TABLE permutations: the cycle's first) */
  int32_t accepted;          /* 1 if every stage accepted and the move was kept */
  int32_t num_stages;        /* bisection stages attempted (the first rejection ends the move) */
  double old_action[PIMC_MAX_LEVELS];
  double new_action[PIMC_MAX_LEVELS];
  double log_sample_ratio[PIMC_MAX_LEVELS];
  double log_u[PIMC_MAX_LEVELS];
  /* TABLE permutation moves (pimc_replica_set_permute); zero otherwise.  cycle holds the
   * chosen cycle's particles (absolute indices, cycle_length of them); forward_t is the
   * probability of choosing it (AttemptPermutation), reverse_t that of the reverse choice
   * (CalcReverseProb), accept_prob and psi the final permutation test (0 where the move
   * ended at a bisection stage). */
  int32_t cycle_length;
  int32_t cycle[PIMC_MAX_ACTIVE];
  int32_t num_cycles;        /* entries of the forward cycle table */
  double forward_t;
  double reverse_t;
  double accept_prob;
  double psi;
} pimc_move_record;

/* Counters of a pimc_bisection_run call, summed over replicas and moves. */
typedef struct pimc_bisection_stats {
  int64_t moves;
  int64_t accepted;
  int64_t stages;
  int64_t pair_evals;        /* pair.slice evaluations: one U(q,z,s2) per (moved, partner, link, mode) */
  int64_t k_evals;           /* k.slice.species rho_k terms evaluated by the long-range action */
  double device_ms;          /* time of the call's compute on the context's stream (device events) */
} pimc_bisection_stats;

/* Runs moves [first_move, first_move + num_moves) on every listed replica.
 * Each replica's moves are a Metropolis chain in move order; replicas are
 * independent.  records (may be NULL) is [nrep][num_moves].  The random
 * numbers of move m of a replica are a pure function of (seed, rng_stream, m),
 * so a chain is reproducible across calls and backends.
 * With records and stats both NULL the call is asynchronous: it returns once the
 * work is queued on the context's stream, and its counters are read with
 * pimc_bisection_stats_take.  Otherwise it returns when the moves are done.
 * Every move first moves the replica's join to its slice2 (PathDataClass::MoveJoin,
 * BisectionBlock.cc:424); with an identity permutation that changes no bead.  A replica
 * whose permutation is not the identity cannot run under a cluster plan
 * (PIMC_ERR_UNSUPPORTED).
 * Replaces the per-move driver BisectionBlockClass::MakeMove
 * (PIMC++/src/Moves/BisectionBlock.cc:310-451) with NoPermuteStageClass
 * (PIMC++/src/Moves/NoPermuteClass.cc:26-68) and the stage loop of
 * MultiStageClass::MakeMove (PIMC++/src/Moves/MultiStage.cc:42-110). */
int pimc_bisection_run(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep,
                       const pimc_bisection_desc* desc, int64_t first_move, int32_t num_moves,
                       int32_t precision, pimc_move_record* records,
                       pimc_bisection_stats* stats);

/* Launch shape of the resident bisection kernel (the context's current plan is
 * used by pimc_bisection_run and captured by pimc_sweep_graph_create). */
typedef struct pimc_bisection_plan {
  int32_t block_threads;     /* threads per replica: 256 or 512 */
  int32_t link_split;        /* threads sharing one partner's links in the short-range loop: 1, 2 or 4 */
  int32_t persistent;        /* 1: a persistent grid takes replicas from a work queue, longest first */
  int32_t cluster;           /* blocks per replica, 1..32 (0 means 1); above 1 the blocks split the
                                partners and k-vectors of every stage and exchange partial sums; the
                                nrep * cluster blocks must all be resident at once; not persistent */
  double trial_ms;           /* pimc_bisection_tune: time of the chosen shape on the trial moves */
} pimc_bisection_plan;

/* The default plan is {256, 1, 0, 1}. */
int pimc_bisection_set_plan(pimc_ctx* ctx, const pimc_bisection_plan* plan);
int pimc_bisection_get_plan(pimc_ctx* ctx, pimc_bisection_plan* plan);

/* Autotuner: times every plan (block_threads x link_split x {one block per
 * replica, persistent, clusters of 2..32 blocks that fit the device at once}) on
 * trial_moves moves [first_move, first_move + trial_moves) of the listed
 * replicas, restores their paths, permutations, rho_k and counters after each
 * trial, and makes the fastest plan the context's plan (also returned in *out,
 * which may be NULL).  TABLE permutation moves run one block of 256 threads per
 * replica, so for them only the link split (1, 2, 4) is tried.  Has no
 * counterpart in the reference, whose CPU loops have no launch shape. */
int pimc_bisection_tune(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep,
                        const pimc_bisection_desc* desc, int64_t first_move, int32_t trial_moves,
                        int32_t precision, pimc_bisection_plan* out);

/* Sweep graphs: one sweep (num_moves moves of every listed replica with the
 * context's plan) captured once as a device graph of kernels; every replay
 * continues the chains, sweep i running moves first_move + i*num_moves onward.
 * Launches are asynchronous; counters are read with pimc_bisection_stats_take.
 * A graph reads the replicas' device state directly: do not change their
 * paths with other calls while a graph's sweeps are queued. */
int pimc_sweep_graph_create(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep,
                            const pimc_bisection_desc* desc, int64_t first_move, int32_t num_moves,
                            int32_t precision, int32_t* graph_id);
int pimc_sweep_graph_launch(pimc_ctx* ctx, int32_t graph_id, int32_t num_sweeps);
int pimc_sweep_graph_destroy(pimc_ctx* ctx, int32_t graph_id);

/* Waits for the context's stream, then stores in *stats the counters every
 * pimc_bisection_run call since the previous take added for the listed
 * replicas, and clears those replicas' counters.  device_ms is the compute time
 * (device events) of all pimc_bisection_run calls, sweep-graph replays and
 * pimc_action_long_range calls on the context since the previous take.
 * (The reference keeps such counters per move object, MoveBaseClass::NumAccepted
 * and TimesCalled, PIMC++/src/Moves/MoveBase.h:70-77, and the time per action,
 * ActionBaseClass::TimeSpent, PIMC++/src/Actions/ActionBase.h:37.) */
int pimc_bisection_stats_take(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep,
                              pimc_bisection_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* PIMC_GPU_H */
