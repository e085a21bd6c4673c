/*
 * sdm_hip.h -- C ABI of the MI355X (gfx950) SDM alchemical force + Langevin step.
 *
 * The reference plugin (rajatkrpal/openmm_sdm_plugin) exposes one OpenMM platform
 * kernel, IntegrateLangevinStepSDMKernel ("IntegrateLangevinStepSDM",
 * openmmapi/include/SDMKernels.h:47-113), created by a platform kernel factory that
 * OpenMM reaches through the plugin's exported registerKernelFactories()
 * (platforms/reference/src/ReferenceSDMKernelFactory.cpp:41-63).  Its step() drives
 * two NonbondedForce evaluations (x and x+d), one bonded evaluation and execute()
 * (openmmapi/src/LangevinIntegratorSDM.cpp:153-183).  This header is the boundary that
 * replaces that kernel: plain pointers and sizes, no C++ or torch types, int status
 * codes (0 = OK) with sdm_last_error() for the message.  Every call is made on the
 * calling thread's current HIP device; a context owns its device buffers and one HIP
 * stream and is not thread-safe.  Units: nm, ps, kJ/mol, e, amu, K.
 *
 * One context holds R >= 1 replicas of the same topology (a lambda ladder); each
 * replica has its own coordinates, velocities, parameters and clock.  R == 1 is the
 * plain single-replica kernel.
 */
#ifndef SDM_HIP_H
#define SDM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define SDM_OK 0
#define SDM_ERR_INVALID -1  /* bad argument (message in sdm_last_error) */
#define SDM_ERR_HIP -2      /* HIP runtime failure */
#define SDM_ERR_STATE -3    /* call not valid in the current state */
#define SDM_ERR_NUMERIC -4  /* non-finite energy or fixed-point overflow flagged */
#define SDM_ERR_CAPACITY -5 /* neighbour-structure capacity exceeded */
#define SDM_ERR_UNSUPPORTED -6

/* ---- method constants: LangevinIntegratorSDM.h:120-122 and :143-145 ---- */
#define SDM_LINEAR_METHOD 0
#define SDM_QUADRATIC_METHOD 1
#define SDM_ILOGISTIC_METHOD 2
#define SDM_NO_SOFTCORE_METHOD 0
#define SDM_TANH_METHOD 1
#define SDM_RATIONAL_METHOD 2

/* ---- per-replica flags reported in sdm_energies.flags ---- */
#define SDM_FLAG_NONFINITE 1  /* an energy or u' was not finite */
#define SDM_FLAG_OVERFLOW 2   /* a force exceeded the 2^31 kJ/mol/nm fixed-point range */
#define SDM_FLAG_CAPACITY 4   /* a neighbour list was truncated */
#define SDM_FLAG_CONSTRAINT 8 /* a constraint cluster did not converge */

/* System description: replaces the OpenMM System + NonbondedForce (group 2) +
   HarmonicBond/AngleForce (group 1) + the integrator's displacement map
   (LangevinIntegratorSDM.h:467-472, captured at ReferenceSDMKernels.cpp:150-154). */
typedef struct {
    int32_t n_atoms;
    const double* charge;  /* [n_atoms] e */
    const double* sigma;   /* [n_atoms] nm, Lorentz-Berthelot */
    const double* epsilon; /* [n_atoms] kJ/mol */
    const double* mass;    /* [n_atoms] amu; 0 = fixed atom */
    int32_t n_exclusions;
    const int32_t* exclusions; /* [2*n_exclusions] excluded nonbonded pairs */
    int32_t n_bonds;
    const int32_t* bond_atoms; /* [2*n_bonds] E = k/2 (r-r0)^2 */
    const double* bond_r0;
    const double* bond_k;
    int32_t n_angles;
    const int32_t* angle_atoms; /* [3*n_angles] E = k/2 (theta-theta0)^2 */
    const double* angle_theta0;
    const double* angle_k;
    const double* displacement; /* [3*n_atoms] state-2 displacement d_i */
    double box[3];              /* rectangular periodic box (ignored if not periodic) */
    int32_t periodic;           /* 1: CutoffPeriodic (minimum image); 0: CutoffNonPeriodic
                                   (no images, as OpenMM's NonbondedForce, the reference's
                                   first example, example/test.py:63) */
    double cutoff;              /* nm; periodic: box >= 2*cutoff in every dimension */
    double eps_rf;              /* reaction-field dielectric */
    int32_t lj_rule;            /* SDM_LJ_ARITHMETIC (0, NonbondedForce's Lorentz-Berthelot
                                   sigma_ij = (s_i + s_j)/2) or SDM_LJ_GEOMETRIC (1,
                                   sigma_ij = sqrt(s_i s_j): the CustomNonbondedForce the
                                   reference's DMS reader builds for OPLS,
                                   desmonddmsfile75.py:780-793); epsilon_ij =
                                   sqrt(e_i e_j) either way.  Geometric runs on the block
                                   Newton-3 pair kernel (sdm_tuning.pair_group 0,
                                   pair_newton3 1). */
} sdm_system_desc;
#define SDM_LJ_ARITHMETIC 0
#define SDM_LJ_GEOMETRIC 1

/* Integrator parameters of one replica: the LangevinIntegratorSDM fields
   (LangevinIntegratorSDM.h:497-521, defaults LangevinIntegratorSDM.cpp:48-85). */
typedef struct {
    int32_t bias_method;     /* setBiasMethod */
    int32_t softcore_method; /* setSoftCoreMethod */
    double lambda;           /* setLambda */
    double gamma;            /* setGamma */
    double wbcoeff;          /* setWBcoeff */
    double w0coeff;          /* setW0coeff */
    double lambda1, lambda2; /* setLambda1/2 */
    double alpha;            /* setAlpha, 1/(kJ/mol) */
    double u0;               /* setU0 */
    double umax, acore, ubcore;
    int32_t nonequilibrium;  /* setNonEquilibrium */
    double noneq_tmax;       /* setNoneqtmax */
    double work;             /* setNoneqWorkvalue */
    double m_l1, m_l2, m_u0, m_w0; /* set*Slope */
    double b_l1, b_l2, b_u0, b_w0; /* set*intercept */
    double temperature;      /* K */
    double friction;         /* 1/ps */
    double step_size;        /* ps */
    uint64_t seed;           /* setRandomNumberSeed */
    double m_alpha;          /* alpha slope while switching: alpha(lambda) = alpha +
                                m_alpha*lambda, with dW/dalpha*m_alpha in the work
                                increment.  0 (the reference, where alpha is constant and
                                its work term is dropped, ReferenceSDMKernels.cpp:295) */
} sdm_params;

/* Observables of one replica for the last evaluated step
   (ReferenceSDMKernels.cpp:264-302). */
typedef struct {
    double U0;         /* State1Energy: group-2 energy at x */
    double U1;         /* State2Energy: group-2 energy at x+d */
    double u;          /* U1 - U0, summed over pairs with d_i != d_j */
    double Ub;         /* BondAndRestraintEnergy: group-1 energy at x */
    double bind_e;     /* getBindE(): u' = SoftCoreF(u) */
    double fp;         /* du'/du */
    double W;          /* bias W(u') */
    double Wp;         /* W'(u') */
    double sp;         /* W' * fp: weight of the state-2 force */
    double pot_energy; /* getPotEnergy(): U0 + W + Ub */
    double lambda, lambda1, lambda2, u0, w0; /* parameters in effect (noneq schedule) */
    double work;       /* getNoneqWorkvalue() after the step */
    double time;       /* clock after the step (ps) */
    int64_t step;      /* steps taken */
    int32_t flags;     /* SDM_FLAG_* (sticky until sdm_clear_flags) */
    int32_t pad;
} sdm_energies;

typedef struct sdm_context sdm_context;

/* Launch configuration of the hot kernels, chosen by sdm_autotune. */
typedef struct {
    int32_t pair_waves_per_group; /* 1, 2 or 4 wave64s per workgroup in the pair kernel */
    int32_t cell_cells_per_dim;   /* sort-grid resolution; 0 = automatic */
    int32_t use_graph;            /* 1: replay steps from captured hipGraphs; 0: launch each
                                     kernel directly.  Either way the bonded terms and the
                                     displaced-state chain run beside the pair kernel on
                                     two side streams. */
    int32_t pair_newton3;         /* block kernel only.  1: each pair once (half tile list,
                                     j-force scatter); 0: each ordered pair (no atomics) */
    int32_t pair_group;           /* 0: block kernel (one 64-atom i-block per wave, shared
                                     j stream); 16 or 8: grouped kernel (the wave's i-block
                                     split into groups of 16 or 8 atoms, each with its own j
                                     stream, each pair once) */
    int32_t sort_interval;        /* steps between full spatial sorts (1: every step); the
                                     steps in between keep the slot order and only refresh
                                     the coordinates and block boxes.  Results differ only
                                     in the FP32 summation order of the pair energies. */
    int32_t energy_every_step;    /* 0: sdm_step evaluates the pair energies (U0, U1,
                                     pot_energy) on the last step of each call only, the
                                     step whose record sdm_get_energies returns (u, W, sp,
                                     the work and the forces are evaluated every step);
                                     1: on every step.  The pair energies are FP64 sums
                                     over the FP64 positions (their own kernel); the
                                     forces are FP32 pair arithmetic. */
    int32_t pair_split;           /* block Newton-3 kernel only: 1, 2, 4 (the default), 8,
                                     16 or 32 waves per i-block, each taking every
                                     pair_split-th tile, so that short tile lists do not
                                     leave the GPU waiting on long ones and small batches
                                     fill it.
                                     Results differ only in rounding: each wave's i forces
                                     and pair energies are summed on their own before they
                                     are added up. */
    int32_t side_fork;            /* where the restraints, force-field terms and the
                                     displaced-state chain run: 0 on two side streams
                                     forked after the tile lists (beside the pair kernel),
                                     1 (the default) forked after the slot order (beside
                                     the tile-list build as well), 2 on the main stream
                                     between the tile lists and the pair kernel, 3 as 1
                                     but with the displaced-state chain joined back
                                     before the pair kernel (beside the tile lists only),
                                     4 as 1 but with the bonds and angles forked with the
                                     rest.  With 0, 1 and 3 the bonds and angles start
                                     with the step on a side stream (beside the spatial
                                     sort, one more fork event per step); with 2 they
                                     run with the rest.  Same results. */
    float list_skin;              /* Verlet skin of the tile lists in nm, 0 to
                                     SDM_MAX_LIST_SKIN (block pair kernels only; 0, the
                                     default, builds them on every evaluation).  Lists
                                     are built for r_c + list_skin and kept while every
                                     atom of the replica has moved less than list_skin / 2
                                     since, decided on the device each step; the pair
                                     kernel still applies the exact cutoff.  With
                                     pair_split 1 results are the same bits as with 0;
                                     otherwise they differ in rounding (the waves' tile
                                     shares change). */
} sdm_tuning;

/* largest sdm_tuning.list_skin (nm); the tile capacity is sized for it */
#define SDM_MAX_LIST_SKIN 0.1

/* Message of the last failing call on this thread ("" if none). */
const char* sdm_last_error(void);

/* Creates a context for n_replicas replicas of `sys` on the current device.
   Replaces IntegrateLangevinStepSDMKernel::initialize (SDMKernels.h:60;
   ReferenceSDMKernels.cpp:144-159, OpenCLSDMKernels.cpp:73-133): masses and the
   displacement map are captured here.  All replicas start with the reference
   constructor defaults (LangevinIntegratorSDM.cpp:48-85), zero coordinates and
   velocities. */
int sdm_create(const sdm_system_desc* sys, int32_t n_replicas, sdm_context** out);
int sdm_destroy(sdm_context* ctx);
int sdm_num_replicas(const sdm_context* ctx, int32_t* n_replicas, int32_t* n_atoms);

/* Parameters: LangevinIntegratorSDM setters/getters (LangevinIntegratorSDM.h:81-464).
   replica = -1 sets every replica.  sdm_get_params returns the device copy, which the
   nonequilibrium schedule and work accumulation update every step. */
int sdm_set_params(sdm_context* ctx, int32_t replica, const sdm_params* p);
int sdm_get_params(sdm_context* ctx, int32_t replica, sdm_params* p);
/* The integrator clock (ContextImpl time and step count). */
int sdm_set_time(sdm_context* ctx, int32_t replica, double time, int64_t step);

/* Replaces updateParametersInContext for the nonbonded per-atom parameters. */
int sdm_set_nonbonded_params(sdm_context* ctx, const double* charge, const double* sigma,
                             const double* epsilon);

/* Restraint forces of SDMUtils (python/SDMUtils.py).  Both kinds are force-group-1 terms:
   evaluated at x (not at x+d), added to the bonded forces and to Ub
   (LangevinIntegratorSDM.cpp:174-176).  Positions are taken as they are, without periodic
   images, as the reference's CustomCentroidBondForce / CustomCompoundBondForce do.
   Units: nm, rad, kJ/mol. */

/* SDMUtils.addRestraintForce (SDMUtils.py:32-158), a CustomCentroidBondForce:
     C*(kfcm/2)*step(d12-tolcm)*(d12-tolcm)^2,  d12 = |c1 - offset - c2|
   with c1, c2 the weighted centres of the ligand and receptor groups, plus, with
   do_angles, C*(kf/2)*(step(dm)*max(0,db)^2 + step(-dm)*max(0,-da)^2) for
   theta = angle(g3,g6,g7) (period pi), phi1 = dihedral(g4,g3,g6,g7) and
   phi2 = dihedral(g3,g6,g7,g8) (period 2 pi), where db, da, dm are theta - hi, theta - lo
   and theta - (lo+hi)/2 wrapped to the period; g3..g5 = rcpt_ref, g6..g8 = lig_ref.
   Dihedrals follow the IUPAC sign convention. */
typedef struct {
    int32_t n_lig, n_rcpt;
    const int32_t* lig_atoms;  /* group g1 */
    const int32_t* rcpt_atoms; /* group g2 */
    const double* lig_weights; /* NULL: the particle masses (the force's default) */
    const double* rcpt_weights;
    double kfcm;               /* kJ/mol/nm^2 */
    double tolcm;              /* nm */
    double offset[3];          /* nm */
    int32_t do_angles;         /* 1: the angle and dihedral terms */
    int32_t rcpt_ref[3];
    int32_t lig_ref[3];
    int32_t pad;
    double kf[3];              /* theta, phi1, phi2: kJ/mol/rad^2 */
    double lo[3], hi[3];       /* flat-bottom windows: center -/+ tolerance, rad */
    double control;            /* SDMRestraintControlParameter (1.0) */
} sdm_centroid_restraint;

/* SDMUtils.addAlignmentForce (SDMUtils.py:161-257): three CustomCompoundBondForces on the
   reference atoms a0..a2 (liga_ref) and b0..b2 (ligb_ref):
     (kfdispl/2)|b0 - offset - a0|^2
   + (ktheta/2)(1 - cos t),  cos t = unit(b1-b0).unit(a1-a0)
   + (kpsi/4)(1 - cos p(b0,b1,b2,a0,a2)) + (kpsi/4)(1 - cos p(a0,a1,a2,b0,b2))
   where cos p(x1..x5) is the cosine between the components of x3-x1 and x5-x4
   perpendicular to x2-x1. */
typedef struct {
    int32_t liga_ref[3];
    int32_t ligb_ref[3];
    double kfdispl;            /* kJ/mol/nm^2 */
    double ktheta, kpsi;       /* kJ/mol */
    double offset[3];          /* nm */
} sdm_alignment_restraint;

/* Further force-field terms of a Desmond DMS system (example/desmonddmsfile75.py:546-850).
   Group 1 (at x, into the bonded forces and Ub):
     periodic torsions   k (1 + cos(n phi - phase))        (PeriodicTorsionForce)
     harmonic impropers  k (phi - phi0)^2                   (CustomTorsionForce, :684-717)
     flat-bottom posre   k (|x - x0| - tol)^2 beyond tol, minimum image (:972-1017)
     harmonic posre      sum_d k_d (x_d - x0_d)^2           (:917-970)
   (phi with the IUPAC sign).  Group 2: the 1-4 exceptions of the NonbondedForce, replacing
   the pair (which must also be an exclusion): 4 eps ((s/r)^12 - (s/r)^6) + k_e qq / r with
   no cutoff, no reaction field and no periodic image (:812-845; OpenMM's exceptions).  Both
   atoms of an exception must have the same displacement (their energy is then the same at
   x and x+d); it is added to U0 and U1 and its force to both states.  Constraints: fixed
   distances kept by the Langevin update (positions after each step, velocities as the
   constrained displacement over dt), as OpenMM's LangevinIntegrator does; solved by SHAKE
   to relative tolerance constraint_tol (the connected atoms of a cluster share one
   thread).  Units nm, rad, kJ/mol, e. */
typedef struct {
    int32_t n_torsions;
    const int32_t* torsion_atoms;   /* [4*n] */
    const int32_t* torsion_n;       /* periodicity */
    const double* torsion_phase;    /* rad */
    const double* torsion_k;        /* kJ/mol */
    int32_t n_impropers;
    const int32_t* improper_atoms;  /* [4*n] */
    const double* improper_phi0;    /* rad */
    const double* improper_k;       /* kJ/mol/rad^2 */
    int32_t n_posre_fb;
    const int32_t* posre_fb_atoms;
    const double* posre_fb_x0;      /* [3*n] nm */
    const double* posre_fb_k;       /* kJ/mol/nm^2 */
    const double* posre_fb_tol;     /* nm */
    int32_t n_posre_harm;
    const int32_t* posre_harm_atoms;
    const double* posre_harm_x0;    /* [3*n] nm */
    const double* posre_harm_k;     /* [3*n] kJ/mol/nm^2 per axis */
    int32_t n_exceptions;
    const int32_t* exception_atoms; /* [2*n] */
    const double* exception_qq;     /* e^2 */
    const double* exception_sigma;  /* nm */
    const double* exception_epsilon;/* kJ/mol */
    int32_t n_constraints;
    const int32_t* constraint_atoms;/* [2*n] */
    const double* constraint_length;/* nm */
    double constraint_tol;          /* relative; 0 = keep the context's tolerance (1e-10
                                       unless sdm_set_constraint_tolerance changed it) */
} sdm_forcefield_terms;

/* Replaces the context's further force-field terms (NULL removes them).  No counterpart
   call in the reference, whose DMS reader adds these forces to the System. */
int sdm_set_forcefield_terms(sdm_context* ctx, const sdm_forcefield_terms* terms);

/* Virtual sites: massless atoms placed from two or three others.  The reference's
   integrator places them after every step (ReferenceStochasticDynamicsSDM.cpp:264,
   OpenCLSDMKernels.cpp:376-377, through OpenMM 7's VirtualSite classes), and its DMS reader
   makes them from the virtual_lc2 / virtual_lc3 / virtual_out3 tables
   (desmonddmsfile75.py _addVirtualSitesToSystem).  Site k is atom atoms[4k] (mass 0),
   placed from atoms[4k+1..4k+3] (the fourth is ignored for SDM_VSITE_AVERAGE2) with
   weights w = weights[3k..3k+2]:
     SDM_VSITE_AVERAGE2      x = w0 x1 + w1 x2                (TwoParticleAverageSite)
     SDM_VSITE_AVERAGE3      x = w0 x1 + w1 x2 + w2 x3        (ThreeParticleAverageSite)
     SDM_VSITE_OUT_OF_PLANE  x = x1 + w0 r12 + w1 r13 + w2 (r12 x r13), r1k = xk - x1
                                                              (OutOfPlaneSite)
   Each step moves the forces on the sites to their parents (the transpose of the site's
   Jacobian) before the Langevin update and places the sites after it; sdm_get_forces
   reports forces moved the same way (zero on the sites).  Parents must have mass and not
   be sites; a site and its parents share one displacement class.  n = 0 removes them.
   No counterpart call in the reference, where the System holds the sites. */
#define SDM_VSITE_AVERAGE2 0
#define SDM_VSITE_AVERAGE3 1
#define SDM_VSITE_OUT_OF_PLANE 2
int sdm_set_virtual_sites(sdm_context* ctx, int32_t n, const int32_t* kind,
                          const int32_t* atoms, const double* weights);

/* Relative tolerance of the position constraints (|r|^2 within tol of d^2).  Replaces
   Integrator::setConstraintTolerance, whose value the reference hands to the constraint
   solver every step (ReferenceStochasticDynamicsSDM.cpp:252, getConstraintTolerance()). */
int sdm_set_constraint_tolerance(sdm_context* ctx, double tol);

/* Centre-of-mass motion removal every `frequency` steps (0 = never, the default): the
   mass-weighted mean velocity is subtracted from every atom with mass, before the
   step's Langevin update, as OpenMM's CMMotionRemover does (which the reference's DMS
   reader adds to every System, desmonddmsfile75.py:538-541). */
int sdm_set_cm_motion_removal(sdm_context* ctx, int32_t frequency);

/* Replaces the context's restraints (n = 0 removes them).  No counterpart call in the
   reference, where SDMUtils adds the forces to the System before the Context is built. */
int sdm_set_restraints(sdm_context* ctx, const sdm_centroid_restraint* centroid,
                       int32_t n_centroid, const sdm_alignment_restraint* alignment,
                       int32_t n_alignment);

/* State (host arrays, [3*n_atoms] doubles, atom order as given to sdm_create). */
int sdm_set_positions(sdm_context* ctx, int32_t replica, const double* xyz);
int sdm_get_positions(sdm_context* ctx, int32_t replica, double* xyz);
int sdm_set_velocities(sdm_context* ctx, int32_t replica, const double* xyz);
int sdm_get_velocities(sdm_context* ctx, int32_t replica, double* xyz);
/* Device-resident variants: d_xyz is a device pointer to [R][n_atoms][3] doubles. */
int sdm_set_positions_device(sdm_context* ctx, const double* d_xyz);
int sdm_get_positions_device(sdm_context* ctx, double* d_xyz);

/* Evaluates the SDM force at the current positions without moving atoms: the two
   state energies, u, the scalar stage and the hybrid force F = F0 + sp*(F1-F0) + Fb.
   Replaces the three calcForcesAndEnergy calls plus the scalar part of execute()
   (LangevinIntegratorSDM.cpp:158-176; ReferenceSDMKernels.cpp:264-318); the clock is
   not advanced. */
int sdm_compute(sdm_context* ctx);

/* nsteps steps of LangevinIntegratorSDM::step (LangevinIntegratorSDM.cpp:153-183) for
   every replica: evaluation, scalar stage, force mix and the Langevin update
   (ReferenceStochasticDynamicsSDM.cpp:131-266 / langevin.cl:7-69). */
int sdm_step(sdm_context* ctx, int32_t nsteps);

/* Blocks until queued work is done, then copies every replica's observables. */
int sdm_get_energies(sdm_context* ctx, sdm_energies* out /* [n_replicas] */);
int sdm_clear_flags(sdm_context* ctx);

/* Forces of the last sdm_compute/sdm_step evaluation, [3*n_atoms] doubles.
   which: 0 = hybrid force, 1 = F0 (group 2 at x), 2 = F1 (group 2 at x+d),
   3 = Fb (group 1).  1-3 require sdm_compute_components. */
#define SDM_FORCE_HYBRID 0
#define SDM_FORCE_STATE1 1
#define SDM_FORCE_STATE2 2
#define SDM_FORCE_BONDED 3
int sdm_compute_components(sdm_context* ctx);
int sdm_get_forces(sdm_context* ctx, int32_t replica, int32_t which, double* out);

/* computeKineticEnergy (SDMKernels.h:77; ReferenceSDMKernels.cpp:108-135,345-347):
   0.5 sum m v^2 with the velocities of constrained atoms projected onto the constraints
   first (on a copy, as the reference's applyToVelocities call does). */
int sdm_kinetic_energy(sdm_context* ctx, double* out /* [n_replicas] */);

/* u' recorded every `every` steps (0 disables) into a per-replica ring of `capacity`:
   the u' evaluated on steps 0, every, 2 every, ... (the step index before the step), the
   exchange record of a replica-exchange ladder.  sdm_get_u_history writes the newest
   min(count held, capacity) records, oldest first, and their number to *count. */
int sdm_set_u_history(sdm_context* ctx, int32_t every, int32_t capacity);
int sdm_get_u_history(sdm_context* ctx, int32_t replica, double* out, int32_t capacity,
                      int32_t* count);

/* Diagnostics for the roofline: exact counts, at the current positions, of
   non-excluded pairs with r < cutoff at x (p0) and of pairs with d_i != d_j and
   r < cutoff at x+d (pdelta). */
int sdm_count_pairs(sdm_context* ctx, int64_t* p0, int64_t* pdelta /* [n_replicas] */);
/* Average duration (ms) of the nonbonded pair kernel as the steps run it (forces only;
   the pair energies have a kernel of their own): nsteps launches back to back, each after the slot
   order's refresh and tile lists and with nothing running beside it, timed with HIP
   events on the context's stream.  Positions do not advance; forces must be recomputed
   afterwards. */
int sdm_time_pair_kernel(sdm_context* ctx, int32_t nsteps, double* avg_ms);

/* Launch-configuration search (project-written autotuner): times each candidate (pair
   kernel, waves per workgroup and per i-block, sort grid) on the context's own replicas
   for `steps` steps, then times the three fastest, each with and without graph replay
   and with each of the five placements of the side work (side_fork), for 3 x `steps`
   steps in two rounds, and keeps the one with the smallest total; with a block pair
   kernel it then times that one with list_skin 0, 0.01, 0.02, 0.03 and 0.04 nm the same
   way.  The state the context had before the call is restored. */
int sdm_autotune(sdm_context* ctx, int32_t steps, sdm_tuning* chosen);
int sdm_get_tuning(sdm_context* ctx, sdm_tuning* t);
int sdm_set_tuning(sdm_context* ctx, const sdm_tuning* t);

/* Adjoint (no counterpart in the reference; parameter gradients of u for force-field
   fitting): for every frame f of d_frames ([nframes][n_atoms][3] device doubles) and
   every selected atom s, writes du/dq, du/dsigma, du/deps to d_grad
   ([nframes][nsel][3] device doubles) and u to d_u ([nframes] device doubles).
   Pair terms are recomputed, never stored.  Only pairs with d_i != d_j contribute.
   Either output may be NULL (u alone is the forward pass, the gradient the backward). */
int sdm_du_dparams_frames(sdm_context* ctx, const double* d_frames, int32_t nframes,
                          const int32_t* sel, int32_t nsel, double* d_grad, double* d_u);

/* sdm_du_dparams_frames without any device allocation: no counterpart in the reference.
   For callers that cannot allocate device memory (code under stream capture, tight
   fitting loops, harnesses that forbid hipMalloc).  sdm_du_dparams_frames stages the
   selection in a buffer it allocates (and grows) on the device; this variant stages it
   in a static device array of SDM_NOALLOC_MAX_SEL entries that every GPU the library
   runs on holds until the process ends.  nsel above SDM_NOALLOC_MAX_SEL is refused with
   SDM_ERR_CAPACITY.  Same arguments, kernels and results.  Every call on one GPU shares
   that GPU's array: calls on one GPU that overlap in time (from several threads, or
   several contexts' streams) race; calls made one after another do not (each returns
   after its kernels have finished). */
#define SDM_NOALLOC_MAX_SEL 16384
int sdm_du_dparams_frames_noalloc(sdm_context* ctx, const double* d_frames, int32_t nframes,
                                  const int32_t* sel, int32_t nsel, double* d_grad,
                                  double* d_u);

/* OpenMM plugin-loader entry points (ReferenceSDMKernelFactory.cpp:41-56).  Without an
   OpenMM runtime in the process they only record that they were called. */
void registerPlatforms(void);
void registerKernelFactories(void);
void registerSDMHIPKernelFactories(void);

#ifdef __cplusplus
}
#endif
#endif
