/*
 * sdm_oracle.h -- FP64 CPU restatement of the SDM alchemical step.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links, loads or calls it.
 *
 * What it restates (reference paths are relative to the plugin's source tree):
 *   orc_softcore          openmmapi/src/LangevinIntegratorSDM.cpp:125-149 (SoftCoreF)
 *   orc_execute_scalars   platforms/reference/src/ReferenceSDMKernels.cpp:202-309
 *                         (noneq schedule, bias W/W', PotEnergy, BindE, work, sp)
 *   orc_langevin_update   platforms/reference/src/ReferenceStochasticDynamicsSDM.cpp:131-266
 *                         with parameters from ReferenceSDMKernels.cpp:320-335
 *   orc_langevin_update_cl  platforms/opencl/src/OpenCLSDMKernels.cpp:331-355 and
 *                         kernels/langevin.cl:7-69 (the friction==0-safe form)
 *   orc_force_mix         ReferenceSDMKernels.cpp:308-318
 *   orc_kinetic_energy    ReferenceSDMKernels.cpp:108-135 (no constraints)
 *   orc_ref_step          openmmapi/src/LangevinIntegratorSDM.cpp:153-183 (one step:
 *                         E1 at x, E2 at x+d, Eb at x, then execute)
 *   orc_nonbonded_*       OpenMM NonbondedForce LJ + reaction field (third-party, not in
 *                         the reference tree; called at LangevinIntegratorSDM.cpp:160,168):
 *                         documented formula, cutoff, no switching, no dispersion
 *                         correction, excluded pairs skipped.
 *   orc_bonded            OpenMM HarmonicBondForce / HarmonicAngleForce (group 1,
 *                         LangevinIntegratorSDM.cpp:176), documented formulas.
 *   orc_sdm_direct        the same U0, U1 = U(x+d), u = U1-U0, F0, F1 as
 *                         orc_ref_step's first two evaluations, summed only over
 *                         pairs whose displacements differ (equal-displacement pairs
 *                         cancel exactly in U(x+d)-U(x)).
 *   orc_du_dparams        d u / d(q_i, sigma_i, eps_i) of displaced atoms (adjoint target).
 *   orc_philox4x32_10     Philox4x32-10 (Salmon et al., SC'11 "Parallel random numbers:
 *                         as easy as 1, 2, 3"); orc_gauss3 = Box-Muller on it.  The
 *                         reference's own RNG is OpenMM-internal and OS-seeded, so
 *                         noise parity with the reference is statistical only.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   softcore / bias       closed forms at special points, FD derivatives, asymptotes.
 *   nonbonded             hand-computed two- and three-atom energies/forces, RF zero at
 *                         the cutoff, brute force vs cell list, Newton's third law,
 *                         F = -grad U by central FD, exclusions.
 *   sdm_direct            d == 0 and d == box vector give u == 0, uniform d gives u == 0,
 *                         u_direct == E2 - E1 of the two-sweep reference form.
 *   force mix             Linear lambda 0/1 give F0+Fb / F1+Fb; F = -grad(U0+W(u'(u))+Ub)
 *                         by central FD.
 *   langevin              T = 0 closed form, friction = 0 quirk, <KE> = 3N/2 kT.
 *   philox                Random123 known-answer vectors.
 *   du_dparams            central FD of u.
 * Against OpenMM's own NonbondedForce binary: parity unpinned (OpenMM is not
 * present; only its documented formula is restated).
 */
#ifndef SDM_ORACLE_H
#define SDM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_ONE_4PI_EPS0 138.935456
#define ORC_BOLTZ (1.380658e-23 * 6.0221367e23 / 1000.0)

enum { ORC_LINEAR = 0, ORC_QUADRATIC = 1, ORC_ILOGISTIC = 2 };
enum { ORC_NOSOFTCORE = 0, ORC_TANH = 1, ORC_RATIONAL = 2 };

typedef struct {
    int bias_method;
    int softcore_method;
    double lambda, gamma, wbcoeff, w0coeff;
    double lambda1, lambda2, alpha, u0;
    double umax, acore, ubcore;
    int nonequilibrium;
    double noneq_tmax, work;
    double m_l1, m_l2, m_u0, m_w0, b_l1, b_l2, b_u0, b_w0;
    double temperature, friction, step_size;
    double m_alpha; /* alpha(lambda) = alpha + m_alpha*lambda while switching (0: constant) */
} orc_params;

typedef struct {
    double bind_e;      /* u' (soft-cored binding energy), BindE */
    double pot_energy;  /* E1 + W + Eb */
    double ebias;       /* W(u') */
    double bfp;         /* W'(u') */
    double fp;          /* du'/du */
    double sp;          /* bfp*fp, the state-2 force weight */
} orc_scalars;

typedef struct {
    int n;
    const double* q;
    const double* sigma;
    const double* eps;
    int n_excl;
    const int* excl; /* 2*n_excl atom indices */
    double box[3];
    int periodic;
    double cutoff;
    double eps_rf;
    int geometric; /* 0: sigma_ij = (s_i + s_j)/2 (NonbondedForce); 1: sqrt(s_i s_j), the
                      CustomNonbondedForce the reference's DMS reader builds for OPLS
                      (desmonddmsfile75.py:780-793) */
} orc_nb;

typedef struct {
    int n_bonds;
    const int* bond_atoms;  /* 2*n_bonds */
    const double* bond_r0;
    const double* bond_k;
    int n_angles;
    const int* angle_atoms; /* 3*n_angles */
    const double* angle_t0;
    const double* angle_k;
} orc_bonded_t;

/* ---- scalar stage ---- */
int orc_softcore(int method, double u, double umax, double a, double ub, double* up, double* fp);
int orc_execute_scalars(orc_params* p, double E1, double E2, double Eb, double time,
                        orc_scalars* out);

/* ---- nonbonded LJ + reaction field ---- */
double orc_pair_energy(double r, double qq, double sig, double eps, double rc, double eps_rf,
                       double* f_over_r);
double orc_nonbonded_brute(const orc_nb* nb, const double* x, double* F, long long* npairs);
double orc_nonbonded_cells(const orc_nb* nb, const double* x, double* F, long long* npairs);
double orc_nonbonded(const orc_nb* nb, const double* x, double* F, long long* npairs);

/* ---- bonded ---- */
double orc_bonded(const orc_bonded_t* b, const double* x, double* F_accum);

/* ---- dual state ---- */
/* two-sweep form: E1,F1 at x; E2,F2 at x+d; Eb,Fb at x (any output may be NULL) */
void orc_sdm_reference_eval(const orc_nb* nb, const orc_bonded_t* b, const double* x,
                            const double* displ, double* E1, double* E2, double* Eb,
                            double* F1, double* F2, double* Fb);
/* direct form: U0, F0 full sweep at x; u and dF = F1-F0 over pairs with d_i != d_j;
   npairs0 = in-cutoff non-excluded pairs at x, npairs_delta = in-cutoff pairs with
   d_i != d_j at x+d */
void orc_sdm_direct(const orc_nb* nb, const double* x, const double* displ, double* U0,
                    double* u, double* F0, double* dF, long long* npairs0,
                    long long* npairs_delta);
void orc_force_mix(int n, double sp, const double* F1, const double* F2, const double* Fb,
                   double* F);
void orc_du_dparams(const orc_nb* nb, const double* x, const double* displ, int nsel,
                    const int* sel, double* dq, double* dsig, double* deps);

/* ---- dynamics ---- */
void orc_langevin_update(int n, double* x, double* v, const double* F, const double* mass,
                         double dt, double friction, double temperature, const double* xi);
void orc_langevin_update_cl(int n, double* x, double* v, const double* F, const double* mass,
                            double dt, double friction, double temperature, const double* xi);
double orc_kinetic_energy(int n, const double* v, const double* mass);

/* ---- random numbers ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_gauss3(uint64_t seed, uint32_t replica, uint64_t step, uint32_t atom, double out[3]);

/* ---- whole reference step ---- */
/* Advances (x, v) by nsteps steps of LangevinIntegratorSDM::step.  time and step are
   the integrator's clock (in/out).  Gaussian noise is orc_gauss3(seed, replica, step, atom).
   last (optional) receives the scalars and energies of the last step.  use_cl_langevin
   selects the friction==0-safe OpenCL form of the update. */
typedef struct {
    double E1, E2, Eb;
    orc_scalars s;
} orc_step_report;
int orc_ref_step(const orc_nb* nb, const orc_bonded_t* b, const double* displ,
                 const double* mass, orc_params* p, double* x, double* v, double* time,
                 uint64_t* step, uint64_t seed, uint32_t replica, int nsteps,
                 int use_cl_langevin, orc_step_report* last);
/* R replicas of orc_ref_step on `threads` OpenMP threads (one replica per thread). */
int orc_ref_step_replicas(const orc_nb* nb, const orc_bonded_t* b, const double* displ,
                          const double* mass, orc_params* params, double* x, double* v,
                          double* time, uint64_t* step, const uint64_t* seed, int R,
                          int nsteps, int use_cl_langevin, int threads);

#ifdef __cplusplus
}
#endif
#endif
