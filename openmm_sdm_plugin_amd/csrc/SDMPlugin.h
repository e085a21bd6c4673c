// SDMPlugin.h -- C++ host API of the SDM alchemical force over the HIP C ABI.
//
// Mirrors the reference plugin's public C++ surface:
//   LangevinIntegratorSDM      openmmapi/include/LangevinIntegratorSDM.h:53-522
//                              (same setters/getters, method constants, defaults from
//                              openmmapi/src/LangevinIntegratorSDM.cpp:48-85, step(),
//                              SoftCoreF, errors "This Integrator is already bound to a
//                              context" / "Unknown soft core method")
//   IntegrateLangevinStepSDMKernel   openmmapi/include/SDMKernels.h:47-113 (abstract)
// and adds the force-level interface the SDM force is driven through:
//   SDMForce / CalcSDMForceKernel    (lambda1/lambda2/alpha/u0/w0, umax/acore/ubcore,
//                                     per-atom displacement map, getPerturbationEnergy)
// OpenMM itself is not a dependency: System and Context below are the minimal stand-ins
// that carry what the reference reads from OpenMM (particles, NonbondedForce LJ+RF
// parameters and exclusions, HarmonicBond/AngleForce terms, the periodic box,
// positions, velocities, time).  Compute runs through libsdm_hip.so.
#pragma once

#include <array>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sdm_hip.h"

namespace SDMPlugin {

using Vec3 = std::array<double, 3>;

class SDMException : public std::runtime_error {
public:
    explicit SDMException(const std::string& m) : std::runtime_error(m) {}
};

// Particles + the group-2 (nonbonded LJ + reaction field) and group-1 (harmonic bonds
// and angles) terms of an OpenMM System.
// OpenMM 7's VirtualSite classes, the three the reference's DMS reader attaches
// (example/desmonddmsfile75.py:877, :892, :907).  A site is a massless particle placed from
// its parents after every step; the forces on it move to the parents.
class VirtualSite {
public:
    virtual ~VirtualSite() = default;
    int getNumParticles() const { return (int)particles_.size(); }
    int getParticle(int i) const;

protected:
    friend class Context;
    VirtualSite(int kind, std::vector<int> particles, const Vec3& w)
        : kind_(kind), particles_(std::move(particles)), w_(w) {}
    int kind_;                    // SDM_VSITE_*
    std::vector<int> particles_;  // parents
    Vec3 w_;
};

// x = w1 x1 + w2 x2
class TwoParticleAverageSite : public VirtualSite {
public:
    TwoParticleAverageSite(int p1, int p2, double w1, double w2)
        : VirtualSite(SDM_VSITE_AVERAGE2, {p1, p2}, {{w1, w2, 0.0}}) {}
    double getWeight(int i) const;
};

// x = w1 x1 + w2 x2 + w3 x3
class ThreeParticleAverageSite : public VirtualSite {
public:
    ThreeParticleAverageSite(int p1, int p2, int p3, double w1, double w2, double w3)
        : VirtualSite(SDM_VSITE_AVERAGE3, {p1, p2, p3}, {{w1, w2, w3}}) {}
    double getWeight(int i) const;
};

// x = x1 + w12 r12 + w13 r13 + wCross (r12 x r13), r1k = xk - x1
class OutOfPlaneSite : public VirtualSite {
public:
    OutOfPlaneSite(int p1, int p2, int p3, double w12, double w13, double wCross)
        : VirtualSite(SDM_VSITE_OUT_OF_PLANE, {p1, p2, p3}, {{w12, w13, wCross}}) {}
    double getWeight12() const { return w_[0]; }
    double getWeight13() const { return w_[1]; }
    double getWeightCross() const { return w_[2]; }
};

class System {
public:
    int addParticle(double mass);
    int getNumParticles() const { return (int)mass_.size(); }
    void setParticleParameters(int i, double charge, double sigma, double epsilon);
    void addExclusion(int a, int b);
    void addBond(int a, int b, double r0, double k);
    void addAngle(int a, int b, int c, double theta0, double k);
    void setPeriodicBoxVectors(const Vec3& box) { box_ = box; }
    // OpenMM NonbondedForce::CutoffPeriodic (the default here) or CutoffNonPeriodic
    enum NonbondedMethod { CutoffNonPeriodic = 1, CutoffPeriodic = 2 };
    void setNonbondedMethod(NonbondedMethod m) { periodic_ = m == CutoffPeriodic; }
    bool usesPeriodicBoundaryConditions() const { return periodic_; }
    void setCutoffDistance(double rc) { cutoff_ = rc; }
    void setReactionFieldDielectric(double e) { eps_rf_ = e; }
    double getCutoffDistance() const { return cutoff_; }
    // OpenMM's System::setVirtualSite: the System takes ownership of site; the particle
    // must have mass 0, its parents mass.  Checked when a Context is created.
    void setVirtualSite(int index, VirtualSite* site);
    bool isVirtualSite(int index) const { return vsites_.count(index) != 0; }
    const VirtualSite& getVirtualSite(int index) const;
    // How the LJ sigma of a pair combines: LorentzBerthelot, (s_i + s_j)/2 as OpenMM's
    // NonbondedForce (the default), or Geometric, sqrt(s_i s_j) as the OPLS
    // CustomNonbondedForce of the reference's DMS reader (desmonddmsfile75.py:780-793);
    // epsilon is sqrt(e_i e_j) either way.
    enum CombiningRule { LorentzBerthelot = SDM_LJ_ARITHMETIC, Geometric = SDM_LJ_GEOMETRIC };
    void setCombiningRule(CombiningRule r) { lj_rule_ = r; }
    CombiningRule getCombiningRule() const { return lj_rule_; }

private:
    friend class Context;
    friend class HIPCalcSDMForceKernel;
    friend class HIPIntegrateLangevinStepSDMKernel;
    std::vector<double> mass_, q_, sig_, eps_;
    std::vector<int> excl_, bonds_, angles_;
    std::vector<double> bond_r0_, bond_k_, ang_t0_, ang_k_;
    Vec3 box_{{0, 0, 0}};
    double cutoff_ = 1.0, eps_rf_ = 78.3;
    bool periodic_ = true;
    CombiningRule lj_rule_ = LorentzBerthelot;
    std::map<int, std::shared_ptr<const VirtualSite>> vsites_;  // shared: Contexts copy
};

class LangevinIntegratorSDM;
class Context;

// Abstract platform kernel of the reference (SDMKernels.h:47-113), with its method names
// and arguments.  execute() takes the three energies that the reference's step() obtains
// from OpenMM's force evaluations (SDMKernels.h:67-68, LangevinIntegratorSDM.cpp:160-181);
// Save/Make/Restore exist for interface compatibility.
class IntegrateLangevinStepSDMKernel {
public:
    static std::string Name() { return "IntegrateLangevinStepSDM"; }
    virtual ~IntegrateLangevinStepSDMKernel() = default;
    virtual void initialize(const System& system, const LangevinIntegratorSDM& integrator) = 0;
    virtual void execute(Context& context, LangevinIntegratorSDM& integrator,
                         double State1Energy, double State2Energy,
                         double BondAndRestraintEnergy) = 0;
    virtual double computeKineticEnergy(Context& context, const LangevinIntegratorSDM& integrator) = 0;
    virtual void SaveState1(Context&, const LangevinIntegratorSDM&) {}
    virtual void SaveState2(Context&, const LangevinIntegratorSDM&) {}
    virtual void RestoreState1(Context&, const LangevinIntegratorSDM&) {}
    virtual void MakeState2(Context&, const LangevinIntegratorSDM&) {}
};

class SDMForce;

// Force-level kernel: energy and hybrid force of the SDM potential at the current
// positions, without stepping.
class CalcSDMForceKernel {
public:
    static std::string Name() { return "CalcSDMForce"; }
    virtual ~CalcSDMForceKernel() = default;
    virtual void initialize(const System& system, const SDMForce& force) = 0;
    virtual double execute(Context& context, bool includeForces, bool includeEnergy) = 0;
    virtual void copyParametersToContext(Context& context, const SDMForce& force) = 0;
};

// Parameters shared by the integrator and the force (LangevinIntegratorSDM fields).
class SDMParameters {
public:
    static const int LinearMethod = SDM_LINEAR_METHOD;
    static const int QuadraticMethod = SDM_QUADRATIC_METHOD;
    static const int ILogisticMethod = SDM_ILOGISTIC_METHOD;
    static const int NoSoftCoreMethod = SDM_NO_SOFTCORE_METHOD;
    static const int TanhMethod = SDM_TANH_METHOD;
    static const int RationalMethod = SDM_RATIONAL_METHOD;

    explicit SDMParameters(int nParticles);
    void setLambda(double v) { p_.lambda = v; }
    double getLambda() const { return p_.lambda; }
    void setBiasMethod(int m) { p_.bias_method = m; }
    int getBiasMethod() const { return p_.bias_method; }
    void setSoftCoreMethod(int m) { p_.softcore_method = m; }
    int getSoftCoreMethod() const { return p_.softcore_method; }
    void setGamma(double v) { p_.gamma = v; }
    double getGamma() const { return p_.gamma; }
    void setWBcoeff(double v) { p_.wbcoeff = v; }
    double getWBcoeff() const { return p_.wbcoeff; }
    void setW0coeff(double v) { p_.w0coeff = v; }
    double getW0coeff() const { return p_.w0coeff; }
    void setLambda1(double v) { p_.lambda1 = v; }
    double getLambda1() const { return p_.lambda1; }
    void setLambda2(double v) { p_.lambda2 = v; }
    double getLambda2() const { return p_.lambda2; }
    void setAlpha(double v) { p_.alpha = v; }
    double getAlpha() const { return p_.alpha; }
    void setU0(double v) { p_.u0 = v; }
    double getU0() const { return p_.u0; }
    double getUmax() const { return p_.umax; }
    void setUmax(double v) { p_.umax = v; }
    double getAcore() const { return p_.acore; }
    void setAcore(double v) { p_.acore = v; }
    double getUbcore() const { return p_.ubcore; }
    void setUbcore(double v) { p_.ubcore = v; }
    void setDisplacement(int atom, double dx, double dy, double dz);
    Vec3 getDisplacement(int atom) const;
    // LangevinIntegratorSDM::SoftCoreF (LangevinIntegratorSDM.cpp:125-149)
    double SoftCoreF(double u, double umax, double a, double ub, double& fp) const;
    const sdm_params& raw() const { return p_; }
    sdm_params& raw() { return p_; }
    const std::vector<double>& displacements() const { return displ_; }

protected:
    sdm_params p_;
    std::vector<double> displ_;
};

class LangevinIntegratorSDM : public SDMParameters {
public:
    LangevinIntegratorSDM(double temperature, double frictionCoeff, double stepSize, int nParticles);
    double getTemperature() const { return p_.temperature; }
    void setTemperature(double t) { p_.temperature = t; }
    double getFriction() const { return p_.friction; }
    void setFriction(double f) { p_.friction = f; }
    double getStepSize() const { return p_.step_size; }
    void setStepSize(double dt) { p_.step_size = dt; }
    int getRandomNumberSeed() const { return (int)p_.seed; }
    void setRandomNumberSeed(int s) { p_.seed = (uint64_t)(uint32_t)s; }
    double getBindE() const { return bindE_; }
    void setBindE(double v) { bindE_ = v; }
    double getPotEnergy() const { return potE_; }
    void setPotEnergy(double v) { potE_ = v; }
    void setNoneqtmax(double v) { p_.noneq_tmax = v; }
    double getNoneqtmax() const { return p_.noneq_tmax; }
    void setNonEquilibrium(int v) { p_.nonequilibrium = v; }
    int getNonEquilibrium() const { return p_.nonequilibrium; }
    void setNoneqWorkvalue(double v) { p_.work = v; }
    double getNoneqWorkvalue() const { return p_.work; }
    void setlambda1Slope(double v) { p_.m_l1 = v; }
    double getlambda1Slope() const { return p_.m_l1; }
    void setlambda2Slope(double v) { p_.m_l2 = v; }
    double getlambda2Slope() const { return p_.m_l2; }
    void setu0Slope(double v) { p_.m_u0 = v; }
    double getu0Slope() const { return p_.m_u0; }
    void setw0Slope(double v) { p_.m_w0 = v; }
    double getw0Slope() const { return p_.m_w0; }
    // no counterpart in the reference (alpha is constant while switching there)
    void setalphaSlope(double v) { p_.m_alpha = v; }
    double getalphaSlope() const { return p_.m_alpha; }
    void setlambda1intercept(double v) { p_.b_l1 = v; }
    double getlambda1intercept() const { return p_.b_l1; }
    void setlambda2intercept(double v) { p_.b_l2 = v; }
    double getlambda2intercept() const { return p_.b_l2; }
    void setu0intercept(double v) { p_.b_u0 = v; }
    double getu0intercept() const { return p_.b_u0; }
    void setw0intercept(double v) { p_.b_w0 = v; }
    double getw0intercept() const { return p_.b_w0; }

    // LangevinIntegratorSDM::step (LangevinIntegratorSDM.cpp:153-183)
    void step(int steps);
    double computeKineticEnergy();

private:
    friend class Context;
    friend class HIPIntegrateLangevinStepSDMKernel;
    void initialize(Context& context);  // LangevinIntegratorSDM.cpp:88-106
    Context* context_ = nullptr;
    std::unique_ptr<IntegrateLangevinStepSDMKernel> kernel_;
    double bindE_ = 0.0, potE_ = 0.0;
};

class SDMForce : public SDMParameters {
public:
    explicit SDMForce(int nParticles) : SDMParameters(nParticles) {}
    // soft-cored perturbation energy u' of the last evaluation (the reference's BindE)
    double getPerturbationEnergy(Context& context) const;
    void updateParametersInContext(Context& context) const;
};

struct State {
    double potentialEnergy = 0, perturbationEnergy = 0, U0 = 0, U1 = 0, time = 0;
    std::vector<Vec3> positions, velocities, forces;
};

// A System on the current HIP device, driven by either a LangevinIntegratorSDM or an
// SDMForce.
class Context {
public:
    Context(const System& system, LangevinIntegratorSDM& integrator);
    Context(const System& system, const SDMForce& force);
    ~Context();
    void setPositions(const std::vector<Vec3>& x);
    void setVelocities(const std::vector<Vec3>& v);
    void setTime(double t);
    State getState(bool positions, bool velocities, bool forces);
    sdm_context* handle() { return h_; }
    const System& getSystem() const { return sys_; }

private:
    friend class HIPIntegrateLangevinStepSDMKernel;
    friend class HIPCalcSDMForceKernel;
    friend class SDMForce;
    void create(const std::vector<double>& displacement);
    System sys_;
    sdm_context* h_ = nullptr;
    std::unique_ptr<CalcSDMForceKernel> force_kernel_;
    const SDMForce* force_ = nullptr;
};

// HIP implementations of the two kernel interfaces (through include/sdm_hip.h).
// HIPIntegrateLangevinStepSDMKernel::execute is one whole step: the dual-state pass
// evaluates U0, U1 and the group-1 energy itself, so it needs no force evaluation
// before it and does not read the three energy arguments (pass NaN); the energies it
// finds are in the integrator's getPotEnergy()/getBindE() afterwards.
class HIPIntegrateLangevinStepSDMKernel : public IntegrateLangevinStepSDMKernel {
public:
    void initialize(const System& system, const LangevinIntegratorSDM& integrator) override;
    void execute(Context& context, LangevinIntegratorSDM& integrator, double State1Energy,
                 double State2Energy, double BondAndRestraintEnergy) override;
    double computeKineticEnergy(Context& context, const LangevinIntegratorSDM& integrator) override;
};

class HIPCalcSDMForceKernel : public CalcSDMForceKernel {
public:
    void initialize(const System& system, const SDMForce& force) override;
    double execute(Context& context, bool includeForces, bool includeEnergy) override;
    void copyParametersToContext(Context& context, const SDMForce& force) override;
};

}  // namespace SDMPlugin
