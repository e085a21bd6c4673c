// SDMPlugin.cpp -- C++ host API of the SDM force over the HIP C ABI (see SDMPlugin.h).
#include "SDMPlugin.h"

#include <cmath>
#include <random>

namespace SDMPlugin {

namespace {
void check(int rc) {
    if (rc != SDM_OK) throw SDMException(sdm_last_error());
}
}  // namespace

// ---------------------------------------------------------------- System ----
int System::addParticle(double mass) {
    mass_.push_back(mass);
    q_.push_back(0.0);
    sig_.push_back(1.0);
    eps_.push_back(0.0);
    return (int)mass_.size() - 1;
}

void System::setParticleParameters(int i, double charge, double sigma, double epsilon) {
    if (i < 0 || i >= (int)mass_.size()) throw SDMException("particle index out of range");
    q_[i] = charge;
    sig_[i] = sigma;
    eps_[i] = epsilon;
}

void System::addExclusion(int a, int b) {
    excl_.push_back(a);
    excl_.push_back(b);
}

void System::addBond(int a, int b, double r0, double k) {
    bonds_.push_back(a);
    bonds_.push_back(b);
    bond_r0_.push_back(r0);
    bond_k_.push_back(k);
}

void System::addAngle(int a, int b, int c, double theta0, double k) {
    angles_.push_back(a);
    angles_.push_back(b);
    angles_.push_back(c);
    ang_t0_.push_back(theta0);
    ang_k_.push_back(k);
}

void System::setVirtualSite(int index, VirtualSite* site) {
    std::shared_ptr<const VirtualSite> owned(site);  // owned before anything can throw
    if (index < 0 || index >= (int)mass_.size()) throw SDMException("particle index out of range");
    if (!site) throw SDMException("null virtual site");
    vsites_[index] = std::move(owned);
}

const VirtualSite& System::getVirtualSite(int index) const {
    auto it = vsites_.find(index);
    if (it == vsites_.end()) throw SDMException("This particle is not a virtual site");
    return *it->second;
}

int VirtualSite::getParticle(int i) const {
    if (i < 0 || i >= (int)particles_.size()) throw SDMException("particle index out of range");
    return particles_[i];
}

double TwoParticleAverageSite::getWeight(int i) const {
    if (i < 0 || i > 1) throw SDMException("weight index out of range");
    return w_[i];
}

double ThreeParticleAverageSite::getWeight(int i) const {
    if (i < 0 || i > 2) throw SDMException("weight index out of range");
    return w_[i];
}

// --------------------------------------------------------- SDMParameters ----
// Defaults of the reference constructor, LangevinIntegratorSDM.cpp:48-85.
SDMParameters::SDMParameters(int nParticles) {
    p_ = sdm_params{};
    p_.bias_method = LinearMethod;
    p_.softcore_method = NoSoftCoreMethod;
    p_.lambda = 1.0;
    p_.gamma = 0.0;
    p_.wbcoeff = 1.0;
    p_.w0coeff = 0.0;
    p_.lambda1 = 1.0;
    p_.lambda2 = 1.0;
    p_.alpha = 1.0;
    p_.u0 = 0.0;
    p_.umax = 200.0;
    p_.acore = 0.25;
    p_.ubcore = 0.0;
    p_.nonequilibrium = 0;
    p_.noneq_tmax = 1.0;
    p_.work = 0.0;
    p_.temperature = 300.0;
    p_.friction = 1.0;
    p_.step_size = 0.001;
    displ_.assign(3 * (size_t)std::max(nParticles, 0), 0.0);
}

void SDMParameters::setDisplacement(int atom, double dx, double dy, double dz) {
    if (atom < 0 || 3 * (size_t)atom >= displ_.size()) throw SDMException("atom index out of range");
    displ_[3 * atom] = dx;
    displ_[3 * atom + 1] = dy;
    displ_[3 * atom + 2] = dz;
}

Vec3 SDMParameters::getDisplacement(int atom) const {
    if (atom < 0 || 3 * (size_t)atom >= displ_.size()) throw SDMException("atom index out of range");
    return {displ_[3 * atom], displ_[3 * atom + 1], displ_[3 * atom + 2]};
}

double SDMParameters::SoftCoreF(double u, double umax, double a, double ub, double& fp) const {
    if (u <= ub || p_.softcore_method == NoSoftCoreMethod) {
        fp = 1.0;
        return u;
    }
    if (p_.softcore_method == TanhMethod) {
        double t = std::tanh((u - ub) / umax);
        fp = 1.0 - t * t;
        return umax * t + ub;
    }
    if (p_.softcore_method == RationalMethod) {
        double gu = (u - ub) / (a * (umax - ub));
        double zeta = 1.0 + 2.0 * gu * (gu + 1.0);
        double zp = std::pow(zeta, a);
        fp = 4.0 * (2.0 * gu + 1.0) / zeta * zp / ((1.0 + zp) * (1.0 + zp));
        return (umax - ub) * (zp - 1.0) / (zp + 1.0) + ub;
    }
    throw SDMException("Unknown soft core method");
}

// ------------------------------------------------- LangevinIntegratorSDM ----
LangevinIntegratorSDM::LangevinIntegratorSDM(double temperature, double frictionCoeff,
                                             double stepSize, int nParticles)
    : SDMParameters(nParticles) {
    p_.temperature = temperature;
    p_.friction = frictionCoeff;
    p_.step_size = stepSize;
    std::random_device rd;
    p_.seed = rd();
}

void LangevinIntegratorSDM::initialize(Context& context) {
    if (context_ != nullptr && context_ != &context)
        throw SDMException("This Integrator is already bound to a context");
    context_ = &context;
    kernel_ = std::make_unique<HIPIntegrateLangevinStepSDMKernel>();
    kernel_->initialize(context.getSystem(), *this);
}

void LangevinIntegratorSDM::step(int steps) {
    if (!context_) throw SDMException("This Integrator is not bound to a context");
    if (p_.softcore_method < 0 || p_.softcore_method > 2) throw SDMException("Unknown soft core method");
    check(sdm_set_params(context_->handle(), 0, &p_));
    check(sdm_step(context_->handle(), steps));
    sdm_energies e;
    check(sdm_get_energies(context_->handle(), &e));
    bindE_ = e.bind_e;
    potE_ = e.pot_energy;
    check(sdm_get_params(context_->handle(), 0, &p_));
}

double LangevinIntegratorSDM::computeKineticEnergy() {
    if (!context_) throw SDMException("This Integrator is not bound to a context");
    return kernel_->computeKineticEnergy(*context_, *this);
}

// ---------------------------------------------------------------- SDMForce ----
double SDMForce::getPerturbationEnergy(Context& context) const {
    if (!context.force_kernel_) throw SDMException("the context is not driven by an SDMForce");
    context.force_kernel_->execute(context, false, true);
    sdm_energies e;
    check(sdm_get_energies(context.handle(), &e));
    return e.bind_e;
}

void SDMForce::updateParametersInContext(Context& context) const {
    if (!context.force_kernel_) throw SDMException("the context is not driven by an SDMForce");
    context.force_kernel_->copyParametersToContext(context, *this);
}

// ----------------------------------------------------------------- Context ----
Context::Context(const System& system, LangevinIntegratorSDM& integrator) : sys_(system) {
    if (integrator.displacements().size() != 3 * (size_t)system.getNumParticles())
        throw SDMException("integrator particle count does not match the system");
    create(integrator.displacements());
    try {
        integrator.initialize(*this);
        check(sdm_set_params(h_, 0, &integrator.raw()));
    } catch (...) {
        sdm_destroy(h_);  // the destructor does not run for a constructor that throws
        h_ = nullptr;
        throw;
    }
}

Context::Context(const System& system, const SDMForce& force) : sys_(system), force_(&force) {
    if (force.displacements().size() != 3 * (size_t)system.getNumParticles())
        throw SDMException("force particle count does not match the system");
    create(force.displacements());
    try {
        force_kernel_ = std::make_unique<HIPCalcSDMForceKernel>();
        force_kernel_->initialize(sys_, force);
        force_kernel_->copyParametersToContext(*this, force);
    } catch (...) {
        sdm_destroy(h_);
        h_ = nullptr;
        throw;
    }
}

Context::~Context() {
    if (h_) sdm_destroy(h_);
}

void Context::create(const std::vector<double>& displacement) {
    sdm_system_desc d{};
    d.n_atoms = sys_.getNumParticles();
    d.charge = sys_.q_.data();
    d.sigma = sys_.sig_.data();
    d.epsilon = sys_.eps_.data();
    d.mass = sys_.mass_.data();
    d.n_exclusions = (int32_t)(sys_.excl_.size() / 2);
    d.exclusions = sys_.excl_.data();
    d.n_bonds = (int32_t)(sys_.bonds_.size() / 2);
    d.bond_atoms = sys_.bonds_.data();
    d.bond_r0 = sys_.bond_r0_.data();
    d.bond_k = sys_.bond_k_.data();
    d.n_angles = (int32_t)(sys_.angles_.size() / 3);
    d.angle_atoms = sys_.angles_.data();
    d.angle_theta0 = sys_.ang_t0_.data();
    d.angle_k = sys_.ang_k_.data();
    d.displacement = displacement.data();
    for (int k = 0; k < 3; k++) d.box[k] = sys_.box_[k];
    d.periodic = sys_.periodic_ ? 1 : 0;
    d.cutoff = sys_.cutoff_;
    d.eps_rf = sys_.eps_rf_;
    d.lj_rule = sys_.lj_rule_;
    check(sdm_create(&d, 1, &h_));
    if (sys_.vsites_.empty()) return;
    std::vector<int32_t> kind, atoms;
    std::vector<double> w;
    for (const auto& [index, site] : sys_.vsites_) {
        kind.push_back(site->kind_);
        atoms.push_back(index);
        for (int k = 0; k < 3; k++)
            atoms.push_back(k < site->getNumParticles() ? site->particles_[k] : -1);
        w.insert(w.end(), site->w_.begin(), site->w_.end());
    }
    int rc = sdm_set_virtual_sites(h_, (int32_t)kind.size(), kind.data(), atoms.data(), w.data());
    if (rc != SDM_OK) {
        std::string msg = sdm_last_error();
        sdm_destroy(h_);  // the constructors' own clean-up covers only what follows create()
        h_ = nullptr;
        throw SDMException(msg);
    }
}

void Context::setPositions(const std::vector<Vec3>& x) {
    if ((int)x.size() != sys_.getNumParticles()) throw SDMException("wrong number of positions");
    check(sdm_set_positions(h_, 0, &x[0][0]));
}

void Context::setVelocities(const std::vector<Vec3>& v) {
    if ((int)v.size() != sys_.getNumParticles()) throw SDMException("wrong number of velocities");
    check(sdm_set_velocities(h_, 0, &v[0][0]));
}

void Context::setTime(double t) { check(sdm_set_time(h_, 0, t, 0)); }

State Context::getState(bool positions, bool velocities, bool forces) {
    State s;
    int n = sys_.getNumParticles();
    if (force_kernel_ && forces) force_kernel_->execute(*this, true, true);
    sdm_energies e;
    check(sdm_get_energies(h_, &e));
    s.potentialEnergy = e.pot_energy;
    s.perturbationEnergy = e.bind_e;
    s.U0 = e.U0;
    s.U1 = e.U1;
    s.time = e.time;
    if (positions) {
        s.positions.resize(n);
        check(sdm_get_positions(h_, 0, &s.positions[0][0]));
    }
    if (velocities) {
        s.velocities.resize(n);
        check(sdm_get_velocities(h_, 0, &s.velocities[0][0]));
    }
    if (forces) {
        s.forces.resize(n);
        check(sdm_get_forces(h_, 0, SDM_FORCE_HYBRID, &s.forces[0][0]));
    }
    return s;
}

// ------------------------------------------------------------ HIP kernels ----
void HIPIntegrateLangevinStepSDMKernel::initialize(const System&, const LangevinIntegratorSDM&) {}

void HIPIntegrateLangevinStepSDMKernel::execute(Context& context, LangevinIntegratorSDM& integrator,
                                                double, double, double) {
    check(sdm_set_params(context.handle(), 0, &integrator.raw()));
    check(sdm_step(context.handle(), 1));
    sdm_energies e;
    check(sdm_get_energies(context.handle(), &e));
    integrator.setBindE(e.bind_e);       // ReferenceSDMKernels.cpp:285
    integrator.setPotEnergy(e.pot_energy);  // :287
    check(sdm_get_params(context.handle(), 0, &integrator.raw()));  // schedule and work
}

double HIPIntegrateLangevinStepSDMKernel::computeKineticEnergy(Context& context,
                                                               const LangevinIntegratorSDM&) {
    double ke = 0.0;
    check(sdm_kinetic_energy(context.handle(), &ke));
    return ke;
}

void HIPCalcSDMForceKernel::initialize(const System&, const SDMForce&) {}

double HIPCalcSDMForceKernel::execute(Context& context, bool, bool) {
    check(sdm_compute(context.handle()));
    sdm_energies e;
    check(sdm_get_energies(context.handle(), &e));
    return e.pot_energy;
}

void HIPCalcSDMForceKernel::copyParametersToContext(Context& context, const SDMForce& force) {
    if (force.getSoftCoreMethod() < 0 || force.getSoftCoreMethod() > 2)
        throw SDMException("Unknown soft core method");
    check(sdm_set_params(context.handle(), 0, &force.raw()));
}

}  // namespace SDMPlugin
