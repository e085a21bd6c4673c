// Host-side checks of the C++ API (SDMPlugin.h) that need no GPU: the System's virtual
// sites (OpenMM's ownership and accessors) and combining rule, and the integrator's
// defaults and setters (LangevinIntegratorSDM.cpp:48-85).  No Context is created.
// Prints "ok" or the first failed check; exit status 0 only when all hold.
#include <cstdio>
#include <string>

#include "SDMPlugin.h"

using namespace SDMPlugin;

static int failures = 0;
#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            std::printf("failed: %s (line %d)\n", #cond, __LINE__);   \
            failures++;                                               \
        }                                                             \
    } while (0)

template <class F>
static std::string error_of(F f) {
    try {
        f();
    } catch (const SDMException& e) {
        return e.what();
    }
    return "";
}

int main() {
    System sys;
    for (int i = 0; i < 6; i++) sys.addParticle(i < 4 ? 16.0 : 0.0);
    CHECK(sys.getCombiningRule() == System::LorentzBerthelot);
    sys.setCombiningRule(System::Geometric);
    CHECK(sys.getCombiningRule() == System::Geometric);
    CHECK(System::Geometric == SDM_LJ_GEOMETRIC && System::LorentzBerthelot == SDM_LJ_ARITHMETIC);

    sys.setVirtualSite(4, new ThreeParticleAverageSite(0, 1, 2, 0.744, 0.128, 0.128));
    sys.setVirtualSite(5, new OutOfPlaneSite(1, 2, 3, 0.25, -0.5, 2.0));
    CHECK(sys.isVirtualSite(4) && sys.isVirtualSite(5) && !sys.isVirtualSite(0));
    const auto& s4 = dynamic_cast<const ThreeParticleAverageSite&>(sys.getVirtualSite(4));
    CHECK(s4.getNumParticles() == 3 && s4.getParticle(0) == 0 && s4.getParticle(2) == 2);
    CHECK(s4.getWeight(0) == 0.744 && s4.getWeight(2) == 0.128);
    const auto& s5 = dynamic_cast<const OutOfPlaneSite&>(sys.getVirtualSite(5));
    CHECK(s5.getWeight12() == 0.25 && s5.getWeight13() == -0.5 && s5.getWeightCross() == 2.0);
    CHECK(s5.getParticle(0) == 1 && s5.getParticle(2) == 3);
    TwoParticleAverageSite two(0, 3, 0.7, 0.3);
    CHECK(two.getNumParticles() == 2 && two.getWeight(1) == 0.3);
    CHECK(!error_of([&] { two.getWeight(2); }).empty());
    CHECK(!error_of([&] { two.getParticle(2); }).empty());
    CHECK(error_of([&] { sys.getVirtualSite(0); }) == "This particle is not a virtual site");
    // an index out of range is refused; the site is owned (and freed) either way
    CHECK(!error_of([&] { sys.setVirtualSite(9, new TwoParticleAverageSite(0, 1, 0.5, 0.5)); }).empty());
    CHECK(!sys.isVirtualSite(9));
    // replacing a site keeps one definition per particle
    sys.setVirtualSite(5, new TwoParticleAverageSite(1, 2, 0.5, 0.5));
    CHECK(sys.getVirtualSite(5).getNumParticles() == 2);
    // Systems copy (a Context keeps its own): the copy shares the sites
    System copy = sys;
    CHECK(copy.isVirtualSite(4) && &copy.getVirtualSite(4) == &sys.getVirtualSite(4));

    // integrator defaults of the reference constructor and the method constants
    LangevinIntegratorSDM integ(300.0, 1.0, 0.001, 6);
    CHECK(integ.getTemperature() == 300.0 && integ.getFriction() == 1.0);
    CHECK(integ.getStepSize() == 0.001);
    CHECK(integ.getLambda() == 1.0 && integ.getLambda1() == 1.0 && integ.getLambda2() == 1.0);
    CHECK(integ.getAlpha() == 1.0 && integ.getU0() == 0.0 && integ.getUmax() == 200.0);
    CHECK(integ.getAcore() == 0.25 && integ.getUbcore() == 0.0);
    CHECK(integ.getBiasMethod() == LangevinIntegratorSDM::LinearMethod);
    CHECK(integ.getSoftCoreMethod() == LangevinIntegratorSDM::NoSoftCoreMethod);
    integ.setLambda1(0.2);
    integ.setU0(110.0);
    CHECK(integ.getLambda1() == 0.2 && integ.getU0() == 110.0);
    CHECK(!error_of([&] { integ.setDisplacement(6, 1.0, 0.0, 0.0); }).empty());
    if (failures == 0) std::printf("ok\n");
    return failures == 0 ? 0 : 1;
}
