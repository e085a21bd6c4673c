// Drives the C++ host API (SDMPlugin.h) the way a reference user drives
// LangevinIntegratorSDM (example/test.py:172-214): read a system, bind an integrator,
// step, read BindE/PotEnergy.  Prints one JSON object.
//   usage: facade_driver <system.txt> <nsteps>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

#include "SDMPlugin.h"

using namespace SDMPlugin;

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: facade_driver <system.txt> <nsteps>\n";
        return 2;
    }
    std::ifstream in(argv[1]);
    int n;
    double rc, erf, bx, by, bz;
    in >> n >> rc >> erf >> bx >> by >> bz;
    System sys;
    std::vector<Vec3> x(n), v(n), d(n);
    for (int i = 0; i < n; i++) {
        double q, s, e, m;
        in >> q >> s >> e >> m >> d[i][0] >> d[i][1] >> d[i][2] >> x[i][0] >> x[i][1] >> x[i][2] >>
            v[i][0] >> v[i][1] >> v[i][2];
        sys.addParticle(m);
        sys.setParticleParameters(i, q, s, e);
    }
    int ne, nb, na;
    in >> ne;
    for (int k = 0; k < ne; k++) {
        int a, b;
        in >> a >> b;
        sys.addExclusion(a, b);
    }
    in >> nb;
    for (int k = 0; k < nb; k++) {
        int a, b;
        double r0, kk;
        in >> a >> b >> r0 >> kk;
        sys.addBond(a, b, r0, kk);
    }
    in >> na;
    for (int k = 0; k < na; k++) {
        int a, b, c;
        double t0, kk;
        in >> a >> b >> c >> t0 >> kk;
        sys.addAngle(a, b, c, t0, kk);
    }
    // optional tail: the combining rule, then virtual sites as
    // "kind site p1 p2 p3 w1 w2 w3" (kind 0: two-particle average, p3 unused; 1: three-
    // particle average; 2: out of plane)
    int lj_rule = 0, nvs = 0;
    if (in >> lj_rule >> nvs) {
        sys.setCombiningRule(lj_rule ? System::Geometric : System::LorentzBerthelot);
        for (int k = 0; k < nvs; k++) {
            int kind, s, p1, p2, p3;
            double w1, w2, w3;
            in >> kind >> s >> p1 >> p2 >> p3 >> w1 >> w2 >> w3;
            if (kind == 0)
                sys.setVirtualSite(s, new TwoParticleAverageSite(p1, p2, w1, w2));
            else if (kind == 1)
                sys.setVirtualSite(s, new ThreeParticleAverageSite(p1, p2, p3, w1, w2, w3));
            else
                sys.setVirtualSite(s, new OutOfPlaneSite(p1, p2, p3, w1, w2, w3));
        }
    }
    sys.setPeriodicBoxVectors({bx, by, bz});
    sys.setCutoffDistance(rc);
    sys.setReactionFieldDielectric(erf);
    int nsteps = std::stoi(argv[2]);
    try {
        // force-level evaluation (SDMForce / CalcSDMForceKernel)
        SDMForce force(n);
        force.setBiasMethod(SDMForce::ILogisticMethod);
        force.setSoftCoreMethod(SDMForce::RationalMethod);
        force.setLambda1(0.2);
        force.setLambda2(0.6);
        force.setAlpha(0.1);
        force.setU0(110.0);
        force.setUmax(200.0);
        force.setAcore(0.0625);
        force.setUbcore(100.0);
        for (int i = 0; i < n; i++) force.setDisplacement(i, d[i][0], d[i][1], d[i][2]);
        Context fctx(sys, force);
        fctx.setPositions(x);
        double uprime = force.getPerturbationEnergy(fctx);
        State fs = fctx.getState(false, false, true);
        double fsum = 0;
        for (auto& f : fs.forces) fsum += f[0] + 2 * f[1] + 3 * f[2];

        // integrator (LangevinIntegratorSDM / IntegrateLangevinStepSDMKernel)
        LangevinIntegratorSDM integ(300.0, 1.0, 0.001, n);
        integ.setBiasMethod(LangevinIntegratorSDM::ILogisticMethod);
        integ.setSoftCoreMethod(LangevinIntegratorSDM::RationalMethod);
        integ.setLambda1(0.2);
        integ.setLambda2(0.6);
        integ.setAlpha(0.1);
        integ.setU0(110.0);
        integ.setUmax(200.0);
        integ.setAcore(0.0625);
        integ.setUbcore(100.0);
        integ.setRandomNumberSeed(4242);
        for (int i = 0; i < n; i++) integ.setDisplacement(i, d[i][0], d[i][1], d[i][2]);
        Context ctx(sys, integ);
        ctx.setPositions(x);
        ctx.setVelocities(v);
        integ.step(nsteps);
        State st = ctx.getState(true, true, false);
        double xsum = 0;
        for (auto& p : st.positions) xsum += p[0] + p[1] + p[2];
        std::string bound_error;
        try {
            Context other(sys, integ);
        } catch (const SDMException& e) {
            bound_error = e.what();
        }
        std::string softcore_error;
        integ.setSoftCoreMethod(7);
        try {
            integ.step(1);
        } catch (const SDMException& e) {
            softcore_error = e.what();
        }
        std::printf(
            "{\"force_uprime\": %.17g, \"force_U0\": %.17g, \"force_checksum\": %.17g, "
            "\"bind_e\": %.17g, \"pot_energy\": %.17g, \"x_checksum\": %.17g, \"time\": %.17g, "
            "\"ke\": %.17g, \"bound_error\": \"%s\", \"softcore_error\": \"%s\"}\n",
            uprime, fs.U0, fsum, integ.getBindE(), integ.getPotEnergy(), xsum, st.time,
            integ.computeKineticEnergy(), bound_error.c_str(), softcore_error.c_str());
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
