/*
 * sdm_oracle.c -- FP64 CPU restatement of the SDM alchemical step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; never shipped or
 * measured as the product).  Function-by-function sources and pins are listed
 * in sdm_oracle.h.
 */
#include "sdm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* scalar stage                                                              */
/* ------------------------------------------------------------------------ */

/* LangevinIntegratorSDM::SoftCoreF, openmmapi/src/LangevinIntegratorSDM.cpp:125-149.
   Returns 0, or -1 for an unknown method (the reference throws). */
int orc_softcore(int method, double u, double umax, double a, double ub, double* up,
                 double* fp) {
    if (u <= ub) {
        *fp = 1.0;
        *up = u;
        return 0;
    }
    if (method == ORC_NOSOFTCORE) {
        *fp = 1.0;
        *up = u;
        return 0;
    }
    if (method == ORC_TANH) {
        double x = (u - ub) / umax;
        double t = tanh(x);
        *fp = 1.0 - t * t;
        *up = umax * t + ub;
        return 0;
    }
    if (method == ORC_RATIONAL) {
        double gu = (u - ub) / (a * (umax - ub));
        double zeta = 1.0 + 2.0 * gu * (gu + 1.0);
        double zetap = pow(zeta, a);
        double s = 4.0 * (2.0 * gu + 1.0) / zeta;
        *fp = s * zetap / pow(1.0 + zetap, 2);
        *up = (umax - ub) * (zetap - 1.0) / (zetap + 1.0) + ub;
        return 0;
    }
    return -1;
}

/* Scalar part of ReferenceIntegrateLangevinStepSDMKernel::execute,
   platforms/reference/src/ReferenceSDMKernels.cpp:202-309, including the
   nonequilibrium schedule (:221-245), the bias (:264-282), the energy bookkeeping
   (:283-287), the work increment (:289-302) and sp (:309).  The locals shadowed
   inside the nonequilibrium block (:236-239) are reproduced: the schedule is
   written back to the parameters, and the outer locals keep their defaults unless
   the ILogistic branch re-reads them. */
int orc_execute_scalars(orc_params* p, double E1, double E2, double Eb, double time,
                        orc_scalars* out) {
    double stepSize = p->step_size;
    double lambdac = p->lambda;
    double dlambdac = 0.0;
    double gamma = 0.0;
    double wbcoeff = lambdac;
    double w0coeff = 0.0;
    double lambda1 = lambdac;
    double lambda2 = lambdac;
    double alpha = 1.0;
    double u0 = 0.0;
    (void)wbcoeff;
    if (p->nonequilibrium == 1) {
        double Ntmax = p->noneq_tmax;
        lambdac = time / Ntmax;
        p->lambda = lambdac;
        p->lambda1 = p->m_l1 * lambdac + p->b_l1;
        p->lambda2 = p->m_l2 * lambdac + p->b_l2;
        p->u0 = p->m_u0 * lambdac + p->b_u0;
        p->w0coeff = p->m_w0 * lambdac + p->b_w0;
        dlambdac = stepSize / Ntmax;
    }
    /* the reference keeps alpha constant; with a slope it follows lambda */
    double alpha_now = p->alpha + (p->nonequilibrium == 1 ? p->m_alpha * lambdac : 0.0);
    int method = p->bias_method;
    if (method == ORC_QUADRATIC) {
        gamma = p->gamma;
        wbcoeff = p->wbcoeff;
        w0coeff = p->w0coeff;
    } else if (method == ORC_ILOGISTIC) {
        lambda1 = p->lambda1;
        lambda2 = p->lambda2;
        alpha = alpha_now;
        u0 = p->u0;
        w0coeff = p->w0coeff;
    }
    double fp, BindE;
    if (orc_softcore(p->softcore_method, E2 - E1, p->umax, p->acore, p->ubcore, &BindE, &fp))
        return -1;
    double bfp = 0.0, ebias = 0.0;
    if (method == ORC_QUADRATIC) {
        ebias = 0.5 * gamma * BindE * BindE + wbcoeff * BindE + w0coeff;
        bfp = gamma * BindE + wbcoeff;
    } else if (method == ORC_ILOGISTIC) {
        double ee = 1.0 + exp(-alpha * (BindE - u0));
        if (alpha > 0) ebias = ((lambda2 - lambda1) / alpha) * log(ee);
        ebias += lambda2 * BindE + w0coeff;
        bfp = (lambda2 - lambda1) / ee + lambda1;
    } else {
        ebias = lambdac * BindE;
        bfp = lambdac;
    }
    if (p->nonequilibrium == 1) {
        double ee = 1.0 + exp(-alpha * (BindE - u0));
        double dwdl1 = -log(ee) / alpha;
        double dwdl2 = BindE + (log(ee) / alpha);
        double dwdu0 = (lambda2 - lambda1) * exp(-alpha * (BindE - u0)) / ee;
        double dwdw0 = 1;
        /* d/dalpha of ((l2-l1)/alpha) ln(ee): the reference writes dwdalpha = 0 (:295),
           which is exact only while alpha does not change with lambda */
        double dwdalpha = (alpha > 0) ? -(lambda2 - lambda1) * log(ee) / (alpha * alpha) -
                                            (BindE - u0) * dwdu0 / alpha
                                      : 0.0;
        double dwdlambda = (dwdl1 * p->m_l1) + (dwdl2 * p->m_l2) + (dwdu0 * p->m_u0) +
                           (dwdw0 * p->m_w0) + dwdalpha * p->m_alpha;
        p->work += dlambdac * dwdlambda;
    }
    out->bind_e = BindE;
    out->pot_energy = E1 + ebias + Eb;
    out->ebias = ebias;
    out->bfp = bfp;
    out->fp = fp;
    out->sp = bfp * fp;
    return 0;
}

/* F_i = sp*F2_i + (1-sp)*F1_i + Fb_i, ReferenceSDMKernels.cpp:308-318. */
void orc_force_mix(int n, double sp, const double* F1, const double* F2, const double* Fb,
                   double* F) {
    for (int i = 0; i < 3 * n; i++) F[i] = sp * F2[i] + (1.0 - sp) * F1[i] + Fb[i];
}

/* ------------------------------------------------------------------------ */
/* nonbonded LJ + reaction field (OpenMM NonbondedForce, CutoffPeriodic /      */
/* CutoffNonPeriodic, documented formula)                                     */
/* ------------------------------------------------------------------------ */

/* Energy of one pair at distance r < rc, with -dE/dr / r in *f_over_r.
   qq = q_i q_j, sig = (s_i+s_j)/2, eps = sqrt(e_i e_j). */
double orc_pair_energy(double r, double qq, double sig, double eps, double rc, double eps_rf,
                       double* f_over_r) {
    double krf = (eps_rf - 1.0) / ((2.0 * eps_rf + 1.0) * rc * rc * rc);
    double crf = 3.0 * eps_rf / ((2.0 * eps_rf + 1.0) * rc);
    double ir = 1.0 / r;
    double ir2 = ir * ir;
    double s2 = sig * sig * ir2;
    double s6 = s2 * s2 * s2;
    double s12 = s6 * s6;
    double e_lj = 4.0 * eps * (s12 - s6);
    double f_lj = 4.0 * eps * (12.0 * s12 - 6.0 * s6) * ir2;
    double kq = ORC_ONE_4PI_EPS0 * qq;
    double e_c = kq * (ir + krf * r * r - crf);
    double f_c = kq * (ir * ir2 - 2.0 * krf);
    *f_over_r = f_lj + f_c;
    return e_lj + e_c;
}

/* per-atom sorted exclusion partner lists (CSR) */
typedef struct {
    int* start;
    int* part;
} excl_csr;

static int cmp_int(const void* a, const void* b) {
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

static void excl_build(const orc_nb* nb, excl_csr* e) {
    int n = nb->n;
    e->start = (int*)calloc((size_t)n + 1, sizeof(int));
    e->part = (int*)malloc(sizeof(int) * (size_t)(2 * nb->n_excl + 1));
    for (int k = 0; k < nb->n_excl; k++) {
        e->start[nb->excl[2 * k] + 1]++;
        e->start[nb->excl[2 * k + 1] + 1]++;
    }
    for (int i = 0; i < n; i++) e->start[i + 1] += e->start[i];
    int* fill = (int*)calloc((size_t)n, sizeof(int));
    for (int k = 0; k < nb->n_excl; k++) {
        int a = nb->excl[2 * k], b = nb->excl[2 * k + 1];
        e->part[e->start[a] + fill[a]++] = b;
        e->part[e->start[b] + fill[b]++] = a;
    }
    for (int i = 0; i < n; i++)
        qsort(e->part + e->start[i], (size_t)(e->start[i + 1] - e->start[i]), sizeof(int),
              cmp_int);
    free(fill);
}

static void excl_free(excl_csr* e) {
    free(e->start);
    free(e->part);
}

static int excl_has(const excl_csr* e, int i, int j) {
    for (int k = e->start[i]; k < e->start[i + 1]; k++)
        if (e->part[k] == j) return 1;
    return 0;
}

static inline double min_image(double d, double L) { return d - L * floor(d / L + 0.5); }

/* the pair's LJ sigma: Lorentz-Berthelot, or geometric (OPLS: the reference's DMS reader
   then moves LJ to CustomNonbondedForce "sigma12=sqrt(sigma1*sigma2)",
   desmonddmsfile75.py:780-793, with epsilon 0 left in NonbondedForce, :798-801) */
static inline double pair_sigma(const orc_nb* nb, int i, int j) {
    return nb->geometric ? sqrt(nb->sigma[i] * nb->sigma[j]) : 0.5 * (nb->sigma[i] + nb->sigma[j]);
}

/* One pair (i, j): adds forces, returns energy; 0 and no force when r >= rc. */
static inline double pair_ij(const orc_nb* nb, const double* xi, const double* xj, int i, int j,
                             double* Fi, double* Fj, int* inside) {
    double d[3];
    for (int k = 0; k < 3; k++) {
        d[k] = xi[k] - xj[k];
        if (nb->periodic) d[k] = min_image(d[k], nb->box[k]);
    }
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    *inside = 0;
    if (r2 >= nb->cutoff * nb->cutoff) return 0.0;
    *inside = 1;
    double fr;
    double e = orc_pair_energy(sqrt(r2), nb->q[i] * nb->q[j], pair_sigma(nb, i, j),
                               sqrt(nb->eps[i] * nb->eps[j]), nb->cutoff, nb->eps_rf, &fr);
    for (int k = 0; k < 3; k++) {
        if (Fi) Fi[k] += fr * d[k];
        if (Fj) Fj[k] -= fr * d[k];
    }
    return e;
}

double orc_nonbonded_brute(const orc_nb* nb, const double* x, double* F, long long* npairs) {
    excl_csr ex;
    excl_build(nb, &ex);
    int n = nb->n;
    if (F) memset(F, 0, sizeof(double) * 3 * (size_t)n);
    double E = 0.0;
    long long cnt = 0;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            if (excl_has(&ex, i, j)) continue;
            int in;
            E += pair_ij(nb, x + 3 * i, x + 3 * j, i, j, F ? F + 3 * i : NULL,
                         F ? F + 3 * j : NULL, &in);
            cnt += in;
        }
    excl_free(&ex);
    if (npairs) *npairs = cnt;
    return E;
}

double orc_nonbonded_cells(const orc_nb* nb, const double* x, double* F, long long* npairs) {
    int n = nb->n;
    int nc[3];
    for (int k = 0; k < 3; k++) {
        nc[k] = nb->periodic ? (int)floor(nb->box[k] / nb->cutoff) : 0;
        if (nc[k] < 3) return orc_nonbonded_brute(nb, x, F, npairs);
    }
    excl_csr ex;
    excl_build(nb, &ex);
    int ncell = nc[0] * nc[1] * nc[2];
    int* head = (int*)malloc(sizeof(int) * (size_t)ncell);
    int* next = (int*)malloc(sizeof(int) * (size_t)n);
    int* cell = (int*)malloc(sizeof(int) * (size_t)n);
    for (int c = 0; c < ncell; c++) head[c] = -1;
    for (int i = n - 1; i >= 0; i--) {
        int ci[3];
        for (int k = 0; k < 3; k++) {
            double w = x[3 * i + k] - nb->box[k] * floor(x[3 * i + k] / nb->box[k]);
            int c = (int)(w / nb->box[k] * nc[k]);
            if (c >= nc[k]) c = nc[k] - 1;
            if (c < 0) c = 0;
            ci[k] = c;
        }
        int c = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
        cell[i] = c;
        next[i] = head[c];
        head[c] = i;
    }
    if (F) memset(F, 0, sizeof(double) * 3 * (size_t)n);
    double E = 0.0;
    long long cnt = 0;
    for (int i = 0; i < n; i++) {
        int c = cell[i];
        int cz = c % nc[2], cy = (c / nc[2]) % nc[1], cx = c / (nc[1] * nc[2]);
        for (int ox = -1; ox <= 1; ox++)
            for (int oy = -1; oy <= 1; oy++)
                for (int oz = -1; oz <= 1; oz++) {
                    int c2 = (((cx + ox + nc[0]) % nc[0]) * nc[1] + (cy + oy + nc[1]) % nc[1]) *
                                 nc[2] +
                             (cz + oz + nc[2]) % nc[2];
                    for (int j = head[c2]; j >= 0; j = next[j]) {
                        if (j <= i) continue;
                        if (excl_has(&ex, i, j)) continue;
                        int in;
                        E += pair_ij(nb, x + 3 * i, x + 3 * j, i, j, F ? F + 3 * i : NULL,
                                     F ? F + 3 * j : NULL, &in);
                        cnt += in;
                    }
                }
    }
    free(head);
    free(next);
    free(cell);
    excl_free(&ex);
    if (npairs) *npairs = cnt;
    return E;
}

double orc_nonbonded(const orc_nb* nb, const double* x, double* F, long long* npairs) {
    return orc_nonbonded_cells(nb, x, F, npairs);
}

/* ------------------------------------------------------------------------ */
/* bonded (OpenMM HarmonicBondForce / HarmonicAngleForce, no PBC)             */
/* ------------------------------------------------------------------------ */
double orc_bonded(const orc_bonded_t* b, const double* x, double* F) {
    double E = 0.0;
    for (int k = 0; k < b->n_bonds; k++) {
        int i = b->bond_atoms[2 * k], j = b->bond_atoms[2 * k + 1];
        double d[3], r2 = 0;
        for (int c = 0; c < 3; c++) {
            d[c] = x[3 * j + c] - x[3 * i + c];
            r2 += d[c] * d[c];
        }
        double r = sqrt(r2);
        double dr = r - b->bond_r0[k];
        E += 0.5 * b->bond_k[k] * dr * dr;
        double g = (r > 0) ? b->bond_k[k] * dr / r : 0.0;
        for (int c = 0; c < 3; c++) {
            F[3 * i + c] += g * d[c];
            F[3 * j + c] -= g * d[c];
        }
    }
    for (int k = 0; k < b->n_angles; k++) {
        int a = b->angle_atoms[3 * k], m = b->angle_atoms[3 * k + 1], c3 = b->angle_atoms[3 * k + 2];
        double v1[3], v2[3], r1 = 0, r2 = 0, dot = 0;
        for (int c = 0; c < 3; c++) {
            v1[c] = x[3 * a + c] - x[3 * m + c];
            v2[c] = x[3 * c3 + c] - x[3 * m + c];
            r1 += v1[c] * v1[c];
            r2 += v2[c] * v2[c];
            dot += v1[c] * v2[c];
        }
        r1 = sqrt(r1);
        r2 = sqrt(r2);
        double cs = dot / (r1 * r2);
        if (cs > 1.0) cs = 1.0;
        if (cs < -1.0) cs = -1.0;
        double th = acos(cs);
        double dth = th - b->angle_t0[k];
        E += 0.5 * b->angle_k[k] * dth * dth;
        double sn = sin(th);
        if (sn < 1e-12) sn = 1e-12;
        /* dE/dtheta * dtheta/dx ; dtheta/dcos = -1/sin */
        double g = b->angle_k[k] * dth / sn; /* = -dE/dcos */
        for (int c = 0; c < 3; c++) {
            double dcos_da = v2[c] / (r1 * r2) - cs * v1[c] / (r1 * r1);
            double dcos_dc = v1[c] / (r1 * r2) - cs * v2[c] / (r2 * r2);
            /* F = -dE/dx = -(dE/dcos) dcos/dx = g * dcos/dx */
            F[3 * a + c] += g * dcos_da;
            F[3 * c3 + c] += g * dcos_dc;
            F[3 * m + c] -= g * (dcos_da + dcos_dc);
        }
    }
    return E;
}

/* ------------------------------------------------------------------------ */
/* dual state                                                                */
/* ------------------------------------------------------------------------ */

/* LangevinIntegratorSDM::step's three evaluations (LangevinIntegratorSDM.cpp:158-176):
   group 2 at x (E1, F1), group 2 at x + d (MakeState2, ReferenceSDMKernels.cpp:192-199;
   E2, F2), group 1 at the restored x (Eb, Fb). */
void orc_sdm_reference_eval(const orc_nb* nb, const orc_bonded_t* b, const double* x,
                            const double* displ, double* E1, double* E2, double* Eb,
                            double* F1, double* F2, double* Fb) {
    int n = nb->n;
    double* y = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    for (int i = 0; i < 3 * n; i++) y[i] = x[i] + displ[i];
    double e1 = orc_nonbonded(nb, x, F1, NULL);
    double e2 = orc_nonbonded(nb, y, F2, NULL);
    if (E1) *E1 = e1;
    if (E2) *E2 = e2;
    free(y);
    if (Fb) memset(Fb, 0, sizeof(double) * 3 * (size_t)n);
    if (b) {
        double* fb = Fb ? Fb : (double*)calloc(3 * (size_t)n, sizeof(double));
        double eb = orc_bonded(b, x, fb);
        if (Eb) *Eb = eb;
        if (!Fb) free(fb);
    } else if (Eb) {
        *Eb = 0.0;
    }
}

static inline int same_displ(const double* displ, int i, int j) {
    return displ[3 * i] == displ[3 * j] && displ[3 * i + 1] == displ[3 * j + 1] &&
           displ[3 * i + 2] == displ[3 * j + 2];
}

static inline int is_displaced(const double* displ, int i) {
    return displ[3 * i] != 0.0 || displ[3 * i + 1] != 0.0 || displ[3 * i + 2] != 0.0;
}

void orc_sdm_direct(const orc_nb* nb, const double* x, const double* displ, double* U0,
                    double* u, double* F0, double* dF, long long* npairs0,
                    long long* npairs_delta) {
    int n = nb->n;
    double e0 = orc_nonbonded(nb, x, F0, npairs0);
    if (U0) *U0 = e0;
    if (dF) memset(dF, 0, sizeof(double) * 3 * (size_t)n);
    excl_csr ex;
    excl_build(nb, &ex);
    double du = 0.0;
    long long cnt = 0;
    double yi[3], yj[3];
    for (int i = 0; i < n; i++) {
        if (!is_displaced(displ, i)) continue;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            if (same_displ(displ, i, j)) continue;
            if (is_displaced(displ, j) && j < i) continue; /* D x D pair once */
            if (excl_has(&ex, i, j)) continue;
            for (int k = 0; k < 3; k++) {
                yi[k] = x[3 * i + k] + displ[3 * i + k];
                yj[k] = x[3 * j + k] + displ[3 * j + k];
            }
            double fyi[3] = {0, 0, 0}, fyj[3] = {0, 0, 0}, fxi[3] = {0, 0, 0}, fxj[3] = {0, 0, 0};
            int iny, inx;
            double ey = pair_ij(nb, yi, yj, i, j, fyi, fyj, &iny);
            double ex0 = pair_ij(nb, x + 3 * i, x + 3 * j, i, j, fxi, fxj, &inx);
            cnt += iny;
            du += ey - ex0;
            if (dF)
                for (int k = 0; k < 3; k++) {
                    dF[3 * i + k] += fyi[k] - fxi[k];
                    dF[3 * j + k] += fyj[k] - fxj[k];
                }
        }
    }
    excl_free(&ex);
    if (u) *u = du;
    if (npairs_delta) *npairs_delta = cnt;
}

/* d u / d(q_s, sigma_s, eps_s) for the selected atoms sel[0..nsel-1]. */
void orc_du_dparams(const orc_nb* nb, const double* x, const double* displ, int nsel,
                    const int* sel, double* dq, double* dsig, double* deps) {
    int n = nb->n;
    int* pos = (int*)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) pos[i] = -1;
    for (int s = 0; s < nsel; s++) {
        pos[sel[s]] = s;
        dq[s] = dsig[s] = deps[s] = 0.0;
    }
    excl_csr ex;
    excl_build(nb, &ex);
    double rc = nb->cutoff, erf_ = nb->eps_rf;
    double krf = (erf_ - 1.0) / ((2.0 * erf_ + 1.0) * rc * rc * rc);
    double crf = 3.0 * erf_ / ((2.0 * erf_ + 1.0) * rc);
    for (int i = 0; i < n; i++) {
        if (!is_displaced(displ, i)) continue;
        for (int j = 0; j < n; j++) {
            if (j == i || same_displ(displ, i, j)) continue;
            if (is_displaced(displ, j) && j < i) continue;
            if (excl_has(&ex, i, j)) continue;
            if (pos[i] < 0 && pos[j] < 0) continue;
            double sig = pair_sigma(nb, i, j);
            double eps = sqrt(nb->eps[i] * nb->eps[j]);
            /* d sigma_ij / d sigma_i: 1/2, or sigma_j / (2 sigma_ij) for the geometric rule */
            double dsi = nb->geometric ? (sig > 0 ? nb->sigma[j] / (2.0 * sig) : 0.0) : 0.5;
            double dsj = nb->geometric ? (sig > 0 ? nb->sigma[i] / (2.0 * sig) : 0.0) : 0.5;
            /* per-geometry partials: g = +1 at y, -1 at x */
            double dEdqq = 0, dEdsig = 0, dEdeps = 0;
            for (int geo = 0; geo < 2; geo++) {
                double d[3], r2 = 0;
                for (int k = 0; k < 3; k++) {
                    double a = x[3 * i + k] + (geo ? displ[3 * i + k] : 0.0);
                    double b = x[3 * j + k] + (geo ? displ[3 * j + k] : 0.0);
                    d[k] = a - b;
                    if (nb->periodic) d[k] = min_image(d[k], nb->box[k]);
                    r2 += d[k] * d[k];
                }
                if (r2 >= rc * rc) continue;
                double r = sqrt(r2), ir = 1.0 / r;
                double sr = sig * ir;
                double s6 = pow(sr, 6), s12 = s6 * s6;
                double sgn = geo ? 1.0 : -1.0;
                dEdqq += sgn * ORC_ONE_4PI_EPS0 * (ir + krf * r2 - crf);
                if (sig > 0) dEdsig += sgn * 4.0 * eps * (12.0 * s12 - 6.0 * s6) / sig;
                dEdeps += sgn * 4.0 * (s12 - s6);
            }
            if (pos[i] >= 0) {
                int s = pos[i];
                dq[s] += dEdqq * nb->q[j];
                dsig[s] += dsi * dEdsig;
                if (eps > 0) deps[s] += dEdeps * nb->eps[j] / (2.0 * eps);
            }
            if (pos[j] >= 0) {
                int s = pos[j];
                dq[s] += dEdqq * nb->q[i];
                dsig[s] += dsj * dEdsig;
                if (eps > 0) deps[s] += dEdeps * nb->eps[i] / (2.0 * eps);
            }
        }
    }
    excl_free(&ex);
    free(pos);
}

/* ------------------------------------------------------------------------ */
/* dynamics                                                                  */
/* ------------------------------------------------------------------------ */

/* ReferenceStochasticDynamicsSDM::update (ReferenceStochasticDynamicsSDM.cpp:131-266)
   with tau = 1/friction (0 when friction == 0, ReferenceSDMKernels.cpp:325); no
   constraints, no virtual sites.  xi holds 3n standard normal draws. */
void orc_langevin_update(int n, double* x, double* v, const double* F, const double* mass,
                         double dt, double friction, double temperature, const double* xi) {
    double tau = (friction == 0.0) ? 0.0 : 1.0 / friction;
    double vscale = exp(-dt / tau);
    double fscale = (1 - vscale) * tau;
    double kT = ORC_BOLTZ * temperature;
    double noisescale = sqrt(2 * kT / tau) * sqrt(0.5 * (1 - vscale * vscale) * tau);
    for (int i = 0; i < n; i++) {
        if (mass[i] == 0.0) continue;
        double im = 1.0 / mass[i];
        double sim = sqrt(im);
        for (int k = 0; k < 3; k++)
            v[3 * i + k] = vscale * v[3 * i + k] + fscale * im * F[3 * i + k] +
                           noisescale * sim * xi[3 * i + k];
    }
    double inv = 1.0 / dt;
    for (int i = 0; i < n; i++) {
        if (mass[i] == 0.0) continue;
        for (int k = 0; k < 3; k++) {
            double xp = x[3 * i + k] + dt * v[3 * i + k];
            v[3 * i + k] = inv * (xp - x[3 * i + k]);
            x[3 * i + k] = xp;
        }
    }
}

/* OpenCL form: OpenCLSDMKernels.cpp:331-355 parameters, langevin.cl:7-69 update (double
   precision branch).  Well defined at friction == 0 (vscale 1, fscale dt, no noise). */
void orc_langevin_update_cl(int n, double* x, double* v, const double* F, const double* mass,
                            double dt, double friction, double temperature, const double* xi) {
    double kT = ORC_BOLTZ * temperature;
    double vscale = exp(-dt * friction);
    double fscale = (friction == 0 ? dt : (1 - vscale) / friction);
    double noisescale = sqrt(kT * (1 - vscale * vscale));
    double inv = 1.0 / dt;
    for (int i = 0; i < n; i++) {
        if (mass[i] == 0.0) continue;
        double im = 1.0 / mass[i];
        double sim = sqrt(im);
        for (int k = 0; k < 3; k++) {
            double vel = vscale * v[3 * i + k] + fscale * im * F[3 * i + k] +
                         noisescale * sim * xi[3 * i + k];
            double delta = dt * vel;
            x[3 * i + k] += delta;
            v[3 * i + k] = inv * delta;
        }
    }
}

/* computeShiftedKineticEnergy without constraints, ReferenceSDMKernels.cpp:108-135. */
double orc_kinetic_energy(int n, const double* v, const double* mass) {
    double e = 0.0;
    for (int i = 0; i < n; i++)
        if (mass[i] > 0)
            e += mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] +
                            v[3 * i + 2] * v[3 * i + 2]);
    return 0.5 * e;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller           */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

void orc_gauss3(uint64_t seed, uint32_t replica, uint64_t step, uint32_t atom, double out[3]) {
    uint32_t ctr[4] = {atom, replica, (uint32_t)step, (uint32_t)(step >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const double twom32 = 2.3283064365386963e-10; /* 2^-32 */
    const double twopi = 6.283185307179586;
    double u1 = ((double)o[0] + 1.0) * twom32;
    double u2 = (double)o[1] * twom32;
    double u3 = ((double)o[2] + 1.0) * twom32;
    double u4 = (double)o[3] * twom32;
    double r1 = sqrt(-2.0 * log(u1));
    double r2 = sqrt(-2.0 * log(u3));
    out[0] = r1 * cos(twopi * u2);
    out[1] = r1 * sin(twopi * u2);
    out[2] = r2 * cos(twopi * u4);
}

/* ------------------------------------------------------------------------ */
/* whole step (LangevinIntegratorSDM::step, LangevinIntegratorSDM.cpp:153-183)  */
/* ------------------------------------------------------------------------ */
int orc_ref_step(const orc_nb* nb, const orc_bonded_t* b, const double* displ,
                 const double* mass, orc_params* p, double* x, double* v, double* time,
                 uint64_t* step, uint64_t seed, uint32_t replica, int nsteps,
                 int use_cl_langevin, orc_step_report* last) {
    int n = nb->n;
    size_t sz = sizeof(double) * 3 * (size_t)n;
    double* F1 = (double*)malloc(sz);
    double* F2 = (double*)malloc(sz);
    double* Fb = (double*)malloc(sz);
    double* F = (double*)malloc(sz);
    double* xi = (double*)malloc(sz);
    int rc = 0;
    for (int s = 0; s < nsteps; s++) {
        double E1, E2, Eb;
        orc_sdm_reference_eval(nb, b, x, displ, &E1, &E2, &Eb, F1, F2, Fb);
        orc_scalars sc;
        if (orc_execute_scalars(p, E1, E2, Eb, *time, &sc)) {
            rc = -1;
            break;
        }
        orc_force_mix(n, sc.sp, F1, F2, Fb, F);
        for (int i = 0; i < n; i++) orc_gauss3(seed, replica, *step, (uint32_t)i, xi + 3 * i);
        if (use_cl_langevin)
            orc_langevin_update_cl(n, x, v, F, mass, p->step_size, p->friction, p->temperature,
                                   xi);
        else
            orc_langevin_update(n, x, v, F, mass, p->step_size, p->friction, p->temperature, xi);
        *time += p->step_size;
        *step += 1;
        if (last) {
            last->E1 = E1;
            last->E2 = E2;
            last->Eb = Eb;
            last->s = sc;
        }
    }
    free(F1);
    free(F2);
    free(Fb);
    free(F);
    free(xi);
    return rc;
}

/* R independent replicas of orc_ref_step on `threads` OpenMP threads, one replica per
   thread at a time (the CPU baseline's one-replica-per-core form).  x, v: [R][3n];
   params, time, step, seed: [R].  Replica r draws the noise of replica index r. */
int orc_ref_step_replicas(const orc_nb* nb, const orc_bonded_t* b, const double* displ,
                          const double* mass, orc_params* params, double* x, double* v,
                          double* time, uint64_t* step, const uint64_t* seed, int R,
                          int nsteps, int use_cl_langevin, int threads) {
    size_t n3 = 3 * (size_t)nb->n;
    int rc = 0;
#pragma omp parallel for num_threads(threads) schedule(static, 1) reduction(| : rc)
    for (int r = 0; r < R; r++)
        rc |= orc_ref_step(nb, b, displ, mass, &params[r], x + r * n3, v + r * n3, &time[r],
                           &step[r], seed[r], (uint32_t)r, nsteps, use_cl_langevin, NULL);
    return rc;
}

