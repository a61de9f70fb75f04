/*
 * ebb_oracle.c -- CPU restatement of EbbCFD's explicit residual-and-update path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (ebbcfd_amd/csrc) never links or calls it.
 *
 * It restates, in the reference's own edge-loop form and operation order:
 *   source/mesh.d:217-643          buildMesh (edges, normals, areas, centroids, ghosts, LSQ)
 *   source/solve.d:175-279         ufvmSetup (IC + boundary binding)
 *   source/solve.d:283-319         updateBoundaryEdge (with its lim[0]/cellIdx[1] quirk)
 *   source/solve.d:322-1122        ufvmSolver (ghosts, gradients, limiter, LP, fluxes, residual)
 *   source/solve.d:27-173          invert, LP
 *   source/flux.d:19-303           convectiveFlux, diffusiveFlux, rusanovFlux, roeFlux
 *   source/euler.d:89-94           gamma, getPressure
 *   source/integrators.d:22-270    Euler, RK4, RK2_TVD, RK2
 *   source/limiters.d:9-37         1-D slope limiters
 *   source/manufacturedsolution.d  solution, solutionGradient, sourceTerm
 *   app/ms.d:131-141               addSourceTerm (source added after the -1/area scaling)
 *   app/solver.d:38-277            runIntegrator (RMS residuals, forces, cflAdjust, NaN scan)
 *   Phobos std.algorithm.sum       pairwise summation used by the RMS residuals
 *
 * Choices where the reference's third-party arithmetic (numd) is not available
 * to pin the last ulp ("parity unpinned" at the ulp level, pinned by the
 * invariant and known-answer tests to ~1e-15):
 *   - Matrix*Vector sums run over every column in ascending order from 0.0;
 *   - normalize(v) = v / sqrt(v0*v0 + v1*v1), component-wise division;
 *   - 2x2 inverse = adjugate * (1.0/det), as the reference's own invert();
 *   - x^^2.0 and x^^2 lower to x*x, x^^3.0 to x*x*x;
 *   - constant sub-expressions are evaluated in IEEE double, left to right.
 * Build with -O2 -ffp-contract=off (no FMA contraction, IEEE div/sqrt).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>

/* The per-cell and per-edge loops of one residual evaluation and of the stage updates run on
   OpenMP threads: each iteration writes only its own cell or edge, min/max reductions are exact
   in any order, and an error reports the first failing index in loop order, so the results are
   those of the serial loops bit for bit (tests/test_oracle_threads.py checks it). */
#define ORC_FIRST(v, kv, x, kind) do { _Pragma("omp critical(orc_first)") { if ((x) < (v)) { (v) = (x); (kv) = (kind); } } } while (0)

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

#define MAXE 6
#define GAMMA 1.4

enum { BT_FULLSTATE = 0, BT_CONSTP, BT_INVWALL, BT_VISCWALL, BT_SYMMETRY, BT_TPINFLOW, BT_DIRICHLET };
enum { FLUX_RUSANOV = 0, FLUX_ROE = 1 };
enum { INT_EULER = 0, INT_RK2, INT_RK2TVD, INT_RK4 };
enum { ORC_OK = 0, ORC_ERR_EDGE_NAN = 1, ORC_ERR_CELL_NAN = 2, ORC_ERR_BC = 3, ORC_ERR_ARG = 4, ORC_ERR_MESH = 5 };

typedef struct {
    uint32_t nodeIdx[2];
    double len, sMax;
    double normal[2], mid[2];
    uint32_t cellIdx[2];
    double q[2][4];
    double flux[4];
    int isBoundary;
    int bGroup;
    int bIdx;
    int bType;
    double bData[4];
    int nbData;
} OEdge;

typedef struct {
    uint32_t edges[MAXE];
    double fluxMultiplier[MAXE];
    uint32_t neighborCells[MAXE];
    double gradMat[2][MAXE];
    double gradient[4][2];
    double lim[4][2];
    double minQ[4], maxQ[4];
    uint32_t nNeighborCells, nEdges;
    double area, d, perim;
    double centroid[2];
    double dt;
} OCell;

typedef struct {
    int flux, integrator, order, limited, localTimestep, multistageLimiting;
    int cflAdjust, viscosity, dynamicDt, mms;
    double dt, tEnd, CFL, lpThresh;
    double ic[4];
    double Pr, R, Re, L, mu;
    int nBC;
    int bcType[32];
    double bcData[32][4];
    int bcNData[32];
    int tagToBc[32];      /* mesh boundary tag index -> config boundary index */
    int nForce;
    int forceTags[32];    /* mesh boundary tag indices, in forceBoundary order */
    int limiter;          /* 0: the reference's scalar limiter; 1: Venkatakrishnan (build option) */
    double venkatK;       /* Venkatakrishnan constant: eps^2 = (K d)^3, d = 2 area / perimeter */
} OrcConfig;

typedef struct {
    uint32_t nNodes, nElem, nCells, nEdges;
    double *nodes;          /* 2*nNodes */
    uint32_t *eoff, *enodes;/* elements, 1-based node ids */
    OCell *cells;
    OEdge *edges;
    double (*q)[4];
    uint32_t *interiorEdges, nInteriorEdges;
    uint32_t *boundaryEdges, nBoundaryEdges;
    uint32_t nTags;
    uint32_t **bGroups, *bGroupLen;
    /* driver scratch */
    double (*R)[4], (*k1)[4], (*k2)[4], (*k3)[4], (*k4)[4], (*tmp)[4];
    double (*src)[4];       /* MMS source at interior centroids */
    int errKind;
    int64_t errIndex;
} OrcMesh;

/* ------------------------------------------------------------------ */
/* source/euler.d:89-94                                                */
/* ------------------------------------------------------------------ */
double orc_get_pressure(const double q[4])
{
    double a = q[1] / q[0];
    double b = q[2] / q[0];
    return (GAMMA - 1) * (q[3] - 0.5 * q[0] * (a * a + b * b));
}

/* ------------------------------------------------------------------ */
/* source/manufacturedsolution.d:27-283                                */
/* ------------------------------------------------------------------ */
static const double ar = 0.9, br = 0.04, cr = -2, dr = 1;
static const double au = 0.1, bu = 0.02, cu = 1, du_ = 1;
static const double av = 0.05, bv = -0.1, cv = 0.7, dv = 1.3;
static const double ap = 1, bp = 0.05, cp = 2, dp = -1;

static double m_rho(double x, double y) { return ar + br * sin(cr * x + dr * y); }
static double m_drdx(double x, double y) { return br * cr * cos(cr * x + dr * y); }
static double m_d2rdx2(double x, double y) { return -br * (cr * cr) * sin(cr * x + dr * y); }
static double m_drdy(double x, double y) { return br * dr * cos(cr * x + dr * y); }
static double m_d2rdy2(double x, double y) { return -br * (dr * dr) * sin(cr * x + dr * y); }
static double m_u(double x, double y) { return au + bu * cos(cu * x + du_ * y); }
static double m_dudx(double x, double y) { return -bu * cu * sin(cu * x + du_ * y); }
static double m_d2udxy(double x, double y) { return -bu * cu * du_ * cos(cu * x + du_ * y); }
static double m_d2udx2(double x, double y) { return -bu * (cu * cu) * cos(cu * x + du_ * y); }
static double m_dudy(double x, double y) { return -bu * du_ * sin(cu * x + du_ * y); }
static double m_d2udy2(double x, double y) { return -bu * (du_ * du_) * cos(cu * x + du_ * y); }
static double m_v(double x, double y) { return av + bv * cos(cv * x + dv * y); }
static double m_dvdx(double x, double y) { return -bv * cv * sin(cv * x + dv * y); }
static double m_dvdy(double x, double y) { return -bv * dv * sin(cv * x + dv * y); }
static double m_d2vdx2(double x, double y) { return -bv * (cv * cv) * cos(cv * x + dv * y); }
static double m_d2vdy2(double x, double y) { return -bv * (dv * dv) * cos(cv * x + dv * y); }
static double m_d2vdxy(double x, double y) { return -bv * cv * dv * cos(cv * x + dv * y); }
static double m_p(double x, double y) { return ap + bp * sin(cp * x + dp * y); }
static double m_dpdx(double x, double y) { return bp * cp * cos(cp * x + dp * y); }
static double m_d2pdx2(double x, double y) { return -bp * (cp * cp) * sin(cp * x + dp * y); }
static double m_dpdy(double x, double y) { return bp * dp * cos(cp * x + dp * y); }
static double m_d2pdy2(double x, double y) { return -bp * (dp * dp) * sin(cp * x + dp * y); }

/* manufacturedsolution.d:244-254 */
void orc_solution(double x, double y, double q[4])
{
    q[0] = m_rho(x, y);
    q[1] = m_rho(x, y) * m_u(x, y);
    q[2] = m_rho(x, y) * m_v(x, y);
    double uu = m_u(x, y), vv = m_v(x, y);
    q[3] = m_p(x, y) / (GAMMA - 1.0) + 0.5 * m_rho(x, y) * (uu * uu + vv * vv);
}

/* manufacturedsolution.d:256-283; grad[4][2] row-major (var, dir) */
void orc_solution_gradient(double x, double y, double g[8])
{
    double r = m_rho(x, y), u = m_u(x, y), v = m_v(x, y);
    double drdx = m_drdx(x, y), drdy = m_drdy(x, y);
    double dudx = m_dudx(x, y), dudy = m_dudy(x, y);
    double dvdx = m_dvdx(x, y), dvdy = m_dvdy(x, y);
    g[0] = drdx;
    g[1] = drdy;
    g[2] = u * drdx + r * dudx;
    g[3] = u * drdy + r * dudy;
    g[4] = v * drdx + r * dvdx;
    g[5] = v * drdy + r * dvdy;
    double dru2dx = 2.0 * r * u * dudx + (u * u) * drdx;
    double dru2dy = 2.0 * r * u * dudy + (u * u) * drdy;
    double drv2dx = 2.0 * r * v * dvdx + (v * v) * drdx;
    double drv2dy = 2.0 * r * v * dvdy + (v * v) * drdy;
    g[6] = 1.0 / (GAMMA - 1.0) * m_dpdx(x, y) + 0.5 * (dru2dx + drv2dx);
    g[7] = 1.0 / (GAMMA - 1.0) * m_dpdy(x, y) + 0.5 * (dru2dy + drv2dy);
}

/* manufacturedsolution.d:158-242 */
void orc_source_term(double x, double y, int viscosity, double Rgas, double mu, double Pr, double S[4])
{
    double q[4], dq[8];
    orc_solution(x, y, q);
    orc_solution_gradient(x, y, dq);
    const double r = q[0], u = q[1] / q[0], v = q[2] / q[0], ru = q[1], rv = q[2];
    const double drdx = dq[0], drdy = dq[1], drudx = dq[2], drvdy = dq[5];
    const double p = orc_get_pressure(q);
    const double dudx = m_dudx(x, y), dudy = m_dudy(x, y), dvdy = m_dvdy(x, y), dvdx = m_dvdx(x, y);
    const double dpdx = m_dpdx(x, y), dpdy = m_dpdy(x, y);
    const double dredx = dq[6], dredy = dq[7], rE = q[3];
    double fx[4], fy[4];
    fx[0] = drudx;
    fx[1] = 2.0 * ru * dudx + (u * u) * drdx + dpdx;
    fx[2] = u * v * drdx + r * u * dvdx + r * v * dudx;
    fx[3] = u * dredx + rE * dudx + u * dpdx + p * dudx;
    fy[0] = drvdy;
    fy[1] = u * v * drdy + r * u * dvdy + r * v * dudy;
    fy[2] = 2.0 * rv * dvdy + (v * v) * drdy + dpdy;
    fy[3] = v * dredy + rE * dvdy + v * dpdy + p * dvdy;
    for (int k = 0; k < 4; k++) S[k] = fx[k] + fy[k];
    if (viscosity) {
        double gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0};
        const double k_ = (GAMMA * Rgas * mu) / ((GAMMA - 1) * Pr);
        const double d2udx2 = m_d2udx2(x, y), d2udy2 = m_d2udy2(x, y), d2udxy = m_d2udxy(x, y);
        const double d2vdx2 = m_d2vdx2(x, y), d2vdy2 = m_d2vdy2(x, y), d2vdxy = m_d2vdxy(x, y);
        const double d2pdx2 = m_d2pdx2(x, y), d2pdy2 = m_d2pdy2(x, y);
        const double d2rdx2 = m_d2rdx2(x, y), d2rdy2 = m_d2rdy2(x, y);
        gx[1] = mu * (2.0 / 3.0) * (2 * d2udx2 - d2vdxy);
        gx[2] = mu * (d2udxy + d2vdx2);
        gx[3] = mu * (2.0 / 3.0) * (2 * (u * d2udx2 + dudx * dudx) - (u * d2vdxy + dudx * dvdy));
        gx[3] += mu * ((v * d2udxy + dvdx * dudy) + (v * d2vdx2 + dvdx * dvdx));
        const double d2Tdx2 = (1.0 / Rgas) * ((1.0 / r) * d2pdx2 + (1.0 / (r * r)) * drdx * dpdx
                                - (1.0 / (r * r * r)) * (r * dpdx - 2.0 * p * drdx) * drdx - (p / (r * r)) * d2rdx2);
        gx[3] += k_ * d2Tdx2;
        gy[1] = mu * (d2udy2 + d2vdxy);
        gy[2] = mu * (2.0 / 3.0) * (2 * d2vdy2 - d2udxy);
        gy[3] = mu * (2.0 / 3.0) * (2.0 * (v * d2vdy2 + dvdy * dvdy) - (v * d2udxy + dvdy * dudx));
        gy[3] += mu * (u * d2udy2 + dudy * dudy + u * d2vdxy + dudy * dvdx);
        const double d2Tdy2 = (1.0 / Rgas) * ((1.0 / r) * d2pdy2 + (1.0 / (r * r)) * drdy * dpdy
                                - (1.0 / (r * r * r)) * (r * dpdy - 2.0 * p * drdy) * drdy - (p / (r * r)) * d2rdy2);
        gy[3] += k_ * d2Tdy2;
        for (int k = 0; k < 4; k++) S[k] -= (gx[k] + gy[k]);
    }
}

/* ------------------------------------------------------------------ */
/* source/flux.d:19-24                                                 */
/* ------------------------------------------------------------------ */
void orc_convective_flux(double p, double u, double v, double rho, double e, const double n[2], double out[4])
{
    double f[4] = {rho * u, p + rho * (u * u), rho * u * v, u * (p + e)};
    double h[4] = {rho * v, rho * u * v, p + rho * (v * v), v * (p + e)};
    for (int k = 0; k < 4; k++) out[k] = f[k] * n[0] + h[k] * n[1];
}

/* numd Matrix(4,4) * Vector(4): every column, ascending, from 0.0 */
static void matvec4(const double A[16], const double x[4], double y[4])
{
    for (int r = 0; r < 4; r++) {
        double s = 0.0;
        for (int c = 0; c < 4; c++) s += A[4 * r + c] * x[c];
        y[r] = s;
    }
}

/* source/flux.d:28-64; dq[4][2] row-major (var, dir) */
void orc_diffusive_flux(double Pr, double mu, const double q[4], const double dq[8], const double n[2], double out[4])
{
    const double rho = q[0], u = q[1] / q[0], v = q[2] / q[0], E = q[3] / q[0];
    const double dqdx[4] = {dq[0], dq[2], dq[4], dq[6]};
    const double dqdy[4] = {dq[1], dq[3], dq[5], dq[7]};
    const double gp = GAMMA / Pr;
    double tmp = (-4.0 / 3.0 + gp) * (u * u) + (-1.0 + gp) * (v * v) - gp * E;
    const double s = mu / rho;
    const double c43 = 4.0 / 3.0, c23 = 2.0 / 3.0, cm43 = -4.0 / 3.0, cm23 = -2.0 / 3.0, c231 = (2.0 / 3.0 - 1.0);
    double A00[16] = {0, 0, 0, 0,
                      cm43 * u, c43, 0, 0,
                      -v, 0, 1, 0,
                      tmp, (c43 - gp) * u, (1.0 - gp) * v, gp};
    double A01[16] = {0, 0, 0, 0,
                      c23 * v, 0, cm23, 0,
                      -u, 1, 0, 0,
                      c231 * u * v, v, cm23 * u, 0};
    double A10[16] = {0, 0, 0, 0,
                      -v, 0, 1, 0,
                      c23 * u, 0, cm23, 0,
                      c231 * u * v, cm23 * v, u, 0};
    tmp = (-1.0 + gp) * (u * u) + (-4.0 / 3.0 + gp) * (v * v) - gp * E;
    double A11[16] = {0, 0, 0, 0,
                      -u, 1, 0, 0,
                      cm43 * v, 0, c43, 0,
                      tmp, (1.0 - gp) * u, (c43 - gp) * v, gp};
    for (int i = 0; i < 16; i++) { A00[i] = s * A00[i]; A01[i] = s * A01[i]; A10[i] = s * A10[i]; A11[i] = s * A11[i]; }
    double a[4], b[4], G0[4], G1[4];
    matvec4(A00, dqdx, a); matvec4(A01, dqdy, b);
    for (int k = 0; k < 4; k++) G0[k] = a[k] + b[k];
    matvec4(A10, dqdx, a); matvec4(A11, dqdy, b);
    for (int k = 0; k < 4; k++) G1[k] = a[k] + b[k];
    for (int k = 0; k < 4; k++) out[k] = G0[k] * n[0] + G1[k] * n[1];
}

static double dmax(double a, double b) { return a < b ? b : a; }

/* Harten entropy fix shared by both fluxes (flux.d:115-131, 259-275) */
static double harten(double lam, double eps)
{
    if (lam < eps) lam = (eps * eps + lam * lam) / (2.0 * eps);
    return lam;
}

/* source/flux.d:66-141 */
void orc_rusanov_flux(const double qL[4], const double qR[4], const double n[2], double out[4], double *sMax)
{
    const double rhoL = qL[0], uL = qL[1] / rhoL, vL = qL[2] / rhoL;
    const double rvL[2] = {rhoL * uL, rhoL * vL};
    const double pL = ((GAMMA - 1) * (qL[3] - 0.5 * rhoL * (uL * uL + vL * vL)));
    const double hL = (qL[3] + pL) / rhoL;
    const double rhoR = qR[0], uR = qR[1] / rhoR, vR = qR[2] / rhoR;
    const double rvR[2] = {rhoR * uR, rhoR * vR};
    const double pR = ((GAMMA - 1) * (qR[3] - 0.5 * rhoR * (uR * uR + vR * vR)));
    const double hR = (qR[3] + pR) / rhoR;

    const double rho = sqrt(rhoL * rhoR);
    double vel[2];
    vel[0] = (sqrt(rhoL) * uL + sqrt(rhoR) * uR) / (sqrt(rhoL) + sqrt(rhoR));
    vel[1] = (sqrt(rhoL) * vL + sqrt(rhoR) * vR) / (sqrt(rhoL) + sqrt(rhoR));
    const double u = vel[0] * n[0] + vel[1] * n[1];
    const double p = (sqrt(rhoL) * pL + sqrt(rhoR) * pR) / (sqrt(rhoL) + sqrt(rhoR));
    const double h = (sqrt(rhoL) * hL + sqrt(rhoR) * hR) / (sqrt(rhoL) + sqrt(rhoR));
    const double a = sqrt(GAMMA * (p / rho));

    const double dRhoV[2] = {rvR[0] - rvL[0], rvR[1] - rvL[1]};
    const double dRho = rhoR - rhoL;
    const double dRhoE = qR[3] - qL[3];

    const double eps = 0.01 * a;
    double lam1 = harten(fabs(u + a), eps);
    double lam2 = harten(fabs(u - a), eps);
    double lam3 = harten(fabs(u), eps);
    double lam4 = harten(fabs(u), eps);
    const double sm = dmax(dmax(dmax(fabs(lam1), fabs(lam2)), fabs(lam3)), fabs(lam4));
    *sMax = sm;

    const double s1 = 0.5 * (sm + sm);
    const double s2 = 0.5 * (sm - sm);
    const double qq = sqrt(vel[0] * vel[0] + vel[1] * vel[1]);
    const double G1 = (GAMMA - 1) * ((qq * qq / 2) * dRho - (vel[0] * dRhoV[0] + vel[1] * dRhoV[1]) + dRhoE);
    const double G2 = -u * dRho + (dRhoV[0] * n[0] + dRhoV[1] * n[1]);
    const double C1 = (G1 / (a * a)) * (s1 - sm) + G2 / a * s2;
    const double C2 = (G1 / a) * s2 + (s1 - sm) * G2;

    double fL[4], fR[4];
    orc_convective_flux(pL, uL, vL, rhoL, qL[3], n, fL);
    orc_convective_flux(pR, uR, vR, rhoR, qR[3], n, fR);
    const double d[4] = {sm * dRho + C1,
                         sm * (dRhoV[0]) + C1 * vel[0] + C2 * n[0],
                         sm * (dRhoV[1]) + C1 * vel[1] + C2 * n[1],
                         sm * dRhoE + C1 * h + C2 * u};
    for (int k = 0; k < 4; k++) out[k] = 0.5 * (fL[k] + fR[k]) - 0.5 * d[k];
}

/* source/flux.d:224-303 */
void orc_roe_flux(const double qL[4], const double qR[4], const double n[2], double out[4], double *sMax)
{
    const double rhoL = qL[0], uL = qL[1] / rhoL, vL = qL[2] / rhoL;
    const double rvL[2] = {rhoL * uL, rhoL * vL};
    const double pL = ((GAMMA - 1) * (qL[3] - 0.5 * rhoL * (uL * uL + vL * vL)));
    const double hL = (qL[3] + pL) / rhoL;
    const double rhoR = qR[0], uR = qR[1] / rhoR, vR = qR[2] / rhoR;
    const double rvR[2] = {rhoR * uR, rhoR * vR};
    const double pR = ((GAMMA - 1) * (qR[3] - 0.5 * rhoR * (uR * uR + vR * vR)));
    const double hR = (qR[3] + pR) / rhoR;

    const double di = sqrt(rhoR / rhoL);
    const double d1 = 1.0 / (1.0 + di);
    double vel[2];
    vel[0] = (di * uR + uL) * d1;
    vel[1] = (di * vR + vL) * d1;
    const double u = vel[0] * n[0] + vel[1] * n[1];
    const double h = (di * hR + hL) * d1;
    const double a = sqrt((GAMMA - 1.0) * (h - 0.5 * (vel[0] * vel[0] + vel[1] * vel[1])));

    const double dRhoV[2] = {rvR[0] - rvL[0], rvR[1] - rvL[1]};
    const double dRho = rhoR - rhoL;
    const double dRhoE = qR[3] - qL[3];

    const double eps = 0.01 * a;
    double lam1 = harten(fabs(u + a), eps);
    double lam2 = harten(fabs(u - a), eps);
    double lam3 = harten(fabs(u), eps);
    double lam4 = harten(fabs(u), eps);
    *sMax = dmax(dmax(dmax(fabs(lam1), fabs(lam2)), fabs(lam3)), fabs(lam4));

    const double s1 = 0.5 * (lam1 + lam2);
    const double s2 = 0.5 * (lam1 - lam2);
    const double qq = sqrt(vel[0] * vel[0] + vel[1] * vel[1]);
    const double G1 = (GAMMA - 1) * (((qq * qq) / 2.0) * dRho - (vel[0] * dRhoV[0] + vel[1] * dRhoV[1]) + dRhoE);
    const double G2 = -u * dRho + (dRhoV[0] * n[0] + dRhoV[1] * n[1]);
    const double C1 = (G1 / (a * a)) * (s1 - lam3) + G2 / a * s2;
    const double C2 = (G1 / a) * s2 + (s1 - lam3) * G2;

    double fL[4], fR[4];
    orc_convective_flux(pL, uL, vL, rhoL, qL[3], n, fL);
    orc_convective_flux(pR, uR, vR, rhoR, qR[3], n, fR);
    const double d[4] = {lam3 * dRho + C1,
                         lam3 * dRhoV[0] + C1 * vel[0] + C2 * n[0],
                         lam3 * dRhoV[1] + C1 * vel[1] + C2 * n[1],
                         lam3 * dRhoE + C1 * h + C2 * u};
    for (int k = 0; k < 4; k++) out[k] = 0.5 * (fL[k] + fR[k]) - 0.5 * d[k];
}

static void flux_dispatch(int which, const double qL[4], const double qR[4], const double n[2], double out[4], double *sMax)
{
    if (which == FLUX_ROE) orc_roe_flux(qL, qR, n, out, sMax);
    else orc_rusanov_flux(qL, qR, n, out, sMax);
}

/* ------------------------------------------------------------------ */
/* source/limiters.d:9-37 (not called by ebb-solve; restated as a set) */
/* ------------------------------------------------------------------ */
static double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }
double orc_limiter1d(int kind, double Sp, double Sm)
{
    switch (kind) {
    case 0: return 0.0;                                                   /* firstOrderUpwindS */
    case 1: return Sp;                                                    /* laxWendroffS */
    case 2: return fmax(0, fmin(fabs(Sm), fabs(Sp))) * sgn(Sm + Sp);      /* minmodS */
    case 3: return (Sm * fabs(Sp) + fabs(Sm) * Sp) / (fabs(Sm) + fabs(Sp) + 10e-12); /* harmonicS */
    case 4: return sqrt((Sm * Sp + fabs(Sm * Sp)) / 2) * sgn(Sm + Sp);    /* geometricS */
    case 5: return fmax(0, fmax(fmin(2 * fabs(Sm), fabs(Sp)), fmin(fabs(Sm), 2 * fabs(Sp)))) * sgn(Sm + Sp); /* superbeeS */
    }
    return NAN;
}

/* ------------------------------------------------------------------ */
/* Phobos std.algorithm.iteration.sum for double[] (pairwise, blocks   */
/* of 16 summed by a fixed tree, a binary-counter stack, then a fold   */
/* from the top of the stack).  Used by app/solver.d:153-168.          */
/* ------------------------------------------------------------------ */
static double sum16(const double *r)
{
    return (((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7])))
         + (((r[8] + r[9]) + (r[10] + r[11])) + ((r[12] + r[13]) + (r[14] + r[15])));
}
static unsigned bsf64(uint64_t v) { return (unsigned)__builtin_ctzll(v); }
double orc_pairwise_sum(const double *data, uint64_t len)
{
    if (len == 0) return 0.0;
    double store[80];
    uint64_t idx = 0;
    uint64_t nb = len / 16;
    for (uint64_t k = 0; k < nb; k++) {
        store[idx] = sum16(data + 16 * k);
        uint64_t lastToKeep = idx - bsf64(k + 1);
        while (idx > lastToKeep) { store[idx - 1] += store[idx]; --idx; }
        ++idx;
    }
    for (uint64_t i = 0; i < len - 16 * nb; i++) {
        store[idx] = data[16 * nb + i];
        uint64_t lastToKeep = idx - bsf64(i + 1);
        while (idx > lastToKeep) { store[idx - 1] += store[idx]; --idx; }
        ++idx;
    }
    double s = store[idx - 1];
    for (int64_t j = (int64_t)idx - 2; j >= 0; j--) s += store[j];
    return 0.0 + s;
}

/* ------------------------------------------------------------------ */
/* Edge hash for buildMesh (mir.sparse N x N map in the reference)     */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t *keys; uint32_t *vals; uint64_t cap; } Hash;
static void hash_init(Hash *h, uint64_t n)
{
    uint64_t cap = 16;
    while (cap < 2 * n + 16) cap <<= 1;
    h->cap = cap;
    h->keys = (uint64_t *)malloc(cap * sizeof(uint64_t));
    h->vals = (uint32_t *)calloc(cap, sizeof(uint32_t));
    for (uint64_t i = 0; i < cap; i++) h->keys[i] = UINT64_MAX;
}
static void hash_free(Hash *h) { free(h->keys); free(h->vals); }
static uint64_t hmix(uint64_t k) { k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33; return k; }
static uint32_t *hash_slot(Hash *h, uint32_t a, uint32_t b)
{
    uint32_t n1 = a > b ? b : a, n2 = a > b ? a : b;
    uint64_t key = ((uint64_t)n1 << 32) | n2;
    uint64_t i = hmix(key) & (h->cap - 1);
    while (h->keys[i] != UINT64_MAX && h->keys[i] != key) i = (i + 1) & (h->cap - 1);
    if (h->keys[i] == UINT64_MAX) { h->keys[i] = key; h->vals[i] = 0; }
    return &h->vals[i];
}
static uint32_t hash_get(const Hash *h, uint32_t a, uint32_t b)
{
    uint32_t n1 = a > b ? b : a, n2 = a > b ? a : b;
    uint64_t key = ((uint64_t)n1 << 32) | n2;
    uint64_t i = hmix(key) & (h->cap - 1);
    while (h->keys[i] != UINT64_MAX) {
        if (h->keys[i] == key) return h->vals[i];
        i = (i + 1) & (h->cap - 1);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* source/mesh.d:217-643 buildMesh (single rank, no comm edges)        */
/* ------------------------------------------------------------------ */
void orc_mesh_free(OrcMesh *m);

OrcMesh *orc_mesh_build(uint32_t nNodes, const double *xy, uint32_t nElem, const uint32_t *eoff,
                        const uint32_t *enodes, uint32_t nB, const uint32_t *bnodes,
                        uint32_t nTags, const uint64_t *bGroupStart)
{
    OrcMesh *m = (OrcMesh *)calloc(1, sizeof(OrcMesh));
    m->nNodes = nNodes; m->nElem = nElem; m->nTags = nTags;
    m->nodes = (double *)malloc(sizeof(double) * 2 * (nNodes ? nNodes : 1));
    memcpy(m->nodes, xy, sizeof(double) * 2 * nNodes);
    m->eoff = (uint32_t *)malloc(sizeof(uint32_t) * (nElem + 1));
    memcpy(m->eoff, eoff, sizeof(uint32_t) * (nElem + 1));
    m->enodes = (uint32_t *)malloc(sizeof(uint32_t) * (eoff[nElem] ? eoff[nElem] : 1));
    memcpy(m->enodes, enodes, sizeof(uint32_t) * eoff[nElem]);

    uint64_t maxEdges = eoff[nElem];
    uint64_t maxCells = nElem + nB + 1;
    m->cells = (OCell *)calloc(maxCells, sizeof(OCell));
    m->edges = (OEdge *)calloc(maxEdges + 1, sizeof(OEdge));
    m->interiorEdges = (uint32_t *)malloc(sizeof(uint32_t) * (maxEdges + 1));
    m->boundaryEdges = (uint32_t *)malloc(sizeof(uint32_t) * (maxEdges + 1));
    m->bGroups = (uint32_t **)calloc(nTags ? nTags : 1, sizeof(uint32_t *));
    m->bGroupLen = (uint32_t *)calloc(nTags ? nTags : 1, sizeof(uint32_t));
    for (uint32_t t = 0; t < nTags; t++) m->bGroups[t] = (uint32_t *)malloc(sizeof(uint32_t) * (nB + 1));

    Hash edgeHash, boundaryHash;
    hash_init(&edgeHash, maxEdges);
    hash_init(&boundaryHash, nB);
    for (uint32_t i = 0; i < nB; i++) *hash_slot(&boundaryHash, bnodes[2 * i], bnodes[2 * i + 1]) = i + 1;

    const double *nd = m->nodes;
    uint32_t nEdges = 0;
    for (uint32_t i = 0; i < nElem; i++) {
        OCell *c = &m->cells[i];
        c->nEdges = eoff[i + 1] - eoff[i];
        if (c->nEdges > MAXE || c->nEdges < 3) { hash_free(&edgeHash); hash_free(&boundaryHash); orc_mesh_free(m); return NULL; }
        for (int j = 0; j < MAXE; j++) { c->edges[j] = 0; c->fluxMultiplier[j] = 1.0; }
        c->area = 0.0; c->d = 0.0; c->perim = 0; c->centroid[0] = 0; c->centroid[1] = 0;
        c->nNeighborCells = 0;
        for (int k = 0; k < 4; k++) { c->lim[k][0] = 1.0; c->lim[k][1] = 1.0; }
        const uint32_t *el = enodes + eoff[i];
        for (uint32_t j = 0; j < c->nEdges; j++) {
            uint32_t n0 = el[j] - 1, n1 = el[(j + 1) % c->nEdges] - 1;
            uint32_t edgeidx;
            uint32_t bval = hash_get(&boundaryHash, n0, n1);
            if (bval == 0) {
                uint32_t *slot = hash_slot(&edgeHash, n0, n1);
                if (*slot != 0) {
                    uint32_t idx = *slot - 1;
                    m->edges[idx].cellIdx[1] = i;
                    edgeidx = idx;
                    c->fluxMultiplier[j] = -1.0;
                } else {
                    OEdge e;
                    memset(&e, 0, sizeof(e));
                    e.nodeIdx[0] = n0; e.nodeIdx[1] = n1;
                    e.cellIdx[0] = i;
                    double ddx = nd[2 * n0] - nd[2 * n1], ddy = nd[2 * n0 + 1] - nd[2 * n1 + 1];
                    e.len = sqrt(ddx * ddx + ddy * ddy);
                    double nx = nd[2 * n1 + 1] - nd[2 * n0 + 1], ny = nd[2 * n0] - nd[2 * n1];
                    double mag = sqrt(nx * nx + ny * ny);
                    e.normal[0] = nx / mag; e.normal[1] = ny / mag;
                    e.mid[0] = 0.5 * (nd[2 * n0] + nd[2 * n1]);
                    e.mid[1] = 0.5 * (nd[2 * n0 + 1] + nd[2 * n1 + 1]);
                    e.bGroup = -1;
                    c->fluxMultiplier[j] = 1.0;
                    m->edges[nEdges] = e;
                    edgeidx = nEdges++;
                    *slot = edgeidx + 1;
                    m->interiorEdges[m->nInteriorEdges++] = edgeidx;
                }
            } else {
                OEdge e;
                memset(&e, 0, sizeof(e));
                e.nodeIdx[0] = n0; e.nodeIdx[1] = n1;
                e.cellIdx[0] = i;
                double ddx = nd[2 * n0] - nd[2 * n1], ddy = nd[2 * n0 + 1] - nd[2 * n1 + 1];
                e.len = sqrt(ddx * ddx + ddy * ddy);
                double inv = 1.0 / e.len;
                e.normal[0] = inv * (nd[2 * n1 + 1] - nd[2 * n0 + 1]);
                e.normal[1] = inv * (nd[2 * n0] - nd[2 * n1]);
                e.mid[0] = 0.5 * (nd[2 * n0] + nd[2 * n1]);
                e.mid[1] = 0.5 * (nd[2 * n0 + 1] + nd[2 * n1 + 1]);
                e.isBoundary = 1;
                /* mesh.d:162-179 isBoundaryEdge group search */
                uint32_t bNodeIdx = bval - 1;
                int bGroup = 0;
                for (uint32_t g = 0; g + 1 < nTags; g++)
                    if (bNodeIdx >= bGroupStart[g] && bNodeIdx < bGroupStart[g + 1]) bGroup = (int)g;
                if (nTags > 0 && bNodeIdx >= bGroupStart[nTags - 1]) bGroup = (int)nTags - 1;
                e.bGroup = bGroup;
                m->bGroups[bGroup][m->bGroupLen[bGroup]++] = nEdges;
                c->fluxMultiplier[j] = 1.0;
                m->edges[nEdges] = e;
                edgeidx = nEdges++;
                m->boundaryEdges[m->nBoundaryEdges++] = edgeidx;
            }
            c->edges[j] = edgeidx;
            c->perim += m->edges[edgeidx].len;
            double x0 = nd[2 * n0], x1 = nd[2 * n1], y0 = nd[2 * n0 + 1], y1 = nd[2 * n1 + 1];
            c->area += x0 * y1 - x1 * y0;
            c->centroid[0] += (x0 + x1) * (x0 * y1 - x1 * y0);
            c->centroid[1] += (y0 + y1) * (x0 * y1 - x1 * y0);
        }
        c->area *= 0.5;
        if (!(c->area > 0)) { hash_free(&edgeHash); hash_free(&boundaryHash); orc_mesh_free(m); return NULL; }
        c->d = (2 * c->area) / c->perim;
        double f = 1 / (6 * c->area);
        c->centroid[0] *= f;
        c->centroid[1] *= f;
    }
    m->nEdges = nEdges;
    hash_free(&edgeHash);
    hash_free(&boundaryHash);

    /* ghost cells, mesh.d:329-381 */
    uint32_t nCells = nElem;
    for (uint32_t b = 0; b < m->nBoundaryEdges; b++) {
        uint32_t be = m->boundaryEdges[b];
        OEdge *e = &m->edges[be];
        OCell *g = &m->cells[nCells];
        memset(g, 0, sizeof(*g));
        for (int j = 0; j < MAXE; j++) { g->edges[j] = 0; g->fluxMultiplier[j] = -1.0; }
        const OCell *oc = &m->cells[e->cellIdx[0]];
        double xr, yr;
        if (nd[2 * e->nodeIdx[0]] != nd[2 * e->nodeIdx[1]]) {
            double x1 = oc->centroid[0], y1 = oc->centroid[1];
            double mm = (nd[2 * e->nodeIdx[0] + 1] - nd[2 * e->nodeIdx[1] + 1]) / (nd[2 * e->nodeIdx[0]] - nd[2 * e->nodeIdx[1]]);
            double x = nd[2 * e->nodeIdx[0]], y = nd[2 * e->nodeIdx[0] + 1];
            double cc = y - mm * x;
            double dd = (x1 + (y1 - cc) * mm) / (1.0 + mm * mm);
            xr = 2.0 * dd - x1;
            yr = 2.0 * dd * mm - y1 + 2.0 * cc;
        } else {
            double x = nd[2 * e->nodeIdx[0]];
            double dd = oc->centroid[0] - x;
            double x1 = oc->centroid[0];
            xr = -(2.0 * dd - x1);
            yr = oc->centroid[1];
        }
        g->centroid[0] = xr; g->centroid[1] = yr;
        g->nNeighborCells = 1;
        for (int k = 0; k < 4; k++) { g->lim[k][0] = 1.0; g->lim[k][1] = 1.0; }
        g->neighborCells[0] = e->cellIdx[0];
        g->edges[0] = be;
        e->cellIdx[1] = nCells;
        nCells++;
    }
    m->nCells = nCells;

    /* neighbours + LSQ gradient matrix, mesh.d:481-642 */
    for (uint32_t i = 0; i < nElem; i++) {
        OCell *c = &m->cells[i];
        double tmpMat[MAXE][2];
        memset(tmpMat, 0, sizeof(tmpMat));
        c->nNeighborCells = 0;
        for (uint32_t j = 0; j < c->nEdges; j++) {
            const OEdge *e = &m->edges[c->edges[j]];
            uint32_t nb = (e->cellIdx[0] == i) ? e->cellIdx[1] : e->cellIdx[0];
            c->neighborCells[c->nNeighborCells] = nb;
            tmpMat[c->nNeighborCells][0] = m->cells[nb].centroid[0] - c->centroid[0];
            tmpMat[c->nNeighborCells][1] = m->cells[nb].centroid[1] - c->centroid[1];
            c->nNeighborCells++;
        }
        /* every edge has cellIdx[1] set (ghost or neighbour), so the
           "Not equal" branch (mesh.d:540-634) never fires */
        double b[2][2];
        for (int r = 0; r < 2; r++)
            for (int cc = 0; cc < 2; cc++) {
                double s = 0.0;
                for (int k = 0; k < MAXE; k++) s += tmpMat[k][r] * tmpMat[k][cc];
                b[r][cc] = s;
            }
        double det = b[0][0] * b[1][1] - b[0][1] * b[1][0];
        double id = 1.0 / det;
        double inv[2][2] = {{b[1][1] * id, -b[0][1] * id}, {-b[1][0] * id, b[0][0] * id}};
        for (int r = 0; r < 2; r++)
            for (int k = 0; k < MAXE; k++)
                c->gradMat[r][k] = inv[r][0] * tmpMat[k][0] + inv[r][1] * tmpMat[k][1];
    }

    size_t C = nCells;
    m->q = (double (*)[4])calloc(C, sizeof(double[4]));
    m->R = (double (*)[4])calloc(C, sizeof(double[4]));
    m->k1 = (double (*)[4])calloc(C, sizeof(double[4]));
    m->k2 = (double (*)[4])calloc(C, sizeof(double[4]));
    m->k3 = (double (*)[4])calloc(C, sizeof(double[4]));
    m->k4 = (double (*)[4])calloc(C, sizeof(double[4]));
    m->tmp = (double (*)[4])calloc(C, sizeof(double[4]));
    m->src = (double (*)[4])calloc(C, sizeof(double[4]));
    return m;
}

void orc_mesh_free(OrcMesh *m)
{
    if (!m) return;
    free(m->nodes); free(m->eoff); free(m->enodes); free(m->cells); free(m->edges);
    free(m->interiorEdges); free(m->boundaryEdges);
    if (m->bGroups) for (uint32_t t = 0; t < m->nTags; t++) free(m->bGroups[t]);
    free(m->bGroups); free(m->bGroupLen);
    free(m->q); free(m->R); free(m->k1); free(m->k2); free(m->k3); free(m->k4); free(m->tmp); free(m->src);
    free(m);
}

/* ------------------------------------------------------------------ */
/* Accessors for the tests                                              */
/* ------------------------------------------------------------------ */
void orc_mesh_counts(const OrcMesh *m, int64_t out[6])
{
    out[0] = m->nCells; out[1] = m->nElem; out[2] = m->nCells - m->nElem;
    out[3] = m->nEdges; out[4] = m->nInteriorEdges; out[5] = m->nBoundaryEdges;
}

void orc_mesh_cells(const OrcMesh *m, double *area, double *d, double *perim, double *centroid,
                    double *gradMat, uint32_t *edges, double *fm, uint32_t *nbrs, uint32_t *nEdges)
{
    for (uint32_t i = 0; i < m->nCells; i++) {
        const OCell *c = &m->cells[i];
        area[i] = c->area; d[i] = c->d; perim[i] = c->perim;
        centroid[2 * i] = c->centroid[0]; centroid[2 * i + 1] = c->centroid[1];
        for (int r = 0; r < 2; r++) for (int k = 0; k < MAXE; k++) gradMat[12 * i + MAXE * r + k] = c->gradMat[r][k];
        for (int k = 0; k < MAXE; k++) { edges[MAXE * i + k] = c->edges[k]; fm[MAXE * i + k] = c->fluxMultiplier[k]; nbrs[MAXE * i + k] = c->neighborCells[k]; }
        nEdges[i] = (i < m->nElem) ? c->nEdges : 0;
    }
}

void orc_mesh_edges(const OrcMesh *m, uint32_t *nodeIdx, uint32_t *cellIdx, double *len, double *normal,
                    double *mid, int32_t *bGroup)
{
    for (uint32_t e = 0; e < m->nEdges; e++) {
        const OEdge *E = &m->edges[e];
        nodeIdx[2 * e] = E->nodeIdx[0]; nodeIdx[2 * e + 1] = E->nodeIdx[1];
        cellIdx[2 * e] = E->cellIdx[0]; cellIdx[2 * e + 1] = E->cellIdx[1];
        len[e] = E->len;
        normal[2 * e] = E->normal[0]; normal[2 * e + 1] = E->normal[1];
        mid[2 * e] = E->mid[0]; mid[2 * e + 1] = E->mid[1];
        bGroup[e] = E->isBoundary ? E->bGroup : -1;
    }
}

void orc_boundary_group(const OrcMesh *m, uint32_t tag, uint32_t *out, uint32_t *n)
{
    *n = m->bGroupLen[tag];
    if (out) memcpy(out, m->bGroups[tag], sizeof(uint32_t) * m->bGroupLen[tag]);
}

void orc_get_q(const OrcMesh *m, double *q) { memcpy(q, m->q, sizeof(double) * 4 * m->nCells); }
void orc_set_q(OrcMesh *m, const double *q) { memcpy(m->q, q, sizeof(double) * 4 * m->nCells); }
void orc_get_edge_data(const OrcMesh *m, double *eq, double *flux, double *sMax)
{
    for (uint32_t e = 0; e < m->nEdges; e++) {
        const OEdge *E = &m->edges[e];
        for (int k = 0; k < 4; k++) { eq[8 * e + k] = E->q[0][k]; eq[8 * e + 4 + k] = E->q[1][k]; flux[4 * e + k] = E->flux[k]; }
        sMax[e] = E->sMax;
    }
}
void orc_get_cell_data(const OrcMesh *m, double *grad, double *lim, double *dtc)
{
    for (uint32_t i = 0; i < m->nCells; i++) {
        const OCell *c = &m->cells[i];
        for (int k = 0; k < 4; k++) {
            grad[8 * i + 2 * k] = c->gradient[k][0]; grad[8 * i + 2 * k + 1] = c->gradient[k][1];
            lim[8 * i + 2 * k] = c->lim[k][0]; lim[8 * i + 2 * k + 1] = c->lim[k][1];
        }
        dtc[i] = c->dt;
    }
}
void orc_last_error(const OrcMesh *m, int *kind, int64_t *index) { *kind = m->errKind; *index = m->errIndex; }

/* ------------------------------------------------------------------ */
/* source/solve.d:175-279 ufvmSetup (fresh start) + app/ms.d:285-298   */
/* ------------------------------------------------------------------ */
static void freestream(double M, double aoaDeg, double rho, double q[4])
{
    /* D evaluates ic[1]*(PI/180) in 80-bit real and cos/sin of a double via real */
    double aoa = (double)((long double)aoaDeg * (3.14159265358979323846264338327950288L / 180.0L));
    double U = 1.0, a = U / M;
    double p = (rho * (a * a)) / GAMMA;
    double u = U * (double)cosl((long double)aoa), v = U * (double)sinl((long double)aoa);
    q[0] = rho; q[1] = rho * u; q[2] = rho * v;
    q[3] = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v);
}

int orc_setup(OrcMesh *m, OrcConfig *cfg)
{
    /* mu is only set for viscous runs (solve.d:186-190); otherwise it keeps D's NaN default */
    cfg->mu = cfg->viscosity ? (cfg->ic[3] * 1.0 * cfg->L) / cfg->Re : NAN;
    double q0[4];
    freestream(cfg->ic[0], cfg->ic[1], cfg->ic[3], q0);
    for (uint32_t i = 0; i < m->nElem; i++) {
        if (cfg->mms) orc_solution(m->cells[i].centroid[0], m->cells[i].centroid[1], m->q[i]);
        else for (int k = 0; k < 4; k++) m->q[i][k] = q0[k];
        if (cfg->mms) orc_source_term(m->cells[i].centroid[0], m->cells[i].centroid[1], cfg->viscosity, cfg->R, cfg->mu, cfg->Pr, m->src[i]);
    }
    for (uint32_t t = 0; t < m->nTags; t++) {
        int bc = cfg->tagToBc[t];
        if (bc < 0 || bc >= cfg->nBC) { m->errKind = ORC_ERR_BC; m->errIndex = t; return ORC_ERR_BC; }
        for (uint32_t j = 0; j < m->bGroupLen[t]; j++) {
            OEdge *e = &m->edges[m->bGroups[t][j]];
            e->bType = cfg->bcType[bc];
            e->bIdx = bc;
            if (e->bType == BT_FULLSTATE) {
                freestream(cfg->bcData[bc][0], cfg->bcData[bc][1], cfg->bcData[bc][3], e->q[1]);
            } else if (e->bType == BT_CONSTP) {
                if (cfg->bcNData[bc] != 1) { m->errKind = ORC_ERR_BC; m->errIndex = t; return ORC_ERR_BC; }
                e->bData[0] = cfg->bcData[bc][0]; e->nbData = 1;
            } else if (e->bType == BT_TPINFLOW) {
                if (cfg->bcNData[bc] != 3) { m->errKind = ORC_ERR_BC; m->errIndex = t; return ORC_ERR_BC; }
                for (int k = 0; k < 3; k++) e->bData[k] = cfg->bcData[bc][k];
                e->nbData = 3;
            } else if (e->bType == BT_DIRICHLET) {
                if (!cfg->mms) { m->errKind = ORC_ERR_BC; m->errIndex = t; return ORC_ERR_BC; }
                orc_solution(e->mid[0], e->mid[1], e->q[1]); /* app/ms.d:282-284 */
            }
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* source/solve.d:27-173 invert + LP                                    */
/* ------------------------------------------------------------------ */
static void invert2(double inv[4], const double mat[4])
{
    double det = mat[0] * mat[3] - mat[1] * mat[2];
    double s = 1.0 / det;
    inv[0] = mat[3] * s; inv[1] = -mat[1] * s; inv[2] = -mat[2] * s; inv[3] = mat[0] * s;
}
static void invert2T(double inv[4], const double mat[4])
{
    double det = mat[0] * mat[3] - mat[1] * mat[2];
    double s = 1.0 / det;
    inv[0] = mat[3] * s; inv[1] = -mat[2] * s; inv[2] = -mat[1] * s; inv[3] = mat[0] * s;
}
void orc_lp(int mrows, const double *A /* mrows x 2 */, const double *b, const double c[2], double xk[2])
{
    int Wk[2] = {0, 1};
    double Ak[4] = {A[2 * Wk[0]], A[2 * Wk[0] + 1], A[2 * Wk[1]], A[2 * Wk[1] + 1]};
    double AkInv[4], lam[2], pk[2], e[2];
    int D[16];
    double gam[16];
    int k = 0;
    for (;;) {
        invert2T(AkInv, Ak);
        lam[0] = 0.0 + AkInv[0] * c[0] + AkInv[1] * c[1];
        lam[1] = 0.0 + AkInv[2] * c[0] + AkInv[3] * c[1];
        if ((lam[0] >= -10e-9) && (lam[1] >= -10e-9)) break;
        int q = (lam[0] < lam[1]) ? Wk[0] : Wk[1];
        if (lam[0] < lam[1]) { e[0] = 1; e[1] = 0; } else { e[0] = 0; e[1] = 1; }
        invert2(AkInv, Ak);
        pk[0] = 0.0 + AkInv[0] * e[0] + AkInv[1] * e[1];
        pk[1] = 0.0 + AkInv[2] * e[0] + AkInv[3] * e[1];
        int Dlen = 0;
        for (int i = 0; i < mrows; i++)
            if (i != Wk[0] && i != Wk[1])
                if (A[2 * i] * pk[0] + A[2 * i + 1] * pk[1] < -1e-11) D[Dlen++] = i;
        double alpha = INFINITY;
        for (int i = 0; i < Dlen; i++) {
            double aixk = A[2 * D[i]] * xk[0] + A[2 * D[i] + 1] * xk[1];
            double aipk = A[2 * D[i]] * pk[0] + A[2 * D[i] + 1] * pk[1];
            gam[i] = (aixk - b[D[i]]) / (-aipk);
            alpha = fmin(alpha, gam[i]);
        }
        int t = 0;
        for (int i = 0; i < Dlen; i++) if (gam[i] == alpha) { t = D[i]; break; }
        if (q == Wk[0]) Wk[0] = t; else if (q == Wk[1]) Wk[1] = t;
        Ak[0] = A[2 * Wk[0]]; Ak[1] = A[2 * Wk[0] + 1]; Ak[2] = A[2 * Wk[1]]; Ak[3] = A[2 * Wk[1] + 1];
        xk[0] = xk[0] + alpha * pk[0];
        xk[1] = xk[1] + alpha * pk[1];
        k++;
        if (k > 15) break;
    }
}

/* ------------------------------------------------------------------ */
/* source/solve.d:283-319                                               */
/* ------------------------------------------------------------------ */
static void update_boundary_edge(OrcMesh *m, double (*q)[4], uint32_t i, const OrcConfig *cfg)
{
    OEdge *e = &m->edges[i];
    if (cfg->order == 1) { for (int k = 0; k < 4; k++) e->q[0][k] = q[e->cellIdx[0]][k]; return; }
    const OCell *c = &m->cells[e->cellIdx[0]];
    const double *qM = q[e->cellIdx[0]];
    double dx = e->mid[0] - c->centroid[0], dy = e->mid[1] - c->centroid[1];
    for (int j = 0; j < 4; j++)
        e->q[0][j] = qM[j] + c->lim[j][0] * c->gradient[j][0] * dx + c->lim[j][1] * c->gradient[j][1] * dy;
    if (orc_get_pressure(e->q[0]) < 0) {
        double lim[2] = {1.0, 1.0};
        for (int j = 0; j < 4; j++) {
            lim[0] = fmin(lim[0], c->lim[j][0]);
            lim[1] = fmin(lim[0], m->cells[e->cellIdx[1]].lim[j][1]);
        }
        for (int j = 0; j < 4; j++)
            e->q[0][j] = qM[j] + lim[0] * c->gradient[j][0] * dx + lim[0] * c->gradient[j][1] * dy;
    }
}

static int has_nan4(const double f[4]) { return isnan(f[0]) || isnan(f[1]) || isnan(f[2]) || isnan(f[3]); }

/* reflection of v about the line through the edge nodes (solve.d:536-552) */
static void reflect_about_edge(const OrcMesh *m, const OEdge *e, const double v[2], double out[2])
{
    double x1 = m->nodes[2 * e->nodeIdx[0]], y1 = m->nodes[2 * e->nodeIdx[0] + 1];
    double x2 = m->nodes[2 * e->nodeIdx[1]], y2 = m->nodes[2 * e->nodeIdx[1] + 1];
    if (x1 != x2) {
        double mm = (y2 - y1) / (x2 - x1);
        double den = 1 + mm * mm;
        double R[4] = {(1 - mm * mm) * 1.0 / den, (2.0 * mm) * 1.0 / den, (2.0 * mm) * 1.0 / den, (mm * mm - 1) * 1.0 / den};
        out[0] = 0.0 + R[0] * v[0] + R[1] * v[1];
        out[1] = 0.0 + R[2] * v[0] + R[3] * v[1];
    } else {
        out[0] = -v[0];
        out[1] = v[1];
    }
}

/* ------------------------------------------------------------------ */
/* source/solve.d:322-1122 ufvmSolver (single rank)                     */
/* ------------------------------------------------------------------ */
int orc_ufvm_solver(OrcMesh *m, const OrcConfig *cfg, double (*R)[4], double (*q)[4],
                    double *newDt, double *Rmax, int limit, int dtUpdate)
{
    const double PrC = cfg->Pr, mu = cfg->mu;
    /* ghost cell states, solve.d:520-689 */
    uint32_t firstG = UINT32_MAX;
    int kindG = 0;
#pragma omp parallel for schedule(static)
    for (uint32_t b = 0; b < m->nBoundaryEdges; b++) {
        uint32_t gi = m->nElem + b;
        const OEdge *e = &m->edges[m->cells[gi].edges[0]];
        uint32_t ci = e->cellIdx[0], g = e->cellIdx[1];
        switch (e->bType) {
        case BT_DIRICHLET:
            orc_solution(m->cells[gi].centroid[0], m->cells[gi].centroid[1], q[gi]);
            break;
        case BT_FULLSTATE:
            for (int k = 0; k < 4; k++) q[g][k] = q[ci][k];
            break;
        case BT_INVWALL:
        case BT_SYMMETRY: {
            double v[2] = {q[ci][1] / q[ci][0], q[ci][2] / q[ci][0]}, nv[2];
            reflect_about_edge(m, e, v, nv);
            q[g][0] = q[ci][0]; q[g][1] = q[ci][0] * nv[0]; q[g][2] = q[ci][0] * nv[1]; q[g][3] = q[ci][3];
            break;
        }
        case BT_VISCWALL:
            q[g][0] = q[ci][0]; q[g][1] = -q[ci][1]; q[g][2] = -q[ci][2]; q[g][3] = q[ci][3];
            break;
        case BT_CONSTP: {
            double rhop = q[ci][0], up = q[ci][1] / q[ci][0], vp = q[ci][2] / q[ci][0];
            double pp = orc_get_pressure(q[ci]);
            double ap_ = sqrt(GAMMA * (pp / rhop));
            double Jp = up + (2.0 * ap_) / (GAMMA - 1);
            double Sp = pp / pow(rhop, GAMMA);
            double pb = e->bData[0];
            double rhob = pow(pb / Sp, 1.0 / GAMMA);
            double ab = sqrt(GAMMA * (pb / rhob));
            double ub = Jp - (2.0 * ab) / (GAMMA - 1);
            double vb = vp;
            double mag = sqrt(ub * ub + vb * vb);
            double rEb = pb / (GAMMA - 1) + 0.5 * rhob * (mag * mag);
            double vpn = up * e->normal[0] + vp * e->normal[1];
            double vbn = ub * e->normal[0] + vb * e->normal[1];
            double Vbg[2] = {up - vpn * e->normal[0] + vbn * e->normal[0], vp - vpn * e->normal[1] + vbn * e->normal[1]};
            q[g][0] = rhob; q[g][1] = rhob * Vbg[0]; q[g][2] = rhob * Vbg[1]; q[g][3] = rEb;
            break;
        }
        case BT_TPINFLOW: {
            /* restated with the reference's four writes to q[g][0] (solve.d:678-681) */
            double Tt = e->bData[0], Pt = e->bData[1];
            double aoa = (double)((long double)e->bData[2] * 3.14159265358979323846264338327950288L / 180.0L);
            double rhop = q[ci][0], up = q[ci][1] / q[ci][0];
            double pp = orc_get_pressure(q[ci]);
            double ap_ = sqrt(GAMMA * (pp / rhop));
            double Jp = up + (2.0 * ap_) / (GAMMA - 1);
            double nin[2] = {(double)cosl((long double)aoa), (double)sinl((long double)aoa)};
            double d = nin[0] * e->normal[0] + nin[1] * e->normal[1];
            double Rg = cfg->R;
            double a = GAMMA * Rg * Tt * (d * d) - 0.5 * (GAMMA - 1) * (Jp * Jp);
            double bq = (4.0 * GAMMA * Rg * Tt * d) / (GAMMA - 1);
            double c = (4.0 * GAMMA * Rg * Tt) / ((GAMMA - 1) * (GAMMA - 1)) - Jp * Jp;
            double M1 = (-bq + sqrt(bq * bq - 4 * a * c)) / (2.0 * a);
            double M2 = (-bq - sqrt(bq * bq - 4 * a * c)) / (2.0 * a);
            double Mb;
            if (M1 >= 0 && M2 >= 0) Mb = fmin(M1, M2);
            else if (M1 >= 0 && M2 < 0) Mb = M1;
            else if (M1 < 0 && M2 >= 0) Mb = M2;
            else { ORC_FIRST(firstG, kindG, b, ORC_ERR_BC); continue; }
            double Tb = Tt / (1 + 0.5 * (GAMMA - 1) * (Mb * Mb));
            double pb = Pt * pow(Tb / Tt, GAMMA / (GAMMA - 1));
            double rhob = pb / (Rg * Tb);
            double ab = sqrt(GAMMA * (pb / rhob));
            double vb[2] = {Mb * ab * nin[0], Mb * ab * nin[1]};
            double mag = sqrt(vb[0] * vb[0] + vb[1] * vb[1]);
            double rEb = pb / (GAMMA - 1) + 0.5 * rhob * (mag * mag);
            q[g][0] = rhob; q[g][0] = rhob * vb[0]; q[g][0] = rhob * vb[1]; q[g][0] = rEb;
            break;
        }
        default:
            ORC_FIRST(firstG, kindG, b, ORC_ERR_BC);
            continue;
        }
    }
    if (firstG != UINT32_MAX) {
        m->errKind = kindG; m->errIndex = m->cells[m->nElem + firstG].edges[0];
        return kindG;
    }

    /* buildGradients, solve.d:324-508 */
    if (cfg->order > 1) {
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < m->nElem; i++) {
            OCell *c = &m->cells[i];
            double du[4][MAXE];
            memset(du, 0, sizeof(du));
            for (int k = 0; k < 4; k++) { c->minQ[k] = q[i][k]; c->maxQ[k] = q[i][k]; c->lim[k][0] = 1.0; c->lim[k][1] = 1.0; }
            for (uint32_t j = 0; j < c->nNeighborCells; j++) {
                uint32_t idx = c->neighborCells[j];
                for (int k = 0; k < 4; k++) {
                    c->minQ[k] = fmin(c->minQ[k], q[idx][k]);
                    c->maxQ[k] = fmax(c->maxQ[k], q[idx][k]);
                    du[k][j] = q[idx][k] - q[i][k];
                }
            }
            for (int j = 0; j < 4; j++)
                for (int r = 0; r < 2; r++) {
                    double s = 0.0;
                    for (int k = 0; k < MAXE; k++) s += c->gradMat[r][k] * du[j][k];
                    c->gradient[j][r] = s;
                }
            if (cfg->limited && limit && cfg->limiter == 1) {
                /* Venkatakrishnan, J. Comput. Phys. 118 (1995), eq. 15 with the smooth min():
                   for the unlimited change D2 to a test point and the bound difference D1,
                   phi = [(D1^2 + eps^2) D2 + 2 D2^2 D1] / [D2 (D1^2 + 2 D2^2 + D1 D2 + eps^2)],
                   at the neighbour centroids the reference's scalar limiter uses; no counterpart
                   in the reference */
                const double kd = cfg->venkatK * c->d, eps2 = kd * kd * kd;
                for (uint32_t j = 0; j < c->nNeighborCells; j++) {
                    const double *cn = m->cells[c->neighborCells[j]].centroid;
                    double dx = cn[0] - c->centroid[0], dy = cn[1] - c->centroid[1];
                    for (int k = 0; k < 4; k++) {
                        const double d2 = c->gradient[k][0] * dx + c->gradient[k][1] * dy;
                        double phi = 1.0;
                        if (d2 != 0.0) {
                            const double d1 = d2 > 0.0 ? c->maxQ[k] - q[i][k] : c->minQ[k] - q[i][k];
                            phi = ((d1 * d1 + eps2) * d2 + 2.0 * (d2 * d2) * d1) /
                                  (d2 * (d1 * d1 + 2.0 * (d2 * d2) + d1 * d2 + eps2));
                        }
                        c->lim[k][0] = fmin(c->lim[k][0], phi);
                        c->lim[k][1] = c->lim[k][0];
                    }
                }
            } else if (cfg->limited && limit) {
                for (uint32_t j = 0; j < c->nNeighborCells; j++) {
                    const double *cn = m->cells[c->neighborCells[j]].centroid;
                    double dx = cn[0] - c->centroid[0], dy = cn[1] - c->centroid[1];
                    for (int k = 0; k < 4; k++) {
                        double s = 1.0;
                        double qE = q[i][k] + (c->gradient[k][0] * dx + c->gradient[k][1] * dy);
                        if (qE < c->minQ[k]) s = (c->minQ[k] - q[i][k]) / (c->gradient[k][0] * dx + c->gradient[k][1] * dy);
                        else if (qE > c->maxQ[k]) s = (c->maxQ[k] - q[i][k]) / (c->gradient[k][0] * dx + c->gradient[k][1] * dy);
                        c->lim[k][0] = fmin(c->lim[k][0], s);
                        c->lim[k][1] = c->lim[k][0];
                    }
                }
                if (cfg->lpThresh > 0) {
                    for (int k = 0; k < 4; k++) {
                        if (c->lim[k][0] < cfg->lpThresh && (c->nNeighborCells == 3 || c->nNeighborCells == 4)) {
                            int rows = 4 + 2 * (int)c->nNeighborCells;
                            double A[24], bb[12], cc[2], xk[2] = {0, 0};
                            memset(A, 0, sizeof(A)); memset(bb, 0, sizeof(bb));
                            cc[0] = -fabs(c->gradient[k][0]); cc[1] = -fabs(c->gradient[k][1]);
                            A[0] = 1; A[1] = 0; A[2] = 0; A[3] = 1; A[4] = -1; A[5] = 0; A[6] = 0; A[7] = -1;
                            bb[0] = 0; bb[1] = 0; bb[2] = -1; bb[3] = -1;
                            for (uint32_t j = 0, ci2 = 0; j < c->nNeighborCells; j++, ci2 += 2) {
                                const double *cn = m->cells[c->neighborCells[j]].centroid;
                                A[2 * (ci2 + 4)] = (cn[0] - c->centroid[0]) * c->gradient[k][0];
                                A[2 * (ci2 + 4) + 1] = (cn[1] - c->centroid[1]) * c->gradient[k][1];
                                bb[ci2 + 4] = c->minQ[k] - q[i][k];
                                A[2 * (ci2 + 5)] = -(cn[0] - c->centroid[0]) * c->gradient[k][0];
                                A[2 * (ci2 + 5) + 1] = -(cn[1] - c->centroid[1]) * c->gradient[k][1];
                                bb[ci2 + 5] = q[i][k] - c->maxQ[k];
                            }
                            orc_lp(rows, A, bb, cc, xk);
                            c->lim[k][0] = xk[0];
                            c->lim[k][1] = xk[1];
                        }
                    }
                }
            }
        }
    }

    /* boundary edges, solve.d:769-968 */
    uint32_t firstB = UINT32_MAX;
    int kindB = 0;
#pragma omp parallel for schedule(static)
    for (uint32_t bi = 0; bi < m->nBoundaryEdges; bi++) {
        uint32_t i = m->boundaryEdges[bi];
        OEdge *e = &m->edges[i];
        update_boundary_edge(m, q, i, cfg);
        int bt = e->bType;
        if (bt == BT_DIRICHLET) {
            OCell *g = &m->cells[e->cellIdx[1]];
            double gg[8];
            orc_solution_gradient(g->centroid[0], g->centroid[1], gg);
            for (int k = 0; k < 4; k++) { g->gradient[k][0] = gg[2 * k]; g->gradient[k][1] = gg[2 * k + 1]; }
            orc_solution(e->mid[0], e->mid[1], e->q[1]);
            bt = BT_FULLSTATE;
        }
        switch (bt) {
        case BT_FULLSTATE: {
            double F[4];
            flux_dispatch(cfg->flux, e->q[0], e->q[1], e->normal, F, &e->sMax);
            if (cfg->viscosity) {
                double qAve[4], dqAve[8], Fv[4];
                const OCell *c0 = &m->cells[e->cellIdx[0]], *c1 = &m->cells[e->cellIdx[1]];
                for (int k = 0; k < 4; k++) qAve[k] = 0.5 * (e->q[0][k] + e->q[1][k]);
                for (int k = 0; k < 4; k++) for (int r = 0; r < 2; r++) dqAve[2 * k + r] = 0.5 * (c0->gradient[k][r] + c1->gradient[k][r]);
                orc_diffusive_flux(PrC, mu, qAve, dqAve, e->normal, Fv);
                for (int k = 0; k < 4; k++) e->flux[k] = F[k] - Fv[k];
            } else {
                for (int k = 0; k < 4; k++) e->flux[k] = F[k];
            }
            break;
        }
        case BT_INVWALL: {
            double inv = 1 / e->q[0][0];
            double velP[2] = {inv * e->q[0][1], inv * e->q[0][2]};
            double vn = velP[0] * e->normal[0] + velP[1] * e->normal[1];
            double t0 = velP[0] - vn * e->normal[0], t1 = velP[1] - vn * e->normal[1];
            double vel = sqrt(t0 * t0 + t1 * t1);
            double p = (GAMMA - 1) * (e->q[0][3] - 0.5 * e->q[0][0] * (vel * vel));
            double a = sqrt(GAMMA * (p / e->q[0][0]));
            if (p < 0) p = 1.0e-12;
            e->flux[0] = 0; e->flux[1] = p * e->normal[0]; e->flux[2] = p * e->normal[1]; e->flux[3] = 0;
            e->sMax = fabs(a);
            for (int k = 0; k < 4; k++) e->q[1][k] = q[e->cellIdx[1]][k];
            break;
        }
        case BT_VISCWALL: {
            double vel = 0.0;
            double p = (GAMMA - 1) * (e->q[0][3] - 0.5 * e->q[0][0] * (vel * vel));
            double a = sqrt(GAMMA * (p / e->q[0][0]));
            if (p < 0) p = 1.0e-12;
            double qL[4], qR[4], qAve[4], dqAve[8], Fv[4];
            for (int k = 0; k < 4; k++) qL[k] = e->q[0][k];
            for (int k = 0; k < 4; k++) qR[k] = qL[k];
            qR[1] *= -1; qR[2] *= -1;
            for (int k = 0; k < 4; k++) e->q[1][k] = qR[k];
            for (int k = 0; k < 4; k++) qAve[k] = 0.5 * (qL[k] + qR[k]);
            qAve[1] = 0.0; qAve[2] = 0.0;
            const OCell *c0 = &m->cells[e->cellIdx[0]];
            for (int k = 0; k < 4; k++) { dqAve[2 * k] = c0->gradient[k][0]; dqAve[2 * k + 1] = c0->gradient[k][1]; }
            dqAve[6] = (qAve[3] / qAve[0]) * dqAve[0];
            dqAve[7] = (qAve[3] / qAve[0]) * dqAve[1];
            orc_diffusive_flux(PrC, mu, qAve, dqAve, e->normal, Fv);
            double P[4] = {0, p * e->normal[0], p * e->normal[1], 0};
            for (int k = 0; k < 4; k++) e->flux[k] = P[k] - Fv[k];
            e->sMax = fabs(a);
            break;
        }
        case BT_CONSTP: {
            const double *qg = q[e->cellIdx[1]];
            double p = orc_get_pressure(qg);
            double u = qg[1] / qg[0], v = qg[2] / qg[0];
            double a = sqrt(GAMMA * (p / qg[0]));
            for (int k = 0; k < 4; k++) e->q[1][k] = qg[k];
            orc_convective_flux(p, u, v, qg[0], qg[3], e->normal, e->flux);
            e->sMax = fabs(a);
            break;
        }
        case BT_SYMMETRY: {
            double v[2] = {e->q[0][1] / e->q[0][0], e->q[0][2] / e->q[0][0]}, nv[2];
            reflect_about_edge(m, e, v, nv);
            e->q[1][0] = e->q[0][0]; e->q[1][1] = e->q[0][0] * nv[0]; e->q[1][2] = e->q[0][0] * nv[1]; e->q[1][3] = e->q[0][3];
            double F[4];
            flux_dispatch(cfg->flux, e->q[0], e->q[1], e->normal, F, &e->sMax);
            if (cfg->viscosity) {
                double qAve[4], Fv[4];
                for (int k = 0; k < 4; k++) qAve[k] = 0.5 * (e->q[0][k] + e->q[1][k]);
                const OCell *c0 = &m->cells[e->cellIdx[0]];
                const double n0 = e->normal[0], n1 = e->normal[1];
                double nO[4] = {n0 * n0, n0 * n1, n1 * n0, n1 * n1};
                double A[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
                double V[4] = {1 - nO[0], 0 - nO[1], 0 - nO[2], 1 - nO[3]};
                A[5] = V[0]; A[6] = V[1]; A[9] = V[2]; A[10] = V[3];
                double gp[8];
                for (int k = 0; k < 4; k++) { gp[2 * k] = c0->gradient[k][0]; gp[2 * k + 1] = c0->gradient[k][1]; }
                double tmp1[8];
                for (int r = 0; r < 4; r++) for (int cc = 0; cc < 2; cc++) {
                    double s = 0.0;
                    for (int k = 0; k < 4; k++) s += A[4 * r + k] * gp[2 * k + cc];
                    tmp1[2 * r + cc] = s;
                }
                double tmp2[4] = {1 - 2 * nO[0], 0 - 2 * nO[1], 0 - 2 * nO[2], 1 - 2 * nO[3]};
                double bGrad[8];
                for (int r = 0; r < 4; r++) for (int cc = 0; cc < 2; cc++) {
                    double s3 = 0.0, s1 = 0.0;
                    for (int k = 0; k < 2; k++) s3 += tmp1[2 * r + k] * tmp2[2 * k + cc];
                    for (int k = 0; k < 2; k++) s1 += gp[2 * r + k] * nO[2 * k + cc];
                    bGrad[2 * r + cc] = s1 + s3;
                }
                orc_diffusive_flux(PrC, mu, qAve, bGrad, e->normal, Fv);
                for (int k = 0; k < 4; k++) e->flux[k] = F[k] - Fv[k];
            } else {
                for (int k = 0; k < 4; k++) e->flux[k] = F[k];
            }
            break;
        }
        case BT_TPINFLOW: {
            for (int k = 0; k < 4; k++) e->q[1][k] = q[e->cellIdx[1]][k];
            double F[4];
            flux_dispatch(cfg->flux, e->q[0], e->q[1], e->normal, F, &e->sMax);
            if (cfg->viscosity) {
                double qAve[4], g1[8], Fv[4];
                for (int k = 0; k < 4; k++) qAve[k] = 0.5 * (e->q[0][k] + e->q[1][k]);
                const OCell *c0 = &m->cells[e->cellIdx[0]];
                for (int k = 0; k < 4; k++) { g1[2 * k] = c0->gradient[k][0]; g1[2 * k + 1] = c0->gradient[k][1]; }
                orc_diffusive_flux(PrC, mu, qAve, g1, e->normal, Fv);
                for (int k = 0; k < 4; k++) e->flux[k] = F[k] - Fv[k];
            } else {
                for (int k = 0; k < 4; k++) e->flux[k] = F[k];
            }
            break;
        }
        default:
            ORC_FIRST(firstB, kindB, bi, ORC_ERR_BC);
            continue;
        }
        if (has_nan4(e->flux)) ORC_FIRST(firstB, kindB, bi, ORC_ERR_EDGE_NAN);
    }
    if (firstB != UINT32_MAX) {
        m->errKind = kindB; m->errIndex = m->boundaryEdges[firstB];
        return kindB;
    }

    /* interior edges, solve.d:970-1033 */
    uint32_t firstI = UINT32_MAX;
    int kindI = 0;
#pragma omp parallel for schedule(static)
    for (uint32_t ii = 0; ii < m->nInteriorEdges; ii++) {
        uint32_t i = m->interiorEdges[ii];
        OEdge *e = &m->edges[i];
        if (cfg->order == 1) {
            for (int k = 0; k < 4; k++) { e->q[0][k] = q[e->cellIdx[0]][k]; e->q[1][k] = q[e->cellIdx[1]][k]; }
        } else {
            for (int s = 0; s < 2; s++) {
                const OCell *c = &m->cells[e->cellIdx[s]];
                const double *qM = q[e->cellIdx[s]];
                double dx = e->mid[0] - c->centroid[0], dy = e->mid[1] - c->centroid[1];
                for (int j = 0; j < 4; j++)
                    e->q[s][j] = qM[j] + c->lim[j][0] * c->gradient[j][0] * dx + c->lim[j][1] * c->gradient[j][1] * dy;
                if (orc_get_pressure(e->q[s]) < 0) {
                    double lim[2] = {1.0, 1.0};
                    for (int j = 0; j < 4; j++) { lim[0] = fmin(lim[0], c->lim[j][0]); lim[1] = fmin(lim[1], c->lim[j][1]); }
                    for (int j = 0; j < 4; j++)
                        e->q[s][j] = qM[j] + lim[0] * c->gradient[j][0] * dx + lim[1] * c->gradient[j][1] * dy;
                }
            }
        }
        double F[4];
        flux_dispatch(cfg->flux, e->q[0], e->q[1], e->normal, F, &e->sMax);
        if (cfg->viscosity) {
            double qAve[4], dqAve[8], Fv[4];
            const OCell *c0 = &m->cells[e->cellIdx[0]], *c1 = &m->cells[e->cellIdx[1]];
            for (int k = 0; k < 4; k++) qAve[k] = 0.5 * (e->q[0][k] + e->q[1][k]);
            for (int k = 0; k < 4; k++) for (int r = 0; r < 2; r++) dqAve[2 * k + r] = 0.5 * (c0->gradient[k][r] + c1->gradient[k][r]);
            orc_diffusive_flux(PrC, mu, qAve, dqAve, e->normal, Fv);
            for (int k = 0; k < 4; k++) e->flux[k] = F[k] - Fv[k];
        } else {
            for (int k = 0; k < 4; k++) e->flux[k] = F[k];
        }
        if (has_nan4(e->flux)) ORC_FIRST(firstI, kindI, ii, ORC_ERR_EDGE_NAN);
    }
    if (firstI != UINT32_MAX) {
        m->errKind = kindI; m->errIndex = m->interiorEdges[firstI];
        return kindI;
    }

    /* residual, solve.d:1080-1118 */
    double ndAll = INFINITY, rmAll = *Rmax;
#pragma omp parallel
    {
    double ndT = INFINITY, rmT = 0;
    double *newDtT = &ndT, *RmaxT = &rmT;
#pragma omp for schedule(static)
    for (uint32_t i = 0; i < m->nElem; i++) {
        OCell *c = &m->cells[i];
        double r[4] = {0, 0, 0, 0};
        double sAve = 0;
        for (uint32_t j = 0; j < c->nEdges; j++) {
            const OEdge *e = &m->edges[c->edges[j]];
            double fl = c->fluxMultiplier[j] * e->len;
            for (int k = 0; k < 4; k++) r[k] += fl * e->flux[k];
            sAve += e->len * e->sMax;
        }
        sAve /= c->perim;
        *newDtT = fmin(*newDtT, (cfg->CFL * c->d) / sAve);
        if (cfg->localTimestep && dtUpdate) {
            c->dt = cfg->CFL * c->d / sAve;
            if (isnan(c->dt)) c->dt = 0;
        }
        for (int k = 0; k < 4; k++) if (fabs(r[k]) > *RmaxT) *RmaxT = fabs(r[k]);
        double s = -(1.0 / c->area);
        for (int k = 0; k < 4; k++) R[i][k] = r[k] * s;
    }
#pragma omp critical(orc_reduce)
    {
        ndAll = fmin(ndAll, ndT);
        if (rmT > rmAll) rmAll = rmT;
    }
    }
    *newDt = ndAll;
    *Rmax = rmAll;
    return ORC_OK;
}

/* solver call as seen by the integrators: ufvmSolver, then the MMS source
   (app/ms.d:131-141, added after the -1/area scaling) when the case is manufactured */
static int stage_residual(OrcMesh *m, const OrcConfig *cfg, double (*R)[4], double (*q)[4],
                          double *newDt, double *Rmax, int limit, int dtUpdate)
{
    int rc = orc_ufvm_solver(m, cfg, R, q, newDt, Rmax, limit, dtUpdate);
    if (rc) return rc;
    if (cfg->mms)
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < m->nElem; i++)
            for (int k = 0; k < 4; k++) R[i][k] = R[i][k] + m->src[i][k];
    return ORC_OK;
}

int orc_residual(OrcMesh *m, const OrcConfig *cfg, double *R, double *newDt, double *Rmax, int limit, int dtUpdate)
{
    return stage_residual(m, cfg, (double (*)[4])R, m->q, newDt, Rmax, limit, dtUpdate);
}

/* ------------------------------------------------------------------ */
/* source/integrators.d:22-270                                          */
/* ------------------------------------------------------------------ */
int orc_step(OrcMesh *m, const OrcConfig *cfg, double *dt, double *Rmax)
{
    double newDt = INFINITY;
    int rc;
    double (*q)[4] = m->q;
    const uint32_t N = m->nElem;
    const int lts = cfg->localTimestep;
    switch (cfg->integrator) {
    case INT_EULER:
        if ((rc = stage_residual(m, cfg, m->R, q, &newDt, Rmax, 1, 1))) return rc;
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < N; i++) {
            double h = lts ? m->cells[i].dt : *dt;
            for (int k = 0; k < 4; k++) q[i][k] = q[i][k] + h * m->R[i][k];
        }
        break;
    case INT_RK4: {
        const double h = *dt;
        if ((rc = stage_residual(m, cfg, m->k1, q, &newDt, Rmax, 1, 1))) return rc;
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < N; i++) for (int k = 0; k < 4; k++) m->tmp[i][k] = q[i][k] + ((h / 2.0) * m->k1[i][k]);
        if ((rc = stage_residual(m, cfg, m->k2, m->tmp, &newDt, Rmax, cfg->multistageLimiting, 0))) return rc;
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < N; i++) for (int k = 0; k < 4; k++) m->tmp[i][k] = q[i][k] + ((h / 2.0) * m->k2[i][k]);
        if ((rc = stage_residual(m, cfg, m->k3, m->tmp, &newDt, Rmax, cfg->multistageLimiting, 0))) return rc;
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < N; i++) for (int k = 0; k < 4; k++) m->tmp[i][k] = q[i][k] + (h * m->k3[i][k]);
        if ((rc = stage_residual(m, cfg, m->k4, m->tmp, &newDt, Rmax, cfg->multistageLimiting, 0))) return rc;
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < N; i++) {
            double hh = lts ? m->cells[i].dt : h;
            for (int k = 0; k < 4; k++)
                q[i][k] = q[i][k] + (hh / 6.0) * (m->k1[i][k] + 2.0 * m->k2[i][k] + 2.0 * m->k3[i][k] + m->k4[i][k]);
        }
        break;
    }
    case INT_RK2TVD: {
        const double h = *dt;
        if ((rc = stage_residual(m, cfg, m->R, q, &newDt, Rmax, 1, lts))) return rc;
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < N; i++) {
            double hh = lts ? m->cells[i].dt : h;
            for (int k = 0; k < 4; k++) m->k1[i][k] = q[i][k] + hh * m->R[i][k];  /* qFE */
        }
        if ((rc = stage_residual(m, cfg, m->k2, m->k1, &newDt, Rmax, 0, !lts))) return rc;
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < N; i++) {
            double hh = lts ? m->cells[i].dt : h;
            for (int k = 0; k < 4; k++) q[i][k] = 0.5 * (q[i][k] + m->k1[i][k] + hh * m->k2[i][k]);
        }
        break;
    }
    case INT_RK2: {
        const double h = *dt;
        if ((rc = stage_residual(m, cfg, m->k1, q, &newDt, Rmax, 1, 1))) return rc;
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < N; i++) for (int k = 0; k < 4; k++) m->tmp[i][k] = q[i][k] + (h * m->k1[i][k]);
        if ((rc = stage_residual(m, cfg, m->k2, m->tmp, &newDt, Rmax, cfg->multistageLimiting, 0))) return rc;
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < N; i++) {
            double hh = lts ? m->cells[i].dt : h;
            for (int k = 0; k < 4; k++) q[i][k] = q[i][k] + (hh / 2.0) * (m->k1[i][k] + m->k2[i][k]);
        }
        break;
    }
    default:
        return ORC_ERR_ARG;
    }
    *dt = newDt;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* app/solver.d:38-277 runIntegrator (single rank), bounded by maxSteps */
/* hist: per step {t, dt, rhoRMS, uRMS, vRMS, ERMS, Rmax, drag, lift, CFL} */
/* state[0..1] in/out: t, dt (t=0, dt=config.dt on a fresh start)      */
/* ------------------------------------------------------------------ */
#define ORC_HIST 10
int orc_run(OrcMesh *m, OrcConfig *cfg, int64_t maxSteps, double *state, double *hist, int64_t *stepsDone)
{
    const uint32_t L = m->nCells, N = m->nElem;
    double *last = (double *)calloc((size_t)4 * L, sizeof(double));
    double *now = (double *)calloc((size_t)4 * L, sizeof(double));
    double *tmp = (double *)calloc(L, sizeof(double));
    double t = state[0], dt = state[1];
    double residRhoLast = INFINITY;
    uint32_t residRhoIncIters = 0;
    int64_t it = 0;
    int rc = ORC_OK;
    long double deg = (long double)cfg->ic[1] * (3.14159265358979323846264338327950288L / 180.0L);
    const double rc00 = (double)cosl(deg), rc01 = -(double)sinl(deg), rc10 = (double)sinl(deg), rc11 = (double)cosl(deg);
    for (uint32_t i = 0; i < N; i++) for (int k = 0; k < 4; k++) last[(size_t)k * L + i] = m->q[i][k];
    while (t < cfg->tEnd && it < maxSteps) {
        double Rmax = 0;
        double newDt = dt;
        rc = orc_step(m, cfg, &newDt, &Rmax);
        if (rc) break;
        if (cfg->dynamicDt) dt = newDt;
        for (uint32_t i = 0; i < N; i++) {
            for (int k = 0; k < 4; k++) now[(size_t)k * L + i] = m->q[i][k];
            if (has_nan4(m->q[i])) { m->errKind = ORC_ERR_CELL_NAN; m->errIndex = i; rc = ORC_ERR_CELL_NAN; break; }
        }
        if (rc) break;
        double rms[4];
        for (int k = 0; k < 4; k++) {
            const double *nw = now + (size_t)k * L, *ls = last + (size_t)k * L;
            for (uint32_t i = 0; i < L; i++) { double dd = nw[i] - ls[i]; tmp[i] = dd * dd; }
            double tmpSum = orc_pairwise_sum(tmp, L);
            double valSum = orc_pairwise_sum(nw, L);
            rms[k] = sqrt((double)N * tmpSum) / valSum;
        }
        memcpy(last, now, sizeof(double) * 4 * (size_t)L);
        /* mesh.d:645-671 computeBoundaryForces */
        double f[2] = {0, 0};
        for (int ft = 0; ft < cfg->nForce; ft++) {
            int tg = cfg->forceTags[ft];
            if (tg < 0) continue;
            for (uint32_t j = 0; j < m->bGroupLen[tg]; j++) {
                const OEdge *e = &m->edges[m->bGroups[tg][j]];
                f[0] += e->flux[1] * e->len;
                f[1] += e->flux[2] * e->len;
            }
        }
        double ld0 = 0.0 + rc00 * f[0] + rc01 * f[1];
        double ld1 = 0.0 + rc10 * f[0] + rc11 * f[1];
        double residMax = dmax(dmax(dmax(rms[0], rms[1]), rms[2]), rms[3]);
        if (cfg->cflAdjust) {
            if (residRhoLast < residMax) residRhoIncIters++;
            else if (residRhoIncIters >= 1) residRhoIncIters--;
            if (residRhoIncIters >= 2000) { cfg->CFL *= 0.99; residRhoIncIters = 0; }
        }
        residRhoLast = residMax;
        if (hist) {
            double *h = hist + ORC_HIST * it;
            h[0] = t; h[1] = dt; h[2] = rms[0]; h[3] = rms[1]; h[4] = rms[2]; h[5] = rms[3];
            h[6] = Rmax; h[7] = ld0; h[8] = ld1; h[9] = cfg->CFL;
        }
        t += dt;
        it++;
    }
    state[0] = t; state[1] = dt;
    *stepsDone = it;
    free(last); free(now); free(tmp);
    return rc;
}
