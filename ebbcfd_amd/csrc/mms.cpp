// Manufactured solution (source/manufacturedsolution.d:27-283) and freestream
// state (source/solve.d:177-184).  Host-only: the solver evaluates these once at
// set-up (ghost Dirichlet states, edge states, ghost gradients, per-cell source),
// so the device never needs sin/cos/pow and the values are bit-identical to the
// CPU restatement.
#include <cmath>

#include "host.hpp"

namespace ebb {

namespace {
const double ar = 0.9, br = 0.04, cr = -2, dr = 1;
const double au = 0.1, bu = 0.02, cu = 1, du = 1;
const double av = 0.05, bv = -0.1, cv = 0.7, dv = 1.3;
const double ap = 1, bp = 0.05, cp = 2, dp = -1;

double rho(double x, double y) { return ar + br * std::sin(cr * x + dr * y); }
double drdx(double x, double y) { return br * cr * std::cos(cr * x + dr * y); }
double d2rdx2(double x, double y) { return -br * (cr * cr) * std::sin(cr * x + dr * y); }
double drdy(double x, double y) { return br * dr * std::cos(cr * x + dr * y); }
double d2rdy2(double x, double y) { return -br * (dr * dr) * std::sin(cr * x + dr * y); }
double uf(double x, double y) { return au + bu * std::cos(cu * x + du * y); }
double dudx(double x, double y) { return -bu * cu * std::sin(cu * x + du * y); }
double d2udxy(double x, double y) { return -bu * cu * du * std::cos(cu * x + du * y); }
double d2udx2(double x, double y) { return -bu * (cu * cu) * std::cos(cu * x + du * y); }
double dudy(double x, double y) { return -bu * du * std::sin(cu * x + du * y); }
double d2udy2(double x, double y) { return -bu * (du * du) * std::cos(cu * x + du * y); }
double vf(double x, double y) { return av + bv * std::cos(cv * x + dv * y); }
double dvdx(double x, double y) { return -bv * cv * std::sin(cv * x + dv * y); }
double dvdy(double x, double y) { return -bv * dv * std::sin(cv * x + dv * y); }
double d2vdx2(double x, double y) { return -bv * (cv * cv) * std::cos(cv * x + dv * y); }
double d2vdy2(double x, double y) { return -bv * (dv * dv) * std::cos(cv * x + dv * y); }
double d2vdxy(double x, double y) { return -bv * cv * dv * std::cos(cv * x + dv * y); }
double pf(double x, double y) { return ap + bp * std::sin(cp * x + dp * y); }
double dpdx(double x, double y) { return bp * cp * std::cos(cp * x + dp * y); }
double d2pdx2(double x, double y) { return -bp * (cp * cp) * std::sin(cp * x + dp * y); }
double dpdy(double x, double y) { return bp * dp * std::cos(cp * x + dp * y); }
double d2pdy2(double x, double y) { return -bp * (dp * dp) * std::sin(cp * x + dp * y); }

double pressure(const double q[4])
{
    const double a = q[1] / q[0], b = q[2] / q[0];
    return (GAMMA - 1) * (q[3] - 0.5 * q[0] * (a * a + b * b));
}
}  // namespace

void mmsSolution(double x, double y, double q[4])
{
    q[0] = rho(x, y);
    q[1] = rho(x, y) * uf(x, y);
    q[2] = rho(x, y) * vf(x, y);
    const double u = uf(x, y), v = vf(x, y);
    q[3] = pf(x, y) / (GAMMA - 1.0) + 0.5 * rho(x, y) * (u * u + v * v);
}

void mmsSolutionGradient(double x, double y, double g[8])
{
    const double r = rho(x, y), u = uf(x, y), v = vf(x, y);
    const double rx = drdx(x, y), ry = drdy(x, y), ux = dudx(x, y), uy = dudy(x, y), vx = dvdx(x, y), vy = dvdy(x, y);
    g[0] = rx;
    g[1] = ry;
    g[2] = u * rx + r * ux;
    g[3] = u * ry + r * uy;
    g[4] = v * rx + r * vx;
    g[5] = v * ry + r * vy;
    const double dru2dx = 2.0 * r * u * ux + (u * u) * rx;
    const double dru2dy = 2.0 * r * u * uy + (u * u) * ry;
    const double drv2dx = 2.0 * r * v * vx + (v * v) * rx;
    const double drv2dy = 2.0 * r * v * vy + (v * v) * ry;
    g[6] = 1.0 / (GAMMA - 1.0) * dpdx(x, y) + 0.5 * (dru2dx + drv2dx);
    g[7] = 1.0 / (GAMMA - 1.0) * dpdy(x, y) + 0.5 * (dru2dy + drv2dy);
}

void mmsSourceTerm(double x, double y, bool viscosity, double Rg, double mu, double Pr, double S[4])
{
    double q[4], dq[8];
    mmsSolution(x, y, q);
    mmsSolutionGradient(x, y, dq);
    const double r = q[0], u = q[1] / q[0], v = q[2] / q[0], ru = q[1], rv = q[2];
    const double rx = dq[0], ry = dq[1], rux = dq[2], rvy = dq[5];
    const double p = pressure(q);
    const double ux = dudx(x, y), uy = dudy(x, y), vy = dvdy(x, y), vx = dvdx(x, y);
    const double px = dpdx(x, y), py = dpdy(x, y);
    const double rEx = dq[6], rEy = dq[7], rE = q[3];
    const double fx[4] = {rux,
                          2.0 * ru * ux + (u * u) * rx + px,
                          u * v * rx + r * u * vx + r * v * ux,
                          u * rEx + rE * ux + u * px + p * ux};
    const double fy[4] = {rvy,
                          u * v * ry + r * u * vy + r * v * uy,
                          2.0 * rv * vy + (v * v) * ry + py,
                          v * rEy + rE * vy + v * py + p * vy};
    for (int k = 0; k < 4; k++) S[k] = fx[k] + fy[k];
    if (viscosity) {
        double gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0};
        const double kk = (GAMMA * Rg * mu) / ((GAMMA - 1) * Pr);
        const double uxx = d2udx2(x, y), uyy = d2udy2(x, y), uxy = d2udxy(x, y);
        const double vxx = d2vdx2(x, y), vyy = d2vdy2(x, y), vxy = d2vdxy(x, y);
        const double pxx = d2pdx2(x, y), pyy = d2pdy2(x, y);
        const double rxx = d2rdx2(x, y), ryy = d2rdy2(x, y);
        gx[1] = mu * (2.0 / 3.0) * (2 * uxx - vxy);
        gx[2] = mu * (uxy + vxx);
        gx[3] = mu * (2.0 / 3.0) * (2 * (u * uxx + ux * ux) - (u * vxy + ux * vy));
        gx[3] += mu * ((v * uxy + vx * uy) + (v * vxx + vx * vx));
        const double Txx = (1.0 / Rg) * ((1.0 / r) * pxx + (1.0 / (r * r)) * rx * px
                           - (1.0 / (r * r * r)) * (r * px - 2.0 * p * rx) * rx - (p / (r * r)) * rxx);
        gx[3] += kk * Txx;
        gy[1] = mu * (uyy + vxy);
        gy[2] = mu * (2.0 / 3.0) * (2 * vyy - uxy);
        gy[3] = mu * (2.0 / 3.0) * (2.0 * (v * vyy + vy * vy) - (v * uxy + vy * ux));
        gy[3] += mu * (u * uyy + uy * uy + u * vxy + uy * vx);
        const double Tyy = (1.0 / Rg) * ((1.0 / r) * pyy + (1.0 / (r * r)) * ry * py
                           - (1.0 / (r * r * r)) * (r * py - 2.0 * p * ry) * ry - (p / (r * r)) * ryy);
        gy[3] += kk * Tyy;
        for (int k = 0; k < 4; k++) S[k] -= (gx[k] + gy[k]);
    }
}

// ic[1]*(PI/180) is evaluated in 80-bit precision by the reference (PI is a real constant)
double aoaRadians(double aoaDeg)
{
    return double((long double)aoaDeg * (3.14159265358979323846264338327950288L / 180.0L));
}

void freestreamState(double M, double aoaDeg, double rho0, double q[4])
{
    const double aoa = aoaRadians(aoaDeg);
    const double U = 1.0, a = U / M;
    const double p = (rho0 * (a * a)) / GAMMA;
    const double u = U * double(cosl((long double)aoa)), v = U * double(sinl((long double)aoa));
    q[0] = rho0;
    q[1] = rho0 * u;
    q[2] = rho0 * v;
    q[3] = p / (GAMMA - 1.0) + 0.5 * rho0 * (u * u + v * v);
}

}  // namespace ebb
