// HIP kernels for the explicit residual-and-update loop, FP64, gfx950.
//
// One residual evaluation (source/solve.d:322-1122) is three launches:
//   k_gradient  per interior cell: ghost states of its boundary faces
//               (solve.d:520-689), LSQ gradient, min/max bounds and the scalar
//               limiter at neighbour centroids (solve.d:324-403), LP vector
//               limiter (solve.d:404-497);
//   k_faces     per face, each computed once: reconstruction at the midpoint with
//               the p<0 fallback (solve.d:283-319, 970-1011), Roe/Rusanov flux
//               (flux.d:66-303), viscous flux (flux.d:28-64), boundary fluxes
//               (solve.d:769-968); stores len*flux and len*sMax per face;
//   k_update    per interior cell: gathers its faces in the cell's own edge
//               order with the owner sign (solve.d:1082-1118), CFL time step
//               and Rmax (block reduction + one ordered-bits atomic per block,
//               exact because min/max are order independent), the MMS source
//               (app/ms.d:131-141) and the integrator stage update
//               (source/integrators.d:22-270) fused in.
// No atomics touch the residual: each cell sums its own faces in a fixed order,
// so results are deterministic and bit-identical to the edge-loop form.
// Built with -ffp-contract=off: every product and sum rounds as written.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "dev.hpp"

namespace ebb {

#define DEV __device__ __forceinline__
constexpr double G_ = 1.4;
constexpr int BLK = 256;
static_assert(uint32_t(BLK) == FACE_TILE_LANES, "face tile padding assumes one lane per workgroup thread");
constexpr int FORCE_LDS = 1024;   // force faces staged per pass by a gradient kernel's force workgroup

enum { BT_FULL = 0, BT_CONSTP, BT_INVWALL, BT_VISCWALL, BT_SYM, BT_TPIN, BT_DIRICHLET };

// workgroup b runs on XCD b % 8; hand each XCD a contiguous slice of the cells
DEV uint32_t xcdBlock(uint32_t b, uint32_t nb)
{
    const uint32_t nx = 8, per = nb / nx, rem = nb % nx, x = b % nx, k = b / nx;
    return x * per + (x < rem ? x : rem) + k;
}

// The same slices, each walked last to first (the k-th workgroup of XCD x takes the k-th
// item from its slice's end; k < the slice's length for every b < nb)
DEV uint32_t xcdBlockBack(uint32_t b, uint32_t nb)
{
    const uint32_t nx = 8, per = nb / nx, rem = nb % nx, x = b % nx, k = b / nx;
    const uint32_t len = per + (x < rem ? 1u : 0u);
    return x * per + (x < rem ? x : rem) + (len - 1 - k);
}

// The face, reconstruction and limiter arithmetic below is written once for the
// real type R it computes in: double on the FP64 path, float on the mixed path
// (extensions.precision = "mixed").  Every literal is converted to R, so the double
// instantiation performs exactly the operations of the FP64 path and the float one
// never promotes to double.
template <typename R>
DEV R pressure(const R q[4])
{
    const R a = q[1] / q[0], b = q[2] / q[0];
    return R(G_ - 1) * (q[3] - R(0.5) * q[0] * (a * a + b * b));
}

template <typename R>
DEV R dmax(R a, R b) { return a < b ? b : a; }

template <typename R>
DEV R harten(R lam, R eps)
{
    if (lam < eps) lam = (eps * eps + lam * lam) / (R(2.0) * eps);
    return lam;
}

// flux.d:19-24
template <typename R>
DEV void convective(R p, R u, R v, R rho, R e, R n0, R n1, R out[4])
{
    out[0] = (rho * u) * n0 + (rho * v) * n1;
    out[1] = (p + rho * (u * u)) * n0 + (rho * u * v) * n1;
    out[2] = (rho * u * v) * n0 + (p + rho * (v * v)) * n1;
    out[3] = (u * (p + e)) * n0 + (v * (p + e)) * n1;
}

// flux.d:224-303
template <typename R>
DEV void roeFlux(const R qL[4], const R qR[4], R n0, R n1, R out[4], R& sMax)
{
    const R gm1 = R(G_ - 1), half = R(0.5);
    const R rhoL = qL[0], uL = qL[1] / rhoL, vL = qL[2] / rhoL;
    const R rvL0 = rhoL * uL, rvL1 = rhoL * vL;
    const R pL = (gm1 * (qL[3] - half * rhoL * (uL * uL + vL * vL)));
    const R hL = (qL[3] + pL) / rhoL;
    const R rhoR = qR[0], uR = qR[1] / rhoR, vR = qR[2] / rhoR;
    const R rvR0 = rhoR * uR, rvR1 = rhoR * vR;
    const R pR = (gm1 * (qR[3] - half * rhoR * (uR * uR + vR * vR)));
    const R hR = (qR[3] + pR) / rhoR;
    const R di = sqrt(rhoR / rhoL);
    const R d1 = R(1.0) / (R(1.0) + di);
    const R vel0 = (di * uR + uL) * d1, vel1 = (di * vR + vL) * d1;
    const R u = vel0 * n0 + vel1 * n1;
    const R h = (di * hR + hL) * d1;
    const R a = sqrt(R(G_ - 1.0) * (h - half * (vel0 * vel0 + vel1 * vel1)));
    const R dRhoV0 = rvR0 - rvL0, dRhoV1 = rvR1 - rvL1;
    const R dRho = rhoR - rhoL, dRhoE = qR[3] - qL[3];
    const R eps = R(0.01) * a;
    const R lam1 = harten<R>(fabs(u + a), eps);
    const R lam2 = harten<R>(fabs(u - a), eps);
    const R lam3 = harten<R>(fabs(u), eps);
    const R lam4 = harten<R>(fabs(u), eps);
    sMax = dmax<R>(dmax<R>(dmax<R>(fabs(lam1), fabs(lam2)), fabs(lam3)), fabs(lam4));
    const R s1 = half * (lam1 + lam2), s2 = half * (lam1 - lam2);
    const R qq = sqrt(vel0 * vel0 + vel1 * vel1);
    const R G1 = gm1 * (((qq * qq) / R(2.0)) * dRho - (vel0 * dRhoV0 + vel1 * dRhoV1) + dRhoE);
    const R G2 = -u * dRho + (dRhoV0 * n0 + dRhoV1 * n1);
    const R C1 = (G1 / (a * a)) * (s1 - lam3) + G2 / a * s2;
    const R C2 = (G1 / a) * s2 + (s1 - lam3) * G2;
    R fL[4], fR[4];
    convective<R>(pL, uL, vL, rhoL, qL[3], n0, n1, fL);
    convective<R>(pR, uR, vR, rhoR, qR[3], n0, n1, fR);
    const R d[4] = {lam3 * dRho + C1, lam3 * dRhoV0 + C1 * vel0 + C2 * n0,
                    lam3 * dRhoV1 + C1 * vel1 + C2 * n1, lam3 * dRhoE + C1 * h + C2 * u};
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = half * (fL[k] + fR[k]) - half * d[k];
}

// flux.d:66-141 (Roe-type averages, C1 = C2 = 0: central flux + sMax jump)
template <typename R>
DEV void rusanovFlux(const R qL[4], const R qR[4], R n0, R n1, R out[4], R& sMax)
{
    const R gm1 = R(G_ - 1), half = R(0.5);
    const R rhoL = qL[0], uL = qL[1] / rhoL, vL = qL[2] / rhoL;
    const R rvL0 = rhoL * uL, rvL1 = rhoL * vL;
    const R pL = (gm1 * (qL[3] - half * rhoL * (uL * uL + vL * vL)));
    const R hL = (qL[3] + pL) / rhoL;
    const R rhoR = qR[0], uR = qR[1] / rhoR, vR = qR[2] / rhoR;
    const R rvR0 = rhoR * uR, rvR1 = rhoR * vR;
    const R pR = (gm1 * (qR[3] - half * rhoR * (uR * uR + vR * vR)));
    const R hR = (qR[3] + pR) / rhoR;
    const R rho = sqrt(rhoL * rhoR);
    const R sL = sqrt(rhoL), sR = sqrt(rhoR);
    const R vel0 = (sL * uL + sR * uR) / (sL + sR);
    const R vel1 = (sL * vL + sR * vR) / (sL + sR);
    const R u = vel0 * n0 + vel1 * n1;
    const R p = (sL * pL + sR * pR) / (sL + sR);
    const R h = (sL * hL + sR * hR) / (sL + sR);
    const R a = sqrt(R(G_) * (p / rho));
    const R dRhoV0 = rvR0 - rvL0, dRhoV1 = rvR1 - rvL1;
    const R dRho = rhoR - rhoL, dRhoE = qR[3] - qL[3];
    const R eps = R(0.01) * a;
    const R lam1 = harten<R>(fabs(u + a), eps);
    const R lam2 = harten<R>(fabs(u - a), eps);
    const R lam3 = harten<R>(fabs(u), eps);
    const R lam4 = harten<R>(fabs(u), eps);
    const R sm = dmax<R>(dmax<R>(dmax<R>(fabs(lam1), fabs(lam2)), fabs(lam3)), fabs(lam4));
    sMax = sm;
    const R s1 = half * (sm + sm), s2 = half * (sm - sm);
    const R qq = sqrt(vel0 * vel0 + vel1 * vel1);
    const R G1 = gm1 * ((qq * qq / R(2)) * dRho - (vel0 * dRhoV0 + vel1 * dRhoV1) + dRhoE);
    const R G2 = -u * dRho + (dRhoV0 * n0 + dRhoV1 * n1);
    const R C1 = (G1 / (a * a)) * (s1 - sm) + G2 / a * s2;
    const R C2 = (G1 / a) * s2 + (s1 - sm) * G2;
    R fL[4], fR[4];
    convective<R>(pL, uL, vL, rhoL, qL[3], n0, n1, fL);
    convective<R>(pR, uR, vR, rhoR, qR[3], n0, n1, fR);
    const R d[4] = {sm * dRho + C1, sm * (dRhoV0) + C1 * vel0 + C2 * n0,
                    sm * (dRhoV1) + C1 * vel1 + C2 * n1, sm * dRhoE + C1 * h + C2 * u};
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = half * (fL[k] + fR[k]) - half * d[k];
}

template <int FLUX, typename R>
DEV void numFlux(const R qL[4], const R qR[4], R n0, R n1, R out[4], R& sMax)
{
    if (FLUX == 1) roeFlux<R>(qL, qR, n0, n1, out, sMax);
    else rusanovFlux<R>(qL, qR, n0, n1, out, sMax);
}

// flux.d:28-64.  The four 4x4 matrices are applied entry by entry in column
// order; their structural zeros only ever add a signed zero, so they are skipped.
// Pr and mu are physical parameters (FP64), converted to R.
template <typename R>
DEV void viscousFlux(double Prd, double mud, const R q[4], const R dq[8], R n0, R n1, R out[4])
{
    const R rho = q[0], u = q[1] / q[0], v = q[2] / q[0], E = q[3] / q[0];
    const R gp = R(G_) / R(Prd);
    const R s = R(mud) / rho;
    const R one = R(1.0);
    const R c43 = R(4.0 / 3.0), c23 = R(2.0 / 3.0), cm43 = R(-4.0 / 3.0), cm23 = R(-2.0 / 3.0), c231 = R(2.0 / 3.0 - 1.0);
    const R t0 = (cm43 + gp) * (u * u) + (-one + gp) * (v * v) - gp * E;
    const R t1 = (-one + gp) * (u * u) + (cm43 + gp) * (v * v) - gp * E;
    const R x0 = dq[0], x1 = dq[2], x2 = dq[4], x3 = dq[6];
    const R y0 = dq[1], y1 = dq[3], y2 = dq[5], y3 = dq[7];
    const R uv = c231 * u * v;
    const R G01 = ((s * (cm43 * u)) * x0 + (s * c43) * x1) + ((s * (c23 * v)) * y0 + (s * cm23) * y2);
    const R G02 = ((s * -v) * x0 + (s * one) * x2) + ((s * -u) * y0 + (s * one) * y1);
    const R G03 = ((s * t0) * x0 + (s * ((c43 - gp) * u)) * x1 + (s * ((one - gp) * v)) * x2 + (s * gp) * x3)
                + ((s * uv) * y0 + (s * v) * y1 + (s * (cm23 * u)) * y2);
    const R G11 = ((s * -v) * x0 + (s * one) * x2) + ((s * -u) * y0 + (s * one) * y1);
    const R G12 = ((s * (c23 * u)) * x0 + (s * cm23) * x2) + ((s * (cm43 * v)) * y0 + (s * c43) * y2);
    const R G13 = ((s * uv) * x0 + (s * (cm23 * v)) * x1 + (s * u) * x2)
                + ((s * t1) * y0 + (s * ((one - gp) * u)) * y1 + (s * ((c43 - gp) * v)) * y2 + (s * gp) * y3);
    out[0] = R(0.0);
    out[1] = G01 * n0 + G11 * n1;
    out[2] = G02 * n0 + G12 * n1;
    out[3] = G03 * n0 + G13 * n1;
}

DEV void ld4(const double4* p, uint32_t i, double o[4])
{
    const double4 t = p[i];
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
}
DEV void st4(double4* p, uint32_t i, const double o[4]) { p[i] = make_double4(o[0], o[1], o[2], o[3]); }
// gradients and limiters are read once, by the next kernel: stored non-temporally so
// they do not displace the neighbour data that kernel's gathers reuse from L2
typedef double ntd2 __attribute__((ext_vector_type(2)));
typedef float ntf4 __attribute__((ext_vector_type(4)));

// the FP32 shadow of a state row (mixed path): the stage's gathers read it in place of the FP64 row
DEV void ld4(const float4* p, uint32_t i, float o[4])
{
    const float4 t = p[i];
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
}
DEV void st4(float4* p, uint32_t i, const float o[4]) { p[i] = make_float4(o[0], o[1], o[2], o[3]); }

// ghost state from the owner state (solve.d:520-689); the owner is the only writer
// false where the reference would stop: a total-conditions inflow with no non-negative Mach
// root (solve.d:661-670)
template <typename R>
DEV bool ghostState(const BFace& b, const R qo[4], R qg[4], const R qgOld[4])
{
    const R g = R(G_), gm1 = R(G_ - 1), half = R(0.5), two = R(2.0);
    switch (b.type) {
    case BT_DIRICHLET:
        qg[0] = R(b.gq[0]); qg[1] = R(b.gq[1]); qg[2] = R(b.gq[2]); qg[3] = R(b.gq[3]);
        break;
    case BT_INVWALL:
    case BT_SYM: {
        const R v0 = qo[1] / qo[0], v1 = qo[2] / qo[0];
        R w0, w1;
        if (!b.vertical) {
            w0 = R(0.0) + R(b.R[0]) * v0 + R(b.R[1]) * v1;
            w1 = R(0.0) + R(b.R[2]) * v0 + R(b.R[3]) * v1;
        } else {
            w0 = -v0;
            w1 = v1;
        }
        qg[0] = qo[0]; qg[1] = qo[0] * w0; qg[2] = qo[0] * w1; qg[3] = qo[3];
        break;
    }
    case BT_VISCWALL:
        qg[0] = qo[0]; qg[1] = -qo[1]; qg[2] = -qo[2]; qg[3] = qo[3];
        break;
    case BT_CONSTP: {
        const R rhop = qo[0], up = qo[1] / qo[0], vp = qo[2] / qo[0];
        const R pp = pressure<R>(qo);
        const R ap = sqrt(g * (pp / rhop));
        const R Jp = up + (two * ap) / gm1;
        const R Sp = pp / pow(rhop, g);
        const R pb = R(b.data[0]);
        const R rhob = pow(pb / Sp, R(1.0) / g);
        const R ab = sqrt(g * (pb / rhob));
        const R ub = Jp - (two * ab) / gm1;
        const R vb = vp;
        const R mag = sqrt(ub * ub + vb * vb);
        const R rEb = pb / gm1 + half * rhob * (mag * mag);
        const R n0 = R(b.nrm[0]), n1 = R(b.nrm[1]);
        const R vpn = up * n0 + vp * n1, vbn = ub * n0 + vb * n1;
        qg[0] = rhob;
        qg[1] = rhob * (up - vpn * n0 + vbn * n0);
        qg[2] = rhob * (vp - vpn * n1 + vbn * n1);
        qg[3] = rEb;
        break;
    }
    case BT_TPIN: {
        // solve.d:625-683, including its four writes to component 0
        const R Tt = R(b.data[0]), Pt = R(b.data[1]), ni0 = R(b.data[2]), ni1 = R(b.data[3]);
        const R rhop = qo[0], up = qo[1] / qo[0];
        const R pp = pressure<R>(qo);
        const R ap = sqrt(g * (pp / rhop));
        const R Jp = up + (two * ap) / gm1;
        const R d = ni0 * R(b.nrm[0]) + ni1 * R(b.nrm[1]);
        const R Rg = R(b.Rgas);
        const R A = g * Rg * Tt * (d * d) - half * gm1 * (Jp * Jp);
        const R B = (R(4.0) * g * Rg * Tt * d) / gm1;
        const R Cc = (R(4.0) * g * Rg * Tt) / (gm1 * gm1) - Jp * Jp;
        const R M1 = (-B + sqrt(B * B - R(4) * A * Cc)) / (two * A);
        const R M2 = (-B - sqrt(B * B - R(4) * A * Cc)) / (two * A);
        R Mb = R(NAN);
        bool rootFound = true;
        if (M1 >= 0 && M2 >= 0) Mb = fmin(M1, M2);
        else if (M1 >= 0 && M2 < 0) Mb = M1;
        else if (M1 < 0 && M2 >= 0) Mb = M2;
        else rootFound = false;   // the reference stops here (enforce)
        const R Tb = Tt / (R(1) + half * gm1 * (Mb * Mb));
        const R pb = Pt * pow(Tb / Tt, g / gm1);
        const R rhob = pb / (Rg * Tb);
        const R ab = sqrt(g * (pb / rhob));
        const R vb0 = Mb * ab * ni0, vb1 = Mb * ab * ni1;
        const R mag = sqrt(vb0 * vb0 + vb1 * vb1);
        qg[0] = pb / gm1 + half * rhob * (mag * mag);
        qg[1] = qgOld[1]; qg[2] = qgOld[2]; qg[3] = qgOld[3];
        return rootFound;
    }
    default:  // FullState
        qg[0] = qo[0]; qg[1] = qo[1]; qg[2] = qo[2]; qg[3] = qo[3];
        break;
    }
    return true;
}

// solve.d:27-173 LP vector limiter (all-inequality simplex, at most 16 pivots)
DEV void lpSolve(int rows, const double* A, const double* b, const double c[2], double xk[2])
{
    int W0 = 0, W1 = 1;
    double Ak[4] = {A[0], A[1], A[2], A[3]};
    double gam[12];
    int D[12];
    for (int k = 0; k <= 15; k++) {
        double det = Ak[0] * Ak[3] - Ak[1] * Ak[2];
        double s = 1.0 / det;
        const double iT0 = Ak[3] * s, iT1 = -Ak[2] * s, iT2 = -Ak[1] * s, iT3 = Ak[0] * s;
        const double lam0 = 0.0 + iT0 * c[0] + iT1 * c[1];
        const double lam1 = 0.0 + iT2 * c[0] + iT3 * c[1];
        if ((lam0 >= -10e-9) && (lam1 >= -10e-9)) break;
        const int q = (lam0 < lam1) ? W0 : W1;
        const double e0 = (lam0 < lam1) ? 1.0 : 0.0, e1 = (lam0 < lam1) ? 0.0 : 1.0;
        det = Ak[0] * Ak[3] - Ak[1] * Ak[2];
        s = 1.0 / det;
        const double i0 = Ak[3] * s, i1 = -Ak[1] * s, i2 = -Ak[2] * s, i3 = Ak[0] * s;
        const double pk0 = 0.0 + i0 * e0 + i1 * e1;
        const double pk1 = 0.0 + i2 * e0 + i3 * e1;
        int Dlen = 0;
        for (int i = 0; i < rows; i++)
            if (i != W0 && i != W1)
                if (A[2 * i] * pk0 + A[2 * i + 1] * pk1 < -1e-11) D[Dlen++] = i;
        double alpha = INFINITY;
        for (int i = 0; i < Dlen; i++) {
            const double aixk = A[2 * D[i]] * xk[0] + A[2 * D[i] + 1] * xk[1];
            const double aipk = A[2 * D[i]] * pk0 + A[2 * D[i] + 1] * pk1;
            gam[i] = (aixk - b[D[i]]) / (-aipk);
            alpha = fmin(alpha, gam[i]);
        }
        int t = 0;
        for (int i = 0; i < Dlen; i++) if (gam[i] == alpha) { t = D[i]; break; }
        if (q == W0) W0 = t; else if (q == W1) W1 = t;
        Ak[0] = A[2 * W0]; Ak[1] = A[2 * W0 + 1]; Ak[2] = A[2 * W1]; Ak[3] = A[2 * W1 + 1];
        xk[0] = xk[0] + alpha * pk0;
        xk[1] = xk[1] + alpha * pk1;
    }
}

// ---------------------------------------------------------------- gradient + limiter
// solve.d:324-497 for one interior cell, given its own state, its S neighbour
// states (ghosts already filled), LSQ weights, centroids and which slots are real
// neighbours (padding slots repeat the cell).  MODE 1: unlimited; 2: scalar
// limiter; 3: scalar limiter then the LP vector limiter.
// Least-squares weights (A^T A)^-1 A^T with A_j = centroid_j - centroid_i, in the
// operation order of the mesh set-up (source/mesh.d; mesh.cpp), so they equal the
// stored ones bit for bit; the padding slots (A_j = 0) add exact zeros, as the
// set-up's zero rows do.
template <int S>
DEV void lsqWeights(double2 ci, const double2 cn[S], double w[2][S])
{
    double ax[S], ay[S];
#pragma unroll
    for (int j = 0; j < S; j++) { ax[j] = cn[j].x - ci.x; ay[j] = cn[j].y - ci.y; }
    double b00 = 0.0, b01 = 0.0, b10 = 0.0, b11 = 0.0;
#pragma unroll
    for (int j = 0; j < S; j++) {
        b00 += ax[j] * ax[j];
        b01 += ax[j] * ay[j];
        b10 += ay[j] * ax[j];
        b11 += ay[j] * ay[j];
    }
    const double det = b00 * b11 - b01 * b10;
    const double id = 1.0 / det;
    const double i00 = b11 * id, i01 = -b01 * id, i10 = -b10 * id, i11 = b00 * id;
#pragma unroll
    for (int j = 0; j < S; j++) {
        w[0][j] = i00 * ax[j] + i01 * ay[j];
        w[1][j] = i10 * ax[j] + i11 * ay[j];
    }
}

// gm: stored weights ((dir, nbr)-major, stride gstride between entries), or null
// to compute them from the centroids.  Weights and centroid offsets are geometry,
// computed in FP64 and converted to R; the gradient, bounds and limiter are R.
// MODE 4: the Venkatakrishnan limiter at the same test points (neighbour centroids): for
// the unlimited increment t = grad . (x_j - x_i) and the bound difference d = max - q_i (t > 0)
// or min - q_i (t < 0), phi_j = [(d^2 + eps2) t + 2 t^2 d] / (t [d^2 + 2 t^2 + d t + eps2]),
// phi = min(1, min_j phi_j); eps2 = (K d_cell)^3 (FP64, passed in).
template <int S, int MODE, typename R>
DEV void gradLimit(const R qi[4], const R qn[S][4], const double* gm, uint32_t gstride, double2 ci,
                   const double2 cn[S], const bool real[S], int limit, double lpThresh, R g[4][2], R lx[4],
                   R ly[4], double eps2d = 0.0)
{
    static_assert(MODE != 3 || sizeof(R) == 8, "the LP vector limiter runs in FP64 only");
    R w[2][S];
    {
        double wd[2][S];
        if (gm) {
#pragma unroll
            for (int r = 0; r < 2; r++)
#pragma unroll
                for (int j = 0; j < S; j++) wd[r][j] = gm[(r * S + j) * gstride];
        } else {
            lsqWeights<S>(ci, cn, wd);
        }
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int j = 0; j < S; j++) w[r][j] = R(wd[r][j]);
    }
    R mn[4], mx[4], du[4][S];
#pragma unroll
    for (int k = 0; k < 4; k++) { mn[k] = qi[k]; mx[k] = qi[k]; }
#pragma unroll
    for (int j = 0; j < S; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            mn[k] = fmin(mn[k], qn[j][k]);
            mx[k] = fmax(mx[k], qn[j][k]);
            du[k][j] = qn[j][k] - qi[k];
        }
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int r = 0; r < 2; r++) {
            R s = R(0.0);
#pragma unroll
            for (int j = 0; j < S; j++) s += w[r][j] * du[k][j];
            g[k][r] = s;
        }
#pragma unroll
    for (int k = 0; k < 4; k++) { lx[k] = R(1.0); ly[k] = R(1.0); }
    if (MODE == 4 && limit) {
        const R eps2 = R(eps2d), two = R(2.0);
#pragma unroll
        for (int j = 0; j < S; j++) {
            const R dx = R(cn[j].x - ci.x), dy = R(cn[j].y - ci.y);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const R t = g[k][0] * dx + g[k][1] * dy;
                R s = R(1.0);
                if (t != R(0.0)) {
                    const R d = t > R(0.0) ? mx[k] - qi[k] : mn[k] - qi[k];
                    s = ((d * d + eps2) * t + two * (t * t) * d) / (t * (d * d + two * (t * t) + d * t + eps2));
                }
                lx[k] = fmin(lx[k], s);
                ly[k] = lx[k];
            }
        }
        return;
    }
    if (MODE >= 2 && MODE <= 3 && limit) {
        R dxs[S], dys[S];
#pragma unroll
        for (int j = 0; j < S; j++) {
            const R dx = R(cn[j].x - ci.x), dy = R(cn[j].y - ci.y);
            dxs[j] = dx; dys[j] = dy;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                R s = R(1.0);
                const R t = g[k][0] * dx + g[k][1] * dy;
                const R qE = qi[k] + t;
                if (qE < mn[k]) s = (mn[k] - qi[k]) / t;
                else if (qE > mx[k]) s = (mx[k] - qi[k]) / t;
                lx[k] = fmin(lx[k], s);
                ly[k] = lx[k];
            }
        }
        if constexpr (MODE == 3) {
            if (lpThresh > 0) {
                int nn = 0;
#pragma unroll
                for (int j = 0; j < S; j++) nn += real[j] ? 1 : 0;
                for (int k = 0; k < 4; k++) {
                    if (lx[k] < lpThresh && (nn == 3 || nn == 4)) {
                        double A[24], bb[12];
                        for (int r = 0; r < 24; r++) A[r] = 0.0;
                        for (int r = 0; r < 12; r++) bb[r] = 0.0;
                        A[0] = 1; A[3] = 1; A[4] = -1; A[7] = -1;
                        bb[2] = -1; bb[3] = -1;
                        int c2 = 0;
                        for (int j = 0; j < S; j++) {
                            if (!real[j]) continue;
                            A[2 * (c2 + 4)] = dxs[j] * g[k][0];
                            A[2 * (c2 + 4) + 1] = dys[j] * g[k][1];
                            bb[c2 + 4] = mn[k] - qi[k];
                            A[2 * (c2 + 5)] = -dxs[j] * g[k][0];
                            A[2 * (c2 + 5) + 1] = -dys[j] * g[k][1];
                            bb[c2 + 5] = qi[k] - mx[k];
                            c2 += 2;
                        }
                        const double cc[2] = {-fabs(g[k][0]), -fabs(g[k][1])};
                        double xk[2] = {0.0, 0.0};
                        lpSolve(4 + 2 * nn, A, bb, cc, xk);
                        lx[k] = xk[0];
                        ly[k] = xk[1];
                    }
                }
            }
        }
    }
}

// the owner's ghost state for boundary record b, with TempPresInflow's untouched components
DEV bool ghostFromOwner(const BFace& bf, const double qo[4], const double4* qbuf, uint32_t gIdx, double qg[4])
{
    double old[4] = {0, 0, 0, 0};
    if (bf.type == BT_TPIN) ld4(qbuf, gIdx, old);
    return ghostState(bf, qo, qg, old);
}

// The boundary force sum of mesh.d:645-671: (len*flux) momentum components of the
// force faces added in order, two interleaved chains, by thread 0 (its result).
// fp: LDS staging of CAP (x, y) pairs.
template <int NT, int CAP>
DEV void orderedForce(const uint32_t* faces, uint32_t nF, const double* lenF, double2* fp, double& f0, double& f1)
{
    f0 = 0; f1 = 0;
    for (uint32_t k0 = 0; k0 < nF; k0 += CAP) {
        const uint32_t cnt = (nF - k0) < uint32_t(CAP) ? (nF - k0) : uint32_t(CAP);
        __syncthreads();
        // all of a thread's face ids, then all of their fluxes, are loaded before any is
        // used: two round trips per pass instead of two per face
        constexpr int FS = (CAP + NT - 1) / NT;
        uint32_t fi[FS];
#pragma unroll
        for (int u = 0; u < FS; u++) {
            const uint32_t x = threadIdx.x + uint32_t(u) * NT;
            fi[u] = x < cnt ? faces[k0 + x] : 0u;
        }
        double2 fv[FS];
#pragma unroll
        for (int u = 0; u < FS; u++) {
            const uint32_t x = threadIdx.x + uint32_t(u) * NT;
            fv[u] = x < cnt ? make_double2(lenF[size_t(fi[u]) * 4 + 1], lenF[size_t(fi[u]) * 4 + 2])
                            : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < FS; u++) {
            const uint32_t x = threadIdx.x + uint32_t(u) * NT;
            if (x < cnt) fp[x] = fv[u];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // the two ordered chains interleave; 32 LDS reads are in flight ahead of the adds
            uint32_t x = 0;
            for (; x + 32 <= cnt; x += 32) {
                double2 v[32];
#pragma unroll
                for (int u = 0; u < 32; u++) v[u] = fp[x + u];
#pragma unroll
                for (int u = 0; u < 32; u++) { f0 += v[u].x; f1 += v[u].y; }
            }
            for (; x < cnt; x++) { const double2 v = fp[x]; f0 += v.x; f1 += v.y; }
        }
    }
}

// a pending force (k_finalize left it): rotated drag/lift into the step's history row
// and the scalars, as k_finalize would have written them
template <int NT, int CAP>
DEV void pendingForce(const ForceArgs& f, DevScalars* sc, double2* fp)
{
    if (!sc->forcePending) return;
    double f0, f1;
    orderedForce<NT, CAP>(f.faces, f.n, f.lenF, fp, f0, f1);
    if (threadIdx.x == 0) {
        const double ld0 = 0.0 + f.rot[0] * f0 + f.rot[1] * f1;
        const double ld1 = 0.0 + f.rot[2] * f0 + f.rot[3] * f1;
        const int64_t slot = sc->forceSlot;
        if (slot >= 0) {
            double* h = f.hist + slot * HIST_FIELDS;
            h[7] = ld0; h[8] = ld1;
        }
        sc->force[0] = ld0; sc->force[1] = ld1;
        sc->forcePending = 0;
    }
}

// MODE 0: ghost states only (order 1); 1: unlimited gradient (no limiter stored);
// 2: scalar limiter; 3: scalar limiter followed by the LP vector limiter (lim
// stored as x,y pairs).
// MIX 0: FP64 throughout.  MIX 1 (mixed precision): states are read from the FP32
// shadows, ghosts, gradients and limiters computed in FP32 (least-squares weights and
// centroid offsets in FP64, converted); gradients and limiters are stored as FP32
// rows (32 + 16 bytes per cell); each ghost is written to the shadow and, widened, to
// the FP64 row.
// LDS of one gradient workgroup: staging of the gradient rows (and of the force faces)
// and of the limiter rows
template <int MODE>
struct GradLds {
    static constexpr int LWS = MODE == 3 ? 8 : ((MODE == 2 || MODE == 4) ? 4 : 0);
    double sg[BLK][9];
    double sl[BLK][LWS + 1];
};

template <int MIX>
using RealOf = typename std::conditional<MIX == 1, float, double>::type;

// a stage-input row: the FP32 shadow on the mixed path
template <int MIX>
DEV void ldStage(const StageArgs& a, uint32_t c, RealOf<MIX> o[4])
{
    if constexpr (MIX == 1) ld4(a.qin32, c, o);
    else ld4(a.qin, c, o);
}
// a ghost row of the stage input; the mixed path also widens it into the FP64 row
template <int MIX>
DEV void stGhostRow(const StageArgs& a, uint32_t c, const RealOf<MIX> o[4])
{
    if constexpr (MIX == 1) {
        st4(a.qin32, c, o);
        a.qin[c] = make_double4(o[0], o[1], o[2], o[3]);
    } else {
        st4(a.qin, c, o);
    }
}

// Gradient and limiter of interior cell i (MODE 1-4) from the stage input, with the ghost
// states of its boundary faces, which are stored to the stage buffer where storeGhosts;
// nbrRow: its S neighbour ids (the mesh's row, or a tile's copy of it).
// Every neighbour row is loaded before any ghost is written, so the loads of all
// neighbours are in flight together.  A ghost slot belongs to this cell alone (the owner of
// its boundary face), so its row still holds the stage buffer's previous ghost state, which
// TempPresInflow keeps in the components it does not write.
template <int S, int MODE, int MIX>
DEV void cellGradient(const DevMesh& m, const PhysParams& p, const StageArgs& a, uint32_t i, const uint32_t* nbrRow,
                      bool storeGhosts, RealOf<MIX> g[4][2], RealOf<MIX> lx[4], RealOf<MIX> ly[4])
{
    using R = RealOf<MIX>;
    R qi[4];
    ldStage<MIX>(a, i, qi);
    uint32_t nb[S];
#pragma unroll
    for (int j = 0; j < S; j++) nb[j] = nbrRow[j];
    R qn[S][4];
    double2 cn[S];
#pragma unroll
    for (int j = 0; j < S; j++) { ldStage<MIX>(a, nb[j], qn[j]); cn[j] = m.cen[nb[j]]; }
    const double2 ci = m.cen[i];
#pragma unroll
    for (int j = 0; j < S; j++)
        if (nb[j] >= m.N) {
            const BFace& bf = m.bface[nb[j] - m.N];
            R old[4] = {0, 0, 0, 0};
            if (bf.type == BT_TPIN) {
#pragma unroll
                for (int k = 0; k < 4; k++) old[k] = qn[j][k];
            }
            if (!ghostState<R>(bf, qi, qn[j], old) && a.recordErrors)
                atomicMin(&a.sc->errBC, (uint64_t(a.stageId) << 32) | (nb[j] - m.N));
            if (storeGhosts) stGhostRow<MIX>(a, nb[j], qn[j]);
        }
    bool real[S];
#pragma unroll
    for (int j = 0; j < S; j++) real[j] = nb[j] != i;
    double eps2 = 0.0;
    if constexpr (MODE == 4) {
        const double kd = p.venkatK * m.cd[i];
        eps2 = kd * kd * kd;
    }
    gradLimit<S, MODE, R>(qi, qn, nullptr, 0, ci, cn, real, a.limit, p.lpThresh, g, lx, ly, eps2);
}

// One workgroup of the gradient kernel on cell block `bid` of its mesh (the solo and the
// batched kernels run this same code).
template <int S, int MODE, int MIX>
DEV void gradientBody(const DevMesh& m, const PhysParams& p, const StageArgs& a, uint32_t bid, GradLds<MODE>& L)
{
    using R = RealOf<MIX>;
    constexpr int LWS = GradLds<MODE>::LWS;
    auto& sg = L.sg;
    auto& sl = L.sl;
    if (a.checkActive && !a.sc->active) return;
    if (bid == 0 && threadIdx.x == 0 && a.first) a.sc->rmaxBits = 0;   // Rmax is per step
    const uint32_t i = bid * BLK + threadIdx.x;
    if (MODE == 0 && i >= m.N) return;
    const bool have = i < m.N;   // no early return below: the staged stores need every thread
    R g[4][2], lx[4], ly[4];
    if (have) {
        if constexpr (MODE == 0) {
            // order 1: the ghost states are the kernel's only output
            R qi[4];
            ldStage<MIX>(a, i, qi);
#pragma unroll
            for (int j = 0; j < S; j++) {
                const uint32_t nb = m.nbr[i * S + j];
                if (nb >= m.N) {
                    const BFace& bf = m.bface[nb - m.N];
                    R qg[4], old[4] = {0, 0, 0, 0};
                    if (bf.type == BT_TPIN) ldStage<MIX>(a, nb, old);
                    if (!ghostState<R>(bf, qi, qg, old) && a.recordErrors)
                        atomicMin(&a.sc->errBC, (uint64_t(a.stageId) << 32) | (nb - m.N));
                    stGhostRow<MIX>(a, nb, qg);
                }
            }
            return;
        } else {
            cellGradient<S, MODE, MIX>(m, p, a, i, m.nbr + size_t(i) * S, true, g, lx, ly);
        }
    }
    if (MODE == 0) return;
    // stage the rows in LDS and store the workgroup's contiguous run of cells with
    // lane-consecutive 16-byte stores
    const uint32_t tid = threadIdx.x;
    const uint32_t base = bid * BLK, n = (m.N - base) < BLK ? (m.N - base) : BLK;
    if constexpr (MIX == 1) {
        float* fg = reinterpret_cast<float*>(&sg[0][0]);   // 8 floats per row
        float* fl = reinterpret_cast<float*>(&sl[0][0]);   // 4 floats per row
        if (have) {
#pragma unroll
            for (int c = 0; c < 4; c++) { fg[8 * tid + 2 * c] = g[c][0]; fg[8 * tid + 2 * c + 1] = g[c][1]; }
            if (MODE == 2 || MODE == 4) {
#pragma unroll
                for (int c = 0; c < 4; c++) fl[4 * tid + c] = lx[c];
            }
        }
        __syncthreads();
        float4* gd = reinterpret_cast<float4*>(a.grad32 + size_t(base) * 8);
        for (uint32_t c = tid; c < 2 * n; c += BLK)
            __builtin_nontemporal_store(*reinterpret_cast<const ntf4*>(fg + 4 * c), reinterpret_cast<ntf4*>(gd) + c);
        if (MODE == 2 || MODE == 4) {
            float4* ld = reinterpret_cast<float4*>(a.lim32 + size_t(base) * 4);
            for (uint32_t c = tid; c < n; c += BLK)
                __builtin_nontemporal_store(*reinterpret_cast<const ntf4*>(fl + 4 * c), reinterpret_cast<ntf4*>(ld) + c);
        }
        return;
    } else {
        if (have) {
#pragma unroll
            for (int c = 0; c < 4; c++) { sg[tid][2 * c] = g[c][0]; sg[tid][2 * c + 1] = g[c][1]; }
            if (MODE == 3) {
#pragma unroll
                for (int c = 0; c < 4; c++) { sl[tid][2 * c] = lx[c]; sl[tid][2 * c + 1] = ly[c]; }
            } else if (MODE == 2 || MODE == 4) {
#pragma unroll
                for (int c = 0; c < 4; c++) sl[tid][c] = lx[c];
            }
        }
        __syncthreads();
        double2* gd = reinterpret_cast<double2*>(a.grad + size_t(base) * 8);
        for (uint32_t c = tid; c < 4 * n; c += BLK) __builtin_nontemporal_store((ntd2){sg[c / 4][2 * (c % 4)], sg[c / 4][2 * (c % 4) + 1]}, reinterpret_cast<ntd2*>(gd) + c);
        if constexpr (LWS > 0) {
            double2* ld = reinterpret_cast<double2*>(a.lim + size_t(base) * LWS);
            constexpr uint32_t PR = LWS / 2;
            for (uint32_t c = tid; c < PR * n; c += BLK) __builtin_nontemporal_store((ntd2){sl[c / PR][2 * (c % PR)], sl[c / PR][2 * (c % PR) + 1]}, reinterpret_cast<ntd2*>(ld) + c);
        }
    }
    // MODE 1 (unlimited stage): every limiter is 1; the face kernels take LP = 2 and read none
}

template <int S, int MODE, int MIX>
__global__ __launch_bounds__(BLK) void k_gradient(DevMesh m, PhysParams p, StageArgs a)
{
    static_assert(BLK * 9 >= 2 * FORCE_LDS, "force staging exceeds the gradient staging rows");
    __shared__ GradLds<MODE> L;
    // with a force boundary, workgroup 0 (dispatched first) sums a pending boundary force
    // of the previous step, whether or not the loop is still active; the others are the
    // cell blocks
    const uint32_t fw = a.force.n ? 1u : 0u;
    if (fw && blockIdx.x == 0) {
        pendingForce<BLK, FORCE_LDS>(a.force, a.sc, reinterpret_cast<double2*>(&L.sg[0][0]));
        return;
    }
    const uint32_t b = blockIdx.x - fw, nb = gridDim.x - fw;
    gradientBody<S, MODE, MIX>(m, p, a, (a.order & 2) ? xcdBlockBack(b, nb) : xcdBlock(b, nb), L);
}

// ---------------------------------------------------------------- faces
// LP 0: scalar limiter (4 per cell); 1: LP vector limiter (x,y pairs); 2: an
// unlimited stage, every limiter 1 (1*g is exactly g, so the products round as with
// stored ones)
template <int LP>
DEV void loadLim(const double* lim, uint32_t c, double lx[4], double ly[4])
{
    if (LP == 2) {
#pragma unroll
        for (int k = 0; k < 4; k++) { lx[k] = 1.0; ly[k] = 1.0; }
    } else if (LP) {
        const double4 a = *reinterpret_cast<const double4*>(lim + size_t(c) * 8);
        const double4 b = *reinterpret_cast<const double4*>(lim + size_t(c) * 8 + 4);
        lx[0] = a.x; ly[0] = a.y; lx[1] = a.z; ly[1] = a.w; lx[2] = b.x; ly[2] = b.y; lx[3] = b.z; ly[3] = b.w;
    } else {
        const double4 a = *reinterpret_cast<const double4*>(lim + size_t(c) * 4);
        lx[0] = ly[0] = a.x; lx[1] = ly[1] = a.y; lx[2] = ly[2] = a.z; lx[3] = ly[3] = a.w;
    }
}
// FP32 rows of the mixed path (no LP vector limiter there)
template <int LP>
DEV void loadLim(const float* lim, uint32_t c, float lx[4], float ly[4])
{
    static_assert(LP != 1, "the LP vector limiter runs in FP64 only");
    if (LP == 2) {
#pragma unroll
        for (int k = 0; k < 4; k++) { lx[k] = 1.0f; ly[k] = 1.0f; }
    } else {
        const float4 a = *reinterpret_cast<const float4*>(lim + size_t(c) * 4);
        lx[0] = ly[0] = a.x; lx[1] = ly[1] = a.y; lx[2] = ly[2] = a.z; lx[3] = ly[3] = a.w;
    }
}

DEV void loadGrad(const double* grad, uint32_t c, double g[8])
{
    const double4 a = *reinterpret_cast<const double4*>(grad + size_t(c) * 8);
    const double4 b = *reinterpret_cast<const double4*>(grad + size_t(c) * 8 + 4);
    g[0] = a.x; g[1] = a.y; g[2] = a.z; g[3] = a.w; g[4] = b.x; g[5] = b.y; g[6] = b.z; g[7] = b.w;
}
DEV void loadGrad(const float* grad, uint32_t c, float g[8])
{
    const float4 a = *reinterpret_cast<const float4*>(grad + size_t(c) * 8);
    const float4 b = *reinterpret_cast<const float4*>(grad + size_t(c) * 8 + 4);
    g[0] = a.x; g[1] = a.y; g[2] = a.z; g[3] = a.w; g[4] = b.x; g[5] = b.y; g[6] = b.z; g[7] = b.w;
}

// solve.d:980-1010 (interior) reconstruction with the p<0 fallback
template <int LP, typename R>
DEV void reconstruct(const R qm[4], const R g[8], const R lx[4], const R ly[4], R dx, R dy, R out[4])
{
#pragma unroll
    for (int j = 0; j < 4; j++) out[j] = qm[j] + lx[j] * g[2 * j] * dx + ly[j] * g[2 * j + 1] * dy;
    if (pressure<R>(out) < 0) {
        R l0 = R(1.0), l1 = R(1.0);
#pragma unroll
        for (int j = 0; j < 4; j++) { l0 = fmin(l0, lx[j]); l1 = fmin(l1, ly[j]); }
#pragma unroll
        for (int j = 0; j < 4; j++) out[j] = qm[j] + l0 * g[2 * j] * dx + l1 * g[2 * j + 1] * dy;
    }
}

// solve.d:970-1033: one interior face from both cells' raw data.  Centroids, midpoint
// and normal are FP64 geometry; the offsets and the normal enter the R arithmetic.
template <int FLUX, int VISC, int ORDER, int LP, typename R>
DEV void interiorFace(const R qa[4], const R ga[8], const R lxa[4], const R lya[4], double2 ca,
                      const R qb[4], const R gb[8], const R lxb[4], const R lyb[4], double2 cb,
                      double2 n, double2 mid, const PhysParams& p, R qL[4], R qR[4], R F[4], R& sMax)
{
    if (ORDER > 1) {
        reconstruct<LP, R>(qa, ga, lxa, lya, R(mid.x - ca.x), R(mid.y - ca.y), qL);
        reconstruct<LP, R>(qb, gb, lxb, lyb, R(mid.x - cb.x), R(mid.y - cb.y), qR);
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) { qL[k] = qa[k]; qR[k] = qb[k]; }
    }
    // the face gradient is formed before the convective flux, so the two cell gradients
    // are not live across it
    R dq[8];
    if (VISC) {
#pragma unroll
        for (int k = 0; k < 8; k++) dq[k] = R(0.5) * (ga[k] + gb[k]);
    }
    numFlux<FLUX, R>(qL, qR, R(n.x), R(n.y), F, sMax);
    if (VISC) {
        R qa2[4], Fv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) qa2[k] = R(0.5) * (qL[k] + qR[k]);
        viscousFlux<R>(p.Pr, p.mu, qa2, dq, R(n.x), R(n.y), Fv);
#pragma unroll
        for (int k = 0; k < 4; k++) F[k] = F[k] - Fv[k];
    }
}

// solve.d:283-319 + 769-968: one boundary face from its owner's data and ghost state
template <int FLUX, int VISC, int ORDER, int LP, typename R>
DEV void boundaryFace(const R qo[4], const R gO[8], const R lx[4], const R ly[4], double2 co,
                      const R qg[4], const BFace& bf, double2 nd, double2 mid, const PhysParams& p,
                      R qL[4], R qR[4], R F[4], R& sMax)
{
    const R g = R(G_), gm1 = R(G_ - 1), half = R(0.5);
    const R nx = R(nd.x), ny = R(nd.y);
    // updateBoundaryEdge: the fallback uses lim[0] for both directions
    if (ORDER > 1) {
        const R dx = R(mid.x - co.x), dy = R(mid.y - co.y);
#pragma unroll
        for (int j = 0; j < 4; j++) qL[j] = qo[j] + lx[j] * gO[2 * j] * dx + ly[j] * gO[2 * j + 1] * dy;
        if (pressure<R>(qL) < 0) {
            R l0 = R(1.0);
#pragma unroll
            for (int j = 0; j < 4; j++) l0 = fmin(l0, lx[j]);
#pragma unroll
            for (int j = 0; j < 4; j++) qL[j] = qo[j] + l0 * gO[2 * j] * dx + l0 * gO[2 * j + 1] * dy;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) qL[j] = qo[j];
    }
    const int type = bf.type;
    if (type == BT_FULL || type == BT_DIRICHLET) {
#pragma unroll
        for (int k = 0; k < 4; k++) qR[k] = R(bf.far[k]);
        numFlux<FLUX, R>(qL, qR, nx, ny, F, sMax);
        if (VISC) {
            R qa[4], dq[8], Fv[4];
#pragma unroll
            for (int k = 0; k < 4; k++) qa[k] = half * (qL[k] + qR[k]);
            if (type == BT_DIRICHLET) {
#pragma unroll
                for (int k = 0; k < 8; k++) dq[k] = half * (gO[k] + R(bf.gg[k]));
            } else {
#pragma unroll
                for (int k = 0; k < 8; k++) dq[k] = half * (gO[k] + R(0.0));
            }
            viscousFlux<R>(p.Pr, p.mu, qa, dq, nx, ny, Fv);
#pragma unroll
            for (int k = 0; k < 4; k++) F[k] = F[k] - Fv[k];
        }
    } else if (type == BT_INVWALL) {
        const R inv = R(1) / qL[0];
        const R v0 = inv * qL[1], v1 = inv * qL[2];
        const R vn = v0 * nx + v1 * ny;
        const R t0 = v0 - vn * nx, t1 = v1 - vn * ny;
        const R vel = sqrt(t0 * t0 + t1 * t1);
        R pw = gm1 * (qL[3] - half * qL[0] * (vel * vel));
        const R aw = sqrt(g * (pw / qL[0]));
        if (pw < 0) pw = R(1.0e-12);
        F[0] = R(0); F[1] = pw * nx; F[2] = pw * ny; F[3] = R(0);
        sMax = fabs(aw);
        for (int k = 0; k < 4; k++) qR[k] = qg[k];
    } else if (type == BT_VISCWALL) {
        const R vel = R(0.0);
        R pw = gm1 * (qL[3] - half * qL[0] * (vel * vel));
        const R aw = sqrt(g * (pw / qL[0]));
        if (pw < 0) pw = R(1.0e-12);
        qR[0] = qL[0]; qR[1] = qL[1] * R(-1); qR[2] = qL[2] * R(-1); qR[3] = qL[3];
        R qa[4], dq[8], Fv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) qa[k] = half * (qL[k] + qR[k]);
        qa[1] = R(0.0); qa[2] = R(0.0);
#pragma unroll
        for (int k = 0; k < 8; k++) dq[k] = gO[k];
        dq[6] = (qa[3] / qa[0]) * dq[0];
        dq[7] = (qa[3] / qa[0]) * dq[1];
        viscousFlux<R>(p.Pr, p.mu, qa, dq, nx, ny, Fv);
        F[0] = R(0) - Fv[0]; F[1] = pw * nx - Fv[1]; F[2] = pw * ny - Fv[2]; F[3] = R(0) - Fv[3];
        sMax = fabs(aw);
    } else if (type == BT_CONSTP) {
        for (int k = 0; k < 4; k++) qR[k] = qg[k];
        const R pg = pressure<R>(qR);
        const R ug = qR[1] / qR[0], vg = qR[2] / qR[0];
        const R ag = sqrt(g * (pg / qR[0]));
        convective<R>(pg, ug, vg, qR[0], qR[3], nx, ny, F);
        sMax = fabs(ag);
    } else if (type == BT_SYM) {
        const R v0 = qL[1] / qL[0], v1 = qL[2] / qL[0];
        R w0, w1;
        if (!bf.vertical) {
            w0 = R(0.0) + R(bf.R[0]) * v0 + R(bf.R[1]) * v1;
            w1 = R(0.0) + R(bf.R[2]) * v0 + R(bf.R[3]) * v1;
        } else {
            w0 = -v0;
            w1 = v1;
        }
        qR[0] = qL[0]; qR[1] = qL[0] * w0; qR[2] = qL[0] * w1; qR[3] = qL[3];
        numFlux<FLUX, R>(qL, qR, nx, ny, F, sMax);
        if (VISC) {
            const R one = R(1), zero = R(0), two = R(2);
            R qa[4], bg[8], Fv[4];
#pragma unroll
            for (int k = 0; k < 4; k++) qa[k] = half * (qL[k] + qR[k]);
            const R o00 = nx * nx, o01 = nx * ny, o10 = ny * nx, o11 = ny * ny;
            const R V00 = one - o00, V01 = zero - o01, V10 = zero - o10, V11 = one - o11;
            const R T00 = one - two * o00, T01 = zero - two * o01, T10 = zero - two * o10, T11 = one - two * o11;
            // tmp1 = A*gradp with A = identity except the momentum block V (4x4 * 4x2, all columns)
            R t1[8];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                t1[0 + c] = zero + one * gO[0 + c] + zero * gO[2 + c] + zero * gO[4 + c] + zero * gO[6 + c];
                t1[2 + c] = zero + zero * gO[0 + c] + V00 * gO[2 + c] + V01 * gO[4 + c] + zero * gO[6 + c];
                t1[4 + c] = zero + zero * gO[0 + c] + V10 * gO[2 + c] + V11 * gO[4 + c] + zero * gO[6 + c];
                t1[6 + c] = zero + zero * gO[0 + c] + zero * gO[2 + c] + zero * gO[4 + c] + one * gO[6 + c];
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const R s30 = zero + t1[2 * r] * T00 + t1[2 * r + 1] * T10;
                const R s31 = zero + t1[2 * r] * T01 + t1[2 * r + 1] * T11;
                const R s10 = zero + gO[2 * r] * o00 + gO[2 * r + 1] * o10;
                const R s11 = zero + gO[2 * r] * o01 + gO[2 * r + 1] * o11;
                bg[2 * r] = s10 + s30;
                bg[2 * r + 1] = s11 + s31;
            }
            viscousFlux<R>(p.Pr, p.mu, qa, bg, nx, ny, Fv);
#pragma unroll
            for (int k = 0; k < 4; k++) F[k] = F[k] - Fv[k];
        }
    } else {  // TempPresInflow
        for (int k = 0; k < 4; k++) qR[k] = qg[k];
        numFlux<FLUX, R>(qL, qR, nx, ny, F, sMax);
        if (VISC) {
            R qa[4], Fv[4];
#pragma unroll
            for (int k = 0; k < 4; k++) qa[k] = half * (qL[k] + qR[k]);
            viscousFlux<R>(p.Pr, p.mu, qa, gO, nx, ny, Fv);
#pragma unroll
            for (int k = 0; k < 4; k++) F[k] = F[k] - Fv[k];
        }
    }
}

// Cell data a face reads: the FP64 rows, or on the mixed path (MIX 1) the FP32
// shadow of the state and the FP32 gradient and limiter rows.
template <int MIX>
struct FaceSrc {
    using R = typename std::conditional<MIX == 1, float, double>::type;
    const StageArgs& a;
    DEV void q(uint32_t c, R o[4]) const
    {
        if constexpr (MIX == 1) ld4(a.qin32, c, o);
        else ld4(a.qin, c, o);
    }
    DEV void grad(uint32_t c, R o[8]) const
    {
        if constexpr (MIX == 1) loadGrad(a.grad32, c, o);
        else loadGrad(a.grad, c, o);
    }
    template <int LP>
    DEV void lim(uint32_t c, R lx[4], R ly[4]) const
    {
        if constexpr (MIX == 1) loadLim<LP>(a.lim32, c, lx, ly);
        else loadLim<LP>(a.lim, c, lx, ly);
    }
};

template <int FLUX, int VISC, int ORDER, int LP, int STORE, int MIX>
__global__ __launch_bounds__(BLK) void k_faces(DevMesh m, PhysParams p, StageArgs a, uint32_t nbi)
{
    using R = typename std::conditional<MIX == 1, float, double>::type;
    if (a.checkActive && !a.sc->active) return;
    const FaceSrc<MIX> src{a};
    R qL[4], qR[4], F[4], sMax;
    uint32_t f;
    double len;
    uint64_t errKey;
    if (blockIdx.x < nbi) {
        f = xcdBlock(blockIdx.x, nbi) * BLK + threadIdx.x;
        if (f >= m.Fi) return;
        const uint2 c = m.fcell[f];
        len = m.flen[f];
        R qa[4], qb[4], ga[8], gb[8], lxa[4], lya[4], lxb[4], lyb[4];
        src.q(c.x, qa);
        src.q(c.y, qb);
        if (ORDER > 1 || VISC) { src.grad(c.x, ga); src.grad(c.y, gb); }
        double2 ca = make_double2(0, 0), cb = make_double2(0, 0), mid = make_double2(0, 0);
        if (ORDER > 1) {
            src.template lim<LP>(c.x, lxa, lya);
            src.template lim<LP>(c.y, lxb, lyb);
            ca = m.cen[c.x]; cb = m.cen[c.y]; mid = m.fmid[f];
        }
        interiorFace<FLUX, VISC, ORDER, LP, R>(qa, ga, lxa, lya, ca, qb, gb, lxb, lyb, cb, m.fn[f], mid, p, qL, qR, F,
                                              sMax);
        errKey = uint64_t(m.Fb) + f;   // boundary edges are checked first in the reference
    } else {
        const uint32_t b = (blockIdx.x - nbi) * BLK + threadIdx.x;
        if (b >= m.Fb) return;
        f = m.Fi + b;
        const uint32_t o = m.fcell[f].x;
        len = m.flen[f];
        R qo[4], qg[4], gO[8], lx[4], ly[4];
        src.q(o, qo);
        src.q(m.N + b, qg);
        if (ORDER > 1 || VISC) src.grad(o, gO);
        double2 co = make_double2(0, 0), mid = make_double2(0, 0);
        if (ORDER > 1) {
            src.template lim<LP>(o, lx, ly);
            co = m.cen[o]; mid = m.fmid[f];
        }
        boundaryFace<FLUX, VISC, ORDER, LP, R>(qo, gO, lx, ly, co, qg, m.bface[b], m.fn[f], mid, p, qL, qR, F, sMax);
        errKey = b;
    }
    if (a.recordErrors && (isnan(F[0]) || isnan(F[1]) || isnan(F[2]) || isnan(F[3])))
        atomicMin(&a.sc->errEdge, (uint64_t(a.stageId) << 32) | errKey);
    *reinterpret_cast<double4*>(a.lenF + size_t(f) * 4) =
        make_double4(len * double(F[0]), len * double(F[1]), len * double(F[2]), len * double(F[3]));
    a.lenS[f] = len * double(sMax);
    if (STORE) {
        a.estate[2 * f] = make_double4(qL[0], qL[1], qL[2], qL[3]);
        a.estate[2 * f + 1] = make_double4(qR[0], qR[1], qR[2], qR[3]);
    }
}

// ---------------------------------------------------------------- residual + update
DEV double warpMin(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
DEV double warpMax(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// solve.d:1092-1117 + the integrator's stage update for interior cell i, from
// its summed face fluxes r and sum of len*sMax
// met: (perimeter, d, -1/area) of the cell, or null to read them from the mesh
// qNew/qOld (optional): the new state and q^n when this stage completes a step
// qnI, accI (optional): the cell's q^n and RK accumulator rows already loaded
// the stage output row, and on the mixed path its FP32 shadow for the next stage's gathers
DEV void storeOut(const StageArgs& a, uint32_t i, const double o[4])
{
    st4(a.qout, i, o);
    if (a.qout32) a.qout32[i] = make_float4(float(o[0]), float(o[1]), float(o[2]), float(o[3]));
}

template <int KIND>
DEV void cellStage(const DevMesh& m, const PhysParams& p, const StageArgs& a, uint32_t i, const double r[4],
                   double sAve, const double* qinI, const double* met, double& dtCand, double& rCand,
                   double* qNew = nullptr, double* qOld = nullptr, const double* qnI = nullptr,
                   const double* accI = nullptr)
{
    auto loadAcc = [&](double v[4]) {
        if (accI) { v[0] = accI[0]; v[1] = accI[1]; v[2] = accI[2]; v[3] = accI[3]; }
        else ld4(a.acc, i, v);
    };
    const double cper = met ? met[0] : m.cper[i], cd = met ? met[1] : m.cd[i], nia = met ? met[2] : m.nia[i];
    sAve /= cper;
    const double CFL = a.sc->CFL;
    const double dtc = (CFL * cd) / sAve;
    if (!isnan(dtc)) dtCand = fmin(dtCand, dtc);
    double hc = 0.0;
    if (p.localTimestep) {
        if (a.dtUpdate) {
            hc = CFL * cd / sAve;
            if (isnan(hc)) hc = 0;
            a.celldt[i] = hc;
        } else {
            hc = a.celldt[i];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) if (fabs(r[k]) > rCand) rCand = fabs(r[k]);
    const double s = nia;
    double R[4];
#pragma unroll
    for (int k = 0; k < 4; k++) R[k] = r[k] * s;
    if (m.src) {
        const double4 sv = m.src[i];
        R[0] = R[0] + sv.x; R[1] = R[1] + sv.y; R[2] = R[2] + sv.z; R[3] = R[3] + sv.w;
    }
    const double dt = a.sc->dt;
    const double h = p.localTimestep ? hc : dt;
    double qn[4], o[4];
    if (KIND != SK_RESIDUAL) {
        if (qnI) { qn[0] = qnI[0]; qn[1] = qnI[1]; qn[2] = qnI[2]; qn[3] = qnI[3]; }
        else ld4(a.qn, i, qn);
    }
    if (KIND == SK_RESIDUAL) {
        st4(a.R, i, R);
    } else if (KIND == SK_EULER) {
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = qn[k] + h * R[k];
        storeOut(a, i, o);
    } else if (KIND == SK_RK2_1) {
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = qn[k] + (dt * R[k]);
        st4(a.acc, i, R);
        storeOut(a, i, o);
    } else if (KIND == SK_RK2_2) {
        double k1[4];
        loadAcc(k1);
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = qn[k] + (h / 2.0) * (k1[k] + R[k]);
        storeOut(a, i, o);
    } else if (KIND == SK_TVD_1) {
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = qn[k] + h * R[k];
        storeOut(a, i, o);
    } else if (KIND == SK_TVD_2) {
        double fe[4];
        if (qinI) { fe[0] = qinI[0]; fe[1] = qinI[1]; fe[2] = qinI[2]; fe[3] = qinI[3]; }
        else ld4(a.qin, i, fe);
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = 0.5 * (qn[k] + fe[k] + h * R[k]);
        storeOut(a, i, o);
    } else if (KIND == SK_RK4_1) {
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = qn[k] + ((dt / 2.0) * R[k]);
        st4(a.acc, i, R);
        storeOut(a, i, o);
    } else if (KIND == SK_RK4_2 || KIND == SK_RK4_3) {
        double ac[4];
        loadAcc(ac);
#pragma unroll
        for (int k = 0; k < 4; k++) ac[k] = ac[k] + 2.0 * R[k];
        const double c = (KIND == SK_RK4_2) ? (dt / 2.0) : dt;
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = qn[k] + (c * R[k]);
        st4(a.acc, i, ac);
        storeOut(a, i, o);
    } else if (KIND == SK_RK4_4) {
        double ac[4];
        loadAcc(ac);
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = qn[k] + (h / 6.0) * (ac[k] + R[k]);
        storeOut(a, i, o);
    }
    if (KIND != SK_RESIDUAL && a.last && a.recordErrors &&
        (isnan(o[0]) || isnan(o[1]) || isnan(o[2]) || isnan(o[3])))
        atomicMin(&a.sc->errCell, (uint64_t(a.stepId) << 32) | m.refId[i]);
    if ((KIND == SK_EULER || KIND == SK_RK2_2 || KIND == SK_TVD_2 || KIND == SK_RK4_4) && qNew) {
#pragma unroll
        for (int k = 0; k < 4; k++) { qNew[k] = o[k]; qOld[k] = qn[k]; }
    }
}

// Leaf sums of the RMS residual's pairwise sum (app/solver.d:153-168), fused into
// the stage that completes a step.  Each group of 16 lanes holds one aligned block
// of 16 consecutive reference cells, lane j its element j (absent cells add
// zeros); the xor butterfly adds exactly like the fixed 16-element tree, and lane
// 0 of the group stores the block's 8 sums: (new - old)^2 and new per variable.
// Every lane of the wave must call this.
DEV void rmsBlockLeaf(bool have, const double qNew[4], const double qOld[4], uint32_t refCell, double* rmsBlocks)
{
    double o[8];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const double x = have ? qNew[v] - qOld[v] : 0.0;
        o[v] = have ? x * x : 0.0;
        o[4 + v] = have ? qNew[v] : 0.0;
    }
#pragma unroll
    for (int mm = 1; mm < 16; mm <<= 1)
#pragma unroll
        for (int q = 0; q < 8; q++) o[q] = o[q] + __shfl_xor(o[q], mm, 64);
    if (have && (threadIdx.x & 15u) == 0) {
        double* dst = rmsBlocks + size_t(refCell / 16) * 8;
        *reinterpret_cast<double4*>(dst) = make_double4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<double4*>(dst + 4) = make_double4(o[4], o[5], o[6], o[7]);
    }
}

// per-workgroup min dt / max |R| partial, combined later by k_reduce_stage (device-scope
// atomics from every workgroup on one address serialise across the XCDs)
DEV void blockPartial(double dtCand, double rCand, double* partial, uint32_t slot)
{
    __shared__ double sMin[BLK / 64], sMax_[BLK / 64];
    dtCand = warpMin(dtCand);
    rCand = warpMax(rCand);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) { sMin[w] = dtCand; sMax_[w] = rCand; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a0 = sMin[0], b0 = sMax_[0];
#pragma unroll
        for (int k = 1; k < BLK / 64; k++) { a0 = fmin(a0, sMin[k]); b0 = fmax(b0, sMax_[k]); }
        partial[2 * slot] = a0;
        partial[2 * slot + 1] = b0;
    }
}

template <int S, int KIND>
__global__ __launch_bounds__(BLK) void k_update(DevMesh m, PhysParams p, StageArgs a)
{
    if (a.checkActive && !a.sc->active) return;
    const uint32_t i = xcdBlock(blockIdx.x, gridDim.x) * BLK + threadIdx.x;
    double dtCand = INFINITY, rCand = 0.0;
    double qNew[4] = {0, 0, 0, 0}, qOld[4] = {0, 0, 0, 0};
    if (i < m.N) {
        double r[4] = {0.0, 0.0, 0.0, 0.0};
        double sAve = 0;
#pragma unroll
        for (int j = 0; j < S; j++) {
            const uint32_t fe = m.cface[i * S + j];
            const uint32_t f = fe & 0x7fffffffu;
            const double4 lf = *reinterpret_cast<const double4*>(a.lenF + size_t(f) * 4);
            if (fe >> 31) { r[0] = r[0] - lf.x; r[1] = r[1] - lf.y; r[2] = r[2] - lf.z; r[3] = r[3] - lf.w; }
            else { r[0] = r[0] + lf.x; r[1] = r[1] + lf.y; r[2] = r[2] + lf.z; r[3] = r[3] + lf.w; }
            sAve += a.lenS[f];
        }
        cellStage<KIND>(m, p, a, i, r, sAve, nullptr, nullptr, dtCand, rCand, qNew, qOld);
    }
    if ((KIND == SK_EULER || KIND == SK_RK2_2 || KIND == SK_TVD_2 || KIND == SK_RK4_4) && a.rmsBlocks)
        rmsBlockLeaf(i < m.N, qNew, qOld, i < m.N ? m.refId[i] : 0u, a.rmsBlocks);
    blockPartial(dtCand, rCand, a.partial, blockIdx.x);
}

// ---------------------------------------------------------------- fused stage
// One workgroup per tile: load the states of own + 2 halo rings into LDS, fill
// the ghosts, compute gradients/limiters for own + ring 1, every face flux of the
// own cells, and the own cells' residual + stage update -- one pass over HBM.
// The per-cell and per-face arithmetic is the same device code as the split
// kernels above, so results are bit-identical.
constexpr int TBLK = 256;
__global__ void k_reduce_stage(const double* __restrict__ partial, uint32_t n, DevScalars* sc, int checkActive);

// Workgroup barrier for data shared through LDS: waits for this wave's LDS
// operations only.  __syncthreads() (and an LDS-scope fence, which covers the
// direct global-to-LDS copies) also waits for every outstanding global access,
// which would stall on the next tile's prefetch.  The "memory" clobber keeps the
// compiler from moving memory accesses across it.
DEV void ldsBarrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Persistent form: a workgroup walks a contiguous run of tiles.  Two LDS buffers
// alternate: while the workgroup computes one tile, the next tile's blob is
// copied into the other buffer by direct global-to-LDS loads.
typedef __attribute__((address_space(3))) void* LdsPtr;

DEV void dmaBlob(const TileArgs& t, const uint4& d1, uint8_t* slot)
{
    const uint32_t n = d1.z, lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint4* src = t.blob + d1.y;
    for (uint32_t g = w; g * 64 < n; g += TBLK / 64) {
        const uint32_t c = g * 64 + lane;
        if (c < n) __builtin_amdgcn_global_load_lds((const void*)(src + c), (LdsPtr)(slot + g * 1024), 16, 0, 0);
    }
}

template <int S, int FLUX, int VISC, int LP>
__global__ __launch_bounds__(TBLK, 2) void k_stage_tile(DevMesh m, PhysParams p, StageArgs a, TileArgs t)
{
    if (a.checkActive && !a.sc->active) return;
    extern __shared__ uint4 ldsv[];
    uint8_t* lds = reinterpret_cast<uint8_t*>(ldsv);
    uint8_t* dyn = lds + 2 * t.slotBytes;
    const uint32_t bid = xcdBlock(blockIdx.x, gridDim.x);
    if (bid == 0 && threadIdx.x == 0 && a.first) a.sc->rmaxBits = 0;   // Rmax is per step
    const uint32_t per = t.nTiles / gridDim.x, rem = t.nTiles % gridDim.x;
    const uint32_t tBeg = bid * per + (bid < rem ? bid : rem), tEnd = tBeg + per + (bid < rem ? 1 : 0);
    constexpr int LW = LP ? 8 : 4;
    const uint32_t tid = threadIdx.x;
    double dtCand = INFINITY, rCand = 0.0;
    uint4 d0 = make_uint4(0, 0, 0, 0), d1 = make_uint4(0, 0, 0, 0);
    if (tBeg < tEnd) {
        d0 = t.desc[2 * tBeg];
        d1 = t.desc[2 * tBeg + 1];
        dmaBlob(t, d1, lds);
    }
    uint32_t par = 0;
    for (uint32_t tile = tBeg; tile < tEnd; tile++) {
        const uint32_t nC = d0.x, nG = d0.y, nQ = d0.z, nL = d0.w, nF = d1.x;
        const TileImg I = tileImage(nC, nG, nQ, nL, nF, S);
        const TileDyn D = tileDyn(nG, nL, LW);
        uint8_t* base = lds + par * t.slotBytes;
        __builtin_amdgcn_s_waitcnt(0);   // this wave's copies of the blob have landed
        ldsBarrier();
        const uint32_t* ids = reinterpret_cast<const uint32_t*>(base + I.ids);
        double* qs = reinterpret_cast<double*>(dyn + D.qs);
        for (uint32_t k = tid; k < nQ; k += TBLK) {
            const double4 v = a.qin[ids[k]];
            qs[4 * k] = v.x; qs[4 * k + 1] = v.y; qs[4 * k + 2] = v.z; qs[4 * k + 3] = v.w;
        }
        const uint4 n0 = reinterpret_cast<const uint4*>(base + I.hdr)[0];
        const uint4 n1 = reinterpret_cast<const uint4*>(base + I.hdr)[1];
        if (tile + 1 < tEnd) dmaBlob(t, n1, lds + (par ^ 1u) * t.slotBytes);   // overlaps the compute below
        ldsBarrier();
        if (nL > nQ) {
            const uint16_t* gown = reinterpret_cast<const uint16_t*>(base + I.gown);
            for (uint32_t k = nQ + tid; k < nL; k += TBLK) {
                const uint32_t g = ids[k], own = gown[k - nQ];
                double qg[4];
                if (!ghostFromOwner(m.bface[g - m.N], qs + 4 * own, a.qin, g, qg) && a.recordErrors)
                    atomicMin(&a.sc->errBC, (uint64_t(a.stageId) << 32) | (g - m.N));
#pragma unroll
                for (int c = 0; c < 4; c++) qs[4 * k + c] = qg[c];
                if (own < nC) st4(a.qin, g, qg);   // one writer per ghost: the tile owning its cell
            }
            ldsBarrier();
        }
        const double2* cs = reinterpret_cast<const double2*>(base + I.cs);
        double* gs = reinterpret_cast<double*>(dyn + D.gs);
        double* ls = reinterpret_cast<double*>(dyn + D.ls);
        if (p.order > 1) {
            const bool lim = p.limited && a.limit;
            const uint16_t* nbS = reinterpret_cast<const uint16_t*>(base + I.nb);
            const double* gmS = reinterpret_cast<const double*>(base + I.gm);
            for (uint32_t k = tid; k < nG; k += TBLK) {
                double qn[S][4];
                double2 cn[S];
                bool real[S];
#pragma unroll
                for (int j = 0; j < S; j++) {
                    const uint32_t nb = nbS[k * S + j];
#pragma unroll
                    for (int c = 0; c < 4; c++) qn[j][c] = qs[4 * nb + c];
                    cn[j] = cs[nb];
                    real[j] = nb != k;
                }
                double g[4][2], lx[4], ly[4];
                const double* gmi = gmS + k;
                if (!lim) gradLimit<S, 1, double>(qs + 4 * k, qn, gmi, nG, cs[k], cn, real, 0, p.lpThresh, g, lx, ly);
                else if (LP) gradLimit<S, 3, double>(qs + 4 * k, qn, gmi, nG, cs[k], cn, real, 1, p.lpThresh, g, lx, ly);
                else gradLimit<S, 2, double>(qs + 4 * k, qn, gmi, nG, cs[k], cn, real, 1, p.lpThresh, g, lx, ly);
#pragma unroll
                for (int c = 0; c < 4; c++) { gs[8 * k + 2 * c] = g[c][0]; gs[8 * k + 2 * c + 1] = g[c][1]; }
                if (LP) {
#pragma unroll
                    for (int c = 0; c < 4; c++) { ls[8 * k + 2 * c] = lx[c]; ls[8 * k + 2 * c + 1] = ly[c]; }
                } else {
#pragma unroll
                    for (int c = 0; c < 4; c++) ls[4 * k + c] = lx[c];
                }
            }
        } else if (VISC) {
            for (uint32_t k = tid; k < 8 * nG; k += TBLK) gs[k] = 0.0;   // order 1: zero gradients
        }
        ldsBarrier();
        double* fs = reinterpret_cast<double*>(base + I.fs);
        const uint2* fcS = reinterpret_cast<const uint2*>(base + I.fc);
        for (uint32_t k = tid; k < nF; k += TBLK) {
            const uint2 fe = fcS[k];
            const uint32_t f = fe.x, la = fe.y & 0xffffu, lb = fe.y >> 16;
            double* slot = fs + 5 * k;
            const double2 n = make_double2(slot[0], slot[1]);
            const double2 mid = make_double2(slot[2], slot[3]);
            const double len = slot[4];
            double qL[4], qR[4], F[4], sMax;
            double la_x[4], la_y[4], lb_x[4], lb_y[4];
            if (p.order > 1) {
                if (LP) {
#pragma unroll
                    for (int c = 0; c < 4; c++) { la_x[c] = ls[8 * la + 2 * c]; la_y[c] = ls[8 * la + 2 * c + 1]; }
                } else {
#pragma unroll
                    for (int c = 0; c < 4; c++) { la_x[c] = ls[4 * la + c]; la_y[c] = la_x[c]; }
                }
            }
            uint64_t errKey;
            if (f < m.Fi) {
                if (p.order > 1) {
                    if (LP) {
#pragma unroll
                        for (int c = 0; c < 4; c++) { lb_x[c] = ls[8 * lb + 2 * c]; lb_y[c] = ls[8 * lb + 2 * c + 1]; }
                    } else {
#pragma unroll
                        for (int c = 0; c < 4; c++) { lb_x[c] = ls[4 * lb + c]; lb_y[c] = lb_x[c]; }
                    }
                }
                const double2 ca = cs[la], cb = cs[lb];
                if (p.order > 1)
                    interiorFace<FLUX, VISC, 2, LP>(qs + 4 * la, gs + 8 * la, la_x, la_y, ca, qs + 4 * lb, gs + 8 * lb,
                                                    lb_x, lb_y, cb, n, mid, p, qL, qR, F, sMax);
                else
                    interiorFace<FLUX, VISC, 1, LP>(qs + 4 * la, gs + 8 * la, la_x, la_y, ca, qs + 4 * lb, gs + 8 * lb,
                                                    lb_x, lb_y, cb, n, mid, p, qL, qR, F, sMax);
                errKey = uint64_t(m.Fb) + f;
            } else {
                const uint32_t b = f - m.Fi;
                const double2 co = cs[la];
                if (p.order > 1)
                    boundaryFace<FLUX, VISC, 2, LP>(qs + 4 * la, gs + 8 * la, la_x, la_y, co, qs + 4 * lb, m.bface[b],
                                                    n, mid, p, qL, qR, F, sMax);
                else
                    boundaryFace<FLUX, VISC, 1, LP>(qs + 4 * la, gs + 8 * la, la_x, la_y, co, qs + 4 * lb, m.bface[b],
                                                    n, mid, p, qL, qR, F, sMax);
                errKey = b;
            }
            if (a.recordErrors && (isnan(F[0]) || isnan(F[1]) || isnan(F[2]) || isnan(F[3])))
                atomicMin(&a.sc->errEdge, (uint64_t(a.stageId) << 32) | errKey);
            slot[0] = len * F[0]; slot[1] = len * F[1]; slot[2] = len * F[2]; slot[3] = len * F[3];
            slot[4] = len * sMax;
            if (f >= m.Fi) {   // boundary faces (each in exactly one tile) feed the force integral
                *reinterpret_cast<double4*>(a.lenF + size_t(f) * 4) =
                    make_double4(len * F[0], len * F[1], len * F[2], len * F[3]);
                a.lenS[f] = len * sMax;
            }
        }
        ldsBarrier();
        const uint16_t* cfS = reinterpret_cast<const uint16_t*>(base + I.cf);
        const double* met = reinterpret_cast<const double*>(base + I.met);
        for (uint32_t k = tid; k < nC; k += TBLK) {
            double r[4] = {0.0, 0.0, 0.0, 0.0};
            double sAve = 0;
#pragma unroll
            for (int j = 0; j < S; j++) {
                const uint32_t fe = cfS[k * S + j];
                const double* lf = fs + 5 * (fe & 0x7fffu);
                if (fe >> 15) { r[0] = r[0] - lf[0]; r[1] = r[1] - lf[1]; r[2] = r[2] - lf[2]; r[3] = r[3] - lf[3]; }
                else { r[0] = r[0] + lf[0]; r[1] = r[1] + lf[1]; r[2] = r[2] + lf[2]; r[3] = r[3] + lf[3]; }
                sAve += lf[4];
            }
            const uint32_t gi = ids[k];
            const double* qi = qs + 4 * k;
            const double* mk = met + 3 * k;
            switch (a.kind) {
            case SK_EULER: cellStage<SK_EULER>(m, p, a, gi, r, sAve, qi, mk, dtCand, rCand); break;
            case SK_RK2_1: cellStage<SK_RK2_1>(m, p, a, gi, r, sAve, qi, mk, dtCand, rCand); break;
            case SK_RK2_2: cellStage<SK_RK2_2>(m, p, a, gi, r, sAve, qi, mk, dtCand, rCand); break;
            case SK_TVD_1: cellStage<SK_TVD_1>(m, p, a, gi, r, sAve, qi, mk, dtCand, rCand); break;
            case SK_TVD_2: cellStage<SK_TVD_2>(m, p, a, gi, r, sAve, qi, mk, dtCand, rCand); break;
            case SK_RK4_1: cellStage<SK_RK4_1>(m, p, a, gi, r, sAve, qi, mk, dtCand, rCand); break;
            case SK_RK4_2: cellStage<SK_RK4_2>(m, p, a, gi, r, sAve, qi, mk, dtCand, rCand); break;
            case SK_RK4_3: cellStage<SK_RK4_3>(m, p, a, gi, r, sAve, qi, mk, dtCand, rCand); break;
            case SK_RK4_4: cellStage<SK_RK4_4>(m, p, a, gi, r, sAve, qi, mk, dtCand, rCand); break;
            default: cellStage<SK_RESIDUAL>(m, p, a, gi, r, sAve, qi, mk, dtCand, rCand); break;
            }
        }
        ldsBarrier();   // the next tile overwrites the LDS image
        d0 = n0;
        d1 = n1;
        par ^= 1u;
    }
    blockPartial(dtCand, rCand, a.partial, blockIdx.x);
}

// ---------------------------------------------------------------- faces + update
// One workgroup per face tile: every face of the tile's cells (states,
// gradients, limiters and ghosts as the gradient kernel left them) into LDS,
// then the cells' residual and stage update from LDS.  Same device code as
// k_faces and k_update, so bit-identical to them.
// Edge length, unit normal and midpoint from the two nodes, in the operation order
// of the mesh set-up (source/mesh.d; mesh.cpp), so they equal the stored ones bit
// for bit.  Interior edges: normal = (y1-y0, x0-x1)/|.|, whose norm is the length
// (the same squares summed in the other order); boundary edges: (1/len) * (...).
DEV void edgeGeometry(bool interior, double2 p0, double2 p1, double2& n, double2& mid, double& len)
{
    const double ddx = p0.x - p1.x, ddy = p0.y - p1.y;
    len = sqrt(ddx * ddx + ddy * ddy);
    const double vx = p1.y - p0.y, vy = p0.x - p1.x;
    if (interior) {
        n = make_double2(vx / len, vy / len);
    } else {
        const double inv = 1.0 / len;
        n = make_double2(inv * vx, inv * vy);
    }
    mid = make_double2(0.5 * (p0.x + p1.x), 0.5 * (p0.y + p1.y));
}

// A tile's face results in LDS.  FP64 path: len*flux and len*sMax per face (40 B).  Mixed
// path: the FP32 flux and sMax and the FP64 length (28 B per face); the cell phase forms the
// same FP64 products len*F, so its sums are those of the 40-byte layout bit for bit, in
// 30 % less LDS (four workgroups per CU at 768-cell tiles instead of three).  `cap`: face
// slots of the largest tile plus the padding face.
template <int MIX>
struct FaceLds;
template <>
struct FaceLds<0> {
    double* v;
    DEV FaceLds(double* base, uint32_t) : v(base) {}
    DEV void put(uint32_t k, double len, const double F[4], double sMax) const
    {
        double* slot = v + 5 * k;
        slot[0] = len * F[0]; slot[1] = len * F[1]; slot[2] = len * F[2]; slot[3] = len * F[3];
        slot[4] = len * sMax;
    }
    DEV void zero(uint32_t k, uint32_t tid) const { if (tid < 5) v[5 * k + tid] = 0.0; }
    DEV void get(uint32_t k, double o[5]) const
    {
        const double* slot = v + 5 * k;
        o[0] = slot[0]; o[1] = slot[1]; o[2] = slot[2]; o[3] = slot[3]; o[4] = slot[4];
    }
};
template <>
struct FaceLds<1> {
    float4* f;
    double* len;
    float* s;
    DEV FaceLds(double* base, uint32_t cap)
        : f(reinterpret_cast<float4*>(base)), len(reinterpret_cast<double*>(f + cap)), s(reinterpret_cast<float*>(len + cap))
    {
    }
    DEV void put(uint32_t k, double l, const float F[4], float sMax) const
    {
        f[k] = make_float4(F[0], F[1], F[2], F[3]);
        len[k] = l;
        s[k] = sMax;
    }
    DEV void zero(uint32_t k, uint32_t tid) const
    {
        if (tid == 0) { f[k] = make_float4(0.f, 0.f, 0.f, 0.f); len[k] = 0.0; s[k] = 0.f; }
    }
    DEV void get(uint32_t k, double o[5]) const
    {
        const float4 F = f[k];
        const double l = len[k];
        o[0] = l * double(F.x); o[1] = l * double(F.y); o[2] = l * double(F.z); o[3] = l * double(F.w);
        o[4] = l * double(s[k]);
    }
};
// The cell phase of a face tile: each own cell sums its faces' results from LDS in its own
// face order and applies the stage update; the workgroup's (min dt, max |R|) pair goes to
// partial slot `pslot`.  giNext/giNext2: the device ids of the first two cell rounds.
template <int S, int MIX>
DEV void tileCellPhase(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t, const uint4& d,
                       uint32_t pslot, const FaceLds<MIX>& fl, uint32_t giNext, uint32_t giNext2)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t nC = d.x;
    // each cell round's per-cell inputs are loaded one round ahead (the first round's
    // before the barrier, which waits for LDS only), from ids loaded two rounds ahead
    // (a load issued after the next round's would wait for it: loads complete in order)
    double metN[3] = {0, 0, 0}, qnN[4] = {0, 0, 0, 0}, qiN[4] = {0, 0, 0, 0}, acN[4] = {0, 0, 0, 0};
    uint32_t refN = 0;
    const bool wantQn = a.kind != SK_RESIDUAL, wantQi = a.kind == SK_TVD_2;
    const bool wantAcc = a.kind == SK_RK2_2 || a.kind == SK_RK4_2 || a.kind == SK_RK4_3 || a.kind == SK_RK4_4;
    const bool wantRef = a.rmsBlocks != nullptr;
    uint32_t cfN[S] = {};
    auto loadCellIn = [&](uint32_t c) {
        metN[0] = m.cper[c]; metN[1] = m.cd[c]; metN[2] = m.nia[c];
        if (wantQn) ld4(a.qn, c, qnN);
        if (wantQi) ld4(a.qin, c, qiN);
        if (wantAcc) ld4(a.acc, c, acN);
        if (wantRef) refN = m.refId[c];
    };
    auto loadCf = [&](uint32_t k) {
#pragma unroll
        for (int j = 0; j < S; j++) cfN[j] = t.cf[size_t(d.z + k) * S + j];
    };
    if (tid < nC) { loadCellIn(giNext); loadCf(tid); }
    ldsBarrier();
    double dtCand = INFINITY, rCand = 0.0;
    // uniform trip count: the RMS leaf sums need every lane of a 16-lane group
    for (uint32_t k0 = 0; k0 < nC; k0 += BLK) {
        const uint32_t k = k0 + tid;
        const bool have = k < nC;
        double qNew[4] = {0, 0, 0, 0}, qOld[4] = {0, 0, 0, 0};
        uint32_t gi = giNext;
        double met[3], qnC[4], qiC[4], acC[4];
#pragma unroll
        for (int x = 0; x < 3; x++) met[x] = metN[x];
#pragma unroll
        for (int x = 0; x < 4; x++) { qnC[x] = qnN[x]; qiC[x] = qiN[x]; acC[x] = acN[x]; }
        const uint32_t refC = refN;
        uint32_t cfC[S];
#pragma unroll
        for (int j = 0; j < S; j++) cfC[j] = cfN[j];
        if (k + BLK < nC) {
            giNext = giNext2;
            if (k + 2 * BLK < nC) giNext2 = t.ids[d.z + k + 2 * BLK];
            loadCellIn(giNext);
            loadCf(k + BLK);
        }
        if (have) {
            double r[4] = {0.0, 0.0, 0.0, 0.0};
            double sAve = 0;
#pragma unroll
            for (int j = 0; j < S; j++) {
                const uint32_t fe = cfC[j];
                double lf[5];
                fl.get(fe & 0x7fffu, lf);
                if (fe >> 15) { r[0] = r[0] - lf[0]; r[1] = r[1] - lf[1]; r[2] = r[2] - lf[2]; r[3] = r[3] - lf[3]; }
                else { r[0] = r[0] + lf[0]; r[1] = r[1] + lf[1]; r[2] = r[2] + lf[2]; r[3] = r[3] + lf[3]; }
                sAve += lf[4];
            }
            switch (a.kind) {
            case SK_EULER: cellStage<SK_EULER>(m, p, a, gi, r, sAve, qiC, met, dtCand, rCand, qNew, qOld, qnC, acC); break;
            case SK_RK2_1: cellStage<SK_RK2_1>(m, p, a, gi, r, sAve, qiC, met, dtCand, rCand, nullptr, nullptr, qnC, acC); break;
            case SK_RK2_2: cellStage<SK_RK2_2>(m, p, a, gi, r, sAve, qiC, met, dtCand, rCand, qNew, qOld, qnC, acC); break;
            case SK_TVD_1: cellStage<SK_TVD_1>(m, p, a, gi, r, sAve, qiC, met, dtCand, rCand, nullptr, nullptr, qnC, acC); break;
            case SK_TVD_2: cellStage<SK_TVD_2>(m, p, a, gi, r, sAve, qiC, met, dtCand, rCand, qNew, qOld, qnC, acC); break;
            case SK_RK4_1: cellStage<SK_RK4_1>(m, p, a, gi, r, sAve, qiC, met, dtCand, rCand, nullptr, nullptr, qnC, acC); break;
            case SK_RK4_2: cellStage<SK_RK4_2>(m, p, a, gi, r, sAve, qiC, met, dtCand, rCand, nullptr, nullptr, qnC, acC); break;
            case SK_RK4_3: cellStage<SK_RK4_3>(m, p, a, gi, r, sAve, qiC, met, dtCand, rCand, nullptr, nullptr, qnC, acC); break;
            case SK_RK4_4: cellStage<SK_RK4_4>(m, p, a, gi, r, sAve, qiC, met, dtCand, rCand, qNew, qOld, qnC, acC); break;
            default: cellStage<SK_RESIDUAL>(m, p, a, gi, r, sAve, qiC, met, dtCand, rCand, nullptr, nullptr, qnC); break;
            }
        }
        if (a.rmsBlocks) rmsBlockLeaf(have, qNew, qOld, have ? refC : 0u, a.rmsBlocks);
    }
    blockPartial(dtCand, rCand, a.partial, pslot);
}

// One workgroup of the face/update kernel on tile `tile` of its mesh, leaving its (min dt,
// max |R|) pair in partial slot `pslot` (the solo and the batched kernels run this same code).
template <int S, int FLUX, int VISC, int ORDER, int LP, int MIX>
DEV void faceTileBody(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t,
                      uint32_t tile, uint32_t pslot, double* fsd)
{
    using R = typename std::conditional<MIX == 1, float, double>::type;
    const FaceSrc<MIX> src{a};
    const FaceLds<MIX> fl(fsd, t.faceStride + 1);
    const uint32_t tid = threadIdx.x;
    // The first two cell rounds' ids are loaded ahead of the faces; the next round's
    // face record is loaded one round ahead, so each round waits for one memory latency
    // (cell data) instead of two (record, then cell data).  The tile's arrays start at
    // fixed strides, so these first loads are issued beside the descriptor's instead of
    // after it (lanes past the tile's counts read padding and do not use it).
    uint32_t giNext = t.ids[size_t(tile) * t.cellStride + tid];
    uint32_t giNext2 = t.ids[size_t(tile) * t.cellStride + tid + BLK];
    uint4 rNext = t.fref[size_t(tile) * t.faceStride + tid];
    const uint4 d = t.desc[tile];
    // the loop's active flag is read after those loads are issued (all in bounds)
    if (a.checkActive && !a.sc->active) return;
    const uint32_t nF = d.y;
    for (uint32_t k = tid; k < nF; k += BLK) {
        const uint4 r = rNext;   // cell a, cell b, node 0, node 1
        if (k + BLK < nF) rNext = t.fref[d.w + k + BLK];
        const bool interior = r.y < m.N;
        // every load of the face is issued before its arithmetic, with no branch between
        // them, so the round waits for one memory latency.  Side b of a boundary face is
        // its ghost (cell N + b), which has a state and a centroid but no gradient or
        // limiter row: the owner's row is read in its place and not used.
        const double2 p0 = m.node[r.z], p1 = m.node[r.w];
        const uint32_t rb = interior ? r.y : r.x;
        R qa[4], qb[4], ga[8], gb[8], lxa[4], lya[4], lxb[4], lyb[4];
        src.q(r.x, qa);
        src.q(r.y, qb);
        if (ORDER > 1 || VISC) { src.grad(r.x, ga); src.grad(rb, gb); }
        double2 ca = make_double2(0, 0), cb = make_double2(0, 0);
        if (ORDER > 1) {
            src.template lim<LP>(r.x, lxa, lya);
            src.template lim<LP>(rb, lxb, lyb);
            ca = m.cen[r.x]; cb = m.cen[r.y];
        }
        double2 n, mid;
        double len;
        edgeGeometry(interior, p0, p1, n, mid, len);
        R qL[4], qR[4], Fr[4], sMr;
        bool isB = false;
        uint32_t b = 0;
        if (interior) {
            interiorFace<FLUX, VISC, ORDER, LP, R>(qa, ga, lxa, lya, ca, qb, gb, lxb, lyb, cb, n, mid, p, qL, qR, Fr,
                                                  sMr);
        } else {
            isB = true;
            b = r.y - m.N;   // the ghost of boundary face b is cell N + b
            boundaryFace<FLUX, VISC, ORDER, LP, R>(qa, ga, lxa, lya, ca, qb, m.bface[b], n, mid, p, qL, qR, Fr, sMr);
        }
        // the residual is accumulated in FP64 from len * flux on either path
        const double F[4] = {double(Fr[0]), double(Fr[1]), double(Fr[2]), double(Fr[3])}, sMax = double(sMr);
        if (isB) {
            // boundary faces (each in exactly one tile) feed the force integral
            const uint32_t f = m.Fi + b;
            *reinterpret_cast<double4*>(a.lenF + size_t(f) * 4) =
                make_double4(len * F[0], len * F[1], len * F[2], len * F[3]);
            a.lenS[f] = len * sMax;
        }
        // boundary edges are checked first in the reference: their keys come first
        if (a.recordErrors && (isnan(F[0]) || isnan(F[1]) || isnan(F[2]) || isnan(F[3])))
            atomicMin(&a.sc->errEdge,
                      (uint64_t(a.stageId) << 32) | (isB ? uint64_t(b) : uint64_t(m.Fb) + t.fid[d.w + k]));
        if constexpr (MIX == 1) fl.put(k, len, Fr, sMr);
        else fl.put(k, len, F, sMax);
    }
    fl.zero(nF, tid);   // padding face
    tileCellPhase<S, MIX>(m, p, a, t, d, pslot, fl, giNext, giNext2);
}

template <int S, int FLUX, int VISC, int ORDER, int LP, int MIX>
__global__ __launch_bounds__(BLK) void k_face_tile(DevMesh m, PhysParams p, StageArgs a, FaceTileArgs t)
{
    extern __shared__ double fsd[];
    const uint32_t tile = (a.order & 1) ? xcdBlockBack(blockIdx.x, gridDim.x) : xcdBlock(blockIdx.x, gridDim.x);
    faceTileBody<S, FLUX, VISC, ORDER, LP, MIX>(m, p, a, t, tile, blockIdx.x, fsd);
}

template <int S, int FLUX, int VISC>
static void faceTileO(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t, hipStream_t st)
{
    const bool lp = p.order > 1 && p.limited && p.lpThresh > 0;
    const dim3 g(t.nTiles), b(BLK);
    if (p.mixed) {   // no LP vector limiter on the mixed path (rejected at set-up)
        if (p.order == 1) hipLaunchKernelGGL((k_face_tile<S, FLUX, VISC, 1, 0, 1>), g, b, t.ldsBytes, st, m, p, a, t);
        else if (!(p.limited && a.limit))
            hipLaunchKernelGGL((k_face_tile<S, FLUX, VISC, 2, 2, 1>), g, b, t.ldsBytes, st, m, p, a, t);
        else hipLaunchKernelGGL((k_face_tile<S, FLUX, VISC, 2, 0, 1>), g, b, t.ldsBytes, st, m, p, a, t);
        return;
    }
    if (p.order == 1) hipLaunchKernelGGL((k_face_tile<S, FLUX, VISC, 1, 0, 0>), g, b, t.ldsBytes, st, m, p, a, t);
    else if (!(p.limited && a.limit))
        hipLaunchKernelGGL((k_face_tile<S, FLUX, VISC, 2, 2, 0>), g, b, t.ldsBytes, st, m, p, a, t);
    else if (lp) hipLaunchKernelGGL((k_face_tile<S, FLUX, VISC, 2, 1, 0>), g, b, t.ldsBytes, st, m, p, a, t);
    else hipLaunchKernelGGL((k_face_tile<S, FLUX, VISC, 2, 0, 0>), g, b, t.ldsBytes, st, m, p, a, t);
}

template <int S>
static void faceTileS(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t, hipStream_t st)
{
    if (p.flux == 1) {
        if (p.viscous) faceTileO<S, 1, 1>(m, p, a, t, st); else faceTileO<S, 1, 0>(m, p, a, t, st);
    } else {
        if (p.viscous) faceTileO<S, 0, 1>(m, p, a, t, st); else faceTileO<S, 0, 0>(m, p, a, t, st);
    }
}

void launchFaceTiles(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t, int S,
                     hipStream_t st)
{
    if (t.nTiles == 0) return;
    if (S == 3) faceTileS<3>(m, p, a, t, st);
    else if (S == 4) faceTileS<4>(m, p, a, t, st);
    else faceTileS<6>(m, p, a, t, st);
    if (!a.deferReduce) hipLaunchKernelGGL(k_reduce_stage, dim3(1), dim3(1024), 0, st, a.partial, t.nTiles, a.sc, a.checkActive);
}

// ---------------------------------------------------------------- gradient tiles
// A stage in two launches that keeps most gradient and limiter rows on chip.
//   k_border_grad  per ring cell (a cell across a face of some other tile): its gradient
//                  and limiter rows to HBM; it also carries the gradient kernel's other
//                  duties (the previous step's boundary force, the per-step Rmax reset);
//   k_grad_tile    per tile: its own cells' ghosts, gradients and limiters into LDS, the
//                  ring cells' rows copied in from the border kernel's output, then every
//                  face of the tile and the cells' update as in k_face_tile.
// Gradients come from cellGradient and faces from interiorFace/boundaryFace, the code of
// the split kernels, on the same inputs, so the stage's results are bit-identical to the
// other paths.  A boundary face takes its ghost state from its owner's row (ghostState)
// instead of reading back the row this launch's gradient phase wrote.

// LDS row of one cell: gradient (x,y per variable) then the 4 scalar limiters
constexpr int GROW = 12;

// One workgroup of the border kernel on border block `bid` of its mesh (the solo and the
// batched kernels run this same code); block 0 also resets the step's Rmax.
template <int S, int MODE, int MIX>
DEV void borderGradBody(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t, uint32_t bid)
{
    using R = RealOf<MIX>;
    if (a.checkActive && !a.sc->active) return;
    if (bid == 0 && threadIdx.x == 0 && a.first) a.sc->rmaxBits = 0;   // Rmax is per step
    const uint32_t k = bid * BLK + threadIdx.x;
    if (k >= t.nBorder) return;
    const uint32_t i = t.border[k];
    R g[4][2], lx[4], ly[4];
    cellGradient<S, MODE, MIX>(m, p, a, i, m.nbr + size_t(i) * S, false, g, lx, ly);   // the owning tile stores the ghosts
    if constexpr (MIX == 1) {
        float4* gd = reinterpret_cast<float4*>(a.grad32 + size_t(i) * 8);
        gd[0] = make_float4(g[0][0], g[0][1], g[1][0], g[1][1]);
        gd[1] = make_float4(g[2][0], g[2][1], g[3][0], g[3][1]);
        if (MODE != 1) *reinterpret_cast<float4*>(a.lim32 + size_t(i) * 4) = make_float4(lx[0], lx[1], lx[2], lx[3]);
    } else {
        double4* gd = reinterpret_cast<double4*>(a.grad + size_t(i) * 8);
        gd[0] = make_double4(g[0][0], g[0][1], g[1][0], g[1][1]);
        gd[1] = make_double4(g[2][0], g[2][1], g[3][0], g[3][1]);
        if (MODE != 1) *reinterpret_cast<double4*>(a.lim + size_t(i) * 4) = make_double4(lx[0], lx[1], lx[2], lx[3]);
    }
}

// Workgroup 0 sums a pending boundary force when there is one (as the gradient kernel's
// force workgroup); the others are border blocks.
template <int S, int MODE, int MIX>
__global__ __launch_bounds__(BLK) void k_border_grad(DevMesh m, PhysParams p, StageArgs a, FaceTileArgs t)
{
    __shared__ double2 fstage[FORCE_LDS];
    const uint32_t fw = a.force.n ? 1u : 0u;
    if (fw && blockIdx.x == 0) {
        pendingForce<BLK, FORCE_LDS>(a.force, a.sc, fstage);
        return;
    }
    borderGradBody<S, MODE, MIX>(m, p, a, t, xcdBlock(blockIdx.x - fw, gridDim.x - fw));
}

// One workgroup of the gradient-tile kernel on tile `tile`.  MODE 1: an unlimited stage
// (no limiter rows; the faces take unit limiters); 2: the scalar limiter; 4: Venkatakrishnan.
template <int S, int FLUX, int VISC, int MODE, int MIX>
DEV void gradTileBody(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t,
                      uint32_t tile, uint32_t pslot, double* lds)
{
    using R = RealOf<MIX>;
    constexpr int LP = MODE == 1 ? 2 : 0;
    const FaceSrc<MIX> src{a};
    const uint32_t tid = threadIdx.x;
    // first loads beside the descriptor's (fixed strides; lanes past the counts read padding)
    const uint32_t gi0 = t.ids[size_t(tile) * t.cellStride + tid];
    uint4 rNext = t.fref[size_t(tile) * t.faceStride + tid];
    uint32_t sNext = t.fslot[size_t(tile) * t.faceStride + tid];
    const uint4 d = t.desc[tile];
    const uint32_t nH = t.nHalo[tile];
    if (a.checkActive && !a.sc->active) return;
    const uint32_t nC = d.x, nF = d.y;
    R* gl = reinterpret_cast<R*>(lds);
    const FaceLds<MIX> fl(lds + t.glDoubles, t.faceStride + 1);
    // the ring's rows first: their loads are in flight while the own gradients compute
    for (uint32_t h = tid; h < nH; h += BLK) {
        const uint32_t y = t.halo[size_t(tile) * t.haloStride + h];
        R gg[8];
        src.grad(y, gg);
        R* row = gl + GROW * (nC + h);
#pragma unroll
        for (int c = 0; c < 8; c++) row[c] = gg[c];
        if (MODE != 1) {
            R lx[4], ly[4];
            src.template lim<0>(y, lx, ly);
#pragma unroll
            for (int c = 0; c < 4; c++) row[8 + c] = lx[c];
        }
    }
    // own cells (ghosts stored: this tile owns them), neighbour ids from the tile's copy of
    // their rows, whose addresses do not wait for the cell ids
    for (uint32_t k = tid; k < nC; k += BLK) {
        const uint32_t i = k == tid ? gi0 : t.ids[d.z + k];
        R g[4][2], lx[4], ly[4];
        cellGradient<S, MODE, MIX>(m, p, a, i, t.nbrT + size_t(d.z + k) * S, true, g, lx, ly);
        R* row = gl + GROW * k;
#pragma unroll
        for (int c = 0; c < 4; c++) { row[2 * c] = g[c][0]; row[2 * c + 1] = g[c][1]; }
        if (MODE != 1) {
#pragma unroll
            for (int c = 0; c < 4; c++) row[8 + c] = lx[c];
        }
    }
    // the first cell rounds' ids for the cell phase
    const uint32_t giNext = gi0;
    const uint32_t giNext2 = t.ids[size_t(tile) * t.cellStride + tid + BLK];
    ldsBarrier();
    // face phase: as k_face_tile, with the rows from LDS
    for (uint32_t k = tid; k < nF; k += BLK) {
        const uint4 r = rNext;   // cell a, cell b, node 0, node 1
        const uint32_t sl = sNext;
        if (k + BLK < nF) { rNext = t.fref[d.w + k + BLK]; sNext = t.fslot[d.w + k + BLK]; }
        const bool interior = r.y < m.N;
        const double2 p0 = m.node[r.z], p1 = m.node[r.w];
        R qa[4], qb[4];
        src.q(r.x, qa);
        src.q(r.y, qb);   // a boundary face: the ghost row, whose TempPresInflow components carry over
        const double2 ca = m.cen[r.x], cb = m.cen[r.y];
        const R* ra = gl + GROW * (sl & 0xffffu);
        const R* rb = gl + GROW * (sl >> 16);
        R ga[8], gb[8], lxa[4], lya[4], lxb[4], lyb[4];
#pragma unroll
        for (int c = 0; c < 8; c++) { ga[c] = ra[c]; gb[c] = rb[c]; }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            lxa[c] = lya[c] = MODE == 1 ? R(1.0) : ra[8 + c];
            lxb[c] = lyb[c] = MODE == 1 ? R(1.0) : rb[8 + c];
        }
        double2 n, mid;
        double len;
        edgeGeometry(interior, p0, p1, n, mid, len);
        R qL[4], qR[4], Fr[4], sMr;
        bool isB = false;
        uint32_t b = 0;
        if (interior) {
            interiorFace<FLUX, VISC, 2, LP, R>(qa, ga, lxa, lya, ca, qb, gb, lxb, lyb, cb, n, mid, p, qL, qR, Fr, sMr);
        } else {
            isB = true;
            b = r.y - m.N;
            const BFace& bf = m.bface[b];
            R qg[4];
            ghostState<R>(bf, qa, qg, qb);   // a failure was recorded by the gradient phase
            boundaryFace<FLUX, VISC, 2, LP, R>(qa, ga, lxa, lya, ca, qg, bf, n, mid, p, qL, qR, Fr, sMr);
        }
        const double F[4] = {double(Fr[0]), double(Fr[1]), double(Fr[2]), double(Fr[3])}, sMax = double(sMr);
        if (isB) {
            const uint32_t f = m.Fi + b;
            *reinterpret_cast<double4*>(a.lenF + size_t(f) * 4) =
                make_double4(len * F[0], len * F[1], len * F[2], len * F[3]);
            a.lenS[f] = len * sMax;
        }
        if (a.recordErrors && (isnan(F[0]) || isnan(F[1]) || isnan(F[2]) || isnan(F[3])))
            atomicMin(&a.sc->errEdge,
                      (uint64_t(a.stageId) << 32) | (isB ? uint64_t(b) : uint64_t(m.Fb) + t.fid[d.w + k]));
        if constexpr (MIX == 1) fl.put(k, len, Fr, sMr);
        else fl.put(k, len, F, sMax);
    }
    fl.zero(nF, tid);   // padding face
    tileCellPhase<S, MIX>(m, p, a, t, d, pslot, fl, giNext, giNext2);
}

// FP64: at most 168 VGPRs, three waves per SIMD (the LDS of three workgroups fits a CU)
template <int S, int FLUX, int VISC, int MODE, int MIX>
__global__ __launch_bounds__(BLK, 3) void k_grad_tile(DevMesh m, PhysParams p, StageArgs a, FaceTileArgs t)
{
    extern __shared__ double glds[];
    gradTileBody<S, FLUX, VISC, MODE, MIX>(m, p, a, t, xcdBlock(blockIdx.x, gridDim.x), blockIdx.x, glds);
}

template <int S, int FLUX, int VISC, int MIX>
static void gradTileM(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t, int mode,
                      int parts, hipStream_t st)
{
    const dim3 g(t.nTiles), b(BLK);
    const uint32_t nb = (t.nBorder + BLK - 1) / BLK;
    const dim3 gb((nb ? nb : 1u) + (a.force.n ? 1u : 0u));   // at least one workgroup: the Rmax reset
    const bool border = parts & 1, tiles = parts & 2;
    switch (mode) {
    case 1:
        if (border) hipLaunchKernelGGL((k_border_grad<S, 1, MIX>), gb, b, 0, st, m, p, a, t);
        if (tiles) hipLaunchKernelGGL((k_grad_tile<S, FLUX, VISC, 1, MIX>), g, b, t.ldsBytes, st, m, p, a, t);
        break;
    case 4:
        if (border) hipLaunchKernelGGL((k_border_grad<S, 4, MIX>), gb, b, 0, st, m, p, a, t);
        if (tiles) hipLaunchKernelGGL((k_grad_tile<S, FLUX, VISC, 4, MIX>), g, b, t.ldsBytes, st, m, p, a, t);
        break;
    default:
        if (border) hipLaunchKernelGGL((k_border_grad<S, 2, MIX>), gb, b, 0, st, m, p, a, t);
        if (tiles) hipLaunchKernelGGL((k_grad_tile<S, FLUX, VISC, 2, MIX>), g, b, t.ldsBytes, st, m, p, a, t);
        break;
    }
}

template <int S>
static void gradTileS(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t, int mode,
                      int parts, hipStream_t st)
{
#define EBB_GT(F, V)                                                      \
    do {                                                                  \
        if (p.mixed) gradTileM<S, F, V, 1>(m, p, a, t, mode, parts, st);  \
        else gradTileM<S, F, V, 0>(m, p, a, t, mode, parts, st);          \
    } while (0)
    if (p.flux == 1) {
        if (p.viscous) EBB_GT(1, 1); else EBB_GT(1, 0);
    } else {
        if (p.viscous) EBB_GT(0, 1); else EBB_GT(0, 0);
    }
#undef EBB_GT
}

void launchGradTiles(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t, int S,
                     hipStream_t st, int parts)
{
    if (t.nTiles == 0) return;
    // order 2 and no LP vector limiter (the set-up keeps other configurations on face tiles)
    const int mode = !(p.limited && a.limit) ? 1 : (p.venkat ? 4 : 2);
    if (S == 3) gradTileS<3>(m, p, a, t, mode, parts, st);
    else if (S == 4) gradTileS<4>(m, p, a, t, mode, parts, st);
    else gradTileS<6>(m, p, a, t, mode, parts, st);
    if ((parts & 2) && !a.deferReduce)
        hipLaunchKernelGGL(k_reduce_stage, dim3(1), dim3(1024), 0, st, a.partial, t.nTiles, a.sc, a.checkActive);
}

template <int S, int FLUX, int VISC, int LP>
static void tileLaunch(const DevMesh& m, const PhysParams& p, const StageArgs& a, const TileArgs& t, hipStream_t st)
{
    hipLaunchKernelGGL((k_stage_tile<S, FLUX, VISC, LP>), dim3(t.nBlocks), dim3(TBLK), t.ldsBytes, st, m, p, a, t);
}

template <int S>
static void tileS(const DevMesh& m, const PhysParams& p, const StageArgs& a, const TileArgs& t, hipStream_t st)
{
    const bool lp = p.order > 1 && p.limited && p.lpThresh > 0;
    if (p.flux == 1) {
        if (p.viscous) { if (lp) tileLaunch<S, 1, 1, 1>(m, p, a, t, st); else tileLaunch<S, 1, 1, 0>(m, p, a, t, st); }
        else { if (lp) tileLaunch<S, 1, 0, 1>(m, p, a, t, st); else tileLaunch<S, 1, 0, 0>(m, p, a, t, st); }
    } else {
        if (p.viscous) { if (lp) tileLaunch<S, 0, 1, 1>(m, p, a, t, st); else tileLaunch<S, 0, 1, 0>(m, p, a, t, st); }
        else { if (lp) tileLaunch<S, 0, 0, 1>(m, p, a, t, st); else tileLaunch<S, 0, 0, 0>(m, p, a, t, st); }
    }
}

void launchStageTiles(const DevMesh& m, const PhysParams& p, const StageArgs& a, const TileArgs& t, int S,
                      hipStream_t st)
{
    if (t.nTiles == 0) return;
    if (S == 3) tileS<3>(m, p, a, t, st);
    else if (S == 4) tileS<4>(m, p, a, t, st);
    else tileS<6>(m, p, a, t, st);
    if (!a.deferReduce) hipLaunchKernelGGL(k_reduce_stage, dim3(1), dim3(1024), 0, st, a.partial, t.nBlocks, a.sc, a.checkActive);
}

// newDt = min over the stage's workgroup partials; Rmax = max with the step's running value
__global__ __launch_bounds__(1024) void k_reduce_stage(const double* __restrict__ partial, uint32_t n,
                                                       DevScalars* sc, int checkActive)
{
    if (checkActive && !sc->active) return;
    __shared__ double sMin[16], sMax_[16];
    double a0 = INFINITY, b0 = 0.0;
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        a0 = fmin(a0, partial[2 * k]);
        b0 = fmax(b0, partial[2 * k + 1]);
    }
    a0 = warpMin(a0);
    b0 = warpMax(b0);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) { sMin[w] = a0; sMax_[w] = b0; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t k = 1; k < blockDim.x / 64; k++) { a0 = fmin(a0, sMin[k]); b0 = fmax(b0, sMax_[k]); }
        a0 = fmin(a0, sMin[0]);
        b0 = fmax(b0, sMax_[0]);
        sc->newDtBits = (unsigned long long)__double_as_longlong(a0);
        const double prev = __longlong_as_double((long long)sc->rmaxBits);
        if (b0 > prev) sc->rmaxBits = (unsigned long long)__double_as_longlong(b0);
    }
}

// ---------------------------------------------------------------- diagnostics
// Phobos std.algorithm.sum over double[] (app/solver.d:153-168): 16-element
// blocks summed by a fixed tree, blocks merged by a binary counter (perfect
// pairwise trees over aligned power-of-two runs), remainder elements likewise,
// and the stack folded from the top.  A workgroup reduces an aligned chunk of
// RMS_CHUNK_BLOCKS blocks level by level; the finalize kernel replays the
// counter over the chunk trees.  Quantities 0..3: sum of (now-last)^2 per
// variable; 4..7: sum of now.
DEV void rmsElement(const DiagArgs& d, uint64_t e, double out[8])
{
    if (e < d.N) {
        const uint32_t dv = d.perm ? d.perm[e] : uint32_t(e);
        double qa[4], qb[4];
        ld4(d.qnew, dv, qa);
        ld4(d.qold, dv, qb);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double x = qa[v] - qb[v];
            out[v] = x * x;
            out[4 + v] = qa[v];
        }
    } else {
#pragma unroll
        for (int v = 0; v < 8; v++) out[v] = 0.0;   // ghost slots hold 0 in both arrays
    }
}

constexpr int RBLK = 1024;   // threads per chunk workgroup (element path)
constexpr int FTILE = 2048;   // force faces staged per pass (RMS kernels)

// Pairwise levels of one chunk: v[q][0..cnt) holds its 16-element block sums.
// Level l's odd leftover is stored as tree l of the chunk (the binary counter's
// stack), mask marks the levels present.  NT threads.
template <int NT>
DEV void chunkLevels(double (*v)[RMS_CHUNK_BLOCKS], uint32_t cnt, double* out, uint32_t* maskOut)
{
    // NT/8 threads per quantity; slot u of a thread is pair l + u * G (registers, no
    // divisions by the level's length)
    constexpr uint32_t G = NT / 8;
    constexpr int PER = int((RMS_CHUNK_BLOCKS / 2 + G - 1) / G);
    const uint32_t qq = threadIdx.x / G, l = threadIdx.x % G;
    uint32_t c = cnt;
    int level = 0;
    uint32_t mask = 0;
    while (c > 0) {
        if (c & 1) {
            if (threadIdx.x < 8) out[threadIdx.x * 16 + level] = v[threadIdx.x][c - 1];
            mask |= 1u << level;
        }
        const uint32_t half = c / 2;
        double tmp[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const uint32_t j = l + uint32_t(u) * G;
            tmp[u] = j < half ? v[qq][2 * j] + v[qq][2 * j + 1] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const uint32_t j = l + uint32_t(u) * G;
            if (j < half) v[qq][j] = tmp[u];
        }
        __syncthreads();
        c = half;
        level++;
    }
    if (threadIdx.x < 8) maskOut[threadIdx.x] = mask;
}

// The boundary force (mesh.d:645-671): len*flux of the force faces summed in the
// reference's edge order by one lane, from LDS-staged values; NT threads stage.
// The stage reductions of a time-loop step, done here instead of after each stage:
// newDt is the last stage's minimum (every stage overwrites it), Rmax the maximum
// over the step's stages; min and max are exact in any order.
template <int NT>
DEV void stageReduce(const DiagArgs& d, double* lds)
{
    // each pass issues all of a thread's loads before using any (the partials were
    // written on all XCDs: every load is a round trip past this XCD's L2)
    constexpr int SLOTS = 24;
    double a0 = INFINITY, b0 = 0.0;
    for (uint32_t s = 0; s < d.nStages; s++) {
        const double2* pp = reinterpret_cast<const double2*>(d.partial + size_t(s) * d.pstride);
        const bool lastStage = s + 1 == d.nStages;
        for (uint32_t base = 0; base < d.nPart; base += NT * SLOTS) {
            double2 pv[SLOTS];
#pragma unroll
            for (int u = 0; u < SLOTS; u++) {
                const uint32_t k = base + threadIdx.x + uint32_t(u) * NT;
                pv[u] = k < d.nPart ? pp[k] : make_double2(INFINITY, 0.0);
            }
#pragma unroll
            for (int u = 0; u < SLOTS; u++) {
                if (lastStage) a0 = fmin(a0, pv[u].x);
                b0 = fmax(b0, pv[u].y);
            }
        }
    }
    a0 = warpMin(a0);
    b0 = warpMax(b0);
    const uint32_t w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) { lds[2 * w] = a0; lds[2 * w + 1] = b0; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t k = 1; k < NT / 64; k++) { a0 = fmin(a0, lds[2 * k]); b0 = fmax(b0, lds[2 * k + 1]); }
        d.sc->newDtBits = (unsigned long long)__double_as_longlong(a0);
        const double prev = __longlong_as_double((long long)d.sc->rmaxBits);
        if (b0 > prev) d.sc->rmaxBits = (unsigned long long)__double_as_longlong(b0);
    }
}
// force workgroup of the RMS kernels: the step's force sums for k_finalize (unless
// they are left to the next gradient kernel)
template <int NT>
DEV void forceSum(const DiagArgs& d, double* fb)
{
    if (d.deferForce) return;
    double f0, f1;
    orderedForce<NT, FTILE>(d.forceFaces, d.nForce, d.lenF, reinterpret_cast<double2*>(fb), f0, f1);
    if (threadIdx.x == 0) { d.sc->fsum[0] = f0; d.sc->fsum[1] = f1; }
}


// RMS sums from the states (element path): one workgroup per chunk of
// RMS_CHUNK_BLOCKS 16-element blocks; 16 consecutive lanes own one block, and
// the xor butterfly over them adds exactly like the fixed tree
// (((r0+r1)+(r2+r3))+...).  The last workgroup sums the boundary force.
__global__ __launch_bounds__(RBLK) void k_rms_chunks(DiagArgs d)
{
    if (!d.sc->active) return;
    __shared__ double v[8][RMS_CHUNK_BLOCKS];
    if (blockIdx.x == gridDim.x - 2) { forceSum<RBLK>(d, &v[0][0]); return; }
    if (blockIdx.x == gridDim.x - 1) { if (d.nStages) stageReduce<RBLK>(d, &v[0][0]); return; }
    const uint64_t nb = uint64_t(d.L) / 16;
    const uint64_t c0 = uint64_t(blockIdx.x) * RMS_CHUNK_BLOCKS;
    const uint32_t cnt = uint32_t((nb - c0) < RMS_CHUNK_BLOCKS ? (nb - c0) : RMS_CHUNK_BLOCKS);
    constexpr uint32_t BPI = RBLK / 16;
    for (uint32_t it = 0; it < RMS_CHUNK_BLOCKS / BPI; it++) {
        const uint32_t lb = it * BPI + threadIdx.x / 16;
        double o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (lb < cnt) rmsElement(d, (c0 + lb) * 16 + (threadIdx.x % 16), o);
#pragma unroll
        for (int m = 1; m < 16; m <<= 1)
#pragma unroll
            for (int q = 0; q < 8; q++) o[q] = o[q] + __shfl_xor(o[q], m, 64);
        if ((threadIdx.x % 16) == 0 && lb < cnt)
#pragma unroll
            for (int q = 0; q < 8; q++) v[q][lb] = o[q];
    }
    __syncthreads();
    chunkLevels<RBLK>(v, cnt, d.chunkTrees + size_t(blockIdx.x) * 8 * 16, d.chunkMask + blockIdx.x * 8);
}

// RMS sums from the leaf sums the last stage stored (blockSums, 8 per block;
// blocks wholly past the interior cells are zero).  Workgroups [0, nfull): one
// full chunk each, only its top tree is needed: each lane adds 4 consecutive
// blocks pairwise, the 64-lane xor butterfly adds levels 3..8 exactly as the
// pairwise tree does (the two orders of each sum are equal), the 4 waves levels
// 9 and 10.  Workgroup nfull: the partial last chunk; nfull+1: the boundary force.
constexpr int RFB = 256;
DEV void rmsBlocksBody(const DiagArgs& d, uint32_t nfull, uint32_t wg, double (*v)[RMS_CHUNK_BLOCKS])
{
    if (!d.sc->active) return;
    const uint64_t nb = uint64_t(d.L) / 16;
    if (wg == nfull + 1) { forceSum<RFB>(d, &v[0][0]); return; }
    if (wg == nfull + 2) { if (d.nStages) stageReduce<RFB>(d, &v[0][0]); return; }
    const uint64_t c0 = uint64_t(wg) * RMS_CHUNK_BLOCKS;
    auto leaf = [&](uint64_t b, double o[8]) {
        if (b * 16 < d.N) {
            const double4 x = *reinterpret_cast<const double4*>(d.blockSums + b * 8);
            const double4 y = *reinterpret_cast<const double4*>(d.blockSums + b * 8 + 4);
            o[0] = x.x; o[1] = x.y; o[2] = x.z; o[3] = x.w; o[4] = y.x; o[5] = y.y; o[6] = y.z; o[7] = y.w;
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++) o[q] = 0.0;
        }
    };
    if (wg == nfull) {   // partial chunk
        const uint32_t cnt = uint32_t(nb - c0);
        for (uint32_t lb = threadIdx.x; lb < cnt; lb += RFB) {
            double o[8];
            leaf(c0 + lb, o);
#pragma unroll
            for (int q = 0; q < 8; q++) v[q][lb] = o[q];
        }
        __syncthreads();
        chunkLevels<RFB>(v, cnt, d.chunkTrees + size_t(nfull) * 8 * 16, d.chunkMask + nfull * 8);
        return;
    }
    double b[4][8], t[8];
#pragma unroll
    for (int u = 0; u < 4; u++) leaf(c0 + 4 * threadIdx.x + u, b[u]);
#pragma unroll
    for (int q = 0; q < 8; q++) t[q] = (b[0][q] + b[1][q]) + (b[2][q] + b[3][q]);
#pragma unroll
    for (int m = 1; m < 64; m <<= 1)
#pragma unroll
        for (int q = 0; q < 8; q++) t[q] = t[q] + __shfl_xor(t[q], m, 64);
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < 8; q++) v[q][w] = t[q];
    __syncthreads();
    if (threadIdx.x < 8) {
        const int q = threadIdx.x;
        d.chunkTrees[size_t(wg) * 128 + q * 16 + 10] = (v[q][0] + v[q][1]) + (v[q][2] + v[q][3]);
    }
}

__global__ __launch_bounds__(RFB) void k_rms_blocks(DiagArgs d, uint32_t nfull)
{
    __shared__ double v[8][RMS_CHUNK_BLOCKS];
    rmsBlocksBody(d, nfull, blockIdx.x, v);
}

// Step epilogue (app/solver.d:122-247), one workgroup.  The chunk trees are
// merged level by level in LDS (a perfect tree over each aligned power-of-two
// run of chunks is exactly the binary counter's merge); the boundary force comes
// summed from the RMS kernel.
constexpr int FIN_T = 256;
constexpr int FIN_CHUNKS = 1024;     // chunk trees merged in LDS per quantity

struct FinLds {
    double v[8][FIN_CHUNKS];
    double trees[8][48];
    int ntree[8];
    double tot[8];
};

DEV void finalizeBody(const DiagArgs& d, FinLds& L)
{
    DevScalars* sc = d.sc;
    auto& v = L.v;
    auto& trees = L.trees;
    auto& ntree = L.ntree;
    auto& tot = L.tot;
    const int tid = threadIdx.x;
    const uint64_t nb = uint64_t(d.L) / 16;
    const uint64_t nfull = nb / RMS_CHUNK_BLOCKS, part = nb % RMS_CHUNK_BLOCKS;
    // Every global read of the kernel is issued here, before the first is used (the
    // active flag included): the previous kernels wrote this data on all XCDs, so each
    // read is a round trip past this XCD's L2, and a chain of them set the kernel's length.
    // FIN_T/8 threads per quantity; slot u of a thread is chunk l + u * G.
    constexpr uint32_t G = FIN_T / 8;
    constexpr int PER = int(FIN_CHUNKS / G), HPER = int(FIN_CHUNKS / 2 / G);
    const uint32_t qq = uint32_t(tid) / G, l = uint32_t(tid) % G;
    const uint32_t n = nfull <= FIN_CHUNKS ? uint32_t(nfull) : 0u;
    double tv[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const uint32_t k = l + uint32_t(u) * G;
        tv[u] = k < n ? d.chunkTrees[size_t(k) * 128 + qq * 16 + 10] : 0.0;
    }
    uint32_t pmask = 0;   // the partial chunk's level trees (thread q < 8: quantity q)
    double ptree[16];
#pragma unroll
    for (int b2 = 0; b2 < 16; b2++) ptree[b2] = 0.0;
    if (tid < 8 && part) {
        pmask = d.chunkMask[nfull * 8 + tid];
#pragma unroll
        for (int b2 = 0; b2 < 16; b2++) ptree[b2] = d.chunkTrees[nfull * 128 + tid * 16 + b2];
    }
    const DevScalars s0 = *sc;
    if (!s0.active) return;
    if (tid < 8) ntree[tid] = 0;
    __syncthreads();
    if (nfull <= FIN_CHUNKS) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const uint32_t k = l + uint32_t(u) * G;
            if (k < n) v[qq][k] = tv[u];
        }
        __syncthreads();
        uint32_t c = n;
        int level = 0;
        uint32_t lmask = 0;
        while (c > 0) {
            if (c & 1) {
                if (tid < 8) trees[tid][47 - level] = v[tid][c - 1];
                lmask |= 1u << level;
            }
            const uint32_t half = c / 2;
            double tmp[HPER];
#pragma unroll
            for (int u = 0; u < HPER; u++) {
                const uint32_t j = l + uint32_t(u) * G;
                tmp[u] = j < half ? v[qq][2 * j] + v[qq][2 * j + 1] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < HPER; u++) {
                const uint32_t j = l + uint32_t(u) * G;
                if (j < half) v[qq][j] = tmp[u];
            }
            __syncthreads();
            c = half;
            level++;
        }
        // chunk-level trees, largest first
        if (tid < 8) {
            int t = 0;
            for (int l = 31; l >= 0; l--)
                if (lmask & (1u << l)) trees[tid][t++] = trees[tid][47 - l];
            ntree[tid] = t;
        }
    } else if (tid < 8) {
        // very large meshes: replay the binary counter over the chunk trees directly, on
        // the quantity's LDS stack (no scratch)
        double* st = trees[tid];
        int idx = 0;
        for (uint64_t k = 0; k < nfull; k++) {
            st[idx] = d.chunkTrees[k * 128 + tid * 16 + 10];
            const int last = idx - __builtin_ctzll(k + 1);
            while (idx > last) { st[idx - 1] += st[idx]; --idx; }
            ++idx;
        }
        ntree[tid] = idx;
    }
    __syncthreads();
    if (tid < 8) {
        const int q = tid;
        int idx = ntree[q];
        double* st = trees[q];
        if (part) {
#pragma unroll
            for (int b2 = 15; b2 >= 0; b2--)
                if (pmask & (1u << b2)) st[idx++] = ptree[b2];
        }
        const uint64_t rem = uint64_t(d.L) - nb * 16;
        for (uint64_t i = 0; i < rem; i++) {
            double o[8];
            rmsElement(d, nb * 16 + i, o);
            double oq = o[0];   // o[q] by selects, so o stays in registers
#pragma unroll
            for (int k = 1; k < 8; k++) oq = q == k ? o[k] : oq;
            st[idx] = oq;
            const int last = idx - __builtin_ctzll(i + 1);
            while (idx > last) { st[idx - 1] += st[idx]; --idx; }
            ++idx;
        }
        double s = 0.0;
        if (idx > 0) {
            s = st[idx - 1];
            for (int j2 = idx - 2; j2 >= 0; j2--) s += st[j2];
        }
        tot[q] = 0.0 + s;
    }
    __syncthreads();
    if (tid != 0) return;
    const double f0 = s0.fsum[0], f1 = s0.fsum[1];   // summed by the RMS kernel's force workgroup (unless deferred)
    double rms[4];
    for (int q = 0; q < 4; q++) rms[q] = sqrt(double(d.N) * tot[q]) / tot[4 + q];
    const double newDt = __longlong_as_double((long long)s0.newDtBits);
    const double Rmax = __longlong_as_double((long long)s0.rmaxBits);
    const double dt = d.dynamicDt ? newDt : s0.dt;
    const double ld0 = 0.0 + d.rot[0] * f0 + d.rot[1] * f1;
    const double ld1 = 0.0 + d.rot[2] * f0 + d.rot[3] * f1;
    const double residMax = dmax(dmax(dmax(rms[0], rms[1]), rms[2]), rms[3]);
    int64_t inc = s0.incIters;
    double CFL = s0.CFL;
    if (d.cflAdjust) {
        if (s0.residLast < residMax) inc++;
        else if (inc >= 1) inc--;
        if (inc >= 2000) { CFL *= 0.99; inc = 0; }
    }
    // the history row: by absolute step within the host's batch, or a ring (batched runs)
    const int64_t slot = d.histRing ? s0.steps % d.histCap : s0.steps - d.histBase;
    if (slot >= 0 && slot < d.histCap) {
        double* h = d.hist + slot * HIST_FIELDS;
        h[0] = s0.t; h[1] = dt; h[2] = rms[0]; h[3] = rms[1]; h[4] = rms[2]; h[5] = rms[3];
        h[6] = Rmax; h[9] = CFL;
        if (!d.deferForce) { h[7] = ld0; h[8] = ld1; }
    }
    const double t1 = s0.t + dt;
    const int64_t steps1 = s0.steps + 1;
    sc->dt = dt;
    sc->incIters = inc;
    sc->CFL = CFL;
    sc->residLast = residMax;
    for (int q = 0; q < 4; q++) sc->rms[q] = rms[q];
    if (d.deferForce) {
        // drag and lift of this step are summed by the next gradient kernel (or at the end of
        // the batch), which writes them into this row and sc->force
        sc->forcePending = 1;
        sc->forceSlot = (slot >= 0 && slot < d.histCap) ? slot : -1;
    } else {
        sc->force[0] = ld0; sc->force[1] = ld1;
    }
    sc->t = t1;
    sc->steps = steps1;
    // convergence stop (extensions.rmsDrop): the rho RMS residual of app/solver.d:153-168 has
    // dropped by the given factor below its value on the first step
    const double rms0 = s0.steps == 0 ? rms[0] : s0.rms0;
    sc->rms0 = rms0;
    const bool converged = d.rmsDrop > 0 && rms[0] <= d.rmsDrop * rms0;
    if (converged) sc->converged = 1;
    if (!(t1 < d.tEnd) || (d.maxSteps >= 0 && steps1 >= d.maxSteps) || s0.errEdge != ~0ull || s0.errCell != ~0ull ||
        s0.errBC != ~0ull || converged)
        sc->active = 0;
}

__global__ __launch_bounds__(FIN_T) void k_finalize(DiagArgs d)
{
    __shared__ FinLds L;
    finalizeBody(d, L);
}

// FP32 shadow of a state array (mixed path): after set-up, an upload or a load
__global__ __launch_bounds__(BLK) void k_shadow(const double4* __restrict__ q, float4* __restrict__ q32, uint32_t n)
{
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const double4 v = q[i];
    q32[i] = make_float4(float(v.x), float(v.y), float(v.z), float(v.w));
}

void launchShadow(const double4* q, float4* q32, uint32_t n, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_shadow, dim3((n + BLK - 1) / BLK), dim3(BLK), 0, st, q, q32, n);
}

__global__ void k_init_scalars(DevScalars* sc, double dt, double t, double CFL)
{
    sc->newDtBits = 0x7FF0000000000000ull;
    sc->rmaxBits = 0;
    sc->errEdge = ~0ull;
    sc->errCell = ~0ull;
    sc->errBC = ~0ull;
    sc->dt = dt;
    sc->t = t;
    sc->CFL = CFL;
    sc->residLast = INFINITY;
    sc->incIters = 0;
    sc->steps = 0;
    sc->active = 1;
    sc->converged = 0;
    sc->rms0 = 0;
    for (int v = 0; v < 4; v++) sc->rms[v] = 0;
    sc->force[0] = sc->force[1] = 0;
    sc->fsum[0] = sc->fsum[1] = 0;
    sc->forcePending = 0;
    sc->pad2 = 0;
    sc->forceSlot = -1;
}

// ---------------------------------------------------------------- launchers
static uint32_t nblk(uint64_t n) { return uint32_t((n + BLK - 1) / BLK); }

template <int S>
static void gradS(const DevMesh& m, const PhysParams& p, const StageArgs& a, hipStream_t st)
{
    const uint32_t g = nblk(m.N) + (a.force.n ? 1u : 0u);   // + the force workgroup
    if (g == 0) return;
    const int mode = p.order == 1 ? 0 : (!(p.limited && a.limit) ? 1 : (p.venkat ? 4 : (p.lpThresh > 0 ? 3 : 2)));
    if (p.mixed) {
        switch (mode) {
        case 0: hipLaunchKernelGGL((k_gradient<S, 0, 1>), dim3(g), dim3(BLK), 0, st, m, p, a); break;
        case 1: hipLaunchKernelGGL((k_gradient<S, 1, 1>), dim3(g), dim3(BLK), 0, st, m, p, a); break;
        case 4: hipLaunchKernelGGL((k_gradient<S, 4, 1>), dim3(g), dim3(BLK), 0, st, m, p, a); break;
        default: hipLaunchKernelGGL((k_gradient<S, 2, 1>), dim3(g), dim3(BLK), 0, st, m, p, a); break;
        }
        return;
    }
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_gradient<S, 0, 0>), dim3(g), dim3(BLK), 0, st, m, p, a); break;
    case 1: hipLaunchKernelGGL((k_gradient<S, 1, 0>), dim3(g), dim3(BLK), 0, st, m, p, a); break;
    case 2: hipLaunchKernelGGL((k_gradient<S, 2, 0>), dim3(g), dim3(BLK), 0, st, m, p, a); break;
    case 4: hipLaunchKernelGGL((k_gradient<S, 4, 0>), dim3(g), dim3(BLK), 0, st, m, p, a); break;
    default: hipLaunchKernelGGL((k_gradient<S, 3, 0>), dim3(g), dim3(BLK), 0, st, m, p, a); break;
    }
}

void launchGradient(const DevMesh& m, const PhysParams& p, const StageArgs& a, int S, hipStream_t st)
{
    if (S == 3) gradS<3>(m, p, a, st);
    else if (S == 4) gradS<4>(m, p, a, st);
    else gradS<6>(m, p, a, st);
}

template <int FLUX, int VISC, int ORDER, int LP, int MIX>
static void facesT(const DevMesh& m, const PhysParams& p, const StageArgs& a, hipStream_t st)
{
    const uint32_t nbi = nblk(m.Fi), nbb = nblk(m.Fb);
    if (nbi + nbb == 0) return;
    if (a.estate)
        hipLaunchKernelGGL((k_faces<FLUX, VISC, ORDER, LP, 1, MIX>), dim3(nbi + nbb), dim3(BLK), 0, st, m, p, a, nbi);
    else
        hipLaunchKernelGGL((k_faces<FLUX, VISC, ORDER, LP, 0, MIX>), dim3(nbi + nbb), dim3(BLK), 0, st, m, p, a, nbi);
}

template <int FLUX, int VISC>
static void facesO(const DevMesh& m, const PhysParams& p, const StageArgs& a, hipStream_t st)
{
    const bool lp = p.order > 1 && p.limited && p.lpThresh > 0;
    if (p.mixed) {
        if (p.order == 1) facesT<FLUX, VISC, 1, 0, 1>(m, p, a, st);
        else if (!(p.limited && a.limit)) facesT<FLUX, VISC, 2, 2, 1>(m, p, a, st);
        else facesT<FLUX, VISC, 2, 0, 1>(m, p, a, st);
        return;
    }
    if (p.order == 1) facesT<FLUX, VISC, 1, 0, 0>(m, p, a, st);
    else if (!(p.limited && a.limit)) facesT<FLUX, VISC, 2, 2, 0>(m, p, a, st);
    else if (lp) facesT<FLUX, VISC, 2, 1, 0>(m, p, a, st);
    else facesT<FLUX, VISC, 2, 0, 0>(m, p, a, st);
}

void launchFaces(const DevMesh& m, const PhysParams& p, const StageArgs& a, hipStream_t st)
{
    if (p.flux == 1) {
        if (p.viscous) facesO<1, 1>(m, p, a, st); else facesO<1, 0>(m, p, a, st);
    } else {
        if (p.viscous) facesO<0, 1>(m, p, a, st); else facesO<0, 0>(m, p, a, st);
    }
}

template <int S>
static void updS(const DevMesh& m, const PhysParams& p, const StageArgs& a, hipStream_t st)
{
    const uint32_t g = nblk(m.N);
    if (g == 0) return;
#define UPD(K) case K: hipLaunchKernelGGL((k_update<S, K>), dim3(g), dim3(BLK), 0, st, m, p, a); break;
    switch (a.kind) {
        UPD(SK_RESIDUAL) UPD(SK_EULER) UPD(SK_RK2_1) UPD(SK_RK2_2) UPD(SK_TVD_1) UPD(SK_TVD_2)
        UPD(SK_RK4_1) UPD(SK_RK4_2) UPD(SK_RK4_3) UPD(SK_RK4_4)
    }
#undef UPD
}

void launchUpdate(const DevMesh& m, const PhysParams& p, const StageArgs& a, int S, hipStream_t st)
{
    if (S == 3) updS<3>(m, p, a, st);
    else if (S == 4) updS<4>(m, p, a, st);
    else updS<6>(m, p, a, st);
    const uint32_t g = nblk(m.N);
    if (g && !a.deferReduce) hipLaunchKernelGGL(k_reduce_stage, dim3(1), dim3(1024), 0, st, a.partial, g, a.sc, a.checkActive);
}

void launchRmsChunks(const DiagArgs& d, hipStream_t st)
{
    const uint64_t nb = uint64_t(d.L) / 16;
    const uint32_t nc = uint32_t((nb + RMS_CHUNK_BLOCKS - 1) / RMS_CHUNK_BLOCKS);
    if (d.blockSums) {
        // full chunks, the partial chunk (its workgroup returns at once when there is none), the
        // force, the stage reductions
        const uint32_t nfull = uint32_t(nb / RMS_CHUNK_BLOCKS);
        hipLaunchKernelGGL(k_rms_blocks, dim3(nfull + 3), dim3(RFB), 0, st, d, nfull);
    } else {
        hipLaunchKernelGGL(k_rms_chunks, dim3(nc + 2), dim3(RBLK), 0, st, d);
    }
}

__global__ __launch_bounds__(BLK) void k_pending_force(ForceArgs f, DevScalars* sc)
{
    __shared__ double2 fsh[FORCE_LDS];
    pendingForce<BLK, FORCE_LDS>(f, sc, fsh);
}

void launchPendingForce(const ForceArgs& f, DevScalars* sc, hipStream_t st)
{
    if (f.n) hipLaunchKernelGGL(k_pending_force, dim3(1), dim3(BLK), 0, st, f, sc);
}

void launchFinalize(const DiagArgs& d, hipStream_t st)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_T), 0, st, d);
}

void launchInitScalars(DevScalars* sc, double dt, double t, double CFL, hipStream_t st)
{
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(1), 0, st, sc, dt, t, CFL);
}

// ---------------------------------------------------------------- batched launches
// One launch per kernel for many meshes.  Workgroup w takes item xcdBlock(w), so each
// XCD gets a contiguous run of items (items are listed mesh by mesh, cell blocks and tiles
// in order); the item names the mesh record and the block, tile or RMS workgroup, and the
// workgroup runs the solo kernel's code on it.
template <int S, int MODE, int MIX>
__global__ __launch_bounds__(BLK) void k_gradient_batch(const MemberArgs* __restrict__ mem, const uint2* __restrict__ items,
                                                        uint32_t parity, uint32_t stage)
{
    __shared__ GradLds<MODE> L;
    const uint2 it = items[xcdBlock(blockIdx.x, gridDim.x)];
    const MemberArgs& M = mem[it.x];
    const StageArgs& a = M.st[parity ^ M.parityFlip][stage];
    if (it.y == FORCE_ITEM) {
        pendingForce<BLK, FORCE_LDS>(a.force, a.sc, reinterpret_cast<double2*>(&L.sg[0][0]));
        return;
    }
    gradientBody<S, MODE, MIX>(M.m, M.p, a, it.y, L);
}

template <int S, int FLUX, int VISC, int ORDER, int LP, int MIX>
__global__ __launch_bounds__(BLK) void k_face_tile_batch(const MemberArgs* __restrict__ mem,
                                                         const uint2* __restrict__ items, uint32_t parity, uint32_t stage)
{
    extern __shared__ double fsd[];
    const uint2 it = items[xcdBlock(blockIdx.x, gridDim.x)];
    const MemberArgs& M = mem[it.x];
    faceTileBody<S, FLUX, VISC, ORDER, LP, MIX>(M.m, M.p, M.st[parity ^ M.parityFlip][stage], M.t, it.y, it.y, fsd);
}

template <int S, int MODE, int MIX>
__global__ __launch_bounds__(BLK) void k_border_grad_batch(const MemberArgs* __restrict__ mem,
                                                           const uint2* __restrict__ items, uint32_t parity, uint32_t stage)
{
    __shared__ double2 fstage[FORCE_LDS];
    const uint2 it = items[xcdBlock(blockIdx.x, gridDim.x)];
    const MemberArgs& M = mem[it.x];
    const StageArgs& a = M.st[parity ^ M.parityFlip][stage];
    if (it.y == FORCE_ITEM) {
        pendingForce<BLK, FORCE_LDS>(a.force, a.sc, fstage);
        return;
    }
    borderGradBody<S, MODE, MIX>(M.m, M.p, a, M.t, it.y);
}

template <int S, int FLUX, int VISC, int MODE, int MIX>
__global__ __launch_bounds__(BLK, 3) void k_grad_tile_batch(const MemberArgs* __restrict__ mem,
                                                         const uint2* __restrict__ items, uint32_t parity, uint32_t stage)
{
    extern __shared__ double glds[];
    const uint2 it = items[xcdBlock(blockIdx.x, gridDim.x)];
    const MemberArgs& M = mem[it.x];
    gradTileBody<S, FLUX, VISC, MODE, MIX>(M.m, M.p, M.st[parity ^ M.parityFlip][stage], M.t, it.y, it.y, glds);
}

__global__ __launch_bounds__(RFB) void k_rms_batch(const MemberArgs* __restrict__ mem, const uint2* __restrict__ items,
                                                   uint32_t parity)
{
    __shared__ double v[8][RMS_CHUNK_BLOCKS];
    const uint2 it = items[blockIdx.x];
    const MemberArgs& M = mem[it.x];
    rmsBlocksBody(M.d[parity ^ M.parityFlip], M.nRmsFull, it.y, v);
}

__global__ __launch_bounds__(FIN_T) void k_finalize_batch(const MemberArgs* __restrict__ mem,
                                                          const uint32_t* __restrict__ members, uint32_t parity)
{
    __shared__ FinLds L;
    const MemberArgs& M = mem[members[blockIdx.x]];
    finalizeBody(M.d[parity ^ M.parityFlip], L);
}

__global__ __launch_bounds__(BLK) void k_pending_force_batch(const MemberArgs* __restrict__ mem,
                                                             const uint32_t* __restrict__ members)
{
    __shared__ double2 fsh[FORCE_LDS];
    const MemberArgs& M = mem[members[blockIdx.x]];
    pendingForce<BLK, FORCE_LDS>(M.f, M.d[0].sc, fsh);
}

__global__ void k_gather_scalars(const MemberArgs* __restrict__ mem, const uint32_t* __restrict__ members, uint32_t n,
                                 DevScalars* __restrict__ out)
{
    // one thread per 8-byte word of a DevScalars record
    constexpr uint32_t W = sizeof(DevScalars) / 8;
    static_assert(sizeof(DevScalars) % 8 == 0, "DevScalars is copied in 8-byte words");
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * W) return;
    const uint32_t k = g / W, w = g % W;
    const uint64_t* src = reinterpret_cast<const uint64_t*>(mem[members[k]].d[0].sc);
    reinterpret_cast<uint64_t*>(out + k)[w] = src[w];
}

template <int S, int MIX>
static void gradBatchS(int mode, const MemberArgs* mem, const uint2* items, uint32_t n, uint32_t par, uint32_t st,
                       hipStream_t s)
{
    const dim3 g(n), b(BLK);
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_gradient_batch<S, 0, MIX>), g, b, 0, s, mem, items, par, st); break;
    case 1: hipLaunchKernelGGL((k_gradient_batch<S, 1, MIX>), g, b, 0, s, mem, items, par, st); break;
    case 4: hipLaunchKernelGGL((k_gradient_batch<S, 4, MIX>), g, b, 0, s, mem, items, par, st); break;
    default: hipLaunchKernelGGL((k_gradient_batch<S, 2, MIX>), g, b, 0, s, mem, items, par, st); break;
    }
}

template <int S, int MIX>
static void borderBatchS(int mode, const MemberArgs* mem, const uint2* items, uint32_t n, uint32_t par, uint32_t st,
                         hipStream_t s)
{
    const dim3 g(n), b(BLK);
    switch (mode) {
    case 1: hipLaunchKernelGGL((k_border_grad_batch<S, 1, MIX>), g, b, 0, s, mem, items, par, st); break;
    case 4: hipLaunchKernelGGL((k_border_grad_batch<S, 4, MIX>), g, b, 0, s, mem, items, par, st); break;
    default: hipLaunchKernelGGL((k_border_grad_batch<S, 2, MIX>), g, b, 0, s, mem, items, par, st); break;
    }
}

void launchGradientBatch(const GroupShape& gs, int limitStage, const MemberArgs* mem, const uint2* items,
                         uint32_t nItems, uint32_t parity, uint32_t stage, hipStream_t st)
{
    if (nItems == 0) return;
    if (gs.gradTiles) {   // border blocks; order 2, no LP vector limiter
        const int mode = !(gs.limited && limitStage) ? 1 : (gs.venkat ? 4 : 2);
        if (gs.mixed) {
            if (gs.S == 3) borderBatchS<3, 1>(mode, mem, items, nItems, parity, stage, st);
            else if (gs.S == 4) borderBatchS<4, 1>(mode, mem, items, nItems, parity, stage, st);
            else borderBatchS<6, 1>(mode, mem, items, nItems, parity, stage, st);
        } else {
            if (gs.S == 3) borderBatchS<3, 0>(mode, mem, items, nItems, parity, stage, st);
            else if (gs.S == 4) borderBatchS<4, 0>(mode, mem, items, nItems, parity, stage, st);
            else borderBatchS<6, 0>(mode, mem, items, nItems, parity, stage, st);
        }
        return;
    }
    // batched groups run without the LP vector limiter (mode 3)
    const int mode = gs.order == 1 ? 0 : (!(gs.limited && limitStage) ? 1 : (gs.venkat ? 4 : 2));
    if (gs.mixed) {
        if (gs.S == 3) gradBatchS<3, 1>(mode, mem, items, nItems, parity, stage, st);
        else if (gs.S == 4) gradBatchS<4, 1>(mode, mem, items, nItems, parity, stage, st);
        else gradBatchS<6, 1>(mode, mem, items, nItems, parity, stage, st);
    } else {
        if (gs.S == 3) gradBatchS<3, 0>(mode, mem, items, nItems, parity, stage, st);
        else if (gs.S == 4) gradBatchS<4, 0>(mode, mem, items, nItems, parity, stage, st);
        else gradBatchS<6, 0>(mode, mem, items, nItems, parity, stage, st);
    }
}

template <int S, int FLUX, int VISC, int MIX>
static void faceBatchO(const GroupShape& gs, int limitStage, const MemberArgs* mem, const uint2* items, uint32_t n,
                       uint32_t lds, uint32_t par, uint32_t stg, hipStream_t s)
{
    if (gs.gradTiles) {
        const dim3 g(n), b(BLK);
        const int mode = !(gs.limited && limitStage) ? 1 : (gs.venkat ? 4 : 2);
        if (mode == 1) hipLaunchKernelGGL((k_grad_tile_batch<S, FLUX, VISC, 1, MIX>), g, b, lds, s, mem, items, par, stg);
        else if (mode == 4) hipLaunchKernelGGL((k_grad_tile_batch<S, FLUX, VISC, 4, MIX>), g, b, lds, s, mem, items, par, stg);
        else hipLaunchKernelGGL((k_grad_tile_batch<S, FLUX, VISC, 2, MIX>), g, b, lds, s, mem, items, par, stg);
        return;
    }
    const dim3 g(n), b(BLK);
    if (gs.order == 1) hipLaunchKernelGGL((k_face_tile_batch<S, FLUX, VISC, 1, 0, MIX>), g, b, lds, s, mem, items, par, stg);
    else if (!(gs.limited && limitStage))
        hipLaunchKernelGGL((k_face_tile_batch<S, FLUX, VISC, 2, 2, MIX>), g, b, lds, s, mem, items, par, stg);
    else hipLaunchKernelGGL((k_face_tile_batch<S, FLUX, VISC, 2, 0, MIX>), g, b, lds, s, mem, items, par, stg);
}

template <int S, int MIX>
static void faceBatchS(const GroupShape& gs, int limitStage, const MemberArgs* mem, const uint2* items, uint32_t n,
                       uint32_t lds, uint32_t par, uint32_t stg, hipStream_t s)
{
    if (gs.flux == 1) {
        if (gs.viscous) faceBatchO<S, 1, 1, MIX>(gs, limitStage, mem, items, n, lds, par, stg, s);
        else faceBatchO<S, 1, 0, MIX>(gs, limitStage, mem, items, n, lds, par, stg, s);
    } else {
        if (gs.viscous) faceBatchO<S, 0, 1, MIX>(gs, limitStage, mem, items, n, lds, par, stg, s);
        else faceBatchO<S, 0, 0, MIX>(gs, limitStage, mem, items, n, lds, par, stg, s);
    }
}

void launchFaceTilesBatch(const GroupShape& gs, int limitStage, const MemberArgs* mem, const uint2* items,
                          uint32_t nItems, uint32_t ldsBytes, uint32_t parity, uint32_t stage, hipStream_t st)
{
    if (nItems == 0) return;
    if (gs.mixed) {
        if (gs.S == 3) faceBatchS<3, 1>(gs, limitStage, mem, items, nItems, ldsBytes, parity, stage, st);
        else if (gs.S == 4) faceBatchS<4, 1>(gs, limitStage, mem, items, nItems, ldsBytes, parity, stage, st);
        else faceBatchS<6, 1>(gs, limitStage, mem, items, nItems, ldsBytes, parity, stage, st);
    } else {
        if (gs.S == 3) faceBatchS<3, 0>(gs, limitStage, mem, items, nItems, ldsBytes, parity, stage, st);
        else if (gs.S == 4) faceBatchS<4, 0>(gs, limitStage, mem, items, nItems, ldsBytes, parity, stage, st);
        else faceBatchS<6, 0>(gs, limitStage, mem, items, nItems, ldsBytes, parity, stage, st);
    }
}

void launchRmsBatch(const MemberArgs* mem, const uint2* items, uint32_t nItems, uint32_t parity, hipStream_t st)
{
    if (nItems) hipLaunchKernelGGL(k_rms_batch, dim3(nItems), dim3(RFB), 0, st, mem, items, parity);
}

void launchFinalizeBatch(const MemberArgs* mem, const uint32_t* members, uint32_t n, uint32_t parity, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_finalize_batch, dim3(n), dim3(FIN_T), 0, st, mem, members, parity);
}

void launchPendingForceBatch(const MemberArgs* mem, const uint32_t* members, uint32_t n, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_pending_force_batch, dim3(n), dim3(BLK), 0, st, mem, members);
}

void launchGatherScalars(const MemberArgs* mem, const uint32_t* members, uint32_t n, DevScalars* out, hipStream_t st)
{
    const uint32_t words = n * uint32_t(sizeof(DevScalars) / 8);
    if (words) hipLaunchKernelGGL(k_gather_scalars, dim3((words + 255) / 256), dim3(256), 0, st, mem, members, n, out);
}

}  // namespace ebb
