// TEST INFRASTRUCTURE — the FP64 CPU oracle for the neXtSIM_DG dynamics hot path.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
// and only as the checker. The product (nextsimdg_amd/) never links or calls it.
//
// Quadrature rules and basis functions in closed form. These restate the
// reference's offline code generators (read as text, not run):
//   dynamics/codegeneration/gaussquadrature.py:7-30   (Gauss-Legendre on [0,1], Lagrange points)
//   dynamics/codegeneration/basisfunctions.py:33-171  (DG, edge and CG1/CG2 Lagrange bases)
//   dynamics/codegeneration/codeGenerationDGinGauss.py:16-431, codeGenerationCGinGauss.py
// whose outputs are the committed tables in dynamics/src/include/codeGeneration*.hpp.
// Index conventions follow those tables: 2-D quadrature index q = gx + g*gy (x fastest),
// CG local dof j = jx + (cg+1)*jy, DG basis order 1, x~, y~, x~^2-1/12, y~^2-1/12, x~y~,
// y~(x~^2-1/12), x~(y~^2-1/12) with x~ = x-1/2.
#pragma once
#include <cmath>

namespace ora {

constexpr int cgdofs(int cg) { return cg == 1 ? 4 : 9; }
constexpr int cg2dgstress(int cg) { return cg == 1 ? 3 : 8; }
// GAUSSPOINTS1D in dynamics/src/include/NextsimDynamics.hpp:262
constexpr int gp1d(int d) { return (d == 8 || d == 6) ? 3 : (d == 3 ? 2 : 1); }
// EDGEDOFS in dynamics/src/include/NextsimDynamics.hpp:268
constexpr int edgedofs(int dg) { return dg == 1 ? 1 : (dg == 3 ? 2 : 3); }

struct Basis {
    double gp[5][4] = {};   // 1-D Gauss points on [0,1], rule with g points in gp[g]
    double gw[5][4] = {};   // 1-D Gauss weights
    double lp[4][3] = {};   // Lagrange points: lp[2] = {0,1}, lp[3] = {0,1/2,1}

    static double dg(int j, double x, double y)
    {
        const double X = x - 0.5, Y = y - 0.5;
        switch (j) {
        case 0: return 1.0;
        case 1: return X;
        case 2: return Y;
        case 3: return X * X - 1.0 / 12.0;
        case 4: return Y * Y - 1.0 / 12.0;
        case 5: return X * Y;
        case 6: return Y * (X * X - 1.0 / 12.0);
        default: return X * (Y * Y - 1.0 / 12.0);
        }
    }
    static double dgx(int j, double x, double y)
    {
        const double X = x - 0.5, Y = y - 0.5;
        switch (j) {
        case 0: return 0.0;
        case 1: return 1.0;
        case 2: return 0.0;
        case 3: return 2.0 * X;
        case 4: return 0.0;
        case 5: return Y;
        case 6: return Y * (2. * X);
        default: return Y * Y - 1.0 / 12.0;
        }
    }
    static double dgy(int j, double x, double y)
    {
        const double X = x - 0.5, Y = y - 0.5;
        switch (j) {
        case 0: return 0.0;
        case 1: return 0.0;
        case 2: return 1.0;
        case 3: return 0.0;
        case 4: return 2. * Y;
        case 5: return X;
        case 6: return X * X - 1.0 / 12.0;
        default: return X * (2. * Y);
        }
    }
    static double edge(int j, double t)
    {
        const double T = t - 0.5;
        return j == 0 ? 1.0 : (j == 1 ? T : T * T - 1.0 / 12.0);
    }
    static double cg1d(int cg, int j, double x)
    {
        if (cg == 1)
            return j == 0 ? 1.0 - x : x;
        if (j == 0)
            return 2.0 * (x - 0.5) * (x - 1.0);
        if (j == 1)
            return 4.0 * x * (1.0 - x);
        return 2.0 * x * (x - 0.5);
    }
    static double cg1d_dx(int cg, int j, double x)
    {
        if (cg == 1)
            return j == 0 ? -1.0 : 1.0;
        if (j == 0)
            return 4.0 * x - 3.0;
        if (j == 1)
            return 4.0 - 8.0 * x;
        return 4.0 * x - 1.0;
    }

    Basis()
    {
        gp[1][0] = 0.5;
        gw[1][0] = 1.0;
        gp[2][0] = 0.5 - std::sqrt(1. / 12.);
        gp[2][1] = 0.5 + std::sqrt(1. / 12.);
        gw[2][0] = gw[2][1] = 0.5;
        gp[3][0] = 0.5 - std::sqrt(3. / 20.);
        gp[3][1] = 0.5;
        gp[3][2] = 0.5 + std::sqrt(3. / 20.);
        gw[3][0] = 5. / 18.;
        gw[3][1] = 8. / 18.;
        gw[3][2] = 5. / 18.;
        const double a = 0.5 * std::sqrt(3. / 7. + 2. / 7. * std::sqrt(6. / 5.));
        const double b = 0.5 * std::sqrt(3. / 7. - 2. / 7. * std::sqrt(6. / 5.));
        gp[4][0] = 0.5 - a;
        gp[4][1] = 0.5 - b;
        gp[4][2] = 0.5 + b;
        gp[4][3] = 0.5 + a;
        gw[4][0] = gw[4][3] = (18. - std::sqrt(30.0)) / 72.;
        gw[4][1] = gw[4][2] = (18. + std::sqrt(30.0)) / 72.;
        lp[2][0] = 0.0;
        lp[2][1] = 1.0;
        lp[3][0] = 0.0;
        lp[3][1] = 0.5;
        lp[3][2] = 1.0;
    }

    // GAUSSWEIGHTS<g>[q]
    double W(int g, int q) const { return gw[g][q % g] * gw[g][q / g]; }
    // PSI<D,g>[i][q] and its reference gradients PSIx/PSIy
    double PSI(int g, int i, int q) const { return dg(i, gp[g][q % g], gp[g][q / g]); }
    double PSIx(int g, int i, int q) const { return dgx(i, gp[g][q % g], gp[g][q / g]); }
    double PSIy(int g, int i, int q) const { return dgy(i, gp[g][q % g], gp[g][q / g]); }
    // PSIe<E,g>[i][q]: edge basis at the g edge Gauss points
    double PSIe(int g, int i, int q) const { return edge(i, gp[g][q]); }
    // PSIe_w<D,g,e>[q][i]: weighted trace of cell basis i on edge e (0 bottom, 1 right, 2 top, 3 left)
    double PSIe_w(int g, int e, int q, int i) const
    {
        const double t = gp[g][q];
        double x, y;
        switch (e) {
        case 0: x = t; y = 0.0; break;
        case 1: x = 1.0; y = t; break;
        case 2: x = t; y = 1.0; break;
        default: x = 0.0; y = t; break;
        }
        return gw[g][q] * dg(i, x, y);
    }
    // PSILagrange<D,g>[i][q] at the g x g Lagrange points
    double PSIL(int g, int i, int q) const { return dg(i, lp[g][q % g], lp[g][q / g]); }
    // PHI<CG,g>[j][q] and gradients PHIx/PHIy
    double PHI(int cg, int g, int j, int q) const
    {
        return cg1d(cg, j % (cg + 1), gp[g][q % g]) * cg1d(cg, j / (cg + 1), gp[g][q / g]);
    }
    double PHIx(int cg, int g, int j, int q) const
    {
        return cg1d_dx(cg, j % (cg + 1), gp[g][q % g]) * cg1d(cg, j / (cg + 1), gp[g][q / g]);
    }
    double PHIy(int cg, int g, int j, int q) const
    {
        return cg1d(cg, j % (cg + 1), gp[g][q % g]) * cg1d_dx(cg, j / (cg + 1), gp[g][q / g]);
    }

    static const Basis& get()
    {
        static const Basis b;
        return b;
    }
};

} // namespace ora
