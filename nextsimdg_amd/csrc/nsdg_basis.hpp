// Compile-time basis and quadrature tables for the gfx950 kernels.
//
// The same tables the reference generates offline into
// dynamics/src/include/codeGenerationDGinGauss.hpp and codeGenerationCGinGauss.hpp
// (generators: dynamics/codegeneration/basisfunctions.py:33-171, gaussquadrature.py:7-23),
// built here by constexpr evaluation so that every fully unrolled contraction in a
// kernel sees them as immediate constants (no memory traffic, no table lookups).
// Conventions: 2-D quadrature index q = gx + G*gy; CG local dof j = jx + (CG+1)*jy;
// DG basis order 1, x~, y~, x~^2-1/12, y~^2-1/12, x~y~, y~(x~^2-1/12), x~(y~^2-1/12).
#pragma once

namespace nsdg {

constexpr int cgdofs(int cg) { return cg == 1 ? 4 : 9; }
constexpr int cg2dgstress(int cg) { return cg == 1 ? 3 : 8; }
constexpr int gp1d(int d) { return (d == 8 || d == 6) ? 3 : (d == 3 ? 2 : 1); }
constexpr int edgedofs(int dg) { return dg == 1 ? 1 : (dg == 3 ? 2 : 3); }

// Gauss-Legendre rules on [0,1]: points 0.5 -+ sqrt(1/12); 0.5 -+ sqrt(3/20), 0.5;
// 0.5 -+ 0.5 sqrt(3/7 -+ 2/7 sqrt(6/5)), as exact double literals.
constexpr double gauss_point(int g, int k)
{
    if (g == 1)
        return 0.5;
    if (g == 2)
        return k == 0 ? 0x1.b0cb174df99c8p-3 : 0x1.93cd3a2c8198ep-1;
    if (g == 3)
        return k == 0 ? 0x1.cda042f0236e0p-4 : (k == 1 ? 0.5 : 0x1.c64bf7a1fb924p-1);
    return k == 0 ? 0x1.1c6490c2719ecp-4
                  : (k == 1 ? 0x1.51ee013116102p-2 : (k == 2 ? 0x1.5708ff6774f7fp-1 : 0x1.dc736de7b1cc2p-1));
}
constexpr double gauss_weight(int g, int k)
{
    if (g == 1)
        return 1.0;
    if (g == 2)
        return 0.5;
    if (g == 3)
        return k == 1 ? 8. / 18. : 5. / 18.;
    return (k == 0 || k == 3) ? 0x1.64340f7e7b66bp-3 : 0x1.4de5f840c24cbp-2;
}
constexpr double lagrange_point(int g, int k) { return g == 2 ? (k == 0 ? 0.0 : 1.0) : 0.5 * k; }

constexpr double dg_basis(int j, double x, double y)
{
    const double X = x - 0.5, Y = y - 0.5;
    return j == 0 ? 1.0
        : j == 1  ? X
        : j == 2  ? Y
        : j == 3  ? X * X - 1.0 / 12.0
        : j == 4  ? Y * Y - 1.0 / 12.0
        : j == 5  ? X * Y
        : j == 6  ? Y * (X * X - 1.0 / 12.0)
                  : X * (Y * Y - 1.0 / 12.0);
}
constexpr double dg_basis_dx(int j, double x, double y)
{
    const double X = x - 0.5, Y = y - 0.5;
    return j == 0 ? 0.0
        : j == 1  ? 1.0
        : j == 2  ? 0.0
        : j == 3  ? 2.0 * X
        : j == 4  ? 0.0
        : j == 5  ? Y
        : j == 6  ? Y * (2. * X)
                  : Y * Y - 1.0 / 12.0;
}
constexpr double dg_basis_dy(int j, double x, double y)
{
    const double X = x - 0.5, Y = y - 0.5;
    return j == 0 ? 0.0
        : j == 1  ? 0.0
        : j == 2  ? 1.0
        : j == 3  ? 0.0
        : j == 4  ? 2. * Y
        : j == 5  ? X
        : j == 6  ? X * X - 1.0 / 12.0
                  : X * (2. * Y);
}
constexpr double edge_basis(int j, double t)
{
    const double T = t - 0.5;
    return j == 0 ? 1.0 : (j == 1 ? T : T * T - 1.0 / 12.0);
}
constexpr double cg_1d(int cg, int j, double x)
{
    return cg == 1 ? (j == 0 ? 1.0 - x : x)
                   : (j == 0 ? 2.0 * (x - 0.5) * (x - 1.0) : (j == 1 ? 4.0 * x * (1.0 - x) : 2.0 * x * (x - 0.5)));
}
constexpr double cg_1d_dx(int cg, int j, double x)
{
    return cg == 1 ? (j == 0 ? -1.0 : 1.0) : (j == 0 ? 4.0 * x - 3.0 : (j == 1 ? 4.0 - 8.0 * x : 4.0 * x - 1.0));
}

template <int R, int C> struct Mat {
    double a[R][C];
    constexpr const double* operator[](int i) const { return a[i]; }
};

// GAUSSWEIGHTS<G>
template <int G> constexpr Mat<1, G * G> make_w()
{
    Mat<1, G * G> m {};
    for (int q = 0; q < G * G; ++q)
        m.a[0][q] = gauss_weight(G, q % G) * gauss_weight(G, q / G);
    return m;
}
// PSI<D,G>, PSIx<D,G>, PSIy<D,G>
template <int D, int G, int WHICH> constexpr Mat<D, G * G> make_psi()
{
    Mat<D, G * G> m {};
    for (int i = 0; i < D; ++i)
        for (int q = 0; q < G * G; ++q) {
            const double x = gauss_point(G, q % G), y = gauss_point(G, q / G);
            m.a[i][q] = WHICH == 0 ? dg_basis(i, x, y) : (WHICH == 1 ? dg_basis_dx(i, x, y) : dg_basis_dy(i, x, y));
        }
    return m;
}
// PSILagrange<D,G>
template <int D, int G> constexpr Mat<D, G * G> make_psil()
{
    Mat<D, G * G> m {};
    for (int i = 0; i < D; ++i)
        for (int q = 0; q < G * G; ++q)
            m.a[i][q] = dg_basis(i, lagrange_point(G, q % G), lagrange_point(G, q / G));
    return m;
}
// PSIe<E,G>
template <int E, int G> constexpr Mat<E, G> make_psie()
{
    Mat<E, G> m {};
    for (int i = 0; i < E; ++i)
        for (int q = 0; q < G; ++q)
            m.a[i][q] = edge_basis(i, gauss_point(G, q));
    return m;
}
// PSIe_w<D,G,side>: [q][i], side 0 bottom, 1 right, 2 top, 3 left
template <int D, int G, int SIDE> constexpr Mat<G, D> make_psiew()
{
    Mat<G, D> m {};
    for (int q = 0; q < G; ++q)
        for (int i = 0; i < D; ++i) {
            const double t = gauss_point(G, q);
            const double x = SIDE == 0 ? t : (SIDE == 1 ? 1.0 : (SIDE == 2 ? t : 0.0));
            const double y = SIDE == 0 ? 0.0 : (SIDE == 1 ? t : (SIDE == 2 ? 1.0 : t));
            m.a[q][i] = gauss_weight(G, q) * dg_basis(i, x, y);
        }
    return m;
}
// PHI<CG,G>, PHIx<CG,G>, PHIy<CG,G>
template <int CG, int G, int WHICH> constexpr Mat<cgdofs(CG), G * G> make_phi()
{
    Mat<cgdofs(CG), G * G> m {};
    for (int j = 0; j < cgdofs(CG); ++j)
        for (int q = 0; q < G * G; ++q) {
            const double x = gauss_point(G, q % G), y = gauss_point(G, q / G);
            const int jx = j % (CG + 1), jy = j / (CG + 1);
            m.a[j][q] = WHICH == 0 ? cg_1d(CG, jx, x) * cg_1d(CG, jy, y)
                : WHICH == 1       ? cg_1d_dx(CG, jx, x) * cg_1d(CG, jy, y)
                                   : cg_1d(CG, jx, x) * cg_1d_dx(CG, jy, y);
        }
    return m;
}

template <int G> struct W {
    static constexpr Mat<1, G * G> v = make_w<G>();
};
template <int D, int G> struct Psi {
    static constexpr Mat<D, G * G> v = make_psi<D, G, 0>();
    static constexpr Mat<D, G * G> x = make_psi<D, G, 1>();
    static constexpr Mat<D, G * G> y = make_psi<D, G, 2>();
};
template <int D, int G> struct PsiL {
    static constexpr Mat<D, G * G> v = make_psil<D, G>();
};
template <int E, int G> struct PsiE {
    static constexpr Mat<E, G> v = make_psie<E, G>();
};
template <int D, int G> struct PsiEW {
    static constexpr Mat<G, D> b = make_psiew<D, G, 0>();
    static constexpr Mat<G, D> r = make_psiew<D, G, 1>();
    static constexpr Mat<G, D> t = make_psiew<D, G, 2>();
    static constexpr Mat<G, D> l = make_psiew<D, G, 3>();
};
// Projection onto the DG stress basis on an affine element: M^-1 (PSI diag(w J)) with the
// element mass M = J diag(m), m_s = sum_q w_q PSI_s(q)^2 (the basis is orthogonal under the
// Gauss rule), so iMJwPSI[s][q] = PSI_s(q) w_q / m_s, independent of the element.
template <int S, int G> constexpr Mat<S, G * G> make_bproj()
{
    Mat<S, G * G> m {};
    for (int s = 0; s < S; ++s) {
        double ms = 0;
        for (int q = 0; q < G * G; ++q) {
            const double p = dg_basis(s, gauss_point(G, q % G), gauss_point(G, q / G));
            ms += gauss_weight(G, q % G) * gauss_weight(G, q / G) * p * p;
        }
        for (int q = 0; q < G * G; ++q)
            m.a[s][q] = dg_basis(s, gauss_point(G, q % G), gauss_point(G, q / G))
                * (gauss_weight(G, q % G) * gauss_weight(G, q / G)) / ms;
    }
    return m;
}
template <int S, int G> struct BProj {
    static constexpr Mat<S, G * G> v = make_bproj<S, G>();
};

template <int CG, int G> struct Phi {
    static constexpr Mat<cgdofs(CG), G * G> v = make_phi<CG, G, 0>();
    static constexpr Mat<cgdofs(CG), G * G> x = make_phi<CG, G, 1>();
    static constexpr Mat<cgdofs(CG), G * G> y = make_phi<CG, G, 2>();
};

} // namespace nsdg
