// TEST INFRASTRUCTURE — FP64 CPU oracle (see ora_basis.hpp header note).
//
// Geometry, parametric maps, DG transport, limiters, interpolations and the mEVP
// dynamics kernel, restated from the reference in plain C++17 + OpenMP (no Eigen).
// Loop order and the OpenMP parallel-for sites (including the two-colour row
// stripes of the scatter loops) follow the reference so that this file can also
// serve as the CPU baseline. DG vectors are row-major N_e x D (ModelArray layout,
// component c of element e at e*D + c); CG vectors have node id ix + (CG*nx+1)*iy.
#pragma once
#include "ora_basis.hpp"
#include "ora_mesh.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace ora {

// ------------------------------------------------------------------ cached tables
template <int D, int G> struct PsiT {
    double v[D][G * G], x[D][G * G], y[D][G * G];
    PsiT()
    {
        const Basis& b = Basis::get();
        for (int i = 0; i < D; ++i)
            for (int q = 0; q < G * G; ++q) {
                v[i][q] = b.PSI(G, i, q);
                x[i][q] = b.PSIx(G, i, q);
                y[i][q] = b.PSIy(G, i, q);
            }
    }
    static const PsiT& get()
    {
        static const PsiT t;
        return t;
    }
};
template <int CG, int G> struct PhiT {
    static constexpr int N = cgdofs(CG);
    double v[N][G * G], x[N][G * G], y[N][G * G];
    PhiT()
    {
        const Basis& b = Basis::get();
        for (int j = 0; j < N; ++j)
            for (int q = 0; q < G * G; ++q) {
                v[j][q] = b.PHI(CG, G, j, q);
                x[j][q] = b.PHIx(CG, G, j, q);
                y[j][q] = b.PHIy(CG, G, j, q);
            }
    }
    static const PhiT& get()
    {
        static const PhiT t;
        return t;
    }
};
template <int G> struct WT {
    double w[G * G];
    WT()
    {
        for (int q = 0; q < G * G; ++q)
            w[q] = Basis::get().W(G, q);
    }
    static const WT& get()
    {
        static const WT t;
        return t;
    }
};
template <int D, int G> struct PsiLT {
    double v[D][G * G];
    PsiLT()
    {
        for (int i = 0; i < D; ++i)
            for (int q = 0; q < G * G; ++q)
                v[i][q] = Basis::get().PSIL(G, i, q);
    }
    static const PsiLT& get()
    {
        static const PsiLT t;
        return t;
    }
};
template <int E, int G> struct PsiET {
    double v[E][G];
    PsiET()
    {
        for (int i = 0; i < E; ++i)
            for (int q = 0; q < G; ++q)
                v[i][q] = Basis::get().PSIe(G, i, q);
    }
    static const PsiET& get()
    {
        static const PsiET t;
        return t;
    }
};
template <int D, int G> struct PsiEWT {
    double v[4][G][D];
    PsiEWT()
    {
        for (int e = 0; e < 4; ++e)
            for (int q = 0; q < G; ++q)
                for (int i = 0; i < D; ++i)
                    v[e][q][i] = Basis::get().PSIe_w(G, e, q, i);
    }
    static const PsiEWT& get()
    {
        static const PsiEWT t;
        return t;
    }
};

// ------------------------------------------------------------------ geometry
// ParametricTools::dxT / dyT / J (dynamics/src/include/ParametricTools.hpp:73-105)
template <int G>
inline void mapGradients(const Mesh& m, size_t eid, double dxT[2][G * G], double dyT[2][G * G])
{
    double c[4][2];
    m.coordinatesOfElement(eid, c);
    const auto& P = PhiT<1, G>::get();
    for (int r = 0; r < 2; ++r)
        for (int q = 0; q < G * G; ++q) {
            double sx = 0, sy = 0;
            for (int k = 0; k < 4; ++k) {
                sx += c[k][r] * P.x[k][q];
                sy += c[k][r] * P.y[k][q];
            }
            dxT[r][q] = sx;
            dyT[r][q] = sy;
        }
}
template <int G> inline void jacobian(const Mesh& m, size_t eid, double J[G * G])
{
    double dxT[2][G * G], dyT[2][G * G];
    mapGradients<G>(m, eid, dxT, dyT);
    for (int q = 0; q < G * G; ++q)
        J[q] = dxT[0][q] * dyT[1][q] - dxT[1][q] * dyT[0][q];
}
// getGaussPointsInElement (ParametricTools.hpp:159-164)
template <int G> inline void gaussPoints(const Mesh& m, size_t eid, double gx[G * G], double gy[G * G])
{
    double c[4][2];
    m.coordinatesOfElement(eid, c);
    const auto& P = PhiT<1, G>::get();
    for (int q = 0; q < G * G; ++q) {
        double sx = 0, sy = 0;
        for (int k = 0; k < 4; ++k) {
            sx += c[k][0] * P.v[k][q];
            sy += c[k][1] * P.v[k][q];
        }
        gx[q] = sx;
        gy[q] = sy;
    }
}
// ParametricTools::massMatrix<D> (ParametricTools.hpp:109-153)
template <int D> inline void massMatrix(const Mesh& m, size_t eid, double M[D][D])
{
    constexpr int G = gp1d(D);
    double J[G * G];
    jacobian<G>(m, eid, J);
    const auto& P = PsiT<D, G>::get();
    const auto& W = WT<G>::get();
    double A[D][G * G];
    for (int i = 0; i < D; ++i)
        for (int q = 0; q < G * G; ++q)
            A[i][q] = P.v[i][q] * (W.w[q] * J[q]);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            double s = 0;
            for (int q = 0; q < G * G; ++q)
                s += A[i][q] * P.v[j][q];
            M[i][j] = s;
        }
}
// NextsimDynamics.hpp:34
constexpr double EarthRadius = 6371000.0;
// SphericalTools::massMatrix<D> (ParametricTools.hpp:189-202): the mass matrix with the
// metric weight cos(latitude) at the Gauss points
template <int D> inline void sphericalMassMatrix(const Mesh& m, size_t eid, double M[D][D])
{
    constexpr int G = gp1d(D);
    double J[G * G], gx[G * G], gy[G * G];
    jacobian<G>(m, eid, J);
    gaussPoints<G>(m, eid, gx, gy);
    const auto& P = PsiT<D, G>::get();
    const auto& W = WT<G>::get();
    double A[D][G * G];
    for (int i = 0; i < D; ++i)
        for (int q = 0; q < G * G; ++q)
            A[i][q] = P.v[i][q] * (W.w[q] * J[q] * std::cos(gy[q]));
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            double s = 0;
            for (int q = 0; q < G * G; ++q)
                s += A[i][q] * P.v[j][q];
            M[i][j] = s;
        }
}
// Dense inverse by Gauss-Jordan elimination with partial pivoting (the reference uses
// Eigen's .inverse(); results agree to rounding).
template <int N> inline void invert(const double A[N][N], double R[N][N])
{
    double a[N][2 * N];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            a[i][j] = A[i][j];
            a[i][N + j] = (i == j) ? 1.0 : 0.0;
        }
    for (int c = 0; c < N; ++c) {
        int p = c;
        for (int r = c + 1; r < N; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c]))
                p = r;
        if (a[p][c] == 0.0)
            throw std::runtime_error("singular element matrix");
        if (p != c)
            for (int j = 0; j < 2 * N; ++j)
                std::swap(a[p][j], a[c][j]);
        const double inv = 1.0 / a[c][c];
        for (int j = 0; j < 2 * N; ++j)
            a[c][j] *= inv;
        for (int r = 0; r < N; ++r) {
            if (r == c || a[r][c] == 0.0)
                continue;
            const double f = a[r][c];
            for (int j = 0; j < 2 * N; ++j)
                a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            R[i][j] = a[i][N + j];
}

// ------------------------------------------------------------------ DG edge traces
// leftedgeofcell / rightedgeofcell / bottomedgeofcell / topedgeofcell
// (dynamics/src/DGTransport.cpp:31-142): cell coefficients restricted to an edge,
// expressed in the edge basis 1, t-1/2, (t-1/2)^2-1/12.
enum Side { BOTTOM = 0, RIGHT = 1, TOP = 2, LEFT = 3 };
template <int DG> inline void trace(const double* c, int side, double* o)
{
    if constexpr (DG == 1) {
        o[0] = c[0];
    } else if constexpr (DG == 3) {
        switch (side) {
        case LEFT: o[0] = c[0] - 0.5 * c[1]; o[1] = c[2]; break;
        case RIGHT: o[0] = c[0] + 0.5 * c[1]; o[1] = c[2]; break;
        case BOTTOM: o[0] = c[0] - 0.5 * c[2]; o[1] = c[1]; break;
        default: o[0] = c[0] + 0.5 * c[2]; o[1] = c[1]; break;
        }
    } else if constexpr (DG == 6) {
        switch (side) {
        case LEFT: o[0] = c[0] - 0.5 * c[1] + 1. / 6. * c[3]; o[1] = c[2] - 0.5 * c[5]; o[2] = c[4]; break;
        case RIGHT: o[0] = c[0] + 0.5 * c[1] + 1. / 6. * c[3]; o[1] = c[2] + 0.5 * c[5]; o[2] = c[4]; break;
        case BOTTOM: o[0] = c[0] - 0.5 * c[2] + 1. / 6. * c[4]; o[1] = c[1] - 0.5 * c[5]; o[2] = c[3]; break;
        default: o[0] = c[0] + 0.5 * c[2] + 1. / 6. * c[4]; o[1] = c[1] + 0.5 * c[5]; o[2] = c[3]; break;
        }
    } else {
        switch (side) {
        case LEFT:
            o[0] = c[0] - 0.5 * c[1] + 1. / 6. * c[3];
            o[1] = c[2] - 0.5 * c[5] + 1. / 6. * c[6];
            o[2] = c[4] - 0.5 * c[7];
            break;
        case RIGHT:
            o[0] = c[0] + 0.5 * c[1] + 1. / 6. * c[3];
            o[1] = c[2] + 0.5 * c[5] + 1. / 6. * c[6];
            o[2] = c[4] + 0.5 * c[7];
            break;
        case BOTTOM:
            o[0] = c[0] - 0.5 * c[2] + 1. / 6. * c[4];
            o[1] = c[1] - 0.5 * c[5] + 1. / 6. * c[7];
            o[2] = c[3] - 0.5 * c[6];
            break;
        default:
            o[0] = c[0] + 0.5 * c[2] + 1. / 6. * c[4];
            o[1] = c[1] + 0.5 * c[5] + 1. / 6. * c[7];
            o[2] = c[3] + 0.5 * c[6];
            break;
        }
    }
}

// ------------------------------------------------------------------ DG transport
// DGTransport<DG> (dynamics/src/include/DGTransport.hpp:27-152, dynamics/src/DGTransport.cpp)
// with its ParametricTransportMap<DG> (dynamics/src/ParametricMap.cpp:13-87).
template <int DG> struct Transport {
    static constexpr int E = edgedofs(DG), G = gp1d(DG), Q = G * G;
    const Mesh& m;
    std::vector<std::array<double, DG * Q>> cellX, cellY;
    std::vector<std::array<double, DG * DG>> invM;
    std::vector<double> velx, vely, nvX, nvY, tmp1, tmp2, tmp3;
    int scheme = 2; // rk1, rk2 (model default, DynamicsKernel.hpp:61) or rk3

    explicit Transport(const Mesh& mesh)
        : m(mesh)
    {
        const size_t ne = m.nelements;
        velx.assign(ne * DG, 0.0);
        vely.assign(ne * DG, 0.0);
        tmp1.assign(ne * DG, 0.0);
        tmp2.assign(ne * DG, 0.0);
        tmp3.assign(ne * DG, 0.0);
        nvX.assign(m.nx * (m.ny + 1) * E, 0.0);
        nvY.assign((m.nx + 1) * m.ny * E, 0.0);
        cellX.resize(ne);
        cellY.resize(ne);
        invM.resize(ne);
        const auto& P = PsiT<DG, G>::get();
        const auto& W = WT<G>::get();
        // InitializeAdvectionCellTerms (ParametricMap.cpp:13-66), Cartesian
#pragma omp parallel for
        for (size_t eid = 0; eid < ne; ++eid) {
            double dxT[2][Q], dyT[2][Q];
            mapGradients<G>(m, eid, dxT, dyT);
            for (int r = 0; r < 2; ++r)
                for (int q = 0; q < Q; ++q) {
                    dxT[r][q] *= W.w[q];
                    dyT[r][q] *= W.w[q];
                }
            for (int i = 0; i < DG; ++i)
                for (int q = 0; q < Q; ++q) {
                    cellX[eid][i * Q + q] = P.x[i][q] * dyT[1][q] - P.y[i][q] * dxT[1][q];
                    cellY[eid][i * Q + q] = P.y[i][q] * dxT[0][q] - P.x[i][q] * dyT[0][q];
                }
            // InitializeInverseDGMassMatrix (ParametricMap.cpp:68-87); on the sphere the
            // cell terms are the Cartesian ones (ParametricMap.cpp:49-62 computes cos(lat)
            // but does not use it) and the inverse mass carries the metric and 1/R
            double M[DG][DG], Mi[DG][DG];
            if (m.spherical)
                sphericalMassMatrix<DG>(m, eid, M);
            else
                massMatrix<DG>(m, eid, M);
            invert<DG>(M, Mi);
            for (int i = 0; i < DG; ++i)
                for (int j = 0; j < DG; ++j)
                    invM[eid][i * DG + j] = m.spherical ? Mi[i][j] / EarthRadius : Mi[i][j];
        }
    }

    // reinitnormalvelocity (DGTransport.cpp:158-252)
    void reinitnormalvelocity()
    {
        std::fill(nvY.begin(), nvY.end(), 0.0);
        std::fill(nvX.begin(), nvX.end(), 0.0);
        const size_t nx = m.nx, ny = m.ny;
#pragma omp parallel for
        for (size_t iy = 0; iy < ny; ++iy) {
            size_t ey = iy * (nx + 1);
            size_t cy = iy * nx;
            for (size_t ix = 0; ix < nx; ++ix, ++ey, ++cy) {
                if (m.landmask[cy] == 0)
                    continue;
                double t[2], tx[E], ty[E];
                m.edgevector(ey, ey + nx + 1, t);
                trace<DG>(&velx[cy * DG], LEFT, tx);
                trace<DG>(&vely[cy * DG], LEFT, ty);
                for (int k = 0; k < E; ++k)
                    nvY[ey * E + k] += 0.5 * (t[1] * tx[k] - t[0] * ty[k]);
                m.edgevector(ey + 1, ey + nx + 2, t);
                trace<DG>(&velx[cy * DG], RIGHT, tx);
                trace<DG>(&vely[cy * DG], RIGHT, ty);
                for (int k = 0; k < E; ++k)
                    nvY[(ey + 1) * E + k] += 0.5 * (t[1] * tx[k] - t[0] * ty[k]);
            }
        }
#pragma omp parallel for
        for (size_t ix = 0; ix < nx; ++ix) {
            size_t cx = ix, n = ix;
            for (size_t iy = 0; iy < ny; ++iy, cx += nx, n += nx + 1) {
                if (m.landmask[cx] == 0)
                    continue;
                double t[2], tx[E], ty[E];
                m.edgevector(n, n + 1, t);
                trace<DG>(&velx[cx * DG], BOTTOM, tx);
                trace<DG>(&vely[cx * DG], BOTTOM, ty);
                for (int k = 0; k < E; ++k)
                    nvX[cx * E + k] += 0.5 * (-t[1] * tx[k] + t[0] * ty[k]);
                m.edgevector(n + nx + 1, n + nx + 2, t);
                trace<DG>(&velx[cx * DG], TOP, tx);
                trace<DG>(&vely[cx * DG], TOP, ty);
                for (int k = 0; k < E; ++k)
                    nvX[(cx + nx) * E + k] += 0.5 * (-t[1] * tx[k] + t[0] * ty[k]);
            }
        }
        for (int seg = 0; seg < 4; ++seg) {
            const auto& dl = m.dirichlet[seg];
#pragma omp parallel for
            for (size_t i = 0; i < dl.size(); ++i) {
                const size_t eid = dl[i], ix = eid % nx, iy = eid / nx;
                double* row;
                if (seg == 0)
                    row = &nvX[(nx * iy + ix) * E];
                else if (seg == 1)
                    row = &nvY[((nx + 1) * iy + ix + 1) * E];
                else if (seg == 2)
                    row = &nvX[(nx * (iy + 1) + ix) * E];
                else
                    row = &nvY[((nx + 1) * iy + ix) * E];
                for (int k = 0; k < E; ++k)
                    row[k] *= 2.0;
            }
        }
    }

    // cell_term (DGTransport.cpp:272-303)
    void cell_term(double dt, double* phiup, const double* phi, size_t eid) const
    {
        if constexpr (DG == 1)
            return;
        if (m.landmask[eid] == 0)
            return;
        const auto& P = PsiT<DG, G>::get();
        double vxg[Q], vyg[Q], pg[Q];
        for (int q = 0; q < Q; ++q) {
            double a = 0, b = 0, c = 0;
            for (int i = 0; i < DG; ++i) {
                a += velx[eid * DG + i] * P.v[i][q];
                b += vely[eid * DG + i] * P.v[i][q];
                c += phi[eid * DG + i] * P.v[i][q];
            }
            vxg[q] = a;
            vyg[q] = b;
            pg[q] = c;
        }
        for (int i = 0; i < DG; ++i) {
            double s = 0;
            for (int q = 0; q < Q; ++q)
                s += dt * (cellX[eid][i * Q + q] * vxg[q] + cellY[eid][i * Q + q] * vyg[q]) * pg[q];
            phiup[eid * DG + i] += s;
        }
    }
    // evaluate an edge vector at the edge Gauss points
    static void edgeEval(const double* ev, double* out)
    {
        const auto& PE = PsiET<E, G>::get();
        for (int q = 0; q < G; ++q) {
            double s = 0;
            for (int k = 0; k < E; ++k)
                s += ev[k] * PE.v[k][q];
            out[q] = s;
        }
    }
    // phiup(c) += sgn * dt * tmp * PSIe_w<side>
    static void edgeAdd(double* up, double sgn, double dt, const double* tmp, int side)
    {
        const auto& PW = PsiEWT<DG, G>::get();
        for (int i = 0; i < DG; ++i) {
            double s = 0;
            for (int q = 0; q < G; ++q)
                s += dt * tmp[q] * PW.v[side][q][i];
            up[i] += sgn * s;
        }
    }
    // edge_term_X / edge_term_Y (DGTransport.cpp:355-433)
    void edge_term(bool isX, double dt, double* phiup, const double* phi, size_t c1, size_t c2, size_t ie) const
    {
        if (m.landmask[c1] == 0 || m.landmask[c2] == 0)
            return;
        double vg[G], t1[E], t2[E], p1[G], p2[G], tmp[G];
        edgeEval(isX ? &nvX[ie * E] : &nvY[ie * E], vg);
        trace<DG>(&phi[c1 * DG], isX ? TOP : RIGHT, t1);
        trace<DG>(&phi[c2 * DG], isX ? BOTTOM : LEFT, t2);
        edgeEval(t1, p1);
        edgeEval(t2, p2);
        for (int q = 0; q < G; ++q)
            tmp[q] = std::max(vg[q], 0.0) * p1[q] + std::min(vg[q], 0.0) * p2[q];
        edgeAdd(&phiup[c1 * DG], -1.0, dt, tmp, isX ? TOP : RIGHT);
        edgeAdd(&phiup[c2 * DG], +1.0, dt, tmp, isX ? BOTTOM : LEFT);
    }
    // boundary_lower / boundary_upper / boundary_left / boundary_right (DGTransport.cpp:306-351).
    // Returns the outflow through the edge: the flux tested with the constant basis function,
    // i.e. what the cell's coefficient 0 loses before the inverse mass.
    double boundary(int side, double dt, double* phiup, const double* phi, size_t c, size_t e) const
    {
        const bool isX = (side == BOTTOM || side == TOP);
        double vg[G], tr[E], pg[G], tmp[G];
        edgeEval(isX ? &nvX[e * E] : &nvY[e * E], vg);
        trace<DG>(&phi[c * DG], side, tr);
        edgeEval(tr, pg);
        const double sgn = (side == BOTTOM || side == LEFT) ? -1.0 : 1.0;
        for (int q = 0; q < G; ++q)
            tmp[q] = pg[q] * std::max(sgn * vg[q], 0.0);
        edgeAdd(&phiup[c * DG], -1.0, dt, tmp, side);
        const auto& PW = PsiEWT<DG, G>::get();
        double o = 0;
        for (int q = 0; q < G; ++q)
            o += dt * tmp[q] * PW.v[side][q][0];
        return o;
    }
    // Outflow through the Dirichlet edges over the steps taken since it was last zeroed, in
    // the units of Tools::MeanValue: sum over RK stages of the stage weight times the
    // stage's outflow (times 1/R on the sphere, where the inverse mass carries it).
    double outflow = 0;
    // DGTransportOperator (DGTransport.cpp:435-512); *out0 receives the Dirichlet outflow
    void op(double dt, const double* phi, double* phiup, double* out0 = nullptr) const
    {
        const size_t nx = m.nx, ny = m.ny, ne = m.nelements;
        std::fill(phiup, phiup + ne * DG, 0.0);
#pragma omp parallel for
        for (size_t eid = 0; eid < ne; ++eid)
            cell_term(dt, phiup, phi, eid);
#pragma omp parallel for
        for (size_t iy = 0; iy < ny; ++iy) {
            size_t ic = iy * nx, ie = iy * (nx + 1) + 1;
            for (size_t i = 0; i + 1 < nx; ++i, ++ic, ++ie)
                edge_term(false, dt, phiup, phi, ic, ic + 1, ie);
        }
#pragma omp parallel for
        for (size_t ix = 0; ix < nx; ++ix) {
            size_t ic = ix, ie = ix + nx;
            for (size_t i = 0; i + 1 < ny; ++i, ic += nx, ie += nx)
                edge_term(true, dt, phiup, phi, ic, ic + nx, ie);
        }
        for (size_t pc = 0; pc < m.periodic.size(); ++pc) {
            const auto& seg = m.periodic[pc];
#pragma omp parallel for
            for (size_t i = 0; i < seg.size(); ++i) {
                if (seg[i][0] == 0)
                    edge_term(true, dt, phiup, phi, seg[i][1], seg[i][2], seg[i][3]);
                else
                    edge_term(false, dt, phiup, phi, seg[i][1], seg[i][2], seg[i][3]);
            }
        }
        double out = 0;
        for (int seg = 0; seg < 4; ++seg) {
            const auto& dl = m.dirichlet[seg];
#pragma omp parallel for reduction(+ : out)
            for (size_t i = 0; i < dl.size(); ++i) {
                const size_t eid = dl[i], ix = eid % nx, iy = eid / nx;
                if (seg == 0)
                    out += boundary(BOTTOM, dt, phiup, phi, eid, nx * iy + ix);
                else if (seg == 1)
                    out += boundary(RIGHT, dt, phiup, phi, eid, (nx + 1) * iy + ix + 1);
                else if (seg == 2)
                    out += boundary(TOP, dt, phiup, phi, eid, nx * (iy + 1) + ix);
                else
                    out += boundary(LEFT, dt, phiup, phi, eid, (nx + 1) * iy + ix);
            }
        }
        if (out0)
            *out0 = m.spherical ? out / EarthRadius : out;
#pragma omp parallel for
        for (size_t eid = 0; eid < ne; ++eid) {
            double r[DG];
            for (int i = 0; i < DG; ++i) {
                double s = 0;
                for (int j = 0; j < DG; ++j)
                    s += invM[eid][i * DG + j] * phiup[eid * DG + j];
                r[i] = s;
            }
            for (int i = 0; i < DG; ++i)
                phiup[eid * DG + i] = r[i];
        }
    }
    // step_rk1 / step_rk2 / step_rk3 (DGTransport.cpp:514-566)
    void step(double dt, double* phi)
    {
        const size_t n = m.nelements * DG;
        double o1 = 0, o2 = 0, o3 = 0;
        if (scheme == 1) {
            op(dt, phi, tmp1.data(), &o1);
            for (size_t i = 0; i < n; ++i)
                phi[i] += tmp1[i];
            outflow += o1;
        } else if (scheme == 2) {
            op(dt, phi, tmp1.data(), &o1);
            for (size_t i = 0; i < n; ++i)
                phi[i] += tmp1[i];
            op(dt, phi, tmp2.data(), &o2);
            for (size_t i = 0; i < n; ++i)
                phi[i] += 0.5 * (tmp2[i] - tmp1[i]);
            outflow += 0.5 * o1 + 0.5 * o2; // phi + (k1 + k2) / 2
        } else if (scheme == 3) {
            op(dt, phi, tmp1.data(), &o1);
            for (size_t i = 0; i < n; ++i)
                tmp1[i] += phi[i];
            op(dt, tmp1.data(), tmp2.data(), &o2);
            for (size_t i = 0; i < n; ++i) {
                tmp2[i] += tmp1[i];
                tmp2[i] *= 0.25;
                tmp2[i] += 0.75 * phi[i];
            }
            op(dt, tmp2.data(), tmp3.data(), &o3);
            for (size_t i = 0; i < n; ++i) {
                tmp3[i] += tmp2[i];
                phi[i] *= 1.0 / 3.0;
                phi[i] += 2.0 / 3.0 * tmp3[i];
            }
            outflow += (o1 + o2) / 6.0 + 2.0 / 3.0 * o3; // phi + (k1 + k2) / 6 + 2 k3 / 3
        } else
            throw std::runtime_error("unknown time stepping scheme");
    }
};

// ------------------------------------------------------------------ limiters
// LimitMax / LimitMin (dynamics/src/include/dgLimit.hpp:16-84)
template <int DG> inline void limitMax(double* dg, size_t n, double max)
{
    if constexpr (DG == 1) {
        for (size_t i = 0; i < n; ++i)
            dg[i] = std::min(dg[i], max);
    } else if constexpr (DG == 3) {
#pragma omp parallel for
        for (size_t i = 0; i < n; ++i) {
            double* c = dg + i * 3;
            c[0] = std::min(max, c[0]);
            const double l0 = 2.0 * std::max(std::fabs(c[1] + c[2]), std::fabs(c[1] - c[2]));
            if (l0 == 0)
                continue;
            if (c[0] + l0 - max > 0) {
                c[1] *= (max - c[0]) / l0;
                c[2] *= (max - c[0]) / l0;
            }
        }
    } else {
        static_assert(DG == 6, "LimitMax defined for DG 1, 3, 6");
        const auto& P = PsiT<6, 3>::get();
#pragma omp parallel for
        for (size_t i = 0; i < n; ++i) {
            double* c = dg + i * 6;
            c[0] = std::min(max, c[0]);
            double mv = -INFINITY;
            for (int q = 0; q < 9; ++q) {
                double s = 0;
                for (int k = 0; k < 6; ++k)
                    s += c[k] * P.v[k][q];
                mv = std::max(mv, s);
            }
            const double l0 = mv - c[0];
            if (mv > max)
                for (int k = 1; k < 6; ++k)
                    c[k] *= (max - c[0]) / l0;
        }
    }
}
template <int DG> inline void limitMin(double* dg, size_t n, double min)
{
    if constexpr (DG == 1) {
        for (size_t i = 0; i < n; ++i)
            dg[i] = std::max(dg[i], min);
    } else if constexpr (DG == 3) {
#pragma omp parallel for
        for (size_t i = 0; i < n; ++i) {
            double* c = dg + i * 3;
            c[0] = std::max(min, c[0]);
            const double l0 = 2.0 * std::max(std::fabs(c[1] + c[2]), std::fabs(c[1] - c[2]));
            if (l0 == 0)
                continue;
            if (c[0] - l0 - min < 0) {
                c[1] *= (c[0] - min) / l0;
                c[2] *= (c[0] - min) / l0;
            }
        }
    } else {
        static_assert(DG == 6, "LimitMin defined for DG 1, 3, 6");
        const auto& P = PsiT<6, 3>::get();
#pragma omp parallel for
        for (size_t i = 0; i < n; ++i) {
            double* c = dg + i * 6;
            c[0] = std::max(min, c[0]);
            double mv = INFINITY;
            for (int q = 0; q < 9; ++q) {
                double s = 0;
                for (int k = 0; k < 6; ++k)
                    s += c[k] * P.v[k][q];
                mv = std::min(mv, s);
            }
            const double l0 = mv - c[0];
            if (mv < min)
                for (int k = 1; k < 6; ++k)
                    c[k] *= (c[0] - min) / l0;
        }
    }
}

// ------------------------------------------------------------------ interpolations
// (dynamics/src/Interpolations.cpp)
using Fn = std::function<double(double, double)>;

// Function2DG (Interpolations.cpp:29-72); on the sphere both the moments and the mass
// matrix carry cos(latitude)
template <int DG> inline void function2DG(const Mesh& m, double* phi, const Fn& f)
{
    constexpr int G = gp1d(DG), Q = G * G;
    std::fill(phi, phi + m.nelements * DG, 0.0);
    const auto& P = PsiT<DG, G>::get();
    const auto& W = WT<G>::get();
#pragma omp parallel for
    for (size_t eid = 0; eid < m.nelements; ++eid) {
        if (!m.landmask[eid])
            continue;
        double gx[Q], gy[Q], J[Q], fv[Q], M[DG][DG], Mi[DG][DG], r[DG];
        gaussPoints<G>(m, eid, gx, gy);
        for (int q = 0; q < Q; ++q)
            fv[q] = f(gx[q], gy[q]);
        jacobian<G>(m, eid, J);
        if (m.spherical)
            sphericalMassMatrix<DG>(m, eid, M);
        else
            massMatrix<DG>(m, eid, M);
        invert<DG>(M, Mi);
        for (int i = 0; i < DG; ++i) {
            double s = 0;
            for (int q = 0; q < Q; ++q)
                s += P.v[i][q] * (m.spherical ? J[q] * W.w[q] * std::cos(gy[q]) : J[q] * W.w[q]) * fv[q];
            r[i] = s;
        }
        for (int i = 0; i < DG; ++i) {
            double s = 0;
            for (int j = 0; j < DG; ++j)
                s += Mi[i][j] * r[j];
            phi[eid * DG + i] = s;
        }
    }
}

template <int CG> inline void cgLocal(const double* cg, size_t cgi, size_t shift, double* l)
{
    if constexpr (CG == 1) {
        l[0] = cg[cgi];
        l[1] = cg[cgi + 1];
        l[2] = cg[cgi + shift];
        l[3] = cg[cgi + 1 + shift];
    } else {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                l[3 * r + c] = cg[cgi + c + r * shift];
    }
}

// CG2DG (Interpolations.cpp:74-122): L2 projection per element; the element mass matrix
// is rebuilt and inverted on every call, as in the reference. On the sphere both sides
// carry cos(latitude) (Interpolations.cpp:99-108).
template <int CG, int DG> inline void CG2DG(const Mesh& m, double* dg, const double* cg)
{
    constexpr int G = gp1d(DG), Q = G * G, N = cgdofs(CG);
    const size_t shift = CG * m.nx + 1;
    const auto& P = PsiT<DG, G>::get();
    const auto& F = PhiT<CG, G>::get();
    const auto& W = WT<G>::get();
#pragma omp parallel for
    for (size_t dgi = 0; dgi < m.nelements; ++dgi) {
        const size_t iy = dgi / m.nx, ix = dgi % m.nx;
        const size_t cgi = CG * shift * iy + CG * ix;
        double l[N], J[Q], t[Q], r[DG], M[DG][DG], Mi[DG][DG], gx[Q], gy[Q];
        cgLocal<CG>(cg, cgi, shift, l);
        jacobian<G>(m, dgi, J);
        if (m.spherical)
            gaussPoints<G>(m, dgi, gx, gy);
        for (int q = 0; q < Q; ++q) {
            double s = 0;
            for (int j = 0; j < N; ++j)
                s += F.v[j][q] * l[j];
            t[q] = m.spherical ? J[q] * W.w[q] * std::cos(gy[q]) * s : J[q] * W.w[q] * s;
        }
        for (int i = 0; i < DG; ++i) {
            double s = 0;
            for (int q = 0; q < Q; ++q)
                s += P.v[i][q] * t[q];
            r[i] = s;
        }
        if (m.spherical)
            sphericalMassMatrix<DG>(m, dgi, M);
        else
            massMatrix<DG>(m, dgi, M);
        invert<DG>(M, Mi);
        for (int i = 0; i < DG; ++i) {
            double s = 0;
            for (int j = 0; j < DG; ++j)
                s += Mi[i][j] * r[j];
            dg[dgi * DG + i] = s;
        }
    }
}

// DG2CG (Interpolations.cpp:129-204): Lagrange-point evaluation with nodal weights,
// two-colour stripes (odd rows first: note the inverted test at :194), then every
// domain-boundary node doubled (corners twice).
template <int CG, int DG> inline void DG2CG(const Mesh& m, double* cgv, const double* dg)
{
    const size_t nx = m.nx, ny = m.ny;
    const size_t row = CG * nx + 1, nn = row * (CG * ny + 1);
    std::fill(cgv, cgv + nn, 0.0);
    const auto& L = PsiLT<DG, CG + 1>::get();
    for (size_t p = 0; p < 2; ++p) {
#pragma omp parallel for
        for (size_t cy = 0; cy < ny; ++cy) {
            if (cy % 2 == p)
                continue;
            size_t c = cy * nx;
            for (size_t cx = 0; cx < nx; ++cx, ++c) {
                constexpr int NL = (CG + 1) * (CG + 1);
                double At[NL];
                for (int q = 0; q < NL; ++q) {
                    double s = 0;
                    for (int i = 0; i < DG; ++i)
                        s += dg[c * DG + i] * L.v[i][q];
                    At[q] = s;
                }
                if constexpr (CG == 1) {
                    const size_t cgi = row * cy + cx;
                    cgv[cgi] += 0.25 * At[0];
                    cgv[cgi + 1] += 0.25 * At[1];
                    cgv[cgi + row] += 0.25 * At[2];
                    cgv[cgi + row + 1] += 0.25 * At[3];
                } else {
                    const size_t cgi = 2 * row * cy + 2 * cx;
                    cgv[cgi] += 0.25 * At[0];
                    cgv[cgi + 1] += 0.5 * At[1];
                    cgv[cgi + 2] += 0.25 * At[2];
                    cgv[cgi + row] += 0.5 * At[3];
                    cgv[cgi + row + 1] += At[4];
                    cgv[cgi + row + 2] += 0.5 * At[5];
                    cgv[cgi + 2 * row] += 0.25 * At[6];
                    cgv[cgi + 2 * row + 1] += 0.5 * At[7];
                    cgv[cgi + 2 * row + 2] += 0.25 * At[8];
                }
            }
        }
    }
    const size_t upperLeft = CG * row * ny;
    for (size_t i = 0; i < CG * nx + 1; ++i) {
        cgv[i] *= 2.0;
        cgv[upperLeft + i] *= 2.0;
    }
    for (size_t i = 0; i < CG * ny + 1; ++i) {
        cgv[i * row] *= 2.0;
        cgv[i * row + CG * nx] *= 2.0;
    }
}

// L2ErrorFunctionDG (Interpolations.cpp:206-247), Cartesian
template <int DG> inline double L2ErrorFunctionDG(const Mesh& m, const double* src, const Fn& f)
{
    constexpr int G = gp1d(DG), Q = G * G;
    const auto& P = PsiT<DG, G>::get();
    const auto& W = WT<G>::get();
    double error = 0;
#pragma omp parallel for reduction(+ : error)
    for (size_t eid = 0; eid < m.nelements; ++eid) {
        if (!m.landmask[eid])
            continue;
        double gx[Q], gy[Q], J[Q];
        gaussPoints<G>(m, eid, gx, gy);
        jacobian<G>(m, eid, J);
        double s = 0;
        for (int q = 0; q < Q; ++q) {
            double v = 0;
            for (int i = 0; i < DG; ++i)
                v += src[eid * DG + i] * P.v[i][q];
            const double d = v - f(gx[q], gy[q]);
            s += J[q] * W.w[q] * (d * d);
        }
        error += s;
    }
    return error;
}

// Tools::MeanValue (dynamics/src/include/Tools.hpp:24-45): 3x3 Gauss regardless of DG;
// on spherical meshes each Gauss point is weighted by cos(latitude) as well (:28-36), so
// the result is the integral over the unit sphere (radians squared). meanValueCell is the
// contribution of element i with coefficients c.
template <int DG> inline double meanValueCell(const Mesh& m, size_t i, const double* c)
{
    const auto& P = PsiT<DG, 3>::get();
    const auto& W = WT<3>::get();
    double J[9], gx[9], gy[9];
    jacobian<3>(m, i, J);
    if (m.spherical)
        gaussPoints<3>(m, i, gx, gy);
    double s = 0;
    for (int q = 0; q < 9; ++q) {
        double v = 0;
        for (int k = 0; k < DG; ++k)
            v += c[k] * P.v[k][q];
        s += m.spherical ? v * std::cos(gy[q]) * (J[q] * W.w[q]) : v * (J[q] * W.w[q]);
    }
    return s;
}
template <int DG> inline double meanValue(const Mesh& m, const double* phi)
{
    double mass = 0;
    for (size_t i = 0; i < m.nelements; ++i)
        mass += meanValueCell<DG>(m, i, phi + i * DG);
    return mass;
}

// CGAveragePeriodic (dynamics/src/include/VectorManipulations.hpp:26-65)
template <int CG> inline void CGAveragePeriodic(const Mesh& m, double* v)
{
    const size_t row = CG * m.nx + 1;
    for (size_t seg = 0; seg < m.periodic.size(); ++seg)
        for (const auto& p : m.periodic[seg]) {
            const size_t eid_lb = p[2], eid_rt = p[1];
            const size_t i0_lb = row * CG * (eid_lb / m.nx) + CG * (eid_lb % m.nx);
            const size_t i0_rt = row * CG * (eid_rt / m.nx) + CG * (eid_rt % m.nx);
            if (p[0] == 0) {
                for (size_t j = 0; j <= CG; ++j) {
                    v[i0_lb + j] = 0.5 * (v[i0_lb + j] + v[i0_rt + CG * row + j]);
                    v[i0_rt + CG * row + j] = v[i0_lb + j];
                }
            } else {
                for (size_t j = 0; j <= CG; ++j) {
                    const size_t i1 = i0_lb + j * row, i2 = i0_rt + CG + j * row;
                    v[i1] = 0.5 * (v[i1] + v[i2]);
                    v[i2] = v[i1];
                }
            }
        }
}

// ------------------------------------------------------------------ parameters
// DynamicsParameters (DynamicsParameters.hpp:14-51) + VPParameters (VPParameters.hpp:15-25)
// + the hard-coded mEVP alpha, beta (VPCGDynamicsKernel.hpp:125-126,
// MEVPStressUpdateStep.hpp:126-127) and nSteps (DynamicsKernel.hpp:187).
struct Params {
    double rho_ice = 900.0, rho_atm = 1.3, rho_ocean = 1026.0;
    double C_atm = 1.2e-3, C_ocean = 5.5e-3;
    double F_atm = 1.2e-3 * 1.3, F_ocean = 5.5e-3 * 1026.0;
    double fc = 1.45842e-4, ocean_turning_angle = 0.0, gravity = 9.81;
    double Pstar = 27500.0, DeltaMin = 2.e-9;
    double alpha = 1500.0, beta = 1500.0;
    int nSteps = 100;
};

// Brittle rheology parameters: MEBParameters defaults (MEBParameters.hpp:38-70).
struct MEBParams {
    double compaction_param = -20.0, nu0 = 1.0 / 3.0, young = 5.96e8, P0 = 10.e3;
    double undamaged_time_relaxation_sigma = 1e7, exponent_relaxation_sigma = 5.0;
    double exponent_compression_factor = 1.5, tan_phi = 0.7, compr_strength = 1e10, C_lab = 2.0e6;
};
// Element size h = sqrt(area) with the quadrilateral area formula of ParametricMesh::area
// (ParametricMesh.hpp:276-299): 0.25 sqrt(4 e f - (b + d - a - c)^2) from the squared
// side lengths a (lower), b (right), c (top), d (left) and diagonals e, f.
inline double elementSize(const Mesh& m, size_t eid)
{
    // ParametricMesh::area reads the stored vertices as they are, without the longitude
    // correction of coordinatesOfElement
    const size_t nid = (eid / m.nx) * (m.nx + 1) + (eid % m.nx);
    const size_t ids[4] = { nid, nid + 1, nid + m.nx + 1, nid + m.nx + 2 }; // LL, LR, UL, UR
    auto sq = [&](int i, int j) {
        const double dx = m.vx[ids[i]] - m.vx[ids[j]], dy = m.vy[ids[i]] - m.vy[ids[j]];
        return dx * dx + dy * dy;
    };
    const double a = sq(0, 1), b = sq(1, 3), cc = sq(2, 3), d = sq(0, 2), e = sq(0, 3), f = sq(1, 2);
    return std::sqrt(0.25 * std::sqrt(4.0 * e * f - (b + d - a - cc) * (b + d - a - cc)));
}

// ------------------------------------------------------------------ momentum map
// ParametricMomentumMap<CG,DG> (dynamics/src/include/ParametricMap.hpp:62-131,
// dynamics/src/ParametricMap.cpp:94-356). On the sphere (x, y) = (longitude, latitude):
// the lumped masses carry cos(latitude), the longitude derivatives 1/R, the latitude
// derivatives cos(latitude)/R, and divM / iMM hold the metric terms of the divergence and
// the strain (ParametricMap.cpp:296-345).
template <int CG> struct MomentumMap {
    static constexpr int S = cg2dgstress(CG), G = gp1d(S), Q = G * G, N = cgdofs(CG);
    const Mesh& m;
    std::vector<double> lumped, lumped1;
    std::vector<std::array<double, N * S>> divS1, divS2, divM; // [j][s]
    std::vector<std::array<double, S * N>> iMgradX, iMgradY, iMM; // [s][j]
    std::vector<std::array<double, S * Q>> iMJwPSI;          // [s][q]
    std::vector<std::array<double, 16>> dX_SSH, dY_SSH;      // [j][k]

    explicit MomentumMap(const Mesh& mesh)
        : m(mesh)
    {
        initLumped();
        initDivS();
    }

    void initLumped()
    {
        const size_t nx = m.nx, ny = m.ny;
        constexpr int GG = (CG == 1 ? 1 : 4);
        const auto& F = PhiT<CG, GG>::get();
        const auto& W = WT<GG>::get();
        const size_t sy = CG * nx + 1;
        lumped.assign(sy * (CG * ny + 1), 0.0);
        for (size_t p = 0; p < 2; ++p) {
#pragma omp parallel for
            for (size_t iy = p; iy < ny; iy += 2)
                for (size_t ix = 0; ix < nx; ++ix) {
                    const size_t eid = nx * iy + ix;
                    double J[GG * GG], Me[N], gx[GG * GG], gy[GG * GG];
                    jacobian<GG>(m, eid, J);
                    if (m.spherical)
                        gaussPoints<GG>(m, eid, gx, gy);
                    for (int q = 0; q < GG * GG; ++q)
                        J[q] = m.spherical ? J[q] * W.w[q] * std::cos(gy[q]) : J[q] * W.w[q];
                    for (int j = 0; j < N; ++j) {
                        double s = 0;
                        for (int q = 0; q < GG * GG; ++q)
                            s += F.v[j][q] * J[q];
                        Me[j] = s;
                    }
                    const size_t n0 = CG * iy * sy + CG * ix;
                    for (int r = 0; r <= CG; ++r)
                        for (int c = 0; c <= CG; ++c)
                            lumped[n0 + c + r * sy] += Me[c + (CG + 1) * r];
                }
        }
        const auto& F1 = PhiT<1, 2>::get();
        const auto& W2 = WT<2>::get();
        const size_t s1 = nx + 1;
        lumped1.assign(s1 * (ny + 1), 0.0);
        for (size_t p = 0; p < 2; ++p) {
#pragma omp parallel for
            for (size_t iy = p; iy < ny; iy += 2)
                for (size_t ix = 0; ix < nx; ++ix) {
                    const size_t eid = nx * iy + ix;
                    double J[4], Me[4], gx[4], gy[4];
                    jacobian<2>(m, eid, J);
                    if (m.spherical)
                        gaussPoints<2>(m, eid, gx, gy);
                    for (int q = 0; q < 4; ++q)
                        J[q] = m.spherical ? J[q] * W2.w[q] * std::cos(gy[q]) : J[q] * W2.w[q];
                    for (int j = 0; j < 4; ++j) {
                        double s = 0;
                        for (int q = 0; q < 4; ++q)
                            s += F1.v[j][q] * J[q];
                        Me[j] = s;
                    }
                    const size_t n0 = iy * s1 + ix;
                    lumped1[n0] += Me[0];
                    lumped1[n0 + 1] += Me[1];
                    lumped1[n0 + s1] += Me[2];
                    lumped1[n0 + 1 + s1] += Me[3];
                }
        }
    }

    void initDivS()
    {
        const size_t ne = m.nelements;
        divS1.resize(ne);
        divS2.resize(ne);
        iMgradX.resize(ne);
        iMgradY.resize(ne);
        iMJwPSI.resize(ne);
        dX_SSH.resize(ne);
        dY_SSH.resize(ne);
        if (m.spherical) {
            divM.resize(ne);
            iMM.resize(ne);
        }
        const double iR = 1.0 / EarthRadius;
        const auto& P = PsiT<S, G>::get();
        const auto& F = PhiT<CG, G>::get();
        const auto& F1 = PhiT<1, G>::get();
        const auto& W = WT<G>::get();
#pragma omp parallel for
        for (size_t eid = 0; eid < ne; ++eid) {
            double Fx[2][Q], Fy[2][Q], J[Q];
            mapGradients<G>(m, eid, Fx, Fy);
            for (int r = 0; r < 2; ++r)
                for (int q = 0; q < Q; ++q) {
                    Fx[r][q] *= W.w[q];
                    Fy[r][q] *= W.w[q];
                }
            jacobian<G>(m, eid, J);
            double cl[Q], sl[Q], gx[Q], gy[Q];
            if (m.spherical) {
                gaussPoints<G>(m, eid, gx, gy);
                for (int q = 0; q < Q; ++q) {
                    cl[q] = std::cos(gy[q]);
                    sl[q] = std::sin(gy[q]);
                }
            }
            double dxc[N][Q], dyc[N][Q], dx1[4][Q], dy1[4][Q];
            for (int j = 0; j < N; ++j)
                for (int q = 0; q < Q; ++q) {
                    dxc[j][q] = F.x[j][q] * Fy[1][q] - F.y[j][q] * Fx[1][q];
                    dyc[j][q] = F.y[j][q] * Fx[0][q] - F.x[j][q] * Fy[0][q];
                }
            for (int j = 0; j < 4; ++j)
                for (int q = 0; q < Q; ++q) {
                    dx1[j][q] = F1.x[j][q] * Fy[1][q] - F1.y[j][q] * Fx[1][q];
                    dy1[j][q] = F1.y[j][q] * Fx[0][q] - F1.x[j][q] * Fy[0][q];
                }
            const bool sph = m.spherical;
            for (int j = 0; j < N; ++j)
                for (int s = 0; s < S; ++s) {
                    double a = 0, b = 0, c = 0;
                    for (int q = 0; q < Q; ++q) {
                        a += dxc[j][q] * P.v[s][q];
                        b += (sph ? dyc[j][q] * cl[q] : dyc[j][q]) * P.v[s][q];
                        if (sph)
                            c += F.v[j][q] * (J[q] * sl[q] * W.w[q]) * P.v[s][q];
                    }
                    divS1[eid][j * S + s] = sph ? a * iR : a;
                    divS2[eid][j * S + s] = sph ? b * iR : b;
                    if (sph)
                        divM[eid][j * S + s] = c * iR;
                }
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    double a = 0, b = 0;
                    for (int q = 0; q < Q; ++q) {
                        a += dx1[j][q] * F1.v[k][q];
                        b += (sph ? dy1[j][q] * cl[q] : dy1[j][q]) * F1.v[k][q];
                    }
                    dX_SSH[eid][j * 4 + k] = sph ? a * iR : a;
                    dY_SSH[eid][j * 4 + k] = sph ? b * iR : b;
                }
            double M[S][S], Mi[S][S];
            if (sph)
                sphericalMassMatrix<S>(m, eid, M);
            else
                massMatrix<S>(m, eid, M);
            invert<S>(M, Mi);
            for (int s = 0; s < S; ++s) {
                for (int j = 0; j < N; ++j) {
                    double a = 0, b = 0;
                    for (int t = 0; t < S; ++t) {
                        a += Mi[s][t] * divS1[eid][j * S + t];
                        b += Mi[s][t] * divS2[eid][j * S + t];
                    }
                    iMgradX[eid][s * N + j] = a;
                    iMgradY[eid][s * N + j] = b;
                    if (sph) {
                        double c = 0;
                        for (int t = 0; t < S; ++t)
                            c += Mi[s][t] * divM[eid][j * S + t];
                        iMM[eid][s * N + j] = c;
                    }
                }
                for (int q = 0; q < Q; ++q) {
                    double a = 0;
                    for (int t = 0; t < S; ++t)
                        a += Mi[s][t] * (P.v[t][q] * (W.w[q] * J[q]));
                    iMJwPSI[eid][s * Q + q] = a;
                }
            }
        }
    }
};

// ------------------------------------------------------------------ mEVP kernel
// MEVPDynamicsKernel<DG> = VPCGDynamicsKernel<DG> (VPCGDynamicsKernel.hpp:22-176) over
// CGDynamicsKernel<DG> (CGDynamicsKernel.cpp:23-449) over DynamicsKernel<DG,S>
// (DynamicsKernel.hpp:37-229) with MEVPStressUpdateStep (MEVPStressUpdateStep.hpp:30-118).
// Strain and stress start at zero (the reference leaves them uninitialised).
template <int CG, int DG> struct Kernel {
    static constexpr int S = cg2dgstress(CG), NG = gp1d(S), Q = NG * NG, N = cgdofs(CG);
    Params params;
    Mesh mesh;
    Transport<DG>* tr = nullptr;
    MomentumMap<CG>* pmap = nullptr;
    // Rheology: 0 mEVP (MEVPDynamicsKernel), 1 BBM (BBMDynamicsKernel,
    // BrittleCGDynamicsKernel.hpp + BBMStressUpdateStep.hpp).
    int rheology = 0;
    MEBParams meb;
    Transport<S>* stressTr = nullptr;
    std::vector<double> damage, avgU, avgV, hElem;
    std::vector<std::array<double, DG * Q>> iMJwPSIdam; // [d][q]
    std::vector<double> hice, cice, ssh, e11, e12, e22, s11, s12, s22;
    std::vector<double> u, v, u0, v0, cgH, cgA, gradX, gradY, dSx, dSy, uOcean, vOcean, uAtmos, vAtmos;
    std::map<std::string, std::vector<double>> extra;
    std::map<std::string, int> extraComp;
    double deltaT = 0;

    ~Kernel()
    {
        delete tr;
        delete pmap;
        delete stressTr;
    }
    size_t ncg() const { return (CG * mesh.nx + 1) * (CG * mesh.ny + 1); }

    // DynamicsKernel::initialise + CGDynamicsKernel::initialise, Cartesian
    void initialise(size_t nx, size_t ny, const double* coords, const double* mask, bool spherical = false)
    {
        mesh.fromCoordsAndMask(nx, ny, coords, mask, spherical);
        finishInit();
    }
    void finishInit()
    {
        delete tr;
        delete pmap;
        tr = new Transport<DG>(mesh);
        tr->scheme = 2;
        pmap = new MomentumMap<CG>(mesh);
        const size_t ne = mesh.nelements;
        hice.assign(ne * DG, 0.0);
        cice.assign(ne * DG, 0.0);
        ssh.assign(ne, 0.0);
        for (auto* x : { &e11, &e12, &e22, &s11, &s12, &s22 })
            x->assign(ne * S, 0.0);
        for (auto* x : { &u, &v, &u0, &v0, &cgH, &cgA, &gradX, &gradY, &dSx, &dSy, &uOcean, &vOcean, &uAtmos, &vAtmos })
            x->assign(ncg(), 0.0);
        if (rheology == 1)
            initBrittle();
    }
    // BrittleCGDynamicsKernel::initialise (BrittleCGDynamicsKernel.hpp:72-87) and the damage
    // projection iMJwPSI_dam (ParametricMap.cpp:290-296). The damage starts undamaged (1).
    void initBrittle()
    {
        const size_t ne = mesh.nelements;
        delete stressTr;
        stressTr = new Transport<S>(mesh);
        stressTr->scheme = 2;
        damage.assign(ne * DG, 0.0);
        for (size_t e = 0; e < ne; ++e)
            damage[e * DG] = 1.0;
        avgU.assign(ncg(), 0.0);
        avgV.assign(ncg(), 0.0);
        hElem.resize(ne);
        iMJwPSIdam.resize(ne);
        const auto& P = PsiT<DG, NG>::get();
        const auto& W = WT<NG>::get();
        for (size_t e = 0; e < ne; ++e) {
            hElem[e] = elementSize(mesh, e);
            double J[Q], M[DG][DG], Mi[DG][DG];
            jacobian<NG>(mesh, e, J);
            if (mesh.spherical)
                sphericalMassMatrix<DG>(mesh, e, M);
            else
                massMatrix<DG>(mesh, e, M);
            invert<DG>(M, Mi);
            for (int d = 0; d < DG; ++d)
                for (int q = 0; q < Q; ++q) {
                    double a = 0;
                    for (int k = 0; k < DG; ++k)
                        a += Mi[d][k] * P.v[k][q] * (W.w[q] * J[q]);
                    iMJwPSIdam[e][d * Q + q] = a;
                }
        }
    }

    // DGModelArray::ma2dg (DGModelArray.hpp:20-32): zero, then all components if the
    // source has D of them, else component 0 only.
    static void ma2dg(const double* data, int ncomp, int D, size_t ne, std::vector<double>& dg)
    {
        dg.assign(ne * D, 0.0);
        if (ncomp == D)
            std::memcpy(dg.data(), data, ne * D * sizeof(double));
        else
            for (size_t e = 0; e < ne; ++e)
                dg[e * D] = data[e];
    }
    // CGDynamicsKernel::setData (CGDynamicsKernel.cpp:53-90) and DynamicsKernel::setData
    void setData(const std::string& name, const double* data, int ncomp)
    {
        const size_t ne = mesh.nelements;
        std::vector<double>* cgt = nullptr;
        if (name == "u")
            cgt = &u;
        else if (name == "v")
            cgt = &v;
        else if (name == "uwind")
            cgt = &uAtmos;
        else if (name == "vwind")
            cgt = &vAtmos;
        else if (name == "uocean")
            cgt = &uOcean;
        else if (name == "vocean")
            cgt = &vOcean;
        if (cgt) {
            std::vector<double> tmp;
            ma2dg(data, ncomp, DG, ne, tmp);
            DG2CG<CG, DG>(mesh, cgt->data(), tmp.data());
        } else if (name == "hice")
            ma2dg(data, ncomp, DG, ne, hice);
        else if (name == "cice")
            ma2dg(data, ncomp, DG, ne, cice);
        else if (name == "ssh")
            ma2dg(data, ncomp, 1, ne, ssh);
        else if (name == "damage" && rheology == 1)
            ma2dg(data, ncomp, DG, ne, damage);
        else if (name == "s11" || name == "s12" || name == "s22" || name == "e11" || name == "e12" || name == "e22")
            ma2dg(data, ncomp, S, ne, stressRef(name));
        else {
            ma2dg(data, ncomp, ncomp, ne, extra[name]);
            extraComp[name] = ncomp;
        }
    }
    std::vector<double>& stressRef(const std::string& n)
    {
        if (n == "s11") return s11;
        if (n == "s12") return s12;
        if (n == "s22") return s22;
        if (n == "e11") return e11;
        if (n == "e12") return e12;
        return e22;
    }
    std::vector<double>* cgRef(const std::string& n)
    {
        if (n == "u") return &u;
        if (n == "v") return &v;
        if (n == "u0") return &u0;
        if (n == "v0") return &v0;
        if (n == "cgH") return &cgH;
        if (n == "cgA") return &cgA;
        if (n == "gradX") return &gradX;
        if (n == "gradY") return &gradY;
        if (n == "dStressX") return &dSx;
        if (n == "dStressY") return &dSy;
        if (n == "uocean") return &uOcean;
        if (n == "vocean") return &vOcean;
        if (n == "uwind") return &uAtmos;
        if (n == "vwind") return &vAtmos;
        if (n == "lumpedcgmass") return &pmap->lumped;
        if (n == "avgU" && rheology == 1) return &avgU;
        if (n == "avgV" && rheology == 1) return &avgV;
        return nullptr;
    }

    // VPCGDynamicsKernel::getIceOceanStress (VPCGDynamicsKernel.hpp:96-120)
    std::vector<double> iceOceanStress(bool x) const
    {
        std::vector<double> t(ncg());
        if (rheology == 1) {
            // BrittleCGDynamicsKernel::getIceOceanStress (BrittleCGDynamicsKernel.hpp:163-182):
            // from the averaged velocity, relative velocity ocean minus ice, turned.
            const double th = params.ocean_turning_angle * 0x1.1df46a2529d39p-6;
            const double ct = std::cos(th), st = std::sin(th);
            for (size_t i = 0; i < t.size(); ++i) {
                const double ur = uOcean[i] - avgU[i], vr = vOcean[i] - avgV[i];
                const double cp = params.F_ocean * std::hypot(ur, vr);
                t[i] = x ? cp * (ur * ct - vr * st) : cp * (vr * ct + ur * st);
            }
            return t;
        }
#pragma omp parallel for
        for (size_t i = 0; i < t.size(); ++i) {
            const double ur = u[i] - uOcean[i], vr = v[i] - vOcean[i];
            const double a = std::sqrt(ur * ur + vr * vr);
            t[i] = params.F_ocean * a * (x ? ur : vr);
        }
        return t;
    }
    // getDG0Data / getDGData (CGDynamicsKernel.cpp:92-118, DynamicsKernel.hpp:119-156)
    void getData(const std::string& name, double* out, int ncomp) const
    {
        const size_t ne = mesh.nelements;
        std::vector<double> tmp;
        const std::vector<double>* src = nullptr;
        int D = DG;
        if (name == "u" || name == "v" || name == "uiostress" || name == "viostress") {
            tmp.assign(ne * DG, 0.0);
            if (name == "u")
                CG2DG<CG, DG>(mesh, tmp.data(), u.data());
            else if (name == "v")
                CG2DG<CG, DG>(mesh, tmp.data(), v.data());
            else {
                auto t = iceOceanStress(name == "uiostress");
                CG2DG<CG, DG>(mesh, tmp.data(), t.data());
            }
            src = &tmp;
        } else if (name == "hice")
            src = &hice;
        else if (name == "cice")
            src = &cice;
        else if (name == "damage" && rheology == 1)
            src = &damage;
        else if (name == "ssh") {
            src = &ssh;
            D = 1;
        } else if (name == "s11" || name == "s12" || name == "s22" || name == "e11" || name == "e12" || name == "e22") {
            src = &const_cast<Kernel*>(this)->stressRef(name);
            D = S;
        } else {
            auto it = extra.find(name);
            if (it == extra.end())
                throw std::out_of_range("unknown field " + name);
            src = &it->second;
            D = extraComp.at(name);
        }
        if (ncomp == D)
            std::memcpy(out, src->data(), ne * D * sizeof(double));
        else
            for (size_t e = 0; e < ne; ++e)
                out[e] = (*src)[e * D];
    }

    // DynamicsKernel::advectionAndLimits (DynamicsKernel.hpp:160-172)
    void advectionAndLimits(double dt)
    {
        // the brittle kernels advect with the velocity averaged over the previous step's
        // substeps (BrittleCGDynamicsKernel.hpp:89-90)
        const std::vector<double>& ua = rheology == 1 ? avgU : u;
        const std::vector<double>& va = rheology == 1 ? avgV : v;
        CG2DG<CG, DG>(mesh, tr->velx.data(), ua.data());
        CG2DG<CG, DG>(mesh, tr->vely.data(), va.data());
        tr->reinitnormalvelocity();
        tr->outflow = 0;
        tr->step(dt, cice.data());
        budget[1] += tr->outflow;
        tr->outflow = 0;
        tr->step(dt, hice.data());
        budget[0] += tr->outflow;
        const std::vector<double> a0 = cice, h0 = hice;
        limitMax<DG>(cice.data(), mesh.nelements, 1.0);
        limitMin<DG>(cice.data(), mesh.nelements, 0.0);
        limitMin<DG>(hice.data(), mesh.nelements, 0.0);
        budget[2] += limiterChange(h0, hice);
        budget[3] += limiterChange(a0, cice);
    }
    // Mass budget of the advection (not part of the reference): Dirichlet outflow of h
    // and A, and the change the limiters make to their Tools::MeanValue, summed over the
    // steps since it was zeroed. Over a step, MeanValue(after) - MeanValue(before) equals
    // limiter change - outflow up to rounding: interior fluxes cancel pairwise.
    double budget[4] = { 0, 0, 0, 0 }; // outflow h, outflow A, limiter h, limiter A
    double limiterChange(const std::vector<double>& before, const std::vector<double>& after) const
    {
        double d = 0;
        for (size_t i = 0; i < mesh.nelements; ++i)
            if (!std::equal(&before[i * DG], &before[i * DG] + DG, &after[i * DG]))
                d += meanValueCell<DG>(mesh, i, &after[i * DG]) - meanValueCell<DG>(mesh, i, &before[i * DG]);
        return d;
    }

    // CGDynamicsKernel::ComputeGradientOfSeaSurfaceHeight (CGDynamicsKernel.cpp:125-258)
    void sshGradient()
    {
        const size_t nx = mesh.nx, ny = mesh.ny, row1 = nx + 1;
        std::vector<double> cgS(row1 * (ny + 1)), uG(cgS.size(), 0.0), vG(cgS.size(), 0.0);
        DG2CG<1, 1>(mesh, cgS.data(), ssh.data());
        for (size_t p = 0; p < 2; ++p) {
#pragma omp parallel for schedule(static)
            for (size_t cy = 0; cy < ny; ++cy) {
                if (cy % 2 != p)
                    continue;
                size_t eid = nx * cy, id = cy * row1;
                for (size_t cx = 0; cx < nx; ++cx, ++eid, ++id) {
                    const double l[4] = { cgS[id], cgS[id + 1], cgS[id + row1], cgS[id + row1 + 1] };
                    const size_t ids[4] = { id, id + 1, id + row1, id + row1 + 1 };
                    for (int j = 0; j < 4; ++j) {
                        double tx = 0, ty = 0;
                        for (int k = 0; k < 4; ++k) {
                            tx += pmap->dX_SSH[eid][j * 4 + k] * l[k];
                            ty += pmap->dY_SSH[eid][j * 4 + k] * l[k];
                        }
                        uG[ids[j]] -= tx;
                        vG[ids[j]] -= ty;
                    }
                }
            }
        }
        for (size_t i = 0; i < uG.size(); ++i) {
            uG[i] /= pmap->lumped1[i];
            vG[i] /= pmap->lumped1[i];
        }
        const size_t topleft = ny * row1;
        for (size_t i = 1; i < nx; ++i) {
            uG[i] = uG[i + row1];
            vG[i] = vG[i + row1];
            uG[topleft + i] = uG[topleft + i - row1];
            vG[topleft + i] = vG[topleft + i - row1];
        }
        for (size_t i = 1; i < ny; ++i) {
            uG[i * row1] = uG[i * row1 + 1];
            vG[i * row1] = vG[i * row1 + 1];
            uG[i * row1 + row1 - 1] = uG[i * row1 + row1 - 2];
            vG[i * row1 + row1 - 1] = vG[i * row1 + row1 - 2];
        }
        uG[0] = uG[row1 + 1];
        vG[0] = vG[row1 + 1];
        uG[nx] = uG[nx + row1 - 1];
        vG[nx] = vG[nx + row1 - 1];
        uG[ny * row1] = uG[(ny - 1) * row1 + 1];
        vG[ny * row1] = vG[(ny - 1) * row1 + 1];
        uG[(ny + 1) * row1 - 1] = uG[ny * row1 - 2];
        vG[(ny + 1) * row1 - 1] = vG[ny * row1 - 2];
        if constexpr (CG == 1) {
            gradX = uG;
            gradY = vG;
        } else {
            const size_t r2 = 2 * nx + 1;
            for (size_t iy = 0; iy <= ny; ++iy)
                for (size_t ix = 0; ix <= nx; ++ix) {
                    const size_t i1 = iy * row1 + ix, i2 = r2 * 2 * iy + 2 * ix;
                    gradX[i2] = uG[i1];
                    gradY[i2] = vG[i1];
                }
            for (size_t iy = 0; iy <= ny; ++iy)
                for (size_t ix = 0; ix < nx; ++ix) {
                    const size_t i1 = row1 * iy + ix, i2 = r2 * 2 * iy + 1 + 2 * ix;
                    gradX[i2] = 0.5 * (uG[i1] + uG[i1 + 1]);
                    gradY[i2] = 0.5 * (vG[i1] + vG[i1 + 1]);
                }
            for (size_t iy = 0; iy < ny; ++iy)
                for (size_t ix = 0; ix <= nx; ++ix) {
                    const size_t i1 = row1 * iy + ix, i2 = r2 * (2 * iy + 1) + 2 * ix;
                    gradX[i2] = 0.5 * (uG[i1] + uG[i1 + row1]);
                    gradY[i2] = 0.5 * (vG[i1] + vG[i1 + row1]);
                }
            for (size_t iy = 0; iy < ny; ++iy)
                for (size_t ix = 0; ix < nx; ++ix) {
                    const size_t i1 = row1 * iy + ix, i2 = r2 * (2 * iy + 1) + 1 + 2 * ix;
                    gradX[i2] = 0.25 * (uG[i1] + uG[i1 + 1] + uG[i1 + row1] + uG[i1 + row1 + 1]);
                    gradY[i2] = 0.25 * (vG[i1] + vG[i1 + 1] + vG[i1 + row1] + vG[i1 + row1 + 1]);
                }
        }
    }

    // CGDynamicsKernel::prepareIteration (CGDynamicsKernel.cpp:260-276)
    void prepareIteration()
    {
        DG2CG<CG, DG>(mesh, cgH.data(), hice.data());
        CGAveragePeriodic<CG>(mesh, cgH.data());
        DG2CG<CG, DG>(mesh, cgA.data(), cice.data());
        CGAveragePeriodic<CG>(mesh, cgA.data());
        sshGradient();
        for (size_t i = 0; i < cgA.size(); ++i) {
            cgA[i] = std::min(cgA[i], 1.0);
            cgA[i] = std::max(cgA[i], 1.e-4);
            cgH[i] = std::max(cgH[i], 1.e-4);
        }
    }

    // CGDynamicsKernel::projectVelocityToStrain (CGDynamicsKernel.cpp:300-337)
    void projectVelocityToStrain()
    {
        const size_t shift = CG * mesh.nx + 1;
#pragma omp parallel for
        for (size_t row = 0; row < mesh.ny; ++row) {
            size_t dgi = mesh.nx * row, cgi = CG * shift * row;
            for (size_t col = 0; col < mesh.nx; ++col, ++dgi, cgi += CG) {
                if (mesh.landmask[dgi] == 0)
                    continue;
                double lu[N], lv[N];
                cgLocal<CG>(u.data(), cgi, shift, lu);
                cgLocal<CG>(v.data(), cgi, shift, lv);
                const auto& gX = pmap->iMgradX[dgi];
                const auto& gY = pmap->iMgradY[dgi];
                for (int s = 0; s < S; ++s) {
                    double a = 0, b = 0, c = 0, d = 0;
                    for (int j = 0; j < N; ++j) {
                        a += gX[s * N + j] * lu[j];
                        b += gY[s * N + j] * lv[j];
                        c += gX[s * N + j] * lv[j];
                        d += gY[s * N + j] * lu[j];
                    }
                    e11[dgi * S + s] = a;
                    e22[dgi * S + s] = b;
                    e12[dgi * S + s] = 0.5 * (c + d);
                    if (mesh.spherical) { // CGDynamicsKernel.cpp:331-334
                        double mv = 0, mu = 0;
                        for (int j = 0; j < N; ++j) {
                            mv += pmap->iMM[dgi][s * N + j] * lv[j];
                            mu += pmap->iMM[dgi][s * N + j] * lu[j];
                        }
                        e11[dgi * S + s] -= mv;
                        e12[dgi * S + s] += 0.5 * mu;
                    }
                }
            }
        }
    }

    // BBMStressUpdateStep::stressUpdateHighOrder (BBMStressUpdateStep.hpp:33-197): at the
    // Gauss points, the elastic increment of the stress, the viscous relaxation factor, the
    // Mohr-Coulomb and compressive failure criterion, the damage and stress relaxation; then
    // the stress projected with iMJwPSI and the damage with iMJwPSI_dam. deltaT is the
    // substep.
    void bbmStressUpdate()
    {
        const MEBParams& P = meb;
        const auto& PA = PsiT<DG, NG>::get();
        const auto& PS = PsiT<S, NG>::get();
        const double dt = deltaT, nu = P.nu0;
#pragma omp parallel for
        for (size_t i = 0; i < mesh.nelements; ++i) {
            double hG[Q], aG[Q], dG[Q], e11G[Q], e12G[Q], e22G[Q], s11G[Q], s12G[Q], s22G[Q];
            for (int q = 0; q < Q; ++q) {
                double h = 0, a = 0, d = 0, x11 = 0, x12 = 0, x22 = 0, y11 = 0, y12 = 0, y22 = 0;
                for (int k = 0; k < DG; ++k) {
                    h += hice[i * DG + k] * PA.v[k][q];
                    a += cice[i * DG + k] * PA.v[k][q];
                    d += damage[i * DG + k] * PA.v[k][q];
                }
                for (int k = 0; k < S; ++k) {
                    x11 += e11[i * S + k] * PS.v[k][q];
                    x12 += e12[i * S + k] * PS.v[k][q];
                    x22 += e22[i * S + k] * PS.v[k][q];
                    y11 += s11[i * S + k] * PS.v[k][q];
                    y12 += s12[i * S + k] * PS.v[k][q];
                    y22 += s22[i * S + k] * PS.v[k][q];
                }
                hG[q] = std::max(h, 0.0);
                aG[q] = std::min(std::max(a, 0.0), 1.0);
                dG[q] = std::min(std::max(d, 1e-12), 1.0);
                e11G[q] = x11;
                e12G[q] = x12;
                e22G[q] = x22;
                s11G[q] = y11;
                s12G[q] = y12;
                s22G[q] = y22;
            }
            const double hE = hElem[i];
            const double scale = std::sqrt(0.1 / hE);
            for (int q = 0; q < Q; ++q) {
                const double sn0 = 0.5 * (s11G[q] + s22G[q]);
                const double expC = std::exp(P.compaction_param * (1.0 - aG[q]));
                const double tv = P.undamaged_time_relaxation_sigma
                    * std::pow(dG[q] * expC, P.exponent_relaxation_sigma - 1);
                const double Pmax = P.P0 * std::pow(hG[q], P.exponent_compression_factor + 1.) * expC;
                const double tildeP = sn0 < 0.0 ? std::min(-Pmax / sn0, 1.0) : 0.0;
                const double mult = tv / (tv + (1. - tildeP) * dt);
                const double elasticity = hG[q] * P.young * dG[q] * expC;
                const double Du = dt * elasticity / (1. - nu * nu);
                s11G[q] += Du * (e11G[q] + nu * e22G[q]);
                s22G[q] += Du * (nu * e11G[q] + e22G[q]);
                s12G[q] += Du * e12G[q] * (1. - nu);
                s11G[q] *= mult;
                s22G[q] *= mult;
                s12G[q] *= mult;
                const double sn = 0.5 * (s11G[q] + s22G[q]);
                const double tau = std::sqrt(0.25 * (s11G[q] - s22G[q]) * (s11G[q] - s22G[q]) + s12G[q] * s12G[q]);
                const double cohesion = P.C_lab * scale * hG[q];
                const double cstr = P.compr_strength * scale * hG[q];
                double dcrit = (tau + P.tan_phi * sn > 0.) ? cohesion / (tau + P.tan_phi * sn) : 1.;
                dcrit = (sn < -cstr) ? -cstr / sn : dcrit;
                dcrit = std::min(dcrit, 1.0);
                const double td = hE * std::sqrt(2. * (1. + nu) * params.rho_ice) / std::sqrt(elasticity);
                dG[q] -= dG[q] * (1. - dcrit) * dt / td;
                s11G[q] -= s11G[q] * (1. - dcrit) * dt / td;
                s12G[q] -= s12G[q] * (1. - dcrit) * dt / td;
                s22G[q] -= s22G[q] * (1. - dcrit) * dt / td;
            }
            const auto& B = pmap->iMJwPSI[i];
            for (int k = 0; k < S; ++k) {
                double a = 0, b = 0, c = 0;
                for (int q = 0; q < Q; ++q) {
                    a += B[k * Q + q] * s11G[q];
                    b += B[k * Q + q] * s12G[q];
                    c += B[k * Q + q] * s22G[q];
                }
                s11[i * S + k] = a;
                s12[i * S + k] = b;
                s22[i * S + k] = c;
            }
            for (int k = 0; k < DG; ++k) {
                double a = 0;
                for (int q = 0; q < Q; ++q)
                    a += iMJwPSIdam[i][k * Q + q] * dG[q];
                damage[i * DG + k] = a;
            }
        }
    }

    // BrittleCGDynamicsKernel::updateMomentum (BrittleCGDynamicsKernel.hpp:200-253): the
    // substep's semi-implicit update with the Coriolis and ocean-turning rotation, written
    // as the reference has it (including the sign of the beta (gradX + tauX) term of v),
    // and the accumulation of the step-averaged velocity.
    void brittleMomentum()
    {
        const Params& p = params;
        const double th = p.ocean_turning_angle * 0x1.1df46a2529d39p-6;
        const double ct = std::cos(th), st = std::sin(th);
        const auto& ML = pmap->lumped;
        const double ns = static_cast<double>(p.nSteps);
#pragma omp parallel for
        for (size_t i = 0; i < u.size(); ++i) {
            const double dtm = deltaT / (p.rho_ice * cgH[i]);
            const double uI = u[i], vI = v[i];
            const double cP = cgA[i] * p.F_ocean * std::hypot(uOcean[i] - uI, vOcean[i] - vI);
            const double al = 1 + dtm * (cP * ct + 0.0);
            const double be = deltaT * p.fc + dtm * cP * st;
            const double rD = 1 / (al * al + be * be);
            const double dA = cgA[i] * p.F_atm * std::hypot(uAtmos[i], vAtmos[i]);
            const double tX = dA * uAtmos[i] + cP * (uOcean[i] * ct - vOcean[i] * st);
            const double tY = dA * vAtmos[i] + cP * (vOcean[i] * ct + uOcean[i] * st);
            const double gX = dSx[i] / ML[i] - p.rho_ice * cgH[i] * p.gravity * gradX[i];
            const double gY = dSy[i] / ML[i] - p.rho_ice * cgH[i] * p.gravity * gradY[i];
            double un = al * uI + be * vI + dtm * (al * (gX + tX) + be * (gY + tY));
            un *= rD;
            double vn = al * vI - be * uI + dtm * (al * (gY + tY) + be * (gX + tX));
            vn *= rD;
            u[i] = un;
            v[i] = vn;
            avgU[i] += un / ns;
            avgV[i] += vn / ns;
        }
    }

    // BrittleCGDynamicsKernel::update (BrittleCGDynamicsKernel.hpp:92-134)
    void updateBBM(double dt)
    {
        const size_t ne = mesh.nelements;
        advectionAndLimits(dt);
        CG2DG<CG, S>(mesh, stressTr->velx.data(), avgU.data());
        CG2DG<CG, S>(mesh, stressTr->vely.data(), avgV.data());
        stressTr->reinitnormalvelocity();
        stressTr->step(dt, s11.data());
        stressTr->step(dt, s12.data());
        stressTr->step(dt, s22.data());
        tr->step(dt, damage.data());
        limitMax<DG>(damage.data(), ne, 1.0);
        limitMin<DG>(damage.data(), ne, 1e-12);
        prepareIteration();
        deltaT = dt / params.nSteps;
        std::fill(avgU.begin(), avgU.end(), 0.0);
        std::fill(avgV.begin(), avgV.end(), 0.0);
        for (int k = 0; k < params.nSteps; ++k) {
            projectVelocityToStrain();
            bbmStressUpdate();
            stressDivergence();
            brittleMomentum();
            applyBoundaries();
        }
    }

    // MEVPStressUpdateStep::stressUpdateHighOrder (MEVPStressUpdateStep.hpp:30-118)
    void stressUpdate()
    {
        if (rheology == 1) {
            bbmStressUpdate();
            return;
        }
        const auto& PA = PsiT<DG, NG>::get();
        const auto& PS = PsiT<S, NG>::get();
        const double alpha = params.alpha;
#pragma omp parallel for
        for (size_t i = 0; i < mesh.nelements; ++i) {
            double hg[Q], ag[Q], g11[Q], g12[Q], g22[Q];
            for (int q = 0; q < Q; ++q) {
                double h = 0, a = 0;
                for (int k = 0; k < DG; ++k) {
                    h += hice[i * DG + k] * PA.v[k][q];
                    a += cice[i * DG + k] * PA.v[k][q];
                }
                hg[q] = std::max(h, 0.0);
                ag[q] = std::min(std::max(a, 0.0), 1.0);
                double x = 0, y = 0, z = 0;
                for (int k = 0; k < S; ++k) {
                    x += e11[i * S + k] * PS.v[k][q];
                    y += e12[i * S + k] * PS.v[k][q];
                    z += e22[i * S + k] * PS.v[k][q];
                }
                g11[q] = x;
                g12[q] = y;
                g22[q] = z;
            }
            double r11[Q], r12[Q], r22[Q];
            for (int q = 0; q < Q; ++q) {
                const double DELTA = std::sqrt(params.DeltaMin * params.DeltaMin
                    + 1.25 * (g11[q] * g11[q] + g22[q] * g22[q]) + 1.50 * g11[q] * g22[q] + g12[q] * g12[q]);
                const double P = params.Pstar * hg[q] * std::exp(-20.0 * (1.0 - ag[q]));
                r11[q] = 1.0 / alpha * (P / 8.0 / DELTA * (5.0 * g11[q] + 3.0 * g22[q]) - 0.5 * P);
                r12[q] = 1.0 / alpha * (P / 4.0 / DELTA * g12[q]);
                r22[q] = 1.0 / alpha * (P / 8.0 / DELTA * (5.0 * g22[q] + 3.0 * g11[q]) - 0.5 * P);
            }
            const auto& B = pmap->iMJwPSI[i];
            for (int s = 0; s < S; ++s) {
                s11[i * S + s] *= (1.0 - 1.0 / alpha);
                s12[i * S + s] *= (1.0 - 1.0 / alpha);
                s22[i * S + s] *= (1.0 - 1.0 / alpha);
                double a = 0, b = 0, c = 0;
                for (int q = 0; q < Q; ++q) {
                    a += B[s * Q + q] * r11[q];
                    b += B[s * Q + q] * r12[q];
                    c += B[s * Q + q] * r22[q];
                }
                s11[i * S + s] += a;
                s12[i * S + s] += b;
                s22[i * S + s] += c;
            }
        }
    }

    // CGDynamicsKernel::dirichletZero (CGDynamicsKernel.cpp:400-437)
    void dirichletZero(std::vector<double>& w) const
    {
        const size_t nx = mesh.nx, row = CG * nx + 1;
        for (int seg = 0; seg < 4; ++seg) {
            const auto& dl = mesh.dirichlet[seg];
#pragma omp parallel for
            for (size_t i = 0; i < dl.size(); ++i) {
                const size_t eid = dl[i], ix = eid % nx, iy = eid / nx;
                for (size_t j = 0; j < CG + 1; ++j) {
                    if (seg == 0)
                        w[iy * CG * row + CG * ix + j] = 0.0;
                    else if (seg == 1)
                        w[iy * CG * row + CG * ix + CG + row * j] = 0.0;
                    else if (seg == 2)
                        w[(iy + 1) * CG * row + CG * ix + j] = 0.0;
                    else
                        w[iy * CG * row + CG * ix + row * j] = 0.0;
                }
            }
        }
    }

    // CGDynamicsKernel::stressDivergence + addStressTensorCell (CGDynamicsKernel.cpp:339-398)
    void stressDivergence()
    {
        std::fill(dSx.begin(), dSx.end(), 0.0);
        std::fill(dSy.begin(), dSy.end(), 0.0);
        const size_t nx = mesh.nx, row = CG * nx + 1;
        for (size_t p = 0; p < 2; ++p) {
#pragma omp parallel for schedule(static)
            for (size_t cy = 0; cy < mesh.ny; ++cy) {
                if (cy % 2 != p)
                    continue;
                size_t c = nx * cy;
                for (size_t cx = 0; cx < nx; ++cx, ++c) {
                    if (mesh.landmask[c] != 1)
                        continue;
                    const auto& A = pmap->divS1[c];
                    const auto& B = pmap->divS2[c];
                    const size_t cg_i = CG * row * cy + CG * cx;
                    for (size_t r = 0; r <= CG; ++r)
                        for (size_t cc = 0; cc <= CG; ++cc) {
                            const int j = cc + (CG + 1) * r;
                            double a = 0, b = 0, x = 0, y = 0;
                            for (int s = 0; s < S; ++s) {
                                a += A[j * S + s] * s11[c * S + s];
                                b += B[j * S + s] * s12[c * S + s];
                                x += A[j * S + s] * s12[c * S + s];
                                y += B[j * S + s] * s22[c * S + s];
                            }
                            double tx = a + b, ty = x + y;
                            if (mesh.spherical) { // CGDynamicsKernel.cpp:348-351
                                double m12 = 0, m11 = 0;
                                for (int s = 0; s < S; ++s) {
                                    m12 += pmap->divM[c][j * S + s] * s12[c * S + s];
                                    m11 += pmap->divM[c][j * S + s] * s11[c * S + s];
                                }
                                tx += m12;
                                ty -= m11;
                            }
                            dSx[cg_i + cc + r * row] -= tx;
                            dSy[cg_i + cc + r * row] -= ty;
                        }
                }
            }
        }
        dirichletZero(dSx);
        dirichletZero(dSy);
        CGAveragePeriodic<CG>(mesh, dSx.data());
        CGAveragePeriodic<CG>(mesh, dSy.data());
    }

    // VPCGDynamicsKernel::updateMomentum (VPCGDynamicsKernel.hpp:132-172). Note the
    // same-component Coriolis terms (-rho h fc u for u, +rho h fc v for v).
    void updateMomentum()
    {
        if (rheology == 1) {
            brittleMomentum();
            return;
        }
        const Params& p = params;
        const double beta = p.beta, SC = 1.0;
        const auto& ML = pmap->lumped;
#pragma omp parallel for
        for (size_t i = 0; i < u.size(); ++i) {
            const double uOcnRel = uOcean[i] - u[i];
            const double vOcnRel = v[i] - vOcean[i];
            const double absatm = std::sqrt(uAtmos[i] * uAtmos[i] + vAtmos[i] * vAtmos[i]);
            const double absocn = std::sqrt(uOcnRel * uOcnRel + vOcnRel * vOcnRel);
            const double ui = u[i], vi = v[i];
            u[i] = (1.0 / (p.rho_ice * cgH[i] / deltaT * (1.0 + beta) + cgA[i] * p.F_ocean * absocn)
                * (p.rho_ice * cgH[i] / deltaT * (beta * ui + u0[i])
                    + cgA[i] * (p.F_atm * absatm * uAtmos[i] + p.F_ocean * absocn * SC * uOcean[i])
                    - p.rho_ice * cgH[i] * p.fc * ui - p.rho_ice * cgH[i] * p.gravity * gradX[i]
                    + dSx[i] / ML[i]));
            v[i] = (1.0 / (p.rho_ice * cgH[i] / deltaT * (1.0 + beta) + cgA[i] * p.F_ocean * absocn)
                * (p.rho_ice * cgH[i] / deltaT * (beta * vi + v0[i])
                    + cgA[i] * (p.F_atm * absatm * vAtmos[i] + p.F_ocean * absocn * SC * vOcean[i])
                    + p.rho_ice * cgH[i] * p.fc * vi - p.rho_ice * cgH[i] * p.gravity * gradY[i]
                    + dSy[i] / ML[i]));
        }
    }
    void applyBoundaries()
    {
        dirichletZero(u);
        dirichletZero(v);
    }
    void subcycle()
    {
        projectVelocityToStrain();
        stressUpdate();
        stressDivergence();
        updateMomentum();
        applyBoundaries();
    }
    // VPCGDynamicsKernel::update (VPCGDynamicsKernel.hpp:63-94)
    void update(double dt)
    {
        if (rheology == 1) {
            updateBBM(dt);
            return;
        }
        advectionAndLimits(dt);
        prepareIteration();
        u0 = u;
        v0 = v;
        deltaT = dt;
        for (int k = 0; k < params.nSteps; ++k)
            subcycle();
    }
};

} // namespace ora
