// mEVP momentum/rheology subcycle and the per-timestep driver on gfx950.
//
// One mEVP subcycle is two bulk-synchronous HBM passes:
//  * element pass: strain rate from the 9 (CG2) / 4 (CG1) velocity nodes, the mEVP
//    stress relaxation at the Gauss points and its projection back to the DG stress
//    basis, in registers; reads u, v, h, A, sigma, writes sigma. Strain never
//    touches HBM.
//  * node pass: a 256-thread block takes a tile of elements plus a one-element halo,
//    computes each element's weak-divergence contributions to its CG nodes into LDS,
//    then every CG node of the tile sums its <= 4 contributions in the reference's
//    two-colour stripe order (atomic-free, deterministic) and applies the FP64
//    implicit velocity update and the Dirichlet zero.
// On a uniform mesh one copy of each element operator is read with wave-uniform
// addresses (scalar loads through the constant cache); otherwise operators are
// streamed per element, k-major. Per-step set-up (CG2DG, normal velocities, the two
// transported fields, limiters, DG2CG, sea-surface-height gradient) are further
// element- or node-parallel kernels. Nothing runs on the host between launches.
//
// Reference: dynamics/src/include/VPCGDynamicsKernel.hpp:63-172,
// dynamics/src/CGDynamicsKernel.cpp:23-444, MEVPStressUpdateStep.hpp:30-118,
// ParametricMap.cpp:94-356, Interpolations.cpp:129-204, DynamicsKernel.hpp:93-172.
#include "nsdg_device.hpp"
#include "nsdg_dyn.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <cmath>
#include <cstring>

namespace nsdg {

// ---------------------------------------------------------------- momentum operators
// Layout (k-major): iMgradX[s*N+j], iMgradY, iMJwPSI[s*Q+q], divS1[j*S+s], divS2,
// dX_SSH[j*4+k], dY_SSH (ParametricMap.hpp:62-131); on spherical meshes K + 2 S N values
// with the metric terms divM[j*S+s] and iMM[s*N+j] appended.
template <int CG> struct MomLayout {
    static constexpr int S = cg2dgstress(CG), G = gp1d(S), Q = G * G, N = cgdofs(CG);
    static constexpr int GX = 0, GY = S * N, B = 2 * S * N, D1 = 2 * S * N + S * Q, D2 = 3 * S * N + S * Q,
                         SX = 4 * S * N + S * Q, SY = SX + 16, K = SY + 16, DM = K, IMM = K + N * S,
                         KS = K + 2 * N * S;
};

// InitializeDivSMatrices (ParametricMap.cpp:209-356). On the sphere (x, y) = (longitude,
// latitude): longitude derivatives carry 1 / R, latitude derivatives cos(latitude) / R,
// the mass matrix cos(latitude), and divM = (PHI_j, PSI_s J sin(latitude) w) / R with
// iMM = M^-1 divM^T hold the metric terms of the divergence and the strain.
template <int CG> __global__ void k_momentum_ops(MeshView m, int nOps, double* ops)
{
    using L = MomLayout<CG>;
    constexpr int S = L::S, G = L::G, Q = L::Q, N = L::N;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nOps)
        return;
    auto put = [&](int k, double v) { ops[static_cast<size_t>(k) * nOps + e] = v; };
    const bool sph = m.spherical != 0;
    const double iR = 1.0 / kEarthRadius;
    double c[4][2], Fx[2][Q], Fy[2][Q], J[Q], cl[Q], sl[Q];
    elementCoords(m, e, c);
    mapGradients<G>(c, Fx, Fy);
    jacobian<G>(c, J);
    if (sph) {
        gaussLatitude<G>(c, sl);
        for (int q = 0; q < Q; ++q) {
            cl[q] = cos(sl[q]);
            sl[q] = sin(sl[q]);
        }
    }
    for (int r = 0; r < 2; ++r)
        for (int q = 0; q < Q; ++q) {
            Fx[r][q] *= W<G>::v[0][q];
            Fy[r][q] *= W<G>::v[0][q];
        }
    // mapped, weighted CG gradients: dx_cg = PHIx Fy1 - PHIy Fx1, dy_cg = PHIy Fx0 - PHIx Fy0
    double d1[N][S], d2[N][S], dm[N][S];
    for (int j = 0; j < N; ++j)
        for (int s = 0; s < S; ++s) {
            double a = 0, b = 0, t = 0;
            for (int q = 0; q < Q; ++q) {
                const double dxc = Phi<CG, G>::x[j][q] * Fy[1][q] - Phi<CG, G>::y[j][q] * Fx[1][q];
                const double dyc = Phi<CG, G>::y[j][q] * Fx[0][q] - Phi<CG, G>::x[j][q] * Fy[0][q];
                a += dxc * Psi<S, G>::v[s][q];
                b += (sph ? dyc * cl[q] : dyc) * Psi<S, G>::v[s][q];
                if (sph)
                    t += Phi<CG, G>::v[j][q] * (J[q] * sl[q] * W<G>::v[0][q]) * Psi<S, G>::v[s][q];
            }
            if (sph) {
                a *= iR;
                b *= iR;
                t *= iR;
                put(L::DM + j * S + s, t);
            }
            d1[j][s] = a;
            d2[j][s] = b;
            dm[j][s] = t;
            put(L::D1 + j * S + s, a);
            put(L::D2 + j * S + s, b);
        }
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) {
            double a = 0, b = 0;
            for (int q = 0; q < Q; ++q) {
                const double dxc = Phi<1, G>::x[j][q] * Fy[1][q] - Phi<1, G>::y[j][q] * Fx[1][q];
                const double dyc = Phi<1, G>::y[j][q] * Fx[0][q] - Phi<1, G>::x[j][q] * Fy[0][q];
                a += dxc * Phi<1, G>::v[k][q];
                b += (sph ? dyc * cl[q] : dyc) * Phi<1, G>::v[k][q];
            }
            put(L::SX + j * 4 + k, sph ? a * iR : a);
            put(L::SY + j * 4 + k, sph ? b * iR : b);
        }
    double M[S][S], Mi[S][S];
    metricMassMatrix<S>(m, c, M);
    invert<S>(M, Mi);
    for (int s = 0; s < S; ++s) {
        for (int j = 0; j < N; ++j) {
            double a = 0, b = 0;
            for (int t = 0; t < S; ++t) {
                a += Mi[s][t] * d1[j][t];
                b += Mi[s][t] * d2[j][t];
            }
            put(L::GX + s * N + j, a);
            put(L::GY + s * N + j, b);
            if (sph) {
                double t = 0;
                for (int r = 0; r < S; ++r)
                    t += Mi[s][r] * dm[j][r];
                put(L::IMM + s * N + j, t);
            }
        }
        for (int q = 0; q < Q; ++q) {
            double a = 0;
            for (int t = 0; t < S; ++t)
                a += Mi[s][t] * (Psi<S, G>::v[t][q] * (W<G>::v[0][q] * J[q]));
            put(L::B + s * Q + q, a);
        }
    }
}

// Per-element geometry of the one-pass subcycle on general meshes, SoA [k][ne]: the
// bilinear map x(xi, eta) = x0 + X1 xi + X2 eta + X3 xi eta (and y), so that at any point
// x_xi = X1 + X3 eta, x_eta = X2 + X3 xi; on the sphere cos and sin of the latitude at the
// Gauss points of the stress rule; last the upper triangle of the element's inverse
// (metric) mass matrix. At a Gauss point (gaussGeom) the weighted map gradients that make
// J grad(phi),
//   J dphi/dx = phi_xi y_eta - phi_eta y_xi,  J dphi/dy = phi_eta x_xi - phi_xi x_eta,
// are a = y_eta w, b = y_xi w, c = x_xi w cos(lat), d = x_eta w cos(lat) (all / R on the
// sphere), with J w and, on the sphere, J w sin(lat) / R. Every operator of
// InitializeDivSMatrices (ParametricMap.cpp:209-356) factors through these:
//   divS1 = sum_q (phi_xi a - phi_eta b) PSI(q),  divS2 = sum_q (phi_eta c - phi_xi d) PSI(q),
//   divM = sum_q phi J w sin(lat) / R PSI(q),  iMgradX/Y = M^-1 divS1/2^T,  iMM = M^-1 divM^T,
//   iMJwPSI = M^-1 PSI diag(J w),
// so 6 + 2 Q + S (S + 1) / 2 values (60 for CG2 on the sphere, 42 on a plane) replace the
// 4 S N + 2 S Q tables (432).
// The geometry pointer as a value the compiler cannot see through: each phase of the
// element update re-reads the values it needs instead of keeping every value of the
// previous phases live in registers.
__device__ __forceinline__ const double* phaseLocal(const double* p)
{
    asm volatile("" : "+s"(p));
    return p;
}
template <int CG> struct GeoLayout {
    static constexpr int S = cg2dgstress(CG), G = gp1d(S), Q = G * G;
    static constexpr int X1 = 0, X2 = 1, X3 = 2, Y1 = 3, Y2 = 4, Y3 = 5, CL = 6, SL = 6 + Q, MI = 6 + 2 * Q;
    static constexpr int K = MI + S * (S + 1) / 2;
    __host__ __device__ static constexpr int mi(int s, int t)
    {
        return s <= t ? MI + s * S - s * (s - 1) / 2 + (t - s) : mi(t, s);
    }
};
template <int CG> __global__ void k_geometry(MeshView m, int ne, double* geo)
{
    using L = GeoLayout<CG>;
    constexpr int S = L::S, G = L::G, Q = L::Q;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    auto put = [&](int k, double v) { geo[static_cast<size_t>(k) * ne + e] = v; };
    double c[4][2]; // LL, LR, UL, UR
    elementCoords(m, e, c);
    put(L::X1, c[1][0] - c[0][0]);
    put(L::X2, c[2][0] - c[0][0]);
    put(L::X3, (c[0][0] - c[1][0]) - (c[2][0] - c[3][0]));
    put(L::Y1, c[1][1] - c[0][1]);
    put(L::Y2, c[2][1] - c[0][1]);
    put(L::Y3, (c[0][1] - c[1][1]) - (c[2][1] - c[3][1]));
    double lat[Q];
    gaussLatitude<G>(c, lat);
    for (int q = 0; q < Q; ++q) {
        put(L::CL + q, m.spherical ? cos(lat[q]) : 1.0);
        put(L::SL + q, m.spherical ? sin(lat[q]) : 0.0);
    }
    double M[S][S], Mi[S][S];
    metricMassMatrix<S>(m, c, M);
    invert<S>(M, Mi);
    for (int s = 0; s < S; ++s)
        for (int t = s; t < S; ++t)
            put(L::mi(s, t), Mi[s][t]);
}
// The bilinear map coefficients of one element and the geometry they give at a Gauss point.
struct MapCoef {
    double x1, x2, x3, y1, y2, y3;
};
template <int CG> __device__ __forceinline__ MapCoef loadMap(const double* g, size_t ne, uint32_t off)
{
    using L = GeoLayout<CG>;
    return { ldAt(g + L::X1 * ne, off), ldAt(g + L::X2 * ne, off), ldAt(g + L::X3 * ne, off),
        ldAt(g + L::Y1 * ne, off), ldAt(g + L::Y2 * ne, off), ldAt(g + L::Y3 * ne, off) };
}
struct GaussGeom {
    double a, b, c, d, jw, js;
};
template <int NG, bool SPH>
__device__ __forceinline__ GaussGeom gaussGeom(const MapCoef& m, int gx, int gy, double cl, double sl)
{
    constexpr double iR = SPH ? 1.0 / kEarthRadius : 1.0;
    const double xi = gauss_point(NG, gx), eta = gauss_point(NG, gy);
    const double w = gauss_weight(NG, gx) * gauss_weight(NG, gy), wr = w * iR;
    const double xxi = m.x1 + m.x3 * eta, xeta = m.x2 + m.x3 * xi;
    const double yxi = m.y1 + m.y3 * eta, yeta = m.y2 + m.y3 * xi;
    const double J = xxi * yeta - xeta * yxi;
    const double wc = SPH ? wr * cl : wr;
    return { yeta * wr, yxi * wr, xxi * wc, xeta * wc, J * w, SPH ? J * w * sl * iR : 0.0 };
}
// y_c = M x_c (c = 1, 2, 3) for the element's symmetric inverse mass matrix, stored as its
// upper triangle (entry k = mi(s, t) - MI given by mAt(k)): each stored entry is read once
// and applied to both of its positions.
template <int CG, class MA>
__device__ __forceinline__ void inverseMassApply(const MA& mAt, const double* x1, const double* x2,
    const double* x3, double* y1, double* y2, double* y3)
{
    using L = GeoLayout<CG>;
    constexpr int S = L::S;
#pragma unroll
    for (int s = 0; s < S; ++s)
        y1[s] = y2[s] = y3[s] = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int t = s; t < S; ++t) {
            const double m = mAt(L::mi(s, t) - L::MI);
            y1[s] += m * x1[t];
            y2[s] += m * x2[t];
            y3[s] += m * x3[t];
            if (t != s) {
                y1[t] += m * x1[s];
                y2[t] += m * x2[s];
                y3[t] += m * x3[s];
            }
        }
}
// Per-element data of the brittle update on general meshes, SoA [k][ne]: the element size
// h = sqrt(area) (ParametricMesh::area, ParametricMesh.hpp:276-299, of the stored vertices,
// whose longitudes it does not correct), then iMJwPSI_dam (ParametricMap.cpp:290-296 and
// 346-351) = M_D^-1 PSI_D diag(w J) at the Gauss points of the stress rule, with the
// metric (cos-latitude) mass matrix on the sphere.
template <int CG, int DG> struct BrittleGeo {
    static constexpr int Q = MomLayout<CG>::Q, H = 0, DAM = 1;
    // then the upper triangle of the inverse (metric) DG mass matrix, for the one-pass kernel
    static constexpr int MDI = 1 + DG * Q, K = MDI + DG * (DG + 1) / 2;
    __host__ __device__ static constexpr int mdi(int s, int t)
    {
        return s <= t ? MDI + s * DG - s * (s - 1) / 2 + (t - s) : mdi(t, s);
    }
};
template <int CG, int DG> __global__ void k_brittle_geometry(MeshView m, int ne, double* out)
{
    using L = BrittleGeo<CG, DG>;
    constexpr int G = MomLayout<CG>::G, Q = L::Q;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    auto put = [&](int k, double v) { out[static_cast<size_t>(k) * ne + e] = v; };
    {
        const int nid = (e / m.nx) * (m.nx + 1) + (e % m.nx);
        const int ids[4] = { nid, nid + 1, nid + m.nx + 1, nid + m.nx + 2 }; // LL, LR, UL, UR
        auto sq = [&](int i, int j) {
            const double dx = m.vxy[2 * ids[i]] - m.vxy[2 * ids[j]], dy = m.vxy[2 * ids[i] + 1] - m.vxy[2 * ids[j] + 1];
            return dx * dx + dy * dy;
        };
        const double a = sq(0, 1), b = sq(1, 3), cc = sq(2, 3), d = sq(0, 2), ed = sq(0, 3), f = sq(1, 2);
        put(L::H, sqrt(0.25 * sqrt(4.0 * ed * f - (b + d - a - cc) * (b + d - a - cc))));
    }
    double c[4][2], J[Q], M[DG][DG], Mi[DG][DG];
    elementCoords(m, e, c);
    jacobian<G>(c, J);
    metricMassMatrix<DG>(m, c, M);
    invert<DG>(M, Mi);
    for (int d = 0; d < DG; ++d)
        for (int q = 0; q < Q; ++q) {
            double a = 0;
            for (int k = 0; k < DG; ++k)
                a += Mi[d][k] * Psi<DG, G>::v[k][q] * (W<G>::v[0][q] * J[q]);
            put(L::DAM + d * Q + q, a);
        }
    for (int d = 0; d < DG; ++d)
        for (int k = d; k < DG; ++k)
            put(L::mdi(d, k), Mi[d][k]);
}

// InitializeLumpedCGMassMatrix (ParametricMap.cpp:94-206): node-centric gather of the
// element integrals of the CG basis (even element rows first, as the stripes add them).
template <int CG, int GG> __global__ void k_lumped_mass(MeshView m, double* out)
{
    constexpr int N = cgdofs(CG);
    const int row = CG * m.nx + 1, nn = row * (CG * m.ny + 1);
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn)
        return;
    double acc = 0;
    forEachAdjacent<CG, true>(n % row, n / row, m.nx, m.ny, [&](int e, int loc) {
        double c[4][2], J[GG * GG], cl[GG * GG];
        elementCoords(m, e, c);
        jacobian<GG>(c, J);
        metricWeight<GG>(m, c, cl); // cos(latitude) on the sphere (ParametricMap.cpp:121-131)
        double me = 0;
        for (int q = 0; q < GG * GG; ++q)
            me += Phi<CG, GG>::v[loc][q] * (J[q] * W<GG>::v[0][q] * cl[q]);
        acc += me;
    });
    (void)N;
    out[n] = acc;
}

// Node tiles: a block owns the CG nodes of the lower-left 2x2 (CG2) node blocks of TX x TY
// elements (the last tile row/column also owns the domain edge) and gathers the weak
// divergence of the tile and a one-element left/bottom halo (HX x HY elements) from LDS.
template <int CG, int TX_, int TY_, int BLOCK = kBlock> struct TileShape {
    static constexpr int TX = TX_, TY = TY_;
    static constexpr int HX = TX + 1, HY = TY + 1, N = cgdofs(CG);
    static constexpr int CW = CG * TX;                        // owned node columns
    static constexpr int TPR = CW < BLOCK ? CW : BLOCK;       // threads per node row
    static constexpr int RS = BLOCK / TPR;                    // node rows per block sweep
    static constexpr int NB = (CG * TY + RS - 1) / RS;        // nodes per thread per sweep
};

// ---------------------------------------------------------------- DG -> CG
// Interpolations::DG2CG (Interpolations.cpp:129-204): Lagrange-point values weighted
// 1/4 (corner), 1/2 (edge midpoint), 1 (centre), gathered odd element rows first,
// doubled on each domain boundary the node lies on. clamp: 0 none, 1 H (>= 1e-4),
// 2 A (<= 1 then >= 1e-4), the limits of prepareIteration (CGDynamicsKernel.cpp:272-275).
template <int CG, int D>
__global__ __launch_bounds__(kBlock) void k_dg2cg(int nx, int ny, int ne, const double* s1, double* d1, int clamp1,
    const double* s2, double* d2, int clamp2)
{
    const int row = CG * nx + 1, nn = row * (CG * ny + 1);
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn)
        return;
    const int nix = n % row, niy = n / row;
    double a1 = 0, a2 = 0;
    forEachAdjacent<CG, false>(nix, niy, nx, ny, [&](int e, int j) {
        const int jx = j % (CG + 1), jy = j / (CG + 1);
        const double w = CG == 1 ? 0.25 : ((jx == 1 && jy == 1) ? 1.0 : ((jx == 1 || jy == 1) ? 0.5 : 0.25));
        double t1 = 0, t2 = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const double l = PsiL<D, CG + 1>::v[i][j];
            t1 += s1[static_cast<size_t>(i) * ne + e] * l;
            if (s2)
                t2 += s2[static_cast<size_t>(i) * ne + e] * l;
        }
        a1 += w * t1;
        a2 += w * t2;
    });
    double f = 1.0;
    if (niy == 0 || niy == CG * ny)
        f *= 2.0;
    if (nix == 0 || nix == CG * nx)
        f *= 2.0;
    auto clampv = [](double x, int c) {
        if (c == 1)
            return fmax(x, 1.e-4);
        if (c == 2)
            return fmax(fmin(x, 1.0), 1.e-4);
        return x;
    };
    d1[n] = clampv(a1 * f, clamp1);
    if (s2)
        d2[n] = clampv(a2 * f, clamp2);
}
// The same, tiled (31 x 7 elements per block, as the one-pass subcycle): every element of
// the tile and of its one-element left/bottom halo evaluates its weighted Lagrange-point
// values into LDS once, one element per thread, then the block's nodes (those of its
// elements' lower-left blocks, plus the domain edge for the last tile row and column) weight
// and sum them in the order of k_dg2cg, with the same operations; each element's
// coefficients are read once (plus the halo) instead of by each of its nodes.
template <int CG, int D>
__global__ __launch_bounds__(kBlock) void k_dg2cg_tiled(int nx, int ny, int ne, const double* s1, double* d1,
    int clamp1, const double* s2, double* d2, int clamp2)
{
    using T = TileShape<CG, 31, 7>;
    constexpr int N = T::N;
    __shared__ double val[T::HX * T::HY][N][2];
    const int tilesX = (nx + T::TX - 1) / T::TX, tilesY = (ny + T::TY - 1) / T::TY;
    const int tile = xcdTile(blockIdx.x, tilesX * tilesY);
    if (tile < 0)
        return;
    const int ex0 = (tile % tilesX) * T::TX, ey0 = (tile / tilesX) * T::TY;
    {
        const int le = threadIdx.x;
        const int ex = ex0 - 1 + le % T::HX, ey = ey0 - 1 + le / T::HX;
        if (ex >= 0 && ex < nx && ey >= 0 && ey < ny) {
            const int e = ex + nx * ey;
            double c1[D], c2[D];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                c1[i] = s1[static_cast<size_t>(i) * ne + e];
                c2[i] = s2 ? s2[static_cast<size_t>(i) * ne + e] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < N; ++j) {
                double t1 = 0, t2 = 0;
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    const double l = PsiL<D, CG + 1>::v[i][j];
                    t1 += c1[i] * l;
                    t2 += c2[i] * l;
                }
                val[le][j][0] = t1;
                val[le][j][1] = t2;
            }
        }
    }
    __syncthreads();
    const int row = CG * nx + 1;
    const int nx0 = CG * ex0, ny0 = CG * ey0;
    const int W = (ex0 + T::TX >= nx) ? row - nx0 : CG * T::TX;
    const int H = (ey0 + T::TY >= ny) ? CG * ny + 1 - ny0 : CG * T::TY;
    auto clampv = [](double x, int c) {
        if (c == 1)
            return fmax(x, 1.e-4);
        if (c == 2)
            return fmax(fmin(x, 1.0), 1.e-4);
        return x;
    };
    for (int idx = threadIdx.x; idx < W * H; idx += kBlock) {
        const int nix = nx0 + idx % W, niy = ny0 + idx / W;
        double a1 = 0, a2 = 0;
        forEachAdjacentXY<CG, false>(nix, niy, nx, ny, [&](int ex, int ey, int j) {
            const int le = (ex - ex0 + 1) + T::HX * (ey - ey0 + 1);
            const int jx = j % (CG + 1), jy = j / (CG + 1);
            const double w = CG == 1 ? 0.25 : ((jx == 1 && jy == 1) ? 1.0 : ((jx == 1 || jy == 1) ? 0.5 : 0.25));
            a1 += w * val[le][j][0];
            a2 += w * val[le][j][1];
        });
        double f = 1.0;
        if (niy == 0 || niy == CG * ny)
            f *= 2.0;
        if (nix == 0 || nix == CG * nx)
            f *= 2.0;
        const int n = nix + row * niy;
        d1[n] = clampv(a1 * f, clamp1);
        if (s2)
            d2[n] = clampv(a2 * f, clamp2);
    }
}
static_assert(TileShape<2, 31, 7>::HX * TileShape<2, 31, 7>::HY == kBlock, "one element per thread");

template <int CG> __global__ void k_clamp_HA(double* H, double* A, int nn)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn)
        return;
    A[n] = fmax(fmin(A[n], 1.0), 1.e-4);
    H[n] = fmax(H[n], 1.e-4);
}

// CGAveragePeriodic (VectorManipulations.hpp:26-65): sequential, as in the reference.
template <int CG> __global__ void k_periodic_average(int nx, const int32_t* ent, int n, double* v1, double* v2)
{
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    const int row = CG * nx + 1;
    for (int k = 0; k < n; ++k) {
        const int type = ent[4 * k], rt = ent[4 * k + 1], lb = ent[4 * k + 2];
        const int i0lb = row * CG * (lb / nx) + CG * (lb % nx);
        const int i0rt = row * CG * (rt / nx) + CG * (rt % nx);
        for (int f = 0; f < 2; ++f) {
            double* v = f == 0 ? v1 : v2;
            if (!v)
                continue;
            for (int j = 0; j <= CG; ++j) {
                const int i1 = type == 0 ? i0lb + j : i0lb + j * row;
                const int i2 = type == 0 ? i0rt + CG * row + j : i0rt + CG + j * row;
                v[i1] = 0.5 * (v[i1] + v[i2]);
                v[i2] = v[i1];
            }
        }
    }
}

// ---------------------------------------------------------------- sea-surface height
// ComputeGradientOfSeaSurfaceHeight (CGDynamicsKernel.cpp:125-258) in three passes:
// CG1 nodal SSH (k_dg2cg<1,1>), the CG1 weak gradient divided by the CG1 lumped mass
// (this kernel), then the boundary extension fused into the CG2 prolongation.
template <bool UNI>
__global__ void k_ssh_gradient(int nx, int ny, const double* cgS, const double* ops, size_t stride, int sx, int sy,
    const double* lumped1, double* gu, double* gv)
{
    const int row = nx + 1, nn = row * (ny + 1);
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn)
        return;
    const Ops<UNI> op { ops, stride };
    double u = 0, v = 0;
    forEachAdjacent<1, true>(n % row, n / row, nx, ny, [&](int e, int j) {
        const int id = (e / nx) * row + e % nx;
        const double l[4] = { cgS[id], cgS[id + 1], cgS[id + row], cgS[id + row + 1] };
        double tx = 0, ty = 0;
        for (int q = 0; q < 4; ++q) {
            tx += op(sx + j * 4 + q, e) * l[q];
            ty += op(sy + j * 4 + q, e) * l[q];
        }
        u -= tx;
        v -= ty;
    });
    gu[n] = u / lumped1[n];
    gv[n] = v / lumped1[n];
}
// Boundary extension of the CG1 gradient (CGDynamicsKernel.cpp:177-203) as an index map.
__device__ __forceinline__ int sshSource(int nix, int niy, int nx, int ny)
{
    int sx = nix, sy = niy;
    if (nix == 0)
        sx = 1;
    else if (nix == nx)
        sx = nx - 1;
    if (niy == 0)
        sy = 1;
    else if (niy == ny)
        sy = ny - 1;
    if ((nix == 0 || nix == nx) && (niy == 0 || niy == ny))
        return sx + (nx + 1) * sy;
    if (niy == 0 || niy == ny)
        return nix + (nx + 1) * sy;
    if (nix == 0 || nix == nx)
        return sx + (nx + 1) * niy;
    return nix + (nx + 1) * niy;
}
template <int CG> __global__ void k_ssh_prolong(int nx, int ny, const double* gu, const double* gv, double* gx, double* gy)
{
    const int row = CG * nx + 1, nn = row * (CG * ny + 1);
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn)
        return;
    const int jx = n % row, jy = n / row;
    auto src = [&](int a, int b) { return sshSource(a, b, nx, ny); };
    if (CG == 1) {
        const int s = src(jx, jy);
        gx[n] = gu[s];
        gy[n] = gv[s];
        return;
    }
    const int ax = jx >> 1, ay = jy >> 1;
    if (!(jx & 1) && !(jy & 1)) {
        const int s = src(ax, ay);
        gx[n] = gu[s];
        gy[n] = gv[s];
    } else if ((jx & 1) && !(jy & 1)) {
        const int s0 = src(ax, ay), s1 = src(ax + 1, ay);
        gx[n] = 0.5 * (gu[s0] + gu[s1]);
        gy[n] = 0.5 * (gv[s0] + gv[s1]);
    } else if (!(jx & 1) && (jy & 1)) {
        const int s0 = src(ax, ay), s1 = src(ax, ay + 1);
        gx[n] = 0.5 * (gu[s0] + gu[s1]);
        gy[n] = 0.5 * (gv[s0] + gv[s1]);
    } else {
        const int s0 = src(ax, ay), s1 = src(ax + 1, ay), s2 = src(ax, ay + 1), s3 = src(ax + 1, ay + 1);
        gx[n] = 0.25 * (gu[s0] + gu[s1] + gu[s2] + gu[s3]);
        gy[n] = 0.25 * (gv[s0] + gv[s1] + gv[s2] + gv[s3]);
    }
}

// ---------------------------------------------------------------- mEVP element pass
struct ElemArgs {
    int nx, ny, ne;
    const uint8_t* ocean;
    const double* u;
    const double* v;
    const double* h; // SoA DG
    const double* A;
    const double* eps; // SoA [3][S][ne] stored strain (land elements keep theirs)
    double* epsOut;
    double* sig; // SoA [3][S][ne]
    const double* ops;
    size_t stride;
    const double* Pg; // SoA [Q][ne] ice strength at the Gauss points (fixed over the subcycles)
    float* sigF;      // FP32 variant: stress and strength in single precision
    const float* PgF;
    uint32_t iter;    // subcycle counter (random bits of the FP32 stochastic rounding)
    int sph;          // spherical mesh: the metric terms iMM of the strain (operator tables)
    double Pstar, DM2, invAlpha, relax;
    double xiX, etaX, xiY, etaY; // inverse map gradient of an affine uniform mesh
};

template <int S, class T, class U> __device__ __forceinline__ void loadTensor(const T* f, size_t ne, int e, U* a, U* b, U* c)
{
#pragma unroll
    for (int s = 0; s < S; ++s) {
        a[s] = f[(0 * S + s) * ne + e];
        b[s] = f[(1 * S + s) * ne + e];
        c[s] = f[(2 * S + s) * ne + e];
    }
}
// The same with scalar field bases and the element's 32-bit byte offset (ldAt).
template <int S, class T> __device__ __forceinline__ void loadTensorAt(const T* f, size_t ne, uint32_t off, T* a, T* b, T* c)
{
#pragma unroll
    for (int s = 0; s < S; ++s) {
        a[s] = ldAt(f + (0 * S + s) * ne, off);
        b[s] = ldAt(f + (1 * S + s) * ne, off);
        c[s] = ldAt(f + (2 * S + s) * ne, off);
    }
}
template <int S, class T> __device__ __forceinline__ void storeTensorAt(T* f, size_t ne, uint32_t off, const T* a, const T* b, const T* c)
{
#pragma unroll
    for (int s = 0; s < S; ++s) {
        fieldAt(f + (0 * S + s) * ne, off) = a[s];
        fieldAt(f + (1 * S + s) * ne, off) = b[s];
        fieldAt(f + (2 * S + s) * ne, off) = c[s];
    }
}
template <int S, class T> __device__ __forceinline__ void storeTensor(T* f, size_t ne, int e, const T* a, const T* b, const T* c)
{
#pragma unroll
    for (int s = 0; s < S; ++s) {
        f[(0 * S + s) * ne + e] = a[s];
        f[(1 * S + s) * ne + e] = b[s];
        f[(2 * S + s) * ne + e] = c[s];
    }
}

// projectVelocityToStrain for one element (CGDynamicsKernel.cpp:300-337)
template <int CG, bool UNI>
__device__ __forceinline__ void strainLocal(const ElemArgs& a, int e, double* e11, double* e12, double* e22)
{
    using L = MomLayout<CG>;
    constexpr int S = L::S, N = L::N;
    const Ops<UNI> op { a.ops, a.stride };
    const int ix = e % a.nx, iy = e / a.nx, row = CG * a.nx + 1;
    const int cgi = CG * row * iy + CG * ix;
    double lu[N], lv[N];
#pragma unroll
    for (int r = 0; r <= CG; ++r)
#pragma unroll
        for (int c = 0; c <= CG; ++c) {
            lu[c + (CG + 1) * r] = a.u[cgi + c + r * row];
            lv[c + (CG + 1) * r] = a.v[cgi + c + r * row];
        }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        double x = 0, y = 0, p = 0, q = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const double gx = op(L::GX + s * N + j, e), gy = op(L::GY + s * N + j, e);
            x += gx * lu[j];
            y += gy * lv[j];
            p += gx * lv[j];
            q += gy * lu[j];
        }
        e11[s] = x;
        e22[s] = y;
        e12[s] = 0.5 * (p + q);
    }
    if constexpr (!UNI) {
        // spherical metric terms (CGDynamicsKernel.cpp:331-334): e11 -= iMM v, e12 += iMM u / 2
        if (a.sph) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                double mv = 0, mu = 0;
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const double im = op(L::IMM + s * N + j, e);
                    mv += im * lv[j];
                    mu += im * lu[j];
                }
                e11[s] -= mv;
                e12[s] += 0.5 * mu;
            }
        }
    }
}

// The 8 DG basis functions at (X, Y) = (x - 1/2, y - 1/2).
__device__ __forceinline__ constexpr void dgBasisAt(double X, double Y, double* b)
{
    b[0] = 1.0;
    b[1] = X;
    b[2] = Y;
    b[3] = X * X - 1.0 / 12.0;
    b[4] = Y * Y - 1.0 / 12.0;
    b[5] = X * Y;
    b[6] = Y * (X * X - 1.0 / 12.0);
    b[7] = X * (Y * Y - 1.0 / 12.0);
}
// Ice strength at a Gauss point from the interpolated h and A, clipped to h >= 0 and
// 0 <= A <= 1: P = P* h exp(-20 (1 - A)) (MEVPStressUpdateStep.hpp:64-72).
__device__ __forceinline__ double iceStrength(double Pstar, double hg, double ag)
{
    hg = fmax(hg, 0.0);
    ag = fmin(fmax(ag, 0.0), 1.0);
    return Pstar * hg * exp(-20.0 * (1.0 - ag));
}
// The strength of element e at its Gauss points, pk[gx + G gy], from its h and A.
template <int CG, int DG>
__device__ __forceinline__ void gaussStrength(const double* h, const double* A, size_t ne, int e, double Pstar,
    double* pk)
{
    constexpr int NG = MomLayout<CG>::G;
    double hc[DG], ac[DG];
#pragma unroll
    for (int i = 0; i < DG; ++i) {
        hc[i] = h[i * ne + e];
        ac[i] = A[i * ne + e];
    }
#pragma unroll
    for (int gy = 0; gy < NG; ++gy)
#pragma unroll
        for (int gx = 0; gx < NG; ++gx) {
            double b[8];
            dgBasisAt(gauss_point(NG, gx) - 0.5, gauss_point(NG, gy) - 0.5, b);
            double hg = 0, ag = 0;
#pragma unroll
            for (int i = 0; i < DG; ++i) {
                hg += hc[i] * b[i];
                ag += ac[i] * b[i];
            }
            pk[gx + NG * gy] = iceStrength(Pstar, hg, ag);
        }
}
// The strength depends only on h and A, which the subcycles do not change: evaluated once
// per timestep at the Gauss points of every element, Pg[q * ne + e], q = gx + G gy.
template <int CG, int DG, class T>
__global__ __launch_bounds__(kBlock) void k_strength(int ne, const double* h, const double* A, double Pstar, T* Pg)
{
    constexpr int Q = MomLayout<CG>::Q;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    double pk[Q];
    gaussStrength<CG, DG>(h, A, ne, e, Pstar, pk);
#pragma unroll
    for (int q = 0; q < Q; ++q)
        Pg[static_cast<size_t>(q) * ne + e] = static_cast<T>(pk[q]);
}

// Element-wise conversion between the FP64 stress and its FP32 copy.
template <class S, class D> __global__ __launch_bounds__(kBlock) void k_cast(const S* src, D* dst, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = static_cast<D>(src[i]);
}

// stressUpdateHighOrder for one element (MEVPStressUpdateStep.hpp:30-118): sigma <-
// (1 - 1/alpha) sigma + iMJwPSI (1/alpha) sigma_VP(strain, h, A) at the Gauss points.
template <int CG, int DG, bool UNI, bool PG, bool VP = false>
__device__ __forceinline__ void stressLocal(const ElemArgs& a, int e, const double* e11, const double* e12,
    const double* e22, double* s11, double* s12, double* s22)
{
    using L = MomLayout<CG>;
    constexpr int S = L::S, NG = L::G, Q = L::Q;
    const Ops<UNI> op { a.ops, a.stride };
    const size_t ne = a.ne;
    double hc[DG], ac[DG];
    if constexpr (!PG) {
#pragma unroll
        for (int i = 0; i < DG; ++i) {
            hc[i] = a.h[i * ne + e];
            ac[i] = a.A[i * ne + e];
        }
    }
    double acc1[S], acc2[S], acc3[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        acc1[s] = acc2[s] = acc3[s] = 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double g11 = 0, g12 = 0, g22 = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            g11 += e11[s] * Psi<S, NG>::v[s][q];
            g12 += e12[s] * Psi<S, NG>::v[s][q];
            g22 += e22[s] * Psi<S, NG>::v[s][q];
        }
        const double delta = sqrt(a.DM2 + 1.25 * (g11 * g11 + g22 * g22) + 1.50 * g11 * g22 + g12 * g12);
        double P;
        if constexpr (PG) {
            P = a.Pg[q * ne + e];
        } else {
            double hg = 0, ag = 0;
#pragma unroll
            for (int i = 0; i < DG; ++i) {
                hg += hc[i] * Psi<DG, NG>::v[i][q];
                ag += ac[i] * Psi<DG, NG>::v[i][q];
            }
            P = iceStrength(a.Pstar, hg, ag);
        }
        const double c = (P * 0.125) / delta; // P/8/Delta
        const double r11 = a.invAlpha * (c * (5.0 * g11 + 3.0 * g22) - 0.5 * P);
        const double r12 = a.invAlpha * ((2.0 * c) * g12);
        const double r22 = a.invAlpha * (c * (5.0 * g22 + 3.0 * g11) - 0.5 * P);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double b = op(L::B + s * Q + q, e);
            acc1[s] += b * r11;
            acc2[s] += b * r12;
            acc3[s] += b * r22;
        }
    }
    if constexpr (VP) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            s11[s] = acc1[s];
            s12[s] = acc2[s];
            s22[s] = acc3[s];
        }
        return;
    }
    // the previous stress is read only now, keeping its 3 S registers free above
    loadTensor<S>(a.sig, ne, e, s11, s12, s22);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        s11[s] = s11[s] * a.relax + acc1[s];
        s12[s] = s12[s] * a.relax + acc2[s];
        s22[s] = s22[s] * a.relax + acc3[s];
    }
}

// Uniform mesh (every element a translate of element 0, hence a parallelogram): every
// element operator reduces to the reference-element basis, evaluated here from 1-D values
// with sum factorisation (constexpr functions folded to literals, no operator loads).
//  * Gauss-point strain: the exact gradient of the CG velocity (the DG stress space holds
//    the CG derivatives, so the reference's L2 projection is exact), from PHI_x/PHI_y in
//    tensor form and the four inverse-map scalars.
//  * Projection back to the stress basis: M^-1 PSI diag(w J) = PSI w / m with
//    m = (1, 1/12, 1/12, 1/180, 1/180, 1/144, 1/2160, 1/2160) on every affine element.
// Land elements take their strain from the stored coefficients (the reference never
// updates it).
__device__ __forceinline__ constexpr double invMassRef(int s)
{
    return s == 0 ? 1.0 : (s <= 2 ? 12.0 : (s <= 4 ? 180.0 : (s == 5 ? 144.0 : 2160.0)));
}
// The relaxed stress: plain in FP64; in FP32 a (sum, compensation) pair whose exact value
// is s - c (initSum: Fast2Sum of a + b, exact when |a| >= |b| or a = 0).
template <class R> struct Acc {
    R s;
    __device__ __forceinline__ void init(R v) { s = v; }
    __device__ __forceinline__ void add(R x) { s += x; }
    __device__ __forceinline__ R get() const { return s; }
};
template <> struct Acc<float> {
    float s, c;
    __device__ __forceinline__ void init(float v)
    {
        s = v;
        c = 0.0f;
    }
    __device__ __forceinline__ void initSum(float a, float b)
    {
        s = a + b;
        c = (s - a) - b;
    }
    __device__ __forceinline__ float get() const { return s - c; }
};

// The mEVP stress relaxation sigma <- (1 - 1/alpha) sigma + ... amplifies a bias in the
// rounding of the stored stress by up to alpha (near a fixed point round-to-nearest makes
// the same error every subcycle). The FP32 stress is therefore stored with unbiased
// stochastic rounding: x is rounded away from zero with probability (|x| - |rz(x)|) / ulp,
// the random bits a hash of (element, coefficient, subcycle), so runs are reproducible.
// One full hash per element and subcycle; from it, one multiply-xorshift round per two
// coefficients (a Weyl step apart), whose low and high halves give each 16 random bits.
// (The hash XORed with a constant per coefficient is uniform too, and as unbiased, but the
// roundings of one element's coefficients are then correlated; measured on the FP32
// variant's tests, that doubles to quadruples the stress error against FP64.) 16 bits
// quantise the rounding probability to 2^-16, a bias of at most 2^-16 of the gap per
// rounding: 0.02 ulp after the alpha-fold accumulation near a fixed point.
__device__ __forceinline__ uint32_t elementSeed(uint32_t e, uint32_t it)
{
    uint32_t h = e * 0x9E3779B1u ^ (it + 1u) * 0xC2B2AE3Du;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}
__device__ __forceinline__ uint32_t coefficientBits(uint32_t seed, uint32_t k)
{
    uint32_t x = seed + ((k >> 1) + 1u) * 0x9E3779B9u;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    return (k & 1u) ? x >> 16 : x & 0xffffu;
}
// The exact value s - c of a pair with |c| at most half an ulp of s lies between s and its
// neighbour g toward s - c; it is rounded to g with probability |c| / |g - s| (g - s is
// exact), which is (|x| - |rz(x)|) / ulp of x = s - c rounded away from zero or toward it.
// c = 0 (s exact, zero included) keeps s. The uniform u in [1, 2) takes the 16 random
// bits as the top of the mantissa of 1.0; (u - 1) |g - s| is then exact.
__device__ __forceinline__ float roundStochastic(const Acc<float>& a, uint32_t bits16)
{
    const uint32_t sb = __float_as_uint(a.s);
    const bool away = ((sb ^ __float_as_uint(a.c)) >> 31) != 0; // the exact value is farther from zero
    const float g = __uint_as_float(away ? sb + 1u : sb - 1u);
    const float u = __uint_as_float((bits16 << 7) | 0x3f800000u);
    const float gap = fabsf(g - a.s);
    return fmaf(u, gap, -gap) < fabsf(a.c) ? g : a.s;
}
template <class R, class C> __device__ __forceinline__ R storeValue(const Acc<C>& a, int e, int k, uint32_t it)
{
    if constexpr (sizeof(R) == 8)
        return a.get();
    else
        return roundStochastic(a, coefficientBits(elementSeed(static_cast<uint32_t>(e), it), static_cast<uint32_t>(k)));
}
__device__ __forceinline__ double rsq(double x) { return rsqrt(x); }
__device__ __forceinline__ float rsq(float x) { return rsqrtf(x); }
template <class R> __device__ __forceinline__ R* stressOf(const ElemArgs& a);
template <> __device__ __forceinline__ double* stressOf<double>(const ElemArgs& a) { return a.sig; }
template <> __device__ __forceinline__ float* stressOf<float>(const ElemArgs& a) { return a.sigF; }
template <class R> __device__ __forceinline__ const R* strengthOf(const ElemArgs& a);
template <> __device__ __forceinline__ const double* strengthOf<double>(const ElemArgs& a) { return a.Pg; }
template <> __device__ __forceinline__ const float* strengthOf<float>(const ElemArgs& a) { return a.PgF; }

// R = double: the FP64 path. R = float: strain, Delta, stress and the projection in FP32;
// the FP64 node velocities enter relative to the element's centre node (their differences,
// which the gradient needs, are formed exactly in FP64 before rounding to FP32).
// VP: sigma <- Pi sigma_VP(eps(u)) outright (no relaxation, the old stress is not read);
// used for the momentum residual and for the stress of a predicted velocity.
// vel(lu, lv) gives the element's node velocities by node row (ocean elements only);
// old(c, s) the previous stress, component c (11, 12, 22), coefficient s.
// str(q) gives the ice strength at Gauss point q (row-major, x fastest); by default it is
// read from the strength table.
template <int CG, int DG, class R, class C, bool VP, class VEL, class OLD, class STR>
__device__ __forceinline__ void elementAffineCore(
    const ElemArgs& a, int e, Acc<C>* acc1, Acc<C>* acc2, Acc<C>* acc3, const VEL& vel, const OLD& old, const STR& str)
{
    using L = MomLayout<CG>;
    constexpr int S = L::S, NG = L::G, N = L::N, K = CG + 1;
    const size_t ne = a.ne;
    const bool oc = ldAt(a.ocean, static_cast<uint32_t>(e)) != 0;
    C buf[3 * S]; // ocean: u (N), v (N) by node row; land: e11, e12, e22 (S each)
    if (oc) {
        double lu[N], lv[N];
        vel(lu, lv);
        const double cu = sizeof(C) == 8 ? 0.0 : lu[N / 2], cv = sizeof(C) == 8 ? 0.0 : lv[N / 2];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            buf[j] = static_cast<C>(lu[j] - cu);
            buf[N + j] = static_cast<C>(lv[j] - cv);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 3 * S; ++k)
            buf[k] = static_cast<C>(a.eps[k * ne + e]);
    }
    const C xiX = static_cast<C>(a.xiX), etaX = static_cast<C>(a.etaX);
    const C xiY = static_cast<C>(a.xiY), etaY = static_cast<C>(a.etaY);
    const C DM2 = static_cast<C>(a.DM2);
    // The projected terms sum_q PSI_s(q) w_q / m_s r(q) of each coefficient, summed plainly:
    // in acc (FP64), in t (FP32).
    C t1[S], t2[S], t3[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        t1[s] = t2[s] = t3[s] = 0;
        if constexpr (sizeof(C) == 8) {
            acc1[s].init(0);
            acc2[s].init(0);
            acc3[s].init(0);
        }
    }
#pragma unroll
    for (int gy = 0; gy < NG; ++gy) {
        // velocity interpolated (Iu, Iv) and differentiated (Du, Dv) along y at this Gauss row
        C Iu[K], Du[K], Iv[K], Dv[K];
        if (oc) {
#pragma unroll
            for (int jx = 0; jx < K; ++jx) {
                C iu = 0, du = 0, iv = 0, dv = 0;
#pragma unroll
                for (int jy = 0; jy < K; ++jy) {
                    const C ph = static_cast<C>(cg_1d(CG, jy, gauss_point(NG, gy)));
                    const C dp = static_cast<C>(cg_1d_dx(CG, jy, gauss_point(NG, gy)));
                    iu += ph * buf[jx + K * jy];
                    du += dp * buf[jx + K * jy];
                    iv += ph * buf[N + jx + K * jy];
                    dv += dp * buf[N + jx + K * jy];
                }
                Iu[jx] = iu;
                Du[jx] = du;
                Iv[jx] = iv;
                Dv[jx] = dv;
            }
        }
#pragma unroll
        for (int gx = 0; gx < NG; ++gx) {
            double bd[8];
            dgBasisAt(gauss_point(NG, gx) - 0.5, gauss_point(NG, gy) - 0.5, bd);
            C g11 = 0, g12 = 0, g22 = 0;
            if (oc) {
                C uxi = 0, ueta = 0, vxi = 0, veta = 0;
#pragma unroll
                for (int jx = 0; jx < K; ++jx) {
                    const C ph = static_cast<C>(cg_1d(CG, jx, gauss_point(NG, gx)));
                    const C dp = static_cast<C>(cg_1d_dx(CG, jx, gauss_point(NG, gx)));
                    uxi += dp * Iu[jx];
                    ueta += ph * Du[jx];
                    vxi += dp * Iv[jx];
                    veta += ph * Dv[jx];
                }
                g11 = uxi * xiX + ueta * etaX;
                g22 = vxi * xiY + veta * etaY;
                g12 = C(0.5) * ((uxi * xiY + ueta * etaY) + (vxi * xiX + veta * etaX));
            } else {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const C b = static_cast<C>(bd[s]);
                    g11 += buf[s] * b;
                    g12 += buf[S + s] * b;
                    g22 += buf[2 * S + s] * b;
                }
            }
            const C X = DM2 + C(1.25) * (g11 * g11 + g22 * g22) + C(1.50) * g11 * g22 + g12 * g12;
            const C P = str(gx + NG * gy);
            const C c = (P * C(0.125)) * rsq(X); // P / (8 Delta)
            const C w = static_cast<C>(gauss_weight(NG, gx) * gauss_weight(NG, gy)) * static_cast<C>(a.invAlpha);
            const C r11 = w * (c * (C(5.0) * g11 + C(3.0) * g22) - C(0.5) * P);
            const C r12 = w * ((C(2.0) * c) * g12);
            const C r22 = w * (c * (C(5.0) * g22 + C(3.0) * g11) - C(0.5) * P);
            // iMJwPSI[s][q] = PSI_s(q) w_q / m_s: w_q is in r, PSI_s(q) / m_s is a literal
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const double bs = bd[s] * invMassRef(s);
                if (bs == 0.0) // basis functions vanishing at this Gauss point (a literal test)
                    continue;
                if constexpr (sizeof(C) == 8) {
                    acc1[s].add(static_cast<C>(bs) * r11);
                    acc2[s].add(static_cast<C>(bs) * r12);
                    acc3[s].add(static_cast<C>(bs) * r22);
                } else {
                    t1[s] += static_cast<C>(bs) * r11;
                    t2[s] += static_cast<C>(bs) * r12;
                    t3[s] += static_cast<C>(bs) * r22;
                }
            }
        }
    }
    // The relaxed previous stress, read only now (3 S registers stay free through the strain).
    // FP64: (1 - 1/alpha) sigma + t, the order of MEVPStressUpdateStep.hpp:78-116. FP32:
    // sigma + d with d = t - sigma/alpha in one FMA (one rounding, and exactly zero at the
    // fixed point t = sigma (1/alpha), with the same rounded 1/alpha that scales the new
    // term; a separately rounded 1 - 1/alpha would move the fixed point by up to 2^-24 alpha
    // relative), then sigma + d as an exact pair, which the stochastic rounding of the
    // stored value rounds without bias. The rounding inside t and d is made on terms of
    // size sigma_VP / alpha and moves the fixed point by as much, relative; a rounding of
    // sigma + d itself would be made on sigma and amplified alpha-fold, which the pair avoids.
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if constexpr (sizeof(C) == 8) {
            if constexpr (!VP) {
                acc1[s].s = old(0, s) * a.relax + acc1[s].s;
                acc2[s].s = old(1, s) * a.relax + acc2[s].s;
                acc3[s].s = old(2, s) * a.relax + acc3[s].s;
            }
        } else if constexpr (VP) {
            acc1[s].init(t1[s]);
            acc2[s].init(t2[s]);
            acc3[s].init(t3[s]);
        } else {
            const C o1 = old(0, s), o2 = old(1, s), o3 = old(2, s);
            const C mInvAlpha = -static_cast<C>(a.invAlpha);
            acc1[s].initSum(o1, fmaf(o1, mInvAlpha, t1[s]));
            acc2[s].initSum(o2, fmaf(o2, mInvAlpha, t2[s]));
            acc3[s].initSum(o3, fmaf(o3, mInvAlpha, t3[s]));
        }
    }
}
template <int CG, int DG, class R, class C, bool VP, class VEL, class OLD>
__device__ __forceinline__ void elementAffineCore(
    const ElemArgs& a, int e, Acc<C>* acc1, Acc<C>* acc2, Acc<C>* acc3, const VEL& vel, const OLD& old)
{
    const R* const Pg = strengthOf<R>(a);
    const size_t ne = a.ne;
    const uint32_t eoff = static_cast<uint32_t>(e) * static_cast<uint32_t>(sizeof(R));
    elementAffineCore<CG, DG, R, C, VP>(
        a, e, acc1, acc2, acc3, vel, old, [&](int q) -> C { return ldAt(Pg + q * ne, eoff); });
}
// The element update from the state in HBM (velocity a.u, a.v; stress of precision R);
// str(q) as in elementAffineCore.
template <int CG, int DG, class R, class C, bool VP, class STR>
__device__ __forceinline__ void elementAffineCompute(
    const ElemArgs& a, int e, Acc<C>* acc1, Acc<C>* acc2, Acc<C>* acc3, const STR& str)
{
    constexpr int S = MomLayout<CG>::S, K = CG + 1;
    const size_t ne = a.ne;
    const R* const sig = stressOf<R>(a);
    const uint32_t eoff = static_cast<uint32_t>(e) * static_cast<uint32_t>(sizeof(R));
    elementAffineCore<CG, DG, R, C, VP>(
        a, e, acc1, acc2, acc3,
        [&](double* lu, double* lv) {
            const int ix = e % a.nx, iy = e / a.nx, row = CG * a.nx + 1;
            const uint32_t noff = static_cast<uint32_t>(CG * row * iy + CG * ix) * 8u;
#pragma unroll
            for (int r = 0; r < K; ++r)
#pragma unroll
                for (int c = 0; c < K; ++c) {
                    lu[c + K * r] = ldAt(a.u + c + r * row, noff);
                    lv[c + K * r] = ldAt(a.v + c + r * row, noff);
                }
        },
        [&](int c, int k) -> C { return ldAt(sig + (c * S + k) * ne, eoff); }, str);
}
template <int CG, int DG, class R, class C, bool VP>
__device__ __forceinline__ void elementAffineCompute(const ElemArgs& a, int e, Acc<C>* acc1, Acc<C>* acc2, Acc<C>* acc3)
{
    const R* const Pg = strengthOf<R>(a);
    const size_t ne = a.ne;
    const uint32_t eoff = static_cast<uint32_t>(e) * static_cast<uint32_t>(sizeof(R));
    elementAffineCompute<CG, DG, R, C, VP>(
        a, e, acc1, acc2, acc3, [&](int q) -> C { return ldAt(Pg + q * ne, eoff); });
}
template <int CG, int DG, class R, class C = R, bool VP = false>
__device__ __forceinline__ void elementAffine(const ElemArgs& a, int e)
{
    constexpr int S = MomLayout<CG>::S;
    const size_t ne = a.ne;
    Acc<C> acc1[S], acc2[S], acc3[S];
    elementAffineCompute<CG, DG, R, C, VP>(a, e, acc1, acc2, acc3);
    R* const sig = stressOf<R>(a);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        sig[(0 * S + s) * ne + e] = storeValue<R>(acc1[s], e, 0 * S + s, a.iter);
        sig[(1 * S + s) * ne + e] = storeValue<R>(acc2[s], e, 1 * S + s, a.iter);
        sig[(2 * S + s) * ne + e] = storeValue<R>(acc3[s], e, 2 * S + s, a.iter);
    }
}

// Fused strain + stress (one HBM pass per subcycle over the elements).
template <int CG, int DG, bool UNI> __global__ __launch_bounds__(kBlock) void k_mevp_element(ElemArgs a)
{
    constexpr int S = cg2dgstress(CG);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.ne)
        return;
    if constexpr (UNI) {
        elementAffine<CG, DG, double>(a, e);
    } else {
        double e11[S], e12[S], e22[S], s11[S], s12[S], s22[S];
        if (a.ocean[e])
            strainLocal<CG, UNI>(a, e, e11, e12, e22);
        else
            loadTensor<S>(a.eps, a.ne, e, e11, e12, e22);
        stressLocal<CG, DG, UNI, true>(a, e, e11, e12, e22, s11, s12, s22);
        storeTensor<S>(a.sig, a.ne, e, s11, s12, s22);
    }
}
// sigma <- Pi sigma_VP(eps(u)) with a.invAlpha = 1 (the converged mEVP stress of u).
template <int CG, int DG, bool UNI> __global__ __launch_bounds__(kBlock) void k_vp_stress(ElemArgs a)
{
    constexpr int S = cg2dgstress(CG);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.ne)
        return;
    if constexpr (UNI) {
        elementAffine<CG, DG, double, double, true>(a, e);
    } else {
        double e11[S], e12[S], e22[S], s11[S], s12[S], s22[S];
        if (a.ocean[e])
            strainLocal<CG, UNI>(a, e, e11, e12, e22);
        else
            loadTensor<S>(a.eps, a.ne, e, e11, e12, e22);
        stressLocal<CG, DG, UNI, true, true>(a, e, e11, e12, e22, s11, s12, s22);
        storeTensor<S>(a.sig, a.ne, e, s11, s12, s22);
    }
}
template <int CG, bool UNI> __global__ __launch_bounds__(kBlock) void k_strain(ElemArgs a)
{
    constexpr int S = cg2dgstress(CG);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.ne || !a.ocean[e])
        return;
    double e11[S], e12[S], e22[S];
    strainLocal<CG, UNI>(a, e, e11, e12, e22);
    storeTensor<S>(a.epsOut, a.ne, e, e11, e12, e22);
}
template <int CG, int DG, bool UNI> __global__ __launch_bounds__(kBlock) void k_stress(ElemArgs a)
{
    constexpr int S = cg2dgstress(CG);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.ne)
        return;
    double e11[S], e12[S], e22[S], s11[S], s12[S], s22[S];
    loadTensor<S>(a.eps, a.ne, e, e11, e12, e22);
    stressLocal<CG, DG, UNI, false>(a, e, e11, e12, e22, s11, s12, s22);
    storeTensor<S>(a.sig, a.ne, e, s11, s12, s22);
}

// ---------------------------------------------------------------- mEVP node pass
struct NodeArgs {
    int nx, ny, ne, nn;
    const uint8_t* ocean;
    const uint8_t* dirNode;
    const double* sig;
    const double* ops;
    size_t stride;
    double* u;
    double* v;
    const double *u0, *v0, *uo, *vo, *ua, *va, *H, *A, *gx, *gy, *ML;
    const double2* ocn; // frozen per-node coefficients, see k_node_coeffs
    const double2* cf;
    const double2* qf;
    double* dSx;
    double* dSy;
    double rho_ice, F_atm, F_ocean, fc, gravity, beta, dt;
    double xiX, etaX, xiY, etaY, detJ; // affine uniform mesh geometry
    int sph;                           // spherical mesh: the metric terms divM of the divergence
};

// updateMomentum for one node (VPCGDynamicsKernel.hpp:132-172), FP64: the node's inputs
// are gathered first (NodeIn) so that a thread can keep several nodes' loads in flight.
struct NodeIn {
    double u, v, uo, vo, ua, va, H, A, ML, u0, v0, gx, gy;
};
__device__ __forceinline__ void loadNode(const NodeArgs& a, int n, NodeIn& in)
{
    in.u = a.u[n];
    in.v = a.v[n];
    in.uo = a.uo[n];
    in.vo = a.vo[n];
    in.ua = a.ua[n];
    in.va = a.va[n];
    in.H = a.H[n];
    in.A = a.A[n];
    in.ML = a.ML[n];
    in.u0 = a.u0[n];
    in.v0 = a.v0[n];
    in.gx = a.gx[n];
    in.gy = a.gy[n];
}
__device__ __forceinline__ void momentumUpdate(const NodeArgs& a, const NodeIn& in, double dSx, double dSy, double& un,
    double& vn)
{
    const double uOcnRel = in.uo - in.u, vOcnRel = in.v - in.vo;
    const double absatm = sqrt(in.ua * in.ua + in.va * in.va);
    const double absocn = sqrt(uOcnRel * uOcnRel + vOcnRel * vOcnRel);
    const double c1 = a.rho_ice * in.H / a.dt;
    const double rhoH = a.rho_ice * in.H;
    const double inv = 1.0 / (c1 * (1.0 + a.beta) + in.A * a.F_ocean * absocn);
    un = inv
        * (c1 * (a.beta * in.u + in.u0) + in.A * (a.F_atm * absatm * in.ua + a.F_ocean * absocn * in.uo)
            - rhoH * a.fc * in.u - rhoH * a.gravity * in.gx + dSx / in.ML);
    vn = inv
        * (c1 * (a.beta * in.v + in.v0) + in.A * (a.F_atm * absatm * in.va + a.F_ocean * absocn * in.vo)
            + rhoH * a.fc * in.v - rhoH * a.gravity * in.gy + dSy / in.ML);
}
__device__ __forceinline__ void momentumNode(const NodeArgs& a, int n, double dSx, double dSy)
{
    NodeIn in;
    loadNode(a, n, in);
    double un, vn;
    momentumUpdate(a, in, dSx, dSy, un, vn);
    a.u[n] = un;
    a.v[n] = vn;
}

// Weak-divergence contributions of one element to its CG nodes:
// tx_j = (divS1 s11 + divS2 s12)_j, ty_j = (divS1 s12 + divS2 s22)_j (CGDynamicsKernel.cpp:339-363).
// Spherical metric terms of addStressTensorCell (CGDynamicsKernel.cpp:348-351):
// tx += divM s12, ty -= divM s11.
template <int CG, class O>
__device__ __forceinline__ void divergenceMetric(const NodeArgs& a, const O& op, int e, const double* s11,
    const double* s12, double* tx, double* ty)
{
    using L = MomLayout<CG>;
    constexpr int S = L::S, N = cgdofs(CG);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double m12 = 0, m11 = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double dm = op(L::DM + j * S + s, e);
            m12 += dm * s12[s];
            m11 += dm * s11[s];
        }
        tx[j] += m12;
        ty[j] -= m11;
    }
}
template <int CG, bool UNI>
__device__ __forceinline__ void divergenceLocal(const NodeArgs& a, int e, double* tx, double* ty)
{
    using L = MomLayout<CG>;
    constexpr int S = L::S, N = cgdofs(CG);
    double s11[S], s12[S], s22[S];
    loadTensor<S>(a.sig, a.ne, e, s11, s12, s22);
    const Ops<UNI> op { a.ops, a.stride };
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double p1 = 0, p2 = 0, p3 = 0, p4 = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double d1 = op(L::D1 + j * S + s, e), d2 = op(L::D2 + j * S + s, e);
            p1 += d1 * s11[s];
            p2 += d2 * s12[s];
            p3 += d1 * s12[s];
            p4 += d2 * s22[s];
        }
        tx[j] = p1 + p2;
        ty[j] = p3 + p4;
    }
    if constexpr (!UNI)
        if (a.sph)
            divergenceMetric<CG>(a, op, e, s11, s12, tx, ty);
}
// The same on a uniform (affine) mesh. With the constant Jacobian the divergence is linear
// in the stress coefficients through two reference-element matrices,
//   Dxi[j][s] = int PHI_j,xi PSI_s,  Deta[j][s] = int PHI_j,eta PSI_s   (over [0,1]^2),
// tx = Dxi p1 + Deta q1 and ty = Dxi p2 + Deta q2, with p1 = |J| (xi_x s11 + xi_y s12),
// q1 = |J| (eta_x s11 + eta_y s12), p2 = |J| (xi_x s12 + xi_y s22), q2 = |J| (eta_x s12 +
// eta_y s22): the reference's Gauss-point form (divS1, divS2: ParametricMap.cpp:269-271) with the
// quadrature carried out in the literals. The matrices are the Gauss rule's values (exact
// for these polynomial degrees), folded to constants; entries that vanish analytically
// are exact zeros and are skipped (40 of 72 nonzero per matrix for CG2).
template <int CG> __device__ __forceinline__ constexpr double divReference(int d, int j, int s)
{
    constexpr int NG = gp1d(cg2dgstress(CG)), K = CG + 1;
    const int jx = j % K, jy = j / K;
    double v = 0.0;
    for (int qy = 0; qy < NG; ++qy)
        for (int qx = 0; qx < NG; ++qx) {
            const double x = gauss_point(NG, qx), y = gauss_point(NG, qy);
            const double dphi = d == 0 ? cg_1d_dx(CG, jx, x) * cg_1d(CG, jy, y) : cg_1d(CG, jx, x) * cg_1d_dx(CG, jy, y);
            v += gauss_weight(NG, qx) * gauss_weight(NG, qy) * dphi * dg_basis(s, x, y);
        }
    return (v < 1e-13 && v > -1e-13) ? 0.0 : v;
}
// t = Dxi p + Deta q with p = xi_x a + xi_y b, q = eta_x a + eta_y b (|J| in the scalars).
template <int CG, class R>
__device__ __forceinline__ void divergenceRows(R xiX, R xiY, R etaX, R etaY, const R* a, const R* b, R* t)
{
    constexpr int S = cg2dgstress(CG), N = cgdofs(CG);
    R p[S], q[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        p[k] = xiX * a[k] + xiY * b[k];
        q[k] = etaX * a[k] + etaY * b[k];
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        R x = 0;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const double dxi = divReference<CG>(0, j, k), deta = divReference<CG>(1, j, k);
            if (dxi != 0.0)
                x += static_cast<R>(dxi) * p[k];
            if (deta != 0.0)
                x += static_cast<R>(deta) * q[k];
        }
        t[j] = x;
    }
}
// Contributions from the stress coefficients of one element (registers).
template <int CG, class R>
__device__ __forceinline__ void divergenceAffineFrom(const NodeArgs& a, const R* s11, const R* s12, const R* s22,
    R* tx, R* ty)
{
    const R xiX = static_cast<R>(a.xiX * a.detJ), etaX = static_cast<R>(a.etaX * a.detJ);
    const R xiY = static_cast<R>(a.xiY * a.detJ), etaY = static_cast<R>(a.etaY * a.detJ);
    // tx from (s11, s12), then ty from (s12, s22): half the intermediate registers live
    divergenceRows<CG>(xiX, xiY, etaX, etaY, s11, s12, tx);
    divergenceRows<CG>(xiX, xiY, etaX, etaY, s12, s22, ty);
}

// The momentum update with the terms that stay fixed over the subcycles of one timestep
// (u0, v0, wind, H, A, SSH gradient, lumped mass) folded once per timestep into
//   ocn = (uo, vo), cf = ML (c1, A F_ocean) with c1 = rho_ice H / dt,
//   qf = ML (c1 u0 + A F_atm |ua| ua - rho_ice H g gx, c1 v0 + A F_atm |ua| va - rho_ice H g gy),
// so that a subcycle reads 8 values per node instead of 13. The update is the one of
// VPCGDynamicsKernel.hpp:132-172 (same-component Coriolis term included) with numerator
// and denominator multiplied by the lumped mass ML, which then divides nothing: the weak
// divergence enters as it is.
__global__ __launch_bounds__(kBlock) void k_node_coeffs(NodeArgs a, double2* ocn, double2* cf, double2* qf)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.nn)
        return;
    const double H = a.H[n], A = a.A[n], ua = a.ua[n], va = a.va[n];
    const double absatm = sqrt(ua * ua + va * va);
    const double c1 = a.rho_ice * H / a.dt, rhoH = a.rho_ice * H;
    const double ML = a.ML[n];
    ocn[n] = make_double2(a.uo[n], a.vo[n]);
    cf[n] = make_double2(ML * c1, ML * (A * a.F_ocean));
    qf[n] = make_double2(ML * (c1 * a.u0[n] + A * (a.F_atm * absatm * ua) - rhoH * a.gravity * a.gx[n]),
        ML * (c1 * a.v0[n] + A * (a.F_atm * absatm * va) - rhoH * a.gravity * a.gy[n]));
}
struct NodeFz {
    double u, v;
    double2 ocn, cf, qf;
};
__device__ __forceinline__ void loadFrozen(const NodeArgs& a, int n, NodeFz& in)
{
    in.u = a.u[n];
    in.v = a.v[n];
    in.ocn = a.ocn[n];
    in.cf = a.cf[n];
    in.qf = a.qf[n];
}
// The same with scalar field bases and 32-bit byte offsets (node n < 2^29).
__device__ __forceinline__ void loadFrozenAt(const NodeArgs& a, uint32_t n, NodeFz& in)
{
    in.u = ldAt(a.u, n * 8u);
    in.v = ldAt(a.v, n * 8u);
    in.ocn = ld2At(a.ocn, n * 16u);
    in.cf = ld2At(a.cf, n * 16u);
    in.qf = ld2At(a.qf, n * 16u);
}
__device__ __forceinline__ void momentumFrozen(const NodeArgs& a, const NodeFz& in, double dSx, double dSy, double& un,
    double& vn)
{
    const double uOcnRel = in.ocn.x - in.u, vOcnRel = in.v - in.ocn.y;
    const double absocn = sqrt(uOcnRel * uOcnRel + vOcnRel * vOcnRel);
    const double c1 = in.cf.x, drag = in.cf.y * absocn; // both times ML
    const double cor = c1 * a.dt * a.fc;                 // ML rho_ice H fc
    const double inv = 1.0 / (c1 * (1.0 + a.beta) + drag);
    un = inv * (c1 * a.beta * in.u + in.qf.x + drag * in.ocn.x - cor * in.u + dSx);
    vn = inv * (c1 * a.beta * in.v + in.qf.y + drag * in.ocn.y + cor * in.v + dSy);
}

// Fused subcycle: 31 x 7 elements, so that tile and halo are exactly one element per thread.
template <int CG> using FusedTile = TileShape<CG, 31, 7>;
static_assert(FusedTile<2>::HX * FusedTile<2>::HY == kBlock, "fused tile must give one element per thread");

// Node update policies of nodeTileUpdate: In is what a node loads before its update;
// kAllNodes runs the update on Dirichlet nodes too (their divergence is zero, their
// velocity is zeroed after the update).
// mEVP (VPCGDynamicsKernel.hpp:132-172) from the frozen coefficients; Dirichlet nodes skip it.
struct MevpNode {
    using In = NodeFz;
    static constexpr bool kAllNodes = false;
    __device__ __forceinline__ void load(const NodeArgs& a, int n, int, int, In& in) const { loadFrozen(a, n, in); }
    __device__ __forceinline__ void update(const NodeArgs& a, const In& in, int, bool, double dx, double dy,
        double& un, double& vn) const
    {
        momentumFrozen(a, in, dx, dy, un, vn);
    }
};

// Phase 2 of a node tile: the owned nodes sum their <= 4 contributions (contrib(le, j, c):
// local element le, its node j, component c; two-colour stripe order, land contributions
// exact zeros) and take the FP64 momentum update of the policy P; u, v are read from
// a.u, a.v and written to uOut, vOut. Each thread updates up to NB
// nodes of one column, all their loads issued first (Dirichlet nodes load theirs too, so
// that no load waits for the Dirichlet flag).
template <int CG, class T, class P = MevpNode, class G>
__device__ __forceinline__ void nodeTileUpdate(const NodeArgs& a, const G& contrib, int ex0, int ey0, double* uOut,
    double* vOut, const P& pol = P {})
{
    const int row = CG * a.nx + 1;
    const int nx0 = CG * ex0, ny0 = CG * ey0;
    const int W = (ex0 + T::TX >= a.nx) ? CG * a.nx + 1 - nx0 : T::CW;
    const int H = (ey0 + T::TY >= a.ny) ? CG * a.ny + 1 - ny0 : CG * T::TY;
    constexpr int NB = T::NB;
    if (threadIdx.x >= T::TPR * T::RS)
        return;
    for (int c = threadIdx.x % T::TPR; c < W; c += T::TPR) {
        for (int r0 = threadIdx.x / T::TPR; r0 < H; r0 += NB * T::RS) {
            typename P::In in[NB];
            bool live[NB], dir[NB];
            const int nix = nx0 + c;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int r = r0 + b * T::RS;
                live[b] = r < H;
                const int n = nix + row * (ny0 + r);
                dir[b] = live[b] ? a.dirNode[n] != 0 : true;
                if (live[b])
                    pol.load(a, n, nix, ny0 + r, in[b]);
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (!live[b])
                    continue;
                const int niy = ny0 + r0 + b * T::RS;
                double un = 0.0, vn = 0.0;
                double dx = 0.0, dy = 0.0;
                if (!dir[b]) {
                    forEachAdjacentXY<CG, true>(nix, niy, a.nx, a.ny, [&](int ex, int ey, int j) {
                        const int le = (ex - ex0 + 1) + T::HX * (ey - ey0 + 1);
                        dx -= static_cast<double>(contrib(le, j, 0));
                        dy -= static_cast<double>(contrib(le, j, 1));
                    });
                }
                if (P::kAllNodes || !dir[b])
                    pol.update(a, in[b], nix + row * niy, dir[b], dx, dy, un, vn);
                uOut[nix + row * niy] = un;
                vOut[nix + row * niy] = vn;
            }
        }
    }
}

// One whole mEVP subcycle in one pass over a uniform mesh (FP64): per tile, phase 1 runs
// the element update (strain, stress, projection) of every element of the tile and its
// left/bottom halo, keeps the new stress in registers, stores it for the tile's own
// elements only (halo elements belong to other tiles, which compute them identically),
// and turns it into weak-divergence contributions in LDS; phase 2 is the node update.
// Stress and velocities are ping-ponged: the pass reads sigma^k, u^k, v^k (e.sig, n.u,
// n.v) and writes sigma^(k+1), u^(k+1), v^(k+1) to other buffers, so no tile reads a
// neighbour's new values. The arithmetic is that of k_mevp_element, k_divergence,
// k_momentum and k_boundaries in sequence.
// R = float: the FP32 variant (stress stored as written by the two-pass FP32 kernels,
// stochastically rounded; the divergence uses the stored values).
struct FusedArgs {
    ElemArgs e;
    NodeArgs n;
    double* sigOut;
    float* sigFOut;
    double* uOut;
    double* vOut;
    const double* geo; // GeoLayout tables of a general mesh (k_mevp_fused_geom)
};
template <int CG, int DG, class R> __global__ __launch_bounds__(kBlock) void k_mevp_fused(FusedArgs f)
{
    using T = FusedTile<CG>;
    constexpr int N = T::N, S = cg2dgstress(CG);
    __shared__ R contrib[T::HX * T::HY][N][2];
    const NodeArgs& a = f.n;
    const int tilesX = (a.nx + T::TX - 1) / T::TX, tilesY = (a.ny + T::TY - 1) / T::TY;
    const int tile = xcdTile(blockIdx.x, tilesX * tilesY);
    if (tile < 0)
        return;
    const int ex0 = (tile % tilesX) * T::TX, ey0 = (tile / tilesX) * T::TY;
    {
        const int le = threadIdx.x; // one element of tile + halo per thread
        const int lx = le % T::HX, ly = le / T::HX;
        const int ex = ex0 - 1 + lx, ey = ey0 - 1 + ly;
        R tx[N], ty[N];
#pragma unroll
        for (int j = 0; j < N; ++j)
            tx[j] = ty[j] = 0;
        if (ex >= 0 && ex < a.nx && ey >= 0 && ey < a.ny) {
            const int e = ex + a.nx * ey;
            Acc<R> acc1[S], acc2[S], acc3[S];
            elementAffineCompute<CG, DG, R, R, false>(f.e, e, acc1, acc2, acc3);
            R s11[S], s12[S], s22[S];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                s11[s] = storeValue<R>(acc1[s], e, 0 * S + s, f.e.iter);
                s12[s] = storeValue<R>(acc2[s], e, 1 * S + s, f.e.iter);
                s22[s] = storeValue<R>(acc3[s], e, 2 * S + s, f.e.iter);
            }
            if (lx >= 1 && ly >= 1) {
                if constexpr (sizeof(R) == 8)
                    storeTensor<S>(f.sigOut, a.ne, e, s11, s12, s22);
                else
                    storeTensor<S>(f.sigFOut, a.ne, e, s11, s12, s22);
            }
            if (a.ocean[e])
                divergenceAffineFrom<CG, R>(a, s11, s12, s22, tx, ty);
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            contrib[le][j][0] = tx[j];
            contrib[le][j][1] = ty[j];
        }
    }
    __syncthreads();
    nodeTileUpdate<CG, T>(
        a, [&](int le, int j, int c) { return contrib[le][j][c]; }, ex0, ey0, f.uOut, f.vOut);
}

// The node values the momentum residual reads.
struct NodeRes {
    double u, v, uo, vo, ua, va, H, A, ML, u0, v0, gx, gy;
};
__device__ __forceinline__ void loadRes(const NodeArgs& a, int n, NodeRes& x)
{
    x.u = a.u[n];
    x.v = a.v[n];
    x.uo = a.uo[n];
    x.vo = a.vo[n];
    x.ua = a.ua[n];
    x.va = a.va[n];
    x.H = a.H[n];
    x.A = a.A[n];
    x.ML = a.ML[n];
    x.u0 = a.u0[n];
    x.v0 = a.v0[n];
    x.gx = a.gx[n];
    x.gy = a.gy[n];
}
// R_x^2 + R_y^2 and (M_L A tau_a)^2 of a node with values x, added to r2 and d2.
__device__ __forceinline__ void residualFrom(const NodeArgs& a, const NodeRes& x, double dSx, double dSy, double& r2,
    double& d2)
{
    const double u = x.u, v = x.v, uo = x.uo, vo = x.vo, ua = x.ua, va = x.va;
    const double H = x.H, A = x.A, ML = x.ML;
    const double absatm = sqrt(ua * ua + va * va);
    const double du = uo - u, dv = v - vo;
    const double absocn = sqrt(du * du + dv * dv);
    const double c1 = a.rho_ice * H / a.dt, rhoH = a.rho_ice * H;
    const double tax = a.F_atm * absatm * ua, tay = a.F_atm * absatm * va;
    const double rx = ML
            * (c1 * (u - x.u0) - A * (tax + a.F_ocean * absocn * (uo - u)) + rhoH * a.fc * u
                + rhoH * a.gravity * x.gx)
        - dSx;
    const double ry = ML
            * (c1 * (v - x.v0) - A * (tay + a.F_ocean * absocn * (vo - v)) - rhoH * a.fc * v
                + rhoH * a.gravity * x.gy)
        - dSy;
    const double fx = ML * A * tax, fy = ML * A * tay;
    r2 += rx * rx + ry * ry;
    d2 += fx * fx + fy * fy;
}
__device__ __forceinline__ void residualNode(const NodeArgs& a, int n, double dSx, double dSy, double& r2, double& d2)
{
    NodeRes x;
    loadRes(a, n, x);
    residualFrom(a, x, dSx, dSy, r2, d2);
}
// The momentum residual in one pass over a uniform mesh without periodic segments: the
// element phase of k_mevp_fused with alpha = 1 and no previous stress (Pi sigma_VP(eps(u)),
// as k_vp_stress), its weak divergence into LDS, then every owned non-Dirichlet node forms
// R from the LDS contributions (in nodeTileUpdate's order) and adds R^2 and
// (M_L A tau_a)^2 to its thread's compensated sums; one (sum, compensation) pair per field
// and block for k_sum_final<2>. Replaces k_strength, k_vp_stress, k_divergence and
// k_residual_nodes: the Gauss-point strength comes from the element's h and A in the
// element phase, and the stress and divergence never touch HBM.
template <int CG, int DG> __global__ __launch_bounds__(kBlock) void k_residual_fused(FusedArgs f, double* part)
{
    using T = FusedTile<CG>;
    constexpr int N = T::N, S = cg2dgstress(CG);
    __shared__ double contrib[T::HX * T::HY][N][2];
    __shared__ double red[4][kBlock];
    const NodeArgs& a = f.n;
    const int tilesX = (a.nx + T::TX - 1) / T::TX, tilesY = (a.ny + T::TY - 1) / T::TY;
    const int tile = xcdTile(blockIdx.x, tilesX * tilesY);
    double sum[2] = { 0.0, 0.0 }, comp[2] = { 0.0, 0.0 };
    if (tile >= 0) {
        const int ex0 = (tile % tilesX) * T::TX, ey0 = (tile / tilesX) * T::TY;
        {
            const int le = threadIdx.x;
            const int ex = ex0 - 1 + le % T::HX, ey = ey0 - 1 + le / T::HX;
            double tx[N], ty[N];
#pragma unroll
            for (int j = 0; j < N; ++j)
                tx[j] = ty[j] = 0.0;
            if (ex >= 0 && ex < a.nx && ey >= 0 && ey < a.ny) {
                const int e = ex + a.nx * ey;
                if (a.ocean[e]) {
                    double pk[MomLayout<CG>::Q];
                    gaussStrength<CG, DG>(f.e.h, f.e.A, f.e.ne, e, f.e.Pstar, pk);
                    Acc<double> acc1[S], acc2[S], acc3[S];
                    elementAffineCompute<CG, DG, double, double, true>(
                        f.e, e, acc1, acc2, acc3, [&](int q) -> double { return pk[q]; });
                    double s11[S], s12[S], s22[S];
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        s11[k] = acc1[k].get();
                        s12[k] = acc2[k].get();
                        s22[k] = acc3[k].get();
                    }
                    divergenceAffineFrom<CG, double>(a, s11, s12, s22, tx, ty);
                }
            }
#pragma unroll
            for (int j = 0; j < N; ++j) {
                contrib[le][j][0] = tx[j];
                contrib[le][j][1] = ty[j];
            }
        }
        __syncthreads();
        const int row = CG * a.nx + 1;
        const int nx0 = CG * ex0, ny0 = CG * ey0;
        const int W = (ex0 + T::TX >= a.nx) ? row - nx0 : CG * T::TX;
        const int H = (ey0 + T::TY >= a.ny) ? CG * a.ny + 1 - ny0 : CG * T::TY;
        // NBB nodes per thread at a time, all their values loaded before any is used
        // (Dirichlet nodes load theirs too and skip the sum); a thread still adds its
        // nodes in increasing order
        constexpr int NBB = 2;
        for (int base = threadIdx.x; base < W * H; base += NBB * kBlock) {
            NodeRes x[NBB];
            bool use[NBB];
#pragma unroll
            for (int bb = 0; bb < NBB; ++bb) {
                const int idx = base + bb * kBlock;
                use[bb] = idx < W * H;
                if (use[bb]) {
                    const int n = (nx0 + idx % W) + row * (ny0 + idx / W);
                    use[bb] = a.dirNode[n] == 0;
                    loadRes(a, n, x[bb]);
                }
            }
#pragma unroll
            for (int bb = 0; bb < NBB; ++bb) {
                if (!use[bb])
                    continue;
                const int idx = base + bb * kBlock;
                const int nix = nx0 + idx % W, niy = ny0 + idx / W;
                double dx = 0.0, dy = 0.0;
                forEachAdjacentXY<CG, true>(nix, niy, a.nx, a.ny, [&](int ex, int ey, int j) {
                    const int le = (ex - ex0 + 1) + T::HX * (ey - ey0 + 1);
                    dx -= contrib[le][j][0];
                    dy -= contrib[le][j][1];
                });
                double r2 = 0.0, d2 = 0.0;
                residualFrom(a, x[bb], dx, dy, r2, d2);
                double t, err;
                twoSum(sum[0], r2, t, err);
                sum[0] = t;
                comp[0] += err;
                twoSum(sum[1], d2, t, err);
                sum[1] = t;
                comp[1] += err;
            }
        }
    }
    blockTwoSum<2>(red, sum, comp);
    if (threadIdx.x == 0)
        for (int k = 0; k < 2; ++k) {
            part[(blockIdx.x * 2 + k) * 2] = sum[k];
            part[(blockIdx.x * 2 + k) * 2 + 1] = comp[k];
        }
}

// Two mEVP subcycles in one pass over a uniform mesh (FP64), the state read from and
// written to HBM once per two subcycles. A tile owns TX x TY = 13 x 13 elements.
// - Sub-step 1 updates the tile's elements with a two-element left/bottom and a
//   one-element right/top halo, 16 x 16, one per thread, from sigma^k, u^k, v^k in HBM;
//   each thread keeps its sigma^(k+1) in registers, and the block updates the CG nodes
//   of the inner 14 x 14 elements (29 x 29 nodes for CG2) into LDS.
// - Sub-step 2 is the one-pass subcycle of those inner elements (the tile and its
//   one-element left/bottom halo) from the registers and LDS: it stores sigma^(k+2) of the
//   owned elements and u^(k+2), v^(k+2) of the owned nodes.
// Sub-step 2 parks sigma^(k+1) in LDS (the sub-step 1 contributions are consumed by then)
// while its strain and stress sums run. Every value is computed by the arithmetic of
// k_mevp_fused, in the same order, from the same operands (a halo value identically in
// every tile that computes it); the compiler may still fuse a multiply-add differently
// in the two kernels, which moves a value by an ulp.
// FP64: 13 x 13 owned elements on 256 threads (51 KB of LDS: three blocks per CU). FP32,
// bound by instruction issue and with half the LDS per value: 29 x 13 on 512 threads, 2.47
// element updates per owned element and two subcycles instead of 2.67.
template <int CG, class R> struct PairTile {
    static constexpr int TX = sizeof(R) == 8 ? 13 : 29, TY = 13;
    static constexpr int W1 = TX + 3, H1 = TY + 3;          // sub-step 1 elements
    static constexpr int W2 = TX + 1, H2 = TY + 1;          // sub-step 2 elements
    static constexpr int NW = CG * W2 + 1, NH = CG * H2 + 1; // the CG nodes of those
    static constexpr int N = cgdofs(CG);
    static constexpr int BLOCK = W1 * H1;                    // one sub-step 1 element per thread
    using Inner = TileShape<CG, TX, TY, BLOCK>;
};
static_assert(PairTile<2, double>::BLOCK == kBlock && PairTile<2, float>::BLOCK == 2 * kBlock,
    "pair tiles of 256 (FP64) and 512 (FP32) threads");
static_assert(PairTile<2, double>::Inner::HX == PairTile<2, double>::W2
        && PairTile<2, double>::Inner::HY == PairTile<2, double>::H2,
    "sub-step 2 is the one-pass tile of the owned elements");
// mEVP node update (as MevpNode) with the velocity entering it from the LDS node window of
// k_mevp_pair (origin x0, y0 in global node coordinates, width nw).
struct MevpNodeWindow {
    using In = NodeFz;
    static constexpr bool kAllNodes = false;
    const double* uS;
    const double* vS;
    int x0, y0, nw;
    __device__ __forceinline__ void load(const NodeArgs& a, int n, int nix, int niy, In& in) const
    {
        const int l = (nix - x0) + nw * (niy - y0);
        in.u = uS[l];
        in.v = vS[l];
        in.ocn = ld2At(a.ocn, static_cast<uint32_t>(n) * 16u);
        in.cf = ld2At(a.cf, static_cast<uint32_t>(n) * 16u);
        in.qf = ld2At(a.qf, static_cast<uint32_t>(n) * 16u);
    }
    __device__ __forceinline__ void update(const NodeArgs& a, const In& in, int, bool, double dx, double dy,
        double& un, double& vn) const
    {
        momentumFrozen(a, in, dx, dy, un, vn);
    }
};
// R = float: the FP32 variant, two of k_mevp_fused<CG, DG, float>'s subcycles in one
// pass (the stored stress of each sub-step stochastically rounded with the subcycle
// indices iter and iter + 1, as two single passes round it; the divergence and sub-step 2
// use the rounded values; the node update stays FP64).
template <int CG, int DG, class R>
__global__ __launch_bounds__((PairTile<CG, R>::BLOCK))
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 8 ? 3 : 1))) void k_mevp_pair(FusedArgs f)
{
    using T = PairTile<CG, R>;
    constexpr int N = T::N, S = cg2dgstress(CG), K = CG + 1;
    constexpr int NWH = T::NW * T::NH, NB = (NWH + T::BLOCK - 1) / T::BLOCK;
    // the sub-step contributions, and sigma^(k+1) parked during sub-step 2, in one array
    constexpr int CWORDS = T::W1 * T::H1 * N * 2, SWORDS = 3 * S * T::W2 * T::H2;
    __shared__ R contribRaw[CWORDS > SWORDS ? CWORDS : SWORDS];
    auto& contrib = *reinterpret_cast<R(*)[T::W1 * T::H1][N][2]>(contribRaw);
    __shared__ double uS[NWH], vS[NWH];
    const NodeArgs& a = f.n;
    const int tilesX = (a.nx + T::TX - 1) / T::TX, tilesY = (a.ny + T::TY - 1) / T::TY;
    const int tile = xcdTile(blockIdx.x, tilesX * tilesY);
    if (tile < 0)
        return;
    const int ex0 = (tile % tilesX) * T::TX, ey0 = (tile / tilesX) * T::TY;
    const int x0 = CG * (ex0 - 1), y0 = CG * (ey0 - 1); // node window origin
    const int row = CG * a.nx + 1;
    const int lx = threadIdx.x % T::W1, ly = threadIdx.x / T::W1;
    const int ex = ex0 - 2 + lx, ey = ey0 - 2 + ly;
    const bool inDomain = ex >= 0 && ex < a.nx && ey >= 0 && ey < a.ny;
    const int e = inDomain ? ex + a.nx * ey : 0;
    R s11[S], s12[S], s22[S]; // sigma^(k+1) of this thread's element
    // the element's Gauss-point strength for sub-step 1. Sub-step 2 reads it again in FP64
    // (from the L2): the 18 VGPRs it frees give the FP64 pass a third block per CU
    // (occupancy 3, 168 VGPRs; LDS 3 x 50 KB). FP32 keeps it in registers.
    constexpr int Q = MomLayout<CG>::Q;
    R pk[Q];
    if (inDomain) {
        const R* const Pg = strengthOf<R>(f.e);
        const uint32_t eoff = static_cast<uint32_t>(e) * static_cast<uint32_t>(sizeof(R));
#pragma unroll
        for (int q = 0; q < Q; ++q)
            pk[q] = ldAt(Pg + static_cast<uint32_t>(q) * static_cast<uint32_t>(f.e.ne), eoff); // < 2^32: ne < 2^29
    }
    const auto kept = [&](int q) -> R { return pk[q]; };
    const auto strength2 = [&](int q) -> R {
        if constexpr (sizeof(R) == 8)
            return ldAt(strengthOf<R>(f.e) + static_cast<uint32_t>(q) * static_cast<uint32_t>(f.e.ne),
                static_cast<uint32_t>(e) * static_cast<uint32_t>(sizeof(R)));
        else
            return pk[q];
    };
    {
        R tx[N], ty[N];
#pragma unroll
        for (int j = 0; j < N; ++j)
            tx[j] = ty[j] = 0;
        if (inDomain) {
            Acc<R> acc1[S], acc2[S], acc3[S];
            elementAffineCompute<CG, DG, R, R, false>(f.e, e, acc1, acc2, acc3, kept);
#pragma unroll
            for (int k = 0; k < S; ++k) {
                s11[k] = storeValue<R>(acc1[k], e, 0 * S + k, f.e.iter);
                s12[k] = storeValue<R>(acc2[k], e, 1 * S + k, f.e.iter);
                s22[k] = storeValue<R>(acc3[k], e, 2 * S + k, f.e.iter);
            }
            if (a.ocean[e])
                divergenceAffineFrom<CG, R>(a, s11, s12, s22, tx, ty);
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            contrib[threadIdx.x][j][0] = tx[j];
            contrib[threadIdx.x][j][1] = ty[j];
        }
    }
    __syncthreads();
    // sub-step 1, the nodes of the inner elements (within the domain) into LDS, in batches
    // of NBB nodes per thread (each batch's frozen terms loaded before its gathers, for
    // Dirichlet nodes too, so that the loads do not wait for the Dirichlet flag)
    constexpr int NBB = 2;
#pragma unroll
    for (int b0 = 0; b0 < NB; b0 += NBB) {
        NodeFz in[NBB];
        bool live[NBB], dir[NBB];
#pragma unroll
        for (int bb = 0; bb < NBB; ++bb) {
            const int i = threadIdx.x + (b0 + bb) * T::BLOCK;
            const int nix = x0 + i % T::NW, niy = y0 + i / T::NW;
            live[bb] = b0 + bb < NB && i < NWH && nix >= 0 && nix <= CG * a.nx && niy >= 0 && niy <= CG * a.ny;
            const int n = nix + row * niy;
            dir[bb] = live[bb] ? a.dirNode[n] != 0 : true;
            if (live[bb])
                loadFrozenAt(a, static_cast<uint32_t>(n), in[bb]);
        }
#pragma unroll
        for (int bb = 0; bb < NBB; ++bb) {
            const int b = b0 + bb;
            if (!live[bb])
                continue;
            const int i = threadIdx.x + b * T::BLOCK;
            const int nix = x0 + i % T::NW, niy = y0 + i / T::NW;
            double un = 0.0, vn = 0.0;
            if (!dir[bb]) {
                double dx = 0.0, dy = 0.0;
                forEachAdjacentXY<CG, true>(nix, niy, a.nx, a.ny, [&](int exa, int eya, int j) {
                    const int le = (exa - ex0 + 2) + T::W1 * (eya - ey0 + 2);
                    dx -= static_cast<double>(contrib[le][j][0]);
                    dy -= static_cast<double>(contrib[le][j][1]);
                });
                momentumFrozen(a, in[bb], dx, dy, un, vn);
            }
            uS[i] = un;
            vS[i] = vn;
        }
    }
    __syncthreads();
    // sub-step 2, the inner elements, from sigma^(k+1) parked in LDS and the LDS velocities
    {
        constexpr int NI = T::W2 * T::H2;
        R* const sigS = contribRaw; // 3 S NI <= SWORDS values
        const bool inner = lx >= 1 && lx <= T::W2 && ly >= 1 && ly <= T::H2;
        const int le2 = (lx - 1) + T::W2 * (ly - 1);
        const bool active = inner && inDomain;
        if (active) {
#pragma unroll
            for (int k = 0; k < S; ++k) {
                sigS[(0 * S + k) * NI + le2] = s11[k];
                sigS[(1 * S + k) * NI + le2] = s12[k];
                sigS[(2 * S + k) * NI + le2] = s22[k];
            }
        }
        R tx[N], ty[N];
#pragma unroll
        for (int j = 0; j < N; ++j)
            tx[j] = ty[j] = 0;
        if (active) {
            const int bx = CG * (ex - ex0 + 1), by = CG * (ey - ey0 + 1); // first node in the window
            Acc<R> acc1[S], acc2[S], acc3[S];
            elementAffineCore<CG, DG, R, R, false>(
                f.e, e, acc1, acc2, acc3,
                [&](double* lu, double* lv) {
#pragma unroll
                    for (int r = 0; r < K; ++r)
#pragma unroll
                        for (int c = 0; c < K; ++c) {
                            lu[c + K * r] = uS[bx + c + T::NW * (by + r)];
                            lv[c + K * r] = vS[bx + c + T::NW * (by + r)];
                        }
                },
                [&](int c, int k) -> R { return sigS[(c * S + k) * NI + le2]; }, strength2);
            R n11[S], n12[S], n22[S];
#pragma unroll
            for (int k = 0; k < S; ++k) {
                n11[k] = storeValue<R>(acc1[k], e, 0 * S + k, f.e.iter + 1);
                n12[k] = storeValue<R>(acc2[k], e, 1 * S + k, f.e.iter + 1);
                n22[k] = storeValue<R>(acc3[k], e, 2 * S + k, f.e.iter + 1);
            }
            if (lx >= 2 && ly >= 2) {
                const uint32_t eoff = static_cast<uint32_t>(e) * static_cast<uint32_t>(sizeof(R));
                if constexpr (sizeof(R) == 8)
                    storeTensorAt<S>(f.sigOut, a.ne, eoff, n11, n12, n22);
                else
                    storeTensorAt<S>(f.sigFOut, a.ne, eoff, n11, n12, n22);
            }
            if (a.ocean[e])
                divergenceAffineFrom<CG, R>(a, n11, n12, n22, tx, ty);
        }
        __syncthreads(); // every parked stress is read before the contributions overwrite it
        if (inner) {
#pragma unroll
            for (int j = 0; j < N; ++j) {
                contrib[le2][j][0] = tx[j];
                contrib[le2][j][1] = ty[j];
            }
        }
    }
    __syncthreads();
    nodeTileUpdate<CG, typename T::Inner, MevpNodeWindow>(
        a, [&](int le, int j, int c) { return contrib[le][j][c]; }, ex0, ey0, f.uOut, f.vOut,
        MevpNodeWindow { uS, vS, x0, y0, T::NW });
}

// The strain coefficients of one element's velocity on a general mesh (CGDynamicsKernel.cpp:
// 300-337, with the spherical metric terms) from its geometry: e = M^-1 sum_q PSI(q) g(q)
// with g the J-weighted velocity gradients at the Gauss points, sum-factorised from the
// 1-D CG basis. mAt(g, k, off) reads entry k of the packed inverse mass (and may keep it).
template <int CG, bool SPH, class MAT>
__device__ __forceinline__ void strainGeom(const ElemArgs& a, const double* geo, int e, double* e11, double* e12,
    double* e22, const MAT& mAt)
{
    using L = GeoLayout<CG>;
    constexpr int S = L::S, NG = L::G, N = cgdofs(CG), K = CG + 1;
    const size_t ne = a.ne;
    const uint32_t off = static_cast<uint32_t>(e) * 8u;
    const double* const g1 = phaseLocal(geo);
    const MapCoef map = loadMap<CG>(g1, ne, off);
    double lu[N], lv[N];
    {
        const int ix = e % a.nx, iy = e / a.nx, row = CG * a.nx + 1;
        const uint32_t noff = static_cast<uint32_t>(CG * row * iy + CG * ix) * 8u;
#pragma unroll
        for (int r = 0; r < K; ++r)
#pragma unroll
            for (int c = 0; c < K; ++c) {
                lu[c + K * r] = ldAt(a.u + c + r * row, noff);
                lv[c + K * r] = ldAt(a.v + c + r * row, noff);
            }
    }
    double t11[S], t12[S], t22[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        t11[s] = t12[s] = t22[s] = 0.0;
#pragma unroll
    for (int gy = 0; gy < NG; ++gy) {
        double Iu[K], Du[K], Iv[K], Dv[K];
#pragma unroll
        for (int jx = 0; jx < K; ++jx) {
            double iu = 0, du = 0, iv = 0, dv = 0;
#pragma unroll
            for (int jy = 0; jy < K; ++jy) {
                const double ph = cg_1d(CG, jy, gauss_point(NG, gy)), dp = cg_1d_dx(CG, jy, gauss_point(NG, gy));
                iu += ph * lu[jx + K * jy];
                du += dp * lu[jx + K * jy];
                iv += ph * lv[jx + K * jy];
                dv += dp * lv[jx + K * jy];
            }
            Iu[jx] = iu;
            Du[jx] = du;
            Iv[jx] = iv;
            Dv[jx] = dv;
        }
#pragma unroll
        for (int gx = 0; gx < NG; ++gx) {
            const int q = gx + NG * gy;
            double uxi = 0, ueta = 0, vxi = 0, veta = 0, uq = 0, vq = 0;
#pragma unroll
            for (int jx = 0; jx < K; ++jx) {
                const double ph = cg_1d(CG, jx, gauss_point(NG, gx)), dp = cg_1d_dx(CG, jx, gauss_point(NG, gx));
                uxi += dp * Iu[jx];
                ueta += ph * Du[jx];
                vxi += dp * Iv[jx];
                veta += ph * Dv[jx];
                if constexpr (SPH) {
                    uq += ph * Iu[jx];
                    vq += ph * Iv[jx];
                }
            }
            const GaussGeom gg = gaussGeom<NG, SPH>(map, gx, gy, SPH ? ldAt(g1 + (L::CL + q) * ne, off) : 1.0,
                SPH ? ldAt(g1 + (L::SL + q) * ne, off) : 0.0);
            double g11 = gg.a * uxi - gg.b * ueta;
            const double g22 = gg.c * veta - gg.d * vxi;
            double g12 = (gg.a * vxi - gg.b * veta) + (gg.c * ueta - gg.d * uxi);
            if constexpr (SPH) {
                g11 -= gg.js * vq;
                g12 += gg.js * uq;
            }
            g12 *= 0.5;
            double bd[8];
            dgBasisAt(gauss_point(NG, gx) - 0.5, gauss_point(NG, gy) - 0.5, bd);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (bd[s] == 0.0) // basis functions vanishing at this Gauss point (a literal test)
                    continue;
                t11[s] += bd[s] * g11;
                t12[s] += bd[s] * g12;
                t22[s] += bd[s] * g22;
            }
        }
    }
    const uint32_t offM = after(after(after(off, t11[S - 1]), t12[S - 1]), t22[S - 1]);
    inverseMassApply<CG>([&](int k) { return mAt(g1, k, offM); }, t11, t12, t22, e11, e12, e22);
}
// The stored strain of the ocean elements from the geometry (the strain output before the
// last subcycle of a run on a general mesh); land elements keep theirs.
template <int CG, bool SPH> __global__ __launch_bounds__(kBlock) void k_strain_geom(ElemArgs a, const double* geo)
{
    using L = GeoLayout<CG>;
    constexpr int S = L::S;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.ne || !a.ocean[e])
        return;
    const size_t ne = a.ne;
    double e11[S], e12[S], e22[S];
    strainGeom<CG, SPH>(a, geo, e, e11, e12, e22, [&](const double* g1, int k, uint32_t offM) {
        return ldAt(g1 + static_cast<size_t>(L::MI + k) * ne, offM);
    });
    storeTensorAt<S>(a.epsOut, ne, static_cast<uint32_t>(e) * 8u, e11, e12, e22);
}
// The element update of one mEVP subcycle on a general mesh from its geometry (GeoLayout):
//  * strain (ocean elements): e = M^-1 sum_q PSI(q) g(q), with g the J-weighted velocity
//    gradients at q, sum-factorised from the 1-D CG basis, and on the sphere the metric
//    terms (CGDynamicsKernel.cpp:300-337: e11 = iMgradX u - iMM v, e22 = iMgradY v,
//    e12 = (iMgradX v + iMgradY u + iMM u) / 2); land elements keep their stored strain;
//  * stress (MEVPStressUpdateStep.hpp:30-118): sigma_VP at the Gauss points from the
//    strain evaluated there, then sigma <- (1 - 1/alpha) sigma + M^-1 sum_q PSI(q) J w
//    sigma_VP(q) / alpha.
// The same operators as the per-element tables of k_momentum_ops, applied in factored form.
// col(k), k < S (S + 1) / 2: a per-thread scratch column (LDS) that keeps the inverse mass
// matrix from the strain to the stress projection.
template <int CG, bool SPH, class COL>
__device__ __forceinline__ void elementGeomUpdate(const ElemArgs& a, const double* geo, int e, double* s11,
    double* s12, double* s22, const COL& col)
{
    using L = GeoLayout<CG>;
    constexpr int S = L::S, NG = L::G;
    const size_t ne = a.ne;
    const uint32_t off = static_cast<uint32_t>(e) * 8u;
    // The strain of the velocity is computed for every element (the nodes of land elements
    // exist too) and replaced by the stored one on land, in a branch taken by whole waves:
    // no divergent control flow around the strain computation.
    double e11[S], e12[S], e22[S];
    strainGeom<CG, SPH>(a, geo, e, e11, e12, e22, [&](const double* g1, int k, uint32_t offM) {
        const double m = ldAt(g1 + static_cast<size_t>(L::MI + k) * ne, offM);
        col(k) = m;
        return m;
    });
    if (__builtin_amdgcn_ballot_w64(!a.ocean[e]) != 0) { // a wave with land: uniform branch
        const bool land = !a.ocean[e];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double x = ldAt(a.eps + (0 * S + s) * ne, off), y = ldAt(a.eps + (1 * S + s) * ne, off),
                         z = ldAt(a.eps + (2 * S + s) * ne, off);
            e11[s] = land ? x : e11[s];
            e12[s] = land ? y : e12[s];
            e22[s] = land ? z : e22[s];
        }
    }
    const double* const g2 = phaseLocal(geo);
    const uint32_t offS = after(after(after(off, e11[S - 1]), e12[S - 1]), e22[S - 1]);
    const MapCoef map2 = loadMap<CG>(g2, ne, offS);
    double u11[S], u12[S], u22[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        u11[s] = u12[s] = u22[s] = 0.0;
#pragma unroll
    for (int q = 0; q < NG * NG; ++q) {
        double bd[8];
        dgBasisAt(gauss_point(NG, q % NG) - 0.5, gauss_point(NG, q / NG) - 0.5, bd);
        double g11 = 0, g12 = 0, g22 = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            g11 += e11[s] * bd[s];
            g12 += e12[s] * bd[s];
            g22 += e22[s] * bd[s];
        }
        const double delta = sqrt(a.DM2 + 1.25 * (g11 * g11 + g22 * g22) + 1.50 * g11 * g22 + g12 * g12);
        const double P = ldAt(a.Pg + q * ne, offS);
        const double c = (P * 0.125) / delta; // P/8/Delta
        const double jw = gaussGeom<NG, false>(map2, q % NG, q / NG, 1.0, 0.0).jw * a.invAlpha;
        const double r11 = jw * (c * (5.0 * g11 + 3.0 * g22) - 0.5 * P);
        const double r12 = jw * ((2.0 * c) * g12);
        const double r22 = jw * (c * (5.0 * g22 + 3.0 * g11) - 0.5 * P);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (bd[s] == 0.0)
                continue;
            u11[s] += bd[s] * r11;
            u12[s] += bd[s] * r12;
            u22[s] += bd[s] * r22;
        }
    }
    // the inverse mass (from the scratch column) and the previous stress are read only now
    const uint32_t offM = after(after(after(off, u11[S - 1]), u12[S - 1]), u22[S - 1]);
    double x11[S], x12[S], x22[S];
    inverseMassApply<CG>([&](int k) { return static_cast<double>(col(k)); }, u11, u12, u22, x11, x12, x22);
    loadTensorAt<S>(a.sig, ne, after(offM, x22[S - 1]), s11, s12, s22);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        s11[s] = s11[s] * a.relax + x11[s];
        s12[s] = s12[s] * a.relax + x12[s];
        s22[s] = s22[s] * a.relax + x22[s];
    }
}
// Weak-divergence contributions of one element (CGDynamicsKernel.cpp:339-398 with the
// spherical terms tx += divM s12, ty -= divM s11) from its geometry and its stress
// coefficients in registers: the stress at the Gauss points, the fluxes
//   Fx = (a s11 - d s12, c s12 - b s11, JS s12), Fy = (a s12 - d s22, c s22 - b s12, -JS s11)
// against (phi_xi, phi_eta, phi), sum-factorised per Gauss row.
template <int CG, bool SPH>
__device__ __forceinline__ void divergenceGeomFrom(const double* geo, size_t ne, int e, const double* s11,
    const double* s12, const double* s22, double* tx, double* ty)
{
    using L = GeoLayout<CG>;
    constexpr int S = L::S, NG = L::G, N = cgdofs(CG), K = CG + 1;
    const double* const g4 = phaseLocal(geo);
    const uint32_t off = after(after(after(static_cast<uint32_t>(e) * 8u, s11[S - 1]), s12[S - 1]), s22[S - 1]);
    const MapCoef map = loadMap<CG>(g4, ne, off);
#pragma unroll
    for (int j = 0; j < N; ++j)
        tx[j] = ty[j] = 0.0;
#pragma unroll
    for (int gy = 0; gy < NG; ++gy) {
        double ax[K], bx[K], ay[K], by[K]; // per node column: sums over the Gauss row
#pragma unroll
        for (int jx = 0; jx < K; ++jx)
            ax[jx] = bx[jx] = ay[jx] = by[jx] = 0.0;
#pragma unroll
        for (int gx = 0; gx < NG; ++gx) {
            const int q = gx + NG * gy;
            double bd[8];
            dgBasisAt(gauss_point(NG, gx) - 0.5, gauss_point(NG, gy) - 0.5, bd);
            double v11 = 0, v12 = 0, v22 = 0;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                v11 += s11[s] * bd[s];
                v12 += s12[s] * bd[s];
                v22 += s22[s] * bd[s];
            }
            const GaussGeom gg = gaussGeom<NG, SPH>(map, gx, gy, SPH ? ldAt(g4 + (L::CL + q) * ne, off) : 1.0,
                SPH ? ldAt(g4 + (L::SL + q) * ne, off) : 0.0);
            const double fxXi = gg.a * v11 - gg.d * v12, fxEta = gg.c * v12 - gg.b * v11;
            const double fyXi = gg.a * v12 - gg.d * v22, fyEta = gg.c * v22 - gg.b * v12;
            double fx0 = 0, fy0 = 0;
            if constexpr (SPH) {
                fx0 = gg.js * v12;
                fy0 = -gg.js * v11;
            }
#pragma unroll
            for (int jx = 0; jx < K; ++jx) {
                const double ph = cg_1d(CG, jx, gauss_point(NG, gx)), dp = cg_1d_dx(CG, jx, gauss_point(NG, gx));
                ax[jx] += dp * fxXi;
                bx[jx] += ph * fxEta;
                ay[jx] += dp * fyXi;
                by[jx] += ph * fyEta;
                if constexpr (SPH) {
                    ax[jx] += ph * fx0;
                    ay[jx] += ph * fy0;
                }
            }
        }
#pragma unroll
        for (int jy = 0; jy < K; ++jy) {
            const double ph = cg_1d(CG, jy, gauss_point(NG, gy)), dp = cg_1d_dx(CG, jy, gauss_point(NG, gy));
#pragma unroll
            for (int jx = 0; jx < K; ++jx) {
                tx[jx + K * jy] += ph * ax[jx] + dp * bx[jx];
                ty[jx + K * jy] += ph * ay[jx] + dp * by[jx];
            }
        }
    }
}
// One whole mEVP subcycle in one pass over a general (non-affine or spherical) mesh without
// periodic segments, FP64: k_mevp_fused's tiles, halo and node phase, with the element
// update and the divergence computed from the streamed geometry instead of literals.
template <int CG, int DG, bool SPH> __global__ __launch_bounds__(kBlock) void k_mevp_fused_geom(FusedArgs f)
{
    using T = FusedTile<CG>;
    constexpr int N = T::N, S = cg2dgstress(CG), NM = S * (S + 1) / 2;
    // One column per local element: first the inverse mass matrix of its update (rows
    // < NM), then its divergence contributions (row 2 j + c: node j, component c). Each
    // thread reads and writes only its own column until the barrier.
    constexpr int kRows = NM > 2 * N ? NM : 2 * N;
    __shared__ double stage[kRows][kBlock];
    const NodeArgs& a = f.n;
    const int tilesX = (a.nx + T::TX - 1) / T::TX, tilesY = (a.ny + T::TY - 1) / T::TY;
    const int tile = xcdTile(blockIdx.x, tilesX * tilesY);
    if (tile < 0)
        return;
    const int ex0 = (tile % tilesX) * T::TX, ey0 = (tile / tilesX) * T::TY;
    {
        const int le = threadIdx.x; // one element of tile + halo per thread
        const int lx = le % T::HX, ly = le / T::HX;
        const int ex = ex0 - 1 + lx, ey = ey0 - 1 + ly;
        // Halo and tile elements outside the domain compute a valid neighbour's update
        // (clamped index) and contribute nothing: no divergent branch around the element.
        const bool inside = ex >= 0 && ex < a.nx && ey >= 0 && ey < a.ny;
        const int e = min(max(ex, 0), a.nx - 1) + a.nx * min(max(ey, 0), a.ny - 1);
        double s11[S], s12[S], s22[S], tx[N], ty[N];
        elementGeomUpdate<CG, SPH>(f.e, f.geo, e, s11, s12, s22, [&](int k) -> double& { return stage[k][le]; });
        if (inside && lx >= 1 && ly >= 1)
            storeTensorAt<S>(f.sigOut, a.ne, static_cast<uint32_t>(e) * 8u, s11, s12, s22);
        divergenceGeomFrom<CG, SPH>(f.geo, a.ne, e, s11, s12, s22, tx, ty);
        const bool adds = inside && a.ocean[e] != 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            stage[2 * j][le] = adds ? tx[j] : 0.0;
            stage[2 * j + 1][le] = adds ? ty[j] : 0.0;
        }
    }
    __syncthreads();
    nodeTileUpdate<CG, T>(
        a, [&](int le, int j, int c) { return stage[2 * j + c][le]; }, ex0, ey0, f.uOut, f.vOut);
}

// Split passes (periodic meshes and per-pass testing).
template <int CG, bool UNI> __global__ __launch_bounds__(kBlock) void k_divergence(NodeArgs a)
{
    using L = MomLayout<CG>;
    constexpr int S = L::S;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.nn)
        return;
    if (a.dirNode[n]) {
        a.dSx[n] = 0.0;
        a.dSy[n] = 0.0;
        return;
    }
    const Ops<UNI> op { a.ops, a.stride };
    const int row = CG * a.nx + 1;
    double dx = 0.0, dy = 0.0;
    forEachAdjacent<CG, true>(n % row, n / row, a.nx, a.ny, [&](int e, int j) {
        if (!a.ocean[e])
            return;
        double s11[S], s12[S], s22[S];
        loadTensor<S>(a.sig, a.ne, e, s11, s12, s22);
        double p1 = 0, p2 = 0, p3 = 0, p4 = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double d1 = op(L::D1 + j * S + s, e), d2 = op(L::D2 + j * S + s, e);
            p1 += d1 * s11[s];
            p2 += d2 * s12[s];
            p3 += d1 * s12[s];
            p4 += d2 * s22[s];
        }
        double tx = p1 + p2, ty = p3 + p4;
        if constexpr (!UNI) {
            if (a.sph) { // CGDynamicsKernel.cpp:348-351
                double m12 = 0, m11 = 0;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const double dm = op(L::DM + j * S + s, e);
                    m12 += dm * s12[s];
                    m11 += dm * s11[s];
                }
                tx += m12;
                ty -= m11;
            }
        }
        dx -= tx;
        dy -= ty;
    });
    a.dSx[n] = dx;
    a.dSy[n] = dy;
}
__global__ void k_momentum(NodeArgs a)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.nn)
        return;
    momentumNode(a, n, a.dSx[n], a.dSy[n]);
}
__global__ void k_boundaries(NodeArgs a)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.nn || !a.dirNode[n])
        return;
    a.u[n] = 0.0;
    a.v[n] = 0.0;
}

// ---------------------------------------------------------------- momentum residual
// Fixed-point residual of updateMomentum with the stress replaced by the converged VP
// stress of the current velocity (dS = the divergence of Pi sigma_VP(eps(u))):
//   R = M_L [c1 (u - u0) - A (tau_a + F_oc |du| (u_o - u)) + rho H f_c u + rho H g dEta/dx] - dS_x
//   R_y likewise with - rho H f_c v (the same-component Coriolis term of the update),
// c1 = rho H / dt, tau_a = F_atm |u_a| u_a. R = 0 exactly where the update
// (VPCGDynamicsKernel.hpp:132-172) leaves u unchanged. Sums of R^2 and of (M_L A tau_a)^2
// over the non-Dirichlet nodes: per-thread FP64 sums, wavefront shuffles, then one
// partial per block; k_sum_final<2> adds the block partials in a fixed tree order
// (deterministic).
constexpr int kResidualBlocks = 512;
constexpr unsigned kTallyBlocks = 1024;
__device__ __forceinline__ double waveSum(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        x += __shfl_xor(x, o, 64);
    return x;
}
__global__ __launch_bounds__(kBlock) void k_residual_nodes(NodeArgs a, const double* dSx, const double* dSy,
    double* part)
{
    __shared__ double red[2][kBlock / 64];
    double r2 = 0.0, d2 = 0.0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < a.nn; n += gridDim.x * blockDim.x) {
        if (a.dirNode[n])
            continue;
        residualNode(a, n, dSx[n], dSy[n], r2, d2);
    }
    r2 = waveSum(r2);
    d2 = waveSum(d2);
    const int w = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0) {
        red[0][w] = r2;
        red[1][w] = d2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sr = 0.0, sd = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) {
            sr += red[0][k];
            sd += red[1][k];
        }
        // the (sum, compensation) layout of k_sum_final<2>, compensation zero
        part[4 * blockIdx.x] = sr;
        part[4 * blockIdx.x + 1] = 0.0;
        part[4 * blockIdx.x + 2] = sd;
        part[4 * blockIdx.x + 3] = 0.0;
    }
}

// ---------------------------------------------------------------- allocation-free sums
// Scratch of nsdg_dyn_mass_noalloc: the two totals at [0, 2), block partials from
// kNoAllocPartials on. 1 MiB per device, held until the process ends; every call on a
// device shares it (k_mass_partial writes every partial it reads back and k_sum_final
// writes both totals, so nothing needs clearing between calls; calls that overlap on one
// device race).
constexpr size_t kNoAllocScratch = size_t(1) << 17;
constexpr size_t kNoAllocPartials = 32;
constexpr size_t kNoAllocMaxElements = (kNoAllocScratch - kNoAllocPartials) / 4 * kBlock;
static __device__ __align__(256) double scratch[kNoAllocScratch];

// Device address of scratch on device dev (each device holds its own copy), looked up
// once per device with that device current.
static double* noAllocScratch(int dev)
{
    static std::mutex mu;
    static std::map<int, double*> ptrs;
    std::lock_guard<std::mutex> lock(mu);
    auto it = ptrs.find(dev);
    if (it != ptrs.end())
        return it->second;
    int cur = -1;
    NSDG_HIP(hipGetDevice(&cur));
    if (cur != dev)
        NSDG_HIP(hipSetDevice(dev));
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
    if (cur != dev)
        NSDG_HIP(hipSetDevice(cur));
    NSDG_HIP(e);
    require(p != nullptr, "no device address for the allocation-free scratch", NSDG_ERR_HIP);
    ptrs[dev] = p;
    return p;
}

// ---------------------------------------------------------------- learned warm start
// Network input and target on the element grid of a CG2 kernel. Element e = ix + nx iy
// owns the CG nodes (2 ix + a, 2 iy + b), a, b in {0, 1}, listed k = a + 2 b; its centre
// node is (2 ix + 1, 2 iy + 1). Scales bring every channel to order one.
constexpr double kScaleU = 0.1;      // m/s, ice velocity and its change
constexpr double kScaleWind = 10.0;  // m/s
constexpr double kScaleOcean = 0.01; // m/s
constexpr double kScaleStress = 1e4; // N/m
__device__ __forceinline__ uint16_t bf16Of(double d)
{
    uint32_t u = __float_as_uint(static_cast<float>(d));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
template <int S>
__global__ __launch_bounds__(kBlock) void k_features(int nx, int ny, const double* h, const double* A,
    const uint8_t* ocean, const double* ua, const double* va, const double* uo, const double* vo, const double* u0,
    const double* v0, const double* sig, double Pstar, uint16_t* x)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int ne = nx * ny;
    if (e >= ne)
        return;
    const int ix = e % nx, iy = e / nx, row = 2 * nx + 1;
    const int centre = (2 * ix + 1) + row * (2 * iy + 1);
    double f[32];
#pragma unroll
    for (int c = 0; c < 32; ++c)
        f[c] = 0.0;
    const double hm = h[e], am = A[e];
    f[0] = hm;
    f[1] = am;
    f[2] = ocean[e] ? 1.0 : 0.0;
    f[3] = ua[centre] / kScaleWind;
    f[4] = va[centre] / kScaleWind;
    f[5] = uo[centre] / kScaleOcean;
    f[6] = vo[centre] / kScaleOcean;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int n = (2 * ix + k % 2) + row * (2 * iy + k / 2);
        f[7 + 2 * k] = u0[n] / kScaleU;
        f[8 + 2 * k] = v0[n] / kScaleU;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        f[15 + c] = sig[static_cast<size_t>(c * S) * ne + e] / kScaleStress;
    f[18] = Pstar * fmax(hm, 0.0) * exp(-20.0 * (1.0 - fmin(fmax(am, 0.0), 1.0))) / kScaleStress;
    uint4* dst = reinterpret_cast<uint4*>(x + static_cast<size_t>(e) * 32);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            w[k] = static_cast<uint32_t>(bf16Of(f[8 * q + 2 * k])) | (static_cast<uint32_t>(bf16Of(f[8 * q + 2 * k + 1])) << 16);
        dst[q] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}
__global__ __launch_bounds__(kBlock) void k_velocity_target(int nx, int ny, const double* u, const double* v,
    const double* u0, const double* v0, const uint8_t* ocean, float* t, float* w)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nx * ny)
        return;
    const int ix = e % nx, iy = e / nx, row = 2 * nx + 1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int n = (2 * ix + k % 2) + row * (2 * iy + k / 2);
        t[8 * static_cast<size_t>(e) + 2 * k] = static_cast<float>((u[n] - u0[n]) / kScaleU);
        t[8 * static_cast<size_t>(e) + 2 * k + 1] = static_cast<float>((v[n] - v0[n]) / kScaleU);
    }
    if (w)
        w[e] = ocean[e] ? 1.0f : 0.0f;
}
__global__ __launch_bounds__(kBlock) void k_apply_velocity(int nx, int ny, const float* y, const double* u0,
    const double* v0, const uint8_t* dirNode, double* u, double* v)
{
    const int row = 2 * nx + 1;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= row * (2 * ny + 1))
        return;
    const int ix = n % row, iy = n / row;
    double un = u0[n], vn = v0[n];
    if (ix < 2 * nx && iy < 2 * ny) {
        const size_t e = static_cast<size_t>(ix / 2) + static_cast<size_t>(nx) * (iy / 2);
        const int k = ix % 2 + 2 * (iy % 2);
        un += kScaleU * y[8 * e + 2 * k];
        vn += kScaleU * y[8 * e + 2 * k + 1];
    }
    if (dirNode[n])
        un = vn = 0.0;
    u[n] = un;
    v[n] = vn;
}

// ---------------------------------------------------------------- outputs and inputs
// getIceOceanStress (VPCGDynamicsKernel.hpp:96-120)
__global__ void k_ice_ocean_stress(int nn, double F_ocean, const double* u, const double* v, const double* uo,
    const double* vo, double* tx, double* ty)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn)
        return;
    const double ur = u[n] - uo[n], vr = v[n] - vo[n];
    const double a = sqrt(ur * ur + vr * vr);
    tx[n] = F_ocean * a * ur;
    ty[n] = F_ocean * a * vr;
}
// Benchmark cyclone (BenchmarkAtmosphere.cpp:29-76) at the cell lower-left corners; with
// several storms, the sum of copies of it, storm i centred at (x0[i], y0[i]) with its vMax.
constexpr int kMaxStorms = 8;
struct Storms {
    int n;
    double x0[kMaxStorms], y0[kMaxStorms], vmax[kMaxStorms];
};
__global__ void k_benchmark_wind(int nx, int ny, double dx, double dy, Storms st, double* uw, double* vw)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nx * ny)
        return;
    const double A = 1e-5, k = 1e-5;
    const double alpha = 72. / 180. * M_PI;
    const double ca = cos(alpha), sa = sin(alpha);
    double u = 0.0, v = 0.0;
    for (int i = 0; i < st.n; ++i) {
        const double xp = (e % nx) * dx - st.x0[i], yp = (e / nx) * dy - st.y0[i];
        const double scale = A * exp(-k * hypot(xp, yp));
        u += -scale * st.vmax[i] * (ca * xp + sa * yp);
        v += -scale * st.vmax[i] * (-sa * xp + ca * yp);
    }
    uw[e] = u;
    vw[e] = v;
}
__global__ void k_zero_higher(double* f, int ne, int nc)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<size_t>(ne) * (nc - 1))
        return;
    f[ne + t] = 0.0;
}

// Element integrals of h and A (Tools::MeanValue, Tools.hpp:24-45: 3x3 Gauss, J w weights
// and cos(latitude) on the sphere, every element), summed with compensation: per block a
// (sum, compensation) pair per field, then k_sum_final<2> over the blocks in a fixed order.
template <int DG>
__global__ __launch_bounds__(kBlock) void k_mass_partial(MeshView m, int ne, const double* h, const double* A, double* part)
{
    __shared__ double sh[4][kBlock];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double sum[2] = { 0.0, 0.0 }, comp[2] = { 0.0, 0.0 };
    if (e < ne) {
        double c[DG];
#pragma unroll
        for (int i = 0; i < DG; ++i)
            c[i] = h[static_cast<size_t>(i) * ne + e];
        sum[0] = elementMean<DG>(m, e, c);
#pragma unroll
        for (int i = 0; i < DG; ++i)
            c[i] = A[static_cast<size_t>(i) * ne + e];
        sum[1] = elementMean<DG>(m, e, c);
    }
    blockTwoSum<2>(sh, sum, comp);
    if (threadIdx.x == 0)
        for (int f = 0; f < 2; ++f) {
            part[(blockIdx.x * 2 + f) * 2] = sum[f];
            part[(blockIdx.x * 2 + f) * 2 + 1] = comp[f];
        }
}
// Compensated sums of NF per-element arrays x[f * n + e] (the advection budget tallies),
// block partials for k_sum_final<NF>; each thread takes the elements t, t + grid, ...
template <int NF>
__global__ __launch_bounds__(kBlock) void k_tally_partial(const double* x, int n, double* part)
{
    __shared__ double sh[2 * NF][kBlock];
    double sum[NF], comp[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f)
        sum[f] = comp[f] = 0.0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            double s, err;
            twoSum(sum[f], x[static_cast<size_t>(f) * n + e], s, err);
            sum[f] = s;
            comp[f] += err;
        }
    blockTwoSum<NF>(sh, sum, comp);
    if (threadIdx.x == 0)
        for (int f = 0; f < NF; ++f) {
            part[(blockIdx.x * NF + f) * 2] = sum[f];
            part[(blockIdx.x * NF + f) * 2 + 1] = comp[f];
        }
}

// ---------------------------------------------------------------- host object
// ---------------------------------------------------------------- brittle (BBM) rheology
// Constants of BBMStressUpdateStep::stressUpdateHighOrder (BBMStressUpdateStep.hpp:40-140)
// for one substep on a uniform mesh: the element size h_e enters through sqrt(0.1 / h_e)
// (cohesion and compressive strength) and h_e sqrt(2 (1 + nu) rho_ice) (damage time).
struct BrittleArgs {
    double dt, nu, young, P0, expP, compaction, tau0, expR, tanPhi, cohesion, compr;
    double duFac;  // dt / (1 - nu^2): the elastic increment is duFac E strain
    double dtOverT; // dt / (h_e sqrt(2 (1 + nu) rho_ice)): dt / t_d = dtOverT sqrt(E)
    const double* damageIn;
    double* damage; // SoA [DG][ne]
    // General meshes: the per-element table of BrittleGeo (null on uniform meshes, whose
    // cohesion, compr and dtOverT above hold for every element) and the factors that turn
    // an element's h into them: cohesion = Clab sqrt(0.1 / h), compr = comprStrength
    // sqrt(0.1 / h), dtOverT = dt / (h tdFac) with tdFac = sqrt(2 (1 + nu) rho_ice).
    const double* geo;
    double Clab, comprStrength, tdFac;
};
// The BBM arithmetic at one Gauss point (BBMStressUpdateStep.hpp:67-140): y11, y12, y22
// hold the interpolated stress on entry and the new one on exit, dg the damage. POW5
// (exponent_relaxation_sigma 5, exponent_compression_factor 1.5, the MEB defaults) writes
// the two powers as products, x^4 = (x^2)^2 and h^2.5 = h^2 sqrt(h).
template <bool POW5 = false>
__device__ __forceinline__ void bbmPoint(const BrittleArgs& b, double hg, double ag, double& dg, double x11, double x12,
    double x22, double& y11, double& y12, double& y22)
{
    const double nu = b.nu, dt = b.dt;
    hg = fmax(hg, 0.0);
    ag = fmin(fmax(ag, 0.0), 1.0);
    dg = fmin(fmax(dg, 1e-12), 1.0);
    const double sn0 = 0.5 * (y11 + y22);
    const double expC = exp(b.compaction * (1.0 - ag));
    double tv, Pmax;
    if constexpr (POW5) {
        const double x = dg * expC, x2 = x * x;
        tv = b.tau0 * (x2 * x2);
        Pmax = b.P0 * (hg * hg * sqrt(hg)) * expC;
    } else {
        tv = b.tau0 * pow(dg * expC, b.expR);
        Pmax = b.P0 * pow(hg, b.expP) * expC;
    }
    const double tildeP = sn0 < 0.0 ? fmin(-Pmax / sn0, 1.0) : 0.0;
    const double mult = tv / (tv + (1.0 - tildeP) * dt);
    const double E = hg * b.young * dg * expC;
    const double Du = b.duFac * E;
    y11 = (y11 + Du * (x11 + nu * x22)) * mult;
    y22 = (y22 + Du * (nu * x11 + x22)) * mult;
    y12 = (y12 + Du * x12 * (1.0 - nu)) * mult;
    const double sn = 0.5 * (y11 + y22);
    const double tau = sqrt(0.25 * (y11 - y22) * (y11 - y22) + y12 * y12);
    const double cohesion = b.cohesion * hg, cstr = b.compr * hg;
    // d_crit = cohesion / (tau + tan(phi) sn) if that denominator is positive, else 1; the
    // compressive branch -cstr / sn replaces it when sn < -cstr: the operands are chosen
    // first, so one division serves both branches
    const bool compressive = sn < -cstr, shear = tau + b.tanPhi * sn > 0.0;
    const double num = compressive ? -cstr : cohesion, den = compressive ? sn : tau + b.tanPhi * sn;
    const double dcrit = fmin(compressive || shear ? num / den : 1.0, 1.0);
    const double relax = (1.0 - dcrit) * (b.dtOverT * sqrt(E));
    dg -= dg * relax;
    y11 -= y11 * relax;
    y12 -= y12 * relax;
    y22 -= y22 * relax;
}

// One element of the BBM stress and damage update. Strain coefficients from the velocity
// (STRAIN, ocean elements; stored) or the stored ones (land; a stress pass alone); at each Gauss point of the stress
// rule the elastic increment, the viscous relaxation (tildeP from the compressive limit
// Pmax), the Mohr-Coulomb / compressive failure d_crit and the damage relaxation over the
// time t_d; then the stress projected with M^-1 PSI diag(w J) from the operator table and
// the damage with psi_d w / m_d (exact for the affine elements of a uniform mesh).
// The new stress goes to acc1..3 and the new damage to accD; with STORE_STRAIN the strain
// of an ocean element is stored as well.
template <int CG, int DG, bool STRAIN, bool STORE_STRAIN, bool UNI>
__device__ __forceinline__ void bbmElement(const ElemArgs& a, const BrittleArgs& bIn, int e, double* acc1,
    double* acc2, double* acc3, double* accD)
{
    using L = MomLayout<CG>;
    using BG = BrittleGeo<CG, DG>;
    constexpr int S = L::S, NG = L::G, Q = L::Q;
    const size_t ne = a.ne;
    const Ops<UNI> op { a.ops, a.stride };
    BrittleArgs b = bIn;
    if constexpr (!UNI) {
        // BBMStressUpdateStep.hpp:134, 158-159 with this element's h
        const double hE = b.geo[BG::H * ne + e], scale = sqrt(0.1 / hE);
        b.cohesion = b.Clab * scale;
        b.compr = b.comprStrength * scale;
        b.dtOverT = b.dt / (hE * b.tdFac);
    }
    double e11[S], e12[S], e22[S];
    if (STRAIN && a.ocean[e]) {
        strainLocal<CG, UNI>(a, e, e11, e12, e22);
        if constexpr (STORE_STRAIN)
            storeTensor<S>(a.epsOut, ne, e, e11, e12, e22);
    } else {
        loadTensor<S>(a.eps, ne, e, e11, e12, e22);
    }
    double s11[S], s12[S], s22[S], hc[DG], ac[DG], dc[DG];
    loadTensor<S>(a.sig, ne, e, s11, s12, s22);
#pragma unroll
    for (int i = 0; i < DG; ++i) {
        hc[i] = a.h[i * ne + e];
        ac[i] = a.A[i * ne + e];
        dc[i] = b.damageIn[i * ne + e];
    }
#pragma unroll
    for (int k = 0; k < S; ++k)
        acc1[k] = acc2[k] = acc3[k] = 0.0;
#pragma unroll
    for (int k = 0; k < DG; ++k)
        accD[k] = 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double hg = 0, ag = 0, dg = 0;
#pragma unroll
        for (int i = 0; i < DG; ++i) {
            hg += hc[i] * Psi<DG, NG>::v[i][q];
            ag += ac[i] * Psi<DG, NG>::v[i][q];
            dg += dc[i] * Psi<DG, NG>::v[i][q];
        }
        double x11 = 0, x12 = 0, x22 = 0, y11 = 0, y12 = 0, y22 = 0;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const double ps = Psi<S, NG>::v[k][q];
            x11 += e11[k] * ps;
            x12 += e12[k] * ps;
            x22 += e22[k] * ps;
            y11 += s11[k] * ps;
            y12 += s12[k] * ps;
            y22 += s22[k] * ps;
        }
        bbmPoint(b, hg, ag, dg, x11, x12, x22, y11, y12, y22);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const double w = op(L::B + k * Q + q, e);
            acc1[k] += w * y11;
            acc2[k] += w * y12;
            acc3[k] += w * y22;
        }
#pragma unroll
        for (int k = 0; k < DG; ++k) {
            if constexpr (UNI)
                accD[k] += (Psi<DG, NG>::v[k][q] * (gauss_weight(NG, q % NG) * gauss_weight(NG, q / NG))
                               * invMassRef(k))
                    * dg;
            else
                accD[k] += b.geo[static_cast<size_t>(BG::DAM + k * Q + q) * ne + e] * dg;
        }
    }
}
// Factors of the DG basis (dgBasisAt): psi_k(X, Y) = f_FX[k](X) g_GY[k](Y) with f, g in
// {1, t, t^2 - 1/12}; on a Gauss row (fixed Y) a DG field is r0 + r1 X + r2 (X^2 - 1/12).
__device__ __forceinline__ constexpr int basisFX(int k) { return k == 1 || k == 5 || k == 7 ? 1 : (k == 3 || k == 6 ? 2 : 0); }
__device__ __forceinline__ constexpr int basisGY(int k) { return k == 2 || k == 5 || k == 6 ? 1 : (k == 4 || k == 7 ? 2 : 0); }
__device__ __forceinline__ constexpr double legendre01(int i, double t) { return i == 0 ? 1.0 : (i == 1 ? t : t * t - 1.0 / 12.0); }

// bbmElement for the one-pass kernel on an affine uniform mesh, with the operators of
// elementAffineCompute: the Gauss-point strain is the exact gradient of the CG velocity by
// sum factorisation (ocean elements; land elements interpolate their stored strain), both
// projections are psi_k(q) w_q / m_k literals, and every DG field is evaluated per Gauss
// row from its three row coefficients. The strain is not stored.
// The caller's per-thread column gp(r), r = 4 q + {0, 1, 2, 3} for Gauss point q, holds
// first the previous stress and damage interpolated there (a pre-pass, one field at a
// time), then the new ones; the projection runs after the Gauss loop in the order q. So
// neither the previous coefficients nor the 30 projection sums are live while the point
// arithmetic runs. Ocean and land elements take separate Gauss loops.
template <int DG, int NG, class GP>
__device__ __forceinline__ void bbmInterpolateToGauss(const double* f, size_t ne, int e, int r, GP& gp)
{
    double c[DG];
#pragma unroll
    for (int i = 0; i < DG; ++i)
        c[i] = f[i * ne + e];
#pragma unroll
    for (int gy = 0; gy < NG; ++gy) {
        const double Y = gauss_point(NG, gy) - 0.5;
        double rc[3] = { 0, 0, 0 };
#pragma unroll
        for (int i = 0; i < DG; ++i)
            rc[basisFX(i)] += c[i] * legendre01(basisGY(i), Y);
#pragma unroll
        for (int gx = 0; gx < NG; ++gx) {
            const double X = gauss_point(NG, gx) - 0.5, X2 = X * X - 1.0 / 12.0;
            gp(4 * (gx + NG * gy) + r) = rc[0] + rc[1] * X + rc[2] * X2;
        }
    }
}
template <int CG, int DG, bool POW5, bool OCEAN, class GP>
__device__ __forceinline__ void bbmGaussLoop(const ElemArgs& a, const BrittleArgs& b, int e, GP& gp)
{
    using L = MomLayout<CG>;
    constexpr int S = L::S, NG = L::G, N = L::N, K = CG + 1;
    const size_t ne = a.ne;
    double buf[OCEAN ? 2 * N : 3 * S]; // ocean: u (N), v (N) by node row; land: e11, e12, e22 (S each)
    if constexpr (OCEAN) {
        const int ix = e % a.nx, iy = e / a.nx, row = CG * a.nx + 1;
        const int cgi = CG * row * iy + CG * ix;
#pragma unroll
        for (int r = 0; r < K; ++r)
#pragma unroll
            for (int c = 0; c < K; ++c) {
                buf[c + K * r] = a.u[cgi + c + r * row];
                buf[N + c + K * r] = a.v[cgi + c + r * row];
            }
    } else {
#pragma unroll
        for (int k = 0; k < 3 * S; ++k)
            buf[k] = a.eps[k * ne + e];
    }
    double hc[DG], ac[DG];
#pragma unroll
    for (int i = 0; i < DG; ++i) {
        hc[i] = a.h[i * ne + e];
        ac[i] = a.A[i * ne + e];
    }
#pragma unroll
    for (int gy = 0; gy < NG; ++gy) {
        const double Y = gauss_point(NG, gy) - 0.5;
        double rh[3] = { 0, 0, 0 }, ra[3] = { 0, 0, 0 };
#pragma unroll
        for (int i = 0; i < DG; ++i) {
            const double g = legendre01(basisGY(i), Y);
            rh[basisFX(i)] += hc[i] * g;
            ra[basisFX(i)] += ac[i] * g;
        }
        double Iu[K], Du[K], Iv[K], Dv[K];
        if constexpr (OCEAN) {
#pragma unroll
            for (int jx = 0; jx < K; ++jx) {
                double iu = 0, du = 0, iv = 0, dv = 0;
#pragma unroll
                for (int jy = 0; jy < K; ++jy) {
                    const double ph = cg_1d(CG, jy, gauss_point(NG, gy)), dp = cg_1d_dx(CG, jy, gauss_point(NG, gy));
                    iu += ph * buf[jx + K * jy];
                    du += dp * buf[jx + K * jy];
                    iv += ph * buf[N + jx + K * jy];
                    dv += dp * buf[N + jx + K * jy];
                }
                Iu[jx] = iu;
                Du[jx] = du;
                Iv[jx] = iv;
                Dv[jx] = dv;
            }
        }
#pragma unroll
        for (int gx = 0; gx < NG; ++gx) {
            const double X = gauss_point(NG, gx) - 0.5, X2 = X * X - 1.0 / 12.0;
            double g11 = 0, g12 = 0, g22 = 0;
            if constexpr (OCEAN) {
                double uxi = 0, ueta = 0, vxi = 0, veta = 0;
#pragma unroll
                for (int jx = 0; jx < K; ++jx) {
                    const double ph = cg_1d(CG, jx, gauss_point(NG, gx)), dp = cg_1d_dx(CG, jx, gauss_point(NG, gx));
                    uxi += dp * Iu[jx];
                    ueta += ph * Du[jx];
                    vxi += dp * Iv[jx];
                    veta += ph * Dv[jx];
                }
                g11 = uxi * a.xiX + ueta * a.etaX;
                g22 = vxi * a.xiY + veta * a.etaY;
                g12 = 0.5 * ((uxi * a.xiY + ueta * a.etaY) + (vxi * a.xiX + veta * a.etaX));
            } else {
                double bd[8];
                dgBasisAt(X, Y, bd);
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    g11 += buf[k] * bd[k];
                    g12 += buf[S + k] * bd[k];
                    g22 += buf[2 * S + k] * bd[k];
                }
            }
            const int q = gx + NG * gy;
            const double hg = rh[0] + rh[1] * X + rh[2] * X2, ag = ra[0] + ra[1] * X + ra[2] * X2;
            double y11 = gp(4 * q + 0), y12 = gp(4 * q + 1), y22 = gp(4 * q + 2), dg = gp(4 * q + 3);
            bbmPoint<POW5>(b, hg, ag, dg, g11, g12, g22, y11, y12, y22);
            gp(4 * q + 0) = y11;
            gp(4 * q + 1) = y12;
            gp(4 * q + 2) = y22;
            gp(4 * q + 3) = dg;
        }
    }
}
template <int CG, int DG, bool POW5, class GP>
__device__ __forceinline__ void bbmElementAffine(const ElemArgs& a, const BrittleArgs& b, int e, double* acc1,
    double* acc2, double* acc3, double* accD, GP&& gp)
{
    using L = MomLayout<CG>;
    constexpr int S = L::S, NG = L::G;
    const size_t ne = a.ne;
    bbmInterpolateToGauss<S, NG>(a.sig + 0 * S * ne, ne, e, 0, gp);
    bbmInterpolateToGauss<S, NG>(a.sig + 1 * S * ne, ne, e, 1, gp);
    bbmInterpolateToGauss<S, NG>(a.sig + 2 * S * ne, ne, e, 2, gp);
    bbmInterpolateToGauss<DG, NG>(b.damageIn, ne, e, 3, gp);
    if (a.ocean[e])
        bbmGaussLoop<CG, DG, POW5, true>(a, b, e, gp);
    else
        bbmGaussLoop<CG, DG, POW5, false>(a, b, e, gp);
#pragma unroll
    for (int k = 0; k < S; ++k)
        acc1[k] = acc2[k] = acc3[k] = 0.0;
#pragma unroll
    for (int k = 0; k < DG; ++k)
        accD[k] = 0.0;
#pragma unroll
    for (int q = 0; q < NG * NG; ++q) {
        const int gx = q % NG, gy = q / NG;
        double bd[8];
        dgBasisAt(gauss_point(NG, gx) - 0.5, gauss_point(NG, gy) - 0.5, bd);
        const double w = gauss_weight(NG, gx) * gauss_weight(NG, gy);
        const double y11 = gp(4 * q + 0), y12 = gp(4 * q + 1), y22 = gp(4 * q + 2), dg = gp(4 * q + 3);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const double c = bd[k] * w * invMassRef(k);
            if (c == 0.0) // basis functions vanishing at this Gauss point (a literal test)
                continue;
            acc1[k] += c * y11;
            acc2[k] += c * y12;
            acc3[k] += c * y22;
        }
#pragma unroll
        for (int k = 0; k < DG; ++k) {
            const double c = bd[k] * w * invMassRef(k);
            if (c == 0.0)
                continue;
            accD[k] += c * dg;
        }
    }
}

template <int CG, int DG, bool STRAIN, bool UNI>
__global__ __launch_bounds__(kBlock) void k_bbm_element(ElemArgs a, BrittleArgs b)
{
    constexpr int S = cg2dgstress(CG);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.ne)
        return;
    double acc1[S], acc2[S], acc3[S], accD[DG];
    bbmElement<CG, DG, STRAIN, STRAIN, UNI>(a, b, e, acc1, acc2, acc3, accD);
    storeTensor<S>(a.sig, a.ne, e, acc1, acc2, acc3);
#pragma unroll
    for (int k = 0; k < DG; ++k)
        b.damage[static_cast<size_t>(k) * a.ne + e] = accD[k];
}

// BrittleCGDynamicsKernel::updateMomentum (BrittleCGDynamicsKernel.hpp:200-253) for one
// node: the semi-implicit substep with the Coriolis term and the ocean drag turned by the
// turning angle, solved as the reference writes it (its v carries +beta (gradX + tauX)),
// and the step-averaged velocity accumulated before the Dirichlet zero. The terms fixed
// over the substeps of a step come from k_brittle_coeffs (a.ocn = (uo, vo), a.cf = (dtm,
// dtm / ML) with dtm = dt / (rho_ice H), a.qf = dtm (dA ua - rho_ice H g gx, dA va -
// rho_ice H g gy) with dA = A F_atm |ua|, af = A F_ocean), so that
//   dtm (gradX + tauX) = dtm / ML dSx + qf.x + dtm cP (uo cos - vo sin), cP = af |uo - u|.
struct BrittleIn {
    double u, v, au, av, af;
    double2 ocn, cf, qf;
};
struct BrittleNode {
    using In = BrittleIn;
    static constexpr bool kAllNodes = true;
    double ct, st, nsteps;
    double* avgU;
    double* avgV;
    const double* af;
    __device__ __forceinline__ void load(const NodeArgs& a, int n, int, int, In& in) const
    {
        in.u = a.u[n];
        in.v = a.v[n];
        in.ocn = a.ocn[n];
        in.cf = a.cf[n];
        in.qf = a.qf[n];
        in.af = af[n];
        in.au = avgU[n];
        in.av = avgV[n];
    }
    // dSx, dSy are zero on Dirichlet nodes (dirichletZero of the divergence)
    __device__ __forceinline__ void update(const NodeArgs& a, const In& in, int n, bool dir, double dSx, double dSy,
        double& un, double& vn) const
    {
        const double dtm = in.cf.x, dtmML = in.cf.y;
        const double dcP = dtm * (in.af * hypot(in.ocn.x - in.u, in.ocn.y - in.v));
        const double al = 1.0 + dcP * ct;
        const double be = a.dt * a.fc + dcP * st;
        const double rD = 1.0 / (al * al + be * be);
        const double GX = dtmML * dSx + in.qf.x + dcP * (in.ocn.x * ct - in.ocn.y * st); // dtm (gradX + tauX)
        const double GY = dtmML * dSy + in.qf.y + dcP * (in.ocn.y * ct + in.ocn.x * st); // dtm (gradY + tauY)
        un = (al * in.u + be * in.v + (al * GX + be * GY)) * rD;
        vn = (al * in.v - be * in.u + (al * GY + be * GX)) * rD;
        avgU[n] = in.au + un / nsteps;
        avgV[n] = in.av + vn / nsteps;
        if (dir) {
            un = 0.0;
            vn = 0.0;
        }
    }
};
// The terms of BrittleNode fixed over the substeps of one step (a.dt = the substep).
__global__ __launch_bounds__(kBlock) void k_brittle_coeffs(NodeArgs a, double2* ocn, double2* cf, double2* qf,
    double* af)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.nn)
        return;
    const double H = a.H[n], A = a.A[n], ua = a.ua[n], va = a.va[n];
    const double dtm = a.dt / (a.rho_ice * H);
    const double dA = A * a.F_atm * hypot(ua, va);
    const double rhoHg = a.rho_ice * H * a.gravity;
    ocn[n] = make_double2(a.uo[n], a.vo[n]);
    cf[n] = make_double2(dtm, dtm / a.ML[n]);
    qf[n] = make_double2(dtm * (dA * ua - rhoHg * a.gx[n]), dtm * (dA * va - rhoHg * a.gy[n]));
    af[n] = A * a.F_ocean;
}
// Split pass: every node, the Dirichlet zero left to k_boundaries.
__global__ __launch_bounds__(kBlock) void k_brittle_momentum(NodeArgs a, BrittleNode pol)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.nn)
        return;
    BrittleIn in;
    pol.load(a, n, 0, 0, in);
    double un, vn;
    pol.update(a, in, n, false, a.dSx[n], a.dSy[n], un, vn);
    a.u[n] = un;
    a.v[n] = vn;
}

// One whole brittle substep in one pass over a uniform mesh without periodic segments,
// laid out as k_mevp_fused: phase 1 updates stress and damage of the tile and its
// left/bottom halo in registers (stores for the tile's own elements), phase 2 sums the
// weak divergence from LDS and runs the brittle node update. Stress, damage and
// velocities are ping-ponged; the averaged velocity is updated in place by the node's
// owner. The arithmetic is that of k_bbm_element, k_divergence, k_brittle_momentum and
// k_boundaries in sequence (the strain is not stored).
struct BrittleFusedArgs {
    ElemArgs e;
    NodeArgs n;
    BrittleArgs b;
    BrittleNode pol;
    double* sigOut;
    double* damageOut;
    double* uOut;
    double* vOut;
    const double* geo; // GeoLayout tables of a general mesh (k_bbm_fused_geom)
};
template <int CG, int DG, bool POW5>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_bbm_fused(BrittleFusedArgs f)
{
    using T = FusedTile<CG>;
    constexpr int N = T::N, S = cg2dgstress(CG), Q = MomLayout<CG>::Q;
    // One column per local element. Rows 4 q + c: the Gauss-point stress and damage of the
    // element update; afterwards rows 2 j + c: its divergence contributions (node j,
    // component c). Each thread reads and writes only its own column.
    constexpr int kRows = 4 * Q > 2 * N ? 4 * Q : 2 * N;
    __shared__ double stage[kRows][kBlock];
    const NodeArgs& a = f.n;
    const int tilesX = (a.nx + T::TX - 1) / T::TX, tilesY = (a.ny + T::TY - 1) / T::TY;
    const int tile = xcdTile(blockIdx.x, tilesX * tilesY);
    if (tile < 0)
        return;
    const int ex0 = (tile % tilesX) * T::TX, ey0 = (tile / tilesX) * T::TY;
    {
        const int le = threadIdx.x; // one element of tile + halo per thread
        const int lx = le % T::HX, ly = le / T::HX;
        const int ex = ex0 - 1 + lx, ey = ey0 - 1 + ly;
        double tx[N], ty[N];
#pragma unroll
        for (int j = 0; j < N; ++j)
            tx[j] = ty[j] = 0;
        if (ex >= 0 && ex < a.nx && ey >= 0 && ey < a.ny) {
            const int e = ex + a.nx * ey;
            double s11[S], s12[S], s22[S], d[DG];
            bbmElementAffine<CG, DG, POW5>(f.e, f.b, e, s11, s12, s22, d, [&](int r) -> double& { return stage[r][le]; });
            if (lx >= 1 && ly >= 1) {
                storeTensor<S>(f.sigOut, a.ne, e, s11, s12, s22);
#pragma unroll
                for (int k = 0; k < DG; ++k)
                    f.damageOut[static_cast<size_t>(k) * a.ne + e] = d[k];
            }
            if (a.ocean[e])
                divergenceAffineFrom<CG, double>(a, s11, s12, s22, tx, ty);
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            stage[2 * j][le] = tx[j];
            stage[2 * j + 1][le] = ty[j];
        }
    }
    __syncthreads();
    nodeTileUpdate<CG, T, BrittleNode>(
        a, [&](int le, int j, int c) { return stage[2 * j + c][le]; }, ex0, ey0, f.uOut, f.vOut, f.pol);
}

// bbmElement for the one-pass kernel on a general mesh, from its geometry (GeoLayout) and
// brittle table (BrittleGeo): the element's h_e sets cohesion, compressive strength and
// damage time; the strain is M^-1 sum_q PSI(q) g(q) as in elementGeomUpdate (ocean
// elements; land elements keep their stored strain) evaluated at the Gauss points; the
// previous stress and damage are staged at the Gauss points in the thread's column col(r),
// r = 4 q + {0, 1, 2, 3}, as in bbmElementAffine; the projections are
// M_S^-1 sum_q PSI(q) J w y(q) and M_D^-1 sum_q PSI_D(q) J w d(q) (iMJwPSI and
// iMJwPSI_dam, BBMStressUpdateStep.hpp:183-193).
template <int CG, int DG, bool POW5, bool SPH, class COL>
__device__ __forceinline__ void bbmElementGeom(const ElemArgs& a, const BrittleArgs& bIn, const double* geo, int e,
    double* s11, double* s12, double* s22, double* dOut, const COL& col)
{
    using L = GeoLayout<CG>;
    using BG = BrittleGeo<CG, DG>;
    constexpr int S = L::S, NG = L::G;
    const size_t ne = a.ne;
    const uint32_t off = static_cast<uint32_t>(e) * 8u;
    BrittleArgs b = bIn;
    {
        // BBMStressUpdateStep.hpp:134, 158-159 with this element's h
        const double hE = ldAt(b.geo + BG::H * ne, off), scale = sqrt(0.1 / hE);
        b.cohesion = b.Clab * scale;
        b.compr = b.comprStrength * scale;
        b.dtOverT = b.dt / (hE * b.tdFac);
    }
    bbmInterpolateToGauss<S, NG>(a.sig + 0 * S * ne, ne, e, 0, col);
    bbmInterpolateToGauss<S, NG>(a.sig + 1 * S * ne, ne, e, 1, col);
    bbmInterpolateToGauss<S, NG>(a.sig + 2 * S * ne, ne, e, 2, col);
    bbmInterpolateToGauss<DG, NG>(b.damageIn, ne, e, 3, col);
    double e11[S], e12[S], e22[S];
    strainGeom<CG, SPH>(a, geo, e, e11, e12, e22, [&](const double* g1, int k, uint32_t offM) {
        return ldAt(g1 + static_cast<size_t>(L::MI + k) * ne, offM);
    });
    if (__builtin_amdgcn_ballot_w64(!a.ocean[e]) != 0) { // a wave with land: uniform branch
        const bool land = !a.ocean[e];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const double x = ldAt(a.eps + (0 * S + k) * ne, off), y = ldAt(a.eps + (1 * S + k) * ne, off),
                         z = ldAt(a.eps + (2 * S + k) * ne, off);
            e11[k] = land ? x : e11[k];
            e12[k] = land ? y : e12[k];
            e22[k] = land ? z : e22[k];
        }
    }
    // the point arithmetic, h and A per Gauss row
    {
        const uint32_t offH = after(after(after(off, e11[S - 1]), e12[S - 1]), e22[S - 1]);
        double hc[DG], ac[DG];
#pragma unroll
        for (int i = 0; i < DG; ++i) {
            hc[i] = ldAt(a.h + i * ne, offH);
            ac[i] = ldAt(a.A + i * ne, offH);
        }
#pragma unroll
        for (int gy = 0; gy < NG; ++gy) {
            const double Y = gauss_point(NG, gy) - 0.5;
            double rh[3] = { 0, 0, 0 }, ra[3] = { 0, 0, 0 };
#pragma unroll
            for (int i = 0; i < DG; ++i) {
                const double g = legendre01(basisGY(i), Y);
                rh[basisFX(i)] += hc[i] * g;
                ra[basisFX(i)] += ac[i] * g;
            }
#pragma unroll
            for (int gx = 0; gx < NG; ++gx) {
                const double X = gauss_point(NG, gx) - 0.5, X2 = X * X - 1.0 / 12.0;
                double bd[8];
                dgBasisAt(X, Y, bd);
                double g11 = 0, g12 = 0, g22 = 0;
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    g11 += e11[k] * bd[k];
                    g12 += e12[k] * bd[k];
                    g22 += e22[k] * bd[k];
                }
                const int q = gx + NG * gy;
                const double hg = rh[0] + rh[1] * X + rh[2] * X2, ag = ra[0] + ra[1] * X + ra[2] * X2;
                double y11 = col(4 * q + 0), y12 = col(4 * q + 1), y22 = col(4 * q + 2), dg = col(4 * q + 3);
                bbmPoint<POW5>(b, hg, ag, dg, g11, g12, g22, y11, y12, y22);
                col(4 * q + 0) = y11;
                col(4 * q + 1) = y12;
                col(4 * q + 2) = y22;
                col(4 * q + 3) = dg;
            }
        }
    }
    // projections
    const double* const g2 = phaseLocal(geo);
    const MapCoef map2 = loadMap<CG>(g2, ne, after(off, static_cast<double>(col(4 * NG * NG - 1))));
    double u11[S], u12[S], u22[S], uD[DG];
#pragma unroll
    for (int k = 0; k < S; ++k)
        u11[k] = u12[k] = u22[k] = 0.0;
#pragma unroll
    for (int k = 0; k < DG; ++k)
        uD[k] = 0.0;
#pragma unroll
    for (int q = 0; q < NG * NG; ++q) {
        double bd[8];
        dgBasisAt(gauss_point(NG, q % NG) - 0.5, gauss_point(NG, q / NG) - 0.5, bd);
        const double jw = gaussGeom<NG, false>(map2, q % NG, q / NG, 1.0, 0.0).jw;
        const double y11 = jw * col(4 * q + 0), y12 = jw * col(4 * q + 1), y22 = jw * col(4 * q + 2),
                     dg = jw * col(4 * q + 3);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            if (bd[k] == 0.0)
                continue;
            u11[k] += bd[k] * y11;
            u12[k] += bd[k] * y12;
            u22[k] += bd[k] * y22;
        }
#pragma unroll
        for (int k = 0; k < DG; ++k) {
            if (bd[k] == 0.0)
                continue;
            uD[k] += bd[k] * dg;
        }
    }
    const uint32_t offM = after(after(after(off, u11[S - 1]), u12[S - 1]), u22[S - 1]);
    inverseMassApply<CG>([&](int k) { return ldAt(g2 + static_cast<size_t>(L::MI + k) * ne, offM); }, u11, u12, u22,
        s11, s12, s22);
#pragma unroll
    for (int k = 0; k < DG; ++k)
        dOut[k] = 0.0;
#pragma unroll
    for (int i = 0; i < DG; ++i)
#pragma unroll
        for (int k = i; k < DG; ++k) {
            const double m = ldAt(b.geo + static_cast<size_t>(BG::mdi(i, k)) * ne, offM);
            dOut[i] += m * uD[k];
            if (k != i)
                dOut[k] += m * uD[i];
        }
}
// One whole brittle substep in one pass over a general mesh without periodic segments:
// k_bbm_fused's tiles, halo, staging and node phase with the element update and the
// divergence computed from the streamed geometry.
template <int CG, int DG, bool POW5, bool SPH> __global__ __launch_bounds__(kBlock) void k_bbm_fused_geom(BrittleFusedArgs f)
{
    using T = FusedTile<CG>;
    constexpr int N = T::N, S = cg2dgstress(CG), Q = MomLayout<CG>::Q;
    // one column per local element: the Gauss-point staging (rows 4 q + c), then the
    // divergence contributions (row 2 j + c)
    constexpr int kRows = 4 * Q > 2 * N ? 4 * Q : 2 * N;
    __shared__ double stage[kRows][kBlock];
    const NodeArgs& a = f.n;
    const int tilesX = (a.nx + T::TX - 1) / T::TX, tilesY = (a.ny + T::TY - 1) / T::TY;
    const int tile = xcdTile(blockIdx.x, tilesX * tilesY);
    if (tile < 0)
        return;
    const int ex0 = (tile % tilesX) * T::TX, ey0 = (tile / tilesX) * T::TY;
    {
        const int le = threadIdx.x; // one element of tile + halo per thread
        const int lx = le % T::HX, ly = le / T::HX;
        const int ex = ex0 - 1 + lx, ey = ey0 - 1 + ly;
        const bool inside = ex >= 0 && ex < a.nx && ey >= 0 && ey < a.ny;
        const int e = min(max(ex, 0), a.nx - 1) + a.nx * min(max(ey, 0), a.ny - 1);
        double s11[S], s12[S], s22[S], d[DG], tx[N], ty[N];
        bbmElementGeom<CG, DG, POW5, SPH>(f.e, f.b, f.geo, e, s11, s12, s22, d,
            [&](int r) -> double& { return stage[r][le]; });
        if (inside && lx >= 1 && ly >= 1) {
            const uint32_t off = static_cast<uint32_t>(e) * 8u;
            storeTensorAt<S>(f.sigOut, a.ne, off, s11, s12, s22);
#pragma unroll
            for (int k = 0; k < DG; ++k)
                fieldAt(f.damageOut + static_cast<size_t>(k) * a.ne, off) = d[k];
        }
        divergenceGeomFrom<CG, SPH>(f.geo, a.ne, e, s11, s12, s22, tx, ty);
        const bool adds = inside && a.ocean[e] != 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            stage[2 * j][le] = adds ? tx[j] : 0.0;
            stage[2 * j + 1][le] = adds ? ty[j] : 0.0;
        }
    }
    __syncthreads();
    nodeTileUpdate<CG, T, BrittleNode>(
        a, [&](int le, int j, int c) { return stage[2 * j + c][le]; }, ex0, ey0, f.uOut, f.vOut, f.pol);
}

// BrittleCGDynamicsKernel::getIceOceanStress (BrittleCGDynamicsKernel.hpp:163-182): from
// the step-averaged velocity, relative velocity ocean minus ice, turned.
__global__ void k_ice_ocean_stress_brittle(int nn, double F_ocean, double ct, double st, const double* au,
    const double* av, const double* uo, const double* vo, double* tx, double* ty)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn)
        return;
    const double ur = uo[n] - au[n], vr = vo[n] - av[n];
    const double cp = F_ocean * hypot(ur, vr);
    tx[n] = cp * (ur * ct - vr * st);
    ty[n] = cp * (vr * ct + ur * st);
}

template <int CG, int DG> struct DynImpl final : DynBase {
    using L = MomLayout<CG>;
    static constexpr int S = L::S, N = L::N;
    nsdg_params p;
    hipStream_t s = nullptr;
    Profiler prof;
    DeviceMesh mesh;
    std::unique_ptr<TransportBase> tr;
    CG2DGOps cg2dg;
    bool ready = false;
    int precision = NSDG_PREC_FP64;
    int pairing = NSDG_PAIR_AUTO;
    size_t ne = 0, nn = 0, nn1 = 0;
    double deltaT = 0;
    DBuf<double> mops, geo, brittleGeo, lumped, lumped1, h, A, ssh, eps, sig;
    DBuf<double> u, v, u0, v0, uo, vo, ua, va, cgH, cgA, gX, gY, dSx, dSy;
    DBuf<double> tmpDG, tmpDG2, tmpCG, tmpCG2, cg1a, cg1b, cg1c, scratch, dg0a, dg0b, part, res;
    DBuf<uint8_t> dirNode;
    DBuf<double2> fzOcn, fzCf, fzQ; // frozen momentum coefficients (k_node_coeffs)
    DBuf<double> strength;          // ice strength at the Gauss points
    DBuf<double> fzAF;              // brittle: A F_ocean per node
    DBuf<float> sigF, strengthF;    // FP32 variant: stress and strength
    DBuf<float> sigFAlt;            // ping-pong partner of sigF (fused subcycle)
    uint32_t subcycleCount = 0;     // element passes run so far (FP32 stochastic rounding)
    DBuf<double> sigVP, rpart, rsum; // momentum residual: VP stress, block partials, sums
    bool budgetOn = false;           // advection mass budget: per-element tallies [4][ne]
    DBuf<double> tally;
    DBuf<double> uStart, vStart, sigStart; // saved start of a timestep (warm-start gate)
    DBuf<double> uAlt, vAlt, sigAlt;       // ping-pong partners of u, v, sigma (fused subcycle)
    std::map<std::string, std::pair<int, std::vector<double>>> passive;
    // brittle rheology (BBMDynamicsKernel): damage (SoA DG), step-averaged velocity,
    // stress transport in the stress space, element size of the uniform mesh
    bool brittle = false;
    nsdg_meb_params meb {};
    DBuf<double> damage, damageAlt, avgU, avgV;
    std::unique_ptr<TransportBase> stressTr;
    CG2DGOps cg2dgS;
    double hElem = 0.0;

    int device = 0; // the device current at creation, which runs the stream's work
    explicit DynImpl(const nsdg_params& prm)
        : p(prm)
    {
        NSDG_HIP(hipGetDevice(&device));
        NSDG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    }
    ~DynImpl() override
    {
        tr.reset();
        stressTr.reset();
        if (s)
            (void)hipStreamDestroy(s);
    }
    Profiler& profiler() override { return prof; }
    const HostMesh& meshHost() const override
    {
        need();
        return mesh.host;
    }
    int cgDegree() const override { return CG; }
    int dgComp() const override { return DG; }
    int stressComp() const override { return S; }
    bool brittleRheology() const override { return brittle; }
    bool initialised() const override { return ready; }
    size_t cgNodes() const override { return nn; }
    void need() const { require(ready, "kernel not initialised", NSDG_ERR_STATE); }
    size_t opStride() const { return mesh.uniform ? 1 : ne; }

    void initialise(const HostMesh& m) override
    {
        ready = false;
        mesh.build(m, s);
        ne = mesh.ne;
        nn = static_cast<size_t>(CG * mesh.nx + 1) * (CG * mesh.ny + 1);
        nn1 = static_cast<size_t>(mesh.nx + 1) * (mesh.ny + 1);
        // the kernels index CG nodes with 32-bit ints (the mesh checks elements and vertices)
        require(nn < (size_t(1) << 31), "too many CG nodes for 32-bit node indices", NSDG_ERR_ARG);
        // the one-pass kernels address element fields and CG nodes with 32-bit byte offsets
        // from scalar field bases (2^29 CG2 nodes: a 11585^2 mesh, ~210 GB of FP64 state)
        require(ne < (size_t(1) << 29), "meshes need fewer than 2^29 elements", NSDG_ERR_ARG);
        require(nn < (size_t(1) << 29), "meshes need fewer than 2^29 CG nodes", NSDG_ERR_ARG);
        tr = makeTransport(DG, mesh, s, prof);
        tr->setScheme(2); // DynamicsKernel.hpp:61
        cg2dg.build(CG, DG, mesh, s);
        const int nOps = mesh.uniform ? 1 : static_cast<int>(ne);
        mops.alloc(static_cast<size_t>(mesh.spherical ? L::KS : L::K) * nOps);
        k_momentum_ops<CG><<<blocksFor(nOps), kBlock, 0, s>>>(view(mesh), nOps, mops.p);
        checkLaunch("momentum_ops");
        if (!mesh.uniform) {
            geo.alloc(static_cast<size_t>(GeoLayout<CG>::K) * ne);
            k_geometry<CG><<<blocksFor(ne), kBlock, 0, s>>>(view(mesh), static_cast<int>(ne), geo.p);
            checkLaunch("geometry");
        }
        lumped.alloc(nn);
        lumped1.alloc(nn1);
        k_lumped_mass<CG, CG == 1 ? 1 : 4><<<blocksFor(nn), kBlock, 0, s>>>(view(mesh), lumped.p);
        checkLaunch("lumped_mass");
        k_lumped_mass<1, 2><<<blocksFor(nn1), kBlock, 0, s>>>(view(mesh), lumped1.p);
        checkLaunch("lumped_mass1");
        for (auto* b : { &h, &A, &tmpDG, &tmpDG2 })
            b->alloc(ne * DG);
        ssh.alloc(ne);
        dg0a.alloc(ne);
        dg0b.alloc(ne);
        eps.alloc(3 * S * ne);
        sig.alloc(3 * S * ne);
        for (auto* b : { &u, &v, &u0, &v0, &uo, &vo, &ua, &va, &cgH, &cgA, &gX, &gY, &dSx, &dSy, &tmpCG, &tmpCG2 })
            b->alloc(nn);
        for (auto* b : { &cg1a, &cg1b, &cg1c })
            b->alloc(nn1);
        for (auto* b : { &fzOcn, &fzCf, &fzQ })
            b->alloc(nn);
        strength.alloc(static_cast<size_t>(L::Q) * ne);
        part.alloc(std::max<size_t>(4 * blocksFor(ne), 8 * kTallyBlocks));
        res.alloc(4);
        if (budgetOn) { // a re-initialised mesh starts a new budget
            tally.alloc(4 * ne);
            deviceFill(tally.p, tally.n, 0.0, s);
        }
        sigVP.alloc(3 * S * ne);
        deviceFill(sigVP.p, sigVP.n, 0.0, s);
        uAlt.alloc(nn);
        vAlt.alloc(nn);
        sigAlt.alloc(3 * S * ne);
        {
            using T = FusedTile<CG>;
            const size_t tiles = xcdGrid(((mesh.nx + T::TX - 1) / T::TX) * ((mesh.ny + T::TY - 1) / T::TY));
            rpart.alloc(4 * std::max<size_t>(kResidualBlocks, tiles));
        }
        rsum.alloc(2);
        for (auto* b : { &h, &A, &tmpDG, &tmpDG2 })
            deviceFill(b->p, b->n, 0.0, s);
        for (auto* b : { &ssh, &eps, &sig, &u, &v, &u0, &v0, &uo, &vo, &ua, &va, &cgH, &cgA, &gX, &gY, &dSx, &dSy })
            deviceFill(b->p, b->n, 0.0, s);
        // Dirichlet CG nodes (dirichletZero, CGDynamicsKernel.cpp:400-437)
        std::vector<uint8_t> dn(nn, 0);
        const size_t nx = mesh.host.nx, row = CG * nx + 1;
        for (int seg = 0; seg < 4; ++seg)
            for (auto eid : mesh.host.dirichlet[seg]) {
                const size_t ix = eid % nx, iy = eid / nx;
                for (size_t j = 0; j <= CG; ++j) {
                    size_t id;
                    if (seg == 0)
                        id = iy * CG * row + CG * ix + j;
                    else if (seg == 1)
                        id = iy * CG * row + CG * ix + CG + row * j;
                    else if (seg == 2)
                        id = (iy + 1) * CG * row + CG * ix + j;
                    else
                        id = iy * CG * row + CG * ix + row * j;
                    dn[id] = 1;
                }
            }
        dirNode.alloc(nn);
        dirNode.upload(dn.data(), nn, s);
        if (brittle)
            initBrittle();
        passive.clear();
        subcycleCount = 0;
        NSDG_HIP(hipStreamSynchronize(s));
        ready = true;
        if (precision == NSDG_PREC_FP32) {
            try {
                checkPrecision();
            } catch (...) {
                precision = NSDG_PREC_FP64;
                throw;
            }
            sigF.release();
            allocPrecision();
        }
    }

    // BrittleCGDynamicsKernel::initialise (BrittleCGDynamicsKernel.hpp:72-87): undamaged
    // ice (damage 1), zero averaged velocity, the stress transport on the stress space.
    void initBrittle()
    {
        stressTr = makeTransport(S, mesh, s, prof);
        stressTr->setScheme(2);
        cg2dgS.build(CG, S, mesh, s);
        damage.alloc(ne * DG);
        damageAlt.alloc(ne * DG);
        deviceFill(damage.p, ne * DG, 0.0, s);
        deviceFill(damage.p, ne, 1.0, s);
        avgU.alloc(nn);
        avgV.alloc(nn);
        fzAF.alloc(nn);
        deviceFill(avgU.p, nn, 0.0, s);
        deviceFill(avgV.p, nn, 0.0, s);
        // sqrt of the quadrilateral area (ParametricMesh.hpp:276-299) of element 0, shared
        // by every translate of it
        const double* c = mesh.host.coords.data();
        const size_t vx = mesh.host.nx + 1;
        const size_t id[4] = { 0, 1, vx, vx + 1 }; // lower left, lower right, upper left, upper right
        auto sq = [&](int i, int j) {
            const double dx = c[2 * id[i]] - c[2 * id[j]], dy = c[2 * id[i] + 1] - c[2 * id[j] + 1];
            return dx * dx + dy * dy;
        };
        const double a = sq(0, 1), b = sq(1, 3), cc = sq(2, 3), d = sq(0, 2), e = sq(0, 3), f = sq(1, 2);
        hElem = std::sqrt(0.25 * std::sqrt(4.0 * e * f - (b + d - a - cc) * (b + d - a - cc)));
        if (!mesh.uniform) {
            // element sizes and damage projections per element
            brittleGeo.alloc(static_cast<size_t>(BrittleGeo<CG, DG>::K) * ne);
            k_brittle_geometry<CG, DG><<<blocksFor(ne), kBlock, 0, s>>>(view(mesh), static_cast<int>(ne), brittleGeo.p);
            checkLaunch("brittle_geometry");
        }
    }
    void setBrittle(const nsdg_meb_params& m) override
    {
        require(!ready, "the rheology is chosen before initialise", NSDG_ERR_STATE);
        require(m.young > 0 && m.nu0 >= 0 && m.nu0 < 1 && m.undamaged_time_relaxation_sigma > 0,
            "MEB parameters need young > 0, 0 <= nu0 < 1 and a positive relaxation time");
        brittle = true;
        meb = m;
    }
    void needMEVP(const char* what) const
    {
        require(!brittle, std::string(what) + " is an mEVP facility", NSDG_ERR_UNSUPPORTED);
    }
    double turnCos() const { return std::cos(p.ocean_turning_angle * 0x1.1df46a2529d39p-6); }
    double turnSin() const { return std::sin(p.ocean_turning_angle * 0x1.1df46a2529d39p-6); }

    // ---- field access
    double* cgRef(const std::string& n)
    {
        if (n == "u") return u.p;
        if (n == "v") return v.p;
        if (n == "u0") return u0.p;
        if (n == "v0") return v0.p;
        if (n == "cgH") return cgH.p;
        if (n == "cgA") return cgA.p;
        if (n == "gradX") return gX.p;
        if (n == "gradY") return gY.p;
        if (n == "dStressX") return dSx.p;
        if (n == "dStressY") return dSy.p;
        if (n == "uocean") return uo.p;
        if (n == "vocean") return vo.p;
        if (n == "uwind") return ua.p;
        if (n == "vwind") return va.p;
        if (n == "lumpedcgmass") return lumped.p;
        if (brittle && n == "avgU") return avgU.p;
        if (brittle && n == "avgV") return avgV.p;
        return nullptr;
    }
    double* tensorRef(const std::string& n)
    {
        if (n == "s11") return sig.p;
        if (n == "s12") return sig.p + S * ne;
        if (n == "s22") return sig.p + 2 * S * ne;
        if (n == "e11") return eps.p;
        if (n == "e12") return eps.p + S * ne;
        if (n == "e22") return eps.p + 2 * S * ne;
        return nullptr;
    }
    // DG2CG of a DG0 host field into a CG node vector (setData of u, v, forcing).
    void dg0ToCG(const double* data, int ncomp, double* dst)
    {
        if (ncomp == DG) {
            uploadSoA(scratch, data, ne, DG, tmpDG.p, s);
            launchDG2CG<DG>(tmpDG.p, dst, 0, nullptr, nullptr, 0);
        } else {
            require(ncomp == 1, "field must have 1 or DG components");
            dg0a.upload(data, ne, s);
            launchDG2CG<1>(dg0a.p, dst, 0, nullptr, nullptr, 0);
        }
    }
    template <int D> void launchDG2CG(const double* s1, double* d1, int c1, const double* s2, double* d2, int c2)
    {
        Timed t(prof, s, "dg2cg");
        using T = TileShape<CG, 31, 7>;
        const unsigned grid = xcdGrid(((mesh.nx + T::TX - 1) / T::TX) * ((mesh.ny + T::TY - 1) / T::TY));
        k_dg2cg_tiled<CG, D><<<grid, kBlock, 0, s>>>(mesh.nx, mesh.ny, static_cast<int>(ne), s1, d1, c1, s2, d2, c2);
        checkLaunch("dg2cg");
    }
    void setField(const std::string& n, const double* data, int ncomp) override
    {
        need();
        require(data != nullptr, "data is NULL");
        if (double* c = (n == "u" || n == "v" || n == "uwind" || n == "vwind" || n == "uocean" || n == "vocean")
                ? cgRef(n)
                : nullptr) {
            dg0ToCG(data, ncomp, c);
        } else if (n == "hice" || n == "cice" || (brittle && n == "damage")) {
            double* f = n == "hice" ? h.p : (n == "cice" ? A.p : damage.p);
            require(ncomp == 1 || ncomp == DG, "hice/cice/damage need 1 or DG components");
            if (ncomp == DG)
                uploadSoA(scratch, data, ne, DG, f, s);
            else {
                deviceFill(f, ne * DG, 0.0, s);
                NSDG_HIP(hipMemcpyAsync(f, data, ne * sizeof(double), hipMemcpyHostToDevice, s));
            }
        } else if (n == "ssh") {
            require(ncomp == 1, "ssh is a DG0 field");
            ssh.upload(data, ne, s);
        } else if (double* t = tensorRef(n)) {
            require(ncomp == 1 || ncomp == S, "stress/strain need 1 or S components");
            if (ncomp == S)
                uploadSoA(scratch, data, ne, S, t, s);
            else {
                deviceFill(t, ne * S, 0.0, s);
                NSDG_HIP(hipMemcpyAsync(t, data, ne * sizeof(double), hipMemcpyHostToDevice, s));
            }
        } else {
            require(ncomp >= 1, "ncomp must be positive");
            passive[n] = { ncomp, std::vector<double>(data, data + ne * ncomp) };
        }
        NSDG_HIP(hipStreamSynchronize(s));
    }
    void getField(const std::string& n, double* out, int ncomp) override
    {
        need();
        require(out != nullptr, "out is NULL");
        if (n == "u" || n == "v" || n == "uiostress" || n == "viostress") {
            require(ncomp == 1 || ncomp == DG, "ncomp must be 1 or DG");
            const double* src = n == "u" ? u.p : v.p;
            if ((n == "uiostress" || n == "viostress") && brittle) {
                k_ice_ocean_stress_brittle<<<blocksFor(nn), kBlock, 0, s>>>(static_cast<int>(nn), p.F_ocean,
                    turnCos(), turnSin(), avgU.p, avgV.p, uo.p, vo.p, tmpCG.p, tmpCG2.p);
                checkLaunch("ice_ocean_stress");
                src = n == "uiostress" ? tmpCG.p : tmpCG2.p;
            } else if (n == "uiostress" || n == "viostress") {
                k_ice_ocean_stress<<<blocksFor(nn), kBlock, 0, s>>>(
                    static_cast<int>(nn), p.F_ocean, u.p, v.p, uo.p, vo.p, tmpCG.p, tmpCG2.p);
                checkLaunch("ice_ocean_stress");
                src = n == "uiostress" ? tmpCG.p : tmpCG2.p;
            }
            cg2dg.apply(mesh, src, tmpDG.p, nullptr, nullptr, s, prof);
            downloadSoA(scratch, tmpDG.p, ne, ncomp == 1 ? 1 : DG, out, s);
        } else if (n == "hice" || n == "cice" || (brittle && n == "damage")) {
            require(ncomp == 1 || ncomp == DG, "ncomp must be 1 or DG");
            downloadSoA(scratch, n == "hice" ? h.p : (n == "cice" ? A.p : damage.p), ne, ncomp, out, s);
        } else if (n == "ssh") {
            require(ncomp == 1, "ssh is a DG0 field");
            downloadSoA(scratch, ssh.p, ne, 1, out, s);
        } else if (const double* t = tensorRef(n)) {
            require(ncomp == 1 || ncomp == S, "ncomp must be 1 or S");
            downloadSoA(scratch, t, ne, ncomp, out, s);
        } else {
            auto it = passive.find(n);
            require(it != passive.end(), "unknown field " + n);
            const int nc = it->second.first;
            const auto& d = it->second.second;
            if (ncomp == nc)
                std::memcpy(out, d.data(), d.size() * sizeof(double));
            else {
                require(ncomp == 1, "ncomp must be 1 or the stored component count");
                for (size_t e = 0; e < ne; ++e)
                    out[e] = d[e * nc];
            }
        }
    }
    void getCG(const std::string& n, double* out) override
    {
        need();
        double* c = cgRef(n);
        require(c != nullptr, "unknown CG vector " + n);
        NSDG_HIP(hipMemcpyAsync(out, c, nn * sizeof(double), hipMemcpyDeviceToHost, s));
        NSDG_HIP(hipStreamSynchronize(s));
    }
    void setCG(const std::string& n, const double* in) override
    {
        need();
        double* c = cgRef(n);
        require(c != nullptr && n != "lumpedcgmass", "unknown or read-only CG vector " + n);
        NSDG_HIP(hipMemcpyAsync(c, in, nn * sizeof(double), hipMemcpyHostToDevice, s));
        NSDG_HIP(hipStreamSynchronize(s));
    }

    // ---- the time step
    // DynamicsKernel::advectionAndLimits (DynamicsKernel.hpp:160-172); the brittle kernel
    // advects with the velocity averaged over the previous step's substeps
    // (BrittleCGDynamicsKernel.hpp:89-90).
    // With the mass budget on, the Dirichlet outflow and the limiters' change of h and A
    // are tallied per element (tally: outflow h, outflow A, limiter h, limiter A).
    void advectAsync(double dt)
    {
        cg2dg.apply(mesh, brittle ? avgU.p : u.p, tr->velx(), brittle ? avgV.p : v.p, tr->vely(), s, prof);
        tr->reinitNormalVelocity();
        double* t = budgetOn ? tally.p : nullptr;
        tr->step(dt, A.p, t ? t + ne : nullptr);
        tr->step(dt, h.p, t);
        tr->limit(A.p, true, 1.0, true, 0.0, t ? t + 3 * ne : nullptr);
        tr->limit(h.p, false, 0.0, true, 0.0, t ? t + 2 * ne : nullptr);
    }
    void prepareIterationAsync()
    {
        const int clampH = mesh.hasPeriodic ? 0 : 1, clampA = mesh.hasPeriodic ? 0 : 2;
        launchDG2CG<DG>(h.p, cgH.p, clampH, A.p, cgA.p, clampA);
        if (mesh.hasPeriodic) {
            for (const auto& seg : mesh.segments) {
                k_periodic_average<CG><<<1, 1, 0, s>>>(
                    mesh.nx, mesh.periodic.p + 4 * seg.first, static_cast<int>(seg.second), cgH.p, cgA.p);
                checkLaunch("periodic_average");
            }
        }
        // sea-surface height gradient
        {
            Timed t(prof, s, "ssh_gradient");
            k_dg2cg<1, 1><<<blocksFor(nn1), kBlock, 0, s>>>(
                mesh.nx, mesh.ny, static_cast<int>(ne), ssh.p, cg1a.p, 0, nullptr, nullptr, 0);
            checkLaunch("ssh_dg2cg");
            auto go = [&](auto kern) {
                kern<<<blocksFor(nn1), kBlock, 0, s>>>(
                    mesh.nx, mesh.ny, cg1a.p, mops.p, opStride(), L::SX, L::SY, lumped1.p, cg1b.p, cg1c.p);
                checkLaunch("ssh_gradient");
            };
            if (mesh.uniform)
                go(k_ssh_gradient<true>);
            else
                go(k_ssh_gradient<false>);
            k_ssh_prolong<CG><<<blocksFor(nn), kBlock, 0, s>>>(mesh.nx, mesh.ny, cg1b.p, cg1c.p, gX.p, gY.p);
            checkLaunch("ssh_prolong");
        }
        if (mesh.hasPeriodic) {
            k_clamp_HA<CG><<<blocksFor(nn), kBlock, 0, s>>>(cgH.p, cgA.p, static_cast<int>(nn));
            checkLaunch("clamp");
        }
    }
    ElemArgs elemArgs()
    {
        ElemArgs a {};
        a.nx = mesh.nx;
        a.ny = mesh.ny;
        a.ne = static_cast<int>(ne);
        a.ocean = mesh.ocean.p;
        a.u = u.p;
        a.v = v.p;
        a.h = h.p;
        a.A = A.p;
        a.eps = eps.p;
        a.epsOut = eps.p;
        a.sig = sig.p;
        a.ops = mops.p;
        a.stride = opStride();
        a.Pg = strength.p;
        a.sigF = sigF.p;
        a.PgF = strengthF.p;
        a.iter = subcycleCount;
        a.sph = mesh.spherical ? 1 : 0;
        a.Pstar = p.Pstar;
        a.DM2 = p.DeltaMin * p.DeltaMin;
        a.invAlpha = 1.0 / p.alpha;
        a.relax = 1.0 - 1.0 / p.alpha;
        a.xiX = mesh.xiX;
        a.etaX = mesh.etaX;
        a.xiY = mesh.xiY;
        a.etaY = mesh.etaY;
        return a;
    }
    NodeArgs nodeArgs()
    {
        NodeArgs a {};
        a.nx = mesh.nx;
        a.ny = mesh.ny;
        a.ne = static_cast<int>(ne);
        a.nn = static_cast<int>(nn);
        a.ocean = mesh.ocean.p;
        a.dirNode = dirNode.p;
        a.sig = sig.p;
        a.ops = mops.p;
        a.stride = opStride();
        a.u = u.p;
        a.v = v.p;
        a.u0 = u0.p;
        a.v0 = v0.p;
        a.uo = uo.p;
        a.vo = vo.p;
        a.ua = ua.p;
        a.va = va.p;
        a.H = cgH.p;
        a.A = cgA.p;
        a.gx = gX.p;
        a.gy = gY.p;
        a.ML = lumped.p;
        a.ocn = fzOcn.p;
        a.cf = fzCf.p;
        a.qf = fzQ.p;
        a.dSx = dSx.p;
        a.dSy = dSy.p;
        a.rho_ice = p.rho_ice;
        a.F_atm = p.F_atm;
        a.F_ocean = p.F_ocean;
        a.fc = p.fc;
        a.gravity = p.gravity;
        a.beta = p.beta;
        a.dt = deltaT;
        a.xiX = mesh.xiX;
        a.etaX = mesh.etaX;
        a.xiY = mesh.xiY;
        a.etaY = mesh.etaY;
        a.detJ = mesh.detJ;
        a.sph = mesh.spherical ? 1 : 0;
        return a;
    }
    // Two-pass subcycle (general meshes, periodic meshes); FP64 only, the FP32 variant always
    // runs the one-pass kernel.
    void elementPass()
    {
        require(precision == NSDG_PREC_FP64, "the two-pass subcycle is FP64", NSDG_ERR_STATE);
        Timed t(prof, s, "mevp_element");
        const ElemArgs a = elemArgs();
        if (mesh.uniform)
            k_mevp_element<CG, DG, true><<<blocksFor(ne), kBlock, 0, s>>>(a);
        else
            k_mevp_element<CG, DG, false><<<blocksFor(ne), kBlock, 0, s>>>(a);
        checkLaunch("mevp_element");
    }
    // Node passes of the two-pass subcycle (meshes with periodic segments): the split
    // sequence in the reference's order, with the periodic averaging of the divergence.
    void nodePass()
    {
        divergenceAsync();
        momentumAsync();
        boundariesAsync();
    }
    void divergenceAsync()
    {
        Timed t(prof, s, "divergence");
        const NodeArgs a = nodeArgs();
        if (mesh.uniform)
            k_divergence<CG, true><<<blocksFor(nn), kBlock, 0, s>>>(a);
        else
            k_divergence<CG, false><<<blocksFor(nn), kBlock, 0, s>>>(a);
        checkLaunch("divergence");
        for (const auto& seg : mesh.segments) {
            k_periodic_average<CG><<<1, 1, 0, s>>>(
                mesh.nx, mesh.periodic.p + 4 * seg.first, static_cast<int>(seg.second), dSx.p, dSy.p);
            checkLaunch("periodic_average");
        }
    }
    void momentumAsync()
    {
        Timed t(prof, s, "momentum");
        k_momentum<<<blocksFor(nn), kBlock, 0, s>>>(nodeArgs());
        checkLaunch("momentum");
    }
    void boundariesAsync()
    {
        k_boundaries<<<blocksFor(nn), kBlock, 0, s>>>(nodeArgs());
        checkLaunch("boundaries");
    }
    void beginIterationAsync(double dt)
    {
        deviceCopy(u0.p, u.p, nn, s);
        deviceCopy(v0.p, v.p, nn, s);
        deltaT = dt;
    }
    // Terms fixed over the subcycles, evaluated from the current fields at the start of
    // every run of subcycles (so set_field between runs is honoured).
    void frozenAsync()
    {
        Timed t(prof, s, "frozen_terms");
        if (precision == NSDG_PREC_FP32)
            k_strength<CG, DG, float><<<blocksFor(ne), kBlock, 0, s>>>(static_cast<int>(ne), h.p, A.p, p.Pstar, strengthF.p);
        else
            k_strength<CG, DG, double><<<blocksFor(ne), kBlock, 0, s>>>(static_cast<int>(ne), h.p, A.p, p.Pstar, strength.p);
        checkLaunch("strength");
        k_node_coeffs<<<blocksFor(nn), kBlock, 0, s>>>(nodeArgs(), fzOcn.p, fzCf.p, fzQ.p);
        checkLaunch("node_coeffs");
    }
    // A run of subcycles: the frozen terms once, the FP32 copy of the stress on entry and
    // back on exit, the passes in between.
    void beginRun()
    {
        frozenAsync();
        if (precision == NSDG_PREC_FP32) {
            k_cast<double, float><<<blocksFor(sig.n), kBlock, 0, s>>>(sig.p, sigF.p, sig.n);
            checkLaunch("stress_to_fp32");
        }
    }
    // One-pass subcycles: uniform meshes without periodic segments.
    bool fusedSubcycle() const { return !mesh.hasPeriodic; }
    template <class X> static void swapBuf(DBuf<X>& x, DBuf<X>& y)
    {
        std::swap(x.p, y.p);
        std::swap(x.n, y.n);
    }
    void fusedPass()
    {
        Timed t(prof, s, "mevp_fused");
        FusedArgs f {};
        f.e = elemArgs();
        f.n = nodeArgs();
        f.sigOut = sigAlt.p;
        f.sigFOut = sigFAlt.p;
        f.uOut = uAlt.p;
        f.vOut = vAlt.p;
        f.geo = geo.p;
        using T = FusedTile<CG>;
        const unsigned grid = xcdGrid(((mesh.nx + T::TX - 1) / T::TX) * ((mesh.ny + T::TY - 1) / T::TY));
        if (!mesh.uniform) {
            require(precision == NSDG_PREC_FP64 && geo.n == static_cast<size_t>(GeoLayout<CG>::K) * ne,
                "general meshes run the FP64 one-pass subcycle from their geometry tables", NSDG_ERR_STATE);
            if (mesh.spherical)
                k_mevp_fused_geom<CG, DG, true><<<grid, kBlock, 0, s>>>(f);
            else
                k_mevp_fused_geom<CG, DG, false><<<grid, kBlock, 0, s>>>(f);
        } else if (precision == NSDG_PREC_FP32) {
            require(sigF.n == 3 * S * ne && sigFAlt.n == sigF.n, "FP32 buffers not allocated", NSDG_ERR_STATE);
            k_mevp_fused<CG, DG, float><<<grid, kBlock, 0, s>>>(f);
        } else {
            k_mevp_fused<CG, DG, double><<<grid, kBlock, 0, s>>>(f);
        }
        checkLaunch("mevp_fused");
        // the new state becomes the current one; the old buffers are the next pass's outputs
        if (precision == NSDG_PREC_FP32)
            swapBuf(sigF, sigFAlt);
        else
            swapBuf(sig, sigAlt);
        swapBuf(u, uAlt);
        swapBuf(v, vAlt);
    }
    // The strain output of the reference is that of the velocity entering the last subcycle
    // (projectVelocityToStrain runs first in every subcycle); the subcycle kernels keep it
    // in registers, so it is stored once, before the last pass of a run.
    void strainAsync()
    {
        Timed t(prof, s, "strain");
        const ElemArgs a = elemArgs();
        if (mesh.uniform)
            k_strain<CG, true><<<blocksFor(ne), kBlock, 0, s>>>(a);
        else if (mesh.spherical)
            k_strain_geom<CG, true><<<blocksFor(ne), kBlock, 0, s>>>(a, geo.p);
        else
            k_strain_geom<CG, false><<<blocksFor(ne), kBlock, 0, s>>>(a, geo.p);
        checkLaunch("strain");
    }
    // Two subcycles per pass: uniform meshes without periodic segments (FP64 or FP32), by default
    // from NSDG_PAIR_MIN_ELEMENTS up (every such mesh: faster per step than k_mevp_fused
    // from 128^2 to 2048^2 except near 320^2, within 3.5% there)
    // (below, the state largely stays in the memory-side cache).
    bool pairSubcycles() const
    {
        const bool can = fusedSubcycle() && mesh.uniform;
        return can && (pairing == NSDG_PAIR_ALWAYS || (pairing == NSDG_PAIR_AUTO && ne >= NSDG_PAIR_MIN_ELEMENTS));
    }
    void pairPass()
    {
        Timed t(prof, s, "mevp_pair");
        FusedArgs f {};
        f.e = elemArgs();
        f.n = nodeArgs();
        f.sigOut = sigAlt.p;
        f.sigFOut = sigFAlt.p;
        f.uOut = uAlt.p;
        f.vOut = vAlt.p;
        auto gridOf = [&](auto tile) {
            using T = decltype(tile);
            return xcdGrid(((mesh.nx + T::TX - 1) / T::TX) * ((mesh.ny + T::TY - 1) / T::TY));
        };
        if (precision == NSDG_PREC_FP32) {
            require(sigF.n == 3 * S * ne && sigFAlt.n == sigF.n, "FP32 buffers not allocated", NSDG_ERR_STATE);
            using T = PairTile<CG, float>;
            k_mevp_pair<CG, DG, float><<<gridOf(T {}), T::BLOCK, 0, s>>>(f);
        } else {
            using T = PairTile<CG, double>;
            k_mevp_pair<CG, DG, double><<<gridOf(T {}), T::BLOCK, 0, s>>>(f);
        }
        checkLaunch("mevp_pair");
        if (precision == NSDG_PREC_FP32)
            swapBuf(sigF, sigFAlt);
        else
            swapBuf(sig, sigAlt);
        swapBuf(u, uAlt);
        swapBuf(v, vAlt);
    }
    void runPasses(int n)
    {
        int k = 0;
        // pairs while a subcycle remains after the pair: the last one runs alone, after the
        // strain output of the velocity entering it
        if (pairSubcycles())
            for (; k + 2 < n; k += 2) {
                pairPass();
                subcycleCount += 2;
            }
        for (; k < n; ++k) {
            if (k == n - 1)
                strainAsync();
            if (fusedSubcycle()) {
                fusedPass();
            } else {
                elementPass();
                nodePass();
            }
            ++subcycleCount;
        }
    }
    void endRun()
    {
        if (precision == NSDG_PREC_FP32) {
            k_cast<float, double><<<blocksFor(sig.n), kBlock, 0, s>>>(sigF.p, sig.p, sig.n);
            checkLaunch("stress_to_fp64");
        }
    }
    void subcyclesAsync(int n)
    {
        if (n <= 0)
            return;
        beginRun();
        runPasses(n);
        endRun();
    }
    // Subcycles in chunks until the momentum residual (which reads only h, A, u, v and the
    // forcing) is at most rTarget or maxN have run. Chunks are at most `every` long; once
    // the residual has fallen over a chunk, the next one ends where its current rate of
    // decrease reaches rTarget (at least one subcycle), so the run stops within a few
    // subcycles of the crossing instead of up to `every` past it; a residual within 1% of
    // the target that did not fall over the last chunk gets chunks of at most 5.
    void subcyclesToResidual(double rTarget, int maxN, int every, double rStart, int* done, double* rFinal) override
    {
        need();
        needMEVP("the residual-driven subcycle loop");
        require(deltaT > 0, "begin_iteration (or update) must set dt first", NSDG_ERR_STATE);
        require(maxN >= 0 && every >= 1, "need max_subcycles >= 0 and check_every >= 1");
        double r = rStart;
        if (!(rStart >= 0.0))
            residual(&r, nullptr, nullptr);
        int n = 0;
        if (r > rTarget && maxN > 0) {
            beginRun();
            double rPrev = r;
            int last = 0;
            while (r > rTarget && n < maxN) {
                int step = std::min(every, maxN - n);
                if (last > 0 && rPrev > r) {
                    const double need = std::ceil((r - rTarget) / ((rPrev - r) / last));
                    if (need < step)
                        step = std::max(1, static_cast<int>(need));
                } else if (last > 0 && r - rTarget <= 0.01 * rTarget) {
                    step = std::min(step, 5); // within 1% of the target but not falling: short chunks
                }
                rPrev = r;
                runPasses(step);
                n += step;
                last = step;
                residual(&r, nullptr, nullptr);
            }
            endRun();
            synchronize();
        }
        if (done)
            *done = n;
        if (rFinal)
            *rFinal = r;
    }

    // ---- brittle substeps (BrittleCGDynamicsKernel.hpp:92-134)
    BrittleArgs brittleArgs() const
    {
        BrittleArgs b {};
        b.dt = deltaT;
        b.nu = meb.nu0;
        b.young = meb.young;
        b.P0 = meb.P0;
        b.expP = meb.exponent_compression_factor + 1.0;
        b.compaction = meb.compaction_param;
        b.tau0 = meb.undamaged_time_relaxation_sigma;
        b.expR = meb.exponent_relaxation_sigma - 1.0;
        b.tanPhi = meb.tan_phi;
        const double scale = std::sqrt(0.1 / hElem);
        b.cohesion = meb.C_lab * scale;
        b.compr = meb.compr_strength * scale;
        b.duFac = deltaT / (1.0 - meb.nu0 * meb.nu0);
        b.dtOverT = deltaT / (hElem * std::sqrt(2.0 * (1.0 + meb.nu0) * p.rho_ice));
        b.damageIn = damage.p;
        b.damage = damage.p;
        b.geo = mesh.uniform ? nullptr : brittleGeo.p;
        b.Clab = meb.C_lab;
        b.comprStrength = meb.compr_strength;
        b.tdFac = std::sqrt(2.0 * (1.0 + meb.nu0) * p.rho_ice);
        return b;
    }
    template <bool STRAIN> void brittleStressAsync()
    {
        Timed t(prof, s, "bbm_element");
        if (mesh.uniform)
            k_bbm_element<CG, DG, STRAIN, true><<<blocksFor(ne), kBlock, 0, s>>>(elemArgs(), brittleArgs());
        else
            k_bbm_element<CG, DG, STRAIN, false><<<blocksFor(ne), kBlock, 0, s>>>(elemArgs(), brittleArgs());
        checkLaunch("bbm_element");
    }
    BrittleNode brittleNode() const
    {
        return BrittleNode { turnCos(), turnSin(), static_cast<double>(p.nsteps), avgU.p, avgV.p, fzAF.p };
    }
    // The node terms fixed over the substeps of a step (from the current fields and deltaT).
    void brittleFrozenAsync()
    {
        Timed t(prof, s, "frozen_terms");
        k_brittle_coeffs<<<blocksFor(nn), kBlock, 0, s>>>(nodeArgs(), fzOcn.p, fzCf.p, fzQ.p, fzAF.p);
        checkLaunch("brittle_coeffs");
    }
    void brittleMomentumAsync()
    {
        Timed t(prof, s, "brittle_momentum");
        k_brittle_momentum<<<blocksFor(nn), kBlock, 0, s>>>(nodeArgs(), brittleNode());
        checkLaunch("brittle_momentum");
    }
    void brittleFusedPass()
    {
        Timed t(prof, s, "bbm_fused");
        BrittleFusedArgs f {};
        f.e = elemArgs();
        f.n = nodeArgs();
        f.b = brittleArgs();
        f.pol = brittleNode();
        f.sigOut = sigAlt.p;
        f.damageOut = damageAlt.p;
        f.uOut = uAlt.p;
        f.vOut = vAlt.p;
        f.geo = geo.p;
        using T = FusedTile<CG>;
        const unsigned grid = xcdGrid(((mesh.nx + T::TX - 1) / T::TX) * ((mesh.ny + T::TY - 1) / T::TY));
        const bool pow5 = meb.exponent_relaxation_sigma == 5.0 && meb.exponent_compression_factor == 1.5;
        if (!mesh.uniform) {
            require(geo.p && brittleGeo.p, "general meshes need their geometry tables", NSDG_ERR_STATE);
            if (mesh.spherical)
                pow5 ? k_bbm_fused_geom<CG, DG, true, true><<<grid, kBlock, 0, s>>>(f)
                     : k_bbm_fused_geom<CG, DG, false, true><<<grid, kBlock, 0, s>>>(f);
            else
                pow5 ? k_bbm_fused_geom<CG, DG, true, false><<<grid, kBlock, 0, s>>>(f)
                     : k_bbm_fused_geom<CG, DG, false, false><<<grid, kBlock, 0, s>>>(f);
        } else if (pow5) {
            k_bbm_fused<CG, DG, true><<<grid, kBlock, 0, s>>>(f);
        } else {
            k_bbm_fused<CG, DG, false><<<grid, kBlock, 0, s>>>(f);
        }
        checkLaunch("bbm_fused");
        swapBuf(sig, sigAlt);
        swapBuf(damage, damageAlt);
        swapBuf(u, uAlt);
        swapBuf(v, vAlt);
    }
    // Substeps: one pass each on meshes without periodic segments (the strain of the last
    // substep's velocity stored before it), else the reference's pass sequence with the
    // per-element operators, sizes and damage projections.
    void brittleSubstepsAsync(int n)
    {
        if (n > 0)
            brittleFrozenAsync();
        for (int k = 0; k < n; ++k) {
            if (!mesh.hasPeriodic) {
                if (k == n - 1)
                    strainAsync();
                brittleFusedPass();
            } else {
                brittleStressAsync<true>();
                divergenceAsync();
                brittleMomentumAsync();
                boundariesAsync();
            }
        }
    }
    // Advection of h and A with the averaged velocity, transport of the stress (stress
    // space, RK2) and of the damage (then limited to [1e-12, 1]), then nsteps substeps of
    // dt / nsteps from a zero average.
    void updateBrittle(double dt)
    {
        advectAsync(dt);
        cg2dgS.apply(mesh, avgU.p, stressTr->velx(), avgV.p, stressTr->vely(), s, prof);
        stressTr->reinitNormalVelocity();
        for (int c = 0; c < 3; ++c)
            stressTr->step(dt, sig.p + static_cast<size_t>(c) * S * ne);
        tr->step(dt, damage.p);
        tr->limit(damage.p, true, 1.0, true, 1e-12);
        prepareIterationAsync();
        deltaT = dt / p.nsteps;
        deviceFill(avgU.p, nn, 0.0, s);
        deviceFill(avgV.p, nn, 0.0, s);
        brittleSubstepsAsync(p.nsteps);
    }

    void update(double dt) override
    {
        need();
        require(dt > 0, "dt must be positive");
        if (brittle) {
            updateBrittle(dt);
            return;
        }
        advectAsync(dt);
        prepareIterationAsync();
        beginIterationAsync(dt);
        subcyclesAsync(p.nsteps);
    }
    void advect(double dt) override
    {
        need();
        advectAsync(dt);
        synchronize();
    }
    void prepareIteration() override
    {
        need();
        prepareIterationAsync();
        synchronize();
    }
    void beginIteration(double dt) override
    {
        need();
        require(dt > 0, "dt must be positive");
        beginIterationAsync(dt);
        synchronize();
    }
    void strain() override
    {
        need();
        const ElemArgs a = elemArgs();
        if (mesh.uniform)
            k_strain<CG, true><<<blocksFor(ne), kBlock, 0, s>>>(a);
        else
            k_strain<CG, false><<<blocksFor(ne), kBlock, 0, s>>>(a);
        checkLaunch("strain");
        synchronize();
    }
    void stress(int repeat) override
    {
        need();
        if (brittle) {
            require(deltaT > 0, "begin_iteration must set dt first", NSDG_ERR_STATE);
            for (int r = 0; r < repeat; ++r)
                brittleStressAsync<false>();
            synchronize();
            return;
        }
        const ElemArgs a = elemArgs();
        for (int r = 0; r < repeat; ++r) {
            if (mesh.uniform)
                k_stress<CG, DG, true><<<blocksFor(ne), kBlock, 0, s>>>(a);
            else
                k_stress<CG, DG, false><<<blocksFor(ne), kBlock, 0, s>>>(a);
            checkLaunch("stress");
        }
        synchronize();
    }
    void divergence() override
    {
        need();
        divergenceAsync();
        synchronize();
    }
    void momentum() override
    {
        need();
        require(deltaT > 0, "begin_iteration must set dt first", NSDG_ERR_STATE);
        if (brittle) {
            brittleFrozenAsync();
            brittleMomentumAsync();
        } else {
            momentumAsync();
        }
        synchronize();
    }
    void boundaries() override
    {
        need();
        boundariesAsync();
        synchronize();
    }
    void subcycles(int n) override
    {
        need();
        require(deltaT > 0, "begin_iteration (or update) must set dt first", NSDG_ERR_STATE);
        if (brittle)
            brittleSubstepsAsync(n);
        else
            subcyclesAsync(n);
    }
    // sigma_out <- Pi sigma_VP(eps(u)) (FP64, from the current h, A, u, v).
    void vpStressAsync(double* out)
    {
        k_strength<CG, DG, double><<<blocksFor(ne), kBlock, 0, s>>>(static_cast<int>(ne), h.p, A.p, p.Pstar, strength.p);
        checkLaunch("strength");
        ElemArgs a = elemArgs();
        a.sig = out;
        a.invAlpha = 1.0;
        a.relax = 0.0;
        if (mesh.uniform)
            k_vp_stress<CG, DG, true><<<blocksFor(ne), kBlock, 0, s>>>(a);
        else
            k_vp_stress<CG, DG, false><<<blocksFor(ne), kBlock, 0, s>>>(a);
        checkLaunch("vp_stress");
    }
    // r = |R| / |M_L A tau_a| over the non-Dirichlet nodes (see k_residual_nodes).
    void residual(double* r, double* normR, double* normD) override
    {
        need();
        needMEVP("the momentum residual");
        require(deltaT > 0, "begin_iteration (or update) must set dt first", NSDG_ERR_STATE);
        if (fusedSubcycle() && mesh.uniform) {
            Timed t(prof, s, "residual");
            FusedArgs f {};
            f.e = elemArgs();
            f.e.invAlpha = 1.0;
            f.e.relax = 0.0;
            f.n = nodeArgs();
            using T = FusedTile<CG>;
            const unsigned grid = xcdGrid(((mesh.nx + T::TX - 1) / T::TX) * ((mesh.ny + T::TY - 1) / T::TY));
            require(rpart.n >= 4 * static_cast<size_t>(grid), "residual partials too small", NSDG_ERR_STATE);
            k_residual_fused<CG, DG><<<grid, kBlock, 0, s>>>(f, rpart.p);
            checkLaunch("residual_fused");
            k_sum_final<2><<<1, kBlock, 0, s>>>(rpart.p, static_cast<int>(grid), rsum.p);
            checkLaunch("residual_final");
        } else {
            Timed t(prof, s, "residual");
            vpStressAsync(sigVP.p);
            NodeArgs a = nodeArgs();
            a.sig = sigVP.p;
            a.dSx = tmpCG.p;
            a.dSy = tmpCG2.p;
            if (mesh.uniform)
                k_divergence<CG, true><<<blocksFor(nn), kBlock, 0, s>>>(a);
            else
                k_divergence<CG, false><<<blocksFor(nn), kBlock, 0, s>>>(a);
            checkLaunch("residual_divergence");
            for (const auto& seg : mesh.segments) {
                k_periodic_average<CG><<<1, 1, 0, s>>>(
                    mesh.nx, mesh.periodic.p + 4 * seg.first, static_cast<int>(seg.second), tmpCG.p, tmpCG2.p);
                checkLaunch("periodic_average");
            }
            k_residual_nodes<<<kResidualBlocks, kBlock, 0, s>>>(a, tmpCG.p, tmpCG2.p, rpart.p);
            checkLaunch("residual_nodes");
            k_sum_final<2><<<1, kBlock, 0, s>>>(rpart.p, kResidualBlocks, rsum.p);
            checkLaunch("residual_final");
        }
        double h2[2];
        rsum.download(h2, 2, s);
        NSDG_HIP(hipStreamSynchronize(s));
        if (prof.on)
            prof.collect();
        const double nr = std::sqrt(h2[0]), nd = std::sqrt(h2[1]);
        if (normR)
            *normR = nr;
        if (normD)
            *normD = nd;
        if (r)
            *r = nd > 0 ? nr / nd : nr;
    }
    int gridNx() const override { return mesh.nx; }
    int gridNy() const override { return mesh.ny; }
    void needCG2() const
    {
        need();
        needMEVP("the learned warm start");
        require(CG == 2, "the learned warm start is built for CG2 velocities", NSDG_ERR_UNSUPPORTED);
    }
    void features(uint16_t* x) override
    {
        needCG2();
        require(x != nullptr && reinterpret_cast<uintptr_t>(x) % 16 == 0, "feature buffer must be 16-byte aligned");
        k_features<S><<<blocksFor(ne), kBlock, 0, s>>>(mesh.nx, mesh.ny, h.p, A.p, mesh.ocean.p, ua.p, va.p, uo.p,
            vo.p, u0.p, v0.p, sig.p, p.Pstar, x);
        checkLaunch("features");
        synchronize();
    }
    void velocityTarget(float* t, float* w) override
    {
        needCG2();
        require(t != nullptr, "target buffer is NULL");
        k_velocity_target<<<blocksFor(ne), kBlock, 0, s>>>(mesh.nx, mesh.ny, u.p, v.p, u0.p, v0.p, mesh.ocean.p, t, w);
        checkLaunch("velocity_target");
        synchronize();
    }
    void applyVelocity(const float* y) override
    {
        needCG2();
        require(y != nullptr, "prediction buffer is NULL");
        k_apply_velocity<<<blocksFor(nn), kBlock, 0, s>>>(mesh.nx, mesh.ny, y, u0.p, v0.p, dirNode.p, u.p, v.p);
        checkLaunch("apply_velocity");
        synchronize();
    }
    void saveStart() override
    {
        need();
        if (uStart.n != nn) {
            uStart.alloc(nn);
            vStart.alloc(nn);
            sigStart.alloc(sig.n);
        }
        deviceCopy(uStart.p, u.p, nn, s);
        deviceCopy(vStart.p, v.p, nn, s);
        deviceCopy(sigStart.p, sig.p, sig.n, s);
        synchronize();
    }
    void restoreStart() override
    {
        need();
        require(uStart.n == nn, "no saved start", NSDG_ERR_STATE);
        deviceCopy(u.p, uStart.p, nn, s);
        deviceCopy(v.p, vStart.p, nn, s);
        deviceCopy(sig.p, sigStart.p, sig.n, s);
        synchronize();
    }
    // sigma <- Pi sigma_VP(eps(u)): the converged mEVP stress of the current velocity.
    void stressFromVelocity() override
    {
        need();
        needMEVP("the viscous-plastic stress of a velocity");
        vpStressAsync(sig.p);
        synchronize();
    }
    // FP32 stress storage and element arithmetic exist for the sum-factorised kernels of
    // uniform meshes with the fused node pass (no periodic segments).
    void checkPrecision() const
    {
        if (precision == NSDG_PREC_FP32)
            needMEVP("the FP32 variant");
        if (precision == NSDG_PREC_FP32)
            require(mesh.uniform && !mesh.hasPeriodic,
                "the FP32 variant needs a uniform (affine) mesh without periodic segments", NSDG_ERR_UNSUPPORTED);
    }
    void allocPrecision()
    {
        if (precision == NSDG_PREC_FP32 && ready && sigF.n != 3 * S * ne) {
            sigF.alloc(3 * S * ne);
            sigFAlt.alloc(3 * S * ne);
            strengthF.alloc(static_cast<size_t>(L::Q) * ne);
        }
    }
    void setPairing(int mode) override
    {
        require(mode == NSDG_PAIR_AUTO || mode == NSDG_PAIR_ALWAYS || mode == NSDG_PAIR_NEVER,
            "pairing must be NSDG_PAIR_AUTO, NSDG_PAIR_ALWAYS or NSDG_PAIR_NEVER");
        pairing = mode;
    }
    void setPrecision(int pr) override
    {
        require(pr == NSDG_PREC_FP64 || pr == NSDG_PREC_FP32, "precision must be NSDG_PREC_FP64 or NSDG_PREC_FP32");
        const int old = precision;
        precision = pr;
        if (ready) {
            try {
                checkPrecision();
            } catch (...) {
                precision = old;
                throw;
            }
            allocPrecision();
        }
    }
    void benchmarkWind(double L, double t, double xoff, double yoff, double vmax) override
    {
        stormWind(L, t, &xoff, &yoff, &vmax, 1);
    }
    // Every storm follows the reference cyclone's track (centre at half the box, moving by
    // a fifth of it per day along the diagonal), shifted by its own offset.
    void stormWind(double L, double t, const double* xoff, const double* yoff, const double* vmax, int n) override
    {
        need();
        require(n >= 1 && n <= kMaxStorms, "between 1 and " + std::to_string(kMaxStorms) + " storms");
        require(xoff && yoff && vmax, "storm arrays are NULL");
        const double dx = L / mesh.nx, dy = L / mesh.ny, tf = t / 86400.0;
        Storms st {};
        st.n = n;
        for (int i = 0; i < n; ++i) {
            st.x0[i] = mesh.nx * dx * 0.5 * (1 + tf / 5.0) + xoff[i];
            st.y0[i] = mesh.ny * dy * 0.5 * (1 + tf / 5.0) + yoff[i];
            st.vmax[i] = vmax[i];
        }
        {
            Timed tm(prof, s, "benchmark_wind");
            k_benchmark_wind<<<blocksFor(ne), kBlock, 0, s>>>(mesh.nx, mesh.ny, dx, dy, st, dg0a.p, dg0b.p);
            checkLaunch("benchmark_wind");
        }
        launchDG2CG<1>(dg0a.p, ua.p, 0, dg0b.p, va.p, 0);
    }
    void resetDG0(const std::string& n) override
    {
        need();
        double* f = nullptr;
        int nc = 0;
        if (n == "hice" || n == "cice") {
            f = n == "hice" ? h.p : A.p;
            nc = DG;
        } else if (double* t = tensorRef(n)) {
            f = t;
            nc = S;
        }
        require(f != nullptr, "reset_dg0 applies to hice, cice and the tensors");
        if (nc > 1) {
            k_zero_higher<<<blocksFor(ne * (nc - 1)), kBlock, 0, s>>>(f, static_cast<int>(ne), nc);
            checkLaunch("zero_higher");
        }
    }
    void mass(double* vol, double* area) override
    {
        need();
        const unsigned nb = blocksFor(ne);
        k_mass_partial<DG><<<nb, kBlock, 0, s>>>(view(mesh), static_cast<int>(ne), h.p, A.p, part.p);
        checkLaunch("mass_partial");
        k_sum_final<2><<<1, kBlock, 0, s>>>(part.p, static_cast<int>(nb), res.p);
        checkLaunch("mass_final");
        double r[2];
        res.download(r, 2, s);
        NSDG_HIP(hipStreamSynchronize(s));
        if (vol)
            *vol = r[0];
        if (area)
            *area = r[1];
    }
    // The sums of mass() with the block partials and totals in the file-scope scratch
    // array instead of the kernel's own buffers: no allocation in the call.
    void massNoAlloc(double* vol, double* area) override
    {
        need();
        require(ne <= kNoAllocMaxElements,
            "nsdg_dyn_mass_noalloc takes at most " + std::to_string(kNoAllocMaxElements) + " elements, the mesh has "
                + std::to_string(ne),
            NSDG_ERR_ARG);
        double* const sp = noAllocScratch(device);
        const unsigned nb = blocksFor(ne);
        k_mass_partial<DG><<<nb, kBlock, 0, s>>>(view(mesh), static_cast<int>(ne), h.p, A.p, sp + kNoAllocPartials);
        checkLaunch("mass_partial");
        k_sum_final<2><<<1, kBlock, 0, s>>>(sp + kNoAllocPartials, static_cast<int>(nb), sp);
        checkLaunch("mass_final");
        double r[2];
        NSDG_HIP(hipMemcpyAsync(r, sp, sizeof r, hipMemcpyDeviceToHost, s));
        NSDG_HIP(hipStreamSynchronize(s));
        if (vol)
            *vol = r[0];
        if (area)
            *area = r[1];
    }
    // ---- advection mass budget (no counterpart in the reference)
    void setBudget(bool on) override
    {
        need();
        if (on && !budgetOn) {
            tally.alloc(4 * ne);
            deviceFill(tally.p, tally.n, 0.0, s);
        }
        if (!on)
            tally.release();
        budgetOn = on;
        NSDG_HIP(hipStreamSynchronize(s));
    }
    void budget(double* out, bool reset) override
    {
        need();
        require(budgetOn, "the mass budget is off (nsdg_dyn_set_budget)", NSDG_ERR_STATE);
        const unsigned nb = std::min<unsigned>(blocksFor(ne), kTallyBlocks);
        k_tally_partial<4><<<nb, kBlock, 0, s>>>(tally.p, static_cast<int>(ne), part.p);
        checkLaunch("tally_partial");
        k_sum_final<4><<<1, kBlock, 0, s>>>(part.p, static_cast<int>(nb), res.p);
        checkLaunch("tally_final");
        double r[4];
        res.download(r, 4, s);
        if (reset)
            deviceFill(tally.p, tally.n, 0.0, s);
        NSDG_HIP(hipStreamSynchronize(s));
        if (out)
            std::copy(r, r + 4, out);
    }
    void synchronize() override
    {
        NSDG_HIP(hipStreamSynchronize(s));
        if (prof.on)
            prof.collect();
    }

    // ---- state blob: header, then hice, cice (DG), sigma (3S), strain (3S), u, v (CG)
    struct Header {
        char magic[8];
        int32_t version, cg, dg, s;
        uint64_t nx, ny;
        double deltaT;
    };
    std::vector<std::pair<double*, size_t>> stateParts()
    {
        std::vector<std::pair<double*, size_t>> parts { { h.p, ne * DG }, { A.p, ne * DG }, { sig.p, 3 * S * ne },
            { eps.p, 3 * S * ne }, { u.p, nn }, { v.p, nn }, { u0.p, nn }, { v0.p, nn } };
        if (brittle) {
            parts.push_back({ damage.p, ne * DG });
            parts.push_back({ avgU.p, nn });
            parts.push_back({ avgV.p, nn });
        }
        return parts;
    }
    std::vector<char> saveState() override
    {
        need();
        size_t total = sizeof(Header);
        for (auto& pr : stateParts())
            total += pr.second * sizeof(double);
        std::vector<char> buf(total);
        Header hd {};
        std::memcpy(hd.magic, "NSDGST01", 8);
        hd.version = brittle ? 2 : 1; // 2: brittle, with damage and averaged velocity
        hd.cg = CG;
        hd.dg = DG;
        hd.s = S;
        hd.nx = mesh.nx;
        hd.ny = mesh.ny;
        hd.deltaT = deltaT;
        std::memcpy(buf.data(), &hd, sizeof hd);
        size_t off = sizeof(Header);
        for (auto& pr : stateParts()) {
            NSDG_HIP(hipMemcpyAsync(buf.data() + off, pr.first, pr.second * sizeof(double), hipMemcpyDeviceToHost, s));
            off += pr.second * sizeof(double);
        }
        NSDG_HIP(hipStreamSynchronize(s));
        return buf;
    }
    void loadState(const char* buf, size_t size) override
    {
        need();
        require(size >= sizeof(Header), "state blob too small");
        Header hd;
        std::memcpy(&hd, buf, sizeof hd);
        require(std::memcmp(hd.magic, "NSDGST01", 8) == 0 && (hd.version == 1 || hd.version == 2),
            "not a state blob of this library");
        require(hd.version == (brittle ? 2 : 1), "state blob of the other rheology");
        require(hd.cg == CG && hd.dg == DG && hd.nx == static_cast<uint64_t>(mesh.nx)
                && hd.ny == static_cast<uint64_t>(mesh.ny),
            "state blob does not match this kernel's degrees or mesh");
        size_t off = sizeof(Header);
        for (auto& pr : stateParts()) {
            require(off + pr.second * sizeof(double) <= size, "state blob truncated");
            NSDG_HIP(hipMemcpyAsync(pr.first, buf + off, pr.second * sizeof(double), hipMemcpyHostToDevice, s));
            off += pr.second * sizeof(double);
        }
        deltaT = hd.deltaT;
        NSDG_HIP(hipStreamSynchronize(s));
    }
};

std::unique_ptr<DynBase> makeDyn(int cg, int dg, const nsdg_params& p)
{
    if (cg == 1 && dg == 3)
        return std::make_unique<DynImpl<1, 3>>(p);
    if (cg == 2 && dg == 6)
        return std::make_unique<DynImpl<2, 6>>(p);
    throw Error(NSDG_ERR_UNSUPPORTED, "supported (cg_degree, dg_adv) pairs: (1, 3) and (2, 6)");
}

} // namespace nsdg
