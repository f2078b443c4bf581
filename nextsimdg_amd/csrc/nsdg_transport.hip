// DG upwind transport on gfx950.
//
// One thread per element gathers everything that changes its coefficients:
// its cell term, the upwind fluxes of its four edges (each interior edge flux is
// evaluated by both neighbours from identical operands, so the pair stays exactly
// conservative without atomics), its Dirichlet outflow terms, the inverse mass
// matrix and the Runge-Kutta stage combination, in one pass over HBM. Periodic
// segments, when present, are applied by a per-entry kernel between the flux pass
// and a finishing pass. Fields are SoA (component-major) on the device.
// Reference: dynamics/src/DGTransport.cpp:158-566, ParametricMap.cpp:13-87,
// dgLimit.hpp:16-84, Interpolations.cpp:74-122.
#include "nsdg_device.hpp"
#include "nsdg_transport.hpp"

#include <cstring>

namespace nsdg {

enum StageMode { ST_RAW = 0, ST_PLUS = 1, ST_PLUS_K = 2, ST_RK2B = 3, ST_RK3B = 4, ST_RK3C = 5 };

struct TrArgs {
    MeshView m;
    int ne, nXe, nYe;
    const double* x;    // operator input (neighbours read it)
    const double* velx;
    const double* vely;
    const double* nvX;
    const double* nvY;
    const double* ops;
    size_t opStride;
    double dt;
    int mode;
    const double* base; // RK3 base field (phi)
    const double* kin;  // RK2 first-stage increment
    double* kout;
    double* out;
    double* raw;        // un-finished increment (periodic path)
    bool withDirichlet;
    double* tally;      // per-element Dirichlet outflow accumulator (mass budget), or null
    double tallyW;      // weight of this stage's outflow in the step (1/R folded in on the sphere)
};

// ---------------------------------------------------------------- operator set-up
// AdvectionCellTermX/Y (ParametricMap.cpp:13-66) and the inverse DG mass matrix
// (:68-87), k-major: cellX[i*Q+q], cellY[DG*Q + i*Q+q], invM[2*DG*Q + i*DG+j].
template <int DG> __global__ void k_transport_ops(MeshView m, int nOps, double* ops)
{
    constexpr int G = gp1d(DG), Q = G * G;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nOps)
        return;
    double c[4][2];
    elementCoords(m, e, c);
    double dxT[2][Q], dyT[2][Q];
    mapGradients<G>(c, dxT, dyT);
    for (int i = 0; i < DG; ++i)
        for (int q = 0; q < Q; ++q) {
            const double w = W<G>::v[0][q];
            const double X = Psi<DG, G>::x[i][q] * (dyT[1][q] * w) - Psi<DG, G>::y[i][q] * (dxT[1][q] * w);
            const double Y = Psi<DG, G>::y[i][q] * (dxT[0][q] * w) - Psi<DG, G>::x[i][q] * (dyT[0][q] * w);
            ops[static_cast<size_t>(i * Q + q) * nOps + e] = X;
            ops[static_cast<size_t>(DG * Q + i * Q + q) * nOps + e] = Y;
        }
    // on the sphere: the cell terms stay the Cartesian ones in (longitude, latitude) and the
    // inverse mass carries cos(latitude) and 1 / R (ParametricMap.cpp:49-80)
    double M[DG][DG], Mi[DG][DG];
    metricMassMatrix<DG>(m, c, M);
    invert<DG>(M, Mi);
    for (int i = 0; i < DG; ++i)
        for (int j = 0; j < DG; ++j)
            ops[static_cast<size_t>(2 * DG * Q + i * DG + j) * nOps + e]
                = m.spherical ? Mi[i][j] / kEarthRadius : Mi[i][j];
}

// Transport operators of a general mesh from its geometry, SoA [k][ne]: the bilinear map
// x(xi, eta) = x0 + X1 xi + X2 eta + X3 xi eta (and y), and the upper triangle of the
// inverse (metric) DG mass matrix (/ R on the sphere). The cell terms factor through the map:
//   sum_q (cellX_i vx + cellY_i vy) phi = sum_q psi_xi,i (y_eta vx - x_eta vy) w phi
//                                            + psi_eta,i (x_xi vy - y_xi vx) w phi,
// so 6 + D (D + 1) / 2 values (27 for D = 6, 42 for D = 8) replace 2 D Q + D^2 (144, 208).
template <int DG> struct TrGeo {
    static constexpr int X1 = 0, X2 = 1, X3 = 2, Y1 = 3, Y2 = 4, Y3 = 5, MI = 6, K = 6 + DG * (DG + 1) / 2;
    __host__ __device__ static constexpr int mi(int i, int j)
    {
        return i <= j ? MI + i * DG - i * (i - 1) / 2 + (j - i) : mi(j, i);
    }
};
template <int DG> __global__ void k_transport_geom(MeshView m, int ne, double* ops)
{
    using L = TrGeo<DG>;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    auto put = [&](int k, double v) { ops[static_cast<size_t>(k) * ne + e] = v; };
    double c[4][2]; // LL, LR, UL, UR
    elementCoords(m, e, c);
    put(L::X1, c[1][0] - c[0][0]);
    put(L::X2, c[2][0] - c[0][0]);
    put(L::X3, (c[0][0] - c[1][0]) - (c[2][0] - c[3][0]));
    put(L::Y1, c[1][1] - c[0][1]);
    put(L::Y2, c[2][1] - c[0][1]);
    put(L::Y3, (c[0][1] - c[1][1]) - (c[2][1] - c[3][1]));
    double M[DG][DG], Mi[DG][DG];
    metricMassMatrix<DG>(m, c, M);
    invert<DG>(M, Mi);
    for (int i = 0; i < DG; ++i)
        for (int j = i; j < DG; ++j)
            put(L::mi(i, j), m.spherical ? Mi[i][j] / kEarthRadius : Mi[i][j]);
}

// ---------------------------------------------------------------- normal velocities
// reinitnormalvelocity (DGTransport.cpp:158-252), one thread per edge: the average of
// the two adjacent ocean cells' traces along the scaled normal, times 2^k where k
// counts the Dirichlet lists naming the edge.
template <int DG> __global__ void k_normal_velocity(MeshView m, const double* vx, const double* vy, int ne,
    double* nvX, double* nvY, int nXe, int nYe)
{
    constexpr int E = edgedofs(DG);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = m.nx, ny = m.ny;
    if (t >= nXe + nYe)
        return;
    double acc[E];
#pragma unroll
    for (int k = 0; k < E; ++k)
        acc[k] = 0.0;
    double cx[DG], cy[DG], tx[E], ty[E];
    if (t < nXe) { // X-edge ie = ix + nx*iy (bottom edge of row iy)
        const int ix = t % nx, iy = t / nx;
        const int n0 = iy * (nx + 1) + ix;
        const double t0 = edgeDx(m, m.vxy[2 * (n0 + 1)] - m.vxy[2 * n0]);
        const double t1 = m.vxy[2 * (n0 + 1) + 1] - m.vxy[2 * n0 + 1];
        // both cells' coefficients are loaded before their ocean flags are tested
        double dx[DG], dy[DG];
        const bool hb = iy > 0, ht = iy < ny;
        if (hb)
#pragma unroll
            for (int i = 0; i < DG; ++i) {
                cx[i] = vx[static_cast<size_t>(i) * ne + (t - nx)];
                cy[i] = vy[static_cast<size_t>(i) * ne + (t - nx)];
            }
        if (ht)
#pragma unroll
            for (int i = 0; i < DG; ++i) {
                dx[i] = vx[static_cast<size_t>(i) * ne + t];
                dy[i] = vy[static_cast<size_t>(i) * ne + t];
            }
        if (hb && m.ocean[t - nx]) {
            trace<DG, 2>(cx, tx);
            trace<DG, 2>(cy, ty);
#pragma unroll
            for (int k = 0; k < E; ++k)
                acc[k] += 0.5 * (-t1 * tx[k] + t0 * ty[k]);
        }
        if (ht && m.ocean[t]) {
#pragma unroll
            for (int i = 0; i < DG; ++i) {
                cx[i] = dx[i];
                cy[i] = dy[i];
            }
            trace<DG, 0>(cx, tx);
            trace<DG, 0>(cy, ty);
#pragma unroll
            for (int k = 0; k < E; ++k)
                acc[k] += 0.5 * (-t1 * tx[k] + t0 * ty[k]);
        }
        const double f = m.factX[t];
#pragma unroll
        for (int k = 0; k < E; ++k)
            nvX[static_cast<size_t>(k) * nXe + t] = acc[k] * f;
    } else { // Y-edge ie = ix + (nx+1)*iy (left edge of column ix)
        const int ie = t - nXe;
        const int ix = ie % (nx + 1), iy = ie / (nx + 1);
        const int n0 = iy * (nx + 1) + ix;
        const double t0 = edgeDx(m, m.vxy[2 * (n0 + nx + 1)] - m.vxy[2 * n0]);
        const double t1 = m.vxy[2 * (n0 + nx + 1) + 1] - m.vxy[2 * n0 + 1];
        const int cl = ix - 1 + nx * iy, cr = ix + nx * iy;
        double dx[DG], dy[DG];
        const bool hl = ix > 0, hr = ix < nx;
        if (hl)
#pragma unroll
            for (int i = 0; i < DG; ++i) {
                cx[i] = vx[static_cast<size_t>(i) * ne + cl];
                cy[i] = vy[static_cast<size_t>(i) * ne + cl];
            }
        if (hr)
#pragma unroll
            for (int i = 0; i < DG; ++i) {
                dx[i] = vx[static_cast<size_t>(i) * ne + cr];
                dy[i] = vy[static_cast<size_t>(i) * ne + cr];
            }
        if (hl && m.ocean[cl]) {
            trace<DG, 1>(cx, tx);
            trace<DG, 1>(cy, ty);
#pragma unroll
            for (int k = 0; k < E; ++k)
                acc[k] += 0.5 * (t1 * tx[k] - t0 * ty[k]);
        }
        if (hr && m.ocean[cr]) {
#pragma unroll
            for (int i = 0; i < DG; ++i) {
                cx[i] = dx[i];
                cy[i] = dy[i];
            }
            trace<DG, 3>(cx, tx);
            trace<DG, 3>(cy, ty);
#pragma unroll
            for (int k = 0; k < E; ++k)
                acc[k] += 0.5 * (t1 * tx[k] - t0 * ty[k]);
        }
        const double f = m.factY[ie];
#pragma unroll
        for (int k = 0; k < E; ++k)
            nvY[static_cast<size_t>(k) * nYe + ie] = acc[k] * f;
    }
}

// ---------------------------------------------------------------- flux helpers
template <int DG> __device__ __forceinline__ void loadCell(const double* f, int ne, int e, double* c)
{
#pragma unroll
    for (int i = 0; i < DG; ++i)
        c[i] = f[static_cast<size_t>(i) * ne + e];
}
template <int E, int G> __device__ __forceinline__ void edgeEval(const double* ev, double* out)
{
#pragma unroll
    for (int q = 0; q < G; ++q) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < E; ++k)
            s += ev[k] * PsiE<E, G>::v[k][q];
        out[q] = s;
    }
}
template <int E> __device__ __forceinline__ void loadEdge(const double* nv, int n, int ie, double* v)
{
#pragma unroll
    for (int k = 0; k < E; ++k)
        v[k] = nv[static_cast<size_t>(k) * n + ie];
}
// up += sgn * sum_q dt * tmp[q] * PSIe_w<side>[q][i]
template <int DG, int SIDE>
__device__ __forceinline__ void edgeAccumulate(double* up, double sgn, double dt, const double* tmp)
{
    constexpr int G = gp1d(DG);
#pragma unroll
    for (int i = 0; i < DG; ++i) {
        double s = 0;
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const double w = SIDE == 0 ? PsiEW<DG, G>::b[q][i]
                : SIDE == 1            ? PsiEW<DG, G>::r[q][i]
                : SIDE == 2            ? PsiEW<DG, G>::t[q][i]
                                       : PsiEW<DG, G>::l[q][i];
            s += dt * tmp[q] * w;
        }
        up[i] += sgn * s;
    }
}
// Upwind flux through an interior edge between c1 (left/bottom, side S1 of c1) and
// c2 (right/top): tmp = max(v,0) phi1 + min(v,0) phi2 (DGTransport.cpp:390-433).
template <int DG, int S1, int S2>
__device__ __forceinline__ void upwind(const double* p1, const double* p2, const double* nv, double* tmp)
{
    constexpr int E = edgedofs(DG), G = gp1d(DG);
    double t1[E], t2[E], g1[G], g2[G], vg[G];
    trace<DG, S1>(p1, t1);
    trace<DG, S2>(p2, t2);
    edgeEval<E, G>(nv, vg);
    edgeEval<E, G>(t1, g1);
    edgeEval<E, G>(t2, g2);
#pragma unroll
    for (int q = 0; q < G; ++q)
        tmp[q] = fmax(vg[q], 0.0) * g1[q] + fmin(vg[q], 0.0) * g2[q];
}
// Dirichlet outflow-only boundary term (DGTransport.cpp:306-351), applied count times.
// Returns the outflow: count times the flux tested with the constant basis function, the
// amount coefficient 0 loses before the inverse mass.
template <int DG, int SIDE>
__device__ __forceinline__ double boundaryTerm(double* up, const double* ph, const double* nv, double dt, int count)
{
    constexpr int E = edgedofs(DG), G = gp1d(DG);
    double tr[E], pg[G], vg[G], tmp[G];
    trace<DG, SIDE>(ph, tr);
    edgeEval<E, G>(tr, pg);
    edgeEval<E, G>(nv, vg);
    constexpr double sgn = (SIDE == 0 || SIDE == 3) ? -1.0 : 1.0;
#pragma unroll
    for (int q = 0; q < G; ++q)
        tmp[q] = pg[q] * fmax(sgn * vg[q], 0.0);
    for (int r = 0; r < count; ++r)
        edgeAccumulate<DG, SIDE>(up, -1.0, dt, tmp);
    double o = 0;
#pragma unroll
    for (int q = 0; q < G; ++q) {
        const double w = SIDE == 0 ? PsiEW<DG, G>::b[q][0]
            : SIDE == 1            ? PsiEW<DG, G>::r[q][0]
            : SIDE == 2            ? PsiEW<DG, G>::t[q][0]
                                   : PsiEW<DG, G>::l[q][0];
        o += dt * tmp[q] * w;
    }
    return count * o;
}
// All Dirichlet terms of element e; with a.tally set, the stage-weighted outflow is added
// to the element's budget accumulator (elements without Dirichlet sides add nothing).
template <int DG> __device__ __forceinline__ void dirichletTerms(const TrArgs& a, int e, const double* ph, double* up)
{
    constexpr int E = edgedofs(DG);
    const uint32_t fl = a.m.eflags[e];
    if (!fl)
        return;
    const int nx = a.m.nx, ix = e % nx, iy = e / nx;
    double nv[E], out = 0;
    if (fl & 0xffu) {
        loadEdge<E>(a.nvX, a.nXe, nx * iy + ix, nv);
        out += boundaryTerm<DG, 0>(up, ph, nv, a.dt, fl & 0xffu);
    }
    if (fl & 0xff00u) {
        loadEdge<E>(a.nvY, a.nYe, (nx + 1) * iy + ix + 1, nv);
        out += boundaryTerm<DG, 1>(up, ph, nv, a.dt, (fl >> 8) & 0xffu);
    }
    if (fl & 0xff0000u) {
        loadEdge<E>(a.nvX, a.nXe, nx * (iy + 1) + ix, nv);
        out += boundaryTerm<DG, 2>(up, ph, nv, a.dt, (fl >> 16) & 0xffu);
    }
    if (fl & 0xff000000u) {
        loadEdge<E>(a.nvY, a.nYe, (nx + 1) * iy + ix, nv);
        out += boundaryTerm<DG, 3>(up, ph, nv, a.dt, fl >> 24);
    }
    if (a.tally)
        a.tally[e] += a.tallyW * out;
}
// Inverse mass matrix and Runge-Kutta stage combination (DGTransport.cpp:509-552).
template <int DG, bool UNI> __device__ __forceinline__ void finishCell(const TrArgs& a, int e, const double* ph, const double* up)
{
    constexpr int Q = gp1d(DG) * gp1d(DG);
    const Ops<UNI> op { a.ops, a.opStride };
    const size_t ne = a.ne;
#pragma unroll
    for (int i = 0; i < DG; ++i) {
        double k = 0;
#pragma unroll
        for (int j = 0; j < DG; ++j) {
            if constexpr (UNI)
                k += op(2 * DG * Q + i * DG + j, e) * up[j];
            else
                k += a.ops[static_cast<size_t>(TrGeo<DG>::mi(i, j)) * ne + e] * up[j];
        }
        const size_t o = i * ne + e;
        double r;
        switch (a.mode) {
        case ST_RAW: r = k; break;
        case ST_PLUS: r = ph[i] + k; break;
        case ST_PLUS_K:
            a.kout[o] = k;
            r = ph[i] + k;
            break;
        case ST_RK2B: r = ph[i] + 0.5 * (k - a.kin[o]); break;
        case ST_RK3B: r = (k + ph[i]) * 0.25 + 0.75 * a.base[o]; break;
        default: r = a.base[o] * (1.0 / 3.0) + 2.0 / 3.0 * (k + ph[i]); break;
        }
        a.out[o] = r;
    }
}

// ---------------------------------------------------------------- DGTransportOperator
// FINISH: apply Dirichlet, inverse mass and the stage here (no periodic segments);
// otherwise leave the raw increment for the periodic and finishing passes.
// One element per thread; the runs of kBlock elements go to the XCDs in contiguous chunks
// (xcdTile), so the rows above and below a run, which its edge fluxes read, were loaded
// into the same XCD's L2 by the neighbouring runs.
template <int DG, bool UNI, bool FINISH> __global__ __launch_bounds__(kBlock) void k_transport_op(TrArgs a)
{
    constexpr int G = gp1d(DG), Q = G * G, E = edgedofs(DG);
    const int run = xcdTile(blockIdx.x, (a.ne + kBlock - 1) / kBlock);
    if (run < 0)
        return;
    const int e = run * kBlock + threadIdx.x;
    if (e >= a.ne)
        return;
    const int nx = a.m.nx, ny = a.m.ny, ix = e % nx, iy = e / nx, ne = a.ne;
    const Ops<UNI> op { a.ops, a.opStride };
    double ph[DG], up[DG];
    loadCell<DG>(a.x, ne, e, ph);
#pragma unroll
    for (int i = 0; i < DG; ++i)
        up[i] = 0.0;
    const bool oc = a.m.ocean[e];
    if constexpr (DG > 1) {
        if (oc) { // cell_term (DGTransport.cpp:278-303)
            double vx[DG], vy[DG], vxg[Q], vyg[Q], pg[Q];
            loadCell<DG>(a.velx, ne, e, vx);
            loadCell<DG>(a.vely, ne, e, vy);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
                for (int i = 0; i < DG; ++i) {
                    s0 += vx[i] * Psi<DG, G>::v[i][q];
                    s1 += vy[i] * Psi<DG, G>::v[i][q];
                    s2 += ph[i] * Psi<DG, G>::v[i][q];
                }
                vxg[q] = s0;
                vyg[q] = s1;
                pg[q] = s2;
            }
            if constexpr (UNI) {
#pragma unroll
                for (int i = 0; i < DG; ++i) {
                    double s = 0;
#pragma unroll
                    for (int q = 0; q < Q; ++q)
                        s += a.dt * (op(i * Q + q, e) * vxg[q] + op(DG * Q + i * Q + q, e) * vyg[q]) * pg[q];
                    up[i] += s;
                }
            } else {
                using L = TrGeo<DG>;
                const size_t sne = static_cast<size_t>(ne);
                const double X1 = a.ops[L::X1 * sne + e], X2 = a.ops[L::X2 * sne + e], X3 = a.ops[L::X3 * sne + e];
                const double Y1 = a.ops[L::Y1 * sne + e], Y2 = a.ops[L::Y2 * sne + e], Y3 = a.ops[L::Y3 * sne + e];
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const double xi = gauss_point(G, q % G), eta = gauss_point(G, q / G);
                    const double w = gauss_weight(G, q % G) * gauss_weight(G, q / G) * a.dt;
                    const double xxi = X1 + X3 * eta, xeta = X2 + X3 * xi, yxi = Y1 + Y3 * eta, yeta = Y2 + Y3 * xi;
                    const double fxi = (yeta * vxg[q] - xeta * vyg[q]) * (w * pg[q]);
                    const double feta = (xxi * vyg[q] - yxi * vxg[q]) * (w * pg[q]);
#pragma unroll
                    for (int i = 0; i < DG; ++i)
                        up[i] += Psi<DG, G>::x[i][q] * fxi + Psi<DG, G>::y[i][q] * feta;
                }
            }
        }
    }
    if (oc) {
        double nb[DG], nv[E], tmp[G];
        if (ix > 0 && a.m.ocean[e - 1]) { // left Y-edge: this cell is c2
            loadCell<DG>(a.x, ne, e - 1, nb);
            loadEdge<E>(a.nvY, a.nYe, iy * (nx + 1) + ix, nv);
            upwind<DG, 1, 3>(nb, ph, nv, tmp);
            edgeAccumulate<DG, 3>(up, 1.0, a.dt, tmp);
        }
        if (ix < nx - 1 && a.m.ocean[e + 1]) { // right Y-edge: this cell is c1
            loadCell<DG>(a.x, ne, e + 1, nb);
            loadEdge<E>(a.nvY, a.nYe, iy * (nx + 1) + ix + 1, nv);
            upwind<DG, 1, 3>(ph, nb, nv, tmp);
            edgeAccumulate<DG, 1>(up, -1.0, a.dt, tmp);
        }
        if (iy > 0 && a.m.ocean[e - nx]) { // bottom X-edge: this cell is c2
            loadCell<DG>(a.x, ne, e - nx, nb);
            loadEdge<E>(a.nvX, a.nXe, ix + nx * iy, nv);
            upwind<DG, 2, 0>(nb, ph, nv, tmp);
            edgeAccumulate<DG, 0>(up, 1.0, a.dt, tmp);
        }
        if (iy < ny - 1 && a.m.ocean[e + nx]) { // top X-edge: this cell is c1
            loadCell<DG>(a.x, ne, e + nx, nb);
            loadEdge<E>(a.nvX, a.nXe, ix + nx * (iy + 1), nv);
            upwind<DG, 2, 0>(ph, nb, nv, tmp);
            edgeAccumulate<DG, 2>(up, -1.0, a.dt, tmp);
        }
    }
    if constexpr (FINISH) {
        dirichletTerms<DG>(a, e, ph, up);
        finishCell<DG, UNI>(a, e, ph, up);
    } else {
#pragma unroll
        for (int i = 0; i < DG; ++i)
            a.raw[static_cast<size_t>(i) * ne + e] = up[i];
    }
}

// Periodic pairs of one segment (DGTransport.cpp:465-481): one thread per entry.
template <int DG> __global__ void k_transport_periodic(TrArgs a, const int32_t* entries, int n)
{
    constexpr int E = edgedofs(DG), G = gp1d(DG);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const int type = entries[4 * k], c1 = entries[4 * k + 1], c2 = entries[4 * k + 2], ie = entries[4 * k + 3];
    if (!a.m.ocean[c1] || !a.m.ocean[c2])
        return;
    double p1[DG], p2[DG], nv[E], tmp[G], u1[DG], u2[DG];
    loadCell<DG>(a.x, a.ne, c1, p1);
    loadCell<DG>(a.x, a.ne, c2, p2);
    loadCell<DG>(a.raw, a.ne, c1, u1);
    loadCell<DG>(a.raw, a.ne, c2, u2);
    if (type == 0) {
        loadEdge<E>(a.nvX, a.nXe, ie, nv);
        upwind<DG, 2, 0>(p1, p2, nv, tmp);
        edgeAccumulate<DG, 2>(u1, -1.0, a.dt, tmp);
        edgeAccumulate<DG, 0>(u2, 1.0, a.dt, tmp);
    } else {
        loadEdge<E>(a.nvY, a.nYe, ie, nv);
        upwind<DG, 1, 3>(p1, p2, nv, tmp);
        edgeAccumulate<DG, 1>(u1, -1.0, a.dt, tmp);
        edgeAccumulate<DG, 3>(u2, 1.0, a.dt, tmp);
    }
#pragma unroll
    for (int i = 0; i < DG; ++i) {
        a.raw[static_cast<size_t>(i) * a.ne + c1] = u1[i];
        a.raw[static_cast<size_t>(i) * a.ne + c2] = u2[i];
    }
}

template <int DG, bool UNI> __global__ __launch_bounds__(kBlock) void k_transport_finish(TrArgs a)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.ne)
        return;
    double ph[DG], up[DG];
    loadCell<DG>(a.x, a.ne, e, ph);
    loadCell<DG>(a.raw, a.ne, e, up);
    dirichletTerms<DG>(a, e, ph, up);
    finishCell<DG, UNI>(a, e, ph, up);
}

// ---------------------------------------------------------------- limiters
// LimitMax then LimitMin (dgLimit.hpp:16-84) on one SoA field. With tally set, the change
// of the element's Tools::MeanValue contribution is added to tally[e] (mass budget).
template <int DG>
__global__ void k_limit(MeshView m, double* f, int ne, bool doMax, double maxv, bool doMin, double minv,
    double* tally)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    double c[DG], c0[DG];
    loadCell<DG>(f, ne, e, c);
#pragma unroll
    for (int i = 0; i < DG; ++i)
        c0[i] = c[i];
    for (int pass = 0; pass < 2; ++pass) {
        const bool isMax = pass == 0;
        if (isMax ? !doMax : !doMin)
            continue;
        const double b = isMax ? maxv : minv;
        c[0] = isMax ? fmin(b, c[0]) : fmax(b, c[0]);
        if constexpr (DG == 3) {
            const double l0 = 2.0 * fmax(fabs(c[1] + c[2]), fabs(c[1] - c[2]));
            if (l0 != 0.0) {
                const bool ex = isMax ? (c[0] + l0 - b > 0) : (c[0] - l0 - b < 0);
                if (ex) {
                    const double s = isMax ? (b - c[0]) / l0 : (c[0] - b) / l0;
                    c[1] *= s;
                    c[2] *= s;
                }
            }
        } else if constexpr (DG == 6) {
            double ext = isMax ? -INFINITY : INFINITY;
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                double s = 0;
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    s += c[k] * Psi<6, 3>::v[k][q];
                ext = isMax ? fmax(ext, s) : fmin(ext, s);
            }
            const double l0 = ext - c[0];
            if (isMax ? (ext > b) : (ext < b)) {
                const double s = isMax ? (b - c[0]) / l0 : (c[0] - b) / l0;
#pragma unroll
                for (int k = 1; k < 6; ++k)
                    c[k] *= s;
            }
        }
    }
    // a cell the limits leave bit for bit as it was is not written back (most cells)
    bool moved = false;
#pragma unroll
    for (int i = 0; i < DG; ++i)
        moved = moved || __double_as_longlong(c[i]) != __double_as_longlong(c0[i]);
    if (!moved)
        return;
#pragma unroll
    for (int i = 0; i < DG; ++i)
        f[static_cast<size_t>(i) * ne + e] = c[i];
    if (tally) {
        bool changed = false;
#pragma unroll
        for (int i = 0; i < DG; ++i)
            changed = changed || c[i] != c0[i];
        if (changed)
            tally[e] += elementMean<DG>(m, e, c) - elementMean<DG>(m, e, c0);
    }
}

// ---------------------------------------------------------------- CG -> DG
// Composite per-element projection op[i*N + j] = (M^-1 PSI diag(J w c) PHI^T)[i][j], c the
// metric weight (cos(latitude) on the sphere, Interpolations.cpp:99-108).
template <int CG, int DG> __global__ void k_cg2dg_ops(MeshView m, int nOps, double* ops)
{
    constexpr int G = gp1d(DG), Q = G * G, N = cgdofs(CG);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nOps)
        return;
    double c[4][2], J[Q], cl[Q], M[DG][DG], Mi[DG][DG];
    elementCoords(m, e, c);
    jacobian<G>(c, J);
    metricWeight<G>(m, c, cl);
    metricMassMatrix<DG>(m, c, M);
    invert<DG>(M, Mi);
    for (int i = 0; i < DG; ++i)
        for (int j = 0; j < N; ++j) {
            double s = 0;
            for (int t = 0; t < DG; ++t) {
                double r = 0;
                for (int q = 0; q < Q; ++q)
                    r += Psi<DG, G>::v[t][q] * (J[q] * W<G>::v[0][q] * cl[q]) * Phi<CG, G>::v[j][q];
                s += Mi[i][t] * r;
            }
            ops[static_cast<size_t>(i * N + j) * nOps + e] = s;
        }
}
template <int CG, int DG, bool UNI>
__global__ __launch_bounds__(kBlock) void k_cg2dg(int nx, int ne, const double* ops, size_t stride, const double* s1,
    double* d1, const double* s2, double* d2)
{
    constexpr int N = cgdofs(CG);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    const Ops<UNI> op { ops, stride };
    const int ix = e % nx, iy = e / nx, row = CG * nx + 1;
    const int cgi = CG * row * iy + CG * ix;
    double l1[N], l2[N];
#pragma unroll
    for (int r = 0; r <= CG; ++r)
#pragma unroll
        for (int cc = 0; cc <= CG; ++cc) {
            l1[cc + (CG + 1) * r] = s1[cgi + cc + r * row];
            if (s2)
                l2[cc + (CG + 1) * r] = s2[cgi + cc + r * row];
        }
#pragma unroll
    for (int i = 0; i < DG; ++i) {
        double a = 0, b = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const double w = op(i * N + j, e);
            a += w * l1[j];
            if (s2)
                b += w * l2[j];
        }
        d1[static_cast<size_t>(i) * ne + e] = a;
        if (s2)
            d2[static_cast<size_t>(i) * ne + e] = b;
    }
}

// ---------------------------------------------------------------- layout helpers
__global__ void k_aos_to_soa(const double* in, size_t n, int nc, double* out)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * nc)
        return;
    const size_t e = t / nc, c = t % nc;
    out[c * n + e] = in[t];
}
__global__ void k_soa_to_aos(const double* in, size_t n, int nc, double* out)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * nc)
        return;
    const size_t e = t / nc, c = t % nc;
    out[t] = in[c * n + e];
}
__global__ void k_fill(double* p, size_t n, double v)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n)
        p[t] = v;
}

void uploadSoA(DBuf<double>& scratch, const double* host, size_t n, int nc, double* dst, hipStream_t s)
{
    if (nc == 1) {
        NSDG_HIP(hipMemcpyAsync(dst, host, n * sizeof(double), hipMemcpyHostToDevice, s));
        return;
    }
    if (scratch.n < n * nc)
        scratch.alloc(n * nc);
    scratch.upload(host, n * nc, s);
    k_aos_to_soa<<<blocksFor(n * nc), kBlock, 0, s>>>(scratch.p, n, nc, dst);
    checkLaunch("aos_to_soa");
}
void downloadSoA(DBuf<double>& scratch, const double* src, size_t n, int nc, double* host, hipStream_t s)
{
    if (nc == 1) {
        NSDG_HIP(hipMemcpyAsync(host, src, n * sizeof(double), hipMemcpyDeviceToHost, s));
        NSDG_HIP(hipStreamSynchronize(s));
        return;
    }
    if (scratch.n < n * nc)
        scratch.alloc(n * nc);
    k_soa_to_aos<<<blocksFor(n * nc), kBlock, 0, s>>>(src, n, nc, scratch.p);
    checkLaunch("soa_to_aos");
    scratch.download(host, n * nc, s);
    NSDG_HIP(hipStreamSynchronize(s));
}
void deviceFill(double* p, size_t n, double v, hipStream_t s)
{
    if (!n)
        return;
    k_fill<<<blocksFor(n), kBlock, 0, s>>>(p, n, v);
    checkLaunch("fill");
}
// Device-to-device copy as a plain kernel of 16-byte accesses (the runtime's blit kernel
// moves these field-sized copies at about half the rate).
__global__ void k_copy(double* dst, const double* src, size_t n)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (2 * t + 1 < n)
        reinterpret_cast<double2*>(dst)[t] = reinterpret_cast<const double2*>(src)[t];
    else if (2 * t < n)
        dst[2 * t] = src[2 * t];
}
void deviceCopy(double* dst, const double* src, size_t n, hipStream_t s)
{
    if (!n)
        return;
    if (((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u) != 0) {
        NSDG_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        return;
    }
    k_copy<<<blocksFor((n + 1) / 2), kBlock, 0, s>>>(dst, src, n);
    checkLaunch("copy");
}

void CG2DGOps::build(int cg_, int dg_, const DeviceMesh& mesh, hipStream_t s)
{
    cg = cg_;
    dg = dg_;
    const int nOps = mesh.uniform ? 1 : static_cast<int>(mesh.ne);
    const int N = cgdofs(cg);
    ops.alloc(static_cast<size_t>(dg) * N * nOps);
    const MeshView mv = view(mesh);
    auto go = [&](auto kern) {
        kern<<<blocksFor(nOps), kBlock, 0, s>>>(mv, nOps, ops.p);
        checkLaunch("cg2dg_ops");
    };
    if (cg == 1 && dg == 1) go(k_cg2dg_ops<1, 1>);
    else if (cg == 1 && dg == 3) go(k_cg2dg_ops<1, 3>);
    else if (cg == 1 && dg == 6) go(k_cg2dg_ops<1, 6>);
    else if (cg == 2 && dg == 1) go(k_cg2dg_ops<2, 1>);
    else if (cg == 2 && dg == 3) go(k_cg2dg_ops<2, 3>);
    else if (cg == 2 && dg == 6) go(k_cg2dg_ops<2, 6>);
    else if (cg == 2 && dg == 8) go(k_cg2dg_ops<2, 8>);
    else throw Error(NSDG_ERR_UNSUPPORTED, "unsupported (cg, dg) for CG2DG");
}

void CG2DGOps::apply(const DeviceMesh& mesh, const double* src, double* dst, const double* src2, double* dst2,
    hipStream_t s, Profiler& p) const
{
    Timed t(p, s, "cg2dg");
    const int ne = static_cast<int>(mesh.ne);
    const size_t stride = mesh.uniform ? 1 : mesh.ne;
    auto go = [&](auto kern) {
        kern<<<blocksFor(ne), kBlock, 0, s>>>(mesh.nx, ne, ops.p, stride, src, dst, src2, dst2);
        checkLaunch("cg2dg");
    };
#define NSDG_CG2DG_CASE(C, D)                        \
    if (cg == C && dg == D) {                        \
        if (mesh.uniform)                            \
            go(k_cg2dg<C, D, true>);                 \
        else                                         \
            go(k_cg2dg<C, D, false>);                \
        return;                                      \
    }
    NSDG_CG2DG_CASE(1, 1)
    NSDG_CG2DG_CASE(1, 3)
    NSDG_CG2DG_CASE(1, 6)
    NSDG_CG2DG_CASE(2, 1)
    NSDG_CG2DG_CASE(2, 3)
    NSDG_CG2DG_CASE(2, 6)
    NSDG_CG2DG_CASE(2, 8)
#undef NSDG_CG2DG_CASE
    throw Error(NSDG_ERR_UNSUPPORTED, "unsupported (cg, dg) for CG2DG");
}

// ---------------------------------------------------------------- transport object
template <int DG> struct TransportImpl final : TransportBase {
    static constexpr int G = gp1d(DG), Q = G * G, E = edgedofs(DG);
    DeviceMesh& mesh;
    hipStream_t s;
    Profiler& prof;
    int scheme = 2;
    int nXe, nYe;
    DBuf<double> ops, vx, vy, nvX, nvY, tmpA, tmpB, kbuf, raw, scratch;
    CG2DGOps cg2dg[3];

    TransportImpl(DeviceMesh& m, hipStream_t st, Profiler& p)
        : mesh(m)
        , s(st)
        , prof(p)
    {
        const size_t ne = mesh.ne;
        nXe = mesh.nx * (mesh.ny + 1);
        nYe = (mesh.nx + 1) * mesh.ny;
        if (mesh.uniform) {
            ops.alloc(static_cast<size_t>(2 * DG * Q + DG * DG));
            k_transport_ops<DG><<<1, kBlock, 0, s>>>(view(mesh), 1, ops.p);
        } else {
            ops.alloc(static_cast<size_t>(TrGeo<DG>::K) * ne);
            k_transport_geom<DG><<<blocksFor(ne), kBlock, 0, s>>>(view(mesh), static_cast<int>(ne), ops.p);
        }
        checkLaunch("transport_ops");
        for (auto* b : { &vx, &vy, &tmpA, &tmpB, &kbuf, &raw })
            b->alloc(ne * DG);
        nvX.alloc(static_cast<size_t>(E) * nXe);
        nvY.alloc(static_cast<size_t>(E) * nYe);
        deviceFill(vx.p, ne * DG, 0.0, s);
        deviceFill(vy.p, ne * DG, 0.0, s);
        deviceFill(nvX.p, nvX.n, 0.0, s);
        deviceFill(nvY.p, nvY.n, 0.0, s);
    }
    int dg() const override { return DG; }
    void setScheme(int rk) override
    {
        require(rk >= 1 && rk <= 3, "time stepping scheme must be rk1, rk2 or rk3");
        scheme = rk;
    }
    double* velx() override { return vx.p; }
    double* vely() override { return vy.p; }
    void setVelocityHost(const double* hx, const double* hy) override
    {
        uploadSoA(scratch, hx, mesh.ne, DG, vx.p, s);
        uploadSoA(scratch, hy, mesh.ne, DG, vy.p, s);
        NSDG_HIP(hipStreamSynchronize(s));
    }
    void prepareAdvectionHost(int cg, const double* u, const double* v) override
    {
        require(cg == 1 || cg == 2, "cg degree must be 1 or 2");
        if (cg2dg[cg].cg != cg)
            cg2dg[cg].build(cg, DG, mesh, s);
        const size_t ncg = static_cast<size_t>(cg * mesh.nx + 1) * (cg * mesh.ny + 1);
        DBuf<double> du, dv;
        du.alloc(ncg);
        dv.alloc(ncg);
        du.upload(u, ncg, s);
        dv.upload(v, ncg, s);
        cg2dg[cg].apply(mesh, du.p, vx.p, dv.p, vy.p, s, prof);
        reinitNormalVelocity();
        NSDG_HIP(hipStreamSynchronize(s));
    }
    void reinitNormalVelocity() override
    {
        Timed t(prof, s, "normal_velocity");
        k_normal_velocity<DG><<<blocksFor(nXe + nYe), kBlock, 0, s>>>(
            view(mesh), vx.p, vy.p, static_cast<int>(mesh.ne), nvX.p, nvY.p, nXe, nYe);
        checkLaunch("normal_velocity");
    }
    TrArgs args(double dt, const double* x) const
    {
        TrArgs a {};
        a.m = view(mesh);
        a.ne = static_cast<int>(mesh.ne);
        a.nXe = nXe;
        a.nYe = nYe;
        a.x = x;
        a.velx = vx.p;
        a.vely = vy.p;
        a.nvX = nvX.p;
        a.nvY = nvY.p;
        a.ops = ops.p;
        a.opStride = mesh.uniform ? 1 : mesh.ne;
        a.dt = dt;
        return a;
    }
    // One DGTransportOperator application fused with its stage combination.
    void stage(TrArgs a)
    {
        const unsigned nb = xcdGrid(static_cast<int>(blocksFor(mesh.ne)));
        if (!mesh.hasPeriodic) {
            Timed t(prof, s, "transport_op");
            if (mesh.uniform)
                k_transport_op<DG, true, true><<<nb, kBlock, 0, s>>>(a);
            else
                k_transport_op<DG, false, true><<<nb, kBlock, 0, s>>>(a);
            checkLaunch("transport_op");
            return;
        }
        a.raw = raw.p;
        {
            Timed t(prof, s, "transport_op");
            if (mesh.uniform)
                k_transport_op<DG, true, false><<<nb, kBlock, 0, s>>>(a);
            else
                k_transport_op<DG, false, false><<<nb, kBlock, 0, s>>>(a);
            checkLaunch("transport_op");
        }
        for (const auto& seg : mesh.segments) {
            k_transport_periodic<DG><<<blocksFor(seg.second), kBlock, 0, s>>>(
                a, mesh.periodic.p + 4 * seg.first, static_cast<int>(seg.second));
            checkLaunch("transport_periodic");
        }
        if (mesh.uniform)
            k_transport_finish<DG, true><<<blocksFor(mesh.ne), kBlock, 0, s>>>(a);
        else
            k_transport_finish<DG, false><<<blocksFor(mesh.ne), kBlock, 0, s>>>(a);
        checkLaunch("transport_finish");
    }
    // step_rk1 / step_rk2 / step_rk3 (DGTransport.cpp:514-552) on a device field. With
    // tally set, each element's Dirichlet outflow over the step (in MeanValue units) is
    // added to tally[e]: the stage outflows weighted as the stages enter the new field.
    void step(double dt, double* phi, double* tally) override
    {
        const size_t n = mesh.ne * DG;
        // on the sphere the inverse mass carries 1/R (ParametricMap.cpp:68-87)
        const double unit = mesh.spherical ? 1.0 / kEarthRadius : 1.0;
        auto withTally = [&](TrArgs a, double w) {
            a.tally = tally;
            a.tallyW = w * unit;
            return a;
        };
        if (scheme == 1) {
            TrArgs a = withTally(args(dt, phi), 1.0);
            a.mode = ST_PLUS;
            a.out = tmpA.p;
            stage(a);
            deviceCopy(phi, tmpA.p, n, s);
        } else if (scheme == 2) { // phi + (k1 + k2) / 2
            TrArgs a = withTally(args(dt, phi), 0.5);
            a.mode = ST_PLUS_K;
            a.kout = kbuf.p;
            a.out = tmpA.p;
            stage(a);
            TrArgs b = withTally(args(dt, tmpA.p), 0.5);
            b.mode = ST_RK2B;
            b.kin = kbuf.p;
            b.out = phi;
            stage(b);
        } else { // phi + (k1 + k2) / 6 + 2 k3 / 3
            TrArgs a = withTally(args(dt, phi), 1.0 / 6.0);
            a.mode = ST_PLUS;
            a.out = tmpA.p;
            stage(a);
            TrArgs b = withTally(args(dt, tmpA.p), 1.0 / 6.0);
            b.mode = ST_RK3B;
            b.base = phi;
            b.out = tmpB.p;
            stage(b);
            TrArgs c = withTally(args(dt, tmpB.p), 2.0 / 3.0);
            c.mode = ST_RK3C;
            c.base = phi;
            c.out = phi;
            stage(c);
        }
    }
    void limit(double* phi, bool doMax, double maxv, bool doMin, double minv, double* tally) override
    {
        Timed t(prof, s, "limit");
        k_limit<DG><<<blocksFor(mesh.ne), kBlock, 0, s>>>(
            view(mesh), phi, static_cast<int>(mesh.ne), doMax, maxv, doMin, minv, tally);
        checkLaunch("limit");
    }
};

std::unique_ptr<TransportBase> makeTransport(int dg, DeviceMesh& mesh, hipStream_t s, Profiler& p)
{
    switch (dg) {
    case 1: return std::make_unique<TransportImpl<1>>(mesh, s, p);
    case 3: return std::make_unique<TransportImpl<3>>(mesh, s, p);
    case 6: return std::make_unique<TransportImpl<6>>(mesh, s, p);
    case 8: return std::make_unique<TransportImpl<8>>(mesh, s, p);
    default: throw Error(NSDG_ERR_UNSUPPORTED, "DG transport degree must be 1, 3, 6 or 8");
    }
}

TransportHandle::~TransportHandle()
{
    t.reset();
    if (stream)
        (void)hipStreamDestroy(stream);
}

} // namespace nsdg
