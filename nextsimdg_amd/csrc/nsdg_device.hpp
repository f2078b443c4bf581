// Device helpers shared by the transport and mEVP kernels: element geometry from
// the vertices, small dense inversion, DG edge traces, operator accessors and the
// CG-node -> adjacent-element enumeration used by every atomic-free gather.
#pragma once
#include "nsdg_internal.hpp"

namespace nsdg {

// ---------------------------------------------------------------- operators
// Per-element operator k of element e, stored k-major (coalesced across elements).
// With UNI the mesh is uniform and one copy serves every element: the address is
// wave-uniform and compile-time, so the load becomes a scalar (constant-cache) load.
template <bool UNI> struct Ops {
    const double* p;
    size_t stride;
    __device__ __forceinline__ double operator()(int k, int e) const
    {
        if constexpr (UNI)
            return p[k];
        else
            return p[static_cast<size_t>(k) * stride + e];
    }
};

// ---------------------------------------------------------------- geometry
// Vertices of element e: LL, LR, UL, UR (ParametricMesh.hpp:218-232).
__device__ __forceinline__ void elementCoords(const MeshView& m, int e, double c[4][2])
{
    const int ix = e % m.nx, iy = e / m.nx;
    const int nid = iy * (m.nx + 1) + ix;
    const int ids[4] = { nid, nid + 1, nid + m.nx + 1, nid + m.nx + 2 };
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        c[k][0] = m.vxy[2 * ids[k]];
        c[k][1] = m.vxy[2 * ids[k] + 1];
    }
}
// Gradient of the bilinear map at the G x G Gauss points: dxT[r][q] = d x_r / d xi,
// dyT[r][q] = d x_r / d eta (ParametricTools.hpp:73-86).
template <int G> __device__ void mapGradients(const double c[4][2], double dxT[2][G * G], double dyT[2][G * G])
{
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int q = 0; q < G * G; ++q) {
            double sx = 0, sy = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                sx += c[k][r] * Phi<1, G>::x[k][q];
                sy += c[k][r] * Phi<1, G>::y[k][q];
            }
            dxT[r][q] = sx;
            dyT[r][q] = sy;
        }
}
template <int G> __device__ void jacobian(const double c[4][2], double J[G * G])
{
    double dxT[2][G * G], dyT[2][G * G];
    mapGradients<G>(c, dxT, dyT);
#pragma unroll
    for (int q = 0; q < G * G; ++q)
        J[q] = dxT[0][q] * dyT[1][q] - dxT[1][q] * dyT[0][q];
}
// Element mass matrix of the D-dof DG space (ParametricTools.hpp:109-153).
template <int D> __device__ void massMatrix(const double c[4][2], double M[D][D])
{
    constexpr int G = gp1d(D), Q = G * G;
    double J[Q];
    jacobian<G>(c, J);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            double s = 0;
            for (int q = 0; q < Q; ++q)
                s += Psi<D, G>::v[i][q] * (W<G>::v[0][q] * J[q]) * Psi<D, G>::v[j][q];
            M[i][j] = s;
        }
}
// In-register Gauss-Jordan inverse with partial pivoting (set-up kernels only).
template <int N> __device__ void invert(double A[N][N], double R[N][N])
{
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            R[i][j] = (i == j) ? 1.0 : 0.0;
    for (int c = 0; c < N; ++c) {
        int p = c;
        for (int r = c + 1; r < N; ++r)
            if (fabs(A[r][c]) > fabs(A[p][c]))
                p = r;
        if (p != c)
            for (int j = 0; j < N; ++j) {
                double t = A[p][j];
                A[p][j] = A[c][j];
                A[c][j] = t;
                t = R[p][j];
                R[p][j] = R[c][j];
                R[c][j] = t;
            }
        const double inv = 1.0 / A[c][c];
        for (int j = 0; j < N; ++j) {
            A[c][j] *= inv;
            R[c][j] *= inv;
        }
        for (int r = 0; r < N; ++r) {
            if (r == c)
                continue;
            const double f = A[r][c];
            if (f == 0.0)
                continue;
            for (int j = 0; j < N; ++j) {
                A[r][j] -= f * A[c][j];
                R[r][j] -= f * R[c][j];
            }
        }
    }
}

// ---------------------------------------------------------------- DG edge traces
// Restriction of the cell basis to one side, in the edge basis 1, t-1/2, (t-1/2)^2-1/12
// (the exact restriction the reference hand-writes in DGTransport.cpp:31-142).
// Side: 0 bottom, 1 right, 2 top, 3 left.
template <int DG, int SIDE> __device__ __forceinline__ void trace(const double* c, double* o)
{
    constexpr double s = (SIDE == 0 || SIDE == 3) ? -0.5 : 0.5; // value of x~ (left/right) or y~
    if constexpr (DG == 1) {
        o[0] = c[0];
    } else if constexpr (SIDE == 1 || SIDE == 3) { // x~ = s, t = y
        if constexpr (DG == 3) {
            o[0] = c[0] + s * c[1];
            o[1] = c[2];
        } else if constexpr (DG == 6) {
            o[0] = c[0] + s * c[1] + 1. / 6. * c[3];
            o[1] = c[2] + s * c[5];
            o[2] = c[4];
        } else {
            o[0] = c[0] + s * c[1] + 1. / 6. * c[3];
            o[1] = c[2] + s * c[5] + 1. / 6. * c[6];
            o[2] = c[4] + s * c[7];
        }
    } else { // y~ = s, t = x
        if constexpr (DG == 3) {
            o[0] = c[0] + s * c[2];
            o[1] = c[1];
        } else if constexpr (DG == 6) {
            o[0] = c[0] + s * c[2] + 1. / 6. * c[4];
            o[1] = c[1] + s * c[5];
            o[2] = c[3];
        } else {
            o[0] = c[0] + s * c[2] + 1. / 6. * c[4];
            o[1] = c[1] + s * c[5] + 1. / 6. * c[7];
            o[2] = c[3] + s * c[6];
        }
    }
}

// ---------------------------------------------------------------- node adjacency
// Elements sharing CG node (nix, niy) of a degree-CG mesh, in the order a two-colour
// row-stripe scatter visits them: rows of one parity first, then the other, columns
// ascending. loc is the element-local CG dof jx + (CG+1) jy of the node.
struct Adj {
    int n;
    int e[4];
    int loc[4];
};
template <int CG> __device__ __forceinline__ Adj adjacent(int nix, int niy, int nx, int ny, bool evenRowsFirst)
{
    int ex[2], lx[2], cx = 0, ey[2], ly[2], cy = 0;
    if (CG == 2 && (nix & 1)) {
        ex[0] = nix >> 1;
        lx[0] = 1;
        cx = 1;
    } else {
        const int c = nix / CG;
        if (c > 0) {
            ex[cx] = c - 1;
            lx[cx++] = CG;
        }
        if (c < nx) {
            ex[cx] = c;
            lx[cx++] = 0;
        }
    }
    if (CG == 2 && (niy & 1)) {
        ey[0] = niy >> 1;
        ly[0] = 1;
        cy = 1;
    } else {
        const int c = niy / CG;
        if (c > 0) {
            ey[cy] = c - 1;
            ly[cy++] = CG;
        }
        if (c < ny) {
            ey[cy] = c;
            ly[cy++] = 0;
        }
    }
    int order[2] = { 0, 1 };
    if (cy == 2) {
        const bool firstIsEven = (ey[0] & 1) == 0;
        if (firstIsEven != evenRowsFirst) {
            order[0] = 1;
            order[1] = 0;
        }
    }
    Adj a;
    a.n = 0;
    for (int r = 0; r < cy; ++r) {
        const int rr = order[r];
        for (int k = 0; k < cx; ++k) {
            a.e[a.n] = ex[k] + nx * ey[rr];
            a.loc[a.n] = lx[k] + (CG + 1) * ly[rr];
            ++a.n;
        }
    }
    return a;
}

} // namespace nsdg
