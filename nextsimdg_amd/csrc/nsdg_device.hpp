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
// Vertices of element e: LL, LR, UL, UR (ParametricMesh.hpp:218-232). On the sphere
// the longitudes are made continuous across the +-pi jump: when one lies more than
// 2 pi / 3 from the first, the negative ones move up by 2 pi (correctlongitude,
// ParametricMesh.hpp:196-210).
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
    if (m.spherical) {
        bool jump = false;
#pragma unroll
        for (int k = 1; k < 4; ++k)
            jump = jump || fabs(c[0][0] - c[k][0]) > 2.0 / 3.0 * M_PI;
        if (jump)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (c[k][0] < 0)
                    c[k][0] += 2.0 * M_PI;
    }
}
// Longitude difference of an edge on the sphere brought back across the +-pi jump
// (edgevector, ParametricMesh.hpp:381-397).
__device__ __forceinline__ double edgeDx(const MeshView& m, double d)
{
    if (m.spherical) {
        if (d > 0.5 * M_PI)
            d -= 2.0 * M_PI;
        if (d < -0.5 * M_PI)
            d += 2.0 * M_PI;
    }
    return d;
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
// Latitude at the G x G Gauss points (getGaussPointsInElement, ParametricTools.hpp:159-164).
template <int G> __device__ void gaussLatitude(const double c[4][2], double lat[G * G])
{
#pragma unroll
    for (int q = 0; q < G * G; ++q) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            s += c[k][1] * Phi<1, G>::v[k][q];
        lat[q] = s;
    }
}
// Metric weight of the element integrals at the G x G Gauss points: cos(latitude) on the
// sphere, 1 otherwise.
template <int G> __device__ void metricWeight(const MeshView& m, const double c[4][2], double w[G * G])
{
    if (m.spherical) {
        gaussLatitude<G>(c, w);
        for (int q = 0; q < G * G; ++q)
            w[q] = cos(w[q]);
    } else {
        for (int q = 0; q < G * G; ++q)
            w[q] = 1.0;
    }
}
// The element mass matrix with the metric weight: SphericalTools::massMatrix
// (ParametricTools.hpp:189-202) on the sphere, the Cartesian one otherwise.
template <int D> __device__ void metricMassMatrix(const MeshView& m, const double c[4][2], double M[D][D])
{
    if (!m.spherical) {
        massMatrix<D>(c, M);
        return;
    }
    constexpr int G = gp1d(D), Q = G * G;
    double J[Q], cl[Q];
    jacobian<G>(c, J);
    metricWeight<G>(m, c, cl);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            double s = 0;
            for (int q = 0; q < Q; ++q)
                s += Psi<D, G>::v[i][q] * (W<G>::v[0][q] * J[q] * cl[q]) * Psi<D, G>::v[j][q];
            M[i][j] = s;
        }
}
// Contribution of element e with DG coefficients c to Tools::MeanValue (Tools.hpp:24-45):
// the field at the 3 x 3 Gauss points (whatever DG is) times J w, and times cos(latitude)
// on the sphere (Tools.hpp:28-36), so spherical totals are in radians squared.
template <int DG> __device__ double elementMean(const MeshView& m, int e, const double* c)
{
    double cc[4][2], J[9], cl[9];
    elementCoords(m, e, cc);
    jacobian<3>(cc, J);
    metricWeight<3>(m, cc, cl);
    double s = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        double v = 0;
#pragma unroll
        for (int i = 0; i < DG; ++i)
            v += c[i] * Psi<DG, 3>::v[i][q];
        s += m.spherical ? v * cl[q] * (J[q] * W<3>::v[0][q]) : v * (J[q] * W<3>::v[0][q]);
    }
    return s;
}

// ---------------------------------------------------------------- compensated sums
// Two-sum: s + err == a + b exactly.
__device__ __forceinline__ void twoSum(double a, double b, double& s, double& err)
{
    s = a + b;
    const double bb = s - a;
    err = (a - (s - bb)) + (b - bb);
}
// Compensated sum of NF values over a block of kBlock threads, in a fixed tree order
// (deterministic). sum[f], comp[f] are each thread's running sum and its compensation;
// thread 0 returns the block's pairs in sum, comp.
template <int NF> __device__ void blockTwoSum(double (*sh)[kBlock], double* sum, double* comp)
{
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        sh[2 * f][threadIdx.x] = sum[f];
        sh[2 * f + 1][threadIdx.x] = comp[f];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(w)) {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                double s, err;
                twoSum(sh[2 * f][threadIdx.x], sh[2 * f][threadIdx.x + w], s, err);
                sh[2 * f][threadIdx.x] = s;
                sh[2 * f + 1][threadIdx.x] += sh[2 * f + 1][threadIdx.x + w] + err;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        sum[f] = sh[2 * f][0];
        comp[f] = sh[2 * f + 1][0];
    }
}
// Final pass over nb block partials laid out part[(b * NF + f) * 2 + {0: sum, 1: comp}]:
// one block; thread t takes the blocks t, t + kBlock, ... in order, then the block tree.
// out[f] = sum + comp. Launch <<<1, kBlock>>>.
template <int NF> __global__ __launch_bounds__(kBlock) void k_sum_final(const double* part, int nb, double* out)
{
    __shared__ double sh[2 * NF][kBlock];
    double sum[NF], comp[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f)
        sum[f] = comp[f] = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            double s, err;
            twoSum(sum[f], part[(b * NF + f) * 2], s, err);
            sum[f] = s;
            comp[f] += err + part[(b * NF + f) * 2 + 1];
        }
    blockTwoSum<NF>(sh, sum, comp);
    if (threadIdx.x == 0)
#pragma unroll
        for (int f = 0; f < NF; ++f)
            out[f] = sum[f] + comp[f];
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
// Elements sharing CG node (nix, niy) of a degree-CG mesh, visited in the order a
// two-colour row-stripe scatter adds them: element rows of one parity first, then the
// other, columns ascending. f(e, loc) receives the element and the node's local CG dof
// jx + (CG+1) jy. Fully unrolled: every candidate lives in registers.
__device__ __forceinline__ void axisCandidates(int cg, int ni, int n, int& c0, int& l0, bool& h0, int& c1, int& l1,
    bool& h1)
{
    if (cg == 2 && (ni & 1)) {
        c0 = ni >> 1;
        l0 = 1;
        h0 = true;
        c1 = 0;
        l1 = 0;
        h1 = false;
        return;
    }
    const int c = ni / cg;
    c0 = c - 1;
    l0 = cg;
    h0 = c > 0;
    c1 = c;
    l1 = 0;
    h1 = c < n;
}
template <int CG, bool EVEN_ROWS_FIRST, class F>
__device__ __forceinline__ void forEachAdjacentXY(int nix, int niy, int nx, int ny, F&& f)
{
    int xa, la, xb, lb, ya, ma, yb, mb;
    bool ha, hb, ga, gb;
    axisCandidates(CG, nix, nx, xa, la, ha, xb, lb, hb);
    axisCandidates(CG, niy, ny, ya, ma, ga, yb, mb, gb);
    if (ga && gb && (((ya & 1) == 0) != EVEN_ROWS_FIRST)) {
        const int ty = ya, tm = ma;
        ya = yb;
        ma = mb;
        yb = ty;
        mb = tm;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const bool hr = r == 0 ? ga : gb;
        const int ey = r == 0 ? ya : yb, my = r == 0 ? ma : mb;
        if (!hr)
            continue;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const bool hc = c == 0 ? ha : hb;
            const int ex = c == 0 ? xa : xb, mx = c == 0 ? la : lb;
            if (!hc)
                continue;
            f(ex, ey, mx + (CG + 1) * my);
        }
    }
}
template <int CG, bool EVEN_ROWS_FIRST, class F>
__device__ __forceinline__ void forEachAdjacent(int nix, int niy, int nx, int ny, F&& f)
{
    forEachAdjacentXY<CG, EVEN_ROWS_FIRST>(nix, niy, nx, ny, [&](int ex, int ey, int loc) { f(ex + nx * ey, loc); });
}

// Block -> tile map that keeps consecutive tiles on one XCD: workgroups are dispatched
// round-robin over the 8 XCDs, so block b runs on XCD b % 8; giving that XCD the
// contiguous tile range [xcd * chunk, (xcd + 1) * chunk) keeps a tile's left and
// lower neighbours (its halo) in the same L2. The grid is padded to 8 * chunk blocks;
// returns -1 for the padding.
// A value of an SoA field addressed as the field's start (uniform: a scalar register pair)
// plus the element's byte offset (32 bits, per lane), in the global address space. The
// start goes through readfirstlane, which the compiler cannot see through, so that it
// does not fold the per-field offsets into per-load 64-bit vector addresses: all fields
// of one element then share one offset register (global_load ... vOff, sBase). The start
// must be the same in every lane (kernel arguments and uniform offsets), and off < 2^32,
// i.e. fewer than 2^29 elements of 8 bytes (the host checks).
template <class T> __device__ __forceinline__ __attribute__((address_space(1))) T& fieldAt(const T* base, uint32_t off)
{
    const uint64_t p = reinterpret_cast<uint64_t>(base);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(p));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(p >> 32));
    using GP = __attribute__((address_space(1))) char*;
    GP b = reinterpret_cast<GP>((static_cast<uint64_t>(hi) << 32) | lo);
    return *reinterpret_cast<__attribute__((address_space(1))) T*>(b + off);
}
// The offset, made to depend on x: loads addressed with the result cannot be scheduled
// before x is computed, which keeps a late group of loads (and its registers) late.
__device__ __forceinline__ uint32_t after(uint32_t off, double x)
{
    asm volatile("" : "+v"(off) : "v"(x));
    return off;
}
template <class T> __device__ __forceinline__ T ldAt(const T* base, uint32_t off) { return fieldAt(base, off); }
// A double2 as its two halves (adjacent: the compiler can issue one 16-byte load).
__device__ __forceinline__ double2 ld2At(const double2* base, uint32_t off)
{
    const double* b = reinterpret_cast<const double*>(base);
    return make_double2(ldAt(b, off), ldAt(b, off + 8u));
}
constexpr int kXCD = 8;
__device__ __forceinline__ int xcdTile(int b, int ntiles)
{
    const int chunk = (ntiles + kXCD - 1) / kXCD;
    const int t = (b % kXCD) * chunk + b / kXCD;
    return t < ntiles ? t : -1;
}
inline unsigned xcdGrid(int ntiles) { return static_cast<unsigned>(kXCD * ((ntiles + kXCD - 1) / kXCD)); }

} // namespace nsdg
