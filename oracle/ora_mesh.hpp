// TEST INFRASTRUCTURE — FP64 CPU oracle (see ora_basis.hpp header note).
//
// Structured parametric quadrilateral mesh, restating
//   dynamics/src/include/ParametricMesh.hpp:42-442 and dynamics/src/ParametricMesh.cpp:8-314.
// Vertices row-major (nx+1)*(ny+1) x 2, node id = ix + (nx+1)*iy. Element id = ix + nx*iy.
// landmask[e] == 1 means ocean/ice (ParametricMesh.cpp:217-223 keeps mask == 1).
// dirichlet[0..3] = bottom, right, top, left element lists, each sorted.
// periodic[seg][k] = {type (0 X-edge, 1 Y-edge), element left/bottom, element right/top, edge id}.
#pragma once
#include <algorithm>
#include <cmath>
#include <array>
#include <cstddef>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace ora {

struct Mesh {
    size_t nx = 0, ny = 0, nelements = 0, nnodes = 0;
    bool spherical = false; // CoordinateSystem == SPHERICAL: vertices are (longitude, latitude) in radians
    std::vector<double> vx, vy;
    std::vector<unsigned char> landmask;
    std::array<std::vector<size_t>, 4> dirichlet;
    std::vector<std::vector<std::array<size_t, 4>>> periodic;

    void init(size_t nx_, size_t ny_)
    {
        nx = nx_;
        ny = ny_;
        nelements = nx * ny;
        nnodes = (nx + 1) * (ny + 1);
        spherical = false;
        vx.assign(nnodes, 0.0);
        vy.assign(nnodes, 0.0);
        landmask.assign(nelements, 1);
        for (auto& d : dirichlet)
            d.clear();
        periodic.clear();
    }

    // correctlongitude (ParametricMesh.hpp:196-210): when a longitude of the element lies
    // more than 2 pi / 3 from the first one, the negative longitudes are moved up by 2 pi.
    template <int N> void correctLongitude(double c[N][2]) const
    {
        bool problem = false;
        for (int i = 1; i < N; ++i)
            if (std::fabs(c[0][0] - c[i][0]) > 2.0 / 3.0 * M_PI) {
                problem = true;
                break;
            }
        if (problem)
            for (int i = 0; i < N; ++i)
                if (c[i][0] < 0)
                    c[i][0] += 2.0 * M_PI;
    }
    // coordinatesOfElement (ParametricMesh.hpp:218-232): LL, LR, UL, UR
    void coordinatesOfElement(size_t eid, double c[4][2]) const
    {
        const size_t nid = (eid / nx) * (nx + 1) + (eid % nx);
        const size_t ids[4] = { nid, nid + 1, nid + nx + 1, nid + nx + 2 };
        for (int k = 0; k < 4; ++k) {
            c[k][0] = vx[ids[k]];
            c[k][1] = vy[ids[k]];
        }
        if (spherical)
            correctLongitude<4>(c);
    }
    // edgevector (ParametricMesh.hpp:381-397): on the sphere the longitude difference is
    // brought back across the +-pi jump
    void edgevector(size_t n1, size_t n2, double t[2]) const
    {
        t[0] = vx[n2] - vx[n1];
        t[1] = vy[n2] - vy[n1];
        if (spherical) {
            if (t[0] > 0.5 * M_PI)
                t[0] -= 2.0 * M_PI;
            if (t[0] < -0.5 * M_PI)
                t[0] += 2.0 * M_PI;
        }
    }
    // RotatePoleToGreenland (ParametricMesh.hpp:137-157): rotation by 40 degrees about the
    // polar axis, then by 15 degrees about the new y axis
    void rotatePoleToGreenland()
    {
        for (size_t i = 0; i < nnodes; ++i) {
            const double x = std::cos(vy[i]) * std::cos(vx[i]);
            const double y = std::cos(vy[i]) * std::sin(vx[i]);
            const double z = std::sin(vy[i]);
            const double aw = 40.0 * M_PI / 180.0;
            const double x1 = std::cos(aw) * x - std::sin(aw) * y;
            const double y1 = std::sin(aw) * x + std::cos(aw) * y;
            const double z1 = z;
            const double bw = 15.0 * M_PI / 180.0;
            const double x2 = std::cos(bw) * x1 - std::sin(bw) * z1;
            const double y2 = y1;
            const double z2 = std::sin(bw) * x1 + std::cos(bw) * z1;
            vy[i] = std::asin(z2);
            vx[i] = std::atan2(y2, x2);
        }
    }

    // coordinatesFromModelArray (ParametricMesh.cpp:198-210): coords is N_v x 2 row-major
    void coordinatesFromArray(size_t nx_, size_t ny_, const double* coords)
    {
        init(nx_, ny_);
        for (size_t i = 0; i < nnodes; ++i) {
            vx[i] = coords[2 * i];
            vy[i] = coords[2 * i + 1];
        }
    }
    // landmaskFromModelArray (ParametricMesh.cpp:217-223)
    void landmaskFromArray(const double* mask)
    {
        for (size_t i = 0; i < nelements; ++i)
            landmask[i] = (mask[i] == 1.) ? 1 : 0;
    }
    void sortDirichlet(int edge) { std::sort(dirichlet[edge].begin(), dirichlet[edge].end()); }
    // dirichletFromMask (ParametricMesh.cpp:228-252)
    void dirichletFromMask()
    {
        const size_t startX[4] = { 0, 0, 0, 1 };
        const size_t stopX[4] = { nx, nx - 1, nx, nx };
        const size_t startY[4] = { 1, 0, 0, 0 };
        const size_t stopY[4] = { ny, ny, ny - 1, ny };
        const long delta[4] = { -static_cast<long>(nx), 1, static_cast<long>(nx), -1 };
        for (int edge = 0; edge < 4; ++edge) {
            for (size_t j = startY[edge]; j < stopY[edge]; ++j)
                for (size_t i = startX[edge]; i < stopX[edge]; ++i) {
                    const size_t idx = i + nx * j;
                    if (!landmask[idx])
                        continue;
                    if (!landmask[static_cast<size_t>(static_cast<long>(idx) + delta[edge])])
                        dirichlet[edge].push_back(idx);
                }
            sortDirichlet(edge);
        }
    }
    // dirichletFromEdge (ParametricMesh.cpp:259-272)
    void dirichletFromEdge(int edge)
    {
        const size_t start[4] = { 0, nx - 1, nelements - nx, 0 };
        const size_t stop[4] = { nx, nelements, nelements, nelements };
        const size_t stride[4] = { 1, nx, 1, nx };
        for (size_t idx = start[edge]; idx < stop[edge]; idx += stride[edge])
            if (landmask[idx])
                dirichlet[edge].push_back(idx);
        sortDirichlet(edge);
    }
    // The sequence DynamicsKernel::initialise runs (DynamicsKernel.hpp:44-57); spherical
    // coordinates arrive in radians and are rotated.
    void fromCoordsAndMask(size_t nx_, size_t ny_, const double* coords, const double* mask, bool sph = false)
    {
        coordinatesFromArray(nx_, ny_, coords);
        spherical = sph;
        if (sph)
            rotatePoleToGreenland();
        landmaskFromArray(mask);
        dirichletFromMask();
        for (int e = 0; e < 4; ++e)
            dirichletFromEdge(e);
    }

    // readmesh (ParametricMesh.cpp:8-190), from the text of a .smesh file
    void readSmesh(const std::string& text)
    {
        std::istringstream IN(text);
        std::string status, version;
        IN >> status >> version;
        if (status != "ParametricMesh" || (version != "1.0" && version != "2.0"))
            throw std::runtime_error("readSmesh: wrong file format");
        size_t nx_, ny_;
        IN >> nx_ >> ny_;
        init(nx_, ny_);
        for (size_t i = 0; i < nnodes; ++i)
            IN >> vx[i] >> vy[i];
        if (version == "1.0") {
            for (size_t i = 0; i < nx; ++i)
                dirichlet[0].push_back(i);
            for (size_t i = 0; i < ny; ++i)
                dirichlet[1].push_back(i * nx + nx - 1);
            for (size_t i = 0; i < nx; ++i)
                dirichlet[2].push_back(i + nx * (ny - 1));
            for (size_t i = 0; i < ny; ++i)
                dirichlet[3].push_back(i * nx);
            return;
        }
        IN >> status;
        if (status != "landmask")
            throw std::runtime_error("readSmesh: landmask expected");
        size_t ne;
        IN >> ne;
        if (ne != 0) {
            landmask.assign(nelements, 0);
            for (size_t i = 0; i < nelements; ++i) {
                int lm;
                IN >> lm;
                if (lm)
                    landmask[i] = 1;
            }
        }
        IN >> status;
        if (status != "dirichlet")
            throw std::runtime_error("readSmesh: dirichlet expected");
        size_t nd;
        IN >> nd;
        for (size_t i = 0; i < nd; ++i) {
            size_t n0, n1;
            IN >> n0 >> n1;
            dirichlet[n1].push_back(n0);
        }
        IN >> status;
        if (status != "periodic")
            throw std::runtime_error("readSmesh: periodic expected");
        IN >> nd;
        periodic.assign(nd, {});
        for (size_t i = 0; i < nd; ++i) {
            size_t nind;
            IN >> nind;
            for (size_t j = 0; j < nind; ++j) {
                size_t n0, n1, n2;
                IN >> n0 >> n1 >> n2;
                if (n2 == 0)
                    periodic[i].push_back({ n2, n0, n1, n0 });
                else
                    periodic[i].push_back({ n2, n0, n1, (n0 / nx) * (nx + 1) + n0 % nx });
            }
        }
    }
};

} // namespace ora
