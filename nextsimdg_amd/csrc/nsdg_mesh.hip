// Host-side mesh set-up (integer boundary lists, flags) and its device upload.
// Behaviour follows ParametricMesh (dynamics/src/ParametricMesh.cpp:198-292).
#include "nsdg_internal.hpp"

#include <algorithm>
#include <cmath>

namespace nsdg {

void HostMesh::create(size_t nx_, size_t ny_, const double* c, bool sph)
{
    require(nx_ >= 2 && ny_ >= 2, "mesh needs at least 2 x 2 elements");
    require(c != nullptr, "coords is NULL");
    nx = nx_;
    ny = ny_;
    spherical = sph;
    coords.assign(c, c + 2 * nv());
    if (sph)
        rotatePoleToGreenland();
    ocean.assign(ne(), 1);
    for (auto& d : dirichlet)
        d.clear();
    periodic.clear();
}

// RotatePoleToGreenland (ParametricMesh.hpp:137-157), applied by DynamicsKernel::initialise
// to spherical meshes (DynamicsKernel.hpp:50-51): 40 degrees about the polar axis, then
// 15 degrees about the new y axis.
void HostMesh::rotatePoleToGreenland()
{
    const double aw = 40.0 * M_PI / 180.0, bw = 15.0 * M_PI / 180.0;
    for (size_t i = 0; i < nv(); ++i) {
        const double lon = coords[2 * i], lat = coords[2 * i + 1];
        const double x = std::cos(lat) * std::cos(lon), y = std::cos(lat) * std::sin(lon), z = std::sin(lat);
        const double x1 = std::cos(aw) * x - std::sin(aw) * y, y1 = std::sin(aw) * x + std::cos(aw) * y, z1 = z;
        const double x2 = std::cos(bw) * x1 - std::sin(bw) * z1, y2 = y1, z2 = std::sin(bw) * x1 + std::cos(bw) * z1;
        coords[2 * i + 1] = std::asin(z2);
        coords[2 * i] = std::atan2(y2, x2);
    }
}

// landmaskFromModelArray (ParametricMesh.cpp:217-223)
void HostMesh::setLandmask(const double* mask)
{
    require(mask != nullptr, "mask is NULL");
    for (size_t i = 0; i < ne(); ++i)
        ocean[i] = (mask[i] == 1.0) ? 1 : 0;
}

// dirichletFromMask (ParametricMesh.cpp:228-252): an ocean element whose neighbour
// across a side is land gets a Dirichlet condition on that side.
void HostMesh::dirichletFromMask()
{
    const long lx = static_cast<long>(nx);
    for (int side = 0; side < 4; ++side) {
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t e = i + nx * j;
                if (!ocean[e])
                    continue;
                long nb;
                if (side == 0) {
                    if (j == 0)
                        continue;
                    nb = static_cast<long>(e) - lx;
                } else if (side == 1) {
                    if (i + 1 == nx)
                        continue;
                    nb = static_cast<long>(e) + 1;
                } else if (side == 2) {
                    if (j + 1 == ny)
                        continue;
                    nb = static_cast<long>(e) + lx;
                } else {
                    if (i == 0)
                        continue;
                    nb = static_cast<long>(e) - 1;
                }
                if (!ocean[nb])
                    dirichlet[side].push_back(e);
            }
        std::sort(dirichlet[side].begin(), dirichlet[side].end());
    }
}

// dirichletFromEdge (ParametricMesh.cpp:259-272): ocean elements on a domain edge.
void HostMesh::dirichletFromEdge(int side)
{
    require(side >= 0 && side < 4, "edge must be 0..3");
    if (side == 0 || side == 2) {
        const size_t j = side == 0 ? 0 : ny - 1;
        for (size_t i = 0; i < nx; ++i)
            if (ocean[i + nx * j])
                dirichlet[side].push_back(i + nx * j);
    } else {
        const size_t i = side == 3 ? 0 : nx - 1;
        for (size_t j = 0; j < ny; ++j)
            if (ocean[i + nx * j])
                dirichlet[side].push_back(i + nx * j);
    }
    std::sort(dirichlet[side].begin(), dirichlet[side].end());
}

void HostMesh::addPeriodic(const uint64_t* entries, size_t n)
{
    std::vector<std::array<uint64_t, 4>> seg(n);
    for (size_t k = 0; k < n; ++k)
        for (int j = 0; j < 4; ++j)
            seg[k][j] = entries[4 * k + j];
    periodic.push_back(std::move(seg));
}

void HostMesh::validate() const
{
    for (int s = 0; s < 4; ++s)
        for (auto e : dirichlet[s])
            require(e < ne(), "Dirichlet element id out of range");
    for (const auto& seg : periodic)
        for (const auto& p : seg) {
            require(p[0] <= 1, "periodic entry type must be 0 or 1");
            require(p[1] < ne() && p[2] < ne(), "periodic element id out of range");
            require(p[3] < (p[0] == 0 ? nx * (ny + 1) : (nx + 1) * ny), "periodic edge id out of range");
        }
}

void DeviceMesh::build(const HostMesh& m, hipStream_t s)
{
    m.validate();
    host = m;
    nx = static_cast<int>(m.nx);
    ny = static_cast<int>(m.ny);
    ne = m.ne();
    require(m.nx < (1u << 30) && m.ny < (1u << 30) && (m.nx + 1) * (m.ny + 1) < (1ull << 31),
        "mesh too large for 32-bit node indices");

    // Uniform detection: every element's vertex offsets equal element 0's to 1e-12 h.
    const double* c = m.coords.data();
    auto off = [&](size_t e, int k, int xy) {
        const size_t nid = (e / m.nx) * (m.nx + 1) + e % m.nx;
        const size_t ids[4] = { nid, nid + 1, nid + m.nx + 1, nid + m.nx + 2 };
        return c[2 * ids[k] + xy] - c[2 * nid + xy];
    };
    const double h = std::max(std::fabs(off(0, 3, 0)), std::fabs(off(0, 3, 1)));
    uniform = h > 0;
    for (size_t e = 0; e < ne && uniform; ++e)
        for (int k = 1; k < 4 && uniform; ++k)
            for (int xy = 0; xy < 2; ++xy)
                if (std::fabs(off(e, k, xy) - off(0, k, xy)) > 1e-12 * h) {
                    uniform = false;
                    break;
                }
    // Affine: element 0 is a parallelogram, so d(x,y)/d(xi,eta) is constant on it.
    affine = uniform && std::fabs(off(0, 3, 0) - off(0, 1, 0) - off(0, 2, 0)) <= 1e-12 * h
        && std::fabs(off(0, 3, 1) - off(0, 1, 1) - off(0, 2, 1)) <= 1e-12 * h;
    if (affine) {
        const double xXi = off(0, 1, 0), xEta = off(0, 2, 0), yXi = off(0, 1, 1), yEta = off(0, 2, 1);
        detJ = xXi * yEta - xEta * yXi;
        affine = detJ > 0;
        xiX = yEta / detJ;
        xiY = -xEta / detJ;
        etaX = -yXi / detJ;
        etaY = xXi / detJ;
    }
    // congruent translates on a structured grid are parallelograms; a degenerate or
    // inverted element 0 leaves the per-element operator path. On the sphere the metric
    // varies with latitude, so no two elements share operators.
    spherical = m.spherical;
    if (spherical)
        affine = false;
    uniform = affine;

    std::vector<uint32_t> fl(ne, 0);
    std::vector<double> fX(m.nx * (m.ny + 1), 1.0), fY((m.nx + 1) * m.ny, 1.0);
    for (int side = 0; side < 4; ++side)
        for (auto e : m.dirichlet[side]) {
            const uint32_t cnt = (fl[e] >> (8 * side)) & 0xffu;
            require(cnt < 255, "Dirichlet list repeats an element too often");
            fl[e] += 1u << (8 * side);
            const size_t ix = e % m.nx, iy = e / m.nx;
            if (side == 0)
                fX[m.nx * iy + ix] *= 2.0;
            else if (side == 1)
                fY[(m.nx + 1) * iy + ix + 1] *= 2.0;
            else if (side == 2)
                fX[m.nx * (iy + 1) + ix] *= 2.0;
            else
                fY[(m.nx + 1) * iy + ix] *= 2.0;
        }
    vxy.alloc(2 * m.nv());
    vxy.upload(m.coords.data(), 2 * m.nv(), s);
    ocean.alloc(ne);
    ocean.upload(m.ocean.data(), ne, s);
    eflags.alloc(ne);
    eflags.upload(fl.data(), ne, s);
    factX.alloc(fX.size());
    factX.upload(fX.data(), fX.size(), s);
    factY.alloc(fY.size());
    factY.upload(fY.data(), fY.size(), s);

    std::vector<int32_t> per;
    segments.clear();
    for (const auto& seg : m.periodic) {
        segments.emplace_back(per.size() / 4, seg.size());
        for (const auto& p : seg)
            for (int j = 0; j < 4; ++j)
                per.push_back(static_cast<int32_t>(p[j]));
    }
    hasPeriodic = !per.empty();
    periodic.alloc(per.size());
    periodic.upload(per.data(), per.size(), s);
    NSDG_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- profiler
void Profiler::begin(const char* name, hipStream_t s, Rec& r)
{
    r.name = name;
    NSDG_HIP(hipEventCreate(&r.a));
    NSDG_HIP(hipEventCreate(&r.b));
    NSDG_HIP(hipEventRecord(r.a, s));
}
void Profiler::end(Rec& r, hipStream_t s)
{
    NSDG_HIP(hipEventRecord(r.b, s));
    pending.push_back(r);
}
void Profiler::collect()
{
    for (auto& r : pending) {
        NSDG_HIP(hipEventSynchronize(r.b));
        float ms = 0;
        NSDG_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        auto& st = stats[r.name];
        st.first += 1;
        st.second += ms;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    pending.clear();
}
Profiler::~Profiler()
{
    for (auto& r : pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
}

void checkLaunch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        throw Error(NSDG_ERR_HIP, std::string("launch of ") + what + " failed: " + hipGetErrorString(e));
}

} // namespace nsdg
