// Internal declarations of the gfx950 dynamics library: errors, device buffers,
// the host-side mesh description and the device views the kernels receive.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/nsdg_dyn.h"
#include "nsdg_basis.hpp"

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace nsdg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m)
        : std::runtime_error(m)
        , code(c)
    {
    }
};

#define NSDG_HIP(expr)                                                                                   \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            throw ::nsdg::Error(NSDG_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

inline void require(bool ok, const std::string& msg, int code = NSDG_ERR_ARG)
{
    if (!ok)
        throw Error(code, msg);
}

constexpr int kBlock = 256;
inline unsigned blocksFor(size_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// Plain device allocation owned by the host object.
template <class T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count)
    {
        release();
        n = count;
        if (n)
            NSDG_HIP(hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)));
    }
    void upload(const T* h, size_t count, hipStream_t s, size_t offset = 0)
    {
        require(offset + count <= n, "upload out of range");
        if (count)
            NSDG_HIP(hipMemcpyAsync(p + offset, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* h, size_t count, hipStream_t s, size_t offset = 0) const
    {
        require(offset + count <= n, "download out of range");
        if (count)
            NSDG_HIP(hipMemcpyAsync(h, p + offset, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// ---------------------------------------------------------------- host mesh
// The structured parametric mesh as the reference's ParametricMesh holds it
// (dynamics/src/include/ParametricMesh.hpp:42-442): vertices, ocean mask, the four
// Dirichlet element lists and the periodic segments.
struct HostMesh {
    size_t nx = 0, ny = 0;
    bool spherical = false;     // coords are (longitude, latitude) in radians, rotated at create
    std::vector<double> coords; // (nx+1)(ny+1) x 2
    std::vector<uint8_t> ocean; // 1 = ocean/ice element
    std::array<std::vector<uint64_t>, 4> dirichlet;
    std::vector<std::vector<std::array<uint64_t, 4>>> periodic;

    size_t ne() const { return nx * ny; }
    size_t nv() const { return (nx + 1) * (ny + 1); }
    void create(size_t nx_, size_t ny_, const double* c, bool sph = false);
    void rotatePoleToGreenland();
    void setLandmask(const double* mask);
    void dirichletFromMask();
    void dirichletFromEdge(int edge);
    void addPeriodic(const uint64_t* entries, size_t n);
    void validate() const;
};

// ---------------------------------------------------------------- device mesh
// Per-element flags: Dirichlet multiplicity per side in bytes (bottom, right, top,
// left) of eflags; per-edge multipliers of the normal velocity (2^count).
struct DeviceMesh {
    int nx = 0, ny = 0;
    size_t ne = 0;
    bool uniform = false;     // all elements congruent translates of element 0
    bool affine = false;      // uniform, and element 0 is a parallelogram (constant Jacobian)
    double xiX = 0, etaX = 0, xiY = 0, etaY = 0, detJ = 0; // inverse map gradient when affine
    bool hasPeriodic = false;
    bool spherical = false;   // (longitude, latitude) vertices: metric terms, never uniform
    DBuf<double> vxy;         // vertices, interleaved x, y
    DBuf<uint8_t> ocean;
    DBuf<uint32_t> eflags;
    DBuf<double> factX, factY;
    DBuf<int32_t> periodic;   // 4 ints per entry, segments concatenated
    std::vector<std::pair<size_t, size_t>> segments; // (first entry, count)
    HostMesh host;            // copy kept for CG node flags and state queries
    void build(const HostMesh& m, hipStream_t s);
};

// View passed by value to kernels.
struct MeshView {
    int nx, ny;
    const double* vxy;
    const uint8_t* ocean;
    const uint32_t* eflags;
    const double* factX;
    const double* factY;
    int spherical;
};
inline MeshView view(const DeviceMesh& m)
{
    return { m.nx, m.ny, m.vxy.p, m.ocean.p, m.eflags.p, m.factX.p, m.factY.p, m.spherical ? 1 : 0 };
}
// NextsimDynamics.hpp:34
constexpr double kEarthRadius = 6371000.0;

// ---------------------------------------------------------------- kernel timing
struct Profiler {
    bool on = false;
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::map<std::string, std::pair<int, double>> stats;
    void begin(const char* name, hipStream_t s, Rec& r);
    void end(Rec& r, hipStream_t s);
    void collect();
    ~Profiler();
};

// RAII bracket for one kernel launch when profiling is on.
struct Timed {
    Profiler& p;
    hipStream_t s;
    Profiler::Rec r;
    bool active;
    Timed(Profiler& p_, hipStream_t s_, const char* name)
        : p(p_)
        , s(s_)
        , active(p_.on)
    {
        if (active)
            p.begin(name, s, r);
    }
    ~Timed()
    {
        if (active)
            p.end(r, s);
    }
};

void checkLaunch(const char* what);

} // namespace nsdg
