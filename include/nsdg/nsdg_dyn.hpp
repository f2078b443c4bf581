// C++17 wrapper over the C ABI of nsdg_dyn.h with the reference's class and method
// names (ParametricMesh, DGTransport, DynamicsKernel with MEVPDynamicsKernel and
// BBMDynamicsKernel), for C++ callers.
// Status codes become nsdg::Error exceptions; buffers are std::vector<double> in the
// ModelArray layout. Header-only; link against libnsdg_hip.so.
#ifndef NSDG_DYN_HPP
#define NSDG_DYN_HPP

#include "../nsdg_dyn.h"

#include <stdexcept>
#include <string>
#include <utility>
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

inline void check(int rc)
{
    if (rc != NSDG_OK)
        throw Error(rc, nsdg_last_error());
}

inline nsdg_params defaultParams()
{
    nsdg_params p;
    check(nsdg_params_default(&p));
    return p;
}

// ParametricMesh (dynamics/src/include/ParametricMesh.hpp:42-442)
class ParametricMesh {
public:
    enum Edge { BOTTOM = 0, RIGHT = 1, TOP = 2, LEFT = 3 };
    ParametricMesh(size_t nx, size_t ny, const std::vector<double>& coords)
        : nx_(nx)
        , ny_(ny)
    {
        check(nsdg_mesh_create(nx, ny, coords.data(), 0, &h_));
    }
    ~ParametricMesh() { nsdg_mesh_destroy(h_); }
    ParametricMesh(const ParametricMesh&) = delete;
    ParametricMesh& operator=(const ParametricMesh&) = delete;
    void landmaskFromModelArray(const std::vector<double>& mask) { check(nsdg_mesh_set_landmask(h_, mask.data())); }
    void dirichletFromMask() { check(nsdg_mesh_dirichlet_from_mask(h_)); }
    void dirichletFromEdge(Edge e) { check(nsdg_mesh_dirichlet_from_edge(h_, e)); }
    std::vector<uint64_t> dirichlet(Edge e) const
    {
        size_t n = 0;
        check(nsdg_mesh_get_dirichlet(h_, e, nullptr, &n));
        std::vector<uint64_t> out(n);
        check(nsdg_mesh_get_dirichlet(h_, e, out.data(), &n));
        return out;
    }
    const nsdg_mesh* handle() const { return h_; }
    size_t nx() const { return nx_; }
    size_t ny() const { return ny_; }

private:
    nsdg_mesh* h_ = nullptr;
    size_t nx_, ny_;
};

// DGTransport<DG> (dynamics/src/include/DGTransport.hpp:27-152)
class DGTransport {
public:
    DGTransport(const ParametricMesh& m, int dg)
    {
        check(nsdg_transport_create(m.handle(), dg, &h_));
    }
    ~DGTransport() { nsdg_transport_destroy(h_); }
    DGTransport(const DGTransport&) = delete;
    DGTransport& operator=(const DGTransport&) = delete;
    void settimesteppingscheme(const std::string& s) { check(nsdg_transport_set_scheme(h_, s.c_str())); }
    void setVelocity(const std::vector<double>& vx, const std::vector<double>& vy)
    {
        check(nsdg_transport_set_velocity(h_, vx.data(), vy.data()));
    }
    void prepareAdvection(int cg, const std::vector<double>& u, const std::vector<double>& v)
    {
        check(nsdg_transport_prepare_advection(h_, cg, u.data(), v.data()));
    }
    void reinitnormalvelocity() { check(nsdg_transport_reinit_normal_velocity(h_)); }
    void step(double dt, std::vector<double>& phi) { check(nsdg_transport_step(h_, dt, phi.data())); }

private:
    nsdg_transport* h_ = nullptr;
};

// DynamicsKernel<DGadvection, DGstress> (dynamics/src/include/DynamicsKernel.hpp:37-229):
// the interface the module calls, shared by the two rheologies below.
class DynamicsKernel {
public:
    virtual ~DynamicsKernel()
    {
        if (h_)
            nsdg_dyn_destroy(h_);
    }
    DynamicsKernel(const DynamicsKernel&) = delete;
    DynamicsKernel& operator=(const DynamicsKernel&) = delete;

    void initialise(size_t nx, size_t ny, const std::vector<double>& coords, bool isSpherical,
        const std::vector<double>& mask)
    {
        nx_ = nx;
        ny_ = ny;
        check(nsdg_dyn_initialise(h_, nx, ny, coords.data(), isSpherical ? 1 : 0, mask.data()));
    }
    void setData(const std::string& name, const std::vector<double>& data)
    {
        check(nsdg_dyn_set_field(h_, name.c_str(), data.data(), static_cast<int>(data.size() / (nx_ * ny_))));
    }
    std::vector<double> getDG0Data(const std::string& name)
    {
        std::vector<double> out(nx_ * ny_);
        check(nsdg_dyn_get_field(h_, name.c_str(), out.data(), 1));
        return out;
    }
    std::vector<double> getDGData(const std::string& name)
    {
        const bool tensor = name.size() == 3 && (name[0] == 's' || name[0] == 'e');
        const int nc = tensor ? s_ : dg_;
        std::vector<double> out(nx_ * ny_ * nc);
        check(nsdg_dyn_get_field(h_, name.c_str(), out.data(), nc));
        return out;
    }
    void update(double dtSeconds) { check(nsdg_dyn_update(h_, dtSeconds)); }
    void synchronize() { check(nsdg_dyn_synchronize(h_)); }
    nsdg_dyn* handle() { return h_; }

protected:
    DynamicsKernel(int cgDegree, int dgAdvection)
        : dg_(dgAdvection)
        , s_(cgDegree == 1 ? 3 : 8)
    {
    }
    nsdg_dyn* h_ = nullptr;

private:
    int dg_, s_;
    size_t nx_ = 0, ny_ = 0;
};

// MEVPDynamicsKernel<DG> (dynamics/src/include/MEVPDynamicsKernel.hpp:18-37)
class MEVPDynamicsKernel : public DynamicsKernel {
public:
    MEVPDynamicsKernel(int cgDegree, int dgAdvection, const nsdg_params& p = defaultParams())
        : DynamicsKernel(cgDegree, dgAdvection)
    {
        check(nsdg_dyn_create(cgDegree, dgAdvection, &p, &h_));
    }
};

inline nsdg_meb_params defaultMEBParams()
{
    nsdg_meb_params m;
    check(nsdg_meb_params_default(&m));
    return m;
}

// BBMDynamicsKernel<DG> (dynamics/src/include/BBMDynamicsKernel.hpp:18-45); the extra
// field "damage" goes through setData / getDGData like hice.
class BBMDynamicsKernel : public DynamicsKernel {
public:
    BBMDynamicsKernel(int cgDegree, int dgAdvection, const nsdg_params& p = defaultParams(),
        const nsdg_meb_params& m = defaultMEBParams())
        : DynamicsKernel(cgDegree, dgAdvection)
    {
        check(nsdg_bbm_create(cgDegree, dgAdvection, &p, &m, &h_));
    }
};

} // namespace nsdg

#endif // NSDG_DYN_HPP
