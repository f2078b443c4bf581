// DG transport on the device: DGTransport<DG> (dynamics/src/include/DGTransport.hpp:27-152,
// dynamics/src/DGTransport.cpp) and the CG -> DG projection it is fed with
// (Interpolations::CG2DG, dynamics/src/Interpolations.cpp:74-122).
#pragma once
#include "nsdg_internal.hpp"

namespace nsdg {

// Runtime-degree interface over TransportImpl<DG>.
struct TransportBase {
    virtual ~TransportBase() = default;
    virtual int dg() const = 0;
    virtual void setScheme(int rk) = 0;
    virtual void setVelocityHost(const double* vx, const double* vy) = 0; // N_e x DG host layout
    virtual void prepareAdvectionHost(int cg, const double* u, const double* v) = 0;
    virtual void reinitNormalVelocity() = 0;
    // SoA field on the device; tally (optional, N_e doubles) accumulates each element's
    // Dirichlet outflow in Tools::MeanValue units
    virtual void step(double dt, double* phiDev, double* tally = nullptr) = 0;
    // tally (optional) accumulates the change of each element's MeanValue contribution
    virtual void limit(double* phiDev, bool doMax, double maxv, bool doMin, double minv, double* tally = nullptr) = 0;
    virtual double* velx() = 0;
    virtual double* vely() = 0;
};

// Owning wrapper used by the C ABI (mesh + transport + stream).
struct TransportHandle {
    hipStream_t stream = nullptr;
    Profiler prof;
    DeviceMesh mesh;
    std::unique_ptr<TransportBase> t;
    DBuf<double> field;
    ~TransportHandle();
};

std::unique_ptr<TransportBase> makeTransport(int dg, DeviceMesh& mesh, hipStream_t s, Profiler& p);

// CG -> DG projection operators and launch (u, v to the advection velocity, output fields).
struct CG2DGOps {
    int cg = 0, dg = 0;
    DBuf<double> ops;
    void build(int cg, int dg, const DeviceMesh& mesh, hipStream_t s);
    // dst: SoA DG x ne; src: CG node vector. Two fields at once when src2/dst2 are set.
    void apply(const DeviceMesh& mesh, const double* src, double* dst, const double* src2, double* dst2,
        hipStream_t s, Profiler& p) const;
};

// Host <-> device layout conversion (ModelArray N x ncomp <-> SoA ncomp x N).
void uploadSoA(DBuf<double>& scratch, const double* host, size_t n, int ncomp, double* dst, hipStream_t s);
void downloadSoA(DBuf<double>& scratch, const double* src, size_t n, int ncomp, double* host, hipStream_t s);
void deviceFill(double* p, size_t n, double v, hipStream_t s);
void deviceCopy(double* dst, const double* src, size_t n, hipStream_t s);

} // namespace nsdg
