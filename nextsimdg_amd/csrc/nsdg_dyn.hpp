// The mEVP dynamics kernel on the device (MEVPDynamicsKernel<DG>,
// dynamics/src/include/MEVPDynamicsKernel.hpp:18-37 and its bases).
#pragma once
#include "nsdg_internal.hpp"
#include "nsdg_transport.hpp"

namespace nsdg {

struct DynBase {
    virtual ~DynBase() = default;
    virtual void initialise(const HostMesh& m) = 0;
    virtual bool initialised() const = 0;
    virtual void setField(const std::string& name, const double* data, int ncomp) = 0;
    virtual void getField(const std::string& name, double* out, int ncomp) = 0;
    virtual void getCG(const std::string& name, double* out) = 0;
    virtual void setCG(const std::string& name, const double* in) = 0;
    virtual size_t cgNodes() const = 0;
    virtual void update(double dt) = 0;
    virtual void advect(double dt) = 0;
    virtual void prepareIteration() = 0;
    virtual void beginIteration(double dt) = 0;
    virtual void strain() = 0;
    virtual void stress(int repeat) = 0;
    virtual void divergence() = 0;
    virtual void momentum() = 0;
    virtual void boundaries() = 0;
    virtual void subcycles(int n) = 0;
    virtual void setPrecision(int p) = 0;
    virtual void setPairing(int mode) = 0;
    virtual void benchmarkWind(double L, double t, double xoff, double yoff, double vmax) = 0;
    virtual void stormWind(double L, double t, const double* xoff, const double* yoff, const double* vmax, int n) = 0;
    virtual void resetDG0(const std::string& name) = 0;
    virtual void mass(double* vol, double* area) = 0;
    virtual void massNoAlloc(double* vol, double* area) = 0;
    virtual void setBudget(bool on) = 0;
    virtual void budget(double* out, bool reset) = 0;
    virtual void residual(double* r, double* normR, double* normD) = 0;
    virtual void stressFromVelocity() = 0;
    virtual void setBrittle(const nsdg_meb_params& m) = 0;
    virtual void subcyclesToResidual(double rTarget, int maxN, int every, double rStart, int* done, double* rFinal) = 0;
    // learned warm start (CG2 element grid; see include/nsdg_cnn.h)
    virtual int gridNx() const = 0;
    virtual int gridNy() const = 0;
    virtual void features(uint16_t* x) = 0;
    virtual void velocityTarget(float* t, float* w) = 0;
    virtual void applyVelocity(const float* y) = 0;
    virtual void saveStart() = 0;
    virtual void restoreStart() = 0;
    virtual void synchronize() = 0;
    virtual std::vector<char> saveState() = 0;
    virtual void loadState(const char* buf, size_t size) = 0;
    virtual Profiler& profiler() = 0;
    // mesh and discretisation, for the restart files
    virtual const HostMesh& meshHost() const = 0;
    virtual int cgDegree() const = 0;
    virtual int dgComp() const = 0;
    virtual int stressComp() const = 0;
    virtual bool brittleRheology() const = 0;
};

// Restart files (netCDF classic, 64-bit offsets) in the layout of ParaGridIO::dumpModelState
// (core/src/ParaGridIO.cpp:272-349); nsdg_restart.cpp.
void writeRestart(DynBase& k, const char* path, double timeSeconds);
void readRestart(DynBase& k, const char* path, double* timeSeconds);

std::unique_ptr<DynBase> makeDyn(int cg, int dg, const nsdg_params& p);

} // namespace nsdg
