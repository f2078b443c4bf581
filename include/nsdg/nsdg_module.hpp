// Module-level mirror of the reference's dynamics modules over the C++ wrapper of
// nsdg_dyn.hpp, for C++ callers that run the kernels inside a model-like time loop:
//   IDynamics            core/src/modules/include/IDynamics.hpp:16-135
//   MEVPDynamics         core/src/modules/DynamicsModule/MEVPDynamics.cpp:28-87
//   BBMDynamics          core/src/modules/DynamicsModule/BBMDynamics.cpp:16-98
//   BenchmarkAtmosphere  physics/src/modules/AtmosphereBoundaryModule/BenchmarkAtmosphere.cpp:29-76
//   BenchmarkOcean       physics/src/modules/OceanBoundaryModule/BenchmarkOcean.cpp:15-37
// The ModelArrayRef store is a ModelStore of named H fields (nx * ny doubles, element
// ix + nx * iy). Every update sets h, A (and the damage) from the store as DG0 fields,
// which zeroes their higher DG moments on the device, runs the kernel and writes the cell
// means back. Header-only; link against libnsdg_hip.so.
#ifndef NSDG_MODULE_HPP
#define NSDG_MODULE_HPP

#include "nsdg_dyn.hpp"

#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace nsdg {

using Field = std::vector<double>;

// Store names of the shared and protected fields (Shared::, Protected:: of ModelArrayRef.hpp)
inline const std::string H_ICE = "H_ICE", C_ICE = "C_ICE", DAMAGE = "DAMAGE", DAMAGE_IN = "DAMAGE0";
inline const std::string WIND_U = "WIND_U", WIND_V = "WIND_V", OCEAN_U = "OCEAN_U", OCEAN_V = "OCEAN_V", SSH = "SSH";
constexpr double kRadians = 0x1.1df46a2529d39p-6; // degrees to radians (MEVPDynamics.cpp:21)
constexpr double kMissingValue = 1.7e38;          // MissingData::defaultValue (MissingData.hpp:18)

struct TimestepTime {
    double start; // seconds since the run's reference time
    double step;  // seconds
};

using ModelStore = std::map<std::string, Field>;

// The initial data map: grid size, vertex coordinates (2 per vertex), the mask, h, A, u,
// v (H fields), optionally damage; "x"/"y" (Cartesian) or "longitude"/"latitude" keys
// select the coordinate system as in IDynamics::checkSpherical.
struct DataMap {
    size_t nx = 0, ny = 0;
    std::map<std::string, Field> fields;
    bool has(const std::string& k) const { return fields.count(k) > 0; }
    const Field& at(const std::string& k) const { return fields.at(k); }
};

class IDynamics {
public:
    virtual ~IDynamics() = default;
    // MEVPDynamics::setData (MEVPDynamics.cpp:37-57)
    virtual void setData(const DataMap& ms)
    {
        bool spherical;
        if (ms.has("longitude") && ms.has("latitude"))
            spherical = true;
        else if (ms.has("x") && ms.has("y"))
            spherical = false;
        else
            throw std::runtime_error("Input data must contain either Cartesian (x, y) or spherical (longitude, "
                                     "latitude) coordinates.");
        Field coords = ms.at("coords");
        if (spherical)
            for (double& c : coords)
                c *= kRadians;
        ocean_.assign(ms.at("mask").size(), false);
        for (size_t i = 0; i < ocean_.size(); ++i)
            ocean_[i] = ms.at("mask")[i] == 1.0;
        kernel_->initialise(ms.nx, ms.ny, coords, spherical, ms.at("mask"));
        uice = ms.at("u");
        vice = ms.at("v");
        damage.assign(ms.nx * ms.ny, usesDamage_ ? 1.0 : 0.0);
        for (const char* n : { "hice", "cice", "u", "v" })
            kernel_->setData(n, ms.at(n));
    }
    // MEVPDynamics::update (MEVPDynamics.cpp:59-87)
    virtual void update(const TimestepTime& tst)
    {
        setInputs();
        kernel_->update(tst.step);
        getOutputs();
    }
    // IDynamics::getState (IDynamics.hpp:40-47): the masked ice velocity
    std::map<std::string, Field> getState() const { return { { "u", mask(uice) }, { "v", mask(vice) } }; }
    // ModelComponent::mask (ModelComponent.cpp:95-107)
    Field mask(const Field& f) const
    {
        Field out(f.size());
        for (size_t i = 0; i < f.size(); ++i)
            out[i] = ocean_[i] ? f[i] : kMissingValue;
        return out;
    }
    bool usesDamage() const { return usesDamage_; }
    DynamicsKernel& kernel() { return *kernel_; }

    Field uice, vice, taux, tauy, damage;

protected:
    IDynamics(ModelStore& store, std::unique_ptr<DynamicsKernel> k, bool usesDamage)
        : store_(store)
        , kernel_(std::move(k))
        , usesDamage_(usesDamage)
    {
    }
    void setInputs()
    {
        kernel_->setData("hice", store_.at(H_ICE));
        kernel_->setData("cice", store_.at(C_ICE));
        kernel_->setData("uwind", store_.at(WIND_U));
        kernel_->setData("vwind", store_.at(WIND_V));
        kernel_->setData("uocean", store_.at(OCEAN_U));
        kernel_->setData("vocean", store_.at(OCEAN_V));
        kernel_->setData("ssh", store_.at(SSH));
    }
    void getOutputs()
    {
        store_[H_ICE] = kernel_->getDG0Data("hice");
        store_[C_ICE] = kernel_->getDG0Data("cice");
        uice = kernel_->getDG0Data("u");
        vice = kernel_->getDG0Data("v");
        taux = kernel_->getDG0Data("uiostress");
        tauy = kernel_->getDG0Data("viostress");
    }
    ModelStore& store_;
    std::unique_ptr<DynamicsKernel> kernel_;
    bool usesDamage_;
    std::vector<bool> ocean_;
};

class MEVPDynamics : public IDynamics {
public:
    explicit MEVPDynamics(ModelStore& store, int cg = 2, int dg = 6, const nsdg_params& p = defaultParams())
        : IDynamics(store, std::make_unique<MEVPDynamicsKernel>(cg, dg, p), false)
    {
    }
};

class BBMDynamics : public IDynamics {
public:
    explicit BBMDynamics(ModelStore& store, int cg = 2, int dg = 6, const nsdg_params& p = defaultParams(),
        const nsdg_meb_params& m = defaultMEBParams())
        : IDynamics(store, std::make_unique<BBMDynamicsKernel>(cg, dg, p, m), true)
    {
    }
    // BBMDynamics::setData (BBMDynamics.cpp:20-56): damage defaults to 1, masked
    void setData(const DataMap& ms) override
    {
        IDynamics::setData(ms);
        damage = ms.has("damage") ? ms.at("damage") : mask(Field(ms.nx * ms.ny, 1.0));
        kernel_->setData("damage", damage);
    }
    // BBMDynamics::update (BBMDynamics.cpp:58-95)
    void update(const TimestepTime& tst) override
    {
        damage = store_.at(DAMAGE_IN);
        setInputs();
        kernel_->setData("damage", damage);
        kernel_->update(tst.step);
        getOutputs();
        damage = kernel_->getDG0Data("damage");
        store_[DAMAGE] = damage;
    }
};

// BenchmarkAtmosphere::update: the cyclone at the cell lower-left corners x = i L / nx
class BenchmarkAtmosphere {
public:
    BenchmarkAtmosphere(ModelStore& store, size_t nx, size_t ny, double L = 512e3, double vMax = 30.0)
        : store_(store)
        , nx_(nx)
        , ny_(ny)
        , L_(L)
        , vMax_(vMax)
    {
    }
    void update(const TimestepTime& tst)
    {
        if (!t0Set_) {
            t0_ = tst.start;
            t0Set_ = true;
        }
        const double frac = (tst.start - t0_) / 86400.0, dx = L_ / nx_, dy = L_ / ny_;
        const double x0 = nx_ * dx * 0.5 * (1 + frac / 5.0), y0 = ny_ * dy * 0.5 * (1 + frac / 5.0);
        const double alpha = 72. / 180. * M_PI, ca = std::cos(alpha), sa = std::sin(alpha);
        Field& u = store_[WIND_U];
        Field& v = store_[WIND_V];
        u.resize(nx_ * ny_);
        v.resize(nx_ * ny_);
        for (size_t j = 0; j < ny_; ++j)
            for (size_t i = 0; i < nx_; ++i) {
                const double xp = i * dx - x0, yp = j * dy - y0;
                const double scale = 1e-5 * std::exp(-1e-5 * std::hypot(xp, yp));
                u[i + nx_ * j] = -scale * vMax_ * (ca * xp + sa * yp);
                v[i + nx_ * j] = -scale * vMax_ * (-sa * xp + ca * yp);
            }
    }

private:
    ModelStore& store_;
    size_t nx_, ny_;
    double L_, vMax_, t0_ = 0;
    bool t0Set_ = false;
};

// BenchmarkOcean: the steady gyre u = 0.01 (2 y / L - 1), v = 0.01 (1 - 2 x / L), flat surface
class BenchmarkOcean {
public:
    BenchmarkOcean(ModelStore& store, size_t nx, size_t ny)
        : store_(store)
        , nx_(nx)
        , ny_(ny)
    {
    }
    void update(const TimestepTime&)
    {
        Field& u = store_[OCEAN_U];
        Field& v = store_[OCEAN_V];
        u.resize(nx_ * ny_);
        v.resize(nx_ * ny_);
        store_[SSH].assign(nx_ * ny_, 0.0);
        for (size_t j = 0; j < ny_; ++j)
            for (size_t i = 0; i < nx_; ++i) {
                u[i + nx_ * j] = 0.01 * (2.0 * j / ny_ - 1.0);
                v[i + nx_ * j] = 0.01 * (1.0 - 2.0 * i / nx_);
            }
    }

private:
    ModelStore& store_;
    size_t nx_, ny_;
};

} // namespace nsdg

#endif // NSDG_MODULE_HPP
