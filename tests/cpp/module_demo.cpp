// C++ caller of the module-level mirror (include/nsdg/nsdg_module.hpp): the benchmark box
// in a model-like time loop (forcing modules, then the dynamics module, each step; the
// damage the brittle module writes is its next input). Prints the sums of the store's
// hice, cice and of the module's u, v, taux, tauy (and damage).
// Usage: module_demo [steps] [mevp|bbm]
#include "nsdg/nsdg_module.hpp"

#include <cstdio>
#include <cstring>
#include <memory>

int main(int argc, char** argv)
{
    const size_t nx = 24, ny = 24;
    const int steps = argc > 1 ? std::atoi(argv[1]) : 3;
    const bool brittle = argc > 2 && std::strcmp(argv[2], "bbm") == 0;
    const double L = 512e3, dx = L / nx, dt = 120.0;
    nsdg::DataMap ms;
    ms.nx = nx;
    ms.ny = ny;
    nsdg::Field coords(2 * (nx + 1) * (ny + 1)), mask(nx * ny, 1.0), h(nx * ny), x(nx * ny), y(nx * ny);
    for (size_t j = 0; j <= ny; ++j)
        for (size_t i = 0; i <= nx; ++i) {
            coords[2 * (i + (nx + 1) * j)] = i * dx;
            coords[2 * (i + (nx + 1) * j) + 1] = j * dx;
        }
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
            const size_t e = i + nx * j;
            if (i == 0 || j == 0 || i == nx - 1 || j == ny - 1)
                mask[e] = 0.0;
            x[e] = i * dx;
            y[e] = j * dx;
            h[e] = 0.3 + 0.005 * (std::sin(60e-3 * x[e] / 1e3) + std::sin(30e-3 * y[e] / 1e3));
        }
    ms.fields = { { "coords", coords }, { "mask", mask }, { "x", x }, { "y", y }, { "hice", h },
        { "cice", nsdg::Field(nx * ny, 1.0) }, { "u", nsdg::Field(nx * ny, 0.0) }, { "v", nsdg::Field(nx * ny, 0.0) } };
    try {
        nsdg::ModelStore store;
        std::unique_ptr<nsdg::IDynamics> dyn;
        if (brittle)
            dyn = std::make_unique<nsdg::BBMDynamics>(store);
        else
            dyn = std::make_unique<nsdg::MEVPDynamics>(store);
        dyn->setData(ms);
        store[nsdg::H_ICE] = h;
        store[nsdg::C_ICE] = ms.at("cice");
        store[nsdg::DAMAGE_IN] = dyn->damage;
        nsdg::BenchmarkAtmosphere atm(store, nx, ny, L);
        nsdg::BenchmarkOcean ocn(store, nx, ny);
        for (int k = 0; k < steps; ++k) {
            const nsdg::TimestepTime tst { k * dt, dt };
            atm.update(tst);
            ocn.update(tst);
            dyn->update(tst);
            if (dyn->usesDamage())
                store[nsdg::DAMAGE_IN] = store[nsdg::DAMAGE];
        }
        // every cell value of every output field, one field per line
        auto print = [](const char* name, const nsdg::Field& f) {
            std::printf("%s", name);
            for (double v : f)
                std::printf(" %.17g", v);
            std::printf("\n");
        };
        print("hice", store[nsdg::H_ICE]);
        print("cice", store[nsdg::C_ICE]);
        print("u", dyn->uice);
        print("v", dyn->vice);
        print("uiostress", dyn->taux);
        print("viostress", dyn->tauy);
        if (brittle)
            print("damage", dyn->damage);
    } catch (const nsdg::Error& e) {
        std::fprintf(stderr, "nsdg error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
