// C++ caller of the gfx950 library through include/nsdg/nsdg_dyn.hpp: one module-level
// step sequence of a small benchmark box; prints the cell means of u, v, hice, cice (and
// damage with the brittle rheology), one field per line. Usage: wrapper_demo [steps] [mevp|bbm]
#include "nsdg/nsdg_dyn.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

int main(int argc, char** argv)
{
    const size_t nx = 24, ny = 24;
    const int steps = argc > 1 ? std::atoi(argv[1]) : 2;
    const bool brittle = argc > 2 && std::strcmp(argv[2], "bbm") == 0;
    const double L = 512e3, dx = L / nx;
    std::vector<double> coords(2 * (nx + 1) * (ny + 1)), mask(nx * ny, 1.0), h(nx * ny), a(nx * ny, 1.0),
        uo(nx * ny), vo(nx * ny), zero(nx * ny, 0.0);
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
            h[e] = 0.3 + 0.005 * (std::sin(6e-5 * i * dx) + std::sin(3e-5 * j * dx));
            uo[e] = 0.01 * (2.0 * j / ny - 1.0);
            vo[e] = 0.01 * (1.0 - 2.0 * i / nx);
        }
    try {
        std::unique_ptr<nsdg::DynamicsKernel> kp;
        if (brittle)
            kp = std::make_unique<nsdg::BBMDynamicsKernel>(2, 6);
        else
            kp = std::make_unique<nsdg::MEVPDynamicsKernel>(2, 6);
        nsdg::DynamicsKernel& k = *kp;
        k.initialise(nx, ny, coords, false, mask);
        k.setData("uocean", uo);
        k.setData("vocean", vo);
        for (int s = 0; s < steps; ++s) {
            k.setData("hice", h);
            k.setData("cice", a);
            k.setData("ssh", zero);
            nsdg::check(nsdg_dyn_benchmark_wind(k.handle(), L, s * 120.0, 0.0, 0.0, 30.0));
            k.update(120.0);
            h = k.getDG0Data("hice");
            a = k.getDG0Data("cice");
        }
        // every cell mean of every output field, one field per line
        auto print = [](const char* name, const std::vector<double>& f) {
            std::printf("%s", name);
            for (double x : f)
                std::printf(" %.17g", x);
            std::printf("\n");
        };
        print("u", k.getDG0Data("u"));
        print("v", k.getDG0Data("v"));
        print("hice", h);
        print("cice", a);
        if (brittle)
            print("damage", k.getDG0Data("damage"));
    } catch (const nsdg::Error& e) {
        std::fprintf(stderr, "nsdg error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
