// TEST INFRASTRUCTURE — FP64 CPU oracle (see ora_basis.hpp header note).
//
// extern "C" surface of the oracle, loaded through ctypes by tests/ and by
// bench.py's cpu_baseline leg. Also holds the analytic test cases of the reference's
// advection tests (dynamics/test/Advection_test.cpp, AdvectionPeriodicBC_test.cpp)
// and the benchmark forcing formulas (physics/src/modules/.../BenchmarkAtmosphere.cpp,
// BenchmarkOcean.cpp) used as inputs.
#ifdef _OPENMP
#include <omp.h>
#endif
#include "ora_core.hpp"

#include <cstdint>
#include <cstdio>
#include <memory>

using namespace ora;

namespace {
thread_local std::string g_err;

template <class F> int guard(F&& f)
{
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// ---- analytic functions of the reference tests
struct AdvCfg {
    double Lx = 409600.0, Ly = 512000.0;
};
const AdvCfg kAdv;
// SmoothBump (Advection_test.cpp:71-84)
double smoothBump(double x, double y)
{
    const double X = x / kAdv.Lx, Y = y / kAdv.Lx;
    const double r = (std::pow(X - 0.25, 2.0) + std::pow(Y - 0.5, 2.0)) / 0.025;
    return r < 1 ? std::exp(-1.0 / (1.0 - r)) : 0.0;
}
// InitialVX / InitialVY (Advection_test.cpp:87-101)
double rotVX(double x, double y) { return (y - 0.5 * kAdv.Lx) * 2.0 * M_PI / kAdv.Lx; }
double rotVY(double x, double y) { return (0.5 * kAdv.Lx - x) * 2.0 * M_PI / kAdv.Lx; }
// Packman and ring velocity (AdvectionPeriodicBC_test.cpp:71-113)
const double R0 = 100000.0, R1 = 250000.0;
double packman(double x, double y)
{
    double r = std::pow((x + 175000.0) / 50000, 2.0) + std::pow(y / 50000.0, 2.0);
    if (r < 1.0)
        return std::exp(1.0) * std::exp(-1.0 / (1.0 - r));
    r = std::pow(x / 50000, 2.0) + std::pow((y - 175000.0) / 50000.0, 2.0);
    if (r < 1.0) {
        if (y > 175000)
            return 1.0;
        if (std::fabs(x) > 15000.0)
            return 1.0;
        return 0.0;
    }
    r = std::sqrt(std::pow((x - 175000.0) / 50000, 2.0) + std::pow(y / 50000.0, 2.0));
    if (r < 1.0)
        return 1.0 - r;
    r = std::sqrt(std::pow(x / 50000, 2.0) + std::pow((y + 175000.) / 50000.0, 2.0));
    if (r < 1.0)
        return 1.0;
    return 0.0;
}
double ringVX(double x, double y) { return y * 2.0 * M_PI / R1; }
double ringVY(double x, double y) { return -x * 2.0 * M_PI / R1; }

Fn fnById(int id)
{
    switch (id) {
    case 0: return smoothBump;
    case 1: return rotVX;
    case 2: return rotVY;
    case 3: return packman;
    case 4: return ringVX;
    case 5: return ringVY;
    default: throw std::runtime_error("unknown test function id");
    }
}

// create_rectanglemesh (Advection_test.cpp:198-241)
void rectangleMesh(Mesh& m, double Lx, double Ly, size_t Nx, size_t Ny, double distort)
{
    m.init(Nx, Ny);
    size_t ii = 0;
    for (size_t iy = 0; iy <= Ny; ++iy)
        for (size_t ix = 0; ix <= Nx; ++ix, ++ii) {
            m.vx[ii] = Lx * ix / Nx + Lx * distort * std::sin(M_PI * ix / Nx * 3.0) * std::sin(M_PI * iy / Ny);
            m.vy[ii] = Ly * iy / Ny + Ly * distort * std::sin(M_PI * iy / Ny * 2.0) * std::sin(M_PI * ix / Nx * 2.0);
        }
    for (size_t i = 0; i < Nx; ++i)
        m.dirichlet[0].push_back(i);
    for (size_t i = 0; i < Ny; ++i)
        m.dirichlet[1].push_back(i * Nx + Nx - 1);
    for (size_t i = 0; i < Nx; ++i)
        m.dirichlet[2].push_back(Nx * (Ny - 1) + i);
    for (size_t i = 0; i < Ny; ++i)
        m.dirichlet[3].push_back(i * Nx);
}
// create_rectanglemesh of the ring test (AdvectionPeriodicBC_test.cpp:210-255), including
// the periodic edge ids exactly as that test writes them.
void ringMesh(Mesh& m, size_t Nx, size_t Ny)
{
    m.init(Nx, Ny);
    size_t ii = 0;
    for (size_t iy = 0; iy <= Ny; ++iy)
        for (size_t ix = 0; ix <= Nx; ++ix, ++ii) {
            const double r = R0 + (R1 - R0) * iy / Ny;
            const double p = -2.0 * M_PI * ix / Nx;
            m.vx[ii] = r * std::cos(p);
            m.vy[ii] = r * std::sin(p);
        }
    for (size_t i = 0; i < Nx; ++i)
        m.dirichlet[0].push_back(i);
    for (size_t i = 0; i < Nx; ++i)
        m.dirichlet[2].push_back(Nx * (Ny - 1) + i);
    m.periodic.assign(1, {});
    for (size_t i = 0; i < Ny; ++i)
        m.periodic[0].push_back({ 1, (i + 1) * Nx - 1, i * Nx, i * (Nx + 1) + i });
}

// ---- runtime dispatch over the template degrees
struct TransportBase {
    virtual ~TransportBase() = default;
    virtual void setScheme(int) = 0;
    virtual void setVelocity(const double*, const double*) = 0;
    virtual void prepareCG(int cg, const double*, const double*) = 0;
    virtual void reinit() = 0;
    virtual void step(double, double*) = 0;
    virtual void op(double, const double*, double*) = 0;
    virtual void normalVel(double*, double*) = 0;
};
template <int DG> struct TransportT : TransportBase {
    Transport<DG> t;
    explicit TransportT(const Mesh& m)
        : t(m)
    {
    }
    void setScheme(int s) override { t.scheme = s; }
    void setVelocity(const double* vx, const double* vy) override
    {
        std::memcpy(t.velx.data(), vx, t.velx.size() * sizeof(double));
        std::memcpy(t.vely.data(), vy, t.vely.size() * sizeof(double));
    }
    void prepareCG(int cg, const double* u, const double* v) override
    {
        if (cg == 1) {
            CG2DG<1, DG>(t.m, t.velx.data(), u);
            CG2DG<1, DG>(t.m, t.vely.data(), v);
        } else {
            CG2DG<2, DG>(t.m, t.velx.data(), u);
            CG2DG<2, DG>(t.m, t.vely.data(), v);
        }
        t.reinitnormalvelocity();
    }
    void reinit() override { t.reinitnormalvelocity(); }
    void step(double dt, double* phi) override { t.step(dt, phi); }
    void op(double dt, const double* phi, double* up) override { t.op(dt, phi, up); }
    void normalVel(double* x, double* y) override
    {
        std::memcpy(x, t.nvX.data(), t.nvX.size() * sizeof(double));
        std::memcpy(y, t.nvY.data(), t.nvY.size() * sizeof(double));
    }
};
struct TransportH {
    Mesh mesh;
    std::unique_ptr<TransportBase> t;
    int dg;
};

struct DynBase {
    virtual ~DynBase() = default;
    virtual Params& params() = 0;
    virtual Mesh& mesh() = 0;
    virtual void initialise(size_t, size_t, const double*, const double*, bool) = 0;
    virtual void finishInit() = 0;
    virtual void setData(const std::string&, const double*, int) = 0;
    virtual void getData(const std::string&, double*, int) = 0;
    virtual int getCG(const std::string&, double*) = 0;
    virtual int setCG(const std::string&, const double*) = 0;
    virtual void update(double) = 0;
    virtual void advection(double) = 0;
    virtual void prepareIteration() = 0;
    virtual void beginIteration(double) = 0;
    virtual void pass(int) = 0;
    virtual size_t ncg() = 0;
    virtual void setRheology(int) = 0;
    virtual MEBParams& meb() = 0;
    virtual void mass(double* out) = 0;
    virtual double* budget() = 0;
};
template <int CG, int DG> struct DynT : DynBase {
    Kernel<CG, DG> k;
    Params& params() override { return k.params; }
    Mesh& mesh() override { return k.mesh; }
    void setRheology(int r) override { k.rheology = r; }
    MEBParams& meb() override { return k.meb; }
    void initialise(size_t nx, size_t ny, const double* c, const double* m, bool sph) override
    {
        k.initialise(nx, ny, c, m, sph);
    }
    void finishInit() override { k.finishInit(); }
    void setData(const std::string& n, const double* d, int c) override { k.setData(n, d, c); }
    void getData(const std::string& n, double* o, int c) override { k.getData(n, o, c); }
    int getCG(const std::string& n, double* o) override
    {
        auto* p = k.cgRef(n);
        if (!p)
            return -1;
        std::memcpy(o, p->data(), p->size() * sizeof(double));
        return 0;
    }
    int setCG(const std::string& n, const double* in) override
    {
        auto* p = k.cgRef(n);
        if (!p)
            return -1;
        std::memcpy(p->data(), in, p->size() * sizeof(double));
        return 0;
    }
    void update(double dt) override { k.update(dt); }
    void advection(double dt) override { k.advectionAndLimits(dt); }
    void prepareIteration() override { k.prepareIteration(); }
    void beginIteration(double dt) override
    {
        k.u0 = k.u;
        k.v0 = k.v;
        k.deltaT = dt;
    }
    void pass(int which) override
    {
        switch (which) {
        case 0: k.projectVelocityToStrain(); break;
        case 1: k.stressUpdate(); break;
        case 2: k.stressDivergence(); break;
        case 3: k.updateMomentum(); break;
        case 4: k.applyBoundaries(); break;
        default: k.subcycle(); break;
        }
    }
    size_t ncg() override { return k.ncg(); }
    void mass(double* out) override
    {
        out[0] = meanValue<DG>(k.mesh, k.hice.data());
        out[1] = meanValue<DG>(k.mesh, k.cice.data());
    }
    double* budget() override { return k.budget; }
};

DynBase* makeDyn(int cg, int dg)
{
    if (cg == 1 && dg == 1) return new DynT<1, 1>();
    if (cg == 1 && dg == 3) return new DynT<1, 3>();
    if (cg == 1 && dg == 6) return new DynT<1, 6>();
    if (cg == 2 && dg == 1) return new DynT<2, 1>();
    if (cg == 2 && dg == 3) return new DynT<2, 3>();
    if (cg == 2 && dg == 6) return new DynT<2, 6>();
    throw std::runtime_error("unsupported (cg, dg) combination");
}
TransportBase* makeTransport(const Mesh& m, int dg)
{
    if (dg == 1) return new TransportT<1>(m);
    if (dg == 3) return new TransportT<3>(m);
    if (dg == 6) return new TransportT<6>(m);
    if (dg == 8) return new TransportT<8>(m);
    throw std::runtime_error("unsupported DG degree");
}

template <class F> auto byDG(int dg, F&& f)
{
    switch (dg) {
    case 1: return f(std::integral_constant<int, 1>());
    case 3: return f(std::integral_constant<int, 3>());
    case 6: return f(std::integral_constant<int, 6>());
    case 8: return f(std::integral_constant<int, 8>());
    default: throw std::runtime_error("unsupported DG degree");
    }
}
} // namespace

extern "C" {

const char* ora_last_error() { return g_err.c_str(); }

// ------------------------------------------------------------------ mesh
void* ora_mesh_new() { return new Mesh(); }
void ora_mesh_free(void* m) { delete static_cast<Mesh*>(m); }
int ora_mesh_from_coords(void* m, size_t nx, size_t ny, const double* coords, const double* mask, int spherical)
{
    return guard([&] { static_cast<Mesh*>(m)->fromCoordsAndMask(nx, ny, coords, mask, spherical != 0); });
}
int ora_mesh_rectangle(void* m, double Lx, double Ly, size_t nx, size_t ny, double distort)
{
    return guard([&] { rectangleMesh(*static_cast<Mesh*>(m), Lx, Ly, nx, ny, distort); });
}
int ora_mesh_ring(void* m, size_t nx, size_t ny)
{
    return guard([&] { ringMesh(*static_cast<Mesh*>(m), nx, ny); });
}
int ora_mesh_read_smesh(void* m, const char* text)
{
    return guard([&] { static_cast<Mesh*>(m)->readSmesh(text); });
}
void ora_mesh_size(void* mp, size_t* nx, size_t* ny)
{
    auto* m = static_cast<Mesh*>(mp);
    *nx = m->nx;
    *ny = m->ny;
}
size_t ora_mesh_dirichlet(void* mp, int edge, uint64_t* out)
{
    auto* m = static_cast<Mesh*>(mp);
    if (out)
        for (size_t i = 0; i < m->dirichlet[edge].size(); ++i)
            out[i] = m->dirichlet[edge][i];
    return m->dirichlet[edge].size();
}
size_t ora_mesh_periodic(void* mp, uint64_t* out)
{
    auto* m = static_cast<Mesh*>(mp);
    size_t k = 0;
    for (size_t s = 0; s < m->periodic.size(); ++s)
        for (const auto& p : m->periodic[s]) {
            if (out) {
                out[5 * k] = s;
                for (int j = 0; j < 4; ++j)
                    out[5 * k + 1 + j] = p[j];
            }
            ++k;
        }
    return k;
}
void ora_mesh_landmask(void* mp, uint8_t* out)
{
    auto* m = static_cast<Mesh*>(mp);
    std::memcpy(out, m->landmask.data(), m->nelements);
}
void ora_mesh_vertices(void* mp, double* out)
{
    auto* m = static_cast<Mesh*>(mp);
    for (size_t i = 0; i < m->nnodes; ++i) {
        out[2 * i] = m->vx[i];
        out[2 * i + 1] = m->vy[i];
    }
}

// ------------------------------------------------------------------ interpolation / tools
int ora_function2dg(void* m, int dg, int fn, double* phi)
{
    return guard([&] { byDG(dg, [&](auto D) { function2DG<D.value>(*static_cast<Mesh*>(m), phi, fnById(fn)); return 0; }); });
}
double ora_l2error_dg(void* m, int dg, int fn, const double* phi)
{
    double r = -1;
    guard([&] { r = byDG(dg, [&](auto D) { return L2ErrorFunctionDG<D.value>(*static_cast<Mesh*>(m), phi, fnById(fn)); }); });
    return r;
}
double ora_mean_value(void* m, int dg, const double* phi)
{
    double r = 0;
    guard([&] { r = byDG(dg, [&](auto D) { return meanValue<D.value>(*static_cast<Mesh*>(m), phi); }); });
    return r;
}
int ora_dg2cg(void* mp, int cg, int dg, const double* src, double* dest)
{
    const Mesh& m = *static_cast<Mesh*>(mp);
    return guard([&] {
        byDG(dg, [&](auto D) {
            if (cg == 1)
                DG2CG<1, D.value>(m, dest, src);
            else
                DG2CG<2, D.value>(m, dest, src);
            return 0;
        });
    });
}
int ora_cg2dg(void* mp, int cg, int dg, const double* src, double* dest)
{
    const Mesh& m = *static_cast<Mesh*>(mp);
    return guard([&] {
        byDG(dg, [&](auto D) {
            if (cg == 1)
                CG2DG<1, D.value>(m, dest, src);
            else
                CG2DG<2, D.value>(m, dest, src);
            return 0;
        });
    });
}
int ora_limit(int dg, size_t n, double* phi, double bound, int ismax)
{
    return guard([&] {
        if (dg == 1)
            ismax ? limitMax<1>(phi, n, bound) : limitMin<1>(phi, n, bound);
        else if (dg == 3)
            ismax ? limitMax<3>(phi, n, bound) : limitMin<3>(phi, n, bound);
        else if (dg == 6)
            ismax ? limitMax<6>(phi, n, bound) : limitMin<6>(phi, n, bound);
        else
            throw std::runtime_error("limiter defined for DG 1, 3, 6");
    });
}

// ------------------------------------------------------------------ transport
void* ora_transport_new(void* mesh, int dg)
{
    TransportH* h = nullptr;
    if (guard([&] {
            h = new TransportH();
            h->mesh = *static_cast<Mesh*>(mesh);
            h->dg = dg;
            h->t.reset(makeTransport(h->mesh, dg));
        })) {
        delete h;
        return nullptr;
    }
    return h;
}
void ora_transport_free(void* t) { delete static_cast<TransportH*>(t); }
int ora_transport_set_scheme(void* t, int rk) { return guard([&] { static_cast<TransportH*>(t)->t->setScheme(rk); }); }
int ora_transport_set_velocity(void* t, const double* vx, const double* vy)
{
    return guard([&] { static_cast<TransportH*>(t)->t->setVelocity(vx, vy); });
}
int ora_transport_prepare_cg(void* t, int cg, const double* u, const double* v)
{
    return guard([&] { static_cast<TransportH*>(t)->t->prepareCG(cg, u, v); });
}
int ora_transport_reinit(void* t) { return guard([&] { static_cast<TransportH*>(t)->t->reinit(); }); }
int ora_transport_step(void* t, double dt, double* phi) { return guard([&] { static_cast<TransportH*>(t)->t->step(dt, phi); }); }
int ora_transport_op(void* t, double dt, const double* phi, double* up)
{
    return guard([&] { static_cast<TransportH*>(t)->t->op(dt, phi, up); });
}
int ora_transport_normalvel(void* t, double* x, double* y) { return guard([&] { static_cast<TransportH*>(t)->t->normalVel(x, y); }); }

// The reference advection regression test, run whole (Advection_test.cpp:105-265):
// returns the L2 error, writes the relative mass loss.
double ora_advection_test(int dg, double distort, int it, double* massloss)
{
    double err = -1;
    guard([&] {
        const size_t Nx = 12 * (1 << it), Ny = 13 * (1 << it);
        const int deg = dg == 1 ? 0 : (dg == 3 ? 1 : 2);
        const size_t NT = 100 * (deg + 1) * (deg + 1) * (1 << it);
        Mesh m;
        rectangleMesh(m, kAdv.Lx, kAdv.Ly, Nx, Ny, distort);
        err = byDG(dg, [&](auto D) {
            constexpr int DGv = D.value;
            Transport<DGv> t(m);
            t.scheme = DGv < 3 ? 2 : 3;
            const double dt = kAdv.Lx / NT;
            std::vector<double> phi(m.nelements * DGv);
            function2DG<DGv>(m, phi.data(), smoothBump);
            const double m0 = meanValue<DGv>(m, phi.data());
            function2DG<DGv>(m, t.velx.data(), rotVX);
            function2DG<DGv>(m, t.vely.data(), rotVY);
            for (size_t iter = 1; iter <= NT; ++iter) {
                t.reinitnormalvelocity();
                t.step(dt, phi.data());
            }
            if (massloss)
                *massloss = (m0 - meanValue<DGv>(m, phi.data())) / m0;
            return std::sqrt(L2ErrorFunctionDG<DGv>(m, phi.data(), smoothBump)) / kAdv.Lx;
        });
    });
    return err;
}
// The reference periodic ring test, run whole (AdvectionPeriodicBC_test.cpp:117-276).
double ora_advection_ring_test(int dg, int it, double* massloss)
{
    double err = -1;
    guard([&] {
        const size_t Nx = 32 * (1 << it), Ny = 4 * (1 << it);
        const int deg = dg == 1 ? 0 : (dg == 3 ? 1 : 2);
        const size_t NT = 50 * (deg + 1) * (deg + 1) * (1 << it);
        Mesh m;
        ringMesh(m, Nx, Ny);
        auto run = [&](auto D) {
            constexpr int DGv = D.value;
            Transport<DGv> t(m);
            t.scheme = DGv < 3 ? 2 : 3;
            const double dt = R1 / NT;
            std::vector<double> phi(m.nelements * DGv);
            function2DG<DGv>(m, phi.data(), packman);
            limitMax<DGv>(phi.data(), m.nelements, 1.0);
            const double m0 = meanValue<DGv>(m, phi.data());
            function2DG<DGv>(m, t.velx.data(), ringVX);
            function2DG<DGv>(m, t.vely.data(), ringVY);
            for (size_t iter = 1; iter <= NT; ++iter) {
                t.reinitnormalvelocity();
                t.step(dt, phi.data());
                limitMax<DGv>(phi.data(), m.nelements, 1.0);
                limitMin<DGv>(phi.data(), m.nelements, 0.0);
            }
            if (massloss)
                *massloss = (m0 - meanValue<DGv>(m, phi.data())) / m0;
            return std::sqrt(L2ErrorFunctionDG<DGv>(m, phi.data(), packman)) / R0;
        };
        if (dg == 1)
            err = run(std::integral_constant<int, 1>());
        else if (dg == 3)
            err = run(std::integral_constant<int, 3>());
        else if (dg == 6)
            err = run(std::integral_constant<int, 6>());
        else
            throw std::runtime_error("ring test defined for DG 1, 3, 6");
    });
    return err;
}

// ------------------------------------------------------------------ dynamics kernel
// params: rho_ice, rho_atm, rho_ocean, C_atm, C_ocean, F_atm, F_ocean, fc,
//         ocean_turning_angle, gravity, Pstar, DeltaMin, alpha, beta, nSteps
void* ora_dyn_new(int cg, int dg, const double* p)
{
    DynBase* d = nullptr;
    if (guard([&] { d = makeDyn(cg, dg); }))
        return nullptr;
    if (p) {
        Params& q = d->params();
        q.rho_ice = p[0];
        q.rho_atm = p[1];
        q.rho_ocean = p[2];
        q.C_atm = p[3];
        q.C_ocean = p[4];
        q.F_atm = p[5];
        q.F_ocean = p[6];
        q.fc = p[7];
        q.ocean_turning_angle = p[8];
        q.gravity = p[9];
        q.Pstar = p[10];
        q.DeltaMin = p[11];
        q.alpha = p[12];
        q.beta = p[13];
        q.nSteps = static_cast<int>(p[14]);
    }
    return d;
}
// The BBM kernel (BBMDynamicsKernel.hpp): as ora_dyn_new, plus the MEB parameters
// compaction_param, nu0, young, P0, undamaged_time_relaxation_sigma,
// exponent_relaxation_sigma, exponent_compression_factor, tan_phi, compr_strength, C_lab
// (NULL: MEBParameters defaults). nSteps in p applies to the brittle substeps.
void* ora_dyn_new_bbm(int cg, int dg, const double* p, const double* meb)
{
    auto* d = static_cast<DynBase*>(ora_dyn_new(cg, dg, p));
    if (!d)
        return nullptr;
    d->setRheology(1);
    if (meb) {
        MEBParams& m = d->meb();
        m.compaction_param = meb[0];
        m.nu0 = meb[1];
        m.young = meb[2];
        m.P0 = meb[3];
        m.undamaged_time_relaxation_sigma = meb[4];
        m.exponent_relaxation_sigma = meb[5];
        m.exponent_compression_factor = meb[6];
        m.tan_phi = meb[7];
        m.compr_strength = meb[8];
        m.C_lab = meb[9];
    }
    return d;
}
void ora_dyn_free(void* d) { delete static_cast<DynBase*>(d); }
// spherical != 0: coords are (longitude, latitude) in radians; the mesh is rotated
// (DynamicsKernel.hpp:47-51)
int ora_dyn_initialise(void* d, size_t nx, size_t ny, const double* coords, const double* mask, int spherical)
{
    return guard([&] { static_cast<DynBase*>(d)->initialise(nx, ny, coords, mask, spherical != 0); });
}
// Initialise from an explicit mesh (landmask, Dirichlet and periodic lists as given).
int ora_dyn_initialise_mesh(void* d, void* mesh)
{
    return guard([&] {
        auto* k = static_cast<DynBase*>(d);
        k->mesh() = *static_cast<Mesh*>(mesh);
        k->finishInit();
    });
}
int ora_dyn_set_field(void* d, const char* name, const double* data, int ncomp)
{
    return guard([&] { static_cast<DynBase*>(d)->setData(name, data, ncomp); });
}
int ora_dyn_get_field(void* d, const char* name, double* out, int ncomp)
{
    return guard([&] { static_cast<DynBase*>(d)->getData(name, out, ncomp); });
}
int ora_dyn_get_cg(void* d, const char* name, double* out) { return static_cast<DynBase*>(d)->getCG(name, out); }
int ora_dyn_set_cg(void* d, const char* name, const double* in) { return static_cast<DynBase*>(d)->setCG(name, in); }
size_t ora_dyn_ncg(void* d) { return static_cast<DynBase*>(d)->ncg(); }
int ora_dyn_update(void* d, double dt) { return guard([&] { static_cast<DynBase*>(d)->update(dt); }); }
int ora_dyn_advection(void* d, double dt) { return guard([&] { static_cast<DynBase*>(d)->advection(dt); }); }
int ora_dyn_prepare_iteration(void* d) { return guard([&] { static_cast<DynBase*>(d)->prepareIteration(); }); }
int ora_dyn_begin_iteration(void* d, double dt) { return guard([&] { static_cast<DynBase*>(d)->beginIteration(dt); }); }
// OpenMP threads of the oracle's parallel loops (for timing on a stated thread count)
int ora_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
// Tools::MeanValue of hice and cice (ice volume and area)
int ora_dyn_mass(void* d, double* out) { return guard([&] { static_cast<DynBase*>(d)->mass(out); }); }
// Advection mass budget since the last reset: outflow h, outflow A, limiter h, limiter A
int ora_dyn_budget(void* d, double* out, int reset)
{
    return guard([&] {
        double* b = static_cast<DynBase*>(d)->budget();
        if (out)
            std::copy(b, b + 4, out);
        if (reset)
            std::fill(b, b + 4, 0.0);
    });
}
// which: 0 strain, 1 stress, 2 divergence, 3 momentum, 4 boundaries, 5 whole subcycle
int ora_dyn_pass(void* d, int which, int repeat)
{
    return guard([&] {
        for (int r = 0; r < repeat; ++r)
            static_cast<DynBase*>(d)->pass(which);
    });
}

// Benchmark cyclone wind at the cell lower-left corners (BenchmarkAtmosphere.cpp:29-76),
// with the centre shifted by (xoff, yoff) metres and the vMax parameter given.
void ora_benchmark_wind(size_t nx, size_t ny, double L, double tsec, double xoff, double yoff, double vmax,
    double* uw, double* vw)
{
    const double dx = L / nx, dy = L / ny, tf = tsec / 86400.0;
    const double A = 1e-5, k = 1e-5, alpha = 72. / 180. * M_PI;
    const double ca = std::cos(alpha), sa = std::sin(alpha);
    const double x0 = nx * dx * 0.5 * (1 + tf / 5.0) + xoff;
    const double y0 = ny * dy * 0.5 * (1 + tf / 5.0) + yoff;
    for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
            const double xp = i * dx - x0, yp = j * dy - y0;
            const double scale = A * std::exp(-k * std::hypot(xp, yp));
            uw[i + nx * j] = -scale * vmax * (ca * xp + sa * yp);
            vw[i + nx * j] = -scale * vmax * (-sa * xp + ca * yp);
        }
}
} // extern "C"

extern "C" {
// Oracle basis tables in the layout of the reference's generated headers (row-major);
// returns the number of values written.
int ora_table(const char* name, int a, int b, int c, double* out)
{
    const Basis& B = Basis::get();
    const std::string n(name);
    int k = 0;
    if (n == "GAUSSWEIGHTS") {
        for (int q = 0; q < a * a; ++q)
            out[k++] = B.W(a, q);
    } else if (n == "PSI" || n == "PSIx" || n == "PSIy") {
        for (int i = 0; i < a; ++i)
            for (int q = 0; q < b * b; ++q)
                out[k++] = n == "PSI" ? B.PSI(b, i, q) : (n == "PSIx" ? B.PSIx(b, i, q) : B.PSIy(b, i, q));
    } else if (n == "PSIe") {
        for (int i = 0; i < a; ++i)
            for (int q = 0; q < b; ++q)
                out[k++] = B.PSIe(b, i, q);
    } else if (n == "PSIe_w") {
        for (int q = 0; q < b; ++q)
            for (int i = 0; i < a; ++i)
                out[k++] = B.PSIe_w(b, c, q, i);
    } else if (n == "PSILagrange") {
        for (int i = 0; i < a; ++i)
            for (int q = 0; q < b * b; ++q)
                out[k++] = B.PSIL(b, i, q);
    } else if (n == "PHI" || n == "PHIx" || n == "PHIy") {
        for (int j = 0; j < cgdofs(a); ++j)
            for (int q = 0; q < b * b; ++q)
                out[k++] = n == "PHI" ? B.PHI(a, b, j, q) : (n == "PHIx" ? B.PHIx(a, b, j, q) : B.PHIy(a, b, j, q));
    } else if (n == "PHI1d") {
        for (int j = 0; j <= a; ++j)
            for (int q = 0; q < b; ++q)
                out[k++] = Basis::cg1d(a, j, B.gp[b][q]);
    } else
        return -1;
    return k;
}
} // extern "C"
