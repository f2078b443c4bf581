// extern "C" entry points of include/nsdg_dyn.h. Every call is guarded: C++
// exceptions become a negative status and a per-thread message.
#include "../../include/nsdg_cnn.h"
#include "nsdg_cnn.hpp"
#include "nsdg_dyn.hpp"
#include "nsdg_transport.hpp"

#include <algorithm>
#include <cstring>

using namespace nsdg;

struct nsdg_mesh {
    HostMesh m;
};
struct nsdg_transport {
    TransportHandle h;
};
struct nsdg_dyn {
    std::unique_ptr<DynBase> k;
    int cg, dg;
};
struct nsdg_cnn {
    std::unique_ptr<cnn::Net> net;
    DBuf<uint16_t> x; // features of one timestep (warm start)
};

namespace {
thread_local std::string g_err;

template <class F> int guard(F&& f)
{
    try {
        f();
        return NSDG_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return NSDG_ERR_ARG;
    } catch (...) {
        g_err = "unknown error";
        return NSDG_ERR_ARG;
    }
}
void needDevice()
{
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1)
        throw Error(NSDG_ERR_HIP, "no HIP device available");
}
template <class T> T* nonNull(T* p, const char* what)
{
    if (!p)
        throw Error(NSDG_ERR_ARG, std::string(what) + " is NULL");
    return p;
}
} // namespace

extern "C" {

const char* nsdg_last_error(void) { return g_err.c_str(); }

int nsdg_params_default(nsdg_params* p)
{
    return guard([&] {
        nonNull(p, "params");
        p->rho_ice = 900.0;
        p->rho_atm = 1.3;
        p->rho_ocean = 1026.0;
        p->C_atm = 1.2e-3;
        p->C_ocean = 5.5e-3;
        p->F_atm = p->C_atm * p->rho_atm;
        p->F_ocean = p->C_ocean * p->rho_ocean;
        p->fc = 1.45842e-4;
        p->ocean_turning_angle = 0.0;
        p->gravity = 9.81;
        p->Pstar = 27500.0;
        p->DeltaMin = 2.e-9;
        p->alpha = 1500.0;
        p->beta = 1500.0;
        p->nsteps = 100;
    });
}

int nsdg_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return NSDG_ERR_HIP;
    return n;
}

// ---------------------------------------------------------------- mesh
int nsdg_mesh_create(size_t nx, size_t ny, const double* coords, int spherical, nsdg_mesh** out)
{
    return guard([&] {
        nonNull(out, "out");
        auto m = std::make_unique<nsdg_mesh>();
        m->m.create(nx, ny, coords, spherical != 0);
        *out = m.release();
    });
}
int nsdg_mesh_set_landmask(nsdg_mesh* m, const double* mask)
{
    return guard([&] { nonNull(m, "mesh")->m.setLandmask(mask); });
}
int nsdg_mesh_dirichlet_from_mask(nsdg_mesh* m)
{
    return guard([&] { nonNull(m, "mesh")->m.dirichletFromMask(); });
}
int nsdg_mesh_dirichlet_from_edge(nsdg_mesh* m, int edge)
{
    return guard([&] { nonNull(m, "mesh")->m.dirichletFromEdge(edge); });
}
int nsdg_mesh_set_dirichlet(nsdg_mesh* m, int edge, const uint64_t* ids, size_t n)
{
    return guard([&] {
        nonNull(m, "mesh");
        require(edge >= 0 && edge < 4, "edge must be 0..3");
        require(ids != nullptr || n == 0, "ids is NULL");
        for (size_t i = 0; i < n; ++i)
            require(ids[i] < m->m.ne(), "Dirichlet element id out of range");
        m->m.dirichlet[edge].assign(ids, ids + n);
    });
}
int nsdg_mesh_add_periodic(nsdg_mesh* m, const uint64_t* entries, size_t n)
{
    return guard([&] {
        nonNull(m, "mesh");
        require(entries != nullptr || n == 0, "entries is NULL");
        m->m.addPeriodic(entries, n);
        m->m.validate();
    });
}
int nsdg_mesh_get_dirichlet(const nsdg_mesh* m, int edge, uint64_t* out, size_t* n)
{
    return guard([&] {
        nonNull(m, "mesh");
        nonNull(n, "n");
        require(edge >= 0 && edge < 4, "edge must be 0..3");
        const auto& d = m->m.dirichlet[edge];
        if (out)
            std::memcpy(out, d.data(), d.size() * sizeof(uint64_t));
        *n = d.size();
    });
}
int nsdg_mesh_destroy(nsdg_mesh* m)
{
    delete m;
    return NSDG_OK;
}

// ---------------------------------------------------------------- transport
int nsdg_transport_create(const nsdg_mesh* m, int dg, nsdg_transport** out)
{
    return guard([&] {
        nonNull(m, "mesh");
        nonNull(out, "out");
        needDevice();
        auto t = std::make_unique<nsdg_transport>();
        NSDG_HIP(hipStreamCreateWithFlags(&t->h.stream, hipStreamNonBlocking));
        t->h.mesh.build(m->m, t->h.stream);
        t->h.t = makeTransport(dg, t->h.mesh, t->h.stream, t->h.prof);
        t->h.field.alloc(t->h.mesh.ne * dg);
        NSDG_HIP(hipStreamSynchronize(t->h.stream));
        *out = t.release();
    });
}
int nsdg_transport_set_scheme(nsdg_transport* t, const char* scheme)
{
    return guard([&] {
        nonNull(t, "transport");
        const std::string s(nonNull(scheme, "scheme"));
        require(s == "rk1" || s == "rk2" || s == "rk3", "scheme must be rk1, rk2 or rk3");
        t->h.t->setScheme(s[2] - '0');
    });
}
int nsdg_transport_set_velocity(nsdg_transport* t, const double* vx, const double* vy)
{
    return guard([&] { nonNull(t, "transport")->h.t->setVelocityHost(nonNull(vx, "vx"), nonNull(vy, "vy")); });
}
int nsdg_transport_prepare_advection(nsdg_transport* t, int cg, const double* u, const double* v)
{
    return guard([&] { nonNull(t, "transport")->h.t->prepareAdvectionHost(cg, nonNull(u, "u"), nonNull(v, "v")); });
}
int nsdg_transport_reinit_normal_velocity(nsdg_transport* t)
{
    return guard([&] {
        nonNull(t, "transport")->h.t->reinitNormalVelocity();
        NSDG_HIP(hipStreamSynchronize(t->h.stream));
    });
}
int nsdg_transport_run(nsdg_transport* t, double dt, double* phi, int nsteps, int limit_flags)
{
    return guard([&] {
        nonNull(t, "transport");
        nonNull(phi, "phi");
        require(nsteps >= 0, "nsteps must be >= 0");
        auto& h = t->h;
        const int dg = h.t->dg();
        DBuf<double> scratch;
        uploadSoA(scratch, phi, h.mesh.ne, dg, h.field.p, h.stream);
        for (int k = 0; k < nsteps; ++k) {
            h.t->reinitNormalVelocity();
            h.t->step(dt, h.field.p);
            if (limit_flags & 3)
                h.t->limit(h.field.p, (limit_flags & 1) != 0, 1.0, (limit_flags & 2) != 0, 0.0);
        }
        downloadSoA(scratch, h.field.p, h.mesh.ne, dg, phi, h.stream);
    });
}
int nsdg_transport_step(nsdg_transport* t, double dt, double* phi)
{
    return guard([&] {
        nonNull(t, "transport");
        nonNull(phi, "phi");
        auto& h = t->h;
        const int dg = h.t->dg();
        DBuf<double> scratch;
        uploadSoA(scratch, phi, h.mesh.ne, dg, h.field.p, h.stream);
        h.t->step(dt, h.field.p);
        downloadSoA(scratch, h.field.p, h.mesh.ne, dg, phi, h.stream);
    });
}
int nsdg_transport_destroy(nsdg_transport* t)
{
    delete t;
    return NSDG_OK;
}

int nsdg_limit(int dg, size_t n, double* phi, double bound, int is_max)
{
    return guard([&] {
        nonNull(phi, "phi");
        require(dg == 1 || dg == 3 || dg == 6, "limiters are defined for dg 1, 3, 6");
        require(n >= 1, "n must be >= 1");
        needDevice();
        TransportHandle h;
        NSDG_HIP(hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking));
        // A strip mesh of at least 2 x 2 elements carries the field in its first n elements.
        const size_t nxm = n < 2 ? 2 : n;
        HostMesh m;
        std::vector<double> coords(2 * (nxm + 1) * 3, 0.0);
        for (size_t j = 0; j < 3; ++j)
            for (size_t i = 0; i <= nxm; ++i) {
                coords[2 * (i + (nxm + 1) * j)] = static_cast<double>(i);
                coords[2 * (i + (nxm + 1) * j) + 1] = static_cast<double>(j);
            }
        m.create(nxm, 2, coords.data());
        h.mesh.build(m, h.stream);
        h.t = makeTransport(dg, h.mesh, h.stream, h.prof);
        DBuf<double> f, scratch;
        f.alloc(h.mesh.ne * dg);
        deviceFill(f.p, f.n, 0.0, h.stream);
        std::vector<double> full(h.mesh.ne * dg, 0.0);
        std::memcpy(full.data(), phi, n * dg * sizeof(double));
        uploadSoA(scratch, full.data(), h.mesh.ne, dg, f.p, h.stream);
        h.t->limit(f.p, is_max != 0, bound, is_max == 0, bound);
        downloadSoA(scratch, f.p, h.mesh.ne, dg, full.data(), h.stream);
        std::memcpy(phi, full.data(), n * dg * sizeof(double));
    });
}

// ---------------------------------------------------------------- dynamics
int nsdg_dyn_create(int cg_degree, int dg_adv, const nsdg_params* params, nsdg_dyn** out)
{
    return guard([&] {
        nonNull(out, "out");
        needDevice();
        nsdg_params p;
        if (params)
            p = *params;
        else
            nsdg_params_default(&p);
        require(p.alpha > 0 && p.beta > 0 && p.nsteps >= 0, "alpha, beta must be > 0 and nsteps >= 0");
        auto d = std::make_unique<nsdg_dyn>();
        d->k = makeDyn(cg_degree, dg_adv, p);
        d->cg = cg_degree;
        d->dg = dg_adv;
        *out = d.release();
    });
}
int nsdg_meb_params_default(nsdg_meb_params* p)
{
    return guard([&] {
        nonNull(p, "params");
        p->compaction_param = -20.0;
        p->nu0 = 1.0 / 3.0;
        p->young = 5.96e8;
        p->P0 = 10.e3;
        p->undamaged_time_relaxation_sigma = 1e7;
        p->exponent_relaxation_sigma = 5.0;
        p->exponent_compression_factor = 1.5;
        p->tan_phi = 0.7;
        p->compr_strength = 1e10;
        p->C_lab = 2.0e6;
    });
}
int nsdg_bbm_create(int cg_degree, int dg_adv, const nsdg_params* params, const nsdg_meb_params* meb, nsdg_dyn** out)
{
    return guard([&] {
        nonNull(out, "out");
        needDevice();
        nsdg_params p;
        if (params)
            p = *params;
        else
            nsdg_params_default(&p);
        require(p.nsteps >= 1, "nsteps must be >= 1");
        nsdg_meb_params m;
        if (meb)
            m = *meb;
        else
            nsdg_meb_params_default(&m);
        auto d = std::make_unique<nsdg_dyn>();
        d->k = makeDyn(cg_degree, dg_adv, p);
        d->k->setBrittle(m);
        d->cg = cg_degree;
        d->dg = dg_adv;
        *out = d.release();
    });
}
int nsdg_dyn_initialise(nsdg_dyn* d, size_t nx, size_t ny, const double* coords, int spherical, const double* mask)
{
    return guard([&] {
        nonNull(d, "dyn");
        HostMesh m;
        m.create(nx, ny, coords, spherical != 0);
        m.setLandmask(mask);
        m.dirichletFromMask();
        for (int e = 0; e < 4; ++e)
            m.dirichletFromEdge(e);
        d->k->initialise(m);
    });
}
int nsdg_dyn_initialise_mesh(nsdg_dyn* d, const nsdg_mesh* m)
{
    return guard([&] { nonNull(d, "dyn")->k->initialise(nonNull(m, "mesh")->m); });
}
int nsdg_dyn_set_field(nsdg_dyn* d, const char* name, const double* data, int ncomp)
{
    return guard([&] { nonNull(d, "dyn")->k->setField(nonNull(name, "name"), data, ncomp); });
}
int nsdg_dyn_get_field(nsdg_dyn* d, const char* name, double* out, int ncomp)
{
    return guard([&] { nonNull(d, "dyn")->k->getField(nonNull(name, "name"), out, ncomp); });
}
int nsdg_dyn_get_cg(nsdg_dyn* d, const char* name, double* out)
{
    return guard([&] { nonNull(d, "dyn")->k->getCG(nonNull(name, "name"), nonNull(out, "out")); });
}
int nsdg_dyn_set_cg(nsdg_dyn* d, const char* name, const double* in)
{
    return guard([&] { nonNull(d, "dyn")->k->setCG(nonNull(name, "name"), nonNull(in, "in")); });
}
size_t nsdg_dyn_cg_nodes(const nsdg_dyn* d) { return d ? d->k->cgNodes() : 0; }
int nsdg_dyn_update(nsdg_dyn* d, double dt)
{
    return guard([&] { nonNull(d, "dyn")->k->update(dt); });
}
int nsdg_dyn_advect(nsdg_dyn* d, double dt)
{
    return guard([&] { nonNull(d, "dyn")->k->advect(dt); });
}
int nsdg_dyn_prepare_iteration(nsdg_dyn* d)
{
    return guard([&] { nonNull(d, "dyn")->k->prepareIteration(); });
}
int nsdg_dyn_begin_iteration(nsdg_dyn* d, double dt)
{
    return guard([&] { nonNull(d, "dyn")->k->beginIteration(dt); });
}
int nsdg_dyn_strain(nsdg_dyn* d)
{
    return guard([&] { nonNull(d, "dyn")->k->strain(); });
}
int nsdg_dyn_stress_update(nsdg_dyn* d, int repeat)
{
    return guard([&] {
        require(repeat >= 0, "repeat must be >= 0");
        nonNull(d, "dyn")->k->stress(repeat);
    });
}
int nsdg_dyn_divergence(nsdg_dyn* d)
{
    return guard([&] { nonNull(d, "dyn")->k->divergence(); });
}
int nsdg_dyn_momentum(nsdg_dyn* d)
{
    return guard([&] { nonNull(d, "dyn")->k->momentum(); });
}
int nsdg_dyn_apply_boundaries(nsdg_dyn* d)
{
    return guard([&] { nonNull(d, "dyn")->k->boundaries(); });
}
int nsdg_dyn_subcycles(nsdg_dyn* d, int n)
{
    return guard([&] {
        require(n >= 0, "n must be >= 0");
        nonNull(d, "dyn")->k->subcycles(n);
    });
}
int nsdg_dyn_set_precision(nsdg_dyn* d, int precision)
{
    return guard([&] { nonNull(d, "dyn")->k->setPrecision(precision); });
}
int nsdg_dyn_set_pairing(nsdg_dyn* d, int mode)
{
    return guard([&] { nonNull(d, "dyn")->k->setPairing(mode); });
}
int nsdg_dyn_benchmark_wind(nsdg_dyn* d, double L, double t, double xoff, double yoff, double vmax)
{
    return guard([&] { nonNull(d, "dyn")->k->benchmarkWind(L, t, xoff, yoff, vmax); });
}
int nsdg_dyn_storm_wind(nsdg_dyn* d, double L, double t, const double* xoff, const double* yoff, const double* vmax,
    int n)
{
    return guard([&] { nonNull(d, "dyn")->k->stormWind(L, t, xoff, yoff, vmax, n); });
}
int nsdg_dyn_reset_dg0(nsdg_dyn* d, const char* name)
{
    return guard([&] { nonNull(d, "dyn")->k->resetDG0(nonNull(name, "name")); });
}
int nsdg_dyn_mass(nsdg_dyn* d, double* vol, double* area)
{
    return guard([&] { nonNull(d, "dyn")->k->mass(vol, area); });
}
int nsdg_dyn_mass_noalloc(nsdg_dyn* d, double* vol, double* area)
{
    return guard([&] { nonNull(d, "dyn")->k->massNoAlloc(vol, area); });
}
int nsdg_dyn_write_restart(nsdg_dyn* d, const char* path, double time_seconds)
{
    return guard([&] { writeRestart(*nonNull(d, "dyn")->k, path, time_seconds); });
}
int nsdg_dyn_read_restart(nsdg_dyn* d, const char* path, double* time_seconds)
{
    return guard([&] { readRestart(*nonNull(d, "dyn")->k, path, time_seconds); });
}
int nsdg_dyn_set_budget(nsdg_dyn* d, int on)
{
    return guard([&] { nonNull(d, "dyn")->k->setBudget(on != 0); });
}
int nsdg_dyn_budget(nsdg_dyn* d, double* out, int reset)
{
    return guard([&] { nonNull(d, "dyn")->k->budget(nonNull(out, "out"), reset != 0); });
}
int nsdg_dyn_residual(nsdg_dyn* d, double* r, double* norm_r, double* norm_forcing)
{
    return guard([&] { nonNull(d, "dyn")->k->residual(r, norm_r, norm_forcing); });
}
int nsdg_dyn_stress_from_velocity(nsdg_dyn* d)
{
    return guard([&] { nonNull(d, "dyn")->k->stressFromVelocity(); });
}
int nsdg_dyn_synchronize(nsdg_dyn* d)
{
    return guard([&] { nonNull(d, "dyn")->k->synchronize(); });
}
int nsdg_dyn_save_state(nsdg_dyn* d, void* buf, size_t* size)
{
    return guard([&] {
        nonNull(d, "dyn");
        nonNull(size, "size");
        const std::vector<char> blob = d->k->saveState();
        if (buf) {
            require(*size >= blob.size(), "buffer too small for the state blob");
            std::memcpy(buf, blob.data(), blob.size());
        }
        *size = blob.size();
    });
}
int nsdg_dyn_load_state(nsdg_dyn* d, const void* buf, size_t size)
{
    return guard([&] { nonNull(d, "dyn")->k->loadState(static_cast<const char*>(nonNull(buf, "buf")), size); });
}
int nsdg_dyn_set_profiling(nsdg_dyn* d, int on)
{
    return guard([&] {
        auto& p = nonNull(d, "dyn")->k->profiler();
        d->k->synchronize();
        p.on = on != 0;
        if (p.on)
            p.stats.clear();
    });
}
int nsdg_dyn_kernel_stats(nsdg_dyn* d, const char* kernel, int* count, double* total_ms)
{
    return guard([&] {
        nonNull(d, "dyn")->k->synchronize();
        auto& st = d->k->profiler().stats;
        auto it = st.find(nonNull(kernel, "kernel"));
        if (count)
            *count = it == st.end() ? 0 : it->second.first;
        if (total_ms)
            *total_ms = it == st.end() ? 0.0 : it->second.second;
    });
}
int nsdg_dyn_destroy(nsdg_dyn* d)
{
    delete d;
    return NSDG_OK;
}

} // extern "C"

// ---------------------------------------------------------------- learned warm start
extern "C" {

nsdg_cnn* nsdg_cnn_create(uint64_t seed)
{
    nsdg_cnn* h = nullptr;
    if (guard([&] {
            needDevice();
            auto p = std::make_unique<nsdg_cnn>();
            p->net = std::make_unique<cnn::Net>(seed);
            h = p.release();
        }))
        return nullptr;
    return h;
}
int nsdg_cnn_destroy(nsdg_cnn* net)
{
    return guard([&] { delete net; });
}
int nsdg_cnn_param_count(nsdg_cnn* net, size_t* count)
{
    return guard([&] {
        nonNull(net, "net");
        *nonNull(count, "count") = cnn::NPARAM;
    });
}
int nsdg_cnn_get_params(nsdg_cnn* net, float* host)
{
    return guard([&] {
        auto& n = *nonNull(net, "net")->net;
        n.params.download(nonNull(host, "host"), cnn::NPARAM, n.s);
        n.sync();
    });
}
int nsdg_cnn_set_params(nsdg_cnn* net, const float* host)
{
    return guard([&] {
        auto& n = *nonNull(net, "net")->net;
        n.params.upload(nonNull(host, "host"), cnn::NPARAM, n.s);
        n.pack();
        n.sync();
    });
}
int nsdg_cnn_forward(nsdg_cnn* net, const uint16_t* x, int n, int ny, int nx, float* y)
{
    return guard([&] {
        auto& nt = *nonNull(net, "net")->net;
        nt.forward(x, n, ny, nx, nonNull(y, "y"), false);
        nt.sync();
    });
}
int nsdg_cnn_loss_grad(nsdg_cnn* net, const uint16_t* x, const float* t, const float* w, int n, int ny, int nx,
    double* loss, float* grad_host)
{
    return guard([&] {
        auto& nt = *nonNull(net, "net")->net;
        const double l = nt.lossGrad(x, t, w, n, ny, nx);
        if (loss)
            *loss = l;
        if (grad_host) {
            nt.grad.download(grad_host, cnn::NPARAM, nt.s);
            nt.sync();
        }
    });
}
int nsdg_cnn_adam(nsdg_cnn* net, float lr)
{
    return guard([&] {
        require(lr > 0.0f, "learning rate must be positive");
        nonNull(net, "net")->net->adam(lr);
    });
}
int nsdg_cnn_set_profiling(nsdg_cnn* net, int on)
{
    return guard([&] {
        auto& nt = *nonNull(net, "net")->net;
        nt.sync();
        nt.prof.on = on != 0;
        if (nt.prof.on)
            nt.prof.stats.clear();
    });
}
int nsdg_cnn_kernel_stats(nsdg_cnn* net, const char* kernel, int* count, double* total_ms)
{
    return guard([&] {
        auto& nt = *nonNull(net, "net")->net;
        nt.sync();
        auto it = nt.prof.stats.find(nonNull(kernel, "kernel"));
        if (count)
            *count = it == nt.prof.stats.end() ? 0 : it->second.first;
        if (total_ms)
            *total_ms = it == nt.prof.stats.end() ? 0.0 : it->second.second;
    });
}

int nsdg_device_alloc(size_t bytes, void** ptr)
{
    return guard([&] {
        needDevice();
        nonNull(ptr, "ptr");
        *ptr = nullptr;
        if (bytes)
            NSDG_HIP(hipMalloc(ptr, bytes));
    });
}
int nsdg_device_free(void* ptr)
{
    return guard([&] {
        if (ptr)
            NSDG_HIP(hipFree(ptr));
    });
}
int nsdg_device_upload(void* dst, const void* src, size_t bytes)
{
    return guard([&] {
        if (bytes)
            NSDG_HIP(hipMemcpy(nonNull(dst, "dst"), nonNull(src, "src"), bytes, hipMemcpyHostToDevice));
    });
}
int nsdg_device_download(void* dst, const void* src, size_t bytes)
{
    return guard([&] {
        if (bytes)
            NSDG_HIP(hipMemcpy(nonNull(dst, "dst"), nonNull(src, "src"), bytes, hipMemcpyDeviceToHost));
    });
}
int nsdg_device_gather(void* dst, const void* src, const int64_t* index, int n, size_t item_bytes)
{
    return guard([&] {
        nonNull(dst, "dst");
        nonNull(src, "src");
        require(n >= 0, "n must be >= 0");
        if (n > 0)
            nonNull(index, "index");
        for (int i = 0; i < n; ++i) {
            require(index[i] >= 0, "negative gather index");
            NSDG_HIP(hipMemcpyAsync(static_cast<char*>(dst) + static_cast<size_t>(i) * item_bytes,
                static_cast<const char*>(src) + static_cast<size_t>(index[i]) * item_bytes, item_bytes,
                hipMemcpyDeviceToDevice, nullptr));
        }
        NSDG_HIP(hipStreamSynchronize(nullptr));
    });
}

int nsdg_dyn_features(nsdg_dyn* d, uint16_t* x)
{
    return guard([&] { nonNull(d, "dyn")->k->features(nonNull(x, "x")); });
}
int nsdg_dyn_velocity_target(nsdg_dyn* d, float* t, float* w)
{
    return guard([&] { nonNull(d, "dyn")->k->velocityTarget(nonNull(t, "t"), w); });
}
int nsdg_dyn_apply_velocity(nsdg_dyn* d, const float* y)
{
    return guard([&] { nonNull(d, "dyn")->k->applyVelocity(nonNull(y, "y")); });
}
int nsdg_dyn_warm_start(nsdg_dyn* d, nsdg_cnn* net, double* r_predicted, double* r_cold, int* used)
{
    return guard([&] {
        auto& k = *nonNull(d, "dyn")->k;
        auto& h = *nonNull(net, "net");
        const int nx = k.gridNx(), ny = k.gridNy();
        require(nx % 16 == 0, "the warm start needs nx to be a multiple of 16", NSDG_ERR_UNSUPPORTED);
        const size_t ne = static_cast<size_t>(nx) * ny;
        if (h.x.n < ne * cnn::C)
            h.x.alloc(ne * cnn::C);
        double rc = 0.0, rp = 0.0;
        k.residual(&rc, nullptr, nullptr);
        k.saveStart();
        k.features(h.x.p);
        h.net->forward(h.x.p, 1, ny, nx, nullptr, false);
        h.net->sync();
        k.applyVelocity(h.net->out.p);
        k.stressFromVelocity();
        k.residual(&rp, nullptr, nullptr);
        const bool keep = rp <= rc;
        if (!keep)
            k.restoreStart();
        if (r_predicted)
            *r_predicted = rp;
        if (r_cold)
            *r_cold = rc;
        if (used)
            *used = keep ? 1 : 0;
    });
}
int nsdg_dyn_subcycles_to_residual(nsdg_dyn* d, double r_target, int max_subcycles, int check_every, double r_start,
    int* done, double* r_final)
{
    return guard([&] {
        nonNull(d, "dyn")->k->subcyclesToResidual(r_target, max_subcycles, check_every, r_start, done, r_final);
    });
}

} // extern "C"
