/*
 * nsdg_dyn.h — C ABI of the MI355X (gfx950) implementation of the neXtSIM_DG
 * dynamics hot path: the mEVP momentum/rheology subcycle and the DG advection of
 * ice thickness and concentration on the structured parametric quadrilateral mesh.
 *
 * Every entry point returns 0 on success and a negative status on failure; the
 * message is then available from nsdg_last_error() (per calling thread). No C++
 * exception crosses this boundary. Host buffers use the reference's ModelArray
 * layout: FP64, element (ix, iy) at row ix + nx*iy with its components
 * contiguous; CG node (ix, iy) at ix + (CG*nx + 1)*iy; vertex (ix, iy) at
 * ix + (nx+1)*iy with its (x, y) pair contiguous. Device state stays resident
 * in HBM between calls. A handle is not re-entrant (one caller thread at a
 * time), as the reference kernel is not.
 *
 * Reference interfaces replaced (file:line in nextsimhub/nextsimdg):
 *   ParametricMesh           dynamics/src/include/ParametricMesh.hpp:42-442
 *   DGTransport<DG>          dynamics/src/include/DGTransport.hpp:27-152
 *   LimitMax / LimitMin      dynamics/src/include/dgLimit.hpp:16-84
 *   DynamicsKernel<DGa,DGs>  dynamics/src/include/DynamicsKernel.hpp:37-229
 *   CGDynamicsKernel<DGa>    dynamics/src/include/CGDynamicsKernel.hpp:26-86
 *   VPCGDynamicsKernel<DGa>  dynamics/src/include/VPCGDynamicsKernel.hpp:22-176
 *   MEVPDynamicsKernel<DGa>  dynamics/src/include/MEVPDynamicsKernel.hpp:18-37
 *   StressUpdateStep         dynamics/src/include/StressUpdateStep.hpp:20-33,
 *                            MEVPStressUpdateStep.hpp:30-118
 */
#ifndef NSDG_DYN_H
#define NSDG_DYN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSDG_OK 0
#define NSDG_ERR_ARG (-1)     /* invalid argument, unknown name, bad size */
#define NSDG_ERR_STATE (-2)   /* call out of order (e.g. update before initialise) */
#define NSDG_ERR_HIP (-3)     /* HIP runtime error */
#define NSDG_ERR_UNSUPPORTED (-4)

/* Physical and rheology parameters: DynamicsParameters (DynamicsParameters.hpp:14-51)
 * and VPParameters (VPParameters.hpp:15-25), plus the mEVP alpha, beta and subcycle
 * count that the reference hard-codes (VPCGDynamicsKernel.hpp:125-126,
 * MEVPStressUpdateStep.hpp:126-127, DynamicsKernel.hpp:187). */
typedef struct nsdg_params {
    double rho_ice, rho_atm, rho_ocean;
    double C_atm, C_ocean, F_atm, F_ocean;
    double fc, ocean_turning_angle, gravity;
    double Pstar, DeltaMin;
    double alpha, beta;
    int32_t nsteps;
} nsdg_params;

/* Arithmetic of the element (rheology) pass. FP64 everywhere is the reference path.
 * NSDG_PREC_FP32 runs strain rates, Delta/viscosities, stress and the basis
 * contractions in FP32 (the relaxed stress sum compensated, the stored stress rounded
 * stochastically), keeping the velocity update, lumped-mass inversion and conservation
 * sums in FP64. */
#define NSDG_PREC_FP64 0
#define NSDG_PREC_FP32 1

typedef struct nsdg_mesh nsdg_mesh;
typedef struct nsdg_transport nsdg_transport;
typedef struct nsdg_dyn nsdg_dyn;

const char* nsdg_last_error(void);
int nsdg_params_default(nsdg_params* p);
/* Number of HIP devices visible; <0 when the HIP runtime is unusable. */
int nsdg_device_count(void);

/* ---------------------------------------------------------------- ParametricMesh
 * coordinatesFromModelArray (ParametricMesh.cpp:198-210): coords is (nx+1)(ny+1) x 2.
 * spherical != 0: (longitude, latitude) in radians, rotated by RotatePoleToGreenland
 * (ParametricMesh.hpp:137-157) as DynamicsKernel::initialise does (DynamicsKernel.hpp:47-51);
 * the operators then carry the metric (SphericalTools, ParametricMap.cpp:49-345). */
int nsdg_mesh_create(size_t nx, size_t ny, const double* coords, int spherical, nsdg_mesh** out);
/* landmaskFromModelArray (ParametricMesh.cpp:217-223): element is ocean iff mask == 1. */
int nsdg_mesh_set_landmask(nsdg_mesh* m, const double* mask);
/* dirichletFromMask (ParametricMesh.cpp:228-252). */
int nsdg_mesh_dirichlet_from_mask(nsdg_mesh* m);
/* dirichletFromEdge (ParametricMesh.cpp:259-272); edge 0 bottom, 1 right, 2 top, 3 left. */
int nsdg_mesh_dirichlet_from_edge(nsdg_mesh* m, int edge);
/* Replace one Dirichlet list (ParametricMesh::dirichlet[edge], ParametricMesh.hpp:70). */
int nsdg_mesh_set_dirichlet(nsdg_mesh* m, int edge, const uint64_t* element_ids, size_t n);
/* Append one periodic segment (ParametricMesh::periodic, ParametricMesh.hpp:79): n entries
 * of 4 values {type (0 X-edge, 1 Y-edge), element left/bottom, element right/top, edge id}. */
int nsdg_mesh_add_periodic(nsdg_mesh* m, const uint64_t* entries, size_t n);
/* Copy out a Dirichlet list; *n is set to its length (out may be NULL). */
int nsdg_mesh_get_dirichlet(const nsdg_mesh* m, int edge, uint64_t* out, size_t* n);
int nsdg_mesh_destroy(nsdg_mesh* m);

/* ---------------------------------------------------------------- DGTransport<DG>
 * DGTransport(const ParametricMesh&) (DGTransport.hpp:71-96); dg in {1, 3, 6, 8}.
 * The mesh is copied to the device; the nsdg_mesh may be destroyed afterwards. */
int nsdg_transport_create(const nsdg_mesh* m, int dg, nsdg_transport** out);
/* settimesteppingscheme (DGTransport.hpp:106-110): "rk1", "rk2" or "rk3". */
int nsdg_transport_set_scheme(nsdg_transport* t, const char* scheme);
/* GetVx() = vx; GetVy() = vy (DGTransport.hpp:100-103): N_e x DG coefficients. */
int nsdg_transport_set_velocity(nsdg_transport* t, const double* vx, const double* vy);
/* prepareAdvection<CG>(u, v) (DGTransport.cpp:261-268): CG2DG of the CG velocity, then
 * reinitnormalvelocity. cg in {1, 2}. */
int nsdg_transport_prepare_advection(nsdg_transport* t, int cg, const double* u, const double* v);
/* reinitnormalvelocity (DGTransport.cpp:158-252). */
int nsdg_transport_reinit_normal_velocity(nsdg_transport* t);
/* step(dt, phi) (DGTransport.cpp:554-566) on a host N_e x DG field, in place. */
int nsdg_transport_step(nsdg_transport* t, double dt, double* phi);
/* nsteps x { reinitnormalvelocity; step(dt, phi); optional limiters } with phi resident
 * on the device between steps. limit_flags bit 0: LimitMax(phi, 1), bit 1: LimitMin(phi, 0),
 * in that order (the loop of AdvectionPeriodicBC_test.cpp:186-198). */
int nsdg_transport_run(nsdg_transport* t, double dt, double* phi, int nsteps, int limit_flags);
int nsdg_transport_destroy(nsdg_transport* t);

/* LimitMax / LimitMin (dgLimit.hpp:16-84) on the GPU for a host N x dg field, dg in {1,3,6}. */
int nsdg_limit(int dg, size_t n, double* phi, double bound, int is_max);

/* ---------------------------------------------------------------- MEVPDynamicsKernel
 * MEVPDynamicsKernel<DG>(params) (MEVPDynamicsKernel.hpp:18-37) for the two builds the
 * reference's CMake produces: (cg_degree 1, dg_adv 3, 3 stress dofs) and
 * (cg_degree 2, dg_adv 6, 8 stress dofs). params may be NULL (defaults). */
int nsdg_dyn_create(int cg_degree, int dg_adv, const nsdg_params* params, nsdg_dyn** out);
/* Brittle rheology parameters, MEBParameters (MEBParameters.hpp:15-70). */
typedef struct nsdg_meb_params {
    double compaction_param, nu0, young, P0, undamaged_time_relaxation_sigma;
    double exponent_relaxation_sigma, exponent_compression_factor, tan_phi, compr_strength, C_lab;
} nsdg_meb_params;
int nsdg_meb_params_default(nsdg_meb_params* p);
/* BBMDynamicsKernel<DG> (BBMDynamicsKernel.hpp:18-45 over BrittleCGDynamicsKernel.hpp:23-256),
 * behind the same handle and calls as nsdg_dyn_create's kernel. update(dt) advects h and A
 * with the velocity averaged over the previous step's substeps, transports the stress
 * (DG 8, RK2) and the damage (then limits it to [1e-12, 1]), and runs params->nsteps
 * brittle substeps of dt / nsteps (strain, BBM stress and damage update, divergence,
 * semi-implicit momentum, Dirichlet zero). Extra fields: "damage" (DG, set_field /
 * get_field; starts undamaged, 1), "avgU", "avgV" (get_cg). The ice-ocean stress comes from
 * the averaged velocity. Any mesh: meshes without periodic segments run one kernel per
 * substep (non-uniform ones from per-element geometry, element sizes and damage mass
 * matrices), periodic meshes the split passes; spherical meshes are projected as the
 * reference projects them, with the cos-latitude mass matrix against weights without
 * cos(latitude). */
int nsdg_bbm_create(int cg_degree, int dg_adv, const nsdg_params* params, const nsdg_meb_params* meb, nsdg_dyn** out);
/* initialise(coords, isSpherical, mask) (DynamicsKernel.hpp:44-75,
 * CGDynamicsKernel.cpp:23-51). Builds the mesh, its Dirichlet lists, the transport and
 * momentum operators on the device. Stress and strain start at zero. With spherical != 0
 * the coordinates are (longitude, latitude) in radians (the module converts degrees,
 * MEVPDynamics.cpp:46-48). Meshes need fewer than 2^29 elements and 2^29 CG nodes (the
 * kernels address fields with 32-bit byte offsets: up to a 11585^2 CG2 mesh, about 210 GB
 * of FP64 state); larger ones are refused with NSDG_ERR_ARG. Non-uniform and spherical
 * meshes run the FP64 subcycle from per-element geometry tables; the FP32 variant stays
 * with uniform meshes. */
int nsdg_dyn_initialise(nsdg_dyn* d, size_t nx, size_t ny, const double* coords, int spherical,
    const double* mask);
/* Initialise from an explicitly built mesh (landmask, Dirichlet and periodic lists as set). */
int nsdg_dyn_initialise_mesh(nsdg_dyn* d, const nsdg_mesh* m);
/* setData(name, ModelArray) (CGDynamicsKernel.cpp:53-90, DynamicsKernel.hpp:93-110).
 * Names: hice, cice, u, v, uwind, vwind, uocean, vocean, ssh, s11, s12, s22; any other
 * name is stored as a passive field. ncomp = 1 gives DG0 semantics (ma2dg zeroes the
 * higher components, DGModelArray.hpp:20-32); ncomp = DG (or 8/3 for stress) sets all. */
int nsdg_dyn_set_field(nsdg_dyn* d, const char* name, const double* data, int ncomp);
/* getDG0Data (ncomp = 1) / getDGData (ncomp = DG or the stress dofs)
 * (CGDynamicsKernel.cpp:92-118, DynamicsKernel.hpp:119-156). Names: hice, cice, u, v,
 * uiostress, viostress, ssh, s11, s12, s22, e11, e12, e22, and stored passive fields. */
int nsdg_dyn_get_field(nsdg_dyn* d, const char* name, double* out, int ncomp);
/* Raw CG node vectors: u, v, u0, v0, cgH, cgA, gradX, gradY, dStressX, dStressY, uocean,
 * vocean, uwind, vwind, lumpedcgmass. */
int nsdg_dyn_get_cg(nsdg_dyn* d, const char* name, double* out);
int nsdg_dyn_set_cg(nsdg_dyn* d, const char* name, const double* in);
size_t nsdg_dyn_cg_nodes(const nsdg_dyn* d);
/* update(TimestepTime) (VPCGDynamicsKernel.hpp:63-94): advection and limiters, then
 * prepareIteration, u0 = u, then nsteps mEVP subcycles. dt is in seconds (the reference
 * passes TimestepTime::step.seconds(), an integer number of seconds). */
int nsdg_dyn_update(nsdg_dyn* d, double dt);
/* Per-pass entries for testing and composition. */
int nsdg_dyn_advect(nsdg_dyn* d, double dt);          /* advectionAndLimits, DynamicsKernel.hpp:160-172 */
int nsdg_dyn_prepare_iteration(nsdg_dyn* d);          /* CGDynamicsKernel.cpp:260-276 */
int nsdg_dyn_begin_iteration(nsdg_dyn* d, double dt); /* u0 = u; v0 = v; deltaT = dt */
int nsdg_dyn_strain(nsdg_dyn* d);                     /* projectVelocityToStrain, CGDynamicsKernel.cpp:300-337 */
int nsdg_dyn_stress_update(nsdg_dyn* d, int repeat);  /* stressUpdateHighOrder, MEVPStressUpdateStep.hpp:30-118 */
int nsdg_dyn_divergence(nsdg_dyn* d);                 /* stressDivergence, CGDynamicsKernel.cpp:365-398 */
int nsdg_dyn_momentum(nsdg_dyn* d);                   /* updateMomentum, VPCGDynamicsKernel.hpp:132-172 */
int nsdg_dyn_apply_boundaries(nsdg_dyn* d);           /* applyBoundaries, CGDynamicsKernel.cpp:439-444 */
/* n fused mEVP subcycles (strain+stress in one element pass; divergence+momentum+
 * Dirichlet in one node pass); equivalent to n x {strain, stress, divergence,
 * momentum, boundaries}. */
int nsdg_dyn_subcycles(nsdg_dyn* d, int n);
/* Element-pass arithmetic: NSDG_PREC_FP64 (default) or NSDG_PREC_FP32. */
int nsdg_dyn_set_precision(nsdg_dyn* d, int precision);
/* How subcycles on uniform meshes without periodic segments are grouped into passes (no
 * counterpart in the reference: a performance control; the results agree to an ulp,
 * within 1e-14 relative). NSDG_PAIR_AUTO (default): two subcycles per pass from
 * NSDG_PAIR_MIN_ELEMENTS elements up (0: on every such mesh; half the launches on small
 * meshes, the state read once per two subcycles on large ones); NSDG_PAIR_ALWAYS;
 * NSDG_PAIR_NEVER: one per pass. The last subcycle of a run always runs alone. */
#define NSDG_PAIR_AUTO 0
#define NSDG_PAIR_ALWAYS 1
#define NSDG_PAIR_NEVER 2
#define NSDG_PAIR_MIN_ELEMENTS 0
int nsdg_dyn_set_pairing(nsdg_dyn* d, int mode);
/* Benchmark cyclone wind (BenchmarkAtmosphere.cpp:29-76) evaluated on the device at the
 * cell lower-left corners x = i L/nx, y = j L/ny for time t (seconds since the start),
 * with the centre shifted by (xoff, yoff) metres and the given vMax, then set as the
 * DG0 uwind/vwind fields (nodal averages, as setData does). */
int nsdg_dyn_benchmark_wind(nsdg_dyn* d, double L, double t_seconds, double xoff, double yoff, double vmax);
/* Several storms (no counterpart in the reference, which has one): the sum of n <= 8
 * copies of that cyclone on its track, storm i shifted by (xoff[i], yoff[i]) metres with
 * vMax vmax[i] (its peak speed is vmax / e, at 100 km from the centre). */
int nsdg_dyn_storm_wind(nsdg_dyn* d, double L, double t_seconds, const double* xoff, const double* yoff,
    const double* vmax, int n);
/* Zero the higher DG moments of a field on the device: the effect of the module-level
 * setData(name, getDG0Data(name)) round trip of MEVPDynamics::update (MEVPDynamics.cpp:64-65). */
int nsdg_dyn_reset_dg0(nsdg_dyn* d, const char* name);
/* FP64 conservation sums (Tools::MeanValue, Tools.hpp:24-45) of hice and cice over all
 * elements (land included, as the reference sums them), with cos(latitude) weights on
 * spherical meshes (Tools.hpp:28-36; totals in radians squared there), accumulated with
 * compensated (two-sum) sums on the device in a fixed order. */
int nsdg_dyn_mass(nsdg_dyn* d, double* ice_volume, double* ice_area);
/* nsdg_dyn_mass without any allocation in the call. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The block partials and
 * the two totals live in a fixed 1 MiB device array of the library (one per device, held
 * until the process ends), so meshes of at most 8,386,560 elements are accepted; larger
 * ones return NSDG_ERR_ARG. Results equal nsdg_dyn_mass bit for bit. Every call on a
 * device shares that array: calls that overlap on one device (from several host threads)
 * race. */
int nsdg_dyn_mass_noalloc(nsdg_dyn* d, double* ice_volume, double* ice_area);
/* Advection mass budget (no counterpart in the reference, which does not conserve mass
 * exactly: Dirichlet edges are outflow-only, DGTransport.cpp:306-351, and the limiters
 * clamp the cell mean, dgLimit.hpp:16-84). With the budget on, every update/advect adds,
 * per element, the Dirichlet outflow of h and of A over the step's RK stages and the
 * change the limiters make to the Tools::MeanValue contribution of h and A.
 * nsdg_dyn_budget writes the totals since the budget was switched on or last reset:
 * out[0] outflow of h, out[1] outflow of A, out[2] limiter change of h, out[3] limiter
 * change of A, in the units of nsdg_dyn_mass, so that over any run of steps
 *   mass(after) - mass(before) = out[2] - out[0]  (h), out[3] - out[1]  (A)
 * up to rounding. reset != 0 zeroes the totals after reading them. */
int nsdg_dyn_set_budget(nsdg_dyn* d, int on);
int nsdg_dyn_budget(nsdg_dyn* d, double* out, int reset);
/* Momentum residual of the current velocity (no counterpart in the reference, which
 * runs a fixed subcycle count): R = M_L [c1 (u - u0) - A (tau_a + F_oc |u_o - u| (u_o - u))
 * + rho H f_c (u | -v) + rho H g grad(eta)] - dS(Pi sigma_VP(eps(u))) per CG node, the
 * fixed-point equation of updateMomentum (VPCGDynamicsKernel.hpp:132-172) with the stress
 * of MEVPStressUpdateStep.hpp:30-118 at alpha = 1. *r = |R|_2 / |M_L A tau_a|_2 over the
 * non-Dirichlet nodes; the two norms go to *norm_r, *norm_forcing (any may be NULL). */
int nsdg_dyn_residual(nsdg_dyn* d, double* r, double* norm_r, double* norm_forcing);
/* stress <- Pi sigma_VP(eps(u)): the converged mEVP stress of the current velocity
 * (a warm start for the stress that matches a given velocity). */
int nsdg_dyn_stress_from_velocity(nsdg_dyn* d);
int nsdg_dyn_synchronize(nsdg_dyn* d);
/* Device state (all DG fields including stress, CG velocities) as an opaque blob in this
 * library's own format. With buf == NULL, *size receives the required size. */
int nsdg_dyn_save_state(nsdg_dyn* d, void* buf, size_t* size);
int nsdg_dyn_load_state(nsdg_dyn* d, const void* buf, size_t size);
/* Restart files in the layout of the reference's restart writer, ParaGridIO::dumpModelState
 * (core/src/ParaGridIO.cpp:272-349): every ModelArray dimension (xdim, ydim, zdim, xvertex,
 * yvertex, x_cg, y_cg, dg_comp, dgstress_comp, ncoords; ModelArrayDetails.cpp:28-41) and
 * FP64 variables with a missing_value attribute, dimensions in the reference's order
 * (components fastest): hice, cice (and damage) as (ydim, xdim, dg_comp); u, v as cell
 * means (ydim, xdim); mask; coords (yvertex, xvertex, ncoords); plus the stress and strain
 * s11..e22 (ydim, xdim, dgstress_comp) and the CG velocities u_cg, v_cg (and avgU_cg,
 * avgV_cg) that restore this library's state exactly. The file is netCDF classic with
 * 64-bit offsets (CDF-2), which has no groups: the reference's "data" group is the root
 * and its metadata (structure type "parametric_rectangular", time) are global attributes.
 * read_restart loads a file written for a kernel of the same mesh size, degrees and
 * rheology (NSDG_ERR_ARG otherwise) and returns its time attribute. */
int nsdg_dyn_write_restart(nsdg_dyn* d, const char* path, double time_seconds);
int nsdg_dyn_read_restart(nsdg_dyn* d, const char* path, double* time_seconds);
/* Per-kernel timing with HIP events on the library's stream. */
int nsdg_dyn_set_profiling(nsdg_dyn* d, int on);
int nsdg_dyn_kernel_stats(nsdg_dyn* d, const char* kernel, int* count, double* total_ms);
int nsdg_dyn_destroy(nsdg_dyn* d);

#ifdef __cplusplus
}
#endif
#endif /* NSDG_DYN_H */
