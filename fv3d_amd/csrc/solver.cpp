// Host runtime of libfv3d_hip.so: the reference's Solver API
// (src/Solver/include/solver/solver_module.h) over device-resident SoA
// buffers, plus the Navier-Stokes plug-in built-ins
// (src/FV3D/src/navier_stokes/) and the fvh_* extension API.
//
// Ownership follows the reference: init_* allocate, free_* release; the
// library owns every device buffer (allocated in init_finite_volume, freed
// in free_finite_volume).  Single host thread, one HIP stream; not
// thread-safe, like the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fv_device.h"
#include "fvh_internal.h"
#include "raw_container.h"
#include "runtime.h"
#include "implicit.h"
#include "krylov.h"
#include "mirror.h"

// ---------------------------------------------------------------------------
// reference globals (solver_module.h:16-18, navier_stokes.c:16-55)
// ---------------------------------------------------------------------------
extern "C" {
Mesh_t *solver_mesh = nullptr;
Variables_t *solver_variables = nullptr;
data_t *solver_data = nullptr;

double cfl_scale = 1.00;
double dfl_scale = 1.00;
double mu_mix = 18.13e-6;
double R_mix = 8.31446261815324 / 28.96e-3;
double Pr = 0.718;
double kappa = 1.4;
double kappa_m1, kappa_p1, s_kappa, s_kappa_m1, s_kappa_p1, cv, cp, kappa_pr,
    lambda;
int ic_rho = -1, ic_rho_u = -1, ic_rho_v = -1, ic_rho_w = -1, ic_rho_e = -1;
int ip_u = -1, ip_v = -1, ip_w = -1, ip_p = -1, ip_T = -1;
}

namespace fvh {

namespace {

std::string g_error;

#define HIP_CHECK(expr)                                                        \
    do                                                                         \
    {                                                                          \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess)                                                  \
            throw std::runtime_error(std::string("HIP error: ") +             \
                                     hipGetErrorString(e_) + " at " +          \
                                     __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

// ---- solver-wide settings (the reference's file-scope statics) ------------
std::string g_title = "fv3d";
int g_i_output_data = -1;
bool g_do_output = false;
bool g_use_restart = false;
std::string g_restart_file;
int g_iter_restart = 0;
double g_t_restart = 0.0;

bool g_is_explicit = true;
int g_max_iter = 1000;
bool g_transient = true;
double g_abort_residual = 1e-10;
double g_t_start = 0.0;
double g_t_end = 0.2;
bool g_viscous_dt = false;

int g_rk_stages = 0;
const double *g_rk_a = nullptr, *g_rk_b = nullptr, *g_rk_g = nullptr;

int g_reconstruction = -1; // FirstOrder / Linear
int g_limiter = -1;        // BarthJespersenn / NoLimter
int g_flux = dev::FluxAUSM;
int g_visc_dt_mode = 0;
bool g_analyze_ready = false;
std::vector<double> g_residual(5, 0.0);

// implicit settings (implicit.c:73-77 defaults)
int g_bdf_scheme = 0;
int g_implicit_solver = 1;
int g_max_iter_inner = 100;
double g_tolerance_lsoe = 1e-12;
int g_max_iter_lsoe = 100;
int g_max_krylov_dims = 15;
int g_max_krylov_restarts = 2;

// hooks (finite_volume.c:13-16, timedisc.c:15-16)
void_calc_flux_ft g_calc_flux = nullptr;
void_calc_exact_ft g_calc_exact = nullptr;
void_update_ft g_update = nullptr;
void_update_gradients_ft g_update_gradients = nullptr;
double_calc_timestep_ft g_calc_timestep = nullptr;

// built-in initial conditions
int g_ic_kind = 2;
double g_ic_params[8] = {0};

int g_verbose = 1;
int g_output_format = 0; // 0: HDF5 when libhdf5 loads, else raw; 1: raw; 2: HDF5
int g_sync_interval = 16;
int g_last_status = 0;
int g_last_iterations = 0;
std::vector<double> g_log;

std::vector<double> g_face_shift;

// LSERK tables, explicit.c:19-39
const double rk_a_euler[1] = {0.0};
const double rk_b_euler[1] = {0.0};
const double rk_g_euler[1] = {1.0};
const double rk_a_rk33[3] = {0., -5. / 9., -153. / 128.};
const double rk_b_rk33[3] = {0., 1. / 3., 3. / 4.};
const double rk_g_rk33[3] = {1. / 3., 15. / 16., 8. / 15.};
const double rk_a_rk45[5] = {0., -567301805773.0 / 1357537059087.0,
                             -2404267990393.0 / 2016746695238.0,
                             -3550918686646.0 / 2091501179385.0,
                             -1275806237668.0 / 842570457699.0};
const double rk_b_rk45[5] = {0., 1432997174477.0 / 9575080441755.0,
                             2526269341429.0 / 6820363962896.0,
                             2006345519317.0 / 3224310063776.0,
                             2802321613138.0 / 2924317926251.0};
const double rk_g_rk45[5] = {1432997174477.0 / 9575080441755.0,
                             5161836677717.0 / 13612068292357.0,
                             1720146321549.0 / 2090206949498.0,
                             3134564353537.0 / 4481467310338.0,
                             2277821191437.0 / 14882151754819.0};

// ---- device state ----------------------------------------------------------
struct Device
{
    bool ready = false;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    dev::DevMesh m{};
    double *U[2] = {nullptr, nullptr};
    int cur = 0;
    // the buffer whose primitives the reference's phi_total holds: after an
    // RK step, the last stage's input (its update hook ran on that state);
    // otherwise the current state
    int uprim = 0;
    // primitive columns of a restart file [N][5] for the first time step
    double *restart_prims = nullptr;
    bool restart_prims_valid = false;
    double *k = nullptr;
    double *R = nullptr;
    double *gphi = nullptr; // [10*B] ghost rows
    double *vg = nullptr;   // [12*N] viscous gradient subset
    double *fs = nullptr;   // [N*maxf+B][8] face states by cell slot, ghosts
    float *fs32 = nullptr;  // the same rows for the FP32 residual
    float *fgA32 = nullptr, *fgB32 = nullptr; // FP32 face records (FP32 residual)
    float *crv32 = nullptr;                   // FP32 distance vectors by slot
    double *grad = nullptr; // [30*N] all gradients (tests only)
    double *mnmx = nullptr; // [20*N] min/max (tests only)
    double *partials = nullptr;
    int partials_cap = 0; // rows of partials
    double *fold = nullptr;
    int fold_cap = 0;
    dev::DevScalars *sc = nullptr;
    dev::DevScalars *sc_host = nullptr; // pinned
    double *log = nullptr;
    int log_cap = 0;
    double *mirror_phi = nullptr; // AoS staging for the host mirror
    double *mirror_dt = nullptr;
    // host-mirror residual path (allocated on first use): gradients
    // [(N+B)][10] x 3, face states [F][10] x 2, fluxes [F][5], residual,
    // distance vectors [F][3] x 2
    double *mg[3] = {nullptr, nullptr, nullptr};
    double *mL = nullptr, *mR = nullptr, *mflux = nullptr, *mres = nullptr;
    double *md1 = nullptr, *md2 = nullptr;
    double vglob = 0.0;
    bool r_valid = false;
    // reference-semantics implicit step (allocated on first use)
    double *imp_old[2] = {nullptr, nullptr}; // BDF history phi_old[0..1]
    double *imp_f = nullptr, *imp_dY = nullptr, *imp_jac = nullptr;
    double *imp_Rp = nullptr, *imp_Rm = nullptr, *imp_Ut = nullptr;
    double *imp_red = nullptr; // reduction partials [1024*64] + results
    double *imp_kv = nullptr;  // Krylov vectors of the linear solve
    int imp_kv_cols = 0;
};
Device g_dev;

template <class T> T *dmalloc(size_t n)
{
    void *p = nullptr;
    if (n == 0)
        n = 1;
    HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    g_dev.allocs.push_back(p);
    return (T *)p;
}

template <class T> T *dupload(const T *h, size_t n)
{
    T *d = dmalloc<T>(n);
    if (n)
        HIP_CHECK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

void free_device()
{
    if (g_dev.stream)
        (void)hipStreamSynchronize(g_dev.stream);
    jfnk_release();
    learned_release_mesh();
    batch_release();
    for (void *p : g_dev.allocs)
        (void)hipFree(p);
    if (g_dev.sc_host)
        (void)hipHostFree(g_dev.sc_host);
    if (g_dev.stream)
        (void)hipStreamDestroy(g_dev.stream);
    g_dev = Device();
}

char *xstrdup(const char *s)
{
    size_t n = std::strlen(s) + 1;
    char *p = (char *)std::malloc(n);
    std::memcpy(p, s, n);
    return p;
}

template <class T> T *xalloc(size_t n)
{
    T *p = (T *)std::calloc(n ? n : 1, sizeof(T));
    if (!p)
        fatal("host allocation failed");
    return p;
}

template <class T> T *xcopy(const std::vector<T> &v)
{
    T *p = xalloc<T>(v.size());
    if (!v.empty())
        std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

// The residual hooks are the built-in Navier-Stokes ones: the fused device
// kernels run.  Otherwise the residual takes the host-mirror path.
bool builtin_hooks()
{
    return g_calc_flux == calc_ns_flux && g_update == ns_update &&
           g_update_gradients == ns_update_gradients;
}

// The time-step hook is the built-in one (computed on the device inside the
// time loop); otherwise the host calls it before every step.
bool builtin_timestep() { return g_calc_timestep == calc_ns_timestep; }

// Time passed to host hooks (the stage time of the current residual).
double g_hook_time = 0.0;

// Iterations of the last linear solve of the implicit step.
int g_last_lsoe = 0;

dev::PhysDev phys()
{
    dev::PhysDev P{};
    P.kappa = kappa;
    P.kappa_m1 = kappa_m1;
    P.s_kappa_m1 = s_kappa_m1;
    P.s_kappa_p1 = s_kappa_p1;
    P.R = R_mix;
    P.mu = mu_mix;
    P.lam = lambda;
    P.cfl = cfl_scale;
    P.dfl = dfl_scale;
    P.kappa_pr = kappa_pr;
    P.visc_dt_mode = g_visc_dt_mode;
    P.bj = (g_reconstruction == Linear && g_limiter == BarthJespersenn) ? 1 : 0;
    return P;
}

bool is_linear() { return g_reconstruction == Linear; }
bool is_viscous() { return mu_mix != 0.0; }
bool needs_gradients() { return is_linear() || is_viscous(); }

void require_ready()
{
    if (!g_dev.ready)
        throw std::runtime_error("finite volume data not initialised "
                                 "(init_finite_volume)");
    if (g_reconstruction < 0 || g_limiter < 0)
        throw std::runtime_error("reconstruction/limiter not initialised");
    if (!g_calc_flux || !g_update || !g_update_gradients || !g_calc_timestep)
        throw std::runtime_error("a flux, update, update-gradients or time-step hook is "
                                 "not set");
}

// The reconstruction's distance vectors of the mesh by cell face slot
// (x_f - x_c, on side 1 through the periodic image of the cell: the
// subtractions of mesh.c:44-56), slots past a cell's face count repeating
// slot 0; r[(c * maxf + s) * 3 + d].
std::vector<double> slot_distances(int maxf)
{
    const int N = g_dev.m.N;
    const Cells_t *cells = solver_mesh->cells;
    const Faces_t *faces = solver_mesh->faces;
    std::vector<double> crv((size_t)N * maxf * 3);
    for (int c = 0; c < N; ++c)
    {
        const int nf = cells->n_faces[c];
        for (int s = 0; s < maxf; ++s)
        {
            const int f = cells->faces[(size_t)c * maxf + (s < nf ? s : 0)];
            const bool own0 = faces->cells[2 * f] == c;
            const double *sh = &g_face_shift[(size_t)3 * f];
            double *r = &crv[((size_t)c * maxf + s) * 3];
            for (int d = 0; d < 3; ++d)
                r[d] = own0 ? faces->x[3 * f + d] - cells->x[3 * c + d]
                            : faces->x[3 * f + d] - (cells->x[3 * c + d] + sh[d]);
        }
    }
    return crv;
}

// FP32 face records of the FP32 residual (layouts in dev::DevMesh), built
// on first use from the host mesh: each value the FP32 rounding of the FP64
// one, the cell-to-face vectors from the FP64 differences (through the
// periodic image on side 1), as the FP32 kernels compute them from the FP64
// records.
void ensure_fp32_geometry()
{
    if (g_dev.m.fgA32)
        return;
    const Mesh_t *mm = solver_mesh;
    const Faces_t *faces = mm->faces;
    const Cells_t *cells = mm->cells;
    const int F = g_dev.m.F, N = g_dev.m.N;
    std::vector<float> a((size_t)F * 12, 0.0f), b((size_t)F * 8, 0.0f);
    for (int f = 0; f < F; ++f)
    {
        float *ra = &a[(size_t)f * 12];
        float *rb = &b[(size_t)f * 8];
        for (int d = 0; d < 3; ++d)
        {
            ra[d] = (float)faces->n[3 * f + d];
            ra[3 + d] = (float)faces->t1[3 * f + d];
            ra[6 + d] = (float)faces->t2[3 * f + d];
            rb[d] = (float)faces->n[3 * f + d];
        }
        ra[9] = (float)faces->area[f];
        rb[3] = (float)faces->area[f];
        rb[4] = (float)faces->lambda[f];
    }
    g_dev.fgA32 = dupload(a.data(), a.size());
    g_dev.fgB32 = dupload(b.data(), b.size());
    g_dev.m.fgA32 = g_dev.fgA32;
    g_dev.m.fgB32 = g_dev.fgB32;
    // the distance vectors by slot, FP32 roundings of the FP64 differences,
    // 4 floats per slot
    const int maxf = g_dev.m.maxf;
    const std::vector<double> crv = slot_distances(maxf);
    std::vector<float> r32((size_t)N * maxf * 4, 0.0f);
    for (size_t q = 0; q < (size_t)N * maxf; ++q)
        for (int d = 0; d < 3; ++d)
            r32[q * 4 + d] = (float)crv[q * 3 + d];
    g_dev.crv32 = dupload(r32.data(), r32.size());
    g_dev.m.crv32 = g_dev.crv32;
}

// face-state rows: one per cell face slot (the cell's reconstructed state on
// that face), then one per boundary face (its ghost state)
size_t face_state_rows()
{
    return (size_t)g_dev.m.N * g_dev.m.maxf + (size_t)g_dev.m.B;
}

void ensure_gradient_buffers()
{
    size_t N = (size_t)g_dev.m.N, F = (size_t)g_dev.m.F;
    if (is_viscous() && !g_dev.vg)
        g_dev.vg = dmalloc<double>(12 * N);
    if (is_linear() && !g_dev.fs)
        g_dev.fs = dmalloc<double>(8 * face_state_rows());
}

dev::GradOut grad_out()
{
    dev::GradOut o{};
    o.vg = is_viscous() ? g_dev.vg : nullptr;
    o.fs = is_linear() ? g_dev.fs : nullptr;
    return o;
}

// Residual of U_in into R (and, for a stage, into U_out/k).
// Per-kernel event timing of enqueue_residual (the benchmark's
// instrumented pass): gradient and residual/stage kernel durations.
struct KernelTimer
{
    hipEvent_t a = nullptr, b = nullptr;
    double acc[3] = {0, 0, 0};
    int cnt[3] = {0, 0, 0};
    void start() { HIP_CHECK(hipEventRecord(a, g_dev.stream)); }
    void stop(int q)
    {
        HIP_CHECK(hipEventRecord(b, g_dev.stream));
        HIP_CHECK(hipEventSynchronize(b));
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, a, b));
        acc[q] += t * 1e3;
        cnt[q]++;
    }
};

// ---- host-mirror residual path ---------------------------------------------
mirror::MirrorMesh mirror_mesh()
{
    const dev::DevMesh &d = g_dev.m;
    mirror::MirrorMesh m{};
    m.N = d.N;
    m.F = d.F;
    m.B = d.B;
    m.maxf = d.maxf;
    m.cf = d.cf;
    m.ncf = d.ncf;
    m.fcell = d.fcell;
    m.fgA = d.fgA;
    m.fgB = d.fgB;
    m.d1 = g_dev.md1;
    m.d2 = g_dev.md2;
    m.vol = d.vol;
    m.svol = d.svol;
    m.bnd_face = d.bnd_face;
    m.bnd_region = d.bnd_region;
    m.region_type = d.region_type;
    return m;
}

void ensure_mirror_buffers()
{
    if (g_dev.mL)
        return;
    const size_t N = g_dev.m.N, F = g_dev.m.F, B = g_dev.m.B;
    for (int d = 0; d < 3; ++d)
        g_dev.mg[d] = dmalloc<double>((N + B) * 10);
    g_dev.mL = dmalloc<double>(F * 10);
    g_dev.mR = dmalloc<double>(F * 10);
    g_dev.mflux = dmalloc<double>(F * 5);
    g_dev.mres = dmalloc<double>(N * 5);
    g_dev.md1 = dupload(solver_mesh->faces->dist_cell_1, F * 3);
    g_dev.md2 = dupload(solver_mesh->faces->dist_cell_2, F * 3);
}

void d2h(double *h, const double *d, size_t n)
{
    HIP_CHECK(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, g_dev.stream));
    HIP_CHECK(hipStreamSynchronize(g_dev.stream));
}

void h2d(double *d, const double *h, size_t n)
{
    HIP_CHECK(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, g_dev.stream));
    HIP_CHECK(hipStreamSynchronize(g_dev.stream));
}

// One residual evaluation in the reference's passes (finite_volume.c:81-124
// with reconstruction.c:142-255 and the hooks): R_out = dU/dt(U_in).  Each
// hook that is not a built-in runs on the host between the device passes,
// on the host mirror (solver_data), which is synchronised before and
// uploaded after the call.  host_all: leave every solver_data array
// (phi_total, the gradients, the face states, the fluxes) current on the
// host, as the reference's residual does.
void mirror_residual(double t, const double *U_in, double *R_out, bool host_all)
{
    ensure_mirror_buffers();
    hipStream_t st = g_dev.stream;
    const int N = g_dev.m.N, F = g_dev.m.F, B = g_dev.m.B;
    const size_t rows = (size_t)(N + B) * 10;
    const mirror::MirrorMesh mm = mirror_mesh();
    const dev::PhysDev P = phys();
    double *phi = g_dev.mirror_phi;
    double **g = g_dev.mg;
    double *hg[3] = {solver_data->grad_phi_total_x, solver_data->grad_phi_total_y,
                     solver_data->grad_phi_total_z};
    bool host_phi = false, host_grad = false;
    // update hook (ns_update: primitives of the domain cells, ghost rows)
    dev::launch_mirror(st, g_dev.m, P, U_in, nullptr, phi, nullptr);
    if (g_update != ns_update)
    {
        d2h(solver_data->phi_total, phi, rows);
        g_update(t);
        h2d(phi, solver_data->phi_total, rows);
        host_phi = true;
    }
    // gradients, then the update-gradients hook on the ghost rows
    mirror::gradients(st, mm, phi, g[0], g[1], g[2]);
    mirror::ghost_raw(st, mm, phi, g[0], g[1], g[2]);
    if (g_update_gradients != ns_update_gradients)
    {
        for (int d = 0; d < 3; ++d)
            d2h(hg[d], g[d], rows);
        g_update_gradients(t);
        for (int d = 0; d < 3; ++d)
            h2d(g[d], hg[d], rows);
        host_grad = true;
    }
    else
        mirror::ghost_gradients(st, mm, g[0], g[1], g[2]);
    mirror::reconstruct(st, mm, P, is_linear(), phi, g[0], g[1], g[2], g_dev.mL, g_dev.mR);
    bool host_faces = false;
    if (g_calc_flux != calc_ns_flux)
    {
        if (!host_phi)
            d2h(solver_data->phi_total, phi, rows);
        if (!host_grad)
            for (int d = 0; d < 3; ++d)
                d2h(hg[d], g[d], rows);
        d2h(solver_data->phi_total_left, g_dev.mL, (size_t)F * 10);
        d2h(solver_data->phi_total_right, g_dev.mR, (size_t)F * 10);
        host_phi = host_grad = host_faces = true;
        g_calc_flux(t);
        h2d(g_dev.mflux, solver_data->flux, (size_t)F * 5);
    }
    else
        mirror::flux(st, mm, P, g_flux, is_viscous(), g_dev.mL, g_dev.mR, g[0], g[1], g[2],
                     g_dev.mflux);
    mirror::sum_fluxes(st, mm, g_dev.mflux, R_out);
    if (host_all)
    {
        if (!host_phi)
            d2h(solver_data->phi_total, phi, rows);
        if (!host_grad)
            for (int d = 0; d < 3; ++d)
                d2h(hg[d], g[d], rows);
        if (!host_faces)
        {
            d2h(solver_data->phi_total_left, g_dev.mL, (size_t)F * 10);
            d2h(solver_data->phi_total_right, g_dev.mR, (size_t)F * 10);
        }
        if (g_calc_flux == calc_ns_flux)
            d2h(solver_data->flux, g_dev.mflux, (size_t)F * 5);
    }
}

void enqueue_residual(const double *U_in, double *U_out, double *R_out,
                      int stage, bool update, bool reduce, dev::DevScalars *sc,
                      KernelTimer *tm = nullptr, bool fp32 = false,
                      const float *U32 = nullptr)
{
    const dev::PhysDev P = phys();
    if (!builtin_hooks())
    {
        // a host hook: the residual through the host mirror, then the stage
        if (fp32)
            throw std::runtime_error("the FP32 residual needs the built-in hooks");
        double *R = R_out ? R_out : g_dev.R;
        mirror_residual(g_hook_time, U_in, R, false);
        if (update)
        {
            dev::StageParams sp{};
            sp.U_in = U_in;
            sp.U_out = U_out;
            sp.k = g_dev.k;
            sp.rk_a = g_rk_a[stage];
            sp.rk_g = g_rk_g[stage];
            sp.stage0 = stage == 0;
            sp.do_update = 1;
            sp.do_reduce = reduce;
            sp.sc = sc;
            sp.partials = g_dev.partials;
            sp.partials_cap = g_dev.partials_cap;
            dev::launch_stage_update(g_dev.stream, g_dev.m, P, sp, R);
        }
        return;
    }
    ensure_gradient_buffers();
    // the FP32 residual keeps its face states and face records in FP32 and
    // may take its state in FP32
    const bool fs32 = fp32 && is_linear();
    if (fs32 && !g_dev.fs32)
        g_dev.fs32 = dmalloc<float>(8 * face_state_rows());
    if (fp32)
        ensure_fp32_geometry();
    else
        U32 = nullptr;
    dev::launch_ghost(g_dev.stream, g_dev.m, P, U_in, g_dev.gphi,
                      is_linear() && !fs32 ? g_dev.fs : nullptr, sc,
                      fs32 ? g_dev.fs32 : nullptr, U32);
    dev::GradOut go = grad_out();
    go.U32 = U32;
    if (fs32)
    {
        go.fs = nullptr;
        go.fs32 = g_dev.fs32;
    }
    if (needs_gradients())
    {
        if (tm)
            tm->start();
        dev::launch_gradient(g_dev.stream, g_dev.m, P, U_in, g_dev.gphi, go,
                             is_linear(), sc, fp32);
        if (tm)
            tm->stop(0);
    }
    dev::StageParams sp{};
    sp.U_in = U_in;
    sp.U_out = U_out;
    sp.k = g_dev.k;
    sp.R_out = R_out;
    sp.vg = g_dev.vg;
    sp.fs = g_dev.fs;
    sp.fs32 = fs32 ? g_dev.fs32 : nullptr;
    sp.U32 = U32;
    sp.gphi = g_dev.gphi;
    sp.rk_a = update ? g_rk_a[stage] : 0.0;
    sp.rk_g = update ? g_rk_g[stage] : 0.0;
    sp.stage0 = stage == 0;
    sp.do_update = update;
    sp.do_reduce = reduce;
    // the last stage of an explicit step: the next step's first stage
    // starts k from R, so this stage's k is never read
    sp.k_dead = update && reduce && stage == g_rk_stages - 1;
    sp.sc = sc;
    sp.partials = g_dev.partials;
    sp.partials_cap = g_dev.partials_cap;
    if (tm)
        tm->start();
    dev::launch_residual(g_dev.stream, g_dev.m, P, sp, g_flux, is_linear(),
                         is_viscous(), fp32);
    if (tm)
        tm->stop(1);
}

// Tiles of the LDS-staged gradient kernel: per tile of dev::kTile
// consecutive cells the sorted halo (face neighbours outside the tile; ghost
// rows as N + boundary) and per cell face slot the neighbour's row in the
// tile's LDS table.  Not built (the plain gather kernel runs) when
// FVH_K1_TILE=0 or when a tile's table would exceed 64 KiB of LDS.
void setup_tile_halo()
{
    const char *e = std::getenv("FVH_K1_TILE");
    if (e && std::atoi(e) == 0)
        return;
    const int N = g_dev.m.N, maxf = g_dev.m.maxf, T = dev::kTile;
    const int nt = (N + T - 1) / T;
    const Cells_t *cells = solver_mesh->cells;
    const Faces_t *faces = solver_mesh->faces;
    std::vector<int> off(nt + 1, 0), halo;
    std::vector<unsigned short> nbl((size_t)N * maxf, 0);
    int hmax = 0;
    std::vector<int> nbs;
    for (int t = 0; t < nt; ++t)
    {
        const int b0 = t * T, b1 = std::min(N, b0 + T);
        nbs.clear();
        for (int c = b0; c < b1; ++c)
            for (int s = 0; s < cells->n_faces[c]; ++s)
            {
                const int f = cells->faces[(size_t)c * maxf + s];
                const int nb = faces->cells[2 * f] == c ? faces->cells[2 * f + 1]
                                                        : faces->cells[2 * f];
                if (nb < b0 || nb >= b1)
                    nbs.push_back(nb);
            }
        std::sort(nbs.begin(), nbs.end());
        nbs.erase(std::unique(nbs.begin(), nbs.end()), nbs.end());
        hmax = std::max(hmax, (int)nbs.size());
        for (int c = b0; c < b1; ++c)
        {
            const int nf = cells->n_faces[c];
            for (int s = 0; s < maxf; ++s)
            {
                const int f = cells->faces[(size_t)c * maxf + (s < nf ? s : 0)];
                const int nb = faces->cells[2 * f] == c ? faces->cells[2 * f + 1]
                                                        : faces->cells[2 * f];
                int r;
                if (nb >= b0 && nb < b1)
                    r = nb - b0;
                else
                    r = T + (int)(std::lower_bound(nbs.begin(), nbs.end(), nb) - nbs.begin());
                nbl[(size_t)c * maxf + s] = (unsigned short)r;
            }
        }
        halo.insert(halo.end(), nbs.begin(), nbs.end());
        off[t + 1] = (int)halo.size();
    }
    const int rows = T + hmax;
    if ((size_t)10 * rows * sizeof(double) > 65536)
        return;
    g_dev.m.tile_halo_off = dupload(off.data(), off.size());
    g_dev.m.tile_halo = dupload(halo.data(), halo.size());
    g_dev.m.nbl = dupload(nbl.data(), nbl.size());
    g_dev.m.tile_rows = rows;
}


// Face tables of the face-centred residual kernel (tetrahedral meshes): per
// cell tile every face of its cells once, a face between two of its cells
// from the side-0 cell, with the face-state rows of its two sides and the
// tile-local slots that receive its contributions.  A cell that is its own
// periodic neighbour keeps one entry per slot, as the lane-per-slot kernel
// evaluates it: its own row left, the face's other slot right.
void tile_faces_table(const std::vector<int> &c0, std::vector<int> &off,
                      std::vector<int4> &ent)
{
    const int N = g_dev.m.N, maxf = g_dev.m.maxf;
    if (maxf != 4)
        fatal("tile faces: tetrahedral meshes only");
    const Cells_t *cells = solver_mesh->cells;
    const Faces_t *faces = solver_mesh->faces;
    auto slot_of = [&](int x, int f, int not_slot) {
        for (int t = 0; t < cells->n_faces[x]; ++t)
            if (cells->faces[(size_t)x * maxf + t] == f && t != not_slot)
                return t;
        fatal("face " + std::to_string(f) + " is missing from the face list of cell " +
              std::to_string(x));
        return -1;
    };
    const int nt = (int)c0.size() - 1;
    off.assign(nt + 1, 0);
    ent.clear();
    ent.reserve((size_t)N * 5 / 2);
    for (int t = 0; t < nt; ++t)
    {
        const int b0 = c0[t], b1 = c0[t + 1];
        if (b0 < 0 || b1 > N || b1 <= b0 || b1 - b0 > dev::kTile)
            fatal("tile faces: a tile must hold 1 .. kTile cells of the mesh");
        auto in_tile = [&](int x) { return x >= b0 && x < b1; };
        for (int c = b0; c < b1; ++c)
            for (int s = 0; s < cells->n_faces[c]; ++s)
            {
                const int f = cells->faces[(size_t)c * maxf + s];
                const int fc0 = faces->cells[2 * f], fc1 = faces->cells[2 * f + 1];
                int rowL, rowR, dA, dB = 0xffff;
                if (fc0 == fc1)
                {
                    rowL = c * 4 + s;
                    rowR = c * 4 + slot_of(c, f, s);
                    dA = (c - b0) * 4 + s;
                }
                else
                {
                    const bool interior = fc1 < N;
                    if (interior && c == fc1 && in_tile(fc0))
                        continue; // evaluated from its side-0 cell
                    const int s0 = c == fc0 ? s : slot_of(fc0, f, -1);
                    rowL = fc0 * 4 + s0;
                    dA = in_tile(fc0) ? (fc0 - b0) * 4 + s0 : 0xffff;
                    if (interior)
                    {
                        const int s1 = c == fc1 ? s : slot_of(fc1, f, -1);
                        rowR = fc1 * 4 + s1;
                        if (in_tile(fc1))
                            dB = (fc1 - b0) * 4 + s1;
                    }
                    else
                        rowR = N * 4 + (fc1 - N);
                }
                ent.push_back(make_int4(f, rowL, rowR, dA | (dB << 16)));
            }
        off[t + 1] = (int)ent.size();
    }
}

// The face tables over tiles of dev::kTile consecutive cells (the explicit
// path's residual kernel); not built with FVH_K2_FACE=0.
// The 8-byte entries of a face table (dev::DevMesh::tile_faces) from the
// {face, left row, right row, slots} ones tile_faces_table builds, each
// checked against its tile's slots; false when the mesh is too large for
// their fields.
bool pack_tile_faces(const std::vector<int> &c0, const std::vector<int> &off,
                     const std::vector<int4> &ent, std::vector<unsigned long long> &out)
{
    const long long rows = (long long)g_dev.m.N * 4 + g_dev.m.B;
    if (dev::kTile * 4 > 512 || g_dev.m.F >= (1 << 26) || rows >= (1LL << 27))
        return false;
    out.resize(ent.size());
    for (size_t t = 0; t + 1 < c0.size(); ++t)
    {
        const long long b4 = (long long)c0[t] * 4;
        for (int e = off[t]; e < off[t + 1]; ++e)
        {
            const int4 r = ent[e];
            const int dA = r.w & 0xffff, dB = (r.w >> 16) & 0xffff;
            unsigned long long q = (unsigned long long)r.x;
            if (dA != 0xffff && dB != 0xffff)
            {
                if (r.y != b4 + dA || r.z != b4 + dB)
                    fatal("tile faces: an entry's rows are not its tile slots");
                q |= (unsigned long long)dA << 26 | (unsigned long long)dB << 35;
            }
            else
            {
                const bool right = dA == 0xffff;
                const int a = right ? dB : dA;
                const int in_row = right ? r.z : r.y, out_row = right ? r.y : r.z;
                if (in_row != b4 + a || out_row < 0)
                    fatal("tile faces: an entry's row is not its tile slot");
                q |= (unsigned long long)a << 26 | (unsigned long long)(right ? 1 : 0) << 35 |
                     (unsigned long long)out_row << 36 | 1ULL << 63;
            }
            out[e] = q;
        }
    }
    return true;
}

void setup_tile_faces()
{
    const char *e = std::getenv("FVH_K2_FACE");
    if (e && std::atoi(e) == 0)
        return;
    const int N = g_dev.m.N, T = dev::kTile;
    if (g_dev.m.maxf != 4)
        return;
    std::vector<int> c0;
    for (int c = 0; c < N; c += T)
        c0.push_back(c);
    c0.push_back(N);
    std::vector<int> off;
    std::vector<int4> ent;
    tile_faces_table(c0, off, ent);
    std::vector<unsigned long long> packed;
    if (!pack_tile_faces(c0, off, ent, packed))
        return; // too large for the 8-byte entries: the lane-per-slot kernel runs
    g_dev.m.tile_face_off = dupload(off.data(), off.size());
    g_dev.m.tile_faces = dupload(packed.data(), packed.size());
}

void sync_mirror(bool with_dt)
{
    const int N = g_dev.m.N, B = g_dev.m.B;
    dev::launch_mirror(g_dev.stream, g_dev.m, phys(), g_dev.U[g_dev.cur],
                       (with_dt && g_dev.r_valid) ? g_dev.R : nullptr,
                       g_dev.mirror_phi, g_dev.mirror_dt, g_dev.U[g_dev.uprim]);
    HIP_CHECK(hipMemcpyAsync(solver_data->phi_total, g_dev.mirror_phi,
                             (size_t)(N + B) * 10 * sizeof(double),
                             hipMemcpyDeviceToHost, g_dev.stream));
    if (with_dt && g_dev.r_valid)
        HIP_CHECK(hipMemcpyAsync(solver_data->phi_dt, g_dev.mirror_dt,
                                 (size_t)N * 5 * sizeof(double),
                                 hipMemcpyDeviceToHost, g_dev.stream));
    HIP_CHECK(hipStreamSynchronize(g_dev.stream));
}

void upload_state_from_mirror()
{
    const int N = g_dev.m.N;
    std::vector<double> aos((size_t)5 * N);
    for (int c = 0; c < N; ++c)
        for (int j = 0; j < 5; ++j)
            aos[(size_t)c * 5 + j] = solver_data->phi_total[(size_t)c * 10 + j];
    HIP_CHECK(hipMemcpy(g_dev.U[g_dev.cur], aos.data(), aos.size() * 8,
                        hipMemcpyHostToDevice));
    g_dev.uprim = g_dev.cur;
    g_dev.restart_prims_valid = false;
    g_dev.r_valid = false;
}

// ---- mesh ------------------------------------------------------------------
void free_mesh_struct(Mesh_t *m)
{
    if (!m)
        return;
    if (m->vertices)
        std::free(m->vertices->x), std::free(m->vertices);
    if (Cells_t *c = m->cells)
    {
        std::free(c->id), std::free(c->type), std::free(c->n_vertices);
        std::free(c->vertices), std::free(c->n_faces), std::free(c->faces);
        std::free(c->x), std::free(c->volume), std::free(c->dx);
        std::free(c->stride), std::free(c);
    }
    if (Faces_t *f = m->faces)
    {
        std::free(f->type), std::free(f->n_vertices), std::free(f->vertices);
        std::free(f->cells), std::free(f->boundary), std::free(f->area);
        std::free(f->lambda), std::free(f->x), std::free(f->n);
        std::free(f->t1), std::free(f->t2), std::free(f->stride);
        std::free(f->dist_cell_1), std::free(f->dist_cell_2);
        std::free(f->internal_faces), std::free(f->boundary_faces);
        std::free(f);
    }
    if (Boundaries_t *b = m->boundaries)
    {
        std::free(b->id), std::free(b->type), std::free(b->n_vertices);
        std::free(b->vertices), std::free(b->face), std::free(b->distance);
        std::free(b->n), std::free(b->t1), std::free(b->t2);
        std::free(b->stride), std::free(b);
    }
    if (Regions_t *r = m->regions)
    {
        for (int i = 0; i < r->n_regions; ++i)
            std::free(r->name[i]);
        std::free(r->name), std::free(r->is_boundary);
        std::free(r->type), std::free(r->function_id), std::free(r->phi_total);
        std::free(r);
    }
    std::free(m->partition);
    std::free(m);
}

// mesh.c:19-83 (with the periodic image of cells[1] for dist_cell_2)
void calc_mesh_metrics(Mesh_t *mesh)
{
    Cells_t *cells = mesh->cells;
    Faces_t *faces = mesh->faces;
    Regions_t *regions = mesh->regions;

    double v = 0.0;
    for (int i = 0; i < cells->n_domain_cells; ++i)
        v += cells->volume[i];
    mesh->local_volume = v;
    mesh->global_volume = v;

    faces->n_internal_faces = 0;
    for (int i = 0; i < faces->n_faces; ++i)
        if (faces->boundary[i] < 0)
            faces->n_internal_faces += 1;
    faces->n_boundary_faces = faces->n_faces - faces->n_internal_faces;

    faces->dist_cell_1 = xalloc<double>((size_t)faces->n_faces * DIM);
    faces->dist_cell_2 = xalloc<double>((size_t)faces->n_faces * DIM);
    faces->internal_faces = xalloc<int>(faces->n_internal_faces);
    faces->boundary_faces = xalloc<int>(faces->n_boundary_faces);

    int k = 0, l = 0;
    for (int i = 0; i < faces->n_faces; ++i)
    {
        int *fc = &faces->cells[i * FACE_CELLS];
        for (int j = 0; j < DIM; ++j)
            faces->dist_cell_1[i * DIM + j] =
                faces->x[i * DIM + j] - cells->x[fc[0] * DIM + j];
        if (fc[1] >= 0)
        {
            for (int j = 0; j < DIM; ++j)
            {
                double xc1 = cells->x[fc[1] * DIM + j] +
                             g_face_shift[(size_t)i * DIM + j];
                faces->dist_cell_2[i * DIM + j] = faces->x[i * DIM + j] - xc1;
            }
            faces->internal_faces[k++] = i;
        }
        else
        {
            for (int j = 0; j < DIM; ++j)
                faces->dist_cell_2[i * DIM + j] = 0.0;
            fc[1] = cells->n_local_cells + faces->boundary[i];
            faces->boundary_faces[l++] = i;
        }
    }
    regions->flow_region = -1;
    for (int i = 0; i < regions->n_regions; ++i)
        if (regions->is_boundary[i] == 0)
        {
            regions->flow_region = i;
            break;
        }
}

} // namespace

void build_tile_faces(const std::vector<int> &c0, std::vector<int> &off,
                      std::vector<int4> &ent)
{
    tile_faces_table(c0, off, ent);
}

bool pack_tile_faces_table(const std::vector<int> &c0, const std::vector<int> &off,
                           const std::vector<int4> &ent, std::vector<unsigned long long> &out)
{
    return pack_tile_faces(c0, off, ent, out);
}

void set_error(const std::string &msg) { g_error = msg; }

void fatal(const std::string &msg)
{
    std::fprintf(stderr, "fv3d_hip: fatal: %s\n", msg.c_str());
    std::fflush(stderr);
    std::abort();
}

const std::vector<double> &face_shift() { return g_face_shift; }

void install_mesh(RawMesh &&m)
{
    if (solver_mesh)
        free_mesh_struct(solver_mesh);
    Mesh_t *mesh = xalloc<Mesh_t>(1);
    mesh->dimension = m.dimension;
    mesh->is_partitioned = 0;
    mesh->partition = nullptr;

    mesh->vertices = xalloc<Vertices_t>(1);
    mesh->vertices->n_vertices = (int)(m.vertices.size() / 3);
    mesh->vertices->x = xcopy(m.vertices);

    const int N = m.n_cells(), F = m.n_faces(), B = m.n_boundaries();
    Cells_t *c = xalloc<Cells_t>(1);
    c->n_global_cells = c->n_local_cells = c->n_domain_cells = N;
    c->max_cell_vertices = m.max_cell_vertices;
    c->max_cell_faces = m.max_cell_faces;
    c->id = xcopy(m.cell_region);
    c->type = xcopy(m.cell_type);
    c->n_vertices = xcopy(m.cell_nv);
    c->vertices = xcopy(m.cell_vertices);
    c->n_faces = xcopy(m.cell_nf);
    c->faces = xcopy(m.cell_faces);
    c->x = xcopy(m.cell_x);
    c->volume = xcopy(m.cell_volume);
    c->dx = xcopy(m.cell_dx);
    c->stride = nullptr;
    mesh->cells = c;

    Faces_t *f = xalloc<Faces_t>(1);
    f->n_global_faces = f->n_faces = F;
    f->max_face_vertices = m.max_face_vertices;
    f->type = xcopy(m.face_type);
    f->n_vertices = xcopy(m.face_nv);
    f->vertices = xcopy(m.face_vertices);
    f->cells = xcopy(m.face_cells);
    f->boundary = xcopy(m.face_boundary);
    f->area = xcopy(m.face_area);
    f->lambda = xcopy(m.face_lambda);
    f->x = xcopy(m.face_x);
    f->n = xcopy(m.face_n);
    f->t1 = xcopy(m.face_t1);
    f->t2 = xcopy(m.face_t2);
    f->stride = nullptr;
    mesh->faces = f;

    Boundaries_t *b = xalloc<Boundaries_t>(1);
    b->n_global_boundaries = b->n_boundaries = B;
    b->max_boundary_vertices = m.max_boundary_vertices;
    b->id = xcopy(m.bnd_region);
    b->type = xcopy(m.bnd_type);
    b->n_vertices = xcopy(m.bnd_nv);
    b->vertices = xcopy(m.bnd_vertices);
    b->face = xcopy(m.bnd_face);
    b->distance = xcopy(m.bnd_distance);
    b->n = xcopy(m.bnd_n);
    b->t1 = xcopy(m.bnd_t1);
    b->t2 = xcopy(m.bnd_t2);
    b->stride = nullptr;
    mesh->boundaries = b;

    Regions_t *r = xalloc<Regions_t>(1);
    r->n_regions = (int)m.region_names.size();
    r->name = xalloc<string_t>(r->n_regions);
    r->max_name_length = 0;
    for (int i = 0; i < r->n_regions; ++i)
    {
        r->name[i] = xstrdup(m.region_names[i].c_str());
        r->max_name_length =
            std::max(r->max_name_length, (int)m.region_names[i].size() + 1);
    }
    r->is_boundary = xcopy(m.region_is_boundary);
    r->type = nullptr;
    r->function_id = nullptr;
    r->phi_total = nullptr;
    mesh->regions = r;

    g_face_shift = std::move(m.face_shift);
    if (g_face_shift.size() != (size_t)F * 3)
        g_face_shift.assign((size_t)F * 3, 0.0);

    calc_mesh_metrics(mesh);
    solver_mesh = mesh;
}

// ---- built-in IC (fvh_set_ic) ----------------------------------------------
static void builtin_exact(int id, double t, double *x, double *phi)
{
    (void)t;
    const int nt = solver_variables->n_tot_variables;
    if (g_ic_kind == 2)
    {
        // calc_ns_exact_function, navier_stokes.c:133-152
        calc_ns_exact_function(id, t, x, phi);
        return;
    }
    for (int j = 0; j < nt; ++j)
        phi[j] = 0.0;
    double rho, u, v, w, p;
    if (g_ic_kind == 0)
    {
        const double a = g_ic_params[0], b = g_ic_params[1];
        rho = a + b * std::sin(2.0 * M_PI * (x[0] + x[1] + x[2]));
        u = v = w = g_ic_params[2];
        p = g_ic_params[3];
    }
    else
    {
        const double p0 = g_ic_params[0], rho0 = g_ic_params[1],
                     U0 = g_ic_params[2];
        const double T0 = p0 / (rho0 * R_mix);
        u = U0 * std::sin(x[0]) * std::cos(x[1]) * std::cos(x[2]);
        v = -U0 * std::cos(x[0]) * std::sin(x[1]) * std::cos(x[2]);
        w = 0.0;
        p = p0 + rho0 * U0 * U0 / 16.0 * (std::cos(2.0 * x[0]) + std::cos(2.0 * x[1])) *
                     (std::cos(2.0 * x[2]) + 2.0);
        rho = p / (R_mix * T0);
    }
    phi[ic_rho] = rho;
    phi[ip_u] = u;
    phi[ip_v] = v;
    phi[ip_w] = w;
    phi[ip_p] = p;
    phi[ip_T] = p / (rho * R_mix);
    // prim_to_con, eos.c:120-127
    phi[ic_rho_u] = phi[ip_u] * phi[ic_rho];
    phi[ic_rho_v] = phi[ip_v] * phi[ic_rho];
    phi[ic_rho_w] = phi[ip_w] * phi[ic_rho];
    double d = 0.0;
    d += phi[ip_u] * phi[ic_rho_u];
    d += phi[ip_v] * phi[ic_rho_v];
    d += phi[ip_w] * phi[ic_rho_w];
    phi[ic_rho_e] = s_kappa_m1 * phi[ip_p] + 0.5 * d;
}

// ---- time loop ---------------------------------------------------------------
namespace {

void print_header()
{
    if (!g_verbose)
        return;
    if (g_is_explicit)
        std::printf("%9s %12s %12s %1s %1s:", "iter", "time", "dt", "V", "O");
    else
        std::printf("%9s %12s %12s %1s %1s %6s %6s:", "iter", "time", "dt",
                    "V", "O", "inner", "lsoe");
    for (int i = 0; i < solver_variables->n_sol_variables; ++i)
        std::printf(" %12s", solver_variables->sol_variables[i].name);
    std::printf("\n");
}

void print_row(int iter, const double *row, bool out, int inner = 0, int lsoe = 0)
{
    if (!g_verbose)
        return;
    if (g_is_explicit)
        std::printf("%09d %12.5e %12.5e %c %c:", iter, row[0], row[1],
                    row[2] != 0.0 ? 'T' : 'F', out ? '*' : ' ');
    else
        std::printf("%09d %12.5e %12.5e %c %c %6d %6d:", iter, row[0], row[1],
                    row[2] != 0.0 ? 'T' : 'F', out ? '*' : ' ', inner, lsoe);
    for (int j = 0; j < 5; ++j)
        std::printf(" %12.5e", row[3 + j]);
    std::printf("\n");
}

// One explicit step (timestep_lserkw2, explicit.c:96-141) + bookkeeping.
// t, dt: the step's time and time step as the host knows them (host hooks
// receive the stage time t + b_s dt; the device kernels read dt from sc).
void enqueue_explicit_step(dev::DevScalars *sc, KernelTimer *tm = nullptr, double t = 0.0,
                           double dt = 0.0)
{
    for (int s = 0; s < g_rk_stages; ++s)
    {
        const bool last = s == g_rk_stages - 1;
        g_hook_time = t + g_rk_b[s] * dt;
        double *Uin = g_dev.U[g_dev.cur];
        double *Uout = g_dev.U[1 - g_dev.cur];
        enqueue_residual(Uin, Uout, last ? g_dev.R : nullptr, s, true, last,
                         sc, tm);
        g_dev.cur = 1 - g_dev.cur;
    }
    if (tm)
        tm->start();
    const int n_partials =
        builtin_hooks() ? dev::residual_grid(g_dev.m) : dev::cell_blocks(g_dev.m.N);
    dev::launch_finalize(g_dev.stream, sc, g_dev.partials, n_partials, g_dev.fold,
                         g_dev.fold_cap, phys(), 1, g_dev.log, g_dev.log_cap);
    if (tm)
        tm->stop(2);
}

// A time-step hook that is not the built-in one, called on the host before a
// step (timedisc.c:164-172): the host mirror is refreshed so that the hook
// reads the current state, and the step is clamped to t_end.  h holds the
// device scalars just read; they are written back with the new dt.
void host_timestep(dev::DevScalars &h)
{
    sync_mirror(false);
    double dt = g_calc_timestep(h.t);
    h.viscous = g_viscous_dt ? 1 : 0;
    if (h.t + dt > h.t_end)
    {
        dt = h.t_end - h.t;
        h.final_pending = 1;
    }
    h.dt = dt;
    HIP_CHECK(hipMemcpyAsync(g_dev.sc, &h, sizeof(h), hipMemcpyHostToDevice, g_dev.stream));
    HIP_CHECK(hipStreamSynchronize(g_dev.stream));
}

void init_scalars(int iter0, double t0, int max_iter, bool transient,
                  double t_end)
{
    dev::DevScalars &h = *g_dev.sc_host;
    std::memset(&h, 0, sizeof(h));
    h.t = t0;
    h.dt = 0.0;
    h.t_end = transient ? t_end : DBL_MAX;
    h.abort_residual = g_abort_residual;
    h.vglob = g_dev.vglob;
    h.iter = iter0;
    h.max_iter_abs = iter0 + max_iter;
    h.transient = transient;
    h.log_base = iter0;
    h.host_dt = builtin_timestep() ? 0 : 1;
    HIP_CHECK(hipMemcpyAsync(g_dev.sc, &h, sizeof(h), hipMemcpyHostToDevice,
                             g_dev.stream));
    // initial dt from the current state (primitives refreshed from U); with
    // a host time-step hook the host sets it before each step
    if (h.host_dt)
        return;
    // after a restart, the first dt takes the stored primitives with the
    // stored density, as the reference's calc_ns_timestep reads phi_total
    dev::launch_dt_partials(g_dev.stream, g_dev.m, phys(), g_dev.U[g_dev.cur],
                            g_dev.partials, g_dev.partials_cap,
                            g_dev.restart_prims_valid ? g_dev.restart_prims : nullptr);
    dev::launch_finalize(g_dev.stream, g_dev.sc, g_dev.partials, dev::dt_partials_grid(g_dev.m), g_dev.fold, g_dev.fold_cap, phys(), 0, nullptr, 0);
}

void read_scalars()
{
    HIP_CHECK(hipMemcpyAsync(g_dev.sc_host, g_dev.sc, sizeof(dev::DevScalars),
                             hipMemcpyDeviceToHost, g_dev.stream));
    HIP_CHECK(hipStreamSynchronize(g_dev.stream));
}

// timedisc.c:134-222, explicit branch, driven from device scalars.
int run_explicit(int max_iter)
{
    require_ready();
    if (g_rk_stages <= 0)
        throw std::runtime_error("explicit scheme not initialised");
    int iter0 = g_use_restart ? g_iter_restart : 0;
    double t0 = g_use_restart ? g_t_restart : g_t_start;

    print_header();
    if (iter0 == 0)
    {
        enqueue_residual(g_dev.U[g_dev.cur], nullptr, g_dev.R, 0, false, false,
                         nullptr);
        g_dev.r_valid = true;
    }
    if (g_do_output && !g_use_restart)
        write_output(iter0, t0);

    if (g_dev.log_cap < max_iter)
    {
        g_dev.log_cap = max_iter;
        g_dev.log = dmalloc<double>((size_t)g_dev.log_cap * 8);
    }
    g_log.assign((size_t)max_iter * 8, 0.0);
    init_scalars(iter0, t0, max_iter, g_transient, g_t_end);

    int printed = 0;
    int status = 0;
    // host hooks need the host to know t and dt at every step
    const bool host_driven = !builtin_hooks() || !builtin_timestep();
    while (true)
    {
        int chunk = std::max(1, g_sync_interval);
        if (g_do_output && g_i_output_data > 0)
        {
            int it = iter0 + printed;
            int to_next = g_i_output_data - (it % g_i_output_data);
            chunk = std::min(chunk, to_next);
        }
        const int cur0 = g_dev.cur, iter_before = iter0 + printed;
        if (host_driven)
        {
            read_scalars();
            dev::DevScalars h = *g_dev.sc_host;
            if (!builtin_timestep())
                host_timestep(h);
            enqueue_explicit_step(g_dev.sc, nullptr, h.t, h.dt);
        }
        else
            for (int i = 0; i < chunk; ++i)
                enqueue_explicit_step(g_dev.sc);
        g_dev.r_valid = true;
        read_scalars();
        const dev::DevScalars &h = *g_dev.sc_host;
        {
            // the steps after the one that ended the run did nothing: the
            // final state is where that step left it (a NaN step ends the
            // run without counting)
            const int steps = h.iter - iter_before + (h.done < 0 ? 1 : 0);
            g_dev.cur = (cur0 + g_rk_stages * steps) & 1;
            if (steps > 0)
            {
                g_dev.uprim = 1 - g_dev.cur;
                g_dev.restart_prims_valid = false;
            }
        }
        int done_rows = h.iter - iter0;
        if (done_rows > printed)
        {
            HIP_CHECK(hipMemcpy(&g_log[(size_t)printed * 8],
                                &g_dev.log[(size_t)printed * 8],
                                (size_t)(done_rows - printed) * 8 * sizeof(double),
                                hipMemcpyDeviceToHost));
            for (int r = printed; r < done_rows; ++r)
            {
                int it = iter0 + r + 1;
                bool last = (r == done_rows - 1) && h.done != 0;
                bool out = last || (g_i_output_data > 0 && it % g_i_output_data == 0);
                print_row(it, &g_log[(size_t)r * 8], out);
                for (int j = 0; j < 5; ++j)
                    g_residual[j] = g_log[(size_t)r * 8 + 3 + j];
                g_viscous_dt = g_log[(size_t)r * 8 + 2] != 0.0;
            }
            printed = done_rows;
        }
        if (h.done < 0)
        {
            status = FVH_ERR_NAN;
            break;
        }
        if (g_do_output && h.done == 0 && g_i_output_data > 0 &&
            (iter0 + printed) % g_i_output_data == 0)
            write_output(iter0 + printed, h.t);
        if (h.done > 0)
        {
            if (g_do_output)
                write_output(iter0 + printed, h.t);
            break;
        }
    }
    g_last_iterations = printed;
    g_last_status = status;
    if (g_verbose)
        std::fflush(stdout);
    return status;
}

// ---- reference-semantics implicit step (implicit.c:224-352) --------------
void ensure_implicit_buffers()
{
    const size_t n = (size_t)5 * g_dev.m.N;
    if (g_dev.imp_f)
        return;
    g_dev.imp_old[0] = dmalloc<double>(n);
    g_dev.imp_old[1] = dmalloc<double>(n);
    g_dev.imp_f = dmalloc<double>(n);
    g_dev.imp_dY = dmalloc<double>(n);
    g_dev.imp_jac = dmalloc<double>(5 * n);
    g_dev.imp_Rp = dmalloc<double>(n);
    g_dev.imp_Rm = dmalloc<double>(n);
    g_dev.imp_Ut = dmalloc<double>(n);
    g_dev.imp_red = dmalloc<double>((size_t)1024 * 64 + 64);
    // dY_n is the next linear solve's initial guess (implicit.c:285-300)
    HIP_CHECK(hipMemset(g_dev.imp_dY, 0, n * sizeof(double)));
}

// Dots <a_j, b> for the columns a_j = A + j*ld, j < k, into host h.
void device_dots(int n, int k, const double *A, size_t ld, const double *b, double *h)
{
    double *out = g_dev.imp_red + (size_t)1024 * 64;
    kry::multidot<double>(g_dev.stream, n, k, A, ld, b, g_dev.imp_red, out);
    HIP_CHECK(hipMemcpyAsync(h, out, (size_t)k * sizeof(double), hipMemcpyDeviceToHost,
                             g_dev.stream));
    HIP_CHECK(hipStreamSynchronize(g_dev.stream));
}

// The implicit step's linear solve J dY = f (implicit.c:285-300) on the
// device, as the oracle restates basec's solve_gmres_n_m /
// solve_bicgstab_n_m: restarted GMRES(max_krylov_dims) with classical
// Gram-Schmidt applied twice and Givens rotations, or BiCGStab; dY holds
// the previous solution as the initial guess; stops at |f - J dY| <= tol
// or max_iter_lsoe iterations.  Returns the iterations.
int lsoe_solve(double tol)
{
    hipStream_t s = g_dev.stream;
    const int N = g_dev.m.N, n = 5 * N;
    const size_t ld = ((size_t)n + 63) / 64 * 64;
    const double *jac = g_dev.imp_jac, *f = g_dev.imp_f;
    double *x = g_dev.imp_dY;
    const int m = std::max(1, std::min(g_max_krylov_dims, kry::kMaxDots - 1));
    const int cols = g_implicit_solver == BiCGStab ? 6 : m + 1;
    if (g_dev.imp_kv_cols < cols)
    {
        g_dev.imp_kv = dmalloc<double>(cols * ld);
        g_dev.imp_kv_cols = cols;
    }
    double *V = g_dev.imp_kv;
    const int max_it = g_max_iter_lsoe;
    int it = 0;
    if (g_implicit_solver == BiCGStab)
    {
        double *r = V, *rh = V + ld, *p = V + 2 * ld, *v = V + 3 * ld, *sv = V + 4 * ld,
               *t = V + 5 * ld;
        imp::block_matvec(s, N, jac, x, r);
        kry::rhs_minus<double>(s, n, f, r);
        HIP_CHECK(hipMemcpyAsync(rh, r, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(double), s));
        HIP_CHECK(hipMemsetAsync(v, 0, n * sizeof(double), s));
        double rho = 1.0, alpha = 1.0, omega = 1.0, d[2];
        device_dots(n, 1, r, ld, r, d);
        double rn = std::sqrt(d[0]);
        while (rn > tol && it < max_it)
        {
            ++it;
            device_dots(n, 1, rh, ld, r, d);
            const double rho1 = d[0];
            const double beta = (rho1 / rho) * (alpha / omega);
            imp::lsoe_elem(s, n, 2, beta, omega, r, v, nullptr, p, nullptr);
            imp::block_matvec(s, N, jac, p, v);
            device_dots(n, 1, rh, ld, v, d);
            alpha = rho1 / d[0];
            imp::lsoe_elem(s, n, 1, alpha, 0.0, r, v, nullptr, sv, nullptr);
            device_dots(n, 1, sv, ld, sv, d);
            const double sn = std::sqrt(d[0]);
            if (sn <= tol)
            {
                imp::lsoe_elem(s, n, 3, alpha, 0.0, p, nullptr, nullptr, x, nullptr);
                break;
            }
            imp::block_matvec(s, N, jac, sv, t);
            device_dots(n, 1, t, ld, sv, d);
            device_dots(n, 1, t, ld, t, d + 1);
            omega = d[0] / d[1];
            imp::lsoe_elem(s, n, 4, alpha, omega, p, sv, t, x, r);
            device_dots(n, 1, r, ld, r, d);
            rn = std::sqrt(d[0]);
            rho = rho1;
        }
        HIP_CHECK(hipStreamSynchronize(s));
        return it;
    }
    std::vector<double> H((size_t)(m + 1) * m, 0.0), g(m + 1), cs(m), sn(m), y(m),
        h1(m + 1), h2(m + 1);
    double *yd = g_dev.imp_red + (size_t)1024 * 64; // reuses the dot results slot
    for (int cycle = 0; cycle <= g_max_krylov_restarts; ++cycle)
    {
        imp::block_matvec(s, N, jac, x, V);
        kry::rhs_minus<double>(s, n, f, V);
        double d;
        device_dots(n, 1, V, ld, V, &d);
        const double beta = std::sqrt(d);
        if (beta <= tol || it >= max_it)
            break;
        imp::lsoe_elem(s, n, 0, beta, 0.0, V, nullptr, nullptr, V, nullptr);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        double resid = beta;
        for (int j = 0; j < m && it < max_it; ++j)
        {
            double *vj1 = V + (size_t)(j + 1) * ld;
            imp::block_matvec(s, N, jac, V + (size_t)j * ld, vj1);
            ++it;
            for (int pass = 0; pass < 2; ++pass)
            {
                std::vector<double> &h = pass == 0 ? h1 : h2;
                device_dots(n, j + 1, V, ld, vj1, h.data());
                HIP_CHECK(hipMemcpyAsync(yd, h.data(), (size_t)(j + 1) * sizeof(double),
                                         hipMemcpyHostToDevice, s));
                kry::multiaxpy<double>(s, n, j + 1, V, ld, yd, vj1);
            }
            for (int i = 0; i <= j; ++i)
                H[(size_t)i * m + j] = h1[i] + h2[i];
            device_dots(n, 1, vj1, ld, vj1, &d);
            const double hn = std::sqrt(d);
            H[(size_t)(j + 1) * m + j] = hn;
            if (hn > 0.0)
                imp::lsoe_elem(s, n, 0, hn, 0.0, vj1, nullptr, nullptr, vj1, nullptr);
            for (int i = 0; i < j; ++i)
            {
                const double a = H[(size_t)i * m + j], c = H[(size_t)(i + 1) * m + j];
                H[(size_t)i * m + j] = cs[i] * a + sn[i] * c;
                H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
            }
            const double a = H[(size_t)j * m + j], c = H[(size_t)(j + 1) * m + j];
            const double r = std::sqrt(a * a + c * c);
            cs[j] = r > 0.0 ? a / r : 1.0;
            sn[j] = r > 0.0 ? c / r : 0.0;
            H[(size_t)j * m + j] = r;
            H[(size_t)(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            resid = std::fabs(g[j + 1]);
            k = j + 1;
            if (resid <= tol)
                break;
        }
        for (int i = k - 1; i >= 0; --i)
        {
            double v = g[i];
            for (int q = i + 1; q < k; ++q)
                v -= H[(size_t)i * m + q] * y[q];
            y[i] = v / H[(size_t)i * m + i];
        }
        if (k > 0)
        {
            HIP_CHECK(hipMemcpyAsync(yd, y.data(), (size_t)k * sizeof(double),
                                     hipMemcpyHostToDevice, s));
            imp::add_combination(s, n, k, V, ld, yd, x);
        }
        if (resid <= tol || it >= max_it)
            break;
    }
    HIP_CHECK(hipStreamSynchronize(s));
    return it;
}

double device_norm2(const double *v, int n)
{
    // |v|_2 through the column sums of squares of v viewed as [n/5][5]
    double sums[5];
    double *out = g_dev.imp_red + (size_t)1024 * 64;
    imp::column_sumsq(g_dev.stream, n / 5, v, g_dev.imp_red, out);
    HIP_CHECK(hipMemcpyAsync(sums, out, sizeof(sums), hipMemcpyDeviceToHost,
                             g_dev.stream));
    HIP_CHECK(hipStreamSynchronize(g_dev.stream));
    return std::sqrt((((sums[0] + sums[1]) + sums[2]) + sums[3]) + sums[4]);
}

// One implicit step from the current state U (R = g_dev.R holds R(U)).
// Returns the inner Newton iterations.
int implicit_step(int iter, double t, double dt, bool transient)
{
    ensure_implicit_buffers();
    hipStream_t s = g_dev.stream;
    const int N = g_dev.m.N, n = 5 * N;
    const int n_bdf = g_bdf_scheme == BDF2 ? 2 : 1;
    const int n_loc = iter < n_bdf ? 1 : n_bdf;
    const bool euler = iter < n_bdf || g_bdf_scheme != BDF2;
    const double a0 = euler ? 0.0 : -1. / 3., a1 = euler ? 0.0 : 1 / 3.;
    const double b = euler ? 1.0 : 2. / 3.;
    const double tpdt = t + dt;
    (void)tpdt; // the built-in hooks are time independent
    double *U = g_dev.U[g_dev.cur];
    double **old = g_dev.imp_old;
    for (int i = n_loc - 1; i > 0; --i)
        HIP_CHECK(hipMemcpyAsync(old[i], old[i - 1], (size_t)n * 8,
                                 hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(old[0], U, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    imp::bdf_rhs(s, n, U, old[0], old[1], n_loc, a0, a1, b, dt, g_dev.R, g_dev.imp_f);
    const double e0 = device_norm2(g_dev.imp_f, n);
    double e_old = e0;
    int inner;
    for (inner = 1; inner <= g_max_iter_inner; ++inner)
    {
        // lumped numerical Jacobian (implicit.c:61-120)
        double sq[5];
        {
            double *out = g_dev.imp_red + (size_t)1024 * 64;
            imp::column_sumsq(s, N, U, g_dev.imp_red, out);
            HIP_CHECK(hipMemcpyAsync(sq, out, sizeof(sq), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        for (int k = 0; k < 5; ++k)
        {
            const double eps = std::sqrt(sq[k]) * 1e-4;
            imp::shift_var(s, N, U, k, 0.5 * eps, g_dev.imp_Ut);
            eval_residual(g_dev.imp_Ut, g_dev.imp_Rp);
            imp::shift_var(s, N, U, k, -(0.5 * eps), g_dev.imp_Ut);
            eval_residual(g_dev.imp_Ut, g_dev.imp_Rm);
            imp::jac_column(s, N, g_dev.imp_Rp, g_dev.imp_Rm, k, b / (eps + 1e-16),
                            1.0 / dt, g_dev.imp_jac);
        }
        if (e_old >= e0)
        {
            g_last_lsoe = lsoe_solve(g_tolerance_lsoe * e_old);
            kry::update(s, n, U, g_dev.imp_dY, 1.0);
            eval_residual(U, g_dev.R);
            imp::bdf_rhs(s, n, U, old[0], old[1], n_loc, a0, a1, b, dt, g_dev.R,
                         g_dev.imp_f);
            e_old = device_norm2(g_dev.imp_f, n);
        }
        else
        {
            // the reference evaluates R at the state the Jacobian loop left
            // in phi_total (variable 4 shifted by -eps/2)
            eval_residual(g_dev.imp_Ut, g_dev.R);
        }
        if (e_old < e0)
            break;
        if (!transient)
            break;
    }
    if (inner >= g_max_iter_inner)
        throw std::runtime_error("implicit step: inner Newton iterations exhausted "
                                 "(max_iter_inner)");
    return inner;
}

// timedisc.c:134-222 with timestep_newton; host-driven (one synchronisation
// per step), bookkeeping on the device scalars like the explicit loop.
int run_implicit(int max_iter)
{
    require_ready();
    int iter0 = g_use_restart ? g_iter_restart : 0;
    double t0 = g_use_restart ? g_t_restart : g_t_start;
    print_header();
    if (iter0 == 0)
    {
        eval_residual(g_dev.U[g_dev.cur], g_dev.R);
        g_dev.r_valid = true;
    }
    if (g_do_output && !g_use_restart)
        write_output(iter0, t0);
    if (g_dev.log_cap < max_iter)
    {
        g_dev.log_cap = max_iter;
        g_dev.log = dmalloc<double>((size_t)g_dev.log_cap * 8);
    }
    g_log.assign((size_t)max_iter * 8, 0.0);
    init_scalars(iter0, t0, max_iter, g_transient, g_t_end);
    read_scalars();
    int status = 0, printed = 0;
    while (true)
    {
        dev::DevScalars h0 = *g_dev.sc_host;
        if (!builtin_timestep())
            host_timestep(h0);
        g_hook_time = h0.t + h0.dt;
        const int inner = implicit_step(h0.iter, h0.t, h0.dt, g_transient);
        g_dev.r_valid = true;
        g_dev.uprim = g_dev.cur;
        g_dev.restart_prims_valid = false;
        imp::step_partials(g_dev.stream, g_dev.m, phys(), g_dev.U[g_dev.cur], g_dev.R,
                           g_dev.partials);
        dev::launch_finalize(g_dev.stream, g_dev.sc, g_dev.partials,
                             dev::cell_blocks(g_dev.m.N), g_dev.fold, g_dev.fold_cap,
                             phys(), 1, g_dev.log, g_dev.log_cap);
        read_scalars();
        const dev::DevScalars &h = *g_dev.sc_host;
        if (h.done < 0)
        {
            status = FVH_ERR_NAN;
            break;
        }
        HIP_CHECK(hipMemcpy(&g_log[(size_t)printed * 8], &g_dev.log[(size_t)printed * 8],
                            8 * sizeof(double), hipMemcpyDeviceToHost));
        const int it = iter0 + printed + 1;
        const bool last = h.done > 0;
        const bool out = last || (g_i_output_data > 0 && it % g_i_output_data == 0);
        print_row(it, &g_log[(size_t)printed * 8], out, inner, g_last_lsoe);
        for (int j = 0; j < 5; ++j)
            g_residual[j] = g_log[(size_t)printed * 8 + 3 + j];
        g_viscous_dt = g_log[(size_t)printed * 8 + 2] != 0.0;
        ++printed;
        if (g_do_output && out)
            write_output(it, h.t);
        if (last)
            break;
    }
    g_last_iterations = printed;
    g_last_status = status;
    if (g_verbose)
        std::fflush(stdout);
    return status;
}

} // namespace

RuntimeView runtime_view(bool allocate)
{
    require_ready();
    if (!builtin_hooks())
        throw std::runtime_error(
            "the Newton-Krylov, learned-preconditioner, batched and allocation-free paths "
            "run the fused device residual and need the built-in hooks (calc_ns_flux, "
            "ns_update, ns_update_gradients)");
    if (allocate)
        ensure_gradient_buffers();
    RuntimeView v{};
    v.stream = g_dev.stream;
    v.m = g_dev.m;
    v.P = phys();
    v.flux = g_flux;
    v.linear = is_linear();
    v.viscous = is_viscous();
    v.U = g_dev.U[g_dev.cur];
    v.vglob = g_dev.vglob;
    const Regions_t *r = solver_mesh->regions;
    v.rho_ref = r->phi_total[(size_t)r->flow_region * 10 + 0];
    v.p_ref = r->phi_total[(size_t)r->flow_region * 10 + 8];
    if (!(v.rho_ref > 0.0) || !(v.p_ref > 0.0))
    {
        v.rho_ref = 1.0;
        v.p_ref = 1.0;
    }
    return v;
}

void eval_residual(const double *U, double *R, bool fp32, const float *U32)
{
    enqueue_residual(U, nullptr, R, 0, false, false, nullptr, nullptr, fp32, U32);
}

void state_changed() { g_dev.r_valid = false; }

StageBuffers stage_buffers()
{
    require_ready();
    if (g_rk_stages <= 0)
        throw std::runtime_error("explicit scheme not initialised");
    StageBuffers b{};
    b.U[0] = g_dev.U[0];
    b.U[1] = g_dev.U[1];
    b.cur = &g_dev.cur;
    b.k = g_dev.k;
    b.R = g_dev.R;
    b.stages = g_rk_stages;
    b.rk_a = g_rk_a;
    b.rk_g = g_rk_g;
    return b;
}

} // namespace fvh

using namespace fvh;

// ===========================================================================
// Solver API (reference names)
// ===========================================================================
extern "C" {

// ---- solver.c -------------------------------------------------------------
void init_solver(cstring_t title) { g_title = title ? title : "fv3d"; }
cstring_t get_simulation_title(void) { return g_title.c_str(); }
void free_solver(void) { g_title = "fv3d"; }

// ---- equation.c:37-169 ------------------------------------------------------
void init_equation(void)
{
    if (solver_variables)
        free_equation();
    solver_variables = xalloc<Variables_t>(1);
}

static void rebuild_tot_variables()
{
    Variables_t *v = solver_variables;
    std::free(v->tot_variables);
    v->n_tot_variables = v->n_sol_variables + v->n_dep_variables;
    v->tot_variables = xalloc<Variable_t *>(v->n_tot_variables);
    for (int i = 0; i < v->n_sol_variables; ++i)
        v->tot_variables[i] = &v->sol_variables[i];
    for (int i = 0; i < v->n_dep_variables; ++i)
        v->tot_variables[v->n_sol_variables + i] = &v->dep_variables[i];
}

int add_sol_variable(cstring_t name)
{
    Variables_t *v = solver_variables;
    if (!v)
        fatal("add_sol_variable before init_equation");
    v->sol_variables = (Variable_t *)std::realloc(
        v->sol_variables, sizeof(Variable_t) * (v->n_sol_variables + 1));
    v->sol_variables[v->n_sol_variables].name = xstrdup(name);
    v->n_sol_variables += 1;
    rebuild_tot_variables();
    return v->n_sol_variables - 1;
}

int add_dep_variable(cstring_t name)
{
    Variables_t *v = solver_variables;
    if (!v)
        fatal("add_dep_variable before init_equation");
    v->dep_variables = (Variable_t *)std::realloc(
        v->dep_variables, sizeof(Variable_t) * (v->n_dep_variables + 1));
    v->dep_variables[v->n_dep_variables].name = xstrdup(name);
    v->n_dep_variables += 1;
    rebuild_tot_variables();
    // dependent variables follow all solution variables (equation.c:153-169)
    return v->n_sol_variables + v->n_dep_variables - 1;
}

void free_equation(void)
{
    Variables_t *v = solver_variables;
    if (!v)
        return;
    for (int i = 0; i < v->n_sol_variables; ++i)
        std::free(v->sol_variables[i].name);
    for (int i = 0; i < v->n_dep_variables; ++i)
        std::free(v->dep_variables[i].name);
    std::free(v->sol_variables);
    std::free(v->dep_variables);
    std::free(v->tot_variables);
    std::free(v);
    solver_variables = nullptr;
}

void print_variables(void)
{
    Variables_t *v = solver_variables;
    if (!v)
        return;
    std::printf("Solution variables (%d):", v->n_sol_variables);
    for (int i = 0; i < v->n_sol_variables; ++i)
        std::printf(" %s", v->sol_variables[i].name);
    std::printf("\nDependent variables (%d):", v->n_dep_variables);
    for (int i = 0; i < v->n_dep_variables; ++i)
        std::printf(" %s", v->dep_variables[i].name);
    std::printf("\n");
}

// ---- mesh.c -------------------------------------------------------------------
void init_mesh(cstring_t mesh_file)
{
    try
    {
        RawMesh m;
        std::string p = mesh_file ? mesh_file : "";
        read_mesh_file(m, p.c_str());
        install_mesh(std::move(m));
    }
    catch (const std::exception &e)
    {
        fatal(e.what());
    }
}

void free_mesh(void)
{
    free_mesh_struct(solver_mesh);
    solver_mesh = nullptr;
    g_face_shift.clear();
}

void print_mesh_info(void)
{
    Mesh_t *m = solver_mesh;
    if (!m)
        return;
    std::printf("Mesh: dimension %d, cells %d, faces %d (internal %d, "
                "boundary %d), boundaries %d, regions %d, volume %.6e\n",
                m->dimension, m->cells->n_domain_cells, m->faces->n_faces,
                m->faces->n_internal_faces, m->faces->n_boundary_faces,
                m->boundaries->n_boundaries, m->regions->n_regions,
                m->global_volume);
    for (int i = 0; i < m->regions->n_regions; ++i)
        std::printf("  region %d: %s%s\n", i, m->regions->name[i],
                    m->regions->is_boundary[i] ? " (boundary)" : "");
}

// ---- finite_volume.c ------------------------------------------------------------
void set_calc_flux(void_calc_flux_ft fun_ptr) { g_calc_flux = fun_ptr; }
void set_exact_function(void_calc_exact_ft fun_ptr) { g_calc_exact = fun_ptr; }
void set_update(void_update_ft fun_ptr) { g_update = fun_ptr; }
void set_update_gradients(void_update_gradients_ft fun_ptr)
{
    g_update_gradients = fun_ptr;
}

// finite_volume.c:23-55 (host mirror) + device buffers; finite_volume.c:189-201
void init_finite_volume(void)
{
    try
    {
        if (!solver_mesh || !solver_variables)
            fatal("init_finite_volume needs the mesh and the variables");
        if (solver_variables->n_sol_variables != 5 ||
            solver_variables->n_tot_variables != 10)
            fatal("the device path implements the 5 + 5 Navier-Stokes variables");
        if (!solver_mesh->regions->phi_total)
            fatal("boundary regions not initialised (navier_stokes_initialize)");
        if (g_dev.ready)
            free_finite_volume();

        const Cells_t *cells = solver_mesh->cells;
        const Faces_t *faces = solver_mesh->faces;
        const Boundaries_t *bnd = solver_mesh->boundaries;
        const Regions_t *reg = solver_mesh->regions;
        const int N = cells->n_domain_cells, F = faces->n_faces,
                  B = bnd->n_boundaries, maxf = cells->max_cell_faces;

        // host mirror: state rows only; gradients, face states and fluxes
        // live on the device (they are recomputed per face there)
        solver_data = xalloc<data_t>(1);
        solver_data->phi_total = xalloc<double>((size_t)(N + B) * 10);
        solver_data->phi_dt = xalloc<double>((size_t)(N + B) * 5);
        // finite_volume.c:39-52 (filled by finite_volume_time_derivative and
        // before host hooks; zero pages until then)
        solver_data->flux = xalloc<double>((size_t)F * 5);
        solver_data->phi_total_left = xalloc<double>((size_t)F * 10);
        solver_data->phi_total_right = xalloc<double>((size_t)F * 10);
        solver_data->grad_phi_total_x = xalloc<double>((size_t)(N + B) * 10);
        solver_data->grad_phi_total_y = xalloc<double>((size_t)(N + B) * 10);
        solver_data->grad_phi_total_z = xalloc<double>((size_t)(N + B) * 10);

        HIP_CHECK(hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking));
        dev::DevMesh &m = g_dev.m;
        m.N = N;
        m.F = F;
        m.B = B;
        m.maxf = maxf;
        // cell face lists and face cells keep the reference's AoS layout
        m.cf = dupload(cells->faces, (size_t)N * maxf);
        m.ncf = dupload(cells->n_faces, N);
        m.fcell = dupload(faces->cells, (size_t)F * 2);
        {
            // the stencil kernels' face list: face and neighbour per slot
            std::vector<int2> cfn((size_t)N * maxf);
            for (int c = 0; c < N; ++c)
            {
                const int nf = cells->n_faces[c];
                for (int s = 0; s < maxf; ++s)
                {
                    const int f = cells->faces[(size_t)c * maxf + (s < nf ? s : 0)];
                    const int c0 = faces->cells[2 * f], c1 = faces->cells[2 * f + 1];
                    cfn[(size_t)c * maxf + s] = c0 == c ? make_int2(f, c1) : make_int2(f, -(c0 + 1));
                }
            }
            m.cfn = dupload(cfn.data(), cfn.size());
            // tetrahedra: the face ids with the side in the sign bit (the
            // tile gradient kernel reads 16 B per cell instead of cfn's 32)
            m.cfo = nullptr;
            if (maxf == 4)
            {
                std::vector<int> cfo(cfn.size());
                for (size_t i = 0; i < cfn.size(); ++i)
                    cfo[i] = cfn[i].y >= 0 ? cfn[i].x : (int)((unsigned)cfn[i].x | 0x80000000u);
                m.cfo = dupload(cfo.data(), cfo.size());
            }
            // the neighbour's slot of each slot's face (a cell that is its
            // own periodic neighbour takes the face's other slot)
            std::vector<unsigned char> nsl((size_t)N * maxf, 0);
            for (int c = 0; c < N; ++c)
            {
                const int nf = cells->n_faces[c];
                for (int s = 0; s < nf; ++s)
                {
                    const int f = cells->faces[(size_t)c * maxf + s];
                    const int c0 = faces->cells[2 * f], c1 = faces->cells[2 * f + 1];
                    const int nb = c0 == c ? c1 : c0;
                    if (nb < 0 || nb >= N)
                        continue;
                    int t = 0;
                    while (t < cells->n_faces[nb] &&
                           (cells->faces[(size_t)nb * maxf + t] != f || (nb == c && t == s)))
                        ++t;
                    if (t == cells->n_faces[nb])
                        fatal("face " + std::to_string(f) + " is missing from the face list of cell " +
                              std::to_string(nb));
                    nsl[(size_t)c * maxf + s] = (unsigned char)t;
                }
            }
            m.nsl = dupload(nsl.data(), nsl.size());
        }
        {
            // face records; periodic shifts as indices into a table of the
            // distinct exact shift vectors (entry 0 = no shift)
            std::vector<double> a((size_t)F * dev::kFgA), b((size_t)F * dev::kFgB);
            std::vector<int> fsh(F, 0);
            std::vector<double> shtab = {0.0, 0.0, 0.0};
            for (int f = 0; f < F; ++f)
            {
                double *ra = &a[(size_t)f * dev::kFgA];
                double *rb = &b[(size_t)f * dev::kFgB];
                for (int d = 0; d < 3; ++d)
                {
                    ra[d] = faces->n[3 * f + d];
                    ra[3 + d] = faces->t1[3 * f + d];
                    ra[6 + d] = faces->t2[3 * f + d];
                    rb[d] = faces->n[3 * f + d];
                }
                ra[9] = faces->area[f];
                rb[3] = faces->area[f];
                rb[4] = faces->lambda[f];
                const double *sh = &g_face_shift[(size_t)3 * f];
                if (sh[0] != 0.0 || sh[1] != 0.0 || sh[2] != 0.0)
                {
                    int k = 1;
                    const int nsh = (int)shtab.size() / 3;
                    for (; k < nsh; ++k)
                        if (std::memcmp(&shtab[3 * k], sh, 3 * sizeof(double)) == 0)
                            break;
                    if (k == nsh)
                        shtab.insert(shtab.end(), sh, sh + 3);
                    fsh[f] = k;
                }
            }
            m.fgA = dupload(a.data(), a.size());
            m.fgB = dupload(b.data(), b.size());
            // 32-B records (n, area) and the interpolation weights apart:
            // the tile gradient kernel reads 40 B per face slot instead of
            // 48, and the face-centred kernel 32 B per face instead of 80
            // when every stored t1 and t2 is bit for bit what it forms from
            // n (the generated and Gmsh meshes; a file may hold other
            // tangents).  FVH_FULL_FRAMES=1 keeps the face-centred kernel
            // on fgA, 2 both kernels on fgA / fgB.
            const char *full_env = std::getenv("FVH_FULL_FRAMES");
            const int full = full_env ? std::atoi(full_env) : 0;
            bool derivable = full == 0;
            for (int f = 0; f < F && derivable; ++f)
            {
                double t1[3], t2[3];
                dev::tangents_from_normal(&faces->n[3 * f], t1, t2);
                derivable = std::memcmp(t1, &faces->t1[3 * f], sizeof t1) == 0 &&
                            std::memcmp(t2, &faces->t2[3 * f], sizeof t2) == 0;
            }
            m.tangents_derived = derivable ? 1 : 0;
            m.fgN = nullptr;
            m.flam = nullptr;
            if (full < 2)
            {
                std::vector<double> nrec((size_t)F * 4);
                for (int f = 0; f < F; ++f)
                {
                    for (int d = 0; d < 3; ++d)
                        nrec[(size_t)f * 4 + d] = faces->n[3 * f + d];
                    nrec[(size_t)f * 4 + 3] = faces->area[f];
                }
                m.fgN = dupload(nrec.data(), nrec.size());
                m.flam = dupload(faces->lambda, (size_t)F);
            }
            else
                m.tangents_derived = 0;
            m.fsh = dupload(fsh.data(), fsh.size());
            m.shtab = dupload(shtab.data(), shtab.size());
        }
        m.cx = dupload(cells->x, (size_t)N * 3);
        {
            // the gradient kernel's distance vectors per face slot
            const std::vector<double> crv = slot_distances(maxf);
            m.crv = dupload(crv.data(), crv.size());
        }
        m.vol = dupload(cells->volume, N);
        {
            std::vector<double> sv(N);
            for (int c = 0; c < N; ++c)
                sv[c] = 1.0 / cells->volume[c];
            m.svol = dupload(sv.data(), N);
        }
        m.cdx = dupload(cells->dx, (size_t)N * 3);
        m.bnd_region = dupload(bnd->id, B);
        m.bnd_face = dupload(bnd->face, B);
        m.region_type = dupload(reg->type, reg->n_regions);
        m.region_phi = dupload(reg->phi_total, (size_t)reg->n_regions * 10);

        g_dev.U[0] = dmalloc<double>((size_t)5 * N);
        g_dev.U[1] = dmalloc<double>((size_t)5 * N);
        g_dev.k = dmalloc<double>((size_t)5 * N);
        g_dev.R = dmalloc<double>((size_t)5 * N);
        g_dev.partials_cap = std::max(dev::residual_grid(m), dev::dt_partials_grid(m));
        g_dev.partials = dmalloc<double>((size_t)g_dev.partials_cap *
                                         dev::kPartialStride);
        g_dev.fold_cap = dev::fold_rows(g_dev.partials_cap);
        g_dev.fold = dmalloc<double>((size_t)g_dev.fold_cap * dev::kPartialStride);
        g_dev.gphi = dmalloc<double>((size_t)10 * B);
        g_dev.sc = dmalloc<dev::DevScalars>(1);
        HIP_CHECK(hipHostMalloc((void **)&g_dev.sc_host, sizeof(dev::DevScalars),
                                hipHostMallocDefault));
        g_dev.mirror_phi = dmalloc<double>((size_t)(N + B) * 10);
        g_dev.mirror_dt = dmalloc<double>((size_t)N * 5);
        g_dev.cur = 0;
        g_dev.vglob = solver_mesh->global_volume;
        g_dev.ready = true;
        noalloc_setup();
        setup_tile_halo();
        setup_tile_faces();

        // set_solution (finite_volume.c:189-201): exact function per domain
        // cell on the host, then the update hook's work on the device
        if (!g_calc_exact)
            fatal("no exact function set");
        for (int i = 0; i < N; ++i)
            g_calc_exact(0, 0.0, &cells->x[(size_t)i * DIM],
                         &solver_data->phi_total[(size_t)i * 10]);
        upload_state_from_mirror();
        sync_mirror(false);
    }
    catch (const std::exception &e)
    {
        fatal(e.what());
    }
}

void free_finite_volume(void)
{
    g_calc_flux = nullptr;
    g_calc_exact = nullptr;
    g_update = nullptr;
    g_update_gradients = nullptr;
    free_device();
    if (solver_data)
    {
        std::free(solver_data->phi_total);
        std::free(solver_data->phi_dt);
        std::free(solver_data->flux);
        std::free(solver_data->phi_total_left);
        std::free(solver_data->phi_total_right);
        std::free(solver_data->grad_phi_total_x);
        std::free(solver_data->grad_phi_total_y);
        std::free(solver_data->grad_phi_total_z);
        std::free(solver_data);
        solver_data = nullptr;
    }
}

// finite_volume.c:81-124 on the device; refreshes the host mirror
void finite_volume_time_derivative(double time)
{
    (void)time;
    try
    {
        require_ready();
        g_hook_time = time;
        g_dev.uprim = g_dev.cur;
        g_dev.restart_prims_valid = false;
        if (builtin_hooks())
        {
            enqueue_residual(g_dev.U[g_dev.cur], nullptr, g_dev.R, 0, false, false,
                             nullptr);
            // the reference's residual also leaves its face states, gradients
            // and fluxes in solver_data: the host-mirror passes fill them
            ensure_mirror_buffers();
            mirror_residual(time, g_dev.U[g_dev.cur], g_dev.mres, true);
        }
        else
            mirror_residual(time, g_dev.U[g_dev.cur], g_dev.R, true);
        g_dev.r_valid = true;
        sync_mirror(true);
    }
    catch (const std::exception &e)
    {
        fatal(e.what());
    }
}

// ---- reconstruction.c / limiter.c --------------------------------------------
void init_reconstruction(reconstruction_type_t t)
{
    if (t != FirstOrder && t != Linear)
        fatal("init_reconstruction: unknown type");
    g_reconstruction = t;
}
void free_reconstruction(void) { g_reconstruction = -1; }
void init_limiter(limiter_type_t t)
{
    if (t != BarthJespersenn && t != NoLimter)
        fatal("init_limiter: unknown type");
    g_limiter = t;
}
void free_limiter(void) { g_limiter = -1; }

// ---- analyze.c ----------------------------------------------------------------
void init_analyze(void)
{
    g_analyze_ready = true;
    std::fill(g_residual.begin(), g_residual.end(), 0.0);
}
void free_analyze(void) { g_analyze_ready = false; }

// analyze.c:37-61 from the last residual on the device
void calc_global_residual(double dt)
{
    try
    {
        require_ready();
        if (!g_dev.r_valid)
            fatal("calc_global_residual: no residual evaluated yet");
        std::vector<double> R((size_t)5 * g_dev.m.N);
        HIP_CHECK(hipMemcpy(R.data(), g_dev.R, R.size() * 8,
                            hipMemcpyDeviceToHost));
        // ordered host sum over the device residual (a reporting path; the
        // time loop uses the device reduction)
        double tmp[5] = {0, 0, 0, 0, 0};
        const int N = g_dev.m.N;
        for (int i = 0; i < N; ++i)
        {
            double v = solver_mesh->cells->volume[i];
            for (int j = 0; j < 5; ++j)
            {
                double r = R[(size_t)i * 5 + j];
                tmp[j] += (r < 0 ? -r : r) * v;
            }
        }
        for (int j = 0; j < 5; ++j)
            g_residual[j] = tmp[j] * (dt / solver_mesh->global_volume);
    }
    catch (const std::exception &e)
    {
        fatal(e.what());
    }
}

// ---- timedisc.c / explicit.c / implicit.c --------------------------------------
void init_timedisc(timedisc_type_t type, int max_iter, bool_t is_transient,
                   double abort_residual, double t_start, double t_end)
{
    if (type == Explicit)
        g_is_explicit = true;
    else if (type == Implicit)
        g_is_explicit = false;
    else
        fatal("init_timedisc: unknown type");
    g_max_iter = max_iter;
    g_transient = is_transient != 0;
    g_abort_residual = abort_residual;
    g_t_start = t_start;
    g_t_end = g_transient ? t_end : DBL_MAX;
}

void free_timedisc(void) { g_calc_timestep = nullptr; }
bool_t is_explicit(void) { return g_is_explicit ? BC_TRUE : BC_FALSE; }
void set_calc_timestep(double_calc_timestep_ft fun_ptr) { g_calc_timestep = fun_ptr; }
void set_viscous_dt(bool_t v) { g_viscous_dt = v != 0; }

void init_explicit(explicit_scheme_t scheme)
{
    switch (scheme)
    {
    case EulerExplicit:
        g_rk_stages = 1, g_rk_a = rk_a_euler, g_rk_b = rk_b_euler,
        g_rk_g = rk_g_euler;
        break;
    case RungeKutta33:
        g_rk_stages = 3, g_rk_a = rk_a_rk33, g_rk_b = rk_b_rk33,
        g_rk_g = rk_g_rk33;
        break;
    case RungeKutta45:
        g_rk_stages = 5, g_rk_a = rk_a_rk45, g_rk_b = rk_b_rk45,
        g_rk_g = rk_g_rk45;
        break;
    default:
        fatal("init_explicit: unknown scheme");
    }
}

void free_explicit(void)
{
    g_rk_stages = 0;
    g_rk_a = g_rk_b = g_rk_g = nullptr;
}

void init_implicit(implicit_scheme_t implicit_scheme,
                   implicit_solver_t implicit_solver, int max_iter_inner,
                   double tolerance_lsoe, int max_iter_lsoe,
                   int max_krylov_dims, int max_krylov_restarts)
{
    if (implicit_scheme != EulerImplicit && implicit_scheme != BDF2)
        fatal("init_implicit: unknown scheme");
    if (implicit_solver != BiCGStab && implicit_solver != GMRes)
        fatal("init_implicit: unknown solver");
    if (max_krylov_dims < 1 || max_krylov_dims > 63)
        fatal("init_implicit: max_krylov_dims must be in 1..63");
    g_bdf_scheme = implicit_scheme;
    g_implicit_solver = implicit_solver;
    g_max_iter_inner = max_iter_inner;
    g_tolerance_lsoe = tolerance_lsoe;
    g_max_iter_lsoe = max_iter_lsoe;
    g_max_krylov_dims = max_krylov_dims;
    g_max_krylov_restarts = max_krylov_restarts;
}

void free_implicit(void) {}

void timedisc(void)
{
    int rc;
    try
    {
        rc = g_is_explicit ? run_explicit(g_max_iter) : run_implicit(g_max_iter);
    }
    catch (const std::exception &e)
    {
        fatal(e.what());
    }
    if (rc == FVH_ERR_NAN)
        fatal("timedisc: NaN or Inf residual");
}

// ---- output.c / restart.c --------------------------------------------------------
void init_output(int i_output_data)
{
    g_i_output_data = i_output_data;
    g_do_output = i_output_data != 0;
}
void free_output(void) {}

namespace
{
bool output_hdf5()
{
    if (g_output_format == 1)
        return false;
    std::string why;
    const bool ok = hdf5_available(&why);
    if (g_output_format == 2 && !ok)
        throw std::runtime_error("HDF5 output requested: " + why);
    return ok;
}

// output.c:20-67 (the header part of every output file)
void write_header(DataWriter &w)
{
    w.attr_int("n_tot_variables", solver_variables->n_tot_variables);
    w.attr_int("n_sol_variables", solver_variables->n_sol_variables);
    std::vector<std::string> tot, sol;
    for (int i = 0; i < solver_variables->n_tot_variables; ++i)
        tot.push_back(solver_variables->tot_variables[i]->name);
    for (int i = 0; i < solver_variables->n_sol_variables; ++i)
        sol.push_back(solver_variables->sol_variables[i].name);
    w.strings("tot_variables", tot);
    w.strings("sol_variables", sol);
}
} // namespace

void create_file_header(cstring_t file_name)
{
    try
    {
        std::unique_ptr<DataWriter> w = open_writer(file_name);
        write_header(*w);
    }
    catch (const std::exception &e)
    {
        fatal(e.what());
    }
}

// output.c:92-191: <title>.%09d.h5 (raw container: .fvs) with the header,
// attrs iter, t and datasets phi_total [N x 10], phi_dt [N x 5]; in implicit
// mode also n_stages and the BDF history phi_old:<i> [N x 5]
void write_output(int iter, double t)
{
    try
    {
        sync_mirror(true);
        char suffix[64];
        const bool h5 = output_hdf5();
        std::snprintf(suffix, sizeof(suffix), h5 ? ".%09d.h5" : ".%09d.fvs", iter);
        std::string path = g_title + suffix;
        std::unique_ptr<DataWriter> wp = h5 ? hdf5_writer(path.c_str())
                                            : std::unique_ptr<DataWriter>(new RawWriter(path.c_str()));
        DataWriter &w = *wp;
        write_header(w);
        w.attr_int("iter", iter);
        w.attr_double("t", t);
        const int N = solver_mesh->cells->n_domain_cells;
        w.doubles("phi_total", solver_data->phi_total, (size_t)N * 10, 10);
        w.doubles("phi_dt", solver_data->phi_dt, (size_t)N * 5, 5);
        if (!g_is_explicit && g_dev.imp_old[0])
        {
            const int n_stages = g_bdf_scheme == BDF2 ? 2 : 1;
            w.attr_int("n_stages", n_stages);
            std::vector<double> h((size_t)N * 5);
            for (int i = 0; i < n_stages; ++i)
            {
                HIP_CHECK(hipMemcpy(h.data(), g_dev.imp_old[i], h.size() * 8,
                                    hipMemcpyDeviceToHost));
                w.doubles("phi_old:" + std::to_string(i), h.data(), h.size(), 5);
            }
        }
    }
    catch (const std::exception &e)
    {
        fatal(e.what());
    }
}

void init_restart(bool_t use_restart, cstring_t restart_file)
{
    g_use_restart = use_restart != 0;
    g_restart_file = restart_file ? restart_file : "";
    if (g_use_restart)
        read_restart_data(g_restart_file.c_str());
}
void free_restart(void) { g_use_restart = false; }

// restart.c:50-178 with the serial dataset names the writer uses
void read_restart_data(cstring_t restart_file)
{
    try
    {
        std::unique_ptr<DataReader> rp = open_reader(restart_file);
        const DataReader &r = *rp;
        g_iter_restart = r.ints("iter")[0];
        g_t_restart = r.doubles("t")[0];
        std::vector<double> phi = r.doubles("phi_total");
        const int N = solver_mesh->cells->n_domain_cells;
        if (phi.size() != (size_t)N * 10)
            fatal("restart file does not match the mesh");
        if (g_dev.ready)
        {
            // phi_total as stored: the conservative columns become the
            // state; the primitive columns (those of the last stage's input
            // when the file was written after an RK step) stay in the host
            // mirror and give the first time step, as in the reference
            std::memcpy(solver_data->phi_total, phi.data(), phi.size() * 8);
            upload_state_from_mirror();
            std::vector<double> pr((size_t)N * 5);
            for (int c = 0; c < N; ++c)
                for (int j = 0; j < 5; ++j)
                    pr[(size_t)c * 5 + j] = phi[(size_t)c * 10 + 5 + j];
            if (!g_dev.restart_prims)
                g_dev.restart_prims = dmalloc<double>((size_t)N * 5);
            HIP_CHECK(hipMemcpy(g_dev.restart_prims, pr.data(), pr.size() * 8,
                                hipMemcpyHostToDevice));
            g_dev.restart_prims_valid = true;
            // phi_dt: the residual of the stored state (the implicit step's
            // first right-hand side)
            if (r.has("phi_dt"))
            {
                std::vector<double> dt = r.doubles("phi_dt");
                if (dt.size() != (size_t)N * 5)
                    fatal("restart file: phi_dt does not match the mesh");
                std::memcpy(solver_data->phi_dt, dt.data(), dt.size() * 8);
                HIP_CHECK(hipMemcpy(g_dev.R, dt.data(), dt.size() * 8, hipMemcpyHostToDevice));
                g_dev.r_valid = true;
            }
            // the BDF history of the implicit scheme (restart.c:87-110)
            if (!g_is_explicit && r.has("n_stages"))
            {
                const int ns = r.ints("n_stages")[0];
                const int want = g_bdf_scheme == BDF2 ? 2 : 1;
                if (ns != want)
                    fatal("restart file: n_stages does not match the implicit scheme");
                ensure_implicit_buffers();
                for (int i = 0; i < ns; ++i)
                {
                    std::vector<double> o = r.doubles("phi_old:" + std::to_string(i));
                    if (o.size() != (size_t)N * 5)
                        fatal("restart file: phi_old does not match the mesh");
                    HIP_CHECK(hipMemcpy(g_dev.imp_old[i], o.data(), o.size() * 8,
                                        hipMemcpyHostToDevice));
                }
            }
        }
    }
    catch (const std::exception &e)
    {
        fatal(e.what());
    }
}

// ---- Navier-Stokes plug-in built-ins ------------------------------------------
// navier_stokes.c:85-124 without the parameter parsing
void navier_stokes_initialize(void)
{
    kappa_m1 = kappa - 1.0;
    kappa_p1 = kappa + 1.0;
    s_kappa = 1.0 / kappa;
    s_kappa_m1 = 1.0 / kappa_m1;
    s_kappa_p1 = 1.0 / kappa_p1;
    cv = R_mix * s_kappa_m1;
    cp = kappa * cv;
    kappa_pr = kappa / Pr;
    lambda = mu_mix * cp / Pr;

    init_equation();
    ic_rho = add_sol_variable("rho");
    ic_rho_u = add_sol_variable("rho_u");
    ic_rho_v = add_sol_variable("rho_v");
    ic_rho_w = add_sol_variable("rho_w");
    ic_rho_e = add_sol_variable("rho_e");
    ip_u = add_dep_variable("u");
    ip_v = add_dep_variable("v");
    ip_w = add_dep_variable("w");
    ip_p = add_dep_variable("p");
    ip_T = add_dep_variable("T");

    set_calc_flux(calc_ns_flux);
    set_calc_timestep(calc_ns_timestep);
    set_exact_function(calc_ns_exact_function);
    set_update(ns_update);
    set_update_gradients(ns_update_gradients);

    if (solver_mesh)
    {
        Regions_t *r = solver_mesh->regions;
        std::free(r->type), std::free(r->function_id), std::free(r->phi_total);
        r->type = xalloc<int>(r->n_regions);
        r->function_id = xalloc<int>(r->n_regions);
        r->phi_total = xalloc<double>((size_t)r->n_regions * 10);
    }
}

// boundary.c:76-121 + parse_primitive_state boundary.c:144-183
void boundary_set_region(int region, int type, const double *prim5)
{
    if (!solver_mesh || !solver_mesh->regions->phi_total)
        fatal("boundary_set_region before navier_stokes_initialize");
    Regions_t *r = solver_mesh->regions;
    if (region < 0 || region >= r->n_regions)
        fatal("boundary_set_region: region out of range");
    if (type == BoundaryFunction)
        fatal("boundary type 'function' is not reachable in the reference "
              "(boundary.c:127 vs navier_stokes.c:142-150)");
    if (type < 0 || type >= BoundaryTypeMax)
        fatal("boundary_set_region: unknown type");
    if (type == BoundaryFlow && region != r->flow_region)
        fatal("the flow state belongs to the flow region");
    r->type[region] = type;
    double *phi = &r->phi_total[(size_t)region * 10];
    if (type == BoundaryFlow || type == BoundaryInflow || type == BoundaryState)
    {
        for (int j = 0; j < 10; ++j)
            phi[j] = 0.0;
        phi[ip_u] = prim5[0];
        phi[ip_v] = prim5[1];
        phi[ip_w] = prim5[2];
        phi[ip_p] = prim5[3];
        phi[ip_T] = prim5[4];
        phi[ic_rho] = phi[ip_p] / (R_mix * phi[ip_T]);
        phi[ic_rho_u] = phi[ip_u] * phi[ic_rho];
        phi[ic_rho_v] = phi[ip_v] * phi[ic_rho];
        phi[ic_rho_w] = phi[ip_w] * phi[ic_rho];
        double d = 0.0;
        d += phi[ip_u] * phi[ic_rho_u];
        d += phi[ip_v] * phi[ic_rho_v];
        d += phi[ip_w] * phi[ic_rho_w];
        phi[ic_rho_e] = s_kappa_m1 * phi[ip_p] + 0.5 * d;
    }
    else if (type == BoundaryIsothermalWall)
    {
        phi[ip_T] = prim5[4];
    }
}

// flux.c:295-312
void flux_initialize_scheme(cstring_t name)
{
    std::string s = name ? name : "";
    if (s == "Rusanov")
        g_flux = dev::FluxRusanov;
    else if (s == "AUSM")
        g_flux = dev::FluxAUSM;
    else
        fatal("unknown flux scheme " + s);
}

// The hooks below are the identities the solver dispatches on; the work of
// each is fused into the device residual kernels.
void calc_ns_flux(double time) { (void)time; }
void ns_update(double time) { (void)time; }
void ns_update_gradients(double time) { (void)time; }

// navier_stokes.c:159-193 on the device (current primitives)
double calc_ns_timestep(double time)
{
    (void)time;
    try
    {
        require_ready();
        dev::launch_dt_partials(g_dev.stream, g_dev.m, phys(),
                                g_dev.U[g_dev.cur], g_dev.partials,
                                g_dev.partials_cap);
        dev::DevScalars &h = *g_dev.sc_host;
        std::memset(&h, 0, sizeof(h));
        h.t_end = DBL_MAX;
        HIP_CHECK(hipMemcpyAsync(g_dev.sc, &h, sizeof(h), hipMemcpyHostToDevice,
                                 g_dev.stream));
        dev::launch_finalize(g_dev.stream, g_dev.sc, g_dev.partials, dev::dt_partials_grid(g_dev.m), g_dev.fold, g_dev.fold_cap, phys(), 0, nullptr, 0);
        read_scalars();
        set_viscous_dt(h.viscous);
        return h.dt;
    }
    catch (const std::exception &e)
    {
        fatal(e.what());
    }
}

// navier_stokes.c:133-152
void calc_ns_exact_function(int id, double t, double *x, double *phi)
{
    (void)t;
    (void)x;
    if (id != BoundaryFlow)
        fatal("calc_ns_exact_function: only the flow region state is defined");
    Regions_t *r = solver_mesh->regions;
    for (int j = 0; j < 10; ++j)
        phi[j] = r->phi_total[(size_t)r->flow_region * 10 + j];
}

// ===========================================================================
// fvh_* extension API
// ===========================================================================
const char *fvh_last_error(void) { return g_error.c_str(); }

#define FVH_TRY(...)                                                           \
    try                                                                        \
    {                                                                          \
        __VA_ARGS__;                                                           \
    }                                                                          \
    catch (const std::exception &e)                                            \
    {                                                                          \
        set_error(e.what());                                                   \
        return FVH_ERR_STATE;                                                  \
    }

int fvh_mesh_box(int nx, int ny, int nz, double x0, double y0, double z0,
                 double lx, double ly, double lz, int tets, int periodic,
                 int tile)
{
    FVH_TRY({
        RawMesh m;
        generate_box(m, nx, ny, nz, x0, y0, z0, lx, ly, lz, tets != 0,
                     periodic, tile);
        install_mesh(std::move(m));
    });
    return FVH_OK;
}

int fvh_mesh_channel(int nx, int ny, int nz, double jitter,
                     unsigned long long seed)
{
    FVH_TRY({
        RawMesh m;
        generate_channel(m, nx, ny, nz, jitter, seed);
        install_mesh(std::move(m));
    });
    return FVH_OK;
}

int fvh_mesh_gmsh(cstring_t path, double sx, double sy, double sz)
{
    FVH_TRY({
        if (!path)
            throw std::runtime_error("fvh_mesh_gmsh: null path");
        RawMesh m;
        const double scale[3] = {sx, sy, sz};
        read_gmsh(m, path, scale);
        install_mesh(std::move(m));
    });
    return FVH_OK;
}

int fvh_mesh_write(cstring_t path)
{
    FVH_TRY({
        if (!solver_mesh)
            throw std::runtime_error("no mesh");
        // reconstruct the file layout (second face cell -1 on boundaries)
        const Mesh_t *mm = solver_mesh;
        RawMesh m;
        m.dimension = mm->dimension;
        const int N = mm->cells->n_domain_cells, F = mm->faces->n_faces,
                  B = mm->boundaries->n_boundaries;
        m.vertices.assign(mm->vertices->x, mm->vertices->x + 3 * mm->vertices->n_vertices);
        m.max_cell_vertices = mm->cells->max_cell_vertices;
        m.max_cell_faces = mm->cells->max_cell_faces;
        m.max_face_vertices = mm->faces->max_face_vertices;
        m.max_boundary_vertices = mm->boundaries->max_boundary_vertices;
        auto vi = [](const int *p, size_t n) { return std::vector<int>(p, p + n); };
        auto vd = [](const double *p, size_t n) { return std::vector<double>(p, p + n); };
        m.cell_region = vi(mm->cells->id, N);
        m.cell_type = vi(mm->cells->type, N);
        m.cell_nv = vi(mm->cells->n_vertices, N);
        m.cell_vertices = vi(mm->cells->vertices, (size_t)N * m.max_cell_vertices);
        m.cell_nf = vi(mm->cells->n_faces, N);
        m.cell_faces = vi(mm->cells->faces, (size_t)N * m.max_cell_faces);
        m.cell_volume = vd(mm->cells->volume, N);
        m.cell_x = vd(mm->cells->x, (size_t)N * 3);
        m.cell_dx = vd(mm->cells->dx, (size_t)N * 3);
        m.face_type = vi(mm->faces->type, F);
        m.face_nv = vi(mm->faces->n_vertices, F);
        m.face_vertices = vi(mm->faces->vertices, (size_t)F * m.max_face_vertices);
        m.face_cells = vi(mm->faces->cells, (size_t)F * 2);
        for (int f = 0; f < F; ++f)
            if (mm->faces->boundary[f] >= 0)
                m.face_cells[2 * f + 1] = -1;
        m.face_boundary = vi(mm->faces->boundary, F);
        m.face_x = vd(mm->faces->x, (size_t)F * 3);
        m.face_n = vd(mm->faces->n, (size_t)F * 3);
        m.face_t1 = vd(mm->faces->t1, (size_t)F * 3);
        m.face_t2 = vd(mm->faces->t2, (size_t)F * 3);
        m.face_area = vd(mm->faces->area, F);
        m.face_lambda = vd(mm->faces->lambda, F);
        m.face_shift = g_face_shift;
        m.bnd_region = vi(mm->boundaries->id, B);
        m.bnd_type = vi(mm->boundaries->type, B);
        m.bnd_nv = vi(mm->boundaries->n_vertices, B);
        m.bnd_vertices = vi(mm->boundaries->vertices, (size_t)B * m.max_boundary_vertices);
        m.bnd_face = vi(mm->boundaries->face, B);
        m.bnd_n = vd(mm->boundaries->n, (size_t)B * 3);
        m.bnd_t1 = vd(mm->boundaries->t1, (size_t)B * 3);
        m.bnd_t2 = vd(mm->boundaries->t2, (size_t)B * 3);
        m.bnd_distance = vd(mm->boundaries->distance, B);
        for (int i = 0; i < mm->regions->n_regions; ++i)
            m.region_names.push_back(mm->regions->name[i]);
        m.region_is_boundary = vi(mm->regions->is_boundary, mm->regions->n_regions);
        if (write_mesh_file(m, path) != 0)
            throw std::runtime_error("cannot write mesh");
    });
    return FVH_OK;
}

int fvh_get_face_shift(double *shift)
{
    if (!solver_mesh)
    {
        set_error("no mesh");
        return FVH_ERR_STATE;
    }
    std::memcpy(shift, g_face_shift.data(), g_face_shift.size() * sizeof(double));
    return FVH_OK;
}

int fvh_set_visc_dt_mode(int mode)
{
    if (mode != 0 && mode != 1)
        return FVH_ERR_ARG;
    g_visc_dt_mode = mode;
    return FVH_OK;
}

int fvh_set_ic(int kind, const double *params, int n_params)
{
    if (kind < 0 || kind > 2 || n_params < 0 || n_params > 8)
        return FVH_ERR_ARG;
    g_ic_kind = kind;
    for (int i = 0; i < n_params; ++i)
        g_ic_params[i] = params[i];
    return FVH_OK;
}

void fvh_exact_function_builtin(int id, double t, double *x, double *phi)
{
    builtin_exact(id, t, x, phi);
}

int fvh_get_state(double *U)
{
    FVH_TRY({
        if (!g_dev.ready)
            throw std::runtime_error("no device state");
        HIP_CHECK(hipStreamSynchronize(g_dev.stream));
        HIP_CHECK(hipMemcpy(U, g_dev.U[g_dev.cur], (size_t)5 * g_dev.m.N * 8,
                            hipMemcpyDeviceToHost));
    });
    return FVH_OK;
}

int fvh_set_state(const double *U)
{
    FVH_TRY({
        if (!g_dev.ready)
            throw std::runtime_error("no device state");
        HIP_CHECK(hipStreamSynchronize(g_dev.stream));
        HIP_CHECK(hipMemcpy(g_dev.U[g_dev.cur], U, (size_t)5 * g_dev.m.N * 8,
                            hipMemcpyHostToDevice));
        g_dev.uprim = g_dev.cur;
        g_dev.restart_prims_valid = false;
        g_dev.r_valid = false;
    });
    return FVH_OK;
}

int fvh_get_phi_dt(double *R)
{
    FVH_TRY({
        if (!g_dev.ready || !g_dev.r_valid)
            throw std::runtime_error("no residual on the device");
        HIP_CHECK(hipStreamSynchronize(g_dev.stream));
        HIP_CHECK(hipMemcpy(R, g_dev.R, (size_t)5 * g_dev.m.N * 8,
                            hipMemcpyDeviceToHost));
    });
    return FVH_OK;
}

int fvh_get_field(int what, double *out)
{
    FVH_TRY({
        if (!g_dev.ready)
            throw std::runtime_error("no device state");
        const int N = g_dev.m.N;
        if (what != 0 && what != 1)
            throw std::runtime_error("unknown field");
        // recompute the gradients and min/max of the current state with the
        // test outputs switched on
        if (!g_dev.grad)
            g_dev.grad = dmalloc<double>((size_t)30 * N);
        if (!g_dev.mnmx)
            g_dev.mnmx = dmalloc<double>((size_t)20 * N);
        ensure_gradient_buffers();
        dev::GradOut o = grad_out();
        o.grad = g_dev.grad;
        o.mnmx = g_dev.mnmx;
        dev::launch_ghost(g_dev.stream, g_dev.m, phys(), g_dev.U[g_dev.cur],
                          g_dev.gphi, o.fs, nullptr);
        dev::launch_gradient(g_dev.stream, g_dev.m, phys(), g_dev.U[g_dev.cur],
                             g_dev.gphi, o, is_linear(), nullptr);
        const double *src = what == 0 ? g_dev.grad : g_dev.mnmx;
        const int rows = what == 0 ? 30 : 20;
        HIP_CHECK(hipStreamSynchronize(g_dev.stream));
        HIP_CHECK(hipMemcpy(out, src, (size_t)rows * N * 8, hipMemcpyDeviceToHost));
    });
    return FVH_OK;
}

int fvh_sync_host(void)
{
    FVH_TRY({ sync_mirror(true); });
    return FVH_OK;
}

int fvh_set_output_format(int format)
{
    if (format < 0 || format > 2)
    {
        set_error("fvh_set_output_format: 0 auto, 1 raw, 2 HDF5");
        return FVH_ERR_ARG;
    }
    g_output_format = format;
    return FVH_OK;
}

int fvh_hdf5_available(void) { return hdf5_available() ? 1 : 0; }

int fvh_set_verbose(int verbose)
{
    g_verbose = verbose;
    return FVH_OK;
}

int fvh_set_sync_interval(int steps)
{
    if (steps < 1)
        return FVH_ERR_ARG;
    g_sync_interval = steps;
    return FVH_OK;
}

int fvh_get_log(double *rows, int max_rows)
{
    int n = std::min(max_rows, g_last_iterations);
    if (n > 0)
        std::memcpy(rows, g_log.data(), (size_t)n * 8 * sizeof(double));
    return n;
}

int fvh_last_status(int *iterations)
{
    if (iterations)
        *iterations = g_last_iterations;
    return g_last_status;
}

int fvh_timedisc(void)
{
    FVH_TRY({ return g_is_explicit ? run_explicit(g_max_iter) : run_implicit(g_max_iter); });
}

int fvh_bench_explicit(int n_steps, int n_warmup, double *seconds,
                       double *k_us)
{
    FVH_TRY({
        require_ready();
        if (g_rk_stages <= 0)
            throw std::runtime_error("explicit scheme not initialised");
        const int total = n_steps + n_warmup;
        if (g_dev.log_cap < total)
        {
            g_dev.log_cap = total;
            g_dev.log = dmalloc<double>((size_t)total * 8);
        }
        init_scalars(0, 0.0, 1 << 30, true, DBL_MAX);
        for (int i = 0; i < n_warmup; ++i)
            enqueue_explicit_step(g_dev.sc);
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipStreamSynchronize(g_dev.stream));
        HIP_CHECK(hipEventRecord(e0, g_dev.stream));
        for (int i = 0; i < n_steps; ++i)
            enqueue_explicit_step(g_dev.sc);
        HIP_CHECK(hipEventRecord(e1, g_dev.stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        *seconds = ms * 1e-3;
        read_scalars();
        if (g_dev.sc_host->done < 0)
            throw std::runtime_error("NaN/Inf residual during the benchmark");

        // instrumented pass: events around each kernel of up to 20 steps
        if (k_us)
        {
            const int ev_steps = std::max(1, std::min(n_steps, 20));
            KernelTimer tm;
            HIP_CHECK(hipEventCreate(&tm.a));
            HIP_CHECK(hipEventCreate(&tm.b));
            for (int st = 0; st < ev_steps; ++st)
                enqueue_explicit_step(g_dev.sc, &tm);
            double *acc = tm.acc;
            int *cnt = tm.cnt;
            hipEvent_t a = tm.a, b = tm.b;
            for (int q = 0; q < 3; ++q)
                k_us[q] = cnt[q] ? acc[q] / cnt[q] : 0.0;
            hipEventDestroy(a);
            hipEventDestroy(b);
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        g_dev.r_valid = true;
    });
    return FVH_OK;
}

int fvh_device_info(char *name, int name_len, int *cu_count,
                    double *hbm_gbps_peak)
{
    FVH_TRY({
        int d = 0;
        HIP_CHECK(hipGetDevice(&d));
        hipDeviceProp_t p;
        HIP_CHECK(hipGetDeviceProperties(&p, d));
        if (name && name_len > 0)
        {
            std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
        }
        if (cu_count)
            *cu_count = p.multiProcessorCount;
        if (hbm_gbps_peak)
            *hbm_gbps_peak = 8000.0; // MI355X published HBM3E peak
    });
    return FVH_OK;
}

} // extern "C"
