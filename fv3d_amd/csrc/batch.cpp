// Batched mode: many independent periodic cases of uneven size in one mesh
// (disjoint cell ranges, one device state), advanced together by the
// explicit RK scheme, every case with its own dt.
//
// Persistent schedule: per stage one launch of the gradient kernel over
// each contiguous run of viscous cases, then one persistent launch of the
// stage kernel per (flux, viscous) class of cases: workgroups take tiles
// (block / lanes consecutive cells of one case) from the launch's work
// queue and compute each with its case's viscosity and dt; the last stage
// writes per-tile partials that a per-case fold turns into the residual
// norms and the next dt.  Group schedule: one launch per (flux, viscous) group and stage, with
// chunked per-case reductions.  Either step is captured once into a HIP
// graph (one per ping-pong parity) and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "batch.h"
#include "fvh_internal.h"
#include "runtime.h"

namespace fvh {

namespace {

struct Group
{
    int flux, viscous, c_begin, c_end;
    double mu;
    int t_begin = 0, t_end = 0; // the group's face-table tiles (face schedule)
};

constexpr int kMaxStages = 5, kClasses = 4, kCounters = kMaxStages * kClasses;

int case_class(int flux, int viscous) { return viscous * 2 + (flux == dev::FluxRusanov); }

struct BatchState
{
    bool ready = false;
    std::vector<int> offsets; // case cell offsets in the union mesh [K+1]
    int K = 0;
    std::vector<Group> groups;
    std::vector<int> case_flux, case_visc;
    std::vector<double> case_mu;
    int n_chunks = 0;
    std::vector<void *> allocs;
    int *cell_case = nullptr, *chunk_begin = nullptr;
    int3 *chunks = nullptr;
    double *partials = nullptr, *d_mu = nullptr, *d_vol = nullptr, *d_dt = nullptr,
           *d_t = nullptr, *d_res = nullptr, *vg = nullptr;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    int steps = 0;
    // persistent schedule
    bool persistent = false; // the per-group schedule measured faster
    int block = 256, grid = 0; // grid 0: min(tiles, 4 x CUs); -1: one workgroup per tile
    int sub_tiles = 1;         // runs of block / lanes cells per tile
    // tiles by class (viscous * 2 + flux: AUSM and Rusanov inviscid, then
    // viscous), each case's tiles consecutive
    int n_tiles = 0, tile_cells = 0, cls_base[kClasses] = {}, cls_tiles[kClasses] = {};
    int *tile_case = nullptr, *tile_begin = nullptr, *case_end = nullptr, *row_lo = nullptr,
        *row_hi = nullptr; // a case's partial rows row_lo[k] .. row_hi[k] - 1
    double *d_lam = nullptr, *tile_partials = nullptr;
    unsigned int *counters = nullptr; // [stage][class]
    unsigned long long *stamps = nullptr; // [stages][n_tiles][2] when timing
    std::vector<std::pair<int, int>> visc_runs; // cell ranges of viscous cases
    std::vector<int> h_tile_lo, h_tile_hi; // a case's tiles
    // face schedule (tetrahedral cases): per group one launch of the
    // face-centred residual kernel over case-aligned tiles of at most
    // dev::kTile cells; the last stage writes kTile / 64 partials rows per
    // tile, and a case's rows are f_row_lo[k] .. f_row_hi[k] - 1
    bool face = false;
    int f_tiles = 0;
    int *f_tile_c0 = nullptr, *f_face_off = nullptr, *f_row_lo = nullptr,
        *f_row_hi = nullptr;
    unsigned long long *f_faces = nullptr; // null: too large for 8-byte entries
    double *f_partials = nullptr;
    std::vector<int> f_case_t0; // a case's tiles f_case_t0[k] .. f_case_t0[k+1] - 1
};


BatchState g_b;
std::vector<int> g_mesh_offsets; // from the last fvh_mesh_batch

template <class X> X *balloc(size_t n)
{
    void *p = nullptr;
    FVH_HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(X)));
    g_b.allocs.push_back(p);
    return static_cast<X *>(p);
}

void release_graphs()
{
    for (int i = 0; i < 2; ++i)
    {
        if (g_b.exec[i])
            (void)hipGraphExecDestroy(g_b.exec[i]);
        if (g_b.graph[i])
            (void)hipGraphDestroy(g_b.graph[i]);
        g_b.exec[i] = nullptr;
        g_b.graph[i] = nullptr;
    }
}

void release()
{
    release_graphs();
    for (void *p : g_b.allocs)
        (void)hipFree(p);
    g_b = BatchState();
}

// Appends src to dst with index offsets (cells, faces, vertices,
// boundaries); regions are those of the first mesh.
void append_mesh(RawMesh &d, const RawMesh &s)
{
    const int vo = (int)(d.vertices.size() / 3), co = d.n_cells(), fo = d.n_faces(),
              bo = d.n_boundaries();
    auto shift = [](std::vector<int> &dst, const std::vector<int> &src, int off) {
        for (int v : src)
            dst.push_back(v >= 0 ? v + off : v);
    };
    auto cat = [](auto &dst, const auto &src) { dst.insert(dst.end(), src.begin(), src.end()); };
    if (co == 0)
    {
        d.dimension = s.dimension;
        d.region_names = s.region_names;
        d.region_is_boundary = s.region_is_boundary;
    }
    if (co > 0 && (d.max_cell_vertices != s.max_cell_vertices ||
                   d.max_cell_faces != s.max_cell_faces ||
                   d.max_face_vertices != s.max_face_vertices ||
                   d.max_boundary_vertices != s.max_boundary_vertices))
        throw std::runtime_error("batch: cases must share the element type");
    d.max_cell_vertices = s.max_cell_vertices;
    d.max_cell_faces = s.max_cell_faces;
    d.max_face_vertices = s.max_face_vertices;
    d.max_boundary_vertices = s.max_boundary_vertices;
    cat(d.vertices, s.vertices);
    cat(d.cell_region, s.cell_region);
    cat(d.cell_type, s.cell_type);
    cat(d.cell_nv, s.cell_nv);
    shift(d.cell_vertices, s.cell_vertices, vo);
    cat(d.cell_nf, s.cell_nf);
    shift(d.cell_faces, s.cell_faces, fo);
    cat(d.cell_volume, s.cell_volume);
    cat(d.cell_x, s.cell_x);
    cat(d.cell_dx, s.cell_dx);
    cat(d.face_type, s.face_type);
    cat(d.face_nv, s.face_nv);
    shift(d.face_vertices, s.face_vertices, vo);
    shift(d.face_cells, s.face_cells, co);
    shift(d.face_boundary, s.face_boundary, bo);
    cat(d.face_x, s.face_x);
    cat(d.face_n, s.face_n);
    cat(d.face_t1, s.face_t1);
    cat(d.face_t2, s.face_t2);
    cat(d.face_area, s.face_area);
    cat(d.face_lambda, s.face_lambda);
    cat(d.face_shift, s.face_shift);
    cat(d.bnd_region, s.bnd_region);
    cat(d.bnd_type, s.bnd_type);
    cat(d.bnd_nv, s.bnd_nv);
    shift(d.bnd_vertices, s.bnd_vertices, vo);
    shift(d.bnd_face, s.bnd_face, fo);
    cat(d.bnd_n, s.bnd_n);
    cat(d.bnd_t1, s.bnd_t1);
    cat(d.bnd_t2, s.bnd_t2);
    cat(d.bnd_distance, s.bnd_distance);
}

// Tiles of the persistent schedule for `block` threads per workgroup (the
// case's cells in order, sub_tiles * block / lanes per tile; the tiles of
// one class together), uploaded to the device.
void build_tiles(const RuntimeView &v)
{
    const int lanes = v.m.maxf <= 4 ? 4 : 8;
    const int run = g_b.block / lanes, gc = run * g_b.sub_tiles;
    std::vector<int> tc, tb, lo(g_b.K), hi(g_b.K), rlo(g_b.K), rhi(g_b.K);
    for (int cls = 0; cls < kClasses; ++cls)
    {
        g_b.cls_base[cls] = (int)tc.size();
        for (int k = 0; k < g_b.K; ++k)
        {
            if (case_class(g_b.case_flux[k], g_b.case_visc[k]) != cls)
                continue;
            lo[k] = (int)tc.size();
            for (int c = g_b.offsets[k]; c < g_b.offsets[k + 1]; c += gc)
            {
                tc.push_back(k);
                tb.push_back(c);
            }
            hi[k] = (int)tc.size();
            // partial rows: one per run of the case's tiles
            rlo[k] = lo[k] * g_b.sub_tiles;
            rhi[k] = rlo[k] + (g_b.offsets[k + 1] - g_b.offsets[k] + run - 1) / run;
        }
        g_b.cls_tiles[cls] = (int)tc.size() - g_b.cls_base[cls];
    }
    g_b.n_tiles = (int)tc.size();
    g_b.tile_cells = gc;
    // earlier tables stay allocated until the batch is released
    g_b.tile_case = balloc<int>(tc.size());
    g_b.tile_begin = balloc<int>(tb.size());
    g_b.row_lo = balloc<int>(g_b.K);
    g_b.row_hi = balloc<int>(g_b.K);
    g_b.tile_partials = balloc<double>((size_t)8 * tc.size() * g_b.sub_tiles);
    g_b.stamps = nullptr; // sized for the tile count: allocated again when timing
    FVH_HIP_CHECK(hipMemcpy(g_b.tile_case, tc.data(), tc.size() * sizeof(int),
                            hipMemcpyHostToDevice));
    FVH_HIP_CHECK(hipMemcpy(g_b.tile_begin, tb.data(), tb.size() * sizeof(int),
                            hipMemcpyHostToDevice));
    FVH_HIP_CHECK(hipMemcpy(g_b.row_lo, rlo.data(), rlo.size() * sizeof(int),
                            hipMemcpyHostToDevice));
    FVH_HIP_CHECK(hipMemcpy(g_b.row_hi, rhi.data(), rhi.size() * sizeof(int),
                            hipMemcpyHostToDevice));
    g_b.h_tile_lo = lo;
    g_b.h_tile_hi = hi;
}

dev::BatchTiles tiles_view(int stage, int cls, unsigned long long *stamps)
{
    dev::BatchTiles bt{};
    bt.tile_base = g_b.cls_base[cls];
    bt.n_tiles = g_b.cls_tiles[cls];
    bt.sub_tiles = g_b.sub_tiles;
    bt.tile_case = g_b.tile_case;
    bt.tile_begin = g_b.tile_begin;
    bt.case_end = g_b.case_end;
    bt.case_mu = g_b.d_mu;
    bt.case_lam = g_b.d_lam;
    bt.counter = g_b.grid < 0 ? nullptr : g_b.counters + stage * kClasses + cls;
    bt.stamps = stamps;
    return bt;
}

int persistent_grid(int n_tiles)
{
    if (g_b.grid < 0)
        return n_tiles; // one workgroup per tile
    int grid = g_b.grid;
    if (grid <= 0)
    {
        int dev = 0, cus = 256;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        grid = 4 * cus;
    }
    return std::max(1, std::min(grid, n_tiles));
}

// One RK step of every case on the persistent schedule; with `stamps`, the
// tiles' wall-clock start and end of stage st go to stamps + st * 2 * tiles.
// The stage launches of the two classes share the stage's state buffers
// (disjoint cells).
void enqueue_step_persistent(const RuntimeView &v, const StageBuffers &sb,
                             unsigned long long *stamps)
{
    hipStream_t s = v.stream;
    const dev::DevMesh &m = v.m;
    for (int st = 0; st < sb.stages; ++st)
    {
        const bool last = st == sb.stages - 1;
        double *Uin = sb.U[*sb.cur], *Uout = sb.U[1 - *sb.cur];
        for (const auto &r : g_b.visc_runs)
        {
            dev::GradOut o{};
            o.vg = g_b.vg;
            o.c_begin = r.first;
            o.c_end = r.second;
            dev::launch_gradient(s, m, v.P, Uin, nullptr, o, 0, nullptr);
        }
        dev::StageParams sp{};
        sp.U_in = Uin;
        sp.U_out = Uout;
        sp.k = sb.k;
        sp.R_out = last ? sb.R : nullptr;
        sp.vg = g_b.vg;
        sp.rk_a = sb.rk_a[st];
        sp.rk_g = sb.rk_g[st];
        sp.stage0 = st == 0;
        sp.k_dead = last; // the next step starts k from R
        sp.do_update = 1;
        sp.do_reduce = last;
        sp.partials = g_b.tile_partials;
        sp.partials_cap = g_b.n_tiles * g_b.sub_tiles;
        sp.c_begin = 0;
        sp.c_end = m.N;
        sp.case_dt = g_b.d_dt;
        sp.cell_case = g_b.cell_case;
        unsigned long long *sst = stamps ? stamps + (size_t)st * 2 * g_b.n_tiles : nullptr;
        for (int cls = 0; cls < kClasses; ++cls)
            if (g_b.cls_tiles[cls] > 0)
                dev::launch_batch_stage(s, m, v.P, sp, tiles_view(st, cls, sst),
                                        cls % 2 ? dev::FluxRusanov : dev::FluxAUSM, cls >= 2,
                                        g_b.block, persistent_grid(g_b.cls_tiles[cls]));
        *sb.cur = 1 - *sb.cur;
    }
    // per-case residuals (dt of the step) and next dt from the tiles'
    // partials, in tile order; then the work-queue counters for the next step
    batch::case_fold(s, 1, v.P, g_b.K, g_b.row_lo, g_b.row_hi, g_b.tile_partials, g_b.d_mu,
                     g_b.d_vol, g_b.d_dt, g_b.d_t, g_b.d_res);
    batch::reset_counters(s, g_b.counters, kCounters);
}

// One RK step of every case on the face schedule: per stage and group, the
// viscous gradients (viscous groups) and the face-centred residual kernel
// over the group's tiles; the last stage's partials per tile, folded per
// case.
void enqueue_step_face(const RuntimeView &v, const StageBuffers &sb,
                       unsigned long long *stamps = nullptr)
{
    hipStream_t s = v.stream;
    dev::DevMesh m = v.m;
    m.tile_c0 = g_b.f_tile_c0;
    m.tile_face_off = g_b.f_face_off;
    m.tile_faces = g_b.f_faces;
    for (int st = 0; st < sb.stages; ++st)
    {
        const bool last = st == sb.stages - 1;
        double *Uin = sb.U[*sb.cur], *Uout = sb.U[1 - *sb.cur];
        for (const Group &g : g_b.groups)
        {
            dev::PhysDev P = v.P;
            P.mu = g.mu;
            P.lam = g.mu * cp / Pr;
            if (g.viscous)
            {
                dev::GradOut o{};
                o.vg = g_b.vg;
                o.c_begin = g.c_begin;
                o.c_end = g.c_end;
                dev::launch_gradient(s, m, P, Uin, nullptr, o, 0, nullptr);
            }
            dev::StageParams sp{};
            sp.U_in = Uin;
            sp.U_out = Uout;
            sp.k = sb.k;
            sp.R_out = last ? sb.R : nullptr;
            sp.vg = g_b.vg;
            sp.rk_a = sb.rk_a[st];
            sp.rk_g = sb.rk_g[st];
            sp.stage0 = st == 0;
            sp.k_dead = last; // the next step starts k from R
            sp.do_update = 1;
            sp.do_reduce = last;
            sp.partials = g_b.f_partials;
            sp.partials_cap = g_b.f_tiles * (dev::kTile / 64);
            sp.c_begin = g.c_begin;
            sp.c_end = g.c_end;
            sp.case_dt = g_b.d_dt;
            sp.cell_case = g_b.cell_case;
            sp.stamps = stamps ? stamps + (size_t)st * 2 * g_b.f_tiles : nullptr;
            dev::launch_residual_tiles(s, m, P, sp, g.flux, false, g.viscous != 0, g.t_begin,
                                       g.t_end - g.t_begin);
        }
        *sb.cur = 1 - *sb.cur;
    }
    batch::case_fold(s, 1, v.P, g_b.K, g_b.f_row_lo, g_b.f_row_hi, g_b.f_partials, g_b.d_mu,
                     g_b.d_vol, g_b.d_dt, g_b.d_t, g_b.d_res);
}

// Enqueues one RK step of every case on the selected schedule (the stages of
// every group, then the per-case residuals and next time steps).
void enqueue_step(const RuntimeView &v, const StageBuffers &sb)
{
    if (g_b.face)
    {
        enqueue_step_face(v, sb);
        return;
    }
    if (g_b.persistent)
    {
        enqueue_step_persistent(v, sb, nullptr);
        return;
    }
    hipStream_t s = v.stream;
    const dev::DevMesh &m = v.m;
    for (int st = 0; st < sb.stages; ++st)
    {
        const bool last = st == sb.stages - 1;
        double *Uin = sb.U[*sb.cur], *Uout = sb.U[1 - *sb.cur];
        for (const Group &g : g_b.groups)
        {
            dev::PhysDev P = v.P;
            P.mu = g.mu;
            P.lam = g.mu * cp / Pr;
            if (g.viscous)
            {
                dev::GradOut o{};
                o.vg = g_b.vg;
                o.c_begin = g.c_begin;
                o.c_end = g.c_end;
                dev::launch_gradient(s, m, P, Uin, nullptr, o, 0, nullptr);
            }
            dev::StageParams sp{};
            sp.U_in = Uin;
            sp.U_out = Uout;
            sp.k = sb.k;
            sp.R_out = last ? sb.R : nullptr;
            sp.vg = g_b.vg;
            sp.rk_a = sb.rk_a[st];
            sp.rk_g = sb.rk_g[st];
            sp.stage0 = st == 0;
            sp.k_dead = last; // the next step starts k from R
            sp.do_update = 1;
            sp.c_begin = g.c_begin;
            sp.c_end = g.c_end;
            sp.case_dt = g_b.d_dt;
            sp.cell_case = g_b.cell_case;
            dev::launch_residual(s, m, P, sp, g.flux, false, g.viscous != 0);
        }
        *sb.cur = 1 - *sb.cur;
    }
    // residual of the last stage, dt from the new density and the last
    // stage's input primitives
    batch::chunk_partials(s, m, v.P, g_b.n_chunks, g_b.chunks, sb.U[*sb.cur],
                          sb.U[1 - *sb.cur], sb.R, g_b.partials);
    batch::case_fold(s, 1, v.P, g_b.K, g_b.chunk_begin, g_b.chunk_begin + 1, g_b.partials,
                     g_b.d_mu, g_b.d_vol, g_b.d_dt, g_b.d_t, g_b.d_res);
}

} // namespace

void batch_release() { release(); }

} // namespace fvh

using namespace fvh;

#define B_TRY(...)                                                             \
    try                                                                        \
    {                                                                          \
        __VA_ARGS__;                                                           \
    }                                                                          \
    catch (const std::exception &e)                                            \
    {                                                                          \
        set_error(e.what());                                                   \
        return FVH_ERR_STATE;                                                  \
    }                                                                          \
    return FVH_OK;

int fvh_mesh_batch(int n_cases, const int *n_hex, double length, int tile)
{
    B_TRY({
        if (n_cases < 1 || !n_hex)
            throw std::runtime_error("fvh_mesh_batch: no cases");
        RawMesh all;
        std::vector<int> off(1, 0);
        for (int k = 0; k < n_cases; ++k)
        {
            if (n_hex[k] < 3)
                throw std::runtime_error("fvh_mesh_batch: a periodic case needs >= 3 "
                                         "hexahedra per side");
            RawMesh m;
            generate_box(m, n_hex[k], n_hex[k], n_hex[k], 0.0, 0.0, 0.0, length, length,
                         length, true, 7, tile);
            append_mesh(all, m);
            off.push_back(all.n_cells());
        }
        release();
        install_mesh(std::move(all));
        g_mesh_offsets = off;
    });
}

int fvh_batch_setup(int n_cases, const int *flux, const int *viscous, const double *mu)
{
    B_TRY({
        if (n_cases + 1 != (int)g_mesh_offsets.size())
            throw std::runtime_error("fvh_batch_setup: case count differs from fvh_mesh_batch");
        RuntimeView v = runtime_view();
        if (v.m.B != 0)
            throw std::runtime_error("fvh_batch_setup: batched cases are periodic");
        if (v.linear)
            throw std::runtime_error("fvh_batch_setup: batched cases use first-order "
                                     "reconstruction");
        release();
        g_b.offsets = g_mesh_offsets;
        g_b.K = n_cases;
        const int N = v.m.N;
        if (g_b.offsets.back() != N)
            throw std::runtime_error("fvh_batch_setup: mesh is not the last fvh_mesh_batch");
        // groups: maximal runs of equal (flux, viscous, mu)
        for (int k = 0; k < n_cases; ++k)
        {
            if (flux[k] != dev::FluxAUSM && flux[k] != dev::FluxRusanov)
                throw std::runtime_error("fvh_batch_setup: flux must be AUSM or Rusanov");
            const int vis = viscous[k] != 0 && mu[k] != 0.0;
            g_b.case_flux.push_back(flux[k]);
            g_b.case_visc.push_back(vis);
            g_b.case_mu.push_back(vis ? mu[k] : 0.0);
            if (!g_b.groups.empty() && g_b.groups.back().flux == flux[k] &&
                g_b.groups.back().viscous == vis && g_b.groups.back().mu == g_b.case_mu[k])
                g_b.groups.back().c_end = g_b.offsets[k + 1];
            else
                g_b.groups.push_back({flux[k], vis, g_b.offsets[k], g_b.offsets[k + 1],
                                      g_b.case_mu[k]});
        }
        // host tables
        std::vector<int> cell_case(N), chunk_begin(1, 0);
        std::vector<int3> chunks;
        std::vector<double> vol(n_cases, 0.0);
        const double *cv = solver_mesh->cells->volume;
        for (int k = 0; k < n_cases; ++k)
        {
            for (int c = g_b.offsets[k]; c < g_b.offsets[k + 1]; ++c)
            {
                cell_case[c] = k;
                vol[k] += cv[c];
            }
            for (int c = g_b.offsets[k]; c < g_b.offsets[k + 1]; c += batch::kChunkCells)
                chunks.push_back(make_int3(k, c, std::min(c + batch::kChunkCells,
                                                          g_b.offsets[k + 1])));
            chunk_begin.push_back((int)chunks.size());
        }
        g_b.n_chunks = (int)chunks.size();
        g_b.cell_case = balloc<int>(N);
        g_b.chunk_begin = balloc<int>(n_cases + 1);
        g_b.chunks = balloc<int3>(chunks.size());
        g_b.partials = balloc<double>((size_t)8 * chunks.size());
        g_b.d_mu = balloc<double>(n_cases);
        g_b.d_vol = balloc<double>(n_cases);
        g_b.d_dt = balloc<double>(n_cases);
        g_b.d_t = balloc<double>(n_cases);
        g_b.d_res = balloc<double>((size_t)5 * n_cases);
        bool any_visc = false;
        for (const Group &g : g_b.groups)
            any_visc = any_visc || g.viscous;
        if (any_visc)
            g_b.vg = balloc<double>((size_t)12 * N);
        FVH_HIP_CHECK(hipMemcpy(g_b.cell_case, cell_case.data(), N * sizeof(int),
                                hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemcpy(g_b.chunk_begin, chunk_begin.data(),
                                chunk_begin.size() * sizeof(int), hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemcpy(g_b.chunks, chunks.data(), chunks.size() * sizeof(int3),
                                hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemcpy(g_b.d_mu, g_b.case_mu.data(), n_cases * sizeof(double),
                                hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemcpy(g_b.d_vol, vol.data(), n_cases * sizeof(double),
                                hipMemcpyHostToDevice));
        // persistent schedule: per-case tables, tiles, counters, and the
        // cell runs of viscous cases (one gradient launch each)
        {
            std::vector<int> cend(n_cases);
            std::vector<double> lam(n_cases);
            for (int k = 0; k < n_cases; ++k)
            {
                cend[k] = g_b.offsets[k + 1];
                lam[k] = g_b.case_mu[k] * cp / Pr;
                if (g_b.case_visc[k])
                {
                    if (!g_b.visc_runs.empty() && g_b.visc_runs.back().second == g_b.offsets[k])
                        g_b.visc_runs.back().second = g_b.offsets[k + 1];
                    else
                        g_b.visc_runs.push_back({g_b.offsets[k], g_b.offsets[k + 1]});
                }
            }
            g_b.case_end = balloc<int>(n_cases);
            g_b.d_lam = balloc<double>(n_cases);
            g_b.counters = balloc<unsigned int>(kCounters);
            FVH_HIP_CHECK(hipMemcpy(g_b.case_end, cend.data(), n_cases * sizeof(int),
                                    hipMemcpyHostToDevice));
            FVH_HIP_CHECK(hipMemcpy(g_b.d_lam, lam.data(), n_cases * sizeof(double),
                                    hipMemcpyHostToDevice));
            batch::reset_counters(v.stream, g_b.counters, kCounters);
            build_tiles(v);
        }
        // face schedule: case-aligned tiles and their face tables
        if (v.m.maxf == 4)
        {
            const int rows = dev::kTile / 64;
            std::vector<int> c0, case_t0(n_cases + 1), rlo(n_cases), rhi(n_cases);
            for (int k = 0; k < n_cases; ++k)
            {
                case_t0[k] = (int)c0.size();
                for (int c = g_b.offsets[k]; c < g_b.offsets[k + 1]; c += dev::kTile)
                    c0.push_back(c);
                rlo[k] = case_t0[k] * rows;
                rhi[k] = (int)c0.size() * rows;
            }
            case_t0[n_cases] = (int)c0.size();
            g_b.f_case_t0 = case_t0;
            g_b.f_tiles = (int)c0.size();
            c0.push_back(N);
            int k = 0;
            for (Group &g : g_b.groups)
            {
                while (g_b.offsets[k] != g.c_begin)
                    ++k;
                g.t_begin = case_t0[k];
                int k1 = k;
                while (g_b.offsets[k1] != g.c_end)
                    ++k1;
                g.t_end = case_t0[k1];
            }
            std::vector<int> off;
            std::vector<int4> ent;
            build_tile_faces(c0, off, ent);
            g_b.f_tile_c0 = balloc<int>(c0.size());
            g_b.f_face_off = balloc<int>(off.size());
            g_b.f_row_lo = balloc<int>(n_cases);
            g_b.f_row_hi = balloc<int>(n_cases);
            g_b.f_partials = balloc<double>((size_t)g_b.f_tiles * rows * dev::kPartialStride);
            FVH_HIP_CHECK(hipMemcpy(g_b.f_tile_c0, c0.data(), c0.size() * sizeof(int),
                                    hipMemcpyHostToDevice));
            FVH_HIP_CHECK(hipMemcpy(g_b.f_face_off, off.data(), off.size() * sizeof(int),
                                    hipMemcpyHostToDevice));
            std::vector<unsigned long long> packed;
            if (pack_tile_faces_table(c0, off, ent, packed))
            {
                g_b.f_faces = balloc<unsigned long long>(packed.size());
                FVH_HIP_CHECK(hipMemcpy(g_b.f_faces, packed.data(),
                                        packed.size() * sizeof(unsigned long long),
                                        hipMemcpyHostToDevice));
            }
            FVH_HIP_CHECK(hipMemcpy(g_b.f_row_lo, rlo.data(), n_cases * sizeof(int),
                                    hipMemcpyHostToDevice));
            FVH_HIP_CHECK(hipMemcpy(g_b.f_row_hi, rhi.data(), n_cases * sizeof(int),
                                    hipMemcpyHostToDevice));
        }
        std::vector<double> zeros((size_t)5 * n_cases, 0.0);
        FVH_HIP_CHECK(hipMemcpy(g_b.d_t, zeros.data(), n_cases * sizeof(double),
                                hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemcpy(g_b.d_res, zeros.data(), zeros.size() * sizeof(double),
                                hipMemcpyHostToDevice));
        // initial time steps from the current state (primitives refreshed)
        batch::chunk_partials(v.stream, v.m, v.P, g_b.n_chunks, g_b.chunks, v.U, v.U, nullptr,
                              g_b.partials);
        batch::case_fold(v.stream, 0, v.P, n_cases, g_b.chunk_begin, g_b.chunk_begin + 1,
                         g_b.partials, g_b.d_mu, g_b.d_vol, g_b.d_dt, g_b.d_t, g_b.d_res);
        FVH_HIP_CHECK(hipStreamSynchronize(v.stream));
        g_b.ready = true;
    });
}

int fvh_batch_run(int steps, int use_graph, double *seconds)
{
    B_TRY({
        if (!g_b.ready)
            throw std::runtime_error("fvh_batch_run: no batch (fvh_batch_setup)");
        RuntimeView v = runtime_view();
        StageBuffers sb = stage_buffers();
        if (sb.stages % 2 == 0 && use_graph)
            throw std::runtime_error("fvh_batch_run: graphs assume an odd stage count");
        hipStream_t s = v.stream;
        if (use_graph)
            for (int par = 0; par < 2; ++par)
            {
                const int want = (*sb.cur + par) & 1;
                if (g_b.exec[want])
                    continue;
                // capture the step that starts from ping-pong buffer `want`
                const int saved = *sb.cur;
                *sb.cur = want;
                FVH_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
                enqueue_step(v, sb);
                FVH_HIP_CHECK(hipStreamEndCapture(s, &g_b.graph[want]));
                FVH_HIP_CHECK(hipGraphInstantiate(&g_b.exec[want], g_b.graph[want], nullptr,
                                                  nullptr, 0));
                *sb.cur = saved;
            }
        hipEvent_t e0, e1;
        FVH_HIP_CHECK(hipEventCreate(&e0));
        FVH_HIP_CHECK(hipEventCreate(&e1));
        FVH_HIP_CHECK(hipEventRecord(e0, s));
        for (int i = 0; i < steps; ++i)
        {
            if (use_graph)
            {
                FVH_HIP_CHECK(hipGraphLaunch(g_b.exec[*sb.cur], s));
                *sb.cur = 1 - *sb.cur; // odd stage count: the step flips the buffers
            }
            else
                enqueue_step(v, sb);
        }
        FVH_HIP_CHECK(hipEventRecord(e1, s));
        FVH_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        FVH_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (seconds)
            *seconds = ms * 1e-3;
        g_b.steps += steps;
        state_changed();
    });
}

int fvh_batch_info(int *n_cases, int *offsets, int max_offsets)
{
    B_TRY({
        if (!g_b.ready)
            throw std::runtime_error("no batch");
        if (n_cases)
            *n_cases = g_b.K;
        if (offsets)
            for (int i = 0; i <= g_b.K && i < max_offsets; ++i)
                offsets[i] = g_b.offsets[i];
    });
}

int fvh_batch_case_scalars(double *dt, double *t, double *res)
{
    B_TRY({
        if (!g_b.ready)
            throw std::runtime_error("no batch");
        FVH_HIP_CHECK(hipDeviceSynchronize());
        if (dt)
            FVH_HIP_CHECK(hipMemcpy(dt, g_b.d_dt, g_b.K * sizeof(double), hipMemcpyDeviceToHost));
        if (t)
            FVH_HIP_CHECK(hipMemcpy(t, g_b.d_t, g_b.K * sizeof(double), hipMemcpyDeviceToHost));
        if (res)
            FVH_HIP_CHECK(hipMemcpy(res, g_b.d_res, (size_t)5 * g_b.K * sizeof(double),
                                    hipMemcpyDeviceToHost));
    });
}

int fvh_batch_set_schedule(int persistent, int block, int grid, int sub_tiles)
{
    B_TRY({
        if (persistent == 2)
        {
            if (!g_b.ready || !g_b.f_faces)
                throw std::runtime_error("fvh_batch_set_schedule: the face schedule needs a "
                                         "batch of tetrahedral cases");
            g_b.face = true;
            g_b.persistent = false;
            release_graphs();
            return FVH_OK;
        }
        if (persistent < 0 || persistent > 2)
            throw std::runtime_error("fvh_batch_set_schedule: schedule must be 0, 1 or 2");
        g_b.face = false;
        if (persistent && block != 128 && block != 256 && block != 512)
            throw std::runtime_error("fvh_batch_set_schedule: block must be 128, 256 or 512");
        if (persistent && (sub_tiles < 1 || sub_tiles > 1024))
            throw std::runtime_error("fvh_batch_set_schedule: sub_tiles must be 1 .. 1024");
        const bool rebuild =
            g_b.ready && persistent && (block != g_b.block || sub_tiles != g_b.sub_tiles);
        g_b.persistent = persistent != 0;
        if (persistent)
        {
            g_b.block = block;
            g_b.sub_tiles = sub_tiles;
        }
        g_b.grid = grid;
        release_graphs();
        if (rebuild)
            build_tiles(runtime_view());
    });
}

int fvh_batch_case_step_times(int steps, double *ms)
{
    B_TRY({
        if (!g_b.ready || !(g_b.persistent || g_b.face))
            throw std::runtime_error("fvh_batch_case_step_times: needs the persistent or the "
                                     "face schedule");
        if (steps < 1 || !ms)
            throw std::runtime_error("fvh_batch_case_step_times: bad arguments");
        RuntimeView v = runtime_view();
        StageBuffers sb = stage_buffers();
        hipStream_t s = v.stream;
        // the schedule's tiles and each case's range of them
        const int tiles = g_b.face ? g_b.f_tiles : g_b.n_tiles;
        const size_t n = (size_t)sb.stages * 2 * tiles;
        if (!g_b.stamps)
            g_b.stamps = balloc<unsigned long long>(
                (size_t)kMaxStages * 2 * std::max(g_b.n_tiles, g_b.f_tiles));
        std::vector<unsigned long long> h(n);
        std::vector<double> acc(g_b.K, 0.0);
        int dev = 0, khz = 100000;
        FVH_HIP_CHECK(hipGetDevice(&dev));
        FVH_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        for (int it = 0; it < steps; ++it)
        {
            if (g_b.face)
                enqueue_step_face(v, sb, g_b.stamps);
            else
                enqueue_step_persistent(v, sb, g_b.stamps);
            FVH_HIP_CHECK(hipMemcpyAsync(h.data(), g_b.stamps, n * sizeof(unsigned long long),
                                         hipMemcpyDeviceToHost, s));
            FVH_HIP_CHECK(hipStreamSynchronize(s));
            // a case's time in a stage: from the start of its first tile to
            // the end of its last, inside the shared launch
            for (int st = 0; st < sb.stages; ++st)
            {
                const unsigned long long *p = h.data() + (size_t)st * 2 * tiles;
                for (int k = 0; k < g_b.K; ++k)
                {
                    unsigned long long lo = ~0ull, hi = 0;
                    const int t0 = g_b.face ? g_b.f_case_t0[k] : g_b.h_tile_lo[k],
                              t1 = g_b.face ? g_b.f_case_t0[k + 1] : g_b.h_tile_hi[k];
                    for (int t = t0; t < t1; ++t)
                    {
                        lo = std::min(lo, p[(size_t)t * 2]);
                        hi = std::max(hi, p[(size_t)t * 2 + 1]);
                    }
                    acc[k] += (double)(hi - lo) / (double)khz;
                }
            }
        }
        for (int k = 0; k < g_b.K; ++k)
            ms[k] = acc[k] / steps;
        g_b.steps += steps;
        state_changed();
    });
}
