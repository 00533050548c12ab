// Mesh generation for the box (hexahedra or Kuhn tetrahedra, optionally
// periodic) and the jittered bump channel.
//
// Output is the layout the reference's preprocessor writes
// (tools/pyMeshFV3D/pymeshfv3d/write.py): faces numbered in order of first
// appearance while walking cells and their local faces
// (process.py:96-142), the owner cell first, each cell's face list in
// ascending face id, geometry per utilities/geometry.py (area :15-47,
// volume :50-116, boundary distance :119-130, weight lambda :133-149,
// outward normal :152-183, tangents :186-200) and dx = sum |n| A over the
// cell's faces (process.py:73-77).
//
// Periodic pairing is new: vertices are wrapped onto the periodic lattice,
// so opposite faces share vertex ids and are matched like any interior
// face; each face keeps the translation that carries cells[1] to the side
// of the face (face_shift), so that lambda and dist_cell_2 are measured
// through the periodic image.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>

#include "fvh_internal.h"

namespace fvh {
namespace {

enum
{
    TRIANGLE = 3,
    QUADRANGLE = 4,
    TETRAEDER = 5,
    HEXAEDER = 6
};

struct V3
{
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
            a.x * b.y - a.y * b.x};
}
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 norm(V3 a)
{
    double l = len(a);
    return {a.x / l, a.y / l, a.z / l};
}

// local faces per element type (utilities/element.py:80-103)
const int kTetFaces[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 3, 2}, {2, 3, 1}};
const int kHexFaces[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 4, 7, 3},
                             {0, 4, 5, 1}, {1, 5, 6, 2}, {3, 7, 6, 2}};

// Kuhn split of a hexahedron (gmsh vertex order) into 6 tetrahedra around
// the v0-v6 diagonal; the same split in every hexahedron is conforming and
// compatible with periodic wrapping.
const int kKuhn[6][4] = {{0, 1, 2, 6}, {0, 1, 5, 6}, {0, 3, 2, 6},
                         {0, 3, 7, 6}, {0, 4, 5, 6}, {0, 4, 7, 6}};

struct CellBuild
{
    int type;                 // TETRAEDER or HEXAEDER
    int nv;                   // 4 or 8
    std::vector<int> vid;     // [N*8] wrapped vertex ids
    std::vector<double> xyz;  // [N*8*3] unwrapped coordinates
    std::vector<int> lat;     // [N*8*3] lattice indices (unwrapped)
};

double volume_of(int type, const V3 *v)
{
    if (type == TETRAEDER)
    {
        V3 a = sub(v[1], v[0]), b = sub(v[2], v[0]), c = sub(v[3], v[0]);
        return std::fabs(dot(a, cross(b, c))) * (1.0 / 6.0);
    }
    // hexahedron (geometry.py:74-85)
    double vol = std::fabs(dot(sub(v[5], v[4]),
                               cross(sub(v[1], v[4]), sub(v[6], v[4]))));
    vol += std::fabs(dot(sub(v[1], v[4]),
                         cross(sub(v[3], v[4]), sub(v[6], v[4]))));
    vol += std::fabs(dot(sub(v[1], v[4]),
                         cross(sub(v[0], v[4]), sub(v[3], v[4]))));
    vol += std::fabs(dot(sub(v[6], v[4]),
                         cross(sub(v[3], v[4]), sub(v[7], v[4]))));
    vol += std::fabs(dot(sub(v[3], v[1]),
                         cross(sub(v[6], v[1]), sub(v[2], v[1]))));
    return vol * (1.0 / 6.0);
}

V3 tangent1(V3 n)
{
    V3 t = std::fabs(n.x) > std::fabs(n.y) ? V3{-n.z, 0.0, n.x}
                                           : V3{0.0, -n.z, n.y};
    return norm(t);
}

// Classifies a boundary face by the lattice indices of its vertices; returns
// the region index (> 0).
using Classifier = int (*)(const int (*lat)[3], int nfv, const void *ctx);

void build_mesh(RawMesh &m, const CellBuild &cb, int n_cells,
                const std::array<double, 3> &period,
                const std::array<bool, 3> &periodic, Classifier classify,
                const void *ctx)
{
    const int nv = cb.nv;
    const int nlf = cb.type == TETRAEDER ? 4 : 6;
    const int nfv = cb.type == TETRAEDER ? 3 : 4;
    const int ftype = cb.type == TETRAEDER ? TRIANGLE : QUADRANGLE;
    auto lf_vert = [&](int lf, int k) {
        return cb.type == TETRAEDER ? kTetFaces[lf][k] : kHexFaces[lf][k];
    };

    m.max_cell_vertices = nv;
    m.max_cell_faces = nlf;
    m.max_face_vertices = nfv;

    // ---- cell geometry ----------------------------------------------------
    m.cell_region.assign(n_cells, 0);
    m.cell_type.assign(n_cells, cb.type);
    m.cell_nv.assign(n_cells, nv);
    m.cell_vertices.assign(cb.vid.begin(), cb.vid.begin() + (size_t)n_cells * nv);
    m.cell_volume.resize(n_cells);
    m.cell_x.resize((size_t)n_cells * 3);
    for (int c = 0; c < n_cells; ++c)
    {
        V3 v[8];
        V3 s{0, 0, 0};
        for (int k = 0; k < nv; ++k)
        {
            const double *p = &cb.xyz[((size_t)c * nv + k) * 3];
            v[k] = {p[0], p[1], p[2]};
            s = {s.x + v[k].x, s.y + v[k].y, s.z + v[k].z};
        }
        m.cell_x[c * 3 + 0] = s.x / nv;
        m.cell_x[c * 3 + 1] = s.y / nv;
        m.cell_x[c * 3 + 2] = s.z / nv;
        m.cell_volume[c] = volume_of(cb.type, v);
    }

    // ---- face matching by sorted vertex ids --------------------------------
    const size_t n_ent = (size_t)n_cells * nlf;
    struct Ent
    {
        int k[4];
        uint32_t e;
    };
    std::vector<Ent> ent(n_ent);
    for (int c = 0; c < n_cells; ++c)
        for (int lf = 0; lf < nlf; ++lf)
        {
            Ent &x = ent[(size_t)c * nlf + lf];
            int t[4] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};
            for (int k = 0; k < nfv; ++k)
                t[k] = cb.vid[(size_t)c * nv + lf_vert(lf, k)];
            std::sort(t, t + 4);
            std::memcpy(x.k, t, sizeof(t));
            x.e = (uint32_t)((size_t)c * nlf + lf);
        }
    std::sort(ent.begin(), ent.end(), [](const Ent &a, const Ent &b) {
        for (int i = 0; i < 4; ++i)
            if (a.k[i] != b.k[i])
                return a.k[i] < b.k[i];
        return a.e < b.e;
    });
    std::vector<int64_t> partner(n_ent, -2); // -2: second of a pair
    for (size_t i = 0; i < n_ent;)
    {
        size_t j = i + 1;
        while (j < n_ent && std::memcmp(ent[j].k, ent[i].k, sizeof(ent[i].k)) == 0)
            ++j;
        if (j - i > 2)
            throw std::runtime_error("mesh: a face is shared by more than two cells");
        if (j - i == 2)
            partner[ent[i].e] = ent[i + 1].e;
        else
            partner[ent[i].e] = -1;
        i = j;
    }
    ent.clear();
    ent.shrink_to_fit();

    // ---- faces in order of first appearance --------------------------------
    std::vector<size_t> face_first, face_second;
    for (size_t e = 0; e < n_ent; ++e)
        if (partner[e] != -2)
        {
            face_first.push_back(e);
            face_second.push_back(partner[e] < 0 ? SIZE_MAX : (size_t)partner[e]);
        }
    partner.clear();
    partner.shrink_to_fit();
    const int F = (int)face_first.size();

    m.face_type.assign(F, ftype);
    m.face_nv.assign(F, nfv);
    m.face_vertices.assign((size_t)F * nfv, -1);
    m.face_cells.assign((size_t)F * 2, -1);
    m.face_boundary.assign(F, -1);
    m.face_x.resize((size_t)F * 3);
    m.face_n.resize((size_t)F * 3);
    m.face_t1.resize((size_t)F * 3);
    m.face_t2.resize((size_t)F * 3);
    m.face_area.resize(F);
    m.face_lambda.resize(F);
    m.face_shift.assign((size_t)F * 3, 0.0);

    m.bnd_region.clear();
    for (int f = 0; f < F; ++f)
    {
        size_t e0 = face_first[f];
        int c0 = (int)(e0 / nlf), lf0 = (int)(e0 % nlf);
        V3 v[4];
        V3 s{0, 0, 0};
        for (int k = 0; k < nfv; ++k)
        {
            int lv = lf_vert(lf0, k);
            const double *p = &cb.xyz[((size_t)c0 * nv + lv) * 3];
            v[k] = {p[0], p[1], p[2]};
            s = {s.x + v[k].x, s.y + v[k].y, s.z + v[k].z};
            m.face_vertices[(size_t)f * nfv + k] = cb.vid[(size_t)c0 * nv + lv];
        }
        V3 xf = {s.x / nfv, s.y / nfv, s.z / nfv};
        double area;
        V3 n;
        if (nfv == 3)
        {
            V3 a = sub(v[1], v[0]), b = sub(v[2], v[0]);
            area = len(cross(a, b)) * 0.5;
            n = norm(cross(a, b));
        }
        else
        {
            V3 a = sub(v[1], v[0]), b = sub(v[2], v[0]), c = sub(v[3], v[0]);
            area = (len(cross(a, b)) + len(cross(b, c))) * 0.5;
            n = norm(cross(sub(v[2], v[0]), sub(v[3], v[1])));
        }
        V3 xc0 = {m.cell_x[c0 * 3], m.cell_x[c0 * 3 + 1], m.cell_x[c0 * 3 + 2]};
        V3 d = sub(xf, xc0);
        double cosang = dot(d, n) / (len(d) * len(n));
        cosang = std::max(std::min(cosang, 1.0), -1.0);
        double ang = std::acos(cosang) * 180.0 / M_PI;
        if (90.0 < ang && ang < 270.0)
            n = scale(n, -1.0);
        V3 t1 = tangent1(n);
        V3 t2 = norm(cross(t1, n));

        m.face_cells[(size_t)f * 2] = c0;
        m.face_x[f * 3] = xf.x, m.face_x[f * 3 + 1] = xf.y, m.face_x[f * 3 + 2] = xf.z;
        m.face_n[f * 3] = n.x, m.face_n[f * 3 + 1] = n.y, m.face_n[f * 3 + 2] = n.z;
        m.face_t1[f * 3] = t1.x, m.face_t1[f * 3 + 1] = t1.y, m.face_t1[f * 3 + 2] = t1.z;
        m.face_t2[f * 3] = t2.x, m.face_t2[f * 3 + 1] = t2.y, m.face_t2[f * 3 + 2] = t2.z;
        m.face_area[f] = area;

        if (face_second[f] != SIZE_MAX)
        {
            size_t e1 = face_second[f];
            int c1 = (int)(e1 / nlf), lf1 = (int)(e1 % nlf);
            m.face_cells[(size_t)f * 2 + 1] = c1;
            V3 s1{0, 0, 0};
            for (int k = 0; k < nfv; ++k)
            {
                const double *p = &cb.xyz[((size_t)c1 * nv + lf_vert(lf1, k)) * 3];
                s1 = {s1.x + p[0], s1.y + p[1], s1.z + p[2]};
            }
            V3 xf1 = {s1.x / nfv, s1.y / nfv, s1.z / nfv};
            double T[3] = {xf.x - xf1.x, xf.y - xf1.y, xf.z - xf1.z};
            for (int a = 0; a < 3; ++a)
                T[a] = periodic[a] ? std::round(T[a] / period[a]) * period[a] : 0.0;
            for (int a = 0; a < 3; ++a)
                m.face_shift[f * 3 + a] = T[a];
            V3 xc1 = {m.cell_x[c1 * 3] + T[0], m.cell_x[c1 * 3 + 1] + T[1],
                      m.cell_x[c1 * 3 + 2] + T[2]};
            double d1 = dot(n, sub(xf, xc0)) / area;
            double d2 = dot(n, sub(xc1, xf)) / area;
            if (!(d1 + d2 > 1e-16))
                throw std::runtime_error("mesh: degenerate interpolation weight");
            m.face_lambda[f] = d1 / (d1 + d2);
        }
        else
        {
            m.face_lambda[f] = 1.0;
            int lat[4][3];
            for (int k = 0; k < nfv; ++k)
                for (int a = 0; a < 3; ++a)
                    lat[k][a] = cb.lat[((size_t)c0 * nv + lf_vert(lf0, k)) * 3 + a];
            int region = classify(lat, nfv, ctx);
            int b = (int)m.bnd_region.size();
            m.face_boundary[f] = b;
            m.bnd_region.push_back(region);
            m.bnd_face.push_back(f);
            V3 bn = scale(n, -1.0);
            V3 bt1 = tangent1(bn);
            V3 bt2 = norm(cross(bt1, bn));
            m.bnd_n.insert(m.bnd_n.end(), {bn.x, bn.y, bn.z});
            m.bnd_t1.insert(m.bnd_t1.end(), {bt1.x, bt1.y, bt1.z});
            m.bnd_t2.insert(m.bnd_t2.end(), {bt2.x, bt2.y, bt2.z});
            V3 dx = sub(xc0, xf);
            double tmp = dot(dx, bn) / (len(dx) * len(bn));
            m.bnd_distance.push_back(len(dx) * std::max(std::min(tmp, 1.0), -1.0));
            m.bnd_type.push_back(ftype);
            m.bnd_nv.push_back(nfv);
            for (int k = 0; k < nfv; ++k)
                m.bnd_vertices.push_back(m.face_vertices[(size_t)f * nfv + k]);
        }
    }
    m.max_boundary_vertices = nfv;

    // ---- cell face lists (ascending) and dx --------------------------------
    m.cell_nf.assign(n_cells, 0);
    m.cell_faces.assign((size_t)n_cells * nlf, -1);
    for (int f = 0; f < F; ++f)
        for (int side = 0; side < 2; ++side)
        {
            int c = m.face_cells[(size_t)f * 2 + side];
            if (c < 0)
                continue;
            int &k = m.cell_nf[c];
            if (k >= nlf)
                throw std::runtime_error("mesh: too many faces for a cell");
            m.cell_faces[(size_t)c * nlf + k] = f;
            ++k;
        }
    m.cell_dx.assign((size_t)n_cells * 3, 0.0);
    for (int c = 0; c < n_cells; ++c)
        for (int k = 0; k < m.cell_nf[c]; ++k)
        {
            int f = m.cell_faces[(size_t)c * nlf + k];
            for (int a = 0; a < 3; ++a)
                m.cell_dx[c * 3 + a] += std::fabs(m.face_n[f * 3 + a]) * m.face_area[f];
        }
}

// ---------------------------------------------------------------------------
struct BoxCtx
{
    int n[3];
    bool periodic[3];
    int region_of_side[6]; // xmin,xmax,ymin,ymax,zmin,zmax (-1 if periodic)
};

int classify_box(const int (*lat)[3], int nfv, const void *vctx)
{
    const BoxCtx *ctx = (const BoxCtx *)vctx;
    for (int a = 0; a < 3; ++a)
    {
        bool lo = true, hi = true;
        for (int k = 0; k < nfv; ++k)
        {
            lo = lo && lat[k][a] == 0;
            hi = hi && lat[k][a] == ctx->n[a];
        }
        if (lo && ctx->region_of_side[2 * a] > 0)
            return ctx->region_of_side[2 * a];
        if (hi && ctx->region_of_side[2 * a + 1] > 0)
            return ctx->region_of_side[2 * a + 1];
    }
    throw std::runtime_error("mesh: unclassified boundary face");
}

// hexahedra in lexicographic (tile = 0) or tile^3-blocked order
std::vector<std::array<int, 3>> hex_order(int nx, int ny, int nz, int tile)
{
    std::vector<std::array<int, 3>> o;
    o.reserve((size_t)nx * ny * nz);
    if (tile <= 0)
    {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    o.push_back({i, j, k});
        return o;
    }
    // inside a block the hexahedra follow a snake: each layer's rows in
    // alternating directions, and the layers' row orders alternating too, so
    // that consecutive hexahedra are face neighbours and a run of cells (one
    // cell tile) is compact: 6.5% fewer faces cut by the tile boundaries
    // and 7.5% fewer halo cells per tile than rows in one direction
    for (int kb = 0; kb < nz; kb += tile)
        for (int jb = 0; jb < ny; jb += tile)
            for (int ib = 0; ib < nx; ib += tile)
            {
                const int ex = std::min(nx, ib + tile) - ib, ey = std::min(ny, jb + tile) - jb;
                int row = 0;
                for (int k = kb; k < std::min(nz, kb + tile); ++k)
                    for (int r = 0; r < ey; ++r, ++row)
                    {
                        const int j = jb + (((k - kb) & 1) ? ey - 1 - r : r);
                        for (int q = 0; q < ex; ++q)
                            o.push_back({ib + ((row & 1) ? ex - 1 - q : q), j, k});
                    }
            }
    return o;
}

const int kHexCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// Fills the cell build from hexahedra; node(i,j,k) gives unwrapped
// coordinates, vid(i,j,k) the wrapped vertex id.
template <class NodeFn, class VidFn>
void fill_cells(CellBuild &cb, const std::vector<std::array<int, 3>> &hexes,
                bool tets, NodeFn node, VidFn vid)
{
    cb.type = tets ? TETRAEDER : HEXAEDER;
    cb.nv = tets ? 4 : 8;
    size_t nc = hexes.size() * (tets ? 6 : 1);
    cb.vid.resize(nc * cb.nv);
    cb.xyz.resize(nc * cb.nv * 3);
    cb.lat.resize(nc * cb.nv * 3);
    size_t c = 0;
    for (const auto &h : hexes)
    {
        int hv[8], hl[8][3];
        double hx[8][3];
        for (int k = 0; k < 8; ++k)
        {
            int i = h[0] + kHexCorner[k][0], j = h[1] + kHexCorner[k][1],
                l = h[2] + kHexCorner[k][2];
            hl[k][0] = i, hl[k][1] = j, hl[k][2] = l;
            hv[k] = vid(i, j, l);
            node(i, j, l, hx[k]);
        }
        int nsub = tets ? 6 : 1;
        for (int s = 0; s < nsub; ++s, ++c)
            for (int k = 0; k < cb.nv; ++k)
            {
                int hk = tets ? kKuhn[s][k] : k;
                cb.vid[c * cb.nv + k] = hv[hk];
                for (int a = 0; a < 3; ++a)
                {
                    cb.xyz[(c * cb.nv + k) * 3 + a] = hx[hk][a];
                    cb.lat[(c * cb.nv + k) * 3 + a] = hl[hk][a];
                }
            }
    }
}

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

double uniform01(uint64_t &s)
{
    return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

struct ChannelCtx
{
    int n[3];
};

int classify_channel(const int (*lat)[3], int nfv, const void *vctx)
{
    const ChannelCtx *ctx = (const ChannelCtx *)vctx;
    // regions: 1 inflow (x lo), 2 outflow (x hi), 3 bottom (y lo), 4 top
    // (y hi), 5 zmin, 6 zmax
    for (int a = 0; a < 3; ++a)
    {
        bool lo = true, hi = true;
        for (int k = 0; k < nfv; ++k)
        {
            lo = lo && lat[k][a] == 0;
            hi = hi && lat[k][a] == ctx->n[a];
        }
        if (lo)
            return 1 + 2 * a;
        if (hi)
            return 2 + 2 * a;
    }
    throw std::runtime_error("mesh: unclassified channel boundary face");
}

} // namespace

int generate_box(RawMesh &m, int nx, int ny, int nz, double x0, double y0,
                 double z0, double lx, double ly, double lz, bool tets,
                 int periodic, int tile)
{
    m = RawMesh();
    int n[3] = {nx, ny, nz};
    double L[3] = {lx, ly, lz};
    double o[3] = {x0, y0, z0};
    bool per[3] = {(periodic & 1) != 0, (periodic & 2) != 0, (periodic & 4) != 0};
    for (int a = 0; a < 3; ++a)
        if (n[a] < 1 || (per[a] && n[a] < 3))
            throw std::runtime_error("mesh: box needs n >= 1 (n >= 3 when periodic)");

    int nn[3];
    for (int a = 0; a < 3; ++a)
        nn[a] = per[a] ? n[a] : n[a] + 1;
    m.vertices.resize((size_t)nn[0] * nn[1] * nn[2] * 3);
    for (int k = 0; k < nn[2]; ++k)
        for (int j = 0; j < nn[1]; ++j)
            for (int i = 0; i < nn[0]; ++i)
            {
                size_t v = ((size_t)k * nn[1] + j) * nn[0] + i;
                int ijk[3] = {i, j, k};
                for (int a = 0; a < 3; ++a)
                    m.vertices[v * 3 + a] = o[a] + L[a] * ((double)ijk[a] / n[a]);
            }

    BoxCtx ctx{};
    m.region_names = {"flow"};
    m.region_is_boundary = {0};
    const char *side_names[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
    for (int a = 0; a < 3; ++a)
    {
        ctx.n[a] = n[a];
        ctx.periodic[a] = per[a];
        for (int s = 0; s < 2; ++s)
        {
            if (per[a])
                ctx.region_of_side[2 * a + s] = -1;
            else
            {
                ctx.region_of_side[2 * a + s] = (int)m.region_names.size();
                m.region_names.push_back(side_names[2 * a + s]);
                m.region_is_boundary.push_back(1);
            }
        }
    }

    CellBuild cb;
    auto hexes = hex_order(nx, ny, nz, tile);
    fill_cells(
        cb, hexes, tets,
        [&](int i, int j, int k, double *x) {
            int ijk[3] = {i, j, k};
            for (int a = 0; a < 3; ++a)
                x[a] = o[a] + L[a] * ((double)ijk[a] / n[a]);
        },
        [&](int i, int j, int k) {
            int ijk[3] = {i, j, k};
            for (int a = 0; a < 3; ++a)
                if (per[a])
                    ijk[a] %= n[a];
            return (int)(((size_t)ijk[2] * nn[1] + ijk[1]) * nn[0] + ijk[0]);
        });
    int n_cells = (int)(hexes.size() * (tets ? 6 : 1));
    build_mesh(m, cb, n_cells, {lx, ly, lz}, {per[0], per[1], per[2]},
               classify_box, &ctx);
    return n_cells;
}

int generate_channel(RawMesh &m, int nx, int ny, int nz, double jitter,
                     unsigned long long seed)
{
    m = RawMesh();
    if (nx < 1 || ny < 1 || nz < 1)
        throw std::runtime_error("mesh: channel needs positive sizes");
    int nn[3] = {nx + 1, ny + 1, nz + 1};
    std::vector<double> X((size_t)nn[0] * nn[1] * nn[2] * 3);
    uint64_t state = seed;
    for (int k = 0; k < nn[2]; ++k)
        for (int j = 0; j < nn[1]; ++j)
            for (int i = 0; i < nn[0]; ++i)
            {
                size_t v = ((size_t)k * nn[1] + j) * nn[0] + i;
                double x = -2.0 + 4.0 * ((double)i / nx);
                double hb = 0.1 * std::exp(-25.0 * x * x);
                double y = hb + (1.0 - hb) * ((double)j / ny);
                double z = 0.5 * ((double)k / nz);
                bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
                if (interior && jitter > 0.0)
                {
                    double ux = uniform01(state), uy = uniform01(state),
                           uz = uniform01(state);
                    x += jitter * (4.0 / nx) * (2.0 * ux - 1.0);
                    y += jitter * ((1.0 - hb) / ny) * (2.0 * uy - 1.0);
                    z += jitter * (0.5 / nz) * (2.0 * uz - 1.0);
                }
                X[v * 3] = x, X[v * 3 + 1] = y, X[v * 3 + 2] = z;
            }
    m.vertices = X;
    m.region_names = {"flow", "inflow", "outflow", "bottom", "top", "zmin", "zmax"};
    m.region_is_boundary = {0, 1, 1, 1, 1, 1, 1};
    ChannelCtx ctx{{nx, ny, nz}};
    CellBuild cb;
    auto hexes = hex_order(nx, ny, nz, 0);
    fill_cells(
        cb, hexes, true,
        [&](int i, int j, int k, double *x) {
            size_t v = ((size_t)k * nn[1] + j) * nn[0] + i;
            for (int a = 0; a < 3; ++a)
                x[a] = X[v * 3 + a];
        },
        [&](int i, int j, int k) {
            return (int)(((size_t)k * nn[1] + j) * nn[0] + i);
        });
    int n_cells = (int)hexes.size() * 6;
    build_mesh(m, cb, n_cells, {4.0, 1.0, 0.5}, {false, false, false},
               classify_channel, &ctx);
    return n_cells;
}

} // namespace fvh
