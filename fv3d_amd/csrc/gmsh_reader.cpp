// Gmsh 4.1 ASCII reader and general mesh preprocessing (mixed tetrahedra,
// hexahedra, prisms and pyramids) into the layout the reference's
// preprocessor writes.
//
// Reader: the sections and checks of tools/pyMeshFV3D/pymeshfv3d/reader/
// gmsh.py:35-139 ($MeshFormat "4.1 0 8", $PhysicalNames with tags 1..n in
// order, $Entities, $Nodes, $Elements; node and element tags 1..n; an
// element's region is its entity's first physical tag - 1).
// Preprocessing (process.py:23-214): 3D elements become cells and 2D
// elements boundaries, both in element-tag order; a cell's local faces follow
// utilities/element.py:58-129; faces are numbered in order of first
// appearance while walking the cells and their local faces, the second cell
// found by the same vertex set and face type; each boundary element is linked
// to the boundary face with its vertex set; geometry follows
// utilities/geometry.py (area :15-47, volume :50-116, distance :119-130,
// weight :133-149, normal :152-183, tangents :186-200) after scaling the
// vertices per axis (process.py:45-48); dx = sum |n| A (process.py:73-77).
//
// Differences: points and lines (0D/1D elements) are ignored instead of
// becoming boundaries that link to no face; a boundary face without a
// boundary element (or a boundary element without a face) is an error; the
// pyramid's fifth face is (v3, v0, v4) where element.py:125 has (v0, v3, v1),
// which is not a face of the pyramid and would leave the neighbour across
// the real face unmatched.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fvh_internal.h"

namespace fvh {
namespace {

enum
{
    POINT = 1,
    LINE = 2,
    TRIANGLE = 3,
    QUADRANGLE = 4,
    TETRAEDER = 5,
    HEXAEDER = 6,
    PRISM = 7,
    PYRAMID = 8
};

// gmsh element type -> (element type, vertex count)
bool gmsh_type(int g, int &type, int &nv)
{
    switch (g)
    {
    case 15: type = POINT, nv = 1; return true;
    case 1: type = LINE, nv = 2; return true;
    case 2: type = TRIANGLE, nv = 3; return true;
    case 3: type = QUADRANGLE, nv = 4; return true;
    case 4: type = TETRAEDER, nv = 4; return true;
    case 5: type = HEXAEDER, nv = 8; return true;
    case 6: type = PRISM, nv = 6; return true;
    case 7: type = PYRAMID, nv = 5; return true;
    default: return false;
    }
}

struct LocalFace
{
    int type, nv, v[4];
};

// local faces (vertex positions in the cell's list) per cell type
int local_faces(int type, LocalFace *out)
{
    auto set = [](LocalFace &f, int t, std::initializer_list<int> v) {
        f.type = t;
        f.nv = (int)v.size();
        int k = 0;
        for (int x : v)
            f.v[k++] = x;
        for (; k < 4; ++k)
            f.v[k] = -1;
    };
    switch (type)
    {
    case TETRAEDER:
        set(out[0], TRIANGLE, {0, 1, 2});
        set(out[1], TRIANGLE, {0, 3, 1});
        set(out[2], TRIANGLE, {0, 3, 2});
        set(out[3], TRIANGLE, {2, 3, 1});
        return 4;
    case HEXAEDER:
        set(out[0], QUADRANGLE, {0, 1, 2, 3});
        set(out[1], QUADRANGLE, {4, 5, 6, 7});
        set(out[2], QUADRANGLE, {0, 4, 7, 3});
        set(out[3], QUADRANGLE, {0, 4, 5, 1});
        set(out[4], QUADRANGLE, {1, 5, 6, 2});
        set(out[5], QUADRANGLE, {3, 7, 6, 2});
        return 6;
    case PRISM:
        set(out[0], QUADRANGLE, {0, 1, 4, 3});
        set(out[1], QUADRANGLE, {1, 2, 5, 4});
        set(out[2], QUADRANGLE, {0, 2, 5, 3});
        set(out[3], TRIANGLE, {0, 1, 2});
        set(out[4], TRIANGLE, {3, 4, 5});
        return 5;
    case PYRAMID:
        set(out[0], QUADRANGLE, {0, 1, 2, 3});
        set(out[1], TRIANGLE, {0, 1, 4});
        set(out[2], TRIANGLE, {1, 2, 4});
        set(out[3], TRIANGLE, {2, 3, 4});
        // element.py:125 lists (0, 3, 1) here, which is no face of the
        // pyramid; the face is (3, 0, 4)
        set(out[4], TRIANGLE, {3, 0, 4});
        return 5;
    default:
        throw std::runtime_error("gmsh: unsupported cell type");
    }
}

struct V3
{
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a)
{
    const double l = len(a);
    return {a.x / l, a.y / l, a.z / l};
}
inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }

double cell_volume(int type, const V3 *v)
{
    double vol;
    switch (type)
    {
    case TETRAEDER:
        vol = std::fabs(dot(sub(v[1], v[0]), cross(sub(v[2], v[0]), sub(v[3], v[0]))));
        break;
    case HEXAEDER:
        vol = std::fabs(dot(sub(v[5], v[4]), cross(sub(v[1], v[4]), sub(v[6], v[4]))));
        vol += std::fabs(dot(sub(v[1], v[4]), cross(sub(v[3], v[4]), sub(v[6], v[4]))));
        vol += std::fabs(dot(sub(v[1], v[4]), cross(sub(v[0], v[4]), sub(v[3], v[4]))));
        vol += std::fabs(dot(sub(v[6], v[4]), cross(sub(v[3], v[4]), sub(v[7], v[4]))));
        vol += std::fabs(dot(sub(v[3], v[1]), cross(sub(v[6], v[1]), sub(v[2], v[1]))));
        break;
    case PRISM:
    {
        const V3 a = sub(v[1], v[0]), b = sub(v[2], v[0]), c = sub(v[3], v[0]),
                 d = sub(v[4], v[0]), e = sub(v[5], v[0]);
        vol = std::fabs(dot(a, cross(b, e)));
        vol += std::fabs(dot(a, cross(e, d)));
        vol += std::fabs(dot(d, cross(e, c)));
        break;
    }
    case PYRAMID:
    {
        const V3 a = sub(v[1], v[0]), b = sub(v[2], v[0]), c = sub(v[3], v[0]),
                 d = sub(v[4], v[0]);
        vol = std::fabs(dot(a, cross(b, d)));
        vol += std::fabs(dot(b, cross(c, d)));
        break;
    }
    default:
        throw std::runtime_error("gmsh: unsupported cell type");
    }
    return vol * (1.0 / 6.0);
}

double face_area(int type, const V3 *v)
{
    if (type == TRIANGLE)
        return len(cross(sub(v[1], v[0]), sub(v[2], v[0]))) * 0.5;
    const V3 a = sub(v[1], v[0]), b = sub(v[2], v[0]), c = sub(v[3], v[0]);
    return (len(cross(a, b)) + len(cross(b, c))) * 0.5;
}

V3 tangent1(V3 n)
{
    return unit(std::fabs(n.x) > std::fabs(n.y) ? V3{-n.z, 0.0, n.x} : V3{0.0, -n.z, n.y});
}

struct Element
{
    int type = 0, region = -1;
    std::vector<int> v;
};

struct Parsed
{
    std::vector<std::string> regions;
    std::vector<double> xyz;
    std::vector<Element> elements;
};

[[noreturn]] void fail(const std::string &what) { throw std::runtime_error("gmsh: " + what); }

std::string next_line(std::istream &in)
{
    std::string s;
    if (!std::getline(in, s))
        fail("unexpected end of file");
    while (!s.empty() && (s.back() == '\r' || s.back() == ' ' || s.back() == '\t'))
        s.pop_back();
    size_t b = 0;
    while (b < s.size() && (s[b] == ' ' || s[b] == '\t'))
        ++b;
    return s.substr(b);
}

void expect(std::istream &in, const char *what)
{
    const std::string s = next_line(in);
    if (s != what)
        fail(std::string("expected '") + what + "', found '" + s + "'");
}

std::vector<long> ints(const std::string &s)
{
    std::vector<long> out;
    std::istringstream is(s);
    long x;
    while (is >> x)
        out.push_back(x);
    return out;
}

Parsed parse(const char *path)
{
    std::ifstream in(path);
    if (!in)
        fail(std::string("cannot open ") + path);
    Parsed p;
    // (dim, entity tag) -> region index
    std::map<std::pair<int, int>, int> entity_region;
    std::string line;
    bool have_nodes = false, have_elements = false;
    while (std::getline(in, line))
    {
        while (!line.empty() && (line.back() == '\r' || line.back() == ' '))
            line.pop_back();
        if (line.empty())
            continue;
        if (line == "$MeshFormat")
        {
            expect(in, "4.1 0 8");
            expect(in, "$EndMeshFormat");
        }
        else if (line == "$PhysicalNames")
        {
            const long n = std::stol(next_line(in));
            for (long i = 0; i < n; ++i)
            {
                const std::string s = next_line(in);
                std::istringstream is(s);
                long dim, tag;
                if (!(is >> dim >> tag))
                    fail("bad physical name line '" + s + "'");
                if (tag != i + 1)
                    fail("physical tags must be 1..n in order");
                const size_t q0 = s.find('"'), q1 = s.rfind('"');
                if (q0 == std::string::npos || q1 == q0)
                    fail("unquoted physical name");
                p.regions.push_back(s.substr(q0 + 1, q1 - q0 - 1));
            }
            expect(in, "$EndPhysicalNames");
        }
        else if (line == "$Entities")
        {
            const std::vector<long> n = ints(next_line(in));
            if (n.size() != 4)
                fail("bad entity counts");
            for (int d = 0; d < 4; ++d)
                for (long j = 0; j < n[d]; ++j)
                {
                    std::istringstream is(next_line(in));
                    long tag;
                    double x;
                    is >> tag;
                    for (int k = 0; k < (d == 0 ? 3 : 6); ++k)
                        is >> x;
                    long ntags = 0, first = 0;
                    is >> ntags;
                    if (ntags > 0)
                        is >> first;
                    if (!is)
                        fail("bad entity line");
                    entity_region[{d, (int)tag}] = ntags > 0 ? (int)first - 1 : -1;
                }
            expect(in, "$EndEntities");
        }
        else if (line == "$Nodes")
        {
            const std::vector<long> h = ints(next_line(in));
            if (h.size() != 4)
                fail("bad node header");
            const long nblocks = h[0], nnodes = h[1];
            p.xyz.assign((size_t)nnodes * 3, NAN);
            for (long b = 0; b < nblocks; ++b)
            {
                const std::vector<long> bh = ints(next_line(in));
                if (bh.size() != 4)
                    fail("bad node block header");
                if (bh[2] != 0)
                    fail("parametric nodes are not supported");
                std::vector<long> tags(bh[3]);
                for (long j = 0; j < bh[3]; ++j)
                    tags[j] = std::stol(next_line(in));
                for (long j = 0; j < bh[3]; ++j)
                {
                    if (tags[j] < 1 || tags[j] > nnodes)
                        fail("node tags must be 1..n");
                    std::istringstream is(next_line(in));
                    double *x = &p.xyz[(size_t)(tags[j] - 1) * 3];
                    if (!(is >> x[0] >> x[1] >> x[2]))
                        fail("bad node coordinates");
                }
            }
            expect(in, "$EndNodes");
            have_nodes = true;
        }
        else if (line == "$Elements")
        {
            const std::vector<long> h = ints(next_line(in));
            if (h.size() != 4)
                fail("bad element header");
            const long nblocks = h[0], nelem = h[1];
            p.elements.assign(nelem, Element());
            for (long b = 0; b < nblocks; ++b)
            {
                const std::vector<long> bh = ints(next_line(in));
                if (bh.size() != 4)
                    fail("bad element block header");
                int type, nv;
                if (!gmsh_type((int)bh[2], type, nv))
                    fail("unsupported element type " + std::to_string(bh[2]));
                auto it = entity_region.find({(int)bh[0], (int)bh[1]});
                if (it == entity_region.end())
                    fail("element block of an unknown entity");
                for (long j = 0; j < bh[3]; ++j)
                {
                    const std::vector<long> e = ints(next_line(in));
                    if ((long)e.size() < 1 + nv || e[0] < 1 || e[0] > nelem)
                        fail("bad element line");
                    Element &el = p.elements[e[0] - 1];
                    el.type = type;
                    el.region = it->second;
                    el.v.resize(nv);
                    for (int k = 0; k < nv; ++k)
                    {
                        if (e[1 + k] < 1 || e[1 + k] > (long)(p.xyz.size() / 3))
                            fail("element node out of range");
                        el.v[k] = (int)e[1 + k] - 1;
                    }
                }
            }
            expect(in, "$EndElements");
            have_elements = true;
        }
        else if (line[0] == '$')
        {
            // skip sections the solver does not use ($Periodic, $NodeData ...)
            const std::string end = "$End" + line.substr(1);
            std::string s;
            do
                s = next_line(in);
            while (s != end);
        }
        else
            fail("unexpected line '" + line + "'");
    }
    if (!have_nodes || !have_elements)
        fail("no $Nodes or $Elements section");
    for (const Element &e : p.elements)
        if (e.type == 0)
            fail("element tags must be 1..n");
    for (double x : p.xyz)
        if (std::isnan(x))
            fail("node tags must be 1..n");
    return p;
}

struct FaceKey
{
    int type, k[4];
    uint32_t ent;
};

bool key_less(const FaceKey &a, const FaceKey &b)
{
    if (a.type != b.type)
        return a.type < b.type;
    for (int i = 0; i < 4; ++i)
        if (a.k[i] != b.k[i])
            return a.k[i] < b.k[i];
    return a.ent < b.ent;
}

bool key_equal(const FaceKey &a, const FaceKey &b)
{
    return a.type == b.type && std::memcmp(a.k, b.k, sizeof(a.k)) == 0;
}

FaceKey make_key(int type, const int *v, int nv, uint32_t ent)
{
    FaceKey f;
    f.type = type;
    int t[4] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};
    for (int i = 0; i < nv; ++i)
        t[i] = v[i];
    std::sort(t, t + 4);
    std::memcpy(f.k, t, sizeof(t));
    f.ent = ent;
    return f;
}

} // namespace

int read_gmsh(RawMesh &m, const char *path, const double scale[3])
{
    Parsed p = parse(path);
    m = RawMesh();
    m.dimension = 3;
    const size_t nvert = p.xyz.size() / 3;
    m.vertices = p.xyz;
    for (size_t i = 0; i < nvert; ++i)
        for (int a = 0; a < 3; ++a)
            m.vertices[i * 3 + a] *= scale[a];
    auto X = [&](int v) {
        return V3{m.vertices[(size_t)v * 3], m.vertices[(size_t)v * 3 + 1],
                  m.vertices[(size_t)v * 3 + 2]};
    };

    std::vector<const Element *> cells, bnds;
    for (const Element &e : p.elements)
    {
        if (e.type >= TETRAEDER)
            cells.push_back(&e);
        else if (e.type == TRIANGLE || e.type == QUADRANGLE)
            bnds.push_back(&e);
    }
    const int N = (int)cells.size();
    if (N == 0)
        fail("no 3D elements");

    // ---- cells ---------------------------------------------------------------
    int maxcv = 0, maxcf = 0;
    std::vector<int> nlf(N);
    std::vector<std::array<LocalFace, 6>> lfs(N);
    for (int c = 0; c < N; ++c)
    {
        nlf[c] = local_faces(cells[c]->type, lfs[c].data());
        maxcv = std::max(maxcv, (int)cells[c]->v.size());
        maxcf = std::max(maxcf, nlf[c]);
    }
    m.max_cell_vertices = maxcv;
    m.max_cell_faces = maxcf;
    m.cell_region.resize(N);
    m.cell_type.resize(N);
    m.cell_nv.resize(N);
    m.cell_vertices.assign((size_t)N * maxcv, -1);
    m.cell_volume.resize(N);
    m.cell_x.resize((size_t)N * 3);
    for (int c = 0; c < N; ++c)
    {
        const Element &e = *cells[c];
        m.cell_region[c] = e.region;
        m.cell_type[c] = e.type;
        m.cell_nv[c] = (int)e.v.size();
        V3 v[8];
        V3 s{0.0, 0.0, 0.0};
        for (size_t k = 0; k < e.v.size(); ++k)
        {
            m.cell_vertices[(size_t)c * maxcv + k] = e.v[k];
            v[k] = X(e.v[k]);
            s = {s.x + v[k].x, s.y + v[k].y, s.z + v[k].z};
        }
        const double nv = (double)e.v.size();
        m.cell_x[(size_t)c * 3] = s.x / nv;
        m.cell_x[(size_t)c * 3 + 1] = s.y / nv;
        m.cell_x[(size_t)c * 3 + 2] = s.z / nv;
        m.cell_volume[c] = cell_volume(e.type, v);
    }

    // ---- faces: match local faces by type and vertex set -------------------
    std::vector<uint32_t> ent_off(N + 1, 0);
    for (int c = 0; c < N; ++c)
        ent_off[c + 1] = ent_off[c] + nlf[c];
    const uint32_t n_ent = ent_off[N];
    std::vector<int> ent_cell(n_ent), ent_lf(n_ent);
    std::vector<FaceKey> keys(n_ent);
    for (int c = 0; c < N; ++c)
        for (int l = 0; l < nlf[c]; ++l)
        {
            const uint32_t e = ent_off[c] + l;
            const LocalFace &lf = lfs[c][l];
            int v[4];
            for (int k = 0; k < lf.nv; ++k)
                v[k] = cells[c]->v[lf.v[k]];
            keys[e] = make_key(lf.type, v, lf.nv, e);
            ent_cell[e] = c;
            ent_lf[e] = l;
        }
    std::sort(keys.begin(), keys.end(), key_less);
    std::vector<int64_t> partner(n_ent, -2); // -2: second entry of a pair
    for (size_t i = 0; i < keys.size();)
    {
        size_t j = i + 1;
        while (j < keys.size() && key_equal(keys[j], keys[i]))
            ++j;
        if (j - i > 2)
            fail("a face is shared by more than two cells");
        partner[keys[i].ent] = j - i == 2 ? (int64_t)keys[i + 1].ent : -1;
        i = j;
    }
    std::vector<uint32_t> first;
    std::vector<int64_t> second;
    for (uint32_t e = 0; e < n_ent; ++e)
        if (partner[e] != -2)
        {
            first.push_back(e);
            second.push_back(partner[e]);
        }
    const int F = (int)first.size();
    int maxfv = 0;
    for (int f = 0; f < F; ++f)
        maxfv = std::max(maxfv, lfs[ent_cell[first[f]]][ent_lf[first[f]]].nv);
    m.max_face_vertices = maxfv;
    m.face_type.resize(F);
    m.face_nv.resize(F);
    m.face_vertices.assign((size_t)F * maxfv, -1);
    m.face_cells.assign((size_t)F * 2, -1);
    m.face_boundary.assign(F, -1);
    m.face_x.resize((size_t)F * 3);
    m.face_n.resize((size_t)F * 3);
    m.face_t1.resize((size_t)F * 3);
    m.face_t2.resize((size_t)F * 3);
    m.face_area.resize(F);
    m.face_lambda.resize(F);
    m.face_shift.assign((size_t)F * 3, 0.0);
    std::vector<FaceKey> bkeys; // boundary faces by vertex set
    for (int f = 0; f < F; ++f)
    {
        const int c0 = ent_cell[first[f]];
        const LocalFace &lf = lfs[c0][ent_lf[first[f]]];
        V3 v[4];
        V3 s{0.0, 0.0, 0.0};
        int vid[4];
        for (int k = 0; k < lf.nv; ++k)
        {
            vid[k] = cells[c0]->v[lf.v[k]];
            m.face_vertices[(size_t)f * maxfv + k] = vid[k];
            v[k] = X(vid[k]);
            s = {s.x + v[k].x, s.y + v[k].y, s.z + v[k].z};
        }
        m.face_type[f] = lf.type;
        m.face_nv[f] = lf.nv;
        const V3 xf = {s.x / lf.nv, s.y / lf.nv, s.z / lf.nv};
        const double area = face_area(lf.type, v);
        V3 n = lf.type == TRIANGLE ? unit(cross(sub(v[1], v[0]), sub(v[2], v[0])))
                                   : unit(cross(sub(v[2], v[0]), sub(v[3], v[1])));
        const V3 xc0 = {m.cell_x[(size_t)c0 * 3], m.cell_x[(size_t)c0 * 3 + 1],
                        m.cell_x[(size_t)c0 * 3 + 2]};
        const V3 d = sub(xf, xc0);
        double cosang = dot(d, n) / (len(d) * len(n));
        cosang = std::max(std::min(cosang, 1.0), -1.0);
        const double ang = std::acos(cosang) * 180.0 / M_PI;
        if (90.0 < ang && ang < 270.0)
            n = neg(n);
        const V3 t1 = tangent1(n), t2 = unit(cross(t1, n));
        m.face_cells[(size_t)f * 2] = c0;
        const double fx[3] = {xf.x, xf.y, xf.z}, fn[3] = {n.x, n.y, n.z},
                     f1[3] = {t1.x, t1.y, t1.z}, f2[3] = {t2.x, t2.y, t2.z};
        for (int a = 0; a < 3; ++a)
        {
            m.face_x[(size_t)f * 3 + a] = fx[a];
            m.face_n[(size_t)f * 3 + a] = fn[a];
            m.face_t1[(size_t)f * 3 + a] = f1[a];
            m.face_t2[(size_t)f * 3 + a] = f2[a];
        }
        m.face_area[f] = area;
        if (second[f] >= 0)
        {
            const int c1 = ent_cell[second[f]];
            m.face_cells[(size_t)f * 2 + 1] = c1;
            const V3 xc1 = {m.cell_x[(size_t)c1 * 3], m.cell_x[(size_t)c1 * 3 + 1],
                            m.cell_x[(size_t)c1 * 3 + 2]};
            const double d1 = dot(n, sub(xf, xc0)) / area;
            const double d2 = dot(n, sub(xc1, xf)) / area;
            if (!(d1 + d2 > 1e-16))
                fail("degenerate interpolation weight");
            m.face_lambda[f] = d1 / (d1 + d2);
        }
        else
        {
            m.face_lambda[f] = 1.0;
            bkeys.push_back(make_key(lf.type, vid, lf.nv, (uint32_t)f));
        }
    }
    std::sort(bkeys.begin(), bkeys.end(), key_less);

    // ---- boundaries in element order, linked to their faces -----------------
    const int B = (int)bnds.size();
    int maxbv = 0;
    for (const Element *e : bnds)
        maxbv = std::max(maxbv, (int)e->v.size());
    m.max_boundary_vertices = maxbv;
    for (int b = 0; b < B; ++b)
    {
        const Element &e = *bnds[b];
        const FaceKey k = make_key(e.type, e.v.data(), (int)e.v.size(), 0);
        auto it = std::lower_bound(bkeys.begin(), bkeys.end(), k, key_less);
        if (it == bkeys.end() || !key_equal(*it, k))
            fail("boundary element " + std::to_string(b) + " matches no boundary face");
        const int f = (int)it->ent;
        if (m.face_boundary[f] >= 0)
            fail("two boundary elements on one face");
        m.face_boundary[f] = b;
        m.bnd_region.push_back(e.region);
        m.bnd_type.push_back(e.type);
        m.bnd_nv.push_back((int)e.v.size());
        for (int q = 0; q < maxbv; ++q)
            m.bnd_vertices.push_back(q < (int)e.v.size() ? e.v[q] : -1);
        m.bnd_face.push_back(f);
        const V3 fn = {m.face_n[(size_t)f * 3], m.face_n[(size_t)f * 3 + 1],
                       m.face_n[(size_t)f * 3 + 2]};
        const V3 bn = neg(fn);
        const V3 bt1 = tangent1(bn), bt2 = unit(cross(bt1, bn));
        m.bnd_n.insert(m.bnd_n.end(), {bn.x, bn.y, bn.z});
        m.bnd_t1.insert(m.bnd_t1.end(), {bt1.x, bt1.y, bt1.z});
        m.bnd_t2.insert(m.bnd_t2.end(), {bt2.x, bt2.y, bt2.z});
        const int c0 = m.face_cells[(size_t)f * 2];
        const V3 dx = {m.cell_x[(size_t)c0 * 3] - m.face_x[(size_t)f * 3],
                       m.cell_x[(size_t)c0 * 3 + 1] - m.face_x[(size_t)f * 3 + 1],
                       m.cell_x[(size_t)c0 * 3 + 2] - m.face_x[(size_t)f * 3 + 2]};
        const double tmp = dot(dx, bn) / (len(dx) * len(bn));
        const double dist = len(dx) * std::max(std::min(tmp, 1.0), -1.0);
        if (dist < 1e-16)
            fail("non-positive boundary distance");
        m.bnd_distance.push_back(dist);
    }
    for (int f = 0; f < F; ++f)
        if (m.face_cells[(size_t)f * 2 + 1] < 0 && m.face_boundary[f] < 0)
            fail("boundary face " + std::to_string(f) + " has no boundary element");

    // ---- regions ------------------------------------------------------------
    m.region_names = p.regions;
    m.region_is_boundary.assign(p.regions.size(), 0);
    for (int r : m.bnd_region)
    {
        if (r < 0 || r >= (int)p.regions.size())
            fail("boundary element without a physical region");
        m.region_is_boundary[r] = 1;
    }
    for (int r : m.cell_region)
        if (r < 0 || r >= (int)p.regions.size())
            fail("cell without a physical region");

    // ---- cell face lists (ascending) and dx ----------------------------------
    m.cell_nf.assign(N, 0);
    m.cell_faces.assign((size_t)N * maxcf, -1);
    for (int f = 0; f < F; ++f)
        for (int side = 0; side < 2; ++side)
        {
            const int c = m.face_cells[(size_t)f * 2 + side];
            if (c < 0)
                continue;
            m.cell_faces[(size_t)c * maxcf + m.cell_nf[c]++] = f;
        }
    m.cell_dx.assign((size_t)N * 3, 0.0);
    for (int c = 0; c < N; ++c)
        for (int k = 0; k < m.cell_nf[c]; ++k)
        {
            const int f = m.cell_faces[(size_t)c * maxcf + k];
            for (int a = 0; a < 3; ++a)
                m.cell_dx[(size_t)c * 3 + a] +=
                    std::fabs(m.face_n[(size_t)f * 3 + a]) * m.face_area[f];
        }
    return N;
}

} // namespace fvh
