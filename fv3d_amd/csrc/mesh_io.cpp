// Mesh and state files: the writer/reader choice (HDF5 for ".h5" names and
// HDF5 signatures, hdf5_io.cpp; otherwise the raw container), the mesh
// layout of tools/pyMeshFV3D/pymeshfv3d/write.py:19-202 read by
// src/Solver/src/mesh/mesh.c:121-630 (write_mesh_file / read_mesh_file),
// and the raw container itself.
//
// The raw little-endian container (.fvm mesh, .fvs state) holds the same
// groups and datasets as the reference's HDF5 files (the mesh layout above
// and the state layout of src/Solver/src/output/output.c:20-191) where
// libhdf5 is not available; dataset names are "GROUP/name", attributes are
// one-element datasets.  Format: "FV3DRAW1", then records {u32 name_len,
// name, u8 type ('i' int32, 'd' float64, 's' string), u32 ndim, u64
// dims[ndim], payload}.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>

#include "fvh_internal.h"
#include "raw_container.h"

namespace fvh {

RawWriter::RawWriter(const char *path)
{
    fp_ = std::fopen(path, "wb");
    if (!fp_)
        throw std::runtime_error(std::string("cannot create ") + path);
    std::fwrite("FV3DRAW1", 1, 8, fp_);
}

RawWriter::~RawWriter()
{
    if (fp_)
        std::fclose(fp_);
}

void RawWriter::record(const std::string &name, char type,
                       const std::vector<uint64_t> &dims, const void *data,
                       size_t bytes)
{
    uint32_t nl = (uint32_t)name.size();
    std::fwrite(&nl, 4, 1, fp_);
    std::fwrite(name.data(), 1, nl, fp_);
    std::fwrite(&type, 1, 1, fp_);
    uint32_t nd = (uint32_t)dims.size();
    std::fwrite(&nd, 4, 1, fp_);
    std::fwrite(dims.data(), 8, nd, fp_);
    if (bytes && std::fwrite(data, 1, bytes, fp_) != bytes)
        throw std::runtime_error("raw container: short write");
}

void RawWriter::ints(const std::string &name, const std::vector<int> &v,
                     uint64_t cols)
{
    uint64_t rows = cols ? v.size() / cols : v.size();
    std::vector<uint64_t> d = cols ? std::vector<uint64_t>{rows, cols}
                                   : std::vector<uint64_t>{rows};
    record(name, 'i', d, v.data(), v.size() * 4);
}

void RawWriter::doubles(const std::string &name, const double *v, size_t n,
                        uint64_t cols)
{
    uint64_t rows = cols ? n / cols : n;
    std::vector<uint64_t> d = cols ? std::vector<uint64_t>{rows, cols}
                                   : std::vector<uint64_t>{rows};
    record(name, 'd', d, v, n * 8);
}

void RawWriter::attr_int(const std::string &name, int v)
{
    record(name, 'i', {1}, &v, 4);
}

void RawWriter::attr_double(const std::string &name, double v)
{
    record(name, 'd', {1}, &v, 8);
}

void RawWriter::strings(const std::string &name,
                        const std::vector<std::string> &v)
{
    std::string blob;
    for (const auto &s : v)
    {
        blob += s;
        blob.push_back('\0');
    }
    record(name, 's', {(uint64_t)v.size()}, blob.data(), blob.size());
}

bool RawReader::open(const char *path)
{
    FILE *fp = std::fopen(path, "rb");
    if (!fp)
        return false;
    char magic[8];
    if (std::fread(magic, 1, 8, fp) != 8 || std::memcmp(magic, "FV3DRAW1", 8))
    {
        std::fclose(fp);
        return false;
    }
    while (true)
    {
        uint32_t nl;
        if (std::fread(&nl, 4, 1, fp) != 1)
            break;
        std::string name(nl, '\0');
        Entry e;
        uint32_t nd;
        if (std::fread(&name[0], 1, nl, fp) != nl ||
            std::fread(&e.type, 1, 1, fp) != 1 || std::fread(&nd, 4, 1, fp) != 1)
            throw std::runtime_error("raw container: truncated header");
        e.dims.resize(nd);
        if (nd && std::fread(e.dims.data(), 8, nd, fp) != nd)
            throw std::runtime_error("raw container: truncated dims");
        size_t count = 1;
        for (auto d : e.dims)
            count *= d;
        size_t bytes;
        if (e.type == 's')
        {
            // strings: NUL-separated blob; length is not stored, scan
            e.data.clear();
            size_t seen = 0;
            while (seen < count)
            {
                int ch = std::fgetc(fp);
                if (ch == EOF)
                    throw std::runtime_error("raw container: truncated strings");
                e.data.push_back((char)ch);
                if (ch == 0)
                    ++seen;
            }
            entries_[name] = std::move(e);
            continue;
        }
        bytes = count * (e.type == 'd' ? 8 : 4);
        e.data.resize(bytes);
        if (bytes && std::fread(&e.data[0], 1, bytes, fp) != bytes)
            throw std::runtime_error("raw container: truncated data");
        entries_[name] = std::move(e);
    }
    std::fclose(fp);
    return true;
}

const RawReader::Entry &RawReader::get(const std::string &name) const
{
    auto it = entries_.find(name);
    if (it == entries_.end())
        throw std::runtime_error("raw container: missing dataset " + name);
    return it->second;
}

bool RawReader::has(const std::string &name) const
{
    return entries_.count(name) != 0;
}

std::vector<int> RawReader::ints(const std::string &name) const
{
    const Entry &e = get(name);
    if (e.type != 'i')
        throw std::runtime_error("raw container: not int32: " + name);
    std::vector<int> v(e.data.size() / 4);
    std::memcpy(v.data(), e.data.data(), e.data.size());
    return v;
}

std::vector<double> RawReader::doubles(const std::string &name) const
{
    const Entry &e = get(name);
    if (e.type != 'd')
        throw std::runtime_error("raw container: not float64: " + name);
    std::vector<double> v(e.data.size() / 8);
    std::memcpy(v.data(), e.data.data(), e.data.size());
    return v;
}

std::vector<std::string> RawReader::strings(const std::string &name) const
{
    const Entry &e = get(name);
    std::vector<std::string> out;
    std::string cur;
    for (char ch : e.data)
    {
        if (ch == 0)
        {
            out.push_back(cur);
            cur.clear();
        }
        else
            cur.push_back(ch);
    }
    return out;
}

std::unique_ptr<DataWriter> open_writer(const char *path)
{
    const std::string p(path);
    if (p.size() > 3 && p.compare(p.size() - 3, 3, ".h5") == 0)
        return hdf5_writer(path);
    return std::unique_ptr<DataWriter>(new RawWriter(path));
}

std::unique_ptr<DataReader> open_reader(const char *path)
{
    FILE *fp = std::fopen(path, "rb");
    if (!fp)
        throw std::runtime_error(std::string("cannot open ") + path);
    unsigned char sig[8] = {0};
    const size_t got = std::fread(sig, 1, 8, fp);
    std::fclose(fp);
    static const unsigned char h5sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    if (got == 8 && std::memcmp(sig, h5sig, 8) == 0)
        return hdf5_reader(path);
    std::unique_ptr<RawReader> r(new RawReader());
    if (!r->open(path))
        throw std::runtime_error(std::string("not an HDF5 file or raw container: ") + path);
    return std::unique_ptr<DataReader>(r.release());
}

int write_mesh_file(const RawMesh &m, const char *path)
{
    std::unique_ptr<DataWriter> wp = open_writer(path);
    DataWriter &w = *wp;
    w.attr_int("dimension", m.dimension);
    w.attr_int("is_partitioned", 0);
    w.attr_int("VERTICES/n_vertices", (int)(m.vertices.size() / 3));
    w.doubles("VERTICES/x", m.vertices.data(), m.vertices.size(), 3);
    w.attr_int("CELLS/n_cells", m.n_cells());
    w.attr_int("CELLS/max_cell_vertices", m.max_cell_vertices);
    w.attr_int("CELLS/max_cell_faces", m.max_cell_faces);
    w.ints("CELLS/id", m.cell_region);
    w.ints("CELLS/type", m.cell_type);
    w.ints("CELLS/n_vertices", m.cell_nv);
    w.ints("CELLS/vertices", m.cell_vertices, m.max_cell_vertices);
    w.ints("CELLS/n_faces", m.cell_nf);
    w.ints("CELLS/faces", m.cell_faces, m.max_cell_faces);
    w.doubles("CELLS/volume", m.cell_volume.data(), m.cell_volume.size());
    w.doubles("CELLS/x", m.cell_x.data(), m.cell_x.size(), 3);
    w.doubles("CELLS/dx", m.cell_dx.data(), m.cell_dx.size(), 3);
    w.attr_int("FACES/n_faces", m.n_faces());
    w.attr_int("FACES/max_face_vertices", m.max_face_vertices);
    w.ints("FACES/type", m.face_type);
    w.ints("FACES/n_vertices", m.face_nv);
    w.ints("FACES/vertices", m.face_vertices, m.max_face_vertices);
    w.ints("FACES/cells", m.face_cells, 2);
    w.ints("FACES/boundary", m.face_boundary);
    w.doubles("FACES/x", m.face_x.data(), m.face_x.size(), 3);
    w.doubles("FACES/n", m.face_n.data(), m.face_n.size(), 3);
    w.doubles("FACES/t1", m.face_t1.data(), m.face_t1.size(), 3);
    w.doubles("FACES/t2", m.face_t2.data(), m.face_t2.size(), 3);
    w.doubles("FACES/area", m.face_area.data(), m.face_area.size());
    w.doubles("FACES/lambda", m.face_lambda.data(), m.face_lambda.size());
    w.doubles("FACES/periodic_shift", m.face_shift.data(), m.face_shift.size(), 3);
    w.attr_int("BOUNDARIES/n_boundaries", m.n_boundaries());
    w.attr_int("BOUNDARIES/max_boundary_vertices", m.max_boundary_vertices);
    w.ints("BOUNDARIES/id", m.bnd_region);
    w.ints("BOUNDARIES/type", m.bnd_type);
    w.ints("BOUNDARIES/n_vertices", m.bnd_nv);
    w.ints("BOUNDARIES/vertices", m.bnd_vertices, m.max_boundary_vertices);
    w.ints("BOUNDARIES/face", m.bnd_face);
    w.doubles("BOUNDARIES/n", m.bnd_n.data(), m.bnd_n.size(), 3);
    w.doubles("BOUNDARIES/t1", m.bnd_t1.data(), m.bnd_t1.size(), 3);
    w.doubles("BOUNDARIES/t2", m.bnd_t2.data(), m.bnd_t2.size(), 3);
    w.doubles("BOUNDARIES/distance", m.bnd_distance.data(), m.bnd_distance.size());
    w.attr_int("REGIONS/n_regions", (int)m.region_names.size());
    w.strings("REGIONS/name", m.region_names);
    w.ints("REGIONS/is_boundary", m.region_is_boundary);
    return 0;
}

int read_mesh_file(RawMesh &m, const char *path)
{
    std::unique_ptr<DataReader> rp = open_reader(path);
    const DataReader &r = *rp;
    m = RawMesh();
    m.dimension = r.ints("dimension")[0];
    m.vertices = r.doubles("VERTICES/x");
    m.max_cell_vertices = r.ints("CELLS/max_cell_vertices")[0];
    m.max_cell_faces = r.ints("CELLS/max_cell_faces")[0];
    m.cell_region = r.ints("CELLS/id");
    m.cell_type = r.ints("CELLS/type");
    m.cell_nv = r.ints("CELLS/n_vertices");
    m.cell_vertices = r.ints("CELLS/vertices");
    m.cell_nf = r.ints("CELLS/n_faces");
    m.cell_faces = r.ints("CELLS/faces");
    m.cell_volume = r.doubles("CELLS/volume");
    m.cell_x = r.doubles("CELLS/x");
    m.cell_dx = r.doubles("CELLS/dx");
    m.max_face_vertices = r.ints("FACES/max_face_vertices")[0];
    m.face_type = r.ints("FACES/type");
    m.face_nv = r.ints("FACES/n_vertices");
    m.face_vertices = r.ints("FACES/vertices");
    m.face_cells = r.ints("FACES/cells");
    m.face_boundary = r.ints("FACES/boundary");
    m.face_x = r.doubles("FACES/x");
    m.face_n = r.doubles("FACES/n");
    m.face_t1 = r.doubles("FACES/t1");
    m.face_t2 = r.doubles("FACES/t2");
    m.face_area = r.doubles("FACES/area");
    m.face_lambda = r.doubles("FACES/lambda");
    if (r.has("FACES/periodic_shift"))
        m.face_shift = r.doubles("FACES/periodic_shift");
    else
        m.face_shift.assign(m.face_area.size() * 3, 0.0);
    m.max_boundary_vertices = r.ints("BOUNDARIES/max_boundary_vertices")[0];
    m.bnd_region = r.ints("BOUNDARIES/id");
    m.bnd_type = r.ints("BOUNDARIES/type");
    m.bnd_nv = r.ints("BOUNDARIES/n_vertices");
    m.bnd_vertices = r.ints("BOUNDARIES/vertices");
    m.bnd_face = r.ints("BOUNDARIES/face");
    m.bnd_n = r.doubles("BOUNDARIES/n");
    m.bnd_t1 = r.doubles("BOUNDARIES/t1");
    m.bnd_t2 = r.doubles("BOUNDARIES/t2");
    m.bnd_distance = r.doubles("BOUNDARIES/distance");
    m.region_names = r.strings("REGIONS/name");
    m.region_is_boundary = r.ints("REGIONS/is_boundary");
    return 0;
}

} // namespace fvh
