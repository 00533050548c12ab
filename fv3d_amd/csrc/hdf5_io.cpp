// HDF5 back end of the mesh/state file interface, in the reference's layout
// (mesh: tools/pyMeshFV3D/pymeshfv3d/write.py:19-202 read by
// src/Solver/src/mesh/mesh.c:121-630; state: src/Solver/src/output/
// output.c:20-191 read by src/Solver/src/restart/restart.c:50-178).
//
// libhdf5 (1.10 ABI) is loaded at run time with dlopen, so the library has
// no link-time dependency on it: FV3D_HDF5_LIB names the shared object,
// else libhdf5.so.103 / libhdf5.so on the loader path, else the conda copy.
// Names "GROUP/x" address dataset or attribute x of group GROUP (created on
// write); names without '/' belong to the file's root group.
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fvh_internal.h"
#include "raw_container.h"

namespace fvh {

namespace {

// HDF5 1.10 C ABI types and constants (H5Ipublic.h, H5public.h,
// H5Fpublic.h, H5Ppublic.h, H5Spublic.h, H5Tpublic.h)
typedef int64_t hid_t;
typedef int herr_t;
typedef int htri_t;
typedef unsigned long long hsize_t;
constexpr unsigned kAccRdonly = 0x0000u, kAccTrunc = 0x0002u;
constexpr hid_t kDefault = 0, kAll = 0;
constexpr int kClassString = 3;

struct H5
{
    void *so = nullptr;
    std::string error;
    herr_t (*open)();
    herr_t (*eset_auto)(hid_t, void *, void *);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*gcreate)(hid_t, const char *, hid_t, hid_t, hid_t);
    hid_t (*gopen)(hid_t, const char *, hid_t);
    herr_t (*gclose)(hid_t);
    htri_t (*lexists)(hid_t, const char *, hid_t);
    htri_t (*aexists)(hid_t, const char *);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    int (*sget_ndims)(hid_t);
    int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
    herr_t (*sclose)(hid_t);
    hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen)(hid_t, const char *, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    hid_t (*dget_space)(hid_t);
    hid_t (*dget_type)(hid_t);
    herr_t (*dclose)(hid_t);
    hid_t (*acreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
    hid_t (*aopen)(hid_t, const char *, hid_t);
    herr_t (*awrite)(hid_t, hid_t, const void *);
    herr_t (*aread)(hid_t, hid_t, void *);
    hid_t (*aget_space)(hid_t);
    hid_t (*aget_type)(hid_t);
    herr_t (*aclose)(hid_t);
    hid_t (*tcopy)(hid_t);
    herr_t (*tset_size)(hid_t, size_t);
    size_t (*tget_size)(hid_t);
    int (*tget_class)(hid_t);
    htri_t (*tis_variable_str)(hid_t);
    herr_t (*tclose)(hid_t);
    const hid_t *native_int, *native_double, *c_s1;
};

H5 *g_h5 = nullptr;

template <class F> void sym(H5 &h, F &f, const char *name)
{
    f = reinterpret_cast<F>(dlsym(h.so, name));
    if (!f && h.error.empty())
        h.error = std::string("libhdf5 lacks ") + name;
}

H5 *load()
{
    if (g_h5)
        return g_h5;
    H5 *h = new H5();
    std::vector<std::string> names;
    if (const char *e = std::getenv("FV3D_HDF5_LIB"))
        names.push_back(e);
    names.push_back("libhdf5.so.103");
    names.push_back("libhdf5.so");
    names.push_back("/opt/conda/lib/libhdf5.so.103");
    for (const std::string &n : names)
        if ((h->so = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL)))
            break;
    if (!h->so)
    {
        h->error = "libhdf5 (1.10) not found (set FV3D_HDF5_LIB)";
        g_h5 = h;
        return h;
    }
    sym(*h, h->open, "H5open");
    sym(*h, h->eset_auto, "H5Eset_auto2");
    sym(*h, h->fcreate, "H5Fcreate");
    sym(*h, h->fopen, "H5Fopen");
    sym(*h, h->fclose, "H5Fclose");
    sym(*h, h->gcreate, "H5Gcreate2");
    sym(*h, h->gopen, "H5Gopen2");
    sym(*h, h->gclose, "H5Gclose");
    sym(*h, h->lexists, "H5Lexists");
    sym(*h, h->aexists, "H5Aexists");
    sym(*h, h->screate_simple, "H5Screate_simple");
    sym(*h, h->sget_ndims, "H5Sget_simple_extent_ndims");
    sym(*h, h->sget_dims, "H5Sget_simple_extent_dims");
    sym(*h, h->sclose, "H5Sclose");
    sym(*h, h->dcreate, "H5Dcreate2");
    sym(*h, h->dopen, "H5Dopen2");
    sym(*h, h->dwrite, "H5Dwrite");
    sym(*h, h->dread, "H5Dread");
    sym(*h, h->dget_space, "H5Dget_space");
    sym(*h, h->dget_type, "H5Dget_type");
    sym(*h, h->dclose, "H5Dclose");
    sym(*h, h->acreate, "H5Acreate2");
    sym(*h, h->aopen, "H5Aopen");
    sym(*h, h->awrite, "H5Awrite");
    sym(*h, h->aread, "H5Aread");
    sym(*h, h->aget_space, "H5Aget_space");
    sym(*h, h->aget_type, "H5Aget_type");
    sym(*h, h->aclose, "H5Aclose");
    sym(*h, h->tcopy, "H5Tcopy");
    sym(*h, h->tset_size, "H5Tset_size");
    sym(*h, h->tget_size, "H5Tget_size");
    sym(*h, h->tget_class, "H5Tget_class");
    sym(*h, h->tis_variable_str, "H5Tis_variable_str");
    sym(*h, h->tclose, "H5Tclose");
    h->native_int = reinterpret_cast<const hid_t *>(dlsym(h->so, "H5T_NATIVE_INT_g"));
    h->native_double = reinterpret_cast<const hid_t *>(dlsym(h->so, "H5T_NATIVE_DOUBLE_g"));
    h->c_s1 = reinterpret_cast<const hid_t *>(dlsym(h->so, "H5T_C_S1_g"));
    if (!h->native_int || !h->native_double || !h->c_s1)
        h->error = "libhdf5 lacks the native type globals";
    if (h->error.empty())
    {
        if (h->open() < 0)
            h->error = "H5open failed";
        else
            h->eset_auto(kDefault, nullptr, nullptr); // errors become return codes
    }
    g_h5 = h;
    return h;
}

H5 &lib()
{
    H5 *h = load();
    if (!h->error.empty())
        throw std::runtime_error("HDF5: " + h->error);
    return *h;
}

void ck(bool ok, const std::string &what)
{
    if (!ok)
        throw std::runtime_error("HDF5: " + what);
}

void split(const std::string &name, std::string &group, std::string &leaf)
{
    const size_t p = name.rfind('/');
    group = p == std::string::npos ? "" : name.substr(0, p);
    leaf = p == std::string::npos ? name : name.substr(p + 1);
}

class H5Writer : public DataWriter
{
  public:
    explicit H5Writer(const char *path) : h_(lib())
    {
        file_ = h_.fcreate(path, kAccTrunc, kDefault, kDefault);
        ck(file_ >= 0, std::string("cannot create ") + path);
    }
    ~H5Writer() override
    {
        if (file_ >= 0)
            h_.fclose(file_);
    }
    void ints(const std::string &name, const std::vector<int> &v, uint64_t cols) override
    {
        dataset(name, *h_.native_int, *h_.native_int, v.data(), v.size(), cols);
    }
    void doubles(const std::string &name, const double *v, size_t n, uint64_t cols) override
    {
        dataset(name, *h_.native_double, *h_.native_double, v, n, cols);
    }
    void attr_int(const std::string &name, int v) override
    {
        attribute(name, *h_.native_int, &v);
    }
    void attr_double(const std::string &name, double v) override
    {
        attribute(name, *h_.native_double, &v);
    }
    void strings(const std::string &name, const std::vector<std::string> &v) override
    {
        size_t len = 1;
        for (const std::string &s : v)
            len = std::max(len, s.size() + 1);
        std::vector<char> buf(len * std::max<size_t>(v.size(), 1), '\0');
        for (size_t i = 0; i < v.size(); ++i)
            std::memcpy(&buf[i * len], v[i].data(), v[i].size());
        const hid_t t = h_.tcopy(*h_.c_s1);
        ck(t >= 0 && h_.tset_size(t, len) >= 0, "string type");
        dataset(name, t, t, buf.data(), v.size(), 0);
        h_.tclose(t);
    }

  private:
    hid_t group(const std::string &g)
    {
        if (g.empty())
            return file_;
        // nested groups: create each level that does not exist
        hid_t cur = file_;
        size_t pos = 0;
        hid_t out = -1;
        while (pos <= g.size())
        {
            const size_t q = g.find('/', pos);
            const std::string part = g.substr(0, q == std::string::npos ? g.size() : q);
            const hid_t id = h_.lexists(file_, part.c_str(), kDefault) > 0
                                 ? h_.gopen(file_, part.c_str(), kDefault)
                                 : h_.gcreate(file_, part.c_str(), kDefault, kDefault, kDefault);
            ck(id >= 0, "group " + part);
            if (out >= 0 && out != file_)
                h_.gclose(out);
            out = id;
            if (q == std::string::npos)
                break;
            pos = q + 1;
        }
        (void)cur;
        return out;
    }
    void dataset(const std::string &name, hid_t ftype, hid_t mtype, const void *data, size_t n,
                 uint64_t cols)
    {
        std::string g, leaf;
        split(name, g, leaf);
        const hid_t gid = group(g);
        hsize_t dims[2] = {(hsize_t)(cols ? n / cols : n), (hsize_t)cols};
        const hid_t sp = h_.screate_simple(cols ? 2 : 1, dims, nullptr);
        ck(sp >= 0, "dataspace " + name);
        const hid_t ds = h_.dcreate(gid, leaf.c_str(), ftype, sp, kDefault, kDefault, kDefault);
        ck(ds >= 0, "dataset " + name);
        ck(n == 0 || h_.dwrite(ds, mtype, kAll, kAll, kDefault, data) >= 0, "write " + name);
        h_.dclose(ds);
        h_.sclose(sp);
        if (gid != file_)
            h_.gclose(gid);
    }
    void attribute(const std::string &name, hid_t type, const void *v)
    {
        std::string g, leaf;
        split(name, g, leaf);
        const hid_t gid = group(g);
        const hsize_t one = 1;
        const hid_t sp = h_.screate_simple(1, &one, nullptr);
        const hid_t a = h_.acreate(gid, leaf.c_str(), type, sp, kDefault, kDefault);
        ck(a >= 0, "attribute " + name);
        ck(h_.awrite(a, type, v) >= 0, "write " + name);
        h_.aclose(a);
        h_.sclose(sp);
        if (gid != file_)
            h_.gclose(gid);
    }
    H5 &h_;
    hid_t file_ = -1;
};

class H5Reader : public DataReader
{
  public:
    explicit H5Reader(const char *path) : h_(lib())
    {
        file_ = h_.fopen(path, kAccRdonly, kDefault);
        ck(file_ >= 0, std::string("cannot open ") + path);
    }
    ~H5Reader() override
    {
        if (file_ >= 0)
            h_.fclose(file_);
    }
    bool has(const std::string &name) const override
    {
        return is_dataset(name) || is_attribute(name);
    }
    std::vector<int> ints(const std::string &name) const override
    {
        std::vector<int> v;
        read(name, *h_.native_int, v);
        return v;
    }
    std::vector<double> doubles(const std::string &name) const override
    {
        std::vector<double> v;
        read(name, *h_.native_double, v);
        return v;
    }
    std::vector<std::string> strings(const std::string &name) const override
    {
        const hid_t ds = h_.dopen(file_, name.c_str(), kDefault);
        ck(ds >= 0, "missing dataset " + name);
        const hid_t ft = h_.dget_type(ds);
        ck(h_.tget_class(ft) == kClassString && h_.tis_variable_str(ft) == 0,
           name + ": not a fixed-length string dataset");
        const size_t len = h_.tget_size(ft);
        const size_t n = count(h_.dget_space(ds), true);
        std::vector<char> buf(len * n + 1, '\0');
        const hid_t mt = h_.tcopy(*h_.c_s1);
        h_.tset_size(mt, len);
        ck(n == 0 || h_.dread(ds, mt, kAll, kAll, kDefault, buf.data()) >= 0, "read " + name);
        h_.tclose(mt);
        h_.tclose(ft);
        h_.dclose(ds);
        std::vector<std::string> out;
        for (size_t i = 0; i < n; ++i)
            out.push_back(std::string(&buf[i * len], strnlen(&buf[i * len], len)));
        return out;
    }

  private:
    bool is_dataset(const std::string &name) const
    {
        // every level of the path must exist for H5Lexists
        size_t pos = 0;
        while (true)
        {
            const size_t q = name.find('/', pos);
            const std::string part = name.substr(0, q);
            if (h_.lexists(file_, part.c_str(), kDefault) <= 0)
                return false;
            if (q == std::string::npos)
                return true;
            pos = q + 1;
        }
    }
    bool is_attribute(const std::string &name) const
    {
        std::string g, leaf;
        split(name, g, leaf);
        if (!g.empty() && !is_dataset(g))
            return false;
        const hid_t gid = g.empty() ? file_ : h_.gopen(file_, g.c_str(), kDefault);
        if (gid < 0)
            return false;
        const bool ok = h_.aexists(gid, leaf.c_str()) > 0;
        if (gid != file_)
            h_.gclose(gid);
        return ok;
    }
    size_t count(hid_t sp, bool close) const
    {
        const int nd = h_.sget_ndims(sp);
        hsize_t dims[8] = {0};
        ck(nd >= 0 && nd <= 8, "dataspace rank");
        h_.sget_dims(sp, dims, nullptr);
        size_t n = 1;
        for (int i = 0; i < nd; ++i)
            n *= (size_t)dims[i];
        if (close)
            h_.sclose(sp);
        return n;
    }
    template <class T> void read(const std::string &name, hid_t mtype, std::vector<T> &v) const
    {
        if (is_dataset(name))
        {
            const hid_t ds = h_.dopen(file_, name.c_str(), kDefault);
            ck(ds >= 0, "open " + name);
            v.resize(count(h_.dget_space(ds), true));
            ck(v.empty() || h_.dread(ds, mtype, kAll, kAll, kDefault, v.data()) >= 0,
               "read " + name);
            h_.dclose(ds);
            return;
        }
        std::string g, leaf;
        split(name, g, leaf);
        ck(is_attribute(name), "missing dataset or attribute " + name);
        const hid_t gid = g.empty() ? file_ : h_.gopen(file_, g.c_str(), kDefault);
        const hid_t a = h_.aopen(gid, leaf.c_str(), kDefault);
        ck(a >= 0, "open attribute " + name);
        v.resize(count(h_.aget_space(a), true));
        ck(h_.aread(a, mtype, v.data()) >= 0, "read attribute " + name);
        h_.aclose(a);
        if (gid != file_)
            h_.gclose(gid);
    }
    H5 &h_;
    hid_t file_ = -1;
};

} // namespace

bool hdf5_available(std::string *why)
{
    H5 *h = load();
    if (why)
        *why = h->error;
    return h->error.empty();
}

std::unique_ptr<DataWriter> hdf5_writer(const char *path)
{
    return std::unique_ptr<DataWriter>(new H5Writer(path));
}

std::unique_ptr<DataReader> hdf5_reader(const char *path)
{
    return std::unique_ptr<DataReader>(new H5Reader(path));
}

} // namespace fvh
