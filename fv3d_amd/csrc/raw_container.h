// Mesh and state file access: the reference's HDF5 layout (groups,
// datasets, attributes named "GROUP/name") behind one reader/writer
// interface, with two back ends: HDF5 through libhdf5 loaded at run time
// (hdf5_io.cpp) and a raw little-endian container (.fvm meshes, .fvs
// states; see mesh_io.cpp for the format) where HDF5 is not available.
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace fvh {

class DataWriter
{
  public:
    virtual ~DataWriter() = default;
    // datasets: [n] or [n/cols][cols]
    virtual void ints(const std::string &name, const std::vector<int> &v,
                      uint64_t cols = 0) = 0;
    virtual void doubles(const std::string &name, const double *v, size_t n,
                         uint64_t cols = 0) = 0;
    // attributes (one value) of the group named before the last '/'
    virtual void attr_int(const std::string &name, int v) = 0;
    virtual void attr_double(const std::string &name, double v) = 0;
    virtual void strings(const std::string &name,
                         const std::vector<std::string> &v) = 0;
};

class DataReader
{
  public:
    virtual ~DataReader() = default;
    virtual bool has(const std::string &name) const = 0;
    // datasets or attributes
    virtual std::vector<int> ints(const std::string &name) const = 0;
    virtual std::vector<double> doubles(const std::string &name) const = 0;
    virtual std::vector<std::string> strings(const std::string &name) const = 0;
};

// Writer by file name: ".h5" -> HDF5, anything else -> raw container.
std::unique_ptr<DataWriter> open_writer(const char *path);
// Reader by file signature (HDF5 or raw container); throws if neither.
std::unique_ptr<DataReader> open_reader(const char *path);

// HDF5 back end (hdf5_io.cpp); available() loads libhdf5 on first use.
bool hdf5_available(std::string *why = nullptr);
std::unique_ptr<DataWriter> hdf5_writer(const char *path);
std::unique_ptr<DataReader> hdf5_reader(const char *path);

class RawWriter : public DataWriter
{
  public:
    explicit RawWriter(const char *path);
    ~RawWriter() override;
    void ints(const std::string &name, const std::vector<int> &v,
              uint64_t cols = 0) override;
    void doubles(const std::string &name, const double *v, size_t n,
                 uint64_t cols = 0) override;
    void attr_int(const std::string &name, int v) override;
    void attr_double(const std::string &name, double v) override;
    void strings(const std::string &name, const std::vector<std::string> &v) override;

  private:
    void record(const std::string &name, char type,
                const std::vector<uint64_t> &dims, const void *data,
                size_t bytes);
    FILE *fp_ = nullptr;
};

class RawReader : public DataReader
{
  public:
    struct Entry
    {
        char type = 0;
        std::vector<uint64_t> dims;
        std::string data;
    };
    bool open(const char *path);
    bool has(const std::string &name) const override;
    const Entry &get(const std::string &name) const;
    std::vector<int> ints(const std::string &name) const override;
    std::vector<double> doubles(const std::string &name) const override;
    std::vector<std::string> strings(const std::string &name) const override;

  private:
    std::map<std::string, Entry> entries_;
};

} // namespace fvh
