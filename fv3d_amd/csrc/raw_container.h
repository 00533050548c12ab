// Raw little-endian container used for .fvm meshes and .fvs states
// (see mesh_io.cpp for the format).
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

namespace fvh {

class RawWriter
{
  public:
    explicit RawWriter(const char *path);
    ~RawWriter();
    void ints(const std::string &name, const std::vector<int> &v,
              uint64_t cols = 0);
    void doubles(const std::string &name, const double *v, size_t n,
                 uint64_t cols = 0);
    void attr_int(const std::string &name, int v);
    void attr_double(const std::string &name, double v);
    void strings(const std::string &name, const std::vector<std::string> &v);

  private:
    void record(const std::string &name, char type,
                const std::vector<uint64_t> &dims, const void *data,
                size_t bytes);
    FILE *fp_ = nullptr;
};

class RawReader
{
  public:
    struct Entry
    {
        char type = 0;
        std::vector<uint64_t> dims;
        std::string data;
    };
    bool open(const char *path);
    bool has(const std::string &name) const;
    const Entry &get(const std::string &name) const;
    std::vector<int> ints(const std::string &name) const;
    std::vector<double> doubles(const std::string &name) const;
    std::vector<std::string> strings(const std::string &name) const;

  private:
    std::map<std::string, Entry> entries_;
};

} // namespace fvh
