// Minimal HDF5 access for the output file, with libhdf5 loaded at run time
// (dlopen), so the library itself has no link-time HDF5 dependency.  The
// dataset shapes and the "unit" attribute follow the reference's writers
// (src/IO/HDF5base.cpp: writeBuffer :74-128, writeSingleNode :222-260,
// writeSingleNodeInt :262-288, writeSingleNodeString :290-333).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace g4mi {

class H5File {
  public:
    // Creates (truncates) `path`; ok() is false and error() says why when
    // libhdf5 cannot be loaded or the file cannot be created.
    explicit H5File(const std::string &path);
    ~H5File();
    bool ok() const { return fid_ >= 0; }
    const std::string &error() const { return err_; }

    // Datasets by slash-separated path below the root; intermediate groups
    // are created on first use.
    // 2-D [nrec][ncol] doubles with a "unit" attribute (writeBuffer).
    bool write_buffer(const std::string &path, const std::string &unit, const std::vector<double> &v,
                      size_t nrec, size_t ncol);
    // 1-D doubles with a "unit" attribute (writeSingleNode).
    bool write_single(const std::string &path, const std::string &unit, const std::vector<double> &v);
    // 1-D ints (writeSingleNodeInt).
    bool write_int(const std::string &path, const std::vector<int> &v);
    // One fixed-length string (writeSingleNodeString).
    bool write_string(const std::string &path, const std::string &s);
    void close();

  private:
    int64_t group_of(const std::string &path, std::string *leaf);
    int64_t fid_ = -1;
    std::string err_;
    std::vector<std::pair<std::string, int64_t>> groups_;
};

// Reads a double dataset (rank 1 or 2) into out; dims[0..1] get its shape
// (dims[1] = 0 for rank 1).  Returns the element count, or -1 with err set.
long long h5_read_doubles(const std::string &file, const std::string &dataset, std::vector<double> &out,
                          long long dims[2], std::string &err);
// Whether libhdf5 can be loaded on this host.
bool h5_available(std::string *err);

}  // namespace g4mi
