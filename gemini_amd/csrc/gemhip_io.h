// gemhip_io.h -- host-side HDF5 file access used by the output and initial-condition
// entry points (implemented in h5io.cpp over the HDF5 C library, loaded at first use).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace gemhip {

// One dataset to write.  dims are C-order (slowest first): the reverse of the Fortran shape.
struct H5Out {
  enum Type { F32 = 0, I32 = 1 };
  const char* name;          // absolute path in the file; intermediate groups are created
  Type type;
  int rank;                  // 0 (scalar) .. 4
  long dims[4];
  const void* data;          // host memory
  hipEvent_t ready;          // if set, waited for before the data are read
};

// Create (truncate) path and write the datasets in order.  comp_lvl > 0 stores datasets of
// rank >= 2 chunked with the shuffle and deflate filters at that level.
bool h5_write(const char* path, const H5Out* ds, int nds, int comp_lvl, std::string& err);

// whether the HDF5 library could be loaded (err says why not)
bool h5_available(std::string& err);

// Read-only file.  Failures set the error string given to the constructor.
class H5In {
 public:
  H5In(const char* path, std::string& err);
  ~H5In();
  H5In(const H5In&) = delete;
  H5In& operator=(const H5In&) = delete;
  bool open() const { return fid_ >= 0; }
  // rank of dataset name and its C-order dims (up to 8), or -1 if absent
  int ndims(const char* name, long* dims) const;
  // Read the whole dataset converted to int (ints) or double.  With mdims, dst is a
  // memory array of rank mrank and C-order dims mdims and the dataset lands in the block
  // that starts at start (a dataset of lower rank fills the trailing dimensions);
  // without, dst receives the dataset contiguously.
  bool read(const char* name, bool ints, void* dst, int mrank = 0, const long* mdims = nullptr,
            const long* start = nullptr) const;

 private:
  int64_t fid_ = -1;
  std::string& err_;
};

}  // namespace gemhip
