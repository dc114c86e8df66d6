// h5io.cpp -- HDF5 files of the reference's input and output formats, through the HDF5 C
// library loaded at first use (dlopen), so the transport library itself has no load-time
// dependency on it.  The library is looked up as $GEMHIP_HDF5_LIB, then libhdf5.so on the
// loader's path, then the copy that ships with the image's conda tree.
//
// Layout convention (that of the reference's Fortran writer): a Fortran array of shape
// (n1, n2, ..., nk) is one dataset of C dimensions (nk, ..., n2, n1) holding the same bytes,
// so the x1-fastest interior blocks are written and read without any transposition.
// Scalars are rank-0 datasets; integers are native 32-bit, reals native 32-bit floats.
#include "gemhip_io.h"

#include <dlfcn.h>
#include <hdf5.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <thread>
#include <vector>

namespace gemhip {
namespace {

// zlib's one-shot compressor (the call HDF5's deflate filter makes), found through the
// HDF5 library's own dependency on zlib
using compress2_fn = int (*)(unsigned char*, unsigned long*, const unsigned char*, unsigned long,
                             int);
using compress_bound_fn = unsigned long (*)(unsigned long);

// Function table of the HDF5 calls used here, resolved by name.  The header declares the
// functions; decltype keeps every pointer's type that of the declaration.
struct H5Api {
  bool ok = false;
  std::string why;
#define H5FN(name) decltype(&::name) name = nullptr;
  H5FN(H5open)
  H5FN(H5get_libversion)
  H5FN(H5Eget_auto2)
  H5FN(H5Eset_auto2)
  H5FN(H5Fcreate)
  H5FN(H5Fopen)
  H5FN(H5Fclose)
  H5FN(H5Gcreate2)
  H5FN(H5Gclose)
  H5FN(H5Lexists)
  H5FN(H5Screate)
  H5FN(H5Screate_simple)
  H5FN(H5Sclose)
  H5FN(H5Sselect_hyperslab)
  H5FN(H5Sget_simple_extent_ndims)
  H5FN(H5Sget_simple_extent_dims)
  H5FN(H5Pcreate)
  H5FN(H5Pclose)
  H5FN(H5Pset_chunk)
  H5FN(H5Pset_shuffle)
  H5FN(H5Pset_deflate)
  H5FN(H5Dcreate2)
  H5FN(H5Dopen2)
  H5FN(H5Dclose)
  H5FN(H5Dget_space)
  H5FN(H5Dwrite)
  H5FN(H5Dread)
#undef H5FN
  // type and property-list class identifiers: library globals, valid after H5open
  const hid_t* native_float = nullptr;
  const hid_t* native_double = nullptr;
  const hid_t* native_int = nullptr;
  const hid_t* dataset_create = nullptr;
  // optional: pre-filtered chunk writes (HDF5 >= 1.10.3) and zlib, for parallel compression
  decltype(&::H5Dwrite_chunk) H5Dwrite_chunk = nullptr;
  compress2_fn z_compress2 = nullptr;
  compress_bound_fn z_bound = nullptr;
};

const H5Api& api() {
  static H5Api a = [] {
    H5Api t;
    void* h = nullptr;
    const char* env = getenv("GEMHIP_HDF5_LIB");
    for (const char* p : {env, "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"}) {
      if (p && *p && (h = dlopen(p, RTLD_NOW | RTLD_LOCAL))) break;
    }
    if (!h) {
      t.why = "the HDF5 library (libhdf5.so) could not be loaded; set GEMHIP_HDF5_LIB to its path";
      return t;
    }
    bool all = true;
    auto sym = [&](const char* name) {
      void* s = dlsym(h, name);
      if (!s) {
        all = false;
        if (t.why.empty()) t.why = std::string("libhdf5 lacks ") + name;
      }
      return s;
    };
#define H5FN(name) t.name = reinterpret_cast<decltype(t.name)>(sym(#name));
    H5FN(H5open)
    H5FN(H5get_libversion)
    H5FN(H5Eget_auto2)
    H5FN(H5Eset_auto2)
    H5FN(H5Fcreate)
    H5FN(H5Fopen)
    H5FN(H5Fclose)
    H5FN(H5Gcreate2)
    H5FN(H5Gclose)
    H5FN(H5Lexists)
    H5FN(H5Screate)
    H5FN(H5Screate_simple)
    H5FN(H5Sclose)
    H5FN(H5Sselect_hyperslab)
    H5FN(H5Sget_simple_extent_ndims)
    H5FN(H5Sget_simple_extent_dims)
    H5FN(H5Pcreate)
    H5FN(H5Pclose)
    H5FN(H5Pset_chunk)
    H5FN(H5Pset_shuffle)
    H5FN(H5Pset_deflate)
    H5FN(H5Dcreate2)
    H5FN(H5Dopen2)
    H5FN(H5Dclose)
    H5FN(H5Dget_space)
    H5FN(H5Dwrite)
    H5FN(H5Dread)
#undef H5FN
    t.native_float = static_cast<const hid_t*>(sym("H5T_NATIVE_FLOAT_g"));
    t.native_double = static_cast<const hid_t*>(sym("H5T_NATIVE_DOUBLE_g"));
    t.native_int = static_cast<const hid_t*>(sym("H5T_NATIVE_INT_g"));
    t.dataset_create = static_cast<const hid_t*>(sym("H5P_CLS_DATASET_CREATE_ID_g"));
    if (!all) return t;
    t.H5Dwrite_chunk = reinterpret_cast<decltype(t.H5Dwrite_chunk)>(dlsym(h, "H5Dwrite_chunk"));
    t.z_compress2 = reinterpret_cast<compress2_fn>(dlsym(h, "compress2"));
    t.z_bound = reinterpret_cast<compress_bound_fn>(dlsym(h, "compressBound"));
    unsigned maj = 0, min = 0, rel = 0;
    if (t.H5open() < 0 || t.H5get_libversion(&maj, &min, &rel) < 0) {
      t.why = "HDF5 library initialisation failed";
      return t;
    }
    if (maj < 1 || (maj == 1 && min < 10)) {   // hid_t is 64-bit from 1.10 on
      t.why = "HDF5 1.10 or newer is needed, found " + std::to_string(maj) + "." +
              std::to_string(min) + "." + std::to_string(rel);
      return t;
    }
    t.ok = true;
    return t;
  }();
  return a;
}

bool have(std::string& err) {
  const H5Api& a = api();
  if (!a.ok) err = a.why;
  return a.ok;
}

// HDF5's automatic error printing is off while a call of this file runs (its failures are
// reported through return codes) and restored afterwards: the setting is process-wide, and
// the host may use the same HDF5 library for its own files.
class QuietErrors {
 public:
  QuietErrors() {
    const H5Api& a = api();
    if (!a.ok) return;
    on_ = a.H5Eget_auto2(H5E_DEFAULT, &func_, &data_) >= 0;
    if (on_) a.H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~QuietErrors() {
    if (on_) api().H5Eset_auto2(H5E_DEFAULT, func_, data_);
  }
  QuietErrors(const QuietErrors&) = delete;
  QuietErrors& operator=(const QuietErrors&) = delete;

 private:
  bool on_ = false;
  H5E_auto2_t func_ = nullptr;
  void* data_ = nullptr;
};

// chunk shape for a compressed dataset: leading dimensions cut to 1 while the chunk is over
// 1 MiB, then the second-fastest dimension cut to fit (whole rows of the fastest one)
void chunk_dims(int rank, const long* dims, size_t elem, hsize_t* ch) {
  for (int i = 0; i < rank; ++i) ch[i] = (hsize_t)(dims[i] > 0 ? dims[i] : 1);
  const size_t cap = 1 << 20;
  auto bytes = [&] {
    size_t b = elem;
    for (int i = 0; i < rank; ++i) b *= ch[i];
    return b;
  };
  for (int i = 0; i < rank - 2 && bytes() > cap; ++i) ch[i] = 1;
  if (bytes() > cap && rank >= 2) {
    const size_t row = elem * ch[rank - 1];
    ch[rank - 2] = (hsize_t)(cap / row > 0 ? cap / row : 1);
  }
}

int io_threads() {
  for (const char* v : {"GEMHIP_IO_THREADS", "OMP_NUM_THREADS"}) {
    const char* e = getenv(v);
    if (e && atoi(e) > 0) return std::min(atoi(e), 64);
  }
  const unsigned h = std::thread::hardware_concurrency();
  return h ? (int)std::min(h, 64u) : 1;
}

// Write a float dataset created with the shuffle and deflate filters by compressing its
// chunks on a thread pool and handing HDF5 the filtered bytes (H5Dwrite_chunk, filter mask
// 0: every filter of the pipeline applied).  Each chunk is gathered (zero-padded where it
// passes the dataset's edge), byte-shuffled as the shuffle filter does (byte b of element
// i to position b*n + i) and compressed with compress2 at the deflate level, which is what
// HDF5's own filters would store for it.
bool write_chunks_parallel(hid_t dset, int rank, const long* dims, const hsize_t* ch,
                           const float* data, int level, std::string& err) {
  const H5Api& a = api();
  long nc[4], nchunks = 1, ce = 1;
  for (int k = 0; k < rank; ++k) {
    nc[k] = (dims[k] + (long)ch[k] - 1) / (long)ch[k];
    nchunks *= nc[k];
    ce *= (long)ch[k];
  }
  const size_t raw = (size_t)ce * sizeof(float);
  std::vector<std::vector<unsigned char>> out((size_t)nchunks);
  std::vector<int> bad((size_t)nchunks, 0);
  const int nt = (int)std::min<long>(io_threads(), nchunks);
  // chunks are taken from a shared counter, so whichever threads start finish them all
  std::atomic<long> next(0);
  auto work = [&]() {
    std::vector<float> buf((size_t)ce);
    std::vector<unsigned char> shuf(raw);
    const unsigned long cap = a.z_bound((unsigned long)raw);
    for (long j = next++; j < nchunks; j = next++) {
      long idx[4], off[4];
      long r = j;
      for (int k = rank - 1; k >= 0; --k) { idx[k] = r % nc[k]; r /= nc[k]; off[k] = idx[k] * (long)ch[k]; }
      // gather rows of the last dimension
      const long rowlen = (long)ch[rank - 1];
      const long valid = std::min(rowlen, dims[rank - 1] - off[rank - 1]);
      const long nrows = ce / rowlen;
      for (long q = 0; q < nrows; ++q) {
        long rem = q, src = 0, stride = 1;
        bool inside = true;
        long pos[4];
        for (int k = rank - 2; k >= 0; --k) { pos[k] = rem % (long)ch[k]; rem /= (long)ch[k]; }
        for (int k = rank - 1; k >= 0; --k) {
          const long c = k == rank - 1 ? off[k] : off[k] + pos[k];
          if (c >= dims[k]) inside = false;
          src += c * stride;
          stride *= dims[k];
        }
        float* dst = buf.data() + q * rowlen;
        if (inside) {
          memcpy(dst, data + src, (size_t)valid * sizeof(float));
          std::fill(dst + valid, dst + rowlen, 0.0f);
        } else {
          std::fill(dst, dst + rowlen, 0.0f);
        }
      }
      const unsigned char* bytes = reinterpret_cast<const unsigned char*>(buf.data());
      const unsigned char* src = bytes;
      if (ce > 1) {
        for (long i = 0; i < ce; ++i)
          for (int b = 0; b < 4; ++b) shuf[(size_t)b * ce + i] = bytes[(size_t)i * 4 + b];
        src = shuf.data();
      }
      std::vector<unsigned char>& o = out[(size_t)j];
      o.resize(cap);
      unsigned long len = cap;
      if (a.z_compress2(o.data(), &len, src, (unsigned long)raw, level) != 0) { bad[(size_t)j] = 1; continue; }
      o.resize(len);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) {
    try {
      pool.emplace_back(work);
    } catch (const std::exception&) {
      break;   // fewer threads: the calling thread still drains the counter
    }
  }
  work();
  for (auto& th : pool) th.join();
  for (long j = 0; j < nchunks; ++j) {
    if (bad[(size_t)j]) { err = "chunk compression failed"; return false; }
    hsize_t off[4];
    long r = j;
    for (int k = rank - 1; k >= 0; --k) { off[k] = (hsize_t)((r % nc[k]) * (long)ch[k]); r /= nc[k]; }
    if (a.H5Dwrite_chunk(dset, H5P_DEFAULT, 0, off, out[(size_t)j].size(), out[(size_t)j].data()) < 0) {
      err = "could not write a compressed chunk";
      return false;
    }
    std::vector<unsigned char>().swap(out[(size_t)j]);
  }
  return true;
}

// the access flags of H5Fpublic.h without the macros' library calls (H5check, H5open)
constexpr unsigned ACC_RDONLY = 0x0000u, ACC_TRUNC = 0x0002u;

}  // namespace

bool h5_write(const char* path, const H5Out* ds, int nds, int comp_lvl, std::string& err) {
  if (!have(err)) return false;
  QuietErrors quiet;
  const H5Api& a = api();
  const hid_t f = a.H5Fcreate(path, ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) {
    err = std::string("could not create ") + path;
    return false;
  }
  bool ok = true;
  std::set<std::string> groups;
  for (int i = 0; i < nds && ok; ++i) {
    const H5Out& d = ds[i];
    std::string name = d.name;
    // intermediate groups ("/time/ymd" needs "/time")
    for (size_t p = name.find('/', 1); p != std::string::npos; p = name.find('/', p + 1)) {
      const std::string gname = name.substr(0, p);
      if (groups.count(gname)) continue;
      const hid_t gid = a.H5Gcreate2(f, gname.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
      if (gid < 0) { err = "could not create group " + gname; ok = false; break; }
      a.H5Gclose(gid);
      groups.insert(gname);
    }
    if (!ok) break;
    hsize_t dims[4];
    for (int k = 0; k < d.rank; ++k) dims[k] = (hsize_t)d.dims[k];
    const hid_t space = d.rank == 0 ? a.H5Screate(H5S_SCALAR) : a.H5Screate_simple(d.rank, dims, nullptr);
    const hid_t type = d.type == H5Out::I32 ? *a.native_int : *a.native_float;
    hid_t dcpl = H5P_DEFAULT;
    if (comp_lvl > 0 && d.rank >= 2) {
      hsize_t ch[4];
      chunk_dims(d.rank, d.dims, 4, ch);
      dcpl = a.H5Pcreate(*a.dataset_create);
      if (dcpl < 0 || a.H5Pset_chunk(dcpl, d.rank, ch) < 0 || a.H5Pset_shuffle(dcpl) < 0 ||
          a.H5Pset_deflate(dcpl, (unsigned)comp_lvl) < 0) {
        err = "could not set up compression for " + name;
        ok = false;
      }
    }
    hid_t dset = -1;
    if (ok && space >= 0) dset = a.H5Dcreate2(f, name.c_str(), type, space, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    if (ok && dset < 0) { err = "could not create dataset " + name; ok = false; }
    if (ok && d.ready && hipEventSynchronize(d.ready) != hipSuccess) {
      err = "waiting for the device copy of " + name + " failed";
      ok = false;
    }
    const bool parallel = ok && dcpl != H5P_DEFAULT && d.type == H5Out::F32 && a.H5Dwrite_chunk &&
                          a.z_compress2 && a.z_bound;
    if (parallel) {
      hsize_t ch[4];
      chunk_dims(d.rank, d.dims, 4, ch);
      if (!write_chunks_parallel(dset, d.rank, d.dims, ch, static_cast<const float*>(d.data),
                                 comp_lvl, err)) {
        err = name + ": " + err;
        ok = false;
      }
    } else if (ok && a.H5Dwrite(dset, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, d.data) < 0) {
      err = "could not write dataset " + name;
      ok = false;
    }
    if (dset >= 0) a.H5Dclose(dset);
    if (dcpl != H5P_DEFAULT && dcpl >= 0) a.H5Pclose(dcpl);
    if (space >= 0) a.H5Sclose(space);
  }
  if (a.H5Fclose(f) < 0 && ok) {
    err = std::string("could not close ") + path;
    ok = false;
  }
  return ok;
}

H5In::H5In(const char* path, std::string& err) : err_(err) {
  if (!have(err)) return;
  QuietErrors quiet;
  fid_ = api().H5Fopen(path, ACC_RDONLY, H5P_DEFAULT);
  if (fid_ < 0) err = std::string("could not open ") + path;
}

H5In::~H5In() {
  QuietErrors quiet;
  if (fid_ >= 0) api().H5Fclose(fid_);
}

int H5In::ndims(const char* name, long* dims) const {
  if (fid_ < 0) return -1;
  QuietErrors quiet;
  const H5Api& a = api();
  if (a.H5Lexists(fid_, name, H5P_DEFAULT) <= 0) return -1;
  const hid_t d = a.H5Dopen2(fid_, name, H5P_DEFAULT);
  if (d < 0) return -1;
  const hid_t s = a.H5Dget_space(d);
  int r = s >= 0 ? a.H5Sget_simple_extent_ndims(s) : -1;
  if (r > 8) r = -1;
  if (r >= 0) {
    hsize_t h[8];
    a.H5Sget_simple_extent_dims(s, h, nullptr);
    for (int k = 0; k < r; ++k) dims[k] = (long)h[k];
  }
  if (s >= 0) a.H5Sclose(s);
  a.H5Dclose(d);
  return r;
}

bool H5In::read(const char* name, bool ints, void* dst, int mrank, const long* mdims,
                const long* start) const {
  if (fid_ < 0) return false;
  QuietErrors quiet;
  const H5Api& a = api();
  long fd[8];
  const int fr = ndims(name, fd);
  if (fr < 0) { err_ = std::string("dataset ") + name + " not found"; return false; }
  const hid_t d = a.H5Dopen2(fid_, name, H5P_DEFAULT);
  hid_t ms = H5S_ALL;
  bool ok = d >= 0;
  if (ok && mdims) {
    // memory space mdims (C order), file extent selected at start
    hsize_t md[8], st[8], ct[8];
    int off = mrank - fr;   // a file dataset of lower rank fills the trailing memory dimensions
    if (off < 0) { err_ = std::string("dataset ") + name + " has too many dimensions"; ok = false; }
    for (int k = 0; ok && k < mrank; ++k) {
      md[k] = (hsize_t)mdims[k];
      st[k] = (hsize_t)start[k];
      ct[k] = k < off ? 1 : (hsize_t)fd[k - off];
      if (st[k] + ct[k] > md[k]) {
        err_ = std::string("dataset ") + name + " does not fit its destination";
        ok = false;
      }
    }
    if (ok) {
      ms = a.H5Screate_simple(mrank, md, nullptr);
      ok = ms >= 0 && a.H5Sselect_hyperslab(ms, H5S_SELECT_SET, st, nullptr, ct, nullptr) >= 0;
    }
  }
  const hid_t mt = ints ? *a.native_int : *a.native_double;
  if (ok && a.H5Dread(d, mt, ms, H5S_ALL, H5P_DEFAULT, dst) < 0) {
    err_ = std::string("could not read dataset ") + name;
    ok = false;
  }
  if (ms != H5S_ALL && ms >= 0) a.H5Sclose(ms);
  if (d >= 0) a.H5Dclose(d);
  return ok;
}

bool h5_available(std::string& err) { return have(err); }

}  // namespace gemhip
