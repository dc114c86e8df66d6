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

#include <cstdlib>
#include <cstring>
#include <set>
#include <string>

namespace gemhip {
namespace {

// Function table of the HDF5 calls used here, resolved by name.  The header declares the
// functions; decltype keeps every pointer's type that of the declaration.
struct H5Api {
  bool ok = false;
  std::string why;
#define H5FN(name) decltype(&::name) name = nullptr;
  H5FN(H5open)
  H5FN(H5get_libversion)
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
    t.H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // errors are reported by return codes
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

// chunk shape for a compressed dataset: whole trailing (fastest) planes up to 1 MiB,
// leading dimensions cut to 1 first
void chunk_dims(int rank, const long* dims, size_t elem, hsize_t* ch) {
  for (int i = 0; i < rank; ++i) ch[i] = (hsize_t)(dims[i] > 0 ? dims[i] : 1);
  const size_t cap = 1 << 20;
  auto bytes = [&] {
    size_t b = elem;
    for (int i = 0; i < rank; ++i) b *= ch[i];
    return b;
  };
  for (int i = 0; i < rank - 1 && bytes() > cap; ++i) ch[i] = 1;
  if (bytes() > cap && rank >= 2) {
    const size_t row = elem * ch[rank - 1];
    ch[rank - 2] = (hsize_t)(cap / row > 0 ? cap / row : 1);
  }
}

// the access flags of H5Fpublic.h without the macros' library calls (H5check, H5open)
constexpr unsigned ACC_RDONLY = 0x0000u, ACC_TRUNC = 0x0002u;

}  // namespace

bool h5_write(const char* path, const H5Out* ds, int nds, int comp_lvl, std::string& err) {
  if (!have(err)) return false;
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
    if (ok && a.H5Dwrite(dset, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, d.data) < 0) {
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
  fid_ = api().H5Fopen(path, ACC_RDONLY, H5P_DEFAULT);
  if (fid_ < 0) err = std::string("could not open ") + path;
}

H5In::~H5In() {
  if (fid_ >= 0) api().H5Fclose(fid_);
}

int H5In::ndims(const char* name, long* dims) const {
  if (fid_ < 0) return -1;
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
