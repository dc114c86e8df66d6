// Test program for the library's HDF5 writer (gemini_amd/csrc/h5io.cpp, compiled in): writes
// a scalar, a 3-D and a 4-D real32 dataset whose chunks end partly outside the dataset, and
// a 1x1 one, at the deflate level given, then the raw values next to the file.
//   h5io_check <file.h5> <level>     (GEMHIP_IO_THREADS sets the compression threads)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../gemini_amd/csrc/gemhip_io.h"

int main(int argc, char** argv) {
  if (argc != 3) { fprintf(stderr, "usage: h5io_check file.h5 level\n"); return 2; }
  const int level = atoi(argv[2]);
  std::vector<float> a(3L * 700 * 600), b(5L * 3 * 450 * 530), c(1, 3.5f);
  for (size_t i = 0; i < a.size(); ++i) a[i] = (float)(i % 977) * 0.5f - (float)(i % 13);
  for (size_t i = 0; i < b.size(); ++i) b[i] = (float)((i * 2654435761u) % 100003) * 1e-3f;
  int iv = 7;
  std::vector<gemhip::H5Out> ds = {
      {"/i", gemhip::H5Out::I32, 0, {0}, &iv, nullptr},
      {"/a", gemhip::H5Out::F32, 3, {3, 700, 600}, a.data(), nullptr},
      {"/g/b", gemhip::H5Out::F32, 4, {5, 3, 450, 530}, b.data(), nullptr},
      {"/c", gemhip::H5Out::F32, 2, {1, 1}, c.data(), nullptr}};
  std::string err;
  if (!gemhip::h5_write(argv[1], ds.data(), (int)ds.size(), level, err)) {
    fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  FILE* f = fopen((std::string(argv[1]) + ".raw").c_str(), "wb");
  if (!f) return 1;
  fwrite(a.data(), 4, a.size(), f);
  fwrite(b.data(), 4, b.size(), f);
  fclose(f);
  return 0;
}
