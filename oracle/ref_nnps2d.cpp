// Driver of the reference's own 2D neighbour search (bench/nnps_c++/nnps.C
// of pankajp/pysph, compiled from where it lies by oracle/Makefile into
// oracle/_ref/).  TEST INFRASTRUCTURE ONLY: tests/ use it to check the
// oracle's neighbour sets against code of the reference.
//
// Input file: int64 n, double radius, n doubles x, n doubles y.
// Output file: for every point as a query, int64 count then count int64
// indices of the points within `radius` (the reference's strict test,
// nnps.C get_nearest_particles), bins of size radius.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "nnps.H"

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s input output\n", argv[0]);
    return 2;
  }
  std::FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 3;
  int64_t n = 0;
  double radius = 0.0;
  if (std::fread(&n, 8, 1, in) != 1 || std::fread(&radius, 8, 1, in) != 1 || n < 0) return 4;
  std::vector<double> x(n), y(n);
  if ((int64_t)std::fread(x.data(), 8, n, in) != n || (int64_t)std::fread(y.data(), 8, n, in) != n)
    return 5;
  std::fclose(in);
  nnps::BinManager bins(&x, &y, radius, radius);
  bins.bin_particles();
  nnps::NNPS search(&bins);
  std::FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 6;
  std::vector<std::size_t> nb;
  for (int64_t i = 0; i < n; ++i) {
    search.get_nearest_particles(x[i], y[i], radius, &nb);
    const int64_t c = (int64_t)nb.size();
    std::fwrite(&c, 8, 1, out);
    for (std::size_t k = 0; k < nb.size(); ++k) {
      const int64_t v = (int64_t)nb[k];
      std::fwrite(&v, 8, 1, out);
    }
  }
  std::fclose(out);
  return 0;
}
