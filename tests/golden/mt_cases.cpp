// Regenerates the random test matrices of the reference's backend tests
// (tests/unit/test_gpu_backend.cpp:26-49 make_problem, :100-104 and :263-267
// B fills) with the same std::mt19937 recipe, so the fixtures hold the exact
// inputs those tests use.  Output: raw little-endian arrays on stdout.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

static void put(const void* p, size_t n) { fwrite(p, 1, n, stdout); }

int main(int argc, char** argv) {
  if (argc < 5) return 1;
  const size_t nrows = std::stoul(argv[1]), ncols = std::stoul(argv[2]);
  const double density = std::stod(argv[3]);
  const uint32_t seed = (uint32_t)std::stoul(argv[4]);
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> val(-4.0, 4.0);
  std::vector<double> values;
  std::vector<int64_t> cols, rp{0};
  for (size_t i = 0; i < nrows; ++i) {
    for (size_t j = 0; j < ncols; ++j) {
      if (std::uniform_real_distribution<double>(0.0, 1.0)(rng) < density) {
        values.push_back(val(rng));
        cols.push_back((int64_t)j);
      }
    }
    rp.push_back((int64_t)values.size());
  }
  std::vector<double> x(ncols);
  for (auto& v : x) v = val(rng);
  int64_t nnz = (int64_t)values.size();
  put(&nnz, 8);
  put(rp.data(), 8 * rp.size());
  put(cols.data(), 8 * cols.size());
  put(values.data(), 8 * values.size());
  put(x.data(), 8 * x.size());
  if (argc >= 8) {  // B matrix: k, seed, lo (range [lo, -lo])
    const size_t k = std::stoul(argv[5]);
    std::mt19937 rb((uint32_t)std::stoul(argv[6]));
    const double lo = std::stod(argv[7]);
    std::uniform_real_distribution<double> bv(lo, -lo);
    std::vector<double> B(ncols * k);
    for (auto& v : B) v = bv(rb);
    put(B.data(), 8 * B.size());
  }
  return 0;
}
