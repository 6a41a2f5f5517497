// Writes known arrays through cleo_amd/host/cleo_zarr.hpp for tests/test_zarr.py, which reads the
// store back with numpy and json: a 2-D (time, gbxindex) array over several chunks with a partial
// last chunk, a 1-D ragged array and its count, a float array with a scale factor, and an array
// flushed while partly filled and then appended to.
// Usage: zarr_writer_check <store dir>
#include <cstdio>
#include <vector>

#include "cleo_zarr.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  namespace cg = cleo_gpu;
  cg::ZarrStore store(argv[1]);
  {
    // 5 rows of 3 gridboxes, chunks of 2 rows: chunks 0.0, 1.0 whole, 2.0 half filled
    cg::ZarrArray<uint64_t> a(store, "mom0", 2, {3}, {"time", "gbxindex"}, "", 1.0);
    for (uint64_t t = 0; t < 5; ++t) {
      const uint64_t row[3] = {10 * t, 10 * t + 1, 10 * t + 2};
      a.append(row, 3);
    }
    // ragged: 4, 0, 7 drops over three observations, chunks of 5
    cg::ZarrArray<uint32_t> count(store, "raggedcount", 5, {}, {"time"}, "", 1.0, "superdroplets");
    cg::ZarrArray<float> r(store, "radius", 5, {}, {"superdroplets"}, "micro-m", 1.0, "superdroplets");
    for (uint32_t n : {4u, 0u, 7u}) {
      count.append(&n, 1);
      std::vector<float> v(n);
      for (uint32_t k = 0; k < n; ++k) v[k] = 0.5f * static_cast<float>(k) + static_cast<float>(n);
      if (n) r.append(v.data(), n);
    }
    cg::ZarrArray<float> m(store, "massmom1", 4, {2}, {"time", "gbxindex"}, "g", cg::kMASS0grams);
    const float x[2] = {1.5f, -2.25f};
    m.append(x, 2);
    m.flush();  // a partial chunk on disk; the next row fills the same chunk
    const float y[2] = {3.0f, 4.0f};
    m.append(y, 2);
  }
  std::printf("ok\n");
  return 0;
}
