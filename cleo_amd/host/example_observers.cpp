// Observers writing CLEO's Zarr dataset from the GPU domain, through the C++ host layer: a small
// coupled domain (nz x nx x ny gridboxes of 25 m, drops given in a file of 72-byte Superdrops, the
// State in a file of 11 doubles per gridbox), Condensation >> CoalBuRe(Long) with motion (Simmel),
// one step of 1 s per observation, the time, gbxindex, mass moments (all drops and raindrops) and
// superdroplets observed at every step.
// Usage: example_observers <drops.bin> <state.bin> <nz> <nx> <ny> <nsteps> <seed> <zarr dir>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "cleo_zarr.hpp"

template <typename T>
static std::vector<T> read_all(const char* path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) {
    std::fprintf(stderr, "cannot read %s\n", path);
    std::exit(2);
  }
  const auto n = static_cast<size_t>(f.tellg()) / sizeof(T);
  std::vector<T> v(n);
  f.seekg(0);
  f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
  return v;
}

int main(int argc, char** argv) {
  namespace cg = cleo_gpu;
  if (argc < 9) {
    std::fprintf(stderr, "usage: example_observers drops.bin state.bin nz nx ny nsteps seed zarrdir\n");
    return 2;
  }
  const auto drops = read_all<cg::Superdrop>(argv[1]);
  auto st = read_all<double>(argv[2]);
  const uint32_t nz = std::atoi(argv[3]), nx = std::atoi(argv[4]), ny = std::atoi(argv[5]);
  const unsigned nsteps = std::atoi(argv[6]);
  const uint64_t seed = std::strtoull(argv[7], nullptr, 10);
  const uint64_t ngbx = static_cast<uint64_t>(nz) * nx * ny;
  if (st.size() != 11 * ngbx) {
    std::fprintf(stderr, "state file holds %zu doubles, expected %llu\n", st.size(),
                 static_cast<unsigned long long>(11 * ngbx));
    return 2;
  }
  const double dx = 25.0 / 1000.0;
  std::vector<double> zh(nz + 1), xh(nx + 1), yh(ny + 1);
  for (uint32_t i = 0; i <= nz; ++i) zh[i] = i * dx;
  for (uint32_t i = 0; i <= nx; ++i) xh[i] = i * dx;
  for (uint32_t i = 0; i <= ny; ++i) yh[i] = i * dx;
  cleo_domain_desc dom{{nz, nx, ny}, zh.data(), xh.data(), yh.data(), {0, 1, 1}};
  cg::GpuDomain domain(drops.size(), dom);
  domain.upload(drops);
  double* s = st.data();
  cleo_gbx_state gs{s, s + ngbx, s + 2 * ngbx, s + 3 * ngbx, s + 4 * ngbx, s + 5 * ngbx,
                    s + 6 * ngbx, s + 7 * ngbx, s + 8 * ngbx, s + 9 * ngbx, s + 10 * ngbx};
  domain.set_state(gs);
  const cleo_solute amsul{1770.0 / (1e5 / (1004.64 * 273.15)), 0.13214 / 0.028966216, 3.0};
  cg::check(domain.ctx(), cleo_set_solute(domain.ctx(), &amsul), "solute");
  domain.set_seed(seed);

  const unsigned step = cg::realtime2step(1.0);
  const auto cond = cg::Condensation(step, cg::step2dimlesstime, true, 50, 1e-6, 1e-8, 0.001, cg::realtime2dimless);
  const auto colls = cg::CoalBuRe(step, cg::step2realtime, cg::LongHydroProb{}, cg::ConstNFrags(5.0),
                                  cg::ConstCoalBuReFlag(0.85));
  const auto motion = cg::CartesianMotion(step, cg::step2dimlesstime, cg::SimmelTerminalVelocity{});
  cg::SDMMethods sdm(step, domain, cond >> colls, motion);

  {
    const size_t maxchunk = 4096;
    cg::ZarrStore store(argv[8]);
    cg::write_gbxindex(store, ngbx);
    cg::TimeObserver time(step, store, maxchunk);
    cg::MassMomentsObserver moms(step, store, maxchunk, ngbx);
    cg::MassMomentsObserver rain(step, store, maxchunk, ngbx, true);
    cg::SuperdropsObserver supers(step, store, maxchunk);
    auto observe = [&](unsigned t) {
      time.observe(domain, t);
      moms.observe(domain, t);
      rain.observe(domain, t);
      supers.observe(domain, t);
    };
    observe(0);
    for (unsigned k = 0; k < nsteps; ++k) {
      sdm.run_step(k * step, (k + 1) * step);
      observe((k + 1) * step);
    }
  }
  std::printf("ok %u observations\n", nsteps + 1);
  return 0;
}
