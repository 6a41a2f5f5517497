// Golovin box model through the C++ host layer, the GPU counterpart of the reference's
// examples/boxmodelcollisions/src/main_golcolls.cpp: one gridbox of 1e6 m^3, superdroplets with
// radii sampled from an exponential-in-volume distribution (mean radius 30.531 um), equal
// multiplicities, collision-coalescence with Golovin's kernel, dt = 1 s.
// Prints "t[s] total_xi in_domain_supers" every `every` seconds.
// Usage: example_golcolls [nsupers=8192] [t_end_s=3600] [every_s=600] [seed=1]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "cleo_sdm.hpp"

int main(int argc, char** argv) {
  namespace cg = cleo_gpu;
  const uint64_t nsupers = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 8192;
  const unsigned t_end = argc > 2 ? std::atoi(argv[2]) : 3600;
  const unsigned every = argc > 3 ? std::atoi(argv[3]) : 600;
  const uint64_t seed = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 1;
  const double side = 100.0 / 1000.0;  // 100 m in COORD0 units
  double zh[2] = {0.0, side}, xh[2] = {0.0, side}, yh[2] = {0.0, side};
  cleo_domain_desc dom{{1, 1, 1}, zh, xh, yh, {0, 1, 1}};
  cg::GpuDomain domain(nsupers, dom);

  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  const double r0 = 30.531e-6, vol0 = 4.0 / 3.0 * M_PI * r0 * r0 * r0;
  std::exponential_distribution<double> expvol(1.0 / vol0);
  const uint64_t xi = static_cast<uint64_t>(std::llround(std::pow(2.0, 23) * 1e6 / nsupers));
  std::vector<cg::Superdrop> supers(nsupers);
  for (uint64_t k = 0; k < nsupers; ++k) {
    cg::Superdrop& d = supers[k];
    d = cg::Superdrop{};
    d.sdgbxindex = 0;
    d.coord3 = u01(gen) * side;
    d.coord1 = u01(gen) * side;
    d.coord2 = u01(gen) * side;
    d.attrs.xi = xi;
    d.attrs.radius = std::cbrt(3.0 * expvol(gen) / (4.0 * M_PI)) / 1e-6;
    d.attrs.msol = 0.0;
    d.sdId = k;
  }
  domain.upload(supers);
  std::vector<double> st(11, 0.0);
  st[0] = 1e6 / 1e9;  // volume / VOL0
  st[1] = 1.0;
  st[2] = 288.15 / 273.15;
  cleo_gbx_state s{&st[0], &st[1], &st[2], &st[3], &st[4], &st[5], &st[6], &st[7], &st[8], &st[9], &st[10]};
  domain.set_state(s);
  domain.set_seed(seed);

  const unsigned collstep = cg::realtime2step(1.0);
  auto colls = cg::CollCoal(collstep, cg::step2realtime, cg::GolovinProb{});
  cg::SDMMethods sdm(collstep, domain, colls, cg::NullMotion{});

  auto report = [&](unsigned t) {
    uint64_t nd = 0;
    const auto out = domain.download(&nd);
    double tot = 0.0;
    for (uint64_t k = 0; k < nd; ++k) tot += static_cast<double>(out[k].attrs.xi);
    std::printf("%u %.17g %llu\n", t, tot, static_cast<unsigned long long>(nd));
  };
  report(0);
  for (unsigned t = 0; t < t_end; t += every) {
    const unsigned n = std::min(every, t_end - t);
    sdm.run_steps(t * collstep, n);
    report(t + n);
  }
  return 0;
}
