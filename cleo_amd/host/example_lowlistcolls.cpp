// Low-List collision box model through the C++ host layer, the GPU counterpart of the reference's
// examples/boxmodelcollisions/src/main_lowlistcolls.cpp with its breakup_config.yaml: one gridbox
// of 100 m x 100 m x 100 m, superdroplets with radii from an exponential-in-volume distribution
// (mean radius 30.531 um, kept within [0.1, 90] um), equal multiplicities for 2^23 droplets per m^3,
// dry radius 1e-16 m, and the microphysics coal >> bu:
//   CollCoal(LowListCoalProb(RogersGKTerminalVelocity)) >> CollBu(LowListBuProb(RogersGK), ConstNFrags(5))
// with a collision step of 1 s.
// Prints "t[s] total_xi in_domain_supers sum(xi r^3)" every `every` seconds.  Given two paths it
// also writes the initial and the final drops there as arrays of the reference's 72-byte Superdrop.
// Usage: example_lowlistcolls [nsupers=8192] [t_end_s=2400] [every_s=600] [seed=1] [init.bin final.bin]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "cleo_sdm.hpp"

static void dump(const char* path, const std::vector<cleo_gpu::Superdrop>& d) {
  FILE* f = std::fopen(path, "wb");
  if (!f || std::fwrite(d.data(), sizeof(cleo_gpu::Superdrop), d.size(), f) != d.size()) {
    std::fprintf(stderr, "cannot write %s\n", path);
    std::exit(2);
  }
  std::fclose(f);
}

int main(int argc, char** argv) {
  namespace cg = cleo_gpu;
  const uint64_t nsupers = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 8192;
  const unsigned t_end = argc > 2 ? std::atoi(argv[2]) : 2400;
  const unsigned every = argc > 3 ? std::atoi(argv[3]) : 600;
  const uint64_t seed = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 1;
  const char* init_path = argc > 6 ? argv[5] : nullptr;
  const char* final_path = argc > 6 ? argv[6] : nullptr;
  const double side = 100.0 / 1000.0;  // 100 m in COORD0 units
  double zh[2] = {0.0, side}, xh[2] = {0.0, side}, yh[2] = {0.0, side};
  cleo_domain_desc dom{{1, 1, 1}, zh, xh, yh, {0, 1, 1}};
  cg::GpuDomain domain(nsupers, dom);

  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  const double r0 = 30.531e-6, vol0 = 4.0 / 3.0 * M_PI * r0 * r0 * r0;
  std::exponential_distribution<double> expvol(1.0 / vol0);
  const uint64_t xi = static_cast<uint64_t>(std::llround(std::pow(2.0, 23) * 1e6 / nsupers));
  // solute of a 1e-16 m dry radius (NaCl), in MASS0 = R0^3 RHO0
  const double rho0 = 1e5 / (1004.64 * 273.15);
  const double msol = 4.0 / 3.0 * M_PI * std::pow(1e-16, 3) * 2016.5 / (1e-18 * rho0);
  std::vector<cg::Superdrop> supers(nsupers);
  for (uint64_t k = 0; k < nsupers; ++k) {
    cg::Superdrop& d = supers[k];
    d = cg::Superdrop{};
    d.sdgbxindex = 0;
    d.coord3 = u01(gen) * side;
    d.coord1 = u01(gen) * side;
    d.coord2 = u01(gen) * side;
    d.attrs.xi = xi;
    const double r = std::cbrt(3.0 * expvol(gen) / (4.0 * M_PI));
    d.attrs.radius = std::fmin(std::fmax(r, 1e-7), 9e-5) / 1e-6;
    d.attrs.msol = msol;
    d.sdId = k;
  }
  domain.upload(supers);
  if (init_path) dump(init_path, domain.download());
  std::vector<double> st(11, 0.0);
  st[0] = 1e6 / 1e9;  // volume / VOL0
  st[1] = 1.0;
  st[2] = 288.15 / 273.15;
  cleo_gbx_state s{&st[0], &st[1], &st[2], &st[3], &st[4], &st[5], &st[6], &st[7], &st[8], &st[9], &st[10]};
  domain.set_state(s);
  domain.set_seed(seed);

  const unsigned collstep = cg::realtime2step(1.0);
  const auto buprob = cg::LowListBuProb(cg::RogersGKTerminalVelocity{});
  const auto nfrags = cg::ConstNFrags(5.0);
  const auto bu = cg::CollBu(collstep, cg::step2realtime, buprob, nfrags);
  const auto coalprob = cg::LowListCoalProb(cg::RogersGKTerminalVelocity{});
  const auto coal = cg::CollCoal(collstep, cg::step2realtime, coalprob);
  cg::SDMMethods sdm(collstep, domain, coal >> bu, cg::NullMotion{});

  auto report = [&](unsigned t) {
    uint64_t nd = 0;
    const auto out = domain.download(&nd);
    double tot = 0.0, m3 = 0.0;
    for (uint64_t k = 0; k < nd; ++k) {
      tot += static_cast<double>(out[k].attrs.xi);
      m3 += static_cast<double>(out[k].attrs.xi) * std::pow(out[k].attrs.radius, 3);
    }
    std::printf("%u %.17g %llu %.17g\n", t, tot, static_cast<unsigned long long>(nd), m3);
  };
  report(0);
  for (unsigned t = 0; t < t_end; t += every) {
    const unsigned n = std::min(every, t_end - t);
    sdm.run_steps(t * collstep, n);
    report(t + n);
  }
  if (final_path) dump(final_path, domain.download());
  return 0;
}
