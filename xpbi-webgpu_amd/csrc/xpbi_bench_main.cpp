// Native driver of the substep over the C ABI alone (no Python): a DP sand
// column of nx x ny x nz particles at spacing r on the floor of its domain
// (default 80 x 160 x 80 at r = 0.005 in [0,1.5]x[0,1]x[0,1.5]: the 1,024,000
// particle column; E 3.5e5 Pa, nu 0.3, rho 2000, phi 30 deg, dt 1/6000), for
// profilers that count hardware events per dispatch.
//
//   xpbi_bench [--settle S] [--steps K] [--iterations I] [--copy-mib M]
//
// Runs S substeps, then K timed substeps, then one copy of M MiB through the
// library's copy kernel (read M MiB, write M MiB: the calibration of byte
// counters).  The lattice jitter comes from splitmix64, not from the scene
// module's counter hash: the same shape, not the same particles.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "xpbi.h"

static uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static int check(int rc, xpbi_sim* sim, const char* what) {
  if (rc != XPBI_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, xpbi_last_error(sim));
    std::exit(1);
  }
  return rc;
}

int main(int argc, char** argv) {
  int settle = 300, steps = 5, iters = 20, copyMiB = 256;
  int nx = 80, ny = 160, nz = 80;
  double r = 0.005, dom[3] = {1.5, 1.0, 1.5};
  for (int k = 1; k + 1 < argc; k += 2) {
    const std::string a = argv[k];
    const int v = std::atoi(argv[k + 1]);
    if (a == "--settle") settle = v;
    else if (a == "--steps") steps = v;
    else if (a == "--iterations") iters = v;
    else if (a == "--copy-mib") copyMiB = v;
    else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return 2;
    }
  }
  const double h = r;
  const uint32_t N = (uint32_t)nx * ny * nz;
  std::vector<float> pos(3ull * N);
  std::vector<uint32_t> mat(N, 0u);
  uint64_t seed = 1;
  const double x0 = 0.5 * (dom[0] - nx * r), z0 = 0.5 * (dom[2] - nz * r);
  size_t q = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const double c[3] = {x0 + (i + 0.5) * r, (j + 0.5) * r, z0 + (k + 0.5) * r};
        for (int d = 0; d < 3; d++) {
          const double u = (double)(splitmix(seed) >> 11) * (1.0 / 9007199254740992.0);
          pos[q++] = (float)(c[d] + (u - 0.5) * 0.1 * r);
        }
      }
  xpbi_scene_desc desc{};
  desc.N = N;
  desc.positions = pos.data();
  desc.matIds = mat.data();
  desc.h = h;
  for (int d = 0; d < 3; d++) {
    desc.domainMin[d] = 0.0;
    desc.domainMax[d] = dom[d];
  }
  desc.device = 0;
  xpbi_sim* sim = nullptr;
  check(xpbi_create(&desc, &sim), nullptr, "xpbi_create");

  // Drucker-Prager sand (the writeMaterials packing: mu, lambda, mass from E, nu, rho)
  const double E = 3.5e5, nu = 0.3, rho = 2000.0, phi = 30.0 * M_PI / 180.0;
  const double sphi = std::sin(phi);
  xpbi_material m{};
  m.mu = (float)(E / (2 * (1 + nu)));
  m.lam = (float)(E * nu / ((1 + nu) * (1 - 2 * nu)));
  m.density = (float)rho;
  m.kind = 1;
  m.p0 = (float)(std::sqrt(2.0 / 3.0) * 2 * sphi / (3 - sphi));
  m.boundFriction = 0.4f;
  m.mass = (float)(rho * h * h * h);
  check(xpbi_write_materials(sim, &m, 1), sim, "xpbi_write_materials");

  xpbi_params p{};
  uint32_t numCells = 0, n = 0;
  check(xpbi_grid_info(sim, p.gridOrigin, &p.cellSize, p.gridDims, &numCells, &n), sim, "xpbi_grid_info");
  p.numCells = numCells;
  p.N = N;
  p.dt = (float)(1.0 / 6000.0);
  p.h = (float)h;
  p.support = (float)(2 * h);
  p.V0 = (float)(h * h * h);
  p.gravity[1] = -9.81f;
  for (int d = 0; d < 3; d++) p.domainMax[d] = (float)dom[d];
  p.xsphC = 0.01f;
  p.omega = 0.5f;
  p.pbfEps = 100.0f;
  p.sCorrK = 1e-4f;
  p.contactOmega = 0.4f;
  p.kDamp = 0.4f;
  check(xpbi_write_params(sim, &p), sim, "xpbi_write_params");

  check(xpbi_step(sim, settle, iters, 0), sim, "xpbi_step");
  float ms = 0.0f;
  check(xpbi_time_substeps(sim, steps, iters, 0, &ms), sim, "xpbi_time_substeps");
  float cms = 0.0f;
  if (copyMiB > 0) check(xpbi_copy_bandwidth(sim, (size_t)copyMiB << 20, 1, &cms), sim, "xpbi_copy_bandwidth");
  // FNV-1a over the final positions and velocities: equal hashes from two builds
  // mean bit-identical states
  uint64_t hash = 1469598103934665603ull;
  std::vector<uint8_t> buf(16ull * N);
  for (int field : {XPBI_FIELD_POS, XPBI_FIELD_VEL}) {
    check(xpbi_read(sim, field, buf.data(), buf.size()), sim, "xpbi_read");
    for (uint8_t byte : buf) hash = (hash ^ byte) * 1099511628211ull;
  }
  std::printf("{\"particles\": %u, \"settle\": %d, \"steps\": %d, \"iterations\": %d, \"ms_per_substep\": %.4f, "
              "\"copy_mib\": %d, \"copy_ms\": %.4f, \"state_fnv1a\": \"%016llx\"}\n",
              N, settle, steps, iters, ms / (steps > 0 ? steps : 1), copyMiB, cms, (unsigned long long)hash);
  xpbi_destroy(sim);
  return 0;
}
