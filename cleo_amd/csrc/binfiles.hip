// binfiles.hip -- CLEO's binary input files: the metadata layout of initialise/readbinary.cpp:22-99,
// superdroplet initial conditions (init_all_supers_from_binary.cpp:57-100), gridbox boundaries
// (gbx_bounds_from_binary.cpp:29-58) and from-file thermodynamics (fromfile_cartesian_dynamics.cpp:
// 22-31).
// Host code only: these calls need neither a context nor a GPU.
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "../../include/cleo_hip.h"

namespace {

// metadata of one variable: first byte, bytes per datapoint, datapoints, type, units, scale
struct VarMeta {
  uint32_t b0, bsize, nvar;
  char vtype, units;
  double scale;
};

// The file's variables from its metadata (readbinary.cpp:22-99: GblMetadata then one VarMetadata
// per variable).  False when the layout does not fit inside the file.
bool read_metadata(std::ifstream& f, uint64_t fsize, std::vector<VarMeta>& out) {
  uint32_t u[4];
  f.seekg(0, std::ios::beg);
  f.read(reinterpret_cast<char*>(u), sizeof(u));
  if (!f) return false;
  const uint64_t charbytes = u[1], nvars = u[2], mbytes = u[3];
  constexpr uint64_t kVarMetaBytes = 3 * sizeof(uint32_t) + 2 + sizeof(double);
  if (mbytes < kVarMetaBytes || sizeof(u) + charbytes + nvars * mbytes > fsize) return false;
  uint64_t pos = sizeof(u) + charbytes;
  out.clear();
  for (uint64_t i = 0; i < nvars; ++i, pos += mbytes) {
    f.seekg(static_cast<std::streamoff>(pos), std::ios::beg);
    uint32_t v[3];
    char c[2];
    double sf;
    f.read(reinterpret_cast<char*>(v), sizeof(v));
    f.read(c, 2);
    f.read(reinterpret_cast<char*>(&sf), sizeof(sf));
    if (!f) return false;
    const VarMeta m{v[0], v[1], v[2], c[0], c[1], sf};
    if (static_cast<uint64_t>(m.b0) + static_cast<uint64_t>(m.bsize) * m.nvar > fsize) return false;
    out.push_back(m);
  }
  return true;
}

// the datapoints of variable m into dst (vector_from_binary, readbinary.hpp:85-92)
template <class T>
bool read_var(std::ifstream& f, const VarMeta& m, T* dst) {
  if (m.bsize != sizeof(T)) return false;
  if (m.nvar == 0) return true;
  f.seekg(m.b0, std::ios::beg);
  f.read(reinterpret_cast<char*>(dst), static_cast<std::streamsize>(m.nvar) * sizeof(T));
  return static_cast<bool>(f);
}

bool open_file(const char* path, std::ifstream& f, uint64_t& fsize) {
  f.open(path, std::ios::in | std::ios::binary);
  if (!f.is_open()) return false;
  f.seekg(0, std::ios::end);
  fsize = static_cast<uint64_t>(f.tellg());
  return true;
}

}  // namespace

int cleo_read_initsupers(const char* path, cleo_supers_soa* out, uint64_t cap, uint64_t* n) {
  if (!path || !n) return CLEO_E_ARG;
  std::ifstream f;
  uint64_t fsize = 0;
  if (!open_file(path, f, fsize)) return CLEO_E_FILE;
  std::vector<VarMeta> meta;
  if (!read_metadata(f, fsize, meta) || meta.size() < 7) return CLEO_E_FILE;
  const uint64_t nd = meta[0].nvar;
  // xi, radius and msol have one datapoint per drop; a coordinate has one per drop or none
  for (int v = 1; v < 4; ++v)
    if (meta[v].nvar != nd) return CLEO_E_FILE;
  for (int v = 4; v < 7; ++v)
    if (meta[v].nvar != nd && meta[v].nvar != 0) return CLEO_E_FILE;
  *n = nd;
  if (!out) return CLEO_OK;
  if (nd > cap) return CLEO_E_ARG;
  if (!out->sdgbxindex || !out->xi || !out->radius || !out->msol || !out->coord3 || !out->coord1 || !out->coord2 ||
      !out->sdid)
    return CLEO_E_ARG;
  double* coords[3] = {out->coord3, out->coord1, out->coord2};
  const bool ok = read_var(f, meta[0], out->sdgbxindex) && read_var(f, meta[1], out->xi) &&
                  read_var(f, meta[2], out->radius) && read_var(f, meta[3], out->msol) &&
                  read_var(f, meta[4], coords[0]) && read_var(f, meta[5], coords[1]) &&
                  read_var(f, meta[6], coords[2]);
  if (!ok) return CLEO_E_FILE;
  for (int a = 0; a < 3; ++a)
    if (meta[4 + a].nvar == 0) std::memset(coords[a], 0, nd * sizeof(double));
  for (uint64_t k = 0; k < nd; ++k) out->sdid[k] = k;
  return CLEO_OK;
}

int cleo_read_initsupers_padded(const char* path, uint64_t maxnsupers, cleo_supers_soa* out, uint64_t* ninit) {
  if (!path || !out || !ninit) return CLEO_E_ARG;
  uint64_t n = 0;
  int rc = cleo_read_initsupers(path, nullptr, 0, &n);
  if (rc != CLEO_OK) return rc;
  // InitSupersFromBinary: maxnsupers < initnsupers is an error (init_supers_from_binary.hpp:75-79)
  if (n > maxnsupers) return CLEO_E_ARG;
  rc = cleo_read_initsupers(path, out, maxnsupers, &n);
  if (rc != CLEO_OK) return rc;
  // add_uninitialised_superdrops_data (init_supers_from_binary.cpp:44-63): out of the domain, NaN
  // attributes and coordinates, and the values signaling_NaN gives integers (0) for xi and sdId
  const double nan = std::numeric_limits<double>::signaling_NaN();
  for (uint64_t k = n; k < maxnsupers; ++k) {
    out->sdgbxindex[k] = CLEO_OOB_GBXINDEX;
    out->coord3[k] = out->coord1[k] = out->coord2[k] = nan;
    out->radius[k] = out->msol[k] = nan;
    out->xi[k] = 0;
    out->sdid[k] = 0;
  }
  *ninit = n;
  return CLEO_OK;
}

int cleo_read_gridfile(const char* path, uint32_t ndims[3], double* zhalf, double* xhalf, double* yhalf) {
  if (!path || !ndims) return CLEO_E_ARG;
  std::ifstream f;
  uint64_t fsize = 0;
  if (!open_file(path, f, fsize)) return CLEO_E_FILE;
  std::vector<VarMeta> meta;
  if (!read_metadata(f, fsize, meta) || meta.size() < 3 || meta[0].nvar != 3) return CLEO_E_FILE;
  uint64_t nd[3];
  if (!read_var(f, meta[0], nd)) return CLEO_E_FILE;
  for (int a = 0; a < 3; ++a)
    if (nd[a] == 0 || nd[a] > 0xFFFFFFFFull) return CLEO_E_FILE;
  // a product that wraps, or more gridboxes than the file can hold bounds for, is not a layout
  uint64_t nzx = 0, ngbx = 0;
  if (__builtin_mul_overflow(nd[0], nd[1], &nzx) || __builtin_mul_overflow(nzx, nd[2], &ngbx) ||
      ngbx > fsize / 48)
    return CLEO_E_FILE;
  // gbx_bounds_from_binary.cpp:38-44: one index and six bounds per gridbox
  if (meta[1].nvar != ngbx || meta[2].nvar != 6 * ngbx) return CLEO_E_FILE;
  std::vector<uint32_t> idx(ngbx);
  std::vector<double> b(6 * ngbx);
  if (!read_var(f, meta[1], idx.data()) || !read_var(f, meta[2], b.data())) return CLEO_E_FILE;
  for (int a = 0; a < 3; ++a) ndims[a] = static_cast<uint32_t>(nd[a]);
  if (!zhalf || !xhalf || !yhalf) return CLEO_OK;
  const uint64_t nz = nd[0], nx = nd[1], ny = nd[2];
  // faces from the gridboxes along each axis from the origin, then every gridbox checked against them
  for (uint64_t k = 0; k < nz; ++k) zhalf[k] = b[6 * k];
  zhalf[nz] = b[6 * (nz - 1) + 1];
  for (uint64_t i = 0; i < nx; ++i) xhalf[i] = b[6 * (nz * i) + 2];
  xhalf[nx] = b[6 * (nz * (nx - 1)) + 3];
  for (uint64_t j = 0; j < ny; ++j) yhalf[j] = b[6 * (nz * nx * j) + 4];
  yhalf[ny] = b[6 * (nz * nx * (ny - 1)) + 5];
  for (uint64_t g = 0; g < ngbx; ++g) {
    const uint64_t k = g % nz, i = (g / nz) % nx, j = g / (nz * nx);
    const double* q = &b[6 * g];
    if (idx[g] != g || q[0] != zhalf[k] || q[1] != zhalf[k + 1] || q[2] != xhalf[i] || q[3] != xhalf[i + 1] ||
        q[4] != yhalf[j] || q[5] != yhalf[j + 1])
      return CLEO_E_FILE;
  }
  return CLEO_OK;
}

int cleo_read_thermofile(const char* path, double* out, uint64_t cap, uint64_t* n) {
  if (!path || !n) return CLEO_E_ARG;
  std::ifstream f;
  uint64_t fsize = 0;
  if (!open_file(path, f, fsize)) return CLEO_E_FILE;
  std::vector<VarMeta> meta;
  if (!read_metadata(f, fsize, meta) || meta.empty()) return CLEO_E_FILE;
  *n = meta[0].nvar;
  if (!out) return meta[0].bsize == sizeof(double) ? CLEO_OK : CLEO_E_FILE;
  if (*n > cap) return CLEO_E_ARG;
  return read_var(f, meta[0], out) ? CLEO_OK : CLEO_E_FILE;
}
