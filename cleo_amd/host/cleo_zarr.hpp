// cleo_zarr.hpp -- observers of the GPU domain that write CLEO's output dataset: a Zarr (version 2)
// group in a directory, readable by xarray / zarr as the reference's output is
// (libs/zarr/fsstore.hpp, zarr_group.hpp, zarr_array.hpp, xarray_zarr_array.hpp, simple_dataset.hpp;
// libs/observers/time_observer.hpp, gbxindex_observer.hpp, massmoments_observer.hpp,
// superdrops_observer.hpp).
//
// Layout, as the reference writes it:
//   .zgroup, .zattrs                 group metadata
//   <name>/.zarray                   {"shape", "chunks", "dtype", "order": "C", "compressor": null,
//                                     "fill_value": null, "filters": null, "zarr_format": 2}
//   <name>/.zattrs                   {"_ARRAY_DIMENSIONS": [...], "units": ..., ["scale_factor": ...],
//                                     ["sample_dimension": ...]} (scale_factor for floating types)
//   <name>/<i>[.<j>]                 raw little-endian chunks, uncompressed, every chunk whole; the
//                                     unfilled tail of the last chunk holds the type's largest value
// Arrays: time (f4, dimensionless time, scale TIME0 s), gbxindex (u4), massmom0 (u8), massmom1 (f4,
// g, scale MASS0 in g), massmom2 (f4, g^2) over (time, gbxindex) with chunks
// [maxchunk / ngbx, ngbx]; the superdroplets as contiguous ragged arrays over "superdroplets"
// (sdId u4, sdgbxindex u4, xi u8, radius f4 in micro-m, msol f4 in g, coord3/1/2 f4 in m) with
// raggedcount (u4 over time) the number of drops each observation added.
#pragma once

#include <cstdio>
#include <cstring>
#include <algorithm>
#include <filesystem>
#include <fstream>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "cleo_sdm.hpp"

namespace cleo_gpu {

// FSStore (fsstore.hpp): keys are paths below the store's directory
class ZarrStore {
 public:
  explicit ZarrStore(const std::filesystem::path& dir) : dir_(dir) {
    std::filesystem::create_directories(dir_);
    put(".zgroup", std::string("{\n  \"zarr_format\": 2\n}"));
    put(".zattrs", std::string("{\n  \"creator\": \"cleo_amd\",\n  \"title\": \"superdroplet model output on one "
                               "GPU, an Xarray and NetCDF compatible Zarr group of arrays\"\n}"));
  }
  void put(const std::string& key, const void* data, size_t nbytes) const {
    const auto path = dir_ / key;
    std::filesystem::create_directories(path.parent_path());
    std::ofstream out(path, std::ios::out | std::ios::binary | std::ios::trunc);
    out.write(static_cast<const char*>(data), static_cast<std::streamsize>(nbytes));
    if (!out.good()) throw std::runtime_error("cannot write " + path.string());
  }
  void put(const std::string& key, const std::string& text) const { put(key, text.data(), text.size()); }

 private:
  std::filesystem::path dir_;
};

template <typename T> struct ZarrType;
template <> struct ZarrType<uint32_t> { static constexpr const char* dtype = "<u4"; static constexpr bool fp = false; };
template <> struct ZarrType<uint64_t> { static constexpr const char* dtype = "<u8"; static constexpr bool fp = false; };
template <> struct ZarrType<float> { static constexpr const char* dtype = "<f4"; static constexpr bool fp = true; };
template <> struct ZarrType<double> { static constexpr const char* dtype = "<f8"; static constexpr bool fp = true; };

inline std::string json_list(const std::vector<size_t>& v) {
  std::string s = "[";
  for (size_t i = 0; i < v.size(); ++i) s += (i ? ", " : "") + std::to_string(v[i]);
  return s + "]";
}
inline std::string json_names(const std::vector<std::string>& v) {
  std::string s = "[";
  for (size_t i = 0; i < v.size(); ++i) s += (i ? ",\"" : "\"") + v[i] + "\"";
  return s + "]";
}

// An array that grows along its outermost dimension, written chunk by chunk (ZarrArray and
// XarrayZarrArray).  `inner` gives the extent of every other dimension; each chunk spans all of
// them (chunkshape = [chunk0, inner...]), so chunk k is named "k" or "k.0".
template <typename T>
class ZarrArray {
 public:
  ZarrArray(const ZarrStore& store, std::string name, size_t chunk0, std::vector<size_t> inner,
            std::vector<std::string> dims, const std::string& units, double scale_factor,
            const std::string& sample_dim = "")
      : store_(store), name_(std::move(name)), chunk0_(chunk0), inner_(std::move(inner)) {
    if (chunk0_ == 0 || dims.size() != inner_.size() + 1) throw std::invalid_argument("zarr array shape");
    row_ = 1;
    for (size_t v : inner_) row_ *= v;
    buf_.assign(chunk0_ * row_, std::numeric_limits<T>::max());
    std::string attrs = "{\n  \"_ARRAY_DIMENSIONS\": " + json_names(dims) + ",\n  \"units\": \"" + units + "\"";
    if (ZarrType<T>::fp) {
      char sf[64];
      std::snprintf(sf, sizeof(sf), "%.*e", std::numeric_limits<float>::digits10, scale_factor);
      attrs += ",\n  \"scale_factor\": " + std::string(sf);
    } else if (scale_factor != 1.0) {
      throw std::invalid_argument("scale_factor cannot be used on non-floating point type");
    }
    if (!sample_dim.empty()) attrs += ",\n  \"sample_dimension\": \"" + sample_dim + "\"";
    store_.put(name_ + "/.zattrs", attrs + "\n}");
    write_zarray();
  }
  ZarrArray(const ZarrArray&) = delete;
  ZarrArray& operator=(const ZarrArray&) = delete;
  ~ZarrArray() {
    try {
      flush();
    } catch (...) {
    }
  }
  // appends n values (whole rows of the inner dimensions)
  void append(const T* data, size_t n) {
    if (n % row_ != 0) throw std::invalid_argument("append whole rows");
    while (n > 0) {
      const size_t k = std::min(n, buf_.size() - fill_);
      std::memcpy(buf_.data() + fill_, data, k * sizeof(T));
      fill_ += k;
      data += k;
      n -= k;
      total_ += k;
      if (fill_ == buf_.size()) {  // a whole chunk: written, and the buffer starts the next one
        write_chunk();
        ++nchunks_;
        std::fill(buf_.begin(), buf_.end(), std::numeric_limits<T>::max());
        fill_ = 0;
      }
    }
    write_zarray();
  }
  // writes the partly filled chunk whole (its tail at the type's largest value); later data goes on
  // filling the same chunk
  void flush() {
    if (fill_ > 0) write_chunk();
    write_zarray();
  }
  size_t size() const { return total_; }

 private:
  void write_chunk() const {
    std::string key = name_ + "/" + std::to_string(nchunks_);
    for (size_t i = 0; i < inner_.size(); ++i) key += ".0";
    store_.put(key, buf_.data(), buf_.size() * sizeof(T));
  }
  void write_zarray() const {
    std::vector<size_t> shape{total_ / row_};
    shape.insert(shape.end(), inner_.begin(), inner_.end());
    std::vector<size_t> chunks{chunk0_};
    chunks.insert(chunks.end(), inner_.begin(), inner_.end());
    store_.put(name_ + "/.zarray", "{\n  \"shape\": " + json_list(shape) + ",\n  \"chunks\": " + json_list(chunks) +
                                       ",\n  \"dtype\": \"" + ZarrType<T>::dtype +
                                       "\",\n  \"order\": \"C\",\n  \"compressor\": null,\n  \"fill_value\": null,"
                                       "\n  \"filters\": null,\n  \"zarr_format\": 2\n}");
  }
  const ZarrStore& store_;
  std::string name_;
  size_t chunk0_, row_ = 1;
  std::vector<size_t> inner_;
  std::vector<T> buf_;
  size_t fill_ = 0, nchunks_ = 0, total_ = 0;
};

// ---- observers (ConstTstepObserver: observe on every multiple of interval) ------------------------
constexpr double kTIME0 = 1000.0;
constexpr double kR0 = 1e-6;
constexpr double kCOORD0 = 1000.0;
constexpr double kMASS0grams = 1e-18 * (1e5 / (1004.64 * 273.15)) * 1000.0;

// time (TimeObserver, time_observer.hpp:60-90) and gridbox index (GbxindexObserver) coordinates
class TimeObserver {
 public:
  TimeObserver(unsigned int interval, const ZarrStore& store, size_t maxchunk)
      : interval_(interval), time_(store, "time", maxchunk, {}, {"time"}, "s", kTIME0) {}
  bool on_step(unsigned int t) const { return t % interval_ == 0; }
  void observe(GpuDomain&, unsigned int t) {
    const float v = static_cast<float>(static_cast<double>(t) / 100000.0);
    time_.append(&v, 1);
  }

 private:
  unsigned int interval_;
  ZarrArray<float> time_;
};

inline void write_gbxindex(const ZarrStore& store, uint64_t ngbx) {
  ZarrArray<uint32_t> a(store, "gbxindex", ngbx, {}, {"gbxindex"}, "", 1.0);
  std::vector<uint32_t> idx(ngbx);
  for (uint64_t g = 0; g < ngbx; ++g) idx[g] = static_cast<uint32_t>(g);
  a.append(idx.data(), ngbx);
}

// MassMomentsObserver (massmoments_observer.hpp:360-390): massmom0/1/2 over (time, gbxindex)
class MassMomentsObserver {
 public:
  MassMomentsObserver(unsigned int interval, const ZarrStore& store, size_t maxchunk, uint64_t ngbx, bool rain = false)
      : interval_(interval),
        rain_(rain),
        m0_(store, rain ? "massmom0_raindrops" : "massmom0", chunk0(maxchunk, ngbx), {ngbx}, {"time", "gbxindex"}, "",
            1.0),
        m1_(store, rain ? "massmom1_raindrops" : "massmom1", chunk0(maxchunk, ngbx), {ngbx}, {"time", "gbxindex"}, "g",
            kMASS0grams),
        m2_(store, rain ? "massmom2_raindrops" : "massmom2", chunk0(maxchunk, ngbx), {ngbx}, {"time", "gbxindex"},
            "g^2", kMASS0grams * kMASS0grams) {}
  bool on_step(unsigned int t) const { return t % interval_ == 0; }
  void observe(GpuDomain& d, unsigned int) {
    const auto m = d.mass_moments(rain_);
    m0_.append(m.mom0.data(), m.mom0.size());
    m1_.append(m.mom1.data(), m.mom1.size());
    m2_.append(m.mom2.data(), m.mom2.size());
  }

 private:
  static size_t chunk0(size_t maxchunk, uint64_t ngbx) { return std::max<size_t>(1, maxchunk / ngbx); }
  unsigned int interval_;
  bool rain_;
  ZarrArray<uint64_t> m0_;
  ZarrArray<float> m1_, m2_;
};

// SuperdropsObserver (superdrops_observer.hpp:480-490): the in-domain drops of every observation
// appended to ragged arrays, and their number to raggedcount
class SuperdropsObserver {
 public:
  SuperdropsObserver(unsigned int interval, const ZarrStore& store, size_t maxchunk)
      : interval_(interval),
        count_(store, "raggedcount", maxchunk, {}, {"time"}, "", 1.0, "superdroplets"),
        id_(store, "sdId", maxchunk, {}, {"superdroplets"}, "", 1.0, "superdroplets"),
        gbx_(store, "sdgbxindex", maxchunk, {}, {"superdroplets"}, "", 1.0, "superdroplets"),
        xi_(store, "xi", maxchunk, {}, {"superdroplets"}, "", 1.0, "superdroplets"),
        r_(store, "radius", maxchunk, {}, {"superdroplets"}, "micro-m", kR0 * 1e6, "superdroplets"),
        m_(store, "msol", maxchunk, {}, {"superdroplets"}, "g", kMASS0grams, "superdroplets"),
        c3_(store, "coord3", maxchunk, {}, {"superdroplets"}, "m", kCOORD0, "superdroplets"),
        c1_(store, "coord1", maxchunk, {}, {"superdroplets"}, "m", kCOORD0, "superdroplets"),
        c2_(store, "coord2", maxchunk, {}, {"superdroplets"}, "m", kCOORD0, "superdroplets") {}
  bool on_step(unsigned int t) const { return t % interval_ == 0; }
  void observe(GpuDomain& d, unsigned int) {
    uint64_t nd = 0;
    const std::vector<Superdrop> s = d.download(&nd);
    std::vector<uint32_t> id(nd), gbx(nd);
    std::vector<uint64_t> xi(nd);
    std::vector<float> r(nd), m(nd), c3(nd), c1(nd), c2(nd);
    for (uint64_t k = 0; k < nd; ++k) {
      id[k] = static_cast<uint32_t>(s[k].sdId);
      gbx[k] = s[k].sdgbxindex;
      xi[k] = s[k].attrs.xi;
      r[k] = static_cast<float>(s[k].attrs.radius);
      m[k] = static_cast<float>(s[k].attrs.msol);
      c3[k] = static_cast<float>(s[k].coord3);
      c1[k] = static_cast<float>(s[k].coord1);
      c2[k] = static_cast<float>(s[k].coord2);
    }
    const uint32_t n32 = static_cast<uint32_t>(nd);
    count_.append(&n32, 1);
    id_.append(id.data(), nd);
    gbx_.append(gbx.data(), nd);
    xi_.append(xi.data(), nd);
    r_.append(r.data(), nd);
    m_.append(m.data(), nd);
    c3_.append(c3.data(), nd);
    c1_.append(c1.data(), nd);
    c2_.append(c2.data(), nd);
  }

 private:
  unsigned int interval_;
  ZarrArray<uint32_t> count_, id_, gbx_;
  ZarrArray<uint64_t> xi_;
  ZarrArray<float> r_, m_, c3_, c1_, c2_;
};

}  // namespace cleo_gpu
