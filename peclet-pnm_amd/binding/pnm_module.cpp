// pnm_module.cpp — the `peclet.pnm._pnm` Python module over the C-ABI of libpnm_hip.so.
//
// Mirrors src/pnm_bindings.cpp of the reference (same names, argument order and meaning,
// same RuntimeError messages for bad input) with pybind11 in place of nanobind, which this
// toolchain does not ship. Differences, each deliberate:
//   - segment_volume / extract_pore_network return the segmentation as a 1-D numpy int32 array
//     (the reference returns a Python list of n ints: ~5 GB at 512^3). Every reference use
//     (np.array(seg), len(set(seg)), extract_topology_gpu(seg, ...)) accepts it.
//   - cases that are undefined behaviour in the reference raise RuntimeError instead:
//     origin/spacing/grad_p with fewer than 3 entries, len(seg) != prod(shape).
//   - the GIL is released around device work.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "pnm_capi.h"

namespace py = pybind11;

namespace {

using farr = py::array_t<float, py::array::c_style | py::array::forcecast>;
using darr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using iarr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;

std::mutex g_mu;
std::vector<pnm_ctx_t> g_all;
int g_device = -1;

int default_device() {
  if (g_device >= 0) return g_device;
  const char* e = std::getenv("PNM_DEVICE");
  return e ? std::atoi(e) : 0;
}

void check(pnm_status s) {
  if (s != PNM_OK) throw std::runtime_error(pnm_last_error());
}

// One context per Python thread, created on first use, destroyed at interpreter exit.
pnm_ctx_t ctx() {
  thread_local pnm_ctx_t c = nullptr;
  thread_local int dev = -1;
  const int want = default_device();
  if (c && dev == want) return c;
  pnm_ctx_t nc = nullptr;
  check(pnm_ctx_create(&nc, want));
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_all.push_back(nc);
  }
  c = nc;
  dev = want;
  return c;
}

void destroy_all() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (pnm_ctx_t c : g_all) pnm_ctx_destroy(c);
  g_all.clear();
}

// C-order (Nz,Ny,Nx) float SDF -> res = (Nx,Ny,Nz)   (pnm_bindings.cpp:57-62)
std::array<int32_t, 3> sdf_res(const farr& a) {
  if (a.ndim() != 3) throw std::runtime_error("SDF array must be 3D (Nz,Ny,Nx)");
  return {(int32_t)a.shape(2), (int32_t)a.shape(1), (int32_t)a.shape(0)};
}

// (pnm_bindings.cpp:65-71)
void check_field(const darr& a, const std::array<int32_t, 3>& res, const char* name) {
  if (a.ndim() != 3 || (int32_t)a.shape(2) != res[0] || (int32_t)a.shape(1) != res[1] ||
      (int32_t)a.shape(0) != res[2])
    throw std::runtime_error(std::string(name) + " must be (Nz,Ny,Nx) matching the SDF");
}

// zyx sequence -> xyz floats (pnm_bindings.cpp:116-118)
std::array<float, 3> zyx_f(const std::vector<double>& v, const char* name) {
  if (v.size() < 3) throw std::runtime_error(std::string(name) + " must have 3 entries (z,y,x)");
  return {(float)v[2], (float)v[1], (float)v[0]};
}

struct Owned {
  void* p = nullptr;
  ~Owned() { pnm_free(p); }
};

py::list pores_list(const pnm_pore_t* p, int64_t n) {
  py::list out(n);
  for (int64_t i = 0; i < n; ++i) out[i] = py::cast(p[i]);
  return out;
}

py::list pairs_list(const int32_t* p, int64_t n) {
  py::list out(n);
  for (int64_t i = 0; i < n; ++i) out[i] = py::make_tuple(p[2 * i], p[2 * i + 1]);
  return out;
}

template <class T>
py::list vec_list(const T* p, int64_t n) {
  py::list out(n);
  for (int64_t i = 0; i < n; ++i) out[i] = py::cast(p[i]);
  return out;
}

void warn_truncated(int64_t total) {
  if (total > PNM_MAX_PORES) {
    auto warnings = py::module_::import("warnings");
    warnings.attr("warn")("extract_pores: " + std::to_string(total) +
                          " peaks found; returning the first " +
                          std::to_string(PNM_MAX_PORES) + " by voxel index (max_pores)");
  }
}

// ---- SDFReader.read_vti: sdf_reader.cpp:28-152 (raw appended VTI, Float32 data after a UInt64
// byte count, header in the first 8 KiB, WholeExtent as point extents) ----
std::string vti_attr(const std::string& content, const std::string& name) {
  const size_t pos = content.find(name + "=\"");
  if (pos == std::string::npos) return "";
  const size_t start = pos + name.size() + 2;
  const size_t end = content.find('"', start);
  if (end == std::string::npos) return "";
  return content.substr(start, end - start);
}

py::tuple read_vti(const std::string& filename) {
  std::ifstream file(filename, std::ios::binary);
  if (!file.is_open()) throw std::runtime_error("Could not open file: " + filename);
  std::vector<char> buffer(8192);
  file.read(buffer.data(), (std::streamsize)buffer.size());
  const std::string header(buffer.data(), (size_t)file.gcount());
  const std::string ext = vti_attr(header, "WholeExtent");
  if (ext.empty()) throw std::runtime_error("Could not find WholeExtent in VTI header");
  int x1 = 0, x2 = -1, y1 = 0, y2 = -1, z1 = 0, z2 = -1;
  std::stringstream se(ext);
  se >> x1 >> x2 >> y1 >> y2 >> z1 >> z2;
  const int64_t nx = (int64_t)x2 - x1 + 1, ny = (int64_t)y2 - y1 + 1, nz = (int64_t)z2 - z1 + 1;
  if (nx < 0 || ny < 0 || nz < 0) throw std::runtime_error("negative WholeExtent in VTI header");
  float org[3] = {0.0f, 0.0f, 0.0f}, spc[3] = {0.0f, 0.0f, 0.0f};
  {
    std::stringstream so(vti_attr(header, "Origin"));
    so >> org[0] >> org[1] >> org[2];
    std::stringstream ss(vti_attr(header, "Spacing"));
    ss >> spc[0] >> spc[1] >> spc[2];
  }
  const std::string tag = "<AppendedData encoding=\"raw\">";
  const size_t tag_pos = header.find(tag);
  if (tag_pos == std::string::npos) throw std::runtime_error("Could not find AppendedData tag");
  const size_t us = header.find('_', tag_pos);
  if (us == std::string::npos)
    throw std::runtime_error("Could not find binary data marker '_' in first 8KB");
  file.clear();
  file.seekg((std::streamoff)(us + 1), std::ios::beg);
  uint64_t data_bytes = 0;
  file.read(reinterpret_cast<char*>(&data_bytes), sizeof(uint64_t));
  if (!file) throw std::runtime_error("Failed to read data_bytes (UInt64) from file.");
  py::print("Read data length:", data_bytes);
  const uint64_t expected = (uint64_t)(nx * ny * nz) * sizeof(float);
  if (data_bytes != expected)
    std::cerr << "Warning: Data length in file (" << data_bytes
              << ") does not match expected size from resolution (" << expected << ")."
              << std::endl;
  py::array_t<float> arr({(py::ssize_t)nz, (py::ssize_t)ny, (py::ssize_t)nx});
  float* dst = arr.mutable_data();
  std::fill(dst, dst + nx * ny * nz, 0.0f);
  // the reference reads data_bytes into a buffer of `expected` bytes; never read past it
  const uint64_t want = data_bytes < expected ? data_bytes : expected;
  file.read(reinterpret_cast<char*>(dst), (std::streamsize)want);
  if ((uint64_t)file.gcount() != want)
    std::cerr << "Warning: Could not read all data. Read " << file.gcount() << " bytes."
              << std::endl;
  py::list o, sp;
  for (int a = 2; a >= 0; --a) {
    o.append((double)org[a]);
    sp.append((double)spc[a]);
  }
  return py::make_tuple(arr, o, sp);
}

struct SDFReaderTag {};

}  // namespace

PYBIND11_MODULE(_pnm, m) {
  m.attr("__doc__") = "pnm — pore-network extraction from SDF geometry (HIP, gfx950)";
  m.attr("execution_space") = py::str(pnm_execution_space());
  py::module_::import("atexit").attr("register")(py::cpp_function([]() { destroy_all(); }));

  m.def(
      "set_device", [](int d) { g_device = d; }, py::arg("device"),
      "Select the HIP device used by subsequent calls on every thread (default: $PNM_DEVICE or 0).");
  m.def("device_count", []() { return pnm_device_count(); });

  // pnm_bindings.cpp:87-102
  py::class_<SDFReaderTag>(m, "SDFReader")
      .def_static("read_vti", &read_vti, py::arg("filename"),
                  "Reads VTI; returns (sdf_3d[nz,ny,nx], origin_zyx, spacing_zyx)");

  py::class_<pnm_pore_t>(m, "Pore")
      .def(py::init<>())
      .def_readwrite("x", &pnm_pore_t::x)
      .def_readwrite("y", &pnm_pore_t::y)
      .def_readwrite("z", &pnm_pore_t::z)
      .def_readwrite("radius", &pnm_pore_t::radius)
      .def("__repr__", [](const pnm_pore_t& p) {
        return "Pore(x=" + std::to_string(p.x) + ", y=" + std::to_string(p.y) +
               ", z=" + std::to_string(p.z) + ", radius=" + std::to_string(p.radius) + ")";
      });

  // pnm_bindings.cpp:110-121
  m.def(
      "extract_pores",
      [](farr sdf, std::vector<double> origin_zyx, std::vector<double> spacing_zyx) {
        auto res = sdf_res(sdf);
        auto org = zyx_f(origin_zyx, "origin_zyx");
        auto spc = zyx_f(spacing_zyx, "spacing_zyx");
        Owned o;
        int64_t n = 0, total = 0;
        {
          py::gil_scoped_release nogil;
          pnm_pore_t* p = nullptr;
          const int64_t nv = (int64_t)res[0] * res[1] * res[2];
          pnm_ctx_t c = nv ? ctx() : nullptr;
          if (nv) check(pnm_extract_pores(c, sdf.data(), res.data(), org.data(), spc.data(), &p,
                                          &n, &total));
          o.p = p;
        }
        warn_truncated(total);
        return pores_list(static_cast<pnm_pore_t*>(o.p), n);
      },
      py::arg("sdf"), py::arg("origin_zyx"), py::arg("spacing_zyx"));

  // pnm_bindings.cpp:123-132
  m.def(
      "segment_volume",
      [](farr sdf, std::vector<double> spacing_zyx) {
        auto res = sdf_res(sdf);
        auto spc = zyx_f(spacing_zyx, "spacing_zyx");
        const int64_t nv = (int64_t)res[0] * res[1] * res[2];
        py::array_t<int32_t> seg(nv);
        int32_t* sp = seg.mutable_data();
        {
          py::gil_scoped_release nogil;
          if (nv) check(pnm_segment_volume(ctx(), sdf.data(), res.data(), spc.data(), sp));
        }
        return seg;
      },
      py::arg("sdf"), py::arg("spacing_zyx"));

  // pnm_bindings.cpp:134-140 (shape is zyx)
  m.def(
      "extract_topology_gpu",
      [](iarr segmentation, std::vector<int64_t> shape) {
        if (shape.size() != 3) throw std::runtime_error("shape must have 3 entries (z,y,x)");
        std::array<int32_t, 3> res{(int32_t)shape[2], (int32_t)shape[1], (int32_t)shape[0]};
        const int64_t nv = shape[0] * shape[1] * shape[2];
        if (segmentation.size() != nv)
          throw std::runtime_error("segmentation length " + std::to_string(segmentation.size()) +
                                   " does not match prod(shape) = " + std::to_string(nv));
        Owned o;
        int64_t n = 0;
        {
          py::gil_scoped_release nogil;
          int32_t* p = nullptr;
          if (nv) check(pnm_extract_topology(ctx(), segmentation.data(), res.data(), &p, &n));
          o.p = p;
        }
        return pairs_list(static_cast<int32_t*>(o.p), n);
      },
      py::arg("segmentation"), py::arg("shape"));

  // pnm_bindings.cpp:143-157
  m.def(
      "extract_pore_network",
      [](farr sdf, std::vector<double> origin_zyx, std::vector<double> spacing_zyx) {
        auto res = sdf_res(sdf);
        auto org = zyx_f(origin_zyx, "origin_zyx");
        auto spc = zyx_f(spacing_zyx, "spacing_zyx");
        const int64_t nv = (int64_t)res[0] * res[1] * res[2];
        py::array_t<int32_t> seg(nv);
        int32_t* sp = seg.mutable_data();
        Owned op, oq;
        int64_t np = 0, total = 0, nq = 0;
        {
          py::gil_scoped_release nogil;
          pnm_pore_t* p = nullptr;
          int32_t* q = nullptr;
          if (nv)
            check(pnm_extract_pore_network(ctx(), sdf.data(), res.data(), org.data(), spc.data(),
                                           &p, &np, &total, sp, &q, &nq));
          op.p = p;
          oq.p = q;
        }
        warn_truncated(total);
        return py::make_tuple(pores_list(static_cast<pnm_pore_t*>(op.p), np), seg,
                              pairs_list(static_cast<int32_t*>(oq.p), nq));
      },
      py::arg("sdf"), py::arg("origin_zyx"), py::arg("spacing_zyx"),
      "Fused extraction (SDF uploaded once, segmentation device-resident across stages): returns "
      "(pores, segmentation_flat, connections).");

  // pnm_bindings.cpp:162-209
  m.def(
      "extract_network_flow",
      [](farr sdf, std::vector<double> origin_zyx, std::vector<double> spacing_zyx, darr u,
         darr v, darr w, darr p, py::object ox, py::object oy, py::object oz,
         std::vector<double> grad_p_zyx) {
        auto res = sdf_res(sdf);
        if (ox.is_none() != oy.is_none() || ox.is_none() != oz.is_none())
          throw std::runtime_error("pass all three openness arrays (ox,oy,oz) or none");
        auto org = zyx_f(origin_zyx, "origin_zyx");
        auto spc = zyx_f(spacing_zyx, "spacing_zyx");
        if (grad_p_zyx.size() < 3)
          throw std::runtime_error("grad_p_zyx must have 3 entries (z,y,x)");
        const double gp[3] = {grad_p_zyx[2], grad_p_zyx[1], grad_p_zyx[0]};
        check_field(u, res, "u");
        check_field(v, res, "v");
        check_field(w, res, "w");
        check_field(p, res, "p");
        const bool ho = !ox.is_none();
        darr oxa, oya, oza;
        if (ho) {
          oxa = ox.cast<darr>();
          oya = oy.cast<darr>();
          oza = oz.cast<darr>();
          check_field(oxa, res, "ox");
          check_field(oya, res, "oy");
          check_field(oza, res, "oz");
        }
        pnm_flow_result_t r{};
        {
          py::gil_scoped_release nogil;
          const int64_t nv = (int64_t)res[0] * res[1] * res[2];
          if (nv)
            check(pnm_extract_network_flow(ctx(), sdf.data(), res.data(), org.data(), spc.data(),
                                           u.data(), v.data(), w.data(), p.data(),
                                           ho ? oxa.data() : nullptr, ho ? oya.data() : nullptr,
                                           ho ? oza.data() : nullptr, gp, &r));
        }
        py::dict d;
        d["pores"] = pores_list(r.pores, r.np);
        d["pore_pressure"] = vec_list(r.pore_pressure, r.np);
        d["pore_residual"] = vec_list(r.pore_residual, r.np);
        d["throats"] = pairs_list(r.throats, r.nt);
        d["throat_flow"] = vec_list(r.throat_flow, r.nt);
        d["throat_area"] = vec_list(r.throat_area, r.nt);
        d["throat_dp"] = vec_list(r.throat_dp, r.nt);
        pnm_flow_result_free(&r);
        return d;
      },
      py::arg("sdf"), py::arg("origin_zyx"), py::arg("spacing_zyx"), py::arg("u"), py::arg("v"),
      py::arg("w"), py::arg("p"), py::arg("ox") = py::none(), py::arg("oy") = py::none(),
      py::arg("oz") = py::none(),
      py::arg("grad_p_zyx") = std::vector<double>{0.0, 0.0, 0.0},
      "Pore-network flow data from a MAC field on the SDF grid: per-pore-label centres, "
      "pressures (trilinear p at the basin peak, periodic) and residuals; per-throat (connected "
      "face patch) flow rates, areas and total-pressure drops.");
}
