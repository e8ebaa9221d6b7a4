// C ABI of libminicfd (include/minicfd.h): handle management, device
// buffers and the System / RungeKutta entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/minicfd.h"
#include "../kernels/fr_hex.hpp"
#include "hex_connectivity.hpp"
#include "tables.hpp"

namespace {

thread_local std::string g_error;

struct Failure {
  int code;
  std::string what;
};

void Check(hipError_t e, const char *where) {
  if (e != hipSuccess) throw Failure{ MCFD_EHIP, std::string(where) + ": " + hipGetErrorString(e) };
}

template <class F>
int Guard(F &&f) {
  try {
    f();
    return MCFD_OK;
  } catch (Failure const &e) {
    g_error = e.what;
    return e.code;
  } catch (std::invalid_argument const &e) {
    g_error = e.what();
    return MCFD_EINVAL;
  } catch (std::exception const &e) {
    g_error = e.what();
    return MCFD_EINVAL;
  }
}

template <class T>
struct DeviceArray {
  T *ptr = nullptr;
  size_t n = 0;
  void Allocate(size_t count) {
    Release();
    if (count == 0) return;
    Check(hipMalloc(&ptr, count * sizeof(T)), "hipMalloc");
    n = count;
  }
  void Upload(const std::vector<T> &v, hipStream_t s) {
    Allocate(v.size());
    if (n) Check(hipMemcpyAsync(ptr, v.data(), n * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
  void Release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  ~DeviceArray() { Release(); }
};

// Inverse of an orientation code (see hex_connectivity.hpp ApplyOrient).
int InverseOrient(int code) {
  const int q = 4;
  for (int c = 0; c < 8; ++c) {
    bool ok = true;
    for (int a = 0; a < q && ok; ++a)
      for (int b = 0; b < q && ok; ++b) {
        int a1, b1, a2, b2;
        mcfd::ApplyOrient(code, q, a, b, &a1, &b1);
        mcfd::ApplyOrient(c, q, a1, b1, &a2, &b2);
        ok = (a2 == a && b2 == b);
      }
    if (ok) return c;
  }
  return -1;
}

}  // namespace

struct mcfd_solver {
  mcfd_scheme_desc scheme;
  int q = 0, np = 0;
  int64_t n_cells = 0;
  double t = 0;
  bool host_only = false;
  bool diag_metric = false;  // every geometry record has a diagonal J*
  hipStream_t stream = nullptr;
  mcfd::HexConnectivity conn;
  mcfd::LineTables lt;
  mcfd::BcTable bc;
  std::vector<double> x1d;
  std::vector<std::array<double, 3>> corner0;  // node 0 of each cell
  DeviceArray<double> u0, u1, u2, res;
  DeviceArray<int32_t> nbr, finfo, geo_id;
  DeviceArray<double> geo, frames;
  DeviceArray<unsigned long long> scratch;

  size_t ColumnSize() const { return size_t(n_cells) * np * 5; }

  void Stage(int stage, double dt) {
    mcfd::FrHexArgs a;
    a.u0 = u0.ptr;
    a.nbr = nbr.ptr; a.finfo = finfo.ptr; a.geo_id = geo_id.ptr;
    a.geo = geo.ptr; a.frames = frames.ptr;
    a.n_cells = n_cells;
    a.dt = dt;
    a.stage = stage;
    a.xcd_remap = 1;
    a.diag_metric = diag_metric ? 1 : 0;
    a.mixed = scheme.precision == MCFD_PRECISION_MIXED ? 1 : 0;
    switch (stage) {
    case 0: a.us = u0.ptr; a.out = res.ptr; break;
    case 1: a.us = u0.ptr; a.out = u1.ptr; break;
    case 2: a.us = u1.ptr; a.out = u2.ptr; break;
    default: a.us = u2.ptr; a.out = u0.ptr; break;
    }
    Check(mcfd::LaunchFrHexStage(q, scheme.riemann, a, lt, bc, stream), "FR stage launch");
  }
};

extern "C" {

const char *mcfd_last_error(void) { return g_error.c_str(); }

int mcfd_create(const mcfd_mesh_desc *mesh, const mcfd_scheme_desc *scheme, const mcfd_bc_desc *bcs,
                int32_t n_bcs, mcfd_solver **out) {
  if (!mesh || !scheme || !out || (n_bcs > 0 && !bcs)) {
    g_error = "null argument";
    return MCFD_EINVAL;
  }
  *out = nullptr;
  auto h = std::make_unique<mcfd_solver>();
  int rc = Guard([&] {
    if (scheme->scheme != MCFD_SCHEME_FR_GENERAL || mesh->cell_type != MCFD_HEXA_8)
      throw Failure{ MCFD_ENOTSUP, "only FR_GENERAL on HEXA_8 is implemented in this build" };
    if (scheme->degree < 1 || scheme->degree > 5) throw std::invalid_argument("degree must be 1..5");
    if (scheme->riemann < 0 || scheme->riemann > 2) throw std::invalid_argument("unknown Riemann solver");
    if (scheme->precision != MCFD_PRECISION_FP64 && scheme->precision != MCFD_PRECISION_MIXED)
      throw std::invalid_argument("unknown precision form");
    if (n_bcs < 0 || n_bcs > mcfd::kMaxBcs) throw std::invalid_argument("too many boundary conditions");
    h->scheme = *scheme;
    h->q = scheme->degree + 1;
    h->np = h->q * h->q * h->q;
    h->n_cells = mesh->n_cells;
    for (int64_t b = 0; b < mesh->n_bfaces; ++b)
      if (mesh->bface_tags[b] >= n_bcs) throw std::invalid_argument("boundary tag without a BC entry");
    h->conn = mcfd::BuildHexConnectivity(*mesh);

    // line tables
    const int q = h->q;
    h->x1d.resize(q);
    std::vector<double> w(q);
    mcfd::GaussRule(scheme->nodes == MCFD_NODES_LOBATTO, q, h->x1d.data(), w.data());
    std::memset(&h->lt, 0, sizeof(h->lt));
    for (int i = 0; i < q; ++i)
      for (int m = 0; m < q; ++m) h->lt.d[i][m] = mcfd::LagrangeDerivative(q, h->x1d.data(), m, h->x1d[i]);
    for (int m = 0; m < q; ++m) {
      h->lt.lend[0][m] = mcfd::LagrangeValue(q, h->x1d.data(), m, -1.0);
      h->lt.lend[1][m] = mcfd::LagrangeValue(q, h->x1d.data(), m, +1.0);
      h->lt.gp[0][m] = mcfd::CorrectionLeftDerivative(scheme->degree, scheme->c_next, h->x1d[m]);
      h->lt.gp[1][m] = mcfd::CorrectionRightDerivative(scheme->degree, scheme->c_next, h->x1d[m]);
    }
    std::memset(&h->bc, 0, sizeof(h->bc));
    for (int b = 0; b < n_bcs; ++b) {
      if (bcs[b].kind < 0 || bcs[b].kind > 3) throw std::invalid_argument("unknown BC kind");
      h->bc.kind[b] = bcs[b].kind;
      for (int k = 0; k < 5; ++k) h->bc.state[b][k] = bcs[b].state[k];
    }

    // geometry records, deduplicated bitwise
    const int64_t nc = mesh->n_cells;
    std::map<std::array<double, mcfd::kGeoStride>, int32_t> geo_ids;
    std::vector<double> geo;
    std::vector<int32_t> geo_id(nc);
    h->corner0.resize(nc);
    for (int64_t c = 0; c < nc; ++c) {
      const auto &j = h->conn.jac[c];
      double d = h->conn.det[c];
      // J* = det(J) J^{-1} is the adjugate of J
      std::array<double, mcfd::kGeoStride> g{};
      g[0] = j[4] * j[8] - j[5] * j[7]; g[1] = j[2] * j[7] - j[1] * j[8]; g[2] = j[1] * j[5] - j[2] * j[4];
      g[3] = j[5] * j[6] - j[3] * j[8]; g[4] = j[0] * j[8] - j[2] * j[6]; g[5] = j[2] * j[3] - j[0] * j[5];
      g[6] = j[3] * j[7] - j[4] * j[6]; g[7] = j[1] * j[6] - j[0] * j[7]; g[8] = j[0] * j[4] - j[1] * j[3];
      g[9] = 1.0 / d;
      for (int f = 0; f < 6; ++f) {
        int dir = mcfd::kHexFaceNormalDir[f];
        double n = std::sqrt(g[dir] * g[dir] + g[3 + dir] * g[3 + dir] + g[6 + dir] * g[6 + dir]);
        g[10 + f] = mcfd::kHexFaceSide[f] ? n : -n;
      }
      g[16] = std::cbrt(8 * d) * 0.5;  // part.hpp:230, volume = 8 det(J)
      auto it = geo_ids.find(g);
      if (it == geo_ids.end()) {
        it = geo_ids.emplace(g, (int32_t)geo_ids.size()).first;
        geo.insert(geo.end(), g.begin(), g.end());
      }
      geo_id[c] = it->second;
      int64_t n0 = mesh->cell_nodes[8 * c];
      h->corner0[c] = { mesh->coords[3 * n0], mesh->coords[3 * n0 + 1], mesh->coords[3 * n0 + 2] };
    }

    h->diag_metric = true;
    for (size_t r = 0; r < geo.size(); r += mcfd::kGeoStride)
      for (int e = 0; e < 9; ++e)
        if (e % 4 != 0 && geo[r + e] != 0.0) h->diag_metric = false;

    // per-cell face records
    std::map<std::array<double, mcfd::kFrameStride>, int32_t> frame_ids;
    std::vector<double> frames;
    std::vector<int32_t> nbr(6 * nc, INT32_MIN), finfo(6 * nc, 0);
    for (const auto &face : h->conn.faces) {
      std::array<double, mcfd::kFrameStride> fr;
      std::copy(face.frame, face.frame + 6, fr.begin());
      auto it = frame_ids.find(fr);
      if (it == frame_ids.end()) {
        it = frame_ids.emplace(fr, (int32_t)frame_ids.size()).first;
        frames.insert(frames.end(), fr.begin(), fr.end());
      }
      int32_t fid = it->second;
      if (fid >= (1 << 24)) throw Failure{ MCFD_ENOTSUP, "too many distinct face frames" };
      int64_t hs = face.holder * 6 + face.holder_face;
      if (face.sharer >= 0) {
        int64_t ss = face.sharer * 6 + face.sharer_face;
        nbr[hs] = (int32_t)face.sharer;
        finfo[hs] = (fid << 7) | (1 << 6) | (face.orient << 3) | face.sharer_face;
        nbr[ss] = (int32_t)face.holder;
        finfo[ss] = (fid << 7) | (InverseOrient(face.orient) << 3) | face.holder_face;
      } else {
        nbr[hs] = -(1 + face.tag);
        finfo[hs] = (fid << 7) | (1 << 6);
      }
    }
    for (auto v : nbr)
      if (v == INT32_MIN) throw std::invalid_argument("cell face without a face record");

    h->host_only = (scheme->flags & MCFD_FLAG_HOST_ONLY) != 0;
    if (h->host_only) return;
    Check(hipStreamCreate(&h->stream), "hipStreamCreate");
    size_t cs = h->ColumnSize();
    h->u0.Allocate(cs); h->u1.Allocate(cs); h->u2.Allocate(cs); h->res.Allocate(cs);
    Check(hipMemsetAsync(h->u0.ptr, 0, cs * sizeof(double), h->stream), "memset");
    h->nbr.Upload(nbr, h->stream);
    h->finfo.Upload(finfo, h->stream);
    h->geo_id.Upload(geo_id, h->stream);
    h->geo.Upload(geo, h->stream);
    h->frames.Upload(frames, h->stream);
    h->scratch.Allocate(1);
    Check(hipStreamSynchronize(h->stream), "create sync");
  });
  if (rc == MCFD_OK) *out = h.release();
  return rc;
}

int mcfd_destroy(mcfd_solver *h) {
  if (!h) return MCFD_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->u0.Release(); h->u1.Release(); h->u2.Release(); h->res.Release();
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return MCFD_OK;
}

int mcfd_sizes(const mcfd_solver *h, int64_t *n_cells, int32_t *n_per_cell, int64_t *column_size) {
  if (!h) { g_error = "null handle"; return MCFD_EINVAL; }
  if (n_cells) *n_cells = h->n_cells;
  if (n_per_cell) *n_per_cell = h->np;
  if (column_size) *column_size = (int64_t)h->ColumnSize();
  return MCFD_OK;
}

int mcfd_set_time(mcfd_solver *h, double t) {
  if (!h) { g_error = "null handle"; return MCFD_EINVAL; }
  h->t = t;
  return MCFD_OK;
}

int mcfd_set_solution_column(mcfd_solver *h, const double *column) {
  if (!h || !column) { g_error = "null argument"; return MCFD_EINVAL; }
  if (h->host_only) { g_error = "handle was created host-only"; return MCFD_ESTATE; }
  return Guard([&] {
    Check(hipMemcpyAsync(h->u0.ptr, column, h->ColumnSize() * sizeof(double), hipMemcpyHostToDevice,
                         h->stream), "set solution");
    Check(hipStreamSynchronize(h->stream), "set solution sync");
  });
}

int mcfd_get_solution_column(mcfd_solver *h, double *column) {
  if (!h || !column) { g_error = "null argument"; return MCFD_EINVAL; }
  if (h->host_only) { g_error = "handle was created host-only"; return MCFD_ESTATE; }
  return Guard([&] {
    Check(hipMemcpyAsync(column, h->u0.ptr, h->ColumnSize() * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream), "get solution");
    Check(hipStreamSynchronize(h->stream), "get solution sync");
  });
}

int mcfd_get_residual_column(mcfd_solver *h, double *column) {
  if (!h || !column) { g_error = "null argument"; return MCFD_EINVAL; }
  if (h->host_only) { g_error = "handle was created host-only"; return MCFD_ESTATE; }
  return Guard([&] {
    h->Stage(0, 0.0);
    Check(hipMemcpyAsync(column, h->res.ptr, h->ColumnSize() * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream), "get residual");
    Check(hipStreamSynchronize(h->stream), "get residual sync");
  });
}

int mcfd_rk3_stage(mcfd_solver *h, int32_t stage, double t, double dt) {
  if (!h || stage < 1 || stage > 3) { g_error = "bad handle or stage"; return MCFD_EINVAL; }
  if (h->host_only) { g_error = "handle was created host-only"; return MCFD_ESTATE; }
  return Guard([&] {
    h->t = t;
    h->Stage(stage, dt);
  });
}

int mcfd_rk3_update(mcfd_solver *h, double t, double dt) {
  if (!h) { g_error = "null handle"; return MCFD_EINVAL; }
  if (h->host_only) { g_error = "handle was created host-only"; return MCFD_ESTATE; }
  return Guard([&] {
    h->t = t;
    h->Stage(1, dt);
    h->Stage(2, dt);
    h->Stage(3, dt);
    h->t = t + dt;
  });
}

int mcfd_get_time_step(mcfd_solver *h, double cfl, double dt_guess, double *dt) {
  if (!h || !dt) { g_error = "null argument"; return MCFD_EINVAL; }
  if (h->host_only) { g_error = "handle was created host-only"; return MCFD_ESTATE; }
  return Guard([&] {
    double init = 1e300;
    unsigned long long bits;
    std::memcpy(&bits, &init, sizeof(bits));
    Check(hipMemcpyAsync(h->scratch.ptr, &bits, sizeof(bits), hipMemcpyHostToDevice, h->stream), "dt init");
    Check(mcfd::LaunchFrHexTimeStep(h->q, h->u0.ptr, h->geo_id.ptr, h->geo.ptr, h->n_cells, h->scratch.ptr,
                                    h->stream), "time step launch");
    Check(hipMemcpyAsync(&bits, h->scratch.ptr, sizeof(bits), hipMemcpyDeviceToHost, h->stream), "dt read");
    Check(hipStreamSynchronize(h->stream), "dt sync");
    double min_dt;
    std::memcpy(&min_dt, &bits, sizeof(min_dt));
    *dt = std::min(cfl * min_dt, dt_guess);
  });
}

int mcfd_sample_points(const mcfd_solver *h, int32_t *n_per_cell, double *xyz) {
  if (!h) { g_error = "null handle"; return MCFD_EINVAL; }
  if (n_per_cell) *n_per_cell = h->np;
  if (!xyz) return MCFD_OK;
  const int q = h->q;
  for (int64_t c = 0; c < h->n_cells; ++c) {
    const auto &j = h->conn.jac[c];
    const auto &x0 = h->corner0[c];
    for (int i = 0; i < q; ++i)
      for (int jj = 0; jj < q; ++jj)
        for (int k = 0; k < q; ++k) {
          double s[3] = { h->x1d[i] + 1, h->x1d[jj] + 1, h->x1d[k] + 1 };
          double *p = xyz + ((c * h->np) + (i * q + jj) * q + k) * 3;
          for (int d = 0; d < 3; ++d) p[d] = x0[d] + s[0] * j[d] + s[1] * j[3 + d] + s[2] * j[6 + d];
        }
  }
  return MCFD_OK;
}

int mcfd_approximate(mcfd_solver *h, const double *values) {
  // Nodal FR: the solution points are the sample points
  // (polynomial/hexahedron.hpp:646-652).
  return mcfd_set_solution_column(h, values);
}

int mcfd_face_count(const mcfd_solver *h, int64_t *n_faces, int64_t *n_interior) {
  if (!h) { g_error = "null handle"; return MCFD_EINVAL; }
  if (n_faces) *n_faces = (int64_t)h->conn.faces.size();
  if (n_interior) *n_interior = h->conn.n_interior;
  return MCFD_OK;
}

int mcfd_faces(const mcfd_solver *h, int64_t *holder, int64_t *sharer, int32_t *holder_face,
               int32_t *sharer_face, int32_t *fp_ijk) {
  if (!h) { g_error = "null handle"; return MCFD_EINVAL; }
  const int q = h->q, nfp = q * q;
  for (size_t f = 0; f < h->conn.faces.size(); ++f) {
    const auto &face = h->conn.faces[f];
    if (holder) holder[f] = face.holder;
    if (sharer) sharer[f] = face.sharer;
    if (holder_face) holder_face[f] = face.holder_face;
    if (sharer_face) sharer_face[f] = face.sharer_face;
    if (!fp_ijk) continue;
    for (int i = 0; i < q; ++i)
      for (int j = 0; j < q; ++j) {
        int a, b;
        mcfd::ApplyOrient(face.quad_to_holder, q, i, j, &a, &b);
        int fh = face.holder_face;
        int32_t *o = fp_ijk + (f * nfp + i * q + j) * 2;
        o[0] = mcfd::HexLineNode(q, fh, a, b, mcfd::kHexFaceSide[fh] ? q - 1 : 0);
        o[1] = -1;
        if (face.sharer >= 0) {
          int a2, b2, fs = face.sharer_face;
          mcfd::ApplyOrient(face.orient, q, a, b, &a2, &b2);
          o[1] = mcfd::HexLineNode(q, fs, a2, b2, mcfd::kHexFaceSide[fs] ? q - 1 : 0);
        }
      }
  }
  return MCFD_OK;
}

int mcfd_synchronize(mcfd_solver *h) {
  if (!h) { g_error = "null handle"; return MCFD_EINVAL; }
  if (h->host_only) { g_error = "handle was created host-only"; return MCFD_ESTATE; }
  return Guard([&] { Check(hipStreamSynchronize(h->stream), "synchronize"); });
}

int mcfd_time_rk3(mcfd_solver *h, double dt, int32_t n_steps, double *total_ms, double *stage_kernel_ms) {
  if (!h || n_steps < 1) { g_error = "bad argument"; return MCFD_EINVAL; }
  if (h->host_only) { g_error = "handle was created host-only"; return MCFD_ESTATE; }
  return Guard([&] {
    // One event pair per stage kernel, all recorded on the handle's stream
    // without host synchronisation in between; read back at the end.
    std::vector<hipEvent_t> ev(2 * 3 * (size_t)n_steps + 2);
    for (auto &e : ev) Check(hipEventCreate(&e), "hipEventCreate");
    try {
      Check(hipEventRecord(ev[0], h->stream), "event");
      for (int s = 0; s < n_steps; ++s)
        for (int stage = 1; stage <= 3; ++stage) {
          size_t i = 2 + 2 * (3 * (size_t)s + stage - 1);
          Check(hipEventRecord(ev[i], h->stream), "event");
          h->Stage(stage, dt);
          Check(hipEventRecord(ev[i + 1], h->stream), "event");
        }
      Check(hipEventRecord(ev[1], h->stream), "event");
      Check(hipEventSynchronize(ev[1]), "event sync");
      float ms = 0;
      double kernel_ms = 0;
      for (size_t i = 2; i < ev.size(); i += 2) {
        Check(hipEventElapsedTime(&ms, ev[i], ev[i + 1]), "elapsed");
        kernel_ms += ms;
      }
      Check(hipEventElapsedTime(&ms, ev[0], ev[1]), "elapsed");
      if (total_ms) *total_ms = ms;
      if (stage_kernel_ms) *stage_kernel_ms = kernel_ms;
    } catch (...) {
      for (auto &e : ev) (void)hipEventDestroy(e);
      throw;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
  });
}

}  // extern "C"
