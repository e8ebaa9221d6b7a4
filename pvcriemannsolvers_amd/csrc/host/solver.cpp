// C ABI of libminicfd (include/minicfd.h): the single-mesh solver handle,
// i.e. the System / RungeKutta entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "../../../include/minicfd.h"
#include "bc_table.hpp"
#include "common.hpp"
#include "dg_group.hpp"
#include "fr_group.hpp"
#include "../kernels/fr_startup.hpp"

namespace mcfd {

std::string &LastError() {
  thread_local std::string error;
  return error;
}

}  // namespace mcfd

using mcfd::Check;
using mcfd::Fail;
using mcfd::Guard;

struct mcfd_solver {
  double t = 0;
  bool host_only = false;
  bool dg = false;  // modal DG (DgGroup) instead of nodal FR (FrHexGroup)
  hipStream_t stream = nullptr;
  mcfd::FrHexGroup group;
  mcfd::DgGroup modal;
  mcfd::DeviceArray<unsigned long long> scratch;  // time-step reduction
  mcfd::BoundaryTable bct;                         // boundary values that vary in space and time
  // The time step (fem.hpp:412-435) of the state an RK step ends with can be
  // folded into its stage-3 pass.  That is done while the caller reads a time
  // step between steps (as the reference's driver does, sourceless.cpp:190-
  // 193); a folded minimum left unread by the next step switches it off.
  bool dt_fresh = false;    // scratch holds the minimum of the current solution
  bool dt_pending = false;  // ... folded in by the last stage 3 and not read yet
  bool fuse_dt_wanted = false;

  // Stage 0 forms the residual at time t; stages 1..3 advance the step that
  // starts at t, whose residuals are formed at t, t + dt and t + dt / 2
  // (temporal/rk.hpp:31-49 with ode.hpp:121-128 setting the time).
  void Stage(int stage, double dt) {
    const double ts = stage == 2 ? t + dt : (stage == 3 ? t + dt / 2 : t);
    const double *bc_points = bct.At(ts, stream);
    if (stage == 1 && dt_pending) fuse_dt_wanted = false;
    // the fused reduction needs the final state of the pass: not with a
    // limiter, which rewrites troubled cells afterwards
    const bool fuse_dt = stage == 3 && fuse_dt_wanted && !(dg && modal.limiter);
    if (stage > 0) dt_fresh = dt_pending = false;
    if (fuse_dt) Check(mcfd::LaunchResetTimeStepBits(scratch.ptr, stream), "time-step reset");
    if (dg) {
      modal.bc_point_state = bc_points;
      modal.Stage(stage, dt, stream, fuse_dt ? scratch.ptr : nullptr);
    } else {
      mcfd::FrHexArgs a = group.Args(stage, dt);
      a.dt_min_bits = fuse_dt ? scratch.ptr : nullptr;
      a.bc_point_state = bc_points;
      Check(mcfd::LaunchFrHexStage(group.q, group.riemann, a, group.lt, stream), "FR stage launch");
    }
    if (fuse_dt) dt_fresh = dt_pending = true;
  }
  size_t ColumnSize() const { return dg ? modal.ColumnSize() : group.ColumnSize(); }
  double *Solution() { return dg ? modal.u0.ptr : group.u0.ptr; }
  double *Residual() { return dg ? modal.res.ptr : group.res.ptr; }
};

#define MCFD_REQUIRE_DEVICE(h)                                                        \
  do {                                                                                \
    if (!(h)) return Fail(MCFD_EINVAL, "null handle");                                \
    if ((h)->host_only) return Fail(MCFD_ESTATE, "handle was created host-only");     \
  } while (0)

extern "C" {

const char *mcfd_last_error(void) { return mcfd::LastError().c_str(); }

int mcfd_create(const mcfd_mesh_desc *mesh, const mcfd_scheme_desc *scheme, const mcfd_bc_desc *bcs,
                int32_t n_bcs, mcfd_solver **out) {
  if (!mesh || !scheme || !out || n_bcs < 0 || (n_bcs > 0 && !bcs)) return Fail(MCFD_EINVAL, "null argument");
  *out = nullptr;
  auto h = std::make_unique<mcfd_solver>();
  int rc = Guard([&] {
    if (scheme->scheme == MCFD_SCHEME_FR_GENERAL && mesh->cell_type != MCFD_HEXA_8)
      throw mcfd::Failure{ MCFD_ENOTSUP, "FR_GENERAL is implemented on HEXA_8" };
    if (scheme->scheme != MCFD_SCHEME_FR_GENERAL && scheme->scheme != MCFD_SCHEME_DG_GENERAL)
      throw mcfd::Failure{ MCFD_ENOTSUP, "unknown scheme" };
    if (scheme->scheme == MCFD_SCHEME_FR_GENERAL && scheme->limiter != MCFD_LIMITER_NONE)
      throw mcfd::Failure{ MCFD_ENOTSUP, "limiters are implemented for DG_GENERAL" };
    std::vector<int32_t> kinds(n_bcs);
    std::vector<double> states(5 * n_bcs);
    for (int b = 0; b < n_bcs; ++b) {
      kinds[b] = bcs[b].kind;
      for (int k = 0; k < 5; ++k) states[5 * b + k] = bcs[b].state[k];
    }
    h->dg = scheme->scheme == MCFD_SCHEME_DG_GENERAL;
    if (h->dg) {
      h->modal.Build(*mesh, *scheme, kinds, states);
    } else {
      h->group.Build(*mesh, *scheme, kinds, states);
    }
    {
      std::vector<int32_t> tags;
      std::vector<double> xyz;
      int nfp;
      if (h->dg) {
        h->modal.BoundaryPoints(&tags, &xyz);
        nfp = h->modal.nfp;
      } else {
        h->group.BoundaryPoints(&tags, &xyz);
        nfp = h->group.q * h->group.q;
      }
      h->bct.Init(nfp, std::move(tags), std::move(xyz), states);
    }
    h->host_only = (scheme->flags & MCFD_FLAG_HOST_ONLY) != 0;
    if (h->host_only) return;
    Check(hipStreamCreate(&h->stream), "hipStreamCreate");
    if (h->dg) {
      h->modal.Upload(h->stream);
    } else {
      h->group.BuildTileOrder(mcfd::FrHexStageTileCells(h->group.q));
      h->group.Upload(h->stream);
    }
    h->scratch.Allocate(1);
    Check(hipStreamSynchronize(h->stream), "create sync");
  });
  if (rc == MCFD_OK) *out = h.release();
  return rc;
}

int mcfd_destroy(mcfd_solver *h) {
  if (!h) return MCFD_OK;
  hipStream_t s = h->stream;
  if (s) (void)hipStreamSynchronize(s);
  delete h;  // releases the device arrays
  if (s) (void)hipStreamDestroy(s);
  return MCFD_OK;
}

int mcfd_sizes(const mcfd_solver *h, int64_t *n_cells, int32_t *n_per_cell, int64_t *column_size) {
  if (!h) return Fail(MCFD_EINVAL, "null handle");
  if (n_cells) *n_cells = h->dg ? h->modal.n_cells : h->group.n_cells;
  if (n_per_cell) *n_per_cell = h->dg ? h->modal.n : h->group.np;
  if (column_size) *column_size = (int64_t)h->ColumnSize();
  return MCFD_OK;
}

int mcfd_set_time(mcfd_solver *h, double t) {
  if (!h) return Fail(MCFD_EINVAL, "null handle");
  h->t = t;
  return MCFD_OK;
}

int mcfd_set_solution_column(mcfd_solver *h, const double *column) {
  MCFD_REQUIRE_DEVICE(h);
  if (!column) return Fail(MCFD_EINVAL, "null column");
  return Guard([&] {
    h->dt_fresh = false;
    Check(hipMemcpyAsync(h->Solution(), column, h->ColumnSize() * sizeof(double),
                         hipMemcpyHostToDevice, h->stream), "set solution");
    if (h->dg) h->modal.SolutionChanged(h->stream);
    Check(hipStreamSynchronize(h->stream), "set solution sync");
  });
}

int mcfd_restore_solution_column(mcfd_solver *h, const double *column) {
  MCFD_REQUIRE_DEVICE(h);
  if (!column) return Fail(MCFD_EINVAL, "null column");
  return Guard([&] {
    h->dt_fresh = false;
    Check(hipMemcpyAsync(h->Solution(), column, h->ColumnSize() * sizeof(double),
                         hipMemcpyHostToDevice, h->stream), "restore solution");
    if (h->dg) h->modal.SolutionChanged(h->stream, /*limit=*/false);
    Check(hipStreamSynchronize(h->stream), "restore solution sync");
  });
}

int mcfd_get_solution_column(mcfd_solver *h, double *column) {
  MCFD_REQUIRE_DEVICE(h);
  if (!column) return Fail(MCFD_EINVAL, "null column");
  return Guard([&] {
    Check(hipMemcpyAsync(column, h->Solution(), h->ColumnSize() * sizeof(double),
                         hipMemcpyDeviceToHost, h->stream), "get solution");
    Check(hipStreamSynchronize(h->stream), "get solution sync");
  });
}

int mcfd_get_residual_column(mcfd_solver *h, double *column) {
  MCFD_REQUIRE_DEVICE(h);
  if (!column) return Fail(MCFD_EINVAL, "null column");
  return Guard([&] {
    h->Stage(0, 0.0);
    Check(hipMemcpyAsync(column, h->Residual(), h->ColumnSize() * sizeof(double),
                         hipMemcpyDeviceToHost, h->stream), "get residual");
    Check(hipStreamSynchronize(h->stream), "get residual sync");
  });
}

int mcfd_rk3_stage(mcfd_solver *h, int32_t stage, double t, double dt) {
  MCFD_REQUIRE_DEVICE(h);
  if (stage < 1 || stage > 3) return Fail(MCFD_EINVAL, "stage must be 1, 2 or 3");
  return Guard([&] {
    h->t = t;
    h->Stage(stage, dt);
  });
}

int mcfd_rk3_update(mcfd_solver *h, double t, double dt) {
  MCFD_REQUIRE_DEVICE(h);
  return Guard([&] {
    h->t = t;
    h->Stage(1, dt);
    h->Stage(2, dt);
    h->Stage(3, dt);
    h->t = t + dt;
  });
}

int mcfd_get_time_step(mcfd_solver *h, double cfl, double dt_guess, double *dt) {
  MCFD_REQUIRE_DEVICE(h);
  if (!dt) return Fail(MCFD_EINVAL, "null output");
  return Guard([&] {
    double min_dt;
    unsigned long long bits;
    h->fuse_dt_wanted = true;
    h->dt_pending = false;
    if (!h->dt_fresh) {  // no RK stage has folded it in since the solution last changed
      Check(mcfd::LaunchResetTimeStepBits(h->scratch.ptr, h->stream), "time-step reset");
      if (h->dg) {
        h->modal.TimeStepBits(h->scratch.ptr, h->stream);
      } else {
        Check(mcfd::LaunchFrHexTimeStep(h->group.q, h->group.u0.ptr, h->group.geo_id.ptr, h->group.geo.ptr,
                                        h->group.n_cells, h->scratch.ptr, h->stream), "time step launch");
      }
      h->dt_fresh = true;
    }
    Check(hipMemcpyAsync(&bits, h->scratch.ptr, sizeof(bits), hipMemcpyDeviceToHost, h->stream), "dt read");
    Check(hipStreamSynchronize(h->stream), "dt sync");
    std::memcpy(&min_dt, &bits, sizeof(min_dt));
    *dt = std::min(cfl * min_dt, dt_guess);
  });
}

int mcfd_sample_points(const mcfd_solver *h, int32_t *n_per_cell, double *xyz) {
  if (!h) return Fail(MCFD_EINVAL, "null handle");
  if (h->dg) {  // the cell quadrature points
    if (n_per_cell) *n_per_cell = h->modal.nq;
    if (xyz) std::copy(h->modal.qp.begin(), h->modal.qp.end(), xyz);
    return MCFD_OK;
  }
  if (n_per_cell) *n_per_cell = h->group.np;
  if (xyz) h->group.SamplePoints(xyz);
  return MCFD_OK;
}

int mcfd_evaluate(mcfd_solver *h, const double *column, int32_t n_local, const double *local, double *xyz,
                  double *values) {
  if (!h || !local || n_local < 1) return Fail(MCFD_EINVAL, "bad argument");
  if (values && !column && h->host_only) return Fail(MCFD_ESTATE, "handle was created host-only: pass a column");
  return Guard([&] {
    std::vector<double> own;
    if (values && !column) {
      own.resize(h->ColumnSize());
      Check(hipMemcpyAsync(own.data(), h->Solution(), own.size() * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream), "get solution");
      Check(hipStreamSynchronize(h->stream), "get solution sync");
      column = own.data();
    }
    if (h->dg) h->modal.Evaluate(column, n_local, local, xyz, values);
    else h->group.Evaluate(column, n_local, local, xyz, values);
  });
}

int mcfd_approximate(mcfd_solver *h, const double *values) {
  MCFD_REQUIRE_DEVICE(h);
  if (!values) return Fail(MCFD_EINVAL, "null values");
  if (!h->dg) {
    // Nodal FR: the solution points are the sample points
    // (polynomial/hexahedron.hpp:646-652).
    return mcfd_set_solution_column(h, values);
  }
  return Guard([&] {
    h->dt_fresh = false;
    h->modal.Approximate(values, h->stream);  // projected on the device, then limited
  });
}

int mcfd_face_count(const mcfd_solver *h, int64_t *n_faces, int64_t *n_interior) {
  if (!h) return Fail(MCFD_EINVAL, "null handle");
  if (n_faces) *n_faces = (int64_t)(h->dg ? h->modal.faces.size() : h->group.conn.faces.size());
  if (n_interior) *n_interior = h->dg ? h->modal.n_interior : h->group.conn.n_interior;
  return MCFD_OK;
}

int mcfd_faces(const mcfd_solver *h, int64_t *holder, int64_t *sharer, int32_t *holder_face,
               int32_t *sharer_face, int32_t *fp_ijk) {
  if (!h) return Fail(MCFD_EINVAL, "null handle");
  if (h->dg) {
    h->modal.Faces(holder, sharer, holder_face, sharer_face);
    return MCFD_OK;
  }
  h->group.Faces(holder, sharer, holder_face, sharer_face, fp_ijk);
  return MCFD_OK;
}

int mcfd_boundary_points(const mcfd_solver *h, int32_t tag, int64_t *n_points, double *xyz) {
  if (!h || !n_points) return Fail(MCFD_EINVAL, "null argument");
  if (tag < 0 || tag >= (int32_t)h->bct.fns.size()) return Fail(MCFD_EINVAL, "boundary tag out of range");
  std::vector<double> p = h->bct.Points(tag);
  *n_points = (int64_t)(p.size() / 3);
  if (xyz) std::copy(p.begin(), p.end(), xyz);
  return MCFD_OK;
}

int mcfd_set_boundary_function(mcfd_solver *h, int32_t tag, mcfd_bc_function f, void *user) {
  MCFD_REQUIRE_DEVICE(h);
  if (tag < 0 || tag >= (int32_t)h->bct.fns.size()) return Fail(MCFD_EINVAL, "boundary tag out of range");
  return Guard([&] { h->bct.SetFunction(tag, f, user); });
}

int mcfd_set_boundary_values(mcfd_solver *h, int32_t tag, const double *states) {
  MCFD_REQUIRE_DEVICE(h);
  if (tag < 0 || tag >= (int32_t)h->bct.fns.size()) return Fail(MCFD_EINVAL, "boundary tag out of range");
  if (!states) return Fail(MCFD_EINVAL, "null states");
  return Guard([&] { h->bct.SetValues(tag, states); });
}

int mcfd_limiter_reconstruct(mcfd_solver *h) {
  MCFD_REQUIRE_DEVICE(h);
  if (!h->dg || !h->modal.limiter) return Fail(MCFD_ESTATE, "the handle has no limiter");
  return Guard([&] {
    h->dt_fresh = false;
    h->modal.Limit(h->modal.u0.ptr, h->modal.ct0.ptr, h->stream);
    Check(hipStreamSynchronize(h->stream), "limiter sync");
  });
}

int mcfd_startup_limiter(mcfd_solver *h, int32_t kind) {
  MCFD_REQUIRE_DEVICE(h);
  if (kind != MCFD_STARTUP_AVERAGE && kind != MCFD_STARTUP_MAJORITY) return Fail(MCFD_EINVAL, "unknown start-up limiter");
  if (h->dg) return Fail(MCFD_ENOTSUP, "the start-up limiters act on nodal (FR_GENERAL) solutions");
  auto &g = h->group;
  // average.hpp:13 and majority.hpp:18 leave low degrees alone
  if (kind == MCFD_STARTUP_MAJORITY && g.degree < 2) return MCFD_OK;
  return Guard([&] {
    h->dt_fresh = false;
    const int q = g.q;
    if (kind == MCFD_STARTUP_AVERAGE) {
      const auto &w = g.w1d;
      std::vector<double> wl;  // local weights in node order (integrator/hexahedron.hpp:66-93)
      for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j)
          for (int k = 0; k < q; ++k) wl.push_back(w[i] * w[j] * w[k]);
      mcfd::DeviceArray<double> d_wl, d_det;
      d_wl.Upload(wl, h->stream);
      d_det.Upload(g.NodeDeterminants(), h->stream);
      Check(mcfd::LaunchFrAverage(g.u0.ptr, d_wl.ptr, d_det.ptr, g.n_cells, g.np, h->stream), "average launch");
      Check(hipStreamSynchronize(h->stream), "average sync");
      return;
    }
    mcfd::DeviceArray<double> value;
    mcfd::DeviceArray<int32_t> failed;
    value.Allocate(size_t(g.n_cells) * 5);
    failed.Upload(std::vector<int32_t>{ 0 }, h->stream);
    Check(mcfd::LaunchFrMajority(g.u0.ptr, value.ptr, failed.ptr, g.n_cells, g.np, h->stream), "majority launch");
    int32_t bad = 0;
    Check(hipMemcpyAsync(&bad, failed.ptr, sizeof(bad), hipMemcpyDeviceToHost, h->stream), "majority flag");
    Check(hipStreamSynchronize(h->stream), "majority sync");
    if (bad) throw mcfd::Failure{ MCFD_EINVAL, "majority limiter: a cell component has no value held by more than half of its nodes" };
    Check(mcfd::LaunchFrSetNodes(g.u0.ptr, value.ptr, g.n_cells, g.np, h->stream), "majority set");
    Check(hipStreamSynchronize(h->stream), "majority sync");
  });
}

int mcfd_limiter_flags(mcfd_solver *h, uint8_t *flags, int64_t *n_troubled) {
  MCFD_REQUIRE_DEVICE(h);
  if (!h->dg || !h->modal.limiter) return Fail(MCFD_ESTATE, "the handle has no limiter");
  return Guard([&] {
    std::vector<uint8_t> f(h->modal.n_cells);
    Check(hipMemcpyAsync(f.data(), h->modal.flags.ptr, f.size(), hipMemcpyDeviceToHost, h->stream), "flags");
    Check(hipStreamSynchronize(h->stream), "flags sync");
    if (flags) std::copy(f.begin(), f.end(), flags);
    if (n_troubled) {
      int64_t n = 0;
      for (auto v : f) n += v;
      *n_troubled = n;
    }
  });
}

int mcfd_synchronize(mcfd_solver *h) {
  MCFD_REQUIRE_DEVICE(h);
  return Guard([&] { Check(hipStreamSynchronize(h->stream), "synchronize"); });
}

int mcfd_time_rk3(mcfd_solver *h, double dt, int32_t n_steps, double *total_ms, double *stage_kernel_ms,
                  double *stage_ms) {
  MCFD_REQUIRE_DEVICE(h);
  if (n_steps < 1) return Fail(MCFD_EINVAL, "n_steps must be positive");
  return Guard([&] {
    // One event pair per stage kernel, all recorded on the handle's stream
    // without host synchronisation in between; read back at the end.
    std::vector<hipEvent_t> ev(2 * 3 * (size_t)n_steps + 2, nullptr);
    struct Events {
      std::vector<hipEvent_t> &v;
      ~Events() { for (auto e : v) if (e) (void)hipEventDestroy(e); }
    } guard{ ev };
    for (auto &e : ev) Check(hipEventCreate(&e), "hipEventCreate");
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
      if (stage_ms) stage_ms[(i - 2) / 2] = ms;
    }
    Check(hipEventElapsedTime(&ms, ev[0], ev[1]), "elapsed");
    if (total_ms) *total_ms = ms;
    if (stage_kernel_ms) *stage_kernel_ms = kernel_ms;
  });
}

}  // extern "C"
