// C ABI of the batched form: many independent cases of uneven size and
// polynomial degree advanced by one persistent launch per RK stage, the three
// stages of a step captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "../../../include/minicfd.h"
#include "common.hpp"
#include "fr_group.hpp"

using mcfd::Check;
using mcfd::Fail;
using mcfd::Guard;

namespace {

struct CaseRef {
  int group;
  int64_t first_cell, n_cells;
};

}  // namespace

struct mcfd_batch {
  int32_t n_cases = 0;
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<mcfd::FrHexGroup>> groups;  // one per degree
  std::vector<int32_t> tiles;                              // tile variant per group
  std::vector<CaseRef> cases;
  std::vector<double> dt;                                  // per case
  mcfd::DeviceArray<double> case_dt;
  mcfd::DeviceArray<unsigned long long> scratch;           // per-case time-step reduction
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int riemann = 0;

  bool mixed = false;
  // 0: one launch per degree group and stage; 1: one persistent launch per
  // stage over all groups.  Chosen by mcfd_batch_autotune.
  int mode = 1;
  // Group tables of the three stages (device) and of autotuning trials.
  mcfd::DeviceArray<mcfd::FrHexBatchGroup> stage_tables, trial_table;

  std::vector<mcfd::FrHexBatchGroup> Groups(int stage, int only_group, int tile, int64_t *n_items,
                                            int *max_lds) const {
    std::vector<mcfd::FrHexBatchGroup> out;
    int64_t items = 0;
    *max_lds = 0;
    for (int g = 0; g < (int)groups.size(); ++g) {
      if (only_group >= 0 && g != only_group) continue;
      mcfd::FrHexBatchGroup dst;
      std::memset(&dst, 0, sizeof(dst));
      const auto &grp = *groups[g];
      dst.args = grp.Args(stage, 0.0);
      dst.args.case_dt = case_dt.ptr;
      dst.lt = grp.lt;
      dst.q = grp.q;
      dst.tile = only_group >= 0 && tile >= 0 ? tile : tiles[g];
      dst.first_item = items;
      const int cells = mcfd::FrHexTileCells(grp.q, dst.tile);
      items += (grp.n_cells + cells - 1) / cells;
      *max_lds = std::max(*max_lds, mcfd::FrHexTileLdsBytes(grp.q, dst.tile));
      out.push_back(dst);
    }
    *n_items = items;
    return out;
  }
  void Launch(const mcfd::FrHexBatchGroup *table, int n_groups, int64_t n_items, int max_lds) {
    mcfd::FrHexBatchArgs a;
    a.group = table;
    a.n_groups = n_groups;
    a.max_tile_lds = max_lds;
    a.n_items = n_items;
    int resident = mcfd::FrHexBatchResidentWorkgroups(riemann, mixed, max_lds);
    if (resident < 1) resident = 1024;
    int wg = (int)std::min<int64_t>(n_items, resident);
    Check(mcfd::LaunchFrHexBatch(riemann, mixed, a, wg, stream), "batched stage launch");
  }
  // Uploads the three stage tables and returns their launch parameters.
  void UploadStages(int64_t *n_items, int *max_lds) {
    std::vector<mcfd::FrHexBatchGroup> all;
    for (int stage = 1; stage <= 3; ++stage) {
      auto t = Groups(stage, -1, -1, n_items, max_lds);
      all.insert(all.end(), t.begin(), t.end());
    }
    stage_tables.Upload(all, stream);
  }
  void ReleaseGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  // One stage of every case in the current mode.
  void Stage(int stage, int64_t n_items, int max_lds) {
    if (mode == 1) {
      Launch(stage_tables.ptr + (stage - 1) * groups.size(), (int)groups.size(), n_items, max_lds);
    } else {
      for (auto &g : groups) {
        auto a = g->Args(stage, 0.0);
        a.case_dt = case_dt.ptr;
        Check(mcfd::LaunchFrHexStage(g->q, riemann, a, g->lt, stream), "group stage launch");
      }
    }
  }
  // Captures the three stages of one SSP-RK3 step.
  void Capture() {
    ReleaseGraph();
    int64_t n_items;
    int max_lds;
    UploadStages(&n_items, &max_lds);
    Check(hipStreamSynchronize(stream), "table upload");
    Check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
    for (int stage = 1; stage <= 3; ++stage) Stage(stage, n_items, max_lds);
    Check(hipStreamEndCapture(stream, &graph), "end capture");
    Check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
  }
  ~mcfd_batch() { ReleaseGraph(); }
};

#define MCFD_BATCH_CASE_CHECK(b, c)                                              \
  do {                                                                           \
    if (!(b)) return Fail(MCFD_EINVAL, "null batch");                            \
    if ((c) < 0 || (c) >= (b)->n_cases) return Fail(MCFD_EINVAL, "case index");  \
  } while (0)

extern "C" {

int mcfd_batch_create(int32_t n_cases, const mcfd_mesh_desc *meshes, const int32_t *degrees,
                      const mcfd_scheme_desc *scheme, const mcfd_bc_desc *bcs, const int32_t *bc_offsets,
                      mcfd_batch **out) {
  if (n_cases < 1 || !meshes || !degrees || !scheme || !bc_offsets || !out)
    return Fail(MCFD_EINVAL, "null argument");
  *out = nullptr;
  auto b = std::make_unique<mcfd_batch>();
  int rc = Guard([&] {
    if (scheme->limiter != MCFD_LIMITER_NONE)
      throw mcfd::Failure{ MCFD_ENOTSUP, "the batched form has no limiter" };
    if (scheme->scheme != MCFD_SCHEME_FR_GENERAL) throw mcfd::Failure{ MCFD_ENOTSUP, "batch implements FR_GENERAL" };
    b->n_cases = n_cases;
    b->riemann = scheme->riemann;
    b->mixed = scheme->precision == MCFD_PRECISION_MIXED;
    const int n_bcs = bc_offsets[n_cases];
    if (n_bcs < 0 || (n_bcs > 0 && !bcs)) throw std::invalid_argument("bad BC table");
    std::vector<int32_t> kinds(n_bcs);
    std::vector<double> states(5 * n_bcs);
    for (int i = 0; i < n_bcs; ++i) {
      kinds[i] = bcs[i].kind;
      for (int k = 0; k < 5; ++k) states[5 * i + k] = bcs[i].state[k];
    }
    std::map<int, std::vector<int>> by_degree;
    for (int c = 0; c < n_cases; ++c) {
      if (degrees[c] < 1 || degrees[c] > 3) throw std::invalid_argument("batched degrees must be 1..3");
      if (meshes[c].cell_type != MCFD_HEXA_8) throw std::invalid_argument("batched cases must be hexahedral");
      if (bc_offsets[c] > bc_offsets[c + 1]) throw std::invalid_argument("bc_offsets must be non-decreasing");
      by_degree[degrees[c]].push_back(c);
    }
    if ((int)by_degree.size() > mcfd::kMaxBatchGroups) throw std::invalid_argument("too many degree groups");
    b->cases.resize(n_cases);
    for (auto &[deg, list] : by_degree) {
      // merged mesh: node, cell and boundary-tag offsets per case
      std::vector<double> coords;
      std::vector<int64_t> cells, bnodes;
      std::vector<int32_t> btags, cell_case;
      int64_t node_off = 0, cell_off = 0;
      const int gi = (int)b->groups.size();
      for (int c : list) {
        const auto &m = meshes[c];
        for (int d = 0; d < 3; ++d)
          if (m.period[d] != 0) throw mcfd::Failure{ MCFD_ENOTSUP, "periodic cases cannot be batched" };
        coords.insert(coords.end(), m.coords, m.coords + 3 * m.n_nodes);
        for (int64_t i = 0; i < 8 * m.n_cells; ++i) cells.push_back(m.cell_nodes[i] + node_off);
        for (int64_t i = 0; i < 4 * m.n_bfaces; ++i) bnodes.push_back(m.bface_nodes[i] + node_off);
        for (int64_t i = 0; i < m.n_bfaces; ++i) {
          int32_t t = m.bface_tags[i];
          if (t < 0 || bc_offsets[c] + t >= bc_offsets[c + 1]) throw std::invalid_argument("boundary tag without a BC entry");
          btags.push_back(bc_offsets[c] + t);
        }
        cell_case.insert(cell_case.end(), m.n_cells, c);
        b->cases[c] = CaseRef{ gi, cell_off, m.n_cells };
        node_off += m.n_nodes;
        cell_off += m.n_cells;
      }
      mcfd_mesh_desc merged{};
      merged.cell_type = MCFD_HEXA_8;
      merged.n_nodes = node_off;
      merged.coords = coords.data();
      merged.n_cells = cell_off;
      merged.cell_nodes = cells.data();
      merged.n_bfaces = (int64_t)btags.size();
      merged.bface_nodes = bnodes.data();
      merged.bface_tags = btags.data();
      mcfd_scheme_desc s = *scheme;
      s.degree = deg;
      auto grp = std::make_unique<mcfd::FrHexGroup>();
      grp->Build(merged, s, kinds, states);
      if (!grp->diag_metric) throw mcfd::Failure{ MCFD_ENOTSUP, "batched cases must be axis-aligned" };
      grp->h_cell_case = std::move(cell_case);
      b->groups.push_back(std::move(grp));
      b->tiles.push_back(0);
    }
    Check(hipStreamCreate(&b->stream), "hipStreamCreate");
    for (auto &g : b->groups) g->Upload(b->stream);
    b->dt.assign(n_cases, 0.0);
    b->case_dt.Upload(b->dt, b->stream);
    b->scratch.Allocate(n_cases);
    Check(hipStreamSynchronize(b->stream), "create sync");
    b->Capture();
  });
  if (rc == MCFD_OK) *out = b.release();
  return rc;
}

int mcfd_batch_destroy(mcfd_batch *b) {
  if (!b) return MCFD_OK;
  hipStream_t s = b->stream;
  if (s) (void)hipStreamSynchronize(s);
  delete b;
  if (s) (void)hipStreamDestroy(s);
  return MCFD_OK;
}

int mcfd_batch_case_sizes(const mcfd_batch *b, int32_t c, int64_t *n_cells, int32_t *n_per_cell) {
  MCFD_BATCH_CASE_CHECK(b, c);
  if (n_cells) *n_cells = b->cases[c].n_cells;
  if (n_per_cell) *n_per_cell = b->groups[b->cases[c].group]->np;
  return MCFD_OK;
}

int mcfd_batch_sample_points(const mcfd_batch *b, int32_t c, double *xyz) {
  MCFD_BATCH_CASE_CHECK(b, c);
  if (!xyz) return Fail(MCFD_EINVAL, "null output");
  const auto &ref = b->cases[c];
  const auto &g = *b->groups[ref.group];
  std::vector<double> all(g.n_cells * g.np * 3);
  g.SamplePoints(all.data());
  std::memcpy(xyz, all.data() + ref.first_cell * g.np * 3, ref.n_cells * g.np * 3 * sizeof(double));
  return MCFD_OK;
}

int mcfd_batch_set_solution_column(mcfd_batch *b, int32_t c, const double *column) {
  MCFD_BATCH_CASE_CHECK(b, c);
  if (!column) return Fail(MCFD_EINVAL, "null column");
  return Guard([&] {
    const auto &ref = b->cases[c];
    const auto &g = *b->groups[ref.group];
    Check(hipMemcpyAsync(g.u0.ptr + ref.first_cell * g.np * 5, column, ref.n_cells * g.np * 5 * sizeof(double),
                         hipMemcpyHostToDevice, b->stream), "set solution");
    Check(hipStreamSynchronize(b->stream), "sync");
  });
}

int mcfd_batch_get_solution_column(mcfd_batch *b, int32_t c, double *column) {
  MCFD_BATCH_CASE_CHECK(b, c);
  if (!column) return Fail(MCFD_EINVAL, "null column");
  return Guard([&] {
    const auto &ref = b->cases[c];
    const auto &g = *b->groups[ref.group];
    Check(hipMemcpyAsync(column, g.u0.ptr + ref.first_cell * g.np * 5, ref.n_cells * g.np * 5 * sizeof(double),
                         hipMemcpyDeviceToHost, b->stream), "get solution");
    Check(hipStreamSynchronize(b->stream), "sync");
  });
}

int mcfd_batch_get_time_steps(mcfd_batch *b, double cfl, double *dt) {
  if (!b || !dt) return Fail(MCFD_EINVAL, "null argument");
  return Guard([&] {
    std::vector<unsigned long long> bits(b->n_cases);
    double init = 1e300;
    std::memcpy(&bits[0], &init, sizeof(init));
    std::fill(bits.begin(), bits.end(), bits[0]);
    Check(hipMemcpyAsync(b->scratch.ptr, bits.data(), bits.size() * sizeof(bits[0]), hipMemcpyHostToDevice,
                         b->stream), "dt init");
    for (int c = 0; c < b->n_cases; ++c) {
      const auto &ref = b->cases[c];
      const auto &g = *b->groups[ref.group];
      Check(mcfd::LaunchFrHexTimeStep(g.q, g.u0.ptr + ref.first_cell * g.np * 5, g.geo_id.ptr + ref.first_cell,
                                      g.geo.ptr, ref.n_cells, b->scratch.ptr + c, b->stream), "time step launch");
    }
    Check(hipMemcpyAsync(bits.data(), b->scratch.ptr, bits.size() * sizeof(bits[0]), hipMemcpyDeviceToHost,
                         b->stream), "dt read");
    Check(hipStreamSynchronize(b->stream), "dt sync");
    for (int c = 0; c < b->n_cases; ++c) {
      double m;
      std::memcpy(&m, &bits[c], sizeof(m));
      dt[c] = cfl * m;
    }
  });
}

int mcfd_batch_set_time_steps(mcfd_batch *b, const double *dt) {
  if (!b || !dt) return Fail(MCFD_EINVAL, "null argument");
  return Guard([&] {
    b->dt.assign(dt, dt + b->n_cases);
    Check(hipMemcpyAsync(b->case_dt.ptr, b->dt.data(), b->dt.size() * sizeof(double), hipMemcpyHostToDevice,
                         b->stream), "set dt");
    Check(hipStreamSynchronize(b->stream), "sync");
  });
}

int mcfd_batch_autotune(mcfd_batch *b, int32_t repeats, int32_t *tiles, int32_t *mode) {
  if (!b || repeats < 1) return Fail(MCFD_EINVAL, "bad argument");
  return Guard([&] {
    hipEvent_t e0, e1;
    Check(hipEventCreate(&e0), "event");
    Check(hipEventCreate(&e1), "event");
    struct Ev { hipEvent_t a, b; ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } guard{ e0, e1 };
    auto time = [&](auto &&launch) {
      launch();  // warm-up
      Check(hipEventRecord(e0, b->stream), "event");
      for (int r = 0; r < repeats; ++r) launch();
      Check(hipEventRecord(e1, b->stream), "event");
      Check(hipEventSynchronize(e1), "event sync");
      float ms;
      Check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      return ms;
    };
    // tile variant per degree group for the persistent form (stage 1 writes
    // only the U1 scratch state)
    for (int g = 0; g < (int)b->groups.size(); ++g) {
      float best = 1e30f;
      int best_tile = 0;
      for (int tile = 0; tile < 3; ++tile) {
        int64_t n_items;
        int max_lds;
        auto table = b->Groups(1, g, tile, &n_items, &max_lds);
        b->trial_table.Upload(table, b->stream);
        float ms = time([&] { b->Launch(b->trial_table.ptr, 1, n_items, max_lds); });
        if (ms < best) { best = ms; best_tile = tile; }
      }
      b->tiles[g] = best_tile;
      if (tiles) tiles[g] = b->groups[g]->q * 16 + best_tile;
    }
    // persistent single launch vs one launch per degree group
    int64_t n_items;
    int max_lds;
    b->UploadStages(&n_items, &max_lds);
    float ms_mode[2];
    for (int m = 0; m < 2; ++m) {
      b->mode = m;
      ms_mode[m] = time([&] { b->Stage(1, n_items, max_lds); });
    }
    b->mode = ms_mode[1] <= ms_mode[0] ? 1 : 0;
    if (mode) *mode = b->mode;
    b->Capture();
  });
}

int mcfd_batch_set_mode(mcfd_batch *b, int32_t mode, const int32_t *tiles) {
  if (!b) return Fail(MCFD_EINVAL, "null batch");
  if (mode != 0 && mode != 1) return Fail(MCFD_EINVAL, "mode must be 0 (per degree group) or 1 (persistent)");
  if (tiles)
    for (size_t g = 0; g < b->groups.size(); ++g)
      if (tiles[g] < 0 || tiles[g] > 2) return Fail(MCFD_EINVAL, "tile variant must be 0, 1 or 2");
  return Guard([&] {
    b->mode = mode;
    if (tiles) std::copy(tiles, tiles + b->groups.size(), b->tiles.begin());
    b->Capture();
  });
}

int mcfd_batch_rk3_steps(mcfd_batch *b, int32_t n_steps) {
  if (!b || n_steps < 0) return Fail(MCFD_EINVAL, "bad argument");
  return Guard([&] {
    for (int s = 0; s < n_steps; ++s) Check(hipGraphLaunch(b->exec, b->stream), "graph launch");
  });
}

int mcfd_batch_synchronize(mcfd_batch *b) {
  if (!b) return Fail(MCFD_EINVAL, "null batch");
  return Guard([&] { Check(hipStreamSynchronize(b->stream), "synchronize"); });
}

int mcfd_batch_time_rk3(mcfd_batch *b, int32_t n_steps, double *total_ms) {
  if (!b || n_steps < 1) return Fail(MCFD_EINVAL, "bad argument");
  return Guard([&] {
    hipEvent_t e0, e1;
    Check(hipEventCreate(&e0), "event");
    Check(hipEventCreate(&e1), "event");
    struct Ev { hipEvent_t a, b; ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } guard{ e0, e1 };
    Check(hipEventRecord(e0, b->stream), "event");
    for (int s = 0; s < n_steps; ++s) Check(hipGraphLaunch(b->exec, b->stream), "graph launch");
    Check(hipEventRecord(e1, b->stream), "event");
    Check(hipEventSynchronize(e1), "event sync");
    float ms;
    Check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    if (total_ms) *total_ms = ms;
  });
}

}  // extern "C"
