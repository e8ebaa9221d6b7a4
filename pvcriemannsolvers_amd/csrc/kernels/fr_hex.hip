// Nodal flux reconstruction on affine hexahedra for CDNA4 (gfx950).
//
// One pass = residual of spatial/fr/general.hpp (volume term :271-288, face
// correction :296-315,409-434, boundary corrections :435-640, det(J) scaling
// :249-268) fused with one SSP-RK3 stage of temporal/rk.hpp:31-49.
//
// Work decomposition: a 256-thread workgroup owns C consecutive cells.  Each
// pass
//   0. loads the cells' states (one thread per solution point) and keeps a
//      skewed structure-of-arrays copy in LDS;
//   1. the three columns of F~ = F J* go through LDS (Q >= 3: all three in
//      one round behind one barrier, while the first face task's neighbour
//      line is already in flight; Q = 2: one column per round), every
//      solution point accumulates its sum-factorised Q-point derivatives
//      and every flux point extrapolates its direction's column to the face;
//   2. per flux point of the cells' 6 faces: own trace (LDS), neighbour trace
//      (the neighbour's line of Q points, gathered through L2), upwind flux
//      and flux gap.  Each interior face is solved by both of its cells with
//      bit-identical inputs, so the scheme stays conservative without
//      atomics or a separate face pass;
//   3. per solution point: lifted gaps, 1/det(J), RK stage in the
//      reference's operation order, store.
// The same body serves the single-mesh launch and the persistent batched
// launch that walks a work list spanning several polynomial degrees.
#include "fr_hex.hpp"

#include <algorithm>

#include "euler.hpp"
#include "fr_common.hpp"

namespace mcfd {

namespace {

using fr::Extrapolate;
using fr::FaceLineIndex;
using fr::LineIndex;
using fr::LocalFluxColumn;
using fr::LogicalBlock;
using fr::SkewedLineIndex;

template <int Q, int Cells = 256 / (Q * Q * Q), int Threads = 256, bool OneRound = false>
struct Shape {
  static constexpr int kPoints = Q * Q * Q;        // solution points per cell
  static constexpr int kFacePoints = Q * Q;        // flux points per face
  static constexpr int kFaceTasks = 6 * Q * Q;     // flux points per cell
  static constexpr int kCells = Cells;             // cells per workgroup
  static constexpr int kThreads = Threads;         // workgroup size
  static_assert(Cells * Q * Q * Q <= Threads, "one thread per solution point");
  static_assert(Threads % 64 == 0 && Threads <= 1024, "whole wavefronts");
  // Skewed LDS position of solution point (i, j, k): i * kSI + j * kSJ + k.
  // With kSJ = Q + 1 and kSI = Q (Q + 1) the Q x Q points of any line family
  // read by one wavefront fall in distinct LDS banks (Q = 4: 20 i + 5 j + k).
  static constexpr int kSJ = Q + 1;
  static constexpr int kSI = Q * (Q + 1);
  static constexpr int kSpan = (Q - 1) * (kSI + kSJ + 1) + 1;
  static constexpr int kPad = kSpan | 1;           // odd row stride per component
  static constexpr int kFacePad = kFaceTasks + 1;
  // per cell: the F~ column(s) in flight (one direction, or all three with
  // OneRound), later the flux gaps
  static constexpr int kColumns = (OneRound ? 15 : 5) * kPad;
  static constexpr int kScratch = kColumns > 5 * kFacePad ? kColumns : 5 * kFacePad;
  static constexpr int kTasks = (kCells * kFaceTasks + Threads - 1) / Threads;  // per thread
  static_assert(kCells >= 1, "at most one workgroup of solution points per cell");
  // states, F~ columns / flux gaps, mixed reference values, face area scales
  static constexpr int kLdsDoubles = kCells * 5 * kPad + kCells * kScratch + kCells * 15 + kCells * 6;
};

// kMixed: the F~ columns cross LDS as FP32 deviations from the cell's first
// solution point and are contracted with FP32 FMAs (each row of the
// derivative matrix sums to zero, so the FP64 reference value drops out of
// the divergence); each direction's FP32 partial is accumulated in FP64.
// States, pressure recovery, traces, the Riemann solve, the flux gaps and
// the RK combination stay FP64.
//
// kOneRound: all three F~ columns cross LDS in one round (three times the
// column space; one barrier instead of six, two in the mixed form); the
// arithmetic and its order are those of the three-round form, so both give
// bitwise the same results.  kPrefetch (with kOneRound): the neighbour lines
// of that many face-task slots are loaded before the passes.
//
// kCurved (trilinear cells): J* and 1/det(J) are read per solution point, the
// flux-point scale per cell flux point and the frame per face flux point (in
// the holder's in-face order) instead of per cell and per face.
template <int Q, int Cells, int Threads, int kRiemann, bool kDiag, bool kMixed, bool kOneRound = false,
          int kPrefetch = 0, bool kCurved = false>
__device__ __forceinline__ void FrHexStageBody(const FrHexArgs &a, const LineTables &lt, int64_t block,
                                               double *lds) {
  using S = Shape<Q, Cells, Threads, kOneRound>;
  constexpr int NP = S::kPoints, NFP = S::kFacePoints, C = S::kCells;
  constexpr int PAD = S::kPad, FPAD = S::kFacePad, T = S::kTasks;
  double *s_u = lds;                          // states, SoA per cell
  double *s_x = lds + C * 5 * PAD;            // F~ column per pass, then flux gaps
  double *s_ref = s_x + C * S::kScratch;      // mixed: FP64 column values at node 0, [d][cell][5]
  double *s_scale = s_ref + C * 15;           // signed face area scales of each cell
  float *s_x32 = reinterpret_cast<float *>(s_x);

  const int tid = threadIdx.x;
  const int64_t c0 = block * C;
  if (c0 >= a.n_cells) return;
  const int nc = (int)min<int64_t>(C, a.n_cells - c0);
  const int64_t base = c0 * NP * 5;

  // ---- node phase: load, primitive, SoA copy of the state
  const bool node_active = tid < nc * NP;
  const int ncell = tid / NP, node = tid - (tid / NP) * NP;
  const int ni = node / (Q * Q), nj = (node / Q) % Q, nk = node % Q;
  const int spos = ni * S::kSI + nj * S::kSJ + nk;  // skewed LDS position
  State cu, pu;
  double u0v[5] = { 0, 0, 0, 0, 0 };
  const double *js = nullptr;
  double inv_det = 0;
  if (node_active) {
#pragma unroll
    for (int k = 0; k < 5; ++k) cu.v[k] = a.us[base + tid * 5 + k];
    pu = FastPrimitive(cu);
    if (a.stage >= 2) {  // U0 of the RK combination, fetched early
#pragma unroll
      for (int k = 0; k < 5; ++k) u0v[k] = a.u0[base + tid * 5 + k];
    }
    if constexpr (kCurved) {
      js = a.geo_node + ((c0 + ncell) * NP + node) * kNodeGeoStride;
    } else {
      js = a.geo + (int64_t)a.geo_id[c0 + ncell] * kGeoStride;
      if (node < 6) s_scale[ncell * 6 + node] = js[10 + node];
    }
    inv_det = js[9];
#pragma unroll
    for (int k = 0; k < 5; ++k) s_u[(ncell * 5 + k) * PAD + spos] = cu.v[k];
  }

  // ---- face task bookkeeping (flux points of the workgroup's cells)
  // Tasks are numbered direction-major: w = ((d * C + cell) * 2 + side) * Q^2
  // + a * Q + b covers the two faces normal to reference direction d.  A
  // wavefront then holds flux points of one direction (of at most two where
  // C * 2 Q^2 is not a multiple of 64), so the extrapolation of pass d runs
  // only in the wavefronts that own direction d instead of in all of them.
  constexpr int PER_DIR = C * 2 * NFP;
  int t_cell[T], t_f[T], t_a[T], t_b[T];
  bool t_on[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int w = tid + Threads * t;
    const int d = w / PER_DIR, rd = w - d * PER_DIR;
    const int cell = rd / (2 * NFP), rc = rd - cell * (2 * NFP);
    const int side = rc / NFP, fp = rc - side * NFP;
    t_on[t] = d < 3 && cell < nc;
    t_cell[t] = cell;
    // local face of (direction, side): x {4, 2}, y {1, 3}, z {0, 5}
    t_f[t] = d == 0 ? (side ? 2 : 4) : (d == 1 ? (side ? 3 : 1) : (side ? 5 : 0));
    t_a[t] = fp / Q;
    t_b[t] = fp - (fp / Q) * Q;
  }
  // neighbour and face-info words of the face tasks, fetched before the
  // volume passes so that the face phase starts its gathers at once
  int32_t t_nb[T], t_info[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    t_nb[t] = t_on[t] ? a.nbr[(c0 + t_cell[t]) * 6 + t_f[t]] : 0;
    t_info[t] = t_on[t] ? a.finfo[(c0 + t_cell[t]) * 6 + t_f[t]] : 0;
  }
  double fn[T][5];

  // mixed: node 0 of each cell publishes its three F~ columns, the FP64
  // references of the FP32 deviations; the first pass's barrier orders them
  if constexpr (kMixed) {
    if (node_active && node == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double col[5];
        LocalFluxColumn<kDiag>(cu, pu, js, d, col);
#pragma unroll
        for (int k = 0; k < 5; ++k) s_ref[(d * C + ncell) * 5 + k] = col[k];
      }
    }
  }

  // neighbour line of face task t (Q nodes of the neighbour cell along the
  // face normal, in the orientation of this cell's flux point)
  auto gather = [&](int t, double (*raw)[5]) {
    const int32_t nb = t_nb[t], info = t_info[t];
    const int ia = t_a[t], ib = t_b[t];
    const int nf = info & 7, orient = (info >> 3) & 7;
    const int s = (orient & 1) ? ib : ia, tt = (orient & 1) ? ia : ib;
    const int na = (orient & 2) ? Q - 1 - s : s;
    const int nbb = (orient & 4) ? Q - 1 - tt : tt;
    const LineIndex ln = FaceLineIndex(Q, nf);
    const double *src = a.us + ((int64_t)nb * NP + na * ln.a_mul + nbb * ln.b_mul) * 5;
#pragma unroll
    for (int m = 0; m < Q; ++m)
#pragma unroll
      for (int k = 0; k < 5; ++k) raw[m][k] = src[m * ln.stride * 5 + k];
  };
  constexpr int PF = kPrefetch < T ? kPrefetch : T;
  double pre[PF > 0 ? PF : 1][Q][5];

  // ---- three passes: one column of F~ at a time through LDS (or all three
  // written at once with kOneRound)
  double div[5] = { 0, 0, 0, 0, 0 };
  if constexpr (kOneRound) {
    if constexpr (kMixed) __syncthreads();  // the FP64 references of the deviations
    if (node_active) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double col[5];
        LocalFluxColumn<kDiag>(cu, pu, js, d, col);
        if constexpr (kMixed) {
#pragma unroll
          for (int k = 0; k < 5; ++k)
            s_x32[ncell * S::kScratch + (5 * d + k) * PAD + spos] = (float)(col[k] - s_ref[(d * C + ncell) * 5 + k]);
        } else {
#pragma unroll
          for (int k = 0; k < 5; ++k) s_x[ncell * S::kScratch + (5 * d + k) * PAD + spos] = col[k];
        }
      }
    }
    __syncthreads();
    // the first face tasks' neighbour lines, in flight during the passes
#pragma unroll
    for (int t = 0; t < PF; ++t)
      if (t_on[t] && t_nb[t] >= 0) gather(t, pre[t]);
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int cb = kOneRound ? 5 * d * PAD : 0;  // column d within the cell's scratch
    if constexpr (!kOneRound) {
      __syncthreads();
      double col[5];
      if (node_active) {
        LocalFluxColumn<kDiag>(cu, pu, js, d, col);
        if constexpr (kMixed) {
#pragma unroll
          for (int k = 0; k < 5; ++k)
            s_x32[ncell * S::kScratch + k * PAD + spos] = (float)(col[k] - s_ref[(d * C + ncell) * 5 + k]);
        } else {
#pragma unroll
          for (int k = 0; k < 5; ++k) s_x[ncell * S::kScratch + k * PAD + spos] = col[k];
        }
      }
      __syncthreads();
    }
    if (node_active) {
      const int id = d == 0 ? ni : (d == 1 ? nj : nk);
      const int stride = d == 0 ? S::kSI : (d == 1 ? S::kSJ : 1);
      const int line0 = spos - id * stride;
      if constexpr (kMixed) {
        const float *sx = s_x32 + ncell * S::kScratch + cb;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          float acc = 0.0f;
#pragma unroll
          for (int m = 0; m < Q; ++m) acc = __fmaf_rn((float)lt.d[id][m], sx[k * PAD + line0 + m * stride], acc);
          div[k] += (double)acc;
        }
      } else {
        const double *sx = s_x + ncell * S::kScratch + cb;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          double acc = div[k];
#pragma unroll
          for (int m = 0; m < Q; ++m) acc = __fma_rn(lt.d[id][m], sx[k * PAD + line0 + m * stride], acc);
          div[k] = acc;
        }
      }
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const LineIndex li = SkewedLineIndex<Q>(t_f[t]);
      if (t_on[t] && li.dir == d) {
        const int l0 = t_a[t] * li.a_mul + t_b[t] * li.b_mul;
        const double *lw = lt.lend[li.side];
        if constexpr (kMixed) {
          const float *sx = s_x32 + t_cell[t] * S::kScratch + cb;
#pragma unroll
          for (int k = 0; k < 5; ++k) {
            float acc = (float)lw[0] * sx[k * PAD + l0];
#pragma unroll
            for (int m = 1; m < Q; ++m) acc = __fmaf_rn((float)lw[m], sx[k * PAD + l0 + m * li.stride], acc);
            fn[t][k] = s_ref[(d * C + t_cell[t]) * 5 + k] + (double)acc;
          }
        } else {
          const double *sx = s_x + t_cell[t] * S::kScratch + cb;
#pragma unroll
          for (int k = 0; k < 5; ++k) {
            double vals[Q];
#pragma unroll
            for (int m = 0; m < Q; ++m) vals[m] = sx[k * PAD + l0 + m * li.stride];
            fn[t][k] = Extrapolate<Q>(lw, vals);
          }
        }
      }
    }
  }

  // ---- flux points: traces, upwind flux, flux gap
  double gap[T][5];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    if (!t_on[t]) continue;
    const int f = t_f[t], ia = t_a[t], ib = t_b[t];
    const LineIndex li = SkewedLineIndex<Q>(f);
    const double *lw = lt.lend[li.side];
    State own;
    {
      const int l0 = ia * li.a_mul + ib * li.b_mul;
      const double *su = s_u + t_cell[t] * 5 * PAD;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        double vals[Q];
#pragma unroll
        for (int m = 0; m < Q; ++m) vals[m] = su[k * PAD + l0 + m * li.stride];
        own.v[k] = Extrapolate<Q>(lw, vals);
      }
    }
    const int32_t nb = t_nb[t];
    const int32_t info = t_info[t];
    const double *fr;
    double scale;
    if constexpr (kCurved) {
      // the face point in the holder's in-face order: this cell's own (a, b),
      // or the neighbour's where the neighbour holds the face
      int cp = ia * Q + ib;
      if (!((info >> 6) & 1)) {
        const int orient = (info >> 3) & 7;
        const int s = (orient & 1) ? ib : ia, tt = (orient & 1) ? ia : ib;
        cp = ((orient & 2) ? Q - 1 - s : s) * Q + ((orient & 4) ? Q - 1 - tt : tt);
      }
      fr = a.frames + ((int64_t)(info >> 7) * NFP + cp) * kFrameStride;
      scale = a.fscale[((c0 + t_cell[t]) * 6 + f) * NFP + ia * Q + ib];
    } else {
      fr = a.frames + (int64_t)(info >> 7) * kFrameStride;
      scale = s_scale[t_cell[t] * 6 + f];
    }
    FaceFrame frame;
#pragma unroll
    for (int d = 0; d < 3; ++d) { frame.n[d] = fr[d]; frame.t[d] = fr[3 + d]; }
    frame.b[0] = frame.n[1] * frame.t[2] - frame.n[2] * frame.t[1];
    frame.b[1] = frame.n[2] * frame.t[0] - frame.n[0] * frame.t[2];
    frame.b[2] = frame.n[0] * frame.t[1] - frame.n[1] * frame.t[0];
    State fu;
    double sgn = 1.0;
    // the exact solver's state of this flux point, one copy per side: both
    // sides solve the same problem with the same inputs every pass, so the
    // two copies stay equal
    const int64_t star_slot = ((c0 + t_cell[t]) * 6 + f) * NFP + ia * Q + ib;
    bool star = kRiemann == kExact ? a.exact_star[star_slot] != 0 : false;
    if (nb >= 0) {
      const bool holder = (info >> 6) & 1;
      const double *nw = lt.lend[FaceLineIndex(Q, info & 7).side];
      double raw[Q][5];
      if (kOneRound && t < PF) {
#pragma unroll
        for (int m = 0; m < Q; ++m)
#pragma unroll
          for (int k = 0; k < 5; ++k) raw[m][k] = pre[t < PF ? t : 0][m][k];
      } else {
        gather(t, raw);
      }
      State other;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        double vals[Q];
#pragma unroll
        for (int m = 0; m < Q; ++m) vals[m] = raw[m][k];
        other.v[k] = Extrapolate<Q>(nw, vals);
      }
      fu = FaceFlux<kRiemann>(frame, holder ? own : other, holder ? other : own, star);
      sgn = holder ? 1.0 : -1.0;
    } else {
      const int tag = -nb - 1;
      const int kind = a.bc_kind[tag];
      // the tag's constant state, or this point's value of the tag's
      // boundary function at the stage time
      const double *gs = a.bc_point_state
          ? a.bc_point_state + ((int64_t)a.bface_of[(c0 + t_cell[t]) * 6 + f] * NFP + ia * Q + ib) * 5
          : a.bc_state + tag * 5;
      State given;
#pragma unroll
      for (int k = 0; k < 5; ++k) given.v[k] = gs[k];
      if (kind == kBcInviscidWall) {
        fu = WallFlux(frame, own);
      } else if (kind == kBcSupersonicOutlet) {
        fu = StateFlux(frame, own);
      } else if (kind == kBcSupersonicInlet) {
        fu = StateFlux(frame, given);
      } else if (kind == kBcSubsonicInlet) {
        fu = SubsonicInletFlux(frame, own, given);
      } else if (kind == kBcSubsonicOutlet) {
        fu = SubsonicOutletFlux(frame, own, given);
      } else {
        fu = FaceFlux<kRiemann>(frame, own, given, star);
      }
    }
    if constexpr (kRiemann == kExact) a.exact_star[star_slot] = star ? 1 : 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) gap[t][k] = fu.v[k] * (sgn * scale) - fn[t][k];
  }
  __syncthreads();  // every read of the F~ columns is done
#pragma unroll
  for (int t = 0; t < T; ++t) {
    if (!t_on[t]) continue;
    const int r = t_f[t] * NFP + t_a[t] * Q + t_b[t];  // face-major gap slot of the node phase
#pragma unroll
    for (int k = 0; k < 5; ++k) s_x[t_cell[t] * S::kScratch + k * FPAD + r] = gap[t][k];
  }
  __syncthreads();

  // ---- node phase: lifted gaps, 1/det(J), RK stage, store
  double speed = 0.0;  // |u| + c of the new state, for a fused time step
  if (node_active) {
    const double *sg = s_x + ncell * S::kScratch;
    const double *su = s_u + ncell * 5 * PAD;
    const double dt = a.case_dt ? a.case_dt[a.cell_case[c0 + ncell]] : a.dt;
    double outv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double *g = sg + k * FPAD;
      const double corr = fr::LiftedGaps<Q>(g, lt, ni, nj, nk);
      const double res = (-div[k] - corr) * inv_det;
      outv[k] = fr::RkOutput(a.stage, res, su[k * PAD + spos], u0v[k], dt);
    }
    // one basic block of stores: merged into 16-byte stores of the node record
#pragma unroll
    for (int k = 0; k < 5; ++k) a.out[base + tid * 5 + k] = outv[k];
    if (a.dt_min_bits) {
      State nw;
#pragma unroll
      for (int k = 0; k < 5; ++k) nw.v[k] = outv[k];
      speed = MaxSpeed(nw);
    }
  }
  if (a.dt_min_bits) {
    // fem.hpp:412-435 on the state just written, as FrHexTimeStepKernel:
    // min over cells of length / max(1e-5, max over nodes of |u| + c)
    __syncthreads();  // every read of the LDS states and gaps is done
    double *s_speed = s_u, *s_cell = s_x;
    if (node_active) s_speed[tid] = speed;
    __syncthreads();
    if (tid < nc) {
      double vmax = 1e-5;
      for (int p = 0; p < NP; ++p) vmax = fmax(vmax, s_speed[tid * NP + p]);
      s_cell[tid] = a.geo[(int64_t)a.geo_id[c0 + tid] * kGeoStride + 16] / vmax;
    }
    __syncthreads();
    if (tid == 0) {
      double local = 1e300;
      for (int c = 0; c < nc; ++c) local = fmin(local, s_cell[c]);
      atomicMin(a.dt_min_bits, (unsigned long long)__double_as_longlong(local));
    }
  }
}

template <int Q, int Cells, int Threads, int kRiemann, bool kDiag, bool kMixed, bool kOneRound, int kPrefetch,
          bool kCurved = false>
__global__ __launch_bounds__(Threads) void FrHexStageKernel(FrHexArgs a, LineTables lt) {
  __shared__ double lds[Shape<Q, Cells, Threads, kOneRound>::kLdsDoubles];
  // The line tables are indexed by lane-varying node indices: read from the
  // kernel arguments they would be vector loads on the critical path of every
  // pass, so they are staged in LDS once (the body's first barrier orders it).
  __shared__ LineTables s_lt;
  constexpr int kTableDoubles = sizeof(LineTables) / sizeof(double);
  if (threadIdx.x < kTableDoubles)
    reinterpret_cast<double *>(&s_lt)[threadIdx.x] = reinterpret_cast<const double *>(&lt)[threadIdx.x];
  int64_t block = LogicalBlock(a.xcd_remap);
  if (a.tile_order) block = a.tile_order[block];
  FrHexStageBody<Q, Cells, Threads, kRiemann, kDiag, kMixed, kOneRound, kPrefetch, kCurved>(a, s_lt, block, lds);
}

// Single-mesh launch shape per Q: one thread per solution point.  Q = 4 and
// 5 run on 128-thread workgroups (2 cells and 1 cell): the same 3 waves per
// SIMD as 256-thread ones, with each barrier waiting on two wavefronts
// instead of four (Q = 4: 0.93 against 0.96 ms per stage at 64^3; Q = 3 as
// 4 cells on 128 threads leaves 20 threads idle and measured no better).  At Q = 4
// the face tasks (96 per cell) fill two slots in half of the wavefronts; 96
// threads per cell (one task each) measured slower on MI355X, since the extra
// wavefronts idle through the volume passes and occupy wave slots that other
// workgroups would fill.
template <int Q> struct StageTile {
  static constexpr int kCells = Q == 4 ? 2 : (Q == 5 ? 1 : Shape<Q>::kCells);
  static constexpr int kThreads = (Q == 4 || Q == 5) ? 128 : 256;
  static constexpr bool kOneRound = Q >= 3;  // at Q = 2 the column space would cost a workgroup per CU
  // face tasks whose neighbour lines are issued before the passes: one, as
  // the registers of the three-waves-per-SIMD budget allow
  static constexpr int kPrefetch = 1;
};

// Tile variants per Q (cells per workgroup) that the batched launch and its
// autotuner choose from.
template <int Q, int V> struct TileOf;
template <> struct TileOf<2, 0> { static constexpr int kCells = 32; };
template <> struct TileOf<2, 1> { static constexpr int kCells = 16; };
template <> struct TileOf<2, 2> { static constexpr int kCells = 8; };
template <> struct TileOf<3, 0> { static constexpr int kCells = 9; };
template <> struct TileOf<3, 1> { static constexpr int kCells = 4; };
template <> struct TileOf<3, 2> { static constexpr int kCells = 2; };
template <> struct TileOf<4, 0> { static constexpr int kCells = 4; };
template <> struct TileOf<4, 1> { static constexpr int kCells = 2; };
template <> struct TileOf<4, 2> { static constexpr int kCells = 1; };

template <int Q, int V>
constexpr int TileLds() { return Shape<Q, TileOf<Q, V>::kCells, 256, (Q >= 3)>::kLdsDoubles; }

// One out-of-line body per (Q, tile): each is register-allocated on its own
// within the kernel's 3-waves-per-SIMD budget, where nine inlined bodies
// would share one allocation and spill.  Q >= 3 bodies pass all three F~
// columns through LDS in one round, as the single-mesh launch does.
template <int Q, int V, int kRiemann, bool kMixed>
__device__ __noinline__ void BatchBody(const FrHexBatchGroup *g, int64_t block, double *lds) {
  FrHexStageBody<Q, TileOf<Q, V>::kCells, 256, kRiemann, true, kMixed, (Q >= 3), 1>(g->args, g->lt, block, lds);
}

template <int Q, int kRiemann, bool kMixed>
__device__ __forceinline__ void BatchItem(const FrHexBatchGroup &g, int64_t block, double *lds) {
  switch (g.tile) {
  case 0: BatchBody<Q, 0, kRiemann, kMixed>(&g, block, lds); break;
  case 1: BatchBody<Q, 1, kRiemann, kMixed>(&g, block, lds); break;
  default: BatchBody<Q, 2, kRiemann, kMixed>(&g, block, lds); break;
  }
}

// Persistent batched pass: workgroup w processes the contiguous item range
// [w * n_items / W, (w + 1) * n_items / W) of the work list (blocks of all
// degree groups in order), so consecutive blocks -- and their neighbour
// cells -- stay in one XCD's L2.  LDS is sized at launch for the largest tile
// in use.
template <int kRiemann, bool kMixed>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void FrHexBatchKernel(FrHexBatchArgs b) {
  extern __shared__ double lds[];
  const int64_t w = blockIdx.x, nw = gridDim.x;
  const int64_t begin = w * b.n_items / nw, end = (w + 1) * b.n_items / nw;
  int gi = 0;
  for (int64_t item = begin; item < end; ++item) {
    // the group table lives in device memory and gi is uniform, so every
    // field is fetched with scalar loads
    while (gi + 1 < b.n_groups && item >= b.group[gi + 1].first_item) ++gi;
    const FrHexBatchGroup &g = b.group[gi];
    const int64_t block = item - g.first_item;
    switch (g.q) {
    case 2: BatchItem<2, kRiemann, kMixed>(g, block, lds); break;
    case 3: BatchItem<3, kRiemann, kMixed>(g, block, lds); break;
    default: BatchItem<4, kRiemann, kMixed>(g, block, lds); break;
    }
    __syncthreads();  // the next item reuses the LDS buffers
  }
}

__global__ void ResetTimeStepBitsKernel(unsigned long long *bits) {
  *bits = (unsigned long long)__double_as_longlong(1e300);
}

template <int Q>
__global__ __launch_bounds__(256) void FrHexTimeStepKernel(const double *u, const int32_t *geo_id,
    const double *geo, int64_t n_cells, unsigned long long *dt_min_bits) {
  constexpr int NP = Q * Q * Q;
  __shared__ double s_min[256];
  double local = 1e300;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < n_cells; c += (int64_t)gridDim.x * 256) {
    double vmax = 1e-5;  // fem.hpp:419, avoids 0 in the divisor
    const double *uc = u + c * NP * 5;
    for (int p = 0; p < NP; ++p) {
      State s;
#pragma unroll
      for (int k = 0; k < 5; ++k) s.v[k] = uc[p * 5 + k];
      vmax = fmax(vmax, MaxSpeed(s));
    }
    local = fmin(local, geo[(int64_t)geo_id[c] * kGeoStride + 16] / vmax);
  }
  s_min[threadIdx.x] = local;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMin(dt_min_bits, (unsigned long long)__double_as_longlong(s_min[0]));
}

template <int Q, int R>
hipError_t Launch(const FrHexArgs &args, const LineTables &lt, hipStream_t stream) {
  constexpr int C = StageTile<Q>::kCells, W = StageTile<Q>::kThreads;
  constexpr bool O = StageTile<Q>::kOneRound;
  constexpr int PF = StageTile<Q>::kPrefetch;
  int64_t blocks = (args.n_cells + C - 1) / C;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks), block(W);
  if (args.curved) {
    if (args.mixed || args.diag_metric) return hipErrorInvalidValue;
    hipLaunchKernelGGL((FrHexStageKernel<Q, C, W, R, false, false, O, PF, true>), grid, block, 0, stream, args, lt);
    return hipGetLastError();
  }
  if (args.mixed) {
    if (args.diag_metric) {
      hipLaunchKernelGGL((FrHexStageKernel<Q, C, W, R, true, true, O, PF>), grid, block, 0, stream, args, lt);
    } else {
      hipLaunchKernelGGL((FrHexStageKernel<Q, C, W, R, false, true, O, PF>), grid, block, 0, stream, args, lt);
    }
  } else {
    if (args.diag_metric) {
      hipLaunchKernelGGL((FrHexStageKernel<Q, C, W, R, true, false, O, PF>), grid, block, 0, stream, args, lt);
    } else {
      hipLaunchKernelGGL((FrHexStageKernel<Q, C, W, R, false, false, O, PF>), grid, block, 0, stream, args, lt);
    }
  }
  return hipGetLastError();
}

template <int Q>
hipError_t LaunchQ(int riemann, const FrHexArgs &args, const LineTables &lt, hipStream_t stream) {
  switch (riemann) {
  case kExact: return Launch<Q, kExact>(args, lt, stream);
  case kHllc: return Launch<Q, kHllc>(args, lt, stream);
  case kRoe: return Launch<Q, kRoe>(args, lt, stream);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t LaunchFrHexStage(int q, int riemann, const FrHexArgs &args, const LineTables &lt,
                            hipStream_t stream) {
  switch (q) {
  case 2: return LaunchQ<2>(riemann, args, lt, stream);
  case 3: return LaunchQ<3>(riemann, args, lt, stream);
  case 4: return LaunchQ<4>(riemann, args, lt, stream);
  case 5: return LaunchQ<5>(riemann, args, lt, stream);
  case 6: return LaunchQ<6>(riemann, args, lt, stream);
  default: return hipErrorInvalidValue;
  }
}

int FrHexTileLdsBytes(int q, int tile) {
  static const int lds[3][3] = { { TileLds<2, 0>(), TileLds<2, 1>(), TileLds<2, 2>() },
                                 { TileLds<3, 0>(), TileLds<3, 1>(), TileLds<3, 2>() },
                                 { TileLds<4, 0>(), TileLds<4, 1>(), TileLds<4, 2>() } };
  if (q < 2 || q > 4 || tile < 0 || tile > 2) return 0;
  return lds[q - 2][tile] * (int)sizeof(double);
}

int FrHexStageTileCells(int q) {
  switch (q) {
  case 2: return StageTile<2>::kCells;
  case 3: return StageTile<3>::kCells;
  case 4: return StageTile<4>::kCells;
  case 5: return StageTile<5>::kCells;
  case 6: return StageTile<6>::kCells;
  default: return 0;
  }
}

int FrHexTileCells(int q, int tile) {
  static const int cells[3][3] = { { 32, 16, 8 }, { 9, 4, 2 }, { 4, 2, 1 } };
  if (q < 2 || q > 4 || tile < 0 || tile > 2) return 0;
  return cells[q - 2][tile];
}

hipError_t LaunchFrHexBatch(int riemann, bool mixed, const FrHexBatchArgs &args, int n_workgroups,
                            hipStream_t stream) {
  if (args.n_groups < 1 || args.n_groups > kMaxBatchGroups || n_workgroups < 1) return hipErrorInvalidValue;
  const dim3 grid((unsigned)n_workgroups), block(256);
  const size_t lds = (size_t)args.max_tile_lds;
  switch (riemann) {
#define MCFD_BATCH_CASE(R)                                                                        \
  case R:                                                                                         \
    if (mixed) hipLaunchKernelGGL((FrHexBatchKernel<R, true>), grid, block, lds, stream, args);   \
    else hipLaunchKernelGGL((FrHexBatchKernel<R, false>), grid, block, lds, stream, args);        \
    break;
  MCFD_BATCH_CASE(kExact) MCFD_BATCH_CASE(kHllc) MCFD_BATCH_CASE(kRoe)
#undef MCFD_BATCH_CASE
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int FrHexBatchResidentWorkgroups(int riemann, bool mixed, int lds_bytes) {
  int per_cu = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (riemann * 2 + (mixed ? 1 : 0)) {
#define MCFD_OCC(R, M) \
  case R * 2 + M: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, FrHexBatchKernel<R, M>, 256, lds_bytes); break;
  MCFD_OCC(kExact, 0) MCFD_OCC(kExact, 1) MCFD_OCC(kHllc, 0) MCFD_OCC(kHllc, 1) MCFD_OCC(kRoe, 0) MCFD_OCC(kRoe, 1)
#undef MCFD_OCC
  default: break;
  }
  int dev = 0, cus = 0;
  if (e != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return std::max(1, per_cu) * cus;
}

hipError_t LaunchResetTimeStepBits(unsigned long long *bits, hipStream_t stream) {
  hipLaunchKernelGGL(ResetTimeStepBitsKernel, dim3(1), dim3(1), 0, stream, bits);
  return hipGetLastError();
}

hipError_t LaunchFrHexTimeStep(int q, const double *u, const int32_t *geo_id, const double *geo,
                               int64_t n_cells, unsigned long long *dt_min_bits, hipStream_t stream) {
  int64_t blocks = std::min<int64_t>((n_cells + 255) / 256, 4096);
  if (blocks < 1) blocks = 1;
  switch (q) {
#define MCFD_TS_CASE(QQ) \
  case QQ: hipLaunchKernelGGL((FrHexTimeStepKernel<QQ>), dim3((unsigned)blocks), dim3(256), 0, stream, \
                              u, geo_id, geo, n_cells, dt_min_bits); break;
  MCFD_TS_CASE(2) MCFD_TS_CASE(3) MCFD_TS_CASE(4) MCFD_TS_CASE(5) MCFD_TS_CASE(6)
#undef MCFD_TS_CASE
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mcfd
