// Modal DG stage kernels for affine hexahedra and tetrahedra (gfx950).
//
// One stage pass is two launches.
//
// DgFaceFluxKernel, one thread per face quadrature point: the holder's and
// the sharer's traces from their Taylor coefficients, the Riemann or boundary
// flux, times the weight -> fhat [face][point][5].  Each face is evaluated
// once, so its two cells read the same flux and the scheme stays
// conservative without atomics.
//
// DgModalStageKernel, a workgroup of 256 threads owning G cells; each cell
// has T = NQ + NF * NFP point tasks (its volume quadrature points and the
// quadrature points of its faces), one thread per task, volume tasks of all
// G cells first.
//   phase 0  the cells' Taylor coefficients and geometry are staged in LDS;
//            the Gram-Schmidt factors and the state / U0 of phase 3 are
//            loaded at the same time, so one round trip serves the pass;
//   phase 1  volume task: state u = Ct T(y) at the point, flux matrix times
//            the global weight -> LDS;  face task: -/+ w Fhat read from fhat
//            and the point's offset Y -> LDS;
//   phase 2  the Taylor moments Rt: a group of lanes per (cell, component),
//            each lane over a subset of the cell's tasks, summed by DPP;
//   phase 3  one thread per (cell, basis function, component): R = Rt S^T,
//            the SSP-RK3 combination (temporal/rk.hpp:31-49 operation order)
//            and the store;
//   phase 4  the new state's Taylor coefficients for the next pass (and,
//            with a limiter, the WENO detector's record of the new state);
//   phase 5  (optional) the time-step reduction of the new state.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dg_modal.hpp"
#include "euler.hpp"

namespace mcfd {
namespace {

// LDS record of one volume task: the weighted flux matrix as F[d][k] (15),
// a pad, then the point's offset y (3) from kSlotY.  A face task: the
// weighted flux (5), then y (3).
constexpr int kSlot = 20;
constexpr int kSlotY = 16;
constexpr int kFaceSlot = 9;

template <bool kHex, int P>
struct DgShape {
  static constexpr int kN = P == 1 ? 4 : 10;
  static constexpr int kNQ = kHex ? (P + 1) * (P + 1) * (P + 1) : 14;
  static constexpr int kNF = kHex ? 6 : 4;
  static constexpr int kNFP = kHex ? (P + 1) * (P + 1) : 6;
  static constexpr int kT = kNQ + kNF * kNFP;
  static constexpr int kG = 256 / kT;
  static constexpr int kPacked = kN * (kN + 1) / 2;
};

// y = J^T eta: offset from the cell centre of the point with local
// coordinates xi = xi_c + eta (J rows are d/dxi, d/deta, d/dzeta).
__device__ __forceinline__ void Offset(const double *j, const double *eta, double *y) {
#pragma unroll
  for (int d = 0; d < 3; ++d) y[d] = __fma_rn(eta[2], j[6 + d], __fma_rn(eta[1], j[3 + d], MulRn(eta[0], j[d])));
}

// u = Ct T(y), Ct stored [N][5].
template <int P>
__device__ __forceinline__ State EvalTaylor(const double *ct, const double *y) {
  constexpr int N = P == 1 ? 4 : 10;
  double t[10];
  t[0] = 1.0; t[1] = y[0]; t[2] = y[1]; t[3] = y[2];
  if constexpr (P >= 2) {
    t[4] = MulRn(y[0], y[0]); t[5] = MulRn(y[0], y[1]); t[6] = MulRn(y[0], y[2]);
    t[7] = MulRn(y[1], y[1]); t[8] = MulRn(y[1], y[2]); t[9] = MulRn(y[2], y[2]);
  }
  State u;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    double acc = ct[k];
#pragma unroll
    for (int j = 1; j < N; ++j) acc = __fma_rn(ct[5 * j + k], t[j], acc);
    u.v[k] = acc;
  }
  return u;
}

// The detector record's share of Taylor item (j, k) of a cell: the rho (k =
// 0) and E (k = 4) coefficients, and with j = 0 the averages u[0][k] *
// S[0][0] (polynomial().average() of limiter/weno.hpp:232-245).
__device__ __forceinline__ void FlagRecordItem(double *rec, int j, int k, double taylor, double u0k, double s00) {
  if (k == 0) rec[j] = taylor;
  if (k == 4) rec[10 + j] = taylor;
  if (j == 0 && (k == 0 || k == 4)) rec[kDgFlagAverage + (k == 4)] = u0k * s00;
}

// Lanes per (cell, component) in the moment phase of a 256-thread
// workgroup of G cells: the largest power of two up to 16 (one DPP row)
// with G * 5 * NPART <= 256.
template <int G>
constexpr int DgMomentParts() {
  int p = 16;
  while (p > 1 && G * 5 * p > 256) p /= 2;
  return p;
}

// v from the lane given by a DPP control (within a row of 16 lanes).
template <int kCtrl>
__device__ __forceinline__ double DppMove(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over each aligned group of NPART lanes (NPART = 1, 2, 4, 8 or 16), by
// a butterfly: lane ^ 1 (quad_perm 1,0,3,2), lane ^ 2 (quad_perm 2,3,0,1),
// the other quad of the half row (row_half_mirror), the other half of the
// row (row_ror 8).  Every lane of a group ends with the same bits, since each
// step adds two values whose order does not matter.  All lanes of the
// wavefront must execute it.
template <int NPART>
__device__ __forceinline__ double GroupSum(double v) {
  if constexpr (NPART >= 2) v += DppMove<0xB1>(v);
  if constexpr (NPART >= 4) v += DppMove<0x4E>(v);
  if constexpr (NPART >= 8) v += DppMove<0x141>(v);
  if constexpr (NPART >= 16) v += DppMove<0x128>(v);
  return v;
}

// The face fluxes w Fhat were computed once per face point by
// DgFaceFluxKernel into a.fhat; the face tasks read them.  Tetrahedra run at
// 5 waves per SIMD (the LDS of 5 workgroups fits in a CU): the few bytes the
// register bound spills cost less than the fifth wave gains.
template <bool kHex, int P>
__global__ __launch_bounds__(256, kHex ? 4 : 5) void DgModalStageKernel(DgModalArgs a) {
  using S = DgShape<kHex, P>;
  constexpr int N = S::kN, NQ = S::kNQ, NF = S::kNF, NFP = S::kNFP, T = S::kT, G = S::kG, PK = S::kPacked;
  constexpr int GS = kDgGeoStride;
  __shared__ double s_own[G * 5 * N + G * GS];  // own Taylor coefficients and geometry; then S
  __shared__ double s_rt[G * 5 * N];       // Taylor moments Rt
  __shared__ double s_vol[G * NQ * kSlot]; // volume tasks: flux x weight, pad, y
  __shared__ double s_face[G * NF * NFP * kFaceSlot];  // face tasks: -/+ w Fhat (5), y; then the new state
  static_assert(G * 5 * N <= G * NF * NFP * kFaceSlot, "LDS aliasing");
  static_assert(G * PK <= G * 5 * N + G * GS, "LDS aliasing");
  static_assert(G * N + G <= G * NQ * kSlot, "LDS aliasing");
  double *s_ct = s_own, *s_geo = s_own + G * 5 * N;
  double *s_s = s_own;  // Gram-Schmidt factors, from phase 2 on
  double *s_new = s_face;
  constexpr int kSIter = (G * PK + 255) / 256, kItems = (G * 5 * N + 255) / 256;

  const int tid = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * G;
  if (c0 >= a.n_cells) return;
  const int nc = (int)min<int64_t>(G, a.n_cells - c0);

  // ---- this thread's point task and its global inputs: issued before the
  // staging loads so that their latency overlaps them.  Tasks are numbered
  // volume points of all G cells first, then face points, so that at most one
  // wavefront holds both kinds and runs both branches.
  const bool is_face = tid >= G * NQ;
  const int tf = tid - G * NQ;
  const int g = is_face ? tf / (NF * NFP) : tid / NQ;
  const int r = is_face ? NQ + (tf - g * (NF * NFP)) : tid - g * NQ;
  const bool active = g < nc && tid < G * T;
  const int lf = is_face ? (r - NQ) / NFP : 0, q = is_face ? (r - NQ) - ((r - NQ) / NFP) * NFP : r;
  const DgTables *tb = a.tables;
  int32_t info = 0;
  double eta[3] = { 0, 0, 0 }, wl = 0;
  double fw[5] = { 0, 0, 0, 0, 0 };  // face task: w Fhat of this face point
  if (active) {
    if (!is_face) {
#pragma unroll
      for (int d = 0; d < 3; ++d) eta[d] = tb->vol_eta[r][d];
      wl = tb->vol_w[r];
    } else {
      info = a.finfo[(c0 + g) * 6 + lf];
      const bool holder = (info >> 7) & 1;
      const int orient = (info >> 3) & 15;
#pragma unroll
      for (int d = 0; d < 3; ++d) eta[d] = tb->face_eta[lf][holder ? 0 : orient][q][d];
      const double *src = a.fhat + ((int64_t)(info >> 8) * NFP + q) * 5;
#pragma unroll
      for (int k = 0; k < 5; ++k) fw[k] = src[k];
    }
  }

  // the Gram-Schmidt factors S and the state and U0 of the phase-3 items,
  // issued with the staging loads so that one round trip serves all of them
  double s_pf[kSIter];
#pragma unroll
  for (int it = 0; it < kSIter; ++it) {
    const int w = tid + 256 * it;
    s_pf[it] = w < nc * PK ? a.basis[c0 * PK + w] : 0.0;
  }
  double us_pf[kItems], u0_pf[kItems];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int w = tid + 256 * it;
    const bool on = w < nc * 5 * N;
    us_pf[it] = on && a.stage >= 1 ? a.us[c0 * 5 * N + w] : 0.0;
    u0_pf[it] = on && a.stage >= 2 ? a.u0[c0 * 5 * N + w] : 0.0;
  }

  // ---- phase 0: stage the cells' data in LDS
  for (int w = tid; w < nc * 5 * N; w += 256) s_ct[w] = a.ct[c0 * 5 * N + w];
  for (int w = tid; w < nc * GS; w += 256) s_geo[w] = a.geo[c0 * GS + w];
  __syncthreads();

  // ---- phase 1: point tasks
  if (active) {
    const double *geo = s_geo + g * GS;
    const double *own_ct = s_ct + g * 5 * N;
    if (!is_face) {
      double *slot = s_vol + (g * NQ + r) * kSlot;
      double y[3];
      Offset(geo, eta, y);
      const State u = EvalTaylor<P>(own_ct, y);
      double f[5][3];
      FluxMatrix(u, f);
      const double w = MulRn(wl, geo[9]);
#pragma unroll
      for (int k = 0; k < 5; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) slot[5 * d + k] = f[k][d] * w;
#pragma unroll
      for (int d = 0; d < 3; ++d) slot[kSlotY + d] = y[d];
    } else {
      // dg/general.hpp:80-88: -w Fhat into the holder, +w Fhat into the sharer
      double *slot = s_face + (g * NF * NFP + (r - NQ)) * kFaceSlot;
      const bool holder = (info >> 7) & 1;
      double y[3];
      Offset(geo, eta, y);
#pragma unroll
      for (int k = 0; k < 5; ++k) slot[k] = holder ? -fw[k] : fw[k];
#pragma unroll
      for (int d = 0; d < 3; ++d) slot[5 + d] = y[d];
    }
  }
  __syncthreads();

  // ---- phase 2: Taylor moments; S goes to the freed LDS of the own data
  // at the end
  // Taylor moments Rt[k][j].  Lane (cell, component, part) accumulates all N
  // moments of its component over the points part, part + NPART, ... of the
  // cell (volume points first, then face points), with the per-point FMA
  // chains of the monomial table (basis/taylor.hpp:204-208):
  //   volume: grad(Y_a Y_b) . F = F_a Y_b + F_b Y_a,  face: Fhat Y_a Y_b;
  // the NPART adjacent lanes then sum their partial moments by a fixed
  // butterfly, so every lane of the group holds the same bits.
  {
    constexpr int NPART = DgMomentParts<G>();
    const int grp = tid / NPART, part = tid - (tid / NPART) * NPART;
    const int gc = grp / 5, k = grp - (grp / 5) * 5;
    double mo[10] = { 0, 0, 0, 0, 0, 0, 0, 0, 0, 0 };
    if (gc < nc) {
      const double *vb = s_vol + gc * NQ * kSlot;
      for (int q = part; q < NQ; q += NPART) {
        const double *r = vb + q * kSlot;
        const double f0 = r[k], f1 = r[5 + k], f2 = r[10 + k];
        const double y0 = r[kSlotY], y1 = r[kSlotY + 1], y2 = r[kSlotY + 2];
        mo[1] += f0;
        mo[2] += f1;
        mo[3] += f2;
        if constexpr (N == 10) {
          mo[4] = __fma_rn(f0, y0, __fma_rn(f0, y0, mo[4]));
          mo[5] = __fma_rn(f1, y0, __fma_rn(f0, y1, mo[5]));
          mo[6] = __fma_rn(f2, y0, __fma_rn(f0, y2, mo[6]));
          mo[7] = __fma_rn(f1, y1, __fma_rn(f1, y1, mo[7]));
          mo[8] = __fma_rn(f2, y1, __fma_rn(f1, y2, mo[8]));
          mo[9] = __fma_rn(f2, y2, __fma_rn(f2, y2, mo[9]));
        }
      }
      const double *fb = s_face + gc * NF * NFP * kFaceSlot;
      for (int q = (part + NPART - NQ % NPART) % NPART; q < NF * NFP; q += NPART) {
        const double *r = fb + q * kFaceSlot;
        const double ff = r[k], y0 = r[5], y1 = r[6], y2 = r[7];
        mo[0] += ff;
        mo[1] = __fma_rn(ff, y0, mo[1]);
        mo[2] = __fma_rn(ff, y1, mo[2]);
        mo[3] = __fma_rn(ff, y2, mo[3]);
        if constexpr (N == 10) {
          mo[4] = __fma_rn(ff, y0 * y0, mo[4]);
          mo[5] = __fma_rn(ff, y0 * y1, mo[5]);
          mo[6] = __fma_rn(ff, y0 * y2, mo[6]);
          mo[7] = __fma_rn(ff, y1 * y1, mo[7]);
          mo[8] = __fma_rn(ff, y1 * y2, mo[8]);
          mo[9] = __fma_rn(ff, y2 * y2, mo[9]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) mo[j] = GroupSum<NPART>(mo[j]);
    if (gc < nc) {
#pragma unroll
      for (int j = 0; j < N; ++j)
        if (j % NPART == part) s_rt[gc * 5 * N + 5 * j + k] = mo[j];
    }
  }
#pragma unroll
  for (int it = 0; it < kSIter; ++it) {
    const int w = tid + 256 * it;
    if (w < nc * PK) s_s[w] = s_pf[it];
  }
  __syncthreads();

  // ---- phase 3: R = Rt S^T, RK stage, store
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int w = tid + 256 * it;
    if (w >= nc * 5 * N) break;
    const int gc = w / (5 * N), rem = w - gc * 5 * N, i = rem / 5, k = rem - (rem / 5) * 5;
    const double *srow = s_s + gc * PK + i * (i + 1) / 2;
    const double *rt = s_rt + gc * 5 * N;
    double res = 0.0;
    for (int j = 0; j <= i; ++j) res = __fma_rn(rt[5 * j + k], srow[j], res);
    const int64_t idx = c0 * 5 * N + w;
    double out;
    if (a.stage == 0) {
      out = res;
    } else {
      const double next = AddRn(us_pf[it], MulRn(res, a.dt));
      if (a.stage == 1) {
        out = next;
      } else if (a.stage == 2) {
        out = AddRn(MulRn(u0_pf[it], 3.0), next) / 4.0;
      } else {
        out = AddRn(MulRn(next, 2.0), u0_pf[it]) / 3.0;
      }
    }
    a.out[idx] = out;
    s_new[w] = out;
  }
  if (!a.ct_out) return;
  __syncthreads();

  // ---- phase 4: Taylor coefficients of the new state for the next pass
  for (int w = tid; w < nc * 5 * N; w += 256) {
    const int gc = w / (5 * N), rem = w - gc * 5 * N, j = rem / 5, k = rem - (rem / 5) * 5;
    const double *sn = s_new + gc * 5 * N;
    const double *sc = s_s + gc * PK;
    double acc = 0.0;
    for (int i = j; i < N; ++i) acc = __fma_rn(sn[5 * i + k], sc[i * (i + 1) / 2 + j], acc);
    a.ct_out[c0 * 5 * N + w] = acc;
    s_rt[w] = acc;  // for the fused time step
    if (a.flag_rec) FlagRecordItem(a.flag_rec + (c0 + gc) * kDgFlagRecord, j, k, acc, sn[k], sc[0]);
  }
  if (!a.dt_min_bits) return;

  // ---- phase 5: fem.hpp:412-435 on the new state, as DgTimeStepKernel: min
  // over cells of length / max(1e-5, max over the first N points of |u| + c)
  __syncthreads();
  double *s_speed = s_vol, *s_cell = s_vol + G * N;
  for (int w = tid; w < nc * N; w += 256) {
    const int gc = w / N, q = w - gc * N;
    double y[3];
    Offset(a.geo + (c0 + gc) * GS, tb->vol_eta[q], y);
    s_speed[w] = MaxSpeed(EvalTaylor<P>(s_rt + gc * 5 * N, y));
  }
  __syncthreads();
  if (tid < nc) {
    double vmax = 1e-5;
    for (int q = 0; q < N; ++q) vmax = fmax(vmax, s_speed[tid * N + q]);
    s_cell[tid] = a.geo[(c0 + tid) * GS + 10] / vmax;
  }
  __syncthreads();
  if (tid == 0) {
    double local = 1e300;
    for (int c = 0; c < nc; ++c) local = fmin(local, s_cell[c]);
    atomicMin(a.dt_min_bits, (unsigned long long)__double_as_longlong(local));
  }
}

// w Fhat at every face quadrature point, once per face (dg/general.hpp:51-55,
// 70-89 and the boundary fluxes :92-202): one thread per (face, point).  The
// point is located in the holder through its canonical order and in the
// sharer through the corner-permutation code, and each trace is evaluated
// from the cell's Taylor coefficients.  The exact solver's iteration needs
// more registers than 4 waves per SIMD allow; the default bound leaves the
// compiler room for it without spills.
template <bool kHex, int P, int kRiemann>
__global__ __launch_bounds__(256) void DgFaceFluxKernel(DgModalArgs a) {
  constexpr int N = P == 1 ? 4 : 10, NFP = DgShape<kHex, P>::kNFP, GS = kDgGeoStride;
  const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (w >= a.n_faces * NFP) return;
  const int64_t f = w / NFP;
  const int q = (int)(w - f * NFP);
  const DgTables *tb = a.tables;
  const int32_t holder = a.face_cells[f * 3], sharer = a.face_cells[f * 3 + 1], pk = a.face_cells[f * 3 + 2];
  const int hf = pk & 7, sf = (pk >> 3) & 7, orient = (pk >> 6) & 15;
  const double *fc = a.faces + f * kDgFaceStride;
  FaceFrame frame;
#pragma unroll
  for (int d = 0; d < 3; ++d) { frame.n[d] = fc[d]; frame.t[d] = fc[3 + d]; }
  frame.b[0] = frame.n[1] * frame.t[2] - frame.n[2] * frame.t[1];
  frame.b[1] = frame.n[2] * frame.t[0] - frame.n[0] * frame.t[2];
  frame.b[2] = frame.n[0] * frame.t[1] - frame.n[1] * frame.t[0];
  const double wq = MulRn(tb->face_w[q], fc[6]);
  double y[3];
  Offset(a.geo + (int64_t)holder * GS, tb->face_eta[hf][0][q], y);
  const State own = EvalTaylor<P>(a.ct + (int64_t)holder * 5 * N, y);
  State flux;
  // the exact solver's state of this face point (ExactFlux), kept across calls
  bool star = kRiemann == kExact ? a.exact_star[w] != 0 : false;
  if (sharer >= 0) {
    double ys[3];
    Offset(a.geo + (int64_t)sharer * GS, tb->face_eta[sf][orient][q], ys);
    const State other = EvalTaylor<P>(a.ct + (int64_t)sharer * 5 * N, ys);
    flux = UpwindFlux<kRiemann>(frame, own, other, star);
  } else {
    const int tag = -sharer - 1;
    const int kind = a.bc_kind[tag];
    // the tag's constant state, or this point's value of the tag's boundary
    // function at the stage time
    const double *gs = a.bc_point_state ? a.bc_point_state + (w - a.n_interior * NFP) * 5 : a.bc_state + tag * 5;
    State given;
#pragma unroll
    for (int k = 0; k < 5; ++k) given.v[k] = gs[k];
    if (kind == kBcInviscidWall) {
      flux = WallFlux(frame, own);
    } else if (kind == kBcSupersonicOutlet) {
      flux = StateFlux(frame, own);
    } else if (kind == kBcSupersonicInlet) {
      flux = StateFlux(frame, given);
    } else if (kind == kBcSubsonicInlet) {
      flux = SubsonicInletFlux(frame, own, given);
    } else if (kind == kBcSubsonicOutlet) {
      flux = SubsonicOutletFlux(frame, own, given);
    } else {
      flux = UpwindFlux<kRiemann>(frame, own, given, star);
    }
  }
  if constexpr (kRiemann == kExact) a.exact_star[w] = star ? 1 : 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) a.fhat[w * 5 + k] = flux.v[k] * wq;
}

// projection.hpp:28-43 (Project) of point values f at the cell quadrature
// points onto the orthonormal basis, one thread per (cell, basis function,
// component): c[i][k] = sum_q (f_q[k] phi_i(x_q)) w_q, with phi = S T(y) and the
// operations in the order of the host-side projection.
template <int P>
__global__ __launch_bounds__(256) void DgApproximateKernel(const double *values, const double *basis,
                                                           const double *geo, const DgTables *tb, int nq,
                                                           int64_t n_cells, double *coeffs) {
  constexpr int N = P == 1 ? 4 : 10, PACKED = N * (N + 1) / 2;
  const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (w >= n_cells * 5 * N) return;
  const int64_t c = w / (5 * N);
  const int r = (int)(w - c * 5 * N), i = r / 5, k = r - (r / 5) * 5;
  const double *g = geo + c * kDgGeoStride;
  const double *s = basis + c * PACKED + i * (i + 1) / 2;
  double acc = 0.0;
  for (int q = 0; q < nq; ++q) {
    const double *e = tb->vol_eta[q];
    double y[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      y[d] = AddRn(AddRn(MulRn(e[0], g[d]), MulRn(e[1], g[3 + d])), MulRn(e[2], g[6 + d]));
    const double t[10] = { 1.0, y[0], y[1], y[2], MulRn(y[0], y[0]), MulRn(y[0], y[1]),
                           MulRn(y[0], y[2]), MulRn(y[1], y[1]), MulRn(y[1], y[2]),
                           MulRn(y[2], y[2]) };
    double phi = 0.0;
    for (int j = 0; j <= i; ++j) phi = AddRn(phi, MulRn(s[j], t[j]));
    const double prod = MulRn(MulRn(values[(c * nq + q) * 5 + k], phi), MulRn(tb->vol_w[q], g[9]));
    acc = AddRn(acc, prod);
  }
  coeffs[w] = acc;
}

template <int P>
__global__ __launch_bounds__(256) void DgTaylorKernel(const double *us, const double *basis, int64_t n_cells,
                                                      double *ct, double *flag_rec) {
  constexpr int N = P == 1 ? 4 : 10, PACKED = N * (N + 1) / 2;
  const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (w >= n_cells * 5 * N) return;
  const int64_t c = w / (5 * N);
  const int rem = (int)(w - c * 5 * N), j = rem / 5, k = rem - (rem / 5) * 5;
  const double *cf = us + c * 5 * N;
  const double *s = basis + c * PACKED;
  double acc = 0.0;
  for (int i = j; i < N; ++i) acc = __fma_rn(cf[5 * i + k], s[i * (i + 1) / 2 + j], acc);
  ct[w] = acc;
  if (flag_rec) FlagRecordItem(flag_rec + c * kDgFlagRecord, j, k, acc, cf[k], s[0]);
}

template <bool kHex, int P>
__global__ __launch_bounds__(256) void DgTimeStepKernel(const double *ct, const double *geo, const DgTables *tb,
                                                        int64_t n_cells, unsigned long long *dt_min_bits) {
  constexpr int N = P == 1 ? 4 : 10;
  __shared__ double s_min[256];
  double local = 1e300;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < n_cells; c += (int64_t)gridDim.x * 256) {
    const double *gc = geo + c * kDgGeoStride;
    double vmax = 1e-5;  // fem.hpp:421
    for (int q = 0; q < N; ++q) {
      double y[3];
      Offset(gc, tb->vol_eta[q], y);
      vmax = fmax(vmax, MaxSpeed(EvalTaylor<P>(ct + c * 5 * N, y)));
    }
    local = fmin(local, gc[10] / vmax);
  }
  s_min[threadIdx.x] = local;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMin(dt_min_bits, (unsigned long long)__double_as_longlong(s_min[0]));
}

// ---------------------------------------------------------------- WENO limiter
constexpr int kLimiterSlots = 7;  // up to 6 neighbours, then the cell itself

// rho and E at offset y from their Taylor coefficients rho_ct[kStride * j],
// e_ct[kStride * j]: the FMA chains of components 0 and 4 in EvalTaylor.
template <int P, int kStride>
__device__ __forceinline__ void EvalTaylorRhoE(const double *rho_ct, const double *e_ct, const double *y, double *rho,
                                               double *e) {
  constexpr int N = P == 1 ? 4 : 10;
  double t[10];
  t[0] = 1.0; t[1] = y[0]; t[2] = y[1]; t[3] = y[2];
  if constexpr (P >= 2) {
    t[4] = MulRn(y[0], y[0]); t[5] = MulRn(y[0], y[1]); t[6] = MulRn(y[0], y[2]);
    t[7] = MulRn(y[1], y[1]); t[8] = MulRn(y[1], y[2]); t[9] = MulRn(y[2], y[2]);
  }
  double a0 = rho_ct[0], a4 = e_ct[0];
#pragma unroll
  for (int j = 1; j < N; ++j) {
    a0 = __fma_rn(rho_ct[kStride * j], t[j], a0);
    a4 = __fma_rn(e_ct[kStride * j], t[j], a4);
  }
  *rho = a0;
  *e = a4;
}

__global__ void DgLimiterResetKernel(int32_t *n_troubled) { *n_troubled = 0; }

// limiter/weno.hpp:230-258 (weno::Eigen::IsNotSmooth) on rho and E, with L
// lanes per cell: lane i evaluates neighbour i's polynomial at the cell's
// centre, and the group's first lane sums the jumps in neighbour order (the
// reference's order, so the flags are bitwise the same) and takes the
// maxima.  Every flagged cell is also appended to the troubled-cell list (one
// atomic per workgroup; the list order is irrelevant, each cell's
// reconstruction only depends on the cell).
constexpr int FlagLanes(bool hex) { return hex ? 8 : 4; }  // lanes per cell, a power of two dividing 64

template <bool kHex, int P>
__global__ __launch_bounds__(256) void DgWenoFlagsKernel(DgLimiterArgs a) {
  constexpr int NA = kHex ? 6 : 4;  // neighbours at most
  constexpr int L = FlagLanes(kHex);
  const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t c = w / L;
  const int i = (int)(w & (L - 1));
  const bool on = c < a.n_cells;
  const bool all = a.mode == 2 || a.mode == 4;
  int na = 0;
  double jump0 = 0.0, jump4 = 0.0, abs0 = 0.0, abs4 = 0.0;  // neighbour i's terms
  const double *rc = a.flag_rec + (on ? c : 0) * kDgFlagRecord;
  if (on && !all) {
    na = a.n_adj[c];
    if (i < na) {
      const int64_t o = a.adj[c * 6 + i];
      const double *ro = a.flag_rec + o * kDgFlagRecord;
      const double y[3] = { rc[kDgFlagCentre] - ro[kDgFlagCentre], rc[kDgFlagCentre + 1] - ro[kDgFlagCentre + 1],
                            rc[kDgFlagCentre + 2] - ro[kDgFlagCentre + 2] };
      double rho, e;
      EvalTaylorRhoE<P, 1>(ro, ro + 10, y, &rho, &e);
      jump0 = fabs(rc[0] - rho);  // u(centre) = Ct T(0)
      jump4 = fabs(rc[10] - e);
      abs0 = fabs(ro[kDgFlagAverage]);
      abs4 = fabs(ro[kDgFlagAverage + 1]);
    }
  }
  const int lane = (int)(threadIdx.x & 63), first = lane & ~(L - 1);
  double sum0 = 0.0, sum4 = 0.0, max0 = 0.0, max4 = 0.0;
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    const double v0 = __shfl(jump0, first + j), v4 = __shfl(jump4, first + j);
    const double m0 = __shfl(abs0, first + j), m4 = __shfl(abs4, first + j);
    if (j < na) {
      sum0 += v0;
      sum4 += v4;
      max0 = j == 0 ? m0 : fmax(max0, m0);
      max4 = j == 0 ? m4 : fmax(max4, m4);
    }
  }
  bool rough = false;
  if (on && i == 0) {
    if (all) {
      rough = true;
    } else {
      double max_abs0 = fmax(1e-9, fabs(rc[kDgFlagAverage])), max_abs4 = fmax(1e-9, fabs(rc[kDgFlagAverage + 1]));
      if (na > 0) {
        max_abs0 = fmax(max_abs0, max0);
        max_abs4 = fmax(max_abs4, max4);
      }
      double divisor = rc[kDgFlagDivisor];  // pow(volume, (P + 1) / 6), set on the host
      divisor *= na;
      const double reference = P < 3 ? 1.0 : 3.0;
      rough = sum0 / max_abs0 / divisor > reference || sum4 / max_abs4 / divisor > reference;
    }
    a.flags[c] = rough ? 1 : 0;
  }
  // append the flagged cells: one atomic per workgroup (flags cluster in
  // space, and late in a run most wavefronts hold some)
  __shared__ int s_count[4];
  __shared__ int s_base;
  const unsigned long long ballot = __ballot(rough);
  const int wave = (int)(threadIdx.x >> 6);
  if (lane == 0) s_count[wave] = (int)__popcll(ballot);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int total = s_count[0] + s_count[1] + s_count[2] + s_count[3];
    s_base = total > 0 ? atomicAdd(a.n_troubled, total) : 0;
  }
  __syncthreads();
  int base = s_base;
  for (int v = 0; v < wave; ++v) base += s_count[v];
  if (rough) a.troubled[base + __popcll(ballot & ((1ull << lane) - 1))] = (int32_t)c;
}

// weno.hpp:74-95 (GetSmoothness) of the scalar expansion c[stride * i]
// (i < N) on the basis `s` of a cell with geometry `gc`.  `vpow` holds the
// cell's volume^(2 o / 3 - 1) for derivative orders o = 1, 2, with the volume
// of OrthoNormal::Measure, 1 / S00^2 (the flag record's scale slots).
// `ys`, `ws` (optional): the cell's quadrature points as offsets from its
// centre and their global weights, already formed (the same arithmetic as
// Offset and vol_w * |det J| below).
template <bool kHex, int P>
__device__ double Smoothness(const double *c, int stride, const double *s, const double *gc, const DgTables *tb,
                             const double *vpow, const double (*ys)[3] = nullptr, const double *ws = nullptr) {
  constexpr int N = P == 1 ? 4 : 10, NQ = DgShape<kHex, P>::kNQ;
  double tc[10];  // Taylor coefficients (c S)[j]
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double acc = 0.0;
    for (int i = j; i < N; ++i) acc = __fma_rn(c[stride * i], s[i * (i + 1) / 2 + j], acc);
    tc[j] = acc;
    // one column of S in flight at a time: scheduled all at once, the 55
    // basis loads would hold twice the registers of the rest of the kernel
    asm volatile("" ::: "memory");
  }
  double integ[10];
#pragma unroll
  for (int j = 0; j < N; ++j) integ[j] = 0.0;
  for (int q = 0; q < NQ; ++q) {
    double y[3];
    if (ys) {
      y[0] = ys[q][0]; y[1] = ys[q][1]; y[2] = ys[q][2];
    } else {
      Offset(gc, tb->vol_eta[q], y);
    }
    const double w = ws ? ws[q] : tb->vol_w[q] * gc[9];
    double pdv[10];  // taylor.hpp:241-281 (GetPartialDerivatives)
    pdv[0] = 0.0;
#pragma unroll
    for (int j = 1; j < N; ++j) pdv[j] = tc[j];
    if constexpr (P >= 2) {
      pdv[1] += tc[4] * (2 * y[0]); pdv[1] += tc[5] * y[1]; pdv[1] += tc[6] * y[2];
      pdv[2] += tc[5] * y[0]; pdv[2] += tc[7] * (2 * y[1]); pdv[2] += tc[8] * y[2];
      pdv[3] += tc[6] * y[0]; pdv[3] += tc[8] * y[1]; pdv[3] += tc[9] * (2 * y[2]);
      pdv[4] += tc[4]; pdv[7] += tc[7]; pdv[9] += tc[9];
    }
#pragma unroll
    for (int j = 0; j < N; ++j) integ[j] += pdv[j] * pdv[j] * w;
  }
  double beta = integ[1];
  beta += integ[2];
  beta += integ[3];
  beta *= vpow[0];
  if constexpr (P >= 2) {
    double b2 = integ[4];
#pragma unroll
    for (int j = 5; j < 10; ++j) b2 += integ[j];
    b2 *= vpow[1] / 4.0;
    beta += b2;
  }
  return beta;
}

// weno.hpp:121-155 (Lazy::Borrow), then the smoothness of every slot.  Slot
// i < n_adj projects neighbour i onto the cell's basis and shifts its average
// onto the cell's; slot n_adj is the cell's own expansion.  weno::Eigen
// takes the projections from a.proj.  weno::Lazy (weno.hpp:157-194) blends
// them here, in LDS, and stages the result and its Taylor coefficients at the
// cell's list position; DgWenoScatterKernel writes them over the state once
// every projection has read its neighbours.  A wavefront works
// on CPW troubled cells at once, LPC = 64 / CPW lanes each: the per-point and
// per-(slot, component) phases of a tetrahedron (14 points, 25 smoothness
// tasks) would leave most of a 64-lane wavefront idle.  Every lane runs the
// same number of loop trips and barriers; a lane group past the end of the
// list only skips the work.
template <bool kHex, int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void DgWenoProjectKernel(DgLimiterArgs a) {
  constexpr int N = P == 1 ? 4 : 10, PACKED = N * (N + 1) / 2, NQ = DgShape<kHex, P>::kNQ;
  constexpr int NA = kHex ? 6 : 4;          // neighbours at most
  constexpr int CPW = kHex ? 1 : 2;         // troubled cells per wavefront
  constexpr int LPC = 64 / CPW;             // lanes per cell
  // Buffers whose lives do not overlap share LDS (12.7 KB per workgroup for
  // tetrahedra, 3 waves per SIMD): the neighbours' Taylor coefficients are
  // dead once their values at the quadrature points are formed, and those
  // values once the projections are.
  __shared__ double s_cp[CPW][NA + 1][5 * N];   // neighbours' Taylor coefficients; then the projections, own last
  __shared__ double s_fv[CPW][NA * NQ * 5];     // neighbour polynomials at the points; then weights and blend
  __shared__ double s_shift[CPW][NA][3];    // own centre - neighbour centre
  __shared__ double s_y[CPW][NQ][3];        // the cell's quadrature points, offsets from its centre
  __shared__ double s_phi[CPW][NQ][N];      // its basis at them
  __shared__ double s_w[CPW][NQ];           // and their global weights
  __shared__ double s_basis[CPW][PACKED];   // the cell's S
  __shared__ double s_vpow[CPW][2];         // its smoothness volume scales
  static_assert((NA + 1) * 5 + 5 * N <= NA * NQ * 5, "weights and blend fit the point values' buffer");
  const int g = threadIdx.x / LPC, lane = threadIdx.x - (threadIdx.x / LPC) * LPC;
  double (*const s_ct)[5 * N] = s_cp[g];    // [NA][5 N], until the point values are formed
  double (*const s_pr)[5 * N] = s_cp[g];    // [NA + 1][5 N], from the projections on
  double (*const s_f)[NQ][5] = reinterpret_cast<double (*)[NQ][5]>(s_fv[g]);  // [NA][NQ][5]
  double (*const s_wt)[5] = reinterpret_cast<double (*)[5]>(s_fv[g]);         // weno::Lazy: [NA + 1][5] weights
  double *const s_fresh = s_fv[g] + (NA + 1) * 5;                             // and the blend [5 N]
  const DgTables *tb = a.tables;
  const int count = *a.n_troubled;
  for (int base = blockIdx.x * CPW; base < count; base += gridDim.x * CPW) {
    const int li = base + g;
    const bool on = li < count;
    const int64_t c = on ? a.troubled[li] : 0;
    const int na = on ? a.n_adj[c] : 0;
    const double *gc = a.geo + c * kDgGeoStride;
    const double *s = s_basis[g];
    for (int e = lane; e < na * 5 * N; e += LPC) {
      const int sl = e / (5 * N), r = e - sl * (5 * N);
      s_ct[sl][r] = a.ct[(int64_t)a.adj[c * 6 + sl] * 5 * N + r];
    }
    for (int e = lane; on && e < PACKED; e += LPC) s_basis[g][e] = a.basis[c * PACKED + e];
    if (on && lane < 2) s_vpow[g][lane] = a.flag_rec[c * kDgFlagRecord + kDgFlagScale1 + lane];
    __syncthreads();
    if (lane < na) {
      const double *go = a.geo + (int64_t)a.adj[c * 6 + lane] * kDgGeoStride;
#pragma unroll
      for (int d = 0; d < 3; ++d) s_shift[g][lane][d] = gc[11 + d] - go[11 + d];
    }
    for (int q = lane; on && q < NQ; q += LPC) {
      double y[3];
      Offset(gc, tb->vol_eta[q], y);
      double t[10];
      t[0] = 1.0; t[1] = y[0]; t[2] = y[1]; t[3] = y[2];
      if constexpr (P >= 2) {
        t[4] = y[0] * y[0]; t[5] = y[0] * y[1]; t[6] = y[0] * y[2];
        t[7] = y[1] * y[1]; t[8] = y[1] * y[2]; t[9] = y[2] * y[2];
      }
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double phi = 0.0;
        for (int j = 0; j <= i; ++j) phi = __fma_rn(s[i * (i + 1) / 2 + j], t[j], phi);
        s_phi[g][q][i] = phi;
        asm volatile("" ::: "memory");  // one row of S in flight at a time
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) s_y[g][q][d] = y[d];
      s_w[g][q] = tb->vol_w[q] * gc[9];
    }
    __syncthreads();
    for (int e = lane; e < na * NQ; e += LPC) {
      const int sl = e / NQ, q = e - sl * NQ;
      const double yo[3] = { s_y[g][q][0] + s_shift[g][sl][0], s_y[g][q][1] + s_shift[g][sl][1],
                             s_y[g][q][2] + s_shift[g][sl][2] };
      const State f = EvalTaylor<P>(s_ct[sl], yo);
#pragma unroll
      for (int k = 0; k < 5; ++k) s_f[sl][q][k] = f.v[k];
    }
    __syncthreads();
    // projection.hpp:28-43 (Project) and ProjectionWrapper::operator+=(my
    // average - its average) on the constant coefficient
    for (int e = lane; e < na * 5 * N; e += LPC) {
      const int sl = e / (5 * N), r = e - sl * (5 * N), i = r / 5, k = r - (r / 5) * 5;
      double acc = 0.0;
      for (int q = 0; q < NQ; ++q) acc += (s_f[sl][q][k] * s_phi[g][q][i]) * s_w[g][q];
      if (i == 0) {
        const double s00 = s[0];
        acc += (a.u[c * 5 * N + k] * s00 - acc * s00) / s00;
      }
      s_pr[sl][r] = acc;
      if (a.mode >= 3) a.proj[(c * kLimiterSlots + sl) * 5 * N + r] = acc;
    }
    for (int r = lane; on && r < 5 * N; r += LPC) {
      const double v = a.u[c * 5 * N + r];
      s_pr[na][r] = v;
      if (a.mode >= 3) a.proj[(c * kLimiterSlots + na) * 5 * N + r] = v;
    }
    if (a.mode < 3) {  // weno::Eigen measures smoothness in characteristic variables
      __syncthreads();
      // weno.hpp:157-194: the nonlinear weight of each (slot, component); the
      // cell's own linear weight is 1 - n_adj w0
      for (int e = lane; on && e < (na + 1) * 5; e += LPC) {
        const int sl = e / 5, k = e - sl * 5;
        double bb = Smoothness<kHex, P>(s_pr[sl] + k, 5, s, gc, tb, s_vpow[g], s_y[g], s_w[g]);
        double wi = a.w0;
        if (sl == na) { wi *= -na; wi += 1.0; }
        bb += a.eps;
        bb *= bb;
        s_wt[sl][k] = wi / bb;
      }
      __syncthreads();
      // the weighted sum, one lane per (coefficient, component)
      double *staged = a.proj + (int64_t)li * 2 * 5 * N;
      for (int r = lane; on && r < 5 * N; r += LPC) {
        const int k = r - (r / 5) * 5;
        double sum = 0.0;
        for (int sl = 0; sl <= na; ++sl) sum += s_wt[sl][k];
        double v = s_pr[na][r] * (s_wt[na][k] / sum);
        for (int sl = 0; sl < na; ++sl) v += s_pr[sl][r] * (s_wt[sl][k] / sum);
        s_fresh[r] = v;
        staged[r] = v;
      }
      __syncthreads();
      // its Taylor coefficients C S
      for (int r = lane; on && r < 5 * N; r += LPC) {
        const int j = r / 5, k = r - (r / 5) * 5;
        double acc = 0.0;
        for (int m = j; m < N; ++m) acc = __fma_rn(s_fresh[5 * m + k], s[m * (m + 1) / 2 + j], acc);
        staged[5 * N + r] = acc;
      }
    }
    __syncthreads();  // the next cells reuse the LDS buffers
  }
}

// weno::Lazy write-back: the staged expansion and Taylor coefficients of
// each troubled cell over its state, after every projection has read the
// old states (the Jacobi order of limiter/reconstruct.hpp:30-40).  One thread
// per (list position, row double), striding over the list.
template <int P>
__global__ __launch_bounds__(256) void DgWenoScatterKernel(DgLimiterArgs a) {
  constexpr int N = P == 1 ? 4 : 10, ROW = 5 * N;
  const int64_t total = (int64_t)*a.n_troubled * 2 * ROW;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int64_t li = e / (2 * ROW);
    const int r = (int)(e - li * 2 * ROW);
    const int64_t c = a.troubled[li];
    const double v = a.proj[e];
    if (r < ROW) a.u[c * ROW + r] = v;
    else a.ct[c * ROW + r - ROW] = v;
  }
}

// riemann/euler/eigen.hpp:26-56: left and right eigenvector matrices of the
// Euler flux Jacobian at state `cv` in the frame (n, t, b), row-major.
__device__ void EigenMatrices(const State &cv, const double *n, const double *t, const double *b,
                              double l[5][5], double r[5][5]) {
  const State pv = ToPrimitive(cv);
  const double ux = pv.v[1], uy = pv.v[2], uz = pv.v[3];
  const double u_nu = ux * n[0] + uy * n[1] + uz * n[2];
  const double u_mu = ux * t[0] + uy * t[1] + uz * t[2];
  const double u_pi = ux * b[0] + uy * b[1] + uz * b[2];
  const double a = sqrt(kGamma * pv.v[4] / pv.v[0]);
  const double ek = (ux * ux + uy * uy + uz * uz) / 2;
  const double e0 = a * a * (1 / kGm1);
  const double h0 = e0 + ek;
  const double u[3] = { ux, uy, uz };
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r[1 + d][0] = u[d] - a * n[d];
    r[1 + d][1] = u[d];
    r[1 + d][2] = t[d];
    r[1 + d][3] = b[d];
    r[1 + d][4] = u[d] + a * n[d];
  }
  r[0][0] = 1; r[0][1] = 1; r[0][2] = 0; r[0][3] = 0; r[0][4] = 1;
  r[4][0] = h0 - a * u_nu; r[4][1] = ek; r[4][2] = u_mu; r[4][3] = u_pi; r[4][4] = h0 + a * u_nu;
  const double b1 = 1 / e0, b2 = b1 * ek;
  l[0][0] = (b2 + u_nu / a) / 2; l[1][0] = 1 - b2; l[2][0] = -u_mu; l[3][0] = -u_pi; l[4][0] = (b2 - u_nu / a) / 2;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    l[0][1 + d] = -(b1 * u[d] + n[d] / a) / 2;
    l[1][1 + d] = b1 * u[d];
    l[2][1 + d] = t[d];
    l[3][1 + d] = b[d];
    l[4][1 + d] = -(b1 * u[d] - n[d] / a) / 2;
  }
  l[0][4] = b1 / 2; l[1][4] = -b1; l[2][4] = 0; l[3][4] = 0; l[4][4] = b1 / 2;
}

__device__ __forceinline__ void FaceFrameOf(const double *fc, double *n, double *t, double *b) {
#pragma unroll
  for (int d = 0; d < 3; ++d) { n[d] = fc[d]; t[d] = fc[3 + d]; }
  b[0] = n[1] * t[2] - n[2] * t[1];
  b[1] = n[2] * t[0] - n[0] * t[2];
  b[2] = n[0] * t[1] - n[1] * t[0];
}

// weno.hpp:292-337 (weno::Eigen::ReconstructOnFace): one thread per
// (troubled cell, adjacent face, characteristic field k).  Every borrowed
// projection is mapped by row k of L (cell average, face frame), its
// smoothness weighs it, and the blended field-k coefficients go to a.eig.
template <bool kHex, int P>
__global__ __launch_bounds__(256) void DgWenoEigenKernel(DgLimiterArgs a) {
  constexpr int N = P == 1 ? 4 : 10, PACKED = N * (N + 1) / 2;
  const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t c = w / 30;
  const int rem = (int)(w - c * 30), f = rem / 5, k = rem - (rem / 5) * 5;
  if (c >= a.n_cells || !a.flags[c]) return;
  const int na = a.n_adj[c];
  if (f >= na) return;
  const double *s = a.basis + c * PACKED;
  const double *gc = a.geo + c * kDgGeoStride;
  const double *pr = a.proj + c * kLimiterSlots * 5 * N;
  State avg;  // Projection::average of the cell's own expansion (projection slot na)
#pragma unroll
  for (int m = 0; m < 5; ++m) avg.v[m] = pr[na * 5 * N + m] * s[0];
  double n[3], t[3], b[3];
  FaceFrameOf(a.faces + (int64_t)a.adj_face[c * 6 + f] * kDgFaceStride, n, t, b);
  double l[5][5], r[5][5];
  EigenMatrices(avg, n, t, b, l, r);
  double wt[kLimiterSlots];
  for (int i = 0; i <= na; ++i) {
    double ch[10];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double v = 0.0;
#pragma unroll
      for (int m = 0; m < 5; ++m) v += l[k][m] * pr[i * 5 * N + 5 * j + m];
      ch[j] = v;
    }
    double beta = Smoothness<kHex, P>(ch, 1, s, gc, a.tables, a.flag_rec + c * kDgFlagRecord + kDgFlagScale1);
    double wi = a.w0;
    if (i == na) { wi *= -na; wi += 1.0; }
    beta += a.eps;
    beta *= beta;
    wt[i] = wi / beta;
  }
  double sum = 0.0;
  for (int i = 0; i <= na; ++i) sum += wt[i];
  double *out = a.eig + ((c * 6 + f) * 5 + k) * N;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double ch_self = 0.0;
#pragma unroll
    for (int m = 0; m < 5; ++m) ch_self += l[k][m] * pr[na * 5 * N + 5 * j + m];
    double v = ch_self * (wt[na] / sum);
    for (int i = 0; i < na; ++i) {
      double ch = 0.0;
#pragma unroll
      for (int m = 0; m < 5; ++m) ch += l[k][m] * pr[i * 5 * N + 5 * j + m];
      v += ch * (wt[i] / sum);
    }
    out[j] = v;
  }
}

// weno.hpp:317-347: back to conservative variables with each face's R, the
// neighbours' volumes as weights; one thread per (troubled cell, component).
template <int P>
__global__ __launch_bounds__(256) void DgWenoEigenCombineKernel(DgLimiterArgs a) {
  constexpr int N = P == 1 ? 4 : 10, PACKED = N * (N + 1) / 2;
  const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t c = w / 5;
  const int m = (int)(w - c * 5);
  if (c >= a.n_cells || !a.flags[c]) return;
  const int na = a.n_adj[c];
  const double *s = a.basis + c * PACKED;
  const double *pr = a.proj + c * kLimiterSlots * 5 * N;
  State avg;  // from the own-projection slot: a.u is written below by sibling threads
#pragma unroll
  for (int q = 0; q < 5; ++q) avg.v[q] = pr[na * 5 * N + q] * s[0];
  double acc[10];
#pragma unroll
  for (int i = 0; i < N; ++i) acc[i] = 0.0;
  double total = 0.0;
  for (int f = 0; f < na; ++f) {
    double n[3], t[3], b[3];
    FaceFrameOf(a.faces + (int64_t)a.adj_face[c * 6 + f] * kDgFaceStride, n, t, b);
    double l[5][5], r[5][5];
    EigenMatrices(avg, n, t, b, l, r);
    const double vol = a.geo[(int64_t)a.adj[c * 6 + f] * kDgGeoStride + 14];
    const double *e = a.eig + (c * 6 + f) * 5 * N;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double v = 0.0;
#pragma unroll
      for (int k = 0; k < 5; ++k) v += r[m][k] * e[k * N + i];
      acc[i] += v * vol;
    }
    total += vol;
  }
  double fresh[10];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    fresh[i] = acc[i] / total;
    a.u[c * 5 * N + 5 * i + m] = fresh[i];
  }
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double v = 0.0;
#pragma unroll
    for (int i = j; i < N; ++i) v = __fma_rn(fresh[i], s[i * (i + 1) / 2 + j], v);
    a.ct[c * 5 * N + 5 * j + m] = v;
  }
}

template <bool kHex, int P, int R>
hipError_t LaunchStage(const DgModalArgs &a, hipStream_t stream) {
  constexpr int G = DgShape<kHex, P>::kG;
  const int64_t blocks = (a.n_cells + G - 1) / G;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  constexpr int NFP = DgShape<kHex, P>::kNFP;
  const int64_t fblocks = (a.n_faces * NFP + 255) / 256;
  if (fblocks > 0x7fffffff || !a.fhat || !a.face_cells) return hipErrorInvalidValue;
  // face fluxes once per face point, then the cell pass
  if (fblocks > 0)
    hipLaunchKernelGGL((DgFaceFluxKernel<kHex, P, R>), dim3((unsigned)fblocks), dim3(256), 0, stream, a);
  hipLaunchKernelGGL((DgModalStageKernel<kHex, P>), dim3((unsigned)blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

template <bool kHex, int P>
hipError_t LaunchStageR(int riemann, const DgModalArgs &a, hipStream_t stream) {
  switch (riemann) {
  case kExact: return LaunchStage<kHex, P, kExact>(a, stream);
  case kHllc: return LaunchStage<kHex, P, kHllc>(a, stream);
  case kRoe: return LaunchStage<kHex, P, kRoe>(a, stream);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t LaunchDgTaylorCoefficients(int degree, const double *us, const double *basis, int64_t n_cells,
                                      double *ct, double *flag_rec, hipStream_t stream) {
  const int n = degree == 1 ? 4 : 10;
  const int64_t blocks = (n_cells * 5 * n + 255) / 256;
  if (blocks < 1) return hipSuccess;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  if (degree == 1) {
    hipLaunchKernelGGL(DgTaylorKernel<1>, dim3((unsigned)blocks), dim3(256), 0, stream, us, basis, n_cells, ct,
                       flag_rec);
  } else if (degree == 2) {
    hipLaunchKernelGGL(DgTaylorKernel<2>, dim3((unsigned)blocks), dim3(256), 0, stream, us, basis, n_cells, ct,
                       flag_rec);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t LaunchDgModalStage(int cell_type, int degree, int riemann, const DgModalArgs &a, hipStream_t stream) {
  if (a.n_cells < 1) return hipSuccess;
  const bool hex = cell_type == 0;
  if (hex && degree == 1) return LaunchStageR<true, 1>(riemann, a, stream);
  if (hex && degree == 2) return LaunchStageR<true, 2>(riemann, a, stream);
  if (!hex && degree == 1) return LaunchStageR<false, 1>(riemann, a, stream);
  if (!hex && degree == 2) return LaunchStageR<false, 2>(riemann, a, stream);
  return hipErrorInvalidValue;
}

hipError_t LaunchDgLimiter(int cell_type, int degree, const DgLimiterArgs &a, hipStream_t stream) {
  if (a.n_cells < 1 || a.mode == 0) return hipSuccess;
  const bool hex = cell_type == 0;
  const int n = degree == 1 ? 4 : 10;
  const int64_t b_flags = (a.n_cells * FlagLanes(hex) + 255) / 256;
  if (b_flags > 0x7fffffff) return hipErrorInvalidValue;
  // 64-thread workgroups striding over the troubled-cell list: two rounds of
  // the 3072 that MI355X holds at once (256 CUs, 12 one-wave workgroups each
  // at 3 waves per SIMD); 8192 left a partial third round (0.827 against
  // 0.814 ms per stage with 44 % of 221 184 tetrahedra flagged)
  const int64_t b_proj = std::min<int64_t>(a.n_cells, 6144);
  const int64_t b_comb = (a.n_cells * 5 + 255) / 256;  // Eigen combine: one thread per (cell, component)
  const int64_t b_lazy = std::min<int64_t>((a.n_cells * 2 * 5 * n + 255) / 256, 4096);  // Lazy write-back
  if (b_comb > 0x7fffffff) return hipErrorInvalidValue;
  const int64_t b_eig = (a.n_cells * 30 + 255) / 256;
  if (b_eig > 0x7fffffff) return hipErrorInvalidValue;
  const bool eigen = a.mode >= 3;
#define MCFD_LIMITER_LAUNCH(P)                                                                          \
  hipLaunchKernelGGL(DgLimiterResetKernel, dim3(1), dim3(1), 0, stream, a.n_troubled);                 \
  if (hex) {                                                                                            \
    hipLaunchKernelGGL((DgWenoFlagsKernel<true, P>), dim3((unsigned)b_flags), dim3(256), 0, stream, a);  \
    hipLaunchKernelGGL((DgWenoProjectKernel<true, P>), dim3((unsigned)b_proj), dim3(64), 0, stream, a);  \
  } else {                                                                                              \
    hipLaunchKernelGGL((DgWenoFlagsKernel<false, P>), dim3((unsigned)b_flags), dim3(256), 0, stream, a); \
    hipLaunchKernelGGL((DgWenoProjectKernel<false, P>), dim3((unsigned)b_proj), dim3(64), 0, stream, a); \
  }                                                                                                     \
  if (!eigen) {                                                                                         \
    hipLaunchKernelGGL(DgWenoScatterKernel<P>, dim3((unsigned)b_lazy), dim3(256), 0, stream, a);        \
  } else {                                                                                              \
    if (hex) {                                                                                          \
      hipLaunchKernelGGL((DgWenoEigenKernel<true, P>), dim3((unsigned)b_eig), dim3(256), 0, stream, a); \
    } else {                                                                                            \
      hipLaunchKernelGGL((DgWenoEigenKernel<false, P>), dim3((unsigned)b_eig), dim3(256), 0, stream, a); \
    }                                                                                                   \
    hipLaunchKernelGGL(DgWenoEigenCombineKernel<P>, dim3((unsigned)b_comb), dim3(256), 0, stream, a);   \
  }
  if (degree == 1) {
    MCFD_LIMITER_LAUNCH(1)
  } else if (degree == 2) {
    MCFD_LIMITER_LAUNCH(2)
  } else {
    return hipErrorInvalidValue;
  }
#undef MCFD_LIMITER_LAUNCH
  return hipGetLastError();
}

hipError_t LaunchDgApproximate(int degree, const double *values, const double *basis, const double *geo,
                               const DgTables *tables, int nq, int64_t n_cells, double *coeffs, hipStream_t stream) {
  const int n = degree == 1 ? 4 : 10;
  const int64_t blocks = (n_cells * 5 * n + 255) / 256;
  if (blocks < 1) return hipSuccess;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  if (degree == 1) {
    hipLaunchKernelGGL(DgApproximateKernel<1>, dim3((unsigned)blocks), dim3(256), 0, stream, values, basis, geo,
                       tables, nq, n_cells, coeffs);
  } else if (degree == 2) {
    hipLaunchKernelGGL(DgApproximateKernel<2>, dim3((unsigned)blocks), dim3(256), 0, stream, values, basis, geo,
                       tables, nq, n_cells, coeffs);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t LaunchDgTimeStep(int cell_type, int degree, const double *ct, const double *geo, const DgTables *tables,
                            int64_t n_cells, unsigned long long *dt_min_bits, hipStream_t stream) {
  const int64_t blocks = std::min<int64_t>((n_cells + 255) / 256, 4096);
  if (blocks < 1) return hipSuccess;
  const dim3 grid((unsigned)blocks), block(256);
  const bool hex = cell_type == 0;
  if (hex && degree == 1) {
    hipLaunchKernelGGL((DgTimeStepKernel<true, 1>), grid, block, 0, stream, ct, geo, tables, n_cells, dt_min_bits);
  } else if (hex && degree == 2) {
    hipLaunchKernelGGL((DgTimeStepKernel<true, 2>), grid, block, 0, stream, ct, geo, tables, n_cells, dt_min_bits);
  } else if (!hex && degree == 1) {
    hipLaunchKernelGGL((DgTimeStepKernel<false, 1>), grid, block, 0, stream, ct, geo, tables, n_cells, dt_min_bits);
  } else if (!hex && degree == 2) {
    hipLaunchKernelGGL((DgTimeStepKernel<false, 2>), grid, block, 0, stream, ct, geo, tables, n_cells, dt_min_bits);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mcfd
