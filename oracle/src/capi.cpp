// Oracle (test infrastructure): C entry points used by tests/ through ctypes,
// by the smoke check and by the CPU-baseline leg of bench.py.  Nothing in the
// product links against this library.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <thread>
#include <vector>

#include "core.hpp"
#include "dg.hpp"
#include "fr.hpp"
#include "mesh.hpp"

using namespace oracle;

namespace {

struct FrHandle {
  Mesh mesh;
  std::unique_ptr<FrScheme> scheme;
};

struct DgHandle {
  Mesh mesh;
  std::unique_ptr<DgScheme> scheme;
};

thread_local std::string g_error;

template <class F>
int Guard(F &&f) {
  try {
    f();
    return 0;
  } catch (std::exception const &e) {
    g_error = e.what();
    return 1;
  }
}

MeshInput MakeInput(int cell_type, int64_t n_nodes, double const *coords, int64_t n_cells,
    int64_t const *cells, int64_t n_bfaces, int64_t const *bface_nodes, int32_t const *bface_tags,
    double const *period) {
  return MeshInput{ cell_type, n_nodes, coords, n_cells, cells, n_bfaces, bface_nodes, bface_tags, period };
}

std::vector<BcEntry> MakeBcs(int n_bc, int32_t const *kinds, double const *states) {
  std::vector<BcEntry> bcs(n_bc);
  for (int b = 0; b < n_bc; ++b) {
    bcs[b].kind = kinds[b];
    for (int k = 0; k < 5; ++k) bcs[b].state[k] = states[5 * b + k];
  }
  return bcs;
}

// temporal/rk.hpp:31-49 with temporal/ode.hpp:121-128, same operation order.
// `set_solution` is the System's SetSolutionColumn side effect on a freshly
// set column (a limiter, spatial/with_limiter.hpp:56-59), or nothing.
template <class Residual, class SetSolution>
void Rk3Step(size_t size, double *u, double dt, Residual &&residual, SetSolution &&set_solution) {
  std::vector<double> u_curr(u, u + size), r(size), u_next(size);
  auto next_solution = [&](double const *state, double *out) {  // Euler::NextSolution
    residual(state, r.data());
    for (size_t i = 0; i < size; ++i) { r[i] *= dt; out[i] = state[i] + r[i]; }
  };
  std::vector<double> u1(size), t(size);
  next_solution(u, u1.data());
  set_solution(u1.data());
  for (size_t i = 0; i < size; ++i) { u_next[i] = u_curr[i]; u_next[i] *= 3; }
  next_solution(u1.data(), t.data());
  for (size_t i = 0; i < size; ++i) { u_next[i] += t[i]; u_next[i] /= 4; }
  set_solution(u_next.data());
  next_solution(u_next.data(), t.data());
  for (size_t i = 0; i < size; ++i) { t[i] *= 2; t[i] += u_curr[i]; t[i] /= 3; u[i] = t[i]; }
  set_solution(u);
}
template <class Residual>
void Rk3Step(size_t size, double *u, double dt, Residual &&residual) {
  Rk3Step(size, u, dt, residual, [](double *) {});
}

// Points of the boundary faces, [boundary face][point][3] in the restatement's
// order, and the given states there ([same][5], or n = 0 for the tags'
// constant states): the caller evaluates the BC functions f(x, t).
template <class Handle>
void BoundaryPoints(Handle *p, double *xyz, int32_t *tags) {
  size_t i = 0, j = 0;
  for (size_t f = p->mesh.n_interior; f < p->mesh.faces.size(); ++f)
    for (auto const &x : p->mesh.faces[f].qp) {
      for (int d = 0; d < 3; ++d) xyz[i++] = x[d];
      if (tags) tags[j++] = p->mesh.faces[f].tag;
    }
}
template <class Handle>
int64_t BoundaryPointCount(Handle *p) {
  int64_t n = 0;
  for (size_t f = p->mesh.n_interior; f < p->mesh.faces.size(); ++f) n += p->mesh.faces[f].qp.size();
  return n;
}
template <class Scheme>
void SetPointStates(Scheme *s, int64_t n, double const *states) {
  s->point_state.resize(n);
  for (int64_t i = 0; i < n; ++i) std::memcpy(s->point_state[i].data(), states + 5 * i, 40);
}
}  // namespace

extern "C" {

char const *oracle_last_error() { return g_error.c_str(); }

// ---------------------------------------------------------------- building blocks
int oracle_quadrature(int lobatto, int q, double *x, double *w) {
  return Guard([&] { if (lobatto) LobattoRule(q, x, w); else LegendreRule(q, x, w); });
}
// out = { g_R(s), g_L(s), g_R'(s), g_L'(s) }
void oracle_vincent(int degree, double c_next, double s, double *out) {
  Vincent v(degree, c_next);
  out[0] = v.RightValue(s); out[1] = v.LeftValue(s);
  out[2] = v.RightDerivative(s); out[3] = v.LeftDerivative(s);
}
// Lagrange basis on `q` nodes x: values and derivatives of all basis functions at s.
void oracle_lagrange(int q, double const *x, double s, double *values, double *derivs) {
  Lagrange1D l;
  l.q = q;
  for (int i = 0; i < q; ++i) l.x[i] = x[i];
  for (int i = 0; i < q; ++i) { values[i] = l.Value(i, s); derivs[i] = l.Derivative(i, s); }
}
void oracle_conservative_to_primitive(double const *cv, double *pv) {
  Val5 c, p;
  std::memcpy(c.data(), cv, 40);
  p = ConservativeToPrimitive(c);
  std::memcpy(pv, p.data(), 40);
}
void oracle_primitive_to_conservative(double const *pv, double *cv) {
  Val5 c, p;
  std::memcpy(p.data(), pv, 40);
  c = PrimitiveToConservative(p);
  std::memcpy(cv, c.data(), 40);
}
void oracle_flux_matrix(double const *cv, double *f15) {
  Val5 c;
  std::memcpy(c.data(), cv, 40);
  Flux53 f = GetFluxMatrix(c);
  for (int k = 0; k < 5; ++k) for (int d = 0; d < 3; ++d) f15[k * 3 + d] = f[k][d];
}
// 1-D exact solver on normal-frame primitives: the t-axis primitive state.
void oracle_exact_primitive(double const *pl, double const *pr, double *out) {
  Val5 l, r;
  std::memcpy(l.data(), pl, 40);
  std::memcpy(r.data(), pr, 40);
  ExactSolver s;
  Val5 o = s.PrimitiveUpwind(l, r);
  std::memcpy(out, o.data(), 40);
}
// Rotated upwind flux of conservative states, frame = {a, b, c} row-wise.
void oracle_riemann_flux(int kind, double const *frame9, double const *cl, double const *cr, double *flux) {
  Frame f;
  for (int d = 0; d < 3; ++d) { f.a[d] = frame9[d]; f.b[d] = frame9[3 + d]; f.c[d] = frame9[6 + d]; }
  Val5 l, r;
  std::memcpy(l.data(), cl, 40);
  std::memcpy(r.data(), cr, 40);
  RotatedEuler re;
  re.kind = RiemannKind(kind);
  Val5 o = re.GetFluxUpwind(f, l, r);
  std::memcpy(flux, o.data(), 40);
}
// The fluxes of mcfd_face_fluxes at n points, each with its own solver
// object whose exact-solver state star_u[i] (exact.hpp:29) is read and
// written back.  kind 0 upwind, 1 inviscid wall, 2 supersonic outlet,
// 3 subsonic inlet and 4 subsonic outlet with the given values in `right`.
void oracle_face_fluxes(int riemann, int kind, int64_t n, double const *frames, double const *left,
    double const *right, double *star_u, double *flux) {
  for (int64_t i = 0; i < n; ++i) {
    Frame f;
    for (int d = 0; d < 3; ++d) {
      f.a[d] = frames[9 * i + d]; f.b[d] = frames[9 * i + 3 + d]; f.c[d] = frames[9 * i + 6 + d];
    }
    Val5 l, r{};
    std::memcpy(l.data(), left + 5 * i, 40);
    if (right) std::memcpy(r.data(), right + 5 * i, 40);
    RotatedEuler re;
    re.kind = RiemannKind(riemann);
    if (star_u) re.exact.star_u = star_u[i];
    Val5 o = kind == 1   ? re.GetFluxOnInviscidWall(f, l)
             : kind == 2 ? re.GetFluxOfState(f, l)
             : kind == 3 ? re.GetFluxOnSubsonicInlet(f, l, r)
             : kind == 4 ? re.GetFluxOnSubsonicOutlet(f, l, r)
                         : re.GetFluxUpwind(f, l, r);
    if (star_u) star_u[i] = re.exact.star_u;
    std::memcpy(flux + 5 * i, o.data(), 40);
  }
}
void oracle_wall_flux(double const *frame9, double const *c, double *flux) {
  Frame f;
  for (int d = 0; d < 3; ++d) { f.a[d] = frame9[d]; f.b[d] = frame9[3 + d]; f.c[d] = frame9[6 + d]; }
  Val5 u;
  std::memcpy(u.data(), c, 40);
  RotatedEuler re;
  re.kind = kHllc;
  Val5 o = re.GetFluxOnInviscidWall(f, u);
  std::memcpy(flux, o.data(), 40);
}

// ---------------------------------------------------------------- nodal FR on hexahedra
int oracle_fr_create(int64_t n_nodes, double const *coords, int64_t n_cells, int64_t const *cells,
    int64_t n_bfaces, int64_t const *bface_nodes, int32_t const *bface_tags, double const *period,
    int q, int lobatto, double c_next, int riemann, int n_bc, int32_t const *bc_kinds,
    double const *bc_states, void **out) {
  return Guard([&] {
    auto h = std::make_unique<FrHandle>();
    h->mesh = BuildHexMesh(MakeInput(kHexa8, n_nodes, coords, n_cells, cells, n_bfaces, bface_nodes,
        bface_tags, period), q, lobatto);
    h->scheme = std::make_unique<FrScheme>(&h->mesh, c_next, riemann, MakeBcs(n_bc, bc_kinds, bc_states));
    *out = h.release();
  });
}
void oracle_fr_destroy(void *h) { delete static_cast<FrHandle *>(h); }
int64_t oracle_fr_boundary_points(void *h, double *xyz, int32_t *tags) {
  auto *p = static_cast<FrHandle *>(h);
  if (xyz) BoundaryPoints(p, xyz, tags);
  return BoundaryPointCount(p);
}
void oracle_fr_set_boundary_states(void *h, int64_t n, double const *states) {
  SetPointStates(static_cast<FrHandle *>(h)->scheme.get(), n, states);
}

int oracle_fr_startup_limiter(void *h, double *u, int kind) {
  return Guard([&] { static_cast<FrHandle *>(h)->scheme->StartupLimiter(u, kind); });
}
// Plants a known mistake in the FR restatement (FrScheme::mutation), so that
// tests can show which pins catch it.
void oracle_fr_set_mutation(void *h, int mutation) { static_cast<FrHandle *>(h)->scheme->mutation = mutation; }
int64_t oracle_fr_n_faces(void *h, int64_t *n_interior) {
  auto *p = static_cast<FrHandle *>(h);
  *n_interior = p->mesh.n_interior;
  return p->mesh.faces.size();
}
// Per face: holder, sharer, local face ids, and the FluxPointCache::ijk of
// every face point on both sides (-1 on the sharer side of boundary faces).
void oracle_fr_faces(void *h, int64_t *holder, int64_t *sharer, int32_t *holder_face,
    int32_t *sharer_face, int32_t *fp_ijk, double *frames) {
  auto *p = static_cast<FrHandle *>(h);
  auto &s = *p->scheme;
  int nfp = s.q * s.q;
  for (size_t f = 0; f < p->mesh.faces.size(); ++f) {
    auto const &face = p->mesh.faces[f];
    holder[f] = face.holder; sharer[f] = face.sharer;
    holder_face[f] = face.holder_face; sharer_face[f] = face.sharer_face;
    for (int i = 0; i < nfp; ++i) {
      auto const &ix = s.fp_index[f * nfp + i];
      fp_ijk[(f * nfp + i) * 2] = s.FluxPointNode(face.holder_face, ix[0], ix[1]);
      fp_ijk[(f * nfp + i) * 2 + 1] = face.sharer >= 0 ? s.FluxPointNode(face.sharer_face, ix[2], ix[3]) : -1;
      auto const &fr = face.frames[i];
      for (int d = 0; d < 3; ++d) {
        frames[(f * nfp + i) * 9 + d] = fr.a[d];
        frames[(f * nfp + i) * 9 + 3 + d] = fr.b[d];
        frames[(f * nfp + i) * 9 + 6 + d] = fr.c[d];
      }
    }
  }
}
// Global coordinates of every solution point, [cell][ijk][xyz].
void oracle_fr_node_coords(void *h, double *xyz) {
  auto *p = static_cast<FrHandle *>(h);
  size_t n = p->scheme->n, i = 0;
  for (auto const &cell : p->mesh.cells)
    for (size_t q = 0; q < n; ++q, ++i)
      for (int d = 0; d < 3; ++d) xyz[3 * i + d] = cell.qp[q][d];
}
// Quadrature weights (global) of every solution point, [cell][ijk].
void oracle_fr_node_weights(void *h, double *w) {
  auto *p = static_cast<FrHandle *>(h);
  size_t n = p->scheme->n, i = 0;
  for (auto const &cell : p->mesh.cells)
    for (size_t q = 0; q < n; ++q) w[i++] = cell.qw[q];
}
void oracle_fr_cell_lengths(void *h, double *len) {
  auto *p = static_cast<FrHandle *>(h);
  for (size_t c = 0; c < p->mesh.cells.size(); ++c) len[c] = p->mesh.cells[c].length;
}
int oracle_fr_residual(void *h, double const *u, double *r) {
  return Guard([&] { static_cast<FrHandle *>(h)->scheme->Residual(u, r); });
}
int oracle_fr_rk3(void *h, double *u, double dt, int n_steps) {
  return Guard([&] {
    auto *p = static_cast<FrHandle *>(h);
    size_t size = p->mesh.cells.size() * p->scheme->n * 5;
    for (int s = 0; s < n_steps; ++s)
      Rk3Step(size, u, dt, [&](double const *x, double *r) { p->scheme->Residual(x, r); });
  });
}
// oracle_fr_rk3 with the residual and the RK combination spread over
// `threads` host threads (bitwise the same states; CPU baseline timing).
int oracle_fr_rk3_threaded(void *h, double *u, double dt, int n_steps, int threads) {
  return Guard([&] {
    auto *p = static_cast<FrHandle *>(h);
    const size_t size = p->mesh.cells.size() * p->scheme->n * 5;
    auto each = [threads, size](auto &&op) {  // elementwise over [0, size) on `threads` threads
      std::vector<std::thread> pool;
      for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
          for (size_t i = size * t / threads; i < size * (t + 1) / threads; ++i) op(i);
        });
      for (auto &th : pool) th.join();
    };
    // Rk3Step with the same operation order per element; the buffers are
    // written before they are read
    auto buffer = [size] { return std::make_unique_for_overwrite<double[]>(size); };
    auto u_curr = buffer(), r = buffer(), u1 = buffer(), u_next = buffer(), t = buffer();
    auto next_solution = [&](double const *state, double *out) {
      p->scheme->ResidualThreaded(state, r.get(), threads);
      each([&](size_t i) { r[i] *= dt; out[i] = state[i] + r[i]; });
    };
    for (int s = 0; s < n_steps; ++s) {
      each([&](size_t i) { u_curr[i] = u[i]; });
      next_solution(u, u1.get());
      each([&](size_t i) { u_next[i] = u_curr[i]; u_next[i] *= 3; });
      next_solution(u1.get(), t.get());
      each([&](size_t i) { u_next[i] += t[i]; u_next[i] /= 4; });
      next_solution(u_next.get(), t.get());
      each([&](size_t i) { t[i] *= 2; t[i] += u_curr[i]; t[i] /= 3; u[i] = t[i]; });
    }
  });
}
double oracle_fr_time_step(void *h, double const *u, double cfl, double dt_guess) {
  return static_cast<FrHandle *>(h)->scheme->TimeStep(u, cfl, dt_guess);
}
// Wall-clock seconds of `n` residual evaluations (CPU baseline timing).
double oracle_fr_time_residuals(void *h, double const *u, double *r, int n) {
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < n; ++i) static_cast<FrHandle *>(h)->scheme->Residual(u, r);
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}


// ---------------------------------------------------------------- modal DG (hexahedra, tetrahedra)
int oracle_dg_create(int cell_type, int64_t n_nodes, double const *coords, int64_t n_cells,
    int64_t const *cells, int64_t n_bfaces, int64_t const *bface_nodes, int32_t const *bface_tags,
    double const *period, int degree, int riemann, int n_bc, int32_t const *bc_kinds,
    double const *bc_states, void **out) {
  return Guard([&] {
    auto h = std::make_unique<DgHandle>();
    h->mesh = BuildMesh(MakeInput(cell_type, n_nodes, coords, n_cells, cells, n_bfaces, bface_nodes,
        bface_tags, period), degree + 1, false);
    h->scheme = std::make_unique<DgScheme>(&h->mesh, degree, riemann, MakeBcs(n_bc, bc_kinds, bc_states));
    *out = h.release();
  });
}
void oracle_dg_destroy(void *h) { delete static_cast<DgHandle *>(h); }
// n_cells, basis functions per cell, quadrature points per cell, faces, interior faces
void oracle_dg_sizes(void *h, int64_t *out5) {
  auto *p = static_cast<DgHandle *>(h);
  out5[0] = p->mesh.cells.size();
  out5[1] = p->scheme->n;
  out5[2] = p->mesh.cells.empty() ? 0 : p->mesh.cells[0].qp.size();
  out5[3] = p->mesh.faces.size();
  out5[4] = p->mesh.n_interior;
}
// Cell quadrature points [cell][q][3], global weights [cell][q], centres [cell][3],
// Gram-Schmidt factors [cell][N][N] and lengths [cell].
void oracle_dg_cells(void *h, double *xyz, double *w, double *centers, double *s, double *lengths) {
  auto *p = static_cast<DgHandle *>(h);
  int n = p->scheme->n;
  size_t iq = 0;
  for (size_t c = 0; c < p->mesh.cells.size(); ++c) {
    auto const &cell = p->mesh.cells[c];
    for (size_t q = 0; q < cell.qp.size(); ++q, ++iq) {
      for (int d = 0; d < 3; ++d) xyz[3 * iq + d] = cell.qp[q][d];
      w[iq] = cell.qw[q];
    }
    for (int d = 0; d < 3; ++d) centers[3 * c + d] = p->scheme->basis[c].center[d];
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) s[(c * n + i) * n + j] = p->scheme->basis[c].s[i][j];
    lengths[c] = cell.length;
  }
}
// Faces: holder, sharer, local face ids, tag; per face point the global
// point [3], weight and frame [9] (face points per face = out of sizes()).
void oracle_dg_faces(void *h, int64_t *holder, int64_t *sharer, int32_t *holder_face, int32_t *sharer_face,
    int32_t *tag, double *xyz, double *w, double *frames) {
  auto *p = static_cast<DgHandle *>(h);
  size_t iq = 0;
  for (size_t f = 0; f < p->mesh.faces.size(); ++f) {
    auto const &face = p->mesh.faces[f];
    holder[f] = face.holder; sharer[f] = face.sharer;
    holder_face[f] = face.holder_face; sharer_face[f] = face.sharer_face; tag[f] = face.tag;
    for (size_t q = 0; q < face.qp.size(); ++q, ++iq) {
      for (int d = 0; d < 3; ++d) xyz[3 * iq + d] = face.qp[q][d];
      w[iq] = face.qw[q];
      auto const &fr = face.frames[q];
      for (int d = 0; d < 3; ++d) {
        frames[9 * iq + d] = fr.a[d]; frames[9 * iq + 3 + d] = fr.b[d]; frames[9 * iq + 6 + d] = fr.c[d];
      }
    }
  }
}
int64_t oracle_dg_boundary_points(void *h, double *xyz, int32_t *tags) {
  auto *p = static_cast<DgHandle *>(h);
  if (xyz) BoundaryPoints(p, xyz, tags);
  return BoundaryPointCount(p);
}
void oracle_dg_set_boundary_states(void *h, int64_t n, double const *states) {
  SetPointStates(static_cast<DgHandle *>(h)->scheme.get(), n, states);
}
int oracle_dg_residual(void *h, double const *u, double *r) {
  return Guard([&] { static_cast<DgHandle *>(h)->scheme->Residual(u, r); });
}
int oracle_dg_rk3(void *h, double *u, double dt, int n_steps) {
  return Guard([&] {
    auto *p = static_cast<DgHandle *>(h);
    size_t size = p->mesh.cells.size() * p->scheme->Stride();
    for (int s = 0; s < n_steps; ++s)
      Rk3Step(size, u, dt, [&](double const *x, double *r) { p->scheme->Residual(x, r); },
              [&](double *x) { p->scheme->Limit(x); });
  });
}
// Limiter of the DG handle (see DgScheme::limiter for the modes).
void oracle_dg_set_limiter(void *h, int mode, double w0, double eps) {
  auto *p = static_cast<DgHandle *>(h);
  p->scheme->limiter = mode; p->scheme->weno_w0 = w0; p->scheme->weno_eps = eps;
}
// Applies the limiter to u in place (as SetSolutionColumn would); flags[cell] gets 1 for
// a reconstructed cell.  Returns the number of troubled cells.
int64_t oracle_dg_limit(void *h, double *u, uint8_t *flags) {
  auto *p = static_cast<DgHandle *>(h);
  int64_t n = p->scheme->Limit(u);
  std::copy(p->scheme->flags.begin(), p->scheme->flags.end(), flags);
  return n;
}
// riemann/euler/eigen.hpp EigenMatrices of a conservative state in a frame
// (rows a, b, c); L and R row-major.
void oracle_eigen_matrices(double const *cv, double const *frame9, double *l, double *r) {
  Val5 c;
  std::memcpy(c.data(), cv, 40);
  Frame f;
  for (int d = 0; d < 3; ++d) { f.a[d] = frame9[d]; f.b[d] = frame9[3 + d]; f.c[d] = frame9[6 + d]; }
  double lm[5][5], rm[5][5];
  DgScheme::EigenMatrices(c, f, lm, rm);
  for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) { l[5 * i + j] = lm[i][j]; r[5 * i + j] = rm[i][j]; }
}
// Flags of the last limiter application (inside rk3 too).
void oracle_dg_last_flags(void *h, uint8_t *flags) {
  auto *p = static_cast<DgHandle *>(h);
  std::copy(p->scheme->flags.begin(), p->scheme->flags.end(), flags);
}
// Smoothness indicator (weno.hpp:74-95) of coefficients on cell c's basis.
void oracle_dg_smoothness(void *h, int64_t c, double const *coeff, double *beta) {
  Val5 b = static_cast<DgHandle *>(h)->scheme->Smoothness(c, coeff);
  for (int k = 0; k < 5; ++k) beta[k] = b[k];
}
double oracle_dg_time_step(void *h, double const *u, double cfl, double dt_guess) {
  return static_cast<DgHandle *>(h)->scheme->TimeStep(u, cfl, dt_guess);
}
void oracle_dg_approximate(void *h, double const *values, double *u) {
  static_cast<DgHandle *>(h)->scheme->Approximate(values, u);
}
// Value of cell c's expansion at a global point.
void oracle_dg_value(void *h, double const *u, int64_t c, double const *x, double *v) {
  Val5 o = static_cast<DgHandle *>(h)->scheme->Value(c, u, Vec3{ x[0], x[1], x[2] });
  for (int k = 0; k < 5; ++k) v[k] = o[k];
}

}  // extern "C"
