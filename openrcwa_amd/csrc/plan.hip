// Batched layer-stack solve: plan construction, execution and the C ABI of
// include/openrcwa_gpu.h.
//
// Pipeline of one execute (every step batched over instances, on one stream):
//   1. convolution matrices of every distinct permittivity grid
//      (FourierSolver2D::solve x3, rcwa/Layer.cpp:99-116), fe33^-1 and the
//      four instance-independent products that P*Q needs
//   2. per chunk of instances:
//      a. P*Q of every patterned layer (rcwa/Layer.cpp:150-220)
//      b. eigen-decomposition (rcwa/WaveEquationSolver.cpp:24-37)
//      c. keff, V = Q W diag(1/keff); homogeneous layers in closed form
//      d. interfaces E = W_a^-1 W_b, H = V_a^-1 V_b (rcwa/RCWASolver.cpp:495-569)
//      e. RTCM cascade from the exit side (rcwa/RCWASolver.cpp:378-492)
//      f. plane-wave response and per-order R/T (rcwa/RCWASolver.cpp:1778-1908)
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "context.h"
#include "rcwa_kernels.h"

using namespace orcwa;

namespace {

constexpr double kPi = 3.1415926535897932;  // rcwa/defns.h:25
constexpr int kMaxPol = 8;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// host-side image of a device blob of tables; resolved after allocation
struct Blob {
  std::vector<char> host;
  template <class T>
  size_t put(const std::vector<T>& v) {
    size_t off = align256(host.size());
    host.resize(off + v.size() * sizeof(T) + 8);
    if (!v.empty()) std::memcpy(host.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

struct StepGroup {
  int count = 0;                 // instances in the group
  // table offsets into the blob
  size_t Wa = 0, Va = 0, Wb = 0, Vb = 0, binv_a = 0, E = 0, H = 0, LU = 0, inst = 0;
  size_t pp = 0, M = 0;          // patterned a: its ProbParam and fe33^-1
};

struct StepTables {
  StepGroup ua, pa, pm;          // a homogeneous / patterned / patterned magnetic
  size_t keff_b = 0, k0d = 0;    // per instance of the chunk
  size_t keff_a = 0;             // per instance: keff of a patterned upper layer, else null
};

struct ChunkTables {
  int i0 = 0, ninst = 0, nprob = 0, nuni = 0;
  int nprob_mag = 0;             // the last nprob_mag problems are magnetic
  size_t mag_mats = 0;           // MagMats of the magnetic problems
  size_t prob_params = 0, inst_of_prob = 0, uni_params = 0, rt_params = 0;
  // patterned problem slots
  size_t pA = 0, pAorig = 0, pW = 0, pV = 0, pZ = 0, pX = 0, pQW = 0;
  size_t pE11 = 0, pE22 = 0;     // grid convolution matrices of each problem
  size_t uV = 0;                 // homogeneous V slots
  // per instance RTCM slots
  size_t Rt[2] = {0, 0};          // R^T of the cascade, alternating by step parity
  size_t E = 0, H = 0;           // interface matrices, then A = E + H, B = E - H
  std::vector<size_t> LU;        // per step: LU of (F + G)^T
  long long prob_base = 0;       // global index of the chunk's first problem
  std::vector<StepTables> steps;
  // host copies for the interfaces of recovered eigen solves: layer of each
  // problem, problem of each (instance, layer) (-1 homogeneous), V and keff
  // of each (instance, layer), H slot of each instance, LU slot of each
  // (instance, step)
  std::vector<int> prob_layer, pidx;
  std::vector<cd*> hV, hK, hH, hLU;
};

}  // namespace

struct orcwa_plan {
  orcwa_gpu_context* ctx = nullptr;
  // every exit of a plan, failed creations included, goes through here
  ~orcwa_plan() {
    if (ctx) --ctx->live_plans;
  }
  int Nx = 0, Ny = 0, nx = 0, ny = 0, h = 0, n = 0, L = 0, count = 0;
  int ncx = 0, ncy = 0, ncell = 0, maxpol = 1;
  double Lx = 0, Ly = 0;
  std::vector<double> cx, cy, thick;
  std::vector<double> ithick;      // [count][L]: each instance's thicknesses
  std::vector<double> lam, kx0, ky0;
  std::vector<int> npol;
  std::vector<cd> pol;             // [count][8][2]
  std::vector<int> gid;            // [count*L]: grid index, -1 homogeneous
  std::vector<int> mgid;           // [count*L]: permeability grid of a magnetic layer, else -1
  std::vector<cd> ueps;            // [count*L]: value of a homogeneous layer
  std::vector<cd> umu;             // [count*L]: its permeability
  std::vector<cd> eps_ref, eps_trn, mu_ref, mu_trn;
  int G = 0;
  std::vector<cd> grid_vals;       // [G*ncell]
  std::vector<int> prob_inst;      // global problem -> instance
  long long nprob_total = 0;

  int chunk = 0, nchunks = 0, pmax = 0, umax = 0;
  std::vector<ChunkTables> chunks;

  // device memory (from the context's cache)
  char* d_ws = nullptr;            // workspace arena
  char* d_blob = nullptr;          // tables
  size_t ws_cap = 0, blob_cap = 0;
  size_t t_ident = 0;              // table {d_ident} for the identity fill
  Blob blob;
  // fixed device regions (pointers inside d_ws)
  double *d_cx = nullptr, *d_cy = nullptr;
  cd* d_grids = nullptr;
  cd* d_strips = nullptr;
  StripTask* d_tasks = nullptr;
  long long* d_strip_base = nullptr;
  int ntasks = 0;
  size_t strip_lds = 0;
  cd* d_ident = nullptr;           // one n x n identity
  // grid matrices: E11, E22, E33, LU, M, S1..S4, tmp  (10 per grid)
  std::vector<cd*> gE11, gE22, gE33, gLU, gM, gS1, gS2, gS3, gS4, gTmp;
  size_t t_gE11 = 0, t_gE22 = 0, t_gE33 = 0, t_gLU = 0, t_gM = 0, t_gS1 = 0, t_gS2 = 0,
         t_gS3 = 0, t_gS4 = 0, t_gTmp = 0, t_gridmats = 0;
  int* d_gipiv = nullptr;
  int* d_ginfo = nullptr;
  // per chunk workspace (shared by all chunks)
  cd* d_lam = nullptr;             // [pmax*n]
  cd* d_keffp = nullptr;           // [pmax*n]
  cd* d_keffu = nullptr;           // [umax*n]
  cd* d_binv = nullptr;            // [umax*4h]
  cd* d_tau = nullptr;             // [pmax*n]
  int* d_ipiv = nullptr;           // [max(chunk, pmax)*n] scratch pivots
  int* d_spiv = nullptr;           // [L-1][chunk*n] pivots of each cascade step
  int* d_perm = nullptr;           // [max(chunk, pmax, G)*n] row permutations of a solve
  cd* d_rtv = nullptr;             // [chunk][4][n] Rud / Tdd columns i0, h+i0
  int* d_pinfo = nullptr;          // [nprob_total]
  double* d_resid = nullptr;       // [nprob_total]
  // outputs
  double* d_ref = nullptr;         // [count][maxpol][h]
  double* d_trn = nullptr;
  cd* d_amp = nullptr;             // [count][8][h]
  int* d_info = nullptr;           // [count]

  // field plans (one instance): per-step interface matrices, the Tdd column
  // chain at the top of every layer, and the instance's layer data
  bool keep = false;
  int gemm_prec = 0;             // GemmPrec of the cascade products (from the context)
  int qr_variant = 0;            // QrVariant and GemmTile: from the context's tuning
  int gemm_tile = 0;             // table for this n (orcwa_plan_autotune)
  // modes-only plans stop after the layer modes (no cascade, no R/T); their
  // layers may each have their own cell edges (one edge set per grid)
  bool modes_only = false;
  bool per_grid_edges = false;
  // P/Q plans (modes-only, one instance): every layer, homogeneous or not,
  // is solved as a patterned problem with P and Q built apart (the mu = 1
  // form of the magnetic builder), and P and Q are kept per layer
  bool keep_pq = false;
  std::vector<cd*> lP, lQ;         // per layer (P/Q plans)
  std::vector<cd*> kP;             // kept copies of P, [L] slots
  std::vector<cd*> kP_src;         // the P slot each copy is taken from
  std::vector<double> grid_cx, grid_cy;  // [G][ncx], [G][ncy] with per-grid edges
  std::vector<cd*> kA, kB;         // [L-1] A = E + H, B = E - H of each cascade step
  cd* kE_src = nullptr;            // the instance's E / H slots
  cd* kH_src = nullptr;
  cd* d_tv = nullptr;              // [L][2][n] chain columns i0, h+i0 at each layer top
  std::vector<cd*> lW, lV, lK;     // per layer: W, V, keff (k0 units)

  hipEvent_t ev[8] = {};
  bool have_events = false;
  int live_graphs = 0;           // orcwa_graph objects that captured this plan
  double phase_ms[6] = {-1, -1, -1, -1, -1, -1};
  double eig_ms_acc = 0.0;

  template <class T>
  T dev(size_t off) const {
    return reinterpret_cast<T>(d_blob + off);
  }
};

namespace {

int fail(const char* msg) {
  fprintf(stderr, "[openrcwa] %s\n", msg);
  return ORCWA_ERR_ARG;
}

ConvGeom geom_of(const orcwa_plan* p) {
  ConvGeom g;
  g.nx = p->nx;
  g.ny = p->ny;
  g.ncx = p->ncx;
  g.ncy = p->ncy;
  g.ncell = p->ncell;
  g.Lx = p->Lx;
  g.Ly = p->Ly;
  g.cx = p->d_cx;
  g.cy = p->d_cy;
  g.sx = p->per_grid_edges ? p->ncx : 0;
  g.sy = p->per_grid_edges ? p->ncy : 0;
  return g;
}

#define HIPCHK(expr)                                                                         \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) {                                                                  \
      fprintf(stderr, "[openrcwa] HIP error %s at %s:%d: %s\n", hipGetErrorName(_e), __FILE__, \
              __LINE__, #expr);                                                              \
      return ORCWA_ERR_HIP;                                                                  \
    }                                                                                        \
  } while (0)

}  // namespace

namespace orcwa {

hipError_t ctx_alloc(orcwa_gpu_context* c, size_t bytes, void** out, size_t* cap) {
  int best = -1;
  for (int i = 0; i < (int)c->cache.size(); ++i)
    if (c->cache[i].bytes >= bytes && (best < 0 || c->cache[i].bytes < c->cache[best].bytes))
      best = i;
  if (best >= 0) {
    *out = c->cache[best].ptr;
    *cap = c->cache[best].bytes;
    c->cached_bytes -= *cap;
    c->cache.erase(c->cache.begin() + best);
    return hipSuccess;
  }
  hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess && !c->cache.empty()) {
    (void)hipGetLastError();
    ctx_trim(c);
    e = hipMalloc(out, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    *cap = 0;
    return e;
  }
  *cap = bytes;
  return hipSuccess;
}

void ctx_release(orcwa_gpu_context* c, void* ptr, size_t cap) {
  if (!ptr) return;
  c->cache.push_back({ptr, cap});
  c->cached_bytes += cap;
  // keep the four largest
  constexpr size_t kKeep = 4;
  while (c->cache.size() > kKeep) {
    int small = 0;
    for (int i = 1; i < (int)c->cache.size(); ++i)
      if (c->cache[i].bytes < c->cache[small].bytes) small = i;
    (void)hipFree(c->cache[small].ptr);
    c->cached_bytes -= c->cache[small].bytes;
    c->cache.erase(c->cache.begin() + small);
  }
}

void ctx_trim(orcwa_gpu_context* c) {
  for (auto& b : c->cache) (void)hipFree(b.ptr);
  c->cache.clear();
  c->cached_bytes = 0;
}

}  // namespace orcwa

extern "C" {

int orcwa_gpu_create(orcwa_gpu_context** ctx, int device) {
  if (!ctx) return fail("orcwa_gpu_create: null context pointer");
  *ctx = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    fprintf(stderr, "[openrcwa] no HIP device available\n");
    return ORCWA_ERR_HIP;
  }
  if (device < 0 || device >= ndev) return fail("orcwa_gpu_create: device index out of range");
  HIPCHK(hipSetDevice(device));
  orcwa_gpu_context* c = new orcwa_gpu_context;
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    fprintf(stderr, "[openrcwa] hipStreamCreate failed\n");
    return ORCWA_ERR_HIP;
  }
  *ctx = c;
  return ORCWA_OK;
}

int orcwa_gpu_set_workspace_limit(orcwa_gpu_context* ctx, size_t bytes) {
  if (!ctx) return fail("orcwa_gpu_set_workspace_limit: null context");
  ctx->workspace_limit = bytes;
  return ORCWA_OK;
}

int orcwa_gpu_set_gemm_precision(orcwa_gpu_context* ctx, int precision) {
  if (!ctx) return fail("orcwa_gpu_set_gemm_precision: null context");
  if (precision < ORCWA_GEMM_FP64 || precision > ORCWA_GEMM_BF16X6)
    return fail("orcwa_gpu_set_gemm_precision: unknown precision");
  ctx->gemm_prec = precision;
  return ORCWA_OK;
}

int orcwa_gpu_destroy(orcwa_gpu_context* ctx) {
  if (!ctx) return ORCWA_OK;
  if (ctx->live_plans > 0)
    return fail("orcwa_gpu_destroy: the context still has live plans (destroy them first)");
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  ctx_trim(ctx);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return ORCWA_OK;
}

int orcwa_plan_destroy(orcwa_plan* p) {
  if (!p) return ORCWA_OK;
  if (p->live_graphs > 0)
    return fail("orcwa_plan_destroy: a graph still uses the plan (destroy the graph first)");
  hipSetDevice(p->ctx->device);
  // the buffers go back to the context once the plan's work is done
  hipStreamSynchronize(p->ctx->stream);
  ctx_release(p->ctx, p->d_ws, p->ws_cap);
  ctx_release(p->ctx, p->d_blob, p->blob_cap);
  if (p->have_events)
    for (auto& e : p->ev) hipEventDestroy(e);
  delete p;
  return ORCWA_OK;
}

static int plan_create(orcwa_gpu_context* ctx, const orcwa_stack* st, const orcwa_instance* inst,
                       int count, bool keep, orcwa_plan** out, bool modes_only = false,
                       const double* layer_cx = nullptr, const double* layer_cy = nullptr,
                       bool keep_pq = false) {
  if (!out) return fail("orcwa_plan_create: null plan pointer");
  *out = nullptr;
  if (!ctx || !st || !inst) return fail("orcwa_plan_create: null argument");
  if (count <= 0) return fail("orcwa_plan_create: count must be positive");
  if (st->nx_harm < 0 || st->ny_harm < 0) return fail("negative harmonic count");
  if (st->n_cx < 2 || st->n_cy < 2 || !st->coord_x || !st->coord_y)
    return fail("each axis needs at least two cell edges");
  if (st->n_layers < 2) return fail("a stack needs at least 2 layers (incidence and exit)");
  if (!st->thickness) return fail("null thickness array");
  const int nx = 2 * st->nx_harm + 1, ny = 2 * st->ny_harm + 1;
  if (nx > 63 || ny > 63) return fail("at most 31 harmonics per axis are supported");
  // the cascade's column chain holds 4 n-vectors in LDS (160 KB per workgroup)
  if (4 * (size_t)(2 * nx * ny) * sizeof(cd) > 160 * 1024)
    return fail("at most 1280 harmonics ((2Nx+1)(2Ny+1)) are supported");
  bool own_edges = (layer_cx != nullptr);
  for (int i = 0; i < count; ++i)
    if (inst[i].coord_x || inst[i].coord_y) own_edges = true;
  // shared edges are strictly increasing; per-layer or per-instance edges are
  // checked one by one below (they may repeat: an empty cell)
  for (int i = 1; i < st->n_cx; ++i)
    if (own_edges ? !(st->coord_x[i] >= st->coord_x[i - 1])
                  : !(st->coord_x[i] > st->coord_x[i - 1]))
      return fail("coord_x must be strictly increasing");
  for (int i = 1; i < st->n_cy; ++i)
    if (own_edges ? !(st->coord_y[i] >= st->coord_y[i - 1])
                  : !(st->coord_y[i] > st->coord_y[i - 1]))
      return fail("coord_y must be strictly increasing");
  if (!(st->coord_x[st->n_cx - 1] > st->coord_x[0]) || !(st->coord_y[st->n_cy - 1] > st->coord_y[0]))
    return fail("the periods must be positive");

  if (keep && count != 1) return fail("a field plan holds one instance");
  if ((layer_cx || layer_cy || modes_only) && count != 1)
    return fail("a modes-only plan holds one instance");
  hipSetDevice(ctx->device);
  orcwa_plan* p = new orcwa_plan;
  p->ctx = ctx;
  p->gemm_prec = ctx->gemm_prec;
  ++ctx->live_plans;
  p->keep = keep;
  p->modes_only = modes_only;
  p->keep_pq = keep_pq && modes_only && keep;
  p->Nx = st->nx_harm;
  p->Ny = st->ny_harm;
  p->nx = nx;
  p->ny = ny;
  p->h = nx * ny;
  p->n = 2 * p->h;
  {
    auto t = ctx->tuning.find(p->n);
    if (t != ctx->tuning.end()) {
      p->qr_variant = t->second.qr;
      p->gemm_tile = t->second.tile;
    }
  }
  p->L = st->n_layers;
  p->count = count;
  p->ncx = st->n_cx;
  p->ncy = st->n_cy;
  p->ncell = (p->ncx - 1) * (p->ncy - 1);
  p->cx.assign(st->coord_x, st->coord_x + p->ncx);
  p->cy.assign(st->coord_y, st->coord_y + p->ncy);
  p->Lx = p->cx.back() - p->cx.front();
  p->Ly = p->cy.back() - p->cy.front();
  p->thick.assign(st->thickness, st->thickness + p->L);
  for (int i = 0; i < count; ++i) {
    const double* t = inst[i].thickness ? inst[i].thickness : st->thickness;
    p->ithick.insert(p->ithick.end(), t, t + p->L);
  }
  const int L = p->L, h = p->h, n = p->n, ncell = p->ncell;

  // ---- instances: classify layers, dedupe grids.  A layer is homogeneous
  // when eps and mu are uniform, a (non-magnetic) problem when eps is
  // patterned and mu == 1, and a magnetic problem otherwise
  // (rcwa/Layer.cpp:66-72, 129-145); eps and mu grids share one list.
  std::map<std::vector<double>, int> grid_index;
  p->lam.resize(count);
  p->kx0.resize(count);
  p->ky0.resize(count);
  p->npol.resize(count);
  p->pol.assign((size_t)count * kMaxPol * 2, mk(0, 0));
  p->gid.assign((size_t)count * L, -1);
  p->mgid.assign((size_t)count * L, -1);
  p->ueps.assign((size_t)count * L, mk(0, 0));
  p->umu.assign((size_t)count * L, mk(1, 0));
  p->eps_ref.resize(count);
  p->eps_trn.resize(count);
  p->mu_ref.resize(count);
  p->mu_trn.resize(count);
  // a grid is its values (and, with per-layer edges, the layer's edges)
  // instances may bring their own cell edges (same counts and periods)
  bool inst_edges = false;
  for (int i = 0; i < count; ++i)
    if (inst[i].coord_x || inst[i].coord_y) inst_edges = true;
  if (inst_edges && layer_cx) {
    delete p;
    return fail("per-layer and per-instance edges cannot be combined");
  }
  p->per_grid_edges = (layer_cx != nullptr) || inst_edges;
  auto check_edges = [&](const double* x, const double* y) {
    for (int i = 1; i < p->ncx; ++i)
      if (!(x[i] >= x[i - 1])) return false;
    for (int i = 1; i < p->ncy; ++i)
      if (!(y[i] >= y[i - 1])) return false;
    return std::fabs((x[p->ncx - 1] - x[0]) - p->Lx) <= 1e-12 * p->Lx &&
           std::fabs((y[p->ncy - 1] - y[0]) - p->Ly) <= 1e-12 * p->Ly;
  };
  // a grid is its values (and, with per-grid edges, its edges)
  auto grid_of = [&](const double* v, const double* ex, const double* ey) {
    std::vector<double> key(v, v + 2 * ncell);
    if (p->per_grid_edges) {
      key.insert(key.end(), ex, ex + p->ncx);
      key.insert(key.end(), ey, ey + p->ncy);
    }
    auto it = grid_index.find(key);
    if (it != grid_index.end()) return it->second;
    const int g = (int)grid_index.size();
    grid_index.emplace(std::move(key), g);
    for (int c = 0; c < ncell; ++c) p->grid_vals.push_back(mk(v[2 * c], v[2 * c + 1]));
    if (p->per_grid_edges) {
      p->grid_cx.insert(p->grid_cx.end(), ex, ex + p->ncx);
      p->grid_cy.insert(p->grid_cy.end(), ey, ey + p->ncy);
    }
    return g;
  };
  if (layer_cx && !layer_cy) {
    delete p;
    return fail("per-layer edges need both axes");
  }
  // the edges of (instance i, layer l); (a moving edge may meet its
  // neighbour: an empty cell adds nothing to the Fourier series)
  auto ex_of = [&](int i, int l) -> const double* {
    if (layer_cx) return layer_cx + (size_t)l * p->ncx;
    return inst[i].coord_x ? inst[i].coord_x : st->coord_x;
  };
  auto ey_of = [&](int i, int l) -> const double* {
    if (layer_cy) return layer_cy + (size_t)l * p->ncy;
    return inst[i].coord_y ? inst[i].coord_y : st->coord_y;
  };
  if (layer_cx)
    for (int l = 0; l < L; ++l)
      if (!check_edges(ex_of(0, l), ey_of(0, l))) {
        delete p;
        return fail("layer edges must be sorted, with the period of the first layer");
      }
  for (int i = 0; i < count && inst_edges; ++i)
    if (!check_edges(ex_of(i, 0), ey_of(i, 0))) {
      delete p;
      return fail("instance edges must be sorted, with the stack's period");
    }
  auto uniform_of = [&](const double* v) {
    for (int c = 1; c < ncell; ++c)
      if (v[2 * c] != v[0] || v[2 * c + 1] != v[1]) return false;
    return true;
  };
  auto has_zero = [&](const double* v) {
    for (int c = 0; c < ncell; ++c)
      if (v[2 * c] == 0.0 && v[2 * c + 1] == 0.0) return true;
    return false;
  };
  for (int i = 0; i < count; ++i) {
    const orcwa_instance& I = inst[i];
    if (!(I.lambda > 0) || !std::isfinite(I.lambda)) {
      delete p;
      return fail("instance wavelength must be positive");
    }
    if (!I.eps || I.n_pol < 1 || I.n_pol > kMaxPol || !I.pol) {
      delete p;
      return fail("instance needs eps and 1..8 polarisations");
    }
    p->lam[i] = I.lambda;
    p->kx0[i] = I.kx0;
    p->ky0[i] = I.ky0;
    p->npol[i] = I.n_pol;
    p->maxpol = std::max(p->maxpol, I.n_pol);
    for (int q = 0; q < I.n_pol; ++q) {
      p->pol[((size_t)i * kMaxPol + q) * 2 + 0] = mk(I.pol[4 * q + 0], I.pol[4 * q + 1]);
      p->pol[((size_t)i * kMaxPol + q) * 2 + 1] = mk(I.pol[4 * q + 2], I.pol[4 * q + 3]);
    }
    for (int l = 0; l < L; ++l) {
      const double* e = I.eps + (size_t)l * ncell * 2;
      const double* m = I.mu ? I.mu + (size_t)l * ncell * 2 : nullptr;
      if (has_zero(e) || (m && has_zero(m))) {
        delete p;
        return fail("zero permittivity or permeability");
      }
      const bool mu_one = !m || (uniform_of(m) && m[0] == 1.0 && m[1] == 0.0);
      const size_t il = (size_t)i * L + l;
      if (p->keep_pq) {
        // P and Q apart: the magnetic builder with a mu = 1 grid
        std::vector<double> ones(2 * (size_t)ncell, 0.0);
        for (int c = 0; c < ncell; ++c) ones[2 * c] = 1.0;
        p->gid[il] = grid_of(e, ex_of(i, l), ey_of(i, l));
        p->mgid[il] = grid_of(m ? m : ones.data(), ex_of(i, l), ey_of(i, l));
      } else if (uniform_of(e) && (!m || uniform_of(m))) {
        p->ueps[il] = mk(e[0], e[1]);
        if (m) p->umu[il] = mk(m[0], m[1]);
      } else {
        p->gid[il] = grid_of(e, ex_of(i, l), ey_of(i, l));
        if (!mu_one) p->mgid[il] = grid_of(m, ex_of(i, l), ey_of(i, l));
      }
    }
    if (!modes_only && (p->gid[(size_t)i * L] >= 0 || p->gid[(size_t)i * L + L - 1] >= 0)) {
      delete p;
      fprintf(stderr,
              "[openrcwa] the incidence and exit half-spaces must be homogeneous (instance %d)\n",
              i);
      return ORCWA_ERR_UNSUPPORTED;
    }
    p->eps_ref[i] = p->ueps[(size_t)i * L];
    p->eps_trn[i] = p->ueps[(size_t)i * L + L - 1];
    p->mu_ref[i] = p->umu[(size_t)i * L];
    p->mu_trn[i] = p->umu[(size_t)i * L + L - 1];
  }
  p->G = (int)grid_index.size();

  // ---- chunking by a workspace budget
  const size_t slot = (size_t)n * n * sizeof(cd);
  int lp_max = 0, lu_max = 0;
  for (int i = 0; i < count; ++i) {
    int a = 0, b = 0;
    for (int l = 0; l < L; ++l) (p->gid[(size_t)i * L + l] >= 0 ? a : b)++;
    lp_max = std::max(lp_max, a);
    lu_max = std::max(lu_max, b);
  }
  // workspace budget: the context's limit (orcwa_gpu_set_workspace_limit),
  // else ORCWA_WORKSPACE_GB, else 70% of the device memory free now (the
  // per-instance slots dominate; a few large chunks keep every batched kernel
  // wide and pay the end-of-launch tail few times)
  double budget_gb = (double)ctx->workspace_limit / 1e9;
  if (!(budget_gb > 0.0))
    if (const char* s = getenv("ORCWA_WORKSPACE_GB")) budget_gb = atof(s);
  if (!(budget_gb > 0.0)) {
    size_t fr = 0, tot = 0;
    // (buffers cached by the context count as free)
    budget_gb = (hipMemGetInfo(&fr, &tot) == hipSuccess)
                    ? 0.7 * (double)(fr + ctx->cached_bytes) / 1e9
                    : 8.0;
  }
  const int islots = L + 3;      // Rt x2, E, H, one LU per cascade step
  const size_t per_inst = (size_t)(7 * lp_max + lu_max + islots) * slot;
  // memory that does not scale with the chunk: per distinct grid the ten h x h
  // convolution products, per instance the outputs
  const double fixed =
      (double)std::max<size_t>(1, grid_index.size()) * (10.0 * h * h + 4.0 * h) * sizeof(cd) +
      (double)count * ((2.0 * kMaxPol + 16.0) * h * sizeof(double) + 64.0);
  long long c_fit = (long long)((budget_gb * 1e9 - fixed) / (double)per_inst);
  c_fit = std::max(1LL, std::min((long long)count, c_fit));
  // equal chunks
  const long long nck = (count + c_fit - 1) / c_fit;
  int chunk = (int)((count + nck - 1) / nck);
  if (const char* s = getenv("ORCWA_CHUNK")) chunk = std::max(1, std::min(count, atoi(s)));
  p->chunk = chunk;
  p->nchunks = (count + chunk - 1) / chunk;

  // ---- problems per chunk
  p->pmax = 0;
  p->umax = 0;
  for (int c = 0; c < p->nchunks; ++c) {
    const int i0 = c * chunk, i1 = std::min(count, i0 + chunk);
    int np = 0, nu = 0;
    for (int i = i0; i < i1; ++i)
      for (int l = 0; l < L; ++l) (p->gid[(size_t)i * L + l] >= 0 ? np : nu)++;
    p->pmax = std::max(p->pmax, np);
    p->umax = std::max(p->umax, nu);
  }

  // ---- workspace layout
  const int G = p->G;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = align256(off);
    off = o + bytes;
    return o;
  };
  const size_t nes = p->per_grid_edges ? (size_t)std::max(1, p->G) : 1;
  const size_t o_cx = take(nes * p->ncx * sizeof(double));
  const size_t o_cy = take(nes * p->ncy * sizeof(double));
  const size_t o_grids = take((size_t)std::max(1, G) * ncell * sizeof(cd));
  // strips: type-0 records per grid (ncy-1 of nx^2+My+Mx), type-1 (ncx-1 of ny^2+Mx)
  const long long rec0 = (long long)nx * nx + (2 * ny - 1) + (2 * nx - 1);
  const long long rec1 = (long long)ny * ny + (2 * nx - 1);
  const long long per_grid_strip = (p->ncy - 1) * rec0 + (p->ncx - 1) * rec1;
  const size_t o_strips = take((size_t)std::max(1, G) * per_grid_strip * sizeof(cd));
  std::vector<StripTask> tasks;
  std::vector<long long> strip_base;
  for (int g = 0; g < G; ++g) {
    long long b0 = g * per_grid_strip, b1 = b0 + (p->ncy - 1) * rec0;
    strip_base.push_back(b0);
    strip_base.push_back(b1);
    for (int i = 0; i < p->ncy - 1; ++i) tasks.push_back({g, 0, i, 0, b0 + i * rec0});
    for (int i = 0; i < p->ncx - 1; ++i) tasks.push_back({g, 1, i, 0, b1 + i * rec1});
  }
  p->ntasks = (int)tasks.size();
  const int mdim = std::max(nx, ny);
  p->strip_lds = (size_t)(2 * mdim - 1 + mdim * mdim + std::max(p->ncx, p->ncy)) * sizeof(cd);
  const size_t o_tasks = take(std::max<size_t>(1, tasks.size()) * sizeof(StripTask));
  const size_t o_sbase = take(std::max<size_t>(1, strip_base.size()) * sizeof(long long));
  const size_t o_ident = take(slot);
  const size_t gslot = (size_t)h * h * sizeof(cd);
  const size_t o_gm = take((size_t)std::max(1, G) * 10 * gslot);
  const size_t o_gipiv = take((size_t)std::max(1, G) * h * sizeof(int));
  const size_t o_ginfo = take((size_t)std::max(1, G) * sizeof(int));
  const int PM = std::max(1, p->pmax), UM = std::max(1, p->umax);
  const size_t o_pslots = take((size_t)PM * 7 * slot);
  const size_t o_uslots = take((size_t)UM * slot);
  const size_t o_islots = take((size_t)chunk * islots * slot);
  const size_t o_spiv = take((size_t)(L - 1) * chunk * n * sizeof(int));
  const size_t o_rtv = take((size_t)chunk * 4 * n * sizeof(cd));
  const size_t o_perm = take((size_t)std::max(std::max(chunk, PM), std::max(1, G)) * n * sizeof(int));
  const size_t o_lam = take((size_t)PM * n * sizeof(cd));
  const size_t o_keffp = take((size_t)PM * n * sizeof(cd));
  const size_t o_keffu = take((size_t)UM * n * sizeof(cd));
  const size_t o_binv = take((size_t)UM * 4 * h * sizeof(cd));
  const size_t o_tau = take((size_t)PM * n * sizeof(cd));
  const size_t o_ipiv = take((size_t)std::max(chunk, PM) * n * sizeof(int));
  p->nprob_total = 0;
  for (int i = 0; i < count; ++i)
    for (int l = 0; l < L; ++l)
      if (p->gid[(size_t)i * L + l] >= 0) ++p->nprob_total;
  const size_t o_pinfo = take((size_t)std::max<long long>(1, p->nprob_total) * sizeof(int));
  const size_t o_resid = take((size_t)std::max<long long>(1, p->nprob_total) * sizeof(double));
  const size_t o_ref = take((size_t)count * p->maxpol * h * sizeof(double));
  const size_t o_trn = take((size_t)count * p->maxpol * h * sizeof(double));
  const size_t o_amp = take((size_t)count * 8 * h * sizeof(cd));
  const size_t o_info = take((size_t)count * sizeof(int));
  const bool keep_cascade = keep && !modes_only;
  const size_t o_keep =
      take(keep_cascade ? 2 * (size_t)(L - 1) * slot + (size_t)L * 2 * n * sizeof(cd) : 0);
  const size_t o_kp = take(p->keep_pq ? (size_t)L * slot : 0);
  const size_t ws_bytes = align256(off);

  if (ctx_alloc(ctx, ws_bytes, (void**)&p->d_ws, &p->ws_cap) != hipSuccess) {
    delete p;
    fprintf(stderr, "[openrcwa] could not allocate %.2f GB of device workspace\n", ws_bytes / 1e9);
    return ORCWA_ERR_NOMEM;
  }
  char* W = p->d_ws;
  p->d_cx = (double*)(W + o_cx);
  p->d_cy = (double*)(W + o_cy);
  p->d_grids = (cd*)(W + o_grids);
  p->d_strips = (cd*)(W + o_strips);
  p->d_tasks = (StripTask*)(W + o_tasks);
  p->d_strip_base = (long long*)(W + o_sbase);
  p->d_ident = (cd*)(W + o_ident);
  p->d_gipiv = (int*)(W + o_gipiv);
  p->d_ginfo = (int*)(W + o_ginfo);
  p->d_lam = (cd*)(W + o_lam);
  p->d_keffp = (cd*)(W + o_keffp);
  p->d_keffu = (cd*)(W + o_keffu);
  p->d_binv = (cd*)(W + o_binv);
  p->d_tau = (cd*)(W + o_tau);
  p->d_ipiv = (int*)(W + o_ipiv);
  p->d_spiv = (int*)(W + o_spiv);
  p->d_rtv = (cd*)(W + o_rtv);
  p->d_perm = (int*)(W + o_perm);
  p->d_pinfo = (int*)(W + o_pinfo);
  p->d_resid = (double*)(W + o_resid);
  p->d_ref = (double*)(W + o_ref);
  p->d_trn = (double*)(W + o_trn);
  p->d_amp = (cd*)(W + o_amp);
  p->d_info = (int*)(W + o_info);
  if (p->keep_pq)
    for (int l = 0; l < L; ++l) p->kP.push_back((cd*)(W + o_kp + (size_t)l * slot));
  if (keep_cascade) {
    for (int si = 0; si < L - 1; ++si) {
      p->kA.push_back((cd*)(W + o_keep + (size_t)(2 * si) * slot));
      p->kB.push_back((cd*)(W + o_keep + (size_t)(2 * si + 1) * slot));
    }
    p->d_tv = (cd*)(W + o_keep + 2 * (size_t)(L - 1) * slot);
  }

  // constant inputs
  hipStream_t s = ctx->stream;
  if (p->per_grid_edges && p->G > 0) {
    HIPCHK(hipMemcpyAsync(p->d_cx, p->grid_cx.data(), p->grid_cx.size() * sizeof(double),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(p->d_cy, p->grid_cy.data(), p->grid_cy.size() * sizeof(double),
                          hipMemcpyHostToDevice, s));
  } else {
    HIPCHK(hipMemcpyAsync(p->d_cx, p->cx.data(), p->ncx * sizeof(double), hipMemcpyHostToDevice,
                          s));
    HIPCHK(hipMemcpyAsync(p->d_cy, p->cy.data(), p->ncy * sizeof(double), hipMemcpyHostToDevice,
                          s));
  }
  if (G > 0) {
    HIPCHK(hipMemcpyAsync(p->d_grids, p->grid_vals.data(), (size_t)G * ncell * sizeof(cd),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(p->d_tasks, tasks.data(), tasks.size() * sizeof(StripTask),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(p->d_strip_base, strip_base.data(), strip_base.size() * sizeof(long long),
                          hipMemcpyHostToDevice, s));
  }

  // ---- tables
  Blob& B = p->blob;
  auto gmat = [&](int g, int k) { return (cd*)(W + o_gm + ((size_t)g * 10 + k) * gslot); };
  for (int g = 0; g < G; ++g) {
    p->gE11.push_back(gmat(g, 0));
    p->gE22.push_back(gmat(g, 1));
    p->gE33.push_back(gmat(g, 2));
    p->gLU.push_back(gmat(g, 3));
    p->gM.push_back(gmat(g, 4));
    p->gS1.push_back(gmat(g, 5));
    p->gS2.push_back(gmat(g, 6));
    p->gS3.push_back(gmat(g, 7));
    p->gS4.push_back(gmat(g, 8));
    p->gTmp.push_back(gmat(g, 9));
  }
  p->t_gE11 = B.put(p->gE11);
  p->t_gE22 = B.put(p->gE22);
  p->t_gE33 = B.put(p->gE33);
  p->t_gLU = B.put(p->gLU);
  p->t_gM = B.put(p->gM);
  p->t_gS1 = B.put(p->gS1);
  p->t_gS2 = B.put(p->gS2);
  p->t_gS3 = B.put(p->gS3);
  p->t_gS4 = B.put(p->gS4);
  p->t_gTmp = B.put(p->gTmp);
  {
    std::vector<GridMats> gm(G);
    for (int g = 0; g < G; ++g)
      gm[g] = {p->gE11[g], p->gE22[g], p->gS1[g], p->gS2[g], p->gS3[g], p->gS4[g]};
    p->t_gridmats = B.put(gm);
  }

  auto pslot = [&](int pi, int k) { return (cd*)(W + o_pslots + ((size_t)pi * 7 + k) * slot); };
  auto uslot = [&](int ui) { return (cd*)(W + o_uslots + (size_t)ui * slot); };
  auto islot = [&](int ii, int k) {
    return (cd*)(W + o_islots + ((size_t)ii * islots + k) * slot);
  };

  long long prob_global = 0;
  for (int c = 0; c < p->nchunks; ++c) {
    ChunkTables ct;
    ct.i0 = c * chunk;
    ct.ninst = std::min(count, ct.i0 + chunk) - ct.i0;
    ct.prob_base = prob_global;
    // (instance, layer) -> problem index or homogeneous index
    std::vector<int> pidx((size_t)ct.ninst * L, -1), uidx((size_t)ct.ninst * L, -1);
    std::vector<ProbParam> pp;
    std::vector<int> inst_of;
    std::vector<UniParam> up;
    // non-magnetic problems first, then the magnetic ones (each group is a
    // contiguous range of the problem tables)
    for (int pass = 0; pass < 2; ++pass)
      for (int ii = 0; ii < ct.ninst; ++ii) {
        const int i = ct.i0 + ii;
        const double inv_k0 = p->lam[i] / (2.0 * kPi);
        for (int l = 0; l < L; ++l) {
          const size_t il = (size_t)i * L + l;
          const int g = p->gid[il];
          if (g >= 0) {
            if ((p->mgid[il] >= 0) != (pass == 1)) continue;
            pidx[(size_t)ii * L + l] = (int)pp.size();
            ct.prob_layer.push_back(l);
            ProbParam q;
            q.kx0 = p->kx0[i];
            q.ky0 = p->ky0[i];
            q.inv_k0 = inv_k0;
            q.grid = g;
            q.mgrid = p->mgid[il];
            pp.push_back(q);
            inst_of.push_back(i);
            p->prob_inst.push_back(i);
            if (pass == 1) ++ct.nprob_mag;
          } else if (pass == 0) {
            uidx[(size_t)ii * L + l] = (int)up.size();
            UniParam u;
            u.eps = p->ueps[il];
            u.mu = p->umu[il];
            u.kx0 = p->kx0[i];
            u.ky0 = p->ky0[i];
            u.inv_k0 = inv_k0;
            u.inst = i;
            u.pad = 0;
            up.push_back(u);
          }
        }
      }
    ct.nprob = (int)pp.size();
    ct.nuni = (int)up.size();
    {
      std::vector<MagMats> mm;
      for (const ProbParam& q : pp)
        if (q.mgrid >= 0)
          mm.push_back({p->gE11[q.grid], p->gE22[q.grid], p->gM[q.grid], p->gE11[q.mgrid],
                        p->gE22[q.mgrid], p->gM[q.mgrid]});
      ct.mag_mats = B.put(mm);
    }
    prob_global += ct.nprob;
    ct.prob_params = B.put(pp);
    ct.inst_of_prob = B.put(inst_of);
    ct.uni_params = B.put(up);
    std::vector<cd*> tA, tAo, tW, tV, tZ, tX, tQW, tUV;
    for (int pi = 0; pi < ct.nprob; ++pi) {
      tA.push_back(pslot(pi, 0));
      tAo.push_back(pslot(pi, 1));
      tW.push_back(pslot(pi, 2));
      tV.push_back(pslot(pi, 3));
      tZ.push_back(pslot(pi, 4));
      tX.push_back(pslot(pi, 5));
      tQW.push_back(pslot(pi, 6));
    }
    for (int ui = 0; ui < ct.nuni; ++ui) tUV.push_back(uslot(ui));
    ct.pA = B.put(tA);
    ct.pAorig = B.put(tAo);
    ct.pW = B.put(tW);
    ct.pV = B.put(tV);
    ct.pZ = B.put(tZ);
    ct.pX = B.put(tX);
    ct.pQW = B.put(tQW);
    ct.uV = B.put(tUV);
    {
      std::vector<cd*> e11, e22;
      for (int pi = 0; pi < ct.nprob; ++pi) {
        e11.push_back(p->gE11[pp[pi].grid]);
        e22.push_back(p->gE22[pp[pi].grid]);
      }
      ct.pE11 = B.put(e11);
      ct.pE22 = B.put(e22);
    }
    {
      std::vector<cd*> tR0, tR1, tE, tH;
      for (int ii = 0; ii < ct.ninst; ++ii) {
        tR0.push_back(islot(ii, 0));
        tR1.push_back(islot(ii, 1));
        tE.push_back(islot(ii, 2));
        tH.push_back(islot(ii, 3));
      }
      ct.Rt[0] = B.put(tR0);
      ct.Rt[1] = B.put(tR1);
      ct.E = B.put(tE);
      ct.H = B.put(tH);
      for (int si = 0; si < L - 1; ++si) {
        std::vector<cd*> tLU;
        for (int ii = 0; ii < ct.ninst; ++ii) tLU.push_back(islot(ii, 4 + si));
        ct.LU.push_back(B.put(tLU));
      }
    }
    // layer references
    auto Wof = [&](int ii, int l) -> cd* {
      int q = pidx[(size_t)ii * L + l];
      return q >= 0 ? pslot(q, 2) : p->d_ident;
    };
    auto Vof = [&](int ii, int l) -> cd* {
      int q = pidx[(size_t)ii * L + l];
      return q >= 0 ? pslot(q, 3) : uslot(uidx[(size_t)ii * L + l]);
    };
    auto Kof = [&](int ii, int l) -> cd* {
      int q = pidx[(size_t)ii * L + l];
      return q >= 0 ? p->d_keffp + (size_t)q * n : p->d_keffu + (size_t)uidx[(size_t)ii * L + l] * n;
    };
    auto Bof = [&](int ii, int l) -> cd* {
      int u = uidx[(size_t)ii * L + l];
      return u >= 0 ? p->d_binv + (size_t)u * 4 * h : nullptr;
    };
    ct.pidx.assign(pidx.begin(), pidx.end());
    for (int ii = 0; ii < ct.ninst; ++ii) {
      for (int l = 0; l < L; ++l) {
        ct.hV.push_back(Vof(ii, l));
        ct.hK.push_back(Kof(ii, l));
      }
      ct.hH.push_back(islot(ii, 3));
      for (int si = 0; si < L - 1; ++si) ct.hLU.push_back(islot(ii, 4 + si));
    }
    if (keep) {
      p->kE_src = islot(0, 2);
      p->kH_src = islot(0, 3);
      for (int l = 0; l < L; ++l) {
        p->lW.push_back(Wof(0, l));
        p->lV.push_back(Vof(0, l));
        p->lK.push_back(Kof(0, l));
        if (p->keep_pq) {
          // P is built into the V slot (and later replaced by V); Q stays in
          // the QW slot
          const int q = pidx[(size_t)l];
          p->kP_src.push_back(pslot(q, 3));
          p->lP.push_back(p->kP[l]);
          p->lQ.push_back(pslot(q, 6));
        }
      }
    }
    for (int sidx = 0; sidx < L - 1; ++sidx) {
      StepTables stt;
      const int idx = L - 1 - sidx, a = idx - 1, b = idx;
      std::vector<cd*> kb;
      std::vector<double> k0d;
      std::vector<cd*> uaWb, uaVb, uaBinv, uaE, uaH;
      std::vector<cd*> paWa, paWb, paVb, paE, paH, paLU, paM, ka;
      std::vector<cd*> pmWa, pmVa, pmWb, pmVb, pmE, pmH, pmLU;
      std::vector<ProbParam> paPP;
      std::vector<int> uaInst, paInst, pmInst;
      for (int ii = 0; ii < ct.ninst; ++ii) {
        const int i = ct.i0 + ii;
        kb.push_back(Kof(ii, b));
        k0d.push_back(2.0 * kPi / p->lam[i] * p->ithick[(size_t)i * L + b]);
        const bool a_mag = pidx[(size_t)ii * L + a] >= 0 && p->mgid[(size_t)i * L + a] >= 0;
        ka.push_back(pidx[(size_t)ii * L + a] >= 0 && !a_mag ? Kof(ii, a) : nullptr);
        if (a_mag) {
          pmWa.push_back(Wof(ii, a));
          pmVa.push_back(Vof(ii, a));
          pmWb.push_back(Wof(ii, b));
          pmVb.push_back(Vof(ii, b));
          pmE.push_back(islot(ii, 2));
          pmH.push_back(islot(ii, 3));
          pmLU.push_back(islot(ii, 4 + sidx));
          pmInst.push_back(i);
        } else if (pidx[(size_t)ii * L + a] < 0) {
          uaWb.push_back(Wof(ii, b));
          uaVb.push_back(Vof(ii, b));
          uaBinv.push_back(Bof(ii, a));
          uaE.push_back(islot(ii, 2));
          uaH.push_back(islot(ii, 3));
          uaInst.push_back(i);
        } else {
          paWa.push_back(Wof(ii, a));
          paWb.push_back(Wof(ii, b));
          paVb.push_back(Vof(ii, b));
          paE.push_back(islot(ii, 2));
          paH.push_back(islot(ii, 3));
          paLU.push_back(islot(ii, 4 + sidx));
          paInst.push_back(i);
          const ProbParam& qa = pp[pidx[(size_t)ii * L + a]];
          paPP.push_back(qa);
          paM.push_back(p->gM[qa.grid]);
        }
      }
      stt.keff_b = B.put(kb);
      stt.k0d = B.put(k0d);
      stt.keff_a = B.put(ka);
      stt.ua.count = (int)uaInst.size();
      stt.ua.Wb = B.put(uaWb);
      stt.ua.Vb = B.put(uaVb);
      stt.ua.binv_a = B.put(uaBinv);
      stt.ua.E = B.put(uaE);
      stt.ua.H = B.put(uaH);
      stt.ua.inst = B.put(uaInst);
      stt.pa.count = (int)paInst.size();
      stt.pa.Wa = B.put(paWa);
      stt.pa.Wb = B.put(paWb);
      stt.pa.Vb = B.put(paVb);
      stt.pa.E = B.put(paE);
      stt.pa.H = B.put(paH);
      stt.pa.LU = B.put(paLU);
      stt.pa.inst = B.put(paInst);
      stt.pa.pp = B.put(paPP);
      stt.pa.M = B.put(paM);
      stt.pm.count = (int)pmInst.size();
      stt.pm.Wa = B.put(pmWa);
      stt.pm.Va = B.put(pmVa);
      stt.pm.Wb = B.put(pmWb);
      stt.pm.Vb = B.put(pmVb);
      stt.pm.E = B.put(pmE);
      stt.pm.H = B.put(pmH);
      stt.pm.LU = B.put(pmLU);
      stt.pm.inst = B.put(pmInst);
      ct.steps.push_back(stt);
    }
    std::vector<RtParam> rp(ct.ninst);
    for (int ii = 0; ii < ct.ninst; ++ii) {
      const int i = ct.i0 + ii;
      RtParam& r = rp[ii];
      std::memset(&r, 0, sizeof(r));
      r.eps_ref = p->eps_ref[i];
      r.eps_trn = p->eps_trn[i];
      r.mu_ref = p->mu_ref[i];
      r.mu_trn = p->mu_trn[i];
      r.kx0 = p->kx0[i];
      r.ky0 = p->ky0[i];
      r.inv_k0 = p->lam[i] / (2.0 * kPi);
      r.npol = p->npol[i];
      for (int q = 0; q < kMaxPol; ++q) {
        r.pol[q][0] = p->pol[((size_t)i * kMaxPol + q) * 2 + 0];
        r.pol[q][1] = p->pol[((size_t)i * kMaxPol + q) * 2 + 1];
      }
    }
    ct.rt_params = B.put(rp);
    p->chunks.push_back(std::move(ct));
  }
  p->t_ident = B.put(std::vector<cd*>{p->d_ident});
  if (ctx_alloc(ctx, std::max<size_t>(256, B.host.size()), (void**)&p->d_blob, &p->blob_cap) !=
      hipSuccess) {
    orcwa_plan_destroy(p);
    return ORCWA_ERR_NOMEM;
  }
  HIPCHK(hipMemcpyAsync(p->d_blob, B.host.data(), B.host.size(), hipMemcpyHostToDevice, s));
  // identity matrix
  set_identity_batched(s, n, p->dev<cd* const*>(p->t_ident), 1);
  for (auto& e : p->ev) HIPCHK(hipEventCreate(&e));
  p->have_events = true;
  HIPCHK(hipStreamSynchronize(s));
  *out = p;
  return ORCWA_OK;
}

int orcwa_plan_create(orcwa_gpu_context* ctx, const orcwa_stack* st, const orcwa_instance* inst,
                      int count, orcwa_plan** out) {
  return plan_create(ctx, st, inst, count, false, out);
}

// The reference's recovery of a failed eigen solve
// (rcwa/MKLEigenSolver.hpp:160-176): problems of the chunk whose residual
// test failed (or whose QR did not converge) are solved again on
// A + diag(1e-11 (i+1) max|A|) and tested against the unperturbed A.  The
// status bits are read back once per chunk; the recovery itself is rare.
static int eig_recover(orcwa_plan* p, const ChunkTables& ct, hipStream_t s,
                       std::vector<int>& recovered) {
  const int np = ct.nprob, n = p->n;
  recovered.clear();
  if (np <= 0) return ORCWA_OK;
  std::vector<int> pin(np);
  HIPCHK(hipMemcpyAsync(pin.data(), p->d_pinfo + ct.prob_base, np * sizeof(int),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  std::vector<int> bad;
  for (int q = 0; q < np; ++q)
    if (pin[q] & (ORCWA_STATUS_EIG_RESIDUAL | ORCWA_STATUS_NOCONV)) bad.push_back(q);
  if (bad.empty()) return ORCWA_OK;
  recovered = bad;
  const int m = (int)bad.size();
  // the problems' slots, read from the host image of the plan's tables
  auto slot = [&](size_t off, int q) {
    return reinterpret_cast<cd* const*>(p->blob.host.data() + off)[q];
  };
  std::vector<cd*> tabs((size_t)5 * m);
  for (int k = 0; k < m; ++k) {
    tabs[k] = slot(ct.pA, bad[k]);
    tabs[m + k] = slot(ct.pAorig, bad[k]);
    tabs[2 * m + k] = slot(ct.pW, bad[k]);
    tabs[3 * m + k] = slot(ct.pZ, bad[k]);
    tabs[4 * m + k] = slot(ct.pX, bad[k]);
  }
  const size_t o_idx = align256(tabs.size() * sizeof(cd*));
  const size_t o_lam = align256(o_idx + m * sizeof(int));
  const size_t o_info = align256(o_lam + (size_t)m * n * sizeof(cd));
  const size_t o_res = align256(o_info + m * sizeof(int));
  const size_t bytes = o_res + m * sizeof(double);
  char* buf = nullptr;
  if (hipMalloc(&buf, bytes) != hipSuccess) return ORCWA_ERR_NOMEM;
  std::vector<char> img(o_lam, 0);
  std::memcpy(img.data(), tabs.data(), tabs.size() * sizeof(cd*));
  std::memcpy(img.data() + o_idx, bad.data(), m * sizeof(int));
  hipError_t e = hipMemcpyAsync(buf, img.data(), img.size(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(buf + o_info, 0, m * sizeof(int), s);
  if (e == hipSuccess) {
    cd* const* tA = reinterpret_cast<cd* const*>(buf);
    const cd* const* tAo = reinterpret_cast<const cd* const*>(buf) + m;
    cd* const* tW = reinterpret_cast<cd* const*>(buf) + 2 * m;
    EigWorkspace ws;
    ws.Z = reinterpret_cast<cd**>(buf) + 3 * m;
    ws.X = reinterpret_cast<cd**>(buf) + 4 * m;
    ws.tau = p->d_tau;
    cd* lam_k = reinterpret_cast<cd*>(buf + o_lam);
    int* info_k = reinterpret_cast<int*>(buf + o_info);
    double* res_k = reinterpret_cast<double*>(buf + o_res);
    eig_perturb_batched(s, n, tAo, tA, m);
    zgeev_batched(s, n, tA, tAo, tW, lam_k, ws, info_k, res_k, ws.X, m, nullptr,
                  p->ctx->eig_tol, p->qr_variant);
    eig_recover_scatter(s, n, m, reinterpret_cast<const int*>(buf + o_idx), lam_k, info_k,
                        res_k, p->d_lam, p->d_pinfo + ct.prob_base, p->d_resid + ct.prob_base);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(buf);
  HIPCHK(e);
  return ORCWA_OK;
}

__global__ void k_scale_rows_keff(int n, cd* const* Hp, const cd* const* Kp) {
  cd* H = Hp[blockIdx.y];
  const cd* kf = Kp[blockIdx.y];
  const long long count = (long long)n * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count;
       e += (long long)gridDim.x * blockDim.x)
    H[e] = H[e] * kf[e % n];
}

// The interface below a non-magnetic patterned layer a is formed as
// H = K_a^-1 W_a^-1 P_a V_b, which equals V_a^-1 V_b when W_a, K_a are the
// eigen solve of P_a Q_a.  A recovered solve is that of A + diag(1e-11 (i+1)
// max|A|) instead, so for the instances whose layer a was recovered H is
// formed as the reference forms it, V_a^-1 V_b with an LU of V_a
// (rcwa/RCWASolver.cpp:537-543), and stored as K_a H to keep the scaling the
// cascade applies.
static int interface_fixup(orcwa_plan* p, const ChunkTables& ct, int si,
                           const std::vector<int>& recovered, hipStream_t s) {
  const int L = p->L, n = p->n, a = L - 2 - si, b = L - 1 - si;
  std::vector<int> sel;
  for (int q : recovered) {
    if (q >= ct.nprob - ct.nprob_mag) continue;   // magnetic layers use V_a^-1 V_b already
    if (ct.prob_layer[q] != a) continue;
    const int i = p->prob_inst[(size_t)ct.prob_base + q];
    sel.push_back(i - ct.i0);
  }
  if (sel.empty()) return ORCWA_OK;
  const int m = (int)sel.size();
  std::vector<cd*> tabs((size_t)5 * m);
  std::vector<int> inst(m);
  for (int k = 0; k < m; ++k) {
    const int ii = sel[k];
    tabs[k] = ct.hV[(size_t)ii * L + a];
    tabs[m + k] = ct.hV[(size_t)ii * L + b];
    tabs[2 * m + k] = ct.hLU[(size_t)ii * (L - 1) + si];
    tabs[3 * m + k] = ct.hH[ii];
    tabs[4 * m + k] = ct.hK[(size_t)ii * L + a];
    inst[k] = ct.i0 + ii;
  }
  const size_t o_inst = align256(tabs.size() * sizeof(cd*));
  const size_t bytes = o_inst + m * sizeof(int);
  char* buf = nullptr;
  if (hipMalloc(&buf, bytes) != hipSuccess) return ORCWA_ERR_NOMEM;
  std::vector<char> img(bytes, 0);
  std::memcpy(img.data(), tabs.data(), tabs.size() * sizeof(cd*));
  std::memcpy(img.data() + o_inst, inst.data(), m * sizeof(int));
  hipError_t e = hipMemcpyAsync(buf, img.data(), bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    cd* const* t = reinterpret_cast<cd* const*>(buf);
    const cd* const* tc = reinterpret_cast<const cd* const*>(buf);
    const long long nn = (long long)n * n;
    copy_batched(s, nn, tc, t + 2 * m, m);       // V_a -> LU slot
    copy_batched(s, nn, tc + m, t + 3 * m, m);   // V_b -> H slot
    zgetrf_batched(s, n, t + 2 * m, p->d_ipiv, p->d_info, reinterpret_cast<const int*>(buf + o_inst),
                   m);
    zgetrs_batched(s, n, n, tc + 2 * m, p->d_ipiv, t + 3 * m, m, p->d_perm);
    hipLaunchKernelGGL(k_scale_rows_keff, dim3(64, m), dim3(256), 0, s, n, t + 3 * m, tc + 4 * m);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(buf);
  HIPCHK(e);
  return ORCWA_OK;
}

__global__ void k_zero_ints(int* a, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    a[i] = 0;
}

static void zero_ints(hipStream_t s, int* a, long long n) {
  const int blocks = (int)std::min<long long>(1024, (n + 255) / 256);
  hipLaunchKernelGGL(k_zero_ints, dim3(std::max(1, blocks)), dim3(256), 0, s, a, n);
}

// The whole batch on stream s.  graph: the form that a caller captures into a
// HIP graph -- device work only (status words cleared by a kernel, no events,
// no host synchronisation, so no in-line eigen recovery: orcwa_graph_sync
// re-runs a plan eagerly when one of its eigen solves failed).
static int plan_run(orcwa_plan* p, hipStream_t s, bool graph) {
  // the plan's GEMM tile for the launches of this thread, restored on return
  struct TileGuard {
    int saved;
    explicit TileGuard(int t) : saved(gemm_tile()) { set_gemm_tile(t); }
    ~TileGuard() { set_gemm_tile(saved); }
  } tile_guard(p->gemm_tile);
  const int n = p->n, h = p->h, G = p->G;
  const ConvGeom g = geom_of(p);
  const cd one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  const long long nn = (long long)n * n;
  auto record = [&](int k) { return graph ? hipSuccess : hipEventRecord(p->ev[k], s); };

  zero_ints(s, p->d_info, p->count);
  zero_ints(s, p->d_pinfo, std::max<long long>(1, p->nprob_total));
  zero_ints(s, p->d_ginfo, std::max(1, G));
  HIPCHK(record(0));

  // ---- 1. convolution matrices and per-grid products
  if (G > 0) {
    launch_conv_strip(s, g, p->d_tasks, p->ntasks, p->d_grids, p->d_strips, p->strip_lds);
    launch_conv_expand(s, g, p->d_strips, p->d_strip_base, p->dev<cd* const*>(p->t_gE11),
                       p->dev<cd* const*>(p->t_gE22), p->dev<cd* const*>(p->t_gE33), G);
    copy_batched(s, (long long)h * h, p->dev<const cd* const*>(p->t_gE33),
                 p->dev<cd* const*>(p->t_gLU), G);
    zgetrf_batched(s, h, p->dev<cd* const*>(p->t_gLU), p->d_gipiv, p->d_ginfo, nullptr, G);
    set_identity_batched(s, h, p->dev<cd* const*>(p->t_gM), G);
    zgetrs_batched(s, h, h, p->dev<const cd* const*>(p->t_gLU), p->d_gipiv,
                   p->dev<cd* const*>(p->t_gM), G, p->d_perm);
    const cd* const* Mt = p->dev<const cd* const*>(p->t_gM);
    zgemm_batched(s, h, h, h, one, Mt, 0, 1, h, 0, p->dev<const cd* const*>(p->t_gE11), 0, 1, h,
                  0, zero, p->dev<cd* const*>(p->t_gS1), 0, 1, h, G);
    launch_scale_rows_order(s, g, 0, p->dev<const cd* const*>(p->t_gE11),
                            p->dev<cd* const*>(p->t_gTmp), G);
    zgemm_batched(s, h, h, h, one, Mt, 0, 1, h, 0, p->dev<const cd* const*>(p->t_gTmp), 0, 1, h,
                  0, zero, p->dev<cd* const*>(p->t_gS2), 0, 1, h, G);
    zgemm_batched(s, h, h, h, one, Mt, 0, 1, h, 0, p->dev<const cd* const*>(p->t_gE22), 0, 1, h,
                  0, zero, p->dev<cd* const*>(p->t_gS3), 0, 1, h, G);
    launch_scale_rows_order(s, g, 1, p->dev<const cd* const*>(p->t_gE22),
                            p->dev<cd* const*>(p->t_gTmp), G);
    zgemm_batched(s, h, h, h, one, Mt, 0, 1, h, 0, p->dev<const cd* const*>(p->t_gTmp), 0, 1, h,
                  0, zero, p->dev<cd* const*>(p->t_gS4), 0, 1, h, G);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(record(1));
  double ms_modes = 0, ms_cascade = 0, ms_rt = 0, ms_eig = 0;

  std::vector<int> recovered;   // problems of the chunk whose eigen solve was recovered
  for (int c = 0; c < p->nchunks; ++c) {
    const ChunkTables& ct = p->chunks[c];
    const int np = ct.nprob, nu = ct.nuni, ni = ct.ninst;
    recovered.clear();
    HIPCHK(record(2));
    // ---- 2a-c. layer modes
    if (np > 0) {
      cd* const* A = p->dev<cd* const*>(ct.pA);
      cd* const* Wp = p->dev<cd* const*>(ct.pW);
      const cd* const* Wc = p->dev<const cd* const*>(ct.pW);
      const int nm = ct.nprob_mag, nq = np - nm;   // magnetic problems at [nq, np)
      launch_build_pq(s, g, p->dev<const ProbParam*>(ct.prob_params),
                      p->dev<const GridMats*>(p->t_gridmats), A, p->dev<cd* const*>(ct.pAorig),
                      nq);
      if (nm > 0) {
        // P in the V slot and Q in the QW slot, PQ = P Q (rcwa/WaveEquationSolver.cpp:30)
        cd* const* Pm = p->dev<cd* const*>(ct.pV) + nq;
        cd* const* Qm = p->dev<cd* const*>(ct.pQW) + nq;
        launch_build_pq_mag(s, g, p->dev<const ProbParam*>(ct.prob_params) + nq,
                            p->dev<const MagMats*>(ct.mag_mats), Pm, Qm, nm);
        zgemm_batched(s, n, n, n, one, (const cd* const*)Pm, 0, 1, n, 0, (const cd* const*)Qm, 0,
                      1, n, 0, zero, A + nq, 0, 1, n, nm);
        if (p->keep_pq)
          for (int l = 0; l < p->L; ++l)
            HIPCHK(hipMemcpyAsync(p->kP[l], p->kP_src[l], nn * sizeof(cd),
                                  hipMemcpyDeviceToDevice, s));
        copy_batched(s, nn, (const cd* const*)(A + nq), p->dev<cd* const*>(ct.pAorig) + nq, nm);
      }
      EigWorkspace ws;
      ws.Z = p->dev<cd**>(ct.pZ);
      ws.X = p->dev<cd**>(ct.pX);
      ws.tau = p->d_tau;
      HIPCHK(record(5));
      zgeev_batched(s, n, A, p->dev<const cd* const*>(ct.pAorig), Wp, p->d_lam, ws,
                    p->d_pinfo + ct.prob_base, p->d_resid + ct.prob_base,
                    p->dev<cd* const*>(ct.pX), np, nullptr, p->ctx->eig_tol, p->qr_variant);
      if (!graph) {
        const int rc = eig_recover(p, ct, s, recovered);
        if (rc != ORCWA_OK) return rc;
      }
      HIPCHK(record(6));
      // QW1 = [E22 W_bottom ; E11 W_top]
      cd* const* QW = p->dev<cd* const*>(ct.pQW);
      if (nq > 0) {
        zgemm_batched(s, h, n, h, one, p->dev<const cd* const*>(ct.pE22), 0, 1, h, 0, Wc, h, 1, n,
                      0, zero, QW, 0, 1, n, nq);
        zgemm_batched(s, h, n, h, one, p->dev<const cd* const*>(ct.pE11), 0, 1, h, 0, Wc, 0, 1, n,
                      0, zero, QW, h, 1, n, nq);
        launch_build_v(s, g, p->dev<const ProbParam*>(ct.prob_params), Wc,
                       p->dev<const cd* const*>(ct.pQW), p->d_lam, p->d_keffp,
                       p->dev<cd* const*>(ct.pV), p->d_info, p->dev<const int*>(ct.inst_of_prob),
                       nq);
      }
      if (nm > 0) {
        // Q W of the magnetic problems into the Z slot, then V = Q W diag(1/keff)
        cd* const* Zm = p->dev<cd* const*>(ct.pZ) + nq;
        zgemm_batched(s, n, n, n, one, p->dev<const cd* const*>(ct.pQW) + nq, 0, 1, n, 0, Wc + nq,
                      0, 1, n, 0, zero, Zm, 0, 1, n, nm);
        launch_build_v_mag(s, n, (const cd* const*)Zm, p->d_lam + (size_t)nq * n,
                           p->d_keffp + (size_t)nq * n, p->dev<cd* const*>(ct.pV) + nq, p->d_info,
                           p->dev<const int*>(ct.inst_of_prob) + nq, nm);
      }
    }
    if (nu > 0)
      launch_uniform(s, g, p->dev<const UniParam*>(ct.uni_params), p->d_keffu,
                     p->dev<cd* const*>(ct.uV), p->d_binv, p->d_info, nu);
    HIPCHK(hipGetLastError());
    HIPCHK(record(3));
    if (p->modes_only) {
      HIPCHK(record(4));
      HIPCHK(record(7));
      continue;
    }

    // ---- 2d-e. interfaces and the RTCM cascade, exit side first
    cd* const* E = p->dev<cd* const*>(ct.E);
    cd* const* H = p->dev<cd* const*>(ct.H);
    auto C = [](cd* const* x) { return (const cd* const*)x; };
    const int i0 = g.ny * (g.nx - 1) / 2 + (g.ny - 1) / 2;
    for (int si = 0; si < p->L - 1; ++si) {
      const StepTables& st = ct.steps[si];
      const bool last = (si == p->L - 2);
      cd* const* M = p->dev<cd* const*>(ct.LU[si]);
      int* piv = p->d_spiv + (size_t)si * p->chunk * n;
      // interfaces E = W_a^-1 W_b, H = V_a^-1 V_b (rcwa/RCWASolver.cpp:495-569)
      if (st.ua.count > 0) {
        const StepGroup& q = st.ua;
        copy_batched(s, nn, p->dev<const cd* const*>(q.Wb), p->dev<cd* const*>(q.E), q.count);
        launch_bdinv_apply(s, h, p->dev<const cd* const*>(q.binv_a),
                           p->dev<const cd* const*>(q.Vb), p->dev<cd* const*>(q.H), q.count);
      }
      if (st.pa.count > 0) {
        // one LU of W_a for both: [E | H'] = W_a^-1 [W_b | P_a V_b], H = K_a^-1 H'
        // (E and H are adjacent slots, so they form one n x 2n right-hand side)
        const StepGroup& q = st.pa;
        cd* const* qLU = p->dev<cd* const*>(q.LU);
        const ProbParam* qpp = p->dev<const ProbParam*>(q.pp);
        const cd* const* qVb = p->dev<const cd* const*>(q.Vb);
        launch_pv_in(s, g, qpp, qVb, qLU, q.count);
        zgemm_batched(s, h, n, h, one, p->dev<const cd* const*>(q.M), 0, 1, h, 0, C(qLU), 0, 1, h,
                      0, zero, qLU, (long long)h * n, 1, h, q.count);
        launch_pv_out(s, g, qpp, C(qLU), (long long)h * n, qVb, p->dev<cd* const*>(q.H), q.count);
        copy_batched(s, nn, p->dev<const cd* const*>(q.Wa), qLU, q.count);
        copy_batched(s, nn, p->dev<const cd* const*>(q.Wb), p->dev<cd* const*>(q.E), q.count);
        zgetrf_batched(s, n, qLU, p->d_ipiv, p->d_info, p->dev<const int*>(q.inst), q.count);
        zgetrs_batched(s, n, 2 * n, C(qLU), p->d_ipiv, p->dev<cd* const*>(q.E), q.count,
                       p->d_perm);
      }
      if (st.pm.count > 0) {
        // a magnetic patterned upper layer: E = W_a^-1 W_b and H = V_a^-1 V_b,
        // one LU each (rcwa/RCWASolver.cpp:537-543)
        const StepGroup& q = st.pm;
        cd* const* qLU = p->dev<cd* const*>(q.LU);
        copy_batched(s, nn, p->dev<const cd* const*>(q.Wa), qLU, q.count);
        copy_batched(s, nn, p->dev<const cd* const*>(q.Wb), p->dev<cd* const*>(q.E), q.count);
        zgetrf_batched(s, n, qLU, p->d_ipiv, p->d_info, p->dev<const int*>(q.inst), q.count);
        zgetrs_batched(s, n, n, C(qLU), p->d_ipiv, p->dev<cd* const*>(q.E), q.count, p->d_perm);
        copy_batched(s, nn, p->dev<const cd* const*>(q.Va), qLU, q.count);
        copy_batched(s, nn, p->dev<const cd* const*>(q.Vb), p->dev<cd* const*>(q.H), q.count);
        zgetrf_batched(s, n, qLU, p->d_ipiv, p->d_info, p->dev<const int*>(q.inst), q.count);
        zgetrs_batched(s, n, n, C(qLU), p->d_ipiv, p->dev<cd* const*>(q.H), q.count, p->d_perm);
      }
      if (!recovered.empty()) {
        const int rc = interface_fixup(p, ct, si, recovered, s);
        if (rc != ORCWA_OK) return rc;
      }
      // A = E + H (in E), B = E - H (in H), and A^T into M, B^T into the next
      // R~^T slot (the two products below add to them)
      cd* const* Rcur = p->dev<cd* const*>(ct.Rt[si & 1]);
      cd* const* Rnext = p->dev<cd* const*>(ct.Rt[(si + 1) & 1]);
      launch_sum_diff_t(s, n, E, H, p->dev<const cd* const*>(st.keff_a), M,
                        last ? nullptr : Rnext, ni);
      if (p->keep) {
        HIPCHK(hipMemcpyAsync(p->kA[si], p->kE_src, nn * sizeof(cd), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(p->kB[si], p->kH_src, nn * sizeof(cd), hipMemcpyDeviceToDevice, s));
      }
      // R~ = phi R phi, kept transposed (the scaling is symmetric in i, j)
      if (si > 0)
        launch_phase_scale(s, n, p->dev<const cd* const*>(st.keff_b),
                           p->dev<const double*>(st.k0d), Rcur, nullptr, p->d_info + ct.i0, ni);
      // M^T = (F + G)^T = A^T + (B R~)^T; the GEMM reads R~ through R~^T and
      // stores its product transposed
      if (si > 0)
        zgemm_batched_prec(p->gemm_prec, s, n, n, n, one, C(H), 0, 1, n, 0, C(Rcur), 0, n, 1, 0,
                           one, M, 0, n, 1, ni);
      // (F - G)^T = B^T + (A R~)^T
      if (!last && si > 0)
        zgemm_batched_prec(p->gemm_prec, s, n, n, n, one, C(E), 0, 1, n, 0, C(Rcur), 0, n, 1, 0,
                           one, Rnext, 0, n, 1, ni);
      zgetrf_batched(s, n, M, piv, p->d_info + ct.i0, nullptr, ni);
      // R_new^T = (F + G)^-T (F - G)^T
      if (!last) zgetrs_batched(s, n, n, C(M), piv, Rnext, ni, p->d_perm);
    }
    {
      // columns i0, h+i0 of Rud and Tdd
      const int sl = p->L - 2;
      // (a field plan keeps the chain's columns at the top of layer 1, 2, ..., L-1)
      const size_t tvb = 2 * (size_t)n * sizeof(cd);
      launch_cascade_tail(s, n, i0, p->dev<const cd* const*>(ct.LU[sl]),
                          p->d_spiv + (size_t)sl * p->chunk * n, C(E), C(H),
                          sl > 0 ? p->dev<const cd* const*>(ct.Rt[sl & 1]) : nullptr, p->d_rtv,
                          ni);
      if (p->keep)
        HIPCHK(hipMemcpyAsync(p->d_tv + 2 * (size_t)n, p->d_rtv + 2 * (size_t)n, tvb,
                              hipMemcpyDeviceToDevice, s));
      for (int sc = sl - 1; sc >= 0; --sc) {
        launch_cascade_chain(s, n, p->dev<const cd* const*>(ct.LU[sc]),
                             p->d_spiv + (size_t)sc * p->chunk * n,
                             p->dev<const cd* const*>(ct.steps[sc + 1].keff_b),
                             p->dev<const double*>(ct.steps[sc + 1].k0d), p->d_rtv, ni);
        if (p->keep)
          HIPCHK(hipMemcpyAsync(p->d_tv + 2 * (size_t)n * (p->L - 1 - sc),
                                p->d_rtv + 2 * (size_t)n, tvb, hipMemcpyDeviceToDevice, s));
      }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(record(4));
    // ---- 2f. plane-wave response
    launch_rt(s, g, p->dev<const RtParam*>(ct.rt_params), p->d_rtv,
              p->d_ref + (size_t)ct.i0 * p->maxpol * h, p->d_trn + (size_t)ct.i0 * p->maxpol * h,
              p->d_amp + (size_t)ct.i0 * 8 * h, p->d_info + ct.i0, p->maxpol, ni);
    HIPCHK(hipGetLastError());
    HIPCHK(record(7));
    if (!graph && (p->nchunks > 1 || c == p->nchunks - 1)) {
      // per-phase timing of this chunk (waits for the chunk; chunks run in order)
      HIPCHK(hipEventSynchronize(p->ev[7]));
      float t = 0;
      if (hipEventElapsedTime(&t, p->ev[2], p->ev[3]) == hipSuccess) ms_modes += t;
      if (np > 0 && hipEventElapsedTime(&t, p->ev[5], p->ev[6]) == hipSuccess) ms_eig += t;
      if (hipEventElapsedTime(&t, p->ev[3], p->ev[4]) == hipSuccess) ms_cascade += t;
      if (hipEventElapsedTime(&t, p->ev[4], p->ev[7]) == hipSuccess) ms_rt += t;
    }
  }
  if (graph) {
    for (double& v : p->phase_ms) v = -1.0;
    return ORCWA_OK;
  }
  float t = 0;
  p->phase_ms[0] = (hipEventElapsedTime(&t, p->ev[0], p->ev[7]) == hipSuccess) ? t : -1;
  p->phase_ms[1] = (hipEventElapsedTime(&t, p->ev[0], p->ev[1]) == hipSuccess) ? t : -1;
  p->phase_ms[2] = ms_modes;
  p->phase_ms[3] = ms_cascade;
  p->phase_ms[4] = ms_rt;
  p->phase_ms[5] = ms_eig;
  return ORCWA_OK;
}

int orcwa_plan_execute(orcwa_plan* p) {
  if (!p) return fail("null plan");
  HIPCHK(hipSetDevice(p->ctx->device));
  return plan_run(p, p->ctx->stream, false);
}

int orcwa_plan_sync(orcwa_plan* p) {
  if (!p) return fail("null plan");
  HIPCHK(hipSetDevice(p->ctx->device));
  HIPCHK(hipStreamSynchronize(p->ctx->stream));
  HIPCHK(hipGetLastError());
  return ORCWA_OK;
}

double orcwa_plan_phase_ms(orcwa_plan* p, int phase) {
  if (!p || phase < 0 || phase > 5) return -1.0;
  return p->phase_ms[phase];
}

int orcwa_plan_results(orcwa_plan* p, orcwa_result* out) {
  if (!p || !out) return fail("null argument");
  HIPCHK(hipSetDevice(p->ctx->device));
  hipStream_t s = p->ctx->stream;
  HIPCHK(hipStreamSynchronize(s));
  const int h = p->h, mp = p->maxpol;
  std::vector<double> ref((size_t)p->count * mp * h), trn((size_t)p->count * mp * h);
  std::vector<cd> amp((size_t)p->count * 8 * h);
  std::vector<int> info(p->count), pinfo(std::max<long long>(1, p->nprob_total));
  HIPCHK(hipMemcpy(ref.data(), p->d_ref, ref.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(trn.data(), p->d_trn, trn.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(amp.data(), p->d_amp, amp.size() * sizeof(cd), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(info.data(), p->d_info, info.size() * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(pinfo.data(), p->d_pinfo, pinfo.size() * sizeof(int), hipMemcpyDeviceToHost));
  int ginfo_any = 0;
  if (p->G > 0) {
    std::vector<int> gi(p->G);
    HIPCHK(hipMemcpy(gi.data(), p->d_ginfo, gi.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int v : gi) ginfo_any |= v;
  }
  for (long long q = 0; q < p->nprob_total; ++q) info[p->prob_inst[q]] |= pinfo[q];
  for (int i = 0; i < p->count; ++i) {
    orcwa_result& r = out[i];
    r.status = info[i] | ginfo_any;
    const int np = p->npol[i];
    if (r.ref)
      std::memcpy(r.ref, ref.data() + (size_t)i * mp * h, (size_t)np * h * sizeof(double));
    if (r.trn)
      std::memcpy(r.trn, trn.data() + (size_t)i * mp * h, (size_t)np * h * sizeof(double));
    if (r.amp) std::memcpy(r.amp, amp.data() + (size_t)i * 8 * h, (size_t)8 * h * sizeof(cd));
  }
  return ORCWA_OK;
}

int orcwa_gpu_solve_batch(orcwa_gpu_context* ctx, const orcwa_stack* stack,
                          const orcwa_instance* inst, int count, orcwa_result* out) {
  if (!out) return fail("null result array");
  orcwa_plan* p = nullptr;
  int rc = orcwa_plan_create(ctx, stack, inst, count, &p);
  if (rc != ORCWA_OK) return rc;
  rc = orcwa_plan_execute(p);
  if (rc == ORCWA_OK) rc = orcwa_plan_sync(p);
  if (rc == ORCWA_OK) rc = orcwa_plan_results(p, out);
  orcwa_plan_destroy(p);
  return rc;
}

// Field image of one instance: the inputCoeffs form of RCWASolver::saveFieldImage
// that the .orcwa driver calls for its first wavelength (rcwa/RCWASolver.cpp:707-1011,
// rcwa/RCWADriver.cpp:342-352).  The instance is solved by a field plan, then
// the mode amplitudes of every interior layer follow from the kept cascade
// data and the image is synthesised on the device.
int orcwa_gpu_field_image(orcwa_gpu_context* ctx, const orcwa_stack* st,
                          const orcwa_instance* inst, const orcwa_field_request* req, int* rows,
                          int* cols, double* field, int* status) {
  if (!ctx || !st || !inst || !req || !rows || !cols) return fail("field image: null argument");
  const int L = st->n_layers;
  const double* thick = inst->thickness ? inst->thickness : st->thickness;
  if (L < 3 || !thick)
    return fail("field image: the stack needs at least one interior layer (3 layers)");
  if (req->component < 0 || req->component > 5) return fail("field image: unknown component");
  if (req->slice < 0 || req->slice > 2) return fail("field image: unknown slice");
  if (req->nres1 < 0) return fail("field image: negative resolution");
  if (req->pol < 0 || req->pol >= inst->n_pol) return fail("field image: no such polarisation");
  if (st->n_cx < 2 || st->n_cy < 2 || !st->coord_x || !st->coord_y)
    return fail("each axis needs at least two cell edges");
  const int nx = 2 * st->nx_harm + 1, ny = 2 * st->ny_harm + 1, h = nx * ny, n = 2 * h;
  if (n > 2048) return fail("field image: at most 1024 harmonics");
  const int nres1 = req->nres1 > 0 ? req->nres1 : 500;       // rcwa/RCWASolver.cpp:763
  const double Lx = st->coord_x[st->n_cx - 1] - st->coord_x[0];
  const double Ly = st->coord_y[st->n_cy - 1] - st->coord_y[0];
  double zint = 0.0;
  for (int l = 1; l < L - 1; ++l) zint += thick[l];
  // image axes (rcwa/RCWASolver.cpp:730-779)
  double s1, l1, s2, l2;
  if (req->slice == 1) {
    s1 = st->coord_x[0], l1 = Lx, s2 = 0.0, l2 = zint;
  } else if (req->slice == 2) {
    s1 = st->coord_y[0], l1 = Ly, s2 = 0.0, l2 = zint;
  } else {
    s1 = st->coord_x[0], l1 = Lx, s2 = st->coord_y[0], l2 = Ly;
  }
  const int nres2 = std::max(1, (int)(nres1 * l2 / l1));
  const double step1 = l1 / nres1, step2 = l2 / nres2;
  // layer of each harmonics row: closed interval [z0, z1] from the top of
  // layer 1 (rcwa/RCWASolver.cpp:788-802)
  const double k0 = 2.0 * kPi / inst->lambda;
  auto locate = [&](double z, FieldRow& fr) {
    double z0 = 0.0;
    for (int j = 1; j < L - 1; ++j) {
      const double z1 = z0 + thick[j];
      if (z0 <= z && z <= z1) {
        fr.layer = j;
        fr.pad = 0;
        fr.k0zp = k0 * (z - z0);
        fr.k0zm = k0 * (z - z1);
        return true;
      }
      z0 = z1;
    }
    return false;
  };
  std::vector<FieldRow> frows;
  if (req->slice == 0) {
    FieldRow fr;
    if (!locate(req->coord, fr)) return fail("field image: z of the XY slice is outside the interior layers");
    frows.push_back(fr);
  } else {
    for (int r = 0; r < nres2; ++r) {
      FieldRow fr;
      if (!locate(s2 + step2 * (r + 0.5), fr)) return fail("field image: row outside the interior layers");
      frows.push_back(fr);
    }
  }
  *rows = nres2;
  *cols = nres1;
  if (!field) return ORCWA_OK;

  orcwa_plan* p = nullptr;
  int rc = plan_create(ctx, st, inst, 1, true, &p);
  if (rc != ORCWA_OK) return rc;
  struct Guard {
    orcwa_plan* p;
    void* buf = nullptr;
    ~Guard() {
      if (buf) hipFree(buf);
      orcwa_plan_destroy(p);
    }
  } guard{p};
  rc = orcwa_plan_execute(p);
  if (rc != ORCWA_OK) return rc;
  hipStream_t s = ctx->stream;
  const int nrh = (int)frows.size();
  std::vector<FieldLayer> fl(L);
  const bool etype = (req->component == 0 || req->component == 1 || req->component == 5);
  for (int l = 0; l < L; ++l) {
    fl[l].M = etype ? p->lW[l] : p->lV[l];
    fl[l].keff = p->lK[l];
    fl[l].m33inv = p->gid[l] >= 0 ? p->gM[p->gid[l]] : nullptr;
    fl[l].eps = p->ueps[l];
  }
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = align256(off);
    off = o + bytes;
    return o;
  };
  const size_t o_cp = take((size_t)L * n * sizeof(cd));
  const size_t o_cm = take((size_t)L * n * sizeof(cd));
  const size_t o_rows = take(frows.size() * sizeof(FieldRow));
  const size_t o_layers = take(fl.size() * sizeof(FieldLayer));
  const size_t o_X = take((size_t)nrh * n * sizeof(cd));
  const size_t o_Hm = take((size_t)nrh * h * sizeof(cd));
  const size_t o_img = take((size_t)nres2 * nres1 * sizeof(cd));
  ctx_trim(ctx);   // buffers cached by earlier plans go back to the device
  if (hipMalloc(&guard.buf, align256(off)) != hipSuccess) return ORCWA_ERR_NOMEM;
  char* B = (char*)guard.buf;
  cd* cp = (cd*)(B + o_cp);
  cd* cm = (cd*)(B + o_cm);
  FieldRow* d_rows = (FieldRow*)(B + o_rows);
  FieldLayer* d_layers = (FieldLayer*)(B + o_layers);
  cd* X = (cd*)(B + o_X);
  cd* Hm = (cd*)(B + o_Hm);
  cd* img = (cd*)(B + o_img);
  HIPCHK(hipMemcpyAsync(d_rows, frows.data(), frows.size() * sizeof(FieldRow),
                        hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_layers, fl.data(), fl.size() * sizeof(FieldLayer),
                        hipMemcpyHostToDevice, s));
  // downward amplitudes for the requested polarisation, then the upward ones
  // from the exit side (no upward wave in the exit half-space)
  const cd px = p->pol[((size_t)0 * kMaxPol + req->pol) * 2 + 0];
  const cd py = p->pol[((size_t)0 * kMaxPol + req->pol) * 2 + 1];
  launch_field_combine(s, n, L, p->d_tv, px, py, cp);
  HIPCHK(hipMemsetAsync(cm + (size_t)(L - 1) * n, 0, (size_t)n * sizeof(cd), s));
  for (int j = L - 2; j >= 1; --j) {
    const int si = L - 2 - j;    // the cascade step of the interface j | j+1
    const double k0d = (j + 1 == L - 1) ? 0.0 : k0 * thick[j + 1];
    launch_field_up(s, n, p->kA[si], p->kB[si], cp + (size_t)(j + 1) * n, cm + (size_t)(j + 1) * n,
                    p->lK[j + 1], k0d, cm + (size_t)j * n);
  }
  launch_field_modes(s, n, d_layers, d_rows, nrh, cp, cm, etype ? 1 : -1, X);
  FieldGeom g;
  g.nx = nx;
  g.ny = ny;
  g.component = req->component;
  g.slice = req->slice;
  g.Lx = Lx;
  g.Ly = Ly;
  g.kx0 = inst->kx0;
  g.ky0 = inst->ky0;
  g.k0 = k0;
  g.start1 = s1;
  g.step1 = step1;
  g.start2 = s2;
  g.step2 = step2;
  g.coord = req->coord;
  launch_field_component(s, g, d_layers, d_rows, nrh, X, Hm);
  launch_field_synth(s, g, Hm, nres2, nres1, img);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(field, img, (size_t)nres2 * nres1 * sizeof(cd), hipMemcpyDeviceToHost));
  if (status) {
    orcwa_result r;
    std::memset(&r, 0, sizeof(r));
    rc = orcwa_plan_results(p, &r);
    if (rc != ORCWA_OK) return rc;
    *status = r.status;
  }
  return ORCWA_OK;
}

}  // extern "C"

namespace orcwa {

int layer_modes(orcwa_gpu_context* ctx, const orcwa_stack* st, const orcwa_instance* inst,
                const double* layer_cx, const double* layer_cy, LayerModes& out, bool want_pq) {
  orcwa_plan* p = nullptr;
  int rc = plan_create(ctx, st, inst, 1, true, &p, true, layer_cx, layer_cy, want_pq);
  if (rc != ORCWA_OK) return rc;
  rc = orcwa_plan_execute(p);
  if (rc == ORCWA_OK) rc = orcwa_plan_sync(p);
  orcwa_result r;
  std::memset(&r, 0, sizeof(r));
  if (rc == ORCWA_OK) rc = orcwa_plan_results(p, &r);
  if (rc != ORCWA_OK) {
    orcwa_plan_destroy(p);
    return rc;
  }
  out.plan = p;
  out.n = p->n;
  out.status = r.status;
  out.W.assign(p->lW.begin(), p->lW.end());
  out.V.assign(p->lV.begin(), p->lV.end());
  out.keff.assign(p->lK.begin(), p->lK.end());
  out.P.assign(p->lP.begin(), p->lP.end());
  out.Q.assign(p->lQ.begin(), p->lQ.end());
  return ORCWA_OK;
}

void layer_modes_release(LayerModes& m) {
  orcwa_plan_destroy(m.plan);
  m.plan = nullptr;
}

}  // namespace orcwa

// ---------------------------------------------------------------------------
// Graphs: the executes of several plans captured once into one HIP graph.
// ---------------------------------------------------------------------------
struct orcwa_graph {
  orcwa_gpu_context* ctx = nullptr;
  std::vector<orcwa_plan*> plans;
  std::vector<hipStream_t> streams;     // the forks of the capture
  std::vector<hipEvent_t> events;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int reruns = 0;
};

int orcwa_graph_destroy(orcwa_graph* gr) {
  if (!gr) return ORCWA_OK;
  hipSetDevice(gr->ctx->device);
  hipStreamSynchronize(gr->ctx->stream);
  if (gr->exec) hipGraphExecDestroy(gr->exec);
  if (gr->graph) hipGraphDestroy(gr->graph);
  for (hipStream_t st : gr->streams) hipStreamDestroy(st);
  for (hipEvent_t e : gr->events) hipEventDestroy(e);
  for (orcwa_plan* p : gr->plans) p->live_graphs--;
  gr->ctx->live_plans--;
  delete gr;
  return ORCWA_OK;
}

int orcwa_graph_create(orcwa_gpu_context* ctx, orcwa_plan* const* plans, int nplans,
                       int n_streams, orcwa_graph** out) {
  if (!ctx || !plans || nplans < 1 || !out || n_streams < 1)
    return fail("orcwa_graph_create: bad argument");
  *out = nullptr;
  for (int i = 0; i < nplans; ++i) {
    if (!plans[i] || plans[i]->ctx != ctx)
      return fail("orcwa_graph_create: every plan must belong to the context");
    if (plans[i]->keep || plans[i]->modes_only)
      return fail("orcwa_graph_create: field and mode plans cannot be captured");
  }
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  orcwa_graph* gr = new orcwa_graph;
  gr->ctx = ctx;
  gr->plans.assign(plans, plans + nplans);
  ctx->live_plans++;
  for (orcwa_plan* p : gr->plans) p->live_graphs++;
  const int ns = std::min(n_streams, nplans);
  auto bail = [&](const char* what) {
    fprintf(stderr, "[openrcwa] orcwa_graph_create: %s\n", what);
    (void)hipGetLastError();
    orcwa_graph_destroy(gr);
    return ORCWA_ERR_HIP;
  };
  // forks at the default priority (streams of other priorities must not be mixed)
  for (int k = 1; k < ns; ++k) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
      return bail("stream creation failed");
    gr->streams.push_back(st);
  }
  for (int k = 0; k < 2 * ns; ++k) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return bail("event creation failed");
    gr->events.push_back(e);
  }
  hipStream_t s0 = ctx->stream;
  if (hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return bail("capture could not begin");
  int rc = ORCWA_OK;
  // fork: every extra stream waits on the origin's start
  hipError_t e = hipEventRecord(gr->events[0], s0);
  for (int k = 1; k < ns && e == hipSuccess; ++k)
    e = hipStreamWaitEvent(gr->streams[k - 1], gr->events[0], 0);
  // plan i on stream i mod ns (the largest plans first spread the work)
  for (int i = 0; i < nplans && e == hipSuccess && rc == ORCWA_OK; ++i) {
    hipStream_t st = (i % ns) == 0 ? s0 : gr->streams[i % ns - 1];
    rc = plan_run(gr->plans[i], st, true);
  }
  // join
  for (int k = 1; k < ns && e == hipSuccess; ++k) {
    e = hipEventRecord(gr->events[ns + k], gr->streams[k - 1]);
    if (e == hipSuccess) e = hipStreamWaitEvent(s0, gr->events[ns + k], 0);
  }
  const hipError_t ec = hipStreamEndCapture(s0, &gr->graph);
  if (rc != ORCWA_OK || e != hipSuccess || ec != hipSuccess) return bail("capture failed");
  if (hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0) != hipSuccess)
    return bail("instantiation failed");
  *out = gr;
  return ORCWA_OK;
}

int orcwa_graph_launch(orcwa_graph* gr) {
  if (!gr) return fail("orcwa_graph_launch: null graph");
  HIPCHK(hipSetDevice(gr->ctx->device));
  HIPCHK(hipGraphLaunch(gr->exec, gr->ctx->stream));
  return ORCWA_OK;
}

int orcwa_graph_sync(orcwa_graph* gr) {
  if (!gr) return fail("orcwa_graph_sync: null graph");
  HIPCHK(hipSetDevice(gr->ctx->device));
  hipStream_t s = gr->ctx->stream;
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipGetLastError());
  // a plan with an eigen solve that failed its residual test (or its QR limit)
  // is run again eagerly, which takes the recovery branch
  gr->reruns = 0;
  for (orcwa_plan* p : gr->plans) {
    std::vector<int> pin((size_t)std::max<long long>(1, p->nprob_total));
    HIPCHK(hipMemcpyAsync(pin.data(), p->d_pinfo, pin.size() * sizeof(int),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    bool bad = false;
    for (int v : pin) bad = bad || (v & (ORCWA_STATUS_EIG_RESIDUAL | ORCWA_STATUS_NOCONV));
    if (bad) {
      const int rc = plan_run(p, s, false);
      if (rc != ORCWA_OK) return rc;
      HIPCHK(hipStreamSynchronize(s));
      gr->reruns++;
    }
  }
  return ORCWA_OK;
}

int orcwa_graph_reruns(const orcwa_graph* gr) { return gr ? gr->reruns : -1; }

// ---------------------------------------------------------------------------
// Autotuning of the kernel variants per layer-matrix size.
// ---------------------------------------------------------------------------
static std::string arch_name(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
  std::string a = prop.gcnArchName;
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}

int orcwa_gpu_set_tuning_file(orcwa_gpu_context* ctx, const char* path, int append) {
  if (!ctx || !path) return fail("orcwa_gpu_set_tuning_file: null argument");
  ctx->tuning_file = path;
  ctx->tuning_append = append != 0;
  FILE* f = fopen(path, "r");
  if (!f) return ORCWA_OK;     // nothing tuned yet
  const std::string arch = arch_name(ctx->device);
  char a[64];
  int n, qr, tile;
  double ms;
  char line[256];
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    if (sscanf(line, "%63s %d %d %d %lf", a, &n, &qr, &tile, &ms) != 5) continue;
    if (arch != a || n < 1 || qr < 0 || qr >= QR_NVARIANTS || tile < 0 || tile >= GEMM_NTILES)
      continue;
    orcwa_gpu_context::Tune t;
    t.qr = qr;
    t.tile = tile;
    t.ms = ms;
    ctx->tuning[n] = t;        // later lines win
  }
  fclose(f);
  return ORCWA_OK;
}

int orcwa_plan_autotune(orcwa_plan* p, int* qr_variant, int* gemm_tile) {
  if (!p) return fail("orcwa_plan_autotune: null plan");
  if (p->keep || p->modes_only) return fail("orcwa_plan_autotune: field and mode plans");
  if (p->live_graphs > 0) return fail("orcwa_plan_autotune: the plan is captured in a graph");
  HIPCHK(hipSetDevice(p->ctx->device));
  auto timed = [&](int qr, int tile, double& ms) -> int {
    p->qr_variant = qr;
    p->gemm_tile = tile;
    const int rc = plan_run(p, p->ctx->stream, false);
    if (rc != ORCWA_OK) return rc;
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    ms = p->phase_ms[0];
    return ORCWA_OK;
  };
  // the QR variants with the default tile, then the tiles with the best QR
  int best_qr = QR_W48, best_tile = GEMM_TILE_64x64;
  double best = 1e300, ms = 0.0;
  for (int q = 0; q < QR_NVARIANTS; ++q) {
    const int rc = timed(q, GEMM_TILE_64x64, ms);
    if (rc != ORCWA_OK) return rc;
    if (ms > 0.0 && ms < best) {
      best = ms;
      best_qr = q;
    }
  }
  for (int t = 1; t < GEMM_NTILES; ++t) {
    const int rc = timed(best_qr, t, ms);
    if (rc != ORCWA_OK) return rc;
    if (ms > 0.0 && ms < best) {
      best = ms;
      best_tile = t;
    }
  }
  p->qr_variant = best_qr;
  p->gemm_tile = best_tile;
  orcwa_gpu_context::Tune t;
  t.qr = best_qr;
  t.tile = best_tile;
  t.ms = best;
  p->ctx->tuning[p->n] = t;
  if (p->ctx->tuning_append && !p->ctx->tuning_file.empty()) {
    if (FILE* f = fopen(p->ctx->tuning_file.c_str(), "a")) {
      fprintf(f, "%s %d %d %d %.3f\n", arch_name(p->ctx->device).c_str(), p->n, best_qr,
              best_tile, best);
      fclose(f);
    }
  }
  if (qr_variant) *qr_variant = best_qr;
  if (gemm_tile) *gemm_tile = best_tile;
  // the last timed execute used another variant: leave the plan's results those of the choice
  const int rc = plan_run(p, p->ctx->stream, false);
  if (rc != ORCWA_OK) return rc;
  HIPCHK(hipStreamSynchronize(p->ctx->stream));
  return ORCWA_OK;
}
