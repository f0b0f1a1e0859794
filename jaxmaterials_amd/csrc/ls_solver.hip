// Host orchestration of the MI355X Lippmann-Schwinger solver and its C ABI.
//
// The reference builds cuFFT plans and ~54 floats/voxel of buffers per call,
// then runs a loop with five device syncs and two blocking copies per iteration
// (cuda/src/lippmann_schwinger.cu:112-234).  Here one solver object holds the
// FFT plans (per-axis twiddle and sin/cos tables), a z-half spectrum for all
// load cases, the strain as base + correction, the FP64-evaluated divergence
// of the base stress and small FP64 reduction buffers.  The loop keeps every
// per-iteration decision on the device (per-load-case active flags) and meets
// the host only every few iterations, where it also decides when to move the
// correction into the base (see ls_kernels.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lippmann_schwinger.h"
#include "ls_kernels.hpp"
#include "ls_passes_reg.hpp"
#include "phasefield.hpp"

using namespace lsk;

#define LS_HIP(expr)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      fprintf(stderr, "liblippmannschwinger: HIP error %s at %s:%d: %s\n", hipGetErrorName(e_), \
              __FILE__, __LINE__, #expr);                                                       \
      return -1;                                                                                \
    }                                                                                           \
  } while (0)

namespace {

// ---------------------------------------------------------------------------
// FFT plan for one axis
struct HostAxis {
  int n = 0;
  std::vector<int> radix;  // empty => direct DFT
  float2* d_tw = nullptr;
  double2* d_tw64 = nullptr;
  float* d_s = nullptr;
  float* d_c = nullptr;
  float2* d_ph = nullptr;  // e^{+i pi k / n}
};

std::vector<int> factorise(int n, bool generic) {
  std::vector<int> r;
  int m = n, e2 = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++e2;
  }
  std::vector<int> p3, p5;
  while (m % 3 == 0) {
    m /= 3;
    p3.push_back(3);
  }
  while (m % 5 == 0) {
    m /= 5;
    p5.push_back(5);
  }
  // m > 1: prime factors above 5, transformed by one generic stage
  // (fft_lds.hpp's generic_stage) after the others; without a factor 2, 3 or
  // 5 at all, or on an axis whose kernels leave the generic stage out, the
  // whole line is one direct DFT
  if (m != 1 && (m == n || !generic)) return {};
  // powers of two as 8s and 4s (16 = 4 x 4), a single 2 only for odd exponents left over
  int e = e2;
  while (e >= 3 && e != 4) {
    r.push_back(8);
    e -= 3;
  }
  while (e >= 2) {
    r.push_back(4);
    e -= 2;
  }
  if (e == 1) r.push_back(2);
  for (int v : p3) r.push_back(v);
  for (int v : p5) r.push_back(v);
  if (m != 1) r.push_back(m);
  return r;
}

// radices with register butterflies; any other radix runs generic_stage
bool butterfly_radix(int R) { return R == 2 || R == 3 || R == 4 || R == 5 || R == 8; }

bool fits(const HostAxis& a, int nlines, int E, int NT) {
  const long long tot = (long long)nlines * a.n;
  if (a.radix.empty()) return (tot + NT - 1) / NT <= E;
  for (int R : a.radix) {
    if (!butterfly_radix(R)) {
      if ((tot + NT - 1) / NT > E) return false;  // one output per work item
      continue;
    }
    const long long bf = tot / R;
    if (((bf + NT - 1) / NT) * R > E) return false;
  }
  return true;
}

AxisPlan to_dev(const HostAxis& a) {
  AxisPlan p;
  p.n = a.n;
  p.nstages = (int)a.radix.size();
  int ns = 1;
  for (int i = 0; i < kMaxStages; ++i) {
    p.radix[i] = i < p.nstages ? a.radix[i] : 0;
    p.f_ns[i] = make_fastdiv(ns);
    p.f_nr[i] = make_fastdiv(i < p.nstages ? a.n / a.radix[i] : 1);
    if (i < p.nstages) ns *= a.radix[i];
  }
  p.f_n = make_fastdiv(a.n);
  p.tw = a.d_tw;
  p.tw64 = a.d_tw64;
  return p;
}

constexpr size_t kMaxLds = 60 * 1024;
constexpr int kMaxTm = 256;  // largest x-pass tile (modes per block); bounds the d_base padding
constexpr int kCUs = 256;     // compute units of an MI355X (fallback when the device query fails)

struct Tile {
  int lines = 0;  // Lz (z), T (y) or Tm (x)
  int E = 0;      // register elements per thread (LDS engine)
  int NT = 0;     // threads per block
  size_t lds = 0;
  int reg = 0;    // 1: register two-stage y kernel (k_ypass_r); 2: three-line x kernel (k_xpass_h)
  int inpl = 0;   // k_xpass_h: transforms in place (half the LDS)
};

// Register two-stage kernels compiled for the axis lengths of the workloads
// (n = A B; lines = T columns for y, TM modes for x).  The first entry of a
// length is its default; the autotuner times all of them against the LDS engine.
#define LS_REG_Y_LIST(F)                                                                         \
  F(256, 16, 16, 16, 256) F(256, 16, 16, 8, 128) F(128, 16, 8, 32, 256) F(128, 16, 8, 48, 384)  \
  F(96, 12, 8, 32, 256) F(96, 12, 8, 64, 512) F(64, 8, 8, 32, 256) F(64, 8, 8, 64, 512)         \
  F(48, 6, 8, 32, 256) F(48, 6, 8, 64, 512) F(32, 8, 4, 64, 256) F(32, 8, 4, 128, 512)          \
  F(160, 5, 32, 32, 256) F(160, 32, 5, 32, 160) F(160, 32, 5, 64, 320) F(160, 5, 32, 64, 320)
// x pass with three-line transforms (k_xpass_h): (n, A, B, TM, NT, mode); mode 0: padded
// LDS, 1: transforms in place, 2: in place with the split stage 2
#define LS_REG_H_LIST(F)                                                                                     \
  F(256, 16, 16, 8, 512, 2) F(256, 32, 8, 8, 256, 1) F(256, 32, 8, 4, 256, 0) F(256, 16, 16, 4, 256, 0)      \
  F(256, 32, 8, 4, 128, 0) F(256, 32, 8, 4, 128, 1) F(256, 16, 16, 4, 256, 2) F(128, 16, 8, 8, 256, 0)      \
  F(128, 16, 8, 4, 256, 0) F(128, 16, 8, 8, 256, 1) F(96, 12, 8, 8, 256, 0) F(96, 12, 8, 4, 128, 0)          \
  F(96, 12, 8, 8, 256, 1) F(64, 8, 8, 8, 256, 0)                                                             \
  F(64, 8, 8, 16, 512, 0) F(64, 8, 8, 16, 512, 1) F(48, 6, 8, 16, 256, 0) F(48, 6, 8, 8, 256, 0)             \
  F(32, 8, 4, 16, 256, 0) F(32, 8, 4, 32, 512, 0) F(160, 32, 5, 8, 256, 1) F(160, 32, 5, 8, 256, 0)          \
  F(160, 32, 5, 16, 512, 1) F(160, 32, 5, 4, 128, 0)

struct RegCfg {
  int n, A, B, lines, NT, inpl;
};
#define LS_AS_CFG(NV, AV, BV, LV, NTV) {NV, AV, BV, LV, NTV, 0},
#define LS_AS_CFG6(NV, AV, BV, LV, NTV, IP) {NV, AV, BV, LV, NTV, IP},
const RegCfg kRegY[] = {LS_REG_Y_LIST(LS_AS_CFG)};
const RegCfg kRegH[] = {LS_REG_H_LIST(LS_AS_CFG6)};
#undef LS_AS_CFG
#undef LS_AS_CFG6

// Register-kernel tiles of an axis: kind 1 = y columns (LDS: T Np complex),
// kind 2 = the three-line x pass (LDS: 3 TM (n + Np) complex).
std::vector<Tile> reg_tiles(const RegCfg* list, size_t count, int n, int kind) {
  std::vector<Tile> out;
  for (size_t i = 0; i < count; ++i) {
    const RegCfg& c = list[i];
    if (c.n != n) continue;
    Tile t;
    t.lines = c.lines;
    t.NT = c.NT;
    t.reg = kind;
    t.E = c.A;  // distinguishes the (A, B) splits of one length
    t.inpl = c.inpl;
    const int np = c.n + c.n / c.B + 1;
    // the in-place x kernel stores its transforms unpadded (swizzled) where the
    // length allows it (SwzLayout)
    const bool swz = c.n % 32 == 0 && (c.B & (c.B - 1)) == 0 && c.B <= 32;
    t.lds = kind == 2 ? (size_t)3 * c.lines * (c.inpl ? (swz ? c.n : np) : c.n + np) * sizeof(float2)
                      : (size_t)c.lines * np * sizeof(float2);
    out.push_back(t);
  }
  return out;
}

struct Budget {
  int E, NT;
};
// preference order: fewest register elements per thread first
const std::vector<Budget> kBudget32 = {{8, 256}, {8, 512}, {16, 512}, {16, 256}, {32, 256}, {64, 256}};
const std::vector<Budget> kBudget64 = {{8, 256}, {8, 512}, {16, 256}, {32, 256}};

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v ? std::atoi(v) : dflt;
}

}  // namespace

struct ls_solver {
  int nx, ny, nz, P;
  long long N, M;
  float Lx, Ly, Lz;
  int nrhs, nmat, depth, m1;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  HostAxis ax[3];
  GridDev gd;
  Tile tz, ty, tx;        // FP32 iteration passes
  Tile tz64, ty64, tx64;  // FP64 base evaluation passes
  int nblk_z = 0, nblk_y = 0, ntiles_y = 0, nblk_x = 0;
  int nblk_z64 = 0, nblk_y64 = 0, ntiles_y64 = 0, nblk_x64 = 0;
  DbLayout dl;
  float2* spec = nullptr;
  double2* spec64 = nullptr;  // one load case
  float* base = nullptr;
  float2* dbase = nullptr;
  double* partials = nullptr;
  float* dc = nullptr;
  RhsDesc* d_rhs = nullptr;
  RhsState* d_st = nullptr;
  RhsState* h_st = nullptr;  // pinned
  int* d_sel = nullptr;
  double* part6 = nullptr;
  double* d_mean6 = nullptr;
  double* h_mean6 = nullptr;  // pinned
  float* d_mm = nullptr;
  float* h_mm = nullptr;  // pinned
  int nblk_stress = 0, chunk_stress = 0;
  int nblk_mm = 0, chunk_mm = 0;
  double rebase_factor = 1e-4;  // LS_REBASE_FACTOR; one rebase per 256^3 solve to 1e-6
  int ncu = kCUs;               // compute units of the device: small grids get about one block per CU
  std::vector<int> h_mat;       // material of each load case (current solve)
  int n_rebase = 0;
  int absolute = 0;  // LS_ABSOLUTE_RESIDUAL flag of the current solve
  // Anderson
  float* aaG = nullptr;
  float2* aaU = nullptr;
  float* aaPl = nullptr;    // Im R of the self-conjugate planes after the inverse y pass
  long long pl_slot = 0;    // floats per slot
  int nblk_p = 0;           // blocks per load case of the plane correction
  long long chunk_p = 0;
  double* aa_alpha = nullptr;
  double* aa_alpha_prev = nullptr;  // weights of the previous mix (rebuild delta)
  double* aa_dots = nullptr;
  double* aa_gram = nullptr;
  int* aa_head = nullptr;
  int* aa_filled = nullptr;
  // conjugate gradients (LS_CG; buffers made at the first such solve)
  int cg = 0;                  // the current solve uses conjugate gradients
  float* cgP = nullptr;        // search directions [R][6][N]
  CgState* d_cg = nullptr;     // [R]
  double* cg_zpart = nullptr;  // [R][nx ny]: z-pass partials of <p, C p>_w
  double* cg_xpart = nullptr;  // [R][ny P]: x-pass partials of the residual's C0 norm
  // captured chunk of iterations between two host synchronisations (LS_GRAPH)
  hipGraphExec_t gexec = nullptr;
  const void* gkey[3] = {nullptr, nullptr, nullptr};
  double gtol[2] = {0.0, 0.0};
  int gsteps = 0, gabs = -1, gcg = -1;
  // timing of the last solve (LS_TIMING set): HIP events on the solver stream
  static constexpr int kKinds = 7;  // z (plain), y, x, other, z Anderson, Anderson Gram row, z conjugate gradients
  double t_ms[kKinds] = {0, 0, 0, 0, 0, 0, 0};
  int t_cnt[kKinds] = {0, 0, 0, 0, 0, 0, 0};
  bool launch_failed = false;  // a pass found no kernel for its tile (set by ypass / xpass)
  int nblk_g = 0;           // blocks per load case of the Gram-row kernel
  long long chunk_g = 0;    // complex entries per block
  bool timing = false;
  std::vector<hipEvent_t> tev;
  std::vector<int> tkind;
  // built in a caller-owned device region (ls_solve_noalloc): no device
  // buffer is freed, host buffers are heap memory, no stream or event is made
  bool arena_mode = false;
  std::vector<std::unique_ptr<unsigned char[]>> host_mem;
};

namespace {

// Where the device buffers of a solver under construction come from: hipMalloc,
// or (while t_arena is set: the no-allocation entry point) consecutive
// 256-byte-aligned pieces of a caller-owned region.  An arena that runs out
// hands out null pointers, keeps counting the bytes the solver needs and
// flags the overflow, so the caller can report the size and give up.
struct Arena {
  unsigned char* base = nullptr;
  size_t cap = 0, off = 0;
  bool overflow = false;
};
thread_local Arena* t_arena = nullptr;

template <typename T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return 0;
  if (t_arena) {
    Arena& A = *t_arena;
    const size_t off = (A.off + 255) & ~(size_t)255;
    A.off = off + count * sizeof(T);
    if (A.off > A.cap)
      A.overflow = true;
    else
      *p = reinterpret_cast<T*>(A.base + off);
    return 0;
  }
  LS_HIP(hipMalloc((void**)p, count * sizeof(T)));
  return 0;
}


// Host buffers the stream copies into: pinned memory, or ordinary heap memory
// owned by the solver when it is built in an arena (no HIP allocation).
template <typename T>
int halloc(ls_solver* s, T** p, size_t count) {
  *p = nullptr;
  if (s->arena_mode) {
    s->host_mem.emplace_back(new unsigned char[std::max<size_t>(1, count * sizeof(T))]);
    *p = reinterpret_cast<T*>(s->host_mem.back().get());
    return 0;
  }
  LS_HIP(hipHostMalloc((void**)p, count * sizeof(T), hipHostMallocDefault));
  return 0;
}

int build_axis(HostAxis& a, int n, bool generic) {
  a.n = n;
  a.radix = factorise(n, generic);
  std::vector<float2> tw(n);
  std::vector<double2> tw64(n);
  for (int t = 0; t < n; ++t) {
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
    tw64[t] = make_double2((double)cosl(ang), (double)sinl(ang));
    tw[t] = make_float2((float)tw64[t].x, (float)tw64[t].y);
  }
  std::vector<float> s(n), c(n);
  for (int k = 0; k < n; ++k) {
    s[k] = (float)std::sin(M_PI * k / n);
    c[k] = (float)std::cos(M_PI * k / n);
  }
  s[0] = 0.f;
  if (n % 2 == 0) c[n / 2] = 0.f;
  std::vector<float2> ph(n);
  for (int k = 0; k < n; ++k) {
    const long double ang = 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
    ph[k] = make_float2((float)cosl(ang), (float)sinl(ang));
  }
  if (n % 2 == 0) ph[n / 2] = make_float2(0.f, 1.f);
  if (dalloc(&a.d_tw, n) || dalloc(&a.d_tw64, n) || dalloc(&a.d_s, n) || dalloc(&a.d_c, n) ||
      dalloc(&a.d_ph, n))
    return -1;
  if (t_arena && t_arena->overflow) return 0;  // sizes counted, nothing to copy into
  LS_HIP(hipMemcpy(a.d_ph, ph.data(), n * sizeof(float2), hipMemcpyHostToDevice));
  LS_HIP(hipMemcpy(a.d_tw, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice));
  LS_HIP(hipMemcpy(a.d_tw64, tw64.data(), n * sizeof(double2), hipMemcpyHostToDevice));
  LS_HIP(hipMemcpy(a.d_s, s.data(), n * sizeof(float), hipMemcpyHostToDevice));
  LS_HIP(hipMemcpy(a.d_c, c.data(), n * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

// Largest tile of at most `want` units whose FFT fits one of the register
// budgets with LDS <= kMaxLds; a unit is `lines_per_unit` FFT lines of
// `bytes_per_unit` LDS bytes.
bool choose_tile(const HostAxis& a, int want, int lines_per_unit, size_t bytes_per_unit,
                 const std::vector<Budget>& budgets, Tile& t) {
  for (int u = std::max(1, want); u >= 1; --u) {
    if ((size_t)u * bytes_per_unit > kMaxLds) continue;
    for (const Budget& b : budgets) {
      if (fits(a, u * lines_per_unit, b.E, b.NT)) {
        t.lines = u;
        t.E = b.E;
        t.NT = b.NT;
        t.lds = (size_t)u * bytes_per_unit;
        return true;
      }
    }
  }
  return false;
}

// Expand CALL(E, NT) for the (E, NT) pair of a tile.
#define LS_DISPATCH32(TILE, CALL)          \
  switch ((TILE).E * 1000 + (TILE).NT) {   \
    case 8256: CALL(8, 256); break;        \
    case 8512: CALL(8, 512); break;        \
    case 16512: CALL(16, 512); break;      \
    case 16256: CALL(16, 256); break;      \
    case 32256: CALL(32, 256); break;      \
    default: CALL(64, 256); break;         \
  }
#define LS_DISPATCH64(TILE, CALL)          \
  switch ((TILE).E * 1000 + (TILE).NT) {   \
    case 8256: CALL(8, 256); break;        \
    case 8512: CALL(8, 512); break;        \
    case 16256: CALL(16, 256); break;      \
    default: CALL(32, 256); break;         \
  }

void mark(ls_solver* s, int kind) {
  if (!s->timing) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s->stream);
  s->tev.push_back(e);
  s->tkind.push_back(kind);
}

// ---- kernel launch dispatch on E ------------------------------------------
AAParams aa_params(const ls_solver* s, int record) {
  AAParams aa;
  aa.G = s->aaG;
  aa.U = s->aaU;
  aa.u_per_rhs = s->dl.per_rhs;
  aa.alpha = s->aa_alpha;
  aa.alpha_prev = s->aa_alpha_prev;
  aa.head = s->aa_head;
  aa.filled = s->aa_filled;
  aa.m1 = s->m1;
  aa.record = (record && s->depth > 0) ? 1 : 0;
  aa.P = s->cgP;
  aa.cg = s->d_cg;
  aa.cg_part = s->cg_zpart;
  return aa;
}

template <int MODE>
void zline(ls_solver* s, const float* lam, const float* mu, float* eps) {
  const AAParams aa = aa_params(s, 0);
  const unsigned grid = (unsigned)s->nblk_z * s->nrhs;
  const size_t lds = s->tz.lds;
  // the damaged law is compiled for the two update modes only
  constexpr bool kDmg = MODE == Z_UPD || MODE == Z_AA;
#define LS_Z(EV, NTV)                                                                                            \
  if (kDmg && s->gd.dmg)                                                                                         \
    k_zline<MODE, EV, NTV, kDmg><<<grid, NTV, lds, s->stream>>>(s->gd, s->d_rhs, s->d_st, lam, mu, eps, s->spec, \
                                                                s->tz.lines, s->nblk_z, aa);                     \
  else                                                                                                           \
    k_zline<MODE, EV, NTV, false><<<grid, NTV, lds, s->stream>>>(s->gd, s->d_rhs, s->d_st, lam, mu, eps, s->spec, \
                                                                 s->tz.lines, s->nblk_z, aa)
  LS_DISPATCH32(s->tz, LS_Z)
#undef LS_Z
  mark(s, MODE == Z_AA ? 4 : MODE == Z_CG ? 6 : 0);
}

// y pass over all active load cases; record (inverse only, Anderson): store
// Im R of the self-conjugate planes for the Gram correction.
void ypass(ls_solver* s, bool inv, bool record = false) {
  AAPlanes pl;
  pl.im = (record && inv && s->depth > 0) ? s->aaPl : nullptr;
  pl.head = s->aa_head;
  pl.per_slot = s->pl_slot;
  pl.m1 = s->m1;
  const unsigned grid = (unsigned)s->nblk_y * s->nrhs;
  const size_t lds = s->ty.lds;
  if (s->ty.reg) {
#define LS_YR(NV, AV, BV, LV, NTV)                                                                      \
  else if (s->ny == NV && s->ty.lines == LV && s->ty.NT == NTV && s->ty.E == AV) {                     \
    if (inv)                                                                                            \
      k_ypass_r<AV, BV, LV, NTV, true><<<grid, NTV, lds, s->stream>>>(s->gd, s->d_st, s->spec, s->ntiles_y, pl); \
    else                                                                                                \
      k_ypass_r<AV, BV, LV, NTV, false><<<grid, NTV, lds, s->stream>>>(s->gd, s->d_st, s->spec, s->ntiles_y, pl); \
  }
    if (false) {
    }
    LS_REG_Y_LIST(LS_YR)
    else {
      fprintf(stderr, "liblippmannschwinger: no register y kernel for ny=%d T=%d\n", s->ny, s->ty.lines);
      s->launch_failed = true;
    }
#undef LS_YR
    mark(s, 1);
    return;
  }
  const FastDiv fT = make_fastdiv(s->ty.lines);
#define LS_Y(EV, NTV)                                                                                               \
  if (inv)                                                                                                          \
    k_ypass<true, EV, NTV, float2><<<grid, NTV, lds, s->stream>>>(s->gd, s->d_st, s->spec, fT, s->ntiles_y, -1, pl); \
  else                                                                                                              \
    k_ypass<false, EV, NTV, float2><<<grid, NTV, lds, s->stream>>>(s->gd, s->d_st, s->spec, fT, s->ntiles_y, -1, pl)
  LS_DISPATCH32(s->ty, LS_Y)
#undef LS_Y
  mark(s, 1);
}

// x pass over all active load cases; record: store the residual's spectral
// form and its Gram row for Anderson (solvers with depth > 0); cg: the
// conjugate-gradient form (running divergence in d_base, C0 norm partials).
void xpass(ls_solver* s, int do_gamma, bool with_base = true, bool record = false, bool cg = false) {
  AAParams aa = aa_params(s, record ? 1 : 0);
  aa.cg_part = s->cg_xpart;
  const int rec = cg ? 2 : aa.record;
  const unsigned grid = (unsigned)s->nblk_x * s->nrhs;
  const size_t lds = s->tx.lds;
  float2* db = with_base ? s->dbase : nullptr;
  if (s->tx.reg == 2) {
#define LS_XH(NV, AV, BV, LV, NTV, IP)                                                                    \
  else if (s->nx == NV && s->tx.lines == LV && s->tx.NT == NTV && s->tx.E == AV && s->tx.inpl == IP) {   \
    if (rec == 2)                                                                                         \
      k_xpass_h<AV, BV, LV, NTV, 2, IP><<<grid, NTV, lds, s->stream>>>(                                    \
          s->gd, s->d_rhs, s->d_st, s->spec, db, s->partials, s->dc, s->nblk_x, s->dl.per_rhs, do_gamma, aa); \
    else if (rec == 1)                                                                                    \
      k_xpass_h<AV, BV, LV, NTV, 1, IP><<<grid, NTV, lds, s->stream>>>(                                    \
          s->gd, s->d_rhs, s->d_st, s->spec, db, s->partials, s->dc, s->nblk_x, s->dl.per_rhs, do_gamma, aa); \
    else                                                                                                  \
      k_xpass_h<AV, BV, LV, NTV, 0, IP><<<grid, NTV, lds, s->stream>>>(                                    \
          s->gd, s->d_rhs, s->d_st, s->spec, db, s->partials, s->dc, s->nblk_x, s->dl.per_rhs, do_gamma, aa); \
  }
    if (false) {
    }
    LS_REG_H_LIST(LS_XH)
    else {
      fprintf(stderr, "liblippmannschwinger: no three-line x kernel for nx=%d TM=%d\n", s->nx, s->tx.lines);
      s->launch_failed = true;
    }
#undef LS_XH
    mark(s, 2);
    return;
  }
  const FastDiv fTm = make_fastdiv(s->tx.lines), fcols = make_fastdiv(6 * s->tx.lines);
#define LS_X(EV, NTV)                                                                                      \
  if (rec == 2)                                                                                             \
    k_xpass<EV, NTV, 2><<<grid, NTV, lds, s->stream>>>(s->gd, s->d_rhs, s->d_st, s->spec, db, s->partials, s->dc,  \
                                                       fTm, fcols, s->nblk_x, s->dl, do_gamma, aa);            \
  else if (rec == 1)                                                                                        \
    k_xpass<EV, NTV, 1><<<grid, NTV, lds, s->stream>>>(s->gd, s->d_rhs, s->d_st, s->spec, db, s->partials, s->dc,  \
                                                       fTm, fcols, s->nblk_x, s->dl, do_gamma, aa);            \
  else                                                                                                      \
    k_xpass<EV, NTV, 0><<<grid, NTV, lds, s->stream>>>(s->gd, s->d_rhs, s->d_st, s->spec, db, s->partials, s->dc,  \
                                                       fTm, fcols, s->nblk_x, s->dl, do_gamma, aa)
  LS_DISPATCH32(s->tx, LS_X)
#undef LS_X
  mark(s, 2);
}

void check(ls_solver* s, double rtol, double atol, int mode) {
  k_check<<<s->nrhs, kThreads, 0, s->stream>>>(s->d_st, s->partials, s->dc, s->nblk_x, rtol, atol, mode,
                                               s->absolute);
  mark(s, 3);
}

// conjugate-gradient scalars (see CgState in ls_kernels.hpp)
void cg_alpha(ls_solver* s) {
  k_cg_alpha<<<s->nrhs, kThreads, 0, s->stream>>>(s->d_st, s->d_cg, s->cg_zpart, s->nblk_z);
  mark(s, 3);
}
void cg_beta(ls_solver* s) {
  k_cg_beta<<<s->nrhs, kThreads, 0, s->stream>>>(s->d_st, s->d_rhs, s->d_cg, s->cg_xpart, s->nblk_x,
                                                 1.0 / (double)s->N);
  mark(s, 3);
}
void cg_reset(ls_solver* s, const int* d_sel) {
  k_cg_reset<<<(s->nrhs + 63) / 64, 64, 0, s->stream>>>(s->d_cg, d_sel, s->nrhs);
}
// apply the pending update of every load case, then clear it
void cg_finish(ls_solver* s, float* delta) {
  dim3 grid((unsigned)((6 * s->N + kThreads - 1) / kThreads), s->nrhs);
  k_cg_finish<<<grid, kThreads, 0, s->stream>>>(s->d_cg, s->N, s->cgP, delta);
  cg_reset(s, nullptr);
  mark(s, 3);
}

// FP64 evaluation of d_base for load case r from s->base (+tau).
void base_eval_one(ls_solver* s, int r, const float* lam, const float* mu, const float* tau) {
#define LS_BZ(EV, NTV)                                                                                          \
  k_base_z<EV, NTV><<<s->nblk_z64, NTV, s->tz64.lds, s->stream>>>(s->gd, s->d_rhs, r, lam, mu, s->base, tau, \
                                                                  s->spec64, s->tz64.lines)
  LS_DISPATCH64(s->tz64, LS_BZ)
#undef LS_BZ
  const FastDiv fT = make_fastdiv(s->ty64.lines);
  const AAPlanes nopl = {nullptr, nullptr, 0, 1};
#define LS_BY(EV, NTV)                                                                                        \
  k_ypass<false, EV, NTV, double2><<<s->nblk_y64, NTV, s->ty64.lds, s->stream>>>(s->gd, s->d_st, s->spec64,   \
                                                                                fT, s->ntiles_y64, r, nopl)
  LS_DISPATCH64(s->ty64, LS_BY)
#undef LS_BY
  const FastDiv fTm = make_fastdiv(s->tx64.lines), fcols = make_fastdiv(6 * s->tx64.lines);
#define LS_BX(EV, NTV)                                                                                         \
  k_base_x<EV, NTV><<<s->nblk_x64, NTV, s->tx64.lds, s->stream>>>(s->gd, r, s->spec64, s->dbase, s->d_st,     \
                                                                  fTm, fcols, s->dl)
  LS_DISPATCH64(s->tx64, LS_BX)
#undef LS_BX
  mark(s, 3);
}

int sync_state(ls_solver* s) {
  LS_HIP(hipMemcpyAsync(s->h_st, s->d_st, s->nrhs * sizeof(RhsState), hipMemcpyDeviceToHost, s->stream));
  LS_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

int put_state(ls_solver* s) {
  LS_HIP(hipMemcpyAsync(s->d_st, s->h_st, s->nrhs * sizeof(RhsState), hipMemcpyHostToDevice, s->stream));
  return 0;
}

int put_sel(ls_solver* s, const std::vector<int>& sel) {
  LS_HIP(hipMemcpyAsync(s->d_sel, sel.data(), s->nrhs * sizeof(int), hipMemcpyHostToDevice, s->stream));
  LS_HIP(hipStreamSynchronize(s->stream));  // sel is a host temporary
  return 0;
}

// Evaluate d_base for every load case active in the host state copy (which
// must be current).
void base_eval_active(ls_solver* s, const float* lam, const float* mu, const float* tau) {
  for (int r = 0; r < s->nrhs; ++r)
    if (s->h_st[r].active) base_eval_one(s, r, lam, mu, tau);
}

// The same for bases that are the uniform eps_bar of every active load case
// (the start of a solve) with Hooke's law and no polarisation: one FP64
// transform of lambda and mu per run of consecutive active load cases of one
// material, sigma_hat = tr(eps_bar) lambda_hat I + 2 mu_hat eps_bar
// (k_base_z<.., true>, k_base_x<.., true>), instead of six stress transforms
// per load case.  Otherwise the per-load-case evaluation.
void base_eval_start(ls_solver* s, const float* lam, const float* mu, const float* tau) {
  if (tau || s->gd.dmg || (int)s->h_mat.size() != s->nrhs) {
    base_eval_active(s, lam, mu, tau);
    return;
  }
  int r = 0;
  while (r < s->nrhs) {
    if (!s->h_st[r].active) {
      ++r;
      continue;
    }
    int q = r + 1;
    while (q < s->nrhs && s->h_st[q].active && s->h_mat[q] == s->h_mat[r]) ++q;
#define LS_BZL(EV, NTV)                                                                                    \
  k_base_z<EV, NTV, true><<<s->nblk_z64, NTV, s->tz64.lds, s->stream>>>(s->gd, s->d_rhs, r, lam, mu, s->base, \
                                                                        nullptr, s->spec64, s->tz64.lines)
    LS_DISPATCH64(s->tz64, LS_BZL)
#undef LS_BZL
    const FastDiv fT = make_fastdiv(s->ty64.lines);
    const AAPlanes nopl = {nullptr, nullptr, 0, 1};
#define LS_BY(EV, NTV)                                                                                        \
  k_ypass<false, EV, NTV, double2><<<s->nblk_y64, NTV, s->ty64.lds, s->stream>>>(s->gd, s->d_st, s->spec64,   \
                                                                                fT, s->ntiles_y64, r, nopl)
    LS_DISPATCH64(s->ty64, LS_BY)
#undef LS_BY
    const FastDiv fTm = make_fastdiv(s->tx64.lines), fcols = make_fastdiv(6 * s->tx64.lines);
#define LS_BXL(EV, NTV)                                                                                        \
  k_base_x<EV, NTV, true><<<s->nblk_x64, NTV, s->tx64.lds, s->stream>>>(s->gd, r, s->spec64, s->dbase, s->d_st, \
                                                                        fTm, fcols, s->dl, s->d_rhs, q - r)
    LS_DISPATCH64(s->tx64, LS_BXL)
#undef LS_BXL
    r = q;
  }
  mark(s, 3);
}

int zero_spec(ls_solver* s, const std::vector<int>& which) {
  for (int r = 0; r < s->nrhs; ++r)
    if (which[r])
      LS_HIP(hipMemsetAsync(s->spec + (size_t)r * s->M * 6, 0, (size_t)s->M * 6 * sizeof(float2), s->stream));
  return 0;
}

int any_active(const ls_solver* s) {
  for (int r = 0; r < s->nrhs; ++r)
    if (s->h_st[r].active) return 1;
  return 0;
}

std::vector<int> active_mask(const ls_solver* s) {
  std::vector<int> m(s->nrhs);
  for (int r = 0; r < s->nrhs; ++r) m[r] = s->h_st[r].active ? 1 : 0;
  return m;
}

void free_solver(ls_solver* s) {
  if (s && s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (!s) return;
  if (s->arena_mode) {
    // device buffers belong to the caller's region, host buffers to host_mem
    for (auto e : s->tev) (void)hipEventDestroy(e);
    delete s;
    return;
  }
  for (auto& a : s->ax) {
    (void)hipFree(a.d_tw);
    (void)hipFree(a.d_tw64);
    (void)hipFree(a.d_s);
    (void)hipFree(a.d_c);
    (void)hipFree(a.d_ph);
  }
  void* dev[] = {s->spec,    s->spec64,  s->base,    s->dbase,    s->partials, s->dc,       s->d_rhs,
                 s->d_st,    s->d_sel,   s->part6,   s->d_mean6,  s->d_mm,     s->aaG,      s->aaU,      s->aaPl,
                 s->aa_alpha, s->aa_alpha_prev, s->aa_dots, s->aa_gram, s->aa_head, s->aa_filled,
                 s->cgP,      s->d_cg,    s->cg_zpart, s->cg_xpart};
  for (void* p : dev) (void)hipFree(p);
  (void)hipHostFree(s->h_st);
  (void)hipHostFree(s->h_mean6);
  (void)hipHostFree(s->h_mm);
  for (auto e : s->tev) (void)hipEventDestroy(e);
  if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// Block counts and the d_base layout of the current FP32 tiles.  The d_base
// stride per load case is fixed for any x tile of at most kMaxTm modes.
void apply_tiles(ls_solver* s) {
  s->nblk_z = (int)((s->nx * (long long)s->ny + s->tz.lines - 1) / s->tz.lines);
  s->ntiles_y = (6 * s->P + s->ty.lines - 1) / s->ty.lines;
  s->nblk_y = s->nx * s->ntiles_y;
  s->nblk_x = (int)(((long long)s->ny * s->P + s->tx.lines - 1) / s->tx.lines);
  s->dl.f_Tm = make_fastdiv(s->tx.lines);
  s->dl.hA = s->tx.reg == 2 ? s->tx.E : 0;
  s->dl.f_A = make_fastdiv(s->dl.hA > 0 ? s->dl.hA : 1);
  s->dl.nx = s->nx;
  // a multiple of 4 complex entries, so every slot of the Anderson y ring
  // starts 32-byte aligned
  s->dl.per_rhs = ((((long long)s->ny * s->P + kMaxTm) * s->nx * 3) + 3) / 4 * 4;
}

int create_impl(ls_solver* s, const int* voxels, const float* extents, int nrhs, int nmat, int depth,
                void* stream) {
  s->nx = voxels[0];
  s->ny = voxels[1];
  s->nz = voxels[2];
  s->Lx = extents[0];
  s->Ly = extents[1];
  s->Lz = extents[2];
  if (s->nx < 1 || s->ny < 1 || s->nz < 1 || s->nx > 2048 || s->ny > 2048 || s->nz > 2048) {
    fprintf(stderr, "liblippmannschwinger: grid %d x %d x %d outside 1..2048 per axis\n", s->nx, s->ny, s->nz);
    return -1;
  }
  if (!(s->Lx > 0.f && s->Ly > 0.f && s->Lz > 0.f)) {
    fprintf(stderr, "liblippmannschwinger: extents must be positive\n");
    return -1;
  }
  if (nrhs < 1 || nmat < 1 || depth < 0 || depth > kMaxDepth1 - 1) {
    fprintf(stderr, "liblippmannschwinger: need nrhs >= 1, nmat >= 1, 0 <= depth <= %d\n", kMaxDepth1 - 1);
    return -1;
  }
  s->P = s->nz / 2 + 1;
  s->N = (long long)s->nx * s->ny * s->nz;
  s->M = (long long)s->nx * s->ny * s->P;
  s->nrhs = nrhs;
  s->nmat = nmat;
  s->depth = depth;
  s->m1 = depth + 1;
  if (const char* rb = std::getenv("LS_REBASE_FACTOR")) s->rebase_factor = std::atof(rb);
  if (stream || s->arena_mode) {
    s->stream = (hipStream_t)stream;
  } else {
    LS_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    s->own_stream = true;
  }
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      s->ncu = cus;
  }
  const int n[3] = {s->nx, s->ny, s->nz};
  // the z-axis kernels are compiled without the generic FFT stage
  for (int d = 0; d < 3; ++d)
    if (build_axis(s->ax[d], n[d], d != 2)) return -1;
  // tiles: aim for ~3-6k complex elements per block in y and x, ~1.5k in z
  // (two lines of six components at 256: 2% faster z passes than four), but
  // on small grids no more lines than leave about one block per CU: there a
  // pass is latency-bound and more, smaller blocks finish sooner
  // (overridable for tuning)
  const int ncu = s->ncu;
  auto per_cu = [ncu](int w, long long units) {
    return std::max(1, (int)std::min<long long>(w, std::max(1LL, units / ncu)));
  };
  const int wz = env_int("LS_TILE_Z", per_cu(std::max(1, 1536 / (3 * s->nz)), (long long)s->nx * s->ny));
  const int wy = env_int("LS_TILE_Y", per_cu(std::max(1, std::min(6 * s->P, 4096 / s->ny)), 6LL * s->nx * s->P));
  const int wx = env_int("LS_TILE_X", per_cu(std::max(1, 3072 / (6 * s->nx)), (long long)s->ny * s->P));
  const int lsz = s->nz + (s->nz >> 4) + 1;  // padded z-line LDS stride
  const bool ok32 = choose_tile(s->ax[2], wz, 3, (size_t)3 * lsz * sizeof(float2), kBudget32, s->tz) &&
                    choose_tile(s->ax[1], wy, 1, (size_t)s->ny * sizeof(float2), kBudget32, s->ty) &&
                    choose_tile(s->ax[0], wx, 6, (size_t)6 * s->nx * sizeof(float2), kBudget32, s->tx);
  const bool ok64 =
      choose_tile(s->ax[2], std::max(1, 2048 / (3 * s->nz)), 3, (size_t)3 * lsz * sizeof(double2), kBudget64, s->tz64) &&
      choose_tile(s->ax[1], std::max(1, std::min(6 * s->P, 3584 / s->ny)), 1, (size_t)s->ny * sizeof(double2), kBudget64,
                  s->ty64) &&
      choose_tile(s->ax[0], std::max(1, 2048 / (6 * s->nx)), 6, (size_t)6 * s->nx * sizeof(double2), kBudget64, s->tx64);
  if (!ok32 || !ok64) {
    fprintf(stderr, "liblippmannschwinger: no FFT tile fits for grid %d x %d x %d\n", s->nx, s->ny, s->nz);
    return -1;
  }
  if (env_int("LS_REG_FFT", 1)) {
    const std::vector<Tile> ry = reg_tiles(kRegY, sizeof(kRegY) / sizeof(kRegY[0]), s->ny, 1);
    const std::vector<Tile> rh = reg_tiles(kRegH, sizeof(kRegH) / sizeof(kRegH[0]), s->nx, 2);
    // LS_REG_Y_INDEX / LS_REG_X_INDEX select another compiled configuration
    // of the length than its default (the first)
    const int iy = env_int("LS_REG_Y_INDEX", 0), ix = env_int("LS_REG_X_INDEX", 0);
    if ((!ry.empty() && (iy < 0 || iy >= (int)ry.size())) || (!rh.empty() && (ix < 0 || ix >= (int)rh.size()))) {
      fprintf(stderr, "liblippmannschwinger: LS_REG_Y_INDEX / LS_REG_X_INDEX out of range\n");
      return -1;
    }
    if (!ry.empty()) s->ty = ry[iy];
    if (!rh.empty()) s->tx = rh[ix];
  }
  if (s->tx.lines > kMaxTm) {
    fprintf(stderr, "liblippmannschwinger: x tile of %d modes exceeds %d\n", s->tx.lines, kMaxTm);
    return -1;
  }
  apply_tiles(s);
  s->nblk_z64 = (int)((s->nx * (long long)s->ny + s->tz64.lines - 1) / s->tz64.lines);
  s->ntiles_y64 = (6 * s->P + s->ty64.lines - 1) / s->ty64.lines;
  s->nblk_y64 = s->nx * s->ntiles_y64;
  s->nblk_x64 = (int)(((long long)s->ny * s->P + s->tx64.lines - 1) / s->tx64.lines);
  GridDev& g = s->gd;
  g.nx = s->nx;
  g.ny = s->ny;
  g.nz = s->nz;
  g.P = s->P;
  g.N = s->N;
  g.M = s->M;
  g.gx = (float)(2.0 * s->nx / (double)s->Lx);
  g.gy = (float)(2.0 * s->ny / (double)s->Ly);
  g.gz = (float)(2.0 * s->nz / (double)s->Lz);
  g.sx = s->ax[0].d_s;
  g.cx = s->ax[0].d_c;
  g.sy = s->ax[1].d_s;
  g.cy = s->ax[1].d_c;
  g.sz = s->ax[2].d_s;
  g.cz = s->ax[2].d_c;
  g.phx = s->ax[0].d_ph;
  g.px = to_dev(s->ax[0]);
  g.py = to_dev(s->ax[1]);
  g.pz = to_dev(s->ax[2]);
  g.f_nz = make_fastdiv(s->nz);
  g.f_P = make_fastdiv(s->P);
  g.lsz = lsz;
  // buffers whose size depends on the FP32 tiles are sized for any tile (so the
  // tiles can change between solves, ls_solver_autotune)
  if (dalloc(&s->spec, (size_t)nrhs * s->M * 6) || dalloc(&s->spec64, (size_t)s->M * 6) ||
      dalloc(&s->base, (size_t)nrhs * 6 * s->N) || dalloc(&s->dbase, (size_t)nrhs * s->dl.per_rhs) ||
      dalloc(&s->partials, (size_t)nrhs * s->ny * s->P) || dalloc(&s->dc, (size_t)nrhs * 6) ||
      dalloc(&s->d_rhs, nrhs) || dalloc(&s->d_st, nrhs) || dalloc(&s->d_sel, nrhs))
    return -1;
  g.base = s->base;
  g.dmg = nullptr;  // Hooke's law until ls_solver_set_damage
  g.kdmg = 0.f;
  if (halloc(s, &s->h_st, nrhs)) return -1;
  s->chunk_stress = 8192;
  s->nblk_stress = (int)((s->N + s->chunk_stress - 1) / s->chunk_stress);
  if (dalloc(&s->part6, (size_t)nrhs * s->nblk_stress * 6) || dalloc(&s->d_mean6, (size_t)nrhs * 6)) return -1;
  if (halloc(s, &s->h_mean6, (size_t)nrhs * 6)) return -1;
  s->chunk_mm = 65536;
  s->nblk_mm = (int)((s->N + s->chunk_mm - 1) / s->chunk_mm);
  if (dalloc(&s->d_mm, (size_t)nmat * s->nblk_mm * 4)) return -1;
  if (halloc(s, &s->h_mm, (size_t)nmat * s->nblk_mm * 4)) return -1;
  if (depth > 0) {
    const size_t hist = (size_t)nrhs * s->m1 * 6 * s->N;
    // y ring in the d_base layout; Gram-row partials per k_aa_gram block
    // 16k complex per block on large grids; on small ones about 512 per block
    // (up to one block per CU), so the latency-bound row finishes sooner
    s->nblk_g = (int)std::min<long long>(
        2048, std::max((s->dl.per_rhs + 16383) / 16384, std::min<long long>(s->ncu, (s->dl.per_rhs + 511) / 512)));
    s->chunk_g = ((s->dl.per_rhs + s->nblk_g - 1) / s->nblk_g + 1) / 2 * 2;
    // self-conjugate planes: k_z = 0, and k_z = nz/2 for even nz
    s->pl_slot = (long long)(s->nz % 2 == 0 ? 2 : 1) * s->nx * s->ny * 6;
    s->nblk_p = (int)std::min<long long>(
        1024, std::max((s->pl_slot + 2047) / 2048, std::min<long long>(s->ncu, (s->pl_slot + 255) / 256)));
    s->chunk_p = (s->pl_slot + s->nblk_p - 1) / s->nblk_p;
    if (dalloc(&s->aaG, hist) || dalloc(&s->aaU, (size_t)nrhs * s->m1 * s->dl.per_rhs) ||
        dalloc(&s->aaPl, (size_t)nrhs * s->m1 * s->pl_slot) || dalloc(&s->aa_alpha, (size_t)nrhs * s->m1) ||
        dalloc(&s->aa_alpha_prev, (size_t)nrhs * s->m1) ||
        dalloc(&s->aa_dots, (size_t)nrhs * (s->nblk_g + s->nblk_p) * s->m1) ||
        dalloc(&s->aa_gram, (size_t)nrhs * s->m1 * s->m1) ||
        dalloc(&s->aa_head, nrhs) || dalloc(&s->aa_filled, nrhs))
      return -1;
  }
  return 0;
}

int reset_aa(ls_solver* s) {
  if (s->depth == 0) return 0;
  std::vector<double> gram((size_t)s->nrhs * s->m1 * s->m1, 0.0);
  for (int r = 0; r < s->nrhs; ++r)
    for (int i = 0; i < s->m1; ++i) gram[((size_t)r * s->m1 + i) * s->m1 + i] = 1.0;
  std::vector<int> head(s->nrhs, s->m1 - 1), filled(s->nrhs, 0);
  LS_HIP(hipMemcpyAsync(s->aa_gram, gram.data(), gram.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
  LS_HIP(hipMemcpyAsync(s->aa_head, head.data(), head.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
  LS_HIP(hipMemcpyAsync(s->aa_filled, filled.data(), filled.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
  // entries of the y ring outside the grid's modes are never written: keep them zero
  LS_HIP(hipMemsetAsync(s->aaU, 0, (size_t)s->nrhs * s->m1 * s->dl.per_rhs * sizeof(float2), s->stream));
  LS_HIP(hipMemsetAsync(s->aaPl, 0, (size_t)s->nrhs * s->m1 * s->pl_slot * sizeof(float), s->stream));
  // no weights yet: the first rebuilt delta is 0
  LS_HIP(hipMemsetAsync(s->aa_alpha, 0, (size_t)s->nrhs * s->m1 * sizeof(double), s->stream));
  LS_HIP(hipMemsetAsync(s->aa_alpha_prev, 0, (size_t)s->nrhs * s->m1 * sizeof(double), s->stream));
  LS_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

void init_fields(ls_solver* s, const int* d_sel, float* delta, const float* src) {
  dim3 grid((unsigned)((6 * s->N + kThreads - 1) / kThreads), s->nrhs);
  k_init_fields<<<grid, kThreads, 0, s->stream>>>(s->d_rhs, d_sel, s->N, s->base, delta, src);
}

// Residual evaluation of the current iterate from the spectrum after the
// forward z and y passes: x pass (residual, Gamma0) and stop test.  With
// Anderson the x pass records the residual's y form (in a new ring slot if
// `advance`, else replacing the newest entry, e.g. after a rebase); the
// weights follow the inverse y pass of the next iteration (aa_weights).
void eval_x(ls_solver* s, double rtol, double atol, int mode, bool advance) {
  if (s->cg) {
    xpass(s, 1, true, false, true);
    check(s, rtol, atol, mode);
    cg_beta(s);
    return;
  }
  if (s->depth > 0 && advance)
    k_aa_advance<<<(s->nrhs + 63) / 64, 64, 0, s->stream>>>(s->d_st, s->aa_head, s->aa_filled, s->aa_gram,
                                                           s->aa_alpha, s->aa_alpha_prev, s->nrhs,
                                                           s->m1);
  xpass(s, 1, true, s->depth > 0);
  check(s, rtol, atol, mode);
}

// Anderson weights of the next z pass from the newest record (x pass y ring +
// plane correction from the inverse y pass): Gram row, then the solve.
void aa_weights(ls_solver* s) {
  const int stride = s->nblk_g + s->nblk_p;
  k_aa_gram<<<dim3((unsigned)s->nblk_g, (unsigned)s->nrhs), kThreads, 0, s->stream>>>(s->d_st, aa_params(s, 1),
                                                                                    s->aa_dots, s->chunk_g, stride);
  AAPlanes pl = {s->aaPl, s->aa_head, s->pl_slot, s->m1};
  k_aa_gram_planes<<<dim3((unsigned)s->nblk_p, (unsigned)s->nrhs), kThreads, 0, s->stream>>>(
      s->d_st, pl, s->aa_filled, s->aa_dots, s->chunk_p, stride, s->nblk_g, 1.0 / ((double)s->nx * s->ny));
  mark(s, 5);
  k_aa_solve<<<s->nrhs, kThreads, 0, s->stream>>>(s->d_st, s->aa_dots, stride, s->m1, s->aa_gram, s->aa_alpha,
                                                  s->aa_head, s->aa_filled);
  mark(s, 3);
}

// One residual evaluation after a z-line forward pass: y fwd, then eval_x.
void evaluate(ls_solver* s, double rtol, double atol, int mode) {
  ypass(s, false);
  eval_x(s, rtol, atol, mode, true);
}

// One iteration for all active load cases (the spectrum holds Gamma0(d) after
// the x pass of the previous check).
void iterate(ls_solver* s, const float* lam, const float* mu, float* eps, double rtol, double atol,
             bool evaluate_after) {
  if (s->cg) {
    // the z pass applies the previous step's update and forms this step's
    // direction; alpha makes this step's update pending (applied by the next
    // z pass or, after the last iteration, by cg_finish)
    ypass(s, true, false);
    zline<Z_CG>(s, lam, mu, eps);
    cg_alpha(s);
    if (evaluate_after) evaluate(s, rtol, atol, 0);
    return;
  }
  ypass(s, true, s->depth > 0);
  if (s->depth == 0) {
    zline<Z_UPD>(s, lam, mu, eps);
  } else {
    aa_weights(s);
    zline<Z_AA>(s, lam, mu, eps);
  }
  if (evaluate_after) evaluate(s, rtol, atol, 0);
}

// `steps` iterations (each followed by its residual evaluation) as one
// replay of a captured HIP graph.  The graph bakes in the kernel arguments, so
// it is captured again when the fields, tolerances or chunk length change.
// The captured work has no memset nodes: iterate() launches kernels only.
int iterate_graph(ls_solver* s, const float* lam, const float* mu, float* eps, double rtol, double atol,
                  int steps) {
  const bool same = s->gexec && s->gkey[0] == lam && s->gkey[1] == mu && s->gkey[2] == eps && s->gtol[0] == rtol &&
                    s->gtol[1] == atol && s->gsteps == steps && s->gabs == s->absolute && s->gcg == s->cg;
  if (!same) {
    if (s->gexec) {
      LS_HIP(hipGraphExecDestroy(s->gexec));
      s->gexec = nullptr;
    }
    hipGraph_t graph = nullptr;
    LS_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < steps; ++k) iterate(s, lam, mu, eps, rtol, atol, true);
    LS_HIP(hipStreamEndCapture(s->stream, &graph));
    const hipError_t e = hipGraphInstantiate(&s->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    LS_HIP(e);
    s->gkey[0] = lam;
    s->gkey[1] = mu;
    s->gkey[2] = eps;
    s->gtol[0] = rtol;
    s->gtol[1] = atol;
    s->gsteps = steps;
    s->gabs = s->absolute;
    s->gcg = s->cg;
  }
  LS_HIP(hipGraphLaunch(s->gexec, s->stream));
  return 0;
}

// The Anderson iteration keeps delta in its ring (k_aa_delta): write it into
// its buffer for the selected load cases.
void aa_write_delta(ls_solver* s, const int* d_sel, float* eps, int last_if_active) {
  if (s->depth == 0 || s->cg) return;
  dim3 grid((unsigned)((6 * s->N + kThreads - 1) / kThreads), s->nrhs);
  k_aa_delta<<<grid, kThreads, 0, s->stream>>>(s->d_st, d_sel, s->N, s->aaG, s->aa_head, s->aa_filled, s->aa_alpha,
                                               s->aa_alpha_prev, s->m1, last_if_active, eps);
}

// Move the corrections of all active load cases into their bases and
// re-evaluate (host state copy must be current).
int rebase(ls_solver* s, const float* lam, const float* mu, const float* tau, float* eps, double rtol,
           double atol) {
  dim3 grid((unsigned)((6 * s->N + kThreads - 1) / kThreads), s->nrhs);
  // conjugate gradients: the pending update first; the recurrence restarts
  // from the exact divergence of the new base
  if (s->cg) cg_finish(s, eps);
  k_rebase<<<grid, kThreads, 0, s->stream>>>(s->d_st, s->N, s->base, eps, s->cg ? nullptr : s->aaG, s->aa_head,
                                              s->aa_filled, s->aa_alpha_prev, s->m1);
  if (s->depth > 0 && !s->cg)
    k_aa_forget<<<(s->nrhs + 63) / 64, 64, 0, s->stream>>>(s->d_st, s->aa_alpha_prev, s->nrhs, s->m1);
  base_eval_active(s, lam, mu, tau);
  if (zero_spec(s, active_mask(s))) return -1;
  eval_x(s, rtol, atol, 2, false);
  ++s->n_rebase;
  return 0;
}

int solve_impl(ls_solver* s, const float* d_lambda, const float* d_mu, const int* rhs_mat, const float* eps_bar,
               const float* d_tau, float* d_eps, float* d_sigma, int warm_start, float rtol, float atol,
               int maxiter, double rel_ref, int verbose, int* iters, double* residual, double* sigma_mean,
               bool legacy_q7, int flags = 0) {
  s->absolute = (flags & 1) ? 1 : 0;
  s->cg = (flags & LS_CG) ? 1 : 0;
  if (maxiter < 0) {
    fprintf(stderr, "liblippmannschwinger: maxiter must be >= 0\n");
    return -1;
  }
  if (s->cg) {
    if (s->gd.dmg) {
      fprintf(stderr, "liblippmannschwinger: conjugate gradients need a linear material law (damage is set)\n");
      return -1;
    }
    if (!s->cgP) {
      if (s->arena_mode) {
        fprintf(stderr, "liblippmannschwinger: conjugate gradients are not available without allocation\n");
        return -1;
      }
      if (dalloc(&s->cgP, (size_t)s->nrhs * 6 * s->N) || dalloc(&s->d_cg, s->nrhs) ||
          dalloc(&s->cg_zpart, (size_t)s->nrhs * s->nx * s->ny) ||
          dalloc(&s->cg_xpart, (size_t)s->nrhs * s->ny * s->P))
        return -1;
    }
    // no direction yet (the first beta is 0, and 0 * p must be 0)
    LS_HIP(hipMemsetAsync(s->cgP, 0, (size_t)s->nrhs * 6 * s->N * sizeof(float), s->stream));
    cg_reset(s, nullptr);
  }
  for (auto e : s->tev) (void)hipEventDestroy(e);
  s->tev.clear();
  s->tkind.clear();
  s->n_rebase = 0;
  s->launch_failed = false;
  // reference medium per material: mid-range of the fields (lippmann_schwinger.cu:172-173)
  {
    dim3 grid(s->nblk_mm, s->nmat);
    k_minmax<<<grid, kThreads, 0, s->stream>>>(d_lambda, d_mu, s->N, s->d_mm, s->chunk_mm);
    LS_HIP(hipGetLastError());
    LS_HIP(hipMemcpyAsync(s->h_mm, s->d_mm, (size_t)s->nmat * s->nblk_mm * 4 * sizeof(float),
                          hipMemcpyDeviceToHost, s->stream));
    LS_HIP(hipStreamSynchronize(s->stream));
  }
  std::vector<float> lam0(s->nmat), mu0(s->nmat);
  for (int m = 0; m < s->nmat; ++m) {
    float a = 3.4e38f, b = -3.4e38f, c = 3.4e38f, d = -3.4e38f;
    for (int i = 0; i < s->nblk_mm; ++i) {
      const float* o = s->h_mm + ((size_t)m * s->nblk_mm + i) * 4;
      a = std::min(a, o[0]);
      b = std::max(b, o[1]);
      c = std::min(c, o[2]);
      d = std::max(d, o[3]);
    }
    lam0[m] = (float)(0.5 * (b + a));
    mu0[m] = (float)(0.5 * (d + c));
    if (!(mu0[m] > 0.f) || !(lam0[m] + 2.f * mu0[m] > 0.f) || !std::isfinite(lam0[m]) ||
        !std::isfinite(mu0[m])) {
      fprintf(stderr, "liblippmannschwinger: invalid reference medium lambda0=%g mu0=%g (material %d)\n",
              lam0[m], mu0[m], m);
      return -1;
    }
  }
  std::vector<RhsDesc> desc(s->nrhs);
  const double invN = 1.0 / (double)s->N;
  for (int r = 0; r < s->nrhs; ++r) {
    const int m = rhs_mat ? rhs_mat[r] : 0;
    if (m < 0 || m >= s->nmat) {
      fprintf(stderr, "liblippmannschwinger: rhs %d refers to material %d of %d\n", r, m, s->nmat);
      return -1;
    }
    RhsDesc& d = desc[r];
    d.mat = m;
    for (int c = 0; c < 6; ++c) d.eps_bar[c] = eps_bar[r * 6 + c];
    const double l0 = lam0[m], m0 = mu0[m];
    d.cA = (float)(-1.0 / m0 * invN);
    d.cB = (float)((l0 + m0) / (m0 * (l0 + 2.0 * m0)) * invN);
    d.lam0 = lam0[m];
    d.mu0 = mu0[m];
  }
  LS_HIP(hipMemcpyAsync(s->d_rhs, desc.data(), s->nrhs * sizeof(RhsDesc), hipMemcpyHostToDevice, s->stream));
  s->h_mat.resize(s->nrhs);
  for (int r = 0; r < s->nrhs; ++r) s->h_mat[r] = desc[r].mat;
  for (int r = 0; r < s->nrhs; ++r) {
    RhsState& q = s->h_st[r];
    memset(&q, 0, sizeof(q));
    q.R0 = rel_ref > 0.0 ? rel_ref : 0.0;
    q.active = 1;
  }
  if (put_state(s)) return -1;
  if (reset_aa(s)) return -1;
  s->timing = !s->arena_mode && (std::getenv("LS_TIMING") != nullptr);
  mark(s, 3);
  const double drtol = rtol, datol = atol;
  std::vector<int> all(s->nrhs, 1);

  // ---- start ----
  if (warm_start) {
    // 1. residual of the uniform start, the normaliser of the tolerance, and
    //    its equilibrium residual norm ||xi . sigma_hat|| = R ||sigma_hat(0)||
    std::vector<double> abs0(s->nrhs, -1.0);
    if (!(rel_ref > 0.0)) {
      init_fields(s, nullptr, nullptr, nullptr);
      base_eval_start(s, d_lambda, d_mu, d_tau);
      if (zero_spec(s, all)) return -1;
      xpass(s, 0);
      check(s, drtol, datol, 1);
      if (sync_state(s)) return -1;
      for (int r = 0; r < s->nrhs; ++r) {
        RhsState& q = s->h_st[r];
        q.R0 = q.R;
        abs0[r] = s->absolute ? q.R : q.R * q.den;
        q.checks = 0;
        q.active = q.status == 0;
      }
      if (put_state(s)) return -1;
    }
    // 2. project the start onto compatible fields with mean eps_bar (FP32),
    //    unless the caller's start is compatible already (LS_COMPATIBLE_START)
    if (!(flags & LS_COMPATIBLE_START)) {
      zline<Z_PROJ>(s, d_lambda, d_mu, d_eps);
      ypass(s, false);
      xpass(s, 1, false);
      ypass(s, true);
      zline<Z_SET>(s, d_lambda, d_mu, d_eps);
    }
    // 3. base = projected start, correction 0; evaluate
    init_fields(s, nullptr, d_eps, d_eps);
    base_eval_active(s, d_lambda, d_mu, d_tau);
    if (zero_spec(s, all)) return -1;
    eval_x(s, drtol, datol, 0, true);
    if (sync_state(s)) return -1;
    // 4. fall back to the uniform start where the warm start is worse: by the
    //    equilibrium residual norms of the two starts (their normalised R
    //    divide by different mean stresses, which differ most for shear load
    //    cases), or by R against the caller's normaliser rel_ref
    std::vector<int> restart(s->nrhs), resume(s->nrhs);
    bool any_restart = false;
    for (int r = 0; r < s->nrhs; ++r) {
      const RhsState& q = s->h_st[r];
      const bool worse = abs0[r] >= 0.0 ? (s->absolute ? q.R : q.R * q.den) > abs0[r] : q.R > q.R0;
      restart[r] = (q.status != 0) || worse;
      resume[r] = !restart[r] && q.active;
      any_restart = any_restart || restart[r];
    }
    if (any_restart) {
      for (int r = 0; r < s->nrhs; ++r) {
        RhsState& q = s->h_st[r];
        q.active = restart[r];
        if (restart[r]) {
          q.checks = 0;
          q.converged = 0;
          q.status = 0;
        }
      }
      if (put_state(s) || put_sel(s, restart)) return -1;
      init_fields(s, s->d_sel, d_eps, nullptr);
      base_eval_start(s, d_lambda, d_mu, d_tau);
      if (zero_spec(s, restart)) return -1;
      if (s->cg) cg_reset(s, s->d_sel);
      eval_x(s, drtol, datol, 0, false);
      if (sync_state(s)) return -1;
      for (int r = 0; r < s->nrhs; ++r)
        if (resume[r]) s->h_st[r].active = 1;
      if (put_state(s)) return -1;
    }
  } else {
    init_fields(s, nullptr, d_eps, nullptr);
    base_eval_start(s, d_lambda, d_mu, d_tau);
    if (zero_spec(s, all)) return -1;
    eval_x(s, drtol, datol, 0, true);
  }
  LS_HIP(hipGetLastError());

  // ---- loop ----
  double R0_print = 0.0;
  if (verbose == 2) {
    if (sync_state(s)) return -1;
    R0_print = s->h_st[0].R0;
    printf("==== Lippmann Schwinger solver ====\n");
    printf("  iteration           ||r||   ||r||/||r_0||\n");
    printf("     %4d          %8.4e  %8.4e\n", 0, s->h_st[0].R, s->h_st[0].R / R0_print);
  }
  const int sync_every = verbose == 2 ? 1 : env_int("LS_SYNC_EVERY", s->N >= (1LL << 21) ? 2 : 8);
  // graph replay of the iteration chunks (not with per-kernel timing, whose events it would capture)
  const bool use_graph = (flags & 2) && !s->timing && sync_every > 1;
  bool q7_done = false;
  std::vector<int> q7(s->nrhs, 0);
  for (int it = 1; it <= maxiter; ++it) {
    const bool last = (it == maxiter);
    if (last || (it - 1) % sync_every == 0) {
      if (sync_state(s)) return -1;
      if (!any_active(s)) break;
      if (verbose == 2 && it > 1)
        printf("     %4d          %8.4e  %8.4e\n", it - 1, s->h_st[0].R, s->h_st[0].R / R0_print);
      // move the correction into the base once the residual has dropped enough
      bool want = false;
      for (int r = 0; r < s->nrhs; ++r) {
        const RhsState& q = s->h_st[r];
        if (q.active && q.R < s->rebase_factor * q.Rbase) want = true;
      }
      if (want && !last) {
        if (rebase(s, d_lambda, d_mu, d_tau, d_eps, drtol, datol)) return -1;
        if (sync_state(s)) return -1;
        if (!any_active(s)) break;
      }
    }
    if (last && legacy_q7 && d_sigma) {
      // reference quirk: when the loop runs out, the returned stress is the one
      // computed before the final update (lippmann_schwinger.cu:192-216, 230-231)
      for (int r = 0; r < s->nrhs; ++r) q7[r] = s->h_st[r].active ? 1 : 0;
      if (put_sel(s, q7)) return -1;
      aa_write_delta(s, s->d_sel, d_eps, 0);
      dim3 grid(s->nblk_stress, s->nrhs);
      k_stress<<<grid, kThreads, 0, s->stream>>>(s->gd, s->d_rhs, s->d_sel, d_lambda, d_mu, s->base, d_eps, 0,
                                                  d_tau, d_sigma, s->part6, s->chunk_stress);
      q7_done = true;
    }
    // a whole chunk up to the next synchronisation as one graph replay
    if (use_graph && (it - 1) % sync_every == 0 && it + sync_every - 1 < maxiter) {
      if (iterate_graph(s, d_lambda, d_mu, d_eps, drtol, datol, sync_every)) return -1;
      it += sync_every - 1;
      continue;
    }
    iterate(s, d_lambda, d_mu, d_eps, drtol, datol, !last);
  }
  LS_HIP(hipGetLastError());
  if (sync_state(s)) return -1;
  if (s->launch_failed) return -1;

  // ---- outputs: eps = base + delta, sigma = C eps + tau, FP64 means ----
  {
    dim3 grid(s->nblk_stress, s->nrhs);
    std::vector<int> sel(s->nrhs, 1);
    if (q7_done)
      for (int r = 0; r < s->nrhs; ++r) sel[r] = q7[r] ? 0 : 1;
    if (put_sel(s, sel)) return -1;
    aa_write_delta(s, nullptr, d_eps, 1);
    if (s->cg) cg_finish(s, d_eps);
    k_stress<<<grid, kThreads, 0, s->stream>>>(s->gd, s->d_rhs, s->d_sel, d_lambda, d_mu, s->base, d_eps, 1, d_tau,
                                                d_sigma, s->part6, s->chunk_stress);
    if (q7_done) {
      // strain of the load cases whose stress was saved before the last update
      if (put_sel(s, q7)) return -1;
      k_stress<<<grid, kThreads, 0, s->stream>>>(s->gd, s->d_rhs, s->d_sel, d_lambda, d_mu, s->base, d_eps, 1,
                                                  d_tau, nullptr, nullptr, s->chunk_stress);
    }
    k_sum6<<<s->nrhs, kThreads, 0, s->stream>>>(s->part6, s->nblk_stress, s->d_mean6, invN);
  }
  LS_HIP(hipGetLastError());
  LS_HIP(hipMemcpyAsync(s->h_mean6, s->d_mean6, s->nrhs * 6 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  LS_HIP(hipStreamSynchronize(s->stream));

  if (s->timing) {
    for (int k = 0; k < ls_solver::kKinds; ++k) {
      s->t_ms[k] = 0.0;
      s->t_cnt[k] = 0;
    }
    for (size_t i = 1; i < s->tev.size(); ++i) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, s->tev[i - 1], s->tev[i]);
      s->t_ms[s->tkind[i]] += ms;
      s->t_cnt[s->tkind[i]] += 1;
    }
  }

  int rc = 0;
  for (int r = 0; r < s->nrhs; ++r) {
    const RhsState& q = s->h_st[r];
    const int it = q.converged ? q.checks - 1 : maxiter;
    if (!q.converged) rc = 1;
    if (iters) iters[r] = it;
    if (residual) residual[r] = q.R0 > 0.0 ? q.R / q.R0 : q.R;
    if (sigma_mean)
      for (int c = 0; c < 6; ++c) sigma_mean[r * 6 + c] = s->h_mean6[r * 6 + c];
    if (q.status != 0) fprintf(stderr, "liblippmannschwinger: non-finite residual (rhs %d)\n", r);
  }
  if (verbose > 0) {
    const RhsState& q = s->h_st[0];
    const int it = q.converged ? q.checks - 1 : maxiter;
    printf(verbose == 1 ? "  LS " : "  ");
    printf(q.converged ? "converged" : "failed to converge");
    printf(" after %4d its, ||r|| = %6.3e ||r||/||r_0|| = %6.3e\n", it, q.R, q.R0 > 0 ? q.R / q.R0 : 0.0);
  }
  return rc;
}

// ---- autotuner -------------------------------------------------------------
// Candidate tiles of one axis: for each wanted unit count and each register
// budget, the largest fitting tile (duplicates dropped).
std::vector<Tile> tile_candidates(const HostAxis& a, const std::vector<int>& wants, int lines_per_unit,
                                  size_t bytes_per_unit, int max_units) {
  std::vector<Tile> out;
  for (int w : wants) {
    for (const Budget& b : kBudget32) {
      Tile t;
      if (!choose_tile(a, std::max(1, std::min(w, max_units)), lines_per_unit, bytes_per_unit, {b}, t)) continue;
      bool dup = false;
      for (const Tile& o : out) dup = dup || (o.lines == t.lines && o.E == t.E && o.NT == t.NT);
      if (!dup) out.push_back(t);
    }
  }
  return out;
}

template <typename F>
float time_launches(ls_solver* s, int reps, F&& launch) {
  // a candidate whose launch is rejected (no kernel for the tile, too much LDS)
  // is reported as -1 and skipped by the sweep
  s->launch_failed = false;
  (void)hipGetLastError();
  launch();  // warm-up
  if (s->launch_failed || hipGetLastError() != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) {
    (void)hipGetLastError();
    s->launch_failed = false;
    return -1.f;
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -2.f;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -2.f;
  }
  (void)hipEventRecord(e0, s->stream);
  for (int i = 0; i < reps; ++i) launch();
  (void)hipEventRecord(e1, s->stream);
  float ms = -2.f;
  if (hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return hipGetLastError() == hipSuccess ? ms / reps : -2.f;
}

// Time each candidate tile of the three FP32 passes on the solver's own
// buffers (zeroed spectrum, strain and d_base; every load case active) and keep
// the fastest.  The z tile is timed with the plain z-line pass; the Anderson
// z-line passes share its LDS and register layout.
int autotune_impl(ls_solver* s, const float* lam, const float* mu, float* eps, int reps, std::string& rep) {
  // a captured iteration chunk holds the launch shapes and the d_base layout
  // of the current tiles, which the sweep changes: capture again at the next
  // graph solve
  if (s->gexec) {
    LS_HIP(hipGraphExecDestroy(s->gexec));
    s->gexec = nullptr;
  }
  const bool timing = s->timing;
  s->timing = false;
  std::vector<RhsDesc> desc(s->nrhs);
  for (auto& d : desc) {
    d.mat = 0;
    for (float& e : d.eps_bar) e = 0.f;
    d.lam0 = 1.f;
    d.mu0 = 1.f;
    d.cA = (float)(-1.0 / (double)s->N);
    d.cB = (float)(2.0 / (3.0 * (double)s->N));
  }
  LS_HIP(hipMemcpyAsync(s->d_rhs, desc.data(), desc.size() * sizeof(RhsDesc), hipMemcpyHostToDevice, s->stream));
  for (int r = 0; r < s->nrhs; ++r) {
    s->h_st[r] = RhsState{};
    s->h_st[r].active = 1;
  }
  if (put_state(s)) return -1;
  LS_HIP(hipMemsetAsync(s->spec, 0, (size_t)s->nrhs * s->M * 6 * sizeof(float2), s->stream));
  LS_HIP(hipMemsetAsync(eps, 0, (size_t)s->nrhs * 6 * s->N * sizeof(float), s->stream));
  LS_HIP(hipMemsetAsync(s->dbase, 0, (size_t)s->nrhs * s->dl.per_rhs * sizeof(float2), s->stream));
  LS_HIP(hipMemsetAsync(s->dc, 0, (size_t)s->nrhs * 6 * sizeof(float), s->stream));
  const int lsz = s->gd.lsz;
  const bool verbose_tune = std::getenv("LS_TUNE_VERBOSE") != nullptr;
  auto sweep = [&](const char* name, Tile& cur, const std::vector<Tile>& cands, auto&& launch) -> int {
    Tile best = cur;
    float best_ms = -1.f;
    char buf[160];
    for (const Tile& t : cands) {
      cur = t;
      apply_tiles(s);
      const float ms = time_launches(s, reps, launch);
      if (ms == -1.f) {
        if (verbose_tune) fprintf(stderr, "  tune %s: candidate rejected at launch\n", name);
        continue;
      }
      if (ms < 0.f) return -1;
      if (verbose_tune)
        fprintf(stderr, "  tune %s: %s%s lines %d E %d NT %d lds %zu: %.4f ms\n", name,
                t.reg == 2 ? "three-line" : (t.reg ? "register" : "lds"), t.inpl ? " in place" : "", t.lines, t.E,
                t.NT, t.lds, ms);
      if (best_ms < 0.f || ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
    cur = best;
    apply_tiles(s);
    snprintf(buf, sizeof(buf), "%s: %s %d lines E %d NT %d %.3f ms (%zu candidates); ", name,
             best.reg == 2 ? "three-line" : (best.reg ? "register" : "lds"), best.lines, best.E, best.NT, best_ms,
             cands.size());
    rep += buf;
    return 0;
  };
  const int wz = std::max(1, 3072 / (3 * s->nz));
  const int wy = std::max(1, std::min(6 * s->P, 4096 / s->ny));
  const int wx = std::max(1, 3072 / (6 * s->nx));
  auto wants = [](int w) { return std::vector<int>{std::max(1, w / 4), std::max(1, w / 2), w, 2 * w, 4 * w}; };
  std::vector<Tile> cy = tile_candidates(s->ax[1], wants(wy), 1, (size_t)s->ny * sizeof(float2), 6 * s->P);
  std::vector<Tile> cx = tile_candidates(s->ax[0], wants(wx), 6, (size_t)6 * s->nx * sizeof(float2), kMaxTm);
  for (const Tile& t : reg_tiles(kRegY, sizeof(kRegY) / sizeof(kRegY[0]), s->ny, 1)) cy.push_back(t);
  for (const Tile& t : reg_tiles(kRegH, sizeof(kRegH) / sizeof(kRegH[0]), s->nx, 2)) cx.push_back(t);
  if (sweep("z", s->tz, tile_candidates(s->ax[2], wants(wz), 3, (size_t)3 * lsz * sizeof(float2), 1 << 20),
            [&] { zline<Z_UPD>(s, lam, mu, eps); }) ||
      sweep("y", s->ty, cy, [&] { ypass(s, true); }) || sweep("x", s->tx, cx, [&] { xpass(s, 1); }))
    return -1;
  s->timing = timing;
  LS_HIP(hipStreamSynchronize(s->stream));
  return 0;
}

// ---- no-allocation solves -------------------------------------------------
// Device memory of ls_solve_noalloc: one array per GPU, held from the
// library's load until the process ends.  Host code never passes `scratch`
// itself (in host code the name is a stand-in, not a device address); it uses
// the address hipGetSymbolAddress returns on each device.
constexpr size_t kScratchBytes = LS_NOALLOC_SCRATCH_BYTES;
constexpr int kScratchDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per device: the array's address and the solver last built in it (kept while
// the grid, extents, load-case count and depth stay the same, so repeated
// calls allocate nothing, not even host memory).
struct ScratchSlot {
  unsigned char* ptr = nullptr;
  ls_solver* s = nullptr;
  int vox[3] = {0, 0, 0};
  float ext[3] = {0.f, 0.f, 0.f};
  int nrhs = 0, depth = -1;
};
ScratchSlot g_scratch[kScratchDevices];
std::mutex g_scratch_mu;  // serialises the calls of all threads of the process

int solve_noalloc_impl(const int* voxels, const float* extents, int nrhs, int depth, const float* d_lambda,
                       const float* d_mu, const float* eps_bar, float* d_eps, float* d_sigma, float rtol,
                       float atol, int maxiter, int* iters, double* residual, double* sigma_mean, void* stream) {
  if (!voxels || !extents || !eps_bar || !d_lambda || !d_mu || !d_eps) {
    fprintf(stderr, "liblippmannschwinger: ls_solve_noalloc: null argument\n");
    return -1;
  }
  int dev = 0;
  LS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) {
    fprintf(stderr, "liblippmannschwinger: ls_solve_noalloc: device %d beyond %d\n", dev, kScratchDevices);
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  ScratchSlot& slot = g_scratch[dev];
  if (!slot.ptr) {
    // once per device, with that device current
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "liblippmannschwinger: ls_solve_noalloc: hipGetSymbolAddress(scratch) failed: %s\n",
              hipGetErrorName(e));
      return -1;
    }
    slot.ptr = static_cast<unsigned char*>(p);
  }
  const bool same = slot.s && slot.nrhs == nrhs && slot.depth == depth &&
                    std::equal(voxels, voxels + 3, slot.vox) && std::equal(extents, extents + 3, slot.ext);
  if (!same) {
    free_solver(slot.s);  // host side only: its device buffers lie in the array
    slot.s = nullptr;
    ls_solver* s = new ls_solver();
    s->arena_mode = true;
    Arena arena;
    arena.base = slot.ptr;
    arena.cap = kScratchBytes;
    t_arena = &arena;
    const int rc = create_impl(s, voxels, extents, nrhs, 1, depth, stream);
    t_arena = nullptr;
    if (rc || arena.overflow) {
      if (!rc)
        fprintf(stderr,
                "liblippmannschwinger: ls_solve_noalloc: grid %d x %d x %d with %d load cases and depth %d needs "
                "%zu bytes of scratch, the library holds %zu\n",
                voxels[0], voxels[1], voxels[2], nrhs, depth, arena.off, kScratchBytes);
      free_solver(s);
      return -1;
    }
    slot.s = s;
    std::copy(voxels, voxels + 3, slot.vox);
    std::copy(extents, extents + 3, slot.ext);
    slot.nrhs = nrhs;
    slot.depth = depth;
  }
  slot.s->stream = (hipStream_t)stream;
  return solve_impl(slot.s, d_lambda, d_mu, nullptr, eps_bar, nullptr, d_eps, d_sigma, 0, rtol, atol, maxiter, 0.0, 0,
                    iters, residual, sigma_mean, false, 0);
}

}  // namespace

// ============================================================================
// C ABI
extern "C" {

int ls_solve_noalloc(const int* voxels, const float* extents, int nrhs, int depth, const float* d_lambda,
                     const float* d_mu, const float* eps_bar, float* d_eps, float* d_sigma, float rtol, float atol,
                     int maxiter, int* iters, double* residual, double* sigma_mean, void* stream) {
  return solve_noalloc_impl(voxels, extents, nrhs, depth, d_lambda, d_mu, eps_bar, d_eps, d_sigma, rtol, atol,
                            maxiter, iters, residual, sigma_mean, stream);
}

ls_solver* ls_solver_create(const int* voxels, const float* extents, int nrhs, int nmat, int depth, void* stream) {
  if (!voxels || !extents) return nullptr;
  ls_solver* s = new ls_solver();
  if (create_impl(s, voxels, extents, nrhs, nmat, depth, stream)) {
    free_solver(s);
    return nullptr;
  }
  return s;
}

void ls_solver_destroy(ls_solver* s) { free_solver(s); }

int ls_solver_phase_field(ls_solver* s, const double* d_H, const double* d_dold, double gc, double lc, double tol,
                          int maxiter, double* d_out, int* iters, double* residual, const double* extents) {
  const double ext[3] = {!s ? 0.0 : extents ? extents[0] : (double)s->Lx,
                         !s ? 0.0 : extents ? extents[1] : (double)s->Ly,
                         !s ? 0.0 : extents ? extents[2] : (double)s->Lz};
  if (!s || !d_H || !d_dold || !d_out || !(gc > 0.0) || !(lc > 0.0) || maxiter < 0 ||
      !(ext[0] > 0.0 && ext[1] > 0.0 && ext[2] > 0.0)) {
    fprintf(stderr, "liblippmannschwinger: ls_solver_phase_field: invalid arguments\n");
    return -1;
  }
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  if (nx > kPfItems || ny > kPfItems || nz > kPfItems) {
    fprintf(stderr, "liblippmannschwinger: ls_solver_phase_field: axis longer than %d\n", kPfItems);
    return -1;
  }
  const long long N = s->N;
  hipStream_t st = s->stream;
  // tiles of the five passes (lines x length <= kPfItems per block; on small
  // grids no more lines than leave about one block per CU)
  // (the z tile is bounded by its padded LDS stride, so L lines take at most
  // kPfItems complex numbers of LDS, like the y and x tiles)
  const int L = std::max(1, std::min({1024, kPfItems / s->gd.lsz, nx * ny / s->ncu}));
  const int nblk_z = (nx * ny + L - 1) / L;
  const size_t lds_z = (size_t)L * s->gd.lsz * sizeof(double2);
  const int Ty = std::max(1, std::min({nz, kPfItems / ny, nx * nz / s->ncu}));
  const int nch_y = (nz + Ty - 1) / Ty;
  const size_t lds_y = (size_t)ny * Ty * sizeof(double2);
  const long long Q = (long long)ny * nz;
  const int Tx = (int)std::max(1LL, std::min({Q, (long long)(kPfItems / nx), Q / s->ncu}));
  const long long nblk_x = (Q + Tx - 1) / Tx;
  const size_t lds_x = (size_t)nx * Tx * sizeof(double2);
  {
    int dev = 0, lds_max = 0;
    LS_HIP(hipGetDevice(&dev));
    LS_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (std::max({lds_z, lds_y, lds_x}) > (size_t)lds_max) {
      fprintf(stderr, "liblippmannschwinger: ls_solver_phase_field: tile needs %zu bytes of LDS, the device has %d\n",
              std::max({lds_z, lds_y, lds_x}), lds_max);
      return -1;
    }
  }
  const int nb_mm = (int)std::min<long long>(1024, (N + kThreads - 1) / kThreads);
  // FP64 symbol tables: sin / cos (pi k / n) with the exact zeros, 2 n / L
  std::vector<double> tab(2 * (size_t)(nx + ny + nz));
  const int ns[3] = {nx, ny, nz};
  size_t off[3];
  {
    size_t o = 0;
    for (int a = 0; a < 3; ++a) {
      off[a] = o;
      const int n = ns[a];
      for (int k = 0; k < n; ++k) {
        tab[o + k] = k == 0 ? 0.0 : std::sin(M_PI * k / n);
        tab[o + n + k] = (n % 2 == 0 && 2 * k == n) ? 0.0 : std::cos(M_PI * k / n);
      }
      o += 2 * (size_t)n;
    }
  }
  double2* C = nullptr;
  double *chi = nullptr, *part = nullptr, *sums = nullptr, *dtab = nullptr;
  int rc = -1;
  do {
    if (dalloc(&C, (size_t)N) || dalloc(&chi, (size_t)N) ||
        dalloc(&part, (size_t)2 * std::max<long long>(nblk_z, nb_mm)) || dalloc(&sums, 4) ||
        dalloc(&dtab, tab.size()))
      break;
    if (hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) break;
    PfTables tb;
    tb.sx = dtab + off[0];
    tb.cx = dtab + off[0] + nx;
    tb.sy = dtab + off[1];
    tb.cy = dtab + off[1] + ny;
    tb.sz = dtab + off[2];
    tb.cz = dtab + off[2] + nz;
    tb.gx = 2.0 * nx / ext[0];
    tb.gy = 2.0 * ny / ext[1];
    tb.gz = 2.0 * nz / ext[2];
    // A0 from min / max of H (A increases with H)
    k_pf_minmax<<<nb_mm, kThreads, 0, st>>>(d_H, N, part);
    std::vector<double> mm(2 * (size_t)nb_mm);
    if (hipMemcpyAsync(mm.data(), part, mm.size() * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      break;
    double hmin = INFINITY, hmax = -INFINITY;
    for (int b = 0; b < nb_mm; ++b) {
      hmin = std::min(hmin, mm[2 * b]);
      hmax = std::max(hmax, mm[2 * b + 1]);
    }
    const double gclc = gc * lc, inv_lc2 = 1.0 / (lc * lc);
    const double A0 = 0.5 * ((inv_lc2 + 2.0 * hmin / gclc) + (inv_lc2 + 2.0 * hmax / gclc));
    k_pf_init<<<nb_mm, kThreads, 0, st>>>(N, d_H, d_dold, inv_lc2, gclc, A0, chi, d_out);
    const FastDiv fy = make_fastdiv(Ty), fx = make_fastdiv(Tx);
    // state[0] stop flag, state[1] iterations (k_pf_check); sums[2] the residual
    int* state = reinterpret_cast<int*>(sums + 3);
    auto one_iteration = [&]() {
      k_pf_z_fwd<<<nblk_z, kPfNT, lds_z, st>>>(s->gd, state, chi, C, L);
      k_pf_y<false><<<nx * nch_y, kPfNT, lds_y, st>>>(s->gd, state, C, fy, nch_y);
      k_pf_x<<<(unsigned)nblk_x, kPfNT, lds_x, st>>>(s->gd, state, tb, C, fx, A0);
      k_pf_y<true><<<nx * nch_y, kPfNT, lds_y, st>>>(s->gd, state, C, fy, nch_y);
      k_pf_z_inv<<<nblk_z, kPfNT, lds_z, st>>>(s->gd, state, C, d_H, inv_lc2, gclc, A0, chi, d_out, part, L);
      k_pf_sum<<<1, kThreads, 0, st>>>(state, part, nblk_z, sums);
      k_pf_check<<<1, 64, 0, st>>>(sums, tol, maxiter, state, sums + 2);
    };
    // a chunk of iterations, each followed by its stop test on the device, as
    // one captured graph (kernels only); the host reads the state once per
    // replay.  The loop runs while res > tol and it < maxiter, res = 1 at the
    // start (lippmann_schwinger.py:540-545).
    const int chunk = std::max(1, env_int("LS_PF_CHUNK", 8));
    double res = 1.0;
    int it = 0;
    bool ok = hipGetLastError() == hipSuccess;
    if (ok && res > tol && maxiter > 0) {
      ok = hipMemsetAsync(state, 0, 2 * sizeof(int), st) == hipSuccess;
      hipGraph_t graph = nullptr;
      hipGraphExec_t gexec = nullptr;
      ok = ok && hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
      if (ok) {
        for (int k = 0; k < chunk; ++k) one_iteration();
        ok = hipStreamEndCapture(st, &graph) == hipSuccess &&
             hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) == hipSuccess;
      }
      if (graph) (void)hipGraphDestroy(graph);
      while (ok) {
        // residual (sums[2]) and the state (sums[3]) in one copy
        double hb[2];
        ok = hipGraphLaunch(gexec, st) == hipSuccess &&
             hipMemcpyAsync(hb, sums + 2, sizeof(hb), hipMemcpyDeviceToHost, st) == hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
        if (!ok) break;
        int h[2];
        std::memcpy(h, &hb[1], sizeof(h));
        res = hb[0];
        it = h[1];
        if (h[0]) break;
      }
      if (gexec) (void)hipGraphExecDestroy(gexec);
    }
    if (!ok) break;
    if (iters) *iters = it;
    if (residual) *residual = res;
    rc = 0;
  } while (0);
  if (C) (void)hipFree(C);
  if (chi) (void)hipFree(chi);
  if (part) (void)hipFree(part);
  if (sums) (void)hipFree(sums);
  if (dtab) (void)hipFree(dtab);
  if (rc != 0) fprintf(stderr, "liblippmannschwinger: ls_solver_phase_field failed\n");
  return rc;
}

int ls_solver_set_damage(ls_solver* s, const float* d_damage, float k) {
  if (!s) return -1;
  s->gd.dmg = d_damage;
  s->gd.kdmg = k;
  if (s->gexec) {
    // a captured iteration bakes in the material law
    (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
  }
  return 0;
}

int ls_solver_solve(ls_solver* s, const float* d_lambda, const float* d_mu, const int* rhs_mat,
                    const float* eps_bar, const float* d_tau, float* d_eps, float* d_sigma, int warm_start,
                    float rtol, float atol, int maxiter, double rel_ref, int flags, int verbose, int* iters,
                    double* residual, double* sigma_mean) {
  if (!s || !d_lambda || !d_mu || !eps_bar || !d_eps) return -1;
  return solve_impl(s, d_lambda, d_mu, rhs_mat, eps_bar, d_tau, d_eps, d_sigma, warm_start, rtol, atol, maxiter,
                    rel_ref, verbose, iters, residual, sigma_mean, false, flags);
}

int ls_solver_autotune(ls_solver* s, const float* d_lambda, const float* d_mu, float* d_eps, int reps, char* report,
                       int report_cap) {
  if (!s || !d_lambda || !d_mu || !d_eps || reps < 1) return -1;
  std::string rep;
  const int rc = autotune_impl(s, d_lambda, d_mu, d_eps, reps, rep);
  if (report && report_cap > 0) snprintf(report, (size_t)report_cap, "%s", rep.c_str());
  return rc;
}

int ls_solver_last_timing(ls_solver* s, double* ms, int* count, int nkinds) {
  if (!s) return -1;
  for (int k = 0; k < nkinds && k < ls_solver::kKinds; ++k) {
    if (ms) ms[k] = s->t_ms[k];
    if (count) count[k] = s->t_cnt[k];
  }
  return s->n_rebase;
}

int lippmann_schwinger_solve(float* lambda, float* mu, float* epsilon_bar, float* epsilon, float* sigma,
                             int* voxels, float* extents, float rtol, float atol, int maxiter, int verbose) {
  if (!lambda || !mu || !epsilon_bar || !epsilon || !sigma || !voxels || !extents) return -1;
  ls_solver* s = ls_solver_create(voxels, extents, 1, 1, 0, nullptr);
  if (!s) return -1;
  const size_t N = (size_t)s->N;
  float *d_lam = nullptr, *d_mu = nullptr, *d_eps = nullptr, *d_sig = nullptr;
  int rc = -1;
  int it = -1;
  do {
    if (hipMalloc(&d_lam, N * sizeof(float)) != hipSuccess || hipMalloc(&d_mu, N * sizeof(float)) != hipSuccess ||
        hipMalloc(&d_eps, 6 * N * sizeof(float)) != hipSuccess ||
        hipMalloc(&d_sig, 6 * N * sizeof(float)) != hipSuccess) {
      fprintf(stderr, "liblippmannschwinger: device allocation failed\n");
      break;
    }
    if (hipMemcpy(d_lam, lambda, N * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_mu, mu, N * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      break;
    int iters = 0;
    rc = solve_impl(s, d_lam, d_mu, nullptr, epsilon_bar, nullptr, d_eps, d_sig, 0, rtol, atol, maxiter, 0.0,
                    verbose, &iters, nullptr, nullptr, true);
    if (rc < 0) break;
    if (hipMemcpy(epsilon, d_eps, 6 * N * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(sigma, d_sig, 6 * N * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = -1;
      break;
    }
    it = iters;
  } while (0);
  (void)hipFree(d_lam);
  (void)hipFree(d_mu);
  (void)hipFree(d_eps);
  (void)hipFree(d_sig);
  ls_solver_destroy(s);
  return rc < 0 ? -1 : it;
}

int lippmann_schwinger_effective_tensor(const float* lambda, const float* mu, const int* voxels,
                                        const float* extents, float rtol, float atol, int maxiter, int depth,
                                        int verbose, double* ceff, int* iters) {
  if (!lambda || !mu || !voxels || !extents || !ceff) return -1;
  ls_solver* s = ls_solver_create(voxels, extents, 6, 1, depth, nullptr);
  if (!s) return -1;
  const size_t N = (size_t)s->N;
  float *d_lam = nullptr, *d_mu = nullptr, *d_eps = nullptr;
  int rc = -1;
  do {
    if (hipMalloc(&d_lam, N * sizeof(float)) != hipSuccess || hipMalloc(&d_mu, N * sizeof(float)) != hipSuccess ||
        hipMalloc(&d_eps, 36 * N * sizeof(float)) != hipSuccess) {
      fprintf(stderr, "liblippmannschwinger: device allocation failed\n");
      break;
    }
    if (hipMemcpy(d_lam, lambda, N * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_mu, mu, N * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      break;
    float eb[36] = {0};
    for (int k = 0; k < 6; ++k) eb[k * 6 + k] = 1.f;
    int mats[6] = {0, 0, 0, 0, 0, 0};
    rc = solve_impl(s, d_lam, d_mu, mats, eb, nullptr, d_eps, nullptr, 0, rtol, atol, maxiter, 0.0, verbose, iters,
                    nullptr, ceff, false);
  } while (0);
  (void)hipFree(d_lam);
  (void)hipFree(d_mu);
  (void)hipFree(d_eps);
  ls_solver_destroy(s);
  return rc;
}

}  // extern "C"
