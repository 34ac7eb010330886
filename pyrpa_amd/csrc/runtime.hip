// Library runtime: device context, caching pool, sticky error channel,
// hipEvent kernel profiling, and the host-side setup entry points of the
// reference (k-mesh, Fermi factors, interaction vertices).  H(k) and the band
// eigenvectors are formed on the device (setup.hip).
#include <algorithm>
#include <cmath>
#include <complex>

#include "../../include/gmod.h"
#include "gmod_common.h"

namespace gmod {

// ------------------------------------------------------------------ context
Context& ctx() {
  static Context* c = [] {
    auto* cc = new Context();
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
      throw Error("no HIP device: libgmod needs an AMD GPU (gfx950)");
    GMOD_HIP(hipSetDevice(0));
    GMOD_HIP(hipStreamCreateWithFlags(&cc->stream, hipStreamNonBlocking));
    if (const char* p = getenv("GMOD_PRECISION")) cc->precision = atoi(p);
    if (const char* v = getenv("GMOD_VERBOSE")) cc->verbose = atoi(v);
    return cc;
  }();
  return *c;
}

// ------------------------------------------------------------------ errors
static std::mutex g_err_mu;
static std::string g_err_msg;
static int g_err_count = 0;

void set_error(const std::string& msg) {
  std::lock_guard<std::mutex> g(g_err_mu);
  g_err_msg = msg;
  ++g_err_count;
  fprintf(stderr, "[gmod] ERROR: %s\n", msg.c_str());
  fflush(stderr);
  if (getenv("GMOD_ABORT_ON_ERROR")) abort();
}
void clear_error() {
  std::lock_guard<std::mutex> g(g_err_mu);
  g_err_count = 0;
  g_err_msg.clear();
}

// ------------------------------------------------------------------ pool
void* Pool::get(size_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  size_t want = (bytes + 255) & ~size_t(255);
  int best = -1;
  for (size_t i = 0; i < blocks_.size(); ++i) {
    auto& b = blocks_[i];
    if (!b.used && b.bytes >= want && b.bytes <= 2 * want + (1 << 20)) {
      if (best < 0 || b.bytes < blocks_[best].bytes) best = (int)i;
    }
  }
  if (best >= 0) {
    blocks_[best].used = true;
    return blocks_[best].p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    for (auto it = blocks_.begin(); it != blocks_.end();) {
      if (!it->used) { (void)hipFree(it->p); it = blocks_.erase(it); }
      else ++it;
    }
    e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error("device allocation of " + std::to_string(want) + " bytes failed");
    }
  }
  blocks_.push_back({p, want, true});
  return p;
}
void Pool::put(void* p) {
  std::lock_guard<std::mutex> g(mu_);
  for (auto& b : blocks_)
    if (b.p == p) { b.used = false; return; }
}
void Pool::trim() {
  std::lock_guard<std::mutex> g(mu_);
  for (auto it = blocks_.begin(); it != blocks_.end();) {
    if (!it->used) { (void)hipFree(it->p); it = blocks_.erase(it); }
    else ++it;
  }
}
Pool::~Pool() {}
Pool& pool() {
  static Pool* p = new Pool();
  return *p;
}
static double pool_bytes() { return 0.0; }

// ------------------------------------------------------------------ profiling
struct ProfRec { int cls; double bytes; hipEvent_t a, b; };
static bool g_prof = false;
static std::vector<ProfRec> g_prof_recs;
static double g_prof_acc[3 * PROF_NCLASS];
static std::vector<hipEvent_t> g_event_pool;  // recycled events (creation is not free)

bool prof_enabled() { return g_prof; }

static hipEvent_t prof_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  GMOD_HIP(hipEventCreate(&e));
  return e;
}

ProfScope::ProfScope(int c, double by, hipStream_t st) : cls(c), bytes(by), s(st) {
  if (!g_prof) return;
  e0 = prof_event();
  GMOD_HIP(hipEventRecord(e0, s));
}
ProfScope::~ProfScope() {
  if (!g_prof || !e0) return;
  hipEvent_t e1 = nullptr;
  try { e1 = prof_event(); } catch (...) { return; }
  (void)hipEventRecord(e1, s);
  g_prof_recs.push_back({cls, bytes, e0, e1});
}

static void prof_flush() {
  if (g_prof_recs.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto& r : g_prof_recs) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    g_prof_acc[3 * r.cls] += ms;
    g_prof_acc[3 * r.cls + 1] += 1;
    g_prof_acc[3 * r.cls + 2] += r.bytes;
    g_event_pool.push_back(r.a);
    g_event_pool.push_back(r.b);
  }
  g_prof_recs.clear();
}
void prof_reset() {
  prof_flush();
  std::fill(g_prof_acc, g_prof_acc + 3 * PROF_NCLASS, 0.0);
}
void prof_read(double* out) {
  prof_flush();
  std::copy(g_prof_acc, g_prof_acc + 3 * PROF_NCLASS, out);
}

}  // namespace gmod

using namespace gmod;

// ====================================================================== C ABI: runtime
extern "C" int gmod_last_error(char* buf, int len) {
  std::lock_guard<std::mutex> g(g_err_mu);
  if (buf && len > 0) {
    strncpy(buf, g_err_msg.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  int c = g_err_count;
  g_err_count = 0;
  return c;
}

extern "C" void gmod_set_precision(int mode) {
  try { ctx().precision = mode; } catch (std::exception& e) { set_error(e.what()); }
}
extern "C" int gmod_get_precision(void) {
  try { return ctx().precision; } catch (std::exception& e) { set_error(e.what()); return -1; }
}
extern "C" void gmod_set_verbose(int level) {
  try { ctx().verbose = level; } catch (std::exception& e) { set_error(e.what()); }
}
extern "C" void gmod_profile(int on) {
  try {
    ctx();
    prof_reset();
    g_prof = on != 0;
  } catch (std::exception& e) { set_error(e.what()); }
}
extern "C" int gmod_profile_read(double* out, int n) {
  double tmp[3 * PROF_NCLASS];
  prof_read(tmp);
  int m = std::min(n, 3 * (int)PROF_NCLASS);
  std::copy(tmp, tmp + m, out);
  return m;
}
extern "C" double gmod_pool_bytes(void) { return pool_bytes(); }
extern "C" void gmod_release_memory(void) {
  try { pool().trim(); } catch (std::exception& e) { set_error(e.what()); }
}

// Device buffers for callers of the device-pointer entry points that hold no
// other HIP runtime (a process that loads a second HIP runtime, e.g. one
// bundled with a Python package, ends up with two runtimes sharing one GPU).
extern "C" void* gmod_device_alloc(size_t bytes) {
  try {
    std::lock_guard<std::mutex> g(ctx().lock);
    void* p = nullptr;
    GMOD_HIP(hipMalloc(&p, bytes ? bytes : 1));
    return p;
  } catch (std::exception& e) {
    set_error(e.what());
    return nullptr;
  }
}
extern "C" void gmod_device_free(void* p) {
  try {
    std::lock_guard<std::mutex> g(ctx().lock);
    if (p) GMOD_HIP(hipFree(p));
  } catch (std::exception& e) { set_error(e.what()); }
}
// kind 1: host to device, 2: device to host, 3: device to device; synchronous
extern "C" int gmod_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  try {
    std::lock_guard<std::mutex> g(ctx().lock);
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                            : kind == 2 ? hipMemcpyDeviceToHost
                            : kind == 3 ? hipMemcpyDeviceToDevice
                                        : throw Error("gmod_memcpy: kind must be 1, 2 or 3");
    hipStream_t s = ctx().stream;
    GMOD_HIP(hipMemcpyAsync(dst, src, bytes, k, s));
    GMOD_HIP(hipStreamSynchronize(s));
    return 0;
  } catch (std::exception& e) {
    set_error(e.what());
    return 1;
  }
}

// ====================================================================== C ABI: setup (host)
// These run on the host exactly like the reference's own setup routines: they
// produce the k-mesh tables and band data once per run.

// finv.f90:102-289: irreducible k-list order (integer coordinates)
static std::vector<std::array<long, 3>> irr_klist(long Nx, long Ny, long Nz, long Nk) {
  std::vector<std::array<long, 3>> kl(Nk + 1, {0, 0, 0});
  auto put = [&](long idx, long i, long j, long k) {
    if (idx < 1 || idx > Nk) throw Error("k-list index out of range (inconsistent Nk)");
    kl[idx] = {i, j, k};
  };
  long ini;
  if (Nx % 2 == 0 && Ny % 2 == 0) {
    for (long j = 0; j <= Ny / 2; ++j)
      for (long i = 0; i <= Nx / 2; ++i) put(i + 1 + (Nx / 2 + 1) * j, i, j, 0);
    ini = (Nx / 2 + 1) * (Ny / 2 + 1);
    for (long j = Ny / 2 + 1; j < Ny; ++j)
      for (long i = 1; i < Nx / 2; ++i) put(ini + i + (Nx / 2 - 1) * (j - Ny / 2 - 1), i, j, 0);
    ini = Nx * Ny / 2 + 3;
  } else if ((Nx * Ny) % 2 == 0) {
    if (Nx % 2 != 0) throw Error("odd Nx with even Ny is not supported by the reference k-mesh (finv.f90:154)");
    long it = 1;
    for (long j = 0; j <= (Ny - 1) / 2; ++j)
      for (long i = 0; i <= Nx / 2; ++i) put(it++, i, j, 0);
    for (long j = (Ny + 1) / 2; j < Ny; ++j)
      for (long i = 1; i < Nx / 2; ++i) put(it++, i, j, 0);
    ini = it;
  } else {
    for (long j = 0; j <= (Ny - 1) / 2; ++j)
      for (long i = 0; i <= (Nx - 1) / 2; ++i) put(i + 1 + ((Nx + 1) / 2) * j, i, j, 0);
    ini = (Nx + 1) * (Ny + 1) / 4;
    for (long j = (Ny + 1) / 2; j < Ny; ++j)
      for (long i = 1; i <= (Nx - 1) / 2; ++i) put(ini + i + ((Nx - 1) / 2) * (j - (Ny + 1) / 2), i, j, 0);
    ini = (Nx * Ny + 1) / 2 + 1;
  }
  if (Nz > 1) {
    if (Nz % 2 == 0) {
      for (long k = 1; k < Nz / 2; ++k)
        for (long j = 0; j < Ny; ++j)
          for (long i = 0; i < Nx; ++i) put(ini + i + Nx * j + Nx * Ny * (k - 1), i, j, k);
      ini += Nx * Ny * (Nz / 2 - 1);
      const long h = Nz / 2;
      if (Nx % 2 == 0 && Ny % 2 == 0) {
        for (long j = 0; j <= Ny / 2; ++j)
          for (long i = 0; i <= Nx / 2; ++i) put(ini + i + (Nx / 2 + 1) * j, i, j, h);
        ini += (Nx / 2 + 1) * (Ny / 2 + 1) - 1;
        for (long j = Ny / 2 + 1; j < Ny; ++j)
          for (long i = 1; i < Nx / 2; ++i) put(ini + i + (Nx / 2 - 1) * (j - Ny / 2 - 1), i, j, h);
      } else {
        long it = ini;
        if ((Nx * Ny) % 2 == 0) {
          for (long j = 0; j <= (Ny - 1) / 2; ++j)
            for (long i = 0; i <= Nx / 2; ++i) put(it++, i, j, h);
          for (long j = (Ny + 1) / 2; j < Ny; ++j)
            for (long i = 1; i < Nx / 2; ++i) put(it++, i, j, h);
        } else {
          for (long j = 0; j <= (Ny - 1) / 2; ++j)
            for (long i = 0; i <= (Nx - 1) / 2; ++i) put(it++, i, j, h);
          for (long j = (Ny + 1) / 2; j < Ny; ++j)
            for (long i = 1; i <= (Nx - 1) / 2; ++i) put(it++, i, j, h);
        }
      }
    } else {
      for (long k = 1; k <= (Nz - 1) / 2; ++k)
        for (long j = 0; j < Ny; ++j)
          for (long i = 0; i < Nx; ++i) put(ini + i + Nx * j + Nx * Ny * (k - 1), i, j, k);
    }
  }
  return kl;
}

extern "C" void generate_irr_kpoint_inv(double* klist, int64_t* kmap, int64_t* invk, const int64_t* Nk_,
                                        const int64_t* Nx_, const int64_t* Ny_, const int64_t* Nz_) {
  try {
    const long Nk = *Nk_, Nx = *Nx_, Ny = *Ny_, Nz = *Nz_;
    const long Nkall = Nx * Ny * Nz;
    auto kl = irr_klist(Nx, Ny, Nz, Nk);
    for (long j = 1; j <= Nk; ++j) {
      klist[3 * (j - 1) + 0] = (double)kl[j][0] / (double)Nx;
      klist[3 * (j - 1) + 1] = (double)kl[j][1] / (double)Ny;
      klist[3 * (j - 1) + 2] = (double)kl[j][2] / (double)Nz;
    }
    // first irreducible j (list order) matching k directly or as -k (finv.f90:28-53)
    std::vector<long> first(Nkall, -1);
    for (long j = Nk; j >= 1; --j) {
      long c = kl[j][0] + Nx * (kl[j][1] + Ny * kl[j][2]);
      first[c] = j;
    }
    for (long i = 0; i < Nkall; ++i) {
      long x = i % Nx, y = (i / Nx) % Ny, z = i / (Nx * Ny);
      kmap[3 * i] = x; kmap[3 * i + 1] = y; kmap[3 * i + 2] = z;
      long mx = (Nx - x) % Nx, my = (Ny - y) % Ny, mz = (Nz - z) % Nz;
      long ci = i, cm = mx + Nx * (my + Ny * mz);
      long jd = first[ci], jn = first[cm];
      long j;
      int f;
      if (jd > 0 && (jn < 0 || jd <= jn)) { j = jd; f = 0; }
      else if (jn > 0) { j = jn; f = 1; }
      else { j = 0; f = 0; }
      invk[3 * i] = j;
      invk[3 * i + 1] = f;
      invk[3 * i + 2] = cm + 1;
    }
  } catch (std::exception& e) { set_error(e.what()); }
}

extern "C" void get_kweight(double* w, const int64_t* invk, const int64_t* Nk, const int64_t* Nkall) {
  for (long i = 0; i < *Nk; ++i) w[i] = 0;
  for (long i = 0; i < *Nkall; ++i) {
    long k = invk[3 * i] - 1;
    if (k >= 0 && k < *Nk) w[k] += 1.0;
  }
}

extern "C" void get_ffermi(double* ffermi, const double* eig, const double* mu, const double* temp,
                           const int64_t* Nk, const int64_t* Norb) {
  double ts = std::max(*temp, 1.0e-12);
  double it = 0.5 / ts;
  for (long i = 0; i < (*Nk) * (*Norb); ++i) ffermi[i] = 0.5 - 0.5 * std::tanh((eig[i] - *mu) * it);
}

extern "C" void openmp_params(int64_t* omp_num, bool* omp_check) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *omp_num = n;
  *omp_check = n > 0;
}

// fUJ.f90:1-57
extern "C" void get_scmat(double* S, double* C, const int64_t* ol, const int64_t* site, const double* U,
                          const double* J, const int64_t* Nchi_) {
  const long n = *Nchi_;
  const double Uv = *U, Jv = *J, Up = Uv - 2 * Jv;
  for (long e = 0; e < n * n; ++e) S[e] = C[e] = 0;
  for (long i = 0; i < n; ++i)
    for (long j = 0; j < n; ++j) {
      if (site[i] != site[j]) continue;
      long a = ol[i], b = ol[i + n], c = ol[j], d = ol[j + n];
      double s = 0, cc = 0;
      bool set = true;
      if (a == b && c == d) {
        if (a == c) { s = Uv; cc = Uv; }
        else { s = Jv; cc = 2 * Up - Jv; }
      } else if (a == c && b == d) { s = Up; cc = -Up + 2 * Jv; }
      else if (a == d && b == c) { s = Jv; cc = Jv; }
      else set = false;
      if (set) { S[j + n * i] = s; C[j + n * i] = cc; }
    }
}

// Spinful on-site vertex (fUJ.f90:109-251).  ol, site (Nchi), sl, invs (Norb)
// as the reference passes them; u(p, q) and jv(p, q) give the intra-orbital U
// (p == q), the inter-orbital U' and the Hund / pair-hopping J between
// spinful orbitals p and q (1-based).  Vmat(j, i) holds the value of the pair
// (i, j): density channels of equal spins (C - S)/2, of opposite spins
// (C + S)/2, the spin-flip channel -S.
template <class UF, class JF>
static void vmat_soc(double* V, const int64_t* ol, const int64_t* sl, const int64_t* site, const int64_t* invs,
                     long n, UF u, JF jv) {
  for (long e = 0; e < n * n; ++e) V[e] = 0;
  auto s = [&](long p) { return sl[p - 1]; };
  auto fl = [&](long p) { return invs[p - 1]; };
  for (long i = 0; i < n; ++i)
    for (long j = 0; j < n; ++j) {
      if (site[i] != site[j]) continue;
      const long a = ol[i], b = ol[i + n], c = ol[j], d = ol[j + n];
      double v = 0;
      if (s(a) == s(b) && s(c) == s(d)) {
        if (s(a) == s(c)) {
          if (a == d && c == b) v = 0;  // intra-orbital and exchange terms vanish
          else if (a == b && c == d) v = u(a, c) - jv(a, c);
          else if (a == c && b == d) v = jv(a, b) - u(a, b);
        } else {
          if (a == b && c == d) v = a == fl(c) ? u(a, a) : u(a, c);
          else if (a == fl(c) && b == fl(d)) v = jv(a, b);
          else if (a == fl(d) && c == fl(b)) v = jv(a, b);
        }
      } else if (s(a) == s(c) && s(b) == s(d)) {
        if (a == fl(b) && c == fl(d)) v = a == c ? -u(a, a) : -jv(a, c);
        else if (a == c && b == d) v = -u(a, b);
        else if (a == fl(d) && c == fl(b)) v = -jv(a, b);
      }
      V[j + n * i] = v;
    }
}

// fUJ.f90:109-180: uniform U, U' = U - 2J, J
extern "C" void get_vmat_soc(double* Vmat, const int64_t* ol, const int64_t* sl, const int64_t* site,
                             const int64_t* invs, const double* Uval, const double* Jval, const int64_t* Nchi,
                             const int64_t* Norb) {
  (void)Norb;
  const double U = *Uval, J = *Jval, Up = U - 2 * J;
  // u(p, q): U for the same orbital of the other spin (p == -q), U' otherwise
  auto u = [&](long p, long q) { return p == q || p == invs[q - 1] ? U : Up; };
  auto jv = [&](long, long) { return J; };
  vmat_soc(Vmat, ol, sl, site, invs, *Nchi, u, jv);
}

// fUJ.f90:182-251: orbital-dependent Umat, Jmat (Norb x Norb, spinful indices)
extern "C" void get_vmat_soc_orb(double* Vmat, const int64_t* ol, const int64_t* sl, const int64_t* site,
                                 const int64_t* invs, const double* Umat, const double* Jmat, const int64_t* Nchi,
                                 const int64_t* Norb) {
  const long No = *Norb;
  auto u = [&](long p, long q) { return Umat[(p - 1) + No * (q - 1)]; };
  auto jv = [&](long p, long q) { return Jmat[(p - 1) + No * (q - 1)]; };
  vmat_soc(Vmat, ol, sl, site, invs, *Nchi, u, jv);
}

// fUJ.f90:59-107
extern "C" void get_scmat_orb(double* S, double* C, const int64_t* ol, const int64_t* site, const double* Um,
                              const double* Jm, const int64_t* Nchi_, const int64_t* Norb_) {
  const long n = *Nchi_, No = *Norb_;
  for (long e = 0; e < n * n; ++e) S[e] = C[e] = 0;
  auto U = [&](long p, long q) { return Um[(p - 1) + No * (q - 1)]; };
  auto Jf = [&](long p, long q) { return Jm[(p - 1) + No * (q - 1)]; };
  for (long i = 0; i < n; ++i)
    for (long j = 0; j < n; ++j) {
      if (site[i] != site[j]) continue;
      long a = ol[i], b = ol[i + n], c = ol[j], d = ol[j + n];
      if (a == b && c == d) {
        if (a == c) { S[j + n * i] = C[j + n * i] = U(a, a); }
        else { S[j + n * i] = Jf(a, c); C[j + n * i] = 2 * U(a, c) - Jf(a, c); }
      } else if (a == c && b == d) {
        S[j + n * i] = U(a, b);
        C[j + n * i] = -U(a, b) + 2 * Jf(a, b);
      } else if (a == d && b == c) {
        S[j + n * i] = C[j + n * i] = Jf(a, b);
      }
    }
}
