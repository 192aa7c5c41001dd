// Host side of the C-ABI (include/crt_hip.h): mesh plan, wavenumbers, device
// buffers, kernel launches and timing.  One HIP stream per handle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/crt_hip.h"
#include "crt_inv.h"
#include "crt_kernels.h"
#include "crt_post.h"

namespace {

constexpr int kErrArg = 1000;
constexpr int kErrHip = 1001;
constexpr int kErrOrder = 1002;
constexpr int kErrBand = 1003;
constexpr int kErrNoAllocSize = 1004;

// Scratch of crt_dmisft_dev_noalloc: the per-datum misfit values psi, for at
// most kNoAllocMaxData data (2 MiB per device, held until the process ends).
// The host never passes `scratch` itself to a kernel: crt_create resolves its
// address on each device once (hipGetSymbolAddress, the device current) into
// noalloc_ptr[device], and the entry point passes that pointer.
constexpr int kNoAllocMaxData = 1 << 18;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocMaxData];
double* noalloc_ptr[kMaxDevices] = {};
std::mutex noalloc_mu;

// Resolve `scratch` on device dev (current); 0 or kErrHip.
int noalloc_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return kErrHip;
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (noalloc_ptr[dev]) return 0;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return kErrHip;
  noalloc_ptr[dev] = static_cast<double*>(ptr);
  return 0;
}

struct Timer {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace

struct crt_handle {
  // mesh (host copies, reference layout)
  int n = 0, mb = 0, E = 0, typanz = 0, elanz = 0, relanz = 0, maxnodes = 0, nimn = 0, nel = 0;
  std::vector<int> typ, nelanz, selanz, nrel, rnr, snr, enr;
  std::vector<double> sx, sy;
  int lsytop = 0, lsink = 0, nsink = 0;
  bool has11 = false, has13 = false;
  // prescribed boundary values (randb2): nodes (1-based) and complex values
  std::vector<int> rwdnr;
  std::vector<double2> rwd;
  double sytop = 0.0;
  // wavenumbers
  int K = 0, swrtr = 1;
  std::vector<double> kwn, kwnwi;
  // stage state
  bool precal_done = false, forward_done = false, volt_done = false, sens_done = false;
  bool dc = false;  // the last forward was the DC twin (crt_forward_dc)
  bool kpt_valid = false;  // "kpt" holds the last forward's kpot transposed
  // wavenumbers per pass of the recomputing mode (crt_set_recompute; 0 =
  // every wavenumber's potentials stored), and whether the last forward
  // stored them all
  int recompute = 0;
  bool kpot_full = false;
  // CG iterations as one captured HIP graph (crt_set_graph), with the kernel
  // arguments and iteration count it was captured for
  bool use_graph = false;
  // distinct electrode pairs of the loaded measurements (sensitivity kernel)
  std::vector<int> pairs_key;
  std::vector<int2> sens_cp, sens_pp, sens_mp;
  std::vector<int> sens_ord;
  std::vector<uint2> sens_goff;  // Gram form: tile-buffer offsets of each measurement's terms
  std::vector<int> sens_gtiles;  // Gram form: the tiles they read
  int sens_form = 0;            // crt_set_sens_form
  // decoupled edges of the triangular smoothness operator (decouplings.dat):
  // element pairs (1-based) and their weight factors
  std::vector<int2> dec_pairs;
  std::vector<double> dec_str;
  // element-to-parameter map (modelmod::mnr) for bpar/bsigma; empty = identity
  int manz_map = 0;
  std::vector<int> mnr_host;
  hipGraph_t cg_graph = nullptr;
  hipGraphExec_t cg_exec = nullptr;
  std::vector<unsigned char> cg_key;
  int sens_mode = -1;
  int nanz = 0;
  std::vector<int> strnr, vnr;
  // device
  hipStream_t stream = nullptr;
  // second stream (default priority, like the first): back substitutions that
  // overlap the factorisation's last round; joined back before the forward ends
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fac = nullptr, ev_join = nullptr;
  int ncu = 0;
  size_t lds_max = 0;
  int dev = 0;  // the device current at crt_create
  hipError_t hip_err = hipSuccess;
  std::map<std::string, void*> buf;
  std::map<std::string, size_t> bufsz;
  int slots = 0;
  int R = 1;
  size_t batch_bytes_max = (size_t)64 << 30;  // device bytes for factors in flight
  // launch choices set by crt_set_tuning (0 = the built-in rule): factor
  // workgroup threads (256 or 512) and back-substitution prefetch depth (8, 16)
  int tune_nt = 0, tune_depth = 8;
  int last_err = 0, last_k = 0, last_l = 0;
  // factor-kernel phase clock (wall-clock ticks summed over workgroups)
  bool kprof = false;
  unsigned long long kprof_ticks[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // timing
  bool timing = false;
  std::vector<Timer> pending;
  std::map<std::string, std::pair<double, int>> times;

  template <class T>
  T* get(const std::string& name, size_t count) {
    size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    auto it = buf.find(name);
    if (it != buf.end() && bufsz[name] >= bytes) return static_cast<T*>(it->second);
    if (it != buf.end()) {
      hipFree(it->second);
      buf.erase(it);
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    buf[name] = p;
    bufsz[name] = bytes;
    return static_cast<T*>(p);
  }
  template <class T>
  T* at(const std::string& name) {
    auto it = buf.find(name);
    return it == buf.end() ? nullptr : static_cast<T*>(it->second);
  }
};

namespace {

// With CRT_VERBOSE_ERRORS set in the environment, a failing HIP call is
// reported on stderr (the call, HIP's error string, the source line) before
// the entry point returns 1001.
void report_hip(const char* what, hipError_t e, int line) {
  static const bool on = std::getenv("CRT_VERBOSE_ERRORS") != nullptr;
  if (on) std::fprintf(stderr, "crt: %s failed: %s (crt_api.cpp:%d)\n", what, hipGetErrorString(e), line);
}

#define HIPCHK(x)                          \
  do {                                     \
    hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) {                \
      h->hip_err = e_;                     \
      report_hip(#x, e_, __LINE__);        \
      return kErrHip;                      \
    }                                      \
  } while (0)

#define NEED(ptr)                 \
  do {                            \
    if (!(ptr)) return kErrHip;   \
  } while (0)

void tstart(crt_handle* h, const char* name) {
  if (!h->timing) return;
  Timer t;
  t.name = name;
  hipEventCreate(&t.a);
  hipEventCreate(&t.b);
  hipEventRecord(t.a, h->stream);
  h->pending.push_back(t);
}
void tstop(crt_handle* h) {
  if (!h->timing || h->pending.empty()) return;
  hipEventRecord(h->pending.back().b, h->stream);
}
// A stage on another stream: returns its slot for tstop_on.
int tstart_on(crt_handle* h, const char* name, hipStream_t s) {
  if (!h->timing) return -1;
  Timer t;
  t.name = name;
  hipEventCreate(&t.a);
  hipEventCreate(&t.b);
  hipEventRecord(t.a, s);
  h->pending.push_back(t);
  return (int)h->pending.size() - 1;
}
void tstop_on(crt_handle* h, int slot, hipStream_t s) {
  if (slot >= 0 && slot < (int)h->pending.size()) hipEventRecord(h->pending[slot].b, s);
}
void tcollect(crt_handle* h) {
  for (auto& t : h->pending) {
    float ms = 0.f;
    hipEventSynchronize(t.b);
    hipEventElapsedTime(&ms, t.a, t.b);
    auto& slot = h->times[t.name];
    slot.first += ms;
    slot.second += 1;
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  h->pending.clear();
}

int finish(crt_handle* h) {
  hipError_t e = hipStreamSynchronize(h->stream);
  tcollect(h);
  if (e != hipSuccess) {
    h->hip_err = e;
    report_hip("hipStreamSynchronize", e, __LINE__);
    return kErrHip;
  }
  e = hipGetLastError();
  if (e != hipSuccess) {
    h->hip_err = e;
    report_hip("hipGetLastError", e, __LINE__);
    return kErrHip;
  }
  return 0;
}

// Device-to-host copy on the handle's stream, waited for.  A synchronous
// hipMemcpy goes through the legacy default stream, which HIP refuses while
// another thread captures a graph on its handle's stream -- and the refusal
// invalidates that capture -- so the entry points never use one.
hipError_t d2h(crt_handle* h, void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return e;
}

int read_err(crt_handle* h) {
  int herr[2] = {0, 0};
  int* d = h->at<int>("err");
  if (hipMemcpyAsync(herr, d, sizeof(herr), hipMemcpyDeviceToHost, h->stream) != hipSuccess) return kErrHip;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return kErrHip;
  if (herr[0]) {
    h->last_err = herr[0];
    if (h->E > 0) {
      h->last_k = herr[1] / h->E + 1;
      h->last_l = herr[1] % h->E + 1;
    }
  }
  return herr[0];
}

int clear_err(crt_handle* h) {
  int* d = h->get<int>("err", 2);
  NEED(d);
  HIPCHK(hipMemsetAsync(d, 0, 2 * sizeof(int), h->stream));
  return 0;
}

template <class T>
int upload(crt_handle* h, const std::string& name, const T* src, size_t count, T** out) {
  T* d = h->get<T>(name, count);
  NEED(d);
  if (count) HIPCHK(hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, h->stream));
  *out = d;
  return 0;
}

// ---- wavenumbers (src/rwaven.f90, gauleg.f90, gaulag.f90, gammln.f90) ----
double gammln(double xx) {
  static const double cof[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                -1.231739572450155, .1208650973866179e-2, -.5395239384953e-5};
  const double stp = 2.5066282746310005;
  double x = xx, y = x, tmp = x + 5.5;
  tmp = (x + 0.5) * std::log(tmp) - tmp;
  double ser = 1.000000000190015;
  for (int j = 0; j < 6; ++j) {
    y = y + 1.0;
    ser = ser + cof[j] / y;
  }
  return tmp + std::log(stp * ser / x);
}

void gauss_legendre(double x1, double x2, double* x, double* w, int n) {
  const double eps = 3.0e-14;
  const int m = (n + 1) / 2;
  const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; ++i) {
    double z = std::cos(3.141592653589793 * ((double)i - .25) / ((double)n + .5));
    double z1, p1, p2, p3, pp;
    do {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * (double)j - 1.0) * z * p2 - ((double)j - 1.0) * p3) / (double)j;
      }
      pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (std::fabs(z - z1) > eps);
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

void gauss_laguerre(double* x, double* w, int n, double alf) {
  const double eps = 3.0e-14;
  double z = 0.0, z1, p1 = 0, p2 = 0, p3, pp = 0;
  for (int i = 1; i <= n; ++i) {
    if (i == 1) {
      z = (1.0 + alf) * (3.0 + .92 * alf) / (1.0 + 2.4 * n + 1.8 * alf);
    } else if (i == 2) {
      z = z + (15.0 + 6.25 * alf) / (1.0 + .9 * alf + 2.5 * n);
    } else {
      double ai = (double)(i - 2);
      z = z + ((1.0 + 2.55 * ai) / (1.9 * ai) + 1.26 * ai * alf / (1.0 + 3.5 * ai)) * (z - x[i - 3]) /
                  (1.0 + .3 * alf);
    }
    for (int its = 1; its <= 10; ++its) {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        p3 = p2;
        p2 = p1;
        p1 = (((double)(2 * j - 1) + alf - z) * p2 - ((double)(j - 1) + alf) * p3) / (double)j;
      }
      pp = ((double)n * p1 - ((double)n + alf) * p2) / z;
      z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) <= eps) break;
    }
    x[i - 1] = z;
    w[i - 1] = -std::exp(gammln(alf + (double)n) - gammln((double)n)) / (pp * (double)n * p2);
  }
}

double node_x(const crt_handle* h, int node1) { return h->sx[h->snr[node1 - 1] - 1]; }
double node_y(const crt_handle* h, int node1) { return h->sy[h->snr[node1 - 1] - 1]; }

}  // namespace

extern "C" {

// Row-assembly plan (kompab's accumulation, src/kompab.f90:79-129), with the
// Dirichlet rows of randb (type-13 nodes) and of randb2 (prescribed nodes) as
// unit rows whose couplings are dropped, and for randb2 the terms that move to
// the right-hand side: b(j) += value_k * a(k,j) for every free node j coupled
// to a prescribed node k (src/randb2.f90:41-71), grouped by j in the order of
// the boundary file, each a(k,j) as the list of its element terms in element
// order.  Uploads the plan arrays; rebuilt when the boundary values change.
static int build_plan(crt_handle* h) {
  struct Term {
    int row, t, e, imn;
  };
  std::vector<Term> terms;
  terms.reserve((size_t)h->nel * 10);
  std::vector<uint8_t> is13(h->n, 0);
  int iel = 0;
  for (int ti = 0; ti < h->typanz; ++ti) {
    const int ty = h->typ[ti], nk = h->selanz[ti];
    for (int j = 0; j < h->nelanz[ti]; ++j, ++iel) {
      if (ty == 13) {
        for (int q = 0; q < nk; ++q) is13[h->nrel[iel + (size_t)q * h->nel] - 1] = 1;
      }
      if (ty > 11) continue;
      int ikl = 0;
      for (int kk = 0; kk < nk; ++kk) {
        const int nzp = h->nrel[iel + (size_t)kk * h->nel];
        for (int ll = 0; ll <= kk; ++ll, ++ikl) {
          const int nnp = h->nrel[iel + (size_t)ll * h->nel];
          if (std::abs(nzp - nnp) > h->mb) return 19;
          const int row = std::max(nzp, nnp) - 1, col = std::min(nzp, nnp) - 1;
          const int code = (ty == 11) ? -((iel - h->elanz) + 1) : iel;
          terms.push_back({row, col - row + h->mb, code, ikl});
        }
      }
    }
  }
  // randb2: first and last position of each prescribed node in the file
  const int nrw = (int)h->rwdnr.size();
  std::vector<int> first_ir(h->n, -1), last_ir(h->n, -1);
  for (int ir = 0; ir < nrw; ++ir) {
    const int k = h->rwdnr[ir] - 1;
    if (first_ir[k] < 0) first_ir[k] = ir;
    last_ir[k] = ir;
  }
  std::vector<uint8_t> unit(h->n, 0);
  for (int r = 0; r < h->n; ++r) unit[r] = is13[r] || last_ir[r] >= 0;
  // moved terms: a(k,j) with k prescribed (and not zeroed before by randb)
  // and j free, used at k's first position in the file
  struct Move {
    int j, ir, e, imn, seq;
  };
  std::vector<Move> moves;
  if (nrw > 0) {
    int seq = 0;
    for (const Term& t : terms) {
      const int r = t.row, c = t.row - h->mb + t.t;
      if (r == c) continue;
      const int ends[2][2] = {{r, c}, {c, r}};
      for (const auto& kj : ends) {
        const int k = kj[0], j = kj[1];
        if (first_ir[k] >= 0 && !is13[k] && !unit[j]) moves.push_back({j, first_ir[k], t.e, t.imn, seq++});
      }
    }
    std::stable_sort(moves.begin(), moves.end(), [](const Move& a, const Move& b) {
      return a.j != b.j ? a.j < b.j : (a.ir != b.ir ? a.ir < b.ir : a.seq < b.seq);
    });
  }
  {
    std::vector<Term> kept;
    kept.reserve(terms.size());
    for (const Term& t : terms) {
      const int col = t.row - h->mb + t.t;
      if (unit[t.row] || unit[col]) continue;
      kept.push_back(t);
    }
    for (int r = 0; r < h->n; ++r)
      if (unit[r]) kept.push_back({r, h->mb, kDirichletUnit, 0});
    terms.swap(kept);
  }
  std::stable_sort(terms.begin(), terms.end(), [](const Term& a, const Term& b) {
    return a.row != b.row ? a.row < b.row : a.t < b.t;
  });
  // row pattern (distinct band offsets, ascending) and the split into area
  // terms (summed once per wavenumber) and mixed-boundary terms (per source)
  std::vector<std::vector<int>> pat(h->n);
  for (const Term& t : terms)
    if (pat[t.row].empty() || pat[t.row].back() != t.t) pat[t.row].push_back(t.t);
  for (int r = 0; r < h->n; ++r)
    if (pat[r].empty() || pat[r].back() != h->mb) pat[r].push_back(h->mb);  // diagonal slot
  int R = 1;
  for (auto& v : pat) R = std::max<int>(R, (int)v.size());
  if (R > kMaxR) return kErrBand;
  h->R = R;
  std::vector<int16_t> pat_t((size_t)h->n * R, -1);
  std::vector<int8_t> pat_diag(h->n, 0);
  for (int r = 0; r < h->n; ++r)
    for (size_t q = 0; q < pat[r].size(); ++q) {
      pat_t[(size_t)r * R + q] = (int16_t)pat[r][q];
      if (pat[r][q] == h->mb) pat_diag[r] = (int8_t)q;
    }
  std::vector<int> rowptr(h->n + 1, 0), te, bptr(h->n + 1, 0), bcode;
  std::vector<int16_t> tt;
  std::vector<int8_t> timn, bslot, bimn;
  for (const Term& t : terms) {
    if (t.e < 0 && t.e != kDirichletUnit) {
      const auto& pr = pat[t.row];
      bptr[t.row + 1]++;
      bslot.push_back((int8_t)(std::find(pr.begin(), pr.end(), t.t) - pr.begin()));
      bcode.push_back(t.e);
      bimn.push_back((int8_t)t.imn);
    } else {
      rowptr[t.row + 1]++;
      tt.push_back((int16_t)t.t);
      timn.push_back((int8_t)t.imn);
      te.push_back(t.e);
    }
  }
  for (int r = 0; r < h->n; ++r) {
    rowptr[r + 1] += rowptr[r];
    bptr[r + 1] += bptr[r];
  }
  if (te.empty()) { tt.push_back(0); timn.push_back(0); te.push_back(0); }
  if (bcode.empty()) { bslot.push_back(0); bcode.push_back(-1); bimn.push_back(0); }
  int* di;
  int16_t* d16;
  int8_t* d8;
  uint8_t* du;
  int rc = 0;
  rc |= upload(h, "rowptr", rowptr.data(), rowptr.size(), &di);
  rc |= upload(h, "term_t", tt.data(), tt.size(), &d16);
  rc |= upload(h, "term_imn", timn.data(), timn.size(), &d8);
  rc |= upload(h, "term_e", te.data(), te.size(), &di);
  rc |= upload(h, "pat_t", pat_t.data(), pat_t.size(), &d16);
  rc |= upload(h, "pat_diag", pat_diag.data(), pat_diag.size(), &d8);
  rc |= upload(h, "bptr", bptr.data(), bptr.size(), &di);
  rc |= upload(h, "bslot", bslot.data(), bslot.size(), &d8);
  rc |= upload(h, "bcode", bcode.data(), bcode.size(), &di);
  rc |= upload(h, "bimn", bimn.data(), bimn.size(), &d8);
  if (h->has13) rc |= upload(h, "isdir", is13.data(), is13.size(), &du);
  if (nrw > 0) {
    std::vector<int> mv_ptr(h->n + 1, 0), mv_ir, mv_e;
    std::vector<int8_t> mv_imn;
    for (const Move& mv : moves) {
      mv_ptr[mv.j + 1]++;
      mv_ir.push_back(mv.ir);
      mv_e.push_back(mv.e);
      mv_imn.push_back((int8_t)mv.imn);
    }
    for (int r = 0; r < h->n; ++r) mv_ptr[r + 1] += mv_ptr[r];
    if (mv_ir.empty()) { mv_ir.push_back(0); mv_e.push_back(0); mv_imn.push_back(0); }
    double2* dz;
    rc |= upload(h, "mv_ptr", mv_ptr.data(), mv_ptr.size(), &di);
    rc |= upload(h, "mv_ir", mv_ir.data(), mv_ir.size(), &di);
    rc |= upload(h, "mv_e", mv_e.data(), mv_e.size(), &di);
    rc |= upload(h, "mv_imn", mv_imn.data(), mv_imn.size(), &d8);
    rc |= upload(h, "rb_last", last_ir.data(), last_ir.size(), &di);
    rc |= upload(h, "rwd", h->rwd.data(), h->rwd.size(), &dz);
  }
  if (rc || hipStreamSynchronize(h->stream) != hipSuccess) return kErrHip;
  return 0;
}

int crt_create(crt_handle** out, const crt_mesh* m) {
  if (!out || !m || m->sanz <= 0 || m->typanz <= 0 || m->mb <= 0 || m->eanz < 0) return kErrArg;
  crt_handle* h = new crt_handle();
  h->n = m->sanz;
  h->mb = m->mb;
  h->E = m->eanz;
  h->typanz = m->typanz;
  h->typ.assign(m->typ, m->typ + m->typanz);
  h->nelanz.assign(m->nelanz, m->nelanz + m->typanz);
  h->selanz.assign(m->selanz, m->selanz + m->typanz);
  bool seen_bnd = false;
  for (int i = 0; i < h->typanz; ++i) {
    const int t = h->typ[i];
    if (t > 10) {
      seen_bnd = true;
      h->relanz += h->nelanz[i];
      if (t == 11) h->has11 = true;
      if (t == 13) h->has13 = true;
      if (h->selanz[i] != 2 || (t != 11 && t != 12 && t != 13)) { delete h; return kErrArg; }
    } else {
      if (seen_bnd) { delete h; return kErrArg; }  // area elements must precede boundaries
      if (!((t == 3 && h->selanz[i] == 3) || ((t == 5 || t == 8) && h->selanz[i] == 4))) {
        delete h;
        return 18;
      }
      h->elanz += h->nelanz[i];
    }
    h->maxnodes = std::max(h->maxnodes, h->selanz[i]);
  }
  h->nel = h->elanz + h->relanz;
  h->nimn = h->maxnodes * (h->maxnodes + 1) / 2;
  h->nrel.assign(m->nrel, m->nrel + (size_t)h->nel * h->maxnodes);
  h->rnr.assign(m->rnr, m->rnr + h->relanz);
  h->snr.assign(m->snr, m->snr + h->n);
  h->sx.assign(m->sx, m->sx + h->n);
  h->sy.assign(m->sy, m->sy + h->n);
  h->enr.assign(m->enr, m->enr + h->E);
  h->lsytop = m->lsytop;
  h->lsink = m->lsink;
  h->nsink = m->nsink;
  for (int v : h->snr)
    if (v < 1 || v > h->n) { delete h; return kErrArg; }
  for (int v : h->enr)
    if (v < 1 || v > h->n) { delete h; return 29; }  // relectr errnr 29
  for (int v : h->rnr)
    if (v < 1 || v > h->elanz) { delete h; return 111; }
  for (size_t i = 0; i < h->nrel.size(); ++i) {
    const int iel = (int)(i % h->nel), q = (int)(i / h->nel);
    int ty = 0, nk = 0, acc = 0;
    for (int t = 0; t < h->typanz; ++t) {
      if (iel < acc + h->nelanz[t]) { ty = h->typ[t]; nk = h->selanz[t]; break; }
      acc += h->nelanz[t];
    }
    (void)ty;
    if (q < nk && (h->nrel[i] < 1 || h->nrel[i] > h->n)) { delete h; return kErrArg; }
  }
  if (h->lsink && (h->nsink < 1 || h->nsink > h->n)) { delete h; return kErrArg; }

  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    delete h;
    return kErrHip;
  }
  h->dev = dev;
  if (noalloc_setup(dev)) {
    delete h;
    return kErrHip;
  }
  h->ncu = prop.multiProcessorCount;
  h->lds_max = prop.sharedMemPerBlock;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return kErrHip;
  }
  if (hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_fac, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess) {
    crt_destroy(h);
    return kErrHip;
  }

  {
    const int rc = build_plan(h);
    if (rc) {
      crt_destroy(h);
      return rc;
    }
  }
  // ---- element / boundary / node arrays for the kernels ----
  std::vector<int> el_type(h->elanz), el_nodes((size_t)h->elanz * 4), el_nk(h->elanz);
  std::vector<int> bnd_type(std::max(h->relanz, 1)), bnd_nodes((size_t)std::max(h->relanz, 1) * 2);
  int iel = 0;
  for (int ti = 0; ti < h->typanz; ++ti) {
    const int ty = h->typ[ti], nk = h->selanz[ti];
    for (int j = 0; j < h->nelanz[ti]; ++j, ++iel) {
      if (ty <= 10) {
        el_type[iel] = ty;
        el_nk[iel] = nk;
        for (int q = 0; q < 4; ++q) el_nodes[(size_t)iel * 4 + q] = h->nrel[iel + (size_t)std::min(q, nk - 1) * h->nel] - 1;
      } else {
        const int rel = iel - h->elanz;
        bnd_type[rel] = ty;
        bnd_nodes[2 * rel] = h->nrel[iel] - 1;
        bnd_nodes[2 * rel + 1] = h->nrel[iel + (size_t)h->nel] - 1;
      }
    }
  }
  std::vector<double> nx(h->n), ny(h->n);
  for (int j = 0; j < h->n; ++j) {
    nx[j] = node_x(h, j + 1);
    ny[j] = node_y(h, j + 1);
  }
  std::vector<int> enr0(std::max(h->E, 1)), rnr0(std::max(h->relanz, 1));
  for (int i = 0; i < h->E; ++i) enr0[i] = h->enr[i] - 1;
  for (int i = 0; i < h->relanz; ++i) rnr0[i] = h->rnr[i] - 1;

  int* di;
  double* dd;
  int rc = 0;
  rc |= upload(h, "el_type", el_type.data(), el_type.size(), &di);
  rc |= upload(h, "el_nodes", el_nodes.data(), el_nodes.size(), &di);
  rc |= upload(h, "el_nk", el_nk.data(), el_nk.size(), &di);
  rc |= upload(h, "bnd_type", bnd_type.data(), bnd_type.size(), &di);
  rc |= upload(h, "bnd_nodes", bnd_nodes.data(), bnd_nodes.size(), &di);
  rc |= upload(h, "nx", nx.data(), nx.size(), &dd);
  rc |= upload(h, "ny", ny.data(), ny.size(), &dd);
  rc |= upload(h, "enr", enr0.data(), enr0.size(), &di);
  rc |= upload(h, "rnr", rnr0.data(), rnr0.size(), &di);
  rc |= clear_err(h);
  if (rc || hipStreamSynchronize(h->stream) != hipSuccess) {
    crt_destroy(h);
    return kErrHip;
  }
  // bsytop (src/bsytop.f90:35-69): running mean of the no-flow element centres
  if (h->lsytop) {
    double st = 0.0;
    iel = 0;
    for (int ti = 0; ti < h->typanz; ++ti) {
      if (h->typ[ti] == 12) {
        const int nk = h->selanz[ti];
        st = 0.0;
        for (int j = 1; j <= h->nelanz[ti]; ++j, ++iel) {
          double sp = 0.0;
          for (int q = 0; q < nk; ++q) sp = sp + node_y(h, h->nrel[iel + (size_t)q * h->nel]);
          sp = sp / (double)nk;
          st = (j == 1) ? sp : (st * (double)(j - 1) + sp) / (double)j;
        }
      } else {
        iel += h->nelanz[ti];
      }
    }
    h->sytop = st;
  }
  *out = h;
  return 0;
}

int crt_destroy(crt_handle* h) {
  if (!h) return 0;
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->stream2) hipStreamSynchronize(h->stream2);
  if (h->cg_exec) hipGraphExecDestroy(h->cg_exec);
  if (h->cg_graph) hipGraphDestroy(h->cg_graph);
  for (auto& kv : h->buf) hipFree(kv.second);
  for (auto& t : h->pending) {
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  if (h->ev_fac) hipEventDestroy(h->ev_fac);
  if (h->ev_join) hipEventDestroy(h->ev_join);
  if (h->stream2) hipStreamDestroy(h->stream2);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int crt_rwaven(crt_handle* h, int32_t nanz, const int32_t* strnr, const int32_t* vnr, int32_t* kwnanz,
               double* kwn, double* kwnwi) {
  if (!h || nanz <= 0 || !strnr || !vnr) return kErrArg;
  double amin = 1e9, amax = 0.0;
  double xke[4] = {0, 0, 0, 0}, yke[4] = {0, 0, 0, 0};
  for (int i = 0; i < nanz; ++i) {
    int el[4];
    el[0] = strnr[i] % 10000;
    el[1] = (strnr[i] - el[0]) / 10000;
    el[2] = vnr[i] % 10000;
    el[3] = (vnr[i] - el[2]) / 10000;
    for (int j = 0; j < 4; ++j) {
      if (el[j] > h->E || el[j] < 0) return 46;
      if (el[j] > 0) {
        xke[j] = node_x(h, h->enr[el[j] - 1]);
        yke[j] = node_y(h, h->enr[el[j] - 1]);
      }
    }
    for (int q = 2; q < 4; ++q) {
      if (el[q] <= 0) continue;
      for (int j = 0; j < 2; ++j) {
        if (el[j] > 0) {
          const double x21 = xke[q] - xke[j], y21 = yke[q] - yke[j];
          const double dum = std::sqrt(x21 * x21 + y21 * y21);
          amin = std::fmin(dum, amin);
          amax = std::fmax(dum, amax);
        }
      }
    }
  }
  if (amin < 1e-12 || amax < 1e-12) return 73;
  const int lanz = 4;
  int ganz = (int)(float)(6.0 * std::log10(amax / amin));  // int(real(...)), single precision
  ganz = std::max(ganz, 2);
  ganz = std::min(ganz, 30 - lanz);
  const int K = ganz + lanz;
  const double kwn0 = 1.0 / (2.0 * amin);
  std::vector<double> x(30, 0.0), w(30, 0.0);
  gauss_legendre(0.0, 1.0, x.data(), w.data(), ganz);
  for (int i = 0; i < ganz; ++i) {
    w[i] = 2.0 * kwn0 * w[i] * x[i];
    x[i] = kwn0 * x[i] * x[i];
  }
  gauss_laguerre(x.data() + ganz, w.data() + ganz, lanz, 0.0);
  for (int i = ganz; i < K; ++i) {
    w[i] = kwn0 * w[i] * std::exp(x[i]);
    x[i] = kwn0 * (x[i] + 1.0);
  }
  h->K = K;
  h->swrtr = 1;
  h->kwn.assign(x.begin(), x.begin() + K);
  h->kwnwi.assign(w.begin(), w.begin() + K);
  h->precal_done = h->forward_done = h->volt_done = h->sens_done = false;
  if (kwnanz) *kwnanz = K;
  if (kwn) std::copy(h->kwn.begin(), h->kwn.end(), kwn);
  if (kwnwi) std::copy(h->kwnwi.begin(), h->kwnwi.end(), kwnwi);
  return 0;
}

int crt_set_wavenumbers(crt_handle* h, int32_t kwnanz, const double* kwn, const double* kwnwi, int32_t swrtr) {
  if (!h || kwnanz <= 0 || kwnanz > 30 || !kwn || !kwnwi) return kErrArg;
  h->K = kwnanz;
  h->swrtr = swrtr ? 1 : 0;
  h->kwn.assign(kwn, kwn + kwnanz);
  h->kwnwi.assign(kwnwi, kwnwi + kwnanz);
  h->precal_done = h->forward_done = h->volt_done = h->sens_done = false;
  return 0;
}

int crt_precal(crt_handle* h, int32_t* lbeta, int32_t* lrandb) {
  if (!h) return kErrArg;
  if (h->K <= 0) return kErrOrder;
  if (h->has11 && h->has13) return 101;
  double *dkwn, *dkwnwi;
  int rc = upload(h, "kwn", h->kwn.data(), h->K, &dkwn);
  rc |= upload(h, "kwnwi", h->kwnwi.data(), h->K, &dkwnwi);
  if (rc) return kErrHip;
  PrecalArgs P;
  P.elanz = h->elanz;
  P.relanz = h->relanz;
  P.eanz = h->E;
  P.K = h->K;
  P.nimn = h->nimn;
  P.el_type = h->at<int>("el_type");
  P.el_nodes = h->at<int>("el_nodes");
  P.bnd_type = h->at<int>("bnd_type");
  P.bnd_nodes = h->at<int>("bnd_nodes");
  P.nx = h->at<double>("nx");
  P.ny = h->at<double>("ny");
  P.enr = h->at<int>("enr");
  P.kwn = dkwn;
  P.sytop = h->sytop;
  P.elbg = h->get<double>("elbg", (size_t)h->K * h->elanz * h->nimn);
  P.relbg = h->get<double>("relbg", (size_t)std::max(h->relanz, 1) * h->nimn);
  P.kg = h->get<double>("kg", (size_t)h->K * h->E * std::max(h->relanz, 1));
  NEED(P.elbg);
  NEED(P.relbg);
  NEED(P.kg);
  if (clear_err(h)) return kErrHip;
  P.err = h->at<int>("err");
  HIPCHK(hipMemsetAsync(P.relbg, 0, sizeof(double) * std::max(h->relanz, 1) * h->nimn, h->stream));
  const int nt = 128;
  tstart(h, "precal_area");
  long na = (long)h->elanz * h->K;
  hipLaunchKernelGGL(crt_precal_area_kernel, dim3((unsigned)((na + nt - 1) / nt)), dim3(nt), 0, h->stream, P);
  tstop(h);
  long nb = (long)h->relanz * h->E * h->K;
  if (nb > 0) {
    tstart(h, "precal_bnd");
    hipLaunchKernelGGL(crt_precal_bnd_kernel, dim3((unsigned)((nb + nt - 1) / nt)), dim3(nt), 0, h->stream, P);
    tstop(h);
  }
  int rc2 = finish(h);
  if (rc2) return rc2;
  int e = read_err(h);
  if (e) return e;
  h->precal_done = true;
  h->forward_done = h->volt_done = h->sens_done = false;
  if (lbeta) *lbeta = h->has11 ? 1 : 0;
  if (lrandb) *lrandb = h->has13 ? 1 : 0;
  return 0;
}

int crt_set_boundary_values(crt_handle* h, int32_t rwdanz, const int32_t* rwdnr, const double* rwd) {
  if (!h || rwdanz < 0 || (rwdanz > 0 && (!rwdnr || !rwd))) return kErrArg;
  for (int i = 0; i < rwdanz; ++i)
    if (rwdnr[i] < 1 || rwdnr[i] > h->n) return 30;  // rrandb errnr 30
  h->rwdnr.assign(rwdnr, rwdnr + rwdanz);
  h->rwd.resize(rwdanz);
  for (int i = 0; i < rwdanz; ++i) h->rwd[i] = make_double2(rwd[2 * i], rwd[2 * i + 1]);
  const int rc = build_plan(h);
  if (rc) return rc;
  h->forward_done = h->volt_done = h->sens_done = false;
  return 0;
}

int crt_get_precal(crt_handle* h, double* elbg, double* relbg, double* kg) {
  if (!h) return kErrArg;
  if (!h->precal_done) return kErrOrder;
  const int K = h->K, el = h->elanz, rl = h->relanz, nimn = h->nimn, E = h->E;
  if (elbg) {
    std::vector<double> t((size_t)K * el * nimn);
    HIPCHK(d2h(h, t.data(), h->at<double>("elbg"), t.size() * 8));
    for (int k = 0; k < K; ++k)
      for (int e = 0; e < el; ++e)
        for (int q = 0; q < nimn; ++q) elbg[e + (size_t)el * (q + (size_t)nimn * k)] = t[((size_t)k * el + e) * nimn + q];
  }
  if (relbg && rl > 0) {
    std::vector<double> t((size_t)rl * nimn);
    HIPCHK(d2h(h, t.data(), h->at<double>("relbg"), t.size() * 8));
    for (int r = 0; r < rl; ++r)
      for (int q = 0; q < nimn; ++q) relbg[r + (size_t)rl * q] = t[(size_t)r * nimn + q];
  }
  if (kg && rl > 0) {
    // device layout [K][E][relanz] is the reference's kg(relanz,eanz,kwnanz)
    HIPCHK(d2h(h, kg, h->at<double>("kg"), (size_t)K * E * rl * 8));
  }
  return 0;
}

// The dynamic-LDS limit of a kernel (hipFuncAttributeMaxDynamicSharedMemorySize)
// belongs to the kernel on its device, not to a handle: handles in flight on
// several threads would lower it under one another's launches if each set its
// own size.  It is raised once per device and kernel to all a workgroup can
// take (the CU's LDS less the kernel's static LDS) and left there.
// *dyn (optional) receives that dynamic limit.
static int open_lds(crt_handle* h, const void* kern, size_t* dyn = nullptr) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> open;
  std::lock_guard<std::mutex> lock(mu);
  auto it = open.find({h->dev, kern});
  if (it == open.end()) {
    hipFuncAttributes fa;
    HIPCHK(hipFuncGetAttributes(&fa, kern));
    const size_t lim = h->lds_max - std::min(h->lds_max, (size_t)fa.sharedSizeBytes);
    HIPCHK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim));
    it = open.emplace(std::make_pair(h->dev, kern), lim).first;
  }
  if (dyn) *dyn = it->second;
  return 0;
}

// The substitutions keep a window of 64 NS rows in NS register slots per lane
// and refill every G rows: G < 64 NS + 1 - mb keeps a refill ahead of its
// first update.  NS = 2 up to mb = 124, NS = 4 up to kMaxBand.  T = double
// for the DC twin's real factors.
static void launch_backsolve(int mb, int nrhs, hipStream_t st, const BacksolveArgs& B, int depth, bool real) {
#define BS(D, G, NS) (real ? crt_backsolve_kernel<D, G, NS, double> : crt_backsolve_kernel<D, G, NS, cplx>)
  void (*k)(BacksolveArgs);
  if (depth == 16 && mb <= 120) k = BS(16, 8, 2);
  else if (depth == 16 && mb > 124) k = BS(16, 8, 4);
  else if (mb <= 120) k = BS(8, 8, 2);
  else if (mb <= 124) k = BS(8, 4, 2);
  else k = BS(8, 8, 4);
#undef BS
  hipLaunchKernelGGL(k, dim3(nrhs), dim3(64), 0, st, B);
}
static void launch_fwdsweep(int mb, int nrhs, hipStream_t st, const FwdArgs& F, bool real) {
#define FS(D, G, NS) (real ? crt_fwdsweep_kernel<D, G, NS, double> : crt_fwdsweep_kernel<D, G, NS, cplx>)
  void (*k)(FwdArgs);
  if (mb <= 120) k = FS(8, 8, 2);
  else if (mb <= 124) k = FS(8, 4, 2);
  else k = FS(8, 8, 4);
#undef FS
  hipLaunchKernelGGL(k, dim3(nrhs), dim3(64), 0, st, F);
}

// The factor kernel's form for a handle's band: the single-model kernel, its
// multi-model twin (null for forms without one), threads per workgroup and
// dynamic LDS.  Tuning (crt_set_tuning) may ask for 256 or 512 threads.  The
// DC twin (h->dc) takes the real-entry kernel of the same shape: half the
// LDS per entry, so narrower bands keep two workgroups per CU.
struct FactorForm {
  void (*kern)(FactorArgs) = nullptr;
  void (*many)(const FactorArgs*, const int2*, int, int*) = nullptr;
  int nt = 256;
  size_t lds = 0;
  bool gring = false;
};

static int choose_factor(const crt_handle* h, FactorForm& F) {
  const int n = h->n, mb = h->mb;
  // the substitutions address one factor with 32-bit byte offsets
  if (mb > kMaxBand || (size_t)n * (mb + 1) * 16 >= (size_t(1) << 31)) return kErrBand;
  // The trailing-window ring lives in LDS up to mb = 124 (when it fits), in a
  // per-workgroup global scratch above (the wide-band form).  The DC twin's
  // real entries take half the LDS per entry.
  const size_t es = h->dc ? 8 : 16;
  const bool gring = mb > 124 || crt_factor_lds_bytes(mb, es, false) > h->lds_max;
  const size_t lds = crt_factor_lds_bytes(mb, es, gring);
  if (lds > h->lds_max) return kErrBand;
  // wave 0 holds the W = mb + kNB panel rows in ROWS slots per lane (256
  // threads: slot 0 holds 32 rows and four copies of the kNB diagonal rows,
  // every further slot 64 rows; 512 threads: 64 rows per slot); the other
  // waves own the trailing window's 4x4 tiles beyond the look-ahead columns,
  // MAXT each.  When one workgroup's LDS leaves no room for a second on the
  // CU, the workgroup takes 512 threads instead of 256 (the wide-band form
  // always).
  const int W = mb + kNB, tw = (mb + 3) / 4;
  const int nrest = tw > 2 ? ((tw - 2) * (tw - 1)) / 2 : 0;
  int nt = gring || 2 * lds > h->lds_max ? 512 : 256;
  if (!gring && (h->tune_nt == 256 || h->tune_nt == 512)) nt = h->tune_nt;
  const int rows = nt == 256 ? (W <= 40 ? 1 : 1 + (W - 40 + 63) / 64) : (W + 63) / 64;
  const int maxt = std::max(1, (nrest + nt - 65) / (nt - 64));
#define FORM(R, T, N, G)                                                   \
  (h->dc ? (F.kern = crt_factor_kernel<R, T, N, G, double>, F.many = nullptr) \
         : (F.kern = crt_factor_kernel<R, T, N, G, cplx>, F.many = crt_factor_many_kernel<R, T, N, G>))
  if (gring) {
    if (rows <= 3 && maxt <= 5) FORM(3, 5, 512, true);
    else if (rows == 4 && maxt <= 5) FORM(4, 5, 512, true);
    else if (rows == 5 && maxt <= 5) FORM(5, 5, 512, true);
  } else if (nt == 256) {
    if (rows == 1 && maxt == 1) FORM(1, 1, 256, false);
    else if (rows == 2 && maxt == 1) FORM(2, 1, 256, false);
    else if (rows == 2 && maxt == 2) FORM(2, 2, 256, false);
    else if (rows == 2 && maxt == 3) FORM(2, 3, 256, false);
    else if (rows == 3 && maxt <= 3) FORM(3, 3, 256, false);
  } else {
    if (rows <= 2 && maxt == 1) FORM(2, 1, 512, false);
    else if (rows <= 2 && maxt == 2) FORM(2, 2, 512, false);
    else if (rows == 3 && maxt <= 2) FORM(3, 2, 512, false);
  }
#undef FORM
  if (!F.kern) return kErrBand;
  F.nt = nt;
  F.lds = lds;
  F.gring = gring;
  return 0;
}

// The model-level fields of the factor and assembly kernels' arguments (the
// per-batch buffers are set by the caller).
static void factor_args(crt_handle* h, FactorArgs& A) {
  std::memset(&A, 0, sizeof(A));
  A.n = h->n;
  A.mb = h->mb;
  A.eanz = h->E;
  A.elanz = h->elanz;
  A.relanz = h->relanz;
  A.nimn = h->nimn;
  A.rowptr = h->at<int>("rowptr");
  A.term_t = h->at<int16_t>("term_t");
  A.term_imn = h->at<int8_t>("term_imn");
  A.term_e = h->at<int>("term_e");
  A.R = h->R;
  A.pat_t = h->at<int16_t>("pat_t");
  A.pat_diag = h->at<int8_t>("pat_diag");
  A.bptr = h->at<int>("bptr");
  A.bslot = h->at<int8_t>("bslot");
  A.bcode = h->at<int>("bcode");
  A.bimn = h->at<int8_t>("bimn");
  A.elbg = h->at<double>("elbg");
  A.relbg = h->at<double>("relbg");
  A.kg = h->at<double>("kg");
  A.sigma = h->at<double2>("sigma");
  A.rnr = h->at<int>("rnr");
  A.enr = h->at<int>("enr");
  A.isdir = h->has13 ? h->at<uint8_t>("isdir") : nullptr;
  A.lsink = h->lsink;
  A.nsink = h->nsink - 1;
  A.dc = h->dc ? 1 : 0;
  A.ring = nullptr;
  A.ring_stride = 0;
  A.sys_stride = 1;
}

// The (k, l) systems of wavenumbers k0 .. k0 + kc - 1: assembly, scaling,
// factorisation and substitutions, kpot of those wavenumbers into kpot_dst
// ([kc][E][n]).  `area` forms the per-wavenumber area sums of every
// wavenumber first (once per conductivity model; a recomputing pass reuses
// them).  Queued on the handle's stream.
static int run_forward_range(crt_handle* h, int k0, int kc, double2* kpot_dst, bool area) {
  const int n = h->n, mb = h->mb, E = h->E, K = h->K;
  FactorForm form;
  if (int rc = choose_factor(h, form)) return rc;
  void (*kern)(FactorArgs) = form.kern;
  const int nt = form.nt;
  const size_t lds = form.lds;
  const bool gring = form.gring;
  if (int rc = open_lds(h, reinterpret_cast<const void*>(kern))) return rc;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, nt, lds) != hipSuccess || per_cu < 1) per_cu = 1;
  const int nsys = kc * E;
  FactorArgs A;
  factor_args(h, A);
  double2* av = h->get<double2>("av", (size_t)K * n * h->R);
  double2* kpot = kpot_dst;
  NEED(av);
  NEED(kpot);
  A.av = av;
  A.err = h->at<int>("err");
  A.prof = nullptr;
  if (h->kprof) {
    A.prof = h->get<unsigned long long>("kprof", 8);  // zeroed by run_forward
    NEED(A.prof);
  }
  // per-wavenumber area sums
  if (area) {
    tstart(h, "area_values");
    hipLaunchKernelGGL(crt_area_values_kernel, dim3((unsigned)(((size_t)K * n + 255) / 256)), dim3(256), 0,
                       h->stream, A, K, av);
    tstop(h);
  }
  // Without mixed boundaries the matrix does not depend on the source: one
  // factor per wavenumber serves all eanz sources (src/fem.f90:372-376,
  // kompb), each source then takes both sweeps.  Otherwise every (k, l) system
  // is factored with its own forward sweep fused in.
  const bool reuse = !h->has11;
  const int nfac = reuse ? kc : nsys;
  const int fac0 = reuse ? k0 : k0 * E;  // first factor (system, or wavenumber when shared)
  const int rpf = reuse ? E : 1;  // right-hand sides per factor
  A.sys_stride = reuse ? E : 1;
  // factors per batch: the scaled rows, L, y and the scaling factors of every
  // factor in the batch stay in device memory until its back substitution
  const bool rb2 = !h->rwdnr.empty();
  const size_t per_sys = (size_t)n * (mb + 1) * 16 + (size_t)n * h->R * 16 + (size_t)n * 24 +
                         (reuse ? (size_t)E * n * 16 + (size_t)n * (mb + 1) * 16 : 0) + (rb2 ? (size_t)n * 16 : 0);
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
  size_t budget = std::min<size_t>(free_b / 2, h->batch_bytes_max);
  const size_t have = h->bufsz.count("L") && h->bufsz.count("S")
                          ? h->bufsz["L"] + h->bufsz["S"] + h->bufsz["y"] + h->bufsz["fak"] +
                                (h->bufsz.count("yr") ? h->bufsz["yr"] : 0) +
                                (h->bufsz.count("Lc") ? h->bufsz["Lc"] : 0)
                          : 0;
  budget = std::max(budget, have);
  int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nfac, budget / per_sys));
  A.L = h->get<double2>("L", (size_t)batch * n * (mb + 1));
  A.S = h->get<double2>("S", (size_t)batch * n * h->R);
  A.y = h->get<double2>("y", (size_t)batch * n);
  A.fak = h->get<double>("fak", (size_t)batch * n);
  NEED(A.L);
  NEED(A.S);
  NEED(A.y);
  NEED(A.fak);
  double2* yr = nullptr;
  double2* Lc = nullptr;
  if (reuse) {
    yr = h->get<double2>("yr", (size_t)batch * E * n);
    Lc = h->get<double2>("Lc", (size_t)batch * n * (mb + 1));
    NEED(yr);
    NEED(Lc);
  }
  BacksolveArgs B;
  B.n = n;
  B.mb = mb;
  B.eanz = E;
  B.rhs_per_fac = rpf;
  B.fstride = A.sys_stride;
  B.L = A.L;
  B.y = reuse ? yr : A.y;
  B.fak = A.fak;
  B.kpot = kpot;
  B.kpot_k0 = k0;
  FwdArgs F;
  F.n = n;
  F.mb = mb;
  F.eanz = E;
  F.Lc = Lc;
  F.fak = A.fak;
  F.enr = A.enr;
  F.isdir = A.isdir;
  F.lsink = A.lsink;
  F.nsink = A.nsink;
  F.y = yr;
  // prescribed boundary values: one assembled right-hand side per system
  A.mv_ptr = A.mv_ir = A.mv_e = A.rb_last = nullptr;
  A.mv_imn = nullptr;
  A.rwd = nullptr;
  A.rhs = nullptr;
  if (rb2) {
    A.mv_ptr = h->at<int>("mv_ptr");
    A.mv_ir = h->at<int>("mv_ir");
    A.mv_e = h->at<int>("mv_e");
    A.mv_imn = h->at<int8_t>("mv_imn");
    A.rb_last = h->at<int>("rb_last");
    A.rwd = h->at<double2>("rwd");
    A.rhs = h->get<double2>("rhs", (size_t)batch * n);
    NEED(A.rhs);
  }
  F.rhs = A.rhs;
  for (int first = fac0; first < fac0 + nfac; first += batch) {
    const int cnt = std::min(batch, fac0 + nfac - first);
    A.sys_first = first;
    A.nsys = cnt;
    const unsigned rgrid = (unsigned)(((size_t)cnt * n + 255) / 256);
    tstart(h, "assemble_scale");
    hipLaunchKernelGGL(crt_scale_kernel, dim3(rgrid), dim3(256), 0, h->stream, A);
    hipLaunchKernelGGL(h->dc ? crt_rows_kernel<double> : crt_rows_kernel<cplx>,
                       dim3((unsigned)(((size_t)cnt * n * h->R + 255) / 256)), dim3(256), 0,
                       h->stream, A);
    if (rb2) hipLaunchKernelGGL(crt_rhs_kernel, dim3(rgrid), dim3(256), 0, h->stream, A);
    tstop(h);
    h->slots = std::min(cnt, h->ncu * per_cu);
    if (gring) {
      A.ring_stride = (size_t)crt_ring_entries(mb);
      A.ring = h->get<double2>("ring", (size_t)h->slots * A.ring_stride);
      NEED(A.ring);
    }
    const int nrhs = cnt * rpf;
    // Own factors per system: the systems of the persistent kernel's full
    // rounds are factored first, then the partial last round, during which
    // the full rounds' back substitutions run on the second stream (a CU in
    // the last round holds fewer factor workgroups and has room for them).
    // Only for forwards of at least two full rounds: smaller ones leave the
    // device to other handles' work, which the cross-stream wait would hold.
    const int full = (cnt / h->slots) * h->slots;
    if (!reuse && full >= 2 * h->slots && full < cnt) {
      for (int part = 0; part < 2; ++part) {
        const int p0 = part ? full : 0, pc = part ? cnt - full : full;
        FactorArgs Ap = A;
        Ap.sys_first = first + p0;
        Ap.nsys = pc;
        // system p0's scaled rows, L and y: entries of 16 bytes, or 8 for
        // the DC twin's real kernels
        const size_t es = h->dc ? 8 : 16;
        auto at_entry = [es](double2* p, size_t entries) {
          return reinterpret_cast<double2*>(reinterpret_cast<char*>(p) + entries * es);
        };
        Ap.L = at_entry(A.L, (size_t)p0 * n * (mb + 1));
        Ap.S = at_entry(A.S, (size_t)p0 * n * h->R);
        Ap.y = at_entry(A.y, (size_t)p0 * n);
        Ap.fak = A.fak + (size_t)p0 * n;
        if (A.rhs) Ap.rhs = A.rhs + (size_t)p0 * n;
        tstart(h, "factor_solve");
        hipLaunchKernelGGL(kern, dim3(std::min(pc, h->slots)), dim3(nt), lds, h->stream, Ap);
        tstop(h);
        BacksolveArgs Bp = B;
        Bp.nrhs = pc;
        Bp.fac_first = first + p0;
        Bp.L = Ap.L;
        Bp.y = Ap.y;
        Bp.fak = Ap.fak;
        hipStream_t bs = part ? h->stream : h->stream2;
        if (!part) {
          HIPCHK(hipEventRecord(h->ev_fac, h->stream));
          HIPCHK(hipStreamWaitEvent(h->stream2, h->ev_fac, 0));
        }
        const int slot = tstart_on(h, "backsolve", bs);
        launch_backsolve(mb, pc, bs, Bp, h->tune_depth, h->dc);
        tstop_on(h, slot, bs);
      }
      HIPCHK(hipEventRecord(h->ev_join, h->stream2));
      HIPCHK(hipStreamWaitEvent(h->stream, h->ev_join, 0));
      continue;
    }
    tstart(h, "factor_solve");
    hipLaunchKernelGGL(kern, dim3(h->slots), dim3(nt), lds, h->stream, A);
    tstop(h);
    if (reuse) {
      F.nrhs = nrhs;
      F.fac_first = first;
      tstart(h, "column_copy");
      const size_t ncopy = (size_t)cnt * n * (mb + 1);
      hipLaunchKernelGGL(h->dc ? crt_lcol_kernel<double> : crt_lcol_kernel<cplx>,
                         dim3((unsigned)((ncopy + 255) / 256)), dim3(256), 0, h->stream,
                         (const double2*)A.L, Lc, n, mb, cnt);
      tstop(h);
      tstart(h, "forward_sweep");
      launch_fwdsweep(mb, nrhs, h->stream, F, h->dc);
      tstop(h);
    }
    B.nrhs = nrhs;
    B.fac_first = first;
    tstart(h, "backsolve");
    launch_backsolve(mb, nrhs, h->stream, B, h->tune_depth, h->dc);
    tstop(h);
  }
  return 0;
}

// The electrode slice of the back-transform for the wavenumbers in kpot
// (k0 .. k0 + kc - 1), carried between passes in "he_acc".
static int rtrafo_elec(crt_handle* h, const double2* kpot, int k0, int kc) {
  const int E = h->E;
  double2* he = h->get<double2>("he", (size_t)E * E);
  double2* acc = h->get<double2>("he_acc", (size_t)E * E);
  NEED(he);
  NEED(acc);
  tstart(h, "rtrafo_elec");
  hipLaunchKernelGGL(crt_rtrafo_elec_kernel, dim3((E * E + 255) / 256), dim3(256), 0, h->stream, kpot,
                     (const double*)h->at<double>("kwnwi"), (const int*)h->at<int>("enr"), h->n, E, k0, kc,
                     h->swrtr, k0 == 0 ? 1 : 0, k0 + kc == h->K ? 1 : 0, acc, he);
  tstop(h);
  return 0;
}

// Wavenumbers per pass of the recomputing mode (crt_set_recompute), 0 when
// every wavenumber's potentials are stored.
static int recompute_pass(const crt_handle* h) { return h->recompute > 0 ? std::min(h->recompute, h->K) : 0; }

// The forward of every (k, l) system: kpot of all wavenumbers stored, or, in
// the recomputing mode, passes of recompute_pass() wavenumbers of which only
// the electrode slice of the back-transform is kept.
static int run_forward(crt_handle* h) {
  const int n = h->n, E = h->E, K = h->K;
  if (clear_err(h)) return kErrHip;
  const int kp = recompute_pass(h);
  if (h->kprof) {
    unsigned long long* prof = h->get<unsigned long long>("kprof", 8);
    NEED(prof);
    HIPCHK(hipMemsetAsync(prof, 0, 8 * sizeof(unsigned long long), h->stream));
  }
  if (!kp) {
    double2* kpot = h->get<double2>("kpot", (size_t)K * E * n);
    NEED(kpot);
    if (int rc = run_forward_range(h, 0, K, kpot, true)) return rc;
    if (int rc = rtrafo_elec(h, kpot, 0, K)) return rc;
  } else {
    double2* kpot = h->get<double2>("kpot_pass", (size_t)kp * E * n);
    NEED(kpot);
    for (int k0 = 0; k0 < K; k0 += kp) {
      const int kc = std::min(kp, K - k0);
      if (int rc = run_forward_range(h, k0, kc, kpot, k0 == 0)) return rc;
      if (int rc = rtrafo_elec(h, kpot, k0, kc)) return rc;
    }
  }
  int rc = finish(h);
  if (rc) return rc;
  int e = read_err(h);
  if (e) return e;
  if (h->kprof) {
    unsigned long long t[8];
    HIPCHK(d2h(h, t, h->at<unsigned long long>("kprof"), sizeof(t)));
    for (int i = 0; i < 8; ++i) h->kprof_ticks[i] += t[i];
  }
  h->forward_done = true;
  h->kpot_full = kp == 0;
  h->volt_done = h->sens_done = false;
  h->kpt_valid = false;
  return 0;
}

// Back-substitution form for a band (the rule of launch_backsolve).
static void (*backsolve_many_form(int mb, int depth))(const BacksolveArgs*, const int2*, int) {
  if (depth == 16 && mb <= 120) return crt_backsolve_many_kernel<16, 8, 2>;
  if (depth == 16 && mb > 124) return crt_backsolve_many_kernel<16, 8, 4>;
  if (mb <= 120) return crt_backsolve_many_kernel<8, 8, 2>;
  if (mb <= 124) return crt_backsolve_many_kernel<8, 4, 2>;
  return crt_backsolve_many_kernel<8, 8, 4>;
}

int crt_forward_many(int32_t nmodels, crt_handle* const* hs, const double* const* sigma_dev) {
  if (nmodels <= 0 || !hs || !sigma_dev) return kErrArg;
  for (int m = 0; m < nmodels; ++m) {
    if (!hs[m] || !sigma_dev[m]) return kErrArg;
    if (hs[m]->dev != hs[0]->dev) return kErrArg;
    for (int j = 0; j < m; ++j)
      if (hs[j] == hs[m]) return kErrArg;
    if (!hs[m]->precal_done) return kErrOrder;
    if (hs[m]->recompute) return kErrArg;  // the batched forward stores every model's potentials
  }
  crt_handle* h0 = hs[0];
  hipStream_t st = h0->stream;
  struct Model {
    crt_handle* h;
    FactorForm form;
    FactorArgs A;
    long long cost;
  };
  std::vector<Model> ms;
  // ---- per model: conductivities, area sums, scaled rows (and right-hand
  // sides) of every system on its own stream; models without mixed
  // boundaries (one factor per wavenumber) run the single-model forward
  for (int m = 0; m < nmodels; ++m) {
    crt_handle* h = hs[m];
    double2* ds = h->get<double2>("sigma", h->elanz);
    NEED(ds);
    if (hipMemcpyAsync(ds, sigma_dev[m], (size_t)h->elanz * 16, hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
      return kErrHip;
    h->dc = false;
    Model M;
    M.h = h;
    if (int rc = choose_factor(h, M.form)) return rc;
    if (!h->has11 || !M.form.many) {
      if (int rc = run_forward(h)) return rc;
      continue;
    }
    const int n = h->n, mb = h->mb, E = h->E, K = h->K, nsys = K * E;
    FactorArgs& A = M.A;
    factor_args(h, A);
    A.nsys = nsys;
    A.sys_first = 0;
    double2* av = h->get<double2>("av", (size_t)K * n * h->R);
    NEED(av);
    NEED(h->get<double2>("kpot", (size_t)K * E * n));
    A.av = av;
    if (clear_err(h)) return kErrHip;
    A.err = h->at<int>("err");
    A.L = h->get<double2>("L", (size_t)nsys * n * (mb + 1));
    A.S = h->get<double2>("S", (size_t)nsys * n * h->R);
    A.y = h->get<double2>("y", (size_t)nsys * n);
    A.fak = h->get<double>("fak", (size_t)nsys * n);
    NEED(A.L);
    NEED(A.S);
    NEED(A.y);
    NEED(A.fak);
    A.mv_ptr = A.mv_ir = A.mv_e = A.rb_last = nullptr;
    A.mv_imn = nullptr;
    A.rwd = nullptr;
    A.rhs = nullptr;
    const bool rb2 = !h->rwdnr.empty();
    if (rb2) {
      A.mv_ptr = h->at<int>("mv_ptr");
      A.mv_ir = h->at<int>("mv_ir");
      A.mv_e = h->at<int>("mv_e");
      A.mv_imn = h->at<int8_t>("mv_imn");
      A.rb_last = h->at<int>("rb_last");
      A.rwd = h->at<double2>("rwd");
      A.rhs = h->get<double2>("rhs", (size_t)nsys * n);
      NEED(A.rhs);
    }
    hipLaunchKernelGGL(crt_area_values_kernel, dim3((unsigned)(((size_t)K * n + 255) / 256)), dim3(256), 0,
                       h->stream, A, K, av);
    const unsigned rgrid = (unsigned)(((size_t)nsys * n + 255) / 256);
    hipLaunchKernelGGL(crt_scale_kernel, dim3(rgrid), dim3(256), 0, h->stream, A);
    hipLaunchKernelGGL(crt_rows_kernel<cplx>, dim3((unsigned)(((size_t)nsys * n * h->R + 255) / 256)), dim3(256), 0,
                       h->stream, A);
    if (rb2) hipLaunchKernelGGL(crt_rhs_kernel, dim3(rgrid), dim3(256), 0, h->stream, A);
    HIPCHK(hipEventRecord(h->ev_fac, h->stream));
    HIPCHK(hipStreamWaitEvent(st, h->ev_fac, 0));
    M.cost = (long long)n * mb * mb;
    ms.push_back(M);
  }
  if (ms.empty()) return 0;
  crt_handle* h = h0;  // HIPCHK records errors on the first handle
  // ---- one persistent factor launch per kernel form over all its models'
  // systems, the costliest models' systems first
  std::vector<std::vector<FactorArgs>> keep_args;
  std::vector<std::vector<int2>> keep_items;
  std::map<const void*, std::vector<int>> bins;
  for (int i = 0; i < (int)ms.size(); ++i) bins[reinterpret_cast<const void*>(ms[i].form.many)].push_back(i);
  int ib = 0;
  const int slot_t = tstart_on(h0, "factor_solve_many", st);
  for (auto& kv : bins) {
    std::vector<int>& idx = kv.second;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ms[a].cost > ms[b].cost; });
    const FactorForm& F = ms[idx[0]].form;
    size_t lds = 0, stride = 0;
    for (int i : idx) {
      lds = std::max(lds, ms[i].form.lds);
      stride = std::max(stride, (size_t)crt_ring_entries(ms[i].h->mb));
    }
    if (int rc = open_lds(h0, reinterpret_cast<const void*>(F.many))) return rc;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, F.many, F.nt, lds) != hipSuccess || per_cu < 1)
      per_cu = 1;
    keep_args.emplace_back();
    keep_items.emplace_back();
    std::vector<FactorArgs>& args = keep_args.back();
    std::vector<int2>& items = keep_items.back();
    for (int j = 0; j < (int)idx.size(); ++j) {
      args.push_back(ms[idx[j]].A);
      for (int sys = 0; sys < ms[idx[j]].A.nsys; ++sys) items.push_back(make_int2(j, sys));
    }
    const int nitems = (int)items.size();
    const int grid = std::min(nitems, h0->ncu * per_cu);
    if (F.gring) {
      double2* ring = h0->get<double2>("many_ring" + std::to_string(ib), (size_t)grid * stride);
      NEED(ring);
      for (auto& a : args) {
        a.ring = ring;
        a.ring_stride = stride;
      }
    }
    FactorArgs* dargs = h0->get<FactorArgs>("many_fargs" + std::to_string(ib), args.size());
    int2* ditems = h0->get<int2>("many_fitems" + std::to_string(ib), items.size());
    int* counter = h0->get<int>("many_counter" + std::to_string(ib), 1);
    NEED(dargs);
    NEED(ditems);
    NEED(counter);
    HIPCHK(hipMemcpyAsync(dargs, args.data(), args.size() * sizeof(FactorArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ditems, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(counter, 0, sizeof(int), st));
    hipLaunchKernelGGL(F.many, dim3(grid), dim3(F.nt), lds, st, (const FactorArgs*)dargs, (const int2*)ditems,
                       nitems, counter);
    ++ib;
  }
  tstop_on(h0, slot_t, st);
  // ---- back substitutions, one launch per form over every right-hand side
  std::map<const void*, std::vector<int>> bbins;
  for (int i = 0; i < (int)ms.size(); ++i)
    bbins[reinterpret_cast<const void*>(backsolve_many_form(ms[i].h->mb, ms[i].h->tune_depth))].push_back(i);
  std::vector<std::vector<BacksolveArgs>> keep_bargs;
  const int slot_b = tstart_on(h0, "backsolve_many", st);
  ib = 0;
  for (auto& kv : bbins) {
    keep_bargs.emplace_back();
    keep_items.emplace_back();
    std::vector<BacksolveArgs>& bargs = keep_bargs.back();
    std::vector<int2>& items = keep_items.back();
    for (int j = 0; j < (int)kv.second.size(); ++j) {
      const Model& M = ms[kv.second[j]];
      BacksolveArgs B;
      B.n = M.h->n;
      B.mb = M.h->mb;
      B.eanz = M.h->E;
      B.nrhs = M.A.nsys;
      B.rhs_per_fac = 1;
      B.fac_first = 0;
      B.fstride = 1;
      B.L = M.A.L;
      B.y = M.A.y;
      B.fak = M.A.fak;
      B.kpot = M.h->at<double2>("kpot");
      B.kpot_k0 = 0;
      bargs.push_back(B);
      for (int r = 0; r < B.nrhs; ++r) items.push_back(make_int2(j, r));
    }
    BacksolveArgs* dargs = h0->get<BacksolveArgs>("many_bargs" + std::to_string(ib), bargs.size());
    int2* ditems = h0->get<int2>("many_bitems" + std::to_string(ib), items.size());
    NEED(dargs);
    NEED(ditems);
    HIPCHK(hipMemcpyAsync(dargs, bargs.data(), bargs.size() * sizeof(BacksolveArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ditems, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, st));
    auto kern = reinterpret_cast<void (*)(const BacksolveArgs*, const int2*, int)>(const_cast<void*>(kv.first));
    hipLaunchKernelGGL(kern, dim3((unsigned)items.size()), dim3(64), 0, st, (const BacksolveArgs*)dargs,
                       (const int2*)ditems, (int)items.size());
    ++ib;
  }
  tstop_on(h0, slot_b, st);
  // ---- electrode rows of the back-transform, per model
  for (const Model& M : ms) {
    crt_handle* hm = M.h;
    double2* he = hm->get<double2>("he", (size_t)hm->E * hm->E);
    NEED(he);
    hipLaunchKernelGGL(crt_rtrafo_elec_kernel, dim3((hm->E * hm->E + 255) / 256), dim3(256), 0, st,
                       (const double2*)hm->at<double2>("kpot"), (const double*)hm->at<double>("kwnwi"),
                       (const int*)M.A.enr, hm->n, hm->E, 0, hm->K, hm->swrtr, 1, 1, he, he);
  }
  if (int rc = finish(h0)) return rc;
  for (const Model& M : ms) {
    const int e = read_err(M.h);
    if (e) return e;
    M.h->forward_done = true;
    M.h->kpot_full = true;
    M.h->volt_done = M.h->sens_done = false;
    M.h->kpt_valid = false;
  }
  return 0;
}

int crt_set_tuning(crt_handle* h, int32_t factor_threads, int32_t backsolve_depth) {
  if (!h) return kErrArg;
  if (factor_threads != 0 && factor_threads != 256 && factor_threads != 512) return kErrArg;
  if (backsolve_depth != 0 && backsolve_depth != 8 && backsolve_depth != 16) return kErrArg;
  h->tune_nt = factor_threads;
  h->tune_depth = backsolve_depth == 0 ? 8 : backsolve_depth;
  return 0;
}

int crt_set_sens_form(crt_handle* h, int32_t form) {
  if (!h || form < 0 || form > 2) return kErrArg;
  if (form == 2 && h->E > kGramMaxE) return kErrArg;
  h->sens_form = form;
  return 0;
}

int crt_set_recompute(crt_handle* h, int32_t wavenumbers_per_pass) {
  if (!h || wavenumbers_per_pass < 0) return kErrArg;
  h->recompute = wavenumbers_per_pass;
  return 0;
}

int crt_forward(crt_handle* h, const double* sigma, double* kpot, double* hpot) {
  if (!h || !sigma) return kErrArg;
  if (!h->precal_done) return kErrOrder;
  if (h->recompute && (kpot || hpot)) return kErrArg;  // the potentials are not stored
  double2* ds;
  if (upload(h, "sigma", reinterpret_cast<const double2*>(sigma), h->elanz, &ds)) return kErrHip;
  h->dc = false;
  int rc = run_forward(h);
  if (rc) return rc;
  const int n = h->n, E = h->E, K = h->K;
  if (kpot) HIPCHK(d2h(h, kpot, h->at<double2>("kpot"), (size_t)K * E * n * 16));
  if (hpot) {
    double2* dh = h->get<double2>("hpot", (size_t)E * n);
    NEED(dh);
    tstart(h, "rtrafo");
    hipLaunchKernelGGL(crt_rtrafo_kernel, dim3((unsigned)(((size_t)n * E + 255) / 256)), dim3(256), 0, h->stream,
                       (const double2*)h->at<double2>("kpot"), (const double*)h->at<double>("kwnwi"), n, E, K,
                       h->swrtr, dh);
    tstop(h);
    rc = finish(h);
    if (rc) return rc;
    HIPCHK(d2h(h, hpot, dh, (size_t)E * n * 16));
  }
  return 0;
}

int crt_forward_dev(crt_handle* h, const double* sigma_dev) {
  if (!h || !sigma_dev) return kErrArg;
  if (!h->precal_done) return kErrOrder;
  double2* ds = h->get<double2>("sigma", h->elanz);
  NEED(ds);
  HIPCHK(hipMemcpyAsync(ds, sigma_dev, (size_t)h->elanz * 16, hipMemcpyDeviceToDevice, h->stream));
  h->dc = false;
  return run_forward(h);
}

// Real parts of n complex device values into a host array, through a device
// staging buffer.
static int copy_real(crt_handle* h, const double2* src, size_t n, double* host) {
  double* d = h->get<double>("dc_real", n);
  NEED(d);
  hipLaunchKernelGGL(crt_real_part_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, src, n, d);
  int rc = finish(h);
  if (rc) return rc;
  HIPCHK(d2h(h, host, d, n * 8));
  return 0;
}

int crt_forward_dc(crt_handle* h, const double* sigma, double* kpotdc, double* hpotdc) {
  if (!h || !sigma) return kErrArg;
  if (!h->precal_done) return kErrOrder;
  if (h->recompute && (kpotdc || hpotdc)) return kErrArg;
  // the caller's sigma stays for bsendc's sigma/V; the assembly takes its real part
  double2* din;
  if (upload(h, "sigma_dc", reinterpret_cast<const double2*>(sigma), h->elanz, &din)) return kErrHip;
  double2* ds = h->get<double2>("sigma", h->elanz);
  NEED(ds);
  hipLaunchKernelGGL(crt_real_sigma_kernel, dim3((h->elanz + 255) / 256), dim3(256), 0, h->stream,
                     (const double2*)din, h->elanz, ds);
  h->dc = true;
  int rc = run_forward(h);
  if (rc) return rc;
  const int n = h->n, E = h->E, K = h->K;
  if (kpotdc) {
    rc = copy_real(h, h->at<double2>("kpot"), (size_t)K * E * n, kpotdc);
    if (rc) return rc;
  }
  if (hpotdc) {
    double2* dh = h->get<double2>("hpot", (size_t)E * n);
    NEED(dh);
    tstart(h, "rtrafo");
    hipLaunchKernelGGL(crt_rtrafo_kernel, dim3((unsigned)(((size_t)n * E + 255) / 256)), dim3(256), 0, h->stream,
                       (const double2*)h->at<double2>("kpot"), (const double*)h->at<double>("kwnwi"), n, E, K,
                       h->swrtr, dh);
    tstop(h);
    rc = copy_real(h, dh, (size_t)E * n, hpotdc);
    if (rc) return rc;
  }
  return 0;
}

static int run_voltages(crt_handle* h, int32_t nanz, const int32_t* strnr, const int32_t* vnr, int32_t mode,
                        const double* kfak) {
  if (nanz <= 0 || !strnr || !vnr) return kErrArg;
  if (!h->forward_done) return kErrOrder;
  for (int i = 0; i < nanz; ++i) {
    const int e1 = strnr[i] % 10000, e2 = strnr[i] / 10000, e3 = vnr[i] % 10000, e4 = vnr[i] / 10000;
    if (e1 < 0 || e1 > h->E || e2 < 0 || e2 > h->E || e3 < 0 || e3 > h->E || e4 < 0 || e4 > h->E) return 46;
  }
  h->nanz = nanz;
  h->strnr.assign(strnr, strnr + nanz);
  h->vnr.assign(vnr, vnr + nanz);
  int *ds, *dv;
  double* dk = nullptr;
  int rc = upload(h, "strnr", strnr, nanz, &ds);
  rc |= upload(h, "vnr", vnr, nanz, &dv);
  if (kfak) rc |= upload(h, "kfak", kfak, nanz, &dk);
  if (rc) return kErrHip;
  double2* vo = h->get<double2>("volt", nanz);
  double2* sa = h->get<double2>("sigmaa", nanz);
  NEED(vo);
  NEED(sa);
  if (clear_err(h)) return kErrHip;
  tstart(h, "volt");
  hipLaunchKernelGGL(crt_volt_kernel, dim3((nanz + 255) / 256), dim3(256), 0, h->stream,
                     (const double2*)h->at<double2>("he"), h->E, nanz, (const int*)ds, (const int*)dv, mode,
                     (const double*)dk, vo, sa, h->at<int>("err"));
  tstop(h);
  rc = finish(h);
  if (rc) return rc;
  int e = read_err(h);
  if (e) return e;
  h->volt_done = true;
  h->sens_done = false;
  return 0;
}

int crt_voltages(crt_handle* h, int32_t nanz, const int32_t* strnr, const int32_t* vnr, int32_t mode,
                 const double* kfak, double* volt, double* sigmaa) {
  if (!h) return kErrArg;
  int rc = run_voltages(h, nanz, strnr, vnr, mode, kfak);
  if (rc) return rc;
  if (volt) HIPCHK(d2h(h, volt, h->at<double2>("volt"), (size_t)nanz * 16));
  if (sigmaa) HIPCHK(d2h(h, sigmaa, h->at<double2>("sigmaa"), (size_t)nanz * 16));
  return 0;
}

int crt_voltages_dev(crt_handle* h, int32_t nanz, const int32_t* strnr, const int32_t* vnr, int32_t mode,
                     double* volt_dev, double* sigmaa_dev) {
  if (!h) return kErrArg;
  int rc = run_voltages(h, nanz, strnr, vnr, mode, nullptr);
  if (rc) return rc;
  if (volt_dev) HIPCHK(hipMemcpyAsync(volt_dev, h->at<double2>("volt"), (size_t)nanz * 16, hipMemcpyDeviceToDevice, h->stream));
  if (sigmaa_dev)
    HIPCHK(hipMemcpyAsync(sigmaa_dev, h->at<double2>("sigmaa"), (size_t)nanz * 16, hipMemcpyDeviceToDevice, h->stream));
  return finish(h);
}

// kpot transposed to [K][n][E] for the element-wise kernels; built once per
// forward
static double2* transposed_kpot(crt_handle* h) {
  double2* kpt = h->get<double2>("kpt", (size_t)h->K * h->n * h->E);
  if (!kpt) return nullptr;
  if (!h->kpt_valid) {
    tstart(h, "kpot_transpose");
    hipLaunchKernelGGL(crt_kpot_transpose_kernel, dim3((h->n + 63) / 64, (h->E + 15) / 16, h->K), dim3(256), 0,
                       h->stream, (const double2*)h->at<double2>("kpot"), h->n, h->E, kpt);
    tstop(h);
    h->kpt_valid = true;
  }
  return kpt;
}

static int run_sens(crt_handle* h, int32_t mode, const double2* sigma, bool real = false) {
  if (!h->kpot_full) return kErrOrder;  // the explicit Jacobian reads every wavenumber's potentials
  SensArgs S;
  S.n = h->n;
  S.elanz = h->elanz;
  S.eanz = h->E;
  S.K = h->K;
  S.nimn = h->nimn;
  S.nanz = h->nanz;
  S.swrtr = h->swrtr;
  S.mode = mode;
  S.el_nk = h->at<int>("el_nk");
  S.el_nodes = h->at<int>("el_nodes");
  S.elbg = h->at<double>("elbg");
  S.kwnwi = h->at<double>("kwnwi");
  S.kpot = h->at<double2>("kpot");
  S.sigma = sigma;
  S.volt = h->at<double2>("volt");
  S.strnr = h->at<int>("strnr");
  S.vnr = h->at<int>("vnr");
  S.sens = h->get<double2>("sens", (size_t)h->nanz * h->elanz);
  NEED(S.sens);
  S.kpt = transposed_kpot(h);
  NEED(S.kpt);
  // distinct current and potential pairs (rebuilt when the measurements change)
  std::vector<int> key(h->strnr);
  key.insert(key.end(), h->vnr.begin(), h->vnr.end());
  if (key != h->pairs_key) {
    std::map<int, int> ci, pi;
    for (int v : h->strnr) ci.emplace(v, 0);
    for (int v : h->vnr) pi.emplace(v, 0);
    h->sens_cp.clear();
    h->sens_pp.clear();
    for (auto& kv : ci) {
      kv.second = (int)h->sens_cp.size();
      h->sens_cp.push_back(make_int2(kv.first % 10000, kv.first / 10000));
    }
    for (auto& kv : pi) {
      kv.second = (int)h->sens_pp.size();
      h->sens_pp.push_back(make_int2(kv.first % 10000, kv.first / 10000));
    }
    h->sens_mp.resize(h->nanz);
    for (int i = 0; i < h->nanz; ++i) h->sens_mp[i] = make_int2(ci[h->strnr[i]], pi[h->vnr[i]]);
    h->sens_ord.resize(h->nanz);
    std::iota(h->sens_ord.begin(), h->sens_ord.end(), 0);
    std::stable_sort(h->sens_ord.begin(), h->sens_ord.end(), [&](int a, int b) {
      const int2 x = h->sens_mp[a], y = h->sens_mp[b];
      return x.x != y.x ? x.x < y.x : x.y < y.y;
    });
    int2 *dc, *dp, *dm;
    int* dord;
    if (upload(h, "sens_cp", h->sens_cp.data(), h->sens_cp.size(), &dc) ||
        upload(h, "sens_pp", h->sens_pp.data(), h->sens_pp.size(), &dp) ||
        upload(h, "sens_mp", h->sens_mp.data(), h->sens_mp.size(), &dm) ||
        upload(h, "sens_ord", h->sens_ord.data(), h->sens_ord.size(), &dord))
      return kErrHip;
    // the Gram form's h = G[p2][c2] - G[p2][c1] - (G[p1][c2] - G[p1][c1])
    // (0-based electrodes, -1 = none) as four tile-buffer offsets (below
    // 2^16 for E <= kGramMaxE; larger E does not take the Gram form)
    h->sens_goff.assign(h->E <= kGramMaxE ? h->nanz : 1, make_uint2(0, 0));
    for (int i = 0; i < h->nanz && h->E <= kGramMaxE; ++i) {
      const int c1 = h->strnr[i] % 10000 - 1, c2 = h->strnr[i] / 10000 - 1;
      const int p1 = h->vnr[i] % 10000 - 1, p2 = h->vnr[i] / 10000 - 1;
      const int E = h->E;
      h->sens_goff[i] = make_uint2((unsigned)crt_gram_offset(E, p2, c2) | (unsigned)crt_gram_offset(E, p2, c1) << 16,
                                   (unsigned)crt_gram_offset(E, p1, c2) | (unsigned)crt_gram_offset(E, p1, c1) << 16);
    }
    // the tiles those terms read (the rest of G is never formed)
    std::vector<char> used(crt_gram_tiles(std::min(h->E, kGramMaxE)) + 1, 0);
    for (const uint2& o : h->sens_goff)
      for (unsigned v : {o.x & 0xffffu, o.x >> 16, o.y & 0xffffu, o.y >> 16}) used[v / 256] = 1;
    h->sens_gtiles.clear();
    for (int x = 0; x + 1 < (int)used.size(); ++x)
      if (used[x]) h->sens_gtiles.push_back(x);
    if (h->sens_gtiles.empty()) h->sens_gtiles.push_back(0);
    uint2* dq;
    int* dt;
    if (upload(h, "sens_goff", h->sens_goff.data(), h->sens_goff.size(), &dq) ||
        upload(h, "sens_gtiles", h->sens_gtiles.data(), h->sens_gtiles.size(), &dt))
      return kErrHip;
    h->pairs_key.swap(key);
  }
  // The Gram form (one E x E product per element and wavenumber on the
  // matrix cores, E <= 128) when the tiles its measurements read hold at most
  // twice as many entries as there are measurements, else the pair form
  // (crt_set_sens_form overrides)
  const int gtiles = (int)h->sens_gtiles.size();
  const bool gram_fits = h->E <= kGramMaxE;
  const bool gram = gram_fits && (h->sens_form == 2 || (h->sens_form == 0 && gtiles * 256 <= 2 * h->nanz));
  if (gram) {
    // tiles per wavefront: ceil(gtiles / 4)
    // (real potentials: the DC twin's, one product per tile)
    void (*k)(SensArgs, const uint2*, const int*, int);
    if (real)
      k = gtiles <= 12 ? crt_sens_gram_kernel<3, true>
          : gtiles <= 24 ? crt_sens_gram_kernel<6, true> : crt_sens_gram_kernel<9, true>;
    else
      k = gtiles <= 12 ? crt_sens_gram_kernel<3, false>
          : gtiles <= 20 ? crt_sens_gram_kernel<5, false>
          : gtiles <= 24 ? crt_sens_gram_kernel<6, false> : crt_sens_gram_kernel<9, false>;
    size_t dyn = 0;
    if (int rc = open_lds(h, reinterpret_cast<const void*>(k), &dyn)) return rc;
    const size_t lds = crt_gram_lds_bytes(h->E);
    if (lds > dyn) return kErrArg;
    tstart(h, "sens");
    hipLaunchKernelGGL(k, dim3((h->elanz + 7) / 8 * 8), dim3(kGramNT), lds, h->stream, S, h->at<uint2>("sens_goff"),
                       h->at<int>("sens_gtiles"), gtiles);
    tstop(h);
    return finish(h);
  }
  SensPairs Q;
  Q.ncp = (int)h->sens_cp.size();
  Q.npp = (int)h->sens_pp.size();
  const size_t lds_pairs = (size_t)(4 * (h->E + 1) + kSensRow * (Q.ncp + Q.npp)) * 16;
  // the pair form: 256 threads, each with a contiguous run of measurements,
  // while its LDS (dynamic and static) leaves room for two workgroups per CU,
  // else 512 in one with the measurements interleaved (each measured the
  // faster in its case; 512 threads with 8 measurements each and two
  // workgroups per CU measured slower: the 128-register bound spills); past
  // the CU's LDS the per-measurement kernel, which holds one row of
  // potentials per element node and electrode
  size_t dyn_pairs = 0, dyn = 0;
  if (int rc = open_lds(h, reinterpret_cast<const void*>(crt_sens_pairs_kernel<false>), &dyn_pairs)) return rc;
  const bool pairs = lds_pairs <= dyn_pairs;
  const bool two = 2 * (lds_pairs + h->lds_max - dyn_pairs) <= h->lds_max;
  if (two)
    if (int rc = open_lds(h, reinterpret_cast<const void*>(crt_sens_pairs_kernel<true>), &dyn_pairs)) return rc;
  const size_t lds = pairs ? lds_pairs : (size_t)4 * h->E * 16;
  if (!pairs) {
    if (int rc = open_lds(h, reinterpret_cast<const void*>(crt_sens_kernel), &dyn)) return rc;
    if (lds > dyn) return kErrArg;
  }
  tstart(h, "sens");
  if (pairs) {
    Q.cp = h->at<int2>("sens_cp");
    Q.pp = h->at<int2>("sens_pp");
    Q.mp = h->at<int2>("sens_mp");
    Q.ord = h->at<int>("sens_ord");
    if (two)
      hipLaunchKernelGGL(crt_sens_pairs_kernel<true>, dim3(h->elanz), dim3(256), lds, h->stream, S, Q);
    else
      hipLaunchKernelGGL(crt_sens_pairs_kernel<false>, dim3(h->elanz), dim3(512), lds, h->stream, S, Q);
  } else {
    hipLaunchKernelGGL(crt_sens_kernel, dim3(h->elanz), dim3(256), lds, h->stream, S);
  }
  tstop(h);
  return finish(h);
}

int crt_sensitivities(crt_handle* h, int32_t mode, double* sens) {
  if (!h || (mode != 0 && mode != 1)) return kErrArg;
  if (!h->volt_done || h->dc) return kErrOrder;
  int rc = run_sens(h, mode, h->at<double2>("sigma"));
  if (rc) return rc;
  h->sens_done = true;
  h->sens_mode = mode;
  if (sens) HIPCHK(d2h(h, sens, h->at<double2>("sens"), (size_t)h->nanz * h->elanz * 16));
  return 0;
}

// bsendc: the bsensi sums of the real potentials times sigma/V, of which the
// real part is kept (src/bsendc.f90:183-184).
int crt_sensitivities_dc(crt_handle* h, double* sensdc) {
  if (!h || !sensdc) return kErrArg;
  if (!h->volt_done || !h->dc) return kErrOrder;
  int rc = run_sens(h, 1, h->at<double2>("sigma_dc"), true);
  if (rc) return rc;
  // the complex J of this call is not the J the CG stage expects
  h->sens_done = false;
  return copy_real(h, h->at<double2>("sens"), (size_t)h->nanz * h->elanz, sensdc);
}

static int run_adjoint(crt_handle* h, int32_t mode, const double* w_dev, const double2* r_dev, double2* g_dev) {
  const int n = h->n, E = h->E, K = h->K, N = h->nanz;
  // more electrodes than the matrix-core GEMM takes: E x kAdjColsHost entries
  // of the coefficient matrix in LDS
  const size_t lds = (size_t)E * kAdjColsHost * 16;
  if (E > 16 * kAdjTilesHost) {
    size_t dyn = 0;
    if (int rc = open_lds(h, reinterpret_cast<const void*>(crt_adj_gemm_kernel), &dyn)) return rc;
    if (lds > dyn) return kErrArg;
  }
  double2* gamma = h->get<double2>("adj_gamma", N);
  double2* C = h->get<double2>("adj_C", (size_t)E * E);
  // potentials of every wavenumber, or of one pass of the recomputing mode
  const int kp = h->kpot_full ? K : std::min(std::max(h->recompute, 1), K);
  double2* Y = h->get<double2>("adj_Y", (size_t)kp * n * E);
  NEED(gamma);
  NEED(C);
  NEED(Y);
  tstart(h, "adj_gamma");
  hipLaunchKernelGGL(crt_adj_gamma_kernel, dim3((N + 255) / 256), dim3(256), 0, h->stream, N, mode, w_dev, r_dev,
                     (const double2*)h->at<double2>("volt"), gamma);
  tstop(h);
  // measurements per potential electrode (those with e3 or e4 = a), ascending
  std::vector<int> qptr(E + 1, 0), qlist;
  for (int i = 0; i < N; ++i) {
    const int e3 = h->vnr[i] % 10000, e4 = h->vnr[i] / 10000;
    if (e3 > 0) qptr[e3]++;
    if (e4 > 0 && e4 != e3) qptr[e4]++;
  }
  for (int a = 0; a < E; ++a) qptr[a + 1] += qptr[a];
  qlist.resize(std::max(qptr[E], 1));
  {
    std::vector<int> fill(qptr.begin(), qptr.end() - 1);
    for (int i = 0; i < N; ++i) {
      const int e3 = h->vnr[i] % 10000, e4 = h->vnr[i] / 10000;
      if (e3 > 0) qlist[fill[e3 - 1]++] = i;
      if (e4 > 0 && e4 != e3) qlist[fill[e4 - 1]++] = i;
    }
  }
  int *dqp, *dql;
  if (upload(h, "adj_qptr", qptr.data(), qptr.size(), &dqp) || upload(h, "adj_qlist", qlist.data(), qlist.size(), &dql))
    return kErrHip;
  tstart(h, "adj_cmat");
  hipLaunchKernelGGL(crt_adj_cmat_kernel, dim3((E * E + 255) / 256), dim3(256), 0, h->stream, N, E,
                     (const int*)h->at<int>("strnr"), (const int*)h->at<int>("vnr"), (const int*)dqp,
                     (const int*)dql, (const double2*)gamma, C);
  tstop(h);
  AdjArgs A;
  A.n = n;
  A.elanz = h->elanz;
  A.eanz = E;
  A.K = K;
  A.nimn = h->nimn;
  A.swrtr = h->swrtr;
  A.mode = mode;
  A.el_nk = h->at<int>("el_nk");
  A.el_nodes = h->at<int>("el_nodes");
  A.elbg = h->at<double>("elbg");
  A.kwnwi = h->at<double>("kwnwi");
  A.Y = Y;
  A.sigma = h->at<double2>("sigma");
  A.grad = g_dev;
  A.acc = h->get<double2>("adj_acc", h->elanz);
  NEED(A.acc);
  // Y = P^T C and the element contraction, over all wavenumbers at once from
  // the stored potentials, or pass by pass with the potentials of each pass
  // recomputed (the same systems as the forward: assembly from the stored
  // area sums, factorisation, substitutions)
  for (int k0 = 0; k0 < K; k0 += kp) {
    const int kc = std::min(kp, K - k0);
    const double2* P;
    if (h->kpot_full) {
      P = h->at<double2>("kpot");
      A.kpt = transposed_kpot(h);
    } else {
      double2* kpass = h->get<double2>("kpot_pass", (size_t)kp * E * n);
      double2* kpt = h->get<double2>("kpt_pass", (size_t)kp * n * E);
      NEED(kpass);
      NEED(kpt);
      if (int rc = run_forward_range(h, k0, kc, kpass, false)) return rc;
      tstart(h, "kpot_transpose");
      hipLaunchKernelGGL(crt_kpot_transpose_kernel, dim3((n + 63) / 64, (E + 15) / 16, kc), dim3(256), 0,
                         h->stream, (const double2*)kpass, n, E, kpt);
      tstop(h);
      P = kpass;
      A.kpt = kpt;
    }
    NEED(A.kpt);
    A.kpot = P;
    tstart(h, "adj_gemm");
    if (E <= 16 * kAdjTilesHost) {
      // C stationary in the wavefronts' registers, 16-node tiles of the
      // potentials through LDS; about four workgroups per CU over all k
      const int ns = E <= 64 ? 16 : (E <= 96 ? 24 : 32);
      void (*kern)(const double2*, const double2*, int, int, double2*) =
          ns == 16 ? crt_adj_gemm_cs_kernel<16> : (ns == 24 ? crt_adj_gemm_cs_kernel<24> : crt_adj_gemm_cs_kernel<32>);
      const int ntiles = (n + 15) / 16;
      const int grid = std::max(1, std::min(ntiles, (1024 + kc - 1) / kc));
      hipLaunchKernelGGL(kern, dim3(grid, kc), dim3(64 * ((E + 15) / 16)), (size_t)2 * 64 * ns * 16, h->stream, P,
                         (const double2*)C, n, E, Y);
    } else {
      hipLaunchKernelGGL(crt_adj_gemm_kernel, dim3((n + 255) / 256, (E + kAdjColsHost - 1) / kAdjColsHost, kc),
                         dim3(256), lds, h->stream, P, (const double2*)C, n, E, Y);
    }
    tstop(h);
    A.k0 = k0;
    A.kc = kc;
    A.first = k0 == 0 ? 1 : 0;
    A.last = k0 + kc == K ? 1 : 0;
    tstart(h, "adj_elem");
    // 8 elements per block, the grid a multiple of 8 blocks (XCD-major mapping)
    const unsigned eblocks = (unsigned)(8 * (((h->elanz + 7) / 8 + 7) / 8));
    hipLaunchKernelGGL(crt_adj_elem_kernel, dim3(eblocks), dim3(256), 0, h->stream, A);
    tstop(h);
  }
  int rc = finish(h);
  if (rc) return rc;
  return h->kpot_full ? 0 : read_err(h);
}

int crt_adjoint_gradient(crt_handle* h, int32_t mode, const double* w, const double* resid, double* grad) {
  if (!h || !w || !resid || !grad || (mode != 0 && mode != 1)) return kErrArg;
  if (!h->volt_done || h->dc) return kErrOrder;
  double* dw;
  double2* dr;
  int rc = upload(h, "adj_w", w, h->nanz, &dw);
  rc |= upload(h, "adj_r", reinterpret_cast<const double2*>(resid), h->nanz, &dr);
  if (rc) return kErrHip;
  double2* dg = h->get<double2>("adj_g", h->elanz);
  NEED(dg);
  rc = run_adjoint(h, mode, dw, dr, dg);
  if (rc) return rc;
  HIPCHK(d2h(h, grad, dg, (size_t)h->elanz * 16));
  return 0;
}

int crt_adjoint_gradient_dev(crt_handle* h, int32_t mode, const double* w_dev, const double* resid_dev,
                             double* grad_dev) {
  if (!h || !w_dev || !resid_dev || !grad_dev || (mode != 0 && mode != 1)) return kErrArg;
  if (!h->volt_done || h->dc) return kErrOrder;
  return run_adjoint(h, mode, w_dev, reinterpret_cast<const double2*>(resid_dev),
                     reinterpret_cast<double2*>(grad_dev));
}

int crt_jhw_r(crt_handle* h, const double* w, const double* resid, double* grad) {
  if (!h || !w || !resid || !grad) return kErrArg;
  if (!h->sens_done) return kErrOrder;
  double* dw;
  double2* dr;
  int rc = upload(h, "adj_w", w, h->nanz, &dw);
  rc |= upload(h, "adj_r", reinterpret_cast<const double2*>(resid), h->nanz, &dr);
  if (rc) return kErrHip;
  double2* dg = h->get<double2>("jhwr_g", h->elanz);
  NEED(dg);
  tstart(h, "jhwr");
  hipLaunchKernelGGL(crt_jhwr_kernel, dim3((unsigned)(((size_t)h->elanz * 64 + 255) / 256)), dim3(256), 0, h->stream,
                     (const double2*)h->at<double2>("sens"), h->nanz, h->elanz, (const double*)dw,
                     (const double2*)dr, dg);
  tstop(h);
  rc = finish(h);
  if (rc) return rc;
  HIPCHK(d2h(h, grad, dg, (size_t)h->elanz * 16));
  return 0;
}

int crt_cg_solve(crt_handle* h, const double* w, const double* dat, const double* sigmaa, double lam, int32_t ltri,
                 int32_t nx, const double* smatm, int32_t smatm_cols, const int32_t* nachbar, int32_t maxnodes,
                 const double* rdiag, const double* rm, int32_t ncgmax, double eps, double* dpar, double* cgfac,
                 double* cgres) {
  if (!h || !w || !dat || !sigmaa || !smatm || !rdiag || !dpar || ncgmax <= 0 || smatm_cols <= 0) return kErrArg;
  if (!h->sens_done) return kErrOrder;
  const bool tri_op = !(ltri == 0 || ltri == 3 || ltri == 4);
  if (ltri == 15) return kErrArg;  // the stochastic (dense) operator is not supported
  if (tri_op && (!nachbar || maxnodes <= 0 || smatm_cols < maxnodes + 1)) return kErrArg;
  if (ltri == 0 && smatm_cols < 3) return kErrArg;
  const int N = h->nanz, M = h->elanz;
  CgArgs G;
  std::memset(&G, 0, sizeof(G));  // the bytes key the captured graph
  G.nanz = N;
  G.manz = M;
  G.ltri = ltri;
  G.nx = nx;
  G.maxnodes = maxnodes;
  G.lam = lam;
  G.eps = eps;
  G.J = h->at<double2>("sens");
  double* dw;
  double* drd;
  double* dsm;
  double2 *ddat, *dsa, *drm = nullptr;
  int* dnb = nullptr;
  int rc = upload(h, "cg_w", w, N, &dw);
  rc |= upload(h, "cg_rdiag", rdiag, M, &drd);
  rc |= upload(h, "cg_smatm", smatm, (size_t)M * smatm_cols, &dsm);
  rc |= upload(h, "cg_dat", reinterpret_cast<const double2*>(dat), N, &ddat);
  rc |= upload(h, "cg_sigmaa", reinterpret_cast<const double2*>(sigmaa), N, &dsa);
  if (rm) rc |= upload(h, "cg_rm", reinterpret_cast<const double2*>(rm), M, &drm);
  if (tri_op) rc |= upload(h, "cg_nachbar", nachbar, (size_t)M * maxnodes, &dnb);
  if (rc) return kErrHip;
  G.w = dw;
  G.rdiag = drd;
  G.smatm = dsm;
  G.nachbar = dnb;
  G.cgfac = h->get<double>("cg_cgfac", M);
  G.bvec = h->get<double2>("cg_bvec", M);
  G.rvec = h->get<double2>("cg_rvec", M);
  G.pvec = h->get<double2>("cg_pvec", M);
  G.ap = h->get<double2>("cg_ap", N);
  G.dpar = h->get<double2>("cg_dpar", M);
  G.cgres = h->get<double>("cg_cgres", ncgmax + 2);
  G.scal = h->get<double>("cg_scal", 8);
  // J (cgfac o p) in column slices: about 2048 workgroups over the row blocks
  {
    const int rb = (N + 255) / 256;
    G.nslice = std::max(1, std::min(std::min(128, M), (2048 + rb - 1) / rb));
    G.chunk = (M + G.nslice - 1) / G.nslice;
    G.nslice = (M + G.chunk - 1) / G.chunk;
  }
  G.part = h->get<double2>("cg_part", (size_t)G.nslice * N);
  NEED(G.part);
  NEED(G.cgfac);
  NEED(G.bvec);
  NEED(G.rvec);
  NEED(G.pvec);
  NEED(G.ap);
  NEED(G.dpar);
  NEED(G.cgres);
  NEED(G.scal);
  HIPCHK(hipMemsetAsync(G.dpar, 0, (size_t)M * 16, h->stream));
  HIPCHK(hipMemsetAsync(G.pvec, 0, (size_t)M * 16, h->stream));
  HIPCHK(hipMemsetAsync(G.scal, 0, 8 * sizeof(double), h->stream));
  HIPCHK(hipMemsetAsync(G.cgres, 0, (ncgmax + 2) * sizeof(double), h->stream));
  tstart(h, "cg_rhs");
  hipLaunchKernelGGL(crt_cg_rhs_kernel, dim3((unsigned)(((size_t)M * 64 + 255) / 256)), dim3(256), 0, h->stream, G,
                     (const double2*)ddat, (const double2*)dsa, (const double2*)drm);
  tstop(h);
  HIPCHK(hipMemcpyAsync(G.rvec, G.bvec, (size_t)M * 16, hipMemcpyDeviceToDevice, h->stream));
  // the iterations: kernels only, the recurrence's scalars and the stopping
  // test stay on the device (later iterations return at once once converged)
  auto iterations = [&](bool timed) {
    for (int it = 0; it < ncgmax; ++it) {
      if (timed) tstart(h, "cg_scalar");
      hipLaunchKernelGGL(crt_cg_scalar_kernel, dim3(1), dim3(1024), 0, h->stream, G, 0, it);
      if (timed) tstop(h);
      if (timed) tstart(h, "cg_bap");
      hipLaunchKernelGGL(crt_cg_bap_kernel, dim3((N + 255) / 256, G.nslice), dim3(256), 0, h->stream, G);
      hipLaunchKernelGGL(crt_cg_bap_sum_kernel, dim3((N + 255) / 256), dim3(256), 0, h->stream, G);
      if (timed) tstop(h);
      if (timed) tstart(h, "cg_bb");
      hipLaunchKernelGGL(crt_cg_bb_kernel, dim3((unsigned)(((size_t)M * 64 + 255) / 256)), dim3(256), 0,
                         h->stream, G);
      if (timed) tstop(h);
      if (timed) tstart(h, "cg_scalar");
      hipLaunchKernelGGL(crt_cg_scalar_kernel, dim3(1), dim3(1024), 0, h->stream, G, 1, it);
      if (timed) tstop(h);
    }
  };
  if (h->use_graph) {
    std::vector<unsigned char> key(sizeof(CgArgs) + sizeof(int));
    std::memcpy(key.data(), &G, sizeof(CgArgs));
    std::memcpy(key.data() + sizeof(CgArgs), &ncgmax, sizeof(int));
    if (!h->cg_exec || key != h->cg_key) {
      if (h->cg_exec) hipGraphExecDestroy(h->cg_exec);
      if (h->cg_graph) hipGraphDestroy(h->cg_graph);
      h->cg_exec = nullptr;
      h->cg_graph = nullptr;
      HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      iterations(false);
      HIPCHK(hipStreamEndCapture(h->stream, &h->cg_graph));
      HIPCHK(hipGraphInstantiate(&h->cg_exec, h->cg_graph, nullptr, nullptr, 0));
      h->cg_key = key;
    }
    tstart(h, "cg_graph");
    HIPCHK(hipGraphLaunch(h->cg_exec, h->stream));
    tstop(h);
  } else {
    iterations(true);
  }
  rc = finish(h);
  if (rc) return rc;
  std::vector<double> cf(M), sc(8);
  std::vector<double> dp(2 * (size_t)M);
  HIPCHK(d2h(h, cf.data(), G.cgfac, M * 8));
  HIPCHK(d2h(h, dp.data(), G.dpar, (size_t)M * 16));
  HIPCHK(d2h(h, sc.data(), G.scal, 64));
  if (cgres) {
    HIPCHK(d2h(h, cgres, G.cgres, (ncgmax + 2) * sizeof(double)));
    cgres[0] = sc[6];
  }
  // dpar = dpar * cgfac (src/update.f90:307)
  for (int j = 0; j < M; ++j) {
    dpar[2 * j] = dp[2 * j] * cf[j];
    dpar[2 * j + 1] = dp[2 * j + 1] * cf[j];
  }
  if (cgfac) std::copy(cf.begin(), cf.end(), cgfac);
  return 0;
}

// ---- smoothness operator (bsmatm), host side, once per inversion ----------
// ltri = 0: the rectangular 5-point stencil of bsmatmreg
// (src/bsmatm_mod.f90:190-362).  Elements are numbered row by row from the
// surface (element (i, j) = (i-1) nx + j), corners in the order lower-left,
// lower-right, upper-right, upper-left.  smatm(:,1) is the diagonal,
// smatm(:,2) the coupling to the right, smatm(:,3) the one downwards;
// horizontal weights alfx * dz / |dx of the centroids|, vertical ones
// alfz * dx / |dz of the centroids|.
static int smooth_rect(crt_handle* h, int nx, int nz, double alfx, double alfz, double* sm) {
  const int manz = h->elanz;
  if (nx < 1 || nz < 1 || nx * nz != manz) return kErrArg;
  for (int t = 0; t < h->typanz; ++t)
    if (h->typ[t] <= 10 && h->selanz[t] != 4) return kErrArg;
  std::fill(sm, sm + (size_t)manz * 3, 0.0);
  auto X = [&](int e, int c) { return node_x(h, h->nrel[e + (size_t)c * h->nel]); };
  auto Y = [&](int e, int c) { return node_y(h, h->nrel[e + (size_t)c * h->nel]); };
  auto cx = [&](int e) { double v = 0.0; for (int c = 0; c < 4; ++c) v = v + X(e, c); return v / 4.0; };
  auto cy = [&](int e) { double v = 0.0; for (int c = 0; c < 4; ++c) v = v + Y(e, c); return v / 4.0; };
  for (int i = 1; i <= nz; ++i)
    for (int j = 1; j <= nx; ++j) {
      const int e = (i - 1) * nx + j - 1;
      const double xm = cx(e), zm = cy(e);
      if (j > 1) sm[e] = 1.0 * alfx * std::fabs(Y(e, 3) - Y(e, 0)) / std::fabs(xm - cx(e - 1));
      if (j < nx) {
        const double w = 1.0 * alfx * std::fabs(Y(e, 2) - Y(e, 1)) / std::fabs(cx(e + 1) - xm);
        sm[e] = sm[e] + w;
        sm[e + (size_t)manz] = -w;
      }
      if (i > 1) sm[e] = sm[e] + 1.0 * alfz * std::fabs(X(e, 2) - X(e, 3)) / std::fabs(cy(e - nx) - zm);
      if (i < nz) {
        const double w = 1.0 * alfz * std::fabs(X(e, 1) - X(e, 0)) / std::fabs(zm - cy(e + nx));
        sm[e] = sm[e] + w;
        sm[e + 2 * (size_t)manz] = -w;
      }
    }
  return 0;
}

// Edge neighbours of the area elements (bnachbar, src/bnachbar.f90:25-66):
// nachbar(i,k) = the element sharing edge k (nodes k and k+1 mod maxnodes)
// of element i, the highest-numbered one if several do, and nachbar(i,
// maxnodes+1) the number of matches; found through a map of edges instead of
// the reference's all-pairs scan.
static void neighbours(crt_handle* h, std::vector<int>& nb) {
  const int manz = h->elanz, mx = h->maxnodes;
  nb.assign((size_t)manz * (mx + 1), 0);
  auto node = [&](int e, int c) { return h->nrel[e + (size_t)c * h->nel]; };
  std::map<std::pair<int, int>, std::vector<int>> edges;
  for (int j = 0; j < manz; ++j)
    for (int c = 0; c < mx; ++c) {
      const int a = node(j, c), b = node(j, (c + 1) % mx);
      edges[{std::min(a, b), std::max(a, b)}].push_back(j);
    }
  for (int i = 0; i < manz; ++i)
    for (int c = 0; c < mx; ++c) {
      const int a = node(i, c), b = node(i, (c + 1) % mx);
      for (int j : edges[{std::min(a, b), std::max(a, b)}]) {
        if (j == i) continue;
        nb[i + (size_t)manz * c] = j + 1;
        nb[i + (size_t)manz * mx] += 1;
      }
    }
}

// ltri = 1, 2: bsmatmtri (src/bsmatm_mod.f90:364-461).  Across edge k of
// element i with neighbour n: w = edge length / centroid distance *
// sqrt((alfx cos a)^2 + (alfz sin a)^2), a the direction of the centroid
// difference, times the factor of every decoupling record that names the pair
// (i, n) in either order (:434-445, applied in record order);
// smatm(i,k) = -w, smatm(i,maxnodes+1) += w.
static void smooth_tri(crt_handle* h, double alfx, double alfz, const std::vector<int>& nb, double* sm) {
  const int manz = h->elanz, mx = h->maxnodes;
  std::vector<double> ex(manz, 0.0), ey(manz, 0.0);
  int iel = 0;
  for (int t = 0; t < h->typanz; ++t)
    for (int j = 0; j < h->nelanz[t]; ++j, ++iel) {
      if (h->typ[t] > 10) continue;
      const int nk = h->selanz[t];
      for (int c = 0; c < nk; ++c) {
        ex[iel] = ex[iel] + node_x(h, h->nrel[iel + (size_t)c * h->nel]);
        ey[iel] = ey[iel] + node_y(h, h->nrel[iel + (size_t)c * h->nel]);
      }
      ex[iel] = ex[iel] / nk;
      ey[iel] = ey[iel] / nk;
    }
  std::fill(sm, sm + (size_t)manz * (mx + 1), 0.0);
  for (int i = 0; i < manz; ++i)
    for (int c = 0; c < mx; ++c) {
      const int a = h->nrel[i + (size_t)c * h->nel], b = h->nrel[i + (size_t)((c + 1) % mx) * h->nel];
      const double dx = node_x(h, a) - node_x(h, b), dy = node_y(h, a) - node_y(h, b);
      const double edglen = std::sqrt(dx * dx + dy * dy);
      const int n = nb[i + (size_t)manz * c];
      if (n <= 0) continue;
      const double px = ex[i] - ex[n - 1], py = ey[i] - ey[n - 1];
      const double dist = std::sqrt(px * px + py * py);
      const double ang = std::atan2(py, px);
      const double ax = alfx * std::cos(ang), az = alfz * std::sin(ang);
      double w = edglen / dist * std::sqrt(ax * ax + az * az);
      for (size_t d = 0; d < h->dec_pairs.size(); ++d) {
        const int2 p = h->dec_pairs[d];
        if ((p.x == i + 1 && p.y == n) || (p.x == n && p.y == i + 1)) w = w * h->dec_str[d];
      }
      sm[i + (size_t)manz * c] = -w;
      sm[i + (size_t)manz * mx] = sm[i + (size_t)manz * mx] + w;
    }
}

int crt_smoothness(crt_handle* h, int32_t ltri, double alfx, double alfz, int32_t nx, int32_t nz,
                   double* smatm, int32_t* nachbar) {
  if (!h || !smatm) return kErrArg;
  if (ltri == 0) return smooth_rect(h, nx, nz, alfx, alfz, smatm);
  if (ltri != 1 && ltri != 2) return kErrArg;
  std::vector<int> nb;
  neighbours(h, nb);
  smooth_tri(h, alfx, alfz, nb, smatm);
  if (nachbar) std::copy(nb.begin(), nb.end(), nachbar);
  return 0;
}

int crt_set_decouplings(crt_handle* h, int32_t decanz, const int32_t* edecoup, const double* edecstr) {
  if (!h || decanz < 0 || (decanz > 0 && (!edecoup || !edecstr))) return kErrArg;
  std::vector<int2> pairs((size_t)decanz);
  for (int d = 0; d < decanz; ++d) {
    const int a = edecoup[d], b = edecoup[d + (size_t)decanz];
    if (a < 1 || a > h->elanz || b < 1 || b > h->elanz) return kErrArg;
    pairs[d] = make_int2(a, b);
  }
  h->dec_pairs = std::move(pairs);
  h->dec_str.assign(edecstr, edecstr + decanz);
  return 0;
}

int crt_set_device_budget(crt_handle* h, uint64_t bytes) {
  if (!h || bytes < ((uint64_t)1 << 20)) return kErrArg;
  h->batch_bytes_max = (size_t)bytes;
  return 0;
}

int crt_set_graph(crt_handle* h, int32_t enable) {
  if (!h) return kErrArg;
  h->use_graph = enable != 0;
  return 0;
}

int crt_error_info(crt_handle* h, int32_t* code, int32_t* k, int32_t* l) {
  if (!h) return kErrArg;
  if (code) *code = h->last_err;
  if (k) *k = h->last_k;
  if (l) *l = h->last_l;
  return 0;
}

int crt_set_timing(crt_handle* h, int32_t enable) {
  if (!h) return kErrArg;
  h->timing = enable != 0;
  h->times.clear();
  return 0;
}

int crt_stage_times(crt_handle* h, int32_t n_max, char* names, double* ms, int32_t* launches) {
  if (!h) return -kErrArg;
  int i = 0;
  for (auto& kv : h->times) {
    if (i >= n_max) break;
    if (names) {
      std::strncpy(names + 48 * i, kv.first.c_str(), 47);
      names[48 * i + 47] = 0;
    }
    if (ms) ms[i] = kv.second.first;
    if (launches) launches[i] = kv.second.second;
    ++i;
  }
  return i;
}

int crt_factor_phase_ticks(crt_handle* h, int32_t enable, uint64_t* ticks8) {
  if (!h) return kErrArg;
  if (ticks8)
    for (int i = 0; i < 8; ++i) ticks8[i] = h->kprof_ticks[i];
  h->kprof = enable != 0;
  for (int i = 0; i < 8; ++i) h->kprof_ticks[i] = 0;
  return 0;
}

int crt_kpot_device(crt_handle* h, const double** kpot_dev) {
  if (!h || !kpot_dev) return kErrArg;
  if (!h->forward_done || !h->kpot_full) return kErrOrder;
  *kpot_dev = reinterpret_cast<const double*>(h->at<double2>("kpot"));
  return 0;
}

}  // extern "C"

// ---- inversion-loop pieces (src/dmisft.f90, src/bpar.f90, src/bsigma.f90,
// src/update.f90:332-351) ----

int crt_dmisft_dev(crt_handle* h, int32_t nanz, const double* dat_dev, const double* sigmaa_dev, double* wmatd_dev,
                   const double* wmatdr_dev, const double* wmatdp_dev, const double* d0_dev, const double* fm0_dev,
                   int32_t flags, double l1min, double* stats_dev, int32_t* wdlok_dev) {
  if (!h || nanz <= 0 || !dat_dev || !sigmaa_dev || !wmatd_dev || !wmatdr_dev || !wmatdp_dev || !stats_dev ||
      !wdlok_dev || (flags & ~15))
    return kErrArg;
  if ((flags & kMisfitLdiff) && (!d0_dev || !fm0_dev)) return kErrArg;
  MisfitArgs A;
  A.nanz = nanz;
  A.flags = flags;
  A.l1min = l1min;
  A.dat = reinterpret_cast<const double2*>(dat_dev);
  A.sigmaa = reinterpret_cast<const double2*>(sigmaa_dev);
  A.wmatd = wmatd_dev;
  A.wmatdr = wmatdr_dev;
  A.wmatdp = wmatdp_dev;
  A.d0 = reinterpret_cast<const double2*>(d0_dev);
  A.fm0 = reinterpret_cast<const double2*>(fm0_dev);
  A.psi = h->get<double>("mis_psi", nanz);
  NEED(A.psi);
  A.wdlok = wdlok_dev;
  A.stats = stats_dev;
  tstart(h, "misfit");
  hipLaunchKernelGGL(crt_misfit_kernel, dim3(1), dim3(kInvThreads), 0, h->stream, A);
  tstop(h);
  return hipGetLastError() == hipSuccess ? 0 : kErrHip;
}

int crt_dmisft_dev_noalloc(crt_handle* h, int32_t nanz, const double* dat_dev, const double* sigmaa_dev,
                           double* wmatd_dev, const double* wmatdr_dev, const double* wmatdp_dev,
                           const double* d0_dev, const double* fm0_dev, int32_t flags, double l1min,
                           double* stats_dev, int32_t* wdlok_dev) {
  if (!h || nanz <= 0 || !dat_dev || !sigmaa_dev || !wmatd_dev || !wmatdr_dev || !wmatdp_dev || !stats_dev ||
      !wdlok_dev || (flags & ~15))
    return kErrArg;
  if ((flags & kMisfitLdiff) && (!d0_dev || !fm0_dev)) return kErrArg;
  if (nanz > kNoAllocMaxData) return kErrNoAllocSize;
  double* psi = (h->dev >= 0 && h->dev < kMaxDevices) ? noalloc_ptr[h->dev] : nullptr;
  if (!psi) return kErrHip;
  MisfitArgs A;
  A.nanz = nanz;
  A.flags = flags;
  A.l1min = l1min;
  A.dat = reinterpret_cast<const double2*>(dat_dev);
  A.sigmaa = reinterpret_cast<const double2*>(sigmaa_dev);
  A.wmatd = wmatd_dev;
  A.wmatdr = wmatdr_dev;
  A.wmatdp = wmatdp_dev;
  A.d0 = reinterpret_cast<const double2*>(d0_dev);
  A.fm0 = reinterpret_cast<const double2*>(fm0_dev);
  // every psi[i] is written before it is read, by the same thread: nothing
  // of the previous call's contents is used
  A.psi = psi;
  A.wdlok = wdlok_dev;
  A.stats = stats_dev;
  hipLaunchKernelGGL(crt_misfit_kernel, dim3(1), dim3(kInvThreads), 0, h->stream, A);
  return hipGetLastError() == hipSuccess ? 0 : kErrHip;
}

int crt_dmisft(crt_handle* h, int32_t nanz, const double* dat, const double* sigmaa, double* wmatd,
               const double* wmatdr, const double* wmatdp, const double* d0, const double* fm0, int32_t flags,
               double l1min, double* stats, int32_t* wdlok) {
  if (!h || nanz <= 0 || !dat || !sigmaa || !wmatd || !wmatdr || !wmatdp || !stats || (flags & ~15))
    return kErrArg;
  const bool ldiff = flags & kMisfitLdiff;
  if (ldiff && (!d0 || !fm0)) return kErrArg;
  double2 *dd, *ds, *d0d = nullptr, *f0d = nullptr;
  double *dw, *dwr, *dwp;
  int rc = upload(h, "mis_dat", reinterpret_cast<const double2*>(dat), nanz, &dd);
  rc |= upload(h, "mis_sig", reinterpret_cast<const double2*>(sigmaa), nanz, &ds);
  rc |= upload(h, "mis_w", wmatd, nanz, &dw);
  rc |= upload(h, "mis_wr", wmatdr, nanz, &dwr);
  rc |= upload(h, "mis_wp", wmatdp, nanz, &dwp);
  if (ldiff) {
    rc |= upload(h, "mis_d0", reinterpret_cast<const double2*>(d0), nanz, &d0d);
    rc |= upload(h, "mis_fm0", reinterpret_cast<const double2*>(fm0), nanz, &f0d);
  }
  if (rc) return kErrHip;
  int* dwd = h->get<int>("mis_wdlok", nanz);
  double* dst = h->get<double>("mis_stats", 6);
  NEED(dwd);
  NEED(dst);
  rc = crt_dmisft_dev(h, nanz, reinterpret_cast<const double*>(dd), reinterpret_cast<const double*>(ds), dw, dwr,
                      dwp, reinterpret_cast<const double*>(d0d), reinterpret_cast<const double*>(f0d), flags, l1min,
                      dst, dwd);
  if (rc) return rc;
  rc = finish(h);
  if (rc) return rc;
  double st[6];
  HIPCHK(d2h(h, st, dst, sizeof(st)));
  if (st[5] == (double)nanz) return 99;  // dmisft :125-130: every datum excluded
  std::memcpy(stats, st, sizeof(st));
  HIPCHK(d2h(h, wmatd, dw, (size_t)nanz * 8));
  if (wdlok) HIPCHK(d2h(h, wdlok, dwd, (size_t)nanz * 4));
  return 0;
}

int crt_set_model_map(crt_handle* h, int32_t manz, const int32_t* mnr) {
  if (!h) return kErrArg;
  const int el = h->elanz;
  if (!mnr) {
    if (manz != el && manz != 0) return kErrArg;
    h->manz_map = el;
    h->mnr_host.clear();
    return 0;
  }
  if (manz <= 0) return kErrArg;
  std::vector<int> last(manz, -1);
  for (int i = 0; i < el; ++i) {
    if (mnr[i] < 1 || mnr[i] > manz) return kErrArg;
    last[mnr[i] - 1] = i;  // bpar's element loop: the last element wins
  }
  int* dm;
  int* dl;
  if (upload(h, "model_mnr", mnr, el, &dm) || upload(h, "model_last", last.data(), manz, &dl)) return kErrHip;
  HIPCHK(hipStreamSynchronize(h->stream));  // the host vector above goes out of scope
  h->manz_map = manz;
  h->mnr_host.assign(mnr, mnr + el);
  return 0;
}

static int model_manz(crt_handle* h) { return h->mnr_host.empty() ? h->elanz : h->manz_map; }

int crt_bpar_dev(crt_handle* h, const double* sigma_dev, double* par_dev) {
  if (!h || !sigma_dev || !par_dev) return kErrArg;
  const int manz = model_manz(h);
  const int* last = h->mnr_host.empty() ? nullptr : h->at<int>("model_last");
  hipLaunchKernelGGL(crt_bpar_kernel, dim3((manz + 255) / 256), dim3(256), 0, h->stream, manz, last,
                     reinterpret_cast<const double2*>(sigma_dev), reinterpret_cast<double2*>(par_dev));
  return hipGetLastError() == hipSuccess ? 0 : kErrHip;
}

int crt_bsigma_dev(crt_handle* h, const double* par_dev, double* sigma_dev) {
  if (!h || !par_dev || !sigma_dev) return kErrArg;
  const int* mnr = h->mnr_host.empty() ? nullptr : h->at<int>("model_mnr");
  hipLaunchKernelGGL(crt_bsigma_kernel, dim3((h->elanz + 255) / 256), dim3(256), 0, h->stream, h->elanz, mnr,
                     reinterpret_cast<const double2*>(par_dev), reinterpret_cast<double2*>(sigma_dev));
  return hipGetLastError() == hipSuccess ? 0 : kErrHip;
}

int crt_bpar(crt_handle* h, const double* sigma, double* par) {
  if (!h || !sigma || !par) return kErrArg;
  const int manz = model_manz(h);
  double2 *ds, *dp;
  if (upload(h, "bp_sigma", reinterpret_cast<const double2*>(sigma), h->elanz, &ds) ||
      upload(h, "bp_par", reinterpret_cast<const double2*>(par), manz, &dp))
    return kErrHip;
  int rc = crt_bpar_dev(h, reinterpret_cast<const double*>(ds), reinterpret_cast<double*>(dp));
  if (rc) return rc;
  rc = finish(h);
  if (rc) return rc;
  HIPCHK(d2h(h, par, dp, (size_t)manz * 16));
  return 0;
}

int crt_bsigma(crt_handle* h, const double* par, double* sigma) {
  if (!h || !par || !sigma) return kErrArg;
  const int manz = model_manz(h);
  double2* dp;
  if (upload(h, "bp_par", reinterpret_cast<const double2*>(par), manz, &dp)) return kErrHip;
  double2* ds = h->get<double2>("bp_sigma", h->elanz);
  NEED(ds);
  int rc = crt_bsigma_dev(h, reinterpret_cast<const double*>(dp), reinterpret_cast<double*>(ds));
  if (rc) return rc;
  rc = finish(h);
  if (rc) return rc;
  HIPCHK(d2h(h, sigma, ds, (size_t)h->elanz * 16));
  return 0;
}

int crt_update_step_dev(crt_handle* h, int32_t manz, double* par_dev, const double* dpar_dev, double step,
                        int32_t lphi0, double* out_dev) {
  if (!h || manz <= 0 || !par_dev || !dpar_dev || !out_dev) return kErrArg;
  hipLaunchKernelGGL(crt_update_step_kernel, dim3(1), dim3(kInvThreads), 0, h->stream, manz,
                     reinterpret_cast<double2*>(par_dev), reinterpret_cast<const double2*>(dpar_dev), step,
                     lphi0 ? 1 : 0, out_dev);
  return hipGetLastError() == hipSuccess ? 0 : kErrHip;
}

int crt_update_step(crt_handle* h, int32_t manz, double* par, const double* dpar, double step, int32_t lphi0,
                    double* out) {
  if (!h || manz <= 0 || !par || !dpar || !out) return kErrArg;
  double2 *dp, *dd;
  if (upload(h, "up_par", reinterpret_cast<const double2*>(par), manz, &dp) ||
      upload(h, "up_dpar", reinterpret_cast<const double2*>(dpar), manz, &dd))
    return kErrHip;
  double* dout = h->get<double>("up_out", 2);
  NEED(dout);
  int rc = crt_update_step_dev(h, manz, reinterpret_cast<double*>(dp), reinterpret_cast<const double*>(dd), step,
                               lphi0, dout);
  if (rc) return rc;
  rc = finish(h);
  if (rc) return rc;
  HIPCHK(d2h(h, par, dp, (size_t)manz * 16));
  HIPCHK(d2h(h, out, dout, 16));
  return 0;
}

// blam0 (src/blam0.f90:69-170) in O(N M): jtj_j = |sum_i conj(J_ij) w_i
// sum_k J_ik| = |(J^H W (J 1))_j| (mode 1: real parts of J, the FPI and DC
// branches), lammax = mean_j jtj_j * 2 / (alfx + alfz) * 5.  The reference's
// triple loop forms the same sums in O(N M^2) and another order.
int crt_blam0(crt_handle* h, const double* w, int32_t mode, double alfx, double alfz, double* lammax) {
  if (!h || !w || !lammax || (mode != 0 && mode != 1) || !(alfx + alfz != 0.0)) return kErrArg;
  if (!h->sens_done) return kErrOrder;
  const int N = h->nanz, M = h->elanz;
  double* dw;
  if (upload(h, "bl_w", w, N, &dw)) return kErrHip;
  double2* part = h->get<double2>("bl_part", (size_t)kRowSlices * N);
  double2* s = h->get<double2>("bl_s", N);
  double2* t = h->get<double2>("bl_t", M);
  NEED(part);
  NEED(s);
  NEED(t);
  const double2* J = h->at<double2>("sens");
  tstart(h, "blam0");
  hipLaunchKernelGGL(crt_rowsum_partial_kernel, dim3((N + 255) / 256, kRowSlices), dim3(256), 0, h->stream, J, N, M,
                     (int)mode, part);
  hipLaunchKernelGGL(crt_rowsum_final_kernel, dim3((N + 255) / 256), dim3(256), 0, h->stream, (const double2*)part,
                     N, s);
  hipLaunchKernelGGL(crt_jhwr_kernel, dim3((unsigned)(((size_t)M * 64 + 255) / 256)), dim3(256), 0, h->stream, J, N,
                     M, (const double*)dw, (const double2*)s, t);
  tstop(h);
  int rc = finish(h);
  if (rc) return rc;
  std::vector<double2> th(M);
  HIPCHK(d2h(h, th.data(), t, (size_t)M * 16));
  double sum = 0.0;
  for (int j = 0; j < M; ++j) sum += mode ? std::fabs(th[j].x) : std::hypot(th[j].x, th[j].y);
  *lammax = sum / (double)M * 2.0 / (alfx + alfz) * 5.0;
  return 0;
}
