// Host runtime of the MI355X ERI library: context set-up (normalisation,
// Boys table, shell pairs, Schwarz screening on the device, quartet lists)
// and the C-ABI entry points declared in include/quax_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "quax_amd.h"
#include "qx_runtime.h"
#include "qx_classes.inc"  // generated; in the build's generated-source directory

static_assert(QX_GEN_LMAX == QX_LMAX_SUPPORTED, "generated kernels and runtime disagree on lmax");
static_assert(QX_NUM_CLASSES == QX_NUM_CLASSES_MAX, "profile slot layout");

namespace {

struct QxError {
  int code;
  std::string msg;
};

#define QX_HIP(call)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      throw QxError{e_ == hipErrorOutOfMemory ? QX_ENOMEM : QX_EHIP,                            \
                    std::string(#call) + ": " + hipGetErrorString(e_)};                         \
  } while (0)

#define QX_REQUIRE(cond, code, msg) \
  do {                              \
    if (!(cond)) throw QxError{code, msg}; \
  } while (0)

int ncart(int l) { return (l + 1) * (l + 2) / 2; }
int nsph(int l, int pure) { return (pure && l >= 2) ? 2 * l + 1 : ncart(l); }
int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }

template <class T>
T* dalloc(qx_ctx* c, size_t n) {
  void* p = nullptr;
  QX_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  c->allocs.push_back(p);
  return static_cast<T*>(p);
}

template <class T>
T* dupload(qx_ctx* c, const std::vector<T>& v) {
  T* p = dalloc<T>(c, v.size());
  if (!v.empty()) QX_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

// scoped temporary device buffer
struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { QX_HIP(hipMalloc(&p, std::max<size_t>(bytes, 8))); }
  ~DevBuf() { if (p) (void)hipFree(p); }
  double* d() const { return static_cast<double*>(p); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

double dfact(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// Libint Shell::renorm convention: x^l normalisation of each primitive,
// then the contraction rescaled to unit self-overlap.
void normalize_shell(int l, int np, const double* a, const double* cin, double* cout) {
  const double pi32 = std::pow(M_PI, 1.5);
  for (int p = 0; p < np; ++p) {
    const double two_a = 2.0 * a[p];
    cout[p] = cin[p] * std::sqrt(std::pow(2.0, l) * std::pow(two_a, l + 1.5) / (pi32 * dfact(2 * l - 1)));
  }
  double s = 0.0;
  for (int p = 0; p < np; ++p)
    for (int q = 0; q < np; ++q)
      s += dfact(2 * l - 1) * pi32 * cout[p] * cout[q] / (std::pow(2.0, l) * std::pow(a[p] + a[q], l + 1.5));
  const double f = 1.0 / std::sqrt(s);
  for (int p = 0; p < np; ++p) cout[p] *= f;
}

// Boys table rows F_0..F_{NM-1}(g/INV_H): series at the top order in long
// double, F_m(T) = exp(-T) sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)),
// then downward recursion F_{m-1} = (2T F_m + exp(-T)) / (2m-1).
std::vector<double> boys_table() {
  std::vector<double> tab((size_t)QX_BOYS_NPTS * QX_BOYS_NM);
  for (int g = 0; g < QX_BOYS_NPTS; ++g) {
    const long double T = g / (long double)QX_BOYS_INV_H;
    const int M = QX_BOYS_NM - 1;
    long double term = 1.0L / (2 * M + 1), sum = term;
    for (int k = 1; k < 4000; ++k) {
      term *= 2.0L * T / (2 * M + 2 * k + 1);
      sum += term;
      if (term < 1e-26L * sum) break;
    }
    const long double et = expl(-T);
    long double f = et * sum;
    tab[(size_t)g * QX_BOYS_NM + M] = (double)f;
    for (int m = M; m > 0; --m) {
      f = (2.0L * T * f + et) / (2 * m - 1);
      tab[(size_t)g * QX_BOYS_NM + m - 1] = (double)f;
    }
  }
  return tab;
}

// real solid harmonics for l = 2 in terms of x^l-normalised Cartesians
// (xx, xy, xz, yy, yz, zz), rows m = -2..2, unit-normalised.
const double kSph2[5 * 6] = {
    0, 1.7320508075688772, 0, 0, 0, 0,                   // m=-2: sqrt3 xy
    0, 0, 0, 0, 1.7320508075688772, 0,                   // m=-1: sqrt3 yz
    -0.5, 0, 0, -0.5, 0, 1.0,                            // m= 0: z^2 - (x^2+y^2)/2
    0, 0, 1.7320508075688772, 0, 0, 0,                   // m=+1: sqrt3 xz
    0.8660254037844386, 0, 0, -0.8660254037844386, 0, 0  // m=+2: sqrt3/2 (x^2-y^2)
};

struct PairRec {
  int a, b;        // oriented shells (l[a] >= l[b])
  int s1, s2;      // reference order (s1 >= s2)
  int cls;
  double Q;
};

void build_pair_arrays(qx_ctx* c, const std::vector<PairRec>& P, QxPairsDev& pr, QxPrimPairsDev& pp, int& npseudo,
                       std::vector<int>* h_first, std::vector<int>* h_n) {
  const int np = (int)P.size();
  std::vector<double> Ax(np), Ay(np), Az(np), ABx(np), ABy(np), ABz(np);
  std::vector<int> sa(np), sb(np), bfa(np), bfb(np), ata(np), atb(np), first(np), cnt(np);
  std::vector<int> ppair;
  std::vector<double> p_, Px, Py, Pz, K, a2, b2;
  for (int i = 0; i < np; ++i) {
    const int a = P[i].a, b = P[i].b;
    const double* A = &c->xyz[3 * c->atom[a]];
    const double* B = &c->xyz[3 * c->atom[b]];
    Ax[i] = A[0]; Ay[i] = A[1]; Az[i] = A[2];
    ABx[i] = A[0] - B[0]; ABy[i] = A[1] - B[1]; ABz[i] = A[2] - B[2];
    sa[i] = a; sb[i] = b;
    bfa[i] = c->bfc[a]; bfb[i] = c->bfc[b];
    ata[i] = c->atom[a]; atb[i] = c->atom[b];
    first[i] = (int)ppair.size();
    const double AB2 = ABx[i] * ABx[i] + ABy[i] * ABy[i] + ABz[i] * ABz[i];
    for (int ia = c->poff[a]; ia < c->poff[a] + c->nprim[a]; ++ia)
      for (int ib = c->poff[b]; ib < c->poff[b] + c->nprim[b]; ++ib) {
        const double al = c->exps[ia], be = c->exps[ib], p = al + be;
        ppair.push_back(i);
        p_.push_back(p);
        Px.push_back((al * A[0] + be * B[0]) / p);
        Py.push_back((al * A[1] + be * B[1]) / p);
        Pz.push_back((al * A[2] + be * B[2]) / p);
        K.push_back(c->coefs[ia] * c->coefs[ib] * std::exp(-al * be / p * AB2));
        a2.push_back(2 * al);
        b2.push_back(2 * be);
      }
    cnt[i] = (int)ppair.size() - first[i];
  }
  npseudo = (int)ppair.size();
  pr.Ax = dupload(c, Ax); pr.Ay = dupload(c, Ay); pr.Az = dupload(c, Az);
  pr.ABx = dupload(c, ABx); pr.ABy = dupload(c, ABy); pr.ABz = dupload(c, ABz);
  pr.sa = dupload(c, sa); pr.sb = dupload(c, sb);
  pr.bfa = dupload(c, bfa); pr.bfb = dupload(c, bfb);
  pr.atom_a = dupload(c, ata); pr.atom_b = dupload(c, atb);
  pr.pp_first = dupload(c, first); pr.pp_n = dupload(c, cnt);
  pp.pair = dupload(c, ppair);
  pp.p = dupload(c, p_); pp.Px = dupload(c, Px); pp.Py = dupload(c, Py); pp.Pz = dupload(c, Pz);
  pp.K = dupload(c, K); pp.a2 = dupload(c, a2); pp.b2 = dupload(c, b2);
  if (h_first) *h_first = first;
  if (h_n) *h_n = cnt;
}

hipEvent_t take_event(qx_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  QX_HIP(hipEventCreate(&e));
  return e;
}

// Runs one kernel launch, bracketed by HIP events on its stream when
// profiling is on.
template <class F>
int timed_launch(qx_ctx* c, int slot, long long pq, hipStream_t s, F&& launch) {
  if (!c->profiling || pq <= 0) return launch();
  hipEvent_t a = take_event(c), b = take_event(c);
  QX_HIP(hipEventRecord(a, s));
  const int rc = launch();
  QX_HIP(hipEventRecord(b, s));
  c->prof_pending.push_back({slot, a, b});
  c->prof_launches[slot] += 1;
  c->prof_pq[slot] += pq;
  return rc;
}

int launch_eri(qx_ctx* c, const QxWorkDev& w, int cls, const QxEpiDev& e, int fp32, hipStream_t s,
               const QxPairsDev& pr, const QxPrimPairsDev& pp) {
  QxArgs a;
  a.pr = pr; a.pp = pp; a.w = w; a.e = e; a.boys = c->d_boys;
  QxLaunch L{fp32, s};
  const QxClassEntry& ent = qx_class_table[cls];
  return timed_launch(c, QX_SLOT_ERI(cls, fp32 ? 1 : 0), w.nquart, s, [&] { return ent.eri(&a, &L); });
}

// One kernel launch of a pass, with its estimated cost (flops).
struct Launch {
  double cost;
  std::function<int(hipStream_t)> go;
};

void ensure_side_streams(qx_ctx* c) {
  const int want = std::max(0, c->nstreams - 1);
  while ((int)c->side.size() < want) {
    hipStream_t t;
    QX_HIP(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
    c->side.push_back(t);
  }
  while ((int)c->fork_ev.size() < want + 1) {
    hipEvent_t e;
    QX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->fork_ev.push_back(e);
  }
}

void check_launch(int rc, const char* what);

// Runs the independent launches of one pass.  With several streams the
// costliest go first, each to the least-loaded stream; the side streams wait
// for the caller's stream before starting and the caller's stream waits for
// all of them at the end, so the pass is ordered on `s` like one kernel.
// With profiling on, each kernel's HIP events sit on the stream it runs on:
// concurrent kernels are timed in place (one stream times each alone).
//
// side = false keeps the pass on the caller's stream.  The passes whose
// kernels need more than ~6 KB of scratch per lane (derivative and
// second-derivative tensors, the tangent kernels: up to 8.3 KB) run there:
// each hardware queue grows its own scratch to the largest per-lane size it
// has run, and growing a side stream's queue past ~8 KB per lane after the
// process already holds scratch on its other queues has been refused by the
// HSA runtime (HSA_STATUS_ERROR_OUT_OF_RESOURCES).  The caller's queue has
// already run the largest kernels, the array-form ones of context set-up.
// The streaming J/K and adjoint-gradient kernels stay under 6 KB per lane.
void run_launches(qx_ctx* c, std::vector<Launch>& v, hipStream_t s, const char* what, bool side) {
  const int ns = side ? std::max(1, std::min<int>(c->nstreams, (int)v.size())) : 1;
  if (ns == 1) {
    for (auto& x : v) check_launch(x.go(s), what);
    return;
  }
  ensure_side_streams(c);
  std::stable_sort(v.begin(), v.end(), [](const Launch& a, const Launch& b) { return a.cost > b.cost; });
  QX_HIP(hipEventRecord(c->fork_ev[0], s));
  for (int k = 0; k < ns - 1; ++k) QX_HIP(hipStreamWaitEvent(c->side[k], c->fork_ev[0], 0));
  std::vector<double> load(ns, 0.0);
  int rc = 0;
  std::exception_ptr thrown;
  for (auto& x : v) {
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[k] += x.cost;
    try {
      rc = x.go(k == 0 ? s : c->side[k - 1]);
    } catch (...) {
      thrown = std::current_exception();
    }
    if (rc != 0 || thrown) break;
  }
  // join even after a failed or throwing launch, so the caller's stream
  // never runs ahead of work already queued on the side streams
  hipError_t join_err = hipSuccess;
  for (int k = 0; k < ns - 1; ++k) {
    hipError_t e = hipEventRecord(c->fork_ev[k + 1], c->side[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, c->fork_ev[k + 1], 0);
    if (e != hipSuccess && join_err == hipSuccess) join_err = e;
  }
  if (thrown) std::rethrow_exception(thrown);
  check_launch(rc, what);
  QX_HIP(join_err);
}

void check_launch(int rc, const char* what) {
  if (rc == -1) throw QxError{QX_EUNSUPPORTED, std::string(what) + ": kernel variant not built"};
  if (rc != 0) throw QxError{QX_EHIP, std::string(what) + ": " + hipGetErrorString((hipError_t)rc)};
}

int class_index(int X, int Y) { return X * (X + 1) / 2 + Y; }

// Positions [ss, se) of fragment f within the class list that starts at
// sorted pair ko and holds kn pairs (sorted by fragment, then Schwarz value).
void frag_segment(const qx_ctx* c, int ko, int kn, int f, int& ss, int& se) {
  const int* fr = c->h_pair_frag.data() + ko;
  ss = (int)(std::lower_bound(fr, fr + kn, f) - fr);
  se = (int)(std::upper_bound(fr, fr + kn, f) - fr);
}

// Admissible kets of bra pair i (position in class X) in class Y: positions
// [ss, ss + kc) of Y's list, i.e. the bra's fragment, Q_i Q_j >= bound (a
// prefix of the descending segment) and canonical j <= i when X == Y.
void ket_window(const qx_ctx* c, int X, int Y, int i, double bound, int& ss, int& kc) {
  const int bo = c->class_pair_off[X], ko = c->class_pair_off[Y], kn = c->class_pair_cnt[Y];
  int se;
  frag_segment(c, ko, kn, c->h_pair_frag[bo + i], ss, se);
  const double Qi = c->h_pair_Q[bo + i];
  int lo = ss, hi = se;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (Qi * c->h_pair_Q[ko + mid] >= bound) lo = mid + 1; else hi = mid;
  }
  kc = lo - ss;
  if (X == Y) kc = std::max(0, std::min(kc, i - ss + 1));
}

void schwarz_and_lists(qx_ctx* c) {
  const int ns = c->nshell;
  // ---- every shell pair, class-sorted, for the Schwarz pass
  std::vector<PairRec> all;
  for (int s1 = 0; s1 < ns; ++s1)
    for (int s2 = 0; s2 <= s1; ++s2) {
      if (c->frag[c->atom[s1]] != c->frag[c->atom[s2]]) continue;  // pairs never span fragments
      PairRec r;
      r.s1 = s1; r.s2 = s2;
      if (c->l[s1] >= c->l[s2]) { r.a = s1; r.b = s2; } else { r.a = s2; r.b = s1; }
      r.cls = pair_class(c->l[r.a], c->l[r.b]);
      r.Q = 0.0;
      all.push_back(r);
    }
  std::stable_sort(all.begin(), all.end(), [](const PairRec& x, const PairRec& y) { return x.cls < y.cls; });
  QxPairsDev pr_all{};
  QxPrimPairsDev pp_all{};
  int np_all = 0;
  const size_t alloc_mark = c->allocs.size();
  std::vector<int> first, cnt;
  build_pair_arrays(c, all, pr_all, pp_all, np_all, &first, &cnt);
  const int npair_all = (int)all.size();
  DevBuf diag((size_t)npair_all * QX_PAIR_STRIDE * sizeof(double));
  check_launch(qxk_fill(diag.d(), (long long)npair_all * QX_PAIR_STRIDE, 0.0, 0), "fill");
  // diagonal lists per pair class: bra = pseudo pairs of the class, kets =
  // the pseudo pairs of the same contracted pair
  const int nclass = pair_class(QX_LMAX_SUPPORTED, QX_LMAX_SUPPORTED) + 1;
  std::vector<std::vector<long long>> pre(nclass);
  std::vector<QxWorkDev> wl(nclass);
  std::vector<DevBuf*> bufs;
  for (int X = 0; X < nclass; ++X) {
    int lo = -1, hi = -1;
    for (int i = 0; i < npair_all; ++i)
      if (all[i].cls == X) { if (lo < 0) lo = i; hi = i + 1; }
    if (lo < 0) continue;
    const int u0 = first[lo], u1 = first[hi - 1] + cnt[hi - 1];
    std::vector<long long> p(u1 - u0 + 1, 0);
    int idx = 0;
    for (int i = lo; i < hi; ++i)
      for (int k = 0; k < cnt[i]; ++k, ++idx) p[idx + 1] = p[idx] + cnt[i];
    DevBuf* b = new DevBuf(p.size() * sizeof(long long));
    bufs.push_back(b);
    QX_HIP(hipMemcpy(b->p, p.data(), p.size() * sizeof(long long), hipMemcpyHostToDevice));
    QxWorkDev w{};
    w.bra_off = u0; w.nbra = u1 - u0; w.ket_off = 0; w.nket = 0;
    w.prefix = static_cast<const long long*>(b->p);
    w.nquart = p.back(); w.store_base = 0; w.diag = 1;
    QxEpiDev e{};
    e.epi = QX_EPI_SCHWARZ; e.out = diag.d(); e.nbf = c->ncart; e.natom = c->natom; e.coord = -1;
    int rc = launch_eri(c, w, class_index(X, X), e, 0, 0, pr_all, pp_all);
    if (rc != 0) {
      for (auto* q : bufs) delete q;
      check_launch(rc, "schwarz");
    }
  }
  hipError_t se = hipDeviceSynchronize();
  for (auto* q : bufs) delete q;
  QX_HIP(se);
  std::vector<double> hd((size_t)npair_all * QX_PAIR_STRIDE);
  QX_HIP(hipMemcpy(hd.data(), diag.p, hd.size() * sizeof(double), hipMemcpyDeviceToHost));
  c->Q.assign((size_t)ns * ns, 0.0);
  double maxq = 0.0;
  for (int i = 0; i < npair_all; ++i) {
    const int na = ncart(c->l[all[i].a]), nb = ncart(c->l[all[i].b]);
    double m = 0.0;
    for (int k = 0; k < na * nb; ++k) m = std::max(m, std::fabs(hd[(size_t)i * QX_PAIR_STRIDE + k]));
    all[i].Q = m;
    c->Q[(size_t)all[i].s1 * ns + all[i].s2] = c->Q[(size_t)all[i].s2 * ns + all[i].s1] = m;
    maxq = std::max(maxq, m);
  }
  // release the all-pairs arrays
  while (c->allocs.size() > alloc_mark) {
    (void)hipFree(c->allocs.back());
    c->allocs.pop_back();
  }
  // ---- screening (libint_interface.cc:215-225)
  const double tol2 = c->tol * c->tol;
  const double cut = tol2 / maxq;
  c->ref_pairs.clear();
  for (int s1 = 0; s1 < ns; ++s1)
    for (int s2 = 0; s2 <= s1; ++s2)
      if (c->Q[(size_t)s1 * ns + s2] >= cut) { c->ref_pairs.push_back(s1); c->ref_pairs.push_back(s2); }
  std::vector<PairRec> scr;
  for (auto& r : all)
    if (r.Q >= cut) scr.push_back(r);
  std::sort(scr.begin(), scr.end(), [c](const PairRec& x, const PairRec& y) {
    if (x.cls != y.cls) return x.cls < y.cls;
    const int fx = c->frag[c->atom[x.a]], fy = c->frag[c->atom[y.a]];
    if (fx != fy) return fx < fy;
    if (x.Q != y.Q) return x.Q > y.Q;
    if (x.s1 != y.s1) return x.s1 < y.s1;
    return x.s2 < y.s2;
  });
  build_pair_arrays(c, scr, c->pr, c->pp, c->npseudo, &c->h_pp_first, &c->h_pp_n);
  c->npairs = (int)scr.size();
  c->h_pair_sa.resize(c->npairs); c->h_pair_sb.resize(c->npairs);
  c->h_pair_class.resize(c->npairs); c->h_pair_Q.resize(c->npairs); c->h_pair_frag.resize(c->npairs);
  for (int i = 0; i < c->npairs; ++i) {
    c->h_pair_sa[i] = scr[i].a; c->h_pair_sb[i] = scr[i].b;
    c->h_pair_class[i] = scr[i].cls; c->h_pair_Q[i] = scr[i].Q;
    c->h_pair_frag[i] = c->frag[c->atom[scr[i].a]];
  }
  c->class_pair_off.assign(nclass, 0); c->class_pair_cnt.assign(nclass, 0);
  for (int i = c->npairs - 1; i >= 0; --i) c->class_pair_off[scr[i].cls] = i;
  for (int i = 0; i < c->npairs; ++i) c->class_pair_cnt[scr[i].cls]++;
  // ---- quartet lists
  c->work.clear();
  c->nq_class.assign(QX_NUM_CLASSES, 0);
  c->ndq_class.assign(QX_NUM_CLASSES, 0);
  // positions (within their class) of pairs whose two shells sit on one atom,
  // per (class, atom), to count single-atom quartets
  std::vector<std::vector<std::vector<int>>> intra(nclass, std::vector<std::vector<int>>(c->natom));
  for (int X = 0; X < nclass; ++X)
    for (int i = 0; i < c->class_pair_cnt[X]; ++i) {
      const int P = c->class_pair_off[X] + i;
      const int A = c->atom[c->h_pair_sa[P]];
      if (A == c->atom[c->h_pair_sb[P]]) intra[X][A].push_back(i);
    }
  std::vector<long long> allpre;
  std::vector<int> allkst;  // per bra pseudo pair: first ket (fragmented contexts)
  std::vector<long long> kst_off;
  c->nquart_total = 0; c->npq_total = 0; c->store_total = 0;
  for (int X = 0; X < nclass; ++X) {
    for (int Y = 0; Y <= X; ++Y) {
      const int bo = c->class_pair_off[X], bn = c->class_pair_cnt[X];
      const int ko = c->class_pair_off[Y], kn = c->class_pair_cnt[Y];
      if (bn == 0 || kn == 0) continue;
      // cumulative pseudo counts over the ket class
      std::vector<long long> kcum(kn + 1, 0);
      for (int j = 0; j < kn; ++j) kcum[j + 1] = kcum[j] + c->h_pp_n[ko + j];
      QxWorkHost wh{};
      wh.cls = class_index(X, Y);
      wh.bra_class = X; wh.ket_class = Y;
      wh.prefix_off = (long long)allpre.size();
      const int u0 = c->h_pp_first[bo];
      const int u1 = c->h_pp_first[bo + bn - 1] + c->h_pp_n[bo + bn - 1];
      allpre.push_back(0);
      kst_off.push_back((long long)allkst.size());
      long long run = 0, ncq = 0;
      for (int i = 0; i < bn; ++i) {
        // kets with Q_i Q_j >= tol^2 form a prefix of the descending list
        // (of the bra's fragment segment)
        int ss, kc;
        ket_window(c, X, Y, i, tol2, ss, kc);
        ncq += kc;
        {
          const int P = bo + i;
          const int A = c->atom[c->h_pair_sa[P]];
          long long single = 0;
          if (A == c->atom[c->h_pair_sb[P]]) {
            const auto& v = intra[Y][A];
            single = (std::lower_bound(v.begin(), v.end(), ss + kc) - v.begin()) -
                     (std::lower_bound(v.begin(), v.end(), ss) - v.begin());
          }
          c->ndq_class[class_index(X, Y)] += kc - single;
        }
        for (int k = 0; k < c->h_pp_n[bo + i]; ++k) {
          run += kcum[ss + kc] - kcum[ss];
          allpre.push_back(run);
          allkst.push_back((int)kcum[ss]);
        }
      }
      wh.w.bra_off = u0; wh.w.nbra = u1 - u0;
      wh.w.ket_off = c->h_pp_first[ko];
      wh.w.nket = (int)kcum[kn];
      wh.w.nquart = run;
      wh.w.diag = 0;
      wh.w.store_base = c->store_total;
      const int N = ncart(c->l[c->h_pair_sa[bo]]) * ncart(c->l[c->h_pair_sb[bo]]) *
                    ncart(c->l[c->h_pair_sa[ko]]) * ncart(c->l[c->h_pair_sb[ko]]);
      c->store_total += run * N;
      c->nquart_total += ncq;
      c->nq_class[class_index(X, Y)] += ncq;
      c->npq_total += run;
      if (run > 0) {
        c->work.push_back(wh);
      } else {
        allpre.resize(wh.prefix_off);
        allkst.resize(kst_off.back());
        kst_off.pop_back();
      }
    }
  }
  c->d_prefix = dupload(c, allpre);
  if (c->fragmented) {
    c->d_kstart = dupload(c, allkst);
    c->d_pair_frag = dupload(c, c->h_pair_frag);
    // Cartesian function range of each fragment id (all kets of a bra pair
    // lie in its fragment, so a block's K' rows span at most this range)
    int fmax = 0;
    for (int f : c->frag) fmax = std::max(fmax, f);
    std::vector<int> lo(fmax + 1, INT32_MAX), hi(fmax + 1, -1);
    for (int sh = 0; sh < c->nshell; ++sh) {
      const int f = c->frag[c->atom[sh]];
      lo[f] = std::min(lo[f], c->bfc[sh]);
      hi[f] = std::max(hi[f], c->bfc[sh] + ncart(c->l[sh]));
    }
    c->kcols = 0;
    for (int f = 0; f <= fmax; ++f) {
      if (hi[f] < 0) lo[f] = 0;
      else c->kcols = std::max(c->kcols, hi[f] - lo[f]);
    }
    c->d_frag_col0 = dupload(c, lo);
  }
  c->h_prefix = allpre;
  c->h_kstart = allkst;
  c->h_kstart_off = kst_off;
  c->h_ppair.assign(c->npseudo, 0);
  for (int i = 0; i < c->npairs; ++i)
    for (int k = 0; k < c->h_pp_n[i]; ++k) c->h_ppair[c->h_pp_first[i] + k] = i;
  for (size_t k = 0; k < c->work.size(); ++k) {
    auto& wh = c->work[k];
    wh.w.prefix = c->d_prefix + wh.prefix_off;
    wh.w.kstart = c->fragmented ? c->d_kstart + kst_off[k] : nullptr;
  }
}

#define QX_JK_ROWS_NMAX 2048

// K' rows of a J/K pass in LDS (each class kernel still decides whether its
// rows fit): over the bra fragment's function range in a fragmented context,
// over all functions otherwise; none when that width exceeds
// QX_JK_ROWS_NMAX, where zeroing and scanning a block's rows would cost more
// than its quartets.
static void set_krows(const qx_ctx* c, QxEpiDev& e) {
  const int width = c->fragmented ? c->kcols : c->ncart;
  e.krows = width <= QX_JK_ROWS_NMAX ? 1 : 0;
  if (c->fragmented) {
    e.kcols = c->kcols;
    e.frag_col0 = c->d_frag_col0;
    e.pair_frag = c->d_pair_frag;
  }
}

// Copies of the 3N gradient accumulators of an adjoint-gradient pass,
// cleared on stream s (the pass's blocks flush into copy block % copies; the
// runtime adds the copies to the caller's gradient after the pass).
#define QX_GRAD_COPIES 64
static double* begin_grad_copies(qx_ctx* c, hipStream_t s) {
  const long long n = (long long)QX_GRAD_COPIES * 3 * c->natom;
  if (!c->gcopies) c->gcopies = dalloc<double>(c, (size_t)n);
  check_launch(qxk_fill(c->gcopies, n, 0.0, s), "fill");
  return c->gcopies;
}

// Work matrices of a J/K or gradient pass on stream s (see qx_ctx::ws).
double* begin_ws(qx_ctx* c, hipStream_t s) {
  if (!c->ws) {
    c->ws = dalloc<double>(c, (size_t)5 * c->ncart * c->ncart);
    QX_HIP(hipEventCreateWithFlags(&c->ws_ev, hipEventDisableTiming));
  }
  if (c->ws_used && !c->capturing) QX_HIP(hipStreamWaitEvent(s, c->ws_ev, 0));
  return c->ws;
}

void end_ws(qx_ctx* c, hipStream_t s) {
  if (c->capturing) return;
  QX_HIP(hipEventRecord(c->ws_ev, s));
  c->ws_used = true;
}

// Mixed precision: for every quartet list, the number of FP64 kets of each
// bra row at a Schwarz-bound threshold.  Kets of each bra pair are sorted by
// descending Q, so for bra i the FP64 kets (Q_i Q_j >= thr^2) are a prefix of
// its admissible kets; a masked launch (w.sel) then runs either part of the
// unchanged list, keeping its block -> bra-row layout.
void build_split(qx_ctx* c, double thr) {
  auto& S = c->split;
  if (S.thr == thr) return;
  QX_HIP(hipDeviceSynchronize());  // no kernel may still read the old counts
  if (S.d_nk64) (void)hipFree(S.d_nk64);
  S = qx_ctx::Split{};
  const double tol2 = c->tol * c->tol, thr2 = thr * thr;
  std::vector<int> nk;                 // per list, per bra row
  std::vector<size_t> off;
  for (const auto& wh : c->work) {
    const int X = wh.bra_class, Y = wh.ket_class;
    const int bo = c->class_pair_off[X], bn = c->class_pair_cnt[X];
    const int ko = c->class_pair_off[Y], kn = c->class_pair_cnt[Y];
    std::vector<long long> kcum(kn + 1, 0);
    for (int j = 0; j < kn; ++j) kcum[j + 1] = kcum[j] + c->h_pp_n[ko + j];
    off.push_back(nk.size());
    long long p64 = 0, p32 = 0, q64 = 0, q32 = 0;
    for (int i = 0; i < bn; ++i) {
      int ss, kc, ss2, k64;
      ket_window(c, X, Y, i, tol2, ss, kc);
      ket_window(c, X, Y, i, thr2, ss2, k64);
      k64 = std::min(kc, k64);
      q64 += k64;
      q32 += kc - k64;
      const long long r64 = kcum[ss + k64] - kcum[ss], rall = kcum[ss + kc] - kcum[ss];
      for (int k = 0; k < c->h_pp_n[bo + i]; ++k) {
        nk.push_back((int)r64);
        p64 += r64;
        p32 += rall - r64;
      }
    }
    if ((long long)(nk.size() - off.back()) != wh.w.nbra)
      throw QxError{QX_EHIP, "mixed-precision rows do not match the quartet list"};
    if (p64 + p32 != wh.w.nquart) throw QxError{QX_EHIP, "mixed-precision split lost quartets"};
    S.pq64.push_back(p64); S.pq32.push_back(p32);
    S.cq64.push_back(q64); S.cq32.push_back(q32);
  }
  QX_HIP(hipMalloc(&S.d_nk64, std::max<size_t>(nk.size(), 1) * sizeof(int)));
  if (!nk.empty()) QX_HIP(hipMemcpy(S.d_nk64, nk.data(), nk.size() * sizeof(int), hipMemcpyHostToDevice));
  S.wm = c->work;
  for (size_t i = 0; i < S.wm.size(); ++i) S.wm[i].w.nk64 = S.d_nk64 + off[i];
  S.thr = thr;
}

void setup_shell_tables(qx_ctx* c) {
  QxShellsDev& s = c->shells;
  s.nshell = c->nshell; s.natom = c->natom; s.ncart = c->ncart; s.nbf = c->nbf; s.pure = c->pure;
  s.l = dupload(c, c->l); s.atom = dupload(c, c->atom); s.nprim = dupload(c, c->nprim);
  s.poff = dupload(c, c->poff); s.bfc = dupload(c, c->bfc); s.bfs = dupload(c, c->bfs);
  s.exps = dupload(c, c->exps); s.coefs = dupload(c, c->coefs);
  s.xyz = dupload(c, c->xyz); s.Z = dupload(c, c->Z);
  std::vector<int> fs, fc, as;
  for (int sh = 0; sh < c->nshell; ++sh) {
    const int l = c->l[sh];
    for (int i = 0; i <= l; ++i)
      for (int j = 0; j <= i; ++j) {
        fs.push_back(sh);
        fc.push_back(((l - i) << 16) | ((i - j) << 8) | j);
      }
    for (int k = 0; k < nsph(l, c->pure); ++k) as.push_back(sh);
  }
  s.f_shell = dupload(c, fs); s.f_comp = dupload(c, fc); s.ao_shell = dupload(c, as);
  // transform blocks: identity for l < 2 (stored for uniform access), solid harmonics for d
  std::vector<double> blk;
  for (int l = 0; l <= QX_LMAX_SUPPORTED + 1; ++l) s.sph_off[l] = 0;
  s.sph_off[0] = 0; blk.push_back(1.0);
  s.sph_off[1] = (int)blk.size();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) blk.push_back(i == j ? 1.0 : 0.0);
  s.sph_off[2] = (int)blk.size();
  if (c->pure) {
    for (int k = 0; k < 30; ++k) blk.push_back(kSph2[k]);
  } else {
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) blk.push_back(i == j ? 1.0 : 0.0);
  }
  s.sph_off[3] = (int)blk.size();
  s.sph = dupload(c, blk);
}

// Work matrices of qx_jk_build_noalloc: one file-scope device array per GPU
// (each device holds its own copy until the process ends), 3 Mi doubles =
// 24 MiB, enough for 5 n_cart^2 matrices up to n_cart = 792 (spherical
// bases) or 2 n_cart^2 up to n_cart = 1254 (Cartesian).  Host code never
// uses the array's name as an address: qx_initialize looks up its device
// address once per device and keeps it in g_scratch_ptr.
#define QX_NOALLOC_DOUBLES (3ull << 20)
static __device__ __align__(256) double scratch[QX_NOALLOC_DOUBLES];

#define QX_MAX_DEVICES 64
std::mutex g_scratch_mu;
double* g_scratch_ptr[QX_MAX_DEVICES] = {};

// Device address of `scratch` on `device`, which must be current.
double* scratch_address(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  QX_REQUIRE(device >= 0 && device < QX_MAX_DEVICES, QX_EINVAL, "device index beyond the scratch table");
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    QX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    QX_REQUIRE(p != nullptr, QX_EHIP, "no device address for the scratch array");
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  return g_scratch_ptr[device];
}

thread_local std::string g_init_error;

int fail(qx_ctx* c, int code, const char* msg) noexcept {
  try {
    if (c) c->err = msg; else g_init_error = msg;
  } catch (...) {
    // the message itself could not be stored; the code still reports the failure
  }
  return code;
}

int fail(qx_ctx* c, const QxError& e) noexcept { return fail(c, e.code, e.msg.c_str()); }

// Runs the body of a C-ABI entry point: every C++ exception becomes a status
// code (nothing crosses the C ABI, so nothing reaches std::terminate).
template <class F>
int guarded(qx_ctx* c, F&& body) noexcept {
  try {
    return body();
  } catch (const QxError& e) {
    return fail(c, e);
  } catch (const std::bad_alloc&) {
    return fail(c, QX_ENOMEM, "host memory allocation failed");
  } catch (const std::exception& e) {
    return fail(c, QX_EHIP, e.what());
  } catch (...) {
    return fail(c, QX_EHIP, "unexpected C++ exception");
  }
}

}  // namespace

// ============================================================== C ABI
extern "C" {

static void init_body(qx_ctx* c, int natom, const double* xyz_bohr, const double* nuclear_charges, int nshell,
                      const int* shell_l, const int* shell_atom, const int* shell_nprim, const double* exps,
                      const double* coefs, int pure, double ints_tol, int device, const int* atom_fragment);

static int init_impl(int natom, const double* xyz_bohr, const double* nuclear_charges, int nshell,
                     const int* shell_l, const int* shell_atom, const int* shell_nprim, const double* exps,
                     const double* coefs, int pure, double ints_tol, int device, const int* atom_fragment,
                     qx_ctx** out) {
  if (!out) return QX_EINVAL;
  *out = nullptr;
  qx_ctx* c = nullptr;
  const int rc = guarded(nullptr, [&]() -> int {
    try {
      c = new qx_ctx();
      for (int k = 0; k < QX_NUM_CLASSES; ++k)
        for (int pass = 0; pass < 2; ++pass) {
          c->fp32_ok[pass][k] = qx_fp32_default[k][pass];
          c->tile[pass][k] = qx_tile_default[k][pass];
          c->gwred[k] = qx_gwred_default[k];
          c->jkrows[k] = 1;
          c->gcopy[k] = qx_gcopy_default[k];
        }
      init_body(c, natom, xyz_bohr, nuclear_charges, nshell, shell_l, shell_atom, shell_nprim, exps, coefs, pure,
                ints_tol, device, atom_fragment);
    } catch (...) {
      if (c) qx_finalize(c);
      c = nullptr;
      throw;
    }
    return QX_OK;
  });
  if (rc == QX_OK) *out = c;
  return rc;
}

static void init_body(qx_ctx* c, int natom, const double* xyz_bohr, const double* nuclear_charges, int nshell,
                      const int* shell_l, const int* shell_atom, const int* shell_nprim, const double* exps,
                      const double* coefs, int pure, double ints_tol, int device, const int* atom_fragment) {
  {
    QX_REQUIRE(natom > 0 && xyz_bohr && nuclear_charges, QX_EINVAL, "natom must be positive with coordinates");
    QX_REQUIRE(nshell > 0 && shell_l && shell_atom && shell_nprim && exps && coefs, QX_EINVAL, "no shells");
    QX_REQUIRE(ints_tol > 0 && std::isfinite(ints_tol), QX_EINVAL, "ints_tol must be positive");
    // host-side tables first (a host allocation failure is reported before
    // any device work starts)
    c->natom = natom;
    c->frag.assign(natom, 0);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw QxError{QX_ENODEVICE, "no HIP device"};
    QX_REQUIRE(device >= 0 && device < ndev, QX_EINVAL, "device index out of range");
    QX_HIP(hipSetDevice(device));
    c->device = device;
    c->scratch = scratch_address(device);
    if (atom_fragment) {
      for (int a = 0; a < natom; ++a) {
        QX_REQUIRE(atom_fragment[a] >= 0, QX_EINVAL, "fragment ids must be non-negative");
        c->frag[a] = atom_fragment[a];
      }
      c->fragmented = true;
    }
    c->xyz.assign(xyz_bohr, xyz_bohr + 3 * natom);
    c->Z.assign(nuclear_charges, nuclear_charges + natom);
    for (double v : c->xyz) QX_REQUIRE(std::isfinite(v), QX_EINVAL, "non-finite coordinate");
    c->nshell = nshell;
    c->pure = pure ? 1 : 0;
    c->tol = ints_tol;
    int tp = 0, nc = 0, nb = 0;
    for (int s = 0; s < nshell; ++s) {
      QX_REQUIRE(shell_l[s] >= 0 && shell_l[s] <= QX_LMAX_SUPPORTED, QX_EUNSUPPORTED,
                 "shell angular momentum above " + std::to_string(QX_LMAX_SUPPORTED) + " is not built");
      QX_REQUIRE(shell_atom[s] >= 0 && shell_atom[s] < natom, QX_EINVAL, "shell atom index out of range");
      QX_REQUIRE(shell_nprim[s] >= 1, QX_EINVAL, "shell without primitives");
      c->l.push_back(shell_l[s]);
      c->atom.push_back(shell_atom[s]);
      c->nprim.push_back(shell_nprim[s]);
      c->poff.push_back(tp);
      c->bfc.push_back(nc);
      c->bfs.push_back(nb);
      tp += shell_nprim[s];
      nc += ncart(shell_l[s]);
      nb += nsph(shell_l[s], c->pure);
      if (shell_nprim[s] != 1) c->k1 = false;
    }
    c->ncart = nc;
    c->nbf = nb;
    c->exps.assign(exps, exps + tp);
    c->coefs.resize(tp);
    for (double v : c->exps) QX_REQUIRE(v > 0 && std::isfinite(v), QX_EINVAL, "exponents must be positive");
    for (int s = 0; s < nshell; ++s)
      normalize_shell(c->l[s], c->nprim[s], &c->exps[c->poff[s]], coefs + c->poff[s], &c->coefs[c->poff[s]]);
    c->d_boys = dupload(c, boys_table());
    setup_shell_tables(c);
    schwarz_and_lists(c);
    QX_HIP(hipDeviceSynchronize());
  }
}

int qx_initialize(int natom, const double* xyz_bohr, const double* nuclear_charges, int nshell,
                  const int* shell_l, const int* shell_atom, const int* shell_nprim, const double* exps,
                  const double* coefs, int pure, double ints_tol, int device, qx_ctx** out) {
  return init_impl(natom, xyz_bohr, nuclear_charges, nshell, shell_l, shell_atom, shell_nprim, exps, coefs, pure,
                   ints_tol, device, nullptr, out);
}

int qx_initialize_fragments(int natom, const double* xyz_bohr, const double* nuclear_charges, int nshell,
                            const int* shell_l, const int* shell_atom, const int* shell_nprim, const double* exps,
                            const double* coefs, int pure, double ints_tol, int device, const int* atom_fragment,
                            qx_ctx** out) {
  if (!atom_fragment) return QX_EINVAL;
  return init_impl(natom, xyz_bohr, nuclear_charges, nshell, shell_l, shell_atom, shell_nprim, exps, coefs, pure,
                   ints_tol, device, atom_fragment, out);
}

int qx_finalize(qx_ctx* c) {
  if (!c) return QX_EINVAL;
  (void)hipSetDevice(c->device);
  for (auto& p : c->prof_pending) {
    (void)hipEventSynchronize(p.stop);
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  (void)hipDeviceSynchronize();
  for (auto t : c->side) (void)hipStreamDestroy(t);
  for (auto e : c->fork_ev) (void)hipEventDestroy(e);
  if (c->ws_ev) (void)hipEventDestroy(c->ws_ev);
  if (c->split.d_nk64) (void)hipFree(c->split.d_nk64);
  for (void* p : c->allocs) (void)hipFree(p);
  delete c;
  return QX_OK;
}

const char* qx_last_error(const qx_ctx* c) { return c ? c->err.c_str() : g_init_error.c_str(); }

int qx_nbf(const qx_ctx* c, int* nbf, int* nc) {
  if (!c) return QX_EINVAL;
  if (nbf) *nbf = c->nbf;
  if (nc) *nc = c->ncart;
  return QX_OK;
}

int qx_schwarz_screening(qx_ctx* c, double* values, int* pairs, int* npairs) {
  if (!c) return QX_EINVAL;
  if (values) std::memcpy(values, c->Q.data(), c->Q.size() * sizeof(double));
  if (pairs) std::memcpy(pairs, c->ref_pairs.data(), c->ref_pairs.size() * sizeof(int));
  if (npairs) *npairs = (int)c->ref_pairs.size() / 2;
  return QX_OK;
}

int qx_num_quartets(const qx_ctx* c, long long* nq) {
  if (!c || !nq) return QX_EINVAL;
  *nq = c->nquart_total;
  return QX_OK;
}

long long qx_how_many_derivs(int natom, int order) {
  if (natom < 0 || order < 0) return -1;
  long long val = 1, fact = 1;
  for (int i = 0; i < order; ++i) {
    val *= 3LL * natom + i;
    fact *= i + 1;
  }
  return val / fact;
}

long long qx_generate_multi_index_lookup(int nparams, int order, int* out, long long cap) {
  if (nparams < 0 || order < 0) return -1;
  std::vector<int> cur(order, 0);
  long long count = 0;
  // combinations with replacement in lexicographic order
  if (order == 0) return 1;
  if (nparams == 0) return 0;
  while (true) {
    if (out) {
      if (count >= cap) return -1;
      for (int k = 0; k < order; ++k) out[count * order + k] = cur[k];
    }
    ++count;
    int k = order - 1;
    while (k >= 0 && cur[k] == nparams - 1) --k;
    if (k < 0) break;
    const int v = cur[k] + 1;
    for (int j = k; j < order; ++j) cur[j] = v;
  }
  return count;
}

static void ao_tensor4_from_cart(qx_ctx* c, const double* cart, double* out, long long outer_slices, hipStream_t s) {
  // four successive index transforms, Cartesian -> AO, per slice
  const long long nc = c->ncart, nb = c->nbf;
  DevBuf t1((size_t)outer_slices * nb * nc * nc * nc * sizeof(double));
  DevBuf t2((size_t)outer_slices * nb * nb * nc * nc * sizeof(double));
  check_launch(qxk_cart_to_ao_axis(&c->shells, cart, t1.d(), outer_slices, (int)nc, (int)nb, nc * nc * nc, s), "sph");
  check_launch(qxk_cart_to_ao_axis(&c->shells, t1.d(), t2.d(), outer_slices * nb, (int)nc, (int)nb, nc * nc, s), "sph");
  check_launch(qxk_cart_to_ao_axis(&c->shells, t2.d(), t1.d(), outer_slices * nb * nb, (int)nc, (int)nb, nc, s), "sph");
  check_launch(qxk_cart_to_ao_axis(&c->shells, t1.d(), out, outer_slices * nb * nb * nb, (int)nc, (int)nb, 1, s), "sph");
  QX_HIP(hipStreamSynchronize(s));
}

int qx_compute_2e_int(qx_ctx* c, double* out, void* stream) {
  if (!c || !out) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nc = c->ncart;
    const size_t n4 = nc * nc * nc * nc;
    DevBuf* tmp = c->pure ? new DevBuf(n4 * sizeof(double)) : nullptr;
    std::unique_ptr<DevBuf> guard(tmp);
    double* dst = c->pure ? tmp->d() : out;
    check_launch(qxk_fill(dst, (long long)n4, 0.0, s), "fill");
    QxEpiDev e{};
    e.epi = QX_EPI_FULL; e.atomic = c->k1 ? 0 : 1; e.nbf = (int)nc; e.natom = c->natom; e.coord = -1; e.out = dst;
    for (auto& wh : c->work) check_launch(launch_eri(c, wh.w, wh.cls, e, 0, s, c->pr, c->pp), "compute_2e_int");
    if (c->pure) ao_tensor4_from_cart(c, dst, out, 1, s);
    return QX_OK;
  });
}

// Derivative kernels (every part) of one quartet list; sel / nq: the masked
// part of the list the launch evaluates and its pseudo-pair quartet count.
static void add_deriv_list(qx_ctx* c, const QxWorkHost& wh, const QxEpiDev& e, int fp32, int sel, long long nq,
                           std::vector<Launch>& out) {
  if (wh.w.nquart <= 0 || nq <= 0) return;
  const QxClassEntry& ent = qx_class_table[wh.cls];
  // adjoint gradient: one launch per group of parts; tensor: one per part
  const bool grad = e.epi == QX_EPI_GRAD;
  const int nl = grad ? ent.ngroups : ent.nparts;
  for (int k = 0; k < nl; ++k) {
    QxArgs a;
    a.pr = c->pr; a.pp = c->pp; a.w = wh.w; a.e = e; a.boys = c->d_boys;
    a.w.sel = sel;
    if (grad) {
      a.e.grad_wred = c->gwred[wh.cls];
      a.e.grad_copies = std::min(std::max(c->gcopy[wh.cls], 1), a.e.grad_copies);
    }
    const int slot = QX_SLOT_DERIV(wh.cls, k, fp32 ? 1 : 0);
    auto fn = grad ? ent.grad[k] : ent.deri[k];
    const double fl = grad ? ent.group_flops[k] : ent.part_flops[k];
    const int block = grad ? c->tile[1][wh.cls] : 0;
    out.push_back({(double)nq * fl, [c, a, fn, fp32, slot, nq, block](hipStream_t st) {
                     QxLaunch L{fp32, st, block};
                     return timed_launch(c, slot, nq, st, [&] { return fn(&a, &L); });
                   }});
  }
}

// J/K kernels (every part) of one quartet list, as add_deriv_list.
static void add_jk_list(qx_ctx* c, const QxWorkHost& wh, const QxEpiDev& e, int fp32, int sel, long long nq,
                        std::vector<Launch>& out) {
  if (wh.w.nquart <= 0 || nq <= 0) return;
  const QxClassEntry& ent = qx_class_table[wh.cls];
  for (int k = 0; k < ent.njk; ++k) {
    QxArgs a;
    a.pr = c->pr; a.pp = c->pp; a.w = wh.w; a.e = e; a.boys = c->d_boys;
    a.w.sel = sel;
    a.e.krows = e.krows && c->jkrows[wh.cls];
    const int slot = QX_SLOT_JK(wh.cls, k, fp32 ? 1 : 0);
    auto fn = ent.jk[k];
    const int block = c->tile[0][wh.cls];
    out.push_back({(double)nq * ent.jk_flops[k], [c, a, fn, fp32, slot, nq, block](hipStream_t st) {
                     QxLaunch L{fp32, st, block};
                     return timed_launch(c, slot, nq, st, [&] { return fn(&a, &L); });
                   }});
  }
}

static void add_deriv_launches(qx_ctx* c, const std::vector<QxWorkHost>& work, const QxEpiDev& e, int fp32,
                               std::vector<Launch>& out) {
  for (const auto& wh : work) add_deriv_list(c, wh, e, fp32, 0, wh.w.nquart, out);
}

static void add_jk_launches(qx_ctx* c, const std::vector<QxWorkHost>& work, const QxEpiDev& e, int fp32,
                            std::vector<Launch>& out) {
  for (const auto& wh : work) add_jk_list(c, wh, e, fp32, 0, wh.w.nquart, out);
}

// Mixed precision, per quartet list i of the split (pass 0 J/K, 1
// gradient): 0 = FP64 kernels, 1 = FP32 kernels, 2 = two launches masked to
// the FP64 and FP32 kets.  FP32 is used in the classes whose policy admits it
// and only when at least QX_MIXED_MIN_FRAC of the list's primitive quartets
// are below the threshold: a masked launch pair runs every wave that holds
// quartets of both kinds twice, which a small FP32 share does not repay.
#define QX_MIXED_MIN_FRAC 0.25
static int mixed_mode(const qx_ctx* c, size_t i, int pass) {
  const auto& S = c->split;
  const long long p64 = S.pq64[i], p32 = S.pq32[i];
  if (!c->fp32_ok[pass][S.wm[i].cls] || p32 < QX_MIXED_MIN_FRAC * (double)(p64 + p32)) return 0;
  return p64 == 0 ? 1 : 2;
}

static void add_mixed_launches(qx_ctx* c, double thr, int pass, const QxEpiDev& e, std::vector<Launch>& out) {
  build_split(c, thr);
  const auto& S = c->split;
  auto add = pass == 0 ? add_jk_list : add_deriv_list;
  for (size_t i = 0; i < S.wm.size(); ++i) {
    const QxWorkHost& wh = S.wm[i];
    switch (mixed_mode(c, i, pass)) {
      case 0: add(c, wh, e, 0, 0, wh.w.nquart, out); break;
      case 1: add(c, wh, e, 1, 0, wh.w.nquart, out); break;
      default:
        add(c, wh, e, 0, 1, S.pq64[i], out);
        add(c, wh, e, 1, 2, S.pq32[i], out);
    }
  }
}

static void launch_deriv_all(qx_ctx* c, const QxEpiDev& e, int fp32, hipStream_t s) {
  std::vector<Launch> v;
  add_deriv_launches(c, c->work, e, fp32, v);
  run_launches(c, v, s, "eri derivative", false);
}

static void deriv_tensor(qx_ctx* c, int coord, double* out, hipStream_t s) {
  const size_t nc = c->ncart;
  const size_t n4 = nc * nc * nc * nc;
  const long long nslices = coord >= 0 ? 1 : 3LL * c->natom;
  std::unique_ptr<DevBuf> tmp(c->pure ? new DevBuf(nslices * n4 * sizeof(double)) : nullptr);
  double* dst = c->pure ? tmp->d() : out;
  check_launch(qxk_fill(dst, (long long)(nslices * n4), 0.0, s), "fill");
  QxEpiDev e{};
  e.epi = QX_EPI_DFULL; e.atomic = 1; e.nbf = (int)nc; e.natom = c->natom; e.coord = coord; e.out = dst;
  launch_deriv_all(c, e, 0, s);
  if (c->pure) ao_tensor4_from_cart(c, dst, out, nslices, s);
}

// Tangent (second-derivative) kernels, every part of every class.
static void add_dert_launches(qx_ctx* c, const QxEpiDev& e, std::vector<Launch>& out) {
  for (const auto& wh : c->work) {
    if (wh.w.nquart <= 0) continue;
    const QxClassEntry& ent = qx_class_table[wh.cls];
    for (int k = 0; k < ent.ntparts; ++k) {
      QxArgs a;
      a.pr = c->pr; a.pp = c->pp; a.w = wh.w; a.e = e; a.boys = c->d_boys;
      auto fn = ent.dert[k];
      const long long nq = wh.w.nquart;
      out.push_back({(double)nq / ent.ntparts, [a, fn](hipStream_t st) {
                       QxLaunch L{0, st};
                       return fn(&a, &L);
                     }});
    }
  }
}

// Index of the coordinate pair (x <= y) among the m(m+1)/2 pairs in
// generate_multi_index_lookup(m, 2) order.
static long long pair_index(int x, int y, int m) {
  return (long long)x * m - (long long)x * (x - 1) / 2 + (y - x);
}

// Second-derivative tensor: for each nuclear coordinate x the tangent
// kernels along e_x give d2(ab|cd)/dx dy for every y; slices y >= x go to
// pair (x, y) (all pairs when want < 0, else only pair `want` into slice 0).
// Quartets on one atom are skipped, like the first-derivative tensor.
static void deriv2_tensor(qx_ctx* c, long long want, double* out, hipStream_t s) {
  const size_t nc = c->ncart;
  const size_t n4 = nc * nc * nc * nc;
  const int m = 3 * c->natom;
  const long long npair = (long long)m * (m + 1) / 2;
  const long long nslices = want >= 0 ? 1 : npair;
  std::unique_ptr<DevBuf> tmp(c->pure ? new DevBuf(nslices * n4 * sizeof(double)) : nullptr);
  double* dst = c->pure ? tmp->d() : out;
  check_launch(qxk_fill(dst, (long long)(nslices * n4), 0.0, s), "fill");
  std::vector<double> tv((size_t)m * m, 0.0);
  std::vector<int> map((size_t)m * m, -1);
  std::vector<int> xs;
  for (int x = 0; x < m; ++x) {
    tv[(size_t)x * m + x] = 1.0;
    bool any = false;
    for (int y = x; y < m; ++y) {
      const long long q = pair_index(x, y, m);
      if (want < 0) map[(size_t)x * m + y] = (int)q;
      else if (q == want) map[(size_t)x * m + y] = 0;
      any = any || map[(size_t)x * m + y] >= 0;
    }
    if (any) xs.push_back(x);
  }
  DevBuf dtv(tv.size() * sizeof(double)), dmap(map.size() * sizeof(int));
  QX_HIP(hipMemcpy(dtv.d(), tv.data(), tv.size() * sizeof(double), hipMemcpyHostToDevice));
  QX_HIP(hipMemcpy(dmap.p, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
  for (int x : xs) {
    QxEpiDev e{};
    e.epi = QX_EPI_DFULL; e.atomic = 1; e.nbf = (int)nc; e.natom = c->natom; e.coord = -1; e.out = dst;
    e.tv = dtv.d() + (size_t)x * m;
    e.slot_map = static_cast<const int*>(dmap.p) + (size_t)x * m;
    std::vector<Launch> v;
    add_dert_launches(c, e, v);
    run_launches(c, v, s, "eri second derivative", false);
  }
  if (c->pure) ao_tensor4_from_cart(c, dst, out, nslices, s);
  QX_HIP(hipStreamSynchronize(s));  // dtv / dmap are freed on return
}

int qx_eri_deriv_core(qx_ctx* c, int order, double* out, void* stream) {
  if (!c || !out) return QX_EINVAL;
  if (order != 1 && order != 2) {
    return fail(c, QX_EUNSUPPORTED, "eri_deriv_core: deriv_order 1 and 2 are built");
  }
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    if (order == 1) deriv_tensor(c, -1, out, (hipStream_t)stream);
    else deriv2_tensor(c, -1, out, (hipStream_t)stream);
    return QX_OK;
  });
}

int qx_compute_2e_deriv(qx_ctx* c, const int* dv, int len, double* out, void* stream) {
  if (!c || !out || !dv) return QX_EINVAL;
  if (len != 3 * c->natom) {
    return fail(c, QX_EINVAL, "compute_2e_deriv: deriv_vec length must be 3*natom");
  }
  int order = 0, coord = -1, c2[2] = {-1, -1}, k = 0;
  for (int i = 0; i < len; ++i) {
    if (dv[i] < 0) return fail(c, QX_EINVAL, "compute_2e_deriv: negative deriv_vec entry");
    order += dv[i];
    if (dv[i]) coord = i;
    for (int r = 0; r < dv[i] && k < 2; ++r) c2[k++] = i;
  }
  if (order > 2) {
    return fail(c, QX_EUNSUPPORTED, "compute_2e_deriv: deriv_vec orders 0, 1 and 2 are built");
  }
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    if (order == 0) {
      // the reference's order-0 call returns buffer 0 of a deriv-0 engine, i.e.
      // the integrals, with single-atom quartets skipped (libint_interface.cc:1180-1185)
      const size_t nc = c->ncart;
      const size_t n4 = nc * nc * nc * nc;
      std::unique_ptr<DevBuf> tmp(c->pure ? new DevBuf(n4 * sizeof(double)) : nullptr);
      double* dst = c->pure ? tmp->d() : out;
      check_launch(qxk_fill(dst, (long long)n4, 0.0, s), "fill");
      QxEpiDev e{};
      e.epi = QX_EPI_FULL; e.atomic = c->k1 ? 0 : 1; e.nbf = (int)nc; e.natom = c->natom; e.coord = -1;
      e.out = dst; e.skip_single = 1;
      for (auto& wh : c->work) check_launch(launch_eri(c, wh.w, wh.cls, e, 0, s, c->pr, c->pp), "compute_2e_deriv");
      if (c->pure) ao_tensor4_from_cart(c, dst, out, 1, s);
      return QX_OK;
    }
    if (order == 1) deriv_tensor(c, coord, out, s);
    else deriv2_tensor(c, pair_index(c2[0], c2[1], len), out, s);
    return QX_OK;
  });
}

int qx_jk_build(qx_ctx* c, const double* D, double* J, double* K, int precision, double thr, void* stream) {
  return qx_jk_build_masked(c, D, J, K, precision, thr, nullptr, stream);
}

int qx_jk_build_masked(qx_ctx* c, const double* D, double* J, double* K, int precision, double thr,
                       const int* frag_skip, void* stream) {
  if (!c || !D || !J || !K) return QX_EINVAL;
  if (frag_skip && !c->fragmented) return QX_EINVAL;
  if (precision != QX_PREC_FP64 && precision != QX_PREC_MIXED) return QX_EINVAL;
  if (precision == QX_PREC_MIXED && !(thr >= 0.0 && std::isfinite(thr))) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nc = c->ncart, n2 = nc * nc;
    double* w = begin_ws(c, s);
    double *Jp = w, *Kp = w + n2, *Dc = w + 2 * n2, *Js = w + 3 * n2, *Ks = w + 4 * n2;
    const double* Dcart = D;
    if (c->pure) {
      check_launch(qxk_ao_to_cart_mat(&c->shells, D, Dc, s), "ao_to_cart");
      Dcart = Dc;
    }
    check_launch(qxk_fill(Jp, (long long)n2, 0.0, s), "fill");
    check_launch(qxk_fill(Kp, (long long)n2, 0.0, s), "fill");
    QxEpiDev e{};
    e.epi = QX_EPI_JK; e.nbf = (int)nc; e.natom = c->natom; e.coord = -1; e.D = Dcart; e.J = Jp; e.K = Kp;
    set_krows(c, e);
    if (frag_skip) {
      e.frag_skip = frag_skip;
      e.pair_frag = c->d_pair_frag;
    }
    std::vector<Launch> v;
    if (precision == QX_PREC_MIXED) {
      add_mixed_launches(c, thr, 0, e, v);
    } else {
      add_jk_launches(c, c->work, e, 0, v);
    }
    run_launches(c, v, s, "jk_build", true);
    if (c->pure) {
      check_launch(qxk_sym_half(Jp, Js, (int)nc, s), "sym");
      check_launch(qxk_sym_half(Kp, Ks, (int)nc, s), "sym");
      check_launch(qxk_cart_to_ao_mat(&c->shells, Js, J, s), "cart_to_ao");
      check_launch(qxk_cart_to_ao_mat(&c->shells, Ks, K, s), "cart_to_ao");
    } else {
      check_launch(qxk_sym_half(Jp, J, (int)nc, s), "sym");
      check_launch(qxk_sym_half(Kp, K, (int)nc, s), "sym");
    }
    end_ws(c, s);
    return QX_OK;
  });
}

// qx_jk_build (FP64) without any device allocation: the J'/K' work
// matrices live in the scratch array, every class kernel runs in turn on the
// caller's stream (no side streams, events or timing), and the work matrices
// are cleared by a fill kernel on every call.  Calls that overlap on one GPU
// share the array and race.
int qx_jk_build_noalloc(qx_ctx* c, const double* D, double* J, double* K, void* stream) {
  if (!c || !D || !J || !K) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nc = c->ncart, n2 = nc * nc;
    const size_t need = (c->pure ? 5 : 2) * n2;
    if (need > QX_NOALLOC_DOUBLES)
      throw QxError{QX_EINVAL, "jk_build_noalloc: " + std::to_string(need) + " doubles of work matrices (" +
                                   std::to_string(nc) + " Cartesian functions) exceed the " +
                                   std::to_string(QX_NOALLOC_DOUBLES) + "-double scratch array"};
    QX_REQUIRE(c->scratch != nullptr, QX_EHIP, "jk_build_noalloc: scratch array not set up");
    double* w = c->scratch;
    double *Jp = w, *Kp = w + n2;
    const double* Dcart = D;
    if (c->pure) {
      check_launch(qxk_ao_to_cart_mat(&c->shells, D, w + 2 * n2, s), "ao_to_cart");
      Dcart = w + 2 * n2;
    }
    check_launch(qxk_fill(Jp, (long long)n2, 0.0, s), "fill");
    check_launch(qxk_fill(Kp, (long long)n2, 0.0, s), "fill");
    QxEpiDev e{};
    e.epi = QX_EPI_JK; e.nbf = (int)nc; e.natom = c->natom; e.coord = -1; e.D = Dcart; e.J = Jp; e.K = Kp;
    set_krows(c, e);
    for (const auto& wh : c->work) {
      if (wh.w.nquart <= 0) continue;
      const QxClassEntry& ent = qx_class_table[wh.cls];
      for (int k = 0; k < ent.njk; ++k) {
        QxArgs a;
        a.pr = c->pr; a.pp = c->pp; a.w = wh.w; a.e = e; a.boys = c->d_boys;
        QxLaunch L{0, s};
        check_launch(ent.jk[k](&a, &L), "jk_build_noalloc");
      }
    }
    if (c->pure) {
      double *Js = w + 3 * n2, *Ks = w + 4 * n2;
      check_launch(qxk_sym_half(Jp, Js, (int)nc, s), "sym");
      check_launch(qxk_sym_half(Kp, Ks, (int)nc, s), "sym");
      check_launch(qxk_cart_to_ao_mat(&c->shells, Js, J, s), "cart_to_ao");
      check_launch(qxk_cart_to_ao_mat(&c->shells, Ks, K, s), "cart_to_ao");
    } else {
      check_launch(qxk_sym_half(Jp, J, (int)nc, s), "sym");
      check_launch(qxk_sym_half(Kp, K, (int)nc, s), "sym");
    }
    return QX_OK;
  });
}

long long qx_jk_noalloc_capacity(void) { return (long long)QX_NOALLOC_DOUBLES; }

int qx_eri_grad(qx_ctx* c, const double* D, double* grad, int precision, double thr, void* stream) {
  if (!c || !D || !grad) return QX_EINVAL;
  if (precision != QX_PREC_FP64 && precision != QX_PREC_MIXED) return QX_EINVAL;
  if (precision == QX_PREC_MIXED && !(thr >= 0.0 && std::isfinite(thr))) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nc = c->ncart, n2 = nc * nc;
    double* w = begin_ws(c, s);
    const double* Dcart = D;
    if (c->pure) {
      check_launch(qxk_ao_to_cart_mat(&c->shells, D, w + 2 * n2, s), "ao_to_cart");
      Dcart = w + 2 * n2;
    }
    QxEpiDev e{};
    e.epi = QX_EPI_GRAD; e.nbf = (int)nc; e.natom = c->natom; e.coord = -1; e.D = Dcart;
    e.grad = begin_grad_copies(c, s);
    e.grad_copies = QX_GRAD_COPIES;
    e.grad_win = c->grad_win;
    std::vector<Launch> v;
    if (precision == QX_PREC_MIXED) {
      add_mixed_launches(c, thr, 1, e, v);
    } else {
      add_deriv_launches(c, c->work, e, 0, v);
    }
    run_launches(c, v, s, "eri_grad", true);
    check_launch(qxk_add_copies(c->gcopies, QX_GRAD_COPIES, 3 * c->natom, grad, s), "gradient copies");
    end_ws(c, s);
    return QX_OK;
  });
}

// Forward-over-reverse piece of the Hessian-vector product: the adjoint
// contraction of the second-derivative integrals along v,
// out[y] += sum_pqrs Gamma(D)_pqrs d/dv d(pq|rs)/dR_y, Gamma = 2 D_pq D_rs -
// D_pr D_qs (tangent kernels, integrals recomputed, nothing stored).
int qx_eri_hess_dir(qx_ctx* c, const double* D, const double* v, double* out, void* stream) {
  if (!c || !D || !v || !out) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nc = c->ncart, n2 = nc * nc;
    double* w = begin_ws(c, s);
    const double* Dcart = D;
    if (c->pure) {
      check_launch(qxk_ao_to_cart_mat(&c->shells, D, w + 2 * n2, s), "ao_to_cart");
      Dcart = w + 2 * n2;
    }
    QxEpiDev e{};
    e.epi = QX_EPI_GRAD; e.nbf = (int)nc; e.natom = c->natom; e.coord = -1; e.D = Dcart; e.tv = v;
    e.grad = begin_grad_copies(c, s);
    e.grad_copies = QX_GRAD_COPIES;
    e.grad_win = c->grad_win;
    std::vector<Launch> L;
    add_dert_launches(c, e, L);
    run_launches(c, L, s, "eri_hess_dir", false);
    check_launch(qxk_add_copies(c->gcopies, QX_GRAD_COPIES, 3 * c->natom, out, s), "gradient copies");
    end_ws(c, s);
    return QX_OK;
  });
}

// J and K of the directional derivative of the integrals along v at fixed D:
// J_pq = sum_rs (d/dv (pq|rs)) D_rs, K_pq = sum_rs (d/dv (pr|qs)) D_rs (the
// derivative Fock matrix of the coupled-perturbed equations).
int qx_jk_deriv_dir(qx_ctx* c, const double* D, const double* v, double* J, double* K, void* stream) {
  if (!c || !D || !v || !J || !K) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nc = c->ncart, n2 = nc * nc;
    double* w = begin_ws(c, s);
    double *Jp = w, *Kp = w + n2, *Dc = w + 2 * n2, *Js = w + 3 * n2, *Ks = w + 4 * n2;
    const double* Dcart = D;
    if (c->pure) {
      check_launch(qxk_ao_to_cart_mat(&c->shells, D, Dc, s), "ao_to_cart");
      Dcart = Dc;
    }
    check_launch(qxk_fill(Jp, (long long)n2, 0.0, s), "fill");
    check_launch(qxk_fill(Kp, (long long)n2, 0.0, s), "fill");
    QxEpiDev e{};
    e.epi = QX_EPI_JKD; e.nbf = (int)nc; e.natom = c->natom; e.coord = -1; e.D = Dcart; e.J = Jp; e.K = Kp;
    e.tv = v;
    std::vector<Launch> L;
    add_deriv_launches(c, c->work, e, 0, L);
    run_launches(c, L, s, "jk_deriv_dir", false);
    if (c->pure) {
      check_launch(qxk_sym_half(Jp, Js, (int)nc, s), "sym");
      check_launch(qxk_sym_half(Kp, Ks, (int)nc, s), "sym");
      check_launch(qxk_cart_to_ao_mat(&c->shells, Js, J, s), "cart_to_ao");
      check_launch(qxk_cart_to_ao_mat(&c->shells, Ks, K, s), "cart_to_ao");
    } else {
      check_launch(qxk_sym_half(Jp, J, (int)nc, s), "sym");
      check_launch(qxk_sym_half(Kp, K, (int)nc, s), "sym");
    }
    end_ws(c, s);
    return QX_OK;
  });
}

int qx_oei(qx_ctx* c, double* S, double* T, double* V, void* stream) {
  if (!c || !S || !T || !V) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nc = c->ncart;
    if (!c->pure) {
      check_launch(qxk_oei(&c->shells, S, T, V, s), "oei");
    } else {
      DevBuf a(nc * nc * 8), b(nc * nc * 8), d(nc * nc * 8);
      check_launch(qxk_oei(&c->shells, a.d(), b.d(), d.d(), s), "oei");
      check_launch(qxk_cart_to_ao_mat(&c->shells, a.d(), S, s), "cart_to_ao");
      check_launch(qxk_cart_to_ao_mat(&c->shells, b.d(), T, s), "cart_to_ao");
      check_launch(qxk_cart_to_ao_mat(&c->shells, d.d(), V, s), "cart_to_ao");
      QX_HIP(hipStreamSynchronize(s));
    }
    return QX_OK;
  });
}

int qx_oei_grad(qx_ctx* c, const double* Pd, const double* Wd, double* grad, void* stream) {
  if (!c || !Pd || !Wd || !grad) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nc = c->ncart;
    if (!c->pure) {
      check_launch(qxk_oei_deriv(&c->shells, Pd, Wd, grad, nullptr, nullptr, nullptr, s), "oei_grad");
    } else {
      DevBuf a(nc * nc * 8), b(nc * nc * 8);
      check_launch(qxk_ao_to_cart_mat(&c->shells, Pd, a.d(), s), "ao_to_cart");
      check_launch(qxk_ao_to_cart_mat(&c->shells, Wd, b.d(), s), "ao_to_cart");
      check_launch(qxk_oei_deriv(&c->shells, a.d(), b.d(), grad, nullptr, nullptr, nullptr, s), "oei_grad");
      QX_HIP(hipStreamSynchronize(s));
    }
    return QX_OK;
  });
}

// One-electron piece of the Hessian-vector product: grad[3N] +=
// sum_pq Pd_pq d/dv d(T+V)_pq/dR + Wd_pq d/dv dS_pq/dR (tangent mode of the
// one-electron gradient kernel; v: 3N device vector).
int qx_oei_hess_dir(qx_ctx* c, const double* Pd, const double* Wd, const double* v, double* grad, void* stream) {
  if (!c || !Pd || !Wd || !v || !grad) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t nc = c->ncart;
    if (!c->pure) {
      check_launch(qxk_oei_hess_dir(&c->shells, Pd, Wd, v, grad, s), "oei_hess_dir");
    } else {
      DevBuf a(nc * nc * 8), b(nc * nc * 8);
      check_launch(qxk_ao_to_cart_mat(&c->shells, Pd, a.d(), s), "ao_to_cart");
      check_launch(qxk_ao_to_cart_mat(&c->shells, Wd, b.d(), s), "ao_to_cart");
      check_launch(qxk_oei_hess_dir(&c->shells, a.d(), b.d(), v, grad, s), "oei_hess_dir");
      QX_HIP(hipStreamSynchronize(s));
    }
    return QX_OK;
  });
}

int qx_oei_deriv(qx_ctx* c, double* dS, double* dT, double* dV, void* stream) {
  if (!c || !dS || !dT || !dV) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    const long long m = 3LL * c->natom, nc2 = (long long)c->ncart * c->ncart, nb2 = (long long)c->nbf * c->nbf;
    std::unique_ptr<DevBuf> tmp(c->pure ? new DevBuf((size_t)(3 * m * nc2) * sizeof(double)) : nullptr);
    double* cS = c->pure ? tmp->d() : dS;
    double* cT = c->pure ? tmp->d() + m * nc2 : dT;
    double* cV = c->pure ? tmp->d() + 2 * m * nc2 : dV;
    check_launch(qxk_fill(cS, m * nc2, 0.0, s), "fill");
    check_launch(qxk_fill(cT, m * nc2, 0.0, s), "fill");
    check_launch(qxk_fill(cV, m * nc2, 0.0, s), "fill");
    check_launch(qxk_oei_deriv(&c->shells, nullptr, nullptr, nullptr, cS, cT, cV, s), "oei_deriv");
    if (c->pure) {
      // every coordinate slice Cartesian -> AO
      for (long long x = 0; x < m; ++x) {
        check_launch(qxk_cart_to_ao_mat(&c->shells, cS + x * nc2, dS + x * nb2, s), "cart_to_ao");
        check_launch(qxk_cart_to_ao_mat(&c->shells, cT + x * nc2, dT + x * nb2, s), "cart_to_ao");
        check_launch(qxk_cart_to_ao_mat(&c->shells, cV + x * nc2, dV + x * nb2, s), "cart_to_ao");
      }
      QX_HIP(hipStreamSynchronize(s));
    }
    return QX_OK;
  });
}

int qx_store_size(const qx_ctx* c, long long* nd) {
  if (!c || !nd) return QX_EINVAL;
  *nd = c->store_total;
  return QX_OK;
}

int qx_eri_store(qx_ctx* c, double* out, void* stream) {
  if (!c || !out) return QX_EINVAL;
  if (!c->k1) {
    return fail(c, QX_EUNSUPPORTED, "eri_store: single-primitive shells only");
  }
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    hipStream_t s = (hipStream_t)stream;
    QxEpiDev e{};
    e.epi = QX_EPI_STORE; e.nbf = c->ncart; e.natom = c->natom; e.coord = -1; e.out = out;
    for (auto& wh : c->work) check_launch(launch_eri(c, wh.w, wh.cls, e, 0, s, c->pr, c->pp), "eri_store");
    return QX_OK;
  });
}

int qx_store_lists(const qx_ctx* c, int* nlists, int* l4, long long* base, long long* nquart) {
  if (!c || !nlists) return QX_EINVAL;
  *nlists = (int)c->work.size();
  for (size_t k = 0; k < c->work.size(); ++k) {
    const QxWorkHost& wh = c->work[k];
    const QxClassEntry& e = qx_class_table[wh.cls];
    if (l4) {
      l4[4 * k] = e.la; l4[4 * k + 1] = e.lb; l4[4 * k + 2] = e.lc; l4[4 * k + 3] = e.ld;
    }
    if (base) base[k] = wh.w.store_base;
    if (nquart) nquart[k] = wh.w.nquart;
  }
  return QX_OK;
}

int qx_store_quartet_pairs(const qx_ctx* c, int list, int* p1, int* p2) {
  if (!c || !p1 || !p2 || list < 0 || list >= (int)c->work.size()) return QX_EINVAL;
  const QxWorkHost& wh = c->work[list];
  const long long* pre = c->h_prefix.data() + wh.prefix_off;
  const int* kst = c->h_kstart.data() + c->h_kstart_off[list];
  for (int lo = 0; lo < wh.w.nbra; ++lo) {
    const int u1 = wh.w.bra_off + lo;
    for (long long t = pre[lo]; t < pre[lo + 1]; ++t) {
      const int u2 = wh.w.ket_off + kst[lo] + (int)(t - pre[lo]);
      p1[t] = c->h_ppair[u1];
      p2[t] = c->h_ppair[u2];
    }
  }
  return QX_OK;
}

int qx_pair_shells(const qx_ctx* c, int* npairs, int* sa, int* sb) {
  if (!c || !npairs) return QX_EINVAL;
  *npairs = c->npairs;
  for (int i = 0; i < c->npairs; ++i) {
    if (sa) sa[i] = c->h_pair_sa[i];
    if (sb) sb[i] = c->h_pair_sb[i];
  }
  return QX_OK;
}

int qx_num_classes(void) { return QX_NUM_CLASSES; }

int qx_profile_slots(void) { return QX_PROFILE_SLOTS; }

int qx_profile_slot(int slot, int* kind, int* cls, int* part, int* fp32) {
  if (slot < 0 || slot >= QX_PROFILE_SLOTS) return QX_EINVAL;
  int k, c, p = 0, f;
  if (slot < 2 * QX_NUM_CLASSES_MAX) {
    k = 0; f = slot / QX_NUM_CLASSES_MAX; c = slot % QX_NUM_CLASSES_MAX;
  } else if (slot < QX_SLOT_JK0) {
    const int r = slot - 2 * QX_NUM_CLASSES_MAX;
    k = 1; p = r % QX_DPARTS_CAP;
    const int fc = r / QX_DPARTS_CAP;
    f = fc / QX_NUM_CLASSES_MAX; c = fc % QX_NUM_CLASSES_MAX;
  } else {
    const int r = slot - QX_SLOT_JK0;
    k = 2; p = r % QX_JKPARTS_CAP;
    const int fc = r / QX_JKPARTS_CAP;
    f = fc / QX_NUM_CLASSES_MAX; c = fc % QX_NUM_CLASSES_MAX;
  }
  if (kind) *kind = k;
  if (cls) *cls = c;
  if (part) *part = p;
  if (fp32) *fp32 = f;
  return QX_OK;
}

int qx_mixed_quartets(qx_ctx* c, double thr, long long* n64, long long* n32) {
  return qx_mixed_pass_quartets(c, thr, -1, n64, n32);
}

int qx_mixed_pass_quartets(qx_ctx* c, double thr, int pass, long long* n64, long long* n32) {
  if (!c || !(thr >= 0.0 && std::isfinite(thr)) || pass < -1 || pass > 1) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    QX_HIP(hipSetDevice(c->device));
    build_split(c, thr);
    const auto& S = c->split;
    long long a = 0, b = 0;
    for (size_t i = 0; i < S.wm.size(); ++i) {
      const bool f32 = pass < 0 || mixed_mode(c, i, pass) != 0;
      a += f32 ? S.cq64[i] : S.cq64[i] + S.cq32[i];
      b += f32 ? S.cq32[i] : 0;
    }
    if (n64) *n64 = a;
    if (n32) *n32 = b;
    return QX_OK;
  });
}

int qx_set_class_tiles(qx_ctx* c, int pass, const int* block) {
  if (!c || pass < 0 || pass > 1) return QX_EINVAL;
  if (block)
    for (int k = 0; k < QX_NUM_CLASSES; ++k)
      if (block[k] < 64 || block[k] > QX_BLOCK_MAX || block[k] % 64) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) c->tile[pass][k] = block ? block[k] : qx_tile_default[k][pass];
  return QX_OK;
}

int qx_set_grad_wave_reduce(qx_ctx* c, const int* flags) {
  if (!c) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) c->gwred[k] = flags ? (flags[k] != 0) : qx_gwred_default[k];
  return QX_OK;
}

int qx_set_grad_copies(qx_ctx* c, const int* copies) {
  if (!c) return QX_EINVAL;
  if (copies)
    for (int k = 0; k < QX_NUM_CLASSES; ++k)
      if (copies[k] < 1 || copies[k] > QX_GRAD_COPIES) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) c->gcopy[k] = copies ? copies[k] : qx_gcopy_default[k];
  return QX_OK;
}

int qx_set_jk_rows(qx_ctx* c, const int* flags) {
  if (!c) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) c->jkrows[k] = flags ? (flags[k] != 0) : 1;
  return QX_OK;
}

int qx_get_jk_rows(const qx_ctx* c, int* flags) {
  if (!c || !flags) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) flags[k] = c->jkrows[k];
  return QX_OK;
}

int qx_set_grad_window(qx_ctx* c, int atoms) {
  if (!c || atoms < 0 || 3 * atoms > QX_GRAD_LDS) return QX_EINVAL;
  c->grad_win = 3 * atoms;
  return QX_OK;
}

int qx_get_grad_window(const qx_ctx* c, int* atoms) {
  if (!c || !atoms) return QX_EINVAL;
  *atoms = (c->grad_win > 0 ? c->grad_win : QX_GRAD_LDS) / 3;
  return QX_OK;
}

int qx_get_grad_copies(const qx_ctx* c, int* copies) {
  if (!c || !copies) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) copies[k] = c->gcopy[k];
  return QX_OK;
}

int qx_get_grad_wave_reduce(const qx_ctx* c, int* flags) {
  if (!c || !flags) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) flags[k] = c->gwred[k];
  return QX_OK;
}

int qx_get_class_tiles(const qx_ctx* c, int pass, int* block) {
  if (!c || !block || pass < 0 || pass > 1) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) block[k] = c->tile[pass][k];
  return QX_OK;
}

int qx_set_fp32_classes(qx_ctx* c, int pass, const int* flags) {
  if (!c || pass < 0 || pass > 1) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) c->fp32_ok[pass][k] = flags ? (flags[k] != 0) : qx_fp32_default[k][pass];
  return QX_OK;
}

int qx_get_fp32_classes(const qx_ctx* c, int pass, int* flags) {
  if (!c || !flags || pass < 0 || pass > 1) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) flags[k] = c->fp32_ok[pass][k];
  return QX_OK;
}

int qx_set_streams(qx_ctx* c, int n) {
  if (!c || n < 1 || n > 8) return QX_EINVAL;
  c->nstreams = n;
  return QX_OK;
}

int qx_class_info(int cls, int* l4, int* nparts) {
  if (cls < 0 || cls >= QX_NUM_CLASSES) return QX_EINVAL;
  const QxClassEntry& e = qx_class_table[cls];
  if (l4) { l4[0] = e.la; l4[1] = e.lb; l4[2] = e.lc; l4[3] = e.ld; }
  if (nparts) *nparts = e.nparts;
  return QX_OK;
}

int qx_class_flops(int cls, double* eri, double* parts, int* ncomb) {
  if (cls < 0 || cls >= QX_NUM_CLASSES) return QX_EINVAL;
  const QxClassEntry& e = qx_class_table[cls];
  if (eri) *eri = e.eri_flops;
  for (int k = 0; k < e.nparts; ++k) {
    if (parts) parts[k] = e.part_flops[k];
    if (ncomb) ncomb[k] = e.part_ncomb[k];
  }
  return QX_OK;
}

int qx_max_parts(void) { return QX_DPARTS_CAP; }

int qx_class_jk(int cls, int* nparts, double* flops) {
  if (cls < 0 || cls >= QX_NUM_CLASSES) return QX_EINVAL;
  const QxClassEntry& e = qx_class_table[cls];
  if (nparts) *nparts = e.njk;
  for (int k = 0; k < e.njk; ++k)
    if (flops) flops[k] = e.jk_flops[k];
  return QX_OK;
}

int qx_class_groups(int cls, int* ngroups, double* flops, int* nparts) {
  if (cls < 0 || cls >= QX_NUM_CLASSES) return QX_EINVAL;
  const QxClassEntry& e = qx_class_table[cls];
  if (ngroups) *ngroups = e.ngroups;
  for (int k = 0; k < e.ngroups; ++k) {
    if (flops) flops[k] = e.group_flops[k];
    if (nparts) nparts[k] = e.group_parts[k];
  }
  return QX_OK;
}

int qx_class_quartets(const qx_ctx* c, long long* eri, long long* deri) {
  if (!c) return QX_EINVAL;
  for (int k = 0; k < QX_NUM_CLASSES; ++k) {
    if (eri) eri[k] = c->nq_class[k];
    if (deri) deri[k] = c->ndq_class[k];
  }
  return QX_OK;
}

int qx_profile_enable(qx_ctx* c, int on) {
  if (!c) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    if (!c->prof_pending.empty()) {
      const int rc = qx_profile_read(c, nullptr, nullptr, nullptr);
      if (rc != QX_OK) return rc;
    }
    c->profiling = on != 0;
    c->prof_ms.assign(QX_PROFILE_SLOTS, 0.0);
    c->prof_launches.assign(QX_PROFILE_SLOTS, 0);
    c->prof_pq.assign(QX_PROFILE_SLOTS, 0);
    return QX_OK;
  });
}

int qx_profile_read(qx_ctx* c, double* ms, long long* launches, long long* pq) {
  if (!c) return QX_EINVAL;
  return guarded(c, [&]() -> int {
    if (c->prof_ms.size() != QX_PROFILE_SLOTS) {
      c->prof_ms.assign(QX_PROFILE_SLOTS, 0.0);
      c->prof_launches.assign(QX_PROFILE_SLOTS, 0);
      c->prof_pq.assign(QX_PROFILE_SLOTS, 0);
    }
    for (auto& p : c->prof_pending) {
      QX_HIP(hipEventSynchronize(p.stop));
      float t = 0.0f;
      QX_HIP(hipEventElapsedTime(&t, p.start, p.stop));
      c->prof_ms[p.slot] += t;
      c->event_pool.push_back(p.start);
      c->event_pool.push_back(p.stop);
    }
    c->prof_pending.clear();
    for (int k = 0; k < QX_PROFILE_SLOTS; ++k) {
      if (ms) ms[k] = c->prof_ms[k];
      if (launches) launches[k] = c->prof_launches[k];
      if (pq) pq[k] = c->prof_pq[k];
    }
    return QX_OK;
  });
}

}  // extern "C"
