// Device-side definitions shared by the generated per-class ERI kernels:
// argument structs, the Boys function, quartet lookup and the epilogues.
//
// Work decomposition.  Screened shell pairs are sorted by class and, within a
// class, by descending Schwarz bound.  Each contracted pair expands into its
// primitive pairs ("pseudo pairs", stored contiguously per pair).  One thread
// evaluates one pseudo-pair quartet; a contracted quartet is the sum of its
// pseudo-pair quartets, so every epilogue that is linear in the integrals
// (J/K, gradient, tensor scatter with atomics) contracts by accumulation.
// For a bra pseudo pair i the admissible ket pseudo pairs form a prefix of the
// ket list (Schwarz-sorted, canonical j <= i within one class), so a work list
// is a running-count array over bra pseudo pairs and thread t finds its bra by
// binary search.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define QX_BLOCK 128      // threads per block of the array-form kernels; default of the streaming ones
#define QX_BLOCK_MAX 256  // largest block of the streaming J/K and gradient kernels (a multiple of 64)
#define QX_TWO_PI_5_2 34.986836655249725  // 2 pi^(5/2)
#define QX_PI 3.14159265358979323846

// Boys table: F_m(T_g) for T_g = g / QX_BOYS_INV_H, g = 0..QX_BOYS_NPTS-1,
// m = 0..QX_BOYS_NM-1.  A seven-term Taylor expansion about the nearest grid
// point (|T - T_g| <= 1/64) gives F_M; lower orders follow by downward
// recursion.  Above QX_BOYS_TMAX: asymptotic F_0 and upward recursion.
#define QX_BOYS_TMAX 36.0
#define QX_BOYS_INV_H 32.0
#define QX_BOYS_NM 24
#define QX_BOYS_NPTS (36 * 32 + 2)
#define QX_BOYS_MMAX (QX_BOYS_NM - 7)

// largest Cartesian shell-pair block (f x f); sizes the Schwarz diagonal rows
#define QX_PAIR_STRIDE 100

enum QxEpi {
  QX_EPI_STORE = 0,    // out[store_base + comp * nquart + t] (single-primitive bases)
  QX_EPI_FULL = 1,     // scatter all distinct permutations into the nbf^4 tensor
  QX_EPI_SCHWARZ = 2,  // diagonal (f1 f2|f1 f2) of pair P into out[P * stride + f1*nb + f2]
  QX_EPI_JK = 3,       // direct J', K' accumulation with the density
  QX_EPI_DFULL = 4,    // derivative tensor scatter (all 3N slices, or one)
  QX_EPI_GRAD = 5,     // adjoint gradient: sum dERI * density pair products
  QX_EPI_JKD = 6       // J', K' of the directional derivative sum_y tv_y d(ab|cd)/dR_y
};

struct QxPairsDev {        // screened shell pairs in sorted order (SoA)
  const double* Ax; const double* Ay; const double* Az;      // centre of shell a
  const double* ABx; const double* ABy; const double* ABz;   // A - B
  const int* sa; const int* sb;                              // shell indices
  const int* bfa; const int* bfb;                            // first Cartesian function
  const int* atom_a; const int* atom_b;
  const int* pp_first; const int* pp_n;                      // pseudo-pair range
};

struct QxPrimPairsDev {    // pseudo pairs (one primitive pair each), SoA
  const int* pair;                                           // owning pair
  const double* p;                                           // alpha + beta
  const double* Px; const double* Py; const double* Pz;      // product centre
  const double* K;      // c_a c_b exp(-alpha beta |AB|^2 / p)
  const double* a2;     // 2 alpha
  const double* b2;     // 2 beta
};

struct QxWorkDev {         // one (bra class, ket class) quartet list
  int bra_off, nbra;         // bra pseudo-pair range
  int ket_off, nket;         // ket pseudo-pair range
  const long long* prefix;   // nbra+1 running quartet counts
  const int* kstart;         // per bra pseudo pair: first admissible ket (null = 0)
  long long nquart;
  long long store_base;      // element offset of this list in the store buffer
  int diag;                  // kets are the pseudo pairs of the bra's own pair
  // Mixed precision over one list: the first nk64[row] kets of each bra row
  // (Schwarz bound >= the FP32 threshold; kets are sorted by descending
  // bound) run in FP64, the rest in FP32.  sel 0: every quartet; 1: the FP64
  // kets only; 2: the FP32 kets only.
  const int* nk64;
  int sel;
};

struct QxEpiDev {
  int epi;
  int atomic;                // STORE/FULL: accumulate with atomics
  int nbf, natom;
  int coord;                 // DFULL: -1 = all 3N slices, else one coordinate
  double* out;
  const double* D;           // density, Cartesian nbf x nbf (JK, GRAD)
  double* J; double* K;      // unsymmetrised J', K' (JK)
  int krows;                 // JK: stage K' rows of a block-uniform bra pair in LDS
  int kcols;                 // JK: width of those LDS rows (widest fragment's function range; 0 = nbf)
  const int* frag_col0;      // JK: first function of each fragment's range (null: 0); with pair_frag
  double* grad;              // GRAD: grad_copies x 3N accumulators (copy = block index % grad_copies)
  int grad_copies;           // GRAD: copies of the 3N accumulators (0 or 1: one)
  int grad_wred;             // GRAD: sum wave-uniform centres across the wavefront before LDS
  int grad_win;              // GRAD: widest LDS atom window in doubles (0 or >= QX_GRAD_LDS: QX_GRAD_LDS)
  const double* tv;          // 3N geometry direction: tangent kernels' seed, JKD weights
  const int* slot_map;       // DFULL: output slice of coordinate y (-1: none); null = coord rule
  int skip_single;           // FULL: leave out quartets with all four centres on one atom
  const int* frag_skip;      // JK: per fragment, nonzero = leave its quartets out (null: none)
  const int* pair_frag;      // JK: fragment of each sorted shell pair (with frag_skip / frag_col0)
};

struct QxArgs {
  QxPairsDev pr;
  QxPrimPairsDev pp;
  QxWorkDev w;
  QxEpiDev e;
  const double* boys;
};

struct QxLaunch {
  int fp32;                  // recursion in FP32 (JK / GRAD epilogues only)
  hipStream_t stream;
  int block;                 // JK / GRAD: threads per block (multiple of 64, <= QX_BLOCK_MAX; 0 = QX_BLOCK)
};

__host__ __device__ inline int qx_launch_block(const QxLaunch* L) { return L->block > 0 ? L->block : QX_BLOCK; }

template <int M>
__device__ __forceinline__ void qx_boys(double T, const double* __restrict__ tab, double (&F)[M + 1]) {
  static_assert(M <= QX_BOYS_MMAX, "Boys order beyond table");
  if (T < QX_BOYS_TMAX) {
    const int g = (int)(T * QX_BOYS_INV_H + 0.5);
    const double d = g * (1.0 / QX_BOYS_INV_H) - T;
    const double* r = tab + g * QX_BOYS_NM + M;
    double f = r[6] * (1.0 / 720.0);
    f = f * d + r[5] * (1.0 / 120.0);
    f = f * d + r[4] * (1.0 / 24.0);
    f = f * d + r[3] * (1.0 / 6.0);
    f = f * d + r[2] * 0.5;
    f = f * d + r[1];
    f = f * d + r[0];
    F[M] = f;
    if constexpr (M > 0) {
      const double e = exp(-T);
      const double t2 = 2.0 * T;
#pragma unroll
      for (int m = M - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * (1.0 / (2 * m + 1));
    }
  } else {
    F[0] = 0.5 * sqrt(QX_PI / T);
    if constexpr (M > 0) {
      const double e = exp(-T);
      const double o2t = 0.5 / T;
#pragma unroll
      for (int m = 0; m < M; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * o2t;
    }
  }
}

// ---- forward-mode (tangent) arithmetic --------------------------------
// A value and its derivative along one geometry direction tv (3N).  Tangent
// kernels run the generated recurrences in this type: the derivative part of
// a first-derivative integral is then a second derivative, contracted with
// tv on one index.
struct qx_dual {
  double v, d;
  __device__ __forceinline__ qx_dual() = default;
  __device__ __forceinline__ constexpr qx_dual(double x) : v(x), d(0.0) {}
  __device__ __forceinline__ constexpr qx_dual(double x, double y) : v(x), d(y) {}
};
__device__ __forceinline__ qx_dual operator+(qx_dual a, qx_dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ qx_dual operator-(qx_dual a, qx_dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ qx_dual operator-(qx_dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ qx_dual operator*(qx_dual a, qx_dual b) { return {a.v * b.v, a.v * b.d + a.d * b.v}; }
__device__ __forceinline__ qx_dual operator*(double a, qx_dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ qx_dual operator*(qx_dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ qx_dual operator+(double a, qx_dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ qx_dual operator+(qx_dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ qx_dual operator-(qx_dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ qx_dual operator-(double a, qx_dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ qx_dual operator/(qx_dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ qx_dual& operator+=(qx_dual& a, qx_dual b) { a = a + b; return a; }
__device__ __forceinline__ qx_dual& operator-=(qx_dual& a, qx_dual b) { a = a - b; return a; }
__device__ __forceinline__ qx_dual& operator*=(qx_dual& a, qx_dual b) { a = a * b; return a; }
// the derivative part (what a tangent kernel writes out)
__device__ __forceinline__ double qx_tan(qx_dual x) { return x.d; }

// F_m(T) along the tangent: dF_m/dT = -F_{m+1}
template <int M>
__device__ __forceinline__ void qx_boys_tan(qx_dual T, const double* __restrict__ tab, qx_dual (&F)[M + 1]) {
  double f[M + 2];
  qx_boys<M + 1>(T.v, tab, f);
#pragma unroll
  for (int m = 0; m <= M; ++m) F[m] = qx_dual(f[m], -T.d * f[m + 1]);
}

// Pair data of pseudo pair u (shell pair P) with its derivative along tv:
// centres move with their atoms, P = (alpha A + beta B) / p,
// K = c_a c_b exp(-alpha beta / p |A - B|^2).
struct QxTanPair {
  qx_dual K, P[3], A[3], AB[3];
};
__device__ __forceinline__ QxTanPair qx_tan_pair(const QxArgs& A_, int P, int u) {
  const double* tv = A_.e.tv;
  const int ta = 3 * A_.pr.atom_a[P], tb = 3 * A_.pr.atom_b[P];
  const double p = A_.pp.p[u], al = 0.5 * A_.pp.a2[u], be = 0.5 * A_.pp.b2[u];
  const double ab[3] = {A_.pr.ABx[P], A_.pr.ABy[P], A_.pr.ABz[P]};
  const double pv[3] = {A_.pp.Px[u], A_.pp.Py[u], A_.pp.Pz[u]};
  const double av[3] = {A_.pr.Ax[P], A_.pr.Ay[P], A_.pr.Az[P]};
  QxTanPair r;
  double abd = 0.0;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const double dA = tv[ta + x], dB = tv[tb + x];
    r.A[x] = qx_dual(av[x], dA);
    r.AB[x] = qx_dual(ab[x], dA - dB);
    r.P[x] = qx_dual(pv[x], (al * dA + be * dB) / p);
    abd += ab[x] * (dA - dB);
  }
  const double K = A_.pp.K[u];
  r.K = qx_dual(K, -2.0 * al * be / p * abd * K);
  return r;
}

// Quartet t of the list: bra / ket pseudo pairs u1 / u2 and their pairs.
// Returns whether the quartet belongs to this launch's precision (w.sel).
__device__ __forceinline__ bool qx_locate(const QxArgs& A_, long long t, int& u1, int& u2, int& P1, int& P2) {
  const long long* pre = A_.w.prefix;
  int lo = 0, hi = A_.w.nbra;  // pre[lo] <= t < pre[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= t) lo = mid; else hi = mid;
  }
  const int local = (int)(t - pre[lo]);
  u1 = A_.w.bra_off + lo;
  P1 = A_.pp.pair[u1];
  u2 = A_.w.diag ? A_.pr.pp_first[P1] + local : A_.w.ket_off + (A_.w.kstart ? A_.w.kstart[lo] : 0) + local;
  P2 = A_.pp.pair[u2];
  return A_.w.sel == 0 || ((local < A_.w.nk64[lo]) == (A_.w.sel == 1));
}

// true when no thread of the block holds a quartet of this launch's
// precision and of a fragment not left out (block-uniform; every thread must
// call it)
__device__ __forceinline__ bool qx_block_idle(const QxArgs& A_, bool active) {
  return (A_.w.sel != 0 || A_.e.frag_skip) && !__syncthreads_or(active ? 1 : 0);
}

__device__ __forceinline__ void qx_atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// true when every active lane of the wavefront holds the same key
__device__ __forceinline__ bool qx_wave_uniform(int key, bool active) {
  int kmin = active ? key : 0x7fffffff, kmax = active ? key : (int)0x80000000;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    kmin = min(kmin, __shfl_xor(kmin, off));
    kmax = max(kmax, __shfl_xor(kmax, off));
  }
  return kmin == kmax;
}

__device__ __forceinline__ double qx_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// lane that performs a wave-reduced store: the first active lane
__device__ __forceinline__ bool qx_wave_leader(bool active) {
  const unsigned long long m = __ballot(active);
  return m != 0 && (int)__lane_id() == __ffsll((long long)m) - 1;
}

__device__ __forceinline__ void qx_put(double* p, double v, bool atomic) {
  if (atomic) qx_atomic_add(p, v); else *p = v;
}

// Degeneracy weight of a canonical quartet: 1/2 for each of a==b, c==d and
// (ab)==(cd), so that summing all 8 index permutations with this weight
// counts every distinct permutation once.
__device__ __forceinline__ double qx_degeneracy(const QxArgs& A_, int P1, int P2) {
  double f = 1.0;
  if (A_.pr.sa[P1] == A_.pr.sb[P1]) f *= 0.5;
  if (A_.pr.sa[P2] == A_.pr.sb[P2]) f *= 0.5;
  if (P1 == P2) f *= 0.5;
  return f;
}

// Writes v to every distinct index permutation of (I J | K L).  Permutations
// that coincide with another element of the same shell block (a==b, c==d,
// (ab)==(cd)) are skipped: that element writes them itself.
__device__ __forceinline__ void qx_scatter8(double* out, size_t n, size_t I, size_t J, size_t K, size_t L, double v,
                                            bool sab, bool scd, bool spair, bool atomic) {
  qx_put(&out[((I * n + J) * n + K) * n + L], v, atomic);
  if (!sab) qx_put(&out[((J * n + I) * n + K) * n + L], v, atomic);
  if (!scd) qx_put(&out[((I * n + J) * n + L) * n + K], v, atomic);
  if (!sab && !scd) qx_put(&out[((J * n + I) * n + L) * n + K], v, atomic);
  if (!spair) {
    qx_put(&out[((K * n + L) * n + I) * n + J], v, atomic);
    if (!scd) qx_put(&out[((L * n + K) * n + I) * n + J], v, atomic);
    if (!sab) qx_put(&out[((K * n + L) * n + J) * n + I], v, atomic);
    if (!sab && !scd) qx_put(&out[((L * n + K) * n + J) * n + I], v, atomic);
  }
}

// ---- streaming-kernel epilogues -----------------------------------------
// Flushes per-thread J'/K' partials: J'_ab += 4f sum_cd (ab|cd) D_cd etc.
// Every lane calls this (converged); inactive lanes hold zeros.
// sk != null: the whole block shares the bra pair, and K'_ac, K'_ad, K'_bc,
// K'_bd go to its LDS rows (a: 0..NA-1, b: NA..NA+NB-1; column j of a row is
// function col0 + j, the bra fragment's range, where every ket of the block
// lies) instead of global atomics.
// Only rows a in [A0, A1) of J'_ab, K'_ac and K'_ad can be non-zero (a J/K
// part covers that range of a components).
template <int NA, int NB, int NC, int ND, int A0, int A1>
__device__ __forceinline__ void qx_jk_flush(const QxArgs& A_, int P1, int P2, bool active, double* sk,
                                            int W, int col0,
                                            const double (&jab)[NA * NB], const double (&jcd)[NC * ND],
                                            const double (&kac)[NA * NC], const double (&kad)[NA * ND],
                                            const double (&kbc)[NB * NC], const double (&kbd)[NB * ND]) {
  const bool uni = qx_wave_uniform(P1, active);  // wave-uniform value
  const bool lead = qx_wave_leader(active);
  const double f = active ? qx_degeneracy(A_, P1, P2) : 0.0;
  const double fj = 4.0 * f, fk = 2.0 * f;
  const int n = A_.e.nbf;
  double* J = A_.e.J;
  double* K = A_.e.K;
  int ia = 0, ib = 0, ic = 0, id = 0;
  if (active) {
    ia = A_.pr.bfa[P1]; ib = A_.pr.bfb[P1]; ic = A_.pr.bfa[P2]; id = A_.pr.bfb[P2];
  }
  if (uni) {
#pragma unroll
    for (int k = A0 * NB; k < A1 * NB; ++k) {
      const double v = qx_wave_sum(fj * jab[k]);
      if (lead) qx_atomic_add(&J[(ia + k / NB) * n + ib + k % NB], v);
    }
  } else if (active) {
#pragma unroll
    for (int k = A0 * NB; k < A1 * NB; ++k) qx_atomic_add(&J[(ia + k / NB) * n + ib + k % NB], fj * jab[k]);
  }
  if (!active) return;
#pragma unroll
  for (int k = 0; k < NC * ND; ++k) qx_atomic_add(&J[(ic + k / ND) * n + id + k % ND], fj * jcd[k]);
  if (sk) {
    const int jc = ic - col0, jd = id - col0;
#pragma unroll
    for (int k = A0 * NC; k < A1 * NC; ++k) atomicAdd(&sk[(k / NC) * W + jc + k % NC], fk * kac[k]);
#pragma unroll
    for (int k = A0 * ND; k < A1 * ND; ++k) atomicAdd(&sk[(k / ND) * W + jd + k % ND], fk * kad[k]);
#pragma unroll
    for (int k = 0; k < NB * NC; ++k) atomicAdd(&sk[(NA + k / NC) * W + jc + k % NC], fk * kbc[k]);
#pragma unroll
    for (int k = 0; k < NB * ND; ++k) atomicAdd(&sk[(NA + k / ND) * W + jd + k % ND], fk * kbd[k]);
    return;
  }
#pragma unroll
  for (int k = A0 * NC; k < A1 * NC; ++k) qx_atomic_add(&K[(ia + k / NC) * n + ic + k % NC], fk * kac[k]);
#pragma unroll
  for (int k = A0 * ND; k < A1 * ND; ++k) qx_atomic_add(&K[(ia + k / ND) * n + id + k % ND], fk * kad[k]);
#pragma unroll
  for (int k = 0; k < NB * NC; ++k) qx_atomic_add(&K[(ib + k / NC) * n + ic + k % NC], fk * kbc[k]);
#pragma unroll
  for (int k = 0; k < NB * ND; ++k) qx_atomic_add(&K[(ib + k / ND) * n + id + k % ND], fk * kbd[k]);
}

// Spreads NCOMB per-(centre, direction) sums to the 12 centre x direction
// slots (d/dD from translational invariance when DINV), weighted by w.
template <int CMASK, int XMASK, int DINV, int NCOMB>
__device__ __forceinline__ void qx_grad_spread(double w, const double (&g)[NCOMB], double (&gv)[12]) {
  int ci = 0;
  double dsum[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int cen = 0; cen < 4; ++cen) {
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      double v = 0.0;
      if ((CMASK & (1 << cen)) && (XMASK & (1 << x))) v = w * g[ci++];
      dsum[x] += v;
      gv[3 * cen + x] = v;
    }
  }
  if (DINV) {
#pragma unroll
    for (int x = 0; x < 3; ++x) gv[9 + x] = -dsum[x];
  }
}

// ---- ERI epilogues ----------------------------------------------------
template <int NA, int NB, int NC, int ND, typename R>
__device__ __forceinline__ void qx_epi_eri(const QxArgs& A_, long long t, int P1, int P2,
                                           const R (&o)[NA * NB * NC * ND], bool active) {
  constexpr int N = NA * NB * NC * ND;
  const int epi = A_.e.epi;
  if (!active) return;
  if (epi == QX_EPI_STORE) {
    double* out = A_.e.out + A_.w.store_base + t;
    const long long nq = A_.w.nquart;
#pragma unroll
    for (int k = 0; k < N; ++k) out[k * nq] = (double)o[k];
  } else if (epi == QX_EPI_SCHWARZ) {
    double* out = A_.e.out + (size_t)P1 * QX_PAIR_STRIDE;
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) qx_atomic_add(&out[i * NB + j], (double)o[((i * NB + j) * NC + i) * ND + j]);
  } else if (epi == QX_EPI_FULL) {
    const size_t n = A_.e.nbf;
    const int ia = A_.pr.bfa[P1], ib = A_.pr.bfb[P1], ic = A_.pr.bfa[P2], id = A_.pr.bfb[P2];
    const bool sab = A_.pr.sa[P1] == A_.pr.sb[P1], scd = A_.pr.sa[P2] == A_.pr.sb[P2], sp = P1 == P2;
    const bool atomic = A_.e.atomic != 0;
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int k = 0; k < NC; ++k)
#pragma unroll
          for (int l = 0; l < ND; ++l)
            qx_scatter8(A_.e.out, n, ia + i, ib + j, ic + k, id + l, (double)o[((i * NB + j) * NC + k) * ND + l],
                        sab, scd, sp, atomic);
  }
}

// ---- derivative epilogues --------------------------------------------
// dv holds NCOMB blocks of N, ordered over centres in CMASK (A=1, B=2, C=4,
// D=8) then Cartesian directions in XMASK.  With DINV the part covers A, B, C
// in all directions and d/dD = -(d/dA + d/dB + d/dC).
__device__ __forceinline__ bool qx_single_atom(const QxArgs& A_, int P1, int P2) {
  const int a = A_.pr.atom_a[P1];
  return a == A_.pr.atom_b[P1] && a == A_.pr.atom_a[P2] && a == A_.pr.atom_b[P2];
}

// Output slice of nuclear coordinate y in a derivative-tensor scatter:
// slot_map[y] when given (-1 = not wanted), else every slice (coord < 0) or
// only coordinate `coord` into slice 0.
__device__ __forceinline__ long long qx_dfull_slice(const QxEpiDev& e, int y) {
  if (e.slot_map) return e.slot_map[y];
  if (e.coord >= 0) return y == e.coord ? 0 : -1;
  return y;
}

// A part covers the output functions with a in [A0, A1), b in [B0, B1) and c
// in [C0, C1): dv holds NCOMB blocks of that part's NEL elements, in
// (a, b, c, d) order.
template <int NA, int NB, int NC, int ND, int A0, int A1, int B0, int B1, int C0, int C1>
struct QxPartShape {
  static constexpr int NEL = (A1 - A0) * (B1 - B0) * (C1 - C0) * ND;
};

template <int NA, int NB, int NC, int ND, int CMASK, int XMASK, int DINV, int NCOMB, int A0, int A1, int B0, int B1,
          int C0, int C1, typename R>
__device__ __forceinline__ void qx_epi_deri(
    const QxArgs& A_, int P1, int P2, const R (&dv)[NCOMB * QxPartShape<NA, NB, NC, ND, A0, A1, B0, B1, C0, C1>::NEL]) {
  constexpr int NEL = QxPartShape<NA, NB, NC, ND, A0, A1, B0, B1, C0, C1>::NEL;
  const int at[4] = {A_.pr.atom_a[P1], A_.pr.atom_b[P1], A_.pr.atom_a[P2], A_.pr.atom_b[P2]};
  const int n = A_.e.nbf;
  const int ia = A_.pr.bfa[P1], ib = A_.pr.bfb[P1], ic = A_.pr.bfa[P2], id = A_.pr.bfb[P2];
  {  // QX_EPI_DFULL
    const size_t nn = n;
    const size_t n4 = nn * nn * nn * nn;
    const bool sab = A_.pr.sa[P1] == A_.pr.sb[P1], scd = A_.pr.sa[P2] == A_.pr.sb[P2], sp = P1 == P2;
#pragma unroll
    for (int i = A0; i < A1; ++i)
#pragma unroll
      for (int j = B0; j < B1; ++j)
#pragma unroll
        for (int k = C0; k < C1; ++k)
#pragma unroll
          for (int l = 0; l < ND; ++l) {
            const int e = ((((i - A0) * (B1 - B0) + (j - B0)) * (C1 - C0)) + (k - C0)) * ND + l;
            double dsum[3] = {0.0, 0.0, 0.0};
            int ci = 0;
#pragma unroll
            for (int cen = 0; cen < 4; ++cen) {
              if (!(CMASK & (1 << cen))) continue;
#pragma unroll
              for (int x = 0; x < 3; ++x) {
                if (!(XMASK & (1 << x))) continue;
                const double v = (double)dv[(ci++) * NEL + e];
                dsum[x] += v;
                const long long sl = qx_dfull_slice(A_.e, 3 * at[cen] + x);
                if (sl < 0) continue;
                qx_scatter8(A_.e.out + (size_t)sl * n4, nn, ia + i, ib + j, ic + k, id + l, v, sab, scd, sp, true);
              }
            }
            if (DINV) {
#pragma unroll
              for (int x = 0; x < 3; ++x) {
                const long long sl = qx_dfull_slice(A_.e, 3 * at[3] + x);
                if (sl < 0) continue;
                qx_scatter8(A_.e.out + (size_t)sl * n4, nn, ia + i, ib + j, ic + k, id + l, -dsum[x], sab, scd, sp,
                            true);
              }
            }
          }
  }
}

// Derivative values of one element e (local index) of a part spread to the
// 12 centre x direction slots; d/dD from translational invariance (DINV).
template <int CMASK, int XMASK, int DINV, int NCOMB, int NEL>
__device__ __forceinline__ void qx_elem_spread(const double (&dv)[NCOMB * NEL], int e, double (&gv)[12]) {
  int ci = 0;
  double dsum[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int cen = 0; cen < 4; ++cen)
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      double v = 0.0;
      if ((CMASK & (1 << cen)) && (XMASK & (1 << x))) v = dv[(ci++) * NEL + e];
      dsum[x] += v;
      gv[3 * cen + x] = v;
    }
  if (DINV) {
#pragma unroll
    for (int x = 0; x < 3; ++x) gv[9 + x] = -dsum[x];
  }
}

// Adjoint contraction of an array of derivative values (tangent kernels):
// gv[12] = 8 f sum_e Gamma_e dv_e, Gamma_ijkl = 2 D_ij D_kl - (D_ik D_jl +
// D_il D_jk) / 2 with the degeneracy weight f of the canonical quartet.
template <int NA, int NB, int NC, int ND, int CMASK, int XMASK, int DINV, int NCOMB, int A0, int A1, int B0, int B1,
          int C0, int C1>
__device__ __forceinline__ void qx_epi_dgrad(
    const QxArgs& A_, int P1, int P2,
    const double (&dv)[NCOMB * QxPartShape<NA, NB, NC, ND, A0, A1, B0, B1, C0, C1>::NEL], double (&gv)[12]) {
  constexpr int NEL = QxPartShape<NA, NB, NC, ND, A0, A1, B0, B1, C0, C1>::NEL;
  const size_t n = A_.e.nbf;
  const double* D = A_.e.D;
  const int ia = A_.pr.bfa[P1], ib = A_.pr.bfb[P1], ic = A_.pr.bfa[P2], id = A_.pr.bfb[P2];
#pragma unroll
  for (int k = 0; k < 12; ++k) gv[k] = 0.0;
  for (int i = A0; i < A1; ++i)
    for (int j = B0; j < B1; ++j)
      for (int k = C0; k < C1; ++k)
        for (int l = 0; l < ND; ++l) {
          const int e = ((((i - A0) * (B1 - B0) + (j - B0)) * (C1 - C0)) + (k - C0)) * ND + l;
          const size_t I = ia + i, J = ib + j, K = ic + k, L = id + l;
          const double G = 2.0 * D[I * n + J] * D[K * n + L] - 0.5 * (D[I * n + K] * D[J * n + L] +
                                                                      D[I * n + L] * D[J * n + K]);
          double ge[12];
          qx_elem_spread<CMASK, XMASK, DINV, NCOMB, NEL>(dv, e, ge);
#pragma unroll
          for (int c = 0; c < 12; ++c) gv[c] += G * ge[c];
        }
  const double w = 8.0 * qx_degeneracy(A_, P1, P2);
#pragma unroll
  for (int k = 0; k < 12; ++k) gv[k] *= w;
}

// J', K' of the directional derivative of each integral, s = sum over the
// 12 centre x direction slots of tv[3 atom + x] d(ab|cd), with the J/K
// kernels' accumulation rule (J'_ab += 4 f s D_cd, K'_ac += 2 f s D_bd, ...).
template <int NA, int NB, int NC, int ND, int CMASK, int XMASK, int DINV, int NCOMB, int A0, int A1, int B0, int B1,
          int C0, int C1>
__device__ __forceinline__ void qx_epi_djk(
    const QxArgs& A_, int P1, int P2,
    const double (&dv)[NCOMB * QxPartShape<NA, NB, NC, ND, A0, A1, B0, B1, C0, C1>::NEL]) {
  constexpr int NEL = QxPartShape<NA, NB, NC, ND, A0, A1, B0, B1, C0, C1>::NEL;
  const size_t n = A_.e.nbf;
  const double* D = A_.e.D;
  const double* tv = A_.e.tv;
  const int at[4] = {A_.pr.atom_a[P1], A_.pr.atom_b[P1], A_.pr.atom_a[P2], A_.pr.atom_b[P2]};
  double w[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) w[k] = tv[3 * at[k / 3] + k % 3];
  const int ia = A_.pr.bfa[P1], ib = A_.pr.bfb[P1], ic = A_.pr.bfa[P2], id = A_.pr.bfb[P2];
  const double f = qx_degeneracy(A_, P1, P2), fj = 4.0 * f, fk = 2.0 * f;
  double* J = A_.e.J;
  double* K = A_.e.K;
  for (int i = A0; i < A1; ++i)
    for (int j = B0; j < B1; ++j)
      for (int k = C0; k < C1; ++k)
        for (int l = 0; l < ND; ++l) {
          const int e = ((((i - A0) * (B1 - B0) + (j - B0)) * (C1 - C0)) + (k - C0)) * ND + l;
          double ge[12];
          qx_elem_spread<CMASK, XMASK, DINV, NCOMB, NEL>(dv, e, ge);
          double s = 0.0;
#pragma unroll
          for (int c = 0; c < 12; ++c) s += w[c] * ge[c];
          if (s == 0.0) continue;
          const size_t I = ia + i, Jx = ib + j, Kx = ic + k, L = id + l;
          qx_atomic_add(&J[I * n + Jx], fj * s * D[Kx * n + L]);
          qx_atomic_add(&J[Kx * n + L], fj * s * D[I * n + Jx]);
          qx_atomic_add(&K[I * n + Kx], fk * s * D[Jx * n + L]);
          qx_atomic_add(&K[I * n + L], fk * s * D[Jx * n + Kx]);
          qx_atomic_add(&K[Jx * n + Kx], fk * s * D[I * n + L]);
          qx_atomic_add(&K[Jx * n + L], fk * s * D[I * n + Kx]);
        }
}

// ---- kernel definitions (instantiated once per generated file) --------
#define QX_GRID(nq) dim3((unsigned)(((nq) + QX_BLOCK - 1) / QX_BLOCK))

#define QX_GRAD_LDS 768  // block-level gradient accumulators (3 x the atoms of a block's window)

// Array-form ERI kernel: store / full-tensor scatter / Schwarz epilogues.
#define QX_DEFINE_ERI_KERNEL(NAME, N, NA, NB, NC, ND)                                             \
  __global__ __launch_bounds__(QX_BLOCK) void qx_k_##NAME(QxArgs A_) {                            \
    const long long t = (long long)blockIdx.x * QX_BLOCK + threadIdx.x;                          \
    if (t >= A_.w.nquart) return;                                                                 \
    int u1, u2, P1, P2;                                                                           \
    qx_locate(A_, t, u1, u2, P1, P2);                                                             \
    if (A_.e.skip_single && qx_single_atom(A_, P1, P2)) return;                                   \
    double o[N];                                                                                  \
    qx_core_##NAME<double>(A_, P1, P2, u1, u2, o);                                                \
    qx_epi_eri<NA, NB, NC, ND, double>(A_, t, P1, P2, o, true);                                   \
  }                                                                                               \
  extern "C" int qx_launch_##NAME(const QxArgs* a, const QxLaunch* L) {                           \
    if (a->w.nquart <= 0) return 0;                                                               \
    if (L->fp32 || a->e.epi == QX_EPI_JK) return -1;                                              \
    hipLaunchKernelGGL(qx_k_##NAME, QX_GRID(a->w.nquart), dim3(QX_BLOCK), 0, L->stream, *a);      \
    return (int)hipGetLastError();                                                                \
  }

// K' rows staged in LDS when a block's quartets share one bra pair: for
// classes with at least QX_JK_ROWS_MIN K' terms per quartet and rows of at
// most QX_JK_LDS_MAX bytes (gfx950 gives a workgroup up to 160 KB of LDS).
// The block's K' updates then cost LDS atomics plus one coalesced pass of
// global atomics over the touched row entries.
#define QX_JK_ROWS_MIN 4
#define QX_JK_LDS_MAX (160 * 1024)

// Streaming J/K kernel: every integral is contracted with the density as
// soon as it is formed; the per-thread J'/K' partials are flushed with
// atomics, J'_ab summed across the wavefront when all its lanes share the
// bra pair, K' through LDS rows when the whole block does.  Cross-lane
// operations and barriers run in converged, block-uniform control flow.
#define QX_DEFINE_JK_KERNEL(NAME, NA, NB, NC, ND, WITH32, A0, A1)                                 \
  template <typename R>                                                                           \
  __global__ __launch_bounds__(QX_BLOCK_MAX) void qx_kj_##NAME(QxArgs A_) {                       \
    HIP_DYNAMIC_SHARED(double, qx_sk)                                                             \
    const int bs_ = (int)blockDim.x;                                                              \
    const long long t = (long long)blockIdx.x * bs_ + threadIdx.x;                                \
    bool active = t < A_.w.nquart;                                                                \
    int u1 = 0, u2 = 0, P1 = -1, P2 = -1;                                                         \
    if (active) active = qx_locate(A_, t, u1, u2, P1, P2);                                        \
    if (active && A_.e.frag_skip && A_.e.frag_skip[A_.e.pair_frag[P1]]) active = false;           \
    if (qx_block_idle(A_, active)) return;                                                        \
    if (!active) P1 = P2 = -1;                                                                    \
    const int n_ = A_.e.nbf;                                                                      \
    const int W_ = A_.e.kcols > 0 ? A_.e.kcols : n_;                                              \
    double* sk = nullptr;                                                                         \
    int ra = 0, rb = 0, c0_ = 0;                                                                  \
    if (A_.e.krows) { /* same result in every thread: block-uniform */                            \
      const long long t0 = (long long)blockIdx.x * bs_;                                           \
      const long long t1 = min(t0 + bs_, A_.w.nquart) - 1;                                        \
      int v1, v2, B0, B1, w2;                                                                     \
      qx_locate(A_, t0, v1, v2, B0, w2);                                                          \
      qx_locate(A_, t1, v1, v2, B1, w2);                                                          \
      if (B0 == B1) {                                                                             \
        sk = qx_sk;                                                                               \
        ra = A_.pr.bfa[B0];                                                                       \
        rb = A_.pr.bfb[B0];                                                                       \
        if (A_.e.frag_col0) c0_ = A_.e.frag_col0[A_.e.pair_frag[B0]];                             \
        for (int i = threadIdx.x; i < (NA + NB) * W_; i += bs_) sk[i] = 0.0;                      \
        __syncthreads();                                                                          \
      }                                                                                           \
    }                                                                                             \
    double jab[NA * NB] = {}, jcd[NC * ND] = {}, kac[NA * NC] = {}, kad[NA * ND] = {},            \
           kbc[NB * NC] = {}, kbd[NB * ND] = {};                                                  \
    if (active) qx_jk_##NAME<R>(A_, P1, P2, u1, u2, jab, jcd, kac, kad, kbc, kbd);                \
    qx_jk_flush<NA, NB, NC, ND, A0, A1>(A_, P1, P2, active, sk, W_, c0_, jab, jcd, kac, kad, kbc, kbd); \
    if (sk) {                                                                                     \
      __syncthreads();                                                                            \
      for (int i = threadIdx.x; i < (NA + NB) * W_; i += bs_) {                                   \
        const double v = sk[i];                                                                   \
        if (v != 0.0) {                                                                           \
          const int r = i / W_, col = c0_ + (i - r * W_);                                         \
          const int row = r < NA ? ra + r : rb + (r - NA);                                        \
          qx_atomic_add(&A_.e.K[(size_t)row * n_ + col], v);                                      \
        }                                                                                         \
      }                                                                                           \
    }                                                                                             \
  }                                                                                               \
  extern "C" int qx_launchj_##NAME(const QxArgs* a, const QxLaunch* L) {                          \
    if (a->w.nquart <= 0) return 0;                                                               \
    if (a->e.epi != QX_EPI_JK) return -1;                                                         \
    QxArgs b = *a;                                                                                \
    const size_t rows_bytes = (size_t)(NA + NB) * (size_t)(a->e.kcols > 0 ? a->e.kcols : a->e.nbf) * sizeof(double); \
    b.e.krows = a->e.krows && ((A1 - A0) * (NC + ND) + NB * NC + NB * ND) >= QX_JK_ROWS_MIN &&    \
                rows_bytes <= QX_JK_LDS_MAX;                                                      \
    const size_t sh = b.e.krows ? rows_bytes : 0;                                                 \
    const int bs = qx_launch_block(L);                                                            \
    if (bs % 64 || bs > QX_BLOCK_MAX) return (int)hipErrorInvalidValue;                           \
    const dim3 grid((unsigned)((a->w.nquart + bs - 1) / bs));                                     \
    if (L->fp32) {                                                                                \
      if (!(WITH32)) return -1;                                                                   \
      if constexpr (WITH32) hipLaunchKernelGGL(qx_kj_##NAME<float>, grid, dim3(bs), sh, L->stream, b); \
    } else {                                                                                      \
      hipLaunchKernelGGL(qx_kj_##NAME<double>, grid, dim3(bs), sh, L->stream, b);                \
    }                                                                                             \
    return (int)hipGetLastError();                                                                \
  }

// Array-form derivative kernel: full derivative-tensor scatter.
#define QX_DEFINE_DERI_KERNEL(NAME, N, NA, NB, NC, ND, CMASK, XMASK, DINV, NCOMB, A0, A1, B0, B1, C0, C1) \
  __global__ __launch_bounds__(QX_BLOCK) void qx_k_##NAME(QxArgs A_) {                            \
    const long long t = (long long)blockIdx.x * QX_BLOCK + threadIdx.x;                          \
    if (t >= A_.w.nquart) return;                                                                 \
    int u1, u2, P1, P2;                                                                           \
    qx_locate(A_, t, u1, u2, P1, P2);                                                             \
    if (qx_single_atom(A_, P1, P2)) return;                                                       \
    using Shape = QxPartShape<NA, NB, NC, ND, A0, A1, B0, B1, C0, C1>;                            \
    static_assert(NCOMB * Shape::NEL * 8 <= 16384,                                                \
                  "per-lane derivative block must stay small (scratch is allocated per lane)");    \
    double dv[NCOMB * Shape::NEL];                                                                \
    qx_core_##NAME<double>(A_, P1, P2, u1, u2, dv);                                               \
    if (A_.e.epi == QX_EPI_JKD)                                                                   \
      qx_epi_djk<NA, NB, NC, ND, CMASK, XMASK, DINV, NCOMB, A0, A1, B0, B1, C0, C1>(A_, P1, P2, dv); \
    else                                                                                          \
      qx_epi_deri<NA, NB, NC, ND, CMASK, XMASK, DINV, NCOMB, A0, A1, B0, B1, C0, C1, double>(A_, P1, P2, dv); \
  }                                                                                               \
  extern "C" int qx_launch_##NAME(const QxArgs* a, const QxLaunch* L) {                           \
    if (a->w.nquart <= 0) return 0;                                                               \
    if (L->fp32 || (a->e.epi != QX_EPI_DFULL && a->e.epi != QX_EPI_JKD)) return -1;               \
    hipLaunchKernelGGL(qx_k_##NAME, QX_GRID(a->w.nquart), dim3(QX_BLOCK), 0, L->stream, *a);      \
    return (int)hipGetLastError();                                                                \
  }

// Tangent array-form derivative kernel: the part's derivative integrals are
// differentiated once more along tv (forward mode, qx_dual recursion), so dv
// holds sum_x tv_x d^2(ab|cd)/dR_x dR_y per centre x direction slot y.
// Epilogues: DFULL (second-derivative tensor slices through slot_map) or
// GRAD (the adjoint contraction with the density pair products, per block in
// LDS over all 3N slots; a context of more than 256 atoms adds to global
// memory instead -- the adjoint kernels' atom windows are left out here, as
// they would lift the largest tangent kernels above 8 KB of scratch per
// lane).  The GRAD path's barriers run in block-uniform control flow.
#define QX_DEFINE_DERT_KERNEL(NAME, N, NA, NB, NC, ND, CMASK, XMASK, DINV, NCOMB, A0, A1, B0, B1, C0, C1) \
  __global__ __launch_bounds__(QX_BLOCK) void qx_kt_##NAME(QxArgs A_) {                           \
    __shared__ double sg[QX_GRAD_LDS];                                                            \
    using Shape = QxPartShape<NA, NB, NC, ND, A0, A1, B0, B1, C0, C1>;                            \
    static_assert(NCOMB * Shape::NEL * 8 <= 16384,                                                \
                  "per-lane derivative block must stay small (scratch is allocated per lane)");    \
    const bool grad = A_.e.epi == QX_EPI_GRAD;                                                    \
    const int ng = 3 * A_.e.natom;                                                                \
    const bool lds = grad && ng <= QX_GRAD_LDS;                                                   \
    if (lds) {                                                                                    \
      for (int i = threadIdx.x; i < ng; i += QX_BLOCK) sg[i] = 0.0;                               \
      __syncthreads();                                                                            \
    }                                                                                             \
    const long long t = (long long)blockIdx.x * QX_BLOCK + threadIdx.x;                          \
    if (t < A_.w.nquart) {                                                                        \
      int u1, u2, P1, P2;                                                                         \
      qx_locate(A_, t, u1, u2, P1, P2);                                                           \
      if (!qx_single_atom(A_, P1, P2)) {                                                          \
        double dv[NCOMB * Shape::NEL];                                                            \
        qx_core_##NAME(A_, P1, P2, u1, u2, dv);                                                   \
        if (grad) {                                                                               \
          double gv[12];                                                                          \
          qx_epi_dgrad<NA, NB, NC, ND, CMASK, XMASK, DINV, NCOMB, A0, A1, B0, B1, C0, C1>(A_, P1, P2, dv, gv); \
          const int at[4] = {A_.pr.atom_a[P1], A_.pr.atom_b[P1], A_.pr.atom_a[P2], A_.pr.atom_b[P2]}; \
          for (int k = 0; k < 12; ++k) {                                                          \
            const int slot = 3 * at[k / 3] + k % 3;                                               \
            if (lds) atomicAdd(&sg[slot], gv[k]);                                                 \
            else qx_atomic_add(&A_.e.grad[slot], gv[k]);                                          \
          }                                                                                       \
        } else {                                                                                  \
          qx_epi_deri<NA, NB, NC, ND, CMASK, XMASK, DINV, NCOMB, A0, A1, B0, B1, C0, C1, double>(A_, P1, P2, dv); \
        }                                                                                         \
      }                                                                                           \
    }                                                                                             \
    if (lds) {                                                                                    \
      __syncthreads();                                                                            \
      qx_grad_flush(A_, sg, 0, ng);                                                               \
    }                                                                                             \
  }                                                                                               \
  extern "C" int qx_launcht_##NAME(const QxArgs* a, const QxLaunch* L) {                          \
    if (a->w.nquart <= 0) return 0;                                                               \
    if (L->fp32 || (a->e.epi != QX_EPI_DFULL && a->e.epi != QX_EPI_GRAD) || !a->e.tv) return -1;  \
    hipLaunchKernelGGL(qx_kt_##NAME, QX_GRID(a->w.nquart), dim3(QX_BLOCK), 0, L->stream, *a);     \
    return (int)hipGetLastError();                                                                \
  }

// Streaming adjoint-gradient kernel body: each derivative integral is
// contracted with the density pair product when formed; per-thread sums over
// the 4 centres x 3 directions go to LDS and are flushed once per block.  A
// generated kernel covers a group of derivative parts of one class (all with
// d/dA, d/dB, d/dC and d/dD by invariance); its grid is (quartet blocks,
// parts in the group) and `part` dispatches on blockIdx.y.  This build emits
// groups of one part.
#define QX_GRAD_NCOMB 9

// Atom window of a block's LDS gradient accumulators: lo = the lowest atom
// that an active lane of the block touches (its centres' atoms at), w = 3 x
// the atoms from lo to the highest one; the block accumulates in LDS when the
// window fits in QX_GRAD_LDS doubles (or the smaller grad_win).  A context
// whose atoms all fit takes the whole gradient as its window, with no
// reduction.  In a fragment context (many molecules, one quartet list) a
// block's quartets lie within one or two molecules, so the window stays a few
// dozen atoms wide however many atoms the context holds.  Called after the
// quartets are evaluated (nothing of it is live across them) by every lane of
// the block; the reduction has two barriers and block-uniform results.
__device__ __forceinline__ bool qx_grad_window(const QxArgs& A_, bool active, const int (&at)[4], int& lo, int& w) {
  const int cap = A_.e.grad_win > 0 && A_.e.grad_win < QX_GRAD_LDS ? A_.e.grad_win : QX_GRAD_LDS;
  if (3 * A_.e.natom <= cap) {  // launch-uniform: every atom fits, no reduction needed
    lo = 0;
    w = 3 * A_.e.natom;
    return true;
  }
  __shared__ int win[2];
  int mn = 0x7fffffff, mx = -1;
  if (active) {
    mn = min(min(at[0], at[1]), min(at[2], at[3]));
    mx = max(max(at[0], at[1]), max(at[2], at[3]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mn = min(mn, __shfl_xor(mn, off));
    mx = max(mx, __shfl_xor(mx, off));
  }
  if (threadIdx.x == 0) {
    win[0] = 0x7fffffff;
    win[1] = -1;
  }
  __syncthreads();
  if (__lane_id() == 0 && mx >= 0) {
    atomicMin(&win[0], mn);
    atomicMax(&win[1], mx);
  }
  __syncthreads();
  const int hi = win[1];
  lo = hi >= 0 ? win[0] : 0;
  w = hi >= 0 ? 3 * (hi - lo + 1) : 0;
  return hi >= 0 && w <= cap;
}

// The gradient copy a block adds to when it does not accumulate in LDS (its
// own of the grad_copies copies, as in qx_grad_flush).
__device__ __forceinline__ double* qx_grad_copy(const QxArgs& A_) {
  const int nc = A_.e.grad_copies > 1 ? A_.e.grad_copies : 1;
  return A_.e.grad + (size_t)(blockIdx.x % nc) * (size_t)(3 * A_.e.natom);
}

// Adds a lane's 12 centre x direction gradient terms gv (zero on inactive
// lanes) to the block's LDS accumulators sg (window from atom lo), or to the
// block's global gradient copy when the window does not fit in LDS.  With
// grad_wred (a per-class choice), a
// centre whose atom is the same on every active lane of the wavefront (the
// bra centres when the wave shares a bra pair) is summed across the wave
// first and added by one lane, so the lanes do not serialise on one LDS
// address; that costs ~200 cross-lane operations per wave, which only the
// cheapest classes repay.  Every lane calls it (converged).
__device__ __forceinline__ void qx_grad_to_block(const QxArgs& A_, double* sg, bool lds, int lo, bool active,
                                                 const int (&at)[4], const double (&gv)[12]) {
  double* const gcp = qx_grad_copy(A_);  // used when the block does not accumulate in LDS
  const int off = lds ? 3 * lo : 0;      // slots are 3 * atom + direction - off
  if (!A_.e.grad_wred) {  // launch-uniform
    if (!active) return;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const int slot = 3 * at[k / 3] + k % 3 - off;
      if (lds) atomicAdd(&sg[slot], gv[k]);
      else qx_atomic_add(&gcp[slot], gv[k]);
    }
    return;
  }
  const bool lead = qx_wave_leader(active);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const bool uni = qx_wave_uniform(at[c], active);  // wave-uniform value
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      const int slot = 3 * at[c] + x - off;
      if (uni) {
        const double v = qx_wave_sum(gv[3 * c + x]);
        if (lead && v != 0.0) {
          if (lds) atomicAdd(&sg[slot], v);
          else qx_atomic_add(&gcp[slot], v);
        }
      } else if (active && gv[3 * c + x] != 0.0) {
        if (lds) atomicAdd(&sg[slot], gv[3 * c + x]);
        else qx_atomic_add(&gcp[slot], gv[3 * c + x]);
      }
    }
  }
}

// Block flush of the LDS accumulators into copy (block index % grad_copies)
// of the global ones: the blocks of a pass spread their atomics over the
// copies instead of all meeting on the same 3N addresses (the runtime sums
// the copies once per pass).
__device__ __forceinline__ void qx_grad_flush(const QxArgs& A_, const double* sg, int lo, int w) {
  double* g = qx_grad_copy(A_) + 3 * lo;
  for (int i = threadIdx.x; i < w; i += blockDim.x)
    if (sg[i] != 0.0) qx_atomic_add(&g[i], sg[i]);
}

template <class Part>
__device__ __forceinline__ void qx_grad_body(const QxArgs& A_, double* sg, Part part) {
  const int bs = (int)blockDim.x;
  const long long t = (long long)blockIdx.x * bs + threadIdx.x;
  int u1 = 0, u2 = 0, P1 = -1, P2 = -1;
  bool active = t < A_.w.nquart;
  if (active) active = qx_locate(A_, t, u1, u2, P1, P2);
  if (qx_block_idle(A_, active)) return;
  if (active && qx_single_atom(A_, P1, P2)) active = false;
  double gv[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) gv[k] = 0.0;
  int at[4] = {0, 0, 0, 0};
  if (active) {
    double g[QX_GRAD_NCOMB];
#pragma unroll
    for (int c = 0; c < QX_GRAD_NCOMB; ++c) g[c] = 0.0;
    part(P1, P2, u1, u2, g);
    qx_grad_spread<7, 7, 1, QX_GRAD_NCOMB>(8.0 * qx_degeneracy(A_, P1, P2), g, gv);
    at[0] = A_.pr.atom_a[P1]; at[1] = A_.pr.atom_b[P1]; at[2] = A_.pr.atom_a[P2]; at[3] = A_.pr.atom_b[P2];
  }
  int lo, w;
  const bool lds = qx_grad_window(A_, active, at, lo, w);
  if (lds) {
    for (int i = threadIdx.x; i < w; i += bs) sg[i] = 0.0;
    __syncthreads();
  }
  qx_grad_to_block(A_, sg, lds, lo, active, at, gv);
  if (lds) {
    __syncthreads();
    qx_grad_flush(A_, sg, lo, w);
  }
}

// Launcher of a generated group kernel qx_kg_NAME with NPARTS parts.
#define QX_DEFINE_GRAD_LAUNCHER(NAME, NPARTS, WITH32)                                             \
  extern "C" int qx_launchg_##NAME(const QxArgs* a, const QxLaunch* L) {                          \
    if (a->w.nquart <= 0) return 0;                                                               \
    if (a->e.epi != QX_EPI_GRAD) return -1;                                                       \
    const int bs = qx_launch_block(L);                                                            \
    if (bs % 64 || bs > QX_BLOCK_MAX) return (int)hipErrorInvalidValue;                           \
    const dim3 grid((unsigned)((a->w.nquart + bs - 1) / bs), NPARTS);                             \
    if (L->fp32) {                                                                                \
      if (!(WITH32)) return -1;                                                                   \
      if constexpr (WITH32) hipLaunchKernelGGL(qx_kg_##NAME<float>, grid, dim3(bs), 0, L->stream, *a); \
    } else {                                                                                      \
      hipLaunchKernelGGL(qx_kg_##NAME<double>, grid, dim3(bs), 0, L->stream, *a);                 \
    }                                                                                             \
    return (int)hipGetLastError();                                                                \
  }

#define QX_DPARTS_CAP 128  // most derivative parts one class may be split into
#define QX_JKPARTS_CAP 8  // most J/K parts (ranges of a components) of one class
#define QX_TPARTS_CAP 256  // most tangent (second-derivative) parts of one class

struct QxClassEntry {
  int la, lb, lc, ld;
  int (*eri)(const QxArgs*, const QxLaunch*);      // array form (store, scatter, Schwarz)
  int njk;
  int (*jk[QX_JKPARTS_CAP])(const QxArgs*, const QxLaunch*);  // streaming J/K, per part
  double jk_flops[QX_JKPARTS_CAP];   // FP operations per primitive quartet, J/K epilogue included
  int nparts;
  int (*deri[QX_DPARTS_CAP])(const QxArgs*, const QxLaunch*);  // array form (derivative tensor), per part
  int ntparts;
  int (*dert[QX_TPARTS_CAP])(const QxArgs*, const QxLaunch*);  // tangent array form (second derivatives), per part
  int ngroups;
  int (*grad[QX_DPARTS_CAP])(const QxArgs*, const QxLaunch*);  // streaming adjoint gradient, per part group
  double eri_flops;                   // algorithmic FP operations of one ERI body (J/K epilogue excluded)
  double part_flops[QX_DPARTS_CAP];   // same per derivative part, with the adjoint epilogue
  int part_ncomb[QX_DPARTS_CAP];      // derivative blocks (centre x direction) per part
  double group_flops[QX_DPARTS_CAP];  // sum of part_flops over each group
  int group_parts[QX_DPARTS_CAP];     // parts per group
};
