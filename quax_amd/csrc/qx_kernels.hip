// Helper kernels of the ERI library: fills, symmetrisation, Cartesian <->
// spherical transforms, and the one-electron integrals (overlap, kinetic,
// nuclear attraction) with their first nuclear derivatives.
//
// One-electron integrals are O(n^2 N) work next to the O(n^4) two-electron
// path, so these kernels favour compact generic code: one thread per ordered
// shell pair, 1D Obara-Saika tables for S and T, an Obara-Saika vertical
// recurrence on the bra for V followed by the binomial transfer
// (a|O|b) = sum_k C(b,k) AB^(b-k) (a+k|O|0) per Cartesian direction.
#include <hip/hip_runtime.h>

#include "qx_runtime.h"

namespace {

__global__ void k_fill(double* p, long long n, double v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// out[i] += sum_c copies[c * n + i]
__global__ void k_add_copies(const double* copies, int ncopy, int n, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  for (int c = 0; c < ncopy; ++c) v += copies[(size_t)c * n + i];
  out[i] += v;
}

__global__ void k_sym_half(const double* in, double* out, int n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * n) return;
  const int i = (int)(t / n), j = (int)(t % n);
  out[t] = 0.5 * (in[(size_t)i * n + j] + in[(size_t)j * n + i]);
}

// AO function k of shell s: row of the transform block; Cartesian function p
// of the same shell: column.
__device__ __forceinline__ double tcoef(const QxShellsDev& sh, int s, int row, int col) {
  const int l = sh.l[s];
  const int nc = (l + 1) * (l + 2) / 2;
  return sh.sph[sh.sph_off[l] + row * nc + col];
}

// out[i][j] = sum_{p in shell(i), q in shell(j)} T[i][p] X[p][q] T[j][q]
__global__ void k_cart_to_ao_mat(QxShellsDev sh, const double* X, double* out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = sh.nbf, nc = sh.ncart;
  if (t >= (long long)nb * nb) return;
  const int i = (int)(t / nb), j = (int)(t % nb);
  const int si = sh.ao_shell[i], sj = sh.ao_shell[j];
  const int ri = i - sh.bfs[si], rj = j - sh.bfs[sj];
  const int li = sh.l[si], lj = sh.l[sj];
  const int ni = (li + 1) * (li + 2) / 2, nj = (lj + 1) * (lj + 2) / 2;
  double acc = 0.0;
  for (int p = 0; p < ni; ++p) {
    const double tp = tcoef(sh, si, ri, p);
    if (tp == 0.0) continue;
    for (int q = 0; q < nj; ++q) acc += tp * X[(size_t)(sh.bfc[si] + p) * nc + sh.bfc[sj] + q] * tcoef(sh, sj, rj, q);
  }
  out[t] = acc;
}

// out[p][q] = sum_{i in shell(p), j in shell(q)} T[i][p] X[i][j] T[j][q]
__global__ void k_ao_to_cart_mat(QxShellsDev sh, const double* X, double* out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = sh.nbf, nc = sh.ncart;
  if (t >= (long long)nc * nc) return;
  const int p = (int)(t / nc), q = (int)(t % nc);
  const int sp = sh.f_shell[p], sq = sh.f_shell[q];
  const int cp = p - sh.bfc[sp], cq = q - sh.bfc[sq];
  const int np = (sh.pure && sh.l[sp] >= 2) ? 2 * sh.l[sp] + 1 : (sh.l[sp] + 1) * (sh.l[sp] + 2) / 2;
  const int nq = (sh.pure && sh.l[sq] >= 2) ? 2 * sh.l[sq] + 1 : (sh.l[sq] + 1) * (sh.l[sq] + 2) / 2;
  double acc = 0.0;
  for (int i = 0; i < np; ++i) {
    const double ti = tcoef(sh, sp, i, cp);
    if (ti == 0.0) continue;
    for (int j = 0; j < nq; ++j) acc += ti * X[(size_t)(sh.bfs[sp] + i) * nb + sh.bfs[sq] + j] * tcoef(sh, sq, j, cq);
  }
  out[t] = acc;
}

// out[o][i][r] = sum_{p in shell(i)} T[i][p] in[o][p][r]
__global__ void k_cart_to_ao_axis(QxShellsDev sh, const double* in, double* out, long long outer, int lin, int lout,
                                  long long inner) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = outer * lout * inner;
  if (t >= total) return;
  const long long r = t % inner;
  const long long rest = t / inner;
  const int i = (int)(rest % lout);
  const long long o = rest / lout;
  const int s = sh.ao_shell[i];
  const int ri = i - sh.bfs[s];
  const int l = sh.l[s];
  const int nc = (l + 1) * (l + 2) / 2;
  double acc = 0.0;
  for (int p = 0; p < nc; ++p) {
    const double tp = tcoef(sh, s, ri, p);
    if (tp != 0.0) acc += tp * in[(o * lin + sh.bfc[s] + p) * inner + r];
  }
  out[t] = acc;
}

// ---------------------------------------------------------------- OEI
#define OE_LMAX 3          // bra/ket l after a derivative shift
#define OE_JMAX (OE_LMAX + 2)
#define OE_LT (2 * OE_LMAX)

__device__ __forceinline__ int d_coff(int l) { return l * (l + 1) * (l + 2) / 6; }
__device__ __forceinline__ int d_gidx(int x, int y, int z) {
  const int i = y + z;
  return d_coff(x + y + z) + i * (i + 1) / 2 + z;
}

__device__ void oe_boys(int M, double T, double* F) {
  // direct series / asymptotic evaluation (one-electron integrals only)
  if (T < 35.0) {
    double term = 1.0 / (2 * M + 1), sum = term;
    for (int k = 1; k < 300; ++k) {
      term *= 2.0 * T / (2 * M + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    const double et = exp(-T);
    F[M] = et * sum;
    for (int m = M; m > 0; --m) F[m - 1] = (2.0 * T * F[m] + et) / (2 * m - 1);
  } else {
    const double et = exp(-T);
    F[0] = 0.5 * sqrt(3.14159265358979323846 / T);
    for (int m = 0; m < M; ++m) F[m + 1] = ((2 * m + 1) * F[m] - et) / (2.0 * T);
  }
}

// The one-electron helpers are templated on the scalar S of the geometric
// quantities: double, or qx_dual for the tangent (forward-mode) kernel that
// differentiates the first-derivative contraction once more along a
// geometry direction.  Exponents stay double.
__device__ __forceinline__ double oe_exp(double x) { return exp(x); }
__device__ __forceinline__ qx_dual oe_exp(qx_dual x) {
  const double e = exp(x.v);
  return qx_dual(e, e * x.d);
}
__device__ __forceinline__ double oe_ipow(double x, int n) { return pow(x, n); }
__device__ __forceinline__ qx_dual oe_ipow(qx_dual x, int n) {
  qx_dual r(1.0);
  for (int i = 0; i < n; ++i) r = r * x;
  return r;
}
// what a contraction accumulates: the value, or the derivative part
__device__ __forceinline__ double oe_out(double x) { return x; }
__device__ __forceinline__ double oe_out(qx_dual x) { return x.d; }

template <typename S>
struct OePrimT {
  double a, b, p;      // exponents, p = a+b
  S K;                 // ca cb exp(-ab/p AB^2)
  S PA[3], PB[3], AB[3];
};
using OePrim = OePrimT<double>;

// 1D overlap tables s[d][i][j] (no prefactor), i <= imax, j <= jmax
template <typename S>
__device__ void oe_s1d(const OePrimT<S>& P, int imax, int jmax, S (*s)[OE_LMAX + 2][OE_JMAX + 2]) {
  const double o2p = 0.5 / P.p;
  for (int d = 0; d < 3; ++d)
    for (int j = 0; j <= jmax; ++j)
      for (int i = 0; i <= imax; ++i) {
        S v;
        if (i == 0 && j == 0) v = S(1.0);
        else if (j == 0) {
          v = P.PA[d] * s[d][i - 1][0];
          if (i > 1) v += o2p * (i - 1) * s[d][i - 2][0];
        } else {
          v = P.PB[d] * s[d][i][j - 1];
          if (i > 0) v += o2p * i * s[d][i - 1][j - 1];
          if (j > 1) v += o2p * (j - 1) * s[d][i][j - 2];
        }
        s[d][i][j] = v;
      }
}

// S and T for component pair (a, b) from the 1D tables
template <typename S>
__device__ __forceinline__ void oe_st(const OePrimT<S>& P, const int* a, const int* b,
                                      S (*s)[OE_LMAX + 2][OE_JMAX + 2], S& Sv, S& T) {
  const double pref = pow(3.14159265358979323846 / P.p, 1.5);
  S sv[3], tv[3];
  for (int d = 0; d < 3; ++d) {
    const int i = a[d], j = b[d];
    sv[d] = s[d][i][j];
    S t = P.b * (2 * j + 1) * s[d][i][j] - 2.0 * P.b * P.b * s[d][i][j + 2];
    if (j > 1) t = t - 0.5 * j * (j - 1) * s[d][i][j - 2];
    tv[d] = t;
  }
  Sv = pref * sv[0] * sv[1] * sv[2];
  T = pref * (tv[0] * sv[1] * sv[2] + sv[0] * tv[1] * sv[2] + sv[0] * sv[1] * tv[2]);
}

// Boys function of a dual argument: dF_m/dT = -F_{m+1}
__device__ void oe_boys(int M, qx_dual T, qx_dual* F) {
  double f[OE_LT + 3];
  oe_boys(M + 1, T.v, f);
  for (int m = 0; m <= M; ++m) F[m] = qx_dual(f[m], -T.d * f[m + 1]);
}

// [e|0]^(0) for |e| <= L toward a unit charge at C (no -Z, includes 2pi/p K)
template <typename S>
__device__ void oe_vrr(const OePrimT<S>& P, const S* Pc, const S* C, int L, S* out /*coff(L+1)*/) {
  S V[OE_LT + 1][56];
  S PC[3] = {Pc[0] - C[0], Pc[1] - C[1], Pc[2] - C[2]};
  const S T = P.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
  S F[OE_LT + 2];
  oe_boys(L, T, F);
  const S pref = 2.0 * 3.14159265358979323846 / P.p * P.K;
  for (int m = 0; m <= L; ++m) V[m][0] = pref * F[m];
  const double o2p = 0.5 / P.p;
  for (int l = 1; l <= L; ++l)
    for (int i = 0; i <= l; ++i)
      for (int j = 0; j <= i; ++j) {
        const int e[3] = {l - i, i - j, j};
        const int ax = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
        int e1[3] = {e[0], e[1], e[2]};
        e1[ax] -= 1;
        const int g = d_gidx(e[0], e[1], e[2]), g1 = d_gidx(e1[0], e1[1], e1[2]);
        int g2 = -1;
        if (e1[ax] > 0) {
          int e2[3] = {e1[0], e1[1], e1[2]};
          e2[ax] -= 1;
          g2 = d_gidx(e2[0], e2[1], e2[2]);
        }
        for (int m = 0; m <= L - l; ++m) {
          S v = P.PA[ax] * V[m][g1] - PC[ax] * V[m + 1][g1];
          if (g2 >= 0) v = v + o2p * e1[ax] * (V[m][g2] - V[m + 1][g2]);
          V[m][g] = v;
        }
      }
  for (int g = 0; g < d_coff(L + 1); ++g) out[g] = V[0][g];
}

__device__ __forceinline__ double binom(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// (a|V_C|b) from the [e|0] table by the binomial transfer
template <typename S>
__device__ S oe_vab(const S* E, const int* a, const int* b, const S* AB) {
  S acc(0.0);
  for (int kx = 0; kx <= b[0]; ++kx)
    for (int ky = 0; ky <= b[1]; ++ky)
      for (int kz = 0; kz <= b[2]; ++kz) {
        const S c = binom(b[0], kx) * binom(b[1], ky) * binom(b[2], kz) * oe_ipow(AB[0], b[0] - kx) *
                    oe_ipow(AB[1], b[1] - ky) * oe_ipow(AB[2], b[2] - kz);
        acc = acc + c * E[d_gidx(a[0] + kx, a[1] + ky, a[2] + kz)];
      }
  return acc;
}

__device__ void unpack(int code, int* c) {
  c[0] = code >> 16;
  c[1] = (code >> 8) & 0xff;
  c[2] = code & 0xff;
}

// Centre coordinate of an atom: its position, moving along tv for duals.
template <typename S>
__device__ __forceinline__ void oe_centre(const QxShellsDev& sh, const double* tv, int atom, S* X);
template <>
__device__ __forceinline__ void oe_centre<double>(const QxShellsDev& sh, const double*, int atom, double* X) {
  for (int d = 0; d < 3; ++d) X[d] = sh.xyz[3 * atom + d];
}
template <>
__device__ __forceinline__ void oe_centre<qx_dual>(const QxShellsDev& sh, const double* tv, int atom, qx_dual* X) {
  for (int d = 0; d < 3; ++d) X[d] = qx_dual(sh.xyz[3 * atom + d], tv[3 * atom + d]);
}

template <typename S>
__device__ OePrimT<S> oe_prim(const QxShellsDev& sh, int sa, int sb, int ia, int ib, const double* tv = nullptr) {
  OePrimT<S> P;
  S A[3], B[3];
  oe_centre<S>(sh, tv, sh.atom[sa], A);
  oe_centre<S>(sh, tv, sh.atom[sb], B);
  P.a = sh.exps[ia];
  P.b = sh.exps[ib];
  P.p = P.a + P.b;
  S AB2(0.0);
  for (int d = 0; d < 3; ++d) {
    const S Pd = (P.a * A[d] + P.b * B[d]) / P.p;
    P.PA[d] = Pd - A[d];
    P.PB[d] = Pd - B[d];
    P.AB[d] = A[d] - B[d];
    AB2 = AB2 + P.AB[d] * P.AB[d];
  }
  P.K = sh.coefs[ia] * sh.coefs[ib] * oe_exp(-P.a * P.b / P.p * AB2);
  return P;
}

// thread per ordered shell pair: S, T, V blocks (Cartesian)
__global__ void k_oei(QxShellsDev sh, double* S, double* T, double* V) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= sh.nshell * sh.nshell) return;
  const int sa = t / sh.nshell, sb = t % sh.nshell;
  const int la = sh.l[sa], lb = sh.l[sb];
  const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
  const int n = sh.ncart;
  double bs[36], bt[36], bv[36];
  for (int k = 0; k < na * nb; ++k) bs[k] = bt[k] = bv[k] = 0.0;
  double s1[3][OE_LMAX + 2][OE_JMAX + 2];
  double E[84];
  for (int ia = sh.poff[sa]; ia < sh.poff[sa] + sh.nprim[sa]; ++ia)
    for (int ib = sh.poff[sb]; ib < sh.poff[sb] + sh.nprim[sb]; ++ib) {
      const OePrim P = oe_prim<double>(sh, sa, sb, ia, ib);
      oe_s1d(P, la, lb + 2, s1);
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
          int a[3], b[3];
          unpack(sh.f_comp[sh.bfc[sa] + i], a);
          unpack(sh.f_comp[sh.bfc[sb] + j], b);
          double s, tk;
          oe_st(P, a, b, s1, s, tk);
          bs[i * nb + j] += P.K * s;
          bt[i * nb + j] += P.K * tk;
        }
      double Pc[3];
      const double* A = sh.xyz + 3 * sh.atom[sa];
      for (int d = 0; d < 3; ++d) Pc[d] = A[d] + P.PA[d];
      for (int at = 0; at < sh.natom; ++at) {
        oe_vrr(P, Pc, sh.xyz + 3 * at, la + lb, E);
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j) {
            int a[3], b[3];
            unpack(sh.f_comp[sh.bfc[sa] + i], a);
            unpack(sh.f_comp[sh.bfc[sb] + j], b);
            bv[i * nb + j] -= sh.Z[at] * oe_vab(E, a, b, P.AB);
          }
      }
    }
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j) {
      const size_t o = (size_t)(sh.bfc[sa] + i) * n + sh.bfc[sb] + j;
      S[o] = bs[i * nb + j];
      T[o] = bt[i * nb + j];
      V[o] = bv[i * nb + j];
    }
}

// Derivatives: d/dA_x of (a|O|b) = 2 alpha (a+1_x|O|b) - a_x (a-1_x|O|b), same on
// b; the nuclear-attraction operator centre follows from translational
// invariance of each nucleus term.  With Pd/Wd/grad the contraction
// grad += Pd.d(T+V) + Wd.dS is accumulated; with dS/dT/dV the full 3N x n x n
// tensors are accumulated.  Thread x takes one ordered shell pair; grid row y
// takes the terms: 0 the overlap and kinetic ones, 1 + C the attraction of
// nucleus C (rows stride by gridDim.y), so a pair's nucleus loop is spread
// over the grid instead of run by one lane.
// S = qx_dual (tangent mode, tv = geometry direction, tensors not written):
// every atom moves along tv and the contraction accumulates the derivative
// part, grad += Pd.d/dtv d(T+V) + Wd.d/dtv dS.
template <typename S>
__global__ void k_oei_deriv(QxShellsDev sh, const double* Pd, const double* Wd, double* grad, double* dS, double* dT,
                            double* dV, const double* tv) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= sh.nshell * sh.nshell) return;
  const int sa = t / sh.nshell, sb = t % sh.nshell;
  const int la = sh.l[sa], lb = sh.l[sb];
  const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
  const int n = sh.ncart;
  const size_t nn = (size_t)n * n;
  const int atA = sh.atom[sa], atB = sh.atom[sb];
  S s1[3][OE_LMAX + 2][OE_JMAX + 2];
  S E[84];
  for (int term = blockIdx.y; term <= sh.natom; term += gridDim.y) {
    double gAB[2][3] = {{0, 0, 0}, {0, 0, 0}};
    for (int ia = sh.poff[sa]; ia < sh.poff[sa] + sh.nprim[sa]; ++ia)
      for (int ib = sh.poff[sb]; ib < sh.poff[sb] + sh.nprim[sb]; ++ib) {
        const OePrimT<S> P = oe_prim<S>(sh, sa, sb, ia, ib, tv);
        if (term == 0) {  // overlap and kinetic
          oe_s1d(P, la + 1, lb + 3, s1);
          for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j) {
              int a[3], b[3];
              unpack(sh.f_comp[sh.bfc[sa] + i], a);
              unpack(sh.f_comp[sh.bfc[sb] + j], b);
              const size_t rc = (size_t)(sh.bfc[sa] + i) * n + sh.bfc[sb] + j;
              const double pd = Pd ? Pd[rc] : 0.0, wd = Wd ? Wd[rc] : 0.0;
              for (int side = 0; side < 2; ++side) {
                const double ex2 = 2.0 * (side == 0 ? P.a : P.b);
                for (int x = 0; x < 3; ++x) {
                  int au[3] = {a[0], a[1], a[2]}, bu[3] = {b[0], b[1], b[2]};
                  if (side == 0) au[x] += 1; else bu[x] += 1;
                  S su, tu;
                  oe_st(P, au, bu, s1, su, tu);
                  S ds = ex2 * su, dt = ex2 * tu;
                  const int q = side == 0 ? a[x] : b[x];
                  if (q > 0) {
                    int ad[3] = {a[0], a[1], a[2]}, bd[3] = {b[0], b[1], b[2]};
                    if (side == 0) ad[x] -= 1; else bd[x] -= 1;
                    S sd, td;
                    oe_st(P, ad, bd, s1, sd, td);
                    ds -= q * sd;
                    dt -= q * td;
                  }
                  ds *= P.K;
                  dt *= P.K;
                  const int at = side == 0 ? atA : atB;
                  if (grad) gAB[side][x] += pd * oe_out(dt) + wd * oe_out(ds);
                  if (dS) {
                    atomicAdd(&dS[(size_t)(3 * at + x) * nn + rc], oe_out(ds));
                    atomicAdd(&dT[(size_t)(3 * at + x) * nn + rc], oe_out(dt));
                  }
                }
              }
            }
          continue;
        }
        // attraction of nucleus atC
        const int atC = term - 1;
        S Pc[3], Ac[3], Cc[3];
        oe_centre<S>(sh, tv, sh.atom[sa], Ac);
        oe_centre<S>(sh, tv, atC, Cc);
        for (int d = 0; d < 3; ++d) Pc[d] = Ac[d] + P.PA[d];
        oe_vrr(P, Pc, Cc, la + lb + 1, E);
        double gC[3] = {0, 0, 0};
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j) {
            int a[3], b[3];
            unpack(sh.f_comp[sh.bfc[sa] + i], a);
            unpack(sh.f_comp[sh.bfc[sb] + j], b);
            const size_t rc = (size_t)(sh.bfc[sa] + i) * n + sh.bfc[sb] + j;
            const double pd = Pd ? Pd[rc] : 0.0;
            for (int side = 0; side < 2; ++side) {
              const double ex2 = 2.0 * (side == 0 ? P.a : P.b);
              for (int x = 0; x < 3; ++x) {
                int au[3] = {a[0], a[1], a[2]}, bu[3] = {b[0], b[1], b[2]};
                if (side == 0) au[x] += 1; else bu[x] += 1;
                S dv = ex2 * oe_vab(E, au, bu, P.AB);
                const int q = side == 0 ? a[x] : b[x];
                if (q > 0) {
                  int ad[3] = {a[0], a[1], a[2]}, bd[3] = {b[0], b[1], b[2]};
                  if (side == 0) ad[x] -= 1; else bd[x] -= 1;
                  dv -= q * oe_vab(E, ad, bd, P.AB);
                }
                dv *= S(-sh.Z[atC]);
                const int at = side == 0 ? atA : atB;
                const double dvo = oe_out(dv);
                if (grad) {
                  gAB[side][x] += pd * dvo;
                  gC[x] -= pd * dvo;
                }
                if (dV) {
                  atomicAdd(&dV[(size_t)(3 * at + x) * nn + rc], dvo);
                  atomicAdd(&dV[(size_t)(3 * atC + x) * nn + rc], -dvo);
                }
              }
            }
          }
        if (grad)
          for (int x = 0; x < 3; ++x)
            if (gC[x] != 0.0) atomicAdd(&grad[3 * atC + x], gC[x]);
      }
    if (grad)
      for (int x = 0; x < 3; ++x) {
        atomicAdd(&grad[3 * atA + x], gAB[0][x]);
        atomicAdd(&grad[3 * atB + x], gAB[1][x]);
      }
  }
}

inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

extern "C" {

int qxk_fill(double* p, long long n, double v, hipStream_t s) {
  if (n <= 0) return 0;
  const long long blocks = std::min<long long>((n + 255) / 256, 1 << 20);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return (int)hipGetLastError();
}

int qxk_add_copies(const double* copies, int ncopy, int n, double* out, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_add_copies, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, copies, ncopy, n, out);
  return (int)hipGetLastError();
}

int qxk_sym_half(const double* in, double* out, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_sym_half, dim3(nblk((long long)n * n, 256)), dim3(256), 0, s, in, out, n);
  return (int)hipGetLastError();
}

int qxk_cart_to_ao_mat(const QxShellsDev* sh, const double* in, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_cart_to_ao_mat, dim3(nblk((long long)sh->nbf * sh->nbf, 256)), dim3(256), 0, s, *sh, in, out);
  return (int)hipGetLastError();
}

int qxk_ao_to_cart_mat(const QxShellsDev* sh, const double* in, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_ao_to_cart_mat, dim3(nblk((long long)sh->ncart * sh->ncart, 256)), dim3(256), 0, s, *sh, in,
                     out);
  return (int)hipGetLastError();
}

int qxk_cart_to_ao_axis(const QxShellsDev* sh, const double* in, double* out, long long outer, int lin, int lout,
                        long long inner, hipStream_t s) {
  const long long total = outer * lout * inner;
  hipLaunchKernelGGL(k_cart_to_ao_axis, dim3(nblk(total, 256)), dim3(256), 0, s, *sh, in, out, outer, lin, lout,
                     inner);
  return (int)hipGetLastError();
}

int qxk_oei(const QxShellsDev* sh, double* S, double* T, double* V, hipStream_t s) {
  hipLaunchKernelGGL(k_oei, dim3(nblk((long long)sh->nshell * sh->nshell, 64)), dim3(64), 0, s, *sh, S, T, V);
  return (int)hipGetLastError();
}

int qxk_oei_deriv(const QxShellsDev* sh, const double* Pd, const double* Wd, double* grad, double* dS, double* dT,
                  double* dV, hipStream_t s) {
  const unsigned rows = (unsigned)std::min(sh->natom + 1, 65535);
  hipLaunchKernelGGL(k_oei_deriv<double>, dim3(nblk((long long)sh->nshell * sh->nshell, 64), rows), dim3(64), 0, s, *sh,
                     Pd, Wd, grad, dS, dT, dV, (const double*)nullptr);
  return (int)hipGetLastError();
}

int qxk_oei_hess_dir(const QxShellsDev* sh, const double* Pd, const double* Wd, const double* tv, double* grad,
                     hipStream_t s) {
  const unsigned rows = (unsigned)std::min(sh->natom + 1, 65535);
  hipLaunchKernelGGL(k_oei_deriv<qx_dual>, dim3(nblk((long long)sh->nshell * sh->nshell, 64), rows), dim3(64), 0, s,
                     *sh, Pd, Wd, grad, (double*)nullptr, (double*)nullptr, (double*)nullptr, tv);
  return (int)hipGetLastError();
}

}  // extern "C"
