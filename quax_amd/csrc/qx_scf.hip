// RHF SCF iterations as one HIP graph per iteration, for one molecule or a
// batch of molecules (fragments of one context).
//
// The reference runs its SCF as one jax.lax.while_loop (quax/methods/
// hartree_fock.py:83-99): J/K from the stored ERI tensor, Fock matrix,
// energy, DIIS error A (F D S - S D F) A, eigh of the orthogonalised Fock
// matrix and the aufbau density.  Here one loop body is a fixed sequence of
// kernels with no host synchronisation, captured once into a hipGraph and
// replayed; the host reads the convergence flags every few replays.
//   * J/K: the integral-direct pass (qx_jk_build) over the context's
//     block-diagonal density;
//   * the dense steps are this file's kernels: a batched FP64 MFMA GEMM
//     (grid z = molecule), per-molecule reductions (no atomics, so nothing
//     needs clearing between replays), DIIS on device (ring of Fock and error
//     matrices, Pulay system solved by one thread per molecule);
//   * the density of the orthogonalised Fock matrix F' comes from
//     trace-correcting purification (Niklasson, Phys. Rev. B 66, 155115
//     (2002), "TC2"): X0 = (e_max - F') / (e_max - e_min) with Gershgorin
//     bounds, then X <- X^2 or 2X - X^2, whichever trace is closer to the
//     occupied count, until tr(X - X^2) < 1e-8 (at most a fixed number of
//     times), then two McWeeny steps X <- 3X^2 - 2X^3.  X converges to the projector
//     onto the nocc lowest eigenvectors of F' -- the aufbau density the
//     reference builds from eigh -- with GEMMs only.
// Molecules are padded to nm functions (zero rows and columns); a molecule
// whose SCF has converged (or hit maxit) is skipped by every dense kernel of
// later replays.  No hipMemsetAsync and no library call is recorded.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "qx_runtime.h"
#include "quax_amd.h"

namespace {

// per-molecule scalars (doubles, device)
enum {
  QS_E = 0,      // energy of the current iteration's density
  QS_ENEW,       // energy computed this iteration
  QS_DE,         // E_prev - E_new
  QS_DRMS,       // RMS of the orthogonalised DIIS error
  QS_EMIN,       // Gershgorin bounds of F'
  QS_EMAX,
  QS_IDEM,       // tr(X - X^2) before the last McWeeny step
  QS_DONE,       // 1 once converged or at maxit
  QS_IT,         // completed iterations
  QS_ENUC,
  QS_NOCC,
  QS_PCONV0,     // purification converged, read by step k from slot k % 2
  QS_PCONV1,     // (written by step k - 1: no block reads what its launch writes)
  QS_N = 16
};

// |tr(X - X^2)| below which the TC2 steps stop: every eigenvalue of X is
// then within ~1e-8 of 0 or 1 (the occupied subspace is separated), and the
// McWeeny steps that follow square the remaining deviations on both sides.
// (A TC2 step squares them on one side only and doubles them on the other,
// so stopping TC2 itself near rounding level leaves ~1e-13 errors, which the
// orthogonaliser amplifies in D = A X A^T.)
constexpr double PURIFY_TOL = 1e-8;
constexpr int MCWEENY_STEPS = 2;

constexpr int SB = 256;  // threads of the reduction and elementwise kernels
// threads of the per-molecule reductions (energy, DIIS dots, Gershgorin
// bounds): one block per molecule, so a block must keep many loads in flight
constexpr int RB = 1024;

// Sum over the RB threads of a block (sh: RB doubles of LDS), in a fixed
// order; every thread gets the result.
__device__ __forceinline__ double block_sum_rb(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = SB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ bool mol_done(const double* sc, int b) { return sc[b * QS_N + QS_DONE] != 0.0; }

// ---- batched FP64 MFMA GEMM ---------------------------------------------
// C[b] = alpha op(A[b]) op(B[b]) + beta C[b], square nm x nm row-major
// blocks with stride nm*nm; op = transpose when TA / TB.  64 x 64 output
// tile per 256-thread workgroup (four wavefronts of 32 x 32, 2 x 2 tiles of
// v_mfma_f64_16x16x4f64), K staged through LDS in steps of 32.  Operand
// layout of the 16x16x4 FP64 MFMA: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator element r of lane l is D[4 r + l / 16][l % 16].
// Molecules flagged done (sc) are skipped when skip is set.
//
// Split K: a single molecule's GEMM has few tiles (25 at n = 280 on a
// 256-CU device), and each tile's four wavefronts walk all of K: the FP64
// MFMAs alone then take ~8 us per tile and the loads of operands that the
// previous kernel just wrote come from memory.  With nsplit > 1, blockIdx.z =
// molecule * nsplit + split: each split walks its own range of K steps and
// stores its partial tile, and k_bgemm_reduce (the next launch, so no fences
// between workgroups) adds the partials in split order -- the same sum
// however the splits were scheduled -- and applies alpha and beta.
typedef double qx_sv4 __attribute__((ext_vector_type(4)));
constexpr int BM = 64, BN = 64, BK = 32, GT = 256, PAD = 2;
constexpr int KSPLIT_MAX = 8;
constexpr int PART_TILE = BM * BN;  // doubles of one partial tile (16 per thread)

// Whether the GEMM of molecule b skips its work (molecule done, or its
// purification already converged for purification step parity pslot).
__device__ __forceinline__ bool gemm_skips(const double* sc, int b, int skip, int pslot) {
  return (skip && mol_done(sc, b)) || (pslot >= 0 && sc[b * QS_N + QS_PCONV0 + pslot] != 0.0);
}

template <bool TA, bool TB>
__global__ __launch_bounds__(GT) void k_bgemm(int nm, const int* nb, const double* __restrict__ A,
                                              const double* __restrict__ B, double* __restrict__ C, double alpha,
                                              double beta, const double* sc, int skip, int pslot, int nsplit,
                                              double* __restrict__ part) {
  const int b = (int)blockIdx.z / nsplit, split = (int)blockIdx.z % nsplit;
  if (gemm_skips(sc, b, skip, pslot)) return;
  const int n = nb[b];
  const long long i0 = (long long)blockIdx.y * BM, j0 = (long long)blockIdx.x * BN;
  if (i0 >= n || j0 >= n) return;  // block-uniform
  const size_t off = (size_t)b * nm * nm;
  A += off; B += off; C += off;
  __shared__ double As[BK][BM + PAD];
  __shared__ double Bs[BK][BN + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  qx_sv4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = qx_sv4{0.0, 0.0, 0.0, 0.0};
  const int nk = (n + BK - 1) / BK, per = (nk + nsplit - 1) / nsplit;  // K steps: all, per split
  const int kbeg = split * per * BK, kend = min(n, (split + 1) * per * BK);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
    for (int r = 0; r < BK * BM / GT; ++r) {
      const int idx = tid + r * GT;
      int i, k;
      if (TA) { k = idx / BM; i = idx % BM; } else { i = idx / BK; k = idx % BK; }
      const long long gi = i0 + i;
      const int gk = k0 + k;
      double v = 0.0;
      if (gi < n && gk < n) v = TA ? A[(size_t)gk * nm + gi] : A[(size_t)gi * nm + gk];
      As[k][i] = v;
    }
#pragma unroll
    for (int r = 0; r < BK * BN / GT; ++r) {
      const int idx = tid + r * GT;
      int j, k;
      if (TB) { j = idx / BK; k = idx % BK; } else { k = idx / BN; j = idx % BN; }
      const long long gj = j0 + j;
      const int gk = k0 + k;
      double v = 0.0;
      if (gj < n && gk < n) v = TB ? B[(size_t)gj * nm + gk] : B[(size_t)gk * nm + gj];
      Bs[k][j] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int ka = kk + (lane >> 4);
      double a[2], bb[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        a[t] = As[ka][wr * 32 + t * 16 + (lane & 15)];
        bb[t] = Bs[ka][wc * 32 + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
          acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta], bb[tb], acc[ta][tb], 0, 0, 0);
    }
    __syncthreads();
  }
  if (nsplit > 1) {  // partial tile out, element q of every thread contiguous across the block
    const size_t tile = ((size_t)b * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    double* mine = part + (tile * nsplit + split) * PART_TILE;
#pragma unroll
    for (int q = 0; q < 16; ++q) mine[q * GT + tid] = acc[q >> 3][(q >> 2) & 1][q & 3];
    return;
  }
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long gi = i0 + wr * 32 + ta * 16 + 4 * r + (lane >> 4);
        const long long gj = j0 + wc * 32 + tb * 16 + (lane & 15);
        if (gi < n && gj < n) {
          double* p = &C[(size_t)gi * nm + gj];
          *p = beta == 0.0 ? alpha * acc[ta][tb][r] : alpha * acc[ta][tb][r] + beta * *p;
        }
      }
}

// C[b] = alpha (sum of the nsplit partial tiles, in split order) + beta C[b];
// grid (tiles, tiles, molecules), the thread -> element map of k_bgemm.
__global__ __launch_bounds__(GT) void k_bgemm_reduce(int nm, const int* nb, double* __restrict__ C, double alpha,
                                                     double beta, const double* sc, int skip, int pslot, int nsplit,
                                                     const double* __restrict__ part) {
  const int b = (int)blockIdx.z;
  if (gemm_skips(sc, b, skip, pslot)) return;
  const int n = nb[b];
  const long long i0 = (long long)blockIdx.y * BM, j0 = (long long)blockIdx.x * BN;
  if (i0 >= n || j0 >= n) return;
  C += (size_t)b * nm * nm;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const size_t tile = ((size_t)b * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const double* all = part + tile * nsplit * PART_TILE;
  // all loads of a split issued together (a thread's 16 x nsplit loads would
  // otherwise run one latency after another)
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = 0.0;
#pragma unroll
  for (int t = 0; t < KSPLIT_MAX; ++t)
    if (t < nsplit) {
      double x[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) x[q] = all[(size_t)t * PART_TILE + q * GT + tid];
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] += x[q];
    }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int ta = q >> 3, tb = (q >> 2) & 1, r = q & 3;
    const long long gi = i0 + wr * 32 + ta * 16 + 4 * r + (lane >> 4);
    const long long gj = j0 + wc * 32 + tb * 16 + (lane & 15);
    if (gi < n && gj < n) {
      double* p = &C[(size_t)gi * nm + gj];
      *p = beta == 0.0 ? alpha * v[q] : alpha * v[q] + beta * *p;
    }
  }
}

// ---- elementwise / reduction kernels --------------------------------------
// grid (ceil(nm*nm / SB), nmol)
#define QS_ELEM_PROLOGUE                                       \
  const int b = blockIdx.y;                                    \
  const long long e = (long long)blockIdx.x * SB + threadIdx.x; \
  if (e >= (long long)nm * nm) return;                         \
  const int i = (int)(e / nm), j = (int)(e % nm);              \
  const size_t o = (size_t)b * nm * nm + e;

// batch <-> block-diagonal full matrix (ntot x ntot) of the context
__global__ void k_gather(const double* full, int ntot, double* bat, int nm, const int* nb, const int* off,
                         const double* sc, int skip) {
  QS_ELEM_PROLOGUE
  if (skip && mol_done(sc, b)) return;
  const int n = nb[b];
  bat[o] = (i < n && j < n) ? full[(size_t)(off[b] + i) * ntot + off[b] + j] : 0.0;
}

__global__ void k_scatter(const double* bat, double* full, int ntot, int nm, const int* nb, const int* off) {
  QS_ELEM_PROLOGUE
  const int n = nb[b];
  if (i < n && j < n) full[(size_t)(off[b] + i) * ntot + off[b] + j] = bat[o];
}

__global__ void k_fock(const double* H, const double* J, const double* K, double* F, int nm, const double* sc) {
  QS_ELEM_PROLOGUE
  (void)i; (void)j;
  if (mol_done(sc, b)) return;
  F[o] = H[o] + 2.0 * J[o] - K[o];
}

// out = M - M^T
__global__ void k_antisym(const double* M, double* out, int nm, const double* sc) {
  QS_ELEM_PROLOGUE
  if (mol_done(sc, b)) return;
  out[o] = M[o] - M[(size_t)b * nm * nm + (size_t)j * nm + i];
}

__global__ void k_copy(const double* src, double* dst, int nm, const double* sc, int skip) {
  QS_ELEM_PROLOGUE
  (void)i; (void)j;
  if (skip && mol_done(sc, b)) return;
  dst[o] = src[o];
}

// F' diagonal += 1e-9 * i / (n - 1) (the reference's spectral shift,
// hartree_fock.py:38-40: diag(linspace(0, 1, n)) * 1e-9)
__global__ void k_shift(double* Fp, int nm, const int* nb, const double* sc) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * SB + threadIdx.x;
  const int n = nb[b];
  if (i >= n || mol_done(sc, b)) return;
  Fp[(size_t)b * nm * nm + (size_t)i * nm + i] += n > 1 ? 1e-9 * (double)i / (double)(n - 1) : 0.0;
}

// energy sum (F + H) D + E_nuc and, when err != null, the RMS of err over
// the molecule's n x n block; grid (nmol), block SB
__global__ __launch_bounds__(RB) void k_energy(const double* F, const double* H, const double* D, const double* err,
                                               int nm, const int* nb, double* sc) {
  __shared__ double sh[RB];
  const int b = blockIdx.x;
  if (mol_done(sc, b)) return;
  const int n = nb[b];
  const size_t base = (size_t)b * nm * nm;
  const long long nn = (long long)n * n;
  double se = 0.0, sr = 0.0;
  for (long long e0 = threadIdx.x; e0 < nn; e0 += 4 * RB) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long e = e0 + (long long)u * RB;
      if (e < nn) {
        const size_t o = base + (size_t)(e / n) * nm + (size_t)(e % n);
        se += (F[o] + H[o]) * D[o];
        if (err) sr += err[o] * err[o];
      }
    }
  }
  se = block_sum_rb(se, sh);
  sr = block_sum_rb(sr, sh);
  if (threadIdx.x == 0) {
    sc[b * QS_N + QS_ENEW] = se + sc[b * QS_N + QS_ENUC];
    if (err) sc[b * QS_N + QS_DRMS] = sqrt(sr / ((double)n * n));
  }
}

// ---- DIIS ----------------------------------------------------------------
// ring of m Fock / error matrices per molecule; iteration `it` stores into
// slot it % m; Bm[b] (m x m) holds the error dot products.
__global__ void k_diis_store(const double* F, const double* err, double* Fh, double* Eh, int nm, int m,
                             const double* sc) {
  QS_ELEM_PROLOGUE
  (void)i; (void)j;
  if (mol_done(sc, b)) return;
  const int slot = (int)sc[b * QS_N + QS_IT] % m;
  const size_t h = ((size_t)b * m + slot) * nm * nm + (size_t)e;
  Fh[h] = F[o];
  Eh[h] = err[o];
}

// grid (m, nmol), block SB: Bm[slot][jj] = Bm[jj][slot] = <E_slot, E_jj>
__global__ __launch_bounds__(RB) void k_diis_dots(const double* Eh, double* Bm, int nm, int m, const int* nb,
                                                  const double* sc) {
  __shared__ double sh[RB];
  const int b = blockIdx.y, jj = blockIdx.x;
  if (mol_done(sc, b)) return;
  const int it = (int)sc[b * QS_N + QS_IT];
  const int count = std::min(it + 1, m), slot = it % m;
  if (jj >= count) return;
  const int n = nb[b];
  const double* x = Eh + ((size_t)b * m + slot) * nm * nm;
  const double* y = Eh + ((size_t)b * m + jj) * nm * nm;
  const long long nn = (long long)n * n;
  double s = 0.0;
  for (long long e0 = threadIdx.x; e0 < nn; e0 += 4 * RB) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long e = e0 + (long long)u * RB;
      if (e < nn) {
        const size_t o = (size_t)(e / n) * nm + (size_t)(e % n);
        s += x[o] * y[o];
      }
    }
  }
  s = block_sum_rb(s, sh);
  if (threadIdx.x == 0) {
    Bm[((size_t)b * m + slot) * m + jj] = s;
    Bm[((size_t)b * m + jj) * m + slot] = s;
  }
}

// Pulay system [B -1; -1 0] [c; l] = [0; -1] over the filled slots, by
// Gaussian elimination with partial pivoting (one thread per molecule, the
// augmented matrix in LDS); fewer than two vectors or a zero pivot:
// c = e_slot (no extrapolation).  grid (nmol), block 64
constexpr int DIIS_MAX = 32;
// Pulay system of molecule b by Gaussian elimination with partial pivoting
// in LDS: one wavefront, the row updates of each pivot step spread over its
// lanes (each row's arithmetic as a serial elimination would do it), pivot
// choice and back substitution by lane 0.
__global__ __launch_bounds__(64) void k_diis_solve(const double* Bm, double* coef, int m, const double* sc) {
  const int b = blockIdx.x;
  if (mol_done(sc, b)) return;
  const int tid = threadIdx.x;
  const int it = (int)sc[b * QS_N + QS_IT];
  const int count = std::min(it + 1, m), slot = it % m;
  double* c = coef + (size_t)b * m;
  __shared__ double a[DIIS_MAX + 1][DIIS_MAX + 2];
  __shared__ int piv, singular;
  for (int k = tid; k < m; k += 64) c[k] = 0.0;
  __syncthreads();
  if (count < 2) {
    if (tid == 0) c[slot] = 1.0;
    return;
  }
  const int q = count + 1;
  for (int idx = tid; idx < q * (q + 1); idx += 64) {
    const int r = idx / (q + 1), k = idx % (q + 1);
    if (k < q)
      a[r][k] = (r < count && k < count) ? Bm[((size_t)b * m + r) * m + k] : (r == count && k == count ? 0.0 : -1.0);
    else
      a[r][q] = r == count ? -1.0 : 0.0;
  }
  __syncthreads();
  for (int k = 0; k < q; ++k) {
    if (tid == 0) {
      int p = k;
      for (int r = k + 1; r < q; ++r)
        if (fabs(a[r][k]) > fabs(a[p][k])) p = r;
      piv = p;
      singular = a[p][k] == 0.0;
    }
    __syncthreads();
    if (singular) {  // block-uniform
      if (tid == 0) c[slot] = 1.0;
      return;
    }
    const int p = piv;
    if (p != k)
      for (int t = tid; t <= q; t += 64) {
        const double v = a[k][t];
        a[k][t] = a[p][t];
        a[p][t] = v;
      }
    __syncthreads();
    for (int r = k + 1 + tid; r < q; r += 64) {
      const double f = a[r][k] / a[k][k];
      for (int t = k; t <= q; ++t) a[r][t] -= f * a[k][t];
    }
    __syncthreads();
  }
  if (tid == 0) {
    // back substitution in place: column q becomes the solution
    for (int r = q - 1; r >= 0; --r) {
      double v = a[r][q];
      for (int t = r + 1; t < q; ++t) v -= a[r][t] * a[t][q];
      a[r][q] = v / a[r][r];
    }
    for (int k = 0; k < count; ++k) c[k] = a[k][q];
  }
}

__global__ void k_diis_combine(const double* Fh, const double* coef, double* Fe, int nm, int m, const double* sc) {
  QS_ELEM_PROLOGUE
  (void)i; (void)j;
  if (mol_done(sc, b)) return;
  double s = 0.0;
  for (int k = 0; k < m; ++k) s += coef[(size_t)b * m + k] * Fh[((size_t)b * m + k) * nm * nm + (size_t)e];
  Fe[o] = s;
}

// ---- purification ----------------------------------------------------------
// Gershgorin bounds of F' (grid nmol, block SB)
__global__ __launch_bounds__(RB) void k_bounds(const double* Fp, int nm, const int* nb, double* sc) {
  __shared__ double lo[RB / 64], hi[RB / 64];
  const int b = blockIdx.x;
  if (mol_done(sc, b)) return;
  const int n = nb[b];
  const double* F = Fp + (size_t)b * nm * nm;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double mn = 1e300, mx = -1e300;
  // a wavefront per row: lanes over the columns, four loads in flight each
  for (int i = wave; i < n; i += RB / 64) {
    const double* row = F + (size_t)i * nm;
    double r = 0.0;
    for (int j0 = lane; j0 < n; j0 += 4 * 64) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = j0 + u * 64;
        if (j < n && j != i) r += fabs(row[j]);
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off);
    const double d = row[i];
    mn = fmin(mn, d - r);
    mx = fmax(mx, d + r);
  }
  if (lane == 0) {
    lo[wave] = mn;
    hi[wave] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = lo[0], z = hi[0];
    for (int w = 1; w < RB / 64; ++w) {
      a = fmin(a, lo[w]);
      z = fmax(z, hi[w]);
    }
    sc[b * QS_N + QS_EMIN] = a;
    sc[b * QS_N + QS_EMAX] = z;
  }
}

__global__ void k_tc2_init(const double* Fp, double* X, int nm, const int* nb, double* sc) {
  QS_ELEM_PROLOGUE
  if (mol_done(sc, b)) return;
  if (e == 0) sc[b * QS_N + QS_PCONV0] = 0.0;
  const int n = nb[b];
  const double emin = sc[b * QS_N + QS_EMIN], emax = sc[b * QS_N + QS_EMAX];
  const double w = emax > emin ? emax - emin : 1.0;
  X[o] = (i < n && j < n) ? ((i == j ? emax : 0.0) - Fp[o]) / w : 0.0;
}

// One TC2 step k: Xout = X2 (= X^2) when |tr X2 - nocc| <= |2 tr X - tr X2 - nocc|,
// else 2 X - X2.  Every block of a molecule sums the same diagonals of the
// unmodified X and X2 (Xout is a third buffer) in the same order, so all
// take the same branch.  Each computed step records tr(X - X2) (the
// idempotency defect before its update); once that is below PURIFY_TOL the
// later steps (and their GEMMs) only carry X over.
__global__ __launch_bounds__(SB) void k_tc2_step(const double* X, const double* X2, double* Xout, int nm,
                                                 const int* nb, double* sc, int k) {
  __shared__ double sh[SB];
  const int b = blockIdx.y;
  if (mol_done(sc, b)) return;
  const int n = nb[b];
  const size_t base = (size_t)b * nm * nm;
  const long long e0 = (long long)blockIdx.x * SB + threadIdx.x;
  double* flag_next = &sc[b * QS_N + QS_PCONV0 + ((k + 1) & 1)];
  if (sc[b * QS_N + QS_PCONV0 + (k & 1)] != 0.0) {  // converged earlier: carry X over
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag_next = 1.0;
    if (e0 < (long long)nm * nm) Xout[base + (size_t)e0] = X[base + (size_t)e0];
    return;
  }
  double t1 = 0.0, t2 = 0.0;
  for (int i = threadIdx.x; i < n; i += SB) {
    t1 += X[base + (size_t)i * nm + i];
    t2 += X2[base + (size_t)i * nm + i];
  }
  t1 = block_sum(t1, sh);
  t2 = block_sum(t2, sh);
  const double nocc = sc[b * QS_N + QS_NOCC];
  const bool sq = fabs(t2 - nocc) <= fabs(2.0 * t1 - t2 - nocc);
  if (blockIdx.x == 0 && threadIdx.x == 0) *flag_next = fabs(t1 - t2) < PURIFY_TOL ? 1.0 : 0.0;
  if (e0 >= (long long)nm * nm) return;
  const size_t o = base + (size_t)e0;
  Xout[o] = sq ? X2[o] : 2.0 * X[o] - X2[o];
}

// McWeeny step Xout = 3 X2 - 2 X3 (X2 = X^2, X3 = X^3): squares the distance
// of every eigenvalue from 0 or 1 and keeps the subspaces.  The last step
// records tr(X - X2) of its input, the defect left (block 0 of each molecule
// writes it; no block reads it).
__global__ __launch_bounds__(SB) void k_mcweeny(const double* X, const double* X2, const double* X3, double* Xout,
                                                int nm, const int* nb, double* sc, int last) {
  __shared__ double sh[SB];
  const int b = blockIdx.y;
  if (mol_done(sc, b)) return;
  const size_t base = (size_t)b * nm * nm;
  if (last) {
    const int n = nb[b];
    double t1 = 0.0, t2 = 0.0;
    for (int i = threadIdx.x; i < n; i += SB) {
      t1 += X[base + (size_t)i * nm + i];
      t2 += X2[base + (size_t)i * nm + i];
    }
    t1 = block_sum(t1, sh);
    t2 = block_sum(t2, sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[b * QS_N + QS_IDEM] = t1 - t2;
  }
  const long long e = (long long)blockIdx.x * SB + threadIdx.x;
  if (e >= (long long)nm * nm) return;
  const size_t o = base + (size_t)e;
  Xout[o] = 3.0 * X2[o] - 2.0 * X3[o];
}

// end of an iteration: energy change, counters, convergence; skip[b] (the
// J/K pass's per-fragment mask) follows the done flag (grid ceil(nmol / 64),
// block 64)
__global__ void k_finish(double* sc, int* skip, int nmol, double conv, int maxit) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nmol || mol_done(sc, b)) return;
  double* s = sc + b * QS_N;
  s[QS_DE] = s[QS_E] - s[QS_ENEW];
  s[QS_E] = s[QS_ENEW];
  s[QS_IT] += 1.0;
  const bool conv_ok = !(fabs(s[QS_DE]) > conv || s[QS_DRMS] > conv);
  if (conv_ok || s[QS_IT] >= maxit) {
    s[QS_DONE] = 1.0;
    skip[b] = 1;
  }
}

}  // namespace

struct qx_scf {
  qx_ctx* ctx = nullptr;
  int nmol = 0, nm = 0, ntot = 0;
  int diis = 0, m = 0, shift = 0, maxit = 0, purify = 0, precision = 0;
  double conv = 0.0, thr = 0.0;
  std::vector<int> h_nb;
  int* nb = nullptr;
  int* off = nullptr;
  int* skip = nullptr;          // per molecule (= fragment id): 1 once done; masks the J/K pass of batches
  // B x nm x nm
  double *H = nullptr, *S = nullptr, *A = nullptr;
  double *D = nullptr, *Dold = nullptr, *Dnew = nullptr, *F = nullptr, *Fe = nullptr, *J = nullptr, *K = nullptr;
  double *T1 = nullptr, *T2 = nullptr, *err = nullptr, *X = nullptr, *Xb = nullptr, *X2 = nullptr;
  double *Fh = nullptr, *Eh = nullptr, *Bm = nullptr, *coef = nullptr;
  // ntot x ntot
  double *Dfull = nullptr, *Jfull = nullptr, *Kfull = nullptr;
  double* sc = nullptr;
  // split-K GEMMs (few tiles per launch): splits and partial tiles
  int ksplit = 1;
  double* gpart = nullptr;
  std::vector<void*> bufs;
  bool initialised = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string err_msg;
};

namespace {

struct ScfError {
  int code;
  std::string msg;
};

#define QS_HIP(x)                                                                           \
  do {                                                                                      \
    const hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                                   \
      throw ScfError{e_ == hipErrorOutOfMemory ? QX_ENOMEM : QX_EHIP,                       \
                     std::string(#x) + ": " + hipGetErrorString(e_)};                       \
  } while (0)

thread_local std::string g_scf_error;

template <class T>
T* salloc(qx_scf* s, size_t n) {
  void* p = nullptr;
  QS_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  s->bufs.push_back(p);
  return static_cast<T*>(p);
}

template <class Fn>
int scf_guarded(qx_scf* s, Fn&& fn) noexcept {
  try {
    return fn();
  } catch (const ScfError& e) {
    (s ? s->err_msg : g_scf_error) = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    (s ? s->err_msg : g_scf_error) = e.what();
    return QX_EHIP;
  } catch (...) {
    (s ? s->err_msg : g_scf_error) = "unknown error";
    return QX_EHIP;
  }
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw ScfError{QX_EHIP, std::string(what) + ": " + hipGetErrorString(e)};
}

dim3 egrid(const qx_scf* s) { return dim3((unsigned)(((long long)s->nm * s->nm + SB - 1) / SB), (unsigned)s->nmol); }

// C = alpha op(A) op(B) + beta C over the batch; pslot >= 0: purification
// step pslot's GEMM (skipped for molecules whose purification converged)
void gemm(qx_scf* s, bool ta, bool tb, const double* A, const double* B, double* C, double alpha, double beta,
          hipStream_t st, int pslot = -1) {
  const int ns = s->ksplit;
  const dim3 grid((unsigned)((s->nm + BN - 1) / BN), (unsigned)((s->nm + BM - 1) / BM), (unsigned)(s->nmol * ns));
  const int ps = pslot < 0 ? -1 : (pslot & 1);
  double* pt = s->gpart;
  if (!ta && !tb) hipLaunchKernelGGL((k_bgemm<false, false>), grid, dim3(GT), 0, st, s->nm, s->nb, A, B, C, alpha, beta, s->sc, 1, ps, ns, pt);
  else if (ta && !tb) hipLaunchKernelGGL((k_bgemm<true, false>), grid, dim3(GT), 0, st, s->nm, s->nb, A, B, C, alpha, beta, s->sc, 1, ps, ns, pt);
  else if (!ta && tb) hipLaunchKernelGGL((k_bgemm<false, true>), grid, dim3(GT), 0, st, s->nm, s->nb, A, B, C, alpha, beta, s->sc, 1, ps, ns, pt);
  else hipLaunchKernelGGL((k_bgemm<true, true>), grid, dim3(GT), 0, st, s->nm, s->nb, A, B, C, alpha, beta, s->sc, 1, ps, ns, pt);
  if (ns > 1)
    hipLaunchKernelGGL(k_bgemm_reduce, dim3(grid.x, grid.y, (unsigned)s->nmol), dim3(GT), 0, st, s->nm, s->nb, C, alpha,
                       beta, s->sc, 1, ps, ns, pt);
  check(hipGetLastError(), "scf gemm");
}

void jk(qx_scf* s, const double* Dsrc, hipStream_t st) {
  hipLaunchKernelGGL(k_scatter, egrid(s), dim3(SB), 0, st, Dsrc, s->Dfull, s->ntot, s->nm, s->nb, s->off);
  check(hipGetLastError(), "scf scatter");
  const int rc = s->nmol > 1
                     ? qx_jk_build_masked(s->ctx, s->Dfull, s->Jfull, s->Kfull, s->precision, s->thr, s->skip, st)
                     : qx_jk_build(s->ctx, s->Dfull, s->Jfull, s->Kfull, s->precision, s->thr, st);
  if (rc != QX_OK) throw ScfError{rc, std::string("scf J/K: ") + qx_last_error(s->ctx)};
  hipLaunchKernelGGL(k_gather, egrid(s), dim3(SB), 0, st, s->Jfull, s->ntot, s->J, s->nm, s->nb, s->off, s->sc, 1);
  hipLaunchKernelGGL(k_gather, egrid(s), dim3(SB), 0, st, s->Kfull, s->ntot, s->K, s->nm, s->nb, s->off, s->sc, 1);
  check(hipGetLastError(), "scf gather");
}

// density of F (Fsrc) into Dnew: F' = A^T F A (+ shift), TC2, D = A X A^T
void density(qx_scf* s, const double* Fsrc, hipStream_t st) {
  gemm(s, false, false, Fsrc, s->A, s->T1, 1.0, 0.0, st);   // F A
  gemm(s, true, false, s->A, s->T1, s->T2, 1.0, 0.0, st);   // A^T F A
  if (s->shift) {
    hipLaunchKernelGGL(k_shift, dim3((unsigned)((s->nm + SB - 1) / SB), (unsigned)s->nmol), dim3(SB), 0, st, s->T2,
                       s->nm, s->nb, s->sc);
  }
  hipLaunchKernelGGL(k_bounds, dim3((unsigned)s->nmol), dim3(RB), 0, st, s->T2, s->nm, s->nb, s->sc);
  hipLaunchKernelGGL(k_tc2_init, egrid(s), dim3(SB), 0, st, s->T2, s->X, s->nm, s->nb, s->sc);
  check(hipGetLastError(), "scf purification setup");
  double* cur = s->X;
  double* nxt = s->Xb;
  for (int k = 0; k < s->purify; ++k) {
    gemm(s, false, false, cur, cur, s->X2, 1.0, 0.0, st, k);
    hipLaunchKernelGGL(k_tc2_step, egrid(s), dim3(SB), 0, st, cur, s->X2, nxt, s->nm, s->nb, s->sc, k);
    check(hipGetLastError(), "scf purification step");
    std::swap(cur, nxt);
  }
  for (int k = 0; k < MCWEENY_STEPS; ++k) {
    gemm(s, false, false, cur, cur, s->X2, 1.0, 0.0, st);
    gemm(s, false, false, s->X2, cur, s->T1, 1.0, 0.0, st);
    hipLaunchKernelGGL(k_mcweeny, egrid(s), dim3(SB), 0, st, cur, s->X2, s->T1, nxt, s->nm, s->nb, s->sc,
                       k == MCWEENY_STEPS - 1 ? 1 : 0);
    check(hipGetLastError(), "scf purification refinement");
    std::swap(cur, nxt);
  }
  gemm(s, false, false, s->A, cur, s->T1, 1.0, 0.0, st);    // A X
  gemm(s, false, true, s->T1, s->A, s->Dnew, 1.0, 0.0, st); // A X A^T
}

// one loop body of the reference's SCF (hartree_fock.py:83-99 with DIIS)
void iteration(qx_scf* s, hipStream_t st) {
  jk(s, s->D, st);
  hipLaunchKernelGGL(k_fock, egrid(s), dim3(SB), 0, st, s->H, s->J, s->K, s->F, s->nm, s->sc);
  check(hipGetLastError(), "scf fock");
  // DIIS error A (F D S - S D F) A of the density that built F (the
  // reference's DIIS_Err, hartree_fock.py:59-62, with A on both sides)
  gemm(s, false, false, s->F, s->D, s->T1, 1.0, 0.0, st);    // F D
  gemm(s, false, false, s->T1, s->S, s->T2, 1.0, 0.0, st);   // F D S
  hipLaunchKernelGGL(k_antisym, egrid(s), dim3(SB), 0, st, s->T2, s->T1, s->nm, s->sc);
  gemm(s, false, false, s->T1, s->A, s->T2, 1.0, 0.0, st);   // e A
  gemm(s, false, false, s->A, s->T2, s->err, 1.0, 0.0, st);  // A e A
  hipLaunchKernelGGL(k_energy, dim3((unsigned)s->nmol), dim3(RB), 0, st, s->F, s->H, s->D, s->err, s->nm, s->nb,
                     s->sc);
  check(hipGetLastError(), "scf energy");
  const double* Fuse = s->F;
  if (s->diis) {
    hipLaunchKernelGGL(k_diis_store, egrid(s), dim3(SB), 0, st, s->F, s->err, s->Fh, s->Eh, s->nm, s->m, s->sc);
    hipLaunchKernelGGL(k_diis_dots, dim3((unsigned)s->m, (unsigned)s->nmol), dim3(RB), 0, st, s->Eh, s->Bm, s->nm,
                       s->m, s->nb, s->sc);
    hipLaunchKernelGGL(k_diis_solve, dim3((unsigned)s->nmol), dim3(64), 0, st, s->Bm, s->coef, s->m, s->sc);
    hipLaunchKernelGGL(k_diis_combine, egrid(s), dim3(SB), 0, st, s->Fh, s->coef, s->Fe, s->nm, s->m, s->sc);
    check(hipGetLastError(), "scf diis");
    Fuse = s->Fe;
  } else {
    hipLaunchKernelGGL(k_copy, egrid(s), dim3(SB), 0, st, s->F, s->Fe, s->nm, s->sc, 1);
  }
  density(s, Fuse, st);
  hipLaunchKernelGGL(k_copy, egrid(s), dim3(SB), 0, st, s->D, s->Dold, s->nm, s->sc, 1);
  hipLaunchKernelGGL(k_copy, egrid(s), dim3(SB), 0, st, s->Dnew, s->D, s->nm, s->sc, 1);
  hipLaunchKernelGGL(k_finish, dim3((unsigned)((s->nmol + 63) / 64)), dim3(64), 0, st, s->sc, s->skip, s->nmol,
                     s->conv, s->maxit);
  check(hipGetLastError(), "scf finish");
}

}  // namespace

extern "C" {

int qx_scf_create(qx_ctx* ctx, int nmol, int nm, const int* nbf, const int* offset, const int* nocc,
                  const double* enuc, const double* H, const double* S, const double* A, int diis, int diis_space,
                  int spectral_shift, double conv, int maxit, int purify_iters, int precision,
                  double fp32_threshold, qx_scf** out) {
  if (!out) return QX_EINVAL;
  *out = nullptr;
  if (!ctx || nmol < 1 || nm < 1 || !nbf || !offset || !nocc || !enuc || !H || !S || !A) return QX_EINVAL;
  if (diis && (diis_space < 2 || diis_space > DIIS_MAX)) return QX_EINVAL;
  if (maxit < 0 || purify_iters < 1 || purify_iters > 1000 || !(conv >= 0.0)) return QX_EINVAL;
  if (precision != QX_PREC_FP64 && precision != QX_PREC_MIXED) return QX_EINVAL;
  int ntot = 0, nb_ = 0;
  if (qx_nbf(ctx, &ntot, &nb_) != QX_OK) return QX_EINVAL;
  for (int b = 0; b < nmol; ++b) {
    if (nbf[b] < 1 || nbf[b] > nm || offset[b] < 0 || offset[b] + nbf[b] > ntot) return QX_EINVAL;
    if (nocc[b] < 0 || nocc[b] > nbf[b]) return QX_EINVAL;
  }
  if (nmol > 1) {
    // a batch: molecule b is fragment b of ctx and owns AOs [offset[b], offset[b] + nbf[b])
    if (!ctx->fragmented) return QX_EINVAL;
    for (int sh = 0; sh < ctx->nshell; ++sh) {
      const int f = ctx->frag[ctx->atom[sh]], l = ctx->l[sh];
      const int size = ctx->pure ? 2 * l + 1 : (l + 1) * (l + 2) / 2;
      if (f < 0 || f >= nmol || ctx->bfs[sh] < offset[f] || ctx->bfs[sh] + size > offset[f] + nbf[f])
        return QX_EINVAL;
    }
  }
  qx_scf* s = new (std::nothrow) qx_scf();
  if (!s) return QX_ENOMEM;
  const int rc = scf_guarded(nullptr, [&]() -> int {
    s->ctx = ctx; s->nmol = nmol; s->nm = nm; s->ntot = ntot;
    s->diis = diis ? 1 : 0; s->m = diis ? diis_space : 1; s->shift = spectral_shift ? 1 : 0;
    s->conv = conv; s->maxit = maxit; s->purify = purify_iters; s->precision = precision; s->thr = fp32_threshold;
    s->h_nb.assign(nbf, nbf + nmol);
    const size_t nb2 = (size_t)nmol * nm * nm;
    s->nb = salloc<int>(s, nmol);
    s->off = salloc<int>(s, nmol);
    s->skip = salloc<int>(s, nmol);
    QS_HIP(hipMemset(s->skip, 0, nmol * sizeof(int)));
    QS_HIP(hipMemcpy(s->nb, nbf, nmol * sizeof(int), hipMemcpyHostToDevice));
    QS_HIP(hipMemcpy(s->off, offset, nmol * sizeof(int), hipMemcpyHostToDevice));
    for (double** p : {&s->H, &s->S, &s->A, &s->D, &s->Dold, &s->Dnew, &s->F, &s->Fe, &s->J, &s->K, &s->T1, &s->T2,
                       &s->err, &s->X, &s->Xb, &s->X2})
      *p = salloc<double>(s, nb2);
    QS_HIP(hipMemcpy(s->H, H, nb2 * sizeof(double), hipMemcpyDeviceToDevice));
    QS_HIP(hipMemcpy(s->S, S, nb2 * sizeof(double), hipMemcpyDeviceToDevice));
    QS_HIP(hipMemcpy(s->A, A, nb2 * sizeof(double), hipMemcpyDeviceToDevice));
    s->Fh = salloc<double>(s, nb2 * s->m);
    s->Eh = salloc<double>(s, nb2 * s->m);
    s->Bm = salloc<double>(s, (size_t)nmol * s->m * s->m);
    s->coef = salloc<double>(s, (size_t)nmol * s->m);
    // every matrix starts at zero: the padding rows and columns, which no
    // GEMM writes, then stay zero through every elementwise step (and the
    // DIIS ring's unfilled slots enter its combination as zeros)
    for (double* p : {s->D, s->Dold, s->Dnew, s->F, s->Fe, s->J, s->K, s->T1, s->T2, s->err, s->X, s->Xb, s->X2})
      QS_HIP(hipMemset(p, 0, nb2 * sizeof(double)));
    // split K when one launch has too few tiles to spread over the CUs
    // (about 256 workgroups wanted): at most KSPLIT_MAX splits, each with at
    // least one K step
    {
      const int tiles = (nm + BM - 1) / BM, nk = (nm + BK - 1) / BK;
      const long long wg = (long long)tiles * tiles * nmol;
      int ns = (int)std::min<long long>(KSPLIT_MAX, std::max<long long>(1, (256 + wg - 1) / wg));
      ns = std::min(ns, nk);
      const int per = (nk + ns - 1) / ns;
      s->ksplit = (nk + per - 1) / per;  // no empty split
      if (s->ksplit > 1) s->gpart = salloc<double>(s, (size_t)wg * s->ksplit * PART_TILE);
    }
    QS_HIP(hipMemset(s->Fh, 0, nb2 * s->m * sizeof(double)));
    QS_HIP(hipMemset(s->Eh, 0, nb2 * s->m * sizeof(double)));
    QS_HIP(hipMemset(s->Bm, 0, (size_t)nmol * s->m * s->m * sizeof(double)));
    QS_HIP(hipMemset(s->coef, 0, (size_t)nmol * s->m * sizeof(double)));
    const size_t nt2 = (size_t)ntot * ntot;
    s->Dfull = salloc<double>(s, nt2);
    s->Jfull = salloc<double>(s, nt2);
    s->Kfull = salloc<double>(s, nt2);
    QS_HIP(hipMemset(s->Dfull, 0, nt2 * sizeof(double)));  // off-diagonal blocks stay zero
    std::vector<double> sc((size_t)nmol * QS_N, 0.0);
    for (int b = 0; b < nmol; ++b) {
      sc[b * QS_N + QS_ENUC] = enuc[b];
      sc[b * QS_N + QS_NOCC] = nocc[b];
      sc[b * QS_N + QS_DE] = 1.0;
      sc[b * QS_N + QS_DRMS] = 1.0;
    }
    s->sc = salloc<double>(s, sc.size());
    QS_HIP(hipMemcpy(s->sc, sc.data(), sc.size() * sizeof(double), hipMemcpyHostToDevice));
    return QX_OK;
  });
  if (rc != QX_OK) {
    qx_scf_destroy(s);
    return rc;
  }
  *out = s;
  return QX_OK;
}

// D0 (B x nm x nm, device): starting density.  The reference's first step
// (hartree_fock.py:66-78): F from D0, E0 = sum (F + H) D0 + E_nuc, the first
// density from F; runs eagerly and leaves the context's J/K workspace,
// streams and mixed-precision lists allocated for capture.
int qx_scf_init(qx_scf* s, const double* D0, void* stream) {
  if (!s || !D0) return QX_EINVAL;
  return scf_guarded(s, [&]() -> int {
    hipStream_t st = (hipStream_t)stream;
    const size_t nb2 = (size_t)s->nmol * s->nm * s->nm;
    QS_HIP(hipMemcpyAsync(s->D, D0, nb2 * sizeof(double), hipMemcpyDeviceToDevice, st));
    jk(s, s->D, st);
    hipLaunchKernelGGL(k_fock, egrid(s), dim3(SB), 0, st, s->H, s->J, s->K, s->F, s->nm, s->sc);
    hipLaunchKernelGGL(k_energy, dim3((unsigned)s->nmol), dim3(RB), 0, st, s->F, s->H, s->D, (const double*)nullptr,
                       s->nm, s->nb, s->sc);
    check(hipGetLastError(), "scf init");
    density(s, s->F, st);
    hipLaunchKernelGGL(k_copy, egrid(s), dim3(SB), 0, st, s->F, s->Fe, s->nm, s->sc, 1);
    hipLaunchKernelGGL(k_copy, egrid(s), dim3(SB), 0, st, s->D, s->Dold, s->nm, s->sc, 1);
    hipLaunchKernelGGL(k_copy, egrid(s), dim3(SB), 0, st, s->Dnew, s->D, s->nm, s->sc, 1);
    // E = E_new; iteration counters stay at zero
    std::vector<double> sc((size_t)s->nmol * QS_N);
    QS_HIP(hipMemcpyAsync(sc.data(), s->sc, sc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    QS_HIP(hipStreamSynchronize(st));
    for (int b = 0; b < s->nmol; ++b) {
      sc[b * QS_N + QS_E] = sc[b * QS_N + QS_ENEW];
      sc[b * QS_N + QS_IT] = 0.0;
      sc[b * QS_N + QS_DONE] = s->maxit == 0 ? 1.0 : 0.0;
      sc[b * QS_N + QS_DE] = 1.0;
      sc[b * QS_N + QS_DRMS] = 1.0;
    }
    QS_HIP(hipMemcpy(s->sc, sc.data(), sc.size() * sizeof(double), hipMemcpyHostToDevice));
    std::vector<int> sk(s->nmol, s->maxit == 0 ? 1 : 0);
    QS_HIP(hipMemcpy(s->skip, sk.data(), sk.size() * sizeof(int), hipMemcpyHostToDevice));
    s->initialised = true;
    return QX_OK;
  });
}

// Captures one iteration into a graph on `stream` (after qx_scf_init).
int qx_scf_capture(qx_scf* s, void* stream) {
  if (!s || !s->initialised) return QX_EINVAL;
  if (s->exec) return QX_OK;
  return scf_guarded(s, [&]() -> int {
    hipStream_t st = (hipStream_t)stream;
    if (s->ctx->profiling) throw ScfError{QX_EINVAL, "capture with kernel profiling on"};
    QS_HIP(hipStreamSynchronize(st));
    QS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    s->ctx->capturing = true;
    try {
      iteration(s, st);
    } catch (...) {
      s->ctx->capturing = false;
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    s->ctx->capturing = false;
    QS_HIP(hipStreamEndCapture(st, &s->graph));
    QS_HIP(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
    return QX_OK;
  });
}

// n iterations: graph replays when use_graph (captures on first use), else
// the same kernels launched directly.  Converged molecules are skipped.
int qx_scf_step(qx_scf* s, int n, int use_graph, void* stream) {
  if (!s || !s->initialised || n < 0) return QX_EINVAL;
  if (use_graph && !s->exec) {
    const int rc = qx_scf_capture(s, stream);
    if (rc != QX_OK) return rc;
  }
  return scf_guarded(s, [&]() -> int {
    hipStream_t st = (hipStream_t)stream;
    for (int k = 0; k < n; ++k) {
      if (use_graph) QS_HIP(hipGraphLaunch(s->exec, st));
      else iteration(s, st);
    }
    return QX_OK;
  });
}

// Synchronises `stream` and reads the per-molecule state (host arrays of
// nmol entries, any may be null).
int qx_scf_status(qx_scf* s, double* energy, double* de, double* drms, int* iterations, int* done,
                  double* idempotency, void* stream) {
  if (!s) return QX_EINVAL;
  return scf_guarded(s, [&]() -> int {
    std::vector<double> sc((size_t)s->nmol * QS_N);
    QS_HIP(hipMemcpyAsync(sc.data(), s->sc, sc.size() * sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)stream));
    QS_HIP(hipStreamSynchronize((hipStream_t)stream));
    for (int b = 0; b < s->nmol; ++b) {
      const double* v = &sc[b * QS_N];
      if (energy) energy[b] = v[QS_E];
      if (de) de[b] = v[QS_DE];
      if (drms) drms[b] = v[QS_DRMS];
      if (iterations) iterations[b] = (int)v[QS_IT];
      if (done) done[b] = v[QS_DONE] != 0.0;
      if (idempotency) idempotency[b] = v[QS_IDEM];
    }
    return QX_OK;
  });
}

// Copies a B x nm x nm state matrix to out (device): 0 = D (the density of
// the last Fock matrix), 1 = D_old (the density that built it), 2 = F (the
// last Fock matrix, DIIS-extrapolated when DIIS is on), 3 = the Fock matrix
// before extrapolation.
int qx_scf_matrix(qx_scf* s, int which, double* out, void* stream) {
  if (!s || !out || which < 0 || which > 3) return QX_EINVAL;
  return scf_guarded(s, [&]() -> int {
    const double* src = which == 0 ? s->D : which == 1 ? s->Dold : which == 2 ? s->Fe : s->F;
    QS_HIP(hipMemcpyAsync(out, src, (size_t)s->nmol * s->nm * s->nm * sizeof(double), hipMemcpyDeviceToDevice,
                          (hipStream_t)stream));
    return QX_OK;
  });
}

void qx_scf_destroy(qx_scf* s) {
  if (!s) return;
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  for (void* p : s->bufs) (void)hipFree(p);
  delete s;
}

const char* qx_scf_last_error(const qx_scf* s) { return s ? s->err_msg.c_str() : g_scf_error.c_str(); }

}  // extern "C"
