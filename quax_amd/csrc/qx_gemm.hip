// FP64 MFMA GEMM for the AO -> MO integral transformation (reference:
// quax/methods/energy_utils.py:46-73; callers mp2.py:29, ccsd.py:32).
//
// One transformation step contracts the last index of a row-major tensor X
// (R rows of n) with a coefficient matrix C (n x m, row-major) and puts the
// new index first:  Y[M, r] = sum_s C[s, M] X[r, s]  -- the reference's
// jnp.tensordot(C, X, axes=[(0,), (3,)]).  As a GEMM, Y (m x R) = C^T X^T:
// both operands are read "transposed", which the kernel does directly from
// LDS tiles (no transposed copies).
//
// Kernel: 64 x 64 output tile per 256-thread workgroup (four 64-wide
// wavefronts, 32 x 32 each as 2 x 2 tiles of v_mfma_f64_16x16x4f64), K in
// steps of 32 staged through LDS.  MFMA operand layout (16x16x4, FP64,
// checked on gfx950 with a one-wavefront probe): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator element r of lane l is
// D[4 r + l / 16][l % 16] (rows interleaved across the four 16-lane groups).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "quax_amd.h"

namespace {

typedef double qx_dv4 __attribute__((ext_vector_type(4)));

constexpr int GM_BM = 64, GM_BN = 64, GM_BK = 32, GM_THREADS = 256;
constexpr int GM_PAD = 2;  // LDS row padding (doubles) against bank conflicts

// out[i * ldo + j] = sum_k opA(i, k) opB(k, j), i < M, j < N, k < K, with
// opA(i, k) = A[k * lda + i] (A stored K x M) and opB(k, j) = B[j * ldb + k]
// (B stored N x K).  Grid: (ceil(N / 64), ceil(M / 64)).
__global__ __launch_bounds__(GM_THREADS) void qx_dgemm_tt_kernel(int M, long long N, int K, const double* __restrict__ A,
                                                                  long long lda, const double* __restrict__ B,
                                                                  long long ldb, double* __restrict__ out,
                                                                  long long ldo) {
  __shared__ double As[GM_BK][GM_BM + GM_PAD];
  __shared__ double Bs[GM_BK][GM_BN + GM_PAD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;  // wave's 32 x 32 quadrant
  const long long i0 = (long long)blockIdx.y * GM_BM;
  const long long j0 = (long long)blockIdx.x * GM_BN;
  qx_dv4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = qx_dv4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += GM_BK) {
    // A tile: rows k (32) x cols i (64), contiguous along i in memory
#pragma unroll
    for (int r = 0; r < GM_BK * GM_BM / GM_THREADS; ++r) {
      const int idx = tid + r * GM_THREADS;
      const int k = idx / GM_BM, i = idx % GM_BM;
      const long long gi = i0 + i;
      const int gk = k0 + k;
      As[k][i] = (gi < M && gk < K) ? A[(long long)gk * lda + gi] : 0.0;
    }
    // B tile: for each output column j, 32 consecutive k
#pragma unroll
    for (int r = 0; r < GM_BK * GM_BN / GM_THREADS; ++r) {
      const int idx = tid + r * GM_THREADS;
      const int j = idx / GM_BK, k = idx % GM_BK;
      const long long gj = j0 + j;
      const int gk = k0 + k;
      Bs[k][j] = (gj < N && gk < K) ? B[gj * ldb + gk] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GM_BK; kk += 4) {
      const int ka = kk + (lane >> 4);
      double a[2], b[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        a[t] = As[ka][wr * 32 + t * 16 + (lane & 15)];
        b[t] = Bs[ka][wc * 32 + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta], b[tb], acc[ta][tb], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long gi = i0 + wr * 32 + ta * 16 + 4 * r + (lane >> 4);
        const long long gj = j0 + wc * 32 + tb * 16 + (lane & 15);
        if (gi < M && gj < N) out[gi * ldo + gj] = acc[ta][tb][r];
      }
}

thread_local std::string g_gemm_error;

int gemm_tt(int M, long long N, int K, const double* A, const double* B, double* out, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return QX_OK;
  const long long gx = (N + GM_BN - 1) / GM_BN, gy = (M + GM_BM - 1) / GM_BM;
  if (gy > 65535 || gx > 0x7fffffffLL) {
    g_gemm_error = "tei transformation: tensor too large for one launch";
    return QX_EINVAL;
  }
  hipLaunchKernelGGL(qx_dgemm_tt_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(GM_THREADS), 0, s, M, N, K, A,
                     (long long)M, B, (long long)K, out, N);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_gemm_error = std::string("tei transformation: ") + hipGetErrorString(e);
    return QX_EHIP;
  }
  return QX_OK;
}

}  // namespace

extern "C" {

long long qx_tei_transform_work(const int* n, const int* m) {
  if (!n || !m) return -1;
  for (int k = 0; k < 4; ++k)
    if (n[k] <= 0 || m[k] <= 0) return -1;
  const long long t1 = (long long)m[3] * n[0] * n[1] * n[2];
  const long long t2 = (long long)m[2] * m[3] * n[0] * n[1];
  const long long t3 = (long long)m[1] * m[2] * m[3] * n[0];
  return std::max(t1, t3) + t2;
}

int qx_partial_tei_transformation(const double* G, const int* n, const double* Ci, const double* Cj,
                                  const double* Ck, const double* Cl, const int* m, double* out, double* work,
                                  void* stream) {
  if (!G || !n || !m || !Ci || !Cj || !Ck || !Cl || !out || !work) {
    g_gemm_error = "tei transformation: null argument";
    return QX_EINVAL;
  }
  const long long wsize = qx_tei_transform_work(n, m);
  if (wsize < 0) {
    g_gemm_error = "tei transformation: dimensions must be positive";
    return QX_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  double* w1 = work;                                                   // T1, then T3
  double* w2 = work + std::max((long long)m[3] * n[0] * n[1] * n[2],  // T2
                               (long long)m[1] * m[2] * m[3] * n[0]);
  // T1[L, p, q, r] = sum_s Cl[s, L] G[p, q, r, s]
  int rc = gemm_tt(m[3], (long long)n[0] * n[1] * n[2], n[3], Cl, G, w1, s);
  // T2[K, L, p, q] = sum_r Ck[r, K] T1[L, p, q, r]
  if (rc == QX_OK) rc = gemm_tt(m[2], (long long)m[3] * n[0] * n[1], n[2], Ck, w1, w2, s);
  // T3[J, K, L, p] = sum_q Cj[q, J] T2[K, L, p, q]
  if (rc == QX_OK) rc = gemm_tt(m[1], (long long)m[2] * m[3] * n[0], n[1], Cj, w2, w1, s);
  // out[I, J, K, L] = sum_p Ci[p, I] T3[J, K, L, p]
  if (rc == QX_OK) rc = gemm_tt(m[0], (long long)m[1] * m[2] * m[3], n[0], Ci, w1, out, s);
  return rc;
}

const char* qx_gemm_last_error(void) { return g_gemm_error.c_str(); }

}  // extern "C"
