// Host-side context of the ERI library and the helper-kernel entry points
// shared between qx_runtime.hip and qx_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "qx_device.h"

#define QX_LMAX_SUPPORTED 2

// Per-shell and per-function tables for the one-electron kernels and the
// spherical transforms.
struct QxShellsDev {
  int nshell, natom, ncart, nbf;
  const int* l; const int* atom; const int* nprim; const int* poff;
  const int* bfc;                         // first Cartesian function of each shell
  const int* bfs;                         // first AO (spherical when pure) function
  const double* exps; const double* coefs;
  const double* xyz;                      // 3*natom
  const double* Z;                        // natom
  const int* f_shell;                     // Cartesian function -> shell
  const int* f_comp;                      // Cartesian function -> packed (lx,ly,lz)
  const int* ao_shell;                    // AO function -> shell
  const double* sph;                      // per-l transform blocks, offsets below
  int sph_off[QX_LMAX_SUPPORTED + 2];
  int pure;
};

struct QxWorkHost {
  int cls;                 // canonical class index into the kernel table
  int bra_class, ket_class;
  QxWorkDev w;             // device view (prefix/kstart point into ctx buffers)
  long long prefix_off;    // offset of this list's prefix in ctx->d_prefix
};

#define QX_NUM_CLASSES_MAX 21  // canonical classes with l <= 2

struct qx_ctx {
  int device = 0;
  int natom = 0;
  std::vector<double> xyz, Z;
  int nshell = 0;
  std::vector<int> l, atom, nprim, poff, bfc, bfs;
  std::vector<double> exps, coefs;
  int pure = 0, nbf = 0, ncart = 0;
  double tol = 1e-12;
  bool k1 = true;          // every shell has one primitive
  // fragment of each atom: shell pairs and quartets never mix fragments, so
  // one context (and one launch per class) serves a batch of molecules
  std::vector<int> frag;
  bool fragmented = false;
  std::string err;

  // device memory owned by the context
  std::vector<void*> allocs;
  double* d_boys = nullptr;
  QxShellsDev shells{};
  // screened pairs + pseudo pairs (sorted)
  QxPairsDev pr{};
  QxPrimPairsDev pp{};
  int npairs = 0, npseudo = 0;
  std::vector<double> Q;                 // Schwarz value per shell pair (ns*ns)
  std::vector<int> ref_pairs;            // reference-order screened pair list
  // host copies of the sorted screened pairs
  std::vector<int> h_pair_sa, h_pair_sb, h_pair_class, h_pp_first, h_pp_n;
  std::vector<double> h_pair_Q;
  std::vector<int> h_pair_frag;
  std::vector<int> class_pair_off, class_pair_cnt;   // per pair class, in sorted pairs
  std::vector<int> class_pp_off, class_pp_cnt;       // per pair class, in pseudo pairs
  // quartet lists (one per canonical class)
  std::vector<QxWorkHost> work;
  long long* d_prefix = nullptr;
  int* d_kstart = nullptr;               // per bra pseudo pair, fragmented contexts only
  // host copies of the list arrays (quartet -> pair maps for callers that
  // read the stored integrals) and of the pseudo pair -> pair map
  std::vector<long long> h_prefix;
  std::vector<int> h_kstart;             // per bra pseudo pair of every list
  std::vector<long long> h_kstart_off;   // per list: offset into h_kstart
  std::vector<int> h_ppair;
  long long nquart_total = 0;            // canonical contracted quartets
  long long npq_total = 0;               // pseudo-pair quartets
  long long store_total = 0;
  std::vector<long long> nq_class, ndq_class;   // contracted quartets per class (all / not single-atom)

  // optional per-kernel timing with HIP events on the launch stream
  bool profiling = false;
  struct ProfEvent {
    int slot;
    hipEvent_t start, stop;
  };
  std::vector<ProfEvent> prof_pending;
  std::vector<hipEvent_t> event_pool;
  std::vector<double> prof_ms;
  std::vector<long long> prof_launches, prof_pq;

  // independent class kernels fan out over this many streams (the caller's
  // stream plus side streams created on first use, default priority)
  int nstreams = 4;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> fork_ev;     // [0] fork, [1..] joins

  // persistent n_cart^2 work matrices of the J/K and gradient passes (J', K',
  // Cartesian D, symmetrised J and K); ws_ev orders a pass after the previous
  // pass that used them, whatever stream each ran on
  double* ws = nullptr;
  // QX_GRAD_COPIES copies of the 3N gradient accumulators that the blocks of
  // an adjoint-gradient pass spread their flushes over (allocated on first use)
  double* gcopies = nullptr;
  // device address of the file-scope scratch array on this context's device
  // (qx_jk_build_noalloc)
  double* scratch = nullptr;
  hipEvent_t ws_ev = nullptr;
  bool ws_used = false;
  // set while a caller captures a J/K pass into a HIP graph: the pass then
  // records and waits on no event outside the capture (the graph's own
  // stream order sequences its passes)
  bool capturing = false;

  // threads per block of the streaming kernels, per pass (0 J/K, 1 adjoint
  // gradient) and class (the tuned quartet tiles)
  int tile[2][QX_NUM_CLASSES_MAX] = {};
  // per class: the adjoint-gradient kernels sum wave-uniform centres across
  // the wavefront before their LDS accumulators
  unsigned char gwred[QX_NUM_CLASSES_MAX] = {};
  // per class: copies of the gradient accumulators its blocks spread their
  // flushes over (1..QX_GRAD_COPIES)
  int gcopy[QX_NUM_CLASSES_MAX] = {};
  // per class: the J/K kernels stage K' rows in LDS (when the kernel and
  // the context allow it at all)
  unsigned char jkrows[QX_NUM_CLASSES_MAX] = {};
  // widest atom window of the adjoint kernels' LDS gradient accumulators,
  // in doubles (0: QX_GRAD_LDS)
  int grad_win = 0;
  int* d_pair_frag = nullptr;            // fragment of each sorted pair (fragmented contexts)
  // fragmented contexts: first Cartesian function of each fragment id's range
  // and the widest range (the width of the J/K kernels' LDS K' rows)
  int* d_frag_col0 = nullptr;
  int kcols = 0;
  // mixed precision: per class and pass (0 J/K, 1 gradient) whether quartets
  // below the threshold run in FP32 (classes whose FP32 kernels are faster)
  unsigned char fp32_ok[2][QX_NUM_CLASSES_MAX] = {};
  // the quartet lists with each bra row's FP64 ket count at one Schwarz-bound
  // threshold (kets are sorted by descending bound, so the FP64 kets of a row
  // are its first nk64 kets)
  struct Split {
    double thr = -1.0;
    std::vector<QxWorkHost> wm;          // c->work with w.nk64 set
    std::vector<long long> pq64, pq32;   // per list: pseudo-pair quartets above / below thr
    std::vector<long long> cq64, cq32;   // per list: contracted quartets above / below thr
    int* d_nk64 = nullptr;
  } split;
};

// profile slots: array-form ERI kernel of class c -> QX_SLOT_ERI(c, fp32);
// derivative part / group k of class c -> QX_SLOT_DERIV(c, k, fp32); J/K
// part k of class c -> QX_SLOT_JK(c, k, fp32)
#define QX_SLOT_ERI(c, f32) ((f32) * QX_NUM_CLASSES_MAX + (c))
#define QX_SLOT_DERIV(c, k, f32) (2 * QX_NUM_CLASSES_MAX + ((f32) * QX_NUM_CLASSES_MAX + (c)) * QX_DPARTS_CAP + (k))
#define QX_SLOT_JK0 (2 * QX_NUM_CLASSES_MAX * (1 + QX_DPARTS_CAP))
#define QX_SLOT_JK(c, k, f32) (QX_SLOT_JK0 + ((f32) * QX_NUM_CLASSES_MAX + (c)) * QX_JKPARTS_CAP + (k))
#define QX_PROFILE_SLOTS (QX_SLOT_JK0 + 2 * QX_NUM_CLASSES_MAX * QX_JKPARTS_CAP)

// helper kernels (qx_kernels.hip)
extern "C" {
int qxk_fill(double* p, long long n, double v, hipStream_t s);
int qxk_sym_half(const double* in, double* out, int n, hipStream_t s);    // out = (in + in^T)/2
int qxk_add_copies(const double* copies, int ncopy, int n, double* out, hipStream_t s);  // out += sum of copies
int qxk_cart_to_ao_mat(const QxShellsDev* sh, const double* in, double* out, hipStream_t s);   // T X T^T
int qxk_ao_to_cart_mat(const QxShellsDev* sh, const double* in, double* out, hipStream_t s);   // T^T X T
int qxk_cart_to_ao_axis(const QxShellsDev* sh, const double* in, double* out, long long outer, int axis_len_in,
                        int axis_len_out, long long inner, hipStream_t s);
int qxk_oei(const QxShellsDev* sh, double* S, double* T, double* V, hipStream_t s);
int qxk_oei_deriv(const QxShellsDev* sh, const double* Pd, const double* Wd, double* grad, double* dS, double* dT,
                  double* dV, hipStream_t s);
// grad += Pd . d/dtv d(T+V)/dR + Wd . d/dtv dS/dR (Cartesian Pd, Wd)
int qxk_oei_hess_dir(const QxShellsDev* sh, const double* Pd, const double* Wd, const double* tv, double* grad,
                     hipStream_t s);
}
