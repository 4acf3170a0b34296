// Runs the C ABI of libquax_hip on the host emulation (see hip/hip_runtime.h)
// for one molecule read from a text file and writes every result as raw
// doubles -- TEST INFRASTRUCTURE.
//
// input: natom; xyz (3*natom); Z (natom); nshell; per shell: l atom nprim;
//        exps; coefs; pure tol; D (nbf*nbf, AO basis)
// output (binary, in order): [G nbf^4 if nbf <= 60] dG (3N nbf^4 if nbf <= 40)
//        J K (nbf^2 each) grad2e (3N) S T V (nbf^2 each) grad1e (3N)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "quax_amd.h"

static void die(qx_ctx* c, int rc, const char* what) {
  std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, qx_last_error(c));
  std::exit(2);
}

int main(int argc, char** argv) {
  if (argc < 3) return 1;
  FILE* f = std::fopen(argv[1], "r");
  int natom, nshell, pure;
  double tol;
  std::fscanf(f, "%d", &natom);
  std::vector<double> xyz(3 * natom), Z(natom);
  for (auto& v : xyz) std::fscanf(f, "%lf", &v);
  for (auto& v : Z) std::fscanf(f, "%lf", &v);
  std::fscanf(f, "%d", &nshell);
  std::vector<int> l(nshell), at(nshell), np(nshell);
  int tp = 0;
  for (int s = 0; s < nshell; ++s) {
    std::fscanf(f, "%d %d %d", &l[s], &at[s], &np[s]);
    tp += np[s];
  }
  std::vector<double> ex(tp), co(tp);
  for (auto& v : ex) std::fscanf(f, "%lf", &v);
  for (auto& v : co) std::fscanf(f, "%lf", &v);
  std::fscanf(f, "%d %lf", &pure, &tol);
  qx_ctx* c = nullptr;
  // EMU_FRAGMENTS="f0 f1 ...": one fragment id per atom (qx_initialize_fragments)
  std::vector<int> frag;
  if (const char* fs = std::getenv("EMU_FRAGMENTS")) {
    char* end = nullptr;
    for (const char* p = fs; *p;) {
      const long v = std::strtol(p, &end, 10);
      if (end == p) break;
      frag.push_back((int)v);
      p = end;
    }
  }
  int rc = !frag.empty()
               ? qx_initialize_fragments(natom, xyz.data(), Z.data(), nshell, l.data(), at.data(), np.data(),
                                         ex.data(), co.data(), pure, tol, 0, frag.data(), &c)
               : qx_initialize(natom, xyz.data(), Z.data(), nshell, l.data(), at.data(), np.data(), ex.data(), co.data(),
                         pure, tol, 0, &c);
  if (rc) die(nullptr, rc, "initialize");
  int nbf, nc;
  qx_nbf(c, &nbf, &nc);
  std::vector<double> D((size_t)nbf * nbf);
  for (auto& v : D) std::fscanf(f, "%lf", &v);
  std::fclose(f);
  FILE* o = std::fopen(argv[2], "wb");
  const size_t n2 = (size_t)nbf * nbf, n4 = n2 * n2;
  if (nbf <= 60) {
    std::vector<double> G(n4);
    if ((rc = qx_compute_2e_int(c, G.data(), nullptr))) die(c, rc, "compute_2e_int");
    std::fwrite(G.data(), 8, G.size(), o);
  }
  if (nbf <= 40) {
    std::vector<double> dG(3 * natom * n4);
    if ((rc = qx_eri_deriv_core(c, 1, dG.data(), nullptr))) die(c, rc, "eri_deriv_core");
    std::fwrite(dG.data(), 8, dG.size(), o);
  }
  std::vector<double> J(n2), K(n2), g(3 * natom, 0.0), S(n2), T(n2), V(n2), g1(3 * natom, 0.0);
  if ((rc = qx_jk_build(c, D.data(), J.data(), K.data(), 0, 0.0, nullptr))) die(c, rc, "jk_build");
  if ((rc = qx_eri_grad(c, D.data(), g.data(), 0, 0.0, nullptr))) die(c, rc, "eri_grad");
  if ((rc = qx_oei(c, S.data(), T.data(), V.data(), nullptr))) die(c, rc, "oei");
  if ((rc = qx_oei_grad(c, D.data(), D.data(), g1.data(), nullptr))) die(c, rc, "oei_grad");
  for (auto* v : {&J, &K, &g, &S, &T, &V, &g1}) std::fwrite(v->data(), 8, v->size(), o);
  // mixed precision, FP32 admitted in every class: every quartet in FP32
  // (threshold above any bound), then a split at 0.05 (both masked launches
  // over each list)
  std::vector<int> all32(qx_num_classes(), 1);
  for (int pass = 0; pass < 2; ++pass)
    if ((rc = qx_set_fp32_classes(c, pass, all32.data()))) die(c, rc, "set_fp32_classes");
  for (double thr : {1e300, 0.05}) {
    std::vector<double> Jm(n2), Km(n2), gm(3 * natom, 0.0);
    if ((rc = qx_jk_build(c, D.data(), Jm.data(), Km.data(), 1, thr, nullptr))) die(c, rc, "jk_build mixed");
    if ((rc = qx_eri_grad(c, D.data(), gm.data(), 1, thr, nullptr))) die(c, rc, "eri_grad mixed");
    for (auto* v : {&Jm, &Km, &gm}) std::fwrite(v->data(), 8, v->size(), o);
  }
  std::fclose(o);
  qx_finalize(c);
  std::fprintf(stderr, "emu: ok\n");
  return 0;
}
