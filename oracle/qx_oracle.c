/*
 * qx_oracle.c -- CPU FP64 restatement of Quax's two-electron integral path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  Nothing in quax_amd/ imports, links or calls it.
 *
 * What it restates (file:line in the Quax reference):
 *   - Shell normalisation as Libint applies it to every shell Quax builds
 *     (libint_interface.cc:48-70 -> libint2::Shell::renorm): primitives get the
 *     x^l normalisation, then the contraction is rescaled to unit self-overlap;
 *     all Cartesian components share that factor.
 *   - schwarz_screening (libint_interface.cc:158-229): Q_ab = max_f |(f1f2|f1f2)|,
 *     pair kept when Q_ab >= tol^2 / max(Q).
 *   - compute_2e_int("eri") (libint_interface.cc:486-647): bra pairs x ket
 *     pairs, skip when Q_ab*Q_cd < tol^2, scatter into 4/2/1 permutations.
 *   - compute_2e_deriv("eri", deriv_vec) (libint_interface.cc:1097-1335) and
 *     eri_deriv_core(1 and 2) (libint_interface.cc:2582-2779): skip
 *     single-atom quartets, sum the shell-centre derivative buffers of the
 *     centres that sit on the differentiated atom(s) -- at order 2 every
 *     (centre, centre) combination, looked up among the 78 buffers.
 *   - The arithmetic Libint supplies (the un-vendored libint2 Engine, any
 *     version: libint_interface.cc:501, 561, 2600, 2648): contracted Cartesian
 *     Coulomb integrals by the Head-Gordon-Pople scheme -- Obara-Saika
 *     vertical recurrence on primitive quartets, contraction, then horizontal
 *     recurrence -- with derivatives from d/dA_i phi_a = 2 alpha phi_{a+1_i} -
 *     a_i phi_{a-1_i} (applied twice for second derivatives) evaluated for all
 *     four centres explicitly (no translational-invariance shortcut, so
 *     invariance is a test, not an input).
 *   - One-electron S, T, V and their first nuclear derivatives (Libint
 *     Operator::overlap/kinetic/nuclear as used by oei_deriv_core,
 *     libint_interface.cc:2403-2579), needed for RHF energies and gradients.
 *
 * Boys function: convergent series in long double for T <= 30 followed by
 * downward recursion; erf + upward recursion above.  Written for clarity, not
 * speed: generic loops over component tables, no code generation.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_LMAX 4                        /* largest shell l accepted */
#define OR_LT (2 * (OR_LMAX + 1))        /* largest bra (or ket) sum incl. derivative shift */
#define OR_MMAX (2 * OR_LT + 1)

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }
static int coff(int l) { return l * (l + 1) * (l + 2) / 6; }
static int cidx(int ly, int lz) { int i = ly + lz; return i * (i + 1) / 2 + lz; }
static int gidx(int lx, int ly, int lz) { return coff(lx + ly + lz) + cidx(ly, lz); }

/* component table for every Cartesian triple with total <= OR_LT */
static int g_comp[/*coff(OR_LT+1)*/ 400][3];
static int g_ready = 0;

static void init_tables(void) {
    if (g_ready) return;
    for (int l = 0; l <= OR_LT; ++l) {
        for (int i = 0; i <= l; ++i) {
            int lx = l - i;
            for (int j = 0; j <= i; ++j) {
                int ly = i - j, lz = j;
                int g = gidx(lx, ly, lz);
                g_comp[g][0] = lx; g_comp[g][1] = ly; g_comp[g][2] = lz;
            }
        }
    }
    g_ready = 1;
}

static int g_shift(int g, int axis, int d) {
    int c[3] = {g_comp[g][0], g_comp[g][1], g_comp[g][2]};
    c[axis] += d;
    if (c[axis] < 0) return -1;
    return gidx(c[0], c[1], c[2]);
}

static int first_axis(int g) {
    if (g_comp[g][0] > 0) return 0;
    if (g_comp[g][1] > 0) return 1;
    return 2;
}

/* ------------------------------------------------------------------ Boys */
void or_boys(int mmax, double T, double* F) {
    long double t = T;
    if (T <= 30.0) {
        long double et = expl(-t);
        long double term = 1.0L / (2 * mmax + 1), sum = term;
        for (int k = 1; k < 2000; ++k) {
            term *= 2.0L * t / (long double)(2 * mmax + 2 * k + 1);
            sum += term;
            if (term < 1e-24L * sum) break;
        }
        long double f = et * sum;
        F[mmax] = (double)f;
        for (int m = mmax - 1; m >= 0; --m) {
            f = (2.0L * t * f + et) / (long double)(2 * m + 1);
            F[m] = (double)f;
        }
    } else {
        long double et = expl(-t);
        long double f = 0.5L * sqrtl(3.14159265358979323846264338327950288L / t) * erfl(sqrtl(t));
        F[0] = (double)f;
        for (int m = 0; m < mmax; ++m) {
            f = ((2 * m + 1) * f - et) / (2.0L * t);
            F[m + 1] = (double)f;
        }
    }
}

/* ------------------------------------------------------------------ basis */
typedef struct {
    int natom;
    double* xyz;      /* 3*natom, bohr */
    double* Zn;       /* nuclear charges */
    int nshell;
    int* l; int* atom; int* nprim; int* poff; int* bf;
    double* alpha; double* coef;   /* coef: normalised (Libint renorm convention) */
    int nbf;
} or_basis;

static double dfact(int n) { /* (n)!! with (-1)!! = 1 */
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

void or_normalize(int l, int nprim, const double* alpha, const double* cin, double* cout) {
    const double pi32 = pow(M_PI, 1.5);
    for (int p = 0; p < nprim; ++p) {
        double ta = 2.0 * alpha[p];
        double n = sqrt(pow(2.0, l) * pow(ta, l + 1.5) / (pi32 * dfact(2 * l - 1)));
        cout[p] = cin[p] * n;
    }
    double norm = 0.0;
    for (int p = 0; p < nprim; ++p)
        for (int q = 0; q < nprim; ++q) {
            double g = alpha[p] + alpha[q];
            norm += dfact(2 * l - 1) * pi32 * cout[p] * cout[q] / (pow(2.0, l) * pow(g, l + 1.5));
        }
    double s = 1.0 / sqrt(norm);
    for (int p = 0; p < nprim; ++p) cout[p] *= s;
}

or_basis* or_basis_create(int natom, const double* xyz, const double* Zn, int nshell,
                          const int* l, const int* atom, const int* nprim,
                          const double* alpha, const double* coef_raw) {
    init_tables();
    or_basis* b = (or_basis*)calloc(1, sizeof(or_basis));
    b->natom = natom;
    b->xyz = (double*)malloc(sizeof(double) * 3 * (natom > 0 ? natom : 1));
    memcpy(b->xyz, xyz, sizeof(double) * 3 * natom);
    b->Zn = (double*)malloc(sizeof(double) * (natom > 0 ? natom : 1));
    memcpy(b->Zn, Zn, sizeof(double) * natom);
    b->nshell = nshell;
    b->l = (int*)malloc(sizeof(int) * nshell);
    b->atom = (int*)malloc(sizeof(int) * nshell);
    b->nprim = (int*)malloc(sizeof(int) * nshell);
    b->poff = (int*)malloc(sizeof(int) * (nshell + 1));
    b->bf = (int*)malloc(sizeof(int) * (nshell + 1));
    int tp = 0, nb = 0;
    for (int s = 0; s < nshell; ++s) {
        b->l[s] = l[s]; b->atom[s] = atom[s]; b->nprim[s] = nprim[s];
        b->poff[s] = tp; tp += nprim[s];
        b->bf[s] = nb; nb += ncart(l[s]);
    }
    b->poff[nshell] = tp; b->bf[nshell] = nb; b->nbf = nb;
    b->alpha = (double*)malloc(sizeof(double) * (tp > 0 ? tp : 1));
    b->coef = (double*)malloc(sizeof(double) * (tp > 0 ? tp : 1));
    memcpy(b->alpha, alpha, sizeof(double) * tp);
    for (int s = 0; s < nshell; ++s)
        or_normalize(l[s], nprim[s], alpha + b->poff[s], coef_raw + b->poff[s], b->coef + b->poff[s]);
    return b;
}

void or_basis_destroy(or_basis* b) {
    if (!b) return;
    free(b->xyz); free(b->Zn); free(b->l); free(b->atom); free(b->nprim); free(b->poff);
    free(b->bf); free(b->alpha); free(b->coef); free(b);
}

int or_nbf(const or_basis* b) { return b->nbf; }

void or_get_coefs(const or_basis* b, double* out) {
    memcpy(out, b->coef, sizeof(double) * b->poff[b->nshell]);
}

static const double* ctr(const or_basis* b, int s) { return b->xyz + 3 * b->atom[s]; }

/* --------------------------------------------------------- two-electron */
/* Obara-Saika vertical recurrence for one primitive quartet.  Fills
 * X[e][f] = [e0|f0]^(0) for every |e| <= Le, |f| <= Lf (component tables in
 * global CCA order), including all prefactors. */
static void vrr_prim(int Le, int Lf, double a, const double* A, double bb, const double* B,
                     double c, const double* C, double d, const double* D, double* V, double* X) {
    const int NE = coff(Le + 1), NF = coff(Lf + 1), Lt = Le + Lf;
    double p = a + bb, q = c + d, pq = p + q;
    double P[3], Q[3], W[3], PA[3], WP[3], QC[3], WQ[3];
    double AB2 = 0, CD2 = 0, PQ2 = 0;
    for (int i = 0; i < 3; ++i) {
        P[i] = (a * A[i] + bb * B[i]) / p;
        Q[i] = (c * C[i] + d * D[i]) / q;
        W[i] = (p * P[i] + q * Q[i]) / pq;
        PA[i] = P[i] - A[i]; WP[i] = W[i] - P[i];
        QC[i] = Q[i] - C[i]; WQ[i] = W[i] - Q[i];
        AB2 += (A[i] - B[i]) * (A[i] - B[i]);
        CD2 += (C[i] - D[i]) * (C[i] - D[i]);
        PQ2 += (P[i] - Q[i]) * (P[i] - Q[i]);
    }
    double rho = p * q / pq;
    double pref = 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(pq)) * exp(-a * bb / p * AB2 - c * d / q * CD2);
    double F[OR_MMAX + 1];
    or_boys(Lt, rho * PQ2, F);
#define VI(m, e, f) V[((size_t)(m) * NE + (e)) * NF + (f)]
    for (int m = 0; m <= Lt; ++m) VI(m, 0, 0) = pref * F[m];
    for (int L = 1; L <= Le; ++L) {
        for (int g = coff(L); g < coff(L + 1); ++g) {
            int i = first_axis(g);
            int e1 = g_shift(g, i, -1), e2 = g_shift(g, i, -2);
            int n1 = g_comp[g][i] - 1;
            for (int m = 0; m <= Lt - L; ++m) {
                double v = PA[i] * VI(m, e1, 0) + WP[i] * VI(m + 1, e1, 0);
                if (n1 > 0) v += n1 / (2.0 * p) * (VI(m, e2, 0) - rho / p * VI(m + 1, e2, 0));
                VI(m, g, 0) = v;
            }
        }
    }
    for (int Lk = 1; Lk <= Lf; ++Lk) {
        for (int gf = coff(Lk); gf < coff(Lk + 1); ++gf) {
            int i = first_axis(gf);
            int f1 = g_shift(gf, i, -1), f2 = g_shift(gf, i, -2);
            int n1 = g_comp[gf][i] - 1;
            for (int ge = 0; ge < NE; ++ge) {
                int le = g_comp[ge][0] + g_comp[ge][1] + g_comp[ge][2];
                int ei = g_comp[ge][i];
                int em = ei > 0 ? g_shift(ge, i, -1) : -1;
                for (int m = 0; m <= Lt - le - Lk; ++m) {
                    double v = QC[i] * VI(m, ge, f1) + WQ[i] * VI(m + 1, ge, f1);
                    if (n1 > 0) v += n1 / (2.0 * q) * (VI(m, ge, f2) - rho / q * VI(m + 1, ge, f2));
                    if (ei > 0) v += ei / (2.0 * pq) * VI(m + 1, em, f1);
                    VI(m, ge, gf) = v;
                }
            }
        }
    }
    for (int e = 0; e < NE; ++e)
        for (int f = 0; f < NF; ++f) X[e * NF + f] = VI(0, e, f);
#undef VI
}

/* Horizontal recurrence on one index pair: input x[g] for every component g
 * with lo <= |g| <= lo+hi (global order), output out[i*ncart(hi)+j] =
 * (lo_i, hi_j) built by (a, b+1_k) = (a+1_k, b) + AB_k (a, b). */
static void hrr1(const double* x, int lo, int hi, const double* AB, double* out, double* Z) {
    const int NA = coff(lo + hi + 1), NB = coff(hi + 1);
    for (int g = coff(lo); g < NA; ++g) Z[g * NB + 0] = x[g];
    for (int bl = 1; bl <= hi; ++bl) {
        for (int gb = coff(bl); gb < coff(bl + 1); ++gb) {
            int k = first_axis(gb);
            int b1 = g_shift(gb, k, -1);
            for (int ga = coff(lo); ga < coff(lo + hi - bl + 1); ++ga) {
                int ap = g_shift(ga, k, +1);
                Z[ga * NB + gb] = Z[ap * NB + b1] + AB[k] * Z[ga * NB + b1];
            }
        }
    }
    const int na = ncart(lo), nb = ncart(hi);
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) out[i * nb + j] = Z[(coff(lo) + i) * NB + coff(hi) + j];
}

/* (ab|cd) from a contracted table X[NE][NF] via ket HRR then bra HRR */
static void hrr_quartet(const double* X, int NE, int NF, int la, int lb, int lc, int ld,
                        const double* AB, const double* CD, double* out) {
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    const int ncd = nc * nd;
    int ZN = coff(OR_LT + 1);
    double* Z = (double*)malloc(sizeof(double) * ZN * coff(OR_LMAX + 2));
    double* T = (double*)calloc((size_t)NE * ncd, sizeof(double));
    double* xf = (double*)malloc(sizeof(double) * (NF > NE ? NF : NE));
    double* tmp = (double*)malloc(sizeof(double) * ncart(OR_LT) * ncart(OR_LT));
    for (int ge = coff(la); ge < coff(la + lb + 1); ++ge) {
        for (int f = 0; f < NF; ++f) xf[f] = X[ge * NF + f];
        hrr1(xf, lc, ld, CD, tmp, Z);
        for (int k = 0; k < ncd; ++k) T[ge * ncd + k] = tmp[k];
    }
    for (int k = 0; k < ncd; ++k) {
        for (int ge = 0; ge < NE; ++ge) xf[ge] = (ge >= coff(la) && ge < coff(la + lb + 1)) ? T[ge * ncd + k] : 0.0;
        hrr1(xf, la, lb, AB, tmp, Z);
        for (int ab = 0; ab < na * nb; ++ab) out[ab * ncd + k] = tmp[ab];
    }
    free(Z); free(T); free(xf); free(tmp);
}

/* Contracted Cartesian shell quartet (ab|cd); out is na*nb*nc*nd row-major. */
void or_shell_quartet(const or_basis* b, int sa, int sb, int sc, int sd, double* out) {
    init_tables();
    int la = b->l[sa], lb = b->l[sb], lc = b->l[sc], ld = b->l[sd];
    int Le = la + lb, Lf = lc + ld;
    const int NE = coff(Le + 1), NF = coff(Lf + 1);
    const double *A = ctr(b, sa), *B = ctr(b, sb), *C = ctr(b, sc), *D = ctr(b, sd);
    double* V = (double*)malloc(sizeof(double) * (Le + Lf + 1) * NE * NF);
    double* X = (double*)malloc(sizeof(double) * NE * NF);
    double* Xc = (double*)calloc((size_t)NE * NF, sizeof(double));
    for (int i = b->poff[sa]; i < b->poff[sa + 1]; ++i)
        for (int j = b->poff[sb]; j < b->poff[sb + 1]; ++j)
            for (int k = b->poff[sc]; k < b->poff[sc + 1]; ++k)
                for (int m = b->poff[sd]; m < b->poff[sd + 1]; ++m) {
                    double w = b->coef[i] * b->coef[j] * b->coef[k] * b->coef[m];
                    vrr_prim(Le, Lf, b->alpha[i], A, b->alpha[j], B, b->alpha[k], C, b->alpha[m], D, V, X);
                    for (int t = 0; t < NE * NF; ++t) Xc[t] += w * X[t];
                }
    double AB[3], CD[3];
    for (int i = 0; i < 3; ++i) { AB[i] = A[i] - B[i]; CD[i] = C[i] - D[i]; }
    hrr_quartet(Xc, NE, NF, la, lb, lc, ld, AB, CD, out);
    free(V); free(X); free(Xc);
}

/* First derivatives of a contracted shell quartet with respect to the 12
 * centre coordinates.  out holds 12 blocks of na*nb*nc*nd, ordered centre-major
 * then x, y, z (Libint's deriv-1 buffer order, libint_interface.cc:2587). */
void or_shell_quartet_deriv1(const or_basis* b, int sa, int sb, int sc, int sd, double* out) {
    init_tables();
    int la = b->l[sa], lb = b->l[sb], lc = b->l[sc], ld = b->l[sd];
    int Le = la + lb + 1, Lf = lc + ld + 1;
    const int NE = coff(Le + 1), NF = coff(Lf + 1);
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    const int nblk = na * nb * nc * nd;
    const double *A = ctr(b, sa), *B = ctr(b, sb), *C = ctr(b, sc), *D = ctr(b, sd);
    double* V = (double*)malloc(sizeof(double) * (Le + Lf + 1) * NE * NF);
    double* X = (double*)malloc(sizeof(double) * NE * NF);
    double* Xw[5];
    for (int k = 0; k < 5; ++k) Xw[k] = (double*)calloc((size_t)NE * NF, sizeof(double));
    for (int i = b->poff[sa]; i < b->poff[sa + 1]; ++i)
        for (int j = b->poff[sb]; j < b->poff[sb + 1]; ++j)
            for (int k = b->poff[sc]; k < b->poff[sc + 1]; ++k)
                for (int m = b->poff[sd]; m < b->poff[sd + 1]; ++m) {
                    double w = b->coef[i] * b->coef[j] * b->coef[k] * b->coef[m];
                    double wt[5] = {w, 2 * b->alpha[i] * w, 2 * b->alpha[j] * w, 2 * b->alpha[k] * w, 2 * b->alpha[m] * w};
                    vrr_prim(Le, Lf, b->alpha[i], A, b->alpha[j], B, b->alpha[k], C, b->alpha[m], D, V, X);
                    for (int s = 0; s < 5; ++s)
                        for (int t = 0; t < NE * NF; ++t) Xw[s][t] += wt[s] * X[t];
                }
    double AB[3], CD[3];
    for (int i = 0; i < 3; ++i) { AB[i] = A[i] - B[i]; CD[i] = C[i] - D[i]; }
    int lmax = OR_LMAX + 1;
    size_t bigblk = (size_t)ncart(lmax) * ncart(lmax) * ncart(lmax) * ncart(lmax);
    double* up = (double*)malloc(sizeof(double) * bigblk);
    double* dn = (double*)malloc(sizeof(double) * bigblk);
    memset(out, 0, sizeof(double) * 12 * nblk);
    int ls[4] = {la, lb, lc, ld};
    for (int ctrI = 0; ctrI < 4; ++ctrI) {
        int l2[4] = {la, lb, lc, ld};
        l2[ctrI] += 1;
        hrr_quartet(Xw[1 + ctrI], NE, NF, l2[0], l2[1], l2[2], l2[3], AB, CD, up);
        int n2[4] = {ncart(l2[0]), ncart(l2[1]), ncart(l2[2]), ncart(l2[3])};
        int have_dn = ls[ctrI] > 0;
        int n3[4] = {na, nb, nc, nd};
        if (have_dn) {
            int l3[4] = {la, lb, lc, ld};
            l3[ctrI] -= 1;
            hrr_quartet(Xw[0], NE, NF, l3[0], l3[1], l3[2], l3[3], AB, CD, dn);
            for (int t = 0; t < 4; ++t) n3[t] = ncart(l3[t]);
        }
        for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib)
        for (int ic = 0; ic < nc; ++ic)
        for (int id = 0; id < nd; ++id) {
            int gl[4] = {coff(la) + ia, coff(lb) + ib, coff(lc) + ic, coff(ld) + id};
            int e = ((ia * nb + ib) * nc + ic) * nd + id;
            for (int x = 0; x < 3; ++x) {
                int gu[4] = {gl[0], gl[1], gl[2], gl[3]};
                gu[ctrI] = g_shift(gl[ctrI], x, +1);
                int iu[4];
                int lu[4] = {la, lb, lc, ld};
                lu[ctrI] += 1;
                for (int t = 0; t < 4; ++t) iu[t] = gu[t] - coff(lu[t]);
                double v = up[((iu[0] * n2[1] + iu[1]) * n2[2] + iu[2]) * n2[3] + iu[3]];
                int q = g_comp[gl[ctrI]][x];
                if (q > 0) {
                    int gd[4] = {gl[0], gl[1], gl[2], gl[3]};
                    gd[ctrI] = g_shift(gl[ctrI], x, -1);
                    int ld4[4] = {la, lb, lc, ld};
                    ld4[ctrI] -= 1;
                    int idn[4];
                    for (int t = 0; t < 4; ++t) idn[t] = gd[t] - coff(ld4[t]);
                    v -= q * dn[((idn[0] * n3[1] + idn[1]) * n3[2] + idn[2]) * n3[3] + idn[3]];
                }
                out[(3 * ctrI + x) * nblk + e] = v;
            }
        }
    }
    free(up); free(dn); free(V); free(X);
    for (int k = 0; k < 5; ++k) free(Xw[k]);
}

/* schwarz_screening (libint_interface.cc:158-229).  vals: nshell*nshell
 * (symmetric), pairs: list of (s1, s2) with s2 <= s1 kept by the tol^2/max
 * filter, in the reference's loop order.  Returns the number of pairs. */
int or_schwarz(const or_basis* b, double tol, double* vals, int* pairs) {
    int ns = b->nshell;
    double maxint = 0.0;
    memset(vals, 0, sizeof(double) * ns * ns);
    for (int s1 = 0; s1 < ns; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2) {
            int n1 = ncart(b->l[s1]), n2 = ncart(b->l[s2]);
            double* buf = (double*)malloc(sizeof(double) * n1 * n2 * n1 * n2);
            or_shell_quartet(b, s1, s2, s1, s2, buf);
            double mx = 0.0;
            for (int f1 = 0; f1 < n1; ++f1)
                for (int f2 = 0; f2 < n2; ++f2) {
                    double v = fabs(buf[f1 * (n1 * n2 * n2 + n2) + f2 * (n1 * n2 + 1)]);
                    if (v > mx) mx = v;
                }
            free(buf);
            if (mx > maxint) maxint = mx;
            vals[s1 * ns + s2] = vals[s2 * ns + s1] = mx;
        }
    double cut = tol * tol / maxint;
    int np = 0;
    for (int s1 = 0; s1 < ns; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2)
            if (vals[s1 * ns + s2] >= cut) { pairs[2 * np] = s1; pairs[2 * np + 1] = s2; ++np; }
    return np;
}

/* compute_2e_int("eri") (libint_interface.cc:486-647). out: nbf^4, zeroed here. */
int or_compute_2e_int(const or_basis* b, double tol, double* out) {
    int ns = b->nshell;
    size_t n = b->nbf;
    double* vals = (double*)malloc(sizeof(double) * ns * ns);
    int* pairs = (int*)malloc(sizeof(int) * ns * (ns + 1));
    int np = or_schwarz(b, tol, vals, pairs);
    double tol2 = tol * tol;
    memset(out, 0, sizeof(double) * n * n * n * n);
    #pragma omp parallel for schedule(dynamic)
    for (int ib = 0; ib < np; ++ib) {
        int p1 = pairs[2 * ib], p2 = pairs[2 * ib + 1];
        int n1 = ncart(b->l[p1]), n2 = ncart(b->l[p2]);
        size_t bf1 = b->bf[p1], bf2 = b->bf[p2];
        double* buf = (double*)malloc(sizeof(double) * 50625);
        for (int ik = 0; ik < np; ++ik) {
            int p3 = pairs[2 * ik], p4 = pairs[2 * ik + 1];
            if (vals[p1 * ns + p2] * vals[p3 * ns + p4] < tol2) continue;
            int n3 = ncart(b->l[p3]), n4 = ncart(b->l[p4]);
            size_t bf3 = b->bf[p3], bf4 = b->bf[p4];
            or_shell_quartet(b, p1, p2, p3, p4, buf);
            for (int f1 = 0, idx = 0; f1 < n1; ++f1)
                for (int f2 = 0; f2 < n2; ++f2)
                    for (int f3 = 0; f3 < n3; ++f3)
                        for (int f4 = 0; f4 < n4; ++f4, ++idx) {
                            size_t i = bf1 + f1, j = bf2 + f2, k = bf3 + f3, l = bf4 + f4;
                            double v = buf[idx];
                            out[((i * n + j) * n + k) * n + l] = v;
                            if (p1 != p2) out[((j * n + i) * n + k) * n + l] = v;
                            if (p3 != p4) out[((i * n + j) * n + l) * n + k] = v;
                            if (p1 != p2 && p3 != p4) out[((j * n + i) * n + l) * n + k] = v;
                        }
        }
        free(buf);
    }
    free(vals); free(pairs);
    return 0;
}

/* Shared body of eri_deriv_core(1) and compute_2e_deriv (order 1).
 * want_coord < 0: all 3N slices into out[(3N) * n^4]; otherwise only the
 * slice for that coordinate into out[n^4]. */
static int deriv1_common(const or_basis* b, double tol, int want_coord, double* out) {
    int ns = b->nshell;
    size_t n = b->nbf;
    size_t n4 = n * n * n * n;
    int ncoord = 3 * b->natom;
    double* vals = (double*)malloc(sizeof(double) * ns * ns);
    int* pairs = (int*)malloc(sizeof(int) * ns * (ns + 1));
    int np = or_schwarz(b, tol, vals, pairs);
    double tol2 = tol * tol;
    memset(out, 0, sizeof(double) * (want_coord < 0 ? ncoord : 1) * n4);
    /* parallel over bra pairs; different bra pairs write disjoint (i, j) rows
     * and their transposes, except when two bra pairs are transposes of each
     * other -- impossible because pairs are s1 >= s2 only. */
    #pragma omp parallel for schedule(dynamic)
    for (int ib = 0; ib < np; ++ib) {
        int p1 = pairs[2 * ib], p2 = pairs[2 * ib + 1];
        int n1 = ncart(b->l[p1]), n2 = ncart(b->l[p2]);
        size_t bf1 = b->bf[p1], bf2 = b->bf[p2];
        double* buf = (double*)malloc(sizeof(double) * 12 * 50625);
        for (int ik = 0; ik < np; ++ik) {
            int p3 = pairs[2 * ik], p4 = pairs[2 * ik + 1];
            if (vals[p1 * ns + p2] * vals[p3 * ns + p4] < tol2) continue;
            int at[4] = {b->atom[p1], b->atom[p2], b->atom[p3], b->atom[p4]};
            if (at[0] == at[1] && at[0] == at[2] && at[0] == at[3]) continue;
            if (want_coord >= 0) {
                int wa = want_coord / 3;
                if (at[0] != wa && at[1] != wa && at[2] != wa && at[3] != wa) continue;
            }
            int n3 = ncart(b->l[p3]), n4s = ncart(b->l[p4]);
            size_t bf3 = b->bf[p3], bf4 = b->bf[p4];
            int nblk = n1 * n2 * n3 * n4s;
            or_shell_quartet_deriv1(b, p1, p2, p3, p4, buf);
            for (int coord = 0; coord < ncoord; ++coord) {
                if (want_coord >= 0 && coord != want_coord) continue;
                int A = coord / 3, x = coord % 3;
                size_t off = want_coord >= 0 ? 0 : (size_t)coord * n4;
                for (int c = 0; c < 4; ++c) {
                    if (at[c] != A) continue;
                    const double* blk = buf + (3 * c + x) * nblk;
                    for (int f1 = 0, idx = 0; f1 < n1; ++f1)
                        for (int f2 = 0; f2 < n2; ++f2)
                            for (int f3 = 0; f3 < n3; ++f3)
                                for (int f4 = 0; f4 < n4s; ++f4, ++idx) {
                                    size_t i = bf1 + f1, j = bf2 + f2, k = bf3 + f3, l = bf4 + f4;
                                    double* e0 = &out[off + ((i * n + j) * n + k) * n + l];
                                    double v = *e0 + blk[idx];
                                    *e0 = v;
                                    if (p1 != p2) out[off + ((j * n + i) * n + k) * n + l] = v;
                                    if (p3 != p4) out[off + ((i * n + j) * n + l) * n + k] = v;
                                    if (p1 != p2 && p3 != p4) out[off + ((j * n + i) * n + l) * n + k] = v;
                                }
                }
            }
        }
        free(buf);
    }
    free(vals); free(pairs);
    return 0;
}

int or_eri_deriv_core1(const or_basis* b, double tol, double* out) {
    return deriv1_common(b, tol, -1, out);
}

/* compute_2e_deriv("eri", 0., deriv_vec) for a one-hot deriv_vec. Returns -1
 * for any other order. */
int or_compute_2e_deriv1(const or_basis* b, double tol, const int* deriv_vec, double* out) {
    int ncoord = 3 * b->natom, order = 0, coord = -1;
    for (int i = 0; i < ncoord; ++i) { order += deriv_vec[i]; if (deriv_vec[i]) coord = i; }
    if (order != 1) return -1;
    return deriv1_common(b, tol, coord, out);
}

/* ---------------------------------------------------- second derivatives */
/* d/dA_x phi_n = 2 alpha phi_{n+1_x} - n_x phi_{n-1_x} applied twice, per
 * primitive quartet: each of the up-to-four terms of a second derivative is
 * a primitive integral of a class with shifted angular momenta, taken from
 * the HRR of one VRR table that covers every shift (bra and ket sums raised
 * by 2).  Written for clarity: the shifted classes are evaluated per
 * primitive quartet and cached by shift. */
typedef struct { double c; int g[4]; } or_term;

/* expand d/dX_x (centre X, axis x) of the product of 4 functions given by
 * component indices g[4] into at most 2 terms */
static int or_apply(const or_term* t, int X, int x, const double* ex2, or_term* out) {
    int k = 0;
    int gu = g_shift(t->g[X], x, +1);
    out[k] = *t; out[k].c = t->c * ex2[X]; out[k].g[X] = gu; ++k;
    int nx = g_comp[t->g[X]][x];
    if (nx > 0) {
        out[k] = *t; out[k].c = -t->c * nx; out[k].g[X] = g_shift(t->g[X], x, -1); ++k;
    }
    return k;
}

static int ltot(int g) { return g_comp[g][0] + g_comp[g][1] + g_comp[g][2]; }

/* Second derivatives of a contracted shell quartet with respect to pairs of
 * the 12 centre coordinates (centre-major, then x, y, z).  out holds 78
 * blocks of na*nb*nc*nd, one per pair i <= j in lexicographic order -- the
 * order of generate_multi_index_lookup(12, 2) that indexes Libint's deriv-2
 * buffers (libint_interface.cc:2587). */
void or_shell_quartet_deriv2(const or_basis* b, int sa, int sb, int sc, int sd, double* out) {
    init_tables();
    int ls[4] = {b->l[sa], b->l[sb], b->l[sc], b->l[sd]};
    int sh[4] = {sa, sb, sc, sd};
    int Le = ls[0] + ls[1] + 2, Lf = ls[2] + ls[3] + 2;
    const int NE = coff(Le + 1), NF = coff(Lf + 1);
    const int nn[4] = {ncart(ls[0]), ncart(ls[1]), ncart(ls[2]), ncart(ls[3])};
    const int nblk = nn[0] * nn[1] * nn[2] * nn[3];
    const double* Cn[4] = {ctr(b, sa), ctr(b, sb), ctr(b, sc), ctr(b, sd)};
    double* V = (double*)malloc(sizeof(double) * (Le + Lf + 1) * NE * NF);
    double* X = (double*)malloc(sizeof(double) * NE * NF);
    /* cache of shifted-class blocks: shift d in -2..2 per centre -> 625 keys */
    double* cache[625];
    memset(cache, 0, sizeof(cache));
    double AB[3], CD[3];
    for (int i = 0; i < 3; ++i) { AB[i] = Cn[0][i] - Cn[1][i]; CD[i] = Cn[2][i] - Cn[3][i]; }
    memset(out, 0, sizeof(double) * 78 * nblk);
    for (int i = b->poff[sa]; i < b->poff[sa + 1]; ++i)
    for (int j = b->poff[sb]; j < b->poff[sb + 1]; ++j)
    for (int k = b->poff[sc]; k < b->poff[sc + 1]; ++k)
    for (int m = b->poff[sd]; m < b->poff[sd + 1]; ++m) {
        const double w = b->coef[i] * b->coef[j] * b->coef[k] * b->coef[m];
        const double ex2[4] = {2 * b->alpha[i], 2 * b->alpha[j], 2 * b->alpha[k], 2 * b->alpha[m]};
        vrr_prim(Le, Lf, b->alpha[i], Cn[0], b->alpha[j], Cn[1], b->alpha[k], Cn[2], b->alpha[m], Cn[3], V, X);
        for (int q = 0; q < 625; ++q) { free(cache[q]); cache[q] = NULL; }
        int pair = 0;
        for (int I = 0; I < 12; ++I)
        for (int J = I; J < 12; ++J, ++pair) {
            const int cI = I / 3, xI = I % 3, cJ = J / 3, xJ = J % 3;
            for (int e = 0; e < nblk; ++e) {
                int rem = e, idx[4];
                for (int t = 3; t >= 0; --t) { idx[t] = rem % nn[t]; rem /= nn[t]; }
                or_term t0;
                t0.c = 1.0;
                for (int t = 0; t < 4; ++t) t0.g[t] = coff(ls[t]) + idx[t];
                or_term t1[2], t2[2];
                int k1 = or_apply(&t0, cI, xI, ex2, t1);
                double acc = 0.0;
                for (int u = 0; u < k1; ++u) {
                    int k2 = or_apply(&t1[u], cJ, xJ, ex2, t2);
                    for (int v = 0; v < k2; ++v) {
                        int l2[4], key = 0;
                        for (int t = 0; t < 4; ++t) { l2[t] = ltot(t2[v].g[t]); key = key * 5 + (l2[t] - ls[t] + 2); }
                        if (!cache[key]) {
                            size_t nb2 = (size_t)ncart(l2[0]) * ncart(l2[1]) * ncart(l2[2]) * ncart(l2[3]);
                            cache[key] = (double*)malloc(sizeof(double) * nb2);
                            hrr_quartet(X, NE, NF, l2[0], l2[1], l2[2], l2[3], AB, CD, cache[key]);
                        }
                        int o = 0;
                        for (int t = 0; t < 4; ++t) o = o * ncart(l2[t]) + (t2[v].g[t] - coff(l2[t]));
                        acc += t2[v].c * cache[key][o];
                    }
                }
                out[(size_t)pair * nblk + e] += w * acc;
            }
        }
    }
    for (int q = 0; q < 625; ++q) free(cache[q]);
    free(V); free(X);
    (void)sh;
}

/* Shared body of eri_deriv_core(2) and compute_2e_deriv at order 2, as
 * libint_interface.cc:2582-2779 (and 1097-1335) do it: for each unique pair
 * (x <= y) of nuclear coordinates, the buffers of every (centre on atom(x),
 * centre on atom(y)) combination -- sorted, looked up in the 78-buffer
 * table -- are summed and scattered into the 4/2/1 permutations; quartets
 * with all four centres on one atom are skipped.  want < 0: every pair,
 * out[(3N(3N+1)/2) * n^4]; otherwise only that pair index, out[n^4]. */
static int deriv2_common(const or_basis* b, double tol, int want, double* out) {
    int ns = b->nshell;
    size_t n = b->nbf;
    size_t n4 = n * n * n * n;
    int ncoord = 3 * b->natom;
    int npair = ncoord * (ncoord + 1) / 2;
    double* vals = (double*)malloc(sizeof(double) * ns * ns);
    int* pairs = (int*)malloc(sizeof(int) * ns * (ns + 1));
    int np = or_schwarz(b, tol, vals, pairs);
    double tol2 = tol * tol;
    int (*plut)[2] = malloc(sizeof(int[2]) * npair);
    for (int x = 0, q = 0; x < ncoord; ++x)
        for (int y = x; y < ncoord; ++y, ++q) { plut[q][0] = x; plut[q][1] = y; }
    int bidx[12][12];
    for (int I = 0, q = 0; I < 12; ++I)
        for (int J = I; J < 12; ++J, ++q) bidx[I][J] = bidx[J][I] = q;
    memset(out, 0, sizeof(double) * (want < 0 ? npair : 1) * n4);
    #pragma omp parallel for schedule(dynamic)
    for (int ib = 0; ib < np; ++ib) {
        int p1 = pairs[2 * ib], p2 = pairs[2 * ib + 1];
        int n1 = ncart(b->l[p1]), n2 = ncart(b->l[p2]);
        size_t bf1 = b->bf[p1], bf2 = b->bf[p2];
        for (int ik = 0; ik < np; ++ik) {
            int p3 = pairs[2 * ik], p4 = pairs[2 * ik + 1];
            if (vals[p1 * ns + p2] * vals[p3 * ns + p4] < tol2) continue;
            int at[4] = {b->atom[p1], b->atom[p2], b->atom[p3], b->atom[p4]};
            if (at[0] == at[1] && at[0] == at[2] && at[0] == at[3]) continue;
            int n3 = ncart(b->l[p3]), n4s = ncart(b->l[p4]);
            size_t bf3 = b->bf[p3], bf4 = b->bf[p4];
            int nblk = n1 * n2 * n3 * n4s;
            double* buf = (double*)malloc(sizeof(double) * 78 * nblk);
            or_shell_quartet_deriv2(b, p1, p2, p3, p4, buf);
            for (int q = 0; q < npair; ++q) {
                if (want >= 0 && q != want) continue;
                int x = plut[q][0], y = plut[q][1];
                size_t off = want >= 0 ? 0 : (size_t)q * n4;
                for (int c1 = 0; c1 < 4; ++c1) {
                    if (at[c1] != x / 3) continue;
                    for (int c2 = 0; c2 < 4; ++c2) {
                        if (at[c2] != y / 3) continue;
                        const double* blk = buf + (size_t)bidx[3 * c1 + x % 3][3 * c2 + y % 3] * nblk;
                        for (int f1 = 0, idx = 0; f1 < n1; ++f1)
                            for (int f2 = 0; f2 < n2; ++f2)
                                for (int f3 = 0; f3 < n3; ++f3)
                                    for (int f4 = 0; f4 < n4s; ++f4, ++idx) {
                                        size_t i = bf1 + f1, j = bf2 + f2, k = bf3 + f3, l = bf4 + f4;
                                        double* e0 = &out[off + ((i * n + j) * n + k) * n + l];
                                        double v = *e0 + blk[idx];
                                        *e0 = v;
                                        if (p1 != p2) out[off + ((j * n + i) * n + k) * n + l] = v;
                                        if (p3 != p4) out[off + ((i * n + j) * n + l) * n + k] = v;
                                        if (p1 != p2 && p3 != p4) out[off + ((j * n + i) * n + l) * n + k] = v;
                                    }
                    }
                }
            }
            free(buf);
        }
    }
    free(vals); free(pairs); free(plut);
    return 0;
}

int or_eri_deriv_core2(const or_basis* b, double tol, double* out) {
    return deriv2_common(b, tol, -1, out);
}

/* compute_2e_deriv("eri", 0., deriv_vec) with sum(deriv_vec) == 2. Returns
 * -1 for any other order. */
int or_compute_2e_deriv2(const or_basis* b, double tol, const int* deriv_vec, double* out) {
    int ncoord = 3 * b->natom, order = 0, c[2] = {-1, -1}, k = 0;
    for (int i = 0; i < ncoord; ++i) {
        order += deriv_vec[i];
        for (int r = 0; r < deriv_vec[i] && k < 2; ++r) c[k++] = i;
    }
    if (order != 2) return -1;
    int q = 0;
    for (int x = 0; x < c[0]; ++x) q += ncoord - x;
    q += c[1] - c[0];
    return deriv2_common(b, tol, q, out);
}

/* Evaluate a list of shell quartets (bench CPU baseline and spot checks).
 * quartets: 4*nq shell indices.  out (optional): concatenated blocks.
 * Returns the sum of all integrals (a checksum that keeps the work live). */
double or_quartet_batch(const or_basis* b, int nq, const int* quartets, int nthreads, double* out,
                        const long long* out_off) {
    double total = 0.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    #pragma omp parallel for schedule(dynamic, 16) reduction(+:total)
    for (int t = 0; t < nq; ++t) {
        const int* s = quartets + 4 * t;
        int nblk = ncart(b->l[s[0]]) * ncart(b->l[s[1]]) * ncart(b->l[s[2]]) * ncart(b->l[s[3]]);
        double local[50625];
        or_shell_quartet(b, s[0], s[1], s[2], s[3], local);
        double acc = 0.0;
        for (int k = 0; k < nblk; ++k) acc += local[k];
        if (out) memcpy(out + out_off[t], local, sizeof(double) * nblk);
        total += acc;
    }
    return total;
}

double or_quartet_batch_deriv1(const or_basis* b, int nq, const int* quartets, int nthreads, double* out,
                               const long long* out_off) {
    double total = 0.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    #pragma omp parallel for schedule(dynamic, 4) reduction(+:total)
    for (int t = 0; t < nq; ++t) {
        const int* s = quartets + 4 * t;
        int nblk = ncart(b->l[s[0]]) * ncart(b->l[s[1]]) * ncart(b->l[s[2]]) * ncart(b->l[s[3]]);
        double* local = (double*)malloc(sizeof(double) * 12 * nblk);
        or_shell_quartet_deriv1(b, s[0], s[1], s[2], s[3], local);
        double acc = 0.0;
        for (int k = 0; k < 12 * nblk; ++k) acc += local[k];
        if (out) memcpy(out + out_off[t], local, sizeof(double) * 12 * nblk);
        free(local);
        total += acc;
    }
    return total;
}

/* --------------------------------------------------------- one-electron */
/* 1D Obara-Saika overlap table S[i][j], i <= imax, j <= jmax (no prefactor) */
static void os1d(double PA, double PB, double oo2p, int imax, int jmax, double* S) {
    int J = jmax + 1;
    for (int j = 0; j <= jmax; ++j)
        for (int i = 0; i <= imax; ++i) {
            double v;
            if (i == 0 && j == 0) v = 1.0;
            else if (j == 0) {
                v = PA * S[(i - 1) * J];
                if (i > 1) v += oo2p * (i - 1) * S[(i - 2) * J];
            } else {
                v = PB * S[i * J + j - 1];
                if (i > 0) v += oo2p * i * S[(i - 1) * J + j - 1];
                if (j > 1) v += oo2p * (j - 1) * S[i * J + j - 2];
            }
            S[i * J + j] = v;
        }
}

/* primitive overlap and kinetic blocks for shells of l = la, lb */
static void prim_st(int la, int lb, double a, const double* A, double bb, const double* B,
                    double* Sout, double* Tout) {
    double p = a + bb, P[3], AB2 = 0;
    for (int i = 0; i < 3; ++i) { P[i] = (a * A[i] + bb * B[i]) / p; AB2 += (A[i] - B[i]) * (A[i] - B[i]); }
    double pref = pow(M_PI / p, 1.5) * exp(-a * bb / p * AB2);
    int imax = la, jmax = lb + 2, J = jmax + 1;
    double S1[3][(OR_LMAX + 2) * (OR_LMAX + 4)];
    for (int i = 0; i < 3; ++i) os1d(P[i] - A[i], P[i] - B[i], 0.5 / p, imax, jmax, S1[i]);
    int na = ncart(la), nb = ncart(lb);
    for (int ia = 0; ia < na; ++ia) {
        int ga = coff(la) + ia;
        for (int ib = 0; ib < nb; ++ib) {
            int gb = coff(lb) + ib;
            double s[3], t[3];
            for (int d = 0; d < 3; ++d) {
                int i = g_comp[ga][d], j = g_comp[gb][d];
                s[d] = S1[d][i * J + j];
                double tv = bb * (2 * j + 1) * S1[d][i * J + j] - 2.0 * bb * bb * S1[d][i * J + j + 2];
                if (j > 1) tv -= 0.5 * j * (j - 1) * S1[d][i * J + j - 2];
                t[d] = tv;
            }
            if (Sout) Sout[ia * nb + ib] = pref * s[0] * s[1] * s[2];
            if (Tout) Tout[ia * nb + ib] = pref * (t[0] * s[1] * s[2] + s[0] * t[1] * s[2] + s[0] * s[1] * t[2]);
        }
    }
}

/* primitive nuclear attraction to a unit positive charge at C: (a|1/r_C|b) */
static void prim_v(int la, int lb, double a, const double* A, double bb, const double* B,
                   const double* Cn, double* out) {
    double p = a + bb, P[3], AB2 = 0, PC2 = 0, PA[3], PC[3], AB[3];
    for (int i = 0; i < 3; ++i) {
        P[i] = (a * A[i] + bb * B[i]) / p;
        AB[i] = A[i] - B[i]; AB2 += AB[i] * AB[i];
        PA[i] = P[i] - A[i]; PC[i] = P[i] - Cn[i]; PC2 += PC[i] * PC[i];
    }
    int Le = la + lb;
    int NE = coff(Le + 1);
    double F[OR_MMAX + 1];
    or_boys(Le, p * PC2, F);
    double pref = 2.0 * M_PI / p * exp(-a * bb / p * AB2);
    double* V = (double*)malloc(sizeof(double) * (Le + 1) * NE);
#define VV(m, e) V[(m) * NE + (e)]
    for (int m = 0; m <= Le; ++m) VV(m, 0) = pref * F[m];
    for (int L = 1; L <= Le; ++L)
        for (int g = coff(L); g < coff(L + 1); ++g) {
            int i = first_axis(g);
            int e1 = g_shift(g, i, -1), e2 = g_shift(g, i, -2);
            int n1 = g_comp[g][i] - 1;
            for (int m = 0; m <= Le - L; ++m) {
                double v = PA[i] * VV(m, e1) - PC[i] * VV(m + 1, e1);
                if (n1 > 0) v += n1 / (2.0 * p) * (VV(m, e2) - VV(m + 1, e2));
                VV(m, g) = v;
            }
        }
    double* Z = (double*)malloc(sizeof(double) * coff(OR_LT + 1) * coff(OR_LMAX + 3));
    double* x = (double*)malloc(sizeof(double) * NE);
    for (int e = 0; e < NE; ++e) x[e] = VV(0, e);
    hrr1(x, la, lb, AB, out, Z);
    free(Z); free(x); free(V);
#undef VV
}

/* S, T, V (nbf x nbf, Cartesian) -- Libint overlap/kinetic/nuclear. */
void or_oei(const or_basis* b, double* S, double* T, double* V) {
    init_tables();
    int n = b->nbf;
    memset(S, 0, sizeof(double) * n * n); memset(T, 0, sizeof(double) * n * n); memset(V, 0, sizeof(double) * n * n);
    double bs[225], bt[225], bv[225];
    for (int s1 = 0; s1 < b->nshell; ++s1)
        for (int s2 = 0; s2 < b->nshell; ++s2) {
            int la = b->l[s1], lb = b->l[s2], na = ncart(la), nb = ncart(lb);
            const double *A = ctr(b, s1), *B = ctr(b, s2);
            for (int i = b->poff[s1]; i < b->poff[s1 + 1]; ++i)
                for (int j = b->poff[s2]; j < b->poff[s2 + 1]; ++j) {
                    double w = b->coef[i] * b->coef[j];
                    prim_st(la, lb, b->alpha[i], A, b->alpha[j], B, bs, bt);
                    for (int k = 0; k < na * nb; ++k) {
                        int r = b->bf[s1] + k / nb, c = b->bf[s2] + k % nb;
                        S[r * n + c] += w * bs[k];
                        T[r * n + c] += w * bt[k];
                    }
                    for (int at = 0; at < b->natom; ++at) {
                        prim_v(la, lb, b->alpha[i], A, b->alpha[j], B, b->xyz + 3 * at, bv);
                        for (int k = 0; k < na * nb; ++k) {
                            int r = b->bf[s1] + k / nb, c = b->bf[s2] + k % nb;
                            V[r * n + c] -= b->Zn[at] * w * bv[k];
                        }
                    }
                }
        }
}

/* First nuclear derivatives of S, T, V: each (3N, nbf, nbf). */
void or_oei_deriv1(const or_basis* b, double* dS, double* dT, double* dV) {
    init_tables();
    int n = b->nbf, N3 = 3 * b->natom;
    size_t nn = (size_t)n * n;
    memset(dS, 0, sizeof(double) * N3 * nn); memset(dT, 0, sizeof(double) * N3 * nn); memset(dV, 0, sizeof(double) * N3 * nn);
    double su[225 * 2], tu[225 * 2], sd[225], td[225], vu[225 * 2], vd[225];
    for (int s1 = 0; s1 < b->nshell; ++s1)
        for (int s2 = 0; s2 < b->nshell; ++s2) {
            int la = b->l[s1], lb = b->l[s2], na = ncart(la), nb = ncart(lb);
            const double *A = ctr(b, s1), *B = ctr(b, s2);
            int atA = b->atom[s1], atB = b->atom[s2];
            for (int i = b->poff[s1]; i < b->poff[s1 + 1]; ++i)
                for (int j = b->poff[s2]; j < b->poff[s2 + 1]; ++j) {
                    double w = b->coef[i] * b->coef[j];
                    double al = b->alpha[i], be = b->alpha[j];
                    for (int side = 0; side < 2; ++side) {
                        /* side 0: derivative on bra centre A; side 1: on ket centre B */
                        int lu_a = la + (side == 0), lu_b = lb + (side == 1);
                        int ld_a = la - (side == 0), ld_b = lb - (side == 1);
                        int have_dn = (side == 0 ? la : lb) > 0;
                        double ex = side == 0 ? al : be;
                        int atom_d = side == 0 ? atA : atB;
                        prim_st(lu_a, lu_b, al, A, be, B, su, tu);
                        if (have_dn) prim_st(ld_a, ld_b, al, A, be, B, sd, td);
                        for (int ia = 0; ia < na; ++ia)
                            for (int ib = 0; ib < nb; ++ib) {
                                int ga = coff(la) + ia, gb = coff(lb) + ib;
                                for (int x = 0; x < 3; ++x) {
                                    int gau = side == 0 ? g_shift(ga, x, 1) : ga;
                                    int gbu = side == 1 ? g_shift(gb, x, 1) : gb;
                                    int ku = (gau - coff(lu_a)) * ncart(lu_b) + (gbu - coff(lu_b));
                                    double vs = 2 * ex * su[ku], vt = 2 * ex * tu[ku];
                                    int q = g_comp[side == 0 ? ga : gb][x];
                                    if (q > 0) {
                                        int gad = side == 0 ? g_shift(ga, x, -1) : ga;
                                        int gbd = side == 1 ? g_shift(gb, x, -1) : gb;
                                        int kd = (gad - coff(ld_a)) * ncart(ld_b) + (gbd - coff(ld_b));
                                        vs -= q * sd[kd]; vt -= q * td[kd];
                                    }
                                    size_t o = (size_t)(3 * atom_d + x) * nn + (size_t)(b->bf[s1] + ia) * n + b->bf[s2] + ib;
                                    dS[o] += w * vs;
                                    dT[o] += w * vt;
                                }
                            }
                        for (int at = 0; at < b->natom; ++at) {
                            const double* Cn = b->xyz + 3 * at;
                            prim_v(lu_a, lu_b, al, A, be, B, Cn, vu);
                            if (have_dn) prim_v(ld_a, ld_b, al, A, be, B, Cn, vd);
                            for (int ia = 0; ia < na; ++ia)
                                for (int ib = 0; ib < nb; ++ib) {
                                    int ga = coff(la) + ia, gb = coff(lb) + ib;
                                    for (int x = 0; x < 3; ++x) {
                                        int gau = side == 0 ? g_shift(ga, x, 1) : ga;
                                        int gbu = side == 1 ? g_shift(gb, x, 1) : gb;
                                        int ku = (gau - coff(lu_a)) * ncart(lu_b) + (gbu - coff(lu_b));
                                        double vv = 2 * ex * vu[ku];
                                        int q = g_comp[side == 0 ? ga : gb][x];
                                        if (q > 0) {
                                            int gad = side == 0 ? g_shift(ga, x, -1) : ga;
                                            int gbd = side == 1 ? g_shift(gb, x, -1) : gb;
                                            int kd = (gad - coff(ld_a)) * ncart(ld_b) + (gbd - coff(ld_b));
                                            vv -= q * vd[kd];
                                        }
                                        vv *= -b->Zn[at] * w;
                                        size_t rc = (size_t)(b->bf[s1] + ia) * n + b->bf[s2] + ib;
                                        dV[(size_t)(3 * atom_d + x) * nn + rc] += vv;
                                        /* operator centre: translational invariance of this term */
                                        dV[(size_t)(3 * at + x) * nn + rc] -= vv;
                                    }
                                }
                        }
                    }
                }
        }
}
