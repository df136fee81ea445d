/*
 * ORACLE -- test infrastructure only.  Nothing in the product (pyscf_ipu_amd)
 * links, loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py do, as the FP64 checker and the CPU baseline.
 *
 * Plain-C FP64 restatement of the integral arithmetic on the reference's hot
 * path.  The reference obtains its ERIs from libcint (PySCF 2.2.1,
 * `mol.intor("int2e_sph", aosym="s8")`, pyscf_ipu/nanoDFT/nanoDFT.py:285) and
 * ships a libcint-derived Rys kernel (pyscf_ipu/electron_repulsion/
 * int2e_sph.cpp:1785-1900 CINTg0_2e, :1475-1600 CINTg0_2e_2d, :2452-2723
 * CINT2e_loop_nopt).  This file follows that algorithm:
 *   pair data (exponent sums, Gaussian product centres, exp prefactor)
 *   -> Rys roots/weights of exp(-T t^2) on [0,1]
 *   -> 2D recurrence G(n,m) over (i+j, k+l) per root and axis
 *   -> horizontal transfer to (i,j,k,l)
 *   -> sum_r w_r Ix Iy Iz, contracted over primitives.
 * The Rys roots are NOT libcint's polynomial fits (int2e_sph.cpp:94-1470):
 * they are computed here from the Boys-function moments
 * F_k(T) = int_0^1 t^(2k) exp(-T t^2) dt with Gautschi's Chebyshev algorithm
 * and a Jacobi eigen-solve of the resulting Jacobi matrix, in long double.
 *
 * Parity pins (tests/test_oracle_integrals.py): Szabo-Ostlund H2 STO-3G
 * overlap / kinetic / nuclear / ERI values and the PyQuante primitive values
 * of pyscf_ipu test/test_integrals.py:24-133; an independent THO/McMurchie-
 * Davidson ERI restated from pyscf_ipu/experimental/integrals.py:186-244;
 * mpmath quadrature of the Rys weight; Boys closed forms.
 *
 * Conventions: libcint atm/bas/env (see pyscf_ipu_amd/gto.py); shells with
 * l <= 2 cartesian; l <= 1 only for spherical output (s: 1 function, p: x,y,z)
 * which is all the STO-3G / 6-31G path needs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ATM_SLOTS 6
#define BAS_SLOTS 8
#define MAXL 2
#define MAXROOTS 5
#define MAXPRIM 16

typedef long double ld;

static const double PI = 3.14159265358979323846;
static const double ANG_FAC[2] = {0.282094791773878143, 0.488602511902919921};

/* ---------------------------------------------------------------- Boys ---- */
/* F_m(T) for m = 0..mmax into f[].  T < 12: series for F_mmax,
 * F_m(T) = exp(-T) sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)), then downward
 * recursion F_{m-1} = (2T F_m + exp(-T)) / (2m-1).  T >= 12: F_0 from erf and
 * upward recursion F_{m+1} = ((2m+1) F_m - exp(-T)) / (2T). */
void oracle_boys(int mmax, double T, double *f)
{
    if (T < 12.0) {
        ld t = T, e = expl(-t);
        ld term = 1.0L / (2 * mmax + 1), sum = term;
        for (int k = 1; k < 200; k++) {
            term *= 2.0L * t / (2 * mmax + 2 * k + 1);
            sum += term;
            if (term < 1e-21L * sum) break;
        }
        ld fm = e * sum;
        f[mmax] = (double)fm;
        for (int m = mmax; m > 0; m--) {
            fm = (2.0L * t * fm + e) / (2 * m - 1);
            f[m - 1] = (double)fm;
        }
    } else {
        ld t = T, e = expl(-t);
        ld f0 = 0.5L * sqrtl((ld)PI / t) * erfl(sqrtl(t));
        f[0] = (double)f0;
        ld fm = f0;
        for (int m = 0; m < mmax; m++) {
            fm = ((2 * m + 1) * fm - e) / (2.0L * t);
            f[m + 1] = (double)fm;
        }
    }
}

static void boys_ld(int mmax, ld T, ld *f)
{
    if (T < 12.0L) {
        ld e = expl(-T);
        ld term = 1.0L / (2 * mmax + 1), sum = term;
        for (int k = 1; k < 300; k++) {
            term *= 2.0L * T / (2 * mmax + 2 * k + 1);
            sum += term;
            if (term < 1e-22L * sum) break;
        }
        f[mmax] = e * sum;
        for (int m = mmax; m > 0; m--) f[m - 1] = (2.0L * T * f[m] + e) / (2 * m - 1);
    } else {
        ld e = expl(-T);
        f[0] = 0.5L * sqrtl((ld)PI / T) * erfl(sqrtl(T));
        for (int m = 0; m < mmax; m++) f[m + 1] = ((2 * m + 1) * f[m] - e) / (2.0L * T);
    }
}

/* ------------------------------------------------------------ Rys roots ---- */
/* Symmetric eigen-solve by cyclic Jacobi rotations (n <= MAXROOTS). */
static void jacobi_eig(int n, ld a[MAXROOTS][MAXROOTS], ld v[MAXROOTS][MAXROOTS])
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) v[i][j] = (i == j);
    for (int sweep = 0; sweep < 60; sweep++) {
        ld off = 0;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) off += a[i][j] * a[i][j];
        if (off < 1e-40L) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                if (fabsl(a[p][q]) < 1e-300L) continue;
                ld theta = (a[q][q] - a[p][p]) / (2 * a[p][q]);
                ld t = (theta >= 0 ? 1 : -1) / (fabsl(theta) + sqrtl(theta * theta + 1));
                ld c = 1 / sqrtl(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    ld akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    ld apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    ld vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
}

/* Roots u_r = t_r^2 in (0,1) and weights w_r (sum w_r = F_0(T)) of the
 * n-point Gauss rule for the weight exp(-T u) / (2 sqrt(u)) on [0,1], whose
 * moments are F_k(T).  Chebyshev algorithm (Gautschi 1982) on the moments,
 * then Golub-Welsch. */
void oracle_rys(int n, double T, double *u, double *w)
{
    ld mu[2 * MAXROOTS];
    boys_ld(2 * n - 1, (ld)T, mu);
    ld alpha[MAXROOTS], beta[MAXROOTS];
    ld sig_prev[2 * MAXROOTS + 1], sig[2 * MAXROOTS + 1], sig_new[2 * MAXROOTS + 1];
    for (int l = 0; l < 2 * n; l++) { sig_prev[l] = 0; sig[l] = mu[l]; }
    alpha[0] = mu[1] / mu[0];
    beta[0] = mu[0];
    for (int k = 1; k < n; k++) {
        for (int l = k; l < 2 * n - k; l++)
            sig_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l];
        alpha[k] = sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1];
        beta[k] = sig_new[k] / sig[k - 1];
        for (int l = 0; l < 2 * n; l++) { sig_prev[l] = sig[l]; sig[l] = (l >= k && l < 2 * n - k) ? sig_new[l] : 0; }
    }
    ld a[MAXROOTS][MAXROOTS] = {{0}}, v[MAXROOTS][MAXROOTS];
    for (int i = 0; i < n; i++) {
        a[i][i] = alpha[i];
        if (i + 1 < n) a[i][i + 1] = a[i + 1][i] = sqrtl(beta[i + 1]);
    }
    jacobi_eig(n, a, v);
    /* sort ascending */
    int idx[MAXROOTS];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (a[idx[j]][idx[j]] < a[idx[i]][idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int i = 0; i < n; i++) {
        u[i] = (double)a[idx[i]][idx[i]];
        w[i] = (double)(mu[0] * v[0][idx[i]] * v[0][idx[i]]);
    }
}

/* -------------------------------------------------------- shell helpers ---- */
typedef struct {
    double r[3];
    int l, nprim;
    double e[MAXPRIM], c[MAXPRIM];
    int ao0;
} Shell;

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

/* cartesian exponents in libcint order: for l=1 x,y,z; l=2 xx,xy,xz,yy,yz,zz */
static void cart_list(int l, int (*lx)[3])
{
    int n = 0;
    for (int ix = l; ix >= 0; ix--)
        for (int iy = l - ix; iy >= 0; iy--) {
            lx[n][0] = ix; lx[n][1] = iy; lx[n][2] = l - ix - iy; n++;
        }
}

static void load_shells(const int *atm, const int *bas, int nbas, const double *env, Shell *sh)
{
    int ao = 0;
    for (int ib = 0; ib < nbas; ib++) {
        const int *b = bas + ib * BAS_SLOTS;
        const double *r = env + atm[b[0] * ATM_SLOTS + 1];
        Shell *s = sh + ib;
        s->r[0] = r[0]; s->r[1] = r[1]; s->r[2] = r[2];
        s->l = b[1];
        s->nprim = b[2];
        for (int k = 0; k < s->nprim; k++) {
            s->e[k] = env[b[5] + k];
            s->c[k] = env[b[6] + k] * (s->l <= 1 ? ANG_FAC[s->l] : 1.0);
        }
        s->ao0 = ao;
        ao += 2 * s->l + 1;
    }
}

/* ------------------------------------------------------------------ ERI ---- */
/* Contracted cartesian (ij|kl) block, out[fi][fj][fk][fl] row-major. */
static void eri_block(const Shell *A, const Shell *B, const Shell *C, const Shell *D, double *out)
{
    int li = A->l, lj = B->l, lk = C->l, ll = D->l;
    int nfi = ncart(li), nfj = ncart(lj), nfk = ncart(lk), nfl = ncart(ll);
    int nout = nfi * nfj * nfk * nfl;
    int nroots = (li + lj + lk + ll) / 2 + 1;
    int nmax = li + lj, mmax = lk + ll;
    int ci[6][3], cj[6][3], ck[6][3], cl[6][3];
    cart_list(li, ci); cart_list(lj, cj); cart_list(lk, ck); cart_list(ll, cl);
    memset(out, 0, sizeof(double) * nout);
    double AB[3], CD[3];
    for (int d = 0; d < 3; d++) { AB[d] = A->r[d] - B->r[d]; CD[d] = C->r[d] - D->r[d]; }
    double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    double rcd2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
    /* G[axis][root][n][m], I[axis][root][i][j][k][l] */
    double G[3][MAXROOTS][2 * MAXL + 1][2 * MAXL + 1];
    double I[3][MAXROOTS][MAXL + 1][MAXL + 1][MAXL + 1][MAXL + 1];
    for (int pa = 0; pa < A->nprim; pa++)
        for (int pb = 0; pb < B->nprim; pb++) {
            double ai = A->e[pa], aj = B->e[pb], aij = ai + aj;
            double kab = exp(-ai * aj / aij * rab2);
            double P[3];
            for (int d = 0; d < 3; d++) P[d] = (ai * A->r[d] + aj * B->r[d]) / aij;
            for (int pc = 0; pc < C->nprim; pc++)
                for (int pd = 0; pd < D->nprim; pd++) {
                    double ak = C->e[pc], al = D->e[pd], akl = ak + al;
                    double kcd = exp(-ak * al / akl * rcd2);
                    double Q[3], PQ[3];
                    for (int d = 0; d < 3; d++) { Q[d] = (ak * C->r[d] + al * D->r[d]) / akl; PQ[d] = P[d] - Q[d]; }
                    double rho = aij * akl / (aij + akl);
                    double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                    double pref = 2.0 * pow(PI, 2.5) / (aij * akl * sqrt(aij + akl)) * kab * kcd
                                  * A->c[pa] * B->c[pb] * C->c[pc] * D->c[pd];
                    double u[MAXROOTS], w[MAXROOTS];
                    oracle_rys(nroots, T, u, w);
                    for (int r = 0; r < nroots; r++) {
                        double t2 = u[r];
                        double b00 = 0.5 * t2 / (aij + akl);
                        double b10 = 0.5 / aij - 0.5 * akl * t2 / (aij * (aij + akl));
                        double b01 = 0.5 / akl - 0.5 * aij * t2 / (akl * (aij + akl));
                        for (int d = 0; d < 3; d++) {
                            double c00 = (P[d] - A->r[d]) - akl * PQ[d] * t2 / (aij + akl);
                            double c0p = (Q[d] - C->r[d]) + aij * PQ[d] * t2 / (aij + akl);
                            double (*g)[2 * MAXL + 1] = G[d][r];
                            g[0][0] = 1.0;
                            for (int n = 0; n < nmax; n++)
                                g[n + 1][0] = c00 * g[n][0] + (n > 0 ? n * b10 * g[n - 1][0] : 0.0);
                            for (int m = 0; m < mmax; m++)
                                for (int n = 0; n <= nmax; n++)
                                    g[n][m + 1] = c0p * g[n][m] + (m > 0 ? m * b01 * g[n][m - 1] : 0.0)
                                                  + (n > 0 ? n * b00 * g[n - 1][m] : 0.0);
                            /* horizontal transfer: j onto i with AB, l onto k with CD */
                            for (int k = 0; k <= lk; k++)
                                for (int l = 0; l <= ll; l++)
                                    for (int i = 0; i <= li; i++)
                                        for (int j = 0; j <= lj; j++) {
                                            /* (i,j) from g(i+j..) by binomial expansion of (x-B)^j = (x-A+AB)^j */
                                            double acc = 0.0;
                                            for (int jj = 0; jj <= j; jj++) {
                                                double bj = 1.0;
                                                for (int q = 0; q < jj; q++) bj = bj * (j - q) / (q + 1);
                                                double pj = pow(AB[d], j - jj);
                                                for (int ll2 = 0; ll2 <= l; ll2++) {
                                                    double bl = 1.0;
                                                    for (int q = 0; q < ll2; q++) bl = bl * (l - q) / (q + 1);
                                                    double pl = pow(CD[d], l - ll2);
                                                    acc += bj * pj * bl * pl * g[i + jj][k + ll2];
                                                }
                                            }
                                            I[d][r][i][j][k][l] = acc;
                                        }
                        }
                    }
                    int o = 0;
                    for (int fi = 0; fi < nfi; fi++)
                        for (int fj = 0; fj < nfj; fj++)
                            for (int fk = 0; fk < nfk; fk++)
                                for (int fl = 0; fl < nfl; fl++, o++) {
                                    double s = 0.0;
                                    for (int r = 0; r < nroots; r++) {
                                        double v = w[r];
                                        for (int d = 0; d < 3; d++)
                                            v *= I[d][r][ci[fi][d]][cj[fj][d]][ck[fk][d]][cl[fl][d]];
                                        s += v;
                                    }
                                    out[o] += pref * s;
                                }
                }
        }
}

static inline int64_t tri(int64_t a) { return a * (a + 1) / 2; }

/* Unique (8-fold symmetric) spherical ERIs of a s/p basis, packed as
 * eri[pq*(pq+1)/2 + rs] with p>=q, r>=s, pq>=rs (PySCF aosym="s8" layout).
 * Returns 0, or -1 for an unsupported shell (l > 1). */
int oracle_eri_s8(const int *atm, int natm, const int *bas, int nbas, const double *env, double *eri)
{
    (void)natm;
    Shell *sh = (Shell *)malloc(sizeof(Shell) * nbas);
    load_shells(atm, bas, nbas, env, sh);
    for (int i = 0; i < nbas; i++)
        if (sh[i].l > 1) { free(sh); return -1; }
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = nbas - 1; i >= 0; i--) {
        double buf[81];
        for (int j = 0; j <= i; j++)
            for (int k = 0; k <= i; k++)
                for (int l = 0; l <= (k == i ? j : k); l++) {
                    eri_block(sh + i, sh + j, sh + k, sh + l, buf);
                    int di = 2 * sh[i].l + 1, dj = 2 * sh[j].l + 1, dk = 2 * sh[k].l + 1, dl = 2 * sh[l].l + 1;
                    int o = 0;
                    for (int a = 0; a < di; a++)
                        for (int b = 0; b < dj; b++)
                            for (int c = 0; c < dk; c++)
                                for (int d = 0; d < dl; d++, o++) {
                                    int64_t p = sh[i].ao0 + a, q = sh[j].ao0 + b, r = sh[k].ao0 + c, s = sh[l].ao0 + d;
                                    /* one writer per unique element: the in-shell filters apply
                                     * only where the two shells of a pair (or the two pairs)
                                     * coincide; otherwise no other image lies in this block. */
                                    if (i == j && q > p) continue;
                                    if (k == l && s > r) continue;
                                    int64_t pq = tri(p) + q, rs = tri(r) + s;
                                    if (i == k && j == l && rs > pq) continue;
                                    int64_t hi = pq > rs ? pq : rs, lo = pq > rs ? rs : pq;
                                    eri[tri(hi) + lo] = buf[o];
                                }
                }
    }
    free(sh);
    return 0;
}

/* Single contracted shell quartet (spherical = cartesian for l<=1), out in
 * (i,j,k,l) row-major order. */
int oracle_eri_quartet(const int *atm, const int *bas, int nbas, const double *env, const int *shls, double *out)
{
    Shell *sh = (Shell *)malloc(sizeof(Shell) * nbas);
    load_shells(atm, bas, nbas, env, sh);
    eri_block(sh + shls[0], sh + shls[1], sh + shls[2], sh + shls[3], out);
    free(sh);
    return 0;
}

/* --------------------------------------------------------- one-electron ---- */
/* 1D overlap table S[i][j] for i <= la+2, j <= lb+2 (Obara-Saika). */
static void os_overlap_1d(double pa, double pb, double p, int imax, int jmax, double S[5][5])
{
    double h = 0.5 / p;
    S[0][0] = 1.0;
    for (int i = 0; i < imax; i++) S[i + 1][0] = pa * S[i][0] + (i > 0 ? i * h * S[i - 1][0] : 0.0);
    for (int j = 0; j < jmax; j++)
        for (int i = 0; i <= imax; i++)
            S[i][j + 1] = pb * S[i][j] + h * ((i > 0 ? i * S[i - 1][j] : 0.0) + (j > 0 ? j * S[i][j - 1] : 0.0));
}

/* Contracted cartesian one-electron blocks <a|O|b> (row-major [fa][fb], l <= 2
 * on either side): overlap S, kinetic T = <a|-nabla^2/2|b>, and V: with nuc < 0
 * the nuclear attraction sum_C -Z_C <a|1/|r-C||b>, with nuc = C only
 * <a|1/|r-C||b> (charge 1, the rinv operator of PySCF's int1e_*rinv). */
static void int1e_block(const Shell *A, const Shell *B, const int *atm, int natm, const double *env, int nuc,
                        double *S, double *T, double *V)
{
    int la = A->l, lb = B->l, na = ncart(la), nb = ncart(lb);
    int ca[6][3], cb[6][3];
    cart_list(la, ca); cart_list(lb, cb);
    memset(S, 0, sizeof(double) * na * nb);
    memset(T, 0, sizeof(double) * na * nb);
    memset(V, 0, sizeof(double) * na * nb);
    double AB[3];
    for (int d = 0; d < 3; d++) AB[d] = A->r[d] - B->r[d];
    double rab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    for (int pa = 0; pa < A->nprim; pa++)
        for (int pb = 0; pb < B->nprim; pb++) {
            double a = A->e[pa], b = B->e[pb], p = a + b, mu = a * b / p;
            double P[3];
            for (int d = 0; d < 3; d++) P[d] = (a * A->r[d] + b * B->r[d]) / p;
            double cc = A->c[pa] * B->c[pb];
            double S1[3][5][5];
            for (int d = 0; d < 3; d++) os_overlap_1d(P[d] - A->r[d], P[d] - B->r[d], p, la + 2, lb + 2, S1[d]);
            double s00 = pow(PI / p, 1.5) * exp(-mu * rab2);
            /* nuclear attraction by Rys quadrature over each nucleus */
            int nr = (la + lb) / 2 + 1;
            double vbuf[6][6];
            memset(vbuf, 0, sizeof(vbuf));
            for (int iat = 0; iat < natm; iat++) {
                if (nuc >= 0 && iat != nuc) continue;
                const double *Cn = env + atm[iat * ATM_SLOTS + 1];
                double Z = nuc >= 0 ? -1.0 : atm[iat * ATM_SLOTS + 0];
                double PC[3];
                for (int d = 0; d < 3; d++) PC[d] = P[d] - Cn[d];
                double Tn = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                double u[MAXROOTS], w[MAXROOTS];
                oracle_rys(nr, Tn, u, w);
                for (int r = 0; r < nr; r++) {
                    double J1[3][5][5];
                    for (int d = 0; d < 3; d++) {
                        double pa_ = P[d] - A->r[d] - u[r] * PC[d];
                        double pb_ = P[d] - B->r[d] - u[r] * PC[d];
                        double h = 0.5 * (1.0 - u[r]) / p;
                        double (*J)[5] = J1[d];
                        J[0][0] = 1.0;
                        for (int i = 0; i < la; i++) J[i + 1][0] = pa_ * J[i][0] + (i > 0 ? i * h * J[i - 1][0] : 0.0);
                        for (int j = 0; j < lb; j++)
                            for (int i = 0; i <= la; i++)
                                J[i][j + 1] = pb_ * J[i][j] + h * ((i > 0 ? i * J[i - 1][j] : 0.0) + (j > 0 ? j * J[i][j - 1] : 0.0));
                    }
                    for (int fa = 0; fa < na; fa++)
                        for (int fb = 0; fb < nb; fb++)
                            vbuf[fa][fb] += -Z * w[r] * J1[0][ca[fa][0]][cb[fb][0]] * J1[1][ca[fa][1]][cb[fb][1]]
                                            * J1[2][ca[fa][2]][cb[fb][2]];
                }
            }
            double vpre = 2.0 * PI / p * exp(-mu * rab2);
            for (int fa = 0; fa < na; fa++)
                for (int fb = 0; fb < nb; fb++) {
                    const int *x = ca[fa], *y = cb[fb];
                    double sx = S1[0][x[0]][y[0]], sy = S1[1][x[1]][y[1]], sz = S1[2][x[2]][y[2]];
                    double tt = 0.0;
                    for (int d = 0; d < 3; d++) {
                        int i = x[d], j = y[d];
                        double t1 = b * (2 * j + 1) * S1[d][i][j] - 2.0 * b * b * S1[d][i][j + 2]
                                    - (j >= 2 ? 0.5 * j * (j - 1) * S1[d][i][j - 2] : 0.0);
                        double others = 1.0;
                        for (int e = 0; e < 3; e++)
                            if (e != d) others *= S1[e][x[e]][y[e]];
                        tt += t1 * others;
                    }
                    S[fa * nb + fb] += cc * s00 * sx * sy * sz;
                    T[fa * nb + fb] += cc * s00 * tt;
                    V[fa * nb + fb] += cc * vpre * vbuf[fa][fb];
                }
        }
}

/* S, T, V (all nuclei) in the spherical AO basis, each nao x nao row-major. */
int oracle_int1e(const int *atm, int natm, const int *bas, int nbas, const double *env, int nao,
                 double *S, double *T, double *V)
{
    Shell *sh = (Shell *)malloc(sizeof(Shell) * nbas);
    load_shells(atm, bas, nbas, env, sh);
    for (int i = 0; i < nbas; i++)
        if (sh[i].l > 1) { free(sh); return -1; }
    memset(S, 0, sizeof(double) * nao * nao);
    memset(T, 0, sizeof(double) * nao * nao);
    memset(V, 0, sizeof(double) * nao * nao);
    double bs[36], bt[36], bv[36];
    for (int ia = 0; ia < nbas; ia++)
        for (int ib = 0; ib < nbas; ib++) {
            const Shell *A = sh + ia, *B = sh + ib;
            int na = ncart(A->l), nb = ncart(B->l);
            int1e_block(A, B, atm, natm, env, -1, bs, bt, bv);
            for (int fa = 0; fa < na; fa++)
                for (int fb = 0; fb < nb; fb++) {
                    int row = A->ao0 + fa, col = B->ao0 + fb;
                    S[row * nao + col] = bs[fa * nb + fb];
                    T[row * nao + col] = bt[fa * nb + fb];
                    V[row * nao + col] = bv[fa * nb + fb];
                }
        }
    free(sh);
    return 0;
}

/* ------------------------------------------------------ derivative integrals ---- */
/* The derivative of a cartesian Gaussian x^ix y^iy z^iz exp(-a r^2) centred at
 * A along the electron coordinate x: ix (i - 1_x) - 2a (i + 1_x).  The two
 * terms are integrals over the same shell with l - 1 (coefficients c) and
 * l + 1 (coefficients 2 a c), the coefficients keeping the s/p angular factor
 * of the original shell. */
static void shell_raised(const Shell *s, Shell *up)
{
    *up = *s;
    up->l = s->l + 1;
    for (int k = 0; k < s->nprim; k++) up->c[k] = 2.0 * s->e[k] * s->c[k];
}

static void shell_lowered(const Shell *s, Shell *dn)
{
    *dn = *s;
    dn->l = s->l - 1;
}

/* index of the cartesian component with exponents e in cart_list(l) order */
static int cart_index(int l, const int *e)
{
    int c[6][3];
    cart_list(l, c);
    for (int k = 0; k < ncart(l); k++)
        if (c[k][0] == e[0] && c[k][1] == e[1] && c[k][2] == e[2]) return k;
    return -1;
}

/* (nabla_x i, j | k l) of the s/p basis, dense [3][nao][nao][nao][nao] in the
 * layout of PySCF's mol.intor("int2e_ip1") (nanoDFT.py:428): the derivative
 * acts on the electron coordinate of the first function, i.e. it is minus the
 * derivative with respect to that function's centre. */
int oracle_eri_ip1(const int *atm, int natm, const int *bas, int nbas, const double *env, int nao, double *out)
{
    (void)natm;
    Shell *sh = (Shell *)malloc(sizeof(Shell) * nbas);
    load_shells(atm, bas, nbas, env, sh);
    for (int i = 0; i < nbas; i++)
        if (sh[i].l > 1) { free(sh); return -1; }
    const int64_t n1 = nao, n2 = n1 * nao, n3 = n2 * nao, n4 = n3 * nao;
    memset(out, 0, sizeof(double) * 3 * n4);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nbas; i++) {
        double up[6 * 27], dn[3 * 27];
        Shell A1, A0;
        shell_raised(sh + i, &A1);
        if (sh[i].l > 0) shell_lowered(sh + i, &A0);
        int ci[3][3];
        cart_list(sh[i].l, ci);
        for (int j = 0; j < nbas; j++)
            for (int k = 0; k < nbas; k++)
                for (int l = 0; l < nbas; l++) {
                    const Shell *B = sh + j, *C = sh + k, *D = sh + l;
                    int nj = ncart(B->l), nk = ncart(C->l), nl = ncart(D->l), rest = nj * nk * nl;
                    eri_block(&A1, B, C, D, up);
                    if (sh[i].l > 0) eri_block(&A0, B, C, D, dn);
                    for (int fi = 0; fi < ncart(sh[i].l); fi++)
                        for (int x = 0; x < 3; x++) {
                            int e[3] = {ci[fi][0], ci[fi][1], ci[fi][2]};
                            e[x] += 1;
                            int iu = cart_index(sh[i].l + 1, e);
                            e[x] -= 2;
                            int id = ci[fi][x] > 0 ? cart_index(sh[i].l - 1, e) : -1;
                            for (int r = 0; r < rest; r++) {
                                int fj = r / (nk * nl), fk = (r / nl) % nk, fl = r % nl;
                                double v = -up[iu * rest + r];
                                if (id >= 0) v += ci[fi][x] * dn[id * rest + r];
                                int64_t p = sh[i].ao0 + fi, q = B->ao0 + fj, rr = C->ao0 + fk, ss = D->ao0 + fl;
                                out[x * n4 + p * n3 + q * n2 + rr * n1 + ss] = v;
                            }
                        }
                }
    }
    free(sh);
    return 0;
}

/* <nabla_x i | O | j> for O = 1 (ipovlp), -nabla^2/2 (ipkin), sum_C -Z_C/|r-C|
 * (ipnuc), each [3][nao][nao] (PySCF int1e_ipovlp / ipkin / ipnuc layout,
 * nanoDFT.py:429-431); with nuc >= 0, `nuc_out` gets <nabla_x i | 1/|r-C| | j>
 * for C = nuc (int1e_iprinv under with_rinv_at_nucleus, nanoDFT.py:441). */
int oracle_int1e_ip(const int *atm, int natm, const int *bas, int nbas, const double *env, int nao, double *ipovlp,
                    double *ipkin, double *ipnuc, int nuc, double *nuc_out)
{
    Shell *sh = (Shell *)malloc(sizeof(Shell) * nbas);
    load_shells(atm, bas, nbas, env, sh);
    for (int i = 0; i < nbas; i++)
        if (sh[i].l > 1) { free(sh); return -1; }
    const int64_t nn = (int64_t)nao * nao;
    if (ipovlp) memset(ipovlp, 0, sizeof(double) * 3 * nn);
    if (ipkin) memset(ipkin, 0, sizeof(double) * 3 * nn);
    if (ipnuc) memset(ipnuc, 0, sizeof(double) * 3 * nn);
    if (nuc_out) memset(nuc_out, 0, sizeof(double) * 3 * nn);
    double us[18], ut[18], uv[18], ds[9], dt[9], dv[9], ur[18], dr[9], tmp_s[18], tmp_t[18];
    for (int ia = 0; ia < nbas; ia++) {
        Shell A1, A0;
        shell_raised(sh + ia, &A1);
        if (sh[ia].l > 0) shell_lowered(sh + ia, &A0);
        int ci[3][3];
        cart_list(sh[ia].l, ci);
        for (int ib = 0; ib < nbas; ib++) {
            const Shell *B = sh + ib;
            int nb = ncart(B->l);
            int1e_block(&A1, B, atm, natm, env, -1, us, ut, uv);
            if (sh[ia].l > 0) int1e_block(&A0, B, atm, natm, env, -1, ds, dt, dv);
            if (nuc >= 0) {
                int1e_block(&A1, B, atm, natm, env, nuc, tmp_s, tmp_t, ur);
                if (sh[ia].l > 0) int1e_block(&A0, B, atm, natm, env, nuc, tmp_s, tmp_t, dr);
            }
            for (int fa = 0; fa < ncart(sh[ia].l); fa++)
                for (int x = 0; x < 3; x++) {
                    int e[3] = {ci[fa][0], ci[fa][1], ci[fa][2]};
                    e[x] += 1;
                    int iu = cart_index(sh[ia].l + 1, e);
                    e[x] -= 2;
                    int id = ci[fa][x] > 0 ? cart_index(sh[ia].l - 1, e) : -1;
                    double lx = ci[fa][x];
                    for (int fb = 0; fb < nb; fb++) {
                        int64_t o = x * nn + (int64_t)(sh[ia].ao0 + fa) * nao + B->ao0 + fb;
                        double vs = -us[iu * nb + fb], vt = -ut[iu * nb + fb], vv = -uv[iu * nb + fb];
                        double vr = nuc >= 0 ? -ur[iu * nb + fb] : 0.0;
                        if (id >= 0) {
                            vs += lx * ds[id * nb + fb];
                            vt += lx * dt[id * nb + fb];
                            vv += lx * dv[id * nb + fb];
                            if (nuc >= 0) vr += lx * dr[id * nb + fb];
                        }
                        if (ipovlp) ipovlp[o] = vs;
                        if (ipkin) ipkin[o] = vt;
                        if (ipnuc) ipnuc[o] = vv;
                        if (nuc_out) nuc_out[o] = vr;
                    }
                }
        }
    }
    free(sh);
    return 0;
}

/* ---------------------------------------------------- sparse symmetric J/K ---- */
/* Restatement of sparse_symmetric_einsum (pyscf_ipu/nanoDFT/
 * sparse_symmetric_ERI.py:86-137) with the 32-entry index table of
 * compute_indices.cpp:54-95: 16 passes of gather dm -> multiply by the
 * multiplicity-divided ERI -> scatter-add, K passes scaled by -hyb/2. */
void oracle_sparse_jk(const double *eri, const int16_t *idx, int64_t nnz, const double *dm, int N, double hyb,
                      double *out)
{
    memset(out, 0, sizeof(double) * N * N);
    for (int sym = 0; sym < 16; sym++) {
        int isK = sym >= 8;
        double scale = isK ? -hyb / 2.0 : 1.0;
        int gsel = sym + isK * 8, ssel = sym + 8 + isK * 8;
        for (int64_t e = 0; e < nnz; e++) {
            uint32_t i = (uint16_t)idx[4 * e], j = (uint16_t)idx[4 * e + 1];
            uint32_t k = (uint16_t)idx[4 * e + 2], l = (uint16_t)idx[4 * e + 3];
            uint32_t t[32] = {i * N + j, j * N + i, i * N + j, j * N + i, k * N + l, l * N + k, k * N + l, l * N + k,
                              k * N + l, k * N + l, l * N + k, l * N + k, i * N + j, i * N + j, j * N + i, j * N + i,
                              k * N + j, k * N + i, l * N + j, l * N + i, i * N + l, i * N + k, j * N + l, j * N + k,
                              i * N + l, j * N + l, i * N + k, j * N + k, k * N + j, l * N + j, k * N + i, l * N + i};
            out[t[ssel]] += scale * dm[t[gsel]] * eri[e];
        }
    }
}
