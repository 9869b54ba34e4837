/*
 * Oracle — TEST INFRASTRUCTURE ONLY.  Plain-C, FP64 restatement of the reference's
 * fractional-step RK3/Crank-Nicolson substep for the periodic plane channel on the
 * turbulence-generator domain (SOLVERRK3_MOM_tg, SOLVERRK3_MOM_tg.f90:13-56).  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and only as the
 * checker; nothing in the product links or calls it.
 *
 * Layout (same as the reference, MEMALLOCT.f90:96-103, x fastest):
 *   halo fields   Q(i, J, k, c) for J = 0..M+1   -> q[((c*N + k)*(M+2) + J)*L + i]
 *   cell fields   RHS(i, J, k)  for J = 1..M     -> rhs[(k*M + J-1)*L + i]
 * i and k are 0-based here (reference I = i+1, K = k+1); J keeps the reference's value.
 * RCCI1/RCCI2/RNDI1/RNDI2 are 1 for the channel (COORDJR.f90:39-42) and are dropped.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 * Pinned by tests/test_oracle_ref.py (against oracle/_ref, the compiled reference) and by
 * tests/test_oracle_math.py (dense solves, DFT definitions, manufactured Poisson solutions).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int L, M, N;
    double dxi, dzi, dxqi, dzqi, volm;
    const double *dyfi;                /* [M]   DYFI(J) at [J-1]              */
    const double *dyci;                /* [M+2] DYCI(J) at [J]                */
    const double *wff, *wfb;           /* [M+2] YCL2ND_WFF/WFB(J) at [J]      */
    const double *amvr, *acvr, *apvr;  /* [3][M+2] A*VR(J, IDR) at [IDR-1][J] */
    const double *amph, *acph, *apph;  /* [M]   A*PH(J) at [J-1]              */
    const double *ak1, *ak3;           /* [L], [N] modified wavenumbers       */
    const double *tgam, *troh, *talp;  /* [4] RK coefficients, index NS       */
} or_mesh;

typedef struct {
    double *q;        /* [3][N][M+2][L] Q_tg          */
    double *pr;       /* [N][M+2][L]    PR_tg         */
    double *convh0;   /* [3][N][M][L]   CONVH0_tg     */
    double *qtmp;     /* [N][M+2][L]    QTMP_tg       */
    double *dph;      /* [N][M+2][L]    DPH_tg        */
    double *rhsllphi; /* [N][M][L]      RHSLLPHI_tg   */
    double *rhs;      /* [N][M][L]      RHS_tg        */
    double dpgrns;    /* last mass-flux correction (RHS_CvLpGpS_tg.f90:179-191) */
} or_state;

#define LL (m->L)
#define MM (m->M)
#define NN (m->N)
#define HX(a, i, J, k) (a)[((size_t)(k) * (MM + 2) + (J)) * LL + (i)]
#define QC(c, i, J, k) s->q[(((size_t)(c) * NN + (k)) * (MM + 2) + (J)) * LL + (i)]
#define CX(a, i, J, k) (a)[((size_t)(k) * MM + (J) - 1) * LL + (i)]
#define H0C(c, i, J, k) s->convh0[(((size_t)(c) * NN + (k)) * MM + (J) - 1) * LL + (i)]

static inline int ip_(const or_mesh *m, int i) { return i == m->L - 1 ? 0 : i + 1; }
static inline int im_(const or_mesh *m, int i) { return i == 0 ? m->L - 1 : i - 1; }
static inline int kp_(const or_mesh *m, int k) { return k == m->N - 1 ? 0 : k + 1; }
static inline int km_(const or_mesh *m, int k) { return k == 0 ? m->N - 1 : k - 1; }

/* CONVECTION_X_tg (CONVECTION_X.f90:13-64), _Y_tg (CONVECTION_Y.f90:13-84, channel branch),
 * _Z_tg (CONVECTION_Z.f90:13-73): divergence-form staggered convection into QTMP, DPH,
 * RHSLLPHI.  XND2CL = YND2CL = ZND2CL = 0.5 (COORDJR.f90:466-468). */
void or_convection(const or_mesh *m, or_state *s)
{
    const double h = 0.5;
    memset(s->qtmp, 0, sizeof(double) * (size_t)NN * (MM + 2) * LL);
    memset(s->dph, 0, sizeof(double) * (size_t)NN * (MM + 2) * LL);
    memset(s->rhsllphi, 0, sizeof(double) * (size_t)NN * MM * LL);
    for (int J = 1; J <= MM; ++J) {
        const int JM = J - 1, JP = J + 1, JJP = J + 1;
        const double coe1 = h * h * m->dxi;
        const double coe2 = m->dyfi[J - 1] * h;
        const double coe3 = h * h * m->dzi;
        for (int k = 0; k < NN; ++k) {
            const int km = km_(m, k), kp = kp_(m, k);
            for (int i = 0; i < LL; ++i) {
                const int ip = ip_(m, i), im = im_(m, i);
                double H11 = ((QC(0, ip, J, k) + QC(0, i, J, k)) * (QC(0, ip, J, k) + QC(0, i, J, k)) -
                              (QC(0, im, J, k) + QC(0, i, J, k)) * (QC(0, im, J, k) + QC(0, i, J, k))) * coe1;
                double H12 = ((QC(1, i, JP, k) + QC(1, im, JP, k)) *
                                  (m->wff[JJP] * QC(0, i, JP, k) + m->wfb[JJP] * QC(0, i, J, k)) -
                              (QC(1, i, J, k) + QC(1, im, J, k)) *
                                  (m->wff[J] * QC(0, i, J, k) + m->wfb[J] * QC(0, i, JM, k))) * coe2;
                double H13 = ((QC(2, i, J, kp) + QC(2, im, J, kp)) * (QC(0, i, J, kp) + QC(0, i, J, k)) -
                              (QC(2, i, J, k) + QC(2, im, J, k)) * (QC(0, i, J, k) + QC(0, i, J, km))) * coe3;
                HX(s->qtmp, i, J, k) = -(H11 + H12 + H13);
            }
        }
    }
    for (int J = 2; J <= MM; ++J) {   /* NYI = 2 on the bottom rank */
        const int JM = J - 1, JP = J + 1;
        const double coe1 = m->dxi * h;
        const double coe2 = m->dyci[J] * h * h;
        const double coe31 = m->dzi * h;
        for (int k = 0; k < NN; ++k) {
            const int km = km_(m, k), kp = kp_(m, k);
            for (int i = 0; i < LL; ++i) {
                const int ip = ip_(m, i), im = im_(m, i);
                double H21 = ((m->wff[J] * QC(0, ip, J, k) + m->wfb[J] * QC(0, ip, JM, k)) *
                                  (QC(1, ip, J, k) + QC(1, i, J, k)) -
                              (m->wff[J] * QC(0, i, J, k) + m->wfb[J] * QC(0, i, JM, k)) *
                                  (QC(1, i, J, k) + QC(1, im, J, k))) * coe1;
                double H22 = ((QC(1, i, JP, k) + QC(1, i, J, k)) * (QC(1, i, JP, k) + QC(1, i, J, k)) -
                              (QC(1, i, J, k) + QC(1, i, JM, k)) * (QC(1, i, J, k) + QC(1, i, JM, k))) * coe2;
                double H23 = ((m->wff[J] * QC(2, i, J, kp) + m->wfb[J] * QC(2, i, JM, kp)) *
                                  (QC(1, i, J, kp) + QC(1, i, J, k)) -
                              (m->wff[J] * QC(2, i, J, k) + m->wfb[J] * QC(2, i, JM, k)) *
                                  (QC(1, i, J, k) + QC(1, i, J, km))) * coe31;
                HX(s->dph, i, J, k) = -(H21 + H22 + H23);
            }
        }
    }
    for (int J = 1; J <= MM; ++J) {
        const int JM = J - 1, JP = J + 1, JJP = J + 1;
        const double coe1 = h * h * m->dxi;
        const double coe2 = m->dyfi[J - 1] * h;
        const double coe3 = h * h * m->dzi;
        for (int k = 0; k < NN; ++k) {
            const int km = km_(m, k), kp = kp_(m, k);
            for (int i = 0; i < LL; ++i) {
                const int ip = ip_(m, i), im = im_(m, i);
                double H31 = ((QC(0, ip, J, k) + QC(0, ip, J, km)) * (QC(2, ip, J, k) + QC(2, i, J, k)) -
                              (QC(0, i, J, k) + QC(0, i, J, km)) * (QC(2, i, J, k) + QC(2, im, J, k))) * coe1;
                double H32 = ((QC(1, i, JP, k) + QC(1, i, JP, km)) *
                                  (m->wff[JJP] * QC(2, i, JP, k) + m->wfb[JJP] * QC(2, i, J, k)) -
                              (QC(1, i, J, k) + QC(1, i, J, km)) *
                                  (m->wff[J] * QC(2, i, J, k) + m->wfb[J] * QC(2, i, JM, k))) * coe2;
                double H33 = ((QC(2, i, J, kp) + QC(2, i, J, k)) * (QC(2, i, J, kp) + QC(2, i, J, k)) -
                              (QC(2, i, J, k) + QC(2, i, J, km)) * (QC(2, i, J, k) + QC(2, i, J, km))) * coe3;
                CX(s->rhsllphi, i, J, k) = -(H31 + H32 + H33);
            }
        }
    }
}

/* RHS_CvLpGpS_tg (RHS_CvLpGpS_tg.f90:12-204), channel, iFlowDriven = 1, NFLOW = 1.
 * Rows J >= NII; for IDR = 2 the bottom row J = 1 of RHS_tg is left as it was. */
void or_rhs(const or_mesh *m, or_state *s, int ns, int idr, double dt, double cvisc)
{
    const int c = idr - 1;
    const int nii = (idr == 2) ? 2 : 1;
    const double *conv = idr == 1 ? s->qtmp : (idr == 2 ? s->dph : NULL);
    const double coe1 = m->talp[ns] * cvisc;   /* M_inlet / D_inlet = 1 (modules.f90:584-586) */
    for (int k = 0; k < NN; ++k) {
        const int km = km_(m, k), kp = kp_(m, k);
        for (int J = nii; J <= MM; ++J) {
            const double ap = m->apvr[c * (MM + 2) + J], ac = m->acvr[c * (MM + 2) + J],
                         am = m->amvr[c * (MM + 2) + J];
            for (int i = 0; i < LL; ++i) {
                const int ip = ip_(m, i), im = im_(m, i);
                const double convh = conv ? HX(conv, i, J, k) : CX(s->rhsllphi, i, J, k);
                double rhsc = m->tgam[ns] * convh + m->troh[ns] * H0C(c, i, J, k);
                double rhsl = coe1 * ((QC(c, ip, J, k) - 2.0 * QC(c, i, J, k) + QC(c, im, J, k)) * m->dxqi +
                                      (QC(c, i, J, kp) - 2.0 * QC(c, i, J, k) + QC(c, i, J, km)) * m->dzqi +
                                      (QC(c, i, J + 1, k) * ap + QC(c, i, J, k) * ac + QC(c, i, J - 1, k) * am));
                H0C(c, i, J, k) = convh;
                CX(s->rhs, i, J, k) = (rhsc + rhsl) * dt;
            }
        }
    }
    const double coe2 = m->talp[ns] * dt;
    for (int k = 0; k < NN; ++k) {
        const int km = km_(m, k);
        for (int J = nii; J <= MM; ++J) {
            const double sucacj = m->dyci[J];
            for (int i = 0; i < LL; ++i) {
                double pgm;
                if (idr == 1) pgm = (HX(s->pr, i, J, k) - HX(s->pr, im_(m, i), J, k)) * m->dxi * coe2;
                else if (idr == 2) pgm = (HX(s->pr, i, J, k) - HX(s->pr, i, J - 1, k)) * sucacj * coe2;
                else pgm = (HX(s->pr, i, J, k) - HX(s->pr, i, J, km)) * m->dzi * coe2;
                CX(s->rhs, i, J, k) = CX(s->rhs, i, J, k) - pgm;
            }
        }
    }
    if (idr == 1) {
        double intg = 0.0;   /* summed I, J, K order as RHS_CvLpGpS_tg.f90:171-178 */
        for (int i = 0; i < LL; ++i)
            for (int J = 1; J <= MM; ++J)
                for (int k = 0; k < NN; ++k) intg = intg + CX(s->rhs, i, J, k) / m->dyfi[J - 1];
        const double dpgrns = intg / m->volm;
        s->dpgrns = dpgrns;
        for (int k = 0; k < NN; ++k)
            for (int i = 0; i < LL; ++i)
                for (int J = 1; J <= MM; ++J) CX(s->rhs, i, J, k) = CX(s->rhs, i, J, k) - dpgrns;
    }
}

/* TDMAIJI_CYC (TDMAIJI_CYC.f90:67-157) for one line of length n with constant-in-line
 * coefficients a (sub), b (diag), c (super), in place on r (stride rs).  Work: 3n doubles. */
void or_tdma_cyc(int n, double a, double b, double c, double *r, long rs, double *work)
{
    double *H0 = work, *G1 = work + n, *G2 = work + 2 * n;
    H0[0] = c / b;
    G1[0] = r[0] / b;
    G2[0] = -a / b;
    for (int i = 1; i <= n - 3; ++i) {
        double pp = 1.0 / (b - H0[i - 1] * a);
        double qq1 = G1[i - 1] * a, qq2 = G2[i - 1] * a;
        H0[i] = c * pp;
        G1[i] = (r[i * rs] - qq1) * pp;
        G2[i] = -qq2 * pp;
    }
    {
        int i = n - 2;
        double pp = 1.0 / (b - H0[i - 1] * a);
        double qq1 = G1[i - 1] * a, qq2 = G2[i - 1] * a;
        G1[i] = (r[i * rs] - qq1) * pp;
        G2[i] = (-c - qq2) * pp;
    }
    for (int i = n - 3; i >= 0; --i) {
        G1[i] = -H0[i] * G1[i + 1] + G1[i];
        G2[i] = -H0[i] * G2[i + 1] + G2[i];
    }
    {
        int i = n - 1;
        double pp = r[i * rs] - c * G1[0] - a * G1[i - 1];
        double qq1 = b + c * G2[0] + a * G2[i - 1];
        r[i * rs] = pp / qq1;
    }
    for (int i = n - 2; i >= 0; --i) r[i * rs] = G1[i] + G2[i] * r[(n - 1) * rs];
}

/* TDMAIJJ_nonCYC (TDMAIJJ_nonCYC.f90:67-123) for one line with per-row coefficients a,b,c
 * and zero boundary values BCJ, in place on r (stride rs).  Work: 2n doubles. */
void or_tdma_noncyc(int n, const double *a, const double *b, const double *c, double *r, long rs,
                    double *work)
{
    double *H = work, *G = work + n;
    H[0] = c[0] / b[0];
    G[0] = r[0] / b[0];
    for (int j = 1; j <= n - 2; ++j) {
        double pp = 1.0 / (b[j] - H[j - 1] * a[j]);
        H[j] = c[j] * pp;
        G[j] = (r[j * rs] - a[j] * G[j - 1]) * pp;
    }
    {
        int j = n - 1;
        double pp = 1.0 / (b[j] - H[j - 1] * a[j]);
        G[j] = (r[j * rs] - a[j] * G[j - 1]) * pp;
        r[j * rs] = G[j];
    }
    for (int j = n - 2; j >= 0; --j) r[j * rs] = G[j] - H[j] * r[(j + 1) * rs];
}

/* The reference's four batched Thomas solvers with full coefficient arrays, on Fortran 2-D
 * arrays A, B, C, R(IS:IS+isz-1, JS:JS+jsz-1) (I fastest); R is overwritten by the solution.
 *   kind 0  TDMAIJI_CYC    (TDMAIJI_CYC.f90:67-157)    cyclic, lines along I (one per J)
 *   kind 1  TDMAIJJ_nonCYC (TDMAIJJ_nonCYC.f90:67-123) lines along J, BCJ(isz, 2)
 *   kind 2  TDMAIJI_nonCYC (TDMAIJI_nonCYC.f90:67-123) lines along I, BCI(jsz, 2)
 *   kind 3  TDMAIJJ_CYC    (TDMAIJJ_CYC.f90:67-157)    cyclic, lines along J (one per I)
 * bc (kinds 1, 2): the boundary values BC(line, 1) before the first and BC(line, 2) after the
 * last element, column-major.  Each line follows the reference's recurrences in its order. */
void or_tdma_lines(int kind, int isz, int jsz, const double *A, const double *B, const double *C, double *R,
                   const double *bc)
{
    const int alongI = (kind == 0 || kind == 2), cyc = (kind == 0 || kind == 3);
    const int n = alongI ? isz : jsz, nl = alongI ? jsz : isz;
    const long es = alongI ? 1 : isz, lstr = alongI ? isz : 1;
    double *H = (double *)malloc(sizeof(double) * 3 * n);
    double *G1 = H + n, *G2 = H + 2 * n;
    for (int l = 0; l < nl; ++l) {
        const double *a = A + l * lstr, *b = B + l * lstr, *c = C + l * lstr;
        double *r = R + l * lstr;
#define E(p, i) p[(long)(i) * es]
        if (cyc) {
            H[0] = E(c, 0) / E(b, 0);
            G1[0] = E(r, 0) / E(b, 0);
            G2[0] = -E(a, 0) / E(b, 0);
            for (int i = 1; i <= n - 3; ++i) {
                const double pp = 1.0 / (E(b, i) - H[i - 1] * E(a, i));
                const double qq1 = G1[i - 1] * E(a, i), qq2 = G2[i - 1] * E(a, i);
                H[i] = E(c, i) * pp;
                G1[i] = (E(r, i) - qq1) * pp;
                G2[i] = -qq2 * pp;
            }
            {
                const int i = n - 2;
                const double pp = 1.0 / (E(b, i) - H[i - 1] * E(a, i));
                const double qq1 = G1[i - 1] * E(a, i), qq2 = G2[i - 1] * E(a, i);
                G1[i] = (E(r, i) - qq1) * pp;
                G2[i] = (-E(c, i) - qq2) * pp;
            }
            for (int i = n - 3; i >= 0; --i) {
                G1[i] = -H[i] * G1[i + 1] + G1[i];
                G2[i] = -H[i] * G2[i + 1] + G2[i];
            }
            {
                const int i = n - 1;
                const double pp = E(r, i) - E(c, i) * G1[0] - E(a, i) * G1[i - 1];
                const double qq1 = E(b, i) + E(c, i) * G2[0] + E(a, i) * G2[i - 1];
                E(r, i) = pp / qq1;
            }
            for (int i = n - 2; i >= 0; --i) E(r, i) = G1[i] + G2[i] * E(r, n - 1);
        } else {
            const double bc1 = bc[l], bc2 = bc[nl + l];
            H[0] = E(c, 0) / E(b, 0);
            G1[0] = (E(r, 0) - E(a, 0) * bc1) / E(b, 0);
            for (int i = 1; i <= n - 2; ++i) {
                const double pp = 1.0 / (E(b, i) - H[i - 1] * E(a, i));
                H[i] = E(c, i) * pp;
                G1[i] = (E(r, i) - E(a, i) * G1[i - 1]) * pp;
            }
            {
                const int i = n - 1;
                const double pp = 1.0 / (E(b, i) - H[i - 1] * E(a, i));
                G1[i] = (E(r, i) - E(c, i) * bc2 - E(a, i) * G1[i - 1]) * pp;
                E(r, i) = G1[i];
            }
            for (int i = n - 2; i >= 0; --i) E(r, i) = G1[i] - H[i] * E(r, i + 1);
        }
#undef E
    }
    free(H);
}

/* MOMFA_tg (MOMFA_tg.f90:15-47): x (MOMFA1_X_tg :51-106) and z (MOMFA2_Z_tg :110-169) cyclic
 * solves over rows J >= N2I, the y solve (MOMFA3_Y_tg :172-239) over all rows J = 1..M, then
 * Q += RHS for rows J >= N2I. */
void or_momfa(const or_mesh *m, or_state *s, int ns, int idr, double dt, double cvisc)
{
    const int c = idr - 1;
    const int n2i = (idr == 2) ? 2 : 1;
    const double facoe = 0.5 * m->talp[ns] * dt * cvisc;
    int nmax = LL > NN ? LL : NN;
    if (MM + 2 > nmax) nmax = MM + 2;
    double *work = (double *)malloc(sizeof(double) * 3 * nmax);
    double *a = (double *)malloc(sizeof(double) * (MM + 2) * 3);
    double *b = a + MM + 2, *cc = b + MM + 2;
    {   /* x: a = c = -FACOE*DXQI, b = 1 - a - c */
        const double co = -facoe * m->dxqi;
        const double bb = 1.0 - co - co;
        for (int k = 0; k < NN; ++k)
            for (int J = n2i; J <= MM; ++J) or_tdma_cyc(LL, co, bb, co, &CX(s->rhs, 0, J, k), 1, work);
    }
    {   /* z: a = c = -FACOE*DZQI (times RNDI2/RCCI2 = 1) */
        const double co = 1.0 * (-facoe * m->dzqi);
        const double bb = 1.0 - co - co;
        for (int i = 0; i < LL; ++i)
            for (int J = n2i; J <= MM; ++J)
                or_tdma_cyc(NN, co, bb, co, &CX(s->rhs, i, J, 0), (long)MM * LL, work);
    }
    {   /* y: b = 1 - FACOE*ACVR, a = -FACOE*AMVR, c = -FACOE*APVR over J = 1..NCL2 */
        for (int J = 1; J <= MM; ++J) {
            b[J - 1] = 1.0 - facoe * m->acvr[c * (MM + 2) + J];
            cc[J - 1] = -facoe * m->apvr[c * (MM + 2) + J];
            a[J - 1] = -facoe * m->amvr[c * (MM + 2) + J];
        }
        for (int k = 0; k < NN; ++k)
            for (int i = 0; i < LL; ++i) or_tdma_noncyc(MM, a, b, cc, &CX(s->rhs, i, 1, k), LL, work);
    }
    for (int k = 0; k < NN; ++k)
        for (int J = n2i; J <= MM; ++J)
            for (int i = 0; i < LL; ++i) QC(c, i, J, k) = CX(s->rhs, i, J, k) + QC(c, i, J, k);
    free(work);
    free(a);
}

/* BC_WALL_Q_tg (BC_WALL.f90:212-255, :646-657): ghost u, w = 0; v = 0 on both walls. */
void or_bc_q(const or_mesh *m, or_state *s)
{
    for (int k = 0; k < NN; ++k)
        for (int i = 0; i < LL; ++i) {
            QC(0, i, MM + 1, k) = 0.0; QC(2, i, MM + 1, k) = 0.0; QC(1, i, MM + 1, k) = 0.0;
            QC(0, i, 0, k) = 0.0; QC(2, i, 0, k) = 0.0; QC(1, i, 1, k) = 0.0; QC(1, i, 0, k) = 0.0;
        }
}

/* BC_WALL_PR / BC_WALL_DPH (BC_WALL.f90:285-411): ghost rows copy the first/last cell. */
void or_bc_copy(const or_mesh *m, double *a)
{
    for (int k = 0; k < NN; ++k)
        for (int i = 0; i < LL; ++i) {
            HX(a, i, MM + 1, k) = HX(a, i, MM, k);
            HX(a, i, 0, k) = HX(a, i, 1, k);
        }
}

/* DIVG_tg (DIVG.f90:15-51). */
void or_divg(const or_mesh *m, or_state *s, int ns, double dt)
{
    const double coe0 = 1.0 / (dt * m->talp[ns]);
    for (int J = 1; J <= MM; ++J) {
        const double coe1 = m->dxi, coe2 = m->dyfi[J - 1];
        for (int k = 0; k < NN; ++k) {
            const int kp = kp_(m, k);
            for (int i = 0; i < LL; ++i) {
                double d = (QC(0, ip_(m, i), J, k) - QC(0, i, J, k)) * coe1 +
                           (QC(1, i, J + 1, k) - QC(1, i, J, k)) * coe2 +
                           (QC(2, i, J, kp) - QC(2, i, J, k)) * m->dzi;
                CX(s->rhsllphi, i, J, k) = d * coe0;
            }
        }
    }
}

/* Complex DFT by mixed-radix decimation in time; sign = -1 forward, +1 backward,
 * unnormalised: X[k] = sum_j x[j] exp(sign * 2 pi i j k / n).  Twiddles are evaluated
 * directly with cos/sin of the reduced angle, so no recurrence error accumulates. */
static void dft_rec(int n, int sign, const double *xr, const double *xi, int stride,
                    double *yr, double *yi, double *tr, double *ti)
{
    if (n == 1) { yr[0] = xr[0]; yi[0] = xi[0]; return; }
    int p = 2;
    while (n % p) ++p;
    const int mlen = n / p;
    for (int r = 0; r < p; ++r)
        dft_rec(mlen, sign, xr + (size_t)r * stride, xi + (size_t)r * stride, stride * p,
                tr + (size_t)r * mlen, ti + (size_t)r * mlen, yr + (size_t)r * mlen, yi + (size_t)r * mlen);
    const double tp = 2.0 * 3.14159265358979323846 / n;
    for (int k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        for (int r = 0; r < p; ++r) {
            long e = ((long)r * k) % n;
            double cr = cos(tp * e), ci = sign * sin(tp * e);
            double ar = tr[(size_t)r * mlen + k % mlen], ai = ti[(size_t)r * mlen + k % mlen];
            sr += ar * cr - ai * ci;
            si += ar * ci + ai * cr;
        }
        yr[k] = sr; yi[k] = si;
    }
}

void or_dft(int n, int sign, const double *xr, const double *xi, double *yr, double *yi)
{
    double *t = (double *)malloc(sizeof(double) * 4 * n);
    double *ar = (double *)malloc(sizeof(double) * 2 * n);
    memcpy(ar, xr, sizeof(double) * n);
    memcpy(ar + n, xi, sizeof(double) * n);
    dft_rec(n, sign, ar, ar + n, 1, yr, yi, t, t + n);
    (void)t;
    free(ar);
    free(t);
}

/* FFT99_POIS3D_periodicxz (possion3d_FFT99.f90:271-505), channel (non-cyclic y):
 * x forward real transform normalised by 1/L (FFT99 ISIGN=-1, FFT99.f90:95-103), z forward
 * complex transform times 1/N (CFFT99 ISIGN=-1, CFFT99.f90:171-200), (kx,kz)=(0,0) gauge:
 * RHS(0,0,J=1) = 0 and row 1 decoupled (possion3d_FFT99.f90:224-231, :341-344), per-mode
 * TDMAIJJ_nonCYC in y for the real and imaginary parts, then the inverse transforms
 * (the imaginary parts of kx = 0 and kx = L/2 are not used, FFT99.f90:84-86). */
void or_poisson(const or_mesh *m, const double *rhs_cells, double *dph_halo)
{
    const int L = LL, M = MM, N = NN, H = L / 2 + 1;
    double *sr = (double *)calloc((size_t)H * M * N, sizeof(double));   /* [k][J-1][kx] */
    double *si = (double *)calloc((size_t)H * M * N, sizeof(double));
    int nmax = L > N ? L : N;
    double *ar = (double *)malloc(sizeof(double) * nmax), *ai = (double *)malloc(sizeof(double) * nmax);
    double *br = (double *)malloc(sizeof(double) * nmax), *bi = (double *)malloc(sizeof(double) * nmax);
#define SP(a, kx, J, k) (a)[((size_t)(k) * M + (J) - 1) * H + (kx)]
    for (int J = 1; J <= M; ++J) {
        for (int k = 0; k < N; ++k) {
            for (int i = 0; i < L; ++i) { ar[i] = rhs_cells[((size_t)k * M + J - 1) * L + i]; ai[i] = 0.0; }
            or_dft(L, -1, ar, ai, br, bi);
            for (int kx = 0; kx < H; ++kx) { SP(sr, kx, J, k) = br[kx] / L; SP(si, kx, J, k) = bi[kx] / L; }
        }
        for (int kx = 0; kx < H; ++kx) {
            for (int k = 0; k < N; ++k) { ar[k] = SP(sr, kx, J, k); ai[k] = SP(si, kx, J, k); }
            or_dft(N, -1, ar, ai, br, bi);
            for (int k = 0; k < N; ++k) { SP(sr, kx, J, k) = br[k] * (1.0 / N); SP(si, kx, J, k) = bi[k] * (1.0 / N); }
        }
    }
    SP(sr, 0, 1, 0) = 0.0;
    SP(si, 0, 1, 0) = 0.0;
    double *a = (double *)malloc(sizeof(double) * M * 3), *b = a + M, *c = b + M;
    double *work = (double *)malloc(sizeof(double) * 2 * M);
    for (int k = 0; k < N; ++k)
        for (int kx = 0; kx < H; ++kx) {
            for (int J = 1; J <= M; ++J) {
                a[J - 1] = m->amph[J - 1];
                c[J - 1] = m->apph[J - 1];
                b[J - 1] = m->acph[J - 1] + (-m->ak1[kx] / 1.0 - m->ak3[k]);
            }
            if (kx == 0 && k == 0) { a[0] = 0.0; c[0] = 0.0; b[0] = m->acph[0] + (-m->ak1[0] / 1.0 - m->ak3[0]); }
            or_tdma_noncyc(M, a, b, c, &SP(sr, kx, 1, k), H, work);
            or_tdma_noncyc(M, a, b, c, &SP(si, kx, 1, k), H, work);
        }
    for (int J = 1; J <= M; ++J) {
        for (int kx = 0; kx < H; ++kx) {
            for (int k = 0; k < N; ++k) { ar[k] = SP(sr, kx, J, k); ai[k] = SP(si, kx, J, k); }
            or_dft(N, +1, ar, ai, br, bi);
            for (int k = 0; k < N; ++k) { SP(sr, kx, J, k) = br[k]; SP(si, kx, J, k) = bi[k]; }
        }
        for (int k = 0; k < N; ++k) {
            for (int kx = 0; kx < H; ++kx) { ar[kx] = SP(sr, kx, J, k); ai[kx] = SP(si, kx, J, k); }
            ai[0] = 0.0;
            ai[L / 2] = 0.0;
            for (int kx = H; kx < L; ++kx) { ar[kx] = ar[L - kx]; ai[kx] = -ai[L - kx]; }
            or_dft(L, +1, ar, ai, br, bi);
            for (int i = 0; i < L; ++i) dph_halo[((size_t)k * (M + 2) + J) * L + i] = br[i];
        }
    }
#undef SP
    free(sr); free(si); free(ar); free(ai); free(br); free(bi); free(a); free(work);
}

/* VELOUPDT_tg (VELO_CORRECTION.f90:14-100). */
void or_veloupdt(const or_mesh *m, or_state *s, int ns, double dt)
{
    double coe1 = dt * m->talp[ns] * m->dxi;
    for (int k = 0; k < NN; ++k)
        for (int J = 1; J <= MM; ++J)
            for (int i = 0; i < LL; ++i)
                QC(0, i, J, k) = QC(0, i, J, k) - (HX(s->dph, i, J, k) - HX(s->dph, im_(m, i), J, k)) * coe1;
    coe1 = dt * m->talp[ns] * m->dzi;
    for (int k = 0; k < NN; ++k)
        for (int J = 1; J <= MM; ++J)
            for (int i = 0; i < LL; ++i)
                QC(2, i, J, k) = QC(2, i, J, k) - (HX(s->dph, i, J, k) - HX(s->dph, i, J, km_(m, k))) * coe1;
    for (int J = 2; J <= MM; ++J) {
        coe1 = dt * m->talp[ns] * m->dyci[J];
        for (int k = 0; k < NN; ++k)
            for (int i = 0; i < LL; ++i)
                QC(1, i, J, k) = QC(1, i, J, k) - (HX(s->dph, i, J, k) - HX(s->dph, i, J - 1, k)) * coe1;
    }
}

/* PRCALC_tg (PRESSURE_CORRECTION.f90:15-36): p += phi - (alpha dt / 2Re) RHSLLPHI. */
void or_prcalc(const or_mesh *m, or_state *s, int ns, double dt, double cvisc)
{
    const double be = 0.5 * m->talp[ns] * dt * cvisc;
    for (int k = 0; k < NN; ++k)
        for (int J = 1; J <= MM; ++J)
            for (int i = 0; i < LL; ++i)
                HX(s->pr, i, J, k) = HX(s->pr, i, J, k) + HX(s->dph, i, J, k) - be * CX(s->rhsllphi, i, J, k);
}

/* SOLVERRK3_MOM_tg (SOLVERRK3_MOM_tg.f90:13-56), one substep NS in {0,1,2,3}. */
void or_substep(const or_mesh *m, or_state *s, int ns, double dt, double cvisc)
{
    or_convection(m, s);
    for (int idr = 1; idr <= 3; ++idr) {
        or_rhs(m, s, ns, idr, dt, cvisc);
        or_momfa(m, s, ns, idr, dt, cvisc);
    }
    or_bc_q(m, s);
    or_divg(m, s, ns, dt);
    memset(s->dph, 0, sizeof(double) * (size_t)NN * (MM + 2) * LL);
    or_poisson(m, s->rhsllphi, s->dph);
    or_bc_copy(m, s->dph);
    or_veloupdt(m, s, ns, dt);
    or_bc_q(m, s);
    or_prcalc(m, s, ns, dt, cvisc);
    or_bc_copy(m, s->pr);
}

/* ======================================================================================
 * Energy equation of the main domain (table mode): SOLVERRK3_ENG_io (SOLVERRK3_ENG_IO.f90:11-82)
 * with RHS_ENERGY_EXPLICIT (RHS_ENERGY_EXPLICIT.f90:11-155), THERM_PROP_UPDATE(IMAIND)
 * (THERM_PROP_UPDATE.f90:124-168) and BC_WALL_THERMAL(IMAIND) (BC_WALL.f90:484-643), channel
 * (RCCI1 = RNDI1 = RCCI2 = 1, COORDJR.f90:39-42), periodic x and z, one y-domain.
 * ====================================================================================== */

typedef struct {
    int n, imin_dh, imax_dh, imin_h, imax_h;   /* 1-based IMIN_/IMAX_ of the reference */
    const double *dh, *h, *t, *d, *m, *k, *cp, *b;
    const double *spl;   /* [7][3][n]: DHH, HT, HD, HM, HK, HCp, HB x (B, C, D) */
} or_table;

/* spline_interpolation_* (thermo_properties_table_searching.f90:678-724 and siblings): range
 * check with tolerance 1e-8 (the reference STOPs; here bit 1 of the result), bisection
 * I = 1, J = N + 1, Horner Y(I) + dx (B(I) + dx (C(I) + dx D(I))). */
static double or_spline(int n, const double *x, const double *y, const double *sb, const double *sc,
                        const double *sd, int imin, int imax, double xv, int *err)
{
    if ((xv - x[imin - 1]) < -1.0e-8 || (xv - x[imax - 1]) > 1.0e-8) *err = 1;
    int I = 1, J = n + 1;
    while (J > I + 1) {
        const int K = (I + J) / 2;
        if (xv < x[K - 1]) J = K;
        else I = K;
    }
    const double dx = xv - x[I - 1];
    return y[I - 1] + dx * (sb[I - 1] + dx * (sc[I - 1] + dx * sd[I - 1]));
}

/* THERM_PROP_UPDATE_FROM_DH, table branch (THERM_PROP_UPDATE.f90:45-52): out = H, T, D, M, K,
 * Cp, B with D from the H->M spline, M from H->K and K from H->D (lines 49-51).  Returns 1 if
 * a lookup left the table range. */
int or_prop_from_dh(const or_table *t, double dh, double *out)
{
    int err = 0;
    const int n = t->n;
#define SP(s, c) (t->spl + ((size_t)(s) * 3 + (c)) * n)
    const double h = or_spline(n, t->dh, t->h, SP(0, 0), SP(0, 1), SP(0, 2), t->imin_dh, t->imax_dh, dh, &err);
    out[0] = h;
    out[1] = or_spline(n, t->h, t->t, SP(1, 0), SP(1, 1), SP(1, 2), t->imin_h, t->imax_h, h, &err);
    out[2] = or_spline(n, t->h, t->m, SP(3, 0), SP(3, 1), SP(3, 2), t->imin_h, t->imax_h, h, &err);
    out[3] = or_spline(n, t->h, t->k, SP(4, 0), SP(4, 1), SP(4, 2), t->imin_h, t->imax_h, h, &err);
    out[4] = or_spline(n, t->h, t->d, SP(2, 0), SP(2, 1), SP(2, 2), t->imin_h, t->imax_h, h, &err);
    out[5] = or_spline(n, t->h, t->cp, SP(5, 0), SP(5, 1), SP(5, 2), t->imin_h, t->imax_h, h, &err);
    out[6] = or_spline(n, t->h, t->b, SP(6, 0), SP(6, 1), SP(6, 2), t->imin_h, t->imax_h, h, &err);
#undef SP
    return err;
}

typedef struct {
    int wall_type[2];        /* bottom, top: 1 fixed heat flux, 2 fixed temperature */
    double cthecd;           /* CTHECD = 1 / (REN PRT0) (Flow_State.f90:79) */
    double dhmin, dhmax;     /* DHmin, DHmax of the table (clamp, SOLVERRK3_ENG_IO.f90:39-42) */
    const double *wallflux;  /* [2][N][L] WALLFLUX(I, iBotWall | iTopWall, K) */
    double wall_state[2][5]; /* fixed-temperature walls: T, H, D, M, K (T_WAL_GV, ...) */
    int top_face;            /* face index of the top ghost-row weights (the reference uses the
                                local row count of the last rank; M on one y-domain) */
} or_eng_par;

typedef struct {
    double *dh, *h, *t, *d, *m, *k, *cp;   /* [N][M+2][L] */
    double *g;                             /* [3][N][M+2][L] G_io */
    double *rhs0;                          /* [N][M][L] RHS_ENERGY0 */
    double *d0, *m0, *dh0;                 /* [N][M+2][L] DENSITY0, Viscousity0, DH0 */
} or_eng;

static int or_eng_cell_props(const or_mesh *m, const or_table *tb, or_eng *e, int i, int J, int k)
{
    double o[7];
    const int err = or_prop_from_dh(tb, HX(e->dh, i, J, k), o);
    HX(e->h, i, J, k) = o[0];
    HX(e->t, i, J, k) = o[1];
    HX(e->d, i, J, k) = o[2];
    HX(e->m, i, J, k) = o[3];
    HX(e->k, i, J, k) = o[4];
    HX(e->cp, i, J, k) = o[5];
    return err;
}

/* BC_WALL_THERMAL(IMAIND): iso-flux walls extrapolate DH to the ghost row and look the
 * properties up there (BC_WALL_Isoflux, BC_WALL.f90:566-643; not BETA); iso-thermal walls take
 * the wall state (BC_WALL_IsoTHERMAL, :519-562). */
int or_eng_walls(const or_mesh *m, const or_table *tb, const or_eng_par *p, or_eng *e)
{
    int err = 0;
    for (int w = 0; w < 2; ++w) {
        const int J = w == 0 ? 0 : MM + 1;
        for (int k = 0; k < NN; ++k)
            for (int i = 0; i < LL; ++i) {
                if (p->wall_type[w] == 1) {
                    if (w == 0)
                        HX(e->dh, i, 0, k) = 2.0 * HX(e->dh, i, 1, k) - m->wff[2] * HX(e->dh, i, 2, k) -
                                             m->wfb[2] * HX(e->dh, i, 1, k);
                    else
                        HX(e->dh, i, J, k) = 2.0 * HX(e->dh, i, MM, k) - m->wff[p->top_face] * HX(e->dh, i, MM, k) -
                                             m->wfb[p->top_face] * HX(e->dh, i, MM - 1, k);
                    err |= or_eng_cell_props(m, tb, e, i, J, k);
                } else {
                    const double *ws = p->wall_state[w];
                    HX(e->t, i, J, k) = ws[0];
                    HX(e->h, i, J, k) = ws[1];
                    HX(e->d, i, J, k) = ws[2];
                    HX(e->m, i, J, k) = ws[3];
                    HX(e->k, i, J, k) = ws[4];
                    HX(e->dh, i, J, k) = HX(e->h, i, J, k) * HX(e->d, i, J, k);
                }
            }
    }
    return err;
}

/* State from DH as the harness loads it: properties of every row 0..M+1 from DH, then the
 * walls, RHS_ENERGY0 = 0. */
int or_eng_init(const or_mesh *m, const or_table *tb, const or_eng_par *p, or_eng *e)
{
    int err = 0;
    for (int k = 0; k < NN; ++k)
        for (int J = 0; J <= MM + 1; ++J)
            for (int i = 0; i < LL; ++i) err |= or_eng_cell_props(m, tb, e, i, J, k);
    err |= or_eng_walls(m, tb, p, e);
    memset(e->rhs0, 0, sizeof(double) * (size_t)LL * MM * NN);
    return err;
}

int or_eng_substep(const or_mesh *m, const or_table *tb, const or_eng_par *p, or_eng *e, int ns, double dt)
{
    const size_t hs = (size_t)NN * (MM + 2) * LL;
    memcpy(e->m0, e->m, sizeof(double) * hs);
    memcpy(e->d0, e->d, sizeof(double) * hs);
    memcpy(e->dh0, e->dh, sizeof(double) * hs);
    const double *g1 = e->g, *g2 = e->g + hs, *g3 = e->g + 2 * hs;
    const double xnd2cl = 0.5, znd2cl = 0.5, one = 1.0;
    const double coe1 = m->dxi * xnd2cl * 0.5;
    const double coe11 = m->dxqi * xnd2cl * p->cthecd;
    const double ca = m->tgam[ns] * dt, cb = m->troh[ns] * dt;
    const double realmin = 1.0e-14;
    for (int J = 1; J <= MM; ++J) {
        const double coe2 = m->dyfi[J - 1] * one * 0.5;
        const double coe3 = m->dzi * znd2cl * one * 0.5;
        const double coe33 = m->dzqi * znd2cl * one * p->cthecd;
        const double coe22 = m->dyfi[J - 1] * one * p->cthecd;
        double coe221, coe222;
        if (J == 1) {
            coe221 = m->dyci[J + 1] / one;
            coe222 = p->wall_type[0] == 1 ? 1.0 / one : m->dyci[J] / one;
        } else if (J == MM) {
            coe222 = m->dyci[J] / one;
            coe221 = p->wall_type[1] == 1 ? 1.0 / one : m->dyci[J + 1] / one;
        } else {
            coe221 = m->dyci[J + 1] / one;
            coe222 = m->dyci[J] / one;
        }
        const double wffp = m->wff[J + 1], wfbp = m->wfb[J + 1], wff = m->wff[J], wfb = m->wfb[J];
        for (int i = 0; i < LL; ++i) {
            const int ip = ip_(m, i), im = im_(m, i);
            for (int k = 0; k < NN; ++k) {
                const int kp = kp_(m, k), km = km_(m, k);
                const double *H = e->h, *T = e->t, *K = e->k;
                double dx_dhu = (HX(H, ip, J, k) + HX(H, i, J, k)) * (HX(g1, ip, J, k) + HX(g1, ip, J, k)) -
                                (HX(H, i, J, k) + HX(H, im, J, k)) * (HX(g1, i, J, k) + HX(g1, i, J, k));
                dx_dhu = dx_dhu * coe1;
                double dy_dhv = (wffp * HX(H, i, J + 1, k) + wfbp * HX(H, i, J, k)) *
                                    (HX(g2, i, J + 1, k) + HX(g2, i, J + 1, k)) -
                                (wff * HX(H, i, J, k) + wfb * HX(H, i, J - 1, k)) * (HX(g2, i, J, k) + HX(g2, i, J, k));
                dy_dhv = dy_dhv * coe2;
                double dz_dhw = (HX(H, i, J, kp) + HX(H, i, J, k)) * (HX(g3, i, J, kp) + HX(g3, i, J, kp)) -
                                (HX(H, i, J, k) + HX(H, i, J, km)) * (HX(g3, i, J, k) + HX(g3, i, J, k));
                dz_dhw = dz_dhw * coe3;
                double dx_kdt = (HX(K, ip, J, k) + HX(K, i, J, k)) * (HX(T, ip, J, k) - HX(T, i, J, k)) -
                                (HX(K, i, J, k) + HX(K, im, J, k)) * (HX(T, i, J, k) - HX(T, im, J, k));
                dx_kdt = dx_kdt * coe11;
                double kdty1 = (wffp * HX(K, i, J + 1, k) + wfbp * HX(K, i, J, k)) *
                               (HX(T, i, J + 1, k) - HX(T, i, J, k)) * coe221;
                double kdty2 = (wff * HX(K, i, J, k) + wfb * HX(K, i, J - 1, k)) *
                               (HX(T, i, J, k) - HX(T, i, J - 1, k)) * coe222;
                if (J == 1 && p->wall_type[0] == 1)
                    kdty2 = -p->wallflux[(size_t)k * LL + i] * coe222;
                else if (J == MM && p->wall_type[1] == 1)
                    kdty1 = -p->wallflux[((size_t)NN + k) * LL + i] * coe221;
                const double dy_kdt = (kdty1 - kdty2) * coe22;
                double dz_kdt = (HX(K, i, J, kp) + HX(K, i, J, k)) * (HX(T, i, J, kp) - HX(T, i, J, k)) -
                                (HX(K, i, J, km) + HX(K, i, J, k)) * (HX(T, i, J, k) - HX(T, i, J, km));
                dz_kdt = dz_kdt * coe33;
                const double r = -dx_dhu - dy_dhv - dz_dhw + dx_kdt + dy_kdt + dz_kdt;
                double dh = HX(e->dh, i, J, k) + ca * r + cb * CX(e->rhs0, i, J, k);
                dh = fmin(p->dhmax - realmin, dh);
                dh = fmax(p->dhmin + realmin, dh);
                HX(e->dh, i, J, k) = dh;
                CX(e->rhs0, i, J, k) = r;
            }
        }
    }
    int err = 0;
    for (int k = 0; k < NN; ++k)
        for (int J = 1; J <= MM; ++J)
            for (int i = 0; i < LL; ++i) err |= or_eng_cell_props(m, tb, e, i, J, k);
    err |= or_eng_walls(m, tb, p, e);
    return err;
}

/* ======================================================================================
 * Main-domain momentum substep with variable density and viscosity, explicit viscous scheme
 * (iVisScheme = 0, the setting of every main-domain template): SOLVERRK3_MOM_io
 * (SOLVERRK3_MOM_IO.f90:11-120) and the staggered properties SOLVE computes before it in a
 * heated run (DENSITY_TIME_DERIVATION, DENSITY_Staggered, MU_Staggered; SOLVE.f90:122-124).
 * Channel (RCCI1 = RCCI2 = RNDI1 = 1), periodic x and z, no inlet (NXI = 1), one y-domain,
 * iWeightedPre = 0.  The Poisson solve is FFT99_POIS3D_periodicxz(IIO), the reference's
 * is_FFT_FFT99 option (SOLVERRK3_MOM_IO.f90:92), with the same gauge as the TG domain.
 * ====================================================================================== */
typedef struct {
    double *g, *g0, *q;      /* [3][N][M+2][L] G_io, G0_io, Q_io                           */
    double *pr, *dph;        /* [N][M+2][L] PR_io, DPH_io (pressure correction)            */
    double *d0, *m0;         /* [N][M+2][L] DENSITY0, Viscousity0                          */
    double *dstg, *mstg;     /* [3][N][M+2][L] D_STG, MU_STG                               */
    double *explt0;          /* [3][N][M][L] EXPLT0_io                                     */
    double *drho;            /* [N][M][L] DrhoDtP                                          */
    double *divu;            /* [N][M][L] DivU_io                                          */
    double *h[3];            /* [N][M][L] QTMP_io, DPH_io, RHSLLPHI_io as momentum terms   */
    double *rhsp;            /* [N][M][L] RHSLLPHI_io as the Poisson right-hand side       */
    double *m1;              /* [N][M+2][L] Viscousity (current; Crank-Nicolson terms)      */
    double *rim;             /* [N][M][L] RHS_io: implicit viscous part, then the increment */
    double *rkeep;           /* NULL or [3][N][M][L]: RHS_io of each component as MOMFA_io gets it */
} or_io;

typedef struct {
    double cvisc;            /* CVISC = 1/REN                                              */
    int gravity;             /* |iGravity|: 0 none, else the direction of the buoyancy term */
    double f_a;              /* F_A (Flow_State.f90:104-123)                               */
    int flow_driven;         /* iFlowDriven: 1 constant mass flux, 2 constant Cf           */
    double cf_given;         /* Cf_Given (iFlowDriven = 2)                                 */
    int vis_scheme;          /* iVisScheme: 0 explicit, 1 Crank-Nicolson                    */
} or_io_par;

#define V3(a, c, i, J, k) (a)[(((size_t)(c) * NN + (k)) * (MM + 2) + (J)) * LL + (i)]
#define C3(a, c, i, J, k) (a)[(((size_t)(c) * NN + (k)) * MM + (J) - 1) * LL + (i)]

/* BC_WALL_G / BC_WALL_Q on the main domain (BC_WALL.f90:212-290, :435-481): ghost rows 0 and
 * M+1 zero, v zero on the bottom wall row J = 1. */
static void or_io_bc_zero(const or_mesh *m, double *a)
{
    for (int k = 0; k < NN; ++k)
        for (int i = 0; i < LL; ++i) {
            for (int c = 0; c < 3; ++c) {
                V3(a, c, i, 0, k) = 0.0;
                V3(a, c, i, MM + 1, k) = 0.0;
            }
            V3(a, 1, i, 1, k) = 0.0;
        }
}

/* DENSITY_TIME_DERIVATION (DIVG.f90:249-317). */
void or_io_drhodt(const or_mesh *m, or_io *s, int ns)
{
    for (int J = 1; J <= MM; ++J) {
        const double coe2 = m->dyfi[J - 1], coe3 = m->dzi;
        for (int i = 0; i < LL; ++i) {
            const int ip = ip_(m, i);
            for (int k = 0; k < NN; ++k) {
                const int kp = kp_(m, k);
                const double divx = (V3(s->g, 0, ip, J, k) - V3(s->g, 0, i, J, k)) * m->dxi;
                const double divx0 = (V3(s->g0, 0, ip, J, k) - V3(s->g0, 0, i, J, k)) * m->dxi;
                const double divy = (V3(s->g, 1, i, J + 1, k) - V3(s->g, 1, i, J, k)) * coe2;
                const double divy0 = (V3(s->g0, 1, i, J + 1, k) - V3(s->g0, 1, i, J, k)) * coe2;
                const double divz = (V3(s->g, 2, i, J, kp) - V3(s->g, 2, i, J, k)) * coe3;
                const double divz0 = (V3(s->g0, 2, i, J, kp) - V3(s->g0, 2, i, J, k)) * coe3;
                CX(s->drho, i, J, k) = -m->tgam[ns] * (divx + divy + divz) - m->troh[ns] * (divx0 + divy0 + divz0);
            }
        }
    }
}

/* DENSITY_Staggered (MASSFLUX_CALC_IO.f90:262-323): rho at the u, v, w points from DENSITY0;
 * v's wall rows take the ghost-row density. */
void or_io_dstg(const or_mesh *m, or_io *s)
{
    const double h = 0.5;   /* XND2CL = ZND2CL (COORDJR.f90:461-463) */
    for (int k = 0; k < NN; ++k) {
        const int km = km_(m, k);
        for (int J = 1; J <= MM; ++J)
            for (int i = 0; i < LL; ++i) {
                V3(s->dstg, 0, i, J, k) = (HX(s->d0, i, J, k) + HX(s->d0, im_(m, i), J, k)) * h;
                V3(s->dstg, 2, i, J, k) = (HX(s->d0, i, J, k) + HX(s->d0, i, J, km)) * h;
            }
        for (int J = 2; J <= MM; ++J)
            for (int i = 0; i < LL; ++i)
                V3(s->dstg, 1, i, J, k) = m->wff[J] * HX(s->d0, i, J, k) + m->wfb[J] * HX(s->d0, i, J - 1, k);
        for (int i = 0; i < LL; ++i) {
            V3(s->dstg, 1, i, 1, k) = HX(s->d0, i, 0, k);
            V3(s->dstg, 1, i, MM + 1, k) = HX(s->d0, i, MM + 1, k);
        }
    }
}

/* MU_Staggered (MASSFLUX_CALC_IO.f90:175-259): mu at the (x,y), (x,z) and (y,z) edges from
 * Viscousity0; the xy and yz edges on the walls take the ghost-row viscosity. */
void or_io_mustg(const or_mesh *m, or_io *s)
{
    const double h = 0.5;
    for (int k = 0; k < NN; ++k) {
        const int km = km_(m, k);
        for (int J = 2; J <= MM; ++J)
            for (int i = 0; i < LL; ++i) {
                const int im = im_(m, i);
                V3(s->mstg, 0, i, J, k) = ((HX(s->m0, i, J, k) + HX(s->m0, im, J, k)) * m->wff[J] +
                                           (HX(s->m0, i, J - 1, k) + HX(s->m0, im, J - 1, k)) * m->wfb[J]) * h;
                V3(s->mstg, 2, i, J, k) = ((HX(s->m0, i, J, km) + HX(s->m0, i, J, k)) * m->wff[J] +
                                           (HX(s->m0, i, J - 1, km) + HX(s->m0, i, J - 1, k)) * m->wfb[J]) * h;
            }
        for (int J = 1; J <= MM; ++J)
            for (int i = 0; i < LL; ++i) {
                const int im = im_(m, i);
                V3(s->mstg, 1, i, J, k) = (HX(s->m0, i, J, k) + HX(s->m0, im, J, k) + HX(s->m0, i, J, km) +
                                           HX(s->m0, im, J, km)) * h * h;
            }
        for (int i = 0; i < LL; ++i) {
            V3(s->mstg, 0, i, 1, k) = HX(s->m0, i, 0, k);
            V3(s->mstg, 2, i, 1, k) = HX(s->m0, i, 0, k);
            V3(s->mstg, 0, i, MM + 1, k) = HX(s->m0, i, MM + 1, k);
            V3(s->mstg, 2, i, MM + 1, k) = HX(s->m0, i, MM + 1, k);
        }
    }
}

/* VELOCITY_CALC_io (MASSFLUX_CALC_IO.f90:55-107): u = G / rho at the staggered points, then
 * the walls (BC_WALL_Q_io). */
void or_io_velocity(const or_mesh *m, or_io *s)
{
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < NN; ++k)
            for (int J = (c == 1 ? 2 : 1); J <= MM; ++J)
                for (int i = 0; i < LL; ++i)
                    V3(s->q, c, i, J, k) = V3(s->g, c, i, J, k) / V3(s->dstg, c, i, J, k);
    or_io_bc_zero(m, s->q);
}

/* CONVECTION_X_io / _Y_io / _Z_io (CONVECTION_X.f90:67-156, CONVECTION_Y.f90:117-272,
 * CONVECTION_Z.f90:106-303): the divergence-form terms with G as the advecting flux. */
void or_io_convection(const or_mesh *m, or_io *s)
{
    const double h = 0.5;
    const double cx1 = m->dxi * h * h, cx3 = m->dzi * h * h;
    const double cy1 = m->dxi * h, cy30 = m->dzi * h;
    const double cz1 = m->dxi * h * h, cz30 = m->dzi * h * h;
    const double *G = s->g, *Q = s->q;
    for (int J = 1; J <= MM; ++J) {
        const double wffp = m->wff[J + 1], wfbp = m->wfb[J + 1], wff = m->wff[J], wfb = m->wfb[J];
        const double cx2 = m->dyfi[J - 1] * h, cy2 = m->dyci[J] * h * h, cz2 = m->dyfi[J - 1] * h;
        for (int k = 0; k < NN; ++k) {
            const int kp = kp_(m, k), km = km_(m, k);
            for (int i = 0; i < LL; ++i) {
                const int ip = ip_(m, i), im = im_(m, i);
                /* x-momentum at (i', J, K) */
                double f = (V3(G, 0, ip, J, k) + V3(G, 0, i, J, k)) * (V3(Q, 0, ip, J, k) + V3(Q, 0, i, J, k));
                double b = (V3(G, 0, i, J, k) + V3(G, 0, im, J, k)) * (V3(Q, 0, i, J, k) + V3(Q, 0, im, J, k));
                const double h11 = (f - b) * cx1;
                f = (V3(G, 1, i, J + 1, k) + V3(G, 1, im, J + 1, k)) *
                    (wffp * V3(Q, 0, i, J + 1, k) + wfbp * V3(Q, 0, i, J, k));
                b = (V3(G, 1, i, J, k) + V3(G, 1, im, J, k)) * (wff * V3(Q, 0, i, J, k) + wfb * V3(Q, 0, i, J - 1, k));
                const double h12 = (f - b) * cx2;
                f = (V3(G, 2, i, J, kp) + V3(G, 2, im, J, kp)) * (V3(Q, 0, i, J, kp) + V3(Q, 0, i, J, k));
                b = (V3(G, 2, i, J, k) + V3(G, 2, im, J, k)) * (V3(Q, 0, i, J, k) + V3(Q, 0, i, J, km));
                const double h13 = (f - b) * cx3;
                CX(s->h[0], i, J, k) = -(h11 + h12 + h13);
                /* y-momentum at (i, j', K), j' >= 2 */
                if (J >= 2) {
                    f = (wff * V3(G, 0, ip, J, k) + wfb * V3(G, 0, ip, J - 1, k)) * (V3(Q, 1, ip, J, k) + V3(Q, 1, i, J, k));
                    b = (wff * V3(G, 0, i, J, k) + wfb * V3(G, 0, i, J - 1, k)) * (V3(Q, 1, i, J, k) + V3(Q, 1, im, J, k));
                    const double h21 = (f - b) * cy1;
                    f = (V3(G, 1, i, J + 1, k) + V3(G, 1, i, J, k)) * (V3(Q, 1, i, J + 1, k) + V3(Q, 1, i, J, k));
                    b = (V3(G, 1, i, J, k) + V3(G, 1, i, J - 1, k)) * (V3(Q, 1, i, J, k) + V3(Q, 1, i, J - 1, k));
                    const double h22 = (f - b) * cy2;
                    f = (wff * V3(G, 2, i, J, kp) + wfb * V3(G, 2, i, J - 1, kp)) * (V3(Q, 1, i, J, kp) + V3(Q, 1, i, J, k));
                    b = (wff * V3(G, 2, i, J, k) + wfb * V3(G, 2, i, J - 1, k)) * (V3(Q, 1, i, J, k) + V3(Q, 1, i, J, km));
                    const double h23 = (f - b) * cy30;
                    CX(s->h[1], i, J, k) = -(h21 + h22 + h23);
                } else {
                    CX(s->h[1], i, J, k) = 0.0;
                }
                /* z-momentum at (i, J, k') */
                f = (V3(G, 0, ip, J, k) + V3(G, 0, ip, J, km)) * (V3(Q, 2, ip, J, k) + V3(Q, 2, i, J, k));
                b = (V3(G, 0, i, J, k) + V3(G, 0, i, J, km)) * (V3(Q, 2, i, J, k) + V3(Q, 2, im, J, k));
                const double h31 = (f - b) * cz1;
                f = (V3(G, 1, i, J + 1, k) + V3(G, 1, i, J + 1, km)) *
                    (wffp * V3(Q, 2, i, J + 1, k) + wfbp * V3(Q, 2, i, J, k));
                b = (V3(G, 1, i, J, k) + V3(G, 1, i, J, km)) * (wff * V3(Q, 2, i, J, k) + wfb * V3(Q, 2, i, J - 1, k));
                const double h32 = (f - b) * cz2;
                f = (V3(G, 2, i, J, kp) + V3(G, 2, i, J, k)) * (V3(Q, 2, i, J, kp) + V3(Q, 2, i, J, k));
                b = (V3(G, 2, i, J, k) + V3(G, 2, i, J, km)) * (V3(Q, 2, i, J, k) + V3(Q, 2, i, J, km));
                const double h33 = (f - b) * cz30;
                CX(s->h[2], i, J, k) = -(h31 + h32 + h33);
            }
        }
    }
}

/* DIVG_U_io (DIVG.f90:140-246): div(u) / 3 at the cell centres. */
void or_io_divu(const or_mesh *m, or_io *s)
{
    for (int J = 1; J <= MM; ++J) {
        const double coe2 = m->dyfi[J - 1], coe3 = m->dzi;
        for (int k = 0; k < NN; ++k) {
            const int kp = kp_(m, k);
            for (int i = 0; i < LL; ++i) {
                const double dx = (V3(s->q, 0, ip_(m, i), J, k) - V3(s->q, 0, i, J, k)) * m->dxi;
                const double dy = (V3(s->q, 1, i, J + 1, k) - V3(s->q, 1, i, J, k)) * coe2;
                const double dz = (V3(s->q, 2, i, J, kp) - V3(s->q, 2, i, J, k)) * coe3;
                CX(s->divu, i, J, k) = (dx + dy + dz) / 3.0;
            }
        }
    }
}

/* VISCOUS_ALL_EXPLT_X_io / _Y_io / _Z_io, channel branch (VISCOUS_X_EXPLICIT_io.f90:13-128,
 * VISCOUS_Y_EXPLICIT_io.f90:14-120, VISCOUS_Z_EXPLICIT_io.f90:13-115): the divergence of the
 * full variable-viscosity stress, added to the convection terms. */
void or_io_viscous(const or_mesh *m, const or_io_par *p, or_io *s)
{
    const double cv = p->cvisc, dxi = m->dxi, dzi = m->dzi;
    const double *Q = s->q, *MU = s->mstg, *V = s->m0;
    for (int J = 1; J <= MM; ++J) {
        for (int k = 0; k < NN; ++k) {
            const int kp = kp_(m, k), km = km_(m, k);
            for (int i = 0; i < LL; ++i) {
                const int ip = ip_(m, i), im = im_(m, i);
                /* x: VISCOUS_X_EXPLICIT_io.f90:44-120 */
                {
                    const double coe1 = 2.0 * dxi * cv, coe2 = m->dyfi[J - 1] * cv, coe3 = dzi * cv;
                    double dudx = (V3(Q, 0, ip, J, k) - V3(Q, 0, i, J, k)) * dxi;
                    const double t11f = 0.5 * (HX(V, i, J, k) + HX(V, i, J, k)) * (dudx - CX(s->divu, i, J, k));
                    dudx = (V3(Q, 0, i, J, k) - V3(Q, 0, im, J, k)) * dxi;
                    const double t11b = 0.5 * (HX(V, im, J, k) + HX(V, im, J, k)) * (dudx - CX(s->divu, im, J, k));
                    const double d11 = (t11f - t11b) * coe1;
                    double dudy = (V3(Q, 0, i, J + 1, k) - V3(Q, 0, i, J, k)) * m->dyci[J + 1];
                    double dvdx = (V3(Q, 1, i, J + 1, k) - V3(Q, 1, im, J + 1, k)) * dxi;
                    const double t12f = V3(MU, 0, i, J + 1, k) * (dudy + dvdx);
                    dudy = (V3(Q, 0, i, J, k) - V3(Q, 0, i, J - 1, k)) * m->dyci[J];
                    dvdx = (V3(Q, 1, i, J, k) - V3(Q, 1, im, J, k)) * dxi;
                    const double t12b = V3(MU, 0, i, J, k) * (dudy + dvdx);
                    const double d12 = (t12f - t12b) * coe2;
                    double dudz = (V3(Q, 0, i, J, kp) - V3(Q, 0, i, J, k)) * dzi;
                    double dwdx = (V3(Q, 2, i, J, kp) - V3(Q, 2, im, J, kp)) * dxi;
                    const double t13f = V3(MU, 1, i, J, kp) * (dudz + dwdx);
                    dudz = (V3(Q, 0, i, J, k) - V3(Q, 0, i, J, km)) * dzi;
                    dwdx = (V3(Q, 2, i, J, k) - V3(Q, 2, im, J, k)) * dxi;
                    const double t13b = V3(MU, 1, i, J, k) * (dudz + dwdx);
                    const double d13 = (t13f - t13b) * coe3;
                    CX(s->h[0], i, J, k) = CX(s->h[0], i, J, k) + d11 + d12 + d13;
                }
                /* y (rows 2..M): VISCOUS_Y_EXPLICIT_io.f90:47-113 */
                if (J >= 2) {
                    const double coe1 = dxi * cv, coe2 = 2.0 * m->dyci[J] * cv, coe3 = dzi * cv;
                    double dvdy = (V3(Q, 1, i, J + 1, k) - V3(Q, 1, i, J, k)) * m->dyfi[J - 1];
                    const double t22f = 0.5 * (HX(V, i, J, k) + HX(V, i, J, k)) * (dvdy - CX(s->divu, i, J, k));
                    dvdy = (V3(Q, 1, i, J, k) - V3(Q, 1, i, J - 1, k)) * m->dyfi[J - 2];
                    const double t22b = 0.5 * (HX(V, i, J - 1, k) + HX(V, i, J - 1, k)) * (dvdy - CX(s->divu, i, J - 1, k));
                    const double d22 = (t22f - t22b) * coe2;
                    double dvdx = (V3(Q, 1, ip, J, k) - V3(Q, 1, i, J, k)) * dxi;
                    double dudy = (V3(Q, 0, ip, J, k) - V3(Q, 0, ip, J - 1, k)) * m->dyci[J];
                    const double t21f = V3(MU, 0, ip, J, k) * (dvdx + dudy);
                    dvdx = (V3(Q, 1, i, J, k) - V3(Q, 1, im, J, k)) * dxi;
                    dudy = (V3(Q, 0, i, J, k) - V3(Q, 0, i, J - 1, k)) * m->dyci[J];
                    const double t21b = V3(MU, 0, i, J, k) * (dvdx + dudy);
                    const double d21 = (t21f - t21b) * coe1;
                    double dvdz = (V3(Q, 1, i, J, kp) - V3(Q, 1, i, J, k)) * dzi;
                    double dwdy = (V3(Q, 2, i, J, kp) - V3(Q, 2, i, J - 1, kp)) * m->dyci[J];
                    const double t23f = V3(MU, 2, i, J, kp) * (dvdz + dwdy);
                    dvdz = (V3(Q, 1, i, J, k) - V3(Q, 1, i, J, km)) * dzi;
                    dwdy = (V3(Q, 2, i, J, k) - V3(Q, 2, i, J - 1, k)) * m->dyci[J];
                    const double t23b = V3(MU, 2, i, J, k) * (dvdz + dwdy);
                    const double d23 = (t23f - t23b) * coe3;
                    CX(s->h[1], i, J, k) = CX(s->h[1], i, J, k) + d21 + d22 + d23;
                }
                /* z: VISCOUS_Z_EXPLICIT_io.f90:44-107 */
                {
                    const double coe1 = dxi * cv, coe2 = m->dyfi[J - 1] * cv, coe3 = 2.0 * dzi * cv;
                    double dwdz = (V3(Q, 2, i, J, kp) - V3(Q, 2, i, J, k)) * dzi;
                    const double t33f = 0.5 * (HX(V, i, J, k) + HX(V, i, J, k)) * (dwdz - CX(s->divu, i, J, k));
                    dwdz = (V3(Q, 2, i, J, k) - V3(Q, 2, i, J, km)) * dzi;
                    const double t33b = 0.5 * (HX(V, i, J, km) + HX(V, i, J, km)) * (dwdz - CX(s->divu, i, J, km));
                    const double d33 = (t33f - t33b) * coe3;
                    double dwdx = (V3(Q, 2, ip, J, k) - V3(Q, 2, i, J, k)) * dxi;
                    double dudz = (V3(Q, 0, ip, J, k) - V3(Q, 0, ip, J, km)) * dzi;
                    const double t31f = V3(MU, 1, ip, J, k) * (dwdx + dudz);
                    dwdx = (V3(Q, 2, i, J, k) - V3(Q, 2, im, J, k)) * dxi;
                    dudz = (V3(Q, 0, i, J, k) - V3(Q, 0, i, J, km)) * dzi;
                    const double t31b = V3(MU, 1, i, J, k) * (dwdx + dudz);
                    const double d31 = (t31f - t31b) * coe1;
                    double dwdy = (V3(Q, 2, i, J + 1, k) - V3(Q, 2, i, J, k)) * m->dyci[J + 1];
                    double dvdz = (V3(Q, 1, i, J + 1, k) - V3(Q, 1, i, J + 1, km)) * dzi;
                    const double t32f = V3(MU, 2, i, J + 1, k) * (dwdy + dvdz);
                    dwdy = (V3(Q, 2, i, J, k) - V3(Q, 2, i, J - 1, k)) * m->dyci[J];
                    dvdz = (V3(Q, 1, i, J, k) - V3(Q, 1, i, J, km)) * dzi;
                    const double t32b = V3(MU, 2, i, J, k) * (dwdy + dvdz);
                    const double d32 = (t32f - t32b) * coe2;
                    CX(s->h[2], i, J, k) = CX(s->h[2], i, J, k) + d31 + d32 + d33;
                }
            }
        }
    }
}

/* RHS_MOM_EXPLICIT_io(NS, IDR), explicit branch (RHS_MOM_EXPLICIT_io.f90:11-304): RK
 * combination with the previous substep's terms, pressure gradient, buoyancy, and for the
 * streamwise component the mass-flux (or given-Cf) correction; the result replaces the
 * momentum term (the increment of G). */
void or_io_rhs(const or_mesh *m, const or_io_par *p, or_io *s, int ns, int idr, double dt)
{
    const int c = idr - 1, j0 = idr == 2 ? 2 : 1;
    const int cn = p->vis_scheme == 1;
    const double coe2i = m->talp[ns] * dt;
    double *H = cn ? s->rim : s->h[c];
    for (int k = 0; k < NN; ++k)
        for (int J = j0; J <= MM; ++J)
            for (int i = 0; i < LL; ++i) {
                const double rc = CX(s->h[c], i, J, k), rl = C3(s->explt0, c, i, J, k);
                if (cn)   /* RHS_MOM_EXPLICIT_io.f90:87-99; the DIVU_io term is zero (or_io_viscous_cn) */
                    CX(H, i, J, k) = (m->tgam[ns] * rc + m->troh[ns] * rl) * dt + coe2i * CX(s->rim, i, J, k);
                else
                    CX(H, i, J, k) = (m->tgam[ns] * rc + m->troh[ns] * rl) * dt;
                C3(s->explt0, c, i, J, k) = rc;
            }
    for (int k = 0; k < NN; ++k)
        for (int J = j0; J <= MM; ++J) {
            double coe2;
            if (idr == 1) coe2 = m->talp[ns] * dt * m->dxi;
            else if (idr == 2) coe2 = m->talp[ns] * dt * m->dyci[J];
            else coe2 = m->talp[ns] * dt * m->dzi;
            for (int i = 0; i < LL; ++i) {
                double pgm;
                if (idr == 1) pgm = (HX(s->pr, i, J, k) - HX(s->pr, im_(m, i), J, k)) * coe2;
                else if (idr == 2) pgm = (HX(s->pr, i, J, k) - HX(s->pr, i, J - 1, k)) * coe2;
                else pgm = (HX(s->pr, i, J, k) - HX(s->pr, i, J, km_(m, k))) * coe2;
                CX(H, i, J, k) = CX(H, i, J, k) - pgm;
            }
        }
    if (p->gravity == idr) {
        const double coe2 = m->talp[ns] * dt * p->f_a;
        for (int k = 0; k < NN; ++k)
            for (int J = j0; J <= MM; ++J)
                for (int i = 0; i < LL; ++i) CX(H, i, J, k) = CX(H, i, J, k) + V3(s->dstg, c, i, J, k) * coe2;
    }
    if (idr == 1) {   /* NFLOW = 1 (modules.f90:45) */
        double dpgrns = 0.0;
        if (p->flow_driven == 1) {
            double sum = 0.0;
            for (int J = 1; J <= MM; ++J)
                for (int i = 0; i < LL; ++i)
                    for (int k = 0; k < NN; ++k) sum = sum + CX(H, i, J, k) / m->dyfi[J - 1];
            dpgrns = sum / m->volm;
        } else if (p->flow_driven == 2) {
            dpgrns = -0.5 * p->cf_given * (m->talp[ns] * dt);
        }
        for (int k = 0; k < NN; ++k)
            for (int J = 1; J <= MM; ++J)
                for (int i = 0; i < LL; ++i) CX(H, i, J, k) = CX(H, i, J, k) - dpgrns;
    }
}

/* MASSFLUX_CALC_io (MASSFLUX_CALC_IO.f90:15-52): G += increments; then the walls. */
void or_io_massflux_calc(const or_mesh *m, or_io *s)
{
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < NN; ++k)
            for (int J = (c == 1 ? 2 : 1); J <= MM; ++J)
                for (int i = 0; i < LL; ++i) V3(s->g, c, i, J, k) = V3(s->g, c, i, J, k) + CX(s->h[c], i, J, k);
    or_io_bc_zero(m, s->g);
}

/* DIVG_io (DIVG.f90:54-137), iWeightedPre = 0: (div G + d rho/dt) / (alpha dt). */
void or_io_divg(const or_mesh *m, or_io *s, int ns, double dt)
{
    const double coe0 = 1.0 / dt / m->talp[ns];
    const double coe4 = 1.0 / dt;
    (void)coe4;
    for (int J = 1; J <= MM; ++J) {
        const double coe1 = m->dxi, coe2 = m->dyfi[J - 1];
        for (int k = 0; k < NN; ++k) {
            const int kp = kp_(m, k);
            for (int i = 0; i < LL; ++i) {
                const double divx = (V3(s->g, 0, ip_(m, i), J, k) - V3(s->g, 0, i, J, k)) * coe1;
                const double divy = (V3(s->g, 1, i, J + 1, k) - V3(s->g, 1, i, J, k)) * coe2;
                const double divz = (V3(s->g, 2, i, J, kp) - V3(s->g, 2, i, J, k)) * m->dzi;
                const double dq = divx + divy + divz + CX(s->drho, i, J, k);
                CX(s->rhsp, i, J, k) = dq * coe0;
            }
        }
    }
}

/* PRCALC_io (PRESSURE_CORRECTION.f90:102-170; with iVisScheme = 1 the viscous term of the
 * pressure, RHSLLPHI_io being the Poisson right-hand side) and MASSFLUX_UPDATE_io
 * (VELO_CORRECTION.f90:103-179): p += phi; G -= alpha dt grad(phi); walls. */
void or_io_correct(const or_mesh *m, const or_io_par *p, or_io *s, int ns, double dt)
{
    const double be = 0.5 * m->talp[ns] * dt * p->cvisc;   /* PRESSURE_CORRECTION.f90:115-122 */
    for (int k = 0; k < NN; ++k)
        for (int J = 1; J <= MM; ++J)
            for (int i = 0; i < LL; ++i) {
                if (p->vis_scheme == 1)
                    HX(s->pr, i, J, k) = HX(s->pr, i, J, k) + HX(s->dph, i, J, k) - be * CX(s->rhsp, i, J, k);
                else
                    HX(s->pr, i, J, k) = HX(s->pr, i, J, k) + HX(s->dph, i, J, k);
            }
    or_bc_copy(m, s->pr);
    double coe1 = dt * m->talp[ns] * m->dxi;
    for (int k = 0; k < NN; ++k)
        for (int J = 1; J <= MM; ++J)
            for (int i = 0; i < LL; ++i)
                V3(s->g, 0, i, J, k) = V3(s->g, 0, i, J, k) - (HX(s->dph, i, J, k) - HX(s->dph, im_(m, i), J, k)) * coe1;
    coe1 = dt * m->talp[ns] * m->dzi;
    for (int k = 0; k < NN; ++k)
        for (int J = 1; J <= MM; ++J)
            for (int i = 0; i < LL; ++i)
                V3(s->g, 2, i, J, k) = V3(s->g, 2, i, J, k) - (HX(s->dph, i, J, k) - HX(s->dph, i, J, km_(m, k))) * coe1;
    for (int J = 2; J <= MM; ++J) {
        coe1 = dt * m->talp[ns] * m->dyci[J];
        for (int k = 0; k < NN; ++k)
            for (int i = 0; i < LL; ++i)
                V3(s->g, 1, i, J, k) = V3(s->g, 1, i, J, k) - (HX(s->dph, i, J, k) - HX(s->dph, i, J - 1, k)) * coe1;
    }
    or_io_bc_zero(m, s->g);
}

/* VISCOUS_PAR_EXPLT_X_io (VISCOUS_X_EXPLICIT_io.f90:131-293), _Y_io (VISCOUS_Y_EXPLICIT_io.f90:
 * 278-434, channel branch), _Z_io (VISCOUS_Z_EXPLICIT_io.f90:300-462, channel branch) for
 * component idr: the viscous stress split after Nicoud (2000).  The part left explicit is
 * added in front of the convection term (QTMP_io / DPH_io / RHSLLPHI_io = EX + conv); the part
 * the line solves treat implicitly goes to RHS_io (rim).  Both read the current viscosity
 * (Viscousity) and MU_STG.  The density-correction parts (the *_AD terms into DIVU_io) are
 * G * DRHOI_STG products, and DRHOI_STG = 1/D_STG - 1/DEN_STG where DENSITY_IMPLICIT_add
 * (MASSFLUX_CALC_IO.f90:332-416) evaluates DEN_STG by D_STG's own expression of the same
 * DENSITY0 (:276-309), right after DENSITY_Staggered (SOLVE.f90:123-126): it is exactly
 * zero, so those parts are +-0 and add nothing. */
void or_io_viscous_cn(const or_mesh *m, const or_io_par *p, or_io *s, int idr)
{
    const double cv = p->cvisc, dxi = m->dxi, dzi = m->dzi;
    const double *Q = s->q, *MU = s->mstg, *V = s->m1;
    const int c = idr - 1;
    memset(s->rim, 0, sizeof(double) * (size_t)NN * MM * LL);   /* RHS_io = 0 */
    for (int J = (idr == 2 ? 2 : 1); J <= MM; ++J) {
        for (int k = 0; k < NN; ++k) {
            const int kp = kp_(m, k), km = km_(m, k);
            for (int i = 0; i < LL; ++i) {
                const int ip = ip_(m, i), im = im_(m, i);
                double ex, imp;
                if (idr == 1) {
                    const double coe1 = -2.0 / 3.0 * dxi * cv, coe2 = m->dyfi[J - 1] * cv, coe3 = dzi * cv;
                    double dudx = (V3(Q, 0, ip, J, k) - V3(Q, 0, i, J, k)) * dxi;
                    double dvdy = (V3(Q, 1, i, J + 1, k) - V3(Q, 1, i, J, k)) * m->dyfi[J - 1];
                    double dwdz = (V3(Q, 2, i, J, kp) - V3(Q, 2, i, J, k)) * dzi;
                    const double t11fe = HX(V, i, J, k) * (dvdy + dwdz), t11fi = HX(V, i, J, k) * dudx;
                    dudx = (V3(Q, 0, i, J, k) - V3(Q, 0, im, J, k)) * dxi;
                    dvdy = (V3(Q, 1, im, J + 1, k) - V3(Q, 1, im, J, k)) * m->dyfi[J - 1];
                    dwdz = (V3(Q, 2, im, J, kp) - V3(Q, 2, im, J, k)) * dzi;
                    const double t11be = HX(V, im, J, k) * (dvdy + dwdz), t11bi = HX(V, im, J, k) * dudx;
                    const double d11e = (t11fe - t11be) * coe1, d11i = (t11fi - t11bi) * coe1 * (-2.0);
                    double vis = V3(MU, 0, i, J + 1, k);
                    double dvdx = (V3(Q, 1, i, J + 1, k) - V3(Q, 1, im, J + 1, k)) * dxi;
                    double dudy = (V3(Q, 0, i, J + 1, k) - V3(Q, 0, i, J, k)) * m->dyci[J + 1];
                    const double t12fe = vis * dvdx, t12fi = vis * dudy;
                    vis = V3(MU, 0, i, J, k);
                    dudy = (V3(Q, 0, i, J, k) - V3(Q, 0, i, J - 1, k)) * m->dyci[J];
                    dvdx = (V3(Q, 1, i, J, k) - V3(Q, 1, im, J, k)) * dxi;
                    const double t12be = vis * dvdx, t12bi = vis * dudy;
                    const double d12e = (t12fe - t12be) * coe2, d12i = (t12fi - t12bi) * coe2;
                    vis = V3(MU, 1, i, J, kp);
                    double dudz = (V3(Q, 0, i, J, kp) - V3(Q, 0, i, J, k)) * dzi;
                    double dwdx = (V3(Q, 2, i, J, kp) - V3(Q, 2, im, J, kp)) * dxi;
                    const double t13fe = vis * dwdx, t13fi = vis * dudz;
                    vis = V3(MU, 1, i, J, k);
                    dudz = (V3(Q, 0, i, J, k) - V3(Q, 0, i, J, km)) * dzi;
                    dwdx = (V3(Q, 2, i, J, k) - V3(Q, 2, im, J, k)) * dxi;
                    const double t13be = vis * dwdx, t13bi = vis * dudz;
                    const double d13e = (t13fe - t13be) * coe3, d13i = (t13fi - t13bi) * coe3;
                    ex = d11e + d12e + d13e;
                    imp = d11i + d12i + d13i;
                } else if (idr == 2) {
                    const double coe1 = dxi * cv, coe3 = dzi * cv, coe2 = -2.0 / 3.0 * m->dyci[J] * cv;
                    double dudx = (V3(Q, 0, ip, J, k) - V3(Q, 0, i, J, k)) * dxi;
                    double dvdy = (V3(Q, 1, i, J + 1, k) - V3(Q, 1, i, J, k)) * m->dyfi[J - 1];
                    double dwdz = (V3(Q, 2, i, J, kp) - V3(Q, 2, i, J, k)) * dzi;
                    const double t22fe = HX(V, i, J, k) * (dudx + dwdz), t22fi = HX(V, i, J, k) * dvdy;
                    dudx = (V3(Q, 0, ip, J - 1, k) - V3(Q, 0, i, J - 1, k)) * dxi;
                    dvdy = (V3(Q, 1, i, J, k) - V3(Q, 1, i, J - 1, k)) * m->dyfi[J - 2];
                    dwdz = (V3(Q, 2, i, J - 1, kp) - V3(Q, 2, i, J - 1, k)) * dzi;
                    const double t22be = HX(V, i, J - 1, k) * (dudx + dwdz), t22bi = HX(V, i, J - 1, k) * dvdy;
                    const double d22e = (t22fe - t22be) * coe2, d22i = (t22fi - t22bi) * coe2 * (-2.0);
                    double vis = V3(MU, 0, ip, J, k);
                    double dvdx = (V3(Q, 1, ip, J, k) - V3(Q, 1, i, J, k)) * dxi;
                    double dudy = (V3(Q, 0, ip, J, k) - V3(Q, 0, ip, J - 1, k)) * m->dyci[J];
                    const double t21fe = vis * dudy, t21fi = vis * dvdx;
                    vis = V3(MU, 0, i, J, k);
                    dvdx = (V3(Q, 1, i, J, k) - V3(Q, 1, im, J, k)) * dxi;
                    dudy = (V3(Q, 0, i, J, k) - V3(Q, 0, i, J - 1, k)) * m->dyci[J];
                    const double t21be = vis * dudy, t21bi = vis * dvdx;
                    const double d21e = (t21fe - t21be) * coe1, d21i = (t21fi - t21bi) * coe1;
                    vis = V3(MU, 2, i, J, kp);
                    double dvdz = (V3(Q, 1, i, J, kp) - V3(Q, 1, i, J, k)) * dzi;
                    double dwdy = (V3(Q, 2, i, J, kp) - V3(Q, 2, i, J - 1, kp)) * m->dyci[J];
                    const double t23fe = vis * dwdy, t23fi = vis * dvdz;
                    vis = V3(MU, 2, i, J, k);
                    dvdz = (V3(Q, 1, i, J, k) - V3(Q, 1, i, J, km)) * dzi;
                    dwdy = (V3(Q, 2, i, J, k) - V3(Q, 2, i, J - 1, k)) * m->dyci[J];
                    const double t23be = vis * dwdy, t23bi = vis * dvdz;
                    const double d23e = (t23fe - t23be) * coe3, d23i = (t23fi - t23bi) * coe3;
                    ex = d21e + d22e + d23e;
                    imp = d21i + d22i + d23i;
                } else {
                    const double coe1 = dxi * cv, coe3 = -2.0 / 3.0 * dzi * cv, coe2 = m->dyfi[J - 1] * cv;
                    double dudx = (V3(Q, 0, ip, J, k) - V3(Q, 0, i, J, k)) * dxi;
                    double dvdy = (V3(Q, 1, i, J + 1, k) - V3(Q, 1, i, J, k)) * m->dyfi[J - 1];
                    double dwdz = (V3(Q, 2, i, J, kp) - V3(Q, 2, i, J, k)) * dzi;
                    const double t33fe = HX(V, i, J, k) * (dudx + dvdy), t33fi = HX(V, i, J, k) * dwdz;
                    dudx = (V3(Q, 0, ip, J, km) - V3(Q, 0, i, J, km)) * dxi;
                    dvdy = (V3(Q, 1, i, J + 1, km) - V3(Q, 1, i, J, km)) * m->dyfi[J - 1];
                    dwdz = (V3(Q, 2, i, J, k) - V3(Q, 2, i, J, km)) * dzi;
                    const double t33be = HX(V, i, J, km) * (dudx + dvdy), t33bi = HX(V, i, J, km) * dwdz;
                    const double d33e = (t33fe - t33be) * coe3, d33i = (t33fi - t33bi) * coe3 * (-2.0);
                    double vis = V3(MU, 1, ip, J, k);
                    double dwdx = (V3(Q, 2, ip, J, k) - V3(Q, 2, i, J, k)) * dxi;
                    double dudz = (V3(Q, 0, ip, J, k) - V3(Q, 0, ip, J, km)) * dzi;
                    const double t31fe = vis * dudz, t31fi = vis * dwdx;
                    vis = V3(MU, 1, i, J, k);
                    dwdx = (V3(Q, 2, i, J, k) - V3(Q, 2, im, J, k)) * dxi;
                    dudz = (V3(Q, 0, i, J, k) - V3(Q, 0, i, J, km)) * dzi;
                    const double t31be = vis * dudz, t31bi = vis * dwdx;
                    const double d31e = (t31fe - t31be) * coe1, d31i = (t31fi - t31bi) * coe1;
                    vis = V3(MU, 2, i, J + 1, k);
                    double dwdy = (V3(Q, 2, i, J + 1, k) - V3(Q, 2, i, J, k)) * m->dyci[J + 1];
                    double dvdz = (V3(Q, 1, i, J + 1, k) - V3(Q, 1, i, J + 1, km)) * dzi;
                    const double t32fe = vis * dvdz, t32fi = vis * dwdy;
                    vis = V3(MU, 2, i, J, k);
                    dwdy = (V3(Q, 2, i, J, k) - V3(Q, 2, i, J - 1, k)) * m->dyci[J];
                    dvdz = (V3(Q, 1, i, J, k) - V3(Q, 1, i, J, km)) * dzi;
                    const double t32be = vis * dvdz, t32bi = vis * dwdy;
                    const double d32e = (t32fe - t32be) * coe2, d32i = (t32fi - t32bi) * coe2;
                    ex = d31e + d32e + d33e;
                    imp = d31i + d32i + d33i;
                }
                CX(s->h[c], i, J, k) = ex + CX(s->h[c], i, J, k);
                CX(s->rim, i, J, k) = imp;
            }
        }
    }
}

/* The Crank-Nicolson wall coefficients of TDMA_COEF (TDMA_COEF.f90:76-87): AMVR1 at the
 * bottom wall, APVRN at the top wall of the u and w y-stencils. */
static void or_cn_wall_coefs(const or_mesh *m, double *amvr1, double *apvrn)
{
    double ucaj = 4.0 / (2.0 / m->dyci[2] + 1.0 / m->dyfi[0]);
    *amvr1 = ucaj * (2.0 * m->dyfi[0]);
    ucaj = 4.0 / (2.0 / m->dyci[MM] + 1.0 / m->dyfi[MM - 1]);
    *apvrn = ucaj * (2.0 * m->dyfi[MM - 1]);
}

/* MOMFA_io(NS, IDR) (MOMFA_io.f90:18-46): the variable-coefficient ADI factors
 * (1 - Ax)(1 - Az)(1 - Ay) dG = RHS_io with FACOE = alpha dt / (2 Re) and coefficients
 * mu / rho_stg per point, MOMFA1_X_io (:77-176, cyclic TDMAIJI_CYC per z-plane over rows
 * N2I..M), MOMFA2_Z_io (:179-265, cyclic along z per x-plane), MOMFA3_Y_io (:268-406,
 * TDMAIJJ_nonCYC along y per z-plane, zero boundary values), then G += RHS_io for rows
 * N2I..M.  The y-solve is the one-rank form: the wall viscosities of the u, w lines come from
 * the line's own z-plane (MOMFA_io.f90:342-352). */
void or_io_momfa(const or_mesh *m, const or_io_par *p, or_io *s, int ns, int idr, double dt)
{
    const int c = idr - 1, n2i = idr == 2 ? 2 : 1, nj = MM - n2i + 1;
    const double facoe = 0.5 * m->talp[ns] * dt * p->cvisc;
    const double *V = s->m1, *MU = s->mstg, *DS = s->dstg;
    double *R = s->rim;
    const size_t nmax = (size_t)(LL > NN ? LL : NN) * (MM > nj ? MM : nj);
    double *A = (double *)malloc(sizeof(double) * 4 * nmax + sizeof(double) * 2 * LL);
    double *B = A + nmax, *Cc = A + 2 * nmax, *F = A + 3 * nmax, *bcz = A + 4 * nmax;
    /* x */
    {
        const double coe1 = facoe * 4.0 / 3.0 * m->dxqi, coe2 = facoe * m->dxqi;
        for (int k = 0; k < NN; ++k) {
            for (int J = n2i; J <= MM; ++J)
                for (int i = 0; i < LL; ++i) {
                    const int im = im_(m, i), ip = ip_(m, i);
                    double amu, apu;
                    if (idr == 1) { amu = -coe1 * HX(V, im, J, k); apu = -coe1 * HX(V, i, J, k); }
                    else if (idr == 2) { amu = -coe2 * V3(MU, 0, i, J, k); apu = -coe2 * V3(MU, 0, ip, J, k); }
                    else { amu = -coe2 * V3(MU, 1, i, J, k); apu = -coe2 * V3(MU, 1, ip, J, k); }
                    const size_t o = (size_t)(J - n2i) * LL + i;
                    A[o] = amu / V3(DS, c, im, J, k);
                    Cc[o] = apu / V3(DS, c, ip, J, k);
                    B[o] = 1.0 - (amu + apu) / V3(DS, c, i, J, k);
                    F[o] = CX(R, i, J, k);
                }
            or_tdma_lines(0, LL, nj, A, B, Cc, F, NULL);
            for (int J = n2i; J <= MM; ++J)
                for (int i = 0; i < LL; ++i) CX(R, i, J, k) = F[(size_t)(J - n2i) * LL + i];
        }
    }
    /* z (RCCI2 = RNDI2 = 1 on the channel) */
    {
        const double coe1 = facoe * 4.0 / 3.0 * m->dzqi, coe2 = facoe * m->dzqi;
        for (int i = 0; i < LL; ++i) {
            for (int J = n2i; J <= MM; ++J)
                for (int k = 0; k < NN; ++k) {
                    const int km = km_(m, k), kp = kp_(m, k);
                    double amu, apu;
                    if (idr == 1) { amu = -coe2 * V3(MU, 1, i, J, k); apu = -coe2 * V3(MU, 1, i, J, kp); }
                    else if (idr == 2) { amu = -coe2 * V3(MU, 2, i, J, k); apu = -coe2 * V3(MU, 2, i, J, kp); }
                    else { amu = -coe1 * HX(V, i, J, km); apu = -coe1 * HX(V, i, J, k); }
                    const size_t o = (size_t)(J - n2i) * NN + k;
                    A[o] = amu / V3(DS, c, i, J, km);
                    Cc[o] = apu / V3(DS, c, i, J, kp);
                    B[o] = 1.0 - (amu + apu) / V3(DS, c, i, J, k);
                    F[o] = CX(R, i, J, k);
                }
            or_tdma_lines(0, NN, nj, A, B, Cc, F, NULL);
            for (int J = n2i; J <= MM; ++J)
                for (int k = 0; k < NN; ++k) CX(R, i, J, k) = F[(size_t)(J - n2i) * NN + k];
        }
    }
    /* y, rows 1..M of every line */
    {
        double amvr1, apvrn;
        or_cn_wall_coefs(m, &amvr1, &apvrn);
        const double coe1 = facoe * 4.0 / 3.0, coe2 = facoe;
        const double *am = m->amvr + (size_t)c * (MM + 2), *ap = m->apvr + (size_t)c * (MM + 2);
        memset(bcz, 0, sizeof(double) * 2 * LL);
        for (int k = 0; k < NN; ++k) {
            for (int J = 1; J <= MM; ++J)
                for (int i = 0; i < LL; ++i) {
                    const size_t o = (size_t)(J - 1) * LL + i;
                    double amu, apu, a, cc, b;
                    if (idr != 2) {
                        const double *FM = MU + (size_t)c * NN * (MM + 2) * LL;
                        if (J == 1) {
                            const double muqt = (HX(V, i, 1, k) + HX(V, i, 0, k)) * 0.5;
                            amu = -coe2 * amvr1 * muqt;
                            apu = -coe2 * ap[J] * HX(FM, i, J + 1, k);
                            a = 0.0;
                            cc = apu / V3(DS, c, i, J + 1, k);
                        } else if (J == MM) {
                            const double muqt = (HX(V, i, MM, k) + HX(V, i, MM + 1, k)) * 0.5;
                            amu = -coe2 * am[J] * HX(FM, i, J, k);
                            apu = -coe2 * apvrn * muqt;
                            a = amu / V3(DS, c, i, J - 1, k);
                            cc = 0.0;
                        } else {
                            amu = -coe2 * am[J] * HX(FM, i, J, k);
                            apu = -coe2 * ap[J] * HX(FM, i, J + 1, k);
                            a = amu / V3(DS, c, i, J - 1, k);
                            cc = apu / V3(DS, c, i, J + 1, k);
                        }
                        b = 1.0 - (amu + apu) / V3(DS, c, i, J, k);
                    } else {
                        amu = -coe1 * am[J] * HX(V, i, J - 1, k);
                        apu = -coe1 * ap[J] * HX(V, i, J, k);
                        if (J == 1) {   /* AMVR(1, 2) = 0: AMU = 0, whatever D_STG holds below the wall */
                            a = 0.0;
                            cc = apu / V3(DS, 1, i, J + 1, k);
                        } else {
                            a = amu / V3(DS, 1, i, J - 1, k);
                            cc = apu / V3(DS, 1, i, J, k);
                        }
                        b = 1.0 - (amu + apu) / V3(DS, 1, i, J, k);
                    }
                    A[o] = a;
                    Cc[o] = cc;
                    B[o] = b;
                    F[o] = CX(R, i, J, k);
                }
            or_tdma_lines(1, LL, MM, A, B, Cc, F, bcz);
            for (int J = 1; J <= MM; ++J)
                for (int i = 0; i < LL; ++i) CX(R, i, J, k) = F[(size_t)(J - 1) * LL + i];
        }
    }
    free(A);
    for (int k = 0; k < NN; ++k)
        for (int J = n2i; J <= MM; ++J)
            for (int i = 0; i < LL; ++i) V3(s->g, c, i, J, k) = CX(R, i, J, k) + V3(s->g, c, i, J, k);
}

/* What SOLVE does between the energy and the momentum substep of stage NS in a heated run
 * (SOLVE.f90:122-124). */
void or_io_props(const or_mesh *m, or_io *s, int ns)
{
    or_io_drhodt(m, s, ns);
    or_io_dstg(m, s);
    or_io_mustg(m, s);
}

/* The momentum part of SOLVERRK3_MOM_io(NS) with iVisScheme = 1 (SOLVERRK3_MOM_IO.f90:25-40,
 * :61-74): per component the split viscous terms, RHS_MOM_EXPLICIT_io and MOMFA_io; then the
 * walls of G (BC_WALL_G_io). */
void or_io_momentum_cn(const or_mesh *m, const or_io_par *p, or_io *s, int ns, double dt)
{
    const size_t hs = (size_t)NN * (MM + 2) * LL;
    memcpy(s->g0, s->g, sizeof(double) * 3 * hs);
    or_io_velocity(m, s);
    or_io_convection(m, s);
    for (int idr = 1; idr <= 3; ++idr) {
        or_io_viscous_cn(m, p, s, idr);
        or_io_rhs(m, p, s, ns, idr, dt);
        if (s->rkeep)
            memcpy(s->rkeep + (size_t)(idr - 1) * NN * MM * LL, s->rim, sizeof(double) * (size_t)NN * MM * LL);
        or_io_momfa(m, p, s, ns, idr, dt);
    }
    or_io_bc_zero(m, s->g);
}

/* SOLVERRK3_MOM_io(NS) (SOLVERRK3_MOM_IO.f90:11-120), either viscous scheme. */
void or_io_substep(const or_mesh *m, const or_io_par *p, or_io *s, int ns, double dt)
{
    const size_t hs = (size_t)NN * (MM + 2) * LL;
    if (p->vis_scheme == 1) {
        or_io_momentum_cn(m, p, s, ns, dt);
    } else {
        memcpy(s->g0, s->g, sizeof(double) * 3 * hs);
        or_io_velocity(m, s);
        or_io_convection(m, s);
        or_io_divu(m, s);
        or_io_viscous(m, p, s);
        for (int idr = 1; idr <= 3; ++idr) or_io_rhs(m, p, s, ns, idr, dt);
        or_io_massflux_calc(m, s);
    }
    or_io_divg(m, s, ns, dt);
    memset(s->dph, 0, sizeof(double) * hs);
    or_poisson(m, s->rhsp, s->dph);
    or_bc_copy(m, s->dph);
    or_io_correct(m, p, s, ns, dt);
    or_io_velocity(m, s);
}
#undef V3
#undef C3

/* ======================================================================================
 * x-z mean statistics of the TG field: PP_MEAN_ZX_FLOW_TG (processing/PP_MEAN_ZX.f90:14-307),
 * channel.  Per row J (out[(J-1)*122 + s]): s = 0..3 U1 (u, v, w, p), 4..6 UP, 7..12 U2
 * (upper triangle, the reference's storage index), 13..22 U3, 23..31 DVDL1 (3 x 3, row
 * major), 32..40 DVDLP, 41..121 DVDL2 (9 x 9, row major).  scal[0] = U1mean_WORK_tg
 * (its / DZI * NCL3 / Area_inlet form), scal[1] = U1maxx_WORK_tg.  Summation order I, then K,
 * as the reference.
 * ====================================================================================== */
void or_mean_zx(const or_mesh *m, const or_state *s, double *out, double *scal)
{
    const double vl = 1.0 / (double)(LL * NN);
    const double coe0 = vl, coe1 = vl * 0.5, coe2 = vl * (0.5 * 0.5), coe3 = vl * (0.5 * 0.5 * 0.5);
    double cog1[3][3] = {{vl * m->dxi, vl * 0.25, vl * 0.25 * m->dzi},
                         {vl * 0.25 * m->dxi, vl, vl * 0.25 * m->dzi},
                         {vl * 0.25 * m->dxi, vl * 0.25, vl * m->dzi}};
    double u1mean = 0.0, u1max = 0.0;
    for (int J = 1; J <= MM; ++J) {
        double *o = out + (size_t)(J - 1) * 122;
        for (int t = 0; t < 122; ++t) o[t] = 0.0;
        for (int i = 0; i < LL; ++i) {
            const int ip = ip_(m, i), im = im_(m, i);
            for (int k = 0; k < NN; ++k) {
                const int kp = kp_(m, k), km = km_(m, k);
                double u[3], ujp[3], ujm[3], ukp[3], ukm[3], uip[3], uim[3];
                u[0] = QC(0, i, J, k) + QC(0, ip, J, k);
                ujp[0] = QC(0, i, J + 1, k) + QC(0, ip, J + 1, k);
                ujm[0] = QC(0, i, J - 1, k) + QC(0, ip, J - 1, k);
                ukp[0] = QC(0, i, J, kp) + QC(0, ip, J, kp);
                ukm[0] = QC(0, i, J, km) + QC(0, ip, J, km);
                u[2] = QC(2, i, J, k) + QC(2, i, J, kp);
                uip[2] = QC(2, ip, J, k) + QC(2, ip, J, kp);
                uim[2] = QC(2, im, J, k) + QC(2, im, J, kp);
                ujp[2] = QC(2, i, J + 1, k) + QC(2, i, J + 1, kp);
                ujm[2] = QC(2, i, J - 1, k) + QC(2, i, J - 1, kp);
                u[1] = QC(1, i, J, k) + QC(1, i, J + 1, k);
                uip[1] = QC(1, ip, J, k) + QC(1, ip, J + 1, k);
                uim[1] = QC(1, im, J, k) + QC(1, im, J + 1, k);
                ukp[1] = QC(1, i, J, kp) + QC(1, i, J + 1, kp);
                ukm[1] = QC(1, i, J, km) + QC(1, i, J + 1, km);
                const double p = HX(s->pr, i, J, k);
                for (int a = 0; a < 3; ++a) {
                    o[a] = o[a] + u[a];
                    o[4 + a] = o[4 + a] + u[a] * p;
                }
                o[3] = o[3] + p;
                for (int a = 0; a < 3; ++a)
                    for (int b = a; b < 3; ++b) {
                        const int l = ((a + 1) * (6 - a)) / 2 + (b + 1) - 3;   /* 1-based */
                        o[7 + l - 1] = o[7 + l - 1] + u[a] * u[b];
                    }
                for (int a = 0; a < 3; ++a)
                    for (int b = a; b < 3; ++b)
                        for (int c = b; c < 3; ++c) {
                            const int l = (a + 1) * (5 - a) + ((b + 1) * (6 - b)) / 2 + (c + 1) - 8;
                            o[13 + l - 1] = o[13 + l - 1] + u[a] * u[b] * u[c];
                        }
                double g[3][3];
                g[0][0] = QC(0, ip, J, k) - QC(0, i, J, k);
                g[0][1] = (ujp[0] - u[0]) * m->dyci[J + 1] + (u[0] - ujm[0]) * m->dyci[J];
                g[0][2] = ukp[0] - ukm[0];
                g[1][0] = uip[1] - uim[1];
                g[1][1] = (QC(1, i, J + 1, k) - QC(1, i, J, k)) * m->dyfi[J - 1];
                g[1][2] = ukp[1] - ukm[1];
                g[2][0] = uip[2] - uim[2];
                g[2][1] = (ujp[2] - u[2]) * m->dyci[J + 1] + (u[2] - ujm[2]) * m->dyci[J];
                g[2][2] = QC(2, i, J, kp) - QC(2, i, J, k);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        o[23 + 3 * a + b] = o[23 + 3 * a + b] + g[a][b];
                        o[32 + 3 * a + b] = o[32 + 3 * a + b] + g[a][b] * p;
                    }
                for (int a = 0; a < 3; ++a)
                    for (int h = 0; h < 3; ++h)
                        for (int b = 0; b < 3; ++b)
                            for (int c = 0; c < 3; ++c) {
                                const int l1 = 3 * a + h, l2 = 3 * b + c;
                                o[41 + 9 * l1 + l2] = o[41 + 9 * l1 + l2] + g[a][h] * g[b][c];
                            }
            }
        }
        o[3] = o[3] * coe0;
        for (int a = 0; a < 3; ++a) o[a] = o[a] * coe1;
        for (int a = 0; a < 3; ++a) o[4 + a] = o[4 + a] * coe1;
        for (int t = 0; t < 6; ++t) o[7 + t] = o[7 + t] * coe2;
        for (int t = 0; t < 10; ++t) o[13 + t] = o[13 + t] * coe3;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                o[23 + 3 * a + b] = o[23 + 3 * a + b] * cog1[a][b];
                o[32 + 3 * a + b] = o[32 + 3 * a + b] * cog1[a][b];
            }
        for (int a = 0; a < 3; ++a)
            for (int h = 0; h < 3; ++h)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c) {
                        const int l1 = 3 * a + h, l2 = 3 * b + c;
                        o[41 + 9 * l1 + l2] = o[41 + 9 * l1 + l2] * (cog1[a][h] * cog1[b][c]);
                    }
        u1mean = u1mean + o[0] / m->dyfi[J - 1];
        u1max = fmax(u1max, fabs(o[0]));
    }
    double area = 0.0;   /* Area_inlet (COORDJR.f90:368-373) */
    for (int J = 1; J <= MM; ++J)
        for (int k = 0; k < NN; ++k) area = area + 1.0 / m->dyfi[J - 1] / m->dzi;
    scal[0] = u1mean / m->dzi * (double)NN / area;
    scal[1] = u1max;
}
