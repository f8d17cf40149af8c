// pa_eig.h -- eigensystem of a time-reversible 4x4 rate matrix, the gBGC
// rate matrix and PhyloAcc-GT's strand-symmetric rate matrix, compiled for
// the host (the model's Q at engine creation) and for the device
// (per-element gBGC classes and per-element base frequencies), so both use
// one solver.
//
// Q = V diag(lam) Vi is found through the symmetric form
// D^1/2 Q D^-1/2 (D = diag of Q's stationary distribution) and cyclic
// Jacobi rotations: the reference calls Armadillo's eig_gen + inv instead
// (src/PhyloAcc-ST/bpp.hpp:357-362, src/PhyloAcc-ST-GBGC/SRC/bpp_c.cpp:1405-1427);
// both give P(t) = V e^{lam t} Vi to rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace pa {

#define PA_EIG_HD __host__ __device__ inline

enum { EIG_OK = 0, EIG_NO_STATIONARY = 1, EIG_NOT_REVERSIBLE = 2, EIG_INACCURATE = 3 };

// Stationary distribution of Q: pi Q = 0, sum pi = 1 (Gaussian elimination
// with partial pivoting); false unless every pi_i > 0.
PA_EIG_HD bool stationary4(const double *Q, double *pi)
{
    double A[4][5];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 4; j++) A[i][j] = Q[j * 4 + i];
        A[i][4] = 0;
    }
    for (int j = 0; j < 4; j++) A[3][j] = 1;
    A[3][4] = 1;
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++) if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
        if (fabs(A[piv][c]) < 1e-300) return false;
        for (int k = 0; k < 5; k++) { const double t = A[c][k]; A[c][k] = A[piv][k]; A[piv][k] = t; }
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            const double f = A[r][c] / A[c][c];
            for (int k = c; k < 5; k++) A[r][k] -= f * A[c][k];
        }
    }
    for (int i = 0; i < 4; i++) pi[i] = A[i][4] / A[i][i];
    for (int i = 0; i < 4; i++) if (!(pi[i] > 0)) return false;
    return true;
}

// Cyclic Jacobi on a symmetric 4x4: A = U diag(w) U^T.
PA_EIG_HD void jacobi4(double A[4][4], double w[4], double U[4][4])
{
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) U[i][j] = i == j;
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = 0;
        for (int p = 0; p < 4; p++) for (int q = p + 1; q < 4; q++) off += A[p][q] * A[p][q];
        if (off < 1e-300) break;
        for (int p = 0; p < 4; p++)
            for (int q = p + 1; q < 4; q++) {
                if (A[p][q] == 0) continue;
                const double th = (A[q][q] - A[p][p]) / (2 * A[p][q]);
                const double t = (th >= 0 ? 1 : -1) / (fabs(th) + sqrt(th * th + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 4; k++) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 4; k++) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 4; k++) {
                    const double ukp = U[k][p], ukq = U[k][q];
                    U[k][p] = c * ukp - s * ukq;
                    U[k][q] = s * ukp + c * ukq;
                }
            }
    }
    for (int i = 0; i < 4; i++) w[i] = A[i][i];
}

// Q = V diag(lam) Vi (row-major, P(t)[i][j] = sum_k V[i][k] e^{lam_k t}
// Vi[k][j] = P_{i->j}) for a time-reversible Q; EIG_* status.
PA_EIG_HD int reversible_eig4(const double *Q, double *lam, double *V, double *Vi)
{
    double pq[4];
    if (!stationary4(Q, pq)) return EIG_NO_STATIONARY;
    double A[4][4], amax = 0, asym = 0;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            A[i][j] = sqrt(pq[i]) * Q[i * 4 + j] / sqrt(pq[j]);
            amax = fmax(amax, fabs(A[i][j]));
        }
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) asym = fmax(asym, fabs(A[i][j] - A[j][i]));
    if (asym > 1e-9 * amax) return EIG_NOT_REVERSIBLE;
    for (int i = 0; i < 4; i++) for (int j = i + 1; j < 4; j++) A[i][j] = A[j][i] = 0.5 * (A[i][j] + A[j][i]);
    double U[4][4];
    jacobi4(A, lam, U);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 4; k++) {
            V[i * 4 + k] = U[i][k] / sqrt(pq[i]);
            Vi[k * 4 + i] = U[i][k] * sqrt(pq[i]);
        }
    double rec = 0, qmax = 0;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += V[i * 4 + k] * lam[k] * Vi[k * 4 + j];
            rec = fmax(rec, fabs(s - Q[i * 4 + j]));
            qmax = fmax(qmax, fabs(Q[i * 4 + j]));
        }
    return rec > 1e-10 * qmax ? EIG_INACCURATE : EIG_OK;
}

// BPP_C::getRate (src/PhyloAcc-ST-GBGC/SRC/bpp_c.cpp:13-29): Q(s, b) from the
// background Q0 (bases A, C, G, T).  Every rate i -> j is scaled by
// x / (1 - e^-x) (1 at x = 0) with x = s + b for A/T -> C/G, s - b for
// C/G -> A/T and s otherwise; the diagonal makes rows sum to 0.  s = b = 0
// returns Q0 unchanged.
PA_EIG_HD double gbgc_fix(double x) { return x == 0 ? 1.0 : x / (1 - exp(-x)); }

PA_EIG_HD void gbgc_rate(const double *Q0, double s, double b, double *Q)
{
    if (s == 0 && b == 0) {
        for (int i = 0; i < 16; i++) Q[i] = Q0[i];
        return;
    }
    for (int i = 0; i < 4; i++) {
        const bool wi = i == 0 || i == 3;        // A, T (weak)
        double row = 0;
        for (int j = 0; j < 4; j++) {
            const bool wj = j == 0 || j == 3;
            const double x = s + (wi && !wj ? b : (!wi && wj ? -b : 0.0));
            Q[i * 4 + j] = Q0[i * 4 + j] * gbgc_fix(x);
            row += Q[i * 4 + j];
        }
        Q[i * 4 + i] -= row;
    }
}

// PhyloAcc-GT's strand-symmetric stationary distribution and rate matrix
// (BPP_C::getQmat, src/PhyloAcc-GT/bpp_c.cpp:160-185, with pi set as in
// BPP_C::sample_pi, bpp_c.cpp:1545-1546): pi = (a, 1/2 - a, 1/2 - a, a) and
// Q[i][j] = inst[k(i, j)] * pi[j] off the diagonal, k over the pairs AC, AG,
// AT, CG, CT, GT; rows sum to 0.
PA_EIG_HD void gt_strand_q(const double *inst, double a, double *Q, double *pi)
{
    pi[0] = pi[3] = a;
    pi[1] = pi[2] = 0.5 - a;
    const int pair[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
    for (int i = 0; i < 4; i++) {
        double row = 0;
        for (int j = 0; j < 4; j++) {
            if (i == j) continue;
            Q[i * 4 + j] = inst[pair[i][j]] * pi[j];
            row += Q[i * 4 + j];
        }
        Q[i * 4 + i] = -row;
    }
}

}  // namespace pa
