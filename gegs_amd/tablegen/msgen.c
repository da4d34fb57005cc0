/*
 * msgen.c -- generator for the screened-Rutherford multiple-scattering
 * q^(2+) surface in the text layout that the reference reads from
 * msnew.data (PhysicsInit.cu:860-906, init_ms_sr).
 *
 * The EGSnrc data file is not available, so the surface is recomputed from
 * its published definition (Kawrakow & Bielajew, NIM B 134 (1998) 325):
 *
 *   single scattering  p(mu) = 2 eta (1+eta) / (1 - mu + 2 eta)^2
 *   Legendre moments   xi_l  = l (Q_{l-1}(x) - x Q_l(x)),  x = 1 + 2 eta
 *   q^(2+) density     F2(mu) = sum_l (l+1/2) c_l P_l(mu) / N2,
 *                      c_l = exp(-lambda)(exp(lambda xi_l) - 1 - lambda xi_l)
 *                      N2  = 1 - exp(-lambda)(1 + lambda)
 *
 * Q_l is the Legendre function of the second kind (minimal solution of the
 * three-term recurrence for x > 1), obtained by Miller's backward recurrence
 * normalised to Q_0 = 0.5 ln((x+1)/(x-1)).
 *
 * Table axes follow the reader in the reference: lambda_i = exp(i * dllamb),
 * i = 0..MAXL_MS, between LAMBMIN_MS = 1 and LAMBMAX_MS = 1e5; the first
 * GS moment G1_j = j * QMAX_MS / MAXQ_MS, j = 0..MAXQ_MS (j = 0 is evaluated at
 * G1 = 1e-4, standing in for the G1 -> 0 limit).  The sampled variable is
 *   u = (1 - mu)(1 + omega2/2) / (1 - mu + omega2)
 * with omega2 the fit used at run time in mscat (electron.cu:196-200).  The
 * u grid (MAXU_MS+1 points) is placed at equal steps of the cumulative
 * distribution; between grid points the density is linear, and the bins are
 * drawn from a Walker alias table (wms, ims), exactly the form consumed by
 * mscat (electron.cu:206-229).
 *
 * Output (stdout) per (i, j), i outer: ums[0..31], fms[0..31] (raw density
 * values at the grid points), wms[0..30], ims[0..30] (1-based), one value
 * per line.  Build: gcc -O2 -o msgen msgen.c -lm
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXL_MS 63
#define MAXQ_MS 7
#define MAXU_MS 31
#define LAMBMIN_MS 1.0
#define LAMBMAX_MS 1e5
#define QMAX_MS 0.5
#define LMAX 40000
#define NFINE 4001

static double q1_of_eta(double eta) {
    return 2.0 * eta * ((1.0 + eta) * log1p(1.0 / eta) - 1.0);
}

static double solve_eta(double q1) {
    /* q1_of_eta is increasing in eta; bisection in log(eta). */
    double lo = -60.0, hi = 10.0;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        if (q1_of_eta(exp(mid)) < q1) lo = mid; else hi = mid;
    }
    return exp(0.5 * (lo + hi));
}

static double omega2_fit(double eta, double lambda) {
    double l = log(lambda);
    double poly;
    if (l < 2.2299)
        poly = 1.347006 + l * (0.209364 - l * (0.45525 - l * (0.50142 - 0.081234 * l)));
    else
        poly = -2.77164 + l * (2.94874 - l * (0.1535754 - l * 0.00552888));
    return eta * (lambda + 4.0) * poly;
}

/* Legendre coefficients xi_l, l = 0..nl-1, of the screened Rutherford law. */
static void sr_moments(double eta, int nl, double *xi) {
    double x = 1.0 + 2.0 * eta;
    double sq = sqrt((x - 1.0) * (x + 1.0));
    /* start well beyond the decay length 1/ln(x + sqrt(x^2-1)) */
    double decay = 1.0 / log(x + sq);
    long start = nl + 60 + (long)(30.0 * decay);
    if (start > 4000000L) start = 4000000L;
    double *q = (double *)malloc(sizeof(double) * (size_t)(nl + 2));
    double qp1 = 0.0, qc = 1e-280;
    for (long l = start; l >= 1; --l) {
        /* Q_{l-1} = ((2l+1) x Q_l - (l+1) Q_{l+1}) / l */
        double qm1 = ((2.0 * l + 1.0) * x * qc - (l + 1.0) * qp1) / (double)l;
        qp1 = qc;
        qc = qm1;
        if (l - 1 <= nl + 1) q[l - 1] = qc;
        if (fabs(qc) > 1e250) {
            qc *= 1e-250;
            qp1 *= 1e-250;
            for (long k = l - 1; k <= nl + 1 && k <= start; ++k) q[k] *= 1e-250;
        }
    }
    double q0 = 0.5 * log((x + 1.0) / (x - 1.0));
    double s = q0 / q[0];
    for (int l = 0; l <= nl + 1; ++l) q[l] *= s;
    xi[0] = 1.0;
    for (int l = 1; l < nl; ++l) xi[l] = l * (q[l - 1] - x * q[l]);
    free(q);
}

static void build_one(double lambda, double g1, double *ums, double *fms,
                      double *wms, int *ims) {
    double q1 = g1 / lambda;
    double eta = solve_eta(q1);
    double om2 = omega2_fit(eta, lambda);
    static double xi[LMAX + 2], c[LMAX + 2];
    /* number of Legendre terms: grow until the coefficients are negligible */
    int nl = 64;
    for (;;) {
        sr_moments(eta, nl, xi);
        double emx = exp(-lambda);
        double tail = 0.0;
        for (int l = 0; l < nl; ++l) {
            double a = lambda * xi[l];
            double cl;
            if (lambda > 60.0) {
                cl = exp(-lambda * (1.0 - xi[l]));
            } else if (fabs(a) < 1e-3) {
                cl = emx * a * a * (0.5 + a / 6.0 + a * a / 24.0);
            } else {
                cl = emx * (expm1(a) - a);
            }
            c[l] = cl;
        }
        for (int l = nl - 8; l < nl; ++l) tail += fabs((l + 0.5) * c[l]);
        if (tail < 1e-10 * (0.5 * c[0]) || nl >= LMAX) break;
        nl *= 2;
        if (nl > LMAX) nl = LMAX;
    }
    double n2 = c[0];
    /* density in u on a fine grid */
    static double uf[NFINE], ff[NFINE], cdf[NFINE];
    double a1 = 1.0 + 0.5 * om2;
    for (int k = 0; k < NFINE; ++k) {
        double u = (double)k / (NFINE - 1);
        double omu = om2 * u / (a1 - u);              /* 1 - mu */
        if (omu > 2.0) omu = 2.0;
        double mu = 1.0 - omu;
        double p0 = 1.0, p1 = mu, sum = 0.5 * c[0];
        if (nl > 1) sum += 1.5 * c[1] * p1;
        for (int l = 2; l < nl; ++l) {
            double p2 = ((2.0 * l - 1.0) * mu * p1 - (l - 1.0) * p0) / l;
            sum += (l + 0.5) * c[l] * p2;
            p0 = p1;
            p1 = p2;
        }
        double fmu = sum / n2;
        double dmu = om2 * a1 / ((a1 - u) * (a1 - u));
        double fu = fmu * dmu;
        if (!(fu > 1e-12)) fu = 1e-12;
        uf[k] = u;
        ff[k] = fu;
    }
    cdf[0] = 0.0;
    for (int k = 1; k < NFINE; ++k)
        cdf[k] = cdf[k - 1] + 0.5 * (ff[k] + ff[k - 1]) * (uf[k] - uf[k - 1]);
    double tot = cdf[NFINE - 1];
    for (int k = 0; k < NFINE; ++k) cdf[k] /= tot;
    /* grid at equal cumulative steps */
    int kk = 0;
    for (int g = 0; g <= MAXU_MS; ++g) {
        double target = (double)g / MAXU_MS;
        double u, f;
        if (g == 0) { u = 0.0; f = ff[0]; }
        else if (g == MAXU_MS) { u = 1.0; f = ff[NFINE - 1]; }
        else {
            while (kk < NFINE - 2 && cdf[kk + 1] < target) ++kk;
            double t = (target - cdf[kk]) / (cdf[kk + 1] - cdf[kk]);
            u = uf[kk] + t * (uf[kk + 1] - uf[kk]);
            f = ff[kk] + t * (ff[kk + 1] - ff[kk]);
        }
        ums[g] = u;
        fms[g] = f / tot;
    }
    /* bin probabilities under the linear-in-bin density */
    double p[MAXU_MS], ps = 0.0;
    for (int g = 0; g < MAXU_MS; ++g) {
        p[g] = 0.5 * (fms[g] + fms[g + 1]) * (ums[g + 1] - ums[g]);
        ps += p[g];
    }
    for (int g = 0; g < MAXU_MS; ++g) p[g] *= (double)MAXU_MS / ps;
    /* Walker alias table: bin g kept with probability wms[g], else ims[g] */
    int small[MAXU_MS], large[MAXU_MS], ns = 0, nlg = 0;
    for (int g = 0; g < MAXU_MS; ++g) {
        wms[g] = 1.0;
        ims[g] = g;
        if (p[g] < 1.0) small[ns++] = g; else large[nlg++] = g;
    }
    while (ns > 0 && nlg > 0) {
        int s = small[--ns];
        int l = large[--nlg];
        wms[s] = p[s];
        ims[s] = l;
        p[l] -= (1.0 - p[s]);
        if (p[l] < 1.0) small[ns++] = l; else large[nlg++] = l;
    }
}

int main(void) {
    double dll = (log(LAMBMAX_MS) - log(LAMBMIN_MS)) / MAXL_MS;
    double dq = QMAX_MS / MAXQ_MS;
    for (int i = 0; i <= MAXL_MS; ++i) {
        double lambda = exp(log(LAMBMIN_MS) + i * dll);
        for (int j = 0; j <= MAXQ_MS; ++j) {
            double g1 = j == 0 ? 1e-4 : j * dq;
            double ums[MAXU_MS + 1], fms[MAXU_MS + 1], wms[MAXU_MS];
            int ims[MAXU_MS];
            build_one(lambda, g1, ums, fms, wms, ims);
            for (int k = 0; k <= MAXU_MS; ++k) printf("%.9e\n", ums[k]);
            for (int k = 0; k <= MAXU_MS; ++k) printf("%.9e\n", fms[k]);
            for (int k = 0; k < MAXU_MS; ++k) printf("%.9e\n", wms[k]);
            for (int k = 0; k < MAXU_MS; ++k) printf("%d\n", ims[k] + 1);
        }
    }
    return 0;
}
