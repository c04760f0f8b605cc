/* best_oracle.c -- CPU FP64 restatement of BEST's dense boundary-element path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it as the checker.  The product (the HIP library
 * in best-hydro_amd/csrc) never links, loads or calls it.
 *
 * Every routine restates the reference operation for operation (same association order,
 * same literals, built with -ffp-contract=off), citing the file:line it follows, so that the
 * matrix entries it produces are the reference's bit for bit where libm agrees:
 *
 *   bo_prep       best/best.f:261-315      centroid shift, triangle areas, area factor
 *   bo_compon     best/compon.f90:19-181   surfaces (vertex matching, union-find), open edges
 *   bo_defgam     best/compon.f90:185-197  per-surface deflation weights
 *   bo_self       best/tgetgr.f:35-145     incenter + exact self block (rotated)
 *   bo_pair       best/triangleint.f:1-247 off-diagonal 3x3 block (21-point Gauss)
 *   bo_tgetgr     best/tgetgr.f:1-173      the full 3N x 3N GR, column-major
 *   bo_entries    hmatrix/best_hmatrix.h:38-40 (bhm_entries) a sub-block of GR
 *   bo_deflate    best/best.f:384-426      normal-field deflation of GR
 *   bo_getrf/rs   LAPACK dgetrf/dgetrs semantics (best/best.f:444, best/besolve.f90:133):
 *                 unblocked right-looking LU with partial pivoting (first index of max |.|)
 *   bo_difco_rhs  best/difco.f:28-44       the six rigid-body right-hand sides
 *   bo_difco_post best/difco.f:51-401      friction tensors, centres, Dtt/Dtr/Drr
 *   bo_strain     best/visco.f:216-243     the five strain-flow right-hand sides
 *   bo_visco_post best/visco.f:25-211      centre of viscosity, CHIPRIME
 *   bo_matin3     best/matin3.f:1-117      3x3 inverse (Numerical Recipes LU)
 *   bo_jacobi     best/jacobi.f:1-96       3x3 cyclic Jacobi eigenvalues
 *   bo_units      best/output.f:16-29,98-167  unit factors and the printed scalars
 *
 * Arrays follow BEST: V1,V2,V3,X are NAR x 3 column-major (v[c*nar+i]); GR is 3N x 3N
 * column-major with leading dimension ld; unknown 3*i+a is component a on triangle i.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, c, n) ((size_t)(c) * (size_t)(n) + (size_t)(i))

/* ---------------------------------------------------------------- geometry prep */

/* best/best.f:261-315.  Shifts the vertices to the centroid of the triangle centroids, computes
 * the triangle areas and, with iarea = 1, scales them by afac = truearea / totarea. */
void bo_prep(int nar, double *v1, double *v2, double *v3, int iarea, double truearea, double *area,
             double *afac_out, double *totarea_out, double *centroid_out) {
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (int j = 0; j < nar; ++j) {
        cx = cx + (v1[IDX(j, 0, nar)] + v2[IDX(j, 0, nar)] + v3[IDX(j, 0, nar)]) / 3.0;
        cy = cy + (v1[IDX(j, 1, nar)] + v2[IDX(j, 1, nar)] + v3[IDX(j, 1, nar)]) / 3.0;
        cz = cz + (v1[IDX(j, 2, nar)] + v2[IDX(j, 2, nar)] + v3[IDX(j, 2, nar)]) / 3.0;
    }
    cx = cx / (double)nar;
    cy = cy / (double)nar;
    cz = cz / (double)nar;
    const double c[3] = {cx, cy, cz};
    for (int n = 0; n < nar; ++n)
        for (int k = 0; k < 3; ++k) {
            v1[IDX(n, k, nar)] = v1[IDX(n, k, nar)] - c[k];
            v2[IDX(n, k, nar)] = v2[IDX(n, k, nar)] - c[k];
            v3[IDX(n, k, nar)] = v3[IDX(n, k, nar)] - c[k];
        }
    double tot = 0.0;
    for (int n = 0; n < nar; ++n) {
        const double ax = v2[IDX(n, 0, nar)] - v1[IDX(n, 0, nar)];
        const double ay = v2[IDX(n, 1, nar)] - v1[IDX(n, 1, nar)];
        const double az = v2[IDX(n, 2, nar)] - v1[IDX(n, 2, nar)];
        const double bx = v3[IDX(n, 0, nar)] - v1[IDX(n, 0, nar)];
        const double by = v3[IDX(n, 1, nar)] - v1[IDX(n, 1, nar)];
        const double bz = v3[IDX(n, 2, nar)] - v1[IDX(n, 2, nar)];
        const double px = by * az - bz * ay;
        const double py = bz * ax - bx * az;
        const double pz = bx * ay - by * ax;
        area[n] = sqrt(px * px + py * py + pz * pz) * 0.5;
        tot = tot + area[n];
    }
    double afac = 1.0;
    if (iarea == 1) {
        afac = truearea / tot;
        for (int j = 0; j < nar; ++j) area[j] = afac * area[j];
    }
    *afac_out = afac;
    *totarea_out = tot;
    centroid_out[0] = cx;
    centroid_out[1] = cy;
    centroid_out[2] = cz;
}

/* ---------------------------------------------------------------- surfaces */

static int bo_key_before(const double *a, const double *b) {
    if (a[0] != b[0]) return a[0] < b[0];
    if (a[1] != b[1]) return a[1] < b[1];
    return a[2] < b[2];
}

/* best/compon.f90:120-181: stable bottom-up merge sort of the columns key[3*m] (0-based idx). */
static void bo_msort3(int m, const double *key, int *idx) {
    int *tmp = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    for (int i = 0; i < m; ++i) idx[i] = i;
    for (int w = 1; w < m; w *= 2) {
        for (int lo = 0; lo < m; lo += 2 * w) {
            int mid = lo + w - 1 < m - 1 ? lo + w - 1 : m - 1;
            int hi = lo + 2 * w - 1 < m - 1 ? lo + 2 * w - 1 : m - 1;
            int i = lo, j = mid + 1, k = lo;
            while (i <= mid && j <= hi) {
                if (bo_key_before(&key[3 * (size_t)idx[j]], &key[3 * (size_t)idx[i]]))
                    tmp[k++] = idx[j++];
                else
                    tmp[k++] = idx[i++];
            }
            while (i <= mid) tmp[k++] = idx[i++];
            while (j <= hi) tmp[k++] = idx[j++];
        }
        memcpy(idx, tmp, sizeof(int) * (size_t)m);
    }
    free(tmp);
}

static int bo_find(int *root, int i) {
    int p = i;
    while (root[p] != p) p = root[p];
    int q = i;
    while (root[q] != p) {
        int t = root[q];
        root[q] = p;
        q = t;
    }
    return p;
}

/* best/compon.f90:19-115.  icomp[n] in 1..ncomp (numbered by first triangle); nopen counts the
 * edges not shared by exactly two triangles. */
void bo_compon(int nar, const double *v1, const double *v2, const double *v3, int *icomp, int *ncomp_out,
               int *nopen_out) {
    const int nv = 3 * nar;
    double *key = (double *)malloc(sizeof(double) * 3 * (size_t)nv);
    int *idx = (int *)malloc(sizeof(int) * (size_t)nv);
    int *vid = (int *)malloc(sizeof(int) * (size_t)nv);
    int *root = (int *)malloc(sizeof(int) * (size_t)nar);
    int *lab = (int *)calloc((size_t)nar, sizeof(int));
    for (int n = 0; n < nar; ++n)
        for (int c = 0; c < 3; ++c) {
            key[3 * (size_t)(3 * n + 0) + c] = v1[IDX(n, c, nar)];
            key[3 * (size_t)(3 * n + 1) + c] = v2[IDX(n, c, nar)];
            key[3 * (size_t)(3 * n + 2) + c] = v3[IDX(n, c, nar)];
        }
    bo_msort3(nv, key, idx);
    int nid = 0;
    for (int j = 0; j < nv; ++j) {
        if (j == 0) {
            nid = 1;
        } else {
            const double *a = &key[3 * (size_t)idx[j]], *b = &key[3 * (size_t)idx[j - 1]];
            if (a[0] != b[0] || a[1] != b[1] || a[2] != b[2]) nid = nid + 1;
        }
        vid[idx[j]] = nid;
    }
    for (int n = 0; n < nar; ++n) root[n] = n;
    for (int j = 1; j < nv; ++j) {
        if (vid[idx[j]] == vid[idx[j - 1]]) {
            int ri = bo_find(root, idx[j] / 3), rj = bo_find(root, idx[j - 1] / 3);
            if (ri != rj) root[ri > rj ? ri : rj] = ri < rj ? ri : rj;
        }
    }
    int ncomp = 0;
    for (int n = 0; n < nar; ++n) {
        int r = bo_find(root, n);
        if (lab[r] == 0) lab[r] = ++ncomp;
        icomp[n] = lab[r];
    }
    for (int n = 0; n < nar; ++n)
        for (int c = 0; c < 3; ++c) {
            int a = vid[3 * n + c], b = vid[3 * n + (c + 1) % 3];
            key[3 * (size_t)(3 * n + c) + 0] = a < b ? a : b;
            key[3 * (size_t)(3 * n + c) + 1] = a < b ? b : a;
            key[3 * (size_t)(3 * n + c) + 2] = 0.0;
        }
    bo_msort3(nv, key, idx);
    int nopen = 0, k = 0;
    for (int j = 1; j <= nv; ++j) {
        if (j == nv) {
            if (j - k != 2) nopen++;
        } else {
            const double *a = &key[3 * (size_t)idx[j]], *b = &key[3 * (size_t)idx[k]];
            if (a[0] != b[0] || a[1] != b[1] || a[2] != b[2]) {
                if (j - k != 2) nopen++;
                k = j;
            }
        }
    }
    *ncomp_out = ncomp;
    *nopen_out = nopen;
    free(key);
    free(idx);
    free(vid);
    free(root);
    free(lab);
}

/* best/compon.f90:185-197. */
void bo_defgam(int nar, int ndef, const int *icomp, const double *xmu, double diagm, double *gdefk) {
    for (int k = 0; k < ndef; ++k) gdefk[k] = 0.0;
    for (int n = 0; n < nar; ++n) {
        const double a = xmu[3 * n], b = xmu[3 * n + 1], c = xmu[3 * n + 2];
        double *g = &gdefk[icomp[n] - 1];
        *g = *g + a * a + b * b + c * c;
    }
    for (int k = 0; k < ndef; ++k) gdefk[k] = diagm / gdefk[k];
}

/* ---------------------------------------------------------------- self block */

static double bo_si(double x) { return 2.0 * (log((1.0 + sqrt(1.0 - x * x)) / x) - sqrt(1.0 - x * x)); }
static double bo_csi(double x) { return 3.0 - 1.0 / x - 2.0 * x; }

/* best/tgetgr.f:35-145.  Triangle n: incenter xc[3] and the 3x3 self block s (row-major,
 * s[3*i+j] = GR(3n+i, 3n+j)). */
void bo_self(int nar, int n, double afac, double cg, const double *v1, const double *v2, const double *v3,
             double *xc, double *s) {
    const double pi = 3.1415926535897932385;
    const double p1x = v1[IDX(n, 0, nar)], p1y = v1[IDX(n, 1, nar)], p1z = v1[IDX(n, 2, nar)];
    const double p2x = v2[IDX(n, 0, nar)], p2y = v2[IDX(n, 1, nar)], p2z = v2[IDX(n, 2, nar)];
    const double p3x = v3[IDX(n, 0, nar)], p3y = v3[IDX(n, 1, nar)], p3z = v3[IDX(n, 2, nar)];
    const double e21x = p2x - p1x, e21y = p2y - p1y, e21z = p2z - p1z;
    const double e31x = p3x - p1x, e31y = p3y - p1y, e31z = p3z - p1z;
    const double e32x = p3x - p2x, e32y = p3y - p2y, e32z = p3z - p2z;
    const double a = sqrt(e21x * e21x + e21y * e21y + e21z * e21z);
    const double b = sqrt(e31x * e31x + e31y * e31y + e31z * e31z);
    const double c = sqrt(e32x * e32x + e32y * e32y + e32z * e32z);
    const double cosC = (e21x * e31x + e21y * e31y + e21z * e31z) / a / b;
    const double t = (a + b - c) / (2. * (1. + cosC));
    xc[0] = p1x + t * (e21x / a + e31x / b);
    xc[1] = p1y + t * (e21y / a + e31y / b);
    xc[2] = p1z + t * (e21z / a + e31z / b);
    double nx = e31y * e21z - e31z * e21y;
    double ny = e31z * e21x - e31x * e21z;
    double nz = e31x * e21y - e31y * e21x;
    const double area2 = sqrt(nx * nx + ny * ny + nz * nz) / 2.;
    nx = .5 * nx / area2;
    ny = .5 * ny / area2;
    nz = .5 * nz / area2;
    const double cc = acos(cosC);
    const double r = t * sin(cc);
    const double cosB = -(e21x * e32x + e21y * e32y + e21z * e32z) / a / c;
    const double bb = acos(cosB);
    const double aa = pi - bb - cc;
    const double t2 = 1. / (2. * cos(0.5 * cc));
    const double dc = (a + b - c) * t2;
    const double db = sqrt(a * a + (b - c) * (b - c) - 2. * a * (b - c) * cosC) * t2;
    const double da = sqrt(b * b + (a - c) * (a - c) - 2. * b * (a - c) * cosC) * t2;
    const double ha = r / da, hb = r / db, hc = r / dc;
    const double dpa = (pi - aa) * (da - r) - 2.0 * (da * acos(ha) - r * log((1. + sqrt(1. - ha * ha)) / ha));
    const double dpb = (pi - bb) * (db - r) - 2.0 * (db * acos(hb) - r * log((1. + sqrt(1. - hb * hb)) / hb));
    const double dpc = (pi - cc) * (dc - r) - 2.0 * (dc * acos(hc) - r * log((1. + sqrt(1. - hc * hc)) / hc));
    const double sca =
        0.5 * r * ((cos(2.0 * cc) + cos(2.0 * bb)) * bo_si(ha) - (sin(2.0 * cc) + sin(2.0 * bb)) * bo_csi(ha));
    const double scb = 0.5 * r * ((1. + cos(2.0 * bb)) * bo_si(hb) + sin(2.0 * bb) * bo_csi(hb));
    const double scc = 0.5 * r * ((1. + cos(2.0 * cc)) * bo_si(hc) + sin(2.0 * cc) * bo_csi(hc));
    const double cca = -sin(bb - cc) * r * (cos(aa) * bo_si(ha) + sin(aa) * bo_csi(ha));
    const double ccb = sin(bb) * r * (cos(bb) * bo_si(hb) + sin(bb) * bo_csi(hb));
    const double ccc = -sin(cc) * r * (cos(cc) * bo_si(hc) + sin(cc) * bo_csi(hc));
    const double cor = sqrt(afac);
    double g[3][3];
    g[0][0] = cor * ((2.0 + cg) * pi * r + (1.0 + 0.5 * cg) * cg * (dpa + dpb + dpc) + cg * 0.25 * (sca + scb + scc));
    g[0][1] = 0.5 * cor * (cca + ccb + ccc) * cg;
    g[1][0] = g[0][1];
    g[1][1] = g[0][0] - 0.5 * cor * cg * (sca + scb + scc);
    g[2][2] = cor * (2.0 * pi * r + (dpa + dpb + dpc));
    g[0][2] = g[1][2] = g[2][0] = g[2][1] = 0.0;
    /* Direction cosines: rows are the edge 1->2, normal x edge, and the unit normal. */
    double rr[3][3];
    rr[0][0] = e21x / a;
    rr[0][1] = e21y / a;
    rr[0][2] = e21z / a;
    rr[2][0] = nx;
    rr[2][1] = ny;
    rr[2][2] = nz;
    rr[1][0] = (ny * e21z - nz * e21y) / a;
    rr[1][1] = (nz * e21x - nx * e21z) / a;
    rr[1][2] = (nx * e21y - ny * e21x) / a;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k)
                acc = acc + rr[k][i] * (g[k][0] * rr[0][j] + g[k][1] * rr[1][j] + g[k][2] * rr[2][j]);
            s[3 * i + j] = acc;
        }
}

/* ---------------------------------------------------------------- pair block */

/* The Gauss-Legendre nodes and weights exactly as best/triangleint.f:66-108 spells them
 * (nodes mapped to [0,1], weights of [-1,1]). */
static const double kNode[21] = {
    3.1239146898052499e-3, 1.6386580716846853e-2, 3.9950332924799586e-2, 7.3318317708341358e-2,
    0.1157800182621610,    0.1664305979012938,    0.2241905820563901,    0.2878289398962806,
    0.3559893415987995,    0.4272190729195525,    0.5,                   0.5727809270804476,
    0.6440106584012005,    0.7121710601037194,    0.7758094179436099,    0.8335694020987062,
    0.8842199817378390,    0.9266816822916586,    0.9600496670752004,    0.9836134192831531,
    0.9968760853101948};
static const double kWeightHalf[11] = {
    1.601722825777433e-2, 3.695378977085249e-2, 5.713442542685720e-2, 7.610011362837930e-2,
    9.344442345603390e-2, 0.1087972991671484,   0.1218314160537285,   0.1322689386333375,
    0.1398873947910731,   0.1445244039899701,   0.1460811336496904};

void bo_gauss_table(double *nodes, double *weights) {
    for (int j = 0; j < 21; ++j) {
        nodes[j] = kNode[j];
        weights[j] = kWeightHalf[j < 11 ? j : 20 - j];
    }
}

static double bo_dot(const double *p, const double *q) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; }

/* best/triangleint.f:1-190 (the integrands h, hx, hy, hxy at :198-247).  The block
 * G(n1 collocation, n2 source) into g (row-major, g[3*i+j] = GR(3n1+i, 3n2+j)).  xc is the
 * incenter of n1; area is the corrected area of n2. */
void bo_pair_xc(int nar, int n2, const double *xc, double cg, const double *v1, const double *v2, const double *v3,
                double area_n2, double *gout) {
    double p1[3], p2[3], p3[3];
    for (int k = 0; k < 3; ++k) {
        p1[k] = v1[IDX(n2, k, nar)];
        p2[k] = v2[IDX(n2, k, nar)];
        p3[k] = v3[IDX(n2, k, nar)];
    }
    double v21[3], v31[3], v32[3], v1c[3], v2c[3], v3c[3];
    for (int k = 0; k < 3; ++k) {
        v21[k] = p2[k] - p1[k];
        v31[k] = p3[k] - p1[k];
        v32[k] = p3[k] - p2[k];
        v1c[k] = p1[k] - xc[k];
        v2c[k] = p2[k] - xc[k];
        v3c[k] = p3[k] - xc[k];
    }
    const double d2121 = bo_dot(v21, v21), d2131 = bo_dot(v21, v31), d2132 = bo_dot(v21, v32);
    const double d211c = bo_dot(v21, v1c), d212c = bo_dot(v21, v2c), d3131 = bo_dot(v31, v31);
    const double d3232 = bo_dot(v32, v32), d3231 = bo_dot(v32, v31), d322c = bo_dot(v32, v2c);
    const double d1c1c = bo_dot(v1c, v1c), d1c31 = bo_dot(v1c, v31), d2c2c = bo_dot(v2c, v2c);
    const double d2c1c = bo_dot(v2c, v1c), d3c31 = bo_dot(v3c, v31), d3c32 = bo_dot(v3c, v32);
    const double d3c3c = bo_dot(v3c, v3c);

    double t0 = 0.0, s0 = 0.0, sx = 0.0, sy = 0.0, sxy = 0.0, syy = 0.0, sxx = 0.0;
    for (int j = 0; j < 21; ++j) {
        const double y = kNode[j];
        const double w = kWeightHalf[j < 11 ? j : 20 - j];
        /* h(y), best/triangleint.f:198-206 */
        const double fh = log((d212c + d2132 * y + sqrt(d2121) * sqrt(d2c2c + 2.0 * y * d322c + y * y * d3232)) /
                              (d211c + d2131 * y + sqrt(d2121) * sqrt(d1c1c + 2.0 * d1c31 * y + d3131 * y * y))) /
                          sqrt(d2121);
        /* hx(y), :208-220 */
        double a = (-d1c31 - y * d2131) / sqrt(d1c1c + 2.0 * y * d211c + y * y * d2121);
        double b = (-y * d3231 + d3c31) / sqrt(y * y * d3232 - 2.0 * y * d3c32 + d3c3c);
        double c = (-d1c31 * d1c31 - 2.0 * y * d1c31 * d2131 - y * y * d2131 * d2131 + d1c1c * d3131 +
                    2.0 * y * d211c * d3131 + y * y * d2121 * d3131);
        const double fhx = (a + b) / c;
        /* hy(y), :222-233 */
        a = (d211c + y * d2131) / sqrt(d1c1c + 2.0 * y * d1c31 + y * y * d3131);
        b = (y * d2132 + d212c) / sqrt(d2c2c + 2.0 * y * d322c + y * y * d3232);
        c = (d211c * d211c - d1c1c * d2121 - 2.0 * y * d1c31 * d2121 + 2.0 * y * d211c * d2131 +
             y * y * d2131 * d2131 - y * y * d2121 * d3131);
        const double fhy = (a - b) / c;
        /* hxy(y), :235-247 */
        a = sqrt(d1c1c + 2.0 * y * d1c31 + y * y * d3131);
        b = (y * d211c - d2c1c - 2.0 * y * d1c31 - y * d2131 - y * y * d3231) /
            sqrt(d2c2c + 2.0 * y * d322c + y * y * d3232);
        c = (-d211c * d211c + d1c1c * d2121 + 2.0 * y * d1c31 * d2121 - 2.0 * y * d211c * d2131 -
             y * y * (d2131 * d2131) + y * y * d2121 * d3131);
        const double fhxy = (a + b) / c;
        t0 = t0 + w * fh;
        s0 = s0 + w * fhy;
        sx = sx + w * fhxy;
        sy = sy + w * y * fhy;
        sxy = sxy + w * y * fhxy;
        syy = syy + w * y * y * fhy;
        sxx = sxx + w * y * y * fhx;
    }
    t0 = 0.5 * t0;
    s0 = 0.5 * s0;
    sx = 0.5 * sx;
    sy = 0.5 * sy;
    sxy = 0.5 * sxy;
    syy = 0.5 * syy;
    sxx = 0.5 * sxx;
    /* t_ab, best/triangleint.f:149-177 */
    double tt[3][3];
    for (int p = 0; p < 3; ++p)
        for (int q = p; q < 3; ++q)
            tt[p][q] = v1c[p] * v1c[q] * s0 + (v1c[p] * v21[q] + v21[p] * v1c[q]) * sx +
                       (v31[p] * v1c[q] + v1c[p] * v31[q]) * sy + (v21[p] * v31[q] + v31[p] * v21[q]) * sxy +
                       v21[p] * v21[q] * sxx + v31[p] * v31[q] * syy;
    /* best/triangleint.f:180-188 */
    const double g11 = 2.0 * area_n2 * (t0 + cg * tt[0][0]);
    const double g21 = 2.0 * area_n2 * tt[0][1] * cg;
    const double g22 = 2.0 * area_n2 * (t0 + cg * tt[1][1]);
    const double g31 = 2.0 * area_n2 * tt[0][2] * cg;
    const double g32 = 2.0 * area_n2 * tt[1][2] * cg;
    const double g33 = 2.0 * area_n2 * (t0 + cg * tt[2][2]);
    gout[0] = g11;
    gout[1] = g21;
    gout[2] = g31;
    gout[3] = g21;
    gout[4] = g22;
    gout[5] = g32;
    gout[6] = g31;
    gout[7] = g32;
    gout[8] = g33;
}

/* Incenters of all triangles (x is NAR x 3 column-major) and the self blocks (9 per triangle,
 * row-major).  best/tgetgr.f:35-145. */
void bo_selfall(int nar, double afac, double cg, const double *v1, const double *v2, const double *v3, double *x,
                double *selfb) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < nar; ++n) {
        double xc[3];
        bo_self(nar, n, afac, cg, v1, v2, v3, xc, &selfb[9 * (size_t)n]);
        for (int k = 0; k < 3; ++k) x[IDX(n, k, nar)] = xc[k];
    }
}

/* GR(i0:i0+m-1, j0:j0+nc-1) before deflation, column-major with leading dimension lda
 * (bhm_entries, hmatrix/best_hmatrix.h:38-40). */
void bo_entries(int nar, double afac, double cg, const double *v1, const double *v2, const double *v3,
                const double *area, const double *x, const double *selfb, int i0, int m, int j0, int nc, double *a,
                int lda) {
    (void)afac;
    const int t1lo = i0 / 3, t1hi = (i0 + m - 1) / 3;
    const int t2lo = j0 / 3, t2hi = (j0 + nc - 1) / 3;
#pragma omp parallel for schedule(dynamic, 4)
    for (int n2 = t2lo; n2 <= t2hi; ++n2) {
        double g[9], xc[3];
        for (int n1 = t1lo; n1 <= t1hi; ++n1) {
            if (n1 == n2) {
                memcpy(g, &selfb[9 * (size_t)n1], sizeof g);
            } else {
                for (int k = 0; k < 3; ++k) xc[k] = x[IDX(n1, k, nar)];
                bo_pair_xc(nar, n2, xc, cg, v1, v2, v3, area[n2], g);
            }
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) {
                    const int i = 3 * n1 + p - i0, j = 3 * n2 + q - j0;
                    if (i >= 0 && i < m && j >= 0 && j < nc) a[(size_t)j * (size_t)lda + (size_t)i] = g[3 * p + q];
                }
        }
    }
}

/* best/tgetgr.f:1-173: X and the full GR (3N x 3N, leading dimension ld). */
void bo_tgetgr(int nar, double afac, double cg, const double *v1, const double *v2, const double *v3, double *x,
               const double *area, double *gr, long ld) {
    double *selfb = (double *)malloc(sizeof(double) * 9 * (size_t)nar);
    bo_selfall(nar, afac, cg, v1, v2, v3, x, selfb);
#pragma omp parallel for schedule(dynamic, 8)
    for (int n1 = 0; n1 < nar; ++n1) {
        double g[9], xc[3];
        for (int k = 0; k < 3; ++k) xc[k] = x[IDX(n1, k, nar)];
        for (int n2 = 0; n2 < nar; ++n2) {
            if (n1 == n2)
                memcpy(g, &selfb[9 * (size_t)n1], sizeof g);
            else
                bo_pair_xc(nar, n2, xc, cg, v1, v2, v3, area[n2], g);
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) gr[(size_t)(3 * n2 + q) * (size_t)ld + (size_t)(3 * n1 + p)] = g[3 * p + q];
        }
    }
    free(selfb);
}

/* ---------------------------------------------------------------- deflation */

/* best/best.f:384-426.  xmu (3N) = area * unit normal (BEST's normal (v3-v1) x (v2-v1)),
 * diagm = mean diagonal entry of GR, gdefk[k] the weight of surface k (ndef = 1: one weight).
 * With gr != NULL adds gdefk(comp(j)) xmu(j) xmu(i) to GR(i,j) for i,j on the same surface. */
void bo_deflate(int nar, const double *v1, const double *v2, const double *v3, const double *area, int ndef,
                const int *icomp, double *gr, long ld, double *xmu, double *diagm_out, double *gdefk) {
    const int n3 = 3 * nar;
    double xmu2 = 0.0, diagm = 0.0;
    for (int n = 0; n < nar; ++n) {
        const double ax = v2[IDX(n, 0, nar)] - v1[IDX(n, 0, nar)];
        const double ay = v2[IDX(n, 1, nar)] - v1[IDX(n, 1, nar)];
        const double az = v2[IDX(n, 2, nar)] - v1[IDX(n, 2, nar)];
        const double bx = v3[IDX(n, 0, nar)] - v1[IDX(n, 0, nar)];
        const double by = v3[IDX(n, 1, nar)] - v1[IDX(n, 1, nar)];
        const double bz = v3[IDX(n, 2, nar)] - v1[IDX(n, 2, nar)];
        const double px = by * az - bz * ay;
        const double py = bz * ax - bx * az;
        const double pz = bx * ay - by * ax;
        const double pl = sqrt(px * px + py * py + pz * pz);
        xmu[3 * n] = area[n] * px / pl;
        xmu[3 * n + 1] = area[n] * py / pl;
        xmu[3 * n + 2] = area[n] * pz / pl;
        xmu2 = xmu2 + xmu[3 * n] * xmu[3 * n] + xmu[3 * n + 1] * xmu[3 * n + 1] + xmu[3 * n + 2] * xmu[3 * n + 2];
        if (gr)
            diagm = diagm + gr[(size_t)(3 * n) * ld + 3 * n] + gr[(size_t)(3 * n + 1) * ld + 3 * n + 1] +
                    gr[(size_t)(3 * n + 2) * ld + 3 * n + 2];
    }
    if (gr) diagm = diagm / (double)n3;
    else diagm = *diagm_out; /* caller supplies the mean diagonal (bhm_diag_mean) */
    if (ndef == 1) {
        gdefk[0] = diagm / xmu2;
    } else {
        bo_defgam(nar, ndef, icomp, xmu, diagm, gdefk);
    }
    *diagm_out = diagm;
    if (!gr) return;
    if (ndef == 1) {
        const double gdef = gdefk[0];
#pragma omp parallel for schedule(static)
        for (int j = 0; j < n3; ++j) {
            const double gdj = gdef * xmu[j];
            double *col = gr + (size_t)j * ld;
            for (int i = 0; i < n3; ++i) col[i] = col[i] + gdj * xmu[i];
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int j = 0; j < n3; ++j) {
            const int k = icomp[j / 3];
            const double gdj = gdefk[k - 1] * xmu[j];
            double *col = gr + (size_t)j * ld;
            for (int i = 0; i < n3; ++i)
                if (icomp[i / 3] == k) col[i] = col[i] + gdj * xmu[i];
        }
    }
}

/* ---------------------------------------------------------------- LU (LAPACK semantics) */

/* Unblocked right-looking LU with partial pivoting (LAPACK dgetf2 semantics: the pivot is the
 * first row of largest |a_ik|, ipiv is 1-based).  Returns info as LAPACK does. */
int bo_getrf(int n, double *a, long lda, int *ipiv) {
    int info = 0;
    for (int k = 0; k < n; ++k) {
        double *ck = a + (size_t)k * lda;
        int p = k;
        double best = fabs(ck[k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(ck[i]) > best) {
                best = fabs(ck[i]);
                p = i;
            }
        ipiv[k] = p + 1;
        if (ck[p] == 0.0) {
            if (info == 0) info = k + 1;
            continue;
        }
        if (p != k)
            for (int j = 0; j < n; ++j) {
                double *cj = a + (size_t)j * lda;
                double t = cj[k];
                cj[k] = cj[p];
                cj[p] = t;
            }
        const double r = 1.0 / ck[k];
        for (int i = k + 1; i < n; ++i) ck[i] *= r;
#pragma omp parallel for schedule(static) if (n - k > 256)
        for (int j = k + 1; j < n; ++j) {
            double *cj = a + (size_t)j * lda;
            const double u = cj[k];
            if (u != 0.0)
                for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * u;
        }
    }
    return info;
}

/* dgetrs('N'): solves with the factors of bo_getrf; b is n x nrhs column-major. */
void bo_getrs(int n, int nrhs, const double *a, long lda, const int *ipiv, double *b, long ldb) {
    for (int r = 0; r < nrhs; ++r) {
        double *x = b + (size_t)r * ldb;
        for (int k = 0; k < n; ++k) {
            int p = ipiv[k] - 1;
            if (p != k) {
                double t = x[k];
                x[k] = x[p];
                x[p] = t;
            }
        }
        for (int j = 0; j < n; ++j) {
            const double xj = x[j];
            if (xj != 0.0) {
                const double *cj = a + (size_t)j * lda;
                for (int i = j + 1; i < n; ++i) x[i] -= xj * cj[i];
            }
        }
        for (int j = n - 1; j >= 0; --j) {
            const double *cj = a + (size_t)j * lda;
            x[j] = x[j] / cj[j];
            const double xj = x[j];
            if (xj != 0.0)
                for (int i = 0; i < j; ++i) x[i] -= xj * cj[i];
        }
    }
}

/* ---------------------------------------------------------------- 3x3 algebra */

/* best/matin3.f:24-88 (LUDCMP3), a is 3x3 column-major (a[i + 3*j]). */
static int bo_ludcmp3(double *a, int *indx) {
    const double tiny = 1.e-20;
    double vv[3];
    int ierr = 0;
    for (int i = 0; i < 3; ++i) {
        double aamax = 0.0;
        for (int j = 0; j < 3; ++j)
            if (fabs(a[i + 3 * j]) > aamax) aamax = fabs(a[i + 3 * j]);
        if (aamax == 0.0) {
            ierr = 1;
            aamax = 1.0;
        }
        vv[i] = 1.0 / aamax;
    }
    if (ierr == 1) return 1;
    int imax = 0;
    for (int j = 0; j < 3; ++j) {
        for (int i = 1; i < j; ++i) { /* rows 2..j-1 (1-based) accumulate; row 1 is left as is */
            double sum = a[i + 3 * j];
            for (int k = 0; k < i; ++k) sum = sum - a[i + 3 * k] * a[k + 3 * j];
            a[i + 3 * j] = sum;
        }
        double aamax = 0.0;
        for (int i = j; i < 3; ++i) {
            double sum = a[i + 3 * j];
            if (j > 0) {
                for (int k = 0; k < j; ++k) sum = sum - a[i + 3 * k] * a[k + 3 * j];
                a[i + 3 * j] = sum;
            }
            const double dum = vv[i] * fabs(sum);
            if (dum >= aamax) {
                imax = i;
                aamax = dum;
            }
        }
        if (j != imax) {
            for (int k = 0; k < 3; ++k) {
                double dum = a[imax + 3 * k];
                a[imax + 3 * k] = a[j + 3 * k];
                a[j + 3 * k] = dum;
            }
            vv[imax] = vv[j];
        }
        indx[j] = imax;
        if (j != 2) {
            if (a[j + 3 * j] == 0.0) a[j + 3 * j] = tiny;
            const double dum = 1.0 / a[j + 3 * j];
            for (int i = j + 1; i < 3; ++i) a[i + 3 * j] = a[i + 3 * j] * dum;
        }
    }
    if (a[2 + 3 * 2] == 0.0) a[2 + 3 * 2] = tiny;
    return 0;
}

/* best/matin3.f:90-117 (LUBKSB3). */
static void bo_lubksb3(const double *a, const int *indx, double *b) {
    int ii = 0;
    for (int i = 0; i < 3; ++i) {
        const int ll = indx[i];
        double sum = b[ll];
        b[ll] = b[i];
        if (ii != 0) {
            for (int j = ii - 1; j < i; ++j) sum = sum - a[i + 3 * j] * b[j];
        } else if (sum != 0.0) {
            ii = 1;
        }
        b[i] = sum;
    }
    for (int i = 2; i >= 0; --i) {
        double sum = b[i];
        for (int j = i + 1; j < 3; ++j) sum = sum - a[i + 3 * j] * b[j];
        b[i] = sum / a[i + 3 * i];
    }
}

/* best/matin3.f:1-22.  y = inverse(a); a is destroyed.  Returns ierr. */
int bo_matin3(double *a, double *y) {
    int indx[3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) y[i + 3 * j] = 0.0;
        y[i + 3 * i] = 1.0;
    }
    int ierr = bo_ludcmp3(a, indx);
    if (ierr == 1) return ierr;
    for (int j = 0; j < 3; ++j) bo_lubksb3(a, indx, &y[3 * j]);
    return 0;
}

/* best/jacobi.f:1-96 for n = np = 3: eigenvalues d of the symmetric matrix held in the upper
 * triangle of a (column-major, destroyed). */
void bo_jacobi(double *a, double *d) {
    const int n = 3;
    double b[3], z[3], v[9];
#define A_(i, j) a[(i) + 3 * (j)]
#define V_(i, j) v[(i) + 3 * (j)]
    for (int ip = 0; ip < n; ++ip) {
        for (int iq = 0; iq < n; ++iq) V_(ip, iq) = 0.;
        V_(ip, ip) = 1.;
    }
    for (int ip = 0; ip < n; ++ip) {
        b[ip] = A_(ip, ip);
        d[ip] = b[ip];
        z[ip] = 0.;
    }
    for (int it = 1; it <= 50; ++it) {
        double sm = 0.;
        for (int ip = 0; ip < n - 1; ++ip)
            for (int iq = ip + 1; iq < n; ++iq) sm = sm + fabs(A_(ip, iq));
        if (sm == 0.) return;
        const double tresh = it < 4 ? 0.2 * sm / (double)(n * n) : 0.;
        for (int ip = 0; ip < n - 1; ++ip)
            for (int iq = ip + 1; iq < n; ++iq) {
                double g = 100. * fabs(A_(ip, iq));
                if (it > 4 && fabs(d[ip]) + g == fabs(d[ip]) && fabs(d[iq]) + g == fabs(d[iq])) {
                    A_(ip, iq) = 0.;
                } else if (fabs(A_(ip, iq)) > tresh) {
                    double h = d[iq] - d[ip];
                    double t;
                    if (fabs(h) + g == fabs(h)) {
                        t = A_(ip, iq) / h;
                    } else {
                        const double theta = 0.5 * h / A_(ip, iq);
                        t = 1. / (fabs(theta) + sqrt(1. + theta * theta));
                        if (theta < 0.) t = -t;
                    }
                    const double c = 1. / sqrt(1 + t * t);
                    const double s = t * c;
                    const double tau = s / (1. + c);
                    h = t * A_(ip, iq);
                    z[ip] = z[ip] - h;
                    z[iq] = z[iq] + h;
                    d[ip] = d[ip] - h;
                    d[iq] = d[iq] + h;
                    A_(ip, iq) = 0.;
                    for (int j = 0; j < ip; ++j) {
                        g = A_(j, ip);
                        h = A_(j, iq);
                        A_(j, ip) = g - s * (h + g * tau);
                        A_(j, iq) = h + s * (g - h * tau);
                    }
                    for (int j = ip + 1; j < iq; ++j) {
                        g = A_(ip, j);
                        h = A_(j, iq);
                        A_(ip, j) = g - s * (h + g * tau);
                        A_(j, iq) = h + s * (g - h * tau);
                    }
                    for (int j = iq + 1; j < n; ++j) {
                        g = A_(ip, j);
                        h = A_(iq, j);
                        A_(ip, j) = g - s * (h + g * tau);
                        A_(iq, j) = h + s * (g - h * tau);
                    }
                    for (int j = 0; j < n; ++j) {
                        g = V_(j, ip);
                        h = V_(j, iq);
                        V_(j, ip) = g - s * (h + g * tau);
                        V_(j, iq) = h + s * (g - h * tau);
                    }
                }
            }
        for (int ip = 0; ip < n; ++ip) {
            b[ip] = b[ip] + z[ip];
            d[ip] = b[ip];
            z[ip] = 0.;
        }
    }
#undef A_
#undef V_
}

/* ---------------------------------------------------------------- DIFCO */

/* best/difco.f:28-44: b is 3N x 6 column-major. */
void bo_difco_rhs(int nar, const double *x, double *b) {
    const size_t n3 = 3 * (size_t)nar;
    for (int n = 0; n < nar; ++n) {
        const size_t i0 = 3 * (size_t)n;
        for (int mx = 0; mx < 6; ++mx) b[i0 + mx * n3] = b[i0 + 1 + mx * n3] = b[i0 + 2 + mx * n3] = 0.0;
        b[i0 + 0 + 0 * n3] = 1.0;
        b[i0 + 1 + 1 * n3] = 1.0;
        b[i0 + 2 + 2 * n3] = 1.0;
        b[i0 + 1 + 3 * n3] = -x[IDX(n, 2, nar)];
        b[i0 + 2 + 3 * n3] = x[IDX(n, 1, nar)];
        b[i0 + 0 + 4 * n3] = x[IDX(n, 2, nar)];
        b[i0 + 2 + 4 * n3] = -x[IDX(n, 0, nar)];
        b[i0 + 0 + 5 * n3] = -x[IDX(n, 1, nar)];
        b[i0 + 1 + 5 * n3] = x[IDX(n, 0, nar)];
    }
}

/* best/difco.f:51-401.  b holds the six solutions.  Outputs (3x3 column-major): dtt, dtr, drr,
 * urc[3], udr[3]; fric[36] = Ktt, Ktr(raw), Krt, Krr after the shift to the centre of
 * resistance (for the frame checks). */
void bo_difco_post(int nar, const double *x, const double *area, double cg, const double *b, double *dtt,
                   double *dtr, double *drr, double *urc, double *udr, double *fric) {
    const size_t n3 = 3 * (size_t)nar;
    double ktt[9], ktr[9], krt[9], krr[9], uu[9], aux[9], kttinv[9];
#define M_(m, i, j) m[(i) + 3 * (j)]
    for (int mx = 0; mx < 6; ++mx) {
        double u[3] = {0.0, 0.0, 0.0}, om[3] = {0.0, 0.0, 0.0};
        if (mx < 3) u[mx] = 1.0;
        else om[mx - 3] = 1.0;
        double ft[3] = {0.0, 0.0, 0.0}, tt[3] = {0.0, 0.0, 0.0};
        for (int n = 0; n < nar; ++n) {
            const size_t i0 = 3 * (size_t)n;
            const double f1 = b[i0 + mx * n3], f2 = b[i0 + 1 + mx * n3], f3 = b[i0 + 2 + mx * n3];
            const double x1 = x[IDX(n, 0, nar)], x2 = x[IDX(n, 1, nar)], x3 = x[IDX(n, 2, nar)];
            const double q1 = (x2 * f3 - x3 * f2), q2 = (x3 * f1 - x1 * f3), q3 = (x1 * f2 - x2 * f1);
            ft[0] = ft[0] - area[n] * f1;
            ft[1] = ft[1] - area[n] * f2;
            ft[2] = ft[2] - area[n] * f3;
            tt[0] = tt[0] - area[n] * q1;
            tt[1] = tt[1] - area[n] * q2;
            tt[2] = tt[2] - area[n] * q3;
        }
        if (mx < 3) {
            const int i = mx;
            const double temp = 1.0 / u[i];
            for (int j = 0; j < 3; ++j) {
                M_(ktt, j, i) = -ft[j] * temp;
                M_(ktr, j, i) = -tt[j] * temp;
            }
        } else {
            const int i = mx - 3;
            const double temp = (1.0) / om[i];
            for (int j = 0; j < 3; ++j) {
                M_(krt, j, i) = -ft[j] * temp;
                M_(krr, j, i) = -tt[j] * temp;
            }
        }
    }
    if (fric) memcpy(fric + 9, ktr, sizeof ktr);
    /* centre of resistance, :141-174 */
    M_(uu, 0, 0) = M_(ktt, 1, 1) + M_(ktt, 2, 2);
    M_(uu, 1, 1) = M_(ktt, 0, 0) + M_(ktt, 2, 2);
    M_(uu, 2, 2) = M_(ktt, 1, 1) + M_(ktt, 0, 0);
    M_(uu, 0, 1) = -M_(ktt, 0, 1);
    M_(uu, 0, 2) = -M_(ktt, 0, 2);
    M_(uu, 2, 1) = -M_(ktt, 1, 2);
    M_(uu, 1, 0) = M_(uu, 0, 1);
    M_(uu, 1, 2) = M_(uu, 2, 1);
    M_(uu, 2, 0) = M_(uu, 0, 2);
    M_(krt, 0, 2) = (M_(krt, 0, 2) + M_(ktr, 2, 0)) * 0.5;
    M_(krt, 2, 0) = (M_(krt, 2, 0) + M_(ktr, 0, 2)) * 0.5;
    M_(krt, 0, 1) = (M_(krt, 0, 1) + M_(ktr, 1, 0)) * 0.5;
    M_(krt, 1, 0) = (M_(krt, 1, 0) + M_(ktr, 0, 1)) * 0.5;
    M_(krt, 1, 2) = (M_(krt, 1, 2) + M_(ktr, 2, 1)) * 0.5;
    M_(krt, 2, 1) = (M_(krt, 2, 1) + M_(ktr, 1, 2)) * 0.5;
    M_(krt, 0, 0) = (M_(krt, 0, 0) + M_(ktr, 0, 0)) * 0.5;
    M_(krt, 1, 1) = (M_(krt, 1, 1) + M_(ktr, 1, 1)) * 0.5;
    M_(krt, 2, 2) = (M_(krt, 2, 2) + M_(ktr, 2, 2)) * 0.5;
    double sur1 = -M_(krt, 2, 1) + M_(krt, 1, 2);
    double sur2 = -M_(krt, 0, 2) + M_(krt, 2, 0);
    double sur3 = -M_(krt, 1, 0) + M_(krt, 0, 1);
    bo_matin3(uu, aux);
    for (int j = 0; j < 3; ++j) urc[j] = (M_(aux, j, 0) * sur1 + M_(aux, j, 1) * sur2 + M_(aux, j, 2) * sur3);
    M_(uu, 2, 0) = -urc[1];
    M_(uu, 2, 1) = urc[0];
    M_(uu, 0, 1) = -urc[2];
    M_(uu, 1, 0) = -M_(uu, 0, 1);
    M_(uu, 1, 2) = -M_(uu, 2, 1);
    M_(uu, 0, 2) = -M_(uu, 2, 0);
    M_(uu, 0, 0) = 0.0;
    M_(uu, 1, 1) = 0.0;
    M_(uu, 2, 2) = 0.0;
    /* shift Krr, Krt to the centre of resistance, :196-214 */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double temp = 0.0;
            for (int k = 0; k < 3; ++k)
                temp = temp + M_(uu, i, k) * (M_(ktt, k, 0) * M_(uu, 0, j) + M_(ktt, k, 1) * M_(uu, 1, j) +
                                              M_(ktt, k, 2) * M_(uu, 2, j));
            const double temp1 = M_(krt, 0, i) * M_(uu, 0, j) + M_(krt, 1, i) * M_(uu, 1, j) + M_(krt, 2, i) * M_(uu, 2, j);
            const double temp2 = M_(uu, i, 0) * M_(krt, 0, j) + M_(uu, i, 1) * M_(krt, 1, j) + M_(uu, i, 2) * M_(krt, 2, j);
            M_(krr, i, j) = M_(krr, i, j) - temp + temp1 - temp2;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            M_(krt, i, j) = M_(krt, i, j) + M_(ktt, i, 0) * M_(uu, 0, j) + M_(ktt, i, 1) * M_(uu, 1, j) +
                            M_(ktt, i, 2) * M_(uu, 2, j);
    if (fric) {
        memcpy(fric, ktt, sizeof ktt);
        memcpy(fric + 18, krt, sizeof krt);
        memcpy(fric + 27, krr, sizeof krr);
    }
    /* Dtt, Drr, :244-330 */
    double tmpm[9];
    memcpy(tmpm, ktt, sizeof tmpm);
    bo_matin3(tmpm, kttinv); /* matin3 destroys its input: ktt itself is reused below as storage */
    memcpy(ktt, tmpm, sizeof tmpm);
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) M_(ktt, i, j) = M_(krr, i, j);
    bo_matin3(ktt, aux);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double temp = 0.0;
            for (int k = 0; k < 3; ++k)
                temp = temp + M_(krt, i, k) * (M_(kttinv, k, 0) * M_(krt, j, 0) + M_(kttinv, k, 1) * M_(krt, j, 1) +
                                               M_(kttinv, k, 2) * M_(krt, j, 2));
            M_(ktt, i, j) = M_(krr, i, j) - temp;
        }
    bo_matin3(ktt, drr);
    if (cg == 0)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M_(drr, i, j) = 4.0 * M_(drr, i, j) / 3.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double temp = 0.0;
            for (int k = 0; k < 3; ++k)
                temp = temp + M_(kttinv, i, k) * (M_(krt, 0, k) * M_(drr, 0, j) + M_(krt, 1, k) * M_(drr, 1, j) +
                                                  M_(krt, 2, k) * M_(drr, 2, j));
            M_(dtr, i, j) = -temp;
            if (cg == 0) M_(dtr, i, j) = 4.0 * M_(dtr, i, j) / 3.0;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double temp = 0.0;
            for (int k = 0; k < 3; ++k)
                temp = temp + M_(kttinv, i, k) * (M_(krt, 0, k) * M_(dtr, j, 0) + M_(krt, 1, k) * M_(dtr, j, 1) +
                                                  M_(krt, 2, k) * M_(dtr, j, 2));
            M_(dtt, i, j) = M_(kttinv, i, j) - temp;
            if (cg == 0) M_(dtt, i, j) = 4.0 * M_(dtt, i, j) / 3.0;
        }
    /* centre of diffusion, :344-401 */
    M_(uu, 0, 0) = M_(drr, 1, 1) + M_(drr, 2, 2);
    M_(uu, 1, 1) = M_(drr, 0, 0) + M_(drr, 2, 2);
    M_(uu, 2, 2) = M_(drr, 1, 1) + M_(drr, 0, 0);
    M_(uu, 0, 1) = -M_(drr, 0, 1);
    M_(uu, 0, 2) = -M_(drr, 0, 2);
    M_(uu, 2, 1) = -M_(drr, 2, 1);
    M_(uu, 1, 0) = M_(uu, 0, 1);
    M_(uu, 1, 2) = M_(uu, 2, 1);
    M_(uu, 2, 0) = M_(uu, 0, 2);
    sur1 = M_(dtr, 2, 1) - M_(dtr, 1, 2);
    sur2 = M_(dtr, 0, 2) - M_(dtr, 2, 0);
    sur3 = M_(dtr, 1, 0) - M_(dtr, 0, 1);
    bo_matin3(uu, aux);
    for (int j = 0; j < 3; ++j) udr[j] = (M_(aux, j, 0) * sur1 + M_(aux, j, 1) * sur2 + M_(aux, j, 2) * sur3);
    M_(uu, 2, 0) = -udr[1];
    M_(uu, 2, 1) = udr[0];
    M_(uu, 0, 1) = -udr[2];
    M_(uu, 1, 0) = -M_(uu, 0, 1);
    M_(uu, 1, 2) = -M_(uu, 2, 1);
    M_(uu, 0, 2) = -M_(uu, 2, 0);
    M_(uu, 0, 0) = 0.0;
    M_(uu, 1, 1) = 0.0;
    M_(uu, 2, 2) = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double temp = 0.0;
            for (int k = 0; k < 3; ++k)
                temp = temp + M_(uu, i, k) * M_(drr, k, 0) * M_(uu, 0, j) + M_(uu, i, k) * M_(drr, k, 1) * M_(uu, 1, j) +
                       M_(uu, i, k) * M_(drr, k, 2) * M_(uu, 2, j);
            const double temp1 = M_(dtr, i, 0) * M_(uu, 0, j) + M_(dtr, i, 1) * M_(uu, 1, j) + M_(dtr, i, 2) * M_(uu, 2, j);
            const double temp2 = -M_(uu, i, 0) * M_(dtr, j, 0) - M_(uu, i, 1) * M_(dtr, j, 1) - M_(uu, i, 2) * M_(dtr, j, 2);
            M_(dtt, i, j) = M_(dtt, i, j) - temp + temp1 + temp2;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            M_(dtr, i, j) = M_(dtr, i, j) - M_(uu, i, 0) * M_(drr, 0, j) - M_(uu, i, 1) * M_(drr, 1, j) -
                            M_(uu, i, 2) * M_(drr, 2, j);
#undef M_
}

/* ---------------------------------------------------------------- VISCO */

/* best/visco.f:216-243: v0 is 3N x 5 column-major. */
void bo_strain(int nar, const double *x, double *v0) {
    const size_t n3 = 3 * (size_t)nar;
    for (int n = 0; n < nar; ++n) {
        const size_t i0 = 3 * (size_t)n;
        const double x1 = x[IDX(n, 0, nar)], x2 = x[IDX(n, 1, nar)], x3 = x[IDX(n, 2, nar)];
        v0[i0 + 0 + 0 * n3] = 0.5 * x2;
        v0[i0 + 1 + 0 * n3] = 0.5 * x1;
        v0[i0 + 2 + 0 * n3] = 0.0;
        v0[i0 + 0 + 1 * n3] = 0.5 * x3;
        v0[i0 + 1 + 1 * n3] = 0.0;
        v0[i0 + 2 + 1 * n3] = 0.5 * x1;
        v0[i0 + 0 + 2 * n3] = 0.0;
        v0[i0 + 1 + 2 * n3] = 0.5 * x3;
        v0[i0 + 2 + 2 * n3] = 0.5 * x2;
        v0[i0 + 0 + 3 * n3] = 0.5 * x1;
        v0[i0 + 1 + 3 * n3] = -0.5 * x2;
        v0[i0 + 2 + 3 * n3] = 0.0;
        v0[i0 + 0 + 4 * n3] = 0.5 * x1;
        v0[i0 + 1 + 4 * n3] = 0.0;
        v0[i0 + 2 + 4 * n3] = -0.5 * x3;
    }
}

/* best/visco.f:25-211.  v0 holds the five strain solutions (modified with icov = 1), b the
 * DIFCO solutions (columns 1-3 used), x the incenters (a private copy is shifted with icov).
 * Returns chiprime; d[3] the centre of viscosity (zeros without icov). */
double bo_visco_post(int nar, const double *xin, const double *area, int icov, double cg, const double *b, double *v0,
                     double *dout) {
    const double pi = 3.1415926535897932385;
    const size_t n3 = 3 * (size_t)nar;
    double *x = (double *)malloc(sizeof(double) * 3 * (size_t)nar);
    memcpy(x, xin, sizeof(double) * 3 * (size_t)nar);
    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
    if (icov == 1) {
        /* G(n,l,k) = 2 v0(3n+k, l); H(n,l,k,m) from the translation solutions, :36-61 */
        double bbm[9] = {0}, c1 = 0.0, c2 = 0.0, c3 = 0.0;
#define G_(l, k) (2.0 * v0[i0 + (k)-1 + ((l)-1) * n3])
        for (int n = 0; n < nar; ++n) {
            const size_t i0 = 3 * (size_t)n;
            double H[5][3][3]; /* H[l][k][m], 0-based */
            for (int k = 0; k < 3; ++k) {
                const double e1 = b[i0 + k + 0 * n3], e2 = b[i0 + k + 1 * n3], e3 = b[i0 + k + 2 * n3];
                H[0][k][0] = e2;
                H[0][k][1] = e1;
                H[0][k][2] = 0.0;
                H[1][k][0] = e3;
                H[1][k][1] = 0.0;
                H[1][k][2] = e1;
                H[2][k][0] = 0.0;
                H[2][k][1] = e3;
                H[2][k][2] = e2;
                H[3][k][0] = e1;
                H[3][k][1] = -e2;
                H[3][k][2] = 0.0;
                H[4][k][0] = e1;
                H[4][k][1] = 0.0;
                H[4][k][2] = -e3;
            }
#define H_(l, k, m) H[(l)-1][(k)-1][(m)-1]
            const double ar = area[n];
            bbm[0] = bbm[0] + ar * (3 * (H_(1, 2, 1) + H_(2, 3, 1)) + 2 * (H_(4, 1, 1) + H_(5, 1, 1)));
            bbm[3] = bbm[3] + ar * (-4 * H_(4, 2, 1) + 3 * (H_(1, 1, 1) + H_(1, 2, 2) + H_(2, 3, 2) + H_(3, 3, 1)) +
                                    2 * (H_(4, 1, 2) + H_(5, 1, 2) + H_(5, 2, 1)));
            bbm[6] = bbm[6] + ar * (-4 * H_(5, 3, 1) + 3 * (H_(1, 2, 3) + H_(2, 1, 1) + H_(2, 3, 3) + H_(3, 2, 1)) +
                                    2 * (H_(4, 1, 3) + H_(4, 3, 1) + H_(5, 1, 3)));
            bbm[4] = bbm[4] + ar * (-4 * H_(4, 2, 2) + 3 * (H_(1, 1, 2) + H_(3, 3, 2)) + 2 * H_(5, 2, 2));
            bbm[7] = bbm[7] + ar * (-4 * H_(4, 2, 3) + 3 * (H_(3, 2, 2) + H_(3, 3, 3) + H_(1, 1, 3) + H_(2, 1, 2)) +
                                    2 * (H_(4, 3, 2) + H_(5, 2, 3)));
            bbm[8] = bbm[8] + ar * (-4 * H_(5, 3, 3) + 3 * (H_(2, 1, 3) + H_(3, 2, 3)) + 2 * H_(4, 3, 3));
            const double x1 = x[IDX(n, 0, nar)], x2 = x[IDX(n, 1, nar)], x3 = x[IDX(n, 2, nar)];
            c1 = c1 + ar * (-3 * G_(1, 2) - 3 * G_(2, 3) - 2 * G_(4, 1) - 2 * G_(5, 1) -
                            x1 * (3 * (H_(1, 2, 1) + H_(2, 3, 1)) + 2 * (H_(4, 1, 1) + H_(5, 1, 1))) -
                            x2 * (3 * (H_(1, 1, 1) + H_(3, 3, 1)) - 4 * H_(4, 2, 1) + 2 * H_(5, 2, 1)) -
                            x3 * (3 * (H_(2, 1, 1) + H_(3, 2, 1)) - 4 * H_(5, 3, 1) + 2 * H_(4, 3, 1)));
            c2 = c2 + ar * (-3 * G_(1, 1) - 3 * G_(3, 3) + 4 * G_(4, 2) - 2 * G_(5, 2) -
                            x1 * (3 * (H_(1, 2, 2) + H_(2, 3, 2)) + 2 * (H_(4, 1, 2) + H_(5, 1, 2))) -
                            x2 * (3 * (H_(1, 1, 2) + H_(3, 3, 2)) - 4 * H_(4, 2, 2) + 2 * H_(5, 2, 2)) -
                            x3 * (3 * (H_(2, 1, 2) + H_(3, 2, 2)) - 4 * H_(5, 3, 2) + 2 * H_(4, 3, 2)));
            c3 = c3 + ar * (-3 * G_(2, 1) - 3 * G_(3, 2) - 2 * G_(4, 3) + 4 * G_(5, 3) -
                            x1 * (3 * (H_(1, 2, 3) + H_(2, 3, 3)) + 2 * (H_(4, 1, 3) + H_(5, 1, 3))) -
                            x2 * (3 * (H_(1, 1, 3) + H_(3, 3, 3)) - 4 * H_(4, 2, 3) + 2 * H_(5, 2, 3)) -
                            x3 * (3 * (H_(2, 1, 3) + H_(3, 2, 3)) - 4 * H_(5, 3, 3) + 2 * H_(4, 3, 3)));
#undef H_
        }
#undef G_
        /* bbm column-major: bbm[i + 3j] = BB(i+1, j+1) */
        bbm[0] = bbm[0] / 15.;
        bbm[3] = bbm[3] / 30.;
        bbm[6] = bbm[6] / 30.;
        bbm[1] = bbm[3];
        bbm[4] = bbm[4] / 15.;
        bbm[7] = bbm[7] / 30.;
        bbm[2] = bbm[6];
        bbm[5] = bbm[7];
        bbm[8] = bbm[8] / 15.;
        c1 = c1 / 30.;
        c2 = c2 / 30.;
        c3 = c3 / 30.;
        double bbi[9];
        bo_matin3(bbm, bbi);
        d1 = bbi[0] * c1 + bbi[3] * c2 + bbi[6] * c3;
        d2 = bbi[1] * c1 + bbi[4] * c2 + bbi[7] * c3;
        d3 = bbi[2] * c1 + bbi[5] * c2 + bbi[8] * c3;
        for (int n = 0; n < nar; ++n) {
            x[IDX(n, 0, nar)] = x[IDX(n, 0, nar)] - d1;
            x[IDX(n, 1, nar)] = x[IDX(n, 1, nar)] - d2;
            x[IDX(n, 2, nar)] = x[IDX(n, 2, nar)] - d3;
        }
        for (size_t i = 0; i < n3; ++i) {
            const double e1 = b[i], e2 = b[i + n3], e3 = b[i + 2 * n3];
            v0[i + 0 * n3] = v0[i + 0 * n3] - 0.5 * (d2 * e1 + d1 * e2);
            v0[i + 1 * n3] = v0[i + 1 * n3] - 0.5 * (d3 * e1 + d1 * e3);
            v0[i + 2 * n3] = v0[i + 2 * n3] - 0.5 * (d3 * e2 + d2 * e3);
            v0[i + 3 * n3] = v0[i + 3 * n3] - 0.5 * (d1 * e1 - d2 * e2);
            v0[i + 4 * n3] = v0[i + 4 * n3] - 0.5 * (d1 * e1 - d3 * e3);
        }
    }
    /* :171-211 */
    double c112 = 0.0, c213 = 0.0, c323 = 0.0, c411 = 0.0, c422 = 0.0, c433 = 0.0, c511 = 0.0, c522 = 0.0,
           c533 = 0.0;
    for (int n = 0; n < nar; ++n) {
        const size_t i0 = 3 * (size_t)n;
        const double x1 = x[IDX(n, 0, nar)], x2 = x[IDX(n, 1, nar)], x3 = x[IDX(n, 2, nar)];
        const double ar = area[n];
#define W_(l, c) v0[i0 + (c)-1 + ((l)-1) * n3]
        c112 = c112 + ar * (W_(1, 1) * x2 + W_(1, 2) * x1);
        c213 = c213 + ar * (W_(2, 1) * x3 + W_(2, 3) * x1);
        c323 = c323 + ar * (W_(3, 2) * x3 + W_(3, 3) * x2);
        c411 = c411 + ar * W_(4, 1) * x1;
        c422 = c422 + ar * W_(4, 2) * x2;
        c433 = c433 + ar * W_(4, 3) * x3;
        c511 = c511 + ar * W_(5, 1) * x1;
        c522 = c522 + ar * W_(5, 2) * x2;
        c533 = c533 + ar * W_(5, 3) * x3;
#undef W_
    }
    c112 = 0.5 * c112;
    c213 = 0.5 * c213;
    c323 = 0.5 * c323;
    double chip = 8 * pi * (0.2 * (c112 + c213 + c323) + (c411 - 2 * c422 + c433 + c511 + c522 - 2 * c533) / 15.0);
    if (cg == 0) chip = 0.75 * chip;
    dout[0] = d1;
    dout[1] = d2;
    dout[2] = d3;
    free(x);
    return chip;
}

/* ---------------------------------------------------------------- units / printed scalars */

/* best/output.f:16-29, :98-167.  From the 3x3 tensors: out[0] = 1/3 Trace DRR(EXACT) (1/s),
 * out[1] = 1/3 Trace DTT(EXACT) (cm^2/s), out[2..4] Drr eigenvalues (1/A^3), out[5..7] Dtt
 * eigenvalues (1/A), out[8] = tunit, out[9] = runit.  drr and dtt are copied, not modified. */
void bo_units(const double *drr_in, const double *dtt_in, double *out) {
    const double temperature = 293.15, eta = 0.01002, pi = 3.141592653;
    const double tunit = (1.38066E-16) * temperature * 1.0E8 / (8.0 * eta * pi);
    const double runit = tunit * 1.0e16;
    double a[9], d[3];
    memcpy(a, drr_in, sizeof a);
    bo_jacobi(a, d);
    double trace = (d[0] + d[1] + d[2]) / 3.;
    out[0] = runit * trace;
    out[2] = d[0];
    out[3] = d[1];
    out[4] = d[2];
    memcpy(a, dtt_in, sizeof a);
    bo_jacobi(a, d);
    trace = (d[0] + d[1] + d[2]) / 3.;
    out[1] = tunit * trace;
    out[5] = d[0];
    out[6] = d[1];
    out[7] = d[2];
    out[8] = tunit;
    out[9] = runit;
}
