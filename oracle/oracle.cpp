// FP64 CPU oracle for the metadynamics CV-and-bias path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
// calls this file: it is the checker used by tests/, by the smoke check and by
// the `cpu_baseline` leg of bench.py.  It restates the reference algorithms in
// plain double precision with FP contraction disabled (built with
// -ffp-contract=off), so every r^2 below is ((dx*dx + dy*dy) + dz*dz) exactly.
//
// Reference files followed (paths relative to the Metad-toy source tree):
//   selection      src/cvs/CVs_Steinhardt.cu:1324-1406  (get_environment_Steinhardt_Q)
//   Y_lm, q_lm     include/CVsets/CV_Steinhardt_math.h:981-1215, src/cvs/CVs_Steinhardt.cu:1410-1532
//   Q_l.AVE        src/cvs/CVs_Steinhardt.cu:824-839, src/utils/CVs_tools.cu:6-18 (Kahan)
//   grid           src/grids/GH_t0_uniformGrid.cu:73-98 (radius), 194-208 (WT height),
//                  226-282 (deposit), 288-461 (Catmull-Rom value/derivative, clamp)
//   chain rule     src/grids/DimSet.cu:83-97 (dV/ds * ds/dv per symbol)
// Parts with no reference implementation (the reference's own gradient is not
// the gradient of its CV, and it has no Lechner-Dellago local Q, no vector-k
// S(k), no hills-sum bias and no MLP) follow their mathematical definitions;
// every such function names the test that pins it in its comment.
//
// Pins (tests/test_oracle.py):
//   Y_lm            scipy.special.sph_harm_y at seeded directions + sum rule
//   q_l             closed-form lattice values (FCC/BCC/SC/HCP)
//   selection       brute force with engineered ties (literal insertion model)
//   gradients       central finite differences of the oracle's own CVs,
//                   translation and rotation invariance
//   grid            closed-form Gaussian sums, Catmull-Rom exactness on
//                   quadratics, reference HILLS fixture
//   S(k)            |A|=N on a perfect lattice at a reciprocal vector, FD
//   MLP             torch float64 autograd on the same weights, FD
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <algorithm>

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Y_lm restated from the reference closed forms (CV_Steinhardt_math.h:981-1215)
// Angles follow CVs_Steinhardt.cu:1449-1480: theta = acos(dz/r), phi = atan2(dy,dx),
// multi-angle sincos.  Output layout per m = 0..l: (Re, Im).
// ---------------------------------------------------------------------------
inline double p2(double v) { return v * v; }
inline double p3(double v) { return v * v * v; }
inline double p4(double v) { return p2(v) * p2(v); }
inline double p5(double v) { return p4(v) * v; }
inline double p6(double v) { return p3(v) * p3(v); }

void ylm_reference(int l, double dx, double dy, double dz, double* Y) {
    double r = std::sqrt(dx * dx + dy * dy + dz * dz);
    double c = dz / r;
    if (c > 1.0) c = 1.0;   // acos domain guard (the reference would return NaN)
    if (c < -1.0) c = -1.0;
    double theta = std::acos(c);
    double phi = std::atan2(dy, dx);
    double st = std::sin(theta), ct = std::cos(theta);
    double sp[7], cp[7], s_t[7], c_t[7];
    for (int k = 1; k <= 6; ++k) {
        sp[k] = std::sin(k * phi); cp[k] = std::cos(k * phi);
        s_t[k] = std::sin(k * theta); c_t[k] = std::cos(k * theta);
    }
    double t;
    if (l == 3) {
        t = 0.1865881662950577 * ct * (-1. + 5. * c_t[2]);
        Y[0] = t; Y[1] = 0.0;
        t = -0.08079504602853766 * (st + 5. * s_t[3]);
        Y[2] = t * cp[1]; Y[3] = t * sp[1];
        t = 1.021985476433282 * ct * p2(st);
        Y[4] = t * cp[2]; Y[5] = t * sp[2];
        t = -0.4172238236327841 * p3(st);
        Y[6] = t * cp[3]; Y[7] = t * sp[3];
    } else if (l == 4) {
        Y[0] = 0.10578554691520430 * (3 - 30 * p2(ct) + 35 * p4(ct)); Y[1] = 0;
        t = 0.47308734787878000 * ct * (-3 + 7 * p2(ct)) * st;
        Y[2] = -t * cp[1]; Y[3] = -t * sp[1];
        t = 0.33452327177864458 * (-1 + 7 * p2(ct)) * p2(st);
        Y[4] = t * cp[2]; Y[5] = t * sp[2];
        t = 1.25167147089835227 * ct * p3(st);
        Y[6] = -t * cp[3]; Y[7] = -t * sp[3];
        t = 0.44253269244498263 * p4(st);
        Y[8] = t * cp[4]; Y[9] = t * sp[4];
    } else if (l == 6) {
        Y[0] = 0.001986537570863388 * (50. + 105. * c_t[2] + 126. * c_t[4] + 231. * c_t[6]); Y[1] = 0.0;
        t = -0.01287423488441065 * (5. * s_t[2] + 12. * s_t[4] + 33. * s_t[6]);
        Y[2] = t * cp[1]; Y[3] = t * sp[1];
        t = 0.04071190536673223 * (35. + 60. * c_t[2] + 33. * c_t[4]) * p2(st);
        Y[4] = t * cp[2]; Y[5] = t * sp[2];
        t = -0.3256952429338579 * ct * (5. + 11. * c_t[2]) * p3(st);
        Y[6] = t * cp[3]; Y[7] = t * sp[3];
        t = 0.178390631426999 * (9. + 11. * c_t[2]) * p4(st);
        Y[8] = t * cp[4]; Y[9] = t * sp[4];
        t = -1.673452458100098 * ct * p5(st);
        Y[10] = t * cp[5]; Y[11] = t * sp[5];
        t = 0.4830841135800662 * p6(st);
        Y[12] = t * cp[6]; Y[13] = t * sp[6];
    }
}

// ---------------------------------------------------------------------------
// Gradient of Y_lm with respect to the pair vector r (exact, no reference
// counterpart: the reference contraction drops terms, see the survey's F2).
// Y_lm(r) = F(n), n = r/|r|, F(n) = c_lm(n_z) (n_x + i n_y)^m with
// c_lm(u) = N_lm (-1)^m d^m P_l/du^m (Condon-Shortley phase, same as above).
// grad_r Y = (I - n n^T) grad_n F / |r|.   Pinned by finite differences of
// ylm_reference in tests/test_oracle.py::test_ylm_gradient_fd.
// ---------------------------------------------------------------------------
struct PolyTable {
    // coef[l][m][k]: coefficient of u^k in c_lm(u)
    double coef[7][7][7];
    PolyTable() {
        std::memset(coef, 0, sizeof(coef));
        for (int l = 0; l <= 6; ++l) {
            // Legendre P_l(u) = 2^-l sum_k (-1)^k C(l,k) C(2l-2k,l) u^(l-2k)
            double P[13] = {0};
            for (int k = 0; 2 * k <= l; ++k) {
                double v = std::pow(-1.0, k) * binom(l, k) * binom(2 * l - 2 * k, l) / std::pow(2.0, l);
                P[l - 2 * k] += v;
            }
            for (int m = 0; m <= l; ++m) {
                double D[13];
                for (int k = 0; k <= l; ++k) D[k] = P[k];
                for (int d = 0; d < m; ++d) {
                    for (int k = 0; k < l; ++k) D[k] = D[k + 1] * (k + 1);
                    D[l] = 0;
                }
                double N = std::sqrt((2 * l + 1) / (4 * kPi) * fact(l - m) / fact(l + m));
                double sgn = (m % 2) ? -1.0 : 1.0;
                for (int k = 0; k <= l; ++k) coef[l][m][k] = sgn * N * D[k];
            }
        }
    }
    static double binom(int n, int k) { return fact(n) / (fact(k) * fact(n - k)); }
    static double fact(int n) { double f = 1; for (int i = 2; i <= n; ++i) f *= i; return f; }
};
const PolyTable& ptab() { static PolyTable t; return t; }

// dY[m*6 + {0..5}] = (dRe/dx, dIm/dx, dRe/dy, dIm/dy, dRe/dz, dIm/dz)
void ylm_grad(int l, double dx, double dy, double dz, double* dY) {
    const PolyTable& T = ptab();
    double r = std::sqrt(dx * dx + dy * dy + dz * dz);
    double nx = dx / r, ny = dy / r, u = dz / r;
    double xr[8], xi[8];          // xi^m = (nx + i ny)^m
    xr[0] = 1; xi[0] = 0;
    for (int m = 1; m <= l; ++m) {
        xr[m] = xr[m - 1] * nx - xi[m - 1] * ny;
        xi[m] = xr[m - 1] * ny + xi[m - 1] * nx;
    }
    for (int m = 0; m <= l; ++m) {
        double c = 0, cd = 0, up = 1;
        for (int k = 0; k <= l; ++k) {
            c += T.coef[l][m][k] * up;
            if (k + 1 <= l) cd += T.coef[l][m][k + 1] * (k + 1) * up;
            up *= u;
        }
        // grad_n F = (c m xi^(m-1), i c m xi^(m-1), c' xi^m)
        double gx_r = 0, gx_i = 0, gy_r = 0, gy_i = 0;
        if (m > 0) {
            gx_r = c * m * xr[m - 1]; gx_i = c * m * xi[m - 1];
            gy_r = -gx_i;             gy_i = gx_r;
        }
        double gz_r = cd * xr[m], gz_i = cd * xi[m];
        double dot_r = nx * gx_r + ny * gy_r + u * gz_r;
        double dot_i = nx * gx_i + ny * gy_i + u * gz_i;
        dY[m * 6 + 0] = (gx_r - dot_r * nx) / r;
        dY[m * 6 + 1] = (gx_i - dot_i * nx) / r;
        dY[m * 6 + 2] = (gy_r - dot_r * ny) / r;
        dY[m * 6 + 3] = (gy_i - dot_i * ny) / r;
        dY[m * 6 + 4] = (gz_r - dot_r * u) / r;
        dY[m * 6 + 5] = (gz_i - dot_i * u) / r;
    }
}

inline int owner_of(int j, int n_local, const int* ghost_owner) {
    return j < n_local ? j : ghost_owner[j - n_local];
}

}  // namespace

extern "C" {

// Oracle build tag: lets a test confirm it loaded this file's build.
int orc_version() { return 1; }
// threads an OpenMP region of the oracle runs on (1 without OpenMP)
int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_ylm(int l, double dx, double dy, double dz, double* Y) { ylm_reference(l, dx, dy, dz, Y); }
void orc_ylm_grad(int l, double dx, double dy, double dz, double* dY) { ylm_grad(l, dx, dy, dz, dY); }
void orc_ylm_poly(int l, double dx, double dy, double dz, double* Y) {
    // polynomial form of the same function (used only to cross-check the table)
    const PolyTable& T = ptab();
    double r = std::sqrt(dx * dx + dy * dy + dz * dz);
    double nx = dx / r, ny = dy / r, u = dz / r;
    double zr = 1, zi = 0;
    for (int m = 0; m <= l; ++m) {
        double c = 0, up = 1;
        for (int k = 0; k <= l; ++k) { c += T.coef[l][m][k] * up; up *= u; }
        Y[2 * m] = c * zr; Y[2 * m + 1] = c * zi;
        double tr = zr * nx - zi * ny, ti = zr * ny + zi * nx;
        zr = tr; zi = ti;
    }
}
void orc_ylm_coef(double* out /*7*7*7*/) { std::memcpy(out, ptab().coef, sizeof(double) * 343); }

// ---------------------------------------------------------------------------
// N-nearest selection: literal restatement of get_environment_Steinhardt_Q
// (src/cvs/CVs_Steinhardt.cu:1324-1406).  For every center c (local index
// centers[c]) scan its CSR list in order, skip r2 > rc2 or r2 < 1e-12, count,
// and insert into N slots with strict '>' carrying the displaced entry
// forward.  Slots start at r2 = box_sum2 = (Lx+Ly+Lz)^2 with tag -1.
// Outputs: sel_idx[c*N+s] (index into x, -1 = empty), sel_pos[c*N+s] (position
// in the center's list, -1 = empty), sel_d[(c*N+s)*4 + {dx,dy,dz,r2}],
// count_rc[c], n_sel[c] = min(count_rc, N).
// ---------------------------------------------------------------------------
void orc_select(int n_center, const int* centers, const int* offsets, const int* idx,
                const double* x, double rc2, int N, double box_sum2,
                int* sel_idx, int* sel_pos, double* sel_d, int* count_rc, int* n_sel) {
    // centers are independent: one thread per run of centers, each center's scan serial
    #pragma omp parallel for schedule(dynamic, 256)
    for (int c = 0; c < n_center; ++c) {
        int i = centers[c];
        double cx = x[3 * i], cy = x[3 * i + 1], cz = x[3 * i + 2];
        int* sidx = sel_idx + (size_t)c * N;
        int* spos = sel_pos + (size_t)c * N;
        double* sd = sel_d + (size_t)c * N * 4;
        for (int s = 0; s < N; ++s) { sd[4 * s + 3] = box_sum2; sidx[s] = -1; spos[s] = -1;
                                      sd[4 * s] = sd[4 * s + 1] = sd[4 * s + 2] = 0.0; }
        int cnt = 0;
        for (int e = offsets[c]; e < offsets[c + 1]; ++e) {
            int j = idx[e];
            double ddx = x[3 * j] - cx, ddy = x[3 * j + 1] - cy, ddz = x[3 * j + 2] - cz;
            double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            if (r2 > rc2 || r2 < 1e-12) continue;
            ++cnt;
            int tag = j, pos = e - offsets[c];
            for (int s = 0; s < N; ++s) {
                if (sd[4 * s + 3] > r2) {
                    double tx = sd[4 * s], ty = sd[4 * s + 1], tz = sd[4 * s + 2], tr = sd[4 * s + 3];
                    int tt = sidx[s], tp = spos[s];
                    sd[4 * s] = ddx; sd[4 * s + 1] = ddy; sd[4 * s + 2] = ddz; sd[4 * s + 3] = r2;
                    sidx[s] = tag; spos[s] = pos;
                    ddx = tx; ddy = ty; ddz = tz; r2 = tr; tag = tt; pos = tp;
                }
            }
        }
        count_rc[c] = cnt;
        n_sel[c] = cnt >= N ? N : cnt;
    }
}

// ---------------------------------------------------------------------------
// q_lm and q_l (CVs_Steinhardt.cu:1410-1532): q_lm = (1/n) sum_j Y_lm(r_ij),
// q_l = sqrt(4pi/(2l+1) (|q_l0|^2 + 2 sum_{m>0} |q_lm|^2)); n = 0 gives 0.
// qlm[c*2(l+1) + 2m + {0,1}], ql[c].
// ---------------------------------------------------------------------------
void orc_steinhardt(int l, int n_center, int N, const double* sel_d, const int* n_sel,
                    double* qlm, double* ql) {
    const int W = 2 * (l + 1);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int c = 0; c < n_center; ++c) {
        double Y[2 * 13];
        double* q = qlm + (size_t)c * W;
        for (int k = 0; k < W; ++k) q[k] = 0.0;
        int n = n_sel[c];
        if (n == 0) { ql[c] = 0.0; continue; }
        for (int s = 0; s < n; ++s) {
            const double* d = sel_d + ((size_t)c * N + s) * 4;
            ylm_reference(l, d[0], d[1], d[2], Y);
            for (int k = 0; k < W; ++k) q[k] += Y[k];
        }
        double inv = 1.0 / (double)n;
        for (int k = 0; k < W; ++k) q[k] *= inv;
        double s2 = q[0] * q[0];
        for (int m = 1; m <= l; ++m) s2 += 2.0 * (q[2 * m] * q[2 * m] + q[2 * m + 1] * q[2 * m + 1]);
        ql[c] = std::sqrt(s2 * 12.56637061435917295385 / double(2 * l + 1));
    }
}

// Kahan mean over flagged entries (CVs_tools.cu:6-18): sum / count.
double orc_kahan_mean(int n, const double* v, const unsigned char* flag, double count) {
    double sum = 0.0, comp = 0.0;
    for (int i = 0; i < n; ++i) {
        if (flag && !flag[i]) continue;
        double y = v[i] - comp;
        double t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    return sum / count;
}

// ---------------------------------------------------------------------------
// Lechner-Dellago local Q_l over the same selection (no reference counterpart;
// the reference's "L" type is switching-weighted, CVs_Steinhardt_Local.cu:1489-1580).
// qbar_lm(i) = (q_lm(i) + sum_{j in S(i)} q_lm(owner(j))) / (n_i + 1)
// lq(i) = sqrt(4pi/(2l+1) sum_m |qbar_lm|^2).  Centers are local atoms
// 0..n_center-1 (center index == local index).  Pinned by: perfect lattice
// (lq == q) and finite differences of its gradient.
// ---------------------------------------------------------------------------
void orc_ld_local(int l, int n_center, int N, int n_local, const int* ghost_owner,
                  const int* sel_idx, const int* n_sel, const double* qlm,
                  double* qbar, double* lq) {
    const int W = 2 * (l + 1);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n_center; ++i) {
        double* qb = qbar + (size_t)i * W;
        for (int k = 0; k < W; ++k) qb[k] = qlm[(size_t)i * W + k];
        int n = n_sel[i];
        for (int s = 0; s < n; ++s) {
            int o = owner_of(sel_idx[(size_t)i * N + s], n_local, ghost_owner);
            for (int k = 0; k < W; ++k) qb[k] += qlm[(size_t)o * W + k];
        }
        double inv = 1.0 / (double)(n + 1);
        for (int k = 0; k < W; ++k) qb[k] *= inv;
        double s2 = qb[0] * qb[0];
        for (int m = 1; m <= l; ++m) s2 += 2.0 * (qb[2 * m] * qb[2 * m] + qb[2 * m + 1] * qb[2 * m + 1]);
        lq[i] = std::sqrt(s2 * 4.0 * kPi / double(2 * l + 1));
    }
}

// ---------------------------------------------------------------------------
// Exact gradient of a per-atom Steinhardt CV written as
//   V = sum_c sum_m mu_m Re(conj(K_cm) q_cm),  mu_0 = 1, mu_{m>0} = 2,
// with q_cm = (1/n_c) sum_{j in S(c)} Y_lm(x_j - x_c), by scatter over the
// selection: pair (c,j) adds  G = (1/n_c) sum_m mu_m Re(conj(K_cm) grad Y_lm)
// to d[owner(j)] and subtracts it from d[c].  K is given per center
// ([c*2(l+1) + 2m + {0,1}]).  Output dV/dx for local atoms (n_local*3, +=).
// Used with K = w * (4pi/(2l+1)) q_cm / (Ng q_l(c)) for Q_l.AVE and with the
// two-hop coefficient of orc_ld_coeff for the local form.
// ---------------------------------------------------------------------------
// centers[c] is the local atom of center c (NULL: center c is atom c).
void orc_grad_from_coeff(int l, int n_center, int N, int n_local, const int* ghost_owner,
                         const int* sel_idx, const double* sel_d, const int* n_sel,
                         const double* K, double* dVdx, const int* centers) {
    const int W = 2 * (l + 1);
    // The pair terms scatter to two atoms: each thread accumulates a static block of
    // centers into its own array, and the arrays are added in thread order.
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<double> acc((size_t)nt * n_local * 3, 0.0);
    #pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* out = acc.data() + (size_t)t * n_local * 3;
        std::vector<double> dY(6 * (l + 1));
        #pragma omp for schedule(static)
        for (int c = 0; c < n_center; ++c) {
            int n = n_sel[c];
            if (n == 0) continue;
            const double* k = K + (size_t)c * W;
            for (int s = 0; s < n; ++s) {
                const double* d = sel_d + ((size_t)c * N + s) * 4;
                ylm_grad(l, d[0], d[1], d[2], dY.data());
                double G[3] = {0, 0, 0};
                for (int m = 0; m <= l; ++m) {
                    double mu = m ? 2.0 : 1.0;
                    for (int a = 0; a < 3; ++a)
                        G[a] += mu * (k[2 * m] * dY[m * 6 + 2 * a] + k[2 * m + 1] * dY[m * 6 + 2 * a + 1]);
                }
                int o = owner_of(sel_idx[(size_t)c * N + s], n_local, ghost_owner);
                const size_t ci = centers ? (size_t)centers[c] : (size_t)c;
                for (int a = 0; a < 3; ++a) {
                    G[a] /= (double)n;
                    out[3 * (size_t)o + a] += G[a];
                    out[3 * ci + a] -= G[a];
                }
            }
        }
        #pragma omp barrier
        #pragma omp for schedule(static)
        for (int64_t q = 0; q < (int64_t)n_local * 3; ++q) {
            double v = 0.0;
            for (int u = 0; u < nt; ++u) v += acc[(size_t)u * n_local * 3 + q];
            dVdx[q] += v;
        }
    }
}

// Coefficients of Q_l.AVE = (1/Ng) sum_c q_l(c), times weight w:
// K_cm = w * (4pi/(2l+1)) q_cm / (Ng q_l(c)); zero when q_l(c) == 0.
void orc_q_coeff(int l, int n_center, const double* qlm, const double* ql, double Ng,
                 double w, double* K) {
    const int W = 2 * (l + 1);
    double P = 4.0 * kPi / double(2 * l + 1);
    for (int c = 0; c < n_center; ++c) {
        double a = ql[c] > 0.0 ? w * P / (Ng * ql[c]) : 0.0;
        for (int k = 0; k < W; ++k) K[(size_t)c * W + k] = a * qlm[(size_t)c * W + k];
    }
}

// Coefficients of LQ_l.AVE = (1/Ng) sum_i lq(i), times weight w (two-hop):
// d lq(i) = (4pi/(2l+1)) / (lq(i)(n_i+1)) sum_m mu_m Re(conj(qbar_im) d(q_im + sum_j q_jm)),
// so atom b in {i} U owner(S(i)) receives beta_i qbar_i, beta_i = w P/(Ng lq(i)(n_i+1)).
void orc_ld_coeff(int l, int n_center, int N, int n_local, const int* ghost_owner,
                  const int* sel_idx, const int* n_sel, const double* qbar, const double* lq,
                  double Ng, double w, double* K) {
    const int W = 2 * (l + 1);
    double P = 4.0 * kPi / double(2 * l + 1);
    for (size_t k = 0; k < (size_t)n_center * W; ++k) K[k] = 0.0;
    for (int i = 0; i < n_center; ++i) {
        int n = n_sel[i];
        if (!(lq[i] > 0.0)) continue;
        double beta = w * P / (Ng * lq[i] * (double)(n + 1));
        for (int k = 0; k < W; ++k) K[(size_t)i * W + k] += beta * qbar[(size_t)i * W + k];
        for (int s = 0; s < n; ++s) {
            int o = owner_of(sel_idx[(size_t)i * N + s], n_local, ghost_owner);
            for (int k = 0; k < W; ++k) K[(size_t)o * W + k] += beta * qbar[(size_t)i * W + k];
        }
    }
}

// ---------------------------------------------------------------------------
// Vector-k structure factor (no reference counterpart: the reference
// STRU_FACTOR is a per-atom Debye sum, CVs_Strufactor.cu:1082-1120).
// A(k) = sum_j exp(i k.r_j), S(k) = |A|^2 / N, CV = (1/K) sum_k S(k).
// dCV/dr_j = -(2/(N K)) sum_k k Im(conj(A(k)) exp(i k.r_j)).
// Pinned by |A| = N for a perfect lattice at a reciprocal vector, S(-k) = S(k)
// and finite differences (tests/test_oracle.py).
// ---------------------------------------------------------------------------
double orc_sk(int n, const double* x, int nk, const double* k, double* A /*2*nk*/, double* S) {
    double acc = 0.0;
    #pragma omp parallel for schedule(static)
    for (int q = 0; q < nk; ++q) {
        double re = 0, im = 0;
        for (int j = 0; j < n; ++j) {
            double ph = k[3 * q] * x[3 * j] + k[3 * q + 1] * x[3 * j + 1] + k[3 * q + 2] * x[3 * j + 2];
            re += std::cos(ph); im += std::sin(ph);
        }
        A[2 * q] = re; A[2 * q + 1] = im;
        S[q] = (re * re + im * im) / n;
    }
    for (int q = 0; q < nk; ++q) acc += S[q];
    return acc / nk;
}

void orc_sk_grad(int n, const double* x, int nk, const double* k, const double* A,
                 double w, double* dVdx /*n*3, +=*/) {
    double scale = -2.0 * w / ((double)n * nk);
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) {
        double g[3] = {0, 0, 0};
        for (int q = 0; q < nk; ++q) {
            double ph = k[3 * q] * x[3 * j] + k[3 * q + 1] * x[3 * j + 1] + k[3 * q + 2] * x[3 * j + 2];
            double c = std::cos(ph), s = std::sin(ph);
            double imv = A[2 * q] * s - A[2 * q + 1] * c;   // Im(conj(A) e^{i ph})
            for (int a = 0; a < 3; ++a) g[a] += k[3 * q + a] * imv;
        }
        for (int a = 0; a < 3; ++a) dVdx[3 * (size_t)j + a] += scale * g[a];
    }
}

// ---------------------------------------------------------------------------
// Uniform bias grid, restated from GH_t0_uniformGrid.cu (1-D and 2-D).
// Row-major, last dimension fastest (:242-258).
// ---------------------------------------------------------------------------
int orc_grid_radius(double sigma, double lo, double hi, int nb) {
    double dcv = (hi - lo) / nb;                                    // :86
    return static_cast<int>(std::ceil(4.0 * sigma / dcv)) + 1;       // :88
}

void orc_grid_deposit(int D, const double* lo, const double* hi, const int* nb, double* grid,
                      const double* s, double w, double sig) {
    int lower[3], upper[3];
    for (int k = 0; k < D; ++k) {
        double dcv = (hi[k] - lo[k]) / nb[k];
        int R = static_cast<int>(std::ceil(4.0 * sig / dcv)) + 1;
        int c = static_cast<int>((s[k] - lo[k]) / dcv);                        // :230
        lower[k] = std::max(0, c - R);
        upper[k] = std::min(nb[k] - 1, c + R);
    }
    if (D == 1) {
        for (long long g = lower[0]; g <= upper[0]; ++g) {
            double ctr = lo[0] + (g + 0.5) * (hi[0] - lo[0]) / nb[0];
            double dx = s[0] - ctr;
            grid[g] += w * std::exp(-0.5 * (dx * dx) / (sig * sig));
        }
    } else if (D == 2) {
        for (long long i = lower[0]; i <= upper[0]; ++i)
            for (long long j = lower[1]; j <= upper[1]; ++j) {
                double d0 = s[0] - (lo[0] + (i + 0.5) * (hi[0] - lo[0]) / nb[0]);
                double d1 = s[1] - (lo[1] + (j + 0.5) * (hi[1] - lo[1]) / nb[1]);
                grid[i * nb[1] + j] += w * std::exp(-0.5 * (d0 * d0 + d1 * d1) / (sig * sig));
            }
    }
}

static inline double cr_value(double p0, double p1, double p2, double p3, double x) {
    return p1 + 0.5 * x * ((p2 - p0) + x * (2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3) +
                           x * x * (-p0 + 3.0 * p1 - 3.0 * p2 + p3));
}
static inline double cr_deriv(double p0, double p1, double p2, double p3, double x) {
    return 0.5 * ((p2 - p0) + 2.0 * x * (2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3) +
                  3.0 * x * x * (-p0 + 3.0 * p1 - 3.0 * p2 + p3));
}
static inline void center_frac_clamp(double cv, double lo, double h, int nb, int& i, double& f) {
    double u = (cv - lo) / h - 0.5;
    i = static_cast<int>(std::floor(u));
    f = u - i;
    if (i < 1) { i = 1; f = 0.0; }
    if (i > nb - 3) { i = nb - 3; f = 1.0; }
}

// V and dV/ds at s (get_bias_energy :428-461, get_dVdcv :312-373); NaN -> 0.
double orc_grid_eval(int D, const double* lo, const double* hi, const int* nb, const double* grid,
                     const double* s, double* dVds) {
    if (D == 1) {
        double h = (hi[0] - lo[0]) / nb[0];
        int i; double x;
        center_frac_clamp(s[0], lo[0], h, nb[0], i, x);
        double p0 = grid[i - 1], p1 = grid[i], p2 = grid[i + 1], p3 = grid[i + 2];
        double d = cr_deriv(p0, p1, p2, p3, x) / h;
        if (std::isnan(d)) d = 0.0;
        if (dVds) dVds[0] = d;
        return cr_value(p0, p1, p2, p3, x);
    }
    double h0 = (hi[0] - lo[0]) / nb[0], h1 = (hi[1] - lo[1]) / nb[1];
    int i0, j0; double dx, dy;
    center_frac_clamp(s[0], lo[0], h0, nb[0], i0, dx);
    center_frac_clamp(s[1], lo[1], h1, nb[1], j0, dy);
    double col[4], dcol[4];
    for (int t = 0; t < 4; ++t) {
        int jj = j0 - 1 + t;
        double p0 = grid[(i0 - 1) * nb[1] + jj], p1 = grid[i0 * nb[1] + jj];
        double p2 = grid[(i0 + 1) * nb[1] + jj], p3 = grid[(i0 + 2) * nb[1] + jj];
        col[t] = cr_value(p0, p1, p2, p3, dx);
        dcol[t] = cr_deriv(p0, p1, p2, p3, dx);
    }
    if (dVds) {
        dVds[0] = cr_value(dcol[0], dcol[1], dcol[2], dcol[3], dy) / h0;
        dVds[1] = cr_deriv(col[0], col[1], col[2], col[3], dy) / h1;
        if (std::isnan(dVds[0]) || std::isnan(dVds[1])) { dVds[0] = 0.0; dVds[1] = 0.0; }
    }
    return cr_value(col[0], col[1], col[2], col[3], dy);
}

// ---------------------------------------------------------------------------
// Sparse-hash bias (Gaussian_Hill_type 1), restated from GH_t1_sparseHash.cu as a
// dense grid of D <= 3 dimensions (a bin the hash never touched reads 0, as a
// missing key does in get_total_bias :385-393).  Deposit: add_to_grid :228-240
// (window as the uniform grid), recursive_add2grid :242-262 (bin centre
// lo + (i + 1/2) dcv), gauss_calc :413-431.  Evaluation: get_bias_energy
// :395-407 and get_dVdcv :283-314 through mixed_recursive_logic :316-366, a
// recursion from dimension 0 to the innermost dimension: Catmull-Rom values
// along every dimension except the derivative one, whose Catmull-Rom
// derivative is divided by that dimension's bin width where it is taken.
// ---------------------------------------------------------------------------
static long long gt1_index(int D, const int* nb, const int* c) {
    long long k = 0;
    for (int d = 0; d < D; ++d) k = k * nb[d] + c[d];
    return k;
}

void orc_gridt1_deposit(int D, const double* lo, const double* hi, const int* nb, double* grid,
                        const double* s, double w, double sig) {
    int lower[3] = {0, 0, 0}, upper[3] = {0, 0, 0};
    double dcv[3] = {1, 1, 1};
    for (int k = 0; k < D; ++k) {
        dcv[k] = (hi[k] - lo[k]) / nb[k];
        int R = static_cast<int>(std::ceil(4.0 * sig / dcv[k])) + 1;
        int c = static_cast<int>((s[k] - lo[k]) / dcv[k]);
        lower[k] = std::max(0, c - R);
        upper[k] = std::min(nb[k] - 1, c + R);
    }
    int c[3] = {0, 0, 0};
    for (c[0] = lower[0]; c[0] <= upper[0]; ++c[0])
        for (c[1] = lower[1]; c[1] <= (D > 1 ? upper[1] : 0); ++c[1])
            for (c[2] = lower[2]; c[2] <= (D > 2 ? upper[2] : 0); ++c[2]) {
                double paw = 0.0;
                for (int d = 0; d < D; ++d) {
                    double dx = s[d] - (lo[d] + (c[d] + 0.5) * dcv[d]);
                    paw += dx * dx;
                }
                grid[gt1_index(D, nb, c)] += w * std::exp(-0.5 * (paw) / (sig * sig));
            }
}

static double gt1_rec(int D, const int* nb, const double* dcv, const double* grid, int cur, int target,
                      int* coord, const double* frac) {
    if (cur == D) return grid[gt1_index(D, nb, coord)];
    double p[4];
    const int origin = coord[cur];
    for (int i = -1; i <= 2; ++i) {
        int t = origin + i;
        if (t < 0) t = 0;
        if (t >= nb[cur]) t = nb[cur] - 1;
        coord[cur] = t;
        p[i + 1] = gt1_rec(D, nb, dcv, grid, cur + 1, target, coord, frac);
    }
    coord[cur] = origin;
    if (cur == target) return cr_deriv(p[0], p[1], p[2], p[3], frac[cur]) / dcv[cur];
    return cr_value(p[0], p[1], p[2], p[3], frac[cur]);
}

double orc_gridt1_eval(int D, const double* lo, const double* hi, const int* nb, const double* grid,
                       const double* s, double* dVds) {
    int base[3];
    double frac[3], dcv[3];
    for (int d = 0; d < D; ++d) {
        dcv[d] = (hi[d] - lo[d]) / nb[d];
        center_frac_clamp(s[d], lo[d], dcv[d], nb[d], base[d], frac[d]);
    }
    if (dVds) {
        for (int k = 0; k < D; ++k) {
            dVds[k] = gt1_rec(D, nb, dcv, grid, 0, k, base, frac);
            if (std::isnan(dVds[k])) dVds[k] = 0.0;
        }
    }
    return gt1_rec(D, nb, dcv, grid, 0, -1, base, frac);
}

// Well-tempered height (add_hill :194-208): w = h0 exp(-V(s)/(T kB (gamma-1))), V = 0 if !WT.
double orc_wt_height(int D, const double* lo, const double* hi, const int* nb, const double* grid,
                     const double* s, double h0, double biasf, double T, double kB, int wt) {
    double V = wt ? orc_grid_eval(D, lo, hi, nb, grid, s, nullptr) : 0.0;
    return h0 * std::exp(-(V) / (T * kB * (biasf - 1.0)));
}

// ---------------------------------------------------------------------------
// Exact FP64 sum of Gaussian hills (any D, per-dimension sigma); no reference
// counterpart (the reference 3-D grid is a no-op, GH_t0_uniformGrid.cu:375-382).
// V(s) = sum_h w_h exp(-sum_d (s_d - c_hd)^2 / (2 sigma_d^2)).  Pinned by FD and
// by agreement with the grid at bin centres (tests/test_oracle.py).
// ---------------------------------------------------------------------------
void orc_hills_sum(int D, int nh, const double* centers, const double* heights,
                   const double* sigma, int np, const double* s, double* V, double* dV) {
    #pragma omp parallel for schedule(static)
    for (int p = 0; p < np; ++p) {
        double v = 0, g[3] = {0, 0, 0};
        for (int h = 0; h < nh; ++h) {
            double e = 0, dd[3];
            for (int d = 0; d < D; ++d) {
                dd[d] = s[p * D + d] - centers[h * D + d];
                e += dd[d] * dd[d] / (2.0 * sigma[d] * sigma[d]);
            }
            double t = heights[h] * std::exp(-e);
            v += t;
            for (int d = 0; d < D; ++d) g[d] -= t * dd[d] / (sigma[d] * sigma[d]);
        }
        V[p] = v;
        if (dV) for (int d = 0; d < D; ++d) dV[p * D + d] = g[d];
    }
}

// ---------------------------------------------------------------------------
// MLP 3 -> H -> H -> 1 with tanh (no reference counterpart).  Row-major
// weights: W1[H][Din], b1[H], W2[H][H], b2[H], W3[H], b3.  Forward value and
// gradient with respect to the input.  Pinned against torch float64 autograd.
// ---------------------------------------------------------------------------
void orc_mlp_forward(int Din, int H, const double* W1, const double* b1, const double* W2,
                     const double* b2, const double* W3, double b3, int np, const double* s,
                     double* V, double* dVds) {
    std::vector<double> h1(H), h2(H), g2(H), g1(H);
    for (int p = 0; p < np; ++p) {
        const double* in = s + (size_t)p * Din;
        for (int a = 0; a < H; ++a) {
            double z = b1[a];
            for (int d = 0; d < Din; ++d) z += W1[a * Din + d] * in[d];
            h1[a] = std::tanh(z);
        }
        for (int a = 0; a < H; ++a) {
            double z = b2[a];
            for (int b = 0; b < H; ++b) z += W2[a * H + b] * h1[b];
            h2[a] = std::tanh(z);
        }
        double v = b3;
        for (int a = 0; a < H; ++a) v += W3[a] * h2[a];
        V[p] = v;
        if (dVds) {
            for (int a = 0; a < H; ++a) g2[a] = W3[a] * (1 - h2[a] * h2[a]);
            for (int b = 0; b < H; ++b) {
                double t = 0;
                for (int a = 0; a < H; ++a) t += g2[a] * W2[a * H + b];
                g1[b] = t * (1 - h1[b] * h1[b]);
            }
            for (int d = 0; d < Din; ++d) {
                double t = 0;
                for (int a = 0; a < H; ++a) t += g1[a] * W1[a * Din + d];
                dVds[(size_t)p * Din + d] = t;
            }
        }
    }
}

// Mean-squared-error loss L = (1/np) sum_p (V(s_p) - y_p)^2 and its gradient
// with respect to every parameter, in the same row-major layout.
double orc_mlp_loss_grad(int Din, int H, const double* W1, const double* b1, const double* W2,
                         const double* b2, const double* W3, double b3, int np, const double* s,
                         const double* y, double* gW1, double* gb1, double* gW2, double* gb2,
                         double* gW3, double* gb3) {
    std::vector<double> h1(H), h2(H), d2(H), d1(H);
    std::memset(gW1, 0, sizeof(double) * H * Din); std::memset(gb1, 0, sizeof(double) * H);
    std::memset(gW2, 0, sizeof(double) * H * H);   std::memset(gb2, 0, sizeof(double) * H);
    std::memset(gW3, 0, sizeof(double) * H);       *gb3 = 0;
    double loss = 0;
    for (int p = 0; p < np; ++p) {
        const double* in = s + (size_t)p * Din;
        for (int a = 0; a < H; ++a) {
            double z = b1[a];
            for (int d = 0; d < Din; ++d) z += W1[a * Din + d] * in[d];
            h1[a] = std::tanh(z);
        }
        for (int a = 0; a < H; ++a) {
            double z = b2[a];
            for (int b = 0; b < H; ++b) z += W2[a * H + b] * h1[b];
            h2[a] = std::tanh(z);
        }
        double v = b3;
        for (int a = 0; a < H; ++a) v += W3[a] * h2[a];
        double r = v - y[p];
        loss += r * r;
        double dv = 2.0 * r / np;
        *gb3 += dv;
        for (int a = 0; a < H; ++a) { gW3[a] += dv * h2[a]; d2[a] = dv * W3[a] * (1 - h2[a] * h2[a]); }
        for (int a = 0; a < H; ++a) {
            gb2[a] += d2[a];
            for (int b = 0; b < H; ++b) gW2[a * H + b] += d2[a] * h1[b];
        }
        for (int b = 0; b < H; ++b) {
            double t = 0;
            for (int a = 0; a < H; ++a) t += d2[a] * W2[a * H + b];
            d1[b] = t * (1 - h1[b] * h1[b]);
        }
        for (int a = 0; a < H; ++a) {
            gb1[a] += d1[a];
            for (int d = 0; d < Din; ++d) gW1[a * Din + d] += d1[a] * in[d];
        }
    }
    return loss / np;
}

}  // extern "C"

// ---------------------------------------------------------------- STEINH L (OpenMP)
// Switching function of a keyword: p = {type, r0, d0, alpha, n, m}, type 0 LINE,
// 1 STEP, 2 FERMI, 3 TANH, 4 RATIONAL (include/utils/zqc_switch_function.h:216-337;
// the same branches as oracle.py's sw_f / sw_df, which tests pin against each other).
static double lq_sw_f(const double* p, double v) {
    const int t = (int)p[0];
    if (t == 1) return v < p[1] ? 1.0 : 0.0;
    if (t == 2) return 1.0 / (1.0 + std::exp(-p[3] * (v - p[1])));
    if (t == 3) return 0.5 * (1.0 + std::tanh(p[3] * (v - p[1])));
    if (t == 4) {
        const int n = (int)p[4], m = (int)p[5];
        const double x = (v - p[2]) / p[1];
        if (2 * n == m) return 1.0 / (1.0 + std::pow(x, n));
        if (n == 2 * m) return 1.0 + std::pow(x, m);
        return (1.0 - std::pow(x, n)) / (1.0 - std::pow(x, m));
    }
    return 1.0;
}
static double lq_sw_df(const double* p, double v) {
    const int t = (int)p[0];
    if (t == 2) { const double e = std::exp(-p[3] * (v - p[1])); return p[3] * e / ((1.0 + e) * (1.0 + e)); }
    if (t == 3) { const double th = std::tanh(p[3] * (v - p[1])); return 0.5 * p[3] * (1.0 - th * th); }
    if (t == 4) {
        const int n = (int)p[4], m = (int)p[5];
        const double r0 = p[1], x = (v - p[2]) / r0;
        if (2 * n == m) { const double xn = std::pow(x, n); return -n * std::pow(x, n - 1) / ((1.0 + xn) * (1.0 + xn)) / r0; }
        if (n == 2 * m) return m * std::pow(x, m - 1) / r0;
        const double xm = std::pow(x, m);
        if (std::fabs(1.0 - xm) < 1e-12) return 0.5 * (n * (n - m)) / m / r0;
        return ((n - m) * std::pow(x, n - 1) * xm + m * std::pow(x, m - 1) - n * std::pow(x, n - 1)) /
               ((1.0 - xm) * (1.0 - xm)) / r0;
    }
    return 0.0;
}

extern "C" {
// The STEINH L CV and its exact gradient (oracle.py lq_cv / lq_grad restated in C with
// OpenMP over atoms; src/cvs/CVs_Steinhardt_Local.cu:1210-1285 environment, :1310-1452
// q_lm, :1490-1579 LQ_l, :694-745 the MEAN_SOLID / NSOLID / FRAC reductions).  mode 0
// MEAN_SOLID, 1 NSOLID, 2 FRAC; group NULL = every local atom; grad (n x 3) may be NULL.
// Sums over atoms are taken in atom order, so the result does not depend on the thread
// count; the gradient's scatter goes through per-thread arrays added in thread order.
double orc_lq(int l, int n, const double* x, const int* off, const int* idx, const int* own,
              const unsigned char* group, const double* swr, const double* swcv, double eps, int mode,
              double* grad) {
    const int L1 = l + 1;
    std::vector<double> S(n, 0.0), q(2 * (size_t)n * L1, 0.0), LQ(2 * (size_t)n * L1, 0.0), ql(n, 0.0);
    std::vector<unsigned char> live(n, 0);
    auto inG = [&](int i) { return group == nullptr || group[i] != 0; };
    auto entry = [&](int a, int e, double* d, double& r) {
        const int j = idx[e];
        d[0] = x[3 * j] - x[3 * a]; d[1] = x[3 * j + 1] - x[3 * a + 1]; d[2] = x[3 * j + 2] - x[3 * a + 2];
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (r2 < 1e-12) return -1.0;
        r = std::sqrt(r2);
        const double s = lq_sw_f(swr, r);
        return s < eps ? -1.0 : s;
    };
    // per-atom S and q_lm
#pragma omp parallel for schedule(static)
    for (int a = 0; a < n; ++a) {
        double Y[14], acc[14] = {0};
        int cnt = 0;
        for (int e = off[a]; e < off[a + 1]; ++e) {
            double d[3], r;
            const double s = entry(a, e, d, r);
            if (s < 0.0) continue;
            ++cnt;
            S[a] += s;
            ylm_reference(l, d[0], d[1], d[2], Y);
            for (int k = 0; k < 2 * L1; ++k) acc[k] += s * Y[k];
        }
        live[a] = cnt > 0 && S[a] >= 1e-12;
        for (int k = 0; k < 2 * L1; ++k) q[(size_t)a * 2 * L1 + k] = live[a] ? acc[k] * (1.0 / S[a]) : 0.0;
    }
    // LQ_lm and LQ_l of the group atoms
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        if (!(inG(i) && live[i])) continue;
        double acc[14];
        for (int k = 0; k < 2 * L1; ++k) acc[k] = q[(size_t)i * 2 * L1 + k];
        for (int e = off[i]; e < off[i + 1]; ++e) {
            double d[3], r;
            const double s = entry(i, e, d, r);
            if (s < 0.0) continue;
            const int o = owner_of(idx[e], n, own);
            for (int k = 0; k < 2 * L1; ++k) acc[k] += s * q[(size_t)o * 2 * L1 + k];
        }
        double s2 = 0.0;
        for (int m = 0; m <= l; ++m) {
            const double re = acc[2 * m] / (S[i] + 1.0), im = acc[2 * m + 1] / (S[i] + 1.0);
            LQ[(size_t)i * 2 * L1 + 2 * m] = re;
            LQ[(size_t)i * 2 * L1 + 2 * m + 1] = im;
            s2 += (m > 0 ? 2.0 : 1.0) * (re * re + im * im);
        }
        ql[i] = std::sqrt(s2 * 4.0 * M_PI / (2 * l + 1));
    }
    double Sq = 0.0, Sf = 0.0;
    int N = 0;
    for (int i = 0; i < n; ++i) {
        if (!inG(i)) continue;
        const double f = lq_sw_f(swcv, ql[i]);
        Sq += ql[i] * f; Sf += f; ++N;
    }
    const double cv = mode == 0 ? (Sf != 0.0 ? Sq / Sf : 0.0) : mode == 1 ? Sf : (N > 0 ? Sf / N : 0.0);
    if (grad == nullptr) return cv;
    // K_i = w_i 4 pi/(2l+1) c_m LQ_lm / LQ_l / (S_i + 1), H = K + scattered sigma K
    std::vector<double> K(2 * (size_t)n * L1, 0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        if (!(inG(i) && live[i] && ql[i] > 0.0)) continue;
        const double f = lq_sw_f(swcv, ql[i]), fp = lq_sw_df(swcv, ql[i]);
        const double w = mode == 0 ? (Sf != 0.0 ? (f + fp * (ql[i] - cv)) / Sf : 0.0) : mode == 1 ? fp : (N > 0 ? fp / N : 0.0);
        for (int m = 0; m <= l; ++m)
            for (int c = 0; c < 2; ++c)
                K[(size_t)i * 2 * L1 + 2 * m + c] = w * 4.0 * M_PI / (2 * l + 1) * (m > 0 ? 2.0 : 1.0) *
                                                     LQ[(size_t)i * 2 * L1 + 2 * m + c] / ql[i] / (S[i] + 1.0);
    }
    std::vector<double> H(K);
    for (int i = 0; i < n; ++i) {
        for (int e = off[i]; e < off[i + 1]; ++e) {
            double d[3], r;
            const double s = entry(i, e, d, r);
            if (s < 0.0) continue;
            const int o = owner_of(idx[e], n, own);
            for (int k = 0; k < 2 * L1; ++k) H[(size_t)o * 2 * L1 + k] += s * K[(size_t)i * 2 * L1 + k];
        }
    }
    const int nt = orc_num_threads();
    std::vector<double> acc((size_t)nt * n * 3, 0.0);
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        double* out = acc.data() + (size_t)tid * n * 3;
#pragma omp for schedule(static)
        for (int a = 0; a < n; ++a) {
            if (!live[a]) continue;
            const double* Ha = &H[(size_t)a * 2 * L1];
            const double* Ka = &K[(size_t)a * 2 * L1];
            const double* qa = &q[(size_t)a * 2 * L1];
            const double* LQa = &LQ[(size_t)a * 2 * L1];
            for (int e = off[a]; e < off[a + 1]; ++e) {
                double d[3], r;
                const double s = entry(a, e, d, r);
                if (s < 0.0) continue;
                const double ds = lq_sw_df(swr, r);
                const int o = owner_of(idx[e], n, own);
                const double* qo = &q[(size_t)o * 2 * L1];
                double Y[14], dY[42];
                ylm_reference(l, d[0], d[1], d[2], Y);
                ylm_grad(l, d[0], d[1], d[2], dY);
                double hy = 0.0, kq = 0.0, hd[3] = {0, 0, 0};
                for (int m = 0; m <= l; ++m) {
                    hy += Ha[2 * m] * (Y[2 * m] - qa[2 * m]) + Ha[2 * m + 1] * (Y[2 * m + 1] - qa[2 * m + 1]);
                    kq += Ka[2 * m] * (qo[2 * m] - LQa[2 * m]) + Ka[2 * m + 1] * (qo[2 * m + 1] - LQa[2 * m + 1]);
                    for (int c = 0; c < 3; ++c) hd[c] += Ha[2 * m] * dY[m * 6 + 2 * c] + Ha[2 * m + 1] * dY[m * 6 + 2 * c + 1];
                }
                for (int c = 0; c < 3; ++c) {
                    const double g = (ds * hy * d[c] / r + s * hd[c]) / S[a] + ds * kq * d[c] / r;
                    out[3 * a + c] -= g;
                    out[3 * o + c] += g;
                }
            }
        }
    }
    for (int k = 0; k < 3 * n; ++k) {
        double v = 0.0;
        for (int t = 0; t < nt; ++t) v += acc[(size_t)t * n * 3 + k];
        grad[k] = v;
    }
    return cv;
}

}  // extern "C"
