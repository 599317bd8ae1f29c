// shsrch.hpp - batched gradient shift search (pftcc_shsrch_grad%minimize,
// simple_pftcc_shsrch_grad.f90:254-338) on the device.
//
// Every (reference, particle) item runs its own L-BFGS-B over the two shift components (bounds
// lims, m = 15, factr 1e7, pgtol 1e-5, Moré-Thuente line search with ftol 0.1, gtol 1e-3,
// xtol 0.1, at most 20 trials; grad_shsrch_new :46-82, opt_spec :22-40,271), as a resumable state
// machine: a step advances an item until it needs the FP64 objective and gradient at a shift
// (gencorr_grad_for_rot_8, evaluated by k_rot8 for all items that asked) or, with the in-plane
// angle callback, the best rotation of the FP32 correlation at a shift (gencorrs + maxloc,
// evaluated by the item mode of k_corr_generic).  Requests are appended to compacted lists, so
// each round launches the evaluators on exactly the items that wait for them.
//
// The L-BFGS-B steps follow the bound-constrained method of Byrd, Lu, Nocedal & Zhu, version 3.0
// (simple_opt_lbfgsb.f90: mainlb :384-917, cauchy :1093-1491, subsm :2381-2530, lnsrlb :2125-2230,
// matupd/formt, dcsrch/dcstep :2532-2830) with the BFGS matrix formed explicitly: for two
// variables B = (the stored corrections applied to theta I) is a 2 x 2 matrix, and the
// generalized Cauchy point and the reduced Newton step are computed from it directly.  The
// positive-definiteness tests of the compact form that make the reference discard its memory
// (formt, formk) are carried out on the same matrices.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfk {

constexpr int LB_M = 15;          // opt_spec%lbfgsb_m
constexpr int LB_MAXLS = 20;      // trials per line search (mainlb :776)
constexpr double LB_EPSMCH = 2.220446049250313e-16;
constexpr double LB_TINY = 1e-10; // is_equal / is_gt_zero tolerance of the reference (DTINY)

enum : int { SH_REQ_NONE = 0, SH_REQ_FG = 1, SH_REQ_CORR = 2, SH_REQ_DONE = 3 };
enum : int {
  SH_RS_START = 0,      // nothing evaluated yet
  SH_RS_INIT_CORR,      // gencorrs at shift 0 (opt_angle)
  SH_RS_INIT_FG,        // objective at shift 0 and the given rotation (no opt_angle)
  SH_RS_FG0,            // first objective of a minimisation
  SH_RS_LS_FG,          // objective at a line-search trial
  SH_RS_NEWX_CORR,      // callback gencorrs at a new iterate
  SH_RS_END_CORR,       // gencorrs after a minimisation (opt_angle cycle test)
  SH_RS_COARSE_CORR,    // coarse_search_opt_angle: gencorrs at a grid point
  SH_RS_COARSE_FG       // coarse_search: objective at a grid point
};
constexpr int SH_COARSE_STEPS = 5;   // coarse_num_steps (simple_pftcc_shsrch_grad.f90:15)
enum : int { LB_FG = 1, LB_NEWX = 2, LB_END = 3 };

struct Dcs {              // Moré-Thuente search state (dcsrch isave/dsave)
  int brackt, stage;
  double finit, ginit, gtest, width, width1, stx, fx, gx, sty, fy, gy, stmin, stmax, stpmin, stpmax;
};

struct ShsItem {
  // the pair and the search settings
  int iref, ip, opt_angle, cur;        // 0-based reference, local particle, flag, rotation (1-based)
  int resume, maxits, cycle, coarse;  // coarse: the 5 x 5 initial grid (coarse_init)
  int ci, cbest_rot, pad0, pad1;       // grid point, best rotation of the grid (opt_angle)
  double cbest, cbx, cby;              // best grid cost and point
  double lo[2], hi[2];
  float x32[2];                        // spec%x (FP32 view of the iterate)
  double lowest_overall;
  // L-BFGS-B
  double x[2], g[2], f;
  double s[LB_M][2], y[LB_M][2];       // correction pairs, oldest first from head (ring)
  int col, head, iupdat, updatd;
  int itr, nfgv, ifun, iback;
  int iwhere[2], wasfree[2];
  double theta, sbgnrm;
  double t[2], r[2], d[2], z[2];
  double fold, gd, gdold, stp, stpmx, dnorm;
  Dcs ls;
  // result (cxy and irot of minimize)
  float out_cc, out_sx, out_sy;
  int out_irot;
};

// ------------------------------------------------------------------ Moré-Thuente (dcstep) ---
__device__ inline void lb_dcstep(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy, double& stp,
                                 double fp, double dp, int& brackt, double stpmin, double stpmax) {
  const double sgnd = dp * (dx / fabs(dx));
  double stpf;
  if (fp > fx) {
    const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    const double s = fmax(fabs(theta), fmax(fabs(dx), fabs(dp)));
    double gamma = s * sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
    if (stp < stx) gamma = -gamma;
    const double p = (gamma - dx) + theta;
    const double q = ((gamma - dx) + gamma) + dp;
    const double stpc = stx + (p / q) * (stp - stx);
    const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
    stpf = fabs(stpc - stx) < fabs(stpq - stx) ? stpc : stpc + (stpq - stpc) / 2.0;
    brackt = 1;
  } else if (sgnd < 0.0) {
    const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    const double s = fmax(fabs(theta), fmax(fabs(dx), fabs(dp)));
    double gamma = s * sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
    if (stp > stx) gamma = -gamma;
    const double p = (gamma - dp) + theta;
    const double q = ((gamma - dp) + gamma) + dx;
    const double stpc = stp + (p / q) * (stx - stp);
    const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
    stpf = fabs(stpc - stp) > fabs(stpq - stp) ? stpc : stpq;
    brackt = 1;
  } else if (fabs(dp) < fabs(dx)) {
    const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    const double s = fmax(fabs(theta), fmax(fabs(dx), fabs(dp)));
    double gamma = s * sqrt(fmax(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
    if (stp > stx) gamma = -gamma;
    const double p = (gamma - dp) + theta;
    const double q = (gamma + (dx - dp)) + gamma;
    const double rr = p / q;
    double stpc;
    if (rr < 0.0 && fabs(gamma) >= LB_TINY) stpc = stp + rr * (stx - stp);   // is_zero(gamma) false
    else if (stp > stx) stpc = stpmax;
    else stpc = stpmin;
    const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
    if (brackt) {
      stpf = fabs(stpc - stp) < fabs(stpq - stp) ? stpc : stpq;
      stpf = stp > stx ? fmin(stp + 0.66 * (sty - stp), stpf) : fmax(stp + 0.66 * (sty - stp), stpf);
    } else {
      stpf = fabs(stpc - stp) > fabs(stpq - stp) ? stpc : stpq;
      stpf = fmax(stpmin, fmin(stpmax, stpf));
    }
  } else {
    if (brackt) {
      const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
      const double s = fmax(fabs(theta), fmax(fabs(dy), fabs(dp)));
      double gamma = s * sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
      if (stp > sty) gamma = -gamma;
      const double p = (gamma - dp) + theta;
      const double q = ((gamma - dp) + gamma) + dy;
      stpf = stp + (p / q) * (sty - stp);
    } else {
      stpf = stp > stx ? stpmax : stpmin;
    }
  }
  if (fp > fx) {
    sty = stp;
    fy = fp;
    dy = dp;
  } else {
    if (sgnd < 0.0) {
      sty = stx;
      fy = fx;
      dy = dx;
    }
    stx = stp;
    fx = fp;
    dx = dp;
  }
  stp = stpf;
}

// dcsrch continuation: returns 0 = evaluate at the new stp, 1 = converged or stopped (warning)
__device__ inline int lb_dcsrch(Dcs& L, double f, double g, double& stp) {
  const double gtol = (double)1e-3f, xtol = 0.1;   // ftol enters through gtest
  const double ftest = L.finit + stp * L.gtest;
  if (L.stage == 1 && f <= ftest && g >= 0.0) L.stage = 2;
  if (L.brackt && (stp <= L.stmin || stp >= L.stmax)) return 1;
  if (L.brackt && L.stmax - L.stmin <= xtol * L.stmax) return 1;
  if (fabs(stp - L.stpmax) < LB_TINY && f <= ftest && g <= L.gtest) return 1;
  if (fabs(stp - L.stpmin) < LB_TINY && (f > ftest || g >= L.gtest)) return 1;
  if (f <= ftest && fabs(g) <= gtol * (-L.ginit)) return 1;
  if (L.stage == 1 && f <= L.fx && f > ftest) {
    const double gt = L.gtest;
    double fxm = L.fx - L.stx * gt, fym = L.fy - L.sty * gt;
    double gxm = L.gx - gt, gym = L.gy - gt;
    lb_dcstep(L.stx, fxm, gxm, L.sty, fym, gym, stp, f - stp * gt, g - gt, L.brackt, L.stmin, L.stmax);
    L.fx = fxm + L.stx * gt;
    L.fy = fym + L.sty * gt;
    L.gx = gxm + gt;
    L.gy = gym + gt;
  } else {
    lb_dcstep(L.stx, L.fx, L.gx, L.sty, L.fy, L.gy, stp, f, g, L.brackt, L.stmin, L.stmax);
  }
  if (L.brackt) {
    if (fabs(L.sty - L.stx) >= 0.66 * L.width1) stp = L.stx + 0.5 * (L.sty - L.stx);
    L.width1 = L.width;
    L.width = fabs(L.sty - L.stx);
    L.stmin = fmin(L.stx, L.sty);
    L.stmax = fmax(L.stx, L.sty);
  } else {
    L.stmin = stp + 1.1 * (stp - L.stx);
    L.stmax = stp + 4.0 * (stp - L.stx);
  }
  stp = fmin(fmax(stp, L.stpmin), L.stpmax);
  if ((L.brackt && (stp <= L.stmin || stp >= L.stmax)) || (L.brackt && L.stmax - L.stmin <= xtol * L.stmax))
    stp = L.stx;
  return 0;
}

// ------------------------------------------------------------------------- L-BFGS-B core ---
__device__ inline int lb_slot(const ShsItem& S, int i) { return (S.head + i) % LB_M; }   // i-th oldest

// B = the stored BFGS corrections applied to theta I, oldest first
__device__ inline void lb_bmat(const ShsItem& S, double B[2][2]) {
  B[0][0] = S.theta;
  B[0][1] = B[1][0] = 0.0;
  B[1][1] = S.theta;
  for (int i = 0; i < S.col; ++i) {
    const int k = lb_slot(S, i);
    const double s0 = S.s[k][0], s1 = S.s[k][1], y0 = S.y[k][0], y1 = S.y[k][1];
    const double b0 = B[0][0] * s0 + B[0][1] * s1, b1 = B[1][0] * s0 + B[1][1] * s1;
    const double sbs = s0 * b0 + s1 * b1, ys = y0 * s0 + y1 * s1;
    B[0][0] += -b0 * b0 / sbs + y0 * y0 / ys;
    B[0][1] += -b0 * b1 / sbs + y0 * y1 / ys;
    B[1][1] += -b1 * b1 / sbs + y1 * y1 / ys;
    B[1][0] = B[0][1];
  }
}

__device__ inline double lb_projgr(const ShsItem& S) {
  double m = 0.0;
  for (int i = 0; i < 2; ++i) {
    double gi = S.g[i];
    if (gi < 0.0) gi = fmax(S.x[i] - S.hi[i], gi);
    else gi = fmin(S.x[i] - S.lo[i], gi);
    m = fmax(m, fabs(gi));
  }
  return m;
}

// in-place Cholesky (dpofa) of the leading k x k block of A (row stride LB_M); false on a
// non-positive pivot
__device__ inline bool lb_chol(double (*A)[LB_M], int k) {
  for (int j = 0; j < k; ++j) {
    double s = A[j][j];
    for (int q = 0; q < j; ++q) s -= A[j][q] * A[j][q];
    if (s <= 0.0) return false;
    const double dj = sqrt(s);
    A[j][j] = dj;
    for (int i = j + 1; i < k; ++i) {
      double v = A[i][j];
      for (int q = 0; q < j; ++q) v -= A[i][q] * A[j][q];
      A[i][j] = v / dj;
    }
  }
  return true;
}

__device__ inline double lb_sy(const ShsItem& S, int i, int j) {   // s_i . y_j (oldest-first indices)
  const int a = lb_slot(S, i), b = lb_slot(S, j);
  return S.s[a][0] * S.y[b][0] + S.s[a][1] * S.y[b][1];
}

// formt (:2050-2085): T = theta S'S + L D^-1 L' must factor
__device__ __noinline__ bool lb_formt_ok(const ShsItem& S) {
  const int c = S.col;
  double T[LB_M][LB_M];
  for (int i = 0; i < c; ++i)
    for (int j = 0; j <= i; ++j) {
      const int a = lb_slot(S, i), b = lb_slot(S, j);
      double v = S.theta * (S.s[a][0] * S.s[b][0] + S.s[a][1] * S.s[b][1]);
      for (int k = 0; k < j; ++k) v += lb_sy(S, i, k) * lb_sy(S, j, k) / lb_sy(S, k, k);
      T[i][j] = v;
    }
  return lb_chol(T, c);
}

// formk (:1552-1750): K11 = Y_Z'Y_Z/theta + D; K12[j][i] = S_Z'Y_Z (i <= j) or -S_A'Y_A (i > j);
// chol(K11) = U'U, X = U'^-1 K12, then chol(theta S_A'S_A + X'X)
__device__ __noinline__ bool lb_formk_ok(const ShsItem& S, const int fr[2]) {
  const int c = S.col;
  double U[LB_M][LB_M], X[LB_M][LB_M];
  for (int i = 0; i < c; ++i)
    for (int j = 0; j <= i; ++j) {
      const int a = lb_slot(S, i), b = lb_slot(S, j);
      double v = 0.0;
      for (int k = 0; k < 2; ++k)
        if (fr[k]) v += S.y[a][k] * S.y[b][k];
      v /= S.theta;
      if (i == j) v += lb_sy(S, i, i);
      U[i][j] = v;
    }
  if (!lb_chol(U, c)) return false;          // lower factor L = U' in U[i][j], j <= i
  // K12 column by column, forward-solved with the lower factor
  for (int col = 0; col < c; ++col) {        // column = s index iy
    const int a = lb_slot(S, col);
    for (int row = 0; row < c; ++row) {      // row = y index jy
      const int b = lb_slot(S, row);
      double v = 0.0;
      if (row >= col) {
        for (int k = 0; k < 2; ++k)
          if (fr[k]) v += S.s[a][k] * S.y[b][k];
      } else {
        for (int k = 0; k < 2; ++k)
          if (!fr[k]) v -= S.s[a][k] * S.y[b][k];
      }
      for (int q = 0; q < row; ++q) v -= U[row][q] * X[q][col];
      X[row][col] = v / U[row][row];
    }
  }
  // theta S_A'S_A + X'X, into U's storage (the factor is no longer needed)
  for (int i = 0; i < c; ++i)
    for (int j = 0; j <= i; ++j) {
      const int a = lb_slot(S, i), b = lb_slot(S, j);
      double v = 0.0;
      for (int k = 0; k < 2; ++k)
        if (!fr[k]) v += S.s[a][k] * S.s[b][k];
      v *= S.theta;
      for (int q = 0; q < c; ++q) v += X[q][i] * X[q][j];
      U[i][j] = v;
    }
  return lb_chol(U, c);
}

// generalized Cauchy point (cauchy, :1093-1491) with explicit B; updates iwhere
__device__ inline void lb_cauchy(ShsItem& S, const double B[2][2], double xcp[2]) {
  xcp[0] = S.x[0];
  xcp[1] = S.x[1];
  if (S.sbgnrm <= 0.0) return;
  double d[2] = {0.0, 0.0};
  double tb[2];
  int ib[2], nbreak = 0, nfree_nobp = 0;
  for (int i = 0; i < 2; ++i) {
    const double neggi = -S.g[i];
    double tl = 0.0, tu = 0.0;
    if (S.iwhere[i] != 3) {
      tl = S.x[i] - S.lo[i];
      tu = S.hi[i] - S.x[i];
      const bool xlower = tl <= 0.0, xupper = tu <= 0.0;
      S.iwhere[i] = 0;
      if (xlower) {
        if (neggi <= 0.0) S.iwhere[i] = 1;
      } else if (xupper) {
        if (neggi >= 0.0) S.iwhere[i] = 2;
      } else if (fabs(neggi) <= 0.0) {
        S.iwhere[i] = -3;
      }
    }
    if (S.iwhere[i] != 0) continue;
    d[i] = neggi;
    if (neggi < 0.0) {
      tb[nbreak] = tl / (-neggi);
      ib[nbreak++] = i;
    } else if (neggi > 0.0) {
      tb[nbreak] = tu / neggi;
      ib[nbreak++] = i;
    } else {
      ++nfree_nobp;
    }
  }
  if (nbreak == 0 && nfree_nobp == 0) return;
  double z[2] = {0.0, 0.0};
  double f1 = S.g[0] * d[0] + S.g[1] * d[1];
  double f2 = d[0] * (B[0][0] * d[0] + B[0][1] * d[1]) + d[1] * (B[1][0] * d[0] + B[1][1] * d[1]);
  const double f2_org = f2;
  double dtm = -f1 / f2;
  double tsum = 0.0;
  if (nbreak > 0) {
    // ascending breakpoints; the first is the earliest minimum (ties keep input order)
    if (nbreak == 2 && tb[1] < tb[0]) {
      const double tt = tb[0];
      tb[0] = tb[1];
      tb[1] = tt;
      const int ii = ib[0];
      ib[0] = ib[1];
      ib[1] = ii;
    }
    double tj = 0.0;
    int nleft = nbreak;
    for (int q = 0; q < nbreak; ++q) {
      const double tj0 = tj;
      tj = tb[q];
      const double dt = tj - tj0;
      if (dtm < dt) break;
      tsum += dt;
      --nleft;
      const int ibp = ib[q];
      const double dibp = d[ibp];
      d[ibp] = 0.0;
      double zibp;
      if (dibp > 0.0) {
        zibp = S.hi[ibp] - S.x[ibp];
        xcp[ibp] = S.hi[ibp];
        S.iwhere[ibp] = 2;
      } else {
        zibp = S.lo[ibp] - S.x[ibp];
        xcp[ibp] = S.lo[ibp];
        S.iwhere[ibp] = 1;
      }
      if (nleft == 0 && nbreak == 2) return;
      for (int i = 0; i < 2; ++i)
        if (d[i] != 0.0) z[i] = tsum * d[i];
      z[ibp] = zibp;
      const double Bd0 = B[0][0] * d[0] + B[0][1] * d[1], Bd1 = B[1][0] * d[0] + B[1][1] * d[1];
      f1 = S.g[0] * d[0] + S.g[1] * d[1] + z[0] * Bd0 + z[1] * Bd1;
      f2 = d[0] * Bd0 + d[1] * Bd1;
      f2 = fmax(LB_EPSMCH * f2_org, f2);
      // with both variables bounded the path ends at the last breakpoint (bnded)
      dtm = nleft > 0 ? -f1 / f2 : 0.0;
    }
  }
  if (dtm <= 0.0) dtm = 0.0;
  tsum += dtm;
  xcp[0] += tsum * d[0];
  xcp[1] += tsum * d[1];
}

// reduced Newton step from xcp over the free variables + the version 3.0 projection rule
// (cmprlb :1493-1550, subsm :2381-2530)
__device__ inline void lb_subsm(const ShsItem& S, const double B[2][2], const double xcp[2], const int fr[2],
                                double out[2]) {
  const double dx0 = xcp[0] - S.x[0], dx1 = xcp[1] - S.x[1];
  const double rr0 = -(S.g[0] + B[0][0] * dx0 + B[0][1] * dx1);
  const double rr1 = -(S.g[1] + B[1][0] * dx0 + B[1][1] * dx1);
  double du[2] = {0.0, 0.0};
  if (fr[0] && fr[1]) {
    const double det = B[0][0] * B[1][1] - B[0][1] * B[1][0];
    du[0] = (B[1][1] * rr0 - B[0][1] * rr1) / det;
    du[1] = (B[0][0] * rr1 - B[1][0] * rr0) / det;
  } else if (fr[0]) {
    du[0] = rr0 / B[0][0];
  } else {
    du[1] = rr1 / B[1][1];
  }
  double xn[2] = {xcp[0], xcp[1]};
  int iword = 0;
  for (int k = 0; k < 2; ++k) {
    if (!fr[k]) continue;
    xn[k] = fmin(S.hi[k], fmax(S.lo[k], xcp[k] + du[k]));
    if (fabs(xn[k] - S.lo[k]) < LB_TINY || fabs(xn[k] - S.hi[k]) < LB_TINY) iword = 1;
  }
  if (iword) {
    const double dd_p = (xn[0] - S.x[0]) * S.g[0] + (xn[1] - S.x[1]) * S.g[1];
    if (dd_p > LB_TINY) {
      // ascent: back to xcp and truncate the Newton step to the box
      xn[0] = xcp[0];
      xn[1] = xcp[1];
      double alpha = 1.0, temp1 = 1.0;
      int ibd = -1;
      for (int k = 0; k < 2; ++k) {
        if (!fr[k]) continue;
        const double dk = du[k];
        if (dk < 0.0) {
          const double temp2 = S.lo[k] - xn[k];
          if (temp2 >= 0.0) temp1 = 0.0;
          else if (dk * alpha < temp2) temp1 = temp2 / dk;
        } else if (dk > 0.0) {
          const double temp2 = S.hi[k] - xn[k];
          if (temp2 <= 0.0) temp1 = 0.0;
          else if (dk * alpha > temp2) temp1 = temp2 / dk;
        }
        if (temp1 < alpha) {
          alpha = temp1;
          ibd = k;
        }
      }
      if (alpha < 1.0) {
        if (du[ibd] > 0.0) {
          xn[ibd] = S.hi[ibd];
          du[ibd] = 0.0;
        } else if (du[ibd] < 0.0) {
          xn[ibd] = S.lo[ibd];
          du[ibd] = 0.0;
        }
      }
      for (int k = 0; k < 2; ++k)
        if (fr[k]) xn[k] += alpha * du[k];
    }
  }
  out[0] = xn[0];
  out[1] = xn[1];
}

__device__ inline void lb_reset(ShsItem& S) {
  S.col = 0;
  S.head = 0;
  S.theta = 1.0;
  S.iupdat = 0;
  S.updatd = 0;
}

// Starts a minimisation from x0 (lbfgsb_minimize + setulb START, active): asks for f, g at x0.
__device__ inline int lb_start(ShsItem& S, double x0, double x1) {
  S.x[0] = fmin(fmax(x0, S.lo[0]), S.hi[0]);
  S.x[1] = fmin(fmax(x1, S.lo[1]), S.hi[1]);
  for (int i = 0; i < 2; ++i) {
    S.iwhere[i] = (S.hi[i] - S.lo[i] <= 0.0) ? 3 : 0;
    S.wasfree[i] = 1;
  }
  lb_reset(S);
  S.itr = 0;
  S.nfgv = 0;
  return LB_FG;
}

// Search direction (label 222) and the first trial of the line search; LB_FG or LB_END.
__device__ __noinline__ int lb_direction(ShsItem& S) {
  for (int guard = 0; guard < 4; ++guard) {
    double B[2][2], xcp[2], zz[2];
    lb_bmat(S, B);
    lb_cauchy(S, B, xcp);
    // freev: free variables at the Cauchy point; wrk when the set changed or memory updated
    int fr[2];
    bool changed = false;
    for (int i = 0; i < 2; ++i) {
      fr[i] = S.iwhere[i] <= 0;
      if (S.itr > 0 && fr[i] != S.wasfree[i]) changed = true;
    }
    const bool wrk = changed || S.updatd;
    S.wasfree[0] = fr[0];
    S.wasfree[1] = fr[1];
    zz[0] = xcp[0];
    zz[1] = xcp[1];
    if ((fr[0] || fr[1]) && S.col > 0) {
      if (wrk && !lb_formk_ok(S, fr)) {
        lb_reset(S);
        continue;
      }
      lb_subsm(S, B, xcp, fr, zz);
    }
    S.z[0] = zz[0];
    S.z[1] = zz[1];
    S.d[0] = zz[0] - S.x[0];
    S.d[1] = zz[1] - S.x[1];
    // lnsrlb start
    const double dtd = S.d[0] * S.d[0] + S.d[1] * S.d[1];
    S.dnorm = sqrt(dtd);
    double stpmx = 1.0e10;
    if (S.itr == 0) {
      stpmx = 1.0;
    } else {
      for (int i = 0; i < 2; ++i) {
        const double a1 = S.d[i];
        if (a1 < 0.0) {
          const double a2 = S.lo[i] - S.x[i];
          if (a2 >= 0.0) stpmx = 0.0;
          else if (a1 * stpmx < a2) stpmx = a2 / a1;
        } else if (a1 > 0.0) {
          const double a2 = S.hi[i] - S.x[i];
          if (a2 <= 0.0) stpmx = 0.0;
          else if (a1 * stpmx > a2) stpmx = a2 / a1;
        }
      }
    }
    S.stpmx = stpmx;
    S.stp = 1.0;                    // boxed
    S.t[0] = S.x[0];
    S.t[1] = S.x[1];
    S.r[0] = S.g[0];
    S.r[1] = S.g[1];
    S.fold = S.f;
    S.ifun = 0;
    S.iback = 0;
    S.gd = S.g[0] * S.d[0] + S.g[1] * S.d[1];
    S.gdold = S.gd;
    bool bad = S.gd >= 0.0 || S.stp > stpmx;
    if (!bad) {
      Dcs& L = S.ls;
      L.stpmin = 0.0;
      L.stpmax = stpmx;
      L.brackt = 0;
      L.stage = 1;
      L.finit = S.f;
      L.ginit = S.gd;
      L.gtest = (double)0.1f * S.gd;
      L.width = stpmx;
      L.width1 = stpmx / 0.5;
      L.stx = 0.0;
      L.fx = S.f;
      L.gx = S.gd;
      L.sty = 0.0;
      L.fy = S.f;
      L.gy = S.gd;
      L.stmin = 0.0;
      L.stmax = S.stp + 4.0 * S.stp;
      // first trial
      S.ifun = 1;
      S.nfgv += 1;
      S.iback = 0;
      S.x[0] = S.z[0];
      S.x[1] = S.z[1];
      return LB_FG;
    }
    // line search cannot start (info = -4): back to the iterate
    S.x[0] = S.t[0];
    S.x[1] = S.t[1];
    S.g[0] = S.r[0];
    S.g[1] = S.r[1];
    S.f = S.fold;
    if (S.col == 0) return LB_END;      // ABNORMAL_TERMINATION_IN_LNSRCH
    lb_reset(S);
  }
  return LB_END;
}

// f, g at a line-search trial: next trial (LB_FG), new iterate (LB_NEWX) or end (failure)
__device__ inline int lb_after_trial(ShsItem& S) {
  S.gd = S.g[0] * S.d[0] + S.g[1] * S.d[1];
  double stp = S.stp;
  const int done = lb_dcsrch(S.ls, S.f, S.gd, stp);
  S.stp = stp;
  if (done) {
    S.itr += 1;
    S.sbgnrm = lb_projgr(S);
    return LB_NEWX;
  }
  S.ifun += 1;
  S.nfgv += 1;
  S.iback = S.ifun - 1;
  if (S.iback >= LB_MAXLS) {
    // too many trials: back to the iterate (the trial was not evaluated)
    S.x[0] = S.t[0];
    S.x[1] = S.t[1];
    S.g[0] = S.r[0];
    S.g[1] = S.r[1];
    S.f = S.fold;
    S.nfgv -= 1;
    if (S.col == 0) return LB_END;
    lb_reset(S);
    return lb_direction(S);
  }
  if (fabs(stp - 1.0) < LB_TINY) {
    S.x[0] = S.z[0];
    S.x[1] = S.z[1];
  } else {
    S.x[0] = stp * S.d[0] + S.t[0];
    S.x[1] = stp * S.d[1] + S.t[1];
  }
  return LB_FG;
}

// after NEW_X was accepted by the caller (label 777): convergence tests, update, next direction
__device__ inline int lb_after_newx(ShsItem& S) {
  if (S.sbgnrm <= 1.0e-5) return LB_END;
  const double ddum0 = fmax(fabs(S.fold), fmax(fabs(S.f), 1.0));
  if (S.fold - S.f <= 1.0e7 * LB_EPSMCH * ddum0) return LB_END;
  const double yv0 = S.g[0] - S.r[0], yv1 = S.g[1] - S.r[1];
  double dr, ddum, sv0, sv1;
  if (fabs(S.stp - 1.0) < LB_TINY) {
    dr = S.gd - S.gdold;
    ddum = -S.gdold;
    sv0 = S.d[0];
    sv1 = S.d[1];
  } else {
    dr = (S.gd - S.gdold) * S.stp;
    sv0 = S.stp * S.d[0];
    sv1 = S.stp * S.d[1];
    ddum = -S.gdold * S.stp;
  }
  if (dr <= LB_EPSMCH * ddum) {
    S.updatd = 0;
    return lb_direction(S);
  }
  S.updatd = 1;
  S.iupdat += 1;
  int slot;
  if (S.col < LB_M) {
    slot = (S.head + S.col) % LB_M;
    S.col += 1;
  } else {
    slot = S.head;
    S.head = (S.head + 1) % LB_M;
  }
  S.s[slot][0] = sv0;
  S.s[slot][1] = sv1;
  S.y[slot][0] = yv0;
  S.y[slot][1] = yv1;
  S.theta = (yv0 * yv0 + yv1 * yv1) / dr;
  if (!lb_formt_ok(S)) lb_reset(S);
  return lb_direction(S);
}

// -------------------------------------------------------------------- the shift search ---
__device__ inline void shs_rotate(const ShsItem& S, float ang_deg, float v0, float v1, float& o0, float& o1) {
  // matmul(v, rotmat2d(ang)) in FP32 (simple_math.f90:1550-1559)
  const float a = ang_deg * 3.14159274f / 180.0f;
  const float c = cosf(a), s = sinf(a);
  o0 = v0 * c + v1 * (-s);
  o1 = v0 * s + v1 * c;
}

// Point c of the coarse grid (coarse_search :340-366): step (hi - lo)/5 in FP64 from the FP32
// limits, lo + step/2 + i step, x outer.
__device__ inline void shs_coarse_point(const ShsItem& S, int c, double& x, double& y) {
  const double stx = (double)((float)S.hi[0] - (float)S.lo[0]) / (double)SH_COARSE_STEPS;
  const double sty = (double)((float)S.hi[1] - (float)S.lo[1]) / (double)SH_COARSE_STEPS;
  const int ix = c / SH_COARSE_STEPS, iy = c % SH_COARSE_STEPS;
  x = (double)(float)S.lo[0] + stx / 2.0 + (double)ix * stx;
  y = (double)(float)S.lo[1] + sty / 2.0 + (double)iy * sty;
}

// Result of the last request: f, g (SH_REQ_FG) or the best rotation and its cc (SH_REQ_CORR).
// Returns the next request.
__device__ inline int shs_advance(ShsItem& S, double f, double g0, double g1, int corr_rot, float corr_cc,
                                  const float* angtab_deg, int max_evals) {
  int lb;
  switch (S.resume) {
    case SH_RS_START:
      S.cycle = 0;
      S.x32[0] = 0.f;
      S.x32[1] = 0.f;
      if (S.opt_angle) {
        S.resume = SH_RS_INIT_CORR;
        return SH_REQ_CORR;
      }
      S.x[0] = 0.0;
      S.x[1] = 0.0;
      S.resume = SH_RS_INIT_FG;
      return SH_REQ_FG;
    case SH_RS_INIT_CORR:
      S.cur = corr_rot;
      S.lowest_overall = -(double)corr_cc;
      if (S.coarse) {
        S.ci = 0;
        S.cbest = 1.0e300;
        S.cbest_rot = 0;
        S.cbx = S.cby = 0.0;
        double gx, gy;
        shs_coarse_point(S, 0, gx, gy);
        S.x32[0] = (float)gx;
        S.x32[1] = (float)gy;
        S.resume = SH_RS_COARSE_CORR;
        return SH_REQ_CORR;
      }
      lb_start(S, 0.0, 0.0);
      S.resume = SH_RS_FG0;
      return SH_REQ_FG;
    case SH_RS_INIT_FG:
      S.lowest_overall = -f;
      if (S.coarse) {
        S.ci = 0;
        S.cbest = 1.0e300;
        S.cbx = S.cby = 0.0;
        shs_coarse_point(S, 0, S.x[0], S.x[1]);
        S.resume = SH_RS_COARSE_FG;
        return SH_REQ_FG;
      }
      lb_start(S, 0.0, 0.0);
      S.resume = SH_RS_FG0;
      return SH_REQ_FG;
    case SH_RS_COARSE_CORR: {
      // coarse_search_opt_angle (:368-395): strict minimum of -max(gencorrs) over the grid
      const double cost = -(double)corr_cc;
      if (cost < S.cbest) {
        double gx, gy;
        shs_coarse_point(S, S.ci, gx, gy);
        S.cbest = cost;
        S.cbest_rot = corr_rot;
        S.cbx = gx;
        S.cby = gy;
      }
      if (++S.ci < SH_COARSE_STEPS * SH_COARSE_STEPS) {
        double gx, gy;
        shs_coarse_point(S, S.ci, gx, gy);
        S.x32[0] = (float)gx;
        S.x32[1] = (float)gy;
        return SH_REQ_CORR;
      }
      S.x32[0] = 0.f;
      S.x32[1] = 0.f;
      if (S.cbest_rot != 0) {
        S.x32[0] = (float)S.cbx;
        S.x32[1] = (float)S.cby;
        S.cur = S.cbest_rot;
      }
      lb_start(S, (double)S.x32[0], (double)S.x32[1]);
      S.resume = SH_RS_FG0;
      return SH_REQ_FG;
    }
    case SH_RS_COARSE_FG: {
      // coarse_search (:340-366): strict minimum of -gencorr_for_rot_8 over the grid
      const double cost = -f;
      if (cost < S.cbest) {
        S.cbest = cost;
        S.cbx = S.x[0];
        S.cby = S.x[1];
      }
      if (++S.ci < SH_COARSE_STEPS * SH_COARSE_STEPS) {
        shs_coarse_point(S, S.ci, S.x[0], S.x[1]);
        return SH_REQ_FG;
      }
      S.x32[0] = 0.f;
      S.x32[1] = 0.f;
      if (S.cbest < S.lowest_overall) {
        S.lowest_overall = S.cbest;
        S.x32[0] = (float)S.cbx;
        S.x32[1] = (float)S.cby;
      }
      lb_start(S, (double)S.x32[0], (double)S.x32[1]);
      S.resume = SH_RS_FG0;
      return SH_REQ_FG;
    }
    case SH_RS_FG0:
      S.f = -f;
      S.g[0] = -g0;
      S.g[1] = -g1;
      S.nfgv = 1;
      S.sbgnrm = lb_projgr(S);
      lb = S.sbgnrm <= 1.0e-5 ? LB_END : lb_direction(S);
      break;
    case SH_RS_LS_FG:
      S.f = -f;
      S.g[0] = -g0;
      S.g[1] = -g1;
      lb = lb_after_trial(S);
      break;
    case SH_RS_NEWX_CORR:
      S.cur = corr_rot;
      lb = S.nfgv > S.maxits ? LB_END : lb_after_newx(S);
      break;
    case SH_RS_END_CORR: {
      const int loc = corr_rot;
      if (loc != S.cur && S.cycle + 1 < max_evals) {
        S.cur = loc;
        S.cycle += 1;
        lb_start(S, (double)S.x32[0], (double)S.x32[1]);
        S.resume = SH_RS_FG0;
        return SH_REQ_FG;
      }
      S.cur = loc;
      const float lowest = -corr_cc;
      if ((double)lowest < S.lowest_overall) {
        S.out_irot = S.cur;
        S.out_cc = -lowest;
        shs_rotate(S, angtab_deg[S.cur - 1], S.x32[0], S.x32[1], S.out_sx, S.out_sy);
      } else {
        S.out_irot = 0;
      }
      return SH_REQ_DONE;
    }
    default:
      return SH_REQ_DONE;
  }
  // continue the minimisation until it needs something from outside
  for (;;) {
    if (lb == LB_FG) {
      S.resume = SH_RS_LS_FG;
      return SH_REQ_FG;
    }
    if (lb == LB_NEWX) {
      if (S.opt_angle) {
        S.x32[0] = (float)S.x[0];
        S.x32[1] = (float)S.x[1];
        S.resume = SH_RS_NEWX_CORR;
        return SH_REQ_CORR;
      }
      lb = S.nfgv > S.maxits ? LB_END : lb_after_newx(S);
      continue;
    }
    // LB_END: the minimisation is over (lowest_cost = real(f), spec%x = real(x_8))
    S.x32[0] = (float)S.x[0];
    S.x32[1] = (float)S.x[1];
    if (S.opt_angle) {
      S.resume = SH_RS_END_CORR;
      return SH_REQ_CORR;
    }
    const float lowest = (float)S.f;
    if ((double)lowest < S.lowest_overall) {
      S.out_irot = S.cur;
      S.out_cc = -lowest;
      shs_rotate(S, angtab_deg[S.cur - 1], S.x32[0], S.x32[1], S.out_sx, S.out_sy);
    } else {
      S.out_irot = 0;
    }
    return SH_REQ_DONE;
  }
}

struct ShsArgs {
  ShsItem* items;
  int nitems;
  const float* angtab_deg;
  int max_evals;
  int nrots;
  // initialisation (init != 0): per item 0-based reference, local particle, 1-based rotation
  int init;
  const int* irefs;
  const int* ips;
  const int* irots;
  float lo0, lo1, hi0, hi1;
  int maxits, opt_angle, coarse_init;
  // consumed requests of the previous round
  int nfg_in, ncorr_in;
  const int* fg_item_in;
  const double* fg_f_in;
  const double* fg_grad_in;        // [n][3]
  const int* corr_item_in;
  const unsigned long long* corr_key_in;
  // produced requests
  int* counters_out;               // [0] FG, [1] CORR
  int* fg_item_out;
  int* fg_ref_out;
  int* fg_ip_out;
  double* fg_sh_out;               // [n][2]
  int* fg_rho_out;
  int* corr_item_out;
  int* corr_ref_out;
  int* corr_ip_out;
  float* corr_sh_out;              // [n][2]
};

__device__ inline void shs_key_decode(unsigned long long k, int nrots, int& rot, float& cc) {
  const unsigned int hi = (unsigned int)(k >> 32);
  const unsigned int bits = (hi & 0x80000000u) ? (hi & 0x7FFFFFFFu) : ~hi;
  cc = __uint_as_float(bits);
  const unsigned int lin = 0xFFFFFFFFu - (unsigned int)(k & 0xFFFFFFFFull);
  rot = (int)(lin % (unsigned)nrots) + 1;
}

// One thread per consumed request (or per item at initialisation).
__global__ void __launch_bounds__(64) k_shs_step(ShsArgs a) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  int item;
  double f = 0.0, g0 = 0.0, g1 = 0.0;
  int rot = 0;
  float cc = 0.f;
  if (a.init) {
    if (pos >= a.nitems) return;
    item = pos;
    ShsItem& S = a.items[item];
    S.iref = a.irefs[pos];
    S.ip = a.ips[pos];
    S.cur = a.irots ? a.irots[pos] : 0;
    S.opt_angle = a.opt_angle;
    S.coarse = a.coarse_init;
    S.maxits = a.maxits;
    S.lo[0] = a.lo0;
    S.lo[1] = a.lo1;
    S.hi[0] = a.hi0;
    S.hi[1] = a.hi1;
    S.resume = SH_RS_START;
    S.out_irot = 0;
    S.out_cc = 0.f;
    S.out_sx = S.out_sy = 0.f;
  } else {
    if (pos >= a.nfg_in + a.ncorr_in) return;
    if (pos < a.nfg_in) {
      item = a.fg_item_in[pos];
      f = a.fg_f_in[pos];
      g0 = a.fg_grad_in[3 * (size_t)pos];
      g1 = a.fg_grad_in[3 * (size_t)pos + 1];
    } else {
      const int q = pos - a.nfg_in;
      item = a.corr_item_in[q];
      shs_key_decode(a.corr_key_in[q], a.nrots, rot, cc);
    }
  }
  ShsItem& S = a.items[item];
  const int rq = shs_advance(S, f, g0, g1, rot, cc, a.angtab_deg, a.max_evals);
  if (rq == SH_REQ_FG) {
    const int p = atomicAdd(&a.counters_out[0], 1);
    a.fg_item_out[p] = item;
    a.fg_ref_out[p] = S.iref;
    a.fg_ip_out[p] = S.ip;
    a.fg_sh_out[2 * (size_t)p] = S.x[0];
    a.fg_sh_out[2 * (size_t)p + 1] = S.x[1];
    a.fg_rho_out[p] = S.cur - 1;
  } else if (rq == SH_REQ_CORR) {
    const int p = atomicAdd(&a.counters_out[1], 1);
    a.corr_item_out[p] = item;
    a.corr_ref_out[p] = S.iref;
    a.corr_ip_out[p] = S.ip;
    a.corr_sh_out[2 * (size_t)p] = S.x32[0];
    a.corr_sh_out[2 * (size_t)p + 1] = S.x32[1];
  }
}

// Gathers the results: cxy[n][3] (cc, shift rotated to the reference frame) and irot (0: no
// better solution than the start).
__global__ void k_shs_out(const ShsItem* __restrict__ items, int n, float* __restrict__ cxy, int* __restrict__ irot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ShsItem& S = items[i];
  cxy[3 * (size_t)i] = S.out_cc;
  cxy[3 * (size_t)i + 1] = S.out_sx;
  cxy[3 * (size_t)i + 2] = S.out_sy;
  irot[i] = S.out_irot;
}

}  // namespace pfk
