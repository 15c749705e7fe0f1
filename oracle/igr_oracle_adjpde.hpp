// CPU ORACLE — TEST INFRASTRUCTURE ONLY (see igr_oracle.hpp).
//
// Plain-loop restatement of the continuous ("optimise-then-discretise")
// adjoint of the IGR Euler system and of the snapshot-storing forward run it
// consumes:
//   IGR_adjoints_2d/adjoint_pde.jl   (2D, conserved order mx, my, rho, E)
//   IGR_adjoints_1d/adjoint_pde.jl   (1D, conserved order mu, rho, E)
//   run_forward_store, IGR_adjoints_2d/forward.jl:111-158 and
//                      IGR_adjoints_1d/forward.jl:137-187
// Each routine cites the lines it follows and keeps the reference's loop and
// operation order.  The adjoint equation in reversed time tau = T - t is
//   d(a)/d(tau) = div F_loc(a) - S_loc(a) + S_Sigma(Pi(a)),
// with Pi the solution of the same SIP system as Sigma, right-hand side
// alpha * mass * (div a_mu + grad a_E . u).
#pragma once
#include "igr_oracle.hpp"

namespace orc {

// Test-only switch that plants one deliberate mistake in this restatement
// (0 = none).  The tests set it to show that each of their checks rejects a
// wrong sign or a dropped term:
//   1  drop the "+ uyy" term of d(divF_mx)/d(mx)  (1D: the "u_x" term of d(divF_mu)/d(mu))
//   2  flip the sign of the adjoint LLF dissipation
//   3  drop div(Pi grad(Sigma) / rho) from the nonlocal rho source
//   4  flip the sign of the nonlocal momentum source
extern thread_local int g_adj_mutation;

// =====================================================================
//                              2D
// =====================================================================

// adjoint_flux_local_x / _y, IGR_adjoints_2d/adjoint_pde.jl:41-78.
// q = (mx, my, rho, E, Sigma), a = (a_mx, a_my, a_rho, a_E).
inline void adj_flux2(int dir, double g, const double q[5], const double a[4], double f[4]) {
  const double mx = q[0], my = q[1], r = q[2], E = q[3], S = q[4];
  const double ux = mx / r, uy = my / r;
  const double P = pressure2(g, mx, my, r, E);
  const double H = E + P + S;
  const double Pmx = -(g - 1) * ux;
  const double Pmy = -(g - 1) * uy;
  const double Pr = (g - 1) * (ux * ux + uy * uy) / 2;
  const double PE = g - 1;
  const double amu = a[0] * mx + a[1] * my;
  if (dir == 0) {
    f[0] = amu / r + a[0] * ux + Pmx * a[0] + a[2] + a[3] * (H / r) + a[3] * Pmx * mx / r;
    f[1] = a[1] * ux + Pmy * a[0] + a[3] * Pmy * mx / r;
    f[2] = -ux * amu / r + Pr * a[0] + a[3] * mx * (Pr / r - H / (r * r));
    f[3] = PE * a[0] + a[3] * (1 + PE) * ux;
  } else {
    f[0] = a[0] * uy + Pmx * a[1] + a[3] * Pmx * my / r;
    f[1] = amu / r + a[1] * uy + Pmy * a[1] + a[2] + a[3] * (H / r) + a[3] * Pmy * my / r;
    f[2] = -uy * amu / r + Pr * a[1] + a[3] * my * (Pr / r - H / (r * r));
    f[3] = PE * a[1] + a[3] * (1 + PE) * uy;
  }
}

// adjoint_llf_flux_x / _y, IGR_adjoints_2d/adjoint_pde.jl:87-120: the
// dissipation has the opposite sign of the forward LLF flux.
inline void adj_llf2(int dir, double g, const double qL[5], const double qR[5], const double aL[4],
                     const double aR[4], double f[4]) {
  const double lam = dir == 0 ? max_(wave_x2(g, qL[0], qL[1], qL[2], qL[3]), wave_x2(g, qR[0], qR[1], qR[2], qR[3]))
                              : max_(wave_y2(g, qL[0], qL[1], qL[2], qL[3]), wave_y2(g, qR[0], qR[1], qR[2], qR[3]));
  double fL[4], fR[4];
  adj_flux2(dir, g, qL, aL, fL);
  adj_flux2(dir, g, qR, aR, fR);
  const double sgn = g_adj_mutation == 2 ? -1.0 : 1.0;
  for (int v = 0; v < 4; ++v) f[v] = (fL[v] + fR[v]) / 2 - sgn * lam * (aL[v] - aR[v]) / 2;
}

// compute_adjoint_source_local!, IGR_adjoints_2d/adjoint_pde.jl:136-212.
// q: 4 fields + S (N each); a: 4 fields; Sl: 4 fields (out).
inline void adj_source_local2(const Grid2& G, double g, const double* q, const double* S, const double* a,
                              double* Sl) {
  const int np = G.np, N = G.N; const Basis& B = G.B; const Mesh2D& M = G.M;
  const double invJx = 1 / M.Jx, invJy = 1 / M.Jy, gm = g - 1;
  const double* Mx = q; const double* My = q + N; const double* R = q + 2 * (size_t)N; const double* En = q + 3 * (size_t)N;
  for (int ey = 0; ey < M.Ney; ++ey)
    for (int ex = 0; ex < M.Nex; ++ex)
      for (int j = 0; j < np; ++j)
        for (int i = 0; i < np; ++i) {
          double mxx = 0, mxy = 0, myx = 0, myy = 0, rx = 0, ry = 0, Ex = 0, Ey = 0, Sx = 0, Sy = 0;
          for (int k = 0; k < np; ++k) {
            const size_t kx = G.at(k, j, ex, ey), ky = G.at(i, k, ex, ey);
            mxx += B.Dik(i, k) * Mx[kx]; mxy += B.Dik(j, k) * Mx[ky];
            myx += B.Dik(i, k) * My[kx]; myy += B.Dik(j, k) * My[ky];
            rx += B.Dik(i, k) * R[kx];   ry += B.Dik(j, k) * R[ky];
            Ex += B.Dik(i, k) * En[kx];  Ey += B.Dik(j, k) * En[ky];
            Sx += B.Dik(i, k) * S[kx];   Sy += B.Dik(j, k) * S[ky];
          }
          mxx *= invJx; myx *= invJx; rx *= invJx; Ex *= invJx; Sx *= invJx;
          mxy *= invJy; myy *= invJy; ry *= invJy; Ey *= invJy; Sy *= invJy;
          const size_t n = G.at(i, j, ex, ey);
          const double r = R[n], mx = Mx[n], my = My[n], en = En[n];
          const double ux = mx / r, uy = my / r;
          const double uxx = (mxx - ux * rx) / r, uxy = (mxy - ux * ry) / r;
          const double uyx = (myx - uy * rx) / r, uyy = (myy - uy * ry) / r;
          (void)uxy; (void)uyx;
          const double P = gm * (en - (mx * mx + my * my) / (2 * r));
          const double H = en + P + S[n];
          const double Pr = gm * (ux * ux + uy * uy) / 2;
          const double Pbx = gm * (Ex - (mx * mxx + my * myx) / r + (mx * mx + my * my) * rx / (2 * (r * r))) + Sx;
          const double Pby = gm * (Ey - (mx * mxy + my * myy) / r + (mx * mx + my * my) * ry / (2 * (r * r))) + Sy;
          const double dPx_dmx = gm * (-mxx / r + mx * rx / (r * r));
          const double dPx_dmy = gm * (-myx / r + my * rx / (r * r));
          const double dPx_dr = gm * ((mx * mxx + my * myx) / (r * r) - (mx * mx + my * my) * rx / (r * r * r));
          const double dPy_dmx = gm * (-mxy / r + mx * ry / (r * r));
          const double dPy_dmy = gm * (-myy / r + my * ry / (r * r));
          const double dPy_dr = gm * ((mx * mxy + my * myy) / (r * r) - (mx * mx + my * my) * ry / (r * r * r));
          const double HEx = Ex + Pbx, HEy = Ey + Pby;
          double dmx_mx = mxx / r + uxx - mx * rx / (r * r) + dPx_dmx;
          if (g_adj_mutation != 1) dmx_mx += uyy;
          const double dmx_my = dPx_dmy + mxy / r - mx * ry / (r * r);
          const double dmx_r = mxx * (-ux / r) + mx * (ux * rx / (r * r) - uxx / r) + dPx_dr +
                               mxy * (-uy / r) + mx * (uy * ry / (r * r) - uyy / r);
          const double dmy_mx = myx / r - my * rx / (r * r) + dPy_dmx;
          const double dmy_my = uxx + myy / r + uyy - my * ry / (r * r) + dPy_dmy;
          const double dmy_r = myx * (-ux / r) + my * (ux * rx / (r * r) - uxx / r) +
                               myy * (-uy / r) + my * (uy * ry / (r * r) - uyy / r) + dPy_dr;
          const double dE_mx = dPx_dmx * ux + HEx / r + (-gm * ux) * uxx + H * (-rx / (r * r)) +
                               dPy_dmx * uy + (-gm * ux) * uyy;
          const double dE_my = dPx_dmy * ux + (-gm * uy) * uxx + dPy_dmy * uy + HEy / r +
                               (-gm * uy) * uyy + H * (-ry / (r * r));
          const double dE_r = dPx_dr * ux + HEx * (-ux / r) + Pr * uxx + H * (ux * rx / (r * r) - uxx / r) +
                              dPy_dr * uy + HEy * (-uy / r) + Pr * uyy + H * (uy * ry / (r * r) - uyy / r);
          const double dE_E = g * (uxx + uyy);
          const double axx = a[n], ayy = a[N + n], aEE = a[3 * (size_t)N + n];
          Sl[n] = axx * dmx_mx + ayy * dmy_mx + aEE * dE_mx;
          Sl[N + n] = axx * dmx_my + ayy * dmy_my + aEE * dE_my;
          Sl[2 * (size_t)N + n] = axx * dmx_r + ayy * dmy_r + aEE * dE_r;
          Sl[3 * (size_t)N + n] = aEE * dE_E;
        }
}

// compute_adjoint_hyperbolic_rhs_local!, IGR_adjoints_2d/adjoint_pde.jl:226-320.
inline void adj_hyp_rhs_local2(const Grid2& G, double g, const double* q, const double* S, const double* a,
                               double* da) {
  const int np = G.np, N = G.N; const Basis& B = G.B; const Mesh2D& M = G.M;
  const double invJx = 1 / M.Jx, invJy = 1 / M.Jy;
  std::vector<double> fx(4 * (size_t)N), fy(4 * (size_t)N);
  auto qn = [&](size_t n, double o[5]) { o[0] = q[n]; o[1] = q[N + n]; o[2] = q[2 * (size_t)N + n]; o[3] = q[3 * (size_t)N + n]; o[4] = S[n]; };
  auto an = [&](size_t n, double o[4]) { for (int v = 0; v < 4; ++v) o[v] = a[(size_t)v * N + n]; };
  for (int n = 0; n < N; ++n) {   // :238-249
    double qq[5], aa[4], f[4];
    qn(n, qq); an(n, aa);
    adj_flux2(0, g, qq, aa, f);
    for (int v = 0; v < 4; ++v) fx[(size_t)v * N + n] = f[v];
    adj_flux2(1, g, qq, aa, f);
    for (int v = 0; v < 4; ++v) fy[(size_t)v * N + n] = f[v];
  }
  for (int ey = 0; ey < M.Ney; ++ey)    // :252-265
    for (int ex = 0; ex < M.Nex; ++ex)
      for (int j = 0; j < np; ++j)
        for (int i = 0; i < np; ++i) {
          double acc[4] = {0, 0, 0, 0};
          for (int k = 0; k < np; ++k)
            for (int v = 0; v < 4; ++v)
              acc[v] += invJx * B.Dik(i, k) * fx[(size_t)v * N + G.at(k, j, ex, ey)] +
                        invJy * B.Dik(j, k) * fy[(size_t)v * N + G.at(i, k, ex, ey)];
          for (int v = 0; v < 4; ++v) da[(size_t)v * N + G.at(i, j, ex, ey)] = acc[v];
        }
  for (int ey = 0; ey < M.Ney; ++ey)    // :268-287
    for (int ex = 0; ex < M.Nex; ++ex) {
      const int exR = ex == M.Nex - 1 ? 0 : ex + 1;
      for (int j = 0; j < np; ++j) {
        const size_t L = G.at(np - 1, j, ex, ey), Rn = G.at(0, j, exR, ey);
        double qL[5], qR[5], aL[4], aR[4], f[4];
        qn(L, qL); qn(Rn, qR); an(L, aL); an(Rn, aR);
        adj_llf2(0, g, qL, qR, aL, aR, f);
        for (int v = 0; v < 4; ++v) da[(size_t)v * N + L] += invJx * (f[v] - fx[(size_t)v * N + L]) / B.w[np - 1];
        for (int v = 0; v < 4; ++v) da[(size_t)v * N + Rn] -= invJx * (f[v] - fx[(size_t)v * N + Rn]) / B.w[0];
      }
    }
  for (int ey = 0; ey < M.Ney; ++ey)    // :290-309
    for (int ex = 0; ex < M.Nex; ++ex) {
      const int eyR = ey == M.Ney - 1 ? 0 : ey + 1;
      for (int i = 0; i < np; ++i) {
        const size_t L = G.at(i, np - 1, ex, ey), Rn = G.at(i, 0, ex, eyR);
        double qL[5], qR[5], aL[4], aR[4], f[4];
        qn(L, qL); qn(Rn, qR); an(L, aL); an(Rn, aR);
        adj_llf2(1, g, qL, qR, aL, aR, f);
        for (int v = 0; v < 4; ++v) da[(size_t)v * N + L] += invJy * (f[v] - fy[(size_t)v * N + L]) / B.w[np - 1];
        for (int v = 0; v < 4; ++v) da[(size_t)v * N + Rn] -= invJy * (f[v] - fy[(size_t)v * N + Rn]) / B.w[0];
      }
    }
  std::vector<double> Sl(4 * (size_t)N);   // :312-318
  adj_source_local2(G, g, q, S, a, Sl.data());
  for (size_t n = 0; n < 4 * (size_t)N; ++n) da[n] -= Sl[n];
}

// compute_adjoint_elliptic_rhs!, IGR_adjoints_2d/adjoint_pde.jl:332-357:
// div(a_mu) + grad(a_E) . u at every node.
inline void adj_elliptic_rhs2(const Grid2& G, const double* a, const double* q, double* b) {
  const int np = G.np, N = G.N; const Basis& B = G.B; const Mesh2D& M = G.M;
  const double invJx = 1 / M.Jx, invJy = 1 / M.Jy;
  for (int ey = 0; ey < M.Ney; ++ey)
    for (int ex = 0; ex < M.Nex; ++ex)
      for (int j = 0; j < np; ++j)
        for (int i = 0; i < np; ++i) {
          double dax = 0, day = 0, dEx = 0, dEy = 0;
          for (int k = 0; k < np; ++k) {
            dax += B.Dik(i, k) * a[G.at(k, j, ex, ey)];
            day += B.Dik(j, k) * a[N + G.at(i, k, ex, ey)];
            dEx += B.Dik(i, k) * a[3 * (size_t)N + G.at(k, j, ex, ey)];
            dEy += B.Dik(j, k) * a[3 * (size_t)N + G.at(i, k, ex, ey)];
          }
          dax *= invJx; dEx *= invJx;
          day *= invJy; dEy *= invJy;
          const size_t n = G.at(i, j, ex, ey);
          const double ux = q[n] / q[2 * (size_t)N + n], uy = q[N + n] / q[2 * (size_t)N + n];
          b[n] = dax + day + dEx * ux + dEy * uy;
        }
}

// compute_nonlocal_source!, IGR_adjoints_2d/adjoint_pde.jl:373-457.
inline void nonlocal_source2(const Grid2& G, const double* Pi, const double* q, const double* S, double* Sn) {
  const int np = G.np, N = G.N; const Basis& B = G.B; const Mesh2D& M = G.M;
  const double invJx = 1 / M.Jx, invJy = 1 / M.Jy;
  const double* R = q + 2 * (size_t)N;
  std::vector<double> ux(N), uy(N), Pdivu(N), Pxx(N), Pyx(N), Pxy(N), Pyy(N), Cx(N), Cy(N), Rtr(N);
  for (int n = 0; n < N; ++n) { ux[n] = q[n] / R[n]; uy[n] = q[N + n] / R[n]; }
  for (int ey = 0; ey < M.Ney; ++ey)     // :395-417
    for (int ex = 0; ex < M.Nex; ++ex)
      for (int j = 0; j < np; ++j)
        for (int i = 0; i < np; ++i) {
          double uxx = 0, uxy = 0, uyx = 0, uyy = 0, Sx = 0, Sy = 0;
          for (int k = 0; k < np; ++k) {
            const size_t kx = G.at(k, j, ex, ey), ky = G.at(i, k, ex, ey);
            uxx += B.Dik(i, k) * ux[kx]; uyx += B.Dik(i, k) * uy[kx];
            uxy += B.Dik(j, k) * ux[ky]; uyy += B.Dik(j, k) * uy[ky];
            Sx += B.Dik(i, k) * S[kx];   Sy += B.Dik(j, k) * S[ky];
          }
          uxx *= invJx; uyx *= invJx; Sx *= invJx;
          uxy *= invJy; uyy *= invJy; Sy *= invJy;
          const double divu = uxx + uyy;
          const size_t n = G.at(i, j, ex, ey);
          const double r = R[n], P = Pi[n];
          Pdivu[n] = P * divu;
          Pxx[n] = P * uxx; Pyx[n] = P * uyx;
          Pxy[n] = P * uxy; Pyy[n] = P * uyy;
          Cx[n] = P * Sx / r; Cy[n] = P * Sy / r;
          Rtr[n] = divu * divu + uxx * uxx + 2 * uxy * uyx + uyy * uyy;
        }
  for (int ey = 0; ey < M.Ney; ++ey)     // :420-455
    for (int ex = 0; ex < M.Nex; ++ex)
      for (int j = 0; j < np; ++j)
        for (int i = 0; i < np; ++i) {
          double dxPd = 0, dyPd = 0, dxPxx = 0, dyPyx = 0, dxPxy = 0, dyPyy = 0, dxCx = 0, dyCy = 0;
          for (int k = 0; k < np; ++k) {
            const size_t kx = G.at(k, j, ex, ey), ky = G.at(i, k, ex, ey);
            dxPd += B.Dik(i, k) * Pdivu[kx];
            dyPd += B.Dik(j, k) * Pdivu[ky];
            dxPxx += B.Dik(i, k) * Pxx[kx];
            dyPyx += B.Dik(j, k) * Pyx[ky];
            dxPxy += B.Dik(i, k) * Pxy[kx];
            dyPyy += B.Dik(j, k) * Pyy[ky];
            dxCx += B.Dik(i, k) * Cx[kx];
            dyCy += B.Dik(j, k) * Cy[ky];
          }
          dxPd *= invJx; dyPd *= invJy;
          dxPxx *= invJx; dyPyx *= invJy;
          dxPxy *= invJx; dyPyy *= invJy;
          dxCx *= invJx; dyCy *= invJy;
          const double Wx = dxPd + dxPxx + dyPyx;
          const double Wy = dyPd + dxPxy + dyPyy;
          const double divC = g_adj_mutation == 3 ? 0.0 : dxCx + dyCy;
          const size_t n = G.at(i, j, ex, ey);
          const double r = R[n], mx = q[n], my = q[N + n];
          const double sm = g_adj_mutation == 4 ? 2.0 : -2.0;
          Sn[n] = sm * Wx / r;
          Sn[N + n] = sm * Wy / r;
          Sn[2 * (size_t)N + n] = (Pi[n] * Rtr[n] + divC) / r + 2 * (Wx * mx + Wy * my) / (r * r);
          Sn[3 * (size_t)N + n] = 0.0;
        }
}

// compute_adjoint_rhs_conservative!, IGR_adjoints_2d/adjoint_pde.jl:471-506.
// Pi is the warm start on entry and the solution on return.
inline int adj_rhs_conservative2(const Grid2& G, double g, double alpha, int n_iter, int solver, const double* q,
                                 const double* S, const double* a, double* Pi, double* da) {
  const int np = G.np, N = G.N; const Basis& B = G.B; const Mesh2D& M = G.M;
  adj_hyp_rhs_local2(G, g, q, S, a, da);
  if (alpha > 0) {
    std::vector<double> b(N), Md(N), x(Pi, Pi + N), Sn(4 * (size_t)N);
    adj_elliptic_rhs2(G, a, q, b.data());
    for (int ey = 0; ey < M.Ney; ++ey)
      for (int ex = 0; ex < M.Nex; ++ex)
        for (int j = 0; j < np; ++j)
          for (int i = 0; i < np; ++i) b[G.at(i, j, ex, ey)] *= alpha * B.w[i] * B.w[j] * M.Jx * M.Jy;
    const double* rho = q + 2 * (size_t)N;
    sip_diag2(G, alpha, rho, Md.data());
    auto A = [&](const double* in, double* out) { apply_sip2(G, alpha, in, rho, out); };
    int rc = run_solver(solver, A, x, b, Md, n_iter, B.p, nullptr);
    if (rc) return rc;
    std::copy(x.begin(), x.end(), Pi);
    nonlocal_source2(G, Pi, q, S, Sn.data());
    for (size_t n = 0; n < 4 * (size_t)N; ++n) da[n] += Sn[n];
  }
  return 0;
}

// run_forward_store, IGR_adjoints_2d/forward.jl:111-158.  q0 (4N) in;
// snaps receives (n_steps + 1) snapshots of (mx, my, rho, E, Sigma), 5N each.
inline int run_forward_store2(const Grid2& G, double g, double alpha, int n_iter, int solver, double dt,
                              double Tend, const double* q0, std::vector<double>& snaps, int* n_steps_out) {
  const int N = G.N;
  const size_t nvN = 4 * (size_t)N;
  std::vector<double> q(q0, q0 + nvN), S(N, 0.0), q1(nvN), q2(nvN), dq(nvN);
  auto push = [&]() {
    snaps.insert(snaps.end(), q.begin(), q.end());
    snaps.insert(snaps.end(), S.begin(), S.end());
  };
  snaps.clear();
  if (alpha > 0) { int rc = solve_elliptic2<double>(G, alpha, n_iter, solver, q.data(), S.data()); if (rc) return rc; }
  push();
  const int n_steps = (int)std::ceil(Tend / dt);
  dt = Tend / n_steps;
  if (n_steps_out) *n_steps_out = n_steps;
  auto rhs = [&](const double* qs, double* d) { return euler_rhs2<double>(G, g, alpha, n_iter, solver, qs, S.data(), d); };
  for (int s = 0; s < n_steps; ++s) {
    int rc = ssp_rk3_step(rhs, nvN, q.data(), dt, q1, q2, dq);
    if (rc) return rc;
    if (alpha > 0) { rc = solve_elliptic2<double>(G, alpha, n_iter, solver, q.data(), S.data()); if (rc) return rc; }
    push();
  }
  return 0;
}

// Terminal condition of run_adjoint_conservative, IGR_adjoints_2d/adjoint_pde.jl:535-576
// (the L2 gradient of J at t = T).  WINDOWED_DENSITY and DISC_KE are the
// objectives the reference does not define; their L2 gradients follow from
// the same quadrature sums as objective2.  prm as in objective2 (window
// centre and width / disc centre and radius).
inline int adj_terminal2(const Grid2& G, int obj, const double* prm, double g, const double* qT, double* a) {
  const int np = G.np, N = G.N; const Basis& B = G.B; const Mesh2D& M = G.M;
  const double* mx = qT; const double* my = qT + N; const double* r = qT + 2 * (size_t)N; const double* E = qT + 3 * (size_t)N;
  for (size_t n = 0; n < 4 * (size_t)N; ++n) a[n] = 0.0;
  double KE = 0, Mt = 0;
  if (obj == OBJ_KINETIC_ENERGY)
    for (int ey = 0; ey < M.Ney; ++ey)
      for (int ex = 0; ex < M.Nex; ++ex)
        for (int j = 0; j < np; ++j)
          for (int i = 0; i < np; ++i) {
            const size_t n = G.at(i, j, ex, ey);
            const double wJ = B.w[i] * B.w[j] * M.Jx * M.Jy;
            KE += wJ * (mx[n] * mx[n] + my[n] * my[n]) / (2 * r[n]);
            Mt += wJ * r[n];
          }
  for (int ey = 0; ey < M.Ney; ++ey)
    for (int ex = 0; ex < M.Nex; ++ex)
      for (int j = 0; j < np; ++j)
        for (int i = 0; i < np; ++i) {
          const size_t n = G.at(i, j, ex, ey);
          const double x = M.x(B, i, ex), y = M.y(B, j, ey);
          const double ux = mx[n] / r[n], uy = my[n] / r[n];
          switch (obj) {
            case OBJ_L2_DENSITY: a[2 * (size_t)N + n] = 2 * r[n]; break;
            case OBJ_L2_PRESSURE: {
              const double P = (g - 1) * (E[n] - (mx[n] * mx[n] + my[n] * my[n]) / (2 * r[n]));
              a[n] = -2 * (g - 1) * P * ux;
              a[N + n] = -2 * (g - 1) * P * uy;
              a[2 * (size_t)N + n] = (g - 1) * P * (ux * ux + uy * uy);
              a[3 * (size_t)N + n] = 2 * (g - 1) * P;
              break;
            }
            case OBJ_WEIGHTED_MOMENTUM:
              a[n] = std::sin(2 * M_PI * x / M.Lx) * std::sin(2 * M_PI * y / M.Ly);
              break;
            case OBJ_KINETIC_ENERGY:
              a[n] = ux / Mt;
              a[N + n] = uy / Mt;
              a[2 * (size_t)N + n] = -(mx[n] * mx[n] + my[n] * my[n]) / (2 * (r[n] * r[n])) / Mt - KE / (Mt * Mt);
              break;
            case OBJ_WINDOWED_KE: {
              const double dx_ = x - prm[0], dy_ = y - prm[1];
              const double w = std::exp(-(dx_ * dx_ + dy_ * dy_) / (prm[2] * prm[2]));
              a[n] = w * ux;
              a[N + n] = w * uy;
              a[2 * (size_t)N + n] = -w * (ux * ux + uy * uy) / 2;
              break;
            }
            case OBJ_WINDOWED_DENSITY: {
              const double dx_ = x - prm[0], dy_ = y - prm[1];
              a[2 * (size_t)N + n] = std::exp(-(dx_ * dx_ + dy_ * dy_) / (prm[2] * prm[2]));
              break;
            }
            case OBJ_DISC_KE: {
              const double dx_ = x - prm[0], dy_ = y - prm[1];
              const double w = (dx_ * dx_ + dy_ * dy_ < prm[2] * prm[2]) ? 1.0 : 0.0;
              a[n] = w * ux;
              a[N + n] = w * uy;
              a[2 * (size_t)N + n] = -w * (ux * ux + uy * uy) / 2;
              break;
            }
            default: return -1;
          }
        }
  return 0;
}

// run_adjoint_conservative, IGR_adjoints_2d/adjoint_pde.jl:522-611.  snaps
// as written by run_forward_store2; a (4N) receives the adjoint at t = 0.
// The step is min(dt, T) as passed (the reference does not re-derive it).
inline int run_adjoint_conservative2(const Grid2& G, const double* snaps, int n_steps, double dt, double Tend,
                                     double g, double alpha, int n_iter, int solver, int obj, const double* prm,
                                     double* a) {
  const int N = G.N;
  const size_t nvN = 4 * (size_t)N, snap = 5 * (size_t)N;
  int rc = adj_terminal2(G, obj, prm, g, snaps + (size_t)n_steps * snap, a);
  if (rc) return rc;
  std::vector<double> Pi(N, 0.0), a1(nvN), a2(nvN), da(nvN);
  for (int n = n_steps; n >= 1; --n) {
    const double h = std::min(dt, Tend);
    const double* qn = snaps + (size_t)n * snap;
    const double* Sn = qn + nvN;
    auto rhs = [&](const double* as, double* d) {
      return adj_rhs_conservative2(G, g, alpha, n_iter, solver, qn, Sn, as, Pi.data(), d);
    };
    rc = rhs(a, da.data()); if (rc) return rc;
    for (size_t k = 0; k < nvN; ++k) a1[k] = a[k] + h * da[k];
    rc = rhs(a1.data(), da.data()); if (rc) return rc;
    for (size_t k = 0; k < nvN; ++k) a2[k] = 3.0 / 4.0 * a[k] + 1.0 / 4.0 * a1[k] + 1.0 / 4.0 * h * da[k];
    rc = rhs(a2.data(), da.data()); if (rc) return rc;
    for (size_t k = 0; k < nvN; ++k) a[k] = 1.0 / 3.0 * a[k] + 2.0 / 3.0 * a2[k] + 2.0 / 3.0 * h * da[k];
  }
  return 0;
}

// =====================================================================
//                              1D
// =====================================================================

// adjoint_flux_local, IGR_adjoints_1d/adjoint_pde.jl:46-66.  q = (mu, rho, E, Sigma).
inline void adj_flux1(double g, const double q[4], const double a[3], double f[3]) {
  const double m = q[0], r = q[1], E = q[2], S = q[3];
  const double u = m / r;
  const double P = pressure1(g, m, r, E);
  const double Pb = P + S;
  const double dP_dE = g - 1;
  const double dP_dm = -(g - 1) * u;
  const double dP_dr = (g - 1) * (u * u) / 2;
  f[0] = a[0] * (2 * u + dP_dm) + a[1] + a[2] * ((E + Pb) / r + dP_dm * u);
  f[1] = a[0] * (-(u * u) + dP_dr) + a[2] * (dP_dr * u - (E + Pb) * u / r);
  f[2] = a[0] * dP_dE + a[2] * (1 + dP_dE) * u;
}

// adjoint_llf_flux_local, IGR_adjoints_1d/adjoint_pde.jl:75-88.
inline void adj_llf1(double g, const double qL[4], const double qR[4], const double aL[3], const double aR[3],
                     double f[3]) {
  const double lam = max_(wave1(g, qL[0], qL[1], qL[2]), wave1(g, qR[0], qR[1], qR[2]));
  double fL[3], fR[3];
  adj_flux1(g, qL, aL, fL);
  adj_flux1(g, qR, aR, fR);
  const double sgn = g_adj_mutation == 2 ? -1.0 : 1.0;
  for (int v = 0; v < 3; ++v) f[v] = (fL[v] + fR[v]) / 2 - sgn * lam * (aL[v] - aR[v]) / 2;
}

// compute_adjoint_source_local!, IGR_adjoints_1d/adjoint_pde.jl:102-176.
inline void adj_source_local1(const Grid1& G, double g, const double* q, const double* S, const double* a,
                              double* Sl) {
  const int np = G.np, N = G.N, Ne = G.M.Ne; const Basis& B = G.B;
  const double invJ = 1 / G.M.J;
  const double* Mu = q; const double* R = q + N; const double* En = q + 2 * (size_t)N;
  for (int e = 0; e < Ne; ++e)
    for (int i = 0; i < np; ++i) {
      double mx = 0, rx = 0, Ex = 0, Sx = 0;
      for (int j = 0; j < np; ++j) {
        const size_t k = G.at(j, e);
        mx += B.Dik(i, j) * Mu[k];
        rx += B.Dik(i, j) * R[k];
        Ex += B.Dik(i, j) * En[k];
        Sx += B.Dik(i, j) * S[k];
      }
      mx *= invJ; rx *= invJ; Ex *= invJ; Sx *= invJ;
      const size_t n = G.at(i, e);
      const double r = R[n], m = Mu[n], En_ = En[n];
      const double u = m / r;
      const double P = pressure1(g, m, r, En_);
      const double Pb = P + S[n];
      const double ux = (mx - u * rx) / r;
      const double dP_dE = g - 1;
      const double dP_dm = -(g - 1) * u;
      const double dP_dr = (g - 1) * (u * u) / 2;
      const double P_x = (g - 1) * (Ex - m * mx / r + (m * m) * rx / (2 * (r * r)));
      const double drP_x = (g - 1) * (m * mx / (r * r) - (m * m) * rx / (r * r * r));
      const double dmP_x = -(g - 1) * ux;
      const double Pb_x = P_x + Sx;
      const double am = a[n], aE = a[2 * (size_t)N + n];
      const double dm_u = 1 / r;
      const double dm_ux = -rx / (r * r);
      const double dr_u = -u / r;
      const double dr_ux = -mx / (r * r) + 2 * m * rx / (r * r * r);
      double dm_divFm = mx * dm_u + m * dm_ux + dmP_x;
      if (g_adj_mutation != 1) dm_divFm = mx * dm_u + ux + m * dm_ux + dmP_x;
      const double dm_divFE = dmP_x * u + (Ex + Pb_x) * dm_u + dP_dm * ux + (En_ + Pb) * dm_ux;
      Sl[n] = am * dm_divFm + aE * dm_divFE;
      const double dr_divFm = mx * dr_u + m * dr_ux + drP_x;
      const double dr_divFE = drP_x * u + (Ex + Pb_x) * dr_u + dP_dr * ux + (En_ + Pb) * dr_ux;
      Sl[N + n] = am * dr_divFm + aE * dr_divFE;
      const double dE_divFE = (1 + dP_dE) * ux;
      Sl[2 * (size_t)N + n] = aE * dE_divFE;
    }
}

// compute_adjoint_hyperbolic_rhs_local!, IGR_adjoints_1d/adjoint_pde.jl:191-264.
inline void adj_hyp_rhs_local1(const Grid1& G, double g, const double* q, const double* S, const double* a,
                               double* da) {
  const int np = G.np, N = G.N, Ne = G.M.Ne; const Basis& B = G.B;
  const double invJ = 1 / G.M.J;
  std::vector<double> f(3 * (size_t)N);
  auto qn = [&](size_t n, double o[4]) { o[0] = q[n]; o[1] = q[N + n]; o[2] = q[2 * (size_t)N + n]; o[3] = S[n]; };
  auto an = [&](size_t n, double o[3]) { for (int v = 0; v < 3; ++v) o[v] = a[(size_t)v * N + n]; };
  for (int n = 0; n < N; ++n) {
    double qq[4], aa[3], ff[3];
    qn(n, qq); an(n, aa);
    adj_flux1(g, qq, aa, ff);
    for (int v = 0; v < 3; ++v) f[(size_t)v * N + n] = ff[v];
  }
  for (int e = 0; e < Ne; ++e)
    for (int i = 0; i < np; ++i) {
      const size_t n = G.at(i, e);
      for (int v = 0; v < 3; ++v) da[(size_t)v * N + n] = 0.0;
      for (int j = 0; j < np; ++j)
        for (int v = 0; v < 3; ++v) da[(size_t)v * N + n] += invJ * B.Dik(i, j) * f[(size_t)v * N + G.at(j, e)];
    }
  for (int e = 0; e < Ne; ++e) {
    const int eL = e == 0 ? Ne - 1 : e - 1, eR = e == Ne - 1 ? 0 : e + 1;
    double qL[4], qR[4], aL[3], aR[3], fR[3], fL[3];
    qn(G.at(np - 1, e), qL); qn(G.at(0, eR), qR); an(G.at(np - 1, e), aL); an(G.at(0, eR), aR);
    adj_llf1(g, qL, qR, aL, aR, fR);
    qn(G.at(np - 1, eL), qL); qn(G.at(0, e), qR); an(G.at(np - 1, eL), aL); an(G.at(0, e), aR);
    adj_llf1(g, qL, qR, aL, aR, fL);
    const size_t nr = G.at(np - 1, e), nl = G.at(0, e);
    for (int v = 0; v < 3; ++v) da[(size_t)v * N + nr] += invJ * (fR[v] - f[(size_t)v * N + nr]) / B.w[np - 1];
    for (int v = 0; v < 3; ++v) da[(size_t)v * N + nl] -= invJ * (fL[v] - f[(size_t)v * N + nl]) / B.w[0];
  }
  std::vector<double> Sl(3 * (size_t)N);
  adj_source_local1(G, g, q, S, a, Sl.data());
  for (size_t n = 0; n < 3 * (size_t)N; ++n) da[n] -= Sl[n];
}

// compute_adjoint_elliptic_rhs!, IGR_adjoints_1d/adjoint_pde.jl:278-301.
inline void adj_elliptic_rhs1(const Grid1& G, const double* a, const double* q, double* b) {
  const int np = G.np, N = G.N, Ne = G.M.Ne; const Basis& B = G.B;
  const double invJ = 1 / G.M.J;
  for (int e = 0; e < Ne; ++e)
    for (int i = 0; i < np; ++i) {
      double dam = 0, daE = 0;
      for (int j = 0; j < np; ++j) {
        dam += B.Dik(i, j) * a[G.at(j, e)];
        daE += B.Dik(i, j) * a[2 * (size_t)N + G.at(j, e)];
      }
      dam *= invJ; daE *= invJ;
      const size_t n = G.at(i, e);
      const double u = q[n] / q[N + n];
      b[n] = dam + daE * u;
    }
}

// compute_nonlocal_source!, IGR_adjoints_1d/adjoint_pde.jl:321-387.
inline void nonlocal_source1(const Grid1& G, const double* Pi, const double* q, const double* S, double* Sn) {
  const int np = G.np, N = G.N, Ne = G.M.Ne; const Basis& B = G.B;
  const double invJ = 1 / G.M.J;
  const double* Mu = q; const double* R = q + N;
  std::vector<double> Pux(N), PSx(N);
  for (int e = 0; e < Ne; ++e)
    for (int i = 0; i < np; ++i) {
      double uxi = 0, Sxi = 0;
      for (int j = 0; j < np; ++j) {
        const size_t k = G.at(j, e);
        const double uj = Mu[k] / R[k];
        uxi += B.Dik(i, j) * uj;
        Sxi += B.Dik(i, j) * S[k];
      }
      uxi *= invJ; Sxi *= invJ;
      const size_t n = G.at(i, e);
      Pux[n] = Pi[n] * uxi;
      PSx[n] = Pi[n] * Sxi / R[n];
    }
  for (int e = 0; e < Ne; ++e)
    for (int i = 0; i < np; ++i) {
      double dPux = 0, dPSx = 0;
      for (int j = 0; j < np; ++j) {
        dPux += B.Dik(i, j) * Pux[G.at(j, e)];
        dPSx += B.Dik(i, j) * PSx[G.at(j, e)];
      }
      dPux *= invJ; dPSx *= invJ;
      if (g_adj_mutation == 3) dPSx = 0.0;
      double uxi = 0;
      for (int j = 0; j < np; ++j) uxi += B.Dik(i, j) * (Mu[G.at(j, e)] / R[G.at(j, e)]);
      uxi *= invJ;
      const size_t n = G.at(i, e);
      const double r = R[n], m = Mu[n];
      const double sm = g_adj_mutation == 4 ? 4.0 : -4.0;
      Sn[n] = sm * dPux / r;
      Sn[N + n] = (2 * Pi[n] * (uxi * uxi) + dPSx) / r + 4 * dPux * m / (r * r);
      Sn[2 * (size_t)N + n] = 0.0;
    }
}

// compute_adjoint_rhs_conservative!, IGR_adjoints_1d/adjoint_pde.jl:406-451.
inline int adj_rhs_conservative1(const Grid1& G, double g, double alpha, int n_iter, int solver, const double* q,
                                 const double* S, const double* a, double* Pi, double* da) {
  const int np = G.np, N = G.N, Ne = G.M.Ne; const Basis& B = G.B;
  adj_hyp_rhs_local1(G, g, q, S, a, da);
  if (alpha > 0) {
    std::vector<double> b(N), Md(N), x(Pi, Pi + N), Sn(3 * (size_t)N);
    adj_elliptic_rhs1(G, a, q, b.data());
    for (int e = 0; e < Ne; ++e)
      for (int i = 0; i < np; ++i) b[G.at(i, e)] *= alpha * B.w[i] * G.M.J;
    const double* rho = q + N;
    sip_diag1(G, alpha, rho, Md.data());
    auto A = [&](const double* in, double* out) { apply_sip1(G, alpha, in, rho, out); };
    int rc = run_solver(solver, A, x, b, Md, n_iter, B.p, nullptr);
    if (rc) return rc;
    std::copy(x.begin(), x.end(), Pi);
    nonlocal_source1(G, Pi, q, S, Sn.data());
    for (size_t n = 0; n < 3 * (size_t)N; ++n) da[n] += Sn[n];
  }
  return 0;
}

// run_forward_store, IGR_adjoints_1d/forward.jl:137-187.  Snapshots of
// (mu, rho, E, Sigma), 4N each.
inline int run_forward_store1(const Grid1& G, double g, double alpha, int n_iter, int solver, double dt,
                              double Tend, const double* q0, std::vector<double>& snaps, int* n_steps_out) {
  const int N = G.N;
  const size_t nvN = 3 * (size_t)N;
  std::vector<double> q(q0, q0 + nvN), S(N, 0.0), q1(nvN), q2(nvN), dq(nvN);
  auto push = [&]() {
    snaps.insert(snaps.end(), q.begin(), q.end());
    snaps.insert(snaps.end(), S.begin(), S.end());
  };
  snaps.clear();
  if (alpha > 0) { int rc = solve_elliptic1<double>(G, alpha, n_iter, solver, q.data(), S.data()); if (rc) return rc; }
  push();
  const int n_steps = (int)std::ceil(Tend / dt);
  dt = Tend / n_steps;
  if (n_steps_out) *n_steps_out = n_steps;
  auto rhs = [&](const double* qs, double* d) { return euler_rhs1<double>(G, g, alpha, n_iter, solver, qs, S.data(), d); };
  for (int s = 0; s < n_steps; ++s) {
    int rc = ssp_rk3_step(rhs, nvN, q.data(), dt, q1, q2, dq);
    if (rc) return rc;
    if (alpha > 0) { rc = solve_elliptic1<double>(G, alpha, n_iter, solver, q.data(), S.data()); if (rc) return rc; }
    push();
  }
  return 0;
}

// Terminal condition, IGR_adjoints_1d/adjoint_pde.jl:479-513.
inline int adj_terminal1(const Grid1& G, int obj, double g, const double* qT, double* a) {
  const int np = G.np, N = G.N, Ne = G.M.Ne; const Basis& B = G.B;
  const double* m = qT; const double* r = qT + N; const double* E = qT + 2 * (size_t)N;
  for (size_t n = 0; n < 3 * (size_t)N; ++n) a[n] = 0.0;
  if (obj == OBJ_L2_DENSITY) {
    for (int n = 0; n < N; ++n) a[N + n] = 2 * r[n];
  } else if (obj == OBJ_L2_PRESSURE) {
    for (int n = 0; n < N; ++n) {
      const double u = m[n] / r[n];
      const double P = (g - 1) * (E[n] - (m[n] * m[n]) / (2 * r[n]));
      a[n] = -2 * (g - 1) * P * u;
      a[N + n] = (g - 1) * P * (u * u);
      a[2 * (size_t)N + n] = 2 * (g - 1) * P;
    }
  } else if (obj == OBJ_WEIGHTED_MOMENTUM) {
    const double Ld = Ne * G.M.dx;
    for (int e = 0; e < Ne; ++e)
      for (int i = 0; i < np; ++i) a[G.at(i, e)] = std::sin(2 * M_PI * G.M.x(B, i, e) / Ld);
  } else if (obj == OBJ_KINETIC_ENERGY) {
    double KE = 0, Mt = 0;
    for (int e = 0; e < Ne; ++e)
      for (int i = 0; i < np; ++i) {
        const double wJ = B.w[i] * G.M.J;
        const size_t n = G.at(i, e);
        KE += wJ * (m[n] * m[n]) / (2 * r[n]);
        Mt += wJ * r[n];
      }
    for (int n = 0; n < N; ++n) {
      a[n] = (m[n] / r[n]) / Mt;
      a[N + n] = -(m[n] * m[n]) / (2 * (r[n] * r[n])) / Mt - KE / (Mt * Mt);
    }
  } else {
    return -1;
  }
  return 0;
}

// run_adjoint_conservative, IGR_adjoints_1d/adjoint_pde.jl:466-552.
inline int run_adjoint_conservative1(const Grid1& G, const double* snaps, int n_steps, double dt, double Tend,
                                     double g, double alpha, int n_iter, int solver, int obj, double* a) {
  const int N = G.N;
  const size_t nvN = 3 * (size_t)N, snap = 4 * (size_t)N;
  int rc = adj_terminal1(G, obj, g, snaps + (size_t)n_steps * snap, a);
  if (rc) return rc;
  std::vector<double> Pi(N, 0.0), a1(nvN), a2(nvN), da(nvN);
  for (int n = n_steps; n >= 1; --n) {
    const double h = std::min(dt, Tend);
    const double* qn = snaps + (size_t)n * snap;
    const double* Sn = qn + nvN;
    auto rhs = [&](const double* as, double* d) {
      return adj_rhs_conservative1(G, g, alpha, n_iter, solver, qn, Sn, as, Pi.data(), d);
    };
    rc = rhs(a, da.data()); if (rc) return rc;
    for (size_t k = 0; k < nvN; ++k) a1[k] = a[k] + h * da[k];
    rc = rhs(a1.data(), da.data()); if (rc) return rc;
    for (size_t k = 0; k < nvN; ++k) a2[k] = 3.0 / 4.0 * a[k] + 1.0 / 4.0 * a1[k] + 1.0 / 4.0 * h * da[k];
    rc = rhs(a2.data(), da.data()); if (rc) return rc;
    for (size_t k = 0; k < nvN; ++k) a[k] = 1.0 / 3.0 * a[k] + 2.0 / 3.0 * a2[k] + 2.0 / 3.0 * h * da[k];
  }
  return 0;
}

}  // namespace orc
