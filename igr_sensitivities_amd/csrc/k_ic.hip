// Initial conditions built on the device, straight into a context's state:
// the reference's IC builders evaluated at every node, one thread per node,
// so a large mesh needs no host-side construction and no upload.
//
//   2D (IGR_adjoints_2d/initial_conditions.jl):
//     IGR_IC_CONSTANT           init_constant_state            :18-25
//     IGR_IC_ISENTROPIC_VORTEX  init_isentropic_vortex         :42-71
//     IGR_IC_TAYLOR_GREEN       build_taylor_green_ic!         :105-121
//     IGR_IC_SEDOV_BLAST        build_sedov_blast_ic!          :136-151
//     IGR_IC_OSEEN_PAIR_BLAST   build_oseen_pair_blast_ic!     :175-232
//   1D (IGR_adjoints_1d/initial_conditions.jl):
//     IGR_IC_CONSTANT           init_constant_state            :52-60
//     IGR_IC_ENTROPY_WAVE_INIT  init_entropy_wave              :24-44
//     IGR_IC_ACOUSTIC_PULSE .. IGR_IC_VELOCITY_PULSE
//                               build_<kind>_ic!               :94-210
//
// Expressions follow the reference's evaluation order (Julia's left-to-right
// * and /, x^2 as x*x, round to nearest with ties to even), so the only
// differences from the reference are those of the device's exp, sin, cos and
// pow.  The optional output dq = d q0 / d A (conservative variables) is the
// tangent seed of the amplitude parameter (ic_dA_weights, :246-258 in 2D and
// :227-244 in 1D, carried through the primitive-to-conservative map).
#include <cmath>
#include "igr_ctx.hpp"
#include "host_common.hpp"

namespace igr {

constexpr int IC_MAX_PRM = 32;
constexpr int OSEEN_NTAB = 4000;     // initial_conditions.jl:183

struct IcArgs {
  int kind, dim, np, Nex, Ney;       // 1D: Nex = N_e, Ney = 1
  double Lx, Ly, dx, dy, Jx, Jy;
  double xi[MAXNP];
  double A, gamma;
  double prm[IC_MAX_PRM];
  int n_prm;
  const double* tab;                 // Oseen enthalpy-defect table
  double* q;                         // nv*N state out
  double* dq;                        // nv*N d q / d A out, or null
  size_t N;
};

// Oseen enthalpy-defect table D(r_i) = (Gamma/2pi)^2 int_{r_i}^{2L} (1 - exp(-s^2/a^2))^2 / s^3 ds
// by the trapezoid rule accumulated from the outer end, in the reference's
// order (initial_conditions.jl:179-193); one thread, since the order is serial.
__global__ void k_oseen_table(double* tab, double gam, double a_core, double M_star, double L) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double two_pi = 2.0 * M_PI;
  const double G = M_star * sqrt(gam) * two_pi * a_core / 0.6383;
  const double r_max = 2.0 * L;
  const double pg = G / two_pi;
  const double pref = pg * pg;
  double acc = 0.0;
  tab[OSEEN_NTAB - 1] = 0.0;
  auto r_at = [&](int i0) { return (double)i0 * r_max / (double)(OSEEN_NTAB - 1); };   // range(0, r_max, length=n_tab)
  auto f = [&](double s) {
    const double e = 1.0 - exp(-(s * s) / (a_core * a_core));
    return e * e / (s * s * s);
  };
  for (int i0 = OSEEN_NTAB - 2; i0 >= 0; --i0) {
    const double s1 = r_at(i0), s2 = r_at(i0 + 1);
    const double f1 = s1 < 1e-12 ? 0.0 : f(s1);
    const double f2 = f(s2);
    acc += 0.5 * (f1 + f2) * (s2 - s1);
    tab[i0] = pref * acc;
  }
}

// defect(r) of initial_conditions.jl:194-200 (1-based t, as the reference)
__device__ __forceinline__ double oseen_defect(const double* tab, double r, double r_max) {
  if (r >= r_max) return 0.0;
  const double t = r / r_max * (double)(OSEEN_NTAB - 1) + 1.0;
  double fi = floor(t);
  fi = fmin(fmax(fi, 1.0), (double)(OSEEN_NTAB - 1));
  const int i = (int)fi;
  const double w = t - fi;
  return (1.0 - w) * tab[i - 1] + w * tab[i];
}

__global__ __launch_bounds__(256) void k_build_ic(IcArgs a) {
  const size_t N = a.N;
  const double g = a.gamma, A = a.A;
  const double two_pi = 2.0 * M_PI;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (size_t)gridDim.x * blockDim.x) {
    if (a.dim == 2) {
      const int i = (int)(n % a.np);
      const int j = (int)((n / a.np) % a.np);
      const size_t e = n / ((size_t)a.np * a.np);
      const int ex = (int)(e % a.Nex), ey = (int)(e / a.Nex);
      const double x = (double)ex * a.dx + a.Jx * (1.0 + a.xi[i]);   // IGR_adjoints_2d/mesh.jl:48-53
      const double y = (double)ey * a.dy + a.Jy * (1.0 + a.xi[j]);
      double rho = 1.0, ux = 0.0, uy = 0.0, p = 1.0;
      double drho = 0.0, dux = 0.0, duy = 0.0, dp = 0.0;
      if (a.kind == IGR_IC_CONSTANT) {
        rho = a.prm[0]; ux = a.prm[1]; uy = a.prm[2]; p = a.prm[3];
      } else if (a.kind == IGR_IC_ISENTROPIC_VORTEX) {
        const double beta = a.prm[0];
        const double ddx = x - a.Lx / 2, ddy = y - a.Ly / 2;
        const double r2 = ddx * ddx + ddy * ddy;
        const double ef = exp((1 - r2) / 2);
        const double du = -ddy * (beta / two_pi) * ef;
        const double dv = ddx * (beta / two_pi) * ef;
        const double dT = -(g - 1) * (beta * beta) / (8 * g * (M_PI * M_PI)) * exp(1 - r2);
        rho = pow(1 + dT, 1 / (g - 1));
        ux = a.prm[1] + du;
        uy = a.prm[2] + dv;
        p = pow(rho, g);
      } else if (a.kind == IGR_IC_TAYLOR_GREEN) {
        const double cx = two_pi * x / a.Lx, cy = two_pi * y / a.Ly;
        const double sx = sin(cx), cxs = cos(cx), sy = sin(cy), cys = cos(cy);
        ux = A * sx * cys;
        uy = -A * cxs * sy;
        dux = sx * cys;
        duy = -cxs * sy;
      } else if (a.kind == IGR_IC_SEDOV_BLAST) {
        const double p_bg = a.prm[0], sg = a.prm[1];
        const double ddx = x - a.Lx / 2, ddy = y - a.Ly / 2;
        const double r2 = ddx * ddx + ddy * ddy;
        const double ee = exp(-r2 / (sg * sg));
        p = p_bg + A * ee;
        dp = ee;
      } else {   // IGR_IC_OSEEN_PAIR_BLAST
        const double a_core = a.prm[0], M_star = a.prm[1], bxc = a.prm[2], byc = a.prm[3], sg = a.prm[4];
        const double L = a.Lx;
        const double G = M_star * sqrt(g) * two_pi * a_core / 0.6383;
        const double r_max = 2.0 * L;
        double h = g / (g - 1);
        const int nv = (a.n_prm - 5) / 3;
        for (int k = 0; k < nv; ++k) {
          const double xc = a.prm[5 + 3 * k], yc = a.prm[6 + 3 * k], sgn = a.prm[7 + 3 * k];
          const double Gk = sgn * G;
          for (int sxi = -1; sxi <= 1; ++sxi)
            for (int syi = -1; syi <= 1; ++syi) {
              const double ddx = x - (xc + sxi * L), ddy = y - (yc + syi * L);
              const double r2 = ddx * ddx + ddy * ddy;
              if (r2 < 1e-24) continue;
              const double vr = Gk / (two_pi * r2) * (1.0 - exp(-r2 / (a_core * a_core)));
              ux += -vr * ddy;
              uy += vr * ddx;
            }
          double ddx = x - xc, ddy = y - yc;
          ddx -= L * rint(ddx / L);
          ddy -= L * rint(ddy / L);
          h -= oseen_defect(a.tab, sqrt(ddx * ddx + ddy * ddy), r_max);
        }
        const double p_vort = pow((g - 1) * h / g, g / (g - 1));
        rho = pow(p_vort, 1.0 / g);
        const double r2b = (x - bxc) * (x - bxc) + (y - byc) * (y - byc);
        const double ee = exp(-r2b / (sg * sg));
        p = p_vort + A * ee;
        dp = ee;
      }
      a.q[n] = rho * ux;
      a.q[N + n] = rho * uy;
      a.q[2 * N + n] = rho;
      a.q[3 * N + n] = p / (g - 1) + 0.5 * rho * (ux * ux + uy * uy);
      if (a.dq) {
        a.dq[n] = rho * dux + ux * drho;
        a.dq[N + n] = rho * duy + uy * drho;
        a.dq[2 * N + n] = drho;
        a.dq[3 * N + n] = dp / (g - 1) + 0.5 * (drho * (ux * ux + uy * uy) + 2.0 * rho * (ux * dux + uy * duy));
      }
    } else {
      const int i = (int)(n % a.np);
      const int e = (int)(n / a.np);
      const double L = a.Lx;
      const double x = (double)e * a.dx + a.Jx * (1.0 + a.xi[i]);    // 1D mesh.jl:26-30
      double rho = 1.0, u = 0.0, p = 1.0, drho = 0.0, du = 0.0, dp = 0.0;
      if (a.kind == IGR_IC_CONSTANT) {
        rho = a.prm[0]; u = a.prm[1]; p = a.prm[2];
      } else if (a.kind == IGR_IC_ENTROPY_WAVE_INIT) {
        rho = 1.0 + 0.2 * sin(two_pi * x / L);
        u = 1.0;
      } else {
        const double s = sin(two_pi * x / L);
        const bool pr = (a.kind == IGR_IC_ACOUSTIC_PULSE || a.kind == IGR_IC_ENTROPY_WAVE ||
                         a.kind == IGR_IC_DENSITY_PULSE);
        const bool pu = (a.kind == IGR_IC_ACOUSTIC_PULSE || a.kind == IGR_IC_VELOCITY_PULSE);
        const bool pp = (a.kind == IGR_IC_ACOUSTIC_PULSE || a.kind == IGR_IC_PRESSURE_PULSE);
        if (pr) { rho = 1 + A * s; drho = s; }
        if (pu) { u = A * s; du = s; }
        if (a.kind == IGR_IC_ENTROPY_WAVE) u = 1.0;
        if (pp) { p = 1 + A * s; dp = s; }
      }
      a.q[n] = rho * u;
      a.q[N + n] = rho;
      a.q[2 * N + n] = p / (g - 1) + 0.5 * rho * (u * u);
      if (a.dq) {
        a.dq[n] = rho * du + u * drho;
        a.dq[N + n] = drho;
        a.dq[2 * N + n] = dp / (g - 1) + 0.5 * (drho * (u * u) + 2.0 * rho * u * du);
      }
    }
  }
}

}  // namespace igr

using namespace igr;

// Reference defaults of the builders' keyword arguments.
static const double IC_DEF_CONST2[4] = {1.0, 1.0, 1.0, 1.0};        // rho0, ux0, uy0, p0
static const double IC_DEF_CONST1[3] = {1.0, 1.0, 1.0};             // rho0, u0, p0
static const double IC_DEF_VORTEX[3] = {5.0, 1.0, 1.0};             // beta, ux0, uy0
static const double IC_DEF_SEDOV[2] = {1.0, 0.1};                   // p_bg, sigma
static const double IC_DEF_OSEEN[11] = {0.05, 0.5, 0.25, 0.50, 0.09,   // a_core, M_star, blast_xc, blast_yc, sigma
                                        0.45, 0.60, 1.0, 0.45, 0.40, -1.0};   // (xc, yc, sign) per vortex

extern "C" int igr_build_ic(igr_ctx* c, int ic, double A, double gamma, const double* prm, int n_prm,
                            double* dq_out) {
  if (!c || !(gamma > 1.0)) return IGR_EINVAL;
  if (n_prm < 0 || n_prm > IC_MAX_PRM || (n_prm > 0 && !prm)) return seterr(*c, IGR_EINVAL, "bad IC parameters");
  const bool two = c->dim == 2;
  const double* def = nullptr;
  int n_def = 0;
  switch (ic) {
    case IGR_IC_CONSTANT: def = two ? IC_DEF_CONST2 : IC_DEF_CONST1; n_def = two ? 4 : 3; break;
    case IGR_IC_ISENTROPIC_VORTEX: if (!two) return seterr(*c, IGR_EINVAL, "Unknown IC type (2D only)");
      def = IC_DEF_VORTEX; n_def = 3; break;
    case IGR_IC_TAYLOR_GREEN: if (!two) return seterr(*c, IGR_EINVAL, "Unknown IC type (2D only)"); break;
    case IGR_IC_SEDOV_BLAST: if (!two) return seterr(*c, IGR_EINVAL, "Unknown IC type (2D only)");
      def = IC_DEF_SEDOV; n_def = 2; break;
    case IGR_IC_OSEEN_PAIR_BLAST: if (!two) return seterr(*c, IGR_EINVAL, "Unknown IC type (2D only)");
      def = IC_DEF_OSEEN; n_def = 11; break;
    case IGR_IC_ENTROPY_WAVE_INIT: case IGR_IC_ACOUSTIC_PULSE: case IGR_IC_ENTROPY_WAVE: case IGR_IC_DENSITY_PULSE:
    case IGR_IC_PRESSURE_PULSE: case IGR_IC_VELOCITY_PULSE:
      if (two) return seterr(*c, IGR_EINVAL, "Unknown IC type (1D only)");
      break;
    default: return seterr(*c, IGR_EINVAL, "Unknown IC type: " + std::to_string(ic));
  }
  IcArgs a{};
  a.kind = ic; a.dim = c->dim; a.np = c->np; a.Nex = c->Nex; a.Ney = two ? c->Ney : 1;
  a.Lx = c->Lx; a.Ly = c->Ly;
  a.dx = c->Lx / c->Nex; a.dy = two ? c->Ly / c->Ney : 1.0;
  a.Jx = a.dx / 2.0; a.Jy = a.dy / 2.0;
  for (int k = 0; k < c->np; ++k) a.xi[k] = c->basis.xi[k];
  a.A = A; a.gamma = gamma; a.N = c->N; a.q = c->q;
  // parameters: the caller's, else the reference's defaults
  const double* src = n_prm > 0 ? prm : def;
  a.n_prm = n_prm > 0 ? n_prm : n_def;
  if (ic == IGR_IC_OSEEN_PAIR_BLAST && (a.n_prm < 5 || (a.n_prm - 5) % 3 != 0))
    return seterr(*c, IGR_EINVAL, "Oseen pair blast parameters: a_core, M_star, blast_xc, blast_yc, sigma, then (xc, yc, sign) per vortex");
  if (def && n_prm > 0 && ic != IGR_IC_OSEEN_PAIR_BLAST && n_prm != n_def)
    return seterr(*c, IGR_EINVAL, "expected " + std::to_string(n_def) + " IC parameters");
  for (int k = 0; k < a.n_prm; ++k) a.prm[k] = src[k];
  int rc = begin_call(*c);
  if (rc) return rc;
  if (ic == IGR_IC_OSEEN_PAIR_BLAST) {
    if (!c->ic_tab) {
      c->ic_tab = c->alloc(OSEEN_NTAB);
      if (!c->ic_tab) return seterr(*c, IGR_ENOMEM, "cannot allocate the enthalpy table");
    }
    hipLaunchKernelGGL(k_oseen_table, dim3(1), dim3(64), 0, c->stream, c->ic_tab, gamma, a.prm[0], a.prm[1], c->Lx);
    a.tab = c->ic_tab;
  }
  if (dq_out) {
    a.dq = c->dq;   // the RHS buffer (nv*N) as staging; no call keeps data in it between calls
  }
  const size_t blocks = std::min<size_t>((c->N + 255) / 256, 65536);
  hipLaunchKernelGGL(k_build_ic, dim3((unsigned)blocks), dim3(256), 0, c->stream, a);
  HIPCHK(*c, hipGetLastError());
  if (dq_out)
    HIPCHK(*c, hipMemcpyAsync(dq_out, c->dq, c->nv * c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return end_call(*c);
}
