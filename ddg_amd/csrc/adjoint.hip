// Parameter gradient of the SIPG form: for per-element diffusion coefficient
// kappa_k and penalty scale t_k (tau_k,f = t_k (N+1)^2 Fscale_k,f, the
// per-element tau_scale of src/ddg/dg2d/sipg.py:263),
//
//   g_kappa[k] = d a(u, lam; u_dir) / d kappa_k      g_tau[k] = d a / d t_k
//
// With A(theta) u = b the solve, L(u) a loss and lam the adjoint solution
// A lam = dL/du (A is symmetric, so the forward operator is reused), the
// loss gradient is dL/dtheta = -d a(u, lam)/d theta (b does not depend on
// kappa or tau).  The form is linear in each kappa_k and each t_k, and only
// the faces of element k carry them:
//
//   d/dt_k     = kappa_k sum_f pw_f P_f
//   d/dkappa_k = vol_k / kappa_k + sum_f ( t_k pw_f P_f - sJw_f X_f )
//
//   P_f = [u]^T m1d_f [lam],   X_f = (Dn u)^T m1d_f [lam] + [u]^T m1d_f (Dn lam)
//   vol_k = kappa_k J_k (grad u)^T M (grad lam)  (the volume term of the form)
//   pw_f  = (N+1)^2 sJw_f sJ_f / J_k   (sJ_f = 2 sJw_f interior, sJw_f Dirichlet)
//
// [.] is the jump seen from element k (Dirichlet ghost: u_dir for u, 0 for
// lam), Dn the kappa-free normal derivative of element k.  Neumann faces
// carry neither parameter.  One thread per element; float64 only.

#include "ddgx_internal.h"

namespace ddgx {

template <int N>
__global__ void __launch_bounds__(128)
param_grad_kernel(const double* __restrict__ tab, const double* __restrict__ rec,
                  const int* __restrict__ nbr, const int* __restrict__ perm,
                  const double* __restrict__ u, const double* __restrict__ lam,
                  const double* __restrict__ udir, const double* __restrict__ kappa,
                  const double* __restrict__ tau, double* __restrict__ gk,
                  double* __restrict__ gt, int K) {
  using S = Shape<N>;
  constexpr int Np = S::Np, Nfp = S::Nfp;
  const int i = blockIdx.x * 128 + threadIdx.x;
  if (i >= K) return;
  const int kg = perm[i];
  const double* Dr = tab + S::oDr;
  const double* Ds = tab + S::oDs;
  const double* Mm = tab + S::oM;
  const double* L1 = tab + S::oL1;
  double rc[DDGX_REC_WORDS];
  for (int w = 0; w < DDGX_REC_WORDS; ++w) rc[w] = rec[(size_t)w * K + i];
  const double kap = kappa != nullptr ? kappa[kg] : 1.0;
  const double t = tau[kg];
  const double* uk = u + (size_t)kg * Np;
  const double* lk = lam + (size_t)kg * Np;
  double ur[Np], us[Np], lr[Np], ls[Np];
  for (int n = 0; n < Np; ++n) ur[n] = us[n] = lr[n] = ls[n] = 0.0;
  for (int j = 0; j < Np; ++j) {
    const double uj = uk[j], lj = lk[j];
    for (int n = 0; n < Np; ++n) {
      const double a = Dr[j * S::RS + n], b = Ds[j * S::RS + n];
      ur[n] += a * uj;
      us[n] += b * uj;
      lr[n] += a * lj;
      ls[n] += b * lj;
    }
  }
  // vol = lr^T M^T (vrr ur + vrs us) + ls^T M^T (vrs ur + vss us)
  double vol = 0.0;
  for (int c = 0; c < Np; ++c) {
    const double wr = rc[R_VRR] * ur[c] + rc[R_VRS] * us[c];
    const double ws = rc[R_VRS] * ur[c] + rc[R_VSS] * us[c];
    double mr = 0.0, ms = 0.0;
    for (int n = 0; n < Np; ++n) {
      mr += Mm[n * S::RS + c] * lr[n];  // (M lr)_c with M stored by columns
      ms += Mm[n * S::RS + c] * ls[n];
    }
    vol += mr * wr + ms * ws;
  }
  double g_k = vol / kap, g_t = 0.0;
  constexpr double NN = (double)((N + 1) * (N + 1));
  for (int f = 0; f < 3; ++f) {
    const int code = nbr[(size_t)f * K + i];
    if (code == CODE_NONE) continue;
    const bool inter = code >= 0;
    const int kn = inter ? nbr_elem(code) : 0, fn = inter ? nbr_face(code) : 0;
    const bool rv = inter && nbr_rev(code);
    double ju[Nfp], jl[Nfp], du[Nfp], dl[Nfp];
    for (int p = 0; p < Nfp; ++p) {
      const int n = fmask<N>(f, p);
      double up = 0.0, lp = 0.0;
      if (inter) {
        const int m = fmask<N>(fn, rv ? N - p : p);
        up = u[(size_t)kn * Np + m];
        lp = lam[(size_t)kn * Np + m];
      } else if (udir != nullptr) {
        up = udir[(size_t)kg * Np + n];
      }
      ju[p] = uk[n] - up;
      jl[p] = lk[n] - lp;
      du[p] = (rc[R_A + f] * ur[n] + rc[R_B + f] * us[n]) / kap;
      dl[p] = (rc[R_A + f] * lr[n] + rc[R_B + f] * ls[n]) / kap;
    }
    double P = 0.0, X = 0.0;
    for (int p = 0; p < Nfp; ++p)
      for (int q = 0; q < Nfp; ++q) {
        const double m = L1[(f * Nfp + p) * S::RSF + q];
        P += ju[p] * m * jl[q];
        X += du[p] * m * jl[q] + ju[p] * m * dl[q];
      }
    const double sjw = rc[R_SJW + f];
    const double sj = inter ? 2.0 * sjw : sjw;
    const double pw = NN * sjw * sj / rc[R_J];
    g_k += t * pw * P - sjw * X;
    g_t += kap * pw * P;
  }
  gk[kg] = g_k;
  gt[kg] = g_t;
}

template <int N>
static int grad_t(const Op& op, const double* u, const double* lam, const double* udir,
                  const double* kappa, const double* tau, double* gk, double* gt,
                  hipStream_t s) {
  const int K = (int)op.K;
  hipLaunchKernelGGL((param_grad_kernel<N>), dim3((K + 127) / 128), dim3(128), 0, s,
                     (const double*)op.tab, (const double*)op.rec, op.nbr, op.perm, u, lam,
                     udir, kappa, tau, gk, gt, K);
  DDGX_HIP(hipGetLastError());
  return DDGX_OK;
}

int launch_param_grad(const Op& op, const double* u, const double* lam, const double* udir,
                      const double* kappa, const double* tau, double* gk, double* gt,
                      hipStream_t s) {
  if (op.dtype != DDGX_F64) {
    set_error("parameter gradients need a float64 operator");
    return DDGX_ERR_INVALID;
  }
  switch (op.N) {
    case 1: return grad_t<1>(op, u, lam, udir, kappa, tau, gk, gt, s);
    case 2: return grad_t<2>(op, u, lam, udir, kappa, tau, gk, gt, s);
    case 3: return grad_t<3>(op, u, lam, udir, kappa, tau, gk, gt, s);
    case 4: return grad_t<4>(op, u, lam, udir, kappa, tau, gk, gt, s);
    case 5: return grad_t<5>(op, u, lam, udir, kappa, tau, gk, gt, s);
    case 6: return grad_t<6>(op, u, lam, udir, kappa, tau, gk, gt, s);
  }
  set_error("polynomial order not built");
  return DDGX_ERR_UNSUPPORTED;
}

}  // namespace ddgx
