// Tapered van der Waals + shielded Coulomb over the full directed edge list,
// QEq self energy, and the final energy reduction.
//
// Energies follow src/xreac/energy.py:213-219: per directed edge
//   vdW  = 0.5 * taper * D * (ev^2 - 2 ev) [+ 0.5 * taper * ecore exp(acore (1 - r/rcore))]
//   Coul = 0.5 * 332.06371 * q_i q_j * shield
// One wave owns one atom's row, so the force on atom i is the row sum
// F_i = sum_e phi'(r_e) v_e / r_e (both directions of a pair contribute half
// of phi each), with no scatter to neighbours.
#include "state.h"
#include "profile.h"

namespace {

constexpr double C_ELE = 332.06371;
constexpr double SELF_CONVERSION = 23.02;

// PART selects the terms: NB_VDW (van der Waals, no charges needed: runs
// beside QEq on the second stream and writes its forces into F_hb, which the
// H-bond kernels then accumulate into), NB_COUL (Coulomb and the QEq self
// energy, after QEq, forces into F_nb), or both into F_nb.  With NB_COUL,
// lane 0 also writes (dE/dq_i, -1) when dEdq is given: the right-hand sides of
// the charge-response adjoint solve.
enum NbPart { NB_VDW = 1, NB_COUL = 2, NB_ALL = 3 };

template <int PART>
__global__ void __launch_bounds__(256) k_nonbonded(Atoms at, FFDev ff, const int* __restrict__ row_ptr,
                                                   const int* __restrict__ nbr_j, const int* __restrict__ nbr_s,
                                                   EnergyAcc acc, double2* __restrict__ dEdq) {
  constexpr bool VDW = (PART & NB_VDW) != 0, COUL = (PART & NB_COUL) != 0;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i < at.n) {
    const CellDev& cell = at.S.sys[__builtin_amdgcn_readfirstlane(at.S.asys[i])].cell;
    const double* g = ff.gen;
    const double rc = g[12];
    const double pw = g[28];
    const double inv_pw = 1.0 / pw;
    const int ti = at.type[i];
    const double qi = COUL ? at.q[i] : 0.0;
    const bool shielded = ff.vdw_type == 1 || ff.vdw_type == 3;
    const bool wall = ff.vdw_type == 2 || ff.vdw_type == 3;
    double ev_sum = 0.0, ec_sum = 0.0, qs_sum = 0.0, F0 = 0.0, F1 = 0.0, F2 = 0.0;
    for (int ed = row_ptr[i] + lane; ed < row_ptr[i + 1]; ed += 64) {
      const int j = nbr_j[ed];
      const int tj = at.type[j];
      const int pij = ti * ff.nt + tj;
      const double* p = ff.pair + pij * PP_COUNT;
      double v[3];
      const double r = xr_edge(at.x, cell, i, j, nbr_s[ed], v);
      double t, dt;
      xr_taper(r, rc, t, dt);
      if (t == 0.0 && dt == 0.0) continue;
      double dtot = 0.0;
      if constexpr (VDW) {
        double f13, df13;
        if (shielded) {
          // f13 = (r^p + (1/gamma_w)^p)^(1/p); d f13/dr = r^(p-1) f13 / base
          const double rp = exp(pw * log(r));
          const double base = rp + ff.vdw_b[pij];
          f13 = exp(log(base) * inv_pw);
          df13 = (rp * f13) / (r * base);
        } else {
          f13 = r;
          df13 = 1.0;
        }
        const double* cv = ff.vdw_c + 4 * pij;
        const double ex = exp(0.5 * p[PP_ALPHA] * (1 - f13 * cv[0]));
        const double ev = p[PP_D] * (ex * ex - 2 * ex);
        const double dev = p[PP_D] * (2 * ex - 2) * ex * cv[1] * df13;
        double phi = t * ev;  // per-direction vdW energy (times 2)
        double dphi = dt * ev + t * dev;
        if (wall) {
          const double ecr = p[PP_ECORE] * exp(p[PP_ACORE] * (1 - r * cv[2]));
          phi += t * ecr;
          dphi += dt * ecr + t * ecr * cv[3];
        }
        ev_sum += phi;
        dtot += dphi;
      }
      if constexpr (COUL) {
        double s, ds;
        xr_shield_tapered(r, t, dt, p[PP_GAMMA], s, ds);
        const double qj = at.q[j];
        const double qq = qi * qj;
        ec_sum += qq * s;
        qs_sum += qj * s;
        dtot += C_ELE * qq * ds;
      }
      dtot /= r;
      F0 += dtot * v[0];
      F1 += dtot * v[1];
      F2 += dtot * v[2];
    }
    if constexpr (VDW) ev_sum = xr_wave_sum(ev_sum);
    if constexpr (COUL) {
      ec_sum = xr_wave_sum(ec_sum);
      if (dEdq) qs_sum = xr_wave_sum(qs_sum);
    }
    F0 = xr_wave_sum(F0);
    F1 = xr_wave_sum(F1);
    F2 = xr_wave_sum(F2);
    if (lane == 0) {
      const double* a = ff.atom + ti * AP_COUNT;
      if constexpr (VDW) xr_energy(acc, CH_VDW, i, 0.5 * ev_sum);
      if constexpr (COUL) {
        xr_energy(acc, CH_COUL, i, 0.5 * C_ELE * ec_sum);
        xr_energy(acc, CH_QEQ, i, SELF_CONVERSION * (a[AP_CHI] * qi + 0.5 * a[AP_ETA] * qi * qi));
        // dE/dq_i = 332.06371 sum_e q_j s_e + 23.02 (chi_i + eta_i q_i)
        if (dEdq) dEdq[i] = make_double2(C_ELE * qs_sum + SELF_CONVERSION * (a[AP_CHI] + a[AP_ETA] * qi), -1.0);
      }
      double* F = PART == NB_VDW ? at.F_hb : at.F_nb;
      F[3 * i] = F0;
      F[3 * i + 1] = F1;
      F[3 * i + 2] = F2;
    }
  }
}

// Charge-response force (full derivative through QEq).  With lambda the
// neutral adjoint solution of [[H, 1], [1^T, 0]] [lambda; nu] = [dE/dq; 0],
// dE/dx = dE/dx|_q - lambda^T (dH/dx) q, so every directed edge adds
// -14.4 ds/dr (lambda_i q_j + lambda_j q_i) v / r to the row atom's force
// (self-image edges cancel pairwise).  The reference obtains the same term by
// differentiating through np.linalg.solve (src/xreac/energy.py:97-102).
__global__ void __launch_bounds__(256) k_charge_response(Atoms at, FFDev ff, const int* __restrict__ row_ptr,
                                                         const int* __restrict__ nbr_j,
                                                         const int* __restrict__ nbr_s,
                                                         const double* __restrict__ lam, double* __restrict__ F) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= at.n) return;
  const CellDev& cell = at.S.sys[__builtin_amdgcn_readfirstlane(at.S.asys[i])].cell;
  const double rc = ff.gen[12];
  const int ti = at.type[i];
  const double li = lam[i], qi = at.q[i];
  double F0 = 0.0, F1 = 0.0, F2 = 0.0;
  for (int ed = row_ptr[i] + lane; ed < row_ptr[i + 1]; ed += 64) {
    const int j = nbr_j[ed];
    const int tj = at.type[j];
    double v[3];
    const double r = xr_edge(at.x, cell, i, j, nbr_s[ed], v);
    double s, ds;
    xr_shield(r, rc, ff.pair[(ti * ff.nt + tj) * PP_COUNT + PP_GAMMA], s, ds);
    const double c = 14.4 * ds * (li * at.q[j] + lam[j] * qi) / r;
    F0 += c * v[0];
    F1 += c * v[1];
    F2 += c * v[2];
  }
  F0 = xr_wave_sum(F0);
  F1 = xr_wave_sum(F1);
  F2 = xr_wave_sum(F2);
  if (lane == 0) {
    F[3 * i] -= F0;
    F[3 * i + 1] -= F1;
    F[3 * i + 2] -= F2;
  }
}

// One block per (system, channel): the channel summed over the system's atoms
// in a fixed order (lane-strided, then a fixed tree), independent of where
// the system sits in the batch.
__global__ void __launch_bounds__(1024) k_energy_reduce(Systems S, EnergyAcc acc, double* __restrict__ chan) {
  const SysDev& sd = S.sys[blockIdx.x];
  const int c = blockIdx.y;
  __shared__ double red[16];
  const double* e = acc.e + (size_t)c * acc.n + sd.atom0;
  // four independent partial sums per thread, so the loads are in flight together
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int t = threadIdx.x;
  for (; t + 3 * 1024 < sd.n; t += 4 * 1024) {
    a0 += e[t];
    a1 += e[t + 1024];
    a2 += e[t + 2048];
    a3 += e[t + 3072];
  }
  for (; t < sd.n; t += 1024) a0 += e[t];
  double s = xr_wave_sum((a0 + a1) + (a2 + a3));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0;
    for (int w = 0; w < 16; ++w) x += red[w];
    chan[(size_t)blockIdx.x * NCH + c] = x;
  }
}

}  // namespace

hipError_t xr_vdw(const Atoms& at, FFDev ff, const int* row_ptr, const int* nbr_j, const int* nbr_s, EnergyAcc acc,
                  hipStream_t s) {
  const dim3 grid((at.n + 3) / 4 > 0 ? (at.n + 3) / 4 : 1);
  hipLaunchKernelGGL(k_nonbonded<NB_VDW>, grid, dim3(256), 0, s, at, ff, row_ptr, nbr_j, nbr_s, acc, nullptr);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_nonbonded(const Atoms& at, FFDev ff, const int* row_ptr, const int* nbr_j, const int* nbr_s,
                        EnergyAcc acc, double2* dEdq, hipStream_t s) {
  const dim3 grid((at.n + 3) / 4 > 0 ? (at.n + 3) / 4 : 1);
  XR_LAUNCH(XT_NONBONDED, k_nonbonded<NB_COUL>, grid, dim3(256), 0, s, at, ff, row_ptr, nbr_j, nbr_s, acc, dEdq);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_charge_response(const Atoms& at, FFDev ff, const int* row_ptr, const int* nbr_j, const int* nbr_s,
                              const double* lam, double* forces, hipStream_t s) {
  const dim3 grid((at.n + 3) / 4 > 0 ? (at.n + 3) / 4 : 1);
  hipLaunchKernelGGL(k_charge_response, grid, dim3(256), 0, s, at, ff, row_ptr, nbr_j, nbr_s, lam, forces);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_energy_reduce(const Systems& S, EnergyAcc acc, double* chan, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_reduce, dim3(S.ns, NCH), dim3(1024), 0, s, S, acc, chan);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}
