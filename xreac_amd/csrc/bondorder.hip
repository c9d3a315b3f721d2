// Bond orders, bond / lone-pair / over-under energies and the bond-order adjoint.
//
// Forward equations restate src/xreac/energy.py:105-212 and 231-246
// (reference line numbers).  Every per-atom sum is taken by the atom's own
// thread over its bond row in canonical order; contributions an atom makes to
// a neighbor's adjoint are left in a per-bond slot and gathered by the
// neighbor through the reverse-bond map, so no floating-point atomics occur.
#include "state.h"

namespace {

constexpr double THB_CUT = 0.001;

__device__ inline const double* AP(const FFDev& ff, int t) { return ff.atom + t * AP_COUNT; }
__device__ inline const double* PPm(const FFDev& ff, int ti, int tj) { return ff.pair + (ti * ff.nt + tj) * PP_COUNT; }

// ------------------------------------------------------------ bond geometry
__global__ void k_bond_geom(Atoms at, Bonds bd, FFDev ff, const int* __restrict__ nbr_j,
                            const int* __restrict__ nbr_s) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= xr_bond_count(bd)) return;
  const int e = bd.edge[b], i = bd.bi[b];
  const int j = nbr_j[e], code = nbr_s[e];
  double v[3];
  const double r = xr_edge(at.x, xr_cell(at.S, i), i, j, code, v);
  const int ti = at.type[i], tj = at.type[j];
  BondRaw raw;
  xr_bond_raw(ff, ti, tj, r, raw);
  const double cut = 0.01 * ff.gen[29];
  bd.bj[b] = j;
  bd.code[b] = code;
  bd.vec[3 * b] = v[0];
  bd.vec[3 * b + 1] = v[1];
  bd.vec[3 * b + 2] = v[2];
  bd.r[b] = r;
  bd.bop[b] = ((raw.sig + raw.pi) + raw.pp) - cut;
  bd.pim[b] = raw.pi;
  bd.ppm[b] = raw.pp;
  bd.dsig[b] = raw.dsig;
  bd.dpi[b] = raw.dpi;
  bd.dpp[b] = raw.dpp;
  const double* p = PPm(ff, ti, tj);
  if (p[PP_OVC] >= 0.001 || p[PP_V13] >= 0.001) {
    unsigned bad = 0;
    if (!isfinite(p[PP_BOC3])) bad |= XR_ERR_BOC3;
    if (!isfinite(p[PP_BOC4])) bad |= XR_ERR_BOC4;
    if (!isfinite(p[PP_BOC5])) bad |= XR_ERR_BOC5;
    if (bad) xr_flag(at.S, i, bad);
  }
}

__global__ void k_bond_reverse(Bonds bd, Systems S) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= xr_bond_count(bd)) return;
  const int i = bd.bi[b], j = bd.bj[b], want = xr_neg_shift(bd.code[b]);
  int found = -1;
  for (int s = bd.ptr[j]; s < bd.ptr[j + 1]; ++s) {
    if (bd.bj[s] == i && bd.code[s] == want) {
      found = s;
      break;
    }
  }
  if (found < 0) {
    xr_flag(S, i, XR_ERR_REVERSE);
    found = b;
  }
  bd.rev[b] = found;
}

__global__ void k_atom_totalp(Atoms at, Bonds bd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= at.n) return;
  double t = 0.0;
  for (int b = bd.ptr[i]; b < bd.ptr[i + 1]; ++b) t += bd.bop[b];
  at.totp[i] = t;
}

// Corrected bond orders (energy.py:150-165).
__global__ void k_bond_correct(Atoms at, Bonds bd, FFDev ff) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= xr_bond_count(bd)) return;
  const double* g = ff.gen;
  const int i = bd.bi[b], j = bd.bj[b];
  const int ti = at.type[i], tj = at.type[j];
  const double* ai = AP(ff, ti);
  const double* aj = AP(ff, tj);
  const double* p = PPm(ff, ti, tj);
  const double bo0 = bd.bop[b];
  const double di = at.totp[i] - ai[AP_VAL], dj = at.totp[j] - aj[AP_VAL];
  const double dbi = at.totp[i] - ai[AP_VALVAL], dbj = at.totp[j] - aj[AP_VALVAL];
  double f1 = 1.0;
  if (p[PP_OVC] >= 0.001) {
    const double f2 = exp(-g[0] * di) + exp(-g[0] * dj);
    const double f3 = -log(0.5 * (exp(-g[1] * di) + exp(-g[1] * dj))) / g[1];
    const double vi = ai[AP_VAL], vj = aj[AP_VAL];
    f1 = 0.5 * ((vi + f2) / (vi + f2 + f3) + (vj + f2) / (vj + f2 + f3));
  }
  double f4 = 1.0, f5 = 1.0, f45 = 1.0;
  if (p[PP_V13] >= 0.001) {
    const double b3 = p[PP_BOC3], b4 = p[PP_BOC4], b5 = p[PP_BOC5];
    f4 = 1 / (1 + exp(-(b4 * (bo0 * bo0) - dbi) * b3 + b5));
    f5 = 1 / (1 + exp(-(b4 * (bo0 * bo0) - dbj) * b3 + b5));
    f45 = f4 * f5;
  }
  const double corr = f1 * f45;
  const double bo = bo0 * corr;
  const double pi = bd.pim[b] * corr * f1;
  const double pp = bd.ppm[b] * corr * f1;
  const double sig = (bo - pi) - pp;
  bd.f1[b] = f1;
  bd.f4[b] = f4;
  bd.f5[b] = f5;
  bd.boc[b] = bo;
  bd.pic[b] = pi;
  bd.ppc[b] = pp;
  bd.BO[b] = bo >= 1e-10 ? bo : 0.0;
  bd.SIG[b] = sig >= 1e-10 ? sig : 0.0;
  bd.PI[b] = pi >= 1e-10 ? pi : 0.0;
  bd.PP[b] = pp >= 1e-10 ? pp : 0.0;
}

// Atom quantities after correction (energy.py:185-199, 295-301) and lone pairs.
__global__ void __launch_bounds__(256) k_atom_forward(Atoms at, Bonds bd, FFDev ff, EnergyAcc acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double elp[1] = {0.0};
  if (i < at.n) {
    const double* g = ff.gen;
    const double* a = AP(ff, at.type[i]);
    double total = 0.0, pipp = 0.0, bo8 = 0.0;
    int cnt = 0;
    for (int b = bd.ptr[i]; b < bd.ptr[i + 1]; ++b) {
      const double B = bd.BO[b];
      total += B;
      pipp += bd.PI[b] + bd.PP[b];
      bo8 += pow(B, 8.0);
      cnt += B > 0.3;
    }
    const double d = total - a[AP_VAL], db = total - a[AP_VALBOC], de = total - a[AP_VALE];
    const double half = trunc(de / 2);
    const double vlpex = de - 2 * half;
    const double t2 = 2 + vlpex;
    const double nlp = exp(-g[15] * (t2 * t2)) - half;
    const double dlp = 0.5 * (a[AP_VALE] - a[AP_VAL]) - nlp;
    const double dlpt = a[AP_MASS] > 21 ? 0.0 : dlp;
    const double ex = exp(-75 * dlp);
    elp[0] = a[AP_PLP2] * dlp / (1 + ex);
    const double a_dlp = a[AP_PLP2] * (1 / (1 + ex) + dlp * 75 * ex / ((1 + ex) * (1 + ex)));
    const double prod = exp(-bo8);
    const double X = vlpex >= 0 ? 0.0 : nlp;
    const double sbo = pipp + (1 - prod) * (-db - g[33] * X);
    double sbo2;
    if (sbo <= 0) sbo2 = 0.0;
    else if (sbo <= 1) sbo2 = pow(sbo, g[16]);
    else if (sbo < 2) sbo2 = 2 - pow(2 - sbo, g[16]);
    else sbo2 = 2.0;
    at.total[i] = total;
    at.d[i] = d;
    at.db[i] = db;
    at.nlp[i] = nlp;
    at.vlpex[i] = vlpex;
    at.dlp[i] = dlp;
    at.dlpt[i] = dlpt;
    at.sbo[i] = sbo;
    at.sbo2[i] = sbo2;
    at.prod[i] = prod;
    at.bond_count[i] = cnt;
    at.A_dlp[i] = a_dlp;
  }
  if (i < at.n) xr_energy(acc, CH_LP, i, elp[0]);
}

// Over/under-coordination, bond energy and the C2 correction; initialises
// the adjoint accumulators (energy.py:194-212, 233-237).
__global__ void __launch_bounds__(256) k_atom_forward2(Atoms at, Bonds bd, FFDev ff, EnergyAcc acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e[3] = {0.0, 0.0, 0.0};  // bond, atom, C2
  if (i < at.n) {
    const double* g = ff.gen;
    const int ti = at.type[i];
    const double* a = AP(ff, ti);
    const double dfvl = a[AP_MASS] <= 21 ? 1.0 : 0.0;
    double ov1 = 0.0, ov2 = 0.0;
    for (int b = bd.ptr[i]; b < bd.ptr[i + 1]; ++b) {
      const int j = bd.bj[b];
      const double* p = PPm(ff, ti, at.type[j]);
      ov1 += p[PP_OVUN1] * p[PP_DES] * bd.BO[b];
      ov2 += (at.d[j] - dfvl * at.dlpt[j]) * (bd.PI[b] + bd.PP[b]);
    }
    const double d = at.d[i], dlpt = at.dlpt[i], val = a[AP_VAL];
    const double e31 = exp(g[31] * ov2);
    const double Dd = 1 + g[32] * e31;
    const double dc = d - dfvl * dlpt / Dd;
    const double p2 = a[AP_POVUN2], p5 = a[AP_POVUN5];
    const double den = dc + val + 1e-8;
    const double e2 = exp(p2 * dc);
    const double over = ov1 * dc / den / (1 + e2);
    const double e6 = exp(g[6] * dc), em2 = exp(-p2 * dc), e9 = exp(g[9] * ov2);
    const double Du = 1 + g[8] * e9;
    const double under = -p5 * (1 - e6) / (1 + em2) / Du;
    e[1] = over + under;
    const double dO_dov1 = dc / den / (1 + e2);
    const double dO_ddc = ov1 * ((val + 1e-8) / (den * den) / (1 + e2) - (dc / den) * p2 * e2 / ((1 + e2) * (1 + e2)));
    const double dU_ddc = -p5 * (-g[6] * e6 / (1 + em2) + (1 - e6) * p2 * em2 / ((1 + em2) * (1 + em2))) / Du;
    const double dU_dov2 = p5 * (1 - e6) / (1 + em2) * g[8] * g[9] * e9 / (Du * Du);
    const double a_dc = dO_ddc + dU_ddc;
    const double a_ov2 = dU_dov2 + a_dc * dfvl * dlpt * g[32] * g[31] * e31 / (Dd * Dd);
    double A_d = a_dc;
    const double A_dlpt = a_dc * (-dfvl / Dd);
    const bool c2 = g[5] > 0.001 && a[AP_CARBON] != 0.0;
    for (int b = bd.ptr[i]; b < bd.ptr[i + 1]; ++b) {
      const int j = bd.bj[b];
      const int tj = at.type[j];
      const double* p = PPm(ff, ti, tj);
      const double SIG = bd.SIG[b];
      const double pw = SIG > 0 ? pow(SIG, p[PP_BE2]) : 0.0;
      const double ex = exp(p[PP_BE1] * (1 - pw));
      e[0] += p[PP_DES] * SIG * ex + p[PP_DEP] * bd.PI[b] + p[PP_DEPP] * bd.PP[b];
      bd.A_SIG[b] = -0.5 * p[PP_DES] * ex * (1 - (SIG > 0 ? p[PP_BE1] * p[PP_BE2] * pw : 0.0));
      const double w = at.d[j] - dfvl * at.dlpt[j];
      bd.A_PI[b] = -0.5 * p[PP_DEP] + a_ov2 * w;
      bd.A_PP[b] = -0.5 * p[PP_DEPP] + a_ov2 * w;
      double A_BO = dO_dov1 * (p[PP_OVUN1] * p[PP_DES]);
      if (c2 && AP(ff, tj)[AP_CARBON] != 0.0) {
        const double vov3 = (bd.BO[b] - d) - 0.040 * pow(d, 4.0);
        if (vov3 > 3) {
          e[2] += g[5] * (vov3 - 3) * (vov3 - 3);
          A_BO += 2 * g[5] * (vov3 - 3);
          A_d += 2 * g[5] * (vov3 - 3) * (-1 - 0.16 * d * d * d);
        }
      }
      bd.A_BO[b] = A_BO;
      bd.A_totfar[b] = 0.0;
      bd.G[3 * b] = bd.G[3 * b + 1] = bd.G[3 * b + 2] = 0.0;
    }
    e[0] *= -0.5;
    at.ov2[i] = ov2;
    at.a_ov2[i] = a_ov2;
    at.A_d[i] = A_d;
    at.A_dlpt[i] = A_dlpt;
    at.A_total[i] = 0.0;
    at.A_db[i] = 0.0;
    at.A_nlp[i] = 0.0;
    at.A_sbo2[i] = 0.0;
  }
  if (i < at.n) {
    xr_energy(acc, CH_BOND, i, e[0]);
    xr_energy(acc, CH_ATOM, i, e[1]);
    xr_energy(acc, CH_C2, i, e[2]);
  }
}

// Triple-bond stabilisation on C-O half edges (energy.py:238-246).  Each end
// of a qualifying pair adds the derivative with respect to its own d and total.
__global__ void __launch_bounds__(256) k_triple(Atoms at, Bonds bd, FFDev ff, EnergyAcc acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e[1] = {0.0};
  if (i < at.n) {
    const double* g = ff.gen;
    const double* ai = AP(ff, at.type[i]);
    const bool all = (int)g[37] == 2;
    for (int b = bd.ptr[i]; b < bd.ptr[i + 1]; ++b) {
      const double B = bd.BO[b];
      if (B < 1.0) continue;
      const int j = bd.bj[b];
      const double* aj = AP(ff, at.type[j]);
      const bool co = (ai[AP_MASS_C12] != 0.0 && aj[AP_MASS_O16] != 0.0) ||
                      (aj[AP_MASS_C12] != 0.0 && ai[AP_MASS_O16] != 0.0);
      if (!(all || co)) continue;
      const bool half = bd.edge[b] < bd.edge[bd.rev[b]];
      const double A = g[10] * exp(-g[7] * ((B - 2.5) * (B - 2.5)));
      const double Ex = exp(g[4] * (at.d[i] + at.d[j]));
      const double Dn = 1 + 25 * Ex;
      const double ei = exp(-g[3] * (at.total[i] - B));
      const double ej = exp(-g[3] * (at.total[j] - B));
      const double T = ei + ej;
      if (half) {
        e[0] += A / Dn * T;
        const double dA = -2 * g[7] * (B - 2.5) * A;
        bd.A_BO[b] += (dA + A * g[3]) * T / Dn;
      }
      at.A_d[i] += -A * T * 25 * g[4] * Ex / (Dn * Dn);
      at.A_total[i] += (A / Dn) * (-g[3] * ei);
    }
  }
  if (i < at.n) xr_energy(acc, CH_TRIPLE, i, e[0]);
}

// ----------------------------------------------------------------- adjoint
// Gather the per-atom adjoints and push d E / d total onto the row's bonds.
__global__ void k_atom_backward(Atoms at, Bonds bd, FFDev ff) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= at.n) return;
  const double* g = ff.gen;
  const double* a = AP(ff, at.type[i]);
  double a_d = at.A_d[i], a_dlpt = at.A_dlpt[i], a_total = at.A_total[i];
  for (int b = bd.ptr[i]; b < bd.ptr[i + 1]; ++b) {
    const int j = bd.bj[b];
    const double w = at.a_ov2[j] * (bd.PI[b] + bd.PP[b]);
    const double dfvl_j = AP(ff, at.type[j])[AP_MASS] <= 21 ? 1.0 : 0.0;
    a_d += w;
    a_dlpt -= w * dfvl_j;
    a_total += bd.A_totfar[bd.rev[b]];
  }
  const double a_dlp = at.A_dlp[i] + (a[AP_MASS] <= 21 ? a_dlpt : 0.0);
  const double sbo = at.sbo[i], prod = at.prod[i], vlpex = at.vlpex[i], nlp = at.nlp[i], db = at.db[i];
  double dsbo2 = 0.0;
  if (sbo > 0 && sbo <= 1) dsbo2 = g[16] * pow(sbo, g[16] - 1);
  else if (sbo > 1 && sbo < 2) dsbo2 = g[16] * pow(2 - sbo, g[16] - 1);
  const double a_sbo = at.A_sbo2[i] * dsbo2;
  const double X = vlpex >= 0 ? 0.0 : nlp;
  const double W = -db - g[33] * X;
  const double a_prod = -a_sbo * W;
  const double a_bo8 = -prod * a_prod;
  const double a_pipp = a_sbo;
  const double a_db = at.A_db[i] - a_sbo * (1 - prod);
  double a_nlp = at.A_nlp[i] - a_dlp;
  if (vlpex < 0) a_nlp += -a_sbo * (1 - prod) * g[33];
  const double t2 = 2 + vlpex;
  const double a_de = a_nlp * (-2 * g[15] * t2 * exp(-g[15] * (t2 * t2)));
  const double aT = a_total + a_d + a_db + a_de;
  for (int b = bd.ptr[i]; b < bd.ptr[i + 1]; ++b) {
    bd.A_BO[b] += aT + a_bo8 * 8 * pow(bd.BO[b], 7.0);
    bd.A_PI[b] += a_pipp;
    bd.A_PP[b] += a_pipp;
  }
}

// Through thresholds and corrections to the uncorrected quantities.
__global__ void k_bond_backward1(Atoms at, Bonds bd, FFDev ff) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= xr_bond_count(bd)) return;
  const double* g = ff.gen;
  const int i = bd.bi[b], j = bd.bj[b];
  const int ti = at.type[i], tj = at.type[j];
  const double* p = PPm(ff, ti, tj);
  const double bo = bd.boc[b], pi = bd.pic[b], pp = bd.ppc[b];
  const double sig = (bo - pi) - pp;
  const double aS = sig >= 1e-10 ? bd.A_SIG[b] : 0.0;
  const double a_bo = (bo >= 1e-10 ? bd.A_BO[b] : 0.0) + aS;
  const double a_pi = (pi >= 1e-10 ? bd.A_PI[b] : 0.0) - aS;
  const double a_pp = (pp >= 1e-10 ? bd.A_PP[b] : 0.0) - aS;
  const double bo0 = bd.bop[b], pim = bd.pim[b], ppm = bd.ppm[b];
  const double f1 = bd.f1[b], f4 = bd.f4[b], f5 = bd.f5[b];
  const bool has13 = p[PP_V13] >= 0.001;
  const double F45 = has13 ? f4 * f5 : 1.0;
  const double corr = f1 * F45;
  double a_bop = a_bo * corr;
  const double a_pim = a_pi * corr * f1;
  const double a_ppm = a_pp * corr * f1;
  const double a_corr = a_bo * bo0 + (a_pi * pim + a_pp * ppm) * f1;
  const double a_f1 = (a_pi * pim + a_pp * ppm) * corr + a_corr * F45;
  const double a_F45 = a_corr * f1;
  double ci = 0.0, cj = 0.0;
  if (p[PP_OVC] >= 0.001) {
    const double* ai = AP(ff, ti);
    const double* aj = AP(ff, tj);
    const double di = at.totp[i] - ai[AP_VAL], dj = at.totp[j] - aj[AP_VAL];
    const double e0i = exp(-g[0] * di), e0j = exp(-g[0] * dj);
    const double e1i = exp(-g[1] * di), e1j = exp(-g[1] * dj);
    const double f2 = e0i + e0j;
    const double f3 = -log(0.5 * (e1i + e1j)) / g[1];
    const double vi = ai[AP_VAL], vj = aj[AP_VAL];
    const double Si = vi + f2 + f3, Sj = vj + f2 + f3;
    const double df1_df2 = 0.5 * (f3 / (Si * Si) + f3 / (Sj * Sj));
    const double df1_df3 = -0.5 * ((vi + f2) / (Si * Si) + (vj + f2) / (Sj * Sj));
    const double s1 = e1i + e1j;
    ci += a_f1 * (df1_df2 * (-g[0] * e0i) + df1_df3 * (e1i / s1));
    cj += a_f1 * (df1_df2 * (-g[0] * e0j) + df1_df3 * (e1j / s1));
  }
  if (has13) {
    const double b3 = p[PP_BOC3], b4 = p[PP_BOC4];
    const double a_z4 = -(a_F45 * f5) * f4 * (1 - f4);
    const double a_z5 = -(a_F45 * f4) * f5 * (1 - f5);
    a_bop += (a_z4 + a_z5) * (-2 * b4 * bo0 * b3);
    ci += a_z4 * b3;
    cj += a_z5 * b3;
  }
  bd.a_bop[b] = a_bop;
  bd.a_pim[b] = a_pim;
  bd.a_ppm[b] = a_ppm;
  bd.c_i[b] = ci;
  bd.c_j[b] = cj;
}

__global__ void k_atom_backward2(Atoms at, Bonds bd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= at.n) return;
  double s = 0.0;
  for (int b = bd.ptr[i]; b < bd.ptr[i + 1]; ++b) s += bd.c_i[b] + bd.c_j[bd.rev[b]];
  at.a_totp[i] = s;
}

__global__ void k_bond_backward2(Atoms at, Bonds bd) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= xr_bond_count(bd)) return;
  const double a_bop = bd.a_bop[b] + at.a_totp[bd.bi[b]];
  const double gr = a_bop * bd.dsig[b] + (a_bop + bd.a_pim[b]) * bd.dpi[b] + (a_bop + bd.a_ppm[b]) * bd.dpp[b];
  const double s = gr / bd.r[b];
  bd.G[3 * b] += s * bd.vec[3 * b];
  bd.G[3 * b + 1] += s * bd.vec[3 * b + 1];
  bd.G[3 * b + 2] += s * bd.vec[3 * b + 2];
}

// F_i = F_nonbonded + F_hbond-acceptor + sum_b (G_b - G_rev(b)).
__global__ void k_force(Atoms at, Bonds bd, double* __restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= at.n) return;
  double f0 = at.F_nb[3 * i] + at.F_hb[3 * i];
  double f1 = at.F_nb[3 * i + 1] + at.F_hb[3 * i + 1];
  double f2 = at.F_nb[3 * i + 2] + at.F_hb[3 * i + 2];
  for (int b = bd.ptr[i]; b < bd.ptr[i + 1]; ++b) {
    const int rb = bd.rev[b];
    f0 += bd.G[3 * b] - bd.G[3 * rb];
    f1 += bd.G[3 * b + 1] - bd.G[3 * rb + 1];
    f2 += bd.G[3 * b + 2] - bd.G[3 * rb + 2];
  }
  F[3 * i] = f0;
  F[3 * i + 1] = f1;
  F[3 * i + 2] = f2;
}

// One wave per block: per-atom and per-bond kernels have a few thousand
// threads, and small blocks spread them over every CU.
constexpr int BLK = 64;
inline dim3 grid(int count) { return dim3(count > 0 ? (count + BLK - 1) / BLK : 1); }

}  // namespace

hipError_t xr_bond_forward(const Atoms& at, const Bonds& bd, FFDev ff, const int* nbr_j, const int* nbr_s,
                           EnergyAcc acc, hipStream_t s) {
  if (bd.cap > 0) {
    hipLaunchKernelGGL(k_bond_geom, grid(bd.cap), dim3(BLK), 0, s, at, bd, ff, nbr_j, nbr_s);
    hipLaunchKernelGGL(k_bond_reverse, grid(bd.cap), dim3(BLK), 0, s, bd, at.S);
  }
  hipLaunchKernelGGL(k_atom_totalp, grid(at.n), dim3(BLK), 0, s, at, bd);
  if (bd.cap > 0) hipLaunchKernelGGL(k_bond_correct, grid(bd.cap), dim3(BLK), 0, s, at, bd, ff);
  hipLaunchKernelGGL(k_atom_forward, grid(at.n), dim3(BLK), 0, s, at, bd, ff, acc);
  hipLaunchKernelGGL(k_atom_forward2, grid(at.n), dim3(BLK), 0, s, at, bd, ff, acc);
  hipLaunchKernelGGL(k_triple, grid(at.n), dim3(BLK), 0, s, at, bd, ff, acc);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_bond_backward(const Atoms& at, const Bonds& bd, FFDev ff, hipStream_t s) {
  hipLaunchKernelGGL(k_atom_backward, grid(at.n), dim3(BLK), 0, s, at, bd, ff);
  if (bd.cap > 0) hipLaunchKernelGGL(k_bond_backward1, grid(bd.cap), dim3(BLK), 0, s, at, bd, ff);
  hipLaunchKernelGGL(k_atom_backward2, grid(at.n), dim3(BLK), 0, s, at, bd);
  if (bd.cap > 0) hipLaunchKernelGGL(k_bond_backward2, grid(bd.cap), dim3(BLK), 0, s, at, bd);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_force_assembly(const Atoms& at, const Bonds& bd, double* forces, hipStream_t s) {
  hipLaunchKernelGGL(k_force, grid(at.n), dim3(BLK), 0, s, at, bd, forces);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}
