// Valence-angle, torsion and hydrogen-bond terms with their adjoints.
//
// Items are enumerated on the device from the bond rows, in the reference's
// selection rules (src/xreac/energy.py:220, 248-370):
//   angles    - centre j, unordered pairs of bonds with BO > 0.001 whose
//               product exceeds 1e-5, every parameter row with |v1| > 0.001;
//   torsions  - half bond j->k with BO > 0.001, left bond at j, right bond at
//               k, excluding the middle edge, its reverse, 3-rings through the
//               same image and products <= 0.001;
//   H-bonds   - hydrogen j, donor bond to a heavy atom with BO >= 0.01 and any
//               heavy-atom edge within min(7.5, rc) other than that same edge.
// An item's derivative is deposited only into slots owned by the thread that
// enumerates it (its own bond row and atom).  Angles are enumerated from both
// of their bonds, torsions from both ends of the middle bond and H-bonds from
// both the hydrogen and the acceptor, so each end writes its own share
// without atomics.
#include "state.h"

namespace {

constexpr double THB_CUT = 0.001;
constexpr double THB_PRODUCT_CUT = 0.00001;
constexpr double DEG = 3.14159265 / 180;

__device__ inline const double* AP(const FFDev& ff, int t) { return ff.atom + t * AP_COUNT; }
__device__ inline double dot3(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
__device__ inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// ------------------------------------------------------------------ angles
// Factors of an angle term that depend on the centre j alone.
struct AngleCentre {
  double db, d, sbo2, esb, ex6, Q, dQ, Cj, dCj, p3, p5;
  int tj;
};

__device__ inline AngleCentre angle_centre(const Atoms& at, const FFDev& ff, int j) {
  const double* g = ff.gen;
  AngleCentre c;
  c.tj = at.type[j];
  const double* aj = AP(ff, c.tj);
  c.db = at.db[j];
  c.d = at.d[j];
  c.sbo2 = at.sbo2[j];
  c.p3 = aj[AP_PVAL3];
  c.p5 = aj[AP_PVAL5];
  const double ex3 = exp(-g[20] * c.d), ex4 = exp(g[21] * c.d);
  c.Q = (2 + ex3) / (1 + ex3 + ex4);
  c.dQ = (-g[20] * ex3 * (1 + ex3 + ex4) - (2 + ex3) * (-g[20] * ex3 + g[21] * ex4)) /
         ((1 + ex3 + ex4) * (1 + ex3 + ex4));
  const double ecj = exp(g[2] * (at.total[j] - aj[AP_VALVAL]));
  c.Cj = 1 / (1 + ecj);
  c.dCj = -g[2] * ecj * c.Cj * c.Cj;
  c.esb = exp(-g[17] * (2 - c.sbo2));
  c.ex6 = exp(g[14] * c.db);
  return c;
}

// One angle (a, j, c) of bonds a < c of the centre, summed over its parameter
// rows: the shares of the two bonds' slots (A_BO, A_totfar, bond-vector
// derivative) and of the centre (energies and adjoints).
struct AnglePair {
  double gA, gC, ta, tc;
  double GA[3], GC[3];
  double e[3];
  double A_db, A_d, A_tot, A_sbo2;
};

// False when the reference's enumeration (energy.py:248-277) skips the pair:
// a bond order at or below 0.001, a product at or below 1e-5, or no row.
__device__ inline bool angle_pair(const Atoms& at, const Bonds& bd, const FFDev& ff, int lammps,
                                  const AngleCentre& cj, int a, int c, AnglePair& o) {
  const double Ba = bd.BO[a], Bc = bd.BO[c];
  if (!(Ba > THB_CUT) || !(Bc > THB_CUT) || Ba * Bc <= THB_PRODUCT_CUT) return false;
  const int nt = ff.nt;
  const int ia = bd.bj[a], kc = bd.bj[c];
  const int key = (at.type[ia] * nt + cj.tj) * nt + at.type[kc];
  const int r0 = ff.ang_off[key], r1 = ff.ang_off[key + 1];
  if (r0 == r1) return false;
  const double* g = ff.gen;
  const double* va = bd.vec + 3 * a;
  const double* vc = bd.vec + 3 * c;
  const double ra = bd.r[a], rc = bd.r[c];
  const double cosv = dot3(va, vc) / (ra * rc);
  double theta, dth;
  if (lammps) {
    const double cc = fmin(fmax(cosv, -1.0), 1.0);
    theta = acos(cc);
    dth = -1.0 / fmax(sin(theta), 1e-5);
  } else {
    const double lo = -1 + 1e-14, hi = 1 - 1e-14;
    const double cc = fmin(fmax(cosv, lo), hi);
    theta = acos(cc);
    dth = (cc == lo || cc == hi) ? 0.0 : -1.0 / sqrt(1 - cc * cc);
  }
  const double bij = Ba - THB_CUT, bjk = Bc - THB_CUT;
  const double tota = at.total[ia], totc = at.total[kc];
  const double p3 = cj.p3, p5 = cj.p5, db = cj.db, esb = cj.esb, ex6 = cj.ex6;
  o = AnglePair{};
  double dE_dcos = 0.0;
  for (int row = r0; row < r1; ++row) {
    const double* pr = ff.ang + row * ANG_W;
    const double th00 = pr[0], v1 = pr[1], v2 = pr[2], coa1 = pr[3], v7 = pr[4], pen1 = pr[5], v4 = pr[6];
    const double theta0 = (180 - th00 * (1 - esb)) * DEG;
    const double pa = pow(bij, v4), pc = pow(bjk, v4);
    const double ea = exp(-p3 * pa), ec = exp(-p3 * pc);
    const double Fa = 1 - ea, Fc = 1 - ec;
    const double f7 = Fa * Fc;
    const double ex7 = exp(-v7 * db);
    const double den8 = 1 + ex6 + ex7;
    const double R = (2 + ex6) / den8;
    const double f8 = p5 - (p5 - 1) * R;
    const double dR = (g[14] * ex6 * den8 - (2 + ex6) * (g[14] * ex6 - v7 * ex7)) / (den8 * den8);
    const double dt = theta0 - theta;
    const double exang = exp(-v2 * (dt * dt));
    const double s = v1 >= 0 ? 1 - exang : -exang;
    o.e[0] += f7 * f8 * v1 * s;
    const double dE_dth = -f7 * f8 * v1 * 2 * v2 * dt * exang;
    dE_dcos += dE_dth * dth;
    o.A_sbo2 += -dE_dth * th00 * g[17] * esb * DEG;
    const double dE_df7 = f8 * v1 * s, dE_df8 = f7 * v1 * s;
    o.gA += dE_df7 * Fc * p3 * v4 * (bij > 0 ? pa / bij : 0.0) * ea;
    o.gC += dE_df7 * Fa * p3 * v4 * (bjk > 0 ? pc / bjk : 0.0) * ec;
    o.A_db += dE_df8 * (-(p5 - 1) * dR);
    // penalty
    const double Ebb = exp(-g[19] * ((bij - 2) * (bij - 2) + (bjk - 2) * (bjk - 2)));
    const double Ep = pen1 * cj.Q * Ebb;
    o.e[1] += Ep;
    o.A_d += pen1 * cj.dQ * Ebb;
    o.gA += Ep * (-2 * g[19] * (bij - 2));
    o.gC += Ep * (-2 * g[19] * (bjk - 2));
    // angle conjugation
    const double E1 = exp(-g[38] * ((tota - bij) * (tota - bij) + (totc - bjk) * (totc - bjk)));
    const double E2 = exp(-g[30] * ((bij - 1.5) * (bij - 1.5) + (bjk - 1.5) * (bjk - 1.5)));
    const double Ec = coa1 * cj.Cj * E1 * E2;
    o.e[2] += Ec;
    o.A_tot += coa1 * cj.dCj * E1 * E2;
    o.ta += Ec * (-2 * g[38] * (tota - bij));
    o.tc += Ec * (-2 * g[38] * (totc - bjk));
    o.gA += Ec * (2 * g[38] * (tota - bij) - 2 * g[30] * (bij - 1.5));
    o.gC += Ec * (2 * g[38] * (totc - bjk) - 2 * g[30] * (bjk - 1.5));
  }
  const double inv = 1.0 / (ra * rc);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o.GA[k] = dE_dcos * (vc[k] * inv - cosv * va[k] / (ra * ra));
    o.GC[k] = dE_dcos * (va[k] * inv - cosv * vc[k] / (rc * rc));
  }
  return true;
}

// One group of ANG_GROUP lanes per centre x; lane l owns bond slot p0 + l
// (rows longer than the group are taken in chunks).  The lane walks every
// other bond y of the row in row order and evaluates the angle of the pair
// (min, max): its own slot takes that angle's share for the bond it owns, in
// the order of the pairs, and the lane of the lower bond adds the centre's
// share, which is then reduced over the group.  Each angle is evaluated by
// both of its lanes: the work is doubled, but spread over a lane per bond
// instead of a thread per centre, which leaves most of the chip idle.  Every
// slot has one writer and a fixed order, so the result is deterministic and
// the same for a system alone or in a batch.
constexpr int ANG_GROUP = 8;
__global__ void __launch_bounds__(64) k_angles(Atoms at, Bonds bd, FFDev ff, int lammps, EnergyAcc acc) {
  const int x = blockIdx.x * (64 / ANG_GROUP) + (threadIdx.x & 63) / ANG_GROUP;
  const int lane = threadIdx.x & (ANG_GROUP - 1);
  if (x >= at.n) return;  // the whole group
  const AngleCentre cj = angle_centre(at, ff, x);
  const int p0 = bd.ptr[x], p1 = bd.ptr[x + 1];
  double e[3] = {0.0, 0.0, 0.0};
  double A_db = 0.0, A_d = 0.0, A_tot = 0.0, A_sbo2 = 0.0;
  for (int chunk = p0; chunk < p1; chunk += ANG_GROUP) {
    const int mine = chunk + lane;
    if (mine >= p1 || !(bd.BO[mine] > THB_CUT)) continue;
    double abo = bd.A_BO[mine], afar = bd.A_totfar[mine];
    double G[3] = {bd.G[3 * mine], bd.G[3 * mine + 1], bd.G[3 * mine + 2]};
    for (int y = p0; y < p1; ++y) {
      if (y == mine) continue;
      const bool low = mine < y;
      AnglePair o;
      if (!angle_pair(at, bd, ff, lammps, cj, low ? mine : y, low ? y : mine, o)) continue;
      abo += low ? o.gA : o.gC;
      afar += low ? o.ta : o.tc;
#pragma unroll
      for (int k = 0; k < 3; ++k) G[k] += low ? o.GA[k] : o.GC[k];
      if (low) {
#pragma unroll
        for (int k = 0; k < 3; ++k) e[k] += o.e[k];
        A_db += o.A_db;
        A_d += o.A_d;
        A_tot += o.A_tot;
        A_sbo2 += o.A_sbo2;
      }
    }
    bd.A_BO[mine] = abo;
    bd.A_totfar[mine] = afar;
#pragma unroll
    for (int k = 0; k < 3; ++k) bd.G[3 * mine + k] = G[k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) e[k] = xr_group_sum<ANG_GROUP>(e[k]);
  A_db = xr_group_sum<ANG_GROUP>(A_db);
  A_d = xr_group_sum<ANG_GROUP>(A_d);
  A_tot = xr_group_sum<ANG_GROUP>(A_tot);
  A_sbo2 = xr_group_sum<ANG_GROUP>(A_sbo2);
  if (lane == 0) {
    at.A_db[x] += A_db;
    at.A_d[x] += A_d;
    at.A_total[x] += A_tot;
    at.A_sbo2[x] += A_sbo2;
    xr_energy(acc, CH_ANGLE, x, e[0]);
    xr_energy(acc, CH_PEN, x, e[1]);
    xr_energy(acc, CH_COA, x, e[2]);
  }
}

// ---------------------------------------------------------------- torsions
struct TorsionGrad {
  double E_tor, E_conj;
  double Gu[3], Gv[3], Gw[3];
  double dB1, dB2, dB3, dPI, dDS;
};

__device__ inline bool torsion_eval(const Bonds& bd, const FFDev& ff, const double* prm, int a, int m, int c,
                                    double dbj, double dbk, TorsionGrad& o) {
  const double* g = ff.gen;
  const double* u = bd.vec + 3 * a;
  const double* v = bd.vec + 3 * m;
  const double* w = bd.vec + 3 * c;
  const double ra = bd.r[a], rm = bd.r[m], rcw = bd.r[c];
  double n1[3], n2[3];
  cross3(u, v, n1);
  cross3(w, v, n2);
  const double A1 = dot3(n1, n1), A2 = dot3(n2, n2);
  const double s1 = A1 / ((ra * rm) * (ra * rm));
  const double s2 = A2 / ((rm * rcw) * (rm * rcw));
  if (s1 < 1e-20 || s2 < 1e-20) return false;
  const double sp = sqrt(s1 * s2);
  const double P = dot3(n1, n2);
  const double sq = sqrt(A1 * A2);
  const double cw = P / sq;
  const double v1 = prm[0], v2 = prm[1], v3 = prm[2], tor1 = prm[3], cot1 = prm[4];
  const double b1 = bd.BO[a] - THB_CUT, b2 = bd.BO[m] - THB_CUT, b3 = bd.BO[c] - THB_CUT;
  const double ds = dbj + dbk;
  const double ex3 = exp(-g[24] * ds), ex4 = exp(g[25] * ds);
  const double den11 = 1 + ex3 + ex4;
  const double f11 = (2 + ex3) / den11;
  const double df11 = (-g[24] * ex3 * den11 - (2 + ex3) * (-g[24] * ex3 + g[25] * ex4)) / (den11 * den11);
  const double PIm = bd.PI[m];
  const double z = 2 - PIm - f11;
  const double ex1 = exp(tor1 * (z * z));
  const double e1 = exp(-g[23] * b1), e2 = exp(-g[23] * b2), e3 = exp(-g[23] * b3);
  const double fn10 = (1 - e1) * (1 - e2) * (1 - e3);
  const double cv = 0.5 * (v1 * (1 + cw) + v2 * ex1 * (2 - 2 * cw * cw) + v3 * (1 + 4 * (cw * cw * cw) - 3 * cw));
  o.E_tor = fn10 * sp * cv;
  const double q1 = b1 - 1.5, q2 = b2 - 1.5, q3 = b3 - 1.5;
  const double fn12 = exp(-g[27] * ((q1 * q1 + q2 * q2) + q3 * q3));
  o.E_conj = cot1 * fn12 * (1 + (cw * cw - 1) * sp);
  const double dcv = 0.5 * (v1 - 4 * v2 * ex1 * cw + v3 * (12 * cw * cw - 3));
  const double dE_dcw = fn10 * sp * dcv + cot1 * fn12 * 2 * cw * sp;
  const double dE_dsp = fn10 * cv + cot1 * fn12 * (cw * cw - 1);
  const double dE_dex1 = fn10 * sp * 0.5 * v2 * (2 - 2 * cw * cw);
  const double dE_dfn10 = sp * cv;
  const double dE_dfn12 = cot1 * (1 + (cw * cw - 1) * sp);
  o.dB1 = dE_dfn10 * g[23] * e1 * (1 - e2) * (1 - e3) + dE_dfn12 * fn12 * (-2 * g[27] * q1);
  o.dB2 = dE_dfn10 * g[23] * e2 * (1 - e1) * (1 - e3) + dE_dfn12 * fn12 * (-2 * g[27] * q2);
  o.dB3 = dE_dfn10 * g[23] * e3 * (1 - e1) * (1 - e2) + dE_dfn12 * fn12 * (-2 * g[27] * q3);
  const double dex1_dz = ex1 * tor1 * 2 * z;  // d ex1 / d z, z = 2 - PI - f11
  o.dPI = -dE_dex1 * dex1_dz;
  o.dDS = -dE_dex1 * dex1_dz * df11;
  const double EP = dE_dcw / sq;
  const double EA1 = dE_dcw * (-0.5 * cw / A1) + dE_dsp * (0.5 * sp / A1);
  const double EA2 = dE_dcw * (-0.5 * cw / A2) + dE_dsp * (0.5 * sp / A2);
  const double Era = -dE_dsp * sp / ra, Erm = -2 * dE_dsp * sp / rm, Erc = -dE_dsp * sp / rcw;
  double vn2[3], vn1[3], n2u[3], n1w[3], n1u[3], n2w[3];
  cross3(v, n2, vn2);
  cross3(v, n1, vn1);
  cross3(n2, u, n2u);
  cross3(n1, w, n1w);
  cross3(n1, u, n1u);
  cross3(n2, w, n2w);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o.Gu[k] = EP * vn2[k] + 2 * EA1 * vn1[k] + Era * u[k] / ra;
    o.Gv[k] = EP * (n2u[k] + n1w[k]) + 2 * EA1 * n1u[k] + 2 * EA2 * n2w[k] + Erm * v[k] / rm;
    o.Gw[k] = EP * vn1[k] + 2 * EA2 * vn2[k] + Erc * w[k] / rcw;
  }
  return true;
}

// Parameter key of torsion (a, m, c) around the half middle bond m = (j -> k),
// or -1 when the reference's enumeration (energy.py:323-344) skips it: a is a
// bond of j other than m, c a bond of k other than rev(m), all three above
// the 0.001 cutoff with product above it, parameters present, and the 3-ring
// i == l through the same image excluded.
__device__ inline int torsion_key(const Atoms& at, const Bonds& bd, const FFDev& ff, int a, int m, int c, int tj,
                                  int tk, int msx, int msy, int msz) {
  if (a == m || c == bd.rev[m]) return -1;
  const double ba = bd.BO[a], bc = bd.BO[c];
  if (!(ba > THB_CUT) || !(bc > THB_CUT)) return -1;
  const int i = bd.bj[a], l = bd.bj[c];
  const int nt = ff.nt;
  const int key = ((at.type[i] * nt + tj) * nt + tk) * nt + at.type[l];
  if (!ff.tor_on[key]) return -1;
  if (ba * bd.BO[m] * bc <= THB_CUT) return -1;
  if (i == l) {
    int asx, asy, asz, csx, csy, csz;
    xr_unpack_shift(bd.code[a], asx, asy, asz);
    xr_unpack_shift(bd.code[c], csx, csy, csz);
    if (asx == msx + csx && asy == msy + csy && asz == msz + csz) return -1;
  }
  return key;
}

// One group of G lanes per atom x: G = TOR_GROUP for atoms with at most
// TOR_GROUP bonds (several atoms per wave), G = 64 for the rest.  Every
// torsion contribution that lands on a bond of x's own row goes to the lane
// owning that slot: as atom j of a half middle bond m, x owns the left bonds
// a (lane of a loops over the right bonds c); as atom k, x owns the right
// bonds c (lane of c loops over the left bonds a).  The middle-bond terms of
// the j role and the energies are reduced over the group.  Each slot has one
// writer and the order is fixed, and which instance takes an atom depends
// only on its own bond count, so the result is deterministic and the same
// for a system alone or in a batch; rows longer than 64 bonds are taken in
// chunks.
constexpr int TOR_GROUP = 8;
template <int G>
__device__ inline void torsion_atom(const Atoms& at, const Bonds& bd, const FFDev& ff, const EnergyAcc& acc, int x,
                                    int lane, int p0, int p1) {
  double e_tor = 0.0, e_conj = 0.0, a_db = 0.0;
  unsigned err = 0;
  for (int chunk = p0; chunk < p1; chunk += G) {
    const int mine = chunk + lane;
    const bool owns = mine < p1;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, abo = 0.0;
    for (int mb = p0; mb < p1; ++mb) {
      if (!(bd.BO[mb] > THB_CUT)) continue;
      const int rv = bd.rev[mb];
      const bool half = bd.edge[mb] < bd.edge[rv];
      const int m = half ? mb : rv;
      const int j = bd.bi[m], k = bd.bj[m];
      const int tj = at.type[j], tk = at.type[k];
      if (!ff.tor_mid[tj * ff.nt + tk]) continue;  // no torsion row has this middle pair (uniform in the group)
      int msx, msy, msz;
      xr_unpack_shift(bd.code[m], msx, msy, msz);
      const double dbj = at.db[j], dbk = at.db[k];
      const double bm = bd.BO[m];
      if (half) {
        // x = j: this lane's left bond a = mine, right bonds c of k.  A lane whose left bond can
        // satisfy neither the parameter test (some row for (t_i, t_j, t_k)) nor the product test
        // with the largest bond order of k skips the loop: both bounds are exact (the product
        // (b_a b_m) b_c is monotone in b_c, rounding included), so no torsion is lost.
        double gm0 = 0.0, gm1 = 0.0, gm2 = 0.0, abm = 0.0, apim = 0.0;
        if (owns && ff.tor_left[(at.type[bd.bj[mine]] * ff.nt + tj) * ff.nt + tk] &&
            bd.BO[mine] * bm * at.bomax[k] > THB_CUT) {
          for (int c = bd.ptr[k]; c < bd.ptr[k + 1]; ++c) {
            const int key = torsion_key(at, bd, ff, mine, m, c, tj, tk, msx, msy, msz);
            if (key < 0) continue;
            TorsionGrad o;
            if (!torsion_eval(bd, ff, ff.tor + key * TOR_W, mine, m, c, dbj, dbk, o)) {
              err |= XR_ERR_COLLINEAR;
              continue;
            }
            e_tor += o.E_tor;
            e_conj += o.E_conj;
            g0 += o.Gu[0];
            g1 += o.Gu[1];
            g2 += o.Gu[2];
            abo += o.dB1;
            gm0 += o.Gv[0];
            gm1 += o.Gv[1];
            gm2 += o.Gv[2];
            abm += o.dB2;
            apim += o.dPI;
            a_db += o.dDS;
          }
        }
        gm0 = xr_group_sum<G>(gm0);
        gm1 = xr_group_sum<G>(gm1);
        gm2 = xr_group_sum<G>(gm2);
        abm = xr_group_sum<G>(abm);
        apim = xr_group_sum<G>(apim);
        if (lane == 0) {
          bd.G[3 * m] += gm0;
          bd.G[3 * m + 1] += gm1;
          bd.G[3 * m + 2] += gm2;
          bd.A_BO[m] += abm;
          bd.A_PI[m] += apim;
        }
      } else if (owns && ff.tor_right[(tj * ff.nt + tk) * ff.nt + at.type[bd.bj[mine]]] &&
                 at.bomax[j] * bm * bd.BO[mine] > THB_CUT) {
        // x = k: this lane's right bond c = mine, left bonds a of j (pruned as above: the product
        // (b_a b_m) b_c is monotone in b_a)
        for (int a = bd.ptr[j]; a < bd.ptr[j + 1]; ++a) {
          const int key = torsion_key(at, bd, ff, a, m, mine, tj, tk, msx, msy, msz);
          if (key < 0) continue;
          TorsionGrad o;
          if (!torsion_eval(bd, ff, ff.tor + key * TOR_W, a, m, mine, dbj, dbk, o)) {
            err |= XR_ERR_COLLINEAR;
            continue;
          }
          g0 += o.Gw[0];
          g1 += o.Gw[1];
          g2 += o.Gw[2];
          abo += o.dB3;
          a_db += o.dDS;
        }
      }
    }
    if (owns) {
      bd.G[3 * mine] += g0;
      bd.G[3 * mine + 1] += g1;
      bd.G[3 * mine + 2] += g2;
      bd.A_BO[mine] += abo;
    }
  }
  e_tor = xr_group_sum<G>(e_tor);
  e_conj = xr_group_sum<G>(e_conj);
  a_db = xr_group_sum<G>(a_db);
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) err |= __shfl_down(err, off, G);
  if (lane == 0) {
    at.A_db[x] += a_db;
    if (err) xr_flag(at.S, x, err);
    xr_energy(acc, CH_TOR, x, e_tor);
    xr_energy(acc, CH_CONJ, x, e_conj);
  }
}

// Atoms with at most TOR_GROUP bonds, eight per wave.  An atom with more is
// appended to `long_atoms` (order immaterial: each atom's work writes only its
// own slots) for k_torsions_long, so that launch does not sweep every atom.
__global__ void __launch_bounds__(256) k_torsions(Atoms at, Bonds bd, FFDev ff, EnergyAcc acc,
                                                  int* __restrict__ long_atoms, unsigned* __restrict__ long_count) {
  constexpr int G = TOR_GROUP;
  const int x = (blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 / G) + ((threadIdx.x & 63) / G);
  const int lane = threadIdx.x & (G - 1);
  if (x >= at.n) return;
  const int p0 = bd.ptr[x], p1 = bd.ptr[x + 1];
  if (p1 - p0 > G) {
    if (lane == 0) long_atoms[atomicAdd(long_count, 1u)] = x;
    return;
  }
  torsion_atom<G>(at, bd, ff, acc, x, lane, p0, p1);
}

// The listed atoms, a whole wave each; every wave strides over the list.
__global__ void __launch_bounds__(256) k_torsions_long(Atoms at, Bonds bd, FFDev ff, EnergyAcc acc,
                                                       const int* __restrict__ long_atoms,
                                                       const unsigned* __restrict__ long_count) {
  const int lane = threadIdx.x & 63;
  const int count = (int)*long_count;
  for (int w = blockIdx.x * 4 + (threadIdx.x >> 6); w < count; w += gridDim.x * 4) {
    const int x = long_atoms[w];
    torsion_atom<64>(at, bd, ff, acc, x, lane, bd.ptr[x], bd.ptr[x + 1]);
  }
}

// ---------------------------------------------------------------- H-bonds
struct HbTerm {
  double E, dBO;
  double Gd[3], Ga[3];
};

__device__ inline void hbond_eval(const double* prm, double BOd, const double* vd, double rd, const double* va,
                                  double ra, HbTerm& o) {
  const double r0 = prm[0], hb1 = prm[1], hb2 = prm[2], hb3 = prm[3];
  const double cosv = dot3(vd, va) / (rd * ra);
  const double cc = fmin(fmax(cosv, -1.0), 1.0);
  const double sin4 = 0.25 * ((1 - cc) * (1 - cc));
  const double eb = exp(-hb2 * BOd);
  const double Ex = exp(-hb3 * (r0 / ra + ra / r0 - 2));
  o.E = hb1 * (1 - eb) * Ex * sin4;
  o.dBO = hb1 * hb2 * eb * Ex * sin4;
  const double dE_ddist = hb1 * (1 - eb) * Ex * sin4 * (-hb3) * (-r0 / (ra * ra) + 1 / r0);
  const double dE_dcos = (cc == -1.0 || cc == 1.0) ? 0.0 : hb1 * (1 - eb) * Ex * (-0.5 * (1 - cc));
  const double inv = 1.0 / (rd * ra);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o.Gd[k] = dE_dcos * (va[k] * inv - cosv * vd[k] / (rd * rd));
    o.Ga[k] = dE_dcos * (vd[k] * inv - cosv * va[k] / (ra * ra)) + dE_ddist * va[k] / ra;
  }
}

// Donor side of hydrogen j (one wave): energy and derivatives.  The row's
// candidates (acceptors within min(7.5, rc), compacted by the edge pass in
// row order) are taken one per lane for each donor bond; a candidate that is
// the donor bond's own edge (same atom and image) is skipped.
__device__ inline void hb_donor(int j, const Atoms& at, const Bonds& bd, const FFDev& ff,
                                const int2* __restrict__ cand, int count, const EnergyAcc& acc) {
  const int lane = threadIdx.x & 63;
  const CellDev& cell = xr_cell(at.S, j);
  const int nt = ff.nt, tj = at.type[j];
  const double hcut = fmin(7.5, ff.gen[12]);
  double etot = 0.0;
  double F[3] = {0, 0, 0};
  for (int d = bd.ptr[j]; d < bd.ptr[j + 1]; ++d) {
    const int kd = bd.bj[d];
    if (AP(ff, at.type[kd])[AP_HBOND] != 2.0 || !(bd.BO[d] >= 0.01)) continue;
    const double* vd = bd.vec + 3 * d;
    const double rd = bd.r[d], BOd = bd.BO[d];
    const int cd = bd.code[d];
    const int kbase = (at.type[kd] * nt + tj) * nt;
    double Gd[3] = {0, 0, 0}, dBO = 0.0;
    for (int t = lane; t < count; t += 64) {
      const int2 c = cand[t];
      const int k = c.x;
      if (k == kd && c.y == cd) continue;
      const int tk = at.type[k];
      if (!ff.hb_on[kbase + tk]) continue;
      double va[3];
      const double ra = xr_edge(at.x, cell, j, k, c.y, va);
      if (!(ra <= hcut)) continue;
      HbTerm o;
      hbond_eval(ff.hb + (kbase + tk) * HB_W, BOd, vd, rd, va, ra, o);
      etot += o.E;
      dBO += o.dBO;
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) {
        Gd[k2] += o.Gd[k2];
        F[k2] += o.Ga[k2];
      }
    }
    dBO = xr_wave_sum(dBO);
#pragma unroll
    for (int k2 = 0; k2 < 3; ++k2) Gd[k2] = xr_wave_sum(Gd[k2]);
    if (lane == 0) {
      bd.A_BO[d] += dBO;
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) bd.G[3 * d + k2] += Gd[k2];
    }
  }
#pragma unroll
  for (int k2 = 0; k2 < 3; ++k2) F[k2] = xr_wave_sum(F[k2]);
  etot = xr_wave_sum(etot);
  if (lane == 0) {
#pragma unroll
    for (int k2 = 0; k2 < 3; ++k2) at.F_hb[3 * j + k2] += F[k2];
    xr_energy(acc, CH_HB, j, etot);
  }
}

// Acceptor side of heavy atom k (one wave): force from the acceptor-vector
// derivative.  Each lane takes one candidate hydrogen of the row (compacted
// by the edge pass) and walks its donor bonds.
__device__ inline void hb_acceptor(int k, const Atoms& at, const Bonds& bd, const FFDev& ff,
                                   const int2* __restrict__ cand, int count) {
  const int lane = threadIdx.x & 63;
  const int tk = at.type[k];
  const CellDev& cell = xr_cell(at.S, k);
  const int nt = ff.nt;
  const double hcut = fmin(7.5, ff.gen[12]);
  double F[3] = {0, 0, 0};
  for (int t = lane; t < count; t += 64) {
    const int2 c = cand[t];
    const int j = c.x;
    const int tj = at.type[j];
    const int code = c.y;
    double vk[3];
    const double ra = xr_edge(at.x, cell, k, j, code, vk);
    if (!(ra <= hcut)) continue;
    const double va[3] = {-vk[0], -vk[1], -vk[2]};
    const int back = xr_neg_shift(code);
    for (int d = bd.ptr[j]; d < bd.ptr[j + 1]; ++d) {
      const int kd = bd.bj[d];
      const int td = at.type[kd];
      if (AP(ff, td)[AP_HBOND] != 2.0 || !(bd.BO[d] >= 0.01)) continue;
      if (kd == k && bd.code[d] == back) continue;  // the acceptor edge is the donor edge
      const int key = (td * nt + tj) * nt + tk;
      if (!ff.hb_on[key]) continue;
      HbTerm o;
      hbond_eval(ff.hb + key * HB_W, bd.BO[d], bd.vec + 3 * d, bd.r[d], va, ra, o);
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) F[k2] -= o.Ga[k2];
    }
  }
#pragma unroll
  for (int k2 = 0; k2 < 3; ++k2) F[k2] = xr_wave_sum(F[k2]);
  if (lane == 0) {
#pragma unroll
    for (int k2 = 0; k2 < 3; ++k2) at.F_hb[3 * k + k2] += F[k2];
  }
}

// One wave per atom: a hydrogen of H-bond type 1 takes its donor side, a
// heavy atom of type 2 its acceptor side.  The two sides write disjoint slots
// (the hydrogen's force and donor bonds, the acceptor's force), so one launch
// holds both.
__global__ void __launch_bounds__(256) k_hbonds(Atoms at, Bonds bd, FFDev ff, const int* __restrict__ row_ptr,
                                                const int2* __restrict__ hb_cand, const int* __restrict__ hb_count,
                                                EnergyAcc acc) {
  const int x = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (x >= at.n) return;
  const double role = AP(ff, at.type[x])[AP_HBOND];
  if (role != 1.0 && role != 2.0) return;
  const int2* cand = hb_cand + row_ptr[x];
  const int count = hb_count[x];
  if (role == 1.0) hb_donor(x, at, bd, ff, cand, count, acc);
  else hb_acceptor(x, at, bd, ff, cand, count);
}

// Largest corrected bond order over each atom's bonds (0 without bonds).
__global__ void k_atom_bomax(Atoms at, Bonds bd) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= at.n) return;
  double m = 0.0;
  for (int b = bd.ptr[x]; b < bd.ptr[x + 1]; ++b) m = fmax(m, bd.BO[b]);
  at.bomax[x] = m;
}

// One wave per block: per-atom and per-bond kernels have a few thousand
// threads, and small blocks spread them over every CU.
constexpr int BLK = 64;
inline dim3 grid(int count) { return dim3(count > 0 ? (count + BLK - 1) / BLK : 1); }

}  // namespace

hipError_t xr_manybody(const Atoms& at, const Bonds& bd, FFDev ff, const int* row_ptr, const int2* hb_cand,
                       const int* hb_count, int lammps_angles, EnergyAcc acc, int* long_atoms,
                       unsigned* long_count, hipStream_t s) {
  const int angle_groups = 64 / ANG_GROUP;
  hipLaunchKernelGGL(k_angles, dim3(at.n > 0 ? (at.n + angle_groups - 1) / angle_groups : 1), dim3(64), 0, s, at, bd,
                     ff, lammps_angles, acc);
  hipLaunchKernelGGL(k_atom_bomax, grid(at.n), dim3(BLK), 0, s, at, bd);
  const dim3 wgrid((at.n + 3) / 4 > 0 ? (at.n + 3) / 4 : 1);
  const int per_block = 4 * (64 / TOR_GROUP);
  hipLaunchKernelGGL(k_torsions, dim3(at.n > 0 ? (at.n + per_block - 1) / per_block : 1), dim3(256), 0, s, at, bd,
                     ff, acc, long_atoms, long_count);
  // at most 4,096 waves stride over the atoms with more than TOR_GROUP bonds
  const int long_blocks = std::max(1, std::min((at.n + 3) / 4, 1024));
  hipLaunchKernelGGL(k_torsions_long, dim3(long_blocks), dim3(256), 0, s, at, bd, ff, acc, long_atoms, long_count);
  hipLaunchKernelGGL(k_hbonds, wgrid, dim3(256), 0, s, at, bd, ff, row_ptr, hb_cand, hb_count, acc);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}
